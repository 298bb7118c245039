"""Host-side libMesh API subset (liblmesh.so): build_cube + DofMap numbering and
the synthetic inputs, bit-identical to the oracle; error handling."""
import numpy as np
import pytest

import libmesh_amd as L
import oracle as O


@pytest.mark.parametrize("t,n", [(L.HEX8, 5), (L.HEX27, 4), (L.TET4, 2), (L.TET10, 2)])
@pytest.mark.parametrize("nv,mask", [(1, L.ALL_FACES), (3, 1 << 4)])
def test_mesh_and_dofs_match_oracle(t, n, nv, mask):
    p = L.Problem(n, t, nv, mask)
    m = O.Mesh(n, t)
    s = O.System(m, nv, mask)
    assert np.array_equal(p.coords, m.coords())
    assert np.array_equal(p.conn, m.conn())
    assert np.array_equal(p.elem_dofs, s.elem_dofs())
    assert np.array_equal(p.constrained, s.constrained())
    assert p.n_dofs == s.n_dofs


def test_synthetic_inputs_match_oracle():
    p = L.Problem(6, L.HEX27)
    p.perturb(0x5EED0001, 0.2)
    m = O.Mesh(6, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    assert np.array_equal(p.coords, m.coords())
    assert np.array_equal(L.lognormal(5000, 0x5EED0002), O.kappa_lognormal(5000, 0x5EED0002))
    k = L.lognormal(200000, 1)
    z = np.log(k)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_bad_arguments():
    with pytest.raises(L.LmhError):
        L.Problem(0, L.HEX27)
    with pytest.raises(L.LmhError):
        L.Problem(2, 99)


def test_batch_recipe():
    """Seeded batch recipe: deterministic, edges in [nmin, nmax] spread
    log-uniformly, all four element types drawn."""
    e1, t1 = L.batch_recipe(0x5EED0003, 64, 8, 48)
    e2, t2 = L.batch_recipe(0x5EED0003, 64, 8, 48)
    assert np.array_equal(e1, e2) and np.array_equal(t1, t2)
    assert e1.min() >= 8 and e1.max() <= 48
    assert set(t1.tolist()) == {L.HEX8, L.HEX27, L.TET4, L.TET10}
    e, t = L.batch_recipe(7, 4000, 8, 48)
    # log-uniform: the median edge is near the geometric mean sqrt(8*48) ~ 19.6
    assert 17 <= np.median(e) <= 22
    assert abs((e <= 12).mean() - np.log(12.5 / 8) / np.log(48 / 8)) < 0.05
    for k in (L.HEX8, L.HEX27, L.TET4, L.TET10):
        assert abs((t == k).mean() - 0.25) < 0.04
    e3, _ = L.batch_recipe(0x5EED0004, 64, 8, 48)
    assert not np.array_equal(e1, e3)
