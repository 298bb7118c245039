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
