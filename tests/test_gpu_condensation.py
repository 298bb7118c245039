"""Static condensation on the device (lmh_condensation_*, StaticCondensation in
include/numerics/static_condensation.h:64) against the oracle's element-by-element
restatement (oracle.StaticCondensation, itself pinned in test_condensation_pins.py).

Bars: reduced pattern and reduced-DoF list bit-exact; reduced values and the
condensed right-hand side within 1e-12 (normwise relative); the condensed CG
solve within 1e-9 of the direct full solve, iterations within 2 of the oracle's
Eigen CG on the oracle's reduced matrix.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import libmesh_amd as L
import oracle as O

pytestmark = pytest.mark.gpu
NPTS = {L.HEX8: 2, L.HEX27: 3, L.TET4: 3, L.TET10: 5}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pair(ctx, t, n, seed=5, frac=0.2):
    p = L.Problem(n, t, 1, L.ALL_FACES)
    s = O.System(O.Mesh(n, t), 1, L.ALL_FACES)
    p.perturb(seed, frac)
    s.mesh.perturb(seed, frac)
    assert np.array_equal(p.coords, s.mesh.coords())
    g = L.GpuProblem.from_problem(ctx, p)
    kappa = L.lognormal(p.n_elem, seed)
    ov, orhs = s.assemble_poisson(NPTS[t], kappa, 1)
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(NPTS[t], kd, 1, b, L.FP64, L.SCATTER_GATHER)
    assert rel(g.values(), ov) < 1e-12
    return p, s, g, kappa, kd, b, ov, orhs


@pytest.mark.parametrize("n", [3, 8])
def test_condensed_operator_and_solve(ctx, n):
    p, s, g, kappa, kd, b, ov, orhs = pair(ctx, L.HEX27, n)
    Ke, _ = s.element_poisson(NPTS[L.HEX27], kappa, 1)
    osc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26])
    cd = g.condensed_dofs()
    assert np.array_equal(np.sort(cd), osc.cdofs) and cd.size == p.n_elem
    sc = L.StaticCondensation(g)
    print(f"n {n} dofs {p.n_dofs} reduced {sc.n_reduced} nnz {g.nnz} -> {sc.nnz}", flush=True)
    assert np.array_equal(sc.reduced_dofs(), osc.rfull)
    rp, ci = sc.pattern()
    assert np.array_equal(rp, osc.S.indptr) and np.array_equal(ci, osc.S.indices)
    assert rel(sc.values(), osc.S.data) < 1e-12
    v_first = sc.values()
    sc.setup()  # values again through the LDS-staged wavefront-per-row kernel
    # the same subtractions in the same order per entry as the first (checking) pass
    assert np.array_equal(sc.values(), v_first)
    gr = L.DeviceArray(ctx, sc.n_reduced)
    sc.forward(b, gr)
    og = osc.forward(orhs)
    assert rel(gr.download(), og) < 1e-12
    # back substitution of the oracle's exact reduced solution
    y = spl.spsolve(osc.S.tocsc(), og)
    yd = L.DeviceArray(ctx, sc.n_reduced, data=y)
    xd = L.DeviceArray(ctx, p.n_dofs)
    sc.backward(b, yd, xd)
    assert rel(xd.download(), osc.backward(orhs, y)) < 1e-12
    # apply(): forward elimination, Jacobi-CG on the reduced matrix, back substitution
    xf = spl.spsolve(sp.csr_matrix((ov, *s.pattern()[::-1]), shape=(p.n_dofs,) * 2).tocsc(), orhs)
    x = L.DeviceArray(ctx, p.n_dofs)
    x.zero()
    its, res, reason = sc.solve(b, x, 1e-12, 5000)
    _, oits, _ = O.cg_eigen(osc.S.indptr.astype(np.int32), osc.S.indices.astype(np.int32), osc.S.data, og, 1e-12,
                            5000)
    fits, _, _ = g.cg(b, L.DeviceArray(ctx, p.n_dofs, data=np.zeros(p.n_dofs)), 1e-12)
    print(f"condensed its {its} oracle {oits} full-system its {fits} res {res:.2e}", flush=True)
    assert reason == L.CONVERGED_RTOL and abs(its - oits) <= 2
    assert rel(x.download(), xf) < 1e-9


def test_setup_after_new_values(ctx):
    p, s, g, kappa, kd, b, ov, orhs = pair(ctx, L.HEX27, 4)
    sc = L.StaticCondensation(g)
    kappa2 = kappa * np.linspace(0.5, 2.0, p.n_elem)
    kd2 = L.DeviceArray(ctx, p.n_elem, data=kappa2)
    g.assemble_poisson(NPTS[L.HEX27], kd2, 1, b, L.FP64, L.SCATTER_GATHER)
    sc.setup()
    Ke, _ = s.element_poisson(NPTS[L.HEX27], kappa2, 1)
    osc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26])
    assert rel(sc.values(), osc.S.data) < 1e-12
    # the reduced operator is an ordinary matrix: the mixed solver runs on it
    L.check(L.hip().lmh_matrix_make_fp32(sc.matrix))
    gr = L.DeviceArray(ctx, sc.n_reduced)
    sc.forward(b, gr)
    y = L.DeviceArray(ctx, sc.n_reduced)
    y.zero()
    its, res, reason = C.c_int64(), C.c_double(), C.c_int()
    stats = np.zeros(3)
    L.check(L.hip().lmh_mixed_solve(sc.matrix, gr.ptr, y.ptr, 1e-10, 0.1, 0, 5000, C.byref(its), C.byref(res),
                                    C.byref(reason), L._ptr(stats)))
    assert reason.value == L.CONVERGED_RTOL
    assert rel(y.download(), spl.spsolve(osc.S.tocsc(), osc.forward(s.assemble_poisson(3, kappa2, 1)[1]))) < 1e-8


def test_no_internal_nodes_and_errors(ctx):
    # TET10 has no internal node: the reduced system is the full one
    p, s, g, kappa, kd, b, ov, orhs = pair(ctx, L.TET10, 3, frac=0.0)
    assert g.condensed_dofs().size == 0
    sc = L.StaticCondensation(g)
    assert sc.n_reduced == p.n_dofs and sc.nnz == g.nnz
    assert np.array_equal(sc.values(), g.values())
    x = L.DeviceArray(ctx, p.n_dofs)
    x.zero()
    its, res, reason = sc.solve(b, x, 1e-10)
    x2 = L.DeviceArray(ctx, p.n_dofs)
    x2.zero()
    its2, _, _ = g.cg(b, x2, 1e-10)
    assert its == its2 and np.array_equal(x.download(), x2.download())
    sc.close()
    # condensed DoFs that couple to each other (nonzero K_ij) are refused; a zero
    # coupling (a Dirichlet neighbour) is allowed
    p, s, g, kappa, kd, b, ov, orhs = pair(ctx, L.HEX27, 3)
    rp, ci = g.pattern()
    i = int(g.condensed_dofs()[0])
    gv = g.values()
    j = int(next(ci[k] for k in range(rp[i], rp[i + 1]) if ci[k] != i and gv[k] != 0.0))
    with pytest.raises(L.LmhError, match="diagonal condensed block"):
        L.StaticCondensation(g, np.array([i, j], dtype=np.int32))
    with pytest.raises(L.LmhError, match="listed twice"):
        L.StaticCondensation(g, np.array([i, i], dtype=np.int32))
    with pytest.raises(L.LmhError, match="out of range"):
        L.StaticCondensation(g, np.array([p.n_dofs], dtype=np.int32))


def _csr_problem(ctx, K, pattern):
    """A matrix with a caller-given pattern (lmh_matrix_create_csr) and K's values on it."""
    import types

    rows = [np.flatnonzero(pattern[i]).astype(np.int32) for i in range(K.shape[0])]
    rp = np.zeros(K.shape[0] + 1, dtype=np.int32)
    rp[1:] = np.cumsum([r.size for r in rows])
    ci = np.concatenate(rows).astype(np.int32)
    h = L._P()
    L.check(L.hip().lmh_matrix_create_csr(ctx.h, K.shape[0], L._ptr(rp), L._ptr(ci), C.byref(h)))
    v = np.ascontiguousarray(np.concatenate([K[i, r] for i, r in enumerate(rows)]))
    L.check(L.hip().lmh_matrix_set_values(h, L._ptr(v)))
    return types.SimpleNamespace(ctx=ctx, matrix=h), rp, ci


def test_structural_zero_coupling_and_structural_fill(ctx):
    """Condensed DoFs 0 and 1 share a structural entry holding an exact zero (the
    condensed block is still diagonal, so it is admitted).  Back substitution
    must not read the other condensed DoF's output: the output vector starts as
    NaN.  A condensed row reaching a column outside a reduced row it couples to
    is refused at creation even when that entry is zero then, since a later
    reassembly could make it nonzero and the refresh kernel has no such check."""
    n = 6
    rng = np.random.default_rng(3)
    M = rng.standard_normal((n, n))
    K = M @ M.T + n * np.eye(n)
    K[0, 1] = K[1, 0] = 0.0
    prob, rp, ci = _csr_problem(ctx, K, np.ones((n, n), dtype=bool))
    sc = L.StaticCondensation(prob, np.array([1, 0], dtype=np.int32))
    S = K[2:, 2:] - K[2:, :2] @ np.linalg.inv(K[:2, :2]) @ K[:2, 2:]
    assert rel(sc.values(), S.ravel()) < 1e-13
    b = rng.standard_normal(n)
    bd = L.DeviceArray(ctx, n, data=b)
    g = L.DeviceArray(ctx, n - 2)
    sc.forward(bd, g)
    assert rel(g.download(), b[2:] - K[2:, :2] @ np.linalg.solve(K[:2, :2], b[:2])) < 1e-13
    yd = L.DeviceArray(ctx, n - 2, data=np.linalg.solve(S, g.download()))
    x = L.DeviceArray(ctx, n, data=np.full(n, np.nan))
    sc.backward(bd, yd, x)
    assert rel(x.download(), np.linalg.solve(K, b)) < 1e-12
    sc.close()
    L.check(L.hip().lmh_matrix_destroy(prob.matrix))
    # row 0 couples to column 5, reduced row 2 (which holds column 0) does not: fill,
    # refused although K[0, 5] is zero now
    pat = np.ones((n, n), dtype=bool)
    pat[2, 5] = pat[5, 2] = False
    K2 = K.copy()
    K2[0, 5] = K2[5, 0] = 0.0
    prob, rp, ci = _csr_problem(ctx, K2, pat)
    with pytest.raises(L.LmhError, match="outside the matrix pattern"):
        L.StaticCondensation(prob, np.array([0], dtype=np.int32))
    L.check(L.hip().lmh_matrix_destroy(prob.matrix))


@pytest.mark.parametrize("n", [7, 400])
def test_refresh_skips_zero_couplings(ctx, n):
    """A reduced row whose structural coupling to a condensed DoF holds an exact
    zero (K_20 = 0): the first pass and every refresh skip that condensed
    column and give the dense Schur complement, bit-identically to each other.
    n = 7 refreshes through the position map; rows of 400 entries exceed its
    limit (384), so n = 400 refreshes through the LDS search kernel."""
    rng = np.random.default_rng(5)
    M = rng.standard_normal((n, n))
    K = M @ M.T + n * np.eye(n)
    K[2, 0] = K[0, 2] = 0.0
    K[0, 1] = K[1, 0] = 0.0
    prob, rp, ci = _csr_problem(ctx, K, np.ones((n, n), dtype=bool))
    sc = L.StaticCondensation(prob, np.array([0, 1], dtype=np.int32))
    S = K[2:, 2:] - K[2:, :2] @ np.linalg.inv(K[:2, :2]) @ K[:2, 2:]
    v_first = sc.values()
    assert rel(v_first, S.ravel()) < 1e-13
    for _ in range(2):
        sc.setup()
        assert np.array_equal(sc.values(), v_first)
    sc.close()
    L.check(L.hip().lmh_matrix_destroy(prob.matrix))
