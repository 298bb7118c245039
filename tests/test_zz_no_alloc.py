"""lmh_cg_solve_noalloc (work vectors in the library's fixed device scratch
array, no allocation) against lmh_cg_solve on the same matrices: two calls in
a row, the second on a smaller problem, so the second call runs on scratch that
still holds the first call's larger vectors.  Bar: the same iteration count,
reason and residual and bit-identical solutions (the same kernels in the same
order), plus the oracle's Eigen CG within 1e-8 on the first problem.  A matrix
above the scratch limit is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import libmesh_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


def solve_both(ctx, n, t, npts, seed, noalloc_first):
    p = L.Problem(n, t)
    p.perturb(seed, 0.2)
    kappa = L.lognormal(p.n_elem, seed + 1)
    g = L.GpuProblem.from_problem(ctx, p)
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(npts, kd, 0, b)
    xa = L.DeviceArray(ctx, p.n_dofs)
    xb = L.DeviceArray(ctx, p.n_dofs)
    # a nonzero initial guess exercises the initial residual in scratch
    x0 = np.random.default_rng(seed).standard_normal(p.n_dofs) * 1e-3
    xa.upload(x0)
    xb.upload(x0)
    print(f"noalloc CG: {n}^3 type {t} dofs {p.n_dofs} noalloc_first {noalloc_first}", flush=True)
    if noalloc_first:
        rb = g.cg_noalloc(b, xb, 1e-10, 5000)
        ra = g.cg(b, xa, 1e-10, 5000)
    else:
        ra = g.cg(b, xa, 1e-10, 5000)
        rb = g.cg_noalloc(b, xb, 1e-10, 5000)
    assert ra[2] == L.CONVERGED_RTOL and ra[1] <= 1e-10, ra
    assert rb == ra, (rb, ra)
    a, c = xa.download(), xb.download()
    assert np.array_equal(a, c), np.abs(a - c).max()
    return p, g, kappa, a, ra


def test_noalloc_cg_matches_cg_twice(ctx):
    p, g, kappa, x, (its, res, _) = solve_both(ctx, 10, L.HEX27, 3, 0x5EED0101, False)
    m = O.Mesh(10, O.HEX27)
    m.perturb(0x5EED0101, 0.2)
    s = O.System(m)
    ov, orhs = s.assemble_poisson(3, kappa, 0)
    rp, ci = s.pattern()
    ox, oits, _ = O.cg_eigen(rp, ci, ov, orhs, 1e-10, 5000)
    assert abs(its - oits) <= 2
    assert np.linalg.norm(x - ox) / np.linalg.norm(ox) < 1e-8
    g.close()
    # second call, smaller problem, fresh matrix (its x windows come with it)
    p2, g2, *_ = solve_both(ctx, 5, L.HEX8, 2, 0x5EED0202, True)
    g2.close()
    # and a third on the same smaller size with the windows already built
    p3, g3, *_ = solve_both(ctx, 6, L.HEX27, 3, 0x5EED0303, False)
    g3.close()


def test_noalloc_cg_refuses_beyond_scratch(ctx):
    n = L.CG_NOALLOC_MAX_ROWS + 1
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    h = C.c_void_p()
    L.check(L.hip().lmh_matrix_create_csr(ctx.h, n, L._ptr(rp), L._ptr(ci), C.byref(h)))
    try:
        ones = np.ones(n)  # held: the call reads it
        L.check(L.hip().lmh_matrix_set_values(h, L._ptr(ones)))
        b = L.DeviceArray(ctx, n, data=np.ones(n))
        x = L.DeviceArray(ctx, n)
        x.zero()
        its, res, reason = C.c_int64(), C.c_double(), C.c_int()
        rc = L.hip().lmh_cg_solve_noalloc(h, b.ptr, x.ptr, 1e-10, 10, C.byref(its), C.byref(res), C.byref(reason))
        assert rc != 0
        assert b"scratch limit" in L.hip().lmh_last_error()
        # the allocating solver takes the same matrix
        L.check(L.hip().lmh_cg_solve(h, b.ptr, x.ptr, 1e-10, 10, C.byref(its), C.byref(res), C.byref(reason)))
        assert reason.value == L.CONVERGED_RTOL and np.array_equal(x.download(), np.ones(n))
        b.free()
        x.free()
    finally:
        L.hip().lmh_matrix_destroy(h)
