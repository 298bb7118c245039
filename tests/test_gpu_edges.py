"""Edge cases of the GPU path against the oracle: single-element meshes (every
DoF of a Q1 element constrained, one free DoF of a Q2 element), boxes with
unequal edge counts, no Dirichlet rows at all, and a zero right-hand side.
Bars as in test_gpu_parity: pattern bit-exact, FP64 entries and rhs within
1e-12 normwise, CG iterations within 2 of the oracle's."""
import numpy as np
import pytest

import libmesh_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

NPTS = {L.HEX8: 2, L.HEX27: 3, L.TET4: 3, L.TET10: 5}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pair(ctx, t, n, ny=None, nz=None, mask=L.ALL_FACES):
    p = L.Problem(n, t, 1, mask, ny, nz)
    s = O.System(O.Mesh(n, t, ny, nz), 1, mask)
    assert np.array_equal(p.elem_dofs, s.elem_dofs())
    return p, s, L.GpuProblem.from_problem(ctx, p)


def check_assembly(ctx, p, s, g, t, f_kind=1):
    rp, ci = s.pattern()
    grp, gci = g.pattern()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    kappa = L.lognormal(p.n_elem, 7)
    ov, orhs = s.assemble_poisson(NPTS[t], kappa, f_kind)
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    for scatter in (L.SCATTER_GATHER, L.SCATTER_COLORED):
        g.assemble_poisson(NPTS[t], kd, f_kind, b, L.FP64, scatter)
        assert rel(g.values(), ov) < 1e-12
        assert rel(b.download(), orhs) <= 1e-12 or np.abs(orhs).max() == 0.0
    return rp, ci, ov, orhs, b


@pytest.mark.parametrize("t", [L.HEX8, L.HEX27, L.TET4, L.TET10])
def test_single_element_meshes(ctx, t):
    p, s, g = pair(ctx, t, 1)
    rp, ci, ov, orhs, b = check_assembly(ctx, p, s, g, t)
    ox, oits, _ = O.cg_eigen(rp, ci, ov, orhs, 1e-10, 100)
    x = L.DeviceArray(ctx, p.n_dofs)
    x.zero()
    its, res, reason = g.cg(b, x, 1e-10, 100)
    assert reason == L.CONVERGED_RTOL and abs(its - oits) <= 2
    assert np.abs(x.download() - ox).max() <= 1e-12 * max(1.0, np.abs(ox).max())


@pytest.mark.parametrize("t,n,ny,nz", [(L.HEX27, 3, 2, 1), (L.HEX8, 5, 1, 3), (L.TET10, 2, 1, 1)])
def test_unequal_edges(ctx, t, n, ny, nz):
    p, s, g = pair(ctx, t, n, ny, nz)
    rp, ci, ov, orhs, b = check_assembly(ctx, p, s, g, t)
    ox, oits, _ = O.cg_eigen(rp, ci, ov, orhs, 1e-10, 2000)
    x = L.DeviceArray(ctx, p.n_dofs)
    x.zero()
    its, res, reason = g.cg(b, x, 1e-10, 2000)
    assert reason == L.CONVERGED_RTOL and abs(its - oits) <= 2
    assert rel(x.download(), ox) < 1e-9


@pytest.mark.parametrize("t", [L.HEX27, L.TET4])
def test_no_dirichlet_rows(ctx, t):
    p, s, g = pair(ctx, t, 2, mask=0)
    assert not p.constrained.any()
    check_assembly(ctx, p, s, g, t)


def test_zero_rhs(ctx):
    """b = 0: Eigen's CG returns x = 0 with 0 iterations (ConjugateGradient.h:
    rhsNorm2 == 0), from any initial guess."""
    p, s, g = pair(ctx, L.HEX27, 2)
    kd = L.DeviceArray(ctx, p.n_elem, data=np.ones(p.n_elem))
    b = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(3, kd, 0, b)
    b.zero()
    x = L.DeviceArray(ctx, p.n_dofs, data=np.random.default_rng(0).standard_normal(p.n_dofs))
    its, res, reason = g.cg(b, x, 1e-10, 100)
    assert its == 0 and reason == L.CONVERGED_RTOL
    assert not x.download().any()


def test_solver_failure_reasons(ctx):
    """Reasons the solvers report instead of a result: an iteration limit
    (DIVERGED_ITS with the residual reached), a NaN or Inf in the right-hand side
    (DIVERGED_NANORINF before the first iteration, as PETSc reports it), and a
    zero operator (DIVERGED_BREAKDOWN: (p, Ap) = 0).  Every solve returns."""
    p, s, g = pair(ctx, L.HEX27, 4)
    kd = L.DeviceArray(ctx, p.n_elem, data=np.ones(p.n_elem))
    b = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(3, kd, 1, b)
    g.make_fp32()
    x = L.DeviceArray(ctx, p.n_dofs)
    x.zero()
    its, res, reason = g.cg(b, x, 1e-14, 3)
    assert its == 3 and reason == L.DIVERGED_ITS and res > 1e-14
    x.zero()
    its, res, reason, st = g.mixed(b, x, 1e-14, 0.1, 0, 3)
    assert reason == L.DIVERGED_ITS and res > 1e-14
    for bad in (np.nan, np.inf, -np.inf):
        # an Inf makes |b|^2 and the threshold rtol^2 |b|^2 infinite: rr <= thr must not pass
        bh = b.download()
        bh[p.n_dofs // 2] = bad
        bn = L.DeviceArray(ctx, p.n_dofs, data=bh)
        for name, solve in (("cg", lambda: g.cg(bn, x, 1e-10, 100)), ("bicgstab", lambda: g.bicgstab(bn, x, 1e-10, 100)),
                            ("mixed", lambda: g.mixed(bn, x, 1e-10, 0.1, 0, 100)[:3]),
                            ("mixed-classic", lambda: g.mixed(bn, x, 1e-10, 1e-5, 1, 100)[:3])):
            x.zero()
            its, res, reason = solve()
            assert reason == L.DIVERGED_NANORINF and its == 0, (bad, name, its, reason)
    g.set_values(np.zeros(g.nnz))
    x.zero()
    its, res, reason = g.cg(b, x, 1e-10, 100)
    assert reason == L.DIVERGED_BREAKDOWN
