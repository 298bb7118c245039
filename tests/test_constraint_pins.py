"""Pins of the oracle's general constraint rows (oracle/lmoracle.cpp:
add_constraint_row, check_for_constraint_loops, process_constraints,
find_connected_dofs, build_constraint_matrix_and_vector, the heterogeneous
element constraint, enforce_constraints_exactly) before it checks the GPU.

  processing   : hand-computed substitution of a chain, a row through a
                 Dirichlet DoF and a value that cancels (dof_map_constraints.C:4341-4497)
  loops        : the reference's own case tests/base/dof_map_test.C:31-44,173-193
                 (0 <- 1, 1 <- 0 with set_error_on_constraint_loop)
  dense path   : for empty rows the dense C^T K C equals the elementwise mask
  pattern      : brute-force union of expanded element cliques
  assembly     : rows of unconstrained DoFs equal the global product C^T K C and
                 C^T (F - K H) built with scipy from the unconstrained assembly;
                 constrained rows are count * (e_i - c_i) and count * value
  solution     : symmetric rows + CG + enforce_constraints_exactly equals the
                 direct solution of the asymmetric system and solves the
                 globally reduced system
  FEM residual : the FEMSystem residual equals the global unconstrained K u - F
                 (scipy) on free rows, 0 on constrained rows, and A u - b when u
                 vanishes on constrained DoFs
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from constraint_cases import X0, X1, Y0, Y1, Z0, Z1, case_rows


def _system(n, t, mask, perturb=True):
    m = O.Mesh(n, t)
    if perturb:
        m.perturb(11, 0.1 if t in (O.TET4, O.TET10) else 0.2)
    return m, O.System(m, 1, mask)


def _add(s, rows):
    for d, row, v in rows:
        s.add_constraint_row(d, row, v)


def test_process_constraints_closed_form():
    m, s = _system(2, O.HEX8, X0, perturb=False)
    cons = s.constrained()
    f = int(np.flatnonzero(cons)[0])  # a Dirichlet DoF (empty row, value 0)
    free = [d for d in range(s.n_dofs) if not cons[d]]
    a, b, c, d, g, h, k = free[:7]
    s.add_constraint_row(a, {b: 0.5, c: 0.25}, 0.1)
    s.add_constraint_row(b, {d: 2.0}, 1.0)
    s.add_constraint_row(g, {f: 3.0, h: 1.0}, 0.5)
    s.add_constraint_row(k, {b: 1.0}, -1.0)
    s.process_constraints(error_on_loop=True)
    r = s.constraint_rows()
    assert r[a] == ({c: 0.25, d: 1.0}, 0.6)
    assert r[b] == ({d: 2.0}, 1.0)
    assert r[g] == ({h: 1.0}, 0.5)
    assert r[k] == ({d: 2.0}, 0.0)  # -1 + 1*1 cancels: the value is dropped
    assert r[f] == ({}, 0.0)
    assert sum(1 for v in r.values() if not v[0] and v[1] == 0.0) == int(cons.sum())


def test_constraint_loop_detection():
    m, s = _system(1, O.HEX8, 0, perturb=False)
    s.add_constraint_row(0, {1: 1.0}, 0.0)
    with pytest.raises(RuntimeError, match="already constrained"):
        s.add_constraint_row(0, {2: 1.0}, 0.0, forbid_overwrite=True)
    s.add_constraint_row(1, {0: 1.0}, 0.0)
    with pytest.raises(RuntimeError, match="loop"):
        s.process_constraints(error_on_loop=True)


@pytest.mark.parametrize("t", [O.HEX8, O.HEX27, O.TET10])
def test_dense_constraint_equals_mask_for_dirichlet(t):
    m, s = _system(2, t, 0x3F)
    K, F = s.element_poisson(3 if t == O.HEX27 else 2, constrain=False)
    Km, Fm = s.element_poisson(3 if t == O.HEX27 else 2, constrain=True)
    ed = s.elem_dofs()
    for e in range(m.n_elem):
        d, Kc, Fc = s.constrain_element(e, K[e], F[e])
        assert np.array_equal(d, ed[e])
        assert np.array_equal(Kc, Km[e]) and np.array_equal(Fc, Fm[e])


def _expanded_pattern(s, rows):
    ed = s.elem_dofs()
    sets = [set() for _ in range(s.n_dofs)]
    for e in range(ed.shape[0]):
        X = set(int(x) for x in ed[e])
        while True:  # find_connected_dofs, brute force
            add = set()
            for x in X:
                if x in rows:
                    add |= set(rows[x][0].keys())
            if add <= X:
                break
            X |= add
        for r in X:
            sets[r] |= X
    rp = np.zeros(s.n_dofs + 1, dtype=np.int64)
    for r in range(s.n_dofs):
        rp[r + 1] = rp[r] + len(sets[r])
    ci = np.concatenate([sorted(x) for x in sets])
    return rp, ci


CASES = [
    ("periodic_x", O.HEX27, 3, Z0 | Y0 | Y1 | Z1),
    ("periodic_xy", O.HEX8, 4, Z0 | Z1),
    ("random", O.TET10, 2, 0x3F),
]


def _build(name, t, n, mask, asym=True):
    m, s = _system(n, t, mask)
    rows = case_rows(name, m.coords(), m.conn(), s.elem_dofs(), s.constrained(), n)
    _add(s, rows)
    s.set_asymmetric_constraint_rows(asym)
    s.process_constraints(error_on_loop=True)
    return m, s


@pytest.mark.parametrize("name,t,n,mask", CASES)
def test_constrained_pattern_brute_force(name, t, n, mask):
    m, s = _build(name, t, n, mask)
    rp, ci = s.pattern()
    brp, bci = _expanded_pattern(s, s.constraint_rows())
    assert np.array_equal(rp, brp) and np.array_equal(ci, bci)


def _global_reduction(m, s, kappa, npts, f_kind=0):
    """Unconstrained assembly on the same mesh and the global C, H."""
    free_sys = O.System(m, 1, 0)
    rpf, cif = free_sys.pattern()
    vf, Ff = free_sys.assemble_poisson(npts, kappa, f_kind)
    N = s.n_dofs
    K = sp.csr_matrix((vf, cif, rpf), shape=(N, N))
    rows = s.constraint_rows()
    I, J, V = [], [], []
    H = np.zeros(N)
    for i in range(N):
        if i in rows:
            for j, c in rows[i][0].items():
                I.append(i)
                J.append(j)
                V.append(c)
            H[i] = rows[i][1]
        else:
            I.append(i)
            J.append(i)
            V.append(1.0)
    Cg = sp.csr_matrix((V, (I, J)), shape=(N, N))
    Ar = (Cg.T @ K @ Cg).tocsr()
    br = Cg.T @ (Ff - K @ H)
    return Ar, br, rows


def _counts(s):
    cnt = np.zeros(s.n_dofs, dtype=np.int64)
    np.add.at(cnt, s.elem_dofs().ravel(), 1)
    return cnt


@pytest.mark.parametrize("name,t,n,mask", CASES)
@pytest.mark.parametrize("asym", [True, False])
def test_constrained_assembly_is_global_reduction(name, t, n, mask, asym):
    m, s = _build(name, t, n, mask, asym)
    npts = 3 if t == O.HEX27 else 2
    kappa = O.kappa_lognormal(m.n_elem, 3)
    vals, rhs = s.assemble_poisson(npts, kappa)
    rp, ci = s.pattern()
    A = sp.csr_matrix((vals, ci, rp), shape=(s.n_dofs,) * 2).toarray()
    Ar, br, rows = _global_reduction(m, s, kappa, npts)
    Ar = Ar.toarray()
    free = np.array([i for i in range(s.n_dofs) if i not in rows])
    scale = np.abs(Ar).max()
    assert np.abs(A[free] - Ar[free]).max() <= 1e-13 * scale
    assert np.abs(rhs[free] - br[free]).max() <= 1e-13 * np.abs(br).max()
    cnt = _counts(s)
    for i, (row, v) in rows.items():
        want = np.zeros(s.n_dofs)
        want[i] = cnt[i]
        if asym:
            for j, c in row.items():
                want[j] = -cnt[i] * c
        # cnt additions of the same entry: exact up to the rounding of the running sum
        tol = 4e-16 * cnt[i] * max(1.0, np.abs(want).max())
        assert np.array_equal(A[i] != 0, want != 0) and np.abs(A[i] - want).max() <= tol, i
        assert abs(rhs[i] - (cnt[i] * v if asym else 0.0)) <= 4e-16 * cnt[i] * abs(v), i


@pytest.mark.parametrize("name,t,n,mask", CASES)
def test_constrained_solution(name, t, n, mask):
    npts_of = {O.HEX27: 3}
    sols = {}
    for asym in (True, False):
        m, s = _build(name, t, n, mask, asym)
        npts = npts_of.get(t, 2)
        kappa = O.kappa_lognormal(m.n_elem, 3)
        vals, rhs = s.assemble_poisson(npts, kappa, f_kind=1)
        rp, ci = s.pattern()
        A = sp.csr_matrix((vals, ci, rp), shape=(s.n_dofs,) * 2).tocsc()
        if asym:
            sols[asym] = spla.spsolve(A, rhs)
        else:
            x, its, err = O.cg_eigen(rp, ci, vals, rhs, 1e-13, 5000)
            assert err < 1e-13
            sols[asym] = s.enforce_constraints(x)
            Ar, br, rows = _global_reduction(m, s, kappa, npts, f_kind=1)
            free = np.array([i for i in range(s.n_dofs) if i not in rows])
            y = spla.spsolve(Ar[free][:, free].tocsc(), br[free])
            assert np.abs(sols[asym][free] - y).max() <= 1e-9 * np.abs(y).max()
            for i, (row, v) in rows.items():
                assert abs(sols[asym][i] - (v + sum(c * sols[asym][j] for j, c in row.items()))) <= 1e-14
    ref = sols[True]
    assert np.abs(sols[False] - ref).max() <= 1e-9 * np.abs(ref).max()


def test_enforce_constraints_rows():
    m, s = _build("periodic_x", O.HEX27, 2, Z0 | Y0 | Y1 | Z1)
    x = np.random.default_rng(1).standard_normal(s.n_dofs)
    rows = s.constraint_rows()
    for hom in (False, True):
        y = s.enforce_constraints(x, homogeneous=hom)
        for i in range(s.n_dofs):
            if i in rows:
                want = (0.0 if hom else rows[i][1]) + sum(c * x[j] for j, c in rows[i][0].items())
                assert abs(y[i] - want) <= 1e-15 * max(1.0, abs(want))
            else:
                assert y[i] == x[i]


@pytest.mark.parametrize("t,n", [(O.HEX27, 3), (O.TET10, 2), (O.HEX8, 4)])
def test_fem_residual(t, n):
    """The FEMSystem residual restatement: for any u, R equals the unconstrained
    global K u - F on unconstrained rows and 0 on constrained ones; for u that
    vanishes on constrained DoFs it equals A u - b of the constrained assembly."""
    m, s = _system(n, t, 0x3F)
    npts = 3 if t == O.HEX27 else 2
    kappa = O.kappa_lognormal(m.n_elem, 3)
    u = np.random.default_rng(2).standard_normal(s.n_dofs)
    R = s.residual_poisson(npts, u, kappa, 1)
    full = O.System(m, 1, 0)
    rpf, cif = full.pattern()
    vf, Ff = full.assemble_poisson(npts, kappa, 1)
    Kf = sp.csr_matrix((vf, cif, rpf), shape=(s.n_dofs,) * 2)
    g = Kf @ u - Ff
    cons = s.constrained().astype(bool)
    scale = np.abs(Kf).max() * np.abs(u).max()
    assert np.abs(R[~cons] - g[~cons]).max() <= 1e-13 * scale
    assert not R[cons].any()
    u[cons] = 0.0
    R0 = s.residual_poisson(npts, u, kappa, 1)
    v, b = s.assemble_poisson(npts, kappa, 1)
    rp, ci = s.pattern()
    A = sp.csr_matrix((v, ci, rp), shape=(s.n_dofs,) * 2)
    assert np.abs(R0 - (A @ u - b)).max() <= 1e-13 * scale


@pytest.mark.parametrize("precond", ["jacobi", "identity"])
@pytest.mark.parametrize("name,t,n,mask", CASES)
def test_bicgstab_restatement(name, t, n, mask, precond):
    """The BiCGSTAB restatement against the reference's own vendored Eigen
    3.3.9 BiCGSTAB (oracle/_ref; DiagonalPreconditioner or
    IdentityPreconditioner) and a direct solve, on the asymmetric
    constraint-row systems that CG cannot solve."""
    m, s = _build(name, t, n, mask, asym=True)
    npts = 3 if t == O.HEX27 else 2
    kappa = O.kappa_lognormal(m.n_elem, 3)
    vals, rhs = s.assemble_poisson(npts, kappa, 1)
    rp, ci = s.pattern()
    x, its, err = O.bicgstab_eigen(rp, ci, vals, rhs, 1e-12, 5000, precond=precond)
    assert err <= 1e-12
    xd = spla.spsolve(sp.csr_matrix((vals, ci, rp), shape=(s.n_dofs,) * 2).tocsc(), rhs)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    ref = O.ref_bicgstab_eigen(rp, ci, vals, rhs, 1e-12, 5000, precond=precond)
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    xr, itr, errr = ref
    assert abs(its - itr) <= 2 and errr <= 1e-12
    assert np.abs(x - xr).max() <= 1e-9 * np.abs(xr).max()
