"""Pins of the CPU oracle (oracle/lmoracle.cpp) before it is trusted as the checker.

Each oracle function is pinned by something that a plausible mistake would fail:
  mesh/numbering  : libMesh's node/element counts, TET4 node formula
                    (mesh_generation.C:1233-1236), first element's DoFs, boundary
                    set sizes, perturbation keeps boundary nodes (tests/mesh/distort.C:58-92)
  sparsity        : nnz closed forms (3n+1)^3 / (8n+1)^3, per-row counts as products
                    of 1D counts, symmetry, brute-force set union
  quadrature      : polynomial exactness (tests/quadrature/quadrature_test.C:150-421),
                    weight sums 8 and 1/6
  shapes          : partition of unity, nodal Kronecker property, zero derivative sum,
                    linear/quadratic reproduction (tests/fe/fe_test.h:69-125)
  element matrix  : Q1 unit-cube closed form h/12*[4, 0, -1, -1], symmetry, zero row sums,
                    affine-branch vs general-branch agreement on a parallelepiped
  constraints     : constrained diagonal = number of incident elements, zero rows/rhs
  CG              : the reference's own vendored Eigen 3.3.9 (oracle/_ref), Laspack
                    restatement vs Eigen restatement, O(h^2) manufactured convergence
  adjoint         : sum_e kappa_e dJ/dkappa_e = -J (J is homogeneous of degree -1 in kappa),
                    restated FD sensitivity vs closed form
  QoI             : J = int_box u of polynomial interpolants in closed form (1, x, xy, xyz),
                    sum dJ/du = |box|, zero derivative off the box's support, a shifted box
                    fails, an unaligned box against an independent affine Gauss sum
No fixture of the reference pins assembled matrices, patterns or CG results
(tests/numerics and tests/systems hold none); those are pinned by the checks above.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n", [1, 2, 5])
def test_mesh_counts(n):
    assert O.Mesh(n, O.HEX8).n_nodes == (n + 1) ** 3
    assert O.Mesh(n, O.HEX27).n_nodes == (2 * n + 1) ** 3
    t4 = O.Mesh(n, O.TET4)
    # mesh_generation.C:1233-1236
    assert t4.n_nodes == 5 * n ** 3 + 2 * 3 * n * n + 3 * n + 1
    assert t4.n_elem == 24 * n ** 3
    t10 = O.Mesh(n, O.TET10)
    # edges of the 24-tet split: Euler for the closed ball V - E + F - T = 1
    T = 24 * n ** 3
    F = (4 * T + 4 * 6 * n * n) // 2  # interior faces shared by 2, boundary 4 per hex face
    E = t4.n_nodes + F - T - 1
    assert t10.n_nodes == t4.n_nodes + E


def test_first_element_numbering():
    for t, npe in ((O.HEX8, 8), (O.HEX27, 27), (O.TET10, 10)):
        m = O.Mesh(3, t)
        assert list(m.conn()[0]) == list(range(npe))  # first touch
        s = O.System(m, 3)
        d = s.elem_dofs()[0]
        assert list(d[:npe]) == [3 * a for a in range(npe)]  # var-major, dof = 3 node + var
        assert list(d[npe:2 * npe]) == [3 * a + 1 for a in range(npe)]


@pytest.mark.parametrize("n", [2, 4])
def test_boundary_set(n):
    s = O.System(O.Mesh(n, O.HEX27))
    assert s.constrained().sum() == (2 * n + 1) ** 3 - (2 * n - 1) ** 3
    s1 = O.System(O.Mesh(n, O.HEX8))
    assert s1.constrained().sum() == (n + 1) ** 3 - (n - 1) ** 3
    # clamped x = 0 face (boundary id 4) for elasticity: 3 dofs per face node
    s3 = O.System(O.Mesh(n, O.HEX27), 3, 1 << 4)
    assert s3.constrained().sum() == 3 * (2 * n + 1) ** 2


def test_perturbation_moves_only_interior():
    m = O.Mesh(4, O.HEX27)
    X0 = m.coords()
    b = m.node_bmask()
    m.perturb(7, 0.2)
    X1 = m.coords()
    assert np.array_equal(X0[b != 0], X1[b != 0])
    d = np.abs(X1 - X0)[b == 0]
    assert (d > 0).all() and d.max() <= 0.2 / 8 + 1e-15
    assert m.min_detj(3) > 0


@pytest.mark.parametrize("t,n,q", [(O.HEX8, 4, 3), (O.HEX27, 3, 8), (O.HEX27, 6, 8)])
def test_sparsity_closed_form(t, n, q):
    s = O.System(O.Mesh(n, t))
    assert s.nnz == (q * n + 1) ** 3
    rp, ci = s.pattern()
    for r in range(s.n_dofs):
        row = ci[rp[r]:rp[r + 1]]
        assert (np.diff(row) > 0).all()
    # symmetric pattern
    rows = np.repeat(np.arange(s.n_dofs), np.diff(rp))
    a = set(zip(rows.tolist(), ci.tolist()))
    assert all((c, r) in a for r, c in a)


def test_sparsity_row_counts_q2():
    """Q2 row counts factor per direction: vertex 5 (interior) / 3 (boundary), mid 3."""
    n = 3
    m = O.Mesh(n, O.HEX27)
    s = O.System(m)
    rp, _ = s.pattern()
    X = m.coords()
    g = np.rint(X * 2 * n).astype(int)

    def c1(i):
        if i % 2 == 1:
            return 3
        return 3 if i in (0, 2 * n) else 5

    expect = np.array([c1(a) * c1(b) * c1(c) for a, b, c in g])
    assert np.array_equal(np.diff(rp), expect)


def test_sparsity_brute_force_union():
    for t in (O.HEX27, O.TET10):
        m = O.Mesh(2, t)
        s = O.System(m, 2)
        ed = s.elem_dofs()
        rows = [set() for _ in range(s.n_dofs)]
        for e in ed:
            for d in e:
                rows[d].update(e.tolist())
        rp, ci = s.pattern()
        for r in range(s.n_dofs):
            assert list(ci[rp[r]:rp[r + 1]]) == sorted(rows[r])


@pytest.mark.parametrize("npts", [2, 3])
def test_qgauss_hex_exactness(npts):
    p, w = O.qgauss_hex(npts)
    assert abs(w.sum() - 8.0) < 1e-14
    deg = 2 * npts - 1
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        exact = np.prod([(1 - (-1) ** (k + 1)) / (k + 1) for k in (a, b, c)])
        got = (w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c).sum()
        assert abs(got - exact) < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_qgauss_tet_exactness(order):
    p, w = O.qgauss_tet(order)
    assert abs(w.sum() - 1.0 / 6) < 1e-15
    for a, b, c in itertools.product(range(order + 1), repeat=3):
        if a + b + c > order:
            continue
        exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
        got = (w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c).sum()
        assert abs(got - exact) < 1e-14


REF_NODES = {
    O.HEX8: np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1],
                      [-1, 1, 1]], float),
    O.TET4: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
}
# HEX27 nodes from the libMesh index tables: index 0 -> -1, 1 -> +1, 2 -> 0
_i27 = [[0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 0, 2, 2, 1, 2, 0, 2, 2],
        [0, 0, 1, 1, 0, 0, 1, 1, 0, 2, 1, 2, 0, 0, 1, 1, 0, 2, 1, 2, 2, 0, 2, 1, 2, 2, 2],
        [0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 2, 2, 2, 2, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1, 2]]
REF_NODES[O.HEX27] = np.array([[{0: -1.0, 1: 1.0, 2: 0.0}[_i27[c][a]] for c in range(3)] for a in range(27)])
_t4 = REF_NODES[O.TET4]
_edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
REF_NODES[O.TET10] = np.vstack([_t4, [(_t4[a] + _t4[b]) / 2 for a, b in _edges]])


@pytest.mark.parametrize("t", [O.HEX8, O.HEX27, O.TET4, O.TET10])
def test_shapes(t):
    nodes = REF_NODES[t]
    phi, dphi = O.shapes(t, nodes)
    assert np.allclose(phi, np.eye(len(nodes)), atol=1e-14)  # Kronecker
    rng = np.random.default_rng(0)
    pts = rng.uniform(0, 0.3, (20, 3)) if t in (O.TET4, O.TET10) else rng.uniform(-1, 1, (20, 3))
    phi, dphi = O.shapes(t, pts)
    assert np.allclose(phi.sum(0), 1.0, atol=1e-14)  # partition of unity
    assert np.allclose(dphi.sum(0), 0.0, atol=1e-13)
    # linear reproduction: sum_a phi_a X_a = x, and its gradient is the identity
    assert np.allclose(phi.T @ nodes, pts, atol=1e-14)
    for c in range(3):
        assert np.allclose(np.einsum("aq,aqr->qr", nodes[:, [c]].repeat(len(pts), 1), dphi), np.eye(3)[c], atol=1e-13)
    if t in (O.HEX27, O.TET10):  # quadratic reproduction
        f = nodes[:, 0] * nodes[:, 1] + nodes[:, 2] ** 2
        assert np.allclose(phi.T @ f, pts[:, 0] * pts[:, 1] + pts[:, 2] ** 2, atol=1e-13)


def test_q1_unit_cube_element_matrix():
    """Trilinear stiffness of an h-cube: h/12 * (4 diag, 0 edge, -1 face diagonal, -1 body diagonal)."""
    n = 2
    s = O.System(O.Mesh(n, O.HEX8), dirichlet_mask=0)
    K, F = s.element_poisson(2, None, 0, 0, 1, constrain=False)
    K = K[0]
    h = 1.0 / n
    X = REF_NODES[O.HEX8]
    for a in range(8):
        for b in range(8):
            d = int(np.sum(X[a] != X[b]))
            expect = {0: 4, 1: 0, 2: -1, 3: -1}[d] * h / 12
            assert abs(K[a, b] - expect) < 1e-15
    assert abs(F.sum() - h ** 3) < 1e-15


@pytest.mark.parametrize("t,npts", [(O.HEX27, 3), (O.TET10, 5), (O.TET4, 3)])
def test_element_matrix_properties(t, npts):
    m = O.Mesh(2, t)
    if t == O.HEX27:
        m.perturb(3, 0.2)
    kap = O.kappa_lognormal(m.n_elem, 9)
    s = O.System(m, dirichlet_mask=0)
    K, F = s.element_poisson(npts, kap, 0, constrain=False)
    assert np.abs(K - K.transpose(0, 2, 1)).max() < 1e-15 * np.abs(K).max()
    assert np.abs(K.sum(2)).max() < 1e-13 * np.abs(K).max()  # constants are in the kernel
    # f = 1: sum_i Fe_i = element volume
    vol = F.sum(1)
    assert abs(vol.sum() - 1.0) < 1e-12


def test_affine_branch_agrees_with_general_branch():
    """On a sheared (affine) HEX27 element the affine shortcut (fe_map.C:1242) and the
    per-qp map must give the same matrix up to rounding."""
    m = O.Mesh(1, O.HEX27)
    X = m.coords()
    Y = X.copy()
    Y[:, 0] += 0.3 * X[:, 1] + 0.1 * X[:, 2]
    m.set_coords(Y)
    s = O.System(m, dirichlet_mask=0)
    Ka, _ = s.element_poisson(3, None, 0, constrain=False)
    Y2 = Y.copy()
    Y2[26, 0] += 1e-9  # breaks the affine test, stays practically the same map
    m.set_coords(Y2)
    Kg, _ = s.element_poisson(3, None, 0, constrain=False)
    assert np.abs(Ka - Kg).max() < 1e-7 * np.abs(Ka).max()


def test_constrained_rows():
    n = 3
    m = O.Mesh(n, O.HEX8)
    s = O.System(m)
    v, b = s.assemble_poisson(2, None, 1)
    rp, ci = s.pattern()
    cons = s.constrained()
    conn = m.conn()
    inc = np.bincount(conn.ravel(), minlength=s.n_dofs)
    for r in np.flatnonzero(cons):
        row = v[rp[r]:rp[r + 1]]
        cols = ci[rp[r]:rp[r + 1]]
        assert row[cols == r][0] == inc[r]  # corner 1, edge 2, face 4
        assert np.all(row[cols != r] == 0.0)
        assert b[r] == 0.0
    # constrained columns of free rows are zero too
    free = np.flatnonzero(cons == 0)
    for r in free[:50]:
        cols = ci[rp[r]:rp[r + 1]]
        assert np.all(v[rp[r]:rp[r + 1]][cons[cols] != 0] == 0.0)


def test_manufactured_convergence_q1():
    errs = []
    for n in (4, 8, 16):
        m = O.Mesh(n, O.HEX8)
        s = O.System(m)
        v, b = s.assemble_poisson(2, None, 1)
        rp, ci = s.pattern()
        x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-12, 5000)
        X = m.coords()
        ue = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1]) * np.sin(np.pi * X[:, 2])
        errs.append(np.abs(x - ue).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.8 < r2 < 4.2, errs


def test_manufactured_convergence_q2():
    errs = []
    for n in (2, 4, 8):
        m = O.Mesh(n, O.HEX27)
        s = O.System(m)
        v, b = s.assemble_poisson(3, None, 1)
        rp, ci = s.pattern()
        x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-12, 5000)
        X = m.coords()
        ue = np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1]) * np.sin(np.pi * X[:, 2])
        errs.append(np.abs(x - ue).max())
    assert errs[1] / errs[2] > 7.0, errs  # nodal error at least O(h^3)


def _c1_system(n):
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m)
    v, b = s.assemble_poisson(3, kappa, 0)
    rp, ci = s.pattern()
    return m, s, kappa, v, b, rp, ci


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("precond", ["jacobi", "identity"])
def test_cg_matches_reference_eigen(precond):
    """The restated CG against the reference's own vendored Eigen 3.3.9
    ConjugateGradient with DiagonalPreconditioner (JACOBI_PRECOND) or
    IdentityPreconditioner (IDENTITY_PRECOND) on the same CSR."""
    m, s, kappa, v, b, rp, ci = _c1_system(6)
    for tol in (1e-6, 1e-10, 1e-12):
        x, its, err = O.cg_eigen(rp, ci, v, b, tol, 5000, precond=precond)
        xr, itr, errr = O.ref_cg_eigen(rp, ci, v, b, tol, 5000, precond=precond)
        assert abs(its - itr) <= 1
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 50 * tol
        assert err < tol and errr < tol


def test_laspack_and_eigen_restatements_agree():
    m, s, kappa, v, b, rp, ci = _c1_system(4)
    x, its, err = O.cg_eigen(rp, ci, v, b, 1e-12, 5000)
    xl, itl, errl = O.cg_laspack(rp, ci, v, b, 1e-12, 5000)
    assert err < 1e-12 and errl < 1e-12
    assert abs(its - itl) <= 2
    assert np.linalg.norm(x - xl) / np.linalg.norm(x) < 1e-9


def test_adjoint_homogeneity_and_fd():
    m, s, kappa, v, b, rp, ci = _c1_system(3)
    u, _, _ = O.cg_eigen(rp, ci, v, b, 1e-15, 20000)
    dq = s.qoi_box_derivative(3)
    z, _, _ = O.cg_eigen(rp, ci, v, dq, 1e-15, 20000)
    J = s.qoi_box(3, u)
    assert abs(J - dq @ u) < 1e-14 * abs(J)
    g = s.adjoint_exact(3, kappa, u, z)
    assert abs((kappa * g).sum() + J) < 1e-10 * abs(J)
    elems = np.arange(0, m.n_elem, 3)
    fd = s.adjoint_sensitivity(3, kappa, u, z, elems)
    assert np.abs(fd - g[elems]).max() < 1e-6 * np.abs(g).max()
    # central FD of J itself on two elements (re-assemble + re-solve)
    for e in (0, m.n_elem // 2):
        d = 1e-6 * kappa[e]
        Js = []
        for sgn in (1, -1):
            kk = kappa.copy()
            kk[e] += sgn * d
            vv, bb = s.assemble_poisson(3, kk, 0)
            uu, _, _ = O.cg_eigen(rp, ci, vv, bb, 1e-15, 20000)
            Js.append(s.qoi_box(3, uu))
        assert abs((Js[0] - Js[1]) / (2 * d) - g[e]) < 1e-6 * abs(g).max()


@pytest.mark.parametrize("t,npts", [(O.HEX8, 2), (O.HEX27, 3), (O.TET4, 3), (O.TET10, 5)])
def test_elasticity_element_rigid_modes_and_trace(t, npts):
    """Linear-elasticity element matrices (systems_of_equations_ex6.C:127-150):
    the six rigid-body modes (3 translations, 3 infinitesimal rotations) are in
    the kernel of the unconstrained Ke on perturbed elements (isoparametric
    Lagrange spaces reproduce linear fields), the sum of the three diagonal
    component blocks is (lambda + 4 mu) times the Poisson element matrix, and
    the body-force vector integrates fz over the unit cube."""
    m = O.Mesh(2, t)
    m.perturb(5, 0.2 if t in (O.HEX8, O.HEX27) else 0.1)  # 0.2 h inverts some TET10 sub-tets
    E, nu, fz = 1.0, 0.3, -1.0
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    s3 = O.System(m, n_vars=3, dirichlet_mask=0)
    s1 = O.System(m, n_vars=1, dirichlet_mask=0)
    K, F = s3.element_elasticity(npts, E, nu, fz, constrain=False)
    Kp, _ = s1.element_poisson(npts, None, 0, constrain=False)
    xyz, conn = m.coords(), m.conn()
    n = conn.shape[1]
    for e in range(m.n_elem):
        X = xyz[conn[e]]
        modes = []
        for c in range(3):
            v = np.zeros(3 * n)
            v[c * n:(c + 1) * n] = 1.0
            modes.append(v)
        for c in range(3):
            w = np.zeros(3)
            w[c] = 1.0
            u = np.cross(w, X)
            modes.append(np.concatenate([u[:, 0], u[:, 1], u[:, 2]]))
        scale = np.abs(K[e]).max()
        for v in modes:
            assert np.abs(K[e] @ v).max() < 1e-12 * scale * np.abs(v).max()
        assert np.abs(K[e] - K[e].T).max() < 1e-14 * scale
        diag = sum(K[e][i * n:(i + 1) * n, i * n:(i + 1) * n] for i in range(3))
        assert np.abs(diag - (lam + 4 * mu) * Kp[e]).max() < 1e-13 * scale
    assert np.abs(F[:, :2 * n]).max() == 0.0
    assert abs(F[:, 2 * n:].sum() - fz) < 1e-12


def test_elasticity_constrained_rows():
    """Clamped x = 0 face (boundary id 4): every DoF of every node on that face is
    constrained; assembled rows of constrained DoFs hold only the diagonal, which
    counts the incident elements, and their rhs entries are zero."""
    m = O.Mesh(3, O.HEX8)
    s = O.System(m, n_vars=3, dirichlet_mask=1 << 4)
    vals, rhs = s.assemble_elasticity(2)
    rp, ci = s.pattern()
    cons = s.constrained().astype(bool)
    assert cons.sum() == 3 * 16
    xyz = m.coords()
    face = np.abs(xyz[:, 0]) < 1e-14
    node_of = np.arange(s.n_dofs) // 3
    assert np.array_equal(cons, face[node_of])
    ed = s.elem_dofs()
    for r in np.nonzero(cons)[0]:
        row = slice(rp[r], rp[r + 1])
        d = vals[row][ci[row] == r][0]
        assert d == float((ed == r).any(axis=1).sum())
        assert np.count_nonzero(vals[row]) == 1
        assert rhs[r] == 0.0


def _nodal(m, s, fn):
    """DoF vector of the nodal interpolant of fn(x, y, z) for one scalar variable."""
    X = m.coords()
    u = np.zeros(s.n_dofs)
    u[s.elem_dofs().ravel()] = fn(*X[m.conn().ravel()].T)
    return u


# J = int_[0.25,0.75]^3 u with the per-qp physical-coordinate inclusion of
# examples/adjoints/adjoints_ex2/L-qoi.C:29-95 (element_qoi / element_qoi_derivative).
# With the box faces on element faces every qp of an element lies inside or
# outside together, so J is the exact integral of the interpolant over the box.
_QOI_CASES = [
    (lambda x, y, z: np.ones_like(x), 0.125),
    (lambda x, y, z: x, 0.0625),          # int x over [.25,.75] = 0.25, times 0.5^2
    (lambda x, y, z: x * y, 0.03125),     # 0.25^2 * 0.5
    (lambda x, y, z: x * y * z, 0.015625),  # 0.25^3 (trilinear: HEX only)
]


@pytest.mark.parametrize("t,n,npts", [(O.HEX27, 4, 3), (O.HEX27, 8, 3), (O.HEX8, 4, 2), (O.TET10, 4, 5)])
def test_qoi_box_closed_forms(t, n, npts):
    m = O.Mesh(n, t)
    s = O.System(m, 1, dirichlet_mask=0)
    cases = _QOI_CASES if t != O.TET10 else _QOI_CASES[:3]
    for fn, exact in cases:
        u = _nodal(m, s, fn)
        assert abs(s.qoi_box(npts, u) - exact) < 1e-14, (fn, exact)
    # dJ/du without constraints: sum_i dq_i = int_box sum_i phi_i = |box| (partition of unity),
    # and dq . u = J for every u (J is linear in u)
    dq = s.qoi_box_derivative(npts)
    assert abs(dq.sum() - 0.125) < 1e-14
    u = _nodal(m, s, lambda x, y, z: x * y)
    assert abs(dq @ u - 0.03125) < 1e-14
    # every DoF whose support misses the box gets exactly 0
    X = m.coords()
    Xd = np.zeros((s.n_dofs, 3))
    Xd[s.elem_dofs().ravel()] = X[m.conn().ravel()]
    far = ((Xd <= 0.25 - 1.0 / n + 1e-12) | (Xd >= 0.75 + 1.0 / n - 1e-12)).any(axis=1)
    assert far.any() and np.all(dq[far] == 0.0)


def test_qoi_box_planted_shift_fails():
    """A box shifted by one element keeps |box|, so u = 1 cannot see it; u = x does."""
    n = 4
    m = O.Mesh(n, O.HEX27)
    s = O.System(m, 1, dirichlet_mask=0)
    u1 = _nodal(m, s, lambda x, y, z: np.ones_like(x))
    ux = _nodal(m, s, lambda x, y, z: x)
    h = 1.0 / n
    assert abs(s.qoi_box(3, u1, 0.25 + h, 0.75 + h) - 0.125) < 1e-14
    assert abs(s.qoi_box(3, ux, 0.25 + h, 0.75 + h) - 0.0625) > 1e-3


def test_qoi_box_unaligned_against_affine_quadrature():
    """Box faces inside elements: J is the Gauss sum over the qps whose physical
    coordinates fall in the box.  On the unperturbed mesh every element is an
    axis-aligned cube of side 1/n, so the physical qps are x0 + (1 + xi) / (2n) and
    JxW = w / (2n)^3; this is summed here independently of the oracle's FE map.
    An inclusion test in reference coordinates, or no test at all, fails it."""
    n = 3
    m = O.Mesh(n, O.HEX27)
    s = O.System(m, 1, dirichlet_mask=0)
    pts, w = O.qgauss_1d(3)
    lo, hi = 0.25, 0.75
    h = 1.0 / n
    # 1D: which qps of each element lie in [lo, hi], with their weights and positions
    q1 = [(e * h + (1 + p) * h / 2, wi * h / 2) for e in range(n) for p, wi in zip(pts, w)]
    ins = [(x, wx) for x, wx in q1 if lo <= x <= hi]
    vol = sum(wx for _, wx in ins) ** 3
    mx = sum(x * wx for x, wx in ins) * sum(wx for _, wx in ins) ** 2
    for fn, exact in ((lambda x, y, z: np.ones_like(x), vol), (lambda x, y, z: x, mx)):
        J = s.qoi_box(3, _nodal(m, s, fn), lo, hi)
        assert abs(J - exact) < 1e-14, (J, exact)
    # differs from the true box integral (0.125): the per-qp rule, not the exact cut
    assert abs(vol - 0.125) > 1e-3
