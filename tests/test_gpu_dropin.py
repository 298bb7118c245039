"""The libMesh-API drop-in classes (HipCsrMatrix, HipLinearSolver, HipAssembler)
driven from C++ exactly as a libMesh application would, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "libmesh_amd", "host", "bin")


def run(args, tmp_path, expect=0):
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=300, cwd=tmp_path)
    assert r.returncode == expect, r.stdout + r.stderr
    return r.stdout


def test_numeric_vector_contract(tmp_path):
    """tests/numerics/numeric_vector_test.h of the reference on HipVector: every
    check through the derived and the base class."""
    out = run(["numeric_vector_contract"], tmp_path)
    assert "FAIL" not in out and "0 failures" in out, out
    for name in ("Localize", "LocalizeToOne", "LocalizeIndices", "Norms", "Operations", "WriteAndRead"):
        assert f"HipVector {name}" in out and f"HipVectorBase {name}" in out, (name, out)


def test_sparse_matrix_contract(tmp_path):
    out = run(["sparse_matrix_contract", os.path.join(ROOT, "tests", "golden", "matrices")], tmp_path)
    assert "FAIL" not in out and "0 failures" in out, out
    assert "testReadMatlab geom_4" in out and "PETSc binary" in out, out
    for what in ("update_sparsity_pattern: element matrices sum", "vector_mult then vector_mult_add",
                 "print_petsc_binary/read_petsc_binary round trip", "zero_rows with a diagonal value"):
        assert what in out, (what, out)


def test_poisson_c0_dropin(tmp_path):
    out = run(["poisson_dropin", 16, 10, tmp_path / "c0", "sine"], tmp_path)
    u = np.fromfile(tmp_path / "c0_u.bin")
    m = O.Mesh(16, O.HEX8)
    s = O.System(m)
    v, b = s.assemble_poisson(2, None, 1)
    rp, ci = s.pattern()
    x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-10, 5000)
    assert np.abs(u - x).max() < 1e-12, out


@pytest.mark.parametrize("mixed", [False, True])
def test_poisson_c1_recipe_dropin_with_adjoint(tmp_path, mixed):
    n = 5
    flags = ["perturb", "kappa", "adjoint"] + (["mixed"] if mixed else [])
    out = run(["poisson_dropin", n, 12, tmp_path / "c1"] + flags, tmp_path)
    u = np.fromfile(tmp_path / "c1_u.bin")
    z = np.fromfile(tmp_path / "c1_z.bin")
    grad = np.fromfile(tmp_path / "c1_grad.bin")
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m)
    v, b = s.assemble_poisson(3, kappa, 0)
    rp, ci = s.pattern()
    true_res = np.linalg.norm(b - O.spmv(rp, ci, v, u)) / np.linalg.norm(b)
    assert true_res <= 1.05e-10, out
    ut, _, _ = O.cg_eigen(rp, ci, v, b, 1e-14, 20000)
    assert np.linalg.norm(u - ut) / np.linalg.norm(ut) < 1e-7
    dq = s.qoi_box_derivative(3)
    zt, _, _ = O.cg_eigen(rp, ci, v, dq, 1e-14, 20000)
    exact = s.adjoint_exact(3, kappa, ut, zt)
    assert np.abs(grad - exact).max() / np.abs(exact).max() < 1e-6
    elems = np.arange(0, m.n_elem, 11)
    fd = s.adjoint_sensitivity(3, kappa, ut, zt, elems)
    assert np.abs(grad[elems] - fd).max() / np.abs(fd).max() < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_static_condensation_dropin(tmp_path, mixed):
    """System::create_static_condensation before init (miscellaneous_ex16.C:
    155-201): the system matrix is a StaticCondensation, assembly fills the
    full operator, the solve runs on the reduced system (the HEX27 centre DoFs
    eliminated) and back-substitutes.  The solution is the oracle's full solve."""
    n = 5
    flags = ["perturb", "kappa", "condense"] + (["mixed"] if mixed else [])
    out = run(["poisson_dropin", n, 12, tmp_path / "sc"] + flags, tmp_path)
    u = np.fromfile(tmp_path / "sc_u.bin")
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m)
    v, b = s.assemble_poisson(3, kappa, 0)
    rp, ci = s.pattern()
    ut, its_full, _ = O.cg_eigen(rp, ci, v, b, 1e-14, 20000)
    assert f"reduced problem size to {s.n_dofs - m.n_elem}" in out, out
    assert np.linalg.norm(u - ut) / np.linalg.norm(ut) < 1e-8, out
    its = int(out.split("iterations ")[1].split()[0])
    assert 0 < its < 20000, out


@pytest.mark.parametrize("bicgstab", [False, True])
def test_poisson_periodic_constraints_dropin(tmp_path, bicgstab):
    """System::Constraint rows (u(x=1) = u(x=0) plus a heterogeneous row):
    symmetric constraint rows, CG and enforce_constraints_exactly; or the
    default asymmetric rows solved by BiCGSTAB.  Both give the oracle's
    constrained solution."""
    from constraint_cases import Y0, Y1, Z0, Z1, periodic_rows, value_row_on

    n = 4
    flags = ["perturb", "kappa", "sine", "periodic"] + (["bicgstab"] if bicgstab else [])
    out = run(["poisson_dropin", n, 12, tmp_path / "per"] + flags, tmp_path)
    u = np.fromfile(tmp_path / "per_u.bin")
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m, 1, Z0 | Y0 | Y1 | Z1)
    taken = set()
    args = (m.coords(), m.conn(), s.elem_dofs(), s.constrained())
    rows = value_row_on(*args, taken, 0, 0.3) + periodic_rows(*args, n, (0,), taken)
    for d, row, v in rows:
        s.add_constraint_row(d, row, v)
    s.set_asymmetric_constraint_rows(False)
    s.process_constraints(error_on_loop=True)
    v, b = s.assemble_poisson(3, kappa, 1)
    rp, ci = s.pattern()
    x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-10, 20000)
    x = s.enforce_constraints(x)
    assert np.abs(u - x).max() <= 1e-8 * np.abs(x).max(), out
    d0 = rows[0][0]
    if bicgstab:
        assert abs(u[d0] - 0.3) <= 1e-8, out
    else:
        assert u[d0] == 0.3, out
        gits = int(out.split("iterations ")[1].split()[0])
        assert abs(gits - its) <= 2, out


def test_degenerate_map_raises(tmp_path):
    """An inverted element raises DegenerateMap (fe_map.C:467-505) through the C ABI."""
    out = run(["poisson_dropin", 2, 12, tmp_path / "bad", "distort"], tmp_path, expect=3)


def test_fem_system_newton_dropin(tmp_path):
    """FEMSystem + SteadySolver + NewtonSolver with the device physics: one
    Newton step solves the linear problem; the solution matches the oracle."""
    n = 4
    out = run(["fem_system_dropin", n, 12, tmp_path / "fem", "perturb", "kappa", "sine", "adjoint"], tmp_path)
    u = np.fromfile(tmp_path / "fem_u.bin")
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m)
    v, b = s.assemble_poisson(3, kappa, 1)
    rp, ci = s.pattern()
    x, _, _ = O.cg_eigen(rp, ci, v, b, 1e-14, 20000)
    assert np.abs(u - x).max() <= 1e-9 * np.abs(x).max(), out
    toks = out.split()
    newton, result = int(toks[toks.index("newton") + 1]), int(toks[toks.index("result") + 1])
    assert newton == 1 and result & 4, out  # CONVERGED_RELATIVE_RESIDUAL
    R = s.residual_poisson(3, u, kappa, 1)
    assert np.linalg.norm(R) <= 1e-9 * np.linalg.norm(b), out
    # adjoint_qoi_parameter_sensitivity through the FEMSystem's Jacobian
    grad = np.fromfile(tmp_path / "fem_grad.bin")
    dq = s.qoi_box_derivative(3)
    zt, _, _ = O.cg_eigen(rp, ci, v, dq, 1e-14, 20000)
    exact = s.adjoint_exact(3, kappa, x, zt)
    assert np.abs(grad - exact).max() / np.abs(exact).max() < 1e-6, out


def test_equation_systems_write_read_dropin(tmp_path):
    """EquationSystems::write / read from C++ (es_io_dropin): two systems on a
    perturbed 3^3 HEX27 cube, one with an additional vector (WRITE_ / READ_ADDITIONAL_DATA).  The ASCII file's values, read here independently,
    are the coordinate polynomials the example set, in node-major order (nodes
    by id, each node's variables); the XDR file decodes to the same bits; both
    read back into fresh systems exactly; a header that does not describe the
    systems is rejected."""
    import libmesh_amd as L
    from es_files import parse_xda, parse_xdr

    n = 3
    out = run(["es_io_dropin", n, tmp_path / "es"], tmp_path)
    assert "xdr_maxdiff 0.000e+00 xda_maxdiff 0.000e+00 mismatch_rejected 1" in out, out
    p = L.Problem(n, L.HEX27)
    p.perturb(0x5EED0001, 0.2)
    x, y, z = np.asarray(p.coords, dtype=np.float64).reshape(-1, 3).T
    a = parse_xda(tmp_path / "es.xda")
    assert a["version"] == "libMesh-1.7.0"
    assert [(s["name"], s["number"], s["type"]) for s in a["systems"]] == [("Poisson", 0, "LinearImplicit"),
                                                                           ("aux", 1, "Basic")]
    assert a["systems"][0]["vars"] == [("u", [], 2, 0)] and a["systems"][0]["vectors"] == [("u_old", 0, 2)]
    assert a["systems"][1]["vars"] == [("a", [], 2, 0), ("b", [], 2, 0)] and a["systems"][1]["vectors"] == []
    u = a["systems"][0]["data"][0]
    assert np.abs(a["systems"][0]["data"][1] - (2 * u + 1)).max() <= 1e-15 * 16
    ab = a["systems"][1]["data"][0]
    assert np.abs(u - (1 + x + 2 * y + 3 * z)).max() <= 1e-15 * 7
    assert np.abs(ab[0::2] - (x * y - z)).max() <= 1e-15 * 2
    assert np.abs(ab[1::2] - (x * x - y * z + 0.125)).max() <= 1e-15 * 2
    b = parse_xdr(tmp_path / "es.xdr")
    for sa, sb in zip(a["systems"], b["systems"]):
        assert sa["vars"] == sb["vars"] and all(np.array_equal(d, e) for d, e in zip(sa["data"], sb["data"]))


def test_poisson_identity_preconditioner_dropin(tmp_path):
    """LinearSolver::set_preconditioner_type(IDENTITY_PRECOND) from C++: the
    solution of the oracle's identity-preconditioned Eigen CG restatement."""
    out = run(["poisson_dropin", 12, 10, tmp_path / "id", "perturb", "kappa", "sine", "identity"], tmp_path)
    u = np.fromfile(tmp_path / "id_u.bin")
    m = O.Mesh(12, O.HEX8)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m)
    v, b = s.assemble_poisson(2, kappa, 1)
    rp, ci = s.pattern()
    x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-10, 5000, precond="identity")
    assert its > 20 and abs(int(out.split("iterations ")[1].split()[0]) - its) <= 2, (its, out)
    assert np.abs(u - x).max() <= 1e-9 * np.abs(x).max(), out


@pytest.mark.parametrize("bicgstab", [False, True])
def test_fem_system_periodic_constraints_dropin(tmp_path, bicgstab):
    """FEMSystem + NewtonSolver with System::Constraint rows (u(x=1) = u(x=0)
    plus a heterogeneous row): symmetric rows with CG, or the default
    asymmetric rows with BiCGSTAB.  The oracle's constrained solution."""
    from constraint_cases import Y0, Y1, Z0, Z1, periodic_rows, value_row_on

    n = 4
    flags = ["perturb", "kappa", "sine", "periodic", "rtol=1e-11"] + (["bicgstab"] if bicgstab else [])
    out = run(["fem_system_dropin", n, 12, tmp_path / "fp"] + flags, tmp_path)
    u = np.fromfile(tmp_path / "fp_u.bin")
    m = O.Mesh(n, O.HEX27)
    m.perturb(0x5EED0001, 0.2)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    s = O.System(m, 1, Z0 | Y0 | Y1 | Z1)
    taken = set()
    args = (m.coords(), m.conn(), s.elem_dofs(), s.constrained())
    rows = value_row_on(*args, taken, 0, 0.3) + periodic_rows(*args, n, (0,), taken)
    for d, row, v in rows:
        s.add_constraint_row(d, row, v)
    s.set_asymmetric_constraint_rows(False)
    s.process_constraints(error_on_loop=True)
    v, b = s.assemble_poisson(3, kappa, 1)
    rp, ci = s.pattern()
    x, its, _ = O.cg_eigen(rp, ci, v, b, 1e-12, 20000)
    x = s.enforce_constraints(x)
    assert np.abs(u - x).max() <= 1e-8 * np.abs(x).max(), out
    assert abs(u[rows[0][0]] - 0.3) <= 1e-12, out
