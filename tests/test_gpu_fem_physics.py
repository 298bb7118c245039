"""FEMSystem with a user physics on the device (libmesh_amd/host/examples/
fem_physics_dropin.hip): the reaction-diffusion -div(kappa grad u) + sigma u^3 = 1
written as a device element_time_derivative, assembled through
FEMSystem::assembly and solved by NewtonSolver, checked against the oracle's
FEMSystem restatement (orc_fem_reaction_diffusion) at 1e-12; and a
three-variable linear elasticity physics checked against the oracle's
elasticity assembly (systems_of_equations_ex6's operator)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "libmesh_amd", "host", "bin", "fem_physics_dropin")

# FEMContext's default rule: QGauss of order 2p+1 (+ extra_quadrature_order);
# the oracle takes points per direction for hexahedra, the order for tets
NPTS = {O.HEX8: 2, O.HEX27: 3, O.TET4: 3, O.TET10: 5}


def run(args, tmp_path):
    print("fem_physics_dropin", *args, flush=True)
    r = subprocess.run([BIN] + [str(a) for a in args], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def oracle_system(n, t, perturb):
    m = O.Mesh(n, t)
    if perturb:
        m.perturb(0x5EED0001, 0.2)
    return m, O.System(m)


CASES = [(4, O.HEX27, True), (5, O.HEX8, True), (3, O.TET10, False), (4, O.TET4, False)]


@pytest.mark.parametrize("n,t,perturb", CASES)
def test_residual_and_jacobian_match_oracle(tmp_path, n, t, perturb):
    m, s = oracle_system(n, t, perturb)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    u = np.random.default_rng(11).standard_normal(s.n_dofs) * 0.3
    u.tofile(tmp_path / "u_in.bin")
    flags = ["kappa", "sigma=50", f"u={tmp_path / 'u_in.bin'}"] + (["perturb"] if perturb else [])
    run([n, t, tmp_path / "rd", "eval"] + flags, tmp_path)
    R = np.fromfile(tmp_path / "rd_R.bin")
    K = np.fromfile(tmp_path / "rd_K.bin")
    Kr, Rr = s.fem_reaction_diffusion(NPTS[t], u, 50.0, kappa)
    assert R.shape == Rr.shape and K.shape == Kr.shape
    assert np.linalg.norm(R - Rr) <= 1e-12 * np.linalg.norm(Rr)
    assert np.max(np.abs(K - Kr)) <= 1e-12 * np.max(np.abs(Kr))
    # constrained rows: residual exactly 0, diagonal the integer count of incident elements
    cons = s.constrained().astype(bool)
    assert np.all(R[cons] == 0.0)
    rp, ci = s.pattern()
    A = sp.csr_matrix((K, ci, rp), shape=(s.n_dofs, s.n_dofs))
    d = A.diagonal()[cons]
    assert np.all(d == np.round(d)) and np.all(d >= 1)


def test_extra_quadrature_order(tmp_path):
    """System::extra_quadrature_order raises the FEMContext rule: HEX8 with +2 is
    QGauss order 5, 3 points per direction."""
    n, t = 4, O.HEX8
    m, s = oracle_system(n, t, True)
    u = np.random.default_rng(12).standard_normal(s.n_dofs) * 0.3
    u.tofile(tmp_path / "u_in.bin")
    run([n, t, tmp_path / "rq", "eval", "perturb", "sigma=20", f"u={tmp_path / 'u_in.bin'}", "extra_q=2"], tmp_path)
    R = np.fromfile(tmp_path / "rq_R.bin")
    _, Rr3 = s.fem_reaction_diffusion(3, u, 20.0, jacobian=False)
    _, Rr2 = s.fem_reaction_diffusion(2, u, 20.0, jacobian=False)
    assert np.linalg.norm(R - Rr3) <= 1e-12 * np.linalg.norm(Rr3)
    assert np.linalg.norm(R - Rr2) > 1e-8 * np.linalg.norm(Rr2)


def cpu_newton(s, npts, sigma, kappa, tol=1e-13, maxit=30):
    u = np.zeros(s.n_dofs)
    rp, ci = s.pattern()
    for _ in range(maxit):
        K, R = s.fem_reaction_diffusion(npts, u, sigma, kappa)
        if np.linalg.norm(R) < tol:
            break
        u -= spla.spsolve(sp.csr_matrix((K, ci, rp), shape=(s.n_dofs, s.n_dofs)).tocsc(), R)
    return u


@pytest.mark.parametrize("n,t,perturb", [(4, O.HEX27, True), (6, O.HEX8, True)])
def test_newton_dropin(tmp_path, n, t, perturb):
    m, s = oracle_system(n, t, perturb)
    kappa = O.kappa_lognormal(m.n_elem, 0x5EED0002)
    flags = ["kappa", "sigma=2000", "rtol=1e-12"] + (["perturb"] if perturb else [])
    out = run([n, t, tmp_path / "nt", "newton"] + flags, tmp_path)
    u = np.fromfile(tmp_path / "nt_u.bin")
    newton_steps = int(out.split("newton ")[1].split()[0])
    assert newton_steps >= 2, out  # the reaction term makes the problem nonlinear
    _, R0 = s.fem_reaction_diffusion(NPTS[t], np.zeros(s.n_dofs), 2000.0, kappa, jacobian=False)
    _, R = s.fem_reaction_diffusion(NPTS[t], u, 2000.0, kappa, jacobian=False)
    assert np.linalg.norm(R) <= 1e-10 * np.linalg.norm(R0), out
    ref = cpu_newton(s, NPTS[t], 2000.0, kappa)
    assert np.linalg.norm(u - ref) <= 1e-9 * np.linalg.norm(ref), out
    # the nonlinearity is visible: the sigma = 0 (Poisson) solution differs
    lin = cpu_newton(s, NPTS[t], 0.0, kappa)
    assert np.linalg.norm(lin - ref) > 1e-2 * np.linalg.norm(ref)


CLAMP = 1 << 4  # boundary id 4: the x = 0 face


@pytest.mark.parametrize("n,t,perturb", [(3, O.HEX8, True), (2, O.HEX27, True), (2, O.TET10, False)])
def test_elasticity_physics_matches_oracle(tmp_path, n, t, perturb):
    """Three variables (var-major element DoFs): the residual K u - F and the
    Jacobian K of the user elasticity physics against the oracle's assembly."""
    m = O.Mesh(n, t)
    if perturb:
        m.perturb(0x5EED0001, 0.2)
    s = O.System(m, 3, CLAMP)
    cons = s.constrained().astype(bool)
    u = np.random.default_rng(13).standard_normal(s.n_dofs) * 0.1
    u[cons] = 0.0
    u.tofile(tmp_path / "u_in.bin")
    run([n, t, tmp_path / "el", "eval", "elasticity", f"u={tmp_path / 'u_in.bin'}"] + (["perturb"] if perturb else []),
        tmp_path)
    R = np.fromfile(tmp_path / "el_R.bin")
    K = np.fromfile(tmp_path / "el_K.bin")
    Kv, F = s.assemble_elasticity(NPTS[t], 1.0, 0.3, -1.0)
    rp, ci = s.pattern()
    A = sp.csr_matrix((Kv, ci, rp), shape=(s.n_dofs, s.n_dofs))
    ref = A @ u - F
    ref[cons] = 0.0
    assert K.shape == Kv.shape
    assert np.max(np.abs(K - Kv)) <= 1e-12 * np.max(np.abs(Kv))
    assert np.linalg.norm(R - ref) <= 1e-12 * np.linalg.norm(ref)


def test_elasticity_physics_newton(tmp_path):
    """The linear elasticity physics through NewtonSolver: one Newton step to
    the oracle's direct solution."""
    n, t = 3, O.HEX27
    m = O.Mesh(n, t)
    m.perturb(0x5EED0001, 0.2)
    s = O.System(m, 3, CLAMP)
    out = run([n, t, tmp_path / "en", "newton", "elasticity", "perturb", "rtol=1e-12"], tmp_path)
    u = np.fromfile(tmp_path / "en_u.bin")
    Kv, F = s.assemble_elasticity(3, 1.0, 0.3, -1.0)
    rp, ci = s.pattern()
    ref = spla.spsolve(sp.csr_matrix((Kv, ci, rp), shape=(s.n_dofs, s.n_dofs)).tocsc(), F)
    assert np.linalg.norm(u - ref) <= 1e-9 * np.linalg.norm(ref), out
