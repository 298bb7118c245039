"""Pins of the oracle's user-physics assembly (orc_fem_reaction_diffusion):
FEMSystem::assembly of -div(kappa grad u) + sigma u^3 = f written as an
element_time_derivative (src/systems/fem_system.C:875-1010,
examples/adjoints/adjoints_ex1/L-shaped.C:58-108).

  sigma = 0      : the residual is K u - F of the separately pinned Poisson
                   assembly (constrained rows 0) and the Jacobian is its K
  finite diff.   : J v against central differences of R along v (R is cubic in
                   the step, so the difference error is known in closed form)
  closed forms   : with no Dirichlet rows and u = c, sum_i R_i = (sigma c^3 - f)
                   |cube| and sum_ij J_ij = 3 sigma c^2 |cube| (partition of
                   unity; the perturbation keeps the cube's volume)
  planted        : a Jacobian reaction factor 2 fails the FD and closed-form
                   pins; a residual reaction u^2 fails the closed form
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O


def _system(n, t, mask=O.ALL_FACES):
    m = O.Mesh(n, t)
    if t in (O.HEX8, O.HEX27):  # the +-0.2h perturbation is for hexahedra (it can fold tets)
        m.perturb(0x5EED0001, 0.2)
    return O.System(m, 1, mask)


def _csr(s, vals):
    rp, ci = s.pattern()
    return sp.csr_matrix((vals, ci, rp), shape=(s.n_dofs, s.n_dofs))


NPTS = {O.HEX8: 2, O.HEX27: 3, O.TET4: 2, O.TET10: 4}


@pytest.mark.parametrize("t", [O.HEX8, O.HEX27, O.TET10])
def test_sigma_zero_is_poisson(t):
    s = _system(3, t)
    rng = np.random.default_rng(5)
    kappa = np.exp(rng.standard_normal(s.mesh.n_elem))
    u = rng.standard_normal(s.n_dofs)
    u[s.constrained().astype(bool)] = 0.0
    vals, R = s.fem_reaction_diffusion(NPTS[t], u, 0.0, kappa)
    Kv, F = s.assemble_poisson(NPTS[t], kappa)
    K = _csr(s, Kv)
    ref = K @ u - F
    ref[s.constrained().astype(bool)] = 0.0
    assert np.linalg.norm(R - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.max(np.abs(vals - Kv)) <= 1e-14 * np.max(np.abs(Kv))


def _fd_error(s, npts, u, v, sigma, kappa, variant=0):
    vals, _ = s.fem_reaction_diffusion(npts, u, sigma, kappa, residual=False, variant=variant)
    Jv = _csr(s, vals) @ v
    h = 1e-3
    _, Rp = s.fem_reaction_diffusion(npts, u + h * v, sigma, kappa, jacobian=False, variant=variant)
    _, Rm = s.fem_reaction_diffusion(npts, u - h * v, sigma, kappa, jacobian=False, variant=variant)
    # R(u + t v) is a cubic in t: (R(h) - R(-h)) / 2h = J v + h^2 sigma M[v^3]
    _, Rv3 = s.fem_reaction_diffusion(npts, v, 1.0, None, jacobian=False)
    _, R0 = s.fem_reaction_diffusion(npts, 0 * v, 1.0, None, jacobian=False)
    _, Rk = s.fem_reaction_diffusion(npts, v, 0.0, None, jacobian=False)
    cubic = (Rv3 - R0) - (Rk - R0)  # sigma = 1, kappa = 1: the reaction term of v alone
    fd = (Rp - Rm) / (2 * h) - h * h * sigma * cubic
    return np.linalg.norm(fd - Jv) / np.linalg.norm(Jv)


@pytest.mark.parametrize("t", [O.HEX8, O.HEX27, O.TET4, O.TET10])
def test_jacobian_against_central_differences(t):
    s = _system(3, t)
    rng = np.random.default_rng(7)
    kappa = np.exp(rng.standard_normal(s.mesh.n_elem))
    u = rng.standard_normal(s.n_dofs)
    v = rng.standard_normal(s.n_dofs)
    cons = s.constrained().astype(bool)
    u[cons] = 0.0
    v[cons] = 0.0
    assert _fd_error(s, NPTS[t], u, v, 5.0, kappa) < 1e-8


def test_planted_jacobian_factor_fails_fd():
    s = _system(3, O.HEX27)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(s.n_dofs)
    v = rng.standard_normal(s.n_dofs)
    cons = s.constrained().astype(bool)
    u[cons] = 0.0
    v[cons] = 0.0
    assert _fd_error(s, 3, u, v, 5.0, None, variant=1) > 1e-3


@pytest.mark.parametrize("t", [O.HEX8, O.HEX27, O.TET10])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_constant_state_closed_forms(t, variant):
    s = _system(2, t, mask=0)
    c, sigma = 0.7, 3.0
    u = np.full(s.n_dofs, c)
    vals, R = s.fem_reaction_diffusion(NPTS[t], u, sigma, None, variant=variant)
    ok_r = abs(R.sum() - (sigma * c**3 - 1.0)) < 1e-12
    ok_j = abs(vals.sum() - 3 * sigma * c**2) < 1e-12
    if variant == 0:
        assert ok_r and ok_j
    elif variant == 1:
        assert ok_r and not ok_j
    else:
        assert not ok_r and ok_j


def test_element_results_sum_to_assembly():
    s = _system(2, O.HEX27)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(s.n_dofs)
    Ke, Fe = s.element_reaction_diffusion(3, u, 2.0)
    vals, R = s.fem_reaction_diffusion(3, u, 2.0)
    ed = s.elem_dofs().reshape(s.mesh.n_elem, -1)
    R2 = np.zeros(s.n_dofs)
    for e in range(s.mesh.n_elem):
        np.add.at(R2, ed[e], Fe[e])
    assert np.allclose(R2, R, rtol=0, atol=1e-14)
    rows = np.repeat(ed[:, :, None], ed.shape[1], 2).ravel()
    cols = np.repeat(ed[:, None, :], ed.shape[1], 1).ravel()
    K2 = sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(s.n_dofs, s.n_dofs))
    assert abs(K2 - _csr(s, vals)).max() < 1e-14
