"""Pins of the oracle's static condensation (oracle.StaticCondensation, the
restatement of src/numerics/static_condensation.C) by what the algebra fixes:

- the sum of element Schur complements equals the global block elimination
  K_uu - K_uc K_cc^-1 K_cu, because each condensed DoF lives on one element;
- the inverse of the Schur complement is the uu block of K^-1 (dense, tiny mesh);
- forward elimination, a direct solve of the reduced system and back
  substitution reproduce the direct solve of the full system;
- the reduced pattern is the full pattern without the condensed rows/columns;
- a block condensation (elasticity, 3 DoFs per internal node) against the same
  global elimination with a block-diagonal K_cc.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle as O


def setup(n, n_vars=1, seed=3):
    s = O.System(O.Mesh(n, O.HEX27), n_vars, O.ALL_FACES)
    s.mesh.perturb(seed, 0.2)
    rp, ci = s.pattern()
    if n_vars == 1:
        kappa = np.exp(np.random.default_rng(seed).standard_normal(s.mesh.n_elem) * 0.5)
        v, b = s.assemble_poisson(3, kappa, 1)
        Ke, _ = s.element_poisson(3, kappa, 1)
    else:
        v, b = s.assemble_elasticity(3, 1.0, 0.3, -1.0)
        Ke, _ = s.element_elasticity(3, 1.0, 0.3, -1.0)
    K = sp.csr_matrix((v, ci, rp), shape=(s.n_dofs, s.n_dofs))
    return s, K, b, Ke


def global_schur(K, u, c, block):
    Kcc = K[c][:, c].toarray()
    Kcc_inv = np.zeros_like(Kcc)
    for q in range(0, c.size, block):  # block-diagonal by construction; checked below
        Kcc_inv[q:q + block, q:q + block] = np.linalg.inv(Kcc[q:q + block, q:q + block])
    assert np.allclose(Kcc @ Kcc_inv, np.eye(c.size), atol=1e-10)
    return (K[u][:, u] - K[u][:, c] @ sp.csr_matrix(Kcc_inv) @ K[c][:, u]).toarray()


@pytest.mark.parametrize("n", [2, 4])
def test_element_sum_is_global_schur(n):
    s, K, b, Ke = setup(n)
    sc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26])
    assert sc.cdofs.size == s.mesh.n_elem
    Sg = global_schur(K, sc.rfull, sc.cdofs, 1)
    assert np.abs(sc.S.toarray() - Sg).max() <= 1e-13 * np.abs(Sg).max()
    # pattern: the full pattern restricted to the uncondensed DoFs
    rp, ci = s.pattern()
    P = sp.csr_matrix((np.ones(ci.size), ci, rp), shape=K.shape)[sc.rfull][:, sc.rfull]
    P.sort_indices()
    assert np.array_equal(P.indptr, sc.S.indptr) and np.array_equal(P.indices, sc.S.indices)
    # solve through the condensation == solve of the full system
    xf = spl.spsolve(K.tocsc(), b)
    x = sc.backward(b, spl.spsolve(sc.S.tocsc(), sc.forward(b)))
    assert np.abs(x - xf).max() <= 1e-12 * np.abs(xf).max()


def test_schur_inverse_is_block_of_inverse():
    s, K, b, Ke = setup(2)
    sc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26])
    Ki = np.linalg.inv(K.toarray())
    Si = np.linalg.inv(sc.S.toarray())
    assert np.abs(Si - Ki[np.ix_(sc.rfull, sc.rfull)]).max() <= 1e-12 * np.abs(Ki).max()


def test_block_condensation_elasticity():
    s, K, b, Ke = setup(2, n_vars=3)
    sc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26, 53, 80])
    # the three components of one internal node are consecutive in sc.cdofs only if
    # numbered so; use the element order of the condensed DoFs for the block inverse
    order = s.elem_dofs()[:, [26, 53, 80]].ravel()
    Sg = global_schur(K, sc.rfull, order, 3)
    assert np.abs(sc.S.toarray() - Sg).max() <= 1e-12 * np.abs(Sg).max()
    xf = spl.spsolve(K.tocsc(), b)
    x = sc.backward(b, spl.spsolve(sc.S.tocsc(), sc.forward(b)))
    assert np.abs(x - xf).max() <= 1e-11 * np.abs(xf).max()


def test_wrong_sign_fails_the_pin():
    """A planted mistake (Auu + Auc Acc^-1 Acu) is caught by the global elimination."""
    s, K, b, Ke = setup(2)
    sc = O.StaticCondensation(s.elem_dofs(), Ke, s.n_dofs, [26])
    Sg = global_schur(K, sc.rfull, sc.cdofs, 1)
    bad = 2 * K[sc.rfull][:, sc.rfull].toarray() - sc.S.toarray()  # Auu + Auc Acc^-1 Acu
    assert np.abs(bad - Sg).max() > 1e-6 * np.abs(Sg).max()
