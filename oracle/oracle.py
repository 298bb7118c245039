"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's vendored
Eigen 3.3.9 CG (_ref/libeigen_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEX8, HEX27, TET4, TET10 = 10, 12, 8, 9
NPE = {HEX8: 8, HEX27: 27, TET4: 4, TET10: 10}
ALL_FACES = 0x3F

_lib = None
_ref = None

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_I = C.c_int


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        sig = {
            "orc_build_cube": (_P, [C.c_int] * 4),
            "orc_mesh_free": (None, [_P]),
            "orc_mesh_n_nodes": (_I64, [_P]),
            "orc_mesh_n_elem": (_I64, [_P]),
            "orc_mesh_nodes_per_elem": (C.c_int, [_P]),
            "orc_mesh_elem_type": (C.c_int, [_P]),
            "orc_mesh_coords": (None, [_P, _P]),
            "orc_mesh_set_coords": (None, [_P, _P]),
            "orc_mesh_conn": (None, [_P, _P]),
            "orc_mesh_node_bmask": (None, [_P, _P]),
            "orc_mesh_side_bid": (None, [_P, _P]),
            "orc_perturb_interior": (None, [_P, C.c_uint64, _D]),
            "orc_kappa_lognormal": (None, [_I64, C.c_uint64, _P]),
            "orc_min_detj": (_D, [_P, C.c_int]),
            "orc_system_create": (_P, [_P, C.c_int, C.c_int, C.c_uint]),
            "orc_system_free": (None, [_P]),
            "orc_n_dofs": (_I64, [_P]),
            "orc_nnz": (_I64, [_P]),
            "orc_dofs_per_elem": (C.c_int, [_P]),
            "orc_elem_dofs": (None, [_P, _P]),
            "orc_constrained": (None, [_P, _P]),
            "orc_pattern": (None, [_P, _P, _P]),
            "orc_add_constraint_row": (C.c_int, [_P, C.c_int32, C.c_int, _P, _P, _D, C.c_int]),
            "orc_process_constraints": (C.c_int, [_P, C.c_int]),
            "orc_set_asymmetric_constraint_rows": (None, [_P, C.c_int]),
            "orc_constraint_counts": (None, [_P, _P, _P]),
            "orc_constraint_rows": (None, [_P, _P, _P, _P, _P, _P]),
            "orc_constrain_element": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, C.c_int]),
            "orc_enforce_constraints": (None, [_P, _P, C.c_int]),
            "orc_residual_poisson": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P]),
            "orc_assemble_poisson": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P]),
            "orc_element_poisson": (C.c_int, [_P, C.c_int, _P, C.c_int, _I64, _I64, C.c_int, _P, _P]),
            "orc_assemble_elasticity": (C.c_int, [_P, C.c_int, _D, _D, _D, _P, _P]),
            "orc_fem_element_reaction_diffusion": (C.c_int, [_P, C.c_int, _P, _D, C.c_int, _P, _I64, _I64, C.c_int,
                                                             C.c_int, _P, _P]),
            "orc_fem_reaction_diffusion": (C.c_int, [_P, C.c_int, _P, _D, C.c_int, _P, C.c_int, _P, _P]),
            "orc_element_elasticity": (C.c_int, [_P, C.c_int, _D, _D, _D, _I64, _I64, C.c_int, _P, _P]),
            "orc_cg_eigen": (None, [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]),
            "orc_cg_laspack": (None, [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]),
            "orc_bicgstab_eigen": (None, [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]),
            "orc_cg_eigen_pc": (None, [_I64, _P, _P, _P, _P, _P, _D, _I64, _I, _P, _P]),
            "orc_bicgstab_eigen_pc": (None, [_I64, _P, _P, _P, _P, _P, _D, _I64, _I, _P, _P]),
            "orc_spmv": (None, [_I64, _P, _P, _P, _P, _P]),
            "orc_qoi_box_derivative": (C.c_int, [_P, C.c_int, _D, _D, _P]),
            "orc_qoi_box": (_D, [_P, C.c_int, _D, _D, _P]),
            "orc_adjoint_sensitivity": (C.c_int, [_P, C.c_int, _P, C.c_int, _P, _P, _P, _I64, _P]),
            "orc_adjoint_exact": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
            "orc_qgauss_1d": (None, [C.c_int, _P, _P]),
            "orc_qgauss_hex": (None, [C.c_int, _P, _P]),
            "orc_qgauss_tet": (C.c_int, [C.c_int, _P, _P]),
            "orc_shapes": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref():
    """The reference's Eigen 3.3.9 CG (None when oracle/_ref was not built)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libeigen_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_eigen_cg.restype = C.c_int
        R.ref_eigen_cg.argtypes = [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]
        R.ref_eigen_bicgstab.restype = C.c_int
        R.ref_eigen_bicgstab.argtypes = [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]
        for name in ("ref_eigen_cg_identity", "ref_eigen_bicgstab_identity"):
            getattr(R, name).restype = C.c_int
            getattr(R, name).argtypes = [_I64, _P, _P, _P, _P, _P, _D, _I64, _P, _P]
        R.ref_eigen_cg_timed.restype = C.c_int
        R.ref_eigen_cg_timed.argtypes = [_I64, _P, _P, _P, _P, _P, _D, _I64, C.c_int, _P, _P, _P]
        R.ref_eigen_spmv.restype = None
        R.ref_eigen_spmv.argtypes = [_I64, _P, _P, _P, _P, _P]
        _ref = R
    return _ref


class Mesh:
    def __init__(self, n, elem_type, ny=None, nz=None):
        L = lib()
        self.h = L.orc_build_cube(n, ny or n, nz or n, elem_type)
        if not self.h:
            raise ValueError("bad mesh arguments")
        self.type = elem_type
        self.n_nodes = L.orc_mesh_n_nodes(self.h)
        self.n_elem = L.orc_mesh_n_elem(self.h)
        self.npe = L.orc_mesh_nodes_per_elem(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def coords(self):
        a = np.empty((self.n_nodes, 3))
        lib().orc_mesh_coords(self.h, _ptr(a))
        return a

    def set_coords(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        lib().orc_mesh_set_coords(self.h, _ptr(a))

    def conn(self):
        a = np.empty((self.n_elem, self.npe), dtype=np.int32)
        lib().orc_mesh_conn(self.h, _ptr(a))
        return a

    def node_bmask(self):
        a = np.empty(self.n_nodes, dtype=np.uint8)
        lib().orc_mesh_node_bmask(self.h, _ptr(a))
        return a

    def side_bid(self):
        ns = 6 if self.type in (HEX8, HEX27) else 4
        a = np.empty((self.n_elem, ns), dtype=np.int8)
        lib().orc_mesh_side_bid(self.h, _ptr(a))
        return a

    def perturb(self, seed, frac):
        lib().orc_perturb_interior(self.h, seed, frac)

    def min_detj(self, npts):
        return lib().orc_min_detj(self.h, npts)


def kappa_lognormal(n_elem, seed):
    a = np.empty(n_elem)
    lib().orc_kappa_lognormal(n_elem, seed, _ptr(a))
    return a


class System:
    def __init__(self, mesh, n_vars=1, dirichlet_mask=ALL_FACES):
        L = lib()
        self.mesh = mesh
        self.h = L.orc_system_create(mesh.h, n_vars, 0, dirichlet_mask)
        self.n_dofs = L.orc_n_dofs(self.h)
        self.nnz = L.orc_nnz(self.h)
        self.dpe = L.orc_dofs_per_elem(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_system_free(self.h)
            self.h = None

    def elem_dofs(self):
        a = np.empty((self.mesh.n_elem, self.dpe), dtype=np.int32)
        lib().orc_elem_dofs(self.h, _ptr(a))
        return a

    def constrained(self):
        a = np.empty(self.n_dofs, dtype=np.uint8)
        lib().orc_constrained(self.h, _ptr(a))
        return a

    def pattern(self):
        rp = np.empty(self.n_dofs + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        lib().orc_pattern(self.h, _ptr(rp), _ptr(ci))
        return rp, ci

    def add_constraint_row(self, dof, row, rhs=0.0, forbid_overwrite=True):
        """DofMap::add_constraint_row; row is {dof: coefficient}."""
        cols = np.array(list(row.keys()), dtype=np.int32)
        coefs = np.array(list(row.values()), dtype=np.float64)
        if lib().orc_add_constraint_row(self.h, int(dof), len(cols), _ptr(cols), _ptr(coefs), float(rhs),
                                        int(forbid_overwrite)) != 0:
            raise RuntimeError(f"DOF {dof} was already constrained")

    def set_asymmetric_constraint_rows(self, on):
        lib().orc_set_asymmetric_constraint_rows(self.h, int(on))

    def process_constraints(self, error_on_loop=False):
        if lib().orc_process_constraints(self.h, int(error_on_loop)) != 0:
            raise RuntimeError("Constraint loop detected")
        self.nnz = lib().orc_nnz(self.h)

    def constraint_rows(self):
        """{dof: ({dof: coefficient}, value)} after the last processing."""
        nr, ne = C.c_int64(), C.c_int64()
        lib().orc_constraint_counts(self.h, C.byref(nr), C.byref(ne))
        d = np.empty(nr.value, dtype=np.int32)
        rp = np.empty(nr.value + 1, dtype=np.int32)
        cols = np.empty(max(ne.value, 1), dtype=np.int32)
        coefs = np.empty(max(ne.value, 1))
        rhs = np.empty(max(nr.value, 1))
        lib().orc_constraint_rows(self.h, _ptr(d), _ptr(rp), _ptr(cols), _ptr(coefs), _ptr(rhs))
        return {int(d[k]): ({int(cols[t]): float(coefs[t]) for t in range(rp[k], rp[k + 1])}, float(rhs[k]))
                for k in range(nr.value)}

    def constrain_element(self, e, K, F, cap=256):
        K = np.ascontiguousarray(K, dtype=np.float64)
        F = np.ascontiguousarray(F, dtype=np.float64)
        d = np.empty(cap, dtype=np.int32)
        Ko = np.empty(cap * cap)
        Fo = np.empty(cap)
        m = lib().orc_constrain_element(self.h, int(e), _ptr(K), _ptr(F), _ptr(d), _ptr(Ko), _ptr(Fo), cap)
        if m < 0:
            raise RuntimeError("expanded element larger than cap")
        return d[:m].copy(), Ko[:m * m].reshape(m, m).copy(), Fo[:m].copy()

    def enforce_constraints(self, x, homogeneous=False):
        x = np.array(x, dtype=np.float64)
        lib().orc_enforce_constraints(self.h, _ptr(x), int(homogeneous))
        return x

    def residual_poisson(self, npts, u, kappa=None, f_kind=0):
        """FEMSystem residual R(u) of the Poisson physics (Dirichlet rows only)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        R = np.empty(self.n_dofs)
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        if lib().orc_residual_poisson(self.h, npts, None if k is None else _ptr(k), f_kind, _ptr(u), _ptr(R)) != 0:
            raise RuntimeError("degenerate map")
        return R

    def assemble_poisson(self, npts, kappa=None, f_kind=0):
        vals = np.empty(self.nnz)
        rhs = np.empty(self.n_dofs)
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        rc = lib().orc_assemble_poisson(self.h, npts, None if k is None else _ptr(k), f_kind, _ptr(vals), _ptr(rhs))
        if rc != 0:
            raise RuntimeError("degenerate map")
        return vals, rhs

    def element_poisson(self, npts, kappa=None, f_kind=0, e0=0, count=None, constrain=True):
        count = self.mesh.n_elem - e0 if count is None else count
        Ke = np.empty((count, self.dpe, self.dpe))
        Fe = np.empty((count, self.dpe))
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        rc = lib().orc_element_poisson(self.h, npts, None if k is None else _ptr(k), f_kind, e0, count,
                                       int(constrain), _ptr(Ke), _ptr(Fe))
        if rc != 0:
            raise RuntimeError("degenerate map")
        return Ke, Fe

    def fem_reaction_diffusion(self, npts, u, sigma, kappa=None, f_kind=0, jacobian=True, residual=True,
                               variant=0):
        """FEMSystem::assembly of -div(kappa grad u) + sigma u^3 = f at u: (CSR
        values of the constrained Jacobian or None, residual or None)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        vals = np.empty(self.nnz) if jacobian else None
        R = np.empty(self.n_dofs) if residual else None
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        rc = lib().orc_fem_reaction_diffusion(self.h, npts, None if k is None else _ptr(k), sigma, f_kind, _ptr(u),
                                              variant, None if vals is None else _ptr(vals),
                                              None if R is None else _ptr(R))
        if rc != 0:
            raise RuntimeError("degenerate map" if rc == -1 else "one variable, Dirichlet rows only")
        return vals, R

    def element_reaction_diffusion(self, npts, u, sigma, kappa=None, f_kind=0, e0=0, count=None, constrain=True,
                                   variant=0):
        count = self.mesh.n_elem - e0 if count is None else count
        u = np.ascontiguousarray(u, dtype=np.float64)
        Ke = np.empty((count, self.dpe, self.dpe))
        Fe = np.empty((count, self.dpe))
        k = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
        rc = lib().orc_fem_element_reaction_diffusion(self.h, npts, None if k is None else _ptr(k), sigma, f_kind,
                                                      _ptr(u), e0, count, int(constrain), variant, _ptr(Ke),
                                                      _ptr(Fe))
        if rc != 0:
            raise RuntimeError("degenerate map" if rc == -1 else "one variable, Dirichlet rows only")
        return Ke, Fe

    def assemble_elasticity(self, npts, E=1.0, nu=0.3, fz=-1.0):
        vals = np.empty(self.nnz)
        rhs = np.empty(self.n_dofs)
        rc = lib().orc_assemble_elasticity(self.h, npts, E, nu, fz, _ptr(vals), _ptr(rhs))
        if rc != 0:
            raise RuntimeError("degenerate map")
        return vals, rhs

    def element_elasticity(self, npts, E=1.0, nu=0.3, fz=-1.0, e0=0, count=None, constrain=True):
        count = self.mesh.n_elem - e0 if count is None else count
        Ke = np.empty((count, self.dpe, self.dpe))
        Fe = np.empty((count, self.dpe))
        rc = lib().orc_element_elasticity(self.h, npts, E, nu, fz, e0, count, int(constrain), _ptr(Ke), _ptr(Fe))
        if rc != 0:
            raise RuntimeError("degenerate map")
        return Ke, Fe

    def qoi_box_derivative(self, npts, lo=0.25, hi=0.75):
        dq = np.empty(self.n_dofs)
        if lib().orc_qoi_box_derivative(self.h, npts, lo, hi, _ptr(dq)) != 0:
            raise RuntimeError("degenerate map")
        return dq

    def qoi_box(self, npts, u, lo=0.25, hi=0.75):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_qoi_box(self.h, npts, lo, hi, _ptr(u))

    def adjoint_sensitivity(self, npts, kappa, u, z, elems, f_kind=0):
        elems = np.ascontiguousarray(elems, dtype=np.int64)
        out = np.empty(len(elems))
        kappa = np.ascontiguousarray(kappa, dtype=np.float64)
        rc = lib().orc_adjoint_sensitivity(self.h, npts, _ptr(kappa), f_kind, _ptr(u), _ptr(z), _ptr(elems),
                                           len(elems), _ptr(out))
        if rc != 0:
            raise RuntimeError("degenerate map")
        return out

    def adjoint_exact(self, npts, kappa, u, z):
        out = np.empty(self.mesh.n_elem)
        kappa = np.ascontiguousarray(kappa, dtype=np.float64)
        if lib().orc_adjoint_exact(self.h, npts, _ptr(kappa), _ptr(u), _ptr(z), _ptr(out)) != 0:
            raise RuntimeError("degenerate map")
        return out


def _cg(fn, rp, ci, vals, b, tol, maxits, x0=None):
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    its = C.c_int64()
    err = C.c_double()
    fn(n, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(np.ascontiguousarray(b)), _ptr(x), tol, maxits, C.byref(its),
       C.byref(err))
    return x, its.value, err.value


PRECOND = {"jacobi": 1, "identity": 0}  # libMesh JACOBI_PRECOND / IDENTITY_PRECOND


def _pc(fn, precond):
    return lambda n, *a: fn(n, *a[:7], PRECOND[precond], *a[7:])


def cg_eigen(rp, ci, vals, b, tol=1e-12, maxits=5000, x0=None, precond="jacobi"):
    return _cg(_pc(lib().orc_cg_eigen_pc, precond), rp, ci, vals, b, tol, maxits, x0)


def cg_laspack(rp, ci, vals, b, tol=1e-12, maxits=5000):
    return _cg(lib().orc_cg_laspack, rp, ci, vals, b, tol, maxits)


def bicgstab_eigen(rp, ci, vals, b, tol=1e-12, maxits=5000, x0=None, precond="jacobi"):
    return _cg(_pc(lib().orc_bicgstab_eigen_pc, precond), rp, ci, vals, b, tol, maxits, x0)


def ref_bicgstab_eigen(rp, ci, vals, b, tol=1e-12, maxits=5000, x0=None, precond="jacobi"):
    R = ref()
    if R is None:
        return None
    fn = R.ref_eigen_bicgstab if precond == "jacobi" else R.ref_eigen_bicgstab_identity
    return _cg(fn, rp, ci, vals, b, tol, maxits, x0)


def ref_cg_eigen(rp, ci, vals, b, tol=1e-12, maxits=5000, x0=None, precond="jacobi"):
    R = ref()
    if R is None:
        return None
    fn = R.ref_eigen_cg if precond == "jacobi" else R.ref_eigen_cg_identity
    return _cg(fn, rp, ci, vals, b, tol, maxits, x0)


def ref_cg_eigen_timed(rp, ci, vals, b, tol, maxits, threads, x0=None):
    """The reference's Eigen CG + Jacobi with only the solve on the clock:
    (x, iterations, error, seconds), or None without oracle/_ref."""
    R = ref()
    if R is None:
        return None
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    its, err, sec = C.c_int64(), C.c_double(), C.c_double()
    R.ref_eigen_cg_timed(n, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(np.ascontiguousarray(b)), _ptr(x), tol, maxits,
                         threads, C.byref(its), C.byref(err), C.byref(sec))
    return x, its.value, err.value, sec.value


def spmv(rp, ci, vals, x):
    y = np.empty(len(rp) - 1)
    lib().orc_spmv(len(rp) - 1, _ptr(rp), _ptr(ci), _ptr(vals), _ptr(np.ascontiguousarray(x)), _ptr(y))
    return y


def qgauss_1d(npts):
    x = np.empty(npts)
    w = np.empty(npts)
    lib().orc_qgauss_1d(npts, _ptr(x), _ptr(w))
    return x, w


def qgauss_hex(npts):
    p = np.empty((npts ** 3, 3))
    w = np.empty(npts ** 3)
    lib().orc_qgauss_hex(npts, _ptr(p), _ptr(w))
    return p, w


def qgauss_tet(order):
    n = lib().orc_qgauss_tet(order, None, None)
    p = np.empty((n, 3))
    w = np.empty(n)
    lib().orc_qgauss_tet(order, _ptr(p), _ptr(w))
    return p, w


def shapes(elem_type, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    nq = len(pts)
    n = NPE[elem_type]
    phi = np.empty((n, nq))
    dphi = np.empty((n, nq, 3))
    lib().orc_shapes(elem_type, 0, nq, _ptr(pts), _ptr(phi), _ptr(dphi))
    return phi, dphi


class StaticCondensation:
    """Element-by-element static condensation, restating libMesh's
    StaticCondensation (src/numerics/static_condensation.C): per element the
    condensed block Acc is factored and Auu - Auc Acc^-1 Acu is added into the
    reduced matrix (:205-211); forward_elimination adds -Auc Acc^-1 b_c to the
    uncondensed part of b (:383-417); backwards_substitution sets
    x_c = Acc^-1 (b_c - Acu x_u) (:419-456).  numpy's LU solve stands in for
    Eigen's partialPivLu.  cond_local: element-local indices condensed in
    every element (the DoFs of internal nodes, static_condensation_dof_map.C:187-199).
    Reduced DoFs are the uncondensed ones in ascending full order."""

    def __init__(self, elem_dofs, Ke, n_dofs, cond_local):
        import scipy.sparse as sp
        ed = np.asarray(elem_dofs, dtype=np.int64)
        dpe = ed.shape[1]
        c = np.asarray(cond_local, dtype=np.int64)
        u = np.setdiff1d(np.arange(dpe), c)
        self.ed, self.c, self.u = ed, c, u
        self.cdofs = np.unique(ed[:, c])
        keep = np.ones(n_dofs, dtype=bool)
        keep[self.cdofs] = False
        self.rfull = np.nonzero(keep)[0]
        self.ridx = np.full(n_dofs, -1, dtype=np.int64)
        self.ridx[self.rfull] = np.arange(self.rfull.size)
        Ke = np.asarray(Ke)
        self.Acc = Ke[:, c][:, :, c]
        self.Acu = Ke[:, c][:, :, u]
        self.Auc = Ke[:, u][:, :, c]
        Auu = Ke[:, u][:, :, u]
        S = Auu - self.Auc @ np.linalg.solve(self.Acc, self.Acu)
        ru = self.ridx[ed[:, u]]
        rows = np.repeat(ru, u.size, axis=1).ravel()
        cols = np.tile(ru, (1, u.size)).ravel()
        m = self.rfull.size
        self.S = sp.csr_matrix((S.ravel(), (rows, cols)), shape=(m, m))
        self.S.sum_duplicates()
        self.S.sort_indices()

    def forward(self, b):
        g = np.asarray(b, dtype=np.float64)[self.rfull].copy()
        bc = b[self.ed[:, self.c]][:, :, None]
        gu = -(self.Auc @ np.linalg.solve(self.Acc, bc))[:, :, 0]
        np.add.at(g, self.ridx[self.ed[:, self.u]], gu)
        return g

    def backward(self, b, y):
        x = np.zeros(self.ridx.size)
        x[self.rfull] = y
        xu = x[self.ed[:, self.u]][:, :, None]
        bc = b[self.ed[:, self.c]][:, :, None]
        x[self.ed[:, self.c]] = np.linalg.solve(self.Acc, bc - self.Acu @ xu)[:, :, 0]
        return x
