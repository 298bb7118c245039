"""libmesh_amd -- MI355X-native libMesh assembly + Jacobi-CG hot path.

Python access to the two native libraries of the package:

* ``liblmhip.so``  -- HIP kernels for gfx950 behind the C ABI of ``include/lmhip.h``
  (mesh/DoF upload, GPU sparsity, element assembly, scatter, SpMV, CG, mixed
  precision, adjoint).
* ``liblmesh.so``  -- the host-side libMesh API subset (``include/lmesh_host.h``):
  ``MeshTools::Generation::build_cube`` and ``DofMap`` numbering.

There is no CPU fallback: importing works without a GPU (so the symbol tables can
be checked), but every compute call goes to the HIP library and raises
``LmhError`` when it fails.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "liblmhip.so")
LIB_HOST = os.path.join(HERE, "liblmesh.so")

HEX8, HEX27, TET4, TET10 = 10, 12, 8, 9
FP64, FP32 = 0, 1
SCATTER_COLORED, SCATTER_GATHER = 0, 1
CONSTRAINT_FORBID_OVERWRITE, CONSTRAINT_ERROR_ON_LOOP, CONSTRAINT_SYMMETRIC_ROWS = 1, 2, 4
ALL_FACES = 0x3F
CONVERGED_RTOL, DIVERGED_ITS = 2, -3
DIVERGED_BREAKDOWN, DIVERGED_NANORINF = -5, -9  # LMH_DIVERGED_* (include/lmhip.h)
PRECOND_IDENTITY, PRECOND_JACOBI = 0, 1  # libMesh PreconditionerType values
# lmh_vector_op operations and lmh_vector_reduce reductions (include/lmhip.h)
(VEC_SET, VEC_ADD_SCALAR, VEC_SCALE, VEC_AXPY, VEC_MUL, VEC_DIV, VEC_RECIPROCAL, VEC_ABS, VEC_PMUL, VEC_PDIV,
 VEC_COPY) = range(11)
(VEC_SUM, VEC_L1, VEC_L2, VEC_LINF, VEC_MIN, VEC_MAX, VEC_DOT, VEC_L1_DIFF, VEC_L2_DIFF) = range(9)
CG_NOALLOC_MAX_ROWS = 4194304  # LMH_CG_NOALLOC_MAX_ROWS (include/lmhip.h)

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_I = C.c_int


class LmhError(RuntimeError):
    pass


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


# name -> (restype, argtypes); the declarations of include/lmhip.h
HIP_API = {
    "lmh_last_error": (C.c_char_p, []),
    "lmh_context_create": (_I, [_I, C.POINTER(_P)]),
    "lmh_context_destroy": (_I, [_P]),
    "lmh_synchronize": (_I, [_P]),
    "lmh_context_stream": (_P, [_P]),
    "lmh_malloc": (_I, [_P, C.c_size_t, C.POINTER(_P)]),
    "lmh_free": (_I, [_P, _P]),
    "lmh_memcpy_h2d": (_I, [_P, _P, _P, C.c_size_t]),
    "lmh_memcpy_d2h": (_I, [_P, _P, _P, C.c_size_t]),
    "lmh_memcpy_d2d": (_I, [_P, _P, _P, C.c_size_t]),
    "lmh_memset": (_I, [_P, _P, _I, C.c_size_t]),
    "lmh_mesh_create": (_I, [_P, _I, _I64, _P, _I64, _P, C.POINTER(_P)]),
    "lmh_mesh_destroy": (_I, [_P]),
    "lmh_mesh_set_coords": (_I, [_P, _P]),
    "lmh_dofmap_create": (_I, [_P, _P, _I, _I64, _P, _P, C.POINTER(_P)]),
    "lmh_dofmap_destroy": (_I, [_P]),
    "lmh_dofmap_set_constraints": (_I, [_P, _I64, _P, _P, _P, _P, _P, _I]),
    "lmh_dofmap_constraint_counts": (_I, [_P, _P, _P, _P]),
    "lmh_dofmap_get_constraints": (_I, [_P, _P, _P, _P, _P, _P]),
    "lmh_enforce_constraints": (_I, [_P, _P, _I]),
    "lmh_assemble_residual_poisson": (_I, [_P, _P, _P, _I, _P, _I, _P, _P]),
    "lmh_axpy": (_I, [_P, _I64, _D, _P, _P]),
    "lmh_vector_op": (_I, [_P, _I, _I64, _D, _P, _P, _P]),
    "lmh_vector_reduce": (_I, [_P, _I, _I64, _P, _P, C.POINTER(_D)]),
    "lmh_bicgstab_solve": (_I, [_P, _P, _P, _D, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I)]),
    "lmh_matrix_create": (_I, [_P, _P, C.POINTER(_P)]),
    "lmh_matrix_create_csr": (_I, [_P, _I64, _P, _P, C.POINTER(_P)]),
    "lmh_matrix_destroy": (_I, [_P]),
    "lmh_matrix_cols": (_I64, [_P]),
    "lmh_matrix_create_csr2": (_I, [_P, _I64, _I64, _P, _P, C.POINTER(_P)]),
    "lmh_matrix_rows": (_I64, [_P]),
    "lmh_matrix_nnz": (_I64, [_P]),
    "lmh_matrix_n_colors": (_I, [_P]),
    "lmh_matrix_get_pattern": (_I, [_P, _P, _P]),
    "lmh_matrix_device_arrays": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "lmh_matrix_zero": (_I, [_P]),
    "lmh_matrix_get_values": (_I, [_P, _P]),
    "lmh_matrix_set_values": (_I, [_P, _P]),
    "lmh_matrix_get_values_fp32": (_I, [_P, _P]),
    "lmh_matrix_add_element_matrices": (_I, [_P, _I64, _I, _P, _P]),
    "lmh_vector_add_element_vectors": (_I, [_P, _P, _I64, _I, _P, _P]),
    "lmh_fem_element_data": (_I, [_P, _P, _I, _P]),
    "lmh_fem_check_flag": (_I, [_P]),
    "lmh_matrix_scatter_elements": (_I, [_P, _P, _P, _P]),
    "lmh_assemble_poisson": (_I, [_P, _P, _P, _I, _P, _I, _P, _I, _I]),
    "lmh_element_poisson": (_I, [_P, _P, _P, _I, _P, _I, _I, _P, _P]),
    "lmh_spmv": (_I, [_P, _P, _P]),
    "lmh_spmv_fp32": (_I, [_P, _P, _P]),
    "lmh_residual": (_I, [_P, _P, _P, _P]),
    "lmh_norm2": (_D, [_P, _P, _I64]),
    "lmh_cg_solve": (_I, [_P, _P, _P, _D, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I)]),
    "lmh_cg_solve_noalloc": (_I, [_P, _P, _P, _D, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I)]),
    "lmh_mixed_solve": (_I, [_P, _P, _P, _D, _D, _I, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I), _P]),
    "lmh_matrix_make_fp32": (_I, [_P]),
    "lmh_matrix_set_fp32_compression": (_I, [_P, _I]),
    "lmh_matrix_fp32_compression": (_I, [_P]),
    "lmh_matrix_window_bits": (_I, [_P]),
    "lmh_matrix_set_preconditioner": (_I, [_P, _I]),
    "lmh_bicgstab_solve_transpose": (_I, [_P, _P, _P, _D, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I)]),
    "lmh_dofmap_condensed_dofs": (_I, [_P, C.POINTER(_I64), _P]),
    "lmh_condensation_create": (_I, [_P, _I64, _P, C.POINTER(_P)]),
    "lmh_condensation_setup": (_I, [_P]),
    "lmh_condensation_reduced_matrix": (_P, [_P]),
    "lmh_condensation_n_condensed": (_I64, [_P]),
    "lmh_condensation_reduced_dofs": (_I, [_P, _P]),
    "lmh_condensation_forward": (_I, [_P, _P, _P]),
    "lmh_condensation_backward": (_I, [_P, _P, _P, _P]),
    "lmh_condensation_solve": (_I, [_P, _P, _P, _D, _I64, C.POINTER(_I64), C.POINTER(_D), C.POINTER(_I)]),
    "lmh_condensation_destroy": (_I, [_P]),
    "lmh_qoi_box_derivative": (_I, [_P, _P, _I, _D, _D, _P]),
    "lmh_adjoint_kappa_gradient": (_I, [_P, _P, _I, _P, _P, _P, _P]),
    "lmh_batch_create": (_I, [_P, _I, _P, C.POINTER(_P)]),
    "lmh_batch_destroy": (_I, [_P]),
    "lmh_batch_size": (_I, [_P]),
    "lmh_batch_matrix": (_I, [_P, _I, C.POINTER(_P)]),
    "lmh_batch_rhs": (_I, [_P, _I, C.POINTER(_P)]),
    "lmh_batch_set_tiles": (_I, [_P, _P, _P]),
    "lmh_batch_get_tiles": (_I, [_P, _P, _P]),
    "lmh_batch_autotune": (_I, [_P, _D, _D, _D, _I, _P]),
    "lmh_batch_assemble_elasticity": (_I, [_P, _D, _D, _D]),
    "lmh_profile_enable": (_I, [_P, _I]),
    "lmh_profile_query": (_I, [_P, C.c_char_p, C.POINTER(_D), C.POINTER(_I64)]),
    "lmh_profile_reset": (_I, [_P]),
}

HOST_API = {
    "lmesh_build_cube": (_P, [_I, _I, _I, _I, _I, C.c_uint]),
    "lmesh_free": (None, [_P]),
    "lmesh_last_error": (C.c_char_p, []),
    "lmesh_perturb": (None, [_P, C.c_uint64, _D]),
    "lmesh_lognormal": (None, [_I64, C.c_uint64, _P]),
    "lmesh_batch_recipe": (None, [C.c_uint64, _I, _I, _I, _P, _P]),
    "lmesh_matrix_read": (_I, [C.c_char_p, C.POINTER(_I64), C.POINTER(_I64), C.POINTER(_I64), _P, _P, _P]),
    "lmesh_matrix_write_matlab": (_I, [C.c_char_p, _I64, _I64, _P, _P, _P]),
    "lmesh_es_file_convert": (_I, [C.c_char_p, C.c_char_p]),
    "lmesh_n_nodes": (_I64, [_P]),
    "lmesh_n_elem": (_I64, [_P]),
    "lmesh_n_dofs": (_I64, [_P]),
    "lmesh_nodes_per_elem": (_I, [_P]),
    "lmesh_dofs_per_elem": (_I, [_P]),
    "lmesh_elem_type": (_I, [_P]),
    "lmesh_coords": (None, [_P, _P]),
    "lmesh_conn": (None, [_P, _P]),
    "lmesh_elem_dofs": (None, [_P, _P]),
    "lmesh_constrained": (None, [_P, _P]),
}

_hip = None
_host = None


def _bind(path, api):
    if not os.path.exists(path):
        raise LmhError(f"native library missing: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in api.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def hip():
    """The HIP C-ABI library (liblmhip.so)."""
    global _hip
    if _hip is None:
        _hip = _bind(LIB_HIP, HIP_API)
    return _hip


def host():
    """The host libMesh-API library (liblmesh.so)."""
    global _host
    if _host is None:
        hip()  # liblmesh.so links liblmhip.so
        _host = _bind(LIB_HOST, HOST_API)
    return _host


def check(rc):
    if rc != 0:
        raise LmhError(f"lmhip error {rc}: {hip().lmh_last_error().decode()}")


# ---------------------------------------------------------------------------
# host problem generation (libMesh build_cube + DofMap)
# ---------------------------------------------------------------------------
class Problem:
    """A generated cube mesh with its DoF map, as host arrays."""

    def __init__(self, n, elem_type, n_vars=1, dirichlet_mask=ALL_FACES, ny=None, nz=None):
        L = host()
        self._h = L.lmesh_build_cube(n, ny or n, nz or n, elem_type, n_vars, dirichlet_mask)
        if not self._h:
            raise LmhError(L.lmesh_last_error().decode())
        self.elem_type = elem_type
        self.n_vars = n_vars
        self._refresh()

    def _refresh(self):
        L = host()
        h = self._h
        self.n_nodes = L.lmesh_n_nodes(h)
        self.n_elem = L.lmesh_n_elem(h)
        self.n_dofs = L.lmesh_n_dofs(h)
        self.npe = L.lmesh_nodes_per_elem(h)
        self.dpe = L.lmesh_dofs_per_elem(h)
        self.coords = np.empty((self.n_nodes, 3))
        L.lmesh_coords(h, _ptr(self.coords))
        self.conn = np.empty((self.n_elem, self.npe), dtype=np.int32)
        L.lmesh_conn(h, _ptr(self.conn))
        self.elem_dofs = np.empty((self.n_elem, self.dpe), dtype=np.int32)
        L.lmesh_elem_dofs(h, _ptr(self.elem_dofs))
        self.constrained = np.empty(self.n_dofs, dtype=np.uint8)
        L.lmesh_constrained(h, _ptr(self.constrained))

    def perturb(self, seed, frac):
        host().lmesh_perturb(self._h, seed, frac)
        self._refresh()

    def __del__(self):
        if getattr(self, "_h", None):
            host().lmesh_free(self._h)
            self._h = None


def lognormal(n, seed):
    out = np.empty(n)
    host().lmesh_lognormal(n, seed, _ptr(out))
    return out


def read_matrix(filename):
    """(m, n, rows, cols, values) of a sparse matrix file (Matlab text, gzipped
    or not, or PETSc binary), entries in file order, 0-based."""
    L = host()
    m, n, nnz = _I64(), _I64(), _I64()
    fn = str(filename).encode()
    if L.lmesh_matrix_read(fn, C.byref(m), C.byref(n), C.byref(nnz), None, None, None) != 0:
        raise LmhError(L.lmesh_last_error().decode())
    i = np.empty(nnz.value, dtype=np.int64)
    j = np.empty(nnz.value, dtype=np.int64)
    v = np.empty(nnz.value)
    if L.lmesh_matrix_read(fn, C.byref(m), C.byref(n), C.byref(nnz), _ptr(i), _ptr(j), _ptr(v)) != 0:
        raise LmhError(L.lmesh_last_error().decode())
    return m.value, n.value, i, j, v


def write_matrix_matlab(filename, m, n, row_ptr, col_idx, vals):
    L = host()
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    if L.lmesh_matrix_write_matlab(str(filename).encode(), m, n, _ptr(rp), _ptr(ci), _ptr(v)) != 0:
        raise LmhError(L.lmesh_last_error().decode())



def convert_equation_systems_file(in_name, out_name):
    """Read an EquationSystems file and write it again (ASCII or XDR chosen by
    each name: ".xdr" selects XDR), header and values unchanged."""
    if host().lmesh_es_file_convert(str(in_name).encode(), str(out_name).encode()) != 0:
        raise LmhError(host().lmesh_last_error().decode())

def batch_recipe(seed, count, nmin=8, nmax=48):
    """(edge, element type) per problem of a seeded batch of cube problems."""
    edge = np.empty(count, dtype=np.int32)
    kind = np.empty(count, dtype=np.int32)
    host().lmesh_batch_recipe(seed, count, nmin, nmax, _ptr(edge), _ptr(kind))
    return edge, kind


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        self._h = _P()
        check(hip().lmh_context_create(device, C.byref(self._h)))

    @property
    def h(self):
        return self._h

    def sync(self):
        check(hip().lmh_synchronize(self._h))

    def close(self):
        if self._h:
            hip().lmh_context_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # profiling (HIP events on the context stream)
    def profile(self, on=True):
        check(hip().lmh_profile_enable(self._h, int(on)))

    def profile_reset(self):
        check(hip().lmh_profile_reset(self._h))

    def profile_query(self, name):
        ms = _D()
        cnt = _I64()
        check(hip().lmh_profile_query(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class DeviceArray:
    """A device allocation owned by a Context."""

    def __init__(self, ctx, n, dtype=np.float64, data=None):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = _P()
        check(hip().lmh_malloc(ctx.h, max(self.nbytes, 1), C.byref(self.ptr)))
        if data is not None:
            self.upload(data)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(hip().lmh_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(hip().lmh_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(hip().lmh_memset(self.ctx.h, self.ptr, 0, self.nbytes))

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        check(hip().lmh_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes))

    def free(self):
        if self.ptr:
            hip().lmh_free(self.ctx.h, self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            if self.ctx.h:
                self.free()
        except Exception:
            pass


class GpuProblem:
    """Mesh + DoF map + CSR matrix on the device (lmh_mesh / lmh_dofmap / lmh_matrix)."""

    def __init__(self, ctx, coords, conn, elem_type, elem_dofs, n_dofs, constrained, n_vars=1, build_matrix=True):
        L = hip()
        self.ctx = ctx
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        edofs = np.ascontiguousarray(elem_dofs, dtype=np.int32)
        cons = np.ascontiguousarray(constrained, dtype=np.uint8)
        self.n_elem, self.npe = conn.shape
        self.n_dofs = int(n_dofs)
        self.dpe = edofs.shape[1]
        self.mesh = _P()
        check(L.lmh_mesh_create(ctx.h, elem_type, coords.shape[0], _ptr(coords), self.n_elem, _ptr(conn),
                                C.byref(self.mesh)))
        self.dofmap = _P()
        check(L.lmh_dofmap_create(ctx.h, self.mesh, n_vars, self.n_dofs, _ptr(edofs), _ptr(cons),
                                  C.byref(self.dofmap)))
        self.matrix = _P()
        if build_matrix:
            check(L.lmh_matrix_create(ctx.h, self.dofmap, C.byref(self.matrix)))
            self.nnz = L.lmh_matrix_nnz(self.matrix)
            self.n_colors = L.lmh_matrix_n_colors(self.matrix)

    def set_constraints(self, rows, forbid_overwrite=False, error_on_loop=False, symmetric_rows=False):
        """DofMap::add_constraint_row for each (dof, {dof: coefficient}, value)
        of `rows` (in order) on top of the Dirichlet rows, then
        process_constraints; the matrix is created again for the new pattern."""
        rows = list(rows)
        dofs = np.array([r[0] for r in rows], dtype=np.int32)
        ptr = np.zeros(len(rows) + 1, dtype=np.int32)
        cols, coefs = [], []
        for k, (_, row, _) in enumerate(rows):
            cols.extend(row.keys())
            coefs.extend(row.values())
            ptr[k + 1] = len(cols)
        cols = np.array(cols if cols else [0], dtype=np.int32)
        coefs = np.array(coefs if coefs else [0.0], dtype=np.float64)
        rhs = np.array([r[2] for r in rows] if rows else [0.0], dtype=np.float64)
        flags = ((CONSTRAINT_FORBID_OVERWRITE if forbid_overwrite else 0) |
                 (CONSTRAINT_ERROR_ON_LOOP if error_on_loop else 0) |
                 (CONSTRAINT_SYMMETRIC_ROWS if symmetric_rows else 0))
        L = hip()
        check(L.lmh_dofmap_set_constraints(self.dofmap, len(rows), _ptr(dofs) if rows else None, _ptr(ptr),
                                           _ptr(cols), _ptr(coefs), _ptr(rhs), flags))
        if self.matrix:
            L.lmh_matrix_destroy(self.matrix)
            self.matrix = _P()
            check(L.lmh_matrix_create(self.ctx.h, self.dofmap, C.byref(self.matrix)))
            self.nnz = L.lmh_matrix_nnz(self.matrix)
            self.n_colors = L.lmh_matrix_n_colors(self.matrix)

    def constraint_rows(self):
        """{dof: ({dof: coefficient}, value)} after processing."""
        nr, ne, ng = _I64(), _I64(), _I64()
        L = hip()
        check(L.lmh_dofmap_constraint_counts(self.dofmap, C.byref(nr), C.byref(ne), C.byref(ng)))
        d = np.empty(max(nr.value, 1), dtype=np.int32)
        ptr = np.empty(nr.value + 1, dtype=np.int32)
        cols = np.empty(max(ne.value, 1), dtype=np.int32)
        coefs = np.empty(max(ne.value, 1))
        rhs = np.empty(max(nr.value, 1))
        check(L.lmh_dofmap_get_constraints(self.dofmap, _ptr(d), _ptr(ptr), _ptr(cols), _ptr(coefs), _ptr(rhs)))
        return {int(d[k]): ({int(cols[t]): float(coefs[t]) for t in range(ptr[k], ptr[k + 1])}, float(rhs[k]))
                for k in range(nr.value)}

    def n_general_elements(self):
        ng = _I64()
        check(hip().lmh_dofmap_constraint_counts(self.dofmap, None, None, C.byref(ng)))
        return ng.value

    def enforce_constraints(self, x_d, homogeneous=False):
        """DofMap::enforce_constraints_exactly on a device vector."""
        check(hip().lmh_enforce_constraints(self.dofmap, x_d.ptr, int(homogeneous)))

    @classmethod
    def from_problem(cls, ctx, p, build_matrix=True):
        return cls(ctx, p.coords, p.conn, p.elem_type, p.elem_dofs, p.n_dofs, p.constrained, p.n_vars, build_matrix)

    def pattern(self):
        rp = np.empty(self.n_dofs + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        check(hip().lmh_matrix_get_pattern(self.matrix, _ptr(rp), _ptr(ci)))
        return rp, ci

    def values(self):
        v = np.empty(self.nnz)
        check(hip().lmh_matrix_get_values(self.matrix, _ptr(v)))
        return v

    def values_fp32(self):
        v = np.empty(self.nnz, dtype=np.float32)
        check(hip().lmh_matrix_get_values_fp32(self.matrix, _ptr(v)))
        return v

    def set_values(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        check(hip().lmh_matrix_set_values(self.matrix, _ptr(v)))

    def assemble_poisson(self, npts, kappa_d, f_kind, rhs_d, precision=FP64, scatter=SCATTER_GATHER):
        check(hip().lmh_assemble_poisson(self.matrix, self.mesh, self.dofmap, npts,
                                         kappa_d.ptr if kappa_d is not None else None, f_kind,
                                         rhs_d.ptr if rhs_d is not None else None, precision, scatter))

    def residual_poisson(self, npts, kappa_d, f_kind, u_d, R_d):
        """FEMSystem residual R(u) = sum_e C^T (K_e u_e - F_e) of the Poisson physics."""
        check(hip().lmh_assemble_residual_poisson(self.matrix, self.mesh, self.dofmap, npts,
                                                  kappa_d.ptr if kappa_d is not None else None, f_kind, u_d.ptr,
                                                  R_d.ptr))

    def element_poisson(self, npts, kappa_d, f_kind, Ke_d, Fe_d, constrain=True):
        check(hip().lmh_element_poisson(self.ctx.h, self.mesh, self.dofmap, npts,
                                        kappa_d.ptr if kappa_d is not None else None, f_kind, int(constrain),
                                        Ke_d.ptr, Fe_d.ptr))

    def cg(self, b_d, x_d, rtol, maxits=5000):
        its = _I64()
        res = _D()
        reason = _I()
        check(hip().lmh_cg_solve(self.matrix, b_d.ptr, x_d.ptr, rtol, maxits, C.byref(its), C.byref(res),
                                 C.byref(reason)))
        return its.value, res.value, reason.value

    def cg_noalloc(self, b_d, x_d, rtol, maxits=5000):
        """cg without device allocation: work vectors in the library's fixed
        scratch array (at most CG_NOALLOC_MAX_ROWS rows)."""
        its = _I64()
        res = _D()
        reason = _I()
        check(hip().lmh_cg_solve_noalloc(self.matrix, b_d.ptr, x_d.ptr, rtol, maxits, C.byref(its), C.byref(res),
                                         C.byref(reason)))
        return its.value, res.value, reason.value

    def condensed_dofs(self):
        """DoFs of the elements' internal nodes (HEX27 node 26), element by element."""
        n = _I64()
        check(hip().lmh_dofmap_condensed_dofs(self.dofmap, C.byref(n), None))
        out = np.empty(n.value, dtype=np.int32)
        if n.value:
            check(hip().lmh_dofmap_condensed_dofs(self.dofmap, C.byref(n), _ptr(out)))
        return out

    def bicgstab(self, b_d, x_d, rtol, maxits=5000, transpose=False):
        its = _I64()
        res = _D()
        reason = _I()
        f = hip().lmh_bicgstab_solve_transpose if transpose else hip().lmh_bicgstab_solve
        check(f(self.matrix, b_d.ptr, x_d.ptr, rtol, maxits, C.byref(its), C.byref(res), C.byref(reason)))
        return its.value, res.value, reason.value

    def make_fp32(self):
        check(hip().lmh_matrix_make_fp32(self.matrix))

    def set_preconditioner(self, precond):
        """PRECOND_JACOBI (default) or PRECOND_IDENTITY for cg / mixed / bicgstab."""
        check(hip().lmh_matrix_set_preconditioner(self.matrix, precond))

    def mixed(self, b_d, x_d, rtol, inner_rtol=0.1, mode=0, maxits=20000):
        its = _I64()
        res = _D()
        reason = _I()
        stats = np.zeros(3)
        check(hip().lmh_mixed_solve(self.matrix, b_d.ptr, x_d.ptr, rtol, inner_rtol, mode, maxits, C.byref(its),
                                    C.byref(res), C.byref(reason), _ptr(stats)))
        return its.value, res.value, reason.value, stats

    def spmv(self, x_d, y_d):
        check(hip().lmh_spmv(self.matrix, x_d.ptr, y_d.ptr))

    def residual(self, b_d, x_d, r_d):
        check(hip().lmh_residual(self.matrix, b_d.ptr, x_d.ptr, r_d.ptr))

    def qoi_box_derivative(self, npts, lo, hi, dq_d):
        check(hip().lmh_qoi_box_derivative(self.mesh, self.dofmap, npts, lo, hi, dq_d.ptr))

    def adjoint_kappa_gradient(self, npts, kappa_d, u_d, lam_d, grad_d):
        check(hip().lmh_adjoint_kappa_gradient(self.mesh, self.dofmap, npts, kappa_d.ptr, u_d.ptr, lam_d.ptr,
                                               grad_d.ptr))

    def close(self):
        L = hip()
        if self.matrix:
            L.lmh_matrix_destroy(self.matrix)
            self.matrix = _P()
        if self.dofmap:
            L.lmh_dofmap_destroy(self.dofmap)
            self.dofmap = _P()
        if self.mesh:
            L.lmh_mesh_destroy(self.mesh)
            self.mesh = _P()

    def __del__(self):
        try:
            if self.ctx.h:
                self.close()
        except Exception:
            pass


class Batch:
    """Batched FP32 elasticity assembly of many GpuProblems (lmh_batch); the
    problems keep their meshes and DoF maps, the batch owns matrices and rhs."""

    TYPES = (HEX8, HEX27, TET4, TET10)

    def __init__(self, ctx, problems):
        self.ctx = ctx
        self.problems = list(problems)
        arr = (_P * len(self.problems))(*[gp.dofmap for gp in self.problems])
        self.h = _P()
        check(hip().lmh_batch_create(ctx.h, len(self.problems), arr, C.byref(self.h)))
        self.matrices = []
        self.rhs_ptrs = []
        for p in range(len(self.problems)):
            A = _P()
            check(hip().lmh_batch_matrix(self.h, p, C.byref(A)))
            r = _P()
            check(hip().lmh_batch_rhs(self.h, p, C.byref(r)))
            self.matrices.append(A)
            self.rhs_ptrs.append(r)

    def tiles(self):
        lpe = np.zeros(4, dtype=np.int32)
        pp = np.zeros(4, dtype=np.int32)
        check(hip().lmh_batch_get_tiles(self.h, _ptr(lpe), _ptr(pp)))
        return lpe, pp

    def set_tiles(self, lpe, passes):
        lpe = np.ascontiguousarray(lpe, dtype=np.int32)
        passes = np.ascontiguousarray(passes, dtype=np.int32)
        check(hip().lmh_batch_set_tiles(self.h, _ptr(lpe), _ptr(passes)))

    def autotune(self, E=1.0, nu=0.3, fz=-1.0, reps=2):
        best = np.zeros(4)
        check(hip().lmh_batch_autotune(self.h, E, nu, fz, reps, _ptr(best)))
        return best

    def assemble(self, E=1.0, nu=0.3, fz=-1.0):
        check(hip().lmh_batch_assemble_elasticity(self.h, E, nu, fz))

    def nnz(self, p):
        return hip().lmh_matrix_nnz(self.matrices[p])

    def pattern(self, p):
        n = self.problems[p].n_dofs
        rp = np.empty(n + 1, dtype=np.int32)
        ci = np.empty(self.nnz(p), dtype=np.int32)
        check(hip().lmh_matrix_get_pattern(self.matrices[p], _ptr(rp), _ptr(ci)))
        return rp, ci

    def values(self, p):
        v = np.empty(self.nnz(p), dtype=np.float32)
        check(hip().lmh_matrix_get_values_fp32(self.matrices[p], _ptr(v)))
        return v

    def rhs(self, p):
        n = self.problems[p].n_dofs
        out = np.empty(n)
        check(hip().lmh_memcpy_d2h(self.ctx.h, _ptr(out), self.rhs_ptrs[p], 8 * n))
        return out

    def close(self):
        if getattr(self, "h", None):
            hip().lmh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StaticCondensation:
    """Static condensation of a GpuProblem's internal-node DoFs on the device
    (StaticCondensation, include/numerics/static_condensation.h:64): the
    reduced matrix (get_condensed_mat), forward elimination, back substitution
    and apply() as one solve.  The reduced DoFs are the uncondensed ones in
    ascending full order."""

    def __init__(self, problem, cond_dofs=None):
        self.problem = problem
        cd = problem.condensed_dofs() if cond_dofs is None else np.ascontiguousarray(cond_dofs, dtype=np.int32)
        self.h = _P()
        check(hip().lmh_condensation_create(problem.matrix, cd.size, _ptr(cd) if cd.size else None, C.byref(self.h)))
        self.matrix = _P(hip().lmh_condensation_reduced_matrix(self.h))
        self.n_condensed = hip().lmh_condensation_n_condensed(self.h)
        self.n_reduced = hip().lmh_matrix_rows(self.matrix)
        self.nnz = hip().lmh_matrix_nnz(self.matrix)

    def close(self):
        """Frees the device data (the full matrix may be gone already); a no-op once the context is gone."""
        if self.h.value and self.problem.ctx.h:
            check(hip().lmh_condensation_destroy(self.h))
        self.h = _P()

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None:
                self.close()
        except Exception:
            pass

    def setup(self):
        check(hip().lmh_condensation_setup(self.h))

    def reduced_dofs(self):
        out = np.empty(self.n_reduced, dtype=np.int32)
        check(hip().lmh_condensation_reduced_dofs(self.h, _ptr(out)))
        return out

    def pattern(self):
        rp = np.empty(self.n_reduced + 1, dtype=np.int32)
        ci = np.empty(self.nnz, dtype=np.int32)
        check(hip().lmh_matrix_get_pattern(self.matrix, _ptr(rp), _ptr(ci)))
        return rp, ci

    def values(self):
        v = np.empty(self.nnz)
        check(hip().lmh_matrix_get_values(self.matrix, _ptr(v)))
        return v

    def forward(self, b_d, g_d):
        check(hip().lmh_condensation_forward(self.h, b_d.ptr, g_d.ptr))

    def backward(self, b_d, y_d, x_d):
        check(hip().lmh_condensation_backward(self.h, b_d.ptr, y_d.ptr, x_d.ptr))

    def solve(self, b_d, x_d, rtol, maxits=5000):
        its, res, reason = _I64(), _D(), _I()
        check(hip().lmh_condensation_solve(self.h, b_d.ptr, x_d.ptr, rtol, maxits, C.byref(its), C.byref(res),
                                           C.byref(reason)))
        return its.value, res.value, reason.value
