"""Row-block SpMV paths that finite-element patterns rarely reach, on matrices
built through the C ABI (lmh_matrix_create_csr):

- blocks of 256 or more short rows (3-7 entries), whose row pointers and epilogue values are
  read from global memory instead of LDS;
- a row longer than one row block (RB_CAP = 2048 entries), reduced on its own;
- the same two through the FP32 x-window kernel (compression on) and the plain
  int32-column FP32 kernel, which must agree bit for bit;
- CG and the mixed solver on a 1-D Laplacian (3 entries per row, ~680 rows per
  block), against the oracle's Eigen restatement.

Bars: FP64 SpMV and residual within 1e-14 of scipy (normwise relative), FP32
within 2e-6; CG iterations within 2 of the oracle's, solutions within 1e-9.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import libmesh_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make(ctx, A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(A.indices, dtype=np.int32)
    h = C.c_void_p()
    L.check(L.hip().lmh_matrix_create_csr(ctx.h, A.shape[0], L._ptr(rp), L._ptr(ci), C.byref(h)))
    v = np.ascontiguousarray(A.data, dtype=np.float64)
    L.check(L.hip().lmh_matrix_set_values(h, L._ptr(v)))
    return h, A, rp, ci, v


def shaped(rng, n, long_row):
    """Rows 0..1499: 3-7 entries within +-3 of the diagonal (blocks of ~340
    rows, a few hundred distinct columns); row 1500 (if long_row): 5000
    entries; the rest: up to 41 entries within +-30 of the diagonal."""
    rows, cols = [], []
    for i in range(n):
        if i < 1500:
            c = {i} | set(np.clip(i + rng.integers(-3, 4, 6), 0, n - 1).tolist())
        elif i == 1500 and long_row:
            c = set(range(0, 5000))
        else:
            c = {i} | set(np.clip(i + rng.integers(-30, 31, 40), 0, n - 1).tolist())
        rows += [i] * len(c)
        cols += sorted(c)
    vals = rng.standard_normal(len(rows))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


@pytest.mark.parametrize("long_row", [True, False])
def test_spmv_short_and_long_rows(ctx, long_row):
    rng = np.random.default_rng(11)
    n = 6000
    h, A, rp, ci, v = make(ctx, shaped(rng, n, long_row))
    try:
        print(f"n {n} nnz {A.nnz} longest row {np.diff(rp).max()}", flush=True)
        xh = rng.standard_normal(n)
        x = L.DeviceArray(ctx, n, data=xh)
        y = L.DeviceArray(ctx, n)
        L.check(L.hip().lmh_spmv(h, x.ptr, y.ptr))
        assert rel(y.download(), A @ xh) < 1e-14
        bh = rng.standard_normal(n)
        b = L.DeviceArray(ctx, n, data=bh)
        r = L.DeviceArray(ctx, n)
        L.check(L.hip().lmh_residual(h, b.ptr, x.ptr, r.ptr))
        assert rel(r.download(), bh - A @ xh) < 1e-14
        # FP32 operator: x windows (when every block fits) against int32 columns
        L.check(L.hip().lmh_matrix_make_fp32(h))
        x32 = L.DeviceArray(ctx, n, np.float32, data=xh.astype(np.float32))
        y1 = L.DeviceArray(ctx, n, np.float32)
        y2 = L.DeviceArray(ctx, n, np.float32)
        L.check(L.hip().lmh_matrix_set_fp32_compression(h, 1))
        windows = L.hip().lmh_matrix_fp32_compression(h)
        assert windows == (0 if long_row else 1)  # a row past one block cannot use windows
        L.check(L.hip().lmh_spmv_fp32(h, x32.ptr, y1.ptr))
        L.check(L.hip().lmh_matrix_set_fp32_compression(h, 0))
        L.check(L.hip().lmh_spmv_fp32(h, x32.ptr, y2.ptr))
        assert np.array_equal(y1.download(), y2.download())
        A32 = sp.csr_matrix((v.astype(np.float32).astype(np.float64), ci, rp), shape=(n, n))
        assert rel(y1.download().astype(np.float64), A32 @ xh.astype(np.float32).astype(np.float64)) < 2e-6
    finally:
        L.hip().lmh_matrix_destroy(h)


def test_cg_and_mixed_on_short_rows(ctx):
    n = 100000
    A = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n) + 1e-3, -np.ones(n - 1)], [-1, 0, 1], format="csr")
    h, A, rp, ci, v = make(ctx, A)
    try:
        rng = np.random.default_rng(5)
        bh = rng.standard_normal(n)
        ox, oits, _ = O.cg_eigen(rp, ci, v, bh, 1e-10, 20000)
        b = L.DeviceArray(ctx, n, data=bh)
        x = L.DeviceArray(ctx, n)
        x.zero()
        its, res, reason = C.c_int64(), C.c_double(), C.c_int()
        L.check(L.hip().lmh_cg_solve(h, b.ptr, x.ptr, 1e-10, 20000, C.byref(its), C.byref(res), C.byref(reason)))
        print(f"cg its {its.value} oracle {oits} res {res.value:.3e}", flush=True)
        assert reason.value == L.CONVERGED_RTOL and abs(its.value - oits) <= 2
        assert rel(x.download(), ox) < 1e-9
        L.check(L.hip().lmh_matrix_make_fp32(h))
        x2 = L.DeviceArray(ctx, n)
        x2.zero()
        stats = np.zeros(3)
        L.check(L.hip().lmh_mixed_solve(h, b.ptr, x2.ptr, 1e-10, 0.1, 0, 40000, C.byref(its), C.byref(res),
                                        C.byref(reason), L._ptr(stats)))
        print(f"mixed its {its.value} refreshes {stats[1]} res {res.value:.3e}", flush=True)
        assert reason.value == L.CONVERGED_RTOL and res.value <= 1e-10
        assert rel(x2.download(), ox) < 1e-7
    finally:
        L.hip().lmh_matrix_destroy(h)


def test_transpose_solve_needs_symmetric_pattern(ctx):
    """lmh_bicgstab_solve_transpose refuses a pattern that lacks a mirror entry."""
    n = 64
    A = sp.diags([np.ones(n - 1), 4.0 * np.ones(n)], [1, 0], format="csr")  # upper bidiagonal
    h, A, rp, ci, v = make(ctx, A)
    try:
        b = L.DeviceArray(ctx, n, data=np.ones(n))
        x = L.DeviceArray(ctx, n)
        x.zero()
        its, res, reason = C.c_int64(), C.c_double(), C.c_int()
        with pytest.raises(L.LmhError, match="structurally symmetric"):
            L.check(L.hip().lmh_bicgstab_solve_transpose(h, b.ptr, x.ptr, 1e-10, 100, C.byref(its), C.byref(res),
                                                         C.byref(reason)))
        # the forward solve of the same matrix is unaffected
        L.check(L.hip().lmh_bicgstab_solve(h, b.ptr, x.ptr, 1e-12, 100, C.byref(its), C.byref(res),
                                           C.byref(reason)))
        assert reason.value == L.CONVERGED_RTOL
        assert rel(x.download(), np.linalg.solve(A.toarray(), np.ones(n))) < 1e-10
    finally:
        L.hip().lmh_matrix_destroy(h)
