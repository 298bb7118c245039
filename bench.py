"""Benchmark: assembly + Jacobi-CG to rtol 1e-10 of the variable-coefficient Q2
Poisson problem on one MI355X, plain FP64 path and mixed-precision path.

Workload (one "step"): 64^3 HEX27 cube, Q2 Lagrange, 3x3x3 Gauss, 2,146,689
DoFs, kappa_e = exp(N(0,1)) per element, interior nodes perturbed by +-0.2 h
(h = node spacing), f = 1, u = 0 on all faces.  A step zeroes the solution,
assembles the operator and right-hand side on the GPU (geometry, element
contraction, constraints, CSR scatter) and solves to |b - A x| / |b| <= 1e-10.
Mesh generation, DoF numbering and the GPU sparsity pattern are setup
(EquationSystems::init), done once before timing.

value = DoFs / (t_assemble + t_solve) / 1e6 of the mixed-precision path
(FP32 inner CG with FP64 reliable-update refinement); the FP64 path is timed the
same way and reported beside it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mdof/s assembly + CG time to 1e-10 (mixed vs FP64), Q2 hex 2.1M dof, 1 GPU"
SEED_PERTURB = 0x5EED0001
SEED_KAPPA = 0x5EED0002
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# hardware-counter summary of the dominant kernel at the 64^3 workload
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_it_spmv16_hex27_64cube_pmc.txt")


def sig(v, digits=6):
    """Floats to `digits` significant digits (residuals like 9.7e-11 keep their value)."""
    if isinstance(v, float):
        return float(f"{v:.{digits}g}")
    if isinstance(v, dict):
        return {k: sig(x, digits) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [sig(x, digits) for x in v]
    return v


def spmv_bytes(nnz, n, val_bytes):
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row_ptr + x read + y write."""
    return (val_bytes + 4) * nnz + 4 * (n + 1) + 2 * val_bytes * n


def spmv_symbol(L, g, fp64):
    """The CG iteration's SpMV kernel for the matrix's column storage (x
    windows with 8- or 16-bit indices, or int32 columns)."""
    bits = L.hip().lmh_matrix_window_bits(g.matrix)
    if bits == 8:
        return f"lmh::k_it_spmv16{'d' if fp64 else ''}<1024, 1, 128>"
    if bits == 16:
        return f"lmh::k_it_spmv16{'d' if fp64 else ''}<2048, 2, 256>"
    return f"lmh::k_it_spmv<{'double' if fp64 else 'float'}>"


def scatter_bytes(nnz, n, n_elem, val_bytes):
    """Compulsory bytes of the assembly scatter: write each CSR value once, read
    col_idx and row_ptr, connectivity, unique coordinates, kappa, write rhs."""
    return val_bytes * nnz + 4 * nnz + 4 * (n + 1) + 4 * 27 * n_elem + 24 * n + 8 * n_elem + 8 * n


def copy_calibration(L, ctx, mib):
    """Time the library's plain copy kernel on `mib` MiB (also the known-size
    kernel the hardware-counter traffic is scaled by)."""
    n = mib * (1 << 20) // 8
    a = L.DeviceArray(ctx, n)
    b = L.DeviceArray(ctx, n)
    a.zero()
    b.copy_from(a)
    ctx.sync()
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(5):
        b.copy_from(a)
    ctx.sync()
    ctx.profile(False)
    ms, c = ctx.profile_query("copy")
    a.free()
    b.free()
    return {"bytes_read": 8 * n, "bytes_written": 8 * n, "avg_ms": round(ms, 5), "launches": c,
            "gbs": round(16 * n / (ms * 1e-3) / 1e9, 1)}


def read_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (FETCH_SIZE
    and WRITE_SIZE per dispatch, each scaled by the copy kernel's reading)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def periodic_x_rows(p, n):
    """Rows u(x=1) = u(x=0) for the unconstrained DoFs of the x = 1 face of a
    build_cube mesh with Dirichlet rows on the y and z faces (boundary nodes
    sit on the lattice, so (y, z) * 2n are integers)."""
    nd = np.full(p.n_nodes, -1, dtype=np.int64)
    nd[p.conn.ravel()] = p.elem_dofs[:, :p.npe].ravel()
    X = p.coords
    key = np.rint(X[:, 1:] * 2 * n).astype(np.int64)
    k1 = key[:, 0] * (2 * n + 1) + key[:, 1]
    near = np.abs(X[:, 0]) < 1e-12
    far = np.abs(X[:, 0] - 1.0) < 1e-12
    lut = np.full((2 * n + 1) ** 2, -1, dtype=np.int64)
    lut[k1[near]] = nd[near]
    fd = nd[far]
    keep = p.constrained[fd] == 0
    return [(int(d), {int(c): 1.0}, 0.0) for d, c in zip(fd[keep], lut[k1[far][keep]])]


def bench_constrained(args, L, ctx, kappa, log):
    """The same problem with general constraint rows: Dirichlet on the y and z
    faces, u(x=1) = u(x=0), symmetric constraint rows; assembly kernels and a
    mixed solve followed by enforce_constraints_exactly."""
    p = L.Problem(args.n, L.HEX27, 1, 1 | 2 | 8 | 32)
    p.perturb(SEED_PERTURB, 0.2)
    rows = periodic_x_rows(p, args.n)
    g = L.GpuProblem.from_problem(ctx, p)
    t0 = time.perf_counter()
    g.set_constraints(rows, error_on_loop=True, symmetric_rows=True)
    ctx.sync()
    t_set = time.perf_counter() - t0
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    x = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(3, kd, 0, b, L.FP64)
    ctx.sync()
    reps = 3
    T0 = time.perf_counter()
    for _ in range(reps):
        g.assemble_poisson(3, kd, 0, b, L.FP64)
    ctx.sync()
    t_asm = (time.perf_counter() - T0) / reps
    ctx.profile_reset()
    ctx.profile(True)
    g.assemble_poisson(3, kd, 0, b, L.FP64)
    ctx.sync()
    ctx.profile(False)
    kern = {}
    for k in ("element_hex27_mfma_fp64", "constrain_general_fp64", "scatter_gather_fp64", "scatter_expanded_fp64"):
        ms, c = ctx.profile_query(k)
        kern[k] = round(ms * c, 4)
    g.assemble_poisson(3, kd, 0, None, L.FP32)
    x.zero()
    ctx.sync()
    T0 = time.perf_counter()
    its, res, reason, st = g.mixed(b, x, args.rtol, args.inner, 0, 20000)
    g.enforce_constraints(x)
    ctx.sync()
    t_solve = time.perf_counter() - T0
    if reason != L.CONVERGED_RTOL:
        raise RuntimeError(f"constrained solve did not converge: its={its} reason={reason}")
    # the default asymmetric constraint rows: a nonsymmetric system, Jacobi-BiCGSTAB
    g.set_constraints(rows, error_on_loop=True, symmetric_rows=False)
    g.assemble_poisson(3, kd, 0, b, L.FP64)
    x.zero()
    ctx.sync()
    T0 = time.perf_counter()
    b_its, b_res, b_reason = g.bicgstab(b, x, args.rtol, 20000)
    ctx.sync()
    t_bcg = time.perf_counter() - T0
    if b_reason != L.CONVERGED_RTOL:
        raise RuntimeError(f"BiCGSTAB did not converge: its={b_its} reason={b_reason}")
    out = {
        "rows": len(rows),
        "general_elements": g.n_general_elements(),
        "asymmetric_rows_bicgstab_ms": round(t_bcg * 1e3, 3),
        "asymmetric_rows_bicgstab_iterations": int(b_its),
        "nnz": g.nnz,
        "set_constraints_s": round(t_set, 3),
        "assembly_fp64_ms": round(t_asm * 1e3, 3),
        "assembly_kernels_ms": kern,
        "mixed_solve_ms": round(t_solve * 1e3, 3),
        "mixed_iterations": int(its),
        "final_rel_residual": res,
        "mdofs": round(p.n_dofs / (t_asm + t_solve) / 1e6, 4),
    }
    g.close()
    return out


def bench_condensed(args, L, ctx, kappa, log):
    """Static condensation of the HEX27 centre-node DoFs on the same problem:
    reduced operator set-up (Schur complement values), then Jacobi-CG and the
    mixed solver on the reduced system with forward elimination and back
    substitution, beside Jacobi-CG on the full system in the same run."""
    p = L.Problem(args.n, L.HEX27, 1, L.ALL_FACES)
    p.perturb(SEED_PERTURB, 0.2)
    g = L.GpuProblem.from_problem(ctx, p)
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    x = L.DeviceArray(ctx, p.n_dofs)
    g.assemble_poisson(3, kd, 0, b, L.FP64)
    ctx.sync()
    T0 = time.perf_counter()
    sc = L.StaticCondensation(g)
    t_create = time.perf_counter() - T0
    reps = 3
    T0 = time.perf_counter()
    for _ in range(reps):
        sc.setup()
    t_setup = (time.perf_counter() - T0) / reps
    ctx.profile_reset()
    ctx.profile(True)
    sc.setup()
    ctx.sync()
    ctx.profile(False)
    ms, c = ctx.profile_query("sc_values")
    k_values = round(ms * c, 4)
    x.zero()
    ctx.sync()
    T0 = time.perf_counter()
    f_its, _, f_reason = g.cg(b, x, args.rtol, 20000)
    ctx.sync()
    t_full = time.perf_counter() - T0
    xf = x.download()
    x.zero()
    ctx.sync()
    T0 = time.perf_counter()
    its, res, reason = sc.solve(b, x, args.rtol, 20000)
    ctx.sync()
    t_cond = time.perf_counter() - T0
    if reason != L.CONVERGED_RTOL or f_reason != L.CONVERGED_RTOL:
        raise RuntimeError(f"condensed solve did not converge: its={its} reason={reason} full={f_reason}")
    diff = float(np.abs(x.download() - xf).max() / np.abs(xf).max())
    # mixed solver on the reduced operator
    L.check(L.hip().lmh_matrix_make_fp32(sc.matrix))
    gr = L.DeviceArray(ctx, sc.n_reduced)
    y = L.DeviceArray(ctx, sc.n_reduced)
    y.zero()
    m_its, m_res, m_reason = L.C.c_int64(), L.C.c_double(), L.C.c_int()
    stats = np.zeros(3)
    ctx.sync()
    T0 = time.perf_counter()
    sc.forward(b, gr)
    L.check(L.hip().lmh_mixed_solve(sc.matrix, gr.ptr, y.ptr, args.rtol, args.inner, 0, 20000, L.C.byref(m_its),
                                    L.C.byref(m_res), L.C.byref(m_reason), L._ptr(stats)))
    sc.backward(b, y, x)
    ctx.sync()
    t_mixed = time.perf_counter() - T0
    if m_reason.value != L.CONVERGED_RTOL:
        raise RuntimeError(f"condensed mixed solve did not converge: reason={m_reason.value}")
    diff_mixed = float(np.abs(x.download() - xf).max() / np.abs(xf).max())
    if diff_mixed > 1e-8 or diff > 1e-8:
        raise RuntimeError(f"condensed solutions differ from the full CG solution: cg {diff:.2e} mixed {diff_mixed:.2e}")
    out = {
        "condensed_dofs": int(sc.n_condensed),
        "reduced_rows": int(sc.n_reduced),
        "nnz_full": int(g.nnz),
        "nnz_reduced": int(sc.nnz),
        "create_host_s": round(t_create, 3),
        "setup_ms": round(t_setup * 1e3, 3),
        "sc_values_kernel_ms": k_values,
        "full_cg_ms": round(t_full * 1e3, 3),
        "full_cg_iterations": int(f_its),
        "condensed_cg_ms": round(t_cond * 1e3, 3),
        "condensed_cg_iterations": int(its),
        "condensed_mixed_ms": round(t_mixed * 1e3, 3),
        "condensed_mixed_iterations": int(m_its.value),
        "max_rel_diff_vs_full_cg": diff,
        "mixed_max_rel_diff_vs_full_cg": diff_mixed,
    }
    sc.close()
    g.close()
    return out


def host_description():
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return cpu_model, usable


def cpu_threads():
    """Threads of the all-core CPU runs: OMP_NUM_THREADS when set, else the CPUs this process may use."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env or host_description()[1]


class OracleRun:
    """The CPU reference path on the bench workload itself (n^3 HEX27, same
    seeds): the oracle's serial element loop (the restated libMesh assembly of
    introduction_ex4.C:313-560 with the kappa-weighted form) for the matrix and
    right-hand side, then the reference's own vendored Eigen 3.3.9
    ConjugateGradient + DiagonalPreconditioner (oracle/_ref, driven as
    EigenSparseLinearSolver::solve does) to the same rtol.  Mesh, DoF numbering
    and the sparsity pattern are setup and stay off the clock, as on the GPU."""

    def __init__(self, n):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        self.O = O
        t0 = time.perf_counter()
        m = O.Mesh(n, O.HEX27)
        m.perturb(SEED_PERTURB, 0.2)
        self.kappa = O.kappa_lognormal(m.n_elem, SEED_KAPPA)
        self.mesh = m
        self.sys = O.System(m)
        self.rp, self.ci = self.sys.pattern()
        self.t_setup = time.perf_counter() - t0
        self.n_dofs = self.sys.n_dofs
        if O.ref() is None:
            raise RuntimeError("oracle/_ref/libeigen_ref.so is not built (the reference tree was absent at build time)")

    def run(self, rtol, threads):
        t0 = time.perf_counter()
        vals, rhs = self.sys.assemble_poisson(3, self.kappa, 0)
        t_asm = time.perf_counter() - t0
        x, its, err, t_solve = self.O.ref_cg_eigen_timed(self.rp, self.ci, vals, rhs, rtol, 20000, threads)
        return {"t_assemble_s": t_asm, "t_solve_s": t_solve, "iterations": its, "rel_residual": err,
                "x": x, "vals": vals, "rhs": rhs}


def cpu_baseline(orc, rtol, log):
    """One run of the CPU reference path on the full workload, all usable cores
    for Eigen's threaded SpMV (the element loop is serial, as in the reference's
    example).  Returns the baseline record and the run (its solution is the FP64
    oracle solution the GPU results are compared with)."""
    threads = cpu_threads()
    model, usable = host_description()
    r = orc.run(rtol, threads)
    log(f"cpu baseline: assembly {r['t_assemble_s']:.2f} s, Eigen CG {r['iterations']} its {r['t_solve_s']:.2f} s "
        f"({threads} threads)")
    rec = {
        "value": orc.n_dofs / (r["t_assemble_s"] + r["t_solve_s"]) / 1e6,
        "unit": "Mdof/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"the full workload ({orc.n_dofs} dofs), one run: assembly by the oracle's serial element loop "
                  f"{r['t_assemble_s']:.2f} s, then the reference's vendored Eigen 3.3.9 Jacobi-CG (oracle/_ref) "
                  f"{r['iterations']} its to rtol {rtol:g} in {r['t_solve_s']:.2f} s on {threads} OpenMP threads "
                  f"(threaded SpMV, serial vector ops); mesh/DoF/pattern setup {orc.t_setup:.2f} s excluded; "
                  f"host {model}, {usable} usable CPUs",
        "t_assemble_s": r["t_assemble_s"],
        "t_solve_s": r["t_solve_s"],
        "iterations": r["iterations"],
        "rel_residual": r["rel_residual"],
    }
    return rec, r


def cpu_full(args, log):
    """Per-phase medians of the CPU reference path at 1 thread and at all
    usable cores (1 warm-up, then `runs` timed runs each); one JSON line."""
    orc = OracleRun(args.n)
    model, usable = host_description()
    allc = cpu_threads()
    out = {"workload": f"{args.n}^3 HEX27 C1 recipe ({orc.n_dofs} dofs)", "host": model, "usable_cpus": usable,
           "setup_s": orc.t_setup, "rtol": args.rtol, "runs": args.cpu_runs, "threads": {}}
    for th in sorted({1, allc}):
        orc.run(args.rtol, th)  # warm-up
        rs = []
        for k in range(args.cpu_runs):
            r = orc.run(args.rtol, th)
            rs.append((r["t_assemble_s"], r["t_solve_s"], r["iterations"]))
            log(f"threads {th} run {k}: assembly {r['t_assemble_s']:.3f} s solve {r['t_solve_s']:.3f} s "
                f"its {r['iterations']}")
        ta = float(np.median([a for a, _, _ in rs]))
        ts = float(np.median([b for _, b, _ in rs]))
        out["threads"][str(th)] = {"assembly_s": [a for a, _, _ in rs], "solve_s": [b for _, b, _ in rs],
                                   "median_assembly_s": ta, "median_solve_s": ts, "iterations": rs[-1][2],
                                   "mdofs": orc.n_dofs / (ta + ts) / 1e6}
    print(json.dumps(sig(out)), flush=True)


SEED_BATCH = 0x5EED0003
CLAMP_X0 = 1 << 4  # boundary id 4: the x = 0 face
TYPE_NAME = {10: "HEX8", 12: "HEX27", 8: "TET4", 9: "TET10"}


def batch_cpu_baseline(nmin, nmax, count):
    """Oracle FP64 elasticity assembly (serial element loop) of the same seeded
    recipe with edges in [nmin, nmax]."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import libmesh_amd as L

    edge, kind = L.batch_recipe(SEED_BATCH, count, nmin, nmax)
    npts = {O.HEX8: 2, O.HEX27: 3, O.TET4: 3, O.TET10: 5}
    ne = 0
    t = 0.0
    for n, k in zip(edge, kind):
        m = O.Mesh(int(n), int(k))
        s = O.System(m, 3, CLAMP_X0)
        t0 = time.perf_counter()
        s.assemble_elasticity(npts[int(k)], 1.0, 0.3, -1.0)
        t += time.perf_counter() - t0
        ne += m.n_elem
    return {"value": ne / t, "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": f"same recipe with edges in [{nmin},{nmax}] ({count} problems, {ne} elements), oracle "
                      f"FP64 element loop + constraints + scatter, {t:.2f}s"}


def bench_batched(args, L, log, ctx, steps, warmup):
    """Batched FP32 elasticity assembly: `batch_count` problems (E=1, nu=0.3,
    body force (0,0,-1), clamped x=0 face), edge N log-uniform in [8, nmax],
    element type uniform over HEX8/HEX27/TET4/TET10 from a seed; one step = one
    hipGraph launch assembling every problem (element kernel + scatter).
    Returns the record (a bench line of its own with --workload batched)."""
    edge, kind = L.batch_recipe(SEED_BATCH, args.batch_count, 8, args.batch_nmax)
    t0 = time.perf_counter()
    gps, n_elem, n_dofs, n_nodes, npe = [], 0, 0, [], []
    for i, (n, k) in enumerate(zip(edge, kind)):
        p = L.Problem(int(n), int(k), 3, CLAMP_X0)
        gps.append(L.GpuProblem.from_problem(ctx, p, build_matrix=False))
        n_elem += p.n_elem
        n_dofs += p.n_dofs
        n_nodes.append(p.n_nodes)
        npe.append(p.npe)
        log(f"problem {i}: {TYPE_NAME[int(k)]} N={int(n)} elements={p.n_elem} dofs={p.n_dofs}")
        del p
    b = L.Batch(ctx, gps)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    nnz = [b.nnz(i) for i in range(len(gps))]
    log(f"setup {t_setup:.1f}s: {n_elem} elements, {n_dofs} dofs, nnz {sum(nnz)}")
    best = b.autotune(1.0, 0.3, -1.0, 2)
    lpe, pp = b.tiles()
    log(f"autotuned lanes/element {lpe.tolist()} passes/tile {pp.tolist()} best ms {best.round(3).tolist()}")
    for _ in range(warmup):
        b.assemble()
    ctx.sync()
    T0 = time.perf_counter()
    for _ in range(steps):
        b.assemble()
    ctx.sync()
    T1 = time.perf_counter()
    # the same launches again with HIP events around each graph launch
    ctx.profile_reset()
    ctx.profile(True)
    for _ in range(steps):
        b.assemble()
    ctx.sync()
    ctx.profile(False)
    gms, gcnt = ctx.profile_query("batch_graph")
    # compulsory bytes: FP32 values written once, col_idx and row_ptr read,
    # connectivity, unique coordinates, FP64 rhs written
    comp = 0
    for i, g in enumerate(gps):
        comp += 4 * nnz[i] + 4 * nnz[i] + 4 * (g.n_dofs + 1) + 4 * npe[i] * g.n_elem + 24 * n_nodes[i] + 8 * g.n_dofs
    gbs = comp / (gms * 1e-3) / 1e9
    types = {TYPE_NAME[t]: int(sum(1 for k in kind if k == t)) for t in (10, 12, 8, 9)}
    out = {
        "metric": "batched FP32 elasticity assembly throughput (mixed HEX8/HEX27/TET4/TET10), elements/s",
        "value": round(n_elem / (gms * 1e-3), 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round((T1 - T0) / steps * 1e3, 3),
        "graph_ms": round(gms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.batch_count} linear-elasticity problems (E=1, nu=0.3, body force (0,0,-1), clamped "
                        f"x=0 face), edge N log-uniform in [8,{args.batch_nmax}], element type uniform over "
                        f"HEX8/HEX27/TET4/TET10, seed {SEED_BATCH:#x}; one hipGraph launch per step",
            "problems_per_type": types,
            "elements": n_elem,
            "dofs": n_dofs,
            "nnz": int(sum(nnz)),
            "edges": [int(n) for n in edge],
        },
        "autotune": {"lanes_per_element": dict(zip(["HEX8", "HEX27", "TET4", "TET10"], lpe.tolist())),
                     "passes_per_tile": dict(zip(["HEX8", "HEX27", "TET4", "TET10"], pp.tolist())),
                     "element_kernel_ms": dict(zip(["HEX8", "HEX27", "TET4", "TET10"], best.round(4).tolist()))},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "compulsory_bytes": int(comp),
                     "kernel": "batch graph (k_batch_elasticity + k_batch_scatter)"},
        "setup_s": round(t_setup, 2),
        "mdofs": round(n_dofs / (gms * 1e-3) / 1e6, 3),
    }
    if args.cpu != "skip":
        log("batched cpu baseline ...")
        out["cpu_baseline"] = batch_cpu_baseline(4, 8, 32)
    b.close()
    for g in gps:
        g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="elements per cube edge")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--inner", type=float, default=0.1, help="reliable-update trigger of the mixed solver")
    ap.add_argument("--scatter", choices=["gather", "colored"], default="gather")
    ap.add_argument("--cpu", choices=["run", "skip"], default="run",
                    help="run: time the CPU reference path once on the full workload (all usable cores) and "
                         "compare every GPU solution with its FP64 solution")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only the CPU reference path: per-phase medians of --cpu-runs runs at 1 thread and at all "
                         "usable cores (no GPU)")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--batched", choices=["run", "skip"], default="run",
                    help="run: add the batched elasticity block (64 problems, one hipGraph launch per step)")
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the copy-kernel calibration (0: skip)")
    ap.add_argument("--workload", choices=["poisson", "batched"], default="poisson",
                    help="poisson: 64^3 HEX27 assembly + CG (the headline); batched: 64 elasticity problems")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-nmax", type=int, default=48)
    ap.add_argument("--quiet", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import libmesh_amd as L

    log = (lambda *a: None) if args.quiet else (lambda *a: print(*a, file=sys.stderr, flush=True))
    if args.cpu_full:
        cpu_full(args, log)
        return
    if args.workload == "batched":
        ctx = L.Context(0)
        print(json.dumps(sig(bench_batched(args, L, log, ctx, args.steps, args.warmup))), flush=True)
        return
    scatter = L.SCATTER_GATHER if args.scatter == "gather" else L.SCATTER_COLORED

    t0 = time.perf_counter()
    p = L.Problem(args.n, L.HEX27)
    p.perturb(SEED_PERTURB, 0.2)
    kappa = L.lognormal(p.n_elem, SEED_KAPPA)
    ctx = L.Context(0)
    g = L.GpuProblem.from_problem(ctx, p)
    kd = L.DeviceArray(ctx, p.n_elem, data=kappa)
    b = L.DeviceArray(ctx, p.n_dofs)
    x = L.DeviceArray(ctx, p.n_dofs)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    N, nnz = p.n_dofs, g.nnz
    log(f"setup {t_setup:.2f}s: dofs={N} nnz={nnz} elements={p.n_elem} colours={g.n_colors}")
    copy = copy_calibration(L, ctx, args.copy_mib) if args.copy_mib > 0 else None
    if copy:
        log(f"copy kernel: {copy['gbs']} GB/s ({args.copy_mib} MiB, avg {copy['avg_ms']} ms)")

    def step(mode):
        """mode fp64: FP64 assembly + FP64 CG.  mixed: FP64 assembly, FP32 inner
        operator rounded from it, mixed solve.  mixed_mfma32: FP64 assembly for
        the outer residual operator plus an FP32-MFMA assembly of the inner
        operator, mixed solve."""
        x.zero()
        ctx.sync()
        ta = time.perf_counter()
        g.assemble_poisson(3, kd, 0, b, L.FP64, scatter)
        if mode == "mixed":
            g.make_fp32()
        elif mode == "mixed_mfma32":
            g.assemble_poisson(3, kd, 0, b, L.FP32, scatter)
        ctx.sync()
        tb = time.perf_counter()
        if mode != "fp64":
            its, res, reason, stats = g.mixed(b, x, args.rtol, args.inner, 0, 20000)
            nref = int(stats[1])
        else:
            its, res, reason = g.cg(b, x, args.rtol, 20000)
            nref = 0
        ctx.sync()
        tc = time.perf_counter()
        if reason != L.CONVERGED_RTOL:
            raise RuntimeError(f"solve did not converge: its={its} res={res} reason={reason}")
        return tb - ta, tc - tb, its, res, nref

    results = {}
    for mode in ("fp64", "mixed_mfma32", "mixed"):
        mixed = mode != "fp64"
        for _ in range(args.warmup):
            step(mode)
        ctx.sync()
        T0 = time.perf_counter()
        rs = [step(mode) for _ in range(args.steps)]
        ctx.sync()
        T1 = time.perf_counter()
        # kernel timing: the same steps again with HIP events around every
        # launch on the solver's stream (the events add a few percent to a
        # step, so the throughput above is timed without them)
        ctx.profile_reset()
        ctx.profile(True)
        Tp0 = time.perf_counter()
        for _ in range(args.steps):
            step(mode)
        ctx.sync()
        Tp1 = time.perf_counter()
        ctx.profile(False)
        kname = "mx_spmv_fp32" if mixed else "cg_spmv_fp64"
        kavg, kcnt = ctx.profile_query(kname)
        ta = float(np.mean([r[0] for r in rs]))
        ts = float(np.mean([r[1] for r in rs]))
        results[mode] = {
            "ms_per_step": (T1 - T0) / args.steps * 1e3,
            "t_assemble_ms": ta * 1e3,
            "t_solve_ms": ts * 1e3,
            "iterations": rs[-1][2],
            "refreshes": rs[-1][4],
            "final_rel_residual": rs[-1][3],
            "mdofs": N / (ta + ts) / 1e6,
            "spmv_kernel": kname,
            "spmv_avg_ms": kavg,
            "spmv_launches": kcnt,
            "ms_per_step_with_kernel_events": (Tp1 - Tp0) / args.steps * 1e3,
            "x": x.download(),
        }
        log(f"{mode}: asm {ta*1e3:.2f} ms, solve {ts*1e3:.1f} ms, its {rs[-1][2]}, refreshes {rs[-1][4]}, "
            f"res {rs[-1][3]:.2e}, {N/(ta+ts)/1e6:.2f} Mdof/s, {kname} avg {kavg*1e3:.1f} us x{kcnt}")

    # adjoint sensitivities (all per-element kappa) of J = integral of u over
    # [0.25,0.75]^3 at the mixed forward solution: QoI derivative, adjoint solve
    # (mixed, same tolerance), backward kernel
    dq = L.DeviceArray(ctx, N)
    lam = L.DeviceArray(ctx, N)
    grad = L.DeviceArray(ctx, p.n_elem)
    lam.zero()
    ctx.sync()
    ctx.profile_reset()
    ctx.profile(True)
    ta0 = time.perf_counter()
    g.qoi_box_derivative(3, 0.25, 0.75, dq)
    a_its, a_res, a_reason, a_st = g.mixed(dq, lam, args.rtol, args.inner, 0, 20000)
    g.adjoint_kappa_gradient(3, kd, x, lam, grad)
    ctx.sync()
    ta1 = time.perf_counter()
    ctx.profile(False)
    gk_ms, _ = ctx.profile_query("adjoint_kappa")
    qd_ms, _ = ctx.profile_query("qoi_box")
    gh = grad.download()
    J = float(dq.download() @ results["mixed"]["x"])
    adjoint = {
        "gradients": int(p.n_elem),
        "total_ms": round((ta1 - ta0) * 1e3, 3),
        "qoi_derivative_ms": round(qd_ms, 4),
        "adjoint_solve_iterations": int(a_its),
        "backward_kernel_ms": round(gk_ms, 4),
        "backward_kernel_tflops_fp64": round(9.7e9 / (gk_ms * 1e-3) / 1e12, 3),
        "J": J,
        "homogeneity_rel_err": abs(float(kappa @ gh) + J) / abs(J),
    }
    log(f"adjoint: {adjoint}")

    # FEMSystem residual R(u) = sum_e C^T (K_e u_e - F_e) at the mixed solution
    R = L.DeviceArray(ctx, N)
    xs = L.DeviceArray(ctx, N, data=results["mixed"]["x"])
    g.residual_poisson(3, kd, 0, xs, R)
    ctx.sync()
    ctx.profile_reset()
    ctx.profile(True)
    g.residual_poisson(3, kd, 0, xs, R)
    ctx.sync()
    ctx.profile(False)
    fem_residual = {k: round(ctx.profile_query(k)[0] * max(ctx.profile_query(k)[1], 1), 4)
                    for k in ("element_hex27_mfma_fp64", "element_residual", "add_element_vectors")}
    bnorm = float(np.linalg.norm(b.download()))
    fem_residual["rel_residual_at_mixed_solution"] = float(np.linalg.norm(R.download()) / bnorm)
    log(f"fem residual: {fem_residual}")
    del R, xs

    constrained = bench_constrained(args, L, ctx, kappa, log)
    log(f"constrained: {constrained}")
    condensed = bench_condensed(args, L, ctx, kappa, log)
    log(f"condensed: {condensed}")

    # assembly kernel timings: mean per assembly over ASM_REPS profiled
    # assemblies (FP64 path, then the FP32 MFMA contraction + FP32 scatter,
    # which writes the FP32 operator only)
    ASM_REPS = 10
    asm_k = {}
    for prec, tag in ((L.FP64, "fp64"), (L.FP32, "fp32")):
        ctx.profile_reset()
        ctx.profile(True)
        for _ in range(ASM_REPS):
            g.assemble_poisson(3, kd, 0, b, prec, scatter)
        ctx.sync()
        ctx.profile(False)
        for k in ("element_hex27_mfma_", "scatter_gather_", "scatter_colored_"):
            ms, c = ctx.profile_query(k + tag)
            if c:
                asm_k[k + tag] = ms * c / ASM_REPS
        if prec == L.FP64:
            rhs_h = b.download()  # the FP64 right-hand side, compared with the oracle's below
    x64 = results["fp64"].pop("x")
    xmx = results["mixed"].pop("x")
    err_mixed_vs_fp64 = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))
    x32m = results["mixed_mfma32"].pop("x")
    err_mfma32_vs_fp64 = float(np.linalg.norm(x32m - x64) / np.linalg.norm(x64))

    mx = results["mixed"]

    def working_ms(r):
        # Launches queued after convergence exit at once (the host polls every
        # 4 iterations); average the kernel time over the launches that did work.
        return r["spmv_avg_ms"] * r["spmv_launches"] / (r["iterations"] * args.steps)

    alg = spmv_bytes(nnz, N, 4)
    achieved = alg / (working_ms(mx) * 1e-3) / 1e9
    alg64 = spmv_bytes(nnz, N, 8)
    ach64 = alg64 / (working_ms(results["fp64"]) * 1e-3) / 1e9
    flops_contract = 27 * 27 * 27 * 3 * 2 * p.n_elem
    traffic = read_traffic(TRAFFIC_SUMMARY) if args.n == 64 else None
    out = {
        "metric": METRIC,
        "value": round(mx["mdofs"], 4),
        "unit": "Mdof/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mx["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3 HEX27 Q2 Poisson -div(kappa grad u)=1, kappa=exp(N(0,1)) per element, "
                        f"interior nodes perturbed +-0.2h, u=0 on boundary, 3x3x3 Gauss, Jacobi-CG to rtol "
                        f"{args.rtol:g}; mixed = FP32 inner CG with FP64 reliable-update refinement",
            "dofs": N,
            "nnz": nnz,
            "elements": p.n_elem,
            "scatter": args.scatter,
            "inner_trigger": args.inner,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": mx["spmv_kernel"],
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_symbol": spmv_symbol(L, g, fp64=False),
            "algorithmic_bytes": alg,
            "avg_launch_ms": round(mx["spmv_avg_ms"], 5),
            "avg_working_launch_ms": round(working_ms(mx), 5),
            "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if traffic else None,
            # HBM bytes the counters saw per launch over the same launch time
            "traffic_frac": (round(traffic / (working_ms(mx) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             if traffic else None),
        },
        "fp64": results["fp64"],
        "mixed": results["mixed"],
        "mixed_mfma32": results["mixed_mfma32"],
        "speedup_mixed_mfma32_vs_fp64": round(results["mixed_mfma32"]["mdofs"] / results["fp64"]["mdofs"], 4),
        "mixed_mfma32_rel_diff_vs_fp64_solution": err_mfma32_vs_fp64,
        "speedup_mixed_vs_fp64": round(mx["mdofs"] / results["fp64"]["mdofs"], 4),
        "mixed_rel_diff_vs_fp64_solution": err_mixed_vs_fp64,
        "fp64_spmv_roofline": {"kernel_symbol": spmv_symbol(L, g, fp64=True), "achieved": round(ach64, 1),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach64 / HBM_PEAK_GBS, 4),
                               "algorithmic_bytes": alg64,
                               "avg_working_launch_ms": round(working_ms(results["fp64"]), 5)},
        "assembly_kernels_ms": {k: round(v, 4) for k, v in asm_k.items()},
        "adjoint": adjoint,
        "constrained": constrained,
        "condensed": condensed,
        "fem_residual_ms": fem_residual,
        "contraction_gflop": flops_contract / 1e9,
        "setup_s": round(t_setup, 3),
    }
    for prec in ("fp64", "fp32"):
        k = f"element_hex27_mfma_{prec}"
        if k in asm_k:
            tf = flops_contract / (asm_k[k] * 1e-3) / 1e12
            out[f"contraction_tflops_{prec}"] = round(tf, 3)
            out[f"contraction_frac_of_{prec}_peak"] = round(tf / MFMA_PEAK_TFLOPS[prec], 4)
    for prec, vb in (("fp64", 8), ("fp32", 4)):
        k = f"scatter_{'gather' if args.scatter == 'gather' else 'colored'}_{prec}"
        if k in asm_k:
            sb = scatter_bytes(nnz, N, p.n_elem, vb)
            gbs = sb / (asm_k[k] * 1e-3) / 1e9
            out[f"scatter_roofline_{prec}"] = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                               "frac": round(gbs / HBM_PEAK_GBS, 4), "compulsory_bytes": sb}
    if copy:
        out["copy_calibration"] = copy
    gvals = g.values() if args.cpu == "run" else None
    g.close()
    del kd, b, x
    if args.batched == "run":
        log("batched elasticity block ...")
        rec = bench_batched(args, L, log, ctx, args.steps, args.warmup)
        out["batched"] = {k: v for k, v in rec.items()
                          if k not in ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")}
    if args.cpu == "run":
        log("cpu baseline (reference path on the full workload) ...")
        orc = OracleRun(args.n)
        rec, r = cpu_baseline(orc, args.rtol, log)
        out["cpu_baseline"] = rec
        xo = r["x"]
        no = float(np.linalg.norm(xo))
        # error of every GPU result against the FP64 oracle: the oracle's own
        # assembly solved by the reference's Eigen CG to the same rtol
        out["rel_err_vs_fp64_oracle"] = {
            "fp64": float(np.linalg.norm(x64 - xo) / no),
            "mixed": float(np.linalg.norm(xmx - xo) / no),
            "mixed_mfma32": float(np.linalg.norm(x32m - xo) / no),
            "oracle_iterations": r["iterations"],
            "oracle_rel_residual": r["rel_residual"],
        }
        out["matrix_rel_diff_vs_oracle"] = float(np.abs(gvals - r["vals"]).max() / np.abs(r["vals"]).max())
        out["rhs_rel_diff_vs_oracle"] = float(np.abs(rhs_h - r["rhs"]).max() / np.abs(r["rhs"]).max())
        log(f"vs oracle: {out['rel_err_vs_fp64_oracle']} matrix {out['matrix_rel_diff_vs_oracle']:.2e}")
    print(json.dumps(sig(out)), flush=True)


if __name__ == "__main__":
    main()
