"""GPU parity of the batched FP32 elasticity assembly (lmh_batch_*) against the
CPU oracle's FP64 restatement of systems_of_equations_ex6.C's assembly.

Bars: every problem's CSR pattern bit-exact; FP32 matrix values within 2e-6
normwise relative of the FP64 oracle (FP32 contraction of FP64 geometry,
max|A - A_ref| / max|A_ref|), constrained diagonals exact integers; FP64
right-hand sides within 1e-12.  Tile shapes change only the work split, not the
arithmetic of any entry, so every tile shape gives bit-identical values; graph
replays give bit-identical values; E enters linearly, so doubling E doubles
every unconstrained entry exactly.
"""
import numpy as np
import pytest

import libmesh_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

NPTS = {L.HEX8: 2, L.HEX27: 3, L.TET4: 3, L.TET10: 5}  # QGauss at default_quadrature_order
CLAMP = 1 << 4  # boundary id 4: the x = 0 face
SEED = 0x5EED0004


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def build(ctx, specs):
    probs, sys_ = [], []
    for k, (n, t, pert) in enumerate(specs):
        p = L.Problem(n, t, 3, CLAMP)
        m = O.Mesh(n, t)
        if pert:
            frac = 0.2 if t in (L.HEX8, L.HEX27) else 0.1
            p.perturb(SEED + k, frac)
            m.perturb(SEED + k, frac)
        probs.append(p)
        sys_.append((m, O.System(m, 3, CLAMP)))
    gps = [L.GpuProblem.from_problem(ctx, p, build_matrix=False) for p in probs]
    return probs, sys_, gps


SPECS = [(3, L.HEX8, False), (2, L.HEX27, True), (2, L.TET4, False), (2, L.TET10, True),
         (4, L.HEX8, True), (3, L.HEX27, False), (3, L.TET4, True), (1, L.TET10, False)]


def test_batch_matches_oracle(ctx):
    probs, sys_, gps = build(ctx, SPECS)
    b = L.Batch(ctx, gps)
    b.assemble(1.0, 0.3, -1.0)
    for k, (p, (m, s)) in enumerate(zip(probs, sys_)):
        rp, ci = b.pattern(k)
        orp, oci = s.pattern()
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci), f"problem {k}: pattern"
        vals, rhs = s.assemble_elasticity(NPTS[p.elem_type], 1.0, 0.3, -1.0)
        v = b.values(k).astype(np.float64)
        assert rel(v, vals) < 2e-6, f"problem {k} ({SPECS[k]}): values {rel(v, vals):.3e}"
        assert rel(b.rhs(k), rhs) < 1e-12, f"problem {k}: rhs"
        cons = s.constrained().astype(bool)
        diag_ref = vals[[orp[r] + np.searchsorted(oci[orp[r]:orp[r + 1]], r) for r in np.nonzero(cons)[0]]]
        diag = v[[rp[r] + np.searchsorted(ci[rp[r]:rp[r + 1]], r) for r in np.nonzero(cons)[0]]]
        assert np.array_equal(diag, diag_ref), f"problem {k}: constrained diagonals"
    b.close()


def test_batch_tiles_and_graph_replays_bit_identical(ctx):
    probs, sys_, gps = build(ctx, SPECS)
    b = L.Batch(ctx, gps)
    b.assemble()
    ref = [b.values(k) for k in range(len(SPECS))]
    rref = [b.rhs(k) for k in range(len(SPECS))]
    lpe0, pp0 = b.tiles()
    # every admissible tile shape: same bits
    for shape in ([32, 256, 32, 64], [64, 256, 64, 256], [256, 256, 256, 256]):
        b.set_tiles(shape, [1, 2, 16, 3])
        b.assemble()
        for k in range(len(SPECS)):
            assert np.array_equal(b.values(k), ref[k]), f"tiles {shape}: problem {k}"
    best = b.autotune(1.0, 0.3, -1.0, reps=1)
    assert all((ms > 0) == (t in [s[1] for s in SPECS]) for ms, t in zip(best, L.Batch.TYPES))
    # replays of the captured graph with eager work in between
    x = L.DeviceArray(ctx, probs[0].n_dofs, np.float32, data=np.ones(probs[0].n_dofs, np.float32))
    y = L.DeviceArray(ctx, probs[0].n_dofs, np.float32)
    for _ in range(3):
        b.assemble()
        L.check(L.hip().lmh_spmv_fp32(b.matrices[0], x.ptr, y.ptr))
        ctx.sync()
        for k in range(len(SPECS)):
            assert np.array_equal(b.values(k), ref[k])
            assert np.array_equal(b.rhs(k), rref[k])
    # E enters linearly: doubling E doubles every unconstrained entry exactly
    b.assemble(2.0, 0.3, -1.0)
    for k, (m, s) in enumerate(sys_):
        rp, ci = b.pattern(k)
        cons = s.constrained().astype(bool)
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        free = ~(cons[rows] | cons[ci])
        v2 = b.values(k)
        assert np.array_equal(v2[free], 2 * ref[k][free])
        assert np.array_equal(v2[~free], ref[k][~free])
    b.close()


def test_batch_recipe_small(ctx):
    """The seeded batch recipe (log-uniform edges, uniform element types) at small
    edges: all problems assembled in one graph launch match the oracle."""
    edge, kind = L.batch_recipe(SEED, 12, 1, 3)
    assert set(kind.tolist()) <= {L.HEX8, L.HEX27, L.TET4, L.TET10}
    specs = [(int(n), int(t), False) for n, t in zip(edge, kind)]
    probs, sys_, gps = build(ctx, specs)
    b = L.Batch(ctx, gps)
    b.autotune()
    b.assemble()
    for k, (p, (m, s)) in enumerate(zip(probs, sys_)):
        vals, rhs = s.assemble_elasticity(NPTS[p.elem_type], 1.0, 0.3, -1.0)
        assert rel(b.values(k).astype(np.float64), vals) < 2e-6
        assert rel(b.rhs(k), rhs) < 1e-12
    b.close()


def test_batch_degenerate_map_raises(ctx):
    p = L.Problem(2, L.HEX27, 3, CLAMP)
    gp_ok = L.GpuProblem.from_problem(ctx, L.Problem(2, L.HEX8, 3, CLAMP), build_matrix=False)
    # push the centre node of element 0 through its +x face: det J < 0
    c = p.conn[0, 26]
    xyz = p.coords.copy()
    xyz[c, 0] += 1.5 / 2
    gp = L.GpuProblem(ctx, xyz, p.conn, p.elem_type, p.elem_dofs, p.n_dofs, p.constrained, 3, build_matrix=False)
    b = L.Batch(ctx, [gp_ok, gp])
    with pytest.raises(L.LmhError, match="degenerate"):
        b.assemble()
    b.close()


def _digest(a):
    import xxhash

    return xxhash.xxh3_64_hexdigest(np.ascontiguousarray(a))


def test_batch_recipe_full_size(ctx):
    """The batched configuration at its real size: 64 problems, edges
    log-uniform in [8, 48], element types uniform over HEX8/HEX27/TET4/TET10
    (about 24 M elements, 43 M DoFs, 3.4 G nonzeros), assembled by one graph
    launch.  On all 64 problems: hex patterns have the closed-form nnz
    9 (3n+1)^3 (HEX8) / 9 (8n+1)^3 (HEX27); every constrained row holds one
    nonzero, a positive integer (the number of incident elements); values are
    bit-identical across graph replays with eager work between them and across
    two tile shapes; doubling E doubles every entry of the unconstrained rows
    exactly (constrained columns there are exact zeros) and leaves constrained
    rows unchanged.  The smallest problem of each element type is compared with
    the oracle (values 2e-6, rhs 1e-12)."""
    import time

    t0 = time.perf_counter()
    edge, kind = L.batch_recipe(0x5EED0003, 64, 8, 48)
    gps, cons, info = [], [], []
    for n, t in zip(edge.tolist(), kind.tolist()):
        p = L.Problem(n, t, 3, CLAMP)
        gps.append(L.GpuProblem.from_problem(ctx, p, build_matrix=False))
        cons.append(p.constrained.astype(bool))
        info.append((n, t))
        del p
    b = L.Batch(ctx, gps)
    print(f"built 64 problems in {time.perf_counter() - t0:.1f} s, nnz {sum(b.nnz(k) for k in range(64))}",
          flush=True)
    b.autotune(1.0, 0.3, -1.0, reps=1)
    lpe, pp = b.tiles()
    b.assemble()
    digests, rdig = [], []
    for k, (n, t) in enumerate(info):
        if t == L.HEX8:
            assert b.nnz(k) == 9 * (3 * n + 1) ** 3, k
        elif t == L.HEX27:
            assert b.nnz(k) == 9 * (8 * n + 1) ** 3, k
        v = b.values(k)
        digests.append(_digest(v))
        rdig.append(_digest(b.rhs(k)))
    print(f"assembled, digests taken ({time.perf_counter() - t0:.1f} s)", flush=True)
    # a second tile shape per element type: twice the lanes per element (more
    # lanes need less LDS per element, so it is admissible too) and other passes
    alt = [min(2 * l, 256) for l in lpe.tolist()]
    alt_pp = [q + 1 for q in pp.tolist()]
    b.set_tiles(alt, alt_pp)
    b.assemble()
    for k in range(64):
        assert _digest(b.values(k)) == digests[k], f"tile shape {alt} {alt_pp}: problem {k} {info[k]}"
    b.set_tiles(lpe, pp)
    # graph replays with eager work between them
    x = L.DeviceArray(ctx, gps[0].n_dofs, np.float32, data=np.ones(gps[0].n_dofs, np.float32))
    y = L.DeviceArray(ctx, gps[0].n_dofs, np.float32)
    for rep in range(2):
        b.assemble()
        L.check(L.hip().lmh_spmv_fp32(b.matrices[0], x.ptr, y.ptr))
        ctx.sync()
        for k in range(64):
            assert _digest(b.values(k)) == digests[k], f"replay {rep}: problem {k}"
            assert _digest(b.rhs(k)) == rdig[k], f"replay {rep}: rhs {k}"
    print(f"tile shapes and replays bit-identical ({time.perf_counter() - t0:.1f} s)", flush=True)
    v1 = [b.values(k) for k in range(64)]
    b.assemble(2.0, 0.3, -1.0)
    for k in range(64):
        rp, _ = b.pattern(k)
        rows_c = np.repeat(cons[k], np.diff(rp))
        v2 = b.values(k)
        assert np.array_equal(v2[~rows_c], 2 * v1[k][~rows_c]), f"E-linearity: problem {k} {info[k]}"
        assert np.array_equal(v2[rows_c], v1[k][rows_c]), f"constrained rows: problem {k}"
        # constrained rows: one nonzero each, a positive integer
        cr = np.flatnonzero(cons[k])
        seg = v1[k][rows_c]
        nz = np.add.reduceat(seg != 0, np.concatenate([[0], np.cumsum(np.diff(rp)[cr])[:-1]]))
        d = np.add.reduceat(seg, np.concatenate([[0], np.cumsum(np.diff(rp)[cr])[:-1]]))
        assert np.all(nz == 1) and np.all(d >= 1) and np.all(d == np.rint(d)), f"constrained diagonals {k}"
        v1[k] = None
    print(f"E-linearity and constrained rows checked ({time.perf_counter() - t0:.1f} s)", flush=True)
    # the smallest problem of each element type against the oracle
    b.assemble()
    for t in (L.HEX8, L.HEX27, L.TET4, L.TET10):
        ks = [k for k in range(64) if info[k][1] == t]
        if not ks:
            continue
        k = min(ks, key=lambda k: info[k][0])
        s = O.System(O.Mesh(info[k][0], t), 3, CLAMP)
        rp, ci = b.pattern(k)
        orp, oci = s.pattern()
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci), f"pattern of problem {k} {info[k]}"
        vals, rhs = s.assemble_elasticity(NPTS[t], 1.0, 0.3, -1.0)
        assert rel(b.values(k).astype(np.float64), vals) < 2e-6, f"values of problem {k} {info[k]}"
        assert rel(b.rhs(k), rhs) < 1e-12, f"rhs of problem {k} {info[k]}"
    print(f"oracle comparisons done ({time.perf_counter() - t0:.1f} s)", flush=True)
    b.close()
    for g in gps:
        g.close()
