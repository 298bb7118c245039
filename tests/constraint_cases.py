"""Constraint-row sets for the constraint tests (shared by the oracle pins and
the GPU parity tests): periodic-style rows that tie the DoFs of one face of the
unit cube to the matching DoFs of the opposite face (u(x=1) = u(x=0), as the
rows of a translation-periodic boundary read for Lagrange elements on matching
meshes), chained rows, and heterogeneous values."""
import numpy as np

# boundary ids of build_cube: 0..5 = z0, y0, x1, y1, x0, z1
Z0, Y0, X1, Y1, X0, Z1 = 1, 2, 4, 8, 16, 32


def node_dofs(conn, elem_dofs, n_nodes, n_vars=1):
    """dof of (node, variable 0) from the element lists."""
    npe = conn.shape[1]
    d = np.full(n_nodes, -1, dtype=np.int64)
    d[conn.ravel()] = elem_dofs[:, :npe].ravel()
    return d


def lattice_key(p, cells):
    """Boundary nodes sit exactly on the lattice (the perturbation moves interior
    nodes only), so their coordinates times 2*cells are integers."""
    return tuple(int(round(c * 2 * cells)) for c in p)


def periodic_rows(coords, conn, elem_dofs, constrained, cells, axes=(0,), taken=None):
    """Rows u(face axis=1) = u(face axis=0) for every DoF on the far face that is
    not constrained already; for several axes the far-far edges chain through
    the rows of the first axis (resolved by process_constraints)."""
    nd = node_dofs(conn, elem_dofs, coords.shape[0])
    taken = set() if taken is None else taken
    rows = []
    for ax in axes:
        near = {}
        for k in range(coords.shape[0]):
            if abs(coords[k, ax]) < 1e-12:
                key = list(lattice_key(coords[k], cells))
                key[ax] = 0
                near[tuple(key)] = nd[k]
        for k in range(coords.shape[0]):
            if abs(coords[k, ax] - 1.0) < 1e-12:
                d = int(nd[k])
                if constrained[d] or d in taken:
                    continue
                key = list(lattice_key(coords[k], cells))
                key[ax] = 0
                partner = int(near[tuple(key)])
                rows.append((d, {partner: 1.0}, 0.0))
                taken.add(d)
    return rows


def free_dofs(constrained, taken):
    return [d for d in range(len(constrained)) if not constrained[d] and d not in taken]


def chain_rows(constrained, taken, rng, count=6):
    """Heterogeneous rows among otherwise free DoFs, each referring to the next
    one's DoF so that substitution has to chain: d0 = 0.5 d1 + 0.25 e0 + 0.1,
    d1 = 2 d2 + 1, ..."""
    free = free_dofs(constrained, taken)
    pick = rng.choice(len(free), size=2 * count + 1, replace=False)
    ds = [free[i] for i in pick[:count + 1]]
    es = [free[i] for i in pick[count + 1:]]
    rows = []
    for k in range(count):
        row = {ds[k + 1]: float(rng.uniform(-1.0, 2.0)), es[k]: float(rng.uniform(-0.5, 0.5))}
        rows.append((ds[k], row, float(rng.uniform(-1.0, 1.0))))
        taken.add(ds[k])
    return rows


def value_row_on(coords, conn, elem_dofs, constrained, taken, face_axis, value):
    """An empty row with a nonzero value on one free DoF of the near face of an
    axis (a heterogeneous Dirichlet-like row that periodic rows then chain
    through)."""
    nd = node_dofs(conn, elem_dofs, coords.shape[0])
    for k in range(coords.shape[0]):
        p = coords[k]
        if abs(p[face_axis]) < 1e-12 and all(1e-9 < p[a] < 1 - 1e-9 for a in range(3) if a != face_axis):
            d = int(nd[k])
            if not constrained[d] and d not in taken:
                taken.add(d)
                return [(d, {}, value)]
    raise AssertionError("no free face DoF")


def case_rows(name, coords, conn, elem_dofs, constrained, cells, seed=5):
    """Named constraint sets used by the tests."""
    rng = np.random.default_rng(seed)
    taken = set()
    if name == "periodic_x":
        # value row on the near face first, so the periodic row of its partner
        # resolves to a pure value
        rows = value_row_on(coords, conn, elem_dofs, constrained, taken, 0, 0.3)
        rows += periodic_rows(coords, conn, elem_dofs, constrained, cells, (0,), taken)
        rows += chain_rows(constrained, taken, rng)
        return rows
    if name == "periodic_xy":
        return periodic_rows(coords, conn, elem_dofs, constrained, cells, (0, 1), taken)
    if name == "random":
        rows = []
        free = free_dofs(constrained, taken)
        pick = rng.choice(len(free), size=min(40, len(free) // 4), replace=False)
        cons = [free[i] for i in pick[:len(pick) // 2]]
        taken.update(cons)
        rest = [d for d in free if d not in taken]
        for d in cons:
            k = int(rng.integers(1, 4))
            row = {int(rest[i]): float(rng.uniform(-1, 1)) for i in rng.choice(len(rest), size=k, replace=False)}
            rows.append((int(d), row, float(rng.uniform(-1, 1)) if rng.random() < 0.5 else 0.0))
        return rows
    raise KeyError(name)
