"""Unstructured simplex meshes written as gmsh format-2 ASCII files for the tests.

The reference's own meshes (test/data/grids/*.msh) are stored as git-lfs
pointers without content, so the tests build their own: a triangulation of
the unit square and a tetrahedralisation of the unit cube with jittered
interior vertices, diagonals chosen at random per cell, shuffled and
non-contiguous node numbers, elements in random order, boundary segments of
the lower dimension, and physical entity 1 / 2 on the two halves x < 1/2 and
x > 1/2 (the vertices on x = 1/2 are not moved, so the halves meet on a plane).
"""
import itertools

import numpy as np


def _write(path, X, cells, phys, bnd, dim):
    rng = np.random.default_rng(11)
    n = X.shape[0]
    ids = rng.permutation(np.arange(1, 3 * n + 1))[:n]  # shuffled, with gaps
    order = rng.permutation(len(cells))
    etype = {2: 2, 3: 4}[dim]
    btype = {2: 1, 3: 2}[dim]
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(n)]
    for i in rng.permutation(n):
        x = [float(v) for v in X[i]] + [0.0] * (3 - dim)
        lines.append(f"{ids[i]} {x[0]!r} {x[1]!r} {x[2]!r}")
    lines += ["$EndNodes", "$Elements", str(len(bnd) + len(cells))]
    k = 1
    for b in bnd:
        lines.append(f"{k} {btype} 2 7 7 " + " ".join(str(ids[v]) for v in b))
        k += 1
    for e in order:
        lines.append(f"{k} {etype} 2 {phys[e]} {phys[e]} " + " ".join(str(ids[v]) for v in cells[e]))
        k += 1
    lines += ["$EndElements", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def square(path, n=8, seed=0):
    """Triangles on [0,1]^2 (n x n cells, two triangles each)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.array([[x, y] for y in g for x in g])
    h = 1.0 / n
    for i, (x, y) in enumerate(X):
        if 0 < x < 1 and 0 < y < 1:
            if abs(x - 0.5) > 1e-12:
                X[i, 0] += rng.uniform(-0.2, 0.2) * h
            X[i, 1] += rng.uniform(-0.2, 0.2) * h
    vid = lambda i, j: j * (n + 1) + i  # noqa: E731
    cells, phys = [], []
    for j in range(n):
        for i in range(n):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            tri = [(a, b, c), (a, c, d)] if rng.random() < 0.5 else [(a, b, d), (b, c, d)]
            for t in tri:
                t = list(t)
                if rng.random() < 0.5:
                    t[1], t[2] = t[2], t[1]  # either orientation
                cells.append(t)
                phys.append(1 if X[t].mean(axis=0)[0] < 0.5 else 2)
    bnd = [(vid(i, 0), vid(i + 1, 0)) for i in range(n)]
    _write(path, X, cells, phys, bnd, 2)
    return len(cells), X.shape[0]


def cube(path, n=4, seed=0):
    """Tetrahedra on [0,1]^3 (n^3 cells, six Kuhn tetrahedra each, axis order
    drawn per cell)."""
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, n + 1)
    X = np.array([[x, y, z] for z in g for y in g for x in g])
    h = 1.0 / n
    for i, p in enumerate(X):
        if np.all((p > 0) & (p < 1)):
            d = rng.uniform(-0.15, 0.15, 3) * h
            if abs(p[0] - 0.5) < 1e-12:
                d[0] = 0.0
            X[i] += d
    vid = lambda i, j, k: (k * (n + 1) + j) * (n + 1) + i  # noqa: E731
    cells, phys = [], []
    for k, j, i in itertools.product(range(n), repeat=3):
        flip = rng.random() < 0.5
        for p in itertools.permutations(range(3)):
            c = [i, j, k]
            t = [vid(*c)]
            for ax in p:
                c[ax] += 1
                t.append(vid(*c))
            if flip:
                t[2], t[3] = t[3], t[2]
            cells.append(t)
            phys.append(1 if X[t].mean(axis=0)[0] < 0.5 else 2)
    bnd = [(vid(i, j, 0), vid(i + 1, j, 0), vid(i + 1, j + 1, 0)) for i in range(n) for j in range(n)]
    _write(path, X, cells, phys, bnd, 3)
    return len(cells), X.shape[0]
