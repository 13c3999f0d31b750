"""Synthetic meshes and openCARP-style mesh files.

The structured cube is a unit box split into hexahedra, each cut into six
P1 tetrahedra along its main diagonal (the Kuhn split).  Node (i, j, k) sits at
``(i*h, j*h, k*h)`` millimetres and has index ``i*N*N + j*N + k``.

The writers produce the text formats the reference's reader consumes
(core/mesh.py:26-63, core/stimulation.py:27-36): ``.pts`` (count line, then
x y z in micrometres), ``.elem`` (count line, then ``Tt a b c d region``),
``.lon`` (header ``1``, then one fibre per element) and ``.vtx`` (count line,
``extra``, then one node index per line).
"""
from itertools import permutations
from pathlib import Path

import numpy as np

# Hex corner offsets (di, dj, dk) indexed by bit pattern 4*di + 2*dj + dk.
_CORNERS = np.array([[(c >> 2) & 1, (c >> 1) & 1, c & 1] for c in range(8)])


def _kuhn_tets():
    """Six corner-index paths 0 -> 7 through the unit cube, one per axis order."""
    tets = []
    for perm in permutations(range(3)):
        path = [0]
        cur = 0
        for ax in perm:
            cur |= 4 >> ax
            path.append(cur)
        tets.append(path)
    return np.array(tets)


def kuhn_cube(N, edge_mm=None, length_mm=10.0):
    """Structured N x N x N node cube of side ``length_mm`` (default 10 mm).

    Returns (nodes (N^3,3) float64 mm, elems (6(N-1)^3, 4) int32).
    """
    h = length_mm / (N - 1) if edge_mm is None else edge_mm
    g = np.arange(N)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    nodes = np.stack([I.ravel() * h, J.ravel() * h, K.ravel() * h], axis=1).astype(np.float64)
    c = np.arange(N - 1)
    ci, cj, ck = np.meshgrid(c, c, c, indexing="ij")
    base = (ci * N * N + cj * N + ck).ravel()
    off = _CORNERS[:, 0] * N * N + _CORNERS[:, 1] * N + _CORNERS[:, 2]
    tets = _kuhn_tets()
    elems = (base[:, None, None] + off[tets][None, :, :]).reshape(-1, 4)
    return nodes, elems.astype(np.int32)


def centroids(nodes, elems):
    return nodes[elems].mean(axis=1)


def rotating_fibres(nodes, elems, length_mm=10.0, theta0_deg=-60.0, span_deg=120.0):
    """Per-element unit fibres (cos t, sin t, 0), t = theta0 + span * z_c / length."""
    zc = centroids(nodes, elems)[:, 2] / length_mm
    th = np.deg2rad(theta0_deg + span_deg * zc)
    return np.stack([np.cos(th), np.sin(th), np.zeros_like(th)], axis=1)


def corner_nodes(N, m):
    """Indices of the m x m x m corner block of nodes (i, j, k < m)."""
    g = np.arange(m)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    return (I * N * N + J * N + K).ravel()


def face_nodes(N, axis=0, index=0):
    """Indices of the nodes on one face of the cube (e.g. x = 0)."""
    g = np.arange(N)
    A, B = np.meshgrid(g, g, indexing="ij")
    idx = [None, None, None]
    others = [a for a in range(3) if a != axis]
    idx[axis] = np.full_like(A, index)
    idx[others[0]] = A
    idx[others[1]] = B
    return (idx[0] * N * N + idx[1] * N + idx[2]).ravel()


def cube_surface_nodes(N):
    """Indices of the nodes on the boundary of the N^3 node cube."""
    g = np.arange(N)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    on = (I == 0) | (I == N - 1) | (J == 0) | (J == N - 1) | (K == 0) | (K == N - 1)
    return (I * N * N + J * N + K)[on].ravel()


def surface_patch(nodes, surface_ids, size, rng):
    """``size`` boundary nodes nearest to a boundary node drawn with ``rng``."""
    surface_ids = np.asarray(surface_ids)
    c = nodes[surface_ids[rng.integers(surface_ids.shape[0])]]
    d = np.linalg.norm(nodes[surface_ids] - c, axis=1)
    return np.sort(surface_ids[np.argsort(d, kind="stable")[:size]])


def cube_workload(N, model="TTP"):
    """Mesh arrays of the large-cube workload at N nodes per side.

    Two conductivity regions split at x = 5 mm (region 0 below, region 1
    above), fibres rotating about z from -60 to +60 degrees.
    """
    nodes, elems = kuhn_cube(N)
    cen = centroids(nodes, elems)
    regions = (cen[:, 0] >= 5.0).astype(np.int32)
    fibres = rotating_fibres(nodes, elems)
    return nodes, elems, regions, fibres


def write_mesh(folder, nodes_mm, elems, regions, fibres, name="mesh", unit=1000):
    """Write ``.pts`` (micrometres), ``.elem`` and ``.lon`` files into ``folder``."""
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    k = elems.shape[1]
    tag = "Tt" if k == 4 else "Tr"
    with open(folder / f"{name}.pts", "w") as f:
        f.write(f"{nodes_mm.shape[0]}\n")
        np.savetxt(f, nodes_mm * unit, fmt="%.17g")
    with open(folder / f"{name}.elem", "w") as f:
        f.write(f"{elems.shape[0]}\n")
        rows = np.concatenate([elems.astype(np.int64), np.asarray(regions, dtype=np.int64)[:, None]], axis=1)
        np.savetxt(f, rows, fmt=tag + " " + " ".join(["%d"] * (k + 1)))
    with open(folder / f"{name}.lon", "w") as f:
        f.write("1\n")
        np.savetxt(f, np.asarray(fibres, dtype=np.float64), fmt="%.17g")
    return folder


def write_vtx(path, node_ids):
    node_ids = np.asarray(node_ids, dtype=np.int64)
    with open(path, "w") as f:
        f.write(f"{node_ids.shape[0]}\nextra\n")
        np.savetxt(f, node_ids, fmt="%d")
    return Path(path)


def jittered_delaunay(n_side, seed=0, length_mm=10.0, jitter=0.3, permute=True):
    """Unstructured tetrahedral mesh of a cube: a jittered lattice of
    ``n_side**3`` points, Delaunay-tetrahedralised (scipy/Qhull), slivers with
    near-zero volume dropped, nodes optionally randomly renumbered.

    Returns (nodes (n,3) mm, elems (E,4) int32).
    """
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    h = length_mm / (n_side - 1)
    g = np.arange(n_side) * h
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    interior = np.all((P > 0) & (P < length_mm), axis=1)
    P[interior] += rng.uniform(-jitter * h, jitter * h, size=(interior.sum(), 3))
    tets = Delaunay(P).simplices.astype(np.int64)
    X = P[tets]
    vol = np.abs(np.einsum("ij,ij->i", X[:, 1] - X[:, 0],
                           np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6
    tets = tets[vol > 1e-6 * h ** 3]
    if permute:
        perm = rng.permutation(P.shape[0])        # new index of old node i
        inv = np.empty_like(perm)
        inv[perm] = np.arange(P.shape[0])
        P = P[inv]
        tets = perm[tets]
    return P, tets.astype(np.int32)


def surface_of(nodes, elems):
    """Boundary triangles of a tetrahedral mesh (faces used by one tet),
    renumbered to the surface nodes.  Returns (nodes, tris int32)."""
    faces = np.concatenate([elems[:, [0, 1, 2]], elems[:, [0, 1, 3]], elems[:, [0, 2, 3]],
                            elems[:, [1, 2, 3]]])
    key = np.sort(faces, axis=1)
    _, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    tri = faces[idx[cnt == 1]]
    used = np.unique(tri)
    remap = -np.ones(nodes.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.shape[0])
    return nodes[used], remap[tri].astype(np.int32)


# Thick ellipsoidal shell: outer semi-axes (x, y, z) in mm and wall thickness.
SHELL_OUTER_MM = (40.0, 30.0, 60.0)
SHELL_WALL_MM = 10.0
# conductivities (il, it) of the shell's four regions: the cube's two sets and
# the same sets scaled by 0.5
SHELL_CONDUCTIVITIES = (((0,), 0.5272, 0.2076), ((1,), 0.9074, 0.3332),
                        ((2,), 0.5 * 0.5272, 0.5 * 0.2076), ((3,), 0.5 * 0.9074, 0.5 * 0.3332))


def shell_depth(p, outer=SHELL_OUTER_MM, wall=SHELL_WALL_MM, iters=40):
    """Transmural coordinate s of points p (m,3): p lies on the ellipsoid of
    semi-axes outer - wall*(1 - s) (s = 0 inner surface, 1 outer), found by
    bisection on s in [-0.5, 1.5]."""
    A = np.asarray(outer, dtype=np.float64)
    lo = np.full(p.shape[0], -0.5)
    hi = np.full(p.shape[0], 1.5)
    for _ in range(iters):
        s = 0.5 * (lo + hi)
        f = ((p / (A - wall * (1.0 - s)[:, None])) ** 2).sum(1) - 1.0
        lo = np.where(f > 0, s, lo)   # point outside that ellipsoid: s larger
        hi = np.where(f > 0, hi, s)
    return 0.5 * (lo + hi)


def shell_fibres(c, outer=SHELL_OUTER_MM, wall=SHELL_WALL_MM, theta0_deg=-60.0, span_deg=120.0):
    """Unit fibres at points c tangent to the shell's ellipsoid layers,
    rotating transmurally from theta0 (inner) to theta0 + span (outer)
    between the circumferential and the apex-base directions."""
    A = np.asarray(outer, dtype=np.float64)
    s = np.clip(shell_depth(c, outer, wall), 0.0, 1.0)
    nrm = c / (A - wall * (1.0 - s)[:, None]) ** 2
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    ec = np.cross(np.array([0.0, 0.0, 1.0]), nrm)
    small = np.linalg.norm(ec, axis=1) < 1e-6           # at the apex / base pole
    ec[small] = np.cross(np.array([1.0, 0.0, 0.0]), nrm[small])
    ec /= np.linalg.norm(ec, axis=1, keepdims=True)
    el = np.cross(nrm, ec)
    th = np.deg2rad(theta0_deg + span_deg * s)[:, None]
    f = np.cos(th) * ec + np.sin(th) * el
    return f / np.linalg.norm(f, axis=1, keepdims=True)


def ellipsoid_shell(n_target=640_000, seed=0, outer=SHELL_OUTER_MM, wall=SHELL_WALL_MM,
                    jitter=0.3, permute=True):
    """Irregular volume mesh: a thick ellipsoidal shell (outer semi-axes
    ``outer`` mm, inner ones ``outer - wall``) from a seeded jittered lattice,
    Delaunay-tetrahedralised (scipy/Qhull), tets kept when their centroid lies
    in the shell, zero-volume slivers dropped, unused lattice points removed,
    nodes randomly renumbered (``permute``).  Lattice spacing is set so the
    mesh has about ``n_target`` nodes.

    Regions by centroid: 2*(z >= 0) + (x >= 0) (conductivities
    ``SHELL_CONDUCTIVITIES``); fibres from ``shell_fibres`` at the centroids.
    Returns (nodes (n,3) mm, elems (E,4) int32, regions (E,) int32,
    fibres (E,3)).
    """
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    A = np.asarray(outer, dtype=np.float64)
    R = A - wall
    volume = 4.0 / 3.0 * np.pi * (np.prod(A) - np.prod(R))
    h = (1.10 * volume / n_target) ** (1.0 / 3.0)     # 1.10: nodes of boundary tets
    ax = [np.arange(-A[i] - h, A[i] + h + 1e-9, h) for i in range(3)]
    P = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
    P = P + rng.uniform(-jitter * h, jitter * h, size=P.shape)
    keep = (((P / (A + h)) ** 2).sum(1) <= 1.0) & (((P / (R - h)) ** 2).sum(1) >= 1.0)
    P = P[keep]
    tets = Delaunay(P).simplices.astype(np.int64)
    X = P[tets]
    c = X.mean(axis=1)
    inside = (((c / A) ** 2).sum(1) <= 1.0) & (((c / R) ** 2).sum(1) >= 1.0)
    vol = np.abs(np.einsum("ij,ij->i", X[:, 1] - X[:, 0],
                           np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))) / 6
    sel = inside & (vol > 1e-6 * h ** 3)
    tets, c = tets[sel], c[sel]
    used = np.unique(tets)
    remap = np.full(P.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.shape[0])
    P, tets = P[used], remap[tets]
    if permute:
        perm = rng.permutation(P.shape[0])        # new index of old node i
        inv = np.empty_like(perm)
        inv[perm] = np.arange(P.shape[0])
        P = P[inv]
        tets = perm[tets]
    regions = (2 * (c[:, 2] >= 0) + (c[:, 0] >= 0)).astype(np.int32)
    return P, tets.astype(np.int32), regions, shell_fibres(c, outer, wall)


def shell_stimuli(nodes, seed=0, n_patches=5, radius_mm=3.0, outer=SHELL_OUTER_MM,
                  wall=SHELL_WALL_MM):
    """Seeded stimulus patches on the inner (endocardial) surface of the
    shell, alternately apical (z < -0.6 c) and basal (z > 0.6 c): node sets
    within ``radius_mm`` of a random inner-surface node; onsets 0 ms for the
    apical patches and 5 ms for the basal ones.  Returns [(ids, start_ms)]."""
    rng = np.random.default_rng(seed)
    s = shell_depth(nodes, outer, wall)
    inner = s < 0.1
    zc = outer[2]
    out = []
    for k in range(n_patches):
        apical = k % 2 == 0
        cand = np.nonzero(inner & ((nodes[:, 2] < -0.6 * zc) if apical else (nodes[:, 2] > 0.6 * zc)))[0]
        centre = nodes[rng.choice(cand)]
        ids = np.nonzero(np.linalg.norm(nodes - centre, axis=1) <= radius_mm)[0]
        out.append((ids, 0.0 if apical else 5.0))
    return out


def shell_workload(n_target=640_000, seed=0):
    """The irregular-volume workload: shell mesh, its four conductivity sets
    and five stimulus patches.  Returns (nodes, elems, regions, fibres,
    conductivities, [(ids, start)])."""
    nodes, elems, regions, fib = ellipsoid_shell(n_target, seed)
    return nodes, elems, regions, fib, SHELL_CONDUCTIVITIES, shell_stimuli(nodes, seed + 1)
