"""Seeded synthetic geometries on an integer voxel lattice.

Every geometry is a set of lattice points ``ijk * Delta`` with the
reference's 18-column neighbour table (utility/voxelization.py:102-133:
column order +x, -x, +y, -y, +z, -z, then the 12 edge neighbours, -1 for a
missing neighbour).  Neighbours come from exact lattice hashing instead of
the reference's KD-tree query, so the table is exact for any spacing.

Geometries:
  * ``sheet``  - the 128 x 128 x 1 sheet of tool/create_geometry.py:41-59
                 (x outer loop, y inner loop, z = 0, 1 mm spacing).
  * ``shell``  - an ellipsoidal left-atrium-like shell with cylindrical vein
                 holes, voxelised at ``Delta`` mm.
  * ``add_scar`` - seeded spherical scar blobs (node_flag 4).
"""
from __future__ import annotations

import numpy as np

# utility/voxelization.py:107-133 offset order.
NEIGHBOR_OFFSETS = np.array([
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
], dtype=np.int64)


def neighbor_table(ijk: np.ndarray, n_cols: int = 18) -> np.ndarray:
    """N x n_cols neighbour ids (-1 = missing) by dense lattice hashing."""
    ijk = np.asarray(ijk, dtype=np.int64)
    n = ijk.shape[0]
    if n == 0:
        return np.full((0, n_cols), -1, dtype=np.int64)
    lo = ijk.min(axis=0) - 1
    dims = ijk.max(axis=0) - lo + 2
    grid = np.full(int(dims[0]) * int(dims[1]) * int(dims[2]), -1, dtype=np.int64)
    rel = ijk - lo
    key = (rel[:, 0] * dims[1] + rel[:, 1]) * dims[2] + rel[:, 2]
    grid[key] = np.arange(n)
    out = np.empty((n, n_cols), dtype=np.int64)
    for c in range(n_cols):
        o = NEIGHBOR_OFFSETS[c]
        kk = ((rel[:, 0] + o[0]) * dims[1] + (rel[:, 1] + o[1])) * dims[2] + (rel[:, 2] + o[2])
        out[:, c] = grid[kk]
    return out


def _geometry(ijk: np.ndarray, delta: float, origin=(0.0, 0.0, 0.0), **extra) -> dict:
    ijk = np.ascontiguousarray(ijk, dtype=np.int32)
    voxel = ijk.astype(np.float64) * delta + np.asarray(origin, dtype=np.float64)
    g = {
        "Delta": delta,
        "voxel": voxel,
        "ijk": ijk,
        "neighbor_id_2d": neighbor_table(ijk),
    }
    g.update(extra)
    return g


def sheet(nx: int = 128, ny: int = 128, delta: float = 1.0) -> dict:
    """2D sheet of tool/create_geometry.py:41-59 (x outer, y inner)."""
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ijk = np.stack([ii.ravel(), jj.ravel(), np.zeros(nx * ny, dtype=np.int64)], axis=1)
    origin = (-(nx // 2) * delta, -(ny // 2) * delta, 0.0)
    return _geometry(ijk, delta, origin, shape=(nx, ny, 1))


def _unit_vectors(rng, n, min_sep_deg=50.0):
    out = []
    tries = 0
    while len(out) < n and tries < 10000:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if all(np.degrees(np.arccos(np.clip(v @ w, -1, 1))) >= min_sep_deg for w in out):
            out.append(v)
        tries += 1
    return np.array(out)


def shell(semi_axes=(25.0, 20.0, 18.0), wall: float = 3.0, delta: float = 0.4,
          n_holes: int = 4, hole_radius: float = 6.0, seed: int = 0,
          chunk: int = 1 << 22) -> dict:
    """Ellipsoidal shell voxelised at ``delta`` mm, cut by vein holes.

    A lattice point p = ijk * delta (cavity centre at the origin) is tissue
    when it lies inside the outer ellipsoid (semi-axes ``semi_axes``), outside
    the inner ellipsoid (semi-axes reduced by ``wall``), and farther than
    ``hole_radius`` from each vein axis (a seeded ray from the origin).
    Voxels are ordered x-outer / z-inner like the reference's slab builder.
    """
    a, b, c = (float(s) for s in semi_axes)
    ai, bi, ci = a - wall, b - wall, c - wall
    rng = np.random.default_rng(seed)
    dirs = _unit_vectors(rng, n_holes)
    hx = int(np.ceil(a / delta)) + 1
    hy = int(np.ceil(b / delta)) + 1
    hz = int(np.ceil(c / delta)) + 1
    xs = np.arange(-hx, hx + 1)
    ys = np.arange(-hy, hy + 1)
    zs = np.arange(-hz, hz + 1)
    keep = []
    ny, nz = len(ys), len(zs)
    total = len(xs) * ny * nz
    for s in range(0, total, chunk):
        lin = np.arange(s, min(total, s + chunk), dtype=np.int64)
        i = xs[lin // (ny * nz)]
        j = ys[(lin // nz) % ny]
        k = zs[lin % nz]
        x, y, z = i * delta, j * delta, k * delta
        qo = (x / a) ** 2 + (y / b) ** 2 + (z / c) ** 2
        qi = (x / ai) ** 2 + (y / bi) ** 2 + (z / ci) ** 2
        m = (qo <= 1.0) & (qi > 1.0)
        if n_holes:
            p = np.stack([x, y, z], axis=1)
            for d in dirs:
                t = p @ d
                r2 = np.einsum("ij,ij->i", p, p) - t * t
                m &= ~((t > 0) & (r2 < hole_radius ** 2))
        keep.append(np.stack([i[m], j[m], k[m]], axis=1))
    ijk = np.concatenate(keep, axis=0)
    g = _geometry(ijk, delta, (0.0, 0.0, 0.0), semi_axes=(a, b, c), wall=wall,
                  hole_dirs=dirs, seed=seed)
    return g


def surface_voxels(geometry: dict) -> np.ndarray:
    """Voxels with at least one missing face neighbour, outer side only."""
    nb = geometry["neighbor_id_2d"][:, :6]
    boundary = np.any(nb < 0, axis=1)
    v = geometry["voxel"]
    if "semi_axes" in geometry:
        a, b, c = geometry["semi_axes"]
        qo = (v[:, 0] / a) ** 2 + (v[:, 1] / b) ** 2 + (v[:, 2] / c) ** 2
        outer = qo > ((a - 0.5 * geometry["wall"]) / a) ** 2
        boundary &= outer
    return np.nonzero(boundary)[0]


def pick_pacing_site(geometry: dict, seed: int) -> int:
    cand = surface_voxels(geometry)
    rng = np.random.default_rng(seed)
    return int(cand[rng.integers(len(cand))])


def pacing_region(geometry: dict, centre: int, radius: float) -> np.ndarray:
    """Voxel ids within ``radius`` mm of voxel ``centre`` (sorted)."""
    v = geometry["voxel"]
    d2 = np.sum((v - v[centre]) ** 2, axis=1)
    return np.nonzero(d2 <= radius * radius)[0]


def add_scar(geometry: dict, n_blobs: int, radius_range=(3.0, 6.0), seed: int = 0,
             node_flag: np.ndarray | None = None) -> np.ndarray:
    """node_flag with seeded spherical scar blobs set to 4 (dense scar)."""
    v = geometry["voxel"]
    n = v.shape[0]
    flag = np.zeros(n, dtype=np.int64) if node_flag is None else node_flag.copy()
    rng = np.random.default_rng(seed)
    for _ in range(n_blobs):
        centre = v[rng.integers(n)]
        r = rng.uniform(*radius_range)
        d2 = np.sum((v - centre) ** 2, axis=1)
        flag[d2 <= r * r] = 4
    return flag


def add_scar_fraction(geometry: dict, n_blobs: int, radius_range=(3.0, 6.0), fraction: float = 0.15,
                      seed: int = 0, node_flag: np.ndarray | None = None) -> np.ndarray:
    """node_flag with n_blobs seeded spherical scar blobs covering ~`fraction` of the voxels.

    Centres and relative radii are seeded as in add_scar (radii drawn in radius_range); one
    common scale of the radii is then found by bisection so that the scarred fraction of the
    wall reaches `fraction`."""
    v = geometry["voxel"]
    n = v.shape[0]
    rng = np.random.default_rng(seed)
    centres = v[rng.integers(n, size=n_blobs)]
    radii = rng.uniform(*radius_range, size=n_blobs)
    d2 = np.stack([np.sum((v - c) ** 2, axis=1) for c in centres])

    def covered(scale):
        return np.any(d2 <= (radii[:, None] * scale) ** 2, axis=0)

    lo, hi = 0.0, 1.0
    while covered(hi).mean() < fraction and hi < 64:
        hi *= 2
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if covered(mid).mean() < fraction:
            lo = mid
        else:
            hi = mid
    flag = np.zeros(n, dtype=np.int64) if node_flag is None else node_flag.copy()
    flag[covered(hi)] = 4
    return flag


def fibonacci_electrodes(n: int, radius: float = 15.0, centre=(0.0, 0.0, 0.0)) -> np.ndarray:
    """n points spread evenly on a sphere (Fibonacci lattice)."""
    k = np.arange(n, dtype=np.float64) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / n)
    theta = np.pi * (1.0 + 5.0 ** 0.5) * k
    pts = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)
    return pts * radius + np.asarray(centre, dtype=np.float64)


def lattice_from_neighbors(neighbor_id_2d: np.ndarray) -> np.ndarray:
    """Recover integer lattice coordinates from face-neighbour links (BFS).

    Used when a geometry dict carries no voxel coordinates.  Each connected
    component is anchored at its lowest voxel id; components are placed
    side by side along x so they never touch.
    """
    nb = np.asarray(neighbor_id_2d)[:, :6]
    n = nb.shape[0]
    ijk = np.zeros((n, 3), dtype=np.int64)
    seen = np.zeros(n, dtype=bool)
    off = NEIGHBOR_OFFSETS[:6]
    shift = 0
    for root in range(n):
        if seen[root]:
            continue
        seen[root] = True
        ijk[root] = (0, 0, 0)
        comp = [root]
        frontier = [root]
        while frontier:
            f = np.asarray(frontier)
            nxt = []
            for c in range(6):
                m = nb[f, c]
                ok = m >= 0
                src, dst = f[ok], m[ok]
                new = ~seen[dst]
                dst_new = dst[new]
                ijk[dst_new] = ijk[src[new]] + off[c]
                seen[dst_new] = True
                nxt.append(np.unique(dst_new))
            frontier = list(np.unique(np.concatenate(nxt))) if nxt else []
            comp.extend(frontier)
        comp = np.asarray(comp)
        ijk[comp] -= ijk[comp].min(axis=0)
        ijk[comp, 0] += shift
        shift = int(ijk[comp, 0].max()) + 3
    return ijk
