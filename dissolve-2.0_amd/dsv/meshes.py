"""Tetrahedral meshes for the transport workloads (inputs, not part of the solve).

The reference reads a gmsh/mmg mesh through DOLFINx (`domain/mesh_utils.py:200-267`,
`meshing/mesh_generator_disc_adaptive.py:86-123`); neither gmsh nor an HDF5 reader is
available here, so every workload builds its own mesh from a seed:

* ``kuhn_cube``      -- unit cube, n^3 cells, 6 tetrahedra per cell, every cell split along
  its main diagonal (the Kuhn/Freudenthal split, conforming across cells as
  `meshing/mesh_generator_stent_adaptive.py:4-6` describes for its structured split).
* ``jitter``         -- interior vertices moved by uniform(+-f*h) per coordinate.
* ``gyroid_lattice`` -- Kuhn grid over [0, 4pi]^3 keeping tets whose centroid satisfies
  sin x cos y + sin y cos z + sin z cos x < level.
* ``graded_disc``    -- cylinder of radius 1, height 0.2; a seeded Delaunay fill whose point
  spacing grades linearly from h_top at z=0.2 to h_bot at z=0, with slivers removed.
* ``locality_order`` -- Morton (Z-order) renumbering of vertices and cells, so that
  contiguous index ranges are compact in space.

All functions return ``(conn int32[C,4], xyz float64[V,3])``.
"""
from __future__ import annotations

import numpy as np

# The six Kuhn tetrahedra of a unit cell: each is the path 0 -> e_a -> e_a+e_b -> (1,1,1)
# for one permutation (a, b, c) of the axes.  Corner k of the cell has offset
# ((k>>0)&1, (k>>1)&1, (k>>2)&1).
_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def _kuhn_local():
    tets = []
    for a, b, _c in _PERMS:
        k0 = 0
        k1 = 1 << a
        k2 = k1 | (1 << b)
        k3 = 7
        tets.append((k0, k1, k2, k3))
    return np.array(tets, dtype=np.int64)


def kuhn_box(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Structured Kuhn split of a box into 6*nx*ny*nz tets; vertex (i,j,k) has index
    i + (nx+1)*(j + (ny+1)*k) and coordinates lo + (i/nx, j/ny, k/nz)*(hi-lo)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    i = np.arange(nx + 1, dtype=np.float64) / nx
    j = np.arange(ny + 1, dtype=np.float64) / ny
    k = np.arange(nz + 1, dtype=np.float64) / nz
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    xyz = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    xyz = lo + xyz * (hi - lo)
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    order = np.lexsort((ci, cj, ck))          # cells ordered k-major like the vertices
    ci, cj, ck = ci[order], cj[order], ck[order]
    corner = []
    for kk in range(8):
        dx, dy, dz = kk & 1, (kk >> 1) & 1, (kk >> 2) & 1
        corner.append((ci + dx) + (nx + 1) * ((cj + dy) + (ny + 1) * (ck + dz)))
    corner = np.stack(corner, axis=1)         # (ncell, 8)
    loc = _kuhn_local()                       # (6, 4)
    conn = corner[:, loc].reshape(-1, 4)
    return conn.astype(np.int32), xyz


def kuhn_cube(n):
    """Unit cube, n^3 cells, 6 tets each (config: 20^3 -> 48,000 tets, 9,261 vertices)."""
    return kuhn_box(n, n, n)


def boundary_vertex_mask(xyz, lo=0.0, hi=1.0, tol=1e-12):
    return np.any((xyz <= lo + tol) | (xyz >= hi - tol), axis=1)


def jitter(xyz, h, frac, rng, movable=None):
    """Move vertices selected by ``movable`` by uniform(-frac*h, frac*h) per coordinate."""
    out = xyz.copy()
    if movable is None:
        movable = np.ones(len(xyz), dtype=bool)
    d = rng.uniform(-frac * h, frac * h, size=(int(movable.sum()), 3))
    out[movable] += d
    return out


def compact(conn, xyz):
    """Drop vertices no cell uses; renumber keeping the relative vertex order."""
    used = np.zeros(len(xyz), dtype=bool)
    used[conn.ravel()] = True
    new_id = np.cumsum(used) - 1
    return new_id[conn].astype(np.int32), xyz[used]


def gyroid_lattice(n=160, level=0.3, extent=4.0 * np.pi):
    """Kuhn grid over [0, extent]^3 filtered to sin x cos y + sin y cos z + sin z cos x < level
    at the tet centroid (config: n=160 -> 14,661,888 tets, 2,599,899 vertices)."""
    conn, xyz = kuhn_box(n, n, n, hi=(extent, extent, extent))
    keep = np.empty(len(conn), dtype=bool)
    step = 1 << 22
    for s in range(0, len(conn), step):
        c = xyz[conn[s:s + step]].mean(axis=1)
        x, y, z = c[:, 0], c[:, 1], c[:, 2]
        g = np.sin(x) * np.cos(y) + np.sin(y) * np.cos(z) + np.sin(z) * np.cos(x)
        keep[s:s + step] = g < level
    return compact(conn[keep], xyz)


def tet_signed_volume(conn, xyz):
    p = xyz[conn]
    e1 = p[:, 1] - p[:, 0]
    e2 = p[:, 2] - p[:, 0]
    e3 = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", e1, np.cross(e2, e3)) / 6.0


def tet_quality(conn, xyz):
    """Normalised volume quality 6*sqrt(2)*|vol| / l_rms^3 (1 for a regular tet, ~0 for slivers)."""
    p = xyz[conn]
    l2 = np.zeros(len(conn))
    for a in range(4):
        for b in range(a + 1, 4):
            d = p[:, a] - p[:, b]
            l2 += np.einsum("ij,ij->i", d, d)
    lrms = np.sqrt(l2 / 6.0)
    return 6.0 * np.sqrt(2.0) * np.abs(tet_signed_volume(conn, xyz)) / lrms ** 3


def graded_disc(seed=0, h_top=0.005, h_bot=0.04, radius=1.0, height=0.2, scale=1.0,
                q_min=0.05, xy_factor=1.0):
    """Seeded graded Delaunay fill of the cylinder r <= radius, 0 <= z <= height.

    Points lie on horizontal layers; the layer spacing and the in-layer spacing both follow
    h(z) = scale*(h_bot + (h_top-h_bot)*z/height).  Each layer is a rotated, jittered
    hexagonal lattice clipped to the disc plus a ring of points on the rim.  The Delaunay
    tets (scipy.spatial.Delaunay) with quality below ``q_min`` are removed (slivers), then
    unused vertices are dropped.  ``scale`` sets the density: scale=1 gives ~2.4M tets and
    455k vertices (seed 0); ``xy_factor`` shrinks the in-layer spacing relative to the layer
    spacing.  The count moves in steps (the top layers hold most points), so the ~800k-tet
    transport workload uses scale=1.47, xy_factor=0.94 (805,903 tets for seed 0).
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)

    def h_of(z):
        return scale * (h_bot + (h_top - h_bot) * z / height)

    zs = [0.0]
    while True:
        z = zs[-1]
        zn = z + h_of(min(z + 0.5 * h_of(z), height))
        if zn >= height - 0.35 * h_of(height):
            break
        zs.append(zn)
    zs.append(height)
    pts = []
    last = len(zs) - 1
    for li, z in enumerate(zs):
        h = xy_factor * h_of(z)
        # rim ring: random phase per layer; interior layers get a z jitter so that no four
        # rim points are co-circular (co-spherical input makes the Delaunay fill very slow)
        nr = max(8, int(np.ceil(2 * np.pi * radius / h)))
        t = (np.arange(nr) + rng.uniform(0.0, 1.0) + rng.uniform(-0.2, 0.2, nr)) * (2 * np.pi / nr)
        zr = np.full(nr, z)
        if 0 < li < last:
            zr += rng.uniform(-0.2, 0.2, nr) * h
        ring = np.stack([radius * np.cos(t), radius * np.sin(t), zr], axis=1)
        # hexagonal lattice inside r < radius - 0.6h, rotated, shifted and jittered
        m = int(np.ceil(radius / h)) + 2
        a = np.arange(-m, m + 1, dtype=np.float64)
        I, J = np.meshgrid(a, a, indexing="ij")
        X = h * (I + 0.5 * J)
        Y = h * (np.sqrt(3.0) / 2.0) * J
        th = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(th), np.sin(th)
        Xr = c * X - s * Y + rng.uniform(-0.5, 0.5) * h
        Yr = s * X + c * Y + rng.uniform(-0.5, 0.5) * h
        Xr = Xr.ravel() + rng.uniform(-0.2, 0.2, Xr.size) * h
        Yr = Yr.ravel() + rng.uniform(-0.2, 0.2, Yr.size) * h
        inside = Xr * Xr + Yr * Yr < (radius - 0.6 * h) ** 2
        lay = np.stack([Xr[inside], Yr[inside], np.full(int(inside.sum()), z)], axis=1)
        if 0 < li < last:
            lay[:, 2] += rng.uniform(-0.35, 0.35, len(lay)) * h
        else:
            # exactly planar end faces make the Delaunay fill degenerate and very slow:
            # bow them outwards by 0.15h at the axis (0 at the rim)
            bow = 0.15 * h * (1.0 - (lay[:, 0] ** 2 + lay[:, 1] ** 2) / radius ** 2)
            lay[:, 2] += bow if li == last else -bow
        pts.append(ring)
        pts.append(lay)
    xyz = np.concatenate(pts, axis=0)
    tri = Delaunay(xyz)
    conn = tri.simplices.astype(np.int64)
    q = tet_quality(conn, xyz)
    conn = conn[q >= q_min]
    return compact(conn.astype(np.int32), xyz)


def _morton_keys(xyz, bits=21):
    lo = xyz.min(axis=0)
    span = np.maximum(xyz.max(axis=0) - lo, 1e-300)
    q = ((xyz - lo) / span.max() * ((1 << bits) - 1)).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
        return v

    return spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))


def locality_order(conn, xyz):
    """Renumber vertices in Morton order and sort cells by their smallest new vertex id."""
    perm = np.argsort(_morton_keys(xyz), kind="stable")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    conn2 = inv[conn]
    order = np.lexsort((conn2.max(axis=1), conn2.min(axis=1)))
    return conn2[order].astype(np.int32), xyz[perm]
