"""Synthetic shell meshes and programmed metrics (pre-processing layer).

The reference builds its input meshes offline in Python
(meshing_scripts/gauss-codazzi-mainardi.py:58 runs distmesh2d, which seeds a
triangular lattice and re-triangulates with scipy Delaunay,
meshing_scripts/distmesh/_distmesh2d.py:147-176) and writes them with
write_VTK (meshing_scripts/mesh_processing_module.py:523-595).  This module is
this project's own generator for the workloads the per-step path is measured
on: a jittered triangular lattice clipped to a disc (plus a boundary ring) or a
square, Delaunay-triangulated, then relabelled along a Hilbert curve so that
patch gathers on the GPU touch nearby memory (the reference relabels with RCM
for the same reason, mesh_processing_module.py:53-140).

Everything here is plain numpy/scipy and runs on the host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Mesh",
    "disc_mesh",
    "square_mesh",
    "hilbert_order",
    "grf",
    "isotropic_abar_info",
    "lce_defect_abar_info",
    "uniform_bbar_info",
]


@dataclass
class Mesh:
    """A flat reference mesh plus the per-triangle programmed data.

    Arrays follow the reference's input file (read_vtk_data.cpp:84-371):
    ``points`` (N,3) reference positions (z = 0), ``tris`` (T,3) vertex ids
    starting at 0, ``abar_info``/``bbar_info`` (T,k) dialling info,
    ``thickness``/``shear_modulus`` (T,), ``constrained`` (N,) 0/1 flags.
    """

    points: np.ndarray
    tris: np.ndarray
    abar_info: np.ndarray | None = None
    bbar_info: np.ndarray | None = None
    thickness: np.ndarray | None = None
    shear_modulus: np.ndarray | None = None
    constrained: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    @property
    def num_nodes(self) -> int:
        return int(self.points.shape[0])

    @property
    def num_tris(self) -> int:
        return int(self.tris.shape[0])

    def centroids(self) -> np.ndarray:
        return self.points[self.tris].mean(axis=1)

    def fill_defaults(self, h: float = 0.01, mu: float = 1.0) -> "Mesh":
        T, N = self.num_tris, self.num_nodes
        if self.abar_info is None:
            self.abar_info = np.tile([1.0, 0.0, 1.0], (T, 1))
        if self.bbar_info is None:
            self.bbar_info = np.zeros((T, 3))
        if self.thickness is None:
            self.thickness = np.full(T, h)
        if self.shear_modulus is None:
            self.shear_modulus = np.full(T, mu)
        if self.constrained is None:
            self.constrained = np.zeros(N, dtype=np.uint8)
        return self


def _hilbert_d(ix: np.ndarray, iy: np.ndarray, order: int) -> np.ndarray:
    """Hilbert-curve index of integer cell coordinates in [0, 2**order)."""
    x = ix.astype(np.int64).copy()
    y = iy.astype(np.int64).copy()
    d = np.zeros_like(x)
    n = 1 << order
    s = n >> 1
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant
        flip = ry == 0
        swap_x = np.where(flip & (rx == 1), n - 1 - x, x)
        swap_y = np.where(flip & (rx == 1), n - 1 - y, y)
        x = np.where(flip, swap_y, swap_x)
        y = np.where(flip, swap_x, swap_y)
        s >>= 1
    return d


def hilbert_order(xy: np.ndarray, order: int = 16) -> np.ndarray:
    """Permutation that sorts 2-D points along a Hilbert curve (stable)."""
    lo = xy.min(axis=0)
    span = float(max(np.ptp(xy[:, 0]), np.ptp(xy[:, 1]), 1e-300))
    cells = np.floor((xy - lo) / span * ((1 << order) - 1)).astype(np.int64)
    key = _hilbert_d(cells[:, 0], cells[:, 1], order)
    return np.argsort(key, kind="stable")


def _lattice(xmin, xmax, ymin, ymax, s, rng, jitter):
    h = s * math.sqrt(3.0) / 2.0
    j0, j1 = int(math.floor(ymin / h)) - 1, int(math.ceil(ymax / h)) + 1
    i0, i1 = int(math.floor(xmin / s)) - 2, int(math.ceil(xmax / s)) + 2
    J, I = np.meshgrid(np.arange(j0, j1 + 1), np.arange(i0, i1 + 1), indexing="ij")
    x = I * s + (J % 2) * (s / 2.0)
    y = J * h
    pts = np.stack([x.ravel(), y.ravel()], axis=1).astype(np.float64)
    if jitter > 0:
        pts += rng.uniform(-jitter * s, jitter * s, pts.shape)
    return pts


def _triangulate(pts2: np.ndarray, min_area_frac: float = 1e-3) -> tuple[np.ndarray, np.ndarray]:
    """Delaunay + Hilbert relabelling of nodes and triangles; CCW triangles."""
    from scipy.spatial import Delaunay

    tri = Delaunay(pts2).simplices.astype(np.int64)
    p = pts2[tri]
    area2 = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 2, 0] - p[:, 0, 0]) * (
        p[:, 1, 1] - p[:, 0, 1]
    )
    # orient counter-clockwise
    neg = area2 < 0
    tri[neg] = tri[neg][:, [0, 2, 1]]
    area2 = np.abs(area2)
    # drop degenerate hull slivers (none expected for a disc ring, but squares
    # with jittered edges can produce them)
    keep = area2 > min_area_frac * np.median(area2)
    tri = tri[keep]
    used = np.zeros(len(pts2), bool)
    used[tri.ravel()] = True
    remap = -np.ones(len(pts2), np.int64)
    remap[used] = np.arange(used.sum())
    pts2 = pts2[used]
    tri = remap[tri]
    # Hilbert relabelling: nodes first, then triangles by centroid
    perm = hilbert_order(pts2)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    pts2 = pts2[perm]
    tri = inv[tri]
    tperm = hilbert_order(pts2[tri].mean(axis=1))
    tri = tri[tperm]
    pts = np.zeros((len(pts2), 3))
    pts[:, :2] = pts2
    return pts, tri.astype(np.int32)


def disc_mesh(spacing: float, radius: float = 1.0, seed: int = 0, jitter: float = 0.1) -> Mesh:
    """Unit-radius (by default) disc: jittered triangular lattice, clipped.

    Lattice points with r < radius - spacing/2 are kept and a boundary ring of
    ceil(2*pi*radius/spacing) equally spaced points is added on r = radius,
    then the set is Delaunay-triangulated.  At spacing 0.04 this gives about
    4.5k triangles / 2.3k nodes; at 0.002 about 1.81M triangles / 0.91M nodes.
    """
    rng = np.random.default_rng(seed)
    pts = _lattice(-radius, radius, -radius, radius, spacing, rng, jitter)
    r = np.hypot(pts[:, 0], pts[:, 1])
    pts = pts[r < radius - spacing / 2.0]
    nb = int(math.ceil(2.0 * math.pi * radius / spacing))
    th = 2.0 * math.pi * np.arange(nb) / nb
    ring = radius * np.stack([np.cos(th), np.sin(th)], axis=1)
    pts, tri = _triangulate(np.vstack([pts, ring]))
    return Mesh(points=pts, tris=tri, meta={"kind": "disc", "spacing": spacing, "seed": seed})


def square_mesh(spacing: float, side: float = 2.0, seed: int = 0, jitter: float = 0.1) -> Mesh:
    """Square sheet [-side/2, side/2]^2: jittered lattice interior + exact boundary."""
    rng = np.random.default_rng(seed)
    a = side / 2.0
    pts = _lattice(-a, a, -a, a, spacing, rng, jitter)
    inside = (np.abs(pts[:, 0]) < a - spacing / 2.0) & (np.abs(pts[:, 1]) < a - spacing / 2.0)
    pts = pts[inside]
    nb = int(math.ceil(side / spacing))
    u = np.linspace(-a, a, nb + 1)[:-1]
    edges = [
        np.stack([u, np.full_like(u, -a)], 1),
        np.stack([np.full_like(u, a), u], 1),
        np.stack([-u, np.full_like(u, a)], 1),
        np.stack([np.full_like(u, -a), -u], 1),
    ]
    pts, tri = _triangulate(np.vstack([pts] + edges))
    return Mesh(points=pts, tris=tri, meta={"kind": "square", "spacing": spacing, "seed": seed})


def grf(xy: np.ndarray, corr_len: float, n_modes: int = 64, seed: int = 0) -> np.ndarray:
    """Unit-variance Gaussian random field from random Fourier modes.

    g(x) = sqrt(2/M) * sum_i cos(k_i . x + phi_i), k_i ~ N(0, corr_len^-2 I),
    phi_i ~ U[0, 2 pi), evaluated at the given 2-D points.
    """
    rng = np.random.default_rng(seed)
    k = rng.normal(0.0, 1.0 / corr_len, size=(n_modes, 2))
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n_modes)
    out = np.zeros(len(xy))
    for i in range(n_modes):  # loop keeps peak memory at O(len(xy))
        out += np.cos(xy @ k[i] + phi[i])
    return out * math.sqrt(2.0 / n_modes)


def isotropic_abar_info(lam: np.ndarray) -> np.ndarray:
    """Protocol-1 abar_info for abar = lambda^2 I (components xx, xy, yy)."""
    lam2 = np.asarray(lam, dtype=np.float64) ** 2
    return np.stack([lam2, np.zeros_like(lam2), lam2], axis=1)


def lce_defect_abar_info(centroids: np.ndarray, lam_final: float | np.ndarray, nu_ot: float,
                         charge: float = 1.0, offset: float = math.pi / 2) -> np.ndarray:
    """Protocol-3 abar_info rows (director angle, final lambda, opto-thermal nu).

    For a +1 azimuthal defect the director angle is phi + pi/2 (charge 1,
    offset pi/2).  The per-triangle programmed metric is then built on the
    device from these rows exactly as the reference's LCE protocol does
    (do_dialling.cpp:104-124).
    """
    phi = np.arctan2(centroids[:, 1], centroids[:, 0])
    ang = charge * phi + offset
    lam = np.broadcast_to(np.asarray(lam_final, dtype=np.float64), ang.shape)
    return np.stack([ang, lam, np.full_like(ang, nu_ot)], axis=1)


def defect_metric_abar_info(centroids: np.ndarray, lam: np.ndarray, nu_ot: float,
                            offset: float = math.pi / 2) -> np.ndarray:
    """Protocol-1 abar_info (xx, xy, yy) of a +1 azimuthal LCE defect at the
    origin with a per-triangle lambda: abar = lambda^2 d d + lambda^(-2 nu) d' d',
    d the director at angle phi + offset, d' its normal (the dial-factor-1
    metric of do_dialling.cpp:104-124)."""
    ang = np.arctan2(centroids[:, 1], centroids[:, 0]) + offset
    c, s = np.cos(ang), np.sin(ang)
    lam = np.asarray(lam, dtype=np.float64)
    l2, lp = lam * lam, lam ** (-2.0 * nu_ot)
    return np.stack([l2 * c * c + lp * s * s, (l2 - lp) * c * s, lp * c * c + l2 * s * s], axis=1)


def uniform_bbar_info(num_tris: int, kappa: float) -> np.ndarray:
    """bbar = kappa I for every triangle."""
    return np.tile([kappa, 0.0, kappa], (num_tris, 1)).astype(np.float64)


def write_vtk(path: str, mesh: Mesh, header: str = "MorphoShell reference mesh") -> None:
    """Write the strict legacy-VTK input the reference reads (read_vtk_data.cpp:84-371)."""
    mesh.fill_defaults()
    N, T = mesh.num_nodes, mesh.num_tris
    fmt = "%.17g"
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\n%s\nASCII\nDATASET POLYDATA\nPOINTS %d double\n" % (header, N))
        np.savetxt(f, mesh.points, fmt=fmt, delimiter=" ")
        f.write("POLYGONS %d %d\n" % (T, 4 * T))
        np.savetxt(f, np.column_stack([np.full(T, 3), mesh.tris]), fmt="%d", delimiter=" ")
        f.write("CELL_DATA %d\nFIELD tri_quantities 5\n" % T)
        f.write("abar_info %d %d double\n" % (mesh.abar_info.shape[1], T))
        np.savetxt(f, mesh.abar_info, fmt=fmt, delimiter=" ")
        f.write("bbar_info %d %d double\n" % (mesh.bbar_info.shape[1], T))
        np.savetxt(f, mesh.bbar_info, fmt=fmt, delimiter=" ")
        f.write("ref_thicknesses 1 %d double\n" % T)
        np.savetxt(f, mesh.thickness, fmt=fmt)
        f.write("ref_shear_moduli 1 %d double\n" % T)
        np.savetxt(f, mesh.shear_modulus, fmt=fmt)
        f.write("tri_tags 1 %d int\n" % T)
        np.savetxt(f, np.zeros(T, dtype=int), fmt="%d")
        f.write("POINT_DATA %d\nFIELD node_quantities 2\n" % N)
        f.write("constraint_indicators 1 %d int\n" % N)
        np.savetxt(f, mesh.constrained.astype(int), fmt="%d")
        f.write("node_tags 1 %d int\n" % N)
        np.savetxt(f, np.zeros(N, dtype=int), fmt="%d")
