"""Synthetic colonies and field inputs for the mechanics path.

* `planar_restart_fields` - height / normal / wall / growth-rate arrays for
  a planar (x-z) state, built the way the reference derives them: the
  column maximum of cell tops (Compute.cpp:256-340), the central-difference
  surface normal (Compute.cpp:466-491) and a rough wall of uniform draws
  (Simulation.cpp:161-168). They are inputs for parity runs, so only their
  shape and plausibility matter.
* `dome_colony` - the 3-D dome-shaped colony of the 2^20-cell kernel bench:
  spherocylinders of radius 0.5 um and cylinder length U[2,4] um in a
  spherical cap (radius 600 um, height 20 um at the centre), orientations
  near the plane with tilt ~ N(0, 15 deg), small overlaps.
"""
from __future__ import annotations

import math

import numpy as np

from .cells import CELL_DTYPE, Params


def grid_shape(p: Params):
    return p.BoxX, p.BoxY, p.BoxZ


def height_normal_from_cells(p: Params, cells: np.ndarray):
    """Column height (max vertex z, floor cellRadius) and surface normal."""
    ref = p.refinementGridHeight
    nx, ny = p.BoxX * ref, p.BoxY * ref
    h = np.full((nx, ny), p.cellRadius)
    if len(cells):
        hx = p.BoxLength / ref
        cm = 0.5 * (cells["p"] + cells["q"])
        top = np.maximum(cells["p"][:, 2], cells["q"][:, 2])
        ix = np.clip(((cm[:, 0] + p.BoxX // 2 * p.BoxLength) / hx).astype(np.int64), 0, nx - 1)
        if ny > 1:
            iy = np.clip(((cm[:, 1] + p.BoxY // 2 * p.BoxLength) / hx).astype(np.int64), 0, ny - 1)
        else:
            iy = np.zeros_like(ix)
        np.maximum.at(h, (ix, iy), top)
    hx = p.BoxLength / ref
    dxh = np.zeros_like(h)
    dxh[1:-1, :] = (h[2:, :] - h[:-2, :]) / (2.0 * hx)
    dyh = np.zeros_like(h)
    if ny > 2:
        dyh[:, 1:-1] = (h[:, 2:] - h[:, :-2]) / (2.0 * hx)
    ln = np.sqrt(dxh * dxh + dyh * dyh + 1.0)
    normal = np.stack([dxh / ln, dyh / ln, -1.0 / ln], axis=-1)
    return h, normal


def synthetic_fields(p: Params, cells: np.ndarray, seed: int = 0, gr_max: float = 0.9):
    """(height, normal, wall, gr1, gr2) arrays in the reference layouts."""
    rng = np.random.default_rng(seed)
    height, normal = height_normal_from_cells(p, cells)
    wall = (rng.random((p.BoxX, p.BoxY)) - 0.5) * p.wall_rough
    x = np.arange(p.BoxX)[:, None, None]
    z = np.arange(p.BoxZ)[None, None, :]
    prof = 0.5 + 0.5 * np.cos(2 * np.pi * x / max(p.BoxX, 1)) * np.exp(-z / max(p.BoxZ, 1))
    gr1 = np.broadcast_to(gr_max * prof, (p.BoxX, p.BoxY, p.BoxZ)).copy()
    gr2 = 0.8 * gr1
    return height, normal, wall, gr1, gr2


def _rot_axis(theta, tilt):
    """Unit axis from in-plane azimuth theta and elevation tilt."""
    ct = np.cos(tilt)
    return np.stack([ct * np.cos(theta), ct * np.sin(theta), np.sin(tilt)], axis=-1)


def _seg_dist(p1, q1, p2, q2):
    """Vectorised segment-segment distance (same clamping rules as the
    reference's min_distance)."""
    v1 = q1 - p1
    v2 = q2 - p2
    r = p1 - p2
    l1 = np.einsum("ij,ij->i", v1, v1)
    l2 = np.einsum("ij,ij->i", v2, v2)
    f = np.einsum("ij,ij->i", v2, r)
    c = np.einsum("ij,ij->i", v1, r)
    b = np.einsum("ij,ij->i", v1, v2)
    den = l1 * l2 - b * b
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(den != 0, np.clip((b * f - c * l2) / den, 0, 1), 0.0)
        tn = b * s + f
        t = np.where(tn < 0, 0.0, np.where(tn > l2, 1.0, tn / l2))
        s = np.where(tn < 0, np.clip(-c / l1, 0, 1), np.where(tn > l2, np.clip((b - c) / l1, 0, 1), s))
    c1 = p1 + v1 * s[:, None]
    c2 = p2 + v2 * t[:, None]
    return np.sqrt(np.einsum("ij,ij->i", c1 - c2, c1 - c2))


def _lattice_sites(rng, radius, height, Rs, a_par, a_perp, a_z, patch=20.0, window=None):
    """Sites (x, y, z, azimuth) of patch-aligned rod lattices inside the cap
    (only patches meeting |x|, |y| <= window when a window is given)."""
    pts = []
    npx = int(math.ceil(2 * radius / patch))
    for ix in range(npx):
        for iy in range(npx):
            x0 = -radius + ix * patch
            y0 = -radius + iy * patch
            xc, yc = x0 + patch / 2, y0 + patch / 2
            if math.hypot(xc, yc) > radius + patch:
                continue
            if window is not None and (x0 > window or x0 + patch < -window or y0 > window or y0 + patch < -window):
                continue
            th = rng.uniform(0, 2 * np.pi)
            u = np.array([np.cos(th), np.sin(th)])
            w = np.array([-u[1], u[0]])
            m = int(patch * 1.5 / a_par) + 1
            k = int(patch * 1.5 / a_perp) + 1
            ii, jj = np.meshgrid(np.arange(-m, m + 1), np.arange(-k, k + 1), indexing="ij")
            off = ii[..., None] * a_par * u + jj[..., None] * a_perp * w
            off = off.reshape(-1, 2) + np.array([xc, yc])
            keep = (off[:, 0] >= x0) & (off[:, 0] < x0 + patch) & (off[:, 1] >= y0) & (off[:, 1] < y0 + patch)
            off = off[keep]
            rr = np.hypot(off[:, 0], off[:, 1])
            off = off[rr < radius]
            if not len(off):
                continue
            rr = np.hypot(off[:, 0], off[:, 1])
            ztop = np.sqrt(np.maximum(Rs * Rs - rr * rr, 0)) - (Rs - height)
            nz = np.maximum(1, np.floor(ztop / a_z).astype(int))
            for lz in range(int(nz.max())):
                sel = nz > lz
                o = off[sel]
                pts.append(np.column_stack([o, np.full(len(o), 0.5 + lz * a_z), np.full(len(o), th)]))
    return np.concatenate(pts)


def dome_colony(
    n: int = 1 << 20,
    radius: float = 600.0,
    height: float = 20.0,
    seed: int = 2,
    cell_radius: float = 0.5,
    max_overlap: float = 0.05,
    tilt_deg: float = 15.0,
):
    """Synthetic dome: returns a CELL_DTYPE array of exactly n cells.

    Centres sit on a jittered lattice inside the cap (pitch 5 um along the
    rod axis, 1.45 um across, 1.5 um vertically, about 10.9 um^3 per cell,
    shrunk by a few percent if the cap would not hold n cells); cells in one
    20 um patch share an in-plane direction and tilt ~N(0, tilt_deg). Every
    pair overlapping by more than max_overlap is then resolved by shortening
    the higher-ID cell (centre kept), repeated until none is left.
    """
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    R = cell_radius
    # spherical cap: z_top(r) = sqrt(Rs^2 - r^2) - (Rs - h), Rs = (a^2 + h^2) / 2h
    Rs = (radius * radius + height * height) / (2 * height)
    scale = 1.0
    while True:
        pts = _lattice_sites(rng, radius, height, Rs, 5.0 * scale, 1.45 * scale, 1.5 * scale)
        # sites of neighbouring patches (different azimuth) closer than 1.4 um
        # cannot be separated by shortening: drop one of each such pair
        close = cKDTree(pts[:, :3]).query_pairs(1.4, output_type="ndarray")
        cross = close[pts[close[:, 0], 3] != pts[close[:, 1], 3]]
        keep = np.ones(len(pts), bool)
        keep[np.maximum(cross[:, 0], cross[:, 1])] = False
        pts = pts[keep]
        if len(pts) >= n:
            break
        scale *= 0.98
        if scale < 0.9:
            raise ValueError(f"dome of radius {radius} cannot hold {n} cells")
    # choose n sites: keep the inner ones (by a random tie-break on radius)
    rr = np.hypot(pts[:, 0], pts[:, 1]) + rng.random(len(pts)) * 1e-3
    pts = pts[np.argsort(rr, kind="stable")[:n]]
    # order IDs by patch-major lattice order (spatially coherent like a grown colony)
    patch = 20.0
    key = np.floor((pts[:, 0] + radius) / patch) * 1e6 + np.floor((pts[:, 1] + radius) / patch) * 1e3 + pts[:, 2]
    pts = pts[np.argsort(key, kind="stable")]
    cm = pts[:, :3].copy()
    cm[:, :2] += rng.uniform(-0.1, 0.1, (n, 2))
    cm[:, 2] += rng.uniform(0.0, 0.05, n)
    theta = pts[:, 3] + rng.normal(0, 0.05, n)
    tilt = np.deg2rad(rng.normal(0, tilt_deg, n))
    # keep the lowest layer's vertices at or above the agar (z >= R - max_overlap)
    L = rng.uniform(2.0, 4.0, n)
    # vertices below the agar (z < R - max_overlap): halve the tilt until none
    for _ in range(30):
        axis = _rot_axis(theta, tilt)
        zlo = cm[:, 2] - 0.5 * L * np.abs(axis[:, 2])
        low = zlo < R - max_overlap
        if not low.any():
            break
        tilt[low] *= 0.5
    # Shortening a rod (centre kept) never brings it closer to another rod,
    # so only pairs that overlap after placement can ever need work.
    thr = 2 * R - max_overlap
    pairs = cKDTree(cm).query_pairs(L.max() + 2 * R, output_type="ndarray")
    for _ in range(60):
        if not len(pairs):
            break
        a, b = pairs[:, 0], pairs[:, 1]
        ha = 0.5 * L[a, None] * axis[a]
        hb = 0.5 * L[b, None] * axis[b]
        d = _seg_dist(cm[a] + ha, cm[a] - ha, cm[b] + hb, cm[b] - hb)
        bad = d < thr
        pairs, d = pairs[bad], d[bad]
        if not len(pairs):
            break
        hi = np.maximum(pairs[:, 0], pairs[:, 1])
        lo = np.minimum(pairs[:, 0], pairs[:, 1])
        excess = thr - d + 1e-3
        shrink = np.zeros(n)
        np.maximum.at(shrink, hi, 2.0 * excess + 0.05)
        stuck = L[hi] <= 0.3  # the higher ID is already minimal: shorten the other
        np.maximum.at(shrink, lo[stuck], 2.0 * excess[stuck] + 0.05)
        L = np.maximum(L - shrink, 0.3)
    half = 0.5 * L[:, None] * axis
    cells = np.zeros(n, CELL_DTYPE)
    cells["p"] = cm + half
    cells["q"] = cm - half
    cells["Length"] = L
    cells["Radius"] = R
    cells["Ldiv"] = 2.0 * np.maximum(L, 2.0)
    cells["Type"] = 1
    cells["Ancestor"] = np.arange(1, n + 1)
    return cells


def dome_params(cells: np.ndarray, bin_size: float = 5.0, max_neighbours: int = 64, extra_cells: int = 0) -> Params:
    """Bin grid covering a dome colony, in the reference's bin convention
    (x index = int(x/BL + BoxX/2), y index = int(y/BL + BoxY/2 + 0.5))."""
    from .cells import reference_defaults

    p = reference_defaults()
    p.cellRadius = 0.5
    p.L_divide = 3.0
    p.k_cc = 200000.0
    p.k_wc = 200000.0
    p.varL = 0.25
    p.varAngle = 0.001
    p.viscosity = 0.1
    p.wall_rough = 0.1
    p.gamma_t = 5000.0
    p.wall_mu = 0.4
    p.cell_mu = 0.1
    p.density_threshold = 0.55
    p.tension = 500.0
    p.DH = -0.01
    p.crossFeeding = 0
    p.planar = 0
    p.refinementGridHeight = 1
    half = np.abs(np.concatenate([cells["p"][:, :2], cells["q"][:, :2]])).max()
    top = max(cells["p"][:, 2].max(), cells["q"][:, 2].max())
    return _with_bins(p, half, top, len(cells) + extra_cells, bin_size, max_neighbours)


def _with_bins(p: Params, half: float, top: float, max_cells: int, bin_size: float, max_neighbours: int) -> Params:
    """Bins of `bin_size` over |x|, |y| <= half and z <= top, with a margin of
    4 bins."""
    p.BoxLength = bin_size
    nb = 2 * int(math.ceil((half + 4 * bin_size) / bin_size)) + 2
    p.BoxX = p.BoxY = nb
    p.BoxZ = int(math.ceil((top + 4 * bin_size) / bin_size))
    p.maxCells = max_cells
    p.maxNeighbours = max_neighbours
    return p


def growing_params(half: float, top: float, max_cells: int, bin_size: float = 5.0, max_neighbours: int = 64) -> Params:
    """The parameters of dome_params for a colony that will grow to fill
    |x|, |y| <= half and z <= top (bins sized for the final colony)."""
    p = dome_params(small_colony(1), bin_size, max_neighbours)
    return _with_bins(p, half, top, max_cells, bin_size, max_neighbours)


def _segment_distance(p1, q1, p2, q2) -> float:
    """Distance between the segments p1-q1 and p2-q2 (closest points by
    clamping the line parameters to [0, 1], each refined against the other)."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0.0, 1.0) if den > 1e-12 else 0.0
    t = (b * s + f) / e
    if t < 0.0:
        t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
    elif t > 1.0:
        t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return float(np.linalg.norm(p1 + s * d1 - (p2 + t * d2)))


def small_colony(n: int = 16, disk: float = 5.0, seed: int = 1, z: float = 0.5, radius: float = 0.5) -> np.ndarray:
    """A founding colony: n rods on the agar plane (centres uniform in the
    disk r <= `disk` at height z, azimuth U[0, 2pi), length U[2, 4] um,
    division length twice the length), placed one by one and redrawn until
    every pair of axes is at least 2 radii apart."""
    rng = np.random.default_rng(seed)
    placed = []
    tries = 0
    while len(placed) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError(f"placed only {len(placed)} of {n} rods")
        r, phi = disk * np.sqrt(rng.random()), rng.uniform(0.0, 2.0 * np.pi)
        th, L = rng.uniform(0.0, 2.0 * np.pi), rng.uniform(2.0, 4.0)
        cm = np.array([r * np.cos(phi), r * np.sin(phi), z])
        half = 0.5 * L * np.array([np.cos(th), np.sin(th), 0.0])
        seg = (cm + half, cm - half, L)
        if all(_segment_distance(seg[0], seg[1], o[0], o[1]) >= 2.0 * radius for o in placed):
            placed.append(seg)
    cells = np.zeros(n, CELL_DTYPE)
    cells["p"] = [s_[0] for s_ in placed]
    cells["q"] = [s_[1] for s_ in placed]
    cells["Length"] = [s_[2] for s_ in placed]
    cells["Radius"] = radius
    cells["Ldiv"] = 2.0 * cells["Length"]
    cells["Type"] = 1
    cells["Ancestor"] = np.arange(1, n + 1)
    return cells


def dome_window_centres(radius: float, height: float, window: float, seed: int = 3, scale: float = 0.99):
    """Centres of mass of the cells of a dome (lattice of dome_colony at the
    given pitch scale) that fall in |x|, |y| <= window: the part of a large
    colony that a nutrient grid narrower than the colony sees."""
    rng = np.random.default_rng(seed)
    Rs = (radius * radius + height * height) / (2 * height)
    pts = _lattice_sites(rng, radius, height, Rs, 5.0 * scale, 1.45 * scale, 1.5 * scale, window=window)
    cm = pts[:, :3].copy()
    cm[:, :2] += rng.uniform(-0.1, 0.1, (len(cm), 2))
    cm[:, 2] += rng.uniform(0.0, 0.05, len(cm))
    inside = (np.abs(cm[:, 0]) <= window) & (np.abs(cm[:, 1]) <= window)
    return cm[inside]


def bench_dome(n: int = 1 << 20, seed: int = 2, radius: float = 600.0, gr_base: float = 0.875,
               dt_h: float = 1.0 / 3600.0):
    """The kernel-bench workload: dome_colony(n, radius, height 20 um), 5 um
    bins, lists of at most 64, room for n/8 daughters; every cell at a
    uniformly spread point of its cycle (division length = length x
    exp(1.585 gr_base dt_h U[1, 1800]), a steady division rate of 1 s
    steps), growth rate gr_base everywhere, height and normal from the
    cells and a rough wall. Returns (params, cells, (height, normal, wall,
    gr1, gr2))."""
    cells = dome_colony(n, radius=radius, height=20.0, seed=seed)
    p = dome_params(cells, bin_size=5.0, max_neighbours=64, extra_cells=max(n // 8, 1024))
    rng = np.random.default_rng(seed)
    gr = 1.5850 * gr_base
    cells["Ldiv"] = cells["Length"] * np.exp(gr * dt_h * rng.uniform(1.0, 1800.0, n))
    cells["GrowthRate"] = gr_base
    height, normal = height_normal_from_cells(p, cells)
    wall = (rng.random((p.BoxX, p.BoxY)) - 0.5) * p.wall_rough
    gr1 = np.full((p.BoxX, p.BoxY, p.BoxZ), gr_base)
    return p, cells, (height, normal, wall, gr1, gr1.copy())
