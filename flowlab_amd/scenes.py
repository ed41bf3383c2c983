"""Scene setup: the label, velocity and dye seeds that feed Simulation.reset().

Vectorised numpy restatements of the reference's scene helpers
(src/scenes/obstacles.ts, testFields.ts, karman.ts, plume.ts) plus the
build-defined generators of the benchmark workloads (seeded with SplitMix64).
Coordinates are WORLD units with h = 1/ny, so the domain is one unit tall
(src/core/simulation.ts:337-365).  Layouts follow src/core/grid.ts:84-141.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import math

import numpy as np

FLUID, AIR, SOLID = 0, 1, 2
DYE_CHANNELS = 3


@dataclass
class Grid:
    nx: int
    ny: int
    h: float

    @property
    def cells(self) -> int:
        return self.nx * self.ny


@dataclass
class DyePatch:
    """A rectangle of prescribed dye in dye-grid cells (src/core/dye.ts:30-36):
    3 value planes then one coverage plane, each nx*ny row-major."""
    i0: int
    j0: int
    nx: int
    ny: int
    data: np.ndarray  # float32, 4*nx*ny


@dataclass
class SceneSpec:
    """src/core/simulation.ts:127-132.  Each part is optional."""
    labels: Optional[Callable] = None      # (g, label) -> None
    seed: Optional[Callable] = None        # (g, u, v) -> None
    dye: Optional[Callable] = None         # (dg, dye[3]) -> None
    dye_source: Optional[Callable] = None  # (dg, g) -> DyePatch


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014): the build's documented PRNG for
    every seeded workload.  next_u64 is the published mixing function;
    uniform() maps the top 53 bits to [0, 1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))

    def uniforms(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        return np.array([self.uniform(lo, hi) for _ in range(n)])


def _centres(n: int, h: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) + 0.5) * h


# ---- labels (src/scenes/obstacles.ts) --------------------------------------

def all_labels(*seeds):
    """obstacles.ts:15-19: later seeds win where they overlap."""
    def f(g, label):
        for s in seeds:
            s(g, label)
    return f


def solid_disk(cx: float, cy: float, r: float):
    """obstacles.ts:30-41: rasterised by cell centre, (x-cx)^2+(y-cy)^2 <= r^2."""
    def f(g: Grid, label: np.ndarray):
        x = _centres(g.nx, g.h)[None, :]
        y = _centres(g.ny, g.h)[:, None]
        m = (x - cx) * (x - cx) + (y - cy) * (y - cy) <= r * r
        label.reshape(g.ny, g.nx)[m] = SOLID
    return f


def solid_rect(x0: float, y0: float, x1: float, y1: float):
    """obstacles.ts:49-60: cell centres with x0<=x<=x1 and y0<=y<=y1."""
    def f(g: Grid, label: np.ndarray):
        x = _centres(g.nx, g.h)[None, :]
        y = _centres(g.ny, g.h)[:, None]
        m = (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)
        label.reshape(g.ny, g.nx)[m] = SOLID
    return f


def open_right():
    """obstacles.ts:138-142: the rightmost column becomes Air."""
    def f(g: Grid, label: np.ndarray):
        label.reshape(g.ny, g.nx)[:, g.nx - 1] = AIR
    return f


def solid_cell_box(i0: int, j0: int, i1: int, j1: int):
    """Cells i0<=i<i1, j0<=j<j1 become Solid (index-space rectangle)."""
    def f(g: Grid, label: np.ndarray):
        label.reshape(g.ny, g.nx)[max(j0, 0):min(j1, g.ny), max(i0, 0):min(i1, g.nx)] = SOLID
    return f


# ---- velocity seeds (src/scenes/testFields.ts) ------------------------------

def add_curl_of_stream(g: Grid, u: np.ndarray, v: np.ndarray, psi: np.ndarray, amp: float):
    """testFields.ts:40-54.  psi is sampled at cell CORNERS, shape (ny+1, nx+1):
    u = amp/h * (psi(i,j+1)-psi(i,j)),  v = -amp/h * (psi(i+1,j)-psi(i,j))."""
    inv = amp / g.h
    U = u.reshape(g.ny, g.nx + 1)
    V = v.reshape(g.ny + 1, g.nx)
    U += inv * (psi[1:, :] - psi[:-1, :])
    V += -inv * (psi[:, 1:] - psi[:, :-1])


def _corner_xy(g: Grid):
    x = (np.arange(g.nx + 1, dtype=np.float64) * g.h)[None, :]
    y = (np.arange(g.ny + 1, dtype=np.float64) * g.h)[:, None]
    return x, y


def add_vortex_pair(g: Grid, u, v, amp=1.0, sigma=0.08):
    """testFields.ts:378-395."""
    k = sigma / (np.sqrt(2.0) * np.exp(-0.5))
    cx = 0.5 * g.nx * g.h
    x, y = _corner_xy(g)

    def blob(bx, by):
        return np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (sigma * sigma))

    add_curl_of_stream(g, u, v, k * (blob(cx - 0.15, 0.3) - blob(cx + 0.15, 0.3)), amp)


def _imul_lcg(seed: int):
    """testFields.ts:401-407: exact uint32 LCG (Math.imul)."""
    s = seed & 0xFFFFFFFF

    def nxt():
        nonlocal s
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        return s / 4294967296.0
    return nxt


def add_vortex_cluster(g: Grid, u, v, amp=1.0, count=14, sigma=0.05, seed=12345):
    """testFields.ts:422-466."""
    k = sigma / (np.sqrt(2.0) * np.exp(-0.5))
    margin = 3.5 * sigma
    span_x = g.nx * g.h - 2 * margin
    span_y = 1 - 2 * margin
    rand = _imul_lcg(seed)
    cx, cy, sign = [], [], []
    for n in range(count):
        cx.append(margin + rand() * span_x)
        cy.append(margin + rand() * span_y)
        sign.append(1.0 if n % 2 == 0 else -1.0)
    x, y = _corner_xy(g)
    psi = np.zeros((g.ny + 1, g.nx + 1))
    for n in range(count):
        dx = x - cx[n]
        dy = y - cy[n]
        psi = psi + sign[n] * np.exp(-(dx * dx + dy * dy) / (sigma * sigma))
    add_curl_of_stream(g, u, v, k * psi, amp)


def make_dye_patch(i0: int, j0: int, nx: int, ny: int, fill) -> DyePatch:
    """dye.ts:46-62.  fill(i, j) -> (rgb[3], coverage) over GLOBAL dye indices,
    vectorised: i, j are broadcast index arrays."""
    j, i = np.meshgrid(np.arange(ny) + j0, np.arange(nx) + i0, indexing="ij")
    rgb, cov = fill(i, j)
    data = np.zeros(4 * nx * ny, np.float32)
    cells = nx * ny
    for c in range(3):
        data[c * cells:(c + 1) * cells] = np.broadcast_to(rgb[c], (ny, nx)).ravel()
    data[3 * cells:] = np.broadcast_to(cov, (ny, nx)).ravel()
    return DyePatch(i0, j0, nx, ny, data)


def heat_disk_source(cx: float, cy: float, r: float, heat=1.0, smoke=1.0):
    """A disk of prescribed heat (channel 0) and smoke (channel 1), coverage 1
    inside by cell centre: the thermal-plume inlet."""
    def src(dg: Grid, g: Grid) -> DyePatch:
        i0 = max(0, int(np.floor((cx - r) / dg.h)) - 1)
        i1 = min(dg.nx, int(np.ceil((cx + r) / dg.h)) + 1)
        j0 = max(0, int(np.floor((cy - r) / dg.h)) - 1)
        j1 = min(dg.ny, int(np.ceil((cy + r) / dg.h)) + 1)

        def fill(i, j):
            x = (i + 0.5) * dg.h
            y = (j + 0.5) * dg.h
            inside = ((x - cx) ** 2 + (y - cy) ** 2 <= r * r).astype(np.float64)
            return (heat * inside, smoke * inside, 0.0 * inside), inside
        return make_dye_patch(i0, j0, i1 - i0, j1 - j0, fill)
    return src


# ---- benchmark workloads ---------------------------------------------------

def thermal_plume_128() -> tuple[Grid, SceneSpec]:
    """128x128 closed box; heat disk of radius 8 cells at bottom centre."""
    g = Grid(128, 128, 1.0 / 128)
    return g, SceneSpec(dye_source=heat_disk_source(0.5, 8 * g.h, 8 * g.h))


def karman_street(nx=2048, ny=1024, seed=20240611, perturb=1e-3) -> tuple[Grid, SceneSpec]:
    """Uniform inflow u=1 past a cylinder of radius 0.05H at (0.25L, 0.5H),
    Air outlet column (zero-gradient outflow), free-slip top and bottom, and
    a seeded 1e-3 perturbation of the inlet column u[0,j] = 1 + 1e-3*xi_j,
    xi ~ U(-1,1) from SplitMix64(seed).  Built from karman.ts:222-238."""
    g = Grid(nx, ny, 1.0 / ny)
    L = nx * g.h
    rng = SplitMix64(seed)
    xi = rng.uniforms(ny, -1.0, 1.0)

    def seed_fn(g, u, v):
        u.fill(1.0)
        u.reshape(g.ny, g.nx + 1)[:, 0] = 1.0 + perturb * xi

    return g, SceneSpec(labels=all_labels(open_right(), solid_disk(0.25 * L, 0.5, 0.05)),
                        seed=seed_fn)


# ---- decaying turbulence among obstacles (batched workload) ------------------

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_array(seed: int, n: int, skip: int = 0) -> np.ndarray:
    """Outputs skip+1 .. skip+n of SplitMix64(seed), vectorised: the k-th output
    is the mixing function of seed + k*gamma (mod 2^64), which is what
    SplitMix64.next_u64 computes one at a time."""
    k = np.arange(skip + 1, skip + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & SplitMix64.MASK) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def splitmix64_uniforms(seed: int, n: int, lo=0.0, hi=1.0) -> np.ndarray:
    """SplitMix64.uniform(lo, hi), n at once."""
    return lo + (hi - lo) * ((splitmix64_array(seed, n) >> np.uint64(11)).astype(np.float64)
                             * (1.0 / (1 << 53)))


def square_lattice(spacing: int, side: int, oi: int, oj: int):
    """Solid squares of `side` cells on a lattice of pitch `spacing` cells,
    offset (oi, oj): cell (i, j) is Solid when (i - oi) mod spacing < side and
    (j - oj) mod spacing < side.  Square obstacles as obstacles.ts:49-60
    rasterises them, in index space."""
    def f(g: Grid, label: np.ndarray):
        L = label.reshape(g.ny, g.nx)
        mi = ((np.arange(g.nx) - oi) % spacing) < side
        mj = ((np.arange(g.ny) - oj) % spacing) < side
        L[np.ix_(mj, mi)] = SOLID
    return f


def spectral_turbulence(seed: int, slope: float = -3.0, rms: float = 1.0):
    """Random solenoidal velocity with energy spectrum E(k) ~ k^slope.

    The stream function is a sine series on the corners, psi(i, j) =
    sum_{m,n} a_mn sin(pi m i / nx) sin(pi n j / ny), so psi = 0 on the walls
    and no face on the boundary carries flow; a_mn = xi_mn k^((slope - 3)/2)
    with k = |(m/nx, n/ny)| and xi ~ U(-1, 1) from SplitMix64(seed), since in
    2D E(k) ~ k^3 |psi(k)|^2.  The velocity is its discrete curl
    (testFields.ts:40-54, divergence-free to rounding), scaled to RMS `rms`
    over all faces."""
    from scipy.fft import idstn

    def f(g: Grid, u: np.ndarray, v: np.ndarray):
        m = np.arange(1, g.nx, dtype=np.float64)[None, :] / g.nx
        n = np.arange(1, g.ny, dtype=np.float64)[:, None] / g.ny
        k = np.sqrt(m * m + n * n)
        xi = splitmix64_uniforms(seed, (g.ny - 1) * (g.nx - 1), -1.0, 1.0).reshape(g.ny - 1,
                                                                                 g.nx - 1)
        a = xi * k ** ((slope - 3.0) / 2.0)
        psi = np.zeros((g.ny + 1, g.nx + 1))
        psi[1:-1, 1:-1] = idstn(a, type=1)
        uu = np.zeros_like(u)
        vv = np.zeros_like(v)
        add_curl_of_stream(g, uu, vv, psi, 1.0)
        s = math.sqrt((np.sum(uu * uu) + np.sum(vv * vv)) / (uu.size + vv.size))
        if s > 0:
            u += uu * (rms / s)
            v += vv * (rms / s)
    return f


BATCH_SHAPES = ((256, 256), (512, 512), (1024, 512), (1024, 1024))


def decaying_turbulence(nx: int, ny: int, seed: int) -> tuple[Grid, SceneSpec]:
    """Closed box (walls all round) with a seeded lattice of square obstacles,
    pitch U[32, 64] cells and area blockage U[0.25, 0.40], and a k^-3
    solenoidal velocity of RMS 1 (spectral_turbulence), all from
    SplitMix64(seed)."""
    g = Grid(nx, ny, 1.0 / ny)
    rng = SplitMix64(seed)
    spacing = rng.randint(32, 64)
    blockage = rng.uniform(0.25, 0.40)
    side = max(1, int(round(spacing * math.sqrt(blockage))))
    oi, oj = rng.randint(0, spacing - 1), rng.randint(0, spacing - 1)
    vel_seed = rng.next_u64()
    return g, SceneSpec(labels=square_lattice(spacing, side, oi, oj),
                        seed=spectral_turbulence(vel_seed))


def turbulence_batch(count: int = 64, seed: int = 20240612,
                     shapes=BATCH_SHAPES) -> list[tuple[Grid, SceneSpec]]:
    """`count` decaying-turbulence scenes with sizes drawn from `shapes` by
    SplitMix64(seed); scene k uses the seed drawn after its shape."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        nx, ny = shapes[rng.randint(0, len(shapes) - 1)]
        out.append(decaying_turbulence(nx, ny, rng.next_u64()))
    return out


# ---- large single domain: many sources and obstacles -------------------------

def _disks(n_cells_x: int, n_cells_y: int, h: float, cx, cy, r, value):
    """Rasterise disks (world units) by cell centre into an (ny, nx) mask-like
    array, each disk evaluated only over its bounding box; later disks win."""
    out = np.zeros((n_cells_y, n_cells_x), np.float32)
    for x0, y0, rr, v in zip(cx, cy, r, value):
        i0 = max(0, int(np.floor((x0 - rr) / h)) - 1)
        i1 = min(n_cells_x, int(np.ceil((x0 + rr) / h)) + 1)
        j0 = max(0, int(np.floor((y0 - rr) / h)) - 1)
        j1 = min(n_cells_y, int(np.ceil((y0 + rr) / h)) + 1)
        if i0 >= i1 or j0 >= j1:
            continue
        x = (np.arange(i0, i1) + 0.5) * h
        y = (np.arange(j0, j1) + 0.5) * h
        m = (x[None, :] - x0) ** 2 + (y[:, None] - y0) ** 2 <= rr * rr
        out[j0:j1, i0:i1][m] = v
    return out


def large_plume(n: int = 8192, seed: int = 20240613, sources: int = 16, obstacles: int = 200,
                cell_scale: float = 1.0):
    """Closed n x n box (walls all round) with `obstacles` seeded circular
    solids (radius U[8, 48] cells) and `sources` buoyant heat disks (radius
    U[16, 64] cells, T ~ U[0.5, 1] in dye channel 0, smoke 1 in channel 1)
    held by one full-domain dye patch; all draws from SplitMix64(seed).
    Sources are placed in the lower half, obstacles anywhere; a source cell
    that is also Solid stays Solid (commitLabels clears its dye).
    cell_scale multiplies every radius in cells: the same draws at
    cell_scale = n / 8192 give the 8192^2 scene's geometry resolved on an
    n x n grid."""
    g = Grid(n, n, 1.0 / n)
    rng = SplitMix64(seed)
    ocx = [rng.uniform(0, 1) for _ in range(obstacles)]
    ocy = [rng.uniform(0, 1) for _ in range(obstacles)]
    orr = [rng.uniform(8, 48) * cell_scale * g.h for _ in range(obstacles)]
    scx = [rng.uniform(0.05, 0.95) for _ in range(sources)]
    scy = [rng.uniform(0.05, 0.5) for _ in range(sources)]
    srr = [rng.uniform(16, 64) * cell_scale * g.h for _ in range(sources)]
    sT = [rng.uniform(0.5, 1.0) for _ in range(sources)]

    def labels(gg: Grid, label: np.ndarray):
        m = _disks(gg.nx, gg.ny, gg.h, ocx, ocy, orr, [1.0] * obstacles)
        label.reshape(gg.ny, gg.nx)[m > 0] = SOLID

    def dye_source(dg: Grid, gg: Grid) -> DyePatch:
        heat = _disks(dg.nx, dg.ny, dg.h, scx, scy, srr, sT)
        cov = _disks(dg.nx, dg.ny, dg.h, scx, scy, srr, [1.0] * sources)
        cells = dg.nx * dg.ny
        data = np.zeros(4 * cells, np.float32)
        data[:cells] = heat.ravel()
        data[cells:2 * cells] = cov.ravel()
        data[3 * cells:] = cov.ravel()
        return DyePatch(0, 0, dg.nx, dg.ny, data)

    return g, SceneSpec(labels=labels, dye_source=dye_source)
