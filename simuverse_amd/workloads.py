"""Seeded synthetic lattices: geometry, materials and initial populations.

Everything here is deterministic from a 64-bit seed through SplitMix64
(Steele, Lea & Flood, OOPSLA 2014), so the GPU runs, the oracle and the
committed fixtures all start from bit-identical inputs.

Scenes
  reference_poiseuille_info  the reference's own default scene
                             (simuverse/src/fluid/lattice.rs:26-98)
  cylinder_channel_*         256x256 channel past a cylinder of radius 16 at
                             (64, 128), inlet ux = 0.05, tau = 0.6
  periodic_disks             periodic box with seeded non-overlapping disks
                             (radius 4..24) covering >= 5 % of the cells
  batch_domains              512 uneven channel domains with their own tau,
                             inlet velocity and 0..8 disks
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .lbm import LATTICE_INFO_DTYPE

SEED = 0x5EED
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# D2Q9 in the reference's numbering (simuverse/src/fluid/mod.rs:22-52).
EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, -1, 0, 1, -1, -1, 1, 1])
INV = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
W_PHYS = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)
W_WGSL = np.array([0.444444] + [0.111111] * 4 + [0.0277777] * 4, dtype=np.float32)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+count-1 of the SplitMix64 stream seeded with seed."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of each SplitMix64 output."""
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class _Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed, self.pos = seed, 0

    def uniform(self, count: int = 1) -> np.ndarray:
        out = uniform01(self.seed, count, self.pos)
        self.pos += count
        return out

    def integers(self, lo: int, hi: int, count: int = 1) -> np.ndarray:
        """Integers in [lo, hi] inclusive."""
        return lo + np.minimum((self.uniform(count) * (hi - lo + 1)).astype(np.int64), hi - lo)


def noise(nx: int, ny: int, amplitude: float = 1e-4, seed: int = SEED) -> np.ndarray:
    """U(-amplitude, amplitude) per (direction, cell), shape (9, ny, nx), float64."""
    u = uniform01(seed, 9 * nx * ny).reshape(9, ny, nx)
    return (2.0 * u - 1.0) * amplitude


# ---------------------------------------------------------------------------
# Reference scenes (parity mode)
# ---------------------------------------------------------------------------


def _reference_frame(nx: int, ny: int, inlet_vx: float) -> tuple[np.ndarray, np.ndarray]:
    """Walls 2 on rows 0 and ny-1, ghosts 7 on columns 0 and nx-1, inlet 3 on
    column 1 (vx = inlet_vx), outlet 5 on column nx-2, bulk 1 elsewhere, in
    the priority order of lattice.rs:80-92.  Returns (info, inner) where
    inner marks the bulk cells that may become obstacles."""
    y, x = np.mgrid[0:ny, 0:nx]
    wall = (y == 0) | (y == ny - 1)
    ghost = ~wall & ((x == 0) | (x == nx - 1))
    inlet = ~wall & ~ghost & (x == 1)
    outlet = ~wall & ~ghost & ~inlet & (x == nx - 2)
    inner = ~wall & ~ghost & ~inlet & ~outlet
    mat = np.full((ny, nx), 1, np.int32)
    mat[wall] = 2
    mat[ghost] = 7
    mat[inlet] = 3
    mat[outlet] = 5
    info = np.zeros((ny, nx), LATTICE_INFO_DTYPE)
    info["material"] = mat
    info["block_iter"] = -1
    info["vx"] = np.where(inlet, np.float32(inlet_vx), np.float32(0.0))
    info["vy"] = 0.0
    return info, inner


def reference_poiseuille_info(nx: int, ny: int, inlet_vx: float = 0.12, radius: float = 28.0) -> np.ndarray:
    """init_lattice_material(.., Poiseuille) of simuverse/src/fluid/lattice.rs:26-98:
    the frame of _reference_frame plus three disks of the given radius
    centred at (nx/7 - r, ny/2), (nx/5, ny/4), (nx/5, 3ny/4), in binary32
    geometry as glam::Vec2 computes it (|p - c| <= r)."""
    f32 = np.float32
    centres = [
        (f32(nx) / f32(7.0) - f32(radius), f32(ny) / f32(2.0)),
        (f32(nx) / f32(5.0), f32(ny) / f32(4.0)),
        (f32(nx) / f32(5.0), f32(ny) * f32(0.75)),
    ]
    info, inner = _reference_frame(nx, ny, inlet_vx)
    y, x = np.mgrid[0:ny, 0:nx]
    xf, yf = x.astype(f32), y.astype(f32)
    obst = np.zeros((ny, nx), bool)
    for cx, cy in centres:
        dx, dy = xf - cx, yf - cy
        obst |= np.sqrt(dx * dx + dy * dy, dtype=f32) <= f32(radius)
    mat = info["material"]
    mat[inner & obst] = 4
    info["material"] = mat
    return info


def cylinder_parity_info(nx: int = 256, ny: int = 256, inlet_vx: float = 0.05, cx: int = 64, cy: int = 128,
                         r: int = 16) -> np.ndarray:
    """The 256x256 cylinder channel in the reference's materials: the frame
    of _reference_frame with inlet vx = inlet_vx and a cylinder 4 of radius r
    at (cx, cy)."""
    info, inner = _reference_frame(nx, ny, inlet_vx)
    y, x = np.mgrid[0:ny, 0:nx]
    cyl = (x - cx) ** 2 + (y - cy) ** 2 <= r * r
    mat = info["material"]
    mat[inner & cyl] = 4
    info["material"] = mat
    return info


def parity_initial_state(info: np.ndarray, amplitude: float = 1e-4, seed: int = SEED) -> tuple[np.ndarray, np.ndarray]:
    """Buffer 0 = w (reference binary32 weights) + noise on fluid cells, 0 on
    solid cells; buffer 1 = 0 -- the reference's init.wgsl with fluid_ty 1,
    plus the seeded noise."""
    ny, nx = info.shape
    solid = np.isin(info["material"], (2, 4))
    f0 = (W_WGSL[:, None, None].astype(np.float64) + noise(nx, ny, amplitude, seed)).astype(np.float32)
    f0[:, solid] = 0.0
    return f0, np.zeros_like(f0)


# ---------------------------------------------------------------------------
# Physics-mode scenes
# ---------------------------------------------------------------------------


def cylinder_channel_solid(nx: int = 256, ny: int = 256, cx: int = 64, cy: int = 128, r: int = 16) -> np.ndarray:
    """Walls on rows 0 and ny-1 and a cylinder (x-cx)^2 + (y-cy)^2 <= r^2."""
    y, x = np.mgrid[0:ny, 0:nx]
    solid = (y == 0) | (y == ny - 1) | ((x - cx) ** 2 + (y - cy) ** 2 <= r * r)
    return solid


def materials_from_solid(solid: np.ndarray) -> np.ndarray:
    """Solid cells -> LBM_MAT_OBSTACLE (4), fluid -> LBM_MAT_BULK (1)."""
    return np.where(solid, 4, 1).astype(np.uint8)


def rest_state(nx: int, ny: int, solid: np.ndarray | None = None, amplitude: float = 0.0,
               seed: int = SEED) -> np.ndarray:
    """f = w (exact weights) + optional seeded noise, float64, solids 0."""
    f = np.broadcast_to(W_PHYS[:, None, None], (9, ny, nx)).copy()
    if amplitude:
        f += noise(nx, ny, amplitude, seed)
    if solid is not None:
        f[:, solid] = 0.0
    return f


@dataclass
class DiskScene:
    nx: int
    ny: int
    solid: np.ndarray          # (ny, nx) bool
    disks: np.ndarray          # (k, 3) int: cx, cy, r
    seed: int

    @property
    def solid_fraction(self) -> float:
        return float(self.solid.mean())


def periodic_disks(nx: int, ny: int, fraction: float = 0.05, rmin: int = 4, rmax: int = 24,
                   seed: int = SEED, max_tries: int = 1_000_000) -> DiskScene:
    """Seeded non-overlapping disks on a periodic lattice, added by rejection
    sampling until at least `fraction` of the cells are solid.

    A candidate (cx, cy, r) is drawn uniformly (cx in [0, nx), cy in [0, ny),
    r in [rmin, rmax]); it is rejected when its periodic centre distance to
    an accepted disk is <= r + r' + 1 (disks never touch).  A cell is solid
    when its periodic distance^2 to a centre is <= r^2."""
    rng = _Stream(seed)
    solid = np.zeros((ny, nx), bool)
    cells = nx * ny
    target = math.ceil(fraction * cells)
    count = 0
    disks: list[tuple[int, int, int]] = []
    cxs = np.zeros(0, np.int64)
    cys = np.zeros(0, np.int64)
    rs = np.zeros(0, np.int64)
    tries = 0
    while count < target:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("disk placement did not reach the requested solid fraction")
        cx = int(rng.integers(0, nx - 1)[0])
        cy = int(rng.integers(0, ny - 1)[0])
        r = int(rng.integers(rmin, rmax)[0])
        if len(disks):
            dx = np.abs(cxs - cx)
            dx = np.minimum(dx, nx - dx)
            dy = np.abs(cys - cy)
            dy = np.minimum(dy, ny - dy)
            if np.any(dx * dx + dy * dy <= (rs + r + 1) ** 2):
                continue
        if 2 * r + 1 > min(nx, ny):
            continue
        oy, ox = np.mgrid[-r:r + 1, -r:r + 1]
        inside = ox * ox + oy * oy <= r * r
        ys = (cy + oy[inside]) % ny
        xs = (cx + ox[inside]) % nx
        solid[ys, xs] = True
        count += int(inside.sum())
        disks.append((cx, cy, r))
        cxs = np.append(cxs, cx)
        cys = np.append(cys, cy)
        rs = np.append(rs, r)
    return DiskScene(nx, ny, solid, np.array(disks, np.int64).reshape(-1, 3), seed)


@dataclass
class SphereScene:
    nx: int
    ny: int
    nz: int
    solid: np.ndarray          # (nz, ny, nx) bool
    spheres: np.ndarray        # (k, 4) int: cx, cy, cz, r
    seed: int

    @property
    def solid_fraction(self) -> float:
        return float(self.solid.mean())


def periodic_spheres(nx: int, ny: int, nz: int, fraction: float = 0.05, rmin: int = 4, rmax: int = 12,
                     seed: int = SEED, max_tries: int = 1_000_000) -> SphereScene:
    """Seeded non-overlapping spheres on a periodic 3-D lattice (the D3Q19
    workload), the 3-D form of periodic_disks: a candidate (cx, cy, cz, r)
    is drawn uniformly, rejected when its periodic centre distance to an
    accepted sphere is <= r + r' + 1, and spheres are added until at least
    `fraction` of the cells are solid."""
    rng = _Stream(seed)
    solid = np.zeros((nz, ny, nx), bool)
    target = math.ceil(fraction * nx * ny * nz)
    count, tries = 0, 0
    cs = np.zeros((0, 3), np.int64)
    rs = np.zeros(0, np.int64)
    dims = np.array([nx, ny, nz])
    while count < target:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("sphere placement did not reach the requested solid fraction")
        c = np.array([int(rng.integers(0, n - 1)[0]) for n in (nx, ny, nz)])
        r = int(rng.integers(rmin, rmax)[0])
        if 2 * r + 1 > min(nx, ny, nz):
            continue
        if len(rs):
            d = np.abs(cs - c)
            d = np.minimum(d, dims - d)
            if np.any((d * d).sum(1) <= (rs + r + 1) ** 2):
                continue
        oz, oy, ox = np.mgrid[-r:r + 1, -r:r + 1, -r:r + 1]
        inside = ox * ox + oy * oy + oz * oz <= r * r
        solid[(c[2] + oz[inside]) % nz, (c[1] + oy[inside]) % ny, (c[0] + ox[inside]) % nx] = True
        count += int(inside.sum())
        cs = np.vstack([cs, c])
        rs = np.append(rs, r)
    return SphereScene(nx, ny, nz, solid, np.hstack([cs, rs[:, None]]), seed)


@dataclass
class BatchDomain:
    nx: int
    ny: int
    tau: float
    inlet_ux: float
    solid: np.ndarray  # (ny, nx) bool: walls on rows 0, ny-1 and the disks
    disks: np.ndarray  # (k, 3) int: cx, cy, r


def batch_domains(count: int = 512, seed: int = SEED, size_min: int = 64, size_max: int = 1024,
                  size_step: int = 32, max_disks: int = 8) -> list[BatchDomain]:
    """Seeded uneven channel domains: nx and ny uniform over the multiples of
    size_step in [size_min, size_max], tau ~ U(0.51, 1.0), inlet ux ~
    U(0.01, 0.08), walls on rows 0 and ny-1, and 0..max_disks disks of radius
    r in [2, max(2, min(nx, ny)/8)] placed away from the walls, the inlet and
    the outlet (rejected when they overlap an earlier disk)."""
    rng = _Stream(seed)
    sizes = np.arange(size_min, size_max + 1, size_step)
    out = []
    for _ in range(count):
        nx = int(sizes[rng.integers(0, len(sizes) - 1)[0]])
        ny = int(sizes[rng.integers(0, len(sizes) - 1)[0]])
        tau = float(0.51 + 0.49 * rng.uniform()[0])
        ux = float(0.01 + 0.07 * rng.uniform()[0])
        k = int(rng.integers(0, max_disks)[0])
        rmax = max(2, min(nx, ny) // 8)
        y, x = np.mgrid[0:ny, 0:nx]
        solid = (y == 0) | (y == ny - 1)
        disks = []
        for _ in range(k):
            r = int(rng.integers(2, rmax)[0])
            cx = int(rng.integers(r + 4, nx - r - 5)[0])
            cy = int(rng.integers(r + 2, ny - r - 3)[0])
            if any((cx - a) ** 2 + (cy - b) ** 2 <= (r + c + 1) ** 2 for a, b, c in disks):
                continue
            disks.append((cx, cy, r))
            solid |= (x - cx) ** 2 + (y - cy) ** 2 <= r * r
        out.append(BatchDomain(nx, ny, tau, ux, solid, np.array(disks, np.int64).reshape(-1, 3)))
    return out


# ---------------------------------------------------------------------------
# Learned-collision workload: a periodic cylinder array, fine and 4x coarse
# ---------------------------------------------------------------------------


def cylinder_array_solid(n: int, per_side: int = 4, radius: int = 128) -> np.ndarray:
    """n x n periodic lattice with per_side^2 cylinders of `radius` centred
    in equal square cells of the array: solid iff (x-cx)^2 + (y-cy)^2 <= r^2."""
    pitch = n // per_side
    y, x = np.mgrid[0:pitch, 0:pitch]
    c = pitch // 2
    tile = (x - c) ** 2 + (y - c) ** 2 <= radius * radius
    return np.tile(tile, (per_side, per_side))


def equilibrium_state(nx: int, ny: int, ux: float, uy: float = 0.0, solid: np.ndarray | None = None) -> np.ndarray:
    """f = feq(rho = 1, u) on fluid cells (binary64), 0 on solid cells."""
    eu = EX * ux + EY * uy
    feq = W_PHYS * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy))
    f = np.broadcast_to(feq[:, None, None], (9, ny, nx)).copy()
    if solid is not None:
        f[:, solid] = 0.0
    return f
