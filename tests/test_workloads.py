"""The configuration generators (simuverse_amd/workloads.py) against their
specifications, on the CPU:

* periodic_disks (the 4096^2 and 8192^2 roofline scenes): radii in
  [4, 24], no two disks within r + r' + 1 of each other (periodic
  distance), at least 5 % solid and less than one more disk beyond, the
  solid count equal to the sum of each disk's lattice-point count (Gauss's
  circle count, computed separately), and on a smaller scene the mask equal
  to the union of disks found by a periodic k-d tree;
* batch_domains (the 512-domain configuration): sizes multiples of 32 in
  [64, 1024], tau in [0.51, 1), inlet ux in [0.01, 0.08), 0..8 disks of
  radius 2..max(2, min(nx, ny)/8) clear of the walls, inlet and outlet and
  of each other, walls on rows 0 and ny - 1, solid = walls + disks exactly,
  and the seeded cell total;
* cylinder_array_solid (the learned-collision scene): 16 cylinders of
  radius n/64 centred in a 4 x 4 array.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from simuverse_amd import workloads as W


def lattice_points_in_disk(r: int) -> int:
    """#{(i, j) integer: i^2 + j^2 <= r^2}."""
    return sum(2 * math.isqrt(r * r - i * i) + 1 for i in range(-r, r + 1))


def periodic_pair_gaps(disks: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """Periodic centre distance^2 minus (r + r' + 1)^2 over all pairs."""
    c, r = disks[:, :2].astype(np.int64), disks[:, 2].astype(np.int64)
    worst = []
    for k in range(0, len(c), 512):
        d = np.abs(c[k:k + 512, None, :] - c[None, :, :])
        d = np.minimum(d, np.array([nx, ny]) - d)
        gap = (d**2).sum(-1) - (r[k:k + 512, None] + r[None, :] + 1) ** 2
        idx = np.arange(k, min(k + 512, len(c)))
        gap[np.arange(len(idx)), idx] = np.iinfo(np.int64).max  # a disk against itself
        worst.append(gap.min())
    return np.array(worst)


@pytest.mark.parametrize("n", [4096, 8192])
def test_roofline_scenes(n):
    scene = W.periodic_disks(n, n)
    d = scene.disks
    assert len(d) > 100
    assert d[:, 2].min() >= 4 and d[:, 2].max() <= 24
    assert np.all((d[:, 0] >= 0) & (d[:, 0] < n) & (d[:, 1] >= 0) & (d[:, 1] < n))
    assert periodic_pair_gaps(d, n, n).min() > 0
    solid = int(scene.solid.sum())
    assert solid == sum(lattice_points_in_disk(int(r)) for r in d[:, 2])
    assert math.ceil(0.05 * n * n) <= solid < math.ceil(0.05 * n * n) + lattice_points_in_disk(24)
    # both radius ends and the middle are drawn
    assert np.bincount(d[:, 2], minlength=25)[4:25].min() > 0


def test_small_scene_mask_is_union_of_disks():
    nx, ny = 320, 256
    scene = W.periodic_disks(nx, ny, fraction=0.08, rmin=3, rmax=11, seed=77)
    d = scene.disks
    tree = cKDTree(d[:, :2].astype(float), boxsize=(nx, ny))
    y, x = np.mgrid[0:ny, 0:nx]
    pts = np.stack([x.ravel(), y.ravel()], 1).astype(float)
    hits = tree.query_ball_point(pts, r=11.0)
    expect = np.zeros(nx * ny, bool)
    for k, near in enumerate(hits):
        for j in near:
            dx = abs(pts[k, 0] - d[j, 0])
            dy = abs(pts[k, 1] - d[j, 1])
            dx, dy = min(dx, nx - dx), min(dy, ny - dy)
            if dx * dx + dy * dy <= d[j, 2] ** 2:
                expect[k] = True
                break
    assert np.array_equal(scene.solid.ravel(), expect)


def test_batch_domains():
    doms = W.batch_domains()
    assert len(doms) == 512
    cells = sum(dd.nx * dd.ny for dd in doms)
    assert cells == 150_377_472  # the seeded total the benchmark quotes
    sizes = np.array([(dd.nx, dd.ny) for dd in doms])
    assert np.all(sizes % 32 == 0) and sizes.min() >= 64 and sizes.max() <= 1024
    assert len(np.unique(sizes)) == 31  # every allowed size is drawn
    taus = np.array([dd.tau for dd in doms])
    uxs = np.array([dd.inlet_ux for dd in doms])
    assert taus.min() >= 0.51 and taus.max() < 1.0 and taus.min() < 0.52 and taus.max() > 0.99
    assert uxs.min() >= 0.01 and uxs.max() < 0.08 and uxs.min() < 0.012 and uxs.max() > 0.078
    ndisk = np.array([len(dd.disks) for dd in doms])
    assert ndisk.min() == 0 and ndisk.max() <= 8
    for dd in doms:
        nx, ny = dd.nx, dd.ny
        y, x = np.mgrid[0:ny, 0:nx]
        expect = (y == 0) | (y == ny - 1)
        rmax = max(2, min(nx, ny) // 8)
        for cx, cy, r in dd.disks:
            assert 2 <= r <= rmax
            assert r + 4 <= cx <= nx - r - 5 and r + 2 <= cy <= ny - r - 3  # clear of inlet, outlet, walls
            expect |= (x - cx) ** 2 + (y - cy) ** 2 <= r * r
        if len(dd.disks) > 1:
            c = dd.disks[:, :2]
            r = dd.disks[:, 2]
            d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
            gap = d2 - (r[:, None] + r[None, :] + 1) ** 2
            np.fill_diagonal(gap, 1)
            assert gap.min() > 0
        assert np.array_equal(dd.solid, expect)


def test_cylinder_array():
    n = 8192
    solid = W.cylinder_array_solid(n, 4, n // 64)
    r, pitch = n // 64, n // 4
    assert int(solid.sum()) == 16 * lattice_points_in_disk(r)
    for j in range(4):
        for i in range(4):
            cx, cy = i * pitch + pitch // 2, j * pitch + pitch // 2
            assert solid[cy, cx] and solid[cy, cx + r] and not solid[cy, cx + r + 1]
            assert solid[cy - r, cx] and not solid[cy - r - 1, cx]
