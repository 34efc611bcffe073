"""The tile-shape autotuners (lbm_autotune, lbm_batch_autotune,
lbm3d_autotune) through the C ABI: the shape they keep fits the lattice and becomes the handle's shape,
the populations the caller had are restored bit for bit, and stepping on
afterwards gives what a handle set to that shape by hand gives."""
import numpy as np
import pytest

from oracle import d3q19 as D
from simuverse_amd import lbm as L
from simuverse_amd import workloads as W
from visible import visible_mask, visible_mask3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gfx950():
    assert L.device_count() >= 1, "no gfx950 device visible"


def channel_solid(nx, ny):
    y, x = np.mgrid[0:ny, 0:nx]
    solid = (y == 0) | (y == ny - 1)
    solid |= (x - nx // 3) ** 2 + (y - ny // 2) ** 2 <= 49
    return solid


@pytest.mark.parametrize("storage", [L.STORAGE_F32, L.STORAGE_F16_DEV], ids=["f32", "f16"])
@pytest.mark.parametrize("nx, ny, x_edge", [(300, 200, 0), (258, 72, 1)], ids=["periodic", "channel"])
def test_autotune_keeps_a_fitting_shape_and_restores_the_state(storage, nx, ny, x_edge):
    print(f"case: autotune {nx}x{ny} x_edge {x_edge} storage {storage}", flush=True)
    if x_edge:
        solid = channel_solid(nx, ny)
        kw = dict(x_edge=L.X_INLET_OUTLET, inlet_ux=0.05, force=(2e-6, 0.0))
    else:
        solid = W.periodic_disks(nx, ny, 0.05, 4, 10, seed=7).solid
        kw = dict(force=(1e-5, 0.0))
    f0 = W.rest_state(nx, ny, solid, amplitude=1e-3).astype(np.float32)

    def fresh():
        lat = L.Lattice(nx, ny, 0.58, storage=storage, **kw)
        lat.write_materials(W.materials_from_solid(solid))
        lat.write_populations(f0, 0)
        lat.run(5)
        return lat

    lat = fresh()
    cur, before = lat.current_buffer(), lat.read_populations()
    v, ms = lat.autotune(steps=6)
    assert L.variant_ok(v, nx, ny, x_edge), v
    assert ms > 0.0 and lat.variant() == v
    assert lat.current_buffer() == cur
    assert np.array_equal(lat.read_populations(), before), "autotune changed the caller's populations"
    lat.run(31)
    tuned = lat.read_populations()
    lat.close()

    ref = fresh()
    ref.set_variant(v)
    ref.run(31)
    vis = visible_mask(solid, periodic_x=not x_edge)
    bad = np.count_nonzero(tuned[vis] != ref.read_populations()[vis])
    ref.close()
    assert bad == 0, f"stepping after autotune differs from shape {v} set by hand at {bad} slots"


def test_d3q19_autotune_keeps_a_fitting_shape_and_restores_the_state():
    print("case: d3q19 autotune 64x48x40", flush=True)
    nx, ny, nz = 64, 48, 40
    solid = W.periodic_spheres(nx, ny, nz, 0.05, 3, 6).solid
    mat = np.where(solid, 4, 1).astype(np.uint8)
    lat = L.Lattice3D(nx, ny, nz, 0.6, force=(1e-5, 0.0, 0.0))
    lat.write_materials(mat)
    lat.run(4)
    before = lat.read_populations()
    v, ms = lat.autotune(steps=4)
    assert 0 <= v < len(L.variant3_shapes()) and ms > 0.0 and lat.variant() == v
    assert np.array_equal(lat.read_populations(), before), "autotune changed the caller's populations"
    lat.run(9)
    tuned = lat.read_populations()
    lat.close()

    ref = L.Lattice3D(nx, ny, nz, 0.6, force=(1e-5, 0.0, 0.0))
    ref.write_materials(mat)
    ref.run(4)
    ref.set_variant(v)
    ref.run(9)
    vis = visible_mask3(solid, D.E)
    bad = np.count_nonzero(tuned[vis] != ref.read_populations()[vis])
    ref.close()
    assert bad == 0, f"stepping after autotune differs from shape {v} set by hand at {bad} slots"


@pytest.mark.parametrize("storage", [L.STORAGE_F32, L.STORAGE_F16_DEV], ids=["f32", "f16"])
def test_batch_autotune_keeps_a_shape_and_restores_the_state(storage):
    print(f"case: batch autotune, storage {storage}", flush=True)
    doms = W.batch_domains(count=6, seed=11, size_min=64, size_max=256)
    spec = [dict(nx=d.nx, ny=d.ny, tau=d.tau, x_edge=L.X_INLET_OUTLET, inlet_ux=d.inlet_ux) for d in doms]

    def fresh():
        b = L.Batch(spec, storage=storage)
        for k, d in enumerate(doms):
            b.write_materials(k, W.materials_from_solid(d.solid))
        b.run(5)
        return b

    b = fresh()
    before = [b.read_populations(k) for k in range(len(doms))]
    v, ms = b.autotune(steps=6)
    assert ms > 0.0 and b.variant() == v
    for k in range(len(doms)):
        assert np.array_equal(b.read_populations(k), before[k]), f"autotune changed domain {k}"
    b.run(31)
    tuned = [b.read_populations(k) for k in range(len(doms))]
    b.close()

    ref = fresh()
    ref.set_variant(v)
    ref.run(31)
    for k, d in enumerate(doms):
        vis = visible_mask(d.solid, periodic_x=False)
        bad = np.count_nonzero(tuned[k][vis] != ref.read_populations(k)[vis])
        assert bad == 0, f"domain {k}: stepping after autotune differs from shape {v} set by hand at {bad} slots"
    ref.close()
