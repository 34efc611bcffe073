"""D3Q19 solver (include/lbm3d.h) through the C ABI against the binary64
D3Q19 oracle: populations and rho/u after 300 steps with spheres and a body
force (binary32 and binary16 deviation storage), every kernel shape and the
graph path bit for bit, the exact Poiseuille profile, the conservation
identities, and the 256^3 benchmark lattice through size-independent
properties."""
import numpy as np
import pytest

from oracle import d3q19 as D
from simuverse_amd import lbm as L
from visible import visible_mask3

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gfx950():
    assert L.device_count() >= 1, "no gfx950 device visible"


def spheres(nx, ny, nz, count, rmin, rmax, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.mgrid[0:nz, 0:ny, 0:nx]
    solid = np.zeros((nz, ny, nx), bool)
    for _ in range(count):
        r = int(rng.integers(rmin, rmax + 1))
        c = rng.integers(0, (nx, ny, nz))
        # periodic distance
        dx = np.minimum(np.abs(x - c[0]), nx - np.abs(x - c[0]))
        dy = np.minimum(np.abs(y - c[1]), ny - np.abs(y - c[1]))
        dz = np.minimum(np.abs(z - c[2]), nz - np.abs(z - c[2]))
        solid |= dx * dx + dy * dy + dz * dz <= r * r
    return solid


def perturbed_rest(shape, amplitude, seed):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    return D.rest_state(nx, ny, nz) * (1.0 + amplitude * rng.uniform(-1, 1, (19,) + shape))


def materials(solid):
    return np.where(solid, 4, 1).astype(np.uint8)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("storage", [L.STORAGE_F32, L.STORAGE_F16_DEV], ids=["f32", "f16"])
def test_d3q19_vs_fp64_oracle(storage):
    """24 x 20 x 16, tau 0.6, force (1e-5, -5e-6, 2e-6), spheres, seeded
    perturbation of 1e-2: 300 steps.  binary32: fluid populations within
    2e-6, rho within 1e-6, u relative L2 < 1e-4; binary16 deviations: rho
    within 1e-4, u relative L2 < 5e-2."""
    nx, ny, nz, tau = 24, 20, 16, 0.6
    g = (1e-5, -5e-6, 2e-6)
    print(f"case: d3q19 {nx}x{ny}x{nz} vs oracle, storage {storage}", flush=True)
    solid = spheres(nx, ny, nz, 4, 2, 4, seed=11)
    f0 = perturbed_rest((nz, ny, nx), 1e-2, seed=12)
    lat = L.Lattice3D(nx, ny, nz, tau, storage=storage, force=g)
    lat.write_materials(materials(solid))
    lat.write_populations(f0.astype(np.float32))
    lat.run(300)
    st = D.D3Q19State(tau, solid, f0.astype(np.float32).astype(np.float64), force=g)
    st.run(300)
    fl = ~solid
    rho1, u1 = lat.macro()
    rho0, u0 = st.macro()
    drho = float(np.max(np.abs(rho1 - rho0)[fl]))
    du = rel_l2(u1[:, fl], u0[:, fl])
    print(f"  rho max {drho:.3e}, u rel L2 {du:.3e}", flush=True)
    if storage == L.STORAGE_F32:
        df = float(np.max(np.abs(lat.read_populations()[:, fl] - st.f[:, fl])))
        print(f"  f max {df:.3e}", flush=True)
        assert df < 2e-6
        assert drho < 1e-6
        assert du < 1e-4
    else:
        assert drho < 1e-4
        assert du < 5e-2
    assert np.all(rho1[solid] == 0) and np.all(u1[:, solid] == 0)
    lat.close()


@pytest.mark.parametrize("shape", [(72, 12, 10), (40, 70, 6)], ids=["72x12x10", "40x70x6"])
def test_d3q19_variants_and_graphs_bit_identical(shape):
    """Every kernel shape, with 100-step graphs or direct launches, gives the
    same populations at every slot a fluid cell can read: 72 x 12 x 10 and
    40 x 70 x 6 (row segments cut by nx for every block size; rows split
    unevenly over the XCDs), spheres, force, 203 steps."""
    nx, ny, nz = shape
    print(f"case: d3q19 variants + graphs {shape}", flush=True)
    solid = spheres(nx, ny, nz, 5, 1, 3, seed=21)
    f0 = perturbed_rest((nz, ny, nx), 1e-2, seed=22).astype(np.float32)
    vis = visible_mask3(solid, D.E)
    outs = []
    shapes = L.variant3_shapes()
    for storage in (L.STORAGE_F32, L.STORAGE_F16_DEV):
        ref = None
        for v in range(len(shapes)):
            lat = L.Lattice3D(nx, ny, nz, 0.55, storage=storage, force=(2e-6, 0.0, -1e-6))
            lat.write_materials(materials(solid))
            lat.write_populations(f0)
            lat.set_variant(v)
            lat.set_graph_chunk(100 if v % 2 == 0 else 0)
            lat.run(203)
            out = lat.read_populations()
            lat.close()
            assert np.isfinite(out[vis]).all()
            if ref is None:
                ref = out
            else:
                bad = np.count_nonzero(out[vis] != ref[vis])
                assert bad == 0, f"storage {storage} variant {v} {shapes[v]} differs at {bad} visible slots"
        outs.append(ref)
    # the two storages agree to the binary16 deviation rounding
    assert rel_l2(outs[1][vis], outs[0][vis]) < 1e-4


def test_d3q19_poiseuille_exact():
    """Force-driven flow between solid planes z = 0 and z = nz - 1 at
    tau = 1/2 + sqrt(3/16): the steady profile is the parabola
    g/(2 nu) (z - 1/2)(nz - 3/2 - z) (exact for BGK with half-way
    bounce-back at this tau); binary32 to 1e-4 relative."""
    nx, ny, nz = 16, 8, 12
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    gx = 1e-5
    print("case: d3q19 Poiseuille", flush=True)
    solid = np.zeros((nz, ny, nx), bool)
    solid[0] = solid[nz - 1] = True
    lat = L.Lattice3D(nx, ny, nz, float(tau), force=(gx, 0.0, 0.0))
    lat.write_materials(materials(solid))
    lat.run(3000)
    _, u = lat.macro()
    z = np.arange(1, nz - 1)
    exact = gx / (2 * nu) * (z - 0.5) * (nz - 1.5 - z)
    prof = u[0, 1:nz - 1].mean(axis=(1, 2))
    err = float(np.max(np.abs(prof - exact)) / exact.max())
    print(f"  profile max rel error {err:.3e}", flush=True)
    assert err < 1e-4
    # uniform in x and y
    assert float(np.max(np.abs(u[0, 1:nz - 1] - prof[:, None, None]))) < 1e-6 * exact.max()
    lat.close()


def test_d3q19_reduce_identities():
    """With spheres and a force: mass is conserved and the first moment
    changes by g N_fluid minus the momentum handed to the solids, step by
    step (binary32 rounding of the populations); the device sums equal the
    oracle's on the same state."""
    nx, ny, nz = 40, 24, 20
    g = np.array([3e-6, 1e-6, -2e-6])
    print("case: d3q19 reduce identities", flush=True)
    solid = spheres(nx, ny, nz, 6, 2, 4, seed=31)
    f0 = perturbed_rest((nz, ny, nx), 1e-3, seed=32).astype(np.float32)
    lat = L.Lattice3D(nx, ny, nz, 0.7, force=tuple(g))
    lat.write_materials(materials(solid))
    lat.write_populations(f0)
    nfl = int((~solid).sum())
    m0 = lat.reduce()[0]
    for _ in range(4):
        r0 = lat.reduce()
        lat.run(1)
        r1 = lat.reduce()
        assert np.max(np.abs((r1[1:4] - r0[1:4]) - (g * nfl - r1[4:7]))) < 1e-5 * nfl * 1e-3
    lat.run(500)
    r = lat.reduce()
    assert abs(r[0] - m0) < 1e-6 * m0
    # device sums against the oracle's on the same (read back) state
    st = D.D3Q19State(0.7, solid, lat.read_populations().astype(np.float64), force=tuple(g))
    ro = st.reduce()
    assert abs(r[0] - ro[0]) < 1e-9 * ro[0]
    assert np.max(np.abs(r[1:] - ro[1:])) < 1e-9 * ro[0]
    lat.close()


def test_d3q19_256_cube_properties():
    """The benchmark lattice (256^3, tau 0.55, force 1e-6, seeded spheres
    of radius 4..12 to 5 % solid): after 200 steps from rest every fluid
    cell is finite, mass is conserved to 1e-6 and the mean velocity points
    along the force."""
    from simuverse_amd import workloads as W

    print("case: d3q19 256^3 properties", flush=True)
    scene = W.periodic_spheres(256, 256, 256)
    lat = L.Lattice3D(256, 256, 256, 0.55, force=(1e-6, 0.0, 0.0))
    lat.write_materials(materials(scene.solid))
    m0 = lat.reduce()[0]
    lat.run(200)
    r = lat.reduce()
    assert np.isfinite(r).all()
    assert abs(r[0] - m0) < 1e-6 * m0
    assert r[1] > 0 and abs(r[2]) < 1e-3 * r[1] and abs(r[3]) < 1e-3 * r[1]
    lat.close()
