"""Parity of the gfx950 library (through its C ABI) with the oracle.

Parity mode is compared with the binary32 restatement of the reference's
WGSL bit for bit; physics mode with the binary64 oracle under the
tolerances written in each test.  Sizes are the ones the oracle finishes in
seconds; the full 4096^2 lattice is checked through properties that do not
depend on size (mass conservation, the momentum identity).
"""
import numpy as np
import pytest

from oracle import oracle as O
from simuverse_amd import lbm as L
from simuverse_amd import workloads as W
from visible import visible_mask

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gfx950():
    n = L.device_count()
    assert n >= 1, "no gfx950 device visible"


def rel_l2(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ---------------------------------------------------------------------------
# Parity mode: reference WGSL semantics, bit for bit
# ---------------------------------------------------------------------------


def run_parity_pair(info, f0, f1, tau, steps, frames_api=False):
    st = O.WgslState(info, tau)
    st.buf[0][:] = f0
    st.buf[1][:] = f1
    cur = st.run(steps)
    ny, nx = info.shape
    lat = L.Lattice(nx, ny, tau, mode=L.MODE_PARITY_WGSL, fluid_ty=1)
    lat.write_lattice_info(info)
    lat.write_populations(f0, 0)
    lat.write_populations(f1, 1)
    if frames_api:
        for _ in range(steps // 2):  # FluidSimulator::compute: step(0), step(1)
            lat.step(0)
            lat.step(1)
    else:
        lat.run(steps, export_macro=True)
    assert lat.current_buffer() == cur
    return st, lat, cur


def test_parity_and_physics_modes_agree_without_clamps():
    """Cross-mode consistency of the two device modes on the case of the
    oracle test of the same name (periodic all-bulk 64 x 48, tau 0.7, a
    vortex with a density wave, 60 steps, every population inside the WGSL
    clamps): the parity-mode mass over the physics-mode mass is the
    truncated weights' closed form (1 - omega (1 - sum w))^t, and the
    velocities agree to 2e-4 (the truncation's drift is 3.5e-5)."""
    from test_oracle_wgsl import bulk_info, moments, smooth_state
    print("case: parity vs physics mode, 64x48 all bulk", flush=True)
    nx, ny, tau, steps = 64, 48, 0.7, 60
    f = smooth_state(nx, ny).astype(np.float32)
    par = L.Lattice(nx, ny, tau, mode=L.MODE_PARITY_WGSL, fluid_ty=1)
    par.write_lattice_info(bulk_info(nx, ny))
    par.write_populations(f, 0)
    par.run(steps)
    phy = L.Lattice(nx, ny, tau)
    phy.write_populations(f, 0)
    phy.run(steps)
    r1, u1, v1 = moments(par.read_populations().astype(np.float64))
    r2, u2, v2 = moments(phy.read_populations().astype(np.float64))
    par.close()
    phy.close()
    omega = float(np.float32(1) / np.float32(tau))
    decay = (1 - omega * (1 - O.wgsl_weights().astype(np.float64).sum())) ** steps
    assert abs((r1 / r2).mean() / decay - 1) < 1e-6
    assert np.abs(r1 / r2 / decay - 1).max() < 5e-6
    du = np.sqrt((((u1 - u2) ** 2 + (v1 - v2) ** 2).sum()) / (u2**2 + v2**2).sum())
    assert du < 2e-4


@pytest.mark.parametrize("nx, ny", [(3, 3), (4, 4), (5, 3), (7, 6), (16, 9)], ids=["3x3", "4x4", "5x3", "7x6", "16x9"])
def test_tiny_periodic_lattices_against_oracles(nx, ny):
    """Brute force on tiny periodic grids, where a two-step tile and its
    one-cell ring wrap around the whole lattice: parity mode bit for bit
    against the WGSL restatement (both buffers, every slot: no solids), and
    physics mode with a body force (and a solid cell when the lattice has
    room), with the default one-step shape, the 128 x 8 two-step tile, the
    512-column two-step row stream and the 64 x 16 and 64 x 24 three-step
    tiles (their two-cell ring wraps around the lattice more than once
    here), against the binary64 oracle; after 1, 2, 7 and 100 steps."""
    from test_oracle_wgsl import bulk_info
    print(f"case: tiny lattice {nx}x{ny}", flush=True)
    rng = np.random.default_rng(100 * nx + ny)
    w = O.wgsl_weights()
    for steps in (1, 2, 7, 100):
        f0 = (w[:, None, None] + rng.uniform(-1e-3, 1e-3, (9, ny, nx))).astype(np.float32)
        st, lat, _ = run_parity_pair(bulk_info(nx, ny), f0, f0.copy(), 0.6, steps)
        for b in (0, 1):
            assert np.array_equal(lat.read_populations(b), st.buf[b]), f"parity, {steps} steps, buffer {b}"
        lat.close()
    solid = np.zeros((ny, nx), bool)
    if nx >= 4 and ny >= 4:
        solid[ny // 2, nx // 2] = True
    f0 = W.W_PHYS[:, None, None] * (1 + rng.uniform(-1e-3, 1e-3, (9, ny, nx)))
    force = (1e-5, -5e-6)
    for steps in (1, 2, 7, 100):
        st = O.PhysState(O.phys_params(nx, ny, 0.7, force=force), solid, f0)
        st.run(steps)
        r0, u0, v0 = st.macro()
        for variant in (None, 18, 34, 41, 46):
            lat = L.Lattice(nx, ny, 0.7, force=force)
            lat.write_materials(W.materials_from_solid(solid))
            lat.write_populations(f0.astype(np.float32), 0)
            if variant is not None:
                lat.set_variant(variant)
            lat.run(steps)
            r1, u1, v1 = lat.macro()
            lat.close()
            fl = ~solid
            assert np.max(np.abs(r1[fl] - r0[fl])) < 1e-6, f"physics rho, {steps} steps, variant {variant}"
            assert np.max(np.hypot(u1 - u0, v1 - v0)[fl]) < 1e-6, f"physics u, {steps} steps, variant {variant}"


def test_parity_c1_cylinder_1000_steps_bit_exact():
    """256x256 channel past a cylinder r=16 at (64,128), tau=0.6, inlet
    accelerate column vx=0.05, f0 = w + U(-1e-4, 1e-4), 1000 steps."""
    print("case: parity C1 256x256 tau=0.6 1000 steps", flush=True)
    info = W.cylinder_parity_info()
    f0, f1 = W.parity_initial_state(info)
    st, lat, cur = run_parity_pair(info, f0, f1, 0.6, 1000)
    solid = np.isin(info["material"], (2, 4))
    vis = visible_mask(solid)
    for b in (cur, cur ^ 1):
        g = lat.read_populations(b)
        mism = np.count_nonzero(g[vis] != st.buf[b][vis])
        assert mism == 0, f"buffer {b}: {mism} visible slots differ, rel L2 {rel_l2(g, st.buf[b], vis):.3e}"
    rho, ux, uy = lat.read_macro()
    np.testing.assert_array_equal(rho, st.macro[..., 2])
    np.testing.assert_array_equal(ux, st.macro[..., 0])
    np.testing.assert_array_equal(uy, st.macro[..., 1])
    # the flow is developed: a non-trivial wake behind the cylinder
    assert np.max(np.abs(uy)) > 1e-3
    # the macro of the current state equals what the next step exports
    r2, u2, v2 = lat.macro()
    st.step(cur)
    np.testing.assert_array_equal(r2, st.macro[..., 2])
    np.testing.assert_array_equal(u2, st.macro[..., 0])
    np.testing.assert_array_equal(v2, st.macro[..., 1])


def test_parity_reference_scene_with_reset_and_edits():
    """The reference's own default scene (init_lattice_material, Poiseuille,
    600x375, tau = 0.56) from init.wgsl's Poiseuille-biased start, driven
    frame by frame like FluidSimulator::compute, with an obstacle click and
    an external-force drag written between frames (d2q9_node.rs:215-300)."""
    print("case: parity reference scene 600x375 tau=0.56", flush=True)
    nx, ny = 600, 375
    tau = float(np.float32(3.0) * np.float32(0.02) + np.float32(0.5))  # d2q9_node.rs:50
    info = W.reference_poiseuille_info(nx, ny)
    st = O.WgslState(info, tau)
    st.init(fluid_ty=0)
    lat = L.Lattice(nx, ny, tau, mode=L.MODE_PARITY_WGSL, fluid_ty=0)
    lat.write_lattice_info(info)
    lat.reset()
    np.testing.assert_array_equal(lat.read_populations(0), st.buf[0])
    np.testing.assert_array_equal(lat.read_populations(1), st.buf[1])

    def frame():
        st.step(0)
        st.step(1)
        lat.step(0)
        lat.step(1)

    for _ in range(40):
        frame()
    # add_obstacle at (300, 180): rows min_y .. min_y + 2r of the info buffer
    cx, cy, r = 300, 180, 28
    rows = slice(cy - r, cy + r)
    yy, xx = np.mgrid[cy - r:cy + r, 0:nx]
    disk = (xx + 0.5 - (cx + 0.5)) ** 2 + (yy + 0.5 - (cy + 0.5)) ** 2 <= r * r
    block = st.info[rows].copy()
    block["material"][disk] = 4
    block["block_iter"][disk] = -1
    block["vx"][disk] = 0.0
    block["vy"][disk] = 0.0
    st.info[rows] = block
    lat.write_lattice_info(block, offset=(cy - r) * nx)
    # add_external_force: material 6, block_iter 90 records along a stroke
    # over fluid cells (a stroke over a solid cell turns it into fluid and
    # exposes slots the reference's boundary pass writes racily)
    for k in range(20):
        x, y = 300 + 3 * k, 60 + k
        rec = np.array([(6, 90, 0.08, 0.04)], dtype=L.LATTICE_INFO_DTYPE)
        st.info[y, x] = rec[0]
        lat.write_lattice_info(rec, offset=y * nx + x)
    for _ in range(60):
        frame()
    solid = np.isin(st.info["material"], (2, 4))
    vis = visible_mask(solid)
    for b in (0, 1):
        g = lat.read_populations(b)
        mism = np.count_nonzero(g[vis] != st.buf[b][vis])
        assert mism == 0, f"buffer {b}: {mism} visible slots differ"
    got = lat.read_lattice_info()
    np.testing.assert_array_equal(got["material"], st.info["material"])
    np.testing.assert_array_equal(got["block_iter"], st.info["block_iter"])
    rho, ux, uy = lat.read_macro()
    np.testing.assert_array_equal(rho, st.macro[..., 2])
    np.testing.assert_array_equal(ux, st.macro[..., 0])
    rgba = lat.read_macro_rgba16f()
    np.testing.assert_array_equal(rgba, st.macro.astype(np.float16))
    # the curl pass on those texels (curl_update.wgsl), bit for bit
    curl = lat.read_curl_rgba16f()
    np.testing.assert_array_equal(curl.view(np.uint16), O.wgsl_curl_rgba16f(st.macro).view(np.uint16))
    assert np.count_nonzero(curl[..., 0] != np.float16(0.5)) > 100


# ---------------------------------------------------------------------------
# Physics mode against the binary64 oracle
# ---------------------------------------------------------------------------


def physics_pair(nx, ny, tau, solid, f0, steps, x_edge=0, inlet_ux=0.0, force=(0.0, 0.0),
                 storage=L.STORAGE_F32, layout=L.LAYOUT_ROWS):
    p = O.phys_params(nx, ny, tau, x_edge=x_edge, inlet_ux=inlet_ux, force=force)
    st = O.PhysState(p, solid, f0)
    st.run(steps)
    lat = L.Lattice(nx, ny, tau, mode=L.MODE_PHYSICS, storage=storage, x_edge=x_edge, inlet_ux=inlet_ux,
                    force=force, layout=layout)
    lat.write_materials(W.materials_from_solid(solid))
    lat.write_populations(f0.astype(np.float32), 0)
    lat.run(steps)
    return st, lat


def macro_errors(st, lat, solid):
    r0, u0, v0 = st.macro()
    r1, u1, v1 = lat.macro()
    fl = ~solid
    du = np.hypot(u1 - u0, v1 - v0)[fl]
    umag = np.hypot(u0, v0)[fl]
    return {
        "rho_l2": rel_l2(r1[fl].astype(np.float64), r0[fl]),
        "rho_max": float(np.max(np.abs(r1[fl] - r0[fl]))),
        "u_l2": float(np.linalg.norm(du) / np.linalg.norm(umag)),
        "u_max": float(np.max(du)),
        "u_scale": float(np.max(umag)),
    }


LAYOUTS = pytest.mark.parametrize("layout", [L.LAYOUT_ROWS, L.LAYOUT_PLANES], ids=["rows", "planes"])


@LAYOUTS
def test_physics_c1_channel_cylinder_vs_fp64(layout):
    """256x256 channel: Zou-He inlet ux=0.05, zero-gradient outlet, walls,
    cylinder r=16 at (64,128), tau=0.6, f0 = w + U(-1e-4,1e-4), 1000 steps.
    binary32 storage and arithmetic against binary64: rho and u relative L2
    <= 1e-5."""
    print("case: physics C1 256x256 1000 steps", flush=True)
    nx = ny = 256
    solid = W.cylinder_channel_solid()
    f0 = W.rest_state(nx, ny, solid, amplitude=1e-4)
    st, lat = physics_pair(nx, ny, 0.6, solid, f0, 1000, x_edge=1, inlet_ux=0.05, layout=layout)
    e = macro_errors(st, lat, solid)
    print("physics C1 errors vs fp64:", e, flush=True)
    assert e["rho_l2"] < 1e-5 and e["u_l2"] < 1e-5
    assert e["u_max"] < 1e-5 * 10


@LAYOUTS
def test_physics_periodic_disks_force_vs_fp64(layout):
    """The roofline scene at 256^2: periodic, Guo force g=1e-6 (x), seeded
    disks r in 4..24 on >= 5 % of the cells, tau = 0.55, 1000 steps."""
    print("case: physics periodic disks 256x256 1000 steps", flush=True)
    nx = ny = 256
    scene = W.periodic_disks(nx, ny)
    f0 = W.rest_state(nx, ny, scene.solid)
    st, lat = physics_pair(nx, ny, 0.55, scene.solid, f0, 1000, force=(1e-6, 0.0), layout=layout)
    e = macro_errors(st, lat, scene.solid)
    print("physics periodic-disk errors vs fp64:", e, flush=True)
    # The force adds 1e-7 per step to deviations of ~1e-3, which binary32
    # resolves to ~1e-4 relative: u carries that floor.
    assert e["rho_l2"] < 1e-6 and e["u_l2"] < 2e-4
    r_gpu = lat.reduce()
    r_ref = st.reduce()
    print("reduce gpu", r_gpu, "fp64", r_ref, flush=True)
    np.testing.assert_allclose(r_gpu[0], r_ref[0], rtol=1e-7)
    assert np.hypot(*(r_gpu[1:3] - r_ref[1:3])) < 1e-4 * np.hypot(*r_ref[1:3])
    assert np.hypot(*(r_gpu[3:] - r_ref[3:])) < 1e-3 * np.hypot(*r_ref[3:])


def test_physics_reduce_identities_small():
    """binary64 reductions on the GPU obey the same exact identities as the
    oracle: P(t+1) - P(t) = g N_fluid - drag(t), mass conserved."""
    print("case: physics reduce identities 96x80", flush=True)
    nx, ny = 96, 80
    scene = W.periodic_disks(nx, ny, 0.08, 3, 6, seed=9)
    g = (2e-5, -1e-5)
    lat = L.Lattice(nx, ny, 0.55, force=g)
    lat.write_materials(W.materials_from_solid(scene.solid))
    lat.write_populations(W.rest_state(nx, ny, scene.solid, amplitude=1e-3).astype(np.float32), 0)
    nfl = int((~scene.solid).sum())
    m0 = lat.reduce()[0]
    for _ in range(4):
        r0 = lat.reduce()
        lat.run(1)
        r1 = lat.reduce()
        # binary32 populations: the identity holds to the rounding of f
        assert abs((r1[1] - r0[1]) - (g[0] * nfl - r1[3])) < 1e-5 * nfl * 1e-3
        assert abs((r1[2] - r0[2]) - (g[1] * nfl - r1[4])) < 1e-5 * nfl * 1e-3
    assert abs(lat.reduce()[0] - m0) < 1e-6 * m0


def edge_channel_solid(nx, ny):
    """Walls on rows 0 and ny-1 and disks touching the inlet column, the
    middle and the outlet columns (so that the outlet column has both fluid
    and solid west neighbours)."""
    y, x = np.mgrid[0:ny, 0:nx]
    solid = (y == 0) | (y == ny - 1)
    for cx, cy, r in ((nx // 3, ny // 2, 9), (nx - 3, ny // 3, 5), (1, 2 * ny // 3, 4)):
        solid |= (x - cx) ** 2 + (y - cy) ** 2 <= r * r
    return solid


@pytest.mark.parametrize("storage", [L.STORAGE_F32, L.STORAGE_F16_DEV], ids=["f32", "f16"])
@pytest.mark.parametrize("nx, ny, x_edge", [(300, 200, 0), (2048, 40, 0), (512, 64, 0), (300, 200, 1),
                                            (258, 72, 1)])
def test_tile_variants_and_graphs_bit_identical(storage, nx, ny, x_edge):
    """Every tile variant that can step the lattice, with and without
    graphs, in both device layouts (row pitch 304 > nx = 300 in the
    row-interleaved one) gives the same populations bit for bit; the
    2048-wide lattice admits the lane-shuffle variants, and every lattice
    the two-step (LDS tile) variants, with tiles cut by the lattice edge
    where nx or ny is not a multiple of the tile (258 leaves a two-column
    last tile for every tile width).  The x_edge 1 cases are channels with
    the Zou-He inlet and the zero-gradient outlet, and disks on the inlet
    and outlet columns.  202 + 1 steps with graph chunks of 0, 100 or 102
    steps (51 two-step launches: an odd number of buffer flips per graph)
    and a final single step (the one-step fallback).  Compared at every
    slot a fluid cell can read: a two-step launch ends in the other buffer,
    whose solid-cell slots that no fluid cell pulls from hold older
    values."""
    print(f"case: variants + graphs {nx}x{ny} x_edge {x_edge} storage {storage}", flush=True)
    if x_edge:
        solid = edge_channel_solid(nx, ny)
        kw = dict(x_edge=L.X_INLET_OUTLET, inlet_ux=0.05, force=(2e-6, 0.0))
    else:
        solid = W.periodic_disks(nx, ny, 0.05, 4, 10, seed=4).solid
        kw = dict(force=(1e-5, 0.0))
    f0 = W.rest_state(nx, ny, solid, amplitude=1e-3).astype(np.float32)
    outs, names = [], []
    nv = len(L.variant_shapes())
    for layout in (L.LAYOUT_ROWS, L.LAYOUT_PLANES):
        for v in range(nv):
            if not L.variant_ok(v, nx, ny, x_edge):
                continue
            lat = L.Lattice(nx, ny, 0.58, layout=layout, storage=storage, **kw)
            lat.write_materials(W.materials_from_solid(solid))
            lat.write_populations(f0, 0)
            lat.set_variant(v)
            lat.set_graph_chunk((0, 100, 102)[v % 3])
            lat.run(202)
            lat.run(1)
            outs.append(lat.read_populations())
            names.append((layout, v, L.variant_shapes()[v]))
            lat.close()
    if nx % 1024 == 0:
        assert any(n[2][3] & 8 for n in names), "no lane-shuffle variant ran"
    assert sum(1 for n in names if n[2][3] & 16) >= 2 * 8, "two-step variants did not run"
    assert sum(1 for n in names if n[2][3] & 128) >= 2 * 4, "row-stream variants did not run"
    assert sum(1 for n in names if n[2][3] & 64) >= 2 * 4, "three-step variants did not run"
    assert np.isfinite(outs[0]).all()
    vis = visible_mask(solid, periodic_x=not x_edge)
    for k, o in enumerate(outs[1:], 1):
        bad = np.count_nonzero(o[vis] != outs[0][vis])
        assert bad == 0, f"variant {names[k]} differs from {names[0]} at {bad} visible slots"


def test_checkpoint_resume_bit_identical():
    print("case: checkpoint/resume 128x96", flush=True)
    nx, ny = 128, 96
    solid = W.cylinder_channel_solid(nx, ny, cx=40, cy=48, r=8)
    f0 = W.rest_state(nx, ny, solid, amplitude=1e-4).astype(np.float32)
    a = L.Lattice(nx, ny, 0.6, x_edge=L.X_INLET_OUTLET, inlet_ux=0.05)
    a.write_materials(W.materials_from_solid(solid))
    a.write_populations(f0, 0)
    a.run(150)
    ref = a.read_populations()
    b = L.Lattice(nx, ny, 0.6, x_edge=L.X_INLET_OUTLET, inlet_ux=0.05)
    b.write_materials(W.materials_from_solid(solid))
    b.write_populations(f0, 0)
    b.run(77)
    ck = b.save_state()
    fck = b.read_populations()
    c = L.Lattice(nx, ny, 0.6, x_edge=L.X_INLET_OUTLET, inlet_ux=0.05)
    c.write_materials(W.materials_from_solid(solid))
    c.load_state(ck)
    c.run(73)
    # every slot a fluid cell can read; solid slots no fluid cell reads keep
    # whatever each handle's reset left in its other buffer
    vis = visible_mask(solid, periodic_x=False)
    assert np.array_equal(c.read_populations()[vis], ref[vis])
    # Through the f-based exchange the state is rounded once (f = w + g in
    # binary32), so the resumed run stays within a few ulp.
    d = L.Lattice(nx, ny, 0.6, x_edge=L.X_INLET_OUTLET, inlet_ux=0.05)
    d.write_materials(W.materials_from_solid(solid))
    d.write_populations(fck, 0)
    d.run(73)
    fl = ~solid
    np.testing.assert_allclose(d.read_populations()[:, fl], ref[:, fl], rtol=0, atol=1e-6)
    # A state carries a header: one saved from a handle whose populations
    # mean something else (another layout, storage scale or size) is
    # rejected, not loaded as wrong values.
    planes = L.Lattice(nx, ny, 0.6, x_edge=L.X_INLET_OUTLET, inlet_ux=0.05, layout=L.LAYOUT_PLANES)
    with pytest.raises(L.LbmError) as ei:
        planes.load_state(ck)
    assert ei.value.code == 1
    assert "does not fit" in str(ei.value) or "bytes" in str(ei.value)
    h14 = L.Lattice(nx, ny, 0.6, storage=L.STORAGE_F16_DEV, dev_shift=14)
    st14 = h14.save_state()
    h10 = L.Lattice(nx, ny, 0.6, storage=L.STORAGE_F16_DEV)
    assert st14.size == h10.save_state().size
    with pytest.raises(L.LbmError) as ei:
        h10.load_state(st14)
    assert "dev_shift 14" in str(ei.value)
    bad = ck.copy()
    bad[:8] = 0
    with pytest.raises(L.LbmError) as ei:
        b.load_state(bad)
    assert "not a state" in str(ei.value)
    for h in (planes, h14, h10):
        h.close()


def test_f16_deviation_storage_vs_fp64():
    """binary16 storage of (f - w) 2^10 with binary32 arithmetic, periodic
    disks + force, 1000 steps: rho relative L2 <= 1e-5, u relative L2 <= 5e-2
    (the flow speed is ~1e-4, so binary16's 2^-11 relative step on the
    deviations is visible)."""
    print("case: f16 storage periodic disks 256x256 1000 steps", flush=True)
    nx = ny = 256
    scene = W.periodic_disks(nx, ny)
    f0 = W.rest_state(nx, ny, scene.solid)
    st, lat = physics_pair(nx, ny, 0.55, scene.solid, f0, 1000, force=(1e-6, 0.0), storage=L.STORAGE_F16_DEV)
    e = macro_errors(st, lat, scene.solid)
    print("f16 errors vs fp64:", e, flush=True)
    assert e["rho_l2"] < 1e-5 and e["u_l2"] < 5e-2


@pytest.mark.parametrize("n, storage, variant", [(4096, L.STORAGE_F32, None), (4096, L.STORAGE_F32, 46),
                                               (8192, L.STORAGE_F16_DEV, 38), (8192, L.STORAGE_F16_DEV, 42)],
                         ids=["4096-f32-default", "4096-f32-three-step", "8192-f16-row-stream",
                              "8192-f16-three-step"])
def test_full_size_mass_and_momentum_properties(n, storage, variant):
    """The roofline lattices at full size (periodic, disks >= 5 %,
    g = 1e-6, tau = 0.55; 4096^2 binary32 with the default one-step shape and
    the three-step 64 x 24 tile the benchmark tunes to, 8192^2 binary16
    deviations with the two shapes its tuning settles on, the row stream and
    the three-step 64 x 16 tile): the binary64 reduction shows mass conserved
    over 100 steps and the momentum identity P(t+1) - P(t) = g N - drag for
    one more step.  The identity holds to the rounding of one step, bounded
    from the state itself (momentum_rounding_bound): every stored deviation
    is rounded once per step, and in the nearly uniform early flow the
    roundings of all cells can share a sign, so the bound adds them up.
    Measured: 1.2e-6 of g N in binary32, 2.2e-3 in binary16, a few percent
    of the bound."""
    print(f"case: {n}x{n} properties 100 steps, storage {storage}, variant {variant}", flush=True)
    scene = W.periodic_disks(n, n)
    g = 1e-6
    lat = L.Lattice(n, n, 0.55, force=(g, 0.0), storage=storage)
    lat.write_materials(W.materials_from_solid(scene.solid))
    lat.reset()
    if variant is not None:
        lat.set_variant(variant)
    nfl = int((~scene.solid).sum())
    r0 = lat.reduce()
    lat.run(100)
    r1 = lat.reduce()
    assert np.isfinite(r1).all()
    assert abs(r1[0] - r0[0]) < 1e-6 * r0[0]
    lat.run(1)
    r2 = lat.reduce()
    residual = abs((r2[1] - r1[1]) - (g * nfl - r2[3]))
    bound = momentum_rounding_bound(lat.read_populations(), scene.solid, storage)
    print(f"  mass drift {abs(r1[0] - r0[0]) / r0[0]:.2e}, momentum identity residual {residual / (g * nfl):.2e} "
          f"of g N, {residual / bound:.2e} of its rounding bound ({bound / (g * nfl):.2e} of g N)", flush=True)
    assert residual <= bound
    assert bound < 0.1 * g * nfl  # the identity is still a sharp check of force and drag


@pytest.mark.parametrize("n, storage, variant", [(4096, L.STORAGE_F32, 46), (8192, L.STORAGE_F16_DEV, 38),
                                               (8192, L.STORAGE_F16_DEV, 42)],
                         ids=["4096-f32-three-step", "8192-f16-row-stream", "8192-f16-three-step"])
def test_full_size_multi_step_shapes_match_one_step(n, storage, variant):
    """At the benchmark's own sizes the shapes it tunes to give the one-step
    kernel's populations bit for bit: 30 steps from the same seeded disk
    lattice (10 three-step or 15 two-step launches against 30 one-step
    launches) and one step more, compared at every slot a fluid cell can
    read."""
    print(f"case: {n}x{n} shape {variant} against one step, storage {storage}", flush=True)
    scene = W.periodic_disks(n, n)
    one_step = next(v for v, s in enumerate(L.variant_shapes()) if not s[3] & (16 | 64))
    outs = []
    for v in (one_step, variant):
        lat = L.Lattice(n, n, 0.55, force=(1e-6, 0.0), storage=storage)
        lat.write_materials(W.materials_from_solid(scene.solid))
        lat.reset()
        lat.set_variant(v)
        lat.run(30)
        lat.run(1)
        outs.append(lat.read_populations())
        lat.close()
    vis = visible_mask(scene.solid, periodic_x=True)
    assert np.isfinite(outs[0][vis]).all()
    bad = np.count_nonzero(outs[1][vis] != outs[0][vis])
    assert bad == 0, f"shape {variant} differs from the one-step kernel at {bad} of {int(vis.sum())} slots"


def momentum_rounding_bound(f, solid, storage, shift=10):
    """Bound on |P_x(t+1) - P_x(t) - (g N - drag)| from one step's rounding,
    from the state after the step (f, dense binary32 populations): each
    stored deviation d_i = f_i - w_i is rounded once -- binary16 (scaled by
    2^shift): relative 2^-11 in the normal range, absolute 2^-25 2^-shift
    below it; binary32: 2^-24 -- and contributes e_ix times its error to the
    momentum; the binary32 collision adds at most a few roundings of 2^-24
    of the deviations, the equilibrium and the velocity it forms (bounded by
    64 2^-24 sum |e_x| |d|).  With every error of one sign:
        u_store S + N 6 a_sub + 64 2^-24 S,   S = sum over fluid cells of
        sum_i |e_ix| |d_i|."""
    fl = ~solid
    S = 0.0
    for i in range(9):
        if W.EX[i] != 0:
            S += float(np.abs(f[i][fl].astype(np.float64) - W.W_PHYS[i]).sum())
    nfl = int(fl.sum())
    if storage == L.STORAGE_F16_DEV:
        return 2.0**-11 * S + nfl * 6 * 2.0 ** (-25 - shift) + 64 * 2.0**-24 * S
    return 2.0**-24 * S + 64 * 2.0**-24 * S