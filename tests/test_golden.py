"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_fixtures.py from the oracle) against the oracle as it is
now (CPU) and against the gfx950 library (GPU)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from simuverse_amd import lbm as L
from simuverse_amd import workloads as W

HERE = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(HERE / name) as z:
        return {k: z[k] for k in z.files}


def parity_info(fx):
    info = np.zeros(fx["info_material"].shape, O.INFO_DTYPE)
    info["material"] = fx["info_material"]
    info["block_iter"] = -1
    info["vx"] = fx["info_vx"]
    return info


def test_oracle_reproduces_parity_fixture():
    fx = load("parity_24x20.npz")
    st = O.WgslState(parity_info(fx), 0.6)
    st.buf[0][:] = fx["f0"]
    st.buf[1][:] = fx["f1"]
    done = 0
    for s in (1, 10, 100):
        st.run(s - done, start_swap=done & 1)
        done = s
        np.testing.assert_array_equal(st.buf[0], fx[f"buf0_{s}"])
        np.testing.assert_array_equal(st.buf[1], fx[f"buf1_{s}"])
        np.testing.assert_array_equal(st.macro, fx[f"macro_{s}"])


@pytest.mark.parametrize("case", ["channel", "box"])
def test_oracle_reproduces_physics_fixture(case):
    fx = load("physics_24x20.npz")
    ny, nx = fx[f"{case}_solid"].shape
    params = (O.phys_params(nx, ny, 0.6, x_edge=1, inlet_ux=0.04) if case == "channel"
              else O.phys_params(nx, ny, 0.55, force=(2e-5, -1e-5)))
    st = O.PhysState(params, fx[f"{case}_solid"], fx[f"{case}_f0"])
    done = 0
    for s in (1, 10, 100):
        st.run(s - done)
        done = s
        np.testing.assert_array_equal(st.f, fx[f"{case}_{s}"])


@pytest.mark.gpu
def test_gpu_matches_parity_fixture_bit_exact():
    fx = load("parity_24x20.npz")
    info = parity_info(fx)
    ny, nx = info.shape
    lat = L.Lattice(nx, ny, 0.6, mode=L.MODE_PARITY_WGSL)
    lat.write_lattice_info(info)
    lat.write_populations(fx["f0"], 0)
    lat.write_populations(fx["f1"], 1)
    solid = np.isin(info["material"], (2, 4))
    done = 0
    for s in (1, 10, 100):
        lat.run(s - done, export_macro=True)
        done = s
        rho, ux, uy = lat.read_macro()
        np.testing.assert_array_equal(np.stack([ux, uy, rho], -1), fx[f"macro_{s}"][..., :3])
        cur = lat.current_buffer()
        got = lat.read_populations(cur)
        np.testing.assert_array_equal(got[:, ~solid], fx[f"buf{cur}_{s}"][:, ~solid])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["channel", "box"])
def test_gpu_matches_physics_fixture(case):
    """binary32 deviation storage against the binary64 fixture: populations
    of fluid cells within 1e-6 relative plus 2e-7 absolute (about two
    binary32 ulp of the density, which every f_i inherits through feq,
    after 100 steps)."""
    fx = load("physics_24x20.npz")
    solid = fx[f"{case}_solid"]
    ny, nx = solid.shape
    kw = dict(x_edge=L.X_INLET_OUTLET, inlet_ux=0.04) if case == "channel" else dict(force=(2e-5, -1e-5))
    lat = L.Lattice(nx, ny, 0.6 if case == "channel" else 0.55, **kw)
    lat.write_materials(W.materials_from_solid(solid))
    lat.write_populations(fx[f"{case}_f0"].astype(np.float32), 0)
    done = 0
    for s in (1, 10, 100):
        lat.run(s - done)
        done = s
        got = lat.read_populations().astype(np.float64)
        np.testing.assert_allclose(got[:, ~solid], fx[f"{case}_{s}"][:, ~solid], rtol=1e-6, atol=2e-7)
