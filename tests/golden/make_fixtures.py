"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

The reference (Rust + WGSL through wgpu) ships no golden vectors for the
LBM path and cannot run in this image, so these fixtures pin regressions of
the oracle and of the GPU kernels against a frozen oracle output; external
parity is argued by the oracle's own pinning tests.

  parity_24x20.npz   reference-semantics channel (walls, ghosts, accelerate
                     inlet column vx = 0.05, cylinder r = 4 at (8, 10)),
                     tau = 0.6, w + U(-1e-4, 1e-4): both buffers and the
                     macro field after 1, 10 and 100 steps (binary32)
  physics_24x20.npz  physics channel (Zou-He inlet 0.04, zero-gradient
                     outlet, same walls and cylinder), tau = 0.6, and a
                     periodic box with a disk and Guo force (2e-5, -1e-5),
                     tau = 0.55: post-collision populations after 1, 10 and
                     100 steps (binary64)

Run:  python tests/golden/make_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import oracle as O  # noqa: E402
from simuverse_amd import workloads as W  # noqa: E402

NX, NY = 24, 20
STEPS = (1, 10, 100)


def parity_fixture():
    info = W.cylinder_parity_info(NX, NY, inlet_vx=0.05, cx=8, cy=10, r=4)
    f0, f1 = W.parity_initial_state(info, seed=7)
    out = {"info_material": info["material"], "info_vx": info["vx"], "f0": f0, "f1": f1}
    st = O.WgslState(info, 0.6)
    st.buf[0][:] = f0
    st.buf[1][:] = f1
    done = 0
    for s in STEPS:
        st.run(s - done, start_swap=done & 1)
        done = s
        out[f"buf0_{s}"] = st.buf[0].copy()
        out[f"buf1_{s}"] = st.buf[1].copy()
        out[f"macro_{s}"] = st.macro.copy()
    return out


def physics_fixture():
    y, x = np.mgrid[0:NY, 0:NX]
    chan = (y == 0) | (y == NY - 1) | ((x - 8) ** 2 + (y - 10) ** 2 <= 16)
    box = (x - 12) ** 2 + (y - 9) ** 2 <= 9
    out = {"channel_solid": chan, "box_solid": box}
    for name, solid, params in (
        ("channel", chan, O.phys_params(NX, NY, 0.6, x_edge=1, inlet_ux=0.04)),
        ("box", box, O.phys_params(NX, NY, 0.55, force=(2e-5, -1e-5))),
    ):
        f0 = W.rest_state(NX, NY, solid, amplitude=1e-4, seed=9)
        out[f"{name}_f0"] = f0
        st = O.PhysState(params, solid, f0)
        done = 0
        for s in STEPS:
            st.run(s - done)
            done = s
            out[f"{name}_{s}"] = st.f.copy()
    return out


if __name__ == "__main__":
    np.savez_compressed(HERE / "parity_24x20.npz", **parity_fixture())
    np.savez_compressed(HERE / "physics_24x20.npz", **physics_fixture())
    for p in sorted(HERE.glob("*.npz")):
        print(p.name, p.stat().st_size)
