"""Storage precision of the physics mode against the binary64 oracle.

The roofline scene's physics (periodic in x and y, tau = 0.55, Guo body
force g = (1e-6, 0), seeded disks of radius 4..24 on >= 5 % of the cells,
from rest) at 512^2, where the CPU oracle runs 10 000 steps in under a
minute: binary32 storage and binary16 deviation storage (f_i - w_i times
2^10), both with binary32 arithmetic, against the binary64 oracle after
1000 and 10 000 steps.  Each checkpoint prints rho and u relative L2 and max
errors over fluid cells; the bounds below are what the storage precision
allows:
  binary32: deviations ~1e-3 carry ~1e-10 rounding, the body force adds
            1e-7 per step, so u (~1e-4 .. 1e-3) keeps ~1e-4 relative;
  binary16: 11 significant bits on deviations (2^-11 ~ 5e-4 relative of
            each), accumulating over the steps to ~1e-3 in u.
Measured on MI355X: binary32 u rel L2 7.6e-5 / 2.7e-4 and binary16
4.5e-4 / 1.3e-3 at 1000 / 10 000 steps (rho 1.3e-7 / 1.8e-7 and
4.3e-7 / 4.1e-6).
"""
import json

import numpy as np
import pytest

from oracle import oracle as O
from simuverse_amd import lbm as L
from simuverse_amd import workloads as W

pytestmark = pytest.mark.gpu


def errors(lat, ref, fl):
    r0, u0, v0 = ref
    r1, u1, v1 = (a.astype(np.float64) for a in lat.macro())
    du = np.hypot(u1 - u0, v1 - v0)[fl]
    umag = np.hypot(u0, v0)[fl]
    return {
        "rho_l2": float(np.linalg.norm(r1[fl] - r0[fl]) / np.linalg.norm(r0[fl])),
        "rho_max": float(np.max(np.abs(r1[fl] - r0[fl]))),
        "u_l2": float(np.linalg.norm(du) / np.linalg.norm(umag)),
        "u_max": float(np.max(du)),
        "u_max_ref": float(np.max(umag)),
    }


def test_f16_and_f32_storage_vs_fp64_at_1000_and_10000_steps():
    n, tau, g = 512, 0.55, 1e-6
    print(f"case: precision study {n}^2, 1000 and 10000 steps", flush=True)
    scene = W.periodic_disks(n, n)
    fl = ~scene.solid
    f0 = W.rest_state(n, n, scene.solid)
    ref = O.PhysState(O.phys_params(n, n, tau, force=(g, 0.0)), scene.solid, f0)
    lats = {}
    for name, storage in (("f32", L.STORAGE_F32), ("f16dev", L.STORAGE_F16_DEV)):
        lat = L.Lattice(n, n, tau, storage=storage, force=(g, 0.0))
        lat.write_materials(W.materials_from_solid(scene.solid))
        lat.write_populations(f0.astype(np.float32), 0)
        lats[name] = lat
    done, rows = 0, []
    for c in (1000, 10000):
        ref.run(c - done)
        m = ref.macro()
        row = {"steps": c}
        for name, lat in lats.items():
            lat.run(c - done)
            row[name] = errors(lat, m, fl)
        done = c
        rows.append(row)
        print("precision", json.dumps(row), flush=True)
    for row in rows:
        assert row["f32"]["rho_l2"] < 1e-6 and row["f32"]["u_l2"] < 1e-3
        assert row["f16dev"]["rho_l2"] < 1e-4 and row["f16dev"]["u_l2"] < 1e-2
        # binary16 storage is the coarser of the two
        assert row["f16dev"]["u_l2"] > row["f32"]["u_l2"]
