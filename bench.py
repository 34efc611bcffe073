#!/usr/bin/env python3
"""Throughput of the D2Q9 hot path on one MI355X.

Workload: plain D2Q9 BGK at 4096x4096 in binary32, tau = 0.55, periodic in
both axes, Guo body force g = (1e-6, 0), seeded non-overlapping disks
(radius 4..24) on >= 5 % of the cells with bounce-back, starting from rest.
One step = one fused collide-stream pass over the whole lattice.  Steps run
as 100-step HIP-graph chunks after the tile shape has been autotuned.

Prints one JSON line: MLUPS (all cells, solid included), the HBM roofline of
the step kernel (72 algorithmic bytes per cell per step over the kernel's
average duration from HIP events on its stream) and a CPU baseline: the
binary32 restatement of the reference's WGSL (collide_stream + boundary
passes, OpenMP over rows) timed on this host on a bounded sample of the same
lattice without the body force, which the reference cannot express.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
  python bench.py --copy-calibration   # plain copy kernel, for byte counters
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS + % HBM roofline, D2Q9 plain vs learned-collision at fixed vel L2 err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_F32 = 72  # 9 populations read + 9 written, binary32


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(workload: str, variant: int):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel calibration), as committed in
    profiles/<workload>_v<variant>_pmc.md for the step-kernel variant used, or
    None."""
    p = ROOT / "profiles" / f"{workload}_v{variant}_pmc.md"
    if not p.exists():
        return None
    text = p.read_text()
    v = re.search(r"variant:\s*(\d+)", text)
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", text)
    if not m or not v or int(v.group(1)) != variant:
        return None
    return float(m.group(1))


def cpu_baseline(scene, budget_s: float = 12.0):
    """The reference's algorithm on the host cores: binary32 restatement of
    collide_stream.wgsl + boundary.wgsl on the same 4096^2 disk lattice
    (materials 1 / 4, no body force), timed over whole steps until about
    budget_s seconds of work."""
    from oracle import oracle as O
    from simuverse_amd import workloads as W

    ny, nx = scene.solid.shape
    info = np.zeros((ny, nx), O.INFO_DTYPE)
    info["material"] = np.where(scene.solid, 4, 1)
    info["block_iter"] = -1
    st = O.WgslState(info, 0.55)
    st.init(fluid_ty=1)
    st.step(0)  # first touch of the buffers
    steps, t0, swap = 0, time.perf_counter(), 1
    while True:
        st.step(swap)
        swap ^= 1
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": round(nx * ny * steps / el / 1e6, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the 4096x4096 disk lattice (materials 1/4, no body force), "
                  f"{el:.1f} s, binary32 WGSL restatement with OpenMP over rows",
    }


def describe_variant(shape):
    """The step-kernel variant as a dict: one-step shapes (block,
    cells_per_thread, vec) or the two-step LDS tile (rows x width)."""
    b, c, v, f = shape
    flags = [name for bit, name in ((1, "nt_store"), (2, "nt_load"), (4, "xcd_rows"), (8, "lane_shuffle"),
                                    (16, "two_step")) if f & bit]
    if f & 16:
        return {"kernel": "k_phys_step2", "block": b, "tile_rows": c, "tile_width": v, "flags": flags,
                "steps_per_launch": 2}
    kern = "k_phys_step_shfl" if f & 8 else ("k_phys_step4" if v == 4 else "k_phys_step")
    return {"kernel": kern, "block": b, "cells_per_thread": c, "vec": v, "flags": flags, "steps_per_launch": 1}


def run_copy_calibration(nbytes: int):
    from simuverse_amd import lbm as L

    ms = L.copy_bandwidth(nbytes, reps=20)
    line = {"kernel": "k_copy16", "bytes_read": nbytes, "bytes_written": nbytes, "ms_per_copy": ms,
            "GBps": 2 * nbytes / ms / 1e6}
    print(json.dumps(line), flush=True)


def run_batch(args):
    """512 seeded uneven channel domains (sizes 64..1024, own tau, inlet ux,
    0..8 disks) stepped by one launch per step (persistent chunks) or per
    two steps (one block per LDS tile), whichever shape the autotuner picks,
    in 100-step HIP graphs; aggregate MLUPS over all cells."""
    from simuverse_amd import lbm as L
    from simuverse_amd import workloads as W

    t = time.perf_counter()
    doms = W.batch_domains()
    log(f"batch: {len(doms)} domains, {sum(d.nx * d.ny for d in doms)} cells ({time.perf_counter()-t:.1f} s)")
    storage = L.STORAGE_F16_DEV if args.f16 else L.STORAGE_F32
    b = L.Batch([dict(nx=d.nx, ny=d.ny, tau=d.tau, x_edge=L.X_INLET_OUTLET, inlet_ux=d.inlet_ux) for d in doms],
                storage=storage)
    for k, d in enumerate(doms):
        b.write_materials(k, W.materials_from_solid(d.solid))
    b.set_graph_chunk(args.graph_chunk)
    if args.variant >= 0:
        b.set_variant(args.variant)
    else:
        v, ms = b.autotune(steps=100)
        log(f"batch autotune: variant {v} {ms:.3f} ms/step")
    b.run(args.warmup)
    b.synchronize()
    K = args.steps
    ms = b.run_timed(K)
    cells = b.cells()
    step_s = ms / 1e3 / K
    bpc = 36 if args.f16 else 72
    achieved = bpc * cells / step_s / 1e9
    variant = b.variant()
    block, vec, tile_h, flags = L.batch_variant_shapes()[variant]
    per_launch = 2 if flags & 16 else 1
    workload = f"d2q9_batch512_{'f16dev' if args.f16 else 'f32'}"
    line = {
        "metric": METRIC,
        "value": round(cells / step_s / 1e6, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms / K, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-deviation storage, f32 math" if args.f16 else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "domains": len(doms), "cells": int(cells), "variant": variant,
                   "batch_variant": {"kernel": "k_batch_step2" if per_launch == 2 else "k_batch_step",
                                     "block": block, "vec": vec, "tile_rows": tile_h, "flags": flags},
                   "graph_chunk": args.graph_chunk, "seed": W.SEED},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(workload, variant),
                     "steps_per_launch": per_launch, "algorithmic_bytes_per_launch": bpc * cells * per_launch},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_d3q19(args):
    """D3Q19 on an n^3 periodic lattice (n = --n3, default 256): tau 0.55,
    Guo force g = (1e-6, 0, 0), seeded non-overlapping spheres of radius
    4..12 to 5 % solid, from rest; the step-kernel shape is autotuned, then
    K steps in 100-step HIP graphs.  Algorithmic traffic 152 B per cell per
    step (19 x 4 B read + 19 x 4 B written; 76 B with binary16 deviations)."""
    from simuverse_amd import lbm as L
    from simuverse_amd import workloads as W

    n = args.n3
    tau, g = 0.55, 1e-6
    t = time.perf_counter()
    scene = W.periodic_spheres(n, n, n)
    log(f"d3q19 scene {n}^3: {len(scene.spheres)} spheres, solid {scene.solid_fraction:.4f} "
        f"({time.perf_counter() - t:.1f} s)")
    storage = L.STORAGE_F16_DEV if args.f16 else L.STORAGE_F32
    lat = L.Lattice3D(n, n, n, tau, storage=storage, force=(g, 0.0, 0.0))
    lat.write_materials(np.where(scene.solid, 4, 1).astype(np.uint8))
    lat.set_graph_chunk(args.graph_chunk)
    if args.variant >= 0:
        lat.set_variant(args.variant)
    elif not args.no_autotune:
        v, ms = lat.autotune(steps=50)
        log(f"d3q19 autotune: variant {v} {L.variant3_shapes()[v]} {ms * 1e3:.1f} us/step")
    variant = lat.variant()
    lat.run(args.warmup)
    lat.synchronize()
    K = args.steps
    ms = lat.run_timed(K)
    cells = n ** 3
    step_s = ms / 1e3 / K
    bpc = 76 if args.f16 else 152
    achieved = bpc * cells / step_s / 1e9
    r = lat.reduce()
    log(f"{K} steps: {ms:.1f} ms, {step_s * 1e6:.2f} us/step, {cells / step_s / 1e6:.0f} MLUPS, {achieved:.0f} GB/s; "
        f"mass {r[0]:.9e} drag ({r[4]:.3e}, {r[5]:.3e}, {r[6]:.3e})")
    if not np.isfinite(r).all():
        raise SystemExit("non-finite state after the timed run")
    workload = f"d3q19_{'f16dev' if args.f16 else 'f32'}_{n}cubed"
    block, cpt, flags = L.variant3_shapes()[variant]
    line = {
        "metric": METRIC,
        "value": round(cells / step_s / 1e6, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms / K, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-deviation storage, f32 math" if args.f16 else "f32",
        "data": "synthetic",
        "config": {"workload": workload, "nx": n, "ny": n, "nz": n, "tau": tau, "body_force": [g, 0.0, 0.0],
                   "solid_fraction": round(scene.solid_fraction, 5), "spheres": int(len(scene.spheres)),
                   "seed": W.SEED, "variant": variant,
                   "tile_variant": {"kernel": "k3_step", "block": block, "cells_per_thread": cpt, "flags": flags},
                   "graph_chunk": args.graph_chunk},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(workload, variant),
                     "steps_per_launch": 1, "algorithmic_bytes_per_launch": bpc * cells},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_learned(args):
    """Learned collision on a 4x coarser lattice against the fine solution.

    Fine: n x n binary32, tau_f = 0.52 (nu = 1/150), periodic 4x4 array of
    cylinders of radius n/64, started from a uniform flow U = 0.05 with a Guo
    body force g_f holding it (Re = U D / nu = 0.05 * n/32 * 150, 1920 at
    n = 8192).  Coarse: (n/4)^2, tau_c = 0.505 (nu_c = nu_f / 4, the same
    flow in acoustic scaling), force 4 g_f, majority-solid restriction.

    1. fine run of N_f steps (timed); every `pair_every` fine steps the state
       and the one 4 fine steps later are restricted and their transition
       goes into the network's sample store (inputs: non-equilibrium part of
       the coarse populations; target: the residual of one coarse BGK step);
    2. Adam training on those samples (binary32, device);
    3. a validation gate: the gain on the correction (1, 1/2, 1/4 or 0, plain
       BGK) with the smallest coarse velocity error at N_f / 2, the end of
       the sampled half of the window;
    4. coarse runs of N_f / 4 steps from the restricted fine initial state,
       plain BGK and BGK + the gated network on the matrix cores (timed);
    5. velocity relative L2 error of each coarse run against the restricted
       fine velocity at N_f, and time-to-solution ratios.
    """
    from simuverse_amd import lbm as L
    from simuverse_amd import workloads as W

    n = args.n if args.n != 4096 else 8192
    nc = n // 4
    tau_f, tau_c, U = 0.52, 0.505, 0.05
    gf = args.learned_force
    gc = (4 * gf, 0.0)
    Nf = args.learned_fine_steps
    solid = W.cylinder_array_solid(n, 4, n // 64)
    log(f"learned: fine {n}^2 tau {tau_f}, coarse {nc}^2 tau {tau_c}, {Nf} fine steps, "
        f"solid {solid.mean():.4f}, Re {U * (n // 32) / ((tau_f - 0.5) / 3):.0f}")
    fine = L.Lattice(n, n, tau_f, force=(gf, 0.0))
    fine.write_materials(W.materials_from_solid(solid))
    fine.write_populations(W.equilibrium_state(n, n, U, 0.0, solid).astype(np.float32), 0)
    del solid
    fine.autotune(steps=10)
    # the timed fine solution: the whole N_f steps
    fc0, solid_c, _, _ = fine.restrict4(tau_c, gc)
    ms_fine = fine.run_timed(Nf)
    _, _, uxr, uyr = fine.restrict4(tau_c, gc)
    log(f"fine: {Nf} steps {ms_fine:.1f} ms ({n * n * Nf / ms_fine / 1e3:.0f} MLUPS)")

    # training data from a second pass over the first half of the window
    t0 = time.perf_counter()
    net = L.Network(seed=7, capacity=args.learned_capacity)
    fine2 = L.Lattice(n, n, tau_f, force=(gf, 0.0))
    solid = W.cylinder_array_solid(n, 4, n // 64)
    fine2.write_materials(W.materials_from_solid(solid))
    fine2.write_populations(W.equilibrium_state(n, n, U, 0.0, solid).astype(np.float32), 0)
    del solid
    fine2.set_variant(fine.variant())
    done, pairs = 0, 0
    net.set_scales(1.0, 1.0)  # normalize_samples sets them from the samples
    stride = max(1, (nc * nc * (Nf // 2 // args.pair_every)) // args.learned_capacity + 1)
    # transitions from the second quarter of the window on (past the start-up
    # transient) to its middle
    for t in range(Nf // 8, Nf // 2, args.pair_every):
        fine2.run(t - done)
        done = t
        a, _, _, _ = fine2.restrict4(tau_c, gc)
        fine2.run(4)
        done += 4
        b, _, _, _ = fine2.restrict4(tau_c, gc)
        net.add_transition(a, b, solid_c, tau_c, gc, stride=stride, offset=pairs % stride)
        pairs += 1
    # the restricted fine velocity at N_f / 2: the validation target
    if done <= Nf // 2:
        fine2.run(Nf // 2 - done)
    _, _, uxh, uyh = fine2.restrict4(tau_c, gc)
    fine2.close()
    rx, rt = net.normalize_samples()
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = net.train(steps=args.learned_train_steps, batch=min(args.learned_batch, net.sample_count()), lr=2e-3,
                     seed=3)
    t_train = time.perf_counter() - t0
    log(f"samples {net.sample_count()} from {pairs} transitions ({t_collect:.1f} s), "
        f"train {args.learned_train_steps} steps ({t_train:.1f} s), loss {loss:.3e}")

    def coarse(learned: bool, steps: int, ref_u):
        lat = L.Lattice(nc, nc, tau_c, force=gc)
        lat.write_materials(W.materials_from_solid(solid_c))
        lat.write_populations(fc0, 0)
        if learned:
            lat.set_learned_collision(net)
        else:
            lat.autotune(steps=10)
            lat.write_populations(fc0, 0)
        ms = lat.run_timed(steps)
        _, ux, uy = lat.macro()
        u, finite = lat.health()
        fl = ~solid_c
        ur, vr = ref_u
        ref = np.hypot(ur, vr)[fl]
        err = float(np.linalg.norm(np.hypot(ux - ur, uy - vr)[fl]) / np.linalg.norm(ref)) if finite else float("inf")
        lat.close()
        return ms, err

    # Validation gate: the correction's gain (a factor on the output scale)
    # is chosen on the first half of the window, coarse runs to N_f / 2
    # against the restricted fine velocity there; 0 is plain BGK.  The
    # reported errors are at N_f, which no sample or gate run has seen.
    s_in, s_out = net.scales()
    gate = {}
    for gain in (1.0, 0.5, 0.25, 0.0):
        if gain > 0.0:
            net.set_scales(s_in, s_out * gain)
        gate[gain] = coarse(gain > 0.0, Nf // 8, (uxh, uyh))[1]
    gain = min(gate, key=lambda k: (gate[k], -k))
    log("validation gate (vel L2 at N_f/2 by gain): " + ", ".join(f"{k:g}: {v:.4e}" for k, v in gate.items())
        + f"; chosen gain {gain:g}")

    ms_plain, e_plain = coarse(False, Nf // 4, (uxr, uyr))
    # the learned kernel at the chosen gain; at gain 0 the gate falls back to
    # plain BGK and the learned kernel is timed at gain 1 for reference
    net.set_scales(s_in, s_out * (gain if gain > 0.0 else 1.0))
    ms_learn, e_learn = coarse(True, Nf // 4, (uxr, uyr))
    e_gated, ms_gated = (e_learn, ms_learn) if gain > 0.0 else (e_plain, ms_plain)
    log(f"coarse plain: {ms_plain:.1f} ms, vel L2 {e_plain:.4e}; learned (gain {gain if gain > 0 else 1:g}): {ms_learn:.1f} ms, "
        f"vel L2 {e_learn:.4e}")
    K = Nf // 4
    line = {
        "metric": METRIC,
        "value": round(nc * nc * K / ms_learn / 1e3, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": 0,
        "ms_per_step": round(ms_learn / K, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, bf16 network on matrix cores (f32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": f"d2q9_learned_{n}_to_{nc}",
            "fine": {"n": n, "tau": tau_f, "steps": Nf, "ms": round(ms_fine, 3)},
            "coarse": {"n": nc, "tau": tau_c, "steps": K},
            "force_fine": gf, "U0": U, "cylinders": "4x4, r = n/64",
            "training": {"pairs": pairs, "samples": net.sample_count(), "steps": args.learned_train_steps,
                         "batch": min(args.learned_batch, net.sample_count()), "final_loss": loss,
                         "collect_s": round(t_collect, 2), "train_s": round(t_train, 2)},
            "validation_gate": {"vel_l2_at_half_window_by_gain": {f"{k:g}": v for k, v in gate.items()},
                                "gain": gain, "vel_l2": e_gated, "coarse_ms": round(ms_gated, 3),
                                "speedup_vs_fine": round(ms_fine / ms_gated, 2)},
            "vel_l2_vs_restricted_fine": {"plain": e_plain, "learned": e_learn},
            "coarse_ms": {"plain": round(ms_plain, 3), "learned": round(ms_learn, 3)},
            "time_to_solution_speedup": {"learned_vs_fine": round(ms_fine / ms_learn, 2),
                                         "learned_incl_training_vs_fine":
                                             round(ms_fine / (ms_learn + 1e3 * (t_collect + t_train)), 3)},
        },
        "roofline": {"bound": "hbm", "achieved": round(72 * nc * nc * K / ms_learn / 1e6, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(72 * nc * nc * K / ms_learn / 1e6 / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--graph-chunk", type=int, default=100, help="steps per HIP graph (0: direct launches)")
    ap.add_argument("--workload", choices=["single", "batch", "learned", "d3q19"], default="single")
    ap.add_argument("--n3", type=int, default=256, help="d3q19 workload: lattice edge")
    ap.add_argument("--f16", action="store_true", help="binary16 deviation storage")
    ap.add_argument("--learned-force", type=float, default=1e-7, help="learned workload: fine body force")
    ap.add_argument("--learned-fine-steps", type=int, default=4000)
    ap.add_argument("--pair-every", type=int, default=100, help="fine steps between training transitions")
    ap.add_argument("--learned-capacity", type=int, default=4_000_000)
    ap.add_argument("--learned-train-steps", type=int, default=300)
    ap.add_argument("--learned-batch", type=int, default=1_000_000)
    ap.add_argument("--layout", choices=["rows", "planes"], default="rows",
                    help="device population layout: row-interleaved directions or one plane per direction")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the lattice would shard by "
              "y-slabs with a one-row halo exchange per step)", file=sys.stderr)
        sys.exit(2)

    from simuverse_amd import lbm as L
    from simuverse_amd import workloads as W

    if args.copy_calibration:
        run_copy_calibration(args.copy_bytes)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "learned":
        run_learned(args)
        return
    if args.workload == "d3q19":
        run_d3q19(args)
        return

    n = args.n
    tau, g = 0.55, 1e-6
    t = time.perf_counter()
    scene = W.periodic_disks(n, n)
    log(f"scene {n}x{n}: {len(scene.disks)} disks, solid {scene.solid_fraction:.4f} ({time.perf_counter()-t:.1f} s)")
    storage = L.STORAGE_F16_DEV if args.f16 else L.STORAGE_F32
    layout = L.LAYOUT_ROWS if args.layout == "rows" else L.LAYOUT_PLANES
    lat = L.Lattice(n, n, tau, mode=L.MODE_PHYSICS, storage=storage, force=(g, 0.0), layout=layout)
    lat.write_materials(W.materials_from_solid(scene.solid))
    lat.reset()
    lat.set_graph_chunk(args.graph_chunk)
    if args.variant >= 0:
        lat.set_variant(args.variant)
    elif not args.no_autotune:
        v, ms = lat.autotune(steps=200)
        log(f"autotune: variant {v} {L.variant_shapes()[v]} {ms*1e3:.1f} us/step")
    variant = lat.variant()

    lat.run(args.warmup)
    lat.synchronize()
    K = args.steps
    wall0 = time.perf_counter()
    ms = lat.run_timed(K)  # HIP events on the library's stream around K steps
    lat.synchronize()
    wall = time.perf_counter() - wall0
    cells = n * n
    sec = ms / 1e3
    mlups = cells * K / sec / 1e6
    step_us = ms * 1e3 / K
    bytes_per_cell = 36 if args.f16 else BYTES_PER_CELL_F32
    achieved = bytes_per_cell * cells / (step_us * 1e-6) / 1e9
    workload = f"d2q9_phys_{'f16dev' if args.f16 else 'f32'}_{n}x{n}"
    r = lat.reduce()
    log(f"{K} steps: {ms:.1f} ms events, {wall*1e3:.1f} ms wall, {step_us:.2f} us/step, {mlups:.0f} MLUPS, "
        f"{achieved:.0f} GB/s; mass {r[0]:.9e} drag ({r[3]:.3e}, {r[4]:.3e})")
    if not np.isfinite(r).all():
        raise SystemExit("non-finite state after the timed run")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene)
        log(f"cpu baseline: {cpu}")

    traffic = traffic_from_profiles(workload, variant)
    per_launch = describe_variant(L.variant_shapes()[variant])["steps_per_launch"]
    line = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms / K, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-deviation storage, f32 math" if args.f16 else "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "nx": n,
            "ny": n,
            "tau": tau,
            "body_force": [g, 0.0],
            "boundaries": "periodic x and y, half-way bounce-back on disks",
            "solid_fraction": round(scene.solid_fraction, 5),
            "disks": int(len(scene.disks)),
            "seed": W.SEED,
            "layout": args.layout,
            "tile_variant": describe_variant(L.variant_shapes()[variant]),
            "graph_chunk": args.graph_chunk,
            "wall_ms": round(wall * 1e3, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "steps_per_launch": per_launch,
            "algorithmic_bytes_per_launch": bytes_per_cell * cells * per_launch,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
