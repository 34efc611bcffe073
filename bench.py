#!/usr/bin/env python3
"""Throughput of the D2Q9 hot path on one MI355X.

Workload: plain D2Q9 BGK at 4096x4096 in binary32, tau = 0.55, periodic in
both axes, Guo body force g = (1e-6, 0), seeded non-overlapping disks
(radius 4..24) on >= 5 % of the cells with bounce-back, starting from rest.
One step = one fused collide-stream pass over the whole lattice.  Steps run
as 100-step HIP-graph chunks after the tile shape has been autotuned.

Prints one JSON line: MLUPS (all cells, solid included), the HBM roofline of
the step kernel, the rho/u error against the binary64 oracle and a CPU
baseline.

  roofline   one launch of the step kernel must read and write the lattice
             once, 72 B per cell in binary32 (36 B with binary16 deviations),
             however many steps it performs; achieved = those bytes over the
             average launch duration (HIP events on the library's stream
             around the timed steps), frac = achieved / 8 TB/s <= 1.
             effective_GBps credits 72 B per cell for every step, what
             one-step launches would have to move: above the peak when a
             launch performs two steps.
  accuracy   the state after the timed run is stepped a further few steps on
             the device and by the binary64 oracle (oracle/phys_oracle.c),
             and rho and u are compared over the fluid cells.
  cpu        the binary32 restatement of the reference's WGSL
             (collide_stream + boundary passes, OpenMP over rows) timed on
             this host on a bounded sample of the same lattice without the
             body force, which the reference cannot express.

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


def traffic_from_profile_key(name: str, key: str):
    """HBM bytes per launch recorded as `key: <bytes>` in profiles/<name>.md
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration), or None."""
    p = ROOT / "profiles" / f"{name}.md"
    if not p.exists():
        return None
    m = re.search(rf"^{re.escape(key)}:\s*([0-9.eE+]+)", p.read_text(), re.M)
    return float(m.group(1)) if m else None


def roofline(bytes_per_cell: int, cells: int, step_s: float, per_launch: int, traffic):
    """The step kernel against the HBM roofline.  A launch that performs
    per_launch steps must still read and write every cell's populations
    once (bytes_per_cell x cells); achieved is that over the average launch
    duration, so frac <= 1.  effective_GBps credits bytes_per_cell for every
    step instead."""
    algo = bytes_per_cell * cells
    launch_s = step_s * per_launch
    achieved = algo / launch_s / 1e9
    effective = algo / step_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "steps_per_launch": per_launch,
            "algorithmic_bytes_per_launch": algo, "launch_us": round(launch_s * 1e6, 3),
            "traffic_frac_of_peak": (round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None),
            "effective_GBps": round(effective, 1), "effective_frac": round(effective / HBM_PEAK_GBS, 4)}


def rho_u_errors(pairs):
    """Relative L2 and largest absolute errors of rho and |u - u_ref| over
    the fluid cells of one or more (device, reference, fluid mask) triples,
    each of (rho, ux, uy[, uz]) arrays."""
    sr = sr0 = su = su0 = 0.0
    mr = mu = umax = 0.0
    for dev, ref, fl in pairs:
        dr = (dev[0] - ref[0])[fl]
        du = np.sqrt(sum((d - r)[fl] ** 2 for d, r in zip(dev[1:], ref[1:])))
        uo = np.sqrt(sum(r[fl] ** 2 for r in ref[1:]))
        sr += float(np.sum(dr ** 2)); sr0 += float(np.sum(ref[0][fl] ** 2))
        su += float(np.sum(du ** 2)); su0 += float(np.sum(uo ** 2))
        mr = max(mr, float(np.max(np.abs(dr)))); mu = max(mu, float(np.max(du))); umax = max(umax, float(np.max(uo)))
    return {"rho_l2_rel": float(np.sqrt(sr / max(sr0, 1e-300))), "rho_max_abs": mr,
            "u_l2_rel": float(np.sqrt(su / max(su0, 1e-300))), "u_max_abs": mu, "u_max": umax}


def accuracy_vs_oracle(lat, solid, tau: float, force, steps: int = 20):
    """rho and u of the device after `steps` more steps from its current
    state, against the binary64 oracle run from the same populations: relative
    L2 and largest absolute error over the fluid cells."""
    from oracle import oracle as O

    ny, nx = solid.shape
    t0 = time.perf_counter()
    f = lat.read_populations()
    ps = O.PhysState(O.phys_params(nx, ny, tau, force=force), solid, f)
    del f
    lat.run(steps)
    r_d, ux_d, uy_d = lat.macro()
    ps.run(steps)
    r_o, ux_o, uy_o = ps.macro()
    del ps
    fl = ~solid
    du = np.hypot(ux_d - ux_o, uy_d - uy_o)[fl]
    uo = np.hypot(ux_o, uy_o)[fl]
    dr = (r_d - r_o)[fl]
    out = {"vs": "binary64 oracle (oracle/phys_oracle.c)", "steps": steps, "from": "the state after the timed run",
           "rho_l2_rel": float(np.linalg.norm(dr) / np.linalg.norm(r_o[fl])),
           "rho_max_abs": float(np.max(np.abs(dr))),
           "u_l2_rel": float(np.linalg.norm(du) / max(np.linalg.norm(uo), 1e-300)),
           "u_max_abs": float(np.max(du)), "u_max": float(np.max(uo)),
           "seconds": round(time.perf_counter() - t0, 1)}
    return out


def accuracy_batch(b, doms, finite, steps: int = 20, every: int = 8):
    """Every `every`-th finite domain of a batch, `steps` more steps on the
    device and by the binary64 oracle from the same populations: rho and u
    errors over the sampled domains' fluid cells together."""
    from oracle import oracle as O
    from simuverse_amd import lbm as L

    t0 = time.perf_counter()
    sample = [k for k in finite[::every]]
    if not sample:
        return None
    states = {}
    for k in sample:
        d = doms[k]
        states[k] = O.PhysState(O.phys_params(d.nx, d.ny, d.tau, x_edge=L.X_INLET_OUTLET, inlet_ux=d.inlet_ux),
                                d.solid, b.read_populations(k))
    b.run(steps)
    pairs = []
    for k in sample:
        states[k].run(steps)
        pairs.append((b.macro(k), states[k].macro(), ~doms[k].solid))
        del states[k]
    out = {"vs": "binary64 oracle (oracle/phys_oracle.c)", "steps": steps, "from": "the state after the timed run",
           "domains": sample, "cells": int(sum(doms[k].nx * doms[k].ny for k in sample))}
    out.update(rho_u_errors(pairs))
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def accuracy_coarse(r, plain_steps: int = 20, learned_steps: int = 2):
    """The coarse lattice of the learned workload from its restricted initial
    state: the plain step (the gate's fallback) `plain_steps` steps and the
    learned step with the trained network at the chosen gain (bf16 operands,
    the default) `learned_steps` steps, each against the binary64 oracle
    (oracle/phys_oracle.c, oracle/learned_oracle.c) from the same
    populations."""
    from oracle import oracle as O
    from simuverse_amd import learned as LC
    from simuverse_amd import lbm as L
    from simuverse_amd import workloads as W

    t0 = time.perf_counter()
    nc = r.solid_c.shape[0]
    fl = ~r.solid_c
    params = O.phys_params(nc, nc, LC.TAU_COARSE, force=r.force_c)
    out = {"vs": "binary64 oracles (oracle/phys_oracle.c, oracle/learned_oracle.c)",
           "from": "the restricted initial state"}
    lat = L.Lattice(nc, nc, LC.TAU_COARSE, force=r.force_c)
    lat.write_materials(W.materials_from_solid(r.solid_c))
    lat.write_populations(r.fc0, 0)
    lat.run(plain_steps)
    ps = O.PhysState(params, r.solid_c, r.fc0)
    ps.run(plain_steps)
    out["plain"] = {"steps": plain_steps, **rho_u_errors([(lat.macro(), ps.macro(), fl)])}
    lat.close()
    del ps
    s_in, s_out = r.net_scales
    net = L.Network(seed=1)
    net.set_params(r.net_params)
    net.set_scales(s_in, s_out)
    lat = L.Lattice(nc, nc, LC.TAU_COARSE, force=r.force_c)
    lat.write_materials(W.materials_from_solid(r.solid_c))
    lat.write_populations(r.fc0, 0)
    lat.set_learned_collision(net)
    lat.run(learned_steps)
    post = O.learned_run(params, r.solid_c, r.fc0, r.net_params, s_in, s_out, learned_steps)
    ref = O.PhysState(params, r.solid_c, post).macro()
    out["learned"] = {"steps": learned_steps, "network": "bf16 operands", "gain": r.net_gain,
                      **rho_u_errors([(lat.macro(), ref, fl)])}
    lat.close()
    net.close()
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def accuracy_d3q19(lat, scene, tau: float, force, steps: int = 5):
    """The 3-D lattice after the timed run, `steps` more steps on the device
    and by the binary64 numpy oracle (oracle/d3q19.py) from the same
    populations: rho and u errors over the fluid cells."""
    from oracle import d3q19 as D3

    t0 = time.perf_counter()
    st = D3.D3Q19State(tau, scene.solid, lat.read_populations(), force=force)
    lat.run(steps)
    st.run(steps)
    rho_o, u_o = st.macro()
    del st
    rho_d, u_d = lat.macro()
    out = {"vs": "binary64 oracle (oracle/d3q19.py)", "steps": steps, "from": "the state after the timed run"}
    out.update(rho_u_errors([((rho_d, *u_d), (rho_o, *u_o), ~scene.solid)]))
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


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
    cells_per_thread, vec), the two- or three-step LDS tile (rows x width)
    or the two-step row stream (strips of block columns, chunks of rows)."""
    b, c, v, f = shape
    flags = [name for bit, name in ((1, "nt_store"), (2, "nt_load"), (4, "xcd_rows"), (8, "lane_shuffle"),
                                    (16, "two_step"), (64, "three_step"), (128, "row_stream")) if f & bit]
    if f & 64:
        return {"kernel": "k_phys_step3", "block": b, "tile_rows": c, "tile_width": v, "flags": flags,
                "steps_per_launch": 3}
    if f & 128:
        return {"kernel": "k_phys_step2_rows", "block": b, "chunk_rows": c, "strip_width": v, "flags": flags,
                "steps_per_launch": 2}
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
    variant = b.variant()
    block, vec, tile_h, flags = L.batch_variant_shapes()[variant]
    per_launch = 3 if flags & 64 else (2 if flags & 16 else 1)
    workload = f"d2q9_batch512_{'f16dev' if args.f16 else 'f32'}"
    roof = roofline(bpc, cells, step_s, per_launch, traffic_from_profiles(workload, variant))
    # Domains whose state is no longer finite (BGK at tau near 1/2 with a
    # fast inlet on a small domain): named, and their cells counted apart.
    diverged = [k for k in range(len(doms)) if not np.isfinite(b.reduce(k)).all()]
    finite_cells = sum(d.nx * d.ny for k, d in enumerate(doms) if k not in set(diverged))
    acc = None
    if args.accuracy_steps > 0:
        acc = accuracy_batch(b, doms, [k for k in range(len(doms)) if k not in set(diverged)], args.accuracy_steps)
        log(f"batch accuracy vs binary64 oracle: {acc}")
    log(f"batch: {K} steps {ms:.1f} ms, {cells / step_s / 1e6:.0f} MLUPS; {len(doms) - len(diverged)} of "
        f"{len(doms)} domains finite after {args.warmup + K} steps; diverged: {diverged}")
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
                   "batch_variant": {"kernel": {3: "k_batch_step3", 2: "k_batch_step2"}.get(per_launch, "k_batch_step"),
                                     "block": block, "vec": vec, "tile_rows": tile_h, "flags": flags},
                   "graph_chunk": args.graph_chunk, "seed": W.SEED,
                   "finite_domains": len(doms) - len(diverged), "diverged_domains": diverged,
                   "finite_cells": int(finite_cells)},
        "roofline": roof,
        "accuracy": acc,
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
    workload = f"d3q19_{'f16dev' if args.f16 else 'f32'}_{n}cubed"
    roof = roofline(bpc, cells, step_s, 1, traffic_from_profiles(workload, variant))
    r = lat.reduce()
    log(f"{K} steps: {ms:.1f} ms, {step_s * 1e6:.2f} us/step, {cells / step_s / 1e6:.0f} MLUPS, {roof['achieved']:.0f} GB/s; "
        f"mass {r[0]:.9e} drag ({r[4]:.3e}, {r[5]:.3e}, {r[6]:.3e})")
    if not np.isfinite(r).all():
        raise SystemExit("non-finite state after the timed run")
    acc = None
    if args.accuracy_steps > 0:
        # the numpy oracle takes about 5 s per step at 256^3
        acc = accuracy_d3q19(lat, scene, tau, (g, 0.0, 0.0), min(args.accuracy_steps, 5))
        log(f"d3q19 accuracy vs binary64 oracle: {acc}")
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
        "roofline": roof,
        "accuracy": acc,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def run_learned(args):
    """Learned collision on a 4x coarser lattice against the fine solution
    (simuverse_amd/learned.py): fine n x n binary32 at tau 0.52 on a periodic
    4x4 cylinder array (Re 1920 at n = 8192), samples from its 4x
    restriction, device training, a validation gate against plain BGK, and
    timed coarse runs at tau 0.505.  The line's value is the learned coarse
    step's MLUPS; the config carries the errors and time-to-solution
    ratios."""
    from simuverse_amd import learned as LC

    n = args.n if args.n != 4096 else 8192
    r = LC.run(n=n, fine_steps=args.learned_fine_steps, force=args.learned_force, pair_every=args.pair_every,
               capacity=args.learned_capacity, train_steps=args.learned_train_steps, batch=args.learned_batch,
               log=log)
    nc, K = n // 4, r.coarse_steps
    acc = None
    if args.accuracy_steps > 0:
        acc = accuracy_coarse(r, args.accuracy_steps, 2)
        log(f"coarse accuracy vs binary64 oracles: {acc}")
    line = {
        "metric": METRIC,
        "value": round(nc * nc * K / r.learned_ms / 1e3, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": 0,
        "ms_per_step": round(r.learned_ms / K, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, bf16 network operands on matrix cores (f32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": f"d2q9_learned_{n}_to_{nc}",
            "fine": {"n": n, "tau": LC.TAU_FINE, "steps": r.fine_steps, "ms": round(r.fine_ms, 3)},
            "coarse": {"n": nc, "tau": LC.TAU_COARSE, "steps": K},
            "force_fine": args.learned_force, "U0": LC.U0, "cylinders": "4x4, r = n/64",
            "reynolds": round(r.reynolds, 1),
            "training": {"pairs": r.pairs, "samples": r.samples, "steps": args.learned_train_steps,
                         "batch": min(args.learned_batch, r.samples), "final_loss": r.final_loss,
                         "collect_s": round(r.collect_s, 2), "train_s": round(r.train_s, 2)},
            "validation_gate": {"vel_l2_at_half_window_by_gain": {f"{k:g}": v for k, v in r.gate_errors.items()},
                                "gain": r.gain, "vel_l2": r.gated_err, "coarse_ms": round(r.gated_ms, 3),
                                "speedup_vs_fine": round(r.fine_ms / r.gated_ms, 2)},
            "vel_l2_vs_restricted_fine": {"plain": r.plain_err, "learned": r.learned_err},
            "coarse_ms": {"plain": round(r.plain_ms, 3), "learned": round(r.learned_ms, 3),
                          "learned_binary16_operands": round(r.learned_f16_ms, 3)},
            "vel_l2_learned_binary16_operands": r.learned_f16_err,
            "time_to_solution_speedup": {"learned_vs_fine": round(r.fine_ms / r.learned_ms, 2),
                                         "learned_incl_training_vs_fine":
                                             round(r.fine_ms / (r.learned_ms + 1e3 * (r.collect_s + r.train_s)), 3)},
        },
        "roofline": roofline(72, nc * nc, r.learned_ms / 1e3 / K, 1,
                             traffic_from_profile_key("d2q9_learned_mfma_2048x2048_trace",
                                                      "hbm_bytes_per_launch_bf16") if nc == 2048 else None),
        "accuracy": acc,
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
    ap.add_argument("--accuracy-steps", type=int, default=20,
                    help="steps after the timed run checked against the binary64 oracle (0: skip)")
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
        print("bench.py measures one GPU: the library has no multi-GPU path (the lattice would shard by "
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
    workload = f"d2q9_phys_{'f16dev' if args.f16 else 'f32'}_{n}x{n}"
    traffic = traffic_from_profiles(workload, variant)
    per_launch = describe_variant(L.variant_shapes()[variant])["steps_per_launch"]
    roof = roofline(bytes_per_cell, cells, step_us * 1e-6, per_launch, traffic)
    r = lat.reduce()
    log(f"{K} steps: {ms:.1f} ms events, {wall*1e3:.1f} ms wall, {step_us:.2f} us/step, {mlups:.0f} MLUPS; "
        f"launch {roof['launch_us']:.1f} us, {roof['achieved']:.0f} GB/s of the bytes a launch must move "
        f"(frac {roof['frac']:.3f}), effective {roof['effective_GBps']:.0f} GB/s; "
        f"mass {r[0]:.9e} drag ({r[3]:.3e}, {r[4]:.3e})")
    if not np.isfinite(r).all():
        raise SystemExit("non-finite state after the timed run")

    acc = None
    if args.accuracy_steps > 0:
        acc = accuracy_vs_oracle(lat, scene.solid, tau, (g, 0.0), args.accuracy_steps)
        log(f"accuracy vs binary64 oracle: {acc}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene)
        log(f"cpu baseline: {cpu}")
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
        "roofline": roof,
        "accuracy": acc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
