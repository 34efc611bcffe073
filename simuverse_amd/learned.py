"""The learned-collision workflow on a periodic cylinder array: a fine plain
run, training samples from its 4x restriction, device training of the
network, a validation gate against plain BGK, and timed coarse runs.

No reference counterpart: jinleili/simuverse has no learned collision.

Flow (n x n fine, (n/4)^2 coarse):

  fine      tau_f = 0.52 (nu = 1/150), 4x4 cylinders of radius n/64, from a
            uniform flow U = 0.05 held by a Guo body force g_f; N_f steps
            (timed), restricted at 0 and at N_f
  samples   a second fine pass over [N_f/8, N_f/2): every `pair_every` steps
            the restricted state and the one 4 fine steps later give one
            transition (lbm_mlp_add_transition), then normalised
  training  Adam on the device (binary32 forward and backward)
  gate      coarse runs to N_f/2 (N_f/8 coarse steps) with the correction's
            output scale times 1, 1/2, 1/4 and 0 (plain BGK); the gain with
            the smallest velocity error against the restricted fine state
            at N_f/2 is kept
  coarse    N_f/4 coarse steps from the restricted initial state, plain BGK
            and BGK + the network at the chosen gain (timed), judged by the
            velocity L2 error against the restricted fine state at N_f; the
            network runs on the matrix cores with bf16 operands (the
            library's default), and once more with binary16 operands for
            comparison
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import lbm as L
from . import workloads as W

TAU_FINE = 0.52
TAU_COARSE = 0.505
U0 = 0.05
GAINS = (1.0, 0.5, 0.25, 0.0)


@dataclass
class LearnedResult:
    n: int
    fine_steps: int
    reynolds: float
    fine_ms: float
    pairs: int = 0
    samples: int = 0
    final_loss: float = float("nan")
    collect_s: float = 0.0
    train_s: float = 0.0
    gate_errors: dict = field(default_factory=dict)
    gain: float = 0.0
    plain_ms: float = 0.0
    plain_err: float = float("nan")
    learned_ms: float = 0.0
    learned_err: float = float("nan")
    learned_f16_ms: float = 0.0
    learned_f16_err: float = float("nan")
    coarse_steps: int = 0
    coarse_health: dict = field(default_factory=dict)
    # the coarse problem and the trained network at the timed learned run's
    # gain, for checks after the workflow
    fc0: np.ndarray | None = None
    solid_c: np.ndarray | None = None
    force_c: tuple = (0.0, 0.0)
    net_params: np.ndarray | None = None
    net_scales: tuple = (1.0, 1.0)
    net_gain: float = 1.0

    @property
    def gated_err(self) -> float:
        return self.learned_err if self.gain > 0.0 else self.plain_err

    @property
    def gated_ms(self) -> float:
        return self.learned_ms if self.gain > 0.0 else self.plain_ms


def velocity_error(ux, uy, ur, vr, fluid) -> float:
    """Relative L2 error of the velocity over the fluid cells."""
    ref = np.hypot(ur, vr)[fluid]
    return float(np.linalg.norm(np.hypot(ux - ur, uy - vr)[fluid]) / np.linalg.norm(ref))


def fine_lattice(n: int, force: float) -> L.Lattice:
    solid = W.cylinder_array_solid(n, 4, n // 64)
    lat = L.Lattice(n, n, TAU_FINE, force=(force, 0.0))
    lat.write_materials(W.materials_from_solid(solid))
    lat.write_populations(W.equilibrium_state(n, n, U0, 0.0, solid).astype(np.float32), 0)
    return lat


def run(n: int = 8192, fine_steps: int = 4000, force: float = 1e-7, pair_every: int = 100,
        capacity: int = 4_000_000, train_steps: int = 300, batch: int = 1_000_000, lr: float = 2e-3,
        log=lambda *a: None) -> LearnedResult:
    """The whole workflow; returns its figures (the network is trained and
    judged inside)."""
    nc = n // 4
    gc = (4 * force, 0.0)
    Nf = fine_steps
    res = LearnedResult(n=n, fine_steps=Nf, reynolds=U0 * (n // 32) / ((TAU_FINE - 0.5) / 3), fine_ms=0.0)
    fine = fine_lattice(n, force)
    fine.autotune(steps=10)
    fc0, solid_c, _, _ = fine.restrict4(TAU_COARSE, gc)
    res.fine_ms = fine.run_timed(Nf)
    _, _, uxr, uyr = fine.restrict4(TAU_COARSE, gc)
    variant = fine.variant()
    fine.close()
    log(f"fine {n}^2: {Nf} steps {res.fine_ms:.1f} ms ({n * n * Nf / res.fine_ms / 1e3:.0f} MLUPS), "
        f"Re {res.reynolds:.0f}")

    # training samples from a second pass over the second quarter of the
    # window (past the start-up transient) to its middle
    t0 = time.perf_counter()
    net = L.Network(seed=7, capacity=capacity)
    net.set_scales(1.0, 1.0)  # normalize_samples sets them from the samples
    fine2 = fine_lattice(n, force)
    fine2.set_variant(variant)
    stride = max(1, (nc * nc * (Nf // 2 // pair_every)) // capacity + 1)
    done = 0
    for t in range(Nf // 8, Nf // 2, pair_every):
        fine2.run(t - done)
        done = t
        a, _, _, _ = fine2.restrict4(TAU_COARSE, gc)
        fine2.run(4)
        done += 4
        b, _, _, _ = fine2.restrict4(TAU_COARSE, gc)
        net.add_transition(a, b, solid_c, TAU_COARSE, gc, stride=stride, offset=res.pairs % stride)
        res.pairs += 1
    if done <= Nf // 2:
        fine2.run(Nf // 2 - done)
    _, _, uxh, uyh = fine2.restrict4(TAU_COARSE, gc)
    fine2.close()
    net.normalize_samples()
    res.samples = net.sample_count()
    res.collect_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    res.final_loss = net.train(steps=train_steps, batch=min(batch, res.samples), lr=lr, seed=3)
    res.train_s = time.perf_counter() - t0
    log(f"samples {res.samples} from {res.pairs} transitions ({res.collect_s:.1f} s), "
        f"train {train_steps} steps ({res.train_s:.1f} s), loss {res.final_loss:.3e}")

    fluid_c = ~solid_c

    def coarse(learned: bool, steps: int, ref_u, precision: int = L.MLP_BF16):
        lat = L.Lattice(nc, nc, TAU_COARSE, force=gc)
        lat.write_materials(W.materials_from_solid(solid_c))
        lat.write_populations(fc0, 0)
        if learned:
            lat.set_learned_collision(net)
            lat.set_learned_precision(precision)
        else:
            lat.autotune(steps=10)
            lat.write_populations(fc0, 0)
        ms = lat.run_timed(steps)
        _, ux, uy = lat.macro()
        u, finite = lat.health()
        lat.close()
        err = velocity_error(ux, uy, ref_u[0], ref_u[1], fluid_c) if finite else float("inf")
        return ms, err, (float(u), bool(finite))

    s_in, s_out = net.scales()
    for gain in GAINS:
        if gain > 0.0:
            net.set_scales(s_in, s_out * gain)
        res.gate_errors[gain] = coarse(gain > 0.0, Nf // 8, (uxh, uyh))[1]
    res.gain = min(res.gate_errors, key=lambda k: (res.gate_errors[k], -k))
    log("validation gate (vel L2 at N_f/2 by gain): "
        + ", ".join(f"{k:g}: {v:.4e}" for k, v in res.gate_errors.items()) + f"; chosen gain {res.gain:g}")
    res.coarse_steps = Nf // 4
    res.plain_ms, res.plain_err, h_plain = coarse(False, res.coarse_steps, (uxr, uyr))
    # the learned kernel at the chosen gain; at gain 0 the gate falls back
    # to plain BGK and the learned kernel is timed at gain 1 for reference
    net.set_scales(s_in, s_out * (res.gain if res.gain > 0.0 else 1.0))
    res.learned_ms, res.learned_err, h_learn = coarse(True, res.coarse_steps, (uxr, uyr))
    res.learned_f16_ms, res.learned_f16_err, _ = coarse(True, res.coarse_steps, (uxr, uyr), L.MLP_F16)
    res.coarse_health = {"plain": h_plain, "learned": h_learn}
    res.fc0, res.solid_c, res.force_c = fc0, solid_c, gc
    res.net_params, res.net_scales = net.params(), net.scales()
    res.net_gain = res.gain if res.gain > 0.0 else 1.0
    net.close()
    log(f"coarse plain: {res.plain_ms:.1f} ms, vel L2 {res.plain_err:.4e}; learned (gain "
        f"{res.gain if res.gain > 0 else 1:g}, bf16 operands): {res.learned_ms:.1f} ms, vel L2 "
        f"{res.learned_err:.4e}; binary16 operands: {res.learned_f16_ms:.1f} ms, vel L2 {res.learned_f16_err:.4e}")
    return res
