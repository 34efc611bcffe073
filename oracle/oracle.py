"""numpy wrappers over the CPU oracle library (oracle/lib/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY.  The tests, the smoke check and the CPU-baseline
leg of the benchmark use this module as the checker; the product code in
simuverse_amd/ never imports it.

  WGSL restatement (binary32)   wgsl_restate.c, follows the reference shaders
  physics oracle (binary64)     phys_oracle.c
  MLP oracle (binary64)         mlp_oracle.c
  learned step (binary64)       learned_oracle.c: the learned collision step,
                                its training targets, the 4x restriction
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "liblbm_oracle.so"

INFO_DTYPE = np.dtype([("material", "<i4"), ("block_iter", "<i4"), ("vx", "<f4"), ("vy", "<f4")])

_lib = None


class PoParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("tau", C.c_double),
        ("x_edge", C.c_int32),
        ("inlet_ux", C.c_double),
        ("gx", C.c_double),
        ("gy", C.c_double),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.c_void_p
        sig = {
            "ws_init": (None, [C.c_int, C.c_int, C.c_int, P, P, P, P]),
            "ws_collide_stream": (None, [C.c_int, C.c_int, C.c_float, P, P, P, P]),
            "ws_boundary": (None, [C.c_int, C.c_int, P, P]),
            "ws_compute_by_pass": (None, [C.c_int, C.c_int, C.c_float, P, P, P, P, C.c_int]),
            "ws_run": (None, [C.c_int, C.c_int, C.c_float, P, P, P, P, C.c_int, C.c_int]),
            "ws_weight": (C.c_float, [C.c_int]),
            "ws_max_value": (C.c_float, [C.c_int]),
            "ws_omega": (C.c_float, [C.c_float]),
            "po_weight": (C.c_double, [C.c_int]),
            "po_gather": (None, [C.POINTER(PoParams), P, P, C.c_int, C.c_int, P]),
            "po_collide": (None, [C.POINTER(PoParams), P, P]),
            "po_step": (None, [C.POINTER(PoParams), P, P, P]),
            "po_run": (None, [C.POINTER(PoParams), P, P, P, C.c_int]),
            "po_macro": (None, [C.POINTER(PoParams), P, P, P, P, P]),
            "po_reduce": (None, [C.POINTER(PoParams), P, P, P]),
            "po_restrict4": (None, [C.c_int, C.c_int, P, P]),
            "mlp_param_count": (C.c_int, []),
            "mlp_gelu": (C.c_double, [C.c_double]),
            "mlp_gelu_grad": (C.c_double, [C.c_double]),
            "mlp_forward": (None, [P, C.c_int, P, P]),
            "mlp_loss_grad": (C.c_double, [P, C.c_int, P, P, P]),
            "mlp_adam": (None, [P, P, P, P, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]),
            "lo_modulation_bound": (C.c_double, [C.c_double]),
            "lo_step": (C.c_int, [C.POINTER(PoParams), P, P, P, P, C.c_double, C.c_double, P]),
            "lo_run": (C.c_int, [C.POINTER(PoParams), P, P, P, P, C.c_double, C.c_double, C.c_int]),
            "lo_transition": (C.c_int, [C.POINTER(PoParams), P, P, P, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_int, P, P]),
            "lo_restrict4": (None, [C.POINTER(PoParams), P, P, C.c_double, C.c_double, C.c_double, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------------------------
# WGSL restatement (parity mode)
# ---------------------------------------------------------------------------


class WgslState:
    """The reference node's buffers: info_buf, two lattice buffers, macro_tex."""

    def __init__(self, info: np.ndarray, tau: float):
        self.info = np.ascontiguousarray(info, dtype=INFO_DTYPE).copy()
        self.ny, self.nx = self.info.shape
        self.tau = np.float32(tau)
        self.buf = [np.zeros((9, self.ny, self.nx), np.float32) for _ in range(2)]
        self.macro = np.zeros((self.ny, self.nx, 4), np.float32)

    def init(self, fluid_ty: int = 1) -> None:
        lib().ws_init(self.nx, self.ny, fluid_ty, _p(self.info), _p(self.buf[0]), _p(self.buf[1]), _p(self.macro))

    def step(self, swap: int) -> None:
        omega = lib().ws_omega(self.tau)
        lib().ws_compute_by_pass(self.nx, self.ny, omega, _p(self.info), _p(self.buf[0]), _p(self.buf[1]),
                                 _p(self.macro), swap)

    def run(self, nsteps: int, start_swap: int = 0) -> int:
        """nsteps alternating steps; returns the buffer holding the result."""
        lib().ws_run(self.nx, self.ny, self.tau, _p(self.info), _p(self.buf[0]), _p(self.buf[1]), _p(self.macro),
                     nsteps, start_swap)
        return (start_swap + nsteps) & 1


def wgsl_curl_rgba16f(macro4: np.ndarray) -> np.ndarray:
    """assets/wgsl/lbm/curl_update.wgsl:12-34 in numpy: macro4 (ny, nx, 4)
    binary32 (ux, uy, rho, w) as the collide pass wrote it; the texture holds
    it as binary16.  curl = ((uy[x+1] - uy[max(x-1,0)]) + ux[max(y-1,0)])
    - ux[y+1] in binary32 (left to right, as the shader writes it), with the
    texel past the last column / row taken as zero; output rgba16float
    (curl * 3.5 + 0.5, 0, 0, 0), float16 (ny, nx, 4)."""
    ux = macro4[..., 0].astype(np.float16).astype(np.float32)
    uy = macro4[..., 1].astype(np.float16).astype(np.float32)
    ny, nx = ux.shape
    uxp = np.zeros((ny + 1, nx), np.float32)
    uxp[:ny] = ux
    uyp = np.zeros((ny, nx + 1), np.float32)
    uyp[:, :nx] = uy
    xs, ys = np.arange(nx), np.arange(ny)
    right = uyp[:, xs + 1]
    left = uyp[:, np.maximum(xs - 1, 0)]
    top = uxp[np.maximum(ys - 1, 0), :]
    bottom = uxp[ys + 1, :]
    curl = ((right - left) + top) - bottom
    out = np.zeros((ny, nx, 4), np.float16)
    out[..., 0] = (curl * np.float32(3.5) + np.float32(0.5)).astype(np.float16)
    return out


def wgsl_weights() -> np.ndarray:
    return np.array([lib().ws_weight(i) for i in range(9)], np.float32)


def wgsl_max_values() -> np.ndarray:
    return np.array([lib().ws_max_value(i) for i in range(9)], np.float32)


# ---------------------------------------------------------------------------
# Physics oracle (binary64)
# ---------------------------------------------------------------------------


def phys_params(nx: int, ny: int, tau: float, x_edge: int = 0, inlet_ux: float = 0.0,
                force: tuple[float, float] = (0.0, 0.0)) -> PoParams:
    return PoParams(nx, ny, tau, x_edge, inlet_ux, force[0], force[1])


class PhysState:
    """Post-collision populations of the fluid cells, binary64."""

    def __init__(self, params: PoParams, solid: np.ndarray, f: np.ndarray):
        self.p = params
        self.solid = np.ascontiguousarray(solid, dtype=np.uint8)
        self.f = np.ascontiguousarray(f, dtype=np.float64).reshape(9, params.ny, params.nx).copy()
        self.f[:, self.solid.astype(bool)] = 0.0
        self._tmp = np.empty_like(self.f)

    def run(self, nsteps: int) -> None:
        lib().po_run(C.byref(self.p), _p(self.solid), _p(self.f), _p(self._tmp), nsteps)

    def macro(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        shape = (self.p.ny, self.p.nx)
        rho, ux, uy = (np.empty(shape) for _ in range(3))
        lib().po_macro(C.byref(self.p), _p(self.solid), _p(self.f), _p(rho), _p(ux), _p(uy))
        return rho, ux, uy

    def reduce(self) -> np.ndarray:
        out = np.zeros(5)
        lib().po_reduce(C.byref(self.p), _p(self.solid), _p(self.f), _p(out))
        return out

    def gather(self, x: int, y: int) -> np.ndarray:
        out = np.zeros(9)
        lib().po_gather(C.byref(self.p), _p(self.solid), _p(self.f), x, y, _p(out))
        return out


def phys_collide(params: PoParams, f: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(9)
    lib().po_collide(C.byref(params), _p(f), _p(out))
    return out


def restrict4(field: np.ndarray) -> np.ndarray:
    field = np.ascontiguousarray(field, dtype=np.float64)
    ny, nx = field.shape
    out = np.zeros((ny // 4, nx // 4))
    lib().po_restrict4(nx, ny, _p(field), _p(out))
    return out


# ---------------------------------------------------------------------------
# MLP oracle (binary64)
# ---------------------------------------------------------------------------


def mlp_param_count() -> int:
    return lib().mlp_param_count()


def mlp_forward(params: np.ndarray, x: np.ndarray) -> np.ndarray:
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 9)
    y = np.zeros_like(x)
    lib().mlp_forward(_p(params), x.shape[0], _p(x), _p(y))
    return y


def mlp_loss_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 9)
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1, 9)
    g = np.zeros_like(params)
    loss = lib().mlp_loss_grad(_p(params), x.shape[0], _p(x), _p(t), _p(g))
    return loss, g


def mlp_adam(params: np.ndarray, grad: np.ndarray, m: np.ndarray, v: np.ndarray, step: int, lr: float,
             beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
    for a in (params, grad, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    lib().mlp_adam(_p(params), _p(grad), _p(m), _p(v), step, lr, beta1, beta2, eps)


# ---------------------------------------------------------------------------
# Learned collision step, its training targets, the 4x restriction (binary64)
# ---------------------------------------------------------------------------


def modulation_bound(omega: float) -> float:
    return lib().lo_modulation_bound(omega)


def _net64(net_params: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(net_params, dtype=np.float64)
    assert p.size == mlp_param_count()
    return p


def learned_step(params: PoParams, solid: np.ndarray, post: np.ndarray, net_params: np.ndarray, s_in: float,
                 s_out: float) -> tuple[np.ndarray, np.ndarray]:
    """One learned step of post-collision populations (9, ny, nx): the next
    post-collision state and the map of corrected cells."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    a = np.ascontiguousarray(post, dtype=np.float64).reshape(9, params.ny, params.nx)
    out = np.empty_like(a)
    learned = np.zeros((params.ny, params.nx), np.uint8)
    rc = lib().lo_step(C.byref(params), _p(solid), _p(a), _p(out), _p(_net64(net_params)), s_in, s_out,
                       _p(learned))
    assert rc == 0, "oracle out of memory"
    return out, learned.astype(bool)


def learned_run(params: PoParams, solid: np.ndarray, post: np.ndarray, net_params: np.ndarray, s_in: float,
                s_out: float, nsteps: int) -> np.ndarray:
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    a = np.ascontiguousarray(post, dtype=np.float64).reshape(9, params.ny, params.nx).copy()
    a[:, solid.astype(bool)] = 0.0
    b = np.empty_like(a)
    rc = lib().lo_run(C.byref(params), _p(solid), _p(a), _p(b), _p(_net64(net_params)), s_in, s_out, nsteps)
    assert rc == 0, "oracle out of memory"
    return a


def transition_samples(s0: np.ndarray, s1: np.ndarray, solid: np.ndarray, tau: float,
                       force: tuple[float, float] = (0.0, 0.0), s_in: float = 1.0, s_out: float = 1.0,
                       stride: int = 1, offset: int = 0, max_samples: int | None = None):
    """Training samples (X, T), each (count, 9), of the transition s0 -> s1
    (pre-collision states, dense (9, ny, nx)) of a periodic lattice."""
    s0 = np.ascontiguousarray(s0, dtype=np.float64)
    s1 = np.ascontiguousarray(s1, dtype=np.float64)
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    ny, nx = solid.shape
    if max_samples is None:
        max_samples = (nx * ny - offset + stride - 1) // stride
    X = np.zeros((max_samples, 9))
    T = np.zeros((max_samples, 9))
    p = phys_params(nx, ny, tau, force=force)
    k = lib().lo_transition(C.byref(p), _p(solid), _p(s0), _p(s1), s_in, s_out, stride, offset, max_samples,
                            _p(X), _p(T))
    return X[:k], T[:k]


def restrict4_state(params: PoParams, solid: np.ndarray, post: np.ndarray, tau_c: float,
                    force_c: tuple[float, float] = (0.0, 0.0)):
    """4x restriction of post-collision populations (9, ny, nx): coarse
    pre-collision populations (9, ny/4, nx/4), the coarse solid map and the
    coarse velocity."""
    solid = np.ascontiguousarray(solid, dtype=np.uint8)
    post = np.ascontiguousarray(post, dtype=np.float64)
    ny, nx = solid.shape
    cy, cx = ny // 4, nx // 4
    fc = np.zeros((9, cy, cx))
    sc = np.zeros((cy, cx), np.uint8)
    ux = np.zeros((cy, cx))
    uy = np.zeros((cy, cx))
    lib().lo_restrict4(C.byref(params), _p(solid), _p(post), tau_c, force_c[0], force_c[1], _p(fc), _p(sc), _p(ux),
                       _p(uy))
    return fc, sc.astype(bool), ux, uy
