"""ctypes binding of ``include/lbm.h`` (the gfx950 D2Q9 library).

The binding mirrors the reference's LBM node (simuverse/src/fluid/
d2q9_node.rs:13-312): ``Lattice`` owns what ``D2Q9Node`` owns, ``step`` is
``compute_by_pass``, ``reset`` is the init shader, ``write_lattice_info`` is
the ``info_buf`` upload.  There is no CPU fallback: loading raises when the
HIP library has not been built, and every call raises ``LbmError`` with the
library's message when it returns a non-zero status.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libsimuverse_lbm.so"

# status codes
LBM_OK = 0
# modes, storage, x edges, materials (include/lbm.h)
MODE_PARITY_WGSL = 0
MODE_PHYSICS = 1
STORAGE_F32 = 0
STORAGE_F16_DEV = 1
X_PERIODIC = 0
X_INLET_OUTLET = 1
LAYOUT_ROWS = 0
LAYOUT_PLANES = 1
MLP_BF16 = 0
MLP_F32 = 1
MLP_F16 = 2
MAT_BULK, MAT_BOUNDARY, MAT_INLET, MAT_OBSTACLE, MAT_OUTLET, MAT_EXTERNAL_FORCE, MAT_GHOST = 1, 2, 3, 4, 5, 6, 7

LATTICE_INFO_DTYPE = np.dtype([("material", "<i4"), ("block_iter", "<i4"), ("vx", "<f4"), ("vy", "<f4")])


class LbmError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"lbm error {code}: {message}")
        self.code = code


class LbmDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("tau", C.c_float),
        ("mode", C.c_int32),
        ("storage", C.c_int32),
        ("fluid_ty", C.c_int32),
        ("x_edge", C.c_int32),
        ("inlet_ux", C.c_float),
        ("force_x", C.c_float),
        ("force_y", C.c_float),
        ("dev_shift", C.c_int32),
        ("layout", C.c_int32),
        ("reserved", C.c_int32 * 4),
    ]


class LatticeInfoRec(C.Structure):
    _fields_ = [("material", C.c_int32), ("block_iter", C.c_int32), ("vx", C.c_float), ("vy", C.c_float)]


# (name, restype, argtypes) of every entry point include/lbm.h declares.
_P = C.c_void_p
_SIGNATURES = [
    ("lbm_abi_version", C.c_int, []),
    ("lbm_last_error", C.c_char_p, []),
    ("lbm_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("lbm_create", C.c_int, [C.POINTER(LbmDesc), C.POINTER(_P)]),
    ("lbm_destroy", C.c_int, [_P]),
    ("lbm_write_lattice_info", C.c_int, [_P, _P, C.c_size_t, C.c_size_t]),
    ("lbm_read_lattice_info", C.c_int, [_P, _P, C.c_size_t, C.c_size_t]),
    ("lbm_write_materials", C.c_int, [_P, _P, C.c_size_t, C.c_size_t]),
    ("lbm_reset", C.c_int, [_P]),
    ("lbm_write_populations", C.c_int, [_P, C.c_int, _P]),
    ("lbm_read_populations", C.c_int, [_P, C.c_int, _P]),
    ("lbm_state_size", C.c_int, [_P, C.POINTER(C.c_size_t)]),
    ("lbm_save_state", C.c_int, [_P, _P, C.c_size_t]),
    ("lbm_load_state", C.c_int, [_P, _P, C.c_size_t]),
    ("lbm_set_tau", C.c_int, [_P, C.c_float]),
    ("lbm_step", C.c_int, [_P, C.c_int, C.c_int]),
    ("lbm_run", C.c_int, [_P, C.c_int, C.c_int]),
    ("lbm_current_buffer", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("lbm_synchronize", C.c_int, [_P]),
    ("lbm_run_timed", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("lbm_read_macro", C.c_int, [_P, _P, _P, _P]),
    ("lbm_read_macro_rgba16f", C.c_int, [_P, _P]),
    ("lbm_read_curl_rgba16f", C.c_int, [_P, _P]),
    ("lbm_macro", C.c_int, [_P, _P, _P, _P]),
    ("lbm_reduce", C.c_int, [_P, _P]),
    ("lbm_variant_count", C.c_int, []),
    ("lbm_variant_info", C.c_int, [C.c_int] + [C.POINTER(C.c_int)] * 4),
    ("lbm_set_variant", C.c_int, [_P, C.c_int]),
    ("lbm_get_variant", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("lbm_autotune", C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("lbm_set_graph_chunk", C.c_int, [_P, C.c_int]),
    ("lbm_copy_bandwidth", C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    ("lbm_batch_create", C.c_int, [_P, C.c_int, C.c_int, C.POINTER(_P)]),
    ("lbm_batch_destroy", C.c_int, [_P]),
    ("lbm_batch_cells", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("lbm_batch_write_materials", C.c_int, [_P, C.c_int, _P]),
    ("lbm_batch_reset", C.c_int, [_P]),
    ("lbm_batch_write_populations", C.c_int, [_P, C.c_int, _P]),
    ("lbm_batch_read_populations", C.c_int, [_P, C.c_int, _P]),
    ("lbm_batch_run", C.c_int, [_P, C.c_int]),
    ("lbm_batch_run_timed", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("lbm_batch_synchronize", C.c_int, [_P]),
    ("lbm_batch_macro", C.c_int, [_P, C.c_int, _P, _P, _P]),
    ("lbm_batch_reduce", C.c_int, [_P, C.c_int, _P]),
    ("lbm_batch_variant_count", C.c_int, []),
    ("lbm_batch_variant_info", C.c_int, [C.c_int] + [C.POINTER(C.c_int)] * 4),
    ("lbm_batch_set_variant", C.c_int, [_P, C.c_int]),
    ("lbm_batch_get_variant", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("lbm_batch_set_graph_chunk", C.c_int, [_P, C.c_int]),
    ("lbm_batch_autotune", C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("lbm_mlp_param_count", C.c_int, []),
    ("lbm_mlp_create", C.c_int, [C.c_uint64, C.c_int, C.POINTER(_P)]),
    ("lbm_mlp_destroy", C.c_int, [_P]),
    ("lbm_mlp_set_params", C.c_int, [_P, _P]),
    ("lbm_mlp_get_params", C.c_int, [_P, _P]),
    ("lbm_mlp_set_scales", C.c_int, [_P, C.c_float, C.c_float]),
    ("lbm_mlp_get_scales", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("lbm_mlp_forward", C.c_int, [_P, _P, C.c_int, _P]),
    ("lbm_mlp_forward_bf16", C.c_int, [_P, _P, C.c_int, _P]),
    ("lbm_mlp_forward_f16", C.c_int, [_P, _P, C.c_int, _P]),
    ("lbm_mlp_loss_grad", C.c_int, [_P, _P, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    ("lbm_mlp_loss_grad_noalloc", C.c_int, [_P, _P, _P, C.c_int, _P, C.POINTER(C.c_double)]),
    ("lbm_mlp_clear_samples", C.c_int, [_P]),
    ("lbm_mlp_sample_count", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("lbm_mlp_read_samples", C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    ("lbm_mlp_add_samples", C.c_int, [_P, _P, _P, C.c_int]),
    ("lbm_mlp_add_transition", C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                         C.c_int, C.c_int, C.POINTER(C.c_int)]),
    ("lbm_mlp_normalize_samples", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("lbm_mlp_train", C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_double)]),
    ("lbm_set_learned_collision", C.c_int, [_P, _P]),
    ("lbm_set_learned_precision", C.c_int, [_P, C.c_int]),
    ("lbm_health", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_int)]),
    ("lbm_run_guarded", C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_int)]),
    ("lbm_restrict4", C.c_int, [_P, C.c_float, C.c_float, C.c_float, _P, _P, _P, _P]),
    # include/lbm3d.h
    ("lbm3d_create", C.c_int, [_P, C.POINTER(_P)]),
    ("lbm3d_destroy", C.c_int, [_P]),
    ("lbm3d_write_materials", C.c_int, [_P, _P]),
    ("lbm3d_reset", C.c_int, [_P]),
    ("lbm3d_write_populations", C.c_int, [_P, _P]),
    ("lbm3d_read_populations", C.c_int, [_P, _P]),
    ("lbm3d_run", C.c_int, [_P, C.c_int]),
    ("lbm3d_run_timed", C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    ("lbm3d_synchronize", C.c_int, [_P]),
    ("lbm3d_macro", C.c_int, [_P, _P, _P, _P, _P]),
    ("lbm3d_reduce", C.c_int, [_P, _P]),
    ("lbm3d_variant_count", C.c_int, []),
    ("lbm3d_variant_info", C.c_int, [C.c_int] + [C.POINTER(C.c_int)] * 3),
    ("lbm3d_set_variant", C.c_int, [_P, C.c_int]),
    ("lbm3d_get_variant", C.c_int, [_P, C.POINTER(C.c_int)]),
    ("lbm3d_autotune", C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("lbm3d_set_graph_chunk", C.c_int, [_P, C.c_int]),
]


class Desc3(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("tau", C.c_float),
        ("storage", C.c_int32),
        ("force_x", C.c_float),
        ("force_y", C.c_float),
        ("force_z", C.c_float),
        ("reserved", C.c_int32 * 8),
    ]


class DomainDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("tau", C.c_float),
        ("x_edge", C.c_int32),
        ("inlet_ux", C.c_float),
        ("force_x", C.c_float),
        ("force_y", C.c_float),
        ("reserved", C.c_int32 * 5),
    ]

_lib = None


def library_path() -> Path:
    return _LIB_PATH


def load_library() -> C.CDLL:
    """Loads the HIP library once; raises when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise ImportError(
            f"{_LIB_PATH} is missing: build the gfx950 library first (python -m simuverse_amd.build); "
            "there is no CPU fallback"
        )
    lib = C.CDLL(str(_LIB_PATH))
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != LBM_OK:
        raise LbmError(rc, load_library().lbm_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def device_count() -> int:
    n = C.c_int(0)
    _check(load_library().lbm_device_count(C.byref(n)))
    return n.value


def variant_shapes() -> list[tuple[int, int, int, int]]:
    """(threads per block, cells per thread, consecutive cells per lane,
    flags: 1 non-temporal stores, 2 non-temporal loads, 4 rows per XCD,
    8 aligned vector loads with x-neighbours from the adjacent lane, 16 two
    steps per launch through an LDS tile of cells_per_thread rows x vec
    cells, 64 three steps per launch through an LDS tile of cells_per_thread
    rows x vec cells with a two-cell ring, 128 the two steps streamed down
    strips of vec columns in chunks of cells_per_thread rows)."""
    lib = load_library()
    out = []
    for v in range(lib.lbm_variant_count()):
        b, c, w, f = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib.lbm_variant_info(v, C.byref(b), C.byref(c), C.byref(w), C.byref(f)))
        out.append((b.value, c.value, w.value, f.value))
    return out


def variant_ok(variant: int, nx: int, ny: int, x_edge: int = X_PERIODIC) -> bool:
    """Whether a tile variant can step an nx x ny lattice (the library's
    own rule: two- and three-step variants need nx, ny >= 3, with the
    inlet/outlet edge a last tile of at least two columns (nx % vec != 1),
    and rows and row stride below 2^24 (ny < 2^24, 9 (nx + 16) < 2^24);
    lane-shuffle variants need periodic x and whole blocks of block*vec
    cells per row; vector variants need nx % vec == 0)."""
    b, c, v, f = variant_shapes()[variant]
    if f & (16 | 64):
        return (nx >= 3 and ny >= 3 and (x_edge == X_PERIODIC or nx % v != 1) and 9 * (nx + 16) < (1 << 24)
                and ny < (1 << 24))
    if f & 8:
        return x_edge == X_PERIODIC and nx % (b * v) == 0
    return v == 1 or nx % v == 0


def copy_bandwidth(nbytes: int, reps: int = 20) -> float:
    """Milliseconds per device copy of nbytes with the plain streaming kernel."""
    ms = C.c_double()
    _check(load_library().lbm_copy_bandwidth(nbytes, reps, C.byref(ms)))
    return ms.value


class Lattice:
    """One D2Q9 domain on the current HIP device."""

    def __init__(
        self,
        nx: int,
        ny: int,
        tau: float,
        mode: int = MODE_PHYSICS,
        storage: int = STORAGE_F32,
        fluid_ty: int = 1,
        x_edge: int = X_PERIODIC,
        inlet_ux: float = 0.0,
        force: tuple[float, float] = (0.0, 0.0),
        dev_shift: int = 0,
        layout: int = LAYOUT_ROWS,
    ):
        self._lib = load_library()
        self.nx, self.ny = int(nx), int(ny)
        self.n = self.nx * self.ny
        self.mode = mode
        d = LbmDesc()
        d.nx, d.ny, d.tau = self.nx, self.ny, tau
        d.mode, d.storage, d.fluid_ty, d.x_edge = mode, storage, fluid_ty, x_edge
        d.inlet_ux, d.force_x, d.force_y, d.dev_shift = inlet_ux, force[0], force[1], dev_shift
        d.layout = layout
        self.desc = d
        h = C.c_void_p()
        _check(self._lib.lbm_create(C.byref(d), C.byref(h)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.lbm_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs -----------------------------------------------------------
    def write_lattice_info(self, info: np.ndarray, offset: int = 0) -> None:
        info = np.ascontiguousarray(info, dtype=LATTICE_INFO_DTYPE).reshape(-1)
        _check(self._lib.lbm_write_lattice_info(self._h, _ptr(info), offset, info.size))

    def read_lattice_info(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=LATTICE_INFO_DTYPE)
        _check(self._lib.lbm_read_lattice_info(self._h, _ptr(out), 0, self.n))
        return out.reshape(self.ny, self.nx)

    def write_materials(self, material: np.ndarray, offset: int = 0) -> None:
        m = np.ascontiguousarray(material, dtype=np.uint8).reshape(-1)
        _check(self._lib.lbm_write_materials(self._h, _ptr(m), offset, m.size))

    def reset(self) -> None:
        _check(self._lib.lbm_reset(self._h))

    def write_populations(self, f: np.ndarray, buf: int | None = None) -> None:
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(9, self.ny, self.nx)
        _check(self._lib.lbm_write_populations(self._h, self.current_buffer() if buf is None else buf, _ptr(f)))

    def read_populations(self, buf: int | None = None) -> np.ndarray:
        out = np.empty((9, self.ny, self.nx), dtype=np.float32)
        _check(self._lib.lbm_read_populations(self._h, self.current_buffer() if buf is None else buf, _ptr(out)))
        return out

    def save_state(self) -> np.ndarray:
        size = C.c_size_t()
        _check(self._lib.lbm_state_size(self._h, C.byref(size)))
        out = np.empty(size.value, np.uint8)
        _check(self._lib.lbm_save_state(self._h, _ptr(out), out.size))
        return out

    def load_state(self, state: np.ndarray) -> None:
        state = np.ascontiguousarray(state, dtype=np.uint8)
        _check(self._lib.lbm_load_state(self._h, _ptr(state), state.size))

    def set_tau(self, tau: float) -> None:
        _check(self._lib.lbm_set_tau(self._h, tau))

    # -- stepping ---------------------------------------------------------
    def step(self, swap: int, export_macro: bool = True) -> None:
        _check(self._lib.lbm_step(self._h, swap, int(export_macro)))

    def run(self, nsteps: int, export_macro: bool = False) -> None:
        _check(self._lib.lbm_run(self._h, nsteps, int(export_macro)))

    def current_buffer(self) -> int:
        b = C.c_int()
        _check(self._lib.lbm_current_buffer(self._h, C.byref(b)))
        return b.value

    def synchronize(self) -> None:
        _check(self._lib.lbm_synchronize(self._h))

    def run_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(self._lib.lbm_run_timed(self._h, nsteps, C.byref(ms)))
        return ms.value

    def set_variant(self, v: int) -> None:
        _check(self._lib.lbm_set_variant(self._h, v))

    def variant(self) -> int:
        v = C.c_int()
        _check(self._lib.lbm_get_variant(self._h, C.byref(v)))
        return v.value

    def autotune(self, steps: int = 20) -> tuple[int, float]:
        v, ms = C.c_int(), C.c_double()
        _check(self._lib.lbm_autotune(self._h, steps, C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def set_graph_chunk(self, chunk: int) -> None:
        _check(self._lib.lbm_set_graph_chunk(self._h, chunk))

    # -- learned collision ------------------------------------------------
    def set_learned_collision(self, net: "Network | None") -> None:
        _check(self._lib.lbm_set_learned_collision(self._h, net.handle if net is not None else None))
        self._net = net  # keep the network alive while attached

    def set_learned_precision(self, precision: int) -> None:
        """MLP_BF16 (matrix cores, bf16 operands; the default), MLP_F16
        (matrix cores, binary16 operands and GELU) or MLP_F32 (vector ALUs)."""
        _check(self._lib.lbm_set_learned_precision(self._h, precision))

    def health(self) -> tuple[float, bool]:
        u, f = C.c_float(), C.c_int()
        _check(self._lib.lbm_health(self._h, C.byref(u), C.byref(f)))
        return u.value, bool(f.value)

    def run_guarded(self, nsteps: int, chunk: int = 100, umax: float = 0.3) -> int:
        falls = C.c_int()
        _check(self._lib.lbm_run_guarded(self._h, nsteps, chunk, umax, C.byref(falls)))
        return falls.value

    def restrict4(self, tau_coarse: float, force_coarse: tuple[float, float] = (0.0, 0.0)):
        """(coarse pre-collision f (9, ny/4, nx/4), coarse solid, ux, uy)."""
        cy, cx = self.ny // 4, self.nx // 4
        f = np.empty((9, cy, cx), np.float32)
        s = np.empty((cy, cx), np.uint8)
        ux, uy = np.empty((cy, cx), np.float32), np.empty((cy, cx), np.float32)
        _check(self._lib.lbm_restrict4(self._h, tau_coarse, force_coarse[0], force_coarse[1], _ptr(f), _ptr(s),
                                       _ptr(ux), _ptr(uy)))
        return f, s.astype(bool), ux, uy

    # -- outputs ----------------------------------------------------------
    def read_macro(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        rho, ux, uy = (np.empty((self.ny, self.nx), np.float32) for _ in range(3))
        _check(self._lib.lbm_read_macro(self._h, _ptr(rho), _ptr(ux), _ptr(uy)))
        return rho, ux, uy

    def read_macro_rgba16f(self) -> np.ndarray:
        out = np.empty((self.ny, self.nx, 4), np.uint16)
        _check(self._lib.lbm_read_macro_rgba16f(self._h, _ptr(out)))
        return out.view(np.float16)

    def read_curl_rgba16f(self) -> np.ndarray:
        """curl_update.wgsl's texels (curl * 3.5 + 0.5, 0, 0, 0) of the
        exported macro field, as float16 (ny, nx, 4)."""
        out = np.empty((self.ny, self.nx, 4), np.uint16)
        _check(self._lib.lbm_read_curl_rgba16f(self._h, _ptr(out)))
        return out.view(np.float16)

    def macro(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        rho, ux, uy = (np.empty((self.ny, self.nx), np.float32) for _ in range(3))
        _check(self._lib.lbm_macro(self._h, _ptr(rho), _ptr(ux), _ptr(uy)))
        return rho, ux, uy

    def reduce(self) -> np.ndarray:
        out = np.zeros(5, np.float64)
        _check(self._lib.lbm_reduce(self._h, _ptr(out)))
        return out


class Batch:
    """Many independent physics-mode domains stepped by one launch."""

    def __init__(self, domains: list[dict], storage: int = STORAGE_F32):
        self._lib = load_library()
        arr = (DomainDesc * len(domains))()
        for k, d in enumerate(domains):
            arr[k].nx, arr[k].ny, arr[k].tau = d["nx"], d["ny"], d["tau"]
            arr[k].x_edge = d.get("x_edge", X_INLET_OUTLET)
            arr[k].inlet_ux = d.get("inlet_ux", 0.0)
            fx, fy = d.get("force", (0.0, 0.0))
            arr[k].force_x, arr[k].force_y = fx, fy
        self.shapes = [(d["ny"], d["nx"]) for d in domains]
        h = C.c_void_p()
        _check(self._lib.lbm_batch_create(arr, len(domains), storage, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.lbm_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self) -> int:
        c = C.c_uint64()
        _check(self._lib.lbm_batch_cells(self._h, C.byref(c)))
        return c.value

    def write_materials(self, d: int, material: np.ndarray) -> None:
        m = np.ascontiguousarray(material, dtype=np.uint8).reshape(-1)
        assert m.size == self.shapes[d][0] * self.shapes[d][1]
        _check(self._lib.lbm_batch_write_materials(self._h, d, _ptr(m)))

    def reset(self) -> None:
        _check(self._lib.lbm_batch_reset(self._h))

    def write_populations(self, d: int, f: np.ndarray) -> None:
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(9, *self.shapes[d])
        _check(self._lib.lbm_batch_write_populations(self._h, d, _ptr(f)))

    def read_populations(self, d: int) -> np.ndarray:
        out = np.empty((9, *self.shapes[d]), np.float32)
        _check(self._lib.lbm_batch_read_populations(self._h, d, _ptr(out)))
        return out

    def run(self, nsteps: int) -> None:
        _check(self._lib.lbm_batch_run(self._h, nsteps))

    def run_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(self._lib.lbm_batch_run_timed(self._h, nsteps, C.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        _check(self._lib.lbm_batch_synchronize(self._h))

    def macro(self, d: int):
        rho, ux, uy = (np.empty(self.shapes[d], np.float32) for _ in range(3))
        _check(self._lib.lbm_batch_macro(self._h, d, _ptr(rho), _ptr(ux), _ptr(uy)))
        return rho, ux, uy

    def reduce(self, d: int) -> np.ndarray:
        out = np.zeros(5)
        _check(self._lib.lbm_batch_reduce(self._h, d, _ptr(out)))
        return out

    def set_variant(self, v: int) -> None:
        _check(self._lib.lbm_batch_set_variant(self._h, v))

    def variant(self) -> int:
        v = C.c_int()
        _check(self._lib.lbm_batch_get_variant(self._h, C.byref(v)))
        return v.value

    def set_graph_chunk(self, chunk: int) -> None:
        _check(self._lib.lbm_batch_set_graph_chunk(self._h, chunk))

    def autotune(self, steps: int = 10) -> tuple[int, float]:
        v, ms = C.c_int(), C.c_double()
        _check(self._lib.lbm_batch_autotune(self._h, steps, C.byref(v), C.byref(ms)))
        return v.value, ms.value


class Network:
    """The learned collision correction's network (9 -> 64 -> 64 -> 9)."""

    def __init__(self, seed: int = 1, capacity: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        _check(self._lib.lbm_mlp_create(seed, capacity, C.byref(h)))
        self._h = h
        self.nparams = self._lib.lbm_mlp_param_count()

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.lbm_mlp_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        _check(self._lib.lbm_mlp_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        _check(self._lib.lbm_mlp_set_params(self._h, _ptr(p)))

    def set_scales(self, s_in: float, s_out: float) -> None:
        _check(self._lib.lbm_mlp_set_scales(self._h, s_in, s_out))

    def scales(self) -> tuple[float, float]:
        a, b = C.c_float(), C.c_float()
        _check(self._lib.lbm_mlp_get_scales(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def forward(self, x: np.ndarray, bf16: bool = False, f16: bool = False) -> np.ndarray:
        """Outputs of n samples: binary32 on the vector ALUs, or on the
        matrix cores with bf16 (bf16=True) or binary16 (f16=True) operands."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        y = np.empty_like(x)
        fn = (self._lib.lbm_mlp_forward_bf16 if bf16 else
              self._lib.lbm_mlp_forward_f16 if f16 else self._lib.lbm_mlp_forward)
        _check(fn(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 9)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _check(self._lib.lbm_mlp_loss_grad(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g

    def loss_grad_noalloc(self, x: np.ndarray, t: np.ndarray) -> tuple[float, np.ndarray]:
        """loss_grad through lbm_mlp_loss_grad_noalloc (no device allocation;
        at most 16384 samples)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 9)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _check(self._lib.lbm_mlp_loss_grad_noalloc(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g),
                                                   C.byref(loss)))
        return loss.value, g

    def clear_samples(self) -> None:
        _check(self._lib.lbm_mlp_clear_samples(self._h))

    def sample_count(self) -> int:
        n = C.c_int()
        _check(self._lib.lbm_mlp_sample_count(self._h, C.byref(n)))
        return n.value

    def read_samples(self, offset: int = 0, count: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        if count is None:
            count = self.sample_count() - offset
        x = np.empty((count, 9), np.float32)
        t = np.empty((count, 9), np.float32)
        _check(self._lib.lbm_mlp_read_samples(self._h, offset, count, _ptr(x), _ptr(t)))
        return x, t

    def add_samples(self, x: np.ndarray, t: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, 9)
        _check(self._lib.lbm_mlp_add_samples(self._h, _ptr(x), _ptr(t), x.shape[0]))

    def add_transition(self, s0: np.ndarray, s1: np.ndarray, solid: np.ndarray, tau: float,
                       force: tuple[float, float] = (0.0, 0.0), stride: int = 1, offset: int = 0) -> int:
        ny, nx = solid.shape
        s0 = np.ascontiguousarray(s0, dtype=np.float32).reshape(9, ny, nx)
        s1 = np.ascontiguousarray(s1, dtype=np.float32).reshape(9, ny, nx)
        sd = np.ascontiguousarray(solid, dtype=np.uint8)
        added = C.c_int()
        _check(self._lib.lbm_mlp_add_transition(self._h, _ptr(s0), _ptr(s1), _ptr(sd), nx, ny, tau, force[0],
                                                 force[1], stride, offset, C.byref(added)))
        return added.value

    def normalize_samples(self) -> tuple[float, float]:
        """Unit-rms stored samples; the factors fold into the scales."""
        rx, rt = C.c_float(), C.c_float()
        _check(self._lib.lbm_mlp_normalize_samples(self._h, C.byref(rx), C.byref(rt)))
        return rx.value, rt.value

    def train(self, steps: int, batch: int, lr: float, seed: int = 7) -> float:
        loss = C.c_double()
        _check(self._lib.lbm_mlp_train(self._h, steps, batch, lr, seed, C.byref(loss)))
        return loss.value


class Lattice3D:
    """One D3Q19 lattice (periodic, bounce-back solids, Guo force) on the
    current HIP device (include/lbm3d.h).  Arrays are (19, nz, ny, nx)
    populations and (nz, ny, nx) fields."""

    def __init__(self, nx: int, ny: int, nz: int, tau: float, storage: int = STORAGE_F32,
                 force: tuple[float, float, float] = (0.0, 0.0, 0.0)):
        self._lib = load_library()
        self.shape = (int(nz), int(ny), int(nx))
        self.n = int(nx) * int(ny) * int(nz)
        d = Desc3()
        d.nx, d.ny, d.nz, d.tau, d.storage = int(nx), int(ny), int(nz), tau, storage
        d.force_x, d.force_y, d.force_z = force
        h = C.c_void_p()
        _check(self._lib.lbm3d_create(C.byref(d), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.lbm3d_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_materials(self, material: np.ndarray) -> None:
        m = np.ascontiguousarray(material, dtype=np.uint8).reshape(-1)
        if m.size != self.n:
            raise ValueError(f"material has {m.size} cells, lattice {self.n}")
        _check(self._lib.lbm3d_write_materials(self._h, _ptr(m)))

    def reset(self) -> None:
        _check(self._lib.lbm3d_reset(self._h))

    def write_populations(self, f: np.ndarray) -> None:
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(19, *self.shape)
        _check(self._lib.lbm3d_write_populations(self._h, _ptr(f)))

    def read_populations(self) -> np.ndarray:
        out = np.empty((19,) + self.shape, np.float32)
        _check(self._lib.lbm3d_read_populations(self._h, _ptr(out)))
        return out

    def run(self, nsteps: int) -> None:
        _check(self._lib.lbm3d_run(self._h, nsteps))

    def run_timed(self, nsteps: int) -> float:
        ms = C.c_double()
        _check(self._lib.lbm3d_run_timed(self._h, nsteps, C.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        _check(self._lib.lbm3d_synchronize(self._h))

    def macro(self) -> tuple[np.ndarray, np.ndarray]:
        """rho (nz, ny, nx) and u (3, nz, ny, nx)."""
        rho = np.empty(self.shape, np.float32)
        u = np.empty((3,) + self.shape, np.float32)
        _check(self._lib.lbm3d_macro(self._h, _ptr(rho), _ptr(u[0]), _ptr(u[1]), _ptr(u[2])))
        return rho, u

    def reduce(self) -> np.ndarray:
        out = np.zeros(7, np.float64)
        _check(self._lib.lbm3d_reduce(self._h, _ptr(out)))
        return out

    def set_variant(self, v: int) -> None:
        _check(self._lib.lbm3d_set_variant(self._h, v))

    def variant(self) -> int:
        v = C.c_int()
        _check(self._lib.lbm3d_get_variant(self._h, C.byref(v)))
        return v.value

    def set_graph_chunk(self, chunk: int) -> None:
        _check(self._lib.lbm3d_set_graph_chunk(self._h, chunk))

    def autotune(self, steps: int = 20) -> tuple[int, float]:
        v, ms = C.c_int(), C.c_double()
        _check(self._lib.lbm3d_autotune(self._h, steps, C.byref(v), C.byref(ms)))
        return v.value, ms.value


def variant3_shapes() -> list[tuple[int, int, int]]:
    """(threads per block, cells per thread, flags: 1 non-temporal stores)
    of the D3Q19 step kernel's shapes."""
    lib = load_library()
    out = []
    for v in range(lib.lbm3d_variant_count()):
        b, c, f = C.c_int(), C.c_int(), C.c_int()
        _check(lib.lbm3d_variant_info(v, C.byref(b), C.byref(c), C.byref(f)))
        out.append((b.value, c.value, f.value))
    return out


def batch_variant_count() -> int:
    return load_library().lbm_batch_variant_count()


def batch_variant_shapes() -> list[tuple[int, int, int, int]]:
    """(threads per block, consecutive cells per lane or multi-step tile
    width, multi-step tile rows (0 for one-step variants), flags: 1
    non-temporal stores, 4 work grouped per XCD, 16 two steps per launch,
    32 persistent blocks take every grid-th chunk, 64 three steps per
    launch)."""
    lib = load_library()
    out = []
    for v in range(lib.lbm_batch_variant_count()):
        b, w, h, f = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(lib.lbm_batch_variant_info(v, C.byref(b), C.byref(w), C.byref(h), C.byref(f)))
        out.append((b.value, w.value, h.value, f.value))
    return out


def exported_symbols() -> list[str]:
    return [name for name, _, _ in _SIGNATURES]

