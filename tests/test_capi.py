"""The C-ABI library loads without a GPU, exports every entry point that
include/lbm.h declares, and rejects bad descriptors before touching a
device."""
import ctypes as C
import re
from pathlib import Path

import pytest

from simuverse_amd import lbm as L

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(lbm_\w+)\s*\(", text, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    lib = L.load_library()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(L.exported_symbols())


def test_abi_version_and_struct_layout():
    lib = L.load_library()
    assert lib.lbm_abi_version() == 1
    assert C.sizeof(L.LbmDesc) == 64
    assert C.sizeof(L.LatticeInfoRec) == 16  # LatticeInfo, fluid/lattice.rs:5-11


@pytest.mark.parametrize(
    "kw, needle",
    [
        (dict(nx=1, ny=8), "both sides"),
        (dict(tau=0.0), "tau"),
        (dict(mode=7), "mode"),
        (dict(mode=L.MODE_PARITY_WGSL, storage=L.STORAGE_F16_DEV), "binary32"),
        (dict(mode=L.MODE_PARITY_WGSL, force=(1e-6, 0.0)), "parity"),
        (dict(x_edge=L.X_INLET_OUTLET, nx=2), "nx >= 3"),
        (dict(dev_shift=20), "dev_shift"),
        (dict(layout=5), "layout"),
    ],
)
def test_invalid_descriptors_are_rejected(kw, needle):
    args = dict(nx=8, ny=8, tau=0.6)
    args.update(kw)
    with pytest.raises(L.LbmError) as ei:
        L.Lattice(**args)
    assert needle in str(ei.value)
    assert ei.value.code in (1, 4)


@pytest.mark.parametrize(
    "kw, needle",
    [
        (dict(nx=2), ">= 3"),
        (dict(nz=0), ">= 3"),
        (dict(tau=0.5), "tau"),
        (dict(storage=9), "storage"),
        (dict(force=(float("nan"), 0.0, 0.0)), "finite"),
        (dict(nx=4096, ny=4096, nz=4096), "32-bit"),
    ],
)
def test_invalid_3d_descriptors_are_rejected(kw, needle):
    args = dict(nx=8, ny=8, nz=8, tau=0.6)
    args.update(kw)
    with pytest.raises(L.LbmError) as ei:
        L.Lattice3D(**args)
    assert needle in str(ei.value)
    assert ei.value.code == 1


def test_3d_struct_layout():
    assert C.sizeof(L.Desc3) == 64


def test_null_handle_is_an_error_not_a_crash():
    lib = L.load_library()
    assert lib.lbm_step(None, 0, 0) == 1
    assert b"null handle" in lib.lbm_last_error()
    assert lib.lbm3d_run(None, 1) == 1
    assert b"null handle" in lib.lbm_last_error()


def test_host_layer_library_loads():
    """The C++ host layer (D2Q9Node / FluidSimulator over the C ABI) is
    built and resolves against the HIP library without a GPU."""
    path = ROOT / "simuverse_amd" / "_lib" / "libsimuverse_host.so"
    assert path.exists(), "libsimuverse_host.so not built"
    C.CDLL(str(path))


def test_tile_fit_rule():
    """The Python mirror of the library's rule for multi-step tile shapes
    (lbm_set_variant refuses a shape the rule rejects): nx, ny >= 3, with
    the inlet/outlet edge a last tile of at least two columns, and rows and
    row stride below 2^24 (the three-step tile's 24-bit row products)."""
    from simuverse_amd import lbm as L

    shapes = L.variant_shapes()
    three = [v for v, s in enumerate(shapes) if s[3] & 64]
    assert three
    for v in three:
        bw = shapes[v][2]
        assert L.variant_ok(v, 4096, 4096)
        assert L.variant_ok(v, 3, 3)
        assert not L.variant_ok(v, 2, 64) and not L.variant_ok(v, 64, 2)
        assert not L.variant_ok(v, bw + 1, 64, L.X_INLET_OUTLET)
        assert L.variant_ok(v, bw + 1, 64, L.X_PERIODIC)
        assert L.variant_ok(v, 3, (1 << 24) - 1) and not L.variant_ok(v, 3, 1 << 24)
        big = (1 << 24) // 9 - 16
        assert L.variant_ok(v, big - 1, 3) and not L.variant_ok(v, big + 1, 3)
