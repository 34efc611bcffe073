"""The C++ host layer (simuverse_amd/host: D2Q9Node, FluidSimulator with the
reference's names) driven frame by frame through clicks, drags, a viscosity
change and a reset, bit for bit against the CPU restatement of the
reference shaders.  The checks live in tests/cpp/test_fluid_simulator.cpp;
this runs the program built by `make -C tests/cpp`."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "_build" / "test_fluid_simulator"


@pytest.mark.gpu
def test_cpp_fluid_simulator_matches_restatement():
    assert BIN.exists(), "tests/cpp/_build/test_fluid_simulator not built (make -C tests/cpp)"
    print("case: C++ FluidSimulator 256x128 vs restatement", flush=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED: 0 failure(s)" in r.stdout
