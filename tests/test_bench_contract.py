"""The benchmark's command-line contract and its profile lookups, on the CPU:
more than one GPU is refused with a message (the path runs on one GPU),
and every committed byte-counter summary names the variant its file name
says and the per-launch traffic bench.py reads from it."""
import importlib.util
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "1 GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_profile_summaries_match_their_names():
    bench = load_bench()
    files = sorted((ROOT / "profiles").glob("*_v*_pmc.md"))
    assert files, "no byte-counter summaries committed"
    for p in files:
        m = re.fullmatch(r"(.+)_v(\d+)_pmc\.md", p.name)
        assert m, p.name
        workload, variant = m.group(1), int(m.group(2))
        text = p.read_text()
        assert f"variant: {variant}" in text, p.name
        assert "rocprofv3 --pmc FETCH_SIZE" in text and "rocprofv3 --pmc WRITE_SIZE" in text, p.name
        traffic = bench.traffic_from_profiles(workload, variant)
        assert traffic is not None and 1e8 < traffic < 1e11, (p.name, traffic)


def test_roofline_fraction_is_per_launch():
    """A two-step launch must still read and write every population once:
    frac counts those bytes over the launch duration and stays <= 1 at the
    measured speeds, while effective_frac credits 72 B per cell for every
    step."""
    bench = load_bench()
    cells = 4096 * 4096
    r = bench.roofline(72, cells, 117.98e-6, 2, 1.2015e9)
    assert r["algorithmic_bytes_per_launch"] == 72 * cells
    assert abs(r["launch_us"] - 235.96) < 1e-6
    assert abs(r["frac"] - 72 * cells / 235.96e-6 / 8e12) < 1e-4 and r["frac"] <= 1.0
    assert abs(r["effective_frac"] - 2 * r["frac"]) < 1e-3
    one = bench.roofline(72, cells, 202.0e-6, 1, None)
    assert one["frac"] == one["effective_frac"] and one["traffic_frac_of_peak"] is None
    # every committed byte-counter summary describes a launch below the peak
    for p in sorted((ROOT / "profiles").glob("d2q9_phys_*_pmc.md")):
        m = re.search(r"roofline frac ([0-9.]+)", p.read_text())
        if m:
            assert float(m.group(1)) <= 1.0, p.name
