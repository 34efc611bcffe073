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


def test_learned_step_traffic_is_recorded():
    """The learned workload's line reads the matrix-core step's measured
    bytes per step from its profile; they sit between one pass of the
    lattice (72 B per cell of the 2048^2 coarse lattice) and twice that."""
    bench = load_bench()
    name = "d2q9_learned_mfma_2048x2048_trace"
    text = (ROOT / "profiles" / f"{name}.md").read_text()
    assert "rocprofv3 --pmc FETCH_SIZE" in text and "rocprofv3 --pmc WRITE_SIZE" in text
    one_pass = 72 * 2048 * 2048
    for key in ("hbm_bytes_per_launch_bf16", "hbm_bytes_per_launch_f16"):
        traffic = bench.traffic_from_profile_key(name, key)
        assert traffic is not None and one_pass <= traffic < 2 * one_pass, (key, traffic)
    assert bench.traffic_from_profile_key(name, "no_such_key") is None


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


def test_rho_u_errors_pool_domains():
    """The accuracy legs pool several domains: the relative L2 errors are
    sums of squares over every sampled fluid cell (not a mean of per-domain
    ratios), solid cells are left out, and u compares the velocity vectors
    (2-D or 3-D)."""
    import numpy as np

    bench = load_bench()
    fl = np.array([[True, True], [True, False]])
    ref1 = (np.full((2, 2), 1.0), np.full((2, 2), 0.1), np.zeros((2, 2)))
    dev1 = (ref1[0] + 1e-3, ref1[1] + 3e-4, ref1[2] + 4e-4)  # |du| = 5e-4
    dev1[0][1, 1] = 99.0  # a solid cell: ignored
    ref2 = (np.full((1, 3), 2.0), np.full((1, 3), 0.2), np.zeros((1, 3)))
    dev2 = (ref2[0], ref2[1], ref2[2])
    out = bench.rho_u_errors([(dev1, ref1, fl), (dev2, ref2, np.ones((1, 3), bool))])
    # rho: 3 cells off by 1e-3 against 3 cells of 1 and 3 of 2
    assert abs(out["rho_l2_rel"] - np.sqrt(3e-6 / (3 + 12))) < 1e-12
    assert abs(out["rho_max_abs"] - 1e-3) < 1e-12
    # u: 3 cells off by 5e-4 against 3 of 0.1 and 3 of 0.2
    assert abs(out["u_l2_rel"] - np.sqrt(3 * 25e-8 / (3 * 0.01 + 3 * 0.04))) < 1e-12
    assert abs(out["u_max_abs"] - 5e-4) < 1e-12 and abs(out["u_max"] - 0.2) < 1e-12
    # three components
    r3 = (np.ones(4), np.zeros(4), np.zeros(4), np.full(4, 0.1))
    d3 = (np.ones(4), np.full(4, 2e-4), np.zeros(4), np.full(4, 0.1))
    o3 = bench.rho_u_errors([(d3, r3, np.ones(4, bool))])
    assert o3["rho_l2_rel"] == 0.0 and abs(o3["u_max_abs"] - 2e-4) < 1e-15
