"""bin/pic_main: the reference-flag-compatible driver.  CPU tests cover argument handling
(no device call); the GPU test runs a small Landau case and compares its diagnostics file
with the same run through the Python context API."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "msc-thesis_amd", "bin", "pic_main")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="bin/pic_main not built")


def run(*args, timeout=120):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def test_help():
    r = run("-h")
    assert r.returncode == 0 and "-P reference|landau|twostream|mcc" in r.stdout
    assert "[-o FILE]" in r.stdout


@pytest.mark.parametrize("args", [["-l", "48"], ["-n", "0"], ["-x", "-3"], ["-p", "f16"],
                                  ["-P", "nope"], ["-P", "mcc", "-p", "f64"]])
def test_invalid_arguments_exit_with_status_1(args):
    r = run(*args)
    assert r.returncode == 1
    assert "Error" in (r.stdout + r.stderr)


def test_skip_gpu_reports_the_configuration():
    """-g (the reference's skip-GPU flag): the run is described and nothing else happens;
    -t/-d give the reference's step count ceil(tmax / dt)."""
    r = run("-l", "64", "-n", "128", "-m", "64", "-t", "1e-5", "-d", "1e-7", "-g")
    assert r.returncode == 0
    assert "grid 64 x 64" in r.stdout and "8192 particles" in r.stdout
    assert "100 steps" in r.stdout and "skipped" in r.stdout


@pytest.mark.gpu
def test_landau_run_matches_context_api(tmp_path):
    from pic_amd import Simulation, _lib, load_landau
    out = tmp_path / "diag.csv"
    r = run("-P", "landau", "-l", "64", "-q", "16", "-N", "40", "-p", "f32", "-f", str(out))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "particle-steps/s" in r.stdout
    got = np.genfromtxt(out, delimiter=",", names=True)
    assert got.shape == (40,)
    L = 4 * math.pi
    parts = load_landau(64, 64, L, L, 16, vth=1.0, alpha=0.01, mode=1, seed=1234)
    cfg = _lib.default_config(nx=64, ny=64, lx=L, ly=L, dt=0.05, precision=_lib.PIC_FP32,
                              seed=1234, capacity=parts[0].size)
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.step(40)
        sim.sync()
        ref = tmp_path / "ref.csv"
        sim.write_diagnostics_csv(ref, 0, 40)
    want = np.genfromtxt(ref, delimiter=",", names=True)
    for name in want.dtype.names:
        assert np.array_equal(got[name], want[name]), name


@pytest.mark.gpu
def test_reference_defaults_match_the_oracle(tmp_path, oracle):
    """pic_main with the reference driver's defaults (thesisCode/collisionless/double/src/
    main.cc:41-72, inc/params.h:21-47: 10 x 10 domain, 10 x 10 particle lattice, tmax 1e-4,
    dt 1e-7, so 1000 steps, FP64; the grid is 32^2, the default power of two) loaded by the
    reference's own initialiser (main.cc:201-244, pic_load_reference): the final positions
    and velocities it writes (-o) against the oracle's FP64 run from the same initial state
    (or_init_reference + or_run_f64)."""
    out = tmp_path / "final.csv"
    r = run("-o", str(out))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "1000 steps" in r.stdout and "100 particles" in r.stdout
    got = np.genfromtxt(out, delimiter=",", names=True)
    n = 100
    x, y, vx, vy = (np.zeros(n) for _ in range(4))
    oracle.or_init_reference(10, 10, 10, 10, 32, 32, 1.0, 1.0, 0.0, 0.0,
                             *(oracle.p(a) for a in (x, y, vx, vy)))
    x0 = x.copy()
    oracle.run_f64(x, y, vx, vy, 32, 32, 10.0, 10.0, 1e-7, -1.0, -(10.0 * 10.0) / n, 1000)
    assert np.abs(x - x0).max() > 0  # the particles moved
    # fixed-point CIC weights (2^-30 of a cell) against the oracle's continuous weights
    for name, want in zip(("x", "y", "vx", "vy"), (x, y, vx, vy)):
        assert np.abs(got[name] - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), name
