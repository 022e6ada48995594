"""The oracle against the committed golden fixtures (numpy/scipy library outputs written
by tests/golden/make_golden.py)."""
import json
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_fft_fixture(oracle):
    f = load("fft_16x8.npz")
    a = f["a"]
    nx, ny = a.shape
    buf = np.zeros(2 * a.size)
    buf[0::2] = a.reshape(-1)
    oracle.or_fft2(nx, ny, oracle.p(buf), -1)
    got = (buf[0::2] + 1j * buf[1::2]).reshape(nx, ny)
    assert np.abs(got - f["fft"]).max() < 1e-12 * np.abs(f["fft"]).max()


@pytest.mark.parametrize("name", ["poisson_64x64.npz", "poisson_32x16.npz"])
def test_poisson_and_field_fixtures(oracle, name):
    f = load(name)
    lx, ly = float(f["lx"]), float(f["ly"])
    phi = oracle.poisson(f["rho"], lx, ly)
    ex, ey = oracle.efield(phi, lx, ly)
    for got, want in ((phi, f["phi"]), (ex, f["ex"]), (ey, f["ey"])):
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_poisson_512_sampled_fixture(oracle):
    f = load("poisson_512_sample.npz")
    rho = np.random.default_rng(512).standard_normal((512, 512))
    # the regenerated input must be the one the fixture was made from
    assert np.allclose(f["rho_checksum"], [rho.sum(), (rho * rho).sum()], rtol=1e-14,
                       atol=1e-9)
    lx = ly = 4 * math.pi
    phi = oracle.poisson(rho, lx, ly)
    ex, ey = oracle.efield(phi, lx, ly)
    idx = f["idx"]
    for got, want in ((phi, f["phi"]), (ex, f["ex"]), (ey, f["ey"])):
        g = got.reshape(-1)[idx]
        assert np.abs(g - want).max() < 1e-11 * np.abs(want).max()


def test_dispersion_roots():
    """The fixture's roots are the textbook values the physics tests fit against."""
    with open(os.path.join(GOLDEN, "landau_root.json")) as fh:
        r = json.load(fh)
    assert abs(r["omega_r"] - 1.4157) < 1e-4 and abs(r["omega_i"] + 0.1534) < 1e-4
    assert abs(r["two_stream_gamma"] - 0.340625) < 1e-6
