"""The oracle against the reference's own deposit, gather and drift code, compiled from
thesisCode/collisionless/double/src/{grid.cc,particle.cc} into oracle/_ref/libref.so
(`make -C oracle ref`), through the documented bridges:

  deposit (grid.cc:46-76): the reference's area weights are the CIC weights of the in-cell
    reflection x~ = (2 cx + 1) dx - x, its charge is q w (not divided by the cell area),
    and node (i, j) sits at (1+i)(ny+2) + (1+j) of a ghost-padded array;
  gather (particle.cc:37-60): the same reflected weights, read from the padded field at
    (i, j) without the ghost offset, times q/m;
  drift (particle.cc:69-106): from a fresh particle (no stored force) x += dt v with a
    single wrap.
Skipped when the reference checkout (and so libref.so) is absent."""
import ctypes
import os

import numpy as np
import pytest

from _oracle import REF_SO

pytestmark = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built")

_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int


@pytest.fixture(scope="module")
def ref():
    lib = ctypes.CDLL(REF_SO)
    lib.ref_charge_assignment.argtypes = [_I, _DP, _DP, _I, _I, _I, _I, _DP]
    lib.ref_force.argtypes = [_I, _DP, _DP, _I, _I, _I, _I, _DP, _DP, _DP, _DP]
    lib.ref_drift.argtypes = [_I, _DP, _DP, _DP, _DP, _I, _I, _I, _I, _D]
    return lib


def p(a):
    return a.ctypes.data_as(_DP)


def reflect(x, d):
    c = np.floor(x / d)
    return (2 * c + 1) * d - x


def interior_positions(rng, n, L, m):
    """Positions away from cell edges (the reflection maps a cell onto itself only inside)."""
    d = L / m
    c = rng.integers(0, m, n)
    return (c + rng.uniform(0.01, 0.99, n)) * d


@pytest.mark.parametrize("xdim,ydim,nx,ny", [(10, 10, 16, 16), (12, 6, 32, 8)])
def test_reference_deposit_is_cic_at_reflected_positions(ref, oracle, xdim, ydim, nx, ny):
    rng = np.random.default_rng(nx + ny)
    n = 4000
    x = interior_positions(rng, n, xdim, nx)
    y = interior_positions(rng, n, ydim, ny)
    pad = np.zeros((nx + 2) * (ny + 2))
    ref.ref_charge_assignment(n, p(x), p(y), xdim, ydim, nx, ny, p(pad))
    got = pad.reshape(nx + 2, ny + 2)[1:nx + 1, 1:ny + 1]
    dx, dy = xdim / nx, ydim / ny
    q = -100.0                                   # the reference's default electronCharge
    want = oracle.deposit_f64(reflect(x, dx), reflect(y, dy), nx, ny, xdim, ydim, q) * dx * dy
    assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()
    assert abs(got.sum() - n * q) < 1e-9 * n * abs(q)


def test_reference_gather_is_cic_at_reflected_positions(ref, oracle):
    xdim, ydim, nx, ny = 10, 10, 16, 16
    rng = np.random.default_rng(3)
    n = 2000
    x = interior_positions(rng, n, xdim, nx)
    y = interior_positions(rng, n, ydim, ny)
    ex, ey = rng.normal(size=(nx, ny)), rng.normal(size=(nx, ny))
    # the reference reads node (i, j) at i*(ny+2)+j: place the field there
    exp_, eyp = np.zeros((nx + 2, ny + 2)), np.zeros((nx + 2, ny + 2))
    exp_[:nx, :ny] = ex
    eyp[:nx, :ny] = ey
    fx, fy = np.zeros(n), np.zeros(n)
    ref.ref_force(n, p(x), p(y), xdim, ydim, nx, ny, p(exp_.reshape(-1)), p(eyp.reshape(-1)),
                  p(fx), p(fy))
    dx, dy = xdim / nx, ydim / ny
    epx, epy = np.zeros(n), np.zeros(n)
    xr, yr = reflect(x, dx), reflect(y, dy)
    oracle.or_gather_f64(n, oracle.p(xr), oracle.p(yr), oracle.p(ex), oracle.p(ey), nx, ny,
                         float(xdim), float(ydim), oracle.p(epx), oracle.p(epy))
    qm = -100.0 / 100.0                          # electronCharge / mass
    assert np.abs(fx - qm * epx).max() < 1e-12
    assert np.abs(fy - qm * epy).max() < 1e-12


def test_reference_drift_matches_oracle_push_without_field(ref, oracle):
    xdim, ydim, nx, ny = 10, 8, 16, 16
    rng = np.random.default_rng(4)
    n = 3000
    x = rng.uniform(0, xdim, n)
    y = rng.uniform(0, ydim, n)
    vx, vy = rng.normal(0, 3, n), rng.normal(0, 3, n)
    dt = 0.1
    xr, yr = x.copy(), y.copy()
    ref.ref_drift(n, p(xr), p(yr), p(vx), p(vy), xdim, ydim, nx, ny, dt)
    xo, yo, vxo, vyo = x.copy(), y.copy(), vx.copy(), vy.copy()
    z = np.zeros((nx, ny))
    oracle.push_f64(xo, yo, vxo, vyo, z, z, float(xdim), float(ydim), dt, -1.0, 1.0)
    # identical except where the reference keeps x == L (its "> L" wrap test)
    assert np.abs(xr - xo).max() < 1e-13 and np.abs(yr - yo).max() < 1e-13
