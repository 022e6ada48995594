"""The oracle's restatement of the reference's own collide arithmetic against the
reference code itself, compiled into oracle/_ref (`make -C oracle ref`):

  MCC  thesisCode/mccCollisions/src/particle.cc:112-127 (Particle::collision, prob 0.8
       hard-coded), called for every particle in order each step;
  DSMC thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:269-350 (cellTracking +
       collisions: nearest partner in the cell, kept when its index is larger, q = 0.2,
       v_i = (m1, m2), v_t = (-m1, -m2)).

Both sides consume the same std::mt19937 gen{10} stream (the reference draws it itself;
the oracle reads it from or_std_uniform_stream); velocities must agree bit for bit.  The
oracle alone is also pinned by the stream's known answer (the 10000th mt19937 output) and by
closed forms of the reset (mean 0, variance 1 per component: 2 x (sum of three U(0,1) -
1.5)).  Skipped where the reference checkout (and so oracle/_ref) is absent."""
import math

import numpy as np
import pytest

from _oracle import load_ref_collide

REF = load_ref_collide()
need_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref collide libraries not built")


def dp(a):
    import ctypes
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stream(oracle, seed, count):
    u = np.zeros(count)
    oracle.or_std_uniform_stream(seed, count, dp(u))
    return u


def test_uniform_stream_known_answer(oracle):
    """libstdc++'s uniform_real_distribution<double> takes two mt19937 words per value,
    (w1 + w2 2^32) / 2^64: value 5000 uses outputs 9999 and 10000, and the 10000th output of
    mt19937 is 4123659995 (ISO C++ [rand.predef])."""
    u = stream(oracle, 5489, 5000)
    w10000 = 4123659995
    lo = u[4999] * 2.0 ** 64 - w10000 * 2.0 ** 32
    assert 0 <= lo < 2 ** 32
    assert np.all((u >= 0) & (u < 1))


def test_mcc_reset_moments(oracle):
    """Every particle collides at prob = 1: the new components are 2(U1+U2+U3-1.5), mean 0
    and variance 1 (Bird's approximate unit normal)."""
    n = 200_000
    vx, vy = np.full(n, 7.0), np.full(n, 7.0)
    u = stream(oracle, 3, 7 * n)
    k = oracle.or_mcc_reference(n, dp(vx), dp(vy), 1.0, dp(u), u.size)
    assert k == 7 * n
    for v in (vx, vy):
        assert abs(v.mean()) < 4 / math.sqrt(n)
        assert abs(v.var() - 1.0) < 4 * math.sqrt(2.0 / n)  # 4 x 3 x 1/12 = 1
        assert v.min() >= -3.0 and v.max() <= 3.0
    # prob = 0: nothing changes, one uniform per particle
    k0 = oracle.or_mcc_reference(n, dp(vx), dp(vy), 0.0, dp(u), u.size)
    assert k0 == n


@need_ref
@pytest.mark.parametrize("n,nsteps", [(1, 3), (5000, 1), (20000, 3)])
def test_mcc_oracle_equals_reference(oracle, n, nsteps):
    rng = np.random.default_rng(n)
    vx0, vy0 = rng.normal(size=n), rng.normal(size=n)
    rvx, rvy = vx0.copy(), vy0.copy()
    REF[0].ref_mcc_collide(n, dp(rvx), dp(rvy), 10, nsteps)
    u = stream(oracle, 10, 7 * n * nsteps + 7)
    ovx, ovy = vx0.copy(), vy0.copy()
    k = 0
    for _ in range(nsteps):
        used = oracle.or_mcc_reference(n, dp(ovx), dp(ovy), 0.8, dp(u[k:]), u.size - k)
        assert used > 0
        k += used
    assert np.array_equal(ovx, rvx) and np.array_equal(ovy, rvy)
    assert np.count_nonzero(ovx != vx0) > 0.7 * n


def dsmc_case(rng, n, nx, L):
    x = rng.uniform(0, L, n)
    y = rng.uniform(0, L, n)
    return x, y, rng.normal(size=n), rng.normal(size=n)


@need_ref
@pytest.mark.parametrize("n,nx,nsteps", [(2, 4, 1), (3000, 16, 1), (20000, 32, 3)])
def test_dsmc_oracle_equals_reference(oracle, n, nx, nsteps):
    import ctypes
    L = float(nx)
    rng = np.random.default_rng(nx)
    x, y, vx0, vy0 = dsmc_case(rng, n, nx, L)
    if n == 2:
        x[:], y[:] = [0.3, 0.6], [0.4, 0.4]  # one cell, one pair
    rvx, rvy = vx0.copy(), vy0.copy()
    REF[1].ref_dsmc_collide(n, dp(x), dp(y), dp(rvx), dp(rvy), nx, nx, 1.0, 1.0, 10, nsteps)
    u = stream(oracle, 10, 7 * n * nsteps + 7)
    ovx, ovy = vx0.copy(), vy0.copy()
    k, tot = 0, 0
    for _ in range(nsteps):
        nc = ctypes.c_int64()
        used = oracle.or_dsmc_reference(n, dp(x), dp(y), dp(ovx), dp(ovy), nx, nx, L, L, 0.2,
                                        dp(u[k:]), u.size - k, ctypes.byref(nc))
        assert used >= 0
        k += used
        tot += nc.value
    assert np.array_equal(ovx, rvx) and np.array_equal(ovy, rvy)
    assert tot > 0 or n == 2
