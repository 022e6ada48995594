"""The oracle's collision stages pinned on the CPU by what the physics and the definitions
fix, independently of the GPU.

DSMC (or_dsmc_step_f32, NTC hard spheres in rounds of disjoint pairs):
  * the slot permutation of a round is a bijection, so the pairs of a round are disjoint;
  * over many rounds every slot meets every other slot about equally often (a broken
    permutation, e.g. pairing neighbours, fails the chi-square bound);
  * accepted collisions per step have mean sum_c 1/2 n_c (n_c - 1) coeff <g> with
    <g> = sqrt(pi) for unit 2D Maxwellians (3 sigma over 8 seeds: a dropped 1/2 or n^2 for
    n(n-1) fails);
  * each collision keeps momentum and energy (FP32 rounding), and an anisotropic
    Maxwellian relaxes towards isotropy;
  * the post-collision relative velocity has a uniform direction.
MCC (or_mcc_step_f32): counts per process within 3 sigma of the null-collision
expectation over 8 seeds, and the elastic recoil energy loss 2 m/M (1 - cos chi) with the
argon mass.
"""
import math

import numpy as np
import pytest

KE, KME = 1.602176634e-19, 9.1093837015e-31
M_AR = 39.948 * 1.66053906660e-27


def perm(oracle, n, cell, step, rnd, seed):
    out = np.zeros(n, np.uint32)
    oracle.or_dsmc_pair_perm(n, cell, step, rnd, seed, out.ctypes.data)
    return out


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 16, 17, 63, 64, 65, 255, 1024, 1025, 3001])
def test_dsmc_round_permutation_is_a_bijection(oracle, n):
    for rnd in range(4):
        p = perm(oracle, n, 12345, 7, rnd, 99)
        assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("n", [5, 8, 13, 33])
def test_dsmc_round_pairing_is_uniform(oracle, n):
    """Partner of slot 0 over 8000 rounds of an n-slot cell: each of the n - 1 partners
    about equally often (chi-square below its 0.999 quantile).  Pairing neighbours, or a
    permutation with too few mixing rounds, fails this."""
    from scipy.stats import chi2 as chi2_dist
    rounds = 8000
    hits = np.zeros(n, np.int64)
    for r in range(rounds):
        p = perm(oracle, n, 3, r // 50, r % 50, 5)
        pos = int(np.nonzero(p == 0)[0][0])  # pairs are (p[2m], p[2m+1]); slot n-1 of an
        if pos == n - 1 and n % 2:           # odd cell sits the round out
            continue
        hits[p[pos ^ 1]] += 1
    tot = hits.sum()
    exp = tot / (n - 1)
    chi2 = ((hits[1:] - exp) ** 2 / exp).sum()
    assert hits[0] == 0
    assert chi2 < chi2_dist.ppf(0.999, n - 2), (chi2, hits)
    if n % 2:  # slot 0 is left out of a round 1/n of the time
        q = 1 / n
        assert abs((rounds - tot) - rounds * q) < 4 * math.sqrt(rounds * q * (1 - q))


def cells(rng, ncells, n, sy=1.0):
    keys = np.sort(rng.integers(0, ncells, n)).astype(np.int32)
    off = np.searchsorted(keys, np.arange(ncells + 1)).astype(np.int32)
    vx = rng.normal(size=n).astype(np.float32)
    vy = (sy * rng.normal(size=n)).astype(np.float32)
    return off, vx, vy


def dsmc(oracle, off, vx, vy, gmax, coeff, seed, step):
    return oracle.or_dsmc_step_f32(off.size - 1, oracle.p(off), oracle.p(vx), oracle.p(vy),
                                   oracle.p(gmax), coeff, seed, step)


def test_dsmc_accepted_rate_over_seeds(oracle):
    """NTC: E[accepted] = sum_c 1/2 n (n-1) coeff <g>, <g> = sqrt(pi); 8 seeds, ensemble
    mean within 3 sigma (Poisson)."""
    rng = np.random.default_rng(6)
    ncells, n, coeff = 1024, 60_000, 5e-4
    tot, mu = 0, 0.0
    for seed in range(8):
        off, vx, vy = cells(rng, ncells, n)
        gmax = np.full(ncells, 8.0, np.float32)
        tot += dsmc(oracle, off, vx, vy, gmax, coeff, 100 + seed, 0)
        c = np.diff(off).astype(np.float64)
        mu += (0.5 * c * (c - 1) * coeff * math.sqrt(math.pi)).sum()
    assert abs(tot - mu) <= 3 * math.sqrt(mu), (tot, mu)


def test_dsmc_conserves_and_relaxes(oracle):
    rng = np.random.default_rng(5)
    ncells, n = 256, 40_000
    off, vx, vy = cells(rng, ncells, n, sy=2.0)  # Ty = 4 Tx
    gmax = np.full(ncells, 6.0, np.float32)
    p0 = np.array([vx.astype(np.float64).sum(), vy.astype(np.float64).sum()])
    e0 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
    r0 = np.var(vy) / np.var(vx)
    ntot = 0
    for s in range(20):
        ntot += dsmc(oracle, off, vx, vy, gmax, 0.01, 7, s)
    p1 = np.array([vx.astype(np.float64).sum(), vy.astype(np.float64).sum()])
    e1 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
    assert ntot > n
    assert np.abs(p1 - p0).max() < 2e-6 * math.sqrt(ntot) * 4   # FP32 rounding per collision
    assert abs(e1 - e0) < 1e-5 * e0
    assert abs(np.var(vy) / np.var(vx) - 1) < 0.2 * abs(r0 - 1)
    assert np.all(gmax >= 6.0)


def test_dsmc_rounds_in_one_large_cell(oracle):
    """One cell of 3001 particles with many rounds per step: momentum and energy kept, and
    the accepted count matches the rate formula (3 sigma)."""
    rng = np.random.default_rng(8)
    n, coeff = 3001, 2e-3
    off = np.array([0, n], np.int32)
    tot, mu = 0, 0.0
    for seed in range(4):
        vx = rng.normal(size=n).astype(np.float32)
        vy = rng.normal(size=n).astype(np.float32)
        e0 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
        gmax = np.array([8.0], np.float32)
        tot += dsmc(oracle, off, vx, vy, gmax, coeff, seed, 3)
        mu += 0.5 * n * (n - 1) * coeff * math.sqrt(math.pi)
        e1 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
        assert abs(e1 - e0) < 1e-5 * e0
    assert 0.5 * n * (n - 1) * coeff * 8.0 > 5 * (n // 2)  # several rounds per step
    assert abs(tot - mu) <= 3 * math.sqrt(mu), (tot, mu)


def test_dsmc_scattering_direction_is_uniform(oracle):
    """Two particles per cell, all candidates accepted (g = gmax): the new relative velocity
    direction is uniform on the circle (first and second trigonometric moments ~ 0)."""
    ncells = 20_000
    off = (2 * np.arange(ncells + 1)).astype(np.int32)
    vx = np.zeros(2 * ncells, np.float32)
    vy = np.zeros(2 * ncells, np.float32)
    vx[0::2], vx[1::2] = 1.0, -1.0  # g = 2 along +x
    gmax = np.full(ncells, 2.0, np.float32)
    # coeff so that M = 1 exactly: 1/2 * 2 * 1 * coeff * 2 = 1
    nacc = dsmc(oracle, off, vx, vy, gmax, 0.5, 3, 0)
    assert nacc == ncells
    gx, gy = vx[0::2] - vx[1::2], vy[0::2] - vy[1::2]
    ang = np.arctan2(gy, gx)
    assert np.allclose(np.hypot(gx, gy), 2.0, atol=1e-6)
    for k in (1, 2):
        assert abs(np.cos(k * ang).mean()) < 4 / math.sqrt(ncells)
        assert abs(np.sin(k * ang).mean()) < 4 / math.sqrt(ncells)
    assert np.allclose(vx[0::2] + vx[1::2], 0.0, atol=1e-6)


def mcc_params(ng=1e21, te=2.0, ne=1e15, numax_dt_target=0.3, mass=M_AR):
    """Normalised-unit MCC parameters as the engine's preset builds them (scaled up to
    nu_max dt = 0.3 so one pass has many events)."""
    eps0 = 8.8541878128e-12
    wp = math.sqrt(ne * KE * KE / (eps0 * KME))
    vu = math.sqrt(te * KE / KME)
    es = np.linspace(1e-3, 1000.0, 200_001)
    sel = 1e-19 / (1 + es / 10)
    sex = np.where(es > 11.5, 2e-21 * (1 - 11.5 / es), 0)
    siz = np.where(es > 15.8, 3e-20 * np.log(np.maximum(es, 15.8) / 15.8) / es, 0)
    numax = (ng * (sel + sex + siz) * np.sqrt(2 * es * KE / KME)).max()
    dt = numax_dt_target / (numax / wp)
    return dict(e_per_v2=0.5 * KME * vu * vu / KE, nu_dt=ng * vu * (1 / wp) * dt,
                p_coll=1 - math.exp(-numax_dt_target), numax_dt=numax_dt_target,
                m2=2 * KME / mass)


def mcc_pass(oracle, p, vx, vy, seed, step, cap=None):
    n = vx.size
    cap = cap or n + n // 2
    arr = [np.zeros(cap, np.float32) for _ in range(4)]
    arr[2][:n], arr[3][:n] = vx, vy
    ids = np.zeros(cap, np.int32)
    ids[:n] = np.arange(n)
    cnt = np.zeros(4, np.int64)
    m = oracle.or_mcc_step_f32(n, cap, *(oracle.p(a) for a in arr), oracle.p(ids), n,
                               p["e_per_v2"], p["nu_dt"], p["p_coll"], p["numax_dt"], p["m2"],
                               seed, step, oracle.p(cnt))
    return arr, ids, cnt, m


def hot_maxwellian(rng, n):
    vx = rng.normal(size=n)
    vy = rng.normal(size=n)
    hot = rng.random(n) < 0.3
    vx[hot] *= 3.5
    vy[hot] *= 3.5
    return vx.astype(np.float32), vy.astype(np.float32)


def test_mcc_rates_over_seeds(oracle):
    """Counts of elastic, excitation and ionisation events summed over 8 seeded passes lie
    within 3 sigma of sum p_coll nu_k(E) / nu_max."""
    p = mcc_params()
    rng = np.random.default_rng(2)
    tot = np.zeros(3)
    mu = np.zeros(3)
    var = np.zeros(3)
    for seed in range(8):
        vx, vy = hot_maxwellian(rng, 50_000)
        _, _, cnt, m = mcc_pass(oracle, p, vx, vy, 1000 + seed, seed)
        assert m - vx.size == cnt[2]
        v2 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2
        e, sp = p["e_per_v2"] * v2, np.sqrt(v2)
        secs = (1e-19 / (1 + e / 10),
                np.where(e > 11.5, 2e-21 * (1 - 11.5 / np.maximum(e, 1e-30)), 0),
                np.where(e > 15.8, 3e-20 * np.log(np.maximum(e, 15.8) / 15.8) / np.maximum(e, 1e-30), 0))
        for k, s in enumerate(secs):
            q = p["p_coll"] * p["nu_dt"] * s * sp / p["numax_dt"]
            mu[k] += q.sum()
            var[k] += (q * (1 - q)).sum()
        tot += cnt[:3]
    for k in range(3):
        assert abs(tot[k] - mu[k]) <= 3 * math.sqrt(var[k]), (k, tot[k], mu[k])


def test_mcc_elastic_recoil_with_argon_mass(oracle):
    """Elastic scattering off argon loses the fraction 2 m/M (1 - cos chi) of the
    electron's energy.  Cold electrons (E well below the excitation threshold) make every
    event elastic; the relative energy loss summed over the scattered electrons equals
    2 m/M sum (1 - cos chi), and with chi uniform its mean is 2 m/M (1 +- a few percent)."""
    p = mcc_params()
    rng = np.random.default_rng(9)
    n = 200_000
    vx = (0.5 * rng.normal(size=n)).astype(np.float32)  # E = v^2 x 1 eV: 0.5 eV mean
    vy = (0.5 * rng.normal(size=n)).astype(np.float32)
    arr, _, cnt, m = mcc_pass(oracle, p, vx, vy, 31, 4)
    assert cnt[1] == cnt[2] == 0 and m == n
    e0 = vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2
    e1 = arr[2][:n].astype(np.float64) ** 2 + arr[3][:n].astype(np.float64) ** 2
    v0 = np.stack([vx, vy], 1).astype(np.float64)
    v1 = np.stack([arr[2][:n], arr[3][:n]], 1).astype(np.float64)
    hit = np.any(v1 != v0, axis=1)
    assert abs(hit.sum() - cnt[0]) <= 2  # a scattering by exactly 0 may leave v unchanged
    cosc = (v0[hit] * v1[hit]).sum(1) / np.sqrt(e0[hit] * e1[hit])
    loss = (e0[hit] - e1[hit]) / e0[hit]
    # per event: loss = 2m/M (1 - cos chi), up to FP32 rounding of the stored velocity
    np.testing.assert_allclose(loss, p["m2"] * (1 - cosc), rtol=0, atol=3e-7)
    assert abs(loss.mean() / p["m2"] - 1) < 0.03
