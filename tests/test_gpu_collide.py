"""GPU collision stages (MCC null-collision, NTC DSMC) against the oracle mirrors, closed-form
rates and conservation laws.

Per-particle parity: with the same Philox streams and the same bit-reproducible sincos/log,
the GPU operators reproduce the oracle's velocities, births and counts bit for bit.
Statistics: collision counts within 3 sigma of the null-collision expectation.
"""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from pic_amd import Simulation, _lib, presets  # noqa: E402
from pic_amd._lib import check, lib  # noqa: E402

DEV = "cuda:0"
KE, KME = 1.602176634e-19, 9.1093837015e-31


def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def mcc_cfg(dt_scale=30.0):
    cfg = _lib.default_config(nx=64, ny=64, lx=64.0, ly=64.0, seed=77)
    check(lib.pic_mcc_preset(ctypes.byref(cfg), 1e21, 2.0, 1e15))
    cfg.dt *= dt_scale  # nu_max dt = 0.3: many events per pass
    return cfg


def mcc_params(cfg):
    numax_dt = lib.pic_mcc_nu_max(ctypes.byref(cfg))
    return dict(e_per_v2=0.5 * KME * cfg.mcc_v_unit ** 2 / KE,
                nu_dt=cfg.mcc_ng * cfg.mcc_v_unit * cfg.mcc_dt_phys * cfg.dt,
                p_coll=1.0 - math.exp(-numax_dt), numax_dt=numax_dt,
                m2=2.0 * KME / cfg.mcc_neutral_mass)


def maxwellian_with_tail(n, seed):
    rng = np.random.default_rng(seed)
    vx = rng.normal(size=n)
    vy = rng.normal(size=n)
    hot = rng.random(n) < 0.3   # a hot tail above the excitation / ionisation thresholds
    vx[hot] *= 3.5
    vy[hot] *= 3.5
    x = rng.uniform(0, 64, n)
    y = rng.uniform(0, 64, n)
    return [a.astype(np.float32) for a in (x, y, vx, vy)]


def run_mcc_op(arrs, cap, cfg, step, ids=None):
    n = arrs[0].size
    ts = [torch.zeros(cap, dtype=torch.float32, device=DEV) for _ in range(4)]
    for t, a in zip(ts, arrs):
        t[:n] = torch.from_numpy(a)
    tid = torch.zeros(cap, dtype=torch.int32, device=DEV)
    tid[:n] = torch.arange(n, dtype=torch.int32) if ids is None else torch.from_numpy(ids)
    cnt = torch.zeros(4, dtype=torch.int64, device=DEV)
    nio = ctypes.c_int64(n)
    check(lib.pic_op_mcc(*(dptr(t) for t in ts), dptr(tid), ctypes.byref(nio), cap,
                         ctypes.byref(cfg), step, dptr(cnt), None))
    torch.cuda.synchronize()
    m = nio.value
    return [t[:m].cpu().numpy() for t in ts], tid[:m].cpu().numpy(), cnt.cpu().numpy()


def test_mcc_matches_oracle_bitwise(oracle):
    cfg = mcc_cfg()
    n, cap = 150_000, 200_000
    arrs = maxwellian_with_tail(n, 1)
    (gx, gy, gvx, gvy), gid, gcnt = run_mcc_op(arrs, cap, cfg, step=5)
    p = mcc_params(cfg)
    ref = [np.zeros(cap, np.float32) for _ in range(4)]
    for r, a in zip(ref, arrs):
        r[:n] = a
    rid = np.zeros(cap, np.int32)
    rid[:n] = np.arange(n)
    rcnt = np.zeros(4, np.int64)
    m = oracle.or_mcc_step_f32(n, cap, *(oracle.p(r) for r in ref), oracle.p(rid), n,
                               p["e_per_v2"], p["nu_dt"], p["p_coll"], p["numax_dt"], p["m2"],
                               cfg.seed, 5, oracle.p(rcnt))
    assert np.array_equal(gcnt, rcnt)
    assert gx.size == m and rcnt[2] == m - n > 0
    for g, r in zip((gx, gy, gvx, gvy), ref):
        assert np.array_equal(g, r[:m])
    assert np.array_equal(gid, rid[:m])


def test_mcc_with_permuted_ids_matches_oracle(oracle):
    """The draws are keyed by particle id, not by slot: particles in a shuffled order, with
    their ids, give the oracle's result on the same arrays (births appended in parent-id
    order)."""
    cfg = mcc_cfg()
    n, cap = 120_001, 160_000
    arrs = maxwellian_with_tail(n, 4)
    ids = np.random.default_rng(5).permutation(n).astype(np.int32)
    (gx, gy, gvx, gvy), gid, gcnt = run_mcc_op(arrs, cap, cfg, step=11, ids=ids)
    p = mcc_params(cfg)
    ref = [np.zeros(cap, np.float32) for _ in range(4)]
    for r, a in zip(ref, arrs):
        r[:n] = a
    rid = np.zeros(cap, np.int32)
    rid[:n] = ids
    rcnt = np.zeros(4, np.int64)
    m = oracle.or_mcc_step_f32(n, cap, *(oracle.p(r) for r in ref), oracle.p(rid), n,
                               p["e_per_v2"], p["nu_dt"], p["p_coll"], p["numax_dt"], p["m2"],
                               cfg.seed, 11, oracle.p(rcnt))
    assert np.array_equal(gcnt, rcnt) and rcnt[:3].sum() > 0
    assert gx.size == m
    for g, r in zip((gx, gy, gvx, gvy), ref):
        assert np.array_equal(g, r[:m])
    assert np.array_equal(gid, rid[:m])


def test_mcc_rates_within_3_sigma():
    """Counts per process summed over 8 seeded passes (different particles, seeds and step
    indices) within 3 sigma of sum p_coll nu_k(E) dt / (nu_max dt)."""
    cfg = mcc_cfg()
    p = mcc_params(cfg)
    n = 200_000
    tot = np.zeros(3)
    mu = np.zeros(3)
    var = np.zeros(3)
    for seed in range(8):
        arrs = maxwellian_with_tail(n, 20 + seed)
        cfg.seed = 1000 + seed
        _, _, cnt = run_mcc_op(arrs, n + n // 2, cfg, step=seed)
        v2 = arrs[2].astype(np.float64) ** 2 + arrs[3].astype(np.float64) ** 2
        e = p["e_per_v2"] * v2
        sp = np.sqrt(v2)
        sel = 1e-19 / (1 + e / 10)
        sex = np.where(e > 11.5, 2e-21 * (1 - 11.5 / np.maximum(e, 1e-30)), 0)
        siz = np.where(e > 15.8, 3e-20 * np.log(np.maximum(e, 15.8) / 15.8) / np.maximum(e, 1e-30), 0)
        for k, s_ in enumerate((sel, sex, siz)):
            prob = p["p_coll"] * p["nu_dt"] * s_ * sp / p["numax_dt"]
            mu[k] += prob.sum()
            var[k] += (prob * (1 - prob)).sum()
        tot += cnt[:3]
    for k in range(3):
        assert abs(tot[k] - mu[k]) <= 3 * math.sqrt(var[k]), (k, tot[k], mu[k])


def test_mcc_energy_bookkeeping():
    """Excitation removes 11.5 eV and ionisation 15.8 eV (shared by two electrons)."""
    cfg = mcc_cfg()
    cfg.mcc_neutral_mass = 1e6  # elastic energy loss 2 m/M -> negligible
    n = 200_000
    arrs = maxwellian_with_tail(n, 3)
    (_, _, vx, vy), _, cnt = run_mcc_op(arrs, n + n // 2, cfg, step=1)
    p = mcc_params(cfg)
    e0 = p["e_per_v2"] * (arrs[2].astype(np.float64) ** 2 + arrs[3].astype(np.float64) ** 2).sum()
    e1 = p["e_per_v2"] * (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
    loss = 11.5 * cnt[1] + 15.8 * cnt[2]
    assert abs((e0 - e1) - loss) <= 1e-5 * e0


def sorted_cells(ncells, n, seed, big=()):
    """Cell-sorted velocities: n particles over ncells random cells, plus cells holding the
    counts in `big` (appended after the others)."""
    rng = np.random.default_rng(seed)
    keys = np.sort(rng.integers(0, ncells, n)).astype(np.int32)
    off = list(np.searchsorted(keys, np.arange(ncells + 1)))
    for b in big:
        off.append(off[-1] + b)
    off = np.asarray(off, np.int32)
    m = int(off[-1])
    vx = rng.normal(size=m).astype(np.float32)
    vy = (2.0 * rng.normal(size=m)).astype(np.float32)  # anisotropic: Ty = 4 Tx
    return off, vx, vy


def run_dsmc_op(off, vx, vy, gmax, coeff, seed, step):
    tv = [torch.from_numpy(a.copy()).to(DEV) for a in (vx, vy)]
    toff = torch.from_numpy(off).to(DEV)
    tid = torch.arange(vx.size, dtype=torch.int32, device=DEV)
    tg = torch.from_numpy(gmax.copy()).to(DEV)
    nc = torch.zeros(1, dtype=torch.int64, device=DEV)
    check(lib.pic_op_dsmc(dptr(tv[0]), dptr(tv[1]), dptr(tid), dptr(toff), off.size - 1, dptr(tg),
                          coeff, seed, step, dptr(nc), None))
    torch.cuda.synchronize()
    return tv[0].cpu().numpy(), tv[1].cpu().numpy(), tg.cpu().numpy(), int(nc.item())


@pytest.mark.parametrize("case", ["cells4096", "big_cells"])
def test_dsmc_matches_oracle_bitwise(oracle, case):
    """Two consecutive passes (gmax carried) bit for bit against the oracle.  big_cells puts
    cells of 1024, 1025 and 3000 particles (many rounds of disjoint pairs per pass) next to
    ordinary cells of ~50."""
    if case == "cells4096":
        ncells, n, big, coeff = 4096, 200_000, (), 0.01
    else:
        ncells, n, big, coeff = 512, 25_000, (1024, 1025, 3000), 2e-3
    off, vx, vy = sorted_cells(ncells, n, 4, big)
    nc = off.size - 1
    gmax = np.full(nc, 4.0, np.float32)
    rvx, rvy, rg = vx.copy(), vy.copy(), gmax.copy()
    for step in (3, 4):
        print(f"dsmc op {case} step {step}", flush=True)
        vx, vy, gmax, gn = run_dsmc_op(off, vx, vy, gmax, coeff, 99, step)
        rn = oracle.or_dsmc_step_f32(nc, oracle.p(off), oracle.p(rvx), oracle.p(rvy),
                                     oracle.p(rg), coeff, 99, step)
        assert gn == rn > 0
        assert np.array_equal(vx, rvx) and np.array_equal(vy, rvy)
        assert np.array_equal(gmax, rg)
    if big:
        cnt = np.diff(off)
        assert (0.5 * cnt * (cnt - 1) * coeff * 4.0).max() > 10 * (cnt.max() // 2)


def test_dsmc_conservation_and_relaxation():
    ncells, n = 1024, 100_000
    off, vx, vy = sorted_cells(ncells, n, 5)
    gmax = np.full(ncells, 6.0, np.float32)
    p0 = np.array([vx.astype(np.float64).sum(), vy.astype(np.float64).sum()])
    e0 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
    tx0, ty0 = np.var(vx), np.var(vy)
    ntot = 0
    for s in range(30):
        vx, vy, gmax, nc = run_dsmc_op(off, vx, vy, gmax, 0.01, 7, s)
        ntot += nc
    p1 = np.array([vx.astype(np.float64).sum(), vy.astype(np.float64).sum()])
    e1 = (vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2).sum()
    assert np.abs(p1 - p0).max() < 1e-3 * math.sqrt(n)   # FP32 rounding of each collision
    assert abs(e1 - e0) < 1e-5 * e0
    # isotropisation: Ty/Tx goes from 4 towards 1
    assert abs(np.var(vy) / np.var(vx) - 1) < 0.5 * abs(ty0 / tx0 - 1)
    assert ntot > 0


def test_dsmc_accepted_rate_over_seeds():
    """NTC: accepted collisions per cell have mean 1/2 N(N-1) coeff <g> (<g> = sqrt(pi) for a
    unit 2D Maxwellian pair); the total over 8 seeded passes lies within 3 sigma (Poisson)
    of the sum of the expectations."""
    ncells, n, coeff = 2048, 200_000, 2e-4
    tot, mu = 0, 0.0
    for seed in range(8):
        rng = np.random.default_rng(60 + seed)
        keys = np.sort(rng.integers(0, ncells, n)).astype(np.int32)
        off = np.searchsorted(keys, np.arange(ncells + 1)).astype(np.int32)
        vx = rng.normal(size=n).astype(np.float32)
        vy = rng.normal(size=n).astype(np.float32)
        gmax = np.full(ncells, 8.0, np.float32)
        _, _, _, nc = run_dsmc_op(off, vx, vy, gmax, coeff, 11 + seed, 0)
        tot += nc
        cnt = np.diff(off).astype(np.float64)
        mu += (0.5 * cnt * (cnt - 1) * coeff * math.sqrt(math.pi)).sum()
    assert abs(tot - mu) <= 3 * math.sqrt(mu), (tot, mu)


@pytest.mark.parametrize("graphs", [True, False])
def test_engine_mcc_run(graphs):
    cfg, parts, _ = presets.mcc_256(ppc=8, n=64)
    cfg.dt *= 10
    with Simulation(cfg) as sim:
        sim.use_graphs(graphs)
        sim.set_particles(*parts)
        n0 = sim.num_particles()
        sim.step(40)
        sim.sync()
        sim.check_finite()
        cnt = sim.collision_counts()
        n1 = sim.num_particles()
        x, y, vx, vy = sim.particles()
    assert cnt[0] > 0 and n1 - n0 == cnt[2]
    assert np.all((x >= 0) & (x < cfg.lx)) and np.all((y >= 0) & (y < cfg.ly))
    assert x.size == n1


def test_engine_mcc_deterministic():
    cfg, parts, _ = presets.mcc_256(ppc=8, n=64)
    cfg.dt *= 10
    outs = []
    for graphs in (True, False):
        with Simulation(cfg) as sim:
            sim.use_graphs(graphs)
            sim.set_particles(*parts)
            sim.step(24)
            outs.append(sim.particles())
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_engine_dsmc_run():
    cfg, parts, _ = presets.dsmc_blob(n=64, mean_ppc=8, coeff=0.05)
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.step(10)
        sim.sync()
        sim.check_finite()
        cnt = sim.collision_counts()
        x, y, vx, vy = sim.particles()
    assert cnt[0] > 0
    assert np.all((x >= 0) & (x < cfg.lx))


@pytest.mark.slow
def test_engine_mcc_at_full_scale():
    """Config 3 at its own size: 256 x 256, 8,388,608 electrons at Te = 2 eV.  The first
    step's counts per collision type lie within 3 sigma of the null-collision expectation
    from the initial velocities (the field of the uniform plasma barely moves them in one
    step); over 50 steps every birth is an ionisation and the state stays finite and in
    the box."""
    import time
    cfg, parts, _ = presets.mcc_256()
    p = mcc_params(cfg)
    vx = np.asarray(parts[2], np.float32).astype(np.float64)
    vy = np.asarray(parts[3], np.float32).astype(np.float64)
    v2 = vx * vx + vy * vy
    e = p["e_per_v2"] * v2
    sp = np.sqrt(v2)
    sel = 1e-19 / (1 + e / 10)
    sex = np.where(e > 11.5, 2e-21 * (1 - 11.5 / np.maximum(e, 1e-30)), 0)
    siz = np.where(e > 15.8, 3e-20 * np.log(np.maximum(e, 15.8) / 15.8) / np.maximum(e, 1e-30), 0)
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        n0 = sim.num_particles()
        sim.step(1)
        sim.sync()
        cnt = sim.collision_counts()
        for k, s in enumerate((sel, sex, siz)):
            prob = p["p_coll"] * p["nu_dt"] * s * sp / p["numax_dt"]
            mu, sd = prob.sum(), math.sqrt((prob * (1 - prob)).sum())
            assert abs(cnt[k] - mu) <= 3 * sd + 1, (k, cnt[k], mu, sd)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(49)
        sim.sync()
        dt = time.perf_counter() - t0
        cnt = sim.collision_counts()
        n1 = sim.num_particles()
        sim.check_finite()
        x, y, _, _ = sim.particles()
    print(f"MCC 256^2: {n0 * 49 / dt:.3e} particle-steps/s over 49 steps (incl. launch), "
          f"counts {cnt}, births {n1 - n0}")
    assert n1 - n0 == cnt[2]
    assert np.all((x >= 0) & (x < cfg.lx)) and np.all((y >= 0) & (y < cfg.ly))


@pytest.mark.slow
def test_engine_dsmc_at_full_scale():
    """The DSMC workload at its own size (256 x 256, a Gaussian blob of 2,097,152 electrons
    over a uniform neutralising background, NTC hard-sphere collisions per cell on the
    cell-sorted state every step).  The blob's field energy is ~2000x the electrons'
    initial kinetic energy, so it expands and heats them and the collision rate climbs to
    many per particle per step.  Checked at this size: the whole step runs (collisions
    accepted, state finite and in the box), and the total energy and the momentum drift
    no more than in the same run without collisions (elastic collisions conserve both;
    the per-collision conservation is pinned at operator level)."""
    import time
    cfg, parts, _ = presets.dsmc_blob()

    def run(collide):
        c = _lib.default_config(nx=cfg.nx, ny=cfg.ny, lx=cfg.lx, ly=cfg.ly, dt=cfg.dt,
                                qm=cfg.qm, precision=cfg.precision, seed=cfg.seed,
                                collide=collide, dsmc_coeff=cfg.dsmc_coeff,
                                dsmc_gmax0=cfg.dsmc_gmax0)
        c.capacity = cfg.capacity
        with Simulation(c) as sim:
            sim.set_particles(*parts)
            n = sim.num_particles()
            sim.step(1)
            sim.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.step(39)
            sim.sync()
            dt = time.perf_counter() - t0
            diag = sim.diagnostics(0, 40)
            cnt = sim.collision_counts()
            sim.check_finite()
            x, y, _, _ = sim.particles()
        e = diag[:, 0] + diag[:, 1]
        drift = np.abs(e - e[0]).max() / e[0]
        pmax = np.abs(diag[:, 2:4]).max() / math.sqrt(2.0 * diag[0, 1] * cfg.lx * cfg.ly / n)
        return n * 39 / dt, cnt, drift, pmax, x, y, n

    rate, cnt, drift, pmax, x, y, n = run(_lib.PIC_COLLIDE_DSMC)
    _, _, drift0, pmax0, _, _, _ = run(_lib.PIC_COLLIDE_NONE)
    print(f"DSMC 256^2: {rate:.3e} particle-steps/s (incl. launch), collisions {cnt[0]}, "
          f"energy drift {drift:.2e} (no collisions {drift0:.2e}), momentum / thermal "
          f"{pmax:.2e} ({pmax0:.2e})")
    assert cnt[0] > n // 4, cnt
    assert drift < 1.1 * drift0 + 1e-4, (drift, drift0)
    assert pmax < 1.1 * pmax0 + 1e-4, (pmax, pmax0)
    assert np.all((x >= 0) & (x < cfg.lx)) and np.all((y >= 0) & (y < cfg.ly))
