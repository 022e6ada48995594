"""Engine runs with collisions against a step-by-step composition of the oracle's stages.

One engine step with collisions is: field of the current charge, gather + kick + drift
(half kick on the first step), deposit of the new positions, then the collision stage on
the stored (half-step) velocities with the step's index: MCC per particle (ionisation
births appended in parent-id order with the next ids), or DSMC per cell on the particles
ordered by cell key cx*ny + cy (true division) and ascending id inside a cell.  The
oracle composition below runs exactly those stages on arrays kept in id order; the
engine returns its particles in id order, so both must agree bit for bit.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pic_amd import Simulation, _lib, load_landau  # noqa: E402
from pic_amd._lib import check, lib  # noqa: E402

KE, KME = 1.602176634e-19, 9.1093837015e-31


def blob(n=64, ppc=8, seed=9, coeff=0.01, charge_scale=1e-3):
    rng = np.random.default_rng(seed)
    npart = n * n * ppc
    x = np.mod(rng.normal(n / 2, n / 6, npart), n)
    y = np.mod(rng.normal(n / 2, n / 6, npart), n)
    vx, vy = rng.normal(0, 1, npart), rng.normal(0, 1, npart)
    cfg = _lib.default_config(nx=n, ny=n, lx=float(n), ly=float(n), dt=0.2, qm=-1.0,
                              precision=_lib.PIC_FP32, seed=seed,
                              collide=_lib.PIC_COLLIDE_DSMC, dsmc_coeff=coeff, dsmc_gmax0=4.0,
                              capacity=npart)
    cfg.qp = -charge_scale * cfg.lx * cfg.ly / npart
    return cfg, (x, y, vx, vy)


def mcc_case(n=64, ppc=8, seed=5):
    cfg = _lib.default_config(nx=n, ny=n, lx=float(n), ly=float(n), precision=_lib.PIC_FP32,
                              seed=seed)
    check(lib.pic_mcc_preset(ctypes.byref(cfg), 1e21, 2.0, 1e15))
    cfg.dt *= 10.0  # nu_max dt = 0.1
    parts = load_landau(n, n, cfg.lx, cfg.ly, ppc, vth=2.0, alpha=0.0, mode=0, seed=seed)
    cfg.capacity = int(parts[0].size * 1.25)
    return cfg, parts


def field_f32(oracle, x, y, cfg, qp):
    acc = oracle.deposit_fixed(x, y, cfg.nx, cfg.ny, cfg.lx, cfg.ly)
    rho = oracle.fixed_to_rho(acc, cfg.nx, cfg.ny, cfg.lx, cfg.ly, qp)
    phi = oracle.poisson(rho, cfg.lx, cfg.ly)
    return oracle.efield(phi, cfg.lx, cfg.ly)


def oracle_run(oracle, cfg, parts, nsteps):
    """Arrays in id order after nsteps engine steps (stored velocities), + counts."""
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    n = parts[0].size
    cap = max(int(cfg.capacity), n)
    x, y, vx, vy = (np.zeros(cap, np.float32) for _ in range(4))
    x[:n], y[:n], vx[:n], vy[:n] = (f32(a) for a in parts)
    ids = np.zeros(cap, np.int32)
    ids[:n] = np.arange(n)
    qp = cfg.qp if cfg.qp != 0 else -(cfg.lx * cfg.ly) / n
    G = cfg.nx * cfg.ny
    gmax = np.full(G, cfg.dsmc_gmax0, np.float32)
    counts = np.zeros(4, np.int64)
    m = n
    if cfg.collide == _lib.PIC_COLLIDE_MCC:
        numax_dt = lib.pic_mcc_nu_max(ctypes.byref(cfg))
        mp = dict(e=0.5 * KME * cfg.mcc_v_unit ** 2 / KE,
                  nu=cfg.mcc_ng * cfg.mcc_v_unit * cfg.mcc_dt_phys * cfg.dt,
                  p=1.0 - math.exp(-numax_dt), numax=numax_dt,
                  m2=2.0 * KME / cfg.mcc_neutral_mass)
    for s in range(nsteps):
        xs, ys, vxs, vys = x[:m], y[:m], vx[:m], vy[:m]  # views
        ex, ey = field_f32(oracle, xs, ys, cfg, qp)
        oracle.push_f32(xs, ys, vxs, vys, ex, ey, cfg.lx, cfg.ly, cfg.dt, cfg.qm,
                        0.5 if s == 0 else 1.0)
        if cfg.collide == _lib.PIC_COLLIDE_DSMC:
            keys = oracle.cell_keys(xs, ys, cfg.nx, cfg.ny, cfg.lx, cfg.ly)
            order = np.argsort(keys, kind="stable")
            off = np.searchsorted(keys[order], np.arange(G + 1)).astype(np.int32)
            svx, svy = np.ascontiguousarray(vxs[order]), np.ascontiguousarray(vys[order])
            counts[0] += oracle.or_dsmc_step_f32(G, oracle.p(off), oracle.p(svx),
                                                 oracle.p(svy), oracle.p(gmax),
                                                 cfg.dsmc_coeff, cfg.seed, s)
            vx[order], vy[order] = svx, svy
        elif cfg.collide == _lib.PIC_COLLIDE_MCC:
            m = oracle.or_mcc_step_f32(m, cap, oracle.p(x), oracle.p(y), oracle.p(vx),
                                       oracle.p(vy), oracle.p(ids), int(ids[:m].max()) + 1,
                                       mp["e"], mp["nu"], mp["p"], mp["numax"], mp["m2"],
                                       cfg.seed, s, oracle.p(counts))
    return [a[:m].copy() for a in (x, y, vx, vy)], counts


def engine_run(cfg, parts, nsteps, graphs=True, tiling=None):
    with Simulation(cfg) as sim:
        sim.use_graphs(graphs)
        if tiling:
            sim.set_tuning(*tiling)
        sim.set_particles(*parts)
        sim.step(nsteps)
        sim.sync()
        out = [np.asarray(a, np.float32) for a in sim.particles(sync=False)]
        cnt = np.asarray(sim.collision_counts())
        # the diagnostics' collision column holds each step's count
        ncoll = sim.diagnostics(0, nsteps)[:, _lib.PIC_DIAG_NCOLL]
        real = cnt[0] if cfg.collide == _lib.PIC_COLLIDE_DSMC else cnt[:3].sum()
        assert ncoll.sum() == real and (real == 0 or np.count_nonzero(ncoll) > 0)
        return out, cnt


@pytest.mark.parametrize("graphs", [True, False])
def test_engine_dsmc_matches_oracle_steps(oracle, graphs):
    cfg, parts = blob(coeff=0.01)
    nsteps = 11  # with graphs: one captured 8-step sequence, then 3 eager steps
    print(f"engine dsmc vs oracle, graphs={graphs}", flush=True)
    got, gcnt = engine_run(cfg, parts, nsteps, graphs)
    want, wcnt = oracle_run(oracle, cfg, parts, nsteps)
    assert gcnt[0] == wcnt[0] > parts[0].size // 4
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_engine_dsmc_is_tiling_invariant():
    """gmax and the random streams are keyed by the row-major cell and a cell's particles are
    in id order under every tiling, so the collisions do not depend on the tiling."""
    cfg, parts = blob(coeff=0.02)
    outs = [engine_run(cfg, parts, 10, tiling=(t, h, 0, 1)) for t, h in ((8, 8), (16, 8), (32, 16))]
    for out, cnt in outs[1:]:
        assert cnt[0] == outs[0][1][0] > 0
        for a, b in zip(out, outs[0][0]):
            assert np.array_equal(a, b)


def test_engine_dsmc_without_collisions_equals_collisionless_run():
    """coeff = 0: the DSMC path (cell sort every step, no candidates) moves the particles
    exactly as the collisionless path (tile sort fused into the push)."""
    cfg, parts = blob(coeff=0.0)
    a, cnt = engine_run(cfg, parts, 12)
    cfg.collide = _lib.PIC_COLLIDE_NONE
    b, _ = engine_run(cfg, parts, 12)
    assert cnt[0] == 0
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("graphs", [True, False])
def test_engine_mcc_matches_oracle_steps(oracle, graphs):
    cfg, parts = mcc_case()
    nsteps = 6
    print(f"engine mcc vs oracle, graphs={graphs}", flush=True)
    got, gcnt = engine_run(cfg, parts, nsteps, graphs)
    want, wcnt = oracle_run(oracle, cfg, parts, nsteps)
    assert np.array_equal(gcnt, wcnt)
    assert wcnt[2] > 0 and got[0].size == parts[0].size + wcnt[2]
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_engine_mcc_births_through_the_fused_sort():
    """MCC births joining the sort fused into the push (sort every 4: the tail items count
    and scatter the particles appended since the last sort, the count step's own births are
    counted by the births kernel) give the same particles, ids and counts as the separate
    sort every step."""
    cfg, parts = mcc_case()
    nsteps = 21
    out = []
    for se in (4, 1):
        with Simulation(cfg) as sim:
            sim.set_tuning(16, 8, 0, se)
            sim.set_particles(*parts)
            assert sim.sort_fused() == (se > 1)
            sim.step(nsteps)
            sim.sync()
            out.append(([np.asarray(a, np.float32) for a in sim.particles(sync=False)],
                        np.asarray(sim.collision_counts())))
    (a, ca), (b, cb) = out
    assert np.array_equal(ca, cb) and ca[2] > 0
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_batch_with_an_mcc_run_equals_runs_alone():
    """A batch holding an MCC run (births through the fused sort's tail items, counted in
    that run's own rows of the batch's count table) and a collisionless run steps each run
    exactly as it steps alone."""
    from pic_amd import Batch
    mcfg, mparts = mcc_case()
    lcfg = _lib.default_config(nx=32, ny=32, lx=4.0 * math.pi, ly=4.0 * math.pi,
                               precision=_lib.PIC_FP32, seed=11)
    lparts = load_landau(32, 32, lcfg.lx, lcfg.ly, 8, vth=1.0, alpha=0.05, mode=1, seed=11)
    lcfg.capacity = lparts[0].size
    specs = [(mcfg, mparts), (lcfg, lparts)]
    nsteps = 13
    alone = []
    for c, p in specs:
        with Simulation(c) as sim:
            sim.set_particles(*p)
            sim.step(nsteps)
            sim.sync()
            alone.append(([np.asarray(a, np.float32) for a in sim.particles(sync=False)],
                          np.asarray(sim.collision_counts())))
    with Batch([c for c, _ in specs]) as b:
        for r, (_, p) in enumerate(specs):
            b.set_particles(*p, run=r)
        assert b.sort_fused()
        b.step(nsteps)
        b.sync()
        for r in range(len(specs)):
            got = [np.asarray(a, np.float32) for a in b.particles(run=r, sync=False)]
            assert np.array_equal(np.asarray(b.collision_counts(run=r)), alone[r][1])
            for u, v in zip(got, alone[r][0]):
                assert np.array_equal(u, v)
    assert alone[0][1][2] > 0  # the MCC run had births
