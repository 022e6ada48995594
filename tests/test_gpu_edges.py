"""Edge cases of the whole step against the FP32 oracle mirror: one particle, particles on
the domain edges, every particle in one cell, non-square grids and domains, the largest
grid, and the inputs the engine must refuse."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import pic_amd  # noqa: E402
from pic_amd import Batch, Simulation, _lib  # noqa: E402

FOUR_PI = 4.0 * math.pi


def run_engine(cfg, parts, nsteps, tuning=None):
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        if tuning:
            sim.set_tuning(**tuning)
        sim.step(nsteps)
        sim.sync()
        out = sim.particles(sync=True)
        diag = sim.diagnostics(0, nsteps)
    return out, diag


def run_mirror(oracle, cfg, parts, nsteps):
    ref = [np.asarray(a, np.float64).astype(np.float32) for a in parts]
    n = ref[0].size
    qp = cfg.qp if cfg.qp != 0.0 else -(cfg.lx * cfg.ly) / n
    diag = oracle.run_f32(*ref, cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.dt, cfg.qm, qp, nsteps)
    return ref, diag


def config(nx, ny, lx, ly, n, dt=0.05):
    return _lib.default_config(nx=nx, ny=ny, lx=lx, ly=ly, dt=dt, precision=_lib.PIC_FP32,
                               seed=11, capacity=n)


def assert_equal_to_mirror(got, ref):
    """The FP32 path reproduces the oracle's FP32 mirror bit for bit (positions, velocities):
    same fixed-point charge, FP64 field within rounding of the oracle's (its FP32 values
    then round alike), same operation order in the push."""
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


def test_one_particle(oracle):
    cfg = config(32, 32, FOUR_PI, FOUR_PI, 1)
    cfg.qp = -0.05   # a light macro-particle (the default would carry the whole box's charge)
    parts = (np.array([1.3]), np.array([7.9]), np.array([0.7]), np.array([-1.1]))
    got, diag = run_engine(cfg, parts, 25)
    ref, diag_r = run_mirror(oracle, cfg, parts, 25)
    assert_equal_to_mirror(got, ref)
    np.testing.assert_allclose(diag[:, 1], diag_r[:, 1], rtol=1e-5)


def test_particles_on_the_domain_edges(oracle):
    """Positions at 0, at L - ulp and one cell from the edges, moving across the periodic
    boundary in both directions."""
    L = FOUR_PI
    below = float(np.nextafter(np.float32(L), np.float32(0)))
    edge = [0.0, below, L / 64, L - L / 64, 0.5 * L]
    xs, ys = np.meshgrid(edge, edge, indexing="ij")
    x = np.tile(xs.ravel(), 40)
    y = np.tile(ys.ravel(), 40)
    rng = np.random.default_rng(3)
    vx = rng.choice([-3.0, 3.0], x.size) + rng.normal(0, 0.1, x.size)
    vy = rng.choice([-3.0, 3.0], x.size) + rng.normal(0, 0.1, x.size)
    cfg = config(64, 64, L, L, x.size)
    got, _ = run_engine(cfg, (x, y, vx, vy), 12)
    ref, _ = run_mirror(oracle, cfg, (x, y, vx, vy), 12)
    assert_equal_to_mirror(got, ref)
    assert got[0].min() >= 0.0 and got[0].max() < L
    assert got[1].min() >= 0.0 and got[1].max() < L


@pytest.mark.parametrize("tuning", [None, dict(tile=8, halo=2, sort_every=2),
                                    dict(tile=16, halo=4, sort_every=1, splits=8)])
def test_every_particle_in_one_cell(oracle, tuning):
    """One tile holds the whole run (the other workgroups are empty), the charge lands on
    four nodes; the cluster then expands across tiles and out of the windows."""
    rng = np.random.default_rng(4)
    n = 40_000
    dx = FOUR_PI / 64
    x = 20 * dx + rng.uniform(0, dx, n)
    y = 33 * dx + rng.uniform(0, dx, n)
    vx = rng.normal(0, 1.0, n)
    vy = rng.normal(0, 1.0, n)
    cfg = config(64, 64, FOUR_PI, FOUR_PI, n)
    got, diag = run_engine(cfg, (x, y, vx, vy), 10, tuning)
    ref, diag_r = run_mirror(oracle, cfg, (x, y, vx, vy), 10)
    assert_equal_to_mirror(got, ref)
    np.testing.assert_allclose(diag[:, 0], diag_r[:, 0], rtol=1e-4)


def test_non_square_grid_and_domain(oracle):
    nx, ny, lx, ly = 32, 128, 2.0 * math.pi, 8.0 * math.pi
    parts = pic_amd.load_landau(nx, ny, lx, ly, 8, vth=1.0, alpha=0.05, mode=1, seed=21)
    cfg = config(nx, ny, lx, ly, parts[0].size)
    got, diag = run_engine(cfg, parts, 20)
    ref, diag_r = run_mirror(oracle, cfg, parts, 20)
    assert_equal_to_mirror(got, ref)
    np.testing.assert_allclose(diag[:, 0], diag_r[:, 0], rtol=1e-5)
    np.testing.assert_allclose(diag[:, 1], diag_r[:, 1], rtol=1e-5)


def test_largest_grid(oracle):
    """2048 x 2048 (the FFT's largest size): initial charge bit-exact and field within
    FP32 rounding of the oracle."""
    n_side = 2048
    rng = np.random.default_rng(5)
    n = 200_000
    parts = (rng.uniform(0, FOUR_PI, n), rng.uniform(0, FOUR_PI, n), np.zeros(n), np.zeros(n))
    cfg = config(n_side, n_side, FOUR_PI, FOUR_PI, n)
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        rho = sim.grid("rho")
        ex = sim.grid("ex")
        sim.step(2)
        sim.sync()
    x32, y32 = parts[0].astype(np.float32), parts[1].astype(np.float32)
    qp = -(cfg.lx * cfg.ly) / n
    acc = oracle.deposit_fixed(x32, y32, n_side, n_side, cfg.lx, cfg.ly)
    rho_r = oracle.fixed_to_rho(acc, n_side, n_side, cfg.lx, cfg.ly, qp) - qp * n / (cfg.lx * cfg.ly)
    assert np.array_equal(rho, rho_r)
    phi_r = oracle.poisson(rho_r, cfg.lx, cfg.ly)
    ex_r, _ = oracle.efield(phi_r, cfg.lx, cfg.ly)
    assert np.abs(ex - ex_r).max() <= 1e-6 * np.abs(ex_r).max()


def test_refused_inputs():
    cfg = config(32, 32, FOUR_PI, FOUR_PI, 16)
    with Simulation(cfg) as sim:
        with pytest.raises(pic_amd.PicError) as e:
            sim.set_particles(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
        assert e.value.code == _lib.PIC_ERR_INVALID_CONFIG
        with pytest.raises(pic_amd.PicError) as e:
            sim.set_tuning(tile=12)
        assert e.value.code == _lib.PIC_ERR_ARG
        with pytest.raises(pic_amd.PicError):
            sim.set_particles(*(np.zeros(17),) * 4)


def test_batch_with_a_one_particle_run(oracle):
    cfgs = [config(32, 32, FOUR_PI, FOUR_PI, 1), config(64, 64, FOUR_PI, FOUR_PI, 64 * 64 * 4)]
    one = (np.array([3.0]), np.array([4.0]), np.array([0.5]), np.array([0.25]))
    many = pic_amd.load_landau(64, 64, FOUR_PI, FOUR_PI, 4, vth=1.0, alpha=0.05, mode=1,
                               seed=8)
    sep = [run_engine(cfgs[0], one, 15)[0], run_engine(cfgs[1], many, 15)[0]]
    with Batch(cfgs) as b:
        b.set_particles(*one, run=0)
        b.set_particles(*many, run=1)
        b.step(15)
        b.sync()
        got = [b.particles(run=0), b.particles(run=1)]
    for s, g in zip(sep, got):
        for u, v in zip(s, g):
            assert np.array_equal(u, v)


def test_fast_particles_in_the_extrapolated_sort_keys(oracle):
    """The tile sorts key on x + lead * v (lead up to (sort_every + 2) / 2 steps): with
    particles crossing the domain several times per step, the extrapolated positions wrap
    many times over.  The order may change, the trajectories may not."""
    rng = np.random.default_rng(6)
    n = 20_000
    L = FOUR_PI
    x = rng.uniform(0, L, n)
    y = rng.uniform(0, L, n)
    vx = rng.normal(0, 1.0, n)
    vy = rng.normal(0, 1.0, n)
    vx[::7] *= 800.0   # ~40 box lengths per step for a seventh of the particles
    vy[::5] *= 800.0
    cfg = config(32, 32, L, L, n)
    base, _ = run_engine(cfg, (x, y, vx, vy), 9, dict(tile=8, halo=2, sort_every=1))
    for tuning in (dict(tile=8, halo=2, sort_every=4), dict(tile=16, halo=4, sort_every=8)):
        got, _ = run_engine(cfg, (x, y, vx, vy), 9, tuning)
        for u, v in zip(base, got):
            assert np.array_equal(u, v)
    ref, _ = run_mirror(oracle, cfg, (x, y, vx, vy), 9)
    assert_equal_to_mirror(base, ref)


def test_fp64_trajectories_do_not_depend_on_the_sort():
    """FP64 particles: the fused sort with extrapolated keys (sort every 4 and 8) and the
    separate sort every step give bit-identical trajectories."""
    parts = pic_amd.load_landau(64, 64, FOUR_PI, FOUR_PI, 8, vth=1.0, alpha=0.05, mode=1,
                                seed=12)
    cfg = _lib.default_config(nx=64, ny=64, lx=FOUR_PI, ly=FOUR_PI, dt=0.05,
                              precision=_lib.PIC_FP64, seed=12, capacity=parts[0].size)
    base, _ = run_engine(cfg, parts, 17, dict(tile=16, halo=8, sort_every=1))
    for tuning in (dict(tile=16, halo=8, sort_every=4), dict(tile=32, halo=8, sort_every=8)):
        got, _ = run_engine(cfg, parts, 17, tuning)
        for u, v in zip(base, got):
            assert np.array_equal(u, v)
