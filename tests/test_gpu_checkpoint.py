"""Checkpoint / restart of a run: a restarted run continues bit for bit (particles, step,
collision counters), mismatched or damaged files are refused."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from pic_amd import Simulation, _lib, checkpoint_info, load_landau  # noqa: E402
from pic_amd._lib import PicError  # noqa: E402

FOUR_PI = 4.0 * math.pi


def landau(precision, n=64, ppc=16):
    parts = load_landau(n, n, FOUR_PI, FOUR_PI, ppc, vth=1.0, alpha=0.05, mode=1, seed=77)
    cfg = _lib.default_config(nx=n, ny=n, lx=FOUR_PI, ly=FOUR_PI, dt=0.05, precision=precision,
                              seed=77, capacity=parts[0].size)
    return cfg, parts


def mcc_small():
    cfg = _lib.default_config(nx=64, ny=64, lx=64.0, ly=64.0, precision=_lib.PIC_FP32, seed=5)
    _lib.check(_lib.lib.pic_mcc_preset(ctypes.byref(cfg), 1e21, 2.0, 1e15))
    cfg.dt *= 10.0      # more collisions per step in a short test (nu_max dt = 0.1)
    parts = load_landau(64, 64, 64.0, 64.0, 16, vth=2.0, alpha=0.0, mode=0, seed=5)
    cfg.capacity = int(parts[0].size * 1.25)
    return cfg, parts


def dsmc_small():
    rng = np.random.default_rng(9)
    n = 64 * 64 * 8
    x = np.mod(rng.normal(32, 10, n), 64.0)
    y = np.mod(rng.normal(32, 10, n), 64.0)
    cfg = _lib.default_config(nx=64, ny=64, lx=64.0, ly=64.0, dt=0.2, precision=_lib.PIC_FP32,
                              seed=3, collide=_lib.PIC_COLLIDE_DSMC, dsmc_coeff=0.05,
                              dsmc_gmax0=4.0, capacity=n)
    return cfg, (x, y, rng.normal(0, 1, n), rng.normal(0, 1, n))


def state(sim):
    n = sim.num_particles()
    return [np.asarray(a).copy() for a in sim.particles(sync=False)], n


@pytest.mark.parametrize("case", ["landau_f32", "landau_f64", "mcc", "dsmc"])
def test_restart_continues_bit_for_bit(tmp_path, case):
    cfg, parts = {"landau_f32": lambda: landau(_lib.PIC_FP32),
                  "landau_f64": lambda: landau(_lib.PIC_FP64),
                  "mcc": mcc_small, "dsmc": dsmc_small}[case]()
    path = tmp_path / "run.ckpt"
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.step(13)
        sim.sync()
        sim.save_checkpoint(path)
        sim.step(11)
        sim.sync()
        want, nwant = state(sim)
        counts = sim.collision_counts()
        assert sim.current_step() == 24
    c2, n_saved, step = checkpoint_info(path)
    assert step == 13 and c2.nx == cfg.nx and c2.precision == cfg.precision
    with Simulation.from_checkpoint(path) as sim2:
        assert sim2.current_step() == 13
        assert sim2.num_particles() == n_saved
        sim2.step(11)
        sim2.sync()
        got, ngot = state(sim2)
        assert sim2.current_step() == 24
        assert sim2.collision_counts() == counts
    assert ngot == nwant
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_dsmc_restart_under_another_tiling(tmp_path):
    """The DSMC per-cell maxima are stored by row-major cell: a run saved under one tiling
    and restored under another continues exactly as the uninterrupted run."""
    cfg, parts = dsmc_small()
    path = tmp_path / "dsmc.ckpt"
    with Simulation(cfg) as sim:
        sim.set_tuning(8, 8, 0, 1)
        sim.set_particles(*parts)
        sim.step(9)
        sim.sync()
        sim.save_checkpoint(path)
        sim.step(9)
        sim.sync()
        want, nwant = state(sim)
        counts = sim.collision_counts()
    with Simulation(cfg) as sim2:
        sim2.set_tuning(32, 16, 0, 1)
        sim2.load_checkpoint(path)
        assert sim2.tuning()["tile"] == 32
        sim2.step(9)
        sim2.sync()
        got, ngot = state(sim2)
        assert sim2.collision_counts() == counts
    assert ngot == nwant and counts[0] > 0
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_checkpoint_larger_than_the_run_is_refused(tmp_path):
    """A checkpoint with more particles than the run's capacity is refused before the
    run is touched."""
    cfg, parts = landau(_lib.PIC_FP32, n=32, ppc=4)
    path = tmp_path / "big.ckpt"
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.step(2)
        sim.save_checkpoint(path)
    small = _lib.default_config(nx=32, ny=32, lx=FOUR_PI, ly=FOUR_PI, dt=0.05,
                                precision=_lib.PIC_FP32, seed=77, capacity=parts[0].size // 2)
    half = [a[: parts[0].size // 2] for a in parts]
    with Simulation(small) as sim:
        sim.set_particles(*half)
        sim.step(1)
        sim.sync()
        before = state(sim)
        with pytest.raises(PicError) as e:
            sim.load_checkpoint(path)
        assert e.value.code == _lib.PIC_ERR_OVERFLOW
        after = state(sim)
        assert sim.current_step() == 1
    for a, b in zip(before[0], after[0]):
        assert np.array_equal(a, b)


def test_mismatched_or_damaged_checkpoint_is_refused(tmp_path):
    cfg, parts = landau(_lib.PIC_FP32, n=32, ppc=4)
    path = tmp_path / "a.ckpt"
    with Simulation(cfg) as sim:
        sim.set_particles(*parts)
        sim.step(3)
        sim.save_checkpoint(path)
    other = _lib.default_config(nx=32, ny=32, lx=FOUR_PI, ly=FOUR_PI, dt=0.1,
                                precision=_lib.PIC_FP32, seed=77, capacity=parts[0].size)
    with Simulation(other) as sim:
        with pytest.raises(PicError) as e:
            sim.load_checkpoint(path)
        assert e.value.code == _lib.PIC_ERR_INVALID_CONFIG
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(bytes(raw))
    with Simulation(cfg) as sim:
        with pytest.raises(PicError) as e:
            sim.load_checkpoint(bad)
        assert e.value.code == _lib.PIC_ERR_ARG
