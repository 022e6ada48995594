"""Named workloads (grid, particles, time step) used by tests, bench.py and the CLI.

All lengths are in Debye lengths and times in inverse plasma frequencies.
"""
import math

import numpy as np

from . import _lib
from .engine import lattice_count, load_landau, load_two_stream

FOUR_PI = 4.0 * math.pi


def two_stream_64(precision=_lib.PIC_FP64, seed=1234):
    """Cold two-stream, 64x64, Lx=Ly=4 pi, 100 particles/cell (409,600), beams +-1 with
    thermal spread 0.01, displacement 1e-3 cos(0.5 x), dt = 0.1, 200 steps."""
    nx = ny = 64
    cfg = _lib.default_config(nx=nx, ny=ny, lx=FOUR_PI, ly=FOUR_PI, dt=0.1, qm=-1.0,
                              precision=precision, seed=seed)
    parts = load_two_stream(nx, ny, FOUR_PI, FOUR_PI, 100, v0=1.0, vth=0.01, amp=1e-3, mode=1,
                            seed=seed)
    cfg.capacity = parts[0].size
    return cfg, parts, 200


def landau_512(precision=_lib.PIC_FP32, seed=1234, ppc=64, n=512):
    """Landau damping, 512x512, k lambda_D = 0.5 (L = 4 pi), 64 particles/cell (16.8M),
    Maxwellian vth = 1, density 1 + 0.01 cos(0.5 x), dt = 0.05, 1000 steps."""
    cfg = _lib.default_config(nx=n, ny=n, lx=FOUR_PI, ly=FOUR_PI, dt=0.05, qm=-1.0,
                              precision=precision, seed=seed)
    parts = load_landau(n, n, FOUR_PI, FOUR_PI, ppc, vth=1.0, alpha=0.01, mode=1, seed=seed)
    cfg.capacity = parts[0].size
    return cfg, parts, 1000


def landau_small(n=64, ppc=64, alpha=0.05, precision=_lib.PIC_FP32, seed=1234, dt=0.05):
    cfg = _lib.default_config(nx=n, ny=n, lx=FOUR_PI, ly=FOUR_PI, dt=dt, qm=-1.0,
                              precision=precision, seed=seed)
    parts = load_landau(n, n, FOUR_PI, FOUR_PI, ppc, vth=1.0, alpha=alpha, mode=1, seed=seed)
    cfg.capacity = parts[0].size
    return cfg, parts


def batch_48(seed=1234, nruns=48):
    """Ragged ensemble: grids from {64, 128, 256}^2, particle counts log-uniform in
    [1e4, 2e6], per-run perturbation amplitude U[1e-3, 1e-2], Landau/two-stream alternating,
    FP32 particles, 300 steps."""
    rng = np.random.default_rng(seed)
    runs = []
    for r in range(nruns):
        n = int(rng.choice([64, 128, 256]))
        npart = int(round(math.exp(rng.uniform(math.log(1e4), math.log(2e6)))))
        amp = float(rng.uniform(1e-3, 1e-2))
        ppc = max(1, int(round(npart / (n * n))))
        cfg = _lib.default_config(nx=n, ny=n, lx=FOUR_PI, ly=FOUR_PI, dt=0.05, qm=-1.0,
                                  precision=_lib.PIC_FP32, seed=seed + r)
        if r % 2 == 0:
            parts = load_landau(n, n, FOUR_PI, FOUR_PI, ppc, vth=1.0, alpha=amp, mode=1,
                                seed=seed + r)
        else:
            cfg.dt = 0.1
            parts = load_two_stream(n, n, FOUR_PI, FOUR_PI, ppc, v0=1.0, vth=0.01, amp=amp,
                                    mode=1, seed=seed + r)
        cfg.capacity = parts[0].size
        runs.append((cfg, parts))
    return runs, 300


def mcc_256(seed=1234, ppc=128, n=256):
    """Electron-neutral MCC on 256x256 with 8,388,608 electrons (128/cell), Te = 2 eV,
    neutrals 1e21 m^-3 at 300 K (argon mass), plasma density 1e15 m^-3; the box is
    256 x 256 Debye lengths at dx = lambda_D and dt is set so nu_max dt_phys = 0.01."""
    cfg = _lib.default_config(nx=n, ny=n, lx=float(n), ly=float(n), qm=-1.0,
                              precision=_lib.PIC_FP32, seed=seed)
    _lib.check(_lib.lib.pic_mcc_preset(cfg, 1e21, 2.0, 1e15))
    # velocities in units of sqrt(Te/m): Maxwellian at Te is N(0, 1) per component
    parts = load_landau(n, n, cfg.lx, cfg.ly, ppc, vth=1.0, alpha=0.0, mode=0, seed=seed)
    cfg.capacity = int(parts[0].size * 1.25)
    cfg.sort_every = 8
    return cfg, parts, 500


def dsmc_blob(seed=1234, n=256, mean_ppc=32, coeff=3e-4, charge_scale=1e-6):
    """Hard-sphere NTC DSMC on a Gaussian density blob (sigma = n/6 cells, mean 32
    particles/cell over the box, 183 per cell at the centre) of Maxwellian particles
    (unit thermal speed per component), dt = 0.2.

    coeff = 3e-4 gives nu dt = (n_c - 1) coeff <g> ~ 0.1 collisions per particle per step
    in the central cell (<g> = sqrt(pi) for unit 2D Maxwellians), the bound DSMC's
    splitting of transport and collisions needs.  The macro-particle charge is
    charge_scale x the plasma value -(lx ly)/N: the engine's Poisson stage neutralises
    with a uniform background (the k = 0 mode), not with a blob of ions, so at full charge
    the blob's space charge (field energy ~1300x its kinetic energy; the ratio scales with
    the charge, as the field energy goes with its square and the macro-particle mass with
    the charge itself) blows it apart within a few plasma periods.  At 1e-6 the field
    energy is ~1.3e-3 of the kinetic energy and the blob relaxes by collisions and free
    streaming while every stage of the step (field solve, push, cell sort, collisions)
    still runs at full size."""
    cfg = _lib.default_config(nx=n, ny=n, lx=float(n), ly=float(n), dt=0.2, qm=-1.0,
                              precision=_lib.PIC_FP32, seed=seed, collide=_lib.PIC_COLLIDE_DSMC,
                              dsmc_coeff=coeff, dsmc_gmax0=4.0)
    rng = np.random.default_rng(seed)
    npart = n * n * mean_ppc
    sig = n / 6.0
    x = np.mod(rng.normal(n / 2, sig, npart), n)
    y = np.mod(rng.normal(n / 2, sig, npart), n)
    vx = rng.normal(0, 1, npart)
    vy = rng.normal(0, 1, npart)
    cfg.capacity = npart
    cfg.qp = -charge_scale * cfg.lx * cfg.ly / npart
    return cfg, (x, y, vx, vy), 500


__all__ = ["two_stream_64", "landau_512", "landau_small", "batch_48", "mcc_256", "dsmc_blob",
           "lattice_count"]
