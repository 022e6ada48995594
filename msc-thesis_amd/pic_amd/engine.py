"""Host-side mirror of the reference's PIC interface over the C ABI.

``Simulation`` is the context API (one run); ``Batch`` steps many runs together.  Arrays
cross the boundary as host float64 numpy arrays (``set_particles`` / ``particles``) or as
device pointers (``*_device``).  Reference interface mirrored:
thesisCode/collisionless/double/src/main.cc:31-194 (driver) and
thesisCode/collisionless/double/src/pic.cu:429-611 (picFloatGpu).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Simulation:
    """One periodic 2D electrostatic PIC run on one GPU."""

    def __init__(self, config=None, **kw):
        cfg = config if config is not None else _lib.default_config()
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.config = cfg
        self._ctx = ctypes.c_void_p()
        check(lib.pic_create(ctypes.byref(cfg), ctypes.byref(self._ctx)))
        self.nruns = 1

    # ------------------------------------------------------------------ lifecycle
    def close(self):
        if self._ctx:
            lib.pic_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ particles
    def set_particles(self, x, y, vx, vy, run=0):
        x, y, vx, vy = (_f64(a) for a in (x, y, vx, vy))
        n = x.size
        if not (y.size == vx.size == vy.size == n):
            raise ValueError("particle arrays differ in length")
        check(lib.pic_set_particles(self._ctx, run, _ptr(x), _ptr(y), _ptr(vx), _ptr(vy), n))

    def set_particles_device(self, x, y, vx, vy, n, run=0):
        check(lib.pic_set_particles_device(self._ctx, run, x, y, vx, vy, n))

    def num_particles(self, run=0):
        n = ctypes.c_int64()
        check(lib.pic_num_particles(self._ctx, run, ctypes.byref(n)))
        return n.value

    def particles(self, run=0, sync=True):
        """(x, y, vx, vy) in id order; synchronous velocities when ``sync``."""
        n = self.num_particles(run)
        out = [np.empty(n) for _ in range(4)]
        check(lib.pic_get_particles(self._ctx, run, *(_ptr(a) for a in out), n, int(sync)))
        return tuple(out)

    def particles_device(self, x, y, vx, vy, n, run=0, sync=True):
        check(lib.pic_get_particles_device(self._ctx, run, x, y, vx, vy, n, int(sync)))

    # ------------------------------------------------------------------ stepping
    def step(self, nsteps=1):
        check(lib.pic_step(self._ctx, int(nsteps)))

    def sync(self):
        check(lib.pic_sync(self._ctx))

    def check_finite(self):
        check(lib.pic_check_finite(self._ctx))

    def use_graphs(self, enable):
        check(lib.pic_use_graphs(self._ctx, int(enable)))

    def current_step(self, run=0):
        s = ctypes.c_int64()
        check(lib.pic_current_step(self._ctx, run, ctypes.byref(s)))
        return s.value

    @property
    def stream(self):
        s = ctypes.c_void_p()
        check(lib.pic_stream(self._ctx, ctypes.byref(s)))
        return s.value

    # ------------------------------------------------------------------ grids, diagnostics
    def grid(self, which, run=0):
        cfg = self._run_cfg(run)
        out = np.empty(cfg.nx * cfg.ny)
        sel = {"rho": _lib.PIC_GRID_RHO, "phi": _lib.PIC_GRID_PHI,
               "ex": _lib.PIC_GRID_EX, "ey": _lib.PIC_GRID_EY}[which]
        check(lib.pic_get_grid(self._ctx, run, sel, _ptr(out)))
        return out.reshape(cfg.nx, cfg.ny)

    def diagnostics(self, first=0, count=None, run=0):
        """Rows [first, first+count) of the per-step diagnostics (PIC_DIAG_* columns)."""
        if count is None:
            count = self.current_step(run) - first
        out = np.zeros((max(count, 0), _lib.PIC_DIAG_COLS))
        if count > 0:
            check(lib.pic_get_diagnostics(self._ctx, run, first, count,
                                          out.ctypes.data_as(_dp)))
        return out

    CSV_HEADER = ("step,time,field_energy,kinetic_energy,total_energy,px,py,mode1,particles,"
                  "collisions")

    def write_diagnostics_csv(self, path, first=0, count=None, run=0):
        """Write diagnostics rows as CSV (the format of bin/pic_main -f): step, time, field
        energy, kinetic energy, their sum, momentum, |rho_hat(1,0)|/(nx ny), particle count,
        collisions of the step.  Values are printed with 17 significant digits."""
        d = self.diagnostics(first, count, run)
        dt = self._run_cfg(run).dt
        with open(path, "w") as f:
            f.write(self.CSV_HEADER + "\n")
            for i, r in enumerate(d):
                s = first + i
                f.write(f"{s},{s * dt:.9g},{r[0]:.17g},{r[1]:.17g},{r[0] + r[1]:.17g},"
                        f"{r[2]:.17g},{r[3]:.17g},{r[4]:.17g},{r[5]:.0f},{r[6]:.0f}\n")

    # ------------------------------------------------------------------ checkpoint / restart
    def save_checkpoint(self, path, run=0):
        """Write a restartable snapshot of one run (see pic_save_checkpoint)."""
        check(lib.pic_save_checkpoint(self._ctx, run, os.fsencode(path)))

    def load_checkpoint(self, path, run=0):
        """Replace a run's particles and state by a snapshot of a matching run."""
        check(lib.pic_load_checkpoint(self._ctx, run, os.fsencode(path)))

    @classmethod
    def from_checkpoint(cls, path):
        """A new single-run context restarted from a snapshot."""
        cfg, n, _ = checkpoint_info(path)
        cfg.capacity = max(cfg.capacity, n)
        sim = cls(cfg)
        sim.load_checkpoint(path)
        return sim

    def collision_counts(self, run=0):
        out = (ctypes.c_int64 * 4)()
        check(lib.pic_collision_counts(self._ctx, run, out))
        return list(out)

    # ------------------------------------------------------------------ tuning / timing
    def tuning(self):
        v = [ctypes.c_int() for _ in range(4)]
        check(lib.pic_get_tuning(self._ctx, *(ctypes.byref(a) for a in v)))
        return dict(zip(("tile", "halo", "splits", "sort_every"), (a.value for a in v)))

    def sort_fused(self):
        """True when the tile sort runs inside the push launches (count on the step before
        a sort step, scatter while pushing on the sort step)."""
        v = ctypes.c_int()
        check(lib.pic_sort_fused(self._ctx, ctypes.byref(v)))
        return bool(v.value)

    def set_tuning(self, tile=0, halo=0, splits=0, sort_every=0):
        check(lib.pic_set_tuning(self._ctx, tile, halo, splits, sort_every))

    def autotune(self, trial_steps=20):
        best = ctypes.c_double()
        check(lib.pic_autotune(self._ctx, trial_steps, ctypes.byref(best)))
        return best.value

    def stage_times(self, nsteps=10):
        out = np.zeros(len(_lib.STAGE_NAMES))
        check(lib.pic_stage_times(self._ctx, nsteps, _ptr(out)))
        return dict(zip(_lib.STAGE_NAMES, out))

    def _run_cfg(self, run):
        return self.config


class Batch(Simulation):
    """Ragged ensemble of independent runs stepped with one launch per stage."""

    def __init__(self, configs):
        arr = (_lib.pic_config * len(configs))(*configs)
        self.configs = list(configs)
        self.config = configs[0]
        self._ctx = ctypes.c_void_p()
        check(lib.pic_batch_create(arr, len(configs), ctypes.byref(self._ctx)))
        self.nruns = len(configs)

    def _run_cfg(self, run):
        return self.configs[run]


def run_compat(x, y, vx, vy, nsteps, config):
    """Whole run in one call (semantics of the reference's picFloatGpu); returns
    (x_out, y_out, gpu_seconds)."""
    x, y, vx, vy = (_f64(a) for a in (x, y, vx, vy))
    n = x.size
    xo, yo = np.empty(n), np.empty(n)
    el = ctypes.c_double()
    check(lib.pic_run_compat(_ptr(x), _ptr(y), _ptr(vx), _ptr(vy), _ptr(xo), _ptr(yo), n,
                             int(nsteps), ctypes.byref(config), ctypes.byref(el)))
    return xo, yo, el.value


def checkpoint_info(path):
    """(pic_config, particle count, completed steps) stored in a checkpoint file."""
    cfg = _lib.pic_config()
    n, step = ctypes.c_int64(), ctypes.c_int64()
    check(lib.pic_checkpoint_info(os.fsencode(path), ctypes.byref(cfg), ctypes.byref(n),
                                  ctypes.byref(step)))
    return cfg, n.value, step.value


# --------------------------------------------------------------------------- loaders
def lattice_count(nx, ny, ppc):
    return int(lib.pic_lattice_count(nx, ny, ppc))


def load_landau(nx, ny, lx, ly, ppc, vth=1.0, alpha=0.01, mode=1, seed=1234):
    n = lattice_count(nx, ny, ppc)
    a = [np.empty(n) for _ in range(4)]
    check(lib.pic_load_landau(nx, ny, lx, ly, ppc, vth, alpha, mode, seed, *(_ptr(v) for v in a)))
    return tuple(a)


def load_two_stream(nx, ny, lx, ly, ppc, v0=1.0, vth=0.01, amp=1e-3, mode=1, seed=1234):
    n = lattice_count(nx, ny, ppc)
    a = [np.empty(n) for _ in range(4)]
    check(lib.pic_load_two_stream(nx, ny, lx, ly, ppc, v0, vth, amp, mode, seed,
                                  *(_ptr(v) for v in a)))
    return tuple(a)


def load_reference(npx, npy, xdim, ydim, nx, ny, vtx=1.0, vty=1.0, vdx=0.0, vdy=0.0):
    n = npx * npy
    a = [np.empty(n) for _ in range(4)]
    check(lib.pic_load_reference(npx, npy, xdim, ydim, nx, ny, vtx, vty, vdx, vdy,
                                 *(_ptr(v) for v in a)))
    return tuple(a)
