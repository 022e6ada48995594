"""Differentiable collisionless run: the hand-written adjoint kernels behind torch.autograd.

The forward map is the engine's collisionless step (deposit, FFT Poisson, E, gather, push;
the reference step order of thesisCode/collisionless/double/src/main.cc:137-172) applied
``nsteps`` times to the initial state

    x0 = xl + sum_m a_m cos(kx_m xl + ky_m yl),  y0 = yl,  v0 = (vx0, vy0)

and the loss is the field energy sum(Ex^2 + Ey^2) dx dy after the last step.  The C ABI
(``pic_adjoint_*``) checkpoints the state every ``checkpoint_every`` steps on the forward
pass and recomputes the steps of each segment on the reverse pass.  Collisions have no
useful derivative and are not part of this path.

torch is used for device memory only: amplitudes and velocities are CUDA tensors whose
pointers cross the C ABI; the arithmetic is in libpicamd.so.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


def _dptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), _dp)


def _vptr(t):
    return ctypes.c_void_p(t.data_ptr())


def quiet_lattice(nx, ny, lx, ly, ppc):
    """Unperturbed positions: an sx x sy sub-lattice in every cell (sx sy = ppc, sy the
    largest divisor of ppc not above its square root), x-major order."""
    sy = max(d for d in range(1, int(math.isqrt(ppc)) + 1) if ppc % d == 0)
    sx = ppc // sy
    dx, dy = lx / nx, ly / ny
    gx = (np.arange(nx)[:, None] + ((np.arange(sx) + 0.5) / sx)[None, :]).reshape(-1) * dx
    gy = (np.arange(ny)[:, None] + ((np.arange(sy) + 0.5) / sy)[None, :]).reshape(-1) * dy
    xl = np.repeat(gx, gy.size)
    yl = np.tile(gy, gx.size)
    return xl, yl


class Adjoint:
    """One differentiable run (owns a ``pic_adjoint`` handle)."""

    def __init__(self, config, xl, yl, modes, nsteps, checkpoint_every=10):
        xl = np.ascontiguousarray(xl, dtype=np.float64)
        yl = np.ascontiguousarray(yl, dtype=np.float64)
        if xl.shape != yl.shape or xl.ndim != 1:
            raise ValueError("xl and yl must be 1-D arrays of one length")
        modes = np.asarray(modes, dtype=np.int32).reshape(-1, 2)
        self.config = config
        self.n = xl.size
        self.nmodes = modes.shape[0]
        self.nsteps = nsteps
        self.dtype = torch.float64 if config.precision == _lib.PIC_FP64 else torch.float32
        self.device = torch.device("cuda", config.device)
        mx = np.ascontiguousarray(modes[:, 0])
        my = np.ascontiguousarray(modes[:, 1])
        self._h = ctypes.c_void_p()
        check(lib.pic_adjoint_create(
            ctypes.byref(config), xl.ctypes.data_as(_dp), yl.ctypes.data_as(_dp), self.n,
            self.nmodes, mx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            my.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nsteps, checkpoint_every,
            ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.pic_adjoint_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_inputs(self, amps, vx0, vy0):
        if amps.dtype != torch.float64 or amps.numel() != self.nmodes:
            raise ValueError("amps must be a float64 tensor of nmodes entries")
        for v in (vx0, vy0):
            if v.dtype != self.dtype or v.numel() != self.n:
                raise ValueError("velocities must have n entries of the run precision")
        for t in (amps, vx0, vy0):
            if t.device != self.device or not t.is_contiguous():
                raise ValueError("inputs must be contiguous tensors on the run's device")

    def forward(self, amps, vx0, vy0):
        """Loss (float) of the run from these parameters; records checkpoints."""
        self._check_inputs(amps, vx0, vy0)
        torch.cuda.current_stream(self.device).synchronize()
        loss = ctypes.c_double()
        check(lib.pic_adjoint_forward(self._h, _dptr(amps), _vptr(vx0), _vptr(vy0),
                                      ctypes.byref(loss)))
        return loss.value

    def backward(self, dloss=1.0):
        """Gradients (amps, vx0, vy0) of dloss * loss at the last forward's parameters."""
        ga = torch.zeros(max(self.nmodes, 1), dtype=torch.float64, device=self.device)
        gx = torch.empty(self.n, dtype=self.dtype, device=self.device)
        gy = torch.empty(self.n, dtype=self.dtype, device=self.device)
        torch.cuda.current_stream(self.device).synchronize()
        check(lib.pic_adjoint_backward(self._h, float(dloss), _dptr(ga), _vptr(gx), _vptr(gy)))
        return ga[: self.nmodes], gx, gy


class FieldEnergyLoss(torch.autograd.Function):
    """loss = FieldEnergyLoss.apply(adjoint, amps, vx0, vy0) -- a 0-d float64 tensor."""

    @staticmethod
    def forward(ctx, adj, amps, vx0, vy0):
        ctx.adj = adj
        val = adj.forward(amps.detach().contiguous(), vx0.detach().contiguous(),
                          vy0.detach().contiguous())
        return torch.tensor(val, dtype=torch.float64, device=amps.device)

    @staticmethod
    def backward(ctx, g):
        ga, gx, gy = ctx.adj.backward(float(g))
        return None, ga, gx, gy


def field_energy_loss(adj, amps, vx0, vy0):
    return FieldEnergyLoss.apply(adj, amps, vx0, vy0)


def default_modes(count=16):
    """(mx, my) of the first ``count`` cosine modes, 1 <= mx <= 4, 0 <= my <= 3."""
    m = [(a, b) for b in range(4) for a in range(1, 5)]
    return np.array(m[:count], dtype=np.int32)


def landau_adjoint_case(nx=128, ny=128, ppc=32, nsteps=100, checkpoint_every=10,
                        precision=_lib.PIC_FP64, vth=1.0, dt=0.1, seed=2024, device=0,
                        nmodes=16):
    """Differentiable Landau setup: L = 4 pi, thermal velocities, seeded amplitudes in [0, 1e-2].

    Returns (Adjoint, amps, vx0, vy0) with the parameters as CUDA tensors.
    """
    lx = ly = 4.0 * math.pi
    rng = np.random.default_rng(seed)
    xl, yl = quiet_lattice(nx, ny, lx, ly, ppc)
    cfg = _lib.default_config()
    cfg.nx, cfg.ny, cfg.lx, cfg.ly, cfg.dt = nx, ny, lx, ly, dt
    cfg.precision = precision
    cfg.device = device
    adj = Adjoint(cfg, xl, yl, default_modes(nmodes), nsteps, checkpoint_every)
    dev = torch.device("cuda", device)
    amps = torch.tensor(rng.uniform(0.0, 1e-2, nmodes), dtype=torch.float64, device=dev)
    vx = rng.normal(0.0, vth, xl.size)
    vy = rng.normal(0.0, vth, xl.size)
    vx0 = torch.tensor(vx, dtype=adj.dtype, device=dev)
    vy0 = torch.tensor(vy, dtype=adj.dtype, device=dev)
    return adj, amps, vx0, vy0


def time_adjoint(nx=128, ny=128, ppc=32, nsteps=100, checkpoint_every=10,
                 precision=_lib.PIC_FP32, reps=3):
    """Wall time of the forward pass (loss, checkpoints) and of the backward pass (segment
    recomputation + reverse sweep); returns both and their ratios."""
    import time
    adj, amps, vx0, vy0 = landau_adjoint_case(nx, ny, ppc, nsteps, checkpoint_every, precision)
    adj.forward(amps, vx0, vy0)
    adj.backward(1.0)
    tf, tb = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        adj.forward(amps, vx0, vy0)
        t1 = time.perf_counter()
        adj.backward(1.0)
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    f, b = min(tf), min(tb)
    adj.close()
    return {"forward_s": f, "backward_s": b, "ratio_backward_over_forward": b / f,
            "ratio_total_over_forward": (f + b) / f,
            "n": nx * ny * ppc, "grid": [nx, ny], "steps": nsteps,
            "checkpoint_every": checkpoint_every,
            "precision": "f64" if precision == _lib.PIC_FP64 else "f32"}
