"""GPU adjoint of the collisionless path: parity with the CPU reverse-mode oracle (itself
pinned against central finite differences in test_oracle_adjoint.py), checkpoint
invariance, the torch.autograd wrapper, and central finite differences of the GPU forward
on the differentiable Landau case (128^2, 32 ppc, 16 modes, 100 steps, checkpoint 10).

Finite differences and CIC: the deposit and gather weights are piecewise linear, so the
loss is piecewise smooth in the parameters.  A central difference of width 2 eps straddles
a cell crossing of some particle with probability ~ eps |dx/da| / dx per crossing; the
difference then picks up the jump of the slope, which makes the FD error O(eps) instead of
O(eps^2).  The tests check that error against the eps it must shrink with, and check the
1e-5 bound on a horizon short enough that no particle crosses a cell edge inside the
difference window.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from pic_amd import _lib  # noqa: E402
from pic_amd.adjoint import (Adjoint, default_modes, field_energy_loss,  # noqa: E402
                             landau_adjoint_case, quiet_lattice)
from pic_amd._lib import PicError  # noqa: E402

FOUR_PI = 4.0 * math.pi

# bars of the config-5 parity test, set from measured values (MI355X):
#   FP64: loss 3.8e-12, amplitude and velocity gradients 1.4e-10 / 2.3e-10 of the largest
#         reference value (the fixed-point CIC weights, 2^-30 of a cell, against the
#         oracle's continuous weights);
#   FP32: loss 1.9e-6, amplitudes 2.1e-3; velocity gradients 2.6e-3 norm-relative (single
#         particles' gradients follow their own FP32 trajectories, which drift from the
#         FP64 ones over 100 thermal steps: up to 0.26 of the largest value).
CONFIG5_TOL = {_lib.PIC_FP64: (1e-10, 1e-9, 1e-9), _lib.PIC_FP32: (5e-6, 5e-3, 1e-2)}


def small_case(seed=5, nx=32, ppc=8):
    xl, yl = quiet_lattice(nx, nx, FOUR_PI, FOUR_PI, ppc)
    rng = np.random.default_rng(seed)
    n = xl.size
    vx, vy = rng.normal(0, 1, n), rng.normal(0, 1, n)
    amps = rng.uniform(0, 1e-2, 16)
    return xl, yl, vx, vy, amps


def make_cfg(nx, precision, dt=0.1):
    cfg = _lib.default_config()
    cfg.nx = cfg.ny = nx
    cfg.lx = cfg.ly = FOUR_PI
    cfg.dt = dt
    cfg.precision = precision
    return cfg


def to_dev(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device="cuda")


def run(adj, amps, vx, vy):
    A = to_dev(amps, torch.float64)
    loss = adj.forward(A, to_dev(vx, adj.dtype), to_dev(vy, adj.dtype))
    ga, gx, gy = adj.backward(1.0)
    return loss, ga.cpu().numpy(), gx.double().cpu().numpy(), gy.double().cpu().numpy()


def rel_to_max(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision,tol_loss,tol_amp,tol_v", [
    # measured 9e-11 (loss), 2.0e-10 (amplitudes), 2.1e-10 (velocities)
    (_lib.PIC_FP64, 5e-10, 1e-9, 1e-9),
    # FP32 trajectories over 20 steps: rounding of positions and velocities amplified by the
    # thermal dynamics; measured 1.5e-7 (loss), 2.4e-6 (amplitudes), 1.6e-3 (velocities)
    (_lib.PIC_FP32, 5e-7, 1e-5, 5e-3),
])
def test_adjoint_matches_oracle(oracle, precision, tol_loss, tol_amp, tol_v):
    xl, yl, vx, vy, amps = small_case()
    nx, nsteps = 32, 20
    qp = -(FOUR_PI * FOUR_PI) / xl.size
    modes = default_modes(16)
    L, ga, gx, gy = oracle.adjoint_f64(xl, yl, vx, vy, modes, amps, nx, nx, FOUR_PI, FOUR_PI,
                                       0.1, -1.0, qp, nsteps)
    adj = Adjoint(make_cfg(nx, precision), xl, yl, modes, nsteps, 5)
    Lg, g1, g2, g3 = run(adj, amps, vx, vy)
    adj.close()
    print(f"32^2 adjoint vs oracle, precision {precision}: loss {abs(Lg - L) / L:.2e}, "
          f"amplitudes {rel_to_max(g1, ga):.2e}, vx {rel_to_max(g2, gx):.2e}, "
          f"vy {rel_to_max(g3, gy):.2e}", flush=True)
    assert abs(Lg - L) / L < tol_loss
    assert rel_to_max(g1, ga) < tol_amp
    assert rel_to_max(g2, gx) < tol_v
    assert rel_to_max(g3, gy) < tol_v


@pytest.fixture(scope="module")
def config5_reference(oracle):
    """The oracle's reverse mode (continuous-weight FP64, whole trajectory kept) on the
    config-5 case itself: 128^2, 32 ppc (524,288 particles), 16 modes, 100 steps."""
    adj, A, VX, VY = landau_adjoint_case(precision=_lib.PIC_FP64)
    adj.close()
    xl, yl = quiet_lattice(128, 128, FOUR_PI, FOUR_PI, 32)
    amps, vx, vy = (t.cpu().numpy() for t in (A, VX, VY))
    qp = -(FOUR_PI * FOUR_PI) / xl.size
    L, ga, gx, gy = oracle.adjoint_f64(xl, yl, vx, vy, default_modes(16), amps, 128, 128,
                                       FOUR_PI, FOUR_PI, 0.1, -1.0, qp, 100)
    return L, ga, gx, gy


@pytest.mark.parametrize("precision,tol_loss,tol_amp,tol_v", [
    # fixed-point CIC weights (2^-30 of a cell) against the oracle's continuous weights
    (_lib.PIC_FP64, CONFIG5_TOL[_lib.PIC_FP64][0], CONFIG5_TOL[_lib.PIC_FP64][1],
     CONFIG5_TOL[_lib.PIC_FP64][2]),
    (_lib.PIC_FP32, CONFIG5_TOL[_lib.PIC_FP32][0], CONFIG5_TOL[_lib.PIC_FP32][1],
     CONFIG5_TOL[_lib.PIC_FP32][2]),
])
def test_config5_adjoint_matches_oracle(config5_reference, precision, tol_loss, tol_amp,
                                        tol_v):
    """Exact reverse-mode parity at config 5's own size (128^2, 32 ppc, 16 modes, 100 steps,
    checkpoints every 10 steps): loss, all 16 amplitude gradients and all 2 x 524,288
    velocity gradients against the oracle's reverse mode."""
    L, ga, gx, gy = config5_reference
    adj, A, VX, VY = landau_adjoint_case(precision=precision)
    loss = adj.forward(A, VX, VY)
    g1, g2, g3 = adj.backward(1.0)
    g1 = g1.cpu().numpy()
    g2, g3 = g2.double().cpu().numpy(), g3.double().cpu().numpy()
    adj.close()
    e = (abs(loss - L) / L, rel_to_max(g1, ga), rel_to_max(g2, gx), rel_to_max(g3, gy))
    nrm = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"config-5 adjoint vs oracle, precision {precision}: loss {e[0]:.2e}, amplitudes "
          f"{e[1]:.2e}, vx {e[2]:.2e}, vy {e[3]:.2e} (norm-relative: amplitudes "
          f"{nrm(g1, ga):.2e}, vx {nrm(g2, gx):.2e}, vy {nrm(g3, gy):.2e})", flush=True)
    assert e[0] < tol_loss
    assert e[1] < tol_amp
    if precision == _lib.PIC_FP64:  # every gradient, relative to the largest
        assert e[2] < tol_v and e[3] < tol_v
    else:                           # the velocity gradient vectors, norm-relative
        assert nrm(g2, gx) < tol_v and nrm(g3, gy) < tol_v


def test_checkpoint_interval_does_not_change_gradient():
    xl, yl, vx, vy, amps = small_case(seed=11)
    modes = default_modes(16)
    res = []
    for K in (1, 7, 20, 50):
        adj = Adjoint(make_cfg(32, _lib.PIC_FP64), xl, yl, modes, 20, K)
        res.append(run(adj, amps, vx, vy))
        adj.close()
    for r in res[1:]:
        assert r[0] == res[0][0], "forward must not depend on the checkpoint interval"
        # the reverse deposit uses FP64 atomics: only the summation order may differ
        for a, b in zip(r[1:], res[0][1:]):
            assert rel_to_max(a, b) < 1e-12


def test_autograd_function_matches_direct_calls():
    xl, yl, vx, vy, amps = small_case(seed=3)
    adj = Adjoint(make_cfg(32, _lib.PIC_FP64), xl, yl, default_modes(16), 10, 4)
    L0, ga, gx, gy = run(adj, amps, vx, vy)
    A = to_dev(amps, torch.float64).requires_grad_(True)
    VX = to_dev(vx, torch.float64).requires_grad_(True)
    VY = to_dev(vy, torch.float64).requires_grad_(True)
    loss = field_energy_loss(adj, A, VX, VY)
    assert loss.dtype == torch.float64 and loss.dim() == 0
    assert float(loss.detach()) == L0
    (3.0 * loss).backward()
    # the reverse deposit uses FP64 atomics: only the summation order differs between calls
    assert rel_to_max(A.grad.cpu().numpy(), 3.0 * ga) < 1e-12
    assert rel_to_max(VX.grad.cpu().numpy(), 3.0 * gx) < 1e-12
    assert rel_to_max(VY.grad.cpu().numpy(), 3.0 * gy) < 1e-12
    adj.close()


def test_repeated_calls_replay_the_same_graphs():
    """forward/backward run as captured graphs: replays with other inputs in between give
    the same loss and gradients as the first call."""
    xl, yl, vx, vy, amps = small_case(seed=17)
    adj = Adjoint(make_cfg(32, _lib.PIC_FP64), xl, yl, default_modes(16), 12, 5)
    first = run(adj, amps, vx, vy)
    other = run(adj, amps * 0.5, vx[::-1].copy(), vy)
    assert other[0] != first[0]
    again = run(adj, amps, vx, vy)
    assert again[0] == first[0]
    for a, b in zip(again[1:], first[1:]):
        assert rel_to_max(a, b) < 1e-12
    adj.close()


def test_adjoint_errors():
    xl, yl, vx, vy, amps = small_case()
    cfg = make_cfg(32, _lib.PIC_FP64)
    cfg.collide = _lib.PIC_COLLIDE_MCC
    with pytest.raises(PicError):
        Adjoint(cfg, xl, yl, default_modes(16), 10, 5)
    adj = Adjoint(make_cfg(32, _lib.PIC_FP64), xl, yl, default_modes(16), 10, 5)
    with pytest.raises(PicError):
        adj.backward(1.0)
    with pytest.raises(ValueError):
        adj.forward(to_dev(amps, torch.float64), to_dev(vx, torch.float32),
                    to_dev(vy, torch.float32))
    adj.close()


def fd_amps(adj64, A, VX, VY, eps):
    fd = np.zeros(A.numel())
    for m in range(A.numel()):
        a = A.clone()
        a[m] += eps
        lp = adj64.forward(a, VX, VY)
        a = A.clone()
        a[m] -= eps
        lm = adj64.forward(a, VX, VY)
        fd[m] = (lp - lm) / (2 * eps)
    return fd


def fd_velocity(adj64, A, VX, VY, i, eps):
    v = VX.clone()
    v[i] += eps
    lp = adj64.forward(A, v, VY)
    v = VX.clone()
    v[i] -= eps
    lm = adj64.forward(A, v, VY)
    return (lp - lm) / (2 * eps)


@pytest.fixture(scope="module")
def landau_fd():
    """FP64 central differences of the 100-step Landau loss at three widths."""
    adj64, A, VX, VY = landau_adjoint_case(precision=_lib.PIC_FP64)
    out = {eps: fd_amps(adj64, A, VX, VY, eps) for eps in (1e-4, 1e-5, 1e-6)}
    idx = [0, 12345, 400000]
    out["v"] = {i: fd_velocity(adj64, A, VX, VY, i, 1e-5) for i in idx}
    adj64.close()
    return out


@pytest.mark.parametrize("precision", [_lib.PIC_FP64, _lib.PIC_FP32])
def test_landau_gradient_vs_central_differences(landau_fd, precision):
    adj, A, VX, VY = landau_adjoint_case(precision=precision)
    adj.forward(A, VX, VY)
    ga, gx, gy = adj.backward(1.0)
    ga = ga.cpu().numpy()
    gx = gx.double().cpu().numpy()
    adj.close()
    errs = {eps: float(np.linalg.norm(ga - landau_fd[eps]) / np.linalg.norm(landau_fd[eps]))
            for eps in (1e-4, 1e-5, 1e-6)}
    print("landau adjoint vs FD, precision", precision, errs)
    # O(eps) cell-crossing error of the difference quotient; measured 1.4e-2, 2-3e-3, 7e-4-2e-3
    assert errs[1e-4] < 3e-2
    assert errs[1e-6] < 4e-3
    assert errs[1e-6] < errs[1e-4] / 4
    for i, fd in landau_fd["v"].items():
        assert abs(gx[i] - fd) <= 1e-2 * abs(fd), (i, gx[i], fd)


# measured 8.9e-6 (FP64) and 9.1e-6 (FP32)
@pytest.mark.parametrize("precision,tol", [(_lib.PIC_FP64, 1e-5), (_lib.PIC_FP32, 1e-5)])
def test_short_horizon_gradient_within_1e5_of_central_differences(precision, tol):
    """Config-5 grid and particles, cold start (no cell crossings inside the FD window):
    FP64 central differences with eps = 1e-4."""
    adj64, A, VX, VY = landau_adjoint_case(precision=_lib.PIC_FP64, vth=0.0, nsteps=5,
                                           checkpoint_every=2)
    fd = fd_amps(adj64, A, VX, VY, 1e-4)
    adj64.close()
    adj, A2, VX2, VY2 = landau_adjoint_case(precision=precision, vth=0.0, nsteps=5,
                                            checkpoint_every=2)
    adj.forward(A2, VX2, VY2)
    ga = adj.backward(1.0)[0].cpu().numpy()
    adj.close()
    err = float(np.linalg.norm(ga - fd) / np.linalg.norm(fd))
    print("short horizon adjoint vs FD, precision", precision, err)
    assert err < tol
