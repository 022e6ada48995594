"""GPU operators of the reference's own collide arithmetic (pic_op_mcc_reference,
pic_op_dsmc_reference) against the oracle's restatement and, where oracle/_ref is built,
against the reference's code itself (thesisCode/mccCollisions/src/particle.cc:112-127;
thesisCode/dsmcCollisions/cellTrackerAdvanced/src/main.cc:269-350), on the same injected
std::mt19937 gen{10} stream: velocities bit for bit, uniforms consumed exactly."""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from pic_amd import _lib  # noqa: E402
from pic_amd._lib import check, lib  # noqa: E402
from _oracle import load_ref_collide  # noqa: E402

DEV = "cuda:0"
REF = load_ref_collide()


def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def hp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def stream(oracle, seed, count):
    u = np.zeros(count)
    oracle.or_std_uniform_stream(seed, count, hp(u))
    return u


def gpu_mcc(vx, vy, prob, u):
    tv = [torch.from_numpy(a.copy()).to(DEV) for a in (vx, vy)]
    tu = torch.from_numpy(u).to(DEV)
    used = ctypes.c_int64()
    check(lib.pic_op_mcc_reference(dptr(tv[0]), dptr(tv[1]), vx.size, prob, dptr(tu), u.size,
                                   ctypes.byref(used), None))
    torch.cuda.synchronize()
    return tv[0].cpu().numpy(), tv[1].cpu().numpy(), used.value


def gpu_dsmc(x, y, vx, vy, nx, L, prob, u):
    tx, ty = (torch.from_numpy(a).to(DEV) for a in (x, y))
    tv = [torch.from_numpy(a.copy()).to(DEV) for a in (vx, vy)]
    tu = torch.from_numpy(u).to(DEV)
    used, nc = ctypes.c_int64(), ctypes.c_int64()
    g = _lib.pic_grid(nx, nx, L, L)
    check(lib.pic_op_dsmc_reference(dptr(tx), dptr(ty), dptr(tv[0]), dptr(tv[1]), x.size,
                                    ctypes.byref(g), prob, dptr(tu), u.size, ctypes.byref(used),
                                    ctypes.byref(nc), None))
    torch.cuda.synchronize()
    return tv[0].cpu().numpy(), tv[1].cpu().numpy(), used.value, nc.value


@pytest.mark.parametrize("prob", [0.8, 0.22, 1.0, 0.0])
def test_mcc_reference_matches_oracle(oracle, prob):
    """Three consecutive passes continuing the stream (the reference's one generator for
    the whole run): velocities and uniforms consumed equal the oracle's."""
    rng = np.random.default_rng(1)
    n = 300_000
    vx, vy = rng.normal(size=n), rng.normal(size=n)
    u = stream(oracle, 10, 7 * n * 3 + 5)
    ovx, ovy = vx.copy(), vy.copy()
    k = 0
    for step in range(3):
        print(f"mcc reference op, prob {prob}, pass {step}", flush=True)
        vx, vy, used = gpu_mcc(vx, vy, prob, u[k:])
        want = oracle.or_mcc_reference(n, hp(ovx), hp(ovy), prob, hp(u[k:]), u.size - k)
        assert used == want > 0
        assert np.array_equal(vx, ovx) and np.array_equal(vy, ovy)
        k += used


@pytest.mark.skipif(REF is None, reason="oracle/_ref collide libraries not built")
def test_mcc_reference_matches_the_reference_code(oracle):
    rng = np.random.default_rng(2)
    n, nsteps = 50_000, 3
    vx, vy = rng.normal(size=n), rng.normal(size=n)
    rvx, rvy = vx.copy(), vy.copy()
    REF[0].ref_mcc_collide(n, hp(rvx), hp(rvy), 10, nsteps)
    u = stream(oracle, 10, 7 * n * nsteps + 7)
    k = 0
    for _ in range(nsteps):
        vx, vy, used = gpu_mcc(vx, vy, 0.8, u[k:])
        k += used
    assert np.array_equal(vx, rvx) and np.array_equal(vy, rvy)


def dsmc_case(rng, n, nx):
    L = float(nx)
    return L, rng.uniform(0, L, n), rng.uniform(0, L, n), rng.normal(size=n), rng.normal(size=n)


@pytest.mark.parametrize("n,nx", [(200_000, 64), (3_000, 1 * 4), (1, 8), (40_000, 16)])
def test_dsmc_reference_matches_oracle(oracle, n, nx):
    rng = np.random.default_rng(n)
    L, x, y, vx, vy = dsmc_case(rng, n, nx)
    u = stream(oracle, 10, 7 * n * 2 + 7)
    ovx, ovy = vx.copy(), vy.copy()
    k = 0
    for step in range(2):
        print(f"dsmc reference op, n={n}, {nx}^2, pass {step}", flush=True)
        vx, vy, used, nc = gpu_dsmc(x, y, vx, vy, nx, L, 0.2, u[k:])
        onc = ctypes.c_int64()
        want = oracle.or_dsmc_reference(n, hp(x), hp(y), hp(ovx), hp(ovy), nx, nx, L, L, 0.2,
                                        hp(u[k:]), u.size - k, ctypes.byref(onc))
        assert used == want and nc == onc.value
        assert np.array_equal(vx, ovx) and np.array_equal(vy, ovy)
        k += used
    if n > 1:
        assert k > 0


@pytest.mark.skipif(REF is None, reason="oracle/_ref collide libraries not built")
def test_dsmc_reference_matches_the_reference_code(oracle):
    rng = np.random.default_rng(7)
    n, nx, nsteps = 30_000, 32, 3
    L, x, y, vx, vy = dsmc_case(rng, n, nx)
    rvx, rvy = vx.copy(), vy.copy()
    REF[1].ref_dsmc_collide(n, hp(x), hp(y), hp(rvx), hp(rvy), nx, nx, 1.0, 1.0, 10, nsteps)
    u = stream(oracle, 10, 7 * n * nsteps + 7)
    k = 0
    for _ in range(nsteps):
        vx, vy, used, _ = gpu_dsmc(x, y, vx, vy, nx, L, 0.2, u[k:])
        k += used
    assert np.array_equal(vx, rvx) and np.array_equal(vy, rvy)


def test_reference_ops_refuse_bad_input(oracle):
    rng = np.random.default_rng(3)
    n = 1000
    vx, vy = rng.normal(size=n), rng.normal(size=n)
    u = stream(oracle, 10, 900)  # fewer than one uniform per particle
    with pytest.raises(_lib.PicError) as e:
        gpu_mcc(vx, vy, 0.5, u)
    assert e.value.code == _lib.PIC_ERR_ARG
    L, x, y, vx, vy = dsmc_case(rng, n, 8)
    x[5] = L  # outside [0, L)
    with pytest.raises(_lib.PicError) as e:
        gpu_dsmc(x, y, vx, vy, 8, L, 0.2, stream(oracle, 10, 7 * n))
    assert e.value.code == _lib.PIC_ERR_ARG
