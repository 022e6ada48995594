"""pic_op_deposit_noalloc (static device scratch, no allocation) against pic_op_deposit.

The two entry points share their kernels; the allocation-free one keeps the fixed-point
accumulators in a static device array that every call on the device reuses, so the
second call of each pair runs on a smaller grid with fewer particles after a larger one
and must not see the first call's sums.  Both results are also checked bit for bit
against the oracle's fixed-point deposit.
"""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from pic_amd import _lib  # noqa: E402
from pic_amd._lib import check, lib  # noqa: E402

DEV = "cuda:0"
FOUR_PI = 4.0 * math.pi


def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def deposit(fn, x, y, prec, nx, ny, lx, ly, qp):
    rho = torch.full((nx * ny,), float("nan"), dtype=torch.float64, device=DEV)
    g = _lib.pic_grid(nx, ny, lx, ly)
    tx, ty = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    check(fn(dptr(tx), dptr(ty), x.size, prec, ctypes.byref(g), qp, dptr(rho), None))
    torch.cuda.synchronize()
    return rho.cpu().numpy()


@pytest.mark.parametrize("dtype,prec", [(np.float32, _lib.PIC_FP32), (np.float64, _lib.PIC_FP64)])
def test_noalloc_deposit_equals_allocating_deposit(oracle, dtype, prec):
    rng = np.random.default_rng(21)
    # a large call, then a smaller one: the second must start from a cleared scratch
    cases = [(1024, 1024, FOUR_PI, FOUR_PI, 3_000_000), (64, 32, FOUR_PI, 3.0, 50_000)]
    for nx, ny, lx, ly, n in cases:
        print(f"noalloc deposit {dtype.__name__} {nx}x{ny} n={n}", flush=True)
        x = rng.uniform(0, lx, n).astype(dtype)
        y = rng.uniform(0, ly, n).astype(dtype)
        qp = -lx * ly / n
        a = deposit(lib.pic_op_deposit, x, y, prec, nx, ny, lx, ly, qp)
        b = deposit(lib.pic_op_deposit_noalloc, x, y, prec, nx, ny, lx, ly, qp)
        assert np.array_equal(a, b)
        if n <= 100_000:
            fixed = oracle.fixed_to_rho(oracle.deposit_fixed(x, y, nx, ny, lx, ly), nx, ny, lx,
                                        ly, qp)
            assert np.array_equal(b.reshape(nx, ny), fixed)
        # total charge is exact in fixed point
        assert abs(b.sum() * (lx / nx) * (ly / ny) - qp * n) <= 1e-12 * abs(qp * n)


def test_noalloc_deposit_empty_and_repeat():
    """n = 0 after a full call gives an all-zero grid; the same call twice gives the same
    bits."""
    rng = np.random.default_rng(5)
    nx = ny = 128
    x = rng.uniform(0, 1.0, 100_000).astype(np.float32)
    y = rng.uniform(0, 1.0, 100_000).astype(np.float32)
    a = deposit(lib.pic_op_deposit_noalloc, x, y, _lib.PIC_FP32, nx, ny, 1.0, 1.0, 1e-5)
    b = deposit(lib.pic_op_deposit_noalloc, x, y, _lib.PIC_FP32, nx, ny, 1.0, 1.0, 1e-5)
    assert np.array_equal(a, b)
    e = deposit(lib.pic_op_deposit_noalloc, x[:0], y[:0], _lib.PIC_FP32, nx, ny, 1.0, 1.0, 1e-5)
    assert np.array_equal(e, np.zeros(nx * ny))


def test_noalloc_deposit_refuses_grids_above_its_scratch():
    """A grid of more than PIC_DEPOSIT_NOALLOC_MAX_NODES nodes is refused before any
    device work."""
    rho = torch.zeros(1, dtype=torch.float64, device=DEV)
    x = torch.zeros(1, dtype=torch.float32, device=DEV)
    g = _lib.pic_grid(2048, 1024, 1.0, 1.0)
    rc = lib.pic_op_deposit_noalloc(dptr(x), dptr(x), 1, _lib.PIC_FP32, ctypes.byref(g), 1.0,
                                    dptr(rho), None)
    assert rc == _lib.PIC_ERR_UNSUPPORTED
    assert b"exceeds" in lib.pic_last_error()
