"""GPU parity of the per-stage operators (C ABI pic_op_*) against the CPU oracle.

Integer outputs (cell keys, sort permutation, counts, offsets, fixed-point charge) must be
bit-exact; FP64 field stages agree to rounding (tolerances stated per test).
"""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import pic_amd  # noqa: E402
from pic_amd import _lib  # noqa: E402
from pic_amd._lib import check, lib  # noqa: E402

DEV = "cuda:0"
FOUR_PI = 4.0 * math.pi


def dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def grid(nx, ny, lx, ly):
    return _lib.pic_grid(nx, ny, lx, ly)


def edge_positions(rng, n, L, dtype):
    x = rng.uniform(0, L, n).astype(dtype)
    # exact cell edges, 0, and the largest value below L
    nx_edges = np.arange(0, 64, dtype=np.float64) * (L / 64)
    x[: nx_edges.size] = nx_edges.astype(dtype)
    x[-1] = np.nextafter(dtype(L), dtype(0))
    x[-2] = dtype(0)
    return x


@pytest.mark.parametrize("dtype,prec", [(np.float32, _lib.PIC_FP32), (np.float64, _lib.PIC_FP64)])
@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 32), (512, 512)])
def test_cell_keys_bit_exact(oracle, dtype, prec, nx, ny):
    rng = np.random.default_rng(1)
    n = 100_003
    x = edge_positions(rng, n, FOUR_PI, dtype)
    y = edge_positions(rng, n, 2.0, dtype)[::-1].copy()
    ref = oracle.cell_keys(x, y, nx, ny, FOUR_PI, 2.0)
    tx, ty = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
    keys = torch.empty(n, dtype=torch.int32, device=DEV)
    g = grid(nx, ny, FOUR_PI, 2.0)
    check(lib.pic_op_cell_keys(dptr(tx), dptr(ty), n, prec, ctypes.byref(g), dptr(keys), None))
    torch.cuda.synchronize()
    assert np.array_equal(keys.cpu().numpy(), ref)


@pytest.mark.parametrize("case", ["uniform", "clustered", "huge_cell", "empty"])
def test_sort_bit_exact(oracle, case):
    rng = np.random.default_rng(7)
    ncells = 4096
    if case == "uniform":
        keys = rng.integers(0, ncells, 300_000).astype(np.int32)
    elif case == "clustered":
        # cells of 1..300 particles: exercises the register and workgroup segment sorts
        keys = np.repeat(np.arange(0, ncells, 7), rng.integers(1, 300, len(range(0, ncells, 7))))
        keys = rng.permutation(keys).astype(np.int32)
    elif case == "huge_cell":
        keys = np.concatenate([np.full(20_000, 5), rng.integers(0, ncells, 1000)])
        keys = rng.permutation(keys).astype(np.int32)
    else:
        keys = np.zeros(0, np.int32)
    n = keys.size
    perm_r, cnt_r, off_r = oracle.stable_sort(keys, ncells)
    tk = torch.from_numpy(keys).to(DEV) if n else torch.zeros(1, dtype=torch.int32, device=DEV)
    perm = torch.full((max(n, 1),), -1, dtype=torch.int32, device=DEV)
    cnt = torch.empty(ncells, dtype=torch.int32, device=DEV)
    off = torch.empty(ncells + 1, dtype=torch.int32, device=DEV)
    check(lib.pic_op_sort(dptr(tk), n, ncells, dptr(perm), dptr(cnt), dptr(off), None))
    torch.cuda.synchronize()
    assert np.array_equal(cnt.cpu().numpy(), cnt_r)
    assert np.array_equal(off.cpu().numpy(), off_r)
    if n:
        assert np.array_equal(perm.cpu().numpy()[:n], perm_r)


def test_sort_rejects_bad_key():
    tk = torch.tensor([0, 3, 99], dtype=torch.int32, device=DEV)
    perm = torch.empty(3, dtype=torch.int32, device=DEV)
    cnt = torch.empty(10, dtype=torch.int32, device=DEV)
    off = torch.empty(11, dtype=torch.int32, device=DEV)
    rc = lib.pic_op_sort(dptr(tk), 3, 10, dptr(perm), dptr(cnt), dptr(off), None)
    assert rc == _lib.PIC_ERR_ARG


@pytest.mark.parametrize("dtype,prec", [(np.float32, _lib.PIC_FP32), (np.float64, _lib.PIC_FP64)])
def test_deposit(oracle, dtype, prec):
    rng = np.random.default_rng(3)
    nx, ny, n = 64, 32, 200_000
    x = rng.uniform(0, FOUR_PI, n).astype(dtype)
    y = rng.uniform(0, 3.0, n).astype(dtype)
    qp = -FOUR_PI * 3.0 / n
    rho = torch.empty(nx * ny, dtype=torch.float64, device=DEV)
    g = grid(nx, ny, FOUR_PI, 3.0)
    tx, ty = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)  # keep alive
    check(lib.pic_op_deposit(dptr(tx), dptr(ty), n, prec, ctypes.byref(g), qp, dptr(rho), None))
    torch.cuda.synchronize()
    got = rho.cpu().numpy().reshape(nx, ny)
    # bit-exact against the fixed-point mirror
    fixed = oracle.fixed_to_rho(oracle.deposit_fixed(x, y, nx, ny, FOUR_PI, 3.0), nx, ny, FOUR_PI,
                                3.0, qp)
    assert np.array_equal(got, fixed)
    # continuous-weight FP64 reference: 1e-5 relative (FP32), 1e-9 (FP64) of max |rho|
    ref = oracle.deposit_f64(x.astype(np.float64), y.astype(np.float64), nx, ny, FOUR_PI, 3.0, qp)
    tol = 1e-5 if dtype == np.float32 else 1e-9
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    # total charge is exact in fixed point
    assert abs(got.sum() * (FOUR_PI / nx) * (3.0 / ny) - qp * n) <= 1e-12 * abs(qp * n)


@pytest.mark.parametrize("nx,ny,lx,ly", [(64, 64, FOUR_PI, FOUR_PI), (256, 64, 10.0, 3.0),
                                         (512, 512, FOUR_PI, FOUR_PI), (8, 2048, 1.0, 50.0)])
def test_poisson_and_efield(oracle, nx, ny, lx, ly):
    rng = np.random.default_rng(11)
    rho = rng.normal(size=(nx, ny))
    phi_r = oracle.poisson(rho, lx, ly)
    ex_r, ey_r = oracle.efield(phi_r, lx, ly)
    trho = torch.from_numpy(rho.ravel()).to(DEV)
    phi = torch.empty(nx * ny, dtype=torch.float64, device=DEV)
    ex = torch.empty_like(phi)
    ey = torch.empty_like(phi)
    g = grid(nx, ny, lx, ly)
    check(lib.pic_op_poisson(dptr(trho), dptr(phi), ctypes.byref(g), None))
    check(lib.pic_op_efield(dptr(phi), dptr(ex), dptr(ey), ctypes.byref(g), None))
    torch.cuda.synchronize()
    got = phi.cpu().numpy().reshape(nx, ny)
    assert np.max(np.abs(got - phi_r)) <= 1e-12 * np.max(np.abs(phi_r))
    assert np.max(np.abs(ex.cpu().numpy().reshape(nx, ny) - ex_r)) <= 1e-12 * np.max(np.abs(ex_r))
    assert np.max(np.abs(ey.cpu().numpy().reshape(nx, ny) - ey_r)) <= 1e-12 * np.max(np.abs(ey_r))


def test_poisson_single_mode_closed_form():
    nx, ny, lx, ly = 128, 64, 10.0, 6.0
    i = np.arange(nx)[:, None] * (lx / nx)
    j = np.arange(ny)[None, :] * (ly / ny)
    kx, ky = 2 * math.pi * 3 / lx, 2 * math.pi * 2 / ly
    rho = np.cos(kx * i + ky * j) * np.ones((nx, ny))
    trho = torch.from_numpy(rho.ravel().copy()).to(DEV)
    phi = torch.empty(nx * ny, dtype=torch.float64, device=DEV)
    check(lib.pic_op_poisson(dptr(trho), dptr(phi), ctypes.byref(grid(nx, ny, lx, ly)), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(phi.cpu().numpy().reshape(nx, ny), rho / (kx * kx + ky * ky),
                               atol=1e-13)


@pytest.mark.parametrize("c", [0.5, 1.0])
def test_gather_push_f32_bit_exact(oracle, c):
    rng = np.random.default_rng(5)
    nx, ny, n = 64, 64, 100_000
    ex = rng.normal(size=(nx, ny))
    ey = rng.normal(size=(nx, ny))
    x = rng.uniform(0, FOUR_PI, n).astype(np.float32)
    y = rng.uniform(0, FOUR_PI, n).astype(np.float32)
    vx = rng.normal(size=n).astype(np.float32) * 3
    vy = rng.normal(size=n).astype(np.float32) * 3
    t = [torch.from_numpy(a.copy()).to(DEV) for a in (x, y, vx, vy)]
    tex, tey = torch.from_numpy(ex.ravel()).to(DEV), torch.from_numpy(ey.ravel()).to(DEV)
    ke = torch.zeros(1, dtype=torch.float64, device=DEV)
    g = grid(nx, ny, FOUR_PI, FOUR_PI)
    check(lib.pic_op_gather_push(*(dptr(a) for a in t), n, _lib.PIC_FP32, dptr(tex), dptr(tey),
                                 ctypes.byref(g), 0.05, -1.0, c, dptr(ke), None))
    torch.cuda.synchronize()
    ke_r = oracle.push_f32(x, y, vx, vy, ex, ey, FOUR_PI, FOUR_PI, 0.05, -1.0, c)
    for a, b in zip(t, (x, y, vx, vy)):
        assert np.array_equal(a.cpu().numpy(), b)
    assert abs(ke.item() - ke_r) <= 1e-12 * ke_r


def test_gather_push_f64_matches_mirror(oracle):
    rng = np.random.default_rng(6)
    nx, ny, n = 32, 64, 50_000
    ex = rng.normal(size=(nx, ny))
    ey = rng.normal(size=(nx, ny))
    arrs = [rng.uniform(0, FOUR_PI, n), rng.uniform(0, 2.0, n), rng.normal(size=n),
            rng.normal(size=n)]
    t = [torch.from_numpy(a.copy()).to(DEV) for a in arrs]
    tex, tey = torch.from_numpy(ex.ravel()).to(DEV), torch.from_numpy(ey.ravel()).to(DEV)
    g = grid(nx, ny, FOUR_PI, 2.0)
    check(lib.pic_op_gather_push(*(dptr(a) for a in t), n, _lib.PIC_FP64, dptr(tex), dptr(tey),
                                 ctypes.byref(g), 0.1, -1.0, 1.0, None, None))
    torch.cuda.synchronize()
    oracle.push_f64(*arrs, ex, ey, FOUR_PI, 2.0, 0.1, -1.0, 1.0, quantized=1)
    for a, b in zip(t, arrs):
        assert np.array_equal(a.cpu().numpy(), b)
