"""Pins of the oracle's reverse-mode restatement (CPU): transposes by dot-product tests,
the whole gradient by central finite differences of the oracle's own FP64 run."""
import math

import numpy as np
import pytest

FOUR_PI = 4.0 * math.pi


def lattice(nx, ny, lx, ly, ppc):
    sy = max(d for d in range(1, int(math.isqrt(ppc)) + 1) if ppc % d == 0)
    sx = ppc // sy
    gx = (np.arange(nx)[:, None] + ((np.arange(sx) + 0.5) / sx)[None, :]).reshape(-1) * lx / nx
    gy = (np.arange(ny)[:, None] + ((np.arange(sy) + 0.5) / sy)[None, :]).reshape(-1) * ly / ny
    return np.repeat(gx, gy.size), np.tile(gy, gx.size)


def modes16():
    return np.array([(a, b) for b in range(4) for a in range(1, 5)], dtype=np.int32)


@pytest.mark.parametrize("nx,ny,lx,ly", [(16, 16, FOUR_PI, FOUR_PI), (32, 8, 3.0, 7.5)])
def test_field_map_transpose(oracle, nx, ny, lx, ly):
    """<M rho, l> = <rho, M^T l> for the map rho -> (Ex, Ey) of or_poisson + or_efield."""
    rng = np.random.default_rng(nx * ny)
    for _ in range(3):
        rho = rng.normal(size=(nx, ny))
        lex, ley = rng.normal(size=(nx, ny)), rng.normal(size=(nx, ny))
        ex, ey = oracle.efield(oracle.poisson(rho, lx, ly), lx, ly)
        lhs = float(np.sum(ex * lex) + np.sum(ey * ley))
        rhs = float(np.sum(rho * oracle.adjoint_mt(lex, ley, lx, ly)))
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.abs(ex * lex).sum())


def test_field_map_transpose_single_mode(oracle):
    """M^T of a single Fourier mode: l = (cos(kx x), 0) gives lrho = -sin(kx dx)/dx / k^2 *
    sin(kx x) (conjugate symbol i sin(k dx)/dx / k^2 of E's -i sin(k dx)/dx / k^2)."""
    nx = ny = 16
    lx = ly = FOUR_PI
    m = 3
    k = 2 * math.pi * m / lx
    dx = lx / nx
    x = (np.arange(nx) * dx)[:, None] * np.ones((1, ny))
    lrho = oracle.adjoint_mt(np.cos(k * x), np.zeros((nx, ny)), lx, ly)
    expect = -math.sin(k * dx) / dx / k ** 2 * np.sin(k * x)
    assert np.allclose(lrho, expect, rtol=0, atol=1e-13)


def test_deposit_is_transpose_of_gather(oracle):
    """sum_nodes rho phi = (qp/dA) sum_p phi(x_p): CIC deposit and gather share weights."""
    rng = np.random.default_rng(7)
    nx, ny, lx, ly = 16, 8, 5.0, 3.0
    n = 500
    x, y = rng.uniform(0, lx, n), rng.uniform(0, ly, n)
    qp = 0.37
    rho = oracle.deposit_f64(x, y, nx, ny, lx, ly, qp)
    phi = rng.normal(size=(nx, ny))
    ep = np.empty(n)
    ep2 = np.empty(n)
    oracle.or_gather_f64(n, oracle.p(x), oracle.p(y), oracle.p(np.ascontiguousarray(phi)),
                         oracle.p(np.ascontiguousarray(phi)), nx, ny, lx, ly, oracle.p(ep),
                         oracle.p(ep2))
    lhs = float(np.sum(rho * phi))
    rhs = qp / (lx / nx * ly / ny) * float(ep.sum())
    assert abs(lhs - rhs) <= 1e-12 * np.abs(rho * phi).sum()


def test_adjoint_loss_equals_field_energy_of_run(oracle):
    """The adjoint's loss is sum(Ex^2+Ey^2) dx dy of or_run_f64's final field (twice its
    field-energy diagnostic, which carries the factor 1/2)."""
    nx = ny = 16
    lx = ly = FOUR_PI
    xl, yl = lattice(nx, ny, lx, ly, 4)
    rng = np.random.default_rng(2)
    n = xl.size
    vx, vy = rng.normal(0, 1, n), rng.normal(0, 1, n)
    modes = modes16()[:3]
    amps = np.array([3e-3, 1e-3, 2e-3])
    qp = -(lx * ly) / n
    nsteps = 6
    L = oracle.adjoint_f64(xl, yl, vx, vy, modes, amps, nx, ny, lx, ly, 0.1, -1.0, qp, nsteps,
                           grad=False)[0]
    x = xl + sum(a * np.cos(2 * math.pi * (mx * xl / lx + my * yl / ly))
                 for a, (mx, my) in zip(amps, modes))
    x = np.mod(x, lx)
    rho = oracle.deposit_f64(x, yl.copy(), nx, ny, lx, ly, qp)
    # step 0 field energy (nsteps = 0 is the initial field)
    ex, ey = oracle.efield(oracle.poisson(rho, lx, ly), lx, ly)
    L0 = oracle.adjoint_f64(xl, yl, vx, vy, modes, amps, nx, ny, lx, ly, 0.1, -1.0, qp, 0,
                            grad=False)[0]
    assert abs(L0 - float(np.sum(ex ** 2 + ey ** 2)) * (lx / nx) * (ly / ny)) <= 1e-14 * L0
    diag = oracle.run_f64(x.copy(), yl.copy(), vx.copy(), vy.copy(), nx, ny, lx, ly, 0.1, -1.0,
                          qp, nsteps + 1)
    assert abs(L - 2.0 * diag[nsteps, 0]) <= 1e-12 * L


@pytest.mark.parametrize("nsteps", [1, 5, 20])
def test_adjoint_gradient_vs_central_differences(oracle, nsteps):
    nx = ny = 16
    lx = ly = FOUR_PI
    xl, yl = lattice(nx, ny, lx, ly, 8)
    rng = np.random.default_rng(1)
    n = xl.size
    vx, vy = rng.normal(0, 1, n), rng.normal(0, 1, n)
    modes = modes16()
    amps = rng.uniform(0, 1e-2, 16)
    qp = -(lx * ly) / n
    args = (nx, ny, lx, ly, 0.1, -1.0, qp, nsteps)
    L, ga, gx, gy = oracle.adjoint_f64(xl, yl, vx, vy, modes, amps, *args)
    eps = 1e-5

    def loss(a=amps, u=vx, w=vy):
        return oracle.adjoint_f64(xl, yl, u, w, modes, a, *args, grad=False)[0]

    for m in (0, 5, 15):
        ap, am = amps.copy(), amps.copy()
        ap[m] += eps
        am[m] -= eps
        fd = (loss(a=ap) - loss(a=am)) / (2 * eps)
        assert abs(fd - ga[m]) <= 1e-7 * np.abs(ga).max(), (m, fd, ga[m])
    for i in (0, 333, 2047):
        up, um = vx.copy(), vx.copy()
        up[i] += eps
        um[i] -= eps
        fd = (loss(u=up) - loss(u=um)) / (2 * eps)
        assert abs(fd - gx[i]) <= 1e-6 * np.abs(gx).max(), (i, fd, gx[i])
        up, um = vy.copy(), vy.copy()
        up[i] += eps
        um[i] -= eps
        fd = (loss(w=up) - loss(w=um)) / (2 * eps)
        assert abs(fd - gy[i]) <= 1e-6 * np.abs(gy).max(), (i, fd, gy[i])


def test_planted_sign_error_is_caught(oracle):
    """A dropped central-difference transpose sign (M^T = -P D instead of P D) flips the
    gradient; the finite-difference check above would see it.  Here the planted variant is
    computed from the pinned forward operators and compared by the dot-product identity."""
    nx = ny = 8
    lx = ly = FOUR_PI
    rng = np.random.default_rng(9)
    rho = rng.normal(size=(nx, ny))
    lex, ley = rng.normal(size=(nx, ny)), rng.normal(size=(nx, ny))
    ex, ey = oracle.efield(oracle.poisson(rho, lx, ly), lx, ly)
    lhs = float(np.sum(ex * lex) + np.sum(ey * ley))
    wrong = -oracle.adjoint_mt(lex, ley, lx, ly)
    assert abs(lhs - float(np.sum(rho * wrong))) > 1e-3 * abs(lhs)
