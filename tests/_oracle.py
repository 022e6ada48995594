"""ctypes binding of oracle/liboracle.so (the CPU checker; test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int32)
_LP = ctypes.POINTER(ctypes.c_int64)
_U32P = ctypes.POINTER(ctypes.c_uint32)

SIG = {
    "or_philox4x32_10": (None, [_U32P, _U32P, _U32P]),
    "or_philox2x32_10": (None, [_U32P, _U32P, _U32P]),
    "or_minstd_rand0_nth": (ctypes.c_uint64, [ctypes.c_uint32, _I]),
    "or_mt19937_nth": (ctypes.c_uint64, [ctypes.c_uint32, _I]),
    "or_init_reference": (None, [_I, _I, _I, _I, _I, _I, _D, _D, _D, _D, _DP, _DP, _DP, _DP]),
    "or_init_two_stream": (None, [_I, _I, _D, _D, _I, _D, _D, _D, _I, ctypes.c_uint64,
                                  _DP, _DP, _DP, _DP]),
    "or_init_landau": (None, [_I, _I, _D, _D, _I, _D, _D, _I, ctypes.c_uint64,
                              _DP, _DP, _DP, _DP]),
    "or_lattice_shape": (None, [_I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "or_cell_keys_f64": (None, [_I64, _DP, _DP, _I, _I, _D, _D, _IP]),
    "or_cell_keys_f32": (None, [_I64, _FP, _FP, _I, _I, _D, _D, _IP]),
    "or_stable_sort": (None, [_I64, _IP, _I, _IP, _IP, _IP]),
    "or_deposit_f64": (None, [_I64, _DP, _DP, _I, _I, _D, _D, _D, _DP]),
    "or_deposit_fixed_f32": (None, [_I64, _FP, _FP, _I, _I, _D, _D, _LP]),
    "or_deposit_fixed_f64": (None, [_I64, _DP, _DP, _I, _I, _D, _D, _LP, _LP]),
    "or_fixed_to_rho_f32": (None, [_I, _I, _D, _D, _D, _LP, _DP]),
    "or_fixed_to_rho_f64": (None, [_I, _I, _D, _D, _D, _LP, _LP, _DP]),
    "or_fft2": (None, [_I, _I, _DP, _I]),
    "or_dft2_brute": (None, [_I, _I, _DP, _DP, _I]),
    "or_poisson": (None, [_I, _I, _D, _D, _DP, _DP]),
    "or_efield": (None, [_I, _I, _D, _D, _DP, _DP, _DP]),
    "or_push_f64": (None, [_I64, _DP, _DP, _DP, _DP, _DP, _DP, _I, _I, _D, _D, _D, _D, _D, _I,
                           _DP, _DP, _DP]),
    "or_push_f32": (None, [_I64, _FP, _FP, _FP, _FP, _FP, _FP, _I, _I, _D, _D, _D, _D, _D,
                           _DP, _DP, _DP]),
    "or_gather_f64": (None, [_I64, _DP, _DP, _DP, _DP, _I, _I, _D, _D, _DP, _DP]),
    "or_run_f64": (None, [_I64, _DP, _DP, _DP, _DP, _I, _I, _D, _D, _D, _D, _D, _I, _DP]),
    "or_run_f32": (None, [_I64, _FP, _FP, _FP, _FP, _I, _I, _D, _D, _D, _D, _D, _I, _DP]),
    "or_run_f32_threads": (None, [_I64, _FP, _FP, _FP, _FP, _I, _I, _D, _D, _D, _D, _D, _I,
                                  _I, _DP]),
    "or_mcc_cross_sections": (None, [_D, _DP, _DP, _DP]),
    "or_mcc_step_f32": (_I64, [_I64, _I64, _FP, _FP, _FP, _FP, _IP, _I64, _D, _D, _D, _D, _D,
                               ctypes.c_uint64, _I64, _LP]),
    "or_dsmc_step_f32": (_I64, [_I, _IP, _FP, _FP, _FP, _D, ctypes.c_uint64, _I64]),
    "or_dsmc_pair_perm": (None, [ctypes.c_uint32, _I64, _I64, _I64, ctypes.c_uint64,
                                 ctypes.c_void_p]),
    "or_std_uniform_stream": (None, [ctypes.c_uint32, _I64, _DP]),
    "or_mcc_reference": (_I64, [_I64, _DP, _DP, _D, _DP, _I64]),
    "or_dsmc_reference": (_I64, [_I64, _DP, _DP, _DP, _DP, _I, _I, _D, _D, _D, _DP, _I64, _LP]),
    "or_sincos2pi": (None, [_D, _DP, _DP]),
    "or_log": (_D, [_D]),
    "or_adjoint_mt": (None, [_I, _I, _D, _D, _DP, _DP, _DP]),
    "or_adjoint_f64": (_D, [_I64, _DP, _DP, _DP, _DP, _I, _IP, _IP, _DP, _I, _I, _D, _D, _D,
                            _D, _D, _I, _DP, _DP, _DP]),
}


class Oracle:
    """Thin numpy-friendly wrapper; every method follows the oracle.h contract."""

    def __init__(self, lib):
        self.lib = lib

    @staticmethod
    def p(a):
        if a is None:
            return None
        t = {np.float64: _DP, np.float32: _FP, np.int32: _IP, np.int64: _LP,
             np.uint32: _U32P}[a.dtype.type]
        return a.ctypes.data_as(t)

    def __getattr__(self, name):
        return getattr(self.lib, name)

    # --- convenience wrappers
    def lattice_shape(self, nx, ny, ppc):
        a, b = _I(), _I()
        self.lib.or_lattice_shape(nx, ny, ppc, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def init_landau(self, nx, ny, lx, ly, ppc, vth=1.0, alpha=0.01, mode=1, seed=1234):
        a, b = self.lattice_shape(nx, ny, ppc)
        out = [np.zeros(a * b) for _ in range(4)]
        self.lib.or_init_landau(nx, ny, lx, ly, ppc, vth, alpha, mode, seed,
                                *(self.p(v) for v in out))
        return out

    def init_two_stream(self, nx, ny, lx, ly, ppc, v0=1.0, vth=0.01, amp=1e-3, mode=1,
                        seed=1234):
        a, b = self.lattice_shape(nx, ny, ppc)
        out = [np.zeros(a * b) for _ in range(4)]
        self.lib.or_init_two_stream(nx, ny, lx, ly, ppc, v0, vth, amp, mode, seed,
                                    *(self.p(v) for v in out))
        return out

    def cell_keys(self, x, y, nx, ny, lx, ly):
        k = np.zeros(x.size, np.int32)
        if x.dtype == np.float32:
            self.lib.or_cell_keys_f32(x.size, self.p(x), self.p(y), nx, ny, lx, ly, self.p(k))
        else:
            self.lib.or_cell_keys_f64(x.size, self.p(x), self.p(y), nx, ny, lx, ly, self.p(k))
        return k

    def stable_sort(self, keys, ncells):
        perm = np.zeros(keys.size, np.int32)
        cnt = np.zeros(ncells, np.int32)
        off = np.zeros(ncells + 1, np.int32)
        self.lib.or_stable_sort(keys.size, self.p(keys), ncells, self.p(perm), self.p(cnt),
                                self.p(off))
        return perm, cnt, off

    def deposit_f64(self, x, y, nx, ny, lx, ly, qp):
        rho = np.zeros(nx * ny)
        self.lib.or_deposit_f64(x.size, self.p(x), self.p(y), nx, ny, lx, ly, qp, self.p(rho))
        return rho.reshape(nx, ny)

    def deposit_fixed(self, x, y, nx, ny, lx, ly):
        if x.dtype == np.float32:
            acc = np.zeros(nx * ny, np.int64)
            self.lib.or_deposit_fixed_f32(x.size, self.p(x), self.p(y), nx, ny, lx, ly,
                                          self.p(acc))
            return acc.reshape(nx, ny)
        hi = np.zeros(nx * ny, np.int64)
        lo = np.zeros(nx * ny, np.int64)
        self.lib.or_deposit_fixed_f64(x.size, self.p(x), self.p(y), nx, ny, lx, ly, self.p(hi),
                                      self.p(lo))
        return hi.reshape(nx, ny), lo.reshape(nx, ny)

    def fixed_to_rho(self, acc, nx, ny, lx, ly, qp):
        rho = np.zeros(nx * ny)
        if isinstance(acc, tuple):
            self.lib.or_fixed_to_rho_f64(nx, ny, lx, ly, qp, self.p(np.ascontiguousarray(acc[0])),
                                         self.p(np.ascontiguousarray(acc[1])), self.p(rho))
        else:
            self.lib.or_fixed_to_rho_f32(nx, ny, lx, ly, qp, self.p(np.ascontiguousarray(acc)),
                                         self.p(rho))
        return rho.reshape(nx, ny)

    def poisson(self, rho, lx, ly):
        nx, ny = rho.shape
        phi = np.zeros(nx * ny)
        self.lib.or_poisson(nx, ny, lx, ly, self.p(np.ascontiguousarray(rho, np.float64)),
                            self.p(phi))
        return phi.reshape(nx, ny)

    def efield(self, phi, lx, ly):
        nx, ny = phi.shape
        ex, ey = np.zeros(nx * ny), np.zeros(nx * ny)
        self.lib.or_efield(nx, ny, lx, ly, self.p(np.ascontiguousarray(phi, np.float64)),
                           self.p(ex), self.p(ey))
        return ex.reshape(nx, ny), ey.reshape(nx, ny)

    def push_f32(self, x, y, vx, vy, ex, ey, lx, ly, dt, qm, c):
        nx, ny = ex.shape
        ke, px, py = _D(), _D(), _D()
        exf = np.ascontiguousarray(ex, np.float32)
        eyf = np.ascontiguousarray(ey, np.float32)
        self.lib.or_push_f32(x.size, self.p(x), self.p(y), self.p(vx), self.p(vy),
                             self.p(exf), self.p(eyf), nx, ny, lx, ly, dt, qm, c,
                             ctypes.byref(ke), ctypes.byref(px), ctypes.byref(py))
        return ke.value

    def push_f64(self, x, y, vx, vy, ex, ey, lx, ly, dt, qm, c, quantized=0):
        nx, ny = ex.shape
        ke, px, py = _D(), _D(), _D()
        self.lib.or_push_f64(x.size, self.p(x), self.p(y), self.p(vx), self.p(vy),
                             self.p(np.ascontiguousarray(ex)), self.p(np.ascontiguousarray(ey)),
                             nx, ny, lx, ly, dt, qm, c, quantized, ctypes.byref(ke),
                             ctypes.byref(px), ctypes.byref(py))
        return ke.value

    def run_f64(self, x, y, vx, vy, nx, ny, lx, ly, dt, qm, qp, nsteps):
        diag = np.zeros((nsteps, 5))
        self.lib.or_run_f64(x.size, self.p(x), self.p(y), self.p(vx), self.p(vy), nx, ny, lx,
                            ly, dt, qm, qp, nsteps, self.p(diag))
        return diag

    def run_f32(self, x, y, vx, vy, nx, ny, lx, ly, dt, qm, qp, nsteps):
        diag = np.zeros((nsteps, 5))
        self.lib.or_run_f32(x.size, self.p(x), self.p(y), self.p(vx), self.p(vy), nx, ny, lx,
                            ly, dt, qm, qp, nsteps, self.p(diag))
        return diag

    def run_f32_threads(self, x, y, vx, vy, nx, ny, lx, ly, dt, qm, qp, nsteps, nthreads):
        diag = np.zeros((nsteps, 5))
        self.lib.or_run_f32_threads(x.size, self.p(x), self.p(y), self.p(vx), self.p(vy), nx,
                                    ny, lx, ly, dt, qm, qp, nsteps, nthreads, self.p(diag))
        return diag

    def adjoint_mt(self, lex, ley, lx, ly):
        nx, ny = lex.shape
        lex = np.ascontiguousarray(lex, dtype=np.float64)
        ley = np.ascontiguousarray(ley, dtype=np.float64)
        out = np.empty((nx, ny))
        self.lib.or_adjoint_mt(nx, ny, lx, ly, self.p(lex), self.p(ley), self.p(out))
        return out

    def adjoint_f64(self, xl, yl, vx0, vy0, modes, amps, nx, ny, lx, ly, dt, qm, qp, nsteps,
                    grad=True):
        """(loss, grad_amps, grad_vx, grad_vy); gradients None when grad is False."""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        xl, yl, vx0, vy0, amps = f(xl), f(yl), f(vx0), f(vy0), f(amps)
        modes = np.asarray(modes, dtype=np.int32).reshape(-1, 2)
        mx = np.ascontiguousarray(modes[:, 0])
        my = np.ascontiguousarray(modes[:, 1])
        n, nm = xl.size, modes.shape[0]
        ga = np.empty(max(nm, 1)) if grad else None
        gx = np.empty(n) if grad else None
        gy = np.empty(n) if grad else None
        loss = self.lib.or_adjoint_f64(n, self.p(xl), self.p(yl), self.p(vx0), self.p(vy0),
                                       nm, self.p(mx), self.p(my), self.p(amps), nx, ny, lx,
                                       ly, dt, qm, qp, nsteps, self.p(ga), self.p(gx),
                                       self.p(gy))
        if not grad:
            return loss, None, None, None
        return loss, ga[:nm], gx, gy


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    if not os.path.exists(ORACLE_SO):
        build()
    lib = ctypes.CDLL(ORACLE_SO)
    for name, (res, args) in SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return Oracle(lib)


REF_MCC_SO = os.path.join(ORACLE_DIR, "_ref", "libref_mcc.so")
REF_DSMC_SO = os.path.join(ORACLE_DIR, "_ref", "libref_dsmc.so")


def load_ref_collide():
    """(mcc, dsmc) ctypes handles of the reference's own collide code, or None if not built."""
    if not (os.path.exists(REF_MCC_SO) and os.path.exists(REF_DSMC_SO)):
        return None
    mcc = ctypes.CDLL(REF_MCC_SO)
    mcc.ref_mcc_collide.argtypes = [_I64, _DP, _DP, ctypes.c_uint32, _I]
    mcc.ref_mcc_collide.restype = None
    dsmc = ctypes.CDLL(REF_DSMC_SO)
    dsmc.ref_dsmc_collide.argtypes = [_I64, _DP, _DP, _DP, _DP, _I, _I, _D, _D,
                                      ctypes.c_uint32, _I]
    dsmc.ref_dsmc_collide.restype = None
    return mcc, dsmc
