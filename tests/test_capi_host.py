"""Host side of the C ABI, on the CPU: the library loads, exports every function the
header declares, validates configurations before touching a device, and its host loaders
reproduce the oracle's loaders bit for bit.  No compute call reaches a GPU here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from pic_amd import _lib
from pic_amd._lib import PicError, check, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pic_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", text))


def test_every_declared_function_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 45
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for n in sorted(names):
        assert hasattr(raw, n), f"{n} declared in pic_capi.h but not exported"
    bound = set(_lib.SIGNATURES) | {"pic_last_error", "pic_version", "pic_config_default",
                                    "pic_mcc_preset", "pic_mcc_nu_max"}
    assert names <= bound, f"unbound: {sorted(names - bound)}"


def test_config_struct_layout_matches_header():
    """The ctypes mirror of pic_config has the header's field order and size."""
    text = open(HEADER).read()
    body = text[text.index("typedef struct pic_config"):]
    body = body[: body.index("} pic_config;")]
    decls = re.findall(r"^\s*(?:int32_t|int64_t|uint64_t|double)\s+([a-z0-9_, ]+);", body,
                       flags=re.M)
    fields = [f.strip() for d in decls for f in d.split(",")]
    assert fields == [f for f, _ in _lib.pic_config._fields_]
    assert ctypes.sizeof(_lib.pic_config) % 8 == 0


def test_version_and_defaults():
    assert lib.pic_version() == 1
    c = _lib.default_config()
    assert c.nx > 0 and c.ny > 0 and c.dt > 0 and c.qm == -1.0
    assert c.precision == _lib.PIC_FP32 and c.collide == _lib.PIC_COLLIDE_NONE


BAD = [
    dict(nx=48), dict(ny=2), dict(nx=4096), dict(lx=0.0), dict(ly=float("inf")),
    dict(dt=-1.0), dict(qm=0.0), dict(precision=7), dict(capacity=-5), dict(collide=9),
    dict(tile=3), dict(tile=64), dict(halo=-1), dict(splits=65), dict(sort_every=-2),
]


@pytest.mark.parametrize("bad", BAD, ids=[next(iter(b)) + "=" + str(next(iter(b.values())))
                                          for b in BAD])
def test_invalid_config_rejected_before_any_device_call(bad):
    cfg = _lib.default_config(**{"nx": 64, "ny": 64, **bad})
    h = ctypes.c_void_p()
    rc = lib.pic_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == _lib.PIC_ERR_INVALID_CONFIG
    assert not h.value
    assert lib.pic_last_error()


def test_collisions_need_fp32():
    cfg = _lib.default_config(nx=64, ny=64, precision=_lib.PIC_FP64,
                              collide=_lib.PIC_COLLIDE_DSMC, dsmc_coeff=1.0, dsmc_gmax0=1.0)
    h = ctypes.c_void_p()
    assert lib.pic_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.PIC_ERR_UNSUPPORTED


def test_argument_errors():
    h = ctypes.c_void_p()
    assert lib.pic_create(None, ctypes.byref(h)) == _lib.PIC_ERR_ARG
    assert lib.pic_step(None, 1) == _lib.PIC_ERR_ARG
    with pytest.raises(PicError) as e:
        check(lib.pic_step(None, 1))
    assert e.value.code == _lib.PIC_ERR_ARG
    cfg = _lib.default_config()
    assert lib.pic_run_compat(None, None, None, None, None, None, 10, 1, ctypes.byref(cfg),
                              None) == _lib.PIC_ERR_ARG


def test_create_without_device_reports_hip_error():
    """With a valid configuration and no visible GPU the context cannot be made: a status,
    never a crash or a silent CPU path."""
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg = _lib.default_config(nx=64, ny=64)
    h = ctypes.c_void_p()
    assert lib.pic_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.PIC_ERR_HIP


@pytest.mark.parametrize("nx,ny,ppc", [(16, 16, 16), (32, 8, 12), (64, 64, 100), (8, 8, 7)])
def test_lattice_count(oracle, nx, ny, ppc):
    from pic_amd import lattice_count
    a, b = oracle.lattice_shape(nx, ny, ppc)
    assert lattice_count(nx, ny, ppc) == a * b == nx * ny * ppc


def same_draws(got, want):
    """Positions bit-exact; velocities to the last bit of the O(1) Box-Muller normal (the two
    builds may pair sin and cos of one angle into a sincos call, whose last bit can differ
    from separate calls)."""
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    for g, w in zip(got[2:], want[2:]):
        assert np.abs(g - w).max() <= 1e-15
        assert np.mean(g == w) > 0.99


def test_landau_loader_matches_oracle(oracle):
    from pic_amd import load_landau
    args = (32, 16, 4 * math.pi, 2 * math.pi, 16)
    got = load_landau(*args, vth=1.0, alpha=0.05, mode=2, seed=99)
    want = oracle.init_landau(*args, vth=1.0, alpha=0.05, mode=2, seed=99)
    same_draws(got, want)


def test_two_stream_loader_matches_oracle(oracle):
    from pic_amd import load_two_stream
    args = (64, 64, 4 * math.pi, 4 * math.pi, 100)
    got = load_two_stream(*args, v0=1.0, vth=0.01, amp=1e-3, mode=1, seed=1234)
    want = oracle.init_two_stream(*args, v0=1.0, vth=0.01, amp=1e-3, mode=1, seed=1234)
    same_draws(got, want)
    vx = got[2]
    assert abs(np.mean(np.abs(vx)) - 1.0) < 1e-3       # two beams at +-1


def test_reference_loader_matches_oracle(oracle):
    from pic_amd import load_reference
    got = load_reference(40, 30, 10, 10, 32, 32, 1.0, 1.0, 0.0, 0.0)
    want = [np.zeros(1200) for _ in range(4)]
    oracle.or_init_reference(40, 30, 10, 10, 32, 32, 1.0, 1.0, 0.0, 0.0,
                             *(oracle.p(w) for w in want))
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    # drift removed: the mean velocity is zero to rounding
    assert abs(got[2].mean()) < 1e-12 and abs(got[3].mean()) < 1e-12


def test_mcc_preset_and_nu_max():
    """nu_max = max_E n_g sigma_T(E) v(E) (0-1000 eV), times the step, with a 1e-4 margin;
    the maximum of sigma_T v for the synthetic cross sections is 9.377e-14 m^3/s."""
    cfg = _lib.default_config(nx=256, ny=256)
    check(lib.pic_mcc_preset(ctypes.byref(cfg), 1e21, 2.0, 1e15))
    assert cfg.mcc_ng == 1e21 and cfg.mcc_v_unit > 0 and cfg.mcc_dt_phys > 0
    e = np.linspace(1e-3, 1000, 2_000_001)
    s = 1e-19 / (1 + e / 10) + np.where(e > 11.5, 2e-21 * (1 - 11.5 / e), 0) + \
        np.where(e > 15.8, 3e-20 * np.log(np.maximum(e, 15.8) / 15.8) / e, 0)
    sv = s * np.sqrt(2 * e * 1.602176634e-19 / 9.1093837015e-31)
    assert abs(sv.max() - 9.377e-14) < 1e-3 * 9.377e-14
    nu_dt = lib.pic_mcc_nu_max(ctypes.byref(cfg))
    # normalised dt times seconds per unit time; the preset picks dt for nu_max dt = 0.01
    assert nu_dt == pytest.approx(1.0001 * 1e21 * sv.max() * cfg.mcc_dt_phys * cfg.dt,
                                  rel=1e-6)
    assert nu_dt == pytest.approx(0.01, rel=1e-9)


def test_checkpoint_info_rejects_missing_and_foreign_files(tmp_path):
    from pic_amd import checkpoint_info
    with pytest.raises(PicError) as e:
        checkpoint_info(tmp_path / "missing.ckpt")
    assert e.value.code == _lib.PIC_ERR_ARG
    junk = tmp_path / "junk.ckpt"
    junk.write_bytes(b"not a checkpoint" * 40)
    with pytest.raises(PicError) as e:
        checkpoint_info(junk)
    assert e.value.code == _lib.PIC_ERR_ARG
