"""C ABI checks that need no GPU: the library loads, exports every entry
point include/periodicflow.h declares, and rejects invalid descriptors
before touching the device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "periodicflow.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pf_\w+)\s*\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    from periodicflow import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpfhip.so not built")
    return _lib.load()


def test_header_and_binding_agree(lib):
    from periodicflow import _lib

    declared = _declared()
    assert len(declared) >= 20
    assert sorted(_lib.EXPORTED) == declared
    for name in declared:
        assert hasattr(lib, name), name


def test_version(lib):
    assert b"gfx950" in lib.pf_version()


def _desc(**kw):
    from periodicflow import _lib

    d = _lib.PfDesc()
    d.dim = 3
    for i in range(3):
        d.n[i] = 16
        d.length[i] = 6.283185307179586
    d.precision = 0
    d.dealias = 0
    d.integrator = 3
    d.nonlinear = 1
    d.noise_type = 0
    d.forcing = 0
    d.batch = 1
    d.viscosity = 0.01
    d.dt = 0.01
    for k, v in kw.items():
        if k == "n":
            for i, x in enumerate(v):
                d.n[i] = x
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw,code,msg", [
    (dict(dim=4), -1, "2D and 3D"),
    (dict(precision=7), -1, "precision"),
    (dict(dealias=5), -1, "dealias"),
    (dict(integrator=9), -1, "integrator"),
    (dict(noise_type=3), -1, "noise type"),
    (dict(batch=0), -1, "batch"),
    (dict(dt=0.0), -1, "dt"),
    (dict(n=(15, 16, 16)), -1, "even"),
    (dict(n=(14, 16, 16)), -4, "transform length"),   # factor 7
    (dict(n=(16, 16, 44)), -4, "transform length"),   # factor 11
    (dict(n=(16, 16, 30)), -4, "transform length"),   # 3/2 rule: padded length 45 is odd
])
def test_plan_create_rejects(lib, kw, code, msg):
    h = C.c_void_p()
    d = _desc(**kw)
    rc = lib.pf_plan_create(C.byref(d), C.byref(h))
    assert rc == code
    assert msg in lib.pf_last_error().decode()
    assert not h.value


def test_null_arguments(lib):
    assert lib.pf_plan_create(None, None) == -1
    assert lib.pf_step(None, None, 1, None, None) == -1
    assert lib.pf_get_step_counter(None) == -1
    assert lib.pf_set_tiles(None, None) == -1
    assert lib.pf_get_tiles(None, None, None) == -1
    assert lib.pf_autotune(None, None, 1, None, None) == -1
    assert lib.pf_plan_destroy(None) == 0


def test_perfmodel_matches_byte_table():
    from periodicflow import perfmodel

    # 256^3, 2/3 rule, FP32, unfused: the stage kernel moves 12 nk complex
    nk = 256 * 256 * 129
    pb = perfmodel.pass_bytes([256] * 3, "2/3", "fp32", fused=False)
    assert pb["stage"] == 12 * nk * 8
    # pruned inverse passes: 171 retained rows per complex axis, 86 entries on the last;
    # the axis-1 inverse reads 5 fields (u, v, w, omega_y, omega_z) and writes 6
    assert pb["inv_axis1"] == (5 * 256 * 171 * 86 + 6 * 256 * 256 * 86) * 8
    fused = perfmodel.pass_bytes([256] * 3, "2/3", "fp32")
    assert fused["fwd_axis0_stage"] == (3 * 256 * 256 * 129 + 9 * nk) * 8
    assert "stage" not in fused
    # 3/2 rule pads the inverse outputs by 3/2 per padded axis, nothing pruned
    pb32 = perfmodel.pass_bytes([64] * 3, "3/2", "fp64")
    assert pb32["inv_axis0_curl"] == (3 * 64 * 64 * 33 + 5 * 96 * 64 * 33) * 16


@pytest.mark.parametrize("kw,msg", [
    (dict(dim=4), "2D and 3D"),
    (dict(nonlinear=0), "nonlinear = 0"),
    (dict(n=(256, 256, 256)), "static scratch"),
    (dict(n=(128, 128, 128), precision=1, dealias=1), "static scratch"),
    (dict(n=(14, 16, 16)), "transform length"),
])
def test_noalloc_rejects_before_touching_the_device(lib, kw, msg):
    """The allocation-free entry point validates the descriptor and checks
    its work memory against the static scratch before any HIP call."""
    d = _desc(**kw)
    x = (C.c_char * 16)()
    rc = lib.pf_rhs_nonlinear_noalloc(C.byref(d), x, x, None)
    assert rc in (-1, -4)
    assert msg in lib.pf_last_error().decode()
    assert lib.pf_rhs_nonlinear_noalloc(None, x, x, None) == -1
