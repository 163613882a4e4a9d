"""Pins the oracle's exactly specified elementary functions (oracle/ora_math.h)
against correctly rounded binary64 references (numpy), with error bounds in
units of the binary32 result's ulp.  A wrong coefficient, sign, reduction
constant or quadrant breaks these bounds by orders of magnitude."""
import numpy as np
import pytest

import oracle


def ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / sp


RS = np.random.default_rng(2024)


@pytest.mark.parametrize("fn,npfn,lo,hi,bound", [
    ("sin", np.sin, -30.0, 30.0, 2.0),
    ("cos", np.cos, -30.0, 30.0, 2.0),
    ("tan", np.tan, -1.5, 1.5, 4.0),
    ("exp", np.exp, -87.0, 88.0, 2.0),
])
def test_trig_exp_ulp_bounds(fn, npfn, lo, hi, bound):
    x = RS.uniform(lo, hi, 400_000).astype(np.float32)
    got = oracle.math_batch(fn, x)
    ref = npfn(x.astype(np.float64))
    m = np.abs(ref) > 1e-3 if fn in ("sin", "cos") else np.ones_like(ref, bool)
    assert ulp_err(got[m], ref[m]).max() <= bound
    if fn in ("sin", "cos"):   # near zeros: absolute error bound
        assert np.max(np.abs(got - ref)) < 2e-7


def test_log_ulp_bounds():
    x = np.exp(RS.uniform(-87, 88, 400_000)).astype(np.float32)
    x = np.concatenate([x, RS.uniform(0.5, 2.0, 200_000).astype(np.float32),
                        np.float32(2.0) ** -np.arange(127, 149, dtype=np.float32)])  # subnormals
    got = oracle.math_batch("log", x)
    ref = np.log(x.astype(np.float64))
    m = np.abs(ref) > 1e-6
    assert ulp_err(got[m], ref[m]).max() <= 3.0


def test_atan2_acos_pow_bounds():
    y = RS.normal(size=300_000).astype(np.float32)
    x = RS.normal(size=300_000).astype(np.float32)
    assert ulp_err(oracle.math_batch("atan2", y, x), np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 4.0
    c = RS.uniform(-1, 1, 300_000).astype(np.float32)
    assert ulp_err(oracle.math_batch("acos", c), np.arccos(c.astype(np.float64))).max() <= 4.0
    b = RS.uniform(1e-6, 1.0, 300_000).astype(np.float32)
    e = RS.uniform(0.0, 1.0, 300_000).astype(np.float32)
    assert ulp_err(oracle.math_batch("pow", b, e), np.power(b.astype(np.float64), e.astype(np.float64))).max() <= 16.0


def test_special_values():
    L = oracle.lib()
    assert L.ora_sin(0.0) == 0.0 and L.ora_cos(0.0) == 1.0
    assert L.ora_exp(0.0) == 1.0 and L.ora_log(1.0) == 0.0
    assert L.ora_exp(-200.0) == 0.0 and np.isinf(L.ora_exp(100.0))
    assert np.isinf(L.ora_log(0.0)) and L.ora_log(0.0) < 0 and np.isnan(L.ora_log(-1.0))
    assert L.ora_atan2(0.0, 1.0) == 0.0
    assert L.ora_atan2(0.0, -1.0) == np.float32(np.pi)
    assert L.ora_atan2(1.0, 0.0) == np.float32(np.pi / 2)
    assert L.ora_acos(1.0) == 0.0 and L.ora_acos(-1.0) == np.float32(np.pi)
    assert L.ora_pow(0.0, 0.5) == 0.0 and L.ora_pow(2.0, 0.0) == 1.0
    # quadrant boundaries of the pi/2 reduction
    for k in range(-8, 9):
        x = np.float32(k * np.pi / 2)
        assert abs(L.ora_sin(x) - np.sin(np.float64(x))) < 2e-7
        assert abs(L.ora_cos(x) - np.cos(np.float64(x))) < 2e-7


def test_gamma_pack_closed_form():
    L = oracle.lib()
    assert L.ora_pack_gamma(0.0, 0.0, 0.0) == 0xFF000000
    assert L.ora_pack_gamma(1.0, 1.0, 1.0) == 0xFFFFFFFF
    assert L.ora_pack_gamma(2.0, -1.0, 0.0) == 0xFF0000FF     # clamp, ABGR packing
    v = L.ora_pack_gamma(0.5, 0.25, 0.125)
    exp = [int(255 * (c ** (1 / 2.2))) for c in (0.5, 0.25, 0.125)]
    assert (v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xFF) == tuple(exp)
