"""The FP32 erfc of the pair kernel (ermod_amd/csrc/ermgpu_erfc.h) against
scipy's double erfc, evaluated on the CPU in the kernel's FP32 operation
order: u = x * fl(1/XMAX), Horner with fused multiply-adds, exp(-x^2) as
2^(fl(x^2) * fl(-log2 e)), one final product.  The bar is the absolute error
(what a pair energy q q' erfc(alpha r) / r sees): the device library's erfcf
measures 1.5e-7 largest and 1.2e-8 rms over the same range on the MI355X.
The fit's error is smooth (not rounding noise), so it adds up over the pairs
of a molecule: on the mixed-solvent systems' Coulomb sums that comes to about
1e-6 kcal/mol per molecule, two orders below the FP32 bar.
"""
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ermod_amd", "csrc",
                   "ermgpu_erfc.h")


def header():
    text = open(HDR).read()
    xmax = float(re.search(r"ERFC_XMAX = ([0-9.eE+-]+)f;", text).group(1))
    deg = int(re.search(r"ERFC_DEG = (\d+);", text).group(1))
    body = text[text.index("ERFC_R[ERFC_DEG + 1] = {") + len("ERFC_R[ERFC_DEG + 1] = {"):]
    body = body[:body.index("}")]
    coef = [float(v.strip().rstrip("f")) for v in body.split(",") if v.strip()]
    assert len(coef) == deg + 1
    return xmax, np.array(coef, dtype=np.float32)


def fma32(a, b, c):
    # one rounding of a*b + c: the product of two floats is exact in double,
    # and the double sum rounded once more to float differs from a true FMA
    # only on double-rounding ties, far below the bars here
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def erfc_fit32(x, xmax, c):
    x = np.asarray(x, np.float32)
    u = (x.astype(np.float64) * np.float64(np.float32(1.0 / xmax))).astype(np.float32)
    r = np.full_like(x, c[-1])
    for ck in c[-2::-1]:
        r = fma32(r, u, ck)
    x2 = (x.astype(np.float64) * x).astype(np.float32)
    arg = (x2.astype(np.float64) * np.float64(np.float32(-1.4426950408889634))).astype(np.float32)
    e = np.exp2(arg.astype(np.float64)).astype(np.float32)
    return (e.astype(np.float64) * r).astype(np.float32)


def test_erfc_fit_absolute_error():
    special = pytest.importorskip("scipy.special")
    xmax, c = header()
    x = np.linspace(0.0, xmax, 1_000_001).astype(np.float32)
    ref = special.erfc(x.astype(np.float64))
    err = np.abs(erfc_fit32(x, xmax, c).astype(np.float64) - ref)
    assert err.max() < 1.3e-7, err.max()
    assert np.sqrt((err ** 2).mean()) < 2.5e-8
    # known values: erfc(0) = 1, erfc(1) = 0.157299207..., erfc(3) = 2.209e-5
    # (the fit bounds the absolute error, so far out it is relatively loose)
    v = erfc_fit32(np.array([0.0, 1.0, 3.0], np.float32), xmax, c).astype(np.float64)
    assert abs(v[0] - 1.0) < 1e-7
    assert abs(v[1] - 0.15729920705028513) < 1e-7
    assert abs(v[2] - 2.209049699858544e-05) < 5e-9


def test_erfc_fit_range_covers_the_workloads():
    """alpha * elecut of the synthetic systems lies inside the fitted range
    (the kernel falls back to erfcf beyond it)."""
    xmax, _ = header()
    from ermod_amd import synth
    for name in ("refs20", "mixed"):
        sy = synth.WORKLOADS[name]()
        s = sy.system
        assert s.screen * s.elecut <= xmax, (name, s.screen, s.elecut)
