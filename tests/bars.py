"""Parity bars shared by the GPU tests.

FP64: per-molecule energies within 1e-9 kcal/mol + 1e-11 relative of the
oracle (summation order, FFT algorithm and the device's erf differ);
counts, edens and ecorr bit-exact with unit weights.

FP32 (pair terms and PME in FP32, per-site and per-molecule sums in FP64):
every per-molecule energy within FP32_ABS = 1e-4 kcal/mol of the FP64 value
wherever |E| <= FP32_EMAX = 1e4 kcal/mol (the last finite bin edge of every
energy coordinate used here).  Above it, where every value falls in the
overflow bin, the energies come from overlapping sites (r^-12) and the
error is bounded relative to |E| (FP32_REL); the bin-flip fraction is
reported.
"""
import numpy as np

TOL64 = dict(rtol=1e-11, atol=1e-9)
FP32_ABS = 1e-4    # kcal/mol, per molecule, |E| <= FP32_EMAX
FP32_EMAX = 1e4    # kcal/mol
FP32_REL = 1e-7    # relative, |E| > FP32_EMAX
BANDS = (10.0, 100.0, 1e3, 1e4)


def fp32_ok(e32, e64):
    """(mask of entries within the FP32 bar, |dE|)."""
    e32, e64 = np.asarray(e32), np.asarray(e64)
    err = np.abs(e32 - e64)
    mag = np.abs(e64)
    ok = np.where(mag <= FP32_EMAX, err <= FP32_ABS, err <= FP32_REL * mag)
    return ok, err


def band_errors(e32, e64):
    """Largest |dE| per magnitude band of |E| (and the count in the band)."""
    e32, e64 = np.asarray(e32), np.asarray(e64)
    err = np.abs(e32 - e64)
    mag = np.abs(e64)
    out, lo = {}, -1.0
    for hi in BANDS + (np.inf,):
        sel = (mag > lo) & (mag <= hi)
        key = f"<={hi:g}" if np.isfinite(hi) else f">{lo:g}"
        out[key] = (float(err[sel].max()) if sel.any() else 0.0, int(sel.sum()))
        lo = hi
    return out


def fp32_report(label, e32, e64, edges=None):
    """Asserts the FP32 bar on every entry and prints the largest error per
    magnitude band (and the bin flips when edges are given); returns
    (max |dE| for |E| <= FP32_EMAX, flips, entries)."""
    e32, e64 = np.asarray(e32), np.asarray(e64)
    ok, err = fp32_ok(e32, e64)
    mag = np.abs(e64)
    inr = mag <= FP32_EMAX
    worst = float(err[inr].max()) if inr.any() else 0.0
    rel = float((err[~inr] / mag[~inr]).max()) if (~inr).any() else 0.0
    flips = -1
    if edges is not None:
        flips = int(np.count_nonzero(np.searchsorted(edges, e32, side="left")
                                     != np.searchsorted(edges, e64, side="left")))
    bands = ", ".join(f"|E|{k}: {v[0]:.2e} ({v[1]})" for k, v in band_errors(e32, e64).items())
    print(f"[fp32] {label}: max |dE| {worst:.3e} kcal/mol where |E| <= {FP32_EMAX:g}, max relative above "
          f"{rel:.2e}; bands {bands}; bin flips {flips}/{err.size}", flush=True)
    assert ok.all(), f"{label}: FP32 error above the bar ({worst:.3e} absolute, {rel:.2e} relative)"
    return worst, flips, err.size
