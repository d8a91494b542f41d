"""Parity bars shared by the GPU tests.

FP64: per-molecule energies within 1e-9 kcal/mol + 1e-11 relative of the
oracle (summation order, FFT algorithm and the device's erf differ);
counts, edens and ecorr bit-exact with unit weights.

FP32 (pair terms and PME in FP32, per-site and per-molecule sums in FP64):
every per-molecule energy within FP32_ABS = 1e-4 kcal/mol of the FP64
value, whatever its magnitude; the bin-flip fraction is reported.
"""
import numpy as np

TOL64 = dict(rtol=1e-11, atol=1e-9)
FP32_ABS = 1e-4  # kcal/mol, per molecule


def fp32_ok(e32, e64):
    """(mask of entries within the FP32 bar, |dE|)."""
    err = np.abs(np.asarray(e32) - np.asarray(e64))
    return err <= FP32_ABS, err


def fp32_report(label, e32, e64, edges=None):
    """Asserts the FP32 bar on every entry and prints the largest error
    below and above |E| = 100 kcal/mol (and the bin flips when edges are
    given); returns (max |dE|, flips, entries)."""
    e32, e64 = np.asarray(e32), np.asarray(e64)
    ok, err = fp32_ok(e32, e64)
    big = np.abs(e64) > 100.0
    lo = err[~big].max() if (~big).any() else 0.0
    hi = err[big].max() if big.any() else 0.0
    flips = -1
    if edges is not None:
        flips = int(np.count_nonzero(np.searchsorted(edges, e32, side="left")
                                     != np.searchsorted(edges, e64, side="left")))
    print(f"[fp32] {label}: max |dE| {err.max():.3e} kcal/mol (|E|<=100: {lo:.3e}, |E|>100: {hi:.3e}, "
          f"{int(big.sum())} entries), bin flips {flips}/{err.size}", flush=True)
    assert ok.all(), f"{label}: FP32 error {err.max():.3e} above {FP32_ABS}"
    return float(err.max()), flips, err.size
