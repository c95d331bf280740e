"""Statistical comparison of two tallies (GPU vs oracle), shared by the parity tests."""
import numpy as np


def bin_stats(total, total_sq, n):
    """Mean and standard error per bin with the reference's estimator
    (WriteReflectanceClassI_and_ClassII, src/octmps_io.cpp:309-318), unclamped: a negative
    variance gives NaN, as in the reference's output."""
    mean = total / n
    var = (total_sq / n - mean * mean) / (n - 1)
    with np.errstate(invalid="ignore"):
        return mean, np.sqrt(var)


def filtered(t, max_rel=0.01):
    """The *_ScattFilt outputs' rule (src/octmps_io.cpp:359-420): a bin whose largest single
    contribution exceeds max_rel of its sum drops that contribution."""
    s, q, mx = t["sum"], t["sumsq"], t["max"]
    cut = mx > max_rel * s
    return {"sum": np.where(cut, s - mx, s), "sumsq": np.where(cut, q - mx * mx, q)}


def chi2_decimated(gpu, ref, n_gpu, n_ref, stride=6):
    """Chi^2 of independent runs over every `stride`-th bin (a detection spans 6 consecutive
    bins, so bins 6 apart share none) of both classes; returns (chi2, ndf)."""
    mg, sg = bin_stats(gpu["sum"], gpu["sumsq"], n_gpu)
    mr, sr = bin_stats(ref["sum"], ref["sumsq"], n_ref)
    mg, sg, mr, sr = (a[:, ::stride] for a in (mg, sg, mr, sr))
    sig2 = sg ** 2 + sr ** 2
    pop = sig2 > 0
    return float((((mg - mr) ** 2)[pop] / sig2[pop]).sum()), int(pop.sum())


def compare(gpu, ref, n):
    """Per-bin z scores over both classes, plus total reflectance of each side.

    Acceptance used by the tests: <= 1% of populated bins beyond 3 sigma_comb, none beyond
    5 sigma_comb, and total R (sum over bins of Class I + Class II means) within 0.5%."""
    mg, sg = bin_stats(gpu["sum"], gpu["sumsq"], n)
    mr, sr = bin_stats(ref["sum"], ref["sumsq"], n)
    # Every bin's SE must be finite on both sides (the exact tally holds every contribution's
    # square, so a negative variance would be a tally fault, not noise).
    assert np.all(np.isfinite(sg)), f"GPU SE not finite in {int((~np.isfinite(sg)).sum())} bins"
    assert np.all(np.isfinite(sr)), f"oracle SE not finite in {int((~np.isfinite(sr)).sum())} bins"
    sig = np.sqrt(sg ** 2 + sr ** 2)
    pop = sig > 0
    z = np.zeros_like(mg)
    z[pop] = (mg[pop] - mr[pop]) / sig[pop]
    same_zero = (~pop) & (mg == mr)
    rg, rr = mg.sum(), mr.sum()
    return {
        "bins": int(pop.sum()),
        "frac_3sigma": float((np.abs(z[pop]) > 3).mean()) if pop.any() else 0.0,
        "max_abs_z": float(np.abs(z[pop]).max()) if pop.any() else 0.0,
        "unpopulated_mismatch": int((~pop & ~same_zero).sum()),
        "R_gpu": float(rg),
        "R_ref": float(rr),
        "R_rel": float(abs(rg - rr) / rr) if rr else float(abs(rg - rr)),
        "Rfilt_rel": _rel(filtered(gpu)["sum"].sum(), filtered(ref)["sum"].sum()),
    }


def _rel(a, b):
    return float(abs(a - b) / b) if b else float(abs(a - b))
