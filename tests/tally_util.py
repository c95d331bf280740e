"""Builds integer tallies (include/octmps.h layout) from exact values, for host-side tests of
the finalisation and the writer."""
import numpy as np

import octmps as O


def digits(values, exp0, ndig):
    """Canonical digits (each < 2^32) of non-negative doubles that are multiples of 2^exp0."""
    out = np.zeros((ndig, len(values)), np.uint64)
    for b, v in enumerate(np.asarray(values, np.float64)):
        num, den = float(v).as_integer_ratio()
        scaled = num * 2 ** (-exp0)
        assert scaled % den == 0, "value is not a multiple of the digit unit"
        V = scaled // den
        assert V >> (32 * ndig) == 0
        for k in range(ndig):
            out[k, b] = (V >> (32 * k)) & 0xFFFFFFFF
    return out


def make_tally(sums, sqs, maxs, cnts, counters):
    """sums, sqs, maxs, cnts: [2, B]; maxs must be FP32 values (as the device stores them)."""
    B = sums.shape[1]
    W = O.TALLY_CLASS_WORDS
    t = np.zeros(O.tally_words(B), np.uint64)
    t[:len(counters)] = counters
    for c in range(2):
        base = 4 + c * W * B
        assert np.array_equal(maxs[c].astype(np.float32).astype(np.float64), maxs[c])
        t[base:base + O.TALLY_SUM_DIGITS * B] = digits(sums[c], O.TALLY_SUM_EXP, O.TALLY_SUM_DIGITS).ravel()
        q0 = base + O.TALLY_SUM_DIGITS * B
        t[q0:q0 + O.TALLY_SQ_DIGITS * B] = digits(sqs[c], O.TALLY_SQ_EXP, O.TALLY_SQ_DIGITS).ravel()
        t[base + (W - 2) * B:base + (W - 1) * B] = cnts[c]
        t[base + (W - 1) * B:base + W * B] = np.ascontiguousarray(maxs[c], np.float64).view(np.uint64)
    return t
