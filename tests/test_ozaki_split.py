"""The digit split behind the int8 Gram (mld_amd/csrc/gram.hip, oz_split_kernel / syrk_i8_kernel),
emulated in NumPy on the oracle's quadratic rows: 7 digits rounded to nearest below 2^e > 2 max|a|
fit int8 (|d| <= 64), and the 34 digit pairs with i + j <= 9 give the Gram to the error of an
FP64 sum (relative to sqrt(G_pp G_qq)), while truncated digits (common sign) or 28 pairs do not."""
import numpy as np

import oracle as O
from mld_amd import synth
from mld_amd.fit import full_to_unique_rows


def rows(nstruct=4):
    rng = np.random.default_rng(5)
    A, w = [], []
    for _ in range(nstruct):
        p, c = synth.bcc_supercell(rng, 2.855, 2, 0.1)
        e, g, nl = O.bso4(p, c, np.zeros(len(p), np.int32), 8, 4.7)
        A.append(full_to_unique_rows(O.rows(e, g, nl, order=2), 55))
        w.append(np.concatenate([[1.0], np.full(3 * len(p), 0.1)]))
    A = np.concatenate(A) * np.sqrt(np.concatenate(w))[:, None]
    # the count and linear columns and every 8th quadratic column (the extended-precision Gram
    # of all 1,596 columns takes minutes)
    return A[:, np.r_[0:56, 56:A.shape[1]:8]]


def ozaki_gram(A, ndig=7, lmax=9, balanced=True):
    cm = np.abs(A).max(0)
    e = np.where(cm > 0, np.floor(np.log2(np.where(cm > 0, cm, 1.0))) + (2 if balanced else 1), 0)
    x = A * 2.0 ** (-e)
    digits = []
    for _ in range(ndig):
        y = x * 128.0
        t = np.rint(y) if balanced else np.trunc(y)
        digits.append(t)
        x = y - t
    bound = max(float(np.abs(d).max()) for d in digits)
    G = np.zeros((A.shape[1],) * 2, dtype=np.longdouble)
    for i, di in enumerate(digits, 1):
        for j, dj in enumerate(digits, 1):
            if i + j <= lmax:
                # integer sums (exact in int64 here, in int32 on the device)
                s = di.astype(np.int64).T @ dj.astype(np.int64)
                G += s.astype(np.longdouble) * np.longdouble(2.0) ** (-7 * (i + j))
    return G * np.outer(2.0 ** e, 2.0 ** e), bound


def test_digit_split_gram_is_fp64_accurate():
    A = rows()
    Gx = A.astype(np.longdouble).T @ A.astype(np.longdouble)
    nrm = np.sqrt(np.outer(np.diag(Gx), np.diag(Gx))).astype(float)

    def err(G):
        return float((np.abs(G - Gx) / nrm).max())

    G, bound = ozaki_gram(A)
    assert bound <= 64
    e_fp64 = err(A.T @ A)
    assert err(G) < 4 * e_fp64 + 1e-15
    Gt, bt = ozaki_gram(A, balanced=False, lmax=8)     # truncated digits: biased diagonal
    assert bt <= 127 and err(Gt) > 100 * e_fp64
    G28, _ = ozaki_gram(A, lmax=8)                      # 28 pairs: one level short
    assert err(G28) > 4 * err(G)
