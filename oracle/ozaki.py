"""CPU model of the split-precision (INT8-slice) GEMM, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
bli_amd_?gemm_split, never as the thing measured. The reference has no split
path. The algorithm is the integer Ozaki scheme (Ozaki, Ogita, Oishi, Rump,
Numer. Algorithms 59, 2012; Ootomo, Ozaki, Yokota, IJHPCA 2024), computed with
the same operations in the same order as the GPU:

  e_i    = frexp exponent of max_p |a_ip|, or 0 for an all-zero row;
  digits   r <- ldexp(a, -e_i); repeat s times: y = 128 r, t = trunc(y),
           r = y - t.  A_t is t as int8;
  S_d    = sum_{t+u=d} A_t B_u, exact (int64 here, int32 on the device);
  v      = sum over classes c = s-1 .. 0 of ldexp(S_{c+2}, -7(c+2)),
           added in FP64 in that order;
  c_ij   = alpha * ldexp(v, e_i + f_j) (+ beta c_ij).

All of these FP64 operations are exact or correctly rounded IEEE operations,
so for beta == 0 and one pass over k the GPU result must equal this model
bit for bit. The model itself is pinned in tests/test_split_model.py against
the double-double oracle and the a-priori truncation bound.
"""
from __future__ import annotations

import numpy as np


def row_exponents(x: np.ndarray) -> np.ndarray:
    """frexp exponent of each row's largest magnitude (0 for zero rows)."""
    mx = np.max(np.abs(x), axis=1) if x.shape[1] else np.zeros(x.shape[0])
    _, e = np.frexp(mx)
    return np.where(mx == 0, 0, e).astype(np.int64)


def mean_gap(x: np.ndarray) -> float:
    """Mean over nonzero rows of (row max exponent - mean entry exponent)."""
    gaps = []
    for row in x:
        nz = row[row != 0]
        if nz.size == 0:
            continue
        _, e = np.frexp(nz.astype(np.float64))
        gaps.append(e.max() - e.mean())
    return float(np.mean(gaps)) if gaps else 0.0


def choose_slices(is_double: bool, gap_bits: float) -> int:
    """Slice-count policy: ceil((significand bits + gap) / 7), at least 2."""
    need = (50.0 if is_double else 22.0) + gap_bits
    return max(2, int(np.ceil(need / 7.0)))


def slices(x: np.ndarray, s: int):
    """Digits of the rows of x (rows x K): list of s int64 arrays, exponents."""
    e = row_exponents(x)
    r = np.ldexp(x.astype(np.float64), -e[:, None])
    out = []
    for _ in range(s):
        y = r * 128.0
        t = np.trunc(y)
        out.append(t.astype(np.int64))
        r = y - t
    return out, e


def split_gemm(a_eff: np.ndarray, b_eff: np.ndarray, s: int, alpha: float = 1.0,
               beta: float = 0.0, c: np.ndarray | None = None) -> np.ndarray:
    """alpha * a_eff @ b_eff + beta * c through s INT8 slices per operand.

    a_eff is m x k (op(A)), b_eff is k x n (op(B)); returns float64 m x n.
    """
    A, ea = slices(a_eff, s)
    B, eb = slices(np.ascontiguousarray(b_eff.T), s)
    m, n = a_eff.shape[0], b_eff.shape[1]
    classes = [np.zeros((m, n), dtype=np.int64) for _ in range(s)]
    for t in range(s):
        for u in range(s - t):
            classes[t + u] += A[t] @ B[u].T
    v = np.zeros((m, n))
    for cidx in range(s - 1, -1, -1):
        v = v + np.ldexp(classes[cidx].astype(np.float64), -7 * (cidx + 2))
    val = alpha * np.ldexp(v, (ea[:, None] + eb[None, :]))
    if beta != 0.0 and c is not None:
        val = val + beta * c.astype(np.float64)
    return val


def truncation_bound(a_eff: np.ndarray, b_eff: np.ndarray, s: int, alpha: float = 1.0) -> float:
    """|alpha| k (s+2) 2^-7s sqrt(sum_i 4^e_i) sqrt(sum_j 4^f_j)."""
    k = a_eff.shape[1]
    ea = row_exponents(a_eff)[np.max(np.abs(a_eff), axis=1) != 0]
    eb = row_exponents(np.ascontiguousarray(b_eff.T))[np.max(np.abs(b_eff), axis=0) != 0]
    sa = float(np.sum(np.ldexp(1.0, 2 * ea)))
    sb = float(np.sum(np.ldexp(1.0, 2 * eb)))
    return abs(alpha) * k * (s + 2) * 2.0 ** (-7 * s) * np.sqrt(sa) * np.sqrt(sb)


def gen_spread(rng: np.random.Generator, shape, sigma: float) -> np.ndarray:
    """u * 2^round(N(0, sigma)), u ~ U[-1, 1]: the exponent-spread workload."""
    u = rng.uniform(-1.0, 1.0, shape)
    if sigma > 0:
        e = np.rint(rng.normal(0.0, sigma, shape)).astype(np.int64)
        return np.ldexp(u, e)
    return u
