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


# ---- the modular scheme (Ozaki scheme II) ---------------------------------------
# Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
# floating-point matrix multiplication using an integer modular technique"
# (2024). Same operations, in the same order, as the library's crt.hip:
#
#   a'_ip  = trunc(ldexp(a_ip, sh - e_i))          (integers, |a'| < 2^sh)
#   r_l    = a' mod p_l as int8 (two's complement for 256, symmetric for odd p)
#   c_l    = (R_A,l R_B,l) mod p_l                 (exact integer products)
#   part_k = sum_l c_l w_lk                        (exact: 39-bit parts)
#   q      = rint((part_0 + part_1) * inv_m)
#   D_k    = part_k - q mpart_k                     (exact)
#   X      = hi + (lo + tail): (hi, lo) = TwoSum(D_0, D_1), tail = sum of
#            D_k for k = K-1 .. 2 in that order
#   c_ij   = alpha * ldexp(X, e_i + f_j - 2 sh) (+ beta c_ij)

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191,
          181, 179, 173, 167, 163, 157, 151, 149, 139, 137, 131, 127, 113, 109, 107, 103]
PART_BITS = 39


def crt_const(n: int) -> dict:
    """Moduli, weights w_l = (M/p_l)((M/p_l)^-1 mod p_l) and M, each split into
    39-bit parts (most significant first) as exact doubles, and 1/M."""
    import math
    ps = MODULI[:n]
    M = math.prod(ps)
    L = M.bit_length()
    parts = -(-L // PART_BITS)
    lo = [max(0, L - PART_BITS * (k + 1)) for k in range(parts)]
    width = [L - PART_BITS * k - lo[k] for k in range(parts)]

    def split(v: int):
        return [math.ldexp(float((v >> lo[k]) & ((1 << width[k]) - 1)), lo[k])
                for k in range(parts)]
    w = []
    for p in ps:
        Ml = M // p
        w.append(split(Ml * pow(Ml % p, -1, p)))
    top_lo = max(0, L - 64)
    inv_m = math.ldexp(1.0 / float(M >> top_lo), -top_lo)
    return {"p": ps, "M": M, "w": w, "mpart": split(M), "inv_m": inv_m, "parts": parts}


def crt_choose(is_double: bool, gap_bits: float, k: int):
    """(moduli, sh): sh = ceil(significand bits + gap); the fewest moduli with
    log2 M >= 2 sh + ceil(log2 k) + 2; (0, sh) if none suffice or k > 2^17."""
    import math
    sh = int(math.ceil((50.0 if is_double else 22.0) + gap_bits))
    if k > (1 << 17):
        return 0, sh
    bits = 2.0 * sh + math.ceil(math.log2(max(k, 2))) + 2.0
    acc = 0.0
    for n, p in enumerate(MODULI, start=1):
        acc += math.log2(p)
        if n >= 2 and acc >= bits:
            return n, sh
    return 0, sh


def crt_shift(n: int, k: int) -> int:
    """Largest sh that n moduli support at depth k."""
    import math
    bits = sum(math.log2(p) for p in MODULI[:n]) - math.ceil(math.log2(max(k, 2))) - 2.0
    return max(1, int(math.floor(bits / 2.0)))


def crt_integers(x: np.ndarray, sh: int):
    """trunc(ldexp(x_ip, sh - e_i)) per row (exact doubles) and the exponents."""
    e = row_exponents(x)
    return np.trunc(np.ldexp(x.astype(np.float64), (sh - e)[:, None])), e


def crt_residues_of(ints: list, shape, p: int) -> np.ndarray:
    """int8-range residues modulo p of Python integers (row-major, shape)."""
    r = np.array([v % p for v in ints], dtype=np.int64).reshape(shape)
    if p == 256:
        return np.where(r >= 128, r - 256, r)
    return np.where(r > (p - 1) // 2, r - p, r)


def crt_residues(t: np.ndarray, p: int) -> np.ndarray:
    """int8-range residues of the integer-valued doubles t modulo p."""
    return crt_residues_of([int(v) for v in t.ravel()], t.shape, p)


def crt_gemm(a_eff: np.ndarray, b_eff: np.ndarray, n: int, sh: int, alpha: float = 1.0,
             beta: float = 0.0, c0: np.ndarray | None = None, exact_out: list | None = None):
    """C = alpha op(A) op(B) (+ beta C0) by the modular scheme with n moduli
    and sh fractional bits. exact_out, if given, receives the exact integer
    products (Python ints) for checking the reconstruction."""
    cc = crt_const(n)
    ta, ea = crt_integers(a_eff, sh)
    tb, eb = crt_integers(b_eff.T, sh)
    m, k = ta.shape
    nn = tb.shape[0]
    ia = [int(v) for v in ta.ravel()]
    ib = [int(v) for v in tb.ravel()]
    prods = []
    for p in cc["p"]:
        ra, rb = crt_residues_of(ia, ta.shape, p), crt_residues_of(ib, tb.shape, p)
        prods.append(np.mod(ra @ rb.T, p).astype(np.float64))
    parts = [np.zeros((m, nn)) for _ in range(cc["parts"])]
    for l, c in enumerate(prods):
        for kk in range(cc["parts"]):
            parts[kk] = parts[kk] + c * cc["w"][l][kk]        # exact
    zero = np.zeros((m, nn))
    q = np.rint((parts[0] + (parts[1] if cc["parts"] > 1 else zero)) * cc["inv_m"])
    D = [parts[kk] - q * cc["mpart"][kk] for kk in range(cc["parts"])]   # each exact
    d1 = D[1] if cc["parts"] > 1 else zero
    hi = D[0] + d1                                             # TwoSum(D_0, D_1)
    bb = hi - D[0]
    lo = (D[0] - (hi - bb)) + (d1 - bb)
    tail = zero
    for kk in range(cc["parts"] - 1, 1, -1):
        tail = tail + D[kk]
    X = hi + (lo + tail)
    if exact_out is not None:
        ra_ = [ia[i * k:(i + 1) * k] for i in range(m)]
        rb_ = [ib[j * k:(j + 1) * k] for j in range(nn)]
        exact_out.append([[sum(x * y for x, y in zip(ra_[i], rb_[j])) for j in range(nn)]
                          for i in range(m)])
    val = alpha * np.ldexp(X, (ea[:, None] + eb[None, :] - 2 * sh))
    if beta == 0.0 or c0 is None:
        return val
    return val + beta * c0


def crt_bound(a_eff: np.ndarray, b_eff: np.ndarray, sh: int, alpha: float = 1.0) -> float:
    """A-priori Frobenius bound |alpha| k 2^(1-sh) (1 + 2^-sh) ||2^e|| ||2^f||."""
    ea = row_exponents(a_eff).astype(np.float64)
    eb = row_exponents(b_eff.T).astype(np.float64)
    nza = np.max(np.abs(a_eff), axis=1) > 0
    nzb = np.max(np.abs(b_eff), axis=0) > 0
    k = a_eff.shape[1]
    return (abs(alpha) * k * 2.0 ** (1 - sh) * (1 + 2.0 ** -sh)
            * np.sqrt(np.sum(np.exp2(2 * ea[nza]))) * np.sqrt(np.sum(np.exp2(2 * eb[nzb]))))
