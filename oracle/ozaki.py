"""Ozaki-scheme split-precision DGEMM, restated in numpy (TEST INFRASTRUCTURE ONLY).

No reference counterpart: CA3DMM has no split-precision path.  This module
is the specification the HIP kernels (ca3dmm_amd/csrc/ozaki.hip) are checked
against bit for bit, and the source of the accuracy study.

Scheme (error-free splitting, Ozaki et al. 2012, with INT8 slices):
  * row i of op(A) is scaled by sigma_i = 2^(e_i + 1), e_i = frexp exponent
    of max_l |a_il|, so |a'_il| = |a_il / sigma_i| <= 1/2; columns of op(B)
    likewise by tau_j.  Zero rows/columns get sigma = 0 (no contribution).
  * slice t = 0..s-1 of a' is q_t = rint(r_t * 2^(7(t+1))) with r_0 = a' and
    r_{t+1} = r_t - q_t * 2^(-7(t+1)); every step is exact in FP64 and
    |q_t| <= 64, so slices are int8 and a' = sum_t q_t 2^(-7(t+1)) + r_s with
    |r_s| <= 2^(-7s-1).
  * diagonal d = 0..s-1: D_d = sum_{t+u=d} Q^A_t Q^B_u, an exact integer
    matrix (INT32 while k*(d+1)*64*64 < 2^31).
  * C_ij = sigma_i tau_j * sum_{d=s-1..0} 2^(-7(d+2)) D_d,ij, the FP64 sum
    taken from the smallest diagonal to the largest (fixed order, so the
    result is reproducible bit for bit); then alpha*C + beta*C_in.
"""
import numpy as np

SLICE_BITS = 7
QMAX = 64


def scale_exponents(X, axis):
    """e such that max |X| along `axis` < 2^e (frexp exponent); -inf marker
    (None -> 0 scale) for all-zero lines.  Returns (exponent int array,
    zero mask)."""
    mx = np.max(np.abs(X), axis=axis)
    _, e = np.frexp(mx)
    return e.astype(np.int64), mx == 0


def split(X, s, axis):
    """Slices of X scaled along `axis` (axis=1: rows of op(A); axis=0:
    columns of op(B)).  Returns (list of s int8 arrays shaped like X,
    scale exponent E = e + 1 per line, zero mask, remainder r_s)."""
    e, zero = scale_exponents(X, axis)
    E = e + 1
    shp = (-1, 1) if axis == 1 else (1, -1)
    r = np.ldexp(X, -E.reshape(shp))
    r = np.where(zero.reshape(shp), 0.0, r)
    slices = []
    for t in range(s):
        w = SLICE_BITS * (t + 1)
        q = np.rint(np.ldexp(r, w))
        assert np.all(np.abs(q) <= QMAX)
        slices.append(q.astype(np.int8))
        r = r - np.ldexp(q, -w)
    return slices, E, zero, r


def diagonals(SA, SB, s):
    """D_d = sum_{t+u=d} SA[t] @ SB[u] as exact integers (computed in FP64,
    exact while k * s * 64 * 64 < 2^53)."""
    out = []
    k = SA[0].shape[1]
    assert k * s * QMAX * QMAX < 2 ** 53
    for d in range(s):
        acc = None
        for t in range(d + 1):
            P = SA[t].astype(np.float64) @ SB[d - t].astype(np.float64)
            acc = P if acc is None else acc + P
        out.append(acc)
    return out


def ozaki_dgemm(opA, opB, s, alpha=1.0, beta=0.0, Cin=None):
    """Reference result of the scheme for op(A) (m x k) and op(B) (k x n)."""
    SA, EA, zA, _ = split(opA, s, axis=1)
    SB, EB, zB, _ = split(opB, s, axis=0)
    D = diagonals(SA, SB, s)
    acc = np.zeros(D[0].shape)
    for d in range(s - 1, -1, -1):
        acc = acc + np.ldexp(D[d], -SLICE_BITS * (d + 2))
    C = np.ldexp(acc, EA.reshape(-1, 1) + EB.reshape(1, -1))
    C = np.where(zA.reshape(-1, 1) | zB.reshape(1, -1), 0.0, C)
    C = alpha * C
    if beta != 0.0:
        C = C + beta * Cin
    return C


def phi_matrix(rnd, shape, phi):
    """Test inputs of the precision study: uniform[-0.5, 0.5] * exp(phi * N(0,1))."""
    return rnd.uniform(-0.5, 0.5, shape) * np.exp(phi * rnd.standard_normal(shape))


def products(s):
    """Number of INT8 slice products of the triangular scheme."""
    return s * (s + 1) // 2


# ---------------------------------------------------------------------------
# 8-bit balanced digits.  Each slice carries a full int8 digit in [-128, 127]
# of base 256.  Top-down rounding cannot produce them (a remainder of exactly
# half a unit would need +128), so the digits are taken bottom-up from the
# exact integer N = rint(a' * 2^(8s)), |a'| <= 253/512: N = M * 2^z with M the
# 53-bit significand, so only a window of at most 9 digits is non-zero; the
# window X = M << (z mod 8) (<= 61 bits) is converted to balanced digits with
# carries, c = ((X + 128) mod 256) - 128, X = (X - c) / 256.
# ---------------------------------------------------------------------------
DIGIT_BITS = 8
TOP_MANT = 253.0 / 256.0


def nonfinite_lines(X, axis):
    """Lines (axis=1: rows, axis=0: columns) holding an Inf or a NaN.  Both
    forms slice such a line as a zero line and write NaN into every C entry
    of its row (op(A)) or column (op(B))."""
    return ~np.all(np.isfinite(X), axis=axis)


def line_scale8(X, axis):
    """Per-line scale exponent of the 8-bit scheme.  With f, e = frexp(max |x|)
    of the line: E = e + 1 when f <= 253/256, else E = e + 2.  Then every
    scaled entry a' = x 2^-E has |a'| <= 253/512, so N = rint(a' 2^(8s)) has
    |N| / 256^(s-1) <= 126.5 and its top balanced digit stays in [-128, 127]
    (the lower digits add at most 128/255 in magnitude).  Returns (E, zero
    mask); a line with a non-finite element counts as zero here."""
    bad = nonfinite_lines(X, axis)
    mx = np.max(np.abs(np.where(np.isfinite(X), X, 0.0)), axis=axis)
    f, e = np.frexp(mx)
    E = np.where(f <= TOP_MANT, e + 1, e + 2).astype(np.int64)
    return E, (mx == 0) | bad


def split8(X, s, axis):
    """Balanced base-256 slices of X scaled per line along `axis` by 2^-E
    (line_scale8).  Returns (list of s int8 arrays, E, zero mask)."""
    E, zero = line_scale8(X, axis)
    shp = (-1, 1) if axis == 1 else (1, -1)
    a = np.ldexp(np.where(zero.reshape(shp), 0.0, X), -E.reshape(shp))
    f, ex = np.frexp(a)                         # a = f * 2^ex, f in [0.5, 1)
    M = np.ldexp(f, 53).astype(np.int64)        # exact 53-bit signed significand
    z = ex.astype(np.int64) - 53 + 8 * s        # N = M * 2^z
    neg = z < 0
    # z < 0: N = rint(M * 2^z) (< 2^53, exact in FP64); window offset 0
    Nsmall = np.rint(np.ldexp(M.astype(np.float64), np.minimum(z, 0))).astype(np.int64)
    off = np.where(neg, 0, z // 8)
    Xw = np.where(neg, Nsmall, M << np.where(neg, 0, z % 8))
    Xw = np.where(a == 0, 0, Xw)
    digits = [np.zeros(X.shape, np.int64) for _ in range(s)]
    for i in range(10):                         # at most 9 non-zero digits + slack
        c = ((Xw + 128) & 255) - 128
        Xw = (Xw - c) >> 8
        j = off + i                             # position from the least significant digit
        t = s - 1 - j                           # slice index (0 = most significant)
        ok = (t >= 0) & (t < s)
        for tt in range(s):
            sel = ok & (t == tt)
            digits[tt] = np.where(sel, c, digits[tt])
    assert np.all(Xw == 0), "digit window overflow"
    out = []
    for d in digits:
        assert np.all((d >= -128) & (d <= 127))
        out.append(d.astype(np.int8))
    return out, E, zero


def reconstruct8(slices, E, axis):
    """sum_t q_t 2^(-8(t+1)) scaled back by 2^E (exact check helper)."""
    s = len(slices)
    acc = np.zeros(slices[0].shape)
    for t in range(s - 1, -1, -1):
        acc = acc + np.ldexp(slices[t].astype(np.float64), -DIGIT_BITS * (t + 1))
    shp = (-1, 1) if axis == 1 else (1, -1)
    return np.ldexp(acc, E.reshape(shp))


def kchunk8(k, s):
    """k-block length over which one diagonal stays exact in INT32:
    s * kchunk * 128 * 128 < 2^31, a multiple of 256, at most k rounded up
    to 256 (the rule of the device kernel, whose LDS ring of four 64-byte
    stages turns a whole number of times per chunk)."""
    kp = -(-max(k, 1) // 256) * 256
    return min(kp, ((2 ** 31 - 1) // (16384 * s)) // 256 * 256)


def ozaki_dgemm8(opA, opB, s, alpha=1.0, beta=0.0, Cin=None):
    """The scheme with 8-bit balanced digits: the reference result the device
    kernels reproduce bit for bit.  The diagonals are added from d = s-1 down
    to 0; within a diagonal, k is processed in INT32-exact chunks in k order."""
    SA, EA, zA = split8(opA, s, axis=1)
    SB, EB, zB = split8(opB, s, axis=0)
    k = opA.shape[1]
    kc = kchunk8(k, s)
    chunks = [slice(c0, min(k, c0 + kc)) for c0 in range(0, max(k, 1), kc)]
    Dc = [diagonals8([x[:, sl] for x in SA], [x[sl, :] for x in SB], s) for sl in chunks]
    acc = np.zeros((opA.shape[0], opB.shape[1]))
    for d in range(s - 1, -1, -1):
        for D in Dc:
            acc = acc + np.ldexp(D[d], -DIGIT_BITS * (d + 2))
    C = np.ldexp(acc, EA.reshape(-1, 1) + EB.reshape(1, -1))          # one rounding
    C = np.where(zA.reshape(-1, 1) | zB.reshape(1, -1), 0.0, C)
    C = np.where(nonfinite_lines(opA, 1).reshape(-1, 1) | nonfinite_lines(opB, 0).reshape(1, -1), np.nan, C)
    C = alpha * C
    if beta != 0.0:
        C = C + beta * Cin
    return C


def diagonals8(SA, SB, s):
    out = []
    k = SA[0].shape[1]
    assert k * s * 128 * 128 < 2 ** 53
    for d in range(s):
        acc = None
        for t in range(d + 1):
            P = SA[t].astype(np.float64) @ SB[d - t].astype(np.float64)
            acc = P if acc is None else acc + P
        out.append(acc)
    return out


# ---------------------------------------------------------------------------
# Ozaki scheme II: the integer-modular (CRT) form on INT8 products
# (Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
# floating-point matrix multiplication using an integer modular technique",
# 2025).  One INT8 product per modulus instead of s(s+1)/2 slice products.
#
#   * every line i of op(A) is scaled by a power of two 2^e_i and rounded to
#     the integer a'_il = rint(a_il 2^e_i) (exact as a double), likewise the
#     columns of op(B) by 2^f_j;
#   * e_i keeps ||a'_i||_2 <= 2^T and f_j keeps ||b'_j||_2 <= 2^T with
#     2^(2T+1) < P = prod p_t, so |C'_ij| = |a'_i . b'_j| < P/2 (Cauchy-Schwarz)
#     and C' is fixed by its residues modulo the N moduli;
#   * residue planes r_t(a') in [-(p_t-1)/2, (p_t-1)/2] ([-128, 127] for 256)
#     are int8; Z_t = R_t(A') R_t(B') is exact in INT32 over k-chunks, and
#     z_t = Z_t mod p_t (balanced);
#   * the moduli are taken in pairs, Q_0 = p_0 p_1, Q_1 = p_2 p_3, ... (for
#     odd N the last modulus stays single), and Garner's algorithm over the
#     Q_i turns (z_t) into the balanced mixed-radix digits V_i (in
#     [-Q_i/2, Q_i/2) for the even Q_0) of the unique x in [-P/2, P/2) with
#     x = C' (mod P), i.e. x = C';
#   * x = V_0 + Q_0 (V_1 + Q_1 (V_2 + ...)) is evaluated in FP64 by Horner
#     from the top digit, one rounded multiply and one rounded add per step
#     (no fused multiply-add), and C_ij = alpha 2^-e_i 2^-f_j x (+ beta C_in).
# Everything up to the digits is exact integer arithmetic, so any correct
# implementation agrees; the Horner sum fixes the rounding, and the device
# kernel follows the same order bit for bit.
#
# The line norm bound is order-independent (integer sums): with E the frexp
# exponent of the line's max |x| and, per element, ex its exponent and
# m4 = ceil(16 |frac|) in [8, 16], x^2 <= m4^2 4^(ex-4), so
# ||line||^2 <= 4^E 2^-40 S with S = sum ceil(m4^2 2^(2(ex-E)+32)) (an
# integer <= k 2^40).  Rounding adds at most sqrt(k)/2 <= 2897/2 to the norm
# (k <= 2^23 <= 2897^2), so ||a'|| <= 2^T holds when
# 2^e ||line|| <= 2^T - 2897/2.  With e = T + 20 - E - h that reads
# 4^(T-h) S <= (2^T - 2897/2)^2, which S <= 4^h - ceil(2897 2^(2h-T))
# implies ((1 - x)^2 >= 1 - 2x); h is the smallest such integer.  (Against
# the plain 2^e ||line|| <= 2^(T-1), this is one more bit for almost every
# line.)
# ---------------------------------------------------------------------------
CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
              223, 217, 211, 199, 197, 193, 191, 181, 179, 173,
              167, 163, 157, 151, 149)
CRT_MIN, CRT_MAX = 4, len(CRT_MODULI)
CRT_KMAX = 1 << 23          # S <= k 2^40 stays below 2^63


def crt_product(N):
    P = 1
    for p in CRT_MODULI[:N]:
        P *= p
    return P


def crt_T(N):
    """Norm budget T: the largest integer with 2^(2T+1) < P, taken from the
    bit length b of P (2^(b-1) <= P): T = (b - 2) // 2."""
    return (crt_product(N).bit_length() - 2) // 2


def crt_line_exponents(X, axis, N):
    """Scale exponents e (int64) and zero mask of the lines of X along
    `axis` (axis=1: rows of op(A); axis=0: columns of op(B)); a line with a
    non-finite element counts as zero here (nonfinite_lines)."""
    T = crt_T(N)
    bad = nonfinite_lines(X, axis)
    X = np.where(np.isfinite(X), X, 0.0)
    X = np.where(bad.reshape((-1, 1) if axis == 1 else (1, -1)), 0.0, X)
    mx = np.max(np.abs(X), axis=axis)
    _, E = np.frexp(mx)
    E = E.astype(np.int64)
    f, ex = np.frexp(np.abs(X))
    m4 = np.ceil(f * 16.0).astype(np.int64)        # 0 for x == 0
    Eb = E.reshape((-1, 1) if axis == 1 else (1, -1))
    sh = (2 * (ex.astype(np.int64) - Eb) + 32)        # <= 32
    m2 = m4 * m4
    up = np.where(sh >= 0, m2 << np.clip(sh, 0, 62), 0)
    dn = np.where(m2 > 0, ((m2 - 1) >> np.clip(-sh, 0, 62)) + 1, 0)
    w = np.where(sh >= 0, up, dn)
    S = np.sum(w, axis=axis, dtype=np.uint64).astype(object)
    h = np.array([crt_scale_h(int(v), T) for v in S.ravel()], dtype=np.int64)
    e = T + 20 - E - h
    return e, (mx == 0) | bad


CRT_SLACK = 2897            # ceil(sqrt(CRT_KMAX)): 2 x the largest rounding slack of a line norm


def crt_scale_h(S, T):
    """The smallest h >= 0 with S <= 4^h - ceil(2897 2^(2h - T)) (Python ints)."""
    h = 0
    while True:
        sh = 2 * h - T
        sub = CRT_SLACK << sh if sh >= 0 else -((-CRT_SLACK) >> (-sh))
        if S <= 4 ** h - sub:
            return h
        h += 1


def crt_scaled(X, axis, N):
    """Integer-valued doubles a' = rint(x 2^e) per line; zero lines 0."""
    e, zero = crt_line_exponents(X, axis, N)
    shp = (-1, 1) if axis == 1 else (1, -1)
    a = np.rint(np.ldexp(np.where(zero.reshape(shp), 0.0, X), np.where(zero, 0, e).reshape(shp)))
    return np.where(zero.reshape(shp), 0.0, a), e, zero


def balanced_mod(x, p):
    """Balanced residue of integer-valued x (numpy int64 or Python int
    arrays): [-(p-1)/2, (p-1)/2] for odd p, [-128, 127] for p = 256."""
    r = np.mod(x, p)
    half = p // 2
    return np.where(r > (half if p % 2 else half - 1), r - p, r)


def crt_residues(Aint, N):
    """Residue planes (int8) of an array of integer-valued doubles, exact:
    |a'| < 2^104 is split into four signed 26-bit pieces first."""
    parts = []
    rem = Aint
    for sh in (78, 52, 26):
        x = np.trunc(np.ldexp(rem, -sh))
        rem = rem - np.ldexp(x, sh)
        parts.append(x.astype(np.int64))
    parts.append(rem.astype(np.int64))
    out = []
    for p in CRT_MODULI[:N]:
        r = sum((x % p) * pow(2, sh, p) for x, sh in zip(parts, (78, 52, 26, 0)))
        out.append(balanced_mod(r, p).astype(np.int8))
    return out


def crt_composites(N):
    """Composite moduli of the reconstruction: Q_i = p_2i p_2i+1, and p_N-1
    alone as the last one for odd N."""
    ps = CRT_MODULI[:N]
    return [ps[i] * (ps[i + 1] if i + 1 < N else 1) for i in range(0, N, 2)]


def crt_garner(Z, N):
    """Balanced mixed-radix digits V_0..V_{C-1} (int64 arrays) over the
    composite moduli of the x in [-P/2, P/2) with x = Z_t (mod p_t) for
    every t; Z: list of residues.  Exact integer arithmetic: the residue
    modulo Q_i comes from its members by the two-modulus CRT, then
    Garner's step V_i = (x - V_0 - V_1 Q_0 - ...) / (Q_0 ... Q_i-1) mod Q_i."""
    ps = CRT_MODULI[:N]
    Qs = crt_composites(N)
    V = []
    for i, Q in enumerate(Qs):
        a = 2 * i
        x = Z[a].astype(np.int64) % ps[a]
        if a + 1 < N:
            pa, pb = ps[a], ps[a + 1]
            x = x + pa * (((Z[a + 1].astype(np.int64) - x) * pow(pa, -1, pb)) % pb)   # in [0, Q)
        for j in range(i):
            x = ((x - V[j]) * pow(Qs[j], -1, Q)) % Q
        V.append(balanced_mod(x, Q))
    return V


def crt_horner(V, N):
    """x = V_0 + Q_0 (V_1 + Q_1 (V_2 + ...)) in FP64, from the top digit,
    a rounded multiply then a rounded add per step."""
    Qs = crt_composites(N)
    x = V[-1].astype(np.float64)
    for i in range(len(Qs) - 2, -1, -1):
        x = x * float(Qs[i])
        x = x + V[i].astype(np.float64)
    return x


def ozaki2_dgemm(opA, opB, N, alpha=1.0, beta=0.0, Cin=None):
    """The CRT form: the reference result the device kernels reproduce bit
    for bit.  Residue products are exact (float64 BLAS on int8 operands:
    |Z| <= k 2^14 < 2^53)."""
    k = opA.shape[1]
    assert CRT_MIN <= N <= CRT_MAX and k <= CRT_KMAX
    Ai, eA, zA = crt_scaled(opA, 1, N)
    Bi, eB, zB = crt_scaled(opB, 0, N)
    RA, RB = crt_residues(Ai, N), crt_residues(Bi, N)
    Z = []
    for t, p in enumerate(CRT_MODULI[:N]):
        P = RA[t].astype(np.float64) @ RB[t].astype(np.float64)
        Z.append(balanced_mod(P.astype(np.int64), p))
    x = crt_horner(crt_garner(Z, N), N)
    C = np.ldexp(x, -(np.where(zA, 0, eA).reshape(-1, 1) + np.where(zB, 0, eB).reshape(1, -1)))   # one rounding
    C = np.where(zA.reshape(-1, 1) | zB.reshape(1, -1), 0.0, C)
    C = np.where(nonfinite_lines(opA, 1).reshape(-1, 1) | nonfinite_lines(opB, 0).reshape(1, -1), np.nan, C)
    C = alpha * C
    if beta != 0.0:
        C = C + beta * Cin
    return C


# ---------------------------------------------------------------------------
# Slice and modulus counts from the operands' spread (no reference
# counterpart; the library's hip_ozaki_choose and hip_ozaki_counts).
#
# S = 4096 deterministic entries (i_s, j_s) of C are sampled.  For each, with
# a = row i_s of op(A) and b = column j_s of op(B):
#   c   = sum |a_l| |b_l| / (||a||_2 ||b||_2)                 (CRT form)
#   c_s = sum |a_l| |b_l| / (max|a| ||b||_1 + max|b| ||a||_1)  (slice form)
# The CRT form's rounding error on an entry scales like 2^-T ||a|| ||b||
# against the native product's u |a|.|b|, so it needs T to grow with
# L = log2 max 1/c; the slice form keeps 8s bits relative to the line
# maxima, so it needs 8s to grow with L_s = log2 max 1/c_s.  The offsets
# are calibrated on uniform[-0.5, 0.5] exp(phi N(0, 1)) inputs, phi in
# {0.5, 2, 4}, at 8192^3 and 65536 x 65536 x 1024 (two seeds each) against
# the modulus and slice counts whose errors first matched the FP64 MFMA
# DGEMM's on the device (4096 sampled entries, double-double reference):
#   N = the smallest modulus count with T(N) >= L + 53.5 + 0.5 log2(k / 1024)
#       (log2 term for k > 1024 only; at most 25),
#   s = the smallest slice count with 8 s >= L_s + 54 + 2.33 log2(k / 1024)
#       (log2 term for k > 1024 only; at most 16).
# On those inputs the rule never asks for fewer moduli or slices than
# measured, and at most two more moduli (phi = 2 needs less than its spread
# suggests) and one more slice.  Both offsets keep a margin because
# cancellation, which the statistic does not see, sets the relative error
# (at 1024^3, phi = 0.5, 14 moduli reach the native |A||B|-relative error
# but not its max relative error).  An estimate, not a bound: entries
# outside the sample can be worse.
# ---------------------------------------------------------------------------
SPREAD_SAMPLES = 4096


def spread_samples(m, n, S=SPREAD_SAMPLES):
    """Sampled entries: i_s = (s 2654435761 + 2135587861) mod 2^32 mod m,
    j_s = (s 2246822519 + 3266489917) mod 2^32 mod n."""
    s = np.arange(S, dtype=np.uint64)
    i = ((s * np.uint64(2654435761) + np.uint64(2135587861)) % np.uint64(2 ** 32)) % np.uint64(m)
    j = ((s * np.uint64(2246822519) + np.uint64(3266489917)) % np.uint64(2 ** 32)) % np.uint64(n)
    return i.astype(np.int64), j.astype(np.int64)


def spread(opA, opB, S=SPREAD_SAMPLES):
    """(L, L_s) of the sampled entries; +inf when a sampled entry has
    |a|.|b| = 0 with a and b both nonzero; 0 when every sample is empty."""
    m, n = opA.shape[0], opB.shape[1]
    i, j = spread_samples(m, n, S)
    a, b = np.abs(opA[i]), np.abs(opB[:, j].T)
    dot = np.sum(a * b, axis=1)
    na, nb = np.sqrt(np.sum(a * a, axis=1)), np.sqrt(np.sum(b * b, axis=1))
    den_s = a.max(axis=1, initial=0.0) * b.sum(axis=1) + b.max(axis=1, initial=0.0) * a.sum(axis=1)
    live = (na > 0) & (nb > 0)
    if not np.any(live):
        return 0.0, 0.0
    with np.errstate(divide="ignore"):
        L = np.max(np.where(live, np.log2(na * nb) - np.log2(dot), -np.inf))
        Ls = np.max(np.where(live, np.log2(den_s) - np.log2(dot), -np.inf))
    return float(L), float(Ls)


def grid_spread(opA, opB, layers, S=SPREAD_SAMPLES):
    """The automatic modulus count of the simulated grid's Ozaki form
    (engine.cpp, choose_ozaki_count): the sampled entries of spread_samples
    over the whole m x n C, sample s taken in k layer l = s mod kp, where
    layers[l] = (k0, kn) is that layer's k range, with a and b over that
    range only.  Returns (L, k_max): the statistic L as in spread (0 when
    every sample is empty) and the longest layer, whose count rule is
    counts_from_spread(L, ., k_max)[1]."""
    m, n = opA.shape[0], opB.shape[1]
    kp = len(layers)
    i, j = spread_samples(m, n, S)
    lay = np.arange(S) % kp
    L = -np.inf
    for l, (k0, kn) in enumerate(layers):
        sel = lay == l
        a = np.abs(opA[i[sel], k0:k0 + kn])
        b = np.abs(opB[k0:k0 + kn][:, j[sel]]).T
        dot = np.sum(a * b, axis=1)
        na, nb = np.sqrt(np.sum(a * a, axis=1)), np.sqrt(np.sum(b * b, axis=1))
        live = (na > 0) & (nb > 0)
        if np.any(live):
            with np.errstate(divide="ignore"):
                L = max(L, float(np.max(np.log2(na[live] * nb[live]) - np.log2(dot[live]))))
    kmax = max((kn for _, kn in layers), default=0)
    return (0.0 if L == -np.inf else L), kmax


def counts_from_spread(L, Ls, k):
    """(slices, moduli) for the spread statistics, as hip_ozaki_counts."""
    N = CRT_MAX
    c = 53.5 + (0.5 * np.log2(k / 1024.0) if k > 1024 else 0.0)
    for cand in range(CRT_MIN, CRT_MAX + 1):
        if crt_T(cand) >= L + c:
            N = cand
            break
    off = 54.0 + (2.33 * np.log2(k / 1024.0) if k > 1024 else 0.0)
    s = 16
    for cand in range(1, 17):
        if 8 * cand >= Ls + off:
            s = cand
            break
    return s, N
