"""The samplers restated over numpy's Philox -- TEST INFRASTRUCTURE ONLY.

Only tests/, the package's smoke check and the bench's CPU baseline may use
this module, and only as a checker.

Random stream: numpy.random.Philox, numpy's own implementation of
Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC 2011), with key = seed.  Block b
of the stream of `seed` is the generator's output for counter (b, 0, 0, 0);
numpy increments its counter before it computes a block, so a generator made
with counter b - 1 (mod 2^256) yields blocks b, b+1, ...  The reservoir block
of step k is counter (k, 1, 0, 0) = k + 2^64.

Mappings, from the reference:
  uniform   include/yell/details/poly.hpp:200-228 -- residue cm takes the next
            `degree` words of the byte stream (64-bit words, or 32-bit words in
            the 32-bit build), masks each to floor(log2 p) + 1 bits, subtracts
            p once if >= p.
  gaussian  include/yell/details/poly.hpp:236-260 -- one signed sample per
            coefficient, times the amplifier, shared by all residues,
            r < 0 ? p + r : r.  The sample is drawn from a cumulative table
            (given by the caller; `cdt_exact` computes the table's exact values)
            with stream word 2i and the sign from the top bit of word 2i+1.
  zo        include/yell/details/poly.hpp:267-279 -- byte i of the stream is r,
            r <= rho ? ((r & 2) ? p-1 : 1) : 0.
  hwt       include/yell/details/poly.hpp:287-330 -- reservoir over slots
            0..hwt-1; step k draws w by the reference's rejection rule
            (accept w <= floor((2^64-1)/k) * k; here the first accepted of the
            four words of step k's block, else the fourth), pos = w mod k,
            slot pos takes k if pos < hwt; sorted hits, the j-th gets +1 if
            stream word j has bit 1 set, else p-1.

Pinned by: numpy's Philox itself (an implementation independent of the
library's), checked against the Random123 known-answer vector for counter 0,
key 0; the Gaussian table against exact values (mpmath, 60 digits); the
mappings by closed-form properties in tests/test_samplers.py (ranges, the
residue-shared pattern, the exact Hamming weight, the moments of each
distribution).
"""
import numpy as np

M64 = (1 << 64) - 1


def stream_words(seed, nwords, start_block=0):
    """64-bit words of blocks start_block, start_block+1, ... of `seed`."""
    nblocks = (nwords + 3) // 4
    g = np.random.Philox(key=int(seed), counter=(start_block - 1) % (1 << 256))
    return g.random_raw(4 * nblocks)[:nwords].astype(np.uint64)


def reservoir_blocks(seed, k0, k1):
    """Blocks (k, 1, 0, 0) for k in [k0, k1), shape (k1 - k0, 4)."""
    if k1 <= k0:
        return np.zeros((0, 4), dtype=np.uint64)
    g = np.random.Philox(key=int(seed), counter=(k0 + (1 << 64) - 1))
    return g.random_raw(4 * (k1 - k0)).astype(np.uint64).reshape(-1, 4)


def _dtype(word):
    return np.uint64 if word == 64 else np.uint32


def uniform(primes, degree, seed, word=64):
    L = len(primes)
    if word == 64:
        w = stream_words(seed, L * degree)
    else:
        w64 = stream_words(seed, (L * degree + 1) // 2)
        w = np.empty(2 * w64.size, dtype=np.uint64)
        w[0::2] = w64 & np.uint64(0xFFFFFFFF)
        w[1::2] = w64 >> np.uint64(32)
        w = w[:L * degree]
    out = np.empty(L * degree, dtype=_dtype(word))
    for cm, p in enumerate(primes):
        p = int(p)
        mask = np.uint64((1 << p.bit_length()) - 1)
        v = w[cm * degree:(cm + 1) * degree] & mask
        v = np.where(v >= np.uint64(p), v - np.uint64(p), v)
        out[cm * degree:(cm + 1) * degree] = v.astype(out.dtype)
    return out


def cdt_exact(sigma):
    """Exact cumulative table: floor(P(|X| <= k) * 2^64) for k < tail, 2^64-1 at
    tail = ceil(13 sigma), P(X = x) proportional to exp(-x^2/(2 sigma^2)) on
    |x| <= tail.  Python ints."""
    import mpmath
    mpmath.mp.dps = 60
    tail = int(mpmath.ceil(13 * mpmath.mpf(sigma)))
    w = [mpmath.exp(-mpmath.mpf(k * k) / (2 * mpmath.mpf(sigma) ** 2)) * (2 if k else 1) for k in range(tail + 1)]
    tot = mpmath.fsum(w)
    out, acc = [], mpmath.mpf(0)
    for k in range(tail + 1):
        acc += w[k] / tot
        v = int(mpmath.floor(acc * mpmath.mpf(2) ** 64))
        out.append(M64 if k == tail or v >= M64 else v)
    return out


def gaussian_signed(degree, cdf, seed):
    """The signed samples (int64) of coefficients 0..degree-1 for table `cdf`."""
    cdf = np.asarray(cdf, dtype=np.uint64)
    tail = cdf.size - 1
    w = stream_words(seed, 2 * degree)
    u, s = w[0::2], w[1::2]
    k = np.minimum(np.searchsorted(cdf, u, side="right"), tail).astype(np.int64)
    neg = (k != 0) & ((s >> np.uint64(63)) == 1)
    return np.where(neg, -k, k)


def gaussian(primes, degree, cdf, seed, amplifier=1, word=64):
    """details/poly.hpp:238-257: r *= amplifier in params::signed_type (int64_t,
    or int32_t in the 32-bit build, params_32.hpp), so the product wraps
    modulo 2^word; then r < 0 ? p + r : r modulo 2^word."""
    r = gaussian_signed(degree, cdf, seed)
    a = r.astype(np.uint64) * np.uint64(amplifier)  # wrapping 64-bit product
    if word == 32:
        a = a & np.uint64(0xFFFFFFFF)               # the int32_t value's bits
    neg = a >= np.uint64(1 << (word - 1))
    out = np.empty(len(primes) * degree, dtype=_dtype(word))
    for cm, p in enumerate(primes):
        v = np.where(neg, a + np.uint64(int(p)), a)
        if word == 32:
            v = v & np.uint64(0xFFFFFFFF)
        out[cm * degree:(cm + 1) * degree] = v.astype(out.dtype)
    return out


def zo(primes, degree, rho, seed, word=64):
    w = stream_words(seed, (degree + 7) // 8)
    r = w.view(np.uint8)[:degree].astype(np.int64)  # little-endian bytes
    out = np.empty(len(primes) * degree, dtype=_dtype(word))
    for cm, p in enumerate(primes):
        p = int(p)
        v = np.where(r <= rho, np.where(r & 2, p - 1, 1), 0)
        out[cm * degree:(cm + 1) * degree] = v.astype(out.dtype)
    return out


def hwt_positions(degree, hwt, seed):
    """Sorted hit positions of the reservoir."""
    hit = list(range(hwt))
    blocks = reservoir_blocks(seed, hwt, degree)
    for i, k in enumerate(range(hwt, degree)):
        lim = (M64 // k) * k
        row = [int(x) for x in blocks[i]]
        w = row[3]
        for c in (2, 1, 0):
            if row[c] <= lim:
                w = row[c]
        pos = w % k
        if pos < hwt:
            hit[pos] = k
    return sorted(hit)


def hwt(primes, degree, weight, seed, word=64):
    pos = np.array(hwt_positions(degree, weight, seed), dtype=np.int64)
    plus = (stream_words(seed, weight) & np.uint64(2)) != 0
    out = np.zeros(len(primes) * degree, dtype=_dtype(word))
    for cm, p in enumerate(primes):
        out[cm * degree + pos] = np.where(plus, 1, int(p) - 1).astype(out.dtype)
    return out
