#!/usr/bin/env python3
"""Generate the NTT-friendly prime tables used by every kernel.

Rule (reproduces the constants of the reference, include/yell/params_64.hpp:19-34):
  * P[i]  : the i-th largest prime of the form 2^62 - k*2^21 + 1 (k = 1, 2, ...),
            so every P[i] - 1 is divisible by 2^21 = 2 * kMaxPolyDegree;
  * Pn[i] : low 64 bits of floor(2^128 / P[i])   (Barrett constant, params_64.hpp:29-31);
  * R[i]  : g^((P[i]-1)/2^21) mod P[i], g the smallest generator of Z/P[i]Z
            (a primitive 2^21-th root of unity, params_64.hpp:33-34).
The 32-bit moduli build (include/yell/params_32.hpp:15-34) follows the same
rule with 291 primes 2^30 - k*2^16 + 1, Pn = low 32 bits of floor(2^64 / P),
primitive 2^16-th roots, and kMaxPolyDegree = 2^15.

Writes params_table.inc and params32_table.inc (C initialiser lists) next to
this script.  Usage: python3 gen_params.py
"""
import os
import sys

LOG_MAX_DEGREE = 20              # kMaxPolyDegree = 2^20
TWO_N_MAX = 1 << (LOG_MAX_DEGREE + 1)


def is_prime(n):
    """Deterministic Miller-Rabin for n < 3.3e24 (bases: first 13 primes)."""
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41)
    for q in small:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prime_factors(n):
    out, f = set(), 2
    while f * f <= n:
        while n % f == 0:
            out.add(f)
            n //= f
        f += 1
    if n > 1:
        out.add(n)
    return out


def smallest_generator(p):
    qs = prime_factors(p - 1)
    g = 2
    while True:
        if all(pow(g, (p - 1) // q, p) != 1 for q in qs):
            return g
        g += 1


def generate(count, bits=62, two_n=TWO_N_MAX, word=64):
    primes, k = [], 1
    while len(primes) < count:
        p = (1 << bits) - k * two_n + 1
        if is_prime(p):
            primes.append(p)
        k += 1
    pn = [((1 << (2 * word)) // p) & ((1 << word) - 1) for p in primes]
    roots = [pow(smallest_generator(p), (p - 1) // two_n, p) for p in primes]
    return primes, pn, roots


def write_table(path, prefix, count, primes, pn, roots, header, fmt, log_max_degree):
    # invkMaxPolyDegree (params_64.hpp:38, params_32.hpp:33): 2^log_max_degree inverted mod P
    invdeg = [pow(1 << log_max_degree, -1, p) for p in primes]
    with open(path, "w") as f:
        f.write(header)
        f.write("#define %s_NUM_MODULI %d\n" % (prefix, count))
        for name, vals in ((prefix + "_TABLE_P", primes), (prefix + "_TABLE_PN", pn),
                           (prefix + "_TABLE_ROOT", roots), (prefix + "_TABLE_INVDEG", invdeg)):
            f.write("#define %s { \\\n" % name)
            for i in range(0, len(vals), 4):
                f.write("  " + ", ".join(fmt % v for v in vals[i:i + 4]) + ", \\\n")
            f.write("}\n")


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    primes, pn, roots = generate(1000)
    write_table(os.path.join(here, "params_table.inc"), "YA", 1000, primes, pn, roots,
                "// Generated by gen_params.py: 1000 primes 2^62 - k*2^21 + 1, descending.\n"
                "// Columns: prime, low64(floor(2^128/prime)), primitive 2^21-th root,\n"
                "// 2^20 inverted mod prime.\n",
                "0x%016xULL", LOG_MAX_DEGREE)
    primes, pn, roots = generate(291, bits=30, two_n=1 << 16, word=32)
    write_table(os.path.join(here, "params32_table.inc"), "YA32", 291, primes, pn, roots,
                "// Generated by gen_params.py: 291 primes 2^30 - k*2^16 + 1, descending.\n"
                "// Columns: prime, low32(floor(2^64/prime)), primitive 2^16-th root,\n"
                "// 2^15 inverted mod prime.\n",
                "0x%08xu", 15)


if __name__ == "__main__":
    main()
