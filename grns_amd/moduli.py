"""Moduli sets in the params.h contract of GRNS (src/params.h:11-39).

The workloads use the k largest primes below 2^31, listed in ascending order like every
set shipped under src/params/ (mixed-radix conversion, src/mrc.cuh:31-42, assumes that
order).  Primality is decided by deterministic Miller-Rabin with the bases that are exact
for all 64-bit integers.
"""
from __future__ import annotations

_MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(v: int) -> bool:
    if v < 2:
        return False
    for p in _MR_BASES:
        if v % p == 0:
            return v == p
    d, s = v - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _MR_BASES:
        x = pow(a, d, v)
        if x in (1, v - 1):
            continue
        for _ in range(s - 1):
            x = x * x % v
            if x == v - 1:
                break
        else:
            return False
    return True


def largest_primes_below(k: int, bound: int = 1 << 31) -> list[int]:
    """The k largest primes strictly below ``bound``, ascending."""
    out = []
    v = bound - 1
    while len(out) < k:
        if is_prime(v):
            out.append(v)
        v -= 1
    return sorted(out)


def product(moduli) -> int:
    m = 1
    for v in moduli:
        m *= v
    return m


def params_header(moduli, threshold: int = 30, rel_error: str = "0.0000001", note: str | None = None) -> str:
    """Text of a params.h for ``moduli`` with the macro names of src/params.h:11-39."""
    n = len(moduli)
    log2m = product(moduli).bit_length() - 1
    vals = ",".join(str(v) for v in moduli)
    note = note or f"{n} largest primes below 2^31, ascending"
    return f"""/*
 *  The set of RNS moduli and other parameters of GRNS
 *  ({note}; floor(log2 M) = {log2m})
 */

#ifndef GRNS_PARAMS_H
#define GRNS_PARAMS_H

#define RNS_MODULI_SIZE ({n})

#define RNS_MODULI_PRODUCT_LOG2 ({log2m})

#define RNS_P2_SCALING_THRESHOLD ({threshold})

#define RNS_EVAL_RELATIVE_ERROR ({rel_error})

#define RNS_MODULI_VALUES {{{vals}}}

constexpr int RNS_MODULI[RNS_MODULI_SIZE] = RNS_MODULI_VALUES;

#define EMPLOY_STD_FMA false

#endif  //GRNS_PARAMS_H
"""


# Sets used by the workloads: 8, 32 and 64 largest primes below 2^31, and the smallest
# shipped set (src/params/moduli-10-69-bit/params.4m.10b.h:16) for exhaustive checks.
SMALL_SET = [3, 5, 7, 11]


def workload_set(n: int) -> list[int]:
    return largest_primes_below(n)
