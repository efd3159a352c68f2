"""Shows that the bit pins of tests/test_oracle_bitpins.py fail on plausible mistakes in the
restatement.  Each mutant is the oracle built with one planted mistake (-DORC_MUTANT=k, the list
is in oracle/grns_oracle.c); the same checks that pass on the real oracle must report mismatches
on it.

Mutant 6 (BoundNear1 with a fused multiply-add in place of x*y+z) is the exception, and the test
says so: for c = 0 both forms give ETA; for 0 < |c| < 2^-917 the product PHI1*|c| is below ETA/2
and both give ETA; above that ETA is less than a quarter of the spacing of the doubles near
PHI1*|c|, so adding it changes the rounding of PHI1*|c| only at a tie, which needs the exact
product to be a midpoint -- and PHI1*|c| = 2^-53 (1 + 2^-52) |c| is a midpoint only for
significands that the +ETA then breaks the same way in both forms.  The two builds are therefore
the same function, and the test checks that they agree bit for bit on the whole pin workload.
"""
from __future__ import annotations

import pytest

from oracle.oracle import Oracle, build_mutant, lib
from tests.bitmodel import Model
from tests.test_oracle_bitpins import (PIN_SETS, check_closed_forms, check_er_exact, check_model, check_mpint_eval,
                                       check_zero_sums)

MUTANTS = {
    1: "tree without the top-level zero pad",
    2: "lower/upper reciprocals swapped",
    3: "refinement by reciprocal products",
    4: "refinement factor without the +1",
    5: "integer part dropped with a directed subtraction",
    7: "sequential summation",
    8: "er_add without directed rounding",
    9: "mpint_mul with the UNIT bounds swapped",
    10: "no zero test before the sums",
    11: "ambiguity branch sets INV_UNIT.low",
    12: "refinement shifts only the upper bound",
    13: "refinement lower bound from the unrefined residues",
}

SETS_USED = ("p8", "p32", "params.h")


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("mutants"))


def _failures(path):
    """Names of the pin checks that report mismatches for the library at `path`."""
    L = lib(path)
    failed = set()
    if check_zero_sums(L):
        failed.add("zero_sums")
    if check_er_exact(L, count=150):
        failed.add("er_exact")
    for name in SETS_USED:
        mod = PIN_SETS[name]
        orc, model = Oracle(mod, lib_path=path), Model(mod)
        if check_closed_forms(orc, model):
            failed.add("closed_forms")
        bad, _, _ = check_model(orc, model, count=200)
        # the value 0 is not in check_model's sample of every set: evaluate it explicitly
        iv = orc.eval1([0] * model.n)
        if bad or (iv["low"]["frac"], iv["upp"]["frac"], iv["low"]["exp"], iv["upp"]["exp"]) != (0, 0, 0, 0):
            failed.add("model")
        if check_mpint_eval(orc, model, count=80):
            failed.add("mpint_eval")
        orc.close()
    return failed


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pins_fail_on_mutant(k, mutant_dir):
    failed = _failures(build_mutant(k, mutant_dir))
    assert failed, "no pin detects mutant %d (%s)" % (k, MUTANTS[k])


def test_fused_fmac_is_the_same_function(mutant_dir):
    assert not _failures(build_mutant(6, mutant_dir))
