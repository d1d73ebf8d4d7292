"""Encoding on the CPU: the oracle's restatement of CKKSCore::encode_polynomial
and the product's host-side encoding constants, both against the reference
engine compiled from its sources (oracle/_ref).  No GPU is needed.

Pinned here:
  * the oracle's slot permutation equals CKKSPrecomputed::pre_permutation_;
  * fhe_params_encode_constants equals CKKSPrecomputed's scale_before_tile,
    log_scale_after_tile and r_mult_scale_after_tile at every level of the
    default and bootstrap presets;
  * oracle_encode, fed the reference's constants, gives the reference's
    plaintext limbs bit for bit: real and complex slots, the top level of the
    one-more-level chain (limb 0 zero), middle and bottom levels, and values
    large enough for the two branches of double_to_rns (above 2^59 and above
    2^62 after scaling);
  * a closed form: the encoding of a constant real vector c is the constant
    polynomial llround(c * scale) (one nonzero coefficient).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference

pytestmark = [pytest.mark.reference,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def _oracle_encode(orc, ref, values, level):
    c = ref.encode_constants(level)
    full = ref.chain_count - 2  # full_level
    mid = full - level
    limbs = ref.chain_count - mid
    first = 1 if c["top_of_one_more_level"] else 0
    enc_mid = mid + first
    return orc.encode(values, c["scale_before_tile"], first, enc_mid, limbs,
                      c["r_mult_scale_after_tile"])


@pytest.mark.parametrize("log_n", [13, 14, 16])
def test_pre_permutation_matches_reference(log_n):
    ref = Reference(log_n, 1 if log_n < 16 else 4)
    try:
        n = 1 << log_n
        mine = np.zeros(n, np.uint64)
        Oracle.lib().oracle_encode_pre_permutation(
            n, mine.ctypes.data)
        assert (mine == ref.pre_permutation()).all()
        # a permutation of the codec indices
        assert sorted(mine.tolist()) == list(range(n))
    finally:
        ref.close()


@pytest.mark.parametrize("log_n,alpha,boot", [(13, 1, 0), (16, 4, 0),
                                              (16, 4, 1)])
def test_encode_constants_match_reference(log_n, alpha, boot):
    ref = Reference(log_n, alpha, bootstrap=boot)
    try:
        kw = dict(scale_bits=54 if boot else 40,
                  small_prime_bits=54 if boot else 40)
        full = ref.chain_count - 2
        tops = 0
        for level in range(full + 1):
            got = fhe_amd.params_encode_constants(log_n, alpha, level, **kw)
            want = ref.encode_constants(level)
            assert got["top_of_one_more_level"] == \
                want["top_of_one_more_level"], level
            assert got["scale_before_tile"] == want["scale_before_tile"], level
            assert got["log_scale_after_tile"] == \
                want["log_scale_after_tile"], level
            assert got["ckks_scale"] == want["ckks_scale"], level
            assert (got["r_mult_scale_after_tile"] ==
                    want["r_mult_scale_after_tile"]).all(), level
            tops += got["top_of_one_more_level"]
        # exactly the top level of the one-more-level chain is special
        assert tops == 1
        assert fhe_amd.params_encode_constants(
            log_n, alpha, full, **kw)["top_of_one_more_level"]
    finally:
        ref.close()


def _cases(slots, rng):
    real = rng.uniform(-1, 1, slots)
    cplx = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    short = np.zeros(slots)
    short[:5] = [0.5, -0.25, 3.0, 1e-3, -7.0]
    big = rng.uniform(-1, 1, slots) * 1e6       # > 2^59 after scaling
    huge = rng.uniform(-1, 1, slots) * 3e7      # some >= 2^62 after scaling
    return [("real", real), ("complex", cplx), ("short", short),
            ("zero", np.zeros(slots)), ("big", big), ("huge", huge)]


@pytest.mark.parametrize("log_n,alpha", [(13, 1)])
def test_oracle_encode_matches_reference(log_n, alpha):
    ref = Reference(log_n, alpha)
    orc = Oracle(log_n, ref.primes, alpha)
    try:
        rng = np.random.default_rng(11)
        full = ref.chain_count - 2
        for level in sorted({full, full - 1, full // 2, 1, 0}):
            for name, values in _cases(ref.slot_count, rng):
                mid, want = ref.encode(values, level)
                assert mid == full - level
                got = _oracle_encode(orc, ref, values, level)
                assert got.shape == want.shape, (name, level)
                assert (got == want).all(), (name, level)
    finally:
        ref.close()


def test_oracle_encode_bootstrap_preset_n16():
    ref = Reference(16, 4, bootstrap=1)
    orc = Oracle(16, ref.primes, 4)
    try:
        rng = np.random.default_rng(5)
        full = ref.chain_count - 2
        for level in (full, 3):
            values = rng.uniform(-1, 1, ref.slot_count) + \
                1j * rng.uniform(-1, 1, ref.slot_count)
            _, want = ref.encode(values, level)
            got = _oracle_encode(orc, ref, values, level)
            assert (got == want).all(), level
    finally:
        ref.close()


def test_constant_vector_closed_form():
    """A constant real vector c encodes to the single coefficient
    round(c * Delta) at X^0 (Delta = the level's plaintext scale), zero
    elsewhere: the canonical embedding maps constants to constants."""
    ref = Reference(13, 1)
    orc = Oracle(13, ref.primes, 1)
    try:
        full = ref.chain_count - 2
        level = full - 2
        mid = full - level
        c = ref.encode_constants(level)
        values = np.full(ref.slot_count, 0.75)
        got = _oracle_encode(orc, ref, values, level)
        # every limb: coefficient 0 holds the same integer mod its prime and
        # all others vanish.  The Montgomery entry with r_mult = R 2^s mod q
        # leaves y 2^s mod q, the plain residue of the scaled integer.
        ints = []
        for l in range(got.shape[0]):
            q = int(orc.primes[mid + l])
            assert (got[l, 1:] % q == 0).all(), l
            ints.append((int(got[l, 0]) % q, q))
        # CRT-consistent: the residue is round(0.75 * ckks_scale) on every
        # limb (ckks_scale = scale_before_tile * 2^log_scale_after_tile)
        # llround: half away from zero, computed exactly
        x = Fraction(0.75 * c["scale_before_tile"])
        delta = math.floor(x + Fraction(1, 2)) << c["log_scale_after_tile"]
        for v, q in ints:
            assert v == delta % q
    finally:
        ref.close()
