"""Decoding on the CPU: the oracle's restatement of CKKSCore's decoder
(rescale loop, BigInt CRT lift, skewer, inverse FFT, slot gather) against
the reference engine compiled from its sources (oracle/_ref), bit for bit.
No GPU is needed.

Pinned here:
  * plaintexts from the reference's encoder at the top level of the
    one-more-level chain, middle and bottom levels, real and complex, with
    the encoder's own scale (at the bottom level the decoder rescales to a
    single limb and takes the make_signed branch);
  * plaintexts from the reference's decryption of its encryptions (noisy,
    every coefficient nonzero; the decoder rescales where ckks_scale / q
    exceeds the base scale),
    whose decode must also equal the reference's decrypt() output;
  * values large enough for multiword CRT words (2^60 scale products);
  * the reference's refusals: a one-limb plaintext and a scale above the
    modulus product;
  * closed form: decode(encode(v)) returns v to within the encoding noise.
"""
import numpy as np
import pytest

from support import Oracle, Reference, ptr

pytestmark = [pytest.mark.reference,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def _check(orc, ref, limbs, level, scale, base):
    mid = ref.chain_count - 2 - level
    real = None
    for cplx in (False, True):
        want = ref.decode(limbs, level, scale, cplx)
        got = orc.decode(limbs, mid, scale, base, cplx)
        assert np.array_equal(got.view(np.float64), want.view(np.float64)), \
            (level, cplx)
        if not cplx:
            real = want
    return real


@pytest.mark.parametrize("log_n,alpha", [(13, 1)])
def test_oracle_decode_encoded(log_n, alpha):
    ref = Reference(log_n, alpha)
    orc = Oracle(log_n, ref.primes, alpha)
    base = 2.0 ** 40
    try:
        rng = np.random.default_rng(21)
        full = ref.chain_count - 2
        for level in sorted({full, full - 1, full // 2, 1, 0}):
            c = ref.encode_constants(level)
            for values in (rng.uniform(-1, 1, ref.slot_count),
                           rng.uniform(-1, 1, ref.slot_count) +
                           1j * rng.uniform(-1, 1, ref.slot_count),
                           rng.uniform(-1, 1, ref.slot_count) * 3e5):
                _, limbs = ref.encode(values, level)
                dec = _check(orc, ref, limbs, level, c["ckks_scale"], base)
                err = np.abs(dec - values.real).max()
                assert err < 1e-3 * max(1, np.abs(values).max()), (level, err)
    finally:
        ref.close()


def test_oracle_decode_decrypted():
    ref = Reference(14, 2)
    orc = Oracle(14, ref.primes, 2)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    try:
        rng = np.random.default_rng(4)
        full = ref.chain_count - 2
        for level in (full, full - 2, 1):
            x = rng.uniform(-1, 1, ref.slot_count)
            ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
            scale = ref.ct_scale(ct)
            mid, limbs = ref.decrypt_limbs(ct, sk)
            assert mid == full - level
            dec = _check(orc, ref, limbs, level, scale, 2.0 ** 40)
            direct = np.zeros(ref.slot_count)
            lib.ref_decrypt(e, ct, sk, ptr(direct))
            assert np.array_equal(dec, direct)
            lib.ref_free_ct(ct)
    finally:
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()


def test_oracle_decode_decrypted_n16():
    """N = 2^16: the twiddle and skewer tables of the reference's build come
    from glibc's sincos; separate sin/cos calls differ in the last bit for
    about a hundred angles at this size, which this case detects."""
    ref = Reference(16, 4)
    orc = Oracle(16, ref.primes, 4)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    try:
        rng = np.random.default_rng(16)
        full = ref.chain_count - 2
        for level in (full, 5):
            x = rng.uniform(-1, 1, ref.slot_count)
            ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
            scale = ref.ct_scale(ct)
            mid, limbs = ref.decrypt_limbs(ct, sk)
            direct = np.zeros(ref.slot_count)
            lib.ref_decrypt(e, ct, sk, ptr(direct))
            got = orc.decode(limbs, mid, scale, 2.0 ** 40)
            assert np.array_equal(got, direct), level
            lib.ref_free_ct(ct)
    finally:
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()


def test_oracle_decode_refusals():
    ref = Reference(13, 1)
    orc = Oracle(13, ref.primes, 1)
    try:
        n = 1 << 13
        one_limb = np.zeros((1, n), np.int64)
        with pytest.raises(ValueError):
            orc.decode(one_limb, ref.chain_count - 1, 2.0 ** 40, 2.0 ** 40)
        two = np.zeros((2, n), np.int64)
        with pytest.raises(ValueError):
            orc.decode(two, ref.chain_count - 2, np.longdouble(2.0) ** 200,
                       2.0 ** 40)
    finally:
        ref.close()
