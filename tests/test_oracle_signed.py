"""The oracle on the reference's signed representatives (CPU): ciphertexts
rescaled from the top of the one-more-level chain carry limbs in (-q, q);
the oracle's transforms, rescale, relinearization, HMult and rotation of such
data equal the compiled reference engine's bit for bit.  This pins the oracle
that the GPU tests on signed inputs compare against."""
import numpy as np
import pytest

from support import Oracle, Reference, ptr

pytestmark = pytest.mark.skipif(not Reference.available(),
                                reason="oracle/_ref not built")


def test_oracle_matches_reference_on_signed_limbs():
    ref = Reference(14, 2, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    orc = Oracle(14, ref.primes, 2)
    top = ref.chain_count - 2
    rng = np.random.default_rng(1)
    handles = []

    def keep(h):
        assert h, lib.ref_last_error()
        handles.append(h)
        return h

    def fresh():
        x = rng.uniform(-1, 1, ref.slot_count)
        return keep(lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, top))

    try:
        s = np.zeros(4, np.int64)
        lib.ref_relin_key_shape(rlk, ptr(s))
        key_mid = int(s[0])
        key = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
        lib.ref_relin_key_get(rlk, ptr(key))
        c0, c1 = fresh(), fresh()
        a = keep(lib.ref_rescale(e, c0))
        b = keep(lib.ref_rescale(e, c1))
        A, B = ref.get(a), ref.get(b)
        assert (A < 0).any() and ref.shape(a)[0] == 1
        # rescale by the 60-bit top prime and the NTT of its signed output
        assert np.array_equal(orc.rescale(ref.get(c0), 0), A)
        na = ref.get(keep(lib.ref_ntt_text(e, a, 0)))
        assert np.array_equal(orc.ntt(A, 1), na)
        # HMult + relin + rescale, and relinearization of the product trio
        want = ref.get(keep(lib.ref_hmult_relin_rescale(e, a, b, rlk)))
        got = orc.hmult(A[None], B[None], key, 1, key_mid)
        assert np.array_equal(got.reshape(want.shape), want)
        trio = ref.get(keep(lib.ref_multiply_ct(e, a, b)))
        relin = ref.get(keep(lib.ref_multiply_relinearize(e, a, b, rlk)))
        assert np.array_equal(orc.relinearize(trio, key, 1, key_mid), relin)
        # rotation of the signed ciphertext
        frk = lib.ref_fixed_rotation_key(e, sk, 3)
        kmid, kd = ref.key_data(frk, 1)
        rot = ref.get(keep(lib.ref_rotate(e, a, frk)))
        assert np.array_equal(orc.rotate(A, kd, 1, kmid, 3), rot)
        lib.ref_free_fixed_rotation_key(frk)
    finally:
        for h in handles:
            lib.ref_free_ct(h)
        lib.ref_free_relin_key(rlk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()
