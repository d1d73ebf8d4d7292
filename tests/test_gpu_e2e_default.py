"""End to end on the device for the reference's default engine (N = 2^14,
one special prime, scale 2^40, 2^13 complex slots): encode -> public-key
encrypt -> multiply + relinearize -> rescale -> decrypt -> decode, all on
the device with no host round trip in between, as the reference's
examples/multiply_vector.cpp:5-38 runs it through its engine (with a
ciphertext-ciphertext product).

Sampling stays on the host (u, e0, e1 are inputs of fhe_encrypt_pk), so the
reference engine is run on the device's own encryptions: every
intermediate (plaintext limbs, product, rescaled product, decrypted limbs)
and the final decoded slots equal the reference's bit for bit, and the
slots are x * y within CKKS noise.
"""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def test_default_engine_end_to_end_matches_reference(gpu_lib):
    ref = Reference(14, 1, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    sk_mid, skd = ref.key_data(sk, 3)
    pk_mid, pkd = ref.key_data(pk, 4)
    key_mid, key = ref.key(rlk)
    ctx = fhe_amd.Context(14, 1, primes=ref.primes)
    handles = []
    arrays = []
    try:
        n, slots = ref.n, ref.slot_count
        assert (n, slots) == (1 << 14, 1 << 13)
        level = ref.chain_count - 2          # the default (top) level
        mid = ctx.info["full_level"] - level
        limbs = ctx.limbs_at(mid)
        rng = np.random.default_rng(20261203)
        x = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
        y = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)

        # device: encode both messages
        pt = ctx.encode(np.stack([x, y]), level)
        arrays.append(pt)
        # device: encrypt with host samples (binary u, rounded Gaussian
        # errors of deviation 3.2, the reference's sampler's shapes)
        u = rng.integers(0, 2, (2, n), dtype=np.int64)
        e0 = np.rint(rng.normal(0, 3.2, (2, n))).astype(np.int64)
        e1 = np.rint(rng.normal(0, 3.2, (2, n))).astype(np.int64)
        dpk = ctx.to_device(pkd)
        du, de0, de1 = ctx.to_device(u), ctx.to_device(e0), ctx.to_device(e1)
        ct = ctx.array((2, 2, limbs, n))
        arrays += [dpk, du, de0, de1, ct]
        ctx.encrypt_pk(pt, dpk, pk_mid, du, de0, de1, ct, mid)
        # device: multiply + relinearize (the engine's prepare_for_multiply
        # rescales the fresh top-level texts first), then rescale
        dk = ctx.to_device(key)
        prod = ctx.array((1, 2, limbs - 1, n))
        arrays += [dk, prod]
        a, b = ct.download()[:1], ct.download()[1:]
        da, db = ctx.to_device(a), ctx.to_device(b)
        arrays += [da, db]
        ctx.multiply(da, level, 0, db, level, 0, prod, key=dk,
                     key_moduli_id=key_mid)
        prod_level = level - 1
        prod_mid = mid + 1
        resc = ctx.array((1, 2, limbs - 2, n))
        arrays.append(resc)
        ctx.rescale(prod, resc, prod_mid)
        # device: the engine's decrypt of a rescaled text (level down after
        # rescale by one level, then decrypt_duo), then decode
        low = ctx.array((1, 2, limbs - 2, n))
        arrays.append(low)
        ctx.level_down(resc, low, prod_level, prod_level - 1, rescaled=True)
        dsk = ctx.to_device(skd[0])
        dec = ctx.array((1, limbs - 2, n))
        arrays += [dsk, dec]
        ctx.decrypt(low, dsk, sk_mid, dec, prod_mid + 1)

        # the reference engine on the device's own encryptions
        _, want_x = ref.encode(x, level)
        _, want_y = ref.encode(y, level)
        assert np.array_equal(pt.download()[0], want_x)
        assert np.array_equal(pt.download()[1], want_y)
        zeros = np.zeros(slots)
        ra = lib.ref_encrypt(e, ptr(zeros), slots, pk, level)
        rb = lib.ref_encrypt(e, ptr(zeros), slots, pk, level)
        handles += [ra, rb]
        ref.set(ra, a[0])
        ref.set(rb, b[0])
        rp = lib.ref_multiply_relinearize(e, ra, rb, rlk)
        handles.append(rp)
        assert np.array_equal(prod.download()[0], ref.get(rp))
        rr = lib.ref_rescale(e, rp)
        handles.append(rr)
        assert np.array_equal(resc.download()[0], ref.get(rr))
        dmid, want_pt = ref.decrypt_limbs(rr, sk)
        assert dmid == prod_mid + 1
        assert np.array_equal(dec.download()[0], want_pt)
        # the decrypted plaintext carries the scale of the level it was
        # brought down to (level_down_ sets the level's standard scale)
        scale = fhe_amd.params_encode_constants(
            14, 1, prod_level - 1, primes=ref.primes)["ckks_scale"]
        got = ctx.decode(dec, dmid, scale, complex_output=True)[0]
        want = ref.decode(want_pt, prod_level - 1, scale, complex_output=True)
        assert np.array_equal(got.real, want.real)
        assert np.array_equal(got.imag, want.imag)
        # the reference engine's own decrypt (real parts) agrees within the
        # 2^-25 absolute tolerance of decrypted slots
        full = np.zeros(slots)
        assert lib.ref_decrypt(e, rr, sk, ptr(full)) == slots
        assert np.abs(got.real - full).max() < 2.0 ** -25
        # the scheme's noise, which varies with the reference's unseeded
        # randomness (the limbs equal the reference's)
        assert np.abs(got - x * y).max() < 1e-4
    finally:
        for h in handles:
            lib.ref_free_ct(h)
        for arr in arrays:
            arr.free()
        ctx.close()
        lib.ref_free_relin_key(rlk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()
