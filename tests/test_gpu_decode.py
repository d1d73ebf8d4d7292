"""GPU parity of decoding, bit-exact against the reference engine
(oracle/_ref) and the oracle:

  * decryption then decoding on the device of reference encryptions equals
    the reference's decrypt() values, at levels with and without the
    decoder's rescale, batched;
  * device encodings decoded on the device equal the reference's decode of
    the same limbs, real and complex, at the top level of the one-more-level
    chain, a middle level and level 0 (rescale to one limb, make_signed);
  * the N = 2^16 default preset at the benchmark's top level;
  * the reference's refusals (one limb, scale above the modulus product).
"""
import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


@pytest.mark.parametrize("log_n,alpha", [(14, 2), (16, 4)])
def test_decrypt_decode_matches_reference(log_n, alpha, gpu_lib):
    ref = Reference(log_n, alpha, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    sk_mid, skd = ref.key_data(sk, 3)
    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
    try:
        dsk = ctx.to_device(skd[0])
        full = ref.chain_count - 2
        levels = [full, full - 1, full - 2, 1] if log_n < 16 else [full, 5]
        rng = np.random.default_rng(log_n)
        for level in levels:
            print(f"decrypt+decode log_n={log_n} level={level}", flush=True)
            cts, want, scales = [], [], []
            for _ in range(3):
                x = rng.uniform(-1, 1, ref.slot_count)
                ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
                cts.append(ref.get(ct))
                scales.append(ref.ct_scale(ct))
                out = np.zeros(ref.slot_count)
                lib.ref_decrypt(e, ct, sk, ptr(out))
                want.append(out)
                lib.ref_free_ct(ct)
            assert len(set(scales)) == 1
            mid = full - level
            dct = ctx.to_device(np.stack(cts))
            pt = ctx.array((3, ref.chain_count - mid, ctx.n))
            ctx.decrypt(dct, dsk, sk_mid, pt, mid)
            got = ctx.decode(pt, mid, scales[0])
            for i in range(3):
                assert np.array_equal(got[i], want[i]), (level, i)
            dct.free()
            pt.free()
        dsk.free()
    finally:
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ctx.close()
        ref.close()


def test_encode_decode_matches_reference(gpu_lib):
    ref = Reference(13, 1)
    ctx = fhe_amd.Context(13, 1, primes=ref.primes)
    orc = Oracle(13, ref.primes, 1)
    try:
        rng = np.random.default_rng(2)
        full = ref.chain_count - 2
        for level in sorted({full, full // 2, 0}):
            c = fhe_amd.params_encode_constants(13, 1, level)
            batch = rng.uniform(-1, 1, (2, ref.slot_count)) + \
                1j * rng.uniform(-1, 1, (2, ref.slot_count))
            pt = ctx.encode(batch, level)
            limbs = pt.download()
            mid = full - level
            for cplx in (False, True):
                print(f"decode level={level} complex={cplx}", flush=True)
                got = ctx.decode(pt, mid, c["ckks_scale"], cplx)
                for i in range(2):
                    want = ref.decode(limbs[i], level, c["ckks_scale"], cplx)
                    assert np.array_equal(got[i], want), (level, cplx, i)
                    orc_out = orc.decode(limbs[i], mid, c["ckks_scale"],
                                         2.0 ** 40, cplx)
                    assert np.array_equal(got[i], orc_out), (level, cplx, i)
            pt.free()
    finally:
        ctx.close()
        ref.close()


def test_decode_refusals(gpu_lib):
    ctx = fhe_amd.Context(13, 1)
    try:
        n = ctx.n
        one = ctx.array((1, 1, n))
        with pytest.raises(fhe_amd.FheError):
            ctx.decode(one, ctx.chain_count - 1, 2.0 ** 40)
        one.free()
        two = ctx.array((1, 2, n))
        with pytest.raises(fhe_amd.FheError):
            ctx.decode(two, ctx.chain_count - 2, np.longdouble(2.0) ** 200)
        two.free()
    finally:
        ctx.close()
