"""GPU parity of the bootstrap's EvalMod, CKKSEngine::round (the Chebyshev /
Paterson-Stockmeyer sine, three double-angle steps and x * 28 - modulo(x)),
composed of the device operations, bit-exact against the reference engine
(oracle/_ref) on the N = 2^16 default preset: a fresh encryption at the top
level (whose first rescale leaves signed limbs) and one a few levels
lower."""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def test_round_matches_reference(gpu_lib):
    ref = Reference(16, 4, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    s = np.zeros(4, np.int64)
    lib.ref_relin_key_shape(rlk, ptr(s))
    key = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
    lib.ref_relin_key_get(rlk, ptr(key))
    ctx = fhe_amd.Context(16, 4, primes=ref.primes)
    try:
        dkey = ctx.to_device(key)
        top = ref.chain_count - 2
        rng = np.random.default_rng(16)
        for level in (top, top - 3):
            print(f"round level={level}", flush=True)
            x = rng.uniform(-0.2, 0.2, ref.slot_count)
            ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
            w = lib.ref_round(e, ct, rlk)
            assert w, lib.ref_last_error()
            want = ref.get(w)
            wshape = ref.shape(w)
            d = ctx.to_device(ref.get(ct))
            out = ctx.array(d.shape)
            lv, fl = ctx.round(d, level, dkey, int(s[0]), out)
            assert (lv, fl & 6) == (wshape[5], wshape[4] & 6)
            # the result's limbs are packed at the start of out
            got = out.download().reshape(-1)[:want.size].reshape(want.shape)
            assert np.array_equal(got, want), level
            for h in (ct, w):
                lib.ref_free_ct(h)
            d.free()
            out.free()
        # a batch of two in one call, each equal to the reference's round
        cts = [lib.ref_encrypt(e, ptr(rng.uniform(-0.2, 0.2,
                                                  ref.slot_count)),
                               ref.slot_count, pk, top - 1) for _ in range(2)]
        ws = [lib.ref_round(e, ct, rlk) for ct in cts]
        wants = [ref.get(w) for w in ws]
        d = ctx.to_device(np.stack([ref.get(ct) for ct in cts]))
        out = ctx.array(d.shape)
        ctx.round(d, top - 1, dkey, int(s[0]), out)
        got = out.download().reshape(-1)[:2 * wants[0].size].reshape(
            (2,) + wants[0].shape)
        for i in range(2):
            assert np.array_equal(got[i], wants[i]), i
        for h in cts + ws:
            lib.ref_free_ct(h)
        d.free()
        out.free()
        dkey.free()
    finally:
        lib.ref_free_relin_key(rlk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ctx.close()
        ref.close()
