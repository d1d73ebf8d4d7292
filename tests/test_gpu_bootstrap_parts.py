"""GPU parity of the bootstrapping building blocks, bit-exact against the
reference engine (oracle/_ref):

  * multiply_imaginary (CKKSEngine::multiply_imaginary) at several levels and
    integer values, N = 2^14 and 2^16;
  * mod_raise (CKKSEngine::mod_raise_) on the N = 2^16 bootstrap preset with
    the reference's small bootstrap key, batched: a level-0 encryption and a
    rescaled ciphertext leveled down to level 0.
"""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


@pytest.mark.parametrize("log_n,alpha", [(14, 2), (16, 4)])
def test_multiply_imaginary_matches_reference(log_n, alpha, gpu_lib):
    ref = Reference(log_n, alpha, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
    try:
        full = ref.chain_count - 2
        rng = np.random.default_rng(log_n)
        levels = [full, 3, 0] if log_n < 16 else [full, 0]
        for level in levels:
            for value in (1, -1, 3):
                print(f"multiply_imaginary log_n={log_n} level={level} "
                      f"value={value}", flush=True)
                x = rng.uniform(-1, 1, ref.slot_count)
                ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
                data = ref.get(ct)
                mid = full - level
                want_ct = lib.ref_multiply_imaginary(e, ct, value)
                want = ref.get(want_ct)
                d = ctx.to_device(data[None])
                out = ctx.array((1,) + data.shape)
                ctx.multiply_imaginary(d, mid, value, out)
                got = out.download()[0]
                assert np.array_equal(got, want), (level, value)
                for h in (ct, want_ct):
                    lib.ref_free_ct(h)
                d.free()
                out.free()
    finally:
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ctx.close()
        ref.close()


def test_mod_raise_matches_reference(gpu_lib):
    ref = Reference(16, 4, bootstrap=1, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    sbk = lib.ref_small_bootstrap_key(e, sk)
    assert sbk, lib.ref_last_error()
    ctx = fhe_amd.Context(16, 4, scale_bits=54, small_prime_bits=54,
                          primes=ref.primes)
    try:
        sparse_mid, sparse = ref.key_data(sbk, 5)
        dense_mid, dense = ref.key_data(sbk, 6)
        dsparse, ddense = ctx.to_device(sparse), ctx.to_device(dense)
        full = ref.chain_count - 2
        rng = np.random.default_rng(3)
        # level-0 ciphertexts as CKKSEngine::mod_raise hands them over: an
        # encryption at level 0 and a rescaled ciphertext leveled down to 0
        x = rng.uniform(-1, 1, ref.slot_count)
        ct0 = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, 0)
        ct2 = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, 2)
        ct2r = lib.ref_rescale(e, ct2)
        ct2d = lib.ref_level_down(e, ct2r, 0)
        # a batch of both through one call
        cts = [ct0, ct2d]
        wants = []
        for ct in cts:
            assert ref.shape(ct)[0] == full, ref.shape(ct)
            want_ct = lib.ref_mod_raise(e, ct, sbk)
            assert want_ct, lib.ref_last_error()
            wants.append(ref.get(want_ct))
            lib.ref_free_ct(want_ct)
        d = ctx.to_device(np.stack([ref.get(ct) for ct in cts]))
        out = ctx.array((len(cts),) + wants[0].shape)
        ctx.mod_raise(d, dsparse, sparse_mid, ddense, dense_mid, out)
        got = out.download()
        for i in range(len(cts)):
            assert np.array_equal(got[i], wants[i]), i
        d.free()
        out.free()
        for h in (ct0, ct2, ct2r, ct2d):
            lib.ref_free_ct(h)
        dsparse.free()
        ddense.free()
    finally:
        lib.ref_free_small_bootstrap_key(sbk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ctx.close()
        ref.close()


@pytest.mark.parametrize("level,flags", [(-1, 0), (10 ** 6, 0), (2, 8),
                                         (2, 1 << 30)])
def test_bootstrap_entry_points_refuse_bad_forms(level, flags, gpu_lib):
    """A level off the chain or unknown flag bits are refused before any
    device work (the key pointers are never dereferenced: the check comes
    first)."""
    with fhe_amd.Context(14, 2) as ctx:
        n, mid = ctx.n, 2
        d = ctx.array((1, 2, ctx.limbs_at(mid), n))
        key = ctx.array((1, 1, n))
        keys = fhe_amd.BootstrapKeys.make(key, 0, key, 0, [(key, 0, 1)], key,
                                          0, key, 0)
        out = ctx.array((1, 2, ctx.limbs_at(0), n))
        with pytest.raises(fhe_amd.FheError, match="out of range|unknown"):
            ctx.bootstrap(d, level, flags, keys, 3, out)
        with pytest.raises(fhe_amd.FheError, match="out of range|unknown"):
            ctx.coeff_to_slot(d, level, flags, keys, 3, out, out)
        with pytest.raises(fhe_amd.FheError, match="out of range|unknown"):
            ctx.slot_to_coeff(d, d, level, flags, keys, 3, out)
        for arr in (d, key, out):
            arr.free()
