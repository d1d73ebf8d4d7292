"""(Collected last: the allocation-free entry point's test.)

The allocation-free HMult + relinearize + rescale
(fhe_hmult_relin_rescale_noalloc, workspace in the library's fixed device
scratch array) against the allocating entry point, bit for bit, on two calls
in a row - the second with a smaller input - on the Gold chain (N = 2^16,
alpha = 4); batches beyond the scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import golden, hmult_inputs

pytestmark = pytest.mark.gpu


def noalloc(ctx, a, b, key, key_mid, out, moduli_id):
    lib = fhe_amd.library()
    count = int(np.prod(a.shape[:-3]))
    return lib.fhe_hmult_relin_rescale_noalloc(
        ctx.handle, count, moduli_id, a.ptr, b.ptr, key.ptr, key_mid, out.ptr)


def test_noalloc_equals_allocating_entry_point(gpu_lib):
    case = [c for c in golden("hmult_golden.json")["cases"]
            if c["name"] == "gold_mid13"][0]
    n = 1 << case["log_n"]
    with fhe_amd.Context(case["log_n"], case["special_prime_count"],
                         primes=case["primes"]) as ctx:
        key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(key_mid)
        rng = np.random.Generator(np.random.PCG64(31))
        key = ctx.array((kp, kc, n))
        for p in range(kp):
            key.upload_at(np.stack([rng.integers(0, int(ctx.primes[key_mid + l])
                                                 * 2, n, dtype=np.uint64)
                                    .astype(np.int64) for l in range(kc)]),
                          p * kc * n)
        # first call: one Gold ciphertext pair at level 24 (25 limbs); second
        # call: a smaller input, two pairs at 11 limbs
        for count, mid in ((1, 13), (2, 27)):
            limbs = ctx.limbs_at(mid)
            A = np.stack([np.stack([np.stack(
                [rng.integers(0, int(ctx.primes[mid + l]), n,
                              dtype=np.uint64).astype(np.int64)
                 for l in range(limbs)]) for _ in range(2)])
                for _ in range(count)])
            B = np.stack([np.stack([np.stack(
                [rng.integers(0, int(ctx.primes[mid + l]), n,
                              dtype=np.uint64).astype(np.int64)
                 for l in range(limbs)]) for _ in range(2)])
                for _ in range(count)])
            da, db = ctx.to_device(A), ctx.to_device(B)
            want = ctx.array((count, 2, limbs - 1, n))
            got = ctx.array((count, 2, limbs - 1, n))
            ctx.hmult_relin_rescale(da, db, key, key_mid, want, mid)
            print(f"no-alloc call: count {count} moduli_id {mid}", flush=True)
            assert noalloc(ctx, da, db, key, key_mid, got, mid) == 0, \
                fhe_amd.library().fhe_last_error()
            assert np.array_equal(got.download(), want.download()), (count,
                                                                     mid)
            for arr in (da, db, want, got):
                arr.free()
        # a batch whose workspace exceeds the scratch is refused
        limbs = ctx.limbs_at(13)
        da = ctx.array((4, 2, limbs, n))
        out = ctx.array((4, 2, limbs - 1, n))
        rc = noalloc(ctx, da, da, key, key_mid, out, 13)
        assert rc == 1
        assert b"scratch" in fhe_amd.library().fhe_last_error()
        for arr in (da, out, key):
            arr.free()
