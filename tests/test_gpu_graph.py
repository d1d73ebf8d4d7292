"""HIP graph capture of the bootstrap (fhe_graph_*): a whole
fhe_bootstrap (RotationKey variant, N = 2^16 bootstrap preset, 2^15 slots,
3 stages) recorded once after an eager call and replayed three times, with
eager work between the replays, gives the eager call's limbs every time;
a call whose set-up is missing (the first call of a shape) is refused while
capturing.  Synthetic keys of the real shapes: one buffer serves every key
(the replay's equality does not depend on key values); the bootstrap's
parity with the reference is test_gpu_bootstrap.py's."""
import numpy as np
import pytest

import fhe_amd
from support import uniform_limbs

pytestmark = pytest.mark.gpu


def test_bootstrap_graph_replays_equal_eager(gpu_lib):
    rng = np.random.Generator(np.random.PCG64(41))
    with fhe_amd.Context(16, 4, scale_bits=54, small_prime_bits=54) as ctx:
        n = ctx.n
        full = ctx.info["full_level"]
        key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(key_mid)
        key = ctx.array((kp, kc, n))
        for p in range(kp):
            key.upload_at(uniform_limbs(rng, ctx.primes, key_mid, kc, n, 2),
                          p * kc * n)
        sp, sc, _ = ctx.key_shape(full)
        sparse = ctx.array((sp, sc, n))
        for p in range(sp):
            sparse.upload_at(uniform_limbs(rng, ctx.primes, full, sc, n, 2),
                             p * sc * n)
        rotations = [(key, key_mid, 1 << i) for i in range(15)]
        keys = fhe_amd.BootstrapKeys.make(key, key_mid, key, key_mid,
                                          rotations, sparse, full, key,
                                          key_mid)
        level = 2
        mid = full - level
        ct = ctx.to_device(np.stack([uniform_limbs(rng, ctx.primes, mid,
                                                   ctx.limbs_at(mid), n)
                                     for _ in range(2)])[None])
        out = ctx.array((1, 2, ctx.chain_count - 1, n))
        other = ctx.array((1, 2, ctx.chain_count - 1, n))
        # a first call of a shape prepares diagonals and buffers: refused
        # while capturing
        ctx.graph_begin()
        try:
            with pytest.raises(fhe_amd.FheError):
                ctx.bootstrap(ct, level, 0, keys, 3, out)
        finally:
            g0 = ctx.graph_end()
        ctx.graph_destroy(g0)
        lv, fl = ctx.bootstrap(ct, level, 0, keys, 3, out)
        ctx.synchronize()
        want = out.download()
        ctx.graph_begin()
        try:
            ctx.bootstrap(ct, level, 0, keys, 3, out)
        finally:
            graph = ctx.graph_end()
        try:
            for rep in range(3):
                out.upload(np.zeros(out.shape, np.int64))
                # eager work on the same stream and pools between replays
                ctx.bootstrap(ct, level, 0, keys, 3, other)
                ctx.synchronize()
                ctx.graph_launch(graph)
                ctx.synchronize()
                assert np.array_equal(out.download(), want), rep
            assert np.array_equal(other.download(), want)
        finally:
            ctx.graph_destroy(graph)
        for arr in (key, sparse, ct, out, other):
            arr.free()
