"""GPU parity of key switching over batches at uneven levels (N = 2^16,
alpha = 8, the workload of 32 ciphertexts drawn from 6..24 limbs with
relinearization and rotations by random slot offsets):

  * real keys from the compiled reference (oracle/_ref): ciphertexts at
    levels drawn from 6..24 limbs, relinearized in one uneven batch and
    rotated by random offsets with the RotationKey, equal the reference's
    relinearize and rotate(ct, RotationKey, delta) limb for limb, and
    decrypt to the rotated product;
  * the full 32-ciphertext batch (synthetic keys and inputs of the real
    shapes): every output equals the ciphertext's own one-ciphertext calls
    (fhe_relinearize, fhe_rotate per fixed key), themselves bit-exact with
    the reference (test_gpu_hmult.py, test_gpu_rotate.py);
  * a HIP graph of the whole step replays to the eager result, three times
    with eager work between the replays; autotuned schedules give the same
    limbs;
  * invalid batches are refused before any launch.
"""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr, uniform_limbs

pytestmark = pytest.mark.gpu

LOG_N, ALPHA = 16, 8


def rotation_bits(delta, slots):
    return [1 << b for b in range(16) if (delta % slots) & (1 << b)]


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built")
def test_uneven_batch_matches_reference(gpu_lib):
    ref = Reference(LOG_N, ALPHA, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    rk = lib.ref_rotation_key(e, sk)
    ctx = fhe_amd.Context(LOG_N, ALPHA, primes=ref.primes)
    handles, arrays = [], []
    try:
        slots = ref.slot_count
        rng = np.random.default_rng(20261203)
        # encryption levels 5..23: the products' trios have 6..24 limbs
        levels = [5, 23, 12, 12, 7, 18, 23, 9, 15, 5, 20, 11]
        deltas = [int(v) for v in rng.integers(1, slots, len(levels))]
        deltas[3] = slots + 3        # normalized to 3
        deltas[5] = -7               # normalized to slots - 7
        key_mid, key = ref.key(rlk)
        dkey = ctx.to_device(key)
        arrays.append(dkey)
        rot_keys, rot_mids, rot_deltas = [], [], []
        for i in range(lib.ref_rotation_key_count(rk)):
            frk = lib.ref_rotation_key_fixed(rk, i)
            mid, data = ref.key_data(frk, 1)
            rot_keys.append(ctx.to_device(data))
            rot_mids.append(mid)
            rot_deltas.append(lib.ref_fixed_rotation_key_delta(frk))
        arrays += rot_keys
        msgs, trios, mids, want_relin, want_rot = [], [], [], [], []
        for level, delta in zip(levels, deltas):
            x = rng.uniform(-1, 1, slots)
            a = lib.ref_encrypt(e, ptr(x), slots, pk, level)
            b = lib.ref_encrypt(e, ptr(x[::-1].copy()), slots, pk, level)
            tr = lib.ref_multiply_ct(e, a, b)
            d = lib.ref_relinearize(e, tr, rlk)
            r = lib.ref_rotate_rotation_key(e, d, rk, delta)
            handles += [a, b, tr, d, r]
            mids.append(ref.shape(tr)[0])
            trios.append(ctx.to_device(ref.get(tr)))
            want_relin.append(ref.get(d))
            want_rot.append((r, ref.get(r)))
            msgs.append(x * x[::-1])
        arrays += trios
        assert sorted(set(ctx.limbs_at(m) for m in mids)) == sorted(
            set(lv + 1 for lv in levels))
        relin = [ctx.array((2, ctx.limbs_at(m), 1 << LOG_N)) for m in mids]
        rot = [ctx.array((2, ctx.limbs_at(m), 1 << LOG_N)) for m in mids]
        arrays += relin + rot
        ctx.relinearize_uneven(trios, mids, dkey, key_mid, relin)
        ctx.rotate_rotation_key_uneven(relin, mids, rot_keys, rot_mids,
                                       rot_deltas, deltas, rot)
        for c in range(len(levels)):
            assert np.array_equal(relin[c].download(), want_relin[c]), c
            got = rot[c].download()
            assert np.array_equal(got, want_rot[c][1]), c
            ref.set(want_rot[c][0], got)
            dec = np.zeros(slots)
            lib.ref_decrypt(e, want_rot[c][0], sk, ptr(dec))
            assert np.abs(dec - np.roll(msgs[c], deltas[c])).max() < 1e-4, c
    finally:
        for h in handles:
            lib.ref_free_ct(h)
        for arr in arrays:
            arr.free()
        ctx.close()
        lib.ref_free_rotation_key(rk)
        lib.ref_free_relin_key(rlk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()


class Batch:
    """32 ciphertexts at limb counts drawn from 6..24 with synthetic keys of
    the real shapes (the work does not depend on key values)."""

    def __init__(self, ctx, count=32, seed=5):
        rng = np.random.Generator(np.random.PCG64(seed))
        n = 1 << LOG_N
        self.n = n
        self.full = ctx.info["chain_count"]
        limbs = rng.integers(6, 25, count)
        self.mids = [int(self.full - l) for l in limbs]
        self.key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(self.key_mid)

        def key():
            d = ctx.array((kp, kc, n))
            for p in range(kp):
                d.upload_at(uniform_limbs(rng, ctx.primes, self.key_mid, kc,
                                          n, 2), p * kc * n)
            return d

        self.rlk = key()
        self.rot_deltas = [1 << b for b in range(LOG_N - 1)]
        self.rot_keys = [key() for _ in self.rot_deltas]
        self.deltas = [int(v) for v in rng.integers(1, n // 2, count)]
        self.trios = [ctx.to_device(np.stack(
            [uniform_limbs(rng, ctx.primes, m, ctx.limbs_at(m), n)
             for _ in range(3)])) for m in self.mids]
        self.relin = [ctx.array((2, ctx.limbs_at(m), n)) for m in self.mids]
        self.rot = [ctx.array((2, ctx.limbs_at(m), n)) for m in self.mids]

    def step(self, ctx):
        ctx.relinearize_uneven(self.trios, self.mids, self.rlk, self.key_mid,
                               self.relin)
        ctx.rotate_rotation_key_uneven(self.relin, self.mids, self.rot_keys,
                                       [self.key_mid] * len(self.rot_keys),
                                       self.rot_deltas, self.deltas, self.rot)

    def free(self):
        for arr in ([self.rlk] + self.rot_keys + self.trios + self.relin +
                    self.rot):
            arr.free()


@pytest.fixture(scope="module")
def batch32():
    ctx = fhe_amd.Context(LOG_N, ALPHA)
    b = Batch(ctx)
    yield ctx, b
    b.free()
    ctx.close()


def single_path(ctx, b, c):
    """Ciphertext c through the one-moduli_id entry points."""
    m, n = b.mids[c], b.n
    limbs = ctx.limbs_at(m)
    one = ctx.to_device(b.trios[c].download()[None])
    out = ctx.array((1, 2, limbs, n))
    ctx.relinearize(one, b.rlk, b.key_mid, out, m)
    relin = out.download()[0]
    for bit in rotation_bits(b.deltas[c], n // 2):
        k = b.rot_deltas.index(bit)
        ctx.rotate(out, b.rot_keys[k], b.key_mid, bit, out, m)
    rot = out.download()[0]
    one.free()
    out.free()
    return relin, rot


def test_uneven_batch32_equals_single_calls(batch32, gpu_lib):
    ctx, b = batch32
    assert len(set(b.mids)) > 10
    b.step(ctx)
    for c in range(len(b.mids)):
        relin, rot = single_path(ctx, b, c)
        assert np.array_equal(b.relin[c].download(), relin), c
        assert np.array_equal(b.rot[c].download(), rot), c


def test_uneven_graph_replays_equal_eager(batch32, gpu_lib):
    ctx, b = batch32
    b.step(ctx)
    ctx.synchronize()
    want = [(r.download(), o.download()) for r, o in zip(b.relin, b.rot)]
    ctx.graph_begin()
    try:
        b.step(ctx)
    finally:
        graph = ctx.graph_end()
    try:
        for rep in range(3):
            # eager work between the replays: clear the outputs
            for arr in b.relin + b.rot:
                arr.upload(np.zeros(arr.shape, np.int64))
            ctx.synchronize()
            ctx.graph_launch(graph)
            ctx.synchronize()
            for c, (r, o) in enumerate(want):
                assert np.array_equal(b.relin[c].download(), r), (rep, c)
                assert np.array_equal(b.rot[c].download(), o), (rep, c)
    finally:
        ctx.graph_destroy(graph)


def test_uneven_autotuned_schedule_same_limbs(gpu_lib):
    ctx = fhe_amd.Context(LOG_N, ALPHA)
    b = Batch(ctx, count=6, seed=9)
    try:
        b.step(ctx)
        want = [(r.download(), o.download()) for r, o in zip(b.relin, b.rot)]
        ctx2 = fhe_amd.Context(LOG_N, ALPHA)
        ctx2.uneven_autotune(True)
        b2 = Batch(ctx2, count=6, seed=9)
        try:
            b2.step(ctx2)
            sched = ctx2.uneven_schedule(True, b2.mids, b2.key_mid)
            assert len(sched) == 8
            for c, (r, o) in enumerate(want):
                assert np.array_equal(b2.relin[c].download(), r), c
                assert np.array_equal(b2.rot[c].download(), o), c
        finally:
            b2.free()
            ctx2.close()
    finally:
        b.free()
        ctx.close()


def test_uneven_invalid_batches_are_rejected(gpu_lib):
    ctx = fhe_amd.Context(LOG_N, ALPHA)
    n = 1 << LOG_N
    key_mid = ctx.key_moduli_id()
    kp, kc, _ = ctx.key_shape(key_mid)
    key = ctx.array((kp, kc, n))
    tr = ctx.array((3, 6, n))
    out = ctx.array((2, 6, n))
    full = ctx.info["chain_count"]
    try:
        with pytest.raises(fhe_amd.FheError):
            ctx.relinearize_uneven([], [], key, key_mid, [])
        with pytest.raises(fhe_amd.FheError):
            ctx.relinearize_uneven([tr], [full], key, key_mid, [out])
        with pytest.raises(fhe_amd.FheError):
            ctx.relinearize_uneven([tr], [key_mid - 1], key, key_mid, [out])
        with pytest.raises(fhe_amd.FheError):
            ctx.relinearize_uneven([tr], [full - 6], key, key_mid, [tr])
        with pytest.raises(fhe_amd.FheError):
            ctx.rotate_rotation_key_uneven([out], [full - 6], [], [], [], [1],
                                           [out])
    finally:
        for arr in (key, tr, out):
            arr.free()
        ctx.close()
