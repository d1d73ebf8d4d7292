"""GPU parity of the key-switching path: HMult + relinearize + rescale,
relinearize alone and rescale alone, bit-exact.

  * every case of tests/golden/hmult_golden.json (digests of the reference
    engine's own output limbs, full Gold size included), inputs regenerated
    from the recorded seeds;
  * batched launches equal the one-ciphertext launches and the oracle;
  * with the compiled reference present (oracle/_ref), real keys and real
    encryptions: GPU limbs equal the reference's and decrypt to x*y.
"""
import hashlib

import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference, golden, hmult_inputs, ptr, uniform_limbs

pytestmark = pytest.mark.gpu

CASES = golden("hmult_golden.json")["cases"]


def digest(arr):
    return [[hashlib.sha256(np.ascontiguousarray(arr[p, l]).tobytes())
             .hexdigest()[:32] for l in range(arr.shape[1])]
            for p in range(arr.shape[0])]


def case_inputs(case):
    n = 1 << case["log_n"]
    A, B, K = hmult_inputs(case["primes"], n, case["moduli_id"], case["limbs"],
                           case["key_moduli_id"], case["key_polys"],
                           case["key_chain"], case["seed"])
    shape = (2, case["limbs"], n)
    return (A.reshape(shape), B.reshape(shape),
            K.reshape(case["key_polys"], case["key_chain"], n))


_contexts = {}


def context_for(case):
    key = (case["log_n"], case["special_prime_count"], tuple(case["primes"]))
    if key not in _contexts:
        _contexts[key] = fhe_amd.Context(case["log_n"],
                                         case["special_prime_count"],
                                         primes=case["primes"])
    return _contexts[key]


@pytest.fixture(scope="module", autouse=True)
def _close_contexts():
    yield
    for ctx in _contexts.values():
        ctx.close()
    _contexts.clear()


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_hmult_matches_reference_digests(case, gpu_lib):
    ctx = context_for(case)
    A, B, K = case_inputs(case)
    assert ctx.key_shape(case["key_moduli_id"]) == K.shape
    da, db, dk = ctx.to_device(A[None]), ctx.to_device(B[None]), ctx.to_device(K)
    out = ctx.array((1, 2, case["limbs"] - 1, ctx.n))
    ctx.hmult_relin_rescale(da, db, dk, case["key_moduli_id"], out,
                            case["moduli_id"])
    got = out.download()[0]
    for arr in (da, db, dk, out):
        arr.free()
    assert digest(got) == case["out_sha256"]


@pytest.mark.parametrize("name", ["bronze_mid4", "silver_mid9", "gold_mid13"])
def test_batched_equals_single_and_oracle(name, gpu_lib):
    """Three different ciphertext pairs in one launch: each output equals its
    one-pair launch; the small ring is also checked against the oracle."""
    case = [c for c in CASES if c["name"] == name][0]
    ctx = context_for(case)
    n, mid, limbs = ctx.n, case["moduli_id"], case["limbs"]
    _, _, K = case_inputs(case)
    rng = np.random.Generator(np.random.PCG64(99))
    batch = 3
    A = np.stack([np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    B = np.stack([np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    dk = ctx.to_device(K)
    da, db = ctx.to_device(A), ctx.to_device(B)
    out = ctx.array((batch, 2, limbs - 1, n))
    ctx.hmult_relin_rescale(da, db, dk, case["key_moduli_id"], out, mid)
    batched = out.download()
    for c in range(batch):
        sa, sb = ctx.to_device(A[c:c + 1]), ctx.to_device(B[c:c + 1])
        so = ctx.array((1, 2, limbs - 1, n))
        ctx.hmult_relin_rescale(sa, sb, dk, case["key_moduli_id"], so, mid)
        assert (so.download()[0] == batched[c]).all(), c
        for arr in (sa, sb, so):
            arr.free()
    if case["log_n"] <= 15:
        orc = Oracle(case["log_n"], case["primes"],
                     case["special_prime_count"])
        for c in range(batch):
            ref = orc.hmult(A[c], B[c], K, mid, case["key_moduli_id"])
            assert (batched[c] == ref).all(), c
    for arr in (da, db, dk, out):
        arr.free()


def test_batch32_benchmark_shape_matches_digests_and_single_launches(gpu_lib):
    """The benchmark's launch: 32 Gold pairs at moduli_id 13 (rescaled level
    24) in one batch.  The recorded pair sits at positions 0, 13 and 31 and
    equals the reference engine's digest there; every position equals its
    own one-pair launch (grid sizes, XCD swizzle and workspace differ)."""
    case = [c for c in CASES if c["name"] == "gold_mid13"][0]
    ctx = context_for(case)
    n, mid, limbs = ctx.n, case["moduli_id"], case["limbs"]
    A0, B0, K = case_inputs(case)
    rng = np.random.Generator(np.random.PCG64(32))
    batch = 32
    A = np.stack([np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    B = np.stack([np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    recorded = (0, 13, 31)
    for c in recorded:
        A[c], B[c] = A0, B0
    dk = ctx.to_device(K)
    da, db = ctx.to_device(A), ctx.to_device(B)
    out = ctx.array((batch, 2, limbs - 1, n))
    ctx.hmult_relin_rescale(da, db, dk, case["key_moduli_id"], out, mid)
    batched = out.download()
    for c in recorded:
        assert digest(batched[c]) == case["out_sha256"], c
    sa, sb = ctx.array((1, 2, limbs, n)), ctx.array((1, 2, limbs, n))
    so = ctx.array((1, 2, limbs - 1, n))
    for c in range(batch):
        sa.upload(A[c:c + 1])
        sb.upload(B[c:c + 1])
        ctx.hmult_relin_rescale(sa, sb, dk, case["key_moduli_id"], so, mid)
        assert (so.download()[0] == batched[c]).all(), c
    for arr in (da, db, dk, out, sa, sb, so):
        arr.free()


@pytest.mark.reference
@pytest.mark.skipif(not Reference.available(),
                    reason="oracle/_ref (compiled reference) not built")
def test_batch32_real_keys_match_reference(gpu_lib):
    """32 pairs of real Gold encryptions (rescaled level 24, the
    benchmark's shape) in one launch: every output equals the reference
    engine's multiply(ct, ct, rlk) + rescale of its pair."""
    ref = Reference(16, 4, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    key_mid, key = ref.key(rlk)
    slots = ref.slot_count
    rng = np.random.default_rng(3232)
    handles = []
    ctx = fhe_amd.Context(16, 4, primes=ref.primes)
    try:
        # 8 rescaled encryptions give 32 distinct ordered pairs
        cts = []
        for _ in range(8):
            x = rng.uniform(-1, 1, slots)
            c0 = lib.ref_encrypt(e, ptr(x), slots, pk, 24)
            c1 = lib.ref_rescale(e, c0)
            handles += [c0, c1]
            cts.append(c1)
        mid = ref.shape(cts[0])[0]
        assert mid == 13
        pairs = [(i, j) for i in range(8) for j in range(8) if i != j][:32]
        data = [ref.get(c) for c in cts]
        A = np.stack([data[i] for i, _ in pairs])
        B = np.stack([data[j] for _, j in pairs])
        da, db, dk = ctx.to_device(A), ctx.to_device(B), ctx.to_device(key)
        limbs = A.shape[2]
        out = ctx.array((32, 2, limbs - 1, ref.n))
        ctx.hmult_relin_rescale(da, db, dk, key_mid, out, mid)
        got = out.download()
        for k, (i, j) in enumerate(pairs):
            w = lib.ref_hmult_relin_rescale(e, cts[i], cts[j], rlk)
            want = ref.get(w)
            lib.ref_free_ct(w)
            assert np.array_equal(got[k], want), (k, i, j)
        for arr in (da, db, dk, out):
            arr.free()
    finally:
        ctx.close()
        for h in handles:
            lib.ref_free_ct(h)
        lib.ref_free_relin_key(rlk)
        lib.ref_free_public_key(pk)
        lib.ref_free_secret_key(sk)
        ref.close()


@pytest.mark.parametrize("name", ["iron_mid1", "silver_mid1", "gold_mid30"])
def test_relinearize_matches_oracle(name, gpu_lib):
    case = [c for c in CASES if c["name"] == name][0]
    ctx = context_for(case)
    orc = Oracle(case["log_n"], case["primes"], case["special_prime_count"])
    n, mid, limbs = ctx.n, case["moduli_id"], case["limbs"]
    _, _, K = case_inputs(case)
    rng = np.random.Generator(np.random.PCG64(5))
    trio = np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n, 2)
                     for _ in range(3)])
    dt, dk = ctx.to_device(trio[None]), ctx.to_device(K)
    out = ctx.array((1, 2, limbs, n))
    ctx.relinearize(dt, dk, case["key_moduli_id"], out, mid)
    got = out.download()[0]
    assert (got == orc.relinearize(trio, K, mid, case["key_moduli_id"])).all()
    assert (dt.download()[0] == trio).all()  # input untouched
    for arr in (dt, dk, out):
        arr.free()


@pytest.mark.parametrize("name", ["iron_mid1", "bronze_mid1", "gold_mid13"])
def test_rescale_matches_oracle(name, gpu_lib):
    case = [c for c in CASES if c["name"] == name][0]
    ctx = context_for(case)
    orc = Oracle(case["log_n"], case["primes"], case["special_prime_count"])
    n, mid, limbs = ctx.n, case["moduli_id"], case["limbs"]
    rng = np.random.Generator(np.random.PCG64(6))
    x = np.stack([uniform_limbs(rng, case["primes"], mid, limbs, n)
                  for _ in range(4)])
    # edge values: zero, q-1, the rounding boundary q0/2 and q0/2 + 1
    q0 = int(case["primes"][mid])
    x[0, :, :4] = 0
    x[1, 0, :3] = [q0 - 1, q0 // 2, q0 // 2 + 1]
    dx = ctx.to_device(x)
    dy = ctx.array((4, limbs - 1, n))
    ctx.rescale(dx, dy, mid)
    assert (dy.download() == orc.rescale(x, mid)).all()
    dx.free()
    dy.free()


def test_invalid_calls_are_rejected(gpu_lib):
    case = [c for c in CASES if c["name"] == "iron_mid1"][0]
    ctx = context_for(case)
    lib = fhe_amd.library()
    d = ctx.array((1, 2, case["limbs"], ctx.n))
    k = ctx.array(ctx.key_shape(case["key_moduli_id"]))
    # empty batch, ciphertext below the key level, too-low level, null key
    assert lib.fhe_hmult_relin_rescale(ctx.handle, 0, 1, d.ptr, d.ptr, k.ptr,
                                       1, d.ptr) == 1
    assert lib.fhe_hmult_relin_rescale(ctx.handle, 1, 0, d.ptr, d.ptr, k.ptr,
                                       1, d.ptr) == 1
    assert lib.fhe_hmult_relin_rescale(ctx.handle, 1, ctx.chain_count - 2,
                                       d.ptr, d.ptr, k.ptr, 1, d.ptr) == 1
    assert lib.fhe_hmult_relin_rescale(ctx.handle, 1, 1, d.ptr, d.ptr, None,
                                       1, d.ptr) == 1
    assert b"level" in lib.fhe_last_error() or lib.fhe_last_error()
    d.free()
    k.free()


@pytest.mark.reference
@pytest.mark.skipif(not Reference.available(),
                    reason="oracle/_ref (compiled reference) not built")
@pytest.mark.parametrize("log_n,alpha,level", [(14, 1, 5), (16, 4, 24)])
def test_real_keys_match_reference_and_decrypt(log_n, alpha, level, gpu_lib):
    """Reference keygen + encryption, GPU HMult+relin+rescale: the limbs equal
    the reference engine's and decrypt (reference decrypt) to x*y."""
    ref = Reference(log_n, alpha, threads=8)
    lib = ref.lib
    e = ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    slots = ref.slot_count
    rng = np.random.default_rng(2024)
    x, y = rng.uniform(-1, 1, slots), rng.uniform(-1, 1, slots)
    a0 = lib.ref_encrypt(e, ptr(x), slots, pk, level)
    b0 = lib.ref_encrypt(e, ptr(y), slots, pk, level)
    a, b = lib.ref_rescale(e, a0), lib.ref_rescale(e, b0)
    mid = ref.shape(a)[0]
    key_mid, key = ref.key(rlk)
    c = lib.ref_hmult_relin_rescale(e, a, b, rlk)
    expect = ref.get(c)

    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
    da, db = ctx.to_device(ref.get(a)[None]), ctx.to_device(ref.get(b)[None])
    dk = ctx.to_device(key)
    out = ctx.array((1,) + expect.shape)
    ctx.hmult_relin_rescale(da, db, dk, key_mid, out, mid)
    got = out.download()[0]
    assert (got == expect).all()

    # identical limbs decrypt identically; check the value itself too
    dec_ref = np.zeros(slots)
    lib.ref_decrypt(e, c, sk, ptr(dec_ref))
    ref.set(c, got)
    dec = np.zeros(slots)
    lib.ref_decrypt(e, c, sk, ptr(dec))
    assert np.abs(dec - dec_ref).max() <= 2 ** -25
    # the scheme's noise, which varies with the reference's unseeded
    # randomness
    assert np.abs(dec - x * y).max() < 1e-4
    for arr in (da, db, dk, out):
        arr.free()
    ctx.close()
    for h in (a0, b0, a, b, c):
        lib.ref_free_ct(h)
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()
