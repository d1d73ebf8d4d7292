"""GPU parity of the scalar operations and level down with the reference
engine (oracle/_ref) on real encryptions, bit-exact, plus the same ops
against the oracle's primitives with the product's multiplier constants.

  fhe_multiply_double  CKKSEngine::multiply(ct, double)
  fhe_multiply_int     CKKSEngine::multiply(ct, int64_t)
  fhe_add_double       CKKSEngine::add(ct, double)
  fhe_level_down       CKKSEngine::level_down(ct, level), rescaled or not
"""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference, ptr

pytestmark = pytest.mark.gpu

needs_ref = pytest.mark.skipif(not Reference.available(),
                               reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def env():
    ref = Reference(14, 2, threads=4)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    ctx = fhe_amd.Context(14, 2, primes=ref.primes)
    info = np.zeros(9, np.int64)
    lib.ref_engine_info(e, ptr(info))
    yield dict(ref=ref, lib=lib, e=e, sk=sk, pk=pk, ctx=ctx,
               full=int(info[4]))
    ctx.close()
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()


def encrypt(env, level, seed, rescaled=False):
    ref, lib = env["ref"], env["lib"]
    x = np.random.default_rng(seed).uniform(-1, 1, ref.slot_count)
    # a rescale keeps the level number and moves to the next moduli_id
    ct = lib.ref_encrypt(env["e"], ptr(x), ref.slot_count, env["pk"], level)
    if rescaled:
        r = lib.ref_rescale(env["e"], ct)
        lib.ref_free_ct(ct)
        ct = r
    return ct


def run_gpu(env, fn, host, out_shape):
    ctx = env["ctx"]
    d = ctx.to_device(host)
    o = ctx.array(out_shape)
    fn(ctx, d, o)
    got = o.download()
    d.free()
    o.free()
    return got


@needs_ref
@pytest.mark.parametrize("level", [1, 3, 6])
@pytest.mark.parametrize("value", [0.5, -1.7, 3.0])
def test_multiply_double_matches_reference(env, level, value, gpu_lib):
    lib, ref = env["lib"], env["ref"]
    ct = encrypt(env, level, level)
    expect_h = lib.ref_multiply_double(env["e"], ct, value)
    expect = ref.get(expect_h)
    got = run_gpu(env, lambda c, d, o: c.multiply_double(d, o, level, value),
                  ref.get(ct), expect.shape)
    assert ref.shape(expect_h)[0] == ref.shape(ct)[0] + 1
    assert (got == expect).all()
    lib.ref_free_ct(ct)
    lib.ref_free_ct(expect_h)


@needs_ref
@pytest.mark.parametrize("value", [3, -5, 1 << 40])
def test_multiply_int_matches_reference(env, value, gpu_lib):
    lib, ref = env["lib"], env["ref"]
    ct = encrypt(env, 4, 11)
    expect_h = lib.ref_multiply_int(env["e"], ct, value)
    expect = ref.get(expect_h)
    mid = ref.shape(ct)[0]
    got = run_gpu(env, lambda c, d, o: c.multiply_int(d, o, mid, value),
                  ref.get(ct), expect.shape)
    assert (got == expect).all()
    lib.ref_free_ct(ct)
    lib.ref_free_ct(expect_h)


@needs_ref
@pytest.mark.parametrize("level", [0, 2, 6])
@pytest.mark.parametrize("value", [0.25, -2.5])
def test_add_double_matches_reference(env, level, value, gpu_lib):
    lib, ref = env["lib"], env["ref"]
    ct = encrypt(env, level, 20 + level)
    expect_h = lib.ref_add_double(env["e"], ct, value)
    expect = ref.get(expect_h)
    got = run_gpu(env, lambda c, d, o: c.add_double(d, o, level, value),
                  ref.get(ct)[None], (1,) + expect.shape)[0]
    assert (got == expect).all()
    lib.ref_free_ct(ct)
    lib.ref_free_ct(expect_h)


@needs_ref
@pytest.mark.parametrize("rescaled", [False, True])
@pytest.mark.parametrize("level,dst", [(6, 5), (6, 2), (3, 0), (1, 0)])
def test_level_down_matches_reference(env, level, dst, rescaled, gpu_lib):
    lib, ref = env["lib"], env["ref"]
    ct = encrypt(env, level, 40 + level, rescaled=rescaled)
    expect_h = lib.ref_level_down(env["e"], ct, dst)
    assert expect_h, lib.ref_last_error()
    expect = ref.get(expect_h)
    got = run_gpu(env, lambda c, d, o: c.level_down(d, o, level, dst,
                                                    rescaled=rescaled),
                  ref.get(ct), expect.shape)
    assert (got == expect).all()
    lib.ref_free_ct(ct)
    lib.ref_free_ct(expect_h)


def test_scalar_ops_match_oracle_primitives(gpu_lib):
    """Without the reference: the same ops equal the oracle's primitives
    (enter-reduce, rescale) applied with the product's constants."""
    log_n, alpha = 13, 1
    primes = fhe_amd.resolve_primes(log_n, alpha)
    ctx = fhe_amd.Context(log_n, alpha, primes=primes)
    orc = Oracle(log_n, primes, alpha)
    lib = Oracle.lib()
    lib.oracle_elementwise_limb.argtypes = [
        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64]
    full = ctx.info["full_level"]
    level = full - 1
    mid = full - level
    limbs = ctx.limbs_at(mid)
    rng = np.random.default_rng(3)
    x = np.stack([np.stack([rng.integers(0, int(primes[mid + l]), ctx.n,
                                         dtype=np.int64)
                            for l in range(limbs)]) for _ in range(2)])

    def enter_reduce(arr, consts, first):
        out = np.empty_like(arr)
        for p in range(arr.shape[0]):
            for l in range(arr.shape[1]):
                q = int(primes[first + l])
                a = np.ascontiguousarray(arr[p, l])
                y = np.empty_like(a)
                lib.oracle_elementwise_limb(20, ptr(a), None, ptr(y), ctx.n,
                                            q, lib.oracle_calculate_k(q),
                                            int(consts[l]))
                out[p, l] = y
        return out

    c = ctx.multiplier("MULT", level, value=0.75)
    expect = orc.rescale(enter_reduce(x, c, mid), mid)
    got = run_gpu({"ctx": ctx},
                  lambda cc, d, o: cc.multiply_double(d, o, level, 0.75), x,
                  expect.shape)
    assert (got == expect).all()
    c = ctx.multiplier("MULT_INT", mid, int_value=-9)
    expect = enter_reduce(x, c, mid)
    got = run_gpu({"ctx": ctx},
                  lambda cc, d, o: cc.multiply_int(d, o, mid, -9), x,
                  expect.shape)
    assert (got == expect).all()
    ctx.close()


@needs_ref
@pytest.mark.parametrize("level", [0, 4])
def test_add_subtract_negate_match_reference(env, level, gpu_lib):
    """CKKSEngine::add / subtract (ct, ct) and negate on INTT-form
    ciphertexts = FHE_EW_ADD_REDUCE / SUB_REDUCE / NEGATE_MAKE_UNSIGNED over
    both polynomials."""
    lib, ref, ctx = env["lib"], env["ref"], env["ctx"]
    a, b = encrypt(env, level, 60 + level), encrypt(env, level, 70 + level)
    mid = ref.shape(a)[0]
    da, db = ctx.to_device(ref.get(a)), ctx.to_device(ref.get(b))
    out = ctx.array(da.shape)
    for op, ref_fn in (("ADD_REDUCE", lambda: lib.ref_add_ct(env["e"], a, b)),
                       ("SUB_REDUCE",
                        lambda: lib.ref_subtract_ct(env["e"], a, b)),
                       ("NEGATE_MAKE_UNSIGNED",
                        lambda: lib.ref_negate(env["e"], a))):
        h = ref_fn()
        assert h, lib.ref_last_error()
        ctx.elementwise(fhe_amd.EW[op], da, out, mid, x2=db)
        assert (out.download() == ref.get(h)).all(), op
        lib.ref_free_ct(h)
    for d in (da, db, out):
        d.free()
    lib.ref_free_ct(a)
    lib.ref_free_ct(b)
