"""GPU parity of the per-coefficient primitives (fhe_elementwise) with the
oracle, bit-exact, on canonical, lazy and negative operands including the
boundary values, for every operation code."""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Oracle, golden, ptr

pytestmark = pytest.mark.gpu

CASE = [c for c in golden("hmult_golden.json")["cases"]
        if c["name"] == "bronze_mid1"][0]


@pytest.fixture(scope="module")
def ctx():
    c = fhe_amd.Context(CASE["log_n"], CASE["special_prime_count"],
                        primes=CASE["primes"])
    yield c
    c.close()


def operands(q, n, rng):
    edges = [0, 1, q - 1, q, q + 1, 2 * q - 1, -1, -q + 1, q // 2, q // 2 + 1]
    pool = np.concatenate([
        rng.integers(0, q, n // 3, dtype=np.int64),
        rng.integers(0, 2 * q, n // 3, dtype=np.int64),
        rng.integers(-q + 1, 0, n - 2 * (n // 3), dtype=np.int64)])
    pool[:len(edges)] = edges
    return rng.permutation(pool)


def oracle_apply(op, x1, x2, y0, primes, mid, consts):
    lib = Oracle.lib()
    lib.oracle_elementwise_limb.argtypes = [
        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64]
    polys, limbs, n = x1.shape
    out = y0.copy()
    for p in range(polys):
        for l in range(limbs):
            q = int(primes[mid + l])
            k = lib.oracle_calculate_k(q)
            a = np.ascontiguousarray(x1[p, l])
            b = np.ascontiguousarray(x2[p, l])
            y = np.ascontiguousarray(out[p, l])
            lib.oracle_elementwise_limb(op, ptr(a), ptr(b), ptr(y), n, q, k,
                                        int(consts[l]))
            out[p, l] = y
    return out


@pytest.mark.parametrize("op", list(range(22)))
def test_elementwise_matches_oracle(op, ctx, gpu_lib):
    rng = np.random.default_rng(1000 + op)
    mid, limbs, polys, n = 2, 5, 3, ctx.n
    primes = CASE["primes"]
    x1 = np.stack([np.stack([operands(int(primes[mid + l]), n, rng)
                             for l in range(limbs)]) for _ in range(polys)])
    x2 = np.stack([np.stack([operands(int(primes[mid + l]), n, rng)
                             for l in range(limbs)]) for _ in range(polys)])
    y0 = np.stack([np.stack([rng.integers(0, 2 * int(primes[mid + l]), n,
                                          dtype=np.int64)
                             for l in range(limbs)]) for _ in range(polys)])
    consts = np.array([int(rng.integers(0, int(primes[mid + l])))
                       for l in range(limbs)], np.int64)
    d1, d2, dy = ctx.to_device(x1), ctx.to_device(x2), ctx.to_device(y0)
    ctx.elementwise(op, d1, dy, mid, x2=d2, constants=consts)
    expect = oracle_apply(op, x1, x2, y0, primes, mid, consts)
    assert (dy.download() == expect).all()
    # in place over the first operand
    ctx.elementwise(op, d1, d1, mid, x2=d2, constants=consts)
    if op != fhe_amd.EW["MULT_ADD"]:
        assert (d1.download() == oracle_apply(op, x1, x2, np.zeros_like(y0),
                                              primes, mid, consts)).all()
    for d in (d1, d2, dy):
        d.free()


def test_elementwise_rejects_missing_operands(ctx, gpu_lib):
    lib = fhe_amd.library()
    d = ctx.array((1, 2, ctx.n))
    assert lib.fhe_elementwise(ctx.handle, fhe_amd.EW["ADD"], d.ptr, None,
                               d.ptr, 1, 2, 0, None) == 1
    assert lib.fhe_elementwise(ctx.handle, fhe_amd.EW["MONT_ENTER"], d.ptr,
                               None, d.ptr, 1, 2, 0, None) == 1
    assert lib.fhe_elementwise(ctx.handle, 99, d.ptr, None, d.ptr, 1, 2, 0,
                               None) == 1
    d.free()
