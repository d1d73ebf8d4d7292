"""GPU parity of the batched NTT / INTT / Montgomery-multiply kernels against
the oracle, bit-exact including lazy representatives in [0, 2q)."""
import numpy as np
import pytest

import fhe_amd
from support import Oracle, golden, uniform_limbs

pytestmark = pytest.mark.gpu


def tier_primes(tier):
    return golden("precomputed_tiers.json")["tests"][tier]["arrays"]["q"]


# (name, log_n, primes, alpha)
RINGS = [
    ("iron", 13, tier_primes("Iron"), 1),
    ("bronze", 14, tier_primes("Bronze"), 1),
    ("silver", 15, tier_primes("Silver"), 2),
    ("gold", 16, tier_primes("Gold"), 4),
]


@pytest.fixture(scope="module", params=RINGS, ids=lambda r: r[0])
def ring(request, gpu_lib):
    name, log_n, primes, alpha = request.param
    ctx = fhe_amd.Context(log_n, alpha, primes=primes)
    orc = Oracle(log_n, primes, alpha)
    yield ctx, orc
    ctx.close()


def random_polys(primes, mid, polys, limbs, n, seed, factor=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.stack([uniform_limbs(rng, primes, mid, limbs, n, factor)
                     for _ in range(polys)])


@pytest.mark.parametrize("enter", [True, False])
def test_ntt_forward_matches_oracle(ring, enter):
    ctx, orc = ring
    mid, limbs, polys = 1, min(6, ctx.chain_count - 1), 3
    x = random_polys(ctx.primes, mid, polys, limbs, ctx.n, 10 + enter,
                     factor=2)
    d = ctx.to_device(x)
    ctx.ntt(d, mid, enter=enter)
    assert (d.download() == orc.ntt(x, mid, enter=enter)).all()
    d.free()


@pytest.mark.parametrize("post", [0, 1, 2])
def test_ntt_inverse_matches_oracle(ring, post):
    ctx, orc = ring
    mid, limbs, polys = 0, min(5, ctx.chain_count), 2
    x = random_polys(ctx.primes, mid, polys, limbs, ctx.n, 20 + post,
                     factor=2)
    d = ctx.to_device(x)
    ctx.intt(d, mid, post=post)
    assert (d.download() == orc.intt(x, mid, post=post)).all()
    d.free()


def test_special_prime_limbs(ring):
    """Extended-basis limbs (Q limbs then the specials) transform with the
    special primes' tables."""
    ctx, orc = ring
    mid = ctx.chain_count - 2
    limbs = 2 + ctx.alpha
    x = random_polys(ctx.primes, mid, 1, limbs, ctx.n, 31)
    d = ctx.to_device(x)
    ctx.ntt(d, mid, enter=True)
    y = d.download()
    assert (y == orc.ntt(x, mid, enter=True)).all()
    ctx.intt(d, mid, post=1)
    assert (d.download() == orc.intt(y, mid, post=1)).all()
    assert (d.download() == x).all()
    d.free()


def test_empty_and_out_of_range_shapes_are_rejected(ring):
    ctx, _ = ring
    d = ctx.array((1, 1, ctx.n))
    lib = fhe_amd.library()
    assert lib.fhe_ntt(ctx.handle, d.ptr, 0, 1, 0, 1) == 1
    assert lib.fhe_ntt(ctx.handle, d.ptr, 1, 1, len(ctx.primes), 1) == 1
    assert lib.fhe_intt(ctx.handle, d.ptr, 1, 1, 0, 7) == 1
    d.free()


def large_ntt_primes(log_n, count, bits=60):
    """The reference's large-prime search: walk down from 2^bits in steps of
    2N (src/arithmetic/prime.cpp:913-930)."""
    from test_oracle_golden import _is_prime
    n = 1 << log_n
    out = []
    x = ((1 << bits) // (2 * n)) * (2 * n) + 1
    while len(out) < count:
        if _is_prime(x):
            out.append(x)
        x -= 2 * n
    return out


def test_batched_ntt_mult_intt_config(gpu_lib):
    """64 polynomials x 24 NTT-friendly ~60-bit primes at N = 2^16: forward
    NTT (to_ntt), pointwise Montgomery multiply, INTT (to_intt), against the
    oracle limb for limb."""
    log_n, polys, limbs = 16, 64, 24
    n = 1 << log_n
    primes = large_ntt_primes(log_n, limbs)
    assert primes[0] == 1152921504606584833  # Gold's base prime
    ctx = fhe_amd.Context(log_n, 1, primes=primes)
    orc = Oracle(log_n, primes, 1)
    a = random_polys(primes, 0, polys, limbs, n, 41)
    b = random_polys(primes, 0, polys, limbs, n, 42)
    da, db = ctx.to_device(a), ctx.to_device(b)
    dz = ctx.array(a.shape)
    ctx.ntt(da, 0, enter=True)
    ctx.ntt(db, 0, enter=True)
    ctx.mont_mult(da, db, dz, 0)
    ctx.intt(dz, 0, post=1)
    z = dz.download()
    an, bn = orc.ntt(a, 0), orc.ntt(b, 0)
    assert (da.download() == an).all()
    assert (db.download() == bn).all()
    assert (z == orc.intt(orc.mont_mult(an, bn, 0), 0, post=1)).all()
    for arr in (da, db, dz):
        arr.free()
    ctx.close()
