"""Bounds of the lazy transform arithmetic used by the representative-free
stages (fhe_amd/csrc/ntt_r16.h): a model of the kLazySmall and kLazyLarge
butterflies in exact integers checks, for worst-case inputs just below 2q,
that every intermediate stays inside the stated bounds (so no 64-bit
wrap-around can occur on the device) and that the residues equal the exact
transform's.  The device results themselves are checked bit-exact against
the oracle by the GPU tests."""
import random

import pytest

M64 = (1 << 64) - 1


def mulhi64(a, b):
    a0, a1, b0, b1 = a & 0xffffffff, a >> 32, b & 0xffffffff, b >> 32
    ll = a0 * b0
    lh = a0 * b1 + (ll >> 32)
    hl = a1 * b0 + (lh & 0xffffffff)
    return (a1 * b1 + (lh >> 32) + (hl >> 32)) & M64


def shoup(y, w, wq, q):
    assert 0 <= y <= M64
    r = (w * y - mulhi64(wq, y) * q) & M64
    assert r < 2 * q
    return r


def ntt_prime(bits, n):
    x = ((1 << bits) // (2 * n)) * (2 * n) + 1
    while True:
        if all(pow(a, x - 1, x) == 1 for a in (2, 3, 5, 7)):
            return x
        x -= 2 * n


def root(q, n):
    for a in range(2, 1000):
        r = pow(a, (q - 1) // (2 * n), q)
        if pow(r, n, q) == q - 1:
            return r


def brev(x, b):
    return int(format(x, f"0{b}b")[::-1], 2)


def forward(x, q, log_n, mode):
    n = 1 << log_n
    psi = root(q, n)
    x, q2, top = list(x), 2 * q, 0
    for s in range(log_n):
        half = n >> (s + 1)
        for g in range(1 << s):
            w = pow(psi, brev((1 << s) + g, log_n), q)
            wq = (w << 64) // q
            for j in range(2 * g * half, 2 * g * half + half):
                a, b = x[j], x[j + half]
                if mode == "exact":
                    t = w * b % q
                    x[j], x[j + half] = (a + t) % q, (a - t) % q
                    continue
                if mode == "large" and a >= q2:
                    a -= q2
                t = shoup(b, w, wq, q)
                x[j], x[j + half] = a + t, a + q2 - t
                top = max(top, x[j], x[j + half])
    return x, top


def inverse(x, q, log_n, mode):
    n = 1 << log_n
    psi_inv = pow(root(q, n), -1, q)
    x, q2, top = list(x), 2 * q, 0
    for t in range(log_n):
        half = 1 << t
        for g in range(n >> (t + 1)):
            w = pow(psi_inv, brev((n >> (t + 1)) + g, log_n), q)
            wq = (w << 64) // q
            for j in range(2 * g * half, 2 * g * half + half):
                a, b = x[j], x[j + half]
                if mode == "exact":
                    x[j], x[j + half] = (a + b) % q, (a - b) * w % q
                    continue
                total, off = a + b, q2
                if mode == "large":
                    total = total - q2 if total >= q2 else total
                else:
                    off = q << (1 + t)
                    assert b < off
                x[j], x[j + half] = total, shoup(a + off - b, w, wq, q)
                top = max(top, x[j], x[j + half])
    return x, top


@pytest.mark.parametrize("bits,mode", [(40, "small"), (41, "small"),
                                       (60, "large"), (54, "large")])
def test_lazy_butterflies_bounds_and_residues(bits, mode):
    log_n = 8
    q = ntt_prime(bits, 1 << log_n)
    rng = random.Random(bits)
    x = [2 * q - 1 - rng.randrange(64) for _ in range(1 << log_n)]
    exact, _ = forward([v % q for v in x], q, log_n, "exact")
    lazy, top = forward(x, q, log_n, mode)
    assert [v % q for v in lazy] == exact
    # forward: below 2q + 2q per stage (small), below 4q (large)
    assert top < ((2 + 2 * log_n) * q if mode == "small" else 4 * q)
    exact, _ = inverse([v % q for v in x], q, log_n, "exact")
    lazy, top = inverse(x, q, log_n, mode)
    assert [v % q for v in lazy] == exact
    # inverse: below 2q * 2^stages (small; 2^59 after 16 stages for q < 2^42)
    assert top < ((2 * q) << log_n if mode == "small" else 2 * q)


@pytest.mark.parametrize("bits", [40, 54, 60])
def test_lazy_moddown_step_matches_reference_formula(bits):
    """ModDown's per-limb steps v = mont(guard(v + 2q - p_j, 2q), s_j^-1 R)
    (signed Montgomery products, then make_unsigned_reduce_2q_to_q) equal
    the Shoup form on v + q_top - p_j, for peeled integers p_j over the range
    (-2^60, 2^62) that the special-prime chain produces (canonical special
    limbs below 2^61, then signed products of values in (-2^61, 2^62) with
    constants below 2^61)."""
    from support import Oracle
    lib = Oracle.lib()
    q = ntt_prime(bits, 1 << 10)
    k = lib.oracle_calculate_k(q)
    top = ((1 << 63) - 1) // q * q
    rng = random.Random(bits)
    R = 1 << 62
    for trial in range(300):
        v = rng.randrange(q)
        steps = []
        for _ in range(4):
            s_inv = rng.randrange(1, q)
            p = rng.choice([rng.randrange(-(1 << 60) + 1, 1 << 62),
                            -(1 << 60) + 1, (1 << 62) - 1, 0])
            steps.append((s_inv, p))
        # the reference's signed chain
        x = v
        for s_inv, p in steps:
            g = x + 2 * q - p
            g = g if g < 2 * q else g - 2 * q
            x = lib.oracle_mont_mult(g, s_inv * R % q, q, k)
        expect = x + q if x < 0 else (x if x < q else x - q)
        # the lazy Shoup chain
        y = v
        for s_inv, p in steps:
            a = (y + top - p) & M64
            assert 0 <= y + top - p < 1 << 64
            y = shoup(a, s_inv, (s_inv << 64) // q, q)
        got = y if y < q else y - q
        assert got == expect, (trial, v, steps)


@pytest.mark.parametrize("bits,mode", [(40, "small"), (41, "small"),
                                       (60, "large"), (54, "large")])
def test_lazy_inverse_epilogue_folds_the_exit(bits, mode):
    """The lazy inverse transforms' epilogue: one Montgomery product with
    N^-1 R^-1 (mont_exit of the table's N^-1 R) in place of the product
    with N^-1 R and the exit.  Over inputs up to the lazy bounds (2^59 for
    primes below 2^42 after 16 stages, 2q otherwise) the result is below
    2q and has the residue of the two-step form."""
    log_n = 16
    q = ntt_prime(bits, 1 << log_n)
    r_bits = 62
    mask = (1 << r_bits) - 1
    k = (-pow(q, -1, 1 << r_bits)) % (1 << r_bits)

    def mont(a, b):
        x = a * b
        s = ((x & mask) * k) & mask
        return (x + s * q) >> r_bits

    ninv = pow(1 << log_n, -1, q) * (1 << r_bits) % q  # N^-1 R
    folded = mont(ninv, 1)                             # mont_exit: N^-1
    top = (1 << 59) if mode == "small" else 2 * q
    rng = random.Random(bits + 1)
    for v in [0, 1, q - 1, q, 2 * q - 1, top - 1] + \
            [rng.randrange(top) for _ in range(2000)]:
        y = mont(v, folded)
        assert y < 2 * q
        assert y % q == mont(mont(v, ninv), 1) % q
