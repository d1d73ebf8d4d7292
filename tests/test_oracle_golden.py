"""Pin the oracle (oracle/ckks_oracle.c) before trusting it.

  1. known-answer vectors from the reference's own tests
     (test/operation/ntt.cpp, test/montgomery/unit.cpp,
      test/ckks/precomputed/default.cpp), extracted to tests/golden/*.json;
  2. HMult+relin+rescale digests produced by the reference engine itself
     (tests/golden/hmult_golden.json);
  3. closed forms a plausible mistake would break: the Montgomery identity,
     NTT = evaluation at the odd powers of psi, INTT o NTT = id, exact key
     switching = floor(sum_p c_p X_p / P), rescale = round(X / q0).
"""
import hashlib
import random

import numpy as np
import pytest

from support import Oracle, Reference, golden, hmult_inputs, ptr


# ---------------------------------------------------------------- 1. KATs

def test_montgomery_known_answers():
    kat = golden("montgomery_vectors.json")["tests"]
    lib = Oracle.lib()
    checked = 0
    for name, t in kat.items():
        a = t["arrays"]
        s = t["scalars"]
        if "mont_mult" not in t["calls"] or not {"x1", "x2"} <= set(a):
            continue
        if "expected" not in a:
            continue
        for x1, x2, e in zip(a["x1"], a["x2"], a["expected"]):
            assert lib.oracle_mont_mult(x1, x2, s["q"], s["k"]) == e, name
            checked += 1
    assert checked >= 2
    # test/montgomery/unit.cpp MultiplyR: x * 2^62 * 2^-62 == x exactly
    q, k = 1099516280833, 2888757134567604223
    for x in (1, 2, 12345, q - 1):
        assert lib.oracle_mont_mult(x, 1 << 62, q, k) == x


def test_calculate_k_known_answers():
    kat = golden("ntt_vectors.json")["tests"]
    lib = Oracle.lib()
    for t in kat.values():
        s = t["scalars"]
        assert lib.oracle_calculate_k(s["q"]) == s["k"]


@pytest.mark.parametrize("name", ["NTT1", "NTT2", "NTT3", "NTTInplace1",
                                  "NTTInplace2", "NTTInplace3"])
def test_ntt_known_answers(name):
    t = golden("ntt_vectors.json")["tests"][name]
    a, s = t["arrays"], t["scalars"]
    x = np.array(a.get("x", a.get("y")), dtype=np.int64)
    root = np.array(a["root"], dtype=np.int64)
    Oracle.lib().oracle_ntt_raw(x.ctypes.data, root.ctypes.data, s["q"],
                                s["k"], s["log_coeff_count"])
    assert x.tolist() == a["expected"]


@pytest.mark.parametrize("name", ["INTT1", "INTT2", "INTT3", "INTTInplace1",
                                  "INTTInplace2", "INTTInplace3"])
def test_intt_known_answers(name):
    t = golden("ntt_vectors.json")["tests"][name]
    a, s = t["arrays"], t["scalars"]
    x = np.array(a.get("x", a.get("y")), dtype=np.int64)
    root = np.array(a["root_inverse"], dtype=np.int64)
    Oracle.lib().oracle_intt_raw(x.ctypes.data, root.ctypes.data, s["q"],
                                 s["k"], s["log_coeff_count"])
    assert x.tolist() == a["expected"]


TIERS = ["Iron", "Bronze", "Silver", "Gold", "BootstrapGold", "Platinum"]
TIER_SHAPE = {"Iron": (13, 1), "Bronze": (14, 1), "Silver": (15, 2),
              "Gold": (16, 4), "BootstrapGold": (16, 4), "Platinum": (17, 6)}


@pytest.mark.parametrize("tier", ["Iron", "Bronze", "Silver"])
def test_precomputed_constants(tier):
    """Per-prime constants, twiddle prefixes and chain constants of every
    tier small enough to derive quickly, against the reference fixture."""
    t = golden("precomputed_tiers.json")["tests"][tier]
    a = t["arrays"]
    log_n, alpha = TIER_SHAPE[tier]
    orc = Oracle(log_n, a["q"], alpha)
    for i in range(len(a["q"])):
        c, root, rinv = orc.constants(i)
        assert c[0] == np.int64(np.uint64(a["k"][i])), (tier, i)
        assert c[1] == a["r_squared"][i]
        assert c[2] == a["coeff_count_inverse"][i]
        if i == 0:
            assert root[:len(a["root"])].tolist() == a["root"]
            assert rinv[:len(a["root_inverse"])].tolist() == a["root_inverse"]


# ------------------------------------------------- 2. reference HMult digests

def _digest(arr):
    return [[hashlib.sha256(arr[p, l].tobytes()).hexdigest()[:32]
             for l in range(arr.shape[1])] for p in range(arr.shape[0])]


def hmult_cases(max_log_n):
    return [c for c in golden("hmult_golden.json")["cases"]
            if c["log_n"] <= max_log_n]


@pytest.mark.parametrize("case", hmult_cases(15), ids=lambda c: c["name"])
def test_oracle_hmult_matches_reference_digests(case):
    orc = Oracle(case["log_n"], case["primes"], case["special_prime_count"])
    n = 1 << case["log_n"]
    A, B, K = hmult_inputs(case["primes"], n, case["moduli_id"], case["limbs"],
                           case["key_moduli_id"], case["key_polys"],
                           case["key_chain"], case["seed"])
    out = orc.hmult(A, B, K, case["moduli_id"], case["key_moduli_id"])
    assert _digest(out) == case["out_sha256"]


def test_oracle_gold_hmult_matches_reference_digest():
    """Gold tier, 8 limbs (the cheapest Gold case) through the oracle."""
    case = [c for c in golden("hmult_golden.json")["cases"]
            if c["name"] == "gold_mid30"][0]
    orc = Oracle(case["log_n"], case["primes"], case["special_prime_count"])
    A, B, K = hmult_inputs(case["primes"], 1 << case["log_n"],
                           case["moduli_id"], case["limbs"],
                           case["key_moduli_id"], case["key_polys"],
                           case["key_chain"], case["seed"])
    out = orc.hmult(A, B, K, case["moduli_id"], case["key_moduli_id"])
    assert _digest(out) == case["out_sha256"]


# ---------------------------------------------------------- 3. closed forms

R = 1 << 62


def small_ring(log_n, alpha, count, bits=40, seed=0):
    """Distinct NTT-friendly primes near 2^bits (q = 1 mod 2N), found by
    brute force, for closed-form checks on tiny rings."""
    n = 1 << log_n
    out = []
    x = ((1 << bits) // (2 * n)) * (2 * n) + 1
    while len(out) < count:
        if _is_prime(x):
            out.append(x)
        x -= 2 * n
    random.Random(seed).shuffle(out)
    return out


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_montgomery_identity_random():
    lib = Oracle.lib()
    rng = random.Random(1)
    for q in small_ring(4, 1, 4, 40) + small_ring(4, 1, 4, 60):
        k = lib.oracle_calculate_k(q)
        assert (q * k) % R == R - 1  # k = -q^-1 mod 2^62
        for _ in range(200):
            a, b = rng.randrange(2 * q), rng.randrange(q)
            r = lib.oracle_mont_mult(a, b, q, k)
            assert 0 <= r < 2 * q and (r * R - a * b) % q == 0


@pytest.mark.parametrize("log_n", [3, 5, 8])
def test_ntt_is_evaluation_at_odd_powers(log_n):
    n = 1 << log_n
    primes = small_ring(log_n, 1, 3, 60)
    orc = Oracle(log_n, primes, 1)
    rng = np.random.default_rng(log_n)
    for i, q in enumerate(primes):
        c, root, _ = orc.constants(i)
        # psi from the twiddle table: root[0] = psi^(N/2) * R, root[N/2-1]
        # = psi^(bitrev(N/2)) = psi * R
        rinv = pow(R, -1, q)
        psi = int(root[n // 2 - 1]) * rinv % q
        assert pow(psi, n, q) == q - 1
        x = rng.integers(0, q, n, dtype=np.uint64).astype(np.int64)
        y = x.copy()
        orc.lib().oracle_ntt_limb(orc.h, i, y.ctypes.data)
        for j in range(n):
            br = int(format(j, f"0{log_n}b")[::-1], 2)
            point = pow(psi, 2 * br + 1, q)
            expect = sum(int(x[t]) * pow(point, t, q) for t in range(n)) % q
            assert 0 <= y[j] < 2 * q and y[j] % q == expect


def test_intt_inverts_ntt():
    log_n = 6
    primes = small_ring(log_n, 1, 4, 40)
    orc = Oracle(log_n, primes, 1)
    rng = np.random.default_rng(3)
    x = np.stack([rng.integers(0, q, 1 << log_n, dtype=np.uint64)
                  for q in primes[:3]]).astype(np.int64)
    y = orc.intt(orc.ntt(x[None], 0), 0, post=2)[0]
    assert (y == x).all()


def _crt(residues, moduli):
    M = 1
    for m in moduli:
        M *= m
    X = 0
    for r, m in zip(residues, moduli):
        Mi = M // m
        X += int(r) * Mi * pow(Mi, -1, m)
    return X % M


@pytest.mark.parametrize("alpha,chain", [(1, 3), (2, 5), (3, 7), (4, 6)])
def test_key_switch_closed_form(alpha, chain):
    """Key = NTT of constant polynomials c_p: the switched value is
    floor((sum_p c_p X_p mod QP) / P) mod q_i, X_p the exact CRT lift of the
    partition's residues."""
    log_n = 3
    n = 1 << log_n
    primes = small_ring(log_n, alpha, chain + alpha, 40, seed=alpha)
    orc = Oracle(log_n, primes, alpha)
    mid = key_mid = 0
    limbs = chain
    parts = orc.partitions(mid, key_mid)
    rng = random.Random(alpha * 10 + chain)
    consts = [rng.randrange(1, 1 << 30) for _ in parts]
    key_polys = 2 * len(parts)
    ext = limbs + alpha
    key = np.zeros((key_polys, ext, n), np.int64)
    for (start, size, kidx), c in zip(parts, consts):
        for i in range(ext):
            key[2 * kidx, i, :] = c * R % primes[i]      # c_p * R, NTT form
            key[2 * kidx + 1, i, :] = 2 * c * R % primes[i]
    x = np.stack([np.array([rng.randrange(q) for _ in range(n)])
                  for q in primes[:limbs]]).astype(np.int64)
    out = orc.key_switch(x, key, mid, key_mid)
    Q = 1
    for q in primes[:limbs]:
        Q *= q
    P = 1
    for q in primes[limbs:]:
        P *= q
    for coef in range(n):
        for half, mult in ((0, 1), (1, 2)):
            Y = 0
            for (start, size, _), c in zip(parts, consts):
                Xp = _crt([x[start + j, coef] for j in range(size)],
                          primes[start:start + size])
                Y += mult * c * Xp
            Y %= Q * P
            expect = Y // P
            for i in range(limbs):
                assert out[half, i, coef] == expect % primes[i]


def test_rescale_closed_form():
    log_n = 3
    n = 1 << log_n
    primes = small_ring(log_n, 1, 6, 40)
    orc = Oracle(log_n, primes, 1)
    chain = 5
    rng = random.Random(7)
    x = np.stack([np.array([rng.randrange(q) for _ in range(n)])
                  for q in primes[:chain]]).astype(np.int64)
    y = orc.rescale(x[None], 0)[0]
    q0 = primes[0]
    for coef in range(n):
        X = _crt(x[:, coef], primes[:chain])
        r = X % q0
        expect = (X - r) // q0 + (1 if r > q0 // 2 else 0)
        for i in range(1, chain):
            assert y[i - 1, coef] == expect % primes[i]


# ------------------------------------------- rotations and conjugation

def rotate_cases(max_log_n):
    return [c for c in golden("rotate_golden.json")["cases"]
            if c["log_n"] <= max_log_n]


@pytest.mark.parametrize("case", rotate_cases(15), ids=lambda c: c["name"])
def test_oracle_rotate_matches_reference_digests(case):
    orc = Oracle(case["log_n"], case["primes"], case["special_prime_count"])
    n = 1 << case["log_n"]
    A, _, K = hmult_inputs(case["primes"], n, case["moduli_id"], case["limbs"],
                           case["key_moduli_id"], case["key_polys"],
                           case["key_chain"], case["seed"])
    if case["op"] == "rotate":
        out = orc.rotate(A, K, case["moduli_id"], case["key_moduli_id"],
                         case["delta"])
    else:
        out = orc.conjugate(A, K, case["moduli_id"], case["key_moduli_id"])
    assert _digest(out) == case["out_sha256"]


def _automorphism_exponent(n, delta):
    """5^-delta mod 2N, or 2N - 1 for the conjugation."""
    two_n = 2 * n
    shift = delta % two_n
    return two_n - 1 if shift == two_n - 1 else pow(5, two_n - shift, two_n)


@pytest.mark.parametrize("delta", [-1, 0, 1, 2, 5, 11])
def test_automorphism_closed_form(delta):
    """y(X) = x(X^g) in Z_q[X]/(X^N + 1), with g = 5^-delta mod 2N: the
    coefficient of X^(g i mod 2N) picks up x_i with sign (-1)^floor(g i / N).
    Checked by evaluation: y(psi^e) = x(psi^(g e)) for every odd e."""
    log_n = 4
    n = 1 << log_n
    primes = small_ring(log_n, 1, 3, 40)
    orc = Oracle(log_n, primes, 1)
    rng = random.Random(delta + 100)
    x = np.array([[rng.randrange(q) for _ in range(n)] for q in primes[:2]],
                 np.int64)
    # coefficient q, a lazy zero, must come out as 0
    x[0, 3] = primes[0]
    y = orc.automorphism(x[None], 0, delta)[0]
    g = _automorphism_exponent(n, delta)
    for l, q in enumerate(primes[:2]):
        assert ((0 <= y[l]) & (y[l] < q)).all()
        psi = _primitive_root_2n(q, n)
        for e in range(1, 2 * n, 2):
            lhs = sum(int(c) * pow(psi, e * i, q) for i, c in enumerate(y[l]))
            rhs = sum(int(c) * pow(psi, g * e * i, q)
                      for i, c in enumerate(x[l]))
            assert lhs % q == rhs % q, (l, e)


def _primitive_root_2n(q, n):
    for a in range(2, q):
        r = pow(a, (q - 1) // (2 * n), q)
        if pow(r, n, q) == q - 1:
            return r
    raise AssertionError("no root")


@pytest.mark.parametrize("alpha,chain,delta", [(1, 3, 1), (2, 5, 3),
                                               (3, 7, -1)])
def test_rotate_closed_form(alpha, chain, delta):
    """rotate = automorphism of (c0 + floor(sum_p c_p X_p / P), ...): the
    rotation with an arbitrary key equals the key switch of c1 plus c0 and
    then the signed permutation; conjugation permutes first."""
    log_n = 3
    n = 1 << log_n
    primes = small_ring(log_n, alpha, chain + alpha, 40, seed=alpha)
    orc = Oracle(log_n, primes, alpha)
    rng = np.random.Generator(np.random.PCG64(alpha * 10 + chain))
    key_mid = 0
    parts = -(-chain // alpha)
    key = np.stack([np.stack([rng.integers(0, 2 * q, n, dtype=np.uint64)
                              .astype(np.int64) for q in primes])
                    for _ in range(2 * parts)])
    ct = np.stack([np.stack([rng.integers(0, q, n, dtype=np.uint64)
                             .astype(np.int64) for q in primes[:chain]])
                   for _ in range(2)])
    qs = np.array(primes[:chain], np.int64)[:, None]
    if delta == -1:
        got = orc.conjugate(ct, key, 0, key_mid)
        perm = orc.automorphism(ct, 0, -1)
        ks = orc.key_switch(perm[1], key, 0, key_mid)
        expect = np.stack([(ks[0] + perm[0]) % qs, ks[1]])
    else:
        got = orc.rotate(ct, key, 0, key_mid, delta)
        ks = orc.key_switch(ct[1], key, 0, key_mid)
        expect = orc.automorphism(np.stack([(ks[0] + ct[0]) % qs, ks[1]]), 0,
                                  delta)
    assert (got == expect).all()


def test_ntt_permutation_commutes_with_automorphism():
    """NTT(x(X^g)) = NTT(x) permuted by support.ntt_permutation, mod q."""
    from support import ntt_permutation
    log_n = 5
    n = 1 << log_n
    primes = small_ring(log_n, 1, 3, 40)
    orc = Oracle(log_n, primes, 1)
    rng = random.Random(3)
    x = np.array([[rng.randrange(q) for _ in range(n)] for q in primes[:2]],
                 np.int64)[None]
    q = np.array(primes[:2], np.int64)[:, None]
    for delta in (-1, 1, 3, 6):
        lhs = orc.ntt(orc.automorphism(x, 0, delta), 0)[0] % q
        rhs = orc.ntt(x, 0)[0][:, ntt_permutation(n, delta)] % q
        assert (lhs == rhs).all(), delta


@pytest.mark.reference
@pytest.mark.skipif(not Reference.available(),
                    reason="oracle/_ref (compiled reference) not built")
def test_permutation_tables_match_reference():
    """src/ckks/encode.cpp:513-557 tables against the closed forms."""
    from support import automorphism_exponent, ntt_permutation, ptr
    lib = Reference.load_lib()
    for n in (16, 1024, 1 << 16):
        for delta in (-1, 0, 1, 2, 7, n // 4, n // 2 - 1):
            perm = np.zeros(n, np.int32)
            sign = np.zeros(n, np.int32)
            assert lib.ref_permutation_and_sign(n, delta, ptr(perm),
                                                ptr(sign)) == 0
            g = automorphism_exponent(n, delta)
            dst = (g * np.arange(n, dtype=np.int64)) % (2 * n)
            assert (perm == dst % n).all()
            assert (sign == np.where(dst >= n, -1, 1)).all()
            assert lib.ref_ntt_permutation(n, delta, ptr(perm)) == 0
            assert (perm == ntt_permutation(n, delta)).all(), (n, delta)


# ------------------------------------------------ per-coefficient primitives

EW_OPS = list(range(22))


def _ew_operands(q, n, rng):
    """Canonical, lazy and negative residues plus the boundary values."""
    edges = [0, 1, q - 1, q, q + 1, 2 * q - 1, -1, -q + 1, q // 2, q // 2 + 1]
    pool = np.concatenate([
        rng.integers(0, q, n // 3, dtype=np.int64),
        rng.integers(0, 2 * q, n // 3, dtype=np.int64),
        rng.integers(-q + 1, 0, n - 2 * (n // 3), dtype=np.int64)])
    pool[:len(edges)] = edges
    return rng.permutation(pool)


@pytest.mark.reference
@pytest.mark.skipif(not Reference.available(),
                    reason="oracle/_ref (compiled reference) not built")
@pytest.mark.parametrize("op", EW_OPS)
def test_elementwise_matches_reference_functions(op):
    """Each oracle primitive equals the reference's own limb function on the
    same operands (src/montgomery/unit.cpp, compound.cpp, scalar.cpp)."""
    from support import ptr
    import ctypes
    lib = Reference.load_lib()
    lib.ref_elementwise.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_int64]
    olib = Oracle.lib()
    olib.oracle_elementwise_limb.argtypes = [
        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64]
    rng = np.random.default_rng(op)
    for q in (1099516280833, 1152921504606584833):  # 40-bit and 60-bit primes
        k = olib.oracle_calculate_k(q)
        n = 600
        x1, x2 = _ew_operands(q, n, rng), _ew_operands(q, n, rng)
        y0 = rng.integers(0, 2 * q, n, dtype=np.int64)
        for c in (int(rng.integers(0, q)), 0, q - 1):
            expect, got = y0.copy(), y0.copy()
            assert lib.ref_elementwise(op, ptr(x1), ptr(x2), ptr(expect), n,
                                       q, k, c) == 0
            olib.oracle_elementwise_limb(op, ptr(x1), ptr(x2), ptr(got), n, q,
                                         k, c)
            assert (got == expect).all(), (op, q, c)


def test_elementwise_closed_forms():
    """Value-level identities of the primitives, independent of the
    reference: products are a*b*2^-62 mod q, sums and differences are the
    residues of a+b and a-b, exit is x*2^-62 mod q."""
    from support import ptr
    import ctypes
    olib = Oracle.lib()
    olib.oracle_elementwise_limb.argtypes = [
        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64]
    q = 1099516280833
    k = olib.oracle_calculate_k(q)
    rinv = pow(1 << 62, -1, q)
    rng = np.random.default_rng(5)
    n = 256
    a = rng.integers(0, 2 * q, n, dtype=np.int64)
    b = rng.integers(0, 2 * q, n, dtype=np.int64)
    c = int(rng.integers(0, q))

    def run(op):
        y = np.zeros(n, np.int64)
        olib.oracle_elementwise_limb(op, ptr(a), ptr(b), ptr(y), n, q, k, c)
        return [int(v) for v in y]

    A, B = [int(v) for v in a], [int(v) for v in b]
    assert all(v % q == x * y * rinv % q for v, x, y in zip(run(0), A, B))
    assert all(v % q == x * c * rinv % q for v, x in zip(run(1), A))
    for op in (5, 14):
        assert all(v % q == (x + y) % q for v, x, y in zip(run(op), A, B))
    for op in (6, 15):
        assert all(v % q == (x - y) % q for v, x, y in zip(run(op), A, B))
    assert all(0 <= v < q for v in run(14) + run(15) + run(13))
    assert all(v % q == x * c * rinv * rinv % q for v, x in zip(run(12), A))
    assert all(v % q == ((x - y) * c * rinv) % q
               for v, x, y in zip(run(10), A, B))


# ------------------------------------------------ encryption / decryption

def _small_signed(rng, n, bound):
    return rng.integers(-bound, bound + 1, n, dtype=np.int64)


@pytest.mark.reference
@pytest.mark.skipif(not Reference.available(),
                    reason="oracle/_ref (compiled reference) not built")
@pytest.mark.parametrize("log_n,alpha,level", [(13, 1, 2), (14, 2, 5)])
def test_oracle_decrypt_and_encrypt_against_reference(log_n, alpha, level):
    """Oracle decrypt equals the reference's decrypt_to_unit_plaintext limb
    for limb on reference encryptions; an oracle encryption of a message m
    with samples (u, e0, e1) decrypts, through the reference, to m plus a
    noise polynomial whose centered coefficients are small and equal across
    limbs (a dropped or mis-signed term would not be)."""
    ref = Reference(log_n, alpha, threads=4)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    sk_mid, skd = ref.key_data(sk, 3)
    pk_mid, pkd = ref.key_data(pk, 4)
    orc = Oracle(log_n, ref.primes, alpha)
    x = np.random.default_rng(4).uniform(-1, 1, ref.slot_count)
    ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
    mid = ref.shape(ct)[0]
    pmid, expect = ref.decrypt_limbs(ct, sk)
    assert pmid == mid
    assert (orc.decrypt(ref.get(ct), skd[0], sk_mid, mid) == expect).all()

    rng = np.random.default_rng(9)
    n = ref.n
    m = _small_signed(rng, n, 1 << 30)
    u = rng.integers(0, 2, n, dtype=np.int64)
    e0, e1 = _small_signed(rng, n, 20), _small_signed(rng, n, 20)
    qs = ref.primes[mid:mid + orc.limbs_at(mid)].astype(np.int64)[:, None]
    pt = np.where(m < 0, m + qs, m)
    oct_ = orc.encrypt_pk(pt, pkd, pk_mid, u, e0, e1, mid)
    ref.set(ct, oct_)
    _, dec = ref.decrypt_limbs(ct, sk)
    noise = (dec - pt) % qs
    noise = np.where(noise > qs // 2, noise - qs, noise)
    assert (noise == noise[0]).all()
    assert np.abs(noise[0]).max() < 1 << 20
    assert np.abs(noise[0]).max() > 0
    lib.ref_free_ct(ct)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()
