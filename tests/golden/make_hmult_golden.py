"""Write HMult+relin+rescale golden vectors from the reference engine.

Run in this container after `make -C oracle ref` (which compiles the
reference from its sources under /root/reference into oracle/_ref/):

    python tests/golden/make_hmult_golden.py

For each case, seeded random inputs (tests/golden/seeded_inputs.py) -- two
rescaled INTT-form ciphertexts at `moduli_id` and a relinearization key at the
key level the reference uses -- are pushed into reference objects and run
through the reference's own multiply(ct, ct, rlk) + rescale(ct).  The fixture
stores the case parameters, the reference prime chain and a SHA-256 digest of
every output limb, so the full-size cases stay a few KB.  Inputs are
regenerated from the seed wherever the fixture is checked.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from seeded_inputs import hmult_inputs  # noqa: E402

REPO = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdesilo_ref.so")

# (name, log_n, special primes, bootstrap preset, moduli ids, seed)
CASES = [
    ("iron", 13, 1, 0, [1], 11),
    ("bronze", 14, 1, 0, [1, 4, 6], 12),
    ("silver", 15, 2, 0, [1, 9, 16], 13),
    ("gold", 16, 4, 0, [1, 13, 30], 14),
    ("n16_alpha8", 16, 8, 0, [8, 14, 26], 15),
]


def load_ref():
    lib = ctypes.CDLL(REF_LIB)
    vp = ctypes.c_void_p
    for name in ["ref_engine_create", "ref_secret_key", "ref_public_key",
                 "ref_relin_key", "ref_encrypt", "ref_hmult_relin_rescale",
                 "ref_rescale"]:
        getattr(lib, name).restype = vp
    lib.ref_last_error.restype = ctypes.c_char_p
    return lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def digest_limbs(arr, polys, limbs, n):
    a = arr.reshape(polys, limbs, n)
    return [[hashlib.sha256(a[p, l].tobytes()).hexdigest()[:32]
             for l in range(limbs)] for p in range(polys)]


def main():
    lib = load_ref()
    vp = ctypes.c_void_p
    out_cases = []
    for name, log_n, alpha, boot, mids, seed in CASES:
        e = lib.ref_engine_create(log_n, alpha, boot, os.cpu_count() or 1)
        if not e:
            raise RuntimeError(lib.ref_last_error().decode())
        info = np.zeros(9, np.int64)
        lib.ref_engine_info(vp(e), ptr(info))
        n, prime_count, chain_count = int(info[0]), int(info[1]), int(info[2])
        primes = np.zeros(prime_count, np.uint64)
        lib.ref_engine_primes(vp(e), ptr(primes))
        sk = lib.ref_secret_key(vp(e))
        pk = lib.ref_public_key(vp(e), vp(sk))
        rlk = lib.ref_relin_key(vp(e), vp(sk))
        kshape = np.zeros(4, np.int64)
        lib.ref_relin_key_shape(vp(rlk), ptr(kshape))
        key_mid, key_polys, key_chain = (int(kshape[0]), int(kshape[1]),
                                         int(kshape[2]))
        slots = int(info[6])
        zeros = np.zeros(slots)
        for mid in mids:
            level = chain_count - mid - 1
            a0 = lib.ref_encrypt(vp(e), ptr(zeros), slots, vp(pk), level)
            b0 = lib.ref_encrypt(vp(e), ptr(zeros), slots, vp(pk), level)
            a = lib.ref_rescale(vp(e), vp(a0))
            b = lib.ref_rescale(vp(e), vp(b0))
            shape = np.zeros(6, np.int64)
            lib.ref_ct_shape(vp(a), ptr(shape))
            assert int(shape[0]) == mid, (name, mid, shape)
            chain = int(shape[2])
            case_seed = seed * 1000 + mid
            A, B, K = hmult_inputs(primes, n, mid, chain, key_mid, key_polys,
                                   key_chain, case_seed)
            lib.ref_ct_set(vp(a), ptr(A))
            lib.ref_ct_set(vp(b), ptr(B))
            lib.ref_relin_key_set(vp(rlk), ptr(K))
            c = lib.ref_hmult_relin_rescale(vp(e), vp(a), vp(b), vp(rlk))
            if not c:
                raise RuntimeError(lib.ref_last_error().decode())
            lib.ref_ct_shape(vp(c), ptr(shape))
            out = np.zeros(2 * (chain - 1) * n, np.int64)
            lib.ref_ct_get(vp(c), ptr(out))
            out_cases.append({
                "name": f"{name}_mid{mid}", "log_n": log_n,
                "special_prime_count": alpha, "bootstrap_preset": boot,
                "primes": [int(q) for q in primes],
                "chain_count": chain_count, "moduli_id": mid,
                "limbs": chain, "key_moduli_id": key_mid,
                "key_polys": key_polys, "key_chain": key_chain,
                "seed": case_seed, "out_moduli_id": int(shape[0]),
                "out_limbs": int(shape[2]),
                "out_sha256": digest_limbs(out, 2, chain - 1, n),
            })
            for h in (a0, b0, a, b, c):
                lib.ref_free_ct(vp(h))
            print(f"{name} mid={mid}: {chain} limbs -> {chain - 1}", flush=True)
        lib.ref_free_relin_key(vp(rlk))
        lib.ref_free_public_key(vp(pk))
        lib.ref_free_secret_key(vp(sk))
        lib.ref_engine_destroy(vp(e))
    with open(os.path.join(HERE, "hmult_golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_hmult_golden.py",
                   "reference_entry": "CKKSEngine::multiply(ct, ct, rlk) "
                   "then CKKSEngine::rescale(ct)", "cases": out_cases}, f,
                  indent=1)


if __name__ == "__main__":
    main()
