"""Write rotation / conjugation golden vectors from the reference engine.

Run in this container after `make -C oracle ref` (which compiles the
reference from its sources under /root/reference into oracle/_ref/):

    python tests/golden/make_rotate_golden.py

For each case, seeded random inputs (tests/golden/seeded_inputs.py) -- one
rescaled INTT-form ciphertext at `moduli_id` and a key-switching key at the
key level the reference uses -- are pushed into reference objects and run
through the reference's own rotate(ct, FixedRotationKey) for each delta and
conjugate(ct, ConjugationKey).  The fixture stores the case parameters, the
prime chain and a SHA-256 digest of every output limb; inputs are
regenerated from the seed wherever the fixture is checked.  The ciphertext is
the `A` array of hmult_inputs(seed) and the key its `K` array.
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

# (name, log_n, special primes, moduli ids, deltas, seed)
CASES = [
    ("iron", 13, 1, [1, 2], [1, 7], 21),
    ("bronze", 14, 1, [4], [1, 1000], 22),
    ("silver", 15, 2, [9], [3, 4096], 23),
    ("gold", 16, 4, [13], [1, 12345], 24),
    ("n16_alpha8", 16, 8, [14], [2], 25),
]

KIND_ROTATION, KIND_CONJUGATION = 1, 2


def load_ref():
    lib = ctypes.CDLL(REF_LIB)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    for name in ["ref_engine_create", "ref_secret_key", "ref_public_key",
                 "ref_encrypt", "ref_rescale", "ref_rotate", "ref_conjugate",
                 "ref_fixed_rotation_key", "ref_conjugation_key"]:
        getattr(lib, name).restype = vp
    lib.ref_fixed_rotation_key.argtypes = [vp, vp, i32]
    lib.ref_conjugation_key.argtypes = [vp, vp]
    lib.ref_rotate.argtypes = [vp, vp, vp]
    lib.ref_conjugate.argtypes = [vp, vp, vp]
    lib.ref_key_shape.argtypes = [vp, i32, vp]
    lib.ref_key_set.argtypes = [vp, i32, vp]
    lib.ref_last_error.restype = ctypes.c_char_p
    return lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def digest_limbs(arr, polys, limbs, n):
    a = arr.reshape(polys, limbs, n)
    return [[hashlib.sha256(a[p, l].tobytes()).hexdigest()[:32]
             for l in range(limbs)] for p in range(polys)]


def run_case(lib, e, primes, n, chain_count, mid, key, kind, seed, op):
    vp = ctypes.c_void_p
    kshape = np.zeros(4, np.int64)
    lib.ref_key_shape(key, kind, ptr(kshape))
    key_mid, key_polys, key_chain = (int(kshape[0]), int(kshape[1]),
                                     int(kshape[2]))
    slots = n // 2
    level = chain_count - mid - 1
    ct0 = lib.ref_encrypt(e, ptr(np.zeros(slots)), slots, vp(op["pk"]), level)
    ct = lib.ref_rescale(e, vp(ct0))
    shape = np.zeros(6, np.int64)
    lib.ref_ct_shape(vp(ct), ptr(shape))
    assert int(shape[0]) == mid, (mid, shape)
    chain = int(shape[2])
    A, _, K = hmult_inputs(primes, n, mid, chain, key_mid, key_polys,
                           key_chain, seed)
    lib.ref_ct_set(vp(ct), ptr(A))
    lib.ref_key_set(key, kind, ptr(K))
    out_ct = op["fn"](vp(ct))
    if not out_ct:
        raise RuntimeError(lib.ref_last_error().decode())
    lib.ref_ct_shape(vp(out_ct), ptr(shape))
    assert int(shape[0]) == mid and int(shape[2]) == chain
    out = np.zeros(2 * chain * n, np.int64)
    lib.ref_ct_get(vp(out_ct), ptr(out))
    for h in (ct0, ct, out_ct):
        lib.ref_free_ct(vp(h))
    return {"moduli_id": mid, "limbs": chain, "key_moduli_id": key_mid,
            "key_polys": key_polys, "key_chain": key_chain, "seed": seed,
            "out_sha256": digest_limbs(out, 2, chain, n)}


def main():
    lib = load_ref()
    vp = ctypes.c_void_p
    out_cases = []
    for name, log_n, alpha, mids, deltas, seed in CASES:
        e = lib.ref_engine_create(log_n, alpha, 0, os.cpu_count() or 1)
        if not e:
            raise RuntimeError(lib.ref_last_error().decode())
        e = vp(e)
        info = np.zeros(9, np.int64)
        lib.ref_engine_info(e, ptr(info))
        n, prime_count, chain_count = int(info[0]), int(info[1]), int(info[2])
        primes = np.zeros(prime_count, np.uint64)
        lib.ref_engine_primes(e, ptr(primes))
        sk = vp(lib.ref_secret_key(e))
        pk = lib.ref_public_key(e, sk)
        common = {"log_n": log_n, "special_prime_count": alpha,
                  "primes": [int(q) for q in primes],
                  "chain_count": chain_count}
        for mid in mids:
            for delta in deltas:
                frk = vp(lib.ref_fixed_rotation_key(e, sk, delta))
                case = run_case(
                    lib, e, primes, n, chain_count, mid, frk, KIND_ROTATION,
                    seed * 1000 + mid * 10,
                    {"pk": pk, "fn": lambda c: lib.ref_rotate(e, c, frk)})
                case.update(common, name=f"{name}_mid{mid}_rot{delta}",
                            op="rotate", delta=delta)
                out_cases.append(case)
                lib.ref_free_fixed_rotation_key(frk)
                print(case["name"], flush=True)
            cjk = vp(lib.ref_conjugation_key(e, sk))
            case = run_case(
                lib, e, primes, n, chain_count, mid, cjk, KIND_CONJUGATION,
                seed * 1000 + mid * 10 + 1,
                {"pk": pk, "fn": lambda c: lib.ref_conjugate(e, c, cjk)})
            case.update(common, name=f"{name}_mid{mid}_conj", op="conjugate",
                        delta=-1)
            out_cases.append(case)
            lib.ref_free_conjugation_key(cjk)
            print(case["name"], flush=True)
        lib.ref_free_public_key(vp(pk))
        lib.ref_free_secret_key(sk)
        lib.ref_engine_destroy(e)
    with open(os.path.join(HERE, "rotate_golden.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_rotate_golden.py",
                   "reference_entry": "CKKSEngine::rotate(ct, "
                   "FixedRotationKey) and CKKSEngine::conjugate(ct, "
                   "ConjugationKey)", "cases": out_cases}, f, indent=1)


if __name__ == "__main__":
    main()
