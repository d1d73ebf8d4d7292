"""Writes bootstrap_key_layout.json: the shapes of the reference engine's
BootstrapKey (CKKSEngine::create_bootstrap_key, src/ckks/engine.cpp:162-222)
on the N = 2^16 bootstrap preset with 3 stages, sizes "medium" and "large"
(inner-loop shift 0 and 1): each stage's level and deltas and each rotation
context's delta and key moduli_id.  Runs the reference built under
oracle/_ref (about a minute per key)."""
import ctypes
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from support import Reference, ptr  # noqa: E402


def layout(ref, sk, size, shift):
    lib, e = ref.lib, ref.e
    bk = lib.ref_bootstrap_key(e, sk, 3, size)
    assert bk, lib.ref_last_error()
    out = {"log_n": 16, "special_prime_count": 4, "stage_count": 3,
           "inner_loop_shift": shift}
    for name, to_slot in (("coeff_to_slot", 1), ("slot_to_coeff", 0)):
        stages = []
        for st in range(3):
            meta = np.zeros(3, np.int32)
            inner = np.zeros(256, np.int32)
            outer = np.zeros(256, np.int32)
            assert lib.ref_bootstrap_key_stage(
                bk, to_slot, st, meta.ctypes.data_as(ctypes.c_void_p),
                inner.ctypes.data_as(ctypes.c_void_p),
                outer.ctypes.data_as(ctypes.c_void_p)) == 0
            stages.append({"level": int(meta[0]),
                           "inner_deltas": inner[:meta[1]].tolist(),
                           "outer_deltas": outer[:meta[2]].tolist()})
        out[name] = stages
    info = np.zeros(4, np.int32)
    lib.ref_bootstrap_key_info(bk, info.ctypes.data_as(ctypes.c_void_p))
    rotations = []
    delta = ctypes.c_int()
    for i in range(int(info[2])):
        frk = lib.ref_bootstrap_key_rotation(bk, i, ctypes.byref(delta))
        s = np.zeros(4, np.int64)
        lib.ref_key_shape(frk, 1, ptr(s))
        rotations.append([delta.value, int(s[0])])
    out["rotations_by_delta"] = rotations
    lib.ref_free_bootstrap_key(bk)
    return out


def main():
    ref = Reference(16, 4, bootstrap=1, threads=8)
    sk = ref.lib.ref_secret_key(ref.e)
    keys = [layout(ref, sk, b"medium", 0), layout(ref, sk, b"large", 1)]
    ref.lib.ref_free_secret_key(sk)
    ref.close()
    with open(os.path.join(HERE, "bootstrap_key_layout.json"), "w") as f:
        json.dump(keys, f, indent=1)


if __name__ == "__main__":
    main()
