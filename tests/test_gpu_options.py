"""Every pipeline choice of fhe_context_set_option gives the same limbs: the
batched Gold-size HMult+relin+rescale under each option equals the default
pipeline's output bit for bit (the default is itself checked against the
reference digests and the oracle in test_gpu_hmult.py)."""
import numpy as np
import pytest

import fhe_amd
from support import golden, hmult_inputs

pytestmark = pytest.mark.gpu

DEFAULTS = {"split_variants": 1, "own_limbs": 1, "overlap_trio": 1,
            "staged_modup": 1,
            "ks_lazy_accumulate": 1, "ks_wide_accumulate": 0,
            "ks_prefetch": 0, "ks_occupancy": 2,
            "lazy_key_switch": 1, "fuse_tensor": 0, "fuse_modup": 0}
OPTIONS = [{"split_variants": 0}, {"own_limbs": 0}, {"overlap_trio": 0},
           {"overlap_trio": 2},
           {"staged_modup": 0},
           {"ks_lazy_accumulate": 0}, {"ks_wide_accumulate": 1},
           {"ks_wide_accumulate": 1, "ks_lazy_accumulate": 0},
           {"ks_prefetch": 1},
           {"ks_occupancy": 3},
           {"lazy_key_switch": 0}, {"fuse_tensor": 1},
           {"staged_modup": 0, "fuse_modup": 1},
           {"split_variants": 0, "own_limbs": 0, "lazy_key_switch": 0}]


def test_options_bit_identical(gpu_lib):
    case = [c for c in golden("hmult_golden.json")["cases"]
            if c["name"] == "gold_mid30"][0]
    n = 1 << case["log_n"]
    limbs = case["limbs"]
    A, B, K = hmult_inputs(case["primes"], n, case["moduli_id"], limbs,
                           case["key_moduli_id"], case["key_polys"],
                           case["key_chain"], case["seed"])
    with fhe_amd.Context(case["log_n"], case["special_prime_count"],
                         primes=case["primes"]) as ctx:
        count = 3
        da = ctx.to_device(np.stack([A] * count))
        db = ctx.to_device(np.stack([B] * count))
        dk = ctx.to_device(K.reshape(case["key_polys"], case["key_chain"], n))
        out = ctx.array((count, 2, limbs - 1, n))

        def run():
            ctx.hmult_relin_rescale(da, db, dk, case["key_moduli_id"], out,
                                    case["moduli_id"])
            return out.download()

        want = run()
        for opts in OPTIONS:
            print(f"options {opts}", flush=True)
            for name, value in opts.items():
                ctx.set_option(name, value)
            got = run()
            for name in opts:
                ctx.set_option(name, DEFAULTS[name])
            assert np.array_equal(got, want), opts
        with pytest.raises(fhe_amd.FheError):
            ctx.set_option("no_such_option", 1)
        for arr in (da, db, dk, out):
            arr.free()
