"""Host constants of the scalar operations: fhe_params_multiplier against
the reference's own CKKSPrecomputed multipliers (oracle/_ref), for the
default and bootstrap presets, every kind, positive / negative / fractional
values and every level where the kind is defined.  No GPU is needed."""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.reference,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

VALUES = [1.0, -1.0, 0.5, -0.37, 3.14159, 1e-3, -123.456, 0.0]
INTS = [1, -1, 7, -123456789, 0]


def ref_multiplier(ref, kind, value, int_value, level, level2):
    lib = ref.lib
    lib.ref_multiplier.restype = ctypes.c_int
    lib.ref_multiplier.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int64,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    out = np.zeros(512, np.int64)
    count = lib.ref_multiplier(ref.e, kind, value, int_value, level, level2,
                               ptr(out))
    return out[:count]


@pytest.mark.parametrize("log_n,alpha,boot", [(13, 1, 0), (14, 2, 0),
                                              (16, 4, 0), (16, 4, 1)])
def test_multipliers_match_reference(log_n, alpha, boot):
    ref = Reference(log_n, alpha, bootstrap=boot)
    kw = dict(scale_bits=54 if boot else 40,
              small_prime_bits=54 if boot else 40)
    info = np.zeros(9, np.int64)
    ref.lib.ref_engine_info(ref.e, ptr(info))
    full = int(info[4])
    primes = fhe_amd.resolve_primes(log_n, alpha, **kw)
    assert (primes == ref.primes).all()
    levels = sorted({0, 1, full // 2, full - 1, full})
    small = log_n < 16
    for level in levels:
        for v in (VALUES if small else VALUES[::3]):
            got = fhe_amd.params_multiplier(log_n, alpha, "ADD", level, v,
                                            **kw)
            assert (got == ref_multiplier(ref, 0, v, 0, level, 0)).all(), \
                ("add", level, v)
            if level >= 1:
                got = fhe_amd.params_multiplier(log_n, alpha, "MULT", level,
                                                v, **kw)
                assert (got == ref_multiplier(ref, 1, v, 0, level, 0)).all(), \
                    ("mult", level, v)
        for iv in (INTS if small else INTS[::2]):
            mid = full - level
            got = fhe_amd.params_multiplier(log_n, alpha, "MULT_INT", mid,
                                            int_value=iv, **kw)
            assert (got == ref_multiplier(ref, 2, 0.0, iv, mid, 0)).all(), \
                ("int", mid, iv)
        dsts = range(level) if small else {0, level // 2, level - 1}
        for dst in sorted(d for d in dsts if 0 <= d < level):
            got = fhe_amd.params_multiplier(log_n, alpha, "LEVEL_DOWN", level,
                                            level2=dst, **kw)
            assert (got == ref_multiplier(ref, 3, 0, 0, level, dst)).all(), \
                ("level_down", level, dst)
            got = fhe_amd.params_multiplier(log_n, alpha,
                                            "LEVEL_DOWN_AFTER_RESCALE", level,
                                            level2=dst, **kw)
            assert (got == ref_multiplier(ref, 4, 0, 0, level, dst)).all(), \
                ("level_down_after_rescale", level, dst)
    ref.close()
