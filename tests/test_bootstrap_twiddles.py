"""The bootstrap's merged butterfly diagonals (host doubles) equal the
reference's get_coeff_to_slot_twiddles / get_slot_to_coeff_twiddles bit for
bit, every stage, for 2^4, 2^9 and 2^15 slots (the last is where glibc's
sincos and separate sin / cos calls first disagree)."""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Reference

pytestmark = pytest.mark.skipif(not Reference.available(),
                                reason="oracle/_ref not built")


@pytest.mark.parametrize("log_slots", [4, 9, 15])
@pytest.mark.parametrize("to_slot", [1, 0])
def test_bootstrap_twiddles_match_reference(log_slots, to_slot):
    ref_lib = Reference.load_lib()
    ref_lib.ref_bootstrap_twiddles.restype = ctypes.c_int
    ref_lib.ref_bootstrap_twiddles.argtypes = [ctypes.c_int] * 4 + \
        [ctypes.c_void_p]
    lib = fhe_amd.library()
    lib.fhe_bootstrap_twiddles.argtypes = [ctypes.c_int] * 4 + [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    slots = 1 << log_slots
    for stage in range(3):
        n = ref_lib.ref_bootstrap_twiddles(log_slots, 3, to_slot, stage, None)
        want = np.zeros((n, slots, 2))
        ref_lib.ref_bootstrap_twiddles(log_slots, 3, to_slot, stage,
                                       want.ctypes.data)
        got = np.zeros((n, slots, 2))
        count = ctypes.c_int64()
        assert lib.fhe_bootstrap_twiddles(log_slots, 3, to_slot, stage,
                                          got.ctypes.data, got.size,
                                          ctypes.byref(count)) == 0
        assert count.value == n
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), stage
