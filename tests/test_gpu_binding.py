"""The reference-side binding compiled (oracle/gpu_engine.cpp, GpuCKKSEngine
built against the reference's headers and library): multiply(ct, ct, rlk)
through the binding runs on the device and returns the same limbs,
moduli_id, level, flag and scale as CKKSEngine::multiply on the same
rescaled encryptions, and decrypts to x * y; get_mode() is "gpu"."""
import ctypes
import os

import pytest

from support import REPO

BINDING = os.path.join(REPO, "oracle", "_ref", "libgpu_engine.so")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(BINDING),
                                 reason="oracle/_ref/libgpu_engine.so not built")]


@pytest.mark.parametrize("log_n,alpha,level", [(14, 1, 5), (16, 4, 20)])
def test_binding_multiply_matches_reference(log_n, alpha, level, gpu_lib):
    lib = ctypes.CDLL(BINDING)
    lib.gpu_engine_multiply_check.argtypes = [
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_char_p, ctypes.c_int]
    lib.gpu_engine_last_error.restype = ctypes.c_char_p
    mode = ctypes.create_string_buffer(16)
    rc = lib.gpu_engine_multiply_check(log_n, alpha, level, 1e-4, mode, 16)
    assert rc == 0, (rc, lib.gpu_engine_last_error())
    assert mode.value == b"gpu"
