"""C ABI checks that need no GPU: the library loads, exports every symbol
include/fhe_amd.h declares, and its host-side parameter derivation reproduces
the reference's prime chains (test/ckks/precomputed/default.cpp)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fhe_amd
from support import REPO, golden

HEADER = os.path.join(REPO, "include", "fhe_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fhe_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(fhe_amd.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = fhe_amd.library()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


# tier -> (log_n, alpha, scale bits, small prime bits)
TIERS = {"Iron": (13, 1, 40, 40), "Bronze": (14, 1, 40, 40),
         "Silver": (15, 2, 40, 40), "Gold": (16, 4, 40, 40),
         "BootstrapGold": (16, 4, 54, 54)}


@pytest.mark.parametrize("tier", sorted(TIERS))
def test_prime_chain_matches_reference(tier):
    log_n, alpha, scale, small = TIERS[tier]
    expected = golden("precomputed_tiers.json")["tests"][tier]["arrays"]["q"]
    got = fhe_amd.resolve_primes(log_n, alpha, scale, small)
    assert [int(q) for q in got] == expected


def test_n17_requires_explicit_primes():
    with pytest.raises(fhe_amd.FheError) as err:
        fhe_amd.resolve_primes(17, 6)
    assert err.value.code == 1 and "explicit" in str(err.value)


@pytest.mark.parametrize("log_n,alpha,scale,small", [
    (12, 1, 40, 40),   # unsupported ring degree
    (16, 0, 40, 40),   # no special prime
    (16, 17, 40, 40),  # too many special primes
    (16, 4, 50, 40),   # unsupported scale
])
def test_invalid_parameters_are_rejected(log_n, alpha, scale, small):
    with pytest.raises(fhe_amd.FheError) as err:
        fhe_amd.resolve_primes(log_n, alpha, scale, small)
    assert err.value.code == 1
    assert fhe_amd.library().fhe_last_error()


def test_null_arguments_are_rejected():
    lib = fhe_amd.library()
    assert lib.fhe_context_create(None, None) == 1
    out = (ctypes.c_uint64 * 4)()
    count = ctypes.c_int()
    p = fhe_amd.make_params(16, 4)
    assert lib.fhe_params_resolve_primes(ctypes.byref(p), out, 4,
                                         ctypes.byref(count)) == 1
