"""Test-only loaders for the checkers under oracle/.

  Oracle     ctypes binding of oracle/liboracle.so, the repo's C restatement
             of the reference's hot-path arithmetic.
  Reference  ctypes binding of oracle/_ref/libdesilo_ref.so, the reference
             engine compiled from its own sources (present only where it was
             built; tests that need it skip otherwise).

Only tests/, the smoke check and bench.py's cpu_baseline leg use these.
"""
import ctypes
import json
import os
import sys

import numpy as np

TESTS = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(TESTS)
GOLDEN = os.path.join(TESTS, "golden")
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libdesilo_ref.so")

sys.path.insert(0, GOLDEN)
from seeded_inputs import hmult_inputs, uniform_limbs  # noqa: E402,F401

VP = ctypes.c_void_p


def ptr(a):
    return a.ctypes.data_as(VP)


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


class Oracle:
    """One oracle context over a prime chain."""
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_LIB):
                raise RuntimeError("oracle/liboracle.so is not built "
                                   "(make -C oracle oracle)")
            lib = ctypes.CDLL(ORACLE_LIB)
            lib.oracle_ctx_create.restype = VP
            lib.oracle_ctx_create.argtypes = [ctypes.c_int, VP, ctypes.c_int,
                                              ctypes.c_int]
            lib.oracle_ctx_destroy.argtypes = [VP]
            i64, u64 = ctypes.c_int64, ctypes.c_uint64
            lib.oracle_mont_mult.restype = i64
            lib.oracle_mont_mult.argtypes = [i64, i64, u64, u64]
            lib.oracle_mont_exit.restype = i64
            lib.oracle_mont_exit.argtypes = [i64, u64, u64]
            lib.oracle_calculate_k.restype = u64
            lib.oracle_calculate_k.argtypes = [u64]
            lib.oracle_ntt_raw.argtypes = [VP, VP, u64, u64, ctypes.c_int]
            lib.oracle_intt_raw.argtypes = [VP, VP, u64, u64, ctypes.c_int]
            for name in ["oracle_ntt_limb", "oracle_intt_limb"]:
                getattr(lib, name).argtypes = [VP, ctypes.c_int, VP]
            lib.oracle_prime_constants.argtypes = [VP, ctypes.c_int, VP, VP,
                                                   VP]
            for name in ["oracle_to_ntt", "oracle_to_intt"]:
                getattr(lib, name).argtypes = [VP, ctypes.c_int, ctypes.c_int,
                                               VP]
            lib.oracle_mont_mult_limbs.argtypes = [VP, ctypes.c_int,
                                                   ctypes.c_int, VP, VP, VP]
            lib.oracle_tensor.argtypes = [VP, ctypes.c_int, ctypes.c_int, VP,
                                          VP, VP]
            lib.oracle_key_switch.argtypes = [VP, ctypes.c_int, VP, VP,
                                              ctypes.c_int, VP]
            lib.oracle_relinearize.argtypes = [VP, ctypes.c_int, VP, VP,
                                               ctypes.c_int, VP]
            lib.oracle_rescale.argtypes = [VP, ctypes.c_int, ctypes.c_int, VP,
                                           VP]
            lib.oracle_hmult_relin_rescale.argtypes = [VP, ctypes.c_int, VP,
                                                       VP, VP, ctypes.c_int,
                                                       VP]
            lib.oracle_partitions.argtypes = [VP, ctypes.c_int, ctypes.c_int,
                                              VP, VP, VP]
            i32 = ctypes.c_int
            lib.oracle_encrypt_pk.argtypes = [VP, i32, VP, VP, i32, VP, VP,
                                              VP, VP]
            lib.oracle_decrypt.argtypes = [VP, i32, VP, VP, i32, VP]
            lib.oracle_automorphism.argtypes = [VP, i32, i32, i32, i32, VP, VP]
            lib.oracle_rotate.argtypes = [VP, i32, VP, VP, i32, i32, VP]
            lib.oracle_conjugate.argtypes = [VP, i32, VP, VP, i32, VP]
            lib.oracle_switch_key.argtypes = [VP, i32, VP, VP, i32, VP]
            lib.oracle_encode.argtypes = [VP, VP, i32, ctypes.c_double, i32,
                                          i32, i32, VP, VP]
            lib.oracle_encode_pre_permutation.argtypes = [i32, VP]
            lib.oracle_decode.restype = i32
            lib.oracle_decode.argtypes = [VP, VP, i32, ctypes.c_longdouble,
                                          ctypes.c_double, i32, VP]
            cls._lib = lib
        return cls._lib

    def __init__(self, log_n, primes, alpha):
        self.log_n = log_n
        self.n = 1 << log_n
        self.primes = np.array([int(q) for q in primes], dtype=np.uint64)
        self.alpha = alpha
        self.chain_count = len(primes) - alpha
        self.h = self.lib().oracle_ctx_create(log_n, ptr(self.primes),
                                              len(primes), alpha)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().oracle_ctx_destroy(self.h)
            self.h = None

    def limbs_at(self, mid):
        return self.chain_count - mid

    def constants(self, i):
        c = np.zeros(3, np.int64)
        r = np.zeros(self.n - 1, np.int64)
        ri = np.zeros(self.n - 1, np.int64)
        self.lib().oracle_prime_constants(self.h, i, ptr(c), ptr(r), ptr(ri))
        return c, r, ri

    def ntt(self, x, mid, enter=True):
        """[polys][limbs][N] -> NTT form (copy)."""
        y = np.array(x, dtype=np.int64, copy=True)
        polys, limbs = _pl(y)
        flat = y.reshape(polys, limbs, self.n)
        for p in range(polys):
            if enter:
                self.lib().oracle_to_ntt(self.h, mid, limbs, ptr(flat[p]))
            else:
                for l in range(limbs):
                    self.lib().oracle_ntt_limb(self.h, mid + l, ptr(flat[p, l]))
        return y

    def intt(self, x, mid, post=1):
        y = np.array(x, dtype=np.int64, copy=True)
        polys, limbs = _pl(y)
        flat = y.reshape(polys, limbs, self.n)
        for p in range(polys):
            if post == 0:
                for l in range(limbs):
                    self.lib().oracle_intt_limb(self.h, mid + l, ptr(flat[p, l]))
            else:
                self.lib().oracle_to_intt(self.h, mid, limbs, ptr(flat[p]))
                if post == 2:
                    q = self.primes[mid:mid + limbs].astype(np.int64)[:, None]
                    flat[p] = np.where(flat[p] < q, flat[p], flat[p] - q)
        return y

    def mont_mult(self, a, b, mid):
        a = np.ascontiguousarray(a, dtype=np.int64)
        b = np.ascontiguousarray(b, dtype=np.int64)
        polys, limbs = _pl(a)
        out = np.empty_like(a)
        fa, fb, fo = (v.reshape(polys, limbs, self.n) for v in (a, b, out))
        for p in range(polys):
            self.lib().oracle_mont_mult_limbs(self.h, mid, limbs, ptr(fa[p]),
                                              ptr(fb[p]), ptr(fo[p]))
        return out

    def key_switch(self, x, key, mid, key_mid):
        limbs = self.limbs_at(mid)
        out = np.zeros((2, limbs, self.n), np.int64)
        self.lib().oracle_key_switch(self.h, mid, ptr(np.ascontiguousarray(x)),
                                     ptr(np.ascontiguousarray(key)), key_mid,
                                     ptr(out))
        return out

    def relinearize(self, trio, key, mid, key_mid):
        limbs = self.limbs_at(mid)
        out = np.zeros((2, limbs, self.n), np.int64)
        self.lib().oracle_relinearize(self.h, mid,
                                      ptr(np.ascontiguousarray(trio)),
                                      ptr(np.ascontiguousarray(key)), key_mid,
                                      ptr(out))
        return out

    def rescale(self, x, mid):
        limbs = self.limbs_at(mid)
        x = np.ascontiguousarray(x, dtype=np.int64)
        polys = x.size // (limbs * self.n)
        out = np.zeros((polys, limbs - 1, self.n), np.int64)
        self.lib().oracle_rescale(self.h, mid, polys, ptr(x), ptr(out))
        return out

    def hmult(self, a, b, key, mid, key_mid):
        limbs = self.limbs_at(mid)
        out = np.zeros((2, limbs - 1, self.n), np.int64)
        self.lib().oracle_hmult_relin_rescale(
            self.h, mid, ptr(np.ascontiguousarray(a)),
            ptr(np.ascontiguousarray(b)), ptr(np.ascontiguousarray(key)),
            key_mid, ptr(out))
        return out

    def automorphism(self, x, mid, delta):
        """Coefficient-domain automorphism of [polys][limbs][N] limbs."""
        x = np.ascontiguousarray(x, dtype=np.int64)
        polys, limbs = _pl(x)
        out = np.empty_like(x)
        self.lib().oracle_automorphism(self.h, mid, limbs, polys, delta,
                                       ptr(x), ptr(out))
        return out

    def rotate(self, ct, key, mid, key_mid, delta):
        out = np.zeros((2, self.limbs_at(mid), self.n), np.int64)
        self.lib().oracle_rotate(self.h, mid, ptr(np.ascontiguousarray(ct)),
                                 ptr(np.ascontiguousarray(key)), key_mid,
                                 delta, ptr(out))
        return out

    def switch_key(self, ct, key, mid, key_mid):
        out = np.zeros((2, self.limbs_at(mid), self.n), np.int64)
        self.lib().oracle_switch_key(self.h, mid,
                                     ptr(np.ascontiguousarray(ct)),
                                     ptr(np.ascontiguousarray(key)), key_mid,
                                     ptr(out))
        return out

    def conjugate(self, ct, key, mid, key_mid):
        out = np.zeros((2, self.limbs_at(mid), self.n), np.int64)
        self.lib().oracle_conjugate(self.h, mid, ptr(np.ascontiguousarray(ct)),
                                    ptr(np.ascontiguousarray(key)), key_mid,
                                    ptr(out))
        return out

    def encrypt_pk(self, pt, pk, pk_mid, u, e0, e1, mid):
        out = np.zeros((2, self.limbs_at(mid), self.n), np.int64)
        c = [np.ascontiguousarray(v, dtype=np.int64)
             for v in (pt, pk, u, e0, e1)]
        self.lib().oracle_encrypt_pk(self.h, mid, ptr(c[0]), ptr(c[1]), pk_mid,
                                     ptr(c[2]), ptr(c[3]), ptr(c[4]),
                                     ptr(out))
        return out

    def decrypt(self, ct, sk, sk_mid, mid):
        out = np.zeros((self.limbs_at(mid), self.n), np.int64)
        self.lib().oracle_decrypt(self.h, mid, ptr(np.ascontiguousarray(ct)),
                                  ptr(np.ascontiguousarray(sk)), sk_mid,
                                  ptr(out))
        return out

    def encode(self, values, scale_before_tile, first, enc_mid, out_limbs,
               r_mult):
        """values: n/2 doubles or complex numbers."""
        values = np.asarray(values)
        if np.iscomplexobj(values):
            flat = np.stack([values.real, values.imag], -1).astype(np.float64)
        else:
            flat = values.astype(np.float64)
        flat = np.ascontiguousarray(flat)
        r = np.ascontiguousarray(r_mult, dtype=np.int64)
        out = np.zeros((out_limbs, self.n), np.int64)
        self.lib().oracle_encode(self.h, ptr(flat),
                                 1 if np.iscomplexobj(values) else 0,
                                 float(scale_before_tile), first, enc_mid,
                                 out_limbs, ptr(r), ptr(out))
        return out

    def decode(self, pt, mid, ckks_scale, base_scale, complex_output=False):
        """Decode plaintext limbs at moduli_id with a long double scale."""
        slots = self.n // 2
        out = np.zeros(slots * (2 if complex_output else 1))
        buf = np.array([ckks_scale], dtype=np.longdouble)
        rc = self.lib().oracle_decode(
            self.h, ptr(np.ascontiguousarray(pt, np.int64)), mid,
            ctypes.c_longdouble.from_buffer(buf), float(base_scale),
            int(complex_output), ptr(out))
        if rc != 0:
            raise ValueError("oracle_decode refused the plaintext")
        if complex_output:
            return out[0::2] + 1j * out[1::2]
        return out

    def partitions(self, mid, key_mid):
        s = np.zeros(128, np.int32)
        z = np.zeros(128, np.int32)
        k = np.zeros(128, np.int32)
        c = self.lib().oracle_partitions(self.h, mid, key_mid, ptr(s), ptr(z),
                                         ptr(k))
        return [(int(s[i]), int(z[i]), int(k[i])) for i in range(c)]


def automorphism_exponent(n, delta):
    """g of X -> X^g for a rotation by delta: 5^-delta mod 2N, 2N - 1 for
    the conjugation (delta = -1 mod 2N)."""
    two_n = 2 * n
    shift = delta % two_n
    return two_n - 1 if shift == two_n - 1 else pow(5, two_n - shift, two_n)


def ntt_permutation(n, delta):
    """out[i] = in[perm[i]] for NTT-form data: slot i holds the value at
    psi^(2 bitrev(i) + 1), so perm[i] = bitrev((g (2 bitrev(i) + 1) mod 2N
    - 1) / 2)."""
    bits = n.bit_length() - 1
    idx = np.arange(n, dtype=np.int64)
    rev = np.zeros(n, np.int64)
    for b in range(bits):
        rev |= ((idx >> b) & 1) << (bits - 1 - b)
    g = automorphism_exponent(n, delta)
    e = (g * (2 * rev + 1)) % (2 * n)
    return rev[(e - 1) // 2]


def _pl(y):
    limbs = y.shape[-2]
    polys = int(np.prod(y.shape[:-2])) if y.ndim > 2 else 1
    return polys, limbs


class Reference:
    """The reference engine compiled from its sources (oracle/_ref)."""

    @staticmethod
    def available():
        return os.path.exists(REF_LIB)

    @staticmethod
    def load_lib():
        lib = ctypes.CDLL(REF_LIB)
        i32 = ctypes.c_int
        # (restype, argtypes) of every shim entry point used by the tests
        sigs = {
            "ref_engine_create": (VP, [i32, i32, i32, i32]),
            "ref_engine_destroy": (None, [VP]),
            "ref_engine_info": (i32, [VP, VP]),
            "ref_engine_primes": (i32, [VP, VP]),
            "ref_secret_key": (VP, [VP]),
            "ref_public_key": (VP, [VP, VP]),
            "ref_relin_key": (VP, [VP, VP]),
            "ref_encrypt": (VP, [VP, VP, i32, VP, i32]),
            "ref_decrypt": (i32, [VP, VP, VP, VP]),
            "ref_rescale": (VP, [VP, VP]),
            "ref_multiply_relinearize": (VP, [VP, VP, VP, VP]),
            "ref_hmult_relin_rescale": (VP, [VP, VP, VP, VP]),
            "ref_ct_shape": (i32, [VP, VP]),
            "ref_ct_get": (i32, [VP, VP]),
            "ref_ct_set": (i32, [VP, VP]),
            "ref_ct_scale": (ctypes.c_double, [VP]),
            "ref_relin_key_shape": (i32, [VP, VP]),
            "ref_relin_key_get": (i32, [VP, VP]),
            "ref_relin_key_set": (i32, [VP, VP]),
            "ref_free_ct": (None, [VP]),
            "ref_free_relin_key": (None, [VP]),
            "ref_free_public_key": (None, [VP]),
            "ref_free_secret_key": (None, [VP]),
            "ref_ntt_limb": (i32, [VP, i32, VP, i32]),
            "ref_prime_constants": (i32, [VP, i32, VP, VP, VP]),
            "ref_fixed_rotation_key": (VP, [VP, VP, i32]),
            "ref_conjugation_key": (VP, [VP, VP]),
            "ref_free_fixed_rotation_key": (None, [VP]),
            "ref_free_conjugation_key": (None, [VP]),
            "ref_key_shape": (i32, [VP, i32, VP]),
            "ref_key_get": (i32, [VP, i32, VP]),
            "ref_key_set": (i32, [VP, i32, VP]),
            "ref_rotate": (VP, [VP, VP, VP]),
            "ref_conjugate": (VP, [VP, VP, VP]),
            "ref_permutation_and_sign": (i32, [i32, i32, VP, VP]),
            "ref_ntt_permutation": (i32, [i32, i32, VP]),
            "ref_multiply_double": (VP, [VP, VP, ctypes.c_double]),
            "ref_multiply_int": (VP, [VP, VP, ctypes.c_int64]),
            "ref_add_double": (VP, [VP, VP, ctypes.c_double]),
            "ref_level_down": (VP, [VP, VP, i32]),
            "ref_negate": (VP, [VP, VP]),
            "ref_add_ct": (VP, [VP, VP, VP]),
            "ref_subtract_ct": (VP, [VP, VP, VP]),
            "ref_decrypt_limbs": (i32, [VP, VP, VP, VP, VP]),
            "ref_encode": (i32, [VP, VP, i32, i32, VP, VP]),
            "ref_encode_constants": (i32, [VP, i32, VP, VP, VP, VP, VP]),
            "ref_pre_permutation": (i32, [VP, VP]),
            "ref_decode": (i32, [VP, VP, i32, VP, i32, VP]),
            "ref_small_bootstrap_key": (VP, [VP, VP]),
            "ref_free_small_bootstrap_key": (None, [VP]),
            "ref_mod_raise": (VP, [VP, VP, VP]),
            "ref_multiply_imaginary": (VP, [VP, VP, ctypes.c_int64]),
            "ref_ct_flag": (i32, [VP]),
            "ref_ct_scale_ld": (i32, [VP, VP]),
            "ref_encode_pt": (VP, [VP, VP, i32]),
            "ref_free_pt": (None, [VP]),
            "ref_pt_shape": (i32, [VP, VP]),
            "ref_pt_get": (i32, [VP, VP]),
            "ref_ntt_text": (VP, [VP, VP, i32]),
            "ref_rescale_text": (VP, [VP, VP, i32]),
            "ref_level_down_pt": (VP, [VP, VP, i32]),
            "ref_multiply_ct": (VP, [VP, VP, VP]),
            "ref_square": (VP, [VP, VP, VP]),
            "ref_multiply_pt": (VP, [VP, VP, VP]),
            "ref_round": (VP, [VP, VP, VP]),
            "ref_relinearize": (VP, [VP, VP, VP]),
            "ref_rotation_key": (VP, [VP, VP]),
            "ref_free_rotation_key": (None, [VP]),
            "ref_rotation_key_count": (i32, [VP]),
            "ref_rotation_key_fixed": (VP, [VP, i32]),
            "ref_rotate_rotation_key": (VP, [VP, VP, VP, i32]),
            "ref_evaluate_polynomial": (VP, [VP, VP, VP, i32, VP]),
            "ref_sum_1d_by_row": (VP, [VP, VP, VP]),
            "ref_coeff_to_slot": (i32, [VP, VP, VP, VP, i32, VP, VP]),
            "ref_slot_to_coeff": (VP, [VP, VP, VP, VP, VP, i32]),
            "ref_bootstrap": (VP, [VP, VP, VP, VP, VP, VP, i32]),
            "ref_fixed_rotation_key_delta": (i32, [VP]),
            "ref_bootstrap_key": (VP, [VP, VP, i32, ctypes.c_char_p]),
            "ref_free_bootstrap_key": (None, [VP]),
            "ref_bootstrap_key_info": (i32, [VP, VP]),
            "ref_bootstrap_key_rotation": (VP, [VP, i32, VP]),
            "ref_bootstrap_key_small": (VP, [VP]),
            "ref_bootstrap_key_stage": (i32, [VP, i32, i32, VP, VP, VP]),
            "ref_bootstrap_key_diagonal": (VP, [VP, i32, i32, i32, i32]),
            "ref_coeff_to_slot_bk": (i32, [VP, VP, VP, VP, VP, VP]),
            "ref_slot_to_coeff_bk": (VP, [VP, VP, VP, VP, VP]),
            "ref_bootstrap_bk": (VP, [VP, VP, VP, VP, VP]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib.ref_last_error.restype = ctypes.c_char_p
        return lib

    def __init__(self, log_n, alpha, bootstrap=0, threads=1):
        lib = self.load_lib()
        self.lib = lib
        self.e = lib.ref_engine_create(log_n, alpha, bootstrap, threads)
        if not self.e:
            raise RuntimeError(lib.ref_last_error().decode())
        info = np.zeros(9, np.int64)
        lib.ref_engine_info(self.e, ptr(info))
        self.n = int(info[0])
        self.prime_count = int(info[1])
        self.chain_count = int(info[2])
        self.slot_count = int(info[6])
        q = np.zeros(self.prime_count, np.uint64)
        lib.ref_engine_primes(self.e, ptr(q))
        self.primes = q

    def close(self):
        if self.e:
            self.lib.ref_engine_destroy(self.e)
            self.e = None

    def shape(self, ct):
        s = np.zeros(6, np.int64)
        self.lib.ref_ct_shape(ct, ptr(s))
        return [int(v) for v in s]

    def get(self, ct):
        mid, polys, chain, n, _, _ = self.shape(ct)
        out = np.zeros((polys, chain, n), np.int64)
        self.lib.ref_ct_get(ct, ptr(out))
        return out

    def pt_shape(self, pt):
        s = np.zeros(6, np.int64)
        self.lib.ref_pt_shape(pt, ptr(s))
        return [int(v) for v in s]

    def pt_get(self, pt):
        mid, polys, chain, n, _, _ = self.pt_shape(pt)
        out = np.zeros((polys, chain, n), np.int64)
        self.lib.ref_pt_get(pt, ptr(out))
        return out

    def set(self, ct, data):
        self.lib.ref_ct_set(ct, ptr(np.ascontiguousarray(data,
                                                         dtype=np.int64)))

    def key_data(self, key, kind):
        """(moduli_id, limbs array) of a key object (kind as ref_key_shape:
        0 relin, 1 rotation, 2 conjugation, 3 secret, 4 public)."""
        s = np.zeros(4, np.int64)
        self.lib.ref_key_shape(key, kind, ptr(s))
        out = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
        self.lib.ref_key_get(key, kind, ptr(out))
        return int(s[0]), out

    def decrypt_limbs(self, ct, sk):
        s = np.zeros(4, np.int64)
        self.lib.ref_decrypt_limbs(self.e, ct, sk, ptr(s), None)
        out = np.zeros((int(s[2]), int(s[3])), np.int64)
        self.lib.ref_decrypt_limbs(self.e, ct, sk, ptr(s), ptr(out))
        return int(s[0]), out

    def encode(self, values, level):
        """CKKSEngine::encode_unit -> (moduli_id, limbs [chain][N])."""
        values = np.asarray(values)
        is_complex = np.iscomplexobj(values)
        if is_complex:
            flat = np.stack([values.real, values.imag], -1).astype(np.float64)
        else:
            flat = values.astype(np.float64)
        flat = np.ascontiguousarray(flat)
        s = np.zeros(4, np.int64)
        rc = self.lib.ref_encode(self.e, ptr(flat), int(is_complex), level,
                                 ptr(s), None)
        if rc != 0 or s[3] == 0:
            raise RuntimeError(self.lib.ref_last_error().decode())
        out = np.zeros((int(s[2]), int(s[3])), np.int64)
        self.lib.ref_encode(self.e, ptr(flat), int(is_complex), level, ptr(s),
                            ptr(out))
        return int(s[0]), out

    def encode_constants(self, level):
        ckks, sbt = np.zeros(1, np.longdouble), np.zeros(1)
        log_after = np.zeros(1, np.int64)
        r = np.zeros(512, np.int64)
        count = np.zeros(1, np.int32)
        top = self.lib.ref_encode_constants(self.e, level, ptr(ckks), ptr(sbt),
                                            ptr(log_after), ptr(r),
                                            ptr(count))
        return dict(ckks_scale=ckks[0], scale_before_tile=float(sbt[0]),
                    log_scale_after_tile=int(log_after[0]),
                    top_of_one_more_level=bool(top),
                    r_mult_scale_after_tile=r[:int(count[0])])

    def decode(self, limbs, level, ckks_scale, complex_output=False):
        """CKKSEngine::decode_unit of plaintext limbs [chain][N] at `level`
        with a long double scale."""
        scale = np.array([ckks_scale], dtype=np.longdouble)
        out = np.zeros(self.slot_count * (2 if complex_output else 1))
        rc = self.lib.ref_decode(self.e,
                                 ptr(np.ascontiguousarray(limbs, np.int64)),
                                 level, ptr(scale), int(complex_output),
                                 ptr(out))
        if rc != self.slot_count:
            raise RuntimeError(self.lib.ref_last_error().decode())
        if complex_output:
            return out[0::2] + 1j * out[1::2]
        return out

    def ct_scale(self, ct):
        out = np.zeros(1, np.longdouble)
        self.lib.ref_ct_scale_ld(ct, ptr(out))
        return out[0]

    def pre_permutation(self):
        out = np.zeros(self.n, np.uint64)
        self.lib.ref_pre_permutation(self.e, ptr(out))
        return out

    def key(self, rlk):
        s = np.zeros(4, np.int64)
        self.lib.ref_relin_key_shape(rlk, ptr(s))
        out = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
        self.lib.ref_relin_key_get(rlk, ptr(out))
        return int(s[0]), out
