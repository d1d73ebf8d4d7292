"""Python binding of the MI355X CKKS backend's C ABI (include/fhe_amd.h).

The product is the native library libfhe_amd.so; this module only loads it
with ctypes and wraps device buffers for tests and the benchmark.  There is
no fallback: if the library is missing or cannot be loaded, import of the
binding raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_PKG, "libfhe_amd.so")

FHE_OK = 0
INFO_FIELDS = ("coeff_count", "prime_count", "chain_count",
               "special_prime_count", "full_level", "max_level",
               "max_key_switching_level", "slot_count", "scale_bits")

# every entry point declared in include/fhe_amd.h
EXPORTED_SYMBOLS = (
    "fhe_context_create", "fhe_context_destroy", "fhe_last_error",
    "fhe_context_info", "fhe_context_primes", "fhe_context_prime_constants",
    "fhe_context_stream", "fhe_synchronize", "fhe_device_alloc",
    "fhe_device_free", "fhe_copy_to_device", "fhe_copy_to_host", "fhe_ntt",
    "fhe_intt", "fhe_mont_mult", "fhe_rescale", "fhe_relinearize",
    "fhe_hmult_relin_rescale", "fhe_params_resolve_primes",
    "fhe_automorphism", "fhe_rotate", "fhe_conjugate", "fhe_profile_enable",
    "fhe_profile_count", "fhe_profile_entry", "fhe_elementwise",
    "fhe_multiplier", "fhe_params_multiplier", "fhe_multiply_double",
    "fhe_multiply_int", "fhe_add_double", "fhe_level_down",
    "fhe_encrypt_pk", "fhe_decrypt", "fhe_rotate_batch", "fhe_switch_key",
    "fhe_encode", "fhe_params_encode_constants", "fhe_decode",
    "fhe_decoded_scale", "fhe_context_set_option", "fhe_multiply_imaginary",
    "fhe_mod_raise", "fhe_multiply", "fhe_multiply_plain", "fhe_add_texts",
    "fhe_round", "fhe_coeff_to_slot", "fhe_slot_to_coeff", "fhe_bootstrap",
    "fhe_bootstrap_twiddles", "fhe_wire_size", "fhe_wire_write",
    "fhe_wire_read_header", "fhe_wire_read", "fhe_modmul_peak",
    "fhe_relinearize_uneven", "fhe_rotate_uneven",
    "fhe_rotate_rotation_key_uneven", "fhe_context_uneven_autotune",
    "fhe_uneven_schedule", "fhe_graph_begin", "fhe_graph_end",
    "fhe_graph_launch", "fhe_graph_destroy", "fhe_evaluate_polynomial",
    "fhe_sum_1d_by_row", "fhe_serve", "fhe_hmult_relin_rescale_noalloc",
    "fhe_bootstrap_key_create", "fhe_bootstrap_key_destroy",
    "fhe_coeff_to_slot_bk", "fhe_slot_to_coeff_bk", "fhe_bootstrap_bk",
    "fhe_bootstrap_key_stage_layout", "fhe_bootstrap_key_rotation_layout",
    "fhe_bootstrap_key_encode_stage",
)

# fhe_wire_* record kinds (FHE_WIRE_* in include/fhe_amd.h)
WIRE = dict(PLAINTEXT=0, CIPHERTEXT=1, KEY=2)


class WireHeader(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("polys", ctypes.c_int),
                ("moduli_id", ctypes.c_int), ("limbs", ctypes.c_int),
                ("level", ctypes.c_int), ("flags", ctypes.c_int),
                ("scale", ctypes.c_longdouble),
                ("payload_bytes", ctypes.c_size_t)]

# data flags of a text (FHE_FLAG_* in include/fhe_amd.h)
FLAG_NTT_FORM = 2
FLAG_RESCALED = 4


# fhe_elementwise operation codes (FHE_EW_* in include/fhe_amd.h)
EW = dict(MONT_MULT=0, MONT_ENTER=1, REDUCE_2Q_TO_Q=2, MAKE_SIGNED=3,
          NEGATE=4, ADD=5, SUBTRACT=6, ENTER_ADD=7, ENTER_ADD_REDUCE=8,
          MULT_ADD=9, SUB_ENTER_REDUCE=10, SUB_ENTER=11, ENTER_EXIT=12,
          ENTER_EXIT_REDUCE=13, ADD_REDUCE=14, SUB_REDUCE=15,
          MAKE_UNSIGNED_REDUCE=16, NEGATE_MAKE_UNSIGNED=17, ADD_SCALAR=18,
          ROUNDER=19, ENTER_REDUCE=20, ADD_SCALAR_REDUCE=21)

# fhe_multiplier kinds (FHE_MULTIPLIER_*)
MULTIPLIER = dict(ADD=0, MULT=1, MULT_INT=2, LEVEL_DOWN=3,
                  LEVEL_DOWN_AFTER_RESCALE=4)


class FheError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"fhe error {code}: {message}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("log_coeff_count", ctypes.c_int32),
                ("special_prime_count", ctypes.c_int32),
                ("scale_bits", ctypes.c_int32),
                ("small_prime_bits", ctypes.c_int32),
                ("one_more_level", ctypes.c_int32),
                ("max_level", ctypes.c_int32),
                ("prime_count", ctypes.c_int32),
                ("primes", ctypes.POINTER(ctypes.c_uint64))]


_lib = None


class BootstrapKeys(ctypes.Structure):
    """fhe_bootstrap_keys: device keys of a bootstrap with a RotationKey."""
    _fields_ = [("relin", ctypes.c_void_p), ("relin_moduli_id", ctypes.c_int),
                ("conjugation", ctypes.c_void_p),
                ("conjugation_moduli_id", ctypes.c_int),
                ("rotation_count", ctypes.c_int),
                ("rotation", ctypes.POINTER(ctypes.c_void_p)),
                ("rotation_moduli_ids", ctypes.POINTER(ctypes.c_int)),
                ("rotation_deltas", ctypes.POINTER(ctypes.c_int)),
                ("to_sparse", ctypes.c_void_p),
                ("to_sparse_moduli_id", ctypes.c_int),
                ("to_dense", ctypes.c_void_p),
                ("to_dense_moduli_id", ctypes.c_int)]

    @classmethod
    def make(cls, relin, relin_mid, conj, conj_mid, rotations, to_sparse,
             sparse_mid, to_dense, dense_mid):
        """rotations: [(device array, moduli_id, delta)] in the key's
        order; the arrays stay owned by the caller."""
        m = len(rotations)
        k = cls()
        k._keep = ((ctypes.c_void_p * m)(*[r[0].ptr.value for r in rotations]),
                   (ctypes.c_int * m)(*[r[1] for r in rotations]),
                   (ctypes.c_int * m)(*[r[2] for r in rotations]))
        k.relin, k.relin_moduli_id = relin.ptr.value, relin_mid
        k.conjugation, k.conjugation_moduli_id = conj.ptr.value, conj_mid
        k.rotation_count = m
        k.rotation = ctypes.cast(k._keep[0], ctypes.POINTER(ctypes.c_void_p))
        k.rotation_moduli_ids = ctypes.cast(k._keep[1],
                                            ctypes.POINTER(ctypes.c_int))
        k.rotation_deltas = ctypes.cast(k._keep[2],
                                        ctypes.POINTER(ctypes.c_int))
        k.to_sparse = to_sparse.ptr.value if to_sparse is not None else None
        k.to_sparse_moduli_id = sparse_mid
        k.to_dense = to_dense.ptr.value if to_dense is not None else None
        k.to_dense_moduli_id = dense_mid
        return k


class PlainMatrix(ctypes.Structure):
    """fhe_plain_matrix: one butterfly stage of a BootstrapKey."""
    _fields_ = [("level", ctypes.c_int), ("inner_count", ctypes.c_int),
                ("outer_count", ctypes.c_int),
                ("inner_deltas", ctypes.POINTER(ctypes.c_int)),
                ("outer_deltas", ctypes.POINTER(ctypes.c_int)),
                ("diagonals", ctypes.POINTER(ctypes.c_void_p))]

    @classmethod
    def make(cls, level, inner_deltas, outer_deltas, diagonals):
        """diagonals: {(outer, inner): device array or address}, every pair
        but (0, 0)."""
        inner, outer = len(inner_deltas), len(outer_deltas)
        table = (ctypes.c_void_p * (inner * outer))()
        for (o, i), d in diagonals.items():
            table[o * inner + i] = _addr(d)
        m = cls()
        m._keep = (_ints(inner_deltas), _ints(outer_deltas), table)
        m.level, m.inner_count, m.outer_count = level, inner, outer
        m.inner_deltas = ctypes.cast(m._keep[0], ctypes.POINTER(ctypes.c_int))
        m.outer_deltas = ctypes.cast(m._keep[1], ctypes.POINTER(ctypes.c_int))
        m.diagonals = ctypes.cast(table, ctypes.POINTER(ctypes.c_void_p))
        return m


class BootstrapKeyParts(ctypes.Structure):
    """fhe_bootstrap_key_parts: a BootstrapKey's parts on the device."""
    _fields_ = [("stage_count", ctypes.c_int),
                ("coeff_to_slot", ctypes.POINTER(PlainMatrix)),
                ("slot_to_coeff", ctypes.POINTER(PlainMatrix)),
                ("rotation_count", ctypes.c_int),
                ("rotation_deltas", ctypes.POINTER(ctypes.c_int)),
                ("rotation_keys", ctypes.POINTER(ctypes.c_void_p)),
                ("rotation_moduli_ids", ctypes.POINTER(ctypes.c_int)),
                ("to_sparse", ctypes.c_void_p),
                ("to_sparse_moduli_id", ctypes.c_int),
                ("to_dense", ctypes.c_void_p),
                ("to_dense_moduli_id", ctypes.c_int)]

    @classmethod
    def make(cls, cts, stc, rotations, to_sparse=None, sparse_mid=0,
             to_dense=None, dense_mid=0):
        """cts, stc: lists of PlainMatrix (either may be empty); rotations:
        [(delta, device key, moduli_id)]."""
        k = cls()
        stages = max(len(cts), len(stc))
        assert not cts or not stc or len(cts) == len(stc)
        ca = (PlainMatrix * max(1, len(cts)))(*cts)
        sa = (PlainMatrix * max(1, len(stc)))(*stc)
        m = len(rotations)
        k._keep = (ca, sa, list(cts) + list(stc),
                   _ints([r[0] for r in rotations]),
                   _addrs([r[1] for r in rotations]),
                   _ints([r[2] for r in rotations]))
        k.stage_count = stages
        k.coeff_to_slot = (ctypes.cast(ca, ctypes.POINTER(PlainMatrix))
                           if cts else None)
        k.slot_to_coeff = (ctypes.cast(sa, ctypes.POINTER(PlainMatrix))
                           if stc else None)
        k.rotation_count = m
        k.rotation_deltas = ctypes.cast(k._keep[3],
                                        ctypes.POINTER(ctypes.c_int))
        k.rotation_keys = ctypes.cast(k._keep[4],
                                      ctypes.POINTER(ctypes.c_void_p))
        k.rotation_moduli_ids = ctypes.cast(k._keep[5],
                                            ctypes.POINTER(ctypes.c_int))
        k.to_sparse = _addr(to_sparse) if to_sparse is not None else None
        k.to_sparse_moduli_id = sparse_mid
        k.to_dense = _addr(to_dense) if to_dense is not None else None
        k.to_dense_moduli_id = dense_mid
        return k


class RotationKey(ctypes.Structure):
    """fhe_rotation_key: a RotationKey's fixed keys on the device."""
    _fields_ = [("count", ctypes.c_int),
                ("keys", ctypes.POINTER(ctypes.c_void_p)),
                ("moduli_ids", ctypes.POINTER(ctypes.c_int)),
                ("deltas", ctypes.POINTER(ctypes.c_int))]

    @classmethod
    def make(cls, keys, moduli_ids, deltas):
        """keys: device arrays (owned by the caller) in the key's order."""
        m = len(keys)
        k = cls()
        k._keep = ((ctypes.c_void_p * m)(*[_addr(x) for x in keys]),
                   (ctypes.c_int * m)(*moduli_ids),
                   (ctypes.c_int * m)(*deltas))
        k.count = m
        k.keys = ctypes.cast(k._keep[0], ctypes.POINTER(ctypes.c_void_p))
        k.moduli_ids = ctypes.cast(k._keep[1], ctypes.POINTER(ctypes.c_int))
        k.deltas = ctypes.cast(k._keep[2], ctypes.POINTER(ctypes.c_int))
        return k


JOB_DOT, JOB_POLY = 0, 1   # FHE_JOB_* in include/fhe_amd.h


class Job(ctypes.Structure):
    """fhe_job: one request of the serving loop."""
    _fields_ = [("engine", ctypes.c_int), ("kind", ctypes.c_int),
                ("level", ctypes.c_int), ("a", ctypes.c_void_p),
                ("b", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("arrival_ms", ctypes.c_double),
                ("admitted_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("out_level", ctypes.c_int),
                ("out_flags", ctypes.c_int)]


class ServeEngine(ctypes.Structure):
    """fhe_serve_engine: a context with its resident keys."""
    _fields_ = [("ctx", ctypes.c_void_p), ("rlk", ctypes.c_void_p),
                ("rlk_moduli_id", ctypes.c_int), ("rotation", RotationKey),
                ("coefficients", ctypes.c_void_p),
                ("coefficient_count", ctypes.c_int)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int), ("max_in_flight", ctypes.c_int),
                ("reserve_bytes", ctypes.c_size_t)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("batches", ctypes.c_int),
                ("max_batch_jobs", ctypes.c_int),
                ("deferred_admissions", ctypes.c_int),
                ("min_free_bytes", ctypes.c_size_t)]


def serve(engines, jobs, max_batch=16, max_in_flight=2,
          reserve_bytes=8 << 30):
    """fhe_serve over `engines` (ServeEngine) and `jobs` (a ctypes array of
    Job, updated in place); returns the ServeStats."""
    e_arr = (ServeEngine * len(engines))(*engines)
    opts = ServeOptions(max_batch, max_in_flight, reserve_bytes)
    stats = ServeStats()
    _check(library().fhe_serve(e_arr, len(engines), jobs, len(jobs),
                               ctypes.byref(opts), ctypes.byref(stats)))
    return stats


def library():
    """Load libfhe_amd.so once; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(
                f"{LIBRARY_PATH} is not built; run python -m fhe_amd.build")
        lib = ctypes.CDLL(LIBRARY_PATH)
        vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
        lib.fhe_last_error.restype = ctypes.c_char_p
        lib.fhe_context_stream.restype = vp
        lib.fhe_context_stream.argtypes = [vp]
        lib.fhe_context_destroy.argtypes = [vp]
        lib.fhe_context_create.argtypes = [ctypes.POINTER(Params),
                                           ctypes.POINTER(vp)]
        lib.fhe_params_resolve_primes.argtypes = [
            ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_uint64), i32,
            ctypes.POINTER(i32)]
        lib.fhe_context_info.argtypes = [vp, vp, i32]
        lib.fhe_context_primes.argtypes = [vp, vp, i32]
        lib.fhe_context_prime_constants.argtypes = [vp, i32, vp, vp, vp]
        lib.fhe_synchronize.argtypes = [vp]
        lib.fhe_device_alloc.argtypes = [vp, sz, ctypes.POINTER(vp)]
        lib.fhe_device_free.argtypes = [vp, vp]
        lib.fhe_copy_to_device.argtypes = [vp, vp, vp, sz]
        lib.fhe_copy_to_host.argtypes = [vp, vp, vp, sz]
        lib.fhe_ntt.argtypes = [vp, vp, i32, i32, i32, i32]
        lib.fhe_intt.argtypes = [vp, vp, i32, i32, i32, i32]
        lib.fhe_mont_mult.argtypes = [vp, vp, vp, vp, i32, i32, i32]
        lib.fhe_rescale.argtypes = [vp, vp, vp, i32, i32]
        lib.fhe_relinearize.argtypes = [vp, i32, i32, vp, vp, i32, vp]
        lib.fhe_hmult_relin_rescale.argtypes = [vp, i32, i32, vp, vp, vp, i32,
                                                vp]
        lib.fhe_automorphism.argtypes = [vp, vp, vp, i32, i32, i32, i32, i32]
        lib.fhe_rotate.argtypes = [vp, i32, i32, vp, vp, i32, i32, vp]
        lib.fhe_conjugate.argtypes = [vp, i32, i32, vp, vp, i32, vp]
        lib.fhe_profile_enable.argtypes = [vp, i32]
        lib.fhe_elementwise.argtypes = [vp, i32, vp, vp, vp, i32, i32, i32, vp]
        lib.fhe_multiplier.argtypes = [vp, i32, ctypes.c_double,
                                       ctypes.c_int64, i32, i32, vp, i32,
                                       ctypes.POINTER(i32)]
        lib.fhe_multiply_double.argtypes = [vp, vp, vp, i32, i32,
                                            ctypes.c_double]
        lib.fhe_multiply_int.argtypes = [vp, vp, vp, i32, i32, ctypes.c_int64]
        lib.fhe_add_double.argtypes = [vp, vp, vp, i32, i32, i32,
                                       ctypes.c_double]
        lib.fhe_level_down.argtypes = [vp, vp, vp, i32, i32, i32, i32]
        lib.fhe_encrypt_pk.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, vp,
                                       vp]
        lib.fhe_decrypt.argtypes = [vp, i32, i32, vp, vp, i32, vp]
        lib.fhe_rotate_batch.argtypes = [vp, i32, i32, vp, i32, vp, vp, vp,
                                         vp]
        lib.fhe_switch_key.argtypes = [vp, i32, i32, vp, vp, i32, vp]
        lib.fhe_relinearize_uneven.argtypes = [vp, i32, vp, vp, vp, i32, vp]
        lib.fhe_rotate_uneven.argtypes = [vp, i32, vp, vp, vp, i32, i32, vp]
        lib.fhe_rotate_rotation_key_uneven.argtypes = [vp, i32, vp, vp, i32,
                                                       vp, vp, vp, vp, vp]
        lib.fhe_context_uneven_autotune.argtypes = [vp, i32]
        lib.fhe_uneven_schedule.argtypes = [vp, i32, i32, vp, i32, vp, vp,
                                            i32]
        ip = ctypes.POINTER(i32)
        lib.fhe_evaluate_polynomial.argtypes = [vp, i32, vp, i32, vp, i32, vp,
                                                i32, vp, ip, ip]
        lib.fhe_sum_1d_by_row.argtypes = [vp, i32, vp, i32, i32,
                                          ctypes.POINTER(RotationKey), vp,
                                          ip, ip]
        lib.fhe_serve.argtypes = [vp, i32, vp, i32, vp, vp]
        lib.fhe_hmult_relin_rescale_noalloc.argtypes = [vp, i32, i32, vp, vp,
                                                        vp, i32, vp]
        lib.fhe_graph_begin.argtypes = [vp]
        lib.fhe_graph_end.argtypes = [vp, ctypes.POINTER(vp)]
        lib.fhe_graph_launch.argtypes = [vp, vp]
        lib.fhe_graph_destroy.argtypes = [vp]
        lib.fhe_graph_destroy.restype = None
        lib.fhe_params_multiplier.argtypes = [
            ctypes.POINTER(Params), i32, ctypes.c_double, ctypes.c_int64, i32,
            i32, vp, i32, ctypes.POINTER(i32)]
        lib.fhe_encode.argtypes = [vp, vp, i32, i32, i32, vp]
        lib.fhe_context_set_option.argtypes = [vp, ctypes.c_char_p, i32]
        lib.fhe_multiply_imaginary.argtypes = [vp, i32, i32, vp,
                                               ctypes.c_int64, vp]
        lib.fhe_mod_raise.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp]
        lib.fhe_multiply.argtypes = [vp, i32, vp, i32, i32, vp, i32, i32, vp,
                                     i32, vp]
        lib.fhe_multiply_plain.argtypes = [vp, i32, vp, i32, i32, i32, vp,
                                           i32, i32, vp]
        lib.fhe_add_texts.argtypes = [vp, i32, vp, i32, i32, i32, vp, i32, i32,
                                      i32, i32, vp]
        lib.fhe_round.argtypes = [vp, i32, vp, i32, vp, i32, vp,
                                  ctypes.POINTER(i32), ctypes.POINTER(i32)]
        sz_p = ctypes.POINTER(ctypes.c_size_t)
        lib.fhe_wire_size.argtypes = [vp, i32, i32, sz_p]
        lib.fhe_wire_write.argtypes = [vp, i32, vp, i32, i32, i32, i32, i32,
                                       ctypes.c_longdouble, vp, sz, sz_p]
        lib.fhe_wire_read_header.argtypes = [vp, vp, sz,
                                             ctypes.POINTER(WireHeader)]
        lib.fhe_wire_read.argtypes = [vp, vp, sz, vp, sz]
        bk = ctypes.POINTER(BootstrapKeys)
        lib.fhe_coeff_to_slot.argtypes = [vp, i32, vp, i32, i32, bk, i32, vp,
                                          vp,
                                          ctypes.POINTER(i32),
                                          ctypes.POINTER(i32)]
        lib.fhe_slot_to_coeff.argtypes = [vp, i32, vp, vp, i32, i32, bk, i32,
                                          vp,
                                          ctypes.POINTER(i32),
                                          ctypes.POINTER(i32)]
        lib.fhe_bootstrap.argtypes = [vp, i32, vp, i32, i32, bk, i32, vp,
                                      ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fhe_bootstrap_key_stage_layout.argtypes = [
            ctypes.POINTER(Params), i32, i32, i32, i32, ctypes.POINTER(i32),
            ctypes.POINTER(i32), ctypes.POINTER(i32), vp, vp, i32]
        lib.fhe_bootstrap_key_rotation_layout.argtypes = [
            ctypes.POINTER(Params), i32, i32, vp, vp, i32,
            ctypes.POINTER(i32)]
        lib.fhe_bootstrap_key_encode_stage.argtypes = [
            vp, i32, i32, i32, i32, vp, ctypes.c_int64, ctypes.POINTER(i32)]
        lib.fhe_bootstrap_key_create.argtypes = [
            vp, ctypes.POINTER(BootstrapKeyParts), ctypes.POINTER(vp)]
        lib.fhe_bootstrap_key_destroy.argtypes = [vp]
        lib.fhe_bootstrap_key_destroy.restype = None
        lib.fhe_coeff_to_slot_bk.argtypes = [
            vp, i32, vp, i32, i32, vp, i32, vp, vp, vp, ctypes.POINTER(i32),
            ctypes.POINTER(i32)]
        lib.fhe_slot_to_coeff_bk.argtypes = [
            vp, i32, vp, vp, i32, i32, vp, i32, vp, vp, ctypes.POINTER(i32),
            ctypes.POINTER(i32)]
        lib.fhe_bootstrap_bk.argtypes = [
            vp, i32, vp, i32, i32, vp, i32, vp, i32, vp, vp,
            ctypes.POINTER(i32), ctypes.POINTER(i32)]
        lib.fhe_decode.argtypes = [vp, vp, i32, i32, ctypes.c_longdouble,
                                   i32, vp]
        lib.fhe_decoded_scale.argtypes = [vp, i32, ctypes.c_longdouble,
                                          ctypes.POINTER(i32), vp]
        lib.fhe_params_encode_constants.argtypes = [
            ctypes.POINTER(Params), i32, vp,
            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
            ctypes.POINTER(i32), vp, i32, ctypes.POINTER(i32)]
        lib.fhe_modmul_peak.argtypes = [vp, i32,
                                        ctypes.POINTER(ctypes.c_double)]
        lib.fhe_profile_count.argtypes = [vp, ctypes.POINTER(i32)]
        lib.fhe_profile_entry.argtypes = [
            vp, i32, ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_double),
            ctypes.POINTER(ctypes.c_int64)]
        _lib = lib
    return _lib


def _addr(x):
    """Device address of a DeviceArray, or an int address as it is."""
    return x.ptr.value if isinstance(x, DeviceArray) else int(x)


def _addrs(xs):
    return (ctypes.c_void_p * len(xs))(*[_addr(x) for x in xs])


def _ints(xs):
    return (ctypes.c_int * len(xs))(*[int(x) for x in xs])


def _longdouble(x):
    """An exact ctypes long double from a numpy long double (ctypes' own
    constructor goes through a Python float)."""
    buf = np.array([x], dtype=np.longdouble)
    return ctypes.c_longdouble.from_buffer(buf)


def _check(code):
    if code != FHE_OK:
        raise FheError(code, library().fhe_last_error().decode())


def make_params(log_n, special_prime_count, scale_bits=40, small_prime_bits=40,
                one_more_level=True, max_level=-1, primes=None):
    p = Params(log_n, special_prime_count, scale_bits, small_prime_bits,
               1 if one_more_level else 0, max_level, 0, None)
    if primes is not None:
        arr = (ctypes.c_uint64 * len(primes))(*[int(q) for q in primes])
        p.prime_count = len(primes)
        p.primes = ctypes.cast(arr, ctypes.POINTER(ctypes.c_uint64))
        p._keep = arr
    return p


def resolve_primes(log_n, special_prime_count, scale_bits=40,
                   small_prime_bits=40, one_more_level=True):
    """Host-only: the prime chain the context would use (no GPU needed)."""
    p = make_params(log_n, special_prime_count, scale_bits, small_prime_bits,
                    one_more_level)
    out = (ctypes.c_uint64 * 512)()
    count = ctypes.c_int(0)
    _check(library().fhe_params_resolve_primes(ctypes.byref(p), out, 512,
                                               ctypes.byref(count)))
    return np.array(out[:count.value], dtype=np.uint64)


def params_multiplier(log_n, special_prime_count, kind, level, value=0.0,
                      int_value=0, level2=0, scale_bits=40,
                      small_prime_bits=40, one_more_level=True, primes=None):
    """Host-only fhe_params_multiplier (no GPU needed)."""
    p = make_params(log_n, special_prime_count, scale_bits, small_prime_bits,
                    one_more_level, -1, primes)
    out = np.zeros(512, np.int64)
    count = ctypes.c_int(0)
    _check(library().fhe_params_multiplier(
        ctypes.byref(p), MULTIPLIER[kind], float(value), int(int_value),
        level, level2, out.ctypes.data, 512, ctypes.byref(count)))
    return out[:count.value]


def params_encode_constants(log_n, special_prime_count, level, scale_bits=40,
                            small_prime_bits=40, one_more_level=True,
                            primes=None):
    """Host-only fhe_params_encode_constants: dict of ckks_scale,
    scale_before_tile, log_scale_after_tile, top_of_one_more_level and
    r_mult_scale_after_tile."""
    p = make_params(log_n, special_prime_count, scale_bits, small_prime_bits,
                    one_more_level, -1, primes)
    out = np.zeros(512, np.int64)
    count = ctypes.c_int(0)
    ckks_scale, sbt = np.zeros(1, np.longdouble), ctypes.c_double()
    log_after, top = ctypes.c_int64(), ctypes.c_int()
    _check(library().fhe_params_encode_constants(
        ctypes.byref(p), level, ckks_scale.ctypes.data, ctypes.byref(sbt),
        ctypes.byref(log_after), ctypes.byref(top), out.ctypes.data, 512,
        ctypes.byref(count)))
    return dict(ckks_scale=ckks_scale[0],
                scale_before_tile=sbt.value,
                log_scale_after_tile=log_after.value,
                top_of_one_more_level=bool(top.value),
                r_mult_scale_after_tile=out[:count.value])


def bootstrap_key_layout(log_n, special_prime_count, stage_count=3,
                         inner_loop_shift=0, scale_bits=54,
                         small_prime_bits=54, one_more_level=True,
                         primes=None):
    """Host-only fhe_bootstrap_key_stage_layout / _rotation_layout: dict
    with "coeff_to_slot" and "slot_to_coeff" (per stage: level, inner and
    outer deltas) and "rotations" ([(delta, moduli_id)] in creation
    order)."""
    p = make_params(log_n, special_prime_count, scale_bits, small_prime_bits,
                    one_more_level, -1, primes)
    lib = library()
    out = {}
    for name, to_slot in (("coeff_to_slot", 1), ("slot_to_coeff", 0)):
        stages = []
        for st in range(stage_count):
            level, ni, no = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            inner = np.zeros(4096, np.int32)
            outer = np.zeros(4096, np.int32)
            _check(lib.fhe_bootstrap_key_stage_layout(
                ctypes.byref(p), stage_count, inner_loop_shift, to_slot, st,
                ctypes.byref(level), ctypes.byref(ni), ctypes.byref(no),
                inner.ctypes.data, outer.ctypes.data, 4096))
            stages.append(dict(level=level.value,
                               inner_deltas=inner[:ni.value].tolist(),
                               outer_deltas=outer[:no.value].tolist()))
        out[name] = stages
    count = ctypes.c_int()
    _check(lib.fhe_bootstrap_key_rotation_layout(
        ctypes.byref(p), stage_count, inner_loop_shift, None, None, 0,
        ctypes.byref(count)))
    deltas = np.zeros(max(1, count.value), np.int32)
    mids = np.zeros(max(1, count.value), np.int32)
    _check(lib.fhe_bootstrap_key_rotation_layout(
        ctypes.byref(p), stage_count, inner_loop_shift, deltas.ctypes.data,
        mids.ctypes.data, count.value, ctypes.byref(count)))
    out["rotations"] = list(zip(deltas[:count.value].tolist(),
                                mids[:count.value].tolist()))
    return out


class DeviceArray:
    """A device int64 buffer owned by a Context."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        ptr = ctypes.c_void_p()
        _check(library().fhe_device_alloc(ctx.handle, self.nbytes,
                                          ctypes.byref(ptr)))
        self.ptr = ptr

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.int64)
        assert host.nbytes == self.nbytes, (host.shape, self.shape)
        _check(library().fhe_copy_to_device(self.ctx.handle, self.ptr,
                                            host.ctypes.data, self.nbytes))
        return self

    def upload_at(self, host, offset):
        """Copy host int64 words into the buffer from word `offset` on."""
        host = np.ascontiguousarray(host, dtype=np.int64)
        assert 0 <= offset and offset * 8 + host.nbytes <= self.nbytes
        _check(library().fhe_copy_to_device(
            self.ctx.handle, ctypes.c_void_p(self.ptr.value + offset * 8),
            host.ctypes.data, host.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.int64)
        _check(library().fhe_copy_to_host(self.ctx.handle, out.ctypes.data,
                                          self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            _check(library().fhe_device_free(self.ctx.handle, self.ptr))
        self.ptr = None


class Context:
    """fhe_context wrapper: parameters + device tables + stream."""

    def __init__(self, log_n, special_prime_count, scale_bits=40,
                 small_prime_bits=40, one_more_level=True, max_level=-1,
                 primes=None):
        self._params = make_params(log_n, special_prime_count, scale_bits,
                                   small_prime_bits, one_more_level, max_level,
                                   primes)
        h = ctypes.c_void_p()
        _check(library().fhe_context_create(ctypes.byref(self._params),
                                            ctypes.byref(h)))
        self.handle = h
        info = np.zeros(len(INFO_FIELDS), np.int64)
        _check(library().fhe_context_info(h, info.ctypes.data, len(info)))
        self.info = dict(zip(INFO_FIELDS, (int(v) for v in info)))
        self.n = self.info["coeff_count"]
        self.log_n = self.n.bit_length() - 1
        self.alpha = self.info["special_prime_count"]
        self.chain_count = self.info["chain_count"]
        q = np.zeros(self.info["prime_count"], np.uint64)
        _check(library().fhe_context_primes(h, q.ctypes.data, len(q)))
        self.primes = q

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            library().fhe_context_destroy(self.handle)
        self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def stream(self):
        return library().fhe_context_stream(self.handle)

    def limbs_at(self, moduli_id):
        return self.chain_count - moduli_id

    def key_moduli_id(self):
        return self.info["full_level"] - self.info["max_key_switching_level"]

    def key_shape(self, key_moduli_id):
        limbs = self.limbs_at(key_moduli_id)
        parts = -(-limbs // self.alpha)
        return (2 * parts, limbs + self.alpha, self.n)

    def array(self, shape):
        return DeviceArray(self, shape)

    def to_device(self, host):
        host = np.ascontiguousarray(host, dtype=np.int64)
        return DeviceArray(self, host.shape).upload(host)

    def synchronize(self):
        _check(library().fhe_synchronize(self.handle))

    def prime_constants(self, index):
        c = np.zeros(3, np.int64)
        root = np.zeros(self.n - 1, np.int64)
        rinv = np.zeros(self.n - 1, np.int64)
        _check(library().fhe_context_prime_constants(
            self.handle, index, c.ctypes.data, root.ctypes.data,
            rinv.ctypes.data))
        return c, root, rinv

    # -- batched primitives on device arrays --------------------------------
    def ntt(self, arr, moduli_id, enter=True):
        polys, limbs = _polys_limbs(arr, self.n)
        _check(library().fhe_ntt(self.handle, arr.ptr, polys, limbs,
                                 moduli_id, 1 if enter else 0))

    def intt(self, arr, moduli_id, post=1):
        polys, limbs = _polys_limbs(arr, self.n)
        _check(library().fhe_intt(self.handle, arr.ptr, polys, limbs,
                                  moduli_id, post))

    def mont_mult(self, a, b, out, moduli_id):
        polys, limbs = _polys_limbs(a, self.n)
        _check(library().fhe_mont_mult(self.handle, a.ptr, b.ptr, out.ptr,
                                       polys, limbs, moduli_id))

    def rescale(self, src, dst, moduli_id):
        polys = int(np.prod(src.shape[:-2]))
        _check(library().fhe_rescale(self.handle, src.ptr, dst.ptr, polys,
                                     moduli_id))

    def relinearize(self, trios, key, key_moduli_id, out, moduli_id):
        count = int(np.prod(trios.shape[:-3]))
        _check(library().fhe_relinearize(self.handle, count, moduli_id,
                                         trios.ptr, key.ptr, key_moduli_id,
                                         out.ptr))

    def hmult_relin_rescale(self, a, b, key, key_moduli_id, out, moduli_id):
        count = int(np.prod(a.shape[:-3]))
        _check(library().fhe_hmult_relin_rescale(self.handle, count, moduli_id,
                                                 a.ptr, b.ptr, key.ptr,
                                                 key_moduli_id, out.ptr))

    def multiply(self, a, a_level, a_flags, b, b_level, b_flags, out,
                 key=None, key_moduli_id=0):
        """fhe_multiply: duos a, b in any form (level, FLAG_* bits); out is
        the NTT-form trio, or with a relinearization key the INTT-form duo,
        at level min(a_level, b_level) - 1."""
        count = int(np.prod(a.shape[:-3]))
        _check(library().fhe_multiply(self.handle, count, a.ptr, a_level,
                                      a_flags, b.ptr, b_level, b_flags,
                                      key.ptr if key is not None else None,
                                      key_moduli_id, out.ptr))

    def multiply_plain(self, ct, ct_level, ct_flags, pt, pt_level, pt_flags,
                       out):
        """fhe_multiply_plain: duos or trios ct [count][polys][limbs][N]
        times plaintexts pt [count][limbs][N]; out in NTT form."""
        count = int(np.prod(ct.shape[:-3]))
        polys = int(ct.shape[-3])
        _check(library().fhe_multiply_plain(self.handle, count, ct.ptr, polys,
                                            ct_level, ct_flags, pt.ptr,
                                            pt_level, pt_flags, out.ptr))

    def add(self, a, a_level, a_flags, b, b_level, b_flags, out,
            subtract=False):
        """fhe_add_texts: a + b (or a - b) of duos or trios in any form."""
        count = int(np.prod(a.shape[:-3]))
        _check(library().fhe_add_texts(self.handle, count, a.ptr,
                                       int(a.shape[-3]), a_level, a_flags,
                                       b.ptr, int(b.shape[-3]), b_level,
                                       b_flags, 1 if subtract else 0,
                                       out.ptr))

    def round(self, ct, level, key, key_moduli_id, out):
        """fhe_round (EvalMod) of one INTT-form, not rescaled duo; returns
        the result's (level, flags)."""
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_round(self.handle, _count(ct), ct.ptr, level,
                                   key.ptr,
                                   key_moduli_id, out.ptr, ctypes.byref(lv),
                                   ctypes.byref(fl)))
        return lv.value, fl.value

    def coeff_to_slot(self, ct, level, flags, keys, stage_count, real, imag):
        """fhe_coeff_to_slot of one duo; returns the parts' (level, flags)."""
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_coeff_to_slot(
            self.handle, _count(ct), ct.ptr, level, flags, ctypes.byref(keys),
            stage_count, real.ptr, imag.ptr, ctypes.byref(lv),
            ctypes.byref(fl)))
        return lv.value, fl.value

    def slot_to_coeff(self, real, imag, level, flags, keys, stage_count,
                      out):
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_slot_to_coeff(
            self.handle, _count(real), real.ptr, imag.ptr, level, flags, ctypes.byref(keys),
            stage_count, out.ptr, ctypes.byref(lv), ctypes.byref(fl)))
        return lv.value, fl.value

    def bootstrap(self, ct, level, flags, keys, stage_count, out):
        """fhe_bootstrap of one INTT-form duo; returns (level, flags)."""
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_bootstrap(
            self.handle, _count(ct), ct.ptr, level, flags, ctypes.byref(keys),
            stage_count, out.ptr, ctypes.byref(lv), ctypes.byref(fl)))
        return lv.value, fl.value

    def bootstrap_key(self, parts):
        """fhe_bootstrap_key_create from BootstrapKeyParts; returns the
        handle (free it with bootstrap_key_destroy before the context)."""
        h = ctypes.c_void_p()
        _check(library().fhe_bootstrap_key_create(
            self.handle, ctypes.byref(parts), ctypes.byref(h)))
        h._parts = parts
        return h

    def bootstrap_key_stage(self, stage_count, inner_loop_shift, to_slot,
                            stage):
        """fhe_bootstrap_key_encode_stage: the stage's encoded diagonals as
        one device array [count][limbs + alpha][N] (outer-major, without
        (0, 0))."""
        count = ctypes.c_int()
        _check(library().fhe_bootstrap_key_encode_stage(
            self.handle, stage_count, inner_loop_shift, to_slot, stage, None,
            0, ctypes.byref(count)))
        level = bootstrap_key_layout(self.log_n, self.alpha, stage_count,
                                     inner_loop_shift,
                                     primes=self.primes)[
            "coeff_to_slot" if to_slot else "slot_to_coeff"][stage]["level"]
        mid = self.info["full_level"] - level + 1
        ext = self.info["prime_count"] - mid
        arr = self.array((count.value, ext, self.n))
        _check(library().fhe_bootstrap_key_encode_stage(
            self.handle, stage_count, inner_loop_shift, to_slot, stage,
            arr.ptr, int(np.prod(arr.shape)), ctypes.byref(count)))
        return arr

    def bootstrap_key_destroy(self, bk):
        library().fhe_bootstrap_key_destroy(bk)

    def coeff_to_slot_bk(self, ct, level, flags, conj, conj_mid, bk, real,
                         imag):
        """fhe_coeff_to_slot_bk; returns the parts' (level, flags)."""
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_coeff_to_slot_bk(
            self.handle, _count(ct), ct.ptr, level, flags, _addr(conj),
            conj_mid, bk, real.ptr, imag.ptr, ctypes.byref(lv),
            ctypes.byref(fl)))
        return lv.value, fl.value

    def slot_to_coeff_bk(self, real, imag, level, flags, conj, conj_mid, bk,
                         out):
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_slot_to_coeff_bk(
            self.handle, _count(real), real.ptr, imag.ptr, level, flags,
            _addr(conj), conj_mid, bk, out.ptr, ctypes.byref(lv),
            ctypes.byref(fl)))
        return lv.value, fl.value

    def bootstrap_bk(self, ct, level, flags, relin, relin_mid, conj, conj_mid,
                     bk, out):
        """fhe_bootstrap_bk of INTT-form duos; returns (level, flags)."""
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_bootstrap_bk(
            self.handle, _count(ct), ct.ptr, level, flags, _addr(relin),
            relin_mid, _addr(conj), conj_mid, bk, out.ptr, ctypes.byref(lv),
            ctypes.byref(fl)))
        return lv.value, fl.value

    def modmul_peak(self, form=0):
        """fhe_modmul_peak: modular products per second of the device (form
        0 Shoup, 1 Montgomery)."""
        v = ctypes.c_double()
        _check(library().fhe_modmul_peak(self.handle, form, ctypes.byref(v)))
        return v.value

    def wire_write(self, kind, arr, moduli_id, level, flags, scale):
        """Serialize a device array [polys][limbs][N] into bytes."""
        polys, limbs = int(np.prod(arr.shape[:-2])), int(arr.shape[-2])
        size = ctypes.c_size_t()
        _check(library().fhe_wire_size(self.handle, polys, limbs,
                                       ctypes.byref(size)))
        buf = (ctypes.c_uint8 * size.value)()
        written = ctypes.c_size_t()
        _check(library().fhe_wire_write(
            self.handle, kind, arr.ptr, polys, moduli_id, limbs, level, flags,
            _longdouble(scale), buf, size.value, ctypes.byref(written)))
        return bytes(buf)

    def wire_header(self, data):
        """The record's header as a dict; the scale as np.longdouble (a
        ctypes c_longdouble field reads back as a double)."""
        h = WireHeader()
        _check(library().fhe_wire_read_header(self.handle, data, len(data),
                                              ctypes.byref(h)))
        raw = (ctypes.c_char * 16).from_address(ctypes.addressof(h) +
                                                WireHeader.scale.offset)
        scale = np.frombuffer(bytes(raw), dtype=np.longdouble, count=1)[0]
        return dict(kind=h.kind, polys=h.polys, moduli_id=h.moduli_id,
                    limbs=h.limbs, level=h.level, flags=h.flags, scale=scale,
                    payload_bytes=h.payload_bytes)

    def wire_read(self, data, out):
        _check(library().fhe_wire_read(self.handle, data, len(data), out.ptr,
                                       out.nbytes))

    def automorphism(self, src, dst, moduli_id, delta, ntt_form=False):
        polys, limbs = _polys_limbs(src, self.n)
        _check(library().fhe_automorphism(self.handle, src.ptr, dst.ptr, polys,
                                          limbs, moduli_id, delta,
                                          1 if ntt_form else 0))

    def rotate(self, cts, key, key_moduli_id, delta, out, moduli_id):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_rotate(self.handle, count, moduli_id, cts.ptr,
                                    key.ptr, key_moduli_id, delta, out.ptr))

    def rotate_batch(self, cts, keys, key_moduli_ids, deltas, outs,
                     moduli_id):
        """Hoisted rotations of cts by every (key, key_moduli_id, delta)."""
        count = int(np.prod(cts.shape[:-3]))
        m = len(keys)
        kp = (ctypes.c_void_p * m)(*[k.ptr.value for k in keys])
        op = (ctypes.c_void_p * m)(*[o.ptr.value for o in outs])
        km = (ctypes.c_int * m)(*key_moduli_ids)
        dl = (ctypes.c_int * m)(*deltas)
        _check(library().fhe_rotate_batch(self.handle, count, moduli_id,
                                          cts.ptr, m, kp, km, dl, op))

    # --- batches at uneven levels: one device buffer per ciphertext, given
    # as DeviceArrays or raw device addresses
    def relinearize_uneven(self, trios, moduli_ids, key, key_moduli_id,
                           outs):
        """fhe_relinearize_uneven: NTT-form trios[c] at moduli_ids[c] into
        INTT-form duos outs[c]."""
        n = len(moduli_ids)
        _check(library().fhe_relinearize_uneven(
            self.handle, n, _ints(moduli_ids), _addrs(trios), key.ptr,
            key_moduli_id, _addrs(outs)))

    def rotate_uneven(self, cts, moduli_ids, key, key_moduli_id, delta, outs):
        """fhe_rotate_uneven: INTT-form duos rotated by one fixed key."""
        n = len(moduli_ids)
        _check(library().fhe_rotate_uneven(
            self.handle, n, _ints(moduli_ids), _addrs(cts), key.ptr,
            key_moduli_id, delta, _addrs(outs)))

    def rotate_rotation_key_uneven(self, cts, moduli_ids, keys,
                                   key_moduli_ids, key_deltas, deltas, outs):
        """fhe_rotate_rotation_key_uneven: ciphertext c rotated by deltas[c]
        with the RotationKey's fixed keys."""
        n = len(moduli_ids)
        _check(library().fhe_rotate_rotation_key_uneven(
            self.handle, n, _ints(moduli_ids), _addrs(cts), len(keys),
            _addrs(keys), _ints(key_moduli_ids), _ints(key_deltas),
            _ints(deltas), _addrs(outs)))

    def uneven_autotune(self, on):
        _check(library().fhe_context_uneven_autotune(self.handle,
                                                     1 if on else 0))

    def uneven_schedule(self, relin, moduli_ids, key_moduli_id):
        """Per-stage (cap, xcd) schedule of a batch shape that has run."""
        caps = (ctypes.c_int * 8)()
        xcds = (ctypes.c_int * 8)()
        _check(library().fhe_uneven_schedule(
            self.handle, 1 if relin else 0, len(moduli_ids),
            _ints(moduli_ids), key_moduli_id, caps, xcds, 8))
        return list(zip(caps, xcds))

    def evaluate_polynomial(self, cts, level, coefficients, rlk, rlk_mid,
                            out):
        """fhe_evaluate_polynomial of INTT-form duos [count][2][limbs][N] at
        `level`; returns the result's (level, flags)."""
        count = int(np.prod(cts.shape[:-3]))
        c = np.ascontiguousarray(coefficients, dtype=np.float64)
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_evaluate_polynomial(
            self.handle, count, cts.ptr, level, c.ctypes.data, len(c),
            rlk.ptr, rlk_mid, out.ptr, ctypes.byref(lv), ctypes.byref(fl)))
        return lv.value, fl.value

    def sum_1d_by_row(self, cts, level, flags, rotation_key, out):
        """fhe_sum_1d_by_row with a RotationKey; returns (level, flags)."""
        count = int(np.prod(cts.shape[:-3]))
        lv, fl = ctypes.c_int(), ctypes.c_int()
        _check(library().fhe_sum_1d_by_row(
            self.handle, count, cts.ptr, level, flags,
            ctypes.byref(rotation_key), out.ptr, ctypes.byref(lv),
            ctypes.byref(fl)))
        return lv.value, fl.value

    def graph_begin(self):
        _check(library().fhe_graph_begin(self.handle))

    def graph_end(self):
        g = ctypes.c_void_p()
        _check(library().fhe_graph_end(self.handle, ctypes.byref(g)))
        return g

    def graph_launch(self, graph):
        _check(library().fhe_graph_launch(self.handle, graph))

    @staticmethod
    def graph_destroy(graph):
        library().fhe_graph_destroy(graph)

    def switch_key(self, cts, key, key_moduli_id, out, moduli_id):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_switch_key(self.handle, count, moduli_id,
                                        cts.ptr, key.ptr, key_moduli_id,
                                        out.ptr))

    def conjugate(self, cts, key, key_moduli_id, out, moduli_id):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_conjugate(self.handle, count, moduli_id, cts.ptr,
                                       key.ptr, key_moduli_id, out.ptr))

    def elementwise(self, op, x1, y, moduli_id, x2=None, constants=None):
        """y = op(x1, x2, constants[limb]) per coefficient (EW_* codes)."""
        polys, limbs = _polys_limbs(x1, self.n)
        c = None
        if constants is not None:
            c = np.ascontiguousarray(constants, dtype=np.int64)
            assert c.shape == (limbs,)
        _check(library().fhe_elementwise(
            self.handle, op, x1.ptr, x2.ptr if x2 is not None else None,
            y.ptr, polys, limbs, moduli_id,
            c.ctypes.data if c is not None else None))

    def multiplier(self, kind, level, value=0.0, int_value=0, level2=0):
        """Per-limb scalars of a CKKSPrecomputed multiplier (MULTIPLIER
        kinds); host-only."""
        prime_count = self.info["prime_count"]
        out = np.zeros(prime_count, np.int64)
        count = ctypes.c_int()
        _check(library().fhe_multiplier(self.handle, MULTIPLIER[kind],
                                        float(value), int(int_value), level,
                                        level2, out.ctypes.data,
                                        prime_count, ctypes.byref(count)))
        return out[:count.value]

    def multiply_double(self, src, dst, level, value):
        polys = int(np.prod(src.shape[:-2]))
        _check(library().fhe_multiply_double(self.handle, src.ptr, dst.ptr,
                                             polys, level, float(value)))

    def multiply_int(self, src, dst, moduli_id, value):
        polys = int(np.prod(src.shape[:-2]))
        _check(library().fhe_multiply_int(self.handle, src.ptr, dst.ptr,
                                          polys, moduli_id, int(value)))

    def add_double(self, src, dst, level, value):
        """src: [count][polys][limbs][N]"""
        count, polys = int(np.prod(src.shape[:-3])), src.shape[-3]
        _check(library().fhe_add_double(self.handle, src.ptr, dst.ptr, count,
                                        polys, level, float(value)))

    def level_down(self, src, dst, level, dst_level, rescaled=False):
        polys = int(np.prod(src.shape[:-2]))
        _check(library().fhe_level_down(self.handle, src.ptr, dst.ptr, polys,
                                        level, 1 if rescaled else 0,
                                        dst_level))

    def encrypt_pk(self, pt, pk, pk_moduli_id, u, e0, e1, out, moduli_id):
        count = int(np.prod(pt.shape[:-2]))
        _check(library().fhe_encrypt_pk(self.handle, count, moduli_id, pt.ptr,
                                        pk.ptr, pk_moduli_id, u.ptr, e0.ptr,
                                        e1.ptr, out.ptr))

    def decrypt(self, cts, sk, sk_moduli_id, out, moduli_id):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_decrypt(self.handle, count, moduli_id, cts.ptr,
                                     sk.ptr, sk_moduli_id, out.ptr))

    def encode(self, values, level, out=None):
        """Encode host slot vectors (real or complex, shape [slots] or
        [count][slots], shorter vectors zero-padded) into plaintext limbs
        [count][limbs][N] at `level`; returns the device array."""
        values = np.asarray(values)
        if values.ndim == 1:
            values = values[None]
        slots = self.n // 2
        if values.shape[-1] > slots:
            raise ValueError("more values than slots")
        is_complex = np.iscomplexobj(values)
        count = values.shape[0]
        if is_complex:
            host = np.zeros((count, slots, 2), np.float64)
            host[:, :values.shape[1], 0] = values.real
            host[:, :values.shape[1], 1] = values.imag
        else:
            host = np.zeros((count, slots), np.float64)
            host[:, :values.shape[1]] = values
        dev = self.to_device(host.view(np.int64))
        limbs = self.info["chain_count"] - (self.info["full_level"] - level)
        if out is None:
            out = self.array((count, limbs, self.n))
        try:
            self.encode_device(dev, is_complex, count, level, out)
            self.synchronize()
        finally:
            dev.free()
        return out

    def encode_device(self, values, complex_input, count, level, out):
        """fhe_encode on a device array of doubles (viewed as int64)."""
        _check(library().fhe_encode(self.handle, values.ptr,
                                    1 if complex_input else 0, count, level,
                                    out.ptr))

    def decode(self, pt, moduli_id, ckks_scale, complex_output=False):
        """Decode device plaintexts [count][limbs][N] at moduli_id with their
        (long double) ckks_scale; returns host values [count][slots]."""
        count = int(np.prod(pt.shape[:-2]))
        slots = self.n // 2
        width = 2 if complex_output else 1
        out = self.array((count, slots * width))
        try:
            self.decode_device(pt, count, moduli_id, ckks_scale,
                               complex_output, out)
            host = out.download().view(np.float64)
        finally:
            out.free()
        if complex_output:
            host = host.reshape(count, slots, 2)
            return host[..., 0] + 1j * host[..., 1]
        return host.reshape(count, slots)

    def decode_device(self, pt, count, moduli_id, ckks_scale, complex_output,
                      out):
        """fhe_decode into a device array of doubles (viewed as int64)."""
        _check(library().fhe_decode(self.handle, pt.ptr, count, moduli_id,
                                    _longdouble(ckks_scale),
                                    1 if complex_output else 0, out.ptr))

    def decoded_scale(self, moduli_id, ckks_scale):
        mid = ctypes.c_int()
        scale = np.zeros(1, np.longdouble)
        _check(library().fhe_decoded_scale(self.handle, moduli_id,
                                           _longdouble(ckks_scale),
                                           ctypes.byref(mid),
                                           scale.ctypes.data))
        return mid.value, scale[0]

    def multiply_imaginary(self, cts, moduli_id, value, out):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_multiply_imaginary(self.handle, count, moduli_id,
                                                cts.ptr, int(value), out.ptr))

    def mod_raise(self, cts, to_sparse, sparse_moduli_id, to_dense,
                  dense_moduli_id, out):
        count = int(np.prod(cts.shape[:-3]))
        _check(library().fhe_mod_raise(self.handle, count, cts.ptr,
                                       to_sparse.ptr, sparse_moduli_id,
                                       to_dense.ptr, dense_moduli_id,
                                       out.ptr))

    def set_option(self, name, value):
        """fhe_context_set_option: choose among equivalent pipelines."""
        _check(library().fhe_context_set_option(self.handle, name.encode(),
                                                int(value)))

    def profile(self, on):
        """Start (clearing totals) or stop stage timing."""
        _check(library().fhe_profile_enable(self.handle, 1 if on else 0))

    def profile_entries(self):
        """{stage name: (total ms, launches)} since profile(True)."""
        lib = library()
        count = ctypes.c_int()
        _check(lib.fhe_profile_count(self.handle, ctypes.byref(count)))
        out = {}
        for i in range(count.value):
            name = ctypes.create_string_buffer(64)
            ms, launches = ctypes.c_double(), ctypes.c_int64()
            _check(lib.fhe_profile_entry(self.handle, i, name, 64,
                                         ctypes.byref(ms),
                                         ctypes.byref(launches)))
            out[name.value.decode()] = (ms.value, launches.value)
        return out


def _count(arr):
    """Ciphertexts in a device array [count][2][limbs][N] or [2][limbs][N]."""
    return int(np.prod(arr.shape[:-3])) if len(arr.shape) > 3 else 1


def _polys_limbs(arr, n):
    assert arr.shape[-1] == n
    limbs = arr.shape[-2]
    polys = int(np.prod(arr.shape[:-2])) if len(arr.shape) > 2 else 1
    return polys, limbs
