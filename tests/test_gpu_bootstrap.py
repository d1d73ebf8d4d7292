"""GPU parity of bootstrapping with a RotationKey on the N = 2^16 bootstrap
preset, bit-exact against the reference engine (oracle/_ref):

  * coeff_to_slot of a raised ciphertext (3 stages: baby-step rotation
    chains, products with the merged diagonals, giant-step rotations,
    conjugation into real and imaginary parts), and slot_to_coeff of the
    two parts;
  * the whole bootstrap (level down, mod_raise, coeff_to_slot, EvalMod of
    both parts, slot_to_coeff, subtraction, normalization) of two ciphertexts
    in one batched call, each equal to the reference's and decrypting to its
    message.
The reference's own run takes tens of seconds on the host cores.
"""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

STAGES = 3


@pytest.fixture(scope="module")
def boot(gpu_lib):
    ref = Reference(16, 4, bootstrap=1, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    conj = lib.ref_conjugation_key(e, sk)
    rk = lib.ref_rotation_key(e, sk)
    sbk = lib.ref_small_bootstrap_key(e, sk)
    assert rk and sbk, lib.ref_last_error()
    ctx = fhe_amd.Context(16, 4, scale_bits=54, small_prime_bits=54,
                          primes=ref.primes)
    arrays = []

    def dev(key, kind):
        mid, data = ref.key_data(key, kind)
        arr = ctx.to_device(data)
        arrays.append(arr)
        return arr, mid

    s = np.zeros(4, np.int64)
    lib.ref_relin_key_shape(rlk, ptr(s))
    relin = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
    lib.ref_relin_key_get(rlk, ptr(relin))
    drelin = ctx.to_device(relin)
    arrays.append(drelin)
    dconj, conj_mid = dev(conj, 2)
    rotations = []
    for i in range(lib.ref_rotation_key_count(rk)):
        frk = lib.ref_rotation_key_fixed(rk, i)
        arr, mid = dev(frk, 1)
        rotations.append((arr, mid, lib.ref_fixed_rotation_key_delta(frk)))
    dsparse, sparse_mid = dev(sbk, 5)
    ddense, dense_mid = dev(sbk, 6)
    keys = fhe_amd.BootstrapKeys.make(drelin, int(s[0]), dconj, conj_mid,
                                      rotations, dsparse, sparse_mid, ddense,
                                      dense_mid)
    yield ref, ctx, keys, dict(sk=sk, pk=pk, rlk=rlk, conj=conj, rk=rk,
                               sbk=sbk)
    for a in arrays:
        a.free()
    ctx.close()
    lib.ref_free_small_bootstrap_key(sbk)
    lib.ref_free_rotation_key(rk)
    lib.ref_free_conjugation_key(conj)
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()


def form(ref, h):
    sh = ref.shape(h)
    return sh[5], sh[4] & 6


def packed(arr, shape):
    return arr.download().reshape(-1)[:int(np.prod(shape))].reshape(shape)


@pytest.mark.timeout(600)
def test_coeff_to_slot_matches_reference(boot):
    ref, ctx, keys, k = boot
    lib, e = ref.lib, ref.e
    x = np.random.default_rng(5).uniform(-1, 1, ref.slot_count)
    ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 0)
    raised = lib.ref_mod_raise(e, ct, k["sbk"])
    real, imag = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.ref_coeff_to_slot(e, raised, k["conj"], k["rk"], STAGES,
                                 ctypes.byref(real), ctypes.byref(imag)) == 0
    want_r, want_i = ref.get(real.value), ref.get(imag.value)
    level, flags = form(ref, raised)
    d = ctx.to_device(ref.get(raised))
    out_r, out_i = ctx.array(d.shape), ctx.array(d.shape)
    lv, fl = ctx.coeff_to_slot(d, level, flags, keys, STAGES, out_r, out_i)
    assert (lv, fl) == form(ref, real.value)
    assert np.array_equal(packed(out_r, want_r.shape), want_r)
    assert np.array_equal(packed(out_i, want_i.shape), want_i)
    # slot_to_coeff of the two parts, against the reference's
    stc = lib.ref_slot_to_coeff(e, real.value, imag.value, k["conj"],
                                k["rk"], STAGES)
    assert stc, lib.ref_last_error()
    want_c = ref.get(stc)
    dr = ctx.to_device(want_r)
    di = ctx.to_device(want_i)
    out_c = ctx.array(d.shape)
    assert ctx.slot_to_coeff(dr, di, lv, fl, keys, STAGES, out_c) == \
        form(ref, stc)
    assert np.array_equal(packed(out_c, want_c.shape), want_c)
    for h in (ct, raised, real.value, imag.value, stc):
        lib.ref_free_ct(h)
    for a in (d, out_r, out_i, dr, di, out_c):
        a.free()


@pytest.mark.timeout(900)
def test_bootstrap_matches_reference(boot):
    """Two ciphertexts bootstrapped in one batched call, each equal to the
    reference's bootstrap of it alone."""
    ref, ctx, keys, k = boot
    lib, e = ref.lib, ref.e
    rng = np.random.default_rng(6)
    xs = [rng.uniform(-1, 1, ref.slot_count) for _ in range(2)]
    cts = [lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 2)
           for x in xs]
    wants, handles = [], []
    for ct in cts:
        h = lib.ref_bootstrap(e, ct, k["rlk"], k["conj"], k["rk"], k["sbk"],
                              STAGES)
        assert h, lib.ref_last_error()
        handles.append(h)
        wants.append(ref.get(h))
    level, flags = form(ref, cts[0])
    d = ctx.to_device(np.stack([ref.get(ct) for ct in cts]))
    out = ctx.array((2, 2, ref.chain_count - 1, ctx.n))
    lv, fl = ctx.bootstrap(d, level, flags, keys, STAGES, out)
    assert (lv, fl) == form(ref, handles[0])
    got = packed(out, (2,) + wants[0].shape)
    for i in range(2):
        assert np.array_equal(got[i], wants[i]), i
        # the device result (equal to the reference's) decrypted: its
        # precision in bits, -log2 of the largest slot error
        ref.set(handles[i], got[i])
        dec = np.zeros(ref.slot_count)
        lib.ref_decrypt(e, handles[i], k["sk"], ptr(dec))
        err = np.abs(dec - xs[i]).max()
        print(f"bootstrap {i}: decrypted max abs error {err:.3e}, "
              f"precision {-np.log2(err):.2f} bits", flush=True)
        assert err < 1e-2
    for h in handles + cts:
        lib.ref_free_ct(h)
    d.free()
    out.free()
