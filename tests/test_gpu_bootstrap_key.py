"""GPU parity of bootstrapping with a BootstrapKey on the N = 2^16 bootstrap
preset (3 stages, size "medium": 8 x 8 diagonals per stage, 38 fixed
rotation keys), bit-exact against the reference engine (oracle/_ref):

  * coeff_to_slot(ct, conj, bk) of a raised ciphertext and slot_to_coeff of
    the two parts: the key's stages (one decomposition shared by the baby
    steps, rotations in the extended basis, one ModDown per stage);
  * the whole bootstrap(ct, rlk, conj, bk) of two ciphertexts in one batched
    call, each equal to the reference's bootstrap of it alone, decrypting to
    its message (precision printed).
The key is the reference's own create_bootstrap_key output uploaded as is
(about 13 GB); building it takes most of a minute on the host cores.
"""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def build_boot(size):
    """The reference engine, its keys and BootstrapKey of `size`, and the
    device copies with the device key (a generator for a fixture)."""
    ref = Reference(16, 4, bootstrap=1, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    conj = lib.ref_conjugation_key(e, sk)
    print(f"building the reference's BootstrapKey ({size.decode()})",
          flush=True)
    bk = lib.ref_bootstrap_key(e, sk, 3, size)
    assert bk, lib.ref_last_error()
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
    relin_mid = int(s[0])
    dconj, conj_mid = dev(conj, 2)
    info = np.zeros(4, np.int32)
    lib.ref_bootstrap_key_info(bk, info.ctypes.data_as(ctypes.c_void_p))
    stage_count, _, rotation_count, _ = (int(v) for v in info)
    rotations = []
    delta = ctypes.c_int()
    for i in range(rotation_count):
        frk = lib.ref_bootstrap_key_rotation(bk, i, ctypes.byref(delta))
        arr, mid = dev(frk, 1)
        rotations.append((delta.value, arr, mid))
    sbk = lib.ref_bootstrap_key_small(bk)
    dsparse, sparse_mid = dev(sbk, 5)
    ddense, dense_mid = dev(sbk, 6)
    stages = {}
    ref_diagonals = {}
    for to_slot in (1, 0):
        mats = []
        for st in range(stage_count):
            meta = np.zeros(3, np.int32)
            inner = np.zeros(256, np.int32)
            outer = np.zeros(256, np.int32)
            assert lib.ref_bootstrap_key_stage(
                bk, to_slot, st, meta.ctypes.data_as(ctypes.c_void_p),
                inner.ctypes.data_as(ctypes.c_void_p),
                outer.ctypes.data_as(ctypes.c_void_p)) == 0
            level, ni, no = (int(v) for v in meta)
            diagonals = {}
            for o in range(no):
                for i in range(ni):
                    p = lib.ref_bootstrap_key_diagonal(bk, to_slot, st, o, i)
                    if p:
                        arr = ctx.to_device(ref.pt_get(p)[0])
                        arrays.append(arr)
                        diagonals[(o, i)] = arr
            ref_diagonals[(to_slot, st)] = diagonals
            mats.append(fhe_amd.PlainMatrix.make(level, inner[:ni],
                                                 outer[:no], diagonals))
        stages[to_slot] = mats
    parts = fhe_amd.BootstrapKeyParts.make(stages[1], stages[0], rotations,
                                           dsparse, sparse_mid, ddense,
                                           dense_mid)
    dbk = ctx.bootstrap_key(parts)
    keys = dict(relin=(drelin, relin_mid), conj=(dconj, conj_mid), bk=dbk)
    yield ref, ctx, keys, dict(sk=sk, pk=pk, rlk=rlk, conj=conj, bk=bk,
                               sbk=sbk, diagonals=ref_diagonals)
    ctx.bootstrap_key_destroy(dbk)
    for a in arrays:
        a.free()
    ctx.close()
    lib.ref_free_bootstrap_key(bk)
    lib.ref_free_conjugation_key(conj)
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()


@pytest.fixture(scope="module")
def boot(gpu_lib):
    yield from build_boot(b"medium")


def form(ref, h):
    sh = ref.shape(h)
    return sh[5], sh[4] & 6


def packed(arr, shape):
    return arr.download().reshape(-1)[:int(np.prod(shape))].reshape(shape)


def differing(got, want):
    """(polynomial, limb) pairs that differ, for the failure message."""
    bad = np.argwhere((got != want).any(axis=-1))
    return [tuple(int(v) for v in b) for b in bad[:8]], len(bad)


@pytest.mark.timeout(900)
def test_coeff_to_slot_bk_matches_reference(boot):
    ref, ctx, keys, k = boot
    lib, e = ref.lib, ref.e
    x = np.random.default_rng(15).uniform(-1, 1, ref.slot_count)
    ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 0)
    raised = lib.ref_mod_raise(e, ct, k["sbk"])
    real, imag = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.ref_coeff_to_slot_bk(e, raised, k["conj"], k["bk"],
                                    ctypes.byref(real),
                                    ctypes.byref(imag)) == 0, \
        lib.ref_last_error()
    want_r, want_i = ref.get(real.value), ref.get(imag.value)
    level, flags = form(ref, raised)
    d = ctx.to_device(ref.get(raised))
    out_r, out_i = ctx.array(d.shape), ctx.array(d.shape)
    conj, conj_mid = keys["conj"]
    print(f"coeff_to_slot_bk level={level} flags={flags}", flush=True)
    lv, fl = ctx.coeff_to_slot_bk(d, level, flags, conj, conj_mid,
                                  keys["bk"], out_r, out_i)
    assert (lv, fl) == form(ref, real.value)
    got_r = packed(out_r, want_r.shape)
    got_i = packed(out_i, want_i.shape)
    assert np.array_equal(got_r, want_r), differing(got_r, want_r)
    assert np.array_equal(got_i, want_i), differing(got_i, want_i)
    # the parts at another level than the slot-to-coeff stages' are refused
    # (the reference reads past its diagonals then)
    dr = ctx.to_device(want_r)
    di = ctx.to_device(want_i)
    out_c = ctx.array(d.shape)
    with pytest.raises(fhe_amd.FheError, match="level differs"):
        ctx.slot_to_coeff_bk(dr, di, lv, fl, conj, conj_mid, keys["bk"],
                             out_c)
    # slot_to_coeff of the two parts leveled down to the stages' level (where
    # EvalMod leaves them in a bootstrap), against the reference's
    meta = np.zeros(3, np.int32)
    assert lib.ref_bootstrap_key_stage(
        k["bk"], 0, 0, meta.ctypes.data_as(ctypes.c_void_p), None, None) == 0
    r2 = lib.ref_level_down(e, real.value, int(meta[0]))
    i2 = lib.ref_level_down(e, imag.value, int(meta[0]))
    assert r2 and i2, lib.ref_last_error()
    stc = lib.ref_slot_to_coeff_bk(e, r2, i2, k["conj"], k["bk"])
    assert stc, lib.ref_last_error()
    want_c = ref.get(stc)
    lv2, fl2 = form(ref, r2)
    dr2 = ctx.to_device(ref.get(r2))
    di2 = ctx.to_device(ref.get(i2))
    print(f"slot_to_coeff_bk level={lv2} flags={fl2}", flush=True)
    assert ctx.slot_to_coeff_bk(dr2, di2, lv2, fl2, conj, conj_mid,
                                keys["bk"], out_c) == form(ref, stc)
    got_c = packed(out_c, want_c.shape)
    assert np.array_equal(got_c, want_c), differing(got_c, want_c)
    for h in (ct, raised, real.value, imag.value, r2, i2, stc):
        lib.ref_free_ct(h)
    for a in (d, out_r, out_i, dr, di, dr2, di2, out_c):
        a.free()


@pytest.mark.timeout(900)
def test_bootstrap_bk_matches_reference(boot):
    """Two ciphertexts bootstrapped in one batched call, each equal to the
    reference's bootstrap of it alone."""
    ref, ctx, keys, k = boot
    lib, e = ref.lib, ref.e
    rng = np.random.default_rng(16)
    xs = [rng.uniform(-1, 1, ref.slot_count) for _ in range(2)]
    cts = [lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 2)
           for x in xs]
    wants, handles = [], []
    for ct in cts:
        h = lib.ref_bootstrap_bk(e, ct, k["rlk"], k["conj"], k["bk"])
        assert h, lib.ref_last_error()
        handles.append(h)
        wants.append(ref.get(h))
    level, flags = form(ref, cts[0])
    d = ctx.to_device(np.stack([ref.get(ct) for ct in cts]))
    out = ctx.array((2, 2, ref.chain_count - 1, ctx.n))
    relin, relin_mid = keys["relin"]
    conj, conj_mid = keys["conj"]
    print("bootstrap_bk of 2", flush=True)
    lv, fl = ctx.bootstrap_bk(d, level, flags, relin, relin_mid, conj,
                              conj_mid, keys["bk"], out)
    assert (lv, fl) == form(ref, handles[0])
    got = packed(out, (2,) + wants[0].shape)
    for i in range(2):
        assert np.array_equal(got[i], wants[i]), (i,
                                                  differing(got[i], wants[i]))
        ref.set(handles[i], got[i])
        dec = np.zeros(ref.slot_count)
        lib.ref_decrypt(e, handles[i], k["sk"], ptr(dec))
        err = np.abs(dec - xs[i]).max()
        print(f"bootstrap_bk {i}: decrypted max abs error {err:.3e}, "
              f"precision {-np.log2(err):.2f} bits", flush=True)
        assert err < 1e-2
    for h in handles + cts:
        lib.ref_free_ct(h)
    d.free()
    out.free()


@pytest.mark.timeout(600)
def test_bootstrap_bk_graph_replays_equal_eager(boot):
    """fhe_bootstrap_bk recorded into a HIP graph after an eager call and
    replayed three times, with eager work between the replays: every replay
    gives the eager call's limbs."""
    ref, ctx, keys, k = boot
    lib, e = ref.lib, ref.e
    x = np.random.default_rng(17).uniform(-1, 1, ref.slot_count)
    ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 2)
    level, flags = form(ref, ct)
    d = ctx.to_device(ref.get(ct)[None])
    out = ctx.array((1, 2, ref.chain_count - 1, ctx.n))
    other = ctx.array(out.shape)
    relin, relin_mid = keys["relin"]
    conj, conj_mid = keys["conj"]

    def run(dst):
        return ctx.bootstrap_bk(d, level, flags, relin, relin_mid, conj,
                                conj_mid, keys["bk"], dst)

    run(out)
    ctx.synchronize()
    want = out.download()
    ctx.graph_begin()
    try:
        run(out)
    finally:
        graph = ctx.graph_end()
    try:
        for rep in range(3):
            out.upload(np.zeros(out.shape, np.int64))
            run(other)
            ctx.synchronize()
            ctx.graph_launch(graph)
            ctx.synchronize()
            assert np.array_equal(out.download(), want), rep
        assert np.array_equal(other.download(), want)
    finally:
        ctx.graph_destroy(graph)
    lib.ref_free_ct(ct)
    for a in (d, out, other):
        a.free()


@pytest.mark.timeout(600)
def test_device_encoded_stages_equal_reference_key(boot):
    """fhe_bootstrap_key_encode_stage (make_bootstrap_stage on the device:
    encode with the special primes, rescale, NTT) gives every stage's
    diagonals equal to the reference key's."""
    ref, ctx, keys, k = boot
    for (to_slot, st), diagonals in sorted(k["diagonals"].items()):
        print(f"stage to_slot={to_slot} {st}", flush=True)
        got = ctx.bootstrap_key_stage(3, 0, to_slot, st)
        host = got.download()
        order = sorted(diagonals)
        assert host.shape[0] == len(order)
        for idx, key in enumerate(order):
            want = diagonals[key].download()
            assert np.array_equal(host[idx], want), (to_slot, st, key)
        got.free()


def test_layout_matches_reference_key(boot):
    """fhe_bootstrap_key_stage_layout / _rotation_layout against the
    reference's BootstrapKey: every stage's level and deltas, every
    rotation context's delta and key level."""
    ref, ctx, keys, k = boot
    lib = ref.lib
    layout = fhe_amd.bootstrap_key_layout(16, 4, 3, primes=ref.primes)
    for name, to_slot in (("coeff_to_slot", 1), ("slot_to_coeff", 0)):
        for st, mine in enumerate(layout[name]):
            meta = np.zeros(3, np.int32)
            inner = np.zeros(256, np.int32)
            outer = np.zeros(256, np.int32)
            assert lib.ref_bootstrap_key_stage(
                k["bk"], to_slot, st, meta.ctypes.data_as(ctypes.c_void_p),
                inner.ctypes.data_as(ctypes.c_void_p),
                outer.ctypes.data_as(ctypes.c_void_p)) == 0
            assert mine["level"] == meta[0], (name, st)
            assert mine["inner_deltas"] == inner[:meta[1]].tolist()
            assert mine["outer_deltas"] == outer[:meta[2]].tolist()
    info = np.zeros(4, np.int32)
    lib.ref_bootstrap_key_info(k["bk"], info.ctypes.data_as(ctypes.c_void_p))
    want = {}
    delta = ctypes.c_int()
    for i in range(int(info[2])):
        frk = lib.ref_bootstrap_key_rotation(k["bk"], i, ctypes.byref(delta))
        s = np.zeros(4, np.int64)
        lib.ref_key_shape(frk, 1, ptr(s))
        want[delta.value] = int(s[0])
    assert dict(layout["rotations"]) == want
    assert len(layout["rotations"]) == len(want)


def test_bootstrap_key_refuses_bad_parts(boot):
    ref, ctx, keys, k = boot
    # a stage whose inner delta has no key
    diag = {(o, i): 1 for o in range(2) for i in range(2) if (o, i) != (0, 0)}
    m = fhe_amd.PlainMatrix.make(25, [0, 12345], [0, 0], diag)
    parts = fhe_amd.BootstrapKeyParts.make([m], [], [])
    with pytest.raises(fhe_amd.FheError, match="no rotation key"):
        ctx.bootstrap_key(parts)
    # complete parts (the diagonal's address is not read at creation), then
    # a missing diagonal
    m = fhe_amd.PlainMatrix.make(25, [0], [0, 0], {(1, 0): 1})
    parts = fhe_amd.BootstrapKeyParts.make([m], [], [])
    ctx.bootstrap_key_destroy(ctx.bootstrap_key(parts))
    m = fhe_amd.PlainMatrix.make(25, [0], [0, 0], {})
    parts = fhe_amd.BootstrapKeyParts.make([m], [], [])
    with pytest.raises(fhe_amd.FheError, match="missing"):
        ctx.bootstrap_key(parts)
