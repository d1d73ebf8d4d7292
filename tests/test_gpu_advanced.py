"""GPU parity of the engine's advanced compositions with the reference
engine (oracle/_ref), bit-exact:

  * evaluate_polynomial with degree-7 double coefficients (power basis by
    squares and products, weighted sum) on a batch of fresh encryptions,
    at the top level and below; the results decrypt to p(x);
  * sum_1d_by_row with the RotationKey (rotate by 1, 2, 4, ... and add),
    at the top of the one-more-level chain and below; the results decrypt
    to the sum of the slots in every slot;
  * the requirements the reference checks are refused.
"""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

COEFFS = [0.25, -1.5, 0.75, 0.5, -0.125, 0.0625, 0.3, -0.2]


@pytest.fixture(scope="module")
def engine():
    ref = Reference(15, 4, threads=16)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    rk = lib.ref_rotation_key(e, sk)
    ctx = fhe_amd.Context(15, 4, primes=ref.primes)
    key_mid, key = ref.key(rlk)
    drlk = ctx.to_device(key)
    rot, mids, deltas = [], [], []
    for i in range(lib.ref_rotation_key_count(rk)):
        frk = lib.ref_rotation_key_fixed(rk, i)
        mid, data = ref.key_data(frk, 1)
        rot.append(ctx.to_device(data))
        mids.append(mid)
        deltas.append(lib.ref_fixed_rotation_key_delta(frk))
    drk = fhe_amd.RotationKey.make(rot, mids, deltas)
    yield dict(ref=ref, lib=lib, e=e, sk=sk, pk=pk, rlk=rlk, rk=rk, ctx=ctx,
               key_mid=key_mid, drlk=drlk, drk=drk)
    for arr in [drlk] + rot:
        arr.free()
    ctx.close()
    lib.ref_free_rotation_key(rk)
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()


def encrypt(env, values, level):
    lib, e, ref = env["lib"], env["e"], env["ref"]
    h = lib.ref_encrypt(e, ptr(np.ascontiguousarray(values)), ref.slot_count,
                        env["pk"], level)
    return h


@pytest.mark.parametrize("level_from_top", [0, 3])
def test_evaluate_polynomial_matches_reference(engine, level_from_top,
                                               gpu_lib):
    ref, lib, e, ctx = engine["ref"], engine["lib"], engine["e"], engine["ctx"]
    level = ctx.info["full_level"] - level_from_top
    slots = ref.slot_count
    rng = np.random.default_rng(11 + level_from_top)
    xs = [rng.uniform(-1, 1, slots) for _ in range(3)]
    handles = [encrypt(engine, x, level) for x in xs]
    try:
        cts = np.stack([ref.get(h) for h in handles])
        din = ctx.to_device(cts)
        out = ctx.array(cts.shape)
        lv, fl = ctx.evaluate_polynomial(din, level, COEFFS, engine["drlk"],
                                         engine["key_mid"], out)
        coeffs = np.array(COEFFS)
        for i, h in enumerate(handles):
            w = lib.ref_evaluate_polynomial(e, h, ptr(coeffs), len(COEFFS),
                                            engine["rlk"])
            want = ref.get(w)
            assert ref.shape(w)[5] == lv and ref.shape(w)[4] == fl
            # results packed: result i at i times its own size
            got = out.download().reshape(-1)[i * want.size:
                                             (i + 1) * want.size].reshape(
                want.shape)
            assert np.array_equal(got, want), i
            dec = np.zeros(slots)
            lib.ref_decrypt(e, w, engine["sk"], ptr(dec))
            p = sum(c * xs[i] ** k for k, c in enumerate(COEFFS))
            assert np.abs(dec - p).max() < 1e-4
            lib.ref_free_ct(w)
        din.free()
        out.free()
    finally:
        for h in handles:
            lib.ref_free_ct(h)


@pytest.mark.parametrize("level_from_top", [0, 5])
def test_sum_1d_by_row_matches_reference(engine, level_from_top, gpu_lib):
    ref, lib, e, ctx = engine["ref"], engine["lib"], engine["e"], engine["ctx"]
    level = ctx.info["full_level"] - level_from_top
    slots = ref.slot_count
    rng = np.random.default_rng(21 + level_from_top)
    xs = [rng.uniform(-1, 1, slots) / slots for _ in range(2)]
    handles = [encrypt(engine, x, level) for x in xs]
    try:
        cts = np.stack([ref.get(h) for h in handles])
        din = ctx.to_device(cts)
        out = ctx.array(cts.shape)
        lv, fl = ctx.sum_1d_by_row(din, level, 0, engine["drk"], out)
        for i, h in enumerate(handles):
            w = lib.ref_sum_1d_by_row(e, h, engine["rk"])
            want = ref.get(w)
            assert ref.shape(w)[5] == lv and ref.shape(w)[4] == fl
            # results packed: result i at i times its own size
            got = out.download().reshape(-1)[i * want.size:
                                             (i + 1) * want.size].reshape(
                want.shape)
            assert np.array_equal(got, want), i
            dec = np.zeros(slots)
            lib.ref_decrypt(e, w, engine["sk"], ptr(dec))
            # the scheme's noise (the limbs equal the reference's above)
            assert np.abs(dec - xs[i].sum()).max() < 1e-4
            lib.ref_free_ct(w)
        din.free()
        out.free()
    finally:
        for h in handles:
            lib.ref_free_ct(h)


def test_requirements_are_refused(engine, gpu_lib):
    ctx = engine["ctx"]
    n = ctx.n
    level = 2
    limbs = ctx.limbs_at(ctx.info["full_level"] - level)
    x = ctx.array((1, 2, limbs, n))
    out = ctx.array((1, 2, limbs, n))
    try:
        # degree 7 needs level 4
        with pytest.raises(fhe_amd.FheError):
            ctx.evaluate_polynomial(x, level, COEFFS, engine["drlk"],
                                    engine["key_mid"], out)
        with pytest.raises(fhe_amd.FheError):
            ctx.evaluate_polynomial(x, level, [1.0], engine["drlk"],
                                    engine["key_mid"], out)
        with pytest.raises(fhe_amd.FheError):
            ctx.evaluate_polynomial(x, level, [1.0, 0.0, 0.0],
                                    engine["drlk"], engine["key_mid"], out)
        with pytest.raises(fhe_amd.FheError):
            ctx.sum_1d_by_row(x, level, fhe_amd.FLAG_NTT_FORM, engine["drk"],
                              out)
    finally:
        x.free()
        out.free()
