"""GPU parity of the engine-level multiplications on inputs in every form,
bit-exact against the reference engine (oracle/_ref):

  * multiply(ct, ct) and multiply(ct, ct, rlk) with both inputs through
    prepare_for_multiply: INTT / NTT form, rescaled or not, equal and
    different levels (level down plain and after rescale), batched;
  * square(ct) with and without the key (a == b);
  * a relinearized product fed back in at its own level (not rescaled);
  * multiply(ct, pt) for duos and trios with plaintexts in every form;
  * refusals: level 0, unknown flags.
"""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

FORM_MASK = fhe_amd.FLAG_NTT_FORM | fhe_amd.FLAG_RESCALED


class Texts:
    """Reference ciphertexts in the requested forms, freed on close."""

    def __init__(self, ref, pk):
        self.ref, self.pk, self.handles = ref, pk, []
        self.rng = np.random.default_rng(11)

    def keep(self, h, plain=False):
        assert h, self.ref.lib.ref_last_error()
        self.handles.append((h, plain))
        return h

    def ct(self, level, rescaled=False, ntt=False):
        lib, e = self.ref.lib, self.ref.e
        x = self.rng.uniform(-1, 1, self.ref.slot_count)
        h = self.keep(lib.ref_encrypt(e, ptr(x), self.ref.slot_count,
                                      self.pk, level))
        if rescaled:
            h = self.keep(lib.ref_rescale_text(e, h, 0))
        if ntt:
            h = self.keep(lib.ref_ntt_text(e, h, 0))
        return h

    def pt(self, level, rescaled=False, ntt=False):
        lib, e = self.ref.lib, self.ref.e
        x = self.rng.uniform(-1, 1, self.ref.slot_count)
        h = self.keep(lib.ref_encode_pt(e, ptr(x), level), True)
        if rescaled:
            h = self.keep(lib.ref_rescale_text(e, h, 1), True)
        if ntt:
            h = self.keep(lib.ref_ntt_text(e, h, 1), True)
        return h

    def form(self, h, plain=False):
        s = self.ref.pt_shape(h) if plain else self.ref.shape(h)
        return s[5], s[4] & FORM_MASK

    def close(self):
        for h, plain in self.handles:
            (self.ref.lib.ref_free_pt if plain else
             self.ref.lib.ref_free_ct)(h)


@pytest.fixture(scope="module")
def setup(gpu_lib):
    ref = Reference(14, 2, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    s = np.zeros(4, np.int64)
    lib.ref_relin_key_shape(rlk, ptr(s))
    key = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
    lib.ref_relin_key_get(rlk, ptr(key))
    ctx = fhe_amd.Context(14, 2, primes=ref.primes)
    texts = Texts(ref, pk)
    dkey = ctx.to_device(key)
    yield ref, ctx, texts, rlk, dkey, int(s[0])
    dkey.free()
    texts.close()
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ctx.close()
    ref.close()


def _run_pairs(setup, pairs, relin):
    """pairs: [(a handle, b handle)], all a of one form, all b of one form;
    one batched device call against the reference per pair."""
    ref, ctx, texts, rlk, dkey, key_mid = setup
    lib, e = ref.lib, ref.e
    la, fa = texts.form(pairs[0][0])
    lb, fb = texts.form(pairs[0][1])
    wants = []
    for a, b in pairs:
        assert texts.form(a) == (la, fa) and texts.form(b) == (lb, fb)
        if a == b:
            w = lib.ref_square(e, a, rlk if relin else None)
        elif relin:
            w = lib.ref_multiply_relinearize(e, a, b, rlk)
        else:
            w = lib.ref_multiply_ct(e, a, b)
        assert w, lib.ref_last_error()
        wants.append(ref.get(w))
        lib.ref_free_ct(w)
    da = ctx.to_device(np.stack([ref.get(a) for a, _ in pairs]))
    db = ctx.to_device(np.stack([ref.get(b) for _, b in pairs]))
    out = ctx.array((len(pairs),) + wants[0].shape)
    ctx.multiply(da, la, fa, db, lb, fb, out,
                 key=dkey if relin else None, key_moduli_id=key_mid)
    got = out.download()
    for h in (da, db, out):
        h.free()
    for i in range(len(pairs)):
        assert np.array_equal(got[i], wants[i]), (la, fa, lb, fb, relin, i)


FORMS = [  # (level offset from the top, rescaled, ntt)
    ((0, False, False), (0, False, False)),
    ((0, True, False), (0, True, False)),
    ((0, True, True), (0, False, True)),
    ((0, False, True), (2, True, False)),
    ((1, True, False), (3, False, True)),
    ((2, False, False), (0, True, True)),
    ((4, True, True), (1, False, False)),
]


@pytest.mark.parametrize("relin", [False, True])
@pytest.mark.parametrize("fa,fb", FORMS)
def test_multiply_forms_match_reference(setup, fa, fb, relin):
    ref, ctx, texts = setup[:3]
    top = ref.chain_count - 2
    print(f"multiply a={fa} b={fb} relin={relin}", flush=True)
    pairs = [(texts.ct(top - fa[0], fa[1], fa[2]),
              texts.ct(top - fb[0], fb[1], fb[2])) for _ in range(2)]
    _run_pairs(setup, pairs, relin)


@pytest.mark.parametrize("relin", [False, True])
@pytest.mark.parametrize("form", [(0, False, False), (1, True, False),
                                  (2, False, True), (3, True, True)])
def test_square_matches_reference(setup, form, relin):
    ref, ctx, texts = setup[:3]
    top = ref.chain_count - 2
    print(f"square {form} relin={relin}", flush=True)
    a = texts.ct(top - form[0], form[1], form[2])
    _run_pairs(setup, [(a, a)], relin)


def test_product_fed_back(setup):
    """A relinearized product (INTT, not rescaled, level L - 1) times a
    fresh rescaled ciphertext at a higher level, then squared."""
    ref, ctx, texts, rlk, dkey, key_mid = setup
    lib, e = ref.lib, ref.e
    top = ref.chain_count - 2
    p = texts.keep(lib.ref_multiply_relinearize(
        e, texts.ct(top), texts.ct(top, True), rlk))
    assert texts.form(p) == (top - 1, 0)
    _run_pairs(setup, [(p, texts.ct(top, True))], True)
    _run_pairs(setup, [(p, p)], True)


@pytest.mark.parametrize("ct_form,pt_form", [
    ((0, False, False), (0, False, False)),
    ((0, True, True), (0, True, False)),
    ((1, False, True), (0, False, True)),
    ((0, True, False), (2, True, True)),
])
@pytest.mark.parametrize("trio", [False, True])
def test_multiply_plain_matches_reference(setup, ct_form, pt_form, trio):
    ref, ctx, texts, rlk = setup[:4]
    lib, e = ref.lib, ref.e
    top = ref.chain_count - 2
    print(f"multiply_plain ct={ct_form} pt={pt_form} trio={trio}", flush=True)
    cts, pts, wants = [], [], []
    for _ in range(2):
        if trio:
            # a trio at top - ct_form[0] - 1: NTT form, not rescaled
            c = texts.keep(lib.ref_multiply_ct(
                e, texts.ct(top - ct_form[0]), texts.ct(top - ct_form[0])))
        else:
            c = texts.ct(top - ct_form[0], ct_form[1], ct_form[2])
        p = texts.pt(top - pt_form[0], pt_form[1], pt_form[2])
        w = lib.ref_multiply_pt(e, c, p)
        assert w, lib.ref_last_error()
        wants.append(ref.get(w))
        lib.ref_free_ct(w)
        cts.append(c)
        pts.append(p)
    lc, fc = texts.form(cts[0])
    lp, fp = texts.form(pts[0], plain=True)
    dct = ctx.to_device(np.stack([ref.get(c) for c in cts]))
    dpt = ctx.to_device(np.stack([ref.pt_get(p)[0] for p in pts]))
    out = ctx.array((2,) + wants[0].shape)
    ctx.multiply_plain(dct, lc, fc, dpt, lp, fp, out)
    got = out.download()
    for h in (dct, dpt, out):
        h.free()
    for i in range(2):
        assert np.array_equal(got[i], wants[i]), i


def test_multiply_refusals(setup):
    ref, ctx = setup[:2]
    n = ctx.n
    top = ref.chain_count - 2
    a = ctx.array((1, 2, ref.chain_count - top, n))
    out = ctx.array((1, 3, ref.chain_count - top + 1, n))
    with pytest.raises(fhe_amd.FheError):
        ctx.multiply(a, 0, 0, a, 0, 0, out)
    with pytest.raises(fhe_amd.FheError):
        ctx.multiply(a, 1, 8, a, 1, 0, out)
    with pytest.raises(fhe_amd.FheError):
        ctx.multiply(a, top + 1, 0, a, top, 0, out)
    a.free()
    out.free()
