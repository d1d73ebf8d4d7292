"""GPU parity of the engine-level addition and subtraction of ciphertexts in
every form, bit-exact against the reference engine (oracle/_ref):
CKKSEngine::add / subtract through match_form (domain of the first text,
level down of the higher one, rescale where only one is rescaled), duo +
duo, trio + duo, duo + trio, trio - duo, duo - trio, batched, including
rescaled top-level (signed) inputs."""
import numpy as np
import pytest

import fhe_amd
from support import Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

MASK = fhe_amd.FLAG_NTT_FORM | fhe_amd.FLAG_RESCALED


@pytest.fixture(scope="module")
def setup(gpu_lib):
    ref = Reference(14, 2, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    ctx = fhe_amd.Context(14, 2, primes=ref.primes)
    handles = []
    rng = np.random.default_rng(21)

    def make(level, rescaled=False, ntt=False, trio=False):
        def enc(lv):
            x = rng.uniform(-1, 1, ref.slot_count)
            h = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, lv)
            handles.append(h)
            return h
        if trio:
            # NTT-form trio, not rescaled, at level - 1
            h = lib.ref_multiply_ct(e, enc(level), enc(level))
            handles.append(h)
            return h
        h = enc(level)
        if rescaled:
            h = lib.ref_rescale_text(e, h, 0)
            handles.append(h)
        if ntt:
            h = lib.ref_ntt_text(e, h, 0)
            handles.append(h)
        return h

    yield ref, ctx, make
    for h in handles:
        lib.ref_free_ct(h)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ctx.close()
    ref.close()


def form(ref, h):
    s = ref.shape(h)
    return s[5], s[4] & MASK


# (level offset from the top, rescaled, ntt, trio) of a and b
CASES = [
    ((1, False, False, False), (1, False, False, False)),
    ((1, True, False, False), (1, False, False, False)),
    ((1, True, True, False), (3, False, False, False)),
    ((2, False, False, False), (1, True, True, False)),
    ((1, False, True, False), (1, True, False, False)),
    ((0, True, False, False), (0, False, False, False)),
    ((0, False, False, True), (1, True, True, False)),
    ((1, True, False, False), (0, False, False, True)),
]


@pytest.mark.parametrize("subtract", [False, True])
@pytest.mark.parametrize("fa,fb", CASES)
def test_add_forms_match_reference(setup, fa, fb, subtract):
    ref, ctx, make = setup
    lib, e = ref.lib, ref.e
    top = ref.chain_count - 2
    print(f"add a={fa} b={fb} subtract={subtract}", flush=True)
    pairs = [(make(top - fa[0], *fa[1:]), make(top - fb[0], *fb[1:]))
             for _ in range(2)]
    wants = []
    for a, b in pairs:
        w = (lib.ref_subtract_ct if subtract else lib.ref_add_ct)(e, a, b)
        assert w, lib.ref_last_error()
        wants.append(ref.get(w))
        lib.ref_free_ct(w)
    la, ga = form(ref, pairs[0][0])
    lb, gb = form(ref, pairs[0][1])
    da = ctx.to_device(np.stack([ref.get(a) for a, _ in pairs]))
    db = ctx.to_device(np.stack([ref.get(b) for _, b in pairs]))
    out = ctx.array((2,) + wants[0].shape)
    ctx.add(da, la, ga, db, lb, gb, out, subtract=subtract)
    got = out.download()
    for h in (da, db, out):
        h.free()
    for i in range(2):
        assert np.array_equal(got[i], wants[i]), i


def test_add_refusals(setup):
    ref, ctx, _ = setup
    a = ctx.array((1, 2, 2, ctx.n))
    with pytest.raises(fhe_amd.FheError):
        ctx.add(a, 0, 16, a, 0, 0, a)
    with pytest.raises(fhe_amd.FheError):
        ctx.add(a, ref.chain_count, 0, a, 0, 0, a)
    a.free()
