"""GPU parity on the reference's signed representatives.  Rescaling the top
of the one-more-level chain divides by its 60-bit prime, so the 40-bit limbs
of the result lie in (-q, q) and are often negative; these are the inputs of
the first multiplication of freshly encrypted ciphertexts.  Bit-exact
against the reference engine (oracle/_ref) and the oracle:

  * forward / inverse NTT of values in (-q, q) (both transform kernels);
  * multiply + relinearize + rescale of rescaled top-level ciphertexts,
    under every pipeline option at N = 2^16;
  * decryption and key switching of rescaled top-level ciphertexts.
"""
import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]

OPTIONS = [{}, {"split_variants": 0}, {"own_limbs": 0}, {"overlap_trio": 0},
           {"overlap_trio": 2},
           {"staged_modup": 0},
           {"ks_lazy_accumulate": 0}, {"lazy_key_switch": 0},
           {"fuse_tensor": 1}, {"staged_modup": 0, "fuse_modup": 1}]
DEFAULTS = {"split_variants": 1, "own_limbs": 1, "overlap_trio": 1,
            "staged_modup": 1,
            "ks_lazy_accumulate": 1, "lazy_key_switch": 1, "fuse_tensor": 0,
            "fuse_modup": 0}


class Setup:
    def __init__(self, log_n, alpha):
        self.ref = ref = Reference(log_n, alpha, threads=8)
        lib, e = ref.lib, ref.e
        self.sk = lib.ref_secret_key(e)
        self.pk = lib.ref_public_key(e, self.sk)
        self.rlk = lib.ref_relin_key(e, self.sk)
        s = np.zeros(4, np.int64)
        lib.ref_relin_key_shape(self.rlk, ptr(s))
        self.key_mid = int(s[0])
        self.key = np.zeros((int(s[1]), int(s[2]), int(s[3])), np.int64)
        lib.ref_relin_key_get(self.rlk, ptr(self.key))
        self.ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
        self.top = ref.chain_count - 2
        self.rng = np.random.default_rng(log_n)
        self.handles = []

    def rescaled_top(self):
        """A rescaled encryption of the top level (moduli_id 1)."""
        lib, e = self.ref.lib, self.ref.e
        x = self.rng.uniform(-1, 1, self.ref.slot_count)
        ct = lib.ref_encrypt(e, ptr(x), self.ref.slot_count, self.pk,
                             self.top)
        r = lib.ref_rescale(e, ct)
        self.handles += [ct, r]
        return r

    def close(self):
        lib = self.ref.lib
        for h in self.handles:
            lib.ref_free_ct(h)
        lib.ref_free_relin_key(self.rlk)
        lib.ref_free_public_key(self.pk)
        lib.ref_free_secret_key(self.sk)
        self.ctx.close()
        self.ref.close()


@pytest.fixture(scope="module", params=[(14, 2), (16, 4)],
                ids=["n14", "n16"])
def setup(request, gpu_lib):
    s = Setup(*request.param)
    yield s
    s.close()


def test_inputs_are_signed(setup):
    a = setup.ref.get(setup.rescaled_top())
    assert setup.ref.shape(setup.handles[-1])[0] == 1
    assert (a < 0).any() and (a > -setup.ref.primes[1:a.shape[1] + 1]
                              .astype(np.int64)[:, None]).all()


@pytest.mark.parametrize("enter", [True, False])
def test_ntt_of_signed_values_matches_oracle(setup, enter):
    ref, ctx = setup.ref, setup.ctx
    orc = Oracle(int(np.log2(ref.n)), ref.primes, ctx.alpha)
    mid, limbs = 1, ref.chain_count - 1
    q = ref.primes[mid:mid + limbs].astype(np.int64)[:, None]
    rng = np.random.default_rng(5)
    # values in (-q, q) on the 40-bit limbs, as rescaling by the top prime
    # leaves them; the 60-bit limbs of such data stay non-negative
    small = q < (1 << 42)
    x = (rng.integers(-(1 << 62), 1 << 62, (2, limbs, ref.n)) % q) - \
        rng.integers(0, 2, (2, limbs, ref.n)) * q * small
    assert (x < 0).any()
    d = ctx.to_device(x)
    ctx.ntt(d, mid, enter=enter)
    y = d.download()
    assert np.array_equal(y, orc.ntt(x, mid, enter=enter))
    d.free()
    # the inverse of the transform's output and of signed values
    for z in (y, x):
        d = ctx.to_device(z)
        ctx.intt(d, mid, post=1)
        assert np.array_equal(d.download(), orc.intt(z, mid, post=1))
        d.free()


def test_hmult_of_rescaled_top_level_matches_reference(setup):
    ref, ctx = setup.ref, setup.ctx
    lib, e = ref.lib, ref.e
    pairs = [(setup.rescaled_top(), setup.rescaled_top()) for _ in range(2)]
    wants = []
    for a, b in pairs:
        w = lib.ref_hmult_relin_rescale(e, a, b, setup.rlk)
        assert w, lib.ref_last_error()
        wants.append(ref.get(w))
        lib.ref_free_ct(w)
    da = ctx.to_device(np.stack([ref.get(a) for a, _ in pairs]))
    db = ctx.to_device(np.stack([ref.get(b) for _, b in pairs]))
    dk = ctx.to_device(setup.key)
    out = ctx.array((2,) + wants[0].shape)
    for opts in OPTIONS:
        print(f"hmult signed n={ref.n} options {opts}", flush=True)
        for name, value in opts.items():
            ctx.set_option(name, value)
        ctx.hmult_relin_rescale(da, db, dk, setup.key_mid, out, 1)
        got = out.download()
        for name in opts:
            ctx.set_option(name, DEFAULTS[name])
        for i in range(2):
            assert np.array_equal(got[i], wants[i]), (opts, i)
    for arr in (da, db, dk, out):
        arr.free()


def test_multiply_relinearize_of_top_level_matches_reference(setup):
    """The engine-level multiply of two fresh top-level encryptions: both
    are rescaled inside, from the 60-bit top prime."""
    ref, ctx = setup.ref, setup.ctx
    lib, e = ref.lib, ref.e
    x = setup.rng.uniform(-1, 1, ref.slot_count)
    a = lib.ref_encrypt(e, ptr(x), ref.slot_count, setup.pk, setup.top)
    b = lib.ref_encrypt(e, ptr(x), ref.slot_count, setup.pk, setup.top)
    setup.handles += [a, b]
    w = lib.ref_multiply_relinearize(e, a, b, setup.rlk)
    want = ref.get(w)
    lib.ref_free_ct(w)
    da, db = ctx.to_device(ref.get(a)[None]), ctx.to_device(ref.get(b)[None])
    dk = ctx.to_device(setup.key)
    out = ctx.array((1,) + want.shape)
    ctx.multiply(da, setup.top, 0, db, setup.top, 0, out, key=dk,
                 key_moduli_id=setup.key_mid)
    assert np.array_equal(out.download()[0], want)
    for arr in (da, db, dk, out):
        arr.free()


def test_decrypt_of_rescaled_top_level_matches_reference(setup):
    """CKKSEngine::decrypt_to_unit_plaintext levels a rescaled ciphertext
    down by one level first (src/ckks/engine_decrypt.cpp:110-113): here
    fhe_level_down after rescale, then fhe_decrypt; fhe_decrypt alone is
    CKKSCore::decrypt_duo, checked against the oracle on the signed input."""
    ref, ctx = setup.ref, setup.ctx
    orc = Oracle(int(np.log2(ref.n)), ref.primes, ctx.alpha)
    sk_mid, skd = ref.key_data(setup.sk, 3)
    cts = [setup.rescaled_top() for _ in range(2)]
    wants = [ref.decrypt_limbs(c, setup.sk) for c in cts]
    data = np.stack([ref.get(c) for c in cts])
    dct = ctx.to_device(data)
    dsk = ctx.to_device(skd[0])
    # core decrypt of the signed limbs at moduli_id 1
    pt = ctx.array((2, ref.chain_count - 1, ref.n))
    ctx.decrypt(dct, dsk, sk_mid, pt, 1)
    got = pt.download()
    for i in range(2):
        assert np.array_equal(got[i], orc.decrypt(data[i], skd[0], sk_mid,
                                                  1)), i
    pt.free()
    # the engine's decrypt: level down after rescale, then decrypt
    mid = wants[0][0]
    low = ctx.array((2, 2, ref.chain_count - mid, ref.n))
    ctx.level_down(dct, low, setup.top, setup.top - 1, rescaled=True)
    pt = ctx.array((2, ref.chain_count - mid, ref.n))
    ctx.decrypt(low, dsk, sk_mid, pt, mid)
    got = pt.download()
    for i in range(2):
        assert np.array_equal(got[i], wants[i][1]), i
    for arr in (dct, dsk, pt, low):
        arr.free()


def test_switch_key_of_rescaled_top_level_matches_oracle(setup):
    ref, ctx = setup.ref, setup.ctx
    orc = Oracle(int(np.log2(ref.n)), ref.primes, ctx.alpha)
    ct = ref.get(setup.rescaled_top())
    want = orc.switch_key(ct, setup.key, 1, setup.key_mid)
    d = ctx.to_device(ct[None])
    dk = ctx.to_device(setup.key)
    out = ctx.array((1,) + want.shape)
    ctx.switch_key(d, dk, setup.key_mid, out, 1)
    assert np.array_equal(out.download()[0], want)
    for arr in (d, dk, out):
        arr.free()
