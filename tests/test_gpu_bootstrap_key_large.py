"""GPU parity of bootstrapping with a "large" BootstrapKey (inner-loop
shift 1: 16 x 4 diagonals per stage, 52 fixed rotation keys) on the N = 2^16
bootstrap preset, bit-exact against the reference engine (oracle/_ref): the
whole bootstrap(ct, rlk, conj, bk) of one ciphertext, and the device-encoded
diagonals of the first coeff-to-slot stage against the reference key's."""
import numpy as np
import pytest

from support import Reference, ptr
from test_gpu_bootstrap_key import build_boot, differing, form, packed

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


@pytest.fixture(scope="module")
def boot_large(gpu_lib):
    yield from build_boot(b"large")


@pytest.mark.timeout(900)
def test_bootstrap_large_key_matches_reference(boot_large):
    ref, ctx, keys, k = boot_large
    lib, e = ref.lib, ref.e
    x = np.random.default_rng(18).uniform(-1, 1, ref.slot_count)
    ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, k["pk"], 2)
    h = lib.ref_bootstrap_bk(e, ct, k["rlk"], k["conj"], k["bk"])
    assert h, lib.ref_last_error()
    want = ref.get(h)
    level, flags = form(ref, ct)
    d = ctx.to_device(ref.get(ct)[None])
    out = ctx.array((1, 2, ref.chain_count - 1, ctx.n))
    relin, relin_mid = keys["relin"]
    conj, conj_mid = keys["conj"]
    print("bootstrap_bk with the large key", flush=True)
    lv, fl = ctx.bootstrap_bk(d, level, flags, relin, relin_mid, conj,
                              conj_mid, keys["bk"], out)
    assert (lv, fl) == form(ref, h)
    got = packed(out, want.shape)
    assert np.array_equal(got, want), differing(got, want)
    ref.set(h, got)
    dec = np.zeros(ref.slot_count)
    lib.ref_decrypt(e, h, k["sk"], ptr(dec))
    err = np.abs(dec - x).max()
    print(f"bootstrap_bk (large key): decrypted max abs error {err:.3e}, "
          f"precision {-np.log2(err):.2f} bits", flush=True)
    assert err < 1e-2
    lib.ref_free_ct(h)
    lib.ref_free_ct(ct)
    d.free()
    out.free()


def test_device_encoded_stage_equals_large_key(boot_large):
    ref, ctx, keys, k = boot_large
    diagonals = k["diagonals"][(1, 0)]
    got = ctx.bootstrap_key_stage(3, 1, 1, 0)
    host = got.download()
    order = sorted(diagonals)
    assert host.shape[0] == len(order) == 63
    for idx, key in enumerate(order):
        assert np.array_equal(host[idx], diagonals[key].download()), key
    got.free()
