"""GPU parity of encryption arithmetic and decryption, bit-exact.

  * fhe_decrypt equals the reference's decrypt_to_unit_plaintext limbs on
    reference encryptions (oracle/_ref), batched;
  * fhe_encrypt_pk with host samples (u, e0, e1) equals the oracle's
    encryption with the same samples, and decrypts to the message plus a
    small noise polynomial.
"""
import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference, ptr

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


@pytest.mark.parametrize("log_n,alpha,levels", [(14, 2, [6, 3, 1]),
                                                (16, 4, [24, 5])])
def test_decrypt_matches_reference(log_n, alpha, levels, gpu_lib):
    ref = Reference(log_n, alpha, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    sk_mid, skd = ref.key_data(sk, 3)
    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
    dsk = ctx.to_device(skd[0])
    rng = np.random.default_rng(log_n)
    for level in levels:
        cts, expect = [], []
        for _ in range(3):
            x = rng.uniform(-1, 1, ref.slot_count)
            ct = lib.ref_encrypt(e, ptr(x), ref.slot_count, pk, level)
            cts.append(ref.get(ct))
            mid, pt = ref.decrypt_limbs(ct, sk)
            expect.append(pt)
            lib.ref_free_ct(ct)
        dct = ctx.to_device(np.stack(cts))
        out = ctx.array((3,) + expect[0].shape)
        ctx.decrypt(dct, dsk, sk_mid, out, mid)
        assert (out.download() == np.stack(expect)).all(), level
        dct.free()
        out.free()
    dsk.free()
    ctx.close()
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()


@pytest.mark.parametrize("log_n,alpha,level", [(14, 2, 4), (16, 4, 24)])
def test_encrypt_matches_oracle_and_decrypts(log_n, alpha, level, gpu_lib):
    ref = Reference(log_n, alpha, threads=8)
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    sk_mid, skd = ref.key_data(sk, 3)
    pk_mid, pkd = ref.key_data(pk, 4)
    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes)
    info = ctx.info
    mid = info["full_level"] - level
    limbs = ctx.limbs_at(mid)
    n = ctx.n
    rng = np.random.default_rng(17)
    count = 2
    m = rng.integers(-(1 << 30), (1 << 30) + 1, (count, n), dtype=np.int64)
    qs = ref.primes[mid:mid + limbs].astype(np.int64)[None, :, None]
    pt = np.where(m[:, None, :] < 0, m[:, None, :] + qs, m[:, None, :])
    pt = np.ascontiguousarray(np.broadcast_to(pt, (count, limbs, n)))
    u = rng.integers(0, 2, (count, n), dtype=np.int64)
    e0 = rng.integers(-20, 21, (count, n), dtype=np.int64)
    e1 = rng.integers(-20, 21, (count, n), dtype=np.int64)
    dpt, dpk = ctx.to_device(pt), ctx.to_device(pkd)
    du, de0, de1 = ctx.to_device(u), ctx.to_device(e0), ctx.to_device(e1)
    out = ctx.array((count, 2, limbs, n))
    ctx.encrypt_pk(dpt, dpk, pk_mid, du, de0, de1, out, mid)
    got = out.download()
    if log_n <= 14:
        orc = Oracle(log_n, ref.primes, alpha)
        for c in range(count):
            expect = orc.encrypt_pk(pt[c], pkd, pk_mid, u[c], e0[c], e1[c],
                                    mid)
            assert (got[c] == expect).all(), c
    dsk = ctx.to_device(skd[0])
    dec = ctx.array((count, limbs, n))
    ctx.decrypt(out, dsk, sk_mid, dec, mid)
    noise = (dec.download() - pt) % qs
    noise = np.where(noise > qs // 2, noise - qs, noise)
    assert (noise == noise[:, :1]).all()
    assert np.abs(noise).max() < 1 << 24
    for d in (dpt, dpk, du, de0, de1, out, dsk, dec):
        d.free()
    ctx.close()
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    ref.close()
