"""GPU parity of encoding, bit-exact against the reference engine's
encode_unit (oracle/_ref) and the oracle's restatement:

  * batches of real and complex slot vectors at the top level of the
    one-more-level chain (limb 0 zero), middle and bottom levels;
  * values that take the double_to_rns branches (above 2^59 and 2^62 after
    scaling), the zero vector and a short zero-padded vector;
  * the N = 2^16 default and bootstrap presets (scale 2^40 and 2^54);
  * a batch whose members take different branches (the 2^59 test is per
    message);
  * invalid levels are refused.
"""
import numpy as np
import pytest

import fhe_amd
from support import Oracle, Reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not Reference.available(),
                                 reason="oracle/_ref not built")]


def _messages(slots, rng):
    short = np.zeros(slots)
    short[:5] = [0.5, -0.25, 3.0, 1e-3, -7.0]
    return [
        rng.uniform(-1, 1, slots),
        rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots),
        short,
        np.zeros(slots),
        rng.uniform(-1, 1, slots) * 1e6,
        rng.uniform(-1, 1, slots) * 3e7,
    ]


@pytest.mark.parametrize("log_n,alpha,boot", [(13, 1, 0), (16, 4, 0),
                                              (16, 4, 1)])
def test_encode_matches_reference(log_n, alpha, boot, gpu_lib):
    ref = Reference(log_n, alpha, bootstrap=boot)
    kw = dict(scale_bits=54 if boot else 40,
              small_prime_bits=54 if boot else 40)
    ctx = fhe_amd.Context(log_n, alpha, primes=ref.primes, **kw)
    try:
        rng = np.random.default_rng(log_n + boot)
        full = ref.chain_count - 2
        levels = sorted({full, full - 1, full // 2, 0}) if log_n < 16 \
            else [full, 2]
        msgs = _messages(ref.slot_count, rng)
        if log_n == 16:
            msgs = msgs[:2] + msgs[4:5]
        for level in levels:
            print(f"encode log_n={log_n} boot={boot} level={level}",
                  flush=True)
            reals = [m for m in msgs if not np.iscomplexobj(m)]
            cplx = [m for m in msgs if np.iscomplexobj(m)]
            # a real batch (mixed branches) and a complex batch
            for batch in (np.stack(reals), np.stack(cplx)):
                out = ctx.encode(batch, level)
                got = out.download()
                out.free()
                for i, m in enumerate(batch):
                    mid, want = ref.encode(m, level)
                    assert got[i].shape == want.shape, (level, i)
                    assert (got[i] == want).all(), (level, i)
    finally:
        ctx.close()
        ref.close()


def test_encode_matches_oracle_batch(gpu_lib):
    """A batch of eight complex messages at one level against the oracle fed
    the product's own host constants."""
    ref = Reference(14, 2)
    ctx = fhe_amd.Context(14, 2, primes=ref.primes)
    orc = Oracle(14, ref.primes, 2)
    try:
        rng = np.random.default_rng(3)
        full = ref.chain_count - 2
        level = full - 3
        batch = rng.normal(size=(8, ref.slot_count)) + \
            1j * rng.normal(size=(8, ref.slot_count))
        out = ctx.encode(batch, level)
        got = out.download()
        out.free()
        c = fhe_amd.params_encode_constants(14, 2, level)
        mid = full - level
        limbs = ref.chain_count - mid
        for i in range(8):
            want = orc.encode(batch[i], c["scale_before_tile"], 0, mid, limbs,
                              c["r_mult_scale_after_tile"])
            assert (got[i] == want).all(), i
    finally:
        ctx.close()
        ref.close()


def test_encode_rejects_bad_level(gpu_lib):
    ctx = fhe_amd.Context(13, 1)
    try:
        full = ctx.info["full_level"]
        for level in (-1, full + 1):
            with pytest.raises(fhe_amd.FheError):
                ctx.encode(np.zeros(ctx.n // 2), level)
    finally:
        ctx.close()
