"""GPU tests of the serving loop (fhe_serve): jobs of two ring degrees
(N = 2^15 and 2^16, alpha = 4), encrypted dot products (multiply + relinearize
+ sum_1d_by_row) and degree-7 polynomial evaluations, starting levels mixed
between 8 and the top, keys and ciphertexts resident on the device:

  * every job's output limbs are the same alone and in the mix (continuous
    batching with staggered arrivals), and equal the direct composition of
    the engine entry points (fhe_multiply + fhe_sum_1d_by_row,
    fhe_evaluate_polynomial), which test_gpu_advanced.py checks against the
    reference engine;
  * the loop reports per-job admission and completion times in order;
  * invalid jobs are refused before any launch.

Keys and inputs are synthetic residues of the real shapes (the work and the
batching do not depend on their values).
"""
import ctypes

import numpy as np
import pytest

import fhe_amd
from support import uniform_limbs

pytestmark = pytest.mark.gpu

COEFFS = [0.25, -1.5, 0.75, 0.5, -0.125, 0.0625, 0.3, -0.2]


class Served:
    """One engine: context, relinearization key, RotationKey, coefficients."""

    def __init__(self, log_n, seed):
        self.ctx = fhe_amd.Context(log_n, 4)
        ctx = self.ctx
        rng = np.random.Generator(np.random.PCG64(seed))
        self.rng = rng
        n = ctx.n
        self.key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(self.key_mid)

        def key():
            d = ctx.array((kp, kc, n))
            for p in range(kp):
                d.upload_at(uniform_limbs(rng, ctx.primes, self.key_mid, kc,
                                          n, 2), p * kc * n)
            return d

        self.rlk = key()
        self.rot = [key() for _ in range(log_n - 1)]
        self.rk = fhe_amd.RotationKey.make(
            self.rot, [self.key_mid] * len(self.rot),
            [1 << b for b in range(log_n - 1)])
        self.coeffs = np.array(COEFFS)
        self.arrays = [self.rlk] + self.rot

    def engine(self):
        e = fhe_amd.ServeEngine()
        e.ctx = self.ctx.handle
        e.rlk = self.rlk.ptr.value
        e.rlk_moduli_id = self.key_mid
        e.rotation = self.rk
        e.coefficients = self.coeffs.ctypes.data
        e.coefficient_count = len(self.coeffs)
        return e

    def duo(self, level):
        ctx = self.ctx
        mid = ctx.info["full_level"] - level
        arr = ctx.to_device(np.stack([uniform_limbs(
            self.rng, ctx.primes, mid, ctx.limbs_at(mid), ctx.n)
            for _ in range(2)]))
        self.arrays.append(arr)
        return arr

    def close(self):
        for arr in self.arrays:
            arr.free()
        self.ctx.close()


@pytest.fixture(scope="module")
def served():
    engines = [Served(15, 1), Served(16, 2)]
    yield engines
    for e in engines:
        e.close()


def make_jobs(served, count, seed):
    """(engine, kind, level, a, b, out) of `count` mixed jobs."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(count):
        g = int(rng.integers(0, 2))
        s = served[g]
        top = s.ctx.info["full_level"]
        level = int(rng.integers(8, top + 1))
        kind = int(rng.integers(0, 2))
        a = s.duo(level)
        b = s.duo(level) if kind == fhe_amd.JOB_DOT else None
        out = s.ctx.array(a.shape)
        s.arrays.append(out)
        specs.append((g, kind, level, a, b, out))
    return specs


def run(served, specs, arrivals):
    jobs = (fhe_amd.Job * len(specs))()
    for j, (g, kind, level, a, b, out) in enumerate(specs):
        jobs[j].engine, jobs[j].kind, jobs[j].level = g, kind, level
        jobs[j].a = a.ptr.value
        jobs[j].b = b.ptr.value if b is not None else None
        jobs[j].out = out.ptr.value
        jobs[j].arrival_ms = arrivals[j]
    stats = fhe_amd.serve([s.engine() for s in served], jobs, max_batch=8)
    for s in served:
        s.ctx.synchronize()
    return jobs, stats


def test_jobs_same_alone_and_in_the_mix(served, gpu_lib):
    specs = make_jobs(served, 24, 3)
    # staggered arrivals over ~40 ms: batches form while others run
    arrivals = list(np.cumsum(np.random.default_rng(4).exponential(1.5, 24)))
    jobs, stats = run(served, specs, arrivals)
    assert stats.batches >= 2
    mix = [out.download() for (_, _, _, _, _, out) in specs]
    for j in range(len(specs)):
        assert jobs[j].admitted_ms >= jobs[j].arrival_ms - 1e-6
        assert jobs[j].done_ms >= jobs[j].admitted_ms
    for j in (0, 5, 11, 17, 23):
        spec = specs[j]
        ctx = served[spec[0]].ctx
        words = 2 * ctx.limbs_at(ctx.info["full_level"] -
                                 jobs[j].out_level) * ctx.n
        spec[5].upload(np.zeros(spec[5].shape, np.int64))
        alone, _ = run(served, [spec], [0.0])
        assert (alone[0].out_level, alone[0].out_flags) == (
            jobs[j].out_level, jobs[j].out_flags)
        assert np.array_equal(spec[5].download().reshape(-1)[:words],
                              mix[j].reshape(-1)[:words]), j


def test_jobs_equal_direct_compositions(served, gpu_lib):
    specs = make_jobs(served, 6, 5)
    jobs, _ = run(served, specs, [0.0] * len(specs))
    for j, (g, kind, level, a, b, out) in enumerate(specs):
        s = served[g]
        ctx = s.ctx
        want = ctx.array(a.shape)
        if kind == fhe_amd.JOB_DOT:
            mid = ctx.info["full_level"] - level
            prod = ctx.array(a.shape)
            ctx.multiply(a, level, 0, b, level, 0, prod, key=s.rlk,
                         key_moduli_id=s.key_mid)
            lv, fl = ctx.sum_1d_by_row(prod, level - 1, 0, s.rk, want)
            prod.free()
        else:
            lv, fl = ctx.evaluate_polynomial(a, level, COEFFS, s.rlk,
                                             s.key_mid, want)
        assert (lv, fl) == (jobs[j].out_level, jobs[j].out_flags), j
        limbs = ctx.limbs_at(ctx.info["full_level"] - lv)
        got = out.download().reshape(-1)[:2 * limbs * ctx.n]
        exp = want.download().reshape(-1)[:2 * limbs * ctx.n]
        assert np.array_equal(got, exp), j
        want.free()


def test_invalid_jobs_are_refused(served, gpu_lib):
    s = served[0]
    a = s.duo(10)
    bad = [(2, fhe_amd.JOB_POLY, 10, a, None, a),      # engine out of range
           (0, 7, 10, a, None, a),                    # unknown kind
           (0, fhe_amd.JOB_DOT, 10, a, None, a),      # DOT without b
           (0, fhe_amd.JOB_POLY, 99, a, None, a)]     # level out of range
    for spec in bad:
        with pytest.raises(fhe_amd.FheError):
            run(served, [spec], [0.0])
