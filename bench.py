"""Benchmark: batched HMult+relin+rescale ciphertexts/s on one MI355X.

Workload (one "step"): one batched fhe_hmult_relin_rescale launch over B
ciphertext pairs of the Gold parameter set (N = 2^16, 38-prime chain with 4
special primes), inputs rescaled at level 24 (25 limbs, moduli_id 13) with a
relinearization key at the key-switching level (moduli_id 1, 10 partitions of
which 7 are used).  Inputs and key are synthetic seeded residues, resident in
HBM before timing starts.

Printed JSON line: metric/value (whole-job ciphertexts/s), plain (B = 1)
throughput beside it, the roofline of the dominant kernel measured live with
HIP events on the library's stream, and the CPU baseline (the reference
engine built from its sources, CPU_PARALLEL on the host cores) on a bounded
sample of the same workload.  "encode_decode" adds the device encoder and
decoder on a batch of complex slot vectors at the same limb count, with the
reference's single-threaded encode_unit / decode_unit beside them.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

LOG_N = 16
ALPHA = 4
MODULI_ID = 13          # rescaled level 24: 25 limbs
HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (the library's own)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [
            ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value


def uniform_limbs(rng, primes, first, limbs, n, factor=1):
    out = np.empty((limbs, n), np.int64)
    for l in range(limbs):
        out[l] = rng.integers(0, int(primes[first + l]) * factor, n,
                              dtype=np.uint64).astype(np.int64)
    return out


def make_inputs(ctx, batch, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    limbs = ctx.limbs_at(MODULI_ID)
    key_mid = ctx.key_moduli_id()
    kp, kc, n = ctx.key_shape(key_mid)
    a = np.stack([np.stack([uniform_limbs(rng, ctx.primes, MODULI_ID, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    b = np.stack([np.stack([uniform_limbs(rng, ctx.primes, MODULI_ID, limbs, n)
                            for _ in range(2)]) for _ in range(batch)])
    key = np.stack([uniform_limbs(rng, ctx.primes, key_mid, kc, n, 2)
                    for _ in range(kp)])
    return a, b, key, key_mid


def time_steps(ctx, ev, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    ctx.synchronize()
    start, stop = ev.event(), ev.event()
    ev.record(start, ctx.stream)
    for _ in range(steps):
        fn()
    ev.record(stop, ctx.stream)
    ms = ev.elapsed_ms(start, stop)
    ctx.synchronize()
    return ms / steps


def host_cpu_share():
    """Host threads this process may use: every core it sees, capped by the
    thread count the machine's launcher grants it (OMP_NUM_THREADS, set to
    the per-GPU CPU share on a shared GPU host)."""
    nproc = os.cpu_count() or 1
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", nproc))
    except ValueError:
        share = nproc
    return nproc, max(1, min(nproc, share))


def _reference_rate(ref, budget_s, cap):
    """HMult+relin+rescale ops/s of the reference engine `ref` on Gold L=24
    fresh encryptions (rescaled inputs, as the device benchmark's)."""
    lib, e = ref.lib, ref.e
    sk = lib.ref_secret_key(e)
    pk = lib.ref_public_key(e, sk)
    rlk = lib.ref_relin_key(e, sk)
    slots = ref.slot_count
    zeros = np.zeros(slots)
    level = ref.chain_count - MODULI_ID - 1
    a0 = lib.ref_encrypt(e, zeros.ctypes.data, slots, pk, level)
    b0 = lib.ref_encrypt(e, zeros.ctypes.data, slots, pk, level)
    a, b = lib.ref_rescale(e, a0), lib.ref_rescale(e, b0)
    done, t0 = 0, time.perf_counter()
    while True:
        c = lib.ref_hmult_relin_rescale(e, a, b, rlk)
        lib.ref_free_ct(c)
        done += 1
        if time.perf_counter() - t0 > budget_s or done >= cap:
            break
    dt = time.perf_counter() - t0
    for h in (a0, b0, a, b):
        lib.ref_free_ct(h)
    lib.ref_free_relin_key(rlk)
    lib.ref_free_public_key(pk)
    lib.ref_free_secret_key(sk)
    return done, dt


def cpu_baseline(budget_s=12.0):
    """The reference engine (oracle/_ref, compiled from its sources) on the
    host's CPU share with CPU_PARALLEL, and on one thread (CPU); falls back
    to the oracle port (1 thread) when absent."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from support import Oracle, Reference
    from seeded_inputs import uniform_limbs
    nproc, threads = host_cpu_share()
    if Reference.available():
        ref = Reference(LOG_N, ALPHA, threads=threads)
        done, dt = _reference_rate(ref, budget_s, 200)
        ref.close()
        ref1 = Reference(LOG_N, ALPHA, threads=1)
        done1, dt1 = _reference_rate(ref1, budget_s, 40)
        ref1.close()
        return {"value": done / dt, "unit": "ciphertexts/s", "cores": threads,
                "nproc": nproc, "kind": "reference",
                "one_thread": {"value": done1 / dt1, "cores": 1,
                               "sample": f"{done1} ops, EngineProcessor::CPU, "
                                         f"{dt1:.1f} s"},
                "sample": f"{done} HMult+relin+rescale at Gold L=24 through "
                          f"CKKSEngine::multiply+rescale, CPU_PARALLEL "
                          f"thread_count={threads} of nproc={nproc} (the "
                          f"host's per-GPU CPU share), {dt:.1f} s"}
    import fhe_amd
    primes = fhe_amd.resolve_primes(LOG_N, ALPHA)
    orc = Oracle(LOG_N, primes, ALPHA)
    rng = np.random.Generator(np.random.PCG64(1))
    limbs = orc.limbs_at(MODULI_ID)
    n = 1 << LOG_N
    a = np.stack([uniform_limbs(rng, primes, MODULI_ID, limbs, n)
                  for _ in range(2)])
    key_mid = 1
    kc = orc.limbs_at(key_mid) + ALPHA
    kp = 2 * (-(-orc.limbs_at(key_mid) // ALPHA))
    key = np.stack([uniform_limbs(rng, primes, key_mid, kc, n, 2)
                    for _ in range(kp)])
    done, t0 = 0, time.perf_counter()
    while True:
        orc.hmult(a, a, key, MODULI_ID, key_mid)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "ciphertexts/s", "cores": 1,
            "kind": "port", "sample": f"{done} oracle HMult at Gold L=24"}


def copy_probe(mib):
    """A plain copy of known size on the device (torch's elementwise copy
    kernel), for calibrating the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.int64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8,
                      "copy_probe_bytes_written": n * 8, "launches": 5}))


def codec_bench(ctx, ev, batch, steps, with_cpu):
    """Encode and decode of `batch` complex slot vectors at the benchmark's
    level (24) on the device, values and plaintexts resident in HBM; with
    with_cpu, the reference engine's encode_unit / decode_unit on one host
    thread (its encoder runs as a single task in every processor mode)."""
    import fhe_amd
    level = ctx.info["full_level"] - MODULI_ID
    slots = ctx.n // 2
    rng = np.random.Generator(np.random.PCG64(3))
    vals = rng.uniform(-1, 1, (batch, slots, 2))
    dv = ctx.to_device(vals.view(np.int64))
    limbs = ctx.limbs_at(MODULI_ID)
    pt = ctx.array((batch, limbs, ctx.n))
    dec = ctx.array((batch, 2 * slots))
    scale = fhe_amd.params_encode_constants(LOG_N, ALPHA, level)["ckks_scale"]
    enc_ms = time_steps(ctx, ev, lambda: ctx.encode_device(dv, True, batch,
                                                           level, pt),
                        steps, 2)
    dec_ms = time_steps(ctx, ev, lambda: ctx.decode_device(pt, batch,
                                                           MODULI_ID, scale,
                                                           True, dec),
                        steps, 2)
    for arr in (dv, pt, dec):
        arr.free()
    out = {"level": level, "batch": batch,
           "encode_messages_per_s": round(batch / (enc_ms / 1e3), 1),
           "decode_messages_per_s": round(batch / (dec_ms / 1e3), 1),
           "encode_ms_per_batch": round(enc_ms, 4),
           "decode_ms_per_batch": round(dec_ms, 4)}
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from support import Reference
        if Reference.available():
            ref = Reference(LOG_N, ALPHA)
            msg = vals[0, :, 0] + 1j * vals[0, :, 1]
            reps, t0 = 0, time.perf_counter()
            while reps < 3 or time.perf_counter() - t0 < 2.0:
                _, limbs_h = ref.encode(msg, level)
                reps += 1
            enc_s = (time.perf_counter() - t0) / reps
            reps, t0 = 0, time.perf_counter()
            while reps < 3 or time.perf_counter() - t0 < 2.0:
                ref.decode(limbs_h, level, scale, True)
                reps += 1
            dec_s = (time.perf_counter() - t0) / reps
            ref.close()
            out["cpu_encode_messages_per_s"] = round(1 / enc_s, 2)
            out["cpu_decode_messages_per_s"] = round(1 / dec_s, 2)
            out["cpu_kind"] = "reference, 1 thread"
    return out


def bootstrap_bench(ev, steps, graph=True):
    """Bootstraps with a RotationKey (fhe_bootstrap: level down,
    mod_raise, coeff_to_slot, EvalMod of both parts, slot_to_coeff,
    normalization) on the N = 2^16 bootstrap preset, wall-clock per
    bootstrap including its host work.  Synthetic keys of the real shapes
    (one random key stands for every key of the key-switching level: the
    work does not depend on key values) and a random level-2 input."""
    import fhe_amd
    rng = np.random.Generator(np.random.PCG64(11))
    with fhe_amd.Context(LOG_N, ALPHA, scale_bits=54,
                         small_prime_bits=54) as ctx:
        full = ctx.info["full_level"]
        key_mid = ctx.key_moduli_id()
        kp, kc, n = ctx.key_shape(key_mid)

        def device_key(mid, polys, chain):
            # one polynomial at a time keeps the host copy small
            d = ctx.array((polys, chain, n))
            for p in range(polys):
                d.upload_at(uniform_limbs(rng, ctx.primes, mid, chain, n),
                            p * chain * n)
            return d

        # every key in its own device buffer, as a real key set is: the
        # relinearization, conjugation and to_dense keys and the 15
        # power-of-two rotation keys (18 x 217 MiB), and to_sparse
        relin, conj, dense = (device_key(key_mid, kp, kc) for _ in range(3))
        rot = [device_key(key_mid, kp, kc) for _ in range(LOG_N - 1)]
        sparse_mid = full
        sp, sc, _ = ctx.key_shape(sparse_mid)
        dsp = device_key(sparse_mid, sp, sc)
        rotations = [(rot[i], key_mid, 1 << i) for i in range(LOG_N - 1)]
        keys = fhe_amd.BootstrapKeys.make(relin, key_mid, conj, key_mid,
                                          rotations, dsp, sparse_mid, dense,
                                          key_mid)
        level = 2
        mid = full - level
        ms = {}
        for batch in (1, 8):
            ct = np.stack([np.stack([uniform_limbs(rng, ctx.primes, mid,
                                                   ctx.limbs_at(mid), n)
                                     for _ in range(2)])
                           for _ in range(batch)])
            dct = ctx.to_device(ct)
            out = ctx.array((batch, 2, ctx.chain_count - 1, n))
            ctx.bootstrap(dct, level, 0, keys, 3, out)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.bootstrap(dct, level, 0, keys, 3, out)
            ctx.synchronize()
            ms[batch] = (time.perf_counter() - t0) * 1e3 / steps
            if batch == 1 and graph:
                # the same bootstrap replayed from a HIP graph
                ctx.graph_begin()
                try:
                    ctx.bootstrap(dct, level, 0, keys, 3, out)
                finally:
                    g = ctx.graph_end()
                ctx.graph_launch(g)
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    ctx.graph_launch(g)
                ctx.synchronize()
                ms["graph"] = (time.perf_counter() - t0) * 1e3 / steps
                ctx.graph_destroy(g)
            for arr in (dct, out):
                arr.free()
        for arr in [relin, conj, dense, dsp] + rot:
            arr.free()
    return {"workload": "bootstrap with a RotationKey (15 power-of-two "
                        "keys), N=2^16 bootstrap preset, 2^15 slots, 3 "
                        "CtS / StC stages, ciphertexts from level 2, one "
                        "per call and 8 per batched call; every key in its "
                        "own device buffer (synthetic values of the real "
                        "shapes)",
            "precision": "bit-identical to the reference engine's bootstrap "
                         "(tests/test_gpu_bootstrap.py), whose decrypted "
                         "precision that test reports",
            "ms_per_bootstrap": round(ms[1], 1),
            "ms_per_bootstrap_graph": (round(ms["graph"], 1) if graph
                                       else None),
            "ms_per_batch_of_8": round(ms[8], 1),
            "ms_per_bootstrap_batched": round(ms[8] / 8, 1), "steps": steps,
            "timing": "wall clock, host work included"}


def bootstrap_bk_bench(steps, stages=3, graph=True, shift=0):
    """Bootstraps with a BootstrapKey (fhe_bootstrap_bk) on the N = 2^16
    bootstrap preset: the key's layout (stage levels, loop counts, deltas,
    38 rotation keys at their creation levels) from
    fhe_bootstrap_key_stage_layout / _rotation_layout, every key and every
    encoded diagonal in its own device buffer with synthetic values of the
    real shapes (the work does not depend on their values)."""
    import fhe_amd
    rng = np.random.Generator(np.random.PCG64(12))
    layout = fhe_amd.bootstrap_key_layout(LOG_N, ALPHA, stages, shift)
    with fhe_amd.Context(LOG_N, ALPHA, scale_bits=54,
                         small_prime_bits=54) as ctx:
        full = ctx.info["full_level"]
        key_mid = ctx.key_moduli_id()
        n = ctx.n
        arrays = []
        fills = {}

        def filled(mid, polys, limbs):
            # one random polynomial per shape, uploaded into every slot
            key = (mid, limbs)
            if key not in fills:
                fills[key] = uniform_limbs(rng, ctx.primes, mid, limbs, n)
            d = ctx.array((polys, limbs, n))
            for p in range(polys):
                d.upload_at(fills[key], p * limbs * n)
            arrays.append(d)
            return d

        def device_key(mid):
            kp, kc, _ = ctx.key_shape(mid)
            return filled(mid, kp, kc)

        relin, conj, dense = (device_key(key_mid) for _ in range(3))
        sparse = device_key(full)
        rotations = [(d, device_key(mid), mid)
                     for d, mid in layout["rotations"]]
        key_bytes = sum(a.nbytes for a in arrays)
        # the stages' diagonals encoded on the device
        # (fhe_bootstrap_key_encode_stage), each in its own slice
        mats = {}
        t0 = time.perf_counter()
        for name, to_slot in (("coeff_to_slot", 1), ("slot_to_coeff", 0)):
            mats[name] = []
            for si, st in enumerate(layout[name]):
                enc = ctx.bootstrap_key_stage(stages, shift, to_slot, si)
                arrays.append(enc)
                inner, outer = len(st["inner_deltas"]), len(st["outer_deltas"])
                step = enc.shape[1] * enc.shape[2] * 8
                pairs = [(o, i) for o in range(outer) for i in range(inner)
                         if (o, i) != (0, 0)]
                diag = {pair: enc.ptr.value + idx * step
                        for idx, pair in enumerate(pairs)}
                mats[name].append(fhe_amd.PlainMatrix.make(
                    st["level"], st["inner_deltas"], st["outer_deltas"],
                    diag))
        ctx.synchronize()
        encode_s = time.perf_counter() - t0
        all_bytes = sum(a.nbytes for a in arrays)
        parts = fhe_amd.BootstrapKeyParts.make(
            mats["coeff_to_slot"], mats["slot_to_coeff"], rotations, sparse,
            full, dense, key_mid)
        bk = ctx.bootstrap_key(parts)
        level = 2
        mid = full - level
        ms = {}
        try:
            for batch in (1, 8):
                ct = np.stack([np.stack([uniform_limbs(
                    rng, ctx.primes, mid, ctx.limbs_at(mid), n)
                    for _ in range(2)]) for _ in range(batch)])
                dct = ctx.to_device(ct)
                out = ctx.array((batch, 2, ctx.chain_count - 1, n))
                ctx.bootstrap_bk(dct, level, 0, relin, key_mid, conj, key_mid,
                                 bk, out)
                ctx.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    ctx.bootstrap_bk(dct, level, 0, relin, key_mid, conj,
                                     key_mid, bk, out)
                ctx.synchronize()
                ms[batch] = (time.perf_counter() - t0) * 1e3 / steps
                if batch == 1 and graph:
                    ctx.graph_begin()
                    try:
                        ctx.bootstrap_bk(dct, level, 0, relin, key_mid, conj,
                                         key_mid, bk, out)
                    finally:
                        g = ctx.graph_end()
                    ctx.graph_launch(g)
                    ctx.synchronize()
                    t0 = time.perf_counter()
                    for _ in range(steps):
                        ctx.graph_launch(g)
                    ctx.synchronize()
                    ms["graph"] = (time.perf_counter() - t0) * 1e3 / steps
                    ctx.graph_destroy(g)
                dct.free()
                out.free()
        finally:
            ctx.bootstrap_key_destroy(bk)
            for a in arrays:
                a.free()
    inner = len(layout["coeff_to_slot"][0]["inner_deltas"])
    outer = len(layout["coeff_to_slot"][0]["outer_deltas"])
    size = "large" if shift else "medium"
    return {"workload": f"bootstrap with a BootstrapKey (size {size}, "
                        f"{stages} CtS / StC stages of {inner} x {outer} "
                        f"diagonals, {len(layout['rotations'])} fixed "
                        "rotation keys), "
                        "N=2^16 bootstrap preset, 2^15 slots, ciphertexts "
                        "from level 2, one per call and 8 per batched call; "
                        "the diagonals encoded on the device "
                        "(fhe_bootstrap_key_encode_stage), every rotation "
                        "key in its own device buffer (synthetic values of "
                        "the real shapes)",
            "key_gb": round(all_bytes / 1e9, 2),
            "rotation_and_small_key_gb": round(key_bytes / 1e9, 2),
            "stage_encoding_s": round(encode_s, 2),
            "precision": "bit-identical to the reference engine's bootstrap "
                         "with a BootstrapKey "
                         "(tests/test_gpu_bootstrap_key.py), whose decrypted "
                         "precision that test reports",
            "ms_per_bootstrap": round(ms[1], 1),
            "ms_per_bootstrap_graph": (round(ms["graph"], 1) if graph
                                       else None),
            "ms_per_batch_of_8": round(ms[8], 1),
            "ms_per_bootstrap_batched": round(ms[8] / 8, 1), "steps": steps,
            "timing": "wall clock, host work included"}


UNEVEN_ALPHA = 8
UNEVEN_ALPHA = 8


def uneven_bench(ev, steps, count=32, autotune=True):
    """Key switching over ciphertexts at uneven levels: `count` NTT-form
    trios at limb counts drawn from 6..24 (N = 2^16, alpha = 8, dnum <= 3)
    are relinearized and each result rotated by its own random slot offset
    with a RotationKey (15 power-of-two keys), one step = the whole batch.
    Timed: the uneven batch (one launch per stage for all levels, eager and
    replayed from a HIP graph), the same batch one level group at a time
    (the same kernels per distinct moduli_id) and one ciphertext at a time
    (fhe_relinearize, fhe_rotate per fixed key).  Synthetic keys and inputs
    of the real shapes (the work does not depend on their values)."""
    import fhe_amd
    rng = np.random.Generator(np.random.PCG64(17))
    with fhe_amd.Context(LOG_N, UNEVEN_ALPHA) as ctx:
        n = ctx.n
        full = ctx.info["chain_count"]
        limbs = [int(v) for v in rng.integers(6, 25, count)]
        mids = [full - l for l in limbs]
        key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(key_mid)

        def key():
            d = ctx.array((kp, kc, n))
            for p in range(kp):
                d.upload_at(uniform_limbs(rng, ctx.primes, key_mid, kc, n, 2),
                            p * kc * n)
            return d

        rlk = key()
        rot_deltas = [1 << b for b in range(LOG_N - 1)]
        rot_keys = [key() for _ in rot_deltas]
        rot_mids = [key_mid] * len(rot_keys)
        deltas = [int(v) for v in rng.integers(1, n // 2, count)]
        trios = [ctx.to_device(np.stack([uniform_limbs(rng, ctx.primes, m, l,
                                                       n)
                                         for _ in range(3)]))
                 for m, l in zip(mids, limbs)]
        relin = [ctx.array((2, l, n)) for l in limbs]
        rot = [ctx.array((2, l, n)) for l in limbs]
        def step(sel=None):
            idx = range(count) if sel is None else sel
            ctx.relinearize_uneven([trios[c] for c in idx],
                                   [mids[c] for c in idx], rlk, key_mid,
                                   [relin[c] for c in idx])
            ctx.rotate_rotation_key_uneven(
                [relin[c] for c in idx], [mids[c] for c in idx], rot_keys,
                rot_mids, rot_deltas, [deltas[c] for c in idx],
                [rot[c] for c in idx])

        groups = {}
        for c, m in enumerate(mids):
            groups.setdefault(m, []).append(c)

        def grouped():
            for sel in groups.values():
                step(sel)

        def plain():
            for c in range(count):
                ctx.relinearize(trios[c], rlk, key_mid, relin[c], mids[c])
                for b in range(LOG_N - 1):
                    if deltas[c] & (1 << b):
                        ctx.rotate(relin[c], rot_keys[b], key_mid, 1 << b,
                                   relin[c], mids[c])

        # default schedules: the batch, one level group at a time, one
        # ciphertext at a time
        t0 = time.perf_counter()
        step()                      # builds the plans
        ctx.synchronize()
        setup_s = time.perf_counter() - t0
        default_ms = time_steps(ctx, ev, step, steps, 1)
        grouped()
        grouped_ms = time_steps(ctx, ev, grouped, max(3, steps // 2), 1)
        plain_ms = time_steps(ctx, ev, plain, max(3, steps // 4), 1)
        # autotuned schedules (tuned on the next call of each plan), eager and
        # replayed from a HIP graph of the whole step
        tune_s = 0.0
        if autotune:
            ctx.uneven_autotune(True)
            t0 = time.perf_counter()
            step()
            ctx.synchronize()
            tune_s = time.perf_counter() - t0
        batched_ms = time_steps(ctx, ev, step, steps, 1)
        ctx.graph_begin()
        try:
            step()
        finally:
            graph = ctx.graph_end()
        graph_ms = time_steps(ctx, ev, lambda: ctx.graph_launch(graph), steps,
                              2)
        ctx.graph_destroy(graph)
        schedule = ctx.uneven_schedule(True, mids, key_mid)
        rotations = sum(bin(d).count("1") for d in deltas)
        for arr in [rlk] + rot_keys + trios + relin + rot:
            arr.free()
    stage_names = ["trio_intt1", "trio_intt2", "digits", "extend",
                   "key_product", "acc_intt2", "moddown", "automorphism"]
    return {
        "workload": f"{count} ciphertexts at limb counts drawn from 6..24 "
                    "(N=2^16, alpha=8, dnum<=3): relinearize, then rotate "
                    "each by its own random slot offset with a RotationKey "
                    f"({rotations} fixed-key rotations per step)",
        "limbs": limbs,
        "level_groups": len(groups),
        "ciphertexts_per_s": round(count / (batched_ms / 1e3), 1),
        "ms_per_step": round(batched_ms, 3),
        "graph_ms_per_step": round(graph_ms, 3),
        "graph_ciphertexts_per_s": round(count / (graph_ms / 1e3), 1),
        "default_schedule_ms_per_step": round(default_ms, 3),
        "grouped_ms_per_step": round(grouped_ms, 3),
        "plain_ms_per_step": round(plain_ms, 3),
        "batched_vs_grouped": round(grouped_ms / default_ms, 3),
        "batched_vs_plain": round(plain_ms / default_ms, 3),
        "comparison": "grouped and plain against the batch with the same "
                      "(default) schedules; ms_per_step is the batch with "
                      "autotuned schedules",
        "key_switches_per_s": round((count + rotations) / (batched_ms / 1e3),
                                    1),
        "autotune": autotune,
        "relin_schedule": {name: {"cap": c, "xcd": x}
                           for name, (c, x) in zip(stage_names, schedule)},
        "first_call_s": round(setup_s, 2),
        "autotune_s": round(tune_s, 2),
    }


POLY_COEFFS = [0.25, -1.5, 0.75, 0.5, -0.125, 0.0625, 0.3, -0.2]


def serve_bench(rate=200.0, seconds=5.0, max_batch=16, seed=23):
    """The serving loop (fhe_serve) under Poisson arrivals at `rate`
    requests/s for `seconds`: each job is an encrypted dot product of two
    2^(logN-1)-slot vectors (multiply + relinearize + sum_1d_by_row) or a
    degree-7 polynomial evaluation (depth 3), on N = 2^15 or 2^16 (alpha = 4),
    starting level uniform in 8..top; keys and input ciphertexts resident on
    the device (synthetic residues of the real shapes).  Reports completed
    jobs/s and the p50 / p99 of arrival-to-completion latency, beside each
    job type's latency alone."""
    import fhe_amd
    rng = np.random.Generator(np.random.PCG64(seed))
    engines = []
    keep = []
    for log_n in (15, 16):
        ctx = fhe_amd.Context(log_n, 4)
        n = ctx.n
        key_mid = ctx.key_moduli_id()
        kp, kc, _ = ctx.key_shape(key_mid)

        def key():
            d = ctx.array((kp, kc, n))
            for p in range(kp):
                d.upload_at(uniform_limbs(rng, ctx.primes, key_mid, kc, n, 2),
                            p * kc * n)
            return d

        rlk = key()
        rot = [key() for _ in range(log_n - 1)]
        rk = fhe_amd.RotationKey.make(rot, [key_mid] * len(rot),
                                      [1 << b for b in range(log_n - 1)])
        coeffs = np.array(POLY_COEFFS)
        top = ctx.info["full_level"]
        inputs = {}
        for level in range(8, top + 1):
            mid = top - level
            inputs[level] = ctx.to_device(np.stack([uniform_limbs(
                rng, ctx.primes, mid, ctx.limbs_at(mid), n)
                for _ in range(2)]))
        outs = [ctx.array((2, ctx.limbs_at(0), n)) for _ in range(64)]
        e = fhe_amd.ServeEngine()
        e.ctx, e.rlk, e.rlk_moduli_id = ctx.handle, rlk.ptr.value, key_mid
        e.rotation = rk
        e.coefficients, e.coefficient_count = coeffs.ctypes.data, len(coeffs)
        engines.append(dict(ctx=ctx, engine=e, top=top, inputs=inputs,
                            outs=outs))
        keep += [rlk, rk, coeffs] + rot
    count = int(rate * seconds)
    arrivals = np.cumsum(rng.exponential(1e3 / rate, count))
    jobs = (fhe_amd.Job * count)()
    kinds = rng.integers(0, 2, count)
    which = rng.integers(0, 2, count)
    for j in range(count):
        g = engines[int(which[j])]
        level = int(rng.integers(8, g["top"] + 1))
        jobs[j].engine, jobs[j].kind, jobs[j].level = int(which[j]), \
            int(kinds[j]), level
        jobs[j].a = jobs[j].b = g["inputs"][level].ptr.value
        jobs[j].out = g["outs"][j % len(g["outs"])].ptr.value
        jobs[j].arrival_ms = float(arrivals[j])

    def alone(kind, g):
        one = (fhe_amd.Job * 1)()
        level = engines[g]["top"]
        one[0].engine, one[0].kind, one[0].level = g, kind, level
        one[0].a = one[0].b = engines[g]["inputs"][level].ptr.value
        one[0].out = engines[g]["outs"][0].ptr.value
        fhe_amd.serve([x["engine"] for x in engines], one)
        fhe_amd.serve([x["engine"] for x in engines], one)
        return round(one[0].done_ms - one[0].arrival_ms, 2)

    # warm: plans, pools and workspaces of every shape
    alone_ms = {f"{k}_n{15 + g}": alone(kk, g)
                for g in (0, 1) for k, kk in (("dot", 0), ("poly", 1))}
    warm = (fhe_amd.Job * min(count, 64))()
    ctypes.memmove(warm, jobs, ctypes.sizeof(warm))
    for j in range(len(warm)):
        warm[j].arrival_ms = 0.0
    fhe_amd.serve([x["engine"] for x in engines], warm, max_batch=max_batch)
    stats = fhe_amd.serve([x["engine"] for x in engines], jobs,
                          max_batch=max_batch)
    for g in engines:
        g["ctx"].synchronize()
    lat = np.array([jobs[j].done_ms - jobs[j].arrival_ms
                    for j in range(count)])
    done = max(jobs[j].done_ms for j in range(count))
    # saturation: the same job mix offered ten times faster, so that jobs
    # queue and batches fill
    sat_rate = rate * 10
    sat_count = min(count * 2, int(sat_rate * 1.0))
    sat = (fhe_amd.Job * sat_count)()
    for j in range(sat_count):
        ctypes.memmove(ctypes.byref(sat, j * ctypes.sizeof(fhe_amd.Job)),
                       ctypes.byref(jobs, (j % count) *
                                    ctypes.sizeof(fhe_amd.Job)),
                       ctypes.sizeof(fhe_amd.Job))
        sat[j].arrival_ms = float(j * 1e3 / sat_rate)
    sat_stats = fhe_amd.serve([x["engine"] for x in engines], sat,
                              max_batch=max_batch)
    for g in engines:
        g["ctx"].synchronize()
    sat_lat = np.array([sat[j].done_ms - sat[j].arrival_ms
                        for j in range(sat_count)])
    sat_done = max(sat[j].done_ms for j in range(sat_count))
    for g in engines:
        for arr in list(g["inputs"].values()) + g["outs"]:
            arr.free()
    for k in keep:
        if hasattr(k, "free"):
            k.free()
    for g in engines:
        g["ctx"].close()
    return {
        "workload": f"Poisson arrivals at {rate:g} requests/s for "
                    f"{seconds:g} s ({count} jobs): encrypted dot products "
                    "(multiply + relinearize + sum_1d_by_row) and degree-7 "
                    "polynomial evaluations, N = 2^15 or 2^16 (alpha 4), "
                    "starting levels uniform in 8..top; keys and inputs "
                    "resident on the device",
        "jobs": count,
        "throughput_jobs_per_s": round(count / (done / 1e3), 1),
        "offered_jobs_per_s": rate,
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
        "latency_ms_max": round(float(lat.max()), 2),
        "alone_latency_ms": alone_ms,
        "batches": stats.batches,
        "mean_batch_jobs": round(count / max(1, stats.batches), 2),
        "max_batch_jobs": stats.max_batch_jobs,
        "deferred_admissions": stats.deferred_admissions,
        "min_free_hbm_gb": round(stats.min_free_bytes / 1e9, 1),
        "max_batch": max_batch,
        "saturation": {
            "offered_jobs_per_s": sat_rate,
            "jobs": sat_count,
            "throughput_jobs_per_s": round(sat_count / (sat_done / 1e3), 1),
            "latency_ms_p50": round(float(np.percentile(sat_lat, 50)), 2),
            "latency_ms_p99": round(float(np.percentile(sat_lat, 99)), 2),
            "batches": sat_stats.batches,
            "mean_batch_jobs": round(sat_count / max(1, sat_stats.batches),
                                     2),
        },
    }


# The unit of work's algorithmic cost (Gold, N = 2^16, l = 25 -> 24,
# alpha = 4): HBM bytes of one HMult+relin+rescale when B ciphertexts share
# one key stream, 8N [4l + 2(l - 1) + 2 ceil(l / alpha)(l + alpha) / B] (read
# both rescaled inputs, write the output, read the key once per batch), and
# its modular products: 436 NTT-equivalents (100 input NTTs, 75 trio INTTs,
# 203 ModUp NTTs, 58 ModDown INTTs) of (N / 2) log2 N butterflies each, one
# product per butterfly, plus about 1.37e8 elementwise products (tensor,
# Garner digits and basis extension, key products, ModDown, rescale).
def unit_bytes(n, limbs, alpha, batch):
    parts = -(-limbs // alpha)
    return 8 * n * (4 * limbs + 2 * (limbs - 1) +
                    2 * parts * (limbs + alpha) / batch)


def unit_modmuls(n, limbs, alpha):
    parts = -(-limbs // alpha)
    ntts = 4 * limbs + 3 * limbs + parts * (limbs + alpha) + 2 * (limbs +
                                                                   alpha)
    log_n = n.bit_length() - 1
    return ntts * (n // 2) * log_n + 1.37e8


# Modular products of one launch of each single-kernel stage (one per
# butterfly, one per pointwise product); shape keys: n, count, limbs, ext
# (= limbs + alpha), parts.
def stage_modmuls(name, s):
    c, e, p, n = s["count"], s["ext"], s["parts"], s["n"]
    half_pass = 4 * n  # butterflies of one 8-bit pass over one limb (N/2 * 8)
    if name == "ks_inner_product_r16":
        # the low forward pass and two key products per extended limb of
        # every partition, the low inverse pass of both accumulators
        return c * p * e * (half_pass + 2 * n) + 2 * c * e * half_pass
    return None


# Kernel name of each single-kernel stage, for the recorded HBM traffic.
STAGE_KERNEL = {"ks_inner_product_r16": "ks_inner_product_r16_kernel"}
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "hmult_gold_b32_traffic.md")


def recorded_figure(field, batch):
    """A figure recorded for the default batch of 32 in the rocprofv3
    counter summary (`<field>: <integer>` lines); None when absent."""
    if batch != 32 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    key = f"{field}:"
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.startswith(key):
                return int(line[len(key):].strip())
    return None


def recorded_traffic(stage, batch):
    """HBM bytes per launch of a stage's kernel from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of known
    size); None when absent."""
    kernel = STAGE_KERNEL.get(stage)
    if kernel is None:
        return None
    return recorded_figure(f"traffic_bytes_per_launch {kernel}", batch)


def valu_issue(stage, batch, avg_ms):
    """The kernel's VALU issue rate: its SQ_INSTS_VALU count per launch
    (recorded) over the launch duration measured here, against the issue
    rate of the modular-product probe (SQ_INSTS_VALU over its duration)."""
    kernel = STAGE_KERNEL.get(stage)
    if kernel is None:
        return None
    insts = recorded_figure(f"valu_insts_per_launch {kernel}", batch)
    ceiling = recorded_figure("valu_issue_ceiling_per_s", batch)
    if insts is None or ceiling is None:
        return None
    rate = insts / (avg_ms / 1e3)
    return {"unit": "VALU wave-instructions/s",
            "insts_per_launch": insts,
            "achieved": round(rate, 0),
            "ceiling": ceiling,
            "frac": round(rate / ceiling, 4)}


def stage_report(stages, step_ms, prof_steps, shape, peaks):
    """Per-stage share of the step and the roofline of the dominant stage
    (the one with the largest total time; one kernel launch per stage
    call), against both bounds: HBM with the unit's algorithmic bytes, and
    the device's measured modular-product rate."""
    total_ms = step_ms * prof_steps
    breakdown = {}
    for name, (ms, launches) in sorted(stages.items(), key=lambda kv: -kv[1][0]):
        breakdown[name] = {"avg_launch_ms": round(ms / launches, 4),
                           "launches": launches,
                           "share": round(ms / total_ms, 4)}
    name = max(stages, key=lambda k: stages[k][0])
    avg = stages[name][0] / stages[name][1]
    n, limbs, batch = shape["n"], shape["limbs"], shape["count"]
    nbytes = unit_bytes(n, limbs, ALPHA, batch) * batch
    gbs = nbytes / (avg / 1e3) / 1e9
    step_gbs = nbytes / (step_ms / 1e3) / 1e9
    kernel_mults = stage_modmuls(name, shape)
    step_mults = unit_modmuls(n, limbs, ALPHA) * batch / (step_ms / 1e3)
    valu = {"unit": "modular products/s",
            "peak_shoup": round(peaks[0], 0),
            "peak_montgomery": round(peaks[1], 0),
            "step_achieved": round(step_mults, 0),
            "step_frac": round(step_mults / peaks[0], 4)}
    if kernel_mults is not None:
        k = kernel_mults / (avg / 1e3)
        valu["kernel_achieved"] = round(k, 0)
        valu["kernel_frac"] = round(k / peaks[0], 4)
    roofline = {
        "bound": "valu",
        "kernel": name,
        "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": recorded_traffic(name, batch),
        "avg_launch_ms": round(avg, 4),
        "algorithmic_bytes": int(nbytes),
        "algorithmic_bytes_basis": "8N[4l+2(l-1)+2ceil(l/a)(l+a)/B] per "
                                   f"ciphertext x B={batch} (l={limbs}, "
                                   f"a={ALPHA})",
        "step_achieved": round(step_gbs, 1),
        "step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "valu": valu,
        "valu_issue": valu_issue(name, batch, avg),
        "binding": "VALU: 62-bit modular products from 32-bit multiply-adds; "
                   "the HBM fractions are of the algorithmic bytes, the "
                   "product fractions of the Shoup-product rate measured "
                   "on this device (fhe_modmul_peak), valu_issue the "
                   "kernel's VALU instructions (rocprofv3 SQ_INSTS_VALU) "
                   "against the issue rate of the product probe",
    }
    return breakdown, roofline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 steps of a batch of 32 keep the device busy for about 2.5 s
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the one-pair (B = 1) comparison, e.g. for a "
                         "profile whose per-kernel averages cover batch B "
                         "launches only")
    ap.add_argument("--no-codec", action="store_true",
                    help="skip the encode / decode measurement")
    ap.add_argument("--no-bootstrap", action="store_true",
                    help="skip the bootstrap measurement")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-level key-switching measurement")
    ap.add_argument("--only-uneven", action="store_true",
                    help="only run the uneven-level key-switching "
                         "measurement and print its JSON object")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving-loop measurement")
    ap.add_argument("--only-serve", action="store_true",
                    help="only run the serving-loop measurement")
    ap.add_argument("--only-bootstrap", action="store_true",
                    help="only run the bootstrap measurement (e.g. under a "
                         "profiler) and print its JSON object")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the HIP-graph replays of the bootstrap legs "
                         "(for runs under rocprofv3, whose tracing of graph "
                         "launches with copy nodes is not supported)")
    ap.add_argument("--only-bootstrap-key", action="store_true",
                    help="only run the BootstrapKey bootstrap measurement "
                         "and print its JSON object")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times "
                         "(the reference kernel for scaling the FETCH_SIZE "
                         "and WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only (--gpus 1): batches shard "
              "across GPUs with replicated keys and no communication, and "
              "that multi-GPU form is not measured.", file=sys.stderr)
        sys.exit(2)

    import fhe_amd
    ev = HipEvents()
    if args.only_bootstrap:
        print(json.dumps(bootstrap_bench(ev, 3, not args.no_graph)))
        return
    if args.only_bootstrap_key:
        print(json.dumps(bootstrap_bk_bench(3, graph=not args.no_graph)))
        print(json.dumps(bootstrap_bk_bench(3, graph=not args.no_graph,
                                            shift=1)))
        return
    if args.only_uneven:
        print(json.dumps(uneven_bench(ev, 10)))
        return
    if args.only_serve:
        print(json.dumps(serve_bench()))
        return
    with fhe_amd.Context(LOG_N, ALPHA) as ctx:
        limbs = ctx.limbs_at(MODULI_ID)
        n = ctx.n
        a, b, key, key_mid = make_inputs(ctx, args.batch, 7)
        da, db, dk = ctx.to_device(a), ctx.to_device(b), ctx.to_device(key)
        out = ctx.array((args.batch, 2, limbs - 1, n))
        step_ms = time_steps(
            ctx, ev, lambda: ctx.hmult_relin_rescale(da, db, dk, key_mid, out,
                                                     MODULI_ID),
            args.steps, args.warmup)
        # plain form: one ciphertext pair per launch
        plain_ms = None
        if not args.no_plain:
            pa, pb = ctx.to_device(a[:1]), ctx.to_device(b[:1])
            pout = ctx.array((1, 2, limbs - 1, n))
            plain_ms = time_steps(
                ctx, ev, lambda: ctx.hmult_relin_rescale(pa, pb, dk, key_mid,
                                                         pout, MODULI_ID),
                max(5, args.steps), 2)
            for arr in (pa, pb, pout):
                arr.free()
        # per-stage launch times, HIP events on the library's stream
        ctx.profile(True)
        prof_steps = max(3, min(args.steps, 10))
        for _ in range(prof_steps):
            ctx.hmult_relin_rescale(da, db, dk, key_mid, out, MODULI_ID)
        stages = ctx.profile_entries()
        ctx.profile(False)
        peaks = (ctx.modmul_peak(0), ctx.modmul_peak(1))
        parts = -(-limbs // ALPHA)
        for arr in (da, db, dk, out):
            arr.free()
        codec = None
        if not args.no_codec:
            codec = codec_bench(ctx, ev, args.batch, max(5, args.steps // 2),
                                not args.no_cpu)

    graph = not args.no_graph
    boot = None if args.no_bootstrap else bootstrap_bench(ev, 3, graph)
    boot_bk = (None if args.no_bootstrap
               else bootstrap_bk_bench(3, graph=graph))
    boot_bk_large = (None if args.no_bootstrap
                     else bootstrap_bk_bench(3, graph=graph, shift=1))
    uneven = None if args.no_uneven else uneven_bench(ev, 10)
    served = None if args.no_serve else serve_bench()
    value = args.batch / (step_ms / 1e3)
    shape = dict(n=n, count=args.batch, limbs=limbs, ext=limbs + ALPHA,
                 parts=parts)
    breakdown, roofline = stage_report(stages, step_ms, prof_steps, shape,
                                       peaks)
    line = {
        "metric": "HMult+relin+rescale ciphertexts/s at N=2^16,L=24 (batched)",
        "value": round(value, 2),
        "unit": "ciphertexts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64 (62-bit Montgomery residues)",
        "data": "synthetic",
        "config": {"workload": "Gold HMult+relin+rescale, N=2^16, L=24 "
                               "(25->24 limbs), alpha=4, dnum=7, batch "
                               f"{args.batch} ciphertext pairs per launch",
                   "batch": args.batch, "log_n": LOG_N, "moduli_id": MODULI_ID},
        "plain_ciphertexts_per_s": (round(1 / (plain_ms / 1e3), 2)
                                    if plain_ms else None),
        "batched_vs_plain": (round(value * plain_ms / 1e3, 3)
                             if plain_ms else None),
        "roofline": roofline,
        "stages": breakdown,
        "encode_decode": codec,
        "bootstrap": boot,
        "bootstrap_bootstrap_key": boot_bk,
        "bootstrap_bootstrap_key_large": boot_bk_large,
        "uneven_key_switch": uneven,
        "serving": served,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
