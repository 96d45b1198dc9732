#!/usr/bin/env python3
"""Benchmark of the NTT / BGV path on one MI355X (prints ONE JSON line).

Headline (`value`): batched negacyclic polynomial products, 512 pairs, N = 2^16,
q = 2^124 + 2^20 + 2^18 + 1 (4 x 32-bit limbs): forward NTT of both inputs, point-wise
Montgomery product, inverse NTT, inputs resident in HBM. Unit: Gcoeff/s = (sum over the
three transforms of N) / time. Beside it the line carries the BGV encryption rate
(1024 ciphertexts, N = 2^14, batched vs the plain per-ciphertext path), the N = 4096
NTT->INTT latency p50/p99, the roofline of the dominant step and a CPU baseline (the
oracle restatement of hmpc's algorithm on the host cores, bounded sample).

Inputs are synthetic: explicit-key ChaCha20 streams with key {42, 0, ..., 0}.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

KEY = [42, 0, 0, 0, 0, 0, 0, 0]
ENC_KEY = [42, 0, 0, 0, 0, 0, 0, 1]
HBM_PEAK_GBS = 8000.0
VALU_CLOCK_GHZ = 2.4  # MI355X peak engine clock
METRIC = "BGV enc/s + NTT Gcoeff/s, batched vs plain on 1 MI355X, bit-exact; p50/p99"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), p))


def time_steps(ctx, stream, fn, steps, warmup):
    """W untimed steps, then K timed steps bracketed by synchronisation; per-step events."""
    for _ in range(warmup):
        fn()
    stream.synchronize()
    ctx.wait()
    events = [ctx.event() for _ in range(steps + 1)]
    t0 = time.perf_counter()
    events[0].record(stream)
    for k in range(steps):
        fn()
        events[k + 1].record(stream)
    stream.synchronize()
    ctx.wait()
    wall = time.perf_counter() - t0
    per = [events[k].elapsed_ms(events[k + 1]) for k in range(steps)]
    total = events[0].elapsed_ms(events[-1])
    return total / steps, per, wall


# algorithmic HBM bytes per launch of each fast kernel, in units of (elements x 4L bytes): what
# the kernel must read and write at least (a, b in, product out; column-pass results in between)
KERNEL_PASSES = {
    "fast persistent jobs": 3,           # polymul: a, b read, product written
    "fast polymul forward columns": 4,   # a, b read; both column results written
    "fast polymul rows": 3,              # both column results read, row result written
    "fast polymul inverse columns": 2,   # row result read, product written
    "fast polymul (whole tile)": 3,
}


# --no-autotune: every modulus keeps the library's default schedule (profiling runs: the tuner times
# its trials, and under a profiler those timings do not reflect the device)
AUTOTUNE = [True]


def tune(mod, log_n, trials, stream):
    if AUTOTUNE[0]:
        return mod.autotune(log_n, trials, stream=stream)
    return dict(mod.schedule(log_n), default=True)


def tune_bgv(mod, log_n, trials, stream):
    if AUTOTUNE[0]:
        return mod.autotune_bgv(log_n, trials, stream=stream)
    return {"default": True}


def bench_polymul(hm, ctx, stream, steps, warmup, n=512, log_n=16):
    mod = hm.Modulus(ctx, hm.Q128A)
    mod.prepare(log_n)
    count = n << log_n
    log(f"[bench] C2 polymul: {n} pairs, N=2^{log_n}, L={mod.limbs}")
    a = mod.sample_uniform(KEY, 0, count)
    b = mod.sample_uniform(KEY, 1, count)
    out = mod.alloc(count)
    tuned = tune(mod, log_n, 64, stream)
    log(f"[bench] autotuned schedule: {tuned}")
    ctx.set_timing(True)
    ms, per, wall = time_steps(ctx, stream, lambda: mod.polymul(log_n, n, a, b, out, stream=stream), steps, warmup)
    kernels = ctx.kernel_times()
    ctx.set_timing(False)
    res = {"mod": mod, "a": a, "b": b, "out": out, "n": n, "log_n": log_n, "ms": ms, "per": per, "wall": wall,
           "schedule": tuned, "kernels": kernels}
    # plain per-op path for the same product: per-stage launches, separate point-wise kernel
    fa, fb = mod.alloc(count), mod.alloc(count)

    def plain():
        mod.ntt(log_n, n, a, fa, plain=True, stream=stream)
        mod.ntt(log_n, n, b, fb, plain=True, stream=stream)
        mod.pointwise(mod.MUL, count, fa, fb, out=fa, stream=stream)
        mod.ntt(log_n, n, fa, out, inverse=True, plain=True, stream=stream)

    pms, _, _ = time_steps(ctx, stream, plain, max(2, steps // 4), 1)
    res["plain_ms"] = pms
    # the timed product again (the plain path wrote `out` last)
    mod.polymul(log_n, n, a, b, out, stream=stream)
    stream.synchronize()
    del fa, fb
    return res


# capability-stream width of the uniform ternary draw (HMPC_BITS_TERNARY in include/hmpc_hip.h)
BITS_TERNARY = 0
# C3's stated randomness: s and u uniform ternary, errors centred binomial with eta = 10 (20 bits)
C3_BITS = (BITS_TERNARY, 20, 20)
# hmpc's lhe::randomness(): u, v, w centred binomial with variances 1/2, 10, 10 (2, 40, 40 bits)
HMPC_BITS = (2, 40, 40)


def keygen(mod, log_n, t_mont, ternary_s=False, e_bits=40):
    """pk = (A, B) with A = B s + t e in NTT form (project-defined key generation: hmpc has none).
    s is centred binomial with variance 1/2 (hmpc's u distribution) or, with ternary_s, uniform
    ternary from the capability stream at nonce {1, 0, 0}; e is centred binomial of e_bits bits."""
    N = 1 << log_n
    B = mod.sample_uniform(KEY, 0, N)
    if ternary_s:
        rng = mod.ctx.rng
        mod.ctx.set_rng(KEY, [1, 0, 0])
        s_coeff = mod.sample_capability("centered_binomial", BITS_TERNARY, "element", log_n, 1)
        mod.ctx.set_rng(*rng)
    else:
        s_coeff = mod.sample_cbd(KEY, 1, 2, N)
    s = mod.ntt(log_n, 1, s_coeff)
    e = mod.ntt(log_n, 1, mod.sample_cbd(KEY, 2, e_bits, N))
    A = mod.pointwise(mod.ADD, N, mod.pointwise(mod.MUL, N, B, s), mod.pointwise_scalar(mod.MUL, N, e, t_mont))
    return A, B, s


def bench_bgv(hm, ctx, stream, steps, warmup, n=1024, log_n=14, n_plain=32):
    mod = hm.Modulus(ctx, hm.Q128B)
    mod.prepare(log_n)
    N = 1 << log_n
    log(f"[bench] C3 BGV: {n} ciphertexts, N=2^{log_n}")
    t_mont = mod.mont(1 << 32)
    A, B, s = keygen(mod, log_n, t_mont, ternary_s=True, e_bits=20)
    msgs = ctx.sample_words(KEY, 3, n * N)
    c0, c1 = mod.alloc(n * N), mod.alloc(n * N)
    tuned = tune_bgv(mod, log_n, 128, stream)  # 128 < resident workgroups: its launches have their own grid
    log(f"[bench] autotuned BGV schedule: {tuned}")

    def timed_batch(A, B, s, bits):
        ctx.set_rng(ENC_KEY, [0, 0, 0])
        ctx.set_timing(True)
        ms, per, _ = time_steps(ctx, stream, lambda: mod.bgv_encrypt(log_n, n, A, B, msgs, t_mont, bits=bits, c0=c0, c1=c1,
                                                                     stream=stream), steps, warmup)
        kernels = ctx.kernel_times()
        ctx.set_timing(False)
        # decrypt-and-compare on the last batch (size-independent property)
        dec = mod.bgv_decrypt(log_n, n, s, c0, c1, stream=stream)
        stream.synchronize()
        return ms, per, kernels, bool(np.array_equal(dec.to_numpy(n * N), msgs.to_numpy(n * N)))

    # hmpc's randomness first, with its own key (s of variance 1/2, e of 40 bits)
    hA, hB, hs = keygen(mod, log_n, t_mont)
    hms, hper, _, hok = timed_batch(hA, hB, hs, HMPC_BITS)
    ms, per, kernels, ok = timed_batch(A, B, s, C3_BITS)
    pc0, pc1 = mod.alloc(n_plain * N), mod.alloc(n_plain * N)
    pms, _, _ = time_steps(ctx, stream, lambda: mod.bgv_encrypt(log_n, n_plain, A, B, msgs, t_mont, bits=C3_BITS, c0=pc0,
                                                               c1=pc1, plain=True, stream=stream), 2, 1)
    # the fused kernels launched once per ciphertext (hmpc_bgv_encrypt with n = 1, one call each)
    views = [(hm.View(ctx, msgs.ptr + 4 * i * N), hm.View(ctx, pc0.ptr + 16 * i * N), hm.View(ctx, pc1.ptr + 16 * i * N))
             for i in range(n_plain)]

    def one_by_one():
        for m_i, c0_i, c1_i in views:
            mod.bgv_encrypt(log_n, 1, A, B, m_i, t_mont, bits=C3_BITS, c0=c0_i, c1=c1_i, stream=stream)

    fms, _, _ = time_steps(ctx, stream, one_by_one, 3, 1)
    return {"n": n, "ms": ms, "per": per, "enc_per_s": n / (ms / 1e3), "decrypt_ok": ok, "schedule": tuned,
            "kernels": kernels, "plain_enc_per_s": n_plain / (pms / 1e3), "plain_sample": n_plain,
            "per_ciphertext_fused_enc_per_s": n_plain / (fms / 1e3),
            "hmpc_randomness": {"bits_u_v_w": list(HMPC_BITS), "s": "centred binomial, variance 1/2",
                                "e_bits": 40, "enc_per_s": n / (hms / 1e3), "ms_per_batch": hms,
                                "p50_ms": percentile(hper, 50), "p99_ms": percentile(hper, 99), "decrypt_ok": hok}}


def bench_c1(hm, ctx, stream, iters=1000):
    mod = hm.Modulus(ctx, hm.Q64)
    mod.prepare(12)
    x = mod.sample_uniform(KEY, 0, 4096)
    y, z = mod.alloc(4096), mod.alloc(4096)
    out = {}
    for name, plain in (("fused", False), ("plain", True)):
        def step():
            mod.ntt(12, 1, x, y, plain=plain, stream=stream)
            mod.ntt(12, 1, y, z, inverse=True, plain=plain, stream=stream)
        _, per, _ = time_steps(ctx, stream, step, iters, 10)
        out[name] = {"p50_us": percentile(per, 50) * 1e3, "p99_us": percentile(per, 99) * 1e3}
    ok = np.array_equal(z.to_numpy((2, 4096)), x.to_numpy((2, 4096)))
    out["roundtrip_ok"] = bool(ok)
    return out


def bench_c4(hm, ctx, stream, jobs=2000, seed=4):
    """Uneven batch: N from 2^12..2^16 (p = .3 .25 .2 .15 .1), q64 / q128a (50/50), product /
    encryption (50/50); autotuned persistent launches replayed from a hipGraph vs the same jobs
    launched one by one (hmpc's per-stage plain path, and the fused kernels per job)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([12, 13, 14, 15, 16], size=jobs, p=[0.3, 0.25, 0.2, 0.15, 0.1])
    qsel = rng.integers(0, 2, size=jobs)
    kind = rng.integers(0, 2, size=jobs)
    mods = {0: hm.Modulus(ctx, hm.Q64), 1: hm.Modulus(ctx, hm.Q128A)}
    log(f"[bench] C4 uneven batch: {jobs} jobs")
    tuned = {}
    keys = {}
    for qi, m in mods.items():
        for ln in range(12, 17):
            m.prepare(ln)
            # trial sizes near the batch's per-class job counts (about 2000 / 20 classes)
            tuned[f"{'q64' if qi == 0 else 'q128a'}/2^{ln}"] = {"polymul": tune(m, ln, 64, stream),
                                                               "bgv": tune_bgv(m, ln, 128, stream)}
            t_mont = m.mont(1 << 32)
            A, B, _ = keygen(m, ln, t_mont)
            keys[(qi, ln)] = (A, B, t_mont)
    bufs = []
    ctx.set_rng(ENC_KEY, [0, 0, 0])
    batch = hm.Batch(ctx)
    coeffs = 0
    for j in range(jobs):
        m, ln, N = mods[int(qsel[j])], int(sizes[j]), 1 << int(sizes[j])
        if kind[j] == 0:
            a = m.sample_uniform(KEY, 1000 + 2 * j, N)
            b = m.sample_uniform(KEY, 1001 + 2 * j, N)
            out = m.alloc(N)
            batch.add_polymul(m, ln, a, b, out)
            bufs.append(("mul", m, ln, (a, b, out)))
            coeffs += 3 * N
        else:
            A, B, t_mont = keys[(int(qsel[j]), ln)]
            msg = ctx.sample_words(KEY, 5000 + j, N)
            c0, c1 = m.alloc(N), m.alloc(N)
            batch.add_bgv_encrypt(m, ln, A, B, msg, t_mont, c0, c1)
            bufs.append(("bgv", m, ln, (A, B, msg, t_mont, c0, c1)))
            coeffs += 3 * N
    batch.run(graph=True, stream=stream)  # build + capture
    stream.synchronize()
    # graph replay and eager launches timed alternately (three rounds of five steps each), so that
    # neither mode is measured alone at the start of the section while the clocks settle; the
    # median round of each mode is reported
    g_rounds, e_rounds, g_per = [], [], []
    for _ in range(3):
        ms, per, _ = time_steps(ctx, stream, lambda: batch.run(graph=True, stream=stream), 5, 1)
        g_rounds.append(ms)
        g_per += per
        ms, _, _ = time_steps(ctx, stream, lambda: batch.run(graph=False, stream=stream), 5, 1)
        e_rounds.append(ms)
    g_ms, e_ms = float(np.median(g_rounds)), float(np.median(e_rounds))
    # per-launch event timing of one eager run (launch probes are not part of the captured graph)
    ctx.set_timing(True)
    batch.run(graph=False, stream=stream)
    stream.synchronize()
    kernels = ctx.kernel_times()
    ctx.set_timing(False)
    info = batch.info()

    def per_job(plain):
        for kind_j, m, ln, t in bufs:
            if kind_j == "mul":
                a, b, out = t
                if plain:
                    N = 1 << ln
                    fa, fb = out, scratch[m]
                    m.ntt(ln, 1, a, fa, plain=True, stream=stream)
                    m.ntt(ln, 1, b, fb, plain=True, stream=stream)
                    m.pointwise(m.MUL, N, fa, fb, out=fa, stream=stream)
                    m.ntt(ln, 1, fa, out, inverse=True, plain=True, stream=stream)
                else:
                    m.polymul(ln, 1, a, b, out, stream=stream)
            else:
                A, B, msg, t_mont, c0, c1 = t
                m.bgv_encrypt(ln, 1, A, B, msg, t_mont, c0=c0, c1=c1, plain=plain, stream=stream)

    scratch = {qi: m.alloc(1 << 16) for qi, m in mods.items()}
    scratch = {m: scratch[qi] for qi, m in mods.items()}
    f_ms, _, _ = time_steps(ctx, stream, lambda: per_job(False), 2, 1)
    p_ms, _, _ = time_steps(ctx, stream, lambda: per_job(True), 1, 0)
    return {"jobs": jobs, "graph_ms": g_ms, "graph_p50_ms": percentile(g_per, 50), "eager_ms": e_ms,
            "graph_rounds_ms": g_rounds, "eager_rounds_ms": e_rounds,
            "per_job_fused_ms": f_ms, "plain_per_job_ms": p_ms, "coeffs": coeffs,
            "gcoeff_per_s": coeffs / (g_ms / 1e3) / 1e9, "batched_vs_plain": p_ms / g_ms,
            "batched_vs_per_job_fused": f_ms / g_ms, "classes": info["classes"],
            "eager_kernels_ms": {k: v[0] * v[1] for k, v in kernels.items()},
            "device_bytes": info["device_bytes"], "schedules": tuned}


def bench_c5(hm, ctx, stream, requests=2000, seed=5, fractions=(0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0),
             operating_requests=100000):
    """Serving: 16 resident BGV keys (q128a, N = 2^13..2^16, four keys per N); each request
    encrypts k ~ U[1, 64] messages under a uniformly chosen key. Peak = closed-loop throughput with
    every request queued at once; then open-loop Poisson arrivals at each fraction of the peak
    rate (`requests` per point: the sweep is reduced below the configured 10^5 unless asked), and
    the configured 10^5 requests at the operating point (the highest load served at >= 80% of the
    peak). Latency runs from the scheduled arrival to completion (client lateness included)."""
    mod = hm.Modulus(ctx, hm.Q128A)
    t_mont = mod.mont(1 << 32)
    log(f"[bench] C5 serving: {requests} requests per load point")
    keys = []
    for ln in (13, 14, 15, 16):
        mod.prepare(ln)
        tune_bgv(mod, ln, 64, stream)
    srv = hm.Server(ctx, budget_bytes=200 << 30, slots=3)
    pk = {}
    for i in range(16):
        ln = 13 + i % 4
        N = 1 << ln
        B = mod.sample_uniform(KEY, 100 + 3 * i, N)
        s = mod.ntt(ln, 1, mod.sample_cbd(KEY, 101 + 3 * i, 2, N))
        e = mod.ntt(ln, 1, mod.sample_cbd(KEY, 102 + 3 * i, 40, N))
        A = mod.pointwise(mod.ADD, N, mod.pointwise(mod.MUL, N, B, s), mod.pointwise_scalar(mod.MUL, N, e, t_mont))
        pk[i] = (A, B, ln)
        keys.append(srv.add_key(mod, ln, A, B, t_mont))
    pool = ctx.sample_words(KEY, 7, 64 << 16)
    rng = np.random.default_rng(seed)

    def draw(n):
        return [(int(rng.integers(0, 16)), int(rng.integers(1, 65))) for _ in range(n)]

    tag = [0]

    def run(reqs, arrivals):
        tag[0] += 1
        lateness, lat, pending = {}, {}, set()
        t0 = time.perf_counter()
        i = 0
        while i < len(reqs) or pending:
            now = time.perf_counter() - t0
            while i < len(reqs) and arrivals[i] <= now:
                key, k = reqs[i]
                rid = srv.submit_encrypt(keys[key], k, pool.ptr, [0, i, tag[0]])  # one nonce per request
                lateness[rid] = now - arrivals[i]
                pending.add(rid)
                i += 1
                now = time.perf_counter() - t0
            got = srv.completed()
            for rid, us in got:
                lat[rid] = us / 1e3 + lateness[rid] * 1e3
                srv.release(rid)
                pending.discard(rid)
            if not got and (i >= len(reqs) or arrivals[i] > now + 50e-6):
                time.sleep(20e-6)
        wall = time.perf_counter() - t0
        return wall, np.array(list(lat.values()))

    # warm-up (pool growth, code objects)
    reqs = draw(200)
    run(reqs, [0.0] * len(reqs))
    reqs = draw(requests)
    cts = sum(k for _, k in reqs)
    wall, _ = run(reqs, [0.0] * len(reqs))
    peak_rps = len(reqs) / wall
    peak_cps = cts / wall
    points = []
    for f in fractions:
        rate = f * peak_rps
        reqs = draw(requests)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
        wall, lat = run(reqs, arrivals)
        done_rate = len(reqs) / max(wall, arrivals[-1])
        points.append({"fraction": f, "offered_req_per_s": rate, "achieved_req_per_s": done_rate,
                       "ct_per_s": sum(k for _, k in reqs) / wall, "p50_ms": percentile(lat, 50),
                       "p99_ms": percentile(lat, 99)})
    ok = [p for p in points if p["achieved_req_per_s"] >= 0.8 * peak_rps]
    op = max(ok, key=lambda p: p["fraction"]) if ok else points[-1]
    # the configured request count (10^5) at the operating point's load
    full = None
    if operating_requests:
        rate = op["fraction"] * peak_rps
        reqs = draw(operating_requests)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
        log(f"[bench] C5 operating point: {operating_requests} requests at {rate:.0f} req/s")
        wall, lat = run(reqs, arrivals)
        full = {"requests": operating_requests, "fraction": op["fraction"], "offered_req_per_s": rate,
                "achieved_req_per_s": len(reqs) / max(wall, arrivals[-1]),
                "ct_per_s": sum(k for _, k in reqs) / wall, "p50_ms": percentile(lat, 50),
                "p99_ms": percentile(lat, 99)}
    st = srv.stats()
    srv.close()
    # plain path: the same kind of requests, one per call through hmpc's launch structure
    sample = draw(8)
    c0, c1 = mod.alloc(64 << 16), mod.alloc(64 << 16)
    stream.synchronize()
    t0 = time.perf_counter()
    n_ct = 0
    for key, k in sample:
        A, B, ln = pk[key]
        mod.bgv_encrypt(ln, k, A, B, pool, t_mont, c0=c0, c1=c1, plain=True, stream=stream)
        n_ct += k
    stream.synchronize()
    plain_ct = n_ct / (time.perf_counter() - t0)
    return {"requests_per_point": requests, "reduced_sweep": requests < 100000, "configured_requests_per_point": 100000,
            "peak_req_per_s": peak_rps, "peak_ct_per_s": peak_cps,
            "operating_point": op, "operating_point_full": full, "points": points, "plain_ct_per_s": plain_ct,
            "plain_sample_requests": len(sample),
            "launches": st["launches"], "budget_bytes": 200 << 30, "slots": 3}


# hmpc's examples/ntt.cpp:22-23 (p, N = 2^20, n = 10) and examples/bgv.cpp:25-29 (q, N = 2^14, p, n = 100), and
# the size of tests/expr/crypto/lhe/enc.cpp (q of 279 bits, p = 0x8822d80623320001, N = 2^16, shape {4, 2})
P_NTT_EXAMPLE = 0x18F25CD9A75CCBD9C146D4ABAEC00001
Q_BGV_EXAMPLE = 0x2FAEADBE7A0195C011AC195AD10269830E8001
P_BGV_EXAMPLE = 0x89D046E0AE640001
Q_ENC_TEST = 0x591F5B834C0D961F67343BCC8902BDEDA2771F54306FF151162FF8D2B40F4194DC0001
P_ENC_TEST = 0x8822D80623320001


def bench_examples(hm, ctx, stream, reps=5):
    """hmpc's own example workloads on their moduli, with argc = 1:
    - examples/ntt.cpp: forward NTT of n = 10 polynomials of N = 2^20 over the 125-bit p, inputs
      argc + N n - i; timed: the transform; checked by the inverse transform;
    - examples/bgv.cpp: encryption of n = 100 messages at N = 2^14 over the 150-bit q with keys
      NTT(argc + n + i), NTT(argc + N - i), messages argc + n - i and the constant p; timed: the
      samplers, transforms and combination, which here also transform the messages (hmpc transforms
      them before its clock starts);
    - tests/expr/crypto/lhe/enc.cpp at its size (q of 279 bits, p = 0x8822d80623320001, N = 2^16,
      8 messages) through the plaintext path: encryption and dec<plaintext>; checked: dec(enc(x)) == x."""
    out = {}
    # examples/ntt.cpp
    p = hm.Modulus(ctx, P_NTT_EXAMPLE)
    log_n, n = 20, 10
    N = 1 << log_n
    p.prepare(log_n)
    words = ((1 + N * n - np.arange(N * n, dtype=np.int64)) & 0xFFFFFFFF).astype(np.uint32)
    w = ctx.upload(words)
    x = p.from_u32(N * n, w)
    y, z = p.alloc(N * n), p.alloc(N * n)
    ms, _, _ = time_steps(ctx, stream, lambda: p.ntt(log_n, n, x, out=y, stream=stream), reps, 1)
    p.ntt(log_n, n, y, out=z, inverse=True, stream=stream)
    stream.synchronize()
    ok = bool(np.array_equal(p.download(z, N * n), p.download(x, N * n)))
    out["ntt_example"] = {"source": "examples/ntt.cpp:22-23", "q_bits": P_NTT_EXAMPLE.bit_length(), "limbs": p.limbs,
                          "N": N, "n": n, "ms": ms, "gcoeff_per_s": n * N / (ms / 1e3) / 1e9,
                          "fast_transform_kernels": p.schedule(log_n)["fast"], "roundtrip_ok": ok}
    del x, y, z, w
    # examples/bgv.cpp
    q = hm.Modulus(ctx, Q_BGV_EXAMPLE)
    log_n, n = 14, 100
    N = 1 << log_n
    q.prepare(log_n)
    argc = 1
    a = q.from_u32(N, ctx.upload(((argc + n + np.arange(N)) & 0xFFFFFFFF).astype(np.uint32)))
    b = q.from_u32(N, ctx.upload(((argc + N - np.arange(N)) & 0xFFFFFFFF).astype(np.uint32)))
    ka = q.ntt(log_n, 1, a)
    kb = q.ntt(log_n, 1, b)
    msgs = ctx.upload(((argc + n - np.arange(N * n, dtype=np.int64)) & 0xFFFFFFFF).astype(np.uint32))
    p_mont = q.mont(P_BGV_EXAMPLE)
    c0, c1 = q.alloc(n * N), q.alloc(n * N)
    ctx.set_rng(ENC_KEY, [0, 0, 0])
    ms, _, _ = time_steps(ctx, stream, lambda: q.bgv_encrypt(log_n, n, kb, ka, msgs, p_mont, c0=c0, c1=c1,
                                                             stream=stream), reps, 1)
    out["bgv_example"] = {"source": "examples/bgv.cpp:25-29", "q_bits": Q_BGV_EXAMPLE.bit_length(), "limbs": q.limbs,
                          "N": N, "n": n, "ms": ms, "enc_per_s": n / (ms / 1e3),
                          "fast_transform_kernels": q.schedule(log_n)["fast"]}
    del c0, c1, msgs
    # tests/expr/crypto/lhe/enc.cpp at its size, through the plaintext path
    qq, pp = hm.Modulus(ctx, Q_ENC_TEST), hm.Modulus(ctx, P_ENC_TEST)
    log_n, n = 16, 8
    N = 1 << log_n
    qq.prepare(log_n)
    pp.prepare(log_n)
    iota_q = np.zeros((qq.limbs, N), np.uint32)
    iota_q[0] = np.arange(N, dtype=np.uint32)
    ka = qq.upload(iota_q)
    one = np.zeros((qq.limbs, N), np.uint32)
    for l_, wv in enumerate(qq.mont(1)):
        one[l_] = wv
    s = qq.upload(one)
    xi = np.zeros((pp.limbs, n * N), np.uint32)
    xi[0] = np.arange(n * N, dtype=np.uint32)
    xd = pp.upload(xi)
    c0, c1 = qq.alloc(n * N), qq.alloc(n * N)
    ctx.set_rng(ENC_KEY, [0, 0, 0])
    ms, _, _ = time_steps(ctx, stream, lambda: qq.bgv_encrypt_plaintext(pp, log_n, n, ka, ka, xd, True, c0=c0, c1=c1,
                                                                        stream=stream), reps, 1)
    dec = pp.alloc(n * N)
    dms, _, _ = time_steps(ctx, stream, lambda: qq.bgv_decrypt_plaintext(pp, log_n, n, s, c0, c1, True, out=dec,
                                                                         stream=stream), reps, 1)
    ok = bool(np.array_equal(pp.download(dec, n * N), xi))
    out["lhe_enc_test"] = {"source": "tests/expr/crypto/lhe/enc.cpp:15-72", "q_bits": Q_ENC_TEST.bit_length(),
                           "q_limbs": qq.limbs, "p_limbs": pp.limbs, "N": N, "n": n, "enc_ms": ms,
                           "enc_per_s": n / (ms / 1e3), "dec_ms": dms, "dec_ok": ok}
    return out


def bench_copy(hm, ctx, stream, nbytes=1 << 30, reps=5):
    """Device-to-device copy of a known size: the achievable-bandwidth reference for this box and
    the calibration of the HBM byte counters (FETCH_SIZE / WRITE_SIZE of this copy are compared
    with its 1 GiB read + 1 GiB write)."""
    src, dst = ctx.buffer(nbytes), ctx.buffer(nbytes)
    lib = hm.load()

    def step():
        hm.runtime.check(lib.hmpc_memcpy_d2d(ctx.handle, dst.ptr, src.ptr, nbytes, stream.handle))

    ms, _, _ = time_steps(ctx, stream, step, reps, 1)
    return {"bytes": nbytes, "ms": ms, "gbs": 2 * nbytes / (ms / 1e3) / 1e9}


def bench_ext(hm, ctx, stream, reps=5):
    """The element-wise operations around the path, each against the HBM roofline with its
    algorithmic bytes (q128a, L = 4, 16 B per element): tensor cipher, ciphertext sum and
    plaintext product over 1024 x 2^14 ciphertexts, fused multiply-add, sum reduction and a
    4096 x 4096 matrix-vector product; and BGV through pinned host buffers (messages up, c0 and c1
    down: the PCIe-inclusive rate)."""
    mod = hm.Modulus(ctx, hm.Q128A)
    L = mod.limbs
    out = {}

    def rate(name, fn, nbytes, extra=None):
        ms, _, _ = time_steps(ctx, stream, fn, reps, 1)
        gbs = nbytes / (ms / 1e3) / 1e9
        out[name] = dict({"ms": ms, "bytes": nbytes, "gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}, **(extra or {}))

    log("[bench] element-wise operations (q128a)")
    count = 1 << 24
    x = mod.sample_uniform(KEY, 40, count)
    y = ctx.buffer(4 * L * count)
    rate("cipher_encrypt", lambda: ctx.cipher_encrypt(L, count, KEY, [1, 0, 0], x, out=y, stream=stream), 2 * 4 * L * count,
         {"elements": count, "chacha_blocks": count * L // 16})
    n_ct = 1024 << 14
    a0, a1 = mod.sample_uniform(KEY, 41, n_ct), mod.sample_uniform(KEY, 42, n_ct)
    b0, b1 = mod.sample_uniform(KEY, 43, n_ct), mod.sample_uniform(KEY, 44, n_ct)
    o0, o1 = mod.alloc(n_ct), mod.alloc(n_ct)
    rate("ciphertext_add", lambda: mod.ciphertext_op(mod.ADD, n_ct, a0, a1, b0, b1, out0=o0, out1=o1, stream=stream),
         3 * 2 * 16 * n_ct, {"ciphertexts": 1024, "N": 1 << 14})
    p = mod.sample_uniform(KEY, 45, 1 << 14)
    rate("ciphertext_mul_plain", lambda: mod.ciphertext_op(mod.MUL, n_ct, a0, a1, p, p, b_count=1 << 14, out0=o0, out1=o1,
                                                           stream=stream), 2 * 2 * 16 * n_ct, {"ciphertexts": 1024})
    del b0, b1
    c = mod.sample_uniform(KEY, 46, count)
    d = mod.sample_uniform(KEY, 47, count)
    f = mod.alloc(count)
    rate("fma", lambda: mod.fma(count, x, c, d, out=f, stream=stream), 4 * 16 * count, {"elements": count})
    r = mod.alloc(1)
    rate("sum", lambda: mod.reduce(mod.ADD, count, x, out=r, stream=stream), 16 * count, {"elements": count})
    rows = cols = 4096
    M = mod.sample_uniform(KEY, 48, rows * cols)
    v = mod.sample_uniform(KEY, 49, cols)
    mv = mod.alloc(rows)
    rate("matvec", lambda: mod.matvec(1, rows, cols, M, v, out=mv, stream=stream), 16 * rows * cols,
         {"rows": rows, "cols": cols})
    del M, a0, a1, o0, o1, c, d, f, x, y
    # BGV with host-resident messages and ciphertexts (pinned staging, async copies on the stream)
    bmod = hm.Modulus(ctx, hm.Q128B)
    log_n, n = 14, 1024
    N = 1 << log_n
    t_mont = bmod.mont(1 << 32)
    A, B, _ = keygen(bmod, log_n, t_mont)
    msgs_h = ctx.host_buffer(4 * n * N)
    msgs_h.array()[:] = ctx.sample_words(KEY, 3, n * N).to_numpy(n * N)
    c0_h, c1_h = ctx.host_buffer(16 * n * N), ctx.host_buffer(16 * n * N)
    msgs_d, c0, c1 = ctx.buffer(4 * n * N), bmod.alloc(n * N), bmod.alloc(n * N)
    tune_bgv(bmod, log_n, 128, stream)

    def host_step():
        msgs_h.to_device(msgs_d, stream=stream)
        bmod.bgv_encrypt(log_n, n, A, B, msgs_d, t_mont, c0=c0, c1=c1, stream=stream)
        c0_h.from_device(c0, stream=stream)
        c1_h.from_device(c1, stream=stream)

    ms, _, _ = time_steps(ctx, stream, host_step, 3, 1)
    out["bgv_host_buffers"] = {"ms": ms, "enc_per_s": n / (ms / 1e3), "pcie_bytes": 4 * n * N + 2 * 16 * n * N,
                               "pcie_gbs": (4 * n * N + 2 * 16 * n * N) / (ms / 1e3) / 1e9}
    return out


def cpu_baseline(pm, seconds_target=10.0):
    """Oracle restatement of hmpc's polymul on the host cores over a bounded sample of
    the same inputs (the first pairs of the batch)."""
    import oracle

    # the GPU's share of the host: 16 threads per GPU, whatever the machine holds in total
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    om = oracle.Modulus(pm["mod"].q)
    log_n = pm["log_n"]
    N = 1 << log_n
    n_total = pm["n"]
    count = n_total * N
    a_all = pm["a"].to_numpy((om.limbs, count))
    b_all = pm["b"].to_numpy((om.limbs, count))
    pairs = 2
    _, secs = om.polymul_timed(log_n, np.ascontiguousarray(a_all[:, : pairs * N]),
                               np.ascontiguousarray(b_all[:, : pairs * N]), threads)
    pairs = int(min(n_total, max(2, pairs * seconds_target / max(secs, 1e-3))))
    a = np.ascontiguousarray(a_all[:, : pairs * N])
    b = np.ascontiguousarray(b_all[:, : pairs * N])
    got, secs = om.polymul_timed(log_n, a, b, threads)
    want = pm["out"].to_numpy((om.limbs, count))[:, : pairs * N]
    return {"value": 3 * pairs * N / secs / 1e9, "unit": "Gcoeff/s", "cores": threads, "kind": "port",
            "sample": f"{pairs} of {n_total} pairs (N=2^{log_n}, q128a), {secs:.2f} s wall",
            "matches_gpu": bool(np.array_equal(got, want)), "cpu_model": cpu_model(),
            "host_logical_cpus": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0))}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for row in f:
                if row.startswith("model name"):
                    return row.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def roofline(kname, kavg, launches, kbytes, achieved, traffic, valu, step_gbs):
    """The bound of the dominant kernel. Its algorithmic HBM rate (bytes per launch / live average
    duration) against 8 TB/s is always reported; when the committed counter summary holds the
    kernel's VALU instruction count and that count needs a larger share of the duration than the
    HBM bytes do, the kernel is VALU-issue bound and the line says so: the achieved rate is then
    wave instructions per second against the peak of one wave instruction per SIMD every 4 cycles
    at the 2.4 GHz peak engine clock (`frac`). Beside it: the engine clock measured under the
    counters (in the counter run, on its box, not this one) and the time the instructions take at
    that clock, to compare with the kernel's duration: the integer load holds the clock near
    2 GHz."""
    hbm = {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS}
    common = {"traffic": traffic, "kernel": kname, "kernel_avg_ms": kavg, "launches": launches,
              "algorithmic_bytes_per_launch": kbytes, "step_hbm_gbs": step_gbs, "valu_issue": valu}
    if valu is not None and valu.get("frac_at_measured_clock", valu["frac"]) > hbm["frac"]:
        rate = valu["wave_instructions"] / (kavg / 1e3) / 1e9
        peak = valu["issue_rate_per_s"] / 1e9
        out = {"bound": "valu", "achieved": rate, "peak": peak, "unit": "Gwave-instructions/s", "frac": rate / peak,
               "clock_ghz": VALU_CLOCK_GHZ, "hbm": hbm, **common}
        if valu.get("clock_ghz_measured"):
            out["counter_clock_ghz"] = valu["clock_ghz_measured"]
            out["issue_ms_at_counter_clock"] = valu["issue_ms_at_measured_clock"]
        return out
    return {"bound": "hbm", **hbm, **common}


def traffic_for(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by a copy kernel of known
    size), or None when no summary covers this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(workload, {}).get(kernel)
    return None if e is None else e.get("bytes_per_launch")


def valu_issue_for(kernel, workload, kernel_ms, cus):
    """VALU issue bound of `kernel` from the committed counter summary (profiles/valu.json:
    SQ_INSTS_VALU wave instructions per dispatch): the time those instructions need at one wave
    instruction per SIMD every 4 cycles at 2.4 GHz, and its fraction of the measured duration;
    where the summary holds the engine clock the kernel ran at under the counters
    (GRBM_GUI_ACTIVE / duration), the same at that clock."""
    path = os.path.join(ROOT, "profiles", "valu.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get(workload, {}).get(kernel)
    if e is None or "SQ_INSTS_VALU" not in e:
        return None
    rate = cus * 4 * VALU_CLOCK_GHZ * 1e9 / 4
    issue_ms = e["SQ_INSTS_VALU"] / rate * 1e3
    out = {"wave_instructions": e["SQ_INSTS_VALU"], "issue_rate_per_s": rate, "issue_ms": issue_ms,
           "frac": issue_ms / kernel_ms}
    if e.get("clock_ghz"):
        ms = e["SQ_INSTS_VALU"] / (cus * e["clock_ghz"] * 1e9) * 1e3
        out.update({"clock_ghz_measured": e["clock_ghz"], "issue_ms_at_measured_clock": ms,
                    "frac_at_measured_clock": ms / kernel_ms})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--only", default="",
                    help="comma list of sections to run: c1,c2,c3,c4,c5,ext,examples,copy (default all)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default schedules (for profiling runs)")
    ap.add_argument("--serving-requests", type=int, default=2000,
                    help="requests per load point of the serving sweep (the full workload uses 100000)")
    ap.add_argument("--serving-operating-requests", type=int, default=100000,
                    help="requests of the run at the serving operating point (0: none)")
    args = ap.parse_args()
    AUTOTUNE[0] = not args.no_autotune
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the workloads shard by independent batch elements); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import hmpc_amd as hm

    sections = set(args.only.split(",")) if args.only else {"c1", "c2", "c3", "c4", "c5", "ext", "examples", "copy"}
    ctx = hm.Context(0, key=KEY)
    stream = ctx.stream()
    info = ctx.info()
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['compute_units']}")

    line = {"metric": METRIC, "unit": "Gcoeff/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic (explicit-key ChaCha20, key 42)"}
    pm = None
    if "c2" in sections:
        pm = bench_polymul(hm, ctx, stream, args.steps, args.warmup)
        n, log_n = pm["n"], pm["log_n"]
        N = 1 << log_n
        L = pm["mod"].limbs
        coeffs = 3 * n * N
        value = coeffs / (pm["ms"] / 1e3) / 1e9
        unit_bytes = N * L * 4 * n  # one pass over the batch
        modmuls = n * (3 * (N // 2) * log_n + N + N // 2)
        # dominant kernel: largest total time among the launches recorded in the timed region
        kname, (klaunch, kavg) = max(pm["kernels"].items(), key=lambda kv: kv[1][0] * kv[1][1])
        kbytes = KERNEL_PASSES.get(kname, 3) * unit_bytes
        achieved = kbytes / (kavg / 1e3) / 1e9
        line.update({
            "value": value, "ms_per_step": pm["ms"],
            "p50_ms": percentile(pm["per"], 50), "p99_ms": percentile(pm["per"], 99),
            "config": {"workload": "batched polymul INTT(NTT(a) o NTT(b)): 512 pairs, N=2^16, "
                                   "q=2^124+2^20+2^18+1 (4x32-bit limbs)", "pairs": n, "N": N, "limbs": L,
                       "schedule": pm["schedule"]},
            "plain_path": {"ms_per_step": pm["plain_ms"], "value": coeffs / (pm["plain_ms"] / 1e3) / 1e9,
                           "speedup_fused": pm["plain_ms"] / pm["ms"]},
            "roofline": roofline(kname, kavg, klaunch, kbytes, achieved, traffic_for(kname, "c2"),
                                 valu_issue_for(kname, "c2", kavg, info["compute_units"]),
                                 3 * unit_bytes / (pm["ms"] / 1e3) / 1e9),
            "kernels_ms": {k: v[1] for k, v in pm["kernels"].items()},
            "int_rate": {"modmul_per_s": modmuls / (pm["ms"] / 1e3), "modmuls_per_step": modmuls},
        })
    if "c3" in sections:
        bg = bench_bgv(hm, ctx, stream, max(3, args.steps // 2), args.warmup)
        line["bgv"] = {"workload": "BGV pk encryption, 1024 ciphertexts, N=2^14, q=0x1...78001 (4 limbs), t=2^32",
                       "randomness": {"s": "uniform ternary", "u": "uniform ternary",
                                      "e_e0_e1": "centred binomial eta=10 (20 bits)", "bits_u_v_w": list(C3_BITS)},
                       "enc_per_s": bg["enc_per_s"], "ms_per_batch": bg["ms"],
                       "p50_ms": percentile(bg["per"], 50), "p99_ms": percentile(bg["per"], 99),
                       "plain_enc_per_s": bg["plain_enc_per_s"], "plain_sample": bg["plain_sample"],
                       "batched_vs_plain": bg["enc_per_s"] / bg["plain_enc_per_s"],
                       "per_ciphertext_fused_enc_per_s": bg["per_ciphertext_fused_enc_per_s"],
                       "batched_vs_per_ciphertext_fused": bg["enc_per_s"] / bg["per_ciphertext_fused_enc_per_s"],
                       "decrypt_ok": bg["decrypt_ok"], "hmpc_randomness": bg["hmpc_randomness"],
                       "schedule": bg["schedule"], "kernels_ms": {k: v[1] for k, v in bg["kernels"].items()}}
    if "c4" in sections:
        line["uneven_batch"] = bench_c4(hm, ctx, stream)
    if "c5" in sections:
        line["serving"] = bench_c5(hm, ctx, stream, requests=args.serving_requests,
                                   operating_requests=args.serving_operating_requests)
    if "c1" in sections:
        line["c1_latency"] = bench_c1(hm, ctx, stream)
    if "ext" in sections:
        line["elementwise"] = bench_ext(hm, ctx, stream)
    if "examples" in sections:
        line["hmpc_examples"] = bench_examples(hm, ctx, stream)
    if "copy" in sections:
        line["copy_reference"] = bench_copy(hm, ctx, stream)
    if pm is not None and not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle, host cores)")
        line["cpu_baseline"] = cpu_baseline(pm, args.cpu_seconds)
    if "value" not in line:
        line.update({"value": None, "ms_per_step": None, "config": {"workload": ",".join(sorted(sections))}})
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
