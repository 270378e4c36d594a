"""Benchmark: HomMul+relin+rescale ops/s at N=2^16 on one MI355X, batched vs
plain, with p50/p99 latency.

Headline workload: N = 2^16, k = 30 Q limbs (63-bit q0 + 29 x 48-bit, prec
48), K = 31 special limbs (55-bit), from the top level 29, uniform random
residues resident in HBM (synthetic ciphertexts and evaluation key of that
shape).  A step of the batched form is one hipGraph replay of B
HomMul+relin+rescale; a step of the plain form is one aoa_hommul_relin_rescale
call.  `value` is the batched throughput.  Reported beside it: the plain
throughput and latency, the ratio, the per-stage time split, the roofline of
the dominant kernel (ModDown), the forward-NTT bandwidth, mixed levels at
N=2^16, 256 mixed-level ciphertexts at N=2^15, rotate-and-sum over 2^15
slots for 64 ciphertexts at N=2^16, the serving loop, and the CPU oracle.

Prints one JSON line.  --profile runs only the headline shape (plain +
batched + stage timing) and a 1 GiB device copy used to calibrate the
FETCH_SIZE / WRITE_SIZE counters, so that a profiler's per-kernel averages
cover exactly the launches the roofline figure is computed from.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, K_Q, K_B, PREC = 65536, 30, 31, 48
HBM_PEAK_GBS = 8000.0
# HBM bytes per k_moddown_d launch at 32 pairs, from two rocprofv3 --pmc runs of
# `bench.py --profile` (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0; the factors make a
# 1 GiB device copy read 1 GiB and write 1 GiB); the run is summarised in
# profiles/moddown_pmc_traffic_n16_k30.md.
MODDOWN_TRAFFIC_BYTES = {32: 3025902773}
# VALU wave-instructions per k_moddown_d launch at 32 pairs (SQ_INSTS_VALU summed
# over the counter instances of one dispatch, rocprofv3 --pmc of
# `bench.py --profile`; profiles/hommul_batched_n16_k30_valu.md), and a
# reference issue rate: MI355X's vector FP64 FMA peak (78.6 TFLOP/s = 2 flops x
# 64 lanes x 256 CUs x 2.4 GHz) counted in wave instructions, one per CU per
# cycle.  64-bit integer multiply-adds issue at about this rate; 32-bit adds
# and logic issue faster (about 1.6 times), so the fraction is a proxy that a
# mix of both can exceed.
MODDOWN_VALU_INSTS = {32: 9.186e8}
VALU_PEAK_INSTS_S = 256 * 2.4e9


def op_bytes(N, L, K):
    """Algorithmic HBM bytes of one HomMul+relin+rescale at L active limbs:
    two input ciphertexts (4L limbs), the evaluation key on the active Q limbs
    and all special limbs (2(L+K)), the rescaled output (2(L-1))."""
    return 8 * N * (8 * L + 2 * K - 2)


def rot_bytes(N, L, K):
    """One rotation (key switch) at L limbs: input ct (2L), rotation key on
    the active Q limbs and the special limbs (2(L+K)), output ct (2L)."""
    return 8 * N * (6 * L + 2 * K)


def moddown_bytes(N, L, K, rescale, addend=False):
    """One ModDown of a QB polynomial: reads L + K input limbs (and an L-limb
    addend if it adds one), writes L (or L-1 after the fused rescale) limbs.
    HomMul with two-pass transforms folds its additions into the key
    products, so its ModDown reads no addend."""
    return 8 * N * (L + K + (L if addend else 0) + (L - 1 if rescale else L))


def ntt_bytes(N, R):
    """One forward or inverse NTT of R limbs reads and writes each residue once."""
    return 16 * N * R


def rand_polys(P, count, nlimbs):
    """count polynomials of nlimbs limbs, uniform residues, generated on the device."""
    import torch
    t = torch.empty((count, nlimbs * P.N), dtype=torch.int64, device="cuda")
    for r in range(nlimbs):
        t[:, r * P.N:(r + 1) * P.N] = torch.randint(0, P.primes[r], (count, P.N), dtype=torch.int64, device="cuda")
    return [t[i] for i in range(count)]


def make_items(P, levels):
    c = rand_polys(P, 4 * len(levels), P.k)
    outs = [P.zeros(P.k) for _ in range(2 * len(levels))]
    return [((c[4 * i], c[4 * i + 1]), (c[4 * i + 2], c[4 * i + 3]), (outs[2 * i], outs[2 * i + 1]), lv)
            for i, lv in enumerate(levels)]


def time_steps(fn, steps, warmup, stream):
    """Warmup untimed calls, then `steps` calls bracketed by a device
    synchronisation on both sides; returns (ms per call from HIP events on
    `stream`, wall seconds)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0


def latencies(fn, n, stream):
    import torch
    lat = []
    for _ in range(n):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))
    return np.array(lat)


def plain_fn(P, evk, items, ws_cache):
    def run():
        for c1, c2, out, lv in items:
            if lv not in ws_cache:
                ws_cache[lv] = P.workspace(lv)
            P.hommul(out, c1, c2, evk, lv, ws_cache[lv])
    return run


def valu_issue(insts, launch_ms):
    """The kernel's VALU issue rate against the chip's: wave-instructions per
    launch (from the counter profile) over the live launch time."""
    if not insts:
        return None
    rate = insts / (launch_ms * 1e-3)
    return {"insts_per_launch": insts, "achieved_insts_s": round(rate / 1e9, 1), "peak_insts_s": VALU_PEAK_INSTS_S / 1e9,
            "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_INSTS_S, 4)}


def batch_options(bt, key, args, stream):
    """Launch options of a batch: tuned now (--tune, recorded in the table),
    else the autotuner table's entry for the workload, else the defaults.
    Returns (options, where they came from)."""
    import torch
    from aoa_amd import tuning
    if args.tune:
        props = torch.cuda.get_device_properties(0)
        dev = "%s (%s)" % (props.name, getattr(props, "gcnArchName", ""))
        _, cands = tuning.tune(bt, key, stream=stream, reps=5, device=dev)
        src = "autotuned in this run (%d candidates)" % len(cands)
    elif tuning.apply(bt, key) is not None:
        src = "autotuner table (aoa_amd/tuned/batch_options.json)"
    else:
        src = "defaults (no table entry)"
    return bt.options(), src


def stage_split(bt, stream, reps):
    """Per-stage milliseconds per batch launch and launches per batch, from
    `reps` eager runs timed with HIP events around each stage launch."""
    tot = {}
    for _ in range(reps):
        for s, (ms, n) in bt.stage_times(stream).items():
            a = tot.setdefault(s, [0.0, 0])
            a[0] += ms
            a[1] += n
    return {s: (v[0] / reps, v[1] // reps) for s, v in tot.items()}


def mixed_n15(args, rng, stream):
    """256 ciphertext pairs at N=2^15, k=20 (prec 48), K=21, levels uniform in
    1..19 (2..20 active limbs): plain calls vs one batched graph replay."""
    import torch
    from aoa_amd import Batch, Params, tuning
    P3 = Params(32768, 20, 0, PREC)
    evk3 = rand_polys(P3, 2, P3.k + P3.kl)
    lv3 = [int(x) for x in rng.integers(1, 20, size=256)]
    it3 = make_items(P3, lv3)
    ws3 = {}
    p3_ms, _ = time_steps(plain_fn(P3, evk3, it3, ws3), 3, 1, stream)
    b3 = Batch()
    b3.add_mul(P3, evk3, it3)
    b3.build()
    opts, src = batch_options(b3, tuning.key(32768, P3.k + P3.kl, 256, "mul"), args, stream)
    b3_ms, _ = time_steps(lambda: b3.launch(graph=True), 5, 2, stream)
    mean_bytes = float(np.mean([op_bytes(32768, lv + 1, P3.kl) for lv in lv3]))
    del b3, it3
    torch.cuda.empty_cache()
    return {"workload": "256 HomMul+relin+rescale, N=2^15, k=20, K=21, levels uniform in 1..19",
            "plain_ops_s": round(256 * 1000.0 / p3_ms, 2), "batched_ops_s": round(256 * 1000.0 / b3_ms, 2),
            "batched_over_plain": round(p3_ms / b3_ms, 3),
            "batched_hbm_frac": round(mean_bytes * 256 / (b3_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "batch_options": opts, "batch_options_source": src}


def sumslots_n16(args, stream):
    """Rotate-and-sum over 2^15 slots (15 left rotations by 2^0..2^14, each
    followed by an add, as ckks_sumslots) for 64 ciphertexts at N=2^16, k=24
    (prec 48), K=25, top level: 15 layers of 64 fused rotate+add items in one
    graph, against plain aoa_rotate_left + add calls on 4 ciphertexts."""
    import torch
    from aoa_amd import ADD, Batch, Params, tuning
    P4 = Params(65536, 24, 0, PREC)
    lv = P4.k - 1
    L = lv + 1
    nct = 64
    steps = [1 << i for i in range(15)]
    rk = rand_polys(P4, 2 * len(steps), P4.k + P4.kl)
    rtk = {s: (rk[2 * i], rk[2 * i + 1]) for i, s in enumerate(steps)}
    cts = rand_polys(P4, 2 * nct, P4.k)
    acc = [P4.zeros(P4.k) for _ in range(2 * nct)]
    bt = Batch()
    for si, s in enumerate(steps):
        items = []
        for c in range(nct):
            src = (cts[2 * c], cts[2 * c + 1]) if si == 0 else (acc[2 * c], acc[2 * c + 1])
            items.append((src, (acc[2 * c], acc[2 * c + 1]), rtk[s], s, lv, True))
        bt.add_rotate(P4, items)
        bt.barrier()
    bt.build()
    # rotate-and-sum accumulates in place: its outputs change from replay to
    # replay, so the tuner times the candidates without comparing outputs
    # (the options' exactness is covered by the reproducible batches)
    opts, src = batch_options(bt, tuning.key(65536, P4.k + P4.kl, nct, "rotsum15"), args, stream)
    b_ms, _ = time_steps(lambda: bt.launch(graph=True), 3, 1, stream)
    del bt
    torch.cuda.empty_cache()
    ws = P4.workspace(lv)
    tmp = [P4.zeros(P4.k), P4.zeros(P4.k)]
    nplain = 4

    def plain():
        for c in range(nplain):
            a = (acc[2 * c], acc[2 * c + 1])
            a[0].copy_(cts[2 * c])
            a[1].copy_(cts[2 * c + 1])
            for s in steps:
                P4.rotate_left(tmp, a, rtk[s], s, lv, ws)
                P4.eltwise(ADD, a[0], a[0], tmp[0], nres=L)
                P4.eltwise(ADD, a[1], a[1], tmp[1], nres=L)

    p_ms, _ = time_steps(plain, 2, 1, stream)
    nrot_b, nrot_p = nct * len(steps), nplain * len(steps)
    r = {"workload": "rotate-and-sum over 2^15 slots (15 rotations+adds) for 64 ciphertexts, N=2^16, k=24, "
                     "K=25, level 23",
         "batched_rotations_s": round(nrot_b * 1000.0 / b_ms, 1),
         "plain_rotations_s": round(nrot_p * 1000.0 / p_ms, 1),
         "batched_over_plain": round((nrot_b / b_ms) / (nrot_p / p_ms), 3),
         "batch_ms": round(b_ms, 3),
         "batch_options": opts, "batch_options_source": src,
         "batched_hbm_frac": round(rot_bytes(65536, L, P4.kl) * nrot_b / (b_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
    del cts, acc, rk, rtk, tmp, ws
    torch.cuda.empty_cache()
    return r


PROC_SHAPES = [(8192, 4), (16384, 8), (32768, 16), (65536, 30)]


def procedure_table(args, stream, niter=50):
    """Plain (one call at a time) latency of each procedure of the path, in
    the style of the reference's benchmark (aoa-ntt/tests/benchmark.cpp:
    1159-1175, 1324-1344: one procedure, NITERATIONS back-to-back calls on one
    stream, mean latency per call), at the serving workload's shapes
    N = 2^13..2^16 with k = 4, 8, 16, 30 (prec 48), kl = k + 1, top level.
    Returns {procedure: {N: microseconds}}."""
    import torch
    from aoa_amd import Params
    out = {}

    def put(name, N, ms):
        out.setdefault(name, {})[str(N)] = round(ms * 1000.0, 2)

    for N_, k_ in PROC_SHAPES:
        P_ = Params(N_, k_, 0, PREC)
        R = P_.k + P_.kl
        lv = P_.k - 1
        x = rand_polys(P_, 2, P_.k)
        qb = rand_polys(P_, 1, R)[0]
        keys = rand_polys(P_, 2, R)
        c3 = [P_.zeros(P_.k), P_.zeros(P_.k)]
        ws = P_.workspace(lv)
        ms, _ = time_steps(lambda: P_.ntt(x[0], 0, P_.k, 0), niter, 3, stream)
        put("NTT", N_, ms)
        ms, _ = time_steps(lambda: P_.ntt(x[0], 0, P_.k, 1), niter, 3, stream)
        put("INTT", N_, ms)
        ms, _ = time_steps(lambda: P_.modup(qb, lv), niter, 3, stream)
        put("ModUp", N_, ms)
        ms, _ = time_steps(lambda: P_.moddown(c3[0], qb, lv), niter, 3, stream)
        put("ModDown", N_, ms)
        ms, _ = time_steps(lambda: P_.rescale(x[0], x[1], lv), niter, 3, stream)
        put("Rescale", N_, ms)
        ms, _ = time_steps(lambda: P_.rotate_left(c3, x, keys, 1, lv, ws), niter, 3, stream)
        put("Rotate", N_, ms)
        ms, _ = time_steps(lambda: P_.hommul(c3, x, x, keys, lv, ws), niter, 3, stream)
        put("Mul", N_, ms)
        del x, qb, keys, c3, ws, P_
        torch.cuda.empty_cache()
    return out


def serving_bench(args, rng):
    """Poisson arrivals; ring dimension uniform in {2^13..2^16} with
    {4, 8, 16, 30} limbs; each request is 1-5 HomMul+relin+rescale (at most
    its level) and 0-4 rotations by power-of-two steps, in random order.
    Keys stay resident; admission by an HBM pool budget.  One run at
    --rate, then a sweep of arrival rates up to past saturation: offered and
    served ops/s, p50 and p99 latency per rate."""
    import torch
    from aoa_amd import Params, Server
    shapes = [(8192, 4), (16384, 8), (32768, 16), (65536, 30)]
    steps = [1, 2, 4, 8, 16]
    sets = []
    for N_, k_ in shapes:
        P_ = Params(N_, k_, 0, PREC)
        evk_ = rand_polys(P_, 2, P_.k + P_.kl)
        rk = rand_polys(P_, 2 * len(steps), P_.k + P_.kl)
        rtk_ = {s: (rk[2 * i], rk[2 * i + 1]) for i, s in enumerate(steps)}
        ins = rand_polys(P_, 8, P_.k)  # 4 input ciphertexts and 4 operands per shape
        sets.append((P_, ins, evk_, rtk_))

    def run(rate, nreq):
        srv = Server(pool_bytes=24 << 30, step_ws_bytes=24 << 30, max_active=256)
        kss = [srv.register(P_, evk_, rtk_) for P_, _, evk_, rtk_ in sets]
        reqs = []
        for _ in range(nreq):
            s = int(rng.integers(0, len(shapes)))
            P_ = sets[s][0]
            nmul = min(int(rng.integers(1, 6)), P_.k - 1)
            nrot = int(rng.integers(0, 5))
            ops = [("mul",)] * nmul + [("rot", int(rng.choice(steps)))] * nrot
            order = rng.permutation(len(ops))
            reqs.append((s, [ops[i] for i in order]))
        gaps = rng.exponential(1.0 / rate, size=nreq)
        outs = [[sets[s][0].zeros(sets[s][0].k), sets[s][0].zeros(sets[s][0].k)] for s, _ in reqs]
        torch.cuda.synchronize()
        srv.start()
        ids = []
        t_next = time.perf_counter()
        for (s, ops), out, gap in zip(reqs, outs, gaps):
            t_next += gap
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
            P_, ins, _, _ = sets[s]
            j = int(rng.integers(0, 4))
            ids.append(srv.submit(kss[s], (ins[2 * j], ins[2 * j + 1]), P_.k - 1, True, (ins[2 * j + 1], ins[2 * j]),
                                  ops, out))
        deadline = time.perf_counter() + 300
        while time.perf_counter() < deadline:
            if all(srv.result(i)[0] >= Server.DONE for i in ids):
                break
            time.sleep(0.002)
        srv.stop()
        res = [srv.result(i) for i in ids]
        st = srv.stats()
        srv.close()
        lat = np.array([r[2] for r in res if r[0] == Server.DONE])
        nops = sum(len(o) for _, o in reqs)
        span = st["span_ms"] / 1000.0
        return {"requests": nreq, "completed": int(st["completed"]),
                "failed": int(sum(r[0] == Server.FAILED for r in res)),
                "arrival_rate_per_s": rate, "offered_ops_per_s": round(rate * nops / nreq, 1),
                "ops": int(st["ops"]), "steps": int(st["steps"]),
                "ops_per_s": round(st["ops"] / span, 1) if span > 0 else None,
                "requests_per_s": round(st["completed"] / span, 1) if span > 0 else None,
                "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
                "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
                "mean_ops_per_request": round(nops / nreq, 2),
                "resident_key_bytes": int(st["key_bytes"])}

    out = run(args.rate, args.requests)
    sweep = []
    for rate in (300, 600, 1200, 2400, 4800, 9600):
        r = run(rate, 400)
        sweep.append({k: r[k] for k in ("arrival_rate_per_s", "offered_ops_per_s", "ops_per_s", "p50_ms", "p99_ms",
                                         "completed", "failed")})
    out["sweep"] = sweep
    # saturation: served throughput stops following the offered load
    out["capacity_ops_per_s"] = max(r["ops_per_s"] or 0 for r in sweep)
    knee = [r["arrival_rate_per_s"] for r in sweep if (r["ops_per_s"] or 0) < 0.9 * r["offered_ops_per_s"]]
    out["saturation_rate_per_s"] = knee[0] if knee else None
    return out


def c1_pipeline(args, stream):
    """One plaintext product end to end at N=2^12, k=4 (prec 52), K=5 on the
    device: encode -> encrypt -> HomMul+relin+rescale -> decrypt -> decode of
    2048 slots (aoa_amd.pipeline), keys and randomness drawn on the device.
    Plain (one C-ABI call per step, from Python) and replayed as one HIP graph;
    beside it the oracle's restatement of the same sequence on the host cores,
    whose decoded values must equal the device's bit for bit."""
    import torch
    from aoa_amd import Params
    from aoa_amd.pipeline import Pipeline
    Nc, kc, klc, precc, slots, seed = 4096, 4, 5, 52, 2048, 42
    P = Params(Nc, kc, klc, precc)
    pipe = Pipeline(P, seed, slots)
    rng = np.random.default_rng(1)
    z1 = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    z2 = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    pipe.set_inputs(z1, z2, 1 << precc)
    nonce = 5
    pipe.run(nonce)
    eager = pipe.result()
    plain_ms, _ = time_steps(lambda: pipe.run(nonce), max(20, args.steps), 3, stream)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        pipe.run(nonce)
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        pipe.run(nonce)
    replays_ok = True
    for _ in range(3):
        pipe.out.zero_()
        g.replay()
        torch.cuda.synchronize()
        replays_ok &= bool(np.array_equal(pipe.result().view(np.uint64), eager.view(np.uint64)))
    graph_ms, _ = time_steps(g.replay, max(50, args.steps * 2), 5, stream)
    res = {"workload": "encode -> encrypt -> HomMul+relin+rescale -> decrypt -> decode, 2048 slots, N=2^12, "
                       "k=4 (prec 52), K=5, scale 2^52; keys, randomness and FFT on the device",
           "plain_ms": round(plain_ms, 4), "graph_ms": round(graph_ms, 4),
           "graph_pipelines_per_s": round(1000.0 / graph_ms, 1), "graph_replays_match_eager": replays_ok,
           "max_err": float(np.max(np.abs(eager - z1 * z2))), "bound": 0.005}
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from orc import Oracle
        O = Oracle(Nc, kc, klc, precc)
        keys = O.c1_keys(seed)
        threads = min(16, os.cpu_count() or 1)
        ts = {}
        for nt in (threads, 1):
            Oracle.threads(nt)
            ref, _ = O.c1_run(keys, z1, z2, 1 << precc, seed, nonce)
            t = []
            for _ in range(5):
                t0 = time.perf_counter()
                O.c1_run(keys, z1, z2, 1 << precc, seed, nonce)
                t.append(time.perf_counter() - t0)
            ts[nt] = float(np.median(t))
        Oracle.threads(threads)
        res["matches_oracle"] = bool(np.array_equal(ref.view(np.uint64), eager.view(np.uint64)))
        res["cpu_ms"] = round(ts[threads] * 1e3, 3)
        res["cpu_threads"] = threads
        res["cpu_1t_ms"] = round(ts[1] * 1e3, 3)
        res["cpu_sample"] = "oracle, same pipeline and inputs, median of 5 runs"
    return res


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                return ln.split(":", 1)[1].strip()
    except Exception:
        pass
    return None


def cpu_baseline(args, P, plain_items, evk, level):
    """The oracle (the reference's algorithm restated in C++ with OpenMP over
    limbs and coefficients) on the same HomMul+relin+rescale input, on the
    host cores: median of `cpu_sample` runs on all threads (at most 16) and of
    `cpu_sample_1` runs on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from orc import Oracle
    threads = min(16, os.cpu_count() or 1)
    O = Oracle(N, K_Q, K_B, PREC)
    c1, c2, _, _ = plain_items[0]
    h = [P.to_host(t) for t in list(c1) + list(c2) + evk]

    def run(nthreads, reps):
        Oracle.threads(nthreads)
        O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)  # warm-up
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)
            ts.append(time.perf_counter() - t)
        return float(np.median(ts))

    t_all = run(threads, args.cpu_sample)
    res = {"value": round(1.0 / t_all, 4), "unit": "ops/s", "cores": threads, "kind": "port",
           "sample": "%d HomMul+relin+rescale (N=2^16, k=30, K=31, level 29) on the oracle, median; %d threads"
                     % (args.cpu_sample, threads),
           "cpu": cpu_model(), "nproc": os.cpu_count()}
    if args.cpu_sample_1 > 0:
        t1 = run(1, args.cpu_sample_1)
        res["single_thread_ops_s"] = round(1.0 / t1, 4)
        res["single_thread_sample"] = "%d runs, median" % args.cpu_sample_1
    # forward and inverse NTT on 30 and 61 limbs (the HomMul's Q and QB
    # polynomials), median of 5 after a warm-up, all threads and one
    ntt = {}
    rng = np.random.default_rng(9)
    for R in (K_Q, K_Q + K_B):
        x = np.concatenate([rng.integers(0, O.primes[r], size=N, dtype=np.uint64) for r in range(R)])
        for nth in (threads, 1):
            Oracle.threads(nth)
            for d, name in ((0, "fwd"), (1, "inv")):
                y = x.copy()
                O.ntt(y, 0, R, d)
                ts = []
                for _ in range(5):
                    y = x.copy()
                    t = time.perf_counter()
                    O.ntt(y, 0, R, d)
                    ts.append(time.perf_counter() - t)
                ntt["%s_%d_limbs_ms_%dt" % (name, R, nth)] = round(1000.0 * float(np.median(ts)), 3)
    res["ntt_ms"] = ntt
    Oracle.threads(threads)
    return res


def calibration_copy():
    """1 GiB device-to-device copy (three times): the kernel whose FETCH_SIZE
    and WRITE_SIZE scale the counter readings to bytes."""
    import torch
    a = torch.empty(1 << 27, dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return 1 << 30


def profile_workload(args, stream):
    """One workload for a counter or trace profile: `plain` runs 20 plain
    HomMul+relin+rescale calls at the headline shape (after 3 untimed ones),
    `c3` the mixed-level N=2^15 batch against its plain calls, `c4` the
    rotate-and-sum batch against its plain rotations; then the calibration
    copy."""
    import torch
    from aoa_amd import Params
    torch.manual_seed(2)
    rng = np.random.default_rng(3)
    if args.profile_workload == "plain":
        P = Params(N, K_Q, K_B, PREC)
        evk = rand_polys(P, 2, K_Q + K_B)
        fn = plain_fn(P, evk, make_items(P, [K_Q - 1]), {})
        ms, _ = time_steps(fn, 20, 3, stream)
        res = {"plain_n16": {"workload": "plain HomMul+relin+rescale, N=2^16, k=30, K=31, level 29",
                             "ms_per_op": round(ms, 4), "ops_s": round(1000.0 / ms, 2)}}
    elif args.profile_workload == "c3":
        res = {"mixed_levels_n15": mixed_n15(args, rng, stream)}
    else:
        res = {"sumslots_n16": sumslots_n16(args, stream)}
    res["calibration_copy_bytes"] = calibration_copy()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32, help="ciphertext pairs per batched step")
    ap.add_argument("--cpu-sample", type=int, default=10, help="HomMuls timed on the CPU oracle, all threads (0: skip)")
    ap.add_argument("--cpu-sample-1", type=int, default=5, help="HomMuls timed on the CPU oracle, one thread")
    ap.add_argument("--no-extra", action="store_true", help="skip the N=2^15 and rotate-and-sum workloads")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving run")
    ap.add_argument("--requests", type=int, default=600, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=300.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--tune", action="store_true",
                    help="run the batch autotuner on the batched workloads first and record the chosen "
                         "options in aoa_amd/tuned/batch_options.json")
    ap.add_argument("--profile", action="store_true",
                    help="headline shape only (plus a 1 GiB copy for counter calibration), for rocprofv3")
    ap.add_argument("--profile-c1", action="store_true",
                    help="the N=2^12 end-to-end pipeline only (no CPU oracle), for rocprofv3")
    ap.add_argument("--profile-workload", choices=["plain", "c3", "c4"],
                    help="one other workload only, for rocprofv3 (plus the 1 GiB calibration copy): plain "
                         "HomMul+relin+rescale calls at the headline shape, the mixed-level N=2^15 batch, or "
                         "the N=2^16 rotate-and-sum batch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from aoa_amd import Batch, Params, tuning

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    stream = torch.cuda.current_stream()
    if args.profile_c1:
        args.cpu_sample = 0
        print(json.dumps({"c1_pipeline_n12": c1_pipeline(args, stream)}))
        return
    if args.profile_workload:
        print(json.dumps(profile_workload(args, stream)))
        return
    torch.manual_seed(2)
    P = Params(N, K_Q, K_B, PREC)
    level = K_Q - 1
    L = level + 1
    evk = rand_polys(P, 2, K_Q + K_B)
    ws_cache = {}

    # ---- plain form (one ciphertext pair per call) -------------------------
    plain_items = make_items(P, [level])
    plain = plain_fn(P, evk, plain_items, ws_cache)
    plain_ms, _ = time_steps(plain, args.steps * 4, args.warmup, stream)
    plain_ops = 1000.0 / plain_ms
    lat = latencies(plain, min(100, args.steps * 2), stream)

    # ---- batched form: the timed region ------------------------------------
    B = args.batch
    items = make_items(P, [level] * B)
    bt = Batch()
    bt.add_mul(P, evk, items)
    bt.build()
    opts, opts_src = batch_options(bt, tuning.key(N, K_Q + K_B, B, "mul"), args, stream)
    batch_ms, wall = time_steps(lambda: bt.launch(graph=True), args.steps, args.warmup, stream)
    batch_ops = B * 1000.0 / batch_ms
    blat = latencies(lambda: bt.launch(graph=True), min(20, args.steps), stream)

    # ---- dominant kernel: ModDown, timed per launch with HIP events --------
    split = stage_split(bt, stream, reps=5)
    md_ms, md_n = split["moddown"]
    md_launch_ms = md_ms / md_n
    md_bytes = B * 2 * moddown_bytes(N, L, K_B, rescale=True)  # one launch covers all B items, 2 polys each
    md_gbs = md_bytes / (md_launch_ms * 1e-3) / 1e9
    del bt
    torch.cuda.empty_cache()

    res = {
        "metric": "HomMul+relin+rescale ops/s at N=2^16 (batched vs plain); p50/p99 latency",
        "value": round(batch_ops, 2),
        "unit": "ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform random residues (ciphertexts and evaluation key, N=2^16, k=30, K=31), on device",
        "config": {"workload": "batched HomMul+relin+rescale, %d ciphertext pairs per step, N=2^16, k=30 "
                               "(prec 48), K=31, level 29" % B,
                   "N": N, "k": K_Q, "K": K_B, "prec": PREC, "level": level, "batch": B,
                   "batch_options": opts, "batch_options_source": opts_src},
        "plain_ops_s": round(plain_ops, 2),
        "batched_over_plain": round(batch_ops / plain_ops, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "batch_p50_ms": round(float(np.percentile(blat, 50)), 4),
        "batch_p99_ms": round(float(np.percentile(blat, 99)), 4),
        "op_hbm_frac": round(op_bytes(N, L, K_B) * batch_ops / (HBM_PEAK_GBS * 1e9), 4),
        "stages_ms_per_step": {s: round(v[0], 4) for s, v in split.items()},
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(md_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(md_gbs / HBM_PEAK_GBS, 4), "traffic": MODDOWN_TRAFFIC_BYTES.get(B),
                     "kernel": "aoa::k_moddown_d (ModDown + rescale, %d items x 2 polys)" % B,
                     "launch_ms": round(md_launch_ms, 4), "algorithmic_bytes": md_bytes,
                     "limiter": "valu",
                     "valu": valu_issue(MODDOWN_VALU_INSTS.get(B), md_launch_ms)},
    }

    if args.profile:
        res["calibration_copy_bytes"] = calibration_copy()
        print(json.dumps(res))
        return

    # ---- mixed levels at N=2^16: batched vs plain on the same items --------
    rng = np.random.default_rng(3)
    mixed_levels = [int(x) for x in rng.integers(1, K_Q, size=B)]
    mitems = make_items(P, mixed_levels)
    mplain_ms, _ = time_steps(plain_fn(P, evk, mitems, ws_cache), max(3, args.steps // 10), 1, stream)
    mb = Batch()
    mb.add_mul(P, evk, mitems)
    mb.build()
    mbatch_ms, _ = time_steps(lambda: mb.launch(graph=True), max(3, args.steps // 5), 2, stream)
    del mb, mitems, items
    torch.cuda.empty_cache()
    res["mixed_levels_n16"] = {"ops": B, "plain_ops_s": round(B * 1000.0 / mplain_ms, 2),
                               "batched_ops_s": round(B * 1000.0 / mbatch_ms, 2),
                               "batched_over_plain": round(mplain_ms / mbatch_ms, 3)}

    # ---- forward NTT of the k+K limbs (two kernels) ------------------------
    R = K_Q + K_B
    x = rand_polys(P, 1, R)[0]
    ntt_ms, _ = time_steps(lambda: P.ntt(x, 0, R, 0), max(20, args.steps), 3, stream)
    ntt_gbs = ntt_bytes(N, R) / (ntt_ms * 1e-3) / 1e9
    res["ntt_fwd_61_limbs"] = {"ms": round(ntt_ms, 4), "achieved_GBs": round(ntt_gbs, 1),
                               "frac": round(ntt_gbs / HBM_PEAK_GBS, 4)}
    del x

    if not args.no_extra:
        res["procedures_us"] = {"note": "plain per-call latency, mean of 50 calls, top level, N=2^13..2^16 with "
                                        "k=4, 8, 16, 30 (prec 48), kl=k+1",
                                **procedure_table(args, stream)}
        res["mixed_levels_n15"] = mixed_n15(args, rng, stream)
        res["sumslots_n16"] = sumslots_n16(args, stream)
        res["c1_pipeline_n12"] = c1_pipeline(args, stream)

    if not args.no_serving:
        res["serving"] = serving_bench(args, rng)
        torch.cuda.empty_cache()

    if args.cpu_sample > 0:
        res["cpu_baseline"] = cpu_baseline(args, P, plain_items, evk, level)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
