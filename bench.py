"""Benchmark: HomMul+relin+rescale ops/s at N=2^16 on one MI355X, batched vs
plain, with p50/p99 latency.

Headline workload (configs[1] shape): N = 2^16, k = 30 Q limbs (63-bit q0 +
29 x 48-bit, prec 48), K = 31 special limbs (55-bit), from the top level 29,
uniform random residues resident in HBM (synthetic ciphertexts and
evaluation key of that shape).  A step of the batched form is one hipGraph
replay of B HomMul+relin+rescale; a step of the plain form is one
aoa_hommul_relin_rescale call.  `value` is the batched throughput; the plain
throughput, their ratio, mixed-level and config-3 comparisons, the NTT
roofline and the CPU oracle baseline are reported beside it.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, K_Q, K_B, PREC = 65536, 30, 31, 48
HBM_PEAK_GBS = 8000.0


def op_bytes(N, L, K):
    """Algorithmic HBM bytes of one HomMul+relin+rescale at L active limbs:
    two input ciphertexts (4L limbs), the evaluation key on the active Q limbs
    and all special limbs (2(L+K)), the rescaled output (2(L-1))."""
    return 8 * N * (8 * L + 2 * K - 2)


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


def serving_bench(args, rng):
    """Config 5: Poisson arrivals; ring dimension uniform in {2^13..2^16} with
    {4, 8, 16, 30} limbs; each request is 1-5 HomMul+relin+rescale (at most
    its level) and 0-4 rotations by power-of-two steps, in random order.
    Keys stay resident; admission by an HBM pool budget."""
    import torch
    from aoa_amd import Params, Server
    shapes = [(8192, 4), (16384, 8), (32768, 16), (65536, 30)]
    steps = [1, 2, 4, 8, 16]
    srv = Server(pool_bytes=24 << 30, step_ws_bytes=24 << 30, max_active=256)
    sets = []
    for N_, k_ in shapes:
        P_ = Params(N_, k_, 0, PREC)
        evk_ = rand_polys(P_, 2, P_.k + P_.kl)
        rk = rand_polys(P_, 2 * len(steps), P_.k + P_.kl)
        rtk_ = {s: (rk[2 * i], rk[2 * i + 1]) for i, s in enumerate(steps)}
        ks = srv.register(P_, evk_, rtk_)
        ins = rand_polys(P_, 8, P_.k)  # 4 input ciphertexts and 4 operands per shape
        sets.append((P_, ks, ins, evk_, rtk_))
    reqs = []
    for _ in range(args.requests):
        s = int(rng.integers(0, len(shapes)))
        P_ = sets[s][0]
        nmul = int(rng.integers(1, 6))
        nmul = min(nmul, P_.k - 1)
        nrot = int(rng.integers(0, 5))
        ops = [("mul",)] * nmul + [("rot", int(rng.choice(steps)))] * nrot
        order = rng.permutation(len(ops))
        reqs.append((s, [ops[i] for i in order]))
    gaps = rng.exponential(1.0 / args.rate, size=args.requests)
    outs = [[sets[s][0].zeros(sets[s][0].k), sets[s][0].zeros(sets[s][0].k)] for s, _ in reqs]
    torch.cuda.synchronize()
    srv.start()
    ids = []
    t0 = time.perf_counter()
    t_next = t0
    for (s, ops), out, gap in zip(reqs, outs, gaps):
        t_next += gap
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        P_, ks, ins, _, _ = sets[s]
        j = int(rng.integers(0, 4))
        ids.append(srv.submit(ks, (ins[2 * j], ins[2 * j + 1]), P_.k - 1, True, (ins[2 * j + 1], ins[2 * j]), ops,
                              out))
    deadline = time.perf_counter() + 300
    while time.perf_counter() < deadline:
        if all(srv.result(i)[0] >= Server.DONE for i in ids):
            break
        time.sleep(0.005)
    srv.stop()
    res = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([r[2] for r in res if r[0] == Server.DONE])
    nops = sum(len(o) for _, o in reqs)
    return {"requests": args.requests, "completed": int(st["completed"]), "failed": int(sum(r[0] == Server.FAILED
                                                                                         for r in res)),
            "arrival_rate_per_s": args.rate, "ops": int(st["ops"]), "steps": int(st["steps"]),
            "ops_per_s": round(st["ops"] / (st["span_ms"] / 1000.0), 1) if st["span_ms"] > 0 else None,
            "requests_per_s": round(st["completed"] / (st["span_ms"] / 1000.0), 1) if st["span_ms"] > 0 else None,
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
            "mean_ops_per_request": round(nops / args.requests, 2),
            "resident_key_bytes": int(st["key_bytes"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=32, help="ciphertext pairs per batched step")
    ap.add_argument("--cpu-sample", type=int, default=1, help="HomMuls timed on the CPU oracle (0: skip)")
    ap.add_argument("--no-c3", action="store_true", help="skip the config-3 comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the config-5 serving run")
    ap.add_argument("--requests", type=int, default=600, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=300.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU baseline, no C3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU and this build has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from aoa_amd import Batch, Params

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    stream = torch.cuda.current_stream()
    torch.manual_seed(2)
    P = Params(N, K_Q, K_B, PREC)
    level = K_Q - 1
    L = level + 1
    evk = rand_polys(P, 2, K_Q + K_B)
    ws_cache = {}

    # ---- plain form at the C2 shape (one ciphertext per call) -------------
    plain_items = make_items(P, [level])
    plain = plain_fn(P, evk, plain_items, ws_cache)
    plain_ms, _ = time_steps(plain, args.steps * 4, args.warmup, stream)
    plain_ops = 1000.0 / plain_ms
    lat = latencies(plain, min(100, args.steps * 2), stream)

    # ---- batched form at the C2 shape --------------------------------------
    B = args.batch
    items = make_items(P, [level] * B)
    bt = Batch()
    bt.add_mul(P, evk, items)
    bt.build()
    batch_ms, wall = time_steps(lambda: bt.launch(graph=True), args.steps, args.warmup, stream)
    batch_ops = B * 1000.0 / batch_ms
    blat = latencies(lambda: bt.launch(graph=True), min(20, args.steps), stream)
    del bt

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

    # ---- dominant kernel roofline: forward NTT of the k+K limbs ------------
    R = K_Q + K_B
    x = rand_polys(P, 1, R)[0]
    nrep = max(20, args.steps)
    ntt_ms, _ = time_steps(lambda: P.ntt(x, 0, R, 0), nrep, 3, stream)
    ntt_gbs = ntt_bytes(N, R) / (ntt_ms * 1e-3) / 1e9

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
        "data": "synthetic: uniform random residues (ciphertexts and evaluation key of the C2 shape), on device",
        "config": {"workload": "batched HomMul+relin+rescale, %d ciphertext pairs per step, N=2^16, k=30 "
                               "(prec 48), K=31, level 29" % B,
                   "N": N, "k": K_Q, "K": K_B, "prec": PREC, "level": level, "batch": B},
        "plain_ops_s": round(plain_ops, 2),
        "batched_over_plain": round(batch_ops / plain_ops, 3),
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "batch_p50_ms": round(float(np.percentile(blat, 50)), 4),
        "batch_p99_ms": round(float(np.percentile(blat, 99)), 4),
        "mixed_levels_n16": {"ops": B, "plain_ops_s": round(B * 1000.0 / mplain_ms, 2),
                             "batched_ops_s": round(B * 1000.0 / mbatch_ms, 2),
                             "batched_over_plain": round(mplain_ms / mbatch_ms, 3)},
        "op_roofline_frac": round(op_bytes(N, L, K_B) * batch_ops / (HBM_PEAK_GBS * 1e9), 4),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(ntt_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ntt_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "forward NTT (column + row pass) of 61 limbs, N=2^16",
                     "launch_ms": round(ntt_ms, 4)},
    }

    # ---- config 3: 256 ciphertexts, N=2^15, 20 limbs, L uniform in [2, 20] --
    if not args.no_c3 and not args.profile:
        P3 = Params(32768, 20, 0, 48)
        evk3 = rand_polys(P3, 2, P3.k + P3.kl)
        lv3 = [int(x) for x in rng.integers(1, 20, size=256)]
        it3 = make_items(P3, lv3)
        ws3 = {}
        p3_ms, _ = time_steps(plain_fn(P3, evk3, it3, ws3), 3, 1, stream)
        b3 = Batch()
        b3.add_mul(P3, evk3, it3)
        b3.build()
        b3_ms, _ = time_steps(lambda: b3.launch(graph=True), 5, 2, stream)
        mean_bytes = float(np.mean([op_bytes(32768, lv + 1, P3.kl) for lv in lv3]))
        res["c3"] = {"ops": 256, "plain_ops_s": round(256 * 1000.0 / p3_ms, 2),
                     "batched_ops_s": round(256 * 1000.0 / b3_ms, 2),
                     "batched_over_plain": round(p3_ms / b3_ms, 3),
                     "batched_roofline_frac": round(mean_bytes * 256 / (b3_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)}
        del b3, it3
        torch.cuda.empty_cache()

    if not args.no_serving and not args.profile:
        res["serving"] = serving_bench(args, rng)
        torch.cuda.empty_cache()

    if args.cpu_sample > 0 and not args.profile:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from orc import Oracle
        threads = min(16, os.cpu_count() or 1)
        Oracle.threads(threads)
        O = Oracle(N, K_Q, K_B, PREC)
        c1, c2, _, _ = plain_items[0]
        h = [P.to_host(t) for t in list(c1) + list(c2) + evk]
        O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)  # warm-up
        t = time.perf_counter()
        for _ in range(args.cpu_sample):
            O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)
        cpu_s = (time.perf_counter() - t) / args.cpu_sample
        res["cpu_baseline"] = {"value": round(1.0 / cpu_s, 4), "unit": "ops/s", "cores": threads, "kind": "port",
                               "sample": "%d HomMul+relin+rescale at the C2 shape on the oracle (the reference's "
                                         "algorithm restated in C++, OpenMP over limbs/coefficients)" % args.cpu_sample}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
