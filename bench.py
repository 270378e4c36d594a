"""Benchmark: HomMul+relin+rescale ops/s at N=2^16 on one MI355X.

Workload (one "step" = one HomMul+relin+rescale of a ciphertext pair):
N = 2^16, k = 30 Q limbs (63-bit q0 + 29 x 48-bit, prec 48), K = 31 special
limbs (55-bit), from the top level 29, uniform random residues in HBM
(synthetic ciphertexts and evaluation key of that shape).

Prints one JSON line.  `value` is the throughput of the batched form when it
is available, else of the plain one-ciphertext-per-call form; both are
reported.  Inputs are resident in HBM before the timed region starts.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=1, help="HomMuls timed on the CPU oracle (0: skip)")
    ap.add_argument("--profile", action="store_true", help="short run for rocprofv3 (no CPU baseline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU and this build has no "
              "multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from aoa_amd import Params

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    P = Params(N, K_Q, K_B, PREC)
    level = K_Q - 1
    L = level + 1
    rng = np.random.default_rng(2)

    def rand_poly(nlimbs):
        out = np.empty(nlimbs * N, dtype=np.uint64)
        for r in range(nlimbs):
            out[r * N:(r + 1) * N] = rng.integers(0, P.primes[r if r < K_Q else r], size=N, dtype=np.uint64)
        return out

    c1 = [P.to_device(rand_poly(K_Q)) for _ in range(2)]
    c2 = [P.to_device(rand_poly(K_Q)) for _ in range(2)]
    evk = [P.to_device(rand_poly(K_Q + K_B)) for _ in range(2)]
    out = [P.zeros(K_Q), P.zeros(K_Q)]
    ws = P.workspace(level)
    stream = torch.cuda.current_stream()

    def plain_step():
        P.hommul(out, c1, c2, evk, level, ws)

    # ---- plain form: throughput -------------------------------------------
    for _ in range(args.warmup):
        plain_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        plain_step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plain_ms = ev0.elapsed_time(ev1) / args.steps
    plain_ops = 1000.0 / plain_ms

    # ---- per-op latency (plain, one op in flight) ---------------------------
    lat = []
    nlat = min(args.steps, 100)
    for _ in range(nlat):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        plain_step()
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat = np.array(lat)

    # ---- dominant-kernel roofline: forward NTT of the k+K limbs of a QB poly --
    R = K_Q + K_B
    x = P.to_device(rand_poly(R))
    nrep = max(20, args.steps // 4)
    for _ in range(3):
        P.ntt(x, 0, R, 0)
        P.ntt(x, 0, R, 1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(nrep):
        P.ntt(x, 0, R, 0)
    e1.record(stream)
    torch.cuda.synchronize()
    ntt_ms = e0.elapsed_time(e1) / nrep
    ntt_gbs = ntt_bytes(N, R) / (ntt_ms * 1e-3) / 1e9

    res = {
        "metric": "HomMul+relin+rescale ops/s at N=2^16 (batched vs plain); p50/p99 latency",
        "value": round(plain_ops, 2),
        "unit": "ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: uniform random residues (ciphertexts and evaluation key of the C2 shape)",
        "config": {"workload": "plain HomMul+relin+rescale, N=2^16, k=30 (prec 48), K=31, level 29",
                   "N": N, "k": K_Q, "K": K_B, "prec": PREC, "level": level},
        "plain_ops_s": round(plain_ops, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "op_roofline_frac": round(op_bytes(N, L, K_B) * plain_ops / (HBM_PEAK_GBS * 1e9), 4),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(ntt_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ntt_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "forward NTT (column + row pass) of 61 limbs",
                     "launch_ms": round(ntt_ms, 4)},
    }

    if args.cpu_sample > 0 and not args.profile:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from orc import Oracle
        threads = min(16, os.cpu_count() or 1)
        Oracle.threads(threads)
        O = Oracle(N, K_Q, K_B, PREC)
        h = [P.to_host(t) for t in c1 + c2 + evk]
        O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)  # warm-up
        t = time.perf_counter()
        for _ in range(args.cpu_sample):
            O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), level)
        cpu_s = (time.perf_counter() - t) / args.cpu_sample
        res["cpu_baseline"] = {"value": round(1.0 / cpu_s, 4), "unit": "ops/s", "cores": threads,
                               "kind": "port",
                               "sample": "%d HomMul+relin+rescale at the same shape on the oracle (the "
                                         "reference's algorithm restated in C++, OpenMP over limbs/"
                                         "coefficients)" % args.cpu_sample}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
