"""Persisted launch options of batches, chosen by the autotuner
(aoa_batch_autotune) and keyed by workload: ring dimension N, residue count
R = k + K, number of items and the kind of item.

The table is a JSON file, by default aoa_amd/tuned/batch_options.json, of
{key: {"ntt_split", "prime_major", "xcd_chunks", "strict_split", "ms",
"candidates", "device", "timings"}}, where timings lists every candidate as
[ntt_split, prime_major, xcd_chunks, strict_split, ms, exact].  It replaces the reference's fixed NTT factor table
(aoa-ntt/include/newckks/cuda/htrans/ntt.h:13-44) with measurements of the
batches themselves.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuned", "batch_options.json")
FIELDS = ("ntt_split", "prime_major", "xcd_chunks", "strict_split")


def key(N, R, items, kind):
    return "N%d_R%d_items%d_%s" % (N, R, items, kind)


def load_table(path=DEFAULT_TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_table(table, path=DEFAULT_TABLE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")
    os.replace(tmp, path)


def apply(batch, k, table=None):
    """Sets the options stored under key k on the batch; returns them, or
    None (defaults kept) when the table has no entry."""
    table = load_table() if table is None else table
    e = table.get(k)
    if e is None:
        return None
    opts = {f: int(e[f]) for f in FIELDS}
    batch.set_options(**opts)
    return opts


def tune(batch, k, stream=None, reps=5, path=DEFAULT_TABLE, device=""):
    """Runs the autotuner on the batch, records the chosen options under k in
    the table at path, and returns (chosen options, all candidates)."""
    cands = batch.autotune(stream=stream, reps=reps)
    chosen = batch.options()
    best = min((c for c in cands if all(c[f] == chosen[f] for f in FIELDS)), key=lambda c: c["ms"])
    table = load_table(path)
    table[k] = dict(chosen, ms=round(best["ms"], 4), candidates=len(cands), device=device,
                    timings=[[c[f] for f in FIELDS] + [round(c["ms"], 4), int(c["exact"])] for c in cands])
    save_table(table, path)
    return chosen, cands
