#!/usr/bin/env python3
"""Spiral PIR answer-path benchmark on one MI355X.

A step is one full server answer for one query against a resident database:
coefficient expansion -> first-dimension database multiply (all n^2
sub-databases) -> from_ntt -> GSW fold, with the database, the query and its
keys already in HBM.  Default workload: (nu1, nu2) = (9, 6), n = 2 (2^15
records of 8 KiB, 2 GiB resident in NTT form).

Prints one JSON line (queries/s) with the database-multiply kernel's HBM
roofline and the CPU oracle's time for the same work on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-pir_amd"))

METRIC = "PIR queries/s + effective DB-scan GB/s, batched vs 1-query; p50/p99 latency"
HBM_PEAK_GBPS = 8000.0
# HBM bytes per k_db_mul launch for one query, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs scaled by a plain copy kernel's readings (FETCH x2.000, WRITE x1.000);
# profiles/db_mul_hbm_traffic_nu9_6.md.  Measured shapes only.
DB_MUL_TRAFFIC_BYTES = {(9, 6): (2142755 + 8192) * 1024}


def db_mul_algorithmic_bytes(shape) -> int:
    """DB residues + selectors + accumulator outputs for one query (u32 words)."""
    db = shape.records * shape.n2 * 2 * 2048 * 4
    sel = shape.rows * 2 * 2 * 2048 * 4
    out = shape.n2 * shape.cols * 2 * 2 * 2048 * 4
    return db + sel + out


def cpu_baseline(shape, q, left, right, vf, sample_cols: int) -> dict:
    """Time the CPU oracle (single thread) on a bounded sample of one query and
    extrapolate to the whole query: full expansion, the database multiply of one
    sub-database restricted to `sample_cols` columns, the fold of one
    sub-database."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.lib()
    t0 = time.perf_counter()
    v = O.expand(q, shape.nu1, shape.nu2, shape.t_gsw, shape.t_exp, left, right)
    t_exp = time.perf_counter() - t0
    g = np.random.default_rng(4242)
    R = shape.rows
    db = np.empty((sample_cols, R, 2, 2048), np.uint32)
    db[:, :, 0] = g.integers(0, 268369921, size=(sample_cols, R, 2048), dtype=np.uint32)
    db[:, :, 1] = g.integers(0, 249561089, size=(sample_cols, R, 2048), dtype=np.uint32)
    sel = np.ascontiguousarray(v[0:2 * R:2])
    nu2s = int(np.log2(sample_cols))
    t0 = time.perf_counter()
    O.db_mul(db, sel, shape.nu1, nu2s)
    t_db = (time.perf_counter() - t0) * (shape.cols / sample_cols) * shape.n2
    accs = g.integers(0, 2**32, size=(shape.cols, 2, 2, 2048), dtype=np.uint32)
    accs[:, :, 0] &= np.uint32(0x00FFFFFF)  # raw values below 2^56 > Q is fine for timing
    vneg = O.folding_neg(vf, shape.nu2, shape.t_gsw)
    t0 = time.perf_counter()
    O.fold(accs, shape.nu2, shape.t_gsw, vf, vneg)
    t_fold = (time.perf_counter() - t0) * shape.n2
    t_q = t_exp + t_db + t_fold
    return {
        "value": 1.0 / t_q,
        "unit": "queries/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle/spiral_oracle.c on the host, 1 thread: full expansion ({t_exp:.2f} s), "
                   f"DB multiply of 1 of {shape.n2} sub-databases x {sample_cols} of {shape.cols} columns "
                   f"(x{shape.n2 * shape.cols // sample_cols} -> {t_db:.2f} s), fold of 1 sub-database "
                   f"(x{shape.n2} -> {t_fold:.2f} s); {t_q:.2f} s per query"),
        "stage_s": {"expand": t_exp, "db_mul": t_db, "fold": t_fold},
    }


def run_batched(torch, Batch, Client, db, shape, records, B: int, steps: int, stream, eager: bool = False) -> dict:
    """B queries (8 clients) answered per hipGraph launch; queries/s over `steps` launches."""
    n_clients = min(8, B)
    clients = [Client(shape, 2100 + k) for k in range(n_clients)]
    keys = [c.keygen() for c in clients]
    g = np.random.default_rng(2200)
    idx = [int(x) for x in g.integers(0, shape.records, size=B)]
    batch = Batch(db, B)
    for slot in range(B):
        c = slot % n_clients
        q, vf = clients[c].query(idx[slot])
        batch.load(slot, q, keys[c][0], keys[c][1], vf)
    sptr = stream.cuda_stream
    batch.run(B, 1, sptr)  # expansion once, so the autotuner can time the database multiply
    tuned = batch.tune(B)
    def launch():
        if eager:
            batch.run(B, 7, sptr)
        else:
            batch.graph_launch(B, sptr)

    launch()  # first graph launch captures the graph
    launch()
    torch.cuda.synchronize()
    resp = batch.responses(B)
    bad = [s for s in range(B) if not np.array_equal(clients[s % n_clients].decode(resp[s]), records[idx[s]])]
    ok = not bad
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        launch()
    stop.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    batch.close()
    return {"B": B, "ms_per_batch": ms, "queries_per_s": B * 1000.0 / ms, "steps": steps,
            "launch": "eager (one launch per kernel)" if eager else "hipGraph replay of the whole stage chain",
            "db_mul_autotune": tuned, "decoded_all_correct": ok, "slots_decoded_wrong": bad}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nu1", type=int, default=9)
    ap.add_argument("--nu2", type=int, default=6)
    ap.add_argument("--index", type=int, default=12345)
    ap.add_argument("--cpu-sample-cols", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=64, help="batched leg: queries per graph launch (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-eager", action="store_true", help="batched leg without hipGraph replay")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by database columns across GPUs, which is described but not built.",
              file=sys.stderr)
        return 2

    import torch
    from spiral_pir import Shape, version
    from spiral_pir.client import Client
    from spiral_pir.server import (Batch, Context, Database, Query, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD,
                                   STAGES_ALL)

    torch.cuda.init()
    if args.copy_mib > 0:
        # A copy kernel of known size: reads and writes copy_mib MiB per dispatch.
        src = torch.empty(args.copy_mib << 18, dtype=torch.int32, device="cuda").fill_(7)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)  # an elementwise kernel, never a DMA copy
        torch.cuda.synchronize()
        del src, dst
    shape = Shape(args.nu1, args.nu2)
    index = args.index % shape.records
    ctx = Context(0)
    records = np.random.default_rng(1001).integers(0, 256, size=(shape.records, shape.n2, 2048), dtype=np.uint8)
    db = Database(ctx, shape)
    t0 = time.perf_counter()
    db.load_bytes(records)
    t_ingest = time.perf_counter() - t0
    client = Client(shape, 2001)
    left, right = client.keygen()
    q, vf = client.query(index)
    query = Query(db)
    query.load(q, left, right, vf)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    for _ in range(args.warmup):
        query.run(STAGES_ALL, sptr)
    torch.cuda.synchronize()
    decoded = client.decode(query.read())
    correct = bool(np.array_equal(decoded, records[index]))
    single_correct = correct

    # Timed region: K whole answers, events around the database multiply.
    K = args.steps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(K):
        query.run(STAGE_EXPAND, sptr)
        ev[k][0].record(stream)
        query.run(STAGE_DB_MUL, sptr)
        ev[k][1].record(stream)
        query.run(STAGE_FOLD, sptr)
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)
    ms_per_step = total_ms / K
    dbmul_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))

    # Stage breakdown (separate pass, events between every stage).
    reps = max(3, min(K, 10))
    st = {"expand": [], "db_mul": [], "fold": []}
    for _ in range(reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record(stream)
        query.run(STAGE_EXPAND, sptr)
        e[1].record(stream)
        query.run(STAGE_DB_MUL, sptr)
        e[2].record(stream)
        query.run(STAGE_FOLD, sptr)
        e[3].record(stream)
        torch.cuda.synchronize()
        st["expand"].append(e[0].elapsed_time(e[1]))
        st["db_mul"].append(e[1].elapsed_time(e[2]))
        st["fold"].append(e[2].elapsed_time(e[3]))
    stage_ms = {k: float(np.median(v)) for k, v in st.items()}

    batched = None
    if args.batch > 0:
        batched = run_batched(torch, Batch, Client, db, shape, records, args.batch, args.batch_steps, stream,
                              eager=args.batch_eager)
        correct = correct and batched["decoded_all_correct"]

    alg_bytes = db_mul_algorithmic_bytes(shape)
    achieved = alg_bytes / (dbmul_ms * 1e-3) / 1e9
    db_scan_gbps = shape.db_device_bytes / (dbmul_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_step,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: uniform random record bytes (PCG64 seed 1001), client keys/query seed 2001",
        "config": {
            "workload": f"single query, (nu1,nu2)=({shape.nu1},{shape.nu2}), n=2 (4 sub-databases), "
                        f"{shape.records} records x {shape.record_bytes} B, d=2048, q1=268369921, "
                        f"q2=249561089, gadget 2^7 x 8, T_GSW=T_EXP=8",
            "records": shape.records,
            "db_resident_bytes": shape.db_device_bytes,
        },
        "db_scan_gbps": db_scan_gbps,
        "db_mul_ms": dbmul_ms,
        "stage_ms": stage_ms,
        "decoded_record_correct": correct,
        "single_query_decoded_correct": single_correct,
        "roofline": {
            "kernel": "k_db_mul (first-dimension database multiply)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": DB_MUL_TRAFFIC_BYTES.get((shape.nu1, shape.nu2)),
            "traffic_unit": "bytes per launch",
        },
        "batched": None if batched is None else {k: v for k, v in batched.items() if k != "decoded_all_correct"},
        "ingest_s": t_ingest,
        "library": version(),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(shape, q, left, right, vf, args.cpu_sample_cols)
    if batched is not None:
        line["batched"]["speedup_vs_single_query"] = batched["queries_per_s"] / line["value"]
    print(json.dumps(line), flush=True)
    query.close()
    db.close()
    ctx.close()
    return 0 if correct else 1


if __name__ == "__main__":
    sys.exit(main())
