#!/usr/bin/env python3
"""Throughput bench of the batched Glynn permanent on one MI355X.

Workload (one "step"): the plain permanent over
  * 65536 complex128 16x16 matrices, each 16 random rows (with repetition) of
    the first 16 columns of a Haar-random 64x64 unitary (QR of a complex
    Gaussian, phases fixed), numpy default_rng(1); and
  * 1024 matrices at n = 24 drawn the same way from the same unitary,
in the fp64 tier (two kernel launches: n=16 and n=24). Inputs are resident in
HBM before timing starts. The dd tier is timed on the same inputs and reported
beside it. Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--dd-steps K2] [--cpu-seconds S]

The value is evals/s of the whole job (65536 + 1024 evaluations per step).
The roofline object describes the dominant kernel (n = 24): canonical flops
2^(n-1) (8n - 4) per evaluation (complex MAC = 8 real flops) over its average
launch duration measured with HIP events on the launch stream, against the
MI355X FP64 vector peak of 78.6 TFLOP/s; executed_tflops / executed_frac give
the FP64 work the kernel issues (hardware counters, EXECUTED_F64) over the
same time. cpu_baseline times the CPU oracle
(a restatement of the reference's cpu_ref Glynn walk) on the host cores on a
bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "evals/s perm/haf n=16,24 fp64+dd; served p50/p99 latency; rel err vs mpmath"
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# HBM bytes of one perm_kernel<24, fp64> launch (1024 x 24x24), measured with
# hardware counters (FETCH_SIZE, WRITE_SIZE) calibrated against a plain copy
TRAFFIC_N24 = 13742848
# Executed FP64 work per evaluation of a cell, measured with hardware counters
# (SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 summed over the cell's dispatches,
# divided by its batch; these kernels' instruction streams do not depend on
# the input values): (flop, FP64 VALU wave instructions). flop counts 64
# lanes per wave instruction, 2 per FMA. Canonical flops count the reference
# algorithm as written; executed flops count what the kernels issue.
EXECUTED_F64 = {
    "perm_fp64_n24": (2.014479e+09, 1.888513e+07),
    "perm_fp64_n16": (5169792, 48586),
    "haf_fp64_n16": (7939584, 65744),
    "haf_fp64_n24": (5.18053e+08, 4206890),
    "lhaf_fp64_n16": (9162176, 78319),
    "lhaf_fp64_n24": (5.749229e+08, 4794570),
    "haf_fp64_config1_n8": (19136, 184.0),
    "tor_fp64_nm8": (60288, 611),
    "tor_fp64_nm12": (505728, 5099),
    "tor_fp64_nm16": (8459776, 85258),
    "tor_fp64_nm20": (1.429188e+08, 1441431),
}
# FP64 VALU wave instructions per second at the 78.6 TFLOP/s peak: one per CU
# per clock (256 CUs, 2.4 GHz; an FMA over 64 lanes is 128 flop)
FP64_ISSUE_PER_S = 256 * 2.4e9


def executed(name: str, evals: int, seconds: float) -> dict:
    """executed TFLOP/s and FP64 issue fraction of a cell, from its counters"""
    if name not in EXECUTED_F64 or seconds <= 0:
        return {}
    flop, insts = EXECUTED_F64[name]
    return {"tflops_executed": flop * evals / seconds / 1e12,
            "fp64_issue_frac": insts * evals / seconds / FP64_ISSUE_PER_S}


def workload():
    from gbskernels_amd.workloads import perm_haar_rows_workload
    return perm_haar_rows_workload()


def perm_flops(n: int) -> float:
    return float(2 ** (n - 1) * (8 * n - 4))


def haf_flops(N: int, loop: bool = False) -> float:
    """canonical power-trace flops of one N x N (loop) hafnian"""
    from math import comb
    n, f = N // 2, 0.0
    for m in range(1, n + 1):
        s = 2.0 * m
        f += comb(n, m) * ((n - 1) * 8 * s ** 3 + n * 2 * s + 4 * n * (n + 1) + 4 * n + 2)
        if loop:
            f += comb(n, m) * n * (8 * s * s + 8 * s + 6)
    return f


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dd-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--haf", type=int, default=1, help="also time haf/lhaf buckets at N=16,24 (0 skips)")
    ap.add_argument("--perm-tiers", type=int, default=1,
                    help="also time the permanent's kappa, certified and certified-dd tiers at n = 16, 20, 24 (0 skips)")
    ap.add_argument("--tor", type=int, default=1,
                    help="also time the torontonian (fp64, certified, certified-dd) at nm = 8, 12, 16, 20 (0 skips)")
    ap.add_argument("--repeated", type=int, default=1,
                    help="also time the repeated-row sieve workload, M = 6, q = 2..6 (0 skips)")
    ap.add_argument("--tor-single", type=int, default=1,
                    help="also time the single-large torontonian at 16..32 modes (0 skips)")
    ap.add_argument("--sampler", type=int, default=1,
                    help="also time the resident GBS sampler on the reference's sampler configurations (0 skips)")
    ap.add_argument("--reference-table", type=int, default=1,
                    help="also time the reference's kernel-only benchmark shapes (B = 4096) against its "
                         "published medians (0 skips)")
    ap.add_argument("--accuracy", type=int, default=1,
                    help="also measure relative error against the mpmath ref tier, physical and "
                         "cancelling inputs, every tier (0 skips)")
    ap.add_argument("--served-requests", type=int, default=40000,
                    help="requests of the served mixed-size workload (0 skips)")
    ap.add_argument("--served-batch", type=int, default=2048, help="requests per batch / window")
    ap.add_argument("--config3-requests", type=int, default=200000,
                    help="ragged haf/lhaf requests through the server (0 skips)")
    ap.add_argument("--config5-requests", type=int, default=40000,
                    help="sampler-stream requests (Poisson 50k req/s) through the server (0 skips)")
    ap.add_argument("--reps", type=int, default=7, help="timed repeats per side-kernel cell (median, IQR)")
    ap.add_argument("--cells", type=lambda t: [c for c in t.split(",") if c], default=None,
                    help="comma-separated side-kernel cells (e.g. haf_fp64_n24,tor_fp64_nm20): time only "
                         "these, print their record and exit (for profiler runs)")
    ap.add_argument("--counter-pass", action="store_true",
                    help="with --cells: run each named cell once, in the order given, with a one-element "
                         "fill kernel after each (splits a hardware-counter run by cell), and exit")
    ap.add_argument("--detail", default="bench_detail.json",
                    help="file receiving the full measurement record (JSON)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (a known-size kernel that "
                         "calibrates hardware byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(instances are independent: several GPUs would each take a share of the batch)", file=sys.stderr)
        return 2

    # one hardware queue per server stream (4 slots x 4 streams): HIP's default
    # of 4 queues makes 16 streams share them, and a light bucket then waits
    # behind a heavy kernel of another slot (sampler stream p50 18 ms -> 6 ms)
    try:
        hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    except ValueError:
        hwq = 4
    if hwq < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    import torch
    from gbskernels_amd import _lib

    L = _lib.lib()
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": src.numel() * 8}))
        return 0
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    if args.cells and args.counter_pass:
        made = {}
        for make in (perm_tier_cells, haf_cells, tor_cells):
            for c in (make(L, torch, dev, stream, headline=True) if make is perm_tier_cells
                      else make(L, torch, dev, stream)):
                if c.name in args.cells:
                    made[c.name] = c
        torch.cuda.synchronize()
        for name in args.cells:
            made[name].go()
            torch.cuda.synchronize()
            torch.full((1,), 7.0, device=dev)  # marker kernel between cells
            torch.cuda.synchronize()
            print(f"bench: counter pass ran {name} (batch {made[name].meta['batch']})", file=sys.stderr, flush=True)
        print(json.dumps({"counter_pass": [[n, made[n].meta["batch"]] for n in args.cells]}))
        return 0
    if args.cells:
        out = {}
        for gi, make in enumerate((perm_tier_cells, haf_cells, tor_cells)):
            print(f"bench: cells {make.__name__}", file=sys.stderr, flush=True)
            made = make(L, torch, dev, stream, headline=True) if make is perm_tier_cells else make(L, torch, dev, stream)
            cells = [c for c in made if c.name in args.cells]
            out.update(run_cells(cells, torch, stream, args.reps, seed=100 + gi))
        print(json.dumps({"cells": out}))
        return 0

    m16, m24 = workload()
    cases = [(16, m16), (24, m24)]
    dev_in, dev_out, dev_ws = {}, {}, {}
    for n, m in cases:
        dev_in[n] = torch.from_numpy(m.view(np.float64)).to(dev)
        dev_out[n] = torch.empty(2 * m.shape[0], dtype=torch.float64, device=dev)
        for tier in (0, 1):
            wsb = int(L.gbs_workspace_bytes(0, tier, n, m.shape[0]))
            dev_ws[(n, tier)] = torch.empty(max(wsb, 256), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def launch(n, tier):
        ws = dev_ws[(n, tier)]
        rc = L.gbs_eval_device(0, tier, dev_in[n].data_ptr(), n, dev_in[n].numel() // (2 * n * n),
                               dev_out[n].data_ptr(), None, ws.data_ptr(), ws.numel(), sp)
        if rc:
            raise RuntimeError(f"gbs_eval_device rc={rc}")

    def run_steps(tier, steps, warmup):
        for _ in range(warmup):
            for n, _m in cases:
                launch(n, tier)
        torch.cuda.synchronize()
        evs = {n: [] for n, _ in cases}
        t0 = time.perf_counter()
        for _ in range(steps):
            for n, _m in cases:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                launch(n, tier)
                e1.record(stream)
                evs[n].append((e0, e1))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        per = {n: np.array([a.elapsed_time(b) for a, b in evs[n]]) for n in evs}
        kern_ms = {n: float(np.mean(v)) for n, v in per.items()}
        return wall, kern_ms, per

    print(f"bench: fp64 warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
    wall, kern_ms, per_ms = run_steps(0, args.steps, args.warmup)
    q = lambda v: (float(np.median(v)), [float(np.percentile(v, 25)), float(np.percentile(v, 75))])
    kern_med = {n: q(v) for n, v in per_ms.items()}
    step_med, step_iqr = q(sum(per_ms.values()))
    evals_per_step = sum(m.shape[0] for _, m in cases)
    value = evals_per_step * args.steps / wall
    out16 = dev_out[16].cpu().numpy().view(np.complex128).copy()
    out24 = dev_out[24].cpu().numpy().view(np.complex128).copy()

    dd_wall, dd_kern_ms, dd16, dd24 = None, None, None, None
    if args.dd_steps > 0:
        print(f"bench: dd steps={args.dd_steps}", file=sys.stderr, flush=True)
        dd_wall, dd_kern_ms, _ = run_steps(1, args.dd_steps, 1)
        dd16 = dev_out[16].cpu().numpy().view(np.complex128).copy()
        dd24 = dev_out[24].cpu().numpy().view(np.complex128).copy()

    dom_n = 24
    dom_flops = perm_flops(dom_n) * m24.shape[0]
    dom_s = kern_ms[dom_n] * 1e-3
    # compute-bound ("mfma" class of the schema), but on the FP64 vector ALU: the
    # Glynn walk has no matrix-product shape; peak = MI355X FP64 vector 78.6 TFLOP/s
    roofline = {"bound": "mfma", "achieved": dom_flops / dom_s / 1e12, "peak": PEAK_FP64 / 1e12,
                "unit": "TFLOP/s", "frac": dom_flops / dom_s / PEAK_FP64, "traffic": TRAFFIC_N24,
                "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE (x2.0) + WRITE_SIZE (x1.0), "
                                "scaled by a 256 MiB copy; profiles/perm_glynn_n24_hbm_traffic.md; "
                                "algorithmic input 9437184 B",
                "kernel": f"perm_kernel<{dom_n},fp64>", "flops_per_launch": dom_flops,
                "avg_launch_ms": kern_ms[dom_n], "pipe": "FP64 VALU (no MFMA)"}
    ex = executed(f"perm_fp64_n{dom_n}", m24.shape[0], dom_s)
    if ex:
        roofline["executed_tflops"] = ex["tflops_executed"]
        roofline["executed_frac"] = ex["tflops_executed"] * 1e12 / PEAK_FP64
        roofline["fp64_issue_frac"] = ex["fp64_issue_frac"]

    per_kernel = {}
    for n, m in cases:
        f = perm_flops(n) * m.shape[0]
        per_kernel[f"perm_fp64_n{n}"] = {"batch": m.shape[0], "avg_ms": kern_ms[n],
                                         "median_ms": kern_med[n][0], "iqr_ms": kern_med[n][1],
                                         "reps": args.steps, "evals_per_s": m.shape[0] / (kern_ms[n] * 1e-3),
                                         "tflops": f / (kern_ms[n] * 1e-3) / 1e12}
        per_kernel[f"perm_fp64_n{n}"].update(executed(f"perm_fp64_n{n}", m.shape[0], kern_ms[n] * 1e-3))
        if dd_kern_ms:
            per_kernel[f"perm_dd_n{n}"] = {"batch": m.shape[0], "avg_ms": dd_kern_ms[n],
                                           "evals_per_s": m.shape[0] / (dd_kern_ms[n] * 1e-3)}

    # the side measurements report a Python-level failure in the detail record
    # instead of losing the headline figure
    def side(name, fn):
        print(f"bench: {name}", file=sys.stderr, flush=True)
        try:
            return fn()
        except Exception as exc:  # noqa: BLE001
            print(f"bench: {name} failed: {exc!r}", file=sys.stderr, flush=True)
            return {"error": repr(exc)}

    cell_groups = []
    if args.perm_tiers:
        cell_groups.append(("perm tiers n=16,20,24", perm_tier_cells))
    if args.haf:
        cell_groups.append(("haf/lhaf n=16,24", haf_cells))
    if args.tor:
        cell_groups.append(("tor nm=8,12,16,20 and haf config 1", tor_cells))
    for gi, (gname, make) in enumerate(cell_groups):
        def timed_group(make=make, gi=gi):
            cells = make(L, torch, dev, stream)
            return run_cells(cells, torch, stream, args.reps, seed=100 + gi)
        per_kernel.update(side(gname, timed_group))
    repeated = side("repeated-row sieve", repeated_bench) if args.repeated else None
    tor_single = side("single-large torontonian", tor_single_bench) if args.tor_single else None
    sampler = side("resident sampler", sampler_bench) if args.sampler else None
    reference = side("reference kernel table", reference_table) if args.reference_table else None
    served = side("served mix", lambda: served_bench(args)) if args.served_requests > 0 else None
    streams = (side("config streams", lambda: config_streams(args))
               if args.config3_requests > 0 or args.config5_requests > 0 else None)
    accuracy = side("accuracy vs mpmath", accuracy_bench) if args.accuracy else None

    cpu, parity = None, None
    if not args.no_cpu:
        print("bench: CPU baseline and CPU agreement gate", file=sys.stderr, flush=True)
        cpu, parity = cpu_baseline(m16, m24, args.cpu_seconds, out16, out24, dd16, dd24)
        if max(parity["n16_fp64_vs_oracle_dd"], parity["n24_fp64_vs_oracle_dd"]) > 1e-10:
            print(f"bench: PARITY FAILURE {parity}", file=sys.stderr)
            return 1

    detail = {
        "roofline": roofline, "cpu_baseline": cpu, "parity": parity, "per_kernel": per_kernel,
        "dd": ({"evals_per_s": evals_per_step * args.dd_steps / dd_wall,
                "ms_per_step": dd_wall / args.dd_steps * 1e3, "steps": args.dd_steps} if dd_wall else None),
        "accuracy": accuracy, "served": served, "streams": streams, "repeated_ab": repeated,
        "tor_single": tor_single, "sampler": sampler, "reference_table": reference,
    }
    detail_path = write_detail(detail, args.detail)
    # the detailed legs as earlier lines (one per leg), the headline last
    for k in ("per_kernel", "accuracy", "served", "streams", "repeated_ab", "tor_single", "sampler",
              "reference_table"):
        if detail[k] is not None:
            print(json.dumps({"bench_detail": k, k: detail[k]}))

    line = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "ms_per_step_median_iqr": [step_med, step_iqr[0], step_iqr[1]], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "perm Glynn fp64: 65536 x 16x16 + 1024 x 24x24 complex128 rows of a "
                               "Haar 64x64 unitary (default_rng(1))",
                   "global_batch": evals_per_step, "tier": "fp64"},
        "roofline": {k: roofline[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic", "kernel",
                                              "avg_launch_ms", "pipe", "executed_tflops", "executed_frac")
                     if k in roofline},
        "cpu_baseline": ({k: cpu[k] for k in ("value", "unit", "cores", "kind", "value_1core", "sample_short")}
                         if cpu else None),
        "dd": ({"evals_per_s": round(detail["dd"]["evals_per_s"], 1)} if dd_wall else None),
        "parity": ({k: float(f"{v:.3g}") for k, v in parity.items() if k != "samples"} if parity else None),
        "served": served_summary(served, streams),
        "vs_reference_kernels": ({k: (round(reference[k], 2) if isinstance(reference[k], float) else reference[k])
                                  for k in ("cells_compared", "geomean_speedup", "min_speedup")}
                                 if isinstance(reference, dict) and "rows" in reference else None),
        "detail": detail_path,
    }
    text = json.dumps(line)
    if len(text) > 2000:  # the driver keeps a short tail: the headline stays one short line
        line.pop("served")
        text = json.dumps(line)
    print(text)
    return 0


def write_detail(detail, path):
    """Writes the full measurement record as JSON; returns the path written, or
    None when it cannot be written."""
    try:
        p = Path(path)
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(detail, indent=1))
        return str(path)
    except OSError as exc:
        print(f"bench: cannot write {path}: {exc!r}", file=sys.stderr)
        return None


def served_summary(served, streams):
    """One-line summary of the served legs: speed-up over the plain per-size
    path with p50 / p99 of the saturated and Poisson runs, and of the
    sampler-stream configuration."""
    out = {}
    if isinstance(served, dict) and "speedup_saturated_vs_plain" in served:
        out["speedup_vs_plain"] = round(served["speedup_saturated_vs_plain"], 3)
        out["identical_values"] = served["identical_values"]
        for k, short in (("served_saturated", "sat"), ("served_poisson_50k", "poisson50k")):
            r = served[k]
            out[short] = [float(f"{r['evals_per_s']:.4g}"), round(r["p50_ms"], 3), round(r["p99_ms"], 3)]
    if isinstance(streams, dict):
        for k, short in (("ragged_haf_lhaf_auto", "config3"), ("sampler_stream_poisson_50k", "config5")):
            if k in streams:
                r = streams[k]
                out[short] = [float(f"{r['evals_per_s']:.4g}"), round(r["p50_ms"], 3), round(r["p99_ms"], 3)]
    if out:
        out["fields"] = "[evals/s, p50 ms, p99 ms]"
    return out or None


def sieve_flops(reps) -> float:
    """Canonical flops of one repeated-row loop hafnian by the sieve as the
    reference kernel writes it (core/repeated.cu:86-132): ceil(T / 2) evaluated
    terms of 4m^2 + 8m + 22K + 20 real flops (sigma = x^T A x over the m nonzero
    counts, mu, the mu-power ladder and the S_N loop; K = N / 2)."""
    n = [int(x) for x in reps if int(x) > 0]
    T = int(np.prod([x + 1 for x in n])) if n else 1
    m, K = len(n), sum(n) // 2
    return (T + 1) // 2 * (4.0 * m * m + 8.0 * m + 22.0 * K + 20.0)


def repeated_bench():
    """The reference's repeated-row sieve benchmark (bench/repeated_ab.py): base
    A = 0.3 (G + G^T), M = 6, gamma = 0, 2048 patterns (q, ..., q), q = 2..6,
    timed like the reference through the public path (H2D, launch, D2H), median
    of 5 after 2 warm-ups; the checksum is compared with the reference's."""
    import gbskernels_amd as gk
    from gbskernels_amd.workloads import repeated_ab_base
    gold = {}
    gpath = Path(__file__).resolve().parent / "tests" / "golden" / "repeated_ab_checksums.json"
    if gpath.exists():
        gold = {r["q"]: r for r in json.loads(gpath.read_text())["rows"]}
    A, gam = repeated_ab_base()
    rows = []
    for q in (2, 3, 4, 5, 6):
        reps = np.full((2048, A.shape[0]), q, np.int32)
        for _ in range(2):
            gk.lhaf_repeated(A, gam, reps[:256])
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            v = gk.lhaf_repeated(A, gam, reps)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        chk = complex(np.sum(v))
        row = {"q": q, "batch": 2048, "sieve_terms": (q + 1) ** A.shape[0], "median_ms": t * 1e3,
               "evals_per_s": 2048 / t, "tflops_canonical": 2048 * sieve_flops(reps[0]) / t / 1e12,
               "checksum": [chk.real, chk.imag]}
        if q in gold:
            ref = complex(*gold[q]["checksum"])
            row["checksum_rel_vs_reference"] = abs(chk - ref) / abs(ref)
            row["reference_public_path_evals_per_s"] = gold[q]["public_path_evals_per_s"]
        rows.append(row)
    return {"workload": "bench/repeated_ab.py: M = 6, gamma = 0, 2048 x (q,...,q), public path", "rows": rows}


def tor_single_bench():
    """One large torontonian per call, the reference's own timing workload
    (core/bench_kernels.cu:368-410): O = 0.5 (G + G^T) + 0.15 I with G = 0.02
    U(-1, 1) (numpy default_rng(4321 + n) here), n = 16, 20, 24, 26 modes and
    32 (the cap), timed through the public path, median of 7 after a warm-up;
    certified and certified-dd at n = 16, 20. The fp64 value of these inputs is
    rounding-dominated (exact ~0.18^n); the certified-dd value is the accurate one."""
    from gbskernels_amd.ext import gbskernels_ext as ext
    ref = {}
    gpath = Path(__file__).resolve().parent / "tests" / "golden" / "tor_single_reference.json"
    if gpath.exists():
        ref = {r["matrix_dim"]: r["evals_per_sec_median"] for r in json.loads(gpath.read_text())["rows"]}
    rows = []
    for n in (16, 20, 24, 26, 32):
        rng = np.random.default_rng(4321 + n)
        G = 0.02 * rng.uniform(-1.0, 1.0, (2 * n, 2 * n))
        O = 0.5 * (G + G.T) + 0.15 * np.eye(2 * n)
        tiers = [("fp64", lambda: ext.tor_single(O, 14))]
        if n <= 20:
            tiers += [("certified", lambda: ext.tor_single_certified(O, 14)),
                      ("certified_dd", lambda: ext.tor_single_ddcertified(O, 14))]
        for tier, fn in tiers:
            reps = 7 if n <= 26 else 3
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                val = fn()
                ts.append(time.perf_counter() - t0)
            t = float(np.median(ts))
            row = {"modes": n, "matrix_dim": 2 * n, "tier": tier, "median_ms": t * 1e3, "evals_per_s": 1.0 / t,
                   "tflops_canonical": tor_flops(n) / t / 1e12, "value": val}
            if tier == "fp64" and 2 * n in ref:
                row["reference_evals_per_s"] = ref[2 * n]
            rows.append(row)
    return {"workload": "core/bench_kernels.cu:368-410 inputs (numpy generator), public path", "rows": rows}


# The reference's kernel-only table (core/bench_kernels.cu, B = 4096): each of
# our tiers against the reference function family it replaces (the best of its
# per-thread / cooperative / warp / small-cap variants at that size), on the
# same input recipes (complex U(-1, 1); symmetric for haf / lhaf / tor, the
# torontonian x 0.05; certified x 0.3, symmetrized by averaging), drawn with
# numpy here. Matrix sides are the reference's `matrix_dim`.
REFERENCE_ROWS = (
    ("perm fp64", 0, 0, "plain", (8, 12, 16, 20), ("perm", "perm_coop", "perm_warp")),
    ("haf fp64", 1, 0, "sym", (8, 12, 16, 20), ("haf", "haf_small", "haf_coop")),
    ("lhaf fp64", 2, 0, "sym", (8, 12, 16), ("lhaf", "lhaf_coop")),
    ("tor fp64, complex O", 3, 0, "tor", (8, 12, 16, 20), ("tor", "tor_coop")),
    ("tor fp64, real O", 3, 0, "tor_real", (8, 12, 16), ("tor_real_chol", "tor_recursive")),
    ("perm dd", 0, 1, "plain", (8, 12, 16), ("perm_dd",)),
    ("haf dd", 1, 1, "sym", (8, 12, 16), ("haf_dd",)),
    ("lhaf dd", 2, 1, "sym", (8, 12), ("lhaf_dd",)),
    ("tor dd", 3, 1, "tor", (8, 12, 16), ("tor_dd",)),
    ("perm certified", 0, 3, "cert", (8, 12, 16), ("perm_certified",)),
    ("haf certified", 1, 3, "cert_sym", (8, 12, 16), ("haf_certified",)),
    ("lhaf certified", 2, 3, "cert_sym", (8, 12), ("lhaf_certified",)),
    ("haf certified-dd", 1, 4, "cert_sym", (8, 12), ("haf_dd_certified",)),
)


def reference_inputs(kind: str, D: int, B: int) -> np.ndarray:
    """B x D x D inputs of one recipe (numpy default_rng(1234 + D))"""
    rng = np.random.default_rng(1234 + D)
    if kind == "tor_real":
        m = 0.05 * rng.uniform(-1.0, 1.0, (B, D, D))
        iu = np.triu_indices(D, 1)
        m[:, iu[1], iu[0]] = m[:, iu[0], iu[1]]
        return m.astype(np.complex128)
    scale = {"plain": 1.0, "sym": 1.0, "tor": 0.05, "cert": 0.3, "cert_sym": 0.3}[kind]
    m = scale * (rng.uniform(-1.0, 1.0, (B, D, D)) + 1j * rng.uniform(-1.0, 1.0, (B, D, D)))
    if kind in ("sym", "tor"):  # the upper triangle copied down
        iu = np.triu_indices(D, 1)
        m[:, iu[1], iu[0]] = m[:, iu[0], iu[1]]
    elif kind == "cert_sym":
        m = 0.5 * (m + m.transpose(0, 2, 1))
    return np.ascontiguousarray(m)


def reference_table():
    """Our kernels on the reference's kernel-only benchmark (B = 4096): the
    reference's protocol (2 untimed warm-ups, wall clock around the launch and
    a device synchronize, median of 7), next to its published medians on an
    RTX 4090 and an A100 (tests/golden/kernel_throughput_reference.json)."""
    import torch
    from gbskernels_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    pub = json.loads((Path(__file__).resolve().parent / "tests" / "golden" /
                      "kernel_throughput_reference.json").read_text())["rows"]
    B, rows = 4096, []
    for label, func, tier, kind, dims, ref_funcs in REFERENCE_ROWS:
        for D in dims:
            m = reference_inputs(kind, D, B)

            def median_wall(mb):
                """the protocol on the batch mb: median wall seconds of launch + synchronize"""
                nb = mb.shape[0]
                d_in = torch.from_numpy(mb.view(np.float64)).to(dev)
                d_out = torch.empty(2 * nb, dtype=torch.float64, device=dev)
                d_aux = torch.empty(nb, dtype=torch.float64, device=dev)
                wsb = int(L.gbs_workspace_bytes(func, tier, D, nb))
                ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device=dev)

                def go():
                    rc = L.gbs_eval_device(func, tier, d_in.data_ptr(), D, nb, d_out.data_ptr(), d_aux.data_ptr(),
                                           ws.data_ptr(), ws.numel(), stream.cuda_stream)
                    if rc:
                        raise RuntimeError(f"gbs_eval_device rc={rc} ({label}, {D})")
                for _ in range(2):
                    go()
                    torch.cuda.synchronize()
                ts = []
                for _ in range(7):
                    t0 = time.perf_counter()
                    go()
                    torch.cuda.synchronize()
                    ts.append(time.perf_counter() - t0)
                fin = float(np.mean(np.isfinite(d_aux.cpu().numpy()))) if tier in (3, 4) else None
                return float(np.median(ts)), fin

            t_b, finite = median_wall(m)
            # the same entry point on one instance: the launch + synchronize floor
            # (launches and the host round trip) that the B = 4096 time includes
            t_1, _ = median_wall(np.ascontiguousarray(m[:1]))
            ours = B / t_b
            best = {}
            for r in pub:
                if r["func"] in ref_funcs and r["matrix_dim"] == D:
                    best[r["gpu"]] = max(best.get(r["gpu"], 0.0), r["evals_per_sec_median"])
            row = {"row": label, "matrix_dim": D, "batch": B, "evals_per_s": ours, "reference": best,
                   "wall_us": t_b * 1e6, "wall_us_batch1": t_1 * 1e6}
            if best:
                row["speedup_vs_best_published"] = ours / max(best.values())
            if finite is not None:
                row["certified_finite"] = finite
            rows.append(row)
    sp = [r["speedup_vs_best_published"] for r in rows if "speedup_vs_best_published" in r]
    return {"protocol": "B = 4096, 2 warm-ups, wall clock launch + synchronize, median of 7 "
                        "(core/bench_kernels.cu:58-119)", "rows": rows,
            "cells_compared": len(sp), "geomean_speedup": float(np.exp(np.mean(np.log(sp)))) if sp else None,
            "min_speedup": float(min(sp)) if sp else None}


def sampler_bench():
    """End-to-end samples/s of the resident chain-rule sampler on the reference's
    sampler-throughput states (bench/sampler_throughput.py:38-49: r ~ U[0.3, 0.6]
    per mode, Haar interferometer, default_rng(0)); median of 7 after a warm-up,
    host reduced-A computation included. The (modes, cutoff, samples) cells are
    the reference's resident rows and its repeated-row sieve rows; its published
    4090 figures for each are quoted beside."""
    from gbskernels_amd.sampling import sample_resident
    ref = {(3, 3, 2000): ("gpu-resident", 3.321e4), (2, 4, 2000): ("gpu-resident", 1.463e4),
           (5, 2, 2000): ("gpu-resident", 8024.0), (6, 6, 300): ("gpu+sieve", 135.3),
           (8, 5, 300): ("gpu+sieve", 23.88), (10, 4, 300): ("gpu+nosieve", 6.631),
           (5, 8, 300): ("gpu+sieve", 628.1)}
    rows = []
    for (m, c, n), (path, rate) in ref.items():
        g = np.random.default_rng(0)
        r = g.uniform(0.3, 0.6, m)
        z = (g.standard_normal((m, m)) + 1j * g.standard_normal((m, m))) / np.sqrt(2.0)
        U, rr = np.linalg.qr(z)
        ph = np.diagonal(rr).copy()
        U = U * (ph / np.abs(ph))
        S = np.block([[U.real, -U.imag], [U.imag, U.real]]) @ np.diag(np.concatenate([np.exp(-r), np.exp(r)]))
        cov = S @ S.T
        sample_resident(cov, n, cutoff=c, seed=0)
        ts = []
        for rep in range(7):
            t0 = time.perf_counter()
            s = sample_resident(cov, n, cutoff=c, seed=rep)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        rows.append({"modes": m, "cutoff": c, "num_samples": n, "samples_per_s": n / t, "median_ms": t * 1e3,
                     "mean_photons": float(s.sum(axis=1).mean()),
                     "reference_4090": {"path": path, "samples_per_s": rate}})
    return {"workload": "bench/sampler_throughput.py states, resident chain (sieve conditionals, Philox draws)",
            "rows": rows}


def accuracy_bench():
    """Relative error |v - exact| / |exact| of every GPU tier against the mpmath
    ref tier at 80 digits (gbskernels_amd's own ref tier: Ryser, matching
    recursion, subset determinants; the error itself is taken in mpmath, so
    double-double errors below 1e-16 show), through the public API and the
    extension module:

    * physical: configuration-style inputs (perm: rows of a Haar 64x64 unitary;
      haf / lhaf: submatrices of U diag(tanh r) U^T, lhaf diagonal CN(0, 0.3^2);
      tor: threshold O of lossy squeezed pairs), the max over K instances per
      size, sizes where mpmath finishes in seconds;
    * cancelling: value = delta x an O(1) factor with O(1) summands,
      delta = 1e-2 .. 1e-14 (workloads.cancellation), with the tier the auto
      ladder chose.
    Certified tiers: fraction of instances whose bound encloses the mpmath
    value, and the largest bound / |exact|. The reference's own published
    errors (4090, one physical instance per size) are quoted beside."""
    import mpmath as mp
    import gbskernels_amd as gk
    from gbskernels_amd import _ref
    from gbskernels_amd.ext import gbskernels_ext as ext
    from gbskernels_amd.workloads import cancellation, gbs_adjacency, haar_unitary, haf_subset, perm_rows, \
        threshold_O

    published = {("perm", 4): (1.7e-16, 4.3e-17), ("perm", 8): (2.3e-15, 7.7e-18), ("perm", 12): (7.4e-15, 3.3e-18),
                 ("haf", 4): (1.0e-16, 1.0e-16), ("haf", 8): (1.2e-14, 2.1e-15), ("haf", 12): (5.5e-15, 6.1e-17),
                 ("lhaf", 4): (1.5e-16, 7.1e-17), ("lhaf", 8): (9.3e-15, 4.5e-16),
                 ("lhaf", 10): (1.9e-14, 8.6e-17)}
    rng = np.random.default_rng(11)
    U64, A32 = haar_unitary(64, rng), gbs_adjacency(rng)
    gen = {"perm": lambda n: perm_rows(U64, n, rng), "haf": lambda n: haf_subset(A32, n, rng, False),
           "lhaf": lambda n: haf_subset(A32, n, rng, True),
           "tor": lambda n: threshold_O(n, rng).astype(np.complex128)}
    sizes = {"perm": ((4, 4), (8, 4), (12, 3)), "haf": ((4, 4), (8, 4), (12, 3)),
             "lhaf": ((4, 4), (8, 4), (10, 3), (12, 2)), "tor": ((2, 4), (4, 4), (6, 4), (8, 3))}

    def errs(vals, exact):
        with mp.workdps(80):
            return [float(abs(mp.mpc(complex(v)) - e) / abs(e)) for v, e in zip(vals, exact)]

    def tiers(func, mats, exact):
        stack = np.ascontiguousarray(np.stack(mats).astype(np.complex128))
        row = {}
        for prec in ("fp64", "dd"):
            v = getattr(gk, f"{func}_batched")(stack, precision=prec, backend="gpu")
            row[f"{prec}_max_rel"] = max(errs(v, exact))
        v, diags = getattr(gk, f"{func}_batched")(stack, precision="auto", backend="gpu", return_diagnostics=True)
        row["auto_max_rel"] = max(errs(v, exact))
        row["auto_tiers"] = sorted({d["tier"] for d in diags})
        for name in ("certified", "dd_certified"):
            v, b = getattr(ext, f"{func}_{name}")(stack)
            with mp.workdps(80):
                enc = [bool(np.isfinite(bb) and abs(mp.mpc(complex(vv)) - e) <= mp.mpf(float(bb)))
                       for vv, bb, e in zip(v, b, exact)]
                width = max(float(mp.mpf(float(bb)) / abs(e)) for bb, e in zip(b, exact))
            row[f"{name}_enclosed"] = f"{sum(enc)}/{len(enc)}"
            row[f"{name}_max_bound_rel"] = width
        return row

    physical = []
    for func, rows in sizes.items():
        for n, K in rows:
            mats = [gen[func](n) for _ in range(K)]
            exact = [_ref.REF[func](m, dps=80) for m in mats]
            row = {"func": func, "n" if func != "tor" else "modes": n, "instances": K}
            row.update(tiers(func, mats, exact))
            if (func, n) in published:
                row["reference_4090_fp64_dd"] = list(published[(func, n)])
            physical.append(row)
    cancelling = []
    for func in ("perm", "haf", "lhaf", "tor"):
        for delta in (1e-2, 1e-6, 1e-10, 1e-14):
            mats = [cancellation(func, delta, rng) for _ in range(2)]
            exact = [_ref.REF[func](m, dps=80) for m in mats]
            row = {"func": func, "delta": delta}
            row.update(tiers(func, mats, exact))
            cancelling.append(row)
    return {"dps": 80, "oracle": "gbskernels_amd ref tier (mpmath)", "physical": physical, "cancelling": cancelling}


class Cell:
    """One timed kernel cell: ``go`` launches it once on ``stream``; ``meta``
    is reported beside the timing; ``post(median_s)`` adds derived fields."""

    def __init__(self, name, go, meta, post=None):
        self.name, self.go, self.meta, self.post = name, go, meta, post


def run_cells(cells, torch, stream, reps, seed):
    """The reference's benchmark hygiene (docs/benchmark_protocol.md:28-36): each
    cell is warmed up once, then timed ``reps`` times with HIP events on the
    launch stream; the cells run in a seeded random order; median and IQR are
    reported."""
    order = np.random.default_rng(seed).permutation(len(cells))
    out = {}
    for i in order:
        c = cells[i]
        c.go()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            c.go()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        med = float(np.median(ms))
        rec = dict(c.meta)
        rec.update({"median_ms": med, "iqr_ms": [float(np.percentile(ms, 25)), float(np.percentile(ms, 75))],
                    "reps": reps, "evals_per_s": c.meta["batch"] / (med * 1e-3)})
        if c.post is not None:
            rec.update(c.post(med * 1e-3))
        rec.update(executed(c.name, c.meta["batch"], med * 1e-3))
        out[c.name] = rec
    return {k: out[k] for k in sorted(out)}


def device_cell(L, torch, dev, stream, name, func, tier, m, meta_extra=None, flops=None):
    """A cell that evaluates the stack ``m`` (B, d, d) with ``gbs_eval_device``
    from HBM-resident buffers; ``flops`` (per evaluation) gives canonical
    TFLOP/s; certified tiers report the finite fraction of their bounds."""
    B, dim = m.shape[0], m.shape[1]
    d_in = torch.from_numpy(np.ascontiguousarray(m).astype(np.complex128).view(np.float64)).to(dev)
    d_out = torch.empty(2 * B, dtype=torch.float64, device=dev)
    d_aux = torch.empty(B, dtype=torch.float64, device=dev)
    wsb = int(L.gbs_workspace_bytes(func, tier, dim, B))
    ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device=dev)

    def go():
        rc = L.gbs_eval_device(func, tier, d_in.data_ptr(), dim, B, d_out.data_ptr(), d_aux.data_ptr(),
                               ws.data_ptr(), ws.numel(), stream.cuda_stream)
        if rc:
            raise RuntimeError(f"gbs_eval_device rc={rc} ({name})")

    def post(t):
        r = {}
        if flops is not None:
            r["tflops_canonical"] = flops * B / t / 1e12
        if tier in (3, 4):
            r["certified_finite"] = float(np.mean(np.isfinite(d_aux.cpu().numpy())))
        return r
    meta = {"batch": B}
    meta.update(meta_extra or {})
    return Cell(name, go, meta, post)


def perm_tier_cells(L, torch, dev, stream, headline=False):
    """The permanent's other tiers (fp64 and dd are the headline's): kappa,
    certified and certified-dd kernels at n = 16 (B = 4096), 20 (B = 512) and
    24 (B = 64), config-2 style rows of a Haar unitary."""
    from gbskernels_amd.workloads import haar_unitary, perm_rows
    from gbskernels_amd.workloads import perm_haar_rows_workload
    cells = []
    if headline:  # the headline's two launches as cells of their own (profiler runs)
        m16, m24 = perm_haar_rows_workload()
        cells += [device_cell(L, torch, dev, stream, "perm_fp64_n24", 0, 0, m24, flops=perm_flops(24)),
                  device_cell(L, torch, dev, stream, "perm_fp64_n16", 0, 0, m16, flops=perm_flops(16))]
    for n, B in ((16, 4096), (20, 512), (24, 64)):
        rng = np.random.default_rng(40 + n)
        U = haar_unitary(64, rng)
        m = np.stack([perm_rows(U, n, rng) for _ in range(B)])
        for tname, tier in (("kappa", 2), ("certified", 3), ("dd_certified", 4)):
            cells.append(device_cell(L, torch, dev, stream, f"perm_{tname}_n{n}", 0, tier, m,
                                     flops=perm_flops(n) * (10 if tier == 4 else 1)))
    return cells


def haf_cells(L, torch, dev, stream):
    """Uniform haf / lhaf buckets at N = 16 (B = 4096) and N = 24 (B = 256),
    config-3 style complex symmetric submatrices, fp64 and dd (canonical flops
    x10 for dd)."""
    from gbskernels_amd.workloads import gbs_adjacency, haf_subset
    A32 = gbs_adjacency(np.random.default_rng(3))
    rng = np.random.default_rng(7)
    cells = []
    for func, code in (("haf", 1), ("lhaf", 2)):
        for N, B in ((16, 4096), (24, 256)):
            m = np.stack([haf_subset(A32, N, rng, func == "lhaf") for _ in range(B)])
            for tier, tname in ((0, "fp64"), (1, "dd")):
                cells.append(device_cell(L, torch, dev, stream, f"{func}_{tname}_n{N}", code, tier, m,
                                         flops=haf_flops(N, func == "lhaf") * (10 if tier else 1)))
    return cells


def tor_flops(nm: int) -> float:
    """canonical recursive-Cholesky torontonian flops: appending the two rows
    (x, p) of the k-th included mode costs sum_{c <= r} (2c + 2) over rows
    r = 2k - 2, 2k - 1, i.e. 8 k^2, plus 4 per subset and 3 per leaf"""
    from math import comb
    return float(sum(comb(nm, k) * (8 * k * k + 4) for k in range(1, nm + 1)) + 3 * 2 ** nm)


def tor_cells(L, torch, dev, stream):
    """Torontonian of threshold-detector O matrices (two-mode squeezed pairs,
    r ~ U[0.3, 1.0], Haar interferometer, loss 0.7), 4096 per mode count
    (512 distinct, tiled; dd and certified-dd at 20 modes: 512), fp64 /
    certified / dd / certified-dd; plus the hafnian of config 1: 4096 real symmetric 8x8
    matrices G + G^T, default_rng(0)."""
    from gbskernels_amd.workloads import threshold_O
    cells = []
    for nm in (8, 12, 16, 20):
        rng = np.random.default_rng(4000 + nm)
        uniq = np.stack([threshold_O(nm, rng) for _ in range(512)])
        m = np.ascontiguousarray(np.tile(uniq, (8, 1, 1)))
        for tier, tname in ((0, "fp64"), (3, "certified"), (1, "dd"), (4, "dd_certified")):
            B = m.shape[0] if (tier not in (1, 4) or nm <= 16) else 512
            cells.append(device_cell(L, torch, dev, stream, f"tor_{tname}_nm{nm}", 3, tier, m[:B],
                                     flops=tor_flops(nm) * (10 if tier in (1, 4) else 1)))
    G = np.random.default_rng(0).standard_normal((4096, 8, 8))
    cells.append(device_cell(L, torch, dev, stream, "haf_fp64_config1_n8", 1, 0, G + G.transpose(0, 2, 1),
                             flops=haf_flops(8)))
    return cells


def served_bench(args):
    """Served mixed-size perm/haf/lhaf (n <= 24, auto tier: fp64 with dd
    re-evaluation when kappa >= 1e8) three ways on the same request set:

    * plain: the per-size launch path -- windows of K requests in arrival
      order, one synchronous call per (func, dim) bucket, then one per bucket of
      escalated items (what Workspace does per bucket), inputs pre-gathered per
      bucket outside the timed region;
    * served, saturated: every request offered at t = 0 to the continuous-
      batching server (max K per batch, 4 slots x 4 streams);
    * served, Poisson 50k req/s: the open-loop stream; p50 / p99 of
      completion - arrival.
    Values must agree bit for bit between the paths."""
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import Batch, Server, pack
    from gbskernels_amd.workloads import served_mix

    n_req, K = args.served_requests, args.served_batch
    print(f"bench: served mix {n_req} requests", file=sys.stderr, flush=True)
    items, arrivals = served_mix(n_req, seed=5)
    func, tier, dim, flat, off = pack(items)
    KAPPA, DD = _lib.TIERS["kappa"], _lib.TIERS["dd"]

    def sub(ix, t):
        ix = np.asarray(ix)
        d = dim[ix]
        m2 = d.astype(np.int64) ** 2
        o = np.concatenate([[0], np.cumsum(m2)[:-1]]).astype(np.int64)
        fl = np.concatenate([flat[off[i]:off[i] + m2[j]] for j, i in enumerate(ix)])
        return (np.ascontiguousarray(func[ix]), np.full(len(ix), t, np.int32), np.ascontiguousarray(d),
                np.ascontiguousarray(fl), o)

    # plain per-size path
    windows = []
    for w0 in range(0, n_req, K):
        groups = {}
        for i in range(w0, min(n_req, w0 + K)):
            groups.setdefault((int(func[i]), int(dim[i])), []).append(i)
        windows.append([(np.array(ix), sub(ix, KAPPA)) for ix in groups.values()])
    plain = np.zeros(n_req, np.complex128)
    dd_cap = {f: _lib.lib().gbs_max_dim(f, DD) for f in range(4)}
    escal = 0
    reps = 3  # each path is timed reps times; the median is reported
    plain_times = []
    with Batch(nstreams=1, use_graphs=False) as b:
        b.run_packed(*windows[0][0][1])  # warm
        for _ in range(reps):
            escal = 0
            t0 = time.perf_counter()
            for groups in windows:
                for ix, packed in groups:
                    v, a = b.run_packed(*packed)
                    kap = a / np.maximum(np.abs(v), 1e-300)
                    risky = (kap >= 1e8) & (dim[ix] <= dd_cap[int(func[ix[0]])])
                    if risky.any():
                        v2, _ = b.run_packed(*sub(ix[risky], DD))
                        v = v.copy()
                        v[risky] = v2
                        escal += int(risky.sum())
                    plain[ix] = v
            plain_times.append(time.perf_counter() - t0)
    t_plain = float(np.median(plain_times))
    nb = sum(len(g) for g in windows)

    auto = np.full(n_req, _lib.GBSK_AUTO, np.int32)

    def serve(s, arr_s):
        m = len(arr_s)
        st0 = s.stats()
        vals, aux, infos = s.replay_packed(func[:m], auto[:m], dim[:m], flat, off[:m], arr_s)
        st = {k: v - st0[k] for k, v in s.stats().items()}
        st["max_in_flight"] = s.stats()["max_in_flight"]
        t_arr = np.array([i.arrival_ns for i in infos], np.int64)
        t_done = np.array([i.done_ns for i in infos], np.int64)
        t0 = t_arr[0] - int(round(arr_s[0] * 1e9))
        lat = (t_done - t_arr) * 1e-9
        return vals, {"evals_per_s": m / ((t_done.max() - t0) * 1e-9),
                      "p50_ms": float(np.percentile(lat, 50)) * 1e3, "p99_ms": float(np.percentile(lat, 99)) * 1e3,
                      "batches": st["batches"], "escalated": st["escalated"], "max_in_flight": st["max_in_flight"]}

    # one long-lived server, as in service: the untimed first pass loads the
    # kernels and sizes the slots' buffers
    with Server(slots=4, streams_per_slot=4, max_batch=K) as s:
        serve(s, np.zeros(n_req))
        sats = [serve(s, np.zeros(n_req)) for _ in range(reps)]
        poi_vals, poi = serve(s, arrivals)
    rates = [x[1]["evals_per_s"] for x in sats]
    sat_vals, sat = sats[int(np.argsort(rates)[len(rates) // 2])]  # the median run
    sat["evals_per_s_runs"] = rates
    same = lambda x, y: bool(np.array_equal(x.view(np.uint64), y.view(np.uint64)))
    plain_rate = n_req / t_plain
    sizes = np.bincount(dim, minlength=25)
    return {
        "workload": f"{n_req} requests, func uniform perm/haf/lhaf, perm n = 1 + Geom(0.15) on 2..24 (rows of a "
                    "Haar 64x64 unitary), haf/lhaf N = 2k, k ~ Geom(0.15) on 1..12 (submatrices of "
                    "U diag(tanh r) U^T, U Haar 32x32; lhaf diagonal CN(0, 0.3^2)); tier auto (fp64, dd when "
                    f"kappa >= 1e8); windows / max batch {K}",
        "plain_per_size": {"evals_per_s": plain_rate, "bucket_calls": nb, "escalated": escal,
                           "evals_per_s_runs": [n_req / t for t in plain_times]},
        "served_saturated": sat,
        "served_poisson_50k": poi,
        "speedup_saturated_vs_plain": sat["evals_per_s"] / plain_rate,
        "identical_values": same(sat_vals, plain) and same(poi_vals, plain)
        and all(same(x[0], plain) for x in sats),
        "size_histogram": {int(k): int(v) for k, v in enumerate(sizes) if v},
    }


def config_streams(args):
    """The ragged-hafnian and served-sampler configurations through the server:

    * ragged haf + lhaf: haf / lhaf uniform, N = 2k with k ~ Geom(0.15) on 1..12,
      submatrices of U diag(tanh r) U^T (U Haar 32x32, r ~ U[0.5, 1.2]), lhaf
      diagonal CN(0, 0.3^2), tier auto (fp64, dd when kappa >= 1e8), all offered
      at once (throughput);
    * served sampler stream: perm / haf / lhaf / tor uniform, sizes geometric(0.15)
      on 2..28 (tor: mode count on 1..14), tiers 70% fp64 / 25% dd / 5% certified,
      Poisson arrivals at 50k req/s (latency)."""
    from gbskernels_amd import _lib
    from gbskernels_amd.serving import Server, pack
    from gbskernels_amd.workloads import served_mix

    def run(items, arr, K):
        f, t, d, flat, off = pack(items)
        with Server(slots=4, streams_per_slot=4, max_batch=K) as s:
            s.replay_packed(f, t, d, flat, off, np.zeros(len(f)))  # untimed: kernels, slot buffers
            st0 = s.stats()
            vals, aux, infos = s.replay_packed(f, t, d, flat, off, arr)
            st = {k: v - st0[k] for k, v in s.stats().items()}
        t_arr = np.array([i.arrival_ns for i in infos], np.int64)
        t_done = np.array([i.done_ns for i in infos], np.int64)
        t0 = t_arr[0] - int(round(arr[0] * 1e9))
        lat = (t_done - t_arr) * 1e-9
        fin = np.isfinite(aux[t == _lib.TIERS["certified"]]) if np.any(t == _lib.TIERS["certified"]) else None
        out = {"requests": len(items), "evals_per_s": len(items) / ((t_done.max() - t0) * 1e-9),
               "p50_ms": float(np.percentile(lat, 50)) * 1e3, "p99_ms": float(np.percentile(lat, 99)) * 1e3,
               "batches": st["batches"], "escalated": st["escalated"]}
        if fin is not None:
            out["certified_finite"] = float(np.mean(fin))
        return out

    res = {}
    n3 = args.config3_requests
    if n3 > 0:
        print(f"bench: ragged haf/lhaf {n3} requests", file=sys.stderr, flush=True)
        items, _ = served_mix(n3, seed=3, funcs=("haf", "lhaf"), nmax=24)
        res["ragged_haf_lhaf_auto"] = run(items, np.zeros(n3), 4096)
    n5 = args.config5_requests
    if n5 > 0:
        print(f"bench: sampler stream {n5} requests", file=sys.stderr, flush=True)
        items, arr = served_mix(n5, seed=5, funcs=("perm", "haf", "lhaf", "tor"), nmax=28,
                                tiers=(("fp64", 0.70), ("dd", 0.25), ("certified", 0.05)))
        res["sampler_stream_poisson_50k"] = run(items, arr, 4096)
    return res


def cpu_baseline(m16, m24, budget_s, out16, out24, dd16, dd24):
    """Oracle Glynn (fp64) on the host cores, OpenMP over instances, on a bounded
    sample; extrapolated to the per-step mix (65536 n=16 + 1024 n=24). Also the
    GPU-vs-CPU agreement gate: the oracle's double-double Glynn walk on a sample
    of the same inputs against the GPU's fp64 and dd outputs. Returns
    (baseline, agreement)."""
    from oracle import oracle as O
    s16 = np.arange(0, 65536, 4099)[:16]
    ref16 = O.perm_dd(m16[s16])
    ref24 = O.perm_dd(m24[:2])
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    parity = {"n16_fp64_vs_oracle_dd": rel(out16[s16], ref16), "n24_fp64_vs_oracle_dd": rel(out24[:2], ref24),
              "samples": [len(s16), 2]}
    if dd16 is not None:
        parity["n16_dd_vs_oracle_dd"] = rel(dd16[s16], ref16)
        parity["n24_dd_vs_oracle_dd"] = rel(dd24[:2], ref24)
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))

    def per_eval(threads, k16, budget24):
        """per-evaluation seconds at n = 16 and n = 24 on ``threads`` threads"""
        O.set_threads(threads)
        O.perm(m16[:threads])  # warm
        t = time.perf_counter()
        O.perm(m16[:k16])
        t16 = (time.perf_counter() - t) / k16
        # n = 24: size the sample to the budget, at least one eval per thread
        est24 = t16 * 2 ** 8 * 1.5
        k24 = int(max(threads, min(256, budget24 / max(est24, 1e-9))))
        k24 = max(threads, (k24 // threads) * threads)
        t = time.perf_counter()
        O.perm(m24[:k24])
        t24 = (time.perf_counter() - t) / k24
        return t16, t24, k24

    t16, t24, k24 = per_eval(cores, 512, budget_s * 0.6)
    s16, s24, j24 = per_eval(1, 64, budget_s * 0.2)
    O.set_threads(cores)
    mix = lambda a, b: (65536 + 1024) / (65536 * a + 1024 * b)
    return {"value": mix(t16, t24), "unit": "evals/s", "cores": cores, "kind": "port",
            "value_1core": mix(s16, s24),
            "sample_short": f"oracle Glynn fp64, {cores} thr: {512} n16 + {k24} n24; 1 thr: 64 n16 + {j24} n24",
            "sample": f"oracle Glynn fp64 (cpu_ref restated in C++), {512} n=16 + {k24} n=24 evals "
                      f"of the same workload, OpenMP {cores} threads, per-eval {t16 * 1e3:.3f} ms (n=16), "
                      f"{t24:.3f} s (n=24); 1 thread: 64 n=16 + {j24} n=24, per-eval {s16 * 1e3:.3f} ms, "
                      f"{s24:.3f} s; both extrapolated to the step mix"}, parity


if __name__ == "__main__":
    sys.exit(main())
