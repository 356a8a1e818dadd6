"""Autotuner of the kernels' work splits and schedules.

``--perm`` (default): the permanent kernels' Gray-code chunk shape (below).
``--tor``: the batched torontonian's schedule per tier family -- up to which
mode count the one-instance-per-wave kernel walks its middle tree breadth-first
(values identical to the depth-first walk) -- and, reported only, the grouped
small-count kernel (a different split: it would move values' last bits, so the
built-in choice changes only with ``--allow-split-change``).
``--haf``: the hafnian's fused small-size kernel -- up to which matrix side it
replaces the per-subset-size launches, and how many workgroups it keeps per
launch (values identical either way).

Permanent chunk shape:

One n x n permanent walks 2^(n-1) Gray terms, cut into LPM lanes (fixed per n)
x W workgroups x 2^k terms per lane. The table ``csrc/perm_tune.hpp`` holds k per
matrix side and tier family (fp64 / kappa / certified share one entry, dd /
certified-dd another, so that the values of a family stay bit-identical); it
replaces the reference's fixed crossovers (``core/host_api.cu:65-72``:
PERM_COOP_CROSSOVER, PERM_COOP_GROUPS). This module measures every candidate k
on the current GPU at two batches, because the split may depend on n only
(never on the batch: a value is the same alone and in any mix):

* a throughput batch that keeps the device busy (65536 matrices at n = 16 down
  to 64 at n = 28 for fp64; an eighth of that for dd);
* one matrix alone, the served case (a request stream rarely holds two large
  permanents of one size in a batch), where large chunks leave the device idle.

Lone-matrix times are compared above a floor ``--latency-ms`` (default
0.25 ms; shorter ones count as equal: a served batch takes milliseconds). A
candidate is admissible when its lone-matrix time stays within 1.25x the
built-in default's (2^9 terms per lane for fp64, 2^6 for dd): tuning may not
make a lone request noticeably slower. Among the admissible candidates whose
batch time is within ``--margin`` (default 2%) of the best, the one with the
fastest lone matrix wins (then the fastest batch); the current entry stays
unless it is outside that band or its lone-matrix time is 1.25x the winner's,
so that timing noise does not churn the table (and the last bits of the values,
which depend on the split).

    python -m gbskernels_amd.autotune              # measure and print
    python -m gbskernels_amd.autotune --write      # ... and rewrite perm_tune.hpp
    python -m gbskernels_amd.autotune --write out.hpp   # ... or write the table elsewhere
    python -m gbskernels_amd.autotune --tor [--write]   # torontonian schedule (tor.hip constants)
    python -m gbskernels_amd.autotune --haf [--write]   # fused small hafnian (haf.hip constants)

Rebuild the library afterwards (``__graft_entry__.build()`` or ``make -C
gbskernels_amd/csrc``) for the new table to be the built-in one. Needs a GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import re
import sys
from pathlib import Path

import numpy as np

TUNE_HEADER = Path(__file__).resolve().parent / "csrc" / "perm_tune.hpp"
TOR_SOURCE = Path(__file__).resolve().parent / "csrc" / "tor.hip"
HAF_SOURCE = Path(__file__).resolve().parent / "csrc" / "haf.hip"
FAMILIES = {0: ("fp64", 0), 1: ("dd", 1)}  # family -> (name, tier code for gbs_eval_device)
LOGMIN = {0: 4, 1: 2}
DEFAULT = {0: 9, 1: 6}  # the untuned entries (log2 Gray terms per lane)


def batch_for(n: int, family: int) -> int:
    b = 2 ** (34 - n) if family == 0 else 2 ** (31 - n)
    return int(min(65536, max(64 if family == 0 else 16, b)))


def _split(L, family: int, n: int) -> tuple[int, int]:
    k, w = ctypes.c_int(), ctypes.c_int()
    rc = L.gbs_perm_split(family, n, ctypes.byref(k), ctypes.byref(w))
    if rc:
        raise RuntimeError(f"gbs_perm_split rc={rc}")
    return k.value, w.value


def _entry(L, family: int, n: int) -> int:
    v = ctypes.c_int()
    rc = L.gbs_perm_get_lane_log2(family, n, ctypes.byref(v))
    if rc:
        raise RuntimeError(f"gbs_perm_get_lane_log2 rc={rc}")
    return v.value


def _set(L, family: int, n: int, k: int) -> None:
    rc = L.gbs_perm_set_lane_log2(family, n, k)
    if rc:
        raise RuntimeError(f"gbs_perm_set_lane_log2({family}, {n}, {k}) rc={rc}")


def candidates(L, family: int, n: int, max_split_log2: int = 12) -> tuple[int, list[int]]:
    """(one-workgroup limit, candidate k values) for side n: the limit down to
    2^max_split_log2 workgroups per matrix, the current entry and the default."""
    old = _entry(L, family, n)
    _set(L, family, n, 27)
    lim = _split(L, family, n)[0]
    _set(L, family, n, old)
    if lim < LOGMIN[family]:
        return lim, []
    lo = max(LOGMIN[family], lim - max_split_log2)
    return lim, sorted(set(range(lo, lim + 1)) | {min(old, lim), min(DEFAULT[family], lim)})


def measure(L, torch, family: int, n: int, ks: list[int], reps: int, B: int) -> dict[int, float]:
    """median kernel ms of one launch of B n x n matrices per candidate k"""
    from .workloads import haar_unitary, perm_rows
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(100 + n)
    U = haar_unitary(64, rng)
    mats = np.stack([perm_rows(U, n, rng) for _ in range(B)])
    d_in = torch.from_numpy(mats.view(np.float64)).to(dev)
    d_out = torch.empty(2 * B, dtype=torch.float64, device=dev)
    tier = FAMILIES[family][1]
    old = _entry(L, family, n)
    times = {}
    try:
        for k in ks:
            _set(L, family, n, k)
            wsb = int(L.gbs_workspace_bytes(0, tier, n, B))
            ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device=dev)

            def go():
                rc = L.gbs_eval_device(0, tier, d_in.data_ptr(), n, B, d_out.data_ptr(), None, ws.data_ptr(),
                                       ws.numel(), stream.cuda_stream)
                if rc:
                    raise RuntimeError(f"gbs_eval_device rc={rc}")
            go()
            torch.cuda.synchronize()
            ms = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                go()
                e1.record(stream)
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
            times[k] = float(np.median(ms))
            del ws
    finally:
        _set(L, family, n, old)
    return times


def choose(tb: dict[int, float], t1: dict[int, float], cur: int, ref: int, latency_ms: float,
           margin: float) -> int:
    """batch-time band of the admissible candidates, fastest lone matrix in it
    (see the module text); ref = the default entry, cur = the current one"""
    lat = {k: max(t, latency_ms) for k, t in t1.items()}  # below the floor: a tie
    cap = 1.25 * lat[ref]
    ok = [k for k in tb if lat[k] <= cap]
    bmin = min(tb[k] for k in ok)
    near = [k for k in ok if tb[k] <= (1.0 + margin) * bmin]
    pick = min(near, key=lambda k: (lat[k], tb[k]))
    if cur in near and lat[cur] <= 1.25 * lat[pick]:
        return cur
    return pick


def write_table(table: dict[int, dict[int, int]], path: Path = TUNE_HEADER, out: Path | None = None) -> None:
    """Rewrite the two rows of kPermLaneLog2 of perm_tune.hpp (other text kept),
    in place or into `out`."""
    text = path.read_text()
    rows = re.findall(r"\{([0-9, ]+)\}", text)
    if len(rows) != 2:
        raise RuntimeError(f"{path}: expected two table rows")
    for family, row in ((0, rows[0]), (1, rows[1])):
        vals = [int(x) for x in row.split(",")]
        for n, k in table.get(family, {}).items():
            vals[n] = k
        text = text.replace("{" + row + "}", "{" + ", ".join(str(v) for v in vals) + "}", 1)
    (out or path).write_text(text)


def _median_ms(torch, stream, go, reps: int) -> float:
    go()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        go()
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return float(np.median(ms))


def _device_call(L, torch, func: int, tier: int, mats: np.ndarray):
    """a closure evaluating the resident stack `mats` with gbs_eval_device"""
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    B, d = mats.shape[0], mats.shape[1]
    d_in = torch.from_numpy(np.ascontiguousarray(mats, dtype=np.complex128).view(np.float64)).to(dev)
    d_out = torch.empty(2 * B, dtype=torch.float64, device=dev)
    d_aux = torch.empty(B, dtype=torch.float64, device=dev)
    wsb = int(L.gbs_workspace_bytes(func, tier, d, B))
    ws = torch.empty(max(wsb, 256), dtype=torch.uint8, device=dev)

    def go():
        rc = L.gbs_eval_device(func, tier, d_in.data_ptr(), d, B, d_out.data_ptr(), d_aux.data_ptr(), ws.data_ptr(),
                               ws.numel(), stream.cuda_stream)
        if rc:
            raise RuntimeError(f"gbs_eval_device rc={rc}")
    return go, stream


def _rewrite_constant(path: Path, name: str, value: int, out: Path | None = None) -> None:
    text = (out or path).read_text() if out and out.exists() else path.read_text()
    new, n = re.subn(rf"(\b{name}\s*=\s*)\d+", rf"\g<1>{value}", text, count=1)
    if n != 1:
        raise RuntimeError(f"{path}: constant {name} not found")
    (out or path).write_text(new)


def tune_tor(L, torch, reps: int, margin: float, allow_split_change: bool, write, sizes=None) -> None:
    """Breadth-first vs depth-first middle walk per family and mode count (4096
    physical instances, fewer at 16+ modes); bfs_max = the largest mode count
    up to which breadth-first is never slower than (1 + margin) x depth-first.
    The grouped kernel is timed beside them and adopted only with
    allow_split_change (it changes the split)."""
    import ctypes
    from .workloads import threshold_O
    sizes = sizes or [4, 6, 8, 10, 12, 14, 16, 18, 20]
    saved = []
    for fam in (0, 1):
        g, b = ctypes.c_int(), ctypes.c_int()
        L.gbs_tor_get_schedule(fam, ctypes.byref(g), ctypes.byref(b))
        saved.append((g.value, b.value))
    chosen = {}
    print("| family | modes | batch | depth-first ms | breadth-first ms | grouped ms |")
    print("|---|---|---|---|---|---|")
    try:
        for fam, tier in ((0, 0), (1, 1)):
            bfs_max, grouped_max, still = 0, 0, True
            for nm in sizes:
                B = 4096 if nm <= 12 else (1024 if nm <= 16 else 256)
                rng = np.random.default_rng(900 + nm)
                O = np.stack([threshold_O(nm, rng) for _ in range(B)]).astype(np.complex128)
                go, stream = _device_call(L, torch, 3, tier, O)
                t = {}
                for label, gm, bm in (("dfs", 0, 0), ("bfs", 0, 32), ("grouped", 10, 32)):
                    L.gbs_tor_set_schedule(fam, gm, bm)
                    t[label] = _median_ms(torch, stream, go, reps)
                print(f"| {FAMILIES[fam][0]} | {nm} | {B} | {t['dfs']:.3f} | {t['bfs']:.3f} | {t['grouped']:.3f} |",
                      flush=True)
                if still and t["bfs"] <= (1.0 + margin) * t["dfs"]:
                    bfs_max = nm
                else:
                    still = False
                if nm <= 10 and t["grouped"] < (1.0 - margin) * min(t["bfs"], t["dfs"]) and grouped_max == nm - 2:
                    grouped_max = nm
            chosen[fam] = (grouped_max if allow_split_change else saved[fam][0], bfs_max)
    finally:
        for fam, (g, b) in enumerate(saved):
            L.gbs_tor_set_schedule(fam, g, b)
    print(f"chosen (grouped_max, bfs_max) per family: {chosen}", file=sys.stderr)
    if write is not None:
        out = Path(write) if write else None
        for fam, suffix in ((0, "F64"), (1, "DD")):
            _rewrite_constant(TOR_SOURCE, f"TOR_GROUPED_MAX_{suffix}", chosen[fam][0], out)
            _rewrite_constant(TOR_SOURCE, f"TOR_BFS_MAX_{suffix}", chosen[fam][1], out)
        print(f"wrote {out or TOR_SOURCE}", file=sys.stderr)


def tune_haf(L, torch, reps: int, margin: float, write) -> None:
    """Fused small-size kernel against the per-subset-size launches for each
    side N <= 12 (4096 instances, fp64): small_max_n = the largest side up to
    which the fused kernel is never slower than (1 + margin) x the other path;
    then the workgroup target at N = 8 (0: the built-in per-bucket cost model)."""
    import ctypes
    from .workloads import gbs_adjacency, haf_subset
    m0, t0 = ctypes.c_int(), ctypes.c_int()
    L.gbs_haf_get_schedule(ctypes.byref(m0), ctypes.byref(t0))
    A = gbs_adjacency(np.random.default_rng(3))
    rng = np.random.default_rng(11)
    small_max, still = 0, True
    print("| N | batch | per-size launches ms | fused ms |")
    print("|---|---|---|---|")
    try:
        for N in (2, 4, 6, 8, 10, 12):
            mats = np.stack([haf_subset(A, N, rng, False) for _ in range(4096)])
            go, stream = _device_call(L, torch, 1, 0, mats)
            L.gbs_haf_set_schedule(0, t0.value)
            tp = _median_ms(torch, stream, go, reps)
            L.gbs_haf_set_schedule(12, t0.value)
            tf = _median_ms(torch, stream, go, reps)
            print(f"| {N} | 4096 | {tp:.4f} | {tf:.4f} |", flush=True)
            if still and tf <= (1.0 + margin) * tp:
                small_max = N
            else:
                still = False
        mats = np.stack([haf_subset(A, 8, rng, False) for _ in range(4096)])
        go, stream = _device_call(L, torch, 1, 0, mats)
        best = None
        print("| workgroup target | N = 8 fused ms |")
        print("|---|---|")
        for tgt in (0, 256, 512, 1024, 2048, 4096):
            L.gbs_haf_set_schedule(12, tgt)
            t = _median_ms(torch, stream, go, reps)
            print(f"| {tgt} | {t:.4f} |", flush=True)
            if best is None or t < best[1]:
                best = (tgt, t)
    finally:
        L.gbs_haf_set_schedule(m0.value, t0.value)
    print(f"chosen small_max_n {small_max}, small_target_wgs {best[0]}", file=sys.stderr)
    if write is not None:
        out = Path(write) if write else None
        _rewrite_constant(HAF_SOURCE, "haf_small_max_n", small_max, out)
        _rewrite_constant(HAF_SOURCE, "haf_small_target_wgs", best[0], out)
        print(f"wrote {out or HAF_SOURCE}", file=sys.stderr)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--families", default="0,1", help="0 = fp64 family, 1 = dd family")
    ap.add_argument("--sizes", default="10-28", help="matrix sides, e.g. 16-28 or 16,20,24")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.02, help="required gain over the current entry")
    ap.add_argument("--latency-ms", type=float, default=0.25,
                    help="one-matrix time any choice may take (or 1.25x the current entry's, if larger)")
    ap.add_argument("--write", nargs="?", const="", default=None, metavar="PATH",
                    help="write the table with the chosen entries (default: rewrite csrc/perm_tune.hpp)")
    ap.add_argument("--tor", action="store_true", help="tune the torontonian schedule instead")
    ap.add_argument("--haf", action="store_true", help="tune the fused small-size hafnian instead")
    ap.add_argument("--allow-split-change", action="store_true",
                    help="(--tor) let the grouped small-count kernel, a different split, become built in")
    args = ap.parse_args(argv)

    import torch
    from . import _lib
    L = _lib.lib()
    if args.tor:
        tune_tor(L, torch, max(args.reps, 5), args.margin, args.allow_split_change, args.write)
        return 0
    if args.haf:
        tune_haf(L, torch, max(args.reps, 10), args.margin, args.write)
        return 0
    if "-" in args.sizes:
        lo, hi = (int(x) for x in args.sizes.split("-"))
        sizes = list(range(lo, hi + 1))
    else:
        sizes = [int(x) for x in args.sizes.split(",")]
    chosen: dict[int, dict[int, int]] = {}
    print("| family | n | batch | entry | limit | k: batch ms / one-matrix ms (workgroups per matrix) | chosen "
          "| batch gain | one-matrix gain |")
    print("|---|---|---|---|---|---|---|---|---|")
    for family in (int(f) for f in args.families.split(",")):
        for n in sizes:
            lim, ks = candidates(L, family, n)
            if len(ks) < 2:
                continue
            entry = _entry(L, family, n)
            cur, ref = min(entry, lim), min(DEFAULT[family], lim)
            # at n >= 10 a wave holds one matrix: skip splits that leave the
            # device with less than one wave per SIMD at the batch (they cannot win)
            B = batch_for(n, family)
            ks = [k for k in ks if k in (cur, ref) or B * 2 ** (lim - k) >= 1024]
            tb = measure(L, torch, family, n, ks, args.reps, B)
            t1 = measure(L, torch, family, n, ks, args.reps, 1)
            pick = choose(tb, t1, cur, ref, args.latency_ms, args.margin)
            chosen.setdefault(family, {})[n] = entry if pick == cur else pick
            cells = ", ".join(f"{k}: {tb[k]:.3f} / {t1[k]:.3f} ({2 ** (lim - k)})" for k in ks)
            print(f"| {FAMILIES[family][0]} | {n} | {B} | {cur} | {lim} | {cells} | {pick} | "
                  f"{tb[cur] / tb[pick]:.3f}x | {t1[cur] / t1[pick]:.3f}x |", flush=True)
    if args.write is not None:
        out = Path(args.write) if args.write else None
        write_table(chosen, out=out)
        print(f"wrote {out or TUNE_HEADER}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
