"""Autotuner of the permanent kernels' Gray-code chunk shape.

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
    args = ap.parse_args(argv)

    import torch
    from . import _lib
    L = _lib.lib()
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
