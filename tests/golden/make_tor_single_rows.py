"""Extract the reference's published single-large torontonian timings into a
committed fixture (data only; no reference code is imported or run).

core/bench_kernels.cu:368-410 times one evaluation of O = 0.5 (G + G^T) + 0.15 I,
G = 0.02 U(-1, 1) (mt19937_64(4321 + n)), n = 16, 20, 24, 26 modes, and records
evals/s = 1 / median seconds. Its checksums are the fp64 values, which at these
sizes are dominated by rounding (the exact values are ~0.18^n), so only the
timings are kept.

    python tests/golden/make_tor_single_rows.py /root/reference
"""
import glob
import json
import sys
from pathlib import Path


def main(ref_root: str) -> None:
    rows = {}
    for f in sorted(glob.glob(str(Path(ref_root) / "results/throughput/throughput_*.json"))):
        d = json.load(open(f))
        gpu = Path(f).name.split("_", 1)[1].rsplit("_", 1)[0]
        for r in d.get("rows", []):
            if r.get("func") == "tor_single":
                cur = rows.setdefault(r["matrix_dim"], {})
                cur[gpu] = max(cur.get(gpu, 0.0), r["evals_per_sec_median"])
    out = {"recipe": "core/bench_kernels.cu:368-410", "rows": [
        {"matrix_dim": k, "evals_per_sec_median": v} for k, v in sorted(rows.items())]}
    dst = Path(__file__).resolve().parent / "tor_single_reference.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} rows to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
