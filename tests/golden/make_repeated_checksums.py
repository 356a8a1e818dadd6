"""Extract the reference's published repeated-row sieve checksums into a
committed fixture.

bench/repeated_ab.py (modes = 3, batch = 2048, seed = 0) evaluates the loop
hafnian of A = 0.3 (G + G^T), G = standard_normal((6, 6)) + 1j standard_normal((6, 6))
from numpy.random.default_rng(0), gamma = 0, every row of the batch the pattern
(q, ..., q), and records checksum = sum of the 2048 values (identical values, so
the pairwise sum is exactly 2048 v). The result artifacts hold the same checksums
on an RTX 4090 and an A100. This script reads that DATA (no reference code is
imported or run) and writes tests/golden/repeated_ab_checksums.json.

Run where the reference checkout is present (its files are not shipped with the tests):
    python tests/golden/make_repeated_checksums.py /root/reference
"""
import glob
import json
import sys
from pathlib import Path


def main(ref_root: str) -> None:
    files = sorted(glob.glob(str(Path(ref_root) / "results/throughput/repeated_ab_*.json")))
    rows, rates = {}, {}
    for f in files:
        d = json.load(open(f))
        gpu = d["provenance"].get("environment", {}).get("gpu", {}).get("name", "?")
        for r in d["rows"]:
            key = r["q"]
            chk = r["sieve"]["checksum"]
            if key in rows and rows[key]["checksum"] != chk:
                raise SystemExit(f"checksums differ between artifacts at q={key}")
            rows[key] = {"q": r["q"], "M": r["M"], "batch": r["batch"], "sieve_terms": r["sieve_terms"],
                         "checksum": chk}
            rates.setdefault(key, {}).setdefault(gpu, []).append(r["sieve"]["evals_per_s"])
    out = {"source": [str(Path(f).relative_to(ref_root)) for f in files],
           "recipe": "bench/repeated_ab.py:45-60: modes 3 (M = 6), batch 2048, seed 0, gamma = 0, reps = q",
           "rows": [dict(rows[q], public_path_evals_per_s={g: max(v) for g, v in rates[q].items()})
                    for q in sorted(rows)]}
    dst = Path(__file__).resolve().parent / "repeated_ab_checksums.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(out['rows'])} rows to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
