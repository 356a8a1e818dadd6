"""Extract the reference's published kernel checksums into a committed fixture.

The reference's own result artifacts carry, per kernel-bench cell, the checksum
sum_b |out_b| over B = 4096 synthetic inputs (13 significant digits). The inputs
are regenerable from the recipe in core/bench_kernels.cu:58-72,128-147
(oracle/ref_inputs.cpp restates it). This script reads that DATA (no reference
code is imported or run) and writes tests/golden/ref_checksums.json.

Run where the reference checkout is present (its files are not shipped with the tests):
    python tests/golden/make_ref_checksums.py /root/reference
"""
import json
import sys
from pathlib import Path

SRC = "results/throughput/throughput_NVIDIA_GeForce_RTX_4090_20260723T012923Z.json"
# cells whose inputs oracle/ref_inputs.cpp regenerates
KEEP = {"perm", "perm_dd", "haf", "haf_dd", "lhaf", "lhaf_dd", "tor", "tor_dd", "tor_recursive"}


def main(ref_root: str) -> None:
    d = json.load(open(Path(ref_root) / SRC))
    rows = []
    for r in d["rows"]:
        if r.get("func") in KEEP:
            rows.append({"func": r["func"], "matrix_dim": r["matrix_dim"], "batch": r["batch"],
                         "checksum": r["checksum"]})
    out = {"source": SRC, "recipe": "core/bench_kernels.cu:58-72,128-147; tor scale 0.05, "
           "symmetric for haf/lhaf/tor; tor_recursive on real inputs", "rows": rows}
    dst = Path(__file__).resolve().parent / "ref_checksums.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} rows to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
