"""Extract the reference's published kernel-only throughput rows into a
committed fixture (data only; no reference code is imported or run).

core/bench_kernels.cu times each batched kernel at B = 4096 on complex
U(-1, 1) inputs (symmetric for haf / lhaf / tor, the torontonian scaled by
0.05; the certified kernels on 0.3 U(-1, 1), symmetrized by averaging), 2
untimed warm-ups, wall clock around launch + synchronize, median of 7
(core/bench_kernels.cu:58-119, :306-363). The fixture keeps, per GPU, every
row's function, matrix side, cooperation width and median evals/s, from the two
runs the published kernel table is quoted on.

    python tests/golden/make_kernel_throughput_rows.py /root/reference
"""
import json
import sys
from pathlib import Path

RUNS = {
    "NVIDIA_GeForce_RTX_4090": "results/throughput/throughput_NVIDIA_GeForce_RTX_4090_20260723T012923Z.json",
    "NVIDIA_A100-SXM4-40GB": "results/throughput/throughput_NVIDIA_A100-SXM4-40GB_20260705T003328Z.json",
}


def main(ref_root: str) -> None:
    rows = []
    for gpu, rel in RUNS.items():
        d = json.load(open(Path(ref_root) / rel))
        for r in d.get("rows", []):
            if r.get("batch") != 4096 or "evals_per_sec_median" not in r:
                continue
            rows.append({"gpu": gpu, "func": r["func"], "matrix_dim": r["matrix_dim"],
                         "groups": r.get("groups"), "evals_per_sec_median": r["evals_per_sec_median"]})
    out = {"recipe": "core/bench_kernels.cu:58-119,306-363", "source": RUNS, "rows": rows}
    dst = Path(__file__).resolve().parent / "kernel_throughput_reference.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} rows to {dst}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
