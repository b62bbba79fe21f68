"""bench.py's accounting on the CPU: the flop count of the two-site matvec against its closed
form and against a brute count of the einsum definition, the per-stage split, the committed
traffic summary covering the kernels the roofline can name, and the one-GPU guard."""
import json
import os
import subprocess
import sys

import numpy as np

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_matvec_flops_closed_form():
    # 4 d^2 D chi^3 + 2 d^4 D^2 chi^2 = 86.74 GFLOP at chi = 1024, D = 5, d = 2
    chi, d, D = 1024, 2, 5
    f = bench.two_site_flops(chi, chi, d, d, D, D)
    assert f == 4 * d * d * D * chi ** 3 + 2 * d ** 4 * D ** 2 * chi ** 2
    assert abs(f / 1e9 - 86.74) < 0.01
    st = bench.stage_flops(chi, chi, d, d, D, D)
    assert st[1] + st[2] + st[4] == f and st[3] == st[1] + st[2]


def test_matvec_flops_brute_count():
    """Multiply-adds of the three contractions counted from the index ranges of each einsum
    (L (p w q) theta (q i j b); S (s t r, w i j) on (p, w i j, b); X2 (p s t r b') R (b' r b))."""
    a, b, d1, d2, wl, wr = 3, 5, 2, 3, 4, 2
    macs = (a * wl) * (d1 * d2 * b) * a              # GEMM1: rows (p w), cols (i j b), sum q
    macs += a * b * (d1 * d2 * wr) * (wl * d1 * d2)  # MPO mix per (p, b)
    macs += (a * d1 * d2) * b * (wr * b)             # GEMM3: rows (p s t), cols b', sum (r b)
    assert bench.two_site_flops(a, b, d1, d2, wl, wr) == 2 * macs


def test_traffic_summary_covers_roofline_kernels():
    names = ["GEMM1+mix theta x L with the MPO mix in the epilogue (gemm1_mix_dma_kernel<float,5,2>)",
             "GEMM3 x R (gemm_dma_kernel<float,double,32,2>)"]
    with open(bench.TRAFFIC_FILE) as f:
        rows = json.load(f)["rows"]
    for name in names:
        t, src = bench.measured_traffic(name, 1024, "mixed")
        assert t is not None and src.endswith("matvec_traffic.json"), name
        row = max((r for r in rows if r["form"] == "mixed" and name.startswith(r["stage"])),
                  key=lambda r: len(r["stage"]))
        assert t == row["fetch_bytes"] + row["write_bytes"]
        # measured traffic is at least the algorithmic bytes (counter rounding aside)
        alg = row["algorithmic_read_bytes"] + row["algorithmic_write_bytes"]
        assert t >= 0.99 * alg
    # the algorithmic bytes of the fused stage: theta (float32) and LT read once, X2 written once
    chi, d, D = 1024, 2, 5
    fused = next(r for r in rows if r["form"] == "mixed" and r["stage"] == "GEMM1+mix")
    assert fused["algorithmic_read_bytes"] == (chi * d * d * chi + chi * chi * D) * 4
    assert fused["algorithmic_write_bytes"] == D * chi * chi * d * d * 4
    assert bench.measured_traffic("GEMM1+mix x", 512, "mixed") == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]
    assert np.isfinite(bench.PEAK["f32"]) and bench.HBM_PEAK == 8000.0
