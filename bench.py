"""Benchmark of the online Kilosort4 template-matching path on one MI355X.

Workload (the drop-in configuration at full probe size): 384 channels, 600 templates, 6 PCs,
nt = 61, 10 s of seeded synthetic int16 at 30 kHz cut into Kilosort batches of NT = 60000 samples
(padded windows of 60122 samples), at most 50 pursuit passes per window, FP32.  One step = one
padded window through the whole device path (CAR, highpass, whitening, projection, U.Xproj,
pursuit, PC features, positions) with the int16 windows already resident in HBM.

Prints one JSON line: x-realtime (seconds of recording per wall second), p50/p99 per-window
latency, the roofline figure of the dominant kernel (measured with HIP events on the library's
stream over the timed steps) and the CPU oracle timed on a bounded sample on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "bf16x3"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each probe stream is "
              "sorted independently on its own GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from livespikesorter_amd import synth
    from livespikesorter_amd.sorter import Sorter, pad_batch

    t0 = time.time()
    s = synth.make(2, seed=args.seed)
    a = s.art
    NT, nt, C, T = a["NT"], a["nt"], s.rec.shape[1], a["U"].shape[0]
    K = a["U"].shape[1]
    nb = (s.rec.shape[0] + NT - 1) // NT
    wins, starts = [], []
    for ib in range(nb):
        w, ws = pad_batch(s.rec, ib, NT, nt)
        wins.append(w)
        starts.append(ws)
    dev = torch.device("cuda:0")
    dwin = torch.from_numpy(np.stack(wins)).to(dev)          # [nb][W'][C] int16 resident in HBM
    Wp = dwin.shape[1]
    print(f"[bench] workload ready in {time.time() - t0:.1f}s: C={C} T={T} windows={nb} W'={Wp}",
          file=sys.stderr, flush=True)

    sorter = Sorter(a, max_window=Wp, max_passes=50, precision=args.precision)
    for i in range(args.warmup):
        sorter.sort_window(dwin[i % nb], starts[i % nb])
    torch.cuda.synchronize()
    lat, stage = [], []
    nsp = 0
    passes = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        out = sorter.sort_window(dwin[i % nb], starts[i % nb])
        lat.append(time.perf_counter() - t1)
        stage.append(sorter.timings())
        nsp += len(out)
        passes.append(out.passes)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    print(f"[bench] {args.steps} windows in {elapsed:.3f}s", file=sys.stderr, flush=True)

    fs = a["fs"]
    xrt = args.steps * NT / fs / elapsed
    lat_ms = np.array(lat) * 1e3
    st_ms = {k: float(np.mean([d[k] for d in stage])) for k in stage[0]}

    # dominant kernel: the larger of the U.Xproj GEMM and the pursuit's update traffic
    gemm_flops = 2.0 * T * K * C * Wp
    gemm_ms = st_ms["scores_gemm"]
    peak = FP32_MFMA_PEAK_TFLOPS if args.precision == "fp32" else BF16_MFMA_PEAK_TFLOPS
    achieved = gemm_flops / (gemm_ms * 1e-3) / 1e12
    roofline = {"kernel": "gemm_f32_kernel (U.Xproj scores)", "bound": "mfma",
                "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "per_launch": f"2*T*K*C*W' = {gemm_flops / 1e9:.1f} GFLOP, avg {gemm_ms:.3f} ms"}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import ks_oracle as ko
        torch.set_num_threads(args.cpu_threads)
        ib = 1 % nb
        X = ko.padded_batch(s.rec, ib, NT, nt, torch.float32)
        c0 = time.perf_counter()
        ref = ko.sort_window(X, a, ko.window_start(ib, NT, nt), torch.float32, max_passes=50)
        csec = time.perf_counter() - c0
        got = sorter.sort_window(dwin[ib], starts[ib])
        f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
        cpu = {"value": round(NT / fs / csec, 4), "unit": "x realtime", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": f"one padded window (batch {ib}, {NT} samples = {NT / fs:.1f} s) through the "
                         f"FP32 torch CPU oracle, {csec:.1f} s; GPU-vs-CPU spike F1 {f1:.5f}"}

    res = {
        "metric": "x-realtime & p99 batch latency, 384ch@30kHz; bf16x3 vs fp32 at spike F1>=0.999",
        "value": round(xrt, 2), "unit": "x realtime", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (seeded probe, templates and spikes; no recordings)",
        "config": {"workload": "384 ch, 600 templates, 6 PCs, nt=61, NT=60000-sample Kilosort windows "
                               "(W'=60122), 10 s recording cycled, <=50 pursuit passes",
                   "n_chan": C, "n_templates": T, "window": int(Wp), "NT": NT, "max_passes": 50,
                   "precision": args.precision},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "spikes_per_window": round(nsp / args.steps, 1),
        "passes_per_window": round(float(np.mean(passes)), 2),
        "stage_ms": {k: round(v, 4) for k, v in st_ms.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
