"""Benchmark of the online Kilosort4 template-matching path on one MI355X.

Default workload (config 5, the metric's batch-latency case): four probes of uneven size (384 ch /
600 templates, 384/250, 128/120, 64/40) at 30 kHz, 60 s each, cut into 3000-sample (100 ms)
batches with 61-sample halos; one step = one graph-replayed batch step over all four probes
(upload-free device samples, CAR, highpass, whitening, projection, U.Xproj, pursuit, residual, PC
features, positions, spike-list readback), bf16x3 GEMMs, tiles and pursuit schedules autotuned
beforehand, a template-refinement round every 10 s.  `value` is the aggregate x-realtime of the
`--steps` timed steps after `--warmup` untimed ones (all remaining batches of the 60 s by
default); p50/p99 are per 100-ms batch step.  Beside it: the same 60 s stream to its end with
the refinement rounds, the plain FP32 path on the same windows (one probe at a time), the same
stream with host samples (PCIe in every step), the pursuit update kernel's roofline (HIP events
around each launch on the library's stream), the FP32 CPU oracle on a sample of the same windows,
and the config-2 figures (384 ch / 600 templates, 2-s Kilosort windows) as `window_2s`.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "update_kernel_traffic_384ch_600t.md")


def _keep(o):
    """What the benchmark keeps of a window's spike list for its checks: sample, template,
    amplitude, column and passes.  Features and positions are read back and converted for every
    window, then dropped, as a consumer that writes them out would (keeping every window's ~3 MB
    would make each call's fresh output arrays page-fault in, ~0.6 ms per window)."""
    from livespikesorter_amd.sorter import SpikeList
    return SpikeList(o.sample, o.template, o.amplitude, None, None, None, o.column, o.passes)


def run_path(sorter, dwin, starts, steps, warmup, profile=False):
    import torch
    nb = len(starts)
    for i in range(warmup):
        sorter.sort_window(dwin[i % nb], starts[i % nb])
    torch.cuda.synchronize()
    if profile:
        sorter.profile_update(True)
    lat, stage, outs = [], [], []
    upd_ms, upd_n = 0.0, 0
    upd_each, upd_gpu, upd_spikes = [], [], []
    t_start = time.perf_counter()
    for i in range(steps):
        t1 = time.perf_counter()
        out = sorter.sort_window(dwin[i % nb], starts[i % nb])
        lat.append(time.perf_counter() - t1)
        stage.append(sorter.timings())
        outs.append(_keep(out))
        if profile:
            ms, n = sorter.update_stats()
            upd_ms += ms
            upd_n += n
            e, g, sp = sorter.update_times()
            upd_each.append(e)
            upd_gpu.append(g)
            upd_spikes.append(sp)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    if profile:
        sorter.profile_update(False)
    return dict(elapsed=elapsed, lat_ms=np.array(lat) * 1e3,
                stage_ms={k: float(np.mean([d[k] for d in stage])) for k in stage[0]},
                outs=outs, upd_ms=upd_ms, upd_n=upd_n, upd_each=upd_each, upd_gpu=upd_gpu, upd_spikes=upd_spikes)


def update_columns_of(sorter):
    from livespikesorter_amd.streaming import update_columns
    return update_columns(sorter)


def run_in_flight(a, sorter, dwin, starts, steps, warmup, k, precision, Wp, seq_outs, tiles):
    """The same windows with k in flight (the sorter plus k-1 more handles with its tuning: tile
    shapes, update grid and layout; one host thread each); checks the spike lists against the
    one-at-a-time run."""
    import torch
    from livespikesorter_amd.sorter import Sorter, sort_windows
    from livespikesorter_amd._lib import lib
    extra = [Sorter(a, max_window=Wp, max_passes=50, precision=precision, max_spikes=1 << 16) for _ in range(k - 1)]
    shapes, grid, cols = tiles
    for so in extra:
        so.set_tiles(*shapes)
        lib().lss_set_tuning(so._h, grid, 0)
        lib().lss_set_update_columns(so._h, cols)
    group = [sorter] + extra
    nb = len(starts)
    wl = [dwin[i % nb] for i in range(steps)]
    sl = [starts[i % nb] for i in range(steps)]
    warm = max(warmup, k)
    sort_windows(group, [dwin[i % nb] for i in range(warm)], [starts[i % nb] for i in range(warm)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = sort_windows(group, wl, sl, consume=lambda i, o: _keep(o))
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    equal = all(np.array_equal(o.sample, q.sample) and np.array_equal(o.template, q.template)
                and np.array_equal(o.amplitude, q.amplitude) for o, q in zip(outs, seq_outs))
    for so in extra:
        so.close()
    return dict(elapsed=el, equal=bool(equal))


def host_cpu() -> dict:
    """The host the CPU baseline ran on: logical CPUs and model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "host_cpus": os.cpu_count()}


def read_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary), or None."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


METRIC = "x-realtime & p99 batch latency, 384ch@30kHz; bf16x3 vs fp32 at spike F1≥0.999"
STREAM_PROBES = [(384, 600), (384, 250), (128, 120), (64, 40)]   # (channels, templates)
STREAM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "update_kernel_traffic_stream_384ch_600t.md")
STREAM_PERSIST_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "persistent_kernel_traffic_stream_384ch_600t.md")
STREAM_GROUP_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "group_persistent_traffic_stream_4probes.md")


def _pct(x, q):
    return round(float(np.percentile(np.asarray(x) * 1e3, q)), 3)


def _update_roofline(kernel, ev_ms, gpu_ms, spikes, bytes_per_spike, traffic, where):
    """Roofline of the pursuit update kernel over the launches that applied spikes: algorithmic
    bytes per launch (bytes_per_spike x spikes) over the average launch duration from the
    workgroups' GPU-clock stamps (the execution time a kernel trace reports); the HIP-event time
    of the same launches, which also holds the dispatch latency, is stated beside it."""
    d_ms = float(gpu_ms.mean())
    per_launch = float(spikes.mean()) * bytes_per_spike
    achieved = per_launch / (d_ms * 1e-3) / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "launch_us": round(d_ms * 1e3, 2), "launch_us_hip_events": round(float(ev_ms.mean()) * 1e3, 2),
            "per_launch": f"{bytes_per_spike / 1e3:.1f} KB per accepted spike x {spikes.mean():.1f} spikes per "
                          f"launch = {per_launch / 1e6:.2f} MB; avg {d_ms * 1e3:.2f} us per launch from the workgroups' "
                          f"GPU-clock stamps (HIP events around each launch: {float(ev_ms.mean()) * 1e3:.2f} us, "
                          f"which includes the dispatch latency) over {len(gpu_ms)} launches that applied spikes "
                          f"({where}, measurement mode)"}


def bench_stream(args):
    """Config 5: four probes of uneven size on one clock, 3000-sample (100 ms) batches with
    nt-sample halos, 60 s of each, one graph-replayed step per batch for all probes, bf16x3 GEMMs,
    a template-refinement round every refine_seconds.  After `warmup` untimed steps exactly `steps`
    steps are timed (the value); the stream then runs to the end of the 60 s (stream_60s, which
    holds the refinement steps).  Beside it: the plain FP32 path (one probe at a time,
    sort_window), the same stream with host samples (PCIe in every step), the update kernel's
    roofline and the FP32 CPU oracle on a sample of the same windows."""
    import gc
    import torch
    from livespikesorter_amd import synth
    from livespikesorter_amd._lib import lib as _l
    from livespikesorter_amd.metrics import f1_same_template
    from livespikesorter_amd.sorter import Sorter
    from livespikesorter_amd.streaming import (StreamingSorter, autotune, autotune_persistent, autotune_stream,
                                               batcher_grid_shares, update_columns)

    batch, fs = 3000, 30000.0
    n = int(args.seconds * fs) + 2 * batch
    t0 = time.time()
    probes = [synth.generate(synth.SynthConfig(n_chan=C, n_units=T, n_samples=n, NT=60000,
                                               nrange=min(32, C), seed=100 + i))
              for i, (C, T) in enumerate(STREAM_PROBES)]
    print(f"[bench] stream workload ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    nbat = (n - 2 * batch) // batch
    warm = max(0, min(args.warmup, nbat - 1))
    K = nbat - warm if args.steps is None else max(1, min(args.steps, nbat - warm))

    def window(rec, k, nt):
        idx = np.clip(np.arange(k * batch - nt, (k + 1) * batch + nt), 0, None)
        return np.ascontiguousarray(rec[idx])

    mk = lambda p, prec: Sorter(p.art, max_window=batch + 2 * p.art["nt"], max_passes=100, precision=prec,
                                max_spikes=1 << 15)
    sorters = [mk(p, args.precision) for p in probes]
    plain_sorters = [mk(p, "fp32") for p in probes]      # the plain per-probe path
    refine_every = int(round(args.refine_seconds * fs / batch)) if args.refine_seconds > 0 else 0
    tuning = []
    shares = batcher_grid_shares(sorters, batch + 2 * probes[0].art["nt"])
    pinned = json.loads(args.pin_tuning) if args.pin_tuning else None
    if pinned is not None and len(pinned) != len(probes):
        raise SystemExit(f"--pin-tuning: {len(pinned)} entries for {len(probes)} probes")
    for i, (s_, s_plain, p, cap) in enumerate(zip(sorters, plain_sorters, probes, shares)):
        if pinned is not None:
            # a schedule printed by an earlier run, replayed as is (profiling runs, whose tracer
            # overhead would steer the autotuner elsewhere)
            tiles, plain_tiles, g, cols, p1, pg, pf = pinned[i]
            s_.set_tiles(*tiles)
            s_plain.set_tiles(*plain_tiles)
            _l().lss_set_tuning(s_._h, g, p1)
            _l().lss_set_update_columns(s_._h, cols)
            s_.set_persistent(pg, pf)
            tuning.append({"tiles": list(tiles), "plain_tiles": list(plain_tiles), "update_grid": g,
                           "update_columns": cols, "first_passes": p1, "grid_share": cap, "persistent_grid": pg,
                           "persistent_after": pf, "pinned": True})
            continue
        w = torch.from_numpy(window(p.rec, 3, p.art["nt"])).to("cuda:0")
        tiles, tile_ms = s_.autotune_tiles(w)
        plain_tiles, _ = s_plain.autotune_tiles(w)
        g, p1, _ = autotune(s_, w)
        pg, pf, pms = autotune_persistent(s_, w, max_grid=cap)
        tuning.append({"tiles": tiles, "tile_ms": tile_ms, "plain_tiles": plain_tiles, "update_grid": g,
                       "update_columns": update_columns(s_),
                       "first_passes": p1, "grid_share": cap, "persistent_grid": pg, "persistent_after": pf,
                       "persistent_ms": {f"{k[0]}/{k[1]}": round(v, 3) for k, v in pms.items()}})
    dev = torch.device("cuda:0")
    drecs = [torch.from_numpy(p.rec).to(dev) for p in probes]     # recordings resident in HBM
    stream_tuning = None
    if pinned is None:
        # the probes' schedules chosen together, on whole streaming steps
        joint, joint_ms = autotune_stream(sorters, drecs, batch)
        for tu, (g, f) in zip(tuning, joint):
            tu["persistent_grid"], tu["persistent_after"] = g, f
        stream_tuning = {"chosen": [list(c) for c in joint],
                         "step_ms": {"/".join(f"{g}:{f}" for g, f in k): round(v, 4) for k, v in joint_ms.items()}}
    w0 = [torch.from_numpy(window(p.rec, 0, p.art["nt"])).to(dev) for p in probes]

    def dwindow(i, k):
        nt = probes[i].art["nt"]
        return w0[i] if k == 0 else drecs[i][k * batch - nt:(k + 1) * batch + nt]

    def fresh_sorters():
        # new handles (a measured run refines the templates of its own) with the same tuning
        out = [mk(p, args.precision) for p in probes]
        for s2, tu in zip(out, tuning):
            s2.set_tiles(*tu["tiles"])
            s2.set_persistent(tu["persistent_grid"], tu["persistent_after"])
            _l().lss_set_tuning(s2._h, tu["update_grid"], tu["first_passes"])
            _l().lss_set_update_columns(s2._h, tu["update_columns"])
        return out

    def group_roofline(k0, nsteps):
        # the grouped persistent pursuit (every probe's passes in one launch: the step's critical
        # path) in measurement mode over steps k0 .. k0 + nsteps - 1 of a stream that runs ahead as
        # value's run does (no refinement); None when the schedule does not run grouped
        srt = fresh_sorters()
        bps = [3 * 4 * s_.T * (2 * s_.nt + 1) for s_ in srt]   # ctc rows read + B rows read and written
        rows = []
        need = (k0 + nsteps + 2) * batch + 64
        with StreamingSorter(srt, batch=batch, use_graph=True, run_ahead=True) as st:
            st.profile(True)
            for i in range(len(probes)):
                st.push(i, drecs[i][:need])
            k = 0
            while k < k0 + nsteps:
                if st.step() is None:
                    break
                if k >= k0:
                    rows.append(st.last_profile())
                k += 1
            grouped_ = st.grouped()
        for s_ in srt:
            s_.close()
        if not grouped_ or not rows:
            return None
        us = np.array([r[0] for r in rows])
        spk = np.array([r[1] for r in rows], np.float64)           # [steps][probes]
        ps = np.array([r[2] for r in rows], np.float64)
        grid = rows[0][3]
        per_launch = float((spk * np.array(bps)).sum(1).mean())
        achieved = per_launch / (float(us.mean()) * 1e-6) / 1e9
        parts = ", ".join(f"{C}/{T}: {b_ / 1e3:.1f} KB x {m_:.1f}" for (C, T), b_, m_ in
                          zip(STREAM_PROBES, bps, spk.mean(0)))
        return {"kernel": f"pursuit_persistent_group_kernel (the four probes' persistent pursuits in one launch, "
                          f"{sum(grid)} workgroups: {'/'.join(str(g_) for g_ in grid)})",
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(STREAM_GROUP_TRAFFIC_FILE),
                "launch_us": round(float(us.mean()), 2),
                "us_per_pass": round(float(us.mean()) / float(ps[:, 0].mean()), 2),
                "per_launch": f"per probe (C/T) bytes per spike applied (12 T (2 nt + 1): ctc rows read, score rows "
                              f"read and written) x spikes per launch: {parts}; {per_launch / 1e6:.2f} MB per launch, "
                              f"the 384/600 probe over {ps[:, 0].mean():.1f} passes; avg {us.mean():.1f} us per launch "
                              f"from the workgroups' GPU-clock stamps over {len(rows)} steps of the timed region "
                              f"(measurement mode: the pursuit tail launched without the graph); each pass is a "
                              f"dependent step ending in a grid barrier, so the kernel is bound by the pass "
                              f"latency, not by HBM"}

    def update_roofline(k0, nwin):
        # the pursuit of the 384-ch / 600-template probe (the step's critical path) in measurement
        # mode over windows k0 .. k0 + nwin - 1, one at a time, on the schedule the step uses: the
        # persistent kernel (the passes after its first per-pass rounds) when the schedule takes
        # it, the update kernel of the per-pass rounds beside it
        so = sorters[0]
        T0, nt0 = so.T, so.nt
        bytes_per_spike = 3 * 4 * T0 * (2 * nt0 + 1)      # ctc rows read + B rows read and written
        so.profile_update(True)
        ev_all, gpu_all, sp_all, pers = [], [], [], []
        for k in range(k0, k0 + nwin):
            so.sort_window(dwindow(0, k), k * batch - nt0)
            ev_, gpu_, sp_ = so.update_times()
            ev_all.append(ev_[sp_ > 0])
            gpu_all.append(gpu_[sp_ > 0])
            sp_all.append(sp_[sp_ > 0])
            pers.append(so.persistent_times())
        so.profile_update(False)
        ev_all, gpu_all, sp_all = np.concatenate(ev_all), np.concatenate(gpu_all), np.concatenate(sp_all)
        upd = None
        if len(sp_all):
            upd = _update_roofline("update_kernel<3, C> (pursuit score update of the 384-ch / 600-template probe, "
                                   f"{update_columns(so)} score column(s) per wavefront)", ev_all, gpu_all, sp_all,
                                   bytes_per_spike, read_traffic(STREAM_TRAFFIC_FILE),
                                   f"{nwin} windows of the timed region, one at a time")
        pers = np.array([p_ for p_ in pers if p_[1] > 0], np.float64)
        if not len(pers):
            return upd, None
        ms, spk, ps, grid = pers[:, 0], pers[:, 1], pers[:, 2], pers[:, 3]
        per_launch = float(spk.mean()) * bytes_per_spike
        achieved = per_launch / (float(ms.mean()) * 1e-3) / 1e9
        roof = {"kernel": f"pursuit_persistent_kernel<3> (the 384-ch / 600-template probe's passes after its first "
                          f"{so.persistent(batch + 2 * nt0)[1]}, {int(grid.mean())} workgroups)",
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(STREAM_PERSIST_TRAFFIC_FILE),
                "launch_us": round(float(ms.mean()) * 1e3, 2),
                "us_per_pass": round(float(ms.sum()) * 1e3 / float(ps.sum()), 2),
                "per_launch": f"{bytes_per_spike / 1e3:.1f} KB per spike it applies x {spk.mean():.1f} spikes per "
                              f"launch = {per_launch / 1e6:.2f} MB over {ps.mean():.1f} passes; avg "
                              f"{ms.mean() * 1e3:.1f} us per launch from the workgroups' GPU-clock stamps over "
                              f"{len(pers)} windows of the timed region, one at a time (measurement mode); each "
                              f"pass is a dependent step ending in a grid barrier, so the kernel is bound by "
                              f"the pass latency, not by HBM"}
        return roof, upd

    if args.roofline_only:
        # counter runs: only the measured pursuit launches (with the schedule pinned or tuned)
        g = group_roofline(args.warmup, 50)
        r, u = update_roofline(args.warmup, 50) if g is None else (None, None)
        return {"roofline": g or r, "roofline_update_kernel": u, "tuning": tuning}

    # plain per-probe path: every batch's windows one probe at a time (device-resident windows)
    for i, (s_, p) in enumerate(zip(plain_sorters, probes)):     # first-window setup, untimed
        s_.sort_window(dwindow(i, 0), -p.art["nt"])
    torch.cuda.synchronize()
    plain, plain_lat = [], []
    for k in range(nbat):
        t1 = time.perf_counter()
        plain.append([s_.sort_window(dwindow(i, k), k * batch - p.art["nt"])
                      for i, (s_, p) in enumerate(zip(plain_sorters, probes))])
        plain_lat.append(time.perf_counter() - t1)

    def run_stream(srt, device_input, run_ahead=True):
        """One pass over the recordings, samples pushed 100 ms at a time, one step per batch.
        run_ahead: the stream holds one batch more than the step needs (a backlog of one batch),
        and each step preprocesses the next batch beside its own pursuit (lss_stream_set_run_ahead);
        otherwise live pacing: each step finds exactly its own batch (and its nt-sample halo).
        Returns per-step latencies, outputs, refinement infos, the wall times of steps [warm, warm+K)
        and [warm, nbat), the steps that ran ahead and whether the steps ran grouped."""
        lat, outs, refines = [], [], []
        src = drecs if device_input else [p.rec for p in probes]
        first = batch + 64 if run_ahead else max(p.art["nt"] for p in probes)
        gc.collect()
        gc.disable()      # no collector pauses of the Python harness inside the timed steps
        try:
            with StreamingSorter(srt, batch=batch, use_graph=True, refine_every=refine_every,
                                 run_ahead=run_ahead) as st:
                for i in range(len(probes)):
                    st.push(i, src[i][:first])
                pos = first
                k = 0
                t_a = t_b = None
                t_k = 0.0
                while k < nbat:
                    if k == warm:
                        torch.cuda.synchronize()
                        t_a = time.perf_counter()
                    for i in range(len(probes)):
                        st.push(i, src[i][pos:pos + batch])
                    pos += batch
                    t1 = time.perf_counter()
                    r = st.step()
                    if r is None:
                        continue
                    lat.append(time.perf_counter() - t1)
                    outs.append(r)
                    k += 1
                    if k == warm + K:
                        torch.cuda.synchronize()
                        t_k = time.perf_counter() - t_a
                    if refine_every and k >= refine_every and k % refine_every == min(refine_every - 1,
                                                                                         len(probes) + 4):
                        # the round's refinements (one per probe, each adopted four steps after its
                        # own) are all in by now
                        refines.append([st.last_refine(i) for i in range(len(probes))])
                torch.cuda.synchronize()
                t_b = time.perf_counter() - t_a
                ahead = st.ran_ahead()
                grouped = st.grouped()
        finally:
            gc.enable()
        return lat, outs, refines, t_k, t_b, ahead, grouped

    # the measured run: recordings resident in HBM, windows assembled device to device
    lat, outs, refines, t_k, t_all, n_ahead, grouped = run_stream(sorters, True)

    # live pacing (no backlog, so no running ahead): the per-batch latency of a live stream
    live_sorters = fresh_sorters()
    lat_live, outs_live, _, t_kl, _, _, _ = run_stream(live_sorters, True, run_ahead=False)
    for k in range(len(outs)):
        for i in range(len(probes)):
            if not (np.array_equal(outs[k][i].sample, outs_live[k][i].sample)
                    and np.array_equal(outs[k][i].amplitude, outs_live[k][i].amplitude)):
                raise SystemExit(f"stream results differ with and without running ahead (batch {k}, probe {i})")
    # the same stream with host samples (each batch crosses PCIe): fresh sorters, same tuning
    host_sorters = fresh_sorters()
    lat_h, _, _, t_kh, t_allh, _, _ = run_stream(host_sorters, False)

    def pooled_f1(k0, k1):
        out = []
        for i in range(len(probes)):
            cat = lambda rr, key: np.concatenate([getattr(rr[k][i], key) for k in range(k0, k1)])
            out.append(f1_same_template(cat(plain, "sample"), cat(plain, "template"),
                                        cat(outs, "sample"), cat(outs, "template"), dt=1))
        return out
    first = min(refine_every, len(outs)) if refine_every else len(outs)
    f1 = min(pooled_f1(0, first))
    f1_all = min(pooled_f1(0, len(outs)))
    P = len(probes)
    sec = lambda nb: nb * batch / fs * P                       # seconds of data in nb steps
    value = sec(K) / t_k
    plain_k = sec(K) / sum(plain_lat[warm:warm + K])
    timed_lat, all_lat = lat[warm:warm + K], lat[warm:]
    n_ref_timed = sum(1 for k in range(warm, warm + K) if refine_every and (k + 1) >= refine_every
                      and ((k + 1) % refine_every) < P)

    roofline, roofline_update = group_roofline(warm, min(K, 50)), None
    if roofline is None:
        roofline, roofline_update = update_roofline(warm, min(K, 50))

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import ks_oracle as ko
        torch.set_num_threads(args.cpu_threads)
        t_art = time.perf_counter()
        ctcs = []
        for p in probes:
            U32, W32 = torch.from_numpy(p.art["U"]), torch.from_numpy(p.art["wPCA"])
            ctcs.append(ko.prepare_matching(U32, W32))
        t_art = time.perf_counter() - t_art
        nb_cpu = min(args.cpu_batches, K)
        per_batch, f1s = [], []
        for k in range(warm, warm + nb_cpu):
            t1 = time.perf_counter()
            refs = []
            for i, p in enumerate(probes):
                nt = p.art["nt"]
                X = torch.from_numpy(window(p.rec, k, nt).T.astype(np.float32))
                refs.append(ko.sort_window(X, p.art, k * batch - nt, dtype=torch.float32, max_passes=100,
                                           ctc=ctcs[i]))
            per_batch.append(time.perf_counter() - t1)
            for i, r in enumerate(refs):
                f1s.append(ko.spike_f1(r.samples, r.templates, plain[k][i].sample, plain[k][i].template, dt=1))
        cpu = {"value": round(sec(nb_cpu) / sum(per_batch), 4), "unit": "x realtime (aggregate over probes)",
               "cores": torch.get_num_threads(), **host_cpu(), "kind": "port",
               "p50_ms": _pct(per_batch, 50), "p99_ms": _pct(per_batch, 99),
               "sample": f"{nb_cpu} batches x {P} probes (100 ms windows of the timed region) through the FP32 torch "
                         f"CPU oracle, {sum(per_batch):.1f} s; ctc of the four probes built once beforehand "
                         f"({t_art:.1f} s, not in the time); spike F1 of the plain FP32 GPU path against it: min "
                         f"{min(f1s):.5f}"}

    res = {
        "metric": METRIC,
        "value": round(value, 2), "unit": "x realtime (aggregate over 4 probes)", "n_gpus": 1,
        "steps": K, "warmup": warm, "ms_per_step": round(t_k / K * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (seeded probes, templates and spikes; no recordings)",
        "config": {"workload": "config 5: 4 probes (384 ch/600 templates, 384/250, 128/120, 64/40) at 30 kHz, "
                               "3000-sample (100 ms) batches with 61-sample halos, one graph-replayed step per "
                               "batch for all probes, bf16x3 GEMMs, autotuned tiles and pursuit schedules, a "
                               "template-refinement round every refine_seconds; recordings resident in HBM",
                   "probes": STREAM_PROBES, "batch": batch, "seconds_per_probe": round(nbat * batch / fs, 2),
                   "refine_seconds": args.refine_seconds},
        "p50_ms": _pct(lat_live[warm:warm + K], 50), "p99_ms": _pct(lat_live[warm:warm + K], 99),
        "max_ms": _pct(lat_live[warm:warm + K], 100),
        "latency_note": "per 100-ms batch (all four probes) of a live stream: each step finds only its own batch "
                        "(no running ahead), over the timed steps",
        "throughput_run": {"step_p50_ms": _pct(timed_lat, 50), "step_p99_ms": _pct(timed_lat, 99),
                           "steps_ran_ahead": int(n_ahead), "steps": len(lat), "grouped_steps": bool(grouped),
                           "live_x_realtime": round(sec(K) / t_kl, 2),
                           "note": "value's run: the stream one batch ahead of the steps (a backlog of one batch), "
                                   "each step preprocessing the next batch beside its own pursuit; results equal "
                                   "those of live pacing (checked)"},
        "refinement_steps_in_timed_region": n_ref_timed,
        "speedup_vs_plain": round(value / plain_k, 3),
        "plain_fp32": {"x_realtime": round(plain_k, 2), "p50_ms": _pct(plain_lat[warm:warm + K], 50),
                       "p99_ms": _pct(plain_lat[warm:warm + K], 99),
                       "note": "FP32 GEMMs, the same windows one probe at a time through lss_sort_window "
                               "(tiles tuned, default pursuit schedule), the timed steps' batches"},
        "stream_60s": {"steps": nbat - warm, "x_realtime": round(sec(nbat - warm) / t_all, 2),
                       "p50_ms": _pct(all_lat, 50), "p99_ms": _pct(all_lat, 99), "max_ms": _pct(all_lat, 100),
                       "slowest_steps": [[int(i) + warm, round(float(all_lat[i]) * 1e3, 3)]
                                         for i in np.argsort(-np.asarray(all_lat))[:6]],
                       "speedup_vs_plain": round((sec(nbat - warm) / t_all) / (sec(nbat - warm) /
                                                                               sum(plain_lat[warm:])), 3),
                       "refinement_rounds": len(refines)},
        "host_input": {"x_realtime": round(sec(K) / t_kh, 2), "p50_ms": _pct(lat_h[warm:warm + K], 50),
                       "p99_ms": _pct(lat_h[warm:warm + K], 99),
                       "stream_60s_x_realtime": round(sec(nbat - warm) / t_allh, 2),
                       "note": "the same stream with samples pushed from host memory: each step uploads its "
                               "windows over PCIe (the ingest-inclusive rate)"},
        "spike_f1_vs_plain_fp32_before_refinement": round(float(f1), 6),
        "spike_f1_vs_plain_fp32_all_batches": round(float(f1_all), 6),
        "f1_note": "per probe, pooled over batches, +-1 sample and same template; the minimum over probes",
        "roofline": roofline,
        "roofline_update_kernel": roofline_update,
        "refinement": {"every_batches": refine_every, "rounds": len(refines),
                       "loss_ratio_after_before": [[round(x["loss_after"] / x["loss_before"], 6) if x["loss_before"]
                                                    else None for x in rr] for rr in refines]},
        "tuning": tuning,
        "stream_tuning": stream_tuning,
        "cpu_baseline": cpu,
    }
    for s_ in sorters + plain_sorters + host_sorters + live_sorters:
        s_.close()
    return res


def bench_window(args, steps, warmup, config=2, in_flight=3, compare="fp32", host_input=True):
    """Config 2 (or 1): 384 ch, 600 templates, 10 s of int16 in NT = 60000 Kilosort windows
    (W' = 60122), <= 50 passes.  One step = one padded window through the whole device path incl.
    the spike-list readback.  Returns the result dict (value: windows in flight on separate
    handles)."""
    import torch
    from livespikesorter_amd import synth
    from livespikesorter_amd.sorter import Sorter, pad_batch
    from livespikesorter_amd.metrics import f1_same_template

    t0 = time.time()
    s = synth.make(config, seed=args.seed)
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
    print(f"[bench] window workload ready in {time.time() - t0:.1f}s: C={C} T={T} windows={nb} W'={Wp}",
          file=sys.stderr, flush=True)

    from livespikesorter_amd.streaming import autotune
    sorter = Sorter(a, max_window=Wp, max_passes=50, precision=args.precision, max_spikes=1 << 16)
    tiles, tile_ms = sorter.autotune_tiles(dwin[1 % nb])
    upd_grid, _, _ = autotune(sorter, dwin[1 % nb])
    tiles = (tiles, upd_grid, update_columns_of(sorter))
    r = run_path(sorter, dwin, starts, steps, warmup)
    fly = run_in_flight(a, sorter, dwin, starts, steps, warmup, in_flight, args.precision, Wp,
                        r["outs"], tiles) if in_flight > 1 else None
    # the roofline's per-launch durations come from a second timed run with an event pair around
    # every update launch (kept out of the run above, whose wall clock is the reported value)
    rp = run_path(sorter, dwin, starts, min(steps, 2 * nb), 0, profile=True)
    print(f"[bench] {args.precision}: {steps} windows in {r['elapsed']:.3f}s", file=sys.stderr, flush=True)
    fs = a["fs"]
    xrt_seq = steps * NT / fs / r["elapsed"]
    xrt = steps * NT / fs / fly["elapsed"] if fly else xrt_seq
    spikes = sum(len(o) for o in r["outs"])
    passes = float(np.mean([o.passes for o in r["outs"]]))

    # dominant kernel: the pursuit update (per accepted spike it reads and writes the B rows of the
    # spike's 2nt+1 columns and reads as many ctc rows: 3 * 4 * T * (2nt+1) bytes), averaged over
    # the launches that applied spikes
    bytes_per_spike = 3 * 4 * T * (2 * nt + 1)
    ev_w = np.concatenate(rp["upd_each"])
    gpu_w = np.concatenate(rp["upd_gpu"])
    sp_w = np.concatenate(rp["upd_spikes"])
    work = sp_w > 0
    traffic = read_traffic() if (C, T) == (384, 600) else None
    from livespikesorter_amd.streaming import update_columns
    roofline = _update_roofline(f"update_kernel<{(T + 255) // 256}, C> (pursuit score update, {update_columns(sorter)} "
                                "score column(s) per wavefront)", ev_w[work], gpu_w[work], sp_w[work],
                                bytes_per_spike, traffic, f"{len(rp['outs'])} windows one at a time")
    gemm_flops = 2.0 * T * K * C * Wp
    gemm_ms = r["stage_ms"]["scores_gemm"]
    peak_gemm = FP32_MFMA_PEAK_TFLOPS if args.precision == "fp32" else BF16_MFMA_PEAK_TFLOPS
    gemm = {"kernel": "U.Xproj GEMM", "bound": "mfma",
            "achieved": round(gemm_flops / (gemm_ms * 1e-3) / 1e12, 2), "peak": peak_gemm, "unit": "TFLOP/s",
            "frac": round(gemm_flops / (gemm_ms * 1e-3) / 1e12 / peak_gemm, 4)}
    if args.precision in ("bf16x3", "fp16x3"):
        gemm["issued_frac"] = round(3 * gemm["frac"], 4)
        gemm["note"] = "achieved/frac count the algorithmic 2TKCW' flops; issued_frac counts the three MFMA products"

    host = None
    if host_input:
        # ingest-inclusive: host int16 windows, uploaded by lss_sort_window in every call, from
        # page-locked memory (the facade's window buffer, lss_host_alloc) and from pageable memory
        def host_rate(hw):
            for i in range(warmup):
                sorter.sort_window(hw[i % nb], starts[i % nb])
            lat_h = []
            t1 = time.perf_counter()
            for i in range(steps):
                t2 = time.perf_counter()
                sorter.sort_window(hw[i % nb], starts[i % nb])
                lat_h.append(time.perf_counter() - t2)
            el = time.perf_counter() - t1
            return {"x_realtime": round(steps * NT / fs / el, 2), "p50_ms": _pct(lat_h, 50),
                    "p99_ms": _pct(lat_h, 99)}

        pinned = [torch.from_numpy(np.ascontiguousarray(w)).pin_memory().numpy() for w in wins]
        host = host_rate(pinned)
        host["pageable"] = host_rate([np.ascontiguousarray(w) for w in wins])
        host["note"] = ("host windows one at a time, each 46 MB window uploaded in the step; page-locked "
                        "windows (pageable: the same from ordinary host memory)")

    cmp = None
    if compare and compare not in ("none", args.precision):
        sorter2 = Sorter(a, max_window=Wp, max_passes=50, precision=compare, max_spikes=1 << 16)
        tiles2, _ = sorter2.autotune_tiles(dwin[1 % nb])
        g2, _, _ = autotune(sorter2, dwin[1 % nb])
        tiles2 = (tiles2, g2, update_columns_of(sorter2))
        r2 = run_path(sorter2, dwin, starts, steps, warmup)
        fly2 = run_in_flight(a, sorter2, dwin, starts, steps, warmup, in_flight, compare, Wp,
                             r2["outs"], tiles2) if in_flight > 1 else None
        f1s = [f1_same_template(o1.sample, o1.template, o2.sample, o2.template, dt=1)
               for o1, o2 in zip(r["outs"][:nb], r2["outs"][:nb])]
        cmp = {"precision": compare,
               "value": round(steps * NT / fs / (fly2 or r2)["elapsed"], 2),
               "sequential_value": round(steps * NT / fs / r2["elapsed"], 2),
               "p50_ms": round(float(np.percentile(r2["lat_ms"], 50)), 3),
               "p99_ms": round(float(np.percentile(r2["lat_ms"], 99)), 3),
               "scores_gemm_ms": round(r2["stage_ms"]["scores_gemm"], 4),
               "whiten_ms": round(r2["stage_ms"]["whiten"], 4),
               "spike_f1_min": round(float(min(f1s)), 6),
               "f1_note": "spike F1 (+-1 sample, same template) of these spike lists against the "
                          "primary precision's, min over the recording's windows",
               "value_ratio": round(xrt / (steps * NT / fs / (fly2 or r2)["elapsed"]), 3)}
        sorter2.close()

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import ks_oracle as ko
        torch.set_num_threads(args.cpu_threads)
        ib = 1 % nb
        X = ko.padded_batch(s.rec, ib, NT, nt, torch.float32)
        c0 = time.perf_counter()
        ctc = ko.prepare_matching(torch.from_numpy(a["U"]), torch.from_numpy(a["wPCA"]))
        t_ctc = time.perf_counter() - c0
        c0 = time.perf_counter()
        ref = ko.sort_window(X, a, ko.window_start(ib, NT, nt), torch.float32, max_passes=50, ctc=ctc)
        csec = time.perf_counter() - c0
        got = sorter.sort_window(dwin[ib], starts[ib])
        f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
        cpu = {"value": round(NT / fs / csec, 4), "unit": "x realtime", "cores": torch.get_num_threads(), **host_cpu(),
               "kind": "port",
               "sample": f"one padded window (batch {ib}, {NT} samples = {NT / fs:.1f} s) through the "
                         f"FP32 torch CPU oracle, {csec:.1f} s (ctc built beforehand, {t_ctc:.1f} s); "
                         f"GPU-vs-CPU spike F1 {f1:.5f}"}
    sorter.close()
    return {
        "metric": METRIC,
        "value": round(xrt, 2), "unit": "x realtime", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round((fly or r)["elapsed"] / steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (seeded probe, templates and spikes; no recordings)",
        "config": {"workload": f"config {config}: {C} ch, {T} templates, {K} PCs, nt={nt}, NT={NT}-sample Kilosort "
                               f"windows (W'={int(Wp)}), {s.rec.shape[0] / a['fs']:.0f} s recording cycled, "
                               f"<=50 pursuit passes, windows resident in HBM",
                   "n_chan": C, "n_templates": T, "window": int(Wp), "NT": NT, "max_passes": 50,
                   "precision": args.precision, "tiles": tiles[0], "tile_ms": tile_ms, "update_grid": tiles[1],
                   "update_columns": tiles[2]},
        "p50_ms": round(float(np.percentile(r["lat_ms"], 50)), 3),
        "p99_ms": round(float(np.percentile(r["lat_ms"], 99)), 3),
        "latency_note": "p50/p99: one window at a time (the live per-batch case)",
        "windows_in_flight": in_flight,
        "sequential": {"value": round(xrt_seq, 2), "ms_per_window": round(r["elapsed"] / steps * 1e3, 3)},
        "in_flight": None if fly is None else {
            "windows": in_flight, "value": round(xrt, 2),
            "ms_per_window": round(fly["elapsed"] / steps * 1e3, 3),
            "spikes_equal_to_sequential": fly["equal"],
            "note": "value: windows spread over this many handles, one host thread each, so one "
                    "window's GEMMs overlap another's pursuit"},
        "host_input": host,
        "spikes_per_window": round(spikes / steps, 1),
        "passes_per_window": round(passes, 2),
        "stage_ms": {k: round(v, 4) for k, v in r["stage_ms"].items()},
        "roofline": roofline,
        "gemm_roofline": gemm,
        "compare": cmp,
        "cpu_baseline": cpu,
    }


def bench_adjoint(args):
    """Template-refinement backward on the drift workload (384 ch, 600 templates, a 2 s window at
    58-60 s of a probe drifting 0-20 um over 60 s): loss and d(loss)/d(U, wPCA) for the forward's
    spike set, the residual recomputed on every call."""
    import torch
    from livespikesorter_amd import synth
    from livespikesorter_amd.refine import SpikeSet
    from livespikesorter_amd.sorter import Sorter, pad_batch

    t0 = time.time()
    s = synth.make(4, seed=args.seed)
    a = s.art
    NT, nt = a["NT"], a["nt"]
    T, K, C = a["U"].shape
    w, ws = pad_batch(s.rec, 0, NT, nt)
    dev = torch.device("cuda:0")
    win = torch.from_numpy(w).to(dev)
    sorter = Sorter(a, max_window=win.shape[0], max_passes=100, drop_negative=False)
    x_wc = sorter.stage_whiten(sorter.stage_filter(win)).contiguous()
    sp = sorter.sort_window(win, ws)
    spikes = SpikeSet.from_spikes(sp)
    U = torch.from_numpy(a["U"]).to(dev).contiguous()
    wp = torch.from_numpy(a["wPCA"]).to(dev).contiguous()
    n = len(spikes)
    print(f"[bench] adjoint workload ready in {time.time() - t0:.1f}s: {n} spikes", file=sys.stderr, flush=True)
    gU, gW = torch.empty_like(U), torch.empty_like(wp)
    for _ in range(args.warmup):
        sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp, gU, gW)
    torch.cuda.synchronize()
    res_ms, adj_ms = [], []
    c0 = time.perf_counter()
    for _ in range(args.steps):
        loss, _, _ = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp, gU, gW)
        tm = sorter.refine_timings()
        res_ms.append(tm["residual"])
        adj_ms.append(tm["adjoint"])
    torch.cuda.synchronize()
    el = time.perf_counter() - c0
    W = win.shape[0]
    # adjoint: every spike's nt x C residual window read once, U read and dU written once
    adj_bytes = 4.0 * (n * nt * C + 2 * T * K * C)
    # residual: X read and R written once, the waves of every spike read once
    res_bytes = 4.0 * (2 * W * C + n * nt * C)
    adj_avg = float(np.mean(adj_ms))
    ach = adj_bytes / (adj_avg * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import ks_oracle as ko
        torch.set_num_threads(args.cpu_threads)
        X32 = x_wc.cpu().T.contiguous()
        U32 = U.cpu().requires_grad_(True)
        W32 = wp.cpu().requires_grad_(True)
        c1 = time.perf_counter()
        lc, _ = ko.residual_energy(X32, spikes.col.cpu().long(), spikes.tmp.cpu().long(), spikes.amp.cpu(), U32, W32)
        torch.autograd.grad(lc, (U32, W32))
        csec = time.perf_counter() - c1
        cpu = {"value": round(csec * 1e3, 1), "unit": "ms per backward", "cores": torch.get_num_threads(), **host_cpu(),
               "kind": "port",
               "sample": f"one loss + torch-autograd gradient of the FP32 oracle on the same 2 s window "
                         f"and {n} spikes ({csec:.1f} s)"}
    res = {
        "metric": "template-refinement backward d|R|^2/d(U, wPCA) on a 2 s drift window",
        "value": round(2.0 / (el / args.steps), 2), "unit": "x realtime (2 s window per backward)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded probe drifting 0-20 um over 60 s; window at 58-60 s)",
        "config": {"workload": "384 ch, 600 templates, 6 PCs, nt=61, 2 s window (W'=60122), spikes fixed from "
                               "the forward pass, residual recomputed per call", "spikes": n, "loss": loss},
        "device_ms": {"residual_p50": round(float(np.median(res_ms)), 4),
                      "adjoint_p50": round(float(np.median(adj_ms)), 4)},
        "roofline": {"kernel": "refine_adjoint_kernel (+ dW reduction)", "bound": "hbm",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                     "per_launch": f"{adj_bytes / 1e6:.1f} MB algorithmic (residual windows of {n} spikes, "
                                   f"U, dU), avg {adj_avg:.3f} ms"},
        "residual_roofline": {"achieved": round(res_bytes / (float(np.mean(res_ms)) * 1e-3) / 1e9, 1),
                              "unit": "GB/s", "algorithmic_MB": round(res_bytes / 1e6, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def bench_precision(args):
    """Config-3 precision report: 384 ch, 1000 templates, 60 s.  For FP32 and the BF16 / FP16 MFMA
    paths with and without the split (x3) correction: throughput over the whole recording, and
    whitening / score / amplitude / spike-list error against the FP64 CPU oracle on one window."""
    import torch
    from livespikesorter_amd import synth
    from livespikesorter_amd.sorter import Sorter, pad_batch
    from oracle import ks_oracle as ko

    t0 = time.time()
    s = synth.make(3, seed=args.seed)
    a = s.art
    NT, nt, C, T = a["NT"], a["nt"], s.rec.shape[1], a["U"].shape[0]
    nb = (s.rec.shape[0] + NT - 1) // NT
    wins, starts = [], []
    for ib in range(nb):
        w, ws = pad_batch(s.rec, ib, NT, nt)
        wins.append(w)
        starts.append(ws)
    dev = torch.device("cuda:0")
    dwin = torch.from_numpy(np.stack(wins)).to(dev)
    Wp = dwin.shape[1]
    print(f"[bench] precision workload ready in {time.time() - t0:.1f}s: C={C} T={T} windows={nb}",
          file=sys.stderr, flush=True)
    # FP64 oracle on one window: filtered/whitened data, scores and the spike list
    torch.set_num_threads(args.cpu_threads)
    ib = 1 % nb
    X = ko.padded_batch(s.rec, ib, NT, nt, torch.float64)
    c0 = time.perf_counter()
    hp = torch.from_numpy(a["hp_filter"]).double()
    wm = torch.from_numpy(a["drift"]).double() @ torch.from_numpy(a["Wrot"]).double()
    xf = ko.filter_batch(X, hp, None)
    xw64 = wm @ xf
    x32 = xw64.float()
    U64, W64 = torch.from_numpy(a["U"]).double(), torch.from_numpy(a["wPCA"]).double()
    B64 = ko.project(x32.double(), U64, W64).numpy().T
    ref = ko.sort_window(X, a, ko.window_start(ib, NT, nt), torch.float64, max_passes=50)
    print(f"[bench] FP64 oracle window in {time.perf_counter() - c0:.1f}s: {len(ref.samples)} spikes",
          file=sys.stderr, flush=True)

    def rel(a_, b_):
        return float(np.abs(np.asarray(a_, np.float64) - b_).max() / np.abs(b_).max())

    variants = {}
    for prec in ["fp32", "bf16", "bf16x3", "fp16", "fp16x3"]:
        so = Sorter(a, max_window=Wp, max_passes=50, precision=prec, max_spikes=1 << 16)
        r = run_path(so, dwin, starts, nb, min(args.warmup, nb))
        ew = rel(so.stage_whiten(xf.float().contiguous().to(dev)).cpu().numpy(), xw64.numpy().T)
        eb = rel(so.stage_scores(x32.T.contiguous().to(dev)).cpu().numpy(), B64)
        got = r["outs"][ib]
        f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
        d0 = {(int(t), int(y)): float(v) for t, y, v in zip(ref.samples, ref.templates, ref.amps)}
        errs = [abs(v - d0[(int(t), int(y))]) / abs(d0[(int(t), int(y))])
                for t, y, v in zip(got.sample, got.template, got.amplitude) if (int(t), int(y)) in d0]
        variants[prec] = {
            "x_realtime": round(nb * NT / a["fs"] / r["elapsed"], 2),
            "p50_ms": round(float(np.percentile(r["lat_ms"], 50)), 3),
            "p99_ms": round(float(np.percentile(r["lat_ms"], 99)), 3),
            "whiten_ms": round(r["stage_ms"]["whiten"], 4), "scores_gemm_ms": round(r["stage_ms"]["scores_gemm"], 4),
            "whitening_rel_err_vs_fp64": float(f"{ew:.3e}"), "scores_rel_err_vs_fp64": float(f"{eb:.3e}"),
            "spike_f1_vs_fp64": round(float(f1), 6),
            "amp_rel_err_vs_fp64": {"median": float(f"{np.median(errs):.3e}") if errs else None,
                                    "max": float(f"{max(errs):.3e}") if errs else None, "matched": len(errs)},
        }
        print(f"[bench] {prec}: {json.dumps(variants[prec])}", file=sys.stderr, flush=True)
        so.close()
    for prec in variants:
        if prec != "fp32":
            variants[prec]["speedup_vs_fp32"] = round(variants[prec]["x_realtime"] / variants["fp32"]["x_realtime"], 3)
    res = {
        "metric": "precision report: score / amplitude / spike-list error vs FP64 oracle and x-realtime, "
                  "384 ch, 1000 templates, 60 s",
        "value": variants["bf16x3"]["x_realtime"], "unit": "x realtime (bf16x3 path)", "n_gpus": 1,
        "steps": nb, "warmup": min(args.warmup, nb),
        "ms_per_step": round(NT / a["fs"] * 1e3 / variants["bf16x3"]["x_realtime"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic (seeded probe, templates and spikes; no recordings)",
        "config": {"workload": "384 ch, 1000 templates, 6 PCs, nt=61, 60 s in NT=60000 Kilosort windows, "
                               "<=50 pursuit passes; errors on window 1 against the FP64 CPU oracle",
                   "n_chan": C, "n_templates": T, "windows": nb},
        "fp64_oracle_spikes": int(len(ref.samples)),
        "variants": variants,
    }
    print(json.dumps(res))


FP64_PEAK_TFLOPS = 78.6


def _event_ms(fn, reps):
    """Per-call ms of fn() between HIP events on torch's current stream (the builders' stream)."""
    import torch
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        out.append(e0.elapsed_time(e1))
    return out


def bench_train(args):
    """Training-side builders in FP64 on the 384-channel drop-in workload (10 s, 5 batches of
    NT = 60000): covariance of a filtered batch (the get_whitening_matrix inner loop), local ZCA
    whitening, iKxx, the per-batch drift matrix and its application to the whitening operator,
    isolated-peak snippets of a whitened batch and the wPCA eigen-solve."""
    import torch
    from livespikesorter_amd import synth, train
    from livespikesorter_amd.sorter import Sorter, pad_batch

    t0 = time.time()
    s = synth.make(2, seed=args.seed)
    a = s.art
    NT, nt = a["NT"], a["nt"]
    C = s.rec.shape[1]
    nb = (s.rec.shape[0] + NT - 1) // NT
    dev = torch.device("cuda:0")
    filt = train.FilterStage(C, a["hp_filter"], a["xc"], a["yc"], NT + 2 * nt, nt=nt)
    xs = [filt(torch.from_numpy(pad_batch(s.rec, j, NT, nt)[0]).to(dev)) for j in range(nb)]
    filt.close()
    sorter = Sorter(a)
    xw = [sorter.stage_whiten(x) for x in xs]
    xc = a["xc"].astype(np.float64)
    yc = a["yc"].astype(np.float64)
    print(f"[bench] train workload ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    reps = max(args.steps, 3)
    CC = torch.zeros((C, C), dtype=torch.float64, device=dev)
    it = iter(range(1 << 30))
    cov_ms = _event_ms(lambda: train.covariance_add(xs[next(it) % nb][:, nt:-nt], CC), reps)
    CC.zero_()
    for x in xs:
        train.covariance_add(x[:, nt:-nt], CC)
    CC /= nb
    zca_ms = _event_ms(lambda: train.whitening_local(CC, xc, yc, 32), max(reps // 5, 3))
    inv = {}

    def _inv():
        inv["iK"], inv["it"] = train.drift_inverse(xc, yc, 20.0, 0.01)
    inv_ms = _event_ms(_inv, max(reps // 10, 3))
    yblk = np.linspace(float(yc.min()) + 200.0, float(yc.max()) - 200.0, 5)
    dsh = np.array([3.0, -1.5, 4.0, 0.5, -2.0])
    dm_ms = _event_ms(lambda: train.get_drift_matrix(xc, yc, dsh, inv["iK"], 20.0, yblk=yblk), reps)
    M = train.get_drift_matrix(xc, yc, dsh, inv["iK"], 20.0, yblk=yblk)

    def _set():
        sorter.set_drift(M)
        torch.cuda.synchronize()
    set_ms = _event_ms(_set, reps)
    it2 = iter(range(1 << 30))
    snip_ms = _event_ms(lambda: train.extract_snippets(xw[next(it2) % nb]), reps)
    parts = [train.extract_snippets(x)[0] for x in xw]
    clips = torch.cat(parts)
    ncl = clips.shape[0]
    pca_ms = _event_ms(lambda: train.wpca(clips, 6), max(reps // 5, 3))
    rep = (500000 + ncl - 1) // max(ncl, 1)
    big = clips.repeat(rep, 1)[:500000].contiguous()
    pca_big_ms = _event_ms(lambda: train.wpca(big, 6), 3)
    flops = float(C) * (C + 1) * NT
    cov_med = float(np.median(cov_ms))
    ach = flops / (cov_med * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import train_oracle as to
        X = xs[0][:, nt:-nt].cpu().numpy()
        c1 = time.perf_counter()
        to.covariance([X])
        cov_cpu = time.perf_counter() - c1
        CCn = CC.cpu().numpy()
        c1 = time.perf_counter()
        to.whitening_local(CCn, xc, yc, 32)
        zca_cpu = time.perf_counter() - c1
        c1 = time.perf_counter()
        iKn = to.drift_inverse(xc, yc, 20.0, 0.01)
        inv_cpu = time.perf_counter() - c1
        c1 = time.perf_counter()
        to.drift_matrix(xc, yc, dsh, iKn, 20.0, yblk=yblk)
        dm_cpu = time.perf_counter() - c1
        c1 = time.perf_counter()
        to.extract_snippets(xw[0].cpu().numpy().T)
        snip_cpu = time.perf_counter() - c1
        c1 = time.perf_counter()
        to.wpca(clips.cpu().numpy(), 6)
        pca_cpu = time.perf_counter() - c1
        cpu = {"value": round(2.0 / cov_cpu, 2), "unit": "x realtime (covariance of one 2 s batch)",
               "cores": int(os.environ.get("OMP_NUM_THREADS", "16")), **host_cpu(), "kind": "port",
               "sample": f"FP64 numpy oracle on the host: covariance of batch 0 {cov_cpu * 1e3:.1f} ms, "
                         f"whitening_local {zca_cpu * 1e3:.0f} ms, iKxx {inv_cpu * 1e3:.0f} ms, drift matrix "
                         f"{dm_cpu * 1e3:.1f} ms, snippets of batch 0 {snip_cpu * 1e3:.0f} ms, wPCA of {ncl} "
                         f"snippets {pca_cpu * 1e3:.1f} ms"}
    res = {
        "metric": "FP64 covariance of filtered 2 s batches, 384 ch (the get_whitening_matrix inner loop)",
        "value": round(2.0 / (cov_med * 1e-3), 1), "unit": "x realtime (2 s batch per call)",
        "n_gpus": 1, "steps": reps, "warmup": 1, "ms_per_step": round(cov_med, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded 384-ch probe, 600 units, 10 s)",
        "config": {"workload": "training-side builders: 384 ch, NT=60000 (5 filtered batches), nrange=32, "
                               "sig_interp=20 um, 5 drift blocks, Th_single_ch=6, 6 PCs"},
        "builders_ms_p50": {
            "covariance_add": round(cov_med, 4),
            "whitening_local_384x32": round(float(np.median(zca_ms)), 3),
            "drift_inverse": round(float(np.median(inv_ms)), 3), "newton_schulz_steps": inv["it"],
            "drift_matrix_5_blocks": round(float(np.median(dm_ms)), 4),
            "set_drift": round(float(np.median(set_ms)), 4),
            "extract_snippets_per_batch": round(float(np.median(snip_ms)), 4),
            f"wpca_{ncl}_snippets": round(float(np.median(pca_ms)), 3),
            "wpca_500000_snippets": round(float(np.median(pca_big_ms)), 3)},
        "roofline": {"kernel": "gemm64_kernel<float,float,true> + reduce64 (symmetric FP64 rank-NT update)",
                     "bound": "fp64", "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                     "per_launch": f"{flops / 1e9:.2f} GFLOP algorithmic (C(C+1) NT), p50 {cov_med:.3f} ms"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (stream: 100-ms batches, default all after the warmup; window: windows, "
                         "default 50)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16", "fp16", "bf16x3", "fp16x3"])
    ap.add_argument("--compare", default="fp32", help="window workload: second precision reported beside the first")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-batches", type=int, default=80, help="stream workload: batches the CPU oracle sorts")
    ap.add_argument("--workload", default="stream", choices=["stream", "window", "adjoint", "precision", "train"],
                    help="stream: config 5, 4 uneven probes, 100 ms batches (default; the config-2 windows are "
                         "reported beside it unless --no-window); window: 384-ch drop-in 2-s windows; "
                         "adjoint: template-refinement backward on a 2 s drift window; "
                         "precision: config-3 precision report (1000 templates, 60 s); "
                         "train: training-side FP64 builders (whitening, drift, wPCA)")
    ap.add_argument("--seconds", type=float, default=60.0, help="stream workload: seconds per probe")
    ap.add_argument("--refine-seconds", type=float, default=10.0,
                    help="stream workload: one template-refinement round every this many seconds (0 = off)")
    ap.add_argument("--no-window", action="store_true", help="stream workload: skip the nested 2-s window figures")
    ap.add_argument("--roofline-only", action="store_true",
                    help="stream workload: only the update-kernel roofline windows (counter runs)")
    ap.add_argument("--pin-tuning", default=None,
                    help="stream workload: JSON list, per probe [tiles, plain_tiles, update_grid, update_columns, "
                         "first_passes, persistent_grid, persistent_after], used instead of the autotuner")
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3],
                    help="window workload: synthetic configuration (1: 64 ch / 32 units, the reference's CPU-runnable "
                         "case; 2: 384 ch / 600 units, default; 3: 384 ch / 1000 units)")
    ap.add_argument("--in-flight", type=int, default=3,
                    help="window workload: windows in flight on separate handles for the value (1 = one at a time)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="instead of the bench, copy a tensor of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each probe stream is "
              "sorted independently on its own GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4}))
        return

    if args.workload == "stream":
        res = bench_stream(args)
        if not args.no_window and not args.roofline_only:
            res["window_2s"] = bench_window(args, steps=20, warmup=5)
        print(json.dumps(res))
        return
    if args.steps is None:
        args.steps = 50
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "train":
        return bench_train(args)
    steps = 50 if args.steps is None else args.steps
    print(json.dumps(bench_window(args, steps, args.warmup, config=args.config, in_flight=args.in_flight,
                                  compare=args.compare)))


if __name__ == "__main__":
    main()
