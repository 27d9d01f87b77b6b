"""Multi-probe streaming sorter (the batching loop of OnlineSpikesV2::runSpikeSorting,
src/SpikeSorter/OnlineSpikesV2.cu:632-820, for several probes on one sample clock), over liblss's
batcher: samples are pushed per probe as they arrive; every step sorts the next `batch` samples of
every probe in one replayed HIP graph (probes as concurrent branches)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import Spikes, check, lib
from .sorter import Sorter, alloc_spikes


class StreamingSorter:
    def __init__(self, sorters: list[Sorter], batch: int = 3000, use_graph: bool = True, refine_every: int = 0,
                 run_ahead: bool = True, grouped: bool = True, pipelined: bool = True):
        """refine_every: one template-refinement step per probe every this many batches (drift
        tracking, lss_refine_templates on that batch's window); 0 = never.  run_ahead: a step whose
        next batch is already pushed preprocesses it beside its own pursuit (lss_stream_set_run_ahead;
        same results).  grouped: all probes' pursuit tails as one launch per stage when every probe's
        pursuit is persistent (lss_stream_set_grouped; same results).  pipelined: a grouped step
        that preprocessed its next batch launches that batch's pursuit before it returns
        (lss_stream_set_pipelined; same results)."""
        self.sorters = list(sorters)
        arr = (ct.c_void_p * len(self.sorters))(*[s._h.value for s in self.sorters])
        h = ct.c_void_p()
        check(lib().lss_stream_create(arr, len(self.sorters), int(batch), int(use_graph), ct.byref(h)),
              "lss_stream_create")
        self._h = h
        self.batch = int(batch)
        # device tensors referenced by the batcher, per probe: (tensor, first sample, end sample);
        # released once the batcher has trimmed them (end <= t0 - nt after a step)
        self._held = {}
        self._recv = [0] * len(self.sorters)
        self._t0 = 0
        if refine_every:
            check(lib().lss_stream_set_refine(h, int(refine_every)), "lss_stream_set_refine")
        check(lib().lss_stream_set_run_ahead(h, int(bool(run_ahead))), "lss_stream_set_run_ahead")
        check(lib().lss_stream_set_grouped(h, int(bool(grouped))), "lss_stream_set_grouped")
        check(lib().lss_stream_set_pipelined(h, int(bool(pipelined))), "lss_stream_set_pipelined")
        arr_out = (Spikes * len(self.sorters))()
        self._bufs = []
        for i, s in enumerate(self.sorters):
            sp, bufs = alloc_spikes(min(s.params.max_spikes, 1 << 16), s.nC, s.K)
            arr_out[i] = sp
            self._bufs.append(bufs)
        self._arr_out = arr_out

    def push(self, probe: int, samples):
        """Host samples (numpy int16 [n][C]) are copied in; device samples (a contiguous torch int16
        tensor on this GPU) are referenced: this object keeps the tensor alive, and the stream must
        keep receiving device samples for that probe."""
        if hasattr(samples, "data_ptr"):
            if not samples.is_contiguous():
                raise ValueError("device samples must be contiguous")
            import torch
            cs = torch.cuda.current_stream()
            # the samples are complete before a step copies them: the batcher's streams are blocking
            # streams, ordered after the work of torch's default (legacy) stream; a stream of the
            # caller's is waited for here
            if cs.cuda_stream != 0:
                cs.synchronize()
            n = int(samples.shape[0])
            check(lib().lss_stream_push_device(self._h, int(probe), samples.data_ptr(), n),
                  "lss_stream_push_device")
            if n:
                self._held.setdefault(probe, []).append((samples, self._recv[probe], self._recv[probe] + n))
            self._recv[probe] += n
            return
        x = np.ascontiguousarray(samples, np.int16)
        check(lib().lss_stream_push(self._h, int(probe), x.ctypes.data, x.shape[0]), "lss_stream_push")
        self._recv[probe] += int(x.shape[0])

    def step(self):
        """Spike lists of the next batch for every probe, or None if samples are missing."""
        ready = ct.c_int32()
        check(lib().lss_stream_step(self._h, self._arr_out, ct.byref(ready)), "lss_stream_step")
        if not ready.value:
            return None
        out = [Sorter._list(self._arr_out[i], self._bufs[i]) for i in range(len(self.sorters))]
        # the batcher keeps samples from t0 - nt on (lss_stream_step): drop the tensors before that
        self._t0 += self.batch
        for probe, held in self._held.items():
            keep_from = max(0, self._t0 - self.sorters[probe].nt)
            while held and held[0][2] <= keep_from:
                held.pop(0)
        return out

    def ran_ahead(self) -> int:
        """Steps so far that preprocessed their next batch (lss_stream_ran_ahead)."""
        n = ct.c_int64()
        check(lib().lss_stream_ran_ahead(self._h, ct.byref(n)), "lss_stream_ran_ahead")
        return n.value

    def pipelined(self) -> int:
        """Steps so far that launched their next batch's pursuit early (lss_stream_pipelined)."""
        n = ct.c_int64()
        check(lib().lss_stream_pipelined(self._h, ct.byref(n)), "lss_stream_pipelined")
        return n.value

    def grouped(self) -> bool:
        """Whether the steps run grouped (lss_stream_grouped; decided at the first step)."""
        g = ct.c_int32()
        check(lib().lss_stream_grouped(self._h, ct.byref(g)), "lss_stream_grouped")
        return bool(g.value)

    def profile(self, on: bool = True):
        """Measurement mode of grouped steps (lss_stream_profile)."""
        check(lib().lss_stream_profile(self._h, int(bool(on))), "lss_stream_profile")

    def last_profile(self):
        """(GPU-clock span of the last step's grouped persistent pursuit in us, [spikes], [passes],
        [workgroups] per probe) (lss_stream_last_profile)."""
        n = len(self.sorters)
        us = ct.c_double()
        sp, ps, gr = (ct.c_int32 * n)(), (ct.c_int32 * n)(), (ct.c_int32 * n)()
        check(lib().lss_stream_last_profile(self._h, ct.byref(us), sp, ps, gr), "lss_stream_last_profile")
        return us.value, list(sp), list(ps), list(gr)

    def held(self, probe: int) -> int:
        """Device tensors of a probe still referenced for the batcher."""
        return len(self._held.get(probe, []))

    def last_ms(self) -> float:
        ms, n = ct.c_float(), ct.c_int64()
        check(lib().lss_stream_stats(self._h, ct.byref(ms), ct.byref(n)), "lss_stream_stats")
        return ms.value

    def breakdown(self) -> dict:
        """Host wall-time split of the last step (ms)."""
        ms = np.zeros(6, np.float32)
        check(lib().lss_stream_breakdown(self._h, ms.ctypes.data), "lss_stream_breakdown")
        return dict(zip(["assemble", "segment_A", "extra_passes", "segment_C", "emit", "total"], ms.tolist()))

    def last_refine(self, probe: int) -> dict:
        """The probe's last refinement step: loss before/after, step length, gradient norm."""
        info = np.zeros(4, np.float64)
        check(lib().lss_stream_last_refine(self._h, int(probe), info.ctypes.data), "lss_stream_last_refine")
        return dict(zip(["loss_before", "loss_after", "step", "grad_norm"], info.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            check(lib().lss_stream_destroy(self._h), "lss_stream_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


UPDATE_GRIDS = [256, 512, 1024, 2048, 4096, 8192]


def autotune(sorter: Sorter, window_dev, reps: int = 2):
    """Picks the update layout (2 or 1 score columns per wavefront) and grid (fastest of
    256..8192 workgroups) and the first-segment pass count for `sorter` on a device window;
    returns (grid, first_passes, {(columns per wavefront, grid): ms})."""
    ms = np.zeros(12, np.float32)
    g, p = ct.c_int32(), ct.c_int32()
    check(lib().lss_autotune(sorter._h, window_dev.data_ptr(), window_dev.shape[0], int(reps), ms.ctypes.data,
                             ct.byref(g), ct.byref(p)), "lss_autotune")
    return g.value, p.value, {(2 - i // 6, UPDATE_GRIDS[i % 6]): float(v) for i, v in enumerate(ms)}


def update_columns(sorter: Sorter) -> int:
    """Score columns per wavefront of the sorter's update kernel (1 or 2)."""
    c = ct.c_int32()
    check(lib().lss_get_update_columns(sorter._h, ct.byref(c)), "lss_get_update_columns")
    return c.value


PERSIST_GRIDS = [0, 64, 128, 256, 384]
PERSIST_FIRST = [0, 2, 4, "P1"]   # P1: the first-segment pass count autotune() chose


def autotune_persistent(sorter: Sorter, window_dev, reps: int = 2, max_grid: int = 0):
    """Picks the pursuit schedule of a short window (<= 4096 samples): per-pass kernels (grid 0) or
    the persistent kernel at 64..384 workgroups (at most max_grid; 0 = the device limit) after 0, 2,
    4 or P1 leading per-pass rounds (P1 from autotune(); after it the persistent kernel usually finds
    the pursuit converged, which makes a streaming step one graph).  Returns (grid, first_passes, {(grid, first): ms}) over the
    candidates that ran."""
    ms = np.zeros(20, np.float32)
    g, f = ct.c_int32(), ct.c_int32()
    check(lib().lss_autotune_persistent(sorter._h, window_dev.data_ptr(), window_dev.shape[0], int(reps),
                                        int(max_grid), ms.ctypes.data, ct.byref(g), ct.byref(f)),
          "lss_autotune_persistent")
    table = {(gg, ff): float(ms[4 * i + j]) for i, gg in enumerate(PERSIST_GRIDS)
             for j, ff in enumerate(PERSIST_FIRST) if ms[4 * i + j] >= 0}
    return g.value, f.value, table


def batcher_grid_shares(sorters: list[Sorter], W: int) -> list[int]:
    """The persistent-pursuit workgroups the batcher allows each probe (lss_stream_create's split:
    probe i gets sqrt(T_i C_i) / sum_j sqrt(T_j C_j) of the device's limit for its kernel)."""
    w = [float(np.sqrt(s.T * s.C)) for s in sorters]
    tot = sum(w)
    return [max(1, int(np.floor(s.persistent(W)[2] * wi / tot))) for s, wi in zip(sorters, w)]


def autotune_stream(sorters: list[Sorter], recordings, batch: int = 3000, steps: int = 12, warmup: int = 3,
                    options=None):
    """Picks every probe's pursuit schedule jointly, by timing whole streaming steps: probes share
    the device inside a step, so the schedule each is fastest with alone (autotune_persistent) is
    not always the fastest together (a persistent pursuit keeps its workgroups for the whole
    pursuit, per-pass kernels take whatever the other probes leave free).  options[i] lists probe
    i's candidate (persistent grid, first per-pass rounds) pairs; by default the probe's current
    setting, (its grid, the first-segment pass count autotune chose) and (its grid, 0): when every
    probe's pursuit is persistent from its first pass the steps run grouped
    (lss_stream_set_grouped).  Every combination (at most
    64) runs `warmup` + `steps` graph-replayed steps of a fresh batcher over the device recordings
    (torch int16 [n][C], one per probe, each at least (warmup + steps + 1) * batch + nt samples); the
    combination with the lowest median step time is set on the sorters.  Results do not depend on
    the schedule.  Returns (chosen [(grid, first)], {combination: median ms})."""
    import itertools
    import statistics
    import torch

    if options is None:
        options = []
        for s in sorters:
            g, f, _ = s.persistent(batch + 2 * s.nt)
            _, p1 = s.tuning()
            opts = [(g, f)]
            if p1 > 0 and (g, p1) not in opts:
                opts.append((g, p1))
            if g > 0 and (g, 0) not in opts:
                opts.append((g, 0))
            options.append(opts)
    combos = list(itertools.product(*options))
    if len(combos) > 64:
        raise ValueError(f"{len(combos)} combinations: pass options with at most 64")
    need = (warmup + steps + 1) * batch
    for s, r in zip(sorters, recordings):
        if r.shape[0] < need + 2 * s.nt:
            raise ValueError(f"a recording of {r.shape[0]} samples is shorter than the {need + 2 * s.nt} needed")
    table = {}
    for combo in combos:
        for s, (g, f) in zip(sorters, combo):
            s.set_persistent(g, f)
        times = []
        with StreamingSorter(sorters, batch=batch, use_graph=True) as st:
            for i, r in enumerate(recordings):
                st.push(i, r[:need + 2 * sorters[i].nt])
            for k in range(warmup + steps):
                torch.cuda.synchronize()
                if st.step() is None:
                    raise RuntimeError("autotune_stream: a step found samples missing")
                if k >= warmup:
                    times.append(st.last_ms())
        table[combo] = statistics.median(times)
    best = min(table, key=table.get)
    for s, (g, f) in zip(sorters, best):
        s.set_persistent(g, f)
    return list(best), table
