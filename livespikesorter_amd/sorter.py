"""Host-side interface of the MI355X sorter, mirroring LiveSpikeSorter's SpikeSorter module.

``Sorter`` is the Python face of one ``OnlineSpikesV2`` (src/SpikeSorter/OnlineSpikesV2.h:29-34):
it is built once from Kilosort4 outputs, owns its device memory, and turns padded int16 windows
into spikes (sample, template, amplitude, PC features, position) - the work of
``kilosortMatchingPursuit`` + ``computeClosestClusters`` (OnlineSpikesV2.cu:823-971) with
Kilosort4's ``run_matching``/``extract`` semantics (kilosort/template_matching.py:28-85, 128-202).
Every call goes through liblss's C ABI; there is no CPU path.

Device arrays are torch CUDA tensors (PyTorch supplies device memory only); host arrays are numpy.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Artifacts, Params, Spikes, check, lib


@dataclass
class SpikeList:
    sample: np.ndarray       # int64 recording sample of each spike
    template: np.ndarray     # int32
    amplitude: np.ndarray    # float32
    features: np.ndarray     # float32 [n][nC][K]
    x: np.ndarray
    y: np.ndarray
    column: np.ndarray       # int32 peak column in the padded window
    passes: int = 0

    def __len__(self):
        return len(self.sample)

    @staticmethod
    def concat(parts):
        if not parts:
            return SpikeList(*(np.zeros(0) for _ in range(7)))
        f = lambda k: np.concatenate([getattr(p, k) for p in parts])
        return SpikeList(f("sample"), f("template"), f("amplitude"), f("features"), f("x"), f("y"),
                         f("column"), sum(p.passes for p in parts))

    def sorted(self):
        o = np.argsort(self.sample, kind="stable")
        return SpikeList(self.sample[o], self.template[o], self.amplitude[o], self.features[o],
                         self.x[o], self.y[o], self.column[o], self.passes)


def alloc_spikes(n: int, nC: int, K: int):
    """An lss_spikes struct over freshly allocated host arrays of capacity n."""
    bufs = dict(sample=np.zeros(n, np.int64), template=np.zeros(n, np.int32),
                amplitude=np.zeros(n, np.float32), features=np.zeros((n, nC, K), np.float32),
                x=np.zeros(n, np.float32), y=np.zeros(n, np.float32), column=np.zeros(n, np.int32))
    s = Spikes()
    s.capacity = n
    s.sample = bufs["sample"].ctypes.data
    s.template_id = bufs["template"].ctypes.data
    s.amplitude = bufs["amplitude"].ctypes.data
    s.features = bufs["features"].ctypes.data
    s.x = bufs["x"].ctypes.data
    s.y = bufs["y"].ctypes.data
    s.column = bufs["column"].ctypes.data
    return s, bufs


TIMING_KEYS = ["car", "highpass", "whiten", "project", "scores_gemm", "pursuit", "features", "total"]


def default_params(**over) -> Params:
    p = Params()
    check(lib().lss_default_params(ct.byref(p)), "lss_default_params")
    for k, v in over.items():
        if k == "precision" and isinstance(v, str):
            v = _lib.PRECISION[v]
        setattr(p, k, v)
    return p


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(t) -> int:
    """Device pointer of a torch tensor (must be contiguous)."""
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def _ready():
    """Waits for torch's pending work on the current stream, so that device tensors handed to the
    library (which works on its own streams) are complete."""
    import torch
    torch.cuda.current_stream().synchronize()


class Sorter:
    """One sorter handle on the current HIP device."""

    def __init__(self, art: dict, max_window: int = 60122, max_passes: int = 50,
                 precision: str = "fp32", max_spikes: int = 1 << 18, do_car: bool = True,
                 drop_negative: bool = True):
        U = _f32(art["U"])
        T, K, C = U.shape
        wpca = _f32(art["wPCA"])
        opt = lambda key, dt: None if art.get(key) is None else np.ascontiguousarray(np.asarray(art[key]), dtype=dt)
        self._keep = dict(U=U, wPCA=wpca, Wrot=_f32(art["Wrot"]),
                          drift=None if art.get("drift") is None else _f32(art["drift"]),
                          hp=_f32(art["hp_filter"]), xc=_f32(art["xc"]), yc=_f32(art["yc"]),
                          iCC=opt("iCC", np.int32), iU=opt("iU", np.int32), Ucc=opt("Ucc", np.float32))
        k = self._keep
        self.params = default_params(
            n_chan=C, n_templates=T, n_pcs=K, nt=wpca.shape[1], nt0min=int(art.get("nt0min", 20)),
            n_nearest=int(art.get("nC", 10)), max_window=int(max_window), max_passes=int(max_passes),
            max_spikes=int(max_spikes), do_car=int(do_car), precision=precision,
            drop_negative=int(drop_negative), th_learned=float(art.get("Th", 8.0)),
            fs=float(art.get("fs", 30000.0)), dup_bins=int(art.get("dup_bins", 7)))
        nC = self.params.n_nearest
        for key, shape in (("iCC", (nC, C)), ("iU", (T,)), ("Ucc", (nC, T, K))):
            if k[key] is not None and k[key].shape != shape:
                raise ValueError(f"{key} must have shape {shape}, not {k[key].shape}")
        a = Artifacts()
        a.U = k["U"].ctypes.data
        a.wPCA = k["wPCA"].ctypes.data
        a.Wrot = k["Wrot"].ctypes.data
        a.drift = None if k["drift"] is None else k["drift"].ctypes.data
        a.hp_filter = k["hp"].ctypes.data
        a.hp_len = len(k["hp"])
        a.xc = k["xc"].ctypes.data
        a.yc = k["yc"].ctypes.data
        for key in ("iCC", "iU", "Ucc"):
            setattr(a, key, None if k[key] is None else k[key].ctypes.data)
        h = ct.c_void_p()
        check(lib().lss_create(ct.byref(self.params), ct.byref(a), ct.byref(h)), "lss_create")
        self._h = h
        self.C, self.T, self.K, self.nt = C, T, K, wpca.shape[1]
        self.nC = self.params.n_nearest

    @classmethod
    def from_dir(cls, path: str, **kw):
        obj = cls.__new__(cls)
        obj.params = default_params(**kw)
        h = ct.c_void_p()
        check(lib().lss_create_from_dir(str(path).encode(), ct.byref(obj.params), ct.byref(h)),
              "lss_create_from_dir")
        obj._h = h
        check(lib().lss_get_params(h, ct.byref(obj.params)), "lss_get_params")
        p = obj.params
        obj.C, obj.T, obj.K, obj.nt, obj.nC = p.n_chan, p.n_templates, p.n_pcs, p.nt, p.n_nearest
        obj._keep = {}
        return obj

    # -------------------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            check(lib().lss_destroy(self._h), "lss_destroy")
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

    # -------------------------------------------------------------------------------------------
    def _out(self, cap):
        """Host output buffers, allocated once per capacity and reused across calls."""
        n = int(cap)
        cached = getattr(self, "_out_cache", None)
        if cached is not None and cached[0] == n:
            return cached[1], cached[2]
        s, bufs = alloc_spikes(n, self.nC, self.K)
        self._out_cache = (n, s, bufs)
        return s, bufs

    @staticmethod
    def _list(s, bufs):
        n = int(s.n)
        return SpikeList(bufs["sample"][:n].copy(), bufs["template"][:n].copy(),
                         bufs["amplitude"][:n].copy(), bufs["features"][:n].copy(),
                         bufs["x"][:n].copy(), bufs["y"][:n].copy(), bufs["column"][:n].copy(),
                         int(s.passes))

    def sort_window(self, window, window_start: int, capacity: int | None = None) -> SpikeList:
        """window: int16 [W][C] numpy (host) or torch CUDA tensor (device)."""
        cap = capacity or self.params.max_spikes
        s, bufs = self._out(cap)
        if isinstance(window, np.ndarray):
            w = np.ascontiguousarray(window, dtype=np.int16)
            st = lib().lss_sort_window(self._h, w.ctypes.data, w.shape[0], 0, int(window_start), ct.byref(s))
        else:
            _ready()
            st = lib().lss_sort_window(self._h, _ptr(window), window.shape[0], 1, int(window_start), ct.byref(s))
        check(st, "lss_sort_window")
        return self._list(s, bufs)

    def sort_recording(self, rec: np.ndarray, NT: int = 60000) -> SpikeList:
        """All NT batches of an int16 recording [n][C], padded as Kilosort does, time-sorted
        (kilosort/template_matching.py:44-85)."""
        rec = np.ascontiguousarray(rec, dtype=np.int16)
        n = rec.shape[0]
        nb = (n + NT - 1) // NT
        win = np.zeros((NT + 2 * self.nt, self.C), np.int16)
        parts = []
        for ib in range(nb):
            ws = ct.c_int64()
            check(lib().lss_pad_batch(rec.ctypes.data, n, self.C, NT, self.nt, ib, win.ctypes.data,
                                      ct.byref(ws)), "lss_pad_batch")
            parts.append(self.sort_window(win, ws.value))
        return SpikeList.concat(parts).sorted()

    def load_info(self) -> dict:
        """Which extraction tables (iCC, iU, Ucc) were given rather than built, and the comparison of
        an oss_input directory's ctc.npy with the FP64 rebuild (max |diff|, max |ctc|)."""
        info = _lib.LoadInfo()
        check(lib().lss_get_load_info(self._h, ct.byref(info)), "lss_get_load_info")
        return {f: getattr(info, f) for f, _ in info._fields_}

    def update_times(self):
        """Measurement mode: (event ms, GPU-clock ms, spikes) per update launch of the last window."""
        n = ct.c_int32()
        check(lib().lss_last_update_times(self._h, None, None, None, 0, ct.byref(n)), "lss_last_update_times")
        ev = np.zeros(n.value, np.float32)
        gpu = np.zeros(n.value, np.float32)
        sp = np.zeros(n.value, np.int32)
        check(lib().lss_last_update_times(self._h, ev.ctypes.data, gpu.ctypes.data, sp.ctypes.data, n.value,
                                          ct.byref(n)), "lss_last_update_times")
        return ev, gpu, sp

    def persistent_times(self):
        """Measurement mode: (GPU-clock ms, spikes applied, passes, workgroups) of the last window's
        persistent pursuit; zeros when the window took the per-pass kernels only."""
        ms, sp, ps, g = ct.c_float(), ct.c_int32(), ct.c_int32(), ct.c_int32()
        check(lib().lss_last_persistent_stats(self._h, ct.byref(ms), ct.byref(sp), ct.byref(ps), ct.byref(g)),
              "lss_last_persistent_stats")
        return ms.value, sp.value, ps.value, g.value

    def timings(self):
        ms = np.zeros(8, np.float32)
        check(lib().lss_last_timings(self._h, ms.ctypes.data, 8), "lss_last_timings")
        return dict(zip(TIMING_KEYS, ms.tolist()))

    def set_tiles(self, whiten: int, project: int, scores: int):
        """Tile shapes of the whitening GEMM, wPCA projection and score GEMM (0..3; identical results)."""
        check(lib().lss_set_tiles(self._h, int(whiten), int(project), int(scores)), "lss_set_tiles")

    def tiles(self):
        t = np.zeros(3, np.int32)
        check(lib().lss_get_tiles(self._h, t.ctypes.data), "lss_get_tiles")
        return tuple(int(x) for x in t)

    def autotune_tiles(self, window_dev, reps: int = 3):
        """Times every tile shape of the three contractions on a device window and keeps the fastest;
        returns ((whiten, project, scores), {stage: [ms per shape]})."""
        ms = np.zeros(12, np.float32)
        best = np.zeros(3, np.int32)
        _ready()
        check(lib().lss_autotune_tiles(self._h, _ptr(window_dev), window_dev.shape[0], int(reps), ms.ctypes.data,
                                       best.ctypes.data), "lss_autotune_tiles")
        names = ["whiten", "project", "scores"]
        return tuple(int(x) for x in best), {names[i]: [round(float(v), 4) for v in ms[4 * i:4 * i + 4]]
                                             for i in range(3)}

    def set_persistent(self, grid: int, first_passes: int = 0):
        """Pursuit of windows <= 4096 samples: the first `first_passes` passes on the per-pass kernels,
        the rest in the persistent kernel of `grid` workgroups (grid 0 = per-pass kernels only)."""
        check(lib().lss_set_persistent(self._h, int(grid), int(first_passes)), "lss_set_persistent")

    def tuning(self):
        """(update grid, first-segment pass count) of the per-pass pursuit (lss_autotune / lss_set_tuning)."""
        g, f = ct.c_int32(), ct.c_int32()
        check(lib().lss_get_tuning(self._h, ct.byref(g), ct.byref(f)), "lss_get_tuning")
        return g.value, f.value

    def persistent(self, W: int):
        """(grid, first_passes, device co-resident limit for windows of W samples; 0 if W is too long)."""
        g, f, m = ct.c_int32(), ct.c_int32(), ct.c_int32()
        check(lib().lss_get_persistent(self._h, int(W), ct.byref(g), ct.byref(f), ct.byref(m)), "lss_get_persistent")
        return g.value, f.value, m.value

    # ---- template refinement -------------------------------------------------------------------
    def residual_energy(self, x_wc, col, tmp, amp, U, wpca) -> float:
        """sum R^2 for the fixed spike set (col sorted ascending); all arguments device tensors:
        x_wc float32 [W][C], col/tmp int32 [n], amp float32 [n], U float32 [T][K][C], wpca [K][nt]."""
        loss = ct.c_double()
        _ready()
        check(lib().lss_residual_energy(self._h, _ptr(x_wc), x_wc.shape[0], _ptr(col), _ptr(tmp), _ptr(amp),
                                        col.numel(), _ptr(U), _ptr(wpca), ct.byref(loss)), "lss_residual_energy")
        return loss.value

    def residual_energy_grad(self, x_wc, col, tmp, amp, U, wpca, grad_U=None, grad_wpca=None):
        """(loss, dL/dU, dL/dwPCA) from the adjoint kernel; the residual is recomputed here."""
        import torch
        gU = torch.empty_like(U) if grad_U is None else grad_U
        gW = torch.empty_like(wpca) if grad_wpca is None else grad_wpca
        loss = ct.c_double()
        _ready()
        check(lib().lss_residual_energy_grad(self._h, _ptr(x_wc), x_wc.shape[0], _ptr(col), _ptr(tmp), _ptr(amp),
                                             col.numel(), _ptr(U), _ptr(wpca), ct.byref(loss), _ptr(gU), _ptr(gW)),
              "lss_residual_energy_grad")
        return loss.value, gU, gW

    def refine_templates(self) -> dict:
        """One refinement step of this sorter's templates on the last window it sorted (adopted only
        if it lowers that window's residual energy); returns loss before/after, step, |gradient|."""
        info = np.zeros(4, np.float64)
        check(lib().lss_refine_templates(self._h, info.ctypes.data), "lss_refine_templates")
        return dict(zip(["loss_before", "loss_after", "step", "grad_norm"], info.tolist()))

    def refine_timings(self):
        """Device ms of the last refinement call: residual (with waves and loss), adjoint, total."""
        ms = np.zeros(3, np.float32)
        check(lib().lss_last_refine_timings(self._h, ms.ctypes.data), "lss_last_refine_timings")
        return dict(zip(["residual", "adjoint", "total"], ms.tolist()))

    def profile_update(self, enable: bool = True):
        """Time each pursuit update launch with HIP events (measurement mode)."""
        check(lib().lss_profile_update(self._h, int(enable)), "lss_profile_update")

    def update_stats(self):
        """(summed update-kernel ms, update launches) of the last window in measurement mode."""
        ms, n = ct.c_float(), ct.c_int32()
        check(lib().lss_last_update_stats(self._h, ct.byref(ms), ct.byref(n)), "lss_last_update_stats")
        return ms.value, n.value

    # ---- stage entry points (torch CUDA tensors) ----------------------------------------------
    def stage_car(self, win_dev, out=None):
        import torch
        W = win_dev.shape[0]
        out = out if out is not None else torch.empty((self.C, W), dtype=torch.float32, device=win_dev.device)
        _ready()
        check(lib().lss_stage_car(self._h, _ptr(win_dev), W, _ptr(out)), "lss_stage_car")
        return out

    def stage_filter(self, win_dev, out=None):
        import torch
        W = win_dev.shape[0]
        out = out if out is not None else torch.empty((self.C, W), dtype=torch.float32, device=win_dev.device)
        _ready()
        check(lib().lss_stage_filter(self._h, _ptr(win_dev), W, _ptr(out)), "lss_stage_filter")
        return out

    def stage_highpass(self, x_cw, out=None):
        import torch
        W = x_cw.shape[1]
        out = out if out is not None else torch.empty_like(x_cw)
        _ready()
        check(lib().lss_stage_highpass(self._h, _ptr(x_cw), W, _ptr(out)), "lss_stage_highpass")
        return out

    def stage_whiten(self, x_cw, out=None):
        import torch
        W = x_cw.shape[1]
        out = out if out is not None else torch.empty((W, self.C), dtype=torch.float32, device=x_cw.device)
        _ready()
        check(lib().lss_stage_whiten(self._h, _ptr(x_cw), W, _ptr(out)), "lss_stage_whiten")
        return out

    def stage_scores(self, x_wc, out=None):
        import torch
        W = x_wc.shape[0]
        out = out if out is not None else torch.empty((W, self.T), dtype=torch.float32, device=x_wc.device)
        _ready()
        check(lib().lss_stage_scores(self._h, _ptr(x_wc), W, _ptr(out)), "lss_stage_scores")
        return out

    def stage_pursuit(self, b_wt, x_wc, window_start: int, capacity: int | None = None) -> SpikeList:
        """Pursuit on scores [W][T] and residual [W][C] (both updated in place)."""
        s, bufs = self._out(capacity or self.params.max_spikes)
        _ready()
        check(lib().lss_stage_pursuit(self._h, _ptr(b_wt), _ptr(x_wc), b_wt.shape[0], int(window_start),
                                      ct.byref(s)), "lss_stage_pursuit")
        return self._list(s, bufs)

    def set_drift(self, M=None):
        """Whitening operator := (M @ Wrot) for the next windows (M: [C][C] FP64 device tensor, e.g.
        from train.get_drift_matrix); None restores Wrot (kilosort/io.py:781-787)."""
        if M is None:
            check(lib().lss_set_drift(self._h, None), "lss_set_drift")
            return
        import torch
        M = M.to(torch.float64).contiguous()
        if tuple(M.shape) != (self.C, self.C) or not M.is_cuda:
            raise ValueError("M must be a [C][C] device tensor")
        torch.cuda.current_stream().synchronize()   # M is complete before the handle's stream reads it
        check(lib().lss_set_drift(self._h, M.data_ptr()), "lss_set_drift")

    def tables(self) -> dict:
        T, C, nt = self.T, self.C, self.nt
        L = 2 * nt + 1
        nm = np.zeros(T, np.float32)
        waves = np.zeros((T, nt, C), np.float32)
        ctc = np.zeros((T, L, T), np.float32)
        wmat = np.zeros((C, C), np.float32)
        check(lib().lss_get_tables(self._h, nm.ctypes.data, waves.ctypes.data, ctc.ctypes.data,
                                   wmat.ctypes.data), "lss_get_tables")
        return dict(nm=nm, waves=waves, ctc=ctc, wmat=wmat)


def sort_windows(sorters, windows, starts, capacity: int | None = None, consume=None) -> list:
    """Windows with up to len(sorters) in flight: window i is sorted by sorters[i % len(sorters)]
    on that sorter's own host thread (ctypes releases the GIL during the library call), so one
    window's GEMMs overlap another's pursuit and each thread copies its spike lists out while the
    GPU runs the others.  Results in window order, equal to sorting each window alone.
    consume(i, spikes), if given, is called on the worker thread as window i finishes and its
    return value is stored in place of the spike list (a writer that keeps nothing returns None)."""
    import threading
    k = len(sorters)
    if k == 0:
        raise ValueError("no sorters")
    if len({id(so) for so in sorters}) != k:
        raise ValueError("sorters must be distinct")
    res = [None] * len(windows)
    errs = []

    def work(j):
        try:
            for i in range(j, len(windows), k):
                out = sorters[j].sort_window(windows[i], starts[i], capacity)
                res[i] = consume(i, out) if consume is not None else out
        except Exception as e:          # re-raised in the caller
            errs.append(e)

    th = [threading.Thread(target=work, args=(j,)) for j in range(k)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise errs[0]
    return res


def sort_windows_native(sorters, windows, starts, capacity: int = 1 << 15, chunk: int = 16) -> list:
    """The same through lss_sort_windows (the C ABI's own threads, one per handle), for C-side
    callers' semantics: output buffers of `capacity` spikes per window, reused over groups of
    `chunk` windows and kept on the first sorter."""
    if not sorters:
        raise ValueError("no sorters")
    on_device = not isinstance(windows[0], np.ndarray)
    hs = (ct.c_void_p * len(sorters))(*[so._h for so in sorters])
    ref = sorters[0]
    group = max(int(chunk), len(sorters))
    need = min(group, len(windows))
    cached = getattr(ref, "_windows_pool", None)
    if cached is None or cached[0] != int(capacity) or len(cached[1]) < need:
        ref._windows_pool = (int(capacity), [alloc_spikes(int(capacity), ref.nC, ref.K) for _ in range(need)])
    pool = ref._windows_pool[1]
    res = []
    for g0 in range(0, len(windows), group):
        ws = windows[g0:g0 + group]
        if on_device:
            _ready()
            ptrs = [_ptr(w) for w in ws]
        else:
            ws = [np.ascontiguousarray(w, dtype=np.int16) for w in ws]
            ptrs = [w.ctypes.data for w in ws]
        n = len(ws)
        wp = (ct.c_void_p * n)(*ptrs)
        Wn = (ct.c_int64 * n)(*[int(w.shape[0]) for w in ws])
        st = (ct.c_int64 * n)(*[int(x) for x in starts[g0:g0 + n]])
        outs = (Spikes * n)(*[pool[i][0] for i in range(n)])
        check(lib().lss_sort_windows(hs, len(sorters), wp, Wn, st, int(on_device), n, outs), "lss_sort_windows")
        for i in range(n):
            res.append(Sorter._list(outs[i], pool[i][1]))
    return res


def pad_batch(rec: np.ndarray, ibatch: int, NT: int, nt: int):
    """Kilosort padded window (int16 [NT+2nt][C]) and its window_start, via the C ABI."""
    rec = np.ascontiguousarray(rec, dtype=np.int16)
    out = np.zeros((NT + 2 * nt, rec.shape[1]), np.int16)
    ws = ct.c_int64()
    check(lib().lss_pad_batch(rec.ctypes.data, rec.shape[0], rec.shape[1], NT, nt, ibatch,
                              out.ctypes.data, ct.byref(ws)), "lss_pad_batch")
    return out, ws.value
