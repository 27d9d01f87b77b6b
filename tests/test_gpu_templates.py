"""More than 1024 templates (lss_params.n_templates up to 2048; Kilosort sizes T from the sort, with
no cap, src/SpikeSorter/OnlineSpikesV2.cu:424-431).  The pursuit kernels hold 4 * NV templates per
lane (NV = 6 at T = 1500, 8 at T = 2048) and run one score column per wavefront; the persistent
pursuit is not used above 1024 templates, so short windows take the per-pass kernels too.

* T = 1500 and 2048, the pursuit given identical FP32 scores, residual and tables: spikes, B and
  the residual bit for bit against the oracle's run_matching (ks/template_matching.py:157-198), on
  an 8000-sample window and on a 3000-sample one (the persistent kernel requested, the per-pass
  kernels used);
* T = 1500: the FP64 tables (nm, waves, ctc) within 1e-6 of the oracle's, the whole window FP32
  against the FP64 oracle (spike F1 >= 0.999 within +-1 sample and same template, matched
  amplitudes within 1e-3 relative) and the streaming batcher equal to single windows."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
NT = 8000
_SYNTH = {}


def _synth(T):
    from livespikesorter_amd import synth
    if T not in _SYNTH:
        _SYNTH[T] = synth.generate(synth.SynthConfig(n_chan=32, n_units=T, n_samples=2 * NT, NT=NT, nt=21, nt0min=7,
                                                     n_pcs=3, nC=6, nrange=16, rate_lo=0.5, rate_hi=2.0,
                                                     seed=40 + T % 7))
    return _SYNTH[T]


def rel_err(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.parametrize("T,W", [(1500, 8042), (1500, 3042), (2048, 8042)])
def test_pursuit_bit_exact_many_templates(T, W):
    from livespikesorter_amd.sorter import Sorter
    print(f"[lss] test_pursuit_bit_exact_many_templates T={T} W={W}", flush=True)
    sy = _synth(T)
    a = sy.art
    nt = a["nt"]
    X = ko.padded_batch(sy.rec, 1, NT, nt, torch.float64)
    ws = ko.window_start(1, NT, nt)
    Wm = torch.as_tensor(a["drift"]).double() @ torch.as_tensor(a["Wrot"]).double()
    x32 = ko.filter_batch(X, torch.from_numpy(a["hp_filter"]).double(), Wm).float()
    if W < x32.shape[1]:
        x32 = x32[:, 1000:1000 + W].contiguous()
        ws = ws + 1000
    U32, w32 = torch.from_numpy(a["U"]), torch.from_numpy(a["wPCA"])
    B32 = ko.project(x32, U32, w32)
    with Sorter(a, max_window=X.shape[1], max_passes=100, max_spikes=1 << 15) as s:
        s.set_persistent(128, 0)            # requested; above 1024 templates the per-pass kernels run
        tb = s.tables()
        ctc = torch.from_numpy(tb["ctc"]).permute(2, 0, 1).contiguous()
        ref = ko.run_matching(x32, U32, w32, ctc, Th=a["Th"], max_passes=100, nm=torch.from_numpy(tb["nm"]),
                              waves=torch.from_numpy(tb["waves"]), B=B32)
        Bd = B32.T.contiguous().to(DEV)
        Xd = x32.T.contiguous().to(DEV)
        got = s.stage_pursuit(Bd, Xd, ws)
        assert got.passes == ref.passes
        keep = ko.spike_samples(ref.st, ws, nt, a["nt0min"]).numpy() >= 0
        st = ref.st.numpy()
        print(f"[lss]   {int(keep.sum())} spikes, {ref.passes} passes", flush=True)
        assert keep.sum() > 50
        np.testing.assert_array_equal(got.column, st[keep, 0])
        np.testing.assert_array_equal(got.template, st[keep, 1])
        np.testing.assert_array_equal(got.amplitude, ref.amps[keep, 0].numpy())
        assert np.array_equal(Bd.cpu().numpy().T, ref.B.numpy())
        assert np.array_equal(Xd.cpu().numpy().T, ref.Xres.numpy())


def test_tables_and_window_1500_templates_vs_fp64_oracle():
    from livespikesorter_amd.sorter import Sorter
    print("[lss] test_tables_and_window_1500_templates_vs_fp64_oracle", flush=True)
    sy = _synth(1500)
    a = sy.art
    X = ko.padded_batch(sy.rec, 1, NT, a["nt"], torch.float64)
    win = np.ascontiguousarray(X.numpy().T.astype(np.int16))
    ws = ko.window_start(1, NT, a["nt"])
    U, W = torch.from_numpy(a["U"]).double(), torch.from_numpy(a["wPCA"]).double()
    ctc = ko.prepare_matching(U, W)
    with Sorter(a, max_window=win.shape[0], max_passes=100, max_spikes=1 << 15) as so:
        tb = so.tables()
        assert rel_err(tb["nm"], ko.template_norms(U).numpy()) <= 1e-6
        assert rel_err(tb["waves"], ko.template_waves(U, W).numpy()) <= 1e-6
        assert rel_err(np.transpose(tb["ctc"], (2, 0, 1)), ctc.numpy()) <= 1e-6
        got = so.sort_window(win, ws)
    ref = ko.sort_window(X, a, ws, dtype=torch.float64, max_passes=100, ctc=ctc)
    f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
    d0 = {(int(t), int(y)): float(v) for t, y, v in zip(ref.samples, ref.templates, ref.amps)}
    errs = [abs(v - d0[(int(t), int(y))]) / abs(d0[(int(t), int(y))])
            for t, y, v in zip(got.sample, got.template, got.amplitude) if (int(t), int(y)) in d0]
    print(f"[lss]   {len(got)} spikes, {len(ref.samples)} in the oracle, F1 {f1:.5f}, "
          f"amp err max {max(errs) if errs else 0:.2e}", flush=True)
    assert len(ref.samples) > 100 and f1 >= 0.999
    assert len(errs) >= 0.99 * len(ref.samples) and max(errs) <= 1e-3


def test_stream_1500_templates_equals_single_windows():
    from livespikesorter_amd.sorter import Sorter
    from livespikesorter_amd.streaming import StreamingSorter
    print("[lss] test_stream_1500_templates_equals_single_windows", flush=True)
    sy = _synth(1500)
    a = sy.art
    nt, batch = a["nt"], 3000
    mk = lambda: Sorter(a, max_window=batch + 2 * nt, max_passes=100, precision="bf16x3", max_spikes=1 << 15)
    with mk() as s, mk() as ref:
        assert s.persistent(batch + 2 * nt)[2] == 0
        with StreamingSorter([s], batch=batch, use_graph=True) as st:
            st.push(0, sy.rec[:3 * batch + nt])
            outs = []
            while True:
                r = st.step()
                if r is None:
                    break
                outs.append(r[0])
        assert len(outs) == 3
        for k, g in enumerate(outs):
            idx = np.clip(np.arange(k * batch - nt, (k + 1) * batch + nt), 0, None)
            r = ref.sort_window(np.ascontiguousarray(sy.rec[idx]), k * batch - nt)
            assert len(r) > 0
            np.testing.assert_array_equal(g.sample, r.sample)
            np.testing.assert_array_equal(g.template, r.template)
            np.testing.assert_array_equal(g.amplitude, r.amplitude)
