"""The template-length limits of lss_params.nt (odd, 3 <= nt <= 127): at nt = 127 the peak test
reads its full 128-column halo, the projection its widest band and the pursuit 255 ctc lags.

* nt = 127, the pursuit given identical FP32 scores, residual and tables: bit for bit against the
  oracle's run_matching (ks/template_matching.py:157-198) on an 8254-sample window (per-pass
  kernels) and a 4000-sample window (persistent kernel, and two per-pass rounds before it); the
  same on 4000-sample windows at nt = 3 (the persistent kernel's direct window scan), 7 and 9 (its
  block-maximum scan with no and with partial runs at the ends);
* nt = 127 and nt = 3, the whole window: FP32 against the FP64 oracle (spike F1 >= 0.999 within
  +-1 sample and same template, matched amplitudes within 1e-3 relative) and bf16x3 against FP32
  (the same bar)."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
NT = 8000
LIMITS = {127: (42, 6), 3: (1, 3), 7: (2, 3), 9: (3, 3)}   # nt: (nt0min, PCs)
_SYNTH = {}


def _synth(nt):
    from livespikesorter_amd import synth
    nt0, K = LIMITS[nt]
    return synth.generate(synth.SynthConfig(n_chan=64, n_units=20, n_samples=2 * NT, NT=NT, nt=nt, nt0min=nt0,
                                            n_pcs=K, rate_lo=20.0, rate_hi=60.0, seed=11))


def _cached(nt):
    if nt not in _SYNTH:
        _SYNTH[nt] = _synth(nt)
    return _SYNTH[nt]


@pytest.fixture(scope="module")
def s127():
    return _cached(127)


def _amp_err(ref_s, ref_t, ref_a, got):
    ref = {(int(s), int(t)): a for s, t, a in zip(ref_s, ref_t, ref_a)}
    err = [abs(a - ref[(int(s), int(t))]) / abs(ref[(int(s), int(t))])
           for s, t, a in zip(got.sample, got.template, got.amplitude) if (int(s), int(t)) in ref]
    return (max(err) if err else 0.0), len(err)


@pytest.mark.parametrize("nt,W,grid,first", [(127, 8254, 0, 0), (127, 4000, 128, 0), (127, 4000, 64, 2),
                                             (3, 4000, 128, 0), (7, 4000, 128, 0), (9, 4000, 64, 2)])
def test_pursuit_bit_exact_nt_limits(nt, W, grid, first):
    from livespikesorter_amd.sorter import Sorter
    print(f"case: nt={nt} pursuit bit-exact W={W} persistent grid {grid} after {first}", flush=True)
    sy = _cached(nt)
    a = sy.art
    X = ko.padded_batch(sy.rec, 1, NT, nt, torch.float64)
    ws = ko.window_start(1, NT, nt)
    Wm = torch.as_tensor(a["drift"]).double() @ torch.as_tensor(a["Wrot"]).double()
    x32 = ko.filter_batch(X, torch.from_numpy(a["hp_filter"]).double(), Wm).float()
    if W < x32.shape[1]:
        x32 = x32[:, 2000:2000 + W].contiguous()
        ws = ws + 2000
    U32, w32 = torch.from_numpy(a["U"]), torch.from_numpy(a["wPCA"])
    B32 = ko.project(x32, U32, w32)
    with Sorter(a, max_window=X.shape[1], max_passes=100, max_spikes=1 << 15) as s:
        s.set_persistent(grid, first)
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
        assert keep.sum() > 20
        np.testing.assert_array_equal(got.column, st[keep, 0])
        np.testing.assert_array_equal(got.template, st[keep, 1])
        np.testing.assert_array_equal(got.amplitude, ref.amps[keep, 0].numpy())
        assert np.array_equal(Bd.cpu().numpy().T, ref.B.numpy())
        assert np.array_equal(Xd.cpu().numpy().T, ref.Xres.numpy())


@pytest.mark.parametrize("nt", [127, 3])
def test_window_at_nt_limit(s127, nt):
    from livespikesorter_amd.sorter import Sorter
    print(f"case: nt={nt} whole window fp32 vs fp64 oracle, bf16x3 vs fp32", flush=True)
    s = s127 if nt == 127 else _cached(nt)
    a = s.art
    X = ko.padded_batch(s.rec, 1, NT, nt, torch.float64)
    win = np.ascontiguousarray(X.numpy().T.astype(np.int16))
    ws = ko.window_start(1, NT, nt)
    ref = ko.sort_window(X, a, ws, dtype=torch.float64, max_passes=100)
    assert len(ref.samples) > 50
    with Sorter(a, max_window=win.shape[0], max_passes=100) as so:
        got = so.sort_window(win, ws)
    f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
    assert f1 >= 0.999, f1
    err, n = _amp_err(ref.samples, ref.templates, ref.amps, got)
    assert n >= 0.99 * len(ref.samples) and err <= 1e-3, (n, err)
    with Sorter(a, max_window=win.shape[0], max_passes=100, precision="bf16x3") as so:
        x3 = so.sort_window(win, ws)
    f1 = ko.spike_f1(got.sample, got.template, x3.sample, x3.template, dt=1)
    assert f1 >= 0.999, f1
    err, n = _amp_err(got.sample, got.template, got.amplitude, x3)
    assert n >= 0.99 * len(got) and err <= 1e-3, (n, err)


def test_residual_runs_of_passes_bit_exact():
    """nt = 3, 256 templates on disjoint channels (ctc zero between templates) and scores that put a
    peak every 4 columns in each of 256 passes: 4-5 spikes of every pass reach each 16-column block
    of the residual kernel, over 1024 in all, so its list is filled run after run of passes.  Spikes,
    B and the residual bit for bit against the oracle's run_matching."""
    from livespikesorter_amd import synth
    from livespikesorter_amd.sorter import Sorter
    print("case: residual runs of passes nt=3 T=256 C=256 W=400, 256 passes", flush=True)
    T = C = 256
    nt, W = 3, 400
    sy = synth.generate(synth.SynthConfig(n_chan=C, n_units=T, n_samples=4000, NT=2000, nt=nt, nt0min=1, n_pcs=2,
                                          nC=4, nrange=16, seed=5))
    a = dict(sy.art)
    U = np.zeros((T, 2, C), np.float32)
    U[np.arange(T), 0, np.arange(T)] = 1.0
    U[np.arange(T), 1, np.arange(T)] = 0.5
    a["U"] = U
    rng = np.random.default_rng(5)
    B = np.ones((T, W), np.float32)
    cols = np.arange(4, W - nt, 4)
    tt, ww = np.meshgrid(np.arange(T), cols, indexing="ij")
    B[:, cols] = (100.0 + 5.0 * ((7 * tt + ww) % 50)).astype(np.float32)
    x32 = torch.from_numpy(rng.standard_normal((C, W)).astype(np.float32))
    B32 = torch.from_numpy(B)
    U32, w32 = torch.from_numpy(U), torch.from_numpy(a["wPCA"])
    ws = 1000
    with Sorter(a, max_window=W, max_passes=256, max_spikes=1 << 15) as s:
        s.set_persistent(0)
        tb = s.tables()
        ctc = torch.from_numpy(tb["ctc"]).permute(2, 0, 1).contiguous()
        ref = ko.run_matching(x32, U32, w32, ctc, Th=a["Th"], max_passes=256, nm=torch.from_numpy(tb["nm"]),
                              waves=torch.from_numpy(tb["waves"]), B=B32)
        Bd = B32.T.contiguous().to(DEV)
        Xd = x32.T.contiguous().to(DEV)
        got = s.stage_pursuit(Bd, Xd, ws)
        # spikes per 16-column block (+-nt/2 reach) over all passes: the residual list is exceeded
        st = ref.st.numpy()
        reach = max(int(np.sum((st[:, 0] >= b - nt // 2) & (st[:, 0] <= b + 15 + nt // 2))) for b in range(0, W, 16))
        print(f"case:   {len(st)} spikes, {ref.passes} passes, at most {reach} reach one block", flush=True)
        assert ref.passes == 256 and reach > 1024
        assert got.passes == ref.passes
        np.testing.assert_array_equal(got.column, st[:, 0])
        np.testing.assert_array_equal(got.template, st[:, 1])
        np.testing.assert_array_equal(got.amplitude, ref.amps[:, 0].numpy())
        assert np.array_equal(Bd.cpu().numpy().T, ref.B.numpy())
        assert np.array_equal(Xd.cpu().numpy().T, ref.Xres.numpy())
