"""Edge cases of the device path: ragged sizes (channel, template and PC counts that are not
multiples of the vector width, a short template) in FP32 and in the split precisions, mean-only
preprocessing (no median reference), a window without spikes, and the window-length limits.  The oracle is the FP64 CPU restatement;
spike lists must agree at F1 >= 0.999 (+-1 sample, same template), amplitudes within 1e-3."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def ragged():
    from livespikesorter_amd import synth
    return synth.generate(synth.SynthConfig(n_chan=37, n_units=13, n_samples=8000, NT=4000, nt=31, nt0min=10,
                                            n_pcs=3, nC=5, nrange=16, rate_lo=20.0, rate_hi=60.0, seed=5))


def _match_amp(r, g):
    ref = {(int(s), int(t)): a for s, t, a in zip(r.samples, r.templates, r.amps)}
    err = [abs(a - ref[(int(s), int(t))]) / abs(ref[(int(s), int(t))])
           for s, t, a in zip(g.sample, g.template, g.amplitude) if (int(s), int(t)) in ref]
    return max(err) if err else 0.0, len(err)


@pytest.mark.parametrize("do_car", [True, False])
def test_ragged_sizes_vs_oracle(ragged, do_car):
    from livespikesorter_amd.sorter import Sorter
    print(f"case: ragged C=37 T=13 K=3 nt=31 do_car={do_car}", flush=True)
    a = ragged.art
    NT, nt = 4000, 31
    win = np.ascontiguousarray(ko.padded_batch(ragged.rec, 1, NT, nt, torch.float64).numpy().T.astype(np.int16))
    ws = ko.window_start(1, NT, nt)
    with Sorter(a, max_window=NT + 2 * nt, max_passes=100, do_car=do_car) as s:
        got = s.sort_window(win, ws)
        car = s.stage_car(torch.from_numpy(win).to(DEV)).cpu().numpy()
    X = torch.from_numpy(win.T.astype(np.float64))
    Xr = X - X.mean(1, keepdim=True)
    if do_car:
        Xr = Xr - torch.median(Xr, 0)[0]
    assert np.abs(car - Xr.numpy()).max() <= 1e-6 * np.abs(Xr.numpy()).max()
    # the oracle's filter with the same reference choice
    Wm = torch.from_numpy(a["drift"].astype(np.float64)) @ torch.from_numpy(a["Wrot"].astype(np.float64))
    Xf = ko.filter_batch(X, torch.from_numpy(a["hp_filter"].astype(np.float64)), Wm, do_car=do_car)
    ctc = ko.prepare_matching(torch.from_numpy(a["U"].astype(np.float64)), torch.from_numpy(a["wPCA"].astype(np.float64)))
    m = ko.run_matching(Xf, torch.from_numpy(a["U"].astype(np.float64)), torch.from_numpy(a["wPCA"].astype(np.float64)),
                        ctc, Th=float(a["Th"]), max_passes=100)
    samples = ko.spike_samples(m.st, ws, nt, int(a["nt0min"]))
    keep = samples >= 0

    class R:
        pass
    r = R()
    r.samples = samples[keep].numpy()
    r.templates = m.st[keep][:, 1].numpy()
    r.amps = m.amps[keep][:, 0].numpy()
    assert len(r.samples) > 20
    f1 = ko.spike_f1(r.samples, r.templates, got.sample, got.template, dt=1)
    assert f1 >= 0.999, f1
    err, n = _match_amp(r, got)
    assert n > 0 and err <= 1e-3


def test_window_without_spikes(ragged):
    from livespikesorter_amd.sorter import Sorter
    print("case: all-zero window", flush=True)
    with Sorter(ragged.art, max_window=4062, max_passes=100) as s:
        sp = s.sort_window(np.zeros((4062, 37), np.int16), 0)
        assert len(sp) == 0 and sp.passes == 0
        # the handle is still usable after an empty window
        win = np.ascontiguousarray(ko.padded_batch(ragged.rec, 0, 4000, 31, torch.float64).numpy().T.astype(np.int16))
        assert len(s.sort_window(win, -31)) > 0


def test_window_length_limits(ragged):
    from livespikesorter_amd._lib import LssError
    from livespikesorter_amd.sorter import Sorter
    print("case: window length limits", flush=True)
    with Sorter(ragged.art, max_window=1000, max_passes=100) as s:
        rec = np.ascontiguousarray(ragged.rec[:1000])
        s.sort_window(rec, 0)                                   # exactly max_window
        s.sort_window(np.ascontiguousarray(rec[:64]), 0)        # the shortest window, 2*nt+2
        with pytest.raises(LssError, match="ARG"):
            s.sort_window(np.ascontiguousarray(rec[:63]), 0)
        with pytest.raises(LssError, match="ARG"):
            s.sort_window(np.zeros((1001, 37), np.int16), 0)


@pytest.mark.parametrize("precision", ["bf16x3", "fp16x3"])
def test_ragged_sizes_split_precision(ragged, precision):
    """The split-precision paths (MFMA projection and score GEMM) at ragged sizes: C = 37 and
    T = 13 not multiples of the 32/64-wide tiles, K = 3 wavefronts, nt = 31: the same spikes as
    the FP32 path (F1 >= 0.999, amplitudes within 1e-3 relative)."""
    from livespikesorter_amd.sorter import Sorter
    print(f"case: ragged C=37 T=13 K=3 nt=31 precision={precision}", flush=True)
    NT, nt = 4000, 31
    win = np.ascontiguousarray(ko.padded_batch(ragged.rec, 1, NT, nt, torch.float64).numpy().T.astype(np.int16))
    ws = ko.window_start(1, NT, nt)
    with Sorter(ragged.art, max_window=NT + 2 * nt, max_passes=100) as s32:
        ref = s32.sort_window(win, ws)
    with Sorter(ragged.art, max_window=NT + 2 * nt, max_passes=100, precision=precision) as s:
        got = s.sort_window(win, ws)
    assert len(ref) > 20
    f1 = ko.spike_f1(ref.sample, ref.template, got.sample, got.template, dt=1)
    assert f1 >= 0.999, f1

    class R:
        pass
    r = R()
    r.samples, r.templates, r.amps = ref.sample, ref.template, ref.amplitude
    err, n = _match_amp(r, got)
    assert n >= 0.99 * len(ref) and err <= 1e-3, (n, err)
