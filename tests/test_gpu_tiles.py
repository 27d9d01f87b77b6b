"""Tile shapes of the three contractions (whitening GEMM, wPCA projection, score GEMM) and their
autotuner: every shape gives bit-identical stage outputs and spike lists, on the 64-channel probe,
on a ragged probe (37 channels, 13 templates: partial tiles on every edge) and on a 100-ms window
of the 384-channel / 600-template probe, for the FP32 path and the split-precision path."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _probe(kind):
    from livespikesorter_amd import synth
    if kind == "c64":
        return synth.generate(synth.SynthConfig(n_chan=64, n_units=24, n_samples=20000, NT=10000, seed=21)), 10122
    if kind == "ragged":
        return synth.generate(synth.SynthConfig(n_chan=37, n_units=13, n_samples=9000, NT=4000, nrange=16, nC=5,
                                                seed=22)), 4122
    return synth.generate(synth.SynthConfig(n_chan=384, n_units=600, n_samples=9000, NT=60000, seed=23)), 3122


@pytest.mark.parametrize("kind", ["c64", "ragged", "c384"])
@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_tile_shapes_are_result_neutral(kind, precision):
    from livespikesorter_amd.sorter import Sorter
    print(f"[lss] test_tile_shapes_are_result_neutral {kind} {precision}", flush=True)
    p, W = _probe(kind)
    nt = p.art["nt"]
    win = torch.from_numpy(np.ascontiguousarray(p.rec[1000:1000 + W])).to(DEV)
    with Sorter(p.art, max_window=W, max_passes=100, precision=precision, max_spikes=1 << 15) as so:
        xf = so.stage_filter(win)
        ref_w = so.stage_whiten(xf).cpu().numpy()
        ref_b = so.stage_scores(torch.from_numpy(ref_w).to(DEV)).cpu().numpy()
        ref = so.sort_window(win, 1000 - nt)
        assert len(ref) > 0
        for t in range(1, 4):
            so.set_tiles(t, t, t)
            assert so.tiles() == (t, t, t)
            np.testing.assert_array_equal(so.stage_whiten(xf).cpu().numpy(), ref_w)
            np.testing.assert_array_equal(so.stage_scores(torch.from_numpy(ref_w).to(DEV)).cpu().numpy(), ref_b)
            got = so.sort_window(win, 1000 - nt)
            for k in ("sample", "template", "amplitude", "features"):
                np.testing.assert_array_equal(getattr(got, k), getattr(ref, k))
        best, ms = so.autotune_tiles(win)
        print(f"[lss]   best {best}, ms {ms}", flush=True)
        assert so.tiles() == best and all(0 <= b < 4 for b in best)
        assert all(len(v) == 4 and min(v) > 0 for v in ms.values())
        got = so.sort_window(win, 1000 - nt)
        np.testing.assert_array_equal(got.amplitude, ref.amplitude)
        with pytest.raises(Exception):
            so.set_tiles(4, 0, 0)
