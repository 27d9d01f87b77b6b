"""Batched low-latency streaming at its full size: four probes of uneven size (384 ch / 600
templates, 384/250, 128/120, 64/40), each generated like the 64-channel workload from its own seed,
3000-sample (100 ms) batches with 61-sample halos, one graph-replayed step per batch over all
probes, bf16x3 GEMMs, the pursuit schedule autotuned per probe, template refinement rounds on.

* Before the first refinement every probe's spikes equal those of sorting the same window alone
  (sample, template, amplitude, PC features: same kernels; tuning does not change results), on
  four batches.
* Against the FP64 CPU oracle on two batches per probe: spike F1 >= 0.999 (+-1 sample, same
  template) and amplitudes within 1e-3 relative over the matched spikes.
* Every probe's refinement step lowers its window's residual energy, and the stream keeps running
  on the refined templates."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

BATCH = 3000
PROBES = [(384, 600), (384, 250), (128, 120), (64, 40)]
REFINE_EVERY = 4
N_BATCHES = 9


def _window(rec, t0, nt):
    idx = np.clip(np.arange(t0 - nt, t0 + BATCH + nt), 0, None)
    return np.ascontiguousarray(rec[idx])


@pytest.fixture(scope="module")
def stream5():
    from livespikesorter_amd import synth
    from livespikesorter_amd.sorter import Sorter
    from livespikesorter_amd.streaming import (StreamingSorter, autotune, autotune_persistent,
                                               batcher_grid_shares)
    print("[lss] generating the four probes", flush=True)
    n = (N_BATCHES + 1) * BATCH + 200
    probes = [synth.generate(synth.SynthConfig(n_chan=C, n_units=T, n_samples=n, NT=60000, nrange=min(32, C),
                                               seed=100 + i)) for i, (C, T) in enumerate(PROBES)]
    mk = lambda p: Sorter(p.art, max_window=BATCH + 2 * p.art["nt"], max_passes=100, precision="bf16x3",
                          max_spikes=1 << 15)
    sorters = [mk(p) for p in probes]
    shares = batcher_grid_shares(sorters, BATCH + 2 * probes[0].art["nt"])
    for s, p, cap in zip(sorters, probes, shares):
        w = torch.from_numpy(_window(p.rec, 3 * BATCH, p.art["nt"])).to("cuda:0")
        g, p1, _ = autotune(s, w)
        pg, pf, _ = autotune_persistent(s, w, max_grid=cap)
        print(f"[lss]   C={s.C} T={s.T}: update grid {g}, first passes {p1}, persistent {pg} after {pf}", flush=True)
    print("[lss] streaming", flush=True)
    out = [[] for _ in probes]
    refine = []
    drec = [torch.from_numpy(p.rec).to("cuda:0") for p in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=True, refine_every=REFINE_EVERY) as st:
        pos = 0
        while len(out[0]) < N_BATCHES:
            for i in range(len(probes)):
                st.push(i, drec[i][pos:pos + 1000])
            pos += 1000
            while True:
                r = st.step()
                if r is None:
                    break
                for i in range(len(probes)):
                    out[i].append(r[i])
        refine = [st.last_refine(i) for i in range(len(probes))]
    # sorters of the same artifacts for the single-window reference (the streamed ones refine)
    refs = [mk(p) for p in probes]
    yield probes, out, refine, refs
    for s in sorters + refs:
        s.close()


def test_config5_stream_equals_single_windows(stream5):
    probes, out, _, refs = stream5
    print("[lss] test_config5_stream_equals_single_windows", flush=True)
    for i, p in enumerate(probes):
        nt = p.art["nt"]
        for k in range(REFINE_EVERY):
            ref = refs[i].sort_window(_window(p.rec, k * BATCH, nt), k * BATCH - nt)
            g = out[i][k]
            assert len(ref) > 0 or i == 3
            np.testing.assert_array_equal(g.sample, ref.sample)
            np.testing.assert_array_equal(g.template, ref.template)
            np.testing.assert_array_equal(g.amplitude, ref.amplitude)
            np.testing.assert_array_equal(g.features, ref.features)


@pytest.mark.parametrize("probe", range(len(PROBES)))
def test_config5_vs_fp64_oracle(stream5, probe):
    probes, out, _, _ = stream5
    p = probes[probe]
    nt = p.art["nt"]
    print(f"[lss] test_config5_vs_fp64_oracle probe {PROBES[probe]}", flush=True)
    for k in (1, 2):
        X = torch.from_numpy(_window(p.rec, k * BATCH, nt).T.astype(np.float64))
        ref = ko.sort_window(X, p.art, k * BATCH - nt, dtype=torch.float64, max_passes=100)
        g = out[probe][k]
        f1 = ko.spike_f1(ref.samples, ref.templates, g.sample, g.template, dt=1)
        d0 = {(int(t), int(y)): float(a) for t, y, a in zip(ref.samples, ref.templates, ref.amps)}
        errs = [abs(a - d0[(int(t), int(y))]) / abs(d0[(int(t), int(y))])
                for t, y, a in zip(g.sample, g.template, g.amplitude) if (int(t), int(y)) in d0]
        print(f"[lss]   batch {k}: {len(g)} spikes, {len(ref.samples)} in the oracle, F1 {f1:.5f}, "
              f"amp err max {max(errs) if errs else 0:.2e}", flush=True)
        assert f1 >= 0.999, (probe, k, f1)
        assert len(errs) >= 0.99 * len(ref.samples) and (not errs or max(errs) <= 1e-3)


def test_config5_refinement_lowers_each_probe_loss(stream5):
    probes, out, refine, _ = stream5
    print(f"[lss] test_config5_refinement_lowers_each_probe_loss {refine}", flush=True)
    assert all(len(o) == N_BATCHES for o in out)
    for i, r in enumerate(refine):
        assert r["loss_before"] > 0 and r["loss_after"] < r["loss_before"], (i, r)
    # batches after the refinement round still find the probe's spikes
    for i in range(len(probes) - 1):
        assert len(out[i][N_BATCHES - 1]) > 0
