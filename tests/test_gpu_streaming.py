"""Multi-probe streaming batcher (graph-captured) against the single-window device path and the
CPU oracle.  The batcher must return exactly the spikes that sorting each window alone returns:
same kernels, and the tuning (persistent or per-pass pursuit, grids, first-segment passes) does
not change results."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

BATCH = 3000


@pytest.fixture(scope="module")
def probes():
    from livespikesorter_amd import synth
    a = synth.generate(synth.SynthConfig(n_chan=64, n_units=24, n_samples=15000, NT=3000, seed=11))
    b = synth.generate(synth.SynthConfig(n_chan=32, n_units=8, n_samples=15000, NT=3000, nrange=16, nC=4, seed=12))
    return [a, b]


def _window(rec, t0, nt):
    idx = np.clip(np.arange(t0 - nt, t0 + BATCH + nt), 0, None)
    return np.ascontiguousarray(rec[idx])


def _stream(probes, sorters, use_graph, chunk_seed=0):
    from livespikesorter_amd.streaming import StreamingSorter
    rng = np.random.default_rng(chunk_seed)
    out = [[] for _ in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=use_graph) as st:
        n = probes[0].rec.shape[0]
        pos = 0
        while pos < n:
            k = int(rng.integers(500, 2500))
            for i, p in enumerate(probes):
                st.push(i, p.rec[pos:pos + k])
            pos += k
            while True:
                r = st.step()
                if r is None:
                    break
                for i in range(len(probes)):
                    out[i].append(r[i])
    return out


def _make_sorters(probes, **kw):
    from livespikesorter_amd.sorter import Sorter
    return [Sorter(p.art, max_window=BATCH + 2 * p.art["nt"], max_passes=100, **kw) for p in probes]


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_stream_matches_single_windows(probes, precision):
    print(f"[lss] test_stream_matches_single_windows {precision}", flush=True)
    sorters = _make_sorters(probes, precision=precision)
    got = _stream(probes, sorters, use_graph=True)
    nb = len(got[0])
    assert nb == (probes[0].rec.shape[0] - 61) // BATCH
    for i, p in enumerate(probes):
        nt = p.art["nt"]
        for k in range(nb):
            t0 = k * BATCH
            ref = sorters[i].sort_window(_window(p.rec, t0, nt), t0 - nt)
            g = got[i][k]
            np.testing.assert_array_equal(g.sample, ref.sample)
            np.testing.assert_array_equal(g.template, ref.template)
            np.testing.assert_array_equal(g.amplitude, ref.amplitude)
            np.testing.assert_array_equal(g.features, ref.features)
    for s in sorters:
        s.close()


def test_graph_and_direct_launches_agree_and_tuning_is_result_neutral(probes):
    print("[lss] test_graph_and_direct_launches_agree_and_tuning_is_result_neutral", flush=True)
    from livespikesorter_amd.streaming import autotune
    s1 = _make_sorters(probes)
    g = _stream(probes, s1, use_graph=True, chunk_seed=1)
    s2 = _make_sorters(probes)
    # tuned: probe 0 with the fastest persistent grid; probe 1 on the per-pass kernels with the
    # fastest update grid and a deliberately short first segment, so that the extra passes between
    # the two graphs run (s1 runs both probes persistent, one graph per step)
    from livespikesorter_amd.streaming import autotune_persistent
    w = torch.from_numpy(_window(probes[0].rec, BATCH, probes[0].art["nt"])).to("cuda:0")
    pg, pf, ptable = autotune_persistent(s2[0], w, max_grid=s2[0].persistent(w.shape[0])[2] // 2)
    print(f"[lss]   persistent autotune probe 0: grid {pg} after {pf}, ms {ptable}", flush=True)
    s2[1].set_persistent(0)
    w1 = torch.from_numpy(_window(probes[1].rec, BATCH, probes[1].art["nt"])).to("cuda:0")
    grid, p1, table = autotune(s2[1], w1)
    print(f"[lss]   autotune probe 1: grid {grid}, first passes {p1}, ms {table}", flush=True)
    from livespikesorter_amd._lib import lib
    assert lib().lss_set_tuning(s2[1]._h, grid, 2) == 0
    d = _stream(probes, s2, use_graph=False, chunk_seed=2)
    for i in range(len(probes)):
        for a, b in zip(g[i], d[i]):
            np.testing.assert_array_equal(a.sample, b.sample)
            np.testing.assert_array_equal(a.amplitude, b.amplitude)
    for s in s1 + s2:
        s.close()


def test_stream_vs_oracle(probes):
    print("[lss] test_stream_vs_oracle", flush=True)
    sorters = _make_sorters(probes)
    got = _stream(probes, sorters, use_graph=True)
    for i, p in enumerate(probes):
        nt = p.art["nt"]
        k = 2
        t0 = k * BATCH
        X = torch.from_numpy(_window(p.rec, t0, nt).T.astype(np.float64))
        ref = ko.sort_window(X, p.art, t0 - nt, dtype=torch.float64, max_passes=100)
        f1 = ko.spike_f1(ref.samples, ref.templates, got[i][k].sample, got[i][k].template, dt=1)
        assert f1 >= 0.999, (i, f1)
    for s in sorters:
        s.close()


def test_stream_refinement_every_two_batches(probes):
    """With refine_every=2 the first two batches are those of the plain stream, every probe's
    refinement lowers its window's residual energy, and the stream keeps running."""
    print("[lss] test_stream_refinement_every_two_batches", flush=True)
    from livespikesorter_amd.streaming import StreamingSorter
    plain = _stream(probes, _make_sorters(probes), use_graph=True)
    sorters = _make_sorters(probes)
    out = [[] for _ in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=True, refine_every=2) as st:
        for i, p in enumerate(probes):
            st.push(i, p.rec)
        while True:
            r = st.step()
            if r is None:
                break
            for i in range(len(probes)):
                out[i].append(r[i])
        infos = [st.last_refine(i) for i in range(len(probes))]
    print(f"[lss]   {infos}", flush=True)
    for i in range(len(probes)):
        assert len(out[i]) == len(plain[i])
        for k in range(2):
            np.testing.assert_array_equal(out[i][k].sample, plain[i][k].sample)
            np.testing.assert_array_equal(out[i][k].amplitude, plain[i][k].amplitude)
        assert infos[i]["loss_after"] < infos[i]["loss_before"]
    for s in sorters:
        s.close()


def test_stream_device_samples_match_host_samples(probes):
    """Samples pushed as device buffers (windows assembled device to device, no upload in the
    graph) give exactly the spikes of the same stream pushed from host memory."""
    print("[lss] test_stream_device_samples_match_host_samples", flush=True)
    from livespikesorter_amd.streaming import StreamingSorter
    host = _stream(probes, _make_sorters(probes), use_graph=True, chunk_seed=5)
    sorters = _make_sorters(probes)
    drec = [torch.from_numpy(p.rec).to("cuda:0") for p in probes]
    rng = np.random.default_rng(5)
    out = [[] for _ in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=True) as st:
        n = probes[0].rec.shape[0]
        pos = 0
        while pos < n:
            k = int(rng.integers(500, 2500))
            for i in range(len(probes)):
                st.push(i, drec[i][pos:pos + k])
            pos += k
            while True:
                r = st.step()
                if r is None:
                    break
                for i in range(len(probes)):
                    out[i].append(r[i])
    for i in range(len(probes)):
        assert len(out[i]) == len(host[i])
        for a, b in zip(out[i], host[i]):
            np.testing.assert_array_equal(a.sample, b.sample)
            np.testing.assert_array_equal(a.template, b.template)
            np.testing.assert_array_equal(a.amplitude, b.amplitude)
            np.testing.assert_array_equal(a.features, b.features)
    for s in sorters:
        s.close()


def test_stream_releases_device_samples(probes):
    """Device sample buffers are referenced only while the batcher still reads them: over many
    steps the number held per probe stays bounded (the batcher keeps samples from t0 - nt on)."""
    print("[lss] test_stream_releases_device_samples", flush=True)
    from livespikesorter_amd.streaming import StreamingSorter
    sorters = _make_sorters(probes)
    drec = [torch.from_numpy(p.rec).to("cuda:0") for p in probes]
    most = 0
    steps = 0
    with StreamingSorter(sorters, batch=BATCH, use_graph=True) as st:
        n = probes[0].rec.shape[0]
        for pos in range(0, n, 500):                     # 6 pushes per batch
            for i in range(len(probes)):
                st.push(i, drec[i][pos:pos + 500])
            while st.step() is not None:
                steps += 1
                most = max(most, max(st.held(i) for i in range(len(probes))))
    print(f"[lss]   {steps} steps, at most {most} buffers held", flush=True)
    assert steps >= 4 and most <= BATCH // 500 + 2
    for s in sorters:
        s.close()


def test_stream_unaffected_by_windows_of_other_lengths(probes):
    """A handle that sorts windows of other lengths between stream steps (each length has its
    own highpass spectrum) leaves the stream's results unchanged."""
    print("[lss] test_stream_unaffected_by_windows_of_other_lengths", flush=True)
    from livespikesorter_amd.streaming import StreamingSorter
    plain = _stream(probes, _make_sorters(probes), use_graph=True, chunk_seed=7)
    sorters = _make_sorters(probes)
    out = [[] for _ in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=True) as st:
        for i, p in enumerate(probes):
            st.push(i, p.rec)
        k = 0
        while True:
            r = st.step()
            if r is None:
                break
            for i, p in enumerate(probes):
                out[i].append(r[i])
                W = 1000 + (700 * k + 131 * i) % 2000            # a different length every time
                sorters[i].sort_window(np.ascontiguousarray(p.rec[:W]), 0)
            k += 1
    for i in range(len(probes)):
        assert len(out[i]) == len(plain[i])
        for a, b in zip(out[i], plain[i]):
            np.testing.assert_array_equal(a.sample, b.sample)
            np.testing.assert_array_equal(a.amplitude, b.amplitude)
            np.testing.assert_array_equal(a.features, b.features)
    for s in sorters:
        s.close()


def test_refinement_rounds_shorter_than_probe_count(probes):
    """refine_every = 1 with two probes: a round still under way carries on, so every probe is
    refined (not only probe 0)."""
    print("[lss] test_refinement_rounds_shorter_than_probe_count", flush=True)
    from livespikesorter_amd.streaming import StreamingSorter
    sorters = _make_sorters(probes)
    with StreamingSorter(sorters, batch=BATCH, use_graph=True, refine_every=1) as st:
        for i, p in enumerate(probes):
            st.push(i, p.rec[:4 * BATCH])
        while st.step() is not None:
            pass
        infos = [st.last_refine(i) for i in range(len(probes))]
    print(f"[lss]   {infos}", flush=True)
    for inf in infos:
        assert inf["loss_before"] > 0 and inf["loss_after"] < inf["loss_before"]
    for s in sorters:
        s.close()


def test_joint_schedule_tuning_is_result_neutral(probes):
    """autotune_stream times every combination of the probes' candidate schedules on whole steps
    and keeps the fastest; a stream on the chosen schedules returns the spikes of the default one."""
    print("[lss] test_joint_schedule_tuning_is_result_neutral", flush=True)
    from livespikesorter_amd.streaming import autotune, autotune_stream
    sorters = _make_sorters(probes, precision="bf16x3")
    ref = _stream(probes, sorters, True)
    drecs = [torch.from_numpy(p.rec).to("cuda:0") for p in probes]
    for s, p in zip(sorters, probes):
        autotune(s, torch.from_numpy(_window(p.rec, BATCH, p.art["nt"])).to("cuda:0"))
    options = [[(0, 0), (64, 0), (64, s.tuning()[1])] for s in sorters]
    best, table = autotune_stream(sorters, drecs, batch=BATCH, steps=2, warmup=1, options=options)
    print(f"[lss]   chosen {best}; {table}", flush=True)
    assert len(table) == 9 and all(v > 0 for v in table.values())
    assert all(tuple(b) in o for b, o in zip(best, options))
    assert [s.persistent(BATCH + 2 * s.nt)[:2] for s in sorters] == [tuple(b) for b in best]
    got = _stream(probes, sorters, True)
    for i in range(len(probes)):
        for g, r in zip(got[i], ref[i]):
            np.testing.assert_array_equal(g.sample, r.sample)
            np.testing.assert_array_equal(g.template, r.template)
            np.testing.assert_array_equal(g.amplitude, r.amplitude)
    for s in sorters:
        s.close()


def _stream_all(probes, sorters, run_ahead, device, refine_every=0, between=None, grouped=True, flags=None,
                use_graph=True, pipelined=True, counts=None):
    """Every sample pushed at once (so each step finds its next batch), all batches stepped;
    between(k) runs after step k.  Returns (outputs per probe, steps that ran ahead); flags, when
    given, gets whether the steps ran grouped."""
    from livespikesorter_amd.streaming import StreamingSorter
    out = [[] for _ in probes]
    src = [torch.from_numpy(p.rec).cuda() for p in probes] if device else [p.rec for p in probes]
    with StreamingSorter(sorters, batch=BATCH, use_graph=use_graph, refine_every=refine_every,
                         run_ahead=run_ahead, grouped=grouped, pipelined=pipelined) as st:
        for i in range(len(probes)):
            st.push(i, src[i])
        k = 0
        while True:
            r = st.step()
            if r is None:
                break
            for i in range(len(probes)):
                out[i].append(r[i])
            if between is not None:
                between(k)
            k += 1
        ahead = st.ran_ahead()
        if flags is not None:
            flags.append(st.grouped())
        if counts is not None:
            counts.append(st.pipelined())
    return out, ahead


def _assert_same(a, b):
    for i in range(len(a)):
        assert len(a[i]) == len(b[i])
        for x, y in zip(a[i], b[i]):
            np.testing.assert_array_equal(x.sample, y.sample)
            np.testing.assert_array_equal(x.template, y.template)
            np.testing.assert_array_equal(x.amplitude, y.amplitude)
            np.testing.assert_array_equal(x.features, y.features)


@pytest.mark.parametrize("device", [True, False])
def test_run_ahead_returns_the_step_by_step_results(probes, device):
    """Steps that preprocess their next batch beside their own pursuit (the second window sets,
    swapped at the next step) return exactly what one batch at a time returns, with template
    refinement on (refining steps do not run ahead; the others do)."""
    print(f"[lss] test_run_ahead_returns_the_step_by_step_results device={device}", flush=True)
    sa, sb = _make_sorters(probes, precision="bf16x3"), _make_sorters(probes, precision="bf16x3")
    got, ahead = _stream_all(probes, sa, True, device, refine_every=2)
    ref, ahead0 = _stream_all(probes, sb, False, device, refine_every=2)
    nb = len(got[0])
    print(f"[lss]   {nb} steps, {ahead} ran ahead", flush=True)
    assert ahead0 == 0 and ahead >= 1
    _assert_same(got, ref)
    for s in sa + sb:
        s.close()


def test_run_ahead_redone_after_a_drift_change(probes):
    """A batch preprocessed ahead with the old whitening is preprocessed again when lss_set_drift
    changes the handle before its step: the results equal those of a stream that does not run
    ahead, with the same change at the same point (and differ from a stream without it)."""
    print("[lss] test_run_ahead_redone_after_a_drift_change", flush=True)
    C = probes[0].art["Wrot"].shape[0]
    rng = np.random.default_rng(5)
    M = torch.from_numpy(np.eye(C) + 0.05 * rng.standard_normal((C, C))).cuda()

    def run(run_ahead, change):
        srt = _make_sorters(probes)
        between = (lambda k: srt[0].set_drift(M) if k == 1 else None) if change else None
        out, ahead = _stream_all(probes, srt, run_ahead, True, between=between)
        for s in srt:
            s.close()
        return out, ahead

    got, ahead = run(True, True)
    ref, _ = run(False, True)
    plain, _ = run(False, False)
    assert ahead >= 2
    _assert_same(got, ref)
    # the change reached batch 2 of probe 0 (the batch preprocessed ahead by step 1)
    assert not (np.array_equal(got[0][2].amplitude, plain[0][2].amplitude)
                and np.array_equal(got[0][2].sample, plain[0][2].sample))


@pytest.fixture(scope="module")
def long_probes():
    from livespikesorter_amd import synth
    a = synth.generate(synth.SynthConfig(n_chan=64, n_units=24, n_samples=13 * BATCH + 200, NT=3000, seed=21))
    b = synth.generate(synth.SynthConfig(n_chan=32, n_units=8, n_samples=13 * BATCH + 200, NT=3000, nrange=16, nC=4,
                                         seed=22))
    return [a, b]


def test_background_refinement_adopted_four_steps_later(long_probes):
    """refine_every=6: probe 0 is refined on batch 5's window in the background and adopts the result
    at the start of step 9.  Its batches then equal: 0-8 those of an unrefined handle, 9-11 those of
    a handle that ran lss_refine_templates right after sorting batch 5 (the same refinement step,
    run in the foreground).  A second stream gives the same lists (no dependence on timing)."""
    print("[lss] test_background_refinement_adopted_four_steps_later", flush=True)
    probes = long_probes
    runs = []
    for _ in range(2):
        sorters = _make_sorters(probes)
        out, _ = _stream_all(probes, sorters, True, True, refine_every=6)
        runs.append(out)
        for s in sorters:
            s.close()
    _assert_same(runs[0], runs[1])
    got = runs[0][0]
    p = probes[0]
    nt = p.art["nt"]
    plain, refined = _make_sorters([p, p])
    for k in range(12):
        t0 = k * BATCH
        w = _window(p.rec, t0, nt)
        ref = (refined if k >= 9 or k <= 5 else plain).sort_window(w, t0 - nt)
        if k == 5:
            info = refined.refine_templates()
            print(f"[lss]   refinement on batch 5: {info}", flush=True)
            assert info["loss_after"] < info["loss_before"]
        np.testing.assert_array_equal(got[k].sample, ref.sample)
        np.testing.assert_array_equal(got[k].template, ref.template)
        np.testing.assert_array_equal(got[k].amplitude, ref.amplitude)
        np.testing.assert_array_equal(got[k].features, ref.features)
    plain.close()
    refined.close()


def test_refinement_calls_refused_while_refining_in_background(long_probes):
    """While a stream's background refinement holds a handle's refinement buffers (from the step
    that starts it to the step that adopts it), a direct refinement call on that handle fails with
    a clear error instead of racing it; once the stream is closed it works."""
    print("[lss] test_refinement_calls_refused_while_refining_in_background", flush=True)
    from livespikesorter_amd._lib import LssError
    from livespikesorter_amd.streaming import StreamingSorter
    probes = long_probes
    sorters = _make_sorters(probes)
    with StreamingSorter(sorters, batch=BATCH, use_graph=True, refine_every=2) as st:
        for i, p in enumerate(probes):
            st.push(i, p.rec)
        for _ in range(2):            # step 1 starts probe 0's refinement (adopted at step 5)
            assert st.step() is not None
        with pytest.raises(LssError, match="background"):
            sorters[0].refine_templates()
    info = sorters[0].refine_templates()
    assert info["loss_before"] > 0
    for s in sorters:
        s.close()


@pytest.mark.parametrize("device,run_ahead,use_graph", [(True, True, True), (False, False, True),
                                                        (True, False, False)])
def test_grouped_steps_equal_per_probe_branches(probes, device, run_ahead, use_graph):
    """A step that runs every probe's pursuit tail as one launch per stage (each probe on its own
    share of the workgroups) returns exactly what a step with a branch per probe returns, with
    refinement on, graphs or direct launches, running ahead or not."""
    print(f"[lss] test_grouped_steps_equal_per_probe_branches device={device} run_ahead={run_ahead} "
          f"graph={use_graph}", flush=True)
    sa, sb = _make_sorters(probes, precision="bf16x3"), _make_sorters(probes, precision="bf16x3")
    fa, fb = [], []
    got, _ = _stream_all(probes, sa, run_ahead, device, refine_every=2, grouped=True, flags=fa,
                         use_graph=use_graph)
    ref, _ = _stream_all(probes, sb, run_ahead, device, refine_every=2, grouped=False, flags=fb,
                         use_graph=use_graph)
    assert fa == [True] and fb == [False]
    _assert_same(got, ref)
    for s in sa + sb:
        s.close()


def test_grouping_not_used_with_per_pass_rounds_and_set_before_the_first_step(probes):
    """A probe whose pursuit starts on the per-pass kernels keeps the per-probe branches (same
    results); grouping cannot change once the stream has stepped."""
    print("[lss] test_grouping_not_used_with_per_pass_rounds_and_set_before_the_first_step", flush=True)
    from livespikesorter_amd._lib import LssError, check, lib
    from livespikesorter_amd.streaming import StreamingSorter
    sa, sb = _make_sorters(probes), _make_sorters(probes)
    sa[1].set_persistent(64, 2)
    fa = []
    got, _ = _stream_all(probes, sa, True, True, flags=fa)
    ref, _ = _stream_all(probes, sb, True, True)
    assert fa == [False]
    _assert_same(got, ref)
    with StreamingSorter(sb, batch=BATCH) as st:
        for i, p in enumerate(probes):
            st.push(i, p.rec[:2 * BATCH])
        assert st.step() is not None
        assert st.grouped()
        with pytest.raises(LssError, match="before the first step"):
            check(lib().lss_stream_set_grouped(st._h, 0), "lss_stream_set_grouped")
    for s in sa + sb:
        s.close()


def test_grouped_measurement_mode(probes):
    """Measurement mode of grouped steps (tail launched without the graph, workgroups stamping the
    GPU clock) returns the same spikes, and reports per step a positive span of the grouped
    persistent pursuit, each probe's spike count as returned, its passes and its workgroups."""
    print("[lss] test_grouped_measurement_mode", flush=True)
    from livespikesorter_amd._lib import LssError
    from livespikesorter_amd.streaming import StreamingSorter
    sa, sb = _make_sorters(probes, precision="bf16x3"), _make_sorters(probes, precision="bf16x3")
    ref, _ = _stream_all(probes, sb, True, True)
    src = [torch.from_numpy(p.rec).cuda() for p in probes]
    got = [[] for _ in probes]
    with StreamingSorter(sa, batch=BATCH) as st:
        with pytest.raises(LssError, match="measurement is off"):
            st.last_profile()
        st.profile(True)
        for i in range(len(probes)):
            st.push(i, src[i])
        k = 0
        while True:
            r = st.step()
            if r is None:
                break
            us, spikes, passes, grid = st.last_profile()
            assert us > 0.0
            n_out = [len(x.sample) for x in r]
            # the first batch's window starts before the recording: spikes there may be dropped
            assert spikes == n_out if k > 0 else all(a >= b for a, b in zip(spikes, n_out))
            k += 1
            assert all(p_ >= 1 for p_ in passes) and all(g > 0 for g in grid)
            for i in range(len(probes)):
                got[i].append(r[i])
        assert st.grouped()
    _assert_same(got, ref)
    for s in sa + sb:
        s.close()


@pytest.mark.parametrize("device", [True, False])
def test_pipelined_steps_equal_step_by_step(probes, device):
    """Steps that launch their next batch's pursuit before returning (pipelined) return exactly what
    steps without it return, with background refinement on (no early launch around a refinement)."""
    print(f"[lss] test_pipelined_steps_equal_step_by_step device={device}", flush=True)
    sa, sb = _make_sorters(probes, precision="bf16x3"), _make_sorters(probes, precision="bf16x3")
    ca, cb = [], []
    got, _ = _stream_all(probes, sa, True, device, refine_every=3, pipelined=True, counts=ca)
    ref, _ = _stream_all(probes, sb, True, device, refine_every=3, pipelined=False, counts=cb)
    print(f"[lss]   {len(got[0])} steps, {ca[0]} launched early", flush=True)
    assert ca[0] >= 1 and cb[0] == 0
    _assert_same(got, ref)
    for s in sa + sb:
        s.close()


def test_pipelined_step_redone_after_a_handle_call(probes):
    """A call that takes one of the stream's handles between two steps (here sorting other windows
    on it, which overwrites the handle's window, scores and spike lists) waits for the pursuit
    launched early and makes the next step sort its batch again: the stream returns what it returns
    without the calls."""
    print("[lss] test_pipelined_step_redone_after_a_handle_call", flush=True)
    sa, sb = _make_sorters(probes), _make_sorters(probes)
    rng = np.random.default_rng(5)

    def between(k):
        i = k % len(probes)
        p = probes[i]
        t = int(rng.integers(0, p.rec.shape[0] - BATCH - 200))
        w = np.ascontiguousarray(p.rec[t:t + BATCH + 2 * p.art["nt"]])
        sa[i].sort_window(w, t)

    ca = []
    got, _ = _stream_all(probes, sa, True, True, between=between, counts=ca)
    ref, _ = _stream_all(probes, sb, True, True, pipelined=False)
    assert ca[0] >= 1
    _assert_same(got, ref)
    for s in sa + sb:
        s.close()
