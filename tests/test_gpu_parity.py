"""Device path (liblss through its C ABI) against the CPU oracle, stage by stage and end to end.

Tolerances (written where they are checked):
* CAR output: max |gpu - fp64| <= 1e-6 of the window's max |x| (exact mean/median up to FP32 rounding);
* highpass, whitening, scores: max |gpu - fp64| <= 1e-5 of the max |reference|;
* pursuit given equal scores: bit-exact spike columns, templates, amplitudes, final B and residual;
* PC features / positions: max |gpu - oracle| <= 1e-4 of the max |oracle feature|;
* whole window: spike F1 >= 0.999 within +-1 sample and same template, matched amplitudes within
  1e-3 relative, against both the FP32 and the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _window(s, ib):
    a = s.art
    X = ko.padded_batch(s.rec, ib, a["NT"], a["nt"], torch.float64)
    win = np.ascontiguousarray(X.numpy().T.astype(np.int16))   # [W][C] int16
    return win, X, ko.window_start(ib, a["NT"], a["nt"])


@pytest.fixture(scope="module")
def sorter(synth1):
    from livespikesorter_amd.sorter import Sorter
    print(f"[lss] creating sorter C={synth1.art['U'].shape[2]} T={synth1.art['U'].shape[0]}", flush=True)
    with Sorter(synth1.art, max_window=60122, max_passes=100) as s:
        yield s


def _wm(a, dtype):
    return torch.as_tensor(a["drift"]).to(dtype) @ torch.as_tensor(a["Wrot"]).to(dtype)


def test_car_stage(sorter, synth1):
    print("[lss] test_car_stage batch 0", flush=True)
    win, X, _ = _window(synth1, 0)
    got = sorter.stage_car(torch.from_numpy(win).to(DEV)).cpu().numpy()
    Xr = X - X.mean(1, keepdim=True)
    ref = (Xr - torch.median(Xr, 0)[0]).numpy()
    assert rel_err(got, ref) <= 1e-6


@pytest.mark.parametrize("W", [60122, 3122, 2999, 1601, 700])
def test_highpass_stage(sorter, W):
    print(f"[lss] test_highpass_stage W={W}", flush=True)
    rng = np.random.default_rng(W)
    x = rng.normal(0, 10, (sorter.C, W)).astype(np.float32)
    got = sorter.stage_highpass(torch.from_numpy(x).to(DEV)).cpu().numpy()
    hp32 = torch.from_numpy(ko.highpass_taps()).float().double()   # the sorter gets float32 taps
    ref = torch.real(torch.fft.ifft(torch.fft.fft(torch.from_numpy(x).double()) *
                                    torch.conj(ko.fft_highpass(hp32, W))))
    ref = torch.fft.fftshift(ref, dim=-1).numpy()
    assert rel_err(got, ref) <= 1e-5


def test_filter_and_whiten_chain(sorter, synth1):
    print("[lss] test_filter_and_whiten_chain batch 1", flush=True)
    a = synth1.art
    win, X, _ = _window(synth1, 1)
    filt = sorter.stage_filter(torch.from_numpy(win).to(DEV))
    xw = sorter.stage_whiten(filt).cpu().numpy()                        # [W][C]
    hp = torch.from_numpy(a["hp_filter"]).double()
    ref_f = ko.filter_batch(X, hp, None)
    assert rel_err(filt.cpu().numpy(), ref_f.numpy()) <= 1e-5
    ref_w = (_wm(a, torch.float64) @ ref_f).numpy().T
    assert rel_err(xw, ref_w) <= 1e-5


def test_scores_stage(sorter, synth1):
    print("[lss] test_scores_stage", flush=True)
    a = synth1.art
    _, X, _ = _window(synth1, 0)
    hp = torch.from_numpy(a["hp_filter"]).double()
    xf = ko.filter_batch(X, hp, _wm(a, torch.float64))
    x32 = xf.float()
    B = sorter.stage_scores(x32.T.contiguous().to(DEV)).cpu().numpy()   # [W][T]
    ref = ko.project(x32.double(), torch.from_numpy(a["U"]).double(),
                     torch.from_numpy(a["wPCA"]).double()).numpy().T
    assert rel_err(B, ref) <= 1e-5


def test_tables_vs_fp64_oracle(sorter, synth1):
    """Load-time tables built on the device in FP64 and stored as FP32: nm, template waves and ctc
    against the oracle's FP64 template_norms / template_waves / prepare_matching
    (template_matching.py:110-121, 141, 197); max |gpu - fp64| <= 1e-6 of max |fp64|."""
    print("[lss] test_tables_vs_fp64_oracle", flush=True)
    a = synth1.art
    tb = sorter.tables()
    U, W = torch.from_numpy(a["U"]).double(), torch.from_numpy(a["wPCA"]).double()
    assert rel_err(tb["nm"], ko.template_norms(U).numpy()) <= 1e-6
    assert rel_err(tb["waves"], ko.template_waves(U, W).numpy()) <= 1e-6
    ctc = ko.prepare_matching(U, W).numpy()                               # [i][j][l]
    assert rel_err(np.transpose(tb["ctc"], (2, 0, 1)), ctc) <= 1e-6


def _pursuit_case(sorter, synth1, ib, max_passes=100):
    a = synth1.art
    _, X, ws = _window(synth1, ib)
    hp = torch.from_numpy(a["hp_filter"]).double()
    x32 = ko.filter_batch(X, hp, _wm(a, torch.float64)).float()
    U32 = torch.from_numpy(a["U"])
    w32 = torch.from_numpy(a["wPCA"])
    B32 = ko.project(x32, U32, w32)                                     # [T][W] fp32
    tb = sorter.tables()
    ctc_ks = torch.from_numpy(tb["ctc"]).permute(2, 0, 1).contiguous()  # [i][j][l]
    return x32, B32, tb, ctc_ks, ws, U32, w32


@pytest.mark.parametrize("W,grid,first", [(60122, 128, 0), (3122, 0, 0), (3122, 32, 0), (3122, 128, 0),
                                          (3122, 64, 2)])
def test_pursuit_bit_exact(sorter, synth1, W, grid, first):
    """Given identical FP32 scores, residual and tables, the device pursuit reproduces the
    reference pursuit (template_matching.py:157-198) bit for bit: long windows and grid 0 on the
    per-pass kernels, short windows on the persistent kernel at 32 and 128 workgroups, and two
    per-pass rounds followed by the persistent kernel."""
    print(f"[lss] test_pursuit_bit_exact batch 0 W={W} persistent grid {grid} after {first}", flush=True)
    sorter.set_persistent(grid, first)
    try:
        _pursuit_bit_exact(sorter, synth1, W)
    finally:
        sorter.set_persistent(128)


@pytest.mark.parametrize("W", [60122, 3122])
def test_pursuit_bit_exact_one_column_per_wavefront(sorter, synth1, W):
    """The update kernel's other layout (one score column per wavefront, twice the work items)
    reproduces the reference pursuit bit for bit as well."""
    from livespikesorter_amd._lib import lib
    print(f"[lss] test_pursuit_bit_exact_one_column_per_wavefront W={W}", flush=True)
    sorter.set_persistent(0)
    assert lib().lss_set_update_columns(sorter._h, 1) == 0
    try:
        _pursuit_bit_exact(sorter, synth1, W)
    finally:
        lib().lss_set_update_columns(sorter._h, 2)
        sorter.set_persistent(128)


def _pursuit_bit_exact(sorter, synth1, W):
    x32, B32, tb, ctc_ks, ws, U32, w32 = _pursuit_case(sorter, synth1, 0)
    if W < x32.shape[1]:
        x32 = x32[:, 5000:5000 + W].contiguous()
        B32 = ko.project(x32, U32, w32)
        ws = ws + 5000
    ref = ko.run_matching(x32, U32, w32, ctc_ks, Th=synth1.art["Th"], max_passes=100,
                          nm=torch.from_numpy(tb["nm"]), waves=torch.from_numpy(tb["waves"]), B=B32)
    Bd = B32.T.contiguous().to(DEV)
    Xd = x32.T.contiguous().to(DEV)
    got = sorter.stage_pursuit(Bd, Xd, ws)
    assert got.passes == ref.passes
    st = ref.st.numpy()
    # drop_negative removes spikes before sample 0 from the device list
    samples = ko.spike_samples(ref.st, ws, synth1.art["nt"], synth1.art["nt0min"]).numpy()
    keep = samples >= 0
    np.testing.assert_array_equal(got.column, st[keep, 0])
    np.testing.assert_array_equal(got.template, st[keep, 1])
    np.testing.assert_array_equal(got.amplitude, ref.amps[keep, 0].numpy())
    # final score matrix ([T][W] in the oracle, [W][T] on the device) and residual
    assert np.array_equal(Bd.cpu().numpy().T, ref.B.numpy())
    assert np.array_equal(Xd.cpu().numpy().T, ref.Xres.numpy())


@pytest.mark.parametrize("grid,first", [(0, 0), (16, 0), (256, 0), (64, 1)])
def test_pursuit_crowded_passes(grid, first):
    """nt = 3, T = 16, W = 4096, random scores and a low threshold: the first pass finds ~600
    peaks, more than the persistent kernel keeps in LDS (512), so it takes its stored-list branch.
    Bit-exact against the reference pursuit, for the per-pass kernels, two persistent grids and one
    per-pass round before the persistent kernel."""
    from livespikesorter_amd.sorter import Sorter
    print(f"[lss] test_pursuit_crowded_passes grid {grid} after {first}", flush=True)
    rng = np.random.default_rng(7)
    C, T, K, nt, W = 32, 16, 2, 3, 4096
    art = dict(U=rng.standard_normal((T, K, C)).astype(np.float32) * 0.3,
               wPCA=rng.standard_normal((K, nt)).astype(np.float32) * 0.5,
               Wrot=np.eye(C, dtype=np.float32), hp_filter=np.zeros(101, np.float32),
               xc=np.arange(C, dtype=np.float32), yc=np.zeros(C, np.float32), nt0min=1, nC=4, Th=0.1)
    s = Sorter(art, max_window=W, max_passes=60, max_spikes=1 << 15)
    try:
        s.set_persistent(grid, first)
        tb = s.tables()
        B32 = torch.from_numpy(rng.standard_normal((T, W)).astype(np.float32))
        x32 = torch.from_numpy(rng.standard_normal((C, W)).astype(np.float32))
        ctc_ks = torch.from_numpy(tb["ctc"]).permute(2, 0, 1).contiguous()
        ref = ko.run_matching(x32, torch.from_numpy(art["U"]), torch.from_numpy(art["wPCA"]), ctc_ks, Th=0.1,
                              max_passes=60, nm=torch.from_numpy(tb["nm"]), waves=torch.from_numpy(tb["waves"]),
                              B=B32.clone())
        assert ref.per_pass[0] > 512, ref.per_pass[:4]
        Bd = B32.T.contiguous().to(DEV)
        Xd = x32.T.contiguous().to(DEV)
        got = s.stage_pursuit(Bd, Xd, 1 << 20)
        print(f"[lss]   passes {got.passes} spikes {len(got.column)} (oracle per pass {ref.per_pass[:4]})",
              flush=True)
        assert got.passes == ref.passes
        np.testing.assert_array_equal(got.column, ref.st[:, 0].numpy())
        np.testing.assert_array_equal(got.template, ref.st[:, 1].numpy())
        np.testing.assert_array_equal(got.amplitude, ref.amps[:, 0].numpy())
        assert np.array_equal(Bd.cpu().numpy().T, ref.B.numpy())
        assert np.array_equal(Xd.cpu().numpy().T, ref.Xres.numpy())
    finally:
        s.close()


def test_features_stage(sorter, synth1):
    print("[lss] test_features_stage batch 1", flush=True)
    a = synth1.art
    x32, B32, tb, ctc_ks, ws, U32, w32 = _pursuit_case(sorter, synth1, 1)
    res = ko.run_matching(x32, U32, w32, ctc_ks, Th=a["Th"], max_passes=100,
                          nm=torch.from_numpy(tb["nm"]), waves=torch.from_numpy(tb["waves"]), B=B32)
    got = sorter.stage_pursuit(B32.T.contiguous().to(DEV), x32.T.contiguous().to(DEV), ws)
    iCC, iU, Ucc = ko.prepare_extract(a["xc"].astype(np.float64), a["yc"].astype(np.float64), U32, a["nC"])
    tF = ko.extract_features(res.Xres, res.st, res.amps, iCC, iU, Ucc, w32).numpy()
    xs, ys = ko.spike_positions(torch.from_numpy(tF), res.st, iCC, iU, a["xc"], a["yc"])
    assert got.features.shape == tF.shape
    assert rel_err(got.features, tF) <= 1e-4
    assert np.abs(got.x - xs.numpy()).max() <= 1e-3
    assert np.abs(got.y - ys.numpy()).max() <= 1e-3


def _match_amp(t0, y0, a0, t1, y1, a1):
    """Relative amplitude error over spikes matched exactly (same sample and template)."""
    d0 = {(int(t), int(y)): float(a) for t, y, a in zip(t0, y0, a0)}
    errs = [abs(a - d0[(int(t), int(y))]) / abs(d0[(int(t), int(y))])
            for t, y, a in zip(t1, y1, a1) if (int(t), int(y)) in d0]
    return max(errs) if errs else 0.0, len(errs)


@pytest.mark.parametrize("ib", [0, 1])
def test_sort_window_vs_oracle(sorter, synth1, ib):
    print(f"[lss] test_sort_window_vs_oracle batch {ib}", flush=True)
    a = synth1.art
    win, X, ws = _window(synth1, ib)
    got = sorter.sort_window(win, ws)
    for dtype in (torch.float32, torch.float64):
        ref = ko.sort_window(X, a, ws, dtype=dtype, max_passes=100)
        f1 = ko.spike_f1(ref.samples, ref.templates, got.sample, got.template, dt=1)
        assert f1 >= 0.999, (dtype, f1, len(ref.samples), len(got.sample))
        err, n = _match_amp(ref.samples, ref.templates, ref.amps, got.sample, got.template, got.amplitude)
        assert n >= 0.99 * len(ref.samples)
        assert err <= 1e-3, (dtype, err)


def test_sort_recording_vs_ground_truth(sorter, synth1):
    print("[lss] test_sort_recording_vs_ground_truth", flush=True)
    got = sorter.sort_recording(synth1.rec, NT=synth1.art["NT"])
    assert np.all(np.diff(got.sample) >= 0)
    f1 = ko.spike_f1(synth1.gt_times, synth1.gt_units, got.sample, got.template, dt=2)
    assert f1 >= 0.9, f1


def test_capacity_error(synth1):
    print("[lss] test_capacity_error", flush=True)
    from livespikesorter_amd._lib import LssError
    from livespikesorter_amd.sorter import Sorter
    win, _, ws = _window(synth1, 0)
    with Sorter(synth1.art, max_window=60122, max_spikes=16) as s:
        with pytest.raises(LssError, match="CAPACITY"):
            s.sort_window(win, ws)
        # the handle stays usable with an output buffer large enough for nothing: window too long
        with pytest.raises(LssError, match="ARG"):
            s.sort_window(np.zeros((60123, win.shape[1]), np.int16), 0)


# ---- reduced-precision contractions (FP32 accumulation) -----------------------------------------
# Error bounds, max |gpu - fp64| / max |fp64| over the window: BF16 operands carry 8 significant
# bits (unit roundoff 2^-9), FP16 11 bits (2^-12), bf16x3 ~16 bits (2^-17 per split term), fp16x3
# ~22 bits where the low half stays in FP16's normal range (smaller values lose bits).
PREC_BOUNDS = {"bf16": 2e-2, "fp16": 3e-3, "bf16x3": 1e-4, "fp16x3": 1e-4}


@pytest.fixture(scope="module")
def lp_sorters(synth1):
    from livespikesorter_amd.sorter import Sorter
    out = {p: Sorter(synth1.art, max_window=60122, max_passes=100, precision=p) for p in PREC_BOUNDS}
    yield out
    for s in out.values():
        s.close()


@pytest.mark.parametrize("prec", list(PREC_BOUNDS))
def test_scores_and_whitening_precision(lp_sorters, synth1, prec):
    print(f"[lss] test_scores_and_whitening_precision {prec}", flush=True)
    a = synth1.art
    so = lp_sorters[prec]
    _, X, _ = _window(synth1, 0)
    hp = torch.from_numpy(a["hp_filter"]).double()
    xf = ko.filter_batch(X, hp, None)                      # before whitening, [C][W]
    xw = so.stage_whiten(xf.float().contiguous().to(DEV)).cpu().numpy()
    ref_w = (_wm(a, torch.float64) @ xf).numpy().T
    ew = rel_err(xw, ref_w)
    x32 = torch.from_numpy(ref_w.T.copy()).float()
    B = so.stage_scores(x32.T.contiguous().to(DEV)).cpu().numpy()
    ref = ko.project(x32.double(), torch.from_numpy(a["U"]).double(),
                     torch.from_numpy(a["wPCA"]).double()).numpy().T
    eb = rel_err(B, ref)
    print(f"[lss]   {prec}: whitening rel err {ew:.2e}, scores rel err {eb:.2e}", flush=True)
    assert ew <= PREC_BOUNDS[prec] and eb <= PREC_BOUNDS[prec]


@pytest.mark.parametrize("prec", list(PREC_BOUNDS))
def test_sort_window_precision_vs_fp32(sorter, lp_sorters, synth1, prec):
    """Spike list of a reduced-precision path against the FP32 device path (+-1 sample, same
    template); the split variants must reach F1 >= 0.999 with amplitudes within 1e-3."""
    print(f"[lss] test_sort_window_precision_vs_fp32 {prec}", flush=True)
    win, _, ws = _window(synth1, 1)
    ref = sorter.sort_window(win, ws)
    got = lp_sorters[prec].sort_window(win, ws)
    f1 = ko.spike_f1(ref.sample, ref.template, got.sample, got.template, dt=1)
    err, n = _match_amp(ref.sample, ref.template, ref.amplitude, got.sample, got.template, got.amplitude)
    print(f"[lss]   {prec}: F1 vs fp32 {f1:.5f}, amp rel err {err:.2e} over {n}", flush=True)
    if prec in ("bf16x3", "fp16x3"):
        assert f1 >= 0.999 and err <= 1e-3
    else:
        assert f1 >= 0.95


def test_from_dir_matches_in_memory_artifacts(sorter, synth1, tmp_path):
    """oss_input export -> lss_create_from_dir gives the same spikes as the in-memory artifacts."""
    print("[lss] test_from_dir_matches_in_memory_artifacts", flush=True)
    from livespikesorter_amd.export import write_oss_input
    from livespikesorter_amd.sorter import Sorter
    d = write_oss_input(synth1.art, tmp_path / "oss_input")
    win, _, ws = _window(synth1, 1)
    ref = sorter.sort_window(win, ws)
    with Sorter.from_dir(d, max_window=60122, max_passes=100) as s2:
        got = s2.sort_window(win, ws)
    np.testing.assert_array_equal(got.sample, ref.sample)
    np.testing.assert_array_equal(got.template, ref.template)
    np.testing.assert_array_equal(got.amplitude, ref.amplitude)


def test_sort_windows_in_flight_equals_one_at_a_time(sorter, synth1):
    """Two windows in flight over two handles (lss_sort_windows, and Python threads over
    lss_sort_window) give each window's spikes exactly as one at a time; repeated handles are
    rejected."""
    print("[lss] test_sort_windows_in_flight_equals_one_at_a_time", flush=True)
    from livespikesorter_amd._lib import LssError
    from livespikesorter_amd.sorter import Sorter, sort_windows, sort_windows_native
    a = synth1.art
    wins, starts = [], []
    for ib in range(2):
        win, _, ws = _window(synth1, ib)
        wins.append(win)
        starts.append(ws)
    order = [0, 1, 1, 0, 1]
    ref = [sorter.sort_window(wins[i], starts[i]) for i in order]
    with Sorter(a, max_window=60122, max_passes=100) as s2:
        for fn, kw in ((sort_windows_native, dict(chunk=3)), (sort_windows, {})):
            got = fn([sorter, s2], [wins[i] for i in order], [starts[i] for i in order], **kw)
            for r, g in zip(ref, got):
                np.testing.assert_array_equal(g.sample, r.sample)
                np.testing.assert_array_equal(g.template, r.template)
                np.testing.assert_array_equal(g.amplitude, r.amplitude)
                np.testing.assert_array_equal(g.features, r.features)
        # a consumer: called once per window with that window's spikes, its results kept in order
        seen = []
        got = sort_windows([sorter, s2], [wins[i] for i in order], [starts[i] for i in order],
                           consume=lambda i, o: (seen.append(i), len(o))[1])
        assert sorted(seen) == list(range(len(order)))
        assert got == [len(r) for r in ref]
        with pytest.raises(LssError):
            sort_windows_native([sorter, sorter], wins, starts)


def test_sort_window_from_page_locked_host_memory(sorter, synth1):
    """A window in lss_host_alloc memory (the facade's page-locked window buffer) sorts exactly as
    the same window from device memory and from pageable host memory."""
    import ctypes as ct
    print("[lss] test_sort_window_from_page_locked_host_memory", flush=True)
    from livespikesorter_amd._lib import check, lib
    win, _, ws = _window(synth1, 1)
    ref = sorter.sort_window(torch.from_numpy(win).to(DEV), ws)
    p = ct.c_void_p()
    check(lib().lss_host_alloc(win.nbytes, ct.byref(p)), "lss_host_alloc")
    try:
        buf = np.ctypeslib.as_array(ct.cast(p, ct.POINTER(ct.c_int16)), shape=win.shape)
        buf[...] = win
        for src in (buf, np.ascontiguousarray(win)):
            got = sorter.sort_window(src, ws)
            assert len(got) > 0
            np.testing.assert_array_equal(got.sample, ref.sample)
            np.testing.assert_array_equal(got.template, ref.template)
            np.testing.assert_array_equal(got.amplitude, ref.amplitude)
            np.testing.assert_array_equal(got.features, ref.features)
    finally:
        check(lib().lss_host_free(p), "lss_host_free")


@pytest.mark.parametrize("grid,first", [(0, 0), (64, 0), (64, 2)])
def test_measurement_mode_accounts_every_spike(sorter, synth1, grid, first):
    """Measurement mode (lss_profile_update) on a 3122-sample window, from a handle's first window
    on: the spikes the per-pass update
    launches report plus those the persistent pursuit reports are the window's spikes, its passes
    split the same way, and the spike list is the one sorting without measurement gives."""
    print(f"[lss] test_measurement_mode_accounts_every_spike grid {grid} first {first}", flush=True)
    from livespikesorter_amd.sorter import Sorter
    win, _, ws = _window(synth1, 1)
    w = torch.from_numpy(np.ascontiguousarray(win[:3122])).to(DEV)
    ref = sorter.sort_window(w, ws)
    # a fresh handle: measurement mode's first window (no update launch timed before it)
    with Sorter(synth1.art, max_window=3122, max_passes=100) as s2:
        s2.set_persistent(grid, first)
        s2.profile_update(True)
        got = s2.sort_window(w, ws)
        ev, gpu, sp = s2.update_times()
        pms, psp, pps, pg = s2.persistent_times()
        s2.profile_update(False)
    np.testing.assert_array_equal(got.sample, ref.sample)
    np.testing.assert_array_equal(got.amplitude, ref.amplitude)
    n_pass = int(np.count_nonzero(sp))
    print(f"[lss]   per-pass spikes {int(sp.sum())} over {n_pass} launches; persistent {psp} spikes, {pps} passes, "
          f"{pms:.3f} ms, {pg} workgroups", flush=True)
    assert len(got) > 0 and int(sp.sum()) + psp == len(got)
    if grid == 0:
        assert pg == 0 and psp == 0 and pms == 0.0 and n_pass == got.passes
    else:
        assert pg == grid and pms > 0 and n_pass == min(first, got.passes) and n_pass + pps == got.passes
        assert np.all(gpu[:n_pass] > 0)
