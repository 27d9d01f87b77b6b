"""Pins the CPU oracle (oracle/ks_oracle.py) against what does not depend on it: the reference's
own known-answer inputs (src/Python/conv1d_test.py:14-43, outputs recomputed by exact brute force),
brute-force loops on tiny inputs, closed forms and invariants.  Each check is chosen so that a
plausible restatement mistake (transposed operand, wrong lag sign, off-by-one pad) fails it."""
from fractions import Fraction

import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

D = torch.float64


# ---- the reference's integer KAT (conv1d_test.py:14-43) -------------------------------------
KAT_BATCH = [[1, 2, 4, 8, 16, 12, 23, 34, 45, 56, 67], [321, 12, 34, 25, 36, 12, 23, 34, 45, 56, 67]]
KAT_WPCA = [[1, 2, 3], [2, 4, 6], [3, 6, 9]]
KAT_WALL3 = [[[1, 2], [3, 4], [5, 6]], [[7, 8], [9, 10], [11, 12]]]
KAT_NM = [42, 45]
KAT_M = 3


def _kat_bruteforce():
    X, W, U = KAT_BATCH, KAT_WPCA, KAT_WALL3
    C, n = len(X), len(X[0])
    K, nt = len(W), len(W[0])
    T = len(U)
    h = nt // 2
    Bp = [[[sum(X[c][w + m - h] * W[k][m] for m in range(nt) if 0 <= w + m - h < n)
            for w in range(n)] for k in range(K)] for c in range(C)]
    B = [[sum(U[t][k][c] * Bp[c][k][w] for k in range(K) for c in range(C)) for w in range(n)]
         for t in range(T)]
    Cf = [[Fraction(max(B[t][w], 0) ** 2, KAT_NM[t]) for w in range(n)] for t in range(T)]
    for t in range(T):
        for w in list(range(KAT_M)) + list(range(n - KAT_M, n)):
            Cf[t][w] = Fraction(0)
    Cfmax = [max(Cf[t][w] for t in range(T)) for w in range(n)]
    imax = [min(t for t in range(T) if Cf[t][w] == Cfmax[w]) for w in range(n)]
    Cmax = [max(Cfmax[max(0, w - KAT_M):min(n, w + KAT_M + 1)]) for w in range(n)]
    return B, Cf, Cfmax, imax, Cmax


def test_kat_projection_and_scores():
    B_bf, Cf_bf, Cfmax_bf, imax_bf, Cmax_bf = _kat_bruteforce()
    X = torch.tensor(KAT_BATCH, dtype=D)
    W = torch.tensor(KAT_WPCA, dtype=D)
    U = torch.tensor(KAT_WALL3, dtype=D)
    B = ko.project(X, U, W)
    assert B.tolist() == B_bf
    Cfmax, imax = ko.score_argmax(B, torch.tensor(KAT_NM, dtype=D), KAT_M)
    assert np.allclose(Cfmax.numpy(), [float(v) for v in Cfmax_bf], rtol=1e-15)
    assert imax.tolist() == imax_bf
    Cmax = torch.nn.functional.max_pool1d(Cfmax[None, None], 2 * KAT_M + 1, stride=1, padding=KAT_M)[0, 0]
    assert np.allclose(Cmax.numpy(), [float(v) for v in Cmax_bf], rtol=1e-15)


# ---- batching (io.py:584-637) ----------------------------------------------------------------
def _pad_bruteforce(rec, ib, NT, nt):
    n, C = rec.shape
    nb = -(-n // NT)
    out = np.zeros((C, NT + 2 * nt))
    for col in range(NT + 2 * nt):
        if ib == 0:
            src = col - nt
            if src < 0:
                src = 0
            if src < min(n, NT + nt):
                out[:, col] = rec[src]
        elif ib == nb - 1:
            start = ib * NT - nt
            last = n - 1
            out[:, col] = rec[min(start + col, last)]
        else:
            out[:, col] = rec[ib * NT - nt + col]
    return out


@pytest.mark.parametrize("n", [50, 47, 30, 12])
def test_padded_batch_bruteforce(n):
    rng = np.random.default_rng(n)
    rec = rng.integers(-500, 500, (n, 3)).astype(np.int16)
    NT, nt = 10, 3
    for ib in range(-(-n // NT)):
        got = ko.padded_batch(rec, ib, NT, nt, D).numpy()
        np.testing.assert_array_equal(got, _pad_bruteforce(rec, ib, NT, nt))


# ---- filtering -------------------------------------------------------------------------------
def test_highpass_taps_invariants():
    h = ko.highpass_taps()
    c = len(h) // 2
    assert len(h) == 30122 and c == 15061
    assert np.abs(h[c + 1:c + 3000] - h[c - 1:c - 3000:-1]).max() < 1e-12     # zero phase
    assert abs(h.sum()) < 1e-9                                                 # DC gain ~ 0
    H = np.abs(np.fft.rfft(h))
    f = np.fft.rfftfreq(len(h), 1 / 30000.0)
    assert np.all(np.abs(H[f > 3000] - 1) < 1e-3)                             # unit gain >> cutoff
    assert np.all(H[(f > 0) & (f < 50)] < 1e-4)                                # stop band
    # filtfilt squares the Butterworth magnitude: |H(300 Hz)| = 1/2
    i = np.argmin(np.abs(f - 300))
    assert abs(H[i] - 0.5) < 2e-2


@pytest.mark.parametrize("W", [400, 401, 30200])
def test_fft_highpass_is_circular_correlation(W):
    """fftshift(ifft(fft(x) conj(fft(f)))) == sum_m f[m] x[(m + i - W//2) mod W] (brute force)."""
    rng = np.random.default_rng(W)
    x = rng.normal(size=(2, W))
    hp = torch.from_numpy(ko.highpass_taps())
    X = torch.from_numpy(x)
    y = torch.fft.fftshift(torch.real(torch.fft.ifft(torch.fft.fft(X) * torch.conj(ko.fft_highpass(hp, W)))), dim=-1)
    ft = len(hp)
    f = np.zeros(W)
    if ft < W:
        pad = (W - ft) // 2
        f[pad:pad + ft] = hp.numpy()
    else:
        crop = (ft - W) // 2
        f = hp.numpy()[crop:crop + W]
    for i in rng.integers(0, W, 25):
        ref = sum(f[m] * x[:, (m + i - W // 2) % W] for m in range(W))
        assert np.allclose(y[:, i].numpy(), ref, atol=1e-10)


def test_filter_mean_median_closed_form():
    rng = np.random.default_rng(1)
    X = torch.from_numpy(rng.normal(size=(6, 40)))
    got = ko.filter_batch(X, None, None).numpy()
    x = X.numpy() - X.numpy().mean(1, keepdims=True)
    med = np.sort(x, 0)[(6 - 1) // 2]                      # lower median of an even count
    np.testing.assert_allclose(got, x - med, atol=1e-14)
    Wm = torch.from_numpy(rng.normal(size=(6, 6)))
    got2 = ko.filter_batch(X, None, Wm).numpy()
    np.testing.assert_allclose(got2, Wm.numpy() @ (x - med), atol=1e-12)


# ---- matching setup -------------------------------------------------------------------------
def _ctc_bruteforce(U, W):
    T, K, C = U.shape
    nt = W.shape[1]
    ctc = np.zeros((T, T, 2 * nt + 1))
    for i in range(T):
        for j in range(T):
            for l in range(2 * nt + 1):
                acc = 0.0
                for k in range(K):
                    for m in range(K):
                        utu = sum(U[i, k, c] * U[j, m, c] for c in range(C))
                        wtw = sum(W[k, nt - l + tau] * W[m, tau] for tau in range(nt) if 0 <= nt - l + tau < nt)
                        acc += utu * wtw
                ctc[i, j, l] = acc
    return ctc


def test_prepare_matching_bruteforce_and_symmetry():
    rng = np.random.default_rng(2)
    U = rng.normal(size=(3, 2, 4))
    W = rng.normal(size=(2, 5))
    ctc = ko.prepare_matching(torch.from_numpy(U), torch.from_numpy(W)).numpy()
    np.testing.assert_allclose(ctc, _ctc_bruteforce(U, W), atol=1e-12)
    np.testing.assert_allclose(ctc, ctc.transpose(1, 0, 2)[:, :, ::-1], atol=1e-12)


def test_ctc_is_score_change_of_a_subtracted_template():
    """B after adding amp * wave_j at column w0 changes by amp * ctc[:, j, l] at w0 + l - nt:
    ties prepare_matching (lag sign, operand order) to project and the residual update."""
    rng = np.random.default_rng(3)
    T, K, C, nt, Wn = 4, 3, 5, 7, 80
    U = torch.from_numpy(rng.normal(size=(T, K, C)))
    Wp = torch.from_numpy(np.linalg.qr(rng.normal(size=(nt, K)))[0].T.copy())
    ctc = ko.prepare_matching(U, Wp)
    waves = ko.template_waves(U, Wp)                               # [T][nt][C]
    j, w0, amp = 2, 40, 1.7
    X = torch.zeros((C, Wn), dtype=D)
    X[:, w0 - nt // 2:w0 + nt // 2 + 1] = amp * waves[j].T
    B = ko.project(X, U, Wp)
    for l in range(2 * nt + 1):
        np.testing.assert_allclose(B[:, w0 + l - nt].numpy(), amp * ctc[:, j, l].numpy(), atol=1e-12)
    # orthonormal wPCA: the zero-lag autocorrelation is the template norm
    nm = ko.template_norms(U)
    np.testing.assert_allclose(torch.diagonal(ctc[:, :, nt]).numpy(), nm.numpy(), rtol=1e-12)


def test_prepare_extract_bruteforce():
    xc = np.array([0.0, 32.0, 0.0, 32.0, 0.0])
    yc = np.array([0.0, 0.0, 20.0, 20.0, 40.0])
    rng = np.random.default_rng(4)
    U = torch.from_numpy(rng.normal(size=(3, 2, 5)))
    iCC, iU, Ucc = ko.prepare_extract(xc, yc, U, 3)
    for c in range(5):
        d = [((xc[j] - xc[c]) ** 2 + (yc[j] - yc[c]) ** 2, j) for j in range(5)]
        assert [j for _, j in sorted(d)][:3] == iCC[:, c].tolist()
    e = (U ** 2).sum(1)
    for t in range(3):
        assert iU[t] == int(np.argmax(e[t].numpy()))
        for j in range(3):
            np.testing.assert_array_equal(Ucc[j, t].numpy(), U[t, :, iCC[j, iU[t]]].numpy())


# ---- pursuit --------------------------------------------------------------------------------
def test_single_spike_recovered_exactly():
    rng = np.random.default_rng(5)
    T, K, C, nt, Wn = 5, 3, 8, 11, 200
    Wp = torch.from_numpy(np.linalg.qr(rng.normal(size=(nt, K)))[0].T.copy())
    U = torch.from_numpy(rng.normal(size=(T, K, C)))
    ctc = ko.prepare_matching(U, Wp)
    waves = ko.template_waves(U, Wp)
    j, iX, amp = 3, 90, 2.5
    X = torch.zeros((C, Wn), dtype=D)
    X[:, iX - nt // 2:iX + nt // 2 + 1] = amp * waves[j].T
    r = ko.run_matching(X, U, Wp, ctc, Th=0.5, max_passes=10)
    assert r.st[0].tolist() == [iX, j]
    assert abs(float(r.amps[0, 0]) - amp) < 1e-12
    assert float(r.Xres.abs().max()) < 1e-12


def test_find_peaks_plateau_and_threshold():
    nt = 3
    Cfmax = torch.tensor([0, 0, 0, 70, 70, 0, 0, 0, 50, 100, 0, 0, 0, 0], dtype=D)
    # both plateau samples are peaks; 100 dominates 50; values must exceed Th^2 = 64
    assert ko.find_peaks(Cfmax, nt, 8.0).tolist() == [3, 4, 9]
    B = torch.tensor([[1.0, 3.0], [3.0, 1.0], [3.0, 3.0]], dtype=D)
    Cfm, im = ko.score_argmax(torch.cat([torch.zeros(3, 4, dtype=D), B, torch.zeros(3, 4, dtype=D)], 1),
                              torch.ones(3, dtype=D), 4)
    assert im[4:6].tolist() == [1, 0]            # first index among equal maxima


def test_sub_windows_duplicates_last_wins():
    A = torch.zeros((1, 6), dtype=D)
    cols = torch.tensor([[1, 2, 3], [2, 3, 4]])
    vals = torch.tensor([[[1.0, 1.0, 1.0], [10.0, 10.0, 10.0]]], dtype=D)
    ko._sub_windows(A, cols, vals)
    assert A[0].tolist() == [0, -1, -10, -10, -10, 0]


# ---- multi-spike update order (ks/template_matching.py:192-198) ------------------------------
def _project_bruteforce(X, U, W):
    """B[t][w] = sum_{k,c} U[t,k,c] sum_m X[c, w+m-nt//2] W[k,m], zero outside [0, n)."""
    C, n = X.shape
    T, K, _ = U.shape
    nt = W.shape[1]
    h = nt // 2
    P = [[[sum(X[c][w + m - h] * W[k][m] for m in range(nt) if 0 <= w + m - h < n)
           for w in range(n)] for k in range(K)] for c in range(C)]
    return np.array([[sum(U[t][k][c] * P[c][k][w] for k in range(K) for c in range(C))
                      for w in range(n)] for t in range(T)])


def _pursuit_bruteforce(X, U, W, Th, max_passes):
    """Greedy pursuit by definition: each pass scores the projection of the current residual,
    takes every column whose score is the maximum of its +-nt neighbourhood and above Th^2
    (ascending), amplitudes from the pass's scores, then subtracts each spike's whole waveform
    from the residual one spike at a time.  No ctc table and no update groups."""
    X = np.array(X, np.float64)
    C, n = X.shape
    T, K, _ = U.shape
    nt = W.shape[1]
    h = nt // 2
    nm = [float(sum(U[t][k][c] ** 2 for k in range(K) for c in range(C))) for t in range(T)]
    wave = [[[sum(U[t][k][c] * W[k][m] for k in range(K)) for c in range(C)] for m in range(nt)]
            for t in range(T)]
    spikes, amps, per_pass = [], [], []
    for _ in range(max_passes):
        B = _project_bruteforce(X, U, W)
        Cfmax, imax = [], []
        for w in range(n):
            cf = [0.0 if (w < nt or w >= n - nt) else max(B[t][w], 0.0) ** 2 / nm[t] for t in range(T)]
            m = max(cf)
            Cfmax.append(m)
            imax.append(cf.index(m))
        peaks = []
        for w in range(n):
            pooled = max(Cfmax[max(0, w - nt):w + nt + 1])
            if pooled > Th ** 2 and abs(pooled - Cfmax[w]) < 1e-9:
                peaks.append(w)
        if not peaks:
            break
        pass_amps = [B[imax[w]][w] / nm[imax[w]] for w in peaks]
        for w, a in zip(peaks, pass_amps):
            for m in range(nt):
                for c in range(C):
                    X[c][w - h + m] -= a * wave[imax[w]][m][c]
            spikes.append((w, imax[w]))
            amps.append(a)
        per_pass.append(len(peaks))
    return spikes, np.array(amps), X, per_pass, _project_bruteforce(X, U, W)


def _crowded_tiny_case(seed=5, C=4, T=3, K=2, nt=9, n=96):
    """Integer data (exact FP64 sums).  A block of nt+1 equal columns at 20.. gives the scores a
    two-column plateau (columns 24, 25: an exact tie, both peaks), a planted template nt < d <= 2nt
    after it puts two peaks of the first pass in overlapping +-nt score windows."""
    rng = np.random.default_rng(seed)
    W = rng.integers(-3, 4, (K, nt)).astype(np.float64)
    U = rng.integers(-3, 4, (T, K, C)).astype(np.float64)
    X = rng.integers(-2, 3, (C, n)).astype(np.float64)
    a, h = 20, nt // 2
    X[:, a:a + nt + 1] = rng.integers(-9, 10, C)[:, None]
    wave = np.einsum("tkc,km->tmc", U, W)
    for w, t, amp in [(a + h + 1 + int(rng.integers(nt + 1, 2 * nt + 1)), int(rng.integers(T)), float(rng.integers(3, 6))),
                      (70 + int(rng.integers(0, 8)), int(rng.integers(T)), float(rng.integers(3, 6)))]:
        if w + h < n - nt:
            X[:, w - h:w + h + 1] += amp * wave[t].T
    return X, U, W


def _match_tiny(X, U, W, Th, passes):
    Xt, Ut, Wt = map(torch.from_numpy, (X, U, W))
    return ko.run_matching(Xt, Ut, Wt, ko.prepare_matching(Ut, Wt), Th=Th, max_passes=passes)


def _check_against_bruteforce(X, U, W, Th=2.0, passes=4):
    sp, amps, Xres, per_pass, B = _pursuit_bruteforce(X, U, W, Th, passes)
    r = _match_tiny(X, U, W, Th, passes)
    assert r.per_pass == per_pass
    assert [tuple(s) for s in r.st.tolist()] == sp
    np.testing.assert_allclose(r.amps[:, 0].numpy(), amps, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(r.Xres.numpy(), Xres, rtol=0, atol=1e-12 * np.abs(Xres).max())
    np.testing.assert_allclose(r.B.numpy(), B, rtol=0, atol=1e-12 * np.abs(B).max())


def test_run_matching_equals_spike_by_spike_pursuit():
    """Two peaks of one pass whose +-nt score windows overlap (nt < d <= 2nt) and an exactly tied
    plateau: run_matching's grouped update equals subtracting every spike's whole waveform."""
    X, U, W = _crowded_tiny_case()
    nt = W.shape[1]
    r = _match_tiny(X, U, W, 2.0, 1)
    first = r.st[:, 0].tolist()
    d = np.diff(first)
    assert 1 in d.tolist()                                      # the tied plateau: both columns
    assert any(nt < x <= 2 * nt for x in d)                     # overlapping score windows
    _check_against_bruteforce(X, U, W)


def test_torch_indexed_subtract_keeps_last_duplicate():
    """The reference's in-place ``A[:, idx] -= v`` (template_matching.py:197-198) with repeated
    columns in idx: on CPU torch the last occurrence's write stays, which ``_sub_windows``
    emulates for exactly tied peaks."""
    rng = np.random.default_rng(21)
    for _ in range(20):
        n, L = int(rng.integers(2, 6)), int(rng.integers(2, 8))
        starts = np.sort(rng.integers(0, 10, n))
        cols = torch.from_numpy(starts[:, None] + np.arange(L))
        A = torch.from_numpy(rng.normal(size=(3, 20)))
        vals = torch.from_numpy(rng.normal(size=(3, n, L)))
        lit = A.clone()
        lit[:, cols] -= vals
        ko._sub_windows(A, cols, vals)
        np.testing.assert_array_equal(A.numpy(), lit.numpy())


def test_pursuit_scores_stay_projection_of_residual():
    """Linearity: after every pass the scores equal the projection of the residual (B = U.Xproj
    of Xres), on a crowded FP64 window (many pairs of one pass nt < d <= 2nt apart)."""
    from livespikesorter_amd import synth
    s = synth.generate(synth.SynthConfig(n_chan=32, n_units=12, n_samples=20000, NT=10000,
                                         nrange=16, nC=4, rate_lo=40.0, rate_hi=80.0, seed=4))
    a = s.art
    X = ko.padded_batch(s.rec, 1, a["NT"], a["nt"], D)
    x = ko.filter_batch(X, torch.from_numpy(a["hp_filter"]), torch.from_numpy(a["Wrot"]).double())
    U, W = torch.from_numpy(a["U"]).double(), torch.from_numpy(a["wPCA"]).double()
    nt = W.shape[1]
    r = ko.run_matching(x, U, W, ko.prepare_matching(U, W), Th=a["Th"], max_passes=100)
    close = 0
    off = 0
    for n in r.per_pass:
        cols = r.st[off:off + n, 0].numpy()
        d = np.diff(cols)
        close += int(((d > nt) & (d <= 2 * nt)).sum())
        off += n
    assert close >= 20, close
    Bp = ko.project(r.Xres, U, W)
    assert float((r.B - Bp).abs().max()) <= 1e-9 * float(Bp.abs().max())


def test_update_checks_catch_a_single_group(monkeypatch):
    """The two checks above fail when the even/odd split is dropped (one group: overlapping
    windows keep only the later spike's write)."""
    monkeypatch.setattr(ko, "UPDATE_GROUPS", 1)
    X, U, W = _crowded_tiny_case()
    with pytest.raises(AssertionError):
        _check_against_bruteforce(X, U, W)
    with pytest.raises(AssertionError):
        test_pursuit_scores_stay_projection_of_residual()


def test_extract_features_and_positions_bruteforce():
    rng = np.random.default_rng(6)
    C, T, K, nt, Wn, nC = 6, 3, 2, 5, 40, 3
    Xres = torch.from_numpy(rng.normal(size=(C, Wn)))
    Wp = torch.from_numpy(rng.normal(size=(K, nt)))
    U = torch.from_numpy(rng.normal(size=(T, K, C)))
    xc = rng.uniform(0, 50, C)
    yc = rng.uniform(0, 50, C)
    iCC, iU, Ucc = ko.prepare_extract(xc, yc, U, nC)
    st = torch.tensor([[10, 1], [25, 2]])
    amps = torch.tensor([[1.5], [0.5]], dtype=D)
    tF = ko.extract_features(Xres, st, amps, iCC, iU, Ucc, Wp).numpy()
    for n in range(2):
        w, t = int(st[n, 0]), int(st[n, 1])
        for j in range(nC):
            ch = int(iCC[j, iU[t]])
            for k in range(K):
                ref = sum(float(Xres[ch, w - nt // 2 + m]) * float(Wp[k, m]) for m in range(nt))
                ref += float(amps[n, 0]) * float(U[t, k, ch])
                assert abs(tF[n, j, k] - ref) < 1e-12
    xs, ys = ko.spike_positions(torch.from_numpy(tF), st, iCC, iU, xc, yc)
    # closed form: all mass on one channel places the spike on that channel
    tF1 = np.zeros_like(tF)
    tF1[:, 1, :] = 1.0
    xs1, ys1 = ko.spike_positions(torch.from_numpy(tF1), st, iCC, iU, xc, yc)
    for n in range(2):
        ch = int(iCC[1, iU[int(st[n, 1])]])
        assert abs(float(xs1[n]) - xc[ch]) < 1e-12 and abs(float(ys1[n]) - yc[ch]) < 1e-12


def test_spike_samples_and_remove_duplicates():
    st = torch.tensor([[61, 0], [100, 1]])
    s = ko.spike_samples(st, -61, 61, 20)
    assert s.tolist() == [61 - 61 - 30 + 20, 100 - 61 - 30 + 20]
    t, c, keep = ko.remove_duplicates(np.array([0, 3, 10, 20, 22]), np.array([1, 1, 1, 2, 2]), dt=7)
    assert keep.tolist() == [True, False, True, True, False]


def test_nmatch_hand_checked():
    n0, m, m0 = ko.nmatch(np.array([10, 20, 30]), np.array([11, 25, 29, 50]), dt=1)
    assert n0 == 2 and m.tolist() == [True, False, True, False]
    assert ko.spike_f1([10, 20], [0, 0], [10, 20], [0, 0]) == 1.0
    assert ko.spike_f1([10, 20], [0, 0], [10, 20], [0, 1]) == 0.5


# ---- whole path on the generator's ground truth ----------------------------------------------
def test_oracle_recovers_ground_truth(small_synth):
    s = small_synth
    a = s.art
    ts, ys = [], []
    for ib in range(ko.n_batches(s.rec.shape[0], a["NT"])):
        X = ko.padded_batch(s.rec, ib, a["NT"], a["nt"], D)
        r = ko.sort_window(X, a, ko.window_start(ib, a["NT"], a["nt"]), D, max_passes=100)
        ts.append(r.samples)
        ys.append(r.templates)
    f1 = ko.spike_f1(s.gt_times, s.gt_units, np.concatenate(ts), np.concatenate(ys), dt=2)
    assert f1 > 0.9, f1


def test_synth_artifacts(small_synth):
    a = small_synth.art
    W = a["wPCA"].astype(np.float64)
    np.testing.assert_allclose(W @ W.T, np.eye(W.shape[0]), atol=1e-6)
    assert small_synth.rec.dtype == np.int16
    assert a["U"].shape == (8, 6, 32)


def test_product_f1_matches_oracle_f1():
    from livespikesorter_amd.metrics import f1_same_template
    rng = np.random.default_rng(9)
    t0 = np.sort(rng.integers(0, 10000, 300))
    y0 = rng.integers(0, 5, 300)
    t1 = t0 + rng.integers(-2, 3, 300)
    y1 = np.where(rng.random(300) < 0.05, (y0 + 1) % 5, y0)
    for dt in (0, 1, 2):
        assert f1_same_template(t0, y0, t1, y1, dt) == ko.spike_f1(t0, y0, t1, y1, dt)


# ---- template refinement loss (residual energy of a fixed spike set) ------------------------
def test_residual_energy_equals_pursuit_residual(small_synth):
    """For the pursuit's own spikes and amplitudes the loss is |Xres|^2 of run_matching."""
    a = small_synth.art
    X = ko.padded_batch(small_synth.rec, 0, a["NT"], a["nt"], D)
    x = ko.filter_batch(X, torch.from_numpy(a["hp_filter"]), torch.from_numpy(a["drift"]).double()
                        @ torch.from_numpy(a["Wrot"]).double())
    U, W = torch.from_numpy(a["U"]).double(), torch.from_numpy(a["wPCA"]).double()
    r = ko.run_matching(x, U, W, ko.prepare_matching(U, W), Th=a["Th"], max_passes=100)
    assert r.st.shape[0] > 20
    loss, R = ko.residual_energy(x, r.st[:, 0], r.st[:, 1], r.amps[:, 0], U, W)
    np.testing.assert_allclose(R.numpy(), r.Xres.numpy(), rtol=0, atol=1e-9)
    assert abs(float(loss) - float((r.Xres ** 2).sum())) <= 1e-9 * float(loss)


def _tiny_refinement_case(seed=5, C=6, T=4, K=3, nt=7, W=90, n=14):
    g = torch.Generator().manual_seed(seed)
    X = torch.randn(C, W, generator=g, dtype=D)
    U = torch.randn(T, K, C, generator=g, dtype=D)
    Wp = torch.randn(K, nt, generator=g, dtype=D)
    col = torch.randint(-2, W + 2, (n,), generator=g)          # includes windows clipped at both ends
    col[:3] = torch.tensor([10, 12, 12])                         # overlaps and a repeated column
    tmp = torch.randint(0, T, (n,), generator=g)
    amp = torch.rand(n, generator=g, dtype=D) + 0.5
    return X, col, tmp, amp, U, Wp


def test_residual_energy_bruteforce():
    X, col, tmp, amp, U, Wp = _tiny_refinement_case()
    C, W = X.shape
    K, nt = Wp.shape
    R = X.numpy().copy()
    for s in range(len(col)):
        for m in range(nt):
            w = int(col[s]) - nt // 2 + m
            if 0 <= w < W:
                for c in range(C):
                    R[c, w] -= float(amp[s]) * sum(float(U[tmp[s], k, c]) * float(Wp[k, m]) for k in range(K))
    loss, Rk = ko.residual_energy(X, col, tmp, amp, U, Wp, chunk=5)
    np.testing.assert_allclose(Rk.numpy(), R, rtol=0, atol=1e-12)
    assert abs(float(loss) - float((R ** 2).sum())) <= 1e-12 * float(loss)


def test_residual_energy_gradient_vs_central_differences():
    """Autograd gradient of the oracle loss against FP64 central differences (h = 1e-4) of the
    full loss.  The loss is quadratic in U (and in wPCA) coordinate-wise, so the differences are
    exact up to rounding."""
    X, col, tmp, amp, U, Wp = _tiny_refinement_case()
    Ug, Wg = U.clone().requires_grad_(True), Wp.clone().requires_grad_(True)
    loss, _ = ko.residual_energy(X, col, tmp, amp, Ug, Wg)
    gU, gW = torch.autograd.grad(loss, (Ug, Wg))
    h = 1e-4
    for flat in range(U.numel()):
        e = torch.zeros(U.numel(), dtype=D)
        e[flat] = h
        e = e.view_as(U)
        fd = (ko.residual_energy(X, col, tmp, amp, U + e, Wp)[0] -
              ko.residual_energy(X, col, tmp, amp, U - e, Wp)[0]) / (2 * h)
        assert abs(float(fd) - float(gU.reshape(-1)[flat])) <= 1e-6 * (1 + abs(float(fd)))
    for flat in range(Wp.numel()):
        e = torch.zeros(Wp.numel(), dtype=D)
        e[flat] = h
        e = e.view_as(Wp)
        fd = (ko.residual_energy(X, col, tmp, amp, U, Wp + e)[0] -
              ko.residual_energy(X, col, tmp, amp, U, Wp - e)[0]) / (2 * h)
        assert abs(float(fd) - float(gW.reshape(-1)[flat])) <= 1e-6 * (1 + abs(float(fd)))


# ---- feature-space cluster assignment (src/Python/OnlineKilosort.py:211-284) -----------------
def test_cluster_centroids_are_ucc_rows():
    """The reference's centroid gather on its file layout (Wall [T][C][P], pc_feature_ind [T][nC]
    = each template's nearest channels) gives the rows of Ucc."""
    rng = np.random.default_rng(11)
    C, T, K, nC = 12, 5, 3, 4
    xc = np.tile([0.0, 32.0], C // 2)
    yc = np.repeat(np.arange(C // 2) * 20.0, 2)
    U = torch.from_numpy(rng.normal(size=(T, K, C)))
    iCC, iU, Ucc = ko.prepare_extract(xc, yc, U, nC)
    Wall = U.numpy().transpose(0, 2, 1)                   # [T][C][P]
    pcfi = np.stack([iCC[:, int(iU[t])].numpy() for t in range(T)])
    cent = ko.cluster_centroids_pca(Wall, pcfi)
    for t in range(T):
        np.testing.assert_array_equal(cent[t], Ucc[:, t, :].numpy().reshape(-1))


def test_assign_to_clusters_is_nearest_centroid_within_R():
    """2<x, mu> - |mu|^2 = |x|^2 - |x - mu|^2, so the label is the nearest centroid in feature
    space among the clusters within R of the spike (brute force with cdist), -1 when none is."""
    from scipy.spatial.distance import cdist
    rng = np.random.default_rng(12)
    N, K, nC, P = 300, 40, 4, 3
    tF = rng.normal(size=(N, nC, P))
    mu = rng.normal(size=(K, nC * P))
    cxy = rng.uniform(0, 400, size=(K, 2))
    cluster_xy = {k: (cxy[k, 0], cxy[k, 1]) for k in range(K) if k % 7 != 3}   # some without a position
    xs, ys = rng.uniform(-100, 500, N), rng.uniform(-100, 500, N)
    R = 80.0
    got = ko.assign_to_clusters(tF, mu, xs, ys, cluster_xy, R)
    d_feat = cdist(tF.reshape(N, -1), mu, "sqeuclidean")
    n_none = 0
    for i in range(N):
        ok = [k for k in cluster_xy if (xs[i] - cxy[k, 0]) ** 2 + (ys[i] - cxy[k, 1]) ** 2 <= R * R]
        if not ok:
            assert got[i] == -1
            n_none += 1
        else:
            assert got[i] == min(ok, key=lambda k: (d_feat[i, k], k))
    assert 0 < n_none < N                                  # both branches exercised
    assert not np.isin(got, [k for k in range(K) if k % 7 == 3]).any()
