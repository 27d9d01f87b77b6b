"""GPU parity of the training-side FP64 builders (liblss through livespikesorter_amd.train) with the
FP64 oracle (oracle/train_oracle.py).  Tolerances: FP64 results agree to 1e-9 relative (different
summation orders and eigen/inverse algorithms, well-conditioned inputs); snippet selection and
clip values are exact (integer decisions on the same FP32 window, copies of its values)."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko
from oracle import train_oracle as to

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _np1(C):
    i = np.arange(C)
    return np.array([43.0, 11.0, 59.0, 27.0])[i % 4], 20.0 * (i // 2)


def _lin2(C):
    i = np.arange(C)
    return 32.0 * (i % 2), 20.0 * (i // 2)


def _spd(rng, n, spread=3.0):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(rng.uniform(-spread, spread, n))
    return (Q * d) @ Q.T


def _relmax(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _spatial_cov(xc, yc, rng):
    d = np.sqrt((xc[:, None] - xc[None, :]) ** 2 + (yc[:, None] - yc[None, :]) ** 2)
    A = np.exp(-d / 60.0)
    return A @ A.T / len(xc) + np.diag(rng.uniform(0.5, 2.0, len(xc)))


def test_covariance_add_vs_oracle():
    from livespikesorter_amd import train
    print("case: covariance C=384 n=60000", flush=True)
    rng = np.random.default_rng(10)
    C, n, nt = 384, 60000, 61
    mix = rng.standard_normal((C, C)).astype(np.float32) * 0.1 + np.eye(C, dtype=np.float32)
    X = (mix @ rng.standard_normal((C, n + 2 * nt)).astype(np.float32)).astype(np.float32)
    x = torch.from_numpy(X).to(DEV)
    cc = torch.zeros((C, C), dtype=torch.float64, device=DEV)
    train.covariance_add(x[:, nt:-nt], cc)
    got = cc.cpu().numpy()
    X64 = X[:, nt:-nt].astype(np.float64)
    ref = X64 @ X64.T / n
    assert _relmax(got, ref) < 1e-12
    assert np.array_equal(got, got.T)
    # accumulation: a second batch adds its own X X^T / n
    train.covariance_add(x[:, :1000], cc)
    X2 = X[:, :1000].astype(np.float64)
    assert _relmax(cc.cpu().numpy(), ref + X2 @ X2.T / 1000) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 7, 32, 61, 96])
def test_whitening_from_covariance_vs_oracle(n):
    from livespikesorter_amd import train
    print(f"case: zca n={n}", flush=True)
    rng = np.random.default_rng(11 + n)
    CC = _spd(rng, n)
    W = train.whitening_from_covariance(torch.from_numpy(CC).to(DEV)).cpu().numpy()
    assert _relmax(W, to.whitening_from_covariance(CC)) < 1e-9


@pytest.mark.parametrize("layout,nrange", [(_np1, 32), (_lin2, 31), (_lin2, 96)])
def test_whitening_local_vs_oracle(layout, nrange):
    from livespikesorter_amd import train
    print(f"case: whitening_local {layout.__name__} nrange={nrange}", flush=True)
    rng = np.random.default_rng(12)
    C = 384
    xc, yc = layout(C)
    CC = _spatial_cov(xc, yc, rng)
    W = train.whitening_local(torch.from_numpy(CC).to(DEV), xc, yc, nrange=nrange).cpu().numpy()
    ref = to.whitening_local(CC, xc, yc, nrange=nrange)
    assert _relmax(W, ref) < 1e-9
    assert np.array_equal(W != 0, ref != 0)


def test_drift_inverse_and_matrix_vs_oracle():
    from livespikesorter_amd import train
    print("case: drift C=384", flush=True)
    C = 384
    xc, yc = _np1(C)
    iK, it = train.drift_inverse(xc, yc, 20.0, 0.01)
    ref = to.drift_inverse(xc, yc, 20.0, 0.01)
    assert _relmax(iK.cpu().numpy(), ref) < 1e-9
    assert it < 80
    M1 = train.get_drift_matrix(xc, yc, 7.25, iK, 20.0).cpu().numpy()
    assert _relmax(M1, to.drift_matrix(xc, yc, 7.25, ref, 20.0)) < 1e-9
    yblk = np.linspace(500.0, 3300.0, 5)
    dsh = np.array([4.0, -3.5, 10.0, 0.0, -8.0])
    M5 = train.get_drift_matrix(xc, yc, dsh, iK, 20.0, yblk=yblk).cpu().numpy()
    assert _relmax(M5, to.drift_matrix(xc, yc, dsh, ref, 20.0, yblk=yblk)) < 1e-9


def test_set_drift_equals_loaded_drift(synth1):
    from livespikesorter_amd.sorter import Sorter
    print("case: set_drift config 1", flush=True)
    art = synth1.art
    with Sorter(art) as a, Sorter(dict(art, drift=None)) as b:
        wa = a.tables()["wmat"]
        assert np.array_equal(b.tables()["wmat"], art["Wrot"])
        b.set_drift(torch.from_numpy(art["drift"].astype(np.float64)).to(DEV))
        torch.cuda.synchronize()
        wb = b.tables()["wmat"]
        # same FP64 product, different summation order: equal up to one FP32 rounding
        ulp = np.spacing(np.abs(wa).astype(np.float32))
        assert np.all(np.abs(wb - wa) <= ulp)
        b.set_drift(None)
        assert np.array_equal(b.tables()["wmat"], art["Wrot"])


def _planted_wc(rng, C, W, n):
    X = rng.standard_normal((C, W)).astype(np.float32) * 1.5
    for _ in range(n):
        c, t = rng.integers(0, C), rng.integers(0, W)
        X[c, t] = rng.choice([-1, 1]) * rng.uniform(6.5, 20)
    return X


def test_extract_snippets_planted_vs_oracle():
    from livespikesorter_amd import train
    print("case: snippets planted C=96 W=20122", flush=True)
    rng = np.random.default_rng(13)
    C, W = 96, 20122
    X = _planted_wc(rng, C, W, 3000)
    x_wc = torch.from_numpy(np.ascontiguousarray(X.T)).to(DEV)
    clips, chan, tim, n = train.extract_snippets(x_wc, nt=61, twav_min=20, Th_single_ch=6.0)
    rc, c, t = to.extract_snippets(X, nt=61, twav_min=20, Th_single_ch=6.0)
    assert n == len(c) > 100
    assert np.array_equal(chan.cpu().numpy(), c) and np.array_equal(tim.cpu().numpy(), t)
    assert np.array_equal(clips.cpu().numpy(), rc)
    # truncation keeps the first rows
    c5, _, _, n5 = train.extract_snippets(x_wc, max_clips=5)
    assert n5 == n and np.array_equal(c5.cpu().numpy(), rc[:5])
    # other windows: loc/long ranges of the brute-force pin
    clips2, chan2, tim2, _ = train.extract_snippets(x_wc, nt=5, twav_min=2, Th_single_ch=6.0, loc_range=(1, 2),
                                                    long_range=(2, 4))
    rc2, c2, t2 = to.extract_snippets(X, nt=5, twav_min=2, Th_single_ch=6.0, loc_range=(1, 2), long_range=(2, 4))
    assert np.array_equal(chan2.cpu().numpy(), c2) and np.array_equal(tim2.cpu().numpy(), t2)
    assert np.array_equal(clips2.cpu().numpy(), rc2)


def test_extract_snippets_whitened_window_vs_oracle(synth1):
    from livespikesorter_amd import train
    from livespikesorter_amd.sorter import Sorter, pad_batch
    print("case: snippets whitened config-1 window", flush=True)
    art = synth1.art
    with Sorter(art) as s:
        win, _ = pad_batch(synth1.rec, 0, 60000, 61)
        x_wc = s.stage_whiten(s.stage_filter(torch.from_numpy(win).to(DEV)))
        clips, chan, tim, n = train.extract_snippets(x_wc)
    X = x_wc.cpu().numpy().T
    rc, c, t = to.extract_snippets(X)
    assert n == len(c) > 10
    assert np.array_equal(chan.cpu().numpy(), c) and np.array_equal(tim.cpu().numpy(), t)
    assert np.array_equal(clips.cpu().numpy(), rc)


def _basis_clips(rng, nt=61, n=20000):
    m = np.arange(nt)
    basis = np.stack([np.exp(-(m - 20) ** 2 / (2 * s ** 2)) * np.cos(f * (m - 20))
                      for s, f in [(2, 0), (4, 0.3), (6, 0.1), (3, 0.8), (8, 0.05), (5, 0.5)]])
    coef = rng.standard_normal((n, 6)) * np.array([5, 3, 2, 1.2, 0.7, 0.4])
    return (coef @ basis + 0.05 * rng.standard_normal((n, nt))).astype(np.float32)


@pytest.mark.parametrize("n_pcs", [3, 6])
def test_wpca_vs_oracle(n_pcs):
    from livespikesorter_amd import train
    print(f"case: wpca n_pcs={n_pcs}", flush=True)
    rng = np.random.default_rng(14)
    clips = _basis_clips(rng)
    w, sv = train.wpca(torch.from_numpy(clips).to(DEV), n_pcs)
    rw, rsv = to.wpca(clips, n_pcs)
    assert np.abs(w.cpu().numpy() - rw).max() < 1e-9
    assert _relmax(sv.cpu().numpy(), rsv) < 1e-12


def test_get_whitening_matrix_end_to_end(synth1):
    """Batch selection (0, nskip, ... < n_batches - 1), the average over batches and the local
    ZCA, against the oracle fed with the library's own filtered batches (exact path) and with the
    oracle's FP64 filter (FP32 filter tolerance)."""
    from livespikesorter_amd import train
    from livespikesorter_amd.sorter import pad_batch
    print("case: get_whitening_matrix config 1, NT=10000 nskip=5", flush=True)
    art = synth1.art
    rec = synth1.rec
    NT, nt, nskip = 10000, 61, 5
    Wrot, CC = train.get_whitening_matrix(rec, art["xc"], art["yc"], art["hp_filter"], NT=NT, nt=nt, nskip=nskip,
                                          nrange=32)
    nb = int(np.ceil(rec.shape[0] / NT))
    js = list(range(0, nb - 1, nskip))
    assert len(js) == 3
    filt = train.FilterStage(rec.shape[1], art["hp_filter"], art["xc"], art["yc"], NT + 2 * nt, nt=nt)
    try:
        Xs = [filt(torch.from_numpy(pad_batch(rec, j, NT, nt)[0]).to(DEV)).cpu().numpy()[:, nt:-nt] for j in js]
    finally:
        filt.close()
    ref_cc = to.covariance(Xs)
    assert _relmax(CC.cpu().numpy(), ref_cc) < 1e-12
    xc = art["xc"].astype(np.float64)
    yc = art["yc"].astype(np.float64)
    assert _relmax(Wrot.cpu().numpy(), to.whitening_local(ref_cc, xc, yc, 32)) < 1e-9
    hp = torch.from_numpy(art["hp_filter"].astype(np.float64))
    Xo = [ko.filter_batch(ko.padded_batch(rec, j, NT, nt, dtype=torch.float64), hp, None).numpy()[:, nt:-nt]
          for j in js]
    cc64 = to.covariance(Xo)
    assert _relmax(CC.cpu().numpy(), cc64) < 1e-4


def test_extract_wpca_end_to_end(synth1):
    from livespikesorter_amd import train
    from livespikesorter_amd.sorter import Sorter, pad_batch
    print("case: extract_wPCA config 1, NT=10000 nskip=2", flush=True)
    art = synth1.art
    NT = 10000
    with Sorter(art, max_window=NT + 122) as s:
        w, sv, n = train.extract_wPCA(s, synth1.rec, NT=NT, nskip=2, n_pcs=6)
        parts = []
        for j in range(0, int(np.ceil(synth1.rec.shape[0] / NT)), 2):
            win, _ = pad_batch(synth1.rec, j, NT, 61)
            x = s.stage_whiten(s.stage_filter(torch.from_numpy(win).to(DEV))).cpu().numpy().T
            parts.append(to.extract_snippets(x)[0])
    clips = np.concatenate(parts)
    assert n == len(clips) > 50
    rw, rsv = to.wpca(clips, 6)
    assert np.abs(w.cpu().numpy() - rw).max() < 1e-9
    assert _relmax(sv.cpu().numpy(), rsv) < 1e-12
