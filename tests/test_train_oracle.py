"""Pins the training-side oracle (oracle/train_oracle.py) against what does not depend on it:
closed forms, brute-force loops, and the library routines the reference calls (torch SVD and
pooling, scipy interp1d, scikit-learn TruncatedSVD).  Each check is chosen so that a plausible
restatement mistake (a transposed operand, a wrong sign or shift direction, an off-by-one window,
a dropped normalisation) fails it."""
import numpy as np
import pytest
import torch

from oracle import train_oracle as to


def _np1(C):
    """Staggered four-column layout (Neuropixels 1.0 style)."""
    i = np.arange(C)
    return np.array([43.0, 11.0, 59.0, 27.0])[i % 4], 20.0 * (i // 2)


def _lin2(C):
    i = np.arange(C)
    return 32.0 * (i % 2), 20.0 * (i // 2)


def _spd(rng, n, spread=3.0):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.exp(rng.uniform(-spread, spread, n))
    return (Q * d) @ Q.T


# ---- covariance ------------------------------------------------------------------------------
def test_covariance_constructed_and_bruteforce():
    rng = np.random.default_rng(0)
    C, n = 6, 64
    Z, _ = np.linalg.qr(rng.standard_normal((n, C)))          # orthonormal columns
    A = rng.standard_normal((C, C))
    X = A @ (Z.T * np.sqrt(n))                                   # X X^T / n = A A^T
    assert np.allclose(to.covariance([X]), A @ A.T, rtol=1e-12, atol=1e-12)
    Xs = [rng.standard_normal((3, 7)) for _ in range(2)]
    bf = np.zeros((3, 3))
    for X in Xs:
        for i in range(3):
            for j in range(3):
                bf[i, j] += sum(X[i, t] * X[j, t] for t in range(7)) / 7
    assert np.allclose(to.covariance(Xs), bf / 2, rtol=1e-13)


# ---- ZCA -------------------------------------------------------------------------------------
def test_whitening_from_covariance_matches_reference_svd_routine():
    rng = np.random.default_rng(1)
    for n in (5, 32, 61):
        CC = _spd(rng, n)
        E, D, _ = torch.linalg.svd(torch.from_numpy(CC))            # the reference's routine, FP64
        ref = ((E / (D + 1e-6) ** 0.5) @ E.T).numpy()
        W = to.whitening_from_covariance(CC)
        assert np.abs(W - ref).max() < 1e-10 * np.abs(ref).max()
        # W CC W = I - eps (CC + eps I)^-1 and W symmetric
        lhs = W @ CC @ W
        rhs = np.eye(n) - 1e-6 * np.linalg.inv(CC + 1e-6 * np.eye(n))
        assert np.abs(lhs - rhs).max() < 1e-9
        assert np.abs(W - W.T).max() < 1e-10 * np.abs(W).max()


def test_whitening_from_covariance_diagonal_closed_form():
    d = np.array([4.0, 1.0, 0.25, 9.0])
    W = to.whitening_from_covariance(np.diag(d), eps=1e-6)
    assert np.allclose(W, np.diag(1 / np.sqrt(d + 1e-6)), rtol=1e-14, atol=1e-15)
    # a negative rounding eigenvalue enters through its magnitude (singular value)
    W2 = to.whitening_from_covariance(np.diag([-1e-9, 1.0]), eps=1e-6)
    assert np.isclose(W2[0, 0], 1 / np.sqrt(1e-9 + 1e-6))


# ---- local whitening -------------------------------------------------------------------------
def test_whitening_local_full_range_is_full_zca():
    rng = np.random.default_rng(2)
    C = 12
    xc, yc = _lin2(C)
    CC = _spd(rng, C, 1.0)
    W = to.whitening_local(CC, xc, yc, nrange=C)
    full = to.whitening_from_covariance(CC)
    assert np.abs(W - full).max() < 1e-10 * np.abs(full).max()
    W2 = to.whitening_local(CC, xc, yc, nrange=100)              # nrange > C takes all channels
    assert np.abs(W2 - full).max() < 1e-10 * np.abs(full).max()


def test_whitening_local_block_diagonal():
    # channels in two groups 1 mm apart: with nrange = group size, each row is that group's ZCA row
    rng = np.random.default_rng(3)
    g = 6
    xc = np.zeros(2 * g)
    yc = np.concatenate([np.arange(g) * 10.0, 1000.0 + np.arange(g) * 10.0])
    A, B = _spd(rng, g, 1.0), _spd(rng, g, 1.0)
    CC = np.zeros((2 * g, 2 * g))
    CC[:g, :g], CC[g:, g:] = A, B
    W = to.whitening_local(CC, xc, yc, nrange=g)
    ref = np.zeros_like(CC)
    ref[:g, :g] = to.whitening_from_covariance(A)
    ref[g:, g:] = to.whitening_from_covariance(B)
    assert np.abs(W - ref).max() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("layout", [_lin2, _np1])
def test_neighbour_sets_match_numpy_default_argsort(layout):
    """The reference sorts with numpy's default (unstable) argsort; at nrange = 32 the chosen sets
    equal the stable sort's on these layouts, so the tie rule does not change the result."""
    xc, yc = layout(384)
    for j in range(384):
        ds = (xc[j] - xc) ** 2 + (yc[j] - yc) ** 2
        assert set(np.argsort(ds)[:32]) == set(to.neighbours(xc, yc, j, 32))
        assert to.neighbours(xc, yc, j, 32)[0] == j


# ---- drift -----------------------------------------------------------------------------------
def test_kernel2d_and_inverse():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 50, (5, 2))
    y = rng.uniform(0, 50, (4, 2))
    K = to.kernel2d(x, y, 7.0)
    for i in range(5):
        for j in range(4):
            d2 = (x[i, 0] - y[j, 0]) ** 2 + (x[i, 1] - y[j, 1]) ** 2
            assert np.isclose(K[i, j], np.exp(-d2 / 98.0), rtol=1e-14)
    xc, yc = _np1(48)
    iK = to.drift_inverse(xc, yc, 20.0, 0.01)
    Kxx = to.kernel2d(to.sites(xc, yc), to.sites(xc, yc), 20.0)
    assert np.abs((Kxx + 0.01 * np.eye(48)) @ iK - np.eye(48)).max() < 1e-10
    lam, E = np.linalg.eigh(Kxx)
    assert np.abs(iK - (E / (lam + 0.01)) @ E.T).max() < 1e-8 * np.abs(iK).max()


def test_drift_shifts_match_scipy_interp1d():
    from scipy.interpolate import interp1d
    yblk = np.array([100.0, 300.0, 500.0, 700.0, 900.0])
    dsh = np.array([3.0, -2.0, 5.5, 1.0, -4.0])
    yc = np.array([-50.0, 0.0, 100.0, 150.0, 300.0, 640.0, 900.0, 1200.0])
    ref = interp1d(yblk, dsh, fill_value="extrapolate", kind="linear")(yc)
    assert np.allclose(to.drift_shifts(yc, dsh, yblk), ref, rtol=0, atol=1e-12)
    # unsorted block centres are sorted first, as interp1d does
    o = np.array([3, 0, 4, 1, 2])
    assert np.allclose(to.drift_shifts(yc, dsh[o], yblk[o]), ref, rtol=0, atol=1e-12)
    assert np.all(to.drift_shifts(yc, [2.5]) == 2.5)


def test_drift_matrix_zero_shift_and_direction():
    xc, yc = _lin2(40)
    iK = to.drift_inverse(xc, yc, 20.0, 0.01)
    M0 = to.drift_matrix(xc, yc, 0.0, iK)
    Kxx = to.kernel2d(to.sites(xc, yc), to.sites(xc, yc), 20.0)
    lam, E = np.linalg.eigh(Kxx)
    assert np.abs(M0 - (E * (lam / (lam + 0.01))) @ E.T).max() < 1e-8
    # a shift of one row pitch (20 um) makes row j sample the site one row below (y - 20)
    M = to.drift_matrix(xc, yc, 20.0, iK)
    j = 20
    below = j - 2                  # same column, 20 um lower
    assert np.argmax(M[j]) == below


# ---- snippets --------------------------------------------------------------------------------
def _torch_isolated(X, nt, th, loc, lng):
    """ks/spikedetect.py:17-43 with torch's pooling routines."""
    import torch.nn.functional as F
    Xt = torch.from_numpy(np.asarray(X, np.float32))
    Xabs = Xt.abs()
    Xmax = F.max_pool2d(Xabs.unsqueeze(0), [2 * loc[0] + 1, 2 * loc[1] + 1], stride=[1, 1],
                        padding=[loc[0], loc[1]])[0]
    ispeak = torch.logical_and(Xmax == Xabs, Xabs > th).float()
    s = F.avg_pool2d(ispeak.unsqueeze(0), [2 * lng[0] + 1, 2 * lng[1] + 1], stride=[1, 1],
                     padding=[lng[0], lng[1]])[0]
    s *= (2 * lng[0] + 1) * (2 * lng[1] + 1)
    iso = (s == 1) * (ispeak == 1)
    iso[:, :nt] = 0
    iso[:, -nt:] = 0
    xy = iso.nonzero()
    return xy[:, 0].numpy(), xy[:, 1].numpy()


def _planted(rng, C, W, n):
    X = rng.standard_normal((C, W)).astype(np.float32) * 1.5
    for _ in range(n):
        c, t = rng.integers(0, C), rng.integers(0, W)
        X[c, t] = rng.choice([-1, 1]) * rng.uniform(6.5, 20)
    return X


def test_isolated_peaks_match_torch_pooling():
    rng = np.random.default_rng(5)
    for (C, W, n) in [(24, 900, 60), (40, 2000, 300), (16, 500, 5)]:
        X = _planted(rng, C, W, n)
        c, t = to.isolated_peaks(X, 61, 6.0)
        rc, rt = _torch_isolated(X, 61, 6.0, (4, 5), (6, 30))
        assert np.array_equal(c, rc) and np.array_equal(t, rt)
        assert len(c) > 0 or n < 10


def test_isolated_peaks_bruteforce_tiny():
    rng = np.random.default_rng(6)
    C, W, nt = 7, 60, 5
    X = _planted(rng, C, W, 12)
    lc, lt, gc, gt = 1, 2, 2, 4
    ab = np.abs(X)
    pk = np.zeros((C, W), bool)
    for c in range(C):
        for t in range(W):
            m = max(ab[cc, tt] for cc in range(max(0, c - lc), min(C, c + lc + 1))
                    for tt in range(max(0, t - lt), min(W, t + lt + 1)))
            pk[c, t] = ab[c, t] == m and ab[c, t] > 6.0
    want = []
    for c in range(C):
        for t in range(nt, W - nt):
            if not pk[c, t]:
                continue
            n = sum(pk[cc, tt] for cc in range(max(0, c - gc), min(C, c + gc + 1))
                    for tt in range(max(0, t - gt), min(W, t + gt + 1)))
            if n == 1:
                want.append((c, t))
    c, t = to.isolated_peaks(X, nt, 6.0, (lc, lt), (gc, gt))
    assert list(zip(c.tolist(), t.tolist())) == want
    clips, c2, t2 = to.extract_snippets(X, nt=nt, twav_min=2, Th_single_ch=6.0, loc_range=(lc, lt),
                                        long_range=(gc, gt))
    for i, (cc, tt) in enumerate(want):
        assert np.array_equal(clips[i], X[cc, tt - 2:tt - 2 + nt])


# ---- wPCA ------------------------------------------------------------------------------------
def test_wpca_matches_truncated_svd():
    from sklearn.decomposition import TruncatedSVD
    rng = np.random.default_rng(7)
    nt, n = 61, 4000
    m = np.arange(nt)
    basis = np.stack([np.exp(-(m - 20) ** 2 / (2 * s ** 2)) * np.cos(f * (m - 20))
                      for s, f in [(2, 0), (4, 0.3), (6, 0.1), (3, 0.8), (8, 0.05), (5, 0.5)]])
    coef = rng.standard_normal((n, 6)) * np.array([5, 3, 2, 1.2, 0.7, 0.4])
    clips = coef @ basis + 0.05 * rng.standard_normal((n, nt))
    comp, sv = to.wpca(clips, 6)
    Q = clips / np.sqrt((clips ** 2).sum(1, keepdims=True))
    for alg in ("arpack", "randomized"):
        ts = TruncatedSVD(n_components=6, algorithm=alg, n_iter=20, random_state=0).fit(Q)
        assert np.abs(comp - ts.components_).max() < 1e-6, alg
        assert np.allclose(sv, ts.singular_values_, rtol=1e-8), alg
    assert np.allclose(comp @ comp.T, np.eye(6), atol=1e-12)
