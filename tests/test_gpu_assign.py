"""Feature-space cluster assignment on the device (lss_cluster_centroids_pca, lss_assign_clusters)
against the oracle's restatement of src/Python/OnlineKilosort.py:211-284.

* centroids: equal to the reference's gather Wall[t, pc_feature_ind[t], :] (exact: a copy).
* labels: equal to the FP64 oracle's for every spike whose best and second-best in-range scores
  differ by more than 1e-4 of the score scale and whose distance to every centroid differs from R
  by more than 1e-3 um (FP32 on the device; the reference also computes in FP32, with its own
  summation order); the excluded near-ties are counted and must be rare.
* a real window's spikes (features and positions from the sorter), random features at D = 60 and
  at the largest D = 256 (the LDS tiles then exceed 64 KB), clusters without a position, no
  cluster in range, and no spikes."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu


def _fp64_reference(tF, mu, xs, ys, cluster_xy, R):
    """Oracle labels plus, per spike, whether it is a near-tie (scores or distance to R)."""
    X = np.asarray(tF, np.float64).reshape(len(tF), -1)
    M = np.asarray(mu, np.float64)
    K = M.shape[0]
    sc = 2 * X @ M.T - (M ** 2).sum(1)
    cx = np.full(K, np.nan)
    cy = np.full(K, np.nan)
    for k, (x, y) in cluster_xy.items():
        cx[k], cy[k] = x, y
    with np.errstate(invalid="ignore"):
        d = np.sqrt((np.asarray(xs)[:, None] - cx) ** 2 + (np.asarray(ys)[:, None] - cy) ** 2)
        near_r = (np.abs(d - R) < 1e-3).any(axis=1)
        mask = d <= R
    s = np.where(mask, sc, -np.inf)
    top2 = -np.sort(-s, axis=1)[:, :2] if K > 1 else np.concatenate([s, np.full_like(s, -np.inf)], 1)
    scale = np.abs(sc).max()
    with np.errstate(invalid="ignore"):
        near_tie = np.isfinite(top2[:, 1]) & (top2[:, 0] - top2[:, 1] <= 1e-4 * scale)
    ref = ko.assign_to_clusters(tF, mu, xs, ys, cluster_xy, R)
    return ref, near_tie | near_r


def _check(tF, mu, xs, ys, cluster_xy, R, max_skip=0.01):
    from livespikesorter_amd.output import assign_to_clusters
    got = assign_to_clusters(tF, mu, xs, ys, cluster_xy, R)
    ref, skip = _fp64_reference(tF, mu, xs, ys, cluster_xy, R)
    assert skip.mean() <= max_skip, skip.mean()
    np.testing.assert_array_equal(got[~skip], ref[~skip])
    return got


def test_centroids_and_labels_of_a_sorted_window(synth1):
    from livespikesorter_amd.output import cluster_centroids_pca
    from livespikesorter_amd.sorter import Sorter
    print("[lss] test_centroids_and_labels_of_a_sorted_window", flush=True)
    a = synth1.art
    X = ko.padded_batch(synth1.rec, 0, a["NT"], a["nt"], torch.float64)
    win = np.ascontiguousarray(X.numpy().T.astype(np.int16))
    with Sorter(a, max_window=win.shape[0], max_passes=100) as so:
        sp = so.sort_window(win, ko.window_start(0, a["NT"], a["nt"]))
        cent = cluster_centroids_pca(so).cpu().numpy()
        nC = so.nC
    U = torch.from_numpy(np.asarray(a["U"], np.float32))
    iCC, iU, Ucc = ko.prepare_extract(np.asarray(a["xc"], np.float64), np.asarray(a["yc"], np.float64), U, nC)
    Wall = U.numpy().transpose(0, 2, 1)
    pcfi = np.stack([iCC[:, int(iU[t])].numpy() for t in range(U.shape[0])])
    np.testing.assert_array_equal(cent, ko.cluster_centroids_pca(Wall, pcfi))
    # clusters at their templates' peak channels; every other one without a position
    xc, yc = np.asarray(a["xc"]), np.asarray(a["yc"])
    cluster_xy = {t: (xc[int(iU[t])], yc[int(iU[t])]) for t in range(U.shape[0]) if t % 5 != 4}
    assert len(sp) > 100
    got = _check(sp.features, cent, sp.x, sp.y, cluster_xy, 80.0)
    assert (got >= 0).mean() > 0.5


@pytest.mark.parametrize("nC,P,K", [(10, 6, 100), (16, 16, 70)])
def test_random_features(nC, P, K):
    print(f"[lss] test_random_features nC={nC} P={P} K={K}", flush=True)
    rng = np.random.default_rng(nC * 100 + K)
    N = 1000
    tF = rng.normal(size=(N, nC, P)).astype(np.float32)
    mu = rng.normal(size=(K, nC * P)).astype(np.float32)
    cxy = rng.uniform(0, 600, size=(K, 2))
    cluster_xy = {k: (float(cxy[k, 0]), float(cxy[k, 1])) for k in range(K) if k % 9 != 2}
    xs = rng.uniform(-50, 650, N).astype(np.float32)
    ys = rng.uniform(-50, 650, N).astype(np.float32)
    got = _check(tF, mu, xs, ys, cluster_xy, 80.0)
    assert (got == -1).any() and (got >= 0).any()
    assert not np.isin(got, [k for k in range(K) if k % 9 == 2]).any()


def test_edge_cases():
    from livespikesorter_amd.output import assign_to_clusters
    print("[lss] test_assign edge cases", flush=True)
    rng = np.random.default_rng(3)
    tF = rng.normal(size=(5, 2, 3)).astype(np.float32)
    mu = rng.normal(size=(4, 6)).astype(np.float32)
    xs = np.zeros(5, np.float32)
    ys = np.zeros(5, np.float32)
    far = {k: (1000.0, 1000.0) for k in range(4)}
    np.testing.assert_array_equal(assign_to_clusters(tF, mu, xs, ys, far, 80.0), -1)
    np.testing.assert_array_equal(assign_to_clusters(tF, mu, xs, ys, {}, 80.0), -1)
    assert len(assign_to_clusters(tF[:0], mu, xs[:0], ys[:0], far, 80.0)) == 0
    # a single cluster in range takes every spike
    np.testing.assert_array_equal(assign_to_clusters(tF, mu, xs, ys, {2: (10.0, 0.0)}, 80.0), 2)
