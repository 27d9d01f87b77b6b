"""The committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py) against the
oracle: the reference's integer KAT outputs (exact), Kilosort's universal temporal basis as a real
wPCA (orthonormal; its ctc zero-lag equals the template norm), and the oracle reproducing its own
recorded window and training-side outputs (regression of the checker itself)."""
from pathlib import Path

import numpy as np
import torch

from oracle import ks_oracle as ko
from oracle import train_oracle as to

G = Path(__file__).resolve().parent / "golden"
D = torch.float64


def test_kat_fixture():
    z = np.load(G / "kat_conv1d.npz")
    X = torch.tensor(z["batch"], dtype=D)
    W = torch.tensor(z["wpca"], dtype=D)
    U = torch.tensor(z["wall3"], dtype=D)
    M = int(z["M"])
    B = ko.project(X, U, W)
    np.testing.assert_array_equal(B.numpy(), z["B"])
    Cfmax, imax = ko.score_argmax(B, torch.tensor(z["nm"], dtype=D), M)
    np.testing.assert_allclose(Cfmax.numpy(), z["Cfmax"], rtol=1e-15)
    np.testing.assert_array_equal(imax.numpy(), z["imax"])
    Cmax = torch.nn.functional.max_pool1d(Cfmax[None, None], 2 * M + 1, stride=1, padding=M)[0, 0]
    np.testing.assert_allclose(Cmax.numpy(), z["Cmax"], rtol=1e-15)


def test_universal_wpca_fixture():
    w = np.load(G / "ks_universal_wpca.npy").astype(np.float64)
    assert w.shape == (6, 61)
    assert np.abs(w @ w.T - np.eye(6)).max() < 1e-5
    rng = np.random.default_rng(0)
    U = torch.from_numpy(rng.standard_normal((5, 6, 12)))
    wo = torch.from_numpy(np.linalg.qr(w.T)[0].T.copy())            # exactly orthonormal rows
    ctc = ko.prepare_matching(U, wo)
    nm = ko.template_norms(U)
    np.testing.assert_allclose(ctc[torch.arange(5), torch.arange(5), 61].numpy(), nm.numpy(), rtol=1e-12)


def test_window_fixture_oracle_reproduces():
    z = np.load(G / "window_small.npz")
    art = {k: z[k] for k in ("U", "wPCA", "Wrot", "drift", "hp_filter", "xc", "yc")}
    art.update(Th=float(z["Th"]), nt0min=int(z["nt0min"]), nC=int(z["nC"]))
    r = ko.sort_window(torch.from_numpy(z["window"].T.copy()), art, int(z["window_start"]), dtype=D, max_passes=100)
    np.testing.assert_array_equal(r.samples, z["sample"])
    np.testing.assert_array_equal(r.templates, z["template"])
    np.testing.assert_allclose(r.amps, z["amplitude"], rtol=1e-10)
    np.testing.assert_allclose(r.tF, z["features"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r.ys, z["y"], rtol=1e-10)
    np.testing.assert_allclose(np.linalg.norm(r.X.numpy(), axis=1), z["whitened_row_norms"], rtol=1e-12)


def test_train_fixture_oracle_reproduces():
    z = np.load(G / "train_small.npz")
    xc, yc = z["xc"], z["yc"]
    CC = to.covariance([z["X"].astype(np.float64)])
    np.testing.assert_allclose(CC, z["CC"], rtol=1e-13)
    np.testing.assert_allclose(to.whitening_local(z["CC"], xc, yc, int(z["nrange"])), z["Wrot_local"], rtol=1e-10,
                               atol=1e-12)
    iK = to.drift_inverse(xc, yc, float(z["sig"]), float(z["lam"]))
    np.testing.assert_allclose(iK, z["iKxx"], rtol=1e-10, atol=1e-12)
    M = to.drift_matrix(xc, yc, z["dshift"], z["iKxx"], float(z["sig"]), yblk=z["yblk"])
    np.testing.assert_allclose(M, z["M"], rtol=1e-10, atol=1e-12)
    clips, c, t = to.extract_snippets(z["planted"])
    np.testing.assert_array_equal(c, z["snip_chan"])
    np.testing.assert_array_equal(t, z["snip_time"])
    np.testing.assert_array_equal(clips, z["snip_clips"])
    comp, sv = to.wpca(z["pca_clips"], 6)
    np.testing.assert_allclose(comp, z["wpca"], atol=1e-12)
    np.testing.assert_allclose(sv, z["sv"], rtol=1e-12)
