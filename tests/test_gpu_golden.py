"""The device path against the committed golden fixtures (tests/golden/): the reference's integer
KAT through the projection + score contraction (exact: every partial sum is an integer below
2^24), the recorded small window through the whole sorter (spike list equal to the FP64 oracle's,
amplitudes within 1e-3 relative, features within 1e-3 of their scale), and the training-side
builders on the recorded inputs (FP64, 1e-9; snippets exact)."""
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

G = Path(__file__).resolve().parent / "golden"
DEV = "cuda:0"


def test_kat_scores_on_device():
    from livespikesorter_amd.sorter import Sorter
    print("case: golden KAT scores", flush=True)
    z = np.load(G / "kat_conv1d.npz")
    U = z["wall3"].astype(np.float32)                   # [T=2][K=3][C=2]
    art = dict(U=U, wPCA=z["wpca"].astype(np.float32), Wrot=np.eye(2, dtype=np.float32), drift=None,
               hp_filter=np.ones(1, np.float32), xc=np.zeros(2, np.float32), yc=np.arange(2, dtype=np.float32),
               nt0min=1, nC=1)
    with Sorter(art, max_window=16, max_spikes=64) as s:
        x = torch.from_numpy(np.ascontiguousarray(z["batch"].T.astype(np.float32))).to(DEV)   # [W][C]
        B = s.stage_scores(x).cpu().numpy()             # [W][T]
    np.testing.assert_array_equal(B.T, z["B"].astype(np.float32))


def test_window_fixture_on_device():
    from livespikesorter_amd.sorter import Sorter
    print("case: golden window (32 ch, 24 units, W'=3122)", flush=True)
    z = np.load(G / "window_small.npz")
    art = {k: z[k] for k in ("U", "wPCA", "Wrot", "drift", "hp_filter", "xc", "yc")}
    art.update(Th=float(z["Th"]), nt0min=int(z["nt0min"]), nC=int(z["nC"]))
    with Sorter(art, max_window=z["window"].shape[0], max_passes=100) as s:
        sp = s.sort_window(z["window"], int(z["window_start"])).sorted()
    o = np.lexsort((z["template"], z["sample"]))
    g = np.lexsort((sp.template, sp.sample))
    np.testing.assert_array_equal(sp.sample[g], z["sample"][o])
    np.testing.assert_array_equal(sp.template[g], z["template"][o])
    amp = z["amplitude"][o]
    assert np.abs(sp.amplitude[g] - amp).max() <= 1e-3 * np.abs(amp).max()
    F = z["features"][o]
    assert np.abs(sp.features[g] - F).max() <= 1e-3 * np.abs(F).max()
    assert np.abs(sp.y[g] - z["y"][o]).max() <= 0.05


def test_train_fixture_on_device():
    from livespikesorter_amd import train
    print("case: golden training builders", flush=True)
    z = np.load(G / "train_small.npz")
    xc, yc = z["xc"], z["yc"]
    C = len(xc)
    cc = torch.zeros((C, C), dtype=torch.float64, device=DEV)
    train.covariance_add(torch.from_numpy(z["X"]).to(DEV), cc)
    np.testing.assert_allclose(cc.cpu().numpy(), z["CC"], rtol=1e-12, atol=1e-13)
    W = train.whitening_local(torch.from_numpy(z["CC"]).to(DEV), xc, yc, int(z["nrange"])).cpu().numpy()
    assert np.abs(W - z["Wrot_local"]).max() < 1e-9 * np.abs(z["Wrot_local"]).max()
    iK, _ = train.drift_inverse(xc, yc, float(z["sig"]), float(z["lam"]))
    assert np.abs(iK.cpu().numpy() - z["iKxx"]).max() < 1e-9 * np.abs(z["iKxx"]).max()
    M = train.get_drift_matrix(xc, yc, z["dshift"], torch.from_numpy(z["iKxx"]).to(DEV), float(z["sig"]),
                               yblk=z["yblk"]).cpu().numpy()
    assert np.abs(M - z["M"]).max() < 1e-9 * np.abs(z["M"]).max()
    xw = torch.from_numpy(np.ascontiguousarray(z["planted"].T)).to(DEV)
    clips, chan, tim, n = train.extract_snippets(xw)
    assert n == len(z["snip_chan"])
    np.testing.assert_array_equal(chan.cpu().numpy(), z["snip_chan"])
    np.testing.assert_array_equal(tim.cpu().numpy(), z["snip_time"])
    np.testing.assert_array_equal(clips.cpu().numpy(), z["snip_clips"])
    w, sv = train.wpca(torch.from_numpy(z["pca_clips"]).to(DEV), 6)
    assert np.abs(w.cpu().numpy() - z["wpca"]).max() < 1e-9
    np.testing.assert_allclose(sv.cpu().numpy(), z["sv"], rtol=1e-12)
