"""Template refinement on the device (lss_residual_energy / lss_residual_energy_grad, the adjoint
kernel) against the CPU oracle, on the drift workload: 384 channels, 600 templates, a 2 s window
at 58-60 s of a probe drifting 0-20 um over 60 s, spikes fixed from the forward pass.

Tolerances (written where they are checked):
* loss: |gpu - fp64| <= 1e-5 |fp64| (FP32 residual, FP64 sum);
* gradients vs the FP64 torch-autograd reference (residual recomputed under
  torch.utils.checkpoint): max |gpu - fp64| <= 1e-4 max |fp64 gradient|, per tensor;
* gradients vs FP64 central differences (h = 1e-4) of the loss on 200 random coordinates
  (192 of U, 8 of wPCA): |gpu - fd| <= 1e-4 max |fp64 gradient| + 1e-5 |fd|;
* the autograd.Function returns the adjoint's gradients bit for bit; two calls agree bit for bit;
* loss vs the pursuit's own residual energy: 1e-5 relative (different FP32 summation order).
"""
import numpy as np
import pytest
import torch
from torch.utils.checkpoint import checkpoint

from oracle import ks_oracle as ko

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def drift_window():
    from livespikesorter_amd import synth
    print("[lss] refine: generating the drift window (384 ch, 600 templates, 2 s)", flush=True)
    s = synth.make(4, seed=4)
    a = s.art
    X = ko.padded_batch(s.rec, 0, a["NT"], a["nt"], torch.float32)
    win = torch.from_numpy(np.ascontiguousarray(X.numpy().T.astype(np.int16))).to(DEV)
    return a, win


@pytest.fixture(scope="module")
def case4(drift_window):
    from livespikesorter_amd.refine import SpikeSet
    from livespikesorter_amd.sorter import Sorter
    a, win = drift_window
    sorter = Sorter(a, max_window=win.shape[0], max_passes=100, drop_negative=False)
    x_wc = sorter.stage_whiten(sorter.stage_filter(win)).contiguous()          # [W][C] float32
    sp = sorter.sort_window(win, ko.window_start(0, a["NT"], a["nt"]))
    spikes = SpikeSet.from_spikes(sp)
    U = torch.from_numpy(a["U"]).to(DEV).contiguous()
    wp = torch.from_numpy(a["wPCA"]).to(DEV).contiguous()
    print(f"[lss] refine: {len(spikes)} spikes in {sp.passes} passes", flush=True)
    yield sorter, x_wc, spikes, U, wp
    sorter.close()


@pytest.fixture(scope="module")
def reference(case4):
    """FP64 loss and autograd gradients on the CPU, the residual recomputed in the backward."""
    sorter, x_wc, spikes, U, wp = case4
    X64 = x_wc.cpu().double().T.contiguous()
    col, tmp = spikes.col.cpu().long(), spikes.tmp.cpu().long()
    amp = spikes.amp.cpu().double()
    U64 = U.cpu().double().requires_grad_(True)
    W64 = wp.cpu().double().requires_grad_(True)
    loss = checkpoint(lambda u, w: ko.residual_energy(X64, col, tmp, amp, u, w)[0], U64, W64, use_reentrant=False)
    gU, gW = torch.autograd.grad(loss, (U64, W64))
    return float(loss), gU.numpy(), gW.numpy(), X64.numpy(), col.numpy(), tmp.numpy(), amp.numpy()


def test_refine_loss_and_gradients_vs_fp64_autograd(case4, reference):
    print("[lss] test_refine_loss_and_gradients_vs_fp64_autograd", flush=True)
    sorter, x_wc, spikes, U, wp = case4
    loss_ref, gU_ref, gW_ref = reference[:3]
    loss, gU, gW = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    print(f"[lss]   loss gpu {loss:.6e} fp64 {loss_ref:.6e}; timings {sorter.refine_timings()}", flush=True)
    assert abs(loss - loss_ref) <= 1e-5 * abs(loss_ref)
    gU, gW = gU.cpu().numpy(), gW.cpu().numpy()
    assert np.abs(gU - gU_ref).max() <= 1e-4 * np.abs(gU_ref).max()
    assert np.abs(gW - gW_ref).max() <= 1e-4 * np.abs(gW_ref).max()
    # the loss-only entry point gives the same loss
    assert sorter.residual_energy(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp) == loss


def _loss_on_columns(X, col, tmp, amp, U, Wp, cols):
    """sum over window columns `cols` of |X[:, w] - sum_s amp_s wave_s[w - col_s + nt/2]|^2,
    recomputed from scratch in FP64 (every other column is unchanged by the perturbations used)."""
    K, nt = Wp.shape
    h = nt // 2
    cols = np.unique(cols)
    cols = cols[(cols >= 0) & (cols < X.shape[1])]
    R = X[:, cols].T.copy()                                    # [ncols][C]
    for m in range(nt):
        w = col - h + m
        i = np.searchsorted(cols, w)
        ok = i < len(cols)
        ok[ok] &= cols[i[ok]] == w[ok]
        if not ok.any():
            continue
        wav = np.einsum("skc,k->sc", U[tmp[ok]], Wp[:, m])
        np.add.at(R, i[ok], -(amp[ok, None] * wav))
    return float((R * R).sum())


def test_refine_gradients_vs_central_differences(case4, reference):
    print("[lss] test_refine_gradients_vs_central_differences (200 coordinates)", flush=True)
    sorter, x_wc, spikes, U, wp = case4
    _, gU_ref, gW_ref, X, col, tmp, amp = reference
    _, gU, gW = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    gU, gW = gU.cpu().numpy(), gW.cpu().numpy()
    U64, W64 = U.cpu().double().numpy(), wp.cpu().double().numpy()
    T, K, C = U64.shape
    nt = W64.shape[1]
    h, step = nt // 2, 1e-4
    rng = np.random.default_rng(44)
    used = np.unique(tmp)
    scaleU, scaleW = np.abs(gU_ref).max(), np.abs(gW_ref).max()
    for _ in range(192):
        t, k, c = int(rng.choice(used)), int(rng.integers(K)), int(rng.integers(C))
        mine = tmp == t
        cols = (col[mine][:, None] - h + np.arange(nt)).ravel()
        Up, Um = U64.copy(), U64.copy()
        Up[t, k, c] += step
        Um[t, k, c] -= step
        fd = (_loss_on_columns(X, col, tmp, amp, Up, W64, cols) -
              _loss_on_columns(X, col, tmp, amp, Um, W64, cols)) / (2 * step)
        assert abs(gU[t, k, c] - fd) <= 1e-4 * scaleU + 1e-5 * abs(fd), (t, k, c, gU[t, k, c], fd)
    for _ in range(8):
        k, m = int(rng.integers(K)), int(rng.integers(nt))
        cols = col - h + m
        Wp, Wm = W64.copy(), W64.copy()
        Wp[k, m] += step
        Wm[k, m] -= step
        fd = (_loss_on_columns(X, col, tmp, amp, U64, Wp, cols) -
              _loss_on_columns(X, col, tmp, amp, U64, Wm, cols)) / (2 * step)
        assert abs(gW[k, m] - fd) <= 1e-4 * scaleW + 1e-5 * abs(fd), (k, m, gW[k, m], fd)


def test_refine_autograd_function_and_determinism(case4):
    print("[lss] test_refine_autograd_function_and_determinism", flush=True)
    from livespikesorter_amd.refine import residual_energy
    sorter, x_wc, spikes, U, wp = case4
    l1, gU1, gW1 = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    l2, gU2, gW2 = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    assert l1 == l2 and torch.equal(gU1, gU2) and torch.equal(gW1, gW2)
    Ur = U.clone().requires_grad_(True)
    Wr = wp.clone().requires_grad_(True)
    loss = residual_energy(sorter, x_wc, spikes, Ur, Wr)
    (2.0 * loss).backward()
    assert float(loss) == l1
    assert torch.equal(Ur.grad, 2.0 * gU1) and torch.equal(Wr.grad, 2.0 * gW1)


def test_refine_loss_is_pursuit_residual_energy(case4):
    """The loss of the forward's own spikes and templates is the energy of the residual the
    device pursuit leaves (same spikes, amplitudes and waves; summation order differs)."""
    print("[lss] test_refine_loss_is_pursuit_residual_energy", flush=True)
    sorter, x_wc, spikes, U, wp = case4
    B = sorter.stage_scores(x_wc)
    Xr = x_wc.clone()
    sp = sorter.stage_pursuit(B, Xr, 0)
    assert len(sp) == len(spikes)
    e_pursuit = float((Xr.double() ** 2).sum())
    loss = sorter.residual_energy(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    assert abs(loss - e_pursuit) <= 1e-5 * e_pursuit


def test_refine_step_lowers_the_loss(case4):
    """A small gradient step on the drift window lowers the residual energy by what the gradient
    predicts to first order (lr |g|^2, here 1e-5 of the loss; the second-order term is ~2% of it)."""
    print("[lss] test_refine_step_lowers_the_loss", flush=True)
    from livespikesorter_amd.refine import refine_step
    sorter, x_wc, spikes, U, wp = case4
    loss0, gU, gW = sorter.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
    g2 = float((gU.double() ** 2).sum() + (gW.double() ** 2).sum())
    lr = 1e-5 * loss0 / g2
    _, U1, W1 = refine_step(sorter, x_wc, spikes, U, wp, lr)
    loss1 = sorter.residual_energy(x_wc, spikes.col, spikes.tmp, spikes.amp, U1.contiguous(), W1.contiguous())
    pred = lr * g2
    print(f"[lss]   loss {loss0:.6e} -> {loss1:.6e}; decrease {loss0 - loss1:.3f}, first order {pred:.3f}", flush=True)
    assert loss1 < loss0
    assert abs((loss0 - loss1) - pred) <= 0.1 * pred


def test_refine_templates_step(drift_window, case4):
    """lss_refine_templates on the sorter's own last window: the loss drops, the loss before equals
    the refinement loss of that window's spikes, and the rebuilt tables are those of
    U0 - step * dL/dU, wPCA0 - step * dL/dwPCA."""
    print("[lss] test_refine_templates_step", flush=True)
    from livespikesorter_amd.sorter import Sorter
    a, win = drift_window
    _, x_wc, spikes, U, wp = case4
    s2 = Sorter(a, max_window=win.shape[0], max_passes=100, drop_negative=False)
    try:
        s2.sort_window(win, ko.window_start(0, a["NT"], a["nt"]))
        info = s2.refine_templates()
        print(f"[lss]   {info}", flush=True)
        loss0, gU, gW = s2.residual_energy_grad(x_wc, spikes.col, spikes.tmp, spikes.amp, U, wp)
        assert info["loss_after"] < info["loss_before"]
        assert abs(info["loss_before"] - loss0) <= 1e-5 * loss0
        U1 = (U.double() - info["step"] * gU.double()).float()
        W1 = (wp.double() - info["step"] * gW.double()).float()
        tb = s2.tables()
        np.testing.assert_allclose(tb["nm"], ko.template_norms(U1.cpu().double()).numpy(), rtol=1e-4)
        ctc1 = ko.prepare_matching(U1.cpu().double(), W1.cpu().double()).numpy()
        assert np.abs(np.transpose(tb["ctc"], (2, 0, 1)) - ctc1).max() <= 1e-4 * np.abs(ctc1).max()
        l1 = s2.residual_energy(x_wc, spikes.col, spikes.tmp, spikes.amp, U1.contiguous(), W1.contiguous())
        assert abs(l1 - info["loss_after"]) <= 1e-4 * l1
        # the refined sorter still sorts the window
        sp2 = s2.sort_window(win, ko.window_start(0, a["NT"], a["nt"]))
        assert len(sp2) > 0.9 * len(spikes)
    finally:
        s2.close()
