"""The oss_input directory as the reference's exporter writes it and OnlineSpikesV2 loads it
(src/Python/run_online_spikes.py:533-576, src/SpikeSorter/OnlineSpikesV2.cu:435-457, 546-610):
iCC.npy / iU.npy / Ucc.npy are used as loaded (not rebuilt), ctc.npy is compared with the FP64
rebuild, misc.txt's duplicate_spike_bins reaches the handle.

The exporter's iCC comes from numpy's default (unstable) argsort, whose order among equidistant
channels is unspecified (kilosort/template_matching.py:22).  The test writes an iCC whose tie order
differs from the stable rule the library uses when it builds iCC itself, and checks the PC features
and positions of the device pursuit against the oracle's extraction with that same iCC (max |gpu -
oracle| <= 1e-4 of the max |oracle feature|, positions within 1e-3 um), and that they differ from
those of the stable rule (so the file is what was used)."""
import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _tie_reversed_icc(xc, yc, nC):
    """iCC [nC][C]: each channel's nearest channels by distance, ties in reverse index order (the
    stable rule puts the lower index first)."""
    C = len(xc)
    out = np.zeros((nC, C), np.int64)
    for c in range(C):
        ds = (xc - xc[c]) ** 2 + (yc - yc[c]) ** 2
        order = sorted(range(C), key=lambda j: (ds[j], -j))
        out[:, c] = order[:nC]
    return out


@pytest.fixture(scope="module")
def case(synth1, tmp_path_factory):
    from livespikesorter_amd.export import write_oss_input
    a = synth1.art
    d = write_oss_input(dict(a, dup_bins=11), tmp_path_factory.mktemp("loader") / "oss_input", with_ctc=True)
    xc, yc = a["xc"].astype(np.float64), a["yc"].astype(np.float64)
    nC = int(a["nC"])
    iCC = _tie_reversed_icc(xc, yc, nC)
    stable = np.load(d / "iCC.npy")
    assert (iCC != stable).any()                      # the tie order differs on this probe
    U = np.asarray(a["U"], np.float32)
    iU = np.argmax((U.astype(np.float64) ** 2).sum(1), -1).astype(np.int64)
    Ucc = np.ascontiguousarray(U[np.arange(U.shape[0])[None, :], :, iCC[:, iU]])   # [nC][T][K]
    np.save(d / "iCC.npy", iCC)                       # int64, as the exporter writes it
    np.save(d / "iU.npy", iU.astype(np.int32))        # any integer dtype is accepted
    np.save(d / "Ucc.npy", Ucc)
    return d, iCC, iU, Ucc


def test_directory_tables_and_misc(case):
    from livespikesorter_amd.sorter import Sorter
    d, _, _, _ = case
    print("[lss] test_directory_tables_and_misc", flush=True)
    with Sorter.from_dir(d, max_window=60122, max_passes=100) as s:
        info = s.load_info()
        assert s.params.dup_bins == 11
    print(f"[lss]   {info}", flush=True)
    assert info["icc_given"] == info["iu_given"] == info["ucc_given"] == info["ctc_in_dir"] == 1
    # the exporter's ctc is the FP64 product rounded to FP32: within FP32 rounding of the rebuild
    assert info["ctc_max_abs_diff"] <= 1e-6 * info["ctc_max_abs"]


def test_directory_icc_is_used_for_features(case, synth1):
    from livespikesorter_amd.sorter import Sorter
    d, iCC, iU, Ucc = case
    a = synth1.art
    print("[lss] test_directory_icc_is_used_for_features batch 1", flush=True)
    X = ko.padded_batch(synth1.rec, 1, a["NT"], a["nt"], torch.float64)
    ws = ko.window_start(1, a["NT"], a["nt"])
    hp = torch.from_numpy(a["hp_filter"]).double()
    wm = torch.as_tensor(a["drift"]).double() @ torch.as_tensor(a["Wrot"]).double()
    x32 = ko.filter_batch(X, hp, wm).float()
    U32, w32 = torch.from_numpy(a["U"]), torch.from_numpy(a["wPCA"])
    B32 = ko.project(x32, U32, w32)
    with Sorter.from_dir(d, max_window=60122, max_passes=100) as s:
        tb = s.tables()
        got = s.stage_pursuit(B32.T.contiguous().to(DEV), x32.T.contiguous().to(DEV), ws)
    with Sorter(a, max_window=60122, max_passes=100) as s2:            # stable-rule iCC
        got_stable = s2.stage_pursuit(B32.T.contiguous().to(DEV), x32.T.contiguous().to(DEV), ws)
    ctc_ks = torch.from_numpy(tb["ctc"]).permute(2, 0, 1).contiguous()
    res = ko.run_matching(x32, U32, w32, ctc_ks, Th=a["Th"], max_passes=100,
                          nm=torch.from_numpy(tb["nm"]), waves=torch.from_numpy(tb["waves"]), B=B32)
    iCC_t, iU_t, Ucc_t = torch.from_numpy(iCC), torch.from_numpy(iU), torch.from_numpy(Ucc)
    tF = ko.extract_features(res.Xres, res.st, res.amps, iCC_t, iU_t, Ucc_t, w32).numpy()
    xs, ys = ko.spike_positions(torch.from_numpy(tF), res.st, iCC_t, iU_t, a["xc"], a["yc"])
    np.testing.assert_array_equal(got.column, res.st[:, 0].numpy())
    assert got.features.shape == tF.shape
    assert rel_err(got.features, tF) <= 1e-4
    assert np.abs(got.x - xs.numpy()).max() <= 1e-3
    assert np.abs(got.y - ys.numpy()).max() <= 1e-3
    # same spikes, different channel order of the features: the file's iCC is what was used
    np.testing.assert_array_equal(got.amplitude, got_stable.amplitude)
    assert rel_err(got_stable.features, tF) > 1e-2
