"""Output step (host-only C ABI functions) and artifact export: checked against brute force and
against the oracle's restatement of Kilosort's duplicate removal (postprocessing.py:10-30)."""
import numpy as np

from oracle import ks_oracle as ko


def test_label_spikes_bruteforce():
    from livespikesorter_amd.output import label_spikes
    rng = np.random.default_rng(0)
    x, y = rng.uniform(0, 40, 500), rng.uniform(0, 400, 500)
    cx, cy = rng.uniform(0, 40, 37), rng.uniform(0, 400, 37)
    got = label_spikes(x, y, cx, cy)
    d = np.sqrt((x[:, None].astype(np.float32) - cx.astype(np.float32)) ** 2 +
                (y[:, None].astype(np.float32) - cy.astype(np.float32)) ** 2)
    np.testing.assert_array_equal(got, np.argmin(d, 1))
    # ties: the first centroid wins
    assert label_spikes([1.0], [0.0], [0.0, 2.0], [0.0, 0.0])[0] == 0


def test_dedup_matches_kilosort_rule_and_carries_state():
    from livespikesorter_amd.output import Deduplicator
    rng = np.random.default_rng(1)
    t = np.sort(rng.integers(0, 5000, 800))
    c = rng.integers(0, 6, 800)
    dd = Deduplicator(6, dt=7)
    s, l, _ = dd(t, c)
    rt, rc, _ = ko.remove_duplicates(t, c, dt=7)
    np.testing.assert_array_equal(s, rt)
    np.testing.assert_array_equal(l, rc)
    # split into two windows with an overlap: the state removes the re-found spikes
    dd2 = Deduplicator(6, dt=7)
    s1, l1, _ = dd2(t[t < 2600], c[t < 2600])
    s2, l2, _ = dd2(t[t >= 2500], c[t >= 2500])
    both = np.concatenate([s1, s2])
    np.testing.assert_array_equal(np.sort(both), rt)


def test_write_spikes_txt(tmp_path):
    from livespikesorter_amd.output import write_spikes_txt
    p = tmp_path / "spikeOutput.txt"
    write_spikes_txt(p, [10, 20], [1, 2], [0.5, 1.25], [40.0, 60.0], recording_offset=1000, append=False)
    write_spikes_txt(p, [30], [3], [2.0], [80.0], recording_offset=1000)
    assert p.read_text().splitlines() == ["1010,1,0.5,40", "1020,2,1.25,60", "1030,3,2,80"]


def test_export_oss_input(tmp_path, small_synth):
    from livespikesorter_amd.export import write_oss_input
    a = small_synth.art
    d = write_oss_input(a, tmp_path / "oss_input")
    U = np.load(d / "Wall3.npy")
    np.testing.assert_array_equal(U, a["U"])
    np.testing.assert_array_equal(np.load(d / "drift_matrix.npy").T, a["drift"])
    misc = dict(l.split(":") for l in (d / "misc.txt").read_text().split())
    assert misc["nt0min"] == "20" and misc["Th_learned"] == "8"
    iCC = np.load(d / "iCC.npy")
    import torch
    iCC_o, iU_o, Ucc_o = ko.prepare_extract(a["xc"].astype(np.float64), a["yc"].astype(np.float64),
                                            torch.from_numpy(a["U"]).double(), int(a["nC"]))
    np.testing.assert_array_equal(iCC, iCC_o.numpy())
    np.testing.assert_array_equal(np.load(d / "iU.npy"), iU_o.numpy())


def test_load_cluster_centroids(tmp_path, small_synth):
    """cluster_centroids.npy as OnlineSpikesV2 reads it ([x1, y1, x2, y2, ...]): written by the
    exporter, read back in C or Fortran order and any numeric dtype; missing, odd-length and
    too-small outputs are reported."""
    import ctypes as ct
    import pytest
    from livespikesorter_amd._lib import LssError, lib
    from livespikesorter_amd.export import write_oss_input
    from livespikesorter_amd.output import load_cluster_centroids
    xy = np.array([[0.0, 10.0], [32.0, 55.5], [16.0, 300.25]], np.float32)
    d = write_oss_input(small_synth.art, tmp_path / "a", cluster_xy=xy)
    cx, cy = load_cluster_centroids(d)
    np.testing.assert_array_equal(cx, xy[:, 0])
    np.testing.assert_array_equal(cy, xy[:, 1])
    np.save(d / "cluster_centroids.npy", np.asfortranarray(xy.astype(np.float64)))
    cx, cy = load_cluster_centroids(d)
    np.testing.assert_array_equal(cx, xy[:, 0])
    np.testing.assert_array_equal(cy, xy[:, 1])
    np.save(d / "cluster_centroids.npy", xy.reshape(-1)[:5])
    with pytest.raises(LssError, match="IO"):
        load_cluster_centroids(d)
    with pytest.raises(LssError, match="IO"):
        load_cluster_centroids(tmp_path / "missing")
    np.save(d / "cluster_centroids.npy", xy)
    n = ct.c_int32()
    buf = np.zeros(2, np.float32)
    st = lib().lss_load_cluster_centroids(str(d).encode(), buf.ctypes.data, buf.ctypes.data, 2, ct.byref(n))
    assert st == 4 and n.value == 3


def test_load_channel_map(tmp_path, small_synth):
    """channelMap.npy (the channels OnlineSpikesV2 hands its DataSocket): written by the exporter
    (the channels 0..C-1 by default, or the given list), read back from int32, int64 and float
    files; missing files, negative or fractional entries and too-small outputs are reported."""
    import ctypes as ct
    import pytest
    from livespikesorter_amd._lib import LssError, lib
    from livespikesorter_amd.export import write_oss_input
    from livespikesorter_amd.output import load_channel_map
    C = np.asarray(small_synth.art["U"]).shape[2]
    d = write_oss_input(small_synth.art, tmp_path / "a")
    np.testing.assert_array_equal(load_channel_map(d), np.arange(C))
    cm = (np.arange(C)[::-1] * 3 + 7).astype(np.int64)
    d = write_oss_input(small_synth.art, tmp_path / "b", channel_map=cm)
    np.testing.assert_array_equal(load_channel_map(d), cm)
    np.save(d / "channelMap.npy", cm.astype(np.float64))
    np.testing.assert_array_equal(load_channel_map(d), cm)
    for bad in (np.array([1, -2, 3]), np.array([0.5, 1.0])):
        np.save(d / "channelMap.npy", bad)
        with pytest.raises(LssError, match="IO"):
            load_channel_map(d)
    with pytest.raises(LssError, match="IO"):
        load_channel_map(tmp_path / "missing")
    with pytest.raises(ValueError):
        write_oss_input(small_synth.art, tmp_path / "c", channel_map=np.arange(C + 1))
    np.save(d / "channelMap.npy", cm.astype(np.int32))
    n = ct.c_int32()
    buf = np.zeros(2, np.int32)
    assert lib().lss_load_channel_map(str(d).encode(), buf.ctypes.data, 2, ct.byref(n)) == 4 and n.value == C
