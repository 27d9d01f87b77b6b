"""The C++ facade with the shape of OnlineSpikesV2 (include/lss_online_spikes.hpp), driven by
tests/online_facade_check over a recorded stream on a simulated clock, against the same batches
sorted through the Python binding: window [start - nt, start + n + nt) per batch (sample 0
repeated before the stream's start), spikes time-ordered, labelled by template (or by the nearest
of the artifact directory's cluster centroids), deduplicated per label with dt = 7 across batches.  The batch schedule (minimum/maximum window, lag tolerance,
small/big skip) is recomputed here from its definition.  Exact equality: same device path, same
windows."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).with_name("online_facade_check")
FS = 30000.0


def _schedule(n_total, first, mn, mx, nt, tb_ms, lag, small):
    now = min(first, n_total)
    processed = now
    out = []
    while True:
        ct = processed + mn + nt
        now = max(now, min(ct, n_total))
        now = min(now + lag, n_total)
        if now < processed + mn + nt:
            break
        ready = now - nt
        allowed = ready - int(float(tb_ms) * FS / 1000.0)
        start, n, skipped = processed, ready - processed, False
        if n > mx:
            if tb_ms == 0:
                start, skipped = ready - mx, True
            elif processed >= allowed:
                pass
            elif small:
                start, skipped = min(allowed, ready - mx), True
            else:
                start, skipped = ready - mx, True
            n = mx
        out.append((start, n, skipped))
        processed = start + n
    return out


def _expected(sorter, rec, batches, nt, n_labels, dt=7, centroids=None):
    from livespikesorter_amd.output import Deduplicator, label_spikes
    dd = Deduplicator(n_labels, dt)
    rows = []
    for start, n, _ in batches:
        first = start - nt
        idx = np.clip(np.arange(first, first + n + 2 * nt), 0, None)
        sp = sorter.sort_window(np.ascontiguousarray(rec[idx]), first)
        pos = np.arange(len(sp), dtype=np.float32)
        lab = sp.template if centroids is None else label_spikes(sp.x, sp.y, *centroids)
        s, l, kept = dd(sp.sample, lab, pos)
        k = kept.astype(np.int64)
        rows.append(np.stack([s.astype(np.float64), l.astype(np.float64),
                              sp.amplitude[k].astype(np.float64), sp.y[k].astype(np.float64)], 1))
    return np.concatenate(rows) if rows else np.zeros((0, 4))


def _read(path):
    lines = [ln.split(",") for ln in Path(path).read_text().splitlines() if ln.strip()]
    if not lines:
        return np.zeros((0, 4))
    return np.array([[int(a), int(b), float(np.float32(c)), float(np.float32(d))] for a, b, c, d in lines])


@pytest.fixture(scope="module")
def facade_inputs(synth1, tmp_path_factory):
    from livespikesorter_amd.export import write_oss_input
    d = tmp_path_factory.mktemp("facade")
    art_dir = write_oss_input(synth1.art, d / "oss_input")
    rec = np.ascontiguousarray(synth1.rec[:30000], np.int16)
    rec.tofile(d / "rec.i16")
    return d, art_dir, rec


# (min window, max window, lag tolerance ms, clock start, clock lead per wait, small skip)
CASES = [(3000, 3000, 100000, 0, 0, 0),      # fixed batches from the stream's start
         (2000, 4000, 100000, 5000, 0, 0),   # started mid-stream
         (2000, 3000, 100000, 0, 2500, 0),   # lagging within tolerance: catch-up batches
         (2000, 3000, 0, 0, 4500, 0),        # no lag tolerated: newest batch
         (2000, 3000, 150, 0, 6000, 1),      # small skip to the allowed lag, then catch-up
         (2000, 3000, 150, 0, 6000, 0)]      # big skip to the newest batch


@pytest.mark.parametrize("mn,mx,tb,first,lag,small", CASES)
def test_facade_matches_python_path(facade_inputs, mn, mx, tb, first, lag, small):
    from livespikesorter_amd.sorter import Sorter
    d, art_dir, rec = facade_inputs
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (livespikesorter_amd.build)")
    out = d / f"out_{mn}_{mx}_{tb}_{first}_{lag}_{small}"
    out.mkdir()
    print(f"case: facade min={mn} max={mx} tb={tb} first={first} lag={lag} small={small}", flush=True)
    r = subprocess.run([str(EXE), str(art_dir), str(d / "rec.i16"), str(rec.shape[1]), str(mn), str(mx), str(out),
                        str(tb), str(first), str(lag), str(small)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    got_batches, got_skips, got_spikes = int(f[1]), int(f[3]), int(f[5])
    nt = 61
    sched = _schedule(rec.shape[0], first, mn, mx, nt, tb, lag, small)
    assert got_batches == len(sched) and got_skips == sum(s for _, _, s in sched)
    assert int(f[7]) == sched[-1][0] + sched[-1][1]
    assert (int(f[9]), int(f[11]), int(f[13])) == (rec.shape[1], 0, rec.shape[1] - 1)   # channelMap.npy
    with Sorter.from_dir(art_dir, max_window=mx + 2 * 127, max_passes=100, max_spikes=1 << 16) as s:
        exp = _expected(s, rec, sched, nt, s.T)
    got = _read(out / "spikeOutput.txt")
    assert got_spikes == len(got) == len(exp) > 0
    np.testing.assert_array_equal(got, exp)
    if lag == 0 and tb > 0:
        # batches tile the stream: no gaps, no overlaps
        ends = [a + b for a, b, _ in sched]
        assert all(sched[i + 1][0] == ends[i] for i in range(len(sched) - 1))


def test_facade_labels_from_directory_centroids(synth1, tmp_path_factory):
    """With cluster_centroids.npy in the artifact directory the facade labels spikes by the nearest
    centroid (OnlineSpikesV2::loadKilosortClusteringData / closestCluster) and deduplicates per
    label with misc.txt's duplicate_spike_bins (OnlineSpikesV2.cu:456); the socket is handed channelMap.npy's channels, and a channel map of the wrong length
    is refused."""
    from livespikesorter_amd.export import write_oss_input
    from livespikesorter_amd.output import load_cluster_centroids
    from livespikesorter_amd.sorter import Sorter
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (livespikesorter_amd.build)")
    print("case: facade with directory centroids", flush=True)
    d = tmp_path_factory.mktemp("facade_c")
    rng = np.random.default_rng(21)
    xc, yc = np.asarray(synth1.art["xc"]), np.asarray(synth1.art["yc"])
    xy = np.stack([rng.uniform(xc.min(), xc.max(), 12), rng.uniform(yc.min(), yc.max(), 12)], 1)
    C = np.asarray(synth1.art["U"]).shape[2]
    cmap = np.arange(C)[::-1] + 100      # the channel list handed to the socket
    # a duplicate window other than the default: the facade takes it from misc.txt
    art_dir = write_oss_input(dict(synth1.art, dup_bins=40), d / "oss_input", cluster_xy=xy, channel_map=cmap)
    rec = np.ascontiguousarray(synth1.rec[:30000], np.int16)
    rec.tofile(d / "rec.i16")
    out = d / "out"
    out.mkdir()
    r = subprocess.run([str(EXE), str(art_dir), str(d / "rec.i16"), str(rec.shape[1]), "3000", "3000", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    assert (int(f[9]), int(f[11]), int(f[13])) == (C, int(cmap[0]), int(cmap[-1]))
    sched = _schedule(rec.shape[0], 0, 3000, 3000, 61, 100000, 0, 0)
    cents = load_cluster_centroids(art_dir)
    with Sorter.from_dir(art_dir, max_window=3000 + 2 * 127, max_passes=100, max_spikes=1 << 16) as s:
        assert s.params.dup_bins == 40
        exp = _expected(s, rec, sched, 61, len(xy), dt=40, centroids=cents)
    got = _read(out / "spikeOutput.txt")
    assert len(got) == len(exp) > 0
    np.testing.assert_array_equal(got, exp)
    assert got[:, 1].max() < len(xy)
    # a channel map of the wrong length is refused, as OnlineSpikesV2::loadChannelMap refuses it
    np.save(art_dir / "channelMap.npy", np.arange(C - 1, dtype=np.int32))
    r = subprocess.run([str(EXE), str(art_dir), str(d / "rec.i16"), str(rec.shape[1]), "3000", "3000", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "channelMap.npy" in r.stderr, (r.returncode, r.stderr)
