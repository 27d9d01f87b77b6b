"""lss_extract_snippets_noalloc (scratch in a fixed device array, no device allocation) against
lss_extract_snippets (scratch from the growing per-device arena), both through the C ABI: clip
values, channels, times and counts bit-identical on consecutive calls of decreasing size (so the
second call runs over the first call's leftover scratch), truncation by max_clips, a whitened
window of the 384-channel probe at the largest size the fixed array takes, and a too-large window
rejected on the host before any GPU work."""
import numpy as np
import pytest
import torch

from oracle import train_oracle as to

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _planted_wc(rng, C, W, n):
    X = rng.standard_normal((C, W)).astype(np.float32) * 1.5
    for _ in range(n):
        c, t = rng.integers(0, C), rng.integers(0, W)
        X[c, t] = rng.choice([-1, 1]) * rng.uniform(6.5, 20)
    return torch.from_numpy(np.ascontiguousarray(X.T)).to(DEV), X


def _same(a, b):
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy())
    assert a[3] == b[3]


def test_noalloc_snippets_equal_arena_snippets():
    from livespikesorter_amd import train
    rng = np.random.default_rng(31)
    # largest first, then two smaller windows: each call runs over the previous call's scratch
    for C, W, n, kw in [(384, 86000, 20000, {}), (96, 20122, 3000, {}), (37, 4001, 300, {}),
                        (96, 20122, 3000, dict(nt=5, twav_min=2, loc_range=(1, 2), long_range=(2, 4)))]:
        print(f"[lss] noalloc snippets C={C} W={W} {kw}", flush=True)
        x_wc, X = _planted_wc(rng, C, W, n)
        ref = train.extract_snippets(x_wc, **kw)
        got = train.extract_snippets(x_wc, no_alloc=True, **kw)
        assert ref[3] > 0
        _same(got, ref)
        # truncation keeps the first rows and still reports the total
        g5 = train.extract_snippets(x_wc, max_clips=5, no_alloc=True, **kw)
        assert g5[3] == ref[3] and np.array_equal(g5[0].cpu().numpy(), ref[0][:5].cpu().numpy())
        if C < 384:
            rc, c, t = to.extract_snippets(X, **kw)
            assert np.array_equal(got[1].cpu().numpy(), c) and np.array_equal(got[2].cpu().numpy(), t)
            assert np.array_equal(got[0].cpu().numpy(), rc)


def test_noalloc_snippets_whitened_window():
    from livespikesorter_amd import synth, train
    from livespikesorter_amd.sorter import Sorter, pad_batch
    print("[lss] noalloc snippets whitened 384-ch window", flush=True)
    p = synth.generate(synth.SynthConfig(n_chan=384, n_units=100, n_samples=60000, NT=60000, seed=32))
    with Sorter(p.art) as s:
        win, _ = pad_batch(p.rec, 0, 60000, p.art["nt"])
        x_wc = s.stage_whiten(s.stage_filter(torch.from_numpy(win).to(DEV)))
        ref = train.extract_snippets(x_wc)
        got = train.extract_snippets(x_wc, no_alloc=True)
        small = train.extract_snippets(x_wc[:3000].contiguous(), no_alloc=True)
        small_ref = train.extract_snippets(x_wc[:3000].contiguous())
    assert ref[3] > 10
    _same(got, ref)
    _same(small, small_ref)


def test_noalloc_snippets_rejects_windows_beyond_the_fixed_scratch():
    from livespikesorter_amd import train
    print("[lss] noalloc snippets size limit", flush=True)
    x_wc = torch.zeros((88000, 384), dtype=torch.float32, device=DEV)
    with pytest.raises(RuntimeError, match="scratch"):
        train.extract_snippets(x_wc, no_alloc=True)
    # the arena entry point takes it, and the fixed one still works afterwards
    assert train.extract_snippets(x_wc)[3] == 0
    y = torch.zeros((1000, 8), dtype=torch.float32, device=DEV)
    y[500, 3] = 9.0
    got = train.extract_snippets(y, no_alloc=True)
    assert got[3] == 1 and int(got[1][0]) == 3 and int(got[2][0]) == 500
