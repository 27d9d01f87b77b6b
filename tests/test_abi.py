"""C ABI checks that need no GPU: the library loads, exports every entry point include/lss.h
declares, and its host-only functions (defaults, batch padding, error reporting) behave."""
import ctypes as ct
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import ks_oracle as ko

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    hdr = (ROOT / "include" / "lss.h").read_text()
    return sorted(set(re.findall(r"\b(lss_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    from livespikesorter_amd import _lib
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.EXPORTS) == set(names)


def test_header_compiles_as_c_and_cpp(tmp_path):
    """include/lss.h is a plain C header (C99, no C++ types) that C++ can include as well."""
    import shutil
    import subprocess
    src = tmp_path / "t.c"
    src.write_text('#include "lss.h"\nint main(void) { return lss_abi_version() > 0 ? 0 : 1; }\n')
    for cc, std in (("gcc", "-std=c99"), ("g++", "-std=c++17")):
        exe = shutil.which(cc)
        if exe is None:
            pytest.skip(f"{cc} not found")
        r = subprocess.run([exe, std, "-Wall", "-Werror", "-fsyntax-only", "-x", "c" if cc == "gcc" else "c++",
                            "-I", str(ROOT / "include"), str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_pinned_host_buffers():
    from livespikesorter_amd import _lib
    L = _lib.lib()
    p = ct.c_void_p()
    if L.lss_host_alloc(0, ct.byref(p)) == 0:
        raise AssertionError("a zero-byte request must be rejected")
    assert L.lss_host_free(None) == 0


def test_default_params():
    from livespikesorter_amd.sorter import default_params
    p = default_params()
    assert (p.nt, p.nt0min, p.n_pcs, p.n_nearest, p.max_passes) == (61, 20, 6, 10, 50)
    assert p.th_learned == 8.0 and p.max_window == 60122 and p.dup_bins == 7


@pytest.mark.parametrize("n", [130000, 120000, 60000, 70])
def test_pad_batch_matches_oracle(n):
    from livespikesorter_amd.sorter import pad_batch
    rng = np.random.default_rng(n)
    rec = rng.integers(-3000, 3000, (n, 5)).astype(np.int16)
    NT, nt = 60000, 61
    for ib in range(-(-n // NT)):
        win, ws = pad_batch(rec, ib, NT, nt)
        ref = ko.padded_batch(rec, ib, NT, nt, torch.float64).numpy().T
        np.testing.assert_array_equal(win, ref.astype(np.int16))
        assert ws == ko.window_start(ib, NT, nt)


def test_errors_are_reported_not_thrown():
    from livespikesorter_amd import _lib
    L = _lib.lib()
    out = np.zeros((10, 2), np.int16)
    rec = np.zeros((5, 2), np.int16)
    ws = ct.c_int64()
    st = L.lss_pad_batch(rec.ctypes.data, 5, 2, 4, 1, 7, out.ctypes.data, ct.byref(ws))
    assert st == 1 and b"ibatch" in L.lss_last_error()
    assert L.lss_destroy(None) == 1
    assert L.lss_abi_version() == 2


def test_argument_checks_before_any_device_work():
    """Entry points taking device pointers or handles reject bad arguments with LSS_ERR_ARG and a
    message, before any HIP call (so these run without a GPU)."""
    from livespikesorter_amd import _lib
    L = _lib.lib()
    p = ct.c_void_p(16)                     # never dereferenced: the checks fail first
    # lss_assign_clusters: negative n, D out of range, no clusters, non-finite R, null labels
    cases = [(-1, 6, 4, 80.0, p), (10, 0, 4, 80.0, p), (10, 257, 4, 80.0, p), (10, 6, 0, 80.0, p),
             (10, 6, 4, float("nan"), p), (10, 6, 4, 80.0, None)]
    for n, D, K, R, lab in cases:
        st = L.lss_assign_clusters(p, n, D, p, K, p, p, p, p, R, lab, None)
        assert st == 1 and b"lss_assign_clusters" in L.lss_last_error(), (n, D, K, R)
    assert L.lss_cluster_centroids_pca(None, p) == 1
    assert L.lss_sort_windows(None, 0, None, None, None, 1, 1, None) == 1
    assert L.lss_sort_window(None, p, 100, 1, 0, None) == 1
    # streaming controls: null stream / output
    assert L.lss_stream_set_run_ahead(None, 1) == 1
    assert L.lss_stream_ran_ahead(None, None) == 1
    assert L.lss_stream_set_grouped(None, 1) == 1
    assert L.lss_stream_grouped(None, None) == 1
    assert L.lss_stream_profile(None, 1) == 1
    assert L.lss_stream_set_pipelined(None, 1) == 1
    assert L.lss_stream_pipelined(None, None) == 1
    assert L.lss_stream_last_profile(None, None, None, None, None) == 1
    assert L.lss_stream_set_refine(None, 10) == 1
    assert L.lss_stream_last_refine(None, 0, None) == 1


def _npy_bytes(header: str, data: bytes) -> bytes:
    h = header.encode()
    pad = 64 - (10 + len(h) + 1) % 64
    h = h + b" " * pad + b"\n"
    return b"\x93NUMPY\x01\x00" + len(h).to_bytes(2, "little") + h + data


MALFORMED = {
    "truncated data": _npy_bytes("{'descr': '<i4', 'fortran_order': False, 'shape': (64,), }", b"\0" * 40),
    "negative shape": _npy_bytes("{'descr': '<i4', 'fortran_order': False, 'shape': (-3,), }", b"\0" * 40),
    "huge shape": _npy_bytes("{'descr': '<f4', 'fortran_order': False, 'shape': (4611686018427387904, 8), }",
                             b"\0" * 40),
    "garbage shape": _npy_bytes("{'descr': '<f4', 'fortran_order': False, 'shape': (1x2,), }", b"\0" * 40),
    "bad dtype": _npy_bytes("{'descr': '<c8', 'fortran_order': False, 'shape': (2,), }", b"\0" * 16),
    "truncated header": b"\x93NUMPY\x01\x00\xff\x7f{'descr'",
    "not npy": b"hello world, not an array",
}


@pytest.mark.parametrize("case", sorted(MALFORMED))
def test_malformed_npy_is_an_io_error(tmp_path, case):
    """A malformed artifact file is reported as LSS_ERR_IO with the file's name (no exception
    crosses the ABI, nothing is allocated from an untrusted size), by the host-only loaders and by
    lss_create_from_dir before any device work."""
    from livespikesorter_amd import _lib
    from livespikesorter_amd.sorter import default_params
    L = _lib.lib()
    (tmp_path / "channelMap.npy").write_bytes(MALFORMED[case])
    (tmp_path / "cluster_centroids.npy").write_bytes(MALFORMED[case])
    n = ct.c_int32(-1)
    assert L.lss_load_channel_map(str(tmp_path).encode(), None, 0, ct.byref(n)) == 3
    assert b"channelMap.npy" in L.lss_last_error()
    assert L.lss_load_cluster_centroids(str(tmp_path).encode(), None, None, 0, ct.byref(n)) == 3
    assert b"cluster_centroids.npy" in L.lss_last_error()
    (tmp_path / "misc.txt").write_text("nt0min:20\nnumNearestChans:10\nTh_learned:8\nduplicate_spike_bins:7\n")
    (tmp_path / "Wall3.npy").write_bytes(MALFORMED[case])
    p = default_params()
    h = ct.c_void_p()
    assert L.lss_create_from_dir(str(tmp_path).encode(), ct.byref(p), ct.byref(h)) == 3
    assert b"Wall3.npy" in L.lss_last_error() and not h.value


def test_npy_loader_reads_dtypes_and_fortran_order(tmp_path):
    """The host loaders accept int32/int64/float arrays in C or Fortran order (numpy's view)."""
    from livespikesorter_amd import _lib
    L = _lib.lib()
    m = np.arange(12, dtype=np.int64)[::-1].copy()
    for arr in (m, m.astype(np.int32), m.astype(np.float64), np.asfortranarray(m.reshape(3, 4)),
                np.asfortranarray(m.reshape(3, 4).astype(np.int32))):
        np.save(tmp_path / "channelMap.npy", arr)
        out = np.zeros(12, np.int32)
        n = ct.c_int32()
        assert L.lss_load_channel_map(str(tmp_path).encode(), out.ctypes.data, 12, ct.byref(n)) == 0
        np.testing.assert_array_equal(out, np.asarray(arr).reshape(-1))
