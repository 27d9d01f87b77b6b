"""ctypes binding of liblss.so (declarations in include/lss.h).

The library is loaded from this package directory; there is no fallback: if it is missing or
fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblss.so"

LSS_OK = 0
STATUS = {0: "LSS_OK", 1: "LSS_ERR_ARG", 2: "LSS_ERR_HIP", 3: "LSS_ERR_IO", 4: "LSS_ERR_CAPACITY",
          5: "LSS_ERR_STATE"}
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x3": 3, "fp16x3": 4}
PERSIST_MAX_WINDOW = 4096

EXPORTS = ["lss_last_error", "lss_abi_version", "lss_default_params", "lss_create",
           "lss_create_from_dir", "lss_destroy", "lss_get_params", "lss_set_stream",
           "lss_sort_window", "lss_pad_batch", "lss_stage_filter", "lss_stage_car",
           "lss_stage_highpass", "lss_stage_whiten", "lss_stage_scores", "lss_stage_pursuit",
           "lss_get_tables", "lss_last_timings", "lss_profile_update", "lss_last_update_stats",
           "lss_label_spikes", "lss_dedup_spikes", "lss_write_spikes_txt",
           "lss_stream_create", "lss_stream_destroy", "lss_stream_push", "lss_stream_step",
           "lss_stream_stats", "lss_stream_breakdown", "lss_autotune", "lss_set_tuning", "lss_get_tuning",
           "lss_set_persistent", "lss_get_persistent", "lss_autotune_persistent",
           "lss_residual_energy", "lss_residual_energy_grad", "lss_last_refine_timings",
           "lss_refine_templates", "lss_stream_set_refine", "lss_stream_set_run_ahead", "lss_stream_ran_ahead",
           "lss_stream_set_grouped", "lss_stream_grouped", "lss_stream_profile", "lss_stream_last_profile",
           "lss_stream_set_pipelined", "lss_stream_pipelined",
           "lss_stream_last_refine",
           "lss_stream_push_device",
           "lss_set_drift", "lss_covariance_add", "lss_whitening_from_covariance", "lss_whitening_local",
           "lss_drift_inverse", "lss_drift_matrix", "lss_extract_snippets", "lss_wpca", "lss_train_release",
           "lss_select_device", "lss_sort_windows", "lss_cluster_centroids_pca", "lss_assign_clusters", "lss_load_cluster_centroids", "lss_load_channel_map",
           "lss_get_load_info", "lss_set_tiles", "lss_get_tiles", "lss_autotune_tiles",
           "lss_last_update_times", "lss_set_update_columns", "lss_get_update_columns",
           "lss_extract_snippets_noalloc", "lss_host_alloc", "lss_host_free", "lss_last_persistent_stats"]


class Params(ct.Structure):
    _fields_ = [("n_chan", ct.c_int32), ("n_templates", ct.c_int32), ("n_pcs", ct.c_int32),
                ("nt", ct.c_int32), ("nt0min", ct.c_int32), ("n_nearest", ct.c_int32),
                ("max_window", ct.c_int32), ("max_passes", ct.c_int32), ("max_spikes", ct.c_int32),
                ("do_car", ct.c_int32), ("precision", ct.c_int32), ("drop_negative", ct.c_int32),
                ("th_learned", ct.c_double), ("fs", ct.c_float), ("dup_bins", ct.c_int32)]


class Artifacts(ct.Structure):
    _fields_ = [("U", ct.c_void_p), ("wPCA", ct.c_void_p), ("Wrot", ct.c_void_p),
                ("drift", ct.c_void_p), ("hp_filter", ct.c_void_p), ("hp_len", ct.c_int32),
                ("xc", ct.c_void_p), ("yc", ct.c_void_p), ("iCC", ct.c_void_p), ("iU", ct.c_void_p),
                ("Ucc", ct.c_void_p)]


class LoadInfo(ct.Structure):
    _fields_ = [("icc_given", ct.c_int32), ("iu_given", ct.c_int32), ("ucc_given", ct.c_int32),
                ("ctc_in_dir", ct.c_int32), ("ctc_max_abs_diff", ct.c_double), ("ctc_max_abs", ct.c_double)]


class Spikes(ct.Structure):
    _fields_ = [("capacity", ct.c_int64), ("n", ct.c_int64), ("passes", ct.c_int32),
                ("sample", ct.c_void_p), ("template_id", ct.c_void_p), ("amplitude", ct.c_void_p),
                ("features", ct.c_void_p), ("x", ct.c_void_p), ("y", ct.c_void_p),
                ("column", ct.c_void_p)]


class LssError(RuntimeError):
    pass


_lib = None


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LssError(f"{LIB_PATH} is not built: run `python -m livespikesorter_amd.build`")
        # torch first: its HIP runtime is then the one liblss binds to, so that torch tensors and
        # the library share one runtime (loading liblss first leaves torch a second one)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ct.CDLL(str(LIB_PATH))
        vp, i32, i64, f64 = ct.c_void_p, ct.c_int32, ct.c_int64, ct.c_double
        sig = {
            "lss_last_error": ([], ct.c_char_p),
            "lss_abi_version": ([], i32),
            "lss_default_params": ([ct.POINTER(Params)], i32),
            "lss_create": ([ct.POINTER(Params), ct.POINTER(Artifacts), ct.POINTER(vp)], i32),
            "lss_create_from_dir": ([ct.c_char_p, ct.POINTER(Params), ct.POINTER(vp)], i32),
            "lss_destroy": ([vp], i32),
            "lss_get_params": ([vp, ct.POINTER(Params)], i32),
            "lss_set_stream": ([vp, vp], i32),
            "lss_sort_window": ([vp, vp, i64, i32, i64, ct.POINTER(Spikes)], i32),
            "lss_pad_batch": ([vp, i64, i32, i32, i32, i64, vp, ct.POINTER(i64)], i32),
            "lss_stage_filter": ([vp, vp, i64, vp], i32),
            "lss_stage_car": ([vp, vp, i64, vp], i32),
            "lss_stage_highpass": ([vp, vp, i64, vp], i32),
            "lss_stage_whiten": ([vp, vp, i64, vp], i32),
            "lss_stage_scores": ([vp, vp, i64, vp], i32),
            "lss_stage_pursuit": ([vp, vp, vp, i64, i64, ct.POINTER(Spikes)], i32),
            "lss_get_tables": ([vp, vp, vp, vp, vp], i32),
            "lss_last_timings": ([vp, vp, i32], i32),
            "lss_profile_update": ([vp, i32], i32),
            "lss_last_update_stats": ([vp, ct.POINTER(ct.c_float), ct.POINTER(i32)], i32),
            "lss_label_spikes": ([vp, vp, i64, vp, vp, i32, vp], i32),
            "lss_dedup_spikes": ([vp, vp, vp, i64, i32, vp, i32, ct.POINTER(i64)], i32),
            "lss_write_spikes_txt": ([ct.c_char_p, vp, vp, vp, vp, i64, i64, i32], i32),
            "lss_stream_create": ([ct.POINTER(vp), i32, i32, i32, ct.POINTER(vp)], i32),
            "lss_stream_destroy": ([vp], i32),
            "lss_stream_push": ([vp, i32, vp, i64], i32),
            "lss_stream_push_device": ([vp, i32, vp, i64], i32),
            "lss_stream_step": ([vp, ct.POINTER(Spikes), ct.POINTER(i32)], i32),
            "lss_stream_stats": ([vp, ct.POINTER(ct.c_float), ct.POINTER(i64)], i32),
            "lss_stream_breakdown": ([vp, vp], i32),
            "lss_autotune": ([vp, vp, i64, i32, vp, ct.POINTER(i32), ct.POINTER(i32)], i32),
            "lss_set_tuning": ([vp, i32, i32], i32),
            "lss_get_tuning": ([vp, ct.POINTER(i32), ct.POINTER(i32)], i32),
            "lss_set_persistent": ([vp, i32, i32], i32),
            "lss_get_persistent": ([vp, i64, ct.POINTER(i32), ct.POINTER(i32), ct.POINTER(i32)], i32),
            "lss_autotune_persistent": ([vp, vp, i64, i32, i32, vp, ct.POINTER(i32), ct.POINTER(i32)], i32),
            "lss_residual_energy": ([vp, vp, i64, vp, vp, vp, i64, vp, vp, ct.POINTER(ct.c_double)], i32),
            "lss_residual_energy_grad": ([vp, vp, i64, vp, vp, vp, i64, vp, vp, ct.POINTER(ct.c_double), vp, vp], i32),
            "lss_last_refine_timings": ([vp, vp], i32),
            "lss_refine_templates": ([vp, vp], i32),
            "lss_stream_set_refine": ([vp, i32], i32),
            "lss_stream_set_run_ahead": ([vp, i32], i32),
            "lss_stream_ran_ahead": ([vp, ct.POINTER(i64)], i32),
    "lss_stream_set_grouped": ([vp, i32], i32),
    "lss_stream_grouped": ([vp, ct.POINTER(i32)], i32),
    "lss_stream_profile": ([vp, i32], i32),
    "lss_stream_set_pipelined": ([vp, i32], i32),
    "lss_stream_pipelined": ([vp, ct.POINTER(i64)], i32),
    "lss_stream_last_profile": ([vp, ct.POINTER(ct.c_double), ct.POINTER(i32), ct.POINTER(i32), ct.POINTER(i32)], i32),
            "lss_stream_last_refine": ([vp, i32, vp], i32),
            "lss_set_drift": ([vp, vp], i32),
            "lss_covariance_add": ([vp, i64, i32, i64, vp, i64, vp], i32),
            "lss_whitening_from_covariance": ([vp, i64, i32, f64, vp, i64, vp], i32),
            "lss_whitening_local": ([vp, i64, i32, vp, vp, i32, f64, vp, i64, vp], i32),
            "lss_drift_inverse": ([vp, vp, i32, f64, f64, vp, i64, ct.POINTER(i32), vp], i32),
            "lss_drift_matrix": ([vp, vp, i32, vp, vp, i32, f64, vp, i64, vp, i64, vp], i32),
            "lss_extract_snippets": ([vp, i64, i32, i64, i32, i32, ct.c_float, vp, vp, vp, i64, vp, vp,
                                      ct.POINTER(i64), vp], i32),
            "lss_extract_snippets_noalloc": ([vp, i64, i32, i64, i32, i32, ct.c_float, vp, vp, vp, i64, vp, vp,
                                              ct.POINTER(i64), vp], i32),
            "lss_wpca": ([vp, i64, i32, i32, vp, vp, vp], i32),
            "lss_train_release": ([], i32),
            "lss_select_device": ([i32], i32),
            "lss_host_alloc": ([ct.c_size_t, ct.POINTER(vp)], i32),
            "lss_last_persistent_stats": ([vp, ct.POINTER(ct.c_float), ct.POINTER(i32), ct.POINTER(i32),
                                           ct.POINTER(i32)], i32),
            "lss_host_free": ([vp], i32),
            "lss_sort_windows": ([ct.POINTER(vp), i32, ct.POINTER(vp), ct.POINTER(i64), ct.POINTER(i64), i32, i32,
                                  ct.POINTER(Spikes)], i32),
            "lss_cluster_centroids_pca": ([vp, vp], i32),
            "lss_assign_clusters": ([vp, i64, i32, vp, i32, vp, vp, vp, vp, f64, vp, vp], i32),
            "lss_load_cluster_centroids": ([ct.c_char_p, vp, vp, i32, ct.POINTER(i32)], i32),
            "lss_load_channel_map": ([ct.c_char_p, vp, i32, ct.POINTER(i32)], i32),
            "lss_get_load_info": ([vp, ct.POINTER(LoadInfo)], i32),
            "lss_set_tiles": ([vp, i32, i32, i32], i32),
            "lss_get_tiles": ([vp, vp], i32),
            "lss_autotune_tiles": ([vp, vp, i64, i32, vp, vp], i32),
            "lss_last_update_times": ([vp, vp, vp, vp, i32, ct.POINTER(i32)], i32),
            "lss_set_update_columns": ([vp, i32], i32),
            "lss_get_update_columns": ([vp, ct.POINTER(i32)], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != LSS_OK:
        msg = lib().lss_last_error().decode(errors="replace")
        raise LssError(f"{what}: {STATUS.get(status, status)}: {msg}")
