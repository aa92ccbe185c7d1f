"""Loader for the native libraries of this package.

``libsr_hip.so`` holds the gfx950 kernels behind the C-ABI declared in
``include/sr_eval.h``; ``libsr_host.so`` holds the host-side PCFG engine
declared in ``include/sr_pcfg.h``.  Both are built in-tree by
``__graft_entry__.build()`` (or ``python -m cudasymbolicregression_amd.build``).

There is no fallback: if a library is missing, importing the module that needs
it raises immediately, and every GPU entry point refuses tensors that are not
on a ROCm device.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_float, c_int, c_int64, c_uint64, c_void_p, POINTER

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libsr_hip.so")
HOST_LIB_PATH = os.path.join(_HERE, "libsr_host.so")

_hip = None
_host = None


class NativeLibraryMissing(ImportError):
    pass


def _load(path: str, what: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{what} not found at {path}. Build it first with "
            "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root "
            "(or `python -m cudasymbolicregression_amd.build`)."
        )
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


MISSING_SYMBOLS: list = []


def _declare(lib: ctypes.CDLL, name: str, restype, argtypes) -> None:
    try:
        fn = getattr(lib, name)
    except AttributeError:
        # Reported by tests/test_abi.py; calling it raises AttributeError.
        MISSING_SYMBOLS.append(name)
        return
    fn.restype = restype
    fn.argtypes = argtypes


def hip() -> ctypes.CDLL:
    """The gfx950 kernel library (loaded once)."""
    global _hip
    if _hip is None:
        lib = _load(HIP_LIB_PATH, "The gfx950 kernel library libsr_hip.so")
        P, I64, I = c_void_p, c_int64, c_int
        _declare(lib, "sr_validate_mb", I, [P, P, I64, I64, P, P])
        for sfx in ("f32", "f64"):
            _declare(lib, f"sr_forward_mb_{sfx}", I, [P, I64, I64, P, P, P, I64, I64, P, P])
            _declare(lib, f"sr_forward_mb_ws_{sfx}", I, [P, I64, I64, P, P, P, I64, I64, P, P, I64, P])
            _declare(lib, f"sr_multi_forward_mb_ws_{sfx}", I, [P, I64, I64, P, P, P, I64, I64, I64, P, P, I64, P])
            _declare(lib, f"sr_backward_mb_rows_{sfx}", I, [P, P, P, P, I64, I64, I64, I64, P, P, P, I64, P, I64, P, P])
            _declare(lib, f"sr_multi_backward_mb_rows_{sfx}", I,
                     [P, P, P, P, I64, I64, I64, I64, I64, P, P, P, I64, P, I64, P, P])
            _declare(lib, f"sr_backward_mb_{sfx}", I, [P, P, P, P, I64, I64, I64, I64, P, P, P, I64, P, P])
            _declare(lib, f"sr_backward_mb_rows_recompute_{sfx}", I,
                     [P, P, P, P, P, P, I64, I64, I64, I64, P, P, P, I64, P, I64, P, P])
            _declare(lib, f"sr_multi_backward_mb_rows_recompute_{sfx}", I,
                     [P, P, P, P, P, P, I64, I64, I64, I64, I64, P, P, P, I64, P, I64, P, P])
            _declare(lib, f"sr_fused_root_{sfx}", I,
                     [P, I64, I64, P, P, P, I64, I64, P, I, I, P, P, P, P, P, P, P, I64, P])
            _declare(lib, f"sr_fused_root_ragged_{sfx}", I,
                     [P, I64, I64, P, P, P, P, P, P, I64, I64, P, I, I, P, P, P, P, P, P, I64, P])
            _declare(lib, f"sr_multi_forward_mb_{sfx}", I, [P, I64, I64, P, P, P, I64, I64, I64, P, P])
            _declare(lib, f"sr_multi_backward_mb_{sfx}", I,
                     [P, P, P, P, I64, I64, I64, I64, I64, P, P, P, I64, P, P])
        _declare(lib, "sr_forward_mb_workspace_bytes", I64, [I64, I64, I64, I64, I])
        _declare(lib, "sr_backward_mb_rows_workspace_bytes", I64, [I64, I64, I64, I64, I])
        _declare(lib, "sr_backward_mb_workspace_bytes", I64, [I64, I64, I64, I])
        _declare(lib, "sr_backward_mb_reads_only", I, [I64, I])
        _declare(lib, "sr_multi_backward_mb_workspace_bytes", I64, [I64, I64, I64, I64, I])
        _declare(lib, "sr_fused_workspace_bytes", I64, [I64, I64, I64, I64, I, I])
        _declare(lib, "sr_fused_num_variants", I, [I, I])
        _declare(lib, "sr_fused_variant_info", I, [I, I, I, POINTER(c_int), POINTER(c_int)])
        for sfx in ("f32", "f64"):
            _declare(lib, f"sr_fused_root_opts_{sfx}", I,
                     [P, I64, I64, P, P, P, I64, I64, P, I, I, P, P, P, P, P, P, P, P, I64, P])
        _declare(lib, "sr_fused_noalloc_max_bytes", I64, [])
        _declare(lib, "sr_pcfg_sample_device", I, [P, P, P, P, I64, I64, I64, I64, P, P, c_uint64, I64, P])
        _declare(lib, "sr_pcfg_infix_to_postfix_device", I, [P, I64, P, P, P, I64, I64, I64, I64, P])
        _declare(lib, "sr_pcfg_get_postfix_children_device", I, [P, P, P, I64, I64, P])
        _declare(lib, "sr_pcfg_postfix_to_infix_device", I, [P, P, P, I64, I64, I64, I64, I64, P])
        _declare(lib, "sr_probe_valu_ops_per_thread", I, [I])
        _declare(lib, "sr_probe_valu_rate", I, [I, I64, I, P, P])
        _declare(lib, "sr_fused_root_noalloc_f32", I,
                 [P, I64, I64, P, P, P, I64, I64, P, I, I, P, P, P, P, P, P, P])
        _declare(lib, "sr_lbfgs_fit", I,
                 [P, P, I64, I64, P, P, P, I64, I64, P, P, P, P, P, P, P, P, P, P, I, I, c_double, I, I, c_double,
                  P, P, I64, P])
        _declare(lib, "sr_lbfgs_fit_multi", I,
                 [P, P, I64, I64, P, P, P, I64, I64, I64, P, P, P, P, P, P, P, P, P, P, I, I, c_double, I, I,
                  c_double, P, P, I64, P])
        _declare(lib, "sr_lbfgs_workspace_bytes", I64, [I64, I64])
        _declare(lib, "sr_fit_num_buckets", I, [])
        _declare(lib, "sr_fit_num_variants", I, [I])
        _declare(lib, "sr_fit_default_variant", I, [I])
        _declare(lib, "sr_fit_variant_info", I, [I, I, POINTER(c_int), POINTER(c_int)])
        F = c_float
        _declare(lib, "sr_mlp_gemm_nt", I, [P, P, P, I64, I64, I64, I64, I64, I, P, P, P, P])
        _declare(lib, "sr_mlp_onehot_gelu", I, [P, P, P, P, I64, I64, I64, I, I, I64, P, P, P, P])
        _declare(lib, "sr_mlp_noop_suffix", I, [P, I, I, I64, I64, P, P])
        _declare(lib, "sr_mlp_transpose", I, [P, P, I64, I64, P])
        _declare(lib, "sr_mlp_gelu_bwd", I, [P, P, I64, P, P, P])
        _declare(lib, "sr_mlp_col_sum_workspace_bytes", I64, [I64, I64])
        _declare(lib, "sr_mlp_col_sum", I, [P, I64, I64, P, P, I64, P])
        _declare(lib, "sr_mlp_adam", I, [P, P, P, P, I64, F, F, F, F, I64, F, P, P])
        _declare(lib, "sr_mlp_to_bf16", I, [P, I64, P, P])
        _declare(lib, "sr_mlp_features", I, [P, I64, I64, I64, I, I, P, I64, I, I, I, P, P])
        _declare(lib, "sr_mlp_mse_head", I, [P, P, P, I64, I64, I64, F, P, P, P, P])
        _declare(lib, "sr_mlp_predict", I, [P, P, I64, I64, P, P])
        _declare(lib, "sr_mlp_select_start", I, [P, P, P, P, I64, I64, P, P, P])
        _hip = lib
    return _hip


def host() -> ctypes.CDLL:
    """The host-side PCFG engine (loaded once)."""
    global _host
    if _host is None:
        lib = _load(HOST_LIB_PATH, "The host PCFG library libsr_host.so")
        P, I64, I = c_void_p, c_int64, c_int
        _declare(lib, "sr_pcfg_sort_rules", None, [P, I64, P])
        _declare(lib, "sr_pcfg_sample", None, [P, P, P, P, P, I64, I64, I64, P, P, P, I64])
        _declare(lib, "sr_pcfg_sample_philox", None, [P, P, P, P, I64, I64, I64, P, P, c_uint64, I64])
        _declare(lib, "sr_philox4x32_10", None, [P, P, P])
        _declare(lib, "sr_pcfg_infix_to_postfix", None, [P, P, P, P, I64, I64, I64, I64])
        _declare(lib, "sr_pcfg_infix_to_prefix", None, [P, P, P, P, I64, I64, I64, I64])
        _declare(lib, "sr_pcfg_postfix_to_infix", None, [P, P, P, I64, I64, I64, I64, I64])
        _declare(lib, "sr_pcfg_prefix_to_infix", None, [P, P, P, I64, I64, I64, I64, I64])
        _declare(lib, "sr_pcfg_prefix_to_postfix", None, [P, P, P, I64, I64])
        _declare(lib, "sr_pcfg_postfix_to_prefix", None, [P, P, P, I64, I64])
        _declare(lib, "sr_pcfg_get_prefix_parent", None, [P, P, P, I64, I64])
        _declare(lib, "sr_pcfg_get_postfix_children", None, [P, P, P, I64, I64])
        _host = lib
    return _host


def ptr(t) -> int:
    """Raw data pointer of a tensor (None -> NULL)."""
    return 0 if t is None else int(t.data_ptr())


def stream_of(device) -> int:
    """Raw hipStream_t of torch's current stream on `device`."""
    import torch

    return int(torch.cuda.current_stream(device).cuda_stream)


def check_status(code: int, what: str) -> None:
    if code != 0:
        if code == -1:
            raise RuntimeError(f"{what}: invalid arguments (sizes out of the supported range)")
        if code == -2:
            raise RuntimeError(f"{what}: workspace too small")
        if code == -3:
            raise RuntimeError(f"{what}: the launch would need more LDS than one workgroup may use")
        raise RuntimeError(f"{what}: HIP error {code}")
