"""ctypes wrapper of the C oracle (oracle/eval_oracle.c) and of the reference's
own CPU units (oracle/_ref/libsrref.so, present only where the reference
could be compiled).

ORACLE — test infrastructure.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may use it, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libsrref.so")

_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        l = ctypes.CDLL(LIB_PATH)
        P, I64 = c_void_p, c_int64
        l.oracle_validate.restype = c_int32
        l.oracle_validate.argtypes = [P, P, I64, I64]
        for s in ("f32", "f64"):
            getattr(l, f"oracle_forward_{s}").argtypes = [P, I64, I64, P, P, P, I64, I64, P]
            getattr(l, f"oracle_backward_{s}").restype = c_int32
            getattr(l, f"oracle_backward_{s}").argtypes = [P, P, P, P, I64, I64, I64, I64, P, P]
            getattr(l, f"oracle_multi_forward_{s}").argtypes = [P, I64, I64, P, P, P, I64, I64, I64, P]
            getattr(l, f"oracle_multi_backward_{s}").restype = c_int32
            getattr(l, f"oracle_multi_backward_{s}").argtypes = [P, P, P, P, I64, I64, I64, I64, I64, P, P]
            for v in ("", "acc64_"):
                getattr(l, f"oracle_fused_root_{v}{s}").argtypes = [P, I64, I64, P, P, P, I64, I64, P, c_int, P, P,
                                                                    P, P, P]
        _lib = l
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        l = ctypes.CDLL(REF_PATH)
        P, I64 = c_void_p, c_int64
        l.ref_validate.restype = c_int32
        l.ref_validate.argtypes = [P, P, I64, I64]
        for s in ("f32", "f64"):
            getattr(l, f"ref_forward_{s}").argtypes = [P, I64, I64, P, P, P, I64, I64, P]
            getattr(l, f"ref_backward_{s}").restype = c_int32
            getattr(l, f"ref_backward_{s}").argtypes = [P, P, P, P, I64, I64, I64, I64, P, P]
        l.ref_multi_forward_f32.argtypes = [P, I64, I64, P, P, P, I64, I64, I64, P]
        l.ref_multi_backward_f32.restype = c_int32
        l.ref_multi_backward_f32.argtypes = [P, P, P, P, I64, I64, I64, I64, I64, P, P]
        l.ref_sample_infix.argtypes = [P, I64, P, P, I64, P, I64, P, I64, I64, I64, P, P, P, I64]
        l.ref_infix_to_postfix.argtypes = [P, I64, P, P, P, I64, I64, I64, I64]
        l.ref_infix_to_prefix.argtypes = [P, I64, P, P, P, I64, I64, I64, I64]
        l.ref_postfix_to_infix.argtypes = [P, P, P, I64, I64, I64, I64, I64]
        l.ref_prefix_to_infix.argtypes = [P, P, P, I64, I64, I64, I64, I64]
        for n in ("ref_prefix_to_postfix", "ref_postfix_to_prefix", "ref_get_prefix_parent",
                  "ref_get_postfix_children"):
            getattr(l, n).argtypes = [P, P, P, I64, I64]
        _ref = l
    return _ref


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _dt(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


# ----------------------------------------------------------------- C oracle
def validate(ops_mb: np.ndarray, ch_mb: np.ndarray, use_ref: bool = False) -> int:
    M, B = ops_mb.shape
    ops_mb = np.ascontiguousarray(ops_mb, np.int64)
    ch_mb = np.ascontiguousarray(ch_mb, np.int64)
    l = ref() if use_ref else lib()
    fn = l.ref_validate if use_ref else l.oracle_validate
    return int(fn(_p(ops_mb), _p(ch_mb), M, B))


def forward(X, ops_mb, ch_mb, C_mb, use_ref: bool = False) -> np.ndarray:
    """Full (M, N, B) cache."""
    X = np.ascontiguousarray(X)
    dt = X.dtype
    ops_mb = np.ascontiguousarray(ops_mb, np.int64)
    ch_mb = np.ascontiguousarray(ch_mb, np.int64)
    C_mb = np.ascontiguousarray(C_mb, dt)
    M, B = ops_mb.shape
    N, nx = X.shape
    cache = np.zeros((M, N, B), dt)
    fn = getattr(ref(), f"ref_forward_{_dt(dt)}") if use_ref else getattr(lib(), f"oracle_forward_{_dt(dt)}")
    fn(_p(X), N, nx, _p(ops_mb), _p(ch_mb), _p(C_mb), M, B, _p(cache))
    return cache


def backward(grad_out, cache, ops_mb, ch_mb, n_x: int, use_ref: bool = False, return_adjoints: bool = False):
    """Returns (grad_C (M,B), grad_X (N,n_x), error code) and, with
    return_adjoints, the accumulated (M, N, B) adjoint tape as a 4th item."""
    dt = cache.dtype
    g = np.array(grad_out, dtype=dt, copy=True, order="C")
    cache = np.ascontiguousarray(cache)
    ops_mb = np.ascontiguousarray(ops_mb, np.int64)
    ch_mb = np.ascontiguousarray(ch_mb, np.int64)
    M, N, B = cache.shape
    gC = np.zeros((M, B), dt)
    gX = np.zeros((N, n_x), dt)
    fn = getattr(ref(), f"ref_backward_{_dt(dt)}") if use_ref else getattr(lib(), f"oracle_backward_{_dt(dt)}")
    e = fn(_p(g), _p(cache), _p(ops_mb), _p(ch_mb), M, B, N, n_x, _p(gC), _p(gX))
    if return_adjoints:
        return gC, gX, int(e), g
    return gC, gX, int(e)


def multi_forward(X, ops_mb, ch_mb, C_mbk, use_ref: bool = False) -> np.ndarray:
    X = np.ascontiguousarray(X)
    dt = X.dtype
    ops_mb = np.ascontiguousarray(ops_mb, np.int64)
    ch_mb = np.ascontiguousarray(ch_mb, np.int64)
    C_mbk = np.ascontiguousarray(C_mbk, dt)
    M, B, K = C_mbk.shape
    N, nx = X.shape
    cache = np.zeros((M, N, B, K), dt)
    fn = ref().ref_multi_forward_f32 if use_ref else getattr(lib(), f"oracle_multi_forward_{_dt(dt)}")
    fn(_p(X), N, nx, _p(ops_mb), _p(ch_mb), _p(C_mbk), M, B, K, _p(cache))
    return cache


def multi_backward(grad_out, cache, ops_mb, ch_mb, n_x: int, use_ref: bool = False):
    dt = cache.dtype
    g = np.array(grad_out, dtype=dt, copy=True, order="C")
    cache = np.ascontiguousarray(cache)
    ops_mb = np.ascontiguousarray(ops_mb, np.int64)
    ch_mb = np.ascontiguousarray(ch_mb, np.int64)
    M, N, B, K = cache.shape
    gC = np.zeros((M, B, K), dt)
    gX = np.zeros((N, n_x), dt)
    fn = ref().ref_multi_backward_f32 if use_ref else getattr(lib(), f"oracle_multi_backward_{_dt(dt)}")
    e = fn(_p(g), _p(cache), _p(ops_mb), _p(ch_mb), M, B, N, n_x, K, _p(gC), _p(gX))
    return gC, gX, int(e)


def fused_root(X, ops_bm, ch_bm, C_bm, dy, mode: int, want_gx: bool = True, acc64: bool = False):
    """Root-only fused forward+backward over batch-first inputs (acc64:
    grad_C summed in double, as the device sums it, then rounded).

    Returns dict(y (B,N), loss (B,) float64, grad_C (B,M), grad_X (N,n_x), err (B,) int32).
    """
    X = np.ascontiguousarray(X)
    dt = X.dtype
    ops_bm = np.ascontiguousarray(ops_bm, np.int64)
    ch_bm = np.ascontiguousarray(ch_bm, np.int64)
    C_bm = np.ascontiguousarray(C_bm, dt)
    dy = np.ascontiguousarray(dy, dt)
    B, M = ops_bm.shape
    N, nx = X.shape
    y = np.zeros((B, N), dt)
    loss = np.zeros(B, np.float64)
    gC = np.zeros((B, M), dt)
    gX = np.zeros((N, nx), dt)
    err = np.zeros(B, np.int32)
    getattr(lib(), f"oracle_fused_root_{'acc64_' if acc64 else ''}{_dt(dt)}")(_p(X), N, nx, _p(ops_bm), _p(ch_bm), _p(C_bm), M, B, _p(dy),
                                                    int(mode), _p(y), _p(loss), _p(gC),
                                                    _p(gX) if want_gx else None, _p(err))
    return dict(y=y, loss=loss, grad_C=gC, grad_X=gX, err=err)
