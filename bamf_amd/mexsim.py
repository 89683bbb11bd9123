"""Calling the mex wrappers (bamf_amd/csrc/mex/*.cpp) without MATLAB.

Each wrapper is built by bamf_amd/build.py into bamf_amd/mex/<name>.so
together with the project's mx runtime (csrc/mex/mx/mx.cpp).  `call` turns
Python values into mxArrays the way MATLAB would hold them (column-major
numeric arrays of the numpy dtype's class, char row vectors, 1 x 1 structs
from dicts, N x 1 struct arrays from lists of dicts), runs the wrapper's
mexFunction and converts its outputs back.

    pgrid, pback, chain, prop, stat = call("cRJMCMC_gpu", 5, data, e_active, mcmc, xs, ys, ...)
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_int, c_size_t, c_void_p
from pathlib import Path

import numpy as np

from ._lib import BamfError, lib

MEX_DIR = Path(__file__).resolve().parent / "mex"

# mxClassID values of mex.h
_CLASS_OF = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.int8): 8, np.dtype(np.uint8): 9,
             np.dtype(np.int16): 10, np.dtype(np.uint16): 11, np.dtype(np.int32): 12, np.dtype(np.uint32): 13,
             np.dtype(np.int64): 14, np.dtype(np.uint64): 15, np.dtype(np.bool_): 3}
_DTYPE_OF = {v: k for k, v in _CLASS_OF.items()}
STRUCT, CHAR = 2, 4

_MAX_OUT = 16
_libs: dict[str, ctypes.CDLL] = {}


def _load(name: str) -> ctypes.CDLL:
    if name in _libs:
        return _libs[name]
    path = MEX_DIR / f"{name}.so"
    if not path.exists():
        raise BamfError(f"{path} is not built (python -m bamf_amd.build)")
    lib()  # libbamf_gpu.so first: the wrapper resolves its symbols there
    L = ctypes.CDLL(str(path))
    L.mxh_numeric.restype = c_void_p
    L.mxh_numeric.argtypes = [c_int, c_int, POINTER(c_size_t), c_void_p]
    L.mxh_char.restype = c_void_p
    L.mxh_char.argtypes = [c_char_p]
    L.mxh_struct.restype = c_void_p
    L.mxh_struct.argtypes = [c_size_t, c_int, POINTER(c_char_p)]
    L.mxh_set_field.restype = c_int
    L.mxh_set_field.argtypes = [c_void_p, c_size_t, c_int, c_void_p]
    L.mxh_call.restype = c_int
    L.mxh_call.argtypes = [c_int, POINTER(c_void_p), c_int, c_int, POINTER(c_void_p)]
    L.mxh_last_error.restype = c_char_p
    for f, r in (("mxh_class", c_int), ("mxh_ndims", c_int), ("mxh_dims", POINTER(c_size_t)),
                 ("mxh_data", c_void_p), ("mxh_nbytes", c_size_t), ("mxh_nfields", c_int)):
        getattr(L, f).restype = r
        getattr(L, f).argtypes = [c_void_p]
    L.mxh_field_name.restype = c_char_p
    L.mxh_field_name.argtypes = [c_void_p, c_int]
    L.mxh_field_at.restype = c_void_p
    L.mxh_field_at.argtypes = [c_void_p, c_size_t, c_int]
    L.mxh_destroy.restype = None
    L.mxh_destroy.argtypes = [c_void_p]
    _libs[name] = L
    return L


def _to_mx(L, v):
    if isinstance(v, dict):
        v = [v]
    if isinstance(v, list) and v and all(isinstance(e, dict) for e in v):
        names = list(v[0].keys())
        arr = (c_char_p * len(names))(*[n.encode() for n in names])
        s = L.mxh_struct(len(v), len(names), arr)
        for i, e in enumerate(v):
            for k, n in enumerate(names):
                if L.mxh_set_field(s, i, k, _to_mx(L, e[n])):
                    raise BamfError(L.mxh_last_error().decode())
        return s
    if isinstance(v, str):
        return L.mxh_char(v.encode())
    a = np.asarray(v, np.float64) if isinstance(v, (bool, int, float)) else np.asarray(v)
    if a.dtype not in _CLASS_OF:
        raise TypeError(f"no mx class for dtype {a.dtype}")
    shape = (1, 1) if a.ndim == 0 else ((1, a.shape[0]) if a.ndim == 1 else a.shape)
    dims = (c_size_t * len(shape))(*shape)
    buf = np.asfortranarray(a.reshape(shape))
    p = L.mxh_numeric(_CLASS_OF[a.dtype], len(shape), dims, buf.ctypes.data_as(c_void_p))
    if not p:
        raise BamfError(L.mxh_last_error().decode())
    return p


def _from_mx(L, p):
    cls = L.mxh_class(p)
    nd = L.mxh_ndims(p)
    dims = tuple(L.mxh_dims(p)[k] for k in range(nd))
    n = int(np.prod(dims))
    if cls == STRUCT:
        nf = L.mxh_nfields(p)
        names = [L.mxh_field_name(p, k).decode() for k in range(nf)]
        out = []
        for i in range(n):
            e = {}
            for k, name in enumerate(names):
                q = L.mxh_field_at(p, i, k)
                e[name] = None if not q else _from_mx(L, q)
            out.append(e)
        return out[0] if dims == (1, 1) else out
    if cls == CHAR:
        raw = ctypes.string_at(L.mxh_data(p), L.mxh_nbytes(p)) if n else b""
        return np.frombuffer(raw, np.uint16).astype(np.uint8).tobytes().decode()
    dt = _DTYPE_OF[cls]
    raw = ctypes.string_at(L.mxh_data(p), L.mxh_nbytes(p)) if n else b""
    return np.frombuffer(raw, dt).reshape(dims, order="F").copy()


def call(name: str, nlhs: int, *args):
    """Runs mex file `name` with MATLAB's calling convention; returns its nlhs
    outputs (a tuple; a single value when nlhs is 1).  A mexErrMsgIdAndTxt in
    the wrapper raises BamfError with its identifier and message."""
    L = _load(name)
    prhs = (c_void_p * max(len(args), 1))()
    made = []
    try:
        for k, a in enumerate(args):
            prhs[k] = _to_mx(L, a)
            made.append(prhs[k])
        nout = max(nlhs, 1)
        # room for every output a wrapper fills (cRJMCMC's fills all five whatever nlhs is,
        # as cRJMCMC.cpp:112-169 does); those past nlhs are dropped
        plhs = (c_void_p * max(nout, _MAX_OUT))()
        if L.mxh_call(nlhs, plhs, len(plhs), len(args), prhs):
            raise BamfError(f"{name}: {L.mxh_last_error().decode()}")
        try:
            outs = tuple(None if not plhs[k] else _from_mx(L, plhs[k]) for k in range(nout))
        finally:
            for k in range(len(plhs)):
                if plhs[k]:
                    L.mxh_destroy(plhs[k])
    finally:
        for p in made:
            L.mxh_destroy(p)
    return outs[0] if nlhs <= 1 else outs
