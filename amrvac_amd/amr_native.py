"""ctypes binding of libamrvac_amr.so, the native AMR tree routines
(include/amrvac_amr.h): the neighbour tables of the leaves
(build_connectivity, src/mod_functions_connectivity.t:55-431) and the 2:1
balance of the regrid flags (proper_nesting,
src/amr/mod_coarsen_refine.t:186-319).  The library is compiled from
csrc/amr_tree.cpp with the host C++ compiler when it is missing."""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libamrvac_amr.so")

_lib = None
_lock = threading.Lock()

_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                from .build import build_host
                build_host(LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            L.amrvac_amr_connectivity.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip, _ip, _ip, _ip]
            L.amrvac_amr_connectivity.restype = ctypes.c_int
            L.amrvac_amr_proper_nesting.argtypes = [ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip, _ip, ctypes.c_int,
                                                   _bp, _bp]
            L.amrvac_amr_proper_nesting.restype = ctypes.c_int
            L.amrvac_amr_last_error.argtypes = []
            L.amrvac_amr_last_error.restype = ctypes.c_char_p
            _lib = L
    return _lib


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().amrvac_amr_last_error().decode()
        if rc == 2:
            raise RuntimeError(msg)
        raise ValueError(msg)


def connectivity(ndim: int, ng1, periodic, level: np.ndarray, ig: np.ndarray):
    """(ntype [nb, 3^nd], nbr [nb, 3^nd], nchild [nb, 3^nd, 2^nd]) of the
    leaves (level [nb], ig [nb, nd], block order)."""
    nb = len(level)
    nn, nc = 3 ** ndim, 2 ** ndim
    ntype = np.zeros((nb, nn), dtype=np.int32)
    nbr = np.zeros((nb, nn), dtype=np.int32)
    nchild = np.zeros((nb, nn, nc), dtype=np.int32)
    _check(lib().amrvac_amr_connectivity(ndim, _i32(ng1), _i32(periodic), nb, _i32(level), _i32(ig).reshape(-1),
                                         ntype.reshape(-1), nbr.reshape(-1), nchild.reshape(-1)))
    return ntype, nbr, nchild


def proper_nesting(ndim: int, ng1, periodic, level: np.ndarray, ig: np.ndarray, order: np.ndarray,
                   refine_max_level: int, refine: np.ndarray, coarsen: np.ndarray) -> None:
    """Balance the refine / coarsen flags (bool [nb], updated in place)."""
    r = np.ascontiguousarray(refine, dtype=np.uint8)
    c = np.ascontiguousarray(coarsen, dtype=np.uint8)
    _check(lib().amrvac_amr_proper_nesting(ndim, _i32(ng1), _i32(periodic), len(level), _i32(level),
                                           _i32(ig).reshape(-1), _i32(order), int(refine_max_level), r, c))
    refine[:] = r.astype(bool)
    coarsen[:] = c.astype(bool)
