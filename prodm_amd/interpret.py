"""Retrieval planning: ProDM's size interpreters and error estimators, run by
the host C++ planners in libprodm_mdr.so (include/prodm/interpreters.hpp).

    sizes  : per level, the byte size of each plane
    errors : per level, the error table (entry k = error after k planes)
    index  : per level, planes already held (updated in place)

Returns the bytes to read per level (GreedyBasedSizeInterpreter.hpp:23-250,
BasicSizeInterpreter.hpp:11-112 in ProDM).
"""
import ctypes as C

import numpy as np

from . import capi

GREEDY, SIGN_EXCLUDE_GREEDY, NEGABINARY_GREEDY, INORDER, ROUND_ROBIN = range(5)
MAX_OB, MAX_HB, SNORM, L2_HB = range(4)


def retrieve_sizes(planner, estimator, sizes, errors, tolerance, index, dtype=np.float32, ndims=3,
                   target_level=None, s=0.0):
    L = len(sizes)
    counts = np.array([len(x) for x in sizes], np.int32)
    for l in range(L):
        if len(errors[l]) != counts[l] + 1:
            raise ValueError(f"level {l}: need {counts[l] + 1} error entries")
    fs = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in sizes]) if L else np.zeros(1, np.uint32))
    fe = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in errors]))
    idx = np.asarray(index, np.uint8).copy()
    out = np.zeros(L, np.uint32)
    tl = L - 1 if target_level is None else target_level
    dt = capi.F64 if np.dtype(dtype) == np.float64 else capi.F32
    rc = capi.lib().prodm_interpret_retrieve_size(planner, estimator, dt, ndims, tl, s, L,
                                                  counts.ctypes.data_as(C.c_void_p), fs.ctypes.data_as(C.c_void_p),
                                                  fe.ctypes.data_as(C.c_void_p), tolerance,
                                                  idx.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    capi.check(rc, "prodm_interpret_retrieve_size")
    if isinstance(index, np.ndarray):
        index[:] = idx
    elif isinstance(index, list):
        index[:] = [int(v) for v in idx]
    return out, idx


def max_estimator_ob_c(ndims, dtype=np.float32):
    dt = capi.F64 if np.dtype(dtype) == np.float64 else capi.F32
    return capi.lib().prodm_max_estimator_ob_constant(ndims, dt)


def recompose_amplification(dims, L, hierarchical=False):
    """A_l, l = 0..L: the largest row 1-norm over every node of the level-l
    block of the recompose operator (exact; include/prodm/qoi.hpp)."""
    out = np.zeros(L + 1)
    d = np.ascontiguousarray(np.asarray(dims, np.uint32))
    rc = capi.lib().prodm_recompose_amplification(d.ctypes.data_as(C.c_void_p), len(dims), L,
                                                  capi.HIERARCHICAL if hierarchical else capi.ORTHOGONAL,
                                                  out.ctypes.data_as(C.c_void_p))
    capi.check(rc, "prodm_recompose_amplification")
    return out


def qoi_retrieve_sizes(sizes, errors, amplification, sample, tolerance, index, rounding=None,
                       qoi=capi.QOI_VELOCITY_MAGNITUDE):
    """Greedy planning on a QoI of several components (QoIGreedyBasedSizeInterpreter):
    sizes/errors/index as retrieve_sizes, one entry per (component, level),
    component-major; sample: (npoints, ncomp) reconstructed values; rounding:
    per-component allowance added to each component's bound. Returns (bytes
    to read, updated index)."""
    amplification = np.ascontiguousarray(np.asarray(amplification, np.float64))
    nl = amplification.size
    nk = len(sizes)
    ncomp = nk // nl
    if ncomp * nl != nk or len(errors) != nk:
        raise ValueError("expected one table per (component, level)")
    counts = np.array([len(x) for x in sizes], np.int32)
    fs = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint32) for x in sizes]))
    fe = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in errors]))
    smp = np.ascontiguousarray(np.asarray(sample, np.float64).reshape(-1, ncomp))
    rnd = None if rounding is None else np.ascontiguousarray(np.asarray(rounding, np.float64))
    idx = np.asarray(index, np.uint8).copy()
    out = np.zeros(nk, np.uint32)
    rc = capi.lib().prodm_qoi_interpret_retrieve_size(
        qoi, ncomp, nl, amplification.ctypes.data_as(C.c_void_p),
        None if rnd is None else rnd.ctypes.data_as(C.c_void_p), counts.ctypes.data_as(C.c_void_p),
        fs.ctypes.data_as(C.c_void_p), fe.ctypes.data_as(C.c_void_p), smp.ctypes.data_as(C.c_void_p), smp.shape[0],
        tolerance, idx.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    capi.check(rc, "prodm_qoi_interpret_retrieve_size")
    return out, idx
