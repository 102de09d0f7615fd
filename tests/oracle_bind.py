"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and the compiled
reference checker (oracle/_ref/libref.so).

Test infrastructure: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py only.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# PRODM_ORACLE_LIB selects another build of the same oracle (e.g. the sanitizer build)
ORACLE_SO = os.environ.get("PRODM_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
ORACLE_MT_SO = os.path.join(ROOT, "oracle", "liboracle_mt.so")  # OpenMP build (all-core CPU baseline)
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _oracle = C.CDLL(ORACLE_SO)
    return _oracle


def use_library(path):
    """Point the bindings at another build of the oracle (ORACLE_MT_SO: the
    OpenMP build, same results); returns the previous library object."""
    global _oracle
    prev = _oracle
    if not os.path.exists(path):
        raise RuntimeError(f"oracle build not found: {path}")
    _oracle = C.CDLL(path)
    return prev


def restore_library(prev):
    global _oracle
    _oracle = prev


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
    return _ref


def _tp(dtype):
    return {np.float32: "f32", np.float64: "f64"}[np.dtype(dtype).type]


def _arr_u32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.uint32))


# ---- geometry ---------------------------------------------------------------
def level_dims(dims, L, lib=None):
    lib = lib or oracle()
    nd = len(dims)
    out = np.zeros((L + 1) * nd, np.uint32)
    (lib.orc_level_dims if lib is oracle() else lib.ref_level_dims)(
        _arr_u32(dims).ctypes.data_as(C.c_void_p), nd, L, out.ctypes.data_as(C.c_void_p))
    return out.reshape(L + 1, nd)


def level_elements(dims, L):
    ld = level_dims(dims, L)
    out = np.zeros(L + 1, np.uint64)
    oracle().orc_level_elements(ld.ctypes.data_as(C.c_void_p), len(dims), L, out.ctypes.data_as(C.c_void_p))
    return out


def geometry_ok(dims, L):
    return bool(oracle().orc_geometry_ok(_arr_u32(dims).ctypes.data_as(C.c_void_p), len(dims), L))


# ---- interleave ---------------------------------------------------------------
def interleave(data, dims, fine, coarse, strides=None, lib="oracle"):
    t = _tp(data.dtype)
    n = int(np.prod(fine, dtype=np.int64) - np.prod(coarse, dtype=np.int64))
    buf = np.zeros(n, data.dtype)
    L = oracle() if lib == "oracle" else ref()
    fn = getattr(L, ("orc_" if lib == "oracle" else "ref_") + "interleave_" + t)
    st = None if strides is None else _arr_u32(strides).ctypes.data_as(C.c_void_p)
    fn(data.ctypes.data_as(C.c_void_p), _arr_u32(dims).ctypes.data_as(C.c_void_p), len(dims),
       _arr_u32(fine).ctypes.data_as(C.c_void_p), _arr_u32(coarse).ctypes.data_as(C.c_void_p), st,
       buf.ctypes.data_as(C.c_void_p))
    return buf


def reposition(buf, data, dims, fine, coarse, strides=None, lib="oracle"):
    t = _tp(buf.dtype)
    L = oracle() if lib == "oracle" else ref()
    fn = getattr(L, ("orc_" if lib == "oracle" else "ref_") + "reposition_" + t)
    st = None if strides is None else _arr_u32(strides).ctypes.data_as(C.c_void_p)
    fn(buf.ctypes.data_as(C.c_void_p), _arr_u32(dims).ctypes.data_as(C.c_void_p), len(dims),
       _arr_u32(fine).ctypes.data_as(C.c_void_p), _arr_u32(coarse).ctypes.data_as(C.c_void_p), st,
       data.ctypes.data_as(C.c_void_p))
    return data


# ---- max / exponent -------------------------------------------------------------
def maxabs(x, lib="oracle"):
    t = _tp(x.dtype)
    if lib == "oracle":
        fn = getattr(oracle(), "orc_maxabs_" + t)
    else:
        fn = getattr(ref(), "ref_maxabs_" + t)
    fn.restype = C.c_float if t == "f32" else C.c_double
    return fn(x.ctypes.data_as(C.c_void_p), C.c_uint64(x.size) if lib == "oracle" else C.c_uint32(x.size))


def frexp_exp(v, dtype):
    t = _tp(dtype)
    fn = getattr(oracle(), "orc_frexp_exp_" + t)
    fn.argtypes = [C.c_float if t == "f32" else C.c_double]
    return fn(v)


# ---- negabinary -------------------------------------------------------------------
def _sdt(dtype):
    return np.uint32 if np.dtype(dtype) == np.float32 else np.uint64


def _block(dtype):
    return 32 if np.dtype(dtype) == np.float32 else 64


def nb_encode(x, exp, nb, lib="oracle"):
    t = _tp(x.dtype)
    n = x.size
    nw = (n - 1) // _block(x.dtype) + 1
    planes = np.zeros(nb * nw, _sdt(x.dtype))
    if lib == "oracle":
        rc = getattr(oracle(), "orc_nb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_uint32(n), C.c_int32(exp),
                                                     nb, planes.ctypes.data_as(C.c_void_p))
    else:
        rc = getattr(ref(), "ref_nb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_int32(n), C.c_int32(exp),
                                                  nb, planes.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return planes.reshape(nb, nw)


def nb_decode(planes, n, exp, start, k, dtype, lib="oracle"):
    """planes: 2-D array (>= k rows); row i is plane start + i."""
    t = _tp(dtype)
    out = np.zeros(n, dtype)
    rows = [np.ascontiguousarray(planes[i]) for i in range(k)]
    ptrs = (C.c_void_p * max(k, 1))(*[r.ctypes.data for r in rows])
    if lib == "oracle":
        rc = getattr(oracle(), "orc_nb_decode_" + t)(ptrs, C.c_uint32(n), C.c_int32(exp), start, k,
                                                     out.ctypes.data_as(C.c_void_p))
    else:
        rc = getattr(ref(), "ref_nb_decode_" + t)(ptrs, C.c_int32(n), C.c_int32(exp), start, k,
                                                  out.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return out


# ---- grouped ----------------------------------------------------------------------
def gb_encode(x, exp, nb, compat=False, lib="oracle"):
    """Returns list of nb byte strings (reference stream format)."""
    t = _tp(x.dtype)
    n = x.size
    fnm = getattr(oracle(), "orc_gb_max_bytes_" + t)
    fnm.restype = C.c_uint64
    cap = fnm(C.c_uint32(n), nb)
    out = np.zeros(cap, np.uint8)
    sizes = np.zeros(nb, np.uint32)
    if lib == "oracle":
        rc = getattr(oracle(), "orc_gb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_uint32(n), C.c_int32(exp),
                                                     nb, int(compat), out.ctypes.data_as(C.c_void_p),
                                                     C.c_uint64(cap), sizes.ctypes.data_as(C.c_void_p))
    else:
        rc = getattr(ref(), "ref_gb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_int32(n), C.c_int32(exp),
                                                  nb, out.ctypes.data_as(C.c_void_p), C.c_uint64(cap),
                                                  sizes.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    res, off = [], 0
    for s in sizes:
        res.append(out[off:off + int(s)].tobytes())
        off += int(s)
    return res


class GBDecoder:
    """Progressive GroupedBP decoder with per-level state (oracle or ref)."""

    def __init__(self, dtype, lib="oracle", max_levels=32):
        self.t = _tp(dtype)
        self.dtype = dtype
        self.lib = lib
        if lib == "oracle":
            fn = oracle().orc_gb_state_new
            fn.restype = C.c_void_p
            self.h = fn(max_levels)
        else:
            fn = getattr(ref(), "ref_gb_new_" + self.t)
            fn.restype = C.c_void_p
            self.h = fn()

    def decode(self, level, streams, n, exp, start, k):
        out = np.zeros(n, self.dtype)
        bufs = [np.frombuffer(s, np.uint8).copy() if len(s) else np.zeros(8, np.uint8) for s in streams[:k]]
        ptrs = (C.c_void_p * max(k, 1))(*[b.ctypes.data for b in bufs])
        if self.lib == "oracle":
            rc = getattr(oracle(), "orc_gb_decode_" + self.t)(C.c_void_p(self.h), level, ptrs, C.c_uint32(n),
                                                              C.c_int32(exp), start, k,
                                                              out.ctypes.data_as(C.c_void_p))
        else:
            rc = getattr(ref(), "ref_gb_decode_" + self.t)(C.c_void_p(self.h), level, ptrs, C.c_int32(n),
                                                           C.c_int32(exp), start, k,
                                                           out.ctypes.data_as(C.c_void_p))
        assert rc == 0, rc
        return out

    def __del__(self):
        try:
            if self.lib == "oracle":
                oracle().orc_gb_state_free(C.c_void_p(self.h))
            else:
                getattr(ref(), "ref_gb_free_" + self.t)(C.c_void_p(self.h))
        except Exception:
            pass


def pb_encode(x, exp, nb, lib="oracle"):
    """PerBit streams (PerBitBPEncoder.hpp:72-146): list of nb byte strings."""
    t = _tp(x.dtype)
    n = x.size
    words = (2 * n + 63) // 64 + 1
    sizes = np.zeros(nb, np.uint32)
    if lib == "oracle":
        bufs = [np.zeros(words, np.uint64) for _ in range(nb)]
        ptrs = (C.c_void_p * nb)(*[b.ctypes.data for b in bufs])
        rc = getattr(oracle(), "orc_pb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_uint32(n), C.c_int32(exp), nb,
                                                     ptrs, sizes.ctypes.data_as(C.c_void_p))
        assert rc == 0, rc
        return [bufs[p].view(np.uint8)[:int(sizes[p])].tobytes() for p in range(nb)]
    cap = nb * (words * 8 + 64)
    out = np.zeros(cap, np.uint8)
    rc = getattr(ref(), "ref_pb_encode_" + t)(x.ctypes.data_as(C.c_void_p), C.c_int32(n), C.c_int32(exp), nb,
                                              out.ctypes.data_as(C.c_void_p), C.c_uint64(cap),
                                              sizes.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    res, off = [], 0
    for sz in sizes:
        res.append(out[off:off + int(sz)].tobytes())
        off += int(sz)
    return res


class PBDecoder:
    """Progressive PerBit decoder with per-level sign state (oracle or ref)."""

    def __init__(self, dtype, lib="oracle", max_levels=32):
        self.t = _tp(dtype)
        self.dtype = dtype
        self.lib = lib
        if lib == "oracle":
            fn = oracle().orc_pb_state_new
            fn.restype = C.c_void_p
            self.h = fn(max_levels)
        else:
            fn = getattr(ref(), "ref_pb_new_" + self.t)
            fn.restype = C.c_void_p
            self.h = fn()

    def decode(self, level, streams, n, exp, start, k):
        out = np.zeros(n, self.dtype)
        # one spare word: the reference's bit reader may load the word after the last bit
        bufs = [np.frombuffer(s + b"\0" * 8, np.uint8).copy() for s in streams[:k]]
        ptrs = (C.c_void_p * max(k, 1))(*[b.ctypes.data for b in bufs])
        if self.lib == "oracle":
            rc = getattr(oracle(), "orc_pb_decode_" + self.t)(C.c_void_p(self.h), level, ptrs, C.c_uint32(n),
                                                              C.c_int32(exp), start, k,
                                                              out.ctypes.data_as(C.c_void_p))
        else:
            rc = getattr(ref(), "ref_pb_decode_" + self.t)(C.c_void_p(self.h), level, ptrs, C.c_int32(n),
                                                           C.c_int32(exp), start, k,
                                                           out.ctypes.data_as(C.c_void_p))
        assert rc == 0, rc
        return out

    def __del__(self):
        try:
            if self.lib == "oracle":
                oracle().orc_pb_state_free(C.c_void_p(self.h))
            else:
                getattr(ref(), "ref_pb_free_" + self.t)(C.c_void_p(self.h))
        except Exception:
            pass


# ---- collectors ---------------------------------------------------------------------
def collect_max(bound, nb, dtype, lib="oracle"):
    t = _tp(dtype)
    out = np.zeros(nb + 1, np.float64)
    fn = getattr(oracle() if lib == "oracle" else ref(), ("orc_" if lib == "oracle" else "ref_") + "collect_max_" + t)
    fn.argtypes = [C.c_float if t == "f32" else C.c_double, C.c_int, C.c_void_p]
    fn(bound, nb, out.ctypes.data_as(C.c_void_p))
    return out


def collect_squared(x, nb, maxerr, mode="exact"):
    t = _tp(x.dtype)
    out = np.zeros(nb + 1, np.float64)
    if mode == "ref":
        fn = ref().ref_collect_squared_f32
    elif mode == "seq":
        fn = getattr(oracle(), "orc_collect_squared_seq_" + t)
    else:
        fn = getattr(oracle(), "orc_collect_squared_" + t)
    fn.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_float if t == "f32" else C.c_double, C.c_void_p]
    fn(x.ctypes.data_as(C.c_void_p), x.size, nb, maxerr, out.ctypes.data_as(C.c_void_p))
    return out


def fused_level_errors(x, exp, nb, enc, mode="exact"):
    """Level-error table of the encoders' level_errors overloads (fp32):
    enc "nb" (NegaBinaryBPEncoder), "gb" (GroupedBPEncoder) or, with mode "ref"
    only, "pb" (PerBitBPEncoder); mode "exact"
    (exactly rounded sums), "seq" (reference order) or "ref" (the compiled
    reference, n >= 32)."""
    assert x.dtype == np.float32
    out = np.zeros(nb + 1, np.float64)
    if mode == "ref":
        fn = getattr(ref(), f"ref_{enc}_level_errors_f32")
        fn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int, C.c_void_p]
    else:
        fn = getattr(oracle(), f"orc_{enc}_level_errors{'_seq' if mode == 'seq' else ''}_f32")
        fn.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_int, C.c_void_p]
    fn(x.ctypes.data_as(C.c_void_p), x.size, int(exp), int(nb), out.ctypes.data_as(C.c_void_p))
    return out


# ---- MGARD ----------------------------------------------------------------------------
def decompose(data, dims, L, strides=None, hierarchical=False):
    t = _tp(data.dtype)
    out = np.ascontiguousarray(data).copy()
    st = None if strides is None else _arr_u32(strides).ctypes.data_as(C.c_void_p)
    rc = getattr(oracle(), ("orc_decompose_hb_" if hierarchical else "orc_decompose_") + t)(out.ctypes.data_as(C.c_void_p), _arr_u32(dims).ctypes.data_as(C.c_void_p),
                                                  len(dims), L, st)
    assert rc == 0, rc
    return out


def recompose(data, dims, L, strides=None, hierarchical=False):
    t = _tp(data.dtype)
    out = np.ascontiguousarray(data).copy()
    st = None if strides is None else _arr_u32(strides).ctypes.data_as(C.c_void_p)
    rc = getattr(oracle(), ("orc_recompose_hb_" if hierarchical else "orc_recompose_") + t)(out.ctypes.data_as(C.c_void_p), _arr_u32(dims).ctypes.data_as(C.c_void_p),
                                                  len(dims), L, st)
    assert rc == 0, rc
    return out


def recompose_adjoint(g, dims, L):
    out = np.ascontiguousarray(g, dtype=np.float64).copy()
    rc = oracle().orc_recompose_adjoint_f64(out.ctypes.data_as(C.c_void_p), _arr_u32(dims).ctypes.data_as(C.c_void_p),
                                            len(dims), L, None)
    assert rc == 0, rc
    return out


# ---- reference-only helpers ---------------------------------------------------------
def ref_interpret(kind, est, nd, L, counts, sizes, errors, tol, index, s_param=0.0):
    counts = np.asarray(counts, np.int32)
    sizes = np.ascontiguousarray(np.concatenate([np.asarray(s, np.uint32) for s in sizes]))
    errors = np.ascontiguousarray(np.concatenate([np.asarray(e, np.float64) for e in errors]))
    idx = np.asarray(index, np.uint8).copy()
    ret = np.zeros(L, np.uint32)
    ref().ref_interpret(C.c_int(kind), C.c_int(est), C.c_int(nd), C.c_double(s_param), C.c_int(L),
                        counts.ctypes.data_as(C.c_void_p), sizes.ctypes.data_as(C.c_void_p),
                        errors.ctypes.data_as(C.c_void_p), C.c_double(tol), idx.ctypes.data_as(C.c_void_p),
                        ret.ctypes.data_as(C.c_void_p))
    return ret, idx


def ref_max_estimator_c(nd):
    fn = ref().ref_max_estimator_ob_c
    fn.restype = C.c_double
    return fn(nd)


def ref_compress_level(streams, kind=1, latter=64):
    nb = len(streams)
    sizes = np.array([len(s) for s in streams], np.uint32)
    inp = np.frombuffer(b"".join(streams), np.uint8).copy() if sum(sizes) else np.zeros(1, np.uint8)
    cap = int(sum(sizes) * 2 + 4096 * nb)
    out = np.zeros(cap, np.uint8)
    osz = np.zeros(nb, np.uint32)
    stop = ref().ref_compress_level(kind, latter, nb, inp.ctypes.data_as(C.c_void_p), sizes.ctypes.data_as(C.c_void_p),
                                    out.ctypes.data_as(C.c_void_p), C.c_uint64(cap), osz.ctypes.data_as(C.c_void_p))
    assert stop >= 0
    res, off = [], 0
    for s in osz:
        res.append(out[off:off + int(s)].tobytes())
        off += int(s)
    return res, stop
