"""Torch-tensor front end of the C ABI (device memory and streams from torch).

Every function here runs the HIP library; inputs must be CUDA (HIP) tensors
on the current device. Shapes follow ProDM: `dims` lists the extents with the
slowest dimension first (DirectInterleaver.hpp:20-21).
"""
import ctypes as C

import numpy as np
import torch

from . import capi
from .capi import check, lib

_TDT = {torch.float32: capi.F32, torch.float64: capi.F64}
_SFX = {torch.float32: "f32", torch.float64: "f64"}
_WORD = {torch.float32: torch.int32, torch.float64: torch.int64}
_NPWORD = {torch.float32: np.uint32, torch.float64: np.uint64}
_BLOCK = {torch.float32: 32, torch.float64: 64}


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _u32(v):
    a = (C.c_uint32 * len(v))(*[int(x) for x in v])
    return a


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _fn(name, dtype):
    return getattr(lib(), f"{name}_{_SFX[dtype]}")


def level_geometry(dims, L):
    nd = len(dims)
    out = (C.c_uint32 * ((L + 1) * nd))()
    el = (C.c_uint64 * (L + 1))()
    ok = lib().prodm_level_geometry(_u32(dims), nd, L, out, el)
    return np.array(out[:], np.uint32).reshape(L + 1, nd), np.array(el[:], np.uint64), bool(ok)


def _ws(nbytes):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")


# ---- MGARD -------------------------------------------------------------------
def decompose(x, dims, L, strides=None, hierarchical=False):
    out = torch.empty_like(x)
    nb = lib().prodm_mgard_workspace_bytes(_u32(dims), len(dims), L, _TDT[x.dtype])
    if nb == 0:
        raise capi.ProdmError("unsupported geometry")
    ws = _ws(nb)
    st = _u32(strides) if strides is not None else None
    name = "prodm_decompose_hierarchical" if hierarchical else "prodm_decompose"
    check(_fn(name, x.dtype)(_ptr(x), _ptr(out), _u32(dims), len(dims), L, st, _ptr(ws), ws.numel(), _stream()),
          "decompose")
    return out


def decompose_hierarchical(x, dims, L, strides=None):
    return decompose(x, dims, L, strides, hierarchical=True)


def recompose_hierarchical(c, dims, L, strides=None):
    return recompose(c, dims, L, strides, hierarchical=True)


def recompose(c, dims, L, strides=None, hierarchical=False):
    out = torch.empty_like(c)
    nb = lib().prodm_mgard_workspace_bytes(_u32(dims), len(dims), L, _TDT[c.dtype])
    if nb == 0:
        raise capi.ProdmError("unsupported geometry")
    ws = _ws(nb)
    st = _u32(strides) if strides is not None else None
    name = "prodm_recompose_hierarchical" if hierarchical else "prodm_recompose"
    check(_fn(name, c.dtype)(_ptr(c), _ptr(out), _u32(dims), len(dims), L, st, _ptr(ws), ws.numel(), _stream()),
          "recompose")
    return out


def recompose_adjoint(g, dims, L, strides=None):
    """dJ/dcoefficients (decompose layout) from dJ/du, the transpose of
    recompose; FP64 inside."""
    out = torch.empty_like(g)
    nb = lib().prodm_recompose_adjoint_workspace_bytes(_u32(dims), len(dims), L)
    if nb == 0:
        raise capi.ProdmError("unsupported geometry")
    ws = _ws(nb)
    st = _u32(strides) if strides is not None else None
    check(_fn("prodm_recompose_adjoint", g.dtype)(_ptr(g), _ptr(out), _u32(dims), len(dims), L, st, _ptr(ws),
                                                  ws.numel(), _stream()), "recompose_adjoint")
    return out


# ---- interleave ------------------------------------------------------------------
def interleave(data, dims, fine, coarse, strides=None):
    n = int(np.prod(fine, dtype=np.int64) - np.prod(coarse, dtype=np.int64))
    buf = torch.empty(n, dtype=data.dtype, device=data.device)
    st = _u32(strides) if strides is not None else None
    check(_fn("prodm_interleave", data.dtype)(_ptr(data), _u32(dims), len(dims), _u32(fine), _u32(coarse), st, _ptr(buf),
                                              _stream()), "interleave")
    return buf


def reposition(buf, data, dims, fine, coarse, strides=None):
    st = _u32(strides) if strides is not None else None
    check(_fn("prodm_reposition", buf.dtype)(_ptr(buf), _u32(dims), len(dims), _u32(fine), _u32(coarse), st, _ptr(data),
                                             _stream()), "reposition")
    return data


def maxabs(x):
    out = C.c_float() if x.dtype == torch.float32 else C.c_double()
    check(_fn("prodm_level_maxabs", x.dtype)(_ptr(x), x.numel(), C.byref(out), _stream()), "maxabs")
    return out.value


# ---- NegaBinary ----------------------------------------------------------------------
def negabinary_encode(x, exp, nb):
    """Returns a (nb, stride) word tensor; plane p = row p[:ceil(n/B)]."""
    n = x.numel()
    B = _BLOCK[x.dtype]
    nw = (n + B - 1) // B
    stride = (nw + 1) & ~1
    planes = torch.zeros((nb, stride), dtype=_WORD[x.dtype], device=x.device)
    check(_fn("prodm_negabinary_encode", x.dtype)(_ptr(x), n, exp, nb, _ptr(planes), stride, _stream()),
          "negabinary encode")
    return planes


def negabinary_decode(plane_rows, n, exp, start, k, dtype):
    out = torch.empty(n, dtype=dtype, device="cuda")
    ptrs = (C.c_void_p * max(k, 1))(*[r.data_ptr() for r in plane_rows[:k]])
    check(_fn("prodm_negabinary_decode", dtype)(ptrs, n, exp, start, k, _ptr(out), _stream()), "negabinary decode")
    return out


def planes_to_numpy(planes, n, dtype):
    """Device plane tensor -> (nb, ceil(n/B)) unsigned numpy array."""
    B = _BLOCK[dtype]
    nw = (n + B - 1) // B
    return planes.cpu().numpy().view(_NPWORD[dtype])[:, :nw].copy()


# ---- GroupedBP ------------------------------------------------------------------------
def grouped_encode(x, exp, nb, compat=False):
    """Returns (rb uint8 tensor, planes word tensor (nb, 2*nblocks), words int32 tensor)."""
    n = x.numel()
    B = _BLOCK[x.dtype]
    nblk = (n - 1) // B + 1
    stride = 2 * nblk
    rb = torch.zeros(nblk, dtype=torch.uint8, device=x.device)
    planes = torch.zeros((nb, stride), dtype=_WORD[x.dtype], device=x.device)
    words = torch.zeros(nb, dtype=torch.int32, device=x.device)
    wsb = lib().prodm_grouped_workspace_bytes(n, nb, _TDT[x.dtype])
    ws = _ws(wsb)
    check(_fn("prodm_grouped_encode", x.dtype)(_ptr(x), n, exp, nb, int(compat), _ptr(rb), _ptr(planes), stride,
                                               _ptr(words), _ptr(ws), ws.numel(), _stream()), "grouped encode")
    return rb, planes, words


def grouped_streams(rb, planes, words, dtype):
    """Assemble the reference byte streams (GroupedBPEncoder.hpp:63-72) on the host."""
    rbh = rb.cpu().numpy()
    ph = planes.cpu().numpy().view(_NPWORD[dtype])
    wh = words.cpu().numpy()
    streams = []
    for p in range(ph.shape[0]):
        body = ph[p, : int(wh[p])].tobytes()
        if p == 0:
            body = np.uint32(rbh.size).tobytes() + rbh.tobytes() + body
        streams.append(body)
    return streams


class GroupedDecoder:
    """Progressive GroupedBP decode with device-resident per-level state."""

    def __init__(self, dtype, max_levels=32):
        self.dtype = dtype
        h = C.c_void_p()
        check(lib().prodm_grouped_state_create(C.byref(h), max_levels, _TDT[dtype]), "grouped state")
        self.h = h

    def decode(self, level, streams, n, exp, start, k):
        """streams: list of device uint8 tensors (reference format)."""
        out = torch.empty(n, dtype=self.dtype, device="cuda")
        ptrs = (C.c_void_p * max(k, 1))(*[s.data_ptr() for s in streams[:k]])
        check(_fn("prodm_grouped_decode", self.dtype)(self.h, level, ptrs, n, exp, start, k, _ptr(out), _stream()),
              "grouped decode")
        return out

    def close(self):
        if self.h:
            lib().prodm_grouped_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- PerBit ---------------------------------------------------------------------------
def perbit_encode(x, exp, nb, compat=False):
    """Returns (streams uint64-word tensor (nb, stride), sizes int32 tensor of
    bytes); stream p is row p[:sizes[p] // 8] (PerBitBPEncoder.hpp:81-143).
    compat reproduces x86's (int64_t) on |v| >= 2^63 or NaN."""
    n = x.numel()
    stride = int(lib().prodm_perbit_stream_words(n))
    streams = torch.empty((nb, max(stride, 1)), dtype=torch.int64, device=x.device)
    sizes = torch.zeros(nb, dtype=torch.int32, device=x.device)
    ws = _ws(lib().prodm_perbit_workspace_bytes(n, nb, _TDT[x.dtype]))
    check(_fn("prodm_perbit_encode", x.dtype)(_ptr(x), n, int(exp), nb, int(compat), _ptr(streams), stride, _ptr(sizes), _ptr(ws),
                                              ws.numel(), _stream()), "perbit encode")
    return streams, sizes


def perbit_streams(streams, sizes):
    """Device encode output -> list of the reference's byte streams."""
    sh = streams.cpu().numpy().view(np.uint64)
    zh = sizes.cpu().numpy()
    return [sh[p, : int(zh[p]) // 8].tobytes() for p in range(sh.shape[0])]


class PerBitDecoder:
    """Progressive PerBit decode (PerBitBPEncoder.hpp:290-382). The format
    has no index of where an element's bits begin, so it decodes on the
    host; the per-level sign state is kept in the library, keyed by level."""

    def __init__(self, dtype):
        self.dtype = dtype
        h = C.c_void_p()
        check(lib().prodm_perbit_state_create(C.byref(h)), "perbit state")
        self.h = h

    def decode(self, level, streams, n, exp, start, k):
        """streams: list of byte strings (stream start + i); returns numpy."""
        npdt = np.float32 if self.dtype == torch.float32 else np.float64
        out = np.empty(n, dtype=npdt)
        bufs = [np.frombuffer(s, dtype=np.uint64) if len(s) else np.zeros(1, np.uint64) for s in streams[:k]]
        ptrs = (C.c_void_p * max(k, 1))(*[b.ctypes.data for b in bufs])
        sizes = (C.c_uint32 * max(k, 1))(*[len(s) for s in streams[:k]])
        check(_fn("prodm_perbit_decode", self.dtype)(self.h, level, ptrs, sizes, n, int(exp), start, k,
                                                     out.ctypes.data), "perbit decode")
        return out

    def close(self):
        if self.h:
            lib().prodm_perbit_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- collectors ---------------------------------------------------------------------------
def collect_max(bound, nb, dtype=torch.float32):
    out = (C.c_double * (nb + 1))()
    _fn("prodm_collect_max", dtype)(bound, nb, out)
    return np.array(out[:])


def collect_squared(x, nb, maxerr):
    n = x.numel()
    wsb = lib().prodm_collect_squared_workspace_bytes(n, nb, _TDT[x.dtype])
    ws = _ws(wsb)
    out = (C.c_double * (nb + 1))()
    check(_fn("prodm_collect_squared", x.dtype)(_ptr(x), n, nb, maxerr, out, _ptr(ws), ws.numel(), _stream()),
          "squared collector")
    return np.array(out[:])


def level_errors(x, exp, nb, encoder=capi.NEGABINARY):
    """Level-error table of the encoders' level_errors overloads (fp32):
    entry j = squared error of keeping the first j planes of the codec."""
    if x.dtype != torch.float32:
        raise ValueError("the encoders' level errors are defined for float32 data")
    n = x.numel()
    ws = _ws(lib().prodm_level_errors_workspace_bytes(n, nb))
    out = (C.c_double * (nb + 1))()
    fn = lib().prodm_negabinary_level_errors_f32 if encoder == capi.NEGABINARY else lib().prodm_grouped_level_errors_f32
    check(fn(_ptr(x), n, int(exp), nb, out, _ptr(ws), ws.numel(), _stream()), "level errors")
    return np.array(out[:])


# ---- pipeline plan ---------------------------------------------------------------------------
class Plan:
    """Device-resident refactor/retrieve of one field (prodm_mdr_plan)."""

    def __init__(self, dims, L, nb, dtype=torch.float32, encoder=capi.NEGABINARY, decomposer=capi.ORTHOGONAL):
        self.dims, self.L, self.nb, self.dtype, self.encoder = list(dims), L, nb, dtype, encoder
        h = C.c_void_p()
        check(lib().prodm_mdr_plan_create(C.byref(h), _u32(dims), len(dims), L, nb, _TDT[dtype], encoder), "plan")
        self.h = h
        if decomposer != capi.ORTHOGONAL:
            check(lib().prodm_mdr_plan_set_decomposer(self.h, decomposer), "plan decomposer")
        self.level_dims, self.level_elements, _ = level_geometry(dims, L)

    @property
    def device_bytes(self):
        return int(lib().prodm_mdr_plan_device_bytes(self.h))

    def refactor(self, x):
        check(lib().prodm_mdr_refactor(self.h, _ptr(x), _stream()), "refactor")

    def level_info(self, level):
        m, e, n = C.c_double(), C.c_int32(), C.c_uint64()
        check(lib().prodm_mdr_level_info(self.h, level, C.byref(m), C.byref(e), C.byref(n)), "level info")
        return m.value, e.value, n.value

    def level_errors(self, level):
        """the plan encoder's level-error table of a level (fp32 plans)"""
        out = (C.c_double * (self.nb + 1))()
        check(lib().prodm_mdr_level_errors(self.h, level, out, _stream()), "plan level errors")
        return np.array(out[:])

    def plane(self, level, p):
        ptr, nbytes = C.c_void_p(), C.c_uint64()
        check(lib().prodm_mdr_plane(self.h, level, p, C.byref(ptr), C.byref(nbytes)), "plane")
        return ptr.value, nbytes.value

    def plane_bytes(self, level, p):
        ptr, nbytes = self.plane(level, p)
        buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
        if nbytes:
            torch.cuda.synchronize()
            _d2d(buf, ptr, int(nbytes))
        return buf[: int(nbytes)].cpu().numpy().tobytes()

    def level_rb(self, level):
        ptr, nblk = C.c_void_p(), C.c_uint64()
        check(lib().prodm_mdr_level_rb(self.h, level, C.byref(ptr), C.byref(nblk)), "level rb")
        nblk = int(nblk.value)
        buf = torch.empty(nblk, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        _d2d(buf, ptr.value, nblk)
        return buf.cpu().numpy()

    def coefficients(self, level):
        ptr = C.c_void_p()
        check(lib().prodm_mdr_level_coefficients(self.h, level, C.byref(ptr)), "coefficients")
        n = int(self.level_elements[level])
        buf = torch.empty(n, dtype=self.dtype, device="cuda")
        torch.cuda.synchronize()
        _d2d(buf, ptr.value, n * buf.element_size())
        return buf

    def retrieve(self, planes_per_level, out=None):
        if out is None:
            out = torch.empty(int(np.prod(self.dims)), dtype=self.dtype, device="cuda")
        k = (C.c_uint8 * (self.L + 1))(*[int(v) for v in planes_per_level])
        check(lib().prodm_mdr_retrieve(self.h, k, _ptr(out), _stream()), "retrieve")
        return out

    def refactor_graph(self, x):
        """refactor through a captured HIP graph (current stream must not be
        the default stream)."""
        check(lib().prodm_mdr_refactor_graph(self.h, _ptr(x), _stream()), "refactor_graph")

    def retrieve_graph(self, planes_per_level, out):
        k = (C.c_uint8 * (self.L + 1))(*[int(v) for v in planes_per_level])
        check(lib().prodm_mdr_retrieve_graph(self.h, k, _ptr(out), _stream()), "retrieve_graph")
        return out

    def plane_stream(self, level, p):
        """Reference-format byte stream of plane p of a level (host bytes)."""
        n = C.c_uint64()
        check(lib().prodm_mdr_plane_stream(self.h, level, p, None, 0, C.byref(n), _stream()), "plane stream size")
        buf = np.zeros(max(int(n.value), 1), np.uint8)
        check(lib().prodm_mdr_plane_stream(self.h, level, p, buf.ctypes.data_as(C.c_void_p), buf.size, C.byref(n),
                                           _stream()), "plane stream")
        return buf[: int(n.value)].tobytes()

    def load_plane_stream(self, level, p, data):
        arr = np.frombuffer(data, np.uint8) if len(data) else np.zeros(1, np.uint8)
        check(lib().prodm_mdr_load_plane_stream(self.h, level, p, arr.ctypes.data_as(C.c_void_p), len(data),
                                                _stream()), "load plane stream")

    def set_level_max(self, level, max_abs):
        check(lib().prodm_mdr_set_level_max(self.h, level, float(max_abs), _stream()), "set level max")

    def profile(self, enable=True):
        check(lib().prodm_mdr_profile(self.h, int(enable)), "profile")

    def stage_times(self):
        """{stage: (ms summed, launch groups)} since the previous call."""
        n = len(capi.STAGES)
        ms = (C.c_double * n)()
        cnt = (C.c_uint32 * n)()
        check(lib().prodm_mdr_stage_times(self.h, ms, cnt, n), "stage times")
        return {capi.STAGES[i]: (ms[i], cnt[i]) for i in range(n)}

    @classmethod
    def borrowed(cls, handle, dims, L, nb, dtype, encoder, owner):
        """A Plan view of a handle owned elsewhere (a batch field); `owner`
        is kept alive and the handle is never destroyed through the view."""
        self = cls.__new__(cls)
        self.dims, self.L, self.nb, self.dtype, self.encoder = list(dims), L, nb, dtype, encoder
        self.h = C.c_void_p(handle)
        self._owner = owner
        self.level_dims, self.level_elements, _ = level_geometry(dims, L)
        return self

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            lib().prodm_mdr_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hip = None


def _d2d(dst_tensor, src_ptr, nbytes):
    """Copy nbytes from a raw device pointer into a tensor; returns once the
    copy is complete. A device-to-device hipMemcpy goes to the null stream
    and may return before the copy ends, and torch's side streams do not
    wait for the null stream: without the device synchronization a kernel
    on such a stream could read the tensor half-copied."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = C.c_int
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.restype = C.c_int
        _hip.hipDeviceSynchronize.argtypes = []
    rc = _hip.hipMemcpy(C.c_void_p(dst_tensor.data_ptr()), C.c_void_p(src_ptr), nbytes, 3)
    if rc != 0:
        raise capi.ProdmError(f"hipMemcpy failed: {rc}")
    rc = _hip.hipDeviceSynchronize()
    if rc != 0:
        raise capi.ProdmError(f"hipDeviceSynchronize after the copy failed: {rc}")
