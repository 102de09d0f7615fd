"""Batched refactor of many fields of uneven shapes (prodm_mdr_batch).

ProDM refactors one field per ComposedRefactor::refactor call
(include/MDR/Refactor/ComposedRefactor.hpp:78-141). A Batch runs the GPU
stages of all its fields together, with one grouped launch per kernel and
level step, and keeps each field's results in its own plan: `batch.plan(f)`
is a Plan view (planes, level maxima, plane streams, retrieve) whose results
are bit-identical to Plan(...).refactor of that field alone.
"""
import ctypes as C

import torch

from . import capi
from .capi import check, lib
from .mdr import Plan, _TDT, _ptr, _stream


class Batch:
    """fields: sequence of (tensor, dims, target_level); each tensor dense on
    the GPU in the batch dtype with prod(dims) elements (dims[0] slowest)."""

    def __init__(self, fields, nb=32, dtype=torch.float32, encoder=capi.NEGABINARY, decomposer=capi.ORTHOGONAL):
        self.nb, self.dtype, self.encoder = nb, dtype, encoder
        self.fields = []
        descs = (capi.FieldDesc * len(fields))()
        for i, (x, dims, L) in enumerate(fields):
            dims = [int(d) for d in dims]
            n = 1
            for d in dims:
                n *= d
            if x.dtype != dtype or not x.is_cuda or not x.is_contiguous() or x.numel() != n:
                raise ValueError(f"field {i}: need a contiguous cuda {dtype} tensor of {n} elements")
            descs[i].d_in = x.data_ptr()
            for k, d in enumerate(dims):
                descs[i].dims[k] = d
            descs[i].ndims = len(dims)
            descs[i].target_level = int(L)
            self.fields.append((x, dims, int(L)))
        h = C.c_void_p()
        check(lib().prodm_mdr_batch_create(C.byref(h), descs, len(fields), nb, _TDT[dtype], encoder), "batch")
        self.h = h
        if decomposer != capi.ORTHOGONAL:
            check(lib().prodm_mdr_batch_set_decomposer(self.h, decomposer), "batch decomposer")

    def __len__(self):
        return len(self.fields)

    @property
    def device_bytes(self):
        return int(lib().prodm_mdr_batch_device_bytes(self.h))

    def set_input(self, f, x):
        _, dims, L = self.fields[f]
        check(lib().prodm_mdr_batch_set_input(self.h, f, _ptr(x)), "batch set input")
        self.fields[f] = (x, dims, L)

    def refactor(self):
        """grouped launches on the current stream"""
        check(lib().prodm_mdr_batched_refactor(self.h, _stream()), "batched refactor")

    def refactor_graph(self):
        """the same sequence through a captured HIP graph (the current stream
        must not be the default stream)"""
        check(lib().prodm_mdr_batched_refactor_graph(self.h, _stream()), "batched refactor graph")

    def autotune(self, cache_path=None, reps=3):
        path = cache_path.encode() if cache_path else None
        check(lib().prodm_mdr_batch_autotune(self.h, path, int(reps), _stream()), "batch autotune")

    def describe(self):
        n = C.c_uint64()
        check(lib().prodm_mdr_batch_describe(self.h, None, 0, C.byref(n)), "batch describe")
        buf = C.create_string_buffer(int(n.value) + 1)
        check(lib().prodm_mdr_batch_describe(self.h, buf, len(buf), C.byref(n)), "batch describe")
        return buf.value.decode()

    def plan(self, f):
        _, dims, L = self.fields[f]
        h = lib().prodm_mdr_batch_plan(self.h, f)
        if not h:
            raise capi.ProdmError(f"batch plan {f}: " + lib().prodm_last_error().decode(errors="replace"))
        return Plan.borrowed(h, dims, L, self.nb, self.dtype, self.encoder, owner=self)

    def close(self):
        if getattr(self, "h", None):
            lib().prodm_mdr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
