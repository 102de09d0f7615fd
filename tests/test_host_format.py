"""ProDM's on-disk layout (metadata.bin + level_<l>.bin) against ProDM's own
code, no GPU. The compiled reference (oracle/_ref/libref.so, built from the
ProDM headers in place) reads with RefactorUtils.hpp's deserialize in the
field order of ComposedReconstructor::load_metadata
(ComposedReconstructor.hpp:133-154), ConcatLevelFileRetriever
(FileRetriever.hpp:15-81, two progressive steps) and
AdaptiveLevelCompressor::decompress_level, and writes with serialize in the
order of ComposedRefactor::write_metadata (ComposedRefactor.hpp:48-67) and
ConcatLevelFileWriter (FileWriter.hpp:13-39).

* The host writer (libprodm_host.so, include/prodm/io.hpp +
  compressors.hpp: the host stage GpuRefactor runs) and the reference's
  writer produce byte-identical files from the same plane streams.
* Each side reads the other's files back to the same metadata fields and
  plane streams, through a progressive two-step retrieval.

The plane streams are the oracle's encodings of a decomposed field
(NegaBinary for float32, GroupedBP for float64, in the reference stream
format), with the level bounds the composers store (x4 for negabinary).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_bind as ob
from fields import bumps_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.environ.get("PRODM_HOST_LIB") or os.path.join(ROOT, "prodm_amd", "libprodm_host.so")

pytestmark = pytest.mark.skipif(not ob.ref_available(), reason="compiled reference not built (oracle/_ref)")


def host():
    if not os.path.exists(HOST):
        pytest.fail("libprodm_host.so is not built: run `python -m prodm_amd.build`")
    return C.CDLL(HOST)


def level_streams(dtype, dims, L, nb):
    """Per level: the raw plane streams and the stored bound."""
    u = bumps_field(dims, seed=7, dtype=dtype)
    dec = ob.decompose(u, dims, L)
    ld = ob.level_dims(dims, L)
    out = []
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(len(dims), np.uint32)
        c = ob.interleave(dec, dims, ld[l], coarse)
        m = float(ob.maxabs(c))
        e = ob.frexp_exp(m, dtype)
        if dtype == np.float32:
            planes = [p.tobytes() for p in ob.nb_encode(c, e, nb)]
            bound = float(np.float32(m) * np.float32(4))
        else:
            planes = ob.gb_encode(c, e, nb)
            bound = m
        out.append((planes, bound))
    return out


def paths(d, L):
    meta = os.path.join(d, "metadata.bin")
    lv = [os.path.join(d, f"level_{l}.bin") for l in range(L + 1)]
    return meta, lv


def cstrs(v):
    arr = (C.c_char_p * len(v))(*[s.encode() for s in v])
    return arr


def write(lib, fn, d, dims, levels, dtype, negabinary, latter, threads=None):
    L = len(levels) - 1
    nb = len(levels[0][0])
    meta, lv = paths(d, L)
    raw = b"".join(b"".join(p) for p, _ in levels)
    buf = np.frombuffer(raw, np.uint8).copy() if raw else np.zeros(1, np.uint8)
    sizes = np.array([len(x) for p, _ in levels for x in p], np.uint32)
    bounds = np.array([b for _, b in levels], np.float64)
    d32 = np.array(dims, np.uint32)
    args = [meta.encode(), cstrs(lv), len(dims), d32.ctypes.data_as(C.c_void_p), L + 1, nb]
    if fn == "prodm_host_write_mdr":
        args += [1 if dtype == np.float64 else 0, bounds.ctypes.data_as(C.c_void_p), int(negabinary), latter,
                 threads or 4]
    else:
        args += [1 if dtype == np.float64 else 0, bounds.ctypes.data_as(C.c_void_p), int(negabinary), latter]
    args += [buf.ctypes.data_as(C.c_void_p), sizes.ctypes.data_as(C.c_void_p)]
    rc = getattr(lib, fn)(*args)
    assert rc == 0, (fn, rc)
    return meta, lv


def read(lib, fn, meta, lv, dtype, latter, k1, cap):
    L1 = len(lv)
    dims = np.zeros(3, np.uint32)
    nd, nl, npl, neg = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    bounds = np.zeros(64)
    lsz = np.zeros(64 * 64, np.uint32)
    stop = np.zeros(64, np.uint8)
    lnum = np.zeros(64, np.uint32)
    out = np.zeros(cap, np.uint8)
    raw = np.zeros(64 * 64, np.uint64)
    got = C.c_uint64()
    k = np.asarray(k1, np.uint8)
    args = [meta.encode(), cstrs(lv), L1, 1 if dtype == np.float64 else 0, latter]
    if fn == "prodm_host_read_mdr":
        args.append(4)
    args += [k.ctypes.data_as(C.c_void_p), dims.ctypes.data_as(C.c_void_p), C.byref(nd), C.byref(nl),
             bounds.ctypes.data_as(C.c_void_p), C.byref(npl), lsz.ctypes.data_as(C.c_void_p),
             stop.ctypes.data_as(C.c_void_p), lnum.ctypes.data_as(C.c_void_p), C.byref(neg),
             out.ctypes.data_as(C.c_void_p), C.c_uint64(cap), raw.ctypes.data_as(C.c_void_p), C.byref(got)]
    fnc = getattr(lib, fn)
    fnc.argtypes = None
    rc = fnc(*args)
    assert rc == 0, (fn, rc)
    nb = npl.value
    planes, off = [], 0
    for l in range(nl.value):
        lvp = []
        for p in range(nb):
            n = int(raw[l * nb + p])
            lvp.append(out[off:off + n].tobytes())
            off += n
        planes.append(lvp)
    return {"dims": dims[:nd.value].tolist(), "bounds": bounds[:nl.value].tolist(),
            "level_sizes": lsz[:nl.value * nb].reshape(nl.value, nb).tolist(), "stopping": stop[:nl.value].tolist(),
            "level_num": lnum[:nl.value].tolist(), "negabinary": neg.value, "planes": planes,
            "retrieved": got.value}


CASES = [(np.float32, [33, 17, 65], 3, 32, 64), (np.float32, [65, 65], 4, 32, 26), (np.float64, [17, 33, 17], 3, 64, 64),
         (np.float64, [129], 5, 40, 20)]


@pytest.mark.parametrize("dtype,dims,L,nb,latter", CASES)
def test_files_interoperate_with_reference(tmp_path, dtype, dims, L, nb, latter):
    levels = level_streams(dtype, dims, L, nb)
    nega = dtype == np.float32
    cap = sum(len(x) for p, _ in levels for x in p) + 64
    k1 = [(l * 7 + 3) % (nb + 1) for l in range(L + 1)]  # uneven first steps, some 0 and some all
    ours, theirs = tmp_path / "ours", tmp_path / "theirs"
    ours.mkdir()
    theirs.mkdir()
    h, r = host(), ob.ref()
    mo, lo = write(h, "prodm_host_write_mdr", str(ours), dims, levels, dtype, nega, latter)
    mr, lr = write(r, "ref_write_mdr", str(theirs), dims, levels, dtype, nega, latter)
    # byte-identical files
    assert open(mo, "rb").read() == open(mr, "rb").read()
    for a, b in zip(lo, lr):
        assert open(a, "rb").read() == open(b, "rb").read(), a
    # each side reads the other's files
    want_bounds = [float(dtype(b)) for _, b in levels]
    for fn, meta, lv in (("ref_read_mdr", mo, lo), ("prodm_host_read_mdr", mr, lr)):
        got = read(r if fn.startswith("ref") else h, fn, meta, lv, dtype, latter, k1, cap)
        assert got["dims"] == list(dims)
        assert got["bounds"] == want_bounds
        assert got["negabinary"] == int(nega)
        assert got["level_num"] == [1] * (L + 1)
        for l in range(L + 1):
            assert got["planes"][l] == levels[l][0], (fn, l)
        assert got["retrieved"] == sum(os.path.getsize(f) for f in lv)
    # the reference and the host reader agree on the stored sizes and stopping indices
    a = read(r, "ref_read_mdr", mo, lo, dtype, latter, k1, cap)
    b = read(h, "prodm_host_read_mdr", mo, lo, dtype, latter, k1, cap)
    assert a["level_sizes"] == b["level_sizes"] and a["stopping"] == b["stopping"]


def test_adaptive_compressor_stops_like_reference():
    """Streams chosen so that the stopping index falls below latter, at
    latter, above it, and nowhere; stored sizes and stopping indices match the
    reference's compressor (the host one skips planes past a stop below
    latter)."""
    rng = np.random.default_rng(3)
    nb = 32
    for stop_at in (5, 20, 26, 30, None):
        streams = []
        for p in range(nb):
            if stop_at is not None and p >= stop_at:
                streams.append(rng.integers(0, 256, 4096, dtype=np.uint8).tobytes())  # incompressible
            else:
                streams.append((np.arange(4096) % (p + 2)).astype(np.uint8).tobytes())
        for latter in (26, 64):
            want, wstop = ob.ref_compress_level(streams, kind=1, latter=latter)
            inp = np.frombuffer(b"".join(streams), np.uint8).copy()
            sizes = np.array([len(s) for s in streams], np.uint32)
            out = np.zeros(len(inp) * 2 + 8192, np.uint8)
            osz = np.zeros(nb, np.uint32)
            lib = host()
            stop = lib.prodm_host_compress_level(1, latter, 8, nb, inp.ctypes.data_as(C.c_void_p),
                                                 sizes.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                                 C.c_uint64(out.size), osz.ctypes.data_as(C.c_void_p))
            assert stop == wstop, (stop_at, latter, stop, wstop)
            got, off = [], 0
            for s in osz:
                got.append(out[off:off + int(s)].tobytes())
                off += int(s)
            assert got == want, (stop_at, latter)
