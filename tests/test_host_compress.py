"""The host lossless stage (include/prodm/compressors.hpp through
libprodm_host.so) against the compiled ProDM level compressors: the same
stored bytes and stopping index for the default and the adaptive
compressor, and the inverse restores every plane."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_bind as ob

HOST = os.environ.get("PRODM_HOST_LIB") or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "prodm_amd", "libprodm_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST):
        pytest.fail("libprodm_host.so is not built: run `python -m prodm_amd.build`")
    return C.CDLL(HOST)


def planes_like(nb, seed, n=20000):
    """Bitplane-like streams: dense low planes, sparse high planes, random tail."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(nb):
        density = min(0.5, 2.0 ** (p - nb // 2))
        bits = rng.random(n * 8) < density
        out.append(np.packbits(bits).tobytes())
    return out


def compress(host, streams, kind, latter, threads):
    nb = len(streams)
    sizes = np.array([len(s) for s in streams], np.uint32)
    inp = np.frombuffer(b"".join(streams), np.uint8).copy()
    cap = int(sizes.sum()) * 2 + 4096 * nb
    out = np.zeros(cap, np.uint8)
    osz = np.zeros(nb, np.uint32)
    stop = host.prodm_host_compress_level(kind, latter, threads, nb, inp.ctypes.data_as(C.c_void_p),
                                          sizes.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                                          C.c_uint64(cap), osz.ctypes.data_as(C.c_void_p))
    assert stop >= 0
    res, off = [], 0
    for s in osz:
        res.append(out[off:off + int(s)].tobytes())
        off += int(s)
    return res, stop


def decompress(host, stored, kind, latter, stop, start, k, raw_total):
    nb = len(stored)
    sizes = np.array([len(s) for s in stored], np.uint32)
    inp = np.frombuffer(b"".join(stored), np.uint8).copy()
    out = np.zeros(raw_total + 16, np.uint8)
    osz = np.zeros(max(k, 1), np.uint32)
    rc = host.prodm_host_decompress_level(kind, latter, 4, nb, stop, inp.ctypes.data_as(C.c_void_p),
                                          sizes.ctypes.data_as(C.c_void_p), start, k, out.ctypes.data_as(C.c_void_p),
                                          C.c_uint64(out.size), osz.ctypes.data_as(C.c_void_p))
    assert rc == 0
    res, off = [], 0
    for s in osz[:k]:
        res.append(out[off:off + int(s)].tobytes())
        off += int(s)
    return res


@pytest.mark.parametrize("kind,latter", [(0, 64), (1, 64), (1, 26), (1, 8)])
@pytest.mark.parametrize("threads", [1, 8])
def test_compress_round_trip(host, kind, latter, threads):
    streams = planes_like(32, kind * 100 + latter)
    stored, stop = compress(host, streams, kind, latter, threads)
    for a, b in ((0, 32), (0, 5), (5, 20), (25, 7)):
        got = decompress(host, stored, kind, latter, stop, a, b, sum(len(s) for s in streams))
        assert got == streams[a:a + b]


@pytest.mark.skipif(not ob.ref_available(), reason="compiled reference checker not built")
@pytest.mark.parametrize("kind,latter", [(0, 64), (1, 64), (1, 26)])
def test_compress_matches_reference(host, kind, latter):
    streams = planes_like(32, 7)
    mine, s1 = compress(host, streams, kind, latter, 8)
    ref, s2 = ob.ref_compress_level(streams, kind=kind, latter=latter)
    assert s1 == s2
    assert mine == ref
