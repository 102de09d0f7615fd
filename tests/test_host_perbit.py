"""The library's host PerBit decoder (prodm_perbit_decode_*) against the
oracle's progressive decode on oracle-encoded streams: bit-identical
outputs and sign state over progressive calls, zero-plane calls, and the
decoder's argument errors (short streams, level size changes)."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from prodm_amd import capi, mdr

TD = {np.float32: torch.float32, np.float64: torch.float64}


def field(n, dt, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-12, 0, n))).astype(dt)
    x[rng.random(n) < 0.1] = 0
    x[: min(n, 3)] = [0.0, -0.0, -1e-30][: min(n, 3)]
    return x


def level_exp(x):
    m = float(np.max(np.abs(x))) if x.size else 0.0
    return int(np.frexp(m)[1])


@pytest.mark.parametrize("dt,nb,n,steps", [
    (np.float32, 32, 1000, [32]),
    (np.float32, 32, 4099, [3, 5, 0, 11, 13]),
    (np.float32, 20, 1, [1, 19]),
    (np.float64, 64, 777, [64]),
    (np.float64, 60, 5000, [7, 0, 21, 32]),
    (np.float64, 40, 65, [40]),
])
def test_host_decoder_matches_oracle(dt, nb, n, steps):
    x = field(n, dt, nb + n)
    e = level_exp(x)
    streams = ob.pb_encode(x, e, nb)
    mine = mdr.PerBitDecoder(TD[dt])
    ref = ob.PBDecoder(dt)
    start = 0
    for k in steps:
        a = mine.decode(3, streams[start:start + k], n, e, start, k)
        b = ref.decode(0, streams[start:start + k], n, e, start, k)
        assert a.tobytes() == b.tobytes(), (start, k)
        start += k
    mine.close()


def test_host_decoder_levels_are_independent():
    xs = [field(n, np.float32, 7 + n) for n in (300, 1200)]
    es = [level_exp(x) for x in xs]
    st = [ob.pb_encode(x, e, 32) for x, e in zip(xs, es)]
    mine = mdr.PerBitDecoder(torch.float32)
    refs = [ob.PBDecoder(np.float32), ob.PBDecoder(np.float32)]
    for start, k in ((0, 6), (6, 10), (16, 16)):
        for lv in (1, 0):
            n = xs[lv].size
            a = mine.decode(lv, st[lv][start:start + k], n, es[lv], start, k)
            b = refs[lv].decode(0, st[lv][start:start + k], n, es[lv], start, k)
            assert a.tobytes() == b.tobytes(), (lv, start, k)


def test_host_decoder_full_planes_reconstruct_truncation():
    # all planes: |y| = trunc(|x| 2^(nb-e)) 2^(e-nb), sign of x (0 for zeros)
    x = field(2000, np.float64, 5)
    e = level_exp(x)
    y = mdr.PerBitDecoder(torch.float64).decode(0, ob.pb_encode(x, e, 52), x.size, e, 0, 52)
    t = np.trunc(np.abs(x) * 2.0 ** (52 - e)) * 2.0 ** (e - 52)
    assert np.array_equal(np.abs(y), t)
    nz = t != 0
    assert np.array_equal(np.signbit(y[nz]), np.signbit(x[nz]))


def test_host_decoder_errors():
    x = field(500, np.float32, 9)
    e = level_exp(x)
    streams = ob.pb_encode(x, e, 16)
    d = mdr.PerBitDecoder(torch.float32)
    with pytest.raises(capi.ProdmError, match="shorter"):
        d.decode(0, [streams[0][:-8]] + streams[1:4], 500, e, 0, 4)
    d2 = mdr.PerBitDecoder(torch.float32)
    d2.decode(0, streams[:2], 500, e, 0, 2)
    with pytest.raises(capi.ProdmError, match="elements"):
        d2.decode(0, streams[2:4], 400, e, 2, 2)
    with pytest.raises(capi.ProdmError, match="plane range"):
        d2.decode(1, streams[:2], 500, e, 31, 2)
    assert not np.any(d2.decode(5, [], 500, e, 0, 0))
