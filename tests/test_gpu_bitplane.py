"""GPU parity: NegaBinary / GroupedBP encode + progressive decode, max-abs,
interleave, collectors -- bit-exact against the CPU oracle on the same inputs."""
import numpy as np
import pytest
import torch

import oracle_bind as ob

pytestmark = pytest.mark.gpu

NS = [1, 5, 31, 32, 33, 63, 64, 65, 100, 1000, 4095, 4096, 4097, 20000, 65537]
TD = {np.float32: torch.float32, np.float64: torch.float64}
UT = {np.float32: np.uint32, np.float64: np.uint64}


def make_data(n, dt, seed, zeros=0.3):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(n) * rng.choice([1e-6, 1e-3, 1.0, 30.0], size=n)).astype(dt)
    x[rng.random(n) < zeros] = 0
    if n > 3:
        x[1] = -0.0
    return x


def bits(a):
    return a.view(UT[a.dtype.type])


def splits(nb):
    return [[nb], [1, nb - 1], [3, 5, nb - 8] if nb > 8 else [1, nb - 1], [2] * (nb // 2)]


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float32, 16), (np.float32, 2), (np.float64, 64),
                                   (np.float64, 40)])
@pytest.mark.parametrize("n", NS)
def test_negabinary_roundtrip(gpu, dt, nb, n):
    from prodm_amd import mdr
    x = make_data(n, dt, 1000 + n)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    ref = ob.nb_encode(x, e, nb)
    dev = mdr.negabinary_encode(torch.from_numpy(x).to(gpu), e, nb)
    got = mdr.planes_to_numpy(dev, n, TD[dt])
    assert np.array_equal(got, ref)
    for sp in splits(nb):
        start = 0
        for k in sp:
            rows = [dev[start + i] for i in range(k)]
            g = mdr.negabinary_decode(rows, n, e, start, k, TD[dt]).cpu().numpy()
            o = ob.nb_decode(ref[start:], n, e, start, k, dt)
            assert np.array_equal(bits(g), bits(o)), (sp, start, k)
            start += k


@pytest.mark.parametrize("dt,nb,compat", [(np.float32, 32, False), (np.float32, 13, False), (np.float64, 64, False),
                                          (np.float64, 64, True), (np.float64, 60, False)])
@pytest.mark.parametrize("n", NS)
def test_grouped_roundtrip(gpu, dt, nb, compat, n):
    from prodm_amd import mdr
    x = make_data(n, dt, 2000 + n)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    ref = ob.gb_encode(x, e, nb, compat=compat)
    rb, planes, words = mdr.grouped_encode(torch.from_numpy(x).to(gpu), e, nb, compat=compat)
    got = mdr.grouped_streams(rb, planes, words, TD[dt])
    assert len(got) == nb
    for p in range(nb):
        assert got[p] == ref[p], f"plane {p}"
    dev_streams = [torch.frombuffer(bytearray(s) if s else bytearray(8), dtype=torch.uint8).to(gpu) for s in ref]
    for sp in splits(nb):
        dec = mdr.GroupedDecoder(TD[dt])
        odec = ob.GBDecoder(dt)
        start = 0
        for k in sp:
            g = dec.decode(3, dev_streams[start:start + k], n, e, start, k).cpu().numpy()
            o = odec.decode(3, ref[start:start + k], n, e, start, k)
            assert np.array_equal(bits(g), bits(o)), (sp, start, k)
            start += k
        dec.close()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_maxabs_and_collect_max(gpu, dt):
    from prodm_amd import mdr
    x = make_data(100003, dt, 5)
    x[17] = np.nan
    m = mdr.maxabs(torch.from_numpy(x).to(gpu))
    assert m == ob.maxabs(x)
    for b in [1.5, 3e-5, 1024.0, m]:
        assert np.array_equal(mdr.collect_max(b, 32, TD[dt]), ob.collect_max(b, 32, dt))


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float32, 16), (np.float64, 52)])
@pytest.mark.parametrize("n", [1, 1000, 65537])
def test_squared_collector_exact(gpu, dt, nb, n):
    from prodm_amd import mdr
    x = make_data(n, dt, 77 + n, zeros=0.1)
    if n > 10:
        x[3] = np.float32(1e-30) if dt == np.float32 else 1e-200  # far below the level max
    m = ob.maxabs(x)
    got = mdr.collect_squared(torch.from_numpy(x).to(gpu), nb, m)
    exp = ob.collect_squared(x, nb, m)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("dims,L", [([33], 3), ([17, 33], 3), ([17, 9, 33], 2), ([65, 65, 65], 4)])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_interleave_reposition(gpu, dims, L, dt):
    from prodm_amd import mdr
    rng = np.random.default_rng(3)
    d = rng.standard_normal(int(np.prod(dims))).astype(dt)
    ld = ob.level_dims(dims, L)
    dd = torch.from_numpy(d).to(gpu)
    back = torch.zeros_like(dd)
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(len(dims), np.uint32)
        g = mdr.interleave(dd, dims, ld[l], coarse).cpu().numpy()
        o = ob.interleave(d, dims, ld[l], coarse)
        assert np.array_equal(g, o)
        mdr.reposition(torch.from_numpy(o).to(gpu), back, dims, ld[l], coarse)
    assert torch.equal(back, dd)
