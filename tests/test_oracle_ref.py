"""The CPU oracle against the compiled ProDM reference (oracle/_ref/libref.so,
built in place from the reference headers by `make -C oracle ref`) over a
seeded sweep of sizes, plane counts and exponents. Skipped where the
reference sources are absent; the golden fixtures cover that case.
"""
import numpy as np
import pytest

import oracle_bind as ob

pytestmark = pytest.mark.skipif(not ob.ref_available(), reason="compiled reference checker not built")


def x_of(n, dt, seed, scale=0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp2(rng.integers(-16, 1, n) + scale)
    x[rng.random(n) < 0.07] = 0
    return x.astype(dt)


@pytest.mark.parametrize("dims,L", [([513, 257, 129], 6), ([100], 3), ([9, 9, 9], 2), ([1, 1, 65], 5)])
def test_geometry(dims, L):
    np.testing.assert_array_equal(ob.level_dims(dims, L), ob.level_dims(dims, L, lib=ob.ref()))


@pytest.mark.parametrize("seed", range(6))
def test_interleave_sweep(seed):
    rng = np.random.default_rng(seed)
    dims = [int(2 ** rng.integers(2, 5) + 1) for _ in range(3)]
    L = int(rng.integers(1, 3))
    dt = np.float32 if seed % 2 else np.float64
    x = rng.standard_normal(int(np.prod(dims))).astype(dt)
    ld = ob.level_dims(dims, L)
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(3, np.uint32)
        np.testing.assert_array_equal(ob.interleave(x, dims, ld[l], coarse),
                                      ob.interleave(x, dims, ld[l], coarse, lib="ref"))
        buf = ob.interleave(x, dims, ld[l], coarse)
        a = ob.reposition(buf, np.zeros_like(x), dims, ld[l], coarse)
        b = ob.reposition(buf, np.zeros_like(x), dims, ld[l], coarse, lib="ref")
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_maxabs(dt):
    x = x_of(5000, dt, 1)
    assert ob.maxabs(x) == ob.maxabs(x, lib="ref")


NB = [(np.float32, 32), (np.float32, 8), (np.float64, 64), (np.float64, 52), (np.float64, 2)]


@pytest.mark.parametrize("dt,nb", NB)
@pytest.mark.parametrize("n", [32, 33, 64, 127, 4099])
@pytest.mark.parametrize("scale", [0, -30, 20])
def test_negabinary(dt, nb, n, scale):
    if n < (32 if dt == np.float32 else 64):
        pytest.skip("the reference's block loop wraps below one block (i < n - block_size, unsigned)")
    x = x_of(n, dt, n + nb, scale)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    p1 = ob.nb_encode(x, e, nb)
    np.testing.assert_array_equal(p1, ob.nb_encode(x, e, nb, lib="ref"))
    for s, k in ((0, nb), (0, 1), (1, nb // 2 - 1), (nb // 2, nb - nb // 2)):
        if k <= 0:
            continue
        a = ob.nb_decode(p1[s:s + k], n, e, s, k, dt)
        b = ob.nb_decode(p1[s:s + k], n, e, s, k, dt, lib="ref")
        np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))


GB = [(np.float32, 32), (np.float32, 5), (np.float64, 63), (np.float64, 64), (np.float64, 30)]


@pytest.mark.parametrize("dt,nb", GB)
@pytest.mark.parametrize("n", [32, 64, 65, 1000, 4097])
@pytest.mark.parametrize("scale", [0, -40])
def test_grouped(dt, nb, n, scale):
    if n < (32 if dt == np.float32 else 64):
        pytest.skip("the reference's block loop wraps below one block (i < n - block_size, unsigned)")
    x = x_of(n, dt, 3 * n + nb, scale)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    compat = dt == np.float64 and nb == 64
    s1 = ob.gb_encode(x, e, nb, compat=compat)
    s2 = ob.gb_encode(x, e, nb, lib="ref")
    assert s1 == s2
    d1, d2 = ob.GBDecoder(dt), ob.GBDecoder(dt, lib="ref")
    cuts = sorted({0, 1, nb // 3, nb // 2, nb})
    for a, b in zip(cuts[:-1], cuts[1:]):
        r1 = d1.decode(0, s1[a:b], n, e, a, b - a)
        r2 = d2.decode(0, s2[a:b], n, e, a, b - a)
        np.testing.assert_array_equal(r1.view(np.uint8), r2.view(np.uint8))


@pytest.mark.parametrize("dt,nb,n", [(np.float32, 32, 3000), (np.float32, 10, 64), (np.float64, 64, 2500),
                                     (np.float64, 52, 7), (np.float32, 31, 1)])
def test_perbit(dt, nb, n):
    x = x_of(n, dt, n + nb)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    a, b = ob.pb_encode(x, e, nb), ob.pb_encode(x, e, nb, lib="ref")
    assert a == b
    d1, d2 = ob.PBDecoder(dt), ob.PBDecoder(dt, lib="ref")
    cuts = sorted({0, 1, nb // 3, nb // 2, nb})
    for s, t in zip(cuts[:-1], cuts[1:]):
        r1 = d1.decode(2, a[s:t], n, e, s, t - s)
        r2 = d2.decode(0, b[s:t], n, e, s, t - s)
        np.testing.assert_array_equal(r1.view(np.uint8), r2.view(np.uint8))


@pytest.mark.parametrize("n,nb", [(2048, 32), (999, 24), (32, 2), (5000, 31)])
@pytest.mark.parametrize("enc", ["nb", "gb"])
def test_fused_level_errors_sequential(n, nb, enc):
    x = x_of(n, np.float32, n + nb)
    e = ob.frexp_exp(ob.maxabs(x), np.float32)
    np.testing.assert_array_equal(ob.fused_level_errors(x, e, nb, enc, mode="seq"),
                                  ob.fused_level_errors(x, e, nb, enc, mode="ref"))



@pytest.mark.parametrize("n,nb", [(2048, 32), (999, 24), (5000, 31)])
def test_perbit_level_errors_are_groupedbp_level_errors(n, nb):
    """PerBitBPEncoder's level-error overload has GroupedBP's per-element
    terms in the same element order: the reference's two tables agree."""
    x = x_of(n, np.float32, 3 * n + nb)
    e = ob.frexp_exp(ob.maxabs(x), np.float32)
    np.testing.assert_array_equal(ob.fused_level_errors(x, e, nb, "pb", mode="ref"),
                                  ob.fused_level_errors(x, e, nb, "gb", mode="ref"))

@pytest.mark.parametrize("n,nb", [(1000, 32), (777, 16), (64, 30)])
def test_squared_sequential(n, nb):
    x = x_of(n, np.float32, n)
    m = np.float32(ob.maxabs(x))
    np.testing.assert_array_equal(ob.collect_squared(x, nb, m, mode="seq"), ob.collect_squared(x, nb, m, mode="ref"))


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float64, 64)])
def test_collect_max(dt, nb):
    for b in (1.0, 2.5, 1e-7):
        np.testing.assert_array_equal(ob.collect_max(dt(b), nb, dt), ob.collect_max(dt(b), nb, dt, lib="ref"))
