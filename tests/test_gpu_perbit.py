"""GPU PerBit encode (prodm_perbit_encode_*) against the oracle: the
streams and their sizes are byte-identical (x86-compatible conversion,
the oracle's), for f32 / f64, one to many workgroups, ragged tails,
special values (zeros, -0, NaN, data above 2^exp); the default conversion
keeps the truncated magnitude at |v| >= 2^63 (fp64, 64 planes), checked
by decoding; the host decoder reconstructs the truncated fixed point from
the GPU's streams; PerBit level errors are GroupedBP's."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from prodm_amd import capi, mdr

pytestmark = pytest.mark.gpu

TD = {np.float32: torch.float32, np.float64: torch.float64}


def field(n, dt, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(n) * np.exp(rng.uniform(-14, 0, n))).astype(dt)
    x[rng.random(n) < 0.05] = 0
    return x


def level_exp(x):
    return ob.frexp_exp(ob.maxabs(x), x.dtype.type)


def encode_gpu(x, e, nb, compat=True):
    streams, sizes = mdr.perbit_encode(torch.from_numpy(x).cuda(), e, nb, compat)
    torch.cuda.synchronize()
    return mdr.perbit_streams(streams, sizes), sizes.cpu().numpy()


@pytest.mark.parametrize("dt,nb,n", [
    (np.float32, 32, 1), (np.float32, 32, 31), (np.float32, 32, 33), (np.float32, 32, 8191),
    (np.float32, 32, 8193), (np.float32, 17, 100003), (np.float32, 1, 5000), (np.float32, 32, 1 << 20),
    (np.float64, 64, 1), (np.float64, 64, 65), (np.float64, 64, 4097), (np.float64, 52, 100003),
    (np.float64, 3, 777), (np.float64, 64, 300000),
])
def test_perbit_encode_matches_oracle(gpu, dt, nb, n):
    x = field(n, dt, n + nb)
    e = level_exp(x)
    mine, sizes = encode_gpu(x, e, nb)
    ref = ob.pb_encode(x, e, nb)
    assert [len(r) for r in ref] == list(sizes)
    for p in range(nb):
        assert mine[p] == ref[p], (p, n, nb)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_perbit_special_values(gpu, dt):
    x = field(20000, dt, 4)
    x[:8] = [0.0, -0.0, np.nan, -np.nan, 1e-38, -1e-38, 0.5, -0.75]
    x[100:140:3] = -x[100:140:3]
    e = level_exp(np.nan_to_num(x))
    for nb in (8 * x.itemsize, 9):
        mine, _ = encode_gpu(x, e, nb)
        assert mine == ob.pb_encode(x, e, nb), nb
    # data above 2^exp: the fixed point keeps only its low nb bits; at
    # |v| >= 2^63 the conversion gives the x86 result
    y = x.copy()
    y[10:20] = [3.0, -5.0, 1e3, -1e6, 2.0 ** 40, -(2.0 ** 62), 1e30, -1e30, np.inf, -np.inf]
    for nb in (8 * x.itemsize, 20):
        mine, _ = encode_gpu(y, 0, nb)
        assert mine == ob.pb_encode(y, 0, nb), nb


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float64, 62), (np.float64, 64)])
def test_perbit_default_conversion(gpu, dt, nb):
    """Default conversion: streams equal the oracle's wherever |v| < 2^63;
    at 64 planes the level's largest elements (|v| in [2^63, 2^64)) keep
    their truncated magnitude instead of x86's 2^63."""
    x = field(30000, dt, nb)
    e = level_exp(x)
    mine, _ = encode_gpu(x, e, nb, compat=False)
    ref = ob.pb_encode(x, e, nb)
    big = np.abs(x.astype(np.float64)) * 2.0 ** (nb - e) >= 2.0 ** 63
    assert (nb == 64) == bool(big.any())
    if nb < 64:
        assert mine == ref
        return
    assert mine != ref
    y = mdr.PerBitDecoder(torch.float64).decode(0, mine, x.size, e, 0, nb)
    # trunc(|x| 2^(nb-e)) 2^(e-nb), in Python integers (the scaling is exact)
    t = np.array([float(int(abs(float(v)) * 2.0 ** (nb - e))) for v in x]) * 2.0 ** (e - nb)
    assert np.array_equal(np.abs(y), t)
    assert np.array_equal(np.abs(y[big]), np.abs(x[big]))


def test_perbit_roundtrip_large(gpu):
    """8 M elements: the host decoder reconstructs trunc(|x| 2^(nb-e)) with
    the signs from the GPU's streams (planes decoded in three steps)."""
    n, nb = 8 << 20, 32
    rng = np.random.default_rng(11)
    x = (rng.standard_normal(n) * 1e-3).astype(np.float32)
    e = level_exp(x)
    streams, _ = encode_gpu(x, e, nb)
    d = mdr.PerBitDecoder(torch.float32)
    acc = np.zeros(n, np.float64)
    for s, k in ((0, 10), (10, 12), (22, 10)):
        acc += d.decode(0, streams[s:s + k], n, e, s, k)
    t = np.trunc(np.abs(x).astype(np.float64) * 2.0 ** (nb - e)) * 2.0 ** (e - nb)
    assert np.array_equal(np.abs(acc), t)
    nz = t != 0
    assert np.array_equal(np.signbit(acc[nz]), np.signbit(x[nz]))
    # every stream holds n bits plus one sign per element whose first 1 it has
    assert sum(len(s) for s in streams) >= nb * n // 8


def test_perbit_level_errors_are_grouped(gpu):
    x = field(50000, np.float32, 8)
    e = level_exp(x)
    got = mdr.level_errors(torch.from_numpy(x).cuda(), e, 32, encoder=capi.GROUPED)
    np.testing.assert_array_equal(got, ob.fused_level_errors(x, e, 32, "gb"))


def test_perbit_encode_rejects_bad_arguments(gpu):
    lib = capi.lib()
    x = torch.zeros(100, device="cuda")
    out = torch.zeros((33, 8), dtype=torch.int64, device="cuda")
    sizes = torch.zeros(33, dtype=torch.int32, device="cuda")
    ws = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    s = mdr._stream()
    args = lambda n, nb, stride: (mdr._ptr(x), n, 0, nb, 0, mdr._ptr(out), stride, mdr._ptr(sizes), mdr._ptr(ws),
                                  ws.numel(), s)
    assert lib.prodm_perbit_encode_f32(*args(0, 8, 8)) != 0        # no elements
    assert lib.prodm_perbit_encode_f32(*args(100, 33, 8)) != 0     # more planes than bits
    assert lib.prodm_perbit_encode_f32(*args(100, 8, 3)) != 0      # stride below ceil(2 n / 64)
    assert lib.prodm_perbit_encode_f32(*args(100, 8, 4)) == 0
    torch.cuda.synchronize()


# ---- the device-resident plan with the PerBit encoder --------------------------------
def _plan_case(dt, dims, L, nb, enc, decomp, seed):
    from fields import bumps_field
    x = torch.from_numpy(bumps_field(dims, seed=seed, dtype=dt)).cuda()
    p = mdr.Plan(dims, L, nb, TD[dt], enc, decomp)
    p.refactor(x)
    torch.cuda.synchronize()
    return x, p


@pytest.mark.parametrize("dt,nb,decomp", [(np.float32, 32, 0), (np.float32, 32, 1), (np.float64, 60, 0),
                                          (np.float64, 64, 1)])
def test_plan_perbit_streams_match_oracle(gpu, dt, nb, decomp):
    """Every level's streams from the plan equal the oracle's PerBit encode
    of that level's coefficients at the level's exponent."""
    dims, L = [33, 65, 33], 4
    _, p = _plan_case(dt, dims, L, nb, capi.PERBIT, decomp, 7)
    for l in range(L + 1):
        c = p.coefficients(l).cpu().numpy()
        _, e, n = p.level_info(l)
        assert n == c.size
        ref = ob.pb_encode(c, e, nb) if nb < 64 else None
        for q in range(nb):
            s = p.plane_stream(l, q)
            assert s == p.plane_bytes(l, q)
            if ref is not None:
                assert s == ref[q], (l, q)
        if ref is None:  # 64 planes: decode every plane on the host and compare with truncation
            y = mdr.PerBitDecoder(TD[dt]).decode(0, [p.plane_stream(l, q) for q in range(nb)], c.size, e, 0, nb)
            t = np.array([float(int(abs(float(v)) * 2.0 ** (nb - e))) for v in c]) * 2.0 ** (e - nb)
            assert np.array_equal(np.abs(y), t), l


@pytest.mark.parametrize("dt,nb,decomp,dims,L", [
    (np.float32, 32, 0, [65, 33, 65], 5), (np.float32, 32, 1, [65, 33, 65], 5), (np.float64, 52, 0, [65, 33, 65], 5),
    (np.float32, 30, 0, [129, 65], 5), (np.float64, 40, 1, [1025], 6),
])
def test_plan_perbit_retrieve_equals_grouped(gpu, dt, nb, decomp, dims, L):
    """PerBit and GroupedBP keep the same magnitude bits and signs: a
    retrieval of the same planes per level is equal value for value. (Where
    no 1 bit is retrieved, GroupedBP's block sign can make a coefficient
    -0.0 and PerBit's +0.0, as in ProDM's decoders, so zeros may differ in
    sign.)"""
    x, pp = _plan_case(dt, dims, L, nb, capi.PERBIT, decomp, 11)
    _, pg = _plan_case(dt, dims, L, nb, capi.GROUPED, decomp, 11)
    for k in ([nb] * (L + 1), [min(nb, 3 + 4 * l) for l in range(L + 1)], [0] + [nb // 2] * L):
        a = pp.retrieve(k)
        b = pg.retrieve(k)
        torch.cuda.synchronize()
        assert torch.equal(a, b), k


def test_plan_perbit_load_streams_and_errors(gpu):
    """Streams moved out and loaded into a fresh plan retrieve the same
    field; level errors are GroupedBP's; graph retrieval is refused."""
    dims, L, nb = [65, 65, 33], 4, 32
    x, p = _plan_case(np.float32, dims, L, nb, capi.PERBIT, 0, 3)
    _, pg = _plan_case(np.float32, dims, L, nb, capi.GROUPED, 0, 3)
    for l in range(L + 1):  # before any retrieval (which decodes into the level buffers)
        np.testing.assert_array_equal(p.level_errors(l), pg.level_errors(l))
    q = mdr.Plan(dims, L, nb, torch.float32, capi.PERBIT)
    for l in range(L + 1):
        mx, _, _ = p.level_info(l)
        q.set_level_max(l, mx)
        for j in range(nb):
            q.load_plane_stream(l, j, p.plane_stream(l, j))
    k = [nb, 30, 24, 18, 12]
    a, b = p.retrieve(k), q.retrieve(k)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    out = torch.empty_like(x)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with pytest.raises(capi.ProdmError):
            p.retrieve_graph(k, out)
    with pytest.raises(capi.ProdmError):
        q.load_plane_stream(0, 0, b"\0" * 12)  # not whole 64-bit words


def test_batched_perbit_equals_plain(gpu):
    from prodm_amd.batched import Batch
    from fields import bumps_field
    shapes = [([65, 65, 33], 4), ([129, 65], 5), ([33, 33, 33], 3)]
    xs = [torch.from_numpy(bumps_field(d, seed=50 + i, dtype=np.float32)).cuda() for i, (d, L) in enumerate(shapes)]
    b = Batch([(x, d, L) for x, (d, L) in zip(xs, shapes)], 32, torch.float32, capi.PERBIT)
    b.refactor()
    torch.cuda.synchronize()
    for f, (d, L) in enumerate(shapes):
        p = mdr.Plan(d, L, 32, torch.float32, capi.PERBIT)
        p.refactor(xs[f])
        torch.cuda.synchronize()
        for l in range(L + 1):
            for q in range(32):
                assert b.plan(f).plane_stream(l, q) == p.plane_stream(l, q), (f, l, q)
