"""Pins of the CPU oracle by what the mathematics fixes, independent of the
oracle's own formulas:

* MGARD orthogonal decomposition (the arithmetic ProDM takes from MGARDx,
  Decomposer/MGARD.hpp:14-35, absent from the sources): after L steps the
  coarsest nodal values are the L2 projection of the fine piecewise
  multilinear field, x = (P^T M P)^-1 P^T M u, with P the assembled
  prolongation and M the assembled fine mass matrix (dense FP64 solve);
  coefficients are u - I u_c at the coefficient nodes; multilinear fields
  give zero coefficients; recompose inverts decompose; the adjoint of
  recompose equals the dense transpose of the assembled recompose matrix.
* interleave/reposition: brute force over the box difference.
* codecs: truncation identities, stream sizes from an independent count of
  recording planes, hand-built single-bit blocks.
* squared collector: exact rational sum of the per-element terms, and the
  table as the squared error of sign-magnitude truncation (fp32 and fp64).
"""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_bind as ob

# ---------------------------------------------------------------- MGARD -----


def prolong_1d(nf):
    """Coarse -> fine linear interpolation on n_f = 2m+1 nodes (n_f = 1: identity)."""
    if nf == 1:
        return np.ones((1, 1))
    nc = (nf + 1) // 2
    P = np.zeros((nf, nc))
    for j in range(nc):
        P[2 * j, j] = 1.0
    for i in range(nc - 1):
        P[2 * i + 1, i] = P[2 * i + 1, i + 1] = 0.5
    return P


def mass_1d(n):
    """Linear finite-element mass matrix, unit spacing (n = 1: [1])."""
    if n == 1:
        return np.ones((1, 1))
    M = np.zeros((n, n))
    for i in range(n - 1):
        M[i, i] += 1 / 3
        M[i + 1, i + 1] += 1 / 3
        M[i, i + 1] += 1 / 6
        M[i + 1, i] += 1 / 6
    return M


def kron_all(mats):
    out = np.ones((1, 1))
    for m in mats:
        out = np.kron(out, m)
    return out


def pad3(dims):
    return [1] * (3 - len(dims)) + list(dims)


def l2_projection(u, dims, L):
    """Nodal values on the level-0 grid of the L2 projection of u (dense)."""
    ld = ob.level_dims(dims, L)
    Ptot = np.eye(int(np.prod(dims)))
    for l in range(L, 0, -1):
        Pl = kron_all([prolong_1d(int(n)) for n in ld[l]])
        Ptot = Ptot @ Pl
    M = kron_all([mass_1d(int(n)) for n in dims])
    A = Ptot.T @ M @ Ptot
    return np.linalg.solve(A, Ptot.T @ (M @ u)), ld[0]


CASES = [([9, 17, 5], 1), ([9, 17, 5], 2), ([17, 9], 3), ([33], 4), ([1, 9, 17], 2), ([5, 5, 9], 2)]


@pytest.mark.parametrize("dims,L", CASES)
def test_coarse_values_are_l2_projection(dims, L):
    rng = np.random.default_rng(1)
    u = rng.standard_normal(int(np.prod(dims)))
    got = ob.decompose(u, dims, L).reshape(dims)
    x, c = l2_projection(u, dims, L)
    box = got[tuple(slice(0, int(n)) for n in c)].ravel()
    np.testing.assert_allclose(box, x, rtol=0, atol=1e-12 * np.max(np.abs(x)))


@pytest.mark.parametrize("dims", [[9, 17, 5], [17, 9], [33], [1, 9, 5]])
def test_coefficients_are_interpolation_residuals(dims):
    """One step: coefficient nodes hold u - I u_c, coarse nodes first along
    each dim (DirectInterleaver.hpp:23-30 reads them as the box difference)."""
    rng = np.random.default_rng(2)
    u = rng.standard_normal(int(np.prod(dims)))
    got = ob.decompose(u, dims, 1).reshape(dims)
    P = kron_all([prolong_1d(n) for n in dims])
    U = u.reshape(dims)
    coarse_idx = [np.arange(0, n, 2) if n > 1 else np.arange(1) for n in dims]
    uc = U[np.ix_(*coarse_idx)].ravel()
    W = (u - P @ uc).reshape(dims)

    def rpos(i, n):
        if n == 1:
            return 0
        m = (n - 1) // 2
        return i // 2 if i % 2 == 0 else m + 1 + i // 2

    for idx in itertools.product(*[range(n) for n in dims]):
        if all(i % 2 == 0 for i in idx):
            continue
        pos = tuple(rpos(i, n) for i, n in zip(idx, dims))
        assert abs(got[pos] - W[idx]) <= 1e-13, (idx, got[pos], W[idx])


@pytest.mark.parametrize("dims,L", [([17, 9, 33], 2), ([33, 17], 3), ([65], 5)])
def test_multilinear_fields_have_zero_coefficients(dims, L):
    axes = [np.linspace(0, 1, n) for n in dims]
    g = np.meshgrid(*axes, indexing="ij")
    u = 0.3 + np.zeros(dims)
    for d in range(len(dims)):
        u = u + (d + 1) * g[d]
    if len(dims) > 1:
        u = u + 2.5 * g[0] * g[1]
    if len(dims) > 2:
        u = u - 1.5 * g[0] * g[1] * g[2] + 0.7 * g[1] * g[2]
    out = ob.decompose(u.ravel(), dims, L).reshape(dims)
    ld = ob.level_dims(dims, L)
    mask = np.ones(dims, bool)
    mask[tuple(slice(0, int(n)) for n in ld[0])] = False
    assert np.max(np.abs(out[mask])) <= 1e-12
    # the level-0 grid holds the field's values at its nodes
    step = [(n - 1) // (c - 1) if c > 1 else 1 for n, c in zip(dims, ld[0])]
    want = u[tuple(slice(0, None, s) for s in step)]
    np.testing.assert_allclose(out[tuple(slice(0, int(n)) for n in ld[0])], want, atol=1e-12)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 1e-6)])
@pytest.mark.parametrize("dims,L", [([33, 17, 65], 3), ([129, 33], 4), ([257], 7), ([1, 33, 33], 4)])
def test_recompose_inverts_decompose(dtype, tol, dims, L):
    rng = np.random.default_rng(3)
    u = rng.standard_normal(int(np.prod(dims))).astype(dtype)
    v = ob.recompose(ob.decompose(u, dims, L), dims, L)
    assert np.max(np.abs(v.astype(np.float64) - u)) <= tol * np.max(np.abs(u))


def test_strided_box_matches_dense_box():
    """Decomposing a box inside a larger array (full strides) equals
    decomposing the box alone, and leaves the rest of the array untouched."""
    rng = np.random.default_rng(4)
    full = [17, 9, 33]
    u = rng.standard_normal(int(np.prod(full)))
    box = [9, 5, 17]
    strides = [9 * 33, 33, 1]
    got = ob.decompose(u, box, 2, strides=strides).reshape(full)
    U = u.reshape(full)
    dense = ob.decompose(np.ascontiguousarray(U[:9, :5, :17]).ravel(), box, 2).reshape(box)
    np.testing.assert_array_equal(got[:9, :5, :17], dense)
    mask = np.ones(full, bool)
    mask[:9, :5, :17] = False
    np.testing.assert_array_equal(got[mask], U[mask])


@pytest.mark.parametrize("dims,L", [([5, 9, 5], 2), ([9, 17], 3), ([17], 4)])
def test_recompose_adjoint_is_dense_transpose(dims, L):
    N = int(np.prod(dims))
    A = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        A[:, j] = ob.recompose(e, dims, L)
    rng = np.random.default_rng(5)
    for _ in range(3):
        g = rng.standard_normal(N)
        np.testing.assert_allclose(ob.recompose_adjoint(g, dims, L), A.T @ g, rtol=0, atol=1e-12 * np.abs(g).sum())


# ----------------------------------------------------------- geometry -------


@pytest.mark.parametrize("dims,L", [([513, 513, 513], 6), ([17, 9, 33], 2), ([100, 37, 8], 2), ([65], 5)])
def test_level_elements_box_count(dims, L):
    ld = ob.level_dims(dims, L)
    for l in range(1, L + 1):
        for d in range(len(dims)):
            assert ld[l - 1][d] == (ld[l][d] >> 1) + 1
    el = ob.level_elements(dims, L)
    prev = 0
    for l in range(L + 1):
        box = int(np.prod(ld[l].astype(np.int64)))
        assert int(el[l]) == box - prev
        prev = box


# ----------------------------------------------------------- interleave -----


@pytest.mark.parametrize("dims,L", [([9, 17, 5], 2), ([17, 33], 3), ([33], 4)])
def test_interleave_brute_force(dims, L):
    rng = np.random.default_rng(6)
    x = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    X = x.reshape(dims)
    ld = ob.level_dims(dims, L)
    y = np.zeros_like(x)
    total = 0
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(len(dims), np.uint32)
        want = [X[idx] for idx in itertools.product(*[range(int(n)) for n in ld[l]])
                if not all(i < int(c) for i, c in zip(idx, coarse))]
        got = ob.interleave(x, dims, ld[l], coarse)
        np.testing.assert_array_equal(got, np.array(want, np.float32))
        ob.reposition(got, y, dims, ld[l], coarse)
        total += got.size
    assert total == x.size
    np.testing.assert_array_equal(y, x)


# ----------------------------------------------------------- max ---------


def test_maxabs_and_exponent():
    rng = np.random.default_rng(7)
    for dt in (np.float32, np.float64):
        x = (rng.standard_normal(1001) * 3).astype(dt)
        m = ob.maxabs(x)
        assert m == np.max(np.abs(x))
        e = ob.frexp_exp(m, dt)
        assert 2.0 ** (e - 1) <= m < 2.0 ** e
        assert e == math.frexp(float(m))[1]
        x[5] = np.nan  # NaN never wins the '>' comparison
        assert ob.maxabs(x) == m


def test_collect_max_closed_form():
    for dt, nb in ((np.float32, 32), (np.float64, 64)):
        for b in (1.0, 0.75, 3.0, 1e-20):
            t = ob.collect_max(dt(b), nb, dt)
            e = math.frexp(float(dt(b)))[1]
            assert t[0] == float(dt(b))
            for k in range(1, nb + 1):
                assert t[k] == 2.0 ** (e - k)


# ----------------------------------------------------------- codecs -------


def codec_x(n, dt, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp2(rng.integers(-10, 1, n))
    x[rng.random(n) < 0.1] = 0
    return x.astype(dt)


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float32, 16), (np.float64, 64), (np.float64, 40)])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 64, 65, 1000])
def test_negabinary_truncation_identity(dt, nb, n):
    x = codec_x(n, dt, n)
    e = ob.frexp_exp(ob.maxabs(x), dt) if np.any(x) else 0
    planes = ob.nb_encode(x, e, nb)
    # nb even: decoding all planes gives trunc(x 2^(nb-e-2)) 2^(e+2-nb)
    q = np.trunc(np.ldexp(x.astype(np.float64), nb - e - 2)).astype(np.int64)
    want = np.ldexp(q.astype(dt), e + 2 - nb).astype(dt)
    got = ob.nb_decode(planes, n, e, 0, nb, dt)
    np.testing.assert_array_equal(got, want)
    # progressive deltas over a split add up to the one-shot result
    a = nb // 3
    d1 = ob.nb_decode(planes[:a], n, e, 0, a, dt).astype(np.float64)
    d2 = ob.nb_decode(planes[a:], n, e, a, nb - a, dt).astype(np.float64)
    np.testing.assert_allclose(d1 + d2, got, rtol=2.0 ** -20, atol=2.0 ** (e - nb + 3))


def gb_rb(x, e, nb, B):
    """Recording plane of each block: nb - bit length of the block's largest magnitude."""
    mag = np.trunc(np.abs(np.ldexp(x.astype(np.float64), nb - e))).astype(np.uint64)
    nblk = (x.size + B - 1) // B
    rb = np.full(nblk, nb, np.int64)
    for b in range(nblk):
        m = int(mag[b * B:(b + 1) * B].max())
        if m:
            rb[b] = nb - m.bit_length()
    return rb


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float32, 12), (np.float64, 60), (np.float64, 64)])
@pytest.mark.parametrize("n", [1, 31, 32, 33, 64, 65, 1000])
def test_grouped_sizes_and_truncation(dt, nb, n):
    B = 32 if dt == np.float32 else 64
    S = B // 8
    x = codec_x(n, dt, 100 + n)
    e = ob.frexp_exp(ob.maxabs(x), dt) if np.any(x) else 0
    streams = ob.gb_encode(x, e, nb)
    rb = gb_rb(x, e, nb, B)
    nblk = rb.size
    for p in range(nb):
        want = S * (int(np.sum(rb <= p)) + int(np.sum(rb == p))) + (4 + nblk if p == 0 else 0)
        assert len(streams[p]) == want, p
    hdr = np.frombuffer(streams[0][:4], np.uint32)[0]
    assert hdr == nblk
    np.testing.assert_array_equal(np.frombuffer(streams[0][4:4 + nblk], np.uint8), rb.astype(np.uint8))
    got = ob.GBDecoder(dt).decode(0, streams, n, e, 0, nb)
    mag = np.trunc(np.abs(np.ldexp(x.astype(np.float64), nb - e)))
    want = (np.sign(x) * np.ldexp(mag, e - nb)).astype(dt)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_single_bit_blocks(dt):
    """One non-zero element per block, at every position: its plane word has
    exactly that bit, the block's sign word is set for negatives."""
    B = 32 if dt == np.float32 else 64
    nb = 24
    x = np.zeros(B * B, dt)
    for b in range(B):
        x[b * B + b] = (-1.0 if b % 2 else 1.0) * 0.5 ** (b % 7)
    x[0] = 1.0
    e = ob.frexp_exp(ob.maxabs(x), dt)
    streams = ob.gb_encode(x, e, nb)
    words = [np.frombuffer(s[(4 + B) if p == 0 else 0:], np.uint32 if B == 32 else np.uint64)
             for p, s in enumerate(streams)]
    pos = [0] * nb
    for b in range(B):
        k = b % 7  # magnitude 2^-k -> first set plane k + (nb - e) - nb ... with e = 1: plane k
        rb = k
        for p in range(rb, nb):
            if p == rb:
                sign = words[p][pos[p]]
                pos[p] += 1
                assert int(sign) == ((1 << b) if b % 2 else 0)
            w = int(words[p][pos[p]])
            pos[p] += 1
            assert w == ((1 << b) if p == rb else 0), (b, p)
    assert pos == [len(w) for w in words]


# ----------------------------------------------------------- squared --------


@pytest.mark.parametrize("nb", [32, 20])
def test_squared_collector_is_exactly_rounded(nb):
    """Each table entry of the exact mode is the correctly rounded sum of the
    per-element terms (a term = the sequential table of that element alone)."""
    x = codec_x(300, np.float32, 9)
    m = np.float32(np.max(np.abs(x)))
    terms = [ob.collect_squared(x[i:i + 1], nb, m, mode="seq") for i in range(x.size)]
    want = [float(sum((Fraction(float(t[k])) for t in terms), Fraction(0))) for k in range(nb + 1)]
    got = ob.collect_squared(x, nb, m, mode="exact")
    np.testing.assert_array_equal(got, np.array(want))


def truncation_table(x, maxerr, nb, dt):
    """SquaredErrorCollector's table from what it measures, without its
    mantissa-bit loop (SquaredErrorCollector.hpp:23-65): entry k is the sum
    over the elements of the squared error of sign-magnitude truncation to
    the weight 2^(le - k), le the exponent of maxerr (maxerr in
    [2^(le-1), 2^le)): T_k(x) = sign(x) floor(|x| 2^(k - le)) 2^(le - k).
    The residual d = x - T_k(x) is exact in the data type; its square is
    rounded once in the data type (the collector squares in T) and the terms
    are summed exactly, then rounded to double. Normal, finite x only."""
    le = math.frexp(float(maxerr))[1]
    tot = [Fraction(0)] * (nb + 1)
    for v in x:
        f = Fraction(float(v))
        if f == 0:
            continue
        a = abs(f)
        for k in range(nb + 1):
            s = Fraction(2) ** (k - le)
            d = a - Fraction(math.floor(a * s)) / s
            dT = dt(float(d))
            assert Fraction(float(dT)) == d  # the residual is exact in T
            tot[k] += Fraction(float(dT * dT))
    return np.array([float(t) for t in tot])


def squared_x(n, dt, seed, spread):
    """Normal numbers of both signs spread over 2^-spread .. 1, zeros and
    exact powers of two among them."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.5, 1.0, n) * np.exp2(-rng.integers(0, spread, n)) * rng.choice([-1.0, 1.0], n)
    x[rng.random(n) < 0.05] = 0
    x[rng.random(n) < 0.05] = np.exp2(-rng.integers(0, spread, 1))[0]
    return x.astype(dt)


@pytest.mark.parametrize("dt,nb,spread", [(np.float64, 64, 70), (np.float64, 60, 40), (np.float64, 40, 90),
                                          (np.float64, 20, 12), (np.float32, 32, 40), (np.float32, 16, 10)])
def test_squared_collector_is_truncation_error(dt, nb, spread):
    """The collector's table (exact and sequential-order sums) against the
    truncation closed form. For fp64 it pins the 64-bit mantissa mask that
    the oracle uses where the reference shifts a 32-bit 1u by up to 51 (B3):
    a 32-bit mask would leave bits 32..51 set and miss those residuals.
    Elements far below 2^(le - nb) (spread > nb) give x^2 to every entry."""
    x = squared_x(200, dt, nb + spread, spread)
    m = dt(np.max(np.abs(x)))
    want = truncation_table(x, m, nb, dt)
    np.testing.assert_array_equal(ob.collect_squared(x, nb, m, mode="exact"), want)
    seq = ob.collect_squared(x, nb, m, mode="seq")
    np.testing.assert_allclose(seq, want, rtol=x.size * 2.0 ** -52, atol=0)
    # the sequential mode adds the float terms in element order, in double
    terms = [truncation_table(x[i:i + 1], m, nb, dt) for i in range(0, x.size, 9)]
    seq9 = ob.collect_squared(np.ascontiguousarray(x[::9]), nb, m, mode="seq")
    acc = np.zeros(nb + 1)
    for t in terms:
        acc = acc + t
    np.testing.assert_array_equal(seq9, acc)


def test_squared_collector_fp64_kat():
    """One element, by hand: x = 1 + 2^-40 (bit 12 of the 52-bit mantissa
    field set), maxerr = x (le = 1), nb = 52: T_k keeps
    weights >= 2^(1-k). Entries k <= 40 hold (2^-40)^2 = 2^-80 (the
    implicit 1 is kept from k = 1 on, and k = 0 keeps nothing: x^2);
    entries k >= 41 hold 0, since 2^-40 = 2^(1-41) is kept from k = 41."""
    x = np.array([1.0 + 2.0 ** -40])
    t = ob.collect_squared(x, 52, x[0], mode="exact")
    assert t[0] == float(Fraction(1 + Fraction(1, 2 ** 40)) ** 2)
    assert all(t[k] == 2.0 ** -80 for k in range(1, 41))
    assert all(t[k] == 0.0 for k in range(41, 53))


# ---------------------------------------------------------------- PerBit ---


def test_perbit_single_element_kat():
    """-1.75 with exponent 1 and 4 planes: magnitude trunc(1.75 * 8) = 14 =
    0b1110; stream 0 holds its leading 1 and then the sign 1 (bits LSB
    first: 0b11), streams 1 and 2 a 1, stream 3 a 0; one word each."""
    x = np.array([-1.75], np.float32)
    s = ob.pb_encode(x, 1, 4)
    words = [int(np.frombuffer(b, np.uint64)[0]) for b in s]
    assert words == [0b11, 1, 1, 0]
    assert ob.PBDecoder(np.float32).decode(0, s, 1, 1, 0, 4)[0] == np.float32(-1.75)


@pytest.mark.parametrize("dt,nb", [(np.float32, 32), (np.float32, 12), (np.float64, 62), (np.float64, 30)])
def test_perbit_stream_lengths_and_truncation(dt, nb):
    """Stream p carries one bit per element plus one sign bit per element
    whose leading 1 is bit nb-1-p (sizes = whole 64-bit words); the full
    decode is sign * trunc(|x| 2^(nb-e)) 2^(e-nb); progressive deltas over
    any split of the planes add up to it. (At 64 planes the reference's
    int64 conversion saturates for |x| >= 2^(e-1), as in GroupedBP; that case
    is pinned by the compiled reference instead.)"""
    x = codec_x(900, dt, 3 + nb)
    e = ob.frexp_exp(ob.maxabs(x), dt)
    streams = ob.pb_encode(x, e, nb)
    xs = x.astype(np.float64) * 2.0 ** (nb - e)
    mag = np.trunc(np.abs(xs)).astype(np.uint64)
    lead = np.full(x.size, -1)
    nz = mag > 0
    lead[nz] = nb - 1 - np.floor(np.log2(mag[nz].astype(np.float64))).astype(int)
    for p in range(nb):
        bits = x.size + int(np.sum(lead == p))
        assert len(streams[p]) == (bits + 63) // 64 * 8, p
    full = ob.PBDecoder(dt).decode(0, streams, x.size, e, 0, nb)
    want = np.where(x < 0, -1.0, 1.0) * mag.astype(np.float64) * 2.0 ** (e - nb)
    np.testing.assert_array_equal(full.astype(np.float64), want.astype(dt).astype(np.float64))
    dec = ob.PBDecoder(dt)
    acc = np.zeros(x.size, np.float64)
    for s, t in ((0, 3), (3, nb // 2), (nb // 2, nb)):
        acc += dec.decode(0, streams[s:t], x.size, e, s, t - s).astype(np.float64)
    np.testing.assert_array_equal(acc, full.astype(np.float64))


# ------------------------------------------------ encoder-fused level errors ---


@pytest.mark.parametrize("enc,nb", [("nb", 32), ("gb", 32), ("nb", 12), ("gb", 18)])
def test_fused_level_errors_exactly_rounded(enc, nb):
    """exact mode = the correctly rounded sum of the per-element tables"""
    x = codec_x(200, np.float32, 21)
    e = ob.frexp_exp(ob.maxabs(x), np.float32)
    terms = [ob.fused_level_errors(x[i:i + 1], e, nb, enc, mode="seq") for i in range(x.size)]
    # per-element tables are already scaled by a power of two: the sum of the
    # scaled terms rounds like the scaled sum of the terms
    want = [float(sum((Fraction(float(t[k])) for t in terms), Fraction(0))) for k in range(nb + 1)]
    np.testing.assert_array_equal(ob.fused_level_errors(x, e, nb, enc, mode="exact"), np.array(want))


@pytest.mark.parametrize("enc,nb", [("nb", 32), ("nb", 16), ("gb", 32), ("gb", 20)])
def test_fused_level_errors_are_plane_prefix_errors(enc, nb):
    """Entry j is the squared error of keeping the first j planes: the sum of
    (x - decode_j(x))^2 with the codec's own decoder (j = 0: sum x^2; j = nb:
    the residual below the last plane). Bins 0 and nb square in float as the
    reference does (relative 2^-23); the others are double."""
    x = codec_x(512, np.float32, 5 + nb)
    e = ob.frexp_exp(ob.maxabs(x), np.float32)
    got = ob.fused_level_errors(x, e, nb, enc, mode="exact")
    if enc == "nb":
        planes = ob.nb_encode(x, e, nb)
        dec = lambda j: ob.nb_decode(planes[:j], x.size, e, 0, j, np.float32)
    else:
        streams = ob.gb_encode(x, e, nb)
        dec = lambda j: ob.GBDecoder(np.float32).decode(0, streams[:j], x.size, e, 0, j)
    for j in range(nb + 1):
        xt = np.zeros_like(x) if j == 0 else dec(j)
        d = x.astype(np.float64) - xt.astype(np.float64)
        want = float(sum((Fraction(float(v)) ** 2 for v in d), Fraction(0)))
        tol = 2.0 ** -22 if j in (0, nb) else 1e-12
        assert abs(got[j] - want) <= tol * want + 1e-300, (j, got[j], want)


# ------------------------------------------------------- hierarchical basis ---


@pytest.mark.parametrize("dims,L", [([9, 17, 5], 2), ([33, 17], 3), ([65], 5)])
def test_hierarchical_keeps_nodal_values(dims, L):
    """The hierarchical basis subtracts the interpolant only: the level-0 grid
    holds the field's own values at its nodes, every coefficient is u - I u_c
    with u_c the original nodal values, and recompose inverts it."""
    rng = np.random.default_rng(8)
    u = rng.standard_normal(int(np.prod(dims)))
    out = ob.decompose(u, dims, L, hierarchical=True).reshape(dims)
    ld = ob.level_dims(dims, L)
    step = [(n - 1) // (c - 1) if c > 1 else 1 for n, c in zip(dims, ld[0])]
    np.testing.assert_array_equal(out[tuple(slice(0, int(n)) for n in ld[0])],
                                  u.reshape(dims)[tuple(slice(0, None, st) for st in step)])
    one = ob.decompose(u, dims, 1, hierarchical=True).reshape(dims)
    P = kron_all([prolong_1d(n) for n in dims])
    U = u.reshape(dims)
    coarse_idx = [np.arange(0, n, 2) if n > 1 else np.arange(1) for n in dims]
    W = (u - P @ U[np.ix_(*coarse_idx)].ravel()).reshape(dims)
    m = [(n - 1) // 2 for n in dims]
    # odd indices sit after the coarse nodes along each dim
    idx = tuple(np.concatenate([np.arange(0, n, 2), np.arange(1, n, 2)]) if n > 1 else np.arange(1) for n in dims)
    Wr = W[np.ix_(*idx)]
    mask = np.ones(dims, bool)
    mask[tuple(slice(0, mm + 1) if n > 1 else slice(0, 1) for mm, n in zip(m, dims))] = False
    np.testing.assert_allclose(one[mask], Wr[mask], rtol=0, atol=1e-13)
    back = ob.recompose(ob.decompose(u, dims, L, hierarchical=True), dims, L, hierarchical=True)
    np.testing.assert_allclose(back, u, rtol=0, atol=1e-14 * np.max(np.abs(u)))
