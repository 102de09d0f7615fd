"""GPU parity of the device-resident refactor / retrieve pipeline
(prodm_mdr_plan): level coefficients against the oracle's decompose +
interleave, bitplanes bit-exact against the oracle encoders fed the same
coefficients, and retrieval error against the requested bound."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fields import bumps_field

pytestmark = pytest.mark.gpu

TD = {np.float32: torch.float32, np.float64: torch.float64}
TOL = {np.float32: 1e-6, np.float64: 1e-14}


def level_coeffs_oracle(u, dims, L):
    dec = ob.decompose(u, dims, L)
    ld = ob.level_dims(dims, L)
    out = []
    for l in range(L + 1):
        coarse = ld[l - 1] if l else np.zeros(len(dims), np.uint32)
        out.append(ob.interleave(dec, dims, ld[l], coarse))
    return out


@pytest.mark.parametrize("dims,L,dt,enc,nb", [
    ([65, 65, 65], 4, np.float32, 0, 32),
    ([65, 65, 65], 4, np.float32, 1, 32),
    ([33, 129], 4, np.float32, 0, 16),
    ([129, 33, 17], 3, np.float64, 1, 60),
    ([257], 6, np.float64, 0, 64),
])
def test_plan_refactor_parity(gpu, dims, L, dt, enc, nb):
    from prodm_amd import mdr
    u = bumps_field(dims, seed=7, dtype=dt)
    plan = mdr.Plan(dims, L, nb, TD[dt], enc)
    plan.refactor(torch.from_numpy(u).to(gpu))
    torch.cuda.synchronize()
    ref = level_coeffs_oracle(u, dims, L)
    scale = max(np.max(np.abs(r)) for r in ref)
    for l in range(L + 1):
        c = plan.coefficients(l).cpu().numpy()
        assert c.size == ref[l].size
        assert np.max(np.abs(c.astype(np.float64) - ref[l])) <= TOL[dt] * scale, l
        m, e, n = plan.level_info(l)
        assert m == ob.maxabs(c) and e == ob.frexp_exp(ob.maxabs(c), dt)
        if enc == 0:
            oplanes = ob.nb_encode(c, e, nb)
            for p in range(nb):
                assert plan.plane_bytes(l, p) == oplanes[p].tobytes(), (l, p)
        else:
            ostreams = ob.gb_encode(c, e, nb)
            rb = plan.level_rb(l)
            assert ostreams[0][4:4 + rb.size] == rb.tobytes()
            for p in range(nb):
                body = plan.plane_bytes(l, p)
                expect = ostreams[p][4 + rb.size:] if p == 0 else ostreams[p]
                assert body == expect, (l, p)
    plan.close()


@pytest.mark.parametrize("enc", [0, 1])
def test_plan_retrieve_bounds(gpu, enc):
    """Retrieve with k planes per level; the L-inf error of the recomposed
    field stays within the MaxErrorEstimatorOB bound sum_l c * 2^(e_l - k_l)
    (MaxErrorEstimator.hpp:13-49, MaxErrorCollector.hpp:14-25)."""
    from prodm_amd import mdr
    dims, L, nb = [65, 65, 65], 4, 32
    u = bumps_field(dims, seed=1)
    plan = mdr.Plan(dims, L, nb, torch.float32, enc)
    x = torch.from_numpy(u).to(gpu)
    plan.refactor(x)
    c = (1.0 + 21.0 * np.sqrt(3) / 8) * 4
    for k in [4, 10, 16, 24, 32]:
        ks = [k] * (L + 1)
        out = plan.retrieve(ks).cpu().numpy().astype(np.float64)
        err = np.max(np.abs(out - u))
        bound = 0.0
        for l in range(L + 1):
            m, e, _ = plan.level_info(l)
            eb = e + 2 if enc == 0 else e  # negabinary bounds are 4 * max
            bound += c * 2.0 ** (eb - k)
        assert err <= bound + 1e-5 * np.max(np.abs(u)), (k, err, bound)
    plan.close()


@pytest.mark.parametrize("dims,L,dt,enc,nb", [
    ([33, 33, 33], 3, np.float32, 0, 32),
    ([33, 33, 33], 3, np.float32, 1, 32),
    ([65, 17], 3, np.float64, 1, 60),
])
def test_plan_streams_round_trip(gpu, dims, L, dt, enc, nb):
    """Plane streams leave the plan in the reference byte format (what
    BitplaneEncoder::encode returns) and, loaded into a fresh plan with the
    level bounds, retrieve the same field bit for bit."""
    from prodm_amd import mdr
    u = bumps_field(dims, seed=3, dtype=dt)
    a = mdr.Plan(dims, L, nb, TD[dt], enc)
    a.refactor(torch.from_numpy(u).to(gpu))
    b = mdr.Plan(dims, L, nb, TD[dt], enc)
    for l in range(L + 1):
        m, e, _ = a.level_info(l)
        c = a.coefficients(l).cpu().numpy()
        want = [p.tobytes() for p in ob.nb_encode(c, e, nb)] if enc == 0 else ob.gb_encode(c, e, nb)
        for p in range(nb):
            s = a.plane_stream(l, p)
            assert s == want[p], (l, p)
            b.load_plane_stream(l, p, s)
        b.set_level_max(l, m)
    for k in (nb, nb // 2, 6):
        ks = [k] * (L + 1)
        x = a.retrieve(ks).cpu().numpy()
        y = b.retrieve(ks).cpu().numpy()
        np.testing.assert_array_equal(x.view(np.uint8), y.view(np.uint8))
    a.close()
    b.close()


@pytest.mark.parametrize("dims,L,dt", [
    ([65, 65, 65], 4, np.float32),
    ([129, 33, 65], 3, np.float32),
    ([257, 129], 5, np.float64),
])
def test_retrieve_paths_bit_identical(gpu, dims, L, dt):
    """The plan's retrieve runs the finest level's decode, load sweep and line
    solves on a second side stream and forms u_c - z in the interpolation;
    with stage timing on, every launch runs on one stream and the last solve
    forms u_c - z. Both give the same field bit for bit."""
    from prodm_amd import mdr
    u = bumps_field(dims, seed=21, dtype=dt)
    nb = 32 if dt == np.float32 else 64
    plan = mdr.Plan(dims, L, nb, TD[dt], 0)
    plan.refactor(torch.from_numpy(u).to(gpu))
    for ks in ([nb] * (L + 1), [nb - 2 * l for l in range(L + 1)][::-1]):
        a = plan.retrieve(ks).clone()
        plan.profile(True)
        b = plan.retrieve(ks).clone()
        plan.profile(False)
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.uint8), b.view(torch.uint8)), ks
    plan.close()
