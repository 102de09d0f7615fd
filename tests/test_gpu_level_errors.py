"""GPU parity of the encoder-fused level errors (the level_errors overloads of
NegaBinaryBPEncoder::encode, NegaBinaryBPEncoder.hpp:65-125, and
GroupedBPEncoder::encode, GroupedBPEncoder.hpp:76-145): the device tables are
bit-identical to the oracle's exactly rounded tables, on codec inputs with
zeros, mixed magnitudes and elements far below 2^-nb of the maximum, and on
the level buffers of a plan."""
import numpy as np
import pytest
import torch

import oracle_bind as ob
from fields import bumps_field

pytestmark = pytest.mark.gpu


def codec_input(n, seed, spread=12, tiny=0.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * np.exp2(rng.integers(-spread, 1, n))
    x[rng.random(n) < 0.05] = 0.0
    if n:
        x[rng.integers(0, n)] = -1.75
    if tiny:
        x[rng.random(n) < tiny] *= 2.0 ** -60
    return x.astype(np.float32)


@pytest.mark.parametrize("enc", [0, 1])
@pytest.mark.parametrize("n,nb,seed,tiny", [
    (1, 32, 1, 0.0), (31, 16, 2, 0.0), (1000, 32, 3, 0.0), (100003, 32, 4, 0.1),
    (65536, 7, 5, 0.0), (8193, 30, 6, 0.02), (300000, 31, 7, 0.0),
])
def test_level_errors_match_oracle(gpu, enc, n, nb, seed, tiny):
    from prodm_amd import mdr
    x = codec_input(n, seed, tiny=tiny)
    e = ob.frexp_exp(ob.maxabs(x), np.float32)
    got = mdr.level_errors(torch.from_numpy(x).to(gpu), e, nb, enc)
    want = ob.fused_level_errors(x, e, nb, "nb" if enc == 0 else "gb", mode="exact")
    np.testing.assert_array_equal(got, want)


def test_level_errors_empty(gpu):
    from prodm_amd import mdr
    got = mdr.level_errors(torch.zeros(0, device=gpu), 0, 32, 0)
    assert np.all(got == 0)


@pytest.mark.parametrize("enc", [0, 1])
def test_plan_level_errors(gpu, enc):
    from prodm_amd import mdr
    dims, L, nb = [65, 65, 65], 4, 32
    u = bumps_field(dims, seed=11, dtype=np.float32)
    plan = mdr.Plan(dims, L, nb, torch.float32, enc)
    plan.refactor(torch.from_numpy(u).to(gpu))
    torch.cuda.synchronize()
    for l in range(L + 1):
        c = plan.coefficients(l).cpu().numpy()
        m, e, _ = plan.level_info(l)
        got = plan.level_errors(l)
        want = ob.fused_level_errors(c, e, nb, "nb" if enc == 0 else "gb", mode="exact")
        np.testing.assert_array_equal(got, want, err_msg=f"level {l}")


def test_level_errors_reject_unsupported(gpu):
    from prodm_amd import capi, mdr
    x = torch.ones(64, device=gpu)
    with pytest.raises(capi.ProdmError):
        mdr.level_errors(x, 1, 33, 0)        # more than 32 planes
    with pytest.raises(ValueError):
        mdr.level_errors(x.double(), 1, 32, 0)  # fp64 data
    plan = mdr.Plan([17, 17], 2, 64, torch.float64, 0)
    plan.refactor(torch.ones(17 * 17, dtype=torch.float64, device=gpu))
    with pytest.raises(capi.ProdmError):
        plan.level_errors(0)
