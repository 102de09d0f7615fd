"""The benchmark configurations beyond the timed one, as GPU parity cases:

* 129^3 float32, 4 levels, 32 bitplanes, GroupedBP (the CPU-runnable case):
  every plane stream bit-exact against the oracle fed the same coefficients,
  and every relative L-inf tolerance 1e-1 .. 1e-5 met by the retrieval the
  greedy planner picks (MaxErrorEstimatorOB + MaxErrorCollector tables).
* 1025 x 1025 x 257 float64, 7 levels, 64 bitplanes, GroupedBP: at full
  size, through size-independent properties: tolerances 1e-4, 1e-8, 1e-12
  met, all 64 planes reproduce the field to the coefficient truncation, and
  the recording planes agree with an independent count on the coefficients.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle_bind as ob
from fields import bumps_field, fourier_field_gpu

pytestmark = pytest.mark.gpu


def plan_planes(plan, L, nb, tol, dtype):
    from prodm_amd import capi, interpret
    tables, sizes = [], []
    for l in range(L + 1):
        m, e, n = plan.level_info(l)
        t = np.zeros(nb + 1)
        if dtype == torch.float32:
            capi.lib().prodm_collect_max_f32(C.c_float(m), nb, t.ctypes.data_as(C.c_void_p))
        else:
            capi.lib().prodm_collect_max_f64(C.c_double(m), nb, t.ctypes.data_as(C.c_void_p))
        tables.append(t)
        sizes.append([max(1, len(plan.plane_stream(l, p))) for p in range(nb)] if n < 4_000_000 else
                     [plan.plane(l, p)[1] or 1 for p in range(nb)])
    idx = np.zeros(L + 1, np.uint8)
    _, idx = interpret.retrieve_sizes(interpret.GREEDY, interpret.MAX_OB, sizes, tables, tol, idx, ndims=3,
                                      dtype=np.float32 if dtype == torch.float32 else np.float64)
    return [int(k) for k in idx]


def test_config0_129_grouped(gpu):
    from prodm_amd import capi, mdr
    dims, L, nb = [129, 129, 129], 4, 32
    u = bumps_field(dims, seed=0, nbumps=8, noise=1e-3)
    plan = mdr.Plan(dims, L, nb, torch.float32, capi.GROUPED)
    x = torch.from_numpy(u).to(gpu)
    plan.refactor(x)
    torch.cuda.synchronize()
    for l in range(L + 1):
        c = plan.coefficients(l).cpu().numpy()
        _, e, _ = plan.level_info(l)
        want = ob.gb_encode(c, e, nb)
        for p in range(nb):
            assert plan.plane_stream(l, p) == want[p], (l, p)
    R = float(u.max() - u.min())
    for rel in (1e-1, 1e-2, 1e-3, 1e-4, 1e-5):
        k = plan_planes(plan, L, nb, rel * R, torch.float32)
        out = plan.retrieve(k).cpu().numpy().astype(np.float64)
        err = float(np.max(np.abs(out - u)))
        print(f"129^3 grouped rel {rel:.0e}: planes {k}, error {err / R:.3e}")
        assert err <= rel * R, (rel, k, err)
    plan.close()


def test_config2_fp64_full_size(gpu):
    from prodm_amd import capi, mdr
    dims, L, nb = [1025, 1025, 257], 7, 64
    u = fourier_field_gpu(dims, seed=2, device=gpu)
    plan = mdr.Plan(dims, L, nb, torch.float64, capi.GROUPED)
    plan.refactor(u)
    torch.cuda.synchronize()
    R = float(u.max() - u.min())
    for rel in (1e-4, 1e-8, 1e-12):
        k = plan_planes(plan, L, nb, rel * R, torch.float64)
        out = plan.retrieve(k)
        err = float(torch.max(torch.abs(out - u)))
        print(f"1025x1025x257 f64 rel {rel:.0e}: planes {k}, error {err / R:.3e}")
        assert err <= rel * R, (rel, k, err)
    # all planes: only the coefficient truncation at 2^(e - 64) remains
    out = plan.retrieve([nb] * (L + 1))
    assert float(torch.max(torch.abs(out - u))) <= 1e-13 * R
    # recording planes of the finest level against a count on its coefficients
    c = plan.coefficients(L)
    m, e, n = plan.level_info(L)
    mag = torch.trunc(torch.abs(torch.ldexp(c, torch.tensor(nb - e, device=gpu))))
    nblk = (n + 63) // 64
    pad = torch.zeros(nblk * 64, dtype=torch.float64, device=gpu)
    pad[:n] = mag
    bmax = pad.view(nblk, 64).max(dim=1).values
    bits = torch.where(bmax > 0, torch.frexp(bmax).exponent.to(torch.float64), torch.zeros_like(bmax))  # bit length
    want = (nb - bits).to(torch.uint8).cpu().numpy()
    np.testing.assert_array_equal(plan.level_rb(L), want)
    plan.close()
