"""The allocation-free GroupedBP encode (prodm_grouped_encode_noalloc_f32)
against prodm_grouped_encode_f32 with a caller workspace, through the C ABI:
the recording-plane bytes, every plane's words and the words per plane are
bit-identical, on two calls in a row (the second with a smaller input, so
the scratch holds the first call's contents), and a request beyond the
scratch size is refused with an error instead of running."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def encode(x, exp, nb, noalloc):
    from prodm_amd import capi
    lib = capi.lib()
    n = x.numel()
    nblk = (n + 31) // 32
    stride = 2 * nblk + 64
    rb = torch.full((nblk,), 0xAB, dtype=torch.uint8, device=x.device)
    planes = torch.zeros(nb * stride, dtype=torch.int32, device=x.device)
    words = torch.zeros(nb, dtype=torch.int32, device=x.device)
    s = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = lib.prodm_grouped_encode_noalloc_f32(x.data_ptr(), n, exp, nb, 0, rb.data_ptr(), planes.data_ptr(),
                                                  stride, words.data_ptr(), s)
    else:
        wsb = lib.prodm_grouped_workspace_bytes(n, nb, capi.F32)
        ws = torch.empty(wsb, dtype=torch.uint8, device=x.device)
        rc = lib.prodm_grouped_encode_f32(x.data_ptr(), n, exp, nb, 0, rb.data_ptr(), planes.data_ptr(), stride,
                                          words.data_ptr(), ws.data_ptr(), wsb, s)
    capi.check(rc, "grouped encode" + (" (no allocation)" if noalloc else ""))
    torch.cuda.synchronize()
    w = words.cpu().numpy()
    p = planes.view(nb, stride).cpu().numpy()
    return rb.cpu().numpy(), w, [p[i, :w[i]].copy() for i in range(nb)]


@pytest.mark.parametrize("sizes", [(5_000_003, 70_001), (1_000_000, 999_937)])
def test_noalloc_grouped_encode_matches(gpu, sizes):
    rng = np.random.default_rng(11)
    for n in sizes:
        print(f"no-allocation grouped encode, n = {n}", flush=True)
        h = (rng.standard_normal(n) * np.exp2(rng.integers(-12, 1, n))).astype(np.float32)
        h[rng.random(n) < 0.05] = 0
        x = torch.from_numpy(h).to(gpu)
        exp = int(np.frexp(np.max(np.abs(h)))[1])
        for nb in (32, 19):
            a = encode(x, exp, nb, noalloc=False)
            b = encode(x, exp, nb, noalloc=True)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
            for pa, pb in zip(a[2], b[2]):
                np.testing.assert_array_equal(pa, pb)


def test_noalloc_grouped_encode_refuses_oversize(gpu):
    from prodm_amd import capi
    lib = capi.lib()
    n = 2_000_000_000  # needs about 64 MB of scratch at 32 planes
    rc = lib.prodm_grouped_encode_noalloc_f32(None, n, 0, 32, 0, None, None, 2 * ((n + 31) // 32), None, None)
    assert rc == -6  # PRODM_ERR_UNSUPPORTED: refused before any device work
    assert b"scratch" in lib.prodm_last_error()
