"""GPU test of the allocation-free dot product (iga_dot_f64_noalloc): equal
to the allocating iga_dot_f64 (and to numpy) to FP64 rounding on two calls in
a row, the second smaller, so the static scratch array is reused with stale
partials from the first call still in it; its size limit is an error."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import igamf  # noqa: E402


def _stream():
    return torch.cuda.current_stream().cuda_stream


def test_dot_noalloc_matches_allocating_dot():
    L = igamf.lib()
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    for seed, n in ((11, 3_000_017), (12, 70_001), (13, 0)):
        a = torch.from_numpy(igamf.rng_normal(seed, n)).cuda()
        b = torch.from_numpy(igamf.rng_normal(seed + 100, n)).cuda()
        ref = C.c_double(0.0)
        if n:  # the allocating form rejects the null pointers of empty vectors
            igamf.check(L.iga_dot_f64(a.data_ptr(), b.data_ptr(), n, C.byref(ref), _stream()))
        rc = L.iga_dot_f64_noalloc(a.data_ptr(), b.data_ptr(), n, out.data_ptr(), _stream())
        assert rc == 0, igamf.IgaError(rc)
        torch.cuda.synchronize()
        got = out[0].item()
        exact = float(np.dot(a.cpu().numpy(), b.cpu().numpy()))
        scale = float(torch.sqrt((a * a).sum() * (b * b).sum()).item()) or 1.0
        assert abs(got - ref.value) <= 1e-13 * scale, (n, got, ref.value)
        assert abs(got - exact) <= 1e-13 * scale, (n, got, exact)
        assert np.isnan(out[1].item())  # nothing written past the one output
    # beyond the scratch array's capacity: an error, no launch (the pointers
    # are valid, only the count is too large)
    a = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    big = 67_108_864 + 1
    rc = L.iga_dot_f64_noalloc(a.data_ptr(), a.data_ptr(), big, out.data_ptr(), _stream())
    assert rc == igamf.IGA_ERR_SIZE_LIMIT
    rc = L.iga_dot_f64_noalloc(a.data_ptr(), a.data_ptr(), 8, None, _stream())
    assert rc == igamf.IGA_ERR_INVALID_ARGUMENT
    torch.cuda.synchronize()
    assert np.isnan(out[1].item())
