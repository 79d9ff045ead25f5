"""Runtime helpers of the C-ABI on the GPU: the pattern fill (AoS and SoA records) and the index
rebasing used when an instance is placed into a concatenated mesh, against numpy."""

import numpy as np
import pytest
import torch

import oppx

pytestmark = pytest.mark.gpu


def stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("dtype,dim,stride_extra", [(np.int32, 1, 0), (np.float64, 3, 0), (np.float64, 3, 5),
                                                    (np.int32, 4, 7), (np.float64, 1, 3)])
def test_fill_pattern(dtype, dim, stride_extra):
    n = 10_007
    pat = (np.arange(dim) * 3 + 1).astype(dtype)
    if stride_extra == 0:  # AoS: n records of dim consecutive values
        dst = torch.full((n * dim + 9,), -5, dtype=torch.from_numpy(pat).dtype, device="cuda")
        oppx.check(oppx.lib().oppx_fill_pattern(dst.data_ptr(), n, dim, pat.itemsize, pat.ctypes.data, 0, stream()),
                   "fill")
        got = dst.cpu().numpy()
        assert np.array_equal(got[:n * dim].reshape(n, dim), np.broadcast_to(pat, (n, dim)))
        assert (got[n * dim:] == -5).all()
    else:  # SoA: component d of record i at d * stride + i
        stride = n + stride_extra
        dst = torch.full((dim * stride,), -5, dtype=torch.from_numpy(pat).dtype, device="cuda")
        oppx.check(oppx.lib().oppx_fill_pattern(dst.data_ptr(), n, dim, pat.itemsize, pat.ctypes.data, stride,
                                                stream()), "fill")
        got = dst.cpu().numpy().reshape(dim, stride)
        assert np.array_equal(got[:, :n], np.broadcast_to(pat[:, None], (dim, n)))
        assert (got[:, n:] == -5).all()


def test_offset_indices():
    rng = np.random.default_rng(3)
    n, rec, first, count, off = 5003, 8, 2, 4, 1000
    a = rng.integers(-3, 500, (n, rec)).astype(np.int32)
    a[rng.random((n, rec)) < 0.05] = oppx.MAX_CELL_INDEX
    d = torch.from_numpy(a.copy()).to("cuda")
    oppx.check(oppx.lib().oppx_offset_indices(d.data_ptr(), n, rec, first, count, off, stream()), "offset")
    exp = a.copy()
    sl = exp[:, first:first + count]
    sl[(sl >= 0) & (sl != oppx.MAX_CELL_INDEX)] += off
    assert np.array_equal(d.cpu().numpy(), exp)
