"""tdm_tdf_from_points_noalloc (no allocation; bit grid in a static per-GPU
array) against tdm_tdf_from_points and the oracle, through the C ABI: two
calls in a row, the second on a smaller cloud and grid, so a stale bit grid
or tile flag left by the first call would show."""
import ctypes

import numpy as np
import pytest

import oracle
import tdm
from tdm import synth

pytestmark = pytest.mark.gpu

VS = np.float32(0.01)
TRUNC = np.float32(0.05)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _run_both(torch, cloud, decode):
    pts = torch.from_numpy(cloud).to("cuda:0")
    dims, origin = oracle.grid_params(cloud, VS, 15)
    print(f"case noalloc n={len(cloud)} dims={dims} decode={decode}", flush=True)
    a = torch.full((dims[2], dims[1], dims[0]), -1.0, device="cuda:0")
    b = torch.full((dims[2], dims[1], dims[0]), -2.0, device="cuda:0")
    tdm.tdf_from_points_noalloc(pts, dims, origin, VS, TRUNC, a, decode)
    tdm.tdf_from_points(pts, dims, origin, VS, TRUNC, b, decode)
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy(), dims, origin


def test_noalloc_equals_allocating_entry_point(torch_gpu):
    torch = torch_gpu
    big = synth.fragment(200_000, seed=3)  # 419 x 321 x 347 voxels, 6.2 MB of bits
    small = synth.fragment(20_000, seed=61) * np.float32(0.5)
    for cloud in (big, small):
        got, want, dims, origin = _run_both(torch, cloud, tdm.DECODE_EXACT)
        assert np.array_equal(_bits(got), _bits(want))
        if len(cloud) < 100_000:
            ref = oracle.compute_tdf(oracle.occupancy(cloud, dims, origin, VS), VS, TRUNC)
            assert np.array_equal(_bits(got), _bits(ref))
    # reference decode below 2^24 voxels is the same computation
    got, want, dims, _ = _run_both(torch, small, tdm.DECODE_REFERENCE)
    assert dims[0] * dims[1] * dims[2] <= 1 << 24
    assert np.array_equal(_bits(got), _bits(want))


def test_noalloc_rejects_oversize_grids_before_any_work(torch_gpu):
    torch = torch_gpu
    L = tdm.lib()
    pts = torch.zeros(1, 3, device="cuda:0")
    out = torch.zeros(1, device="cuda:0")
    # 1024 x 1024 x 160 voxels: 20 MiB of bits, above the 16 MiB array
    dims = (ctypes.c_int * 3)(1024, 1024, 160)
    org = (ctypes.c_float * 3)(0, 0, 0)
    st = L.tdm_tdf_from_points_noalloc(ctypes.c_void_p(pts.data_ptr()), 1, dims, org,
                                       ctypes.c_float(0.01), ctypes.c_float(0.05), 0,
                                       ctypes.c_void_p(out.data_ptr()), None)
    assert st == tdm.ERR_UNSUPPORTED and b"static scratch" in L.tdm_last_error()
    # reference decode above 2^24 voxels needs the allocating entry point
    dims = (ctypes.c_int * 3)(300, 300, 200)
    st = L.tdm_tdf_from_points_noalloc(ctypes.c_void_p(pts.data_ptr()), 1, dims, org,
                                       ctypes.c_float(0.01), ctypes.c_float(0.05), 1,
                                       ctypes.c_void_p(out.data_ptr()), None)
    assert st == tdm.ERR_UNSUPPORTED and b"reference decode" in L.tdm_last_error()
