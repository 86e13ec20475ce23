"""Python face of the CPU oracle (oracle.c) plus the FP64 network oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product path never imports it.

- liboracle.so: restatements of the reference kernels in C (see oracle.c for
  the reference lines each follows), built with gcc -ffp-contract=off.
- net_forward_fp64: the descriptor network in float64 with torch's CPU
  conv3d / relu / max_pool3d (library routines) -- cuDNN's arithmetic is
  closed source, so the FP64 definition is the network oracle.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import c_float, c_int, c_long, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TDM_ORACLE_LIB names another build of the oracle (the sanitizer build)
LIB_PATH = os.environ.get("TDM_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_invert_matrix.restype = c_int
    return _lib


def _ptr(a: np.ndarray) -> c_void_p:
    return a.ctypes.data_as(c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


# ---- TSDF -------------------------------------------------------------------------


def integrate(cam_K, cam2base, depth, dims, origin, voxel_size, trunc, tsdf, weight) -> None:
    """In place on float32 grids [dz, dy, dx] (depth-fusion/demo.cu:18-68)."""
    K, P, D = _f32(cam_K), _f32(cam2base), _f32(depth)
    assert tsdf.dtype == np.float32 and weight.dtype == np.float32
    assert tsdf.flags.c_contiguous and weight.flags.c_contiguous
    H, W = D.shape
    lib().orc_integrate(_ptr(K), _ptr(P), _ptr(D), c_int(H), c_int(W), c_int(dims[0]),
                        c_int(dims[1]), c_int(dims[2]), c_float(origin[0]), c_float(origin[1]),
                        c_float(origin[2]), c_float(voxel_size), c_float(trunc), _ptr(tsdf),
                        _ptr(weight))


def new_grid(dims):
    dx, dy, dz = dims
    return np.ones((dz, dy, dx), np.float32), np.zeros((dz, dy, dx), np.float32)


def multiply_matrix(a, b) -> np.ndarray:
    out = np.zeros(16, np.float32)
    lib().orc_multiply_matrix(_ptr(_f32(a).reshape(16)), _ptr(_f32(b).reshape(16)), _ptr(out))
    return out.reshape(4, 4)


def invert_matrix(m) -> np.ndarray:
    out = np.zeros(16, np.float32)
    ok = lib().orc_invert_matrix(_ptr(_f32(m).reshape(16)), _ptr(out))
    if not ok:
        raise ValueError("singular matrix")
    return out.reshape(4, 4)


def depth_from_mm(mm: np.ndarray) -> np.ndarray:
    mm = np.ascontiguousarray(mm, dtype=np.uint16)
    out = np.zeros(mm.shape, np.float32)
    lib().orc_depth_from_mm(_ptr(mm), c_long(mm.size), _ptr(out))
    return out


# ---- TDF ---------------------------------------------------------------------------


def grid_params(pts, voxel_size=0.01, pad=15):
    p = _f32(pts)
    dims = np.zeros(3, np.int32)
    origin = np.zeros(3, np.float32)
    lib().orc_grid_params(_ptr(p), c_long(len(p)), c_float(voxel_size), c_int(pad), _ptr(dims),
                          _ptr(origin))
    return [int(d) for d in dims], origin


def occupancy(pts, dims, origin, voxel_size) -> np.ndarray:
    p = _f32(pts)
    occ = np.zeros((dims[2], dims[1], dims[0]), np.float32)
    lib().orc_occupancy(_ptr(p), c_long(len(p)), _ptr(np.array(dims, np.int32)),
                        _ptr(_f32(origin)), c_float(voxel_size), _ptr(occ))
    return occ


def compute_tdf(occ: np.ndarray, voxel_size=0.01, trunc=0.05, decode="exact") -> np.ndarray:
    occ = _f32(occ)
    dz, dy, dx = occ.shape
    tdf = np.zeros_like(occ)
    fn = lib().orc_compute_tdf_ref if decode == "reference" else lib().orc_compute_tdf_exact
    fn(_ptr(occ), _ptr(tdf), c_int(dx), c_int(dy), c_int(dz), c_float(voxel_size), c_float(trunc))
    return tdf


def float_decode(v: int, dx: int, dy: int):
    out = np.zeros(3, np.int32)
    lib().orc_float_decode(c_long(v), c_int(dx), c_int(dy), _ptr(out))
    return tuple(int(a) for a in out)


def keypoint_grid(kp, origin, voxel_size) -> np.ndarray:
    k = _f32(kp).reshape(-1, 3)
    out = np.zeros(k.shape, np.int32)
    lib().orc_keypoint_grid(_ptr(k), c_long(len(k)), _ptr(_f32(origin)), c_float(voxel_size), _ptr(out))
    return out


def extract_patches(tdf: np.ndarray, kp_grid, size=30) -> np.ndarray:
    t = _f32(tdf)
    dz, dy, dx = t.shape
    kg = np.ascontiguousarray(kp_grid, dtype=np.int32).reshape(-1, 3)
    out = np.zeros((len(kg), size, size, size), np.float32)
    lib().orc_extract_patches(_ptr(t), c_int(dx), c_int(dy), c_int(dz), _ptr(kg), c_long(len(kg)),
                              c_int(size), _ptr(out))
    return out


# ---- network -------------------------------------------------------------------------


def conv3d(x: np.ndarray, w: np.ndarray, b: np.ndarray | None, relu: bool) -> np.ndarray:
    """float32 direct convolution, NCDHW (core/marvin.hpp:4635-4709 semantics)."""
    x = _f32(x)
    w = _f32(w)
    nb, cin, d = x.shape[0], x.shape[1], x.shape[2]
    cout, k = w.shape[0], w.shape[2]
    od = d - k + 1
    out = np.zeros((nb, cout, od, od, od), np.float32)
    bp = _ptr(_f32(b).reshape(-1)) if b is not None else c_void_p(0)
    lib().orc_conv3d(_ptr(x), c_long(nb), c_int(cin), c_int(d), _ptr(w), bp, c_int(cout), c_int(k),
                     c_int(1 if relu else 0), _ptr(out))
    return out


def maxpool3d(x: np.ndarray, w=2, s=2) -> np.ndarray:
    x = _f32(x)
    nb, c, d = x.shape[0], x.shape[1], x.shape[2]
    od = 1 + int(np.ceil((d - w) / s))
    out = np.zeros((nb, c, od, od, od), np.float32)
    lib().orc_maxpool3d(_ptr(x), c_long(nb), c_int(c), c_int(d), c_int(w), c_int(s), _ptr(out))
    return out


def net_forward_fp32(patches: np.ndarray, weights: dict, shapes, upto: int | None = None):
    """float32 C loop nest over the whole net; returns the output of layer
    `upto` (default: last), NCDHW.  shapes: tdm.netdef.layer_shapes()."""
    x = _f32(patches).reshape(len(patches), 1, *patches.shape[-3:])
    last = len(shapes) - 1 if upto is None else upto
    for i, (name, _cin, _cout, _k, _d, _od, pooled) in enumerate(shapes[: last + 1]):
        relu = i < len(shapes) - 1
        x = conv3d(x, weights[name + ".weight"], weights[name + ".bias"].reshape(-1), relu)
        if pooled is not None:
            x = maxpool3d(x)
    return x


def net_forward_fp64(patches: np.ndarray, weights: dict, shapes, upto: int | None = None):
    """float64 torch (CPU) forward; returns a list of every layer's output up to
    `upto` (after ReLU and pool), as float64 numpy NCDHW."""
    import torch
    import torch.nn.functional as F

    x = torch.from_numpy(np.asarray(patches, dtype=np.float64).reshape(len(patches), 1, *patches.shape[-3:]))
    outs = []
    last = len(shapes) - 1 if upto is None else upto
    with torch.no_grad():
        for i, (name, _cin, _cout, _k, _d, _od, pooled) in enumerate(shapes[: last + 1]):
            w = torch.from_numpy(np.asarray(weights[name + ".weight"], dtype=np.float64))
            b = torch.from_numpy(np.asarray(weights[name + ".bias"], dtype=np.float64).reshape(-1))
            x = F.conv3d(x, w, b)
            if i < len(shapes) - 1:
                x = F.relu(x)
            if pooled is not None:
                x = F.max_pool3d(x, 2, 2, ceil_mode=True)
            outs.append(x.numpy().copy())
    return outs


def compute_tdf_ref_at(occ: np.ndarray, idx: np.ndarray, voxel_size=0.01, trunc=0.05) -> np.ndarray:
    occ = _f32(occ)
    dz, dy, dx = occ.shape
    ix = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros(len(ix), np.float32)
    lib().orc_compute_tdf_ref_at(_ptr(occ), c_int(dx), c_int(dy), c_int(dz), c_float(voxel_size),
                                 c_float(trunc), _ptr(ix), c_long(len(ix)), _ptr(out))
    return out


def float_decode_all(n: int, dx: int, dy: int):
    """numpy float32 emulation of core/demo.cu:32-34 for voxel indices 0..n-1
    (independent of the C restatement; used to pin it)."""
    v = np.arange(n, dtype=np.int64).astype(np.float32)
    fdx, fdy = np.float32(dx), np.float32(dy)
    z = np.floor(v / (fdx * fdy)).astype(np.int64)
    zf = z.astype(np.float32)
    y = np.floor((v - (zf * fdx * fdy)) / fdx).astype(np.int64)
    yf = y.astype(np.float32)
    x = (v - (zf * fdx * fdy) - (yf * fdx)).astype(np.float32)
    x = np.trunc(x).astype(np.int64)
    return x, y, z


def compute_tdf_range(occ: np.ndarray, tdf: np.ndarray, v0: int, v1: int, voxel_size=0.01,
                      trunc=0.05) -> None:
    """Exact-decode TDF of voxels [v0, v1) into tdf (whole grid, in place)."""
    occ = _f32(occ)
    assert tdf.dtype == np.float32 and tdf.flags.c_contiguous and tdf.shape == occ.shape
    dz, dy, dx = occ.shape
    lib().orc_compute_tdf_exact_range(_ptr(occ), _ptr(tdf), c_int(dx), c_int(dy), c_int(dz),
                                      c_float(voxel_size), c_float(trunc), c_long(v0), c_long(v1))


def surface_points(tsdf: np.ndarray, weight: np.ndarray, origin, voxel_size, tsdf_thresh=0.2,
                   weight_thresh=1.0) -> np.ndarray:
    """SaveVoxelGrid2SurfacePointCloud (depth-fusion/utils.hpp:14-56): the
    voxels with |tsdf| < tsdf_thresh and weight > weight_thresh in linear
    index order, as origin + (float)index * voxel_size (float32, no FMA)."""
    dz, dy, dx = tsdf.shape
    t = tsdf.reshape(-1)
    w = weight.reshape(-1)
    idx = np.nonzero((np.abs(t) < np.float32(tsdf_thresh)) & (w > np.float32(weight_thresh)))[0]
    z = idx // (dx * dy)
    y = (idx - z * dx * dy) // dx
    x = idx - z * dx * dy - y * dx
    o = np.asarray(origin, np.float32)
    vs = np.float32(voxel_size)
    return np.stack([o[0] + x.astype(np.float32) * vs, o[1] + y.astype(np.float32) * vs,
                     o[2] + z.astype(np.float32) * vs], axis=1)


def tdf_accurate(pts, dims, origin, voxel_size, trunc) -> np.ndarray:
    """Get1NNSearchDist (training/match.hpp:13-51) on the demo grid: for each
    voxel centre origin + idx*vs (float32), d = sqrtf((dx^2 + dy^2) + dz^2)
    to the nearest point in float32, tdf = 1 - min(d, trunc)/trunc.  Every
    point counts, inside the grid or not (the kernel loops over all the
    points it is given, training/match.hpp:40-48).  The candidates of each
    voxel come from a float64 KD-tree query with a margin; the distances
    themselves are evaluated in float32 in the reference's order."""
    from scipy.spatial import cKDTree

    pts = np.asarray(pts, np.float32)
    dx, dy, dz = dims
    o = np.asarray(origin, np.float32)
    vs, tr = np.float32(voxel_size), np.float32(trunc)
    z, y, x = np.meshgrid(np.arange(dz), np.arange(dy), np.arange(dx), indexing="ij")
    c = np.stack([o[0] + x.astype(np.float32) * vs, o[1] + y.astype(np.float32) * vs,
                  o[2] + z.astype(np.float32) * vs], -1).reshape(-1, 3)
    out = np.zeros(len(c), np.float32)
    if len(pts) == 0:
        return out.reshape(dz, dy, dx)
    tree = cKDTree(pts.astype(np.float64))
    cand = tree.query_ball_point(c.astype(np.float64), float(tr) * 1.0001 + 1e-6)
    for i, lst in enumerate(cand):
        if not lst:
            continue
        q = pts[lst]
        e = c[i] - q
        d2 = (e[:, 0] * e[:, 0] + e[:, 1] * e[:, 1]) + e[:, 2] * e[:, 2]
        d = np.sqrt(d2.min())
        out[i] = np.float32(1.0) - np.minimum(d, tr) / tr
    return out.reshape(dz, dy, dx)
