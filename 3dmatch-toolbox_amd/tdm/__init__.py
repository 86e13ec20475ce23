"""ctypes bindings of lib3dmatch_amd.so (the C ABI declared in include/tdm.h).

Device arrays are torch tensors on a ROCm device (torch is used only for
device memory and streams); every compute call goes to the HIP kernels in
lib3dmatch_amd.so.  There is no CPU fallback: if the library is missing,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_void_p

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
# TDM_LIB names another build of the library (the host sanitizer build)
LIB_PATH = os.environ.get("TDM_LIB") or os.path.join(PKG_DIR, "lib3dmatch_amd.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "tdm.h")
BIN_DIR = os.path.join(PKG_DIR, "bin")

OK = 0
ERR_INVALID_ARG = 1
ERR_IO = 2
ERR_FORMAT = 3
ERR_HIP = 4
ERR_NOMEM = 5
ERR_UNSUPPORTED = 6
ERR_NO_DEVICE = 7

PREC_FP32 = 0
PREC_BF16 = 1
PREC_FP16 = 2
PREC_FP8 = 3
DECODE_EXACT = 0
DECODE_REFERENCE = 1


class TdmError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"tdm error {status}: {message}")
        self.status = status


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.tdm_status_string.restype = c_char_p
        L.tdm_last_error.restype = c_char_p
        L.tdm_net_layer_name.restype = c_char_p
        L.tdm_free.restype = None
        _lib = L
    return _lib


def header_symbols() -> list[str]:
    """Function names declared in include/tdm.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(tdm_[a-z0-9_]+)\s*\(", text)))


def _check(status: int) -> None:
    if status != OK:
        msg = lib().tdm_last_error().decode(errors="replace")
        raise TdmError(status, msg)


def _p(t) -> c_void_p:
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return c_void_p(0)
    return c_void_p(t.data_ptr())


def _stream(stream=None) -> c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


def _f(x) -> c_float:
    return c_float(float(x))


def _int3(v):
    return (c_int * 3)(*[int(a) for a in v])


def _float3(v):
    return (c_float * 3)(*[float(a) for a in np.asarray(v, dtype=np.float32)])


# ---- TSDF -----------------------------------------------------------------------


def tsdf_init(tsdf, weight, stream=None) -> None:
    _check(lib().tdm_tsdf_init(None, _p(tsdf), _p(weight), c_size_t(tsdf.numel()), _stream(stream)))


def integrate(cam_K, cam2base, depth, dims, origin, voxel_size, trunc, tsdf, weight, stream=None):
    """One frame (Integrate, depth-fusion/demo.cu:18-68).  depth: [H, W]."""
    H, W = depth.shape[-2:]
    _check(lib().tdm_integrate(None, _p(cam_K), _p(cam2base), _p(depth), c_int(H), c_int(W),
                               c_int(dims[0]), c_int(dims[1]), c_int(dims[2]), _f(origin[0]),
                               _f(origin[1]), _f(origin[2]), _f(voxel_size), _f(trunc),
                               _p(tsdf), _p(weight), _stream(stream)))


def integrate_frames(cam_K, cam2base, depths, dims, origin, voxel_size, trunc, tsdf, weight,
                     stream=None):
    """All frames in one pass.  cam2base: [F, 16]; depths: [F, H, W]."""
    F, H, W = depths.shape
    _check(lib().tdm_integrate_frames(None, _p(cam_K), _p(cam2base), _p(depths), c_int(F), c_int(H),
                                      c_int(W), c_int(dims[0]), c_int(dims[1]), c_int(dims[2]),
                                      _f(origin[0]), _f(origin[1]), _f(origin[2]),
                                      _f(voxel_size), _f(trunc), _p(tsdf), _p(weight),
                                      _stream(stream)))


def integrate_scenes(cam_K, cam2base, depths, dims, origins, voxel_size, trunc, tsdf, weight,
                     stream=None):
    """Many scenes in one launch.  cam2base [S, F, 16]; depths [S, F, H, W];
    origins [S, 3]; tsdf/weight [S, N] (all device)."""
    S, F, H, W = depths.shape
    _check(lib().tdm_integrate_scenes(None, _p(cam_K), _p(cam2base), _p(depths), c_int(S), c_int(F),
                                      c_int(H), c_int(W), c_int(dims[0]), c_int(dims[1]),
                                      c_int(dims[2]), _p(origins), _f(voxel_size), _f(trunc),
                                      _p(tsdf), _p(weight), _stream(stream)))


def surface_points(tsdf, weight, dims, origin, voxel_size, pts=None, capacity=0,
                   tsdf_thresh=0.2, weight_thresh=1.0, stream=None) -> int:
    """GPU surface extraction (depth-fusion/utils.hpp:14-56) into pts
    [capacity, 3]; returns the total point count (synchronizes)."""
    n = c_size_t()
    _check(lib().tdm_surface_points(None, _p(tsdf), _p(weight), _int3(dims), _float3(origin),
                                    _f(voxel_size), _f(tsdf_thresh), _f(weight_thresh), _p(pts),
                                    c_size_t(capacity), ctypes.byref(n), _stream(stream)))
    return int(n.value)


class Fusion:
    """tdm_fusion: depth PNG frames fused with overlapped decode/copy/integrate."""

    def __init__(self, cam_K, height, width, dims, origin, voxel_size, trunc, threads=0):
        K = (c_float * 9)(*[float(v) for v in np.asarray(cam_K, np.float32).reshape(9)])
        self._h = c_void_p()
        _check(lib().tdm_fusion_create(None, K, c_int(height), c_int(width), _int3(dims),
                                       _float3(origin), _f(voxel_size), _f(trunc), c_int(threads),
                                       ctypes.byref(self._h)))
        self.dims = list(dims)

    def add_png(self, path: str, cam2base) -> None:
        P = (c_float * 16)(*[float(v) for v in np.asarray(cam2base, np.float32).reshape(16)])
        _check(lib().tdm_fusion_add_png(self._h, path.encode(), P))

    def wait(self) -> None:
        _check(lib().tdm_fusion_wait(self._h))

    def copy_grids(self, tsdf, weight, stream=None) -> None:
        """Waits, then copies the fused grids into device tensors."""
        _check(lib().tdm_fusion_copy_grids(self._h, _p(tsdf), _p(weight), _stream(stream)))

    def close(self) -> None:
        if self._h:
            lib().tdm_fusion_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def depth_from_mm(mm, out, stream=None) -> None:
    _check(lib().tdm_depth_from_mm(None, _p(mm), c_size_t(mm.numel()), _p(out), _stream(stream)))


# ---- TDF ------------------------------------------------------------------------


def grid_params(pts, voxel_size=0.01, padding=15, stream=None):
    """(dims[3] ints, origin[3] float32) for a device point tensor [N, 3]."""
    dims = (c_int * 3)()
    origin = (c_float * 3)()
    _check(lib().tdm_grid_params(None, _p(pts), c_int(pts.shape[0]), _f(voxel_size),
                                 c_int(padding), dims, origin, _stream(stream)))
    return [int(d) for d in dims], np.array(list(origin), dtype=np.float32)


def occupancy(pts, dims, origin, voxel_size, occ, stream=None) -> None:
    _check(lib().tdm_occupancy(None, _p(pts), c_int(pts.shape[0]), _int3(dims), _float3(origin),
                               _f(voxel_size), _p(occ), _stream(stream)))


def compute_tdf(occ, tdf, dims, voxel_size, trunc, decode=DECODE_EXACT, stream=None) -> None:
    _check(lib().tdm_compute_tdf(None, _p(occ), _p(tdf), c_int(dims[0]), c_int(dims[1]),
                                 c_int(dims[2]), _f(voxel_size), _f(trunc), c_int(decode),
                                 _stream(stream)))


def tdf_from_points(pts, dims, origin, voxel_size, trunc, tdf, decode=DECODE_EXACT, stream=None):
    _check(lib().tdm_tdf_from_points(None, _p(pts), c_int(pts.shape[0]), _int3(dims),
                                     _float3(origin), _f(voxel_size), _f(trunc), c_int(decode),
                                     _p(tdf), _stream(stream)))


def tdf_from_points_noalloc(pts, dims, origin, voxel_size, trunc, tdf, decode=DECODE_EXACT,
                            stream=None):
    """tdm_tdf_from_points_noalloc: the same TDF with the bit grid in the
    library's static per-GPU scratch (no allocation; calls on one GPU must not
    overlap)."""
    _check(lib().tdm_tdf_from_points_noalloc(_p(pts), c_int(pts.shape[0]), _int3(dims),
                                             _float3(origin), _f(voxel_size), _f(trunc),
                                             c_int(decode), _p(tdf), _stream(stream)))


NOALLOC_SCRATCH_BYTES = 16 * 1024 * 1024


def tdf_accurate(pts, dims, origin, voxel_size, trunc, tdf, stream=None) -> None:
    """Voxel-centre to nearest-point TDF (training/match.hpp:13-51)."""
    _check(lib().tdm_tdf_accurate(None, _p(pts), c_int(pts.shape[0]), _int3(dims),
                                  _float3(origin), _f(voxel_size), _f(trunc), _p(tdf),
                                  _stream(stream)))


class _CloudRef(ctypes.Structure):
    _fields_ = [("pts", c_void_p), ("tdf", c_void_p), ("num_pts", c_int), ("dim_x", c_int),
                ("dim_y", c_int), ("dim_z", c_int), ("origin_x", c_float), ("origin_y", c_float),
                ("origin_z", c_float), ("reserved", c_int)]


class TdfPlan:
    """tdm_tdf_plan: the TDF of many clouds in three launches per run
    (capturable in a hipGraph).  clouds: list of (pts tensor [n,3], tdf tensor,
    dims, origin); the tensors must outlive the plan."""

    def __init__(self, clouds, voxel_size=0.01, trunc=0.05, decode=DECODE_EXACT):
        arr = (_CloudRef * len(clouds))()
        for i, (pts, tdf, dims, origin) in enumerate(clouds):
            o = np.asarray(origin, np.float32)
            arr[i] = _CloudRef(pts.data_ptr() if pts.numel() else None, tdf.data_ptr(),
                               int(pts.shape[0]), int(dims[0]), int(dims[1]), int(dims[2]),
                               float(o[0]), float(o[1]), float(o[2]), 0)
        self._keep = clouds
        self._h = c_void_p()
        _check(lib().tdm_tdf_plan_create(None, arr, c_int(len(clouds)), _f(voxel_size), _f(trunc),
                                         c_int(decode), ctypes.byref(self._h)))

    def run(self, stream=None) -> None:
        _check(lib().tdm_tdf_plan_run(self._h, _stream(stream)))

    def close(self) -> None:
        if self._h:
            lib().tdm_tdf_plan_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def keypoints_to_grid(kp: np.ndarray, origin, voxel_size) -> np.ndarray:
    kp = np.ascontiguousarray(kp, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(kp.shape, dtype=np.int32)
    _check(lib().tdm_keypoints_to_grid(kp.ctypes.data_as(c_void_p), c_int(kp.shape[0]),
                                       _float3(origin), _f(voxel_size),
                                       out.ctypes.data_as(c_void_p)))
    return out


def extract_patches(tdf, dims, kp_grid, out, patch_size=30, stream=None) -> None:
    _check(lib().tdm_extract_patches(None, _p(tdf), c_int(dims[0]), c_int(dims[1]), c_int(dims[2]),
                                     _p(kp_grid), c_int(kp_grid.shape[0]), c_int(patch_size),
                                     _p(out), _stream(stream)))


# ---- registration -----------------------------------------------------------------


def match_nn(desc_a, desc_b, nn_ab=None, nn_ba=None, dist_ab=None, dist_ba=None, refined=None,
             stream=None) -> None:
    """Nearest neighbours both ways (knnsearch, core/demo.m:49-52) into the
    given device int32/float32 tensors; refined: device int32[2] or None."""
    na, dim = desc_a.shape
    nb = desc_b.shape[0]
    assert desc_b.shape[1] == dim
    _check(lib().tdm_match_nn(None, _p(desc_a), c_int(na), _p(desc_b), c_int(nb), c_int(dim),
                              _p(nn_ab), _p(dist_ab), _p(nn_ba), _p(dist_ba), _p(refined),
                              _stream(stream)))


def mutual_pairs(nn_ab, nn_ba, pairs, num_pairs, stream=None) -> None:
    """Mutual pairs (core/demo.m:53-55): pairs int32 [nb, 2], num_pairs int32 [1]."""
    _check(lib().tdm_mutual_pairs(None, _p(nn_ab), c_int(nn_ab.shape[0]), _p(nn_ba),
                                  c_int(nn_ba.shape[0]), _p(pairs), _p(num_pairs),
                                  _stream(stream)))


def ransac_rt(kp1, kp2, rt, inliers, stats, pairs=None, num_pairs=None, max_pairs=None,
              threshold=0.05, max_trials=50000, seed=0, trial_counts=None, stream=None) -> None:
    """ransacfitRt (core/external/sfm/ransacfitRt.m) on the device.
    rt: float64 [12]; inliers: int32 [max_pairs]; stats: int32 [4]."""
    if max_pairs is None:
        max_pairs = pairs.shape[0] if pairs is not None else kp1.shape[0]
    _check(lib().tdm_ransac_rt(None, _p(kp1), _p(kp2), _p(pairs), _p(num_pairs),
                               c_int(max_pairs), ctypes.c_double(threshold),
                               ctypes.c_long(max_trials), ctypes.c_uint64(seed), _p(rt),
                               _p(inliers), _p(stats), _p(trial_counts), _stream(stream)))


def estimate_rt(x1, x2, rt, stream=None) -> None:
    """estimateRt (core/external/sfm/estimateRt.m): rt float64 [12]."""
    _check(lib().tdm_estimate_rt(None, _p(x1), _p(x2), c_int(x1.shape[0]), _p(rt),
                                 _stream(stream)))


def alignment_ratio(pts1, pts2, rt, radius=0.05, stream=None):
    """ratioAligned of register2Fragments.m:90-97 on the device: (fraction of
    pts1 within radius of pts2 moved by rt, fraction of moved pts2 within
    radius of pts1).  pts1, pts2: device [n, 3] float; rt: host 3x4 float64."""
    r = np.ascontiguousarray(np.asarray(rt, np.float64).reshape(12))
    out = (ctypes.c_double * 2)()
    _check(lib().tdm_alignment_ratio(None, _p(pts1), c_int(pts1.shape[0]), _p(pts2),
                                     c_int(pts2.shape[0]), r.ctypes.data_as(c_void_p),
                                     _f(radius), out, _stream(stream)))
    return out[0], out[1]


class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", c_size_t), ("max_inflight", c_int), ("keypoints", c_int),
                ("max_pass_keypoints", c_int), ("voxel_size", c_float),
                ("trunc_margin", c_float), ("padding", c_int),
                ("ransac_threshold", ctypes.c_double), ("ransac_max_trials", ctypes.c_long)]


class _RegRequest(ctypes.Structure):
    _fields_ = [("points1", c_void_p), ("num_points1", c_int), ("points2", c_void_p),
                ("num_points2", c_int), ("keypoints1", c_void_p), ("keypoints2", c_void_p),
                ("num_keypoints", c_int), ("seed", ctypes.c_uint64)]


class _RegResult(ctypes.Structure):
    _fields_ = [("rt", ctypes.c_double * 12), ("status", c_int), ("num_matches", c_int),
                ("num_inliers", c_int), ("trials", c_int), ("fallback", c_int),
                ("pass_keypoints", c_int), ("queue_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double)]


class Server:
    """tdm_server: fragment-pair registration requests with continuous
    batching of descriptor passes and admission by a device-memory budget."""

    def __init__(self, net: "Net", device: int = 0, **cfg):
        c = ServerConfig()
        lib().tdm_server_config_init(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.config = c
        self._net = net  # the server drives it; keep it alive
        h = c_void_p()
        _check(lib().tdm_server_create(net._h, c_int(device), ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def submit(self, points1: np.ndarray, points2: np.ndarray, keypoints1=None, keypoints2=None,
               seed: int = 0) -> int:
        p1 = np.ascontiguousarray(points1, dtype=np.float32).reshape(-1, 3)
        p2 = np.ascontiguousarray(points2, dtype=np.float32).reshape(-1, 3)
        r = _RegRequest(p1.ctypes.data, len(p1), p2.ctypes.data, len(p2), None, None, 0, seed)
        keep = []
        if keypoints1 is not None:
            k1 = np.ascontiguousarray(keypoints1, dtype=np.int32)
            k2 = np.ascontiguousarray(keypoints2, dtype=np.int32)
            assert len(k1) == len(k2)
            r.keypoints1, r.keypoints2, r.num_keypoints = k1.ctypes.data, k2.ctypes.data, len(k1)
            keep = [k1, k2]
        t = ctypes.c_uint64()
        _check(lib().tdm_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        del keep
        return int(t.value)

    def wait(self, ticket: int) -> dict:
        res = _RegResult()
        _check(lib().tdm_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(res)))
        return {"rt": np.array(res.rt[:]).reshape(3, 4), "status": res.status,
                "num_matches": res.num_matches, "num_inliers": res.num_inliers,
                "trials": res.trials, "fallback": res.fallback,
                "pass_keypoints": res.pass_keypoints, "queue_ms": res.queue_ms,
                "latency_ms": res.latency_ms}

    def stats(self) -> dict:
        s = (ctypes.c_longlong * 6)()
        _check(lib().tdm_server_stats(self._h, s))
        keys = ["peak_bytes", "budget", "max_inflight", "passes", "completed", "max_pass_requests"]
        return dict(zip(keys, [int(v) for v in s]))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(lib().tdm_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid_table(grids, device="cuda:0"):
    """Device table of tdm_grid_ref entries for (tdf tensor, dims) pairs; the
    tensors must stay alive while the table is used."""
    import torch

    raw = np.zeros((len(grids), 3), dtype=np.int64)  # 24-byte entries
    for i, (t, d) in enumerate(grids):
        raw[i, 0] = t.data_ptr()
        dims = np.array([d[0], d[1], d[2], 0], dtype=np.int32)
        raw[i, 1:] = dims.view(np.int64)
    return torch.from_numpy(raw).to(device)


# ---- network --------------------------------------------------------------------


class Net:
    """tdm_net: Marvin JSON network + .marvin weights on one GPU."""

    def __init__(self, json_path: str, precision: int = PREC_FP32, max_batch: int = 0):
        self._h = c_void_p()
        _check(lib().tdm_net_create(None, json_path.encode(), c_int(precision), c_int(max_batch),
                                    ctypes.byref(self._h)))
        self.precision = precision
        self.max_batch = max_batch if max_batch > 0 else 256

    def close(self) -> None:
        if self._h:
            lib().tdm_net_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def output_dim(self) -> int:
        return lib().tdm_net_output_dim(self._h)

    @property
    def layer_names(self) -> list[str]:
        L = lib()
        return [L.tdm_net_layer_name(self._h, i).decode() for i in range(L.tdm_net_num_layers(self._h))]

    def layer_info(self, index: int) -> dict:
        info = (c_int * 8)()
        _check(lib().tdm_net_layer_info(self._h, c_int(index), info))
        keys = ("cin", "cout", "window", "in_extent", "out_extent", "relu", "pool", "final_extent")
        return dict(zip(keys, list(info)))

    def load_weights(self, marvin_path: str) -> None:
        _check(lib().tdm_net_load_weights(self._h, marvin_path.encode()))

    def set_layer(self, name: str, weight: np.ndarray, bias: np.ndarray) -> None:
        w = np.ascontiguousarray(weight, dtype=np.float32)
        b = np.ascontiguousarray(bias, dtype=np.float32)
        _check(lib().tdm_net_set_layer(self._h, name.encode(), w.ctypes.data_as(c_void_p),
                                       c_size_t(w.size), b.ctypes.data_as(c_void_p),
                                       c_size_t(b.size)))

    def set_profiling(self, enable: bool) -> None:
        _check(lib().tdm_net_set_profiling(self._h, c_int(1 if enable else 0)))

    def layer_stats(self, index: int):
        """(total ms, passes, patches) of layer `index` since profiling was enabled."""
        ms, n, p = ctypes.c_double(), ctypes.c_long(), ctypes.c_long()
        _check(lib().tdm_net_layer_stats(self._h, c_int(index), ctypes.byref(ms), ctypes.byref(n),
                                         ctypes.byref(p)))
        return ms.value, n.value, p.value

    def forward(self, data, feat, stream=None) -> None:
        """data: [n, 1, P, P, P] float device tensor; feat: [n, D] float."""
        _check(lib().tdm_net_forward(self._h, _p(data), c_int(data.shape[0]), _p(feat),
                                     _stream(stream)))

    def forward_grid(self, tdf, dims, kp_grid, feat, stream=None) -> None:
        _check(lib().tdm_net_forward_grid(self._h, _p(tdf), c_int(dims[0]), c_int(dims[1]),
                                          c_int(dims[2]), _p(kp_grid), c_int(kp_grid.shape[0]),
                                          _p(feat), _stream(stream)))

    def calibrate(self, ref: "Net", data, stream=None) -> list[float]:
        """fp8 nets: per-layer activation scales from `ref` (a float/16-bit
        net with the same weights) on patches `data` [n, 30, 30, 30]."""
        _check(lib().tdm_net_calibrate(self._h, ref._h, _p(data), c_int(data.shape[0]),
                                       _stream(stream)))
        return [self.act_scale(i) for i in range(len(self.layer_names))]

    def act_scale(self, index: int) -> float:
        v = c_float()
        _check(lib().tdm_net_act_scale(self._h, c_int(index), ctypes.byref(v)))
        return float(v.value)

    def set_act_scale(self, index: int, scale: float) -> None:
        _check(lib().tdm_net_set_act_scale(self._h, c_int(index), _f(scale)))

    def autotune(self, num: int = 0, stream=None) -> list[int]:
        """Times each convolution's tile variants and keeps the fastest;
        returns the chosen variant per layer."""
        _check(lib().tdm_net_autotune(self._h, c_int(num or self.max_batch), _stream(stream)))
        return [self.layer_variant(i) for i in range(len(self.layer_names))]

    def layer_variant(self, index: int) -> int:
        return int(lib().tdm_net_layer_variant(self._h, c_int(index)))

    def set_layer_variant(self, index: int, variant: int) -> None:
        _check(lib().tdm_net_set_layer_variant(self._h, c_int(index), c_int(variant)))

    def forward_grids(self, table, kp_grid, grid_index, feat, stream=None) -> None:
        """table: device tensor from grid_table(); kp_grid int32 [n, 3] and
        grid_index int32 [n] on the device."""
        _check(lib().tdm_net_forward_grids(self._h, _p(table), _p(kp_grid), _p(grid_index),
                                           c_int(kp_grid.shape[0]), _p(feat), _stream(stream)))

    def forward_layer(self, data, layer_index: int, out, stream=None) -> None:
        _check(lib().tdm_net_forward_layer(self._h, _p(data), c_int(data.shape[0]),
                                           c_int(layer_index), _p(out), c_size_t(out.numel()),
                                           _stream(stream)))


# ---- formats (host) ----------------------------------------------------------------


def _take_floats(ptr, count) -> np.ndarray:
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, POINTER(c_float)), shape=(count,)).copy()
    lib().tdm_free(ptr)
    return arr


def read_ply_points(path: str) -> np.ndarray:
    ptr, n = c_void_p(), c_int()
    _check(lib().tdm_read_ply_points(path.encode(), ctypes.byref(ptr), ctypes.byref(n)))
    return _take_floats(ptr, n.value * 3).reshape(-1, 3)


def write_ply_points(path: str, pts: np.ndarray) -> None:
    pts = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    _check(lib().tdm_write_ply_points(path.encode(), pts.ctypes.data_as(c_void_p), c_int(len(pts))))


def write_keypts_bin(path: str, kp: np.ndarray) -> None:
    kp = np.ascontiguousarray(kp, dtype=np.float32).reshape(-1, 3)
    _check(lib().tdm_write_keypts_bin(path.encode(), kp.ctypes.data_as(c_void_p), c_int(len(kp))))


def read_keypts_bin(path: str) -> np.ndarray:
    ptr, n = c_void_p(), c_int()
    _check(lib().tdm_read_keypts_bin(path.encode(), ctypes.byref(ptr), ctypes.byref(n)))
    return _take_floats(ptr, n.value * 3).reshape(-1, 3)


def write_desc_bin(path: str, desc: np.ndarray) -> None:
    desc = np.ascontiguousarray(desc, dtype=np.float32)
    _check(lib().tdm_write_desc_bin(path.encode(), desc.ctypes.data_as(c_void_p),
                                    c_int(desc.shape[0]), c_int(desc.shape[1])))


def read_desc_bin(path: str) -> np.ndarray:
    ptr, n, d = c_void_p(), c_int(), c_int()
    _check(lib().tdm_read_desc_bin(path.encode(), ctypes.byref(ptr), ctypes.byref(n), ctypes.byref(d)))
    return _take_floats(ptr, n.value * d.value).reshape(n.value, d.value)


def write_grid_bin(path: str, grid: np.ndarray, origin, voxel_size, trunc) -> None:
    """grid: [dz, dy, dx] float32 (x fastest)."""
    g = np.ascontiguousarray(grid, dtype=np.float32)
    dims = (g.shape[2], g.shape[1], g.shape[0])
    _check(lib().tdm_write_grid_bin(path.encode(), _int3(dims), _float3(origin), _f(voxel_size),
                                    _f(trunc), g.ctypes.data_as(c_void_p)))


def read_grid_bin(path: str):
    dims = (c_int * 3)()
    origin = (c_float * 3)()
    vs, tr, ptr = c_float(), c_float(), c_void_p()
    _check(lib().tdm_read_grid_bin(path.encode(), dims, origin, ctypes.byref(vs), ctypes.byref(tr),
                                   ctypes.byref(ptr)))
    dx, dy, dz = list(dims)
    grid = _take_floats(ptr, dx * dy * dz).reshape(dz, dy, dx)
    return grid, np.array(list(origin), dtype=np.float32), vs.value, tr.value


def read_patch_volumes(path: str) -> np.ndarray:
    ptr, n = c_void_p(), c_int()
    dims = (c_int * 3)()
    _check(lib().tdm_read_patch_volumes(path.encode(), ctypes.byref(ptr), ctypes.byref(n), dims))
    d = list(dims)
    return _take_floats(ptr, n.value * d[0] * d[1] * d[2]).reshape(n.value, d[2], d[1], d[0])


def write_feat_bin(path: str, feat: np.ndarray) -> None:
    feat = np.ascontiguousarray(feat, dtype=np.float32)
    _check(lib().tdm_write_feat_bin(path.encode(), feat.ctypes.data_as(c_void_p),
                                    c_int(feat.shape[0]), c_int(feat.shape[1])))


def load_matrix(path: str, rows: int, cols: int) -> np.ndarray:
    out = np.zeros(rows * cols, dtype=np.float32)
    _check(lib().tdm_load_matrix(path.encode(), c_int(rows), c_int(cols), out.ctypes.data_as(c_void_p)))
    return out.reshape(rows, cols)


def read_depth_png(path: str, height: int = 480, width: int = 640) -> np.ndarray:
    out = np.zeros(height * width, dtype=np.float32)
    _check(lib().tdm_read_depth_png(path.encode(), c_int(height), c_int(width),
                                    out.ctypes.data_as(c_void_p)))
    return out.reshape(height, width)


def read_png_u16(path: str) -> np.ndarray:
    ptr, h, w = c_void_p(), c_int(), c_int()
    _check(lib().tdm_read_png_u16(path.encode(), ctypes.byref(ptr), ctypes.byref(h), ctypes.byref(w)))
    arr = np.ctypeslib.as_array(ctypes.cast(ptr, POINTER(ctypes.c_uint16)),
                                shape=(h.value * w.value,)).copy()
    lib().tdm_free(ptr)
    return arr.reshape(h.value, w.value)


def multiply_matrix(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(16)
    b = np.ascontiguousarray(b, dtype=np.float32).reshape(16)
    out = np.zeros(16, dtype=np.float32)
    _check(lib().tdm_multiply_matrix(a.ctypes.data_as(c_void_p), b.ctypes.data_as(c_void_p),
                                     out.ctypes.data_as(c_void_p)))
    return out.reshape(4, 4)


def invert_matrix(m: np.ndarray) -> np.ndarray:
    m = np.ascontiguousarray(m, dtype=np.float32).reshape(16)
    out = np.zeros(16, dtype=np.float32)
    _check(lib().tdm_invert_matrix(m.ctypes.data_as(c_void_p), out.ctypes.data_as(c_void_p)))
    return out.reshape(4, 4)


def save_surface_ply(path: str, tsdf: np.ndarray, weight: np.ndarray, origin, voxel_size,
                     tsdf_thresh=0.2, weight_thresh=1.0) -> int:
    t = np.ascontiguousarray(tsdf, dtype=np.float32)
    w = np.ascontiguousarray(weight, dtype=np.float32)
    dims = (t.shape[2], t.shape[1], t.shape[0])
    n = c_int()
    _check(lib().tdm_save_surface_ply(path.encode(), _int3(dims), _f(voxel_size), _float3(origin),
                                      t.ctypes.data_as(c_void_p), w.ctypes.data_as(c_void_p),
                                      _f(tsdf_thresh), _f(weight_thresh), ctypes.byref(n)))
    return n.value


def marvin_write(path: str, tensors: list[tuple[str, np.ndarray]]) -> None:
    """Float tensors as a .marvin file (name, array) in order."""
    n = len(tensors)
    names = (c_char_p * n)(*[name.encode() for name, _ in tensors])
    arrays = [np.ascontiguousarray(a, dtype=np.float32) for _, a in tensors]
    ndims = (c_int * n)(*[a.ndim for a in arrays])
    dim_arrays = [(c_int * max(a.ndim, 1))(*a.shape) for a in arrays]
    dims = (POINTER(c_int) * n)(*[ctypes.cast(d, POINTER(c_int)) for d in dim_arrays])
    data = (c_void_p * n)(*[a.ctypes.data for a in arrays])
    _check(lib().tdm_marvin_write(path.encode(), c_int(n), names, ndims, dims, data))


def marvin_count(path: str):
    nt, nv = c_int(), c_size_t()
    _check(lib().tdm_marvin_count(path.encode(), ctypes.byref(nt), ctypes.byref(nv)))
    return nt.value, nv.value
