"""Seeded synthetic inputs for the fusion and descriptor workloads.

Scene: the inside of an axis-aligned room [-2,2] x [-2,2] x [0,3] m (z up)
holding 20 spheres of radius U[0.1, 0.5] m.  Cameras sit on a horizontal
circle of radius 1.5 m at height 1.5 m around the room axis and look at
(0, 0, 1.5).  Depth frames are ray-cast z-depth through pixel centres
((u - cx)/fx, (v - cy)/fy, 1), the projection Integrate inverts
(depth-fusion/demo.cu:46-47), with noise N(0, (0.0015 z^2)^2) m, 5 % dropouts
and millimetre quantisation to uint16, as the sample frames store depth
(depth-fusion/utils.hpp:74-86).
"""
from __future__ import annotations

import numpy as np

FX = FY = 585.0
CX, CY = 320.0, 240.0
H, W = 480, 640
ROOM_LO = np.array([-2.0, -2.0, 0.0])
ROOM_HI = np.array([2.0, 2.0, 3.0])


def intrinsics() -> np.ndarray:
    return np.array([[FX, 0, CX], [0, FY, CY], [0, 0, 1]], dtype=np.float32)


def spheres(seed: int = 1, count: int = 20, cam_radius: float = 1.5, cam_height: float = 1.5):
    """[count, 4] (centre xyz, radius); spheres stay inside the room and off the
    camera circle."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        r = rng.uniform(0.1, 0.5)
        c = rng.uniform(ROOM_LO + r, ROOM_HI - r)
        # distance from the centre to the camera circle
        rho = np.hypot(c[0], c[1])
        d_circle = np.hypot(rho - cam_radius, c[2] - cam_height)
        if d_circle > r + 0.1:
            out.append([c[0], c[1], c[2], r])
    return np.array(out)


def camera_poses(n: int = 50, radius: float = 1.5, height: float = 1.5) -> np.ndarray:
    """[n, 4, 4] float64 camera-to-world poses (x right, y down, z forward)."""
    poses = np.zeros((n, 4, 4))
    up = np.array([0.0, 0.0, 1.0])
    for i in range(n):
        th = 2 * np.pi * i / n
        c = np.array([radius * np.cos(th), radius * np.sin(th), height])
        f = np.array([0.0, 0.0, height]) - c
        f /= np.linalg.norm(f)
        x = np.cross(f, up)
        x /= np.linalg.norm(x)
        y = np.cross(f, x)
        poses[i, :3, 0], poses[i, :3, 1], poses[i, :3, 2], poses[i, :3, 3] = x, y, f, c
        poses[i, 3, 3] = 1.0
    return poses


def _pixel_rays():
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    return np.stack([(u - CX) / FX, (v - CY) / FY, np.ones_like(u)], axis=-1).reshape(-1, 3)


def raycast_depth(pose: np.ndarray, balls: np.ndarray) -> np.ndarray:
    """Noise-free z-depth [H, W] in metres of one camera."""
    d_cam = _pixel_rays()
    R, c = pose[:3, :3], pose[:3, 3]
    d = d_cam @ R.T
    with np.errstate(divide="ignore", invalid="ignore"):
        t_hi = np.where(d > 0, (ROOM_HI - c) / d, np.where(d < 0, (ROOM_LO - c) / d, np.inf))
    t = t_hi.min(axis=1)
    dd = np.einsum("ij,ij->i", d, d)
    for s in balls:
        oc = c - s[:3]
        b = d @ oc
        disc = b * b - dd * (oc @ oc - s[3] ** 2)
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0.0))
        t0 = (-b - sq) / dd
        t1 = (-b + sq) / dd
        hit = np.where(t0 > 1e-6, t0, np.where(t1 > 1e-6, t1, np.inf))
        t = np.where(ok, np.minimum(t, hit), t)
    return t.reshape(H, W)  # d_cam has z = 1, so the ray parameter is z-depth


def depth_frames(n: int = 50, scene_seed: int = 1, noise_seed: int = 2,
                 dropout: float = 0.05, count: int | None = None):
    """(depth_mm [count, H, W] uint16, cam2world [count, 4, 4] float64): the
    first `count` (default n) of n cameras spaced 2*pi/n apart."""
    balls = spheres(scene_seed)
    poses = camera_poses(n)
    count = n if count is None else count
    poses = poses[:count]
    rng = np.random.default_rng(noise_seed)
    frames = np.zeros((count, H, W), dtype=np.uint16)
    for i in range(count):
        z = raycast_depth(poses[i], balls)
        z = z + rng.standard_normal(z.shape) * (0.0015 * z * z)
        z[rng.random(z.shape) < dropout] = 0.0
        frames[i] = np.clip(np.round(z * 1000.0), 0, 65535).astype(np.uint16)
    return frames, poses


def relative_poses(cam2world: np.ndarray, base: int = 0) -> np.ndarray:
    """cam2base = inv(base2world) @ cam2world, float32 [n, 4, 4]."""
    inv = np.linalg.inv(cam2world[base])
    return np.einsum("ij,njk->nik", inv, cam2world).astype(np.float32)


def depth_to_metres(mm: np.ndarray) -> np.ndarray:
    """(float)mm / 1000.0f with > 6 m set to 0 (depth-fusion/utils.hpp:80-84)."""
    d = mm.astype(np.float32) / np.float32(1000.0)
    d[d > np.float32(6.0)] = 0.0
    return d


# Fusion grid of the TSDF workload: 256^3 at 1 cm in the first camera's frame,
# centred 1.5 m ahead of it (on the room centre).
TSDF_DIMS = (256, 256, 256)
TSDF_ORIGIN = np.array([-1.28, -1.28, 0.22], dtype=np.float32)
TSDF_VOXEL = np.float32(0.01)
TSDF_TRUNC = np.float32(0.05)


def fragment(n_points: int = 200_000, view: int = 0, n_views: int = 50, scene_seed: int = 1,
             noise_seed: int = 2, seed: int = 3, sigma: float = 0.002) -> np.ndarray:
    """[n_points, 3] float32 cloud: one view back-projected in its camera frame
    (core/demo.m's recipe), resampled and jittered with N(0, sigma^2)."""
    frames, _ = depth_frames(n_views, scene_seed, noise_seed, count=view + 1)
    z = frames[view].astype(np.float64) / 1000.0
    v, u = np.nonzero(z > 0)
    zz = z[v, u]
    pts = np.stack([(u - CX) / FX * zz, (v - CY) / FY * zz, zz], axis=1)
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(pts), size=n_points, replace=len(pts) < n_points)
    pts = pts[idx] + rng.standard_normal((n_points, 3)) * sigma
    return pts.astype(np.float32)


def keypoint_indices(n_points: int, count: int = 5000, seed: int = 4) -> np.ndarray:
    """Uniform random point indices: floor(U[0, N - 1e-4)) (core/demo.cu:169-170,
    core/utils.hpp:15-20), from a seeded generator."""
    rng = np.random.default_rng(seed)
    return np.floor(rng.uniform(0.0, n_points - 1e-4, size=count)).astype(np.int64)


_VIEW_CACHE: dict = {}


def _view_points(view: int, scene_seed: int) -> np.ndarray:
    """Noise-free back-projected points [N, 3] (camera frame) of one view."""
    key = (view, scene_seed)
    if key not in _VIEW_CACHE:
        z = raycast_depth(camera_poses(50)[view], spheres(scene_seed))
        v, u = np.nonzero(np.isfinite(z) & (z > 0) & (z < 6))
        zz = z[v, u]
        _VIEW_CACHE[key] = np.stack([(u - CX) / FX * zz, (v - CY) / FY * zz, zz], axis=1)
    return _VIEW_CACHE[key]


def fragment_pair(seed: int, n1: int | None = None, n2: int | None = None,
                  overlap: float | None = None, max_angle: float = np.pi / 3,
                  max_shift: float = 1.0, sigma: float = 0.002, scene_seed: int = 1):
    """One fragment-pair registration request: two parts of a ray-cast view
    that share `overlap` of each part's extent (default U[0.5, 0.8]), each
    resampled to its point count (default U[20k, 500k]) with N(0, sigma^2)
    noise; fragment 2 is then moved by a rigid motion (angle U[0, max_angle]
    about a uniform axis, shift of length U[0, max_shift]).  Returns
    (pts1, pts2, R, t) with pts1 = R pts2 + t on the shared surface."""
    rng = np.random.default_rng(seed)
    P = _view_points(int(rng.integers(0, 50)), scene_seed)
    o = rng.uniform(0.5, 0.8) if overlap is None else overlap
    a = 1.0 / (2.0 - o)  # each part spans a quantile range of width a; they share 2a - 1
    key = P[:, 0]
    lo, hi = np.quantile(key, [1.0 - a, a])
    parts = [P[key <= hi], P[key >= lo]]
    counts = [n1, n2]
    out = []
    for part, n in zip(parts, counts):
        n = int(rng.integers(20_000, 500_001)) if n is None else n
        idx = rng.choice(len(part), size=n, replace=len(part) < n)
        out.append(part[idx] + rng.standard_normal((n, 3)) * sigma)
    axis = rng.standard_normal(3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0.0, max_angle)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
    d = rng.standard_normal(3)
    t = d / np.linalg.norm(d) * rng.uniform(0.0, max_shift)
    pts2 = (out[1] - t) @ R  # R^T (x - t) for row vectors
    return out[0].astype(np.float32), pts2.astype(np.float32), R, t


def cropped_fragment(seed: int, n_points: int, extent: float, sigma: float = 0.002,
                     scene_seed: int = 1) -> np.ndarray:
    """[n_points, 3] float32 cloud: the points of a random view inside a cube
    of side `extent` around a random surface point, resampled to n_points
    with N(0, sigma^2) noise (fragments of the batched workload)."""
    rng = np.random.default_rng(seed)
    P = _view_points(int(rng.integers(0, 50)), scene_seed)
    for _ in range(20):
        c = P[rng.integers(0, len(P))]
        keep = np.all(np.abs(P - c) <= extent / 2, axis=1)
        if keep.sum() >= 2000:
            break
    part = P[keep]
    idx = rng.choice(len(part), size=n_points, replace=len(part) < n_points)
    return (part[idx] + rng.standard_normal((n_points, 3)) * sigma).astype(np.float32)


def depth_frames_torch(torch, n: int = 50, scene_seed: int = 1, noise_seed: int = 2,
                       dropout: float = 0.05, device="cuda:0"):
    """depth_frames on the GPU with torch (same scene, cameras, noise model
    and millimetre quantisation; its own random stream): (depth in metres
    [n, H, W] float32 as tdm_depth_from_mm would give, cam2world [n, 4, 4]
    float64 numpy).  For large synthetic batches."""
    balls = torch.tensor(spheres(scene_seed), dtype=torch.float64, device=device)
    poses = camera_poses(n)
    rays = torch.tensor(_pixel_rays(), dtype=torch.float64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(noise_seed)
    lo = torch.tensor(ROOM_LO, dtype=torch.float64, device=device)
    hi = torch.tensor(ROOM_HI, dtype=torch.float64, device=device)
    out = torch.empty(n, H, W, dtype=torch.float32, device=device)
    for i in range(n):
        R = torch.tensor(poses[i, :3, :3], dtype=torch.float64, device=device)
        c = torch.tensor(poses[i, :3, 3], dtype=torch.float64, device=device)
        d = rays @ R.T
        big = torch.full_like(d, float("inf"))
        t_hi = torch.where(d > 0, (hi - c) / d, torch.where(d < 0, (lo - c) / d, big))
        t = t_hi.min(dim=1).values
        dd = (d * d).sum(1)
        for s in balls:
            oc = c - s[:3]
            b = d @ oc
            disc = b * b - dd * ((oc * oc).sum() - s[3] ** 2)
            ok = disc >= 0
            sq = torch.sqrt(torch.clamp(disc, min=0.0))
            t0 = (-b - sq) / dd
            t1 = (-b + sq) / dd
            hit = torch.where(t0 > 1e-6, t0, torch.where(t1 > 1e-6, t1, torch.full_like(t0, float("inf"))))
            t = torch.where(ok, torch.minimum(t, hit), t)
        z = t.reshape(H, W)
        z = z + torch.randn(z.shape, generator=gen, device=device, dtype=torch.float64) * (0.0015 * z * z)
        z = torch.where(torch.rand(z.shape, generator=gen, device=device) < dropout, torch.zeros_like(z), z)
        mm = torch.clamp(torch.round(z * 1000.0), 0, 65535)
        m = (mm.float() / 1000.0)
        out[i] = torch.where(m > 6.0, torch.zeros_like(m), m)
    return out, poses


def write_png16(path: str, mm: np.ndarray) -> None:
    """16-bit grayscale PNG (filter 0 on every row) of a [H, W] uint16 image,
    the format of the reference's frame-%06d.depth.png files."""
    import struct
    import zlib

    h, w = mm.shape
    raw = b"".join(b"\x00" + mm[r].astype(">u2").tobytes() for r in range(h))

    def chunk(tag, data):
        c = struct.pack(">I", len(data)) + tag + data
        return c + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 16, 0, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))

