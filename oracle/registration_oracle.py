"""CPU restatement of the registration back end: descriptor nearest
neighbours both ways, the mutual filter, estimateRt and ransacfitRt.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's CPU baseline)
as the checker.  The product path never imports it.

Reference behaviour restated here:
- knnsearch(KDTreeSearcher(B), A) both ways and the mutual filter:
  core/demo.m:49-56, evaluation/geometric-registration/register2Fragments.m:
  49-55.  A KD-tree search in double is exact, so the oracle is the exact
  argmin of the squared distance (ties to the lower index).
- estimateRt -> estimateRigidTransform: core/external/sfm/estimateRt.m,
  core/external/estimateRigidTransform/estimateRigidTransform.m:43-72
  (Horn's quaternion: B = sum A_i' A_i, quat = V(:,4) of svd(B)),
  quat2rot.m, crossTimesMatrix.m.
- ransacfitRt: core/external/sfm/ransacfitRt.m:20-64 (s = 3, euc3Ddist with
  `abs(d) < t`, refit on the inliers), core/external/peter/ransac.m:113-231
  (p = 0.99, `>=` update, N = max(10, log(1-p)/log(pNoOutliers)), stop when
  trialcount > maxTrials).  The trial samples come from the counter-based
  generator below (MATLAB's RandStream cannot be reproduced).
- ratioAligned: evaluation/geometric-registration/register2Fragments.m:68-69,
  90-97 (fragment 2 moved by the estimate in single precision, then the
  fraction of each fragment's points with a point of the other closer than
  0.05).  MATLAB's sgemm order and its KD-tree's distance arithmetic are not
  reproducible; the restatement fixes a left-to-right float32 order.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

EPS = 2.220446049250313e-16  # MATLAB eps = 2^-52
MASK64 = (1 << 64) - 1


# ---- nearest neighbours ------------------------------------------------------------


def nn_exact(Q: np.ndarray, R: np.ndarray, chunk: int = 512):
    """For each row of Q the index of the nearest row of R (exact squared
    distance, ties to the lower index) and the Euclidean distance.

    Candidates are screened with the fp64 expansion |q|^2 + |r|^2 - 2 q.r,
    whose error is far below the 1e-9 relative screen; the survivors are
    compared with direct fp64 sums of squared differences."""
    Q = np.asarray(Q, np.float64)
    R = np.asarray(R, np.float64)
    qq = (Q * Q).sum(1)
    rr = (R * R).sum(1)
    rmax = math.sqrt(rr.max())
    idx = np.empty(len(Q), np.int64)
    dist = np.empty(len(Q), np.float64)
    for c0 in range(0, len(Q), chunk):
        q = Q[c0:c0 + chunk]
        D = qq[c0:c0 + chunk, None] + rr[None, :] - 2.0 * (q @ R.T)
        m = D.min(1)
        tol = 1e-9 * (np.sqrt(qq[c0:c0 + chunk]) + rmax) ** 2
        for k in range(len(q)):
            cand = np.nonzero(D[k] <= m[k] + 2.0 * tol[k])[0]
            diff = R[cand] - q[k]
            ex = (diff * diff).sum(1)
            w = int(np.argmin(ex))  # first of equal values: lowest index
            idx[c0 + k] = cand[w]
            dist[c0 + k] = math.sqrt(ex[w])
    return idx, dist


def nn_fraction(Q: np.ndarray, R: np.ndarray) -> np.ndarray:
    """Brute force in exact rational arithmetic (tiny inputs only)."""
    Qf = [[Fraction(float(v)) for v in row] for row in Q]
    Rf = [[Fraction(float(v)) for v in row] for row in R]
    out = []
    for q in Qf:
        best, bi = None, -1
        for j, r in enumerate(Rf):
            d = sum((a - b) * (a - b) for a, b in zip(q, r))
            if best is None or d < best:
                best, bi = d, j
        out.append(bi)
    return np.array(out, np.int64)


def mutual_pairs(nn_ab: np.ndarray, nn_ba: np.ndarray) -> np.ndarray:
    """(nn_ba[j], j) for every j with nn_ab[nn_ba[j]] == j, ascending j
    (core/demo.m:53-55)."""
    j = np.arange(len(nn_ba))
    keep = nn_ab[nn_ba] == j
    return np.stack([nn_ba[keep], j[keep]], 1).astype(np.int64)


# ---- rigid fit ---------------------------------------------------------------------


def quat2rot(q: np.ndarray) -> np.ndarray:
    """quat2rot.m for a batch of quaternions [..., 4] (q0 scalar)."""
    q0, q1, q2, q3 = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3
    R[..., 0, 1] = 2 * (q1 * q2 - q0 * q3)
    R[..., 0, 2] = 2 * (q1 * q3 + q0 * q2)
    R[..., 1, 0] = 2 * (q1 * q2 + q0 * q3)
    R[..., 1, 1] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3
    R[..., 1, 2] = 2 * (q2 * q3 - q0 * q1)
    R[..., 2, 0] = 2 * (q1 * q3 - q0 * q2)
    R[..., 2, 1] = 2 * (q2 * q3 + q0 * q1)
    R[..., 2, 2] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3
    return R


def estimate_rt(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """[R|t] (3x4) with x = R y + t, following estimateRigidTransform.m:
    centroids, A_i = [0, (y_i - x_i)'; x_i - y_i, [y_i + x_i]x] on centred
    points, B = sum A_i' A_i, quat = V(:,4) of svd(B), T = T3 T2 T1.
    Batched over leading axes: x, y [..., n, 3]."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = x.shape[-2]
    xc = x.sum(-2) / n
    yc = y.sum(-2) / n
    xs = x - xc[..., None, :]
    ys = y - yc[..., None, :]
    d = ys - xs
    s = ys + xs
    A = np.zeros(x.shape[:-1] + (4, 4))
    A[..., 0, 1:] = d
    A[..., 1:, 0] = -d
    A[..., 1, 2] = -s[..., 2]
    A[..., 1, 3] = s[..., 1]
    A[..., 2, 1] = s[..., 2]
    A[..., 2, 3] = -s[..., 0]
    A[..., 3, 1] = -s[..., 1]
    A[..., 3, 2] = s[..., 0]
    B = np.einsum("...nji,...njk->...ik", A, A)
    _, _, Vh = np.linalg.svd(B)
    quat = Vh[..., 3, :]
    rot = quat2rot(quat)
    t = xc - np.einsum("...ij,...j->...i", rot, yc)
    return np.concatenate([rot, t[..., :, None]], -1)


def kabsch(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Least-squares proper rotation by the SVD of the cross-covariance
    (Kabsch/Umeyama): an independent closed form of the same optimum."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    xc, yc = x.mean(0), y.mean(0)
    H = (y - yc).T @ (x - xc)
    U, _, Vt = np.linalg.svd(H)
    D = np.diag([1.0, 1.0, np.sign(np.linalg.det(Vt.T @ U.T))])
    R = Vt.T @ D @ U.T
    return np.concatenate([R, (xc - R @ yc)[:, None]], 1)


# ---- RANSAC ----------------------------------------------------------------------


def mix64(z: np.ndarray) -> np.ndarray:
    """splitmix64 output function on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def sample3(seed: int, trials: int, n: int) -> np.ndarray:
    """Three distinct indices of [0, n) per trial: r_k = mix64(mix64(seed) +
    3 t + k); i0 = r0 % n, i1 = r1 % (n-1) skipping i0, i2 = r2 % (n-2)
    skipping both."""
    base = int(mix64(np.array([seed & MASK64], np.uint64))[0])
    t = np.arange(trials, dtype=np.uint64)
    with np.errstate(over="ignore"):
        b = np.uint64(base) + np.uint64(3) * t
        r0, r1, r2 = mix64(b), mix64(b + np.uint64(1)), mix64(b + np.uint64(2))
    i0 = (r0 % np.uint64(n)).astype(np.int64)
    i1 = (r1 % np.uint64(n - 1)).astype(np.int64)
    i1 += i1 >= i0
    lo, hi = np.minimum(i0, i1), np.maximum(i0, i1)
    i2 = (r2 % np.uint64(n - 2)).astype(np.int64)
    i2 += i2 >= lo
    i2 += i2 >= hi
    return np.stack([i0, i1, i2], 1)


def inlier_mask(rt: np.ndarray, x: np.ndarray, y: np.ndarray, thr: float) -> np.ndarray:
    """euc3Ddist (ransacfitRt.m:40-62): sqrt(sum((x - (R y + t)).^2)) < t,
    for rt [T, 3, 4] against all pairs -> [T, n] booleans.  Evaluated as
    ((R0 y0 + R1 y1) + R2 y2) + t and ((e0^2 + e1^2) + e2^2)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    s = None
    for r in range(3):
        p = (rt[:, r, 0, None] * y[None, :, 0] + rt[:, r, 1, None] * y[None, :, 1]
             + rt[:, r, 2, None] * y[None, :, 2] + rt[:, r, 3, None])
        e = x[None, :, r] - p
        s = e * e if s is None else s + e * e
    return np.sqrt(s) < thr


def trial_counts(x, y, thr, max_trials, seed, chunk=2048):
    """Inlier count of every trial 0..max_trials and the trials' fits."""
    n = len(x)
    T = max_trials + 1
    idx = sample3(seed, T, n)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    counts = np.empty(T, np.int64)
    fits = np.empty((T, 3, 4))
    for c0 in range(0, T, chunk):
        sl = idx[c0:c0 + chunk]
        rt = estimate_rt(x[sl], y[sl])
        fits[c0:c0 + chunk] = rt
        counts[c0:c0 + chunk] = inlier_mask(rt, x, y, thr).sum(1)
    return counts, fits


def ransac_replay(counts, n: int, max_trials: int, p: float = 0.99):
    """The trial loop of ransac.m:125-221 over precomputed inlier counts:
    returns (kept trial, its inlier count, trials run)."""
    bestscore = 0
    best = None
    N = 1.0
    trialcount = 0
    while N > trialcount:
        ninliers = int(counts[trialcount])
        if ninliers >= bestscore:
            bestscore = ninliers
            best = trialcount
            frac = ninliers / n
            pno = 1.0 - frac ** 3
            pno = max(EPS, pno)
            pno = min(1.0 - EPS, pno)
            N = math.log(1.0 - p) / math.log(pno)
            N = max(N, 10.0)
        trialcount += 1
        if trialcount > max_trials:
            break
    return best, bestscore, trialcount


def ransac_fit_rt(x, y, thr=0.05, max_trials=50000, seed=0):
    """ransacfitRt(x, thr, 0) with this module's samples.  Returns
    (rt 3x4, inlier indices, kept trial, trials run, counts)."""
    n = len(x)
    if n < 3:
        return np.hstack([np.eye(3), np.zeros((3, 1))]), np.zeros(0, np.int64), -1, 0, None
    if n == 3:
        return estimate_rt(x, y), np.arange(3), -1, 0, None
    counts, fits = trial_counts(x, y, thr, max_trials, seed)
    best, _, run = ransac_replay(counts, n, max_trials)
    inl = np.nonzero(inlier_mask(fits[best:best + 1], x, y, thr)[0])[0]
    return estimate_rt(np.asarray(x)[inl], np.asarray(y)[inl]), inl, best, run, counts


def random_rotation(rng: np.random.Generator, max_angle=None) -> np.ndarray:
    """Rotation about a uniform axis by a uniform angle in [0, max_angle]
    (or a uniformly random rotation when max_angle is None)."""
    if max_angle is None:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        return quat2rot(q)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = rng.uniform(0, max_angle)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K


def move_points(pts: np.ndarray, rt: np.ndarray) -> np.ndarray:
    """Fragment 2 moved by [R|t] in float32 (register2Fragments.m:68: MATLAB
    multiplies the double matrix by the single cloud in single), each
    coordinate ((r0*x + r1*y) + r2*z) + t without contraction."""
    R = np.asarray(rt, np.float64).reshape(3, 4)[:, :3].astype(np.float32)
    t = np.asarray(rt, np.float64).reshape(3, 4)[:, 3].astype(np.float32)
    p = np.asarray(pts, np.float32)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([((R[i, 0] * x + R[i, 1] * y) + R[i, 2] * z) + t[i] for i in range(3)], 1)


def fraction_within(queries: np.ndarray, targets: np.ndarray, radius: float) -> float:
    """Fraction of queries whose nearest target is closer than radius, the
    distance being sqrtf((dx*dx + dy*dy) + dz*dz) in float32
    (register2Fragments.m:91-96 with multiQueryKNNSearchImpl's 1-NN).  A
    float64 KD-tree (scipy cKDTree) settles the clear cases; queries whose
    nearest float64 distance lies within 1e-5 of the radius are decided by the
    float32 distance to every target in reach."""
    from scipy.spatial import cKDTree

    q = np.asarray(queries, np.float32)
    t = np.asarray(targets, np.float32)
    r32 = np.float32(radius)
    tree = cKDTree(t.astype(np.float64))
    d, _ = tree.query(q.astype(np.float64), k=1, workers=-1)
    slack = 1e-5 * max(1.0, float(radius))
    hit = d < float(r32) - slack
    unsure = np.nonzero(np.abs(d - float(r32)) <= slack)[0]
    if len(unsure):
        cands = tree.query_ball_point(q[unsure].astype(np.float64), r=float(r32) + 2 * slack,
                                       workers=-1)
        for k, i in enumerate(unsure):
            c = np.asarray(cands[k], dtype=np.int64)
            if len(c) == 0:
                continue
            dd = q[i][None, :] - t[c]
            d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
            hit[i] = bool(np.any(np.sqrt(d2) < r32))
    return float(np.count_nonzero(hit)) / len(q)


def fraction_within_brute(queries, targets, radius) -> float:
    """The same by brute force over every pair in float32 (small inputs)."""
    q = np.asarray(queries, np.float32)
    t = np.asarray(targets, np.float32)
    hits = 0
    for i in range(len(q)):
        dd = q[i][None, :] - t
        d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
        hits += bool(np.any(np.sqrt(d2) < np.float32(radius)))
    return hits / len(q)


def alignment_ratio(pts1, pts2, rt, radius=0.05):
    """ratioAligned of register2Fragments.m:90-97."""
    moved = move_points(pts2, rt)
    return fraction_within(pts1, moved, radius), fraction_within(moved, pts1, radius)

