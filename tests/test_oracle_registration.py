"""Pins of the registration oracle (nearest neighbours, mutual filter, rigid
fit, RANSAC loop) against exact arithmetic, an independent library search,
closed forms and hand-derived trial counts.

The reference's own implementation (MATLAB knnsearch / ransac.m) cannot run
here and holds no test vectors, so these checks are what pins the oracle:
- nearest neighbours: exact rational brute force (fractions) on tiny sets,
  including exact ties, and scipy's cKDTree (an exact KD-tree search, like
  KDTreeSearcher) on larger ones;
- estimateRt: exact recovery of a known rigid motion, and agreement with the
  SVD (Kabsch) closed form of the same least-squares optimum on noisy data;
- the trial loop of ransac.m: stop points and kept trials derived by hand
  from N = max(10, log(0.01)/log(1 - w^3)).
"""
import math

import numpy as np
import pytest

import registration_oracle as ro


def test_nn_matches_exact_rationals_with_ties():
    rng = np.random.default_rng(0)
    Q = rng.integers(-3, 4, size=(14, 5)).astype(np.float32) * np.float32(0.25)
    R = rng.integers(-3, 4, size=(19, 5)).astype(np.float32) * np.float32(0.25)
    R[7] = R[3]  # exact duplicate: ties go to index 3
    Q[0] = R[3]  # distance 0 to rows 3 and 7
    idx, dist = ro.nn_exact(Q, R)
    assert np.array_equal(idx, ro.nn_fraction(Q, R))
    assert idx[0] == 3 and dist[0] == 0.0
    d_ref = np.sqrt(((Q.astype(np.float64) - R[idx].astype(np.float64)) ** 2).sum(1))
    assert np.allclose(dist, d_ref, rtol=0, atol=1e-12)


def test_nn_matches_exact_rationals_random_floats():
    rng = np.random.default_rng(1)
    Q = rng.normal(size=(10, 7)).astype(np.float32)
    R = rng.normal(size=(25, 7)).astype(np.float32)
    idx, _ = ro.nn_exact(Q, R)
    assert np.array_equal(idx, ro.nn_fraction(Q, R))


def test_nn_matches_kdtree():
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(2)
    Q = rng.normal(size=(400, 32)).astype(np.float32)
    R = rng.normal(size=(650, 32)).astype(np.float32)
    idx, dist = ro.nn_exact(Q, R)
    kd_d, kd_i = cKDTree(R.astype(np.float64)).query(Q.astype(np.float64), k=1)
    assert np.array_equal(idx, kd_i)
    assert np.allclose(dist, kd_d, rtol=1e-12)


def test_mutual_pairs_of_a_permutation():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(300, 16)).astype(np.float32)
    perm = rng.permutation(300)
    B = A[perm] + rng.normal(scale=1e-4, size=A.shape).astype(np.float32)
    ab, _ = ro.nn_exact(A, B)
    ba, _ = ro.nn_exact(B, A)
    pairs = ro.mutual_pairs(ab, ba)
    assert len(pairs) == 300
    assert np.array_equal(pairs[:, 1], np.arange(300))
    assert np.array_equal(pairs[:, 0], perm)


def test_mutual_pairs_filter():
    ab = np.array([2, 0, 1, 1])
    ba = np.array([1, 3, 0])  # j=0 -> i=1 (ab[1]=0 ok); j=1 -> i=3 (ab[3]=1 ok); j=2 -> i=0 (ab[0]=2 ok)
    assert ro.mutual_pairs(ab, ba).tolist() == [[1, 0], [3, 1], [0, 2]]
    ba2 = np.array([1, 2, 2])  # j=1 -> i=2, ab[2]=1 ok; j=2 -> i=2, ab[2]=1 != 2
    assert ro.mutual_pairs(ab, ba2).tolist() == [[1, 0], [2, 1]]


def _motion(seed):
    rng = np.random.default_rng(seed)
    R = ro.random_rotation(rng)
    t = rng.uniform(-1, 1, 3)
    return rng, R, t


def test_estimate_rt_exact_recovery():
    rng, R, t = _motion(4)
    y = rng.uniform(-1, 1, (50, 3))
    x = y @ R.T + t
    rt = ro.estimate_rt(x, y)
    assert np.abs(rt[:, :3] - R).max() < 1e-12
    assert np.abs(rt[:, 3] - t).max() < 1e-12


def test_estimate_rt_three_points_and_batch():
    rng, R, t = _motion(5)
    y = rng.uniform(-1, 1, (6, 3, 3))
    x = y @ R.T + t
    rt = ro.estimate_rt(x, y)
    assert rt.shape == (6, 3, 4)
    assert np.abs(rt[:, :, :3] - R).max() < 1e-10
    assert np.abs(rt[:, :, 3] - t).max() < 1e-10


def test_estimate_rt_equals_kabsch_on_noisy_data():
    for seed in range(6, 12):
        rng, R, t = _motion(seed)
        y = rng.uniform(-1, 1, (200, 3))
        x = y @ R.T + t + rng.normal(scale=0.05, size=(200, 3))
        a = ro.estimate_rt(x, y)
        b = ro.kabsch(x, y)
        assert np.abs(a - b).max() < 1e-10
        assert abs(np.linalg.det(a[:, :3]) - 1) < 1e-12


def test_sample3_distinct_and_uniform():
    s = ro.sample3(123, 60000, 10)
    assert s.min() >= 0 and s.max() < 10
    assert np.all((s[:, 0] != s[:, 1]) & (s[:, 0] != s[:, 2]) & (s[:, 1] != s[:, 2]))
    for k in range(3):
        h = np.bincount(s[:, k], minlength=10) / 60000
        assert np.abs(h - 0.1).max() < 0.01
    # deterministic in the seed, different across seeds
    assert np.array_equal(s[:100], ro.sample3(123, 100, 10))
    assert not np.array_equal(s[:100], ro.sample3(124, 100, 10))


def test_replay_closed_forms():
    n = 100
    # every trial sees all points: N = 10 after the first trial, so 10 trials
    # run and `>=` keeps the last of them
    assert ro.ransac_replay(np.full(60, n), n, 50000) == (9, n, 10)
    # half the points: N = log(0.01)/log(0.875) = 34.49 -> 35 trials
    assert ro.ransac_replay(np.full(100, 50), n, 50000) == (34, 50, 35)
    # no inliers ever: pNoOutliers clamps to 1 - eps and the cap stops the loop
    # after max_trials + 1 trials, keeping the last
    assert ro.ransac_replay(np.zeros(21, np.int64), n, 20) == (20, 0, 21)
    # an early good trial then worse ones: the good one is kept
    c = np.full(100, 10)
    c[3] = 80
    N = math.log(0.01) / math.log(1 - 0.8 ** 3)
    assert N < 10
    assert ro.ransac_replay(c, n, 50000) == (3, 80, 10)
    # a later tie with the best replaces it (>=)
    c[7] = 80
    assert ro.ransac_replay(c, n, 50000) == (7, 80, 10)


def test_ransac_recovers_motion_with_outliers():
    rng, R, t = _motion(13)
    n = 200
    y = rng.uniform(-1, 1, (n, 3))
    x = y @ R.T + t
    out = rng.random(n) < 0.4
    x[out] = rng.uniform(-2, 2, (out.sum(), 3))
    rt, inl, best, run, counts = ro.ransac_fit_rt(x, y, 0.05, 5000, seed=1)
    assert np.array_equal(inl, np.nonzero(~out)[0])
    assert np.abs(rt[:, :3] - R).max() < 1e-9 and np.abs(rt[:, 3] - t).max() < 1e-9
    assert counts[best] == (~out).sum()
    assert 10 <= run <= 5001


def test_ransac_small_sets():
    x = np.eye(3)
    rt, inl, best, run, _ = ro.ransac_fit_rt(x[:2], x[:2])
    assert np.array_equal(rt, np.hstack([np.eye(3), np.zeros((3, 1))])) and len(inl) == 0
    rt, inl, best, run, _ = ro.ransac_fit_rt(x, x)
    assert list(inl) == [0, 1, 2] and run == 0
    assert np.abs(rt - np.hstack([np.eye(3), np.zeros((3, 1))])).max() < 1e-12


# ---- ratioAligned (register2Fragments.m:90-97) ---------------------------------------


def test_alignment_fraction_pinned_by_brute_force_and_closed_forms():
    rng = np.random.default_rng(71)
    a = rng.random((300, 3), dtype=np.float32)
    b = rng.random((250, 3), dtype=np.float32) * np.float32(1.2)
    for r in (0.02, 0.05, 0.2):
        assert ro.fraction_within(a, b, r) == ro.fraction_within_brute(a, b, r)
    # identity transform of the same cloud: every point finds itself
    eye = np.hstack([np.eye(3), np.zeros((3, 1))])
    assert ro.alignment_ratio(a, a, eye) == (1.0, 1.0)
    # a copy shifted by more than the radius: nothing aligns
    far = np.hstack([np.eye(3), [[10.0], [0.0], [0.0]]])
    assert ro.alignment_ratio(a, a, far) == (0.0, 0.0)
    # exactly at the radius is not closer than it (sqrtf(fl(x*x)) == |x|)
    q = np.array([[0, 0, 0]], np.float32)
    t = np.array([[0.05, 0, 0]], np.float32)
    assert ro.fraction_within(q, t, 0.05) == 0.0
    t2 = np.array([[np.nextafter(np.float32(0.05), np.float32(0))], [0], [0]], np.float32).T
    assert ro.fraction_within(q, t2, 0.05) == 1.0
    # a planted mistake (<= instead of <) is caught by that boundary
    dd = q[0] - t
    d2 = (dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]) + dd[:, 2] * dd[:, 2]
    assert np.sqrt(d2)[0] <= np.float32(0.05) and not np.sqrt(d2)[0] < np.float32(0.05)


def test_move_points_matches_float64_transform():
    rng = np.random.default_rng(72)
    p = rng.random((1000, 3), dtype=np.float32) * 4 - 2
    R = ro.random_rotation(rng, max_angle=np.pi / 3)
    rt = np.hstack([R, rng.random((3, 1)) - 0.5])
    got = ro.move_points(p, rt)
    want = p.astype(np.float64) @ rt[:, :3].T + rt[:, 3]
    assert np.abs(got - want).max() < 1e-5
