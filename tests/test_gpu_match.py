"""GPU parity of the registration back end against the registration oracle:
nearest neighbours both ways (bit-exact indices), mutual pairs (exact),
estimateRt (fp64, 1e-10) and ransacfitRt (exact per-trial inlier counts,
kept trial, trials run and inlier set; fitted [R|t] within 1e-9).
"""
import numpy as np
import pytest

import registration_oracle as ro
import tdm
from tdm import synth

pytestmark = pytest.mark.gpu


def _dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to("cuda:0")


def _match(torch, A, B):
    na, nb = len(A), len(B)
    dA, dB = _dev(torch, A, np.float32), _dev(torch, B, np.float32)
    ab = torch.empty(na, dtype=torch.int32, device="cuda:0")
    ba = torch.empty(nb, dtype=torch.int32, device="cuda:0")
    dab = torch.empty(na, dtype=torch.float32, device="cuda:0")
    dba = torch.empty(nb, dtype=torch.float32, device="cuda:0")
    ref = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    tdm.match_nn(dA, dB, ab, ba, dab, dba, ref)
    torch.cuda.synchronize()
    return (ab.cpu().numpy(), ba.cpu().numpy(), dab.cpu().numpy(), dba.cpu().numpy(),
            ref.cpu().numpy())


def _check_nn(torch, A, B):
    ab, ba, dab, dba, refined = _match(torch, A, B)
    oab, odab = ro.nn_exact(A, B)
    oba, odba = ro.nn_exact(B, A)
    assert np.array_equal(ab, oab), np.nonzero(ab != oab)[0][:10]
    assert np.array_equal(ba, oba), np.nonzero(ba != oba)[0][:10]
    assert np.allclose(dab, odab, rtol=1e-6, atol=1e-30)
    assert np.allclose(dba, odba, rtol=1e-6, atol=1e-30)
    return ab, ba, refined


def test_nn_random_descriptors(torch_gpu):
    rng = np.random.default_rng(0)
    A = rng.normal(size=(700, 512)).astype(np.float32)
    B = rng.normal(size=(900, 512)).astype(np.float32)
    B[:300] = A[:300] + rng.normal(scale=0.3, size=(300, 512)).astype(np.float32)
    _, _, refined = _check_nn(torch_gpu, A, B)
    print("refined queries", refined, flush=True)


@pytest.mark.parametrize("na,nb,dim", [(1, 1, 4), (1, 300, 36), (129, 257, 512),
                                       (300, 1, 2048), (255, 130, 68)])
def test_nn_ragged_shapes(torch_gpu, na, nb, dim):
    rng = np.random.default_rng(na * 7 + nb + dim)
    A = rng.normal(size=(na, dim)).astype(np.float32)
    B = rng.normal(size=(nb, dim)).astype(np.float32)
    _check_nn(torch_gpu, A, B)


def test_nn_ties_and_near_ties(torch_gpu):
    """Exact duplicates go to the lower index; near ties (relative gap far
    below the fp32 error bound) are settled by the fp64 search."""
    rng = np.random.default_rng(1)
    A = rng.normal(size=(200, 512)).astype(np.float32)
    B = rng.normal(size=(400, 512)).astype(np.float32)
    B[350:] = B[:50]  # duplicates of rows 0..49
    A[:50] = B[:50] + rng.normal(scale=0.01, size=(50, 512)).astype(np.float32)
    # near ties: two candidates mirrored about the query
    d = rng.normal(scale=0.05, size=(50, 512)).astype(np.float32)
    A[50:100] = rng.normal(size=(50, 512)).astype(np.float32)
    B[100:150] = A[50:100] + d
    B[150:200] = A[50:100] - d
    ab, ba, refined = _check_nn(torch_gpu, A, B)
    assert np.all(ab[:50] < 350)
    assert refined[0] >= 40, refined  # the mirrored pairs cannot be certified in fp32


def test_nn_identical_sets(torch_gpu):
    rng = np.random.default_rng(2)
    A = rng.normal(size=(500, 512)).astype(np.float32)
    ab, ba, _ = _check_nn(torch_gpu, A, A.copy())
    assert np.array_equal(ab, np.arange(500)) and np.array_equal(ba, np.arange(500))


def test_mutual_pairs(torch_gpu):
    rng = np.random.default_rng(3)
    A = rng.normal(size=(1000, 512)).astype(np.float32)
    B = rng.normal(size=(1200, 512)).astype(np.float32)
    B[:600] = A[200:800] + rng.normal(scale=0.5, size=(600, 512)).astype(np.float32)
    torch = torch_gpu
    dA, dB = _dev(torch, A), _dev(torch, B)
    ab = torch.empty(1000, dtype=torch.int32, device="cuda:0")
    ba = torch.empty(1200, dtype=torch.int32, device="cuda:0")
    tdm.match_nn(dA, dB, ab, ba)
    pairs = torch.full((1200, 2), -7, dtype=torch.int32, device="cuda:0")
    npairs = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    tdm.mutual_pairs(ab, ba, pairs, npairs)
    torch.cuda.synchronize()
    n = int(npairs.item())
    oab, _ = ro.nn_exact(A, B)
    oba, _ = ro.nn_exact(B, A)
    ref = ro.mutual_pairs(oab, oba)
    assert n == len(ref) and n > 300
    assert np.array_equal(pairs[:n].cpu().numpy(), ref)


def _rigid_data(seed, n, outlier_frac, noise):
    rng = np.random.default_rng(seed)
    R = ro.random_rotation(rng, max_angle=np.pi / 3)
    t = rng.uniform(-1, 1, 3)
    y = rng.uniform(-1.5, 1.5, (n, 3))
    x = y @ R.T + t + rng.normal(scale=noise, size=(n, 3))
    out = rng.random(n) < outlier_frac
    x[out] = rng.uniform(-2.5, 2.5, (out.sum(), 3))
    return x.astype(np.float32), y.astype(np.float32), R, t, out


def test_estimate_rt(torch_gpu):
    torch = torch_gpu
    for seed, n, noise in [(10, 3, 0.0), (11, 50, 0.0), (12, 1000, 0.02), (13, 5000, 0.05)]:
        x, y, R, t, _ = _rigid_data(seed, n, 0.0, noise)
        rt = torch.zeros(12, dtype=torch.float64, device="cuda:0")
        tdm.estimate_rt(_dev(torch, x), _dev(torch, y), rt)
        got = rt.cpu().numpy().reshape(3, 4)
        want = ro.estimate_rt(x, y)
        assert np.abs(got - want).max() < 1e-10, (seed, np.abs(got - want).max())
        if noise == 0.0:
            assert np.abs(got[:, :3] - R).max() < 1e-5 and np.abs(got[:, 3] - t).max() < 1e-5


def _ransac_gpu(torch, x, y, max_trials, seed, thr=0.05, pairs=None, npairs=None, max_pairs=None):
    n = len(x) if max_pairs is None else max_pairs
    rt = torch.zeros(12, dtype=torch.float64, device="cuda:0")
    inl = torch.full((max(n, 1),), -1, dtype=torch.int32, device="cuda:0")
    stats = torch.full((4,), -9, dtype=torch.int32, device="cuda:0")
    counts = torch.zeros(max_trials + 1, dtype=torch.int32, device="cuda:0")
    tdm.ransac_rt(_dev(torch, x), _dev(torch, y), rt, inl, stats, pairs=pairs, num_pairs=npairs,
                  max_pairs=n, threshold=thr, max_trials=max_trials, seed=seed,
                  trial_counts=counts)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    return (rt.cpu().numpy().reshape(3, 4), inl.cpu().numpy()[:st[0]], st,
            counts.cpu().numpy())


@pytest.mark.parametrize("seed,n,outliers,noise,max_trials",
                         [(20, 300, 0.5, 0.01, 2000), (21, 1000, 0.8, 0.02, 5000),
                          (22, 64, 0.2, 0.0, 300), (23, 2000, 0.95, 0.01, 50000)])
def test_ransac_matches_oracle(torch_gpu, seed, n, outliers, noise, max_trials):
    x, y, R, t, out = _rigid_data(seed, n, outliers, noise)
    rt, inl, st, counts = _ransac_gpu(torch_gpu, x, y, max_trials, seed)
    ort, oinl, obest, orun, ocounts = ro.ransac_fit_rt(x, y, 0.05, max_trials, seed)
    assert np.array_equal(counts, ocounts), np.nonzero(counts != ocounts)[0][:10]
    assert st[1] == obest and st[2] == orun and st[3] == 0
    assert st[0] == len(oinl) and np.array_equal(inl, oinl)
    assert np.abs(rt - ort).max() < 1e-9
    print(f"seed {seed}: trials {st[2]}, inliers {st[0]}/{n}", flush=True)
    if noise == 0.0:
        assert np.abs(rt[:, :3] - R).max() < 1e-5


def test_ransac_on_mutual_pairs_device_count(torch_gpu):
    """pairs and their count straight from tdm_mutual_pairs (no host sync)."""
    torch = torch_gpu
    rng = np.random.default_rng(30)
    x, y, R, t, out = _rigid_data(31, 800, 0.3, 0.005)
    kp1 = x
    perm = rng.permutation(800)
    kp2 = y[perm]  # pair k = (k, inverse perm)
    inv = np.argsort(perm)
    # descriptors that make the true correspondences mutual nearest neighbours
    D1 = rng.normal(size=(800, 64)).astype(np.float32)
    D2 = D1[perm] + rng.normal(scale=0.01, size=(800, 64)).astype(np.float32)
    ab = torch.empty(800, dtype=torch.int32, device="cuda:0")
    ba = torch.empty(800, dtype=torch.int32, device="cuda:0")
    tdm.match_nn(_dev(torch, D1), _dev(torch, D2), ab, ba)
    pairs = torch.empty((800, 2), dtype=torch.int32, device="cuda:0")
    npairs = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    tdm.mutual_pairs(ab, ba, pairs, npairs)
    rt, inl, st, counts = _ransac_gpu(torch, kp1, kp2, 3000, 5, pairs=pairs, npairs=npairs,
                                      max_pairs=800)
    p = pairs.cpu().numpy()[: int(npairs.item())]
    assert len(p) == 800 and np.array_equal(p[:, 0], perm)
    ort, oinl, obest, orun, ocounts = ro.ransac_fit_rt(kp1[p[:, 0]], kp2[p[:, 1]], 0.05, 3000, 5)
    assert np.array_equal(counts, ocounts)
    assert st[1] == obest and st[2] == orun and np.array_equal(inl, oinl)
    assert np.abs(rt - ort).max() < 1e-9
    assert np.abs(rt[:, :3] - R).max() < 1e-2
    del inv


def test_ransac_small_sets(torch_gpu):
    x, y, R, t, _ = _rigid_data(40, 3, 0.0, 0.0)
    rt, inl, st, _ = _ransac_gpu(torch_gpu, x[:2], y[:2], 100, 1)
    assert st[3] == 1 and st[0] == 0
    assert np.array_equal(rt, np.hstack([np.eye(3), np.zeros((3, 1))]))
    rt, inl, st, _ = _ransac_gpu(torch_gpu, x, y, 100, 1)
    assert st[3] == 0 and st[0] == 3 and list(inl) == [0, 1, 2] and st[2] == 0
    assert np.abs(rt - ro.estimate_rt(x, y)).max() < 1e-10


def test_nn_common_offset_and_dense_clusters(torch_gpu):
    """Descriptors sharing a large common component (centring makes the
    fp32 pass decisive) and a dense cluster of near-equidistant candidates
    (the fp64 tile search decides)."""
    rng = np.random.default_rng(4)
    base = rng.normal(size=512).astype(np.float32) * 40
    A = base + rng.normal(scale=0.5, size=(600, 512)).astype(np.float32)
    B = base + rng.normal(scale=0.5, size=(700, 512)).astype(np.float32)
    B[:200] = A[:200] + rng.normal(scale=0.05, size=(200, 512)).astype(np.float32)
    B[650:] = B[600] + rng.normal(scale=1e-4, size=(50, 512)).astype(np.float32)
    _, _, refined = _check_nn(torch_gpu, A, B)
    print("refined queries", refined, flush=True)


# ---- ratioAligned (register2Fragments.m:90-97) ---------------------------------------


def _ratio(torch, p1, p2, rt, radius=0.05):
    return tdm.alignment_ratio(_dev(torch, p1, np.float32), _dev(torch, p2, np.float32), rt, radius)


def test_alignment_ratio_fragment_pair_vs_oracle(torch_gpu):
    """A fragment pair of the serving workload (300k + 250k points) under its
    true motion and under a perturbed one: both ratios equal the oracle's
    (counts are exact, so the fractions are bit-identical)."""
    p1, p2, R, t = synth.fragment_pair(77, n1=300_000, n2=250_000)
    true_rt = np.hstack([R, t[:, None]])
    rng = np.random.default_rng(5)
    off_rt = np.hstack([ro.random_rotation(rng, max_angle=0.05) @ R, (t + 0.03)[:, None]])
    for rt in (true_rt, off_rt):
        got = _ratio(torch_gpu, p1, p2, rt)
        want = ro.alignment_ratio(p1, p2, rt)
        print("ratios", got, want, flush=True)
        assert got == want
    assert _ratio(torch_gpu, p1, p2, true_rt)[0] > 0.3


def test_alignment_ratio_boundaries_and_wide_boxes(torch_gpu):
    rng = np.random.default_rng(9)
    eye = np.hstack([np.eye(3), np.zeros((3, 1))])
    # pairs exactly at the radius (not closer) and one float below it
    q = np.zeros((2, 3), np.float32)
    q[1, 1] = 3.0
    t = np.array([[0.05, 0, 0], [0, 3.0, 0.05]], np.float32)
    assert _ratio(torch_gpu, q, t, eye) == ro.alignment_ratio(q, t, eye) == (0.0, 0.0)
    t[0, 0] = np.nextafter(np.float32(0.05), np.float32(0))
    assert _ratio(torch_gpu, q, t, eye) == ro.alignment_ratio(q, t, eye) == (0.5, 0.5)
    # far outliers stretch the boxes (wider cells), a thin slab, clustered points
    a = rng.random((20_000, 3), dtype=np.float32)
    a[:5] *= np.float32(5000.0)
    b = a[rng.permutation(len(a))[:15_000]] + rng.normal(0, 0.03, (15_000, 3)).astype(np.float32)
    slab = np.stack([rng.random(30_000) * 800, rng.random(30_000) * 0.1, rng.random(30_000) * 0.1],
                    1).astype(np.float32)
    slab2 = slab + rng.normal(0, 0.04, slab.shape).astype(np.float32)
    clus = (rng.normal(0, 0.01, (10_000, 3))).astype(np.float32)
    for p1, p2 in ((a, b), (slab, slab2), (clus, clus[::-1].copy())):
        R = ro.random_rotation(rng, max_angle=0.02)
        rt = np.hstack([R, rng.normal(0, 0.01, (3, 1))])
        got = _ratio(torch_gpu, p1, p2, rt)
        want = ro.alignment_ratio(p1, p2, rt)
        print("case", len(p1), len(p2), got, want, flush=True)
        assert got == want
