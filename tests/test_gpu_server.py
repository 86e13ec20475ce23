"""Registration server on the GPU: continuous batching gives each request the
same result alone and in a mix (bit-identical [R|t], match and inlier
counts), agrees with the same pipeline run entry point by entry point, and
admits by its device-memory budget."""
import numpy as np
import pytest

import tdm
from tdm import netdef, synth

pytestmark = pytest.mark.gpu

NK = 800        # keypoints per fragment (small, for test time)
TRIALS = 4000   # RANSAC trial cap


@pytest.fixture(scope="module")
def net16(tmp_path_factory):
    d = tmp_path_factory.mktemp("srv")
    js, wt = str(d / "net.json"), str(d / "w.marvin")
    netdef.write_net_json(js)
    tdm.marvin_write(wt, netdef.he_normal_weights(5))
    net = tdm.Net(js, tdm.PREC_FP16, max_batch=256)
    net.load_weights(wt)
    return net


@pytest.fixture(scope="module")
def requests():
    out = []
    for k in range(6):
        p1, p2, R, t = synth.fragment_pair(100 + k, n1=20_000 + 5_000 * k, n2=30_000 - 2_000 * k,
                                           max_angle=0.3, max_shift=0.3)
        out.append((p1, p2, R, t, 1000 + k))
    return out


def _server(net, **kw):
    cfg = dict(keypoints=NK, max_pass_keypoints=4 * NK, ransac_max_trials=TRIALS)
    cfg.update(kw)
    return tdm.Server(net, **cfg)


def test_alone_equals_in_mix(torch_gpu, net16, requests):
    srv = _server(net16, max_inflight=1, memory_budget=2 << 30)
    alone = []
    for p1, p2, _, _, seed in requests:
        alone.append(srv.wait(srv.submit(p1, p2, seed=seed)))
    st_alone = srv.stats()
    srv.close()
    assert st_alone["max_inflight"] == 1 and st_alone["max_pass_requests"] == 1

    srv = _server(net16, max_inflight=6, memory_budget=2 << 30)
    tickets = [srv.submit(p1, p2, seed=seed) for p1, p2, _, _, seed in requests]
    mixed = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    print("mix stats", st, flush=True)
    assert st["completed"] == len(requests)
    assert st["max_inflight"] > 1 and st["max_pass_requests"] > 1  # passes were shared
    for a, b in zip(alone, mixed):
        assert a["status"] == 0 and b["status"] == 0
        assert np.array_equal(a["rt"], b["rt"])
        assert (a["num_matches"], a["num_inliers"], a["trials"]) == \
            (b["num_matches"], b["num_inliers"], b["trials"])
        assert b["pass_keypoints"] >= a["pass_keypoints"]


def test_server_equals_entry_points(torch_gpu, net16, requests):
    """The server's result equals the pipeline run through the individual
    entry points (TDF, multi-grid descriptors, nearest neighbours, mutual
    pairs, RANSAC) with the server's keypoint choice."""
    torch = torch_gpu
    p1, p2, _, _, seed = requests[2]
    rng = np.random.default_rng(5)
    k1 = rng.integers(0, len(p1), NK).astype(np.int32)
    k2 = rng.integers(0, len(p2), NK).astype(np.int32)
    srv = _server(net16)
    res = srv.wait(srv.submit(p1, p2, k1, k2, seed=seed))
    srv.close()

    feats, kps = [], []
    for pts, kidx in ((p1, k1), (p2, k2)):
        dp = torch.from_numpy(pts).to("cuda:0")
        dims, origin = tdm.grid_params(dp, 0.01, 15)
        tdf = torch.empty(dims[2], dims[1], dims[0], dtype=torch.float32, device="cuda:0")
        tdm.tdf_from_points(dp, dims, origin, 0.01, 0.05, tdf)
        kg = tdm.keypoints_to_grid(pts[kidx], origin, 0.01)
        f = torch.empty(NK, 512, dtype=torch.float32, device="cuda:0")
        net16.forward_grid(tdf, dims, torch.from_numpy(kg).to("cuda:0"), f)
        feats.append(f)
        kps.append(torch.from_numpy(np.ascontiguousarray(pts[kidx])).to("cuda:0"))
    ab = torch.empty(NK, dtype=torch.int32, device="cuda:0")
    ba = torch.empty(NK, dtype=torch.int32, device="cuda:0")
    tdm.match_nn(feats[0], feats[1], ab, ba)
    pairs = torch.empty((NK, 2), dtype=torch.int32, device="cuda:0")
    npairs = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    tdm.mutual_pairs(ab, ba, pairs, npairs)
    rt = torch.zeros(12, dtype=torch.float64, device="cuda:0")
    inl = torch.empty(NK, dtype=torch.int32, device="cuda:0")
    stats = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    tdm.ransac_rt(kps[0], kps[1], rt, inl, stats, pairs=pairs, num_pairs=npairs, max_pairs=NK,
                  threshold=0.05, max_trials=TRIALS, seed=seed)
    torch.cuda.synchronize()
    st = stats.cpu().numpy()
    assert res["status"] == 0
    assert res["num_matches"] == int(npairs.item())
    assert res["num_inliers"] == st[0] and res["trials"] == st[2]
    assert np.array_equal(res["rt"], rt.cpu().numpy().reshape(3, 4))


def test_admission_by_memory_budget(torch_gpu, net16, requests):
    p1, p2, _, _, seed = requests[0]
    srv = _server(net16, memory_budget=2 << 30)
    srv.wait(srv.submit(p1, p2, seed=seed))
    one = srv.stats()["peak_bytes"]  # the block of this one request
    srv.close()
    # room for one such request at a time: everything still completes, one by one
    budget = int(one * 1.5)
    srv = _server(net16, memory_budget=budget, max_inflight=4)
    tickets = [srv.submit(p1, p2, seed=seed + i) for i in range(4)]
    out = [srv.wait(t) for t in tickets]
    st = srv.stats()
    assert all(r["status"] == 0 for r in out)
    assert st["max_inflight"] == 1 and st["peak_bytes"] <= budget
    # a request larger than the whole budget is refused at submit
    with pytest.raises(tdm.TdmError) as e:
        big = np.concatenate([p1, p1 + np.float32(3.0)])
        srv.submit(big, p2, seed=1)
    assert e.value.status == tdm.ERR_NOMEM
    srv.close()
