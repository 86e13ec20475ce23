// Registration server: continuous batching of fragment-pair registration
// requests on one GPU (the register2Fragments.m pipeline with the
// descriptors of core/demo.cu), admission by a device-memory budget.
//
// Request life: QUEUED (copied host points, grid sizes known) -> ADMITTED
// (arena block and a slot: stream, context and pinned staging) with its two
// TDF grids built on the slot stream -> in a descriptor PASS on the net's
// stream -> MATCHING (mutual NN + RANSAC on the slot stream) -> DONE (result
// read back through the slot's pinned buffer).  One scheduler thread issues
// every HIP call; submit() and wait() only touch host queues.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "pipeline.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// First-fit allocator over one device allocation; blocks are 256-byte
// aligned and coalesce on release.
class Arena {
public:
    void init(char *base, size_t size) {
        base_ = base;
        size_ = size;
        free_.clear();
        free_[0] = size;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = align256(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align256(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *at(size_t off) const { return base_ + off; }
    size_t peak() const { return peak_; }
    size_t size() const { return size_; }

private:
    char *base_ = nullptr;
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Device layout of one request inside its arena block.
struct Layout {
    size_t pts[2], kp[2], kpg[2], tdf[2], desc[2], nn_ab, nn_ba, pairs, inliers, result;
    size_t total;
};

// The part of a request read back at the end: rt, RANSAC stats, pair count.
struct DevResult {
    double rt[12];
    int stats[4];
    int num_pairs;
    int pad[3];
};

enum class State { Queued, Admitted, InPass, Matching, Done };

struct Request {
    uint64_t id = 0;
    std::vector<float> pts[2];
    std::vector<int> kidx[2];
    std::vector<float> kp[2];
    std::vector<int> kpg[2];
    int dims[2][3];
    float origin[2][3];
    uint64_t seed = 0;
    int nk = 0;
    Layout lay{};
    size_t off = 0;
    int slot = -1;
    State state = State::Queued;
    hipEvent_t pass_done = nullptr;  // event of the pass it ran in (not owned)
    int pass_keypoints = 0;
    Clock::time_point t_submit, t_admit;
    tdm_reg_result result{};
};

struct Slot {
    hipStream_t stream = nullptr;
    tdm_ctx *ctx = nullptr;
    hipEvent_t ready = nullptr;  // TDF grids built
    hipEvent_t done = nullptr;   // result copied to `res`
    unsigned char *staging = nullptr;
    size_t staging_bytes = 0;
    DevResult *res = nullptr;  // pinned
    Request *req = nullptr;
};

// One descriptor pass: pinned host inputs, device inputs/outputs, an event.
struct Pass {
    unsigned char *host = nullptr;  // pinned [table | kp | gid]
    unsigned char *dev = nullptr;   // device copy of the same
    float *feat = nullptr;          // device descriptors
    hipEvent_t done = nullptr;
    bool busy = false;
};

Layout make_layout(int n1, int n2, const int d1[3], const int d2[3], int nk, int dim) {
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += align256(bytes);
        return at;
    };
    const int n[2] = {n1, n2};
    const int *d[2] = {d1, d2};
    for (int f = 0; f < 2; ++f) {
        L.pts[f] = take((size_t)n[f] * 12);
        L.kp[f] = take((size_t)nk * 12);
        L.kpg[f] = take((size_t)nk * 12);
        L.tdf[f] = take((size_t)d[f][0] * d[f][1] * d[f][2] * 4);
        L.desc[f] = take((size_t)nk * dim * 4);
    }
    L.nn_ab = take((size_t)nk * 4);
    L.nn_ba = take((size_t)nk * 4);
    L.pairs = take((size_t)nk * 8);
    L.inliers = take((size_t)nk * 4);
    L.result = take(sizeof(DevResult));
    L.total = o;
    return L;
}

}  // namespace

struct tdm_server {
    tdm_net *net = nullptr;
    int device = 0;
    tdm_server_config cfg{};
    int dim = 512;
    char *arena_mem = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    Pass passes[2];
    hipStream_t net_stream = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stopping = false;
    std::thread worker;

    // scheduler-thread state
    std::vector<Request *> ready_for_pass, matching;
    int inflight = 0;
    // counters read by tdm_server_stats under `mu`; written under it too
    long long max_inflight_seen = 0, passes_run = 0, completed = 0, max_pass_requests = 0;

    void run();
    bool admit_one();
    int start_request(Request *r, Slot &s);
    bool launch_pass();
    int start_matching(Request *r);
    bool poll();
    void finish(Request *r, int status);
};

namespace {

// Host half of a request, done in submit(): copies, grid sizes, keypoints.
int prepare(const tdm_server_config &cfg, const tdm_reg_request *q, Request *r) {
    const float *src[2] = {q->points1, q->points2};
    const int n[2] = {q->num_points1, q->num_points2};
    const int *kidx[2] = {q->keypoints1, q->keypoints2};
    const bool given = q->keypoints1 && q->keypoints2;
    TDM_CHECK(!!q->keypoints1 == !!q->keypoints2, TDM_ERR_INVALID_ARG,
              "give keypoints for both fragments or for neither");
    r->nk = given && q->num_keypoints > 0 ? q->num_keypoints : cfg.keypoints;
    TDM_CHECK(r->nk >= 1, TDM_ERR_INVALID_ARG, "no keypoints");
    r->seed = q->seed;
    for (int f = 0; f < 2; ++f) {
        TDM_CHECK(src[f] && n[f] > 0, TDM_ERR_INVALID_ARG, "fragment %d is empty", f + 1);
        r->pts[f].assign(src[f], src[f] + (size_t)n[f] * 3);
        // bounding box (core/demo.cu:102-115), then the grid rule
        float box[6] = {src[f][0], src[f][1], src[f][2], src[f][0], src[f][1], src[f][2]};
        for (long i = 0; i < n[f]; ++i)
            for (int a = 0; a < 3; ++a) {
                box[a] = std::min(box[a], src[f][3 * i + a]);
                box[3 + a] = std::max(box[3 + a], src[f][3 * i + a]);
            }
        TDM_TRY(tdm::grid_from_box(box, cfg.voxel_size, cfg.padding, r->dims[f], r->origin[f]));
        r->kidx[f].resize(r->nk);
        for (int k = 0; k < r->nk; ++k) {
            int idx;
            if (given) {
                idx = kidx[f][k];
                TDM_CHECK(idx >= 0 && idx < n[f], TDM_ERR_INVALID_ARG,
                          "keypoint index %d out of range", idx);
            } else {
                // floor(U[0, N - 1e-4)) (core/demo.cu:169-170, core/utils.hpp:15-20)
                const uint64_t z = mix64(mix64(q->seed ^ (0x5EEDull << 32)) + 2ull * k + f);
                const double u = (double)(z >> 11) * 0x1.0p-53;
                idx = std::min(n[f] - 1, (int)std::floor(u * ((double)n[f] - 1e-4)));
            }
            r->kidx[f][k] = idx;
        }
        r->kp[f].resize((size_t)r->nk * 3);
        for (int k = 0; k < r->nk; ++k)
            for (int a = 0; a < 3; ++a) r->kp[f][3 * k + a] = src[f][3 * (size_t)r->kidx[f][k] + a];
        r->kpg[f].resize((size_t)r->nk * 3);
        TDM_TRY(tdm_keypoints_to_grid(r->kp[f].data(), r->nk, r->origin[f], cfg.voxel_size,
                                      r->kpg[f].data()));
    }
    return TDM_OK;
}

}  // namespace

void tdm_server::finish(Request *r, int status) {
    r->result.status = status;
    r->result.latency_ms = ms_between(r->t_submit, Clock::now());
    if (status != TDM_OK) {
        // work that may still read or write the request's block must end
        // before the block is reused
        if (r->slot >= 0) (void)hipStreamSynchronize(slots[r->slot].stream);
        (void)hipStreamSynchronize(net_stream);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        if (r->slot >= 0) {
            slots[r->slot].req = nullptr;
            r->slot = -1;
            arena.release(r->off, r->lay.total);
            --inflight;
        }
        r->state = State::Done;
        ++completed;
    }
    cv_done.notify_all();
}

int tdm_server::start_request(Request *r, Slot &s) {
    hipStream_t st = s.stream;
    // pinned staging: points, keypoints, keypoint grid coordinates
    size_t need = 0;
    for (int f = 0; f < 2; ++f)
        need += r->pts[f].size() * 4 + r->kp[f].size() * 4 + r->kpg[f].size() * 4;
    if (need > s.staging_bytes) {
        // the slot is idle (its last request finished), so the old buffer is free
        if (s.staging) TDM_HIP(hipHostFree(s.staging));
        s.staging = nullptr;
        s.staging_bytes = 0;
        TDM_HIP(hipHostMalloc((void **)&s.staging, need, hipHostMallocDefault));
        s.staging_bytes = need;
    }
    size_t o = 0;
    auto up = [&](const void *src, size_t bytes, size_t dev_off) -> int {
        std::memcpy(s.staging + o, src, bytes);
        TDM_HIP(hipMemcpyAsync(arena.at(r->off + dev_off), s.staging + o, bytes,
                               hipMemcpyHostToDevice, st));
        o += bytes;
        return TDM_OK;
    };
    for (int f = 0; f < 2; ++f) {
        TDM_TRY(up(r->pts[f].data(), r->pts[f].size() * 4, r->lay.pts[f]));
        TDM_TRY(up(r->kp[f].data(), r->kp[f].size() * 4, r->lay.kp[f]));
        TDM_TRY(up(r->kpg[f].data(), r->kpg[f].size() * 4, r->lay.kpg[f]));
    }
    for (int f = 0; f < 2; ++f)
        TDM_TRY(tdm_tdf_from_points(s.ctx, reinterpret_cast<float *>(arena.at(r->off + r->lay.pts[f])),
                                    (int)(r->pts[f].size() / 3), r->dims[f], r->origin[f],
                                    cfg.voxel_size, cfg.trunc_margin, TDM_DECODE_EXACT,
                                    reinterpret_cast<float *>(arena.at(r->off + r->lay.tdf[f])),
                                    st));
    TDM_HIP(hipEventRecord(s.ready, st));
    return TDM_OK;
}

bool tdm_server::admit_one() {
    Request *r = nullptr;
    int si = -1;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (queue.empty()) return false;
        for (int i = 0; i < (int)slots.size(); ++i)
            if (!slots[i].req) {
                si = i;
                break;
            }
        if (si < 0) return false;
        r = queue.front();
        if (!arena.alloc(r->lay.total, &r->off)) return false;  // wait for memory (FIFO)
        queue.pop_front();
        r->state = State::Admitted;
    }
    r->t_admit = Clock::now();
    r->result.queue_ms = ms_between(r->t_submit, r->t_admit);
    {
        std::lock_guard<std::mutex> lk(mu);
        r->slot = si;
        slots[si].req = r;
        ++inflight;
        max_inflight_seen = std::max<long long>(max_inflight_seen, inflight);
    }
    const int st = start_request(r, slots[si]);
    if (st != TDM_OK) {
        finish(r, st);
        return true;
    }
    ready_for_pass.push_back(r);
    return true;
}

bool tdm_server::launch_pass() {
    if (ready_for_pass.empty()) return false;
    Pass *p = nullptr;
    for (auto &c : passes)
        if (!c.busy) {
            p = &c;
            break;
        }
    if (!p) return false;
    // take requests in arrival order up to the pass size (at least one)
    std::vector<Request *> take;
    int total = 0;
    size_t i = 0;
    for (; i < ready_for_pass.size(); ++i) {
        Request *r = ready_for_pass[i];
        if (!take.empty() && total + 2 * r->nk > cfg.max_pass_keypoints) break;
        take.push_back(r);
        total += 2 * r->nk;
    }
    ready_for_pass.erase(ready_for_pass.begin(), ready_for_pass.begin() + i);
    const int ng = 2 * (int)take.size();
    const size_t tbytes = align256((size_t)ng * sizeof(tdm_grid_ref));
    const size_t kbytes = align256((size_t)total * 12);
    auto fail_all = [&](int st) {
        for (Request *r : take) finish(r, st);
        return true;
    };
    // host inputs (the pass buffers were sized for max_pass_keypoints or for
    // the largest single request, whichever is bigger; see tdm_server_create)
    tdm_grid_ref *table = reinterpret_cast<tdm_grid_ref *>(p->host);
    int *kp = reinterpret_cast<int *>(p->host + tbytes);
    int *gid = reinterpret_cast<int *>(p->host + tbytes + kbytes);
    int at = 0;
    for (size_t q = 0; q < take.size(); ++q) {
        Request *r = take[q];
        for (int f = 0; f < 2; ++f) {
            const int g = 2 * (int)q + f;
            table[g].tdf = reinterpret_cast<const float *>(arena.at(r->off + r->lay.tdf[f]));
            table[g].dim_x = r->dims[f][0];
            table[g].dim_y = r->dims[f][1];
            table[g].dim_z = r->dims[f][2];
            table[g].reserved = 0;
            std::memcpy(kp + 3 * (size_t)at, r->kpg[f].data(), (size_t)r->nk * 12);
            for (int k = 0; k < r->nk; ++k) gid[at + k] = g;
            at += r->nk;
        }
        if (hipStreamWaitEvent(net_stream, slots[r->slot].ready, 0) != hipSuccess)
            return fail_all(TDM_ERR_HIP);
    }
    const size_t bytes = tbytes + kbytes + (size_t)total * 4;
    if (hipMemcpyAsync(p->dev, p->host, bytes, hipMemcpyHostToDevice, net_stream) != hipSuccess)
        return fail_all(TDM_ERR_HIP);
    const int st = tdm_net_forward_grids(
        net, reinterpret_cast<const tdm_grid_ref *>(p->dev),
        reinterpret_cast<const int *>(p->dev + tbytes),
        reinterpret_cast<const int *>(p->dev + tbytes + kbytes), total, p->feat, net_stream);
    if (st != TDM_OK) return fail_all(st);
    at = 0;
    for (Request *r : take)
        for (int f = 0; f < 2; ++f) {
            if (hipMemcpyAsync(arena.at(r->off + r->lay.desc[f]), p->feat + (size_t)at * dim,
                               (size_t)r->nk * dim * 4, hipMemcpyDeviceToDevice,
                               net_stream) != hipSuccess)
                return fail_all(TDM_ERR_HIP);
            at += r->nk;
        }
    if (hipEventRecord(p->done, net_stream) != hipSuccess) return fail_all(TDM_ERR_HIP);
    p->busy = true;
    {
        std::lock_guard<std::mutex> lk(mu);
        ++passes_run;
        max_pass_requests = std::max<long long>(max_pass_requests, (long long)take.size());
    }
    for (Request *r : take) {
        r->pass_done = p->done;
        r->pass_keypoints = total;
        r->state = State::InPass;
        const int s2 = start_matching(r);
        if (s2 != TDM_OK)
            finish(r, s2);
        else
            matching.push_back(r);
    }
    return true;
}

int tdm_server::start_matching(Request *r) {
    Slot &s = slots[r->slot];
    hipStream_t st = s.stream;
    TDM_HIP(hipStreamWaitEvent(st, r->pass_done, 0));
    auto F = [&](size_t o) { return reinterpret_cast<float *>(arena.at(r->off + o)); };
    auto I = [&](size_t o) { return reinterpret_cast<int *>(arena.at(r->off + o)); };
    DevResult *res = reinterpret_cast<DevResult *>(arena.at(r->off + r->lay.result));
    TDM_TRY(tdm_match_nn(s.ctx, F(r->lay.desc[0]), r->nk, F(r->lay.desc[1]), r->nk, dim,
                         I(r->lay.nn_ab), nullptr, I(r->lay.nn_ba), nullptr, nullptr, st));
    TDM_TRY(tdm_mutual_pairs(s.ctx, I(r->lay.nn_ab), r->nk, I(r->lay.nn_ba), r->nk,
                             I(r->lay.pairs), &res->num_pairs, st));
    TDM_TRY(tdm_ransac_rt(s.ctx, F(r->lay.kp[0]), F(r->lay.kp[1]), I(r->lay.pairs),
                          &res->num_pairs, r->nk, cfg.ransac_threshold, cfg.ransac_max_trials,
                          r->seed, res->rt, I(r->lay.inliers), res->stats, nullptr, st));
    TDM_HIP(hipMemcpyAsync(s.res, res, sizeof(DevResult), hipMemcpyDeviceToHost, st));
    TDM_HIP(hipEventRecord(s.done, st));
    r->state = State::Matching;
    return TDM_OK;
}

bool tdm_server::poll() {
    bool progressed = false;
    for (auto &p : passes)
        if (p.busy && hipEventQuery(p.done) == hipSuccess) {
            p.busy = false;
            progressed = true;
        }
    for (size_t i = 0; i < matching.size();) {
        Request *r = matching[i];
        const hipError_t q = hipEventQuery(slots[r->slot].done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        const DevResult &d = *slots[r->slot].res;
        std::memcpy(r->result.rt, d.rt, sizeof d.rt);
        r->result.num_inliers = d.stats[0];
        r->result.trials = d.stats[2];
        r->result.fallback = d.stats[3];
        r->result.num_matches = d.num_pairs;
        r->result.pass_keypoints = r->pass_keypoints;
        matching.erase(matching.begin() + i);
        finish(r, q == hipSuccess ? TDM_OK : TDM_ERR_HIP);
        progressed = true;
    }
    return progressed;
}

void tdm_server::run() {
    (void)hipSetDevice(device);
    for (;;) {
        bool progressed = false;
        while (admit_one()) progressed = true;
        if (launch_pass()) progressed = true;
        if (poll()) progressed = true;
        if (progressed) continue;
        std::unique_lock<std::mutex> lk(mu);
        if (stopping && queue.empty() && inflight == 0) break;
        if (inflight == 0 && queue.empty()) {
            cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
        } else {
            // GPU work in flight: check again shortly
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

namespace {

// Device resources of a new server (released by tdm_server_destroy, also
// after a partial failure here).
int server_init(tdm_server *s) {
    const tdm_server_config &c = s->cfg;
    s->dim = tdm_net_output_dim(s->net);
    TDM_CHECK(s->dim > 0 && s->dim % 4 == 0, TDM_ERR_INVALID_ARG, "net output width %d", s->dim);
    TDM_HIP(hipMalloc((void **)&s->arena_mem, c.memory_budget));
    s->arena.init(s->arena_mem, c.memory_budget);
    TDM_HIP(hipStreamCreateWithFlags(&s->net_stream, hipStreamNonBlocking));
    s->slots.resize(c.max_inflight);
    for (auto &sl : s->slots) {
        TDM_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        TDM_TRY(tdm_ctx_create(s->device, &sl.ctx));
        TDM_HIP(hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming));
        TDM_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        TDM_HIP(hipHostMalloc((void **)&sl.res, sizeof(DevResult), hipHostMallocDefault));
        // matching and RANSAC scratch up front, so requests do not allocate
        void *p = nullptr;
        TDM_TRY(tdm::ctx_scratch(sl.ctx, 2, tdm::match_scratch_bytes(c.keypoints, c.keypoints, s->dim),
                                 &p, sl.stream));
        TDM_TRY(tdm::ctx_scratch(sl.ctx, 3, tdm::ransac_scratch_bytes(c.ransac_max_trials), &p,
                                 sl.stream));
    }
    // a pass holds max_pass_keypoints, or one request's 2 x keypoints
    const int maxk = std::max(c.max_pass_keypoints, 2 * c.keypoints);
    const int maxg = 2 * std::max(1, maxk / 2);
    const size_t pbytes = align256((size_t)maxg * sizeof(tdm_grid_ref)) +
                          align256((size_t)maxk * 12) + (size_t)maxk * 4;
    for (auto &p : s->passes) {
        TDM_HIP(hipHostMalloc((void **)&p.host, pbytes, hipHostMallocDefault));
        TDM_HIP(hipMalloc((void **)&p.dev, pbytes));
        TDM_HIP(hipMalloc((void **)&p.feat, (size_t)maxk * s->dim * 4));
        TDM_HIP(hipEventCreateWithFlags(&p.done, hipEventDisableTiming));
    }
    TDM_HIP(hipDeviceSynchronize());
    return TDM_OK;
}

}  // namespace

extern "C" {

void tdm_server_config_init(tdm_server_config *cfg) {
    if (!cfg) return;
    cfg->memory_budget = (size_t)4 << 30;
    cfg->max_inflight = 3;  // 3 request streams + the net's = HIP's 4 hardware queues
    cfg->keypoints = 5000;
    cfg->max_pass_keypoints = 20000;
    cfg->voxel_size = 0.01f;
    cfg->trunc_margin = 0.01f * 5;
    cfg->padding = 15;
    cfg->ransac_threshold = 0.05;
    cfg->ransac_max_trials = 50000;
}

int tdm_server_create(tdm_net *net, int device, const tdm_server_config *cfg,
                      tdm_server **out) {
    tdm::clear_error();
    TDM_CHECK(net && out, TDM_ERR_INVALID_ARG, "NULL argument");
    tdm_server_config c;
    tdm_server_config_init(&c);
    if (cfg) c = *cfg;
    TDM_CHECK(c.max_inflight >= 1 && c.keypoints >= 1 && c.max_pass_keypoints >= 2 &&
                  c.voxel_size > 0 && c.trunc_margin > 0 && c.padding >= 0 &&
                  c.memory_budget > 0,
              TDM_ERR_INVALID_ARG, "bad server configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return tdm::fail(TDM_ERR_NO_DEVICE, "no HIP device");
    TDM_CHECK(device >= 0 && device < ndev, TDM_ERR_INVALID_ARG, "bad device %d", device);
    TDM_HIP(hipSetDevice(device));
    tdm_server *s = new tdm_server();
    s->net = net;
    s->device = device;
    s->cfg = c;
    const int st = server_init(s);
    if (st != TDM_OK) {
        tdm_server_destroy(s);
        return st;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return TDM_OK;
}

int tdm_server_submit(tdm_server *srv, const tdm_reg_request *req, uint64_t *ticket) {
    tdm::clear_error();
    TDM_CHECK(srv && req && ticket, TDM_ERR_INVALID_ARG, "NULL argument");
    auto r = std::make_unique<Request>();
    r->t_submit = Clock::now();
    TDM_TRY(prepare(srv->cfg, req, r.get()));
    TDM_CHECK(!req->keypoints1 || r->nk <= srv->cfg.keypoints, TDM_ERR_INVALID_ARG,
              "%d keypoints exceed the configured %d", r->nk, srv->cfg.keypoints);
    r->lay = make_layout((int)(r->pts[0].size() / 3), (int)(r->pts[1].size() / 3), r->dims[0],
                         r->dims[1], r->nk, srv->dim);
    TDM_CHECK(r->lay.total <= srv->arena.size(), TDM_ERR_NOMEM,
              "request needs %zu bytes, more than the %zu-byte budget", r->lay.total,
              srv->arena.size());
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        TDM_CHECK(!srv->stopping, TDM_ERR_INVALID_ARG, "server is stopping");
        r->id = srv->next_id++;
        *ticket = r->id;
        srv->queue.push_back(r.get());
        srv->all.emplace(r->id, std::move(r));
    }
    srv->cv_work.notify_one();
    return TDM_OK;
}

int tdm_server_wait(tdm_server *srv, uint64_t ticket, tdm_reg_result *out) {
    tdm::clear_error();
    TDM_CHECK(srv && out, TDM_ERR_INVALID_ARG, "NULL argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(ticket);
    TDM_CHECK(it != srv->all.end(), TDM_ERR_INVALID_ARG, "unknown ticket %llu",
              (unsigned long long)ticket);
    Request *r = it->second.get();
    srv->cv_done.wait(lk, [&] { return r->state == State::Done; });
    *out = r->result;
    srv->all.erase(it);
    return TDM_OK;
}

int tdm_server_stats(tdm_server *srv, long long stats[6]) {
    tdm::clear_error();
    TDM_CHECK(srv && stats, TDM_ERR_INVALID_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    stats[0] = (long long)srv->arena.peak();
    stats[1] = (long long)srv->arena.size();
    stats[2] = srv->max_inflight_seen;
    stats[3] = srv->passes_run;
    stats[4] = srv->completed;
    stats[5] = srv->max_pass_requests;
    return TDM_OK;
}

int tdm_server_destroy(tdm_server *srv) {
    tdm::clear_error();
    if (!srv) return TDM_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->device);
    (void)hipDeviceSynchronize();
    for (auto &sl : srv->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.ctx) tdm_ctx_destroy(sl.ctx);
        if (sl.ready) (void)hipEventDestroy(sl.ready);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.staging) (void)hipHostFree(sl.staging);
        if (sl.res) (void)hipHostFree(sl.res);
    }
    for (auto &p : srv->passes) {
        if (p.host) (void)hipHostFree(p.host);
        if (p.dev) (void)hipFree(p.dev);
        if (p.feat) (void)hipFree(p.feat);
        if (p.done) (void)hipEventDestroy(p.done);
    }
    if (srv->net_stream) (void)hipStreamDestroy(srv->net_stream);
    if (srv->arena_mem) (void)hipFree(srv->arena_mem);
    delete srv;
    return TDM_OK;
}

}  // extern "C"
