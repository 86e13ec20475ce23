// The descriptor network: Marvin JSON -> a chain of convolutions, .marvin
// weights, and the forward pass over patches (core/marvin.hpp:7434-7679).
//
// Layers are resolved from the JSON by following response names: every
// Convolution produces a response; an Activation (ReLU) whose input is that
// response (in place in core/3dmatch-net-test.json) is folded into the
// convolution's epilogue, and a max Pooling is fused into it as well when its
// window is 2x2x2 with stride 2 and the extent is even.  Activations live in
// two ping-pong device buffers sized for `max_batch` patches, NDHWC, in the
// storage type of the precision (float or bf16); the last layer writes float
// descriptors straight into the caller's buffer.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "conv.hpp"
#include "json.hpp"
#include "marvin_io.hpp"

namespace {

struct Layer {
    std::string name;
    std::string in_name, out_name;
    int cin = 0, cout = 0, k = 3;
    int in_d = 0, out_d = 0;  // extents before pooling
    bool relu = false;
    bool pool = false;       // max pool window 2 stride 2
    bool pool_fused = false;
    int pooled_d = 0;
    std::vector<float> w, b;  // host copies (float, reference layout)
    bool has_w = false, has_b = false;
    bool dirty = false;      // host weights newer than the device copy
    void *w_dev = nullptr;   // packed [N][taps][C] (conv1: [N][KP])
    float *b_dev = nullptr;
    int variant = -1;        // conv tile variant (-1: default; see tdm_net_autotune)
    // fp8: per output channel weight scale (host), scale of this layer's
    // output activations (set by tdm_net_calibrate), and the device
    // dequantization scale per output channel (weight scale x input scale)
    std::vector<float> wscale;
    float act_scale = 0.0f;
    float *colscale_dev = nullptr;
    int final_d() const { return pool ? pooled_d : out_d; }
    size_t out_elems() const { return (size_t)cout * final_d() * final_d() * final_d(); }
};

}  // namespace

struct tdm_net {
    tdm_ctx *ctx = nullptr;
    int device = 0;
    tdm::DType dt = tdm::DType::F32;
    int max_batch = 256;
    int patch = 30;
    int json_batch = 50;
    std::vector<Layer> layers;
    void *act[2] = {nullptr, nullptr};
    size_t act_bytes = 0;
    float *partial = nullptr;
    size_t partial_bytes = 0;
    // Per-layer timing (tdm_net_set_profiling): events recorded on the stream
    // around each layer's launches, resolved when statistics are read.
    bool profiling = false;
    struct Timed {
        int layer;
        int patches;
        hipEvent_t start, stop;
    };
    std::vector<Timed> pending;
    std::vector<double> layer_ms;
    std::vector<long> layer_launches, layer_patches;
};

namespace {

int read_text(const char *path, std::string &out) {
    std::ifstream in(path, std::ios::binary);
    TDM_CHECK(in, TDM_ERR_IO, "cannot open network file %s", path);
    std::stringstream ss;
    ss << in.rdbuf();
    out = ss.str();
    return TDM_OK;
}

int int_list(const tdm::Json *j, std::vector<int> &out) {
    out.clear();
    if (!j) return TDM_OK;
    TDM_CHECK(j->kind == tdm::Json::Array, TDM_ERR_FORMAT, "expected a number list");
    for (auto &v : j->arr) {
        TDM_CHECK(v.kind == tdm::Json::Number, TDM_ERR_FORMAT, "expected a number");
        out.push_back((int)v.num);
    }
    return TDM_OK;
}

std::string first_name(const tdm::Json *j) {
    if (!j) return "";
    if (j->kind == tdm::Json::String) return j->str;
    if (j->kind == tdm::Json::Array && !j->arr.empty() && j->arr[0].kind == tdm::Json::String)
        return j->arr[0].str;
    return "";
}

std::string str_field(const tdm::Json &o, const char *key) {
    const tdm::Json *v = o.get(key);
    return v && v->kind == tdm::Json::String ? v->str : "";
}

// Builds the layer chain of the test phase (core/marvin.hpp:7434-7527).
int build_layers(const tdm::Json &root, tdm_net &net) {
    const tdm::Json *layers = root.get("layers");
    TDM_CHECK(layers && layers->kind == tdm::Json::Array, TDM_ERR_FORMAT, "no \"layers\" array");
    std::string data_name;
    int data_c = 0, data_d = 0;
    for (const tdm::Json &l : layers->arr) {
        TDM_CHECK(l.kind == tdm::Json::Object, TDM_ERR_FORMAT, "layer is not an object");
        const std::string type = str_field(l, "type");
        const std::string name = str_field(l, "name");
        const std::string phase = str_field(l, "phase");
        if (phase == "Training") continue;
        const std::string in = first_name(l.get("in"));
        const std::string out = first_name(l.get("out"));
        if (type == "PlaceHolderData" || type == "MemoryData" || type == "DiskData") {
            std::vector<int> dim;
            TDM_TRY(int_list(l.get("dim"), dim));
            TDM_CHECK(dim.size() == 5 && dim[2] == dim[3] && dim[3] == dim[4], TDM_ERR_UNSUPPORTED,
                      "data layer %s: need dim [N, C, P, P, P]", name.c_str());
            data_name = out;
            net.json_batch = dim[0];
            data_c = dim[1];
            data_d = dim[2];
        } else if (type == "Convolution") {
            Layer c;
            c.name = name;
            c.in_name = in;
            c.out_name = out;
            const tdm::Json *no = l.get("num_output");
            TDM_CHECK(no && no->kind == tdm::Json::Number, TDM_ERR_FORMAT, "%s: no num_output",
                      name.c_str());
            c.cout = (int)no->num;
            std::vector<int> win, pad, stride;
            TDM_TRY(int_list(l.get("window"), win));
            TDM_TRY(int_list(l.get("padding"), pad));
            TDM_TRY(int_list(l.get("stride"), stride));
            TDM_CHECK(win.size() == 3 && win[0] == win[1] && win[1] == win[2] &&
                          (win[0] == 3 || win[0] == 1),
                      TDM_ERR_UNSUPPORTED, "%s: only cubic 3x3x3 / 1x1x1 windows", name.c_str());
            for (int v : pad) TDM_CHECK(v == 0, TDM_ERR_UNSUPPORTED, "%s: padding", name.c_str());
            for (int v : stride) TDM_CHECK(v == 1, TDM_ERR_UNSUPPORTED, "%s: stride", name.c_str());
            c.k = win[0];
            if (net.layers.empty()) {
                TDM_CHECK(in == data_name && !data_name.empty(), TDM_ERR_UNSUPPORTED,
                          "%s: first convolution must read the data layer", name.c_str());
                c.cin = data_c;
                c.in_d = data_d;
            } else {
                const Layer &p = net.layers.back();
                TDM_CHECK(in == p.out_name, TDM_ERR_UNSUPPORTED,
                          "%s: input %s is not the previous layer's output", name.c_str(),
                          in.c_str());
                c.cin = p.cout;
                c.in_d = p.final_d();
            }
            c.out_d = c.in_d - c.k + 1;
            TDM_CHECK(c.out_d >= 1, TDM_ERR_UNSUPPORTED, "%s: input extent too small", name.c_str());
            net.layers.push_back(c);
        } else if (type == "Activation") {
            TDM_CHECK(str_field(l, "mode") == "ReLU", TDM_ERR_UNSUPPORTED,
                      "%s: only ReLU activations", name.c_str());
            TDM_CHECK(!net.layers.empty() && in == net.layers.back().out_name &&
                          !net.layers.back().pool,
                      TDM_ERR_UNSUPPORTED, "%s: ReLU must follow a convolution", name.c_str());
            net.layers.back().relu = true;
            net.layers.back().out_name = out;
        } else if (type == "Pooling") {
            const std::string mode = str_field(l, "mode");
            TDM_CHECK(mode.empty() || mode == "max", TDM_ERR_UNSUPPORTED, "%s: only max pooling",
                      name.c_str());
            std::vector<int> win, stride, pad;
            TDM_TRY(int_list(l.get("window"), win));
            TDM_TRY(int_list(l.get("stride"), stride));
            TDM_TRY(int_list(l.get("padding"), pad));
            if (stride.empty()) stride = win;
            TDM_CHECK(win.size() == 3 && win == std::vector<int>({2, 2, 2}) &&
                          stride == std::vector<int>({2, 2, 2}),
                      TDM_ERR_UNSUPPORTED, "%s: only 2x2x2 stride-2 pooling", name.c_str());
            for (int v : pad) TDM_CHECK(v == 0, TDM_ERR_UNSUPPORTED, "%s: padding", name.c_str());
            TDM_CHECK(!net.layers.empty() && in == net.layers.back().out_name &&
                          !net.layers.back().pool,
                      TDM_ERR_UNSUPPORTED, "%s: pooling must follow a convolution", name.c_str());
            Layer &p = net.layers.back();
            p.pool = true;
            // core/marvin.hpp:5560-5562: 1 + ceil((E - 2)/2).
            p.pooled_d = 1 + (p.out_d - 2 + 1) / 2;
            p.out_name = out;
        } else {
            return tdm::fail(TDM_ERR_UNSUPPORTED, "layer %s: type %s is not supported",
                             name.c_str(), type.c_str());
        }
    }
    TDM_CHECK(!net.layers.empty(), TDM_ERR_FORMAT, "network has no convolution");
    TDM_CHECK(net.layers[0].cin == 1, TDM_ERR_UNSUPPORTED, "data must have one channel");
    TDM_CHECK(net.layers.back().final_d() == 1, TDM_ERR_UNSUPPORTED,
              "the last layer must reduce a patch to 1x1x1");
    net.patch = net.layers[0].in_d;
    for (size_t i = 0; i < net.layers.size(); ++i) {
        Layer &l = net.layers[i];
        TDM_CHECK(l.cout % 64 == 0, TDM_ERR_UNSUPPORTED, "%s: %d outputs (need a multiple of 64)",
                  l.name.c_str(), l.cout);
        if (i == 0) {
            TDM_CHECK(l.k == 3, TDM_ERR_UNSUPPORTED, "first convolution must be 3x3x3");
        } else {
            TDM_CHECK(l.cin % 32 == 0, TDM_ERR_UNSUPPORTED, "%s: %d inputs (need a multiple of 32)",
                      l.name.c_str(), l.cin);
        }
        l.pool_fused = l.pool && l.out_d % 2 == 0 && i > 0;
    }
    return TDM_OK;
}

// Packs host weights [O][I][k^3] into the device layout of the precision.
// fp8: each output channel is scaled so its largest |w| maps to the largest
// e4m3 value, quantized on the device, and the dequantization scale of the
// channel (weight scale x the input activations' scale) goes to colscale.
int upload_layer(tdm_net &net, Layer &l, size_t li) {
    const int taps = l.k * l.k * l.k;
    // the first convolution of an fp8 net computes in fp16
    const tdm::DType wdt = (net.dt == tdm::DType::F8 && li == 0) ? tdm::DType::F16 : net.dt;
    const size_t esz = tdm::dtype_size(wdt);
    std::vector<float> packed;
    int row = 0;
    if (li == 0) {
        row = tdm::conv1_taps_padded(wdt);
        packed.assign((size_t)l.cout * row, 0.0f);
        for (int o = 0; o < l.cout; ++o)
            for (int t = 0; t < taps; ++t) packed[(size_t)o * row + t] = l.w[(size_t)o * taps + t];
    } else {
        row = taps * l.cin;
        packed.resize((size_t)l.cout * row);
        for (int o = 0; o < l.cout; ++o)
            for (int c = 0; c < l.cin; ++c)
                for (int t = 0; t < taps; ++t)
                    packed[(size_t)o * row + (size_t)t * l.cin + c] =
                        l.w[((size_t)o * l.cin + c) * taps + t];
    }
    const size_t bytes = packed.size() * esz;
    if (!l.w_dev) TDM_HIP(hipMalloc(&l.w_dev, bytes));
    if (!l.b_dev) TDM_HIP(hipMalloc(&l.b_dev, sizeof(float) * l.cout));
    if (wdt == tdm::DType::F32) {
        TDM_HIP(hipMemcpy(l.w_dev, packed.data(), bytes, hipMemcpyHostToDevice));
    } else if (wdt == tdm::DType::F8) {
        l.wscale.assign(l.cout, 1.0f);
        std::vector<float> inv(l.cout, 1.0f), cs(l.cout);
        for (int o = 0; o < l.cout; ++o) {
            float m = 0.0f;
            for (int k = 0; k < row; ++k) m = std::max(m, std::fabs(packed[(size_t)o * row + k]));
            if (m > 0.0f) {
                l.wscale[o] = m / tdm::kFp8Max;
                inv[o] = tdm::kFp8Max / m;
            }
            cs[o] = l.wscale[o] * net.layers[li - 1].act_scale;
        }
        float *tmp = nullptr, *dinv = nullptr;
        TDM_HIP(hipMalloc(&tmp, packed.size() * sizeof(float)));
        int st = TDM_OK;
        if (hipMalloc(&dinv, sizeof(float) * l.cout) != hipSuccess) st = TDM_ERR_HIP;
        if (st == TDM_OK &&
            (hipMemcpy(tmp, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice) !=
                 hipSuccess ||
             hipMemcpy(dinv, inv.data(), sizeof(float) * l.cout, hipMemcpyHostToDevice) !=
                 hipSuccess))
            st = TDM_ERR_HIP;
        if (st == TDM_OK) st = tdm::launch_quantize_fp8(tmp, l.cout, row, dinv, l.w_dev, nullptr);
        if (st == TDM_OK && hipDeviceSynchronize() != hipSuccess) st = TDM_ERR_HIP;
        (void)hipFree(tmp);
        if (dinv) (void)hipFree(dinv);
        if (st != TDM_OK) return tdm::fail(st, "%s: fp8 weight quantization failed", l.name.c_str());
        if (!l.colscale_dev) TDM_HIP(hipMalloc(&l.colscale_dev, sizeof(float) * l.cout));
        TDM_HIP(hipMemcpy(l.colscale_dev, cs.data(), sizeof(float) * l.cout, hipMemcpyHostToDevice));
    } else {
        std::vector<uint16_t> h(packed.size());
        if (wdt == tdm::DType::F16)
            for (size_t i = 0; i < h.size(); ++i) h[i] = tdm::float_to_f16(packed[i]);
        else
            for (size_t i = 0; i < h.size(); ++i) h[i] = tdm::float_to_bf16(packed[i]);
        TDM_HIP(hipMemcpy(l.w_dev, h.data(), bytes, hipMemcpyHostToDevice));
    }
    TDM_HIP(hipMemcpy(l.b_dev, l.b.data(), sizeof(float) * l.cout, hipMemcpyHostToDevice));
    return TDM_OK;
}

int ensure_buffers(tdm_net &net) {
    if (net.act[0]) return TDM_OK;
    const size_t esz = tdm::dtype_size(net.dt);
    size_t most = 0, most_partial = 0;
    for (size_t i = 0; i < net.layers.size(); ++i) {
        const Layer &l = net.layers[i];
        size_t pre = (size_t)l.cout * l.out_d * l.out_d * l.out_d;
        most = std::max(most, l.pool_fused ? l.out_elems() : pre);
        most = std::max(most, l.out_elems());
    }
    // Float outputs (last layer, per-layer dumps) may need 4-byte elements.
    net.act_bytes = std::max(most * esz, (size_t)net.layers.back().cout * 4) * net.max_batch;
    TDM_HIP(hipMalloc(&net.act[0], net.act_bytes));
    TDM_HIP(hipMalloc(&net.act[1], net.act_bytes));
    // Split-K partials for small-row layers: up to 64 splits of batch rows.
    for (const Layer &l : net.layers) {
        size_t rows = (size_t)net.max_batch * l.out_d * l.out_d * l.out_d;
        if (rows <= (size_t)net.max_batch * 27) most_partial = std::max(most_partial, rows * l.cout);
    }
    net.partial_bytes = most_partial * 64 * sizeof(float);
    if (net.partial_bytes) TDM_HIP(hipMalloc(&net.partial, net.partial_bytes));
    return TDM_OK;
}

// Reduction splits for a layer with few rows, to give the GPU >= ~2048 waves.
// Chosen from max_batch, never from the rows of the current pass, so a
// patch's descriptor does not depend on how many patches share its pass.
int choose_splits(const tdm_net &net, const Layer &l, size_t nq_chunks) {
    const long rows = (long)net.max_batch * l.out_d * l.out_d * l.out_d;
    const long waves = ((rows + 63) / 64) * (l.cout / 64);
    if (waves >= 2048 || l.pool) return 1;
    long s = (2048 + waves - 1) / waves;
    s = std::min<long>(s, 64);
    s = std::min<long>(s, (long)nq_chunks);
    const size_t need = (size_t)s * rows * l.cout * sizeof(float);
    if (need > net.partial_bytes) s = (long)(net.partial_bytes / ((size_t)rows * l.cout * sizeof(float)));
    return (int)std::max<long>(s, 1);
}

// Multiplier applied to a layer's outputs before fp8 conversion.
float out_scale(const tdm_net &net, const Layer &l) {
    return net.dt == tdm::DType::F8 && l.act_scale > 0.0f ? 1.0f / l.act_scale : 1.0f;
}

// Runs layers [0, upto] for `n` patches of one pass.  The input of the first
// layer comes from `src`.  When upto is the last layer, its float output goes
// to `final_out`; otherwise the output stays in *result (storage type).
int run_layer(tdm_net &net, int li, const tdm::VolumeSource &src, int n, float *final_out,
              int last, int &cur, hipStream_t s);

int run_pass(tdm_net &net, const tdm::VolumeSource &src, int n, int upto, float *final_out,
             const void **result, hipStream_t s) {
    int cur = 0;
    const int last = (int)net.layers.size() - 1;
    for (int li = 0; li <= upto; ++li) {
        tdm_net::Timed timed{li, n, nullptr, nullptr};
        if (net.profiling) {
            TDM_HIP(hipEventCreate(&timed.start));
            TDM_HIP(hipEventCreate(&timed.stop));
            TDM_HIP(hipEventRecord(timed.start, s));
        }
        const int st = run_layer(net, li, src, n, final_out, last, cur, s);
        if (net.profiling) {
            TDM_HIP(hipEventRecord(timed.stop, s));
            net.pending.push_back(timed);
        }
        TDM_TRY(st);
    }
    if (result) *result = net.act[cur];
    return TDM_OK;
}

// One layer of run_pass; `cur` is the ping-pong buffer holding its input and
// is advanced past its output.
int run_layer(tdm_net &net, int li, const tdm::VolumeSource &src, int n, float *final_out,
              int last, int &cur, hipStream_t s) {
    {
        Layer &l = net.layers[li];
        const bool to_final = li == last && final_out;
        void *out = to_final ? static_cast<void *>(final_out) : net.act[1 - cur];
        // fp8 nets keep the last layer (the descriptor) in float
        const bool out_f32 =
            to_final || net.dt == tdm::DType::F32 || (net.dt == tdm::DType::F8 && li == last);
        if (li == 0) {
            void *dst = l.pool && !l.pool_fused ? net.act[1 - cur] : out;
            TDM_TRY(tdm::launch_conv1(net.dt, out_f32, src, l.cout, l.relu, n, l.w_dev, l.b_dev,
                                      dst, s, out_scale(net, l)));
            cur = 1 - cur;
            if (l.pool) {
                TDM_TRY(tdm::launch_maxpool(net.dt, net.act[cur], n, l.cout, l.out_d, net.act[1 - cur], s));
                cur = 1 - cur;
            }
            return TDM_OK;
        }
        tdm::ConvDesc d;
        d.C = l.cin;
        d.N = l.cout;
        d.D = l.in_d;
        d.OD = l.out_d;
        d.ksize = l.k;
        d.relu = l.relu;
        d.pool = l.pool_fused;
        const size_t nq = (size_t)tdm::conv_reduction_steps(net.dt, d);
        const int splits = choose_splits(net, l, nq);
        void *dst = (l.pool && !l.pool_fused) ? net.act[1 - cur] : out;
        TDM_TRY(tdm::launch_conv(net.dt, out_f32, d, n, net.act[cur], l.w_dev, l.b_dev, dst,
                                 splits, net.partial, l.variant, s, l.colscale_dev,
                                 out_scale(net, l)));
        cur = 1 - cur;
        if (l.pool && !l.pool_fused) {
            TDM_CHECK(!to_final, TDM_ERR_UNSUPPORTED, "pooling after the last layer");
            TDM_TRY(tdm::launch_maxpool(net.dt, net.act[cur], n, l.cout, l.out_d, net.act[1 - cur], s));
            cur = 1 - cur;
        }
    }
    return TDM_OK;
}

int resolve_timing(tdm_net &net) {
    const size_t L = net.layers.size();
    net.layer_ms.resize(L, 0.0);
    net.layer_launches.resize(L, 0);
    net.layer_patches.resize(L, 0);
    for (auto &t : net.pending) {
        float ms = 0.0f;
        TDM_HIP(hipEventSynchronize(t.stop));
        TDM_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        net.layer_ms[t.layer] += ms;
        net.layer_launches[t.layer] += 1;
        net.layer_patches[t.layer] += t.patches;
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    net.pending.clear();
    return TDM_OK;
}

// The device is bound at the first forward (a net can be built and given
// weights without a GPU); later calls make it current again.
int bind_device(tdm_net *net) {
    (void)hipGetLastError();  // a caller's stale error is not this call's failure
    if (net->device < 0) {
        tdm_ctx *c = net->ctx;
        TDM_TRY(tdm::resolve_ctx(c, &c));
        net->ctx = c;
        net->device = c->device;
    }
    TDM_HIP(hipSetDevice(net->device));
    return TDM_OK;
}

int check_ready(tdm_net *net) {
    TDM_CHECK(net, TDM_ERR_INVALID_ARG, "null net");
    for (const Layer &l : net->layers)
        TDM_CHECK(l.has_w && l.has_b, TDM_ERR_INVALID_ARG, "layer %s has no weights loaded",
                  l.name.c_str());
    if (net->dt == tdm::DType::F8)
        for (size_t li = 0; li + 1 < net->layers.size(); ++li)
            TDM_CHECK(net->layers[li].act_scale > 0.0f, TDM_ERR_INVALID_ARG,
                      "fp8 net: activation scales not set (call tdm_net_calibrate)");
    TDM_TRY(bind_device(net));
    for (size_t li = 0; li < net->layers.size(); ++li)
        if (net->layers[li].dirty) {
            // Work queued earlier may still read the old weights.
            TDM_HIP(hipDeviceSynchronize());
            TDM_TRY(upload_layer(*net, net->layers[li], li));
            net->layers[li].dirty = false;
        }
    TDM_TRY(ensure_buffers(*net));
    return TDM_OK;
}

// Host copy of one layer's weights; uploaded (packed) at the next forward.
int set_layer(tdm_net *net, Layer &l, const float *w, size_t wn, const float *b, size_t bn) {
    const size_t want_w = (size_t)l.cout * l.cin * l.k * l.k * l.k;
    TDM_CHECK(wn == want_w && bn == (size_t)l.cout, TDM_ERR_INVALID_ARG,
              "%s: expected %zu weights and %d biases, got %zu and %zu", l.name.c_str(), want_w,
              l.cout, wn, bn);
    (void)net;
    l.w.assign(w, w + wn);
    l.b.assign(b, b + bn);
    l.has_w = l.has_b = true;
    l.dirty = true;
    return TDM_OK;
}

}  // namespace

extern "C" {

int tdm_net_create(tdm_ctx *ctx, const char *json_path, int precision, int max_batch,
                   tdm_net **out) {
    tdm::clear_error();
    TDM_CHECK(json_path && out, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(precision == TDM_PREC_FP32 || precision == TDM_PREC_BF16 ||
                  precision == TDM_PREC_FP16 || precision == TDM_PREC_FP8,
              TDM_ERR_INVALID_ARG, "unknown precision %d", precision);
    TDM_CHECK(max_batch >= 0 && max_batch <= 65536, TDM_ERR_INVALID_ARG, "bad max_batch");
    std::string text, err;
    TDM_TRY(read_text(json_path, text));
    tdm::Json root;
    TDM_CHECK(tdm::json_parse(text, root, err), TDM_ERR_FORMAT, "%s: %s", json_path, err.c_str());
    tdm_net *net = new tdm_net();
    net->ctx = ctx;
    net->device = ctx ? ctx->device : -1;
    net->dt = precision == TDM_PREC_FP32   ? tdm::DType::F32
              : precision == TDM_PREC_FP16 ? tdm::DType::F16
              : precision == TDM_PREC_FP8  ? tdm::DType::F8
                                           : tdm::DType::BF16;
    net->max_batch = max_batch ? max_batch : 256;
    int s = build_layers(root, *net);
    if (s != TDM_OK) {
        delete net;
        return s;
    }
    *out = net;
    return TDM_OK;
}

int tdm_net_destroy(tdm_net *net) {
    if (!net) return TDM_OK;
    if (net->device >= 0) {
        (void)hipSetDevice(net->device);
        (void)hipDeviceSynchronize();
        (void)resolve_timing(*net);
    }
    for (Layer &l : net->layers) {
        if (l.w_dev) (void)hipFree(l.w_dev);
        if (l.b_dev) (void)hipFree(l.b_dev);
        if (l.colscale_dev) (void)hipFree(l.colscale_dev);
    }
    for (void *p : net->act)
        if (p) (void)hipFree(p);
    if (net->partial) (void)hipFree(net->partial);
    delete net;
    return TDM_OK;
}

int tdm_net_load_weights(tdm_net *net, const char *marvin_path) {
    tdm::clear_error();
    TDM_CHECK(net && marvin_path, TDM_ERR_INVALID_ARG, "null pointer argument");
    std::vector<tdm::MarvinTensor> ts;
    TDM_TRY(tdm::marvin_read(marvin_path, ts));
    for (size_t li = 0; li < net->layers.size(); ++li) {
        Layer &l = net->layers[li];
        const tdm::MarvinTensor *w = nullptr, *b = nullptr;
        const size_t want_w = (size_t)l.cout * l.cin * l.k * l.k * l.k;
        for (const auto &t : ts) {
            // core/marvin.hpp:3321-3377: match by name; a numel mismatch is skipped.
            if (t.name == l.name + ".weight") {
                if (t.data.size() == want_w) w = &t;
                else std::fprintf(stderr, "[Warning] %s.weight is found but not loaded because the numels are mismatched\n", l.name.c_str());
            }
            if (t.name == l.name + ".bias") {
                if (t.data.size() == (size_t)l.cout) b = &t;
                else std::fprintf(stderr, "[Warning] %s.bias is found but not loaded because the numels are mismatched\n", l.name.c_str());
            }
        }
        if (w && b) TDM_TRY(set_layer(net, l, w->data.data(), w->data.size(), b->data.data(), b->data.size()));
        else if (w || b) {
            // Keep the other half if it was set earlier.
            std::vector<float> wv = w ? w->data : l.w, bv = b ? b->data : l.b;
            if (wv.size() == want_w && bv.size() == (size_t)l.cout)
                TDM_TRY(set_layer(net, l, wv.data(), wv.size(), bv.data(), bv.size()));
        }
    }
    return TDM_OK;
}

int tdm_net_set_layer(tdm_net *net, const char *layer_name, const float *weight,
                      size_t weight_count, const float *bias, size_t bias_count) {
    tdm::clear_error();
    TDM_CHECK(net && layer_name && weight && bias, TDM_ERR_INVALID_ARG, "null pointer argument");
    for (size_t li = 0; li < net->layers.size(); ++li)
        if (net->layers[li].name == layer_name)
            return set_layer(net, net->layers[li], weight, weight_count, bias, bias_count);
    return tdm::fail(TDM_ERR_INVALID_ARG, "no layer named %s", layer_name);
}

int tdm_net_set_profiling(tdm_net *net, int enable) {
    tdm::clear_error();
    TDM_CHECK(net, TDM_ERR_INVALID_ARG, "null net");
    if (net->device >= 0) {
        TDM_HIP(hipSetDevice(net->device));
        TDM_TRY(resolve_timing(*net));
    }
    net->layer_ms.resize(net->layers.size());
    net->layer_launches.resize(net->layers.size());
    net->layer_patches.resize(net->layers.size());
    net->profiling = enable != 0;
    std::fill(net->layer_ms.begin(), net->layer_ms.end(), 0.0);
    std::fill(net->layer_launches.begin(), net->layer_launches.end(), 0);
    std::fill(net->layer_patches.begin(), net->layer_patches.end(), 0);
    return TDM_OK;
}

int tdm_net_layer_stats(tdm_net *net, int layer_index, double *total_ms, long *launches,
                        long *patches) {
    tdm::clear_error();
    TDM_CHECK(net && total_ms && launches && patches, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(layer_index >= 0 && layer_index < (int)net->layers.size(), TDM_ERR_INVALID_ARG,
              "layer index %d out of range", layer_index);
    if (net->device >= 0) {
        TDM_HIP(hipSetDevice(net->device));
        TDM_TRY(resolve_timing(*net));
    }
    net->layer_ms.resize(net->layers.size());
    net->layer_launches.resize(net->layers.size());
    net->layer_patches.resize(net->layers.size());
    *total_ms = net->layer_ms[layer_index];
    *launches = net->layer_launches[layer_index];
    *patches = net->layer_patches[layer_index];
    return TDM_OK;
}

int tdm_net_output_dim(const tdm_net *net) { return net && !net->layers.empty() ? net->layers.back().cout : 0; }

int tdm_net_num_layers(const tdm_net *net) { return net ? (int)net->layers.size() : 0; }

int tdm_net_layer_info(const tdm_net *net, int index, int info[8]) {
    tdm::clear_error();
    TDM_CHECK(net && info, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(index >= 0 && index < (int)net->layers.size(), TDM_ERR_INVALID_ARG,
              "layer index %d out of range", index);
    const Layer &l = net->layers[index];
    const int v[8] = {l.cin, l.cout, l.k, l.in_d, l.out_d, l.relu ? 1 : 0, l.pool ? 1 : 0, l.final_d()};
    std::memcpy(info, v, sizeof v);
    return TDM_OK;
}

const char *tdm_net_layer_name(const tdm_net *net, int index) {
    if (!net || index < 0 || index >= (int)net->layers.size()) return nullptr;
    return net->layers[index].name.c_str();
}

int tdm_net_forward(tdm_net *net, const float *data, int num, float *feat, void *stream) {
    tdm::clear_error();
    TDM_CHECK(num >= 0, TDM_ERR_INVALID_ARG, "negative count");
    TDM_CHECK(num == 0 || (data && feat), TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_TRY(check_ready(net));
    hipStream_t s = tdm::as_stream(stream);
    const long vol = (long)net->patch * net->patch * net->patch;
    const int D = net->layers.back().cout;
    for (int start = 0; start < num; start += net->max_batch) {
        const int n = std::min(net->max_batch, num - start);
        tdm::VolumeSource src;
        src.base = data + (long)start * vol;
        src.P = net->patch;
        TDM_TRY(run_pass(*net, src, n, (int)net->layers.size() - 1, feat + (long)start * D, nullptr, s));
    }
    return TDM_OK;
}

int tdm_net_forward_grid(tdm_net *net, const float *tdf, int dim_x, int dim_y, int dim_z,
                         const int *kp_grid, int num, float *feat, void *stream) {
    tdm::clear_error();
    TDM_CHECK(num >= 0, TDM_ERR_INVALID_ARG, "negative count");
    TDM_CHECK(num == 0 || (tdf && kp_grid && feat), TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(dim_x > 0 && dim_y > 0 && dim_z > 0, TDM_ERR_INVALID_ARG, "bad grid");
    TDM_TRY(check_ready(net));
    hipStream_t s = tdm::as_stream(stream);
    const int D = net->layers.back().cout;
    for (int start = 0; start < num; start += net->max_batch) {
        const int n = std::min(net->max_batch, num - start);
        tdm::VolumeSource src;
        src.base = tdf;
        src.kp = kp_grid + (long)start * 3;
        src.dims[0] = dim_x;
        src.dims[1] = dim_y;
        src.dims[2] = dim_z;
        src.P = net->patch;
        TDM_TRY(run_pass(*net, src, n, (int)net->layers.size() - 1, feat + (long)start * D, nullptr, s));
    }
    return TDM_OK;
}

int tdm_net_forward_grids(tdm_net *net, const tdm_grid_ref *grids, const int *kp_grid,
                          const int *grid_index, int num, float *feat, void *stream) {
    static_assert(sizeof(tdm_grid_ref) == sizeof(tdm::GridRef) &&
                      offsetof(tdm_grid_ref, dim_x) == offsetof(tdm::GridRef, dims),
                  "tdm_grid_ref and GridRef layouts differ");
    tdm::clear_error();
    TDM_CHECK(num >= 0, TDM_ERR_INVALID_ARG, "negative count");
    TDM_CHECK(num == 0 || (grids && kp_grid && grid_index && feat), TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_TRY(check_ready(net));
    hipStream_t s = tdm::as_stream(stream);
    const int D = net->layers.back().cout;
    for (int start = 0; start < num; start += net->max_batch) {
        const int n = std::min(net->max_batch, num - start);
        tdm::VolumeSource src;
        src.kp = kp_grid + (long)start * 3;
        src.grids = reinterpret_cast<const tdm::GridRef *>(grids);
        src.gid = grid_index + start;
        src.P = net->patch;
        TDM_TRY(run_pass(*net, src, n, (int)net->layers.size() - 1, feat + (long)start * D, nullptr, s));
    }
    return TDM_OK;
}

int tdm_net_autotune(tdm_net *net, int num, void *stream) {
    tdm::clear_error();
    TDM_TRY(check_ready(net));
    TDM_CHECK(num >= 1, TDM_ERR_INVALID_ARG, "batch %d", num);
    const int n = std::min(num, net->max_batch);
    hipStream_t s = tdm::as_stream(stream);
    // finite inputs for the timing runs (results are discarded)
    TDM_HIP(hipMemsetAsync(net->act[0], 0, net->act_bytes, s));
    hipEvent_t e0, e1;
    TDM_HIP(hipEventCreate(&e0));
    TDM_HIP(hipEventCreate(&e1));
    int st = TDM_OK;
    for (size_t li = 1; li < net->layers.size() && st == TDM_OK; ++li) {
        Layer &l = net->layers[li];
        tdm::ConvDesc d;
        d.C = l.cin;
        d.N = l.cout;
        d.D = l.in_d;
        d.OD = l.out_d;
        d.ksize = l.k;
        d.relu = l.relu;
        d.pool = l.pool_fused;
        if (!tdm::conv_uses_lds(net->dt, d)) continue;
        const size_t nq = (size_t)tdm::conv_reduction_steps(net->dt, d);
        const int splits = choose_splits(*net, l, nq);
        const bool out_f32 = net->dt == tdm::DType::F32 || li + 1 == net->layers.size();
        auto run = [&](int v) {
            return tdm::launch_conv(net->dt, out_f32, d, n, net->act[0], l.w_dev, l.b_dev,
                                    net->act[1], splits, net->partial, v, s, l.colscale_dev,
                                    out_scale(*net, l));
        };
        // Three rounds over the candidates, each timing one launch per
        // candidate, and every candidate's fastest launch decides: drifting
        // clocks then weigh on all candidates alike, not on the last ones.
        std::vector<int> cand;
        for (int v = 0; v < tdm::kConvVariants; ++v)
            if (tdm::conv_variant_valid(net->dt, v, d)) cand.push_back(v);
        std::vector<float> fastest(cand.size(), 3.4e38f);
        for (size_t c = 0; c < cand.size() && st == TDM_OK; ++c) st = run(cand[c]);  // warm-up
        for (int r = 0; r < 3 && st == TDM_OK; ++r) {
            for (size_t c = 0; c < cand.size() && st == TDM_OK; ++c) {
                if (hipEventRecord(e0, s) != hipSuccess) st = TDM_ERR_HIP;
                if (st == TDM_OK) st = run(cand[c]);
                if (st == TDM_OK && hipEventRecord(e1, s) != hipSuccess) st = TDM_ERR_HIP;
                float ms = 0.0f;
                if (st == TDM_OK && (hipEventSynchronize(e1) != hipSuccess ||
                                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    st = TDM_ERR_HIP;
                if (st == TDM_OK) fastest[c] = std::min(fastest[c], ms);
            }
        }
        int best_v = tdm::conv_default_variant(net->dt, d);
        float best = 3.4e38f;
        for (size_t c = 0; c < cand.size(); ++c)
            if (fastest[c] < best) {
                best = fastest[c];
                best_v = cand[c];
            }
        l.variant = best_v;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st == TDM_ERR_HIP) return tdm::fail(st, "autotune timing failed");
    return st;
}

int tdm_net_calibrate(tdm_net *net, tdm_net *ref, const float *data, int num, void *stream) {
    tdm::clear_error();
    TDM_CHECK(net && ref && data && num >= 1, TDM_ERR_INVALID_ARG, "bad arguments");
    TDM_CHECK(net->dt == tdm::DType::F8, TDM_ERR_INVALID_ARG, "calibration is for fp8 nets");
    TDM_CHECK(ref->dt != tdm::DType::F8, TDM_ERR_INVALID_ARG, "the reference net must not be fp8");
    TDM_CHECK(ref->layers.size() == net->layers.size(), TDM_ERR_INVALID_ARG,
              "nets have different layers");
    size_t most = 0;
    for (size_t i = 0; i < net->layers.size(); ++i) {
        const Layer &a = net->layers[i], &b = ref->layers[i];
        TDM_CHECK(a.cin == b.cin && a.cout == b.cout && a.k == b.k && a.in_d == b.in_d &&
                      a.pool == b.pool,
                  TDM_ERR_INVALID_ARG, "layer %zu differs between the nets", i);
        most = std::max(most, a.out_elems());
    }
    hipStream_t s = tdm::as_stream(stream);
    float *tmp = nullptr;
    unsigned *amax = nullptr;
    TDM_HIP(hipMalloc(&tmp, most * num * sizeof(float)));
    int st = TDM_OK;
    if (hipMalloc(&amax, sizeof(unsigned)) != hipSuccess) st = TDM_ERR_HIP;
    for (size_t li = 0; li + 1 < net->layers.size() && st == TDM_OK; ++li) {
        const size_t count = net->layers[li].out_elems() * num;
        st = tdm_net_forward_layer(ref, data, num, (int)li, tmp, count, stream);
        if (st == TDM_OK && hipMemsetAsync(amax, 0, sizeof(unsigned), s) != hipSuccess)
            st = TDM_ERR_HIP;
        if (st == TDM_OK) st = tdm::launch_absmax(tmp, (long)count, amax, s);
        unsigned bits = 0;
        if (st == TDM_OK &&
            (hipMemcpyAsync(&bits, amax, sizeof bits, hipMemcpyDeviceToHost, s) != hipSuccess ||
             hipStreamSynchronize(s) != hipSuccess))
            st = TDM_ERR_HIP;
        float m;
        std::memcpy(&m, &bits, sizeof m);
        net->layers[li].act_scale = (st == TDM_OK && m > 0.0f) ? m / tdm::kFp8Max : 1.0f;
    }
    (void)hipFree(tmp);
    if (amax) (void)hipFree(amax);
    if (st != TDM_OK) return st == TDM_ERR_HIP ? tdm::fail(st, "calibration failed") : st;
    for (Layer &l : net->layers) l.dirty = l.has_w;  // re-upload with the new scales
    return TDM_OK;
}

int tdm_net_act_scale(const tdm_net *net, int index, float *scale) {
    tdm::clear_error();
    TDM_CHECK(net && scale && index >= 0 && index < (int)net->layers.size(), TDM_ERR_INVALID_ARG,
              "bad arguments");
    *scale = net->layers[index].act_scale;
    return TDM_OK;
}

int tdm_net_set_act_scale(tdm_net *net, int index, float scale) {
    tdm::clear_error();
    TDM_CHECK(net && index >= 0 && index < (int)net->layers.size() && scale > 0.0f,
              TDM_ERR_INVALID_ARG, "bad arguments");
    net->layers[index].act_scale = scale;
    for (Layer &l : net->layers) l.dirty = l.has_w;
    return TDM_OK;
}

int tdm_net_layer_variant(const tdm_net *net, int index) {
    if (!net || index < 0 || index >= (int)net->layers.size()) return -2;
    return net->layers[index].variant;
}

int tdm_net_set_layer_variant(tdm_net *net, int index, int variant) {
    tdm::clear_error();
    TDM_CHECK(net && index >= 0 && index < (int)net->layers.size(), TDM_ERR_INVALID_ARG,
              "bad layer index");
    TDM_CHECK(variant >= -1 && variant < tdm::kConvVariants, TDM_ERR_INVALID_ARG,
              "bad variant %d", variant);
    net->layers[index].variant = variant;
    return TDM_OK;
}

int tdm_net_forward_layer(tdm_net *net, const float *data, int num, int layer_index, float *out,
                          size_t out_count, void *stream) {
    tdm::clear_error();
    TDM_CHECK(net && data && out && num > 0, TDM_ERR_INVALID_ARG, "bad arguments");
    TDM_CHECK(layer_index >= 0 && layer_index < (int)net->layers.size(), TDM_ERR_INVALID_ARG,
              "layer index %d out of range", layer_index);
    TDM_TRY(check_ready(net));
    const Layer &l = net->layers[layer_index];
    const size_t per = l.out_elems();
    TDM_CHECK(out_count >= per * num, TDM_ERR_INVALID_ARG, "output buffer too small");
    hipStream_t s = tdm::as_stream(stream);
    const long vol = (long)net->patch * net->patch * net->patch;
    for (int start = 0; start < num; start += net->max_batch) {
        const int n = std::min(net->max_batch, num - start);
        tdm::VolumeSource src;
        src.base = data + (long)start * vol;
        src.P = net->patch;
        const void *res = nullptr;
        TDM_TRY(run_pass(*net, src, n, layer_index, nullptr, &res, s));
        const bool last_f32 =
            net->dt == tdm::DType::F8 && layer_index + 1 == (int)net->layers.size();
        TDM_TRY(tdm::launch_to_ncdhw(net->dt, last_f32, res, n, l.cout, l.final_d(),
                                     out + (long)start * per, s, l.act_scale));
    }
    return TDM_OK;
}

}  // extern "C"
