// Realignment screen, host side: the engine's per-region screen step (candidate list -> device
// features -> network scores -> keep flags) and the network's trainer, with their C entry points.
// None of this has a counterpart in the reference; exact mode never calls it.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "rabbitvar_amd.h"
#include "rv_engine.h"
#include "rv_realign.h"
#include "rv_region.h"
#include "rv_screen.h"
#include "rv_screen_kernels.h"

namespace rv {

#define RV_HIPE(x)                                                                                         \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) return fail(RV_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

rv_status Engine::screen_plan(const PileupResult& pr, ScreenPlan& plan, bool want_features, bool events_done) {
    const auto l0 = std::chrono::steady_clock::now();
    if (events_done)
        screen_candidates_finish(pr.P, plan);
    else
        screen_candidates(pr.P, plan);
    stats.screen_list_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - l0).count();
    const size_t n = plan.cand.size();
    if (n == 0) return RV_OK;
    // neighbourhood tables, packed
    size_t np = 0, nc = 0;
    for (int t = 0; t < ST_N; t++) {
        np += plan.tpos[t].size();
        nc += plan.tcum[t].size();
    }
    std::vector<int32_t> pos;
    std::vector<int64_t> cum;
    pos.reserve(np + 1);
    cum.reserve(nc);
    size_t po[ST_N], co[ST_N];
    for (int t = 0; t < ST_N; t++) {
        po[t] = pos.size();
        co[t] = cum.size();
        pos.insert(pos.end(), plan.tpos[t].begin(), plan.tpos[t].end());
        cum.insert(cum.end(), plan.tcum[t].begin(), plan.tcum[t].end());
    }
    pos.push_back(0);  // no table is read past its end; keeps the buffer non-empty
    auto* dcand = (ScreenCand*)grow(b_scr_cand, sizeof(ScreenCand) * n);
    auto* dpos = (int32_t*)grow(b_scr_pos, sizeof(int32_t) * pos.size());
    auto* dcum = (int64_t*)grow(b_scr_cum, sizeof(int64_t) * cum.size());
    auto* dfeat = (float*)grow(b_scr_feat, sizeof(float) * SCR_F * n);
    auto* dscore = (float*)grow(b_scr_score, sizeof(float) * n);
    if (!dcand || !dpos || !dcum || !dfeat || !dscore) return fail(RV_ERR_OOM, "device allocation for the screen failed");
    // synchronous copies: the host vectors are pageable and short-lived
    RV_HIPE(hipStreamSynchronize(stream));
    RV_HIPE(hipMemcpy(dcand, plan.cand.data(), sizeof(ScreenCand) * n, hipMemcpyHostToDevice));
    RV_HIPE(hipMemcpy(dpos, pos.data(), sizeof(int32_t) * pos.size(), hipMemcpyHostToDevice));
    RV_HIPE(hipMemcpy(dcum, cum.data(), sizeof(int64_t) * cum.size(), hipMemcpyHostToDevice));
    ScreenIn S{};
    S.cand = dcand;
    S.n = (int)n;
    for (int t = 0; t < ST_N; t++) {
        S.tpos[t] = dpos + po[t];
        S.tcum[t] = dcum + co[t];
        S.tn[t] = (int)plan.tpos[t].size();
    }
    S.rows = rows;
    S.dense_lo = pr.P.dense_lo;
    S.dense_n = pr.P.dense_n;
    S.ref = ref_dev;
    S.ref_lo = pr.W.ref_lo;
    S.ref_hi = pr.W.ref_lo + (int)pr.ref_win.size() - 1;
    S.feat = dfeat;
    RV_HIPE(launch_screen_features(S, stream));
    if (screen_on) {
        RV_HIPE(launch_mlp_forward((const float*)b_scr_params.p, dfeat, (int64_t)n, dscore, stream));
        std::vector<float> score(n);
        RV_HIPE(hipMemcpyAsync(score.data(), dscore, sizeof(float) * n, hipMemcpyDeviceToHost, stream));
        RV_HIPE(hipStreamSynchronize(stream));
        plan.keep.resize(n);
        for (size_t i = 0; i < n; i++) plan.keep[i] = score[i] >= screen_threshold ? 1 : 0;
    }
    if (want_features) {
        collect_feat.resize(n * SCR_F);
        RV_HIPE(hipMemcpyAsync(collect_feat.data(), dfeat, sizeof(float) * SCR_F * n, hipMemcpyDeviceToHost, stream));
    }
    RV_HIPE(hipStreamSynchronize(stream));
    return RV_OK;
}

// ------------------------------------------------------------------ trainer
struct Mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr, *loss = nullptr;
    float* X = nullptr;
    uint8_t* Y = nullptr;
    int32_t* idx = nullptr;
    int64_t n = 0;
    int64_t step = 0;

    rv_status fail(rv_status s, const std::string& m) {
        err = m;
        return s;
    }
    ~Mlp() {
        for (void* p : {(void*)P, (void*)G, (void*)M, (void*)V, (void*)loss, (void*)X, (void*)Y, (void*)idx})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    rv_status open(int dev, const float* init, uint64_t seed) {
        device = dev;
        RV_HIPE(hipSetDevice(dev));
        RV_HIPE(hipStreamCreate(&stream));
        const size_t pb = sizeof(float) * SCR_NPARAM;
        RV_HIPE(hipMalloc((void**)&P, pb));
        RV_HIPE(hipMalloc((void**)&G, pb));
        RV_HIPE(hipMalloc((void**)&M, pb));
        RV_HIPE(hipMalloc((void**)&V, pb));
        RV_HIPE(hipMalloc((void**)&loss, sizeof(float)));
        std::vector<float> h(SCR_NPARAM, 0.f);
        if (init) {
            std::copy(init, init + SCR_NPARAM, h.begin());
        } else {
            // Glorot-uniform weights, zero biases
            std::mt19937_64 g(seed);
            auto fill = [&](int off, int fin, int fout) {
                const double a = std::sqrt(6.0 / (fin + fout));
                std::uniform_real_distribution<double> u(-a, a);
                for (int i = 0; i < fin * fout; i++) h[(size_t)(off + i)] = (float)u(g);
            };
            fill(SCR_W1, SCR_F, SCR_H);
            fill(SCR_W2, SCR_H, SCR_H);
            fill(SCR_W3, SCR_H, 1);
        }
        RV_HIPE(hipMemcpy(P, h.data(), pb, hipMemcpyHostToDevice));
        RV_HIPE(hipMemset(G, 0, pb));
        RV_HIPE(hipMemset(M, 0, pb));
        RV_HIPE(hipMemset(V, 0, pb));
        return RV_OK;
    }
    rv_status load(const float* x, const uint8_t* y, int64_t cnt) {
        for (void** p : {(void**)&X, (void**)&Y, (void**)&idx})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        n = cnt;
        if (cnt <= 0) return RV_OK;
        RV_HIPE(hipMalloc((void**)&X, sizeof(float) * SCR_F * (size_t)cnt));
        RV_HIPE(hipMalloc((void**)&Y, (size_t)cnt));
        RV_HIPE(hipMalloc((void**)&idx, sizeof(int32_t) * (size_t)cnt));
        RV_HIPE(hipMemcpy(X, x, sizeof(float) * SCR_F * (size_t)cnt, hipMemcpyHostToDevice));
        if (y)
            RV_HIPE(hipMemcpy(Y, y, (size_t)cnt, hipMemcpyHostToDevice));
        else
            RV_HIPE(hipMemset(Y, 0, (size_t)cnt));
        return RV_OK;
    }
    rv_status upload_order(const std::vector<int32_t>& order) {
        RV_HIPE(hipMemcpyAsync(idx, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice, stream));
        RV_HIPE(hipStreamSynchronize(stream));  // `order` is pageable
        return RV_OK;
    }
    rv_status gradient(float pos_weight, float* grad_out, float* loss_out) {
        if (n <= 0) return fail(RV_ERR_ARG, "no samples loaded");
        std::vector<int32_t> order((size_t)n);
        for (int64_t i = 0; i < n; i++) order[(size_t)i] = (int32_t)i;
        rv_status s = upload_order(order);
        if (s != RV_OK) return s;
        RV_HIPE(hipMemsetAsync(G, 0, sizeof(float) * SCR_NPARAM, stream));
        RV_HIPE(hipMemsetAsync(loss, 0, sizeof(float), stream));
        RV_HIPE(launch_mlp_grad(P, X, Y, idx, n, 1.f / (float)n, pos_weight, G, loss, stream));
        if (grad_out) RV_HIPE(hipMemcpyAsync(grad_out, G, sizeof(float) * SCR_NPARAM, hipMemcpyDeviceToHost, stream));
        if (loss_out) RV_HIPE(hipMemcpyAsync(loss_out, loss, sizeof(float), hipMemcpyDeviceToHost, stream));
        RV_HIPE(hipMemsetAsync(G, 0, sizeof(float) * SCR_NPARAM, stream));
        RV_HIPE(hipStreamSynchronize(stream));
        return RV_OK;
    }
    rv_status train(int epochs, int batch, float lr, float pos_weight, uint64_t seed, float* epoch_loss) {
        if (n <= 0) return fail(RV_ERR_ARG, "no samples loaded");
        if (epochs <= 0 || batch <= 0 || !(lr > 0.f)) return fail(RV_ERR_ARG, "epochs, batch and lr must be positive");
        std::vector<int32_t> order((size_t)n);
        for (int64_t i = 0; i < n; i++) order[(size_t)i] = (int32_t)i;
        std::mt19937_64 g(seed);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int ep = 0; ep < epochs; ep++) {
            std::shuffle(order.begin(), order.end(), g);
            rv_status s = upload_order(order);
            if (s != RV_OK) return s;
            RV_HIPE(hipMemsetAsync(loss, 0, sizeof(float), stream));
            int64_t nbatches = 0;
            for (int64_t b0 = 0; b0 < n; b0 += batch) {
                const int64_t nb = std::min<int64_t>(batch, n - b0);
                RV_HIPE(launch_mlp_grad(P, X, Y, idx + b0, nb, 1.f / (float)nb, pos_weight, G, loss, stream));
                step++;
                const float c1 = 1.f / (1.f - std::pow(b1, (float)step)), c2 = 1.f / (1.f - std::pow(b2, (float)step));
                RV_HIPE(launch_adam(P, G, M, V, SCR_NPARAM, lr, b1, b2, eps, c1, c2, stream));
                nbatches++;
            }
            float l = 0.f;
            RV_HIPE(hipMemcpyAsync(&l, loss, sizeof(float), hipMemcpyDeviceToHost, stream));
            RV_HIPE(hipStreamSynchronize(stream));
            if (epoch_loss) epoch_loss[ep] = l / (float)nbatches;
        }
        return RV_OK;
    }
    rv_status predict(const float* x, int64_t cnt, float* out) {
        if (cnt <= 0) return RV_OK;
        float *dx = nullptr, *dy = nullptr;
        struct Free {
            float*& a;
            float*& b;
            ~Free() {
                if (a) (void)hipFree(a);
                if (b) (void)hipFree(b);
            }
        } fr{dx, dy};
        RV_HIPE(hipMalloc((void**)&dx, sizeof(float) * SCR_F * (size_t)cnt));
        RV_HIPE(hipMalloc((void**)&dy, sizeof(float) * (size_t)cnt));
        RV_HIPE(hipMemcpy(dx, x, sizeof(float) * SCR_F * (size_t)cnt, hipMemcpyHostToDevice));
        RV_HIPE(launch_mlp_forward(P, dx, cnt, dy, stream));
        RV_HIPE(hipMemcpyAsync(out, dy, sizeof(float) * (size_t)cnt, hipMemcpyDeviceToHost, stream));
        RV_HIPE(hipStreamSynchronize(stream));
        return RV_OK;
    }
};

}  // namespace rv

struct rv_mlp {
    rv::Mlp m;
};

struct rv_engine;
namespace rv {
Engine* engine_of(rv_engine* e);  // rv_capi.cpp
}

extern "C" {

size_t rv_mlp_param_count(void) { return (size_t)rv::SCR_NPARAM; }

rv_status rv_mlp_open(int device, const float* init_params, uint64_t seed, rv_mlp** out) {
    if (!out) return RV_ERR_ARG;
    *out = nullptr;
    rv_mlp* m = new (std::nothrow) rv_mlp;
    if (!m) return RV_ERR_OOM;
    rv_status s = m->m.open(device, init_params, seed);
    if (s != RV_OK) {
        delete m;
        return s;
    }
    *out = m;
    return RV_OK;
}

rv_status rv_mlp_load(rv_mlp* m, const float* features, const uint8_t* labels, size_t n) {
    if (!m || (n && !features) || n > (size_t)INT32_MAX) return RV_ERR_ARG;
    return m->m.load(features, labels, (int64_t)n);
}

rv_status rv_mlp_train(rv_mlp* m, int epochs, int batch, float lr, float pos_weight, uint64_t seed, float* epoch_loss) {
    if (!m) return RV_ERR_ARG;
    return m->m.train(epochs, batch, lr, pos_weight, seed, epoch_loss);
}

rv_status rv_mlp_gradient(rv_mlp* m, float pos_weight, float* grad, float* loss) {
    if (!m) return RV_ERR_ARG;
    return m->m.gradient(pos_weight, grad, loss);
}

rv_status rv_mlp_predict(rv_mlp* m, const float* features, size_t n, float* scores) {
    if (!m || (n && (!features || !scores))) return RV_ERR_ARG;
    return m->m.predict(features, (int64_t)n, scores);
}

rv_status rv_mlp_get_params(rv_mlp* m, float* params) {
    if (!m || !params) return RV_ERR_ARG;
    hipError_t e = hipMemcpy(params, m->m.P, sizeof(float) * rv::SCR_NPARAM, hipMemcpyDeviceToHost);
    return e == hipSuccess ? RV_OK : RV_ERR_DEVICE;
}

const char* rv_mlp_last_error(const rv_mlp* m) { return m ? m->m.err.c_str() : ""; }

void rv_mlp_close(rv_mlp* m) { delete m; }

rv_status rv_set_screen(rv_engine* eh, const float* params, size_t n_params, float threshold) {
    rv::Engine* e = rv::engine_of(eh);
    if (!e) return RV_ERR_ARG;
    if (!params || threshold <= 0.f) {
        e->screen_on = false;
        return RV_OK;
    }
    if (n_params != (size_t)rv::SCR_NPARAM) return e->fail(RV_ERR_ARG, "screen model: wrong parameter count");
    void* d = e->grow(e->b_scr_params, sizeof(float) * rv::SCR_NPARAM);
    if (!d) return e->fail(RV_ERR_OOM, "device allocation for the screen model failed");
    if (hipMemcpy(d, params, sizeof(float) * rv::SCR_NPARAM, hipMemcpyHostToDevice) != hipSuccess)
        return e->fail(RV_ERR_DEVICE, "screen model upload failed");
    e->screen_on = true;
    e->screen_threshold = threshold;
    return RV_OK;
}

rv_status rv_screen_dataset(rv_engine* eh, const rv_region* region, const rv_reads* reads, const rv_contig* contig,
                            float** features, uint8_t** labels, int32_t** kinds, size_t* n) {
    rv::Engine* e = rv::engine_of(eh);
    if (!e || !region || !reads || !contig || !features || !labels || !kinds || !n) return RV_ERR_ARG;
    *features = nullptr;
    *labels = nullptr;
    *kinds = nullptr;
    *n = 0;
    const bool was_on = e->screen_on;
    e->screen_on = false;  // labels come from the exact realigner
    e->screen_collect = true;
    rv::CallSet cs;
    rv_status s;
    try {
        s = rv::run_region(*e, *region, *reads, *contig, cs);
    } catch (const std::exception& ex) {
        s = e->fail(RV_ERR_DEVICE, ex.what());
    }
    e->screen_collect = false;
    e->screen_on = was_on;
    if (s != RV_OK) return s;
    const size_t m = e->collect_label.size();
    if (m) {
        *features = (float*)malloc(sizeof(float) * rv::SCR_F * m);
        *labels = (uint8_t*)malloc(m);
        *kinds = (int32_t*)malloc(sizeof(int32_t) * m);
        if (!*features || !*labels || !*kinds) {
            free(*features);
            free(*labels);
            free(*kinds);
            *features = nullptr;
            *labels = nullptr;
            *kinds = nullptr;
            return RV_ERR_OOM;
        }
        memcpy(*features, e->collect_feat.data(), sizeof(float) * rv::SCR_F * m);
        memcpy(*labels, e->collect_label.data(), m);
        memcpy(*kinds, e->collect_kind.data(), sizeof(int32_t) * m);
    }
    *n = m;
    return RV_OK;
}

}  // extern "C"
