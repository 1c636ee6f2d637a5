// select.hip -- the learned kernel-variant selector of a single plan (host code).
//
// AUTO's hand rule (solve.hip: default_variant) looks at three properties of a factor.  A
// selector network, trained offline on the autotuner's timings of every variant over generated
// factors (sptrsv_amd/tune_single.py), reads the 24 features the batched selector uses (size,
// level count and widths, row lengths, offsets, supernodes, log2 m) and scores every variant;
// once installed with sptrsv_selector_set, AUTO takes the best-scored variant among those valid
// for the plan, m and precision, and the hand rule stays the fallback (no network installed, or
// no valid variant scored).  The network is 24 -> h1 -> h2 -> K with ReLU, the layout of the
// GPU-trained MLP (mlp.hip); its forward pass here is a few thousand multiply-adds on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "plan.h"

namespace sptrsv {

namespace {

struct Net {
    int in = 0, h1 = 0, h2 = 0, out = 0;
    std::vector<float> params, mu, sd;
    std::vector<int> classes;  // output index -> SPTRSV_VAR_*
};
std::mutex g_net_mutex;
std::shared_ptr<const Net> g_net;  // null: no network installed

// The factor's 24 features (feature 22, log2 m, filled per call); computed once per plan from
// its CSR and levels.
int plan_features_cached(sptrsv_plan_s* P, int m, float* f) {
    if (!P->feat_ok) {
        const int n = P->n;
        const int64_t nnz = P->nnz;
        std::vector<int> rp32((size_t)n + 1), col((size_t)nnz), lev((size_t)n);
        std::vector<double> val((size_t)nnz);
        SPT_HIP(hipStreamSynchronize(P->stream));
        SPT_HIP(hipMemcpy(rp32.data(), P->rowptr, rp32.size() * sizeof(int), hipMemcpyDeviceToHost));
        SPT_HIP(hipMemcpy(col.data(), P->col, col.size() * sizeof(int), hipMemcpyDeviceToHost));
        SPT_HIP(hipMemcpy(val.data(), P->val, val.size() * sizeof(double), hipMemcpyDeviceToHost));
        SPT_HIP(hipMemcpy(lev.data(), P->levels, lev.size() * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int64_t> rp(rp32.begin(), rp32.end());
        system_features(rp.data(), col.data(), val.data(), lev, 0, n, 1, P->feat);
        P->feat_ok = true;
    }
    for (int k = 0; k < 24; ++k) f[k] = P->feat[k];
    f[22] = (float)std::log2((double)m);
    return SPTRSV_OK;
}

// Variants that can run this solve.  SUPERNODAL needs a supernodal layout: one the caller built,
// or relaxed supernodes where the hand rule would build them too (built here on first use, so
// this is asked only of the best-scored variant).
bool variant_valid(sptrsv_plan_s* P, int v, int m, int precision, bool exact) {
    const bool fp64 = precision == SPTRSV_FP64;
    switch (v) {
        case SPTRSV_VAR_LEVEL_PULL:
        case SPTRSV_VAR_ROW_PULL:
        case SPTRSV_VAR_LEVELSET: return true;
        case SPTRSV_VAR_CSC_PUSH: return !exact;
        case SPTRSV_VAR_CHAIN: return !exact && m == 1;
        case SPTRSV_VAR_TILE: return fp64 && m == 1;
        case SPTRSV_VAR_BLOCK: return fp64 && (m == 1 || (P->grid && m <= 3 && block_supports(P, m)));
        case SPTRSV_VAR_PANEL: return fp64 && panel_supports(P);
        case SPTRSV_VAR_SUPERNODAL: return !exact && fp64 && (P->sn > 0 || relaxed_supernodes_pay(P));
        default: return false;
    }
}

}  // namespace

// AUTO's choice by the installed network; false when none is installed or none of the variants
// it scores is valid here.
bool selector_pick(sptrsv_plan_s* P, int m, int precision, bool exact, int* variant) {
    std::shared_ptr<const Net> held;  // the network stays alive while it is read, even if replaced
    {
        std::lock_guard<std::mutex> lk(g_net_mutex);
        held = g_net;
    }
    if (!held) return false;
    const Net& net = *held;
    float x[24];
    if (net.in != 24 || plan_features_cached(P, m, x) != SPTRSV_OK) return false;
    for (int k = 0; k < 24; ++k) x[k] = (x[k] - net.mu[(size_t)k]) / net.sd[(size_t)k];
    const float* W1 = net.params.data();
    const float* b1 = W1 + (size_t)net.in * net.h1;
    const float* W2 = b1 + net.h1;
    const float* b2 = W2 + (size_t)net.h1 * net.h2;
    const float* W3 = b2 + net.h2;
    const float* b3 = W3 + (size_t)net.h2 * net.out;
    std::vector<float> h1((size_t)net.h1), h2((size_t)net.h2);
    for (int j = 0; j < net.h1; ++j) {
        float s = b1[j];
        for (int i = 0; i < net.in; ++i) s += x[i] * W1[(size_t)i * net.h1 + j];
        h1[(size_t)j] = std::max(s, 0.0f);
    }
    for (int j = 0; j < net.h2; ++j) {
        float s = b2[j];
        for (int i = 0; i < net.h1; ++i) s += h1[(size_t)i] * W2[(size_t)i * net.h2 + j];
        h2[(size_t)j] = std::max(s, 0.0f);
    }
    std::vector<std::pair<float, int>> score((size_t)net.out);
    for (int k = 0; k < net.out; ++k) {
        float s = b3[k];
        for (int i = 0; i < net.h2; ++i) s += h2[(size_t)i] * W3[(size_t)i * net.out + k];
        score[(size_t)k] = {s, k};
    }
    // best score first (ties: the lower output index); the first valid variant is the choice
    std::stable_sort(score.begin(), score.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    for (const auto& [s, k] : score) {
        const int v = net.classes[(size_t)k];
        if (variant_valid(P, v, m, precision, exact)) {
            *variant = v;
            return true;
        }
    }
    return false;
}

}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_plan_features(sptrsv_plan_t P, int m, float* features) {
    if (!P || !features || m < 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad features arguments");
    return plan_features_cached(P, m, features);
}

int sptrsv_selector_set(int in, int h1, int h2, int out, const float* params, const float* mu, const float* sd,
                        const int* classes) {
    if (in != 24 || h1 < 1 || h2 < 1 || out < 1 || !params || !mu || !sd || !classes)
        return fail(SPTRSV_ERR_BAD_INPUT, "selector: sizes 24 x h1 x h2 x out and all arrays required");
    for (int k = 0; k < out; ++k)
        if (classes[k] < 1 || classes[k] >= SPTRSV_VAR_COUNT) return fail(SPTRSV_ERR_BAD_INPUT, "selector: bad variant id");
    for (int k = 0; k < in; ++k)
        if (!(sd[k] > 0.0f) || !std::isfinite(sd[k]) || !std::isfinite(mu[k]))
            return fail(SPTRSV_ERR_BAD_INPUT, "selector: feature scales must be finite, sd > 0");
    auto net = std::make_shared<Net>();
    net->in = in;
    net->h1 = h1;
    net->h2 = h2;
    net->out = out;
    const size_t np = (size_t)in * h1 + h1 + (size_t)h1 * h2 + h2 + (size_t)h2 * out + out;
    net->params.assign(params, params + np);
    net->mu.assign(mu, mu + in);
    net->sd.assign(sd, sd + in);
    net->classes.assign(classes, classes + out);
    std::lock_guard<std::mutex> lk(g_net_mutex);
    g_net = std::move(net);
    return SPTRSV_OK;
}

int sptrsv_selector_clear(void) {
    std::lock_guard<std::mutex> lk(g_net_mutex);
    g_net.reset();
    return SPTRSV_OK;
}

}  // extern "C"
