// mlp.hip -- the kernel selector's network: in -> h1 -> h2 -> out, ReLU, softmax cross-entropy,
// forward, backward and Adam on the GPU (FP32).  Sizes are small (24 -> 64 -> 64 -> 5), so the
// kernels are plain VALU code with one thread per output element; every reduction runs over
// the batch in a fixed order inside one thread, so gradients are deterministic.
//
// Parameter layout (one device buffer): W1[in][h1], b1[h1], W2[h1][h2], b2[h2], W3[h2][out], b3[out].
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "common.h"

struct sptrsv_mlp_s {
    int in = 0, h1 = 0, h2 = 0, out = 0;
    int nparams = 0;
    float* params = nullptr;
    float* m1 = nullptr;  // Adam moments
    float* m2 = nullptr;
    float* grad = nullptr;
    // work buffers (grown on demand)
    int cap = 0;
    float *x = nullptr, *a1 = nullptr, *a2 = nullptr, *z = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr,
          *mask = nullptr, *loss = nullptr;
    int* y = nullptr;
    hipStream_t stream = nullptr;
};

namespace sptrsv {
namespace {

struct Off {
    int w1, b1, w2, b2, w3, b3, total;
};
Off offsets(const sptrsv_mlp_s* M) {
    Off o;
    o.w1 = 0;
    o.b1 = o.w1 + M->in * M->h1;
    o.w2 = o.b1 + M->h1;
    o.b2 = o.w2 + M->h1 * M->h2;
    o.w3 = o.b2 + M->h2;
    o.b3 = o.w3 + M->h2 * M->out;
    o.total = o.b3 + M->out;
    return o;
}

// Y[b][j] = act(sum_i X[b][i] W[i][j] + c[j]), i ascending
__global__ void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ c,
                            float* __restrict__ Y, int B, int nin, int nout, int relu) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)B * nout) return;
    const int b = (int)(q / nout), j = (int)(q % nout);
    float acc = c[j];
    for (int i = 0; i < nin; ++i) acc = fmaf(X[(size_t)b * nin + i], W[(size_t)i * nout + j], acc);
    Y[q] = relu ? fmaxf(acc, 0.f) : acc;
}

// masked softmax cross-entropy: per-sample loss and dZ = (p - onehot(y)) / B over allowed classes
__global__ void k_softmax_ce(const float* __restrict__ Z, const int* __restrict__ y, const float* __restrict__ mask,
                             float* __restrict__ dZ, float* __restrict__ loss, int B, int K) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float* z = Z + (size_t)b * K;
    const float* mk = mask + (size_t)b * K;
    float mx = -INFINITY;
    for (int k = 0; k < K; ++k)
        if (mk[k] > 0.f) mx = fmaxf(mx, z[k]);
    float se = 0.f;
    for (int k = 0; k < K; ++k)
        if (mk[k] > 0.f) se += expf(z[k] - mx);
    const float lse = mx + logf(se);
    loss[b] = lse - z[y[b]];
    for (int k = 0; k < K; ++k) {
        const float p = mk[k] > 0.f ? expf(z[k] - lse) : 0.f;
        dZ[(size_t)b * K + k] = (p - (k == y[b] ? 1.f : 0.f)) / (float)B;
    }
}

// dW[i][j] = sum_b X[b][i] dY[b][j] (b ascending); db[j] = sum_b dY[b][j]
__global__ void k_dense_bwd_w(const float* __restrict__ X, const float* __restrict__ dY, float* __restrict__ dW,
                              float* __restrict__ db, int B, int nin, int nout) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nin * nout) {
        const int i = q / nout, j = q % nout;
        float acc = 0.f;
        for (int b = 0; b < B; ++b) acc = fmaf(X[(size_t)b * nin + i], dY[(size_t)b * nout + j], acc);
        dW[q] = acc;
    } else if (q < nin * nout + nout) {
        const int j = q - nin * nout;
        float acc = 0.f;
        for (int b = 0; b < B; ++b) acc += dY[(size_t)b * nout + j];
        db[j] = acc;
    }
}

// dX[b][i] = (sum_j dY[b][j] W[i][j]) * [A[b][i] > 0]   (A = the layer input, post-ReLU)
__global__ void k_dense_bwd_x(const float* __restrict__ dY, const float* __restrict__ W, const float* __restrict__ A,
                              float* __restrict__ dX, int B, int nin, int nout) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (int64_t)B * nin) return;
    const int b = (int)(q / nin), i = (int)(q % nin);
    float acc = 0.f;
    for (int j = 0; j < nout; ++j) acc = fmaf(dY[(size_t)b * nout + j], W[(size_t)i * nout + j], acc);
    dX[q] = A[q] > 0.f ? acc : 0.f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       int n, float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    m1[q] = beta1 * m1[q] + (1.f - beta1) * gq;
    m2[q] = beta2 * m2[q] + (1.f - beta2) * gq * gq;
    p[q] -= lr * (m1[q] / c1) / (sqrtf(m2[q] / c2) + eps);
}

inline unsigned blocks(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

int ensure_cap(sptrsv_mlp_s* M, int B) {
    if (B <= M->cap) return SPTRSV_OK;
    float** fb[] = {&M->x, &M->a1, &M->a2, &M->z, &M->d3, &M->d2, &M->d1, &M->mask, &M->loss};
    for (float** p : fb)
        if (*p) SPT_HIP(hipFree(*p));
    if (M->y) SPT_HIP(hipFree(M->y));
    SPT_HIP(hipMalloc((void**)&M->x, (size_t)B * M->in * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->a1, (size_t)B * M->h1 * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->a2, (size_t)B * M->h2 * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->z, (size_t)B * M->out * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->d3, (size_t)B * M->out * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->d2, (size_t)B * M->h2 * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->d1, (size_t)B * M->h1 * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->mask, (size_t)B * M->out * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->loss, (size_t)B * sizeof(float)));
    SPT_HIP(hipMalloc((void**)&M->y, (size_t)B * sizeof(int)));
    M->cap = B;
    return SPTRSV_OK;
}

void forward(sptrsv_mlp_s* M, int B) {
    const Off o = offsets(M);
    const float* P = M->params;
    k_dense_fwd<<<blocks((int64_t)B * M->h1), 256, 0, M->stream>>>(M->x, P + o.w1, P + o.b1, M->a1, B, M->in, M->h1, 1);
    k_dense_fwd<<<blocks((int64_t)B * M->h2), 256, 0, M->stream>>>(M->a1, P + o.w2, P + o.b2, M->a2, B, M->h1, M->h2, 1);
    k_dense_fwd<<<blocks((int64_t)B * M->out), 256, 0, M->stream>>>(M->a2, P + o.w3, P + o.b3, M->z, B, M->h2, M->out, 0);
}

void backward(sptrsv_mlp_s* M, int B) {
    const Off o = offsets(M);
    const float* P = M->params;
    float* G = M->grad;
    k_softmax_ce<<<blocks(B), 256, 0, M->stream>>>(M->z, M->y, M->mask, M->d3, M->loss, B, M->out);
    k_dense_bwd_w<<<blocks(M->h2 * M->out + M->out), 256, 0, M->stream>>>(M->a2, M->d3, G + o.w3, G + o.b3, B, M->h2, M->out);
    k_dense_bwd_x<<<blocks((int64_t)B * M->h2), 256, 0, M->stream>>>(M->d3, P + o.w3, M->a2, M->d2, B, M->h2, M->out);
    k_dense_bwd_w<<<blocks(M->h1 * M->h2 + M->h2), 256, 0, M->stream>>>(M->a1, M->d2, G + o.w2, G + o.b2, B, M->h1, M->h2);
    k_dense_bwd_x<<<blocks((int64_t)B * M->h1), 256, 0, M->stream>>>(M->d2, P + o.w2, M->a1, M->d1, B, M->h1, M->h2);
    k_dense_bwd_w<<<blocks(M->in * M->h1 + M->h1), 256, 0, M->stream>>>(M->x, M->d1, G + o.w1, G + o.b1, B, M->in, M->h1);
}

int upload_batch(sptrsv_mlp_s* M, const float* X, const int* y, const float* mask, int B) {
    SPT_CHECK(ensure_cap(M, B));
    SPT_HIP(hipMemcpy(M->x, X, (size_t)B * M->in * sizeof(float), hipMemcpyHostToDevice));
    if (y) SPT_HIP(hipMemcpy(M->y, y, (size_t)B * sizeof(int), hipMemcpyHostToDevice));
    if (mask) {
        SPT_HIP(hipMemcpy(M->mask, mask, (size_t)B * M->out * sizeof(float), hipMemcpyHostToDevice));
    } else {
        std::vector<float> ones((size_t)B * M->out, 1.f);
        SPT_HIP(hipMemcpy(M->mask, ones.data(), ones.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return SPTRSV_OK;
}

}  // namespace
}  // namespace sptrsv

using namespace sptrsv;

extern "C" {

int sptrsv_mlp_create(int in, int h1, int h2, int out, unsigned seed, sptrsv_mlp_t* mlp) {
    if (!mlp || in <= 0 || h1 <= 0 || h2 <= 0 || out <= 1) return fail(SPTRSV_ERR_BAD_INPUT, "bad MLP sizes");
    SPT_CHECK(ensure_device());
    auto* M = new sptrsv_mlp_s();
    M->in = in;
    M->h1 = h1;
    M->h2 = h2;
    M->out = out;
    const Off o = offsets(M);
    M->nparams = o.total;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) from a fixed LCG, biases zero
    std::vector<float> p((size_t)o.total, 0.f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto uni = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((double)(st >> 11) / 9007199254740992.0) * 2.f - 1.f;
    };
    auto fill = [&](int off, int count, int fan_in) {
        const float s = 1.f / std::sqrt((float)fan_in);
        for (int k = 0; k < count; ++k) p[(size_t)(off + k)] = s * uni();
    };
    fill(o.w1, in * h1, in);
    fill(o.w2, h1 * h2, h1);
    fill(o.w3, h2 * out, h2);
    hipError_t e = hipMalloc((void**)&M->params, (size_t)o.total * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M->m1, (size_t)o.total * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M->m2, (size_t)o.total * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&M->grad, (size_t)o.total * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(M->params, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice);
    std::vector<float> zero((size_t)o.total, 0.f);
    if (e == hipSuccess) e = hipMemcpy(M->m1, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(M->m2, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        sptrsv_mlp_destroy(M);
        return fail(SPTRSV_ERR_HIP, std::string("MLP setup: ") + hipGetErrorString(e));
    }
    *mlp = M;
    return SPTRSV_OK;
}

int sptrsv_mlp_destroy(sptrsv_mlp_t M) {
    if (!M) return SPTRSV_OK;
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    void* bufs[] = {M->params, M->m1, M->m2, M->grad, M->x, M->a1, M->a2, M->z, M->d3, M->d2, M->d1, M->mask, M->loss, M->y};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
    return SPTRSV_OK;
}

int sptrsv_mlp_num_params(sptrsv_mlp_t M, int* count) {
    if (!M || !count) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    *count = M->nparams;
    return SPTRSV_OK;
}

int sptrsv_mlp_get_params(sptrsv_mlp_t M, float* host) {
    if (!M || !host) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    SPT_HIP(hipStreamSynchronize(M->stream));
    SPT_HIP(hipMemcpy(host, M->params, (size_t)M->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

int sptrsv_mlp_set_params(sptrsv_mlp_t M, const float* host) {
    if (!M || !host) return fail(SPTRSV_ERR_BAD_INPUT, "null argument");
    SPT_HIP(hipStreamSynchronize(M->stream));
    SPT_HIP(hipMemcpy(M->params, host, (size_t)M->nparams * sizeof(float), hipMemcpyHostToDevice));
    return SPTRSV_OK;
}

// Forward + backward on a host batch: per-sample losses (B) and the gradient (nparams) of the
// mean loss.  mask (B x out, 1 = allowed class) may be NULL.
int sptrsv_mlp_loss_grad(sptrsv_mlp_t M, const float* X, const int* y, const float* mask, int B, float* losses,
                         float* grad) {
    if (!M || !X || !y || B <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "bad loss_grad arguments");
    for (int b = 0; b < B; ++b)
        if (y[b] < 0 || y[b] >= M->out) return fail(SPTRSV_ERR_BAD_INPUT, "label out of range");
    SPT_CHECK(upload_batch(M, X, y, mask, B));
    forward(M, B);
    backward(M, B);
    SPT_HIP(hipGetLastError());
    SPT_HIP(hipStreamSynchronize(M->stream));
    if (losses) SPT_HIP(hipMemcpy(losses, M->loss, (size_t)B * sizeof(float), hipMemcpyDeviceToHost));
    if (grad) SPT_HIP(hipMemcpy(grad, M->grad, (size_t)M->nparams * sizeof(float), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

// Logits (B x out) of a host batch.
int sptrsv_mlp_forward(sptrsv_mlp_t M, const float* X, int B, float* logits) {
    if (!M || !X || !logits || B <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "bad forward arguments");
    SPT_CHECK(upload_batch(M, X, nullptr, nullptr, B));
    forward(M, B);
    SPT_HIP(hipGetLastError());
    SPT_HIP(hipStreamSynchronize(M->stream));
    SPT_HIP(hipMemcpy(logits, M->z, (size_t)B * M->out * sizeof(float), hipMemcpyDeviceToHost));
    return SPTRSV_OK;
}

// Full-batch Adam training on the device: `epochs` forward/backward/update passes, the data
// uploaded once.  *final_loss = mean loss of the last pass (before its update).
int sptrsv_mlp_train(sptrsv_mlp_t M, const float* X, const int* y, const float* mask, int B, int epochs, float lr,
                     float* final_loss) {
    if (!M || !X || !y || B <= 0 || epochs <= 0) return fail(SPTRSV_ERR_BAD_INPUT, "bad train arguments");
    for (int b = 0; b < B; ++b)
        if (y[b] < 0 || y[b] >= M->out) return fail(SPTRSV_ERR_BAD_INPUT, "label out of range");
    SPT_CHECK(upload_batch(M, X, y, mask, B));
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float c1 = 1.f, c2 = 1.f;
    for (int e = 0; e < epochs; ++e) {
        forward(M, B);
        backward(M, B);
        c1 *= beta1;
        c2 *= beta2;
        k_adam<<<blocks(M->nparams), 256, 0, M->stream>>>(M->params, M->grad, M->m1, M->m2, M->nparams, lr, beta1,
                                                           beta2, eps, 1.f - c1, 1.f - c2);
    }
    SPT_HIP(hipGetLastError());
    SPT_HIP(hipStreamSynchronize(M->stream));
    if (final_loss) {
        std::vector<float> l((size_t)B);
        SPT_HIP(hipMemcpy(l.data(), M->loss, l.size() * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : l) s += v;
        *final_loss = (float)(s / B);
    }
    return SPTRSV_OK;
}

}  // extern "C"
