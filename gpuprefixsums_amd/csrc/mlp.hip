// A small fully connected cost model trained and evaluated on the GPU:
//   z1 = W1 x + b1, a1 = relu(z1)      (hidden H)
//   z2 = W2 a1 + b2, a2 = relu(z2)     (hidden H)
//   y  = W3 a2 + b3                    (scalar: predicted log2 time)
// trained with mean squared error and Adam. Parameters are one flat FP32
// array [W1 (H x D, row-major), b1 (H), W2 (H x H), b2 (H), W3 (H), b3 (1)].
// Layers are a few thousand multiply-adds per sample, so every kernel is a
// plain one-thread-per-output loop (deterministic summation order, no
// atomics); matrix cores would not pay at this size. Weight gradients sum
// over the batch, so they are split into fixed chunks of samples whose
// partial sums are added in chunk order: thousands of threads instead of
// one per weight, and the same order on every run.
//
// No reference counterpart (an extension beside the reference's scans). Compiled with
// -ffp-contract=off so results are reproducible against the CPU oracle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "handle.hpp"

using gps::fail;
using gps::hip_fail;

namespace {

constexpr int kThreads = 256;

inline dim3 grid_of(uint64_t n) { return dim3((uint32_t)((n + kThreads - 1) / kThreads)); }

// Y[b, o] = act(b[o] + sum_i W[o, i] X[b, i]); Z keeps the pre-activation.
__global__ void linear_fwd(const float* X, const float* W, const float* bias, float* Z, float* Y, uint32_t B,
                           uint32_t Din, uint32_t Dout, int relu) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)B * Dout) return;
    const uint32_t b = (uint32_t)(t / Dout), o = (uint32_t)(t % Dout);
    float acc = bias[o];
    const float* x = X + (uint64_t)b * Din;
    const float* w = W + (uint64_t)o * Din;
    for (uint32_t i = 0; i < Din; ++i) acc += w[i] * x[i];
    if (Z) Z[t] = acc;
    Y[t] = relu ? fmaxf(acc, 0.f) : acc;
}

// dY = 2 (y - target) / B; per-block partial sums of (y - target)^2.
__global__ void mse_grad(const float* y, const float* target, float* dy, float* partial, uint32_t B) {
    __shared__ float s[kThreads];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.f;
    if (i < B) {
        const float d = y[i] - target[i];
        dy[i] = 2.f * d / (float)B;
        e = d * d;
    }
    s[threadIdx.x] = e;
    __syncthreads();
    for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void sum_partials(const float* partial, uint32_t n, float scale, float* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.f;
        for (uint32_t i = 0; i < n; ++i) s += partial[i];
        *out = s * scale;
    }
}

// Weight gradients in two deterministic passes. The batch is cut into chunks
// of kChunk samples; pass 1 writes, per chunk c and output t,
//   part[c, t] = sum_{b in chunk c} dZ[b, o] X[b, i]   (t = o Din + i), or
//   part[c, t] = sum_{b in chunk c} dZ[b, o]           (t = Dout Din + o, the bias),
// in increasing b; pass 2 sums the chunks of each output in increasing c.
// The summation order is fixed, so results do not depend on scheduling.
constexpr uint32_t kChunk = 128;

inline uint32_t chunks_of(uint32_t B) { return (B + kChunk - 1) / kChunk; }

__global__ void linear_bwd_w_part(const float* dZ, const float* X, float* part, uint32_t B, uint32_t Din,
                                  uint32_t Dout) {
    const uint64_t outs = (uint64_t)Dout * Din + Dout;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= outs) return;
    const uint32_t c = blockIdx.y;
    const uint32_t b0 = c * kChunk, b1 = min(B, b0 + kChunk);
    float acc = 0.f;
    if (t < (uint64_t)Dout * Din) {
        const uint32_t o = (uint32_t)(t / Din), i = (uint32_t)(t % Din);
        for (uint32_t b = b0; b < b1; ++b) acc += dZ[(uint64_t)b * Dout + o] * X[(uint64_t)b * Din + i];
    } else {
        const uint32_t o = (uint32_t)(t - (uint64_t)Dout * Din);
        for (uint32_t b = b0; b < b1; ++b) acc += dZ[(uint64_t)b * Dout + o];
    }
    part[(uint64_t)c * outs + t] = acc;
}

// dW (Dout x Din) and db (Dout) follow each other in the flat gradient array.
__global__ void linear_bwd_w_sum(const float* part, float* dWb, uint64_t outs, uint32_t chunks) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= outs) return;
    float acc = 0.f;
    for (uint32_t c = 0; c < chunks; ++c) acc += part[(uint64_t)c * outs + t];
    dWb[t] = acc;
}

// dZprev[b, i] = (sum_o dZ[b, o] W[o, i]) * (Zprev[b, i] > 0)   (through relu).
__global__ void linear_bwd_x(const float* dZ, const float* W, const float* Zprev, float* dZprev, uint32_t B,
                             uint32_t Din, uint32_t Dout) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)B * Din) return;
    const uint32_t b = (uint32_t)(t / Din), i = (uint32_t)(t % Din);
    float acc = 0.f;
    for (uint32_t o = 0; o < Dout; ++o) acc += dZ[(uint64_t)b * Dout + o] * W[(uint64_t)o * Din + i];
    dZprev[t] = Zprev[t] > 0.f ? acc : 0.f;
}

// Adam with decoupled weight decay (AdamW); step is 1-based.
__global__ void adam_step(float* p, const float* g, float* m, float* v, uint64_t count, float lr, float beta1,
                          float beta2, float eps, float wd, float bc1, float bc2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    p[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
}

}  // namespace

struct gps_mlp {
    uint32_t D = 0, H = 0, max_batch = 0;
    uint64_t count = 0;
    uint64_t step = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    float *x = nullptr, *t = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr;
    float *dy = nullptr, *dz2 = nullptr, *dz1 = nullptr, *partial = nullptr, *loss = nullptr;
    float* wpart = nullptr;  // per-chunk weight-gradient partial sums

    float* W1() const { return params; }
    float* b1() const { return params + (uint64_t)H * D; }
    float* W2() const { return b1() + H; }
    float* b2() const { return W2() + (uint64_t)H * H; }
    float* W3() const { return b2() + H; }
    float* b3() const { return W3() + H; }
    float* gW1() const { return grads; }
    float* gb1() const { return grads + (uint64_t)H * D; }
    float* gW2() const { return gb1() + H; }
    float* gb2() const { return gW2() + (uint64_t)H * H; }
    float* gW3() const { return gb2() + H; }
    float* gb3() const { return gW3() + H; }
};

namespace {

uint64_t param_count(uint32_t D, uint32_t H) { return (uint64_t)H * D + H + (uint64_t)H * H + H + H + 1; }

void free_mlp(gps_mlp* m) {
    if (!m) return;
    float* ptrs[] = {m->params, m->grads, m->m, m->v,  m->x,       m->t,    m->z1,   m->a1, m->z2,
                     m->a2,     m->y,     m->dy, m->dz2, m->dz1, m->partial, m->loss, m->wpart};
    for (float* p : ptrs) (void)hipFree(p);
    delete m;
}

hipError_t forward(gps_mlp* m, uint32_t B) {
    const uint32_t D = m->D, H = m->H;
    hipLaunchKernelGGL(linear_fwd, grid_of((uint64_t)B * H), dim3(kThreads), 0, nullptr, m->x, m->W1(), m->b1(),
                       m->z1, m->a1, B, D, H, 1);
    hipLaunchKernelGGL(linear_fwd, grid_of((uint64_t)B * H), dim3(kThreads), 0, nullptr, m->a1, m->W2(), m->b2(),
                       m->z2, m->a2, B, H, H, 1);
    hipLaunchKernelGGL(linear_fwd, grid_of(B), dim3(kThreads), 0, nullptr, m->a2, m->W3(), m->b3(), nullptr, m->y,
                       B, H, 1u, 0);
    return hipGetLastError();
}

// dW and db of one layer (Din inputs, Dout outputs) into the flat gradient
// array at dWb, through the per-chunk partials.
void linear_bwd_w(gps_mlp* m, const float* dZ, const float* X, float* dWb, uint32_t B, uint32_t Din, uint32_t Dout) {
    const uint64_t outs = (uint64_t)Dout * Din + Dout;
    const uint32_t chunks = chunks_of(B);
    const dim3 g1((uint32_t)((outs + kThreads - 1) / kThreads), chunks);
    hipLaunchKernelGGL(linear_bwd_w_part, g1, dim3(kThreads), 0, nullptr, dZ, X, m->wpart, B, Din, Dout);
    hipLaunchKernelGGL(linear_bwd_w_sum, grid_of(outs), dim3(kThreads), 0, nullptr, m->wpart, dWb, outs, chunks);
}

// Loss (device scalar at `loss`) and all parameter gradients for the batch in m->x / m->t.
hipError_t loss_and_grads(gps_mlp* m, uint32_t B, float* loss) {
    const uint32_t D = m->D, H = m->H;
    hipError_t e = forward(m, B);
    if (e != hipSuccess) return e;
    const dim3 gb = grid_of(B);
    hipLaunchKernelGGL(mse_grad, gb, dim3(kThreads), 0, nullptr, m->y, m->t, m->dy, m->partial, B);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(64), 0, nullptr, m->partial, gb.x, 1.f / (float)B, loss);
    // layer 3: dz3 = dy
    linear_bwd_w(m, m->dy, m->a2, m->gW3(), B, H, 1u);
    hipLaunchKernelGGL(linear_bwd_x, grid_of((uint64_t)B * H), dim3(kThreads), 0, nullptr, m->dy, m->W3(), m->z2,
                       m->dz2, B, H, 1u);
    // layer 2
    linear_bwd_w(m, m->dz2, m->a1, m->gW2(), B, H, H);
    hipLaunchKernelGGL(linear_bwd_x, grid_of((uint64_t)B * H), dim3(kThreads), 0, nullptr, m->dz2, m->W2(), m->z1,
                       m->dz1, B, H, H);
    // layer 1 (no gradient into x)
    linear_bwd_w(m, m->dz1, m->x, m->gW1(), B, D, H);
    return hipGetLastError();
}

int stage_batch(gps_mlp* m, const float* h_x, const float* h_t, uint32_t batch) {
    if (!m || !h_x || batch == 0) return fail(GPS_ERR_INVALID, "bad MLP arguments");
    if (batch > m->max_batch) return fail(GPS_ERR_TOO_LARGE, "batch %u exceeds max_batch %u", batch, m->max_batch);
    HIP_TRY(hipMemcpy(m->x, h_x, (size_t)batch * m->D * sizeof(float), hipMemcpyHostToDevice));
    if (h_t) HIP_TRY(hipMemcpy(m->t, h_t, (size_t)batch * sizeof(float), hipMemcpyHostToDevice));
    return GPS_OK;
}

}  // namespace

extern "C" {

uint64_t gps_mlp_param_count(uint32_t in_dim, uint32_t hidden) { return param_count(in_dim, hidden); }

int gps_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t max_batch, gps_mlp** out) {
    if (!out) return fail(GPS_ERR_INVALID, "null output pointer");
    *out = nullptr;
    if (!in_dim || !hidden || !max_batch || in_dim > 4096 || hidden > 4096)
        return fail(GPS_ERR_INVALID, "bad MLP dimensions");
    if (chunks_of(max_batch) > 65535u)  // the chunk index is the weight-gradient grid's y dimension
        return fail(GPS_ERR_TOO_LARGE, "max_batch %u exceeds %u samples", max_batch, 65535u * kChunk);
    gps_mlp* m = new (std::nothrow) gps_mlp();
    if (!m) return fail(GPS_ERR_INVALID, "out of host memory");
    m->D = in_dim;
    m->H = hidden;
    m->max_batch = max_batch;
    m->count = param_count(in_dim, hidden);
    const uint64_t B = max_batch;
    const uint32_t nblocks = (max_batch + kThreads - 1) / kThreads;
    struct A {
        float** p;
        uint64_t n;
    } allocs[] = {{&m->params, m->count}, {&m->grads, m->count}, {&m->m, m->count},    {&m->v, m->count},
                  {&m->x, B * in_dim},    {&m->t, B},            {&m->z1, B * hidden}, {&m->a1, B * hidden},
                  {&m->z2, B * hidden},   {&m->a2, B * hidden},  {&m->y, B},           {&m->dy, B},
                  {&m->dz2, B * hidden},  {&m->dz1, B * hidden}, {&m->partial, nblocks}, {&m->loss, 1},
                  {&m->wpart, ((uint64_t)hidden * std::max(in_dim, hidden) + hidden) * chunks_of(max_batch)}};
    for (auto& a : allocs) {
        const hipError_t e = hipMalloc(a.p, a.n * sizeof(float));
        if (e != hipSuccess) {
            free_mlp(m);
            return hip_fail(e, "gps_mlp_create");
        }
    }
    hipError_t e = hipMemset(m->params, 0, m->count * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->m, 0, m->count * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->v, 0, m->count * sizeof(float));
    if (e != hipSuccess) {
        free_mlp(m);
        return hip_fail(e, "gps_mlp_create");
    }
    *out = m;
    return GPS_OK;
}

int gps_mlp_destroy(gps_mlp* m) {
    if (!m) return fail(GPS_ERR_INVALID, "null MLP");
    free_mlp(m);
    return GPS_OK;
}

int gps_mlp_set_params(gps_mlp* m, const float* h_params) {
    if (!m || !h_params) return fail(GPS_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpy(m->params, h_params, m->count * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->m, 0, m->count * sizeof(float)));
    HIP_TRY(hipMemset(m->v, 0, m->count * sizeof(float)));
    m->step = 0;
    return GPS_OK;
}

int gps_mlp_get_params(gps_mlp* m, float* h_params) {
    if (!m || !h_params) return fail(GPS_ERR_INVALID, "null argument");
    HIP_TRY(hipMemcpy(h_params, m->params, m->count * sizeof(float), hipMemcpyDeviceToHost));
    return GPS_OK;
}

int gps_mlp_forward(gps_mlp* m, const float* h_x, uint32_t batch, float* h_y) {
    if (!h_y) return fail(GPS_ERR_INVALID, "null output");
    int rc = stage_batch(m, h_x, nullptr, batch);
    if (rc) return rc;
    HIP_TRY(forward(m, batch));
    HIP_TRY(hipMemcpy(h_y, m->y, (size_t)batch * sizeof(float), hipMemcpyDeviceToHost));
    return GPS_OK;
}

int gps_mlp_loss_grad(gps_mlp* m, const float* h_x, const float* h_t, uint32_t batch, float* h_loss,
                      float* h_grad) {
    if (!h_t) return fail(GPS_ERR_INVALID, "null targets");
    int rc = stage_batch(m, h_x, h_t, batch);
    if (rc) return rc;
    HIP_TRY(loss_and_grads(m, batch, m->loss));
    if (h_loss) HIP_TRY(hipMemcpy(h_loss, m->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (h_grad) HIP_TRY(hipMemcpy(h_grad, m->grads, m->count * sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipDeviceSynchronize());
    return GPS_OK;
}

int gps_mlp_train(gps_mlp* m, const float* h_x, const float* h_t, uint32_t batch, uint32_t epochs, float lr,
                  float weight_decay, float* h_loss_curve) {
    int rc = stage_batch(m, h_x, h_t, batch);
    if (rc) return rc;
    if (!h_t) return fail(GPS_ERR_INVALID, "null targets");
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    // The epoch losses stay on the device until the end (one copy, no
    // per-epoch synchronisation).
    float* d_curve = nullptr;
    if (h_loss_curve && epochs) HIP_TRY(hipMalloc(&d_curve, (size_t)epochs * sizeof(float)));
    hipError_t e = hipSuccess;
    for (uint32_t ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        e = loss_and_grads(m, batch, d_curve ? d_curve + ep : m->loss);
        if (e != hipSuccess) break;
        ++m->step;
        const float bc1 = 1.f - std::pow(beta1, (float)m->step);
        const float bc2 = 1.f - std::pow(beta2, (float)m->step);
        hipLaunchKernelGGL(adam_step, grid_of(m->count), dim3(kThreads), 0, nullptr, m->params, m->grads, m->m,
                           m->v, m->count, lr, beta1, beta2, eps, weight_decay, bc1, bc2);
        e = hipGetLastError();
    }
    if (e == hipSuccess && d_curve)
        e = hipMemcpy(h_loss_curve, d_curve, (size_t)epochs * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (d_curve) (void)hipFree(d_curve);
    if (e != hipSuccess) return hip_fail(e, "gps_mlp_train");
    return GPS_OK;
}

}  // extern "C"
