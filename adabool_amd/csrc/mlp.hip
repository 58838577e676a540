// mlp.hip -- FP32 cost model for per-query dispatch: 12 -> 64 -> 64 -> 2
// (ReLU after the two hidden layers), predicting log execution time of the
// dense and the sparse path.  Inputs are standardized by per-feature mean and
// standard deviation kept with the model.
//
// Kernels (all hand-written, FP32):
//   k_mlp_decide  -- one wave per query, one lane per hidden unit, weights
//                    staged transposed in LDS; writes the chosen path (and
//                    optionally the two predictions).
//   k_mlp_step    -- one work-group runs a whole minibatch step: forward with
//                    activations in a global scratch, backward layer by layer
//                    (dW = dZ^T A reductions split over the 256 threads), then
//                    Adam with bias correction, or only the gradient.
// Loss: mean over samples and outputs of 0.5 * (y_hat - y)^2.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"

namespace ab {

// parameter offsets
constexpr int oW1 = 0, ob1 = oW1 + kH1 * kF, oW2 = ob1 + kH1, ob2 = oW2 + kH2 * kH1, oW3 = ob2 + kH2,
              ob3 = oW3 + kO * kH2;
static_assert(ob3 + kO == kParams, "parameter layout");

constexpr int kStepNT = 256;
constexpr int kMaxBatch = 4096;

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

// One sample forward; p in LDS.  Returns outputs; optionally keeps z1/z2.
__device__ __forceinline__ void mlp_fwd1(const float *__restrict__ p, const float *__restrict__ x, float *z1out,
                                         float *z2out, float (&y)[kO])
{
    float h1[kH1];
#pragma unroll
    for (int j = 0; j < kH1; ++j) {
        float s = p[ob1 + j];
#pragma unroll
        for (int f = 0; f < kF; ++f) s = fmaf(p[oW1 + j * kF + f], x[f], s);
        if (z1out) z1out[j] = s;
        h1[j] = relu(s);
    }
    y[0] = p[ob3];
    y[1] = p[ob3 + 1];
    for (int j = 0; j < kH2; ++j) {
        float s = p[ob2 + j];
#pragma unroll
        for (int i = 0; i < kH1; ++i) s = fmaf(p[oW2 + j * kH1 + i], h1[i], s);
        if (z2out) z2out[j] = s;
        const float h = relu(s);
        y[0] = fmaf(p[oW3 + j], h, y[0]);
        y[1] = fmaf(p[oW3 + kH2 + j], h, y[1]);
    }
}

__device__ __forceinline__ void load_x(const float *__restrict__ X, uint32_t row, const float *__restrict__ norm,
                                       float (&x)[kF])
{
#pragma unroll
    for (int f = 0; f < kF; ++f) x[f] = (X[(uint64_t)row * kF + f] - norm[f]) / norm[kF + f];
}

// One wave per query, lane j computing hidden unit j of both layers (and
// lanes 0..kO-1 the outputs).  Every unit accumulates its fmaf terms in the
// same order as mlp_fwd1, so the predictions are bit-identical to it.  The
// weight matrices are staged transposed (W^T[i][j] at i * rows + j) so the 64
// lanes of a wave read 64 consecutive floats: no LDS bank conflicts.
constexpr int kDecideQ = 4;  // queries (waves) per work-group
static_assert(kH1 == 64 && kH2 == 64 && kO <= 64, "one lane per hidden unit");
__global__ __launch_bounds__(64 * kDecideQ) void k_mlp_decide(const float *__restrict__ params,
                                                             const float *__restrict__ norm,
                                                             const float *__restrict__ X, uint32_t n, uint8_t *sel,
                                                             float *pred)
{
    __shared__ float w1t[kF * kH1];   // w1t[f * kH1 + j] = W1[j][f]
    __shared__ float w2t[kH1 * kH2];  // w2t[i * kH2 + j] = W2[j][i]
    __shared__ float nm[2 * kF];
    __shared__ float hb[kDecideQ][64];
    for (int i = threadIdx.x; i < kF * kH1; i += 64 * kDecideQ) {
        const int j = i / kF, f = i - j * kF;
        w1t[f * kH1 + j] = params[oW1 + i];
    }
    for (int i = threadIdx.x; i < kH1 * kH2; i += 64 * kDecideQ) {
        const int j = i / kH1, k = i - j * kH1;
        w2t[k * kH2 + j] = params[oW2 + i];
    }
    if (threadIdx.x < 2 * kF) nm[threadIdx.x] = norm[threadIdx.x];
    __syncthreads();
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * kDecideQ + wave;
    if (q >= n) return;  // no work-group barrier below this point
    float x[kF];
    load_x(X, q, nm, x);
    float *h = hb[wave];
    {
        float s = params[ob1 + lane];
#pragma unroll
        for (int f = 0; f < kF; ++f) s = fmaf(w1t[f * kH1 + lane], x[f], s);
        h[lane] = relu(s);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float s = params[ob2 + lane];
#pragma unroll 16
    for (int i = 0; i < kH1; ++i) s = fmaf(w2t[i * kH2 + lane], h[i], s);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    h[lane] = relu(s);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float y = 0.f;
    if (lane < (uint32_t)kO) {
        y = params[ob3 + lane];
#pragma unroll 16
        for (int j = 0; j < kH2; ++j) y = fmaf(params[oW3 + lane * kH2 + j], h[j], y);
    }
    const float y0 = __shfl(y, 0, 64), y1 = __shfl(y, 1, 64);
    if (lane == 0) {
        if (sel) sel[q] = y0 < y1 ? (uint8_t)AB_PATH_DENSE : (uint8_t)AB_PATH_SPARSE;
        if (pred) {
            pred[2ull * q] = y0;
            pred[2ull * q + 1] = y1;
        }
    }
}

// One minibatch (rows idx[0..B)) forward + backward; adam != 0 also updates
// params (m, v moments) with step t.  grad receives the gradient, loss the
// batch loss.  scratch: B * (kH1 + kH2 + kO) floats.
__global__ __launch_bounds__(kStepNT) void k_mlp_step(float *__restrict__ params, float *__restrict__ m,
                                                      float *__restrict__ v, const float *__restrict__ norm,
                                                      const float *__restrict__ X, const float *__restrict__ T,
                                                      const uint32_t *__restrict__ idx, uint32_t B,
                                                      float *__restrict__ scratch, float *__restrict__ grad,
                                                      float *__restrict__ loss, int adam, float lr, float b1,
                                                      float b2, float eps, int t)
{
    __shared__ float p[kParams];
    __shared__ float nm[2 * kF];
    __shared__ float red[kStepNT];
    for (int i = threadIdx.x; i < kParams; i += kStepNT) p[i] = params[i];
    if (threadIdx.x < 2 * kF) nm[threadIdx.x] = norm[threadIdx.x];
    __syncthreads();
    float *Z1 = scratch, *Z2 = Z1 + (size_t)B * kH1, *DY = Z2 + (size_t)B * kH2;
    const float scale = 1.0f / ((float)B * kO);
    // forward
    float lsum = 0.f;
    for (uint32_t s = threadIdx.x; s < B; s += kStepNT) {
        float x[kF], y[kO];
        load_x(X, idx[s], nm, x);
        mlp_fwd1(p, x, Z1 + (size_t)s * kH1, Z2 + (size_t)s * kH2, y);
#pragma unroll
        for (int o = 0; o < kO; ++o) {
            const float e = y[o] - T[(uint64_t)idx[s] * kO + o];
            lsum += 0.5f * e * e;
            DY[(size_t)s * kO + o] = e * scale;
        }
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = 0.f;
        for (int i = 0; i < kStepNT; ++i) a += red[i];
        *loss = a * scale;
    }
    // layer 3: dW3[o][i] = sum_s DY[s][o] * relu(Z2[s][i]); db3
    for (int e = threadIdx.x; e < kO * kH2 + kO; e += kStepNT) {
        float a = 0.f;
        if (e < kO * kH2) {
            const int o = e / kH2, i = e % kH2;
            for (uint32_t s = 0; s < B; ++s) a = fmaf(DY[(size_t)s * kO + o], relu(Z2[(size_t)s * kH2 + i]), a);
            grad[oW3 + e] = a;
        } else {
            const int o = e - kO * kH2;
            for (uint32_t s = 0; s < B; ++s) a += DY[(size_t)s * kO + o];
            grad[ob3 + o] = a;
        }
    }
    __syncthreads();
    // dZ2 = (Z2 > 0) * W3^T DY, in place
    for (uint32_t s = threadIdx.x; s < B; s += kStepNT) {
        float *z = Z2 + (size_t)s * kH2;
        const float d0 = DY[(size_t)s * kO], d1 = DY[(size_t)s * kO + 1];
        for (int i = 0; i < kH2; ++i) z[i] = z[i] > 0.f ? fmaf(p[oW3 + i], d0, p[oW3 + kH2 + i] * d1) : 0.f;
    }
    __syncthreads();
    // layer 2: dW2[j][i] = sum_s dZ2[s][j] * relu(Z1[s][i]); db2
    for (int e = threadIdx.x; e < kH2 * kH1 + kH2; e += kStepNT) {
        float a = 0.f;
        if (e < kH2 * kH1) {
            const int j = e / kH1, i = e % kH1;
            for (uint32_t s = 0; s < B; ++s) a = fmaf(Z2[(size_t)s * kH2 + j], relu(Z1[(size_t)s * kH1 + i]), a);
            grad[oW2 + e] = a;
        } else {
            const int j = e - kH2 * kH1;
            for (uint32_t s = 0; s < B; ++s) a += Z2[(size_t)s * kH2 + j];
            grad[ob2 + j] = a;
        }
    }
    __syncthreads();
    // dZ1 = (Z1 > 0) * W2^T dZ2, in place
    for (uint32_t s = threadIdx.x; s < B; s += kStepNT) {
        float *z = Z1 + (size_t)s * kH1;
        const float *d2 = Z2 + (size_t)s * kH2;
        for (int i = 0; i < kH1; ++i) {
            float a = 0.f;
            for (int j = 0; j < kH2; ++j) a = fmaf(p[oW2 + j * kH1 + i], d2[j], a);
            z[i] = z[i] > 0.f ? a : 0.f;
        }
    }
    __syncthreads();
    // layer 1: dW1[j][f] = sum_s dZ1[s][j] * x[s][f]; db1
    for (int e = threadIdx.x; e < kH1 * kF + kH1; e += kStepNT) {
        float a = 0.f;
        if (e < kH1 * kF) {
            const int j = e / kF, f = e % kF;
            for (uint32_t s = 0; s < B; ++s) {
                const float xf = (X[(uint64_t)idx[s] * kF + f] - nm[f]) / nm[kF + f];
                a = fmaf(Z1[(size_t)s * kH1 + j], xf, a);
            }
            grad[oW1 + e] = a;
        } else {
            const int j = e - kH1 * kF;
            for (uint32_t s = 0; s < B; ++s) a += Z1[(size_t)s * kH1 + j];
            grad[ob1 + j] = a;
        }
    }
    if (!adam) return;
    __syncthreads();
    __threadfence_block();
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    for (int i = threadIdx.x; i < kParams; i += kStepNT) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        params[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
ab_status mlp_decide(ab_ctx *c, int slot, const float *d_feat, uint32_t nq, uint8_t *d_sel)
{
    const ab_ctx::Model &md = c->mlp[slot];
    if (!md.loaded) {
        set_error("cost model not loaded");
        return AB_INVALID_ARG;
    }
    if (!nq) return AB_OK;
    hipLaunchKernelGGL(k_mlp_decide, dim3((nq + kDecideQ - 1) / kDecideQ), dim3(64 * kDecideQ), 0, c->stream, (const float *)md.params.p,
                       (const float *)md.norm.p, d_feat, nq, d_sel, (float *)nullptr);
    return hipGetLastError() == hipSuccess ? AB_OK : AB_DEVICE_ERROR;
}

static void host_fwd(const float *p, const float *norm, const float *f, float (&y)[kO])
{
    float x[kF], h1[kH1];
    for (int i = 0; i < kF; ++i) x[i] = (f[i] - norm[i]) / norm[kF + i];
    for (int j = 0; j < kH1; ++j) {
        float s = p[ob1 + j];
        for (int i = 0; i < kF; ++i) s = fmaf(p[oW1 + j * kF + i], x[i], s);
        h1[j] = s > 0 ? s : 0;
    }
    y[0] = p[ob3];
    y[1] = p[ob3 + 1];
    for (int j = 0; j < kH2; ++j) {
        float s = p[ob2 + j];
        for (int i = 0; i < kH1; ++i) s = fmaf(p[oW2 + j * kH1 + i], h1[i], s);
        const float h = s > 0 ? s : 0;
        y[0] = fmaf(p[oW3 + j], h, y[0]);
        y[1] = fmaf(p[oW3 + kH2 + j], h, y[1]);
    }
}

int mlp_decide_host(const ab_ctx *c, int slot, const float *f)
{
    float y[kO];
    const std::vector<float> &h = c->mlp[slot].host;
    host_fwd(h.data(), h.data() + kParams, f, y);
    return y[0] < y[1] ? AB_PATH_DENSE : AB_PATH_SPARSE;
}

}  // namespace ab

using namespace ab;

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));            \
            return e_ == hipErrorOutOfMemory ? AB_OOM : AB_DEVICE_ERROR;              \
        }                                                                             \
    } while (0)
#define AB_TRY(expr)                \
    do {                            \
        ab_status s_ = (expr);      \
        if (s_ != AB_OK) return s_; \
    } while (0)

extern "C" uint32_t ab_mlp_nparams(void) { return (uint32_t)kParams; }

static bool bad_slot(int slot) { return slot < 0 || slot > 1; }

extern "C" ab_status ab_mlp_set(ab_ctx *c, int slot, const float *params, const float *mean, const float *stdv)
{
    if (!c || !params || !mean || !stdv || bad_slot(slot)) {
        set_error("ab_mlp_set: NULL argument or bad slot");
        return AB_INVALID_ARG;
    }
    for (int i = 0; i < kF; ++i)
        if (!(stdv[i] > 0) || !std::isfinite(mean[i])) {
            set_error("ab_mlp_set: standard deviations must be positive and means finite");
            return AB_INVALID_ARG;
        }
    for (int i = 0; i < kParams; ++i)
        if (!std::isfinite(params[i])) {
            set_error("ab_mlp_set: non-finite parameter");
            return AB_INVALID_ARG;
        }
    ab_ctx::Model &md = c->mlp[slot];
    md.host.assign(params, params + kParams);
    md.host.insert(md.host.end(), mean, mean + kF);
    md.host.insert(md.host.end(), stdv, stdv + kF);
    AB_TRY(md.params.ensure(kParams * 4));
    AB_TRY(md.norm.ensure(2 * kF * 4));
    HIP_TRY(hipMemcpy(md.params.p, params, kParams * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(md.norm.p, md.host.data() + kParams, 2 * kF * 4, hipMemcpyHostToDevice));
    md.loaded = true;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_mlp_get(ab_ctx *c, int slot, float *params, float *mean, float *stdv)
{
    if (!c || bad_slot(slot) || !c->mlp[slot].loaded) {
        set_error("ab_mlp_get: no model loaded in that slot");
        return AB_INVALID_ARG;
    }
    const ab_ctx::Model &md = c->mlp[slot];
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (params) HIP_TRY(hipMemcpy(params, md.params.p, kParams * 4, hipMemcpyDeviceToHost));
    if (mean) memcpy(mean, md.host.data() + kParams, kF * 4);
    if (stdv) memcpy(stdv, md.host.data() + kParams + kF, kF * 4);
    return AB_OK;
}

extern "C" ab_status ab_mlp_clear(ab_ctx *c, int slot)
{
    if (!c || bad_slot(slot)) {
        set_error("ab_mlp_clear: NULL context or bad slot");
        return AB_INVALID_ARG;
    }
    c->mlp[slot].loaded = false;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_mlp_predict(ab_ctx *c, int slot, const float *h_x, uint32_t n, float *h_y, uint8_t *h_sel)
{
    if (!c || bad_slot(slot) || !c->mlp[slot].loaded || (n && (!h_x || !h_y))) {
        set_error("ab_mlp_predict: bad argument or no model");
        return AB_INVALID_ARG;
    }
    const ab_ctx::Model &md = c->mlp[slot];
    if (!n) return AB_OK;
    DevBuf x, y, s;
    AB_TRY(x.ensure((size_t)n * kF * 4));
    AB_TRY(y.ensure((size_t)n * kO * 4));
    AB_TRY(s.ensure(n));
    ab_status st = AB_OK;
    if (hipMemcpyAsync(x.p, h_x, (size_t)n * kF * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = AB_DEVICE_ERROR;
    if (st == AB_OK) {
        hipLaunchKernelGGL(k_mlp_decide, dim3((n + kDecideQ - 1) / kDecideQ), dim3(64 * kDecideQ), 0, c->stream, (const float *)md.params.p,
                           (const float *)md.norm.p, (const float *)x.p, n, (uint8_t *)s.p, (float *)y.p);
        if (hipMemcpyAsync(h_y, y.p, (size_t)n * kO * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            (h_sel && hipMemcpyAsync(h_sel, s.p, n, hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            st = AB_DEVICE_ERROR;
    }
    if (st != AB_OK) set_error("ab_mlp_predict: device error");
    x.release();
    y.release();
    s.release();
    return st;
}

// Gradient of the loss over all n samples at the loaded parameters (the
// backward kernel in gradient-only mode), for checks against a CPU reference.
extern "C" ab_status ab_mlp_grad(ab_ctx *c, int slot, const float *h_x, const float *h_y, uint32_t n,
                                 float *h_grad, float *h_loss)
{
    if (!c || bad_slot(slot) || !c->mlp[slot].loaded || !n || n > (uint32_t)kMaxBatch || !h_x || !h_y || !h_grad) {
        set_error("ab_mlp_grad: bad argument (1..4096 samples, model loaded)");
        return AB_INVALID_ARG;
    }
    ab_ctx::Model &md = c->mlp[slot];
    DevBuf x, y, idx, scr, g, l;
    std::vector<uint32_t> id(n);
    for (uint32_t i = 0; i < n; ++i) id[i] = i;
    AB_TRY(x.ensure((size_t)n * kF * 4));
    AB_TRY(y.ensure((size_t)n * kO * 4));
    AB_TRY(idx.ensure((size_t)n * 4));
    AB_TRY(scr.ensure((size_t)n * (kH1 + kH2 + kO) * 4));
    AB_TRY(g.ensure(kParams * 4));
    AB_TRY(l.ensure(4));
    ab_status st = AB_OK;
    if (hipMemcpy(x.p, h_x, (size_t)n * kF * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.p, h_y, (size_t)n * kO * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(idx.p, id.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
        st = AB_DEVICE_ERROR;
    if (st == AB_OK) {
        hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(kStepNT), 0, c->stream, (float *)md.params.p, (float *)nullptr,
                           (float *)nullptr, (const float *)md.norm.p, (const float *)x.p, (const float *)y.p,
                           (const uint32_t *)idx.p, n, (float *)scr.p, (float *)g.p, (float *)l.p, 0, 0.f, 0.f, 0.f,
                           0.f, 0);
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(h_grad, g.p, kParams * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (h_loss && hipMemcpy(h_loss, l.p, 4, hipMemcpyDeviceToHost) != hipSuccess))
            st = AB_DEVICE_ERROR;
    }
    if (st != AB_OK) set_error("ab_mlp_grad: device error");
    for (DevBuf *b : {&x, &y, &idx, &scr, &g, &l}) b->release();
    return st;
}

// Trains from the loaded parameters (or a seeded initialization when none is
// loaded) with Adam on shuffled minibatches; mean/std are taken from the data
// when no model is loaded.  loss_curve (epochs floats) gets the mean
// minibatch loss of each epoch.
extern "C" ab_status ab_mlp_train(ab_ctx *c, int slot, const float *h_x, const float *h_y, uint32_t n,
                                  uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss_curve)
{
    if (!c || bad_slot(slot) || !h_x || !h_y || n == 0 || batch == 0 || batch > (uint32_t)kMaxBatch || !(lr > 0)) {
        set_error("ab_mlp_train: bad argument");
        return AB_INVALID_ARG;
    }
    ab_ctx::Model &md = c->mlp[slot];
    uint64_t st64 = seed ? seed : 0x9E3779B97F4A7C15ull;
    auto rnd = [&]() {
        st64 ^= st64 << 13;
        st64 ^= st64 >> 7;
        st64 ^= st64 << 17;
        return st64;
    };
    if (!md.loaded) {
        std::vector<float> p(kParams, 0.f), mean(kF, 0.f), sd(kF, 1.f);
        for (int f = 0; f < kF; ++f) {
            double s = 0, s2 = 0;
            for (uint32_t i = 0; i < n; ++i) {
                s += h_x[(size_t)i * kF + f];
                s2 += (double)h_x[(size_t)i * kF + f] * h_x[(size_t)i * kF + f];
            }
            const double mu = s / n, var = s2 / n - mu * mu;
            mean[f] = (float)mu;
            sd[f] = var > 1e-12 ? (float)std::sqrt(var) : 1.f;
        }
        // He-uniform initialization of the weights, zero biases
        auto init = [&](int off, int fan_in, int count) {
            const double lim = std::sqrt(6.0 / fan_in);
            for (int i = 0; i < count; ++i) p[off + i] = (float)((((rnd() >> 11) * (1.0 / 9007199254740992.0)) * 2 - 1) * lim);
        };
        init(oW1, kF, kH1 * kF);
        init(oW2, kH1, kH2 * kH1);
        init(oW3, kH2, kO * kH2);
        // output biases at the target means
        for (int o = 0; o < kO; ++o) {
            double s = 0;
            for (uint32_t i = 0; i < n; ++i) s += h_y[(size_t)i * kO + o];
            p[ob3 + o] = (float)(s / n);
        }
        AB_TRY(ab_mlp_set(c, slot, p.data(), mean.data(), sd.data()));
    }
    DevBuf x, y, idx, scr, g, l, m, v;
    AB_TRY(x.ensure((size_t)n * kF * 4));
    AB_TRY(y.ensure((size_t)n * kO * 4));
    AB_TRY(idx.ensure((size_t)n * 4));
    AB_TRY(scr.ensure((size_t)batch * (kH1 + kH2 + kO) * 4));
    AB_TRY(g.ensure(kParams * 4));
    const uint32_t nb = (n + batch - 1) / batch;
    AB_TRY(l.ensure((size_t)nb * 4));
    AB_TRY(m.ensure(kParams * 4));
    AB_TRY(v.ensure(kParams * 4));
    HIP_TRY(hipMemcpy(x.p, h_x, (size_t)n * kF * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(y.p, h_y, (size_t)n * kO * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m.p, 0, kParams * 4));
    HIP_TRY(hipMemset(v.p, 0, kParams * 4));
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    std::vector<float> lb(nb);
    int t = 0;
    for (uint32_t e = 0; e < epochs; ++e) {
        for (uint32_t i = n; i > 1; --i) std::swap(perm[i - 1], perm[rnd() % i]);
        HIP_TRY(hipMemcpyAsync(idx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        for (uint32_t b = 0; b < nb; ++b) {
            const uint32_t b0 = b * batch, bs = std::min(batch, n - b0);
            ++t;
            hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(kStepNT), 0, c->stream, (float *)md.params.p, (float *)m.p,
                               (float *)v.p, (const float *)md.norm.p, (const float *)x.p, (const float *)y.p,
                               (const uint32_t *)idx.p + b0, bs, (float *)scr.p, (float *)g.p, (float *)l.p + b, 1, lr,
                               0.9f, 0.999f, 1e-8f, t);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(lb.data(), l.p, (size_t)nb * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // perm is rewritten next epoch
        if (loss_curve) {
            double s = 0;
            for (uint32_t b = 0; b < nb; ++b) s += lb[b];
            loss_curve[e] = (float)(s / nb);
        }
    }
    // refresh the host copy used by single-query dispatch
    HIP_TRY(hipMemcpy(md.host.data(), md.params.p, kParams * 4, hipMemcpyDeviceToHost));
    c->cfg_gen++;
    for (DevBuf *bf : {&x, &y, &idx, &scr, &g, &l, &m, &v}) bf->release();
    return AB_OK;
}
