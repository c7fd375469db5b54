// lulesh_mlp.hip -- training and evaluation of the learned-Q network
// (the lulesh_mlp_* entry points of include/lulesh_mi355x.h).
//
// Kernels:
//   k_mlp_pack       fp32 master parameters -> bf16 MFMA fragments
//   k_mlp_eval       forward pass over an array of samples
//   k_mlp_grad       forward + backward pass of 64 samples per wave, loss
//                    0.5 * mean over samples of sum over outputs of
//                    w_o (y - t)^2; each wave writes its partial gradient
//   k_mlp_adam       sums the partial gradients in wave order (the result
//                    does not depend on the launch) and takes an Adam step
// The backward pass runs its products on the matrix cores as well: dW3 =
// H2^T dY, dW2 = H1^T dH2 and dW1 = X^T dH1 contract over the wave's 64
// samples, dH1 = dH2 W2^T over the hidden units; dH2 = dY W3^T (K = 2) is
// formed per lane.  Bias gradients are summed in fp32 from the fp32 deltas.
#include "lulesh_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "lulesh_mlp.hpp"

using namespace lulesh;

namespace {

constexpr int kWaves = 4;  // waves per block
constexpr int kBlockThreads = 64 * kWaves;

// The partial gradients of lulesh_mlp_gradient_noalloc: one row of
// kMlpParams + 1 floats per wave of 64 samples, for up to
// LULESH_MLP_NOALLOC_MAX samples (a whole number of blocks); 5.0 MiB.  Each
// GPU the library runs on holds its own copy until the process ends.  In
// host code the name stands for no device address: host code uses the
// per-device pointer that noalloc_scratch() resolves.
constexpr int kNoallocWaves = LULESH_MLP_NOALLOC_MAX / 64;
static_assert(LULESH_MLP_NOALLOC_MAX % kBlockThreads == 0,
              "the no-allocation buffer holds whole blocks");
static __device__ __align__(256) float scratch[size_t(kNoallocWaves) * (kMlpParams + 1)];

__global__ void k_mlp_pack(const float *__restrict__ p, const float *__restrict__ norm,
                           MlpFrags *F)
{
    const int t = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    // W1 fragments: B[k][n] = W1[k][n] for k < 12, 0 for the padding rows
    if (t < 4 * 64) {
        const int nt = t / 64, lane = t % 64;
        const int n = 16 * nt + (lane & 15), k0 = 4 * (lane >> 4);
        uint16_t v[4];
        for (int i = 0; i < 4; ++i) {
            const int k = k0 + i;
            v[i] = (k < kMlpIn) ? to_bf16(p[kOffW1 + k * kMlpHidden + n]) : 0;
        }
        F->w1[nt][lane] = pack4(v[0], v[1], v[2], v[3]);
    }
    if (t < 16 * 64) {
        const int kt = t / 256, nt = (t / 64) % 4, lane = t % 64;
        const int n = 16 * nt + (lane & 15), k0 = 16 * kt + 4 * (lane >> 4);
        uint16_t v[4], w[4];
        for (int i = 0; i < 4; ++i) {
            v[i] = to_bf16(p[kOffW2 + (k0 + i) * kMlpHidden + n]);
            w[i] = to_bf16(p[kOffW2 + n * kMlpHidden + (k0 + i)]);  // W2^T[k][n]
        }
        F->w2[kt][nt][lane] = pack4(v[0], v[1], v[2], v[3]);
        F->w2t[kt][nt][lane] = pack4(w[0], w[1], w[2], w[3]);
    }
    if (t < 4 * 64) {
        const int kt = t / 64, lane = t % 64;
        const int n = lane & 15, k0 = 16 * kt + 4 * (lane >> 4);
        uint16_t v[4];
        for (int i = 0; i < 4; ++i) {
            v[i] = (n < kMlpOut) ? to_bf16(p[kOffW3 + (k0 + i) * kMlpOut + n]) : 0;
        }
        F->w3[kt][lane] = pack4(v[0], v[1], v[2], v[3]);
    }
    if (t < kMlpHidden) {
        F->b1[t] = p[kOffB1 + t];
        F->b2[t] = p[kOffB2 + t];
        for (int o = 0; o < kMlpOut; ++o) {
            F->w3f[t][o] = p[kOffW3 + t * kMlpOut + o];
        }
    }
    if (t < kMlpOut) {
        F->b3[t] = p[kOffB3 + t];
    }
    if (t < kMlpIn) {
        F->in_mean[t] = norm[t];
        F->in_inv_std[t] = norm[kMlpIn + t];
    }
}

// standardised, bf16-rounded inputs of sample `s` into the wave's x row
__device__ __forceinline__ void stage_inputs(const MlpFrags &F, MlpWaveLds &L, int lane,
                                             const float *__restrict__ x, int s, bool valid)
{
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        float z = 0.f;
        if (valid && i < kMlpIn) {
            z = (x[size_t(s) * kMlpIn + i] - F.in_mean[i]) * F.in_inv_std[i];
        }
        L.x[lane][i] = to_bf16(z);
    }
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(kBlockThreads) void k_mlp_eval(const MlpFrags *__restrict__ Fp,
                                                            const float *__restrict__ x,
                                                            int n, float *__restrict__ y)
{
    __shared__ MlpWaveLds lds[kWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    MlpWaveLds &L = lds[wave];
    const MlpFrags &F = *Fp;
    const int s = (int(blockIdx.x) * kWaves + wave) * 64 + lane;
    stage_inputs(F, L, lane, x, s, s < n);
    mlp_forward_wave(F, L, lane);
    if (s < n) {
        for (int o = 0; o < kMlpOut; ++o) {
            y[size_t(s) * kMlpOut + o] = L.y[lane][o];
        }
    }
}

// One wave = 64 samples idx[base + lane] (idx == nullptr: base + lane);
// samples at or beyond n are masked.  part: [waves][kMlpParams + 1], the
// last entry the wave's loss.
__global__ __launch_bounds__(kBlockThreads) void k_mlp_grad(
    const MlpFrags *__restrict__ Fp, const float *__restrict__ x,
    const float *__restrict__ t, const int *__restrict__ idx, int n, int batch_base,
    int batch_size, float w0, float w1, float *__restrict__ part)
{
    __shared__ MlpWaveLds lds[kWaves];
    __shared__ float dh_f32[kWaves][64][kMlpHidden + 1];  // fp32 deltas for the biases
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gwave = int(blockIdx.x) * kWaves + wave;
    MlpWaveLds &L = lds[wave];
    const MlpFrags &F = *Fp;
    const int j = gwave * 64 + lane;  // position in the batch
    const bool valid = j < batch_size && batch_base + j < n;
    const int s = valid ? (idx ? idx[batch_base + j] : batch_base + j) : 0;
    stage_inputs(F, L, lane, x, s, valid);
    mlp_forward_wave(F, L, lane);

    // loss and output deltas, scaled by 1 / batch_size
    const float inv_b = 1.0f / float(batch_size);
    float dy[kMlpOut];
    float loss = 0.f;
    {
        const float wo[2] = {w0, w1};
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) {
            const float r = valid ? (L.y[lane][o] - t[size_t(s) * kMlpOut + o]) : 0.f;
            loss += 0.5f * wo[o] * r * r;
            dy[o] = wo[o] * r * inv_b;
        }
    }
    float *out = part + size_t(gwave) * (kMlpParams + 1);
    const int col = lane & 15;
    const int rg = 4 * (lane >> 4);

    // dY as a bf16 tile in the x slice (64 x 16, columns 0..1 used): the
    // inputs are re-read below only through their own copy
    uint16_t xrow[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        xrow[i] = L.x[lane][i];
    }
    __builtin_amdgcn_wave_barrier();

    // dW3 = H2^T dY and db3 = sum dY
    {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            L.x[lane][i] = (i < kMlpOut) ? to_bf16(dy[i]) : 0;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            mfma_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                acc = mfma16(load_at<kHPad>(L.h2, 16 * mt, 16 * kt, lane),
                             load_b<16>(L.x, 16 * kt, 0, lane), acc);
            }
            if (col < kMlpOut) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    out[kOffW3 + (16 * mt + rg + i) * kMlpOut + col] = acc[i];
                }
            }
        }
#pragma unroll
        for (int o = 0; o < kMlpOut; ++o) {
            float v = dy[o];
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                v += __shfl_xor(v, off, 64);
            }
            if (lane == 0) {
                out[kOffB3 + o] = v;
            }
        }
        float lv = loss;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lv += __shfl_xor(lv, off, 64);
        }
        if (lane == 0) {
            out[kMlpParams] = lv;
        }
    }

    // dH2 = dY W3^T masked by ReLU'(H2), per lane (row = sample), in place
    // of H2; fp32 copy for db2
#pragma unroll 4
    for (int c = 0; c < kMlpHidden; ++c) {
        const float g = dy[0] * F.w3f[c][0] + dy[1] * F.w3f[c][1];
        const float d = (from_bf16(L.h2[lane][c]) > 0.f) ? g : 0.f;
        dh_f32[wave][lane][c] = d;
        L.h2[lane][c] = to_bf16(d);
    }
    __builtin_amdgcn_wave_barrier();
    {
        float v = 0.f;
        for (int r = 0; r < 64; ++r) {
            v += dh_f32[wave][r][lane];
        }
        out[kOffB2 + lane] = v;
    }
    // dW2 = H1^T dH2
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mfma_v4f acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            acc[nt] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            const mfma_v4s a = load_at<kHPad>(L.h1, 16 * mt, 16 * kt, lane);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                acc[nt] = mfma16(a, load_b<kHPad>(L.h2, 16 * kt, 16 * nt, lane), acc[nt]);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                out[kOffW2 + (16 * mt + rg + i) * kMlpHidden + 16 * nt + col] = acc[nt][i];
            }
        }
    }
    // dH1 = dH2 W2^T masked by ReLU'(H1), written over H1 once every tile
    // is formed
    {
        mfma_v4f acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                acc[mt][nt] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const mfma_v4s a = load_a<kHPad>(L.h2, 16 * mt, 16 * kt, lane);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    acc[mt][nt] = mfma16(a, F.w2t[kt][nt][lane], acc[mt][nt]);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = 16 * mt + rg + i, c = 16 * nt + col;
                    const float d = (from_bf16(L.h1[r][c]) > 0.f) ? acc[mt][nt][i] : 0.f;
                    dh_f32[wave][r][c] = d;
                    L.h1[r][c] = to_bf16(d);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    {
        float v = 0.f;
        for (int r = 0; r < 64; ++r) {
            v += dh_f32[wave][r][lane];
        }
        out[kOffB1 + lane] = v;
    }
    // dW1 = X^T dH1 (X restored from the lane's copy)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        L.x[lane][i] = xrow[i];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        mfma_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            acc = mfma16(load_at<16>(L.x, 0, 16 * kt, lane),
                         load_b<kHPad>(L.h1, 16 * kt, 16 * nt, lane), acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = rg + i;
            if (k < kMlpIn) {
                out[kOffW1 + k * kMlpHidden + 16 * nt + col] = acc[i];
            }
        }
    }
}

// grad[p] = sum over waves of part[w][p] (fixed order); Adam step on params
// when lr > 0.  loss_out[0] = summed loss / batch.
__global__ void k_mlp_adam(const float *__restrict__ part, int nwaves, float *grad,
                           float *params, float *m, float *v, int step, float lr,
                           float beta1, float beta2, float eps, int batch,
                           float *loss_out)
{
    const int p = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (p > kMlpParams) {
        return;
    }
    float g = 0.f;
    for (int w = 0; w < nwaves; ++w) {
        g += part[size_t(w) * (kMlpParams + 1) + p];
    }
    if (p == kMlpParams) {
        loss_out[0] = g / float(batch);
        return;
    }
    grad[p] = g;
    if (lr > 0.f) {
        const float mm = beta1 * m[p] + (1.f - beta1) * g;
        const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        const float mh = mm / (1.f - powf(beta1, float(step)));
        const float vh = vv / (1.f - powf(beta2, float(step)));
        params[p] -= lr * mh / (sqrtf(vh) + eps);
    }
}

}  // namespace

struct lulesh_mlp {
    float *d_params = nullptr;  // fp32 master weights
    float *d_norm = nullptr;    // in_mean[12], in_inv_std[12]
    float *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    float *d_loss = nullptr;
    MlpFrags *d_frags = nullptr;
    float *d_part = nullptr;
    size_t part_waves = 0;
    int step = 0;
    int device = -1;  // the device current at lulesh_mlp_create
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

int mfail(lulesh_mlp *n, int code, const std::string &msg)
{
    if (n) n->err = msg;
    return code;
}

#define MHIP_TRY(n, expr)                                                    \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            return mfail((n), LULESH_HIP_ERROR,                              \
                         std::string(#expr ": ") + hipGetErrorString(e_));   \
        }                                                                    \
    } while (0)

int repack(lulesh_mlp *n)
{
    hipLaunchKernelGGL(k_mlp_pack, dim3(4), dim3(256), 0, n->stream, n->d_params,
                       n->d_norm, n->d_frags);
    MHIP_TRY(n, hipGetLastError());
    return LULESH_OK;
}

int ensure_part(lulesh_mlp *n, size_t waves)
{
    if (waves <= n->part_waves) {
        return LULESH_OK;
    }
    MHIP_TRY(n, hipStreamSynchronize(n->stream));
    if (n->d_part) MHIP_TRY(n, hipFree(n->d_part));
    n->d_part = nullptr;
    MHIP_TRY(n, hipMalloc(&n->d_part, waves * (kMlpParams + 1) * sizeof(float)));
    n->part_waves = waves;
    return LULESH_OK;
}

// device copies of host arrays, freed by the caller
template <typename T>
int upload(lulesh_mlp *n, T **dst, const T *src, size_t count)
{
    MHIP_TRY(n, hipMalloc(dst, count * sizeof(T)));
    MHIP_TRY(n, hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice,
                               n->stream));
    return LULESH_OK;
}

// one batch: partial gradients, then their sum (and an Adam step if lr > 0)
int run_batch(lulesh_mlp *n, const float *dx, const float *dt, const int *didx, int count,
              int base, int batch, const float *w, float lr)
{
    const int waves = (batch + 63) / 64;
    const int blocks = (waves + kWaves - 1) / kWaves;
    int rc = ensure_part(n, size_t(blocks) * kWaves);
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kBlockThreads), 0, n->stream,
                       n->d_frags, dx, dt, didx, count, base, batch, w[0], w[1],
                       n->d_part);
    MHIP_TRY(n, hipGetLastError());
    if (lr > 0.f) {
        ++n->step;
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0,
                       n->stream, n->d_part, blocks * kWaves, n->d_grad, n->d_params,
                       n->d_m, n->d_v, n->step, lr, 0.9f, 0.999f, 1e-8f, batch,
                       n->d_loss);
    MHIP_TRY(n, hipGetLastError());
    if (lr > 0.f) {
        return repack(n);
    }
    return LULESH_OK;
}

// The device address of `scratch` on the current device, resolved once per
// device (hipGetSymbolAddress with that device current) and kept; nullptr
// and the HIP error when it cannot be resolved.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

float *noalloc_scratch(int device, hipError_t *err)
{
    *err = hipSuccess;
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        *err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) {
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<float *>(ptr);
    }
    return g_scratch_ptr[device];
}

}  // namespace

const MlpFrags *lulesh_mlp_device_frags(const lulesh_mlp *n) { return n ? n->d_frags : nullptr; }

extern "C" {

int lulesh_mlp_param_count(void) { return kMlpParams; }

int lulesh_mlp_create(const float *params, const float *in_mean, const float *in_std,
                      lulesh_mlp **out)
{
    if (!params || !in_mean || !in_std || !out) {
        return LULESH_BAD_ARG;
    }
    *out = nullptr;
    for (int i = 0; i < kMlpIn; ++i) {
        if (!(in_std[i] > 0.f) || !std::isfinite(in_mean[i])) {
            return LULESH_BAD_ARG;
        }
    }
    for (int i = 0; i < kMlpParams; ++i) {
        if (!std::isfinite(params[i])) {
            return LULESH_BAD_ARG;
        }
    }
    lulesh_mlp *n = new (std::nothrow) lulesh_mlp();
    if (!n) {
        return LULESH_BAD_ARG;
    }
    float norm[2 * kMlpIn];
    for (int i = 0; i < kMlpIn; ++i) {
        norm[i] = in_mean[i];
        norm[kMlpIn + i] = 1.0f / in_std[i];
    }
    const size_t pb = kMlpParams * sizeof(float);
    bool ok = hipGetDevice(&n->device) == hipSuccess &&
              hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&n->d_params, pb) == hipSuccess &&
              hipMalloc(&n->d_m, pb) == hipSuccess && hipMalloc(&n->d_v, pb) == hipSuccess &&
              hipMalloc(&n->d_grad, pb) == hipSuccess &&
              hipMalloc(&n->d_loss, sizeof(float)) == hipSuccess &&
              hipMalloc(&n->d_norm, sizeof(norm)) == hipSuccess &&
              hipMalloc(&n->d_frags, sizeof(MlpFrags)) == hipSuccess;
    ok = ok && hipMemcpy(n->d_params, params, pb, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(n->d_norm, norm, sizeof(norm), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemsetAsync(n->d_m, 0, pb, n->stream) == hipSuccess &&
         hipMemsetAsync(n->d_v, 0, pb, n->stream) == hipSuccess;
    if (!ok || repack(n) != LULESH_OK || hipStreamSynchronize(n->stream) != hipSuccess) {
        lulesh_mlp_destroy(n);
        return LULESH_HIP_ERROR;
    }
    *out = n;
    return LULESH_OK;
}

void lulesh_mlp_destroy(lulesh_mlp *n)
{
    if (!n) {
        return;
    }
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    float *bufs[] = {n->d_params, n->d_norm, n->d_m, n->d_v, n->d_grad, n->d_loss, n->d_part};
    for (float *b : bufs) {
        if (b) (void)hipFree(b);
    }
    if (n->d_frags) (void)hipFree(n->d_frags);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

int lulesh_mlp_get_params(lulesh_mlp *n, float *params)
{
    if (!n || !params) {
        return LULESH_BAD_ARG;
    }
    MHIP_TRY(n, hipStreamSynchronize(n->stream));
    MHIP_TRY(n, hipMemcpy(params, n->d_params, kMlpParams * sizeof(float),
                          hipMemcpyDeviceToHost));
    return LULESH_OK;
}

int lulesh_mlp_eval(lulesh_mlp *n, const float *x, int count, float *y)
{
    if (!n || !x || !y || count < 0) {
        return LULESH_BAD_ARG;
    }
    if (count == 0) {
        return LULESH_OK;
    }
    float *dx = nullptr, *dy = nullptr;
    int rc = upload(n, &dx, x, size_t(count) * kMlpIn);
    if (!rc && hipMalloc(&dy, size_t(count) * kMlpOut * sizeof(float)) != hipSuccess) {
        rc = mfail(n, LULESH_HIP_ERROR, "hipMalloc failed");
    }
    if (!rc) {
        const int blocks = (count + kBlockThreads - 1) / kBlockThreads;
        hipLaunchKernelGGL(k_mlp_eval, dim3(blocks), dim3(kBlockThreads), 0, n->stream,
                           n->d_frags, dx, count, dy);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(y, dy, size_t(count) * kMlpOut * sizeof(float),
                           hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
            hipStreamSynchronize(n->stream) != hipSuccess) {
            rc = mfail(n, LULESH_HIP_ERROR, "mlp eval failed");
        }
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return rc;
}

int lulesh_mlp_gradient(lulesh_mlp *n, const float *x, const float *t, int count,
                        const float *out_weight, float *grad, double *loss)
{
    if (!n || !x || !t || !grad || count < 1) {
        return LULESH_BAD_ARG;
    }
    const float w[2] = {out_weight ? out_weight[0] : 1.f, out_weight ? out_weight[1] : 1.f};
    float *dx = nullptr, *dt = nullptr;
    int rc = upload(n, &dx, x, size_t(count) * kMlpIn);
    if (!rc) rc = upload(n, &dt, t, size_t(count) * kMlpOut);
    if (!rc) rc = run_batch(n, dx, dt, nullptr, count, 0, count, w, 0.f);
    float l = 0.f;
    if (!rc && (hipMemcpyAsync(grad, n->d_grad, kMlpParams * sizeof(float),
                               hipMemcpyDeviceToHost, n->stream) != hipSuccess ||
                hipMemcpyAsync(&l, n->d_loss, sizeof(float), hipMemcpyDeviceToHost,
                               n->stream) != hipSuccess ||
                hipStreamSynchronize(n->stream) != hipSuccess)) {
        rc = mfail(n, LULESH_HIP_ERROR, "mlp gradient failed");
    }
    if (loss) *loss = l;
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    return rc;
}

int lulesh_mlp_gradient_noalloc(lulesh_mlp *n, const float *x, const float *t, int count,
                                const float *out_weight, float *grad, float *loss,
                                void *stream)
{
    if (!n || !x || !t || !grad || !loss || count < 1) {
        return LULESH_BAD_ARG;
    }
    if (count > LULESH_MLP_NOALLOC_MAX) {
        return mfail(n, LULESH_BAD_ARG,
                     "lulesh_mlp_gradient_noalloc: " + std::to_string(count) +
                         " samples exceed the maximum of " +
                         std::to_string(LULESH_MLP_NOALLOC_MAX) + " per call");
    }
    int device = -1;
    MHIP_TRY(n, hipGetDevice(&device));
    if (device != n->device) {
        return mfail(n, LULESH_BAD_ARG,
                     "lulesh_mlp_gradient_noalloc: the current device is not the "
                     "network's");
    }
    hipError_t e = hipSuccess;
    float *part = noalloc_scratch(device, &e);
    if (!part) {
        return mfail(n, LULESH_HIP_ERROR,
                     std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    const float w0 = out_weight ? out_weight[0] : 1.f;
    const float w1 = out_weight ? out_weight[1] : 1.f;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // the kernels of run_batch without an Adam step: every wave of the grid
    // writes its whole row of partials (masked samples add zeros), and the
    // sum reads exactly those rows, so no row has to be cleared first
    const int waves = (count + 63) / 64;
    const int blocks = (waves + kWaves - 1) / kWaves;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(kBlockThreads), 0, st, n->d_frags, x,
                       t, nullptr, count, 0, count, w0, w1, part);
    MHIP_TRY(n, hipGetLastError());
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, st,
                       part, blocks * kWaves, grad, n->d_params, n->d_m, n->d_v, n->step,
                       0.f, 0.9f, 0.999f, 1e-8f, count, loss);
    MHIP_TRY(n, hipGetLastError());
    return LULESH_OK;
}

int lulesh_mlp_train(lulesh_mlp *n, const float *x, const float *t, int count,
                     const float *out_weight, int epochs, int batch, float lr,
                     unsigned long long seed, double *final_loss)
{
    if (!n || !x || !t || count < 1 || epochs < 1 || batch < 64 || !(lr > 0.f)) {
        return LULESH_BAD_ARG;
    }
    const float w[2] = {out_weight ? out_weight[0] : 1.f, out_weight ? out_weight[1] : 1.f};
    float *dx = nullptr, *dt = nullptr;
    int *didx = nullptr;
    int rc = upload(n, &dx, x, size_t(count) * kMlpIn);
    if (!rc) rc = upload(n, &dt, t, size_t(count) * kMlpOut);
    std::vector<int> order(static_cast<size_t>(count));
    for (int i = 0; i < count; ++i) order[size_t(i)] = i;
    if (!rc && hipMalloc(&didx, size_t(count) * sizeof(int)) != hipSuccess) {
        rc = mfail(n, LULESH_HIP_ERROR, "hipMalloc failed");
    }
    uint64_t state = seed;
    auto next = [&]() {  // splitmix64
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    double epoch_loss = 0.0;
    for (int e = 0; e < epochs && !rc; ++e) {
        // Fisher-Yates shuffle of the sample order
        for (int i = count - 1; i > 0; --i) {
            const int j = int(next() % uint64_t(i + 1));
            std::swap(order[size_t(i)], order[size_t(j)]);
        }
        if (hipMemcpyAsync(didx, order.data(), order.size() * sizeof(int),
                           hipMemcpyHostToDevice, n->stream) != hipSuccess) {
            rc = mfail(n, LULESH_HIP_ERROR, "upload failed");
            break;
        }
        std::vector<float> losses;
        for (int base = 0; base + batch <= count || (base == 0 && count < batch);
             base += batch) {
            const int b = std::min(batch, count - base);
            if ((rc = run_batch(n, dx, dt, didx, count, base, b, w, lr))) break;
            float l = 0.f;
            if (hipMemcpyAsync(&l, n->d_loss, sizeof(float), hipMemcpyDeviceToHost,
                               n->stream) != hipSuccess) {
                rc = mfail(n, LULESH_HIP_ERROR, "loss copy failed");
                break;
            }
            MHIP_TRY(n, hipStreamSynchronize(n->stream));
            losses.push_back(l);
        }
        double sum = 0.0;
        for (float l : losses) sum += l;
        epoch_loss = losses.empty() ? 0.0 : sum / double(losses.size());
    }
    if (final_loss) *final_loss = epoch_loss;
    (void)hipStreamSynchronize(n->stream);
    if (dx) (void)hipFree(dx);
    if (dt) (void)hipFree(dt);
    if (didx) (void)hipFree(didx);
    return rc;
}

const char *lulesh_mlp_last_error(const lulesh_mlp *n)
{
    return n ? n->err.c_str() : "null network";
}

}  // extern "C"
