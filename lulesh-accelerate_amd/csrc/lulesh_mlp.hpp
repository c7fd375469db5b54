// lulesh_mlp.hpp -- the 12-64-64-2 network of the learned-Q closure on the
// matrix cores of gfx950.
//
// One wavefront evaluates 64 samples (lane l = sample l).  Every layer is a
// bf16 x bf16 -> fp32 product on v_mfma_f32_16x16x16_bf16: 16 x 16 output
// tiles, K = 16 per instruction.  The fragment layout of that instruction
// (64 lanes, g = lane / 16, r = lane % 16):
//   A (M x K):  lane holds A[r][4g .. 4g+3]
//   B (K x N):  lane holds B[4g .. 4g+3][r]
//   D (M x N):  lane holds D[4g + i][r], i = 0..3
// Weights are kept in fp32 (the trainer's master copy) and packed into bf16
// B fragments, so that every lane loads its fragment with one 8-byte read.
// Activations go through the wave's own LDS slice (bf16, rows padded to
// kHPad values against bank conflicts).
//
// Parameter vector (fp32, kMlpParams values), in this order:
//   W1[12][64] (input-major), b1[64], W2[64][64], b2[64], W3[64][2], b3[2].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lulesh {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHidden;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpOut;
constexpr int kMlpParams = kOffB3 + kMlpOut;  // 5122

typedef short mfma_v4s __attribute__((ext_vector_type(4)));
typedef float mfma_v4f __attribute__((ext_vector_type(4)));

// bf16 B fragments of the packed network, 8 bytes per lane each
struct MlpFrags {
    mfma_v4s w1[4][64];      // [n tile][lane]: W1 padded to 16 x 64
    mfma_v4s w2[4][4][64];   // [k tile][n tile][lane]
    mfma_v4s w2t[4][4][64];  // W2 transposed, for the backward pass
    mfma_v4s w3[4][64];      // [k tile][lane]: W3 padded to 64 x 16
    float b1[kMlpHidden];
    float b2[kMlpHidden];
    float b3[kMlpOut];
    float w3f[kMlpHidden][kMlpOut];  // fp32 W3 for the backward pass
    // input standardisation: z = (x - mean) * inv_std
    float in_mean[kMlpIn];
    float in_inv_std[kMlpIn];
};

constexpr int kHPad = 68;  // bf16 values per activation row in LDS

// LDS of one wave
struct MlpWaveLds {
    uint16_t x[64][16];
    uint16_t h1[64][kHPad];
    uint16_t h2[64][kHPad];
    float y[64][kMlpOut];
};

// fp32 -> bf16, round to nearest even (finite inputs)
__host__ __device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__host__ __device__ __forceinline__ float from_bf16(uint16_t h)
{
    const uint32_t u = uint32_t(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ mfma_v4f mfma16(mfma_v4s a, mfma_v4s b, mfma_v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ mfma_v4s pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d)
{
    mfma_v4s v;
    v[0] = short(a);
    v[1] = short(b);
    v[2] = short(c);
    v[3] = short(d);
    return v;
}

// A fragment from a row-major bf16 tile in LDS: rows r0 + (lane % 16),
// columns c0 + 4g .. c0 + 4g + 3
template <int LD>
__device__ __forceinline__ mfma_v4s load_a(const uint16_t (*m)[LD], int r0, int c0, int lane)
{
    const uint16_t *p = &m[r0 + (lane & 15)][c0 + 4 * (lane >> 4)];
    return pack4(p[0], p[1], p[2], p[3]);
}

// A fragment of the transpose of a row-major tile: A[i][k] = m[k][i]
template <int LD>
__device__ __forceinline__ mfma_v4s load_at(const uint16_t (*m)[LD], int i0, int k0, int lane)
{
    const int i = i0 + (lane & 15);
    const int k = k0 + 4 * (lane >> 4);
    return pack4(m[k][i], m[k + 1][i], m[k + 2][i], m[k + 3][i]);
}

// B fragment from a row-major tile: B[k][n] = m[k][n]
template <int LD>
__device__ __forceinline__ mfma_v4s load_b(const uint16_t (*m)[LD], int k0, int n0, int lane)
{
    const int n = n0 + (lane & 15);
    const int k = k0 + 4 * (lane >> 4);
    return pack4(m[k][n], m[k + 1][n], m[k + 2][n], m[k + 3][n]);
}

// Forward pass of the wave's 64 samples.  Each lane has written its sample's
// standardised inputs (bf16, zero-padded to 16) to L.x[lane] beforehand;
// afterwards L.h1 / L.h2 hold the post-ReLU activations (bf16) and L.y the
// outputs (fp32).  All 64 lanes must call it.
__device__ __forceinline__ void mlp_forward_wave(const MlpFrags &F, MlpWaveLds &L, int lane)
{
    const int col = lane & 15;
    const int rg = 4 * (lane >> 4);
    // layer 1: x[64 x 16] . W1[16 x 64]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const mfma_v4s a = load_a<16>(L.x, 16 * mt, 0, lane);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            mfma_v4f acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(a, F.w1[nt][lane], acc);
            const int n = 16 * nt + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float h = acc[i] + F.b1[n];
                L.h1[16 * mt + rg + i][n] = to_bf16(h > 0.f ? h : 0.f);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    // layer 2: h1[64 x 64] . W2[64 x 64]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mfma_v4f acc[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            acc[nt] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            const mfma_v4s a = load_a<kHPad>(L.h1, 16 * mt, 16 * kt, lane);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                acc[nt] = mfma16(a, F.w2[kt][nt][lane], acc[nt]);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int n = 16 * nt + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float h = acc[nt][i] + F.b2[n];
                L.h2[16 * mt + rg + i][n] = to_bf16(h > 0.f ? h : 0.f);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    // layer 3: h2[64 x 64] . W3[64 x 16 (2 used)]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        mfma_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            acc = mfma16(load_a<kHPad>(L.h2, 16 * mt, 16 * kt, lane), F.w3[kt][lane], acc);
        }
        if (col < kMlpOut) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                L.y[16 * mt + rg + i][col] = acc[i] + F.b3[col];
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

}  // namespace lulesh

// the packed fragments of a network (lulesh_mlp.hip)
struct lulesh_mlp;
const lulesh::MlpFrags *lulesh_mlp_device_frags(const lulesh_mlp *n);
