// mlp.hpp -- the learned Gaussian proposal's network on CDNA4 matrix cores.
//
// A 6 -> 64 -> 64 -> 8 tanh MLP maps (previous state, observation) features
// to the proposal's mean offsets and log-scale offsets (the reference has no
// learned proposal; the model it corrects is the PF of
// user_guide/cpp/pf_seq.cpp:70-81, :111-122).
//
// Layout. Every layer is evaluated transposed, Z^T = W H^T: A = a weight
// tile (output features x input features), B = an activation tile (input
// features x 16 samples), accumulator C/D = output features x samples. The
// first layer (6 inputs) uses v_mfma_f32_16x16x16_bf16, where lane l holds
// A[i][k] for i = l % 16, k = 4 (l / 16) + v and B[k][j] for j = l % 16,
// k = 4 (l / 16) + v (v = 0..3); the 64-wide layers use gfx950's
// v_mfma_f32_16x16x32_bf16, where lane l holds A[i][k] and B[k][j] for
// k = 8 (l / 16) + v (v = 0..7). Both write D[i][j] for j = l % 16,
// i = 4 (l / 16) + v (v = 0..3). A contraction index may be numbered in any
// order shared by A and B, so the K = 32 slot 8 g + v of half kt of a
// 64-wide input is hidden unit kslot(kt, g, v): rows 4 g + v of its first
// 16-row tile for v < 4, of its second tile for v >= 4 -- exactly the rows
// lane group g holds of the previous layer's two D tiles. Activations thus
// never leave registers in the forward pass. Weight tiles are pre-swizzled
// into per-lane fragments (Frags) by a pack kernel after every optimiser
// step. Accumulation is FP32; inputs, weights and activations are bfloat16
// (round to nearest even, the hardware conversion); tanh comes from the
// hardware exp2 and reciprocal (tanh_fast).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsmch {
namespace mlp {

constexpr int kIn = 6;
constexpr int kH = 64;
constexpr int kOut = 8;
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kH * kIn;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kOut * kH;
constexpr int kParams = oB3 + kOut;   // 5128
constexpr int kGradLen = kParams + 1; // + loss

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4h __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Hidden unit of K = 32 slot (g, v) of half kt of a 64-wide input.
__host__ __device__ constexpr int kslot(int kt, int g, int v)
{
    return 32 * kt + (v < 4 ? 4 * g + v : 16 + 4 * g + (v - 4));
}

// What the forward pass reads (the learned move stages this in LDS).
struct FwdFrags {
    bf16x4 f1[4][64];      // W1 as A: row tile rt (K = inputs, padded to 16)
    bf16x8 f2[4][2][64];   // W2 as A: [rt][kt], K slots by kslot
    bf16x8 f3[2][64];      // W3 as A: [kt] (outputs padded to 16 rows)
    float b1[kH];
    float b2[kH];
    float b3[16];          // padded with zeros
    float shift[8];        // input normalisation (f - shift) * scale
    float scale[8];
};

struct Frags {
    FwdFrags fwd;
    bf16x4 g3[4][64];      // W3^T as A: [rt] (K = outputs padded to 16)
    bf16x4 g2[4][4][64];   // W2^T as A: [rt][kt]
};

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u)
        return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b)
{
    return __builtin_bit_cast(float, static_cast<uint32_t>(b) << 16);
}

// Round to nearest even, by the hardware conversion (the same results as
// f2bf for every non-NaN input).
__device__ inline bf16x4 pack4(f32x4 v)
{
    const bf16x4h h{static_cast<__bf16>(v[0]), static_cast<__bf16>(v[1]),
        static_cast<__bf16>(v[2]), static_cast<__bf16>(v[3])};
    return __builtin_bit_cast(bf16x4, h);
}

// The K = 32 operand of two 4-row fragments (slots v < 4, then v >= 4).
__device__ inline bf16x8 cat8(bf16x4 lo, bf16x4 hi)
{
    const bf16x4h a = __builtin_bit_cast(bf16x4h, lo);
    const bf16x4h b = __builtin_bit_cast(bf16x4h, hi);
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline f32x4 load4(const float *p)
{
    return f32x4{p[0], p[1], p[2], p[3]};
}

// tanh x = 1 - 2 / (2^(2 x log2 e) + 1) with v_exp_f32 and v_rcp_f32. With
// both within about one ulp, the absolute error stays below 1e-6 for every
// x (at most about 0.5 of exp's relative error, plus 2 ulp of the
// reciprocal, plus the final rounding); large |x| gives exactly +-1, NaN
// stays NaN. The activations are rounded to bfloat16 (relative 2^-9) right
// after, so this error is below their resolution wherever |tanh| > 2e-4.
__device__ inline float tanh_fast(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * 2.88539008177792681f);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

__device__ inline f32x4 tanh4(f32x4 v)
{
    return f32x4{tanh_fast(v[0]), tanh_fast(v[1]), tanh_fast(v[2]), tanh_fast(v[3])};
}

// Forward pass of one 16-sample column tile. x: this lane's input
// fragment (features 4 (l / 16) + v of sample l % 16, already normalised
// and rounded). Returns the FP32 activations (D layout) and the outputs:
// lanes 0-15 hold o[0..3], lanes 16-31 o[4..7] of sample l % 16.
struct Fwd {
    f32x4 h1[4];
    f32x4 h2[4];
    f32x4 o;
};

__device__ inline void forward_tile(const FwdFrags *__restrict__ F, bf16x4 x,
    Fwd &r)
{
    const int l = threadIdx.x & 63;
    const int fo = 4 * (l >> 4);
    bf16x4 h1b[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        f32x4 acc = load4(&F->b1[rt * 16 + fo]);
        acc = mfma16(F->f1[rt][l], x, acc);
        r.h1[rt] = tanh4(acc);
        h1b[rt] = pack4(r.h1[rt]);
    }
    const bf16x8 h1k[2] = {cat8(h1b[0], h1b[1]), cat8(h1b[2], h1b[3])};
    bf16x4 h2b[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
        f32x4 acc = load4(&F->b2[rt * 16 + fo]);
#pragma unroll
        for (int kt = 0; kt < 2; ++kt)
            acc = mfma32(F->f2[rt][kt][l], h1k[kt], acc);
        r.h2[rt] = tanh4(acc);
        h2b[rt] = pack4(r.h2[rt]);
    }
    f32x4 acc = load4(&F->b3[fo]);
    acc = mfma32(F->f3[0][l], cat8(h2b[0], h2b[1]), acc);
    acc = mfma32(F->f3[1][l], cat8(h2b[2], h2b[3]), acc);
    r.o = acc;
}

// Normalised, rounded input fragment of a lane from the sample's raw
// features f[6] (lanes 32-63 carry the zero padding rows).
__device__ inline bf16x4 input_frag(const FwdFrags *__restrict__ F,
    const float (&f)[6])
{
    const int l = threadIdx.x & 63;
    const int fo = 4 * (l >> 4);
    f32x4 v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = fo + q;
        if (k < kIn)
            v[q] = (f[k] - F->shift[k]) * F->scale[k];
    }
    return pack4(v);
}

} // namespace mlp

// ---- host side (mlp.hip)

struct MlpSet {
    float *params = nullptr;    // FP32 master parameters [kParams]
    float *adam_m = nullptr;
    float *adam_v = nullptr;
    mlp::Frags *frags = nullptr;
    float *part = nullptr;      // per-wave gradient partials
    long long part_rows = 0;
    float *loss_hist = nullptr; // device losses of the last train call
    long long loss_cap = 0;
    long long step = 0;         // Adam step counter
    float shift[8] = {0};
    float scale[8] = {1, 1, 1, 1, 1, 1, 1, 1};
    hipStream_t stream = nullptr;
};

int mlp_create(MlpSet **out, hipStream_t st);
void mlp_destroy(MlpSet *m);
int mlp_set_params(MlpSet *m, const float *host_params, const float *shift,
    const float *scale);
int mlp_get_params(MlpSet *m, float *host_params);
int mlp_forward(MlpSet *m, long long n, const float *feat, float *out);
int mlp_loss_grad(MlpSet *m, long long n, const float *feat,
    const float *target, float *grad_out, float *loss_out);
int mlp_train(MlpSet *m, long long n, const float *feat, const float *target,
    long long batch, long long steps, float lr, float *loss_host);

} // namespace vsmch
