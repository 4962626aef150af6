// mlp.hip -- training and evaluation of the learned proposal's network:
// fragment packing, forward, the loss gradient (forward + backward on the
// matrix cores, weight gradients reduced per wave), and Adam.
//
// Loss per sample (mean over the batch): sum_d [ls_d + r_d^2 / 2] with
// r_d = (t_d - mu_d) exp(-ls_d), where (mu, ls) = outputs (0..3, 4..7) and
// t the standardised transition (x_t - m) / s of the proposal's target.
//
// Backward. With the transposed layout the output gradient g_o (D layout)
// is the B operand of g_h2 = W3^T g_o, and so on down the network (A =
// pre-swizzled W^T tiles). The weight gradients dW = sum_s g(s) h(s)^T
// contract over samples, so the per-tile gradients and activations go
// through LDS as [feature][16 samples] bfloat16 arrays, from which both the
// A (feature rows, sample K) and B (sample K, feature columns) fragments
// are single 8-byte reads.
#include "mlp.hpp"
#include "internal.hpp"

#include <math.h>
#include <vector>

namespace vsmch {

using mlp::bf16x4;
using mlp::f32x4;
using mlp::Frags;
using mlp::kGradLen;
using mlp::kH;
using mlp::kIn;
using mlp::kOut;
using mlp::kParams;

namespace {

constexpr int kMB = 256;   // threads per block (4 waves)

// One thread per lane-entry of the fragment arrays, plus biases.
__global__ void __launch_bounds__(kMB) k_pack(const float *__restrict__ p,
    const float *__restrict__ shift, const float *__restrict__ scale,
    Frags *__restrict__ F)
{
    const int id = blockIdx.x * kMB + threadIdx.x;
    if (id >= (4 + 16 + 4 + 4 + 16 + 1) * 64)
        return;
    const int l = id & 63;
    const int li = l & 15;
    const int kq = 4 * (l >> 4);
    int e = id >> 6;     // entry index over [f1 4 | f2 16 | f3 4 | g3 4 | g2 16]
    bf16x4 v{0, 0, 0, 0};
    const int g = l >> 4;
    if (e < 4) {
        const int rt = e;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (kq + q < kIn)
                v[q] = mlp::f2bf(p[mlp::oW1 + (rt * 16 + li) * kIn + kq + q]);
        F->fwd.f1[rt][l] = v;
    } else if ((e -= 4) < 16) {
        // K = 32 fragments: 8 rt x kt pairs, the other 8 entries idle
        if (e < 8) {
            const int rt = e >> 1, kt = e & 1;
            mlp::bf16x8 w;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                w[q] = __builtin_bit_cast(__bf16,
                    mlp::f2bf(p[mlp::oW2 + (rt * 16 + li) * kH + mlp::kslot(kt, g, q)]));
            F->fwd.f2[rt][kt][l] = w;
        }
    } else if ((e -= 16) < 4) {
        if (e < 2) {
            const int kt = e;
            mlp::bf16x8 w;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                w[q] = __builtin_bit_cast(__bf16, li < kOut ?
                    mlp::f2bf(p[mlp::oW3 + li * kH + mlp::kslot(kt, g, q)]) :
                    static_cast<uint16_t>(0));
            F->fwd.f3[kt][l] = w;
        }
    } else if ((e -= 4) < 4) {
        const int rt = e;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (kq + q < kOut)
                v[q] = mlp::f2bf(p[mlp::oW3 + (kq + q) * kH + rt * 16 + li]);
        F->g3[rt][l] = v;
    } else if ((e -= 4) < 16) {
        const int rt = e >> 2, kt = e & 3;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v[q] = mlp::f2bf(p[mlp::oW2 + (kt * 16 + kq + q) * kH + rt * 16 + li]);
        F->g2[rt][kt][l] = v;
    } else if ((e -= 16) == 0) {
        // biases and input normalisation (64 threads)
        F->fwd.b1[l] = p[mlp::oB1 + l];
        F->fwd.b2[l] = p[mlp::oB2 + l];
        if (l < 16)
            F->fwd.b3[l] = l < kOut ? p[mlp::oB3 + l] : 0.0f;
        if (l < 8) {
            F->fwd.shift[l] = l < kIn ? shift[l] : 0.0f;
            F->fwd.scale[l] = l < kIn ? scale[l] : 0.0f;
        }
    }
}

constexpr int kPackThreads = (4 + 16 + 4 + 4 + 16 + 1) * 64;

// Raw features of sample s (row of 8 floats, 6 used).
__device__ inline void load_feat(const float *feat, long long s, bool valid,
    float (&f)[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k)
        f[k] = valid ? feat[s * 8 + k] : 0.0f;
}

// Forward over n samples: out[s][8].
__global__ void __launch_bounds__(kMB) k_forward(const Frags *__restrict__ F,
    long long n, const float *__restrict__ feat, float *__restrict__ out)
{
    const int l = threadIdx.x & 63;
    const int fo = 4 * (l >> 4);
    const long long waves = static_cast<long long>(gridDim.x) * (kMB / 64);
    const long long w0 = static_cast<long long>(blockIdx.x) * (kMB / 64) +
        (threadIdx.x >> 6);
    for (long long tile = w0; tile * 16 < n; tile += waves) {
        const long long s = tile * 16 + (l & 15);
        const bool valid = s < n;
        float f[6];
        load_feat(feat, s, valid, f);
        mlp::Fwd r;
        mlp::forward_tile(&F->fwd, mlp::input_frag(&F->fwd, f), r);
        if (valid && fo < kOut) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                out[s * 8 + fo + v] = r.o[v];
        }
    }
}

struct alignas(16) GradLds {
    uint16_t xin[16][16];
    uint16_t go[16][16];
    uint16_t h1[64][16];
    uint16_t h2[64][16];
    uint16_t gz1[64][16];
    uint16_t gz2[64][16];
};

__device__ inline void put_tile(uint16_t (*a)[16], int rt, f32x4 v)
{
    const int l = threadIdx.x & 63;
    const int fo = 4 * (l >> 4);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        a[rt * 16 + fo + q][l & 15] = mlp::f2bf(v[q]);
}

// A / B fragment of rows [t * 16, t * 16 + 16) over the 16 samples.
__device__ inline bf16x4 get_frag(const uint16_t (*a)[16], int t)
{
    const int l = threadIdx.x & 63;
    return *reinterpret_cast<const bf16x4 *>(&a[t * 16 + (l & 15)][4 * (l >> 4)]);
}

__device__ inline float group16_sum(float x)
{
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    return x;
}

// Loss gradient over the batch {off + j * stride : j < batch} of n samples;
// each wave writes one partial row of kGradLen floats (gradient, loss).
__global__ void __launch_bounds__(kMB) k_grad(const Frags *__restrict__ F,
    long long batch, long long off, long long stride,
    const float *__restrict__ feat, const float *__restrict__ target,
    float inv_b, float *__restrict__ part)
{
    __shared__ GradLds lds_all[kMB / 64];
    const int wave = threadIdx.x >> 6;
    GradLds &L = lds_all[wave];
    const int l = threadIdx.x & 63;
    const int fo = 4 * (l >> 4);
    f32x4 dW1[4], dW2[4][4], dW3[4], db1[4], db2[4], db3;
    const f32x4 z4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        dW1[a] = z4;
        dW3[a] = z4;
        db1[a] = z4;
        db2[a] = z4;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            dW2[a][b] = z4;
    }
    db3 = z4;
    float loss = 0.0f;
    for (long long chunk = blockIdx.x; chunk * 64 < batch; chunk += gridDim.x) {
        const long long j = chunk * 64 + wave * 16 + (l & 15);
        const bool valid = j < batch;
        const long long s = off + j * stride;
        float f[6];
        load_feat(feat, s, valid, f);
        const bf16x4 xin = mlp::input_frag(&F->fwd, f);
        mlp::Fwd r;
        mlp::forward_tile(&F->fwd, xin, r);
        // output gradient (lanes 0-15: means, 16-31: log-scales)
        f32x4 other;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            other[q] = __shfl(r.o[q], l < 16 ? l + 16 : (l < 32 ? l - 16 : l), 64);
        f32x4 go = z4;
        if (valid && l < 32) {
            const f32x4 mu = l < 16 ? r.o : other;
            const f32x4 ls = l < 16 ? other : r.o;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float t = target[s * 4 + q];
                const float e = expf(-ls[q]);
                const float rr = (t - mu[q]) * e;
                if (l < 16) {
                    go[q] = -rr * e * inv_b;
                    loss += (ls[q] + 0.5f * rr * rr) * inv_b;
                } else {
                    go[q] = (1.0f - rr * rr) * inv_b;
                }
            }
        }
        const bf16x4 gob = mlp::pack4(go);
        db3 += go;
        // g_z2 = (W3^T g_o) (1 - h2^2)
        f32x4 gz2[4];
        bf16x4 gz2b[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            f32x4 acc = mlp::mfma16(F->g3[rt][l], gob, z4);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[q] *= 1.0f - r.h2[rt][q] * r.h2[rt][q];
            gz2[rt] = acc;
            gz2b[rt] = mlp::pack4(acc);
            db2[rt] += acc;
        }
        // g_z1 = (W2^T g_z2) (1 - h1^2)
        f32x4 gz1[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            f32x4 acc = z4;
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                acc = mlp::mfma16(F->g2[rt][kt][l], gz2b[kt], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                acc[q] *= 1.0f - r.h1[rt][q] * r.h1[rt][q];
            gz1[rt] = acc;
            db1[rt] += acc;
        }
        // stage [feature][sample] tiles for the sample contractions
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            L.xin[fo + q][l & 15] = static_cast<uint16_t>(xin[q]);
            L.go[fo + q][l & 15] = static_cast<uint16_t>(gob[q]);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            put_tile(L.h1, rt, r.h1[rt]);
            put_tile(L.h2, rt, r.h2[rt]);
            put_tile(L.gz1, rt, gz1[rt]);
            put_tile(L.gz2, rt, gz2[rt]);
        }
        __syncthreads();
        const bf16x4 a_go = get_frag(L.go, 0);
        const bf16x4 b_in = get_frag(L.xin, 0);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
            dW3[ct] = mlp::mfma16(a_go, get_frag(L.h2, ct), dW3[ct]);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const bf16x4 a2 = get_frag(L.gz2, rt);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                dW2[rt][ct] = mlp::mfma16(a2, get_frag(L.h1, ct), dW2[rt][ct]);
            dW1[rt] = mlp::mfma16(get_frag(L.gz1, rt), b_in, dW1[rt]);
        }
        __syncthreads();
    }
    // write this wave's partial row
    float *row = part + (static_cast<long long>(blockIdx.x) * (kMB / 64) + wave) *
        kGradLen;
    const int li = l & 15;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = rt * 16 + fo + q;
            if (li < kIn)
                row[mlp::oW1 + i * kIn + li] = dW1[rt][q];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                row[mlp::oW2 + i * kH + ct * 16 + li] = dW2[rt][ct][q];
            const float s1 = group16_sum(db1[rt][q]);
            const float s2 = group16_sum(db2[rt][q]);
            if (li == 0) {
                row[mlp::oB1 + i] = s1;
                row[mlp::oB2 + i] = s2;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = fo + q;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
            if (i < kOut)
                row[mlp::oW3 + i * kH + ct * 16 + li] = dW3[ct][q];
        const float s3 = group16_sum(db3[q]);
        if (li == 0 && i < kOut)
            row[mlp::oB3 + i] = s3;
    }
    // loss: sum over the wave
    float ls = loss;
    for (int m = 32; m >= 1; m >>= 1)
        ls += __shfl_xor(ls, m, 64);
    if (l == 0)
        row[kParams] = ls;
}

// Sums the partial rows; mode 0 writes the gradient (grad[kParams], loss at
// grad[kParams]); mode 1 applies one Adam step with bias corrections c1 =
// 1 - beta1^t, c2 = 1 - beta2^t and records the loss in loss_hist[idx].
__global__ void __launch_bounds__(kMB) k_reduce(const float *__restrict__ part,
    long long rows, int mode, float *__restrict__ params, float *__restrict__ m,
    float *__restrict__ v, float lr, float beta1, float beta2, float c1,
    float c2, float eps, float *__restrict__ grad, float *__restrict__ loss_hist,
    long long idx)
{
    const int p = blockIdx.x * kMB + threadIdx.x;
    if (p >= kGradLen)
        return;
    float g = 0.0f;
    for (long long r = 0; r < rows; ++r)
        g += part[r * kGradLen + p];
    if (p == kParams) {
        if (mode == 0)
            grad[p] = g;
        else if (loss_hist)
            loss_hist[idx] = g;
        return;
    }
    if (mode == 0) {
        grad[p] = g;
        return;
    }
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

constexpr int kGradBlocksMax = 256;

int pack(MlpSet *m)
{
    float *sh = m->adam_m + kParams;   // shift / scale live after the moments
    hipLaunchKernelGGL(k_pack, dim3((kPackThreads + kMB - 1) / kMB), dim3(kMB), 0, m->stream,
        m->params, sh, sh + 8, m->frags);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int ensure_part(MlpSet *m, long long rows)
{
    if (rows <= m->part_rows)
        return kOk;
    if (m->part)
        VSMCH_HIP(hipFree(m->part));
    m->part = nullptr;
    VSMCH_HIP(hipMalloc(&m->part, sizeof(float) * kGradLen * rows));
    m->part_rows = rows;
    return kOk;
}

long long grad_blocks(long long batch)
{
    const long long c = (batch + 63) / 64;
    return c < kGradBlocksMax ? (c < 1 ? 1 : c) : kGradBlocksMax;
}

} // namespace

int mlp_create(MlpSet **out, hipStream_t st)
{
    MlpSet *m = new MlpSet();
    m->stream = st;
    // params | adam m (+ shift, scale) | adam v
    if (hipMalloc(&m->params, sizeof(float) * kParams) != hipSuccess ||
        hipMalloc(&m->adam_m, sizeof(float) * (kParams + 16)) != hipSuccess ||
        hipMalloc(&m->adam_v, sizeof(float) * kParams) != hipSuccess ||
        hipMalloc(&m->frags, sizeof(Frags)) != hipSuccess) {
        set_error("mlp_create: device allocation failed");
        mlp_destroy(m);
        return kErrAlloc;
    }
    std::vector<float> zero(kParams + 16, 0.0f);
    for (int k = 0; k < 8; ++k)
        zero[kParams + 8 + k] = 1.0f;
    VSMCH_HIP(hipMemcpyAsync(m->params, zero.data(), sizeof(float) * kParams,
        hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipMemcpyAsync(m->adam_m, zero.data(), sizeof(float) * (kParams + 16),
        hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipMemcpyAsync(m->adam_v, zero.data(), sizeof(float) * kParams,
        hipMemcpyHostToDevice, st));
    VSMCH_HIP(hipStreamSynchronize(st));
    int rc = pack(m);
    if (rc != kOk) {
        mlp_destroy(m);
        return rc;
    }
    *out = m;
    return kOk;
}

void mlp_destroy(MlpSet *m)
{
    if (!m)
        return;
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->params);
    (void)hipFree(m->adam_m);
    (void)hipFree(m->adam_v);
    (void)hipFree(m->frags);
    (void)hipFree(m->part);
    (void)hipFree(m->loss_hist);
    delete m;
}

int mlp_set_params(MlpSet *m, const float *hp, const float *shift,
    const float *scale)
{
    std::vector<float> z(kParams, 0.0f);
    if (hp)
        VSMCH_HIP(hipMemcpyAsync(m->params, hp, sizeof(float) * kParams,
            hipMemcpyHostToDevice, m->stream));
    float ss[16];
    for (int k = 0; k < 8; ++k) {
        ss[k] = (shift && k < kIn) ? shift[k] : (k < kIn ? m->shift[k] : 0.0f);
        ss[8 + k] = (scale && k < kIn) ? scale[k] : (k < kIn ? m->scale[k] : 0.0f);
        m->shift[k] = ss[k];
        m->scale[k] = ss[8 + k];
    }
    VSMCH_HIP(hipMemcpyAsync(m->adam_m + kParams, ss, sizeof(ss),
        hipMemcpyHostToDevice, m->stream));
    // a new parameter vector restarts the optimiser
    VSMCH_HIP(hipMemcpyAsync(m->adam_m, z.data(), sizeof(float) * kParams,
        hipMemcpyHostToDevice, m->stream));
    VSMCH_HIP(hipMemcpyAsync(m->adam_v, z.data(), sizeof(float) * kParams,
        hipMemcpyHostToDevice, m->stream));
    m->step = 0;
    int rc = pack(m);
    VSMCH_HIP(hipStreamSynchronize(m->stream));
    return rc;
}

int mlp_get_params(MlpSet *m, float *hp)
{
    VSMCH_HIP(hipMemcpyAsync(hp, m->params, sizeof(float) * kParams,
        hipMemcpyDeviceToHost, m->stream));
    VSMCH_HIP(hipStreamSynchronize(m->stream));
    return kOk;
}

int mlp_forward(MlpSet *m, long long n, const float *feat, float *out)
{
    if (n < 1)
        return kOk;
    const long long tiles = (n + 15) / 16;
    long long blocks = (tiles + 3) / 4;
    blocks = blocks > 4096 ? 4096 : blocks;
    hipLaunchKernelGGL(k_forward, dim3(blocks), dim3(kMB), 0, m->stream,
        m->frags, n, feat, out);
    VSMCH_HIP(hipGetLastError());
    return kOk;
}

int mlp_loss_grad(MlpSet *m, long long n, const float *feat,
    const float *target, float *grad_out, float *loss_out)
{
    if (n < 1) {
        set_error("mlp_loss_grad: empty batch");
        return kErrInvalid;
    }
    const long long blocks = grad_blocks(n);
    int rc = ensure_part(m, blocks * (kMB / 64));
    if (rc != kOk)
        return rc;
    hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(kMB), 0, m->stream, m->frags,
        n, 0LL, 1LL, feat, target, 1.0f / static_cast<float>(n), m->part);
    hipLaunchKernelGGL(k_reduce, dim3((kGradLen + kMB - 1) / kMB), dim3(kMB), 0,
        m->stream, m->part, blocks * (kMB / 64), 0, m->params, m->adam_m,
        m->adam_v, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 0.0f, grad_out,
        static_cast<float *>(nullptr), 0LL);
    VSMCH_HIP(hipGetLastError());
    if (loss_out) {
        VSMCH_HIP(hipMemcpyAsync(loss_out, grad_out + kParams, sizeof(float),
            hipMemcpyDeviceToHost, m->stream));
        VSMCH_HIP(hipStreamSynchronize(m->stream));
    }
    return kOk;
}

int mlp_train(MlpSet *m, long long n, const float *feat, const float *target,
    long long batch, long long steps, float lr, float *loss_host)
{
    if (n < 1 || batch < 1 || batch > n || steps < 0) {
        set_error("mlp_train: need 1 <= batch <= n and steps >= 0");
        return kErrInvalid;
    }
    if (steps > m->loss_cap) {
        if (m->loss_hist)
            VSMCH_HIP(hipFree(m->loss_hist));
        m->loss_hist = nullptr;
        VSMCH_HIP(hipMalloc(&m->loss_hist, sizeof(float) * (steps > 0 ? steps : 1)));
        m->loss_cap = steps;
    }
    const long long blocks = grad_blocks(batch);
    int rc = ensure_part(m, blocks * (kMB / 64));
    if (rc != kOk)
        return rc;
    // batches {b + j * nb : j < batch}: every sample once per nb steps,
    // each batch spread over the whole data set
    const long long nb = n / batch;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    for (long long it = 0; it < steps; ++it) {
        const long long b = it % nb;
        ++m->step;
        const float c1 = 1.0f - powf(beta1, static_cast<float>(m->step));
        const float c2 = 1.0f - powf(beta2, static_cast<float>(m->step));
        hipLaunchKernelGGL(k_grad, dim3(blocks), dim3(kMB), 0, m->stream,
            m->frags, batch, b, nb, feat, target,
            1.0f / static_cast<float>(batch), m->part);
        hipLaunchKernelGGL(k_reduce, dim3((kGradLen + kMB - 1) / kMB), dim3(kMB),
            0, m->stream, m->part, blocks * (kMB / 64), 1, m->params, m->adam_m,
            m->adam_v, lr, beta1, beta2, c1, c2, eps,
            static_cast<float *>(nullptr), m->loss_hist, it);
        rc = pack(m);
        if (rc != kOk)
            return rc;
    }
    VSMCH_HIP(hipGetLastError());
    if (loss_host && steps > 0) {
        VSMCH_HIP(hipMemcpyAsync(loss_host, m->loss_hist, sizeof(float) * steps,
            hipMemcpyDeviceToHost, m->stream));
    }
    VSMCH_HIP(hipStreamSynchronize(m->stream));
    return kOk;
}

} // namespace vsmch
