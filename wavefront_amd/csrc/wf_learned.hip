// Learned thin absorbing boundary: the per-face MLP (BF16 weights on MFMA,
// FP32 accumulate), its training kernels (forward, backward, Adam), and the
// face hook that gathers the boundary stencil each half step and either
// predicts the ghost plane or records training samples.
//
// MFMA layout of v_mfma_f32_16x16x16_bf16 (D = A.B + C, 16x16x16): lane l
// holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and D[4(l/16)+j][l%16],
// j = 0..3.  Layers are evaluated transposed (features x samples): a wave
// carries 16 samples as the 16 columns, so a layer's output D is already in
// the B layout the next layer consumes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "wf_internal.h"
#include "wf_learned.h"

namespace wf {
namespace {

__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Four floats to bf16, round to nearest even: two v_cvt_pk_bf16_f32 (the
// same bits as bf16_rne for finite inputs).
__device__ __forceinline__ bfv4 to_bf4(fv4 v)
{
    const bf16x2 lo = __builtin_convertvector((f32x2{v.x, v.y}), bf16x2);
    const bf16x2 hi = __builtin_convertvector((f32x2{v.z, v.w}), bf16x2);
    const uint32_t a = __builtin_bit_cast(uint32_t, lo), b = __builtin_bit_cast(uint32_t, hi);
    bfv4 r;
    r.x = (short)(a & 0xffffu);
    r.y = (short)(a >> 16);
    r.z = (short)(b & 0xffffu);
    r.w = (short)(b >> 16);
    return r;
}

__device__ __forceinline__ fv4 mfma(bfv4 a, bfv4 b, fv4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// tanh(x) = sign(x) (1 - 2 / (exp(2|x|) + 1)), with the hardware exp2 and
// reciprocal (v_exp_f32, v_rcp_f32: about 1 ulp each).  Large |x|: exp2
// gives +inf, the reciprocal 0, tanh +-1.
__device__ __forceinline__ float tanh_f(float x)
{
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * fabsf(x));  // 2 log2(e) |x|
    const float t = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    return copysignf(t, x);
}

// Hidden activation: hardtanh, clamp(z, -1, 1) (one v_med3_f32; odd and
// bounded like tanh at a twelfth of its cost), and its derivative from the
// activation's value: 1 inside (-1, 1), 0 where it clamps.
__device__ __forceinline__ float act_f(float z) { return __builtin_amdgcn_fmed3f(z, -1.0f, 1.0f); }
__device__ __forceinline__ fv4 act4(fv4 z) { return fv4{act_f(z.x), act_f(z.y), act_f(z.z), act_f(z.w)}; }
__device__ __forceinline__ float dact_f(float a) { return fabsf(a) < 1.0f ? 1.0f : 0.0f; }
__device__ __forceinline__ fv4 dact4(fv4 a) { return fv4{dact_f(a.x), dact_f(a.y), dact_f(a.z), dact_f(a.w)}; }

// The Higdon head's fixed factor and damping, and the time smoothing of
// its adaptive coefficient: wlin[0..3] = kf, gf, ga, beta.
struct HigdonFix {
    float kf, gf, ga;
    float beta;  // 0 or >= 1: none
};


// The network in registers: A operands and the per-lane biases.
struct MlpRegs {
    bfv4 A1[4], A2[4][4], A3[4];
    fv4 b1[4], b2[4];
    fv4 wl;
    HigdonFix hf;  // wlin[0..2]: the Higdon head's fixed factor
    float b3;
    int head;
};

__device__ __forceinline__ void load_regs(const MlpPacked *__restrict__ w, int lane, MlpRegs &r)
{
    const int rb = 4 * (lane >> 4);
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        r.A1[o] = w->a1[o][lane];
        r.A3[o] = w->a3[o][lane];
        r.b1[o] = fv4{w->b1[16 * o + rb], w->b1[16 * o + rb + 1], w->b1[16 * o + rb + 2],
                      w->b1[16 * o + rb + 3]};
        r.b2[o] = fv4{w->b2[16 * o + rb], w->b2[16 * o + rb + 1], w->b2[16 * o + rb + 2],
                      w->b2[16 * o + rb + 3]};
#pragma unroll
        for (int k = 0; k < 4; ++k) r.A2[o][k] = w->a2[o][k][lane];
    }
    r.wl = fv4{w->wlin[rb], w->wlin[rb + 1], w->wlin[rb + 2], w->wlin[rb + 3]};
    r.hf = HigdonFix{w->wlin[0], w->wlin[1], w->wlin[2], w->wlin[3]};
    r.b3 = w->b3;
    r.head = w->head;
}

// Sign that makes the one-way head's network input even: the sign of u0, or
// of u4 when u0 = 0 (same value on the four lanes of a sample).
__device__ __forceinline__ float input_sign(fv4 u, int sl)
{
    const float u0 = __shfl(u.x, sl), u4 = __shfl(u.x, sl + 16);
    return (u0 > 0.f || (u0 == 0.f && u4 >= 0.f)) ? 1.0f : -1.0f;
}

// The scaled history and stencil values a head combines, on lanes 0..15
// (sample = lane): P0, P1, P1(n-1), P1(n-2), g(n-1), g(n-2), P0(n-1),
// P0(n-2).  Every lane takes part in the shuffles.
struct HeadIn {
    float u0, u1, u10, u11, u12, u13, u14, u15;
};
__device__ __forceinline__ HeadIn head_in(fv4 u, int sl)
{
    HeadIn h;
    h.u0 = __shfl(u.x, sl);
    h.u1 = __shfl(u.y, sl);
    h.u10 = __shfl(u.z, sl + 32);
    h.u11 = __shfl(u.w, sl + 32);
    h.u12 = __shfl(u.x, sl + 48);
    h.u13 = __shfl(u.y, sl + 48);
    h.u14 = __shfl(u.z, sl + 48);
    h.u15 = __shfl(u.w, sl + 48);
    return h;
}

// The Higdon head's terms: ghost / s = A + ka B (the product's expansion
// in the data u, with the history h).
__device__ __forceinline__ void higdon_ab(const HeadIn &h, const HigdonFix &c, float &A, float &B)
{
    const float kf = c.kf, gf = c.gf, ga = c.ga;
    A = kf * ((h.u0 - gf * h.u12) - ga * h.u10) + (gf + ga) * h.u14 +
        gf * ga * (kf * h.u15 - h.u11);
    B = ((h.u0 - ga * h.u12) - kf * h.u1) + (kf * (ga + gf)) * h.u14 +
        gf * (ga * (h.u15 - kf * h.u13) - h.u10);
}

// Prediction in scaled units from the network's MLP output m and linear
// skip p (lanes 0..15); c: the Higdon head's fixed factor.
__device__ __forceinline__ float head_out(int head, float m, float p, const HeadIn &h,
                                          const HigdonFix &c)
{
    if (head == kHeadDirect) return m + p;
    if (head == kHeadOneWay) return h.u14 + tanh_f(m + p) * (h.u0 - h.u12);
    if (head == kHeadDiff) return h.u0 + (m + p);
    if (head == kHeadHigdon) {
        float A, B;
        higdon_ab(h, c, A, B);
        return A + tanh_f(m) * B;
    }
    return p + m * (h.u0 - h.u14);  // kHeadDelta
}

struct Acts {
    fv4 a1[4], a2[4];
    bfv4 a1b[4], a2b[4];  // the activations' leading bf16 parts
};

__device__ __forceinline__ fv4 from_bf4(bfv4 b)
{
    return fv4{bf16_f((unsigned short)b.x), bf16_f((unsigned short)b.y),
               bf16_f((unsigned short)b.z), bf16_f((unsigned short)b.w)};
}

// v as a pair of bf16 values, hi + lo: hi = v rounded, lo = the remainder
// rounded (|v - hi - lo| <= 2^-17 |v|).
__device__ __forceinline__ void split_bf4(fv4 v, bfv4 &hi, bfv4 &lo)
{
    hi = to_bf4(v);
    lo = to_bf4(v - from_bf4(hi));
}

// Forward pass for the wave's 16 samples.  ui: this lane's 4 MLP inputs,
// uf: its 4 linear-skip inputs (features 4(l/16)..+3 of sample l%16).
// pairs: inputs and hidden activations enter each product as bf16 pairs
// (two MFMAs against the same bf16 weights), so the network is a function of
// its inputs to about FP32 precision -- for the heads whose network output
// is itself part of the ghost value (direct, delta, difference: one bf16
// rounding moves it by up to 2^-9 of the terms, in steps that do not vary
// smoothly with the fields); otherwise single bf16 values (the one-way and
// Higdon heads, where the output sets a bounded coefficient).  Returns the
// MLP output and sets p to the skip, valid on lanes 0..15 (sample = lane);
// ub gets the inputs' leading bf16 parts.
__device__ __forceinline__ bool head_pairs(int head) { return head != kHeadOneWay && head != kHeadHigdon; }

__device__ __forceinline__ float mlp_forward(const MlpRegs &r, fv4 ui, fv4 uf, Acts &t, float &p,
                                             bfv4 &ub, bool pairs)
{
    const fv4 zero = {0.f, 0.f, 0.f, 0.f};
    bfv4 ul, a1l[4], a2l[4];
    split_bf4(ui, ub, ul);
    // biases enter as the accumulator of the first product of each layer
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        fv4 z = mfma(r.A1[o], ub, r.b1[o]);
        if (pairs) z = mfma(r.A1[o], ul, z);
        t.a1[o] = act4(z);
        split_bf4(t.a1[o], t.a1b[o], a1l[o]);
    }
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        fv4 z = r.b2[o];
#pragma unroll
        for (int k = 0; k < 4; ++k) z = mfma(r.A2[o][k], t.a1b[k], z);
        if (pairs) {
#pragma unroll
            for (int k = 0; k < 4; ++k) z = mfma(r.A2[o][k], a1l[k], z);
        }
        t.a2[o] = act4(z);
        split_bf4(t.a2[o], t.a2b[o], a2l[o]);
    }
    fv4 z3 = zero;
#pragma unroll
    for (int k = 0; k < 4; ++k) z3 = mfma(r.A3[k], t.a2b[k], z3);
    if (pairs) {
#pragma unroll
        for (int k = 0; k < 4; ++k) z3 = mfma(r.A3[k], a2l[k], z3);
    }
    // linear skip, summed over the four lanes that hold one sample
    p = r.wl.x * uf.x + r.wl.y * uf.y + r.wl.z * uf.z + r.wl.w * uf.w;
    p = p + __shfl_xor(p, 16);
    p = p + __shfl_xor(p, 32);
    return z3.x + r.b3;
}

// The difference head's inputs (kHeadDiff): this lane's four of sample sl,
// lane group q4 = inputs 4 q4 .. 4 q4 + 3.  Every lane takes part in the
// shuffle.
__device__ __forceinline__ fv4 diff_in(fv4 u, int sl)
{
    const float p0 = __shfl(u.x, sl), p1 = __shfl(u.y, sl);
    const int q4 = (threadIdx.x & 63) >> 4;
    if (q4 < 2) return fv4{u.x - u.y, u.y - u.z, u.z - u.w, 0.f};
    if (q4 == 2) return fv4{u.x, u.y, u.z - p1, u.w - u.z};
    return fv4{u.x - p0, u.y - u.x, u.z - p0, u.w - u.z};
}

// MLP input of a sample: the scaled inputs, sign-normalized for the heads
// whose network must be even (one-way, delta), their differences for the
// difference head; skip input: sign-normalized for the one-way head (its
// skip is inside the tanh), the differences for the difference head.
__device__ __forceinline__ fv4 mlp_in(int head, fv4 u, int sl)
{
    if (head == kHeadDirect) return u;
    if (head == kHeadDiff) return diff_in(u, sl);
    return u * input_sign(u, sl);
}
__device__ __forceinline__ fv4 skip_in(int head, fv4 u, fv4 ui)
{
    return head == kHeadOneWay || head == kHeadDiff ? ui : u;
}

// FP32 parameters -> per-lane MFMA operands.  One wave.
__global__ void k_pack_mlp(const float *__restrict__ p, int head, MlpPacked *__restrict__ w)
{
    const int l = threadIdx.x, c = l & 15, rb = 4 * (l >> 4);
    auto bf = [](float v) { return (short)bf16_rne(v); };
    for (int o = 0; o < 4; ++o) {
        bfv4 v;
        v.x = bf(p[kOffW1 + (rb + 0) * kMlpHid + 16 * o + c]);
        v.y = bf(p[kOffW1 + (rb + 1) * kMlpHid + 16 * o + c]);
        v.z = bf(p[kOffW1 + (rb + 2) * kMlpHid + 16 * o + c]);
        v.w = bf(p[kOffW1 + (rb + 3) * kMlpHid + 16 * o + c]);
        w->a1[o][l] = v;
        for (int k = 0; k < 4; ++k) {
            bfv4 u;
            u.x = bf(p[kOffW2 + (16 * k + rb + 0) * kMlpHid + 16 * o + c]);
            u.y = bf(p[kOffW2 + (16 * k + rb + 1) * kMlpHid + 16 * o + c]);
            u.z = bf(p[kOffW2 + (16 * k + rb + 2) * kMlpHid + 16 * o + c]);
            u.w = bf(p[kOffW2 + (16 * k + rb + 3) * kMlpHid + 16 * o + c]);
            w->a2[o][k][l] = u;
            // backward operand: block (i = k, o), A[i][k'] = W2[16k + c][16o + rb + j]
            bfv4 b;
            b.x = bf(p[kOffW2 + (16 * k + c) * kMlpHid + 16 * o + rb + 0]);
            b.y = bf(p[kOffW2 + (16 * k + c) * kMlpHid + 16 * o + rb + 1]);
            b.z = bf(p[kOffW2 + (16 * k + c) * kMlpHid + 16 * o + rb + 2]);
            b.w = bf(p[kOffW2 + (16 * k + c) * kMlpHid + 16 * o + rb + 3]);
            w->a2b[k][o][l] = b;
        }
        bfv4 a3 = {0, 0, 0, 0};
        if (c == 0) {
            a3.x = bf(p[kOffW3 + 16 * o + rb + 0]);
            a3.y = bf(p[kOffW3 + 16 * o + rb + 1]);
            a3.z = bf(p[kOffW3 + 16 * o + rb + 2]);
            a3.w = bf(p[kOffW3 + 16 * o + rb + 3]);
        }
        w->a3[o][l] = a3;
    }
    w->b1[l] = p[kOffB1 + l];
    w->b2[l] = p[kOffB2 + l];
    w->w3r[l] = bf16_f(bf16_rne(p[kOffW3 + l]));
    if (l < kMlpIn) w->wlin[l] = p[kOffWL + l];
    if (l == 0) {
        w->b3 = p[kOffB3];
        w->head = head;
    }
}

// Prediction for a matrix of already-scaled inputs (tests and evaluation).
__global__ __launch_bounds__(256) void k_mlp_predict(const float *__restrict__ u, int n,
                                                     const MlpPacked *__restrict__ w,
                                                     float *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int ngroups = (n + 15) / 16;
    MlpRegs r;
    load_regs(w, lane, r);
    const int sl = lane & 15, rb = 4 * (lane >> 4);
    for (int grp = wave; grp < ngroups; grp += nwaves) {
        const int s = grp * 16 + sl;
        fv4 uf = {0.f, 0.f, 0.f, 0.f};
        if (s < n) uf = *reinterpret_cast<const fv4 *>(u + (size_t)s * kMlpIn + rb);
        const fv4 ui = mlp_in(r.head, uf, sl);
        Acts t;
        float sk;
        bfv4 ub;
        const float m = mlp_forward(r, ui, skip_in(r.head, uf, ui), t, sk, ub, head_pairs(r.head));
        const float p = head_out(r.head, m, sk, head_in(uf, sl), r.hf);
        if (lane < 16 && s < n) out[s] = p;
    }
}

// 16x16 bf16 tile transpose through LDS: writes X (D/B layout: lane holds
// X[4(l/16)+j][l%16]) and returns X^T's operand: lane gets X[l%16][4(l/16)+j].
__device__ __forceinline__ bfv4 transpose16(unsigned short *tile, bfv4 x, int lane)
{
    const int c = lane & 15, rb = 4 * (lane >> 4);
    tile[(rb + 0) * 16 + c] = (unsigned short)x.x;
    tile[(rb + 1) * 16 + c] = (unsigned short)x.y;
    tile[(rb + 2) * 16 + c] = (unsigned short)x.z;
    tile[(rb + 3) * 16 + c] = (unsigned short)x.w;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    const bfv4 r = *reinterpret_cast<const bfv4 *>(tile + c * 16 + rb);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    return r;
}

__device__ __forceinline__ float sum16(float v)
{
    v = v + __shfl_xor(v, 1);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 8);
    return v;
}

// Sum over the 64 lanes by a fixed butterfly (every lane gets the total).
__device__ __forceinline__ float wave_sum(float v)
{
    v = v + __shfl_xor(v, 1);
    v = v + __shfl_xor(v, 2);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 16);
    v = v + __shfl_xor(v, 32);
    return v;
}

__device__ __forceinline__ fv4 sum16v(fv4 v) { return fv4{sum16(v.x), sum16(v.y), sum16(v.z), sum16(v.w)}; }

// Mean-squared-error gradient of the whole sample set: every wave loops
// over groups of 16 samples and writes its partial gradient (and partial
// squared error) as one row of `partials` ([nwaves][kMlpParams + 1]).
__global__ __launch_bounds__(256) void k_mlp_train(const float *__restrict__ u,
                                                   const float *__restrict__ tgt, int n,
                                                   float inv_n, const MlpPacked *__restrict__ w,
                                                   float *__restrict__ partials)
{
    __shared__ unsigned short tiles[4][256];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int ngroups = (n + 15) / 16;
    const int sl = lane & 15, rb = 4 * (lane >> 4);
    unsigned short *tile = tiles[wv];

    MlpRegs r;
    load_regs(w, lane, r);
    bfv4 A2B[4][4];
    fv4 w3r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int o = 0; o < 4; ++o) A2B[i][o] = w->a2b[i][o][lane];
        w3r[i] = fv4{w->w3r[16 * i + rb], w->w3r[16 * i + rb + 1], w->w3r[16 * i + rb + 2],
                     w->w3r[16 * i + rb + 3]};
    }
    const fv4 zero = {0.f, 0.f, 0.f, 0.f};
    fv4 gW1[4], gW2[4][4], gb1[4], gb2[4], gw3[4], gwl = zero;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gW1[i] = gb1[i] = gb2[i] = gw3[i] = zero;
#pragma unroll
        for (int o = 0; o < 4; ++o) gW2[i][o] = zero;
    }
    float gb3 = 0.f, loss = 0.f;

    for (int grp = wave; grp < ngroups; grp += nwaves) {
        const int s = grp * 16 + sl;
        const bool live = s < n;
        fv4 uf = zero;
        float t = 0.f;
        if (live) {
            uf = *reinterpret_cast<const fv4 *>(u + (size_t)s * kMlpIn + rb);
            t = tgt[s];
        }
        const fv4 ui = mlp_in(r.head, uf, sl);
        const fv4 us = skip_in(r.head, uf, ui);
        bfv4 ub;
        Acts a;
        float sk;
        const float m = mlp_forward(r, ui, us, a, sk, ub, head_pairs(r.head));
        const HeadIn hv = head_in(uf, sl);
        const float pred = head_out(r.head, m, sk, hv, r.hf);
        const float res = (lane < 16 && live) ? pred - t : 0.f;
        loss = loss + res * res;
        float g = 2.0f * res * inv_n;  // dL/dpred on lanes 0..15
        float gm = g, gp = g;           // dL/dm, dL/dp
        if (r.head == kHeadOneWay) {
            const float th = tanh_f(m + sk);
            gm = gp = g * (1.0f - th * th) * (hv.u0 - hv.u12);
        } else if (r.head == kHeadDelta) {
            gm = g * (hv.u0 - hv.u14);
        } else if (r.head == kHeadHigdon) {
            float A, B;
            higdon_ab(hv, r.hf, A, B);
            const float th = tanh_f(m);
            gm = g * (1.0f - th * th) * B;
            gp = 0.f;  // the skip is not used (wlin[0..3] hold the fixed factor)
        }
        const float gs = __shfl(gm, sl);  // broadcast to the sample's four lanes
        if (lane < 16) gb3 = gb3 + gm;
        gwl = gwl + us * __shfl(gp, sl);
        // layer 3 (rank one): dW3 = a2 g, da2 = W3 g
        fv4 dz2[4];
        bfv4 dz2b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            gw3[k] = gw3[k] + a.a2[k] * gs;
            const fv4 da2 = w3r[k] * gs;
            dz2[k] = da2 * dact4(a.a2[k]);
            gb2[k] = gb2[k] + dz2[k];
            dz2b[k] = to_bf4(dz2[k]);
        }
        // dW2 = a1 dz2^T: K = the 16 samples, operands transposed via LDS
        bfv4 a1T[4], dz2T[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a1T[k] = transpose16(tile, a.a1b[k], lane);
            dz2T[k] = transpose16(tile, dz2b[k], lane);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int o = 0; o < 4; ++o) gW2[i][o] = mfma(a1T[i], dz2T[o], gW2[i][o]);
        // da1 = W2 dz2 on MFMA; dz1 = da1 act'(a1)
        bfv4 dz1b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            fv4 d = zero;
#pragma unroll
            for (int o = 0; o < 4; ++o) d = mfma(A2B[i][o], dz2b[o], d);
            const fv4 dz1 = d * dact4(a.a1[i]);
            gb1[i] = gb1[i] + dz1;
            dz1b[i] = to_bf4(dz1);
        }
        // dW1 = u dz1^T
        const bfv4 uT = transpose16(tile, ub, lane);
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const bfv4 dz1T = transpose16(tile, dz1b[o], lane);
            gW1[o] = mfma(uT, dz1T, gW1[o]);
        }
    }

    // sum the per-sample partials over the 16 lanes that share features
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        gb1[i] = sum16v(gb1[i]);
        gb2[i] = sum16v(gb2[i]);
        gw3[i] = sum16v(gw3[i]);
    }
    gwl = sum16v(gwl);
    gb3 = sum16(gb3);
    loss = sum16(loss);

    float *out = partials + (size_t)wave * (kMlpParams + 1);
#pragma unroll
    for (int o = 0; o < 4; ++o) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            out[kOffW1 + (rb + j) * kMlpHid + 16 * o + sl] = gW1[o][j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[kOffW2 + (16 * i + rb + j) * kMlpHid + 16 * o + sl] = gW2[i][o][j];
        }
    }
    if (sl == 0) {
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                out[kOffB1 + 16 * o + rb + j] = gb1[o][j];
                out[kOffB2 + 16 * o + rb + j] = gb2[o][j];
                out[kOffW3 + 16 * o + rb + j] = gw3[o][j];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) out[kOffWL + rb + j] = gwl[j];
    }
    if (lane == 0) {
        out[kOffB3] = gb3;
        out[kMlpParams] = loss;
    }
}

// Sums the wave rows in a fixed order, in two stages so that enough lanes
// run: stage one sums chunk blockIdx.y of `per` consecutive rows into row
// blockIdx.y of `out`; stage two (gridDim.y = 1, per = rows) sums those.
// Row kMlpParams of a column is the squared error.
__global__ void k_mlp_reduce(const float *__restrict__ partials, int rows, int per,
                             float *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    const int r0 = blockIdx.y * per, r1 = min(r0 + per, rows);
    float s = 0.f;
    for (int r = r0; r < r1; ++r) s = s + partials[(size_t)r * (kMlpParams + 1) + p];
    out[(size_t)blockIdx.y * (kMlpParams + 1) + p] = s;
}

// Adam; the linear skip (wlin) moves at lr_skip instead of lr, and with
// `odd` the biases stay where they are (zero: the network is then an odd
// function of its inputs, like the linear field update it stands in for).
__global__ void k_mlp_adam(float *__restrict__ prm, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float lr, float lr_skip, int odd,
                           float beta1, float beta2, float eps, float bc1, float bc2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    if (p >= kOffWL) lr = lr_skip;
    const bool bias = (p >= kOffB1 && p < kOffW2) || (p >= kOffB2 && p < kOffW3) || p == kOffB3;
    if (odd && bias) return;
    const float g = grad[p];
    const float mn = beta1 * m[p] + (1.0f - beta1) * g;
    const float vn = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mn;
    v[p] = vn;
    prm[p] = prm[p] - lr * (mn / bc1) / (sqrtf(vn / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Face hook
// ---------------------------------------------------------------------------
// Face coordinates (normal n, fast f, slow s) -> linear cell index.  The
// fast axis is the lower-numbered tangential axis: a = 0 (y, z), 1 (x, z),
// 2 (x, y).
__device__ __forceinline__ size_t face_cell(const Geo &g, int a, int n, int f, int s)
{
    const int x = a == 0 ? n : f;
    const int y = a == 0 ? f : (a == 1 ? n : s);
    const int z = a == 2 ? n : s;
    return ((size_t)z * g.ny + y) * g.nx + x;
}

// Pointer selection by a run-time index without indexing the by-value
// kernel argument (which would spill it to scratch).
__device__ __forceinline__ const float *field_ptr(const Ptrs &p, int c)
{
    switch (c) {
    case 0: return p.f[0];
    case 1: return p.f[1];
    case 2: return p.f[2];
    case 3: return p.f[3];
    case 4: return p.f[4];
    default: return p.f[5];
    }
}

// Passive limit on a predicted ghost value (scaled units: every field value
// divided by the sample's input scale s).
//
// The Yee scheme conserves two staggered energies, W^n = sum eps E^n.E^n / 2 +
// mu H^{n-1/2}.H^{n+1/2} / 2 and V^{n+1/2} = sum mu H^{n+1/2}.H^{n+1/2} / 2 +
// eps E^n.E^{n+1} / 2 (cell volumes and material coefficients folded in);
// with ghost planes they change only by the flux through them.  Each ghost
// value is paired with the boundary field its curl term updates (Q(0)), and
// with sigma = +1 for polarization 0 and -1 for polarization 1 the energy
// leaving through it in one step is
//   low face (ghost H_g^{n+1/2}, boundary E, W-form):
//       f = sigma H_g (Q(0)^n + Q(0)^{n+1}) / 2,
//       Q(0)^{n+1} = ca Q(0)^n + cb (R - sigma H_g);
//   high face (ghost E_g^n, boundary H, V-form):
//       f = sigma E_g (Q(0)^{n-1/2} + Q(0)^{n+1/2}) / 2,
//       Q(0)^{n+1/2} = da Q(0)^{n-1/2} - db (R + sigma E_g),
// R being the boundary cell's curl without the ghost term.  Both are
// f(g) = sigma B g - q g^2 / 2 with B = ((1 + ca) Q(0) + cb R) / 2, q = cb
// (low) or B = ((1 + da) Q(0) - db R) / 2, q = db (high): a concave quadratic
// in the new ghost value alone.  (A constraint on the time average of two
// successive ghost values, the W-form of the high face, would admit a ghost
// that alternates in sign every step, which no flux sees; summed over steps
// the V-form flux equals the W-form flux up to one instantaneous term.)
//
// The face keeps one energy credit C, the energy that has left through it so
// far (exact ghosts of a thick-CPML run never take it below zero, but a single
// position can: reactive energy leaves at one position and returns at
// another).  Each of the face's N = 2 x positions ghost values may draw C / N:
// a prediction with f(g) < -C / N is moved to the nearest value with
// f(g) = -C / N, inside the interval between the quadratic's roots, which
// always contains 0.  The face's total flux in the step is then at least -C,
// so C stays non-negative and the face never returns more energy than it has
// absorbed, whatever the network predicts.  Second, gamma > 0 bounds
// |ghost| <= gamma max_d |P(d)|.  The boundary cell on an edge shared with
// another limited face reads both faces' ghosts; such values get only the
// magnitude bound here and are limited in pairs by k_edge_limit.  Exact for
// faces without a CPML on the boundary cell (the learned domains have none).
// Returns the limited prediction on lanes 0..15, its flux (raw units) in
// `flux`, and the statistics: steps limited by the energy bound and by the
// magnitude bound, and the sum of squared corrections.
__device__ __forceinline__ float limit_ghost(const Geo &g, const Ptrs &es, const Ptrs &hs,
                                             const FaceHook &h, int pol, int pos, bool live,
                                             int lane, fv4 u, float s, float rs,
                                             const float *mem, float pred, float &flux,
                                             float *mem_stats)
{
    const int sl = lane & 15;
    const int a = h.axis, t1 = (a + 1) % 3, t2 = (a + 2) % 3;
    // every lane takes part in the shuffle; lanes 0..15 hold P(0..3) in u
    const float q0 = __shfl(u.x, sl + 16);  // Q(0)
    const float pm = fmaxf(fmaxf(fabsf(u.x), fabsf(u.y)), fmaxf(fabsf(u.z), fabsf(u.w)));
    flux = 0.f;
    mem_stats[0] = mem[kMemStats + 1];
    mem_stats[1] = mem[kMemStats + 2];
    mem_stats[2] = mem[kMemStats + 3];
    if (lane >= 16 || !live) return pred;
    const float sg = pol ? -1.0f : 1.0f;
    // this value's share of the credit, in scaled units (inf for a tiny s:
    // no bound; 0 stays 0)
    const float c0 = fmaxf(h.credit[0] - h.retain * h.credit[1], 0.f) / (float)(2 * h.nfast * h.nslow);
    const float cs = c0 > 0.f ? (c0 * rs) * rs : 0.f;
    const float inf = __builtin_inff();
    const int f = h.fast0 + pos % h.nfast, sc = h.slow0 + pos / h.nfast;
    const size_t cell = face_cell(g, a, h.b0, f, sc);
    // the normal component beside the boundary cell along the tangential
    // axis t of this polarization's curl: one cell back (low face, H) or
    // forward (high face, E)
    const int t = pol ? t2 : t1;
    const int ct = t == (t1 < t2 ? t1 : t2) ? f : sc;
    const int nt = t == 0 ? g.nx : (t == 1 ? g.ny : g.nz);
    const size_t st = t == 0 ? 1 : (t == 1 ? (size_t)g.nx : (size_t)g.nx * g.ny);
    const bool wrap = g.periodic >> t & 1u;
    if (!wrap && ct == (h.hi ? nt - 1 : 0) && (h.predicted >> (2 * t + h.hi) & 1u)) {
        // on an edge with another limited face: k_edge_limit takes the pair
        const float gl = h.gamma > 0.f ? fminf(fmaxf(pred, -h.gamma * pm), h.gamma * pm) : pred;
        mem_stats[1] += gl != pred ? 1.f : 0.f;
        return gl;
    }
    const float4 cf = es.coef[es.id[cell]];
    float B, cq;
    if (h.hi) {
        const float *Ea = field_ptr(es, a);
        const float e0 = Ea[cell];
        const float e1 = ct + 1 < nt ? Ea[cell + st] : (wrap ? Ea[cell - (size_t)(nt - 1) * st] : 0.f);
        const float d = (e1 - e0) * rs;
        const float R = pol ? d + u.x : -u.x - d;  // u.x = P(0) = E_t1 / E_t2 at the boundary
        B = 0.5f * ((1.0f + cf.z) * q0 - cf.w * R);
        cq = cf.w;
    } else {
        const float *Ha = field_ptr(hs, 3 + a);
        const float h0 = Ha[cell];
        const float hm = ct > 0 ? Ha[cell - st] : (wrap ? Ha[cell + (size_t)(nt - 1) * st] : 0.f);
        const float d = (h0 - hm) * rs;
        const float R = pol ? d - u.x : u.x - d;  // u.x = P(0) = H_t1 / H_t2 at the boundary
        B = 0.5f * ((1.0f + cf.x) * q0 + cf.y * R);
        cq = cf.y;
    }
    // f(g) = sg B g - cq g^2 / 2 >= -cs  <=>  g between the roots
    float lo = -inf, hi = inf;
    if (cq > 0.f) {
        const float disc = sqrtf(B * B + 2.0f * cq * cs);
        lo = (sg * B - disc) / cq;
        hi = (sg * B + disc) / cq;
    }
    float glo = lo, ghi = hi;
    if (h.gamma > 0.f) {
        glo = fmaxf(lo, -h.gamma * pm);
        ghi = fminf(hi, h.gamma * pm);
    }
    const float gl = glo <= ghi ? fminf(fmaxf(pred, glo), ghi) : 0.f;
    flux = (sg * B * gl - 0.5f * cq * gl * gl) * (s * s);
    const bool by_energy = pred < lo || pred > hi, by_mag = !by_energy && gl != pred;
    mem_stats[0] += by_energy ? 1.f : 0.f;
    mem_stats[1] += by_mag ? 1.f : 0.f;
    const float dg = (gl - pred) * s;
    mem_stats[2] += dg * dg;
    return gl;
}

// Inputs per sample (face position, polarization): P = the component whose
// ghost value the boundary cell's update reads, Q = its partner in the
// curl, d = depth from the boundary cell inward:
//   0..3   P(d)                     4..7  Q(d)
//   8, 9   zero (unused)
//   10, 11 P(1)(n-1), P(1)(n-2)                    (history)
//   12..15 g(n-1), g(n-2), P(0)(n-1), P(0)(n-2)   (history)
// High faces (before the H phase): P = E_t1 / E_t2, Q = H_t2 / H_t1, ghost
// E at b0+1.  Low faces (before the E phase): P = H_t1 / H_t2, Q = E_t2 /
// E_t1, ghost H at b0-1.  t1 = (a+1)%3, t2 = (a+2)%3.
// PAIRS: the network of the predicted faces takes bf16 pairs (its head);
// LIMITS: some face of the launch is limited.  Both are the same for the
// faces of one launch; each combination is its own kernel, so the register
// budget of the others does not hold a lighter one back.
template <bool PAIRS, bool LIMITS>
__device__ __forceinline__ void face_hook(const Geo &g, const Ptrs &es, const Ptrs &hs,
                                          const FaceHook &h, const uint64_t *__restrict__ step)
{
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    const int npos = h.nfast * h.nslow;
    const int gpp = (npos + 15) / 16;  // groups per polarization
    const int a = h.axis, t1 = (a + 1) % 3, t2 = (a + 2) % 3;
    const int q4 = lane >> 4, sl = lane & 15;
    const int dir = h.hi ? -1 : 1;

    // per polarization: the P and Q fields, chosen once (the switches stay
    // out of the loop); high faces: P = E_t1 / E_t2, Q = H_t2 / H_t1; low
    // faces: P = H_t1 / H_t2, Q = E_t2 / E_t1
    const int P0 = (h.hi ? 0 : 3) + t1, P1 = (h.hi ? 0 : 3) + t2;
    const int Q0 = (h.hi ? 3 : 0) + t2, Q1 = (h.hi ? 3 : 0) + t1;
    const float *const Pf_0 = h.hi ? field_ptr(es, P0) : field_ptr(hs, P0);
    const float *const Pf_1 = h.hi ? field_ptr(es, P1) : field_ptr(hs, P1);
    const float *const Qf_0 = h.hi ? field_ptr(hs, Q0) : field_ptr(es, Q0);
    const float *const Qf_1 = h.hi ? field_ptr(hs, Q1) : field_ptr(es, Q1);
    // step of one cell inward along the normal, in the field arrays
    const long sa = (long)dir * (a == 0 ? 1l : (a == 1 ? (long)g.nx : (long)g.nx * g.ny));

    MlpRegs r;
    if (h.mode == 0) load_regs(h.net, lane, r);
    long row = -1;
    if (h.mode == 1) {
        const uint64_t k = *step - h.emit_base;
        if (k % h.emit_every == 0 && k / h.emit_every < h.emit_rows) row = (long)(k / h.emit_every);
    }

    if (h.mode == 2 && !(LIMITS && h.passive)) {
        // replay: ghost values recorded by a thick-CPML run, step by step
        const uint64_t k = *step - h.emit_base;
        for (int grp = wave; grp < 2 * gpp; grp += nwaves) {
            const int pol = grp / gpp;
            const int pos = (grp - pol * gpp) * 16 + sl;
            if (lane < 16 && pos < npos) {
                const float v = k < h.emit_rows
                                    ? h.samples[((k * 2 + pol) * npos + pos) * kSampleStride + 17]
                                    : 0.f;
                (pol == 0 ? h.ghost[0] : h.ghost[1])[pos] = v;
            }
        }
        return;
    }

    // This lane's four inputs of group grp (zero past the face).
    auto gather = [&](int grp) __attribute__((always_inline)) -> fv4 {
        fv4 x = {0.f, 0.f, 0.f, 0.f};
        const int pol = grp / gpp;
        const int pos = (grp - pol * gpp) * 16 + sl;
        if (pos >= npos) return x;
        const int f = h.fast0 + pos % h.nfast, sc = h.slow0 + pos / h.nfast;
        if (q4 < 2) {
            // P or Q at depths 0..3
            const float *Pp = pol ? Pf_1 : Pf_0, *Qp = pol ? Qf_1 : Qf_0;  // uniform
            const float *F = q4 == 0 ? Pp : Qp;
            const long c0 = (long)face_cell(g, a, h.b0, f, sc);
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = F[c0 + j * sa];
        } else if (q4 == 2) {
            // inputs 8..11: zero, zero and the history of P(1)
            const float2 q = *reinterpret_cast<const float2 *>(
                h.mem + ((size_t)pol * npos + pos) * kMemStride + 4);
            x[2] = q.x;
            x[3] = q.y;
        } else {
            x = *reinterpret_cast<const fv4 *>(h.mem + ((size_t)pol * npos + pos) * kMemStride);
        }
        return x;
    };

    // Each wave takes groups wave, wave + nwaves, ...; the next group's
    // inputs are loaded before the current one is evaluated.
    float wflux = 0.f;  // this wave's flux through its ghosts in this step
    fv4 xn = wave < 2 * gpp ? gather(wave) : fv4{0.f, 0.f, 0.f, 0.f};
    for (int grp = wave; grp < 2 * gpp; grp += nwaves) {
        const fv4 x = xn;
        if (grp + nwaves < 2 * gpp) xn = gather(grp + nwaves);
        const int pol = grp / gpp;
        const int pos = (grp - pol * gpp) * 16 + sl;
        const bool live = pos < npos;
        float *mem = h.mem + ((size_t)pol * npos + (live ? pos : 0)) * kMemStride;
        // scale: the largest magnitude of the 16 inputs (no squares, so no
        // underflow for the tiny fields ahead of a wave)
        float s = fmaxf(fmaxf(fabsf(x.x), fabsf(x.y)), fmaxf(fabsf(x.z), fabsf(x.w)));
        s = fmaxf(s, __shfl_xor(s, 16));
        s = fmaxf(s, __shfl_xor(s, 32));
        // at least 2^-100 (a normal number whose reciprocal is finite and
        // exact): the first values of a wave front reaching the face can be
        // denormal, and their reciprocal would overflow
        s = fmaxf(s, 0x1p-100f);
        const float rs = 1.0f / s;
        const fv4 u = x * rs;
        // history on lanes 0..15: their own P(0) and the memory lanes' values
        const float g1 = __shfl(x.x, sl + 48), p1 = __shfl(x.z, sl + 48);
        const float q1 = __shfl(x.z, sl + 32);  // P1(n-1)
        float gnew, mem_stats[3] = {0.f, 0.f, 0.f}, flux = 0.f, dk = 0.f;
        if (h.mode == 0) {
            Acts t;
            const fv4 ui = mlp_in(r.head, u, sl);
            float sk;
            bfv4 ub;
            const float m = mlp_forward(r, ui, skip_in(r.head, u, ui), t, sk, ub, PAIRS);
            const HeadIn hv = head_in(u, sl);
            float pred = head_out(r.head, m, sk, hv, r.hf);
            if (r.head == kHeadHigdon && r.hf.beta > 0.f && r.hf.beta < 1.f) {
                // the adaptive coefficient's departure from its base
                // tanh(b3), smoothed in time at this position
                float A, B;
                higdon_ab(hv, r.hf, A, B);
                const float kb = tanh_f(r.b3);
                const float ki = tanh_f(m);
                dk = (1.0f - r.hf.beta) * mem[6] + r.hf.beta * (ki - kb);
                pred = A + (kb + dk) * B;
            }
            if (LIMITS && h.passive) pred = limit_ghost(g, es, hs, h, pol, pos, live, lane, u, s, rs, mem, pred, flux,
                                               mem_stats);
            gnew = s * pred;
            if (lane < 16 && live) (pol == 0 ? h.ghost[0] : h.ghost[1])[pos] = gnew;
        } else if (h.mode == 2) {
            // replay through the limits: the recorded ghost in place of the
            // network's prediction (a check that the exact ghosts of a
            // thick-CPML run pass the limits untouched)
            const uint64_t k = *step - h.emit_base;
            float pred = 0.f;
            if (live && k < h.emit_rows)
                pred = h.samples[((k * 2 + pol) * npos + pos) * kSampleStride + 17] * rs;
            pred = limit_ghost(g, es, hs, h, pol, pos, live, lane, u, s, rs, mem, pred, flux,
                               mem_stats);
            gnew = s * pred;
            if (lane < 16 && live) (pol == 0 ? h.ghost[0] : h.ghost[1])[pos] = gnew;
        } else {
            gnew = 0.f;
            if (live) {
                const int f = h.fast0 + pos % h.nfast, sc = h.slow0 + pos / h.nfast;
                gnew = (pol ? Pf_1 : Pf_0)[(long)face_cell(g, a, h.b0, f, sc) - sa];
            }
            if (row >= 0 && live) {
                float *o = h.samples + (((size_t)row * 2 + pol) * npos + pos) * kSampleStride;
                *reinterpret_cast<fv4 *>(o + 4 * q4) = u;  // 80-byte records: 16-byte aligned
                if (lane < 16) {
                    o[16] = gnew * rs;
                    o[17] = gnew;  // raw ghost value (replay)
                    o[18] = s;
                }
            }
        }
        wflux += sum16(flux);  // the 16 samples' flux, summed in lane order
        if (lane < 16 && live) {
            *reinterpret_cast<fv4 *>(mem) = fv4{gnew, g1, x.x, p1};
            *reinterpret_cast<fv4 *>(mem + 4) = fv4{x.y, q1, dk, 0.f};
            if (LIMITS && h.passive && h.mode != 1)
                *reinterpret_cast<fv4 *>(mem + kMemStats) =
                    fv4{flux, mem_stats[0], mem_stats[1], mem_stats[2]};
        }
    }
    if (LIMITS && h.passive && h.mode != 1) {
        // the workgroup's flux partial, its four waves summed in order
        __shared__ float red[4];
        if (lane == 0) red[threadIdx.x >> 6] = wflux;
        __syncthreads();
        if (threadIdx.x == 0) h.parts[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// Edges shared by two limited faces of one side.  The boundary cell on such
// an edge (its component along the edge, c) reads one ghost of each face:
// face A on axis a = k (polarization 0) and face B on axis b = k + 1
// (polarization 1), c = k + 2.  Its update holds both as v = g_A - g_B (low
// faces: E_c^{n+1} = ca E_c + cb (R0 - v), R0 = H_b - H_a at the cell; high
// faces: H_c^{n+1/2} = da H_c - db (R0 + v), R0 = E_a - E_b), so the energy
// leaving through the pair is f(v) = B v - q v^2 / 2 with B and q as for one
// ghost in limit_ghost, and only v matters: v is bounded by the pair's share
// of the two credits (C_A / N_A + C_B / N_B), the two predictions move by
// equal and opposite amounts (the smallest change that reaches the bound),
// and half of f goes to each face's flux.  One thread per edge cell,
// blockIdx.y = k; each workgroup writes its flux partial for face A (slot
// kFluxParts + blockIdx.x) and face B (slot kFluxParts + kEdgeParts +
// blockIdx.x).  Launched only for families whose two faces are limited.
__global__ __launch_bounds__(256) void k_edge_limit(Geo g, Ptrs es, Ptrs hs, FaceHook hx,
                                                    FaceHook hy, FaceHook hz, int hi)
{
    const int k = blockIdx.y;
    const int a = k, b = (k + 1) % 3, c = (k + 2) % 3;
    const FaceHook &A = a == 0 ? hx : (a == 1 ? hy : hz);
    const FaceHook &B = b == 0 ? hx : (b == 1 ? hy : hz);
    if (A.mode < 0 || A.mode == 1 || B.mode < 0 || B.mode == 1 || !A.passive || !B.passive) return;
    const int n[3] = {g.nx, g.ny, g.nz};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float fr = 0.f;
    if (i < n[c]) {
        int xyz[3];
        xyz[a] = A.b0;
        xyz[b] = B.b0;
        xyz[c] = i;
        const size_t cell = ((size_t)xyz[2] * g.ny + xyz[1]) * g.nx + xyz[0];
        auto pos_on = [&](const FaceHook &F, int ax) {
            const int u1 = (ax + 1) % 3, u2 = (ax + 2) % 3;
            const int fa = u1 < u2 ? u1 : u2, sa = u1 < u2 ? u2 : u1;
            return (xyz[fa] - F.fast0) + F.nfast * (xyz[sa] - F.slow0);
        };
        const int pA = pos_on(A, a), pB = pos_on(B, b);
        float *mA = A.mem + (size_t)pA * kMemStride;                         // polarization 0
        float *mB = B.mem + ((size_t)B.nfast * B.nslow + pB) * kMemStride;  // polarization 1
        const float gA = A.ghost[0][pA], gB = B.ghost[1][pB];
        const float4 cf = es.coef[es.id[cell]];
        float q, f1, f2, uc, uq;
        if (hi) {
            q = field_ptr(hs, 3 + c)[cell];
            f1 = field_ptr(es, a)[cell];
            f2 = field_ptr(es, b)[cell];
            uc = cf.z;
            uq = cf.w;
        } else {
            q = field_ptr(es, c)[cell];
            f1 = field_ptr(hs, 3 + b)[cell];
            f2 = field_ptr(hs, 3 + a)[cell];
            uc = cf.x;
            uq = cf.y;
        }
        float s = fmaxf(fmaxf(fabsf(gA), fabsf(gB)), fmaxf(fabsf(q), fmaxf(fabsf(f1), fabsf(f2))));
        s = fmaxf(s, 0x1p-100f);
        const float rs = 1.0f / s;
        const float r0 = (f1 - f2) * rs;
        const float Bv = hi ? 0.5f * ((1.0f + uc) * (q * rs) - uq * r0)
                            : 0.5f * ((1.0f + uc) * (q * rs) + uq * r0);
        const float c0 = fmaxf(A.credit[0] - A.retain * A.credit[1], 0.f) / (float)(2 * A.nfast * A.nslow) +
                         fmaxf(B.credit[0] - B.retain * B.credit[1], 0.f) / (float)(2 * B.nfast * B.nslow);
        const float cs = c0 > 0.f ? (c0 * rs) * rs : 0.f;
        const float v = (gA - gB) * rs;
        float vt = v;
        if (uq > 0.f) {
            const float disc = sqrtf(Bv * Bv + 2.0f * uq * cs);
            vt = fminf(fmaxf(v, (Bv - disc) / uq), (Bv + disc) / uq);
        }
        const float d = 0.5f * (vt - v) * s;
        const float nA = gA + d, nB = gB - d;
        fr = (Bv * vt - 0.5f * uq * vt * vt) * (s * s);
        A.ghost[0][pA] = nA;
        B.ghost[1][pB] = nB;
        mA[0] = nA;
        mB[0] = nB;
        mA[kMemStats] = 0.5f * fr;
        mB[kMemStats] = 0.5f * fr;
        if (vt != v) {
            mA[kMemStats + 1] += 1.f;
            mB[kMemStats + 1] += 1.f;
            mA[kMemStats + 3] += d * d;
            mB[kMemStats + 3] += d * d;
        }
    }
    // the workgroup's flux, summed in a fixed order (waves, then lanes)
    __shared__ float red[4];
    const float w = wave_sum(fr);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = ((red[0] + red[1]) + red[2]) + red[3];
        A.parts[kFluxParts + blockIdx.x] = 0.5f * t;
        B.parts[kFluxParts + kEdgeParts + blockIdx.x] = 0.5f * t;
    }
}

// The faces' credits after a step: C = max(C + F, 0) and A = A + max(F, 0),
// F = the step's flux (its partials summed in a fixed order: lanes stride
// over them, then a fixed butterfly, so the result does not depend on
// timing).  The limits allow a step's flux down to -(C - retain A), so
// C - retain A stays >= 0: at most (1 - retain) of the absorbed energy can
// come back.
// One wave per face (blockIdx.x picks the hook); nparts hook partials, and
// the edge partials of the families the face took part in (na as A, nb as B
// workgroups; 0 = none).
struct CreditJob {
    float *credit, *parts;
    int na, nb;
};
__global__ __launch_bounds__(64) void k_credit_reduce(CreditJob j0, CreditJob j1, CreditJob j2,
                                                       int nparts)
{
    const CreditJob j = blockIdx.x == 0 ? j0 : (blockIdx.x == 1 ? j1 : j2);
    if (!j.credit) return;
    const int lane = threadIdx.x;
    float t = 0.f;
    for (int i = lane; i < nparts; i += 64) t += j.parts[i];
    for (int i = lane; i < j.na; i += 64) t += j.parts[kFluxParts + i];
    for (int i = lane; i < j.nb; i += 64) t += j.parts[kFluxParts + kEdgeParts + i];
    t = wave_sum(t);
    if (lane == 0) {
        j.credit[0] = fmaxf(j.credit[0] + t, 0.f);  // net energy out so far
        j.credit[1] = j.credit[1] + fmaxf(t, 0.f);  // energy absorbed in steps of net outflow
    }
}

// Up to three faces of one side in one launch: blockIdx.y picks the face.
template <bool PAIRS, bool LIMITS>
__global__ __launch_bounds__(256) void k_face_hooks(Geo g, Ptrs es, Ptrs hs, FaceHook h0,
                                                    FaceHook h1, FaceHook h2,
                                                    const uint64_t *__restrict__ step)
{
    const int k = blockIdx.y;
    const FaceHook h = k == 0 ? h0 : (k == 1 ? h1 : h2);
    face_hook<PAIRS, LIMITS>(g, es, hs, h, step);
}

}  // namespace

hipError_t pack_mlp(const float *params, int head, MlpPacked *net, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, st, params, head, net);
    return hipGetLastError();
}

hipError_t launch_face_hooks(const Geo &g, const Ptrs &esrc, const Ptrs &hsrc,
                             const FaceHook *const *hooks, int n, const uint64_t *step,
                             hipStream_t st)
{
    if (n < 1) return hipSuccess;
    if (n > 3) return hipErrorInvalidValue;
    // One resident wave set: the kernel holds 3 waves per SIMD (about 130
    // VGPRs), i.e. three 4-wave blocks per CU, shared by the n faces; every
    // wave then loads the network once and strides over many groups.
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (!cus[dev] &&
        hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return hipErrorInvalidDevice;
    const int cap = std::min(kFluxParts, std::max(1, 3 * cus[dev] / n));
    int blocks = 1;
    for (int k = 0; k < n; ++k) {
        const int npos = hooks[k]->nfast * hooks[k]->nslow;
        const int groups = 2 * ((npos + 15) / 16);
        blocks = std::max(blocks, std::min((groups + 3) / 4, cap));
    }
    const FaceHook &h0 = *hooks[0], &h1 = *hooks[n > 1 ? 1 : 0], &h2 = *hooks[n > 2 ? 2 : 0];
    bool pairs = false, limits = false;
    for (int k = 0; k < n; ++k) {
        const FaceHook &h = *hooks[k];
        pairs = pairs || (h.mode == 0 && h.head != kHeadOneWay && h.head != kHeadHigdon);
        limits = limits || (h.mode != 1 && h.passive);
    }
    const dim3 grid(blocks, n), block(256);
    if (pairs && limits)
        hipLaunchKernelGGL((k_face_hooks<true, true>), grid, block, 0, st, g, esrc, hsrc, h0, h1, h2, step);
    else if (pairs)
        hipLaunchKernelGGL((k_face_hooks<true, false>), grid, block, 0, st, g, esrc, hsrc, h0, h1, h2, step);
    else if (limits)
        hipLaunchKernelGGL((k_face_hooks<false, true>), grid, block, 0, st, g, esrc, hsrc, h0, h1, h2, step);
    else
        hipLaunchKernelGGL((k_face_hooks<false, false>), grid, block, 0, st, g, esrc, hsrc, h0, h1, h2, step);
    // limited faces: the edges shared by two of them, then the credits
    FaceHook ax[3];
    bool lim[3] = {false, false, false};
    for (int k = 0; k < 3; ++k) ax[k].mode = -1;
    for (int k = 0; k < n; ++k) {
        const FaceHook &h = *hooks[k];
        ax[h.axis] = h;
        lim[h.axis] = h.mode != 1 && h.passive;
    }
    if (!lim[0] && !lim[1] && !lim[2]) return hipGetLastError();
    const int nd[3] = {g.nx, g.ny, g.nz};
    int eb[3] = {0, 0, 0};  // edge workgroups of family k (faces on axes k and k + 1)
    for (int k = 0; k < 3; ++k)
        if (lim[k] && lim[(k + 1) % 3]) eb[k] = (nd[(k + 2) % 3] + 255) / 256;
    const int emax = std::max(eb[0], std::max(eb[1], eb[2]));
    if (emax > kEdgeParts) return hipErrorInvalidValue;
    if (emax)
        hipLaunchKernelGGL(k_edge_limit, dim3(emax, 3), dim3(256), 0, st, g, esrc, hsrc, ax[0],
                           ax[1], ax[2], hooks[0]->hi);
    CreditJob job[3] = {};
    for (int k = 0; k < n; ++k) {
        const FaceHook &h = *hooks[k];
        if (!lim[h.axis]) continue;
        job[k] = CreditJob{h.credit, h.parts, eb[h.axis], eb[(h.axis + 2) % 3]};
    }
    hipLaunchKernelGGL(k_credit_reduce, dim3(n), dim3(64), 0, st, job[0], job[1], job[2], blocks);
    return hipGetLastError();
}

}  // namespace wf

using namespace wf;

#define WF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(WF_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

// ===========================================================================
// Trainer
// ===========================================================================
struct wf_mlp_trainer {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0;
    float *d_u = nullptr, *d_t = nullptr;
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr, *d_part = nullptr;
    float *d_part2 = nullptr;  // per-chunk sums of d_part's rows
    MlpPacked *d_w = nullptr;
    int blocks = 0;
    int head = kHeadDirect;
    uint64_t adam_t = 0;
};

namespace {

constexpr int kTrainBlocks = 256;  // 1024 waves
constexpr int kReduceChunks = 32;  // first reduction stage: 32 rows per chunk

template <typename T>
int dalloc(T **p, size_t bytes)
{
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(p), bytes) != hipSuccess) {
        *p = nullptr;
        return fail(WF_ERR_OOM, "hipMalloc(%zu bytes) failed", bytes);
    }
    return WF_OK;
}

int trainer_grad(wf_mlp_trainer *t)
{
    hipLaunchKernelGGL(k_mlp_train, dim3(t->blocks), dim3(256), 0, t->stream, t->d_u, t->d_t,
                       (int)t->n, 1.0f / (float)t->n, t->d_w, t->d_part);
    const int rows = t->blocks * 4, per = (rows + kReduceChunks - 1) / kReduceChunks;
    const unsigned cols = (kMlpParams + 1 + 255) / 256;
    hipLaunchKernelGGL(k_mlp_reduce, dim3(cols, kReduceChunks), dim3(256), 0, t->stream, t->d_part,
                       rows, per, t->d_part2);
    hipLaunchKernelGGL(k_mlp_reduce, dim3(cols, 1), dim3(256), 0, t->stream, t->d_part2,
                       kReduceChunks, kReduceChunks, t->d_grad);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

}  // namespace

extern "C" {

uint32_t wf_mlp_param_count(void) { return (uint32_t)kMlpParams; }

void wf_mlp_trainer_destroy(wf_mlp_trainer *t)
{
    if (!t) return;
    (void)hipSetDevice(t->device);
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    for (float *p : {t->d_u, t->d_t, t->d_p, t->d_m, t->d_v, t->d_grad, t->d_part, t->d_part2})
        if (p) (void)hipFree(p);
    if (t->d_w) (void)hipFree(t->d_w);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int wf_mlp_trainer_create(int device, uint64_t n, const float *inputs, const float *targets,
                          wf_mlp_trainer **out)
{
    clear_error();
    if (!out || !inputs || !targets || n == 0) return fail(WF_ERR_INVALID, "wf_mlp_trainer_create: bad arguments");
    if (n > 0x7fffffffull) return fail(WF_ERR_INVALID, "wf_mlp_trainer_create: too many samples");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(WF_ERR_HIP, "no HIP device %d available", device);
    wf_mlp_trainer *t = new wf_mlp_trainer;
    t->device = device;
    t->n = n;
    t->blocks = kTrainBlocks;
    int rc = WF_OK;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        rc = fail(WF_ERR_HIP, "wf_mlp_trainer_create: stream");
    if (!rc) rc = dalloc(&t->d_u, n * kMlpIn * sizeof(float));
    if (!rc) rc = dalloc(&t->d_t, n * sizeof(float));
    if (!rc) rc = dalloc(&t->d_p, kMlpParams * sizeof(float));
    if (!rc) rc = dalloc(&t->d_m, kMlpParams * sizeof(float));
    if (!rc) rc = dalloc(&t->d_v, kMlpParams * sizeof(float));
    if (!rc) rc = dalloc(&t->d_grad, (kMlpParams + 1) * sizeof(float));
    if (!rc) rc = dalloc(&t->d_part, (size_t)t->blocks * 4 * (kMlpParams + 1) * sizeof(float));
    if (!rc) rc = dalloc(&t->d_part2, (size_t)kReduceChunks * (kMlpParams + 1) * sizeof(float));
    if (!rc) rc = dalloc(&t->d_w, sizeof(MlpPacked));
    if (!rc && (hipMemcpy(t->d_u, inputs, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(t->d_t, targets, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(t->d_p, 0, kMlpParams * sizeof(float)) != hipSuccess ||
                hipMemset(t->d_m, 0, kMlpParams * sizeof(float)) != hipSuccess ||
                hipMemset(t->d_v, 0, kMlpParams * sizeof(float)) != hipSuccess))
        rc = fail(WF_ERR_HIP, "wf_mlp_trainer_create: upload failed");
    if (!rc) {
        hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, t->stream, t->d_p, t->head, t->d_w);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(t->stream) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_mlp_trainer_create: pack failed");
    }
    if (rc) {
        wf_mlp_trainer_destroy(t);
        return rc;
    }
    *out = t;
    return WF_OK;
}

int wf_mlp_set_params(wf_mlp_trainer *t, const float *params)
{
    clear_error();
    if (!t || !params) return fail(WF_ERR_INVALID, "wf_mlp_set_params: null pointer");
    WF_HIP(hipSetDevice(t->device));
    WF_HIP(hipMemcpy(t->d_p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    WF_HIP(hipMemset(t->d_m, 0, kMlpParams * sizeof(float)));
    WF_HIP(hipMemset(t->d_v, 0, kMlpParams * sizeof(float)));
    t->adam_t = 0;
    hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, t->stream, t->d_p, t->head, t->d_w);
    WF_HIP(hipGetLastError());
    WF_HIP(hipStreamSynchronize(t->stream));
    return WF_OK;
}

int wf_mlp_set_head(wf_mlp_trainer *t, int head)
{
    clear_error();
    if (!t) return fail(WF_ERR_INVALID, "wf_mlp_set_head: null trainer");
    if (!valid_head(head))
        return fail(WF_ERR_INVALID, "wf_mlp_set_head: unknown head %d", head);
    WF_HIP(hipSetDevice(t->device));
    t->head = head;
    hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, t->stream, t->d_p, t->head, t->d_w);
    WF_HIP(hipGetLastError());
    WF_HIP(hipStreamSynchronize(t->stream));
    return WF_OK;
}

int wf_mlp_get_params(wf_mlp_trainer *t, float *params)
{
    clear_error();
    if (!t || !params) return fail(WF_ERR_INVALID, "wf_mlp_get_params: null pointer");
    WF_HIP(hipSetDevice(t->device));
    WF_HIP(hipStreamSynchronize(t->stream));
    WF_HIP(hipMemcpy(params, t->d_p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return WF_OK;
}

int wf_mlp_gradient(wf_mlp_trainer *t, float *grad, double *loss)
{
    clear_error();
    if (!t || !grad) return fail(WF_ERR_INVALID, "wf_mlp_gradient: null pointer");
    WF_HIP(hipSetDevice(t->device));
    if (int rc = trainer_grad(t)) return rc;
    std::vector<float> g(kMlpParams + 1);
    WF_HIP(hipMemcpyAsync(g.data(), t->d_grad, (kMlpParams + 1) * sizeof(float), hipMemcpyDeviceToHost,
                          t->stream));
    WF_HIP(hipStreamSynchronize(t->stream));
    std::memcpy(grad, g.data(), kMlpParams * sizeof(float));
    if (loss) *loss = (double)g[kMlpParams] / (double)t->n;
    return WF_OK;
}

int wf_mlp_train(wf_mlp_trainer *t, uint32_t iters, float lr, float lr_skip, int odd, double *loss)
{
    clear_error();
    if (!t) return fail(WF_ERR_INVALID, "wf_mlp_train: null trainer");
    WF_HIP(hipSetDevice(t->device));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t i = 0; i < iters; ++i) {
        if (int rc = trainer_grad(t)) return rc;
        t->adam_t += 1;
        const float bc1 = 1.0f - std::pow(b1, (float)t->adam_t);
        const float bc2 = 1.0f - std::pow(b2, (float)t->adam_t);
        hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, t->stream, t->d_p,
                           t->d_m, t->d_v, t->d_grad, lr, lr_skip, odd, b1, b2, eps, bc1, bc2);
        hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, t->stream, t->d_p, t->head, t->d_w);
        WF_HIP(hipGetLastError());
    }
    if (loss) {
        // loss at the final parameters
        if (int rc = trainer_grad(t)) return rc;
        float l = 0.f;
        WF_HIP(hipMemcpyAsync(&l, t->d_grad + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, t->stream));
        WF_HIP(hipStreamSynchronize(t->stream));
        *loss = (double)l / (double)t->n;
    }
    WF_HIP(hipStreamSynchronize(t->stream));
    return WF_OK;
}

int wf_mlp_predict(wf_mlp_trainer *t, uint64_t n, const float *inputs, float *outputs)
{
    clear_error();
    if (!t || !inputs || !outputs) return fail(WF_ERR_INVALID, "wf_mlp_predict: null pointer");
    if (n == 0) return WF_OK;
    if (n > 0x7fffffffull) return fail(WF_ERR_INVALID, "wf_mlp_predict: too many samples");
    WF_HIP(hipSetDevice(t->device));
    float *du = nullptr, *dy = nullptr;
    if (int rc = dalloc(&du, n * kMlpIn * sizeof(float))) return rc;
    if (int rc = dalloc(&dy, n * sizeof(float))) {
        (void)hipFree(du);
        return rc;
    }
    int rc = WF_OK;
    if (hipMemcpy(du, inputs, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(WF_ERR_HIP, "wf_mlp_predict: upload");
    if (!rc) {
        const uint64_t groups = (n + 15) / 16;
        unsigned blocks = (unsigned)std::min<uint64_t>((groups + 3) / 4, 1024);
        hipLaunchKernelGGL(k_mlp_predict, dim3(blocks), dim3(256), 0, t->stream, du, (int)n, t->d_w, dy);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(t->stream) != hipSuccess ||
            hipMemcpy(outputs, dy, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_mlp_predict: kernel failed");
    }
    (void)hipFree(du);
    (void)hipFree(dy);
    return rc;
}

}  // extern "C"

// ===========================================================================
// Allocation-free evaluation of a face network
// ===========================================================================
// wf_mlp_predict_noalloc stages the parameters, the packed network, the
// inputs and the outputs in one static device array instead of device
// memory of its own: the same k_pack_mlp and k_mlp_predict as wf_mlp_predict,
// for callers that may not allocate.  The array's layout, in bytes:
//   [0, kScrNet)            FP32 parameters (kMlpParams floats)
//   [kScrNet, kScrIn)       MlpPacked network
//   [kScrIn, kScrOut)       inputs, kNoAllocMaxSamples x 16 floats
//   [kScrOut, kScratchBytes) outputs, kNoAllocMaxSamples floats
// Each device holds its own copy of the array until the process ends.
namespace {

constexpr uint64_t kNoAllocMaxSamples = 65536;
constexpr size_t round256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kScrNet = round256(kMlpParams * sizeof(float));
constexpr size_t kScrIn = kScrNet + round256(sizeof(MlpPacked));
constexpr size_t kScrOut = kScrIn + kNoAllocMaxSamples * kMlpIn * sizeof(float);
constexpr size_t kScratchBytes = kScrOut + kNoAllocMaxSamples * sizeof(float);

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScrMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kScrMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once per
// device.  In host code the array's name is only a stand-in: kernels and
// copies get this address.
int scratch_on(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(WF_ERR_HIP, "wf_mlp_predict_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s",
                        device, hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[device];
    return WF_OK;
}

}  // namespace

extern "C" {

uint64_t wf_mlp_predict_noalloc_max(void) { return kNoAllocMaxSamples; }

int wf_mlp_predict_noalloc(int device, const float *params, int head, uint64_t n,
                           const float *inputs, float *outputs)
{
    clear_error();
    if (!params || (n && (!inputs || !outputs)))
        return fail(WF_ERR_INVALID, "wf_mlp_predict_noalloc: null pointer");
    if (!valid_head(head))
        return fail(WF_ERR_INVALID, "wf_mlp_predict_noalloc: unknown head %d", head);
    if (n > kNoAllocMaxSamples)
        return fail(WF_ERR_INVALID,
                    "wf_mlp_predict_noalloc: %llu samples exceed the static scratch's maximum of %llu",
                    (unsigned long long)n, (unsigned long long)kNoAllocMaxSamples);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev ||
        device >= kScrMaxDevices)
        return fail(WF_ERR_HIP, "no HIP device %d available", device);
    WF_HIP(hipSetDevice(device));
    unsigned char *base = nullptr;
    if (int rc = scratch_on(device, &base)) return rc;
    float *d_p = reinterpret_cast<float *>(base);
    MlpPacked *d_w = reinterpret_cast<MlpPacked *>(base + kScrNet);
    float *d_u = reinterpret_cast<float *>(base + kScrIn);
    float *d_y = reinterpret_cast<float *>(base + kScrOut);
    // every byte the kernels read is written first in this call: parameters,
    // packed network (all fields k_mlp_predict loads) and the n input rows
    WF_HIP(hipMemcpyAsync(d_p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, nullptr));
    hipLaunchKernelGGL(k_pack_mlp, dim3(1), dim3(64), 0, nullptr, d_p, head, d_w);
    WF_HIP(hipGetLastError());
    if (n) {
        WF_HIP(hipMemcpyAsync(d_u, inputs, n * kMlpIn * sizeof(float), hipMemcpyHostToDevice, nullptr));
        const uint64_t groups = (n + 15) / 16;
        const unsigned blocks = (unsigned)std::min<uint64_t>((groups + 3) / 4, 1024);
        hipLaunchKernelGGL(k_mlp_predict, dim3(blocks), dim3(256), 0, nullptr, d_u, (int)n, d_w, d_y);
        WF_HIP(hipGetLastError());
        WF_HIP(hipMemcpyAsync(outputs, d_y, n * sizeof(float), hipMemcpyDeviceToHost, nullptr));
    }
    WF_HIP(hipStreamSynchronize(nullptr));
    return WF_OK;
}

}  // extern "C"
