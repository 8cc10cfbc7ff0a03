// Learned thin absorbing boundary: device structures shared by the engine
// and the MLP kernels.
//
// A face of the domain is terminated by a few-cell layer (the first cells
// of a thick CPML profile) plus a learned outer boundary condition: the
// tangential fields just outside the domain (the "ghost" plane the boundary
// cell's curl update reads, zero in the reference's PEC-backed boundary,
// src/engine.rs:83) are predicted each half step by a small per-face MLP
// from the 4-cell-deep E/H/psi stencil at the face and a short history.
//
// MLP: 16 inputs -> 64 -> 64 -> 1 with hardtanh, BF16 weights and activations on
// MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, plus an FP32
// linear skip from the inputs.  Inputs are scaled by their largest
// magnitude s and the output is scaled back by s, so the predictor is
// homogeneous of degree 1 like the field equations it stands in for.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wf_kernels.h"

namespace wf {

constexpr int kMlpIn = 16;
constexpr int kMlpHid = 64;
// FP32 parameter vector: W1 [16][64] | b1 [64] | W2 [64][64] | b2 [64] |
// W3 [64] | b3 | wlin [16]   (W[in][out], row-major)
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid;
constexpr int kOffWL = kOffB3 + 1;
constexpr int kMlpParams = kOffWL + kMlpIn;  // 5329
// A recorded sample: 16 scaled inputs, target / s, target, s, 1 float padding.
constexpr int kSampleStride = 20;
// Floats per (polarization, position) of a face's memory: the history
// g(n-1), g(n-2), P0(n-1), P0(n-2), P1(n-1), P1(n-2), two unused, then from
// kMemStats the last step's flux through the ghost, the steps limited by the
// energy bound and by the magnitude bound, and the sum of squared
// corrections.
constexpr int kMemStride = 12;
constexpr int kMemStats = 8;
// Flux partials per face and step: one per hook workgroup (at most
// kFluxParts), then kEdgeParts per edge family the face belongs to (two).
constexpr int kFluxParts = 1024;
constexpr int kEdgeParts = 64;
constexpr int kPartsFloats = kFluxParts + 2 * kEdgeParts;

typedef short bfv4 __attribute__((ext_vector_type(4)));  // 4 x bf16 (raw bits)
typedef float fv4 __attribute__((ext_vector_type(4)));

// Weights pre-arranged as per-lane MFMA operands (bf16), plus FP32 biases.
struct MlpPacked {
    bfv4 a1[4][64];      // layer 1, A = W1^T block o: lane holds W1[4(l/16)+j][16o + l%16]
    bfv4 a2[4][4][64];   // layer 2, A = W2^T block (o, k)
    bfv4 a3[4][64];      // layer 3, A = W3^T (row 0 only) block k
    bfv4 a2b[4][4][64];  // backward, A = W2 block (i, o): W2[16i + l%16][16o + 4(l/16)+j]
    float b1[kMlpHid], b2[kMlpHid];
    float w3r[kMlpHid];  // W3 rounded to bf16 (the values the forward uses)
    float wlin[kMlpIn];
    float b3;
    int head;  // output head (kHead*)
    float pad[2];
};

// Output heads.  Direct: the network output is the ghost value / s.
// One-way: the output y sets the coefficient of a first-order one-way
// boundary update, ghost / s = P0(n-1) + tanh(y) (P0(n) - g(n-1)) (in scaled
// inputs u14 + tanh(y) (u0 - u12)); |tanh y| < 1 keeps the recursion on the
// ghost contractive.  The network then sees sign-normalized inputs (sign of
// u0, or of u4 when u0 = 0), so y is even and the ghost odd in the fields.
// Delta: the linear skip plus the MLP's output times the boundary field's
// step, ghost / s = skip(u) + m(sigma u) (P0(n) - P0(n-1)) (scaled: u0 -
// u14), with the MLP on sign-normalized inputs (even, so the product is odd
// in the fields).  A static field gets the skip alone: the network cannot
// add a response at zero frequency.
// Difference: ghost / s = P0(n) + skip(d) + m(d), where network and skip
// see differences that vanish for a field constant in space and time:
//   d0..2 = P(0)-P(1), P(1)-P(2), P(2)-P(3);  d4..6 = the same of Q;
//   d8, d9 = inputs 8, 9;  d10, d11 = P1(n-1)-P1, P1(n-2)-P1(n-1);
//   d12..15 = g(n-1)-P0, g(n-2)-g(n-1), P0(n-1)-P0, P0(n-2)-P0(n-1)
// (d3 = d7 = 0).  A static uniform field is held exactly (ghost = P0)
// whatever the weights, so rounding them cannot move the recursion's gain
// at zero frequency above one.
constexpr int kHeadDirect = 0;
constexpr int kHeadOneWay = 1;
constexpr int kHeadDelta = 2;
// Higdon: the second-order Higdon update, the product of two first-order
// one-way factors B(k, g) = (1 + k g T) - X (g T + k) (T one step back in
// time, X one cell inward, g <= 1 a damping of the factor's memory): a
// fixed factor B(kf, gf) and an adaptive one B(ka, ga) with ka = tanh(y),
// y the network's output on sign-normalized inputs (even, so the ghost is
// odd in the fields); kf, gf, ga = wlin[0..2], the other skip weights
// unused but wlin[3] = beta: in the face hooks, 0 < beta < 1 smooths the
// adaptive coefficient in time per face position, ka(n) = kb + d(n),
// d(n) = (1 - beta) d(n-1) + beta (tanh(y) - kb), kb = tanh(b3) (a
// coefficient that follows the fields slowly cannot pump energy into the
// grid at their frequencies).  Solved for the ghost, ghost / s = A + ka B with
//   A = kf (u0 - gf u12 - ga u10) + (gf + ga) u14 + gf ga (kf u15 - u11)
//   B = u0 - ga u12 - kf u1 + kf (ga + gf) u14 + gf (ga (u15 - kf u13) - u10)
// (u0 = P0, u1 = P1, u10, u11 = P1(n-1), P1(n-2), u12, u13 = g(n-1), g(n-2),
// u14, u15 = P0(n-1), P0(n-2)).
constexpr int kHeadDiff = 3;
constexpr int kHeadHigdon = 4;
constexpr bool valid_head(int h) { return h >= kHeadDirect && h <= kHeadHigdon; }

// One face hook: a face of the domain whose ghost plane is predicted
// (mode 0), or a virtual face inside a thick-CPML run whose stencils and true
// ghost values are recorded as training samples (mode 1).
struct FaceHook {
    int axis, hi;       // normal axis, 1 = high face
    int b0;             // normal coordinate of the depth-0 (boundary) cell
    int fast0, nfast;   // range on the lower-numbered tangential axis
    int slow0, nslow;   // range on the higher-numbered tangential axis
    int na, nf, ns;     // axis lengths: normal, fast, slow
    int mode;           // 0 predict, 1 record, 2 replay recorded ghosts
    int head;           // mode 0: the network's output head (kHead*)
    int passive;        // modes 0, 2: limit the ghosts so the face stays passive
    unsigned predicted; // faces (bit f) whose ghost planes are predicted or replayed
    float *credit;      // passive: the face's energy credit (one float)
    float *parts;       // passive: per-step flux partials, kFluxParts from the hook
                        // blocks, then 2 x kEdgeParts from the edge blocks
    float gamma;        // passive: |ghost| <= gamma max_d |P(d)| (0 = no bound)
    float retain;       // passive: net outflow >= retain x the absorbed energy (0..1)
    float *ghost[2];    // mode 0: ghost planes [nslow][nfast], polarization 0 / 1
    float *mem;         // [2][npos][kMemStride]: history and statistics (kMemStats)
    float *samples;     // modes 1, 2: [rows][2][npos][kSampleStride]
    uint64_t emit_base, emit_every, emit_rows;
    const MlpPacked *net;
};

// n (1..3) faces in one launch; the hooks must not write what another
// reads (true for the faces of one side: each writes its own ghost planes
// and history, and reads only the fields).
hipError_t launch_face_hooks(const Geo &g, const Ptrs &esrc, const Ptrs &hsrc,
                             const FaceHook *const *hooks, int n, const uint64_t *step,
                             hipStream_t st);
// FP32 parameter vector (device) -> packed network (device).
hipError_t pack_mlp(const float *params, int head, MlpPacked *net, hipStream_t st);

}  // namespace wf
