// sc_mlp.h -- the learned surrogate of the patch attraction of a pair.
//
// Network: 12 body-frame features -> 64 -> 64 -> 64 -> 1, SiLU after each
// hidden layer.  Inference inside the Monte Carlo sweep runs one wave per
// tile of 32 pairs on MFMA (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32
// accumulation); training uses FP32 kernels with FP32 master weights.
//
// The surrogate pair energy used by the delayed-acceptance first stage is
//   s(a,b) = WCA(d_seg) + [d_seg <= rcut] * MLP(features(a,b))
// with the exact (cheap) segment distance and repulsion, and the network in
// place of the branch-heavy patch-wedge attraction (paire.h:466-827,
// paire.cpp:244-301).  The exact energy remains the second stage.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "sc_pair.h"

namespace scg {

constexpr int MLP_NF = 12;     // features
constexpr int MLP_KP = 16;     // features (+ the bias slot) padded to the MFMA K
constexpr int MLP_NH = 64;     // hidden width

// FP32 master weights, flat: W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_NF * MLP_NH;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_NH;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_NH * MLP_NH;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_NH;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_NH * MLP_NH;
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_NH;
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_NH;
constexpr int MLP_NPARAM = MLP_OFF_B4 + 1;   // 9217

SC_HD float silu(float z) { return z / (1.0f + expf(-z)); }
SC_HD float silu_grad(float z) {
    float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Features of pair (a moving as part1, b) in a's body frame (e1 = dir,
// e2 = patch direction, e3 = e1 x e2); r = pos_a - pos_b (Cuboid::image
// convention), d = closest segment distance.  lsc, rcut, pdis, pswitch: scale
// constants of the pair type.
SC_HD void mlp_features(const V3<float>& adir, const V3<float>& apd, const V3<float>& bdir, const V3<float>& bpd,
                        const V3<float>& r, float d, float lsc, float rcut, float pdis, float pswitch, float* f) {
    V3<float> e3 = cross(adir, apd);
    V3<float> v = mk<float>(-r.x, -r.y, -r.z);
    float il = 1.0f / lsc;
    f[0] = dot(v, adir) * il;
    f[1] = dot(v, apd) * il;
    f[2] = dot(v, e3) * il;
    f[3] = dot(bdir, adir);
    f[4] = dot(bdir, apd);
    f[5] = dot(bdir, e3);
    f[6] = dot(bpd, adir);
    f[7] = dot(bpd, apd);
    f[8] = dot(bpd, e3);
    f[9] = d / rcut;
    f[10] = sqrtf(dot(v, v)) * il;
    float w;
    if (d < pdis) w = 1.0f;
    else if (d > pdis + pswitch) w = 0.0f;
    else {
        float c = cosf(1.5707963267948966f * (d - pdis) / pswitch);
        w = c * c;
    }
    f[11] = w;
}

// FP32 forward of one sample (reference path for tests and training)
SC_HD float mlp_forward_f32(const float* W, const float* x) {
    float h1[MLP_NH], h2[MLP_NH];
    for (int n = 0; n < MLP_NH; n++) {
        float z = W[MLP_OFF_B1 + n];
        for (int k = 0; k < MLP_NF; k++) z += x[k] * W[MLP_OFF_W1 + k * MLP_NH + n];
        h1[n] = silu(z);
    }
    for (int n = 0; n < MLP_NH; n++) {
        float z = W[MLP_OFF_B2 + n];
        for (int k = 0; k < MLP_NH; k++) z += h1[k] * W[MLP_OFF_W2 + k * MLP_NH + n];
        h2[n] = silu(z);
    }
    float out = W[MLP_OFF_B4];
    for (int n = 0; n < MLP_NH; n++) {
        float z = W[MLP_OFF_B3 + n];
        for (int k = 0; k < MLP_NH; k++) z += h2[k] * W[MLP_OFF_W3 + k * MLP_NH + n];
        out += silu(z) * W[MLP_OFF_W4 + n];
    }
    return out;
}

// ---------------------------------------------------------------- MFMA tile
// Inference runs the transposed product H^T = W^T X^T on
// v_mfma_f32_32x32x16_bf16 (D[i][j] = sum_k A[i][k] B[k][j], 32x32 output,
// K = 16).  Operand layout: lane l holds A[l%32][8*(l/32)+0..7] and
// B[8*(l/32)+0..7][l%32]; output register r of lane l is
// D[8*(r/4) + 4*(l/32) + r%4][l%32].
// With A = a 32-unit block of W^T and B = X^T (32 samples), lane l ends up
// holding sample l%32 and hidden units u(kg, blk, r) = 32*blk + 8*(r/4) +
// 4*kg + r%4 (kg = l/32).  Those are exactly the 8-value K slices the next
// layer's B operand needs, if that layer's K index is permuted as
//   k(ks, kg, j) = 32*(ks/2) + 8*(2*(ks%2) + j/4) + 4*kg + j%4
// (K step ks = 0..3, slot j = 0..7): registers c_blk[0..7] and c_blk[8..15]
// become the B operands of K steps 2*blk and 2*blk+1.  So activations never
// leave registers; the weights are packed in that permuted order once, after
// training.  The first layer's bias rides in feature slot 12 (= 1.0), the
// other biases initialise the accumulators; the output layer is a per-lane
// dot over the lane's 32 units plus one cross-half exchange.
//
// Packed image (global memory), all per-lane slices contiguous:
//   A1 [blk 2][lane 64] bf16x8        W1ext[8*kg+j][32*blk+l%32]
//   A2 [blk 2][ks 4][lane 64] bf16x8  W2[k(ks,kg,j)][32*blk+l%32]
//   A3 likewise for W3
//   C  [3][kg 2][blk 2][16] float     b2[u], b3[u], w4[u] with u = u(kg,blk,r)
//   b4, 3 floats of padding
constexpr int MLP_A1_OFF = 0;
constexpr int MLP_A2_OFF = MLP_A1_OFF + 2 * 64 * 16;
constexpr int MLP_A3_OFF = MLP_A2_OFF + 8 * 64 * 16;
constexpr int MLP_C_OFF = MLP_A3_OFF + 8 * 64 * 16;
constexpr int MLP_C_BYTES = 4 * (3 * 64 + 4);
constexpr int MLP_IMG_BYTES = MLP_C_OFF + MLP_C_BYTES;   // 19216
constexpr int MLP_LDS_BYTES = MLP_C_BYTES;                // per wave: the constants
static_assert(MLP_C_OFF % 16 == 0 && MLP_C_BYTES % 16 == 0, "16-byte copies");

SC_HD int mlp_unit(int kg, int blk, int r) { return 32 * blk + 8 * (r >> 2) + 4 * kg + (r & 3); }
SC_HD int mlp_kperm(int ks, int kg, int j) { return 32 * (ks >> 1) + 8 * (2 * (ks & 1) + (j >> 2)) + 4 * kg + (j & 3); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// the A operands of one lane
struct MlpRegs {
    bf16x8 a1[2];
    bf16x8 a2[8];   // [blk*4 + ks]
    bf16x8 a3[8];
};

// copy the constants into LDS (`lds`, MLP_LDS_BYTES) and this lane's operands into registers
__device__ __forceinline__ void mlp_setup(const unsigned char* __restrict__ img, float* lds, int lane, MlpRegs& R) {
    for (int i = lane; i < MLP_C_BYTES / 16; i += 64) ((uint4*)lds)[i] = ((const uint4*)(img + MLP_C_OFF))[i];
    const bf16x8* a1 = (const bf16x8*)(img + MLP_A1_OFF);
    const bf16x8* a2 = (const bf16x8*)(img + MLP_A2_OFF);
    const bf16x8* a3 = (const bf16x8*)(img + MLP_A3_OFF);
#pragma unroll
    for (int b = 0; b < 2; b++) R.a1[b] = a1[b * 64 + lane];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        R.a2[q] = a2[q * 64 + lane];
        R.a3[q] = a3[q * 64 + lane];
    }
}

// SiLU for inference: hardware exp2 and reciprocal
__device__ __forceinline__ float silu_fast(float z) {
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ __forceinline__ bf16x8 mlp_pack8(const floatx16& c, int half) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = (__bf16)silu_fast(c[8 * half + j]);
    return v;
}

__device__ __forceinline__ floatx16 mlp_load_c(const float* p) {
    floatx16 c;
    const float4* q = (const float4*)p;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float4 v = q[i];
        c[4 * i] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
    }
    return c;
}

// Network output for sample l%32 of a 32-sample tile (returned on lanes l and
// l+32).  x: this lane's feature slots 8*kg..8*kg+7 of its sample, BF16
// (slot 12 = 1.0, slots 13-15 = 0).  cst: the constants in LDS.  Samples
// are independent: a sample's output does not depend on its tile mates.
__device__ __forceinline__ float mlp_tile(const MlpRegs& R, const float* cst, bf16x8 x, int lane) {
    int kg = lane >> 5;
    floatx16 z = {};
    floatx16 c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.a1[0], x, z, 0, 0, 0);
    floatx16 c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.a1[1], x, z, 0, 0, 0);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        bf16x8 h0 = mlp_pack8(c0, 0), h1 = mlp_pack8(c0, 1), h2 = mlp_pack8(c1, 0), h3 = mlp_pack8(c1, 1);
        const bf16x8* a = layer == 0 ? R.a2 : R.a3;
        const float* cb = cst + layer * 64 + kg * 32;
        c0 = mlp_load_c(cb);
        c1 = mlp_load_c(cb + 16);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], h0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[4], h0, c1, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], h1, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[5], h1, c1, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2], h2, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[6], h2, c1, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[3], h3, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[7], h3, c1, 0, 0, 0);
    }
    const float* w4 = cst + 128 + kg * 32;
    float part = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; r++) part += silu_fast(c0[r]) * w4[r] + silu_fast(c1[r]) * w4[16 + r];
    part += __shfl_xor(part, 32);
    return part + cst[192];
}

// BF16 feature slice of this lane from the 12 features of its sample
__device__ __forceinline__ bf16x8 mlp_slots(const float* f, int kg) {
    bf16x8 x;
    if (kg == 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = (__bf16)f[j];
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = (__bf16)f[8 + j];
        x[4] = (__bf16)1.0f;
        x[5] = x[6] = x[7] = (__bf16)0.0f;
    }
    return x;
}

}  // namespace scg
