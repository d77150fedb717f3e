// sc_mlp.h -- the learned surrogate of the patch attraction of a pair.
//
// Network: 12 body-frame features -> 64 -> 64 -> 64 -> 1, SiLU after each
// hidden layer.  Inference inside the Monte Carlo sweep runs tiles of 16
// pairs on MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation), the hidden units of a tile split between the waves of the
// sweep's block; training uses FP32 kernels with FP32 master weights.
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

// cos of an argument in [0, pi/2] for the features: the hardware's on the
// device (v_cos_f32, ~1e-6 absolute there), libm's on the host
SC_HD float mlp_cos(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __cosf(x);
#else
    return cosf(x);
#endif
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
        float c = mlp_cos(1.5707963267948966f * (d - pdis) / pswitch);
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
// v_mfma_f32_16x16x32_bf16 (D[i][j] = sum_k A[i][k] B[k][j], 16x16 output,
// K = 32), one tile of 16 samples per wave.  Operand layout: lane l holds
// A[l%16][8*(l/16)+0..7] and B[8*(l/16)+0..7][l%16]; output register r of
// lane l is D[4*(l/16) + r][l%16].
// With A = a 16-unit block of W^T and B = X^T, lane l (group kg = l/16) ends
// up holding sample l%16 and hidden units u(kg, blk, r) = 16*blk + 4*kg + r
// of the four blocks.  Those are exactly the 8-value K slices the next
// layer's B operand needs, if that layer's K index is permuted as
//   k(ks, kg, j) = 16*(2*ks + j/4) + 4*kg + j%4
// (K step ks = 0..1, slot j = 0..7): the registers of blocks 2ks and 2ks+1
// become the B operand of K step ks.  So activations never leave registers;
// the weights are packed in that permuted order once, after training.  The
// first layer's K is the 12 features, the bias (slot 12 = 1.0) and zeros;
// the other biases initialise the accumulators; the output layer is a
// per-lane dot over the lane's 16 units plus two cross-group exchanges.
//
// Packed image (global memory), all per-lane slices contiguous:
//   A1 [blk 4][lane 64] bf16x8        W1ext[8*kg+j][16*blk+l%16]
//   A2 [blk 4][ks 2][lane 64] bf16x8  W2[k(ks,kg,j)][16*blk+l%16]
//   A3 likewise for W3
//   C  [3][kg 4][blk 4][4] float      b2[u], b3[u], w4[u] with u = u(kg,blk,r)
//   b4, 3 floats of padding
constexpr int MLP_TILE = 16;   // samples per tile
constexpr int MLP_A1_OFF = 0;
constexpr int MLP_A2_OFF = MLP_A1_OFF + 4 * 64 * 16;
constexpr int MLP_A3_OFF = MLP_A2_OFF + 8 * 64 * 16;
constexpr int MLP_C_OFF = MLP_A3_OFF + 8 * 64 * 16;
constexpr int MLP_C_BYTES = 4 * (3 * 64 + 4);
constexpr int MLP_IMG_BYTES = MLP_C_OFF + MLP_C_BYTES;   // 21264
constexpr int MLP_LDS_BYTES = MLP_C_BYTES;                // per wave: the constants
static_assert(MLP_C_OFF % 16 == 0 && MLP_C_BYTES % 16 == 0, "16-byte copies");

SC_HD int mlp_unit(int kg, int blk, int r) { return 16 * blk + 4 * kg + r; }
SC_HD int mlp_kperm(int ks, int kg, int j) { return 16 * (2 * ks + (j >> 2)) + 4 * kg + (j & 3); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// The A operands of one lane for BPW of the four 16-unit blocks of every
// layer (blocks b0 .. b0+BPW-1): a whole network (BPW = 4, 80 VGPRs) for one
// wave that evaluates a tile alone, or its share when the waves of a block
// split the hidden units between them.
template <int BPW>
struct MlpRegsW {
    bf16x8 a1[BPW];       // [blk]
    bf16x8 a2[2 * BPW];   // [blk*2 + ks]
    bf16x8 a3[2 * BPW];
};
using MlpRegs = MlpRegsW<4>;

// copy the constants into LDS (`lds`, MLP_LDS_BYTES) and this lane's
// operands of blocks b0 .. b0+BPW-1 into registers
template <int BPW>
__device__ __forceinline__ void mlp_setup_w(const unsigned char* __restrict__ img, float* lds, int lane, int b0,
                                            MlpRegsW<BPW>& R) {
    for (int i = lane; i < MLP_C_BYTES / 16; i += 64) ((uint4*)lds)[i] = ((const uint4*)(img + MLP_C_OFF))[i];
    const bf16x8* a1 = (const bf16x8*)(img + MLP_A1_OFF);
    const bf16x8* a2 = (const bf16x8*)(img + MLP_A2_OFF);
    const bf16x8* a3 = (const bf16x8*)(img + MLP_A3_OFF);
#pragma unroll
    for (int b = 0; b < BPW; b++) {
        R.a1[b] = a1[(b0 + b) * 64 + lane];
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            R.a2[2 * b + ks] = a2[(2 * (b0 + b) + ks) * 64 + lane];
            R.a3[2 * b + ks] = a3[(2 * (b0 + b) + ks) * 64 + lane];
        }
    }
}
__device__ __forceinline__ void mlp_setup(const unsigned char* __restrict__ img, float* lds, int lane, MlpRegs& R) {
    mlp_setup_w<4>(img, lds, lane, 0, R);
}
// this lane's operands of all four unit blocks from the packed image in
// global memory (the constants stay where mlp_setup put them)
__device__ __forceinline__ void mlp_load_ops(const unsigned char* __restrict__ img, int lane, MlpRegs& R) {
    const bf16x8* a1 = (const bf16x8*)(img + MLP_A1_OFF);
    const bf16x8* a2 = (const bf16x8*)(img + MLP_A2_OFF);
    const bf16x8* a3 = (const bf16x8*)(img + MLP_A3_OFF);
#pragma unroll
    for (int b = 0; b < 4; b++) R.a1[b] = a1[b * 64 + lane];
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

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x4 mlp_act4(const floatx4& c) {
    bf16x4 v;
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = (__bf16)silu_fast(c[j]);
    return v;
}

__device__ __forceinline__ bf16x8 mlp_join(const bf16x4& a, const bf16x4& b) {
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// LDS exchange of a block's activations between the waves of a split tile:
// two buffers (one per hidden layer that feeds another) of [blk 4][lane 64]
constexpr int MLP_HX_BYTES = 2 * 4 * 64 * (int)sizeof(bf16x4);   // 4 KiB

// Network output for sample l%16 of a 16-sample tile, evaluated by NW waves
// together (NW = 1, 2 or 4): wave wv computes the unit blocks
// wv*4/NW .. (wv+1)*4/NW - 1 of every layer and returns its share of the
// output (the output layer's dot over its units, plus the output bias on
// wave 0), so that the shares of the waves sum to the network's output; for
// NW = 1 that is the whole output.  The result is on lanes l, l+16, l+32,
// l+48.  With NW > 1 each hidden layer's activations meet in `hx`
// (MLP_HX_BYTES of LDS) behind a barrier of the block, so every wave of the
// block must call this the same number of times.  x: this lane's feature
// slots 8*kg..8*kg+7 of its sample, BF16 (slot 12 = 1.0, the rest of 12..31
// zero), the same on every wave.  cst: the constants in LDS.  Samples are
// independent: a sample's output does not depend on its tile mates.
template <int NW>
__device__ __forceinline__ float mlp_tile_w(const MlpRegsW<4 / NW>& R, const float* cst, bf16x4* hx, bf16x8 x,
                                            int lane, int wv) {
    constexpr int BPW = 4 / NW;
    const int kg = lane >> 4, b0 = wv * BPW;
    floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    floatx4 c[BPW];
#pragma unroll
    for (int b = 0; b < BPW; b++) c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.a1[b], x, z, 0, 0, 0);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        // this layer's activations as the next layer's two K slices
        // (blocks 0,1 and 2,3)
        bf16x8 h0, h1;
        if constexpr (NW == 1) {
            h0 = mlp_join(mlp_act4(c[0]), mlp_act4(c[1]));
            h1 = mlp_join(mlp_act4(c[2]), mlp_act4(c[3]));
        } else {
            bf16x4* buf = hx + layer * 256;
#pragma unroll
            for (int b = 0; b < BPW; b++) buf[(b0 + b) * 64 + lane] = mlp_act4(c[b]);
            __syncthreads();
            h0 = mlp_join(buf[lane], buf[64 + lane]);
            h1 = mlp_join(buf[128 + lane], buf[192 + lane]);
        }
        const bf16x8* a = layer == 0 ? R.a2 : R.a3;
        const floatx4* cb = (const floatx4*)(cst + layer * 64 + kg * 16);
#pragma unroll
        for (int b = 0; b < BPW; b++) {
            c[b] = cb[b0 + b];
            c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2 * b], h0, c[b], 0, 0, 0);
            c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2 * b + 1], h1, c[b], 0, 0, 0);
        }
    }
    const floatx4* w4 = (const floatx4*)(cst + 128 + kg * 16);
    float part = 0.0f;
#pragma unroll
    for (int b = 0; b < BPW; b++) {
        floatx4 w = w4[b0 + b];
#pragma unroll
        for (int r = 0; r < 4; r++) part += silu_fast(c[b][r]) * w[r];
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return wv == 0 ? part + cst[192] : part;
}

// the whole network on one wave
__device__ __forceinline__ float mlp_tile(const MlpRegs& R, const float* cst, bf16x8 x, int lane) {
    return mlp_tile_w<1>(R, cst, nullptr, x, lane, 0);
}

// The whole network on one wave with the operands read from the packed
// image in global memory layer by layer (at most the 32 VGPRs of one hidden
// layer live at a time, not all 80): for a wave that evaluates tiles while
// its register budget is shared with other work.
__device__ __forceinline__ float mlp_tile_g(const unsigned char* __restrict__ img, const float* cst, bf16x8 x,
                                            int lane) {
    const int kg = lane >> 4;
    const bf16x8* a1 = (const bf16x8*)(img + MLP_A1_OFF);
    floatx4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    floatx4 c[4];
#pragma unroll
    for (int b = 0; b < 4; b++) c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[b * 64 + lane], x, z, 0, 0, 0);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        // nothing moves across: a layer's operand loads stay after the
        // previous layer (the scheduler would otherwise hoist all 80 VGPRs)
        __builtin_amdgcn_sched_barrier(0);
        bf16x8 h0 = mlp_join(mlp_act4(c[0]), mlp_act4(c[1]));
        bf16x8 h1 = mlp_join(mlp_act4(c[2]), mlp_act4(c[3]));
        const bf16x8* a = (const bf16x8*)(img + (layer == 0 ? MLP_A2_OFF : MLP_A3_OFF));
        const floatx4* cb = (const floatx4*)(cst + layer * 64 + kg * 16);
#pragma unroll
        for (int b = 0; b < 4; b++) {
            c[b] = cb[b];
            c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[(2 * b) * 64 + lane], h0, c[b], 0, 0, 0);
            c[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[(2 * b + 1) * 64 + lane], h1, c[b], 0, 0, 0);
        }
    }
    const floatx4* w4 = (const floatx4*)(cst + 128 + kg * 16);
    float part = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        floatx4 w = w4[b];
#pragma unroll
        for (int r = 0; r < 4; r++) part += silu_fast(c[b][r]) * w[r];
    }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    return part + cst[192];
}

// BF16 feature slice of this lane from the 12 features of its sample
__device__ __forceinline__ bf16x8 mlp_slots(const float* f, int kg) {
    bf16x8 x;
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = (__bf16)0.0f;
    if (kg == 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = (__bf16)f[j];
    } else if (kg == 1) {
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = (__bf16)f[8 + j];
        x[4] = (__bf16)1.0f;
    }
    return x;
}

}  // namespace scg
