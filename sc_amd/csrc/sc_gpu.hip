// sc_gpu.hip -- gfx950 kernels and the C ABI (include/sc_gpu.h) of the
// patchy-spherocylinder Monte Carlo hot path.
//
// Device state (one context = one system):
//   P  uint4  (ux, uy, uz, dense type)  positions as 32-bit fractions of the box
//   O  float4 (dir.x, dir.y, dir.z, pd0.x)
//   Q  float2 (pd0.y, pd0.z)
//   ID int    original (reference) index of the particle in this slot
// 40 bytes of record + 4 of index per particle, kept sorted by cell after
// every cell-list build (double buffered), so a cell is a contiguous range.
// Derived vectors (patch sides, chiral axes, second patch) are recomputed on
// the fly from dir/pd0 and the per-type constants.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

// Per-expression FMA contraction for the whole file (the pair header's
// reason: every kernel must round a pair term alike).
#pragma clang fp contract(on)

#include "../../include/sc_gpu.h"
#include "sc_host.h"
#include "sc_mlp.h"
#include "sc_pair.h"
#include "sc_rng.h"

using namespace scg;

typedef PairParamsT<float> PP32;
typedef PairParamsT<double> PP64;
typedef TypeParamsT<float> TP32;
typedef TypeParamsT<double> TP64;

#define WAVE 64
#ifndef SWEEP_TABLES_IN_LDS
#define SWEEP_TABLES_IN_LDS 1
#endif
#define SWEEP_TABLE_MAXT 4   // systems of up to this many types keep the tables in LDS
#define MAX_TYPES 16

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ int wave_prefix(unsigned long long mask) {
    return (int)__popcll(mask & ((1ull << lane_id()) - 1ull));
}

// Helpers for kernels whose blocks are NW waves working on one task (NW = 1:
// plain wave code, no barriers).  Both take every thread of the block.
// Sum over the block, the same value on every thread; waves combine in a
// fixed order, so the result is deterministic.  red: NW doubles of LDS.
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    if constexpr (NW == 1) {
        return v;
    } else {
        if (lane_id() == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NW; k++) s += red[k];
        __syncthreads();
        return s;
    }
}
// Wave sum in a fixed tree through DPP lane moves instead of LDS shuffles:
// pairs of lanes (xor 1, xor 2), quads (half-row mirror) and halves (row
// mirror) inside each row of 16 lanes, then the four row sums read back in
// order.  The same value on every lane.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_lane(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_move<0xb1>(v);    // quad_perm [1,0,3,2]
    v += dpp_move<0x4e>(v);    // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);   // row_half_mirror
    v += dpp_move<0x140>(v);   // row_mirror
    return (row_lane(v, 0) + row_lane(v, 16)) + (row_lane(v, 32) + row_lane(v, 48));
}
// Two block sums of a sequence of block-uniform calls behind one barrier:
// consecutive calls alternate between two halves of red (par flips), so a
// call's writes never meet the previous call's reads.  red: 4 NW doubles.
template <int NW>
__device__ __forceinline__ void block_sum2(double a, double b, double* red, int& par, double& sa, double& sb) {
    a = wave_sum_dpp(a);
    b = wave_sum_dpp(b);
    if constexpr (NW == 1) {
        sa = a;
        sb = b;
    } else {
        double* r = red + 2 * NW * par;
        par ^= 1;
        if (lane_id() == 0) {
            r[2 * (threadIdx.x >> 6)] = a;
            r[2 * (threadIdx.x >> 6) + 1] = b;
        }
        __syncthreads();
        double x = 0.0, y = 0.0;
#pragma unroll
        for (int k = 0; k < NW; k++) {
            x += r[2 * k];
            y += r[2 * k + 1];
        }
        sa = x;
        sb = y;
    }
}
// K stable compaction steps at once (the items of set 0, then those of set 1,
// ...), one barrier.  Like block_sum2, consecutive calls alternate between two
// halves of cnt (CNT_HALF ints each, K NW <= CNT_HALF; block_compact1 shares
// cnt and par).  cnt: 2 CNT_HALF ints.
#define CNT_HALF 16
template <int NW, int K>
__device__ __forceinline__ void block_compactK(const bool* f, int& total, int* cnt, int& par, int* p) {
    static_assert(NW == 1 || K * NW <= CNT_HALF, "compaction counts exceed their half");
    unsigned long long m[K];
#pragma unroll
    for (int h = 0; h < K; h++) m[h] = __ballot(f[h]);
    if constexpr (NW == 1) {
#pragma unroll
        for (int h = 0; h < K; h++) {
            p[h] = total + wave_prefix(m[h]);
            total += __popcll(m[h]);
        }
    } else {
        int* c = cnt + CNT_HALF * par;
        par ^= 1;
        int w = threadIdx.x >> 6;
        if (lane_id() == 0) {
#pragma unroll
            for (int h = 0; h < K; h++) c[h * NW + w] = __popcll(m[h]);
        }
        __syncthreads();
        int base = total;
#pragma unroll
        for (int h = 0; h < K; h++) {
            int before = 0, all = 0;
#pragma unroll
            for (int k = 0; k < NW; k++) {
                int ck = c[h * NW + k];
                before += k < w ? ck : 0;
                all += ck;
            }
            p[h] = base + before + wave_prefix(m[h]);
            base += all;
        }
        total = base;
    }
}
// block_compact with one barrier, alternating halves of cnt as
// block_compactK, with which it shares cnt and par.
template <int NW>
__device__ __forceinline__ int block_compact1(bool flag, int& total, int* cnt, int& par) {
    unsigned long long m = __ballot(flag);
    if constexpr (NW == 1) {
        int pos = total + wave_prefix(m);
        total += __popcll(m);
        return pos;
    } else {
        int* c = cnt + CNT_HALF * par;
        par ^= 1;
        int w = threadIdx.x >> 6;
        if (lane_id() == 0) c[w] = __popcll(m);
        __syncthreads();
        int base = total, all = 0;
#pragma unroll
        for (int k = 0; k < NW; k++) {
            int ck = c[k];
            base += k < w ? ck : 0;
            all += ck;
        }
        total += all;
        return base + wave_prefix(m);
    }
}
// Stable compaction step over a chunk of NW*64 items: the output index of
// this thread's item (when flag), in thread order; total advances by the
// chunk's count.  cnt: NW ints of LDS.
template <int NW>
__device__ __forceinline__ int block_compact(bool flag, int& total, int* cnt) {
    unsigned long long m = __ballot(flag);
    if constexpr (NW == 1) {
        int pos = total + wave_prefix(m);
        total += __popcll(m);
        return pos;
    } else {
        int w = threadIdx.x >> 6;
        if (lane_id() == 0) cnt[w] = __popcll(m);
        __syncthreads();
        int base = total, all = 0;
#pragma unroll
        for (int k = 0; k < NW; k++) {
            int c = cnt[k];
            base += k < w ? c : 0;
            all += c;
        }
        __syncthreads();
        total += all;
        return base + wave_prefix(m);
    }
}

// Workgroups are dealt to the 8 XCDs round-robin by ID (XCD = ID mod 8), each
// XCD with its own L2.  This bijection of [0, n) gives XCD k a contiguous run
// of logical indices, so with z-major cell order each XCD sweeps one slab and
// the overlapping stencils of its cells share its L2.
__device__ __forceinline__ int xcd_slab_index(int b, int n) {
    int q = n >> 3, rem = n & 7, k = b & 7;
    return k * q + min(k, rem) + (b >> 3);
}

__device__ __forceinline__ int axis_cell(uint32_t u, uint32_t s, int nc) {
    return (int)(((uint64_t)(uint32_t)(u - s) * (uint64_t)nc) >> 32);
}

struct Grid {
    int nc[3];
    uint32_t sh[3];
    int r;   // stencil radius
};

__device__ __forceinline__ int cell_of_u(const Grid& g, uint32_t ux, uint32_t uy, uint32_t uz) {
    return axis_cell(ux, g.sh[0], g.nc[0]) + g.nc[0] * (axis_cell(uy, g.sh[1], g.nc[1]) + g.nc[1] * axis_cell(uz, g.sh[2], g.nc[2]));
}

// particle from its stored record (FP32 or FP64 working precision)
template <typename R>
__host__ __device__ __forceinline__ DPart<R> load_part(const float4 o, const float2 q, int t, const TypeParamsT<R>* tp) {
    DPart<R> p;
    p.dir = mk<R>(o.x, o.y, o.z);
    p.pd[0] = mk<R>(o.w, q.x, q.y);
    p.pd[1] = p.pd[0];
    p.sd[0] = p.sd[1] = p.sd[2] = p.sd[3] = p.pd[0];
    p.t = t;
    derive<R>(p, tp[t]);
    return p;
}

// patch k of a particle from its stored record: view(load_part(...), k),
// computing only patch k's vectors (4 quaternion rotations for a two-patch
// chiral type instead of derive's 7) with derive's expressions, so the values
// are the same.  Types without patches take the full derivation.
template <typename R>
__host__ __device__ __forceinline__ PatchView<R> load_view(const float4 o, const float2 q, int t,
                                                           const TypeParamsT<R>* tp, int k) {
    const TypeParamsT<R>& T = tp[t];
    const int g = T.geotype;
    if (g >= 30 || g == G_SCA || g == G_SCN) return view(load_part<R>(o, q, t, tp), k);
    V3<R> dir = normalised(mk<R>(o.x, o.y, o.z));
    V3<R> pd = normalised(perp(mk<R>(o.w, q.x, q.y), dir));
    R hc = T.pcoshalfi0, hs = T.psinhalfi0;
    if (k) {   // derive's pd[1] (two-patch types only: k = 1 exists for them alone)
        pd = normalised(perp(qrot(pd, dir, T.csec, T.ssec), dir));
        hc = T.pcoshalfi2;
        hs = T.psinhalfi2;
    }
    PatchView<R> v;
    v.ax = is_chiral(g) ? qrot(dir, pd, T.chcos, T.chsin) : dir;
    v.pd = pd;
    v.s0 = qrot(pd, v.ax, hc, hs);
    v.s1 = qrot(pd, v.ax, hc, -hs);
    return v;
}

// ---------------------------------------------------------------- cell list
__global__ void k_cell_count(int n, const uint4* __restrict__ P, Grid g, int* __restrict__ cell_tmp,
                             int* __restrict__ counts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint4 p = P[k];
    int c = cell_of_u(g, p.x, p.y, p.z);
    cell_tmp[k] = c;
    atomicAdd(&counts[c], 1);
}

// exclusive scan of counts[0..ncell) into starts[0..ncell] by one block of
// 1024 threads, in chunks of 8192 staged through LDS (coalesced loads, each
// thread scans 8 contiguous entries, wave prefix sums via shuffles); also the
// maximum occupancy and the number of non-empty cells.
__global__ __launch_bounds__(1024) void k_scan(int ncell, const int* __restrict__ counts, int* __restrict__ starts,
                                                int* __restrict__ maxocc, int* __restrict__ nonempty) {
    __shared__ int tile[8192];
    __shared__ int wsum[16];
    __shared__ int wmax[16], wne[16];
    int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int carry = 0, mx = 0, ne = 0;
    for (int c0 = 0; c0 < ncell; c0 += 8192) {
        for (int i = 0; i < 8; i++) {
            int idx = c0 + t + 1024 * i;
            tile[t + 1024 * i] = idx < ncell ? counts[idx] : 0;
        }
        __syncthreads();
        int v[8], sum = 0;
        for (int i = 0; i < 8; i++) {
            v[i] = tile[8 * t + i];
            sum += v[i];
            mx = max(mx, v[i]);
            ne += v[i] > 0 ? 1 : 0;
        }
        int incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        int wpre = 0, total = 0;
        for (int w = 0; w < 16; w++) {
            if (w < wv) wpre += wsum[w];
            total += wsum[w];
        }
        int run = carry + wpre + incl - sum;
        for (int i = 0; i < 8; i++) {
            int idx = c0 + 8 * t + i;
            if (idx < ncell) starts[idx] = run;
            run += v[i];
        }
        carry += total;
        __syncthreads();
    }
    for (int o = 32; o >= 1; o >>= 1) {
        mx = max(mx, __shfl_xor(mx, o, 64));
        ne += __shfl_xor(ne, o, 64);
    }
    if (lane == 0) { wmax[wv] = mx; wne[wv] = ne; }
    __syncthreads();
    if (t == 0) {
        int m = 0, e = 0;
        for (int w = 0; w < 16; w++) { m = max(m, wmax[w]); e += wne[w]; }
        *maxocc = m;
        *nonempty = e;
        starts[ncell] = carry;
    }
}

__global__ void k_cell_scatter(int n, const int* __restrict__ cell_tmp, const int* __restrict__ starts,
                               int* __restrict__ fill, int* __restrict__ src_of_pos) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int c = cell_tmp[k];
    int pos = starts[c] + atomicAdd(&fill[c], 1);
    src_of_pos[pos] = k;
}

// order each cell by original index (deterministic, matches a stable CPU
// sort): one wave per cell; a cell of up to 64 particles is ranked in
// registers (lane k's rank = how many of the cell's indices are smaller;
// indices are distinct), a larger one (grids of one or two cells per axis) is
// insertion-sorted by its first lane
#define CELL_SORT_CELLS 4   // cells (waves) per 256-thread block
#define CELL_SORT_CHUNKS 8  // cells of up to 512 particles sort in registers; larger ones by one lane
__global__ __launch_bounds__(WAVE * CELL_SORT_CELLS) void k_cell_sort(int ncell, const int* __restrict__ starts,
                                                                     int* __restrict__ src_of_pos,
                                                                     const int* __restrict__ ID) {
    int c = blockIdx.x * CELL_SORT_CELLS + (int)(threadIdx.x >> 6);
    int lane = (int)(threadIdx.x & (WAVE - 1));
    if (c >= ncell) return;
    int b = starts[c], e = starts[c + 1], n = e - b;
    if (n <= 1) return;
    if (n <= WAVE) {
        int v = lane < n ? src_of_pos[b + lane] : 0;
        int key = lane < n ? ID[v] : 0x7fffffff;
        int rank = 0;
        for (int k = 0; k < n; k++) rank += __shfl(key, k, WAVE) < key ? 1 : 0;
        // every lane's read above precedes its write: the write needs the ranks
        if (lane < n) src_of_pos[b + rank] = v;
        return;
    }
    if (n <= CELL_SORT_CHUNKS * WAVE) {
        // the same rank sort over up to CELL_SORT_CHUNKS chunks of 64 held in
        // registers (cells of small replicas, whose axes may hold one cell)
        const int nch = (n + WAVE - 1) / WAVE;
        int vv[CELL_SORT_CHUNKS], kk[CELL_SORT_CHUNKS], rk[CELL_SORT_CHUNKS];
#pragma unroll
        for (int m = 0; m < CELL_SORT_CHUNKS; m++) {
            int i = m * WAVE + lane;
            vv[m] = (m < nch && i < n) ? src_of_pos[b + i] : 0;
            kk[m] = (m < nch && i < n) ? ID[vv[m]] : 0x7fffffff;   // IDs are distinct; padding ranks last
            rk[m] = 0;
        }
        for (int m2 = 0; m2 < nch; m2++)
            for (int k = 0; k < WAVE; k++) {
                int key = __shfl(kk[m2], k, WAVE);
#pragma unroll
                for (int m = 0; m < CELL_SORT_CHUNKS; m++) rk[m] += key < kk[m] ? 1 : 0;
            }
#pragma unroll
        for (int m = 0; m < CELL_SORT_CHUNKS; m++) {
            int i = m * WAVE + lane;
            if (m < nch && i < n) src_of_pos[b + rk[m]] = vv[m];
        }
        return;
    }
    if (lane != 0) return;
    for (int i = b + 1; i < e; i++) {
        int v = src_of_pos[i], key = ID[v], j = i - 1;
        while (j >= b && ID[src_of_pos[j]] > key) { src_of_pos[j + 1] = src_of_pos[j]; j--; }
        src_of_pos[j + 1] = v;
    }
}

__global__ void k_permute(int n, const int* __restrict__ src_of_pos, const uint4* __restrict__ Pi,
                          const float4* __restrict__ Oi, const float2* __restrict__ Qi, const int* __restrict__ IDi,
                          uint4* __restrict__ Po, float4* __restrict__ Oo, float2* __restrict__ Qo,
                          int* __restrict__ IDo, int* __restrict__ slot) {
    int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= n) return;
    int k = src_of_pos[pos];
    Po[pos] = Pi[k];
    Oo[pos] = Oi[k];
    Qo[pos] = Qi[k];
    int id = IDi[k];
    IDo[pos] = id;
    slot[id] = pos;
}

// ------------------------------------------------------------------ staging
// LDS image of the particles of the (2r+1)^3 stencil around a cell.
struct Stage {
    uint4* u;
    float4* o;
    float2* q;
    int* id;
    int n;
    int h0, h1;   // home cell range inside the stage
};

// Stencil cells of home cell (cx,cy,cz): offsets -r..r per axis, reduced to
// distinct cells when an axis has fewer than 2r+1 cells.
__device__ __forceinline__ int stencil_cell(const Grid& g, int cx, int cy, int cz, int idx, int& is_home) {
    int w0 = min(2 * g.r + 1, g.nc[0]), w1 = min(2 * g.r + 1, g.nc[1]);
    int dx = idx % w0, dy = (idx / w0) % w1, dz = idx / (w0 * w1);
    int x = cx - g.r + dx, y = cy - g.r + dy, z = cz - g.r + dz;
    x = ((x % g.nc[0]) + g.nc[0]) % g.nc[0];
    y = ((y % g.nc[1]) + g.nc[1]) % g.nc[1];
    z = ((z % g.nc[2]) + g.nc[2]) % g.nc[2];
    is_home = (x == cx && y == cy && z == cz);
    return x + g.nc[0] * (y + g.nc[1] * z);
}

__device__ __forceinline__ int stencil_size(const Grid& g) {
    return min(2 * g.r + 1, g.nc[0]) * min(2 * g.r + 1, g.nc[1]) * min(2 * g.r + 1, g.nc[2]);
}

// cooperative copy by the block (NW waves; the stencil prefix by wave 0);
// sten holds 2*ns+2 ints.  Returns false on overflow.
template <int NW>
__device__ bool stage_load(Stage& S, int cap, const Grid& g, int cx, int cy, int cz, const int* __restrict__ starts,
                           const uint4* __restrict__ P, const float4* __restrict__ O, const float2* __restrict__ Q,
                           const int* __restrict__ ID, int* sten) {
    int ns = stencil_size(g);
    int* pre = sten;          // ns+1 prefix of counts
    int* cst = sten + ns + 1; // cell start per stencil entry
    int tid = (int)threadIdx.x;
    if (tid < WAVE) {
        int lane = tid;
        int home_idx = -1;
        // counts -> prefix (serial over chunks of 64)
        int carry = 0;
        for (int b = 0; b < ns; b += WAVE) {
            int i = b + lane, cnt = 0, st = 0, hm = 0;
            if (i < ns) {
                int c = stencil_cell(g, cx, cy, cz, i, hm);
                st = starts[c];
                cnt = starts[c + 1] - st;
                cst[i] = st;
            }
            // inclusive wave scan of cnt
            int v = cnt;
            for (int o = 1; o < WAVE; o <<= 1) {
                int t = __shfl_up(v, o, WAVE);
                if (lane >= o) v += t;
            }
            if (i < ns) pre[i] = carry + v - cnt;
            unsigned long long hmask = __ballot(i < ns && hm);
            if (hmask) home_idx = b + __ffsll((long long)hmask) - 1;
            carry += __shfl(v, WAVE - 1, WAVE);
        }
        if (lane == 0) {
            pre[ns] = carry;
            sten[2 * ns + 1] = home_idx;
        }
    }
    __syncthreads();
    int total = pre[ns], home_idx = sten[2 * ns + 1];
    S.n = total;
    S.h0 = pre[home_idx];
    S.h1 = pre[home_idx + 1];
    if (total > cap) return false;
    for (int e = tid; e < total; e += WAVE * NW) {
        // binary search the stencil entry holding element e
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= e) lo = mid; else hi = mid - 1;
        }
        int src = cst[lo] + (e - pre[lo]);
        S.u[e] = P[src];
        S.o[e] = O[src];
        S.q[e] = Q[src];
        S.id[e] = ID[src];
    }
    __syncthreads();
    return true;
}

struct SysConst {
    float L[3];
    float Lh[3], Ll[3];   // box edge as FP32 head + tail (fx_image_c)
    double Ld[3];
    float sqmaxcut;
    int nt;
};

// ------------------------------------------------------------ bonded chains
// Constants of a molecule type's connectivity list (ParticleVector::getConlist,
// scOOP/structures/Conf.h:97-147) and its angle terms (MoleculeParams,
// moleculeparams.h:16-32).  A negative angle constant marks an absent term.
struct MolP {
    double c0, eq0, mod0, mod1, sp;   // nearest-neighbour bond (bond1 / bondd / bondh)
    double c1, eq1;                   // second-neighbour bond (bond2)
    double a1c, a1eq, a2c, a2eq;      // angle1, angle2
};

// Chain word of a particle, by ID: bit p (0..3) set when it is bonded to the
// particle at ID + cw_rel(p); CW_SAME_NEXT when ID + 1 belongs to the same
// molecule; bits 8.. the molecule type.  0 for atomic molecules.
#define CW_SAME_NEXT 16
__host__ __device__ __forceinline__ int cw_rel(int p) { return p == 0 ? -1 : p == 1 ? 1 : p == 2 ? -2 : 2; }
__host__ __device__ __forceinline__ int cw_bit(int rel) {
    return rel == -1 ? 0 : rel == 1 ? 1 : rel == -2 ? 2 : rel == 2 ? 3 : -1;
}

// chain tables handed to the energy and sweep kernels (cw == nullptr: no bonds)
struct ChainTab {
    const int* cw;
    const MolP* mol;
    const PP64* pp64;
};

// EBond::harmonicPotential, paire.h:202-204
__device__ __forceinline__ double harm(double a, double eq, double k) { return k * (a - eq) * (a - eq) * 0.5; }

// pos_a - pos_b, minimum image, FP64 from the fixed-point fractions
__device__ __forceinline__ V3<double> fx_r12(uint4 a, uint4 b, const double* L) {
    return fx_image<double>(a.x, a.y, a.z, b.x, b.y, b.z, L[0], L[1], L[2]);
}

__device__ __forceinline__ V3<double> unit64(float x, float y, float z) { return normalised(mk<double>(x, y, z)); }

// AngleSc::angleEnergyAngle2 (paire.h:335-350): the angle between the patch
// directions of a and b projected on the planes normal to dir_a x dir_b.
// Parallel axes leave the angle undefined (0/0 in the reference): no term.
__device__ __forceinline__ bool angle2_of(const V3<double>& da, const V3<double>& pa, const V3<double>& db,
                                          const V3<double>& pb, double& ang) {
    V3<double> ax = cross(da, db), x1 = cross(da, ax), x2 = cross(db, ax);
    double v1x = dot(x1, pa), v1y = dot(ax, pa), v2x = dot(x2, pb), v2y = dot(ax, pb);
    double den = sqrt((v1x * v1x + v1y * v1y) * (v2x * v2x + v2y * v2y));
    if (!(den > 0.0)) return false;
    ang = acos(fmin(1.0, fmax(-1.0, (v1x * v2x + v1y * v2y) / den)));
    return true;
}

// Bond + angle energy of particle 1 with its chain partner 2 (ID_2 = ID_1 +
// rel) for part1 = 1: HarmonicSc + AngleSc for the SpheroCylinder and MixSpSc
// functors, HarmonicSp for Sphere, nothing for a pair without a functor
// (paire.h:221-353, functor table paire.cpp:6-80).  r12 = pos1 - pos2; d, q
// are unit axes and patch directions.  FP64.  Where rounding takes the
// argument of an acos outside [-1, 1] (NaN in the reference) it is clamped.
__device__ double bond_energy(const MolP& m, const PP64& P, int rel, const V3<double>& r12, const V3<double>& d1,
                              const V3<double>& q1, const V3<double>& d2, const V3<double>& q2) {
    bool nearp = rel == 1 || rel == -1;
    bool sph = P.kind == K_SPH_WCA || P.kind == K_SPH_TAYLOR;
    bool scf = (P.kind >= 0 && P.kind <= K_SCA) || P.kind >= K_SCASPA;
    if (!sph && !scf) return 0.0;
    if (!nearp || sph) {   // HarmonicSp; HarmonicSc for second neighbours: centre distance
        double dist = sqrt(dot(r12, r12));
        return nearp ? harm(dist, m.eq0, m.c0) : harm(dist, m.eq1, m.c1);
    }
    bool prev = rel == -1;
    // HarmonicSc: the bonded ends (part1's +dir end to a previous partner's
    // -dir end, and the reverse), spheres at sp along their axis
    double h1 = P.g0 < 30 ? (P.hl0 + (prev ? m.mod0 : m.mod1)) * (prev ? 1.0 : -1.0) : m.sp;
    double h2 = P.g1 < 30 ? (P.hl1 + (prev ? m.mod1 : m.mod0)) * (prev ? -1.0 : 1.0) : m.sp;
    V3<double> b = mk<double>(r12.x + d1.x * h1 - d2.x * h2, r12.y + d1.y * h1 - d2.y * h2,
                              r12.z + d1.z * h1 - d2.z * h2);
    double e = harm(len(b), m.eq0, m.c0);
    if (m.a1c >= 0) {   // AngleSc angle1: a sphere's axis points from its centre to the partner's bonded end
        V3<double> v1 = d1, v2 = d2;
        if (P.g0 >= 30) {
            double h = P.hl1 * (prev ? -1.0 : 1.0);
            v1 = normalised(mk<double>(d2.x * h - r12.x, d2.y * h - r12.y, d2.z * h - r12.z));
        }
        if (P.g1 >= 30) {
            double h = P.hl0 * (prev ? 1.0 : -1.0);
            v2 = normalised(mk<double>(r12.x + d1.x * h, r12.y + d1.y * h, r12.z + d1.z * h));
        }
        e += harm(acos(fmin(1.0, fmax(-1.0, dot(v1, v2)))), m.a1eq, m.a1c);
    }
    if (m.a2c >= 0 && P.g0 < 30 && P.g1 < 30) {   // AngleSc angle2, (earlier, later) in chain order
        double ang;
        bool ok = prev ? angle2_of(d2, q2, d1, q1, ang) : angle2_of(d1, q1, d2, q2, ang);
        if (ok) e += harm(ang, m.a2eq, m.a2c);
    }
    return e;
}

// Bond energy of the particle at slot a (ID ia, record ua/oa/qa) with those of
// its partners given by `sel` (bits of its chain word), partners read from
// global memory through slot[].  part1 = the particle.
__device__ double bonds_global(const ChainTab& ct, int nt, const double* L, int ia, uint4 ua, float4 oa, float2 qa,
                               int sel, const int* __restrict__ slot, const uint4* __restrict__ P,
                               const float4* __restrict__ O, const float2* __restrict__ Q) {
    int cw = ct.cw[ia];
    int mask = cw & sel & 15;
    if (!mask) return 0.0;
    const MolP& m = ct.mol[cw >> 8];
    V3<double> d1 = unit64(oa.x, oa.y, oa.z), q1 = unit64(oa.w, qa.x, qa.y);
    double e = 0.0;
    for (int p = 0; p < 4; p++) {
        if (!((mask >> p) & 1)) continue;
        int b = slot[ia + cw_rel(p)];
        uint4 ub = P[b];
        float4 ob = O[b];
        float2 qb = Q[b];
        e += bond_energy(m, ct.pp64[ua.w * nt + ub.w], cw_rel(p), fx_r12(ua, ub, L), d1, q1, unit64(ob.x, ob.y, ob.z),
                         unit64(ob.w, qb.x, qb.y));
    }
    return e;
}

// Whole-system bond energy: one thread per ID, its partners of smaller ID
// (part1 = the larger ID, as TotalEMatrix::allToAll pairs (i, j < i)).
__global__ void k_bond_all(int n, ChainTab ct, SysConst sc, const int* __restrict__ slot, const uint4* __restrict__ P,
                           const float4* __restrict__ O, const float2* __restrict__ Q, double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = slot[i];
    out[i] = bonds_global(ct, sc.nt, sc.Ld, i, P[a], O[a], Q[a], 1 | 4, slot, P, O, Q);
}

// ------------------------------------------------------------ energy (all)
// One block of NW waves per cell; pairs (i in home cell, j in stencil) with
// ID_j < ID_i flattened over the block's threads (wave w takes the chunks
// w, w + NW, ...).  The cheap stage (image, cutoff, segment distance, WCA)
// runs for every candidate; pairs inside the attraction range are queued in
// the wave's LDS queue and evaluated 64 at a time so the branch-heavy patch
// code runs with full lanes.  FP32 pair terms, FP64 accumulation; per-cell
// partial sums combined in a fixed order.
#define QCAP 128
#define ENERGY_NW 4

// the LDS writes of a wave are visible to its own later reads (a wave's LDS
// operations complete in order; this keeps the compiler from moving them)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// The synchronisation of one trial's threads: the block, or (WS) only the
// calling wave, when the waves of a block evaluate trials of their own.
template <bool WS>
__device__ __forceinline__ void tsync() {
    if constexpr (WS) wave_lds_sync();
    else __syncthreads();
}

template <int NW>
__global__ __launch_bounds__(WAVE * NW) void k_energy_all(Grid g, SysConst sc, const int* __restrict__ starts,
                                                          const uint4* __restrict__ P, const float4* __restrict__ O,
                                                          const float2* __restrict__ Q, const int* __restrict__ ID,
                                                          const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                                                          int cap, double* __restrict__ partial, int* __restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    int cell = blockIdx.x;
    int cx = cell % g.nc[0], cy = (cell / g.nc[0]) % g.nc[1], cz = cell / (g.nc[0] * g.nc[1]);
    Stage S;
    S.u = (uint4*)smem;
    S.o = (float4*)(S.u + cap);
    S.q = (float2*)(S.o + cap);
    S.id = (int*)(S.q + cap);
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1), w = tid >> 6;
    int* sten = S.id + cap;                              // 2*125+2 ints
    double* red = (double*)(sten + 2 * 125 + 2);         // NW doubles (8-byte aligned: cap is a multiple of 64)
    int* queue = (int*)(red + NW) + w * 2 * QCAP;        // this wave's QCAP (hi, j) entries
    if (starts[cell + 1] == starts[cell]) {
        if (tid == 0) partial[cell] = 0.0;
        return;
    }
    if (!stage_load<NW>(S, cap, g, cx, cy, cz, starts, P, O, Q, ID, sten)) {
        if (tid == 0) { atomicOr(flag, 1); partial[cell] = 0.0; }
        return;
    }
    int nh = S.h1 - S.h0, ns = S.n;
    long total = (long)nh * ns;
    double acc = 0.0;
    int qn = 0;   // wave-uniform queue length
    for (long base = (long)w * WAVE; base < total; base += (long)WAVE * NW) {
        long pidx = base + lane;
        bool need = false;
        int hi_ = 0, j = 0;
        if (pidx < total) {
            hi_ = (int)(pidx / ns);
            j = (int)(pidx - (long)hi_ * ns);
            int i = S.h0 + hi_;
            if (S.id[j] < S.id[i]) {
                uint4 a = S.u[i], b = S.u[j];
                V3<float> r = fx_image_c(a.x, a.y, a.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
                float r2 = dot(r, r);
                if (r2 <= sc.sqmaxcut) {
                    const PP32& P2 = pp[a.w * sc.nt + b.w];
                    if (P2.kind <= K_SCA && P2.kind >= 0) {
                        DPart<float> pa, pb;
                        pa.dir = mk<float>(S.o[i].x, S.o[i].y, S.o[i].z);
                        pb.dir = mk<float>(S.o[j].x, S.o[j].y, S.o[j].z);
                        pa.dir = normalised(pa.dir);
                        pb.dir = normalised(pb.dir);
                        float erep, d2;
                        need = scsc_cheap(P2, pa, pb, r, erep, d2);
                        acc += (double)erep;
                    } else if (P2.kind == K_SPH_WCA || P2.kind == K_SPH_TAYLOR) {
                        acc += (double)sphere_pair(P2, sqrtf(r2));
                    } else if (P2.kind != K_NONE) {
                        need = true;   // sphere - spherocylinder: evaluated whole in the queue
                    }
                }
            }
        }
        unsigned long long m = __ballot(need);
        if (need) {
            int pos = qn + wave_prefix(m);
            queue[2 * pos] = hi_;
            queue[2 * pos + 1] = j;
        }
        qn += __popcll(m);
        wave_lds_sync();
        bool last = base + (long)WAVE * NW >= total;
        if (qn >= WAVE || (last && qn > 0)) {
            // drain whole batches of 64 (everything at the end)
            int take = last ? qn : (qn / WAVE) * WAVE;
            for (int e0 = 0; e0 < take; e0 += WAVE) {
                int e = e0 + lane;
                if (e < take) {
                    int i = S.h0 + queue[2 * e], jj = queue[2 * e + 1];
                    uint4 a = S.u[i], b = S.u[jj];
                    V3<float> r = fx_image_c(a.x, a.y, a.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
                    const PP32& P2 = pp[a.w * sc.nt + b.w];
                    DPart<float> pa = load_part<float>(S.o[i], S.q[i], a.w, tp);
                    DPart<float> pb = load_part<float>(S.o[jj], S.q[jj], b.w, tp);
                    if (P2.kind <= K_SCA) acc += scsc_attr_d(P2, pa, pb, r);
                    else acc += (double)spsc(P2, pa, pb, r);
                }
            }
            wave_lds_sync();
            int rest = qn - take;   // < 64
            int c0 = 0, c1 = 0;
            if (lane < rest) { c0 = queue[2 * (take + lane)]; c1 = queue[2 * (take + lane) + 1]; }
            wave_lds_sync();
            if (lane < rest) { queue[2 * lane] = c0; queue[2 * lane + 1] = c1; }
            wave_lds_sync();
            qn = rest;
        }
    }
    double s = block_sum<NW>(acc, red);
    if (tid == 0) partial[cell] = s;
}

// deterministic final reduction of per-cell partial sums
__global__ __launch_bounds__(1024) void k_reduce(int n, const double* __restrict__ v, double* __restrict__ out) {
    __shared__ double sh[1024];
    int t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < n; i += 1024) s += v[i];
    sh[t] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    if (t == 0) *out = sh[0];
}

// --------------------------------------------------------- energy (queries)
// One wave per query particle: its energy with every particle of the stencil
// around its cell except those of ID ex_lo..ex_hi (the particle itself, or
// its whole molecule for mol2others), plus its bonds for oneToAll (not for
// mol2others, whose connectivity list is empty).  The query state is given
// explicitly (trial states) or taken from a slot.
struct Query {
    uint32_t u[3];
    int t;
    float4 o;
    float2 q;
    int id;            // the query particle's ID
    int ex_lo, ex_hi;  // IDs left out
    int bonds;         // add the bonds with its chain partners
};

__global__ __launch_bounds__(WAVE) void k_energy_query(int nq, const Query* __restrict__ qs, const int* __restrict__ qslot,
                                                       Grid g, SysConst sc, const int* __restrict__ starts,
                                                       const uint4* __restrict__ P, const float4* __restrict__ O,
                                                       const float2* __restrict__ Q, const int* __restrict__ ID,
                                                       const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                                                       ChainTab ct, const int* __restrict__ slot,
                                                       double* __restrict__ out) {
    int qi = blockIdx.x;
    if (qi >= nq) return;
    int lane = lane_id();
    Query q;
    if (qslot) {
        int s = qslot[qi];
        uint4 p = P[s];
        q.u[0] = p.x; q.u[1] = p.y; q.u[2] = p.z; q.t = p.w;
        q.o = O[s]; q.q = Q[s]; q.id = q.ex_lo = q.ex_hi = ID[s];
        q.bonds = 1;
    } else {
        q = qs[qi];
    }
    DPart<float> pa = load_part<float>(q.o, q.q, q.t, tp);
    int c = cell_of_u(g, q.u[0], q.u[1], q.u[2]);
    int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
    int ns = stencil_size(g);
    double acc = 0.0;
    for (int si = 0; si < ns; si++) {
        int hm;
        int cc = stencil_cell(g, cx, cy, cz, si, hm);
        int b = starts[cc], e = starts[cc + 1];
        for (int k = b + lane; k < e; k += WAVE) {
            int idk = ID[k];
            if (idk >= q.ex_lo && idk <= q.ex_hi) continue;
            uint4 pb4 = P[k];
            V3<float> r = fx_image_c(q.u[0], q.u[1], q.u[2], pb4.x, pb4.y, pb4.z, sc.Lh, sc.Ll);
            if (dot(r, r) > sc.sqmaxcut) continue;
            DPart<float> pb = load_part<float>(O[k], Q[k], pb4.w, tp);
            // the pair in its canonical order, the higher ID as part1 (allToAll's)
            if (idk > q.id)
                acc += pair_energy_d<float>(pp[pb4.w * sc.nt + q.t], pb, pa, mk<float>(-r.x, -r.y, -r.z), sc.sqmaxcut);
            else
                acc += pair_energy_d<float>(pp[q.t * sc.nt + pb4.w], pa, pb, r, sc.sqmaxcut);
        }
    }
    // chain partners, wherever they are
    if (ct.cw && lane == 0 && q.bonds)
        acc += bonds_global(ct, sc.nt, sc.Ld, q.id, make_uint4(q.u[0], q.u[1], q.u[2], (uint32_t)q.t), q.o, q.q,
                            15, slot, P, O, Q);
    double s = wave_sum(acc);
    if (lane == 0) out[qi] = s;
}

// ------------------------------------------------------------ pair energies
// p2p(i, j) with i's connectivity list: the non-bonded functor, plus the bond
// and angle terms when j is one of i's chain partners
template <typename R>
__global__ void k_pairs(int m, const int* __restrict__ pi, const int* __restrict__ pj, const int* __restrict__ slot,
                        SysConst sc, const uint4* __restrict__ P, const float4* __restrict__ O,
                        const float2* __restrict__ Q, const PairParamsT<R>* __restrict__ pp,
                        const TypeParamsT<R>* __restrict__ tp, R sqmaxcut, ChainTab ct, double* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int a = slot[pi[k]], b = slot[pj[k]];
    uint4 ua = P[a], ub = P[b];
    DPart<R> pa = load_part<R>(O[a], Q[a], ua.w, tp);
    DPart<R> pb = load_part<R>(O[b], Q[b], ub.w, tp);
    V3<R> r;
    if (sizeof(R) == 8)
        r = fx_image<R>(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, (R)sc.Ld[0], (R)sc.Ld[1], (R)sc.Ld[2]);
    else {
        V3<float> rf = fx_image_c(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, sc.Lh, sc.Ll);
        r = mk<R>((R)rf.x, (R)rf.y, (R)rf.z);
    }
    // FP32: the terms added in FP64, as every FP32 kernel adds them; FP64: the
    // reference's own order of the sum
    double e = sizeof(R) == 8 ? (double)pair_energy<R>(pp[ua.w * sc.nt + ub.w], pa, pb, r, sqmaxcut)
                              : pair_energy_d<R>(pp[ua.w * sc.nt + ub.w], pa, pb, r, sqmaxcut);
    if (ct.cw) {
        int bit = cw_bit(pj[k] - pi[k]), cw = ct.cw[pi[k]];
        if (bit >= 0 && ((cw >> bit) & 1))
            e += bond_energy(ct.mol[cw >> 8], ct.pp64[ua.w * sc.nt + ub.w], cw_rel(bit), fx_r12(ua, ub, sc.Ld),
                             unit64(O[a].x, O[a].y, O[a].z), unit64(O[a].w, Q[a].x, Q[a].y),
                             unit64(O[b].x, O[b].y, O[b].z), unit64(O[b].w, Q[b].x, Q[b].y));
    }
    out[k] = e;
}

// ---------------------------------------------------------------- overlaps
// overlap <=> closest distance < sigma_ij; FP32 first, FP64 when within
// 1e-6 sigma of contact
#define OVERLAP_RECHECK 1e-5
__global__ void k_overlap(int m, const int* __restrict__ pi, const int* __restrict__ pj, const int* __restrict__ slot,
                          SysConst sc, const uint4* __restrict__ P, const float4* __restrict__ O,
                          const float2* __restrict__ Q, const PP64* __restrict__ pp64, const TP64* __restrict__ tp64,
                          unsigned char* __restrict__ out, double* __restrict__ dmin) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int a = slot[pi[k]], b = slot[pj[k]];
    uint4 ua = P[a], ub = P[b];
    const PP64& Pd = pp64[ua.w * sc.nt + ub.w];
    float sig = (float)(Pd.rcutwca / 1.122462048309373);   // sigma_ij = rcutwca / 2^(1/6)
    double sigd = Pd.rcutwca / 1.122462048309373;
    auto dist_of = [&](auto zero) -> double {
        typedef decltype(zero) R;
        V3<R> r = fx_image<R>(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, (R)sc.Ld[0], (R)sc.Ld[1], (R)sc.Ld[2]);
        V3<R> da = normalised(mk<R>(O[a].x, O[a].y, O[a].z)), db = normalised(mk<R>(O[b].x, O[b].y, O[b].z));
        bool s1 = Pd.g0 >= 30, s2 = Pd.g1 >= 30;
        if (s1 && s2) return (double)len(r);
        if (s1 || s2) {
            R contt;
            V3<R> dv;
            R d2 = s1 ? sp_dist<R>(mk<R>(-r.x, -r.y, -r.z), db, (R)Pd.hl1, contt, dv) : sp_dist<R>(r, da, (R)Pd.hl0, contt, dv);
            return sqrt((double)d2);
        }
        V3<R> dv = min_dist(da, db, (R)Pd.hl0, (R)Pd.hl1, r);
        return sqrt((double)dot(dv, dv));
    };
    // FP32 distance; within OVERLAP_RECHECK sigma of contact, where the FP32
    // error (~1e-7 of the separation, below 1e-6 sigma at stencil distances)
    // could put it on the wrong side, the FP64 distance decides
    double d = dist_of(0.0f);
    if (fabs(d - (double)sig) < OVERLAP_RECHECK * sigd) d = dist_of(0.0);
    out[k] = (d < sigd) ? 1 : 0;
    if (dmin) dmin[k] = d;
}

// ------------------------------------------------------------------- sweep
struct MoveParams {
    float transmx, rot_half;
    double temper;
    uint32_t key0, key1;
    uint32_t pass_lo, pass_hi;   // global pass counter
    int ox, oy, oz;              // colour offset
    int p;                       // colour period (r+1)
    int nact[3];                 // active cells per axis
    int trial_mode;              // 0: trials = cell occupancy; 1: dithered mean occupancy
    int n;                       // particles
    int hcap;                    // cells of up to hcap (<= HCAP) particles cache their derived vectors
    int visit;                   // trials per visit of a cell: visit x its occupancy (mode 0) or mean occupancy (1)
};

__device__ __forceinline__ V3<float> unit_from(float a, float b) {
    float z = 1.0f - 2.0f * a;
    float s = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float ph = 6.283185307179586f * b;
    return mk<float>(s * cosf(ph), s * sinf(ph), z);
}

// Moving particle's derived vectors kept in LDS during a trial (27 floats per
// state, DPart order: dir, pd[2], sd[4], ax[2]).
#define MV_STRIDE 28

// All 27 floats from one thread: members chosen by lane (a branch or a
// select per lane) become an indexed stack copy of the particle.
__device__ __forceinline__ void mv_put(float* mv, int off, const V3<float>& v) {
    mv[off] = v.x;
    mv[off + 1] = v.y;
    mv[off + 2] = v.z;
}
__device__ __forceinline__ void mv_store(float* mv, const DPart<float>& p) {
    mv_put(mv, 0, p.dir);
    mv_put(mv, 3, p.pd[0]);
    mv_put(mv, 6, p.pd[1]);
    mv_put(mv, 9, p.sd[0]);
    mv_put(mv, 12, p.sd[1]);
    mv_put(mv, 15, p.sd[2]);
    mv_put(mv, 18, p.sd[3]);
    mv_put(mv, 21, p.ax[0]);
    mv_put(mv, 24, p.ax[1]);
}
__device__ __forceinline__ V3<float> mv_vec(const float* mv, int off) { return mk<float>(mv[off], mv[off + 1], mv[off + 2]); }
__device__ __forceinline__ PatchView<float> mv_view(const float* mv, int k) {
    PatchView<float> v;
    v.pd = mv_vec(mv, 3 + 3 * k);
    v.s0 = mv_vec(mv, 9 + 6 * k);
    v.s1 = mv_vec(mv, 12 + 6 * k);
    v.ax = mv_vec(mv, 21 + 3 * k);
    return v;
}
__device__ __forceinline__ DPart<float> mv_part(const float* mv, int t, int g) {
    DPart<float> p;
    p.dir = mv_vec(mv, 0);
    p.pd[0] = mv_vec(mv, 3); p.pd[1] = mv_vec(mv, 6);
    p.sd[0] = mv_vec(mv, 9); p.sd[1] = mv_vec(mv, 12); p.sd[2] = mv_vec(mv, 15); p.sd[3] = mv_vec(mv, 18);
    p.ax[0] = mv_vec(mv, 21); p.ax[1] = mv_vec(mv, 24);
    p.t = t;
    p.g = g;
    return p;
}

// Energy of the moving particle (as part1, oneToAllTrial, totalenergycalculator.h:383-415)
// at its old and new state against every other staged particle, in three
// compacted phases so that every phase runs with full lanes:
//   A  candidates inside sqmaxcut at either state -> near list
//   B  (near, state) items: segment distance + WCA -> expensive list
//   C  (expensive, patch pair) items: wedge intersections + attraction
struct TrialScratch {
    int* nl;     // near list, cap entries
    int* xl;     // expensive list, 2*cap entries
    float* xd;   // segment distance of each expensive entry, 2*cap
    float* mv;   // 2 * MV_STRIDE floats
    int* bp;     // staged index of each chain partner (4 ints)
    int* cnt;    // block compaction counts (2 CNT_HALF ints)
    double* red; // block sums (16 doubles)
    bf16x4* hx;  // the network's activation exchange between the waves (MLP_HX_BYTES)
    double* netd; // the network screen's energy change, from the wave that computed it
};

// A trial's proposal and its Metropolis thresholds, drawn for TRIAL_CHUNK
// trials at a time by the lanes of the first wave, so that the chain of
// trials does not wait on the RNG.  The thresholds are -T ln u: the test
// exp(-dE/T) > u of moveTry (movecreator.h:175-187) is dE < -T ln u.
struct __align__(16) TrialRec {
    int k;          // staged index of the particle
    int translate;  // 1: translation by (dx, dy, dz); 0: rotation
    int w[5];       // dx, dy, dz (fixed point) | axis x, y, z, cos, sin of the half angle
    int pad;
    double thr;     // final stage
    double thr1;    // first screen of the delayed acceptance
    double thr2;    // second screen
    double pad2;
};
static_assert(sizeof(TrialRec) == 64, "trial record size");
#define TRIAL_CHUNK 64
// LDS after the trial scratch: chain partners, compaction counts, block sums
// (8-aligned), trial records (16-aligned)
// derived vectors of the home cell's particles in their current state (the
// old state of every trial), MV_STRIDE floats each, for cells of up to HCAP
#define HCAP 64
#define SWEEP_SMALL_LDS (16 + 4 * 2 * CNT_HALF + 8 + 128 + 8 + TRIAL_CHUNK * 64 + HCAP * MV_STRIDE * 4)

// sweep counters: STAT_REP replicas of 8 counters, STAT_STRIDE apart (128 B)
#define STAT_REP 64
#define STAT_STRIDE 16

// delayed-acceptance modes of the sweep
#define DA_OFF 0   // exact Metropolis
#define DA_NET 1   // screens: exact cheap terms, then the network's attraction
#define DA_REP 2   // screen: exact cheap terms only

struct MlpConst {
    float lsc, rcut, pdis, pswitch;
    int tnet;   // dense type of the network's pair (the screen's other pairs count 0)
    int minx;   // the network screens trials with at least minx (pair, state) entries in range
};

// Phase A of a trial: every staged particle j != k inside sqmaxcut of the
// moving particle's old or new position goes to the near list T.nl with its
// state flags (bit 28: old, 29: new), in stencil order.  The first SB threads
// of the block scan, PA entries each per round; the others only take part in
// the compaction barriers.  With chain bonds it records the staged index of
// each bonded partner.  tid is the thread's index among the trial's threads;
// entries beyond ncap are counted in nn but not stored.
template <int NW, int PA>
__device__ __forceinline__ void scan_stencil(const Stage& S, int k, const uint4 uo, const uint4 un, bool moved,
                                             const SysConst& sc, int bmask, int idk, const TrialScratch& T, int SB,
                                             bool scanner, int& nn, int& cpar, int ncap, int tid) {
    for (int base = 0; base < S.n; base += PA * SB) {
        int fl[PA], jv[PA];
#pragma unroll
        for (int h = 0; h < PA; h++) {
            int j = base + h * SB + tid, f = 0;
            if (scanner && j < S.n && j != k) {
                uint4 b = S.u[j];
                V3<float> ro = fx_image_c(uo.x, uo.y, uo.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
                bool io = dot(ro, ro) <= sc.sqmaxcut, in_ = io;
                if (moved) {
                    V3<float> rn = fx_image_c(un.x, un.y, un.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
                    in_ = dot(rn, rn) <= sc.sqmaxcut;
                }
                f = (io ? 1 : 0) | (in_ ? 2 : 0);
                if (bmask) {
                    int bit = cw_bit(S.id[j] - idk);
                    if (bit >= 0 && ((bmask >> bit) & 1)) T.bp[bit] = j;
                }
            }
            fl[h] = f;
            jv[h] = j;
        }
        bool fb[PA];
        int pv[PA];
#pragma unroll
        for (int h = 0; h < PA; h++) fb[h] = fl[h] != 0;
        block_compactK<NW, PA>(fb, nn, T.cnt, cpar, pv);
#pragma unroll
        for (int h = 0; h < PA; h++)
            if (fl[h] && pv[h] < ncap) T.nl[pv[h]] = jv[h] | (fl[h] << 28);
    }
}

// The network's 12 features of expensive-list entry e (moving particle as
// part1 at the entry's state, the neighbour as part2); false, with zero
// features, for entries past nx and for pairs of other types (they screen
// with 0).
__device__ __forceinline__ bool pair_features(const Stage& S, const TrialScratch& T, int nx, int e, int t,
                                              const uint4 uo, const uint4 un, const float* mvo, const float* mvn,
                                              const SysConst& sc, const MlpConst& mc, float* f, int& which) {
    bool net = false;
    which = 0;
    if (e < nx) {
        int w = T.xl[e];
        int jj = w & 0x3fffffff;
        which = (w >> 30) & 1;
        uint4 b = S.u[jj];
        net = t == mc.tnet && (int)b.w == mc.tnet;
        if (net) {
            V3<float> r = which ? fx_image_c(un.x, un.y, un.z, b.x, b.y, b.z, sc.Lh, sc.Ll)
                                : fx_image_c(uo.x, uo.y, uo.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
            const float* mv = which ? mvn : mvo;
            float4 ob = S.o[jj];
            float2 qb = S.q[jj];
            mlp_features(mv_vec(mv, 0), mv_vec(mv, 3), mk<float>(ob.x, ob.y, ob.z), mk<float>(ob.w, qb.x, qb.y), r,
                         T.xd[e], mc.lsc, mc.rcut, mc.pdis, mc.pswitch, f);
        }
    }
    if (!net)
        for (int q = 0; q < MLP_NF; q++) f[q] = 0.0f;
    return net;
}

// Delayed acceptance (Christen & Fox 2005, in stages): with DA != DA_OFF
// the move first faces Metropolis screens on cheaper energy changes, each
// accepting with min(1, exp(-(its change)/T)); returns false at the first
// screen that rejects (stage = the number of screens passed).  Otherwise eo/en
// are the exact energies of the two states and des the sum of the screened
// changes, which the final stage divides out: min(1, exp(-(dE - des)/T)).
// Chain bonds are exact cheap terms (FP64, in the first screen); a move whose
// bonded partner lies outside the staged stencil is rejected (stage -1): the
// partner might be moving in another cell of the pass, and the rejection does
// not depend on the move (the particle stays in its cell).
template <int DA, int NW, bool WS = false>
__device__ __forceinline__ bool trial_energies(const Stage& S, int k, const uint4 uo, const uint4 un,
                                               const float4 oo, const float2 qo, const float4 on, const float2 qn,
                                               const SysConst& sc,
                                               const PP32* __restrict__ pp, const TP32* __restrict__ tp, int ppmax,
                                               const TrialScratch& T, const float* ml, const MlpRegsW<4 / NW>& mr,
                                               const unsigned char* __restrict__ mlp_img,
                                               const MlpConst& mc, const ChainTab& ct, double thr1, double thr2,
                                               const float* hvk, bool translate, double& eo, double& en,
                                               double& des, int& stage, int ncap, int xcap) {
    static_assert(!WS || NW == 1, "a wave's own trial is a one-wave trial");
    const int tid = WS ? lane_id() : (int)threadIdx.x, lane = tid & (WAVE - 1), BS = WAVE * NW;
    int t = uo.w;
    bool moved = (uo.x != un.x) || (uo.y != un.y) || (uo.z != un.z);
    const int idk = S.id[k];
    int cwk = 0;
    if (ct.cw) cwk = ct.cw[idk];
    int bmask = cwk & 15;
    // parities of the one-barrier helpers; a trial starts after a barrier
    // that follows every read of the previous trial's helpers
    int cpar = 0, rpar = 0;
    // Derived vectors of both states (DPart order, MV_STRIDE floats each):
    // the old state's from the home cell's cache when there is one (else
    // derived into T.mv); a translation keeps them for the new state; for a
    // rotation the last wave derives the new state's into T.mv + MV_STRIDE
    // while the other waves scan the stencil (phase A needs positions only).
    const int wv = tid >> 6;
    const float* mvo = hvk;
    const float* mvn = T.mv + MV_STRIDE;
    V3<float> dold;
    if (hvk) {
        dold = mk<float>(hvk[0], hvk[1], hvk[2]);
    } else {
        DPart<float> po = load_part<float>(oo, qo, t, tp);
        if (tid == 0) mv_store(T.mv, po);
        dold = po.dir;
        mvo = T.mv;
    }
    if (translate) mvn = mvo;
    if (!translate && (NW == 1 || wv == NW - 1)) {
        DPart<float> pn = load_part<float>(on, qn, t, tp);
        if (lane == 0) mv_store(T.mv + MV_STRIDE, pn);
    }
    if (bmask && tid < 4) T.bp[tid] = -1;
    if (bmask) tsync<WS>();
    // phase A: the staged particles inside sqmaxcut at either state -> near
    // list, in stencil order (one barrier per round); all waves scan a
    // translation, the waves but the last one a rotation; a single wave takes
    // eight entries per lane (one round for a stencil of up to 512)
    int nn = 0;
    if (NW > 1 && !translate)
        scan_stencil<NW, NW >= 4 ? 3 : 8>(S, k, uo, un, moved, sc, bmask, idk, T, (NW - 1) * WAVE, wv < NW - 1, nn,
                                         cpar, ncap, tid);
    else
        scan_stencil<NW, NW >= 4 ? 2 : (NW == 1 ? 8 : 4)>(S, k, uo, un, moved, sc, bmask, idk, T, NW * WAVE, true,
                                                           nn, cpar, ncap, tid);
    tsync<WS>();
    if (nn > ncap) {   // a wave's near list is full: the trial cannot be evaluated here
        stage = -2;
        return false;
    }
    const V3<float> dnew = translate ? dold : mv_vec(mvn, 0);
    double ao = 0.0, an = 0.0;
    if (bmask) {
        int found = 0;
        for (int q = 0; q < 4; q++) found |= (T.bp[q] >= 0 ? 1 : 0) << q;
        if (found != bmask) {
            stage = -1;
            return false;
        }
        // threads 0-3: the old state with partners 0-3, threads 4-7 the new one
        int q = tid & 3, which = (tid >> 2) & 1;
        if (tid < 8 && ((bmask >> q) & 1)) {
            int j = T.bp[q];
            uint4 b = S.u[j];
            float4 ob = S.o[j];
            float2 qb = S.q[j];
            const float* mv = which ? mvn : mvo;
            double eb = bond_energy(ct.mol[cwk >> 8], ct.pp64[t * sc.nt + b.w], cw_rel(q),
                                    fx_r12(which ? un : uo, b, sc.Ld), unit64(mv[0], mv[1], mv[2]),
                                    unit64(mv[3], mv[4], mv[5]), unit64(ob.x, ob.y, ob.z), unit64(ob.w, qb.x, qb.y));
            if (which) an += eb; else ao += eb;
        }
    }
    // phase B
    int nx = 0;
    for (int base = 0; base < 2 * nn; base += BS) {
        int it = base + tid, jj = 0, which = 0;
        bool need = false;
        float dseg2 = 0.0f;
        if (it < 2 * nn) {
            int e = T.nl[it >> 1];
            which = it & 1;
            jj = e & 0x0fffffff;
            if ((e >> (28 + which)) & 1) {
                uint4 b = S.u[jj];
                // the pair in its canonical order, the higher ID as part1 (the
                // order of allToAll): the pair term is then one function of
                // the pair's state whichever particle moves, and the accepted
                // changes add up to allToAll's differences term by term
                const bool sw = S.id[jj] > idk;
                const PP32& P2 = sw ? pp[b.w * sc.nt + t] : pp[t * sc.nt + b.w];
                V3<float> r = which ? fx_image_c(un.x, un.y, un.z, b.x, b.y, b.z, sc.Lh, sc.Ll)
                                    : fx_image_c(uo.x, uo.y, uo.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
                if (sw) r = mk<float>(-r.x, -r.y, -r.z);
                float ev = 0.0f;
                if (P2.kind <= K_SCA && P2.kind >= 0) {
                    V3<float> dm = which ? dnew : dold, db = normalised(mk<float>(S.o[jj].x, S.o[jj].y, S.o[jj].z));
                    DPart<float> pa, pb;
                    pa.dir = sw ? db : dm;
                    pb.dir = sw ? dm : db;
                    need = scsc_cheap(P2, pa, pb, r, ev, dseg2);
                } else if (P2.kind == K_SPH_WCA || P2.kind == K_SPH_TAYLOR) {
                    ev = sphere_pair(P2, sqrtf(dot(r, r)));
                } else if (P2.kind != K_NONE) {
                    need = true;
                }
                if (which) an += (double)ev; else ao += (double)ev;
            }
        }
        int pos = block_compact1<NW>(need, nx, T.cnt, cpar);
        if (need && pos < xcap) {
            T.xl[pos] = jj | (which << 30);
            if constexpr (DA == DA_NET) T.xd[pos] = sqrtf(dseg2);
        }
    }
    tsync<WS>();
    if (nx > xcap) {   // the expensive list is full (only with a wave's own bounded lists)
        stage = -2;
        return false;
    }
    // screens of the delayed acceptance: (1) the exact cheap terms already
    // summed (repulsion, sphere pairs), (2) with DA_NET the network's patch
    // attraction, tiles of 16 pairs on MFMA.  With four waves the network runs
    // on the last wave while the other three compute phase C, and both meet
    // at the final sum: a move the network rejects is rejected there (its
    // phase C result is dropped), so the screen's decisions, and the chain of
    // states, are those of evaluating it first.  With fewer waves the
    // network's hidden units are split between the waves and it runs before
    // phase C.
    constexpr bool NET_PAR = DA == DA_NET && NW == 4;
    des = 0.0;
    stage = 2;
    double rep = 0.0;
    if constexpr (DA != DA_OFF) {
        double rn, ro;
        block_sum2<NW>(an, ao, T.red, rpar, rn, ro);
        rep = rn - ro;
        stage = 0;
        if (!(rep <= 0.0 || rep < thr1)) return false;
        des = rep;
        if constexpr (DA == DA_NET && !NET_PAR) {
            stage = 1;
            if (nx >= mc.minx) {   // block-uniform: the entries of both states, the same for the reverse move
            double so = 0.0, sn = 0.0;
            const int wv = tid >> 6;
            for (int t0 = 0; t0 < nx; t0 += MLP_TILE) {   // block-uniform: every wave in every tile
                float f[MLP_NF];
                int which = 0;
                bool net = pair_features(S, T, nx, t0 + (lane & (MLP_TILE - 1)), t, uo, un, mvo, mvn, sc, mc, f,
                                         which);
                float out;
                if constexpr (WS)   // a wave of its own: operands from global memory, as the last wave's screen
                    out = mlp_tile_g(mlp_img, ml, mlp_slots(f, lane >> 4), lane);
                else
                    out = mlp_tile_w<NW>(mr, ml, T.hx, mlp_slots(f, lane >> 4), lane, wv);
                if (lane < MLP_TILE && net) {
                    if (which) sn += (double)out; else so += (double)out;
                }
            }
            double dn, d0;
            block_sum2<NW>(sn, so, T.red, rpar, dn, d0);
            double ds = dn - d0;
            if (!(ds <= 0.0 || ds < thr2)) return false;
            des = rep + ds;
            }
        }
        // 2: through every screen; 3: past the first, the network not run
        stage = (DA == DA_NET && nx < mc.minx) ? 3 : 2;
    }
    // phase C: (expensive pair, patch pair, side) items.  The two wedge
    // intersections of a patch pair run on neighbouring lanes (side = lane
    // parity) and meet through a lane swap; the even lane adds the attraction.
    const int CB = NET_PAR ? 3 * WAVE : BS;   // threads of phase C
    if (NET_PAR && (tid >> 6) == 3) {
        // the network screen on the last wave: the whole network per tile,
        // its operands read from global memory (L1/L2) layer by layer
        double so = 0.0, sn = 0.0;
        for (int t0 = 0; t0 < (nx >= mc.minx ? nx : 0); t0 += MLP_TILE) {
            float f[MLP_NF];
            int which = 0;
            bool net = pair_features(S, T, nx, t0 + (lane & (MLP_TILE - 1)), t, uo, un, mvo, mvn, sc, mc, f, which);
            float out = mlp_tile_g(mlp_img, ml, mlp_slots(f, lane >> 4), lane);
            if (lane < MLP_TILE && net) {
                if (which) sn += (double)out; else so += (double)out;
            }
        }
        double ds = wave_sum_dpp(sn) - wave_sum_dpp(so);
        if (lane == 0) *T.netd = ds;
    } else {
    int items = nx * ppmax * 2;
    // a wave of its own whose items take more than one pass: each lane takes
    // both sides of its patch pair, one after the other (half the passes, and
    // the view, set-up and attraction once per patch pair; wave-uniform, and
    // the same values: each side is the same patch_side call)
    const bool dual = WS && items > CB;
    const int nit = dual ? items / 2 : items;
    for (int base = 0; base < nit; base += CB) {
        int it = dual ? 2 * (base + tid) : base + tid, which = 0, k1 = 0, k2 = 0, side = it & 1;
        bool hit = false, hit2 = false, pair_item = false;
        float x1 = 0.0f, x2 = 0.0f, ev = 0.0f, z1 = 0.0f, z2 = 0.0f;
        const PP32* P2 = pp;
        V3<float> r = mk<float>(0.0f, 0.0f, 0.0f);
        PatchView<float> va, vb;
        if (it < items) {
            int e = it / (2 * ppmax), mpp = (it - e * 2 * ppmax) >> 1;
            int w = T.xl[e];
            int jj = w & 0x3fffffff;
            which = (w >> 30) & 1;
            uint4 b = S.u[jj];
            const bool sw = S.id[jj] > idk;   // canonical order, as in phase B
            P2 = sw ? pp + b.w * sc.nt + t : pp + t * sc.nt + b.w;
            r = which ? fx_image_c(un.x, un.y, un.z, b.x, b.y, b.z, sc.Lh, sc.Ll)
                      : fx_image_c(uo.x, uo.y, uo.z, b.x, b.y, b.z, sc.Lh, sc.Ll);
            if (sw) r = mk<float>(-r.x, -r.y, -r.z);
            const float* mv = which ? mvn : mvo;
            if (P2->kind <= K_SCA) {
                if (mpp < npatch_pairs(*P2)) {
                    patch_pair_index(*P2, mpp, k1, k2);   // patches of part1 and part2
                    PatchView<float> vm = mv_view(mv, sw ? k2 : k1),
                                     vn = load_view<float>(S.o[jj], S.q[jj], b.w, tp, sw ? k1 : k2);
                    va = sw ? vn : vm;
                    vb = sw ? vm : vn;
                    hit = patch_side<float>(*P2, va, vb, r, k1, k2, side, x1, x2);
                    if (dual) hit2 = patch_side<float>(*P2, va, vb, r, k1, k2, 1, z1, z2);
                    pair_item = true;
                }
            } else if (mpp == 0 && side == 0) {
                DPart<float> pb = load_part<float>(S.o[jj], S.q[jj], b.w, tp);
                DPart<float> pm = mv_part(mv, t, tp[t].geotype);
                ev = sw ? spsc(*P2, pb, pm, r) : spsc(*P2, pm, pb, r);
            }
        }
        // the other side's result (lane ^ 1 holds item it ^ 1: CB is even)
        float y1 = z1, y2 = z2;
        int oh = hit2 ? 1 : 0;
        if (!dual) {
            y1 = __shfl_xor(x1, 1);
            y2 = __shfl_xor(x2, 1);
            oh = __shfl_xor(hit ? 1 : 0, 1);
        }
        if (pair_item && side == 0 && hit && oh) ev = patch_atr<float>(*P2, va, vb, r, k1, k2, x1, x2, y1, y2);
        if (which) an += (double)ev; else ao += (double)ev;
    }
    }
    block_sum2<NW>(ao, an, T.red, rpar, eo, en);
    if constexpr (NET_PAR) {
        stage = 1;
        double ds = *T.netd;   // written before the sum's barrier (0 when the network did not run)
        if (!(ds <= 0.0 || ds < thr2)) return false;
        des = rep + ds;
        stage = nx < mc.minx ? 3 : 2;
    }
    return true;
}

// One active cell of a checkerboard pass (logical index b of the colour's
// active cells): stage the stencil, run the cell's trial moves, write back.
// The proposal and thresholds of trial tt of a cell: counter (trial, cell,
// pass, stream), streams 0 (particle, move kind, translation), 1 (rotation),
// 2 (final and first-screen uniforms); the second screen's uniform uses
// stream 2 with the trial's top bit set.
template <int DA>
__device__ __forceinline__ TrialRec draw_trial(U4 ctr, int tt, const MoveParams& mp, const Stage& S, int nh,
                                               const TP32* __restrict__ tp, float Lx, float Ly, float Lz) {
    U4 c = ctr;
    c.x = (uint32_t)tt;
    c.w = mp.pass_hi << 2;
    U4 ra = philox(c, mp.key0, mp.key1);
    TrialRec R;
    R.k = S.h0 + (int)(((uint64_t)ra.x * (uint64_t)nh) >> 32);
    int g0 = tp[S.u[R.k].w].geotype;
    R.translate = (u01f(ra.y) < 0.5f) || (g0 >= 30);
    if (R.translate) {
        V3<float> d = unit_from(u01f(ra.z), u01f(ra.w));
        R.w[0] = (int32_t)rintf(d.x * mp.transmx * (4294967296.0f / Lx));
        R.w[1] = (int32_t)rintf(d.y * mp.transmx * (4294967296.0f / Ly));
        R.w[2] = (int32_t)rintf(d.z * mp.transmx * (4294967296.0f / Lz));
        R.w[3] = R.w[4] = 0;
    } else {
        c.w = (mp.pass_hi << 2) | 1u;
        U4 rb = philox(c, mp.key0, mp.key1);
        V3<float> axv = unit_from(u01f(rb.x), u01f(rb.y));
        float ang = mp.rot_half * u01f(rb.z);
        float vc = cosf(ang), vs = sqrtf(fmaxf(0.0f, 1.0f - vc * vc));
        if (u01f(rb.w) >= 0.5f) vs = -vs;   // Particle::pscRotate direction choice (particle.h:194-196)
        R.w[0] = __float_as_int(axv.x);
        R.w[1] = __float_as_int(axv.y);
        R.w[2] = __float_as_int(axv.z);
        R.w[3] = __float_as_int(vc);
        R.w[4] = __float_as_int(vs);
    }
    R.pad = 0;
    c.w = (mp.pass_hi << 2) | 2u;
    U4 rc = philox(c, mp.key0, mp.key1);
    R.thr = -mp.temper * log(u01d(rc.x, rc.y));
    R.thr1 = R.thr2 = R.pad2 = 0.0;
    if constexpr (DA != DA_OFF) {   // the screens' uniforms
        R.thr1 = -mp.temper * log(u01d(rc.z, rc.w));
        c.x |= 0x80000000u;
        U4 rt = philox(c, mp.key0, mp.key1);
        R.thr2 = -mp.temper * log(u01d(rt.x, rt.y));
    }
    return R;
}

// The trial state of record R from the particle's state (uo, oo, qo): a
// translation by R's fixed-point step (partDisplace), or the quaternion
// rotation of axis and patch direction (pscRotate), renormalised in FP32.
__device__ __forceinline__ void propose(const TrialRec& R, uint4 uo, float4 oo, float2 qo, uint4& un, float4& on,
                                        float2& qn) {
    un = uo;
    on = oo;
    qn = qo;
    if (R.translate) {
        un.x = uo.x + (uint32_t)R.w[0];
        un.y = uo.y + (uint32_t)R.w[1];
        un.z = uo.z + (uint32_t)R.w[2];
    } else {
        V3<float> axv = mk<float>(__int_as_float(R.w[0]), __int_as_float(R.w[1]), __int_as_float(R.w[2]));
        float vc = __int_as_float(R.w[3]), vs = __int_as_float(R.w[4]);
        V3<float> d = qrot(mk<float>(oo.x, oo.y, oo.z), axv, vc, vs);
        V3<float> pq = qrot(mk<float>(oo.w, qo.x, qo.y), axv, vc, vs);
        d = normalised(d);
        pq = normalised(perp(pq, d));
        on = make_float4(d.x, d.y, d.z, pq.x);
        qn = make_float2(pq.y, pq.z);
    }
}

template <int DA, int NW>
__device__ __forceinline__ void sweep_cell(int b, const Grid& g, const SysConst& sc, const MoveParams& mp,
                                           const int* __restrict__ starts, uint4* __restrict__ P,
                                           float4* __restrict__ O, float2* __restrict__ Q, const int* __restrict__ ID,
                                           const PP32* __restrict__ pp, const TP32* __restrict__ tp, int cap,
                                           int ppmax, const int* __restrict__ nonempty,
                                           const unsigned char* __restrict__ mlp_img, const MlpConst& mc,
                                           const ChainTab& ct, unsigned long long* __restrict__ stats,
                                           double* __restrict__ de_sum, int* __restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    int ax = b % mp.nact[0], ay = (b / mp.nact[0]) % mp.nact[1], az = b / (mp.nact[0] * mp.nact[1]);
    int cx = mp.ox + mp.p * ax, cy = mp.oy + mp.p * ay, cz = mp.oz + mp.p * az;
    int cell = cx + g.nc[0] * (cy + g.nc[1] * cz);
    int nh = starts[cell + 1] - starts[cell];
    if (nh == 0) return;
    Stage S;
    S.u = (uint4*)smem;
    S.o = (float4*)(S.u + cap);
    S.q = (float2*)(S.o + cap);
    S.id = (int*)(S.q + cap);
    TrialScratch T;
    T.nl = S.id + cap;
    T.xl = T.nl + cap;
    T.xd = (float*)(T.xl + 2 * cap);
    T.mv = DA == DA_NET ? T.xd + 2 * cap : (float*)(T.xl + 2 * cap);
    int* sten = (int*)(T.mv + 2 * MV_STRIDE);
    T.bp = sten + 2 * 125 + 2;
    T.cnt = T.bp + 4;
    T.red = (double*)(((uintptr_t)(T.cnt + 2 * CNT_HALF) + 7) & ~(uintptr_t)7);
    TrialRec* recs = (TrialRec*)(((uintptr_t)(T.red + 16) + 15) & ~(uintptr_t)15);
    float* hv = (float*)(recs + TRIAL_CHUNK);
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1);
    float* ml = nullptr;
    MlpRegsW<4 / NW> mr;   // this wave's share of the network's operands
    T.hx = nullptr;
    T.netd = nullptr;
    uintptr_t tail = (uintptr_t)(hv + HCAP * MV_STRIDE);
    if constexpr (DA == DA_NET) {
        ml = (float*)((tail + 15) & ~(uintptr_t)15);
        if constexpr (NW == 4)   // the last wave evaluates whole tiles, operands from global memory
            for (int i = lane; i < MLP_C_BYTES / 16; i += 64) ((uint4*)ml)[i] = ((const uint4*)(mlp_img + MLP_C_OFF))[i];
        else
            mlp_setup_w<4 / NW>(mlp_img, ml, lane, (tid >> 6) * (4 / NW), mr);
        tail = (uintptr_t)ml + MLP_LDS_BYTES;
        T.hx = (bf16x4*)((tail + 15) & ~(uintptr_t)15);
        tail = (uintptr_t)T.hx + MLP_HX_BYTES;
        T.netd = (double*)tail;
        tail += 16;
    }
#if SWEEP_TABLES_IN_LDS
    // the pair and type parameter tables in LDS: every pair term reads them,
    // and a global (L1/L2) load per field would stall the few waves of a CU
    if (sc.nt <= SWEEP_TABLE_MAXT) {
        PP32* ppl = (PP32*)((tail + 15) & ~(uintptr_t)15);
        TP32* tpl = (TP32*)(ppl + sc.nt * sc.nt);
        const int w4 = (int)(sizeof(PP32) / 4), t4 = (int)(sizeof(TP32) / 4);
        for (int e = tid; e < sc.nt * sc.nt * w4; e += WAVE * NW) ((int*)ppl)[e] = ((const int*)pp)[e];
        for (int e = tid; e < sc.nt * t4; e += WAVE * NW) ((int*)tpl)[e] = ((const int*)tp)[e];
        pp = ppl;
        tp = tpl;
    }
#endif
    if (!stage_load<NW>(S, cap, g, cx, cy, cz, starts, P, O, Q, ID, sten)) {
        if (tid == 0) atomicOr(flag, 1);
        return;
    }
    const float Lx = sc.L[0], Ly = sc.L[1], Lz = sc.L[2];
    U4 ctr;
    ctr.y = (uint32_t)cell;
    ctr.z = mp.pass_lo;
    // trials in this cell: its occupancy (every particle once per sweep on
    // average), or with trial_mode 1 the mean occupancy of the non-empty
    // cells, dithered, so the longest chain of a pass is not the fullest cell
    int ntrial = nh * mp.visit;
    if (mp.trial_mode == 1) {
        ctr.x = 0xffffffffu;
        ctr.w = (mp.pass_hi << 2) | 3u;
        U4 rd = philox(ctr, mp.key0, mp.key1);
        float mean = (float)mp.n * (float)mp.visit / (float)max(1, *nonempty);
        int fl = (int)floorf(mean);
        ntrial = fl + ((u01f(rd.x) < mean - (float)fl) ? 1 : 0);
    }
    unsigned n_tacc = 0, n_ttry = 0, n_racc = 0, n_rtry = 0, n_left = 0, n_s1try = 0, n_s1acc = 0, n_s2try = 0,
             n_s2acc = 0, n_bond = 0;   // per block: far below 2^32
    // derived vectors of the home particles (read after the first trial's barrier)
    const bool hcache = nh <= min(HCAP, mp.hcap);
    if (hcache)
        for (int i = tid; i < nh; i += WAVE * NW)
            mv_store(hv + i * MV_STRIDE, load_part<float>(S.o[S.h0 + i], S.q[S.h0 + i], S.u[S.h0 + i].w, tp));
    double dsum = 0.0;
    for (int t = 0; t < ntrial; t++) {
        if ((t & (TRIAL_CHUNK - 1)) == 0) {
            // the next TRIAL_CHUNK trials' proposals and thresholds, one per lane
            // of the first wave (the same draws as trial by trial: counter
            // (trial, cell, pass, stream))
            __syncthreads();
            int tt = t + tid;
            if (tid < TRIAL_CHUNK && tt < ntrial) {
                recs[tid] = draw_trial<DA>(ctr, tt, mp, S, nh, tp, Lx, Ly, Lz);
            }
            __syncthreads();
        }
        const TrialRec& R = recs[t & (TRIAL_CHUNK - 1)];
        int k = R.k;
        uint4 uo = S.u[k];
        float4 oo = S.o[k];
        float2 qo = S.q[k];
        bool translate = R.translate != 0;
        uint4 un;
        float4 on;
        float2 qn;
        propose(R, uo, oo, qo, un, on, qn);
        if (translate) {
            n_ttry++;
            if (cell_of_u(g, un.x, un.y, un.z) != cell) {   // leaving the active cell: rejected
                n_left++;
                continue;
            }
        } else {
            n_rtry++;
        }
        double eo, en, des;
        bool pass;
        int stage;
        const float* hvk = hcache ? hv + (k - S.h0) * MV_STRIDE : nullptr;
        pass = trial_energies<DA, NW>(S, k, uo, un, oo, qo, on, qn, sc, pp, tp, ppmax, T, ml, mr, mlp_img, mc, ct,
                                      R.thr1,
                                      R.thr2, hvk, translate, eo, en, des, stage, cap, 2 * cap);
        if (stage < 0) {   // a chain partner outside the stencil
            n_bond++;
            __syncthreads();
            continue;
        }
        if constexpr (DA != DA_OFF) {
            n_s1try++;
            if (stage >= 1) n_s1acc++;
            if constexpr (DA == DA_NET) {
                if (stage == 1 || stage == 2) n_s2try++;   // 3: the network did not run
                if (stage == 2) n_s2acc++;
            }
            if (!pass) { __syncthreads(); continue; }
        }
        double de = en - eo;
        // MoveCreator::moveTry (movecreator.h:175-187); with the surrogate the
        // second stage accepts with min(1, exp(-(dE - dE_s)/T)) (Christen & Fox 2005)
        double de2 = de - des;
        bool acc = de2 <= 0.0 || de2 < R.thr;
        if (acc) {
            if (translate) n_tacc++; else n_racc++;
            dsum += de;
            if (tid == 0) {
                S.u[k] = un;
                S.o[k] = on;
                S.q[k] = qn;
            }
            if (hcache && !translate && tid < 27) hv[(k - S.h0) * MV_STRIDE + tid] = T.mv[MV_STRIDE + tid];
        }
        __syncthreads();
    }
    // write the home cell back (same slots it was staged from)
    int gs = starts[cell];
    for (int e = tid; e < nh; e += WAVE * NW) {
        P[gs + e] = S.u[S.h0 + e];
        O[gs + e] = S.o[S.h0 + e];
        Q[gs + e] = S.q[S.h0 + e];
    }
    // counters: one atomic per lane into one of STAT_REP replicas (a cache
    // line each), so a thousand waves ending together do not queue on one line
    int rep = (int)(blockIdx.x & (STAT_REP - 1));
    if (tid < 10) {
        unsigned v = tid == 0 ? n_tacc : tid == 1 ? n_ttry : tid == 2 ? n_racc : tid == 3 ? n_rtry
                   : tid == 4 ? n_left : tid == 5 ? n_s1acc : tid == 6 ? n_s1try : tid == 7 ? n_s2acc
                   : tid == 8 ? n_s2try : n_bond;
        if (v) atomicAdd(&stats[rep * STAT_STRIDE + tid], (unsigned long long)v);
    } else if (tid == 10) {
        atomicAdd(&de_sum[rep * STAT_STRIDE], dsum);
    }
}

// NW waves per block work on one active cell (NW = 1, 2 or 4).  Every
// variant keeps to 256 registers per lane, so that two blocks of four waves
// share a CU (two waves per SIMD): a pass has ~2 active cells per CU.
template <int DA, int NW>
__global__ __launch_bounds__(WAVE * NW, 2) void k_sweep_pass(Grid g, SysConst sc, MoveParams mp, const int* __restrict__ starts,
                                                     uint4* __restrict__ P, float4* __restrict__ O,
                                                     float2* __restrict__ Q, const int* __restrict__ ID,
                                                     const PP32* __restrict__ pp, const TP32* __restrict__ tp, int cap,
                                                     int ppmax, const int* __restrict__ nonempty,
                                                     const unsigned char* __restrict__ mlp_img, MlpConst mc,
                                                     ChainTab ct, unsigned long long* __restrict__ stats,
                                                     double* __restrict__ de_sum, int* __restrict__ flag) {
    sweep_cell<DA, NW>(xcd_slab_index((int)blockIdx.x, (int)gridDim.x), g, sc, mp, starts, P, O, Q, ID, pp, tp, cap,
                       ppmax, nonempty, mlp_img, mc, ct, stats, de_sum, flag);
}

// ------------------------------------------------- speculative trials
// The chain of trials of a cell is sequential, and one trial leaves most of
// a block's lanes idle (its phases hold a few dozen items).  In this variant
// each of the block's SPEC_W waves evaluates one trial of the next SPEC_W
// against the cell's current state, all at once; wave 0 then takes the
// results in trial order and keeps each one unless it depends on an earlier
// trial of the round that was accepted (the same particle, chain partners, or
// two particles whose pair term can be non-zero in any of their old/new
// states, by the pair code's own cutoff tests), where the round stops: that
// trial and the later ones are evaluated again in the next round, against the
// updated state.  A kept
// trial saw exactly the state it would have seen in sequence, so the chain
// of states and decisions is the sequential sweep's (only the FP64 sums run
// in another order, as between the 1- and 4-wave sweeps).
#define SPEC_W 4
#define SPEC_LCAP 320   // near-list entries of one wave's trial (a stencil of ~450 has ~60 within maxcut)
#define SPEC_XCAP 192   // its (pair, state) entries within the attraction's range (typically 10-20)

struct __align__(16) SpecRes {
    uint4 uo, un;    // position at the round's start and proposed
    float4 on;       // proposed orientation record
    float2 qn;
    int k, id;       // staged index, particle ID
    double de;       // E_new - E_old (kept trials that reached the final stage)
    int status;      // 0 evaluated, 1 left the cell, 2 chain partner outside the stencil, 3 a list full
    int translate, acc, stage;
    int dep;         // bit i: the trial depends on the round's trial i, should that one be accepted
};

// LDS of the speculative sweep: the stage, then SPEC_W per-wave trial
// scratch blocks, then what the waves share (byte offsets; per-wave fields
// relative to the wave's block).
struct SpecLayout {
    size_t nl, xl, xd, mv, bp, cnt, red, netd, wstride;
    size_t wbase, sten, misc, res, recs, hv, ml, tab, end;
};
__host__ __device__ inline size_t spec_align(size_t x, size_t a) { return (x + a - 1) / a * a; }
__host__ __device__ inline SpecLayout spec_layout(int cap, int lcap, int xcap, bool net, size_t tab_bytes) {
    SpecLayout L;
    L.nl = 0;
    L.xl = L.nl + 4 * (size_t)lcap;
    L.xd = L.xl + 4 * (size_t)xcap;
    L.mv = spec_align(L.xd + (net ? 4 * (size_t)xcap : 0), 16);
    L.bp = L.mv + 4 * 2 * MV_STRIDE;
    L.cnt = L.bp + 16;
    L.red = spec_align(L.cnt + 4 * 2 * CNT_HALF, 16);
    L.netd = L.red + 8 * 16;
    L.wstride = spec_align(L.netd + 16, 16);
    L.wbase = spec_align((size_t)cap * (16 + 16 + 8 + 4), 16);   // the stage: u, o, q, id
    size_t o = L.wbase + SPEC_W * L.wstride;
    L.sten = o;
    o += 4 * (2 * 125 + 2);
    L.misc = spec_align(o, 16);
    o = L.misc + 4 * 16;
    L.res = spec_align(o, 16);
    o = L.res + SPEC_W * sizeof(SpecRes);
    L.recs = spec_align(o, 16);
    o = L.recs + TRIAL_CHUNK * sizeof(TrialRec);
    L.hv = spec_align(o, 16);
    o = L.hv + 4 * (size_t)HCAP * MV_STRIDE;
    L.ml = spec_align(o, 16);
    o = L.ml + (net ? MLP_C_BYTES : 0);
    L.tab = spec_align(o, 16);
    L.end = L.tab + tab_bytes;
    return L;
}

// Can the pair of particles 1 (position u1, orientation record o1, ID id1)
// and 2 have a non-zero pair term?  The canonical evaluation's own FP32 tests
// (higher ID as part1): no beyond sqmaxcut (phase A), none for a type pair
// without a functor, and for spherocylinder pairs none when the segment
// distance is beyond both the attraction's and the repulsion's range (phase
// B's scsc_cheap; a relative margin of 1e-5 covers the K_SCA term's
// square-root test).  Sphere pairs: any pair inside sqmaxcut.
__device__ __forceinline__ bool pair_active(uint4 u1, float4 o1, int id1, uint4 u2, float4 o2, int id2,
                                            const SysConst& sc, const PP32* __restrict__ pp) {
    if (id1 < id2) {
        uint4 tu = u1; u1 = u2; u2 = tu;
        float4 to = o1; o1 = o2; o2 = to;
    }
    V3<float> r = fx_image_c(u1.x, u1.y, u1.z, u2.x, u2.y, u2.z, sc.Lh, sc.Ll);
    if (dot(r, r) > sc.sqmaxcut) return false;
    const PP32& P = pp[u1.w * sc.nt + u2.w];
    if (P.kind == K_NONE) return false;
    if (P.kind > K_SCA || P.kind < 0) return true;
    V3<float> dv = min_dist(normalised(mk<float>(o1.x, o1.y, o1.z)), normalised(mk<float>(o2.x, o2.y, o2.z)), P.hl0,
                            P.hl1, r);
    float d2 = dot(dv, dv);
    return d2 <= fmaxf(P.rcutSq, P.rcutwcaSq) * 1.00001f;
}

// Every trial's dependences of a round, on one wave (wave 0, whose own trial
// has none, so the other waves start their trials at once).  Bit i of
// deps[q]: trial q and the round's earlier trial i would see each other's
// move: the same particle, chain
// partners (bond and angle terms reach beyond maxcut), or a pair that can
// interact in any of their old/new states; the relation is symmetric.  Lanes
// 4p .. 4p+3 take the p-th pair (q, i) of (1,0) (2,0) (2,1) (3,0) (3,1)
// (3,2) and its four state combinations.
__device__ __forceinline__ void spec_deps_all(const TrialRec* recs, int t, int W, const Stage& S, const SysConst& sc,
                                              const PP32* __restrict__ pp, bool bonded, int* deps) {
    static_assert(SPEC_W <= 4, "six trial pairs on 24 lanes");
    const int lane = lane_id(), p = lane >> 2, c = lane & 3;
    const int q = p < 1 ? 1 : (p < 3 ? 2 : 3), i = p - (q * (q - 1)) / 2;
    bool d = false;
    if (p < 6 && q < W) {
        const TrialRec& Rq = recs[(t + q) & (TRIAL_CHUNK - 1)];
        const TrialRec& Ra = recs[(t + i) & (TRIAL_CHUNK - 1)];
        const int k = Rq.k, ka = Ra.k;
        if (ka == k || (bonded && abs(S.id[ka] - S.id[k]) <= 2)) {
            d = true;
        } else {
            uint4 uo = S.u[k], un, uoa = S.u[ka], una;
            float4 oo = S.o[k], on, ooa = S.o[ka], ona;
            float2 qo = S.q[k], qn, qoa = S.q[ka], qna;
            propose(Rq, uo, oo, qo, un, on, qn);
            propose(Ra, uoa, ooa, qoa, una, ona, qna);
            d = pair_active((c & 1) ? una : uoa, (c & 1) ? ona : ooa, S.id[ka], (c & 2) ? un : uo, (c & 2) ? on : oo,
                            S.id[k], sc, pp);
        }
    }
    const unsigned long long m = __ballot(d);
#pragma unroll
    for (int qq = 1; qq < SPEC_W; qq++) {
        int dep = 0;
#pragma unroll
        for (int ii = 0; ii < qq; ii++)
            if ((m >> (4 * ((qq * (qq - 1)) / 2 + ii))) & 15ull) dep |= 1 << ii;
        deps[qq] = dep;
    }
}

// LT: the pair and type tables are copied into LDS (systems of up to
// SWEEP_TABLE_MAXT types).  A compile-time choice, so their loads are LDS
// loads: chosen at run time, the pointer could be either, and every table
// load would be a flat load that waits on both memory counters.
template <int DA, bool LT>
__device__ __forceinline__ void sweep_cell_spec(int b, const Grid& g, const SysConst& sc, const MoveParams& mp,
                                                const int* __restrict__ starts, uint4* __restrict__ P,
                                                float4* __restrict__ O, float2* __restrict__ Q,
                                                const int* __restrict__ ID, const PP32* __restrict__ pp,
                                                const TP32* __restrict__ tp, int cap, int lcap, int ppmax,
                                                const int* __restrict__ nonempty,
                                                const unsigned char* __restrict__ mlp_img, const MlpConst& mc,
                                                const ChainTab& ct, unsigned long long* __restrict__ stats,
                                                double* __restrict__ de_sum, int* __restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool NET = DA == DA_NET;
    int ax = b % mp.nact[0], ay = (b / mp.nact[0]) % mp.nact[1], az = b / (mp.nact[0] * mp.nact[1]);
    int cx = mp.ox + mp.p * ax, cy = mp.oy + mp.p * ay, cz = mp.oz + mp.p * az;
    int cell = cx + g.nc[0] * (cy + g.nc[1] * cz);
    int nh = starts[cell + 1] - starts[cell];
    if (nh == 0) return;
    const SpecLayout L = spec_layout(cap, lcap, SPEC_XCAP, NET, 0);
    Stage S;
    S.u = (uint4*)smem;
    S.o = (float4*)(S.u + cap);
    S.q = (float2*)(S.o + cap);
    S.id = (int*)(S.q + cap);
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1), w = tid >> 6;
    unsigned char* wb = smem + L.wbase + (size_t)w * L.wstride;
    TrialScratch T;
    T.nl = (int*)(wb + L.nl);
    T.xl = (int*)(wb + L.xl);
    T.xd = (float*)(wb + L.xd);
    T.mv = (float*)(wb + L.mv);
    T.bp = (int*)(wb + L.bp);
    T.cnt = (int*)(wb + L.cnt);
    T.red = (double*)(wb + L.red);
    T.hx = nullptr;
    T.netd = (double*)(wb + L.netd);
    int* sten = (int*)(smem + L.sten);
    int* misc = (int*)(smem + L.misc);
    SpecRes* res = (SpecRes*)(smem + L.res);
    TrialRec* recs = (TrialRec*)(smem + L.recs);
    float* hv = (float*)(smem + L.hv);
    float* ml = nullptr;
    MlpRegsW<4> mr;   // not read: a wave's network screen takes its operands from global memory
    if constexpr (NET) {
        ml = (float*)(smem + L.ml);
        for (int i = tid; i < MLP_C_BYTES / 16; i += WAVE * SPEC_W) ((uint4*)ml)[i] = ((const uint4*)(mlp_img + MLP_C_OFF))[i];
    }
    if constexpr (LT) {
        PP32* ppl = (PP32*)(smem + L.tab);
        TP32* tpl = (TP32*)(ppl + sc.nt * sc.nt);
        const int w4 = (int)(sizeof(PP32) / 4), t4 = (int)(sizeof(TP32) / 4);
        for (int e = tid; e < sc.nt * sc.nt * w4; e += WAVE * SPEC_W) ((int*)ppl)[e] = ((const int*)pp)[e];
        for (int e = tid; e < sc.nt * t4; e += WAVE * SPEC_W) ((int*)tpl)[e] = ((const int*)tp)[e];
        pp = ppl;
        tp = tpl;
    }
    if (!stage_load<SPEC_W>(S, cap, g, cx, cy, cz, starts, P, O, Q, ID, sten)) {
        if (tid == 0) atomicOr(flag, 1);
        return;
    }
    const float Lx = sc.L[0], Ly = sc.L[1], Lz = sc.L[2];
    U4 ctr;
    ctr.y = (uint32_t)cell;
    ctr.z = mp.pass_lo;
    int ntrial = nh * mp.visit;
    if (mp.trial_mode == 1) {
        ctr.x = 0xffffffffu;
        ctr.w = (mp.pass_hi << 2) | 3u;
        U4 rd = philox(ctr, mp.key0, mp.key1);
        float mean = (float)mp.n * (float)mp.visit / (float)max(1, *nonempty);
        int fl = (int)floorf(mean);
        ntrial = fl + ((u01f(rd.x) < mean - (float)fl) ? 1 : 0);
    }
    const bool hcache = nh <= min(HCAP, mp.hcap);
    if (hcache)
        for (int i = tid; i < nh; i += WAVE * SPEC_W)
            mv_store(hv + i * MV_STRIDE, load_part<float>(S.o[S.h0 + i], S.q[S.h0 + i], S.u[S.h0 + i].w, tp));
    // wave 0 keeps the counters of the kept trials (the same on all its lanes)
    unsigned n_tacc = 0, n_ttry = 0, n_racc = 0, n_rtry = 0, n_left = 0, n_s1try = 0, n_s1acc = 0, n_s2try = 0,
             n_s2acc = 0, n_bond = 0;
    double dsum = 0.0;
    const bool bonded = ct.cw != nullptr;
    int t = 0, drawn = 0;
    while (t < ntrial) {   // t is block-uniform (from LDS)
        if (t >= drawn) {
            __syncthreads();   // the previous records are no longer read
            int tt = drawn + tid;
            if (tid < TRIAL_CHUNK && tt < ntrial) recs[tid] = draw_trial<DA>(ctr, tt, mp, S, nh, tp, Lx, Ly, Lz);
            drawn += TRIAL_CHUNK;
            __syncthreads();
        }
        const int W = min(SPEC_W, min(ntrial - t, drawn - t));
        if (w < W) {
            if (w == 0) {   // every trial's dependences, then its own trial
                int deps[SPEC_W];
                deps[0] = 0;
                spec_deps_all(recs, t, W, S, sc, pp, bonded, deps);
                if (lane == 0) {
#pragma unroll
                    for (int q = 0; q < SPEC_W; q++)
                        if (q < W) res[q].dep = deps[q];
                }
            }
            const TrialRec& R = recs[(t + w) & (TRIAL_CHUNK - 1)];
            int k = R.k;
            uint4 uo = S.u[k];
            float4 oo = S.o[k];
            float2 qo = S.q[k];
            bool translate = R.translate != 0;
            uint4 un;
            float4 on;
            float2 qn;
            propose(R, uo, oo, qo, un, on, qn);
            int status = 0, stage = 0;
            bool acc = false;
            double de = 0.0;
            if (translate && cell_of_u(g, un.x, un.y, un.z) != cell) {   // leaving the active cell: rejected
                status = 1;
            } else {
                double eo, en, des;
                const float* hvk = hcache ? hv + (k - S.h0) * MV_STRIDE : nullptr;
                bool pass = trial_energies<DA, 1, true>(S, k, uo, un, oo, qo, on, qn, sc, pp, tp, ppmax, T, ml, mr,
                                                        mlp_img, mc, ct, R.thr1, R.thr2, hvk, translate, eo, en, des,
                                                        stage, lcap, SPEC_XCAP);
                if (stage == -1) {
                    status = 2;
                } else if (stage == -2) {
                    status = 3;
                } else if (pass) {
                    de = en - eo;
                    // MoveCreator::moveTry (movecreator.h:175-187), with the screens' changes divided out
                    double de2 = de - des;
                    acc = de2 <= 0.0 || de2 < R.thr;
                }
            }
            if (lane == 0) {
                SpecRes& r = res[w];
                r.uo = uo;
                r.un = un;
                r.on = on;
                r.qn = qn;
                r.k = k;
                r.id = S.id[k];
                r.de = de;
                r.status = status;
                r.translate = translate ? 1 : 0;
                r.acc = acc ? 1 : 0;
                r.stage = stage;
            }
        }
        __syncthreads();
        if (w == 0) {
            // wave 0 takes the round's results, a lane each, in trial order:
            // kept are the trials before the first that depends on an earlier
            // accepted one (or overflowed); the counters of the kept trials are
            // ballot counts; each kept accepted trial commits its own particle
            // (distinct particles: the same particle is a dependence)
            const bool valid = lane < W;
            const SpecRes& a = res[valid ? lane : 0];
            const int st = valid ? a.status : 0, dep = valid ? a.dep : 0, stg = valid ? a.stage : 0;
            const bool acc = valid && a.acc, tr = valid && a.translate;
            const unsigned long long accm = __ballot(acc);
            const bool depends = valid && (dep & (int)accm & ((1 << lane) - 1)) != 0;
            const bool full = valid && st == 3;
            const unsigned long long fullm = __ballot(full);
            const unsigned long long stopm = __ballot(depends || full);
            const int first = stopm ? __ffsll((long long)stopm) - 1 : W;
            const bool kept = valid && lane < first;
            const bool left = kept && tr && st == 1, bond = kept && !left && st == 2;
            const bool ev = kept && !left && !bond;   // reached the acceptance decision
            n_ttry += __popcll(__ballot(kept && tr));
            n_rtry += __popcll(__ballot(kept && !tr));
            n_left += __popcll(__ballot(left));
            n_bond += __popcll(__ballot(bond));
            if constexpr (DA != DA_OFF) {
                n_s1try += __popcll(__ballot(ev));
                n_s1acc += __popcll(__ballot(ev && stg >= 1));
                if constexpr (DA == DA_NET) {
                    n_s2try += __popcll(__ballot(ev && (stg == 1 || stg == 2)));   // 3: the network did not run
                    n_s2acc += __popcll(__ballot(ev && stg == 2));
                }
            }
            const bool com = ev && acc;
            n_tacc += __popcll(__ballot(com && tr));
            n_racc += __popcll(__ballot(com && !tr));
            if (com) {
                S.u[a.k] = a.un;
                S.o[a.k] = a.on;
                S.q[a.k] = a.qn;
            }
            const double dei = com ? a.de : 0.0;
#pragma unroll
            for (int q = 0; q < SPEC_W; q++) dsum += __shfl(dei, q, WAVE);   // in trial order
            if (lane == 0) {
                misc[0] = first;
                misc[1] = (first < W && ((fullm >> first) & 1ull)) ? 1 : 0;
            }
        }
        __syncthreads();
        const int nfin = misc[0];
        if (hcache && w < nfin && lane < 27) {   // the derived vectors of a kept, accepted rotation
            const SpecRes& a = res[w];
            if (a.acc && !a.translate) hv[(a.k - S.h0) * MV_STRIDE + lane] = T.mv[MV_STRIDE + lane];
        }
        if (misc[1]) {   // a trial's near list overflowed: stop the cell, report
            if (tid == 0) atomicOr(flag, 16);
            break;
        }
        t += nfin;
        __syncthreads();
    }
    __syncthreads();
    // write the home cell back (same slots it was staged from)
    int gs = starts[cell];
    for (int e = tid; e < nh; e += WAVE * SPEC_W) {
        P[gs + e] = S.u[S.h0 + e];
        O[gs + e] = S.o[S.h0 + e];
        Q[gs + e] = S.q[S.h0 + e];
    }
    if (tid == 0) {
        int rep = (int)(blockIdx.x & (STAT_REP - 1));
        const unsigned v[10] = {n_tacc, n_ttry, n_racc, n_rtry, n_left, n_s1acc, n_s1try, n_s2acc, n_s2try, n_bond};
        for (int q = 0; q < 10; q++)
            if (v[q]) atomicAdd(&stats[rep * STAT_STRIDE + q], (unsigned long long)v[q]);
        atomicAdd(&de_sum[rep * STAT_STRIDE], dsum);
    }
}

template <int DA, bool LT>
__global__ __launch_bounds__(WAVE * SPEC_W, 2) void k_sweep_spec(Grid g, SysConst sc, MoveParams mp,
                                                             const int* __restrict__ starts, uint4* __restrict__ P,
                                                             float4* __restrict__ O, float2* __restrict__ Q,
                                                             const int* __restrict__ ID,
                                                             const PP32* __restrict__ pp,
                                                             const TP32* __restrict__ tp, int cap, int lcap,
                                                             int ppmax, const int* __restrict__ nonempty,
                                                             const unsigned char* __restrict__ mlp_img, MlpConst mc,
                                                             ChainTab ct, unsigned long long* __restrict__ stats,
                                                             double* __restrict__ de_sum, int* __restrict__ flag) {
    sweep_cell_spec<DA, LT>(xcd_slab_index((int)blockIdx.x, (int)gridDim.x), g, sc, mp, starts, P, O, Q, ID, pp, tp, cap,
                        lcap, ppmax, nonempty, mlp_img, mc, ct, stats, de_sum, flag);
}

// Per-replica parameters of one checkerboard pass of a replica batch.
struct RepPass {
    Grid g;
    SysConst sc;
    MoveParams mp;
    int cell_off;   // the replica's first cell in the batch's cell arrays
    int blk_off;    // the replica's first block in this pass's launch
    int rep;        // replica index (counters, non-empty cells)
};

// One colour pass of every replica of a batch in one launch: the launch's
// blocks are the replicas' active cells back to back; a block finds its
// replica by binary search over blk_off (a replica without active cells
// shares its offset with the next one, and the search takes the last).
template <int DA, int NW>
__global__ __launch_bounds__(WAVE * NW, 2) void k_sweep_batch(const RepPass* __restrict__ rp, int nrep,
                                                      const int* __restrict__ starts, uint4* __restrict__ P,
                                                      float4* __restrict__ O, float2* __restrict__ Q,
                                                      const int* __restrict__ ID, const PP32* __restrict__ pp,
                                                      const TP32* __restrict__ tp, int cap, int ppmax,
                                                      const int* __restrict__ nonempty,
                                                      const unsigned char* __restrict__ mlp_img, MlpConst mc,
                                                      ChainTab ct, unsigned long long* __restrict__ stats,
                                                      double* __restrict__ de_sum, int* __restrict__ flag) {
    int lb = xcd_slab_index((int)blockIdx.x, (int)gridDim.x);
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rp[mid].blk_off <= lb) lo = mid; else hi = mid - 1;
    }
    const RepPass R = rp[lo];
    sweep_cell<DA, NW>(lb - R.blk_off, R.g, R.sc, R.mp, starts + R.cell_off, P, O, Q, ID, pp, tp, cap, ppmax,
                       nonempty + R.rep, mlp_img, mc, ct, stats + (size_t)R.rep * STAT_REP * STAT_STRIDE,
                       de_sum + (size_t)R.rep * STAT_REP * STAT_STRIDE, flag);
}

// The speculative sweep of every replica's active cells of one colour pass.
template <int DA, bool LT>
__global__ __launch_bounds__(WAVE * SPEC_W, 2) void k_sweep_batch_spec(const RepPass* __restrict__ rp, int nrep,
                                                                   const int* __restrict__ starts,
                                                                   uint4* __restrict__ P, float4* __restrict__ O,
                                                                   float2* __restrict__ Q,
                                                                   const int* __restrict__ ID,
                                                                   const PP32* __restrict__ pp,
                                                                   const TP32* __restrict__ tp, int cap, int lcap,
                                                                   int ppmax, const int* __restrict__ nonempty,
                                                                   const unsigned char* __restrict__ mlp_img,
                                                                   MlpConst mc, ChainTab ct,
                                                                   unsigned long long* __restrict__ stats,
                                                                   double* __restrict__ de_sum,
                                                                   int* __restrict__ flag) {
    int lb = xcd_slab_index((int)blockIdx.x, (int)gridDim.x);
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rp[mid].blk_off <= lb) lo = mid; else hi = mid - 1;
    }
    const RepPass R = rp[lo];
    sweep_cell_spec<DA, LT>(lb - R.blk_off, R.g, R.sc, R.mp, starts + R.cell_off, P, O, Q, ID, pp, tp, cap, lcap, ppmax,
                        nonempty + R.rep, mlp_img, mc, ct, stats + (size_t)R.rep * STAT_REP * STAT_STRIDE,
                        de_sum + (size_t)R.rep * STAT_REP * STAT_STRIDE, flag);
}

// cell list of a replica batch: the cell of every particle in its replica's
// grid, offset into the batch's cell arrays.  The replica of a particle is
// found from its index (replicas are index ranges part_off..), by binary
// search; after the build the replicas' cells, and so their particles, lie
// contiguous and in replica order.
struct RepGrid {
    Grid g;
    int cell_off;
    int part_off;
};
__global__ void k_cell_count_batch(int n, int nrep, const RepGrid* __restrict__ rg, const uint4* __restrict__ P,
                                   const int* __restrict__ ID, int* __restrict__ cell_tmp, int* __restrict__ counts) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int id = ID[k];
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rg[mid].part_off <= id) lo = mid; else hi = mid - 1;
    }
    uint4 u = P[k];
    int c = rg[lo].cell_off + cell_of_u(rg[lo].g, u.x, u.y, u.z);
    cell_tmp[k] = c;
    atomicAdd(&counts[c], 1);
}

// non-empty cells of each replica (one block per replica)
__global__ void k_nonempty_batch(const RepGrid* __restrict__ rg, int nrep, int ncell_total,
                                 const int* __restrict__ starts, int* __restrict__ nonempty) {
    int r = blockIdx.x;
    int c0 = rg[r].cell_off, c1 = r + 1 < nrep ? rg[r + 1].cell_off : ncell_total;
    int cnt = 0;
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x) cnt += starts[c + 1] > starts[c] ? 1 : 0;
    double v = wave_sum((double)cnt);
    __shared__ int part[16];
    if (lane_id() == 0) part[threadIdx.x >> 6] = (int)v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
        nonempty[r] = t;
    }
}

// ---------------------------------------------------------------- surrogate
// Training data: random pairs of one type pair with uniform orientations and
// closest segment distance d in [0, dmax]; a fraction `near_frac` is drawn with
// d <= 1.1 rcut where the attraction lives.  Target: the exact FP32 patch
// attraction (0 beyond rcut); also the exact full pair energy.
__global__ void k_mlp_gen(long ns, uint32_t key0, uint32_t key1, const PP32* __restrict__ ppair,
                          const TP32* __restrict__ tpa, const TP32* __restrict__ tpb, MlpConst mc, float dmax,
                          float near_frac, float* __restrict__ feats, float* __restrict__ target,
                          float* __restrict__ efull) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const PP32& P = *ppair;
    float R = P.hl0 + P.hl1 + dmax;
    U4 c;
    c.x = (uint32_t)s; c.y = (uint32_t)(s >> 32); c.z = 0; c.w = 0x5a17u;
    U4 h = philox(c, key0, key1);
    bool want_near = u01f(h.x) < near_frac;
    float dlim = want_near ? fminf(dmax, 1.1f * P.rcut) : dmax;
    V3<float> r, da, qa, db, qb;
    float d = 0.0f;
    for (int attempt = 0; attempt < 256; attempt++) {
        c.z = (uint32_t)attempt;
        U4 u = philox(c, key0, key1);
        c.w = 0x5a18u;
        U4 w = philox(c, key0, key1);
        c.w = 0x5a17u;
        // centre offset uniform in a ball of radius R
        V3<float> dir = unit_from(u01f(u.x), u01f(u.y));
        float rad = R * cbrtf(u01f(u.z));
        r = mk<float>(dir.x * rad, dir.y * rad, dir.z * rad);
        da = unit_from(u01f(u.w), u01f(w.x));
        db = unit_from(u01f(w.y), u01f(w.z));
        // patch directions uniform around the axes
        V3<float> ref = fabsf(da.x) < 0.9f ? mk<float>(1, 0, 0) : mk<float>(0, 1, 0);
        V3<float> e1 = normalised(perp(ref, da)), e2 = cross(da, e1);
        float ph = 6.283185307179586f * u01f(w.w);
        qa = mk<float>(cosf(ph) * e1.x + sinf(ph) * e2.x, cosf(ph) * e1.y + sinf(ph) * e2.y, cosf(ph) * e1.z + sinf(ph) * e2.z);
        ref = fabsf(db.x) < 0.9f ? mk<float>(1, 0, 0) : mk<float>(0, 1, 0);
        e1 = normalised(perp(ref, db));
        e2 = cross(db, e1);
        ph = 6.283185307179586f * u01f(u.x ^ w.y);
        qb = mk<float>(cosf(ph) * e1.x + sinf(ph) * e2.x, cosf(ph) * e1.y + sinf(ph) * e2.y, cosf(ph) * e1.z + sinf(ph) * e2.z);
        V3<float> dv = min_dist(da, db, P.hl0, P.hl1, r);
        d = sqrtf(dot(dv, dv));
        if (d <= dlim) break;
    }
    DPart<float> pa, pb;
    pa.dir = da; pa.pd[0] = qa; pa.pd[1] = qa; pa.sd[0] = pa.sd[1] = pa.sd[2] = pa.sd[3] = qa; pa.t = 0;
    pb.dir = db; pb.pd[0] = qb; pb.pd[1] = qb; pb.sd[0] = pb.sd[1] = pb.sd[2] = pb.sd[3] = qb; pb.t = 0;
    derive<float>(pa, *tpa);
    derive<float>(pb, *tpb);
    float erep, d2;
    bool att = scsc_cheap(P, pa, pb, r, erep, d2);
    float ea = att ? scsc_attr(P, pa, pb, r) : 0.0f;
    float f[MLP_NF];
    mlp_features(pa.dir, pa.pd[0], pb.dir, pb.pd[0], r, sqrtf(d2), mc.lsc, mc.rcut, mc.pdis, mc.pswitch, f);
    for (int q = 0; q < MLP_NF; q++) feats[s * MLP_NF + q] = f[q];
    target[s] = ea;
    efull[s] = erep + ea;
}

// ------------------------------------------------------------- training
// One training step is two launches.  k_mlp_fwdbwd: each block takes TR_S
// samples through the forward pass and back, with the weights and every
// activation in LDS, and writes its share of the gradient (and of the
// squared error) to its own row of `part`.  k_mlp_reduce_adam sums the rows
// in a fixed order (the gradient is deterministic) and takes the Adam step.
// FP32 throughout.  The hidden weight matrices sit in LDS with a row stride
// of 65 floats, so the backward pass, whose lanes walk down a column, hits
// 64 different banks.  Eight waves per block (two per SIMD) hide the LDS
// latency; lane j of wave w owns column j of rows w, w + 8, ...
#define TR_S 32
#define TR_NW 8
#define TR_WS 65
constexpr int TL_W1 = 0, TL_B1 = TL_W1 + MLP_NF * MLP_NH, TL_W2 = TL_B1 + MLP_NH, TL_B2 = TL_W2 + MLP_NH * TR_WS,
              TL_W3 = TL_B2 + MLP_NH, TL_B3 = TL_W3 + MLP_NH * TR_WS, TL_W4 = TL_B3 + MLP_NH, TL_B4 = TL_W4 + MLP_NH,
              TL_X = TL_B4 + 4, TL_Z1 = TL_X + TR_S * MLP_NF, TL_H1 = TL_Z1 + TR_S * MLP_NH,
              TL_Z2 = TL_H1 + TR_S * MLP_NH, TL_H2 = TL_Z2 + TR_S * MLP_NH, TL_Z3 = TL_H2 + TR_S * MLP_NH,
              TL_H3 = TL_Z3 + TR_S * MLP_NH, TL_DO = TL_H3 + TR_S * MLP_NH, TL_LS = TL_DO + TR_S,
              TL_END = TL_LS + TR_NW;
constexpr size_t TRAIN_LDS_BYTES = sizeof(float) * TL_END;
// a block's row of `part`: the gradient, then its sum of squared errors (rows 256-byte aligned)
constexpr int TR_PART = ((MLP_NPARAM + 1 + 63) / 64) * 64;
constexpr int TR_ROWS = TR_S / TR_NW;   // sample rows per thread

// z = b + in W (K inputs, W row stride ws), h = SiLU(z), for the block's TR_S
// x 64 outputs; the inputs of a row are read four at a time (K a multiple
// of 4, rows 16-byte aligned)
template <int K>
__device__ __forceinline__ void tr_dense(const float* in, const float* w, int ws, const float* b, float* z, float* h) {
    int tid = (int)threadIdx.x, j = tid & 63, s0 = tid >> 6;
    float acc[TR_ROWS];
#pragma unroll
    for (int r = 0; r < TR_ROWS; r++) acc[r] = b[j];
    for (int k = 0; k < K; k += 4) {
        float w0 = w[k * ws + j], w1 = w[(k + 1) * ws + j], w2 = w[(k + 2) * ws + j], w3 = w[(k + 3) * ws + j];
#pragma unroll
        for (int r = 0; r < TR_ROWS; r++) {
            float4 x = *(const float4*)&in[(s0 + TR_NW * r) * K + k];
            acc[r] += x.x * w0 + x.y * w1 + x.z * w2 + x.w * w3;
        }
    }
#pragma unroll
    for (int r = 0; r < TR_ROWS; r++) {
        z[(s0 + TR_NW * r) * MLP_NH + j] = acc[r];
        h[(s0 + TR_NW * r) * MLP_NH + j] = silu(acc[r]);
    }
}

// dZprev = (dZ W^T) * SiLU'(Zprev), in place over Zprev; lane k = tid & 63
__device__ __forceinline__ void tr_back(const float* dz, const float* w, float* zprev) {
    int tid = (int)threadIdx.x, k = tid & 63, s0 = tid >> 6;
    float acc[TR_ROWS];
#pragma unroll
    for (int r = 0; r < TR_ROWS; r++) acc[r] = 0.0f;
    for (int j = 0; j < MLP_NH; j += 4) {
        float w0 = w[k * TR_WS + j], w1 = w[k * TR_WS + j + 1], w2 = w[k * TR_WS + j + 2], w3 = w[k * TR_WS + j + 3];
#pragma unroll
        for (int r = 0; r < TR_ROWS; r++) {
            float4 d = *(const float4*)&dz[(s0 + TR_NW * r) * MLP_NH + j];
            acc[r] += d.x * w0 + d.y * w1 + d.z * w2 + d.w * w3;
        }
    }
#pragma unroll
    for (int r = 0; r < TR_ROWS; r++) {
        float& zp = zprev[(s0 + TR_NW * r) * MLP_NH + k];
        zp = acc[r] * silu_grad(zp);
    }
}

// gradient block dW[k][j] = sum_s in[s][k] dz[s][j] (K rows) and db[j] =
// sum_s dz[s][j]: lane j; wave w takes rows w R ... w R + R - 1 (read four
// at a time) when K is a multiple of 32, else rows w, w + 8, ...
template <int K>
__device__ __forceinline__ void tr_grad(const float* in, const float* dz, float* out_w, float* out_b) {
    int tid = (int)threadIdx.x, j = tid & 63, wv = tid >> 6;
    constexpr bool blocked = K % (4 * TR_NW) == 0;
    constexpr int R = blocked ? K / TR_NW : (K + TR_NW - 1) / TR_NW;
    float acc[R];
#pragma unroll
    for (int i = 0; i < R; i++) acc[i] = 0.0f;
    float bs = 0.0f;
    for (int s = 0; s < TR_S; s++) {
        float d = dz[s * MLP_NH + j];
        bs += d;
        if constexpr (blocked) {
#pragma unroll
            for (int i = 0; i < R; i += 4) {
                float4 x = *(const float4*)&in[s * K + wv * R + i];
                acc[i] += x.x * d;
                acc[i + 1] += x.y * d;
                acc[i + 2] += x.z * d;
                acc[i + 3] += x.w * d;
            }
        } else {
#pragma unroll
            for (int i = 0; i < R; i++)
                if (wv + TR_NW * i < K) acc[i] += in[s * K + wv + TR_NW * i] * d;
        }
    }
#pragma unroll
    for (int i = 0; i < R; i++) {
        int k = blocked ? wv * R + i : wv + TR_NW * i;
        if (k < K) out_w[k * MLP_NH + j] = acc[i];
    }
    if (wv == 0) out_b[j] = bs;
}

__device__ __forceinline__ long tr_row(long a, long b, long n, int i) {
    return (long)(((unsigned long long)a * (unsigned long long)i + (unsigned long long)b) % (unsigned long long)n);
}

// forward and backward of samples rows(i) = (a i + b) mod n of X / Y, i in
// [blockIdx.x TR_S, ...) below B; dOut = 2 (out - y) / B
__global__ __launch_bounds__(WAVE * TR_NW) void k_mlp_fwdbwd(int B, long n, long a, long b,
                                                             const float* __restrict__ X, const float* __restrict__ Y,
                                                             const float* __restrict__ W, float* __restrict__ part) {
    extern __shared__ __align__(16) float sm[];
    constexpr int NT = WAVE * TR_NW;
    const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int s0 = blockIdx.x * TR_S, ns = min(TR_S, B - s0);
    // weights (hidden matrices to the padded stride)
    for (int i = tid; i < MLP_NF * MLP_NH + MLP_NH; i += NT) sm[TL_W1 + i] = W[MLP_OFF_W1 + i];
    for (int i = tid; i < MLP_NH * MLP_NH; i += NT) {
        int k = i >> 6, j = i & 63;
        sm[TL_W2 + k * TR_WS + j] = W[MLP_OFF_W2 + i];
        sm[TL_W3 + k * TR_WS + j] = W[MLP_OFF_W3 + i];
    }
    if (tid < MLP_NH) {
        sm[TL_B2 + tid] = W[MLP_OFF_B2 + tid];
        sm[TL_B3 + tid] = W[MLP_OFF_B3 + tid];
        sm[TL_W4 + tid] = W[MLP_OFF_W4 + tid];
    }
    if (tid == 0) sm[TL_B4] = W[MLP_OFF_B4];
    // samples (rows past the batch are zero and carry no error)
    for (int i = tid; i < TR_S * MLP_NF; i += NT) {
        int s = i / MLP_NF, q = i - s * MLP_NF;
        sm[TL_X + i] = s < ns ? X[tr_row(a, b, n, s0 + s) * MLP_NF + q] : 0.0f;
    }
    __syncthreads();
    tr_dense<MLP_NF>(sm + TL_X, sm + TL_W1, MLP_NH, sm + TL_B1, sm + TL_Z1, sm + TL_H1);
    __syncthreads();
    tr_dense<MLP_NH>(sm + TL_H1, sm + TL_W2, TR_WS, sm + TL_B2, sm + TL_Z2, sm + TL_H2);
    __syncthreads();
    tr_dense<MLP_NH>(sm + TL_H2, sm + TL_W3, TR_WS, sm + TL_B3, sm + TL_Z3, sm + TL_H3);
    __syncthreads();
    // output layer and error: wave w takes rows w, w + 8, ..., its lanes
    // split the dot product over the 64 units
    float l = 0.0f;
    for (int s = wv; s < TR_S; s += TR_NW) {
        float o = sm[TL_H3 + s * MLP_NH + lane] * sm[TL_W4 + lane];
#pragma unroll
        for (int o2 = 32; o2 >= 1; o2 >>= 1) o += __shfl_xor(o, o2, WAVE);
        if (lane == 0) {
            float e = s < ns ? o + sm[TL_B4] - Y[tr_row(a, b, n, s0 + s)] : 0.0f;
            sm[TL_DO + s] = 2.0f * e / (float)B;
            l += e * e;
        }
    }
    if (lane == 0) sm[TL_LS + wv] = l;
    __syncthreads();
    float* prow = part + (size_t)blockIdx.x * TR_PART;
    if (tid == 0) {
        float ls = 0.0f;
#pragma unroll
        for (int k = 0; k < TR_NW; k++) ls += sm[TL_LS + k];
        prow[MLP_NPARAM] = ls;
    }
    {   // dZ3 = dOut w4 SiLU'(Z3), in place
        int j = lane;
        for (int s = wv; s < TR_S; s += TR_NW) {
            float& z = sm[TL_Z3 + s * MLP_NH + j];
            z = sm[TL_DO + s] * sm[TL_W4 + j] * silu_grad(z);
        }
    }
    // gradient of the output layer (wave 0): four independent sums per lane
    if (wv == 0) {
        float g[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gb = 0.0f;
        for (int s = 0; s < TR_S; s += 4) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float d = sm[TL_DO + s + q];
                g[q] += sm[TL_H3 + (s + q) * MLP_NH + lane] * d;
                gb += d;
            }
        }
        prow[MLP_OFF_W4 + lane] = (g[0] + g[1]) + (g[2] + g[3]);
        if (lane == 0) prow[MLP_OFF_B4] = gb;
    }
    __syncthreads();
    tr_grad<MLP_NH>(sm + TL_H2, sm + TL_Z3, prow + MLP_OFF_W3, prow + MLP_OFF_B3);
    tr_back(sm + TL_Z3, sm + TL_W3, sm + TL_Z2);   // Z2 <- dZ2
    __syncthreads();
    tr_grad<MLP_NH>(sm + TL_H1, sm + TL_Z2, prow + MLP_OFF_W2, prow + MLP_OFF_B2);
    tr_back(sm + TL_Z2, sm + TL_W2, sm + TL_Z1);   // Z1 <- dZ1
    __syncthreads();
    tr_grad<MLP_NF>(sm + TL_X, sm + TL_Z1, prow + MLP_OFF_W1, prow + MLP_OFF_B1);
}

// g[p] = sum of the blocks' rows; with lr > 0 also the Adam step; the squared
// errors, divided by B, are added to *loss.  One block of 16 waves per 64
// entries: lane = entry, wave w sums rows w, w + 16, ... (all its loads in
// flight), the waves combine in a fixed order (the sum is deterministic).
#define TR_RNW 16
__global__ __launch_bounds__(WAVE * TR_RNW) void k_mlp_reduce_adam(int nblk, int B, const float* __restrict__ part,
                                                                   float* __restrict__ g, float* __restrict__ w,
                                                                   float* __restrict__ m, float* __restrict__ v,
                                                                   float lr, float b1, float b2, float eps, float c1,
                                                                   float c2, double* __restrict__ loss) {
    __shared__ double red[TR_RNW][64];
    int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    int p = blockIdx.x * 64 + lane;
    bool ok = p <= MLP_NPARAM;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    int k = wv;
    for (; k + 3 * TR_RNW < nblk; k += 4 * TR_RNW) {
        float x[4];
#pragma unroll
        for (int q = 0; q < 4; q++) x[q] = ok ? part[(size_t)(k + TR_RNW * q) * TR_PART + p] : 0.0f;
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += (double)x[q];
    }
    for (; k < nblk; k += TR_RNW) acc[0] += ok ? (double)part[(size_t)k * TR_PART + p] : 0.0;
    red[wv][lane] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (wv != 0 || !ok) return;
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < TR_RNW; q++) sum += red[q][lane];
    if (p == MLP_NPARAM) {
        if (loss) *loss += sum / (double)B;
        return;
    }
    float gi = (float)sum;
    g[p] = gi;
    if (lr > 0.0f) {
        float mi = b1 * m[p] + (1.0f - b1) * gi, vi = b2 * v[p] + (1.0f - b2) * gi * gi;
        m[p] = mi;
        v[p] = vi;
        w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// FP32 master weights -> packed BF16 MFMA image (operand order of mlp_tile)
__global__ void k_mlp_pack(const float* __restrict__ W, unsigned char* __restrict__ img) {
    int t = threadIdx.x;   // blockDim.x a multiple of 64: one thread per (operand, lane)
    int lane = t & 63, m = lane & 15, kg = lane >> 4;
    for (int q = t >> 6; q < 20; q += blockDim.x >> 6) {
        bf16x8 v;
        if (q < 4) {   // A1 block q (first-layer bias in K slot 12)
            for (int j = 0; j < 8; j++) {
                int k = 8 * kg + j, u = 16 * q + m;
                float w = k < MLP_NF ? W[MLP_OFF_W1 + k * MLP_NH + u] : (k == MLP_NF ? W[MLP_OFF_B1 + u] : 0.0f);
                v[j] = (__bf16)w;
            }
            ((bf16x8*)(img + MLP_A1_OFF))[q * 64 + lane] = v;
        } else {
            int qq = q - 4, layer = qq >> 3, blk = (qq >> 1) & 3, ks = qq & 1;
            int off = layer == 0 ? MLP_OFF_W2 : MLP_OFF_W3;
            for (int j = 0; j < 8; j++) v[j] = (__bf16)W[off + mlp_kperm(ks, kg, j) * MLP_NH + 16 * blk + m];
            ((bf16x8*)(img + (layer == 0 ? MLP_A2_OFF : MLP_A3_OFF)))[(blk * 2 + ks) * 64 + lane] = v;
        }
    }
    float* C = (float*)(img + MLP_C_OFF);
    for (int i = t; i < 3 * 64; i += blockDim.x) {
        int layer = i / 64, rest = i - 64 * layer;
        int g = rest / 16, blk = (rest / 4) & 3, r = rest & 3;
        int u = mlp_unit(g, blk, r);
        C[i] = layer == 0 ? W[MLP_OFF_B2 + u] : (layer == 1 ? W[MLP_OFF_B3 + u] : W[MLP_OFF_W4 + u]);
    }
    if (t == 0) { C[192] = W[MLP_OFF_B4]; C[193] = C[194] = C[195] = 0.0f; }
}

// batched inference: one wave per tile of 16 feature rows; bf16 != 0 uses the
// MFMA tile (the sweep's path), otherwise the FP32 per-sample forward
__global__ __launch_bounds__(WAVE) void k_mlp_infer(int m, const float* __restrict__ X, const unsigned char* __restrict__ img,
                                                    const float* __restrict__ W, int bf16, float* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    int lane = lane_id();
    int row = blockIdx.x * MLP_TILE + (lane & (MLP_TILE - 1));
    if (!bf16) {
        if (lane < MLP_TILE && row < m) out[row] = mlp_forward_f32(W, X + (long)row * MLP_NF);
        return;
    }
    MlpRegs R;
    float* cst = (float*)smem;
    mlp_setup(img, cst, lane, R);
    float f[MLP_NF];
    for (int q = 0; q < MLP_NF; q++) f[q] = row < m ? X[(long)row * MLP_NF + q] : 0.0f;
    __syncthreads();
    float o = mlp_tile(R, cst, mlp_slots(f, lane >> 4), lane);
    if (lane < MLP_TILE && row < m) out[row] = o;
}

// ------------------------------------------------------- pairs and clusters
// Squared minimum-image distance in FP64 from fixed-point fractions, with the
// CPU brute force's operation order and no contraction into FMA, so the
// neighbour test is bit-identical to it.
__device__ __forceinline__ double fx_r2_fp64(uint4 a, uint4 b, double bx, double by, double bz) {
#pragma clang fp contract(off)
    const double s = 2.3283064365386962890625e-10;   // 2^-32
    double dx = bx * ((double)(int32_t)(a.x - b.x) * s);
    double dy = by * ((double)(int32_t)(a.y - b.y) * s);
    double dz = bz * ((double)(int32_t)(a.z - b.z) * s);
    double r2 = dx * dx;
    r2 = r2 + dy * dy;
    r2 = r2 + dz * dz;
    return r2;
}

// neighbour cells of a cell along one axis without repeats (grids of 1 or 2
// cells wrap onto themselves)
__device__ __forceinline__ int axis_offsets(int nc, int* off) {
    if (nc == 1) { off[0] = 0; return 1; }
    if (nc == 2) { off[0] = 0; off[1] = 1; return 2; }
    off[0] = -1; off[1] = 0; off[2] = 1;
    return 3;
}

// Pairs (ID_i < ID_j) within rc over a cell list with cells >= rc wide: one
// thread per particle (sorted position), counting (write == nullptr) or
// writing at its offset.  Within a particle the pairs come in cell-list order.
__global__ void k_pair_list(int n, const uint4* __restrict__ P, const int* __restrict__ ID, const int* __restrict__ starts,
                            Grid g, double rc2, double bx, double by, double bz, int* __restrict__ count,
                            const int* __restrict__ offset, int* __restrict__ pi, int* __restrict__ pj) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    uint4 ua = P[a];
    int ida = ID[a];
    int cx = axis_cell(ua.x, g.sh[0], g.nc[0]), cy = axis_cell(ua.y, g.sh[1], g.nc[1]), cz = axis_cell(ua.z, g.sh[2], g.nc[2]);
    int ox[3], oy[3], oz[3];
    int nx = axis_offsets(g.nc[0], ox), ny = axis_offsets(g.nc[1], oy), nz = axis_offsets(g.nc[2], oz);
    int k = 0, base = offset ? offset[a] : 0;
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int c = ((cx + ox[ix] + g.nc[0]) % g.nc[0]) +
                        g.nc[0] * (((cy + oy[iy] + g.nc[1]) % g.nc[1]) + g.nc[1] * ((cz + oz[iz] + g.nc[2]) % g.nc[2]));
                for (int b = starts[c]; b < starts[c + 1]; b++) {
                    int idb = ID[b];
                    if (idb <= ida) continue;
                    if (fx_r2_fp64(ua, P[b], bx, by, bz) <= rc2) {
                        if (offset) {
                            pi[base + k] = ida;
                            pj[base + k] = idb;
                        }
                        k++;
                    }
                }
            }
    if (!offset) count[a] = k;
}

// Connected components of the graph {pairs with energy <= thr}: hook the
// larger root under the smaller, then halve paths; repeated until no label
// changes, each particle ends labelled with the smallest index of its
// component (the fixed point ClusterSampler::genClusterList reaches,
// clust.cpp:105-163).
__global__ void k_cluster_hook(long m, const int* __restrict__ pi, const int* __restrict__ pj, const double* __restrict__ e,
                               double thr, int* __restrict__ label, int* __restrict__ changed) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= m || !(e[k] <= thr)) return;
    int a = label[pi[k]], b = label[pj[k]];
    if (a == b) return;
    int hi = a > b ? a : b, lo = a < b ? a : b;
    atomicMin(&label[hi], lo);
    *changed = 1;
}
// ClusterSampler::sameCluster (clust.cpp:86-90): particles of one molecule
// share a cluster; consecutive members are hooked like a bonded pair.
__global__ void k_cluster_hook_mol(int n, const int* __restrict__ cw, int* __restrict__ label,
                                   int* __restrict__ changed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !(cw[i] & CW_SAME_NEXT)) return;
    int a = label[i], b = label[i + 1];
    if (a == b) return;
    int hi = a > b ? a : b, lo = a < b ? a : b;
    atomicMin(&label[hi], lo);
    *changed = 1;
}
__global__ void k_cluster_jump(int n, int* __restrict__ label, int* __restrict__ changed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int l = label[i], ll = label[l];
    if (ll != l) {
        label[i] = ll;
        *changed = 1;
    }
}
__global__ void k_iota(int n, int* __restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

// Updater::alignmentOrder (updater.cpp:466-482): per pair within the list,
// dir_i . dir_j when the centre distance is < 1.5 (strict), else 0; FP64
// axes renormalised from the stored FP32 ones
__global__ void k_align_pairs(long m, const int* __restrict__ pi, const int* __restrict__ pj,
                              const int* __restrict__ slot, const uint4* __restrict__ P, const float4* __restrict__ O,
                              double bx, double by, double bz, double* __restrict__ out) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= m) return;
    int a = slot[pi[k]], b = slot[pj[k]];
    double v = 0.0;
    if (fx_r2_fp64(P[a], P[b], bx, by, bz) < 2.25) {
        float4 oa = O[a], ob = O[b];
        v = dot(unit64(oa.x, oa.y, oa.z), unit64(ob.x, ob.y, ob.z));
    }
    out[k] = v;
}

// Q = sum over particles of u u^T (6 unique components, FP64), per block
__global__ void k_nematic_q(int n, const float4* __restrict__ O, double* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double q[6] = {0, 0, 0, 0, 0, 0};
    if (i < n) {
        float4 o = O[i];
        double x = o.x, y = o.y, z = o.z, inv = 1.0 / sqrt(x * x + y * y + z * z);
        x *= inv; y *= inv; z *= inv;
        q[0] = x * x; q[1] = y * y; q[2] = z * z; q[3] = x * y; q[4] = x * z; q[5] = y * z;
    }
    for (int c = 0; c < 6; c++) {
        double v = wave_sum(q[c]);
        if (lane_id() == 0) atomicAdd(&out[c], v);
    }
}

// ------------------------------------------------------------- chain moves
// Whole-molecule moves (MoveCreator::chainMove, movecreator.cpp:305-328;
// chainDisplace :1075-1137; chainRotate + clusterRotate :1139-1305) on a
// coarse checkerboard of their own.  A molecule is anchored at its centre
// (clusterCM, :1307-1324: weighted by Ia_param::volume, taken here on the
// coordinates unwrapped by the minimum image from its first particle); its
// extent is the largest centre-to-particle distance.  Coarse cells are at
// least 2 rho + maxcut wide and coloured 2x2x2: two molecules moved in one
// pass have their particles within rho of anchors in different active
// cells, before and after every move, so they stay beyond maxcut of each
// other and each active cell runs an independent Metropolis chain.  A move
// is rejected when the molecule is wider than rho or its anchor would leave
// its coarse cell (the reverse move meets the same conditions).
#define CHAIN_MAXM 64     // particles per molecule
#define CHAIN_NW 8        // waves per block of the pass (each old/new item of a 3-particle molecule its own)
#define CHAIN_LCAP 4096   // molecules per coarse cell

struct ChainMoveParams {
    double temper, chain_mx, chain_angle, rho2;
    double frac;             // trials per molecule per visit of its cell
    uint32_t key0, key1;
    uint32_t pass_lo, pass_hi;
    uint32_t stamp;          // marks the particles moved in this pass
    int ox, oy, oz;          // colour of the pass
    int pa[3];               // colour period per axis (1 on an axis of one cell)
    int nact[3];             // active coarse cells per axis
    int nchain;
    uint32_t reach[3];       // rho + maxcut as box fractions (fixed point), the staged region's margin
};

__device__ __forceinline__ double u01_32(uint32_t x) { return (double)x * 2.3283064365386963e-10; }

__device__ __forceinline__ V3<double> unit_from_d(double a, double b) {
    double z = 1.0 - 2.0 * a;
    double s = sqrt(fmax(0.0, 1.0 - z * z));
    double ph = 6.283185307179586 * b;
    return mk<double>(s * cos(ph), s * sin(ph), z);
}

// Centre of a molecule relative to its particle 0 (real units, FP64) and the
// square of its extent.  Every lane of the wave passes its particle (lanes
// without one pass valid = false) and gets the same result.
struct ChainGeom {
    double c[3];
    double ext2;
};
__device__ ChainGeom chain_geom(uint4 ua, uint4 u0, bool valid, double v, const double* L) {
    const double k = 2.3283064365386963e-10;
    double dx = valid ? (double)(int32_t)(ua.x - u0.x) * (L[0] * k) : 0.0;
    double dy = valid ? (double)(int32_t)(ua.y - u0.y) * (L[1] * k) : 0.0;
    double dz = valid ? (double)(int32_t)(ua.z - u0.z) * (L[2] * k) : 0.0;
    double wv = valid ? v : 0.0;
    double vs = wave_sum(wv);
    ChainGeom G;
    G.c[0] = wave_sum(wv * dx) / vs;
    G.c[1] = wave_sum(wv * dy) / vs;
    G.c[2] = wave_sum(wv * dz) / vs;
    double ex = dx - G.c[0], ey = dy - G.c[1], ez = dz - G.c[2];
    double e2 = valid ? ex * ex + ey * ey + ez * ez : 0.0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) e2 = fmax(e2, __shfl_xor(e2, o, WAVE));
    G.ext2 = e2;
    return G;
}

// the centre as a fixed-point position
__device__ __forceinline__ uint4 chain_anchor_u(uint4 u0, const ChainGeom& G, const double* L) {
    return make_uint4(u0.x + (uint32_t)(long long)rint(G.c[0] * (4294967296.0 / L[0])),
                      u0.y + (uint32_t)(long long)rint(G.c[1] * (4294967296.0 / L[1])),
                      u0.z + (uint32_t)(long long)rint(G.c[2] * (4294967296.0 / L[2])), 0u);
}

__device__ __forceinline__ bool chain_cell_active(const Grid& gc, const ChainMoveParams& cp, int cc) {
    int x = cc % gc.nc[0], y = (cc / gc.nc[0]) % gc.nc[1], z = cc / (gc.nc[0] * gc.nc[1]);
    return x % cp.pa[0] == cp.ox && y % cp.pa[1] == cp.oy && z % cp.pa[2] == cp.oz;
}

// One wave per molecule: its coarse cell (counted per cell), and for each of
// its particles the cell while the molecule can move (extent <= rho), else -1.
__global__ __launch_bounds__(256) void k_chain_anchor(int nchain, const int* __restrict__ cfirst,
                                                      const int* __restrict__ csize, const int* __restrict__ slot,
                                                      const uint4* __restrict__ P, const double* __restrict__ vol,
                                                      SysConst sc, Grid gc, double rho2, int* __restrict__ ccell_chain,
                                                      int* __restrict__ ccell_id, int* __restrict__ ccount) {
    int c = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & (WAVE - 1));
    if (c >= nchain) return;
    int f = cfirst[c], m = csize[c];
    bool v = lane < m;
    uint4 u0 = P[slot[f]];
    uint4 ua = v ? P[slot[f + lane]] : u0;
    ChainGeom G = chain_geom(ua, u0, v, vol[ua.w], sc.Ld);
    uint4 an = chain_anchor_u(u0, G, sc.Ld);
    int cell = cell_of_u(gc, an.x, an.y, an.z);
    if (lane == 0) {
        ccell_chain[c] = cell;
        atomicAdd(&ccount[cell], 1);
    }
    if (v) ccell_id[f + lane] = G.ext2 <= rho2 ? cell : -1;
}

// the pair functor of a queued pair as a call of its own (its registers are
// not added to those of the surrounding scan)
__device__ __attribute__((noinline)) double chain_pair(const PP32* __restrict__ pp, const TP32* __restrict__ tp, int nt,
                                                       float sqmaxcut, uint4 ua, float4 oa, float2 qa, uint4 ub,
                                                       float4 ob, float2 qb, V3<float> r, bool sw) {
    DPart<float> pa = load_part<float>(oa, qa, ua.w, tp);
    DPart<float> pb = load_part<float>(ob, qb, ub.w, tp);
    // the pair in its canonical order, the higher ID as part1 (allToAll's)
    if (sw) return pair_energy_d<float>(pp[ub.w * nt + ua.w], pb, pa, mk<float>(-r.x, -r.y, -r.z), sqmaxcut);
    return pair_energy_d<float>(pp[ua.w * nt + ub.w], pa, pb, r, sqmaxcut);
}

// Energy of one particle of molecule [f, f + m) at the state (ua, pa) with
// everything outside its molecule (the wave's lanes hold partial sums): the
// stencil of its fine cell, without the particles of molecules that may move
// in another active coarse cell (they lie beyond maxcut) and without those
// this block moved earlier in the pass (their cell list entries are stale:
// they are taken from mlist instead).  Each lane takes four candidates per
// step: positions and IDs are loaded together, the molecule checks only for
// candidates inside sqmaxcut.  Candidates that pass are queued and evaluated
// 64 at a time.
#define CHAIN_QCAP (WAVE + 4 * WAVE)
__device__ double chain_item_energy(uint4 ua, float4 oa, float2 qa, int idk, int f, int m, int cell, const Grid& g,
                                    const Grid& gc, const SysConst& sc, const ChainMoveParams& cp,
                                    const int* __restrict__ starts, const uint4* P, const float4* O, const float2* Q,
                                    const int* __restrict__ ID, const int* __restrict__ slot,
                                    const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                                    const int* __restrict__ cfirst, const int* __restrict__ ccell_id,
                                    const uint32_t* stamp, const int* mlist, const int* mpre, int nm, int* pre,
                                    int* cbeg, int* queue) {
    int lane = lane_id();
    int c = cell_of_u(g, ua.x, ua.y, ua.z);
    int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
    int ns = stencil_size(g);
    int carry = 0;
    for (int b0 = 0; b0 < ns; b0 += WAVE) {
        int i = b0 + lane, cnt = 0, hm = 0;
        if (i < ns) {
            int cc = stencil_cell(g, cx, cy, cz, i, hm);
            int st = starts[cc];
            cnt = starts[cc + 1] - st;
            cbeg[i] = st;
        }
        int v = cnt;
        for (int o = 1; o < WAVE; o <<= 1) {
            int t = __shfl_up(v, o, WAVE);
            if (lane >= o) v += t;
        }
        if (i < ns) pre[i] = carry + v - cnt;
        carry += __shfl(v, WAVE - 1, WAVE);
    }
    wave_lds_sync();
    int total = carry, nmb = nm > 0 ? mpre[nm] : 0, all = total + nmb;
    double acc = 0.0;
    int qn = 0;
    for (int base = 0; base < all; base += 4 * WAVE) {
        int s[4], id[4];
        uint4 ub[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int idx = base + u * WAVE + lane;
            s[u] = -1;
            id[u] = -1;
            if (idx < total) {
                int lo = 0, hi = ns;
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (pre[mid] <= idx) lo = mid; else hi = mid;
                }
                s[u] = cbeg[lo] + idx - pre[lo];
                id[u] = ID[s[u]];
            } else if (idx < all) {
                int j = idx - total, lo = 0, hi = nm;
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (mpre[mid] <= j) lo = mid; else hi = mid;
                }
                id[u] = cfirst[mlist[lo]] + j - mpre[lo];
                s[u] = slot[id[u]];
            }
            if (s[u] >= 0) ub[u] = P[s[u]];
        }
        bool need[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int idx = base + u * WAVE + lane;
            need[u] = false;
            if (s[u] >= 0 && !(id[u] >= f && id[u] < f + m)) {
                V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ub[u].x, ub[u].y, ub[u].z, sc.Lh, sc.Ll);
                need[u] = dot(r, r) <= sc.sqmaxcut;
                if (need[u] && idx < total) {   // a cell-list entry: is it one that moves elsewhere, or moved here?
                    int cc = ccell_id[id[u]];
                    if (cc >= 0 && cc != cell && chain_cell_active(gc, cp, cc)) need[u] = false;
                    else if (stamp[id[u]] == cp.stamp) need[u] = false;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            unsigned long long msk = __ballot(need[u]);
            if (need[u]) queue[qn + wave_prefix(msk)] = s[u];
            qn += __popcll(msk);
        }
        wave_lds_sync();
        bool last = base + 4 * WAVE >= all;
        if (qn >= WAVE || (last && qn > 0)) {
            int take = last ? qn : (qn / WAVE) * WAVE;
            for (int e0 = 0; e0 < take; e0 += WAVE) {
                int e = e0 + lane;
                if (e < take) {
                    int sb = queue[e];
                    uint4 ubb = P[sb];
                    V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ubb.x, ubb.y, ubb.z, sc.Lh, sc.Ll);
                    acc += chain_pair(pp, tp, sc.nt, sc.sqmaxcut, ua, oa, qa, ubb, O[sb], Q[sb], r, ID[sb] > idk);
                }
            }
            wave_lds_sync();
            int rest = qn - take;
            int cq = lane < rest ? queue[take + lane] : 0;
            wave_lds_sync();
            if (lane < rest) queue[lane] = cq;
            wave_lds_sync();
            qn = rest;
        }
    }
    return acc;
}

// Wave 0 of a chain-move block: the trial state (Un, On, Qn) of the molecule's
// m particles (Uo, Oo, Qo) under the move drawn from ra (and the stream ctr),
// and the reject code (0: none, 1: wider than rho, 2: anchor leaves the
// coarse cell); every lane returns the same code.
__device__ int chain_trial_state(bool disp, U4 ra, U4 ctr, const ChainMoveParams& cp, const SysConst& sc,
                                 const Grid& gc, int cell, int m, const uint4* Uo, const float4* Oo,
                                 const float2* Qo, uint4* Un, float4* On, float2* Qn,
                                 const double* __restrict__ vol) {
    int lane = lane_id();
    bool v = lane < m;
    int a = v ? lane : 0;
    uint4 u0 = Uo[0], ua = Uo[a];
    float4 oa = Oo[a];
    float2 qa = Qo[a];
    double va = vol[ua.w];
    ChainGeom G = chain_geom(ua, u0, v, va, sc.Ld);
    uint4 un = ua;
    float4 on = oa;
    float2 qn = qa;
    if (disp) {
        // |dr| = chain_mx exactly, uniform direction (chainDisplace, :1088-1091)
        V3<double> d = unit_from_d(u01_32(ra.z), u01_32(ra.w));
        un.x = ua.x + (uint32_t)(long long)rint(d.x * cp.chain_mx * (4294967296.0 / sc.Ld[0]));
        un.y = ua.y + (uint32_t)(long long)rint(d.y * cp.chain_mx * (4294967296.0 / sc.Ld[1]));
        un.z = ua.z + (uint32_t)(long long)rint(d.z * cp.chain_mx * (4294967296.0 / sc.Ld[2]));
    } else {
        // clusterRotate (:1258-1305): quaternion (cos a, sin a * axis),
        // a = chain_angle * U, random sign; positions about the centre
        // in real units, axes and patch directions with them
        U4 c1 = ctr;
        c1.w |= 1u;
        U4 rb = philox(c1, cp.key0, cp.key1);
        V3<double> axv = unit_from_d(u01_32(rb.x), u01_32(rb.y));
        double vc = cos(cp.chain_angle * u01d(rb.z, rb.w));
        double vs = sqrt(fmax(0.0, 1.0 - vc * vc));
        if (ra.y & 1u) vs = -vs;
        const double k32 = 2.3283064365386963e-10;
        V3<double> o = mk<double>((double)(int32_t)(ua.x - u0.x) * (sc.Ld[0] * k32) - G.c[0],
                                  (double)(int32_t)(ua.y - u0.y) * (sc.Ld[1] * k32) - G.c[1],
                                  (double)(int32_t)(ua.z - u0.z) * (sc.Ld[2] * k32) - G.c[2]);
        V3<double> o2 = qrot(o, axv, vc, vs);
        un.x = u0.x + (uint32_t)(long long)rint((G.c[0] + o2.x) * (4294967296.0 / sc.Ld[0]));
        un.y = u0.y + (uint32_t)(long long)rint((G.c[1] + o2.y) * (4294967296.0 / sc.Ld[1]));
        un.z = u0.z + (uint32_t)(long long)rint((G.c[2] + o2.z) * (4294967296.0 / sc.Ld[2]));
        V3<double> d = normalised(qrot(mk<double>(oa.x, oa.y, oa.z), axv, vc, vs));
        V3<double> pq = qrot(mk<double>(oa.w, qa.x, qa.y), axv, vc, vs);
        pq = normalised(perp(pq, d));
        on = make_float4((float)d.x, (float)d.y, (float)d.z, (float)pq.x);
        qn = make_float2((float)pq.y, (float)pq.z);
    }
    if (v) {
        Un[lane] = un;
        On[lane] = on;
        Qn[lane] = qn;
    }
    uint4 un0 = make_uint4(__shfl(un.x, 0, WAVE), __shfl(un.y, 0, WAVE), __shfl(un.z, 0, WAVE), 0u);
    ChainGeom Gn = chain_geom(un, un0, v, va, sc.Ld);
    uint4 an = chain_anchor_u(un0, Gn, sc.Ld);
    if (G.ext2 > cp.rho2 || Gn.ext2 > cp.rho2) return 1;
    if (cell_of_u(gc, an.x, an.y, an.z) != cell) return 2;
    return 0;
}

// molecules of each coarse cell into index order (the scatter's atomics
// leave them in arrival order): one block per cell, a rank sort in LDS
__global__ __launch_bounds__(256) void k_chain_sort(int ncc, const int* __restrict__ cstart, int* __restrict__ clist) {
    __shared__ int v[CHAIN_LCAP];
    int c = blockIdx.x, tid = (int)threadIdx.x;
    if (c >= ncc) return;
    int b = cstart[c], n = cstart[c + 1] - b;
    if (n <= 1 || n > CHAIN_LCAP) return;   // too many: k_chain_pass reports it
    for (int i = tid; i < n; i += 256) v[i] = clist[b + i];
    __syncthreads();
    for (int i = tid; i < n; i += 256) {
        int key = v[i], rank = 0;
        for (int j = 0; j < n; j++) rank += v[j] < key ? 1 : 0;
        clist[b + rank] = key;   // keys are distinct molecule indices
    }
}

// LDS of k_chain_pass
__host__ __device__ constexpr size_t chain_pass_lds(int nw) {
    return (size_t)CHAIN_MAXM * (2 * 16 + 2 * 16 + 2 * 8) + 2 * nw * 8 +
           sizeof(int) * ((size_t)3 * CHAIN_LCAP + 1 + (size_t)nw * (256 + CHAIN_QCAP) + nw + 4);
}

// One block of NW waves per active coarse cell: trials on the cell's
// molecules, one after another.  Every thread draws the same numbers; wave 0
// builds the trial state; the energy items (particle, old/new state) are
// spread over the waves.
// The molecule's own energy at the state (U, O, Q) of its m particles (IDs
// f .. f+m-1): its non-bonded pairs and its bonds, pair (a, b < a) with a as
// part1 (allToAll's order), this thread's share (pairs tid, tid + stride, ...).
// A rigid move leaves it unchanged in exact arithmetic, but the stored state
// is rounded (fixed-point positions, FP32 axes); the chain trial adds its
// change to mol2others' so that the accepted changes add up to allToAll's.
__device__ __attribute__((noinline)) double chain_intra(int f, int m, const uint4* U, const float4* O, const float2* Q, int tid, int stride,
                              const SysConst& sc, const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                              const ChainTab& ct) {
    double e = 0.0;
    int np = m * (m - 1) / 2;
    for (int p = tid; p < np; p += stride) {
        int a = 1;
        while ((a + 1) * a / 2 <= p) a++;   // pair p = (a, b) with a (a - 1) / 2 <= p < (a + 1) a / 2
        int b = p - a * (a - 1) / 2;
        uint4 ua = U[a], ub = U[b];
        V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, sc.Lh, sc.Ll);
        if (dot(r, r) <= sc.sqmaxcut) e += chain_pair(pp, tp, sc.nt, sc.sqmaxcut, ua, O[a], Q[a], ub, O[b], Q[b], r, false);
        if (ct.cw && a - b <= 2) {
            int cw = ct.cw[f + a], bit = cw_bit(b - a);
            if (bit >= 0 && ((cw >> bit) & 1))
                e += bond_energy(ct.mol[cw >> 8], ct.pp64[ua.w * sc.nt + ub.w], b - a, fx_r12(ua, ub, sc.Ld),
                                 unit64(O[a].x, O[a].y, O[a].z), unit64(O[a].w, Q[a].x, Q[a].y),
                                 unit64(O[b].x, O[b].y, O[b].z), unit64(O[b].w, Q[b].x, Q[b].y));
        }
    }
    return e;
}

template <int NW>
__global__ __launch_bounds__(WAVE * NW) void k_chain_pass(Grid g, Grid gc, SysConst sc, ChainMoveParams cp,
                                                          const int* __restrict__ starts, uint4* P, float4* O,
                                                          float2* Q, const int* __restrict__ ID,
                                                          const int* __restrict__ slot, const PP32* __restrict__ pp,
                                                          const TP32* __restrict__ tp, ChainTab ct,
                                                          const double* __restrict__ vol,
                                                          const int* __restrict__ cfirst, const int* __restrict__ csize,
                                                          const int* __restrict__ cstart,
                                                          const int* __restrict__ clist_g,
                                                          const int* __restrict__ ccell_id, uint32_t* stamp,
                                                          unsigned long long* __restrict__ stats,
                                                          double* __restrict__ de_sum, int* __restrict__ flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint4* Uo = (uint4*)smem;
    uint4* Un = Uo + CHAIN_MAXM;
    float4* Oo = (float4*)(Un + CHAIN_MAXM);
    float4* On = Oo + CHAIN_MAXM;
    float2* Qo = (float2*)(On + CHAIN_MAXM);
    float2* Qn = Qo + CHAIN_MAXM;
    double* red = (double*)(Qn + CHAIN_MAXM);   // 2 NW
    int* clist = (int*)(red + 2 * NW);          // the cell's molecules
    int* mlist = clist + CHAIN_LCAP;            // molecules moved in this pass
    int* mpre = mlist + CHAIN_LCAP;             // particle prefix of mlist (CHAIN_LCAP + 1)
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1), w = tid >> 6;
    int* pre = mpre + CHAIN_LCAP + 1 + w * 256; // this wave's stencil prefix and cell starts
    int* cbeg = pre + 128;
    int* queue = mpre + CHAIN_LCAP + 1 + NW * 256 + w * CHAIN_QCAP;
    int* misc = mpre + CHAIN_LCAP + 1 + NW * (256 + CHAIN_QCAP);   // NW counts, reject code
    int b = blockIdx.x;
    int bx = b % cp.nact[0], by = (b / cp.nact[0]) % cp.nact[1], bz = b / (cp.nact[0] * cp.nact[1]);
    int cell = (bx * cp.pa[0] + cp.ox) + gc.nc[0] * ((by * cp.pa[1] + cp.oy) + gc.nc[1] * (bz * cp.pa[2] + cp.oz));
    // the cell's molecules in index order; the sign bit marks those moved in this pass
    int c0 = cstart[cell], total = cstart[cell + 1] - c0;
    if (total > CHAIN_LCAP) {
        if (tid == 0) atomicOr(flag, 4);
        return;
    }
    if (total == 0) return;
    for (int i = tid; i < total; i += WAVE * NW) clist[i] = clist_g[c0 + i];
    if (tid == 0) mpre[0] = 0;
    __syncthreads();
    U4 ctr;
    ctr.y = (uint32_t)cell;
    ctr.z = cp.pass_lo;
    // trials: the cell's share of the sweep's chain moves, randomly rounded
    ctr.x = 0xffffffffu;
    ctr.w = (cp.pass_hi << 2) | 3u;
    U4 rd = philox(ctr, cp.key0, cp.key1);
    double mean = (double)total * cp.frac;
    int ntrial = (int)floor(mean) + (u01d(rd.x, rd.y) < mean - floor(mean) ? 1 : 0);
    unsigned long long n_dacc = 0, n_dtry = 0, n_racc = 0, n_rtry = 0, n_rej = 0;
    double dsum = 0.0;
    int nm = 0;
    for (int t = 0; t < ntrial; t++) {
        ctr.x = (uint32_t)t;
        ctr.w = cp.pass_hi << 2;
        U4 ra = philox(ctr, cp.key0, cp.key1);
        int pick = (int)(((uint64_t)ra.x * (uint64_t)total) >> 32);
        int k = clist[pick] & 0x7fffffff;
        bool fresh = clist[pick] >= 0;
        bool disp = u01f(ra.y) < 0.5f;
        int f = cfirst[k], m = csize[k];
        if (disp) n_dtry++; else n_rtry++;
        if (tid < m) {
            int s = slot[f + tid];
            Uo[tid] = P[s];
            Oo[tid] = O[s];
            Qo[tid] = Q[s];
        }
        __syncthreads();
        if (w == 0) {
            int rej = chain_trial_state(disp, ra, ctr, cp, sc, gc, cell, m, Uo, Oo, Qo, Un, On, Qn, vol);
            if (lane == 0) misc[NW] = rej;
        }
        __syncthreads();
        if (misc[NW] != 0) {
            n_rej++;
            __syncthreads();
            continue;
        }
        ctr.w = (cp.pass_hi << 2) | 2u;
        U4 rc = philox(ctr, cp.key0, cp.key1);   // the acceptance uniform, drawn ahead of the energies
        double acc_o = 0.0, acc_n = 0.0;
        // items: (particle, state) for mol2others, then the molecule's own
        // pairs and bonds at each state (waves without a particle item take them)
        for (int it = w; it < 2 * m + (m > 1 ? 2 : 0); it += NW) {
            if (it >= 2 * m) {
                bool st = it > 2 * m;
                double e = chain_intra(f, m, st ? Un : Uo, st ? On : Oo, st ? Qn : Qo, lane, WAVE, sc, pp, tp, ct);
                if (st) acc_n += e; else acc_o += e;
                continue;
            }
            int a = it >> 1;
            bool st = (it & 1) != 0;
            uint4 ua = st ? Un[a] : Uo[a];
            double e = chain_item_energy(ua, st ? On[a] : Oo[a], st ? Qn[a] : Qo[a], f + a, f, m, cell, g, gc, sc, cp,
                                         starts,
                                         P, O, Q, ID, slot, pp, tp, cfirst, ccell_id, stamp, mlist, mpre, nm, pre, cbeg,
                                         queue);
            if (st) acc_n += e; else acc_o += e;
        }

        double eo, en;
        int rpar = 0;   // one paired sum per trial, after the previous trial's barrier
        block_sum2<NW>(acc_o, acc_n, red, rpar, eo, en);
        double de = en - eo;
        // moveTry (movecreator.h:175-187) on mol2others before and after
        bool acc = de <= 0.0 || exp(-de / cp.temper) > u01d(rc.x, rc.y);
        if (acc) {
            if (disp) n_dacc++; else n_racc++;
            dsum += de;
            if (tid < m) {
                int s = slot[f + tid];
                P[s] = Un[tid];
                O[s] = On[tid];
                Q[s] = Qn[tid];
                stamp[f + tid] = cp.stamp;
            }
            if (fresh) {
                if (tid == 0) {
                    mlist[nm] = k;
                    mpre[nm + 1] = mpre[nm] + m;
                    clist[pick] = k | (int)0x80000000;
                }
                nm++;
            }
        }
        __syncthreads();
    }
    int rep = (int)(blockIdx.x & (STAT_REP - 1));
    if (tid < 5) {
        unsigned long long v = tid == 0 ? n_dacc : tid == 1 ? n_dtry : tid == 2 ? n_racc : tid == 3 ? n_rtry : n_rej;
        if (v) atomicAdd(&stats[rep * STAT_STRIDE + 10 + tid], v);
    } else if (tid == 5) {
        atomicAdd(&de_sum[rep * STAT_STRIDE], dsum);
    }
}

// ---- staged chain pass
// Each block first copies into LDS every particle within rho + maxcut of its
// coarse cell (the region its molecules' particles and their neighbours can
// occupy during the pass), leaving out the particles of molecules that may
// move in another active cell (they stay beyond maxcut).  The trials then
// read nothing but LDS; an accepted move updates the staged copy and global
// memory.  The region is found through the fine cell list (whole cells
// overlapping it, filtered by position).  A block whose region holds more
// than `cap` particles makes no move in this pass (its trials are counted as
// rejected, and the block in chain_skipped).
#define CHAIN_ST_LCAP 512    // molecules per coarse cell
#define CHAIN_ST_OCAP 1024   // particles of the cell's movable molecules
#define CHAIN_ST_RCAP 1024   // fine cells overlapping a region
#define CHAIN_LDS_BUDGET 163840
#define CHAIN_OWN 0x40000000      // stage ID flag: a particle that may move in this pass
#define CHAIN_IDMASK 0x00ffffff

struct ChainStLayout {
    size_t su, so, sq, red, ints, sid, total;
};
__host__ __device__ inline ChainStLayout chain_st_layout(int nw, int cap) {
    ChainStLayout L;
    size_t o = 4 * CHAIN_MAXM * 16;   // Uo Un Oo On
    L.su = o; o += (size_t)cap * 16;
    L.so = o; o += (size_t)cap * 16;
    o += 2 * CHAIN_MAXM * 8;          // Qo Qn
    L.sq = o; o += (size_t)cap * 8;
    L.red = o; o += (size_t)2 * nw * 8;
    // bidx, clist, spre (stage start of each region cell), then the region
    // prefix (staging) sharing its space with the queues and stencil ranges (trials)
    L.ints = o;
    size_t shared = std::max((size_t)2 * CHAIN_ST_RCAP + 1, (size_t)nw * (CHAIN_QCAP + 64));
    o += sizeof(int) * ((size_t)CHAIN_MAXM + CHAIN_ST_LCAP + CHAIN_ST_OCAP + CHAIN_ST_RCAP + 1 + shared + nw + 4);
    L.sid = o; o += (size_t)cap * 4;
    L.total = o;
    return L;
}

// energy of one molecule particle at (ua, pa) with the staged particles
// outside its molecule (wave-level; lanes hold partial sums): the staged
// particles of the fine cells around it (every partner within maxcut lies
// in a region cell; stencil cells outside the region hold none), except the
// particles that may move in this pass, which may have left the cell they
// are staged under: those are taken from `own`, wherever they are now.
__device__ double chain_item_energy_st(uint4 ua, float4 oa, float2 qa, int idk, int f, int m, const Grid& g,
                                       const int* f0,
                                       const int* fn, const int* spre, const int* own, int nown, const SysConst& sc,
                                       const PP32* __restrict__ pp, const TP32* __restrict__ tp, const uint4* su,
                                       const float4* so, const float2* sq, const int* sid, int* queue, int* spr) {
    int lane = lane_id();
    int c = cell_of_u(g, ua.x, ua.y, ua.z);
    int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
    int ns = stencil_size(g);   // 27 at most: r = 1 grids only reach this kernel
    int beg = 0, cnt = 0;
    if (lane < ns) {
        int hm;
        int sc_ = stencil_cell(g, cx, cy, cz, lane, hm);
        int x = sc_ % g.nc[0], y = (sc_ / g.nc[0]) % g.nc[1], z = sc_ / (g.nc[0] * g.nc[1]);
        int rx = (x - f0[0] + g.nc[0]) % g.nc[0], ry = (y - f0[1] + g.nc[1]) % g.nc[1],
            rz = (z - f0[2] + g.nc[2]) % g.nc[2];
        if (rx < fn[0] && ry < fn[1] && rz < fn[2]) {
            int rc = rx + fn[0] * (ry + fn[1] * rz);
            beg = spre[rc];
            cnt = spre[rc + 1] - beg;
        }
    }
    int v = cnt;
    for (int o = 1; o < WAVE; o <<= 1) {
        int t = __shfl_up(v, o, WAVE);
        if (lane >= o) v += t;
    }
    int total = __shfl(v, WAVE - 1, WAVE), all = total + nown;
    if (lane < 32) {
        spr[lane] = v - cnt;
        spr[32 + lane] = beg;
    }
    wave_lds_sync();
    double acc = 0.0;
    int qn = 0;
    for (int base = 0; base < all; base += 4 * WAVE) {
        bool need[4];
        int si[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int idx = base + u * WAVE + lane;
            need[u] = false;
            si[u] = 0;
            bool ok = false;
            if (idx < total) {
                int lo = 0, hi = ns;
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (spr[mid] <= idx) lo = mid; else hi = mid;
                }
                si[u] = spr[32 + lo] + idx - spr[lo];
                ok = (sid[si[u]] & CHAIN_OWN) == 0;
            } else if (idx < all) {
                si[u] = own[idx - total];
                int id = sid[si[u]] & CHAIN_IDMASK;
                ok = !(id >= f && id < f + m);
            }
            if (ok) {
                uint4 ub = su[si[u]];
                V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, sc.Lh, sc.Ll);
                need[u] = dot(r, r) <= sc.sqmaxcut;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            unsigned long long msk = __ballot(need[u]);
            if (need[u]) queue[qn + wave_prefix(msk)] = si[u];
            qn += __popcll(msk);
        }
        wave_lds_sync();
        bool last = base + 4 * WAVE >= all;
        if (qn >= WAVE || (last && qn > 0)) {
            int take = last ? qn : (qn / WAVE) * WAVE;
            for (int e0 = 0; e0 < take; e0 += WAVE) {
                int e = e0 + lane;
                if (e < take) {
                    int sb = queue[e];
                    uint4 ub = su[sb];
                    V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, sc.Lh, sc.Ll);
                    acc += chain_pair(pp, tp, sc.nt, sc.sqmaxcut, ua, oa, qa, ub, so[sb], sq[sb], r,
                                      (sid[sb] & CHAIN_IDMASK) > idk);
                }
            }
            wave_lds_sync();
            int rest = qn - take;
            int cq = lane < rest ? queue[take + lane] : 0;
            wave_lds_sync();
            if (lane < rest) queue[lane] = cq;
            wave_lds_sync();
            qn = rest;
        }
    }
    return acc;
}

template <int NW>
__global__ __launch_bounds__(WAVE * NW) void k_chain_pass_st(Grid g, Grid gc, SysConst sc, ChainMoveParams cp,
                                                             int cap, const int* __restrict__ starts, uint4* P,
                                                             float4* O, float2* Q, const int* __restrict__ ID,
                                                             const int* __restrict__ slot,
                                                             const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                                                             ChainTab ct, const double* __restrict__ vol,
                                                             const int* __restrict__ cfirst,
                                                             const int* __restrict__ csize,
                                                             const int* __restrict__ cstart,
                                                             const int* __restrict__ clist_g,
                                                             const int* __restrict__ ccell_id,
                                                             unsigned long long* __restrict__ stats,
                                                             double* __restrict__ de_sum) {
    extern __shared__ __align__(16) unsigned char smem[];
    const ChainStLayout L = chain_st_layout(NW, cap);
    uint4* Uo = (uint4*)smem;
    uint4* Un = Uo + CHAIN_MAXM;
    float4* Oo = (float4*)(Un + CHAIN_MAXM);
    float4* On = Oo + CHAIN_MAXM;
    uint4* su = (uint4*)(smem + L.su);
    float4* so = (float4*)(smem + L.so);
    float2* Qo = (float2*)(smem + L.so + (size_t)cap * 16);
    float2* Qn = Qo + CHAIN_MAXM;
    float2* sq = (float2*)(smem + L.sq);
    double* red = (double*)(smem + L.red);
    int* bidx = (int*)(smem + L.ints);   // stage index of the trial molecule's particles
    int* clist = bidx + CHAIN_MAXM;
    int* own = clist + CHAIN_ST_LCAP;    // stage indices of the particles of the cell's movable molecules
    int* spre = own + CHAIN_ST_OCAP;     // stage index of each region cell's first particle (+ the end)
    int* rpre = spre + CHAIN_ST_RCAP + 1; // region cells: prefix of counts, cell starts (staging only)
    int* rbeg = rpre + CHAIN_ST_RCAP + 1;
    const int tid = (int)threadIdx.x, lane = tid & (WAVE - 1), w = tid >> 6;
    int* queue = rpre + w * (CHAIN_QCAP + 64);   // the same space, once the stage is built
    int* spr = queue + CHAIN_QCAP;               // this wave's stencil: 32 prefix + 32 starts
    int* misc = rpre + (2 * CHAIN_ST_RCAP + 1 > NW * (CHAIN_QCAP + 64) ? 2 * CHAIN_ST_RCAP + 1 : NW * (CHAIN_QCAP + 64));
    int* sid = (int*)(smem + L.sid);
    int b = blockIdx.x;
    int bx = b % cp.nact[0], by = (b / cp.nact[0]) % cp.nact[1], bz = b / (cp.nact[0] * cp.nact[1]);
    int gxyz[3] = {bx * cp.pa[0] + cp.ox, by * cp.pa[1] + cp.oy, bz * cp.pa[2] + cp.oz};
    int cell = gxyz[0] + gc.nc[0] * (gxyz[1] + gc.nc[1] * gxyz[2]);
    int c0 = cstart[cell], total = cstart[cell + 1] - c0;
    if (total == 0) return;
    // the region: the coarse cell widened by cp.reach on each side; the fine
    // cells that overlap it
    uint32_t rlo[3];
    uint64_t rwid[3];
    int f0[3], fn[3];
    for (int k = 0; k < 3; k++) {
        uint64_t nc = (uint64_t)gc.nc[k];
        uint32_t lo = gc.sh[k] + (uint32_t)(((uint64_t)gxyz[k] * 4294967296ull + nc - 1) / nc);
        uint32_t hi = gc.sh[k] + (uint32_t)(((uint64_t)(gxyz[k] + 1) * 4294967296ull + nc - 1) / nc);
        uint64_t wid = (uint64_t)(uint32_t)(hi - lo);
        if (gc.nc[k] == 1) wid = 4294967296ull;
        rwid[k] = wid + 2ull * cp.reach[k];
        rlo[k] = lo - cp.reach[k];
        uint64_t cw = 4294967296ull / (uint64_t)g.nc[k];
        if (rwid[k] + 2 * cw >= 4294967296ull) {
            rwid[k] = 4294967296ull;
            f0[k] = 0;
            fn[k] = g.nc[k];
        } else {
            f0[k] = axis_cell(rlo[k], g.sh[k], g.nc[k]);
            int f1 = axis_cell(rlo[k] + (uint32_t)(rwid[k] - 1), g.sh[k], g.nc[k]);
            fn[k] = (f1 - f0[k] + g.nc[k]) % g.nc[k] + 1;
        }
    }
    int nrc = fn[0] * fn[1] * fn[2];
    bool over = total > CHAIN_ST_LCAP || nrc > CHAIN_ST_RCAP;
    int nst = 0, nown_s = 0;
    if (!over) {
        for (int i = tid; i < total; i += WAVE * NW) clist[i] = clist_g[c0 + i];
        for (int rc = tid; rc < nrc; rc += WAVE * NW) {
            spre[rc] = -1;
            int ix = (f0[0] + rc % fn[0]) % g.nc[0], iy = (f0[1] + (rc / fn[0]) % fn[1]) % g.nc[1],
                iz = (f0[2] + rc / (fn[0] * fn[1])) % g.nc[2];
            int c = ix + g.nc[0] * (iy + g.nc[1] * iz);
            int st = starts[c];
            rbeg[rc] = st;
            rpre[rc] = starts[c + 1] - st;
        }
        __syncthreads();
        if (w == 0) {   // counts -> exclusive prefix
            int carry = 0;
            for (int b0 = 0; b0 < nrc; b0 += WAVE) {
                int i = b0 + lane, cnt = i < nrc ? rpre[i] : 0, v = cnt;
                for (int o = 1; o < WAVE; o <<= 1) {
                    int t = __shfl_up(v, o, WAVE);
                    if (lane >= o) v += t;
                }
                if (i < nrc) rpre[i] = carry + v - cnt;
                carry += __shfl(v, WAVE - 1, WAVE);
            }
            if (lane == 0) rpre[nrc] = carry;
        }
        __syncthreads();
        // four candidates per thread per step (their loads in flight
        // together), one block scan of the per-thread counts per step; the
        // stage keeps the candidates' order
        int T = rpre[nrc];
        constexpr int NT = WAVE * NW;
        for (int base = 0; base < T; base += 4 * NT) {
            int k[4], id[4], rcq[4];
            uint4 u[4];
            bool in[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int idx = base + 4 * tid + q;
                in[q] = false;
                k[q] = 0;
                rcq[q] = -1;
                if (idx < T) {
                    int lo = 0, hi = nrc;
                    while (hi - lo > 1) {
                        int mid = (lo + hi) >> 1;
                        if (rpre[mid] <= idx) lo = mid; else hi = mid;
                    }
                    if (rpre[lo] == idx) rcq[q] = lo;   // the first candidate of its region cell
                    k[q] = rbeg[lo] + idx - rpre[lo];
                    u[q] = P[k[q]];
                    id[q] = ID[k[q]];
                    in[q] = (uint64_t)(uint32_t)(u[q].x - rlo[0]) < rwid[0] &&
                            (uint64_t)(uint32_t)(u[q].y - rlo[1]) < rwid[1] &&
                            (uint64_t)(uint32_t)(u[q].z - rlo[2]) < rwid[2];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (in[q]) {
                    int cc = ccell_id[id[q]];
                    if (cc >= 0 && cc != cell && chain_cell_active(gc, cp, cc)) in[q] = false;
                    if (cc == cell) id[q] |= CHAIN_OWN;   // may move in this pass
                }
            }
            int cnt = (int)in[0] + (int)in[1] + (int)in[2] + (int)in[3];
            // exclusive block scan of cnt: lanes, then the waves in order
            int incl = cnt;
            for (int o = 1; o < WAVE; o <<= 1) {
                int t = __shfl_up(incl, o, WAVE);
                if (lane >= o) incl += t;
            }
            if (lane == WAVE - 1) misc[w] = incl;
            __syncthreads();
            int wbase = 0, all = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                int c = misc[q];
                wbase += q < w ? c : 0;
                all += c;
            }
            __syncthreads();
            int pos = nst + wbase + incl - cnt;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (rcq[q] >= 0) spre[rcq[q]] = pos;
                if (in[q]) {
                    if (pos < cap) {
                        su[pos] = u[q];
                        so[pos] = O[k[q]];
                        sq[pos] = Q[k[q]];
                        sid[pos] = id[q];
                    }
                    pos++;
                }
            }
            nst += all;
        }
        over = nst > cap;
        __syncthreads();
        if (w == 0) {   // region cells without candidates start where the next one does
            int carry = nst;
            for (int b0 = ((nrc - 1) / WAVE) * WAVE; b0 >= 0; b0 -= WAVE) {
                int i = b0 + lane;
                int v = (i < nrc && spre[i] >= 0) ? spre[i] : 0x7fffffff;
                for (int o = 1; o < WAVE; o <<= 1) {
                    int t = __shfl_down(v, o, WAVE);
                    if (lane + o < WAVE) v = min(v, t);
                }
                v = min(v, carry);
                if (i < nrc) spre[i] = v;
                carry = __shfl(v, 0, WAVE);
            }
            if (lane == 0) spre[nrc] = nst;
        }
        __syncthreads();
        // the particles that may move in this pass, in stage order
        int nown = 0;
        if (!over) {
            for (int b0 = 0; b0 < nst; b0 += WAVE * NW) {
                int i = b0 + tid;
                bool fl = i < nst && (sid[i] & CHAIN_OWN) != 0;
                int pos = block_compact<NW>(fl, nown, misc);
                if (fl && pos < CHAIN_ST_OCAP) own[pos] = i;
            }
        }
        over = over || nown > CHAIN_ST_OCAP;
        nown_s = nown;
    }
    U4 ctr;
    ctr.y = (uint32_t)cell;
    ctr.z = cp.pass_lo;
    ctr.x = 0xffffffffu;
    ctr.w = (cp.pass_hi << 2) | 3u;
    U4 rd = philox(ctr, cp.key0, cp.key1);
    double mean = (double)total * cp.frac;
    int ntrial = (int)floor(mean) + (u01d(rd.x, rd.y) < mean - floor(mean) ? 1 : 0);
    unsigned long long n_dacc = 0, n_dtry = 0, n_racc = 0, n_rtry = 0, n_rej = 0;
    double dsum = 0.0;
    __syncthreads();
    for (int t = 0; t < ntrial; t++) {
        ctr.x = (uint32_t)t;
        ctr.w = cp.pass_hi << 2;
        U4 ra = philox(ctr, cp.key0, cp.key1);
        bool disp = u01f(ra.y) < 0.5f;
        if (disp) n_dtry++; else n_rtry++;
        if (over) {
            n_rej++;
            continue;
        }
        int k = clist[(int)(((uint64_t)ra.x * (uint64_t)total) >> 32)];
        int f = cfirst[k], m = csize[k];
        if (tid < m) bidx[tid] = -1;
        __syncthreads();
        for (int i = tid; i < nown_s; i += WAVE * NW) {   // a movable molecule's particles are all in `own`
            int id = sid[own[i]] & CHAIN_IDMASK;
            if (id >= f && id < f + m) bidx[id - f] = own[i];
        }
        __syncthreads();
        if (w == 0) {
            // a particle outside the region means the molecule is wider than rho
            bool miss = __ballot(lane < m && bidx[lane] < 0) != 0ull;
            int rej = 1;
            if (!miss) {
                if (lane < m) {
                    int i = bidx[lane];
                    Uo[lane] = su[i];
                    Oo[lane] = so[i];
                    Qo[lane] = sq[i];
                }
                wave_lds_sync();
                rej = chain_trial_state(disp, ra, ctr, cp, sc, gc, cell, m, Uo, Oo, Qo, Un, On, Qn, vol);
            }
            if (lane == 0) misc[NW] = rej;
        }
        __syncthreads();
        if (misc[NW] != 0) {
            n_rej++;
            __syncthreads();
            continue;
        }
        ctr.w = (cp.pass_hi << 2) | 2u;
        U4 rc = philox(ctr, cp.key0, cp.key1);   // the acceptance uniform, drawn ahead of the energies
        double acc_o = 0.0, acc_n = 0.0;
        // items: (particle, state) for mol2others, then the molecule's own
        // pairs and bonds at each state (waves without a particle item take them)
        for (int it = w; it < 2 * m + (m > 1 ? 2 : 0); it += NW) {
            if (it >= 2 * m) {
                bool st = it > 2 * m;
                double e = chain_intra(f, m, st ? Un : Uo, st ? On : Oo, st ? Qn : Qo, lane, WAVE, sc, pp, tp, ct);
                if (st) acc_n += e; else acc_o += e;
                continue;
            }
            int a = it >> 1;
            bool st = (it & 1) != 0;
            uint4 ua = st ? Un[a] : Uo[a];
            double e = chain_item_energy_st(ua, st ? On[a] : Oo[a], st ? Qn[a] : Qo[a], f + a, f, m, g, f0, fn, spre,
                                            own,
                                            nown_s, sc, pp, tp, su, so, sq, sid, queue, spr);
            if (st) acc_n += e; else acc_o += e;
        }

        double eo, en;
        int rpar = 0;   // one paired sum per trial, after the previous trial's barrier
        block_sum2<NW>(acc_o, acc_n, red, rpar, eo, en);
        double de = en - eo;
        // moveTry (movecreator.h:175-187) on mol2others before and after
        bool acc = de <= 0.0 || exp(-de / cp.temper) > u01d(rc.x, rc.y);
        if (acc) {
            if (disp) n_dacc++; else n_racc++;
            dsum += de;
            if (tid < m) {
                int i = bidx[tid], s = slot[f + tid];
                su[i] = Un[tid];
                so[i] = On[tid];
                sq[i] = Qn[tid];
                P[s] = Un[tid];
                O[s] = On[tid];
                Q[s] = Qn[tid];
            }
        }
        __syncthreads();
    }
    int rep = (int)(blockIdx.x & (STAT_REP - 1));
    if (tid < 5) {
        unsigned long long v = tid == 0 ? n_dacc : tid == 1 ? n_dtry : tid == 2 ? n_racc : tid == 3 ? n_rtry : n_rej;
        if (v) atomicAdd(&stats[rep * STAT_STRIDE + 10 + tid], v);
    } else if (tid == 5) {
        atomicAdd(&de_sum[rep * STAT_STRIDE], dsum);
    } else if (tid == 6 && over) {
        atomicAdd(&stats[rep * STAT_STRIDE + 15], 1ull);
    }
}

// plain streaming copy: the calibration kernel for HBM byte counters
__global__ void k_copy(long n16, const uint4* __restrict__ a, uint4* __restrict__ b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n16; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// ------------------------------------------------------------------ context
// A ring of pinned host buffers for the per-sweep parameter uploads of
// replica batches: a copy from pinned memory is truly asynchronous, so the
// host prepares the next sweep while the device runs this one (a pageable
// copy waits for the stream and leaves the device idle while the host
// launches); each slot's event keeps the host from rewriting a buffer whose
// copy has not run yet.
#define PIN_SLOTS 4
struct PinRing {
    unsigned char* host[PIN_SLOTS] = {};
    size_t cap[PIN_SLOTS] = {};
    hipEvent_t ev[PIN_SLOTS] = {};
    bool used[PIN_SLOTS] = {};
    int next = 0;
};

struct sc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    const char* last_call = "";   // the HIP call or kernel launch issued last
    schost::Topology topo;
    bool have_topo = false;
    schost::Options opts;
    bool have_opts = false;
    int ref2dense[schost::MAXT];
    std::vector<int> dense2ref;
    int nt = 0;
    PP32* d_pp32 = nullptr;
    PP64* d_pp64 = nullptr;
    TP32* d_tp32 = nullptr;
    TP64* d_tp64 = nullptr;
    double maxcut = 0, sqmaxcut = 0;
    double minvol = 1.0;
    // state
    int n = 0;
    double box[3] = {0, 0, 0};
    bool have_state = false;
    std::vector<int> ptypes;   // reference types, original order
    uint4* P[2] = {nullptr, nullptr};
    float4* O[2] = {nullptr, nullptr};
    float2* Q[2] = {nullptr, nullptr};
    int* ID[2] = {nullptr, nullptr};
    int cur = 0;
    int* slot = nullptr;
    int* cell_tmp = nullptr;
    int* src_of_pos = nullptr;
    // grid
    Grid g;
    int ncell = 0, ncell_alloc = 0;
    int* starts = nullptr;
    int* counts = nullptr;
    int* fill = nullptr;
    double* partial = nullptr;
    bool grid_valid = false;
    // scratch
    double* d_out = nullptr;
    int d_out_cap = 0;
    int* d_flag = nullptr;
    int* d_maxocc = nullptr;
    int* d_nonempty = nullptr;
    int trial_mode = 0;
    int ppmax = 1;
    unsigned long long* d_stats = nullptr;
    double* d_dsum = nullptr;
    Query* d_q = nullptr;
    int* d_i = nullptr;
    int* d_j = nullptr;
    int ij_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_kernel_ms = 0;
    long last_launches = 0;
    bool profile = false;               // time every sweep-pass launch with events
    std::vector<hipEvent_t> kev;        // event pairs for the launches of one call
    PinRing pin_rg, pin_rp;             // uploads of the batch grids and pass parameters
    // surrogate
    float* mlp_w = nullptr;             // FP32 master weights (MLP_NPARAM)
    float* mlp_g = nullptr;
    float* mlp_m = nullptr;
    float* mlp_v = nullptr;
    unsigned char* mlp_img = nullptr;   // packed BF16 image for MFMA
    bool mlp_has = false;
    int mlp_type = -1;                  // reference type the network was trained for
    MlpConst mlp_c{};
    long adam_t = 0;
    float* tr_x = nullptr;
    float* tr_y = nullptr;
    float* tr_e = nullptr;
    long tr_n = 0;
    float* bx = nullptr;                // a batch handed in by the caller: features, then targets
    int act_batch = 0;
    float* d_part = nullptr;            // per block of a training step: gradient + squared error
    int part_blocks = 0;
    double* d_loss = nullptr;
    int da_mode = DA_OFF;              // delayed-acceptance mode of sc_sweep
    // pair lists, clusters, order parameter
    int* d_pcount = nullptr;            // per sorted particle: pair count, then offsets (n + 1)
    int pcount_cap = 0;
    int* d_pi = nullptr;
    int* d_pj = nullptr;
    double* d_pe = nullptr;
    long pair_cap = 0;
    int* d_label = nullptr;
    int label_cap = 0;
    double* d_q6 = nullptr;
    int* d_changed = nullptr;
    // replica batch (nrep > 0): the loaded particles are the replicas back to back
    int nrep = 0;
    std::vector<int> rep_n, rep_off;
    std::vector<std::array<double, 3>> rep_box;
    std::vector<RepGrid> rep_grid;
    std::vector<int> rep_ncell;
    int rep_ncell_total = 0;
    RepGrid* d_rg = nullptr;
    RepPass* d_rp = nullptr;
    int rp_cap = 0;
    int* d_rep_nonempty = nullptr;
    unsigned long long* d_rep_stats = nullptr;
    double* d_rep_dsum = nullptr;
    double* d_rep_e = nullptr;
    int rep_alloc = 0;
    int tuned_r = 2;                    // stencil radius chosen by the autotuner
    // bonded chains ([Molecules] with more than one particle)
    bool bonds = false;                 // some particle has a chain partner
    bool molecules = false;             // some molecule has more than one particle
    int* d_cw = nullptr;                // chain word per ID
    MolP* d_mol = nullptr;              // per molecule type
    double* d_bond = nullptr;           // per-ID bond energies (n)
    int bond_cap = 0;
    // whole-system energy carried from move to move (the counterpart of
    // TotalEMatrix::allToAll, which sums the cached pair energies): set by a
    // full evaluation, advanced by the sweeps' accepted dE, dropped when the
    // state changes otherwise
    bool e_valid = false;
    double e_track = 0.0;
    int e_age = 0;                      // volume moves since the last full evaluation
    int sweep_nw = 0;                   // waves per active cell in the sweep kernels (1, 2, 4; 0 = by mode)
    int spec = 1;                       // single-system sweep: speculative trials, one per wave (sweep_nw 0)
    int net_minx = -1;                  // network screen: least (pair, state) entries (< 0: one phase-C round)
    int home_cache = HCAP;              // cell size limit of the sweep's derived-vector cache (0: off)
    int visit = 1;                      // colour passes of a sweep = colours / visit, visit x the trials each
    int chain_visit = 1;                // the same for the chain-move passes (the largest divisor of their
                                        // colour count not above it)
    int chain_visit_eff = 0;            // the factor the last sweep's chain passes used
    unsigned char* noalloc = nullptr;   // this device's address of `scratch` (sc_pair_energies_noalloc)
    int max_active = 0;                 // sweep grids: at most this many active cells per colour (0: no limit)
    int max_active_default = 0;         // two per CU of the device
    // whole-molecule moves of sc_sweep
    std::vector<int> chain_first, chain_size;   // molecules of more than one particle
    int nchain_dev = 0;                 // molecules uploaded below
    int* d_cfirst = nullptr;
    int* d_csize = nullptr;
    int* d_ccell_chain = nullptr;       // per molecule: coarse cell of its anchor
    int* d_ccount = nullptr;            // coarse cells: count, fill, CSR start, molecules by cell
    int* d_cfill = nullptr;
    int* d_cstart = nullptr;
    int* d_clist = nullptr;
    int ccell_alloc = 0;
    int* d_ccell_id = nullptr;          // per ID: that cell while the molecule can move, else -1
    uint32_t* d_stamp = nullptr;        // per ID: stamp of the pass that last moved it
    double* d_vol = nullptr;            // Ia_param::volume per dense type
    uint32_t chain_stamp = 0;
    double chainprob = 0.0, chain_mx = 0.0, chain_angle = 0.0, chain_rho = 0.0;
    bool chain_staged = false;          // the last chain passes ran with the LDS stage
    int chain_period = 0;               // and this colour period
};

static ChainTab chain_tab(const sc_ctx* ctx) {
    ChainTab t;
    t.cw = ctx->bonds ? ctx->d_cw : nullptr;
    t.mol = ctx->d_mol;
    t.pp64 = ctx->d_pp64;
    return t;
}

// Every HIP call of a library function goes through HIPCK, which records the
// call (for error messages) and turns a failure into SC_EHIP.
#define HIPCK(x)                                                                       \
    do {                                                                               \
        ctx->last_call = #x;                                                           \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);                 \
            return SC_EHIP;                                                            \
        }                                                                              \
    } while (0)
#define SC_LAUNCH(kernel, ...)                                                         \
    do {                                                                               \
        ctx->last_call = #kernel;                                                      \
        hipLaunchKernelGGL(kernel, __VA_ARGS__);                                       \
    } while (0)

// copy `bytes` from host memory `src` to device `dst` on the context's
// stream through the next slot of ring r (see PinRing)
static int pin_upload(sc_ctx* ctx, PinRing& r, void* dst, const void* src, size_t bytes) {
    int s = r.next;
    r.next = (s + 1) % PIN_SLOTS;
    if (r.used[s]) HIPCK(hipEventSynchronize(r.ev[s]));
    if (bytes > r.cap[s]) {
        if (r.host[s]) HIPCK(hipHostFree(r.host[s]));
        r.host[s] = nullptr;
        r.cap[s] = 0;
        HIPCK(hipHostMalloc((void**)&r.host[s], bytes));
        r.cap[s] = bytes;
    }
    if (!r.ev[s]) HIPCK(hipEventCreateWithFlags(&r.ev[s], hipEventDisableTiming));
    std::memcpy(r.host[s], src, bytes);
    HIPCK(hipMemcpyAsync(dst, r.host[s], bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipEventRecord(r.ev[s], ctx->stream));
    r.used[s] = true;
    return SC_OK;
}

static int fail(sc_ctx* ctx, int code, const std::string& m) {
    ctx->err = m;
    return code;
}

template <typename T>
static int dalloc(sc_ctx* ctx, T** p, size_t count) {
    if (*p) { hipFree(*p); *p = nullptr; }
    if (count == 0) count = 1;
    HIPCK(hipMalloc((void**)p, count * sizeof(T)));
    return SC_OK;
}

// --- parameter tables -------------------------------------------------------
template <typename R>
static void fill_pair(PairParamsT<R>& o, const schost::Ia& p, const schost::Ia& tr) {
    memset(&o, 0, sizeof(o));
    int g = p.geotype[0], h = p.geotype[1];
    auto psc = [](int x) { return x == 12 || x == 14 || x == 16 || x == 18; };
    auto cpsc = [](int x) { return x == 13 || x == 15 || x == 17 || x == 19; };
    // functor table of PairE::initIntFCE (paire.cpp:6-80), later rules win
    int kind = K_NONE;
    if ((cpsc(g) && psc(h)) || (psc(g) && cpsc(h))) kind = K_MIXED;
    if (cpsc(g) && cpsc(h)) kind = K_CPSC;
    if (psc(g) && psc(h)) kind = K_PSC;
    if (g == 10 && h == 10) kind = K_SCN;
    if (g == 11 && h == 11) kind = K_SCA;
    if (g == 30 || h == 30) kind = K_SPH_WCA;
    if (g == 31 && h == 31) kind = K_SPH_TAYLOR;
    if ((g == 11 && h == 31) || (g == 31 && h == 11)) kind = K_SCASPA;
    if ((psc(g) && (h == 31 || h == 30)) || ((g == 31 || g == 30) && psc(h))) kind = K_PSCSPA;
    if ((cpsc(g) && (h == 31 || h == 30)) || ((g == 31 || g == 30) && cpsc(h))) kind = K_CPSCSPA;
    o.kind = kind;
    o.g0 = g;
    o.g1 = h;
    o.exclude = p.exclude ? 1 : 0;
    o.fT = is_two(g);
    o.sT = is_two(h);
    o.isp1 = g < 30;
    const schost::Ia& s = o.isp1 ? p : tr;   // iaParam of MixSpSc (SC first)
    int sg = s.geotype[0];
    if (kind == K_PSCSPA) { o.sp_chiral = sg == 14 || sg == 18; o.sp_sec = sg == 16 || sg == 18; }
    if (kind == K_CPSCSPA) { o.sp_chiral = sg == 15 || sg == 19; o.sp_sec = sg == 17 || sg == 19; }
    o.hl0 = (R)p.half_len[0];
    o.hl1 = (R)p.half_len[1];
    o.eps = (R)p.epsilon; o.A = (R)p.A; o.B = (R)p.B; o.pdis = (R)p.pdis; o.pswitch = (R)p.pswitch;
    o.pswitchINV = (R)p.pswitchINV; o.rcut = (R)p.rcut; o.rcutSq = (R)p.rcutSq; o.rcutwca = (R)p.rcutwca;
    o.rcutwcaSq = (R)p.rcutwcaSq;
    for (int k = 0; k < 4; k++) { o.pcangl[k] = (R)p.pcangl[k]; o.pcanglsw[k] = (R)p.pcanglsw[k]; }
    o.parallel = (R)p.parallel;
    o.t_hl0 = (R)tr.half_len[0]; o.t_eps = (R)tr.epsilon; o.t_pdis = (R)tr.pdis; o.t_pswitch = (R)tr.pswitch;
    o.t_rcut = (R)tr.rcut; o.t_rcutSq = (R)tr.rcutSq; o.t_rcutwcaSq = (R)tr.rcutwcaSq;
    o.t_pcangl0 = (R)tr.pcangl[0]; o.t_pcanglsw0 = (R)tr.pcanglsw[0]; o.t_exclude = tr.exclude ? 1 : 0;
}

template <typename R>
static void fill_type(TypeParamsT<R>& o, const schost::Ia& d) {
    memset(&o, 0, sizeof(o));
    o.geotype = d.geotype[0];
    o.pcoshalfi0 = (R)d.pcoshalfi[0]; o.psinhalfi0 = (R)d.psinhalfi[0];
    o.pcoshalfi2 = (R)d.pcoshalfi[2]; o.psinhalfi2 = (R)d.psinhalfi[2];
    o.csec = (R)d.csecpatchrot[0]; o.ssec = (R)d.ssecpatchrot[0];
    o.chcos = (R)d.chiral_cos[0]; o.chsin = (R)d.chiral_sin[0];
}

static int upload_tables(sc_ctx* ctx) {
    // dense types: those referenced by the system, in increasing type number
    for (int i = 0; i < schost::MAXT; i++) ctx->ref2dense[i] = -1;
    ctx->dense2ref.clear();
    std::vector<int> used(schost::MAXT, 0);
    for (int t : ctx->ptypes) if (t >= 0 && t < schost::MAXT) used[t] = 1;
    for (int t = 0; t < schost::MAXT; t++)
        if (used[t]) {
            if (ctx->topo.ia[t][t].geotype[0] == 0) return fail(ctx, SC_EINVAL, "particle type " + std::to_string(t) + " is not defined in [Types]");
            ctx->ref2dense[t] = (int)ctx->dense2ref.size();
            ctx->dense2ref.push_back(t);
        }
    ctx->nt = (int)ctx->dense2ref.size();
    if (ctx->nt > MAX_TYPES) return fail(ctx, SC_ECAPACITY, "more than 16 particle types in one system");
    int nt = ctx->nt;
    std::vector<PP32> h32(nt * nt);
    std::vector<PP64> h64(nt * nt);
    std::vector<TP32> t32(nt);
    std::vector<TP64> t64(nt);
    ctx->minvol = 1e30;
    for (int a = 0; a < nt; a++) {
        const schost::Ia& d = ctx->topo.ia[ctx->dense2ref[a]][ctx->dense2ref[a]];
        fill_type(t32[a], d);
        fill_type(t64[a], d);
        double geomv = 3.141592653589793 * pow(d.sigma, 3) / 6.0 + 3.141592653589793 * d.sigma * d.sigma * d.len[0] / 4.0;
        ctx->minvol = std::min(ctx->minvol, geomv > 0 ? geomv : 0.5);
        for (int b = 0; b < nt; b++) {
            const schost::Ia& p = ctx->topo.ia[ctx->dense2ref[a]][ctx->dense2ref[b]];
            const schost::Ia& tr = ctx->topo.ia[ctx->dense2ref[b]][ctx->dense2ref[a]];
            fill_pair(h32[a * nt + b], p, tr);
            fill_pair(h64[a * nt + b], p, tr);
        }
    }
    ctx->ppmax = 1;
    for (const PP32& q : h32)
        if (q.kind >= 0 && q.kind <= K_SCA) ctx->ppmax = std::max(ctx->ppmax, npatch_pairs(q));
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_pp32, h32.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_pp64, h64.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_tp32, t32.size()))) return rc;
    if ((rc = dalloc(ctx, &ctx->d_tp64, t64.size()))) return rc;
    HIPCK(hipMemcpy(ctx->d_pp32, h32.data(), h32.size() * sizeof(PP32), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->d_pp64, h64.data(), h64.size() * sizeof(PP64), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->d_tp32, t32.data(), t32.size() * sizeof(TP32), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->d_tp64, t64.data(), t64.size() * sizeof(TP64), hipMemcpyHostToDevice));
    ctx->maxcut = ctx->topo.maxcut;
    ctx->sqmaxcut = ctx->topo.sqmaxcut;
    return SC_OK;
}

// Chain words and molecule constants from the topology: particles in
// [System] order, molecule after molecule (Inicializer::setParticlesParamss,
// inicializer.cpp:409-450); the connectivity of ParticleVector::getConlist
// (Conf.h:97-147).  A configuration of another layout cannot carry chains.
static int upload_chains(sc_ctx* ctx) {
    ctx->bonds = ctx->molecules = false;
    ctx->chain_first.clear();
    ctx->chain_size.clear();
    ctx->nchain_dev = 0;
    const schost::Topology& T = ctx->topo;
    bool multi = false;
    for (const auto& m : T.molecules) multi |= m.types.size() > 1;
    if (!multi) return SC_OK;
    std::vector<int> cw;
    std::vector<int> types;
    for (const auto& sy : T.system)
        for (size_t mi = 0; mi < T.molecules.size(); mi++) {
            const schost::Molecule& m = T.molecules[mi];
            if (m.name != sy.first) continue;
            int size = (int)m.types.size();
            bool nearb = size > 1 && (m.bond1c >= 0.0 || m.bonddc >= 0.0 || m.bondhc >= 0.0);
            bool second = size > 1 && m.bond2c >= 0.0;
            for (long c = 0; c < sy.second; c++)
                for (int pos = 0; pos < size; pos++) {
                    if (pos == 0 && size > 1) {   // ParticleVector::getChain, Conf.h:150-163
                        ctx->chain_first.push_back((int)cw.size());
                        ctx->chain_size.push_back(size);
                    }
                    int w = (int)mi << 8;
                    if (nearb && pos > 0) w |= 1;
                    if (nearb && pos + 1 < size) w |= 2;
                    if (second && pos > 1) w |= 4;
                    if (second && pos + 2 < size) w |= 8;
                    if (pos + 1 < size) w |= CW_SAME_NEXT;
                    cw.push_back(w);
                    types.push_back(m.types[pos]);
                }
        }
    if ((int)cw.size() != ctx->n || types != ctx->ptypes)
        return fail(ctx, SC_EINVAL, "the topology has chains: the configuration must follow its [System] layout");
    std::vector<MolP> mp(T.molecules.size());
    for (size_t mi = 0; mi < T.molecules.size(); mi++) {
        const schost::Molecule& m = T.molecules[mi];
        MolP& q = mp[mi];
        q.c0 = q.eq0 = q.mod0 = q.mod1 = q.sp = q.c1 = q.eq1 = 0.0;
        if (m.bond1c >= 0.0) { q.eq0 = m.bond1eq; q.c0 = m.bond1c; }
        if (m.bonddc >= 0.0) { q.eq0 = 0.0; q.c0 = m.bonddc; q.mod0 = m.bonddeq; }
        if (m.bondhc >= 0.0) { q.eq0 = 0.0; q.c0 = m.bondhc; q.mod0 = q.mod1 = q.sp = m.bondheq; }
        if (m.bond2c >= 0.0) { q.eq1 = m.bond2eq; q.c1 = m.bond2c; }
        q.a1c = m.angle1c; q.a1eq = m.angle1eq;
        q.a2c = m.angle2c; q.a2eq = m.angle2eq;
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_cw, cw.size())) || (rc = dalloc(ctx, &ctx->d_mol, mp.size())) ||
        (rc = dalloc(ctx, &ctx->d_bond, cw.size())))
        return rc;
    HIPCK(hipMemcpy(ctx->d_cw, cw.data(), cw.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->d_mol, mp.data(), mp.size() * sizeof(MolP), hipMemcpyHostToDevice));
    ctx->molecules = true;
    for (int w : cw) ctx->bonds |= (w & 15) != 0;
    return SC_OK;
}

static SysConst sysconst(const sc_ctx* ctx, const double* box) {
    SysConst s;
    for (int k = 0; k < 3; k++) {
        s.L[k] = (float)box[k];
        s.Ld[k] = box[k];
        s.Lh[k] = (float)box[k];
        s.Ll[k] = (float)(box[k] - (double)s.Lh[k]);
    }
    s.sqmaxcut = (float)ctx->sqmaxcut;
    s.nt = ctx->nt;
    return s;
}

// --- cell list --------------------------------------------------------------
// Cell list with cells at least wmin wide; with period p > 1 the cell count
// per axis is a multiple of p (the checkerboard's colour period) or 1.
static int build_grid_w(sc_ctx* ctx, int r, const uint32_t shift[3], const double* box, double wmin, int p,
                        int cap_active = 0);

static int build_grid(sc_ctx* ctx, int r, const uint32_t shift[3], const double* box, int cap_active = 0) {
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (r < 1 || r > 2) return fail(ctx, SC_EINVAL, "stencil radius must be 1 or 2");
    return build_grid_w(ctx, r, shift, box, ctx->maxcut / r, r + 1, cap_active);
}

// cells per axis of a grid with cells >= wmin wide and colour period p
static int grid_axis_cells(double edge, double wmin, int p) {
    int nc = (int)floor(edge / wmin);
    if (nc > 1024) nc = 1024;
    if (nc >= p) nc = (nc / p) * p;   // a multiple of the colour period
    else nc = 1;                      // too small for a checkerboard: one cell
    if (nc < 1) nc = 1;
    return nc;
}

static int build_grid_w(sc_ctx* ctx, int r, const uint32_t shift[3], const double* box, double wmin, int p,
                        int cap_active) {
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    Grid g;
    g.r = r;
    for (int k = 0; k < 3; k++) {
        g.nc[k] = grid_axis_cells(box[k], wmin, p);
        g.sh[k] = shift ? shift[k] : 0u;
    }
    // with cap_active > 0, coarser cells (fewer along the axis with the most)
    // until a colour has at most cap_active cells: a pass whose active cells
    // all fit the GPU at once beats one that runs them in two rounds
    if (cap_active > 0)
        for (;;) {
            long act = 1;
            int kmax = -1;
            for (int k = 0; k < 3; k++) {
                act *= g.nc[k] >= p ? g.nc[k] / p : 1;
                if (g.nc[k] >= 2 * p && (kmax < 0 || g.nc[k] > g.nc[kmax])) kmax = k;
            }
            if (act <= cap_active || kmax < 0) break;
            g.nc[kmax] -= p;
        }
    int ncell = g.nc[0] * g.nc[1] * g.nc[2];
    if (ncell > ctx->ncell_alloc) {
        int rc;
        if ((rc = dalloc(ctx, &ctx->starts, ncell + 1))) return rc;
        if ((rc = dalloc(ctx, &ctx->counts, ncell))) return rc;
        if ((rc = dalloc(ctx, &ctx->fill, ncell))) return rc;
        if ((rc = dalloc(ctx, &ctx->partial, ncell))) return rc;
        ctx->ncell_alloc = ncell;
    }
    ctx->g = g;
    ctx->ncell = ncell;
    int n = ctx->n, c = ctx->cur, nb = (n + 255) / 256;
    HIPCK(hipMemsetAsync(ctx->counts, 0, sizeof(int) * ncell, ctx->stream));
    HIPCK(hipMemsetAsync(ctx->fill, 0, sizeof(int) * ncell, ctx->stream));
    if (n > 0) hipLaunchKernelGGL(k_cell_count, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->P[c], g, ctx->cell_tmp, ctx->counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, ncell, ctx->counts, ctx->starts, ctx->d_maxocc,
                       ctx->d_nonempty);
    if (n > 0) {
        hipLaunchKernelGGL(k_cell_scatter, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->cell_tmp, ctx->starts, ctx->fill, ctx->src_of_pos);
        hipLaunchKernelGGL(k_cell_sort, dim3((ncell + CELL_SORT_CELLS - 1) / CELL_SORT_CELLS), dim3(WAVE * CELL_SORT_CELLS), 0,
                           ctx->stream, ncell, ctx->starts, ctx->src_of_pos, ctx->ID[c]);
        hipLaunchKernelGGL(k_permute, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->src_of_pos, ctx->P[c], ctx->O[c], ctx->Q[c],
                           ctx->ID[c], ctx->P[c ^ 1], ctx->O[c ^ 1], ctx->Q[c ^ 1], ctx->ID[c ^ 1], ctx->slot);
    }
    HIPCK(hipGetLastError());
    ctx->cur ^= 1;
    ctx->grid_valid = true;
    return SC_OK;
}

// gfx950 gives a workgroup up to 160 KiB of LDS (as much as a CU has).  The
// sweep keeps a block's image under 80 KiB, so two blocks share a CU (the
// register budget allows two waves per SIMD); a stencil holds at most
// STAGE_CAP_MAX records (~2.5x the mean of the densest systems built here).
// The one-wave whole-system energy kernel keeps to ENERGY_CAP_MAX records
// (~52 KiB), so three of its blocks share a CU.
#define SWEEP_LDS_BUDGET 81920
#define STAGE_CAP_MAX 1408
#define ENERGY_CAP_MAX 1152

// stage capacity for a stencil: ~3x the mean occupancy (local packing cannot
// exceed ~2.6x the mean at the fluid densities used), bounded by the system
static int stage_cap_n(int n, double minvol, const Grid& g, const double* box, int capmax = STAGE_CAP_MAX) {
    double cellvol = (box[0] / g.nc[0]) * (box[1] / g.nc[1]) * (box[2] / g.nc[2]);
    int ns = std::min(2 * g.r + 1, g.nc[0]) * std::min(2 * g.r + 1, g.nc[1]) * std::min(2 * g.r + 1, g.nc[2]);
    double mean = (double)n / (double)(g.nc[0] * g.nc[1] * g.nc[2]) * ns;
    double pack = 0.74 * cellvol * ns / minvol;   // close packing of the stencil volume
    int cap = (int)std::min(std::max(3.0 * mean + 64.0, 128.0), pack + 64.0);
    cap = std::max(cap, 64);
    cap = std::min(cap, n + 64);
    cap = std::min(cap, capmax);
    return ((cap + 63) / 64) * 64;
}

static int stage_cap(const sc_ctx* ctx, const Grid& g, const double* box, int capmax = STAGE_CAP_MAX) {
    double cellvol = (box[0] / g.nc[0]) * (box[1] / g.nc[1]) * (box[2] / g.nc[2]);
    int ns = std::min(2 * g.r + 1, g.nc[0]) * std::min(2 * g.r + 1, g.nc[1]) * std::min(2 * g.r + 1, g.nc[2]);
    double mean = (double)ctx->n / (double)(g.nc[0] * g.nc[1] * g.nc[2]) * ns;
    double pack = 0.74 * cellvol * ns / ctx->minvol;   // close packing of the stencil volume
    int cap = (int)std::min(std::max(3.0 * mean + 64.0, 128.0), pack + 64.0);
    cap = std::max(cap, 64);
    cap = std::min(cap, ctx->n + 64);
    cap = std::min(cap, capmax);
    return ((cap + 63) / 64) * 64;
}

static size_t stage_bytes(int cap) { return (size_t)cap * (16 + 16 + 8 + 4); }

// LDS of k_energy_all: the stage, the stencil scratch, the block sums and one
// pair queue per wave
static size_t energy_lds_bytes(int cap) {
    return stage_bytes(cap) + sizeof(int) * (2 * 125 + 2) + sizeof(double) * ENERGY_NW +
           sizeof(int) * 2 * QCAP * ENERGY_NW;
}


// LDS for the sweep's copy of the parameter tables (with alignment slack)
static size_t sweep_table_bytes(int nt) {
    if (!SWEEP_TABLES_IN_LDS || nt > SWEEP_TABLE_MAXT) return 0;
    return (size_t)nt * nt * sizeof(PP32) + (size_t)nt * sizeof(TP32) + 16;
}

static int check_flag(sc_ctx* ctx, const char* what) {
    int f = 0;
    HIPCK(hipMemcpyAsync(&f, ctx->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (f) {
        HIPCK(hipMemset(ctx->d_flag, 0, sizeof(int)));
        if (f & 4) return fail(ctx, SC_ECAPACITY, std::string(what) + ": a coarse cell held more than 4096 molecules");
        if (f & 16)
            return fail(ctx, SC_ECAPACITY, std::string(what) + ": more than 320 particles within maxcut, or 192 "
                                                               "(pair, state) entries within the attraction's range, "
                                                               "of a moving particle (the speculative sweep's lists); "
                                                               "sc_set_speculation(ctx, 0) evaluates such systems");
        return fail(ctx, SC_ECAPACITY, std::string(what) + ": a stencil held more particles than the stage capacity");
    }
    return SC_OK;
}

static int ensure_out(sc_ctx* ctx, int m) {
    if (m > ctx->d_out_cap) {
        int rc;
        if ((rc = dalloc(ctx, &ctx->d_out, m))) return rc;
        ctx->d_out_cap = m;
    }
    return SC_OK;
}

static int ensure_ij(sc_ctx* ctx, int m) {
    if (m > ctx->ij_cap) {
        int rc;
        if ((rc = dalloc(ctx, &ctx->d_i, m))) return rc;
        if ((rc = dalloc(ctx, &ctx->d_j, m))) return rc;
        if ((rc = dalloc(ctx, &ctx->d_q, m))) return rc;
        ctx->ij_cap = m;
    }
    return SC_OK;
}

// Stencil radius of the whole-system energy's grid.  Radius 1 (cells >=
// maxcut, 27-cell stencil) stages each record for ~16 home particles of a
// cell; radius 2 (cells >= maxcut/2) culls more candidate pairs but a record
// staged in LDS serves only ~2 home particles.
#ifndef ENERGY_R
#define ENERGY_R 1
#endif

// The grid of the whole-system energy for box `box`.  Positions are box
// fractions: an unshifted grid of that radius with the cell counts the box
// asks for already holds the right cells.
static int energy_grid(sc_ctx* ctx, const double* box) {
    bool reuse = ctx->grid_valid && ctx->g.r == ENERGY_R;
    for (int k = 0; k < 3 && reuse; k++)
        reuse = ctx->g.sh[k] == 0u &&
                ctx->g.nc[k] == grid_axis_cells(box[k], ctx->maxcut / ENERGY_R, ENERGY_R + 1);
    return reuse ? SC_OK : build_grid(ctx, ENERGY_R, nullptr, box);
}

static int energy_all_impl(sc_ctx* ctx, const double* box, double* e) {
    if (ctx->nrep > 0) return fail(ctx, SC_ESTATE, "the context holds a replica batch: use sc_batch_energies");
    int rc = energy_grid(ctx, box);
    if (rc) return rc;
    const Grid& g = ctx->g;
    int cap = stage_cap(ctx, g, box, ENERGY_CAP_MAX);
    size_t lds = energy_lds_bytes(cap);
    if ((rc = ensure_out(ctx, 2))) return rc;
    SysConst sc = sysconst(ctx, box);
    int c = ctx->cur;
    HIPCK(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(k_energy_all<ENERGY_NW>, dim3(ctx->ncell), dim3(WAVE * ENERGY_NW), lds, ctx->stream, g, sc, ctx->starts, ctx->P[c],
                       ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32, cap, ctx->partial, ctx->d_flag);
    HIPCK(hipEventRecord(ctx->ev1, ctx->stream));
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, ctx->stream, ctx->ncell, ctx->partial, ctx->d_out);
    double eb[2] = {0.0, 0.0};
    if (ctx->bonds) {   // chain bonds: FP64 per particle, its partners of smaller ID
        hipLaunchKernelGGL(k_bond_all, dim3((ctx->n + 255) / 256), dim3(256), 0, ctx->stream, ctx->n, chain_tab(ctx),
                           sc, ctx->slot, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_bond);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, ctx->stream, ctx->n, ctx->d_bond, ctx->d_out + 1);
    }
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(eb, ctx->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = check_flag(ctx, "energy_all"))) return rc;
    *e = ctx->bonds ? eb[0] + eb[1] : eb[0];
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    ctx->last_kernel_ms = ms;
    ctx->last_launches = 1;
    return SC_OK;
}

// ------------------------------------------------------------------ the ABI
// No C++ exception may leave the C ABI (callers are C, ctypes, cgo...): each
// entry point runs its body under a handler that turns one into SC_EHIP with
// the exception's text and the HIP call that was in progress.
static int guard_fail(sc_ctx* ctx, const char* fn, const char* what) {
    if (ctx) ctx->err = std::string(fn) + ": C++ exception '" + what + "' during " + ctx->last_call;
    return SC_EHIP;
}
#define SC_GUARD(ctx, call)                                                            \
    try {                                                                              \
        return call;                                                                   \
    } catch (const std::exception& e_) {                                               \
        return guard_fail(ctx, __func__, e_.what());                                   \
    } catch (...) {                                                                    \
        return guard_fail(ctx, __func__, "unknown");                                   \
    }

extern "C" {

// Staging of sc_pair_energies_noalloc, the entry point that allocates
// nothing: the pair indices i and j (int32) and the energies (double) of up to
// SC_NOALLOC_MAX_PAIRS pairs, 1 MiB in all, in one file-scope device array
// that every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is a
// stand-in, not a device address): sc_create looks up each device's own
// address once, with that device current, and contexts keep it.
static __device__ __align__(256) unsigned char scratch[(size_t)SC_NOALLOC_MAX_PAIRS * 16];
#define SC_SCRATCH_DEVICES 64
static std::mutex scratch_mu;
static unsigned char* scratch_addr[SC_SCRATCH_DEVICES];

// the device address of `scratch` on the current device `device`
static hipError_t scratch_lookup(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (device < 0 || device >= SC_SCRATCH_DEVICES) return hipErrorInvalidDevice;
    if (!scratch_addr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_addr[device] = (unsigned char*)p;
    }
    *out = scratch_addr[device];
    return hipSuccess;
}

static void sc_destroy_impl(sc_ctx* ctx);
static int sc_create_impl(sc_ctx** out, int device) {
    if (!out) return SC_EINVAL;
    *out = nullptr;
    if (device == -1) {
        // host-only context: input formats and topology queries; every
        // device call on it returns SC_ENODEV
        sc_ctx* ctx = new sc_ctx();
        ctx->device = -1;
        *out = ctx;
        return SC_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return SC_ENODEV;
    sc_ctx* ctx = new sc_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return SC_ENODEV;
    }
    hipEventCreate(&ctx->ev0);
    hipEventCreate(&ctx->ev1);
    if (hipMalloc((void**)&ctx->d_flag, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_maxocc, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_nonempty, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_stats, STAT_REP * STAT_STRIDE * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_dsum, STAT_REP * STAT_STRIDE * sizeof(double)) != hipSuccess) {
        delete ctx;
        return SC_EHIP;
    }
    hipMemset(ctx->d_flag, 0, sizeof(int));
    // two blocks of the sweep per CU (their register budget): a colour of at
    // most two active cells per CU runs in one round
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->max_active = 2 * cus;
    ctx->max_active_default = ctx->max_active;
    if (scratch_lookup(device, &ctx->noalloc) != hipSuccess) {
        sc_destroy_impl(ctx);
        return SC_EHIP;
    }
    *out = ctx;
    return SC_OK;
}

int sc_create(sc_ctx** out, int device) {
    SC_GUARD(nullptr, sc_create_impl(out, device))
}

static void sc_destroy_impl(sc_ctx* ctx) {
    if (!ctx) return;
    if (ctx->device < 0) { delete ctx; return; }
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {ctx->d_pp32, ctx->d_pp64, ctx->d_tp32, ctx->d_tp64, ctx->P[0], ctx->P[1], ctx->O[0], ctx->O[1],
                    ctx->Q[0], ctx->Q[1], ctx->ID[0], ctx->ID[1], ctx->slot, ctx->cell_tmp, ctx->src_of_pos,
                    ctx->starts, ctx->counts, ctx->fill, ctx->partial, ctx->d_out, ctx->d_flag, ctx->d_maxocc,
                    ctx->d_nonempty, ctx->d_stats, ctx->d_dsum, ctx->d_q, ctx->d_i, ctx->d_j, ctx->mlp_w,
                    ctx->mlp_g, ctx->mlp_m, ctx->mlp_v, ctx->mlp_img, ctx->tr_x, ctx->tr_y, ctx->tr_e, ctx->d_part,
                    ctx->bx, ctx->d_loss, ctx->d_pcount, ctx->d_pi, ctx->d_pj, ctx->d_pe, ctx->d_label, ctx->d_q6,
                    ctx->d_changed, ctx->d_rg, ctx->d_rp, ctx->d_rep_nonempty, ctx->d_rep_stats, ctx->d_rep_dsum,
                    ctx->d_rep_e, ctx->d_cw, ctx->d_mol, ctx->d_bond, ctx->d_cfirst, ctx->d_csize,
                    ctx->d_ccell_chain, ctx->d_ccell_id, ctx->d_stamp, ctx->d_vol, ctx->d_ccount, ctx->d_cfill,
                    ctx->d_cstart, ctx->d_clist};
    for (void* p : ptrs)
        if (p) hipFree(p);
    for (hipEvent_t e : ctx->kev) hipEventDestroy(e);
    for (PinRing* r : {&ctx->pin_rg, &ctx->pin_rp})
        for (int k = 0; k < PIN_SLOTS; k++) {
            if (r->host[k]) hipHostFree(r->host[k]);
            if (r->ev[k]) hipEventDestroy(r->ev[k]);
        }
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

void sc_destroy(sc_ctx* ctx) {
    try {
        sc_destroy_impl(ctx);
    } catch (...) {
    }
}

const char* sc_last_error(const sc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static int sc_load_topology_impl(sc_ctx* ctx, const char* text) {
    if (!ctx || !text) return SC_EINVAL;
    std::string e = ctx->topo.parse(text);
    if (!e.empty()) { ctx->have_topo = false; return fail(ctx, SC_EINVAL, e); }
    ctx->have_topo = true;
    ctx->have_state = false;
    ctx->ptypes = ctx->topo.particleTypes();
    return SC_OK;
}

int sc_load_topology(sc_ctx* ctx, const char* text) {
    SC_GUARD(ctx, sc_load_topology_impl(ctx, text))
}

static int sc_load_options_impl(sc_ctx* ctx, const char* text) {
    if (!ctx || !text) return SC_EINVAL;
    std::string e = ctx->opts.parse(text);
    if (!e.empty()) return fail(ctx, SC_EINVAL, e);
    ctx->have_opts = true;
    return SC_OK;
}

int sc_load_options(sc_ctx* ctx, const char* text) {
    SC_GUARD(ctx, sc_load_options_impl(ctx, text))
}

int sc_get_option(const sc_ctx* ctx, const char* key, double* v) {
    if (!ctx || !key || !v) return SC_EINVAL;
    return ctx->opts.get(key, *v) ? SC_OK : SC_EINVAL;
}

int sc_num_particles(const sc_ctx* ctx) { return ctx ? (int)ctx->ptypes.size() : 0; }

int sc_particle_types(const sc_ctx* ctx, int32_t* types, int cap) {
    if (!ctx) return SC_EINVAL;
    int n = (int)ctx->ptypes.size();
    for (int i = 0; i < n && i < cap; i++) types[i] = ctx->ptypes[i];
    return n;
}

int sc_maxcut(const sc_ctx* ctx, double* m) {
    if (!ctx || !m || !ctx->have_topo) return SC_ESTATE;
    *m = ctx->topo.maxcut;
    return SC_OK;
}

int sc_pair_params(const sc_ctx* ctx, int ti, int tj, double* out) {
    if (!ctx || !out || !ctx->have_topo || ti < 0 || tj < 0 || ti >= schost::MAXT || tj >= schost::MAXT) return SC_EINVAL;
    schost::iaRow(ctx->topo.ia[ti][tj], out);
    return SC_OK;
}

static int sc_set_config_impl(sc_ctx* ctx, int n, const double box[3], const double* pos, const double* dir, const double* pd0,
                  const int32_t* types) {
    if (!ctx || n < 0 || !box || (n > 0 && (!pos || !dir || !pd0))) return SC_EINVAL;
    if (!ctx->have_topo) return fail(ctx, SC_ESTATE, "load a topology first");
    if (ctx->device < 0) return fail(ctx, SC_ENODEV, "host-only context: no device state");
    if (n >= (1 << 24)) return fail(ctx, SC_ECAPACITY, "too many particles");
    HIPCK(hipSetDevice(ctx->device));
    ctx->nrep = 0;   // a new configuration ends any replica batch
    if (types) ctx->ptypes.assign(types, types + n);
    if ((int)ctx->ptypes.size() != n) return fail(ctx, SC_EINVAL, "particle count differs from the topology's [System]");
    int rc;
    if ((rc = upload_tables(ctx))) return rc;
    for (int k = 0; k < 3; k++) {
        if (!(box[k] > 0)) return fail(ctx, SC_EINVAL, "box edges must be positive");
        ctx->box[k] = box[k];
    }
    if (n != ctx->n || !ctx->P[0]) {
        for (int b = 0; b < 2; b++) {
            if ((rc = dalloc(ctx, &ctx->P[b], n))) return rc;
            if ((rc = dalloc(ctx, &ctx->O[b], n))) return rc;
            if ((rc = dalloc(ctx, &ctx->Q[b], n))) return rc;
            if ((rc = dalloc(ctx, &ctx->ID[b], n))) return rc;
        }
        if ((rc = dalloc(ctx, &ctx->slot, n))) return rc;
        if ((rc = dalloc(ctx, &ctx->cell_tmp, n))) return rc;
        if ((rc = dalloc(ctx, &ctx->src_of_pos, n))) return rc;
    }
    ctx->n = n;
    std::vector<uint4> hp(n);
    std::vector<float4> ho(n);
    std::vector<float2> hq(n);
    std::vector<int> hid(n);
    for (int i = 0; i < n; i++) {
        uint32_t u[3];
        for (int k = 0; k < 3; k++) {
            double s = pos[3 * i + k];
            s -= floor(s);
            double v = nearbyint(s * 4294967296.0);
            u[k] = (uint32_t)(uint64_t)(v >= 4294967296.0 ? 0.0 : v);
        }
        // Inicializer::initConfig normalisation (inicializer.cpp:209-225)
        double d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]}, q[3] = {pd0[3 * i], pd0[3 * i + 1], pd0[3 * i + 2]};
        double t = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (t != 0.0) { t = 1.0 / t; d[0] *= t; d[1] *= t; d[2] *= t; }
        double dp = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
        q[0] -= dp * d[0]; q[1] -= dp * d[1]; q[2] -= dp * d[2];
        t = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        if (t != 0.0) { t = 1.0 / t; q[0] *= t; q[1] *= t; q[2] *= t; }
        int dt = ctx->ref2dense[ctx->ptypes[i]];
        hp[i] = make_uint4(u[0], u[1], u[2], (uint32_t)dt);
        ho[i] = make_float4((float)d[0], (float)d[1], (float)d[2], (float)q[0]);
        hq[i] = make_float2((float)q[1], (float)q[2]);
        hid[i] = i;
    }
    ctx->cur = 0;
    if (n > 0) {
        HIPCK(hipMemcpy(ctx->P[0], hp.data(), n * sizeof(uint4), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->O[0], ho.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->Q[0], hq.data(), n * sizeof(float2), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->ID[0], hid.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->slot, hid.data(), n * sizeof(int), hipMemcpyHostToDevice));
    }
    if ((rc = upload_chains(ctx))) return rc;
    ctx->e_valid = false;
    ctx->have_state = true;
    ctx->grid_valid = false;
    return build_grid(ctx, 2, nullptr, ctx->box);
}

int sc_set_config(sc_ctx* ctx, int n, const double box[3], const double* pos, const double* dir, const double* pd0,
                  const int32_t* types) {
    SC_GUARD(ctx, sc_set_config_impl(ctx, n, box, pos, dir, pd0, types))
}

static int sc_load_config_impl(sc_ctx* ctx, const char* text) {
    if (!ctx || !text) return SC_EINVAL;
    if (!ctx->have_topo) return fail(ctx, SC_ESTATE, "load a topology first");
    int n = (int)ctx->ptypes.size();
    double box[3];
    std::vector<double> pos, dir, pd0;
    std::string e = schost::parseConfig(text, n, box, pos, dir, pd0);
    if (!e.empty()) return fail(ctx, SC_EINVAL, e);
    std::vector<int32_t> ty(ctx->ptypes.begin(), ctx->ptypes.end());
    return sc_set_config(ctx, n, box, pos.data(), dir.data(), pd0.data(), ty.data());
}

int sc_load_config(sc_ctx* ctx, const char* text) {
    SC_GUARD(ctx, sc_load_config_impl(ctx, text))
}

static int download(sc_ctx* ctx, std::vector<uint4>& hp, std::vector<float4>& ho, std::vector<float2>& hq) {
    int n = ctx->n, c = ctx->cur;
    std::vector<int> sl(n);
    std::vector<uint4> sp(n);
    std::vector<float4> so(n);
    std::vector<float2> sq(n);
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (n > 0) {
        HIPCK(hipMemcpy(sl.data(), ctx->slot, n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(sp.data(), ctx->P[c], n * sizeof(uint4), hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(so.data(), ctx->O[c], n * sizeof(float4), hipMemcpyDeviceToHost));
        HIPCK(hipMemcpy(sq.data(), ctx->Q[c], n * sizeof(float2), hipMemcpyDeviceToHost));
    }
    hp.resize(n); ho.resize(n); hq.resize(n);
    for (int i = 0; i < n; i++) { hp[i] = sp[sl[i]]; ho[i] = so[sl[i]]; hq[i] = sq[sl[i]]; }
    return SC_OK;
}

static int sc_get_config_impl(sc_ctx* ctx, double box[3], double* pos, double* dir, double* pd0) {
    if (!ctx) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    std::vector<uint4> hp; std::vector<float4> ho; std::vector<float2> hq;
    int rc = download(ctx, hp, ho, hq);
    if (rc) return rc;
    if (box) for (int k = 0; k < 3; k++) box[k] = ctx->box[k];
    const double s = 1.0 / 4294967296.0;
    for (int i = 0; i < ctx->n; i++) {
        if (pos) { pos[3 * i] = hp[i].x * s; pos[3 * i + 1] = hp[i].y * s; pos[3 * i + 2] = hp[i].z * s; }
        if (dir) { dir[3 * i] = ho[i].x; dir[3 * i + 1] = ho[i].y; dir[3 * i + 2] = ho[i].z; }
        if (pd0) { pd0[3 * i] = ho[i].w; pd0[3 * i + 1] = hq[i].x; pd0[3 * i + 2] = hq[i].y; }
    }
    return SC_OK;
}

int sc_get_config(sc_ctx* ctx, double box[3], double* pos, double* dir, double* pd0) {
    SC_GUARD(ctx, sc_get_config_impl(ctx, box, pos, dir, pd0))
}

static int sc_get_fixed_positions_impl(sc_ctx* ctx, uint32_t* u) {
    if (!ctx || !u) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    std::vector<uint4> hp; std::vector<float4> ho; std::vector<float2> hq;
    int rc = download(ctx, hp, ho, hq);
    if (rc) return rc;
    for (int i = 0; i < ctx->n; i++) { u[3 * i] = hp[i].x; u[3 * i + 1] = hp[i].y; u[3 * i + 2] = hp[i].z; }
    return SC_OK;
}

int sc_get_fixed_positions(sc_ctx* ctx, uint32_t* u) {
    SC_GUARD(ctx, sc_get_fixed_positions_impl(ctx, u))
}

static long sc_write_config_impl(sc_ctx* ctx, char* buf, long cap) {
    if (!ctx) return SC_EINVAL;
    int n = ctx->n;
    std::vector<double> pos(3 * n), dir(3 * n), pd0(3 * n);
    double box[3];
    int rc = sc_get_config(ctx, box, pos.data(), dir.data(), pd0.data());
    if (rc) return rc;
    std::vector<int> mol = ctx->topo.particleMolecules();
    std::string s = schost::drawConfig(box, pos, dir, pd0, mol);
    if (buf && cap > (long)s.size()) memcpy(buf, s.c_str(), s.size() + 1);
    return (long)s.size();
}

long sc_write_config(sc_ctx* ctx, char* buf, long cap) {
    SC_GUARD(ctx, sc_write_config_impl(ctx, buf, cap))
}

static int sc_build_cells_impl(sc_ctx* ctx, int r, const uint32_t shift[3], int32_t* cell_of, int32_t* cell_start,
                   int32_t* sorted, int32_t ncell_out[3]) {
    if (!ctx) return SC_EINVAL;
    HIPCK(hipSetDevice(ctx->device));
    int rc = build_grid(ctx, r, shift, ctx->box);
    if (rc) return rc;
    int n = ctx->n, nc = ctx->ncell, c = ctx->cur;
    if (ncell_out) for (int k = 0; k < 3; k++) ncell_out[k] = ctx->g.nc[k];
    std::vector<int> st(nc + 1), ids(n);
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipMemcpy(st.data(), ctx->starts, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (n > 0) HIPCK(hipMemcpy(ids.data(), ctx->ID[c], n * sizeof(int), hipMemcpyDeviceToHost));
    if (cell_start) for (int i = 0; i <= nc; i++) cell_start[i] = st[i];
    if (sorted) for (int i = 0; i < n; i++) sorted[i] = ids[i];
    if (cell_of)
        for (int cc = 0; cc < nc; cc++)
            for (int k = st[cc]; k < st[cc + 1]; k++) cell_of[ids[k]] = cc;
    return SC_OK;
}

int sc_build_cells(sc_ctx* ctx, int r, const uint32_t shift[3], int32_t* cell_of, int32_t* cell_start,
                   int32_t* sorted, int32_t ncell_out[3]) {
    SC_GUARD(ctx, sc_build_cells_impl(ctx, r, shift, cell_of, cell_start, sorted, ncell_out))
}

// All pairs within rc on the device: cell list with cells >= rc, count pass,
// scan, write pass.  Leaves the pairs (IDs, i < j) in d_pi / d_pj.
static int device_pairs(sc_ctx* ctx, double rc, long* total) {
    int n = ctx->n;
    *total = 0;
    if (n < 2) return SC_OK;
    HIPCK(hipSetDevice(ctx->device));
    double vol = ctx->box[0] * ctx->box[1] * ctx->box[2];
    double w = std::max(rc, cbrt(2.0 * vol / n));   // cells >= rc holding a few particles
    uint32_t zero[3] = {0, 0, 0};
    int rc_ = build_grid_w(ctx, 1, zero, ctx->box, w, 1);
    if (rc_) return rc_;
    if (n + 1 > ctx->pcount_cap) {
        if ((rc_ = dalloc(ctx, &ctx->d_pcount, (size_t)n + 1)) || (rc_ = dalloc(ctx, &ctx->d_label, (size_t)n + 1)))
            return rc_;
        ctx->pcount_cap = n + 1;
    }
    int c = ctx->cur, nb = (n + 255) / 256;
    const Grid& g = ctx->g;
    double rc2 = rc * rc;
    // counts into d_label (scratch), offsets into d_pcount
    SC_LAUNCH(k_pair_list, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->P[c], ctx->ID[c], ctx->starts, g, rc2, ctx->box[0],
              ctx->box[1], ctx->box[2], ctx->d_label, (const int*)nullptr, (int*)nullptr, (int*)nullptr);
    SC_LAUNCH(k_scan, dim3(1), dim3(1024), 0, ctx->stream, n, ctx->d_label, ctx->d_pcount, ctx->d_maxocc,
              ctx->d_nonempty);
    int tot = 0;
    HIPCK(hipMemcpyAsync(&tot, ctx->d_pcount + n, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (tot < 0) return fail(ctx, SC_ECAPACITY, "more than 2^31 pairs");
    if (tot > ctx->pair_cap) {
        if ((rc_ = dalloc(ctx, &ctx->d_pi, (size_t)tot)) || (rc_ = dalloc(ctx, &ctx->d_pj, (size_t)tot)) ||
            (rc_ = dalloc(ctx, &ctx->d_pe, (size_t)tot)))
            return rc_;
        ctx->pair_cap = tot;
    }
    if (tot > 0)
        SC_LAUNCH(k_pair_list, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->P[c], ctx->ID[c], ctx->starts, g, rc2,
                  ctx->box[0], ctx->box[1], ctx->box[2], (int*)nullptr, (const int*)ctx->d_pcount, ctx->d_pi, ctx->d_pj);
    HIPCK(hipGetLastError());
    *total = tot;
    return SC_OK;
}

static int sc_neighbor_pairs_impl(sc_ctx* ctx, double rc, int32_t* pairs, long cap, long* count) {
    if (!ctx || !count || !(rc > 0)) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    long tot = 0;
    int rc_ = device_pairs(ctx, rc, &tot);
    if (rc_) return rc_;
    *count = tot;
    if (!pairs || tot == 0) return SC_OK;
    // returned in lexicographic (i, j) order
    std::vector<int32_t> hi(tot), hj(tot);
    HIPCK(hipMemcpyAsync(hi.data(), ctx->d_pi, tot * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipMemcpyAsync(hj.data(), ctx->d_pj, tot * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    std::vector<uint64_t> key(tot);
    for (long k = 0; k < tot; k++) key[k] = ((uint64_t)(uint32_t)hi[k] << 32) | (uint32_t)hj[k];
    std::sort(key.begin(), key.end());
    long m = std::min(cap, tot);
    for (long k = 0; k < m; k++) {
        pairs[2 * k] = (int32_t)(key[k] >> 32);
        pairs[2 * k + 1] = (int32_t)(key[k] & 0xffffffffu);
    }
    return SC_OK;
}

// ClusterSampler::sameCluster (clust.cpp:51-99): two particles share a
// cluster when they belong to one molecule or their pair energy is <=
// threshold (-0.10 in the reference); labels as ClusterSampler::genClusterList
// (clust.cpp:105-163) leaves them: the smallest index of the component.
// Pair energies in FP64 (k_pairs<double>) over all pairs within maxcut.
static int sc_cluster_labels_impl(sc_ctx* ctx, double threshold, int32_t* labels, int* nclusters) {
    if (!ctx || !nclusters) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    int n = ctx->n;
    long m = 0;
    int rc_ = device_pairs(ctx, ctx->maxcut, &m);
    if (rc_) return rc_;
    if (n + 1 > ctx->pcount_cap) {
        if ((rc_ = dalloc(ctx, &ctx->d_pcount, (size_t)n + 1)) || (rc_ = dalloc(ctx, &ctx->d_label, (size_t)n + 1)))
            return rc_;
        ctx->pcount_cap = n + 1;
    }
    if (!ctx->d_changed && (rc_ = dalloc(ctx, &ctx->d_changed, 1))) return rc_;
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    if (m > 0)
        SC_LAUNCH(k_pairs<double>, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, ctx->stream, (int)m, ctx->d_pi,
                  ctx->d_pj, ctx->slot, sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp64, ctx->d_tp64, ctx->sqmaxcut,
                  chain_tab(ctx), ctx->d_pe);
    SC_LAUNCH(k_iota, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_label);
    for (int it = 0; it < 4 * n + 64; it++) {
        int changed = 0;
        HIPCK(hipMemsetAsync(ctx->d_changed, 0, sizeof(int), ctx->stream));
        if (m > 0)
            SC_LAUNCH(k_cluster_hook, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream, m, ctx->d_pi,
                      ctx->d_pj, ctx->d_pe, threshold, ctx->d_label, ctx->d_changed);
        if (ctx->molecules)
            SC_LAUNCH(k_cluster_hook_mol, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_cw, ctx->d_label,
                      ctx->d_changed);
        SC_LAUNCH(k_cluster_jump, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->d_label, ctx->d_changed);
        HIPCK(hipMemcpyAsync(&changed, ctx->d_changed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCK(hipStreamSynchronize(ctx->stream));
        if (!changed) break;
    }
    std::vector<int32_t> lab(n);
    HIPCK(hipMemcpy(lab.data(), ctx->d_label, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    int nc = 0;
    for (int i = 0; i < n; i++) nc += lab[i] == i ? 1 : 0;
    *nclusters = nc;
    if (labels) memcpy(labels, lab.data(), n * sizeof(int32_t));
    return SC_OK;
}

// nematic order parameter of the particle axes: Q = (3 <u u> - I) / 2 and
// S2 = -2 x its middle eigenvalue (the convention of Tests/sc35.c:8563-8603);
// the director is the eigenvector of the largest eigenvalue
static void jacobi3(double a[3][3], double ev[3], double vec[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) vec[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off < 1e-30) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (fabs(a[p][q]) < 1e-300) continue;
                double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; k++) {
                    double kp = a[k][p], kq = a[k][q];
                    a[k][p] = cs * kp - sn * kq;
                    a[k][q] = sn * kp + cs * kq;
                }
                for (int k = 0; k < 3; k++) {
                    double pk = a[p][k], qk = a[q][k];
                    a[p][k] = cs * pk - sn * qk;
                    a[q][k] = sn * pk + cs * qk;
                }
                for (int k = 0; k < 3; k++) {
                    double kp = vec[k][p], kq = vec[k][q];
                    vec[k][p] = cs * kp - sn * kq;
                    vec[k][q] = sn * kp + cs * kq;
                }
            }
    }
    for (int i = 0; i < 3; i++) ev[i] = a[i][i];
}

static int sc_nematic_order_impl(sc_ctx* ctx, double* s2, double director[3]) {
    if (!ctx || !s2) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    int n = ctx->n, rc_;
    if (n <= 0) { *s2 = 0.0; return SC_OK; }
    HIPCK(hipSetDevice(ctx->device));
    if (!ctx->d_q6 && (rc_ = dalloc(ctx, &ctx->d_q6, 6))) return rc_;
    HIPCK(hipMemsetAsync(ctx->d_q6, 0, 6 * sizeof(double), ctx->stream));
    SC_LAUNCH(k_nematic_q, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, ctx->O[ctx->cur], ctx->d_q6);
    HIPCK(hipGetLastError());
    double q6[6];
    HIPCK(hipMemcpyAsync(q6, ctx->d_q6, sizeof(q6), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    double q[3][3] = {{q6[0], q6[3], q6[4]}, {q6[3], q6[1], q6[5]}, {q6[4], q6[5], q6[2]}}, ev[3], vec[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) q[a][b] = (q[a][b] * 3.0 / n - (a == b ? 1.0 : 0.0)) / 2.0;
    jacobi3(q, ev, vec);
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int x, int y) { return ev[x] < ev[y]; });
    *s2 = -2.0 * ev[order[1]];
    if (director)
        for (int k = 0; k < 3; k++) director[k] = vec[k][order[2]];
    return SC_OK;
}

int sc_neighbor_pairs(sc_ctx* ctx, double rc, int32_t* pairs, long cap, long* count) {
    SC_GUARD(ctx, sc_neighbor_pairs_impl(ctx, rc, pairs, cap, count))
}

int sc_cluster_labels(sc_ctx* ctx, double threshold, int32_t* labels, int* nclusters) {
    SC_GUARD(ctx, sc_cluster_labels_impl(ctx, threshold, labels, nclusters))
}

int sc_nematic_order(sc_ctx* ctx, double* s2, double director[3]) {
    SC_GUARD(ctx, sc_nematic_order_impl(ctx, s2, director))
}

static int sc_alignment_order_impl(sc_ctx* ctx, double* out) {
    if (!ctx || !out) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    long m = 0;
    int rc = device_pairs(ctx, 1.5, &m);
    if (rc) return rc;
    *out = 0.0;
    if (m == 0) return SC_OK;
    if ((rc = ensure_out(ctx, 1))) return rc;
    int c = ctx->cur;
    SC_LAUNCH(k_align_pairs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream, m, ctx->d_pi, ctx->d_pj,
              ctx->slot, ctx->P[c], ctx->O[c], ctx->box[0], ctx->box[1], ctx->box[2], ctx->d_pe);
    SC_LAUNCH(k_reduce, dim3(1), dim3(1024), 0, ctx->stream, (int)m, ctx->d_pe, ctx->d_out);
    HIPCK(hipMemcpyAsync(out, ctx->d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int sc_alignment_order(sc_ctx* ctx, double* out) {
    SC_GUARD(ctx, sc_alignment_order_impl(ctx, out))
}

static long put_text(const std::string& s, char* buf, long cap) {
    if (buf && cap > (long)s.size()) memcpy(buf, s.c_str(), s.size() + 1);
    return (long)s.size();
}

static long sc_write_movie_frame_impl(sc_ctx* ctx, long sweep, char* buf, long cap) {
    if (!ctx) return SC_EINVAL;
    int n = ctx->n;
    std::vector<double> pos(3 * n), dir(3 * n), pd0(3 * n);
    double box[3];
    int rc = sc_get_config(ctx, box, pos.data(), dir.data(), pd0.data());
    if (rc) return rc;
    return put_text(schost::movieFrame(sweep, box, pos, dir, pd0, ctx->topo.particleMolecules()), buf, cap);
}

long sc_write_movie_frame(sc_ctx* ctx, long sweep, char* buf, long cap) {
    SC_GUARD(ctx, sc_write_movie_frame_impl(ctx, sweep, buf, cap))
}

static long sc_write_stat_line_impl(sc_ctx* ctx, long sweep, char* buf, long cap) {
    if (!ctx) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    return put_text(schost::statLine(sweep, ctx->box[0] * ctx->box[1] * ctx->box[2]), buf, cap);
}

long sc_write_stat_line(sc_ctx* ctx, long sweep, char* buf, long cap) {
    SC_GUARD(ctx, sc_write_stat_line_impl(ctx, sweep, buf, cap))
}

static long sc_write_energy_line_impl(sc_ctx* ctx, long sweep, char* buf, long cap) {
    if (!ctx) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    double e = 0, al = 0;
    int rc = energy_all_impl(ctx, ctx->box, &e);
    if (rc) return rc;
    if ((rc = sc_alignment_order_impl(ctx, &al))) return rc;
    return put_text(schost::energyLine(sweep, e, al), buf, cap);
}

long sc_write_energy_line(sc_ctx* ctx, long sweep, char* buf, long cap) {
    SC_GUARD(ctx, sc_write_energy_line_impl(ctx, sweep, buf, cap))
}

static int sc_pair_energies_impl(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64);

// ClusterSampler::writeCluster (clust.cpp:7-49): labels as sc_cluster_labels;
// cluster energies as calcClusterEnergies (clust.h:100-111), the sum of p2p
// (smaller index as part1) over the pairs of each cluster: pairs within
// maxcut plus bonded pairs beyond it (FP64 pair energies).
static int sc_write_clusters_impl(sc_ctx* ctx, double threshold, long sweep, char* stat_buf, long stat_cap,
                                  char* cl_buf, long cl_cap, long* stat_len, long* cl_len) {
    if (!ctx || !stat_len || !cl_len) return SC_EINVAL;
    int n = ctx->n, ncl = 0;
    std::vector<int32_t> lab(n);
    int rc = sc_cluster_labels_impl(ctx, threshold, lab.data(), &ncl);
    if (rc) return rc;
    // pairs within maxcut and their FP64 energies, as the labelling left them
    long m = 0;
    if ((rc = device_pairs(ctx, ctx->maxcut, &m))) return rc;
    std::vector<int32_t> hi(m), hj(m);
    std::vector<double> he(m);
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    if (m > 0) {
        SC_LAUNCH(k_pairs<double>, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, ctx->stream, (int)m, ctx->d_pi,
                  ctx->d_pj, ctx->slot, sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp64, ctx->d_tp64, ctx->sqmaxcut,
                  chain_tab(ctx), ctx->d_pe);
        HIPCK(hipMemcpyAsync(hi.data(), ctx->d_pi, m * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPCK(hipMemcpyAsync(hj.data(), ctx->d_pj, m * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIPCK(hipMemcpyAsync(he.data(), ctx->d_pe, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCK(hipStreamSynchronize(ctx->stream));
    }
    std::vector<double> eroot(n, 0.0);
    std::vector<char> listed;
    std::vector<int32_t> bi, bj;
    if (ctx->bonds) {
        // bonded pairs (i, i+1), (i, i+2) that the list lacks
        std::vector<int> cw(n);
        HIPCK(hipMemcpy(cw.data(), ctx->d_cw, n * sizeof(int), hipMemcpyDeviceToHost));
        listed.assign(2 * (size_t)n, 0);
        for (long k = 0; k < m; k++) {
            int d = hj[k] - hi[k];
            if (d == 1 || d == 2) listed[2 * (size_t)hi[k] + (d - 1)] = 1;
        }
        for (int i = 0; i < n; i++)
            for (int d = 1; d <= 2; d++)
                if (((cw[i] >> cw_bit(d)) & 1) && !listed[2 * (size_t)i + (d - 1)]) {
                    bi.push_back(i);
                    bj.push_back(i + d);
                }
    }
    for (long k = 0; k < m; k++)
        if (lab[hi[k]] == lab[hj[k]]) eroot[lab[hi[k]]] += he[k];
    if (!bi.empty()) {
        std::vector<double> be(bi.size());
        if ((rc = sc_pair_energies_impl(ctx, (int)bi.size(), bi.data(), bj.data(), be.data(), 1))) return rc;
        for (size_t k = 0; k < bi.size(); k++)
            if (lab[bi[k]] == lab[bj[k]]) eroot[lab[bi[k]]] += be[k];
    }
    std::vector<int> labv(lab.begin(), lab.end());
    std::string st, cl;
    schost::clusterFiles(sweep, labv, eroot, st, cl);
    *stat_len = put_text(st, stat_buf, stat_cap);
    *cl_len = put_text(cl, cl_buf, cl_cap);
    return SC_OK;
}

int sc_write_clusters(sc_ctx* ctx, double threshold, long sweep, char* stat_buf, long stat_cap, char* cl_buf,
                      long cl_cap, long* stat_len, long* cl_len) {
    SC_GUARD(ctx, sc_write_clusters_impl(ctx, threshold, sweep, stat_buf, stat_cap, cl_buf, cl_cap, stat_len, cl_len))
}

static int sc_pair_energies_impl(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64) {
    if (!ctx || m < 0 || (m > 0 && (!i || !j || !out))) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    for (int k = 0; k < m; k++)
        if (i[k] < 0 || i[k] >= ctx->n || j[k] < 0 || j[k] >= ctx->n) return fail(ctx, SC_EINVAL, "particle index out of range");
    if (m == 0) return SC_OK;
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_ij(ctx, m)) || (rc = ensure_out(ctx, m))) return rc;
    HIPCK(hipMemcpyAsync(ctx->d_i, i, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->d_j, j, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    if (fp64)
        hipLaunchKernelGGL(k_pairs<double>, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, ctx->d_i, ctx->d_j,
                           ctx->slot, sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp64, ctx->d_tp64, ctx->sqmaxcut,
                           chain_tab(ctx), ctx->d_out);
    else
        hipLaunchKernelGGL(k_pairs<float>, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, ctx->d_i, ctx->d_j,
                           ctx->slot, sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp32, ctx->d_tp32, (float)ctx->sqmaxcut,
                           chain_tab(ctx), ctx->d_out);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, ctx->d_out, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int sc_pair_energies(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64) {
    SC_GUARD(ctx, sc_pair_energies_impl(ctx, m, i, j, out, fp64))
}

// sc_pair_energies through this device's `scratch` instead of the context's
// staging buffers: no allocation of any kind.  The kernel and its results
// are those of sc_pair_energies.  Calls that overlap on one GPU share the
// staging and race; nothing in it needs clearing (a call reads back only
// what its kernel wrote).
static int sc_pair_energies_noalloc_impl(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out,
                                         int fp64) {
    if (!ctx || m < 0 || (m > 0 && (!i || !j || !out))) return SC_EINVAL;
    if (ctx->device < 0) return SC_ENODEV;
    if (m > SC_NOALLOC_MAX_PAIRS)
        return fail(ctx, SC_ECAPACITY, "sc_pair_energies_noalloc: " + std::to_string(m) + " pairs exceed " +
                                           std::to_string(SC_NOALLOC_MAX_PAIRS));
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    for (int k = 0; k < m; k++)
        if (i[k] < 0 || i[k] >= ctx->n || j[k] < 0 || j[k] >= ctx->n) return fail(ctx, SC_EINVAL, "particle index out of range");
    if (m == 0) return SC_OK;
    if (!ctx->noalloc) return fail(ctx, SC_ESTATE, "no staging address for this device");
    HIPCK(hipSetDevice(ctx->device));
    int* di = (int*)ctx->noalloc;
    int* dj = di + SC_NOALLOC_MAX_PAIRS;
    double* dout = (double*)(ctx->noalloc + (size_t)SC_NOALLOC_MAX_PAIRS * 8);
    HIPCK(hipMemcpyAsync(di, i, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(dj, j, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    if (fp64)
        hipLaunchKernelGGL(k_pairs<double>, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, di, dj, ctx->slot, sc,
                           ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp64, ctx->d_tp64, ctx->sqmaxcut, chain_tab(ctx),
                           dout);
    else
        hipLaunchKernelGGL(k_pairs<float>, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, di, dj, ctx->slot, sc,
                           ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp32, ctx->d_tp32, (float)ctx->sqmaxcut,
                           chain_tab(ctx), dout);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, dout, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

int sc_pair_energies_noalloc(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, double* out, int fp64) {
    SC_GUARD(ctx, sc_pair_energies_noalloc_impl(ctx, m, i, j, out, fp64))
}

static int sc_energy_all_impl(sc_ctx* ctx, double* e) {
    if (!ctx || !e) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    HIPCK(hipSetDevice(ctx->device));
    return energy_all_impl(ctx, ctx->box, e);
}

int sc_energy_all(sc_ctx* ctx, double* e) {
    SC_GUARD(ctx, sc_energy_all_impl(ctx, e))
}

static int sc_energy_all_trial_impl(sc_ctx* ctx, const double box_new[3], double* e) {
    if (!ctx || !e || !box_new) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    for (int k = 0; k < 3; k++)
        if (!(box_new[k] > 2.0 * ctx->maxcut * 0.5)) return fail(ctx, SC_EINVAL, "trial box smaller than the interaction range");
    HIPCK(hipSetDevice(ctx->device));
    return energy_all_impl(ctx, box_new, e);
}

int sc_energy_all_trial(sc_ctx* ctx, const double box_new[3], double* e) {
    SC_GUARD(ctx, sc_energy_all_trial_impl(ctx, box_new, e))
}

static int sc_commit_box_impl(sc_ctx* ctx, const double box_new[3]) {
    if (!ctx || !box_new) return SC_EINVAL;
    for (int k = 0; k < 3; k++) {
        if (!(box_new[k] > 0)) return fail(ctx, SC_EINVAL, "box edges must be positive");
        ctx->box[k] = box_new[k];
    }
    ctx->e_valid = false;
    return build_grid(ctx, 2, nullptr, ctx->box);
}

int sc_commit_box(sc_ctx* ctx, const double box_new[3]) {
    SC_GUARD(ctx, sc_commit_box_impl(ctx, box_new))
}

static int run_queries(sc_ctx* ctx, int nq, const Query* hq, const int* slots_dev, double* out) {
    int rc;
    if ((rc = ensure_out(ctx, nq)) || (rc = ensure_ij(ctx, nq))) return rc;
    if (!ctx->grid_valid) {
        if ((rc = build_grid(ctx, 2, nullptr, ctx->box))) return rc;
    }
    if (hq) HIPCK(hipMemcpyAsync(ctx->d_q, hq, nq * sizeof(Query), hipMemcpyHostToDevice, ctx->stream));
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    hipLaunchKernelGGL(k_energy_query, dim3(nq), dim3(WAVE), 0, ctx->stream, nq, hq ? ctx->d_q : nullptr, slots_dev,
                       ctx->g, sc, ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32,
                       chain_tab(ctx), ctx->slot, ctx->d_out);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(out, ctx->d_out, nq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

static int sc_energy_one_impl(sc_ctx* ctx, int target, double* e) {
    if (!ctx || !e) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (target < 0 || target >= ctx->n) return fail(ctx, SC_EINVAL, "target out of range");
    HIPCK(hipSetDevice(ctx->device));
    int s = 0;
    HIPCK(hipMemcpy(&s, ctx->slot + target, sizeof(int), hipMemcpyDeviceToHost));
    int rc;
    if ((rc = ensure_ij(ctx, 1))) return rc;
    HIPCK(hipMemcpy(ctx->d_i, &s, sizeof(int), hipMemcpyHostToDevice));
    return run_queries(ctx, 1, nullptr, ctx->d_i, e);
}

int sc_energy_one(sc_ctx* ctx, int target, double* e) {
    SC_GUARD(ctx, sc_energy_one_impl(ctx, target, e))
}

static int sc_energy_per_particle_impl(sc_ctx* ctx, double* out) {
    if (!ctx || !out) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    HIPCK(hipSetDevice(ctx->device));
    int n = ctx->n, rc;
    if (n == 0) return SC_OK;
    if ((rc = ensure_ij(ctx, n))) return rc;
    if (!ctx->grid_valid && (rc = build_grid(ctx, 2, nullptr, ctx->box))) return rc;
    // queries in slot order (query q = slot q), results mapped back through slot
    std::vector<int> sl(n), iota(n);
    for (int k = 0; k < n; k++) iota[k] = k;
    HIPCK(hipMemcpy(ctx->d_i, iota.data(), n * sizeof(int), hipMemcpyHostToDevice));
    std::vector<double> bySlot(n);
    if ((rc = run_queries(ctx, n, nullptr, ctx->d_i, bySlot.data()))) return rc;
    HIPCK(hipMemcpy(sl.data(), ctx->slot, n * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[i] = bySlot[sl[i]];
    return SC_OK;
}

int sc_energy_per_particle(sc_ctx* ctx, double* out) {
    SC_GUARD(ctx, sc_energy_per_particle_impl(ctx, out))
}

static Query make_query(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3]) {
    Query q;
    for (int k = 0; k < 3; k++) {
        double s = pos[k] - floor(pos[k]);
        double v = nearbyint(s * 4294967296.0);
        q.u[k] = (uint32_t)(uint64_t)(v >= 4294967296.0 ? 0.0 : v);
    }
    double d[3] = {dir[0], dir[1], dir[2]}, p[3] = {pd0[0], pd0[1], pd0[2]};
    double t = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (t != 0.0) { t = 1.0 / t; d[0] *= t; d[1] *= t; d[2] *= t; }
    double dp = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
    p[0] -= dp * d[0]; p[1] -= dp * d[1]; p[2] -= dp * d[2];
    t = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (t != 0.0) { t = 1.0 / t; p[0] *= t; p[1] *= t; p[2] *= t; }
    q.t = ctx->ref2dense[ctx->ptypes[target]];
    q.o = make_float4((float)d[0], (float)d[1], (float)d[2], (float)p[0]);
    q.q = make_float2((float)p[1], (float)p[2]);
    q.id = q.ex_lo = q.ex_hi = target;
    q.bonds = 1;
    return q;
}

static int sc_energy_one_trial_impl(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3],
                        double* e) {
    if (!ctx || !e || !pos || !dir || !pd0) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (target < 0 || target >= ctx->n) return fail(ctx, SC_EINVAL, "target out of range");
    HIPCK(hipSetDevice(ctx->device));
    Query q = make_query(ctx, target, pos, dir, pd0);
    return run_queries(ctx, 1, &q, nullptr, e);
}

int sc_energy_one_trial(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3],
                        double* e) {
    SC_GUARD(ctx, sc_energy_one_trial_impl(ctx, target, pos, dir, pd0, e))
}

static int sc_commit_particle_impl(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3]) {
    if (!ctx || !pos || !dir || !pd0) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (target < 0 || target >= ctx->n) return fail(ctx, SC_EINVAL, "target out of range");
    HIPCK(hipSetDevice(ctx->device));
    Query q = make_query(ctx, target, pos, dir, pd0);
    int s = 0, c = ctx->cur;
    HIPCK(hipMemcpy(&s, ctx->slot + target, sizeof(int), hipMemcpyDeviceToHost));
    uint4 p = make_uint4(q.u[0], q.u[1], q.u[2], (uint32_t)q.t);
    HIPCK(hipMemcpy(ctx->P[c] + s, &p, sizeof(uint4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->O[c] + s, &q.o, sizeof(float4), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->Q[c] + s, &q.q, sizeof(float2), hipMemcpyHostToDevice));
    ctx->e_valid = false;
    // the particle may have changed cell: rebuild the cell list
    return build_grid(ctx, ctx->grid_valid ? ctx->g.r : 2, nullptr, ctx->box);
}

int sc_commit_particle(sc_ctx* ctx, int target, const double pos[3], const double dir[3], const double pd0[3]) {
    SC_GUARD(ctx, sc_commit_particle_impl(ctx, target, pos, dir, pd0))
}

// mol2others / mol2othersTrial: one query per member of [first, first +
// size), each leaving the whole molecule out; the stored state of the
// members (pos == nullptr) or the given trial states
static int sc_energy_mol_impl(sc_ctx* ctx, int first, int size, const double* pos, const double* dir,
                              const double* pd0, double* e) {
    if (!ctx || !e || size < 1 || (pos && (!dir || !pd0))) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (first < 0 || first + size > ctx->n) return fail(ctx, SC_EINVAL, "molecule out of range");
    HIPCK(hipSetDevice(ctx->device));
    std::vector<Query> q(size);
    if (pos) {
        for (int a = 0; a < size; a++) q[a] = make_query(ctx, first + a, pos + 3 * a, dir + 3 * a, pd0 + 3 * a);
    } else {
        std::vector<int> sl(size);
        HIPCK(hipMemcpy(sl.data(), ctx->slot + first, size * sizeof(int), hipMemcpyDeviceToHost));
        int c = ctx->cur;
        for (int a = 0; a < size; a++) {
            uint4 p;
            HIPCK(hipMemcpy(&p, ctx->P[c] + sl[a], sizeof(uint4), hipMemcpyDeviceToHost));
            HIPCK(hipMemcpy(&q[a].o, ctx->O[c] + sl[a], sizeof(float4), hipMemcpyDeviceToHost));
            HIPCK(hipMemcpy(&q[a].q, ctx->Q[c] + sl[a], sizeof(float2), hipMemcpyDeviceToHost));
            q[a].u[0] = p.x; q[a].u[1] = p.y; q[a].u[2] = p.z;
            q[a].t = (int)p.w;
            q[a].id = first + a;
        }
    }
    for (int a = 0; a < size; a++) {
        q[a].ex_lo = first;
        q[a].ex_hi = first + size - 1;
        q[a].bonds = 0;
    }
    std::vector<double> out(size);
    int rc = run_queries(ctx, size, q.data(), nullptr, out.data());
    if (rc) return rc;
    double s = 0.0;
    for (double v : out) s += v;
    *e = s;
    return SC_OK;
}

int sc_energy_mol(sc_ctx* ctx, int first, int size, double* e) {
    SC_GUARD(ctx, sc_energy_mol_impl(ctx, first, size, nullptr, nullptr, nullptr, e))
}

int sc_energy_mol_trial(sc_ctx* ctx, int first, int size, const double* pos, const double* dir, const double* pd0,
                        double* e) {
    if (!pos) return SC_EINVAL;
    SC_GUARD(ctx, sc_energy_mol_impl(ctx, first, size, pos, dir, pd0, e))
}

static int sc_commit_mol_impl(sc_ctx* ctx, int first, int size, const double* pos, const double* dir,
                              const double* pd0) {
    if (!ctx || !pos || !dir || !pd0 || size < 1) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (first < 0 || first + size > ctx->n) return fail(ctx, SC_EINVAL, "molecule out of range");
    HIPCK(hipSetDevice(ctx->device));
    std::vector<int> sl(size);
    HIPCK(hipMemcpy(sl.data(), ctx->slot + first, size * sizeof(int), hipMemcpyDeviceToHost));
    int c = ctx->cur;
    for (int a = 0; a < size; a++) {
        Query q = make_query(ctx, first + a, pos + 3 * a, dir + 3 * a, pd0 + 3 * a);
        uint4 p = make_uint4(q.u[0], q.u[1], q.u[2], (uint32_t)q.t);
        HIPCK(hipMemcpy(ctx->P[c] + sl[a], &p, sizeof(uint4), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->O[c] + sl[a], &q.o, sizeof(float4), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(ctx->Q[c] + sl[a], &q.q, sizeof(float2), hipMemcpyHostToDevice));
    }
    ctx->e_valid = false;
    return build_grid(ctx, ctx->grid_valid ? ctx->g.r : 2, nullptr, ctx->box);
}

int sc_commit_mol(sc_ctx* ctx, int first, int size, const double* pos, const double* dir, const double* pd0) {
    SC_GUARD(ctx, sc_commit_mol_impl(ctx, first, size, pos, dir, pd0))
}

static int sc_overlap_flags_impl(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, uint8_t* out, double* dmin) {
    if (!ctx || m < 0 || (m > 0 && (!i || !j || !out))) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    for (int k = 0; k < m; k++)
        if (i[k] < 0 || i[k] >= ctx->n || j[k] < 0 || j[k] >= ctx->n) return fail(ctx, SC_EINVAL, "particle index out of range");
    if (m == 0) return SC_OK;
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_ij(ctx, m)) || (rc = ensure_out(ctx, m))) return rc;
    unsigned char* d_flags = nullptr;
    HIPCK(hipMalloc((void**)&d_flags, m));
    HIPCK(hipMemcpyAsync(ctx->d_i, i, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipMemcpyAsync(ctx->d_j, j, m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    SysConst sc = sysconst(ctx, ctx->box);
    int c = ctx->cur;
    hipLaunchKernelGGL(k_overlap, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, ctx->d_i, ctx->d_j, ctx->slot,
                       sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp64, ctx->d_tp64, d_flags, ctx->d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_flags, m, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && dmin) e = hipMemcpyAsync(dmin, ctx->d_out, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(d_flags);
    if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
    return SC_OK;
}

int sc_overlap_flags(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, uint8_t* out, double* dmin) {
    SC_GUARD(ctx, sc_overlap_flags_impl(ctx, m, i, j, out, dmin))
}

static int sc_overlap_count_impl(sc_ctx* ctx, long* count) {
    if (!ctx || !count) return SC_EINVAL;
    long np = 0;
    int rc = sc_neighbor_pairs(ctx, ctx->maxcut, nullptr, 0, &np);
    if (rc) return rc;
    std::vector<int32_t> pr(2 * np);
    if ((rc = sc_neighbor_pairs(ctx, ctx->maxcut, pr.data(), np, &np))) return rc;
    std::vector<int32_t> a(np), b(np);
    for (long k = 0; k < np; k++) { a[k] = pr[2 * k]; b[k] = pr[2 * k + 1]; }
    std::vector<uint8_t> f(np);
    if ((rc = sc_overlap_flags(ctx, (int)np, a.data(), b.data(), f.data(), nullptr))) return rc;
    long c = 0;
    for (long k = 0; k < np; k++) c += f[k];
    *count = c;
    return SC_OK;
}

int sc_overlap_count(sc_ctx* ctx, long* count) {
    SC_GUARD(ctx, sc_overlap_count_impl(ctx, count))
}

// host-side stream of the sweep schedule (grid shifts, colour orders)
static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Grid shift and colour order of sweep `sw` under `seed`: the shift is three
// random 32-bit fractions, the order a random permutation of the p^3 colours.
struct SweepPlan {
    uint32_t shift[3];
    int order[27];
};
static void sweep_plan(uint64_t seed, uint64_t sw, int p, SweepPlan& pl) {
    uint64_t h = mix64(seed ^ mix64(sw * 0x632BE59BD9B4E019ull));
    pl.shift[0] = (uint32_t)h;
    pl.shift[1] = (uint32_t)(h >> 32);
    pl.shift[2] = (uint32_t)mix64(h);
    int ncol = p * p * p;
    for (int k = 0; k < ncol; k++) pl.order[k] = k;
    uint64_t hs = mix64(h ^ 0xA5A5A5A5ull);
    for (int k = ncol - 1; k > 0; k--) {
        hs = mix64(hs);
        int j = (int)(hs % (uint64_t)(k + 1));
        std::swap(pl.order[k], pl.order[j]);
    }
}

// ---- whole-molecule moves of the sweep ---------------------------------------
static int chain_prepare(sc_ctx* ctx) {
    int nch = (int)ctx->chain_first.size();
    if (ctx->nchain_dev == nch && ctx->d_ccell_id) return SC_OK;
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_cfirst, nch)) || (rc = dalloc(ctx, &ctx->d_csize, nch)) ||
        (rc = dalloc(ctx, &ctx->d_ccell_chain, nch)) || (rc = dalloc(ctx, &ctx->d_ccell_id, ctx->n)) ||
        (rc = dalloc(ctx, &ctx->d_stamp, ctx->n)) || (rc = dalloc(ctx, &ctx->d_vol, ctx->nt)) ||
        (rc = dalloc(ctx, &ctx->d_clist, nch)))
        return rc;
    HIPCK(hipMemcpy(ctx->d_cfirst, ctx->chain_first.data(), nch * sizeof(int), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->d_csize, ctx->chain_size.data(), nch * sizeof(int), hipMemcpyHostToDevice));
    HIPCK(hipMemset(ctx->d_ccell_id, 0xff, ctx->n * sizeof(int)));   // atomic particles: -1 for good
    HIPCK(hipMemset(ctx->d_stamp, 0, ctx->n * sizeof(uint32_t)));
    std::vector<double> vol(ctx->nt);
    for (int d = 0; d < ctx->nt; d++) {
        int r = ctx->dense2ref[d];
        vol[d] = ctx->topo.ia[r][r].volume;
    }
    HIPCK(hipMemcpy(ctx->d_vol, vol.data(), ctx->nt * sizeof(double), hipMemcpyHostToDevice));
    ctx->chain_stamp = 0;
    ctx->nchain_dev = nch;
    return SC_OK;
}

// largest centre-to-particle distance over the molecules (the centre of
// k_chain_anchor, in FP64 on the host)
static int chain_max_extent(sc_ctx* ctx, double* out) {
    std::vector<uint4> hp;
    std::vector<float4> ho;
    std::vector<float2> hq;
    int rc = download(ctx, hp, ho, hq);
    if (rc) return rc;
    const double k = 2.3283064365386963e-10;
    double mx = 0.0;
    for (size_t c = 0; c < ctx->chain_first.size(); c++) {
        int f = ctx->chain_first[c], m = ctx->chain_size[c];
        std::vector<double> d(3 * m);
        double vs = 0.0, cen[3] = {0, 0, 0};
        for (int a = 0; a < m; a++) {
            const uint4& u = hp[f + a];
            const uint4& u0 = hp[f];
            int r = ctx->dense2ref[u.w];
            double v = ctx->topo.ia[r][r].volume;
            d[3 * a] = (double)(int32_t)(u.x - u0.x) * (ctx->box[0] * k);
            d[3 * a + 1] = (double)(int32_t)(u.y - u0.y) * (ctx->box[1] * k);
            d[3 * a + 2] = (double)(int32_t)(u.z - u0.z) * (ctx->box[2] * k);
            for (int j = 0; j < 3; j++) cen[j] += v * d[3 * a + j];
            vs += v;
        }
        for (int a = 0; a < m; a++) {
            double e2 = 0.0;
            for (int j = 0; j < 3; j++) {
                double x = d[3 * a + j] - cen[j] / vs;
                e2 += x * x;
            }
            mx = std::max(mx, sqrt(e2));
        }
    }
    *out = mx;
    return SC_OK;
}

static int sc_set_chain_moves_impl(sc_ctx* ctx, double chainprob, double chain_mx, double chain_angle, double extent) {
    if (!ctx || !(chainprob >= 0.0 && chainprob <= 1.0) || !(chain_mx >= 0.0) || !(chain_angle >= 0.0) ||
        !(chain_angle <= 1.5707963267948966) || !(extent >= 0.0))
        return SC_EINVAL;
    if (chainprob == 0.0) {
        ctx->chainprob = 0.0;
        return SC_OK;
    }
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (ctx->nrep > 0) return fail(ctx, SC_ESTATE, "replica batches hold atomic molecules only");
    for (int m : ctx->chain_size)
        if (m > CHAIN_MAXM) return fail(ctx, SC_ECAPACITY, "chain moves take molecules of at most 64 particles");
    HIPCK(hipSetDevice(ctx->device));
    double rho = extent;
    if (rho == 0.0) {
        int rc = chain_max_extent(ctx, &rho);
        if (rc) return rc;
        rho = 1.2 * rho + 0.01;
    }
    ctx->chainprob = chainprob;
    ctx->chain_mx = chain_mx;
    ctx->chain_angle = chain_angle;
    ctx->chain_rho = rho;
    return SC_OK;
}

int sc_set_chain_moves(sc_ctx* ctx, double chainprob, double chain_mx, double chain_angle, double extent) {
    SC_GUARD(ctx, sc_set_chain_moves_impl(ctx, chainprob, chain_mx, chain_angle, extent))
}

int sc_get_chain_extent(sc_ctx* ctx, double* extent) {
    if (!ctx || !extent) return SC_EINVAL;
    *extent = ctx->chain_rho;
    return SC_OK;
}

static int sc_set_chain_visit_factor_impl(sc_ctx* ctx, int v) {
    if (!ctx || v < 1 || v > 27) return SC_EINVAL;
    ctx->chain_visit = v;
    return SC_OK;
}

int sc_set_chain_visit_factor(sc_ctx* ctx, int v) {
    SC_GUARD(ctx, sc_set_chain_visit_factor_impl(ctx, v))
}

int sc_chain_visit_info(const sc_ctx* ctx, int* visit) {
    if (!ctx || !visit) return SC_EINVAL;
    *visit = ctx->chain_visit_eff;
    return SC_OK;
}

int sc_chain_pass_info(const sc_ctx* ctx, int* period, int* staged) {
    if (!ctx || !period || !staged) return SC_EINVAL;
    *period = ctx->chain_period;
    *staged = ctx->chain_staged ? 1 : 0;
    return SC_OK;
}

// The chain-move passes of sweep `sw`: anchors on a coarse grid with its own
// random shift, then one pass per colour (random order), each followed by a
// rebuild of the fine cell list (moved particles may have changed cell).
static int chain_passes(sc_ctx* ctx, uint64_t seed, uint64_t sw, double temper, int r, const uint32_t* shift) {
    int nch = (int)ctx->chain_first.size();
    if (nch == 0 || ctx->chainprob == 0.0) return SC_OK;
    int rc;
    if ((rc = chain_prepare(ctx))) return rc;
    // coarse cells: 2 rho + maxcut.  Box edges: that much (no molecule reaches
    // its own periodic image) and more than 4 rho (the minimum image from a
    // molecule's first particle unwraps it)
    double wmin = 2.0 * ctx->chain_rho + ctx->maxcut;
    for (int k = 0; k < 3; k++)
        if (!(ctx->box[k] >= wmin) || !(ctx->box[k] > 4.0 * ctx->chain_rho))
            return fail(ctx, SC_EINVAL,
                        "chain moves need box edges of at least 2 x molecule extent + maxcut and above 4 x the "
                        "extent (a molecule must not reach its own periodic image)");
    // colour period p: same-colour cells are p - 1 cells apart, so cells of
    // wmin / (p - 1); the period with the most active cells per pass (the
    // passes run their cells' trials in sequence)
    int p = 2;
    long best = 0;
    for (int q = 2; q <= 3; q++) {
        long act = 1;
        for (int k = 0; k < 3; k++) {
            int nc = grid_axis_cells(ctx->box[k], wmin / (q - 1), q);
            act *= nc >= q ? nc / q : 1;
        }
        if (act > best) { best = act; p = q; }
    }
    uint64_t cseed = mix64(seed ^ 0x6368616E6D6F7665ull);
    SweepPlan plan;
    sweep_plan(cseed, sw, p, plan);
    // visit factor: the passes visit the first 1/v of the colours in the
    // sweep's random order, each with v times the trials (v the largest
    // divisor of the colour count not above the context's chain factor); a
    // move never leaves its anchor's coarse cell, so a cell's staged
    // neighbourhood holds for all its trials
    Grid gc;
    gc.r = 1;
    for (int k = 0; k < 3; k++) {
        gc.nc[k] = grid_axis_cells(ctx->box[k], wmin / (p - 1), p);
        gc.sh[k] = plan.shift[k];
    }
    int ncc = gc.nc[0] * gc.nc[1] * gc.nc[2];
    if (ncc > ctx->ccell_alloc) {
        if ((rc = dalloc(ctx, &ctx->d_ccount, ncc)) || (rc = dalloc(ctx, &ctx->d_cfill, ncc)) ||
            (rc = dalloc(ctx, &ctx->d_cstart, ncc + 3)))   // + the scan's occupancy outputs
            return rc;
        ctx->ccell_alloc = ncc;
    }
    HIPCK(hipMemsetAsync(ctx->d_ccount, 0, ncc * sizeof(int), ctx->stream));
    HIPCK(hipMemsetAsync(ctx->d_cfill, 0, ncc * sizeof(int), ctx->stream));
    SysConst sc = sysconst(ctx, ctx->box);
    double rho2 = ctx->chain_rho * ctx->chain_rho;
    int c = ctx->cur;
    hipLaunchKernelGGL(k_chain_anchor, dim3((nch + 3) / 4), dim3(256), 0, ctx->stream, nch, ctx->d_cfirst,
                       ctx->d_csize, ctx->slot, ctx->P[c], ctx->d_vol, sc, gc, rho2, ctx->d_ccell_chain,
                       ctx->d_ccell_id, ctx->d_ccount);
    // molecules by coarse cell, in index order
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, ncc, ctx->d_ccount, ctx->d_cstart,
                       ctx->d_cstart + ncc + 1, ctx->d_cstart + ncc + 2);
    hipLaunchKernelGGL(k_cell_scatter, dim3((nch + 255) / 256), dim3(256), 0, ctx->stream, nch, ctx->d_ccell_chain,
                       ctx->d_cstart, ctx->d_cfill, ctx->d_clist);
    hipLaunchKernelGGL(k_chain_sort, dim3(ncc), dim3(256), 0, ctx->stream, ncc, ctx->d_cstart, ctx->d_clist);
    HIPCK(hipGetLastError());
    ChainMoveParams cp;
    cp.temper = temper;
    cp.chain_mx = ctx->chain_mx;
    cp.chain_angle = ctx->chain_angle;
    cp.rho2 = rho2;
    // MoveCreator::chainMove is tried with probability chainprob at each of
    // the N steps of a sweep (updater.cpp:206-217): chainprob N trials
    int pa[3], ncv = 1;
    for (int k = 0; k < 3; k++) {
        pa[k] = gc.nc[k] >= p ? p : 1;
        ncv *= pa[k];
    }
    int vis = 1;
    for (int v = 1; v <= std::min(ctx->chain_visit, ncv); v++)
        if (ncv % v == 0) vis = v;
    ctx->chain_visit_eff = vis;
    cp.frac = ctx->chainprob * (double)ctx->n / (double)nch * (double)vis;
    cp.key0 = (uint32_t)cseed;
    cp.key1 = (uint32_t)(cseed >> 32);
    cp.nchain = nch;
    for (int k = 0; k < 3; k++) {
        cp.pa[k] = pa[k];
        cp.nact[k] = gc.nc[k] / pa[k];
    }
    size_t lds = chain_pass_lds(CHAIN_NW);
    // the staged pass when a region (the coarse cell widened by rho + maxcut)
    // holds at most `cap` particles with a margin of 1.12 over the mean density
    int cap = 64;
    while (chain_st_layout(CHAIN_NW, cap + 64).total <= CHAIN_LDS_BUDGET) cap += 64;
    double reach = ctx->chain_rho + ctx->maxcut, vreg = 1.0;
    long nrc = 1;
    for (int k = 0; k < 3; k++) {
        double wk = ctx->box[k] / gc.nc[k], ext = std::min(ctx->box[k], wk + 2.0 * reach);
        vreg *= ext;
        nrc *= std::min((long)ctx->g.nc[k], (long)ceil(ext / (ctx->box[k] / ctx->g.nc[k])) + 1);
        cp.reach[k] = (uint32_t)std::min(4294967295.0, ceil(reach / ctx->box[k] * 4294967296.0) + 1.0);
    }
    double boxv = ctx->box[0] * ctx->box[1] * ctx->box[2];
    // and when a cell's molecules and their particles fit their lists (1.25 x the mean)
    double mol_cell = (double)nch / (double)ncc, part_mol = 0.0;
    for (int m : ctx->chain_size) part_mol += m;
    part_mol /= (double)nch;
    bool staged = 1.12 * (double)ctx->n * vreg / boxv <= (double)cap && nrc <= CHAIN_ST_RCAP && ctx->g.r == 1 &&
                  1.25 * mol_cell <= CHAIN_ST_LCAP && 1.25 * mol_cell * part_mol <= CHAIN_ST_OCAP;
    size_t lds_st = chain_st_layout(CHAIN_NW, cap).total;
    ctx->chain_staged = staged;
    ctx->chain_period = p;
    int visited = 0;
    for (int ci = 0; ci < p * p * p && visited < ncv / vis; ci++) {
        int col = plan.order[ci];
        cp.ox = col % p;
        cp.oy = (col / p) % p;
        cp.oz = col / (p * p);
        if (cp.ox >= pa[0] || cp.oy >= pa[1] || cp.oz >= pa[2]) continue;
        visited++;
        uint64_t pass = sw * 32 + (uint64_t)ci;
        cp.pass_lo = (uint32_t)pass;
        cp.pass_hi = (uint32_t)(pass >> 32);
        cp.stamp = ++ctx->chain_stamp;
        if (cp.stamp == 0) {   // wrapped: clear the stamps once
            HIPCK(hipMemsetAsync(ctx->d_stamp, 0, ctx->n * sizeof(uint32_t), ctx->stream));
            cp.stamp = ctx->chain_stamp = 1;
        }
        c = ctx->cur;
        int nblk = cp.nact[0] * cp.nact[1] * cp.nact[2];
        if (staged)
            hipLaunchKernelGGL((k_chain_pass_st<CHAIN_NW>), dim3(nblk), dim3(WAVE * CHAIN_NW), lds_st, ctx->stream,
                               ctx->g, gc, sc, cp, cap, ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c],
                               ctx->slot, ctx->d_pp32, ctx->d_tp32, chain_tab(ctx), ctx->d_vol, ctx->d_cfirst,
                               ctx->d_csize, ctx->d_cstart, ctx->d_clist, ctx->d_ccell_id, ctx->d_stats, ctx->d_dsum);
        else
            hipLaunchKernelGGL((k_chain_pass<CHAIN_NW>), dim3(nblk), dim3(WAVE * CHAIN_NW), lds, ctx->stream, ctx->g,
                               gc, sc, cp, ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->slot,
                               ctx->d_pp32, ctx->d_tp32, chain_tab(ctx), ctx->d_vol, ctx->d_cfirst, ctx->d_csize,
                               ctx->d_cstart, ctx->d_clist, ctx->d_ccell_id, ctx->d_stamp, ctx->d_stats, ctx->d_dsum,
                               ctx->d_flag);
        HIPCK(hipGetLastError());
        if ((rc = build_grid(ctx, r, shift, ctx->box, ctx->max_active))) return rc;
    }
    return SC_OK;
}

static int sc_sweep_impl(sc_ctx* ctx, int nsweeps, double temper, double transmx, double rot_half, int r, uint64_t seed,
             uint64_t counter, sc_sweep_stats* stats) {
    if (!ctx || nsweeps < 0 || !(temper > 0) || r < 0 || r > 2) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (ctx->nrep > 0) return fail(ctx, SC_ESTATE, "the context holds a replica batch: use sc_batch_sweep");
    if (r == 0) r = ctx->tuned_r;
    HIPCK(hipSetDevice(ctx->device));
    bool carried = ctx->e_valid;   // the carried energy survives only a sweep that completes
    ctx->e_valid = false;
    HIPCK(hipMemsetAsync(ctx->d_stats, 0, STAT_REP * STAT_STRIDE * sizeof(unsigned long long), ctx->stream));
    HIPCK(hipMemsetAsync(ctx->d_dsum, 0, STAT_REP * STAT_STRIDE * sizeof(double), ctx->stream));
    int p = r + 1;
    long launches = 0;
    HIPCK(hipEventRecord(ctx->ev0, ctx->stream));
    for (int s = 0; s < nsweeps; s++) {
        uint64_t sw = counter + (uint64_t)s;
        SweepPlan plan;
        sweep_plan(seed, sw, p, plan);
        const uint32_t* shift = plan.shift;
        int rc = build_grid(ctx, r, shift, ctx->box, ctx->max_active);
        if (rc) return rc;
        const Grid& g = ctx->g;
        int cap = stage_cap(ctx, g, ctx->box);
        int mode = ctx->da_mode;
        bool da = mode == DA_NET;
        // the speculative variant: the stage that fits beside SPEC_W trials' lists
        bool spec = ctx->spec && ctx->sweep_nw == 0;
        const bool spec_lt = sweep_table_bytes(ctx->nt) > 0;   // the tables fit in LDS beside the stage
        int cap_spec = cap;
        size_t lds_spec = 0;
        if (spec) {
            size_t tabb = sweep_table_bytes(ctx->nt);
            long room = (long)SWEEP_LDS_BUDGET - (long)spec_layout(0, SPEC_LCAP, SPEC_XCAP, da, tabb).end - 16;
            int cmax = room > 0 ? (int)(room / (long)stage_bytes(1)) : 0;
            if (cap_spec > cmax) cap_spec = (cmax / 64) * 64;
            spec = cap_spec >= 64;
            if (spec) lds_spec = spec_layout(cap_spec, SPEC_LCAP, SPEC_XCAP, da, tabb).end;
        }
        if (da && (ctx->mlp_type < 0 || ctx->ref2dense[ctx->mlp_type] < 0))
            return fail(ctx, SC_EINVAL, "the network screen needs the network's particle type in the system");
        MlpConst mlpc = ctx->mlp_c;
        // the network screens trials whose phase C would take more than one
        // wave's round (2 sides x ppmax patch pairs per entry), unless set
        mlpc.minx = ctx->net_minx >= 0 ? ctx->net_minx : WAVE / (2 * std::max(1, ctx->ppmax)) + 1;
        mlpc.tnet = ctx->mlp_type >= 0 ? ctx->ref2dense[ctx->mlp_type] : -1;
        // LDS: stage + near/expensive lists + segment distances + moving particle
        // + stencil scratch (+ the network and its activation tile)
        // (+ chain partners, block compaction counts, block sums, trial records)
        size_t fixed = sizeof(float) * 2 * MV_STRIDE + sizeof(int) * (2 * 125 + 2) + SWEEP_SMALL_LDS +
                       (da ? MLP_LDS_BYTES + MLP_HX_BYTES + 48 : 0) +
                       sweep_table_bytes(ctx->nt);
        size_t per = stage_bytes(1) + sizeof(int) * 3 + (da ? sizeof(float) * 2 : 0);
        int capmax = (int)((SWEEP_LDS_BUDGET - fixed) / per);
        if (cap > capmax) cap = (capmax / 64) * 64;
        size_t lds = per * cap + fixed;
        int ncol = p * p * p;
        int pa[3];
        for (int k = 0; k < 3; k++) pa[k] = (g.nc[k] >= p) ? p : 1;
        // the colours of this sweep in its random order (a single-cell axis
        // has period 1: colour offsets along it stay 0); with a visit factor
        // v the sweep takes the first 1/v of them, each with v times the trials
        int cols[27], ncv = 0;
        for (int ci = 0; ci < ncol; ci++) {
            int col = plan.order[ci];
            if (col % p < pa[0] && (col / p) % p < pa[1] && col / (p * p) < pa[2]) cols[ncv++] = col;
        }
        if (ncv % ctx->visit) return fail(ctx, SC_EINVAL, "the visit factor must divide the number of colours");
        ncv /= ctx->visit;
        SysConst sc = sysconst(ctx, ctx->box);
        int c = ctx->cur;
        for (int ci = 0; ci < ncv; ci++) {
            int col = cols[ci];
            MoveParams mp;
            mp.ox = col % p; mp.oy = (col / p) % p; mp.oz = col / (p * p);
            mp.p = p;
            for (int k = 0; k < 3; k++) mp.nact[k] = g.nc[k] / pa[k];
            mp.trial_mode = ctx->trial_mode;
            mp.hcap = ctx->home_cache;
            mp.visit = ctx->visit;
            mp.n = ctx->n;
            mp.transmx = (float)transmx;
            mp.rot_half = (float)rot_half;
            mp.temper = temper;
            mp.key0 = (uint32_t)seed;
            mp.key1 = (uint32_t)(seed >> 32);
            uint64_t pass = sw * 32 + (uint64_t)ci;
            mp.pass_lo = (uint32_t)pass;
            mp.pass_hi = (uint32_t)(pass >> 32);
            int nblk = mp.nact[0] * mp.nact[1] * mp.nact[2];
            if (ctx->profile) {
                while ((long)ctx->kev.size() < 2 * (launches + 1)) {
                    hipEvent_t e;
                    HIPCK(hipEventCreate(&e));
                    ctx->kev.push_back(e);
                }
                HIPCK(hipEventRecord(ctx->kev[2 * launches], ctx->stream));
            }
#define SWEEP_PASS(DAM, NWV, IMG, MC)                                                                          \
    hipLaunchKernelGGL((k_sweep_pass<DAM, NWV>), dim3(nblk), dim3(WAVE * NWV), lds, ctx->stream, g, sc, mp,         \
                       ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32, cap,     \
                       ctx->ppmax, ctx->d_nonempty, IMG, MC, chain_tab(ctx), ctx->d_stats, ctx->d_dsum, ctx->d_flag)
#define SPEC_PASS(DAM, IMG, MC)                                                                                \
    if (spec_lt) SPEC_PASS_LT(DAM, true, IMG, MC); else SPEC_PASS_LT(DAM, false, IMG, MC)
#define SPEC_PASS_LT(DAM, LT, IMG, MC)                                                                         \
    hipLaunchKernelGGL((k_sweep_spec<DAM, LT>), dim3(nblk), dim3(WAVE * SPEC_W), lds_spec, ctx->stream, g, sc, mp,     \
                       ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32,          \
                       cap_spec, SPEC_LCAP, ctx->ppmax, ctx->d_nonempty, IMG, MC, chain_tab(ctx), ctx->d_stats,    \
                       ctx->d_dsum, ctx->d_flag)
            // four waves per cell in every mode (the network screen splits its
            // hidden units, and so its MFMA operands, between the four waves)
            int nw = ctx->sweep_nw ? ctx->sweep_nw : 4;
            if (spec) {
                if (mode == DA_NET) SPEC_PASS(DA_NET, ctx->mlp_img, mlpc);
                else if (mode == DA_REP) SPEC_PASS(DA_REP, nullptr, ctx->mlp_c);
                else SPEC_PASS(DA_OFF, nullptr, ctx->mlp_c);
            } else if (mode == DA_NET) {
                if (nw == 4) SWEEP_PASS(DA_NET, 4, ctx->mlp_img, mlpc);
                else if (nw == 2) SWEEP_PASS(DA_NET, 2, ctx->mlp_img, mlpc);
                else SWEEP_PASS(DA_NET, 1, ctx->mlp_img, mlpc);
            } else if (mode == DA_REP) {
                if (nw == 4) SWEEP_PASS(DA_REP, 4, nullptr, ctx->mlp_c);
                else if (nw == 2) SWEEP_PASS(DA_REP, 2, nullptr, ctx->mlp_c);
                else SWEEP_PASS(DA_REP, 1, nullptr, ctx->mlp_c);
            } else {
                if (nw == 4) SWEEP_PASS(DA_OFF, 4, nullptr, ctx->mlp_c);
                else if (nw == 2) SWEEP_PASS(DA_OFF, 2, nullptr, ctx->mlp_c);
                else SWEEP_PASS(DA_OFF, 1, nullptr, ctx->mlp_c);
            }
#undef SWEEP_PASS
#undef SPEC_PASS
#undef SPEC_PASS_LT
            if (ctx->profile) HIPCK(hipEventRecord(ctx->kev[2 * launches + 1], ctx->stream));
            launches++;
        }
        HIPCK(hipGetLastError());
        if ((rc = chain_passes(ctx, seed, sw, temper, r, shift))) return rc;
    }
    HIPCK(hipEventRecord(ctx->ev1, ctx->stream));
    unsigned long long hs[STAT_STRIDE];
    double ds = 0;
    std::vector<unsigned long long> hrep(STAT_REP * STAT_STRIDE);
    std::vector<double> drep(STAT_REP * STAT_STRIDE);
    HIPCK(hipMemcpyAsync(hrep.data(), ctx->d_stats, hrep.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         ctx->stream));
    HIPCK(hipMemcpyAsync(drep.data(), ctx->d_dsum, drep.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    int rc = check_flag(ctx, "sweep");
    if (rc) return rc;
    for (int j = 0; j < STAT_STRIDE; j++) {
        hs[j] = 0;
        for (int q = 0; q < STAT_REP; q++) hs[j] += hrep[q * STAT_STRIDE + j];
    }
    for (int q = 0; q < STAT_REP; q++) ds += drep[q * STAT_STRIDE];
    if (carried) {
        ctx->e_track += ds;
        ctx->e_valid = true;
    }
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    ctx->last_launches = launches;
    ctx->last_kernel_ms = 0;
    if (ctx->profile)
        for (long k = 0; k < launches; k++) {
            float km = 0;
            HIPCK(hipEventElapsedTime(&km, ctx->kev[2 * k], ctx->kev[2 * k + 1]));
            ctx->last_kernel_ms += km;
        }
    if (stats) {
        stats->trans_acc = hs[0];
        stats->trans_try = hs[1];
        stats->rot_acc = hs[2];
        stats->rot_try = hs[3];
        stats->left_cell = hs[4];
        stats->stage1_acc = hs[5];
        stats->stage1_try = hs[6];
        stats->stage2_acc = hs[7];
        stats->stage2_try = hs[8];
        stats->energy_change = ds;
        stats->bond_rejected = hs[9];
        stats->chain_disp_acc = hs[10];
        stats->chain_disp_try = hs[11];
        stats->chain_rot_acc = hs[12];
        stats->chain_rot_try = hs[13];
        stats->chain_rejected = hs[14];
        stats->chain_cells_skipped = hs[15];
    }
    return SC_OK;
}

int sc_sweep(sc_ctx* ctx, int nsweeps, double temper, double transmx, double rot_half, int r, uint64_t seed,
             uint64_t counter, sc_sweep_stats* stats) {
    SC_GUARD(ctx, sc_sweep_impl(ctx, nsweeps, temper, transmx, rot_half, r, seed, counter, stats))
}

// ------------------------------------------------------------ replica batch
// A batch is the loaded configuration cut into replicas of their own size and
// box (particles back to back, positions as fractions of their own box): the
// parallel-tempering replicas of MoveCreator::replicaExchangeMove
// (movecreator.cpp:555-760), which the reference runs as one MPI process
// each.  Every replica has its own cell grid; a checkerboard pass of all
// replicas is one launch.

// RNG key of replica r under a sweep seed (a replica swept on its own with
// this seed follows the same trajectory)
static uint64_t replica_seed(uint64_t seed, int r) { return mix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(r + 1)); }

static bool batch_active(const sc_ctx* ctx) { return ctx->nrep > 0; }

static int batch_alloc(sc_ctx* ctx) {
    int R = ctx->nrep, rc;
    if (R > ctx->rep_alloc) {
        if ((rc = dalloc(ctx, &ctx->d_rg, R)) || (rc = dalloc(ctx, &ctx->d_rep_nonempty, R)) ||
            (rc = dalloc(ctx, &ctx->d_rep_stats, (size_t)R * STAT_REP * STAT_STRIDE)) ||
            (rc = dalloc(ctx, &ctx->d_rep_dsum, (size_t)R * STAT_REP * STAT_STRIDE)) ||
            (rc = dalloc(ctx, &ctx->d_rep_e, R)))
            return rc;
        ctx->rep_alloc = R;
    }
    return SC_OK;
}

// cell lists of every replica: grid of replica r shifted by shifts[r], cells
// at least maxcut / r wide, a multiple of r + 1 per axis (or one cell)
static int batch_build_grid(sc_ctx* ctx, int r, const std::vector<std::array<uint32_t, 3>>* shifts) {
    int R = ctx->nrep, p = r + 1, rc;
    double wmin = ctx->maxcut / r;
    ctx->rep_grid.resize(R);
    ctx->rep_ncell.resize(R);
    int total = 0;
    for (int q = 0; q < R; q++) {
        Grid g;
        g.r = r;
        for (int k = 0; k < 3; k++) {
            int nc = (int)floor(ctx->rep_box[q][k] / wmin);
            if (nc > 1024) nc = 1024;
            if (nc >= p) nc = (nc / p) * p;
            else nc = 1;
            g.nc[k] = nc;
            g.sh[k] = shifts ? (*shifts)[q][k] : 0u;
        }
        ctx->rep_grid[q].g = g;
        ctx->rep_grid[q].cell_off = total;
        ctx->rep_grid[q].part_off = ctx->rep_off[q];
        ctx->rep_ncell[q] = g.nc[0] * g.nc[1] * g.nc[2];
        total += ctx->rep_ncell[q];
    }
    ctx->rep_ncell_total = total;
    if (total > ctx->ncell_alloc) {
        if ((rc = dalloc(ctx, &ctx->starts, total + 1)) || (rc = dalloc(ctx, &ctx->counts, total)) ||
            (rc = dalloc(ctx, &ctx->fill, total)) || (rc = dalloc(ctx, &ctx->partial, total)))
            return rc;
        ctx->ncell_alloc = total;
    }
    if ((rc = batch_alloc(ctx))) return rc;
    int n = ctx->n, c = ctx->cur, nb = (n + 255) / 256;
    if ((rc = pin_upload(ctx, ctx->pin_rg, ctx->d_rg, ctx->rep_grid.data(), R * sizeof(RepGrid)))) return rc;
    HIPCK(hipMemsetAsync(ctx->counts, 0, sizeof(int) * total, ctx->stream));
    HIPCK(hipMemsetAsync(ctx->fill, 0, sizeof(int) * total, ctx->stream));
    SC_LAUNCH(k_cell_count_batch, dim3(nb), dim3(256), 0, ctx->stream, n, R, ctx->d_rg, ctx->P[c], ctx->ID[c],
              ctx->cell_tmp, ctx->counts);
    SC_LAUNCH(k_scan, dim3(1), dim3(1024), 0, ctx->stream, total, ctx->counts, ctx->starts, ctx->d_maxocc,
              ctx->d_nonempty);
    SC_LAUNCH(k_cell_scatter, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->cell_tmp, ctx->starts, ctx->fill,
              ctx->src_of_pos);
    SC_LAUNCH(k_cell_sort, dim3((total + CELL_SORT_CELLS - 1) / CELL_SORT_CELLS), dim3(WAVE * CELL_SORT_CELLS), 0,
              ctx->stream, total, ctx->starts, ctx->src_of_pos, ctx->ID[c]);
    SC_LAUNCH(k_permute, dim3(nb), dim3(256), 0, ctx->stream, n, ctx->src_of_pos, ctx->P[c], ctx->O[c], ctx->Q[c],
              ctx->ID[c], ctx->P[c ^ 1], ctx->O[c ^ 1], ctx->Q[c ^ 1], ctx->ID[c ^ 1], ctx->slot);
    SC_LAUNCH(k_nonempty_batch, dim3(R), dim3(256), 0, ctx->stream, ctx->d_rg, R, total, ctx->starts,
              ctx->d_rep_nonempty);
    HIPCK(hipGetLastError());
    ctx->cur ^= 1;
    ctx->grid_valid = false;   // the single-system grid no longer describes the arrays
    return SC_OK;
}

static SysConst sysconst_box(const sc_ctx* ctx, const std::array<double, 3>& b) {
    double box[3] = {b[0], b[1], b[2]};
    return sysconst(ctx, box);
}

static int sc_batch_set_impl(sc_ctx* ctx, int nrep, const int32_t* n_per_rep, const double* boxes) {
    if (!ctx || nrep < 1 || !n_per_rep || !boxes) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "load the replicas' particles with sc_set_config first");
    long tot = 0;
    for (int q = 0; q < nrep; q++) {
        if (n_per_rep[q] < 1) return fail(ctx, SC_EINVAL, "every replica needs a particle");
        for (int k = 0; k < 3; k++)
            if (!(boxes[3 * q + k] > 0)) return fail(ctx, SC_EINVAL, "box edges must be positive");
        tot += n_per_rep[q];
    }
    if (tot != ctx->n) return fail(ctx, SC_EINVAL, "replica sizes do not add up to the loaded particles");
    if (ctx->molecules) return fail(ctx, SC_EINVAL, "replica batches hold atomic molecules only");
    ctx->e_valid = false;
    ctx->nrep = nrep;
    ctx->rep_n.assign(n_per_rep, n_per_rep + nrep);
    ctx->rep_off.resize(nrep);
    ctx->rep_box.resize(nrep);
    long off = 0;
    for (int q = 0; q < nrep; q++) {
        ctx->rep_off[q] = (int)off;
        off += n_per_rep[q];
        for (int k = 0; k < 3; k++) ctx->rep_box[q][k] = boxes[3 * q + k];
    }
    HIPCK(hipSetDevice(ctx->device));
    return batch_build_grid(ctx, ctx->tuned_r, nullptr);
}

// nsweeps sweeps of every replica at its temperature temper[r]
static int sc_batch_sweep_impl(sc_ctx* ctx, int nsweeps, const double* temper, double transmx, double rot_half, int r,
                               uint64_t seed, uint64_t counter, sc_sweep_stats* stats) {
    if (!ctx || nsweeps < 0 || !temper || r < 0 || r > 2) return SC_EINVAL;
    if (!batch_active(ctx)) return fail(ctx, SC_ESTATE, "no replica batch (sc_batch_set)");
    if (r == 0) r = ctx->tuned_r;
    int R = ctx->nrep;
    for (int q = 0; q < R; q++)
        if (!(temper[q] > 0)) return fail(ctx, SC_EINVAL, "temperatures must be positive");
    int mode = ctx->da_mode;
    if (mode == DA_NET) return fail(ctx, SC_EINVAL, "the network screen runs on single systems only");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = batch_alloc(ctx))) return rc;
    HIPCK(hipMemsetAsync(ctx->d_rep_stats, 0, (size_t)R * STAT_REP * STAT_STRIDE * sizeof(unsigned long long),
                         ctx->stream));
    HIPCK(hipMemsetAsync(ctx->d_rep_dsum, 0, (size_t)R * STAT_REP * STAT_STRIDE * sizeof(double), ctx->stream));
    int p = r + 1, ncol = p * p * p;
    if ((long)ncol * R > ctx->rp_cap) {
        if ((rc = dalloc(ctx, &ctx->d_rp, (size_t)ncol * R))) return rc;
        ctx->rp_cap = ncol * R;
    }
    std::vector<RepPass> hp((size_t)ncol * R);
    std::vector<int> nblk(ncol);
    std::vector<SweepPlan> plans(R);
    std::vector<std::array<uint32_t, 3>> shifts(R);
    std::vector<std::array<int, 27>> vcols(R);
    std::vector<int> nvis(R), vfac(R);
    long launches = 0;
    HIPCK(hipEventRecord(ctx->ev0, ctx->stream));
    for (int sidx = 0; sidx < nsweeps; sidx++) {
        uint64_t sw = counter + (uint64_t)sidx;
        for (int q = 0; q < R; q++) {
            sweep_plan(replica_seed(seed, q), sw, p, plans[q]);
            for (int k = 0; k < 3; k++) shifts[q][k] = plans[q].shift[k];
        }
        if ((rc = batch_build_grid(ctx, r, &shifts))) return rc;
        int cap = 64;
        for (int q = 0; q < R; q++) {
            double b[3] = {ctx->rep_box[q][0], ctx->rep_box[q][1], ctx->rep_box[q][2]};
            cap = std::max(cap, stage_cap_n(ctx->rep_n[q], ctx->minvol, ctx->rep_grid[q].g, b));
        }
        // the speculative variant (block waves 0): the stage beside SPEC_W trials' lists
        bool spec = ctx->spec && ctx->sweep_nw == 0;
        const bool spec_lt = sweep_table_bytes(ctx->nt) > 0;   // the tables fit in LDS beside the stage
        int cap_spec = cap;
        size_t lds_spec = 0;
        if (spec) {
            size_t tabb = sweep_table_bytes(ctx->nt);
            long room = (long)SWEEP_LDS_BUDGET - (long)spec_layout(0, SPEC_LCAP, SPEC_XCAP, false, tabb).end - 16;
            int cmax = room > 0 ? (int)(room / (long)stage_bytes(1)) : 0;
            if (cap_spec > cmax) cap_spec = (cmax / 64) * 64;
            spec = cap_spec >= 64;
            if (spec) lds_spec = spec_layout(cap_spec, SPEC_LCAP, SPEC_XCAP, false, tabb).end;
        }
        size_t fixed = sizeof(float) * 2 * MV_STRIDE + sizeof(int) * (2 * 125 + 2) + SWEEP_SMALL_LDS +
                       sweep_table_bytes(ctx->nt);
        size_t per = stage_bytes(1) + sizeof(int) * 3;
        int capmax = (int)((SWEEP_LDS_BUDGET - fixed) / per);
        if (cap > capmax) cap = (capmax / 64) * 64;
        size_t lds = per * cap + fixed;
        // each replica's colours in its random order (an axis with a single
        // cell has period 1), and its visit factor: the context's where it
        // divides the replica's number of colours, else 1 (as sc_sweep on the
        // replica alone, which refuses the other factors)
        int nlaunch = 0;
        for (int q = 0; q < R; q++) {
            const Grid& g = ctx->rep_grid[q].g;
            int pa[3], nv = 0;
            for (int k = 0; k < 3; k++) pa[k] = (g.nc[k] >= p) ? p : 1;
            for (int ci = 0; ci < ncol; ci++) {
                int col = plans[q].order[ci];
                if (col % p < pa[0] && (col / p) % p < pa[1] && col / (p * p) < pa[2]) vcols[q][nv++] = col;
            }
            vfac[q] = nv % ctx->visit == 0 ? ctx->visit : 1;
            nvis[q] = nv / vfac[q];
            nlaunch = std::max(nlaunch, nvis[q]);
        }
        for (int ci = 0; ci < nlaunch; ci++) {
            int boff = 0;
            for (int q = 0; q < R; q++) {
                RepPass& P_ = hp[(size_t)ci * R + q];
                const Grid& g = ctx->rep_grid[q].g;
                P_.g = g;
                P_.sc = sysconst_box(ctx, ctx->rep_box[q]);
                P_.cell_off = ctx->rep_grid[q].cell_off;
                P_.rep = q;
                P_.blk_off = boff;
                MoveParams& mp = P_.mp;
                bool skip = ci >= nvis[q];   // this replica's sweep is done
                int col = skip ? 0 : vcols[q][ci];
                int pa[3];
                for (int k = 0; k < 3; k++) pa[k] = (g.nc[k] >= p) ? p : 1;
                mp.ox = col % p; mp.oy = (col / p) % p; mp.oz = col / (p * p);
                mp.p = p;
                for (int k = 0; k < 3; k++) mp.nact[k] = g.nc[k] / pa[k];
                mp.trial_mode = ctx->trial_mode;
                mp.hcap = ctx->home_cache;
                mp.visit = vfac[q];
                mp.n = ctx->rep_n[q];
                mp.transmx = (float)transmx;
                mp.rot_half = (float)rot_half;
                mp.temper = temper[q];
                uint64_t rs = replica_seed(seed, q);
                mp.key0 = (uint32_t)rs;
                mp.key1 = (uint32_t)(rs >> 32);
                uint64_t pass = sw * 32 + (uint64_t)ci;
                mp.pass_lo = (uint32_t)pass;
                mp.pass_hi = (uint32_t)(pass >> 32);
                if (!skip) boff += mp.nact[0] * mp.nact[1] * mp.nact[2];
            }
            nblk[ci] = boff;
        }
        if ((rc = pin_upload(ctx, ctx->pin_rp, ctx->d_rp, hp.data(), (size_t)nlaunch * R * sizeof(RepPass)))) return rc;
        int c = ctx->cur;
        for (int ci = 0; ci < nlaunch; ci++) {
            if (nblk[ci] == 0) continue;
            const RepPass* rp = ctx->d_rp + (size_t)ci * R;
            if (ctx->profile) {
                while ((long)ctx->kev.size() < 2 * (launches + 1)) {
                    hipEvent_t e;
                    HIPCK(hipEventCreate(&e));
                    ctx->kev.push_back(e);
                }
                HIPCK(hipEventRecord(ctx->kev[2 * launches], ctx->stream));
            }
#define SWEEP_BATCH(DAM, NWV)                                                                                  \
    SC_LAUNCH((k_sweep_batch<DAM, NWV>), dim3(nblk[ci]), dim3(WAVE * NWV), lds, ctx->stream, rp, R, ctx->starts,     \
              ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32, cap, ctx->ppmax,                \
              ctx->d_rep_nonempty, nullptr, ctx->mlp_c, ChainTab{nullptr, nullptr, ctx->d_pp64}, ctx->d_rep_stats,  \
              ctx->d_rep_dsum, ctx->d_flag)
#define SPEC_BATCH(DAM)                                                                                        \
    if (spec_lt) SPEC_BATCH_LT(DAM, true); else SPEC_BATCH_LT(DAM, false)
#define SPEC_BATCH_LT(DAM, LT)                                                                                 \
    SC_LAUNCH((k_sweep_batch_spec<DAM, LT>), dim3(nblk[ci]), dim3(WAVE * SPEC_W), lds_spec, ctx->stream, rp, R,           \
              ctx->starts, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32, cap_spec, SPEC_LCAP,  \
              ctx->ppmax, ctx->d_rep_nonempty, nullptr, ctx->mlp_c, ChainTab{nullptr, nullptr, ctx->d_pp64},           \
              ctx->d_rep_stats, ctx->d_rep_dsum, ctx->d_flag)
            int nw = ctx->sweep_nw ? ctx->sweep_nw : 4;
            if (spec) {
                if (mode == DA_REP) SPEC_BATCH(DA_REP);
                else SPEC_BATCH(DA_OFF);
            } else if (mode == DA_REP) {
                if (nw == 4) SWEEP_BATCH(DA_REP, 4);
                else if (nw == 2) SWEEP_BATCH(DA_REP, 2);
                else SWEEP_BATCH(DA_REP, 1);
            } else {
                if (nw == 4) SWEEP_BATCH(DA_OFF, 4);
                else if (nw == 2) SWEEP_BATCH(DA_OFF, 2);
                else SWEEP_BATCH(DA_OFF, 1);
            }
#undef SWEEP_BATCH
#undef SPEC_BATCH
#undef SPEC_BATCH_LT
            if (ctx->profile) HIPCK(hipEventRecord(ctx->kev[2 * launches + 1], ctx->stream));
            launches++;
        }
        HIPCK(hipGetLastError());
    }
    HIPCK(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<unsigned long long> hrep((size_t)R * STAT_REP * STAT_STRIDE);
    std::vector<double> drep((size_t)R * STAT_REP * STAT_STRIDE);
    HIPCK(hipMemcpyAsync(hrep.data(), ctx->d_rep_stats, hrep.size() * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipMemcpyAsync(drep.data(), ctx->d_rep_dsum, drep.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
    if ((rc = check_flag(ctx, "batch sweep"))) return rc;
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    ctx->last_launches = launches;
    ctx->last_kernel_ms = 0;
    if (ctx->profile)
        for (long q = 0; q < launches; q++) {
            float km = 0;
            HIPCK(hipEventElapsedTime(&km, ctx->kev[2 * q], ctx->kev[2 * q + 1]));
            ctx->last_kernel_ms += km;
        }
    if (stats)
        for (int q = 0; q < R; q++) {
            unsigned long long hs[STAT_STRIDE] = {0};
            double ds = 0;
            for (int j = 0; j < STAT_STRIDE; j++)
                for (int t = 0; t < STAT_REP; t++) hs[j] += hrep[((size_t)q * STAT_REP + t) * STAT_STRIDE + j];
            for (int t = 0; t < STAT_REP; t++) ds += drep[((size_t)q * STAT_REP + t) * STAT_STRIDE];
            sc_sweep_stats& st = stats[q];
            st.trans_acc = hs[0]; st.trans_try = hs[1]; st.rot_acc = hs[2]; st.rot_try = hs[3]; st.left_cell = hs[4];
            st.stage1_acc = hs[5]; st.stage1_try = hs[6]; st.stage2_acc = hs[7]; st.stage2_try = hs[8];
            st.energy_change = ds;
            st.bond_rejected = hs[9];
        }
    return SC_OK;
}

// whole-system energy of every replica (TotalEMatrix::allToAll per replica,
// totalenergycalculator.h:338-353): one launch per replica over its cells
static int sc_batch_energies_impl(sc_ctx* ctx, double* e) {
    if (!ctx || !e) return SC_EINVAL;
    if (!batch_active(ctx)) return fail(ctx, SC_ESTATE, "no replica batch (sc_batch_set)");
    HIPCK(hipSetDevice(ctx->device));
    int R = ctx->nrep, rc;
    if ((rc = batch_alloc(ctx))) return rc;
    int c = ctx->cur;
    for (int q = 0; q < R; q++) {
        const Grid& g = ctx->rep_grid[q].g;
        double b[3] = {ctx->rep_box[q][0], ctx->rep_box[q][1], ctx->rep_box[q][2]};
        int cap = stage_cap_n(ctx->rep_n[q], ctx->minvol, g, b, ENERGY_CAP_MAX);
        size_t lds = energy_lds_bytes(cap);
        int co = ctx->rep_grid[q].cell_off, nc = ctx->rep_ncell[q];
        SysConst sc = sysconst(ctx, b);
        SC_LAUNCH(k_energy_all<ENERGY_NW>, dim3(nc), dim3(WAVE * ENERGY_NW), lds, ctx->stream, g, sc, ctx->starts + co, ctx->P[c], ctx->O[c],
                  ctx->Q[c], ctx->ID[c], ctx->d_pp32, ctx->d_tp32, cap, ctx->partial + co, ctx->d_flag);
        SC_LAUNCH(k_reduce, dim3(1), dim3(1024), 0, ctx->stream, nc, ctx->partial + co, ctx->d_rep_e + q);
    }
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(e, ctx->d_rep_e, R * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return check_flag(ctx, "batch energies");
}

// One replica-exchange step (MoveCreator::replicaExchangeMove,
// movecreator.cpp:555-760): replicas ordered by temperature; neighbours
// (k, k+1) with k % 2 == parity attempt to swap temperatures, accepted if
// change = (1/T_k - 1/T_k+1) (E_k - E_k+1) > 0 or U < exp(change) (the
// canonical criterion of :695-713).  temper (per replica) is updated in
// place; accepted[k] = 1 for the swapped pairs (by rank k in the ladder).
static int sc_batch_exchange_impl(sc_ctx* ctx, double* temper, const double* energy, int parity, uint64_t seed,
                                  uint64_t counter, int32_t* accepted) {
    if (!ctx || !temper || !energy) return SC_EINVAL;
    if (!batch_active(ctx)) return fail(ctx, SC_ESTATE, "no replica batch (sc_batch_set)");
    int R = ctx->nrep;
    std::vector<int> rank(R);
    for (int q = 0; q < R; q++) rank[q] = q;
    std::stable_sort(rank.begin(), rank.end(), [&](int a, int b) { return temper[a] < temper[b]; });
    uint64_t h = mix64(seed ^ mix64(counter * 0xD6E8FEB86659FD93ull + 3));
    for (int k = 0; k + 1 < R; k++) {
        if (accepted) accepted[k] = 0;
        if ((k & 1) != (parity & 1)) continue;
        int a = rank[k], b = rank[k + 1];
        h = mix64(h);
        double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
        double change = (1.0 / temper[a] - 1.0 / temper[b]) * (energy[a] - energy[b]);
        if (change > 0 || u < exp(change)) {
            std::swap(temper[a], temper[b]);
            if (accepted) accepted[k] = 1;
        }
    }
    return SC_OK;
}

// the configuration of replica r (positions as fractions of its own box)
static int sc_batch_get_config_impl(sc_ctx* ctx, int r, double* pos, double* dir, double* pd0) {
    if (!ctx || !pos || !dir || !pd0) return SC_EINVAL;
    if (!batch_active(ctx) || r < 0 || r >= ctx->nrep) return fail(ctx, SC_EINVAL, "no such replica");
    int n = ctx->n;
    std::vector<double> P(3 * (size_t)n), D(3 * (size_t)n), Qv(3 * (size_t)n);
    double box[3];
    int rc = sc_get_config_impl(ctx, box, P.data(), D.data(), Qv.data());
    if (rc) return rc;
    size_t o = (size_t)ctx->rep_off[r] * 3, m = (size_t)ctx->rep_n[r] * 3;
    memcpy(pos, P.data() + o, m * sizeof(double));
    memcpy(dir, D.data() + o, m * sizeof(double));
    memcpy(pd0, Qv.data() + o, m * sizeof(double));
    return SC_OK;
}

int sc_batch_set(sc_ctx* ctx, int nrep, const int32_t* n_per_rep, const double* boxes) {
    SC_GUARD(ctx, sc_batch_set_impl(ctx, nrep, n_per_rep, boxes))
}
int sc_batch_sweep(sc_ctx* ctx, int nsweeps, const double* temper, double transmx, double rot_half, int r,
                   uint64_t seed, uint64_t counter, sc_sweep_stats* stats) {
    SC_GUARD(ctx, sc_batch_sweep_impl(ctx, nsweeps, temper, transmx, rot_half, r, seed, counter, stats))
}
int sc_batch_energies(sc_ctx* ctx, double* e) { SC_GUARD(ctx, sc_batch_energies_impl(ctx, e)) }
int sc_batch_exchange(sc_ctx* ctx, double* temper, const double* energy, int parity, uint64_t seed, uint64_t counter,
                      int32_t* accepted) {
    SC_GUARD(ctx, sc_batch_exchange_impl(ctx, temper, energy, parity, seed, counter, accepted))
}
int sc_batch_get_config(sc_ctx* ctx, int r, double* pos, double* dir, double* pd0) {
    SC_GUARD(ctx, sc_batch_get_config_impl(ctx, r, pos, dir, pd0))
}
uint64_t sc_replica_seed(uint64_t seed, int r) { return replica_seed(seed, r); }

// ------------------------------------------------------------ autotuner
// Times the sweep in each launch shape on the loaded state (restored after
// every trial) and keeps the fastest: stencil radius r = 1 (cells >= maxcut,
// 2x2x2 colours) or 2 (cells >= maxcut/2, 3x3x3 colours), crossed with the
// trial count per active cell (its occupancy, or the dithered mean).  Works
// on a single system and on a replica batch (temper: one temperature, or one
// per replica).
static int sc_autotune_impl(sc_ctx* ctx, int sweeps, const double* temper, double transmx, double rot_half,
                            uint64_t seed, double* rates, int* r_best, int* mode_best) {
    if (!ctx || sweeps < 1 || !temper) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    HIPCK(hipSetDevice(ctx->device));
    int n = ctx->n, c0 = ctx->cur, rc;
    uint4* sp = nullptr; float4* so = nullptr; float2* sq = nullptr; int* si = nullptr; int* ss = nullptr;
    HIPCK(hipMalloc((void**)&sp, n * sizeof(uint4)));
    HIPCK(hipMalloc((void**)&so, n * sizeof(float4)));
    HIPCK(hipMalloc((void**)&sq, n * sizeof(float2)));
    HIPCK(hipMalloc((void**)&si, n * sizeof(int)));
    HIPCK(hipMalloc((void**)&ss, n * sizeof(int)));
    auto save = [&](bool back) -> int {
        if (!back) ctx->cur = c0;
        uint4* P_ = ctx->P[c0]; float4* O_ = ctx->O[c0]; float2* Q_ = ctx->Q[c0]; int* I_ = ctx->ID[c0];
        hipMemcpyKind k = hipMemcpyDeviceToDevice;
        HIPCK(hipMemcpyAsync(back ? sp : P_, back ? P_ : sp, n * sizeof(uint4), k, ctx->stream));
        HIPCK(hipMemcpyAsync(back ? so : O_, back ? O_ : so, n * sizeof(float4), k, ctx->stream));
        HIPCK(hipMemcpyAsync(back ? sq : Q_, back ? Q_ : sq, n * sizeof(float2), k, ctx->stream));
        HIPCK(hipMemcpyAsync(back ? si : I_, back ? I_ : si, n * sizeof(int), k, ctx->stream));
        HIPCK(hipMemcpyAsync(back ? ss : ctx->slot, back ? ctx->slot : ss, n * sizeof(int), k, ctx->stream));
        HIPCK(hipStreamSynchronize(ctx->stream));
        return SC_OK;
    };
    int saved_mode = ctx->trial_mode, best_r = ctx->tuned_r, best_m = saved_mode;
    double best = -1.0;
    rc = save(true);
    for (int cand = 0; cand < 4 && !rc; cand++) {
        int r = 1 + cand / 2, m = cand % 2;
        ctx->trial_mode = m;
        double tries = 0;
        float ms = 0;
        if (batch_active(ctx)) {
            std::vector<sc_sweep_stats> st(ctx->nrep);
            if ((rc = sc_batch_sweep_impl(ctx, 1, temper, transmx, rot_half, r, seed, 1, st.data()))) break;
            if ((rc = sc_batch_sweep_impl(ctx, sweeps, temper, transmx, rot_half, r, seed, 2, st.data()))) break;
            for (auto& x : st) tries += (double)(x.trans_try + x.rot_try);
        } else {
            sc_sweep_stats st;
            if ((rc = sc_sweep_impl(ctx, 1, temper[0], transmx, rot_half, r, seed, 1, &st))) break;
            if ((rc = sc_sweep_impl(ctx, sweeps, temper[0], transmx, rot_half, r, seed, 2, &st))) break;
            tries = (double)(st.trans_try + st.rot_try);
        }
        ms = (float)ctx->last_ms;
        double rate = tries / (double)n / (ms * 1e-3);
        if (rates) rates[cand] = rate;
        if (rate > best) { best = rate; best_r = r; best_m = m; }
        if ((rc = save(false))) break;
    }
    hipFree(sp); hipFree(so); hipFree(sq); hipFree(si); hipFree(ss);
    ctx->trial_mode = saved_mode;
    if (rc) return rc;
    ctx->tuned_r = best_r;
    ctx->trial_mode = best_m;
    if (r_best) *r_best = best_r;
    if (mode_best) *mode_best = best_m;
    if (batch_active(ctx)) return batch_build_grid(ctx, best_r, nullptr);
    return build_grid(ctx, best_r, nullptr, ctx->box);
}

int sc_autotune(sc_ctx* ctx, int sweeps, const double* temper, double transmx, double rot_half, uint64_t seed,
                double* rates, int* r_best, int* mode_best) {
    SC_GUARD(ctx, sc_autotune_impl(ctx, sweeps, temper, transmx, rot_half, seed, rates, r_best, mode_best))
}

static int sc_volume_move_impl(sc_ctx* ctx, int ptype, double press, double temper, double edge_mx, uint64_t seed,
                   uint64_t counter, int* accepted, double box_out[3]) {
    if (!ctx || !accepted || !(temper > 0)) return SC_EINVAL;
    if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
    if (ctx->nrep > 0) return fail(ctx, SC_ESTATE, "volume moves act on single systems, not replica batches");
    if (ptype != 0 && ptype != 1) return fail(ctx, SC_EINVAL, "volume move: ptype 0 and 1 are supported");
    // MoveCreator::pressureMove (movecreator.cpp:330-421)
    uint64_t h = mix64(seed ^ mix64(counter * 0xD1B54A32D192ED03ull + 7));
    auto u01 = [&]() { h = mix64(h); return (double)(h >> 11) * (1.0 / 9007199254740992.0); };
    double e_old = 0, e_new = 0;
    int rc = SC_OK;
    // E_old: the energy carried through the sweeps (as the reference's
    // allToAll sums its cached pair energies), re-evaluated in full every 16
    // volume moves and whenever it is not known
    if (ctx->e_valid && ctx->e_age < 16) {
        e_old = ctx->e_track;
        ctx->e_age++;
    } else {
        if ((rc = energy_all_impl(ctx, ctx->box, &e_old))) return rc;
        ctx->e_track = e_old;
        ctx->e_valid = true;
        ctx->e_age = 0;
    }
    double nb[3] = {ctx->box[0], ctx->box[1], ctx->box[2]};
    double enermove;
    double npart = (double)ctx->n;
    if (ptype == 0) {
        double rs = u01();
        int side = rs < 1.0 / 3.0 ? 0 : (rs < 2.0 / 3.0 ? 1 : 2);
        double area = nb[(side + 1) % 3] * nb[(side + 2) % 3];
        double old = nb[side];
        nb[side] += edge_mx * (u01() - 0.5);
        if (nb[side] <= 0.0) { *accepted = 0; if (box_out) for (int k = 0; k < 3; k++) box_out[k] = ctx->box[k]; return SC_OK; }
        enermove = press * area * (nb[side] - old) - npart * temper * log(nb[side] / old);
    } else {
        double psch = edge_mx * (u01() - 0.5);
        double pvol = nb[0] * nb[1] * nb[2];
        for (int k = 0; k < 3; k++) nb[k] += psch;
        double pvoln = nb[0] * nb[1] * nb[2];
        enermove = press * (pvoln - pvol) - npart * temper * log(pvoln / pvol);
    }
    for (int k = 0; k < 3; k++)
        if (!(nb[k] > ctx->maxcut)) { *accepted = 0; if (box_out) for (int q = 0; q < 3; q++) box_out[q] = ctx->box[q]; return SC_OK; }
    if ((rc = energy_all_impl(ctx, nb, &e_new))) return rc;
    enermove += e_new;
    bool acc = (enermove <= e_old) || (exp(-1.0 * (enermove - e_old) / temper) > u01());
    *accepted = acc ? 1 : 0;
    if (acc) {
        for (int k = 0; k < 3; k++) ctx->box[k] = nb[k];
        ctx->e_track = e_new;
        ctx->e_age = 0;
    }
    if ((rc = energy_grid(ctx, ctx->box))) return rc;
    if (box_out) for (int k = 0; k < 3; k++) box_out[k] = ctx->box[k];
    return SC_OK;
}

static int sc_carried_energy_impl(sc_ctx* ctx, double* e, int* age) {
    if (!ctx || !e) return SC_EINVAL;
    if (!ctx->e_valid) return fail(ctx, SC_ESTATE, "no carried energy: no volume move since the state changed");
    *e = ctx->e_track;
    if (age) *age = ctx->e_age;
    return SC_OK;
}

int sc_carried_energy(sc_ctx* ctx, double* e, int* age) { SC_GUARD(ctx, sc_carried_energy_impl(ctx, e, age)) }

int sc_volume_move(sc_ctx* ctx, int ptype, double press, double temper, double edge_mx, uint64_t seed,
                   uint64_t counter, int* accepted, double box_out[3]) {
    SC_GUARD(ctx, sc_volume_move_impl(ctx, ptype, press, temper, edge_mx, seed, counter, accepted, box_out))
}

static int sc_copy_probe_impl(sc_ctx* ctx, long bytes, int reps) {
    if (!ctx || bytes <= 0 || reps < 1) return SC_EINVAL;
    if (ctx->device < 0) return fail(ctx, SC_ENODEV, "host-only context");
    HIPCK(hipSetDevice(ctx->device));
    long n16 = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    HIPCK(hipMalloc((void**)&a, n16 * 16));
    if (hipMalloc((void**)&b, n16 * 16) != hipSuccess) { hipFree(a); return fail(ctx, SC_EHIP, "copy probe alloc"); }
    hipError_t e = hipMemsetAsync(a, 1, n16 * 16, ctx->stream);
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, ctx->stream, n16, a, b);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
    return SC_OK;
}

int sc_copy_probe(sc_ctx* ctx, long bytes, int reps) {
    SC_GUARD(ctx, sc_copy_probe_impl(ctx, bytes, reps))
}

}  // extern "C"

// ------------------------------------------------------------ surrogate ABI
static_assert(MLP_NPARAM == SC_MLP_NPARAM && MLP_NF == SC_MLP_NF, "surrogate sizes of the public header");
static MlpConst mlp_const_for(const schost::Ia& p) {
    MlpConst c;
    c.lsc = (float)(p.half_len[0] + p.half_len[1] + p.rcut);
    c.rcut = (float)p.rcut;
    c.pdis = (float)p.pdis;
    c.pswitch = (float)p.pswitch;
    c.tnet = -1;
    c.minx = 0;
    return c;
}

static int mlp_alloc(sc_ctx* ctx) {
    if (ctx->mlp_w) return SC_OK;
    int rc;
    if ((rc = dalloc(ctx, &ctx->mlp_w, MLP_NPARAM)) || (rc = dalloc(ctx, &ctx->mlp_g, MLP_NPARAM)) ||
        (rc = dalloc(ctx, &ctx->mlp_m, MLP_NPARAM)) || (rc = dalloc(ctx, &ctx->mlp_v, MLP_NPARAM)) ||
        (rc = dalloc(ctx, &ctx->mlp_img, MLP_IMG_BYTES)) || (rc = dalloc(ctx, &ctx->d_loss, 1)))
        return rc;
    return SC_OK;
}

static int mlp_repack(sc_ctx* ctx) {
    SC_LAUNCH(k_mlp_pack, dim3(1), dim3(256), 0, ctx->stream, ctx->mlp_w, ctx->mlp_img);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(ctx->stream));
    return SC_OK;
}

static int sc_mlp_init_impl(sc_ctx* ctx, int type, uint64_t seed) {
    if (!ctx) return SC_EINVAL;
    if (ctx->device < 0) return fail(ctx, SC_ENODEV, "host-only context");
    if (!ctx->have_topo || type < 0 || type >= schost::MAXT || ctx->topo.ia[type][type].geotype[0] == 0)
        return fail(ctx, SC_EINVAL, "sc_mlp_init: type not defined in the topology");
    {
        PP32 hp;
        fill_pair(hp, ctx->topo.ia[type][type], ctx->topo.ia[type][type]);
        if (hp.kind != K_PSC && hp.kind != K_CPSC)
            return fail(ctx, SC_EINVAL, "sc_mlp_init: the surrogate models patchy spherocylinders (PSC/CPSC families)");
    }
    HIPCK(hipSetDevice(ctx->device));
    int rc = mlp_alloc(ctx);
    if (rc) return rc;
    // Kaiming-uniform weights (bound sqrt(6/fan_in) scaled for SiLU), zero biases
    std::vector<float> w(MLP_NPARAM, 0.0f);
    uint64_t h = mix64(seed + 99);
    auto u = [&]() { h = mix64(h); return (double)(h >> 11) * (1.0 / 9007199254740992.0); };
    auto fill = [&](int off, int fan_in, int count) {
        double b = sqrt(6.0 / fan_in);
        for (int i = 0; i < count; i++) w[off + i] = (float)((2.0 * u() - 1.0) * b);
    };
    fill(MLP_OFF_W1, MLP_NF, MLP_NF * MLP_NH);
    fill(MLP_OFF_W2, MLP_NH, MLP_NH * MLP_NH);
    fill(MLP_OFF_W3, MLP_NH, MLP_NH * MLP_NH);
    fill(MLP_OFF_W4, MLP_NH, MLP_NH);
    for (int i = 0; i < MLP_NH; i++) w[MLP_OFF_W4 + i] *= 0.1f;
    HIPCK(hipMemcpy(ctx->mlp_w, w.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemset(ctx->mlp_m, 0, MLP_NPARAM * sizeof(float)));
    HIPCK(hipMemset(ctx->mlp_v, 0, MLP_NPARAM * sizeof(float)));
    ctx->adam_t = 0;
    ctx->mlp_type = type;
    ctx->mlp_c = mlp_const_for(ctx->topo.ia[type][type]);
    ctx->mlp_has = true;
    return mlp_repack(ctx);
}

extern "C" int sc_mlp_init(sc_ctx* ctx, int type, uint64_t seed) {
    SC_GUARD(ctx, sc_mlp_init_impl(ctx, type, seed))
}

static int sc_mlp_set_weights_impl(sc_ctx* ctx, const float* w, int n) {
    if (!ctx || !w || n != MLP_NPARAM) return SC_EINVAL;
    if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "call sc_mlp_init first");
    HIPCK(hipMemcpy(ctx->mlp_w, w, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
    return mlp_repack(ctx);
}

extern "C" int sc_mlp_set_weights(sc_ctx* ctx, const float* w, int n) {
    SC_GUARD(ctx, sc_mlp_set_weights_impl(ctx, w, n))
}

static int sc_mlp_get_weights_impl(sc_ctx* ctx, float* w, int n) {
    if (!ctx || !w || n != MLP_NPARAM) return SC_EINVAL;
    if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "call sc_mlp_init first");
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipMemcpy(w, ctx->mlp_w, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return SC_OK;
}

extern "C" int sc_mlp_get_weights(sc_ctx* ctx, float* w, int n) {
    SC_GUARD(ctx, sc_mlp_get_weights_impl(ctx, w, n))
}

static int single_pair_tables(sc_ctx* ctx, int type, PP32** pp, TP32** tp) {
    PP32 hp;
    TP32 ht;
    fill_pair(hp, ctx->topo.ia[type][type], ctx->topo.ia[type][type]);
    fill_type(ht, ctx->topo.ia[type][type]);
    HIPCK(hipMalloc((void**)pp, sizeof(PP32)));
    HIPCK(hipMalloc((void**)tp, sizeof(TP32)));
    HIPCK(hipMemcpy(*pp, &hp, sizeof(PP32), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(*tp, &ht, sizeof(TP32), hipMemcpyHostToDevice));
    return SC_OK;
}

static int sc_mlp_generate_impl(sc_ctx* ctx, long ns, uint64_t seed, double dmax, double near_frac) {
    if (!ctx || ns <= 0 || !(dmax > 0)) return SC_EINVAL;
    if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "call sc_mlp_init first");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if (ns != ctx->tr_n) {
        if ((rc = dalloc(ctx, &ctx->tr_x, ns * MLP_NF)) || (rc = dalloc(ctx, &ctx->tr_y, ns)) ||
            (rc = dalloc(ctx, &ctx->tr_e, ns)))
            return rc;
        ctx->tr_n = ns;
    }
    PP32* pp = nullptr;
    TP32* tp = nullptr;
    if ((rc = single_pair_tables(ctx, ctx->mlp_type, &pp, &tp))) return rc;
    SC_LAUNCH(k_mlp_gen, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, ns, (uint32_t)seed,
                       (uint32_t)(seed >> 32), pp, tp, tp, ctx->mlp_c, (float)dmax, (float)near_frac, ctx->tr_x,
                       ctx->tr_y, ctx->tr_e);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    hipFree(pp);
    hipFree(tp);
    if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
    return SC_OK;
}

extern "C" int sc_mlp_generate(sc_ctx* ctx, long ns, uint64_t seed, double dmax, double near_frac) {
    SC_GUARD(ctx, sc_mlp_generate_impl(ctx, ns, seed, dmax, near_frac))
}

static int sc_mlp_training_data_impl(sc_ctx* ctx, long first, long count, float* feats, float* target, float* efull) {
    if (!ctx || first < 0 || count < 0 || first + count > ctx->tr_n) return SC_EINVAL;
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (feats) HIPCK(hipMemcpy(feats, ctx->tr_x + first * MLP_NF, count * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
    if (target) HIPCK(hipMemcpy(target, ctx->tr_y + first, count * sizeof(float), hipMemcpyDeviceToHost));
    if (efull) HIPCK(hipMemcpy(efull, ctx->tr_e + first, count * sizeof(float), hipMemcpyDeviceToHost));
    return SC_OK;
}

extern "C" int sc_mlp_training_data(sc_ctx* ctx, long first, long count, float* feats, float* target, float* efull) {
    SC_GUARD(ctx, sc_mlp_training_data_impl(ctx, first, count, feats, target, efull))
}

static int ensure_act(sc_ctx* ctx, int B) {
    int nblk = (B + TR_S - 1) / TR_S, rc;
    if (B > ctx->act_batch) {
        if ((rc = dalloc(ctx, &ctx->bx, (size_t)B * (MLP_NF + 1)))) return rc;
        ctx->act_batch = B;
    }
    if (nblk > ctx->part_blocks) {
        if ((rc = dalloc(ctx, &ctx->d_part, (size_t)nblk * TR_PART))) return rc;
        ctx->part_blocks = nblk;
    }
    return SC_OK;
}

// One step on B samples, rows (a i + b) mod n of X / Y: gradient into mlp_g;
// with lr > 0 also the Adam step on mlp_w; the batch's mean squared error is
// added to *loss (device, may be null)
static int mlp_step(sc_ctx* ctx, int B, long n, long a, long b, const float* X, const float* Y, float lr, float c1,
                    float c2, double* loss) {
    int nblk = (B + TR_S - 1) / TR_S;
    SC_LAUNCH(k_mlp_fwdbwd, dim3(nblk), dim3(WAVE * TR_NW), TRAIN_LDS_BYTES, ctx->stream, B, n, a, b, X, Y,
              ctx->mlp_w, ctx->d_part);
    SC_LAUNCH(k_mlp_reduce_adam, dim3((MLP_NPARAM + 1 + 63) / 64), dim3(WAVE * TR_RNW), 0, ctx->stream, nblk, B,
              ctx->d_part,
              ctx->mlp_g, ctx->mlp_w, ctx->mlp_m, ctx->mlp_v, lr, 0.9f, 0.999f, 1e-8f, c1, c2, loss);
    HIPCK(hipGetLastError());
    return SC_OK;
}

static int sc_mlp_gradients_impl(sc_ctx* ctx, int B, const float* x, const float* y, float* grads, int ngrads,
                                 double* loss) {
    if (!ctx || B <= 0 || !x || !y || !grads || ngrads != MLP_NPARAM) return SC_EINVAL;
    if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "call sc_mlp_init first");
    HIPCK(hipSetDevice(ctx->device));
    int rc = ensure_act(ctx, B);
    if (rc) return rc;
    HIPCK(hipMemcpy(ctx->bx, x, (size_t)B * MLP_NF * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->bx + (size_t)B * MLP_NF, y, (size_t)B * sizeof(float), hipMemcpyHostToDevice));
    HIPCK(hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
    if ((rc = mlp_step(ctx, B, B, 1, 0, ctx->bx, ctx->bx + (size_t)B * MLP_NF, 0.0f, 1.0f, 1.0f, ctx->d_loss)))
        return rc;
    double l = 0;
    HIPCK(hipMemcpyAsync(&l, ctx->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipMemcpyAsync(grads, ctx->mlp_g, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return SC_OK;
}

extern "C" int sc_mlp_gradients(sc_ctx* ctx, int B, const float* x, const float* y, float* grads, int ngrads,
                                double* loss) {
    SC_GUARD(ctx, sc_mlp_gradients_impl(ctx, B, x, y, grads, ngrads, loss))
}

extern "C" int sc_mlp_nparam(void) { return MLP_NPARAM; }

static long gcdl(long a, long b) { while (b) { long t = a % b; a = b; b = t; } return a; }

static int sc_mlp_train_impl(sc_ctx* ctx, int epochs, int batch, double lr, double lr_final, uint64_t seed,
                            double val_frac, double* train_loss, double* val_rmse) {
    if (!ctx || epochs < 1 || batch < 32 || !(lr > 0)) return SC_EINVAL;
    if (!ctx->mlp_has || ctx->tr_n <= 0) return fail(ctx, SC_ESTATE, "generate training data first");
    HIPCK(hipSetDevice(ctx->device));
    long nval = (long)(ctx->tr_n * val_frac), ntr = ctx->tr_n - nval;
    if (ntr < batch) return fail(ctx, SC_EINVAL, "training set smaller than a batch");
    int rc = ensure_act(ctx, batch);
    if (rc) return rc;
    long steps_per_epoch = ntr / batch, total = steps_per_epoch * epochs, step = 0;
    uint64_t h = mix64(seed);
    HIPCK(hipMemsetAsync(ctx->d_loss, 0, sizeof(double), ctx->stream));
    for (int ep = 0; ep < epochs; ep++) {
        h = mix64(h);
        long a = (long)(h % (uint64_t)ntr) | 1;
        while (gcdl(a, ntr) != 1) a += 2;
        h = mix64(h);
        long b0 = (long)(h % (uint64_t)ntr);
        for (long st = 0; st < steps_per_epoch; st++, step++) {
            // batch rows (a i + b) mod ntr: a permutation of the training rows per epoch
            long bs = (long)((b0 + (unsigned long long)a * (unsigned long long)(st * batch)) % (unsigned long long)ntr);
            // cosine decay from lr to lr_final
            double lrs = lr_final + 0.5 * (lr - lr_final) * (1.0 + cos(3.141592653589793 * (double)step / (double)total));
            ctx->adam_t++;
            float c1 = (float)(1.0 - pow(0.9, (double)ctx->adam_t)), c2 = (float)(1.0 - pow(0.999, (double)ctx->adam_t));
            if ((rc = mlp_step(ctx, batch, ntr, a, bs, ctx->tr_x, ctx->tr_y, (float)lrs, c1, c2,
                               ep == epochs - 1 ? ctx->d_loss : nullptr)))
                return rc;
        }
    }
    HIPCK(hipGetLastError());
    if ((rc = mlp_repack(ctx))) return rc;
    double lsum = 0.0;
    HIPCK(hipMemcpyAsync(&lsum, ctx->d_loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
    if (train_loss) *train_loss = steps_per_epoch ? lsum / (double)steps_per_epoch : 0.0;
    if (val_rmse) {
        *val_rmse = 0.0;
        if (nval > 0) {
            std::vector<float> xv((size_t)nval * MLP_NF), yv(nval), ov(nval);
            HIPCK(hipMemcpy(xv.data(), ctx->tr_x + ntr * MLP_NF, xv.size() * sizeof(float), hipMemcpyDeviceToHost));
            HIPCK(hipMemcpy(yv.data(), ctx->tr_y + ntr, nval * sizeof(float), hipMemcpyDeviceToHost));
            float* dx = nullptr;
            float* dout = nullptr;
            HIPCK(hipMalloc((void**)&dx, xv.size() * sizeof(float)));
            HIPCK(hipMalloc((void**)&dout, nval * sizeof(float)));
            HIPCK(hipMemcpy(dx, xv.data(), xv.size() * sizeof(float), hipMemcpyHostToDevice));
            SC_LAUNCH(k_mlp_infer, dim3((unsigned)((nval + MLP_TILE - 1) / MLP_TILE)), dim3(WAVE), MLP_LDS_BYTES, ctx->stream,
                               (int)nval, dx, ctx->mlp_img, ctx->mlp_w, 1, dout);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpy(ov.data(), dout, nval * sizeof(float), hipMemcpyDeviceToHost);
            hipFree(dx);
            hipFree(dout);
            if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
            double ss = 0;
            for (long i = 0; i < nval; i++) ss += (double)(ov[i] - yv[i]) * (ov[i] - yv[i]);
            *val_rmse = sqrt(ss / nval);
        }
    }
    return SC_OK;
}

extern "C" int sc_mlp_train(sc_ctx* ctx, int epochs, int batch, double lr, double lr_final, uint64_t seed,
                            double val_frac, double* train_loss, double* val_rmse) {
    SC_GUARD(ctx, sc_mlp_train_impl(ctx, epochs, batch, lr, lr_final, seed, val_frac, train_loss, val_rmse))
}

static int sc_mlp_forward_impl(sc_ctx* ctx, int m, const float* x, float* out, int bf16) {
    if (!ctx || m <= 0 || !x || !out) return SC_EINVAL;
    if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "call sc_mlp_init first");
    HIPCK(hipSetDevice(ctx->device));
    float* dx = nullptr;
    float* dout = nullptr;
    HIPCK(hipMalloc((void**)&dx, (size_t)m * MLP_NF * sizeof(float)));
    if (hipMalloc((void**)&dout, (size_t)m * sizeof(float)) != hipSuccess) { hipFree(dx); return fail(ctx, SC_EHIP, "alloc"); }
    hipError_t e = hipMemcpy(dx, x, (size_t)m * MLP_NF * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        SC_LAUNCH(k_mlp_infer, dim3((m + MLP_TILE - 1) / MLP_TILE), dim3(WAVE), bf16 ? MLP_LDS_BYTES : 0, ctx->stream, m, dx,
                           ctx->mlp_img, ctx->mlp_w, bf16, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, (size_t)m * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(dx);
    hipFree(dout);
    if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
    return SC_OK;
}

extern "C" int sc_mlp_forward(sc_ctx* ctx, int m, const float* x, float* out, int bf16) {
    SC_GUARD(ctx, sc_mlp_forward_impl(ctx, m, x, out, bf16))
}

// features and exact attraction of pairs (i as part1) of the current state
__global__ void k_pair_features(int m, const int* __restrict__ pi, const int* __restrict__ pj, const int* __restrict__ slot,
                                SysConst sc, const uint4* __restrict__ P, const float4* __restrict__ O,
                                const float2* __restrict__ Q, const PP32* __restrict__ pp, const TP32* __restrict__ tp,
                                MlpConst mc, float* __restrict__ feats, float* __restrict__ eattr) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    int a = slot[pi[k]], b = slot[pj[k]];
    uint4 ua = P[a], ub = P[b];
    DPart<float> pa = load_part<float>(O[a], Q[a], ua.w, tp);
    DPart<float> pb = load_part<float>(O[b], Q[b], ub.w, tp);
    V3<float> r = fx_image_c(ua.x, ua.y, ua.z, ub.x, ub.y, ub.z, sc.Lh, sc.Ll);
    const PP32& P2 = pp[ua.w * sc.nt + ub.w];
    float erep = 0.0f, d2 = 0.0f, ea = 0.0f;
    if (P2.kind <= K_SCA && P2.kind >= 0) {
        bool att = scsc_cheap(P2, pa, pb, r, erep, d2);
        if (att) ea = scsc_attr(P2, pa, pb, r);
    }
    float f[MLP_NF];
    mlp_features(pa.dir, pa.pd[0], mk<float>(O[b].x, O[b].y, O[b].z), mk<float>(O[b].w, Q[b].x, Q[b].y), r, sqrtf(d2),
                 mc.lsc, mc.rcut, mc.pdis, mc.pswitch, f);
    for (int q = 0; q < MLP_NF; q++) feats[(long)k * MLP_NF + q] = f[q];
    eattr[k] = ea;
}

static int sc_mlp_pair_features_impl(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, float* feats, float* eattr) {
    if (!ctx || m <= 0 || !i || !j || !feats || !eattr) return SC_EINVAL;
    if (!ctx->have_state || !ctx->mlp_has) return fail(ctx, SC_ESTATE, "needs a configuration and a network");
    for (int k = 0; k < m; k++)
        if (i[k] < 0 || i[k] >= ctx->n || j[k] < 0 || j[k] >= ctx->n) return fail(ctx, SC_EINVAL, "index out of range");
    HIPCK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_ij(ctx, m))) return rc;
    float *df = nullptr, *de = nullptr;
    HIPCK(hipMalloc((void**)&df, (size_t)m * MLP_NF * sizeof(float)));
    if (hipMalloc((void**)&de, (size_t)m * sizeof(float)) != hipSuccess) { hipFree(df); return fail(ctx, SC_EHIP, "alloc"); }
    hipError_t e = hipMemcpy(ctx->d_i, i, m * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ctx->d_j, j, m * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        SysConst sc = sysconst(ctx, ctx->box);
        int c = ctx->cur;
        SC_LAUNCH(k_pair_features, dim3((m + 127) / 128), dim3(128), 0, ctx->stream, m, ctx->d_i, ctx->d_j,
                           ctx->slot, sc, ctx->P[c], ctx->O[c], ctx->Q[c], ctx->d_pp32, ctx->d_tp32, ctx->mlp_c, df, de);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(feats, df, (size_t)m * MLP_NF * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(eattr, de, (size_t)m * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(df);
    hipFree(de);
    if (e != hipSuccess) return fail(ctx, SC_EHIP, hipGetErrorString(e));
    return SC_OK;
}

extern "C" int sc_mlp_pair_features(sc_ctx* ctx, int m, const int32_t* i, const int32_t* j, float* feats, float* eattr) {
    SC_GUARD(ctx, sc_mlp_pair_features_impl(ctx, m, i, j, feats, eattr))
}

static int sc_set_surrogate_impl(sc_ctx* ctx, int mode) {
    if (!ctx || mode < DA_OFF || mode > DA_REP) return SC_EINVAL;
    if (mode == DA_NET) {
        if (!ctx->mlp_has) return fail(ctx, SC_ESTATE, "no surrogate network");
        if (!ctx->have_state) return fail(ctx, SC_ESTATE, "no configuration loaded");
        if (ctx->mlp_type < 0 || ctx->ref2dense[ctx->mlp_type] < 0)
            return fail(ctx, SC_EINVAL, "the network screen needs the network's particle type in the system");
    }
    ctx->da_mode = mode;
    return SC_OK;
}

extern "C" int sc_set_surrogate(sc_ctx* ctx, int mode) {
    SC_GUARD(ctx, sc_set_surrogate_impl(ctx, mode))
}

extern "C" {

static int sc_set_trial_mode_impl(sc_ctx* ctx, int mode) {
    if (!ctx || mode < 0 || mode > 1) return SC_EINVAL;
    ctx->trial_mode = mode;
    return SC_OK;
}

int sc_set_trial_mode(sc_ctx* ctx, int mode) {
    SC_GUARD(ctx, sc_set_trial_mode_impl(ctx, mode))
}

static int sc_set_block_waves_impl(sc_ctx* ctx, int nw) {
    if (!ctx || (nw != 0 && nw != 1 && nw != 2 && nw != 4)) return SC_EINVAL;
    ctx->sweep_nw = nw;
    return SC_OK;
}

int sc_set_block_waves(sc_ctx* ctx, int nw) {
    SC_GUARD(ctx, sc_set_block_waves_impl(ctx, nw))
}

static int sc_set_speculation_impl(sc_ctx* ctx, int on) {
    if (!ctx || (on != 0 && on != 1)) return SC_EINVAL;
    ctx->spec = on;
    return SC_OK;
}

int sc_set_speculation(sc_ctx* ctx, int on) { SC_GUARD(ctx, sc_set_speculation_impl(ctx, on)) }

static int sc_set_network_min_entries_impl(sc_ctx* ctx, int m) {
    if (!ctx) return SC_EINVAL;
    ctx->net_minx = m < 0 ? -1 : m;
    return SC_OK;
}

int sc_set_network_min_entries(sc_ctx* ctx, int m) { SC_GUARD(ctx, sc_set_network_min_entries_impl(ctx, m)) }

static int sc_set_home_cache_impl(sc_ctx* ctx, int cap) {
    if (!ctx || cap < 0) return SC_EINVAL;
    ctx->home_cache = cap < HCAP ? cap : HCAP;
    return SC_OK;
}

int sc_set_home_cache(sc_ctx* ctx, int cap) {
    SC_GUARD(ctx, sc_set_home_cache_impl(ctx, cap))
}

static int sc_set_max_active_cells_impl(sc_ctx* ctx, int n) {
    if (!ctx) return SC_EINVAL;
    ctx->max_active = n < 0 ? ctx->max_active_default : n;
    return SC_OK;
}

int sc_set_max_active_cells(sc_ctx* ctx, int n) {
    SC_GUARD(ctx, sc_set_max_active_cells_impl(ctx, n))
}

static int sc_set_visit_factor_impl(sc_ctx* ctx, int v) {
    if (!ctx || v < 1 || v > 27) return SC_EINVAL;
    ctx->visit = v;
    return SC_OK;
}

int sc_set_visit_factor(sc_ctx* ctx, int v) {
    SC_GUARD(ctx, sc_set_visit_factor_impl(ctx, v))
}

static int sc_set_profiling_impl(sc_ctx* ctx, int on) {
    if (!ctx) return SC_EINVAL;
    ctx->profile = on != 0;
    return SC_OK;
}

int sc_set_profiling(sc_ctx* ctx, int on) {
    SC_GUARD(ctx, sc_set_profiling_impl(ctx, on))
}

int sc_last_timing(const sc_ctx* ctx, double* ms, long* launches, double* kernel_ms) {
    if (!ctx) return SC_EINVAL;
    if (ms) *ms = ctx->last_ms;
    if (launches) *launches = ctx->last_launches;
    if (kernel_ms) *kernel_ms = ctx->last_kernel_ms;
    return SC_OK;
}

}  // extern "C"

// ------------------------------------------------------- synthetic inputs
// Random sequential addition (host, FP64): particles are inserted one by one
// at uniform positions with uniform orientations (patch direction uniform
// around the axis) and kept if their closest distance to every particle
// already placed is >= sigma_ij (the hard-core criterion of sc_overlap_flags).
// Used to build seeded, overlap-free starting configurations.
static int sc_generate_rsa_impl(sc_ctx* ctx, int n, const int32_t* types, const double box[3], uint64_t seed,
                               long max_attempts, double* pos, double* dir, double* pd0) {
    if (!ctx || n < 0 || !types || !box || !pos || !dir || !pd0) return SC_EINVAL;
    if (!ctx->have_topo) return fail(ctx, SC_ESTATE, "load a topology first");
    const schost::Topology& T = ctx->topo;
    double wmin = 0;
    for (int i = 0; i < n; i++) {
        const schost::Ia& a = T.ia[types[i]][types[i]];
        if (a.geotype[0] == 0) return fail(ctx, SC_EINVAL, "undefined type in sc_generate_rsa");
        wmin = std::max(wmin, a.len[0] + a.sigma);
    }
    int nc[3];
    for (int k = 0; k < 3; k++) nc[k] = std::max(1, (int)floor(box[k] / std::max(wmin, 1e-9)));
    int ncell = nc[0] * nc[1] * nc[2];
    std::vector<std::vector<int>> cells(ncell);
    uint64_t h = mix64(seed + 0x1234567ull);
    auto u01 = [&]() { h = mix64(h); return (double)(h >> 11) * (1.0 / 9007199254740992.0); };
    auto cell_of = [&](const double* s) {
        int c[3];
        for (int k = 0; k < 3; k++) { c[k] = (int)(s[k] * nc[k]); if (c[k] >= nc[k]) c[k] = nc[k] - 1; }
        return c[0] + nc[0] * (c[1] + nc[1] * c[2]);
    };
    long attempts = 0;
    for (int i = 0; i < n; i++) {
        const schost::Ia& ai = T.ia[types[i]][types[i]];
        bool placed = false;
        while (!placed) {
            if (++attempts > max_attempts) return fail(ctx, SC_ECAPACITY, "random sequential addition ran out of attempts at particle " + std::to_string(i));
            double s[3] = {u01(), u01(), u01()};
            double z = 1 - 2 * u01(), ph = 6.283185307179586 * u01(), sr = sqrt(std::max(0.0, 1 - z * z));
            V3<double> d = mk<double>(sr * cos(ph), sr * sin(ph), z);
            // a unit vector perpendicular to d, rotated by a uniform angle about d
            V3<double> ref = fabs(d.x) < 0.9 ? mk<double>(1, 0, 0) : mk<double>(0, 1, 0);
            V3<double> e1 = normalised(perp(ref, d)), e2 = cross(d, e1);
            double psi = 6.283185307179586 * u01();
            V3<double> q = mk<double>(cos(psi) * e1.x + sin(psi) * e2.x, cos(psi) * e1.y + sin(psi) * e2.y,
                                      cos(psi) * e1.z + sin(psi) * e2.z);
            int c = cell_of(s);
            int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
            bool ok = true;
            for (int dz = -1; dz <= 1 && ok; dz++)
                for (int dy = -1; dy <= 1 && ok; dy++)
                    for (int dx = -1; dx <= 1 && ok; dx++) {
                        if ((nc[0] < 3 && dx == 1 && nc[0] == 2) || (nc[0] == 1 && dx != 0)) continue;
                        if ((nc[1] < 3 && dy == 1 && nc[1] == 2) || (nc[1] == 1 && dy != 0)) continue;
                        if ((nc[2] < 3 && dz == 1 && nc[2] == 2) || (nc[2] == 1 && dz != 0)) continue;
                        int x = (cx + dx + nc[0]) % nc[0], y = (cy + dy + nc[1]) % nc[1], zz = (cz + dz + nc[2]) % nc[2];
                        for (int j : cells[x + nc[0] * (y + nc[1] * zz)]) {
                            const schost::Ia& p = T.ia[types[i]][types[j]];
                            V3<double> r;
                            double dd[3];
                            for (int k = 0; k < 3; k++) {
                                double df = s[k] - pos[3 * j + k];
                                dd[k] = box[k] * (df - nearbyint(df));
                            }
                            r = mk<double>(dd[0], dd[1], dd[2]);
                            V3<double> dj = mk<double>(dir[3 * j], dir[3 * j + 1], dir[3 * j + 2]);
                            bool si = ai.geotype[0] >= 30, sj = T.ia[types[j]][types[j]].geotype[0] >= 30;
                            double d2;
                            if (si && sj) d2 = dot(r, r);
                            else if (!si && !sj) { V3<double> v = min_dist(d, dj, p.half_len[0], p.half_len[1], r); d2 = dot(v, v); }
                            else {
                                double contt;
                                V3<double> dv;
                                d2 = si ? sp_dist<double>(mk<double>(-r.x, -r.y, -r.z), dj, p.half_len[1], contt, dv)
                                        : sp_dist<double>(r, d, p.half_len[0], contt, dv);
                            }
                            if (d2 < p.sigma * p.sigma) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            for (int k = 0; k < 3; k++) pos[3 * i + k] = s[k];
            dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
            pd0[3 * i] = q.x; pd0[3 * i + 1] = q.y; pd0[3 * i + 2] = q.z;
            cells[c].push_back(i);
            placed = true;
        }
    }
    return SC_OK;
}

extern "C" int sc_generate_rsa(sc_ctx* ctx, int n, const int32_t* types, const double box[3], uint64_t seed,
                               long max_attempts, double* pos, double* dir, double* pd0) {
    SC_GUARD(ctx, sc_generate_rsa_impl(ctx, n, types, box, seed, max_attempts, pos, dir, pd0))
}
