// Device helpers shared by the lattice-filter, mean-field and adjoint kernels.
//
// Lane groups.  A group of L lanes owns one item (a splat chunk, a vertex, a pixel);
// lane l holds label l, so a row of L floats is one contiguous segment.
//   WAVE = true  (L <= 64): G = 64 / L groups per 64-wide wave, the block is 4 waves.
//                 A group never spans waves; group-wide max / sum / peer reads are
//                 ds_bpermute shuffles (no block barrier).  Lanes G*L..63 are idle.
//   WAVE = false (L > 64): block / L groups per block of 256, 512 or 1024 threads;
//                 group-wide operations go through LDS with __syncthreads().
// Every group-wide operation must be reached by all threads of the block.
// Reductions have a fixed order (a pairwise tree in wave mode, label order in block
// mode), so results do not vary from run to run.
#pragma once

#include "dcrf_internal.h"

#pragma clang fp contract(off)

namespace dcrf {

constexpr int kWaveBlock = 256;  // threads per block in wave mode (4 waves)
constexpr int kMaxD1 = DCRF_MAX_DIM + 1;
constexpr int kXcds = 8;         // MI355X: 8 XCDs, each with its own L2

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD b % 8).
// Renumber them so that each XCD processes one contiguous range of blocks, hence of
// items: neighbouring vertices / pixels, whose rows the gathers share, then meet in
// the same L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int x = b % kXcds, i = b / kXcds;
    const int q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + i;
}

template <bool WAVE>
struct LaneGroup {
    int L, g, l, base, t, per;
    bool on;
    float staged;
    float* sh;

    // per: groups per wave (WAVE) or per block; sh: block-shared floats (block mode)
    __device__ __forceinline__ LaneGroup(int L_, int per_, float* sh_) : L(L_), per(per_), sh(sh_) {
        t = threadIdx.x;
        const int lane = WAVE ? (t & 63) : t;
        g = lane / L;
        l = lane - g * L;
        on = g < per;
        if (!on) {  // idle lanes shadow group 0 so that every shuffle index stays valid
            g = 0;
            l = l < L ? l : 0;
        }
        base = WAVE ? (t & ~63) + g * L : g * L;
        staged = 0.0f;
    }
    // Index of this group's item.
    __device__ __forceinline__ int item() const {
        const int b = xcd_block(blockIdx.x, gridDim.x);
        return WAVE ? (b * (kWaveBlock / 64) + (t >> 6)) * per + g : b * per + g;
    }
    // Items per block.
    __device__ static __forceinline__ int items_per_block(int per) { return WAVE ? (kWaveBlock / 64) * per : per; }

    __device__ __forceinline__ void stage(float x) {
        if (WAVE) {
            staged = x;
        } else {
            sh[t] = x;
            __syncthreads();
        }
    }
    // Staged value of lane b of this group.
    __device__ __forceinline__ float peer(int b) const {
        if (WAVE) return __shfl(staged, (base + b) & 63);
        return sh[base + b];
    }
    __device__ __forceinline__ void unstage() {
        if (!WAVE) __syncthreads();
    }

    __device__ __forceinline__ float all_max(float x) {
        if (WAVE) {
            // ring doubling: after the step with offset k a lane holds the max of 2k
            // consecutive lanes (cyclically); max is idempotent, so overlap is harmless
            for (int k = 1; k < L; k <<= 1) {
                int s = l + k;
                if (s >= L) s -= L;
                x = fmaxf(x, __shfl(x, (base + s) & 63));
            }
            return x;
        }
        stage(x);
        float m = peer(0);
        for (int b = 1; b < L; b++) m = fmaxf(m, peer(b));
        unstage();
        return m;
    }

    __device__ __forceinline__ float all_sum(float x) {
        if (WAVE) {
            // pairwise tree towards lane 0 of the group, then a broadcast
            for (int k = 1; k < L; k <<= 1) {
                const float y = __shfl(x, (base + l + k) & 63);
                if ((l & (2 * k - 1)) == 0 && l + k < L) x = x + y;
            }
            return __shfl(x, base & 63);
        }
        stage(x);
        float s = 0.0f;
        for (int b = 0; b < L; b++) s = s + peer(b);
        unstage();
        return s;
    }
};

// Slice of term T for pixel p, label l: sum_j sw_j * vals[vtx_j][l] in j order.  The
// (vertex, weight) pairs are loaded first, then all vertex rows, so the d+1 gathers
// are in flight together.
__device__ __forceinline__ float slice_one(const SliceTerm& T, int p, int l, int L) {
    const int d1 = T.d1;
    const int* vt = T.vtx + (size_t)p * d1;
    const float* sw = T.sw + (size_t)p * d1;
    int v[kMaxD1];
    float w[kMaxD1], x[kMaxD1];
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) {
            v[j] = vt[j];
            w[j] = sw[j];
        }
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) x[j] = T.vals[(size_t)v[j] * L + l];
    float m = 0.0f;
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) {
            const float prod = w[j] * x[j];
            m = m + prod;
        }
    return m;
}

// Compatibility transform of one message value (labelcompatibility.cpp:46-48, 66-69,
// 82-84).  Matrix reads the pixel's whole message vector from its group.
template <bool WAVE>
__device__ __forceinline__ float compat_apply(const SliceTerm& T, float m, LaneGroup<WAVE>& G) {
    if (T.ckind == DCRF_COMPAT_POTTS) {
        const float nw = -T.cw[0];
        return nw * m;
    }
    if (T.ckind == DCRF_COMPAT_DIAGONAL) return T.cw[G.l] * m;
    G.stage(m);
    const float* row = T.cw + (size_t)G.l * G.L;
    float c = 0.0f;
    for (int b = 0; b < G.L; b++) {
        const float prod = row[b] * G.peer(b);
        c = c + prod;
    }
    G.unstage();
    return c;
}

}  // namespace dcrf
