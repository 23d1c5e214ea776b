// Device helpers shared by the lattice-filter, mean-field and adjoint kernels.
//
// Lane groups.  A group of S = L / R lanes owns one item (a splat chunk, a vertex, a
// pixel); lane l owns the R consecutive labels [l*R, l*R + R) of the item's row, so a
// row of L floats is read and written as S contiguous R-float pieces (dwordx R
// accesses).  R is the largest of 4, 3, 2 dividing L, else 1 (choose_labels_per_lane):
// no lane holds a partial piece, so no access is conditional; per-item work (addresses,
// reductions, exp) is shared by R labels.
//   WAVE = true  (S <= 64): 64 / S groups per 64-wide wave, the block is 4 waves.  A
//                 group never spans waves; group-wide max / sum / peer reads are
//                 ds_bpermute shuffles with no block barrier.
//   WAVE = false (S > 64): block / S groups per block of 256, 512 or 1024 threads;
//                 group-wide operations go through LDS with __syncthreads().
// Every group-wide operation must be reached by all threads of the block.
// Reductions have a fixed order (the lane's labels in order, then a pairwise tree over
// lanes in wave mode or lane order in block mode): results do not vary between runs.
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

// Position of pixel p (row-major, image W pixels wide) when the image is stored in
// 8 x 8 pixel blocks, blocks row-major, pixels row-major inside a block.
__device__ __forceinline__ int q_block_pos(int p, int W) {
    const int y = p / W, x = p - y * W;
    const int nbx = (W + 7) >> 3;
    return (((y >> 3) * nbx + (x >> 3)) << 6) + ((y & 7) << 3) + (x & 7);
}

// Last i in [0, n) with base[i] <= q (base ascending, base[0] <= q).
__device__ __forceinline__ int last_le(const int* __restrict__ base, int n, int q) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Pixels of tile t: rows of tw pixels starting at p0, p0 + stride, ... (th rows).
// 2-D instances are cut into side x side squares (row-major over the image), other
// instances into runs of side*side pixels.
__device__ __forceinline__ void tile_rect(const TileGeom& g, int t, int& p0, int& tw, int& th, int& stride) {
    const int i = g.n_inst > 1 ? last_le(g.tile_base, g.n_inst, t) : 0;
    const int lt = t - g.tile_base[i];
    const int Ni = g.pix_base[i + 1] - g.pix_base[i];
    if (g.is2d) {
        const int W = g.W[i], H = Ni / W;
        const int ntx = (W + g.side - 1) / g.side;
        const int ty = lt / ntx, tx = lt - ty * ntx;
        const int x0 = tx * g.side, y0 = ty * g.side;
        tw = min(g.side, W - x0);
        th = min(g.side, H - y0);
        p0 = g.pix_base[i] + y0 * W + x0;
        stride = W;
    } else {
        const int run = g.side * g.side;
        p0 = g.pix_base[i] + lt * run;
        tw = min(run, Ni - lt * run);
        th = 1;
        stride = 0;
    }
}

// Index of pixel p inside its tile (row-major in the tile).
__device__ __forceinline__ int tile_local(const TileGeom& g, int p) {
    const int i = g.n_inst > 1 ? last_le(g.pix_base, g.n_inst, p) : 0;
    const int local = p - g.pix_base[i];
    if (g.is2d) {
        const int W = g.W[i];
        const int y = local / W, x = local - y * W;
        const int x0 = (x / g.side) * g.side;
        const int tw = min(g.side, W - x0);
        return (y % g.side) * tw + (x - x0);
    }
    return local % (g.side * g.side);
}

// R consecutive floats at 4-byte alignment: one global_load/store_dwordxR.
template <int R>
struct __attribute__((aligned(4))) Piece {
    float v[R];
};

// Labels k0..k0+R-1 of a row: one R-float access (R divides L, so every lane's piece
// is whole and no access is conditional: loads issued in a row stay independent).
template <int R>
__device__ __forceinline__ void load_piece(const float* __restrict__ row, int k0, int L, float (&x)[R]) {
    (void)L;
    const Piece<R> p = *reinterpret_cast<const Piece<R>*>(row + k0);
#pragma unroll
    for (int r = 0; r < R; r++) x[r] = p.v[r];
}

template <int R>
__device__ __forceinline__ void store_piece(float* __restrict__ row, int k0, int L, const float (&x)[R]) {
    (void)L;
    Piece<R> p;
#pragma unroll
    for (int r = 0; r < R; r++) p.v[r] = x[r];
    *reinterpret_cast<Piece<R>*>(row + k0) = p;
}

template <bool WAVE, int R>
struct LaneGroup {
    int L, S, g, l, k0, base, t, per;
    bool on;
    float* sh;  // block mode: block * R floats

    // per: groups per wave (WAVE) or per block
    __device__ __forceinline__ LaneGroup(int L_, int S_, int per_, float* sh_) : L(L_), S(S_), per(per_), sh(sh_) {
        t = threadIdx.x;
        const int lane = WAVE ? (t & 63) : t;
        g = lane / S;
        l = lane - g * S;
        on = g < per;
        if (!on) {  // idle lanes shadow group 0 so that every shuffle index stays valid
            g = 0;
            l = l < S ? l : 0;
        }
        k0 = l * R;
        base = WAVE ? (t & ~63) + g * S : g * S;
    }
    // Index of this group's item in block b of a grid.
    __device__ __forceinline__ int item_in(int b) const {
        return WAVE ? (b * (kWaveBlock / 64) + (t >> 6)) * per + g : b * per + g;
    }
    // Index of this group's item (blocks taken in XCD-contiguous order).
    __device__ __forceinline__ int item() const { return item_in(xcd_block(blockIdx.x, gridDim.x)); }
    __device__ static __forceinline__ int items_per_block(int per) { return WAVE ? (kWaveBlock / 64) * per : per; }
    __device__ __forceinline__ bool has(int) const { return true; }  // R divides L: every slot is a label

    // Value of lane s of this group (all lanes must call).
    __device__ __forceinline__ float lane_value(float x, int s) {
        if (WAVE) return __shfl(x, (base + s) & 63);
        sh[t] = x;
        __syncthreads();
        const float v = sh[base + s];
        __syncthreads();
        return v;
    }

    __device__ __forceinline__ float all_max(const float (&x)[R]) {
        float m = x[0];
#pragma unroll
        for (int r = 1; r < R; r++)
            if (has(r)) m = fmaxf(m, x[r]);
        if (WAVE) {
            // ring doubling over the S lanes; max is idempotent, so overlap is harmless
            for (int k = 1; k < S; k <<= 1) {
                int s = l + k;
                if (s >= S) s -= S;
                m = fmaxf(m, __shfl(m, (base + s) & 63));
            }
            return m;
        }
        sh[t] = m;
        __syncthreads();
        float v = sh[base];
        for (int s = 1; s < S; s++) v = fmaxf(v, sh[base + s]);
        __syncthreads();
        return v;
    }

    __device__ __forceinline__ float all_sum(const float (&x)[R]) {
        float a = x[0];
#pragma unroll
        for (int r = 1; r < R; r++)
            if (has(r)) a = a + x[r];
        if (WAVE) {
            // pairwise tree over the S lanes towards lane 0 of the group, then a broadcast
            for (int k = 1; k < S; k <<= 1) {
                const float y = __shfl(a, (base + l + k) & 63);
                if ((l & (2 * k - 1)) == 0 && l + k < S) a = a + y;
            }
            return __shfl(a, base & 63);
        }
        sh[t] = a;
        __syncthreads();
        float v = 0.0f;
        for (int s = 0; s < S; s++) v = v + sh[base + s];
        __syncthreads();
        return v;
    }
};

// The adjoint kernels' per-block compatibility statistic (SliceTerm::gstat): the sum of
// each label's products over the block's groups, in a fixed order.  stat_put leaves the
// thread's R products (0 in lanes without a pixel) in slot `slot` of the region sst
// (blockDim.x * R floats per slot).  stat_finish is called by every thread of the block
// after all slots are put, at the kernel's end:
//   WAVE: after a barrier, one thread per (slot, wave, label) adds the wave's groups in
//   group order into the wave's row of L floats (after the slots), then, after a second
//   barrier, the block's waves are added in wave order;
//   block mode: after a barrier, the block's groups are added in group order.
// It writes the sums to the term's row of this block.
template <int R>
__device__ __forceinline__ void stat_put(float* sst, int slot, const float (&v)[R]) {
    float* d = sst + ((size_t)slot * blockDim.x + threadIdx.x) * R;
#pragma unroll
    for (int r = 0; r < R; r++) d[r] = v[r];
}

// NT: the terms to look the slots up in (st.t[0 .. NT-1]).
template <bool WAVE, int R, int NT = DCRF_MAX_TERMS>
__device__ __forceinline__ void stat_finish(const SliceTable& st, const LaneGroup<WAVE, R>& G, float* sst) {
    const int L = G.L;
    float* wrow = sst + (size_t)st.nstat * blockDim.x * R;  // WAVE: [slot][wave][L]
    __syncthreads();
    if (WAVE) {  // one thread per (slot, wave, label): the wave's groups in group order
        constexpr int W = kWaveBlock / 64;
        for (int j = threadIdx.x; j < st.nstat * W * L; j += blockDim.x) {
            const int slot = j / (W * L), rem = j - slot * (W * L);
            const int w = rem / L, lab = rem - w * L;
            const int ln = lab / R, rr = lab - ln * R;
            const float* src = sst + (size_t)slot * blockDim.x * R + (size_t)(w * 64 + ln) * R + rr;
            float s = 0.0f;
            for (int gi = 0; gi < G.per; gi++) s = s + src[(size_t)gi * G.S * R];
            wrow[j] = s;
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < st.nstat * L; j += blockDim.x) {
        const int slot = j / L, lab = j - slot * L;
        float s = 0.0f;
        if (WAVE) {
            const float* src = wrow + (size_t)slot * (kWaveBlock / 64) * L + lab;
            for (int w = 0; w < kWaveBlock / 64; w++) s = s + src[(size_t)w * L];
        } else {
            const int ln = lab / R, rr = lab - ln * R;  // lane and register holding label lab
            const float* src = sst + (size_t)slot * blockDim.x * R;
            for (int gi = 0; gi < G.per; gi++) s = s + src[(gi * G.S + ln) * R + rr];
        }
        float* out = nullptr;
#pragma unroll
        for (int k = 0; k < NT; k++)
            if (k < st.nterms && st.t[k].gslot == slot) out = st.t[k].gstat;
        out[(size_t)blockIdx.x * L + lab] = s;
    }
}

// Floats of the statistic region for nstat terms.
__host__ __device__ inline size_t stat_lds_floats(bool wave, int nstat, int L, int block, int R) {
    return (size_t)nstat * block * R + (wave ? (size_t)nstat * (kWaveBlock / 64) * L : 0);
}

// Slice of term T for pixel p: m[r] = sum_j sw_j * vals[vtx_j][k0 + r] in j order.  The
// (vertex, weight) pairs are loaded first, then all vertex rows, so the d+1 gathers are
// in flight together.
template <int R>
__device__ __forceinline__ void slice_piece(const SliceTerm& T, int p, int k0, int L, float (&m)[R]) {
    const int d1 = T.d1;
    const int* vt = T.vtx + (size_t)p * d1;
    const float* sw = T.sw + (size_t)p * d1;
    int v[kMaxD1];
    float w[kMaxD1];
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) {
            v[j] = vt[j];
            w[j] = sw[j];
        }
    float x[kMaxD1][R];
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) load_piece<R>(T.vals + (size_t)v[j] * T.ldv, k0, L, x[j]);
#pragma unroll
    for (int r = 0; r < R; r++) m[r] = 0.0f;
#pragma unroll
    for (int j = 0; j < kMaxD1; j++)
        if (j < d1) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float prod = w[j] * x[j][r];
                m[r] = m[r] + prod;
            }
        }
}

// Copy the Matrix compatibilities that have an LDS slot (cw_lds >= 0) into the compat
// region shc.  Every thread of the block must call it (one barrier when anything is
// staged; the condition is uniform).
__device__ __forceinline__ void stage_compat(const SliceTable& st, int L, float* shc) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < DCRF_MAX_TERMS; k++)
        if (k < st.nterms && st.t[k].cw_lds >= 0) {
            any = true;
            const float* src = st.t[k].cw;
            float* dst = shc + st.t[k].cw_lds;
            const int ldw = st.t[k].ldw;
            if (ldw == L) {
                for (int i = threadIdx.x; i < L * L; i += blockDim.x) dst[i] = src[i];
            } else {  // padded rows; the pad columns read as 0
                for (int i = threadIdx.x; i < L * ldw; i += blockDim.x) {
                    const int row = i / ldw, col = i - row * ldw;
                    dst[i] = col < L ? src[row * L + col] : 0.0f;
                }
            }
        }
    if (any) __syncthreads();
}

// c[r] = sum_b cw[k0 + r][b] m_b over the pixel's whole message vector (read from its
// group in label order), one fused multiply-add per weight.
template <bool WAVE, int R>
__device__ __forceinline__ void matrix_piece(const float* __restrict__ cw, int ldw, const float (&m)[R],
                                             float (&c)[R], LaneGroup<WAVE, R>& G) {
    const int L = G.L;
#pragma unroll
    for (int r = 0; r < R; r++) c[r] = 0.0f;
    for (int s = 0; s < G.S; s++) {
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const float mb = G.lane_value(m[rr], s);
            const int b = s * R + rr;
            if (b < L) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (G.has(r)) c[r] = fmaf(cw[(size_t)(G.k0 + r) * ldw + b], mb, c[r]);
            }
        }
    }
}

// Compatibility transform of the lane's message values (labelcompatibility.cpp:46-48,
// 66-69, 82-84).  Matrix reads the pixel's whole message vector from its group, in
// label order, and its weights from the LDS copy when there is one.  All lanes must
// call (Matrix synchronises in block mode).
template <bool WAVE, int R>
__device__ __forceinline__ void compat_piece(const SliceTerm& T, const float (&m)[R], float (&c)[R],
                                             LaneGroup<WAVE, R>& G, const float* shc) {
    if (T.ckind == DCRF_COMPAT_POTTS) {
        const float nw = -T.cw[0];
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = nw * m[r];
        return;
    }
    if (T.ckind == DCRF_COMPAT_DIAGONAL) {
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = G.has(r) ? T.cw[G.k0 + r] * m[r] : 0.0f;
        return;
    }
    // two copies of the loop, one per weight source, so that each pointer keeps its
    // address space (one select between them would make every weight read a flat load)
    if (T.cw_lds >= 0)
        matrix_piece<WAVE, R>(shc + T.cw_lds, T.ldw, m, c, G);
    else
        matrix_piece<WAVE, R>(T.cw, G.L, m, c, G);
}

// matrix_piece with the pixel's message vector already in LDS (mrow, 16-byte aligned)
// and the weights in LDS rows of stride ldw (a multiple of 4, 16-byte aligned): 16-byte
// reads of both, no cross-lane traffic; the same fused multiply-adds in the same (label)
// order, so the results are bit-identical to matrix_piece's.
template <int R>
__device__ __forceinline__ void matrix_piece_rows(const float* __restrict__ w, int ldw,
                                                  const float* __restrict__ mrow, int L, int k0, float (&c)[R]) {
#pragma unroll
    for (int r = 0; r < R; r++) c[r] = 0.0f;
    int b = 0;
#pragma unroll 1
    for (; b + 4 <= L; b += 4) {
        const float4 mv = *reinterpret_cast<const float4*>(mrow + b);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float4 wv = *reinterpret_cast<const float4*>(w + (size_t)(k0 + r) * ldw + b);
            c[r] = fmaf(wv.x, mv.x, c[r]);
            c[r] = fmaf(wv.y, mv.y, c[r]);
            c[r] = fmaf(wv.z, mv.z, c[r]);
            c[r] = fmaf(wv.w, mv.w, c[r]);
        }
    }
#pragma unroll 1
    for (; b < L; b++) {
        const float mv = mrow[b];
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = fmaf(w[(size_t)(k0 + r) * ldw + b], mv, c[r]);
    }
}

// sumAndNormalize (densecrf.cpp:107-115) of x = Q .* g over the lane group:
// b = sum(x) q - x.
template <bool WAVE, int R>
__device__ __forceinline__ void sum_normalize(const float (&q)[R], const float (&gv)[R], float (&b)[R],
                                              LaneGroup<WAVE, R>& G) {
    float x[R];
#pragma unroll
    for (int r = 0; r < R; r++) x[r] = q[r] * gv[r];
    const float s = G.all_sum(x);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const float sq = s * q[r];
        b[r] = sq - x[r];
    }
}

}  // namespace dcrf

// Launch a lane-group kernel template <WAVE, R> (DCRF_DISPATCH) or <WAVE, R, C>
// (DCRF_DISPATCH_T) for the shape sh: one instantiation per (R, wave) pair.
#define DCRF_DISPATCH(sh, kernel, grid, shm, s, ...)                                     \
    do {                                                                                \
        const unsigned g_ = (grid);                                                     \
        const size_t m_ = (shm);                                                        \
        switch ((sh).R * 2 + ((sh).wave ? 1 : 0)) {                                     \
            case 3: kernel<true, 1><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 2: kernel<false, 1><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 5: kernel<true, 2><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 4: kernel<false, 2><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 7: kernel<true, 3><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 6: kernel<false, 3><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 9: kernel<true, 4><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            default: kernel<false, 4><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break; \
        }                                                                               \
        DCRF_HIP_CHECK(hipGetLastError());                                              \
    } while (0)

// As DCRF_DISPATCH for kernels with a third template argument CWL (compiled-in LDS
// staging of Matrix compatibilities): CWL = (cwl != 0).
#define DCRF_DISPATCH_T(sh, kernel, C, grid, shm, s, ...)                                   \
    do {                                                                                   \
        const unsigned g_ = (grid);                                                        \
        const size_t m_ = (shm);                                                           \
        switch ((sh).R * 2 + ((sh).wave ? 1 : 0)) {                                        \
            case 3: kernel<true, 1, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 2: kernel<false, 1, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 5: kernel<true, 2, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 4: kernel<false, 2, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 7: kernel<true, 3, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            case 6: kernel<false, 3, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;  \
            case 9: kernel<true, 4, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break;   \
            default: kernel<false, 4, C><<<g_, (sh).block, m_, (s)>>>(__VA_ARGS__); break; \
        }                                                                                  \
        DCRF_HIP_CHECK(hipGetLastError());                                                 \
    } while (0)
#define DCRF_DISPATCH_CWL(sh, cwl, kernel, grid, shm, s, ...)                     \
    do {                                                                         \
        if (cwl)                                                                 \
            DCRF_DISPATCH_T(sh, kernel, true, grid, shm, s, __VA_ARGS__);        \
        else                                                                     \
            DCRF_DISPATCH_T(sh, kernel, false, grid, shm, s, __VA_ARGS__);       \
    } while (0)
