// Lattice filtering and the fused mean-field update for gfx950.
//
// Replaces Permutohedral::compute (pydensecrf/densecrf/src/permutohedral.cpp:476-610),
// DenseKernel::filter (pairwise.cpp:63-80), LabelCompatibility::apply
// (labelcompatibility.cpp:46-87) and expAndNormalize (densecrf.cpp:98-106).
//
// Thread mapping: a lane group of S = L / R consecutive threads handles one item (a
// splat chunk, a lattice vertex or a pixel), lane l owning the R labels [l*R, l*R+R),
// so every vertex/pixel row of L floats is read and written as S contiguous R-float
// pieces (dcrf_device.h).  All reductions run in a fixed order: results are
// deterministic from run to run.
//   splat    each chunk (<= C contributors of one vertex, pixel ascending) is summed
//            sequentially; vertices with several chunks write partial rows that the
//            combine kernel adds in chunk order.  No atomics.
//   blur     one launch per step; step s covers the terms whose lattice has axis s
//            (a prefix of the vertex space, which is ordered by decreasing d):
//            new = old + 0.5 * (n1 + n2), the reference's update (:556-569).
//   meanfield per pixel: slice every term (sum_j w_j*alpha*norm_p * row_j), apply the
//            compatibility, subtract from -U, softmax over the lane group.
#include "dcrf_internal.h"
#include "dcrf_device.h"

#include <cmath>
#include <type_traits>

#pragma clang fp contract(off)

namespace dcrf {

// Labels per lane: the largest R in {4, 3, 2} that divides L, else 1 (so every lane's
// piece of a row is whole: no partial lanes, no conditional accesses).
int choose_labels_per_lane(int L) {
    for (int R = 4; R >= 2; R--)
        if (L % R == 0) return R;
    return 1;
}

LaunchShape shape_for(int L) { return shape_for_r(L, choose_labels_per_lane(L)); }

LaunchShape shape_for_r(int L, int R) {
    const int S = (L + R - 1) / R;
    if (S <= 64) return LaunchShape{kWaveBlock, 64 / S, true, (kWaveBlock / 64) * (64 / S), R, S};
    const int block = S <= 256 ? 256 : (S <= 512 ? 512 : 1024);
    return LaunchShape{block, block / S, false, block / S, R, S};
}

namespace {

constexpr int kThreads = 256;
constexpr int kBatch = 8;  // contributors / partial rows loaded together
inline unsigned blocks_for(size_t n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

__global__ void k_fill(float* p, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Splat: each group sums one chunk (<= C contributors of a vertex, in entry order)
// sequentially.  Entries are (pixel, weight bits) pairs.  A chunk with out < 0 continues
// the chunk before it in the same wave (same vertex): WAVE mode adds it to that chunk's
// sum instead of storing it.
//   WAVE mode: the chunks of a wave are consecutive and their entry ranges adjoin (the
//   layout guarantees chunk c+1 begins where chunk c ends), so the wave loads its whole
//   entry range cooperatively (64 consecutive pairs per load instruction) into an LDS
//   window of kWin pairs, and each group reads its pairs from there: one cache line
//   serves 16 pairs instead of one load per lane and pair.  Input rows are gathered
//   kBatch at a time.
//   block mode (a group wider than a wave): the pairs are loaded by every lane.
constexpr int kWin = 512;  // entry pairs per wave window

template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_splat(const Chunk* __restrict__ chunks, int n,
                                                const int2* __restrict__ ent, const float* __restrict__ in, int ldi,
                                                float* __restrict__ vals, int ldv, int L, int S, int per) {
    LaneGroup<WAVE, R> G(L, S, per, nullptr);
    const int c = G.item();
    const int k0 = G.k0;
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    if (WAVE) {
        extern __shared__ int2 win_all[];
        int2* win = win_all + (threadIdx.x >> 6) * (kWin + kBatch);
        const int lane = threadIdx.x & 63;
        const int cw0 = c - G.g;  // the wave's first chunk
        if (cw0 >= n) return;     // the whole wave is idle (no block barrier below)
        const bool mine = G.on && c < n;
        const Chunk ch = chunks[mine ? c : cw0];
        const int r0 = chunks[cw0].beg, r1 = chunks[min(cw0 + per - 1, n - 1)].end;
        for (int w0 = r0; w0 < r1; w0 += kWin) {  // wave-uniform
            const int wn = min(kWin, r1 - w0);
            // all loads first (clamped, unconditional), then the stores: the loads stay
            // in flight together
            int2 t[kWin / 64];
#pragma unroll
            for (int u = 0; u < kWin / 64; u++) t[u] = ent[w0 + min(lane + 64 * u, wn - 1)];
            __builtin_amdgcn_wave_barrier();  // the previous window's reads come first
#pragma unroll
            for (int u = 0; u < kWin / 64; u++)
                if (lane + 64 * u < wn) win[lane + 64 * u] = t[u];
            if (lane < kBatch) win[wn + lane] = make_int2(0, 0);  // padding: pixel 0, weight 0
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (mine) {
                const int a = max(ch.beg, w0) - w0, b = min(ch.end, w0 + wn) - w0;
                for (int i = a; i < b; i += kBatch) {
                    int2 e[kBatch];
                    float xx[kBatch][R];
#pragma unroll
                    for (int u = 0; u < kBatch; u++) e[u] = win[i + u];  // i + u < wn + kBatch
#pragma unroll
                    for (int u = 0; u < kBatch; u++) load_piece<R>(in + (size_t)e[u].x * ldi, k0, L, xx[u]);
#pragma unroll
                    for (int u = 0; u < kBatch; u++)
                        if (i + u < b) {
                            const float wu = __int_as_float(e[u].y);
#pragma unroll
                            for (int r = 0; r < R; r++) {
                                const float prod = wu * xx[u][r];
                                acc[r] = acc[r] + prod;
                            }
                        }
                }
            }
        }
        // A run of the wave's chunks with one vertex (the layout gives the later ones
        // out = -1): their sums are folded into the run's first group, right to left, in
        // a fixed order.  Lists without such runs skip this.
        const int nv = __shfl(ch.v, (lane + S) & 63);
        const int nout = __shfl(ch.out, (lane + S) & 63);
        const bool cont = mine && G.g + 1 < per && c + 1 < n && nout < 0 && nv == ch.v;
        if (__any(cont)) {
            for (int j = per - 2; j >= 0; j--) {  // wave-uniform
                float y[R];
#pragma unroll
                for (int r = 0; r < R; r++) y[r] = __shfl(acc[r], (lane + S) & 63);
                if (cont && G.g == j)
#pragma unroll
                    for (int r = 0; r < R; r++) acc[r] = acc[r] + y[r];
            }
        }
        if (mine && ch.out >= 0) store_piece<R>(vals + (size_t)ch.out * ldv, k0, L, acc);
        return;
    }
    if (!G.on || c >= n) return;
    const Chunk ch = chunks[c];
    for (int i = ch.beg; i < ch.end; i += kBatch) {
        int2 e[kBatch];
        float xx[kBatch][R];
#pragma unroll
        for (int u = 0; u < kBatch; u++) e[u] = ent[min(i + u, ch.end - 1)];
#pragma unroll
        for (int u = 0; u < kBatch; u++) load_piece<R>(in + (size_t)e[u].x * ldi, k0, L, xx[u]);
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (i + u < ch.end) {
                const float wu = __int_as_float(e[u].y);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const float prod = wu * xx[u][r];
                    acc[r] = acc[r] + prod;
                }
            }
    }
    store_piece<R>(vals + (size_t)ch.out * ldv, k0, L, acc);
}

// Tiled splat: block = one tile.  Stage 1 copies the tile's input rows, its entries
// (pixel index in the tile, weight) and its segments (beg, end, out) into LDS, every
// thread keeping kStage loads in flight; stage 2 gives the segments to the lane groups
// of the block, which sum them from LDS in entry order (the same order, hence the same
// bits, as k_splat).
constexpr int kStage = 32;

template <int R>
__global__ void __launch_bounds__(1024) k_splat_tiled(TiledSplat ts, const int2* __restrict__ ent,
                                                            const float* __restrict__ in, float* __restrict__ vals,
                                                            int ldv, int L, int S, int per, int q_max) {
    extern __shared__ float sh[];
    const int t = xcd_block(blockIdx.x, gridDim.x);
    int p0, tw, th, stride;
    tile_rect(ts.geom, t, p0, tw, th, stride);
    const int e0 = ts.tile_e[t], ne = ts.tile_e[t + 1] - e0;
    const int c0 = ts.tile_seg[t], ns = ts.tile_seg[t + 1] - c0;
    const int rowf = tw * L;
    const int n1 = th * rowf, n3 = n1 + 2 * ne, n4 = n3 + 3 * ns;
    const float inv_rowf = 1.0f / (float)rowf;
    const long long rskip = (long long)stride * L - rowf;  // gap between two tile rows in `in`
    const unsigned* __restrict__ qsrc = reinterpret_cast<const unsigned*>(in) + (size_t)p0 * L;
    const unsigned* __restrict__ esrc = reinterpret_cast<const unsigned*>(ent + e0);
    const unsigned* __restrict__ csrc = reinterpret_cast<const unsigned*>(ts.chunks + c0);
    unsigned* shu = reinterpret_cast<unsigned*>(sh);
    // LDS: [rows: q_max floats][entries: e_max (index, weight) pairs][segments: 3 s_max]
    const int nt = blockDim.x;
    // every load is unconditional (index clamped to the last element, source chosen by
    // select), so the kStage loads of a thread are in flight together; stores follow
    for (int b = threadIdx.x; b < n4; b += kStage * nt) {
        unsigned x[kStage];
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int i = min(b + u * nt, n4 - 1);
            int r = (int)((float)i * inv_rowf);  // row of the tile (exact within one for i < 2^24)
            r -= r * rowf > i ? 1 : 0;
            r += (r + 1) * rowf <= i ? 1 : 0;
            const int j = i - n3;
            const int c = j / 3;
            const unsigned* q = qsrc + (i + (long long)r * rskip);
            const unsigned* e = esrc + (i - n1);
            const unsigned* h = csrc + (4 * c + 1 + (j - 3 * c));  // Chunk {v, beg, end, out}
            x[u] = *(i < n1 ? q : (i < n3 ? e : h));
        }
#pragma unroll
        for (int u = 0; u < kStage; u++) {
            const int i = b + u * nt;
            if (i < n4) {
                const int off = i < n1 ? 0 : (i < n3 ? q_max - n1 : q_max + 2 * ts.e_max - n3);
                shu[i + off] = x[u];
            }
        }
    }
    __syncthreads();
    const float* tq = sh;
    const int2* te = reinterpret_cast<const int2*>(sh + q_max);
    const int* tsg = reinterpret_cast<const int*>(sh + q_max + 2 * ts.e_max);
    LaneGroup<true, R> G(L, S, per, nullptr);
    if (!G.on) return;  // no barrier below
    const int groups = (nt / 64) * per;
    const int k0 = G.k0;
    for (int c = (threadIdx.x >> 6) * per + G.g; c < ns; c += groups) {
        const int b = tsg[3 * c] - e0, e = tsg[3 * c + 1] - e0, out = tsg[3 * c + 2];
        float acc[R];
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] = 0.0f;
        int i = b;
        for (; i + 4 <= e; i += 4) {
            int pp[4];
            float ww[4], xx[4][R];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int2 x = te[i + u];
                pp[u] = x.x;
                ww[u] = __int_as_float(x.y);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) load_piece<R>(tq + pp[u] * L, k0, L, xx[u]);
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const float prod = ww[u] * xx[u][r];
                    acc[r] = acc[r] + prod;
                }
        }
        for (; i < e; i++) {
            float xx[R];
            const int2 x = te[i];
            const float ww = __int_as_float(x.y);
            load_piece<R>(tq + x.x * L, k0, L, xx);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float prod = ww * xx[r];
                acc[r] = acc[r] + prod;
            }
        }
        store_piece<R>(vals + (size_t)out * ldv, k0, L, acc);
    }
}

template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_combine(const Combine* __restrict__ cb, int n, float* __restrict__ vals,
                                                  int ldv, int L, int S, int per) {
    LaneGroup<WAVE, R> G(L, S, per, nullptr);
    const int c = G.item();
    if (!G.on || c >= n) return;
    const int k0 = G.k0;
    const Combine rc = cb[c];
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    for (int i = 0; i < rc.count; i += kBatch) {  // count >= 2; loads clamped to the last row
        float xx[kBatch][R];
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            load_piece<R>(vals + (size_t)(rc.first + min(i + u, rc.count - 1)) * ldv, k0, L, xx[u]);
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (i + u < rc.count)
#pragma unroll
                for (int r = 0; r < R; r++) acc[r] = acc[r] + xx[u][r];
    }
    store_piece<R>(vals + (size_t)rc.v * ldv, k0, L, acc);
}

// One-launch recombination (wave mode).  Items are lane-group sums of cb[c].count
// partial rows from cb[c].first.  An item with pad == 0 is a whole vertex and stores
// its sum as row cb[c].v (count 0: a filler, stores nothing).  A vertex with more rows
// fills one whole wave: its items (pad = the number of pieces, the fillers after them
// count 0) sum one piece each; the pieces go to the wave's LDS rows and group 0 adds them
// in piece order and stores the vertex row.  A wave holds either light items or one
// heavy vertex, so the branch is uniform per wave.
template <int R>
__global__ void __launch_bounds__(kWaveBlock) k_combine_waves(const Combine* __restrict__ cb, int n,
                                                              float* __restrict__ vals, int ldv, int L, int S,
                                                              int per) {
    extern __shared__ float rows_lds[];  // per wave: per rows of L floats
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int c = G.item();
    const int cw = c - G.g;  // the wave's first item
    if (cw >= n) return;     // whole waves only: no item of this wave exists
    const bool live = G.on && c < n;
    const Combine rc = live ? cb[c] : Combine{0, 0, 0, 0};
    const int k0 = G.k0;
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    for (int i = 0; i < rc.count; i += kBatch) {  // loads clamped to the last row
        float xx[kBatch][R];
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            load_piece<R>(vals + (size_t)(rc.first + min(i + u, rc.count - 1)) * ldv, k0, L, xx[u]);
#pragma unroll
        for (int u = 0; u < kBatch; u++)
            if (i + u < rc.count)
#pragma unroll
                for (int r = 0; r < R; r++) acc[r] = acc[r] + xx[u][r];
    }
    const Combine head = cb[cw];  // uniform over the wave
    if (head.pad == 0) {
        if (rc.count > 0) store_piece<R>(vals + (size_t)rc.v * ldv, k0, L, acc);
        return;
    }
    float* wrows = rows_lds + (size_t)(threadIdx.x >> 6) * per * L;
    if (G.on && G.g < head.pad) store_piece<R>(wrows + (size_t)G.g * L, k0, L, acc);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (G.on && G.g == 0) {
        float sum[R];
#pragma unroll
        for (int r = 0; r < R; r++) sum[r] = 0.0f;
        for (int j = 0; j < head.pad; j++) {
            float x[R];
            load_piece<R>(wrows + (size_t)j * L, k0, L, x);
#pragma unroll
            for (int r = 0; r < R; r++) sum[r] = sum[r] + x[r];
        }
        store_piece<R>(vals + (size_t)head.v * ldv, k0, L, sum);
    }
}

template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_blur(const float* __restrict__ src, float* __restrict__ dst, int v0,
                                               int Vp, int ldv, int L, int S, int per, BlurTable bt, int step,
                                               int reverse) {
    LaneGroup<WAVE, R> G(L, S, per, nullptr);
    const int v = v0 + G.item();  // vertices [v0, Vp)
    if (!G.on || v >= Vp) return;
    const int k0 = G.k0;
    // terms are ordered by vbase; pick the last one starting at or before v
    int vb = bt.t[0].vbase, V = bt.t[0].V, d = bt.t[0].d;
    const int2* nbr = bt.t[0].nbr;
#pragma unroll
    for (int k = 1; k < DCRF_MAX_TERMS; k++) {
        if (k < bt.nterms && v >= bt.t[k].vbase) {
            vb = bt.t[k].vbase;
            V = bt.t[k].V;
            d = bt.t[k].d;
            nbr = bt.t[k].nbr;
        }
    }
    const int j = reverse ? d - step : step;
    float o[R], a[R], b[R], out[R];
    const int2 nb = nbr[(size_t)j * V + (v - vb)];
    load_piece<R>(src + (size_t)v * ldv, k0, L, o);
    load_piece<R>(src + (size_t)nb.x * ldv, k0, L, a);
    load_piece<R>(src + (size_t)nb.y * ldv, k0, L, b);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const float s2 = a[r] + b[r];
        const float h = 0.5f * s2;
        out[r] = o[r] + h;
    }
    store_piece<R>(dst + (size_t)v * ldv, k0, L, out);
}

// A pass of K = 1 or 2 consecutive blur steps (axes a_1 .. a_K) of one term.  Each step
// is x' = x + 0.5 (x_n1 + x_n2) in k_blur's operation order; the pass table lists per
// vertex v the 3^K rows of the step tree: level K holds v and its two neighbours along
// a_K, each expanded along a_{K-1}, and so on, zero row for an absent vertex (whose
// value is 0 at every level, as in the intermediate buffers of single steps).  The
// kernel loads the 3^K rows together and folds them level by level: the same bits as
// K launches of k_blur, with K-1 fewer launches and no intermediate rows written.
__device__ __forceinline__ float blur_fold(float o, float a, float b) {
    const float s2 = a + b;
    const float h = 0.5f * s2;
    return o + h;
}

// zr: the zero row (absent neighbours point to it; it holds zeros in every vertex
// buffer), whose loads are skipped: a group that has it takes 0.0f instead, the value
// the load would give, and its lanes issue no memory access.  Half of the rows of a
// sparse (bilateral) lattice's pass are absent neighbours.
template <int K, int R>
__device__ __forceinline__ void blur_pass_vertex(const float* __restrict__ src, float* __restrict__ dst,
                                                 const int* __restrict__ e, int v, int ldv, int k0, int L, int zr) {
    constexpr int NR = K == 1 ? 3 : (K == 2 ? 9 : 27);
    int idx[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) idx[i] = e[i];
    float x[NR][R];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        if (idx[i] != zr) {
            load_piece<R>(src + (size_t)idx[i] * ldv, k0, L, x[i]);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) x[i][r] = 0.0f;
        }
    }
    float out[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        float y[NR];
#pragma unroll
        for (int i = 0; i < NR; i++) y[i] = x[i][r];
        int n = NR;
#pragma unroll
        for (int lev = 0; lev < K; lev++) {
#pragma unroll
            for (int i = 0; i < NR / 3; i++)
                if (i < n / 3) y[i] = blur_fold(y[3 * i], y[3 * i + 1], y[3 * i + 2]);
            n /= 3;
        }
        out[r] = y[0];
    }
    store_piece<R>(dst + (size_t)v * ldv, k0, L, out);
}

// A pass of three steps: the 27-row tree as its three 9-row subtrees (top-level child c
// at rows 9c .. 9c + 8), one after the other, so that at most 9 rows are held at once
// (a pass of three steps over all 27 rows would need ~80 VGPRs at three labels per
// lane and lower the occupancy of every pass of the launch).  Same folds, in the same
// order, as blur_pass_vertex<3>.
template <int R>
__device__ __forceinline__ void blur_pass_vertex3(const float* __restrict__ src, float* __restrict__ dst,
                                                  const int* __restrict__ e, int v, int ldv, int k0, int L, int zr) {
    float s[3][R];
#pragma unroll 1
    for (int c = 0; c < 3; c++) {
        int idx[9];
#pragma unroll
        for (int i = 0; i < 9; i++) idx[i] = e[9 * c + i];
        float x[9][R];
#pragma unroll
        for (int i = 0; i < 9; i++) {
            if (idx[i] != zr) {
                load_piece<R>(src + (size_t)idx[i] * ldv, k0, L, x[i]);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) x[i][r] = 0.0f;
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float a = blur_fold(x[0][r], x[1][r], x[2][r]);
            const float b = blur_fold(x[3][r], x[4][r], x[5][r]);
            const float d = blur_fold(x[6][r], x[7][r], x[8][r]);
            const float t = blur_fold(a, b, d);
            // (a runtime index into s would put it in scratch memory)
            if (c == 0) s[0][r] = t;
            else if (c == 1) s[1][r] = t;
            else s[2][r] = t;
        }
    }
    float out[R];
#pragma unroll
    for (int r = 0; r < R; r++) out[r] = blur_fold(s[0][r], s[1][r], s[2][r]);
    store_piece<R>(dst + (size_t)v * ldv, k0, L, out);
}

template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_blur_pass(const float* __restrict__ src, float* __restrict__ dst, int v0,
                                                    int Vp, int ldv, int L, int S, int per, PassTable pt) {
    LaneGroup<WAVE, R> G(L, S, per, nullptr);
    const int v = v0 + G.item();  // vertices [v0, Vp)
    if (!G.on || v >= Vp) return;
    const int k0 = G.k0;
    int vb = pt.t[0].vbase, steps = pt.t[0].steps;
    const int* tab = pt.t[0].tab;
#pragma unroll
    for (int k = 1; k < DCRF_MAX_TERMS; k++) {
        if (k < pt.nterms && v >= pt.t[k].vbase) {
            vb = pt.t[k].vbase;
            tab = pt.t[k].tab;
            steps = pt.t[k].steps;
        }
    }
    const size_t lv = (size_t)(v - vb);
    if (steps == 2)
        blur_pass_vertex<2, R>(src, dst, tab + lv * 9, v, ldv, k0, L, pt.zero_row);
    else if (steps == 3)
        blur_pass_vertex3<R>(src, dst, tab + lv * 27, v, ldv, k0, L, pt.zero_row);
    else
        blur_pass_vertex<1, R>(src, dst, tab + lv * 3, v, ldv, k0, L, pt.zero_row);
}

// Pass table of one term: entry (i_K, .., i_1) of vertex v at v*3^K + sum i_j 3^(j-1),
// expanding along axis ax[j-1] at level j (ax[K-1] first, the pass's last step).
__global__ void k_pass_table(const int2* __restrict__ nbr, int V, int vbase, int zero_row, int K, int ax0, int ax1,
                             int ax2, int* __restrict__ tab) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int ax[3] = {ax0, ax1, ax2};
    int cur[27], nxt[27];
    cur[0] = vbase + v;
    int n = 1;
    for (int lev = K - 1; lev >= 0; lev--) {
        // expand every row of the current level along axis ax[lev]: (u, n1(u), n2(u))
        for (int i = 0; i < n; i++) {
            const int u = cur[i];
            const int2 nb = u == zero_row ? make_int2(zero_row, zero_row) : nbr[(size_t)ax[lev] * V + (u - vbase)];
            nxt[3 * i] = u;
            nxt[3 * i + 1] = nb.x;
            nxt[3 * i + 2] = nb.y;
        }
        n *= 3;
        for (int i = 0; i < n; i++) cur[i] = nxt[i];
    }
    for (int i = 0; i < n; i++) tab[(size_t)v * n + i] = cur[i];
}

// Mean-field update of one pixel: E = -U - sum_k C_k(w_k slice_k), Q = softmax(-E)
// (densecrf.cpp:122-129, with the per-term weight w_k = 1 unless set).
template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_meanfield(const float* __restrict__ U, float* __restrict__ Q, int N,
                                                    int L, int S, int per, SliceTable st, float* __restrict__ tmp1,
                                                    float* __restrict__ tmp2) {
    extern __shared__ float sh[];
    float* shc = sh + blockDim.x;  // compat region after the group scratch
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    float e[R], last[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        e[r] = 0.0f;
        last[r] = 0.0f;
    }
    if (active && U) {
        load_piece<R>(U + (size_t)p * L, k0, L, e);
#pragma unroll
        for (int r = 0; r < R; r++) e[r] = -e[r];
    }
#pragma unroll
    for (int k = 0; k < DCRF_MAX_TERMS; k++) {
        if (k < st.nterms) {
            const SliceTerm& T = st.t[k];
            float m[R], c[R];
            if (active) {
                slice_piece<R>(T, p, k0, L, m);
                if (T.mout) store_piece<R>(T.mout + (size_t)p * L, k0, L, m);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) m[r] = 0.0f;
            }
            if (T.kw) {
                const float kw = T.kw[0];
#pragma unroll
                for (int r = 0; r < R; r++) m[r] = m[r] * kw;
            }
            compat_piece<WAVE, R>(T, m, c, G, shc);
#pragma unroll
            for (int r = 0; r < R; r++) {
                e[r] = e[r] - c[r];
                last[r] = c[r];
            }
        }
    }
    const float mx = G.all_max(e);
    float x[R];
#pragma unroll
    for (int r = 0; r < R; r++) x[r] = expf(e[r] - mx);
    const float s = G.all_sum(x);
    if (active) {
        float q[R];
#pragma unroll
        for (int r = 0; r < R; r++) q[r] = x[r] / s;
        store_piece<R>(Q + (size_t)p * L, k0, L, q);
        if (tmp1) store_piece<R>(tmp1 + (size_t)p * L, k0, L, e);
        if (tmp2) store_piece<R>(tmp2 + (size_t)p * L, k0, L, last);
    }
}

// Term k's message in a fused pixel update: the weighted sum of its gathered rows (entry
// order, as slice_piece) into raw, and raw times the term weight into m.  The caller
// stores raw to the term's `mout` (as k_meanfield does) only after its last read of
// the terms' weights: once a kernel may have stored to global memory, the compiler
// reads those weights through the vector memory path and waits for every load in
// flight (the row gathers among them).
template <int R, int D0, int D1>
__device__ __forceinline__ void term_message(const SliceTerm& T, int k, const int2 (&e)[D0 + D1],
                                             const float (&x)[D0 + D1][R], float (&raw)[R], float (&m)[R]) {
    constexpr int DS = D0 + D1;
    const int jb = k ? D0 : 0, je = k ? DS : D0;
#pragma unroll
    for (int r = 0; r < R; r++) raw[r] = 0.0f;
#pragma unroll
    for (int j = jb; j < je; j++) {
        const float w = __int_as_float(e[j].y);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float prod = w * x[j][r];
            raw[r] = raw[r] + prod;
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) m[r] = raw[r];
    if (T.kw) {
        const float kw = T.kw[0];
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = m[r] * kw;
    }
}

// The compatibility transforms c[k] of the pixel's NT term messages.  ROWS: the Matrix
// terms whose weights have an LDS copy (16-byte aligned rows) first write the group's
// message vector into the wave's message region `msg` (NT rows of (L + 3) & ~3 floats
// per group), then read it back whole with matrix_piece_rows: no cross-lane shuffles.
// The other terms go through compat_piece.  Same results either way.
template <int R, int NT, bool ROWS>
__device__ __forceinline__ void compat_terms(LaneGroup<true, R>& G, const SliceTable& st, const float (&mk)[NT][R],
                                             float (&c)[NT][R], const float* shc, float* msg) {
    const int ldm = (G.L + 3) & ~3;
    if (ROWS) {
        bool any = false;
#pragma unroll
        for (int k = 0; k < NT; k++)
            if (st.t[k].ckind == DCRF_COMPAT_MATRIX && st.t[k].cw_lds >= 0) {
                any = true;
                if (G.on) {  // idle lanes shadow group 0: they must not write its row
                    float* row = msg + ((size_t)k * G.per + G.g) * ldm + G.k0;
#pragma unroll
                    for (int r = 0; r < R; r++) row[r] = mk[k][r];
                }
            }
        if (any) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
#pragma unroll
    for (int k = 0; k < NT; k++) {
        const SliceTerm& T = st.t[k];
        if (ROWS && T.ckind == DCRF_COMPAT_MATRIX && T.cw_lds >= 0)
            matrix_piece_rows<R>(shc + T.cw_lds, T.ldw, msg + ((size_t)k * G.per + G.g) * ldm, G.L, G.k0, c[k]);
        else
            compat_piece<true, R>(T, mk[k], c[k], G, shc);
    }
}

// The per-pixel tail of the fused mean-field kernels: the terms' compatibility
// messages subtracted from -U, then the softmax over the lane group, stored when
// `active`.
template <int R, int D0, int D1, bool ROWS>
__device__ __forceinline__ void finish_pixel(LaneGroup<true, R>& G, const SliceTable& st, const int2 (&e)[D0 + D1],
                                             const float (&x)[D0 + D1][R], const float (&u)[R], bool active,
                                             float* __restrict__ qrow, const float* shc, int p, float* msg) {
    constexpr int NT = D1 ? 2 : 1;
    float raw[NT][R], mk[NT][R], c[NT][R];
#pragma unroll
    for (int k = 0; k < NT; k++) term_message<R, D0, D1>(st.t[k], k, e, x, raw[k], mk[k]);
    compat_terms<R, NT, ROWS>(G, st, mk, c, shc, msg);
    float en[R];
#pragma unroll
    for (int r = 0; r < R; r++) en[r] = -u[r];
#pragma unroll
    for (int k = 0; k < NT; k++)
#pragma unroll
        for (int r = 0; r < R; r++) en[r] = en[r] - c[k][r];
    const float mx = G.all_max(en);
    float y[R];
#pragma unroll
    for (int r = 0; r < R; r++) y[r] = expf(en[r] - mx);
    const float sum = G.all_sum(y);
    if (active) {
        float q[R];
#pragma unroll
        for (int r = 0; r < R; r++) q[r] = y[r] / sum;
        store_piece<R>(qrow, G.k0, G.L, q);
#pragma unroll
        for (int k = 0; k < NT; k++)
            if (st.t[k].mout) store_piece<R>(st.t[k].mout + (size_t)p * G.L, G.k0, G.L, raw[k]);
    }
}

// One (vertex, weight bits) pair by a nontemporal 8-byte load (x in the low word).
__device__ __forceinline__ int2 load_pair_nt(const int2* p) {
    const long long v = __builtin_nontemporal_load(reinterpret_cast<const long long*>(p));
    return make_int2((int)(unsigned)(v & 0xffffffffll), (int)(v >> 32));
}

// Stage the slice pairs of the wave's pixels (consecutive: one contiguous run of
// per * DS pairs, 64 pairs per load instruction) into the wave's LDS slice wp; the
// wave's groups then read their own.  `p` is the group's pixel.
template <int R, int DS>
__device__ __forceinline__ void stage_wave_pairs(const LaneGroup<true, R>& G, int p, int N,
                                                 const int2* __restrict__ slc, int2* wp) {
    const int per = G.per;
    const int pw0 = p - G.g;  // the wave's first pixel
    const int np = max(0, min(per, N - pw0)) * DS;
    const int lane = threadIdx.x & 63;
    // two loads per lane per round (clamped, unconditional), then the stores.  The pairs
    // are read once per iteration: nontemporal loads keep them from displacing the Q rows
    // this kernel writes, which the next splat gathers (at 1024x768x21 the splat went from
    // 72.5 to 58.3 us; nontemporal loads of the splat's own entries or of the unary row
    // measured slower)
    const size_t base = (size_t)pw0 * DS;
    const int last = max(np - 1, 0);
    for (int i0 = 0; i0 < np; i0 += 128) {
        const int2 t0 = load_pair_nt(slc + base + min(i0 + lane, last));
        const int2 t1 = load_pair_nt(slc + base + min(i0 + 64 + lane, last));
        if (i0 + lane < np) wp[i0 + lane] = t0;
        if (i0 + 64 + lane < np) wp[i0 + 64 + lane] = t1;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Mean-field update for models of one or two terms with d+1 = D0 (and D1) lattice
// entries per pixel: slc holds every pixel's (vertex, weight bits) pairs of both terms
// side by side, so all D0 + D1 vertex rows are gathered after one dependent load.  Same
// arithmetic, in the same order, as k_meanfield.
template <int R, int D0, int D1, bool CWL>
__global__ void __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(8)))  // <= 64 VGPRs
k_meanfield_fused(const float* __restrict__ U, float* __restrict__ Q,
                                                                int ldq, int N, int L, int S, int per,
                                                                SliceTable st, const int2* __restrict__ slc,
                                                                int cwf, int bw) {
    constexpr int DS = D0 + D1;
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    const int pc = active ? p : 0;  // idle lanes read pixel 0 (in bounds; results unused)
    extern __shared__ __align__(16) int2 pairs_all[];
    int2* wp = pairs_all + (threadIdx.x >> 6) * (per * DS);
    // CWL: some Matrix term has an LDS copy of its weights (staging costs a barrier, so
    // it is compiled in only then)
    float* shc = reinterpret_cast<float*>(pairs_all + (blockDim.x >> 6) * (per * DS));
    // the unary row does not depend on the slice pairs: its load is in flight while they
    // are staged (not at four labels per lane, where the longer-lived row would spill)
    constexpr bool kEarlyU = R <= 3;
    float u[R];
    auto load_u = [&]() {
        if (U) {
            load_piece<R>(U + (size_t)pc * L, k0, L, u);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) u[r] = 0.0f;
        }
    };
    if (kEarlyU) load_u();
    if (CWL) stage_compat(st, L, shc);
    stage_wave_pairs<R, DS>(G, p, N, slc, wp);
    int2 e[DS];
#pragma unroll
    for (int j = 0; j < DS; j++) e[j] = active ? wp[G.g * DS + j] : make_int2(0, 0);
    if (!kEarlyU) load_u();
    float x[DS][R];
#pragma unroll
    for (int j = 0; j < DS; j++) {
        const SliceTerm& T = j < D0 ? st.t[0] : st.t[1];
        load_piece<R>(T.vals + (size_t)e[j].x * T.ldv, k0, L, x[j]);
    }
    // CWL: the wave's message rows follow the compat region
    float* msg = shc + cwf + (size_t)(threadIdx.x >> 6) * (D1 ? 2 : 1) * per * ((L + 3) & ~3);
    // bw > 0: Q is written in 8 x 8 pixel blocks of an image bw pixels wide (q_block_pos)
    const int qpos = bw > 0 ? q_block_pos(pc, bw) : p;
    finish_pixel<R, D0, D1, CWL>(G, st, e, x, u, active, Q + (size_t)qpos * ldq, shc, p, msg);
}

// The adjoint step of k_adjoint (adjoint.hip) over the same interleaved slice pairs:
// g = sum_k C_k(w_k slice_k(filtered b)), b = sumAndNormalize(Q .* g, Q), dU += b, and
// the terms' statistic rows (SliceTerm::gstat).  The pairs carry the transposed filter's
// slice weights.  Same arithmetic, in the same order, as k_adjoint.
template <int R, int D0, int D1, bool CWL>
__global__ void __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_adjoint_fused(const float* __restrict__ Q, float* __restrict__ b, float* __restrict__ dU, int N, int L, int S,
                int per, SliceTable st, const int2* __restrict__ slc, int cwf) {
    constexpr int DS = D0 + D1;
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    const int pc = active ? p : 0;
    extern __shared__ __align__(16) int2 pairs_all[];
    int2* wp = pairs_all + (threadIdx.x >> 6) * (per * DS);
    float* shc = reinterpret_cast<float*>(pairs_all + (blockDim.x >> 6) * (per * DS));
    float q[R];
    load_piece<R>(Q + (size_t)pc * L, k0, L, q);  // in flight while the pairs are staged
    if (CWL) stage_compat(st, L, shc);
    stage_wave_pairs<R, DS>(G, p, N, slc, wp);
    int2 e[DS];
#pragma unroll
    for (int j = 0; j < DS; j++) e[j] = active ? wp[G.g * DS + j] : make_int2(0, 0);
    float x[DS][R];
#pragma unroll
    for (int j = 0; j < DS; j++) {
        const SliceTerm& T = j < D0 ? st.t[0] : st.t[1];
        load_piece<R>(T.vals + (size_t)e[j].x * T.ldv, k0, L, x[j]);
    }
    constexpr int NT = D1 ? 2 : 1;
    float raw[NT][R], mk[NT][R], c[NT][R];  // (the adjoint stores no raw messages)
#pragma unroll
    for (int k = 0; k < NT; k++) term_message<R, D0, D1>(st.t[k], k, e, x, raw[k], mk[k]);
    // statistic region after the message rows (present with a compat region)
    float* sst = shc + cwf + (cwf ? (blockDim.x >> 6) * NT * per * ((L + 3) & ~3) : 0);
#pragma unroll
    for (int k = 0; k < NT; k++)
        if (st.t[k].gslot >= 0) {
            float v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = active ? raw[k][r] * q[r] : 0.0f;
            stat_put<R>(sst, st.t[k].gslot, v);
        }
    float* msg = shc + cwf + (size_t)(threadIdx.x >> 6) * NT * per * ((L + 3) & ~3);
    compat_terms<R, NT, CWL>(G, st, mk, c, shc, msg);
    float gsum[R];
#pragma unroll
    for (int r = 0; r < R; r++) gsum[r] = 0.0f;
#pragma unroll
    for (int k = 0; k < NT; k++)
#pragma unroll
        for (int r = 0; r < R; r++) gsum[r] = gsum[r] + c[k][r];
    float bn[R];
    sum_normalize<true, R>(q, gsum, bn, G);
    if (active) {
        float du[R];
        load_piece<R>(dU + (size_t)p * L, k0, L, du);
#pragma unroll
        for (int r = 0; r < R; r++) du[r] = du[r] + bn[r];
        store_piece<R>(b + (size_t)p * L, k0, L, bn);
        store_piece<R>(dU + (size_t)p * L, k0, L, du);
    }
    if (st.nstat) stat_finish<true, R, NT>(st, G, sst);  // uniform condition
}

// The same update over pixel runs (2-D rectangles whose rows hold whole waves of
// pixels), one workgroup per run: the run's distinct vertex rows (a few hundred, against
// P*(D0+D1) gathers) are staged in LDS once, then every pixel's slice reads them from
// there (rs.loc holds the pairs with the vertex replaced by its index in the run's
// list, in run-major pixel order).
template <int R, int D0, int D1>
__global__ void __launch_bounds__(kWaveBlock) k_meanfield_runs(const float* __restrict__ U, float* __restrict__ Q,
                                                               int N, int L, int S, int per, SliceTable st,
                                                               RunSliceDev rs) {
    (void)N;  // the runs cover the pixels exactly (rs.pix_base)
    constexpr int DS = D0 + D1;
    extern __shared__ float lds[];
    float* tq = lds;  // rs.rows_lds floats: the run's rows
    int2* tp = reinterpret_cast<int2*>(lds + rs.rows_lds);  // rs.P * DS pairs: the run's slice pairs
    float* shc = reinterpret_cast<float*>(tp + (size_t)rs.P * DS);
    stage_compat(st, L, shc);
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int k0 = G.k0;
    const int run = xcd_block(blockIdx.x, gridDim.x);
    const int k0r = rs.pix_base[run], np = rs.pix_base[run + 1] - k0r;  // run-major positions
    const int u0 = rs.base[run], nu = rs.base[run + 1] - u0;
    const int items = (blockDim.x >> 6) * per;
    const int gi = (threadIdx.x >> 6) * per + G.g;
    // stage the rows (group gi takes rows gi, gi + items, ...) and the pairs (all threads,
    // consecutive words): loads first, then stores
    for (int q0 = 0; q0 < nu; q0 += 4 * items) {
        float xs[4][R];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int q = min(q0 + b * items + gi, nu - 1);
            const int tag = rs.rows[u0 + q];
            load_piece<R>(((tag & 1) ? rs.v1 : rs.v0) + (size_t)(tag >> 1) * st.t[0].ldv, k0, L, xs[b]);
        }
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int q = q0 + b * items + gi;
            if (G.on && q < nu) store_piece<R>(tq + (size_t)q * L, k0, L, xs[b]);
        }
    }
    {
        const int npairs = np * DS;
        const int2* src = rs.loc + (size_t)k0r * DS;
        for (int i0 = 0; i0 < npairs; i0 += 4 * (int)blockDim.x) {
            int2 t[4];
#pragma unroll
            for (int b = 0; b < 4; b++) t[b] = src[min(i0 + b * (int)blockDim.x + (int)threadIdx.x, npairs - 1)];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int i = i0 + b * (int)blockDim.x + (int)threadIdx.x;
                if (i < npairs) tp[i] = t[b];
            }
        }
    }
    __syncthreads();
    // pixel passes: each pass's unary rows are loaded one pass ahead
    auto pixel = [&](int b0) { return rs.rpix[b0 + gi < np && G.on ? k0r + b0 + gi : k0r]; };
    float un[R];
    int pn = pixel(0);
    if (U) {
        load_piece<R>(U + (size_t)pn * L, k0, L, un);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) un[r] = 0.0f;
    }
    for (int b0 = 0; b0 < np; b0 += items) {
        const bool active = G.on && b0 + gi < np;
        const int p = pn;
        float u[R];
#pragma unroll
        for (int r = 0; r < R; r++) u[r] = un[r];
        if (b0 + items < np) {
            pn = pixel(b0 + items);
            if (U) load_piece<R>(U + (size_t)pn * L, k0, L, un);
        }
        int2 e[DS];
#pragma unroll
        for (int j = 0; j < DS; j++) e[j] = active ? tp[(b0 + gi) * DS + j] : make_int2(0, 0);
        float x[DS][R];
#pragma unroll
        for (int j = 0; j < DS; j++) load_piece<R>(tq + (size_t)e[j].x * L, k0, L, x[j]);
        finish_pixel<R, D0, D1, false>(G, st, e, x, u, active, Q + (size_t)p * L, shc, p, nullptr);
    }
}

// slc[p][off + j] = (vtx[p][j], bits of sw[p][j]) for j < d1; rows of ds pairs.
__global__ void k_interleave_slice(const int* __restrict__ vtx, const float* __restrict__ sw, int N, int d1, int ds,
                                   int off, int2* __restrict__ slc) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * d1) return;
    const size_t p = i / d1, j = i - p * d1;
    slc[p * ds + off + j] = make_int2(vtx[i], __float_as_int(sw[i]));
}

// out = compat_which(slice_which) (ckind < 0: the plain slice).
template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_slice_term(float* __restrict__ out, int N, int L, int S, int per,
                                                     SliceTable st, int which) {
    extern __shared__ float sh[];
    float* shc = sh + blockDim.x;
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    SliceTerm T = st.t[0];
#pragma unroll
    for (int k = 1; k < DCRF_MAX_TERMS; k++)
        if (k == which) T = st.t[k];
    float m[R], c[R];
    if (active) {
        slice_piece<R>(T, p, G.k0, L, m);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = 0.0f;
    }
    if (T.kw) {
        const float kw = T.kw[0];
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = m[r] * kw;
    }
    if (T.ckind >= 0) {
        compat_piece<WAVE, R>(T, m, c, G, shc);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = m[r];
    }
    if (active) store_piece<R>(out + (size_t)p * L, G.k0, L, c);
}

// KL divergence terms per pixel, reduced per block in double in a fixed order
// (densecrf.cpp:214-235: q log max(q,1e-20) + u q + sum_k q * pairwise_k(Q)).
template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_kl(const float* __restrict__ Q, const float* __restrict__ U, int N, int L,
                                             int S, int per, SliceTable st, double* __restrict__ partial) {
    extern __shared__ float sh[];
    __shared__ double red[1024];
    float* shc = sh + blockDim.x;
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    float q[R], u[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        q[r] = 0.0f;
        u[r] = 0.0f;
    }
    double acc = 0.0;
    if (active) {
        load_piece<R>(Q + (size_t)p * L, k0, L, q);
        if (U) load_piece<R>(U + (size_t)p * L, k0, L, u);
#pragma unroll
        for (int r = 0; r < R; r++)
            if (G.has(r)) {
                const float qc = q[r] > 1e-20f ? q[r] : 1e-20f;
                acc += (double)q[r] * log((double)qc);
                if (U) acc += (double)(u[r] * q[r]);
            }
    }
#pragma unroll
    for (int k = 0; k < DCRF_MAX_TERMS; k++) {
        if (k < st.nterms) {
            const SliceTerm& T = st.t[k];
            float m[R], c[R];
            if (active) {
                slice_piece<R>(T, p, k0, L, m);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) m[r] = 0.0f;
            }
            if (T.kw) {
                const float kw = T.kw[0];
#pragma unroll
                for (int r = 0; r < R; r++) m[r] = m[r] * kw;
            }
            compat_piece<WAVE, R>(T, m, c, G, shc);
            if (active)
#pragma unroll
                for (int r = 0; r < R; r++)
                    if (G.has(r)) acc += (double)(q[r] * c[r]);
        }
    }
    const int t = threadIdx.x;
    red[t] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) partial[blockIdx.x] = red[0];
}

// DenseKernel::initLattice normalisers (pairwise.cpp:54-61), in double like the reference.
__global__ void k_norm_finish(float* norm, int N, int ntype) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double x = (double)norm[i] + 1e-20;
    if (ntype == DCRF_NORMALIZE_SYMMETRIC) norm[i] = (float)(1.0 / sqrt(x));
    else if (ntype == DCRF_NORMALIZE_BEFORE || ntype == DCRF_NORMALIZE_AFTER) norm[i] = (float)(1.0 / x);
}

__global__ void k_remap_nbr(const int* __restrict__ src, int2* __restrict__ dst, int n, int vbase,
                            int zero_row) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a = src[2 * (size_t)i], b = src[2 * (size_t)i + 1];
    dst[i] = make_int2(a >= 0 ? a + vbase : zero_row, b >= 0 ? b + vbase : zero_row);
}

__global__ void k_entry_gather(const int* __restrict__ csr_e, int E, int d1, const float* __restrict__ bary,
                               const float* __restrict__ norm, int use_norm, const int* __restrict__ tile_ptr,
                               const int* __restrict__ base, int n_tiles, TileGeom g, int local, int pix_off,
                               int2* __restrict__ ent) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= E) return;
    const int e = csr_e[j];
    const int p = e / d1;
    const int pos = base ? base[last_le(tile_ptr, n_tiles, j)] + j : j;
    const float b = bary[e];
    const float w = use_norm ? b * norm[p] : b;
    ent[pos] = make_int2((local ? tile_local(g, p) : p) + pix_off, __float_as_int(w));
}

__global__ void k_slice_weights(const int* __restrict__ vtx, const float* __restrict__ bary,
                                const float* __restrict__ norm, int use_norm, float alpha, int N, int d1,
                                int vbase, int* __restrict__ gvtx, float* __restrict__ sw) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * d1) return;
    const int p = (int)(i / d1);
    gvtx[i] = vtx[i] + vbase;
    const float wa = bary[i] * alpha;
    sw[i] = use_norm ? wa * norm[p] : wa;
}

// DenseCRF::currentMap (densecrf.cpp:202-211): out[p] = the first label of the largest
// Q[p][l] (a later label replaces it only when strictly larger, as Eigen's maxCoeff
// visitor does), one thread per pixel.
__global__ void k_argmax(const float* __restrict__ Q, int ldq, int N, int L, int16_t* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const float* row = Q + (size_t)p * ldq;
    float best = row[0];
    int arg = 0;
    for (int l = 1; l < L; l++) {
        const float v = row[l];
        if (v > best) {
            best = v;
            arg = l;
        }
    }
    out[p] = (int16_t)arg;
}

// out[c][r] = in[r][c] for an rows x cols row-major input, 32x32 tiles through LDS.
__global__ void k_transpose(const float* __restrict__ in, float* __restrict__ out, int rows, int cols) {
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    for (int k = 0; k < 32; k += 8) {
        const int r = by + ty + k, c = bx + tx;
        if (r < rows && c < cols) tile[ty + k][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = 0; k < 32; k += 8) {
        const int c = bx + ty + k, r = by + tx;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][ty + k];
    }
}

}  // namespace

void launch_fill(float* p, float v, size_t n, hipStream_t s) {
    if (!n) return;
    k_fill<<<blocks_for(n), kThreads, 0, s>>>(p, v, n);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_argmax(const float* Q, int ldq, int N, int L, int16_t* out, hipStream_t s) {
    if (N <= 0) return;
    k_argmax<<<blocks_for(N), kThreads, 0, s>>>(Q, ldq, N, L, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_transpose(const float* in, float* out, int rows, int cols, hipStream_t s) {
    if (!rows || !cols) return;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    k_transpose<<<grid, 256, 0, s>>>(in, out, rows, cols);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_splat(const Chunk* chunks, int n_chunks, const int2* ent, const float* in, int ldi, float* vals,
                  int ldv, int L, hipStream_t s) {
    if (!n_chunks) return;
    const LaunchShape sh = shape_for(L);
    const size_t shm = sh.wave ? (size_t)(kWaveBlock / 64) * (kWin + kBatch) * sizeof(int2) : 0;
    DCRF_DISPATCH(sh, k_splat, blocks_for(n_chunks, sh.items), shm, s, chunks, n_chunks, ent, in, ldi, vals, ldv, L,
                  sh.S, sh.per);
}

size_t tiled_splat_lds(int L, int side, int e_max, int s_max, size_t lds_cap) {
    const LaunchShape sh = shape_for(L);
    if (!sh.wave) return 0;
    const size_t words = (((size_t)side * side * L + 1) & ~(size_t)1) + 2 * (size_t)e_max + 3 * (size_t)s_max;
    const size_t bytes = words * sizeof(float);
    return bytes <= lds_cap ? bytes : 0;
}

// Workgroup size of the tiled splat: 1024 threads (16 waves of lane groups) once a tile
// takes more than a quarter of the workgroup LDS limit, else 256.
int tiled_splat_block(size_t lds, size_t lds_cap) { return lds * 4 > lds_cap ? 1024 : kWaveBlock; }

void launch_splat_tiled(const TiledSplat& ts, const int2* ent, const float* in, float* vals, int ldv, int L,
                        size_t lds_cap, hipStream_t s) {
    if (!ts.n_tiles) return;
    const LaunchShape sh = shape_for(L);
    const size_t lds = tiled_splat_lds(L, ts.geom.side, ts.e_max, ts.s_max, lds_cap);
    if (!lds) throw Error{DCRF_ERR_VALUE, "tiled splat does not fit the LDS of a block"};
    const int q_max = (ts.geom.side * ts.geom.side * L + 1) & ~1;  // the entry pairs start 8-byte aligned
    const int block = tiled_splat_block(lds, lds_cap);
    // dynamic LDS beyond the default launch limit must be granted per kernel
    static size_t granted[5] = {};
    auto grant = [&](const void* fn, int r) {
        if (lds > granted[r]) {
            DCRF_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            granted[r] = lds;
        }
    };
    switch (sh.R) {
        case 1: grant(reinterpret_cast<const void*>(&k_splat_tiled<1>), 1); break;
        case 2: grant(reinterpret_cast<const void*>(&k_splat_tiled<2>), 2); break;
        case 3: grant(reinterpret_cast<const void*>(&k_splat_tiled<3>), 3); break;
        default: grant(reinterpret_cast<const void*>(&k_splat_tiled<4>), 4); break;
    }
    switch (sh.R) {
        case 1: k_splat_tiled<1><<<ts.n_tiles, block, lds, s>>>(ts, ent, in, vals, ldv, L, sh.S, sh.per, q_max); break;
        case 2: k_splat_tiled<2><<<ts.n_tiles, block, lds, s>>>(ts, ent, in, vals, ldv, L, sh.S, sh.per, q_max); break;
        case 3: k_splat_tiled<3><<<ts.n_tiles, block, lds, s>>>(ts, ent, in, vals, ldv, L, sh.S, sh.per, q_max); break;
        default: k_splat_tiled<4><<<ts.n_tiles, block, lds, s>>>(ts, ent, in, vals, ldv, L, sh.S, sh.per, q_max); break;
    }
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_combine(const Combine* c, int n, float* vals, int ldv, int L, hipStream_t s) {
    if (!n) return;
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH(sh, k_combine, blocks_for(n, sh.items), 0, s, c, n, vals, ldv, L, sh.S, sh.per);
}

void launch_combine_waves(const Combine* c, int n, float* vals, int ldv, int L, hipStream_t s) {
    if (!n) return;
    const LaunchShape sh = shape_for(L);
    if (!sh.wave) throw Error{DCRF_ERR_STATE, "one-launch combine: wave-mode shapes only"};
    const size_t lds = sizeof(float) * (kWaveBlock / 64) * sh.per * L;
    const unsigned grid = blocks_for(n, sh.items);
    switch (sh.R) {
        case 1: k_combine_waves<1><<<grid, kWaveBlock, lds, s>>>(c, n, vals, ldv, L, sh.S, sh.per); break;
        case 2: k_combine_waves<2><<<grid, kWaveBlock, lds, s>>>(c, n, vals, ldv, L, sh.S, sh.per); break;
        case 3: k_combine_waves<3><<<grid, kWaveBlock, lds, s>>>(c, n, vals, ldv, L, sh.S, sh.per); break;
        default: k_combine_waves<4><<<grid, kWaveBlock, lds, s>>>(c, n, vals, ldv, L, sh.S, sh.per); break;
    }
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_blur(const float* src, float* dst, int v0, int Vp, int ldv, int L, const BlurTable& bt, int step,
                 int reverse, hipStream_t s) {
    if (Vp <= v0) return;
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH(sh, k_blur, blocks_for(Vp - v0, sh.items), 0, s, src, dst, v0, Vp, ldv, L, sh.S, sh.per, bt, step,
                  reverse);
}

void launch_blur_pass(const float* src, float* dst, int v0, int Vp, int ldv, int L, const PassTable& pt,
                      hipStream_t s, int R) {
    if (Vp <= v0) return;
    // R: labels per lane (0 = the default for L); the blur is elementwise per label, so
    // any lane split gives the same bits
    const LaunchShape sh = R > 0 && L % R == 0 ? shape_for_r(L, R) : shape_for(L);
    DCRF_DISPATCH(sh, k_blur_pass, blocks_for(Vp - v0, sh.items), 0, s, src, dst, v0, Vp, ldv, L, sh.S, sh.per, pt);
}

void launch_pass_table(const int2* nbr, int V, int vbase, int zero_row, int K, const int* axes, int* tab,
                       hipStream_t s) {
    if (V <= 0) return;
    if (K < 1 || K > 3) throw Error{DCRF_ERR_STATE, "blur pass: 1 to 3 steps"};
    k_pass_table<<<blocks_for(V), kThreads, 0, s>>>(nbr, V, vbase, zero_row, K, axes[0], K > 1 ? axes[1] : 0,
                                                    K > 2 ? axes[2] : 0, tab);
    DCRF_HIP_CHECK(hipGetLastError());
}

bool fused_meanfield_ok(int L, int nterms, int d1_0, int d1_1) {
    if (!shape_for(L).wave) return false;
    if (nterms == 1) return d1_0 == 3 || d1_0 == 6;
    if (nterms == 2) return (d1_0 == 3 && d1_1 == 6) || (d1_0 == 6 && d1_1 == 3);
    return false;
}

void launch_interleave_slice(const int* vtx, const float* sw, int N, int d1, int ds, int off, int2* slc,
                             hipStream_t s) {
    if (!N) return;
    k_interleave_slice<<<blocks_for((size_t)N * d1), kThreads, 0, s>>>(vtx, sw, N, d1, ds, off, slc);
    DCRF_HIP_CHECK(hipGetLastError());
}

// dynamic LDS of the fused kernels: the wave's slice pairs, the compat region (cwf
// floats), with a compat region the message rows (NT per group, (L + 3) & ~3 floats),
// and the adjoint's statistic region (stat_lds_floats)
inline size_t fused_lds_bytes(const LaunchShape& sh, int ds, int nt, int L, int cwf, int nstat = 0) {
    const size_t waves = kWaveBlock / 64;
    return waves * sh.per * ds * sizeof(int2) + (size_t)cwf * sizeof(float) +
           (cwf ? waves * nt * sh.per * (size_t)((L + 3) & ~3) * sizeof(float) : 0) +
           stat_lds_floats(true, nstat, L, kWaveBlock, sh.R) * sizeof(float);
}

#define DCRF_FUSED(R_, D0_, D1_)                                                                          \
    (cwf ? k_meanfield_fused<R_, D0_, D1_, true> : k_meanfield_fused<R_, D0_, D1_, false>)<<<                 \
        blocks_for(N, sh.items), kWaveBlock, fused_lds_bytes(sh, D0_ + D1_, D1_ ? 2 : 1, L, cwf), s>>>(       \
        U, Q, ldq, N, L, sh.S, sh.per, st, slc, cwf, bw)

template <int R>
void launch_fused_r(const LaunchShape& sh, const float* U, float* Q, int ldq, int N, int L, const SliceTable& st,
                    const int2* slc, int cwf, int bw, hipStream_t s) {
    const int a = st.t[0].d1, b = st.nterms > 1 ? st.t[1].d1 : 0;
    if (a == 3 && b == 0) DCRF_FUSED(R, 3, 0);
    else if (a == 6 && b == 0) DCRF_FUSED(R, 6, 0);
    else if (a == 3 && b == 6) DCRF_FUSED(R, 3, 6);
    else if (a == 6 && b == 3) DCRF_FUSED(R, 6, 3);
    else throw Error{DCRF_ERR_VALUE, "no fused mean-field kernel for these terms"};
    DCRF_HIP_CHECK(hipGetLastError());
}
#undef DCRF_FUSED

#define DCRF_ADJ_FUSED(R_, D0_, D1_)                                                                      \
    (cwf ? k_adjoint_fused<R_, D0_, D1_, true> : k_adjoint_fused<R_, D0_, D1_, false>)<<<                   \
        blocks_for(N, sh.items), kWaveBlock, fused_lds_bytes(sh, D0_ + D1_, D1_ ? 2 : 1, L, cwf, st.nstat),  \
        s>>>(Q, b, dU, N, L, sh.S, sh.per, st, slc, cwf)

template <int R>
void launch_adj_fused_r(const LaunchShape& sh, const float* Q, float* b, float* dU, int N, int L,
                        const SliceTable& st, const int2* slc, int cwf, hipStream_t s) {
    const int a = st.t[0].d1, c = st.nterms > 1 ? st.t[1].d1 : 0;
    if (a == 3 && c == 0) DCRF_ADJ_FUSED(R, 3, 0);
    else if (a == 6 && c == 0) DCRF_ADJ_FUSED(R, 6, 0);
    else if (a == 3 && c == 6) DCRF_ADJ_FUSED(R, 3, 6);
    else if (a == 6 && c == 3) DCRF_ADJ_FUSED(R, 6, 3);
    else throw Error{DCRF_ERR_VALUE, "no fused adjoint kernel for these terms"};
    DCRF_HIP_CHECK(hipGetLastError());
}
#undef DCRF_ADJ_FUSED

#define DCRF_RUNS(R_, D0_, D1_)                                                                           \
    k_meanfield_runs<R_, D0_, D1_><<<rs.n_runs, kWaveBlock, lds, s>>>(U, Q, N, L, sh.S, sh.per, st, rs)

template <int R>
void launch_runs_r(const LaunchShape& sh, const float* U, float* Q, int N, int L, const SliceTable& st,
                   const RunSliceDev& rs, size_t lds, hipStream_t s) {
    const int a = st.t[0].d1, b = st.nterms > 1 ? st.t[1].d1 : 0;
    if (a == 3 && b == 0) DCRF_RUNS(R, 3, 0);
    else if (a == 6 && b == 0) DCRF_RUNS(R, 6, 0);
    else if (a == 3 && b == 6) DCRF_RUNS(R, 3, 6);
    else if (a == 6 && b == 3) DCRF_RUNS(R, 6, 3);
    else throw Error{DCRF_ERR_VALUE, "no run mean-field kernel for these terms"};
    DCRF_HIP_CHECK(hipGetLastError());
}
#undef DCRF_RUNS

size_t meanfield_runs_lds(int L, int ds, int P, int rows_lds, int cwf) {
    (void)L;
    return (size_t)rows_lds * sizeof(float) + (size_t)P * ds * sizeof(int2) + (size_t)cwf * sizeof(float);
}

void launch_meanfield_runs(const float* U, float* Q, int N, int L, const SliceTable& st_in, const RunSliceDev& rs,
                           hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    if (!fused_meanfield_ok(L, st.nterms, st.t[0].d1, st.nterms > 1 ? st.t[1].d1 : 0))
        throw Error{DCRF_ERR_VALUE, "no run mean-field kernel for these terms"};
    const int ds = st.t[0].d1 + (st.nterms > 1 ? st.t[1].d1 : 0);
    const size_t lds = meanfield_runs_lds(L, ds, rs.P, rs.rows_lds, cwf);
    static size_t granted[5] = {};
    auto grant = [&](const void* fn, int r) {
        if (lds > granted[r]) {
            DCRF_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            granted[r] = lds;
        }
    };
    const int a = st.t[0].d1, b = st.nterms > 1 ? st.t[1].d1 : 0;
    auto fn_for = [&](auto rtag) -> const void* {
        constexpr int R_ = decltype(rtag)::value;
        if (a == 3 && b == 0) return reinterpret_cast<const void*>(&k_meanfield_runs<R_, 3, 0>);
        if (a == 6 && b == 0) return reinterpret_cast<const void*>(&k_meanfield_runs<R_, 6, 0>);
        if (a == 3 && b == 6) return reinterpret_cast<const void*>(&k_meanfield_runs<R_, 3, 6>);
        return reinterpret_cast<const void*>(&k_meanfield_runs<R_, 6, 3>);
    };
    switch (sh.R) {
        case 1: grant(fn_for(std::integral_constant<int, 1>{}), 1); launch_runs_r<1>(sh, U, Q, N, L, st, rs, lds, s); break;
        case 2: grant(fn_for(std::integral_constant<int, 2>{}), 2); launch_runs_r<2>(sh, U, Q, N, L, st, rs, lds, s); break;
        case 3: grant(fn_for(std::integral_constant<int, 3>{}), 3); launch_runs_r<3>(sh, U, Q, N, L, st, rs, lds, s); break;
        default: grant(fn_for(std::integral_constant<int, 4>{}), 4); launch_runs_r<4>(sh, U, Q, N, L, st, rs, lds, s); break;
    }
}

void launch_meanfield_fused(const float* U, float* Q, int ldq, int N, int L, const SliceTable& st_in,
                            const int2* slc, hipStream_t s, int bw) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L, true);  // 16-byte rows for matrix_piece_rows
    const LaunchShape sh = shape_for(L);
    if (!fused_meanfield_ok(L, st.nterms, st.t[0].d1, st.nterms > 1 ? st.t[1].d1 : 0))
        throw Error{DCRF_ERR_VALUE, "no fused mean-field kernel for these terms"};
    switch (sh.R) {
        case 1: launch_fused_r<1>(sh, U, Q, ldq, N, L, st, slc, cwf, bw, s); break;
        case 2: launch_fused_r<2>(sh, U, Q, ldq, N, L, st, slc, cwf, bw, s); break;
        case 3: launch_fused_r<3>(sh, U, Q, ldq, N, L, st, slc, cwf, bw, s); break;
        default: launch_fused_r<4>(sh, U, Q, ldq, N, L, st, slc, cwf, bw, s); break;
    }
}

namespace {
// Entry (pixel, weight) -> (its 8 x 8 block position, weight) for an image W pixels wide.
__global__ void k_entry_block(const int2* __restrict__ ent, size_t E, int W, int2* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < E) {
        const int2 e = ent[i];
        out[i] = make_int2(q_block_pos(e.x, W), e.y);
    }
}
}  // namespace

void launch_entry_block(const int2* ent, size_t E, int W, int2* out, hipStream_t s) {
    if (!E) return;
    k_entry_block<<<blocks_for(E), kThreads, 0, s>>>(ent, E, W, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_adjoint_fused(const float* Q, float* b, float* dU, int N, int L, const SliceTable& st_in,
                          const int2* slc, hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L, true);  // 16-byte rows for matrix_piece_rows
    const LaunchShape sh = shape_for(L);
    if (!fused_meanfield_ok(L, st.nterms, st.t[0].d1, st.nterms > 1 ? st.t[1].d1 : 0))
        throw Error{DCRF_ERR_VALUE, "no fused adjoint kernel for these terms"};
    switch (sh.R) {
        case 1: launch_adj_fused_r<1>(sh, Q, b, dU, N, L, st, slc, cwf, s); break;
        case 2: launch_adj_fused_r<2>(sh, Q, b, dU, N, L, st, slc, cwf, s); break;
        case 3: launch_adj_fused_r<3>(sh, Q, b, dU, N, L, st, slc, cwf, s); break;
        default: launch_adj_fused_r<4>(sh, Q, b, dU, N, L, st, slc, cwf, s); break;
    }
}

int assign_compat_lds(SliceTable& st, int L, bool pad) {
    const int ldw = pad ? (L + 3) & ~3 : L;
    int total = 0;
    for (int k = 0; k < st.nterms; k++) {
        st.t[k].cw_lds = -1;
        st.t[k].ldw = ldw;
    }
    for (int k = 0; k < st.nterms; k++)
        if (st.t[k].ckind == DCRF_COMPAT_MATRIX && total + L * ldw <= kCompatLdsFloats) {
            st.t[k].cw_lds = total;
            total += L * ldw;
        }
    return total;
}

void launch_meanfield(const float* U, float* Q, int N, int L, const SliceTable& st_in, float* tmp1, float* tmp2,
                      hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH_CWL(sh, cwf, k_meanfield, blocks_for(N, sh.items), sizeof(float) * (sh.block + cwf), s, U, Q, N,
                      L, sh.S, sh.per, st, tmp1, tmp2);
}

void launch_softmax_neg(const float* U, float* Q, int N, int L, hipStream_t s) {
    SliceTable st{};
    st.nterms = 0;
    launch_meanfield(U, Q, N, L, st, nullptr, nullptr, s);
}

void launch_slice_term(float* out, int N, int L, const SliceTable& st_in, int which, hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH_CWL(sh, cwf, k_slice_term, blocks_for(N, sh.items), sizeof(float) * (sh.block + cwf), s, out, N,
                      L, sh.S, sh.per, st, which);
}

int launch_kl(const float* Q, const float* U, int N, int L, const SliceTable& st_in, double* partials,
              hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    const int nb = (int)blocks_for(N, sh.items);
    DCRF_DISPATCH_CWL(sh, cwf, k_kl, nb, sizeof(float) * (sh.block + cwf), s, Q, U, N, L, sh.S, sh.per, st,
                      partials);
    return nb;
}

void launch_norm_finish(float* norm, int N, int ntype, hipStream_t s) {
    k_norm_finish<<<blocks_for(N), kThreads, 0, s>>>(norm, N, ntype);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_remap_nbr(const int* src, int2* dst, int n, int vbase, int zero_row, hipStream_t s) {
    if (!n) return;
    k_remap_nbr<<<blocks_for(n), kThreads, 0, s>>>(src, dst, n, vbase, zero_row);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_entry_gather(const int* csr_e, int E, int d1, const float* bary, const float* norm, int use_norm,
                         const int* tile_ptr, const int* base, int n_tiles, const TileGeom* local, int2* ent_out,
                         hipStream_t s, int pix_off) {
    if (!E) return;
    const TileGeom g = local ? *local : TileGeom{};
    k_entry_gather<<<blocks_for(E), kThreads, 0, s>>>(csr_e, E, d1, bary, norm, use_norm, tile_ptr, base, n_tiles, g,
                                                      local ? 1 : 0, pix_off, ent_out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_slice_weights(const int* vtx, const float* bary, const float* norm, int use_norm, float alpha, int N,
                          int d1, int vbase, int* gvtx, float* sw, hipStream_t s) {
    if (!N) return;
    k_slice_weights<<<blocks_for((size_t)N * d1), kThreads, 0, s>>>(vtx, bary, norm, use_norm, alpha, N, d1, vbase,
                                                                    gvtx, sw);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
