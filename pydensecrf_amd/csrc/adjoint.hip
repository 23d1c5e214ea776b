// Adjoint mean-field for gfx950: the backward of E_{t+1} = U + sum_k C_k(w_k K~_k Q_t),
// Q_t = softmax(-E_t), following DenseCRF::gradient (pydensecrf/densecrf/src/
// densecrf.cpp:238-297) with the reference's conventions:
//   b_n   = sumAndNormalize(Q_n .* dL/dQ_n, Q_n)                  (= dL/dE_n, :257-258)
//   for t = n-1 .. 0:
//     G_k  += b_{t+1} (K~_k Q_t)^T         compatibility / kernel-weight statistics
//                                          (labelcompatibility.cpp:57-61, 76-78, 101-108)
//     g_t   = sum_k w_k K~_k^T C_k^T b_{t+1}   transposed filter (reverse blur order,
//                                          pairwise.cpp:179-183, permutohedral.cpp:556)
//     b_t   = sumAndNormalize(Q_t .* g_t, Q_t)                     (:290)
//   dL/dU = sum_t b_t                                              (:261-262, 293-294)
// K~_k Q_t comes from the forward pass's message cache, so the backward runs one
// filter per term per iteration (the reference re-filters, pairwise.cpp:193).
// Reductions run in a fixed order (per-block float partials, then a double sum over
// blocks): results are deterministic.
#include "dcrf_internal.h"
#include "dcrf_device.h"

#pragma clang fp contract(off)

namespace dcrf {
namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(size_t n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_sn_first(const float* __restrict__ Q, const float* __restrict__ dQ,
                                                   float* __restrict__ b, float* __restrict__ dU, int N, int L,
                                                   int S, int per) {
    extern __shared__ float sh[];
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    float q[R], gv[R], bn[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        q[r] = 0.0f;
        gv[r] = 0.0f;
    }
    if (active) {
        load_piece<R>(Q + (size_t)p * L, G.k0, L, q);
        load_piece<R>(dQ + (size_t)p * L, G.k0, L, gv);
    }
    sum_normalize<WAVE, R>(q, gv, bn, G);
    if (active) {
        store_piece<R>(b + (size_t)p * L, G.k0, L, bn);
        store_piece<R>(dU + (size_t)p * L, G.k0, L, bn);
    }
}

template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_adjoint(const float* __restrict__ Q, float* __restrict__ b,
                                                  float* __restrict__ dU, int N, int L, int S, int per,
                                                  SliceTable st) {
    extern __shared__ float sh[];
    float* shc = sh + blockDim.x;
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    float gsum[R];
#pragma unroll
    for (int r = 0; r < R; r++) gsum[r] = 0.0f;
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
            compat_piece<WAVE, R>(T, m, c, G, shc);  // C^T = C (W is symmetric)
#pragma unroll
            for (int r = 0; r < R; r++) gsum[r] = gsum[r] + c[r];
        }
    }
    float q[R], bn[R];
#pragma unroll
    for (int r = 0; r < R; r++) q[r] = 0.0f;
    if (active) load_piece<R>(Q + (size_t)p * L, k0, L, q);
    sum_normalize<WAVE, R>(q, gsum, bn, G);
    if (active) {
        float du[R];
        load_piece<R>(dU + (size_t)p * L, k0, L, du);
#pragma unroll
        for (int r = 0; r < R; r++) du[r] = du[r] + bn[r];
        store_piece<R>(b + (size_t)p * L, k0, L, bn);
        store_piece<R>(dU + (size_t)p * L, k0, L, du);
    }
}

// partial[block][label] = sum over the block's pixels of b .* M (groups stride over the
// pixels; the block's groups are then added in group order per label).  The grid covers
// `iters` iterations of nb blocks each: block i * nb + j takes iteration i's operands
// b + i * bstride and M + i * mstride.
template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_grad_diag(const float* __restrict__ b, const float* __restrict__ M, int N,
                                                    int L, int S, int per, int nb, size_t bstride, size_t mstride,
                                                    float* __restrict__ partial) {
    extern __shared__ float sh[];  // block * R floats
    const int it = blockIdx.x / nb, blk = blockIdx.x - it * nb;
    b += (size_t)it * bstride;
    M += (size_t)it * mstride;
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int items = LaneGroup<WAVE, R>::items_per_block(per);
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; r++) acc[r] = 0.0f;
    if (G.on) {
        for (int p = G.item_in(blk); p < N; p += nb * items) {
            float x[R], y[R];
            load_piece<R>(b + (size_t)p * L, G.k0, L, x);
            load_piece<R>(M + (size_t)p * L, G.k0, L, y);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float prod = x[r] * y[r];
                acc[r] = acc[r] + prod;
            }
        }
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int r = 0; r < R; r++) sh[t * R + r] = G.on ? acc[r] : 0.0f;
    __syncthreads();
    // labels strided over the block's threads (a block-mode group may hold more labels
    // than the block has threads)
    for (int lab = t; lab < L; lab += blockDim.x) {
        const int ln = lab / R, rr = lab - ln * R;  // lane and slot holding label lab
        float s = 0.0f;
        for (int gi = 0; gi < items; gi++) {
            const int lane0 = WAVE ? (gi / per) * 64 + (gi % per) * S : gi * S;
            s = s + sh[(lane0 + ln) * R + rr];
        }
        partial[(size_t)blockIdx.x * L + lab] = s;
    }
}

// The L x L statistic G[a][c] = sum_t sum_p b_t[p][a] M_t[p][c] is a tall-skinny GEMM
// (K = iterations x pixels): one wave per (iteration, pixel range, 32 x 32 output tile)
// accumulates it on the matrix cores with v_mfma_f32_32x32x2_f32, two pixels per
// instruction, kMfmaBatch pixel pairs of operands in flight; lane l supplies
// A[l % 32][l / 32] = b[p + l/32][a] and B[l / 32][l % 32] = M[p + l/32][c] (labels past
// L read as 0).  Iteration t's operands are b + t * bstride and M + t * mstride.  The
// tile goes to partial[t * ranges + range][a * L + c]; k_reduce_cols adds them in double.
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kMfmaBatch = 8;

__global__ void __launch_bounds__(256) k_grad_mfma(const float* __restrict__ b, const float* __restrict__ M, int N,
                                                   int L, int tiles1, int ppw, int ranges, size_t bstride,
                                                   size_t mstride, int n_waves, float* __restrict__ partial) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n_waves) return;  // whole waves only; no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const int tiles2 = tiles1 * tiles1;
    const int it_range = w / tiles2, tile = w - it_range * tiles2;
    const int it = it_range / ranges, range = it_range - it * ranges;
    b += (size_t)it * bstride;
    M += (size_t)it * mstride;
    const int ta = tile / tiles1, tc = tile - ta * tiles1;
    const int i = lane & 31, k = lane >> 5;
    const int a = ta * 32 + i, c = tc * 32 + i;
    const int ac = min(a, L - 1), cc = min(c, L - 1);  // clamped: loads stay in bounds
    const float ma = a < L ? 1.0f : 0.0f, mc = c < L ? 1.0f : 0.0f;
    const int p0 = range * ppw, p1 = min(N, p0 + ppw);
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int p = p0; p < p1; p += 2 * kMfmaBatch) {
        float av[kMfmaBatch], bv[kMfmaBatch];
#pragma unroll
        for (int u = 0; u < kMfmaBatch; u++) {
            const int pp = p + 2 * u + k;
            const int pc = min(pp, p1 - 1);
            av[u] = b[(size_t)pc * L + ac];
            bv[u] = M[(size_t)pc * L + cc];
        }
#pragma unroll
        for (int u = 0; u < kMfmaBatch; u++) {
            const float live = p + 2 * u + k < p1 ? 1.0f : 0.0f;  // pixels past the range add 0
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u] * (ma * live), bv[u] * mc, acc, 0, 0, 0);
        }
    }
    float* out = partial + (size_t)it_range * L * L;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = ta * 32 + 8 * (r >> 2) + 4 * k + (r & 3), col = tc * 32 + i;
        if (row < L && col < L) out[(size_t)row * L + col] = acc[r];
    }
}

// acc[off + j] += sum over k < nblocks of partial[k * stride + j], in double: one block
// per column j, a fixed split over the threads, then a fixed tree (deterministic).
__global__ void k_reduce_cols(const float* __restrict__ partial, int nblocks, int stride, double* __restrict__ acc,
                              int off) {
    __shared__ double red[256];
    const int j = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int k = t; k < nblocks; k += blockDim.x) s += (double)partial[(size_t)k * stride + j];
    red[t] = s;
    __syncthreads();
    for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) acc[off + j] += red[0];
}

__global__ void k_fill_double(double* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Gradients of one term from its statistic G (diagonal [L] for Potts/Diagonal, full
// [L*L] for Matrix): raw compatibility parameters and the term weight.
__global__ void k_compat_finish(const double* __restrict__ G, int ckind, int L, const float* __restrict__ cw,
                                const float* __restrict__ kw, float* __restrict__ dcompat,
                                float* __restrict__ dkw) {
    const int t = threadIdx.x;
    const double w = kw ? (double)kw[0] : 1.0;
    if (ckind == DCRF_COMPAT_POTTS) {
        if (t == 0) {
            double s = 0.0;
            for (int a = 0; a < L; a++) s += G[a];
            if (dcompat) dcompat[0] = (float)(-w * s);
            if (dkw) dkw[0] = (float)(-(double)cw[0] * s);
        }
    } else if (ckind == DCRF_COMPAT_DIAGONAL) {
        for (int a = t; a < L; a += blockDim.x)
            if (dcompat) dcompat[a] = (float)(w * G[a]);
        if (t == 0 && dkw) {
            double s = 0.0;
            for (int a = 0; a < L; a++) s += (double)cw[a] * G[a];
            dkw[0] = (float)s;
        }
    } else {
        for (int j = t; j < L * L; j += blockDim.x) {
            const int a = j / L, c = j - a * L;
            if (dcompat) dcompat[j] = (float)(w * 0.5 * (G[a * L + c] + G[c * L + a]));
        }
        if (t == 0 && dkw) {
            double s = 0.0;
            for (int j = 0; j < L * L; j++) s += (double)cw[j] * G[j];
            dkw[0] = (float)s;
        }
    }
}

// Compatibility buffer from raw parameters: Matrix W = (m + m^T) / 2
// (labelcompatibility.cpp:79); Potts / Diagonal copied.
__global__ void k_set_compat(const float* __restrict__ raw, int ckind, int L, float* __restrict__ cw) {
    const int n = ckind == DCRF_COMPAT_POTTS ? 1 : (ckind == DCRF_COMPAT_DIAGONAL ? L : L * L);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        if (ckind == DCRF_COMPAT_MATRIX) {
            const int a = j / L, c = j - a * L;
            const float s = raw[a * L + c] + raw[c * L + a];
            cw[j] = 0.5f * s;
        } else {
            cw[j] = raw[j];
        }
    }
}

}  // namespace

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

void launch_sn_first(const float* Q, const float* dQ, float* b, float* dU, int N, int L, hipStream_t s) {
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH(sh, k_sn_first, blocks_for(N, sh.items), sizeof(float) * sh.block, s, Q, dQ, b, dU, N, L, sh.S,
                  sh.per);
}

void launch_adjoint(const float* Q, float* b, float* dU, int N, int L, const SliceTable& st_in, hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH_CWL(sh, cwf, k_adjoint, blocks_for(N, sh.items), sizeof(float) * (sh.block + cwf), s, Q, b, dU,
                      N, L, sh.S, sh.per, st);
}

int grad_partial_blocks(int N, int L) {
    const LaunchShape sh = shape_for(L);
    const int need = (N + sh.items - 1) / sh.items;
    return need < 1024 ? need : 1024;
}

namespace {
// Pixel ranges per iteration of the matrix-core statistic: about kStatWaves waves over
// all iterations, at least 16 pixels per range; an even range length keeps the MFMA
// pixel pairs aligned.
constexpr long long kStatWaves = 4096;
void stat_ranges(int N, int L, int iters, int* ppw_out, long long* ranges_out) {
    const int tiles1 = (L + 31) / 32;
    const long long per_iter = std::max<long long>(1, kStatWaves / ((long long)iters * tiles1 * tiles1));
    long long ranges = std::max<long long>(1, std::min<long long>(per_iter, (N + 15) / 16));
    int ppw = (int)((N + ranges - 1) / ranges);
    ppw = (ppw + 1) & ~1;
    *ppw_out = ppw;
    *ranges_out = (N + ppw - 1) / ppw;
}
}  // namespace

size_t grad_partial_floats(int N, int L, int iters, bool full) {
    if (!full) return (size_t)iters * grad_partial_blocks(N, L) * L;
    int ppw;
    long long ranges;
    stat_ranges(N, L, iters, &ppw, &ranges);
    return (size_t)iters * ranges * L * L;
}

void launch_grad_stat(const float* b, size_t bstride, const float* M, size_t mstride, int iters, int N, int L,
                      bool full, float* partial, double* acc, hipStream_t s) {
    if (iters < 1) return;
    if (!full) {
        const int nb = grad_partial_blocks(N, L);
        const LaunchShape sh = shape_for(L);
        DCRF_DISPATCH(sh, k_grad_diag, (unsigned)iters * nb, sizeof(float) * sh.block * sh.R, s, b, M, N, L, sh.S,
                      sh.per, nb, bstride, mstride, partial);
        const long long nblk = (long long)iters * nb;
        const int threads = nblk >= 256 ? 256 : (nblk >= 64 ? 64 : 32);
        k_reduce_cols<<<L, threads, 0, s>>>(partial, (int)nblk, L, acc, 0);
        DCRF_HIP_CHECK(hipGetLastError());
        return;
    }
    const long long LL = (long long)L * L;
    int ppw;
    long long ranges;
    stat_ranges(N, L, iters, &ppw, &ranges);
    const int tiles1 = (L + 31) / 32;
    const long long waves = (long long)iters * ranges * tiles1 * tiles1;
    if (waves > 0x7fffffffll / 64 || (long long)iters * ranges > 0x7fffffffll)
        throw Error{DCRF_ERR_VALUE, "gradient statistic too large"};
    k_grad_mfma<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(b, M, N, L, tiles1, ppw, (int)ranges, bstride, mstride,
                                                            (int)waves, partial);
    DCRF_HIP_CHECK(hipGetLastError());
    const long long nblk = (long long)iters * ranges;
    const int threads = nblk >= 256 ? 256 : (nblk >= 64 ? 64 : 32);
    k_reduce_cols<<<(unsigned)LL, threads, 0, s>>>(partial, (int)nblk, (int)LL, acc, 0);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_fill_double(double* p, size_t n, hipStream_t s) {
    if (!n) return;
    k_fill_double<<<blocks_for(n), kThreads, 0, s>>>(p, n);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_compat_finish(const double* G, int ckind, int L, const float* cw, const float* kw, float* dcompat,
                          float* dkw, hipStream_t s) {
    k_compat_finish<<<1, kThreads, 0, s>>>(G, ckind, L, cw, kw, dcompat, dkw);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_set_compat(const float* raw, int ckind, int L, float* cw, hipStream_t s) {
    const int n = ckind == DCRF_COMPAT_POTTS ? 1 : (ckind == DCRF_COMPAT_DIAGONAL ? L : L * L);
    k_set_compat<<<blocks_for(n), kThreads, 0, s>>>(raw, ckind, L, cw);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
