// Adjoint mean-field for gfx950: the backward of E_{t+1} = U + sum_k C_k(w_k K~_k Q_t),
// Q_t = softmax(-E_t), following DenseCRF::gradient (pydensecrf/densecrf/src/
// densecrf.cpp:238-297) with the reference's conventions:
//   b_n   = sumAndNormalize(Q_n .* dL/dQ_n, Q_n)                  (= dL/dE_n, :257-258)
//   for t = n-1 .. 0:
//     G_k  += b_{t+1} (K~_k Q_t)^T         compatibility / kernel-weight statistics
//                                          (labelcompatibility.cpp:57-61, 76-78, 101-108);
//                                          Potts / Diagonal take only its diagonal, summed
//                                          as (K~_k^T b_{t+1}) .* Q_t inside the adjoint
//                                          kernel (the transposed message is at hand there)
//     g_t   = sum_k w_k K~_k^T C_k^T b_{t+1}   transposed filter (reverse blur order,
//                                          pairwise.cpp:179-183, permutohedral.cpp:556)
//     b_t   = sumAndNormalize(Q_t .* g_t, Q_t)                     (:290)
//   dL/dU = sum_t b_t                                              (:261-262, 293-294)
// Matrix terms take K~_k Q_t from the forward pass's message cache; either way the
// backward runs one filter per term per iteration (the reference also re-filters Q_t
// for the statistic, pairwise.cpp:193).
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

// Terms with a statistic row (SliceTerm::gstat) also leave the products of their raw
// transposed message and Q in the statistic slots (after the compat region), summed per
// block at the end: sum_p (K~^T b)[p] .* Q[p] = sum_p b[p] .* (K~ Q)[p], the diagonal of
// the compatibility gradient's statistic, with no message cache.
template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_adjoint(const float* __restrict__ Q, float* __restrict__ b,
                                                  float* __restrict__ dU, int N, int L, int S, int per,
                                                  SliceTable st, int cwf) {
    extern __shared__ float sh[];
    float* shc = sh + blockDim.x;
    float* sst = shc + cwf;
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int k0 = G.k0;
    float q[R];
#pragma unroll
    for (int r = 0; r < R; r++) q[r] = 0.0f;
    if (active) load_piece<R>(Q + (size_t)p * L, k0, L, q);
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
            if (T.gslot >= 0) {
                float v[R];
#pragma unroll
                for (int r = 0; r < R; r++) v[r] = active ? m[r] * q[r] : 0.0f;
                stat_put<R>(sst, T.gslot, v);
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
    float bn[R];
    sum_normalize<WAVE, R>(q, gsum, bn, G);
    if (active) {
        float du[R];
        load_piece<R>(dU + (size_t)p * L, k0, L, du);
#pragma unroll
        for (int r = 0; r < R; r++) du[r] = du[r] + bn[r];
        store_piece<R>(b + (size_t)p * L, k0, L, bn);
        store_piece<R>(dU + (size_t)p * L, k0, L, du);
    }
    if (st.nstat) stat_finish<WAVE, R>(st, G, sst);  // uniform condition
}

// The L x L statistic G[a][c] = sum_t sum_p b_t[p][a] M_t[p][c] is a tall-skinny GEMM
// (K = iterations x pixels): one wave per (iteration, pixel range, 32 x 32 output tile)
// accumulates it on the matrix cores with v_mfma_f32_32x32x2_f32, two pixels per
// instruction, kMfmaBatch pixel pairs of operands in flight; lane l supplies
// A[l % 32][l / 32] = b[p + l/32][a] and B[l / 32][l % 32] = M[p + l/32][c] (labels past
// L read as 0).  Iteration t's operands are b + t * bstride and M + t * mstride.  The
// tile goes to partial[t * ranges + range][a * L + c]; launch_sum_rows adds them in double.
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

// acc[j] += sum over r < rows of part[r * cols + j], in double and a fixed order.
// Stage 1: block i takes rows [i * rpb, (i + 1) * rpb); the columns go in passes of
// cpb = min(cols, 256), tpc = 256 / cpb threads per column, thread t adding column
// c0 + t % cpb of every tpc-th row from row t / cpb (consecutive threads read
// consecutive floats; four rows in flight, added in row order); the column's threads
// are then added in thread order into mid[i * cols + j].  Stage 2: one block per
// column adds mid over the blocks (strided over its threads, then a fixed tree).
constexpr int kRowBlocks = 1024;
constexpr int kRowsMin = 16;  // rows per first-stage block, at least

__global__ void __launch_bounds__(256) k_rows_stage1(const float* __restrict__ part, long long rows, int cols,
                                                      long long rpb, double* __restrict__ mid) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    const int cpb = cols < 256 ? cols : 256, tpc = 256 / cpb;
    const long long r0 = (long long)blockIdx.x * rpb, r1 = min(rows, r0 + rpb);
    const int sub = t / cpb;
    for (int c0 = 0; c0 < cols; c0 += cpb) {
        const int c = c0 + t % cpb;
        double acc = 0.0;
        if (sub < tpc && c < cols) {
            const float* src = part + c;
            long long r = r0 + sub;
            for (; r + 3 * tpc < r1; r += 4 * tpc) {
                const float a0 = src[r * cols], a1 = src[(r + tpc) * cols];
                const float a2 = src[(r + 2 * tpc) * cols], a3 = src[(r + 3 * tpc) * cols];
                acc += (double)a0;
                acc += (double)a1;
                acc += (double)a2;
                acc += (double)a3;
            }
            for (; r < r1; r += tpc) acc += (double)src[r * cols];
        }
        red[t] = acc;
        __syncthreads();
        if (t < cpb && c0 + t < cols) {
            double v = 0.0;
            for (int i = 0; i < tpc; i++) v += red[i * cpb + t];
            mid[(size_t)blockIdx.x * cols + c0 + t] = v;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_rows_stage2(const double* __restrict__ mid, int nb, int cols,
                                                      double* __restrict__ acc) {
    __shared__ double red[256];
    const int c = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (int i = t; i < nb; i += blockDim.x) v += mid[(size_t)i * cols + c];
    red[t] = v;
    __syncthreads();
    for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) acc[c] += red[0];
}

__global__ void k_fill_double(double* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Gradients of one term from its statistic G (diagonal [L] for Potts/Diagonal, full
// [L*L] for Matrix): raw compatibility parameters and the term weight.  One block of
// 256 threads; the term-weight sum is strided over the threads and added by a fixed
// tree (deterministic).
__global__ void __launch_bounds__(256) k_compat_finish(const double* __restrict__ G, int ckind, int L,
                                                        const float* __restrict__ cw, const float* __restrict__ kw,
                                                        float* __restrict__ dcompat, float* __restrict__ dkw) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    const double w = kw ? (double)kw[0] : 1.0;
    const int n = ckind == DCRF_COMPAT_MATRIX ? L * L : L;
    double s = 0.0;
    for (int j = t; j < n; j += blockDim.x) {
        if (ckind == DCRF_COMPAT_POTTS) {
            s += G[j];
        } else if (ckind == DCRF_COMPAT_DIAGONAL) {
            if (dcompat) dcompat[j] = (float)(w * G[j]);
            s += (double)cw[j] * G[j];
        } else {
            const int a = j / L, c = j - a * L;
            if (dcompat) dcompat[j] = (float)(w * 0.5 * (G[a * L + c] + G[c * L + a]));
            s += (double)cw[j] * G[j];
        }
    }
    red[t] = s;
    __syncthreads();
    for (int h = blockDim.x >> 1; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) {
        if (ckind == DCRF_COMPAT_POTTS) {
            if (dcompat) dcompat[0] = (float)(-w * red[0]);
            if (dkw) dkw[0] = (float)(-(double)cw[0] * red[0]);
        } else if (dkw) {
            dkw[0] = (float)red[0];
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

void launch_sn_first(const float* Q, const float* dQ, float* b, float* dU, int N, int L, hipStream_t s) {
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH(sh, k_sn_first, blocks_for(N, sh.items), sizeof(float) * sh.block, s, Q, dQ, b, dU, N, L, sh.S,
                  sh.per);
}

void launch_adjoint(const float* Q, float* b, float* dU, int N, int L, const SliceTable& st_in, hipStream_t s) {
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    const size_t lds =
        sizeof(float) * ((size_t)sh.block + cwf + stat_lds_floats(sh.wave, st.nstat, L, sh.block, sh.R));
    DCRF_DISPATCH_CWL(sh, cwf, k_adjoint, blocks_for(N, sh.items), lds, s, Q, b, dU, N, L, sh.S, sh.per, st, cwf);
}

int adjoint_blocks(int N, int L) {
    const LaunchShape sh = shape_for(L);
    return (int)blocks_for(N, sh.items);
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

size_t grad_partial_floats(int N, int L, int iters) {
    int ppw;
    long long ranges;
    stat_ranges(N, L, iters, &ppw, &ranges);
    return (size_t)iters * ranges * L * L;
}

size_t rows_mid_doubles(int cols) { return (size_t)kRowBlocks * cols; }

void launch_sum_rows(const float* part, long long rows, int cols, double* mid, double* acc, hipStream_t s) {
    if (rows < 1 || cols < 1) return;
    const long long nb = std::max<long long>(1, std::min<long long>(kRowBlocks, rows / kRowsMin));
    const long long rpb = (rows + nb - 1) / nb;
    const int nbu = (int)((rows + rpb - 1) / rpb);  // blocks with rows
    k_rows_stage1<<<nbu, 256, 0, s>>>(part, rows, cols, rpb, mid);
    DCRF_HIP_CHECK(hipGetLastError());
    k_rows_stage2<<<cols, 256, 0, s>>>(mid, nbu, cols, acc);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_grad_stat(const float* b, size_t bstride, const float* M, size_t mstride, int iters, int N, int L,
                      float* partial, double* mid, double* acc, hipStream_t s) {
    if (iters < 1) return;
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
    launch_sum_rows(partial, (long long)iters * ranges, L * L, mid, acc, s);
}

void launch_fill_double(double* p, size_t n, hipStream_t s) {
    if (!n) return;
    k_fill_double<<<blocks_for(n), kThreads, 0, s>>>(p, n);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_compat_finish(const double* G, int ckind, int L, const float* cw, const float* kw, float* dcompat,
                          float* dkw, hipStream_t s) {
    k_compat_finish<<<1, 256, 0, s>>>(G, ckind, L, cw, kw, dcompat, dkw);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_set_compat(const float* raw, int ckind, int L, float* cw, hipStream_t s) {
    const int n = ckind == DCRF_COMPAT_POTTS ? 1 : (ckind == DCRF_COMPAT_DIAGONAL ? L : L * L);
    k_set_compat<<<blocks_for(n), kThreads, 0, s>>>(raw, ckind, L, cw);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
