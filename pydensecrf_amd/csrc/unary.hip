// LogisticUnaryEnergy on the matrix cores (pydensecrf/densecrf/src/unary.cpp:48-68):
//   U = L f            L [M labels][F features], f [F][N]      (get, :50-52)
//   dL/dL = b f^T      b = dL/dU summed over iterations        (gradient, :64-68)
// Device layouts: U [N][M] label fastest (the rest of the library's), L as given
// [M][F]; the features both as given [F][N] (the forward's operand: a feature row is
// contiguous over pixels) and transposed to [N][F] (the gradient's: a pixel row is
// contiguous over features).
//
// Both products run on v_mfma_f32_32x32x2_f32 (fp32 in, fp32 accumulate, the precision
// of the reference's Eigen product).  The forward is one wave per 32 pixels x 32
// labels tile with K = F (the K order inside an MFMA sum is free as long as A and B
// use the same assignment: here instruction j of a batch of 16 takes feature
// c0 + 2j + lane/32).  The gradient is a tall-skinny GEMM
// (K = N pixels): one wave per (32 x 32 output tile, pixel range), the ranges added in
// double in a fixed order afterwards (deterministic).
#include <algorithm>
#include <cmath>

#include "dcrf_internal.h"

namespace dcrf {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kThreads = 256;
constexpr int kGramBatch = 8;  // pixel pairs of operands in flight per wave

__global__ void __launch_bounds__(256) k_logistic_mfma(const float* __restrict__ fFN, const float* __restrict__ Lm,
                                                       int N, int F, int M, int tilesM, int n_waves,
                                                       float* __restrict__ U) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n_waves) return;  // whole waves only; no barrier in this kernel
    const int lane = threadIdx.x & 63;
    const int i = lane & 31, h = lane >> 5;
    const int tn = w / tilesM, tm = w - tn * tilesM;
    const int n = tn * 32 + i, m = tm * 32 + i;
    const int nc = min(n, N - 1), mc = min(m, M - 1);  // clamped: every load stays in bounds
    const float* brow = Lm + (size_t)mc * F;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    // instruction j of a batch takes features c0 + 2j + h: the 32 lanes of a half read 32
    // consecutive pixels of one feature row (one 128-byte access); a batch of 16
    // features keeps 8 such loads in flight
    for (int c0 = 0; c0 < F; c0 += 16) {
        float a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 2 * j + h;
            const int cc = min(c, F - 1);
            a[j] = fFN[(size_t)cc * N + nc] * (c < F ? 1.0f : 0.0f);  // features past F add 0
            b[j] = brow[cc];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j], b[j], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = tn * 32 + 8 * (r >> 2) + 4 * h + (r & 3), col = tm * 32 + i;
        if (row < N && col < M) U[(size_t)row * M + col] = acc[r];
    }
}

// k_logistic_mfma for M <= 32 labels (one label tile): the block first copies L into
// LDS (rows padded to F + 1 floats: the 32 lanes of a half read 32 different rows of one
// column without bank conflicts), so the B operand comes from LDS instead of a 32-row
// gather per instruction.  Same products in the same order as k_logistic_mfma.
__global__ void __launch_bounds__(256) k_logistic_mfma_lds(const float* __restrict__ fFN,
                                                           const float* __restrict__ Lm, int N, int F, int M,
                                                           int n_waves, float* __restrict__ U) {
    extern __shared__ float sL[];  // M rows of F + 1 floats
    const int ldl = F + 1;
    for (int q = threadIdx.x; q < M * F; q += blockDim.x) {
        const int r = q / F, c = q - r * F;
        sL[r * ldl + c] = Lm[q];
    }
    __syncthreads();
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n_waves) return;  // after the block's only barrier
    const int lane = threadIdx.x & 63;
    const int i = lane & 31, h = lane >> 5;
    const int n = w * 32 + i;
    const int nc = min(n, N - 1), mc = min(i, M - 1);
    const float* brow = sL + mc * ldl;
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int c0 = 0; c0 < F; c0 += 16) {
        float a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = c0 + 2 * j + h;
            const int cc = min(c, F - 1);
            a[j] = fFN[(size_t)cc * N + nc] * (c < F ? 1.0f : 0.0f);  // features past F add 0
            b[j] = brow[cc];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j], b[j], acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = w * 32 + 8 * (r >> 2) + 4 * h + (r & 3), col = i;
        if (row < N && col < M) U[(size_t)row * M + col] = acc[r];
    }
}

// G[a][c] = sum_p A[p][a] B[p][c] over the pixel range of each wave: lane l supplies
// A[l % 32][l / 32] = A[p + l/32][a] and B[l / 32][l % 32] = B[p + l/32][c] (columns past
// LA / LB and pixels past the range read as 0); partial[range][a * LB + c].
__global__ void __launch_bounds__(256) k_gram_mfma(const float* __restrict__ A, int LA, const float* __restrict__ B,
                                                   int LB, int N, int tilesB, int tiles, int ppw, int n_waves,
                                                   float* __restrict__ partial) {
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= n_waves) return;
    const int lane = threadIdx.x & 63;
    const int range = w / tiles, tile = w - range * tiles;
    const int ta = tile / tilesB, tc = tile - ta * tilesB;
    const int i = lane & 31, k = lane >> 5;
    const int a = ta * 32 + i, c = tc * 32 + i;
    const int ac = min(a, LA - 1), cc = min(c, LB - 1);
    const float ma = a < LA ? 1.0f : 0.0f, mc = c < LB ? 1.0f : 0.0f;
    const int p0 = range * ppw, p1 = min(N, p0 + ppw);
    floatx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int p = p0; p < p1; p += 2 * kGramBatch) {
        float av[kGramBatch], bv[kGramBatch];
#pragma unroll
        for (int u = 0; u < kGramBatch; u++) {
            const int pc = min(p + 2 * u + k, p1 - 1);
            av[u] = A[(size_t)pc * LA + ac];
            bv[u] = B[(size_t)pc * LB + cc];
        }
#pragma unroll
        for (int u = 0; u < kGramBatch; u++) {
            const float live = p + 2 * u + k < p1 ? 1.0f : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u] * (ma * live), bv[u] * mc, acc, 0, 0, 0);
        }
    }
    float* out = partial + (size_t)range * LA * LB;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = ta * 32 + 8 * (r >> 2) + 4 * k + (r & 3), col = tc * 32 + i;
        if (row < LA && col < LB) out[(size_t)row * LB + col] = acc[r];
    }
}

// out[j] = sum over ranges of partial[k * stride + j] in double: one block per column, a
// fixed split over the threads and a fixed tree.
__global__ void k_gram_reduce(const float* __restrict__ partial, int nranges, int stride, double* __restrict__ out) {
    __shared__ double red[256];
    const int j = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int k = t; k < nranges; k += blockDim.x) s += (double)partial[(size_t)k * stride + j];
    red[t] = s;
    __syncthreads();
    for (int hh = blockDim.x >> 1; hh > 0; hh >>= 1) {
        if (t < hh) red[t] += red[t + hh];
        __syncthreads();
    }
    if (t == 0) out[j] = red[0];
}

// utils.unary_from_softmax (pydensecrf/utils.py:55-79) fused with the transpose into
// the library's layout: U[n][l] = -log(clip(scale sm[l][n] + (1 - scale) / L)), in
// double like numpy's float64 path, rounded once to float.  A 32 x 32 tile per block
// through LDS: reads coalesced along pixels, writes along labels.
__global__ void k_unary_from_softmax(const float* __restrict__ sm, int L, int N, double scale, int use_scale,
                                     double clip, int use_clip, float* __restrict__ U) {
    __shared__ float tile[32][33];
    const int n0 = blockIdx.x * 32, l0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int l = l0 + k, n = n0 + tx;
        float v = 0.0f;
        if (l < L && n < N) {
            double x = sm[(size_t)l * N + n];
            if (use_scale) x = scale * x + (1.0 - scale) * (1.0 / L);
            if (use_clip) x = x < clip ? clip : (x > 1.0 ? 1.0 : x);
            v = (float)(-log(x));
        }
        tile[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int n = n0 + k, l = l0 + tx;
        if (n < N && l < L) U[(size_t)n * L + l] = tile[tx][k];
    }
}

// utils.unary_from_labels (pydensecrf/utils.py:6-45): U[n][l] = p_energy for the
// annotated label (label - 1 with zero_unsure), n_energy elsewhere, and -log(1/L) for
// every label of an unsure (0) pixel.
__global__ void k_unary_from_labels(const int* __restrict__ labels, int N, int L, float p_energy, float n_energy,
                                    float u_energy, int zero_unsure, float* __restrict__ U) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * L) return;
    const int n = (int)(i / L), l = (int)(i - (size_t)n * L);
    const int lab = labels[n];
    float v;
    if (zero_unsure && lab == 0)
        v = u_energy;
    else
        v = (zero_unsure ? lab - 1 : lab) == l ? p_energy : n_energy;
    U[i] = v;
}

}  // namespace

void launch_unary_from_softmax(const float* sm, int L, int N, double scale, bool use_scale, double clip,
                               bool use_clip, float* U, hipStream_t s) {
    if (N <= 0 || L <= 0) return;
    const dim3 grid((unsigned)((N + 31) / 32), (unsigned)((L + 31) / 32));
    k_unary_from_softmax<<<grid, kThreads, 0, s>>>(sm, L, N, scale, use_scale ? 1 : 0, clip, use_clip ? 1 : 0, U);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_unary_from_labels(const int* labels, int N, int L, float p_energy, float n_energy, float u_energy,
                              bool zero_unsure, float* U, hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n) return;
    k_unary_from_labels<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(
        labels, N, L, p_energy, n_energy, u_energy, zero_unsure ? 1 : 0, U);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_logistic_unary(const float* fFN, const float* Lm, int N, int F, int M, float* U, hipStream_t s) {
    if (N <= 0 || M <= 0) return;
    if (F <= 0) throw Error{DCRF_ERR_VALUE, "logistic unary: no features"};
    const int tilesM = (M + 31) / 32;
    const long long waves = (long long)((N + 31) / 32) * tilesM;
    if (waves > 0x7fffffffll / 64) throw Error{DCRF_ERR_VALUE, "logistic unary too large"};
    const size_t lds = (size_t)M * (F + 1) * sizeof(float);
    if (tilesM == 1 && lds <= 32 * 1024) {
        k_logistic_mfma_lds<<<(unsigned)((waves + 3) / 4), kThreads, lds, s>>>(fFN, Lm, N, F, M, (int)waves, U);
    } else {
        k_logistic_mfma<<<(unsigned)((waves + 3) / 4), kThreads, 0, s>>>(fFN, Lm, N, F, M, tilesM, (int)waves, U);
    }
    DCRF_HIP_CHECK(hipGetLastError());
}

size_t gram_partial_floats(int LA, int LB, int N) {
    const long long per = (long long)LA * LB;
    const long long ranges = std::max<long long>(1, std::min<long long>(((long long)1 << 22) / per, (N + 15) / 16));
    return (size_t)(ranges * per);
}

void launch_gram(const float* A, int LA, const float* B, int LB, int N, float* partial, double* out, hipStream_t s) {
    if (LA <= 0 || LB <= 0) return;
    const long long per = (long long)LA * LB;
    long long ranges = (long long)(gram_partial_floats(LA, LB, N) / per);
    int ppw = (int)((N + ranges - 1) / ranges);
    ppw = std::max(2, (ppw + 1) & ~1);  // even: the MFMA pixel pairs stay aligned
    ranges = (N + ppw - 1) / ppw;
    const int tilesA = (LA + 31) / 32, tilesB = (LB + 31) / 32;
    const long long waves = ranges * tilesA * tilesB;
    if (waves > 0x7fffffffll / 64) throw Error{DCRF_ERR_VALUE, "gram product too large"};
    if (waves > 0) {
        k_gram_mfma<<<(unsigned)((waves + 3) / 4), kThreads, 0, s>>>(A, LA, B, LB, N, tilesB, tilesA * tilesB, ppw,
                                                                     (int)waves, partial);
        DCRF_HIP_CHECK(hipGetLastError());
    }
    const int threads = ranges >= 256 ? 256 : (ranges >= 64 ? 64 : 32);
    k_gram_reduce<<<(unsigned)per, threads, 0, s>>>(partial, (int)ranges, (int)per, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
