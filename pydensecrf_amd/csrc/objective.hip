// The learning objectives on the device (pydensecrf/densecrf/src/objective.cpp:37-108):
// the objective's value and dL/dQ from the final Q [N][L] and the ground truth, so that
// DenseCRF::gradient's objective step needs no copy of Q to the host.
//   LogLikelihood(gt, robust)  r = sum_i log(max(Q_gt + robust, 1e-20)) / N, dQ_gt = 1/QQ/N
//   Hamming(gt, w)             r = sum_i w_gt Q_gt,                          dQ_gt = w_gt
//   IntersectionOverUnion(gt)  r = sum_l in_l / un_l / L with in_l, un_l in double,
//                              dQ(l, i) = 1 / (un_l L) for l = gt_i, else -in_l / (un_l^2 L)
// Pixels whose label is outside [0, L) take no part (dQ = 0).  Sums over pixels are
// double, per block in a fixed tree and then over blocks in block order: results are
// deterministic (the order differs from the reference's sequential loop).
#include "dcrf_internal.h"

namespace dcrf {
namespace {

constexpr int kThreads = 256;

// LogLikelihood / Hamming: one thread per pixel, the block's sum of r's terms into
// partial[block].
__global__ void __launch_bounds__(kThreads) k_obj_pixel(const float* __restrict__ Q, const int16_t* __restrict__ gt,
                                                        int N, int L, int kind, float param,
                                                        const float* __restrict__ w, float* __restrict__ dQ,
                                                        double* __restrict__ partial) {
    __shared__ double red[kThreads];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kThreads + t;
    double v = 0.0;
    if (i < N) {
        const int g = gt[i];
        if (g >= 0 && g < L) {
            const size_t at = (size_t)i * L + g;
            if (kind == DCRF_OBJECTIVE_LOG_LIKELIHOOD) {
                const float qr = Q[at] + param;
                const float QQ = qr > 1e-20f ? qr : 1e-20f;
                v = log((double)QQ) / N;
                dQ[at] = 1.0f / QQ / (float)N;
            } else {
                const float QQ = w[g] * Q[at];
                v = QQ;
                dQ[at] = w[g];
            }
        }
    }
    red[t] = v;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    if (t == 0) partial[blockIdx.x] = red[0];
}

// IoU statistics: labels in chunks of LB = min(L, 256) (blockIdx.y); thread t of a
// block keeps label chunk * LB + t % LB for the pixels t / LB, t / LB + P, ... of the
// block's range (P = 256 / LB pixels per round, reads coalesced over [N][L] when the
// chunk is the whole label range); the threads of one label are then added in thread
// order into partial[block][2 l] (in_l) and [2 l + 1] (un_l without its 1e-20 start).
__global__ void __launch_bounds__(kThreads) k_iou_stats(const float* __restrict__ Q, const int16_t* __restrict__ gt,
                                                        int N, int L, int ppb, double* __restrict__ partial) {
    __shared__ double s_in[kThreads], s_un[kThreads];
    const int t = threadIdx.x;
    const int LB = min(L, kThreads);
    const int P = kThreads / LB;  // pixels per round
    const int lc = t % LB, po = t / LB;
    const int l = blockIdx.y * LB + lc;
    const int p0 = blockIdx.x * ppb, p1 = min(N, p0 + ppb);
    double in = 0.0, un = 0.0;
    if (po < P && l < L) {
        for (int i = p0 + po; i < p1; i += P) {
            const int g = gt[i];
            if (g < 0 || g >= L) continue;
            const float q = Q[(size_t)i * L + l];
            if (l == g) {
                in += q;
                un += 1.0;
            } else {
                un += q;
            }
        }
    }
    s_in[t] = in;
    s_un[t] = un;
    __syncthreads();
    if (t < LB && l < L) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < P; k++) {
            a += s_in[k * LB + t];
            b += s_un[k * LB + t];
        }
        partial[(size_t)blockIdx.x * 2 * L + 2 * l] = a;
        partial[(size_t)blockIdx.x * 2 * L + 2 * l + 1] = b;
    }
}

// stats[2 l] = in_l, stats[2 l + 1] = un_l = 1e-20 + the blocks' sums, in block order.
__global__ void k_iou_final(const double* __restrict__ partial, int nblocks, int L, double* __restrict__ stats) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 2 * L) return;
    double s = (j & 1) ? 1e-20 : 0.0;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * 2 * L + j];
    stats[j] = s;
}

__global__ void k_iou_grad(const int16_t* __restrict__ gt, const double* __restrict__ stats, int N, int L,
                           float* __restrict__ dQ) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)N * L) return;
    const int i = (int)(e / L), l = (int)(e - (size_t)i * L);
    const int g = gt[i];
    float v = 0.0f;
    if (g >= 0 && g < L) {
        const double in = stats[2 * l], un = stats[2 * l + 1];
        v = l == g ? (float)(1.0 / (un * L)) : (float)(-in / (un * un * L));
    }
    dQ[e] = v;
}

}  // namespace

int objective_partials(int kind, int N, int L) {
    if (kind == DCRF_OBJECTIVE_IOU) {
        const int nb = std::max(1, std::min(1024, (N + 4095) / 4096));
        return nb * 2 * L + 2 * L;  // block partials, then the final statistics
    }
    return std::max(1, (N + kThreads - 1) / kThreads);
}

void launch_objective(int kind, const float* Q, const int16_t* gt, int N, int L, float param, const float* w,
                      float* dQ, double* partial, hipStream_t s) {
    if (N <= 0) return;
    if (kind == DCRF_OBJECTIVE_IOU) {
        const int nb = std::max(1, std::min(1024, (N + 4095) / 4096));
        const int ppb = (N + nb - 1) / nb;
        const int chunks = (L + kThreads - 1) / kThreads;
        k_iou_stats<<<dim3((unsigned)nb, (unsigned)chunks), kThreads, 0, s>>>(Q, gt, N, L, ppb, partial);
        DCRF_HIP_CHECK(hipGetLastError());
        double* stats = partial + (size_t)nb * 2 * L;
        k_iou_final<<<(unsigned)((2 * L + kThreads - 1) / kThreads), kThreads, 0, s>>>(partial, nb, L, stats);
        DCRF_HIP_CHECK(hipGetLastError());
        const size_t n = (size_t)N * L;
        k_iou_grad<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(gt, stats, N, L, dQ);
        DCRF_HIP_CHECK(hipGetLastError());
        return;
    }
    launch_fill(dQ, 0.0f, (size_t)N * L, s);
    k_obj_pixel<<<(unsigned)((N + kThreads - 1) / kThreads), kThreads, 0, s>>>(Q, gt, N, L, kind, param, w, dQ,
                                                                                partial);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
