// Kernel-parameter (feature-scale) gradient for gfx950, following
// PairwisePotential::kernelGradient (pydensecrf/densecrf/src/pairwise.cpp:202-207),
// DenseKernel::gradient / featureGradient (:87-114, :152-163) and
// Permutohedral::gradient (permutohedral.cpp:612-695).
//
// Per pairwise term and reverse iteration t, with b the adjoint before the update and
// lbl = C (w Q_t) (the compatibility applied to Q_t, times the term weight):
//   fg = featureGradient(b, lbl)                 [N][d] float
//        NO_NORMALIZATION   G(b, lbl)
//        SYMMETRIC          -G(0.5 (b K(lbl n) + K^T(b n) lbl) n^3, 1) + G(b n, lbl n)
//        AFTER              -G((b K(lbl)) n^2, 1) + G(b n, lbl)
//        BEFORE             -G((K^T(b) lbl) n^2, 1) + G(b, lbl n)
//   DIAG kernel: dtheta_j += sum_i f_ij fg_ij;  FULL: dP(r, c) += sum_i fg_ir f_ic
// K is the term's lattice filter without normaliser (splat, d+1 blur steps, slice
// scaled by 1/(1+2^-d)), K^T the same with the blur order reversed.  G(a, c) is the
// lattice gradient of a^T K c with respect to the features: the lattice splats a and
// blurs forward (dir 0), splats c and blurs backward (dir 1), and for each point turns
// its d+1 vertex rows, rotated by rank, into differences (times 1/(1+2^-d)/(d+1)) that
// are elevated onto the feature axes and paired with c (dir 0) or a (dir 1).
// The splat of the all-ones input in dir 1 of G(., 1) is the same for every label and
// every iteration: it is blurred once per lattice with one value per vertex and read
// with a zero label stride.
// Reductions over points run in double in a fixed order (per-block partials, then an
// ordered sum over blocks): results are deterministic.
#include <algorithm>
#include <cmath>

#include "dcrf_internal.h"
#include "dcrf_device.h"

#pragma clang fp contract(off)

namespace dcrf {
namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(size_t n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

// lbl = w * compat(Q): Potts -c Q, Diagonal c_l Q_l, Matrix W Q (labelcompatibility.cpp:
// 31-33, 46-48, 66-69).  One thread per (pixel, label).
__global__ void k_compat_rows(const float* __restrict__ Q, int N, int L, int ckind, const float* __restrict__ cw,
                              const float* __restrict__ kw, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * L) return;
    const int l = (int)(i % L);
    const size_t row = i - l;
    const float w = kw ? kw[0] : 1.0f;
    float v;
    if (ckind == DCRF_COMPAT_POTTS) {
        v = -cw[0] * (Q[i] * w);
    } else if (ckind == DCRF_COMPAT_DIAGONAL) {
        v = cw[l] * (Q[i] * w);
    } else {
        float s = 0.0f;
        for (int j = 0; j < L; j++) {
            const float m = cw[(size_t)l * L + j] * (Q[row + j] * w);
            s = s + m;
        }
        v = s;
    }
    out[i] = v;
}

// Inputs of featureGradient (pairwise.cpp:88-114) before the filters: a_n = b n,
// c_n = lbl n (either may be skipped: null output).
__global__ void k_kg_scale(const float* __restrict__ b, const float* __restrict__ lbl, const float* __restrict__ norm,
                           int N, int L, float* __restrict__ bn, float* __restrict__ ln) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * L) return;
    const float n = norm[i / L];
    if (bn) bn[i] = b[i] * n;
    if (ln) ln[i] = lbl[i] * n;
}

// The first argument of the norm-derivative term, in place of fa (or into x):
//   SYMMETRIC  x = (0.5 (b fb + fa lbl)) n^3      (pairwise.cpp:90-97)
//   AFTER      x = (b fb) n^2                     (:98-105)
//   BEFORE     x = (fa lbl) n^2                   (:106-113)
__global__ void k_kg_x(const float* __restrict__ b, const float* __restrict__ lbl, const float* __restrict__ fa,
                       const float* __restrict__ fb, const float* __restrict__ norm, int N, int L, int ntype,
                       float* __restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * L) return;
    const float n = norm[i / L];
    const float n2 = n * n;
    float v;
    if (ntype == DCRF_NORMALIZE_SYMMETRIC) {
        const float p1 = b[i] * fb[i], p2 = fa[i] * lbl[i];
        const float sum = p1 + p2;
        const float h = 0.5f * sum;
        const float n3 = n2 * n;
        v = h * n3;
    } else if (ntype == DCRF_NORMALIZE_AFTER) {
        const float p = b[i] * fb[i];
        v = p * n2;
    } else {
        const float p = fa[i] * lbl[i];
        v = p * n2;
    }
    x[i] = v;
}

struct GradScale {
    float sf[DCRF_MAX_DIM];  // elevation scale of axis j: sqrt(2/3)(d+1)/sqrt((j+2)(j+1))
    float alpha;             // 1/(1+2^-d)/(d+1)
};

// Slicing gradient of one direction (permutohedral.cpp:662-692), one thread per point:
//   for each label k: u_r = vals[vtx[p][r]][k] (r = 0..d), ra_j = alpha u_{r0} - alpha
//   u_{r1} with r0 = d - rank[p][j], r1 = r0 + 1 mod d + 1; then for j = 1..d
//   grad_j += c[p][k] * sf_{j-1} (sm - j ra_j), sm = ra_0 + .. + ra_{j-1}.
// vals rows have stride ldv and label stride ks (0: one value per vertex for every
// label); c null means all ones.  df[p][j-1] = grad_j (accumulate: += grad_j).
template <int D>
__global__ void __launch_bounds__(kThreads) k_slice_grad(const float* __restrict__ vals, int ldv, int ks,
                                                         const int* __restrict__ vtx, const int* __restrict__ rank,
                                                         const float* __restrict__ c, int N, int L, GradScale gs,
                                                         int accumulate, float* __restrict__ df) {
    constexpr int D1 = D + 1;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    int o0[D1], o1[D1];
    {
        int v[D1];
#pragma unroll
        for (int r = 0; r < D1; r++) v[r] = vtx[(size_t)p * D1 + r];
#pragma unroll
        for (int j = 0; j < D1; j++) {
            const int r0 = D - rank[(size_t)p * D1 + j];
            const int r1 = r0 + 1 > D ? 0 : r0 + 1;
            // r0, r1 in 0..D: select from the registers without dynamic indexing
            int a = v[0], b = v[0];
#pragma unroll
            for (int r = 1; r < D1; r++) {
                a = r0 == r ? v[r] : a;
                b = r1 == r ? v[r] : b;
            }
            o0[j] = a;
            o1[j] = b;
        }
    }
    float grad[D];
#pragma unroll
    for (int j = 0; j < D; j++) grad[j] = 0.0f;
    const float* crow = c ? c + (size_t)p * L : nullptr;
    for (int k = 0; k < L; k++) {
        const size_t col = (size_t)k * ks;
        float ra[D1];
#pragma unroll
        for (int j = 0; j < D1; j++) {
            const float u0 = gs.alpha * vals[(size_t)o0[j] * ldv + col];
            const float u1 = gs.alpha * vals[(size_t)o1[j] * ldv + col];
            ra[j] = u0 - u1;
        }
        const float ck = crow ? crow[k] : 1.0f;
        float sm = ra[0];
#pragma unroll
        for (int j = 1; j <= D; j++) {
            const float jr = (float)j * ra[j];
            const float diff = sm - jr;
            const float v = gs.sf[j - 1] * diff;
            const float prod = ck * v;
            grad[j - 1] = grad[j - 1] + prod;
            sm = sm + ra[j];
        }
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        float* o = df + (size_t)p * D + j;
        *o = accumulate ? *o + grad[j] : grad[j];
    }
}

// Per-block partial sums of the parameter gradient (pairwise.cpp:157-161):
//   fg = (r ? -r + g : g) [N][d];  DIAG: out[j] = sum_p f[q(p)][j] fg[p][j];
//   FULL: out[c*d + r] = sum_p fg[p][r] f[q(p)][c]   (column-major, like the reference).
// f is the term's feature table in the padded point layout: q(p) = pad_base[i] + p -
// pix_base[i] for p in instance i.
__global__ void __launch_bounds__(kThreads) k_kparam_partial(const float* __restrict__ r, const float* __restrict__ g,
                                                             const float* __restrict__ f, const int* __restrict__ pix_base,
                                                             const int* __restrict__ pad_base, int n_inst, int N, int d,
                                                             int full, double* __restrict__ partial) {
    __shared__ double red[kThreads];
    const int np = full ? d * d : d;
    const int stride = gridDim.x * blockDim.x;
    for (int q = 0; q < np; q++) {
        const int row = full ? q % d : q, colf = full ? q / d : q;
        double s = 0.0;
        for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += stride) {
            int inst = 0;
            if (n_inst > 1) {
                int lo = 0, hi = n_inst - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (pix_base[mid] <= p) lo = mid; else hi = mid - 1;
                }
                inst = lo;
            }
            const int fq = pad_base[inst] + (p - pix_base[inst]);
            const float gv = g[(size_t)p * d + row];
            const float fgv = r ? -r[(size_t)p * d + row] + gv : gv;
            s += (double)fgv * (double)f[(size_t)fq * d + colf];
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = kThreads / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * np + q] = red[0];
        __syncthreads();
    }
}

// acc[q] += sum over blocks of partial[b][q], blocks in order.
__global__ void k_kparam_final(const double* __restrict__ partial, int nb, int np, double* __restrict__ acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += partial[(size_t)b * np + q];
    acc[q] += s;
}

__global__ void k_kparam_store(const double* __restrict__ acc, int n, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)acc[i];
}

// Scaled features for DenseKernel::setParameters (pairwise.cpp:141-151): DIAG
// out[q][j] = p_j f[q][j]; FULL out[q][j] = sum_c P(j, c) f[q][c], P column-major.
__global__ void k_scale_features(const float* __restrict__ f, int npad, int d, int full, const float* __restrict__ P,
                                 float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)npad * d) return;
    const int j = (int)(i % d);
    const float* fr = f + (i - j);
    float v;
    if (!full) {
        v = P[j] * fr[j];
    } else {
        float s = 0.0f;
        for (int c = 0; c < d; c++) {
            const float m = P[(size_t)c * d + j] * fr[c];
            s = s + m;
        }
        v = s;
    }
    out[i] = v;
}

}  // namespace

constexpr int kKparamBlocks = 512;

size_t kparam_partial_doubles(int d) { return (size_t)kKparamBlocks * d * d; }

void launch_compat_rows(const float* Q, int N, int L, int ckind, const float* cw, const float* kw, float* out,
                        hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n) return;
    k_compat_rows<<<blocks_for(n), kThreads, 0, s>>>(Q, N, L, ckind, cw, kw, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kg_scale(const float* b, const float* lbl, const float* norm, int N, int L, float* bn, float* ln,
                     hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n || (!bn && !ln)) return;
    k_kg_scale<<<blocks_for(n), kThreads, 0, s>>>(b, lbl, norm, N, L, bn, ln);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kg_x(const float* b, const float* lbl, const float* fa, const float* fb, const float* norm, int N, int L,
                 int ntype, float* x, hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n) return;
    k_kg_x<<<blocks_for(n), kThreads, 0, s>>>(b, lbl, fa, fb, norm, N, L, ntype, x);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_slice_grad(const float* vals, int ldv, int ks, const int* vtx, const int* rank, const float* c, int N,
                       int L, int d, int accumulate, float* df, hipStream_t s) {
    if (!N) return;
    GradScale gs{};
    // the reference's constants (permutohedral.cpp:620-628): double, stored as float
    const float inv_std_dev = (float)(std::sqrt(2.0 / 3.0) * (d + 1));
    for (int i = 0; i < d; i++) gs.sf[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    gs.alpha = 1.0f / (1 + std::pow(2.0f, (float)-d)) / (float)(d + 1);
    const unsigned nb = blocks_for(N);
#define DCRF_SG(DD)                                                                                              \
    case DD:                                                                                                     \
        k_slice_grad<DD><<<nb, kThreads, 0, s>>>(vals, ldv, ks, vtx, rank, c, N, L, gs, accumulate, df); \
        break;
    switch (d) {
        DCRF_SG(1) DCRF_SG(2) DCRF_SG(3) DCRF_SG(4) DCRF_SG(5) DCRF_SG(6) DCRF_SG(7) DCRF_SG(8)
        DCRF_SG(9) DCRF_SG(10) DCRF_SG(11) DCRF_SG(12) DCRF_SG(13) DCRF_SG(14) DCRF_SG(15)
        default:
            throw Error{DCRF_ERR_VALUE, "feature dimension out of range for the kernel gradient"};
    }
#undef DCRF_SG
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kparam_reduce(const float* r, const float* g, const float* f, const int* pix_base, const int* pad_base,
                          int n_inst, int N, int d, bool full, double* partial, double* acc, hipStream_t s) {
    const int np = full ? d * d : d;
    if (!np) return;
    const int nb = std::max(1, std::min(kKparamBlocks, (int)blocks_for(N)));
    k_kparam_partial<<<nb, kThreads, 0, s>>>(r, g, f, pix_base, pad_base, n_inst, N, d, full ? 1 : 0, partial);
    DCRF_HIP_CHECK(hipGetLastError());
    k_kparam_final<<<blocks_for(np, 64), 64, 0, s>>>(partial, nb, np, acc);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kparam_store(const double* acc, int n, float* out, hipStream_t s) {
    if (n <= 0) return;
    k_kparam_store<<<blocks_for(n), kThreads, 0, s>>>(acc, n, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_scale_features(const float* f, int npad, int d, bool full, const float* P, float* out, hipStream_t s) {
    const size_t n = (size_t)npad * d;
    if (!n) return;
    k_scale_features<<<blocks_for(n), kThreads, 0, s>>>(f, npad, d, full ? 1 : 0, P, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
