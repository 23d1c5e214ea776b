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

// The two filter inputs of featureGradient (pairwise.cpp:88-114) side by side,
// P[p] = [a | c] (2L floats): a = b n (SYMMETRIC, AFTER) or b, c = lbl n (SYMMETRIC,
// BEFORE) or lbl.  One splat and the two blur orders of P give K a, K^T a, K c, K^T c.
__global__ void k_kg_pack(const float* __restrict__ b, const float* __restrict__ lbl, const float* __restrict__ norm,
                          int N, int L, int ntype, float* __restrict__ P) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * L) return;
    const size_t p = i / L;
    const int k = (int)(i - p * L);
    const float n = norm[p];
    const bool na = ntype == DCRF_NORMALIZE_SYMMETRIC || ntype == DCRF_NORMALIZE_AFTER;
    const bool nc = ntype == DCRF_NORMALIZE_SYMMETRIC || ntype == DCRF_NORMALIZE_BEFORE;
    float* row = P + p * 2 * L;
    row[k] = na ? b[i] * n : b[i];
    row[L + k] = nc ? lbl[i] * n : lbl[i];
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
// label); c (row stride ldc) null means all ones.  df[p][j-1] = grad_j (accumulate:
// += grad_j).
template <int D>
__global__ void __launch_bounds__(kThreads) k_slice_grad(const float* __restrict__ vals, int ldv, int ks,
                                                         const int* __restrict__ vtx, const int* __restrict__ rank,
                                                         const float* __restrict__ c, int ldc, int N, int L,
                                                         GradScale gs, int accumulate, float* __restrict__ df) {
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
    const float* crow = c ? c + (size_t)p * ldc : nullptr;
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

// The same with S lanes per point, R consecutive labels per lane (S <= 64: the lane
// groups of one wave), for d <= 6: each lane loads its pieces of the d+1 vertex rows
// once, forms the rotated differences by select, and the d gradients are summed over
// the lanes of the group.  The label order of the sum differs from the reference's
// sequential one (float tolerance).
template <int R, int D>
__global__ void __launch_bounds__(kWaveBlock) k_slice_grad_lanes(const float* __restrict__ vals, int ldv, int ks,
                                                                 const int* __restrict__ vtx,
                                                                 const int* __restrict__ rank,
                                                                 const float* __restrict__ c, int ldc, int N, int L,
                                                                 int S, int per, GradScale gs, int accumulate,
                                                                 float* __restrict__ df) {
    constexpr int D1 = D + 1;
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int p = G.item();
    const bool active = G.on && p < N;
    const int pp = active ? p : 0;
    int v[D1], r0[D1];
#pragma unroll
    for (int q = 0; q < D1; q++) {
        v[q] = vtx[(size_t)pp * D1 + q];
        r0[q] = D - rank[(size_t)pp * D1 + q];
    }
    // au[q][r] = alpha * vals[v[q]][k0 + r]
    float au[D1][R];
#pragma unroll
    for (int q = 0; q < D1; q++) {
        if (ks) {
            load_piece<R>(vals + (size_t)v[q] * ldv, G.k0, L, au[q]);
        } else {
            const float x = vals[(size_t)v[q] * ldv];
#pragma unroll
            for (int r = 0; r < R; r++) au[q][r] = x;
        }
#pragma unroll
        for (int r = 0; r < R; r++) au[q][r] = gs.alpha * au[q][r];
    }
    float cv[R];
    if (c) {
        load_piece<R>(c + (size_t)pp * ldc, G.k0, L, cv);
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) cv[r] = 1.0f;
    }
    float grad[D];
#pragma unroll
    for (int j = 0; j < D; j++) grad[j] = 0.0f;
#pragma unroll
    for (int r = 0; r < R; r++) {
        float ra[D1];
#pragma unroll
        for (int j = 0; j < D1; j++) {
            const int a = r0[j], b = a + 1 > D ? 0 : a + 1;
            float x0 = au[0][r], x1 = au[0][r];
#pragma unroll
            for (int q = 1; q < D1; q++) {
                x0 = a == q ? au[q][r] : x0;
                x1 = b == q ? au[q][r] : x1;
            }
            ra[j] = x0 - x1;
        }
        float sm = ra[0];
#pragma unroll
        for (int j = 1; j <= D; j++) {
            const float jr = (float)j * ra[j];
            const float diff = sm - jr;
            const float e = gs.sf[j - 1] * diff;
            const float prod = cv[r] * e;
            grad[j - 1] = grad[j - 1] + prod;
            sm = sm + ra[j];
        }
    }
#pragma unroll
    for (int j = 0; j < D; j++) {
        float x[R];
        x[0] = grad[j];
#pragma unroll
        for (int r = 1; r < R; r++) x[r] = 0.0f;
        grad[j] = G.all_sum(x);  // every lane takes part in the shuffles
    }
    if (active && G.l == 0) {
#pragma unroll
        for (int j = 0; j < D; j++) {
            float* o = df + (size_t)p * D + j;
            *o = accumulate ? *o + grad[j] : grad[j];
        }
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

// acc[q] += sum over blocks of partial[b][q]: one workgroup per parameter, a fixed
// strided assignment of blocks to threads and a fixed tree.
__global__ void __launch_bounds__(kThreads) k_kparam_final(const double* __restrict__ partial, int nb, int np,
                                                           double* __restrict__ acc) {
    __shared__ double red[kThreads];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kThreads) s += partial[(size_t)b * np + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[q] += red[0];
}

__global__ void k_iota(int* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
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

void launch_kg_pack(const float* b, const float* lbl, const float* norm, int N, int L, int ntype, float* P,
                    hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n) return;
    k_kg_pack<<<blocks_for(n), kThreads, 0, s>>>(b, lbl, norm, N, L, ntype, P);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kg_x(const float* b, const float* lbl, const float* fa, const float* fb, const float* norm, int N, int L,
                 int ntype, float* x, hipStream_t s) {
    const size_t n = (size_t)N * L;
    if (!n) return;
    k_kg_x<<<blocks_for(n), kThreads, 0, s>>>(b, lbl, fa, fb, norm, N, L, ntype, x);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_iota(int* out, int n, hipStream_t s) {
    if (n <= 0) return;
    k_iota<<<blocks_for(n), kThreads, 0, s>>>(out, n);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_slice_grad(const float* vals, int ldv, int ks, const int* vtx, const int* rank, const float* c, int ldc,
                       int N, int L, int d, int accumulate, float* df, hipStream_t s) {
    if (!N) return;
    GradScale gs{};
    // the reference's constants (permutohedral.cpp:620-628): double, stored as float
    const float inv_std_dev = (float)(std::sqrt(2.0 / 3.0) * (d + 1));
    for (int i = 0; i < d; i++) gs.sf[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    gs.alpha = 1.0f / (1 + std::pow(2.0f, (float)-d)) / (float)(d + 1);
    const LaunchShape sh = shape_for(L);
    if (sh.wave && d <= 6) {
        const unsigned g = blocks_for(N, sh.items);
#define DCRF_SGL(RR, DD)                                                                                       \
    case RR * 8 + DD:                                                                                          \
        k_slice_grad_lanes<RR, DD><<<g, kWaveBlock, 0, s>>>(vals, ldv, ks, vtx, rank, c, ldc, N, L, sh.S, sh.per, \
                                                           gs, accumulate, df);                                \
        break;
        switch (sh.R * 8 + d) {
            DCRF_SGL(1, 1) DCRF_SGL(1, 2) DCRF_SGL(1, 3) DCRF_SGL(1, 4) DCRF_SGL(1, 5) DCRF_SGL(1, 6)
            DCRF_SGL(2, 1) DCRF_SGL(2, 2) DCRF_SGL(2, 3) DCRF_SGL(2, 4) DCRF_SGL(2, 5) DCRF_SGL(2, 6)
            DCRF_SGL(3, 1) DCRF_SGL(3, 2) DCRF_SGL(3, 3) DCRF_SGL(3, 4) DCRF_SGL(3, 5) DCRF_SGL(3, 6)
            DCRF_SGL(4, 1) DCRF_SGL(4, 2) DCRF_SGL(4, 3) DCRF_SGL(4, 4) DCRF_SGL(4, 5) DCRF_SGL(4, 6)
            default:
                throw Error{DCRF_ERR_STATE, "slice gradient: no lane-group kernel for this shape"};
        }
#undef DCRF_SGL
        DCRF_HIP_CHECK(hipGetLastError());
        return;
    }
    const unsigned nb = blocks_for(N);
#define DCRF_SG(DD)                                                                                              \
    case DD:                                                                                                     \
        k_slice_grad<DD><<<nb, kThreads, 0, s>>>(vals, ldv, ks, vtx, rank, c, ldc, N, L, gs, accumulate, df); \
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
    k_kparam_final<<<np, kThreads, 0, s>>>(partial, nb, np, acc);
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
