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

// lbl = C (w Q_t), the term's compatibility applied to its weighted Q_t
// (PairwisePotential::kernelGradient, pairwise.cpp:202-207; st.t[0] carries the
// compatibility and the weight), and the two filter inputs of featureGradient
// (pairwise.cpp:88-114) side by side, P[p] = [a | c] (2L floats): a = b n (SYMMETRIC,
// AFTER) or b, c = lbl n (SYMMETRIC, BEFORE) or lbl.  One splat and the two blur orders
// of P give K a, K^T a, K c, K^T c.  lbl is stored only where the caller gives a buffer
// (the normaliser input x reads it).
template <bool WAVE, int R, bool CWL>
__global__ void __launch_bounds__(1024) k_kg_prepare(const float* __restrict__ Qt, const float* __restrict__ b,
                                                     const float* __restrict__ norm, int N, int L, int S, int per,
                                                     SliceTable st, int ntype, float* __restrict__ lbl,
                                                     float* __restrict__ P) {
    extern __shared__ float sh[];
    float* shc = sh + blockDim.x;
    if (CWL) stage_compat(st, L, shc);
    LaneGroup<WAVE, R> G(L, S, per, sh);
    const int p = G.item();
    const bool active = G.on && p < N;
    const SliceTerm& T = st.t[0];
    float m[R], bv[R], c[R];
    float n = 0.0f;
    if (active) {
        load_piece<R>(Qt + (size_t)p * L, G.k0, L, m);
        load_piece<R>(b + (size_t)p * L, G.k0, L, bv);
        n = norm[p];
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = bv[r] = 0.0f;
    }
    if (T.kw) {
        const float kw = T.kw[0];
#pragma unroll
        for (int r = 0; r < R; r++) m[r] = m[r] * kw;
    }
    compat_piece<WAVE, R>(T, m, c, G, shc);  // every lane: Matrix reads its group
    if (!active) return;
    if (lbl) store_piece<R>(lbl + (size_t)p * L, G.k0, L, c);
    const bool na = ntype == DCRF_NORMALIZE_SYMMETRIC || ntype == DCRF_NORMALIZE_AFTER;
    const bool nc = ntype == DCRF_NORMALIZE_SYMMETRIC || ntype == DCRF_NORMALIZE_BEFORE;
    float av[R], cv[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        av[r] = na ? bv[r] * n : bv[r];
        cv[r] = nc ? c[r] * n : c[r];
    }
    float* row = P + (size_t)p * 2 * L;
    store_piece<R>(row, G.k0, L, av);
    store_piece<R>(row + L, G.k0, L, cv);
}

// The first argument x of featureGradient's norm-derivative term, with fa = K^T a
// (st.t[0]: the reverse-blurred rows of P) and fb = K c (st.t[1]: the forward-blurred
// rows, from label L on) sliced in the same pass; a slice whose rows are null is not
// needed by the normalisation type:
//   SYMMETRIC  x = (0.5 (b fb + fa lbl)) n^3      (pairwise.cpp:90-97)
//   AFTER      x = (b fb) n^2                     (:98-105)
//   BEFORE     x = (fa lbl) n^2                   (:106-113)
template <bool WAVE, int R>
__global__ void __launch_bounds__(1024) k_kg_x(const float* __restrict__ b, const float* __restrict__ lbl,
                                               const float* __restrict__ norm, int N, int L, int S, int per,
                                               SliceTable st, int ntype, float* __restrict__ x) {
    LaneGroup<WAVE, R> G(L, S, per, nullptr);  // no group-wide operation here
    const int p = G.item();
    if (!G.on || p >= N) return;
    float fa[R] = {}, fb[R] = {}, bv[R] = {}, lv[R] = {}, v[R];
    const bool sym = ntype == DCRF_NORMALIZE_SYMMETRIC, after = ntype == DCRF_NORMALIZE_AFTER;
    if (st.t[0].vals) slice_piece<R>(st.t[0], p, G.k0, L, fa);
    if (st.t[1].vals) slice_piece<R>(st.t[1], p, G.k0, L, fb);
    if (sym || after) load_piece<R>(b + (size_t)p * L, G.k0, L, bv);
    if (!after) load_piece<R>(lbl + (size_t)p * L, G.k0, L, lv);
    const float n = norm[p];
    const float n2 = n * n;
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (sym) {
            const float p1 = bv[r] * fb[r], p2 = fa[r] * lv[r];
            const float sum = p1 + p2;
            const float h = 0.5f * sum;
            const float n3 = n2 * n;
            v[r] = h * n3;
        } else if (after) {
            const float q = bv[r] * fb[r];
            v[r] = q * n2;
        } else {
            const float q = fa[r] * lv[r];
            v[r] = q * n2;
        }
    }
    store_piece<R>(x + (size_t)p * L, G.k0, L, v);
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

// The lane's pieces of the point's d+1 vertex rows for one direction of the slicing
// gradient: au[q][r] = vals[v[q]][k0 + r] (row stride ldv), or with per_vertex the one
// value per vertex for every label.
template <int R, int D>
__device__ __forceinline__ void grad_rows(const float* __restrict__ vals, int ldv, bool per_vertex,
                                          const int (&v)[D + 1], int k0, int L, float (&au)[D + 1][R]) {
#pragma unroll
    for (int q = 0; q < D + 1; q++) {
        if (!per_vertex) {
            load_piece<R>(vals + (size_t)v[q] * ldv, k0, L, au[q]);
        } else {
            const float x = vals[v[q]];
#pragma unroll
            for (int r = 0; r < R; r++) au[q][r] = x;
        }
    }
}

// The lane's share of one direction of the slicing gradient (the loop body of
// k_slice_grad) for its R consecutive labels, from the rows grad_rows loaded: the rotated
// differences of alpha * au are formed by select, elevated and paired with cv; grad[j]
// sums the lane's labels in order.  The lane group adds the shares afterwards.
template <int R, int D>
__device__ __forceinline__ void grad_share(const float (&au_in)[D + 1][R], const int (&r0)[D + 1],
                                           const float (&cv)[R], const GradScale& gs, float* grad) {
    constexpr int D1 = D + 1;
    float au[D1][R];
#pragma unroll
    for (int q = 0; q < D1; q++)
#pragma unroll
        for (int r = 0; r < R; r++) au[q][r] = gs.alpha * au_in[q][r];
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
}

// The operands of k_kg_grad.
struct KgGrad {
    const float* fwd;   // forward-blurred rows of P (stride 2L): K a in labels 0..L-1
    const float* rev;   // reverse-blurred rows of P (stride 2L): K^T c in labels L..2L-1
    const float* P;     // [N][2L] = [a | c]
    const float* xr;    // forward-blurred rows of x (stride L); null: no normaliser term
    const float* ones;  // the reverse-blurred all-ones lattice, one value per vertex
    const float* x;     // [N][L]
    const float* f;     // the term's features in the padded point layout, d per point
    const int* pix_base;
    const int* pad_base;
    int n_inst;
    int full;  // FULL_KERNEL: d*d parameters, else d
};

// The whole per-point part of one term's kernel gradient in one pass, S lanes per point
// and R labels per lane (S <= 64: lane groups inside one wave), d <= 6:
//   g = G(a, c) = dir 0 (K a rows, paired with c) + dir 1 (K^T c rows, paired with a),
//   r = G(x, 1) = dir 0 (K x rows, paired with ones) + dir 1 (K^T 1, paired with x),
//   fg = -r + g (g alone without a normaliser term: x and its rows are not read),
// each direction summed over the group's lanes (a pairwise tree towards the group's
// first lane) before the two are added; the 4d lane shares go through one tree together.
// The parameter products (pairwise.cpp:157-161), DIAG f[q(p)][j] fg[j], FULL fg[r]
// f[q(p)][c] at c*d + r (column-major, like the reference), are added in double by one
// lane per parameter.  Per round the group leaders leave fg and the point's feature row
// in LDS, and lane q of the wave adds its product for the wave's points in point order,
// over the block's `rounds` consecutive blocks' worth of points.  At the end the waves'
// sums are added in wave order, and the block adds them to partial[blockIdx.x][q]: one
// row per block sums all reverse iterations, in iteration order.  The label order of
// each direction's sum differs from the reference's sequential one (float tolerance).
template <int R, int D>
__global__ void __launch_bounds__(kWaveBlock) k_kg_grad(const int* __restrict__ vtx, const int* __restrict__ rank,
                                                        KgGrad a, int N, int L, int S, int per, int rounds,
                                                        GradScale gs, double* __restrict__ partial) {
    constexpr int D1 = D + 1, W = kWaveBlock / 64;
    __shared__ float fgl[W][64][D];  // per wave: fg of each group's point
    __shared__ float ffl[W][64][D];  // and the point's feature row
    __shared__ double wsum[W][D * D];
    const int np = a.full ? D * D : D;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    LaneGroup<true, R> G(L, S, per, nullptr);
    const int L2 = 2 * L;
    const bool nx = a.xr != nullptr;
    // lane q < np: parameter q = (row, col) of the products fg[row] * f[col]
    const int prow = a.full ? lane % D : lane, pcol = a.full ? lane / D : lane;
    double acc = 0.0;
    const int b0 = xcd_block(blockIdx.x, gridDim.x) * rounds;
    for (int i = 0; i < rounds; i++) {  // uniform over the block: every lane takes part in every shuffle
        const int p = G.item_in(b0 + i);
        const bool active = G.on && p < N;
        const int pp = active ? p : 0;
        const bool lead = active && G.l == 0;
        int v[D1], r0[D1];
#pragma unroll
        for (int q = 0; q < D1; q++) {
            v[q] = vtx[(size_t)pp * D1 + q];
            r0[q] = D - rank[(size_t)pp * D1 + q];
        }
        // the leader's feature row of the point, for the products below
        float fv[D];
        {
            int fq = 0;
            if (lead) {
                const int inst = a.n_inst > 1 ? last_le(a.pix_base, a.n_inst, p) : 0;
                fq = a.pad_base[inst] + (p - a.pix_base[inst]);
            }
#pragma unroll
            for (int j = 0; j < D; j++) fv[j] = lead ? a.f[(size_t)fq * D + j] : 0.0f;
        }
        // shares [dir 0 of G(a, c) | dir 1 | dir 0 of G(x, 1) | dir 1], d each; a
        // direction's rows are loaded where its share is formed
        float sh[4 * D];
        {
            float u[D1][R], cv[R];
            load_piece<R>(a.P + (size_t)pp * L2 + L, G.k0, L, cv);
            grad_rows<R, D>(a.fwd, L2, false, v, G.k0, L, u);
            grad_share<R, D>(u, r0, cv, gs, sh);
            load_piece<R>(a.P + (size_t)pp * L2, G.k0, L, cv);
            grad_rows<R, D>(a.rev + L, L2, false, v, G.k0, L, u);
            grad_share<R, D>(u, r0, cv, gs, sh + D);
            if (nx) {
#pragma unroll
                for (int r = 0; r < R; r++) cv[r] = 1.0f;
                grad_rows<R, D>(a.xr, L, false, v, G.k0, L, u);
                grad_share<R, D>(u, r0, cv, gs, sh + 2 * D);
                load_piece<R>(a.x + (size_t)pp * L, G.k0, L, cv);
                grad_rows<R, D>(a.ones, 1, true, v, G.k0, L, u);
                grad_share<R, D>(u, r0, cv, gs, sh + 3 * D);
            } else {
#pragma unroll
                for (int j = 2 * D; j < 4 * D; j++) sh[j] = 0.0f;
            }
        }
        // the group's sums in its first lane (the tree of LaneGroup::all_sum)
        for (int k = 1; k < S; k <<= 1) {
            float y[4 * D];
#pragma unroll
            for (int j = 0; j < 4 * D; j++) y[j] = __shfl(sh[j], (G.base + G.l + k) & 63);
            if ((G.l & (2 * k - 1)) == 0 && G.l + k < S) {
#pragma unroll
                for (int j = 0; j < 4 * D; j++) sh[j] = sh[j] + y[j];
            }
        }
        if (lead) {
#pragma unroll
            for (int j = 0; j < D; j++) {
                const float g = sh[j] + sh[D + j];
                const float r = sh[2 * D + j] + sh[3 * D + j];
                fgl[w][G.g][j] = nx ? -r + g : g;
                ffl[w][G.g][j] = fv[j];
            }
        }
        __syncthreads();
        if (lane < np) {
            const int ng = min(per, N - (G.item_in(b0 + i) - G.g));  // the wave's points in this round
            for (int gi = 0; gi < ng; gi++) {
                float fgv = fgl[w][gi][0], fcv = ffl[w][gi][0];
#pragma unroll
                for (int j = 1; j < D; j++) {
                    fgv = prow == j ? fgl[w][gi][j] : fgv;
                    fcv = pcol == j ? ffl[w][gi][j] : fcv;
                }
                acc += (double)fgv * (double)fcv;
            }
        }
        __syncthreads();
    }
    if (lane < np) wsum[w][lane] = acc;
    __syncthreads();
    for (int q = threadIdx.x; q < np; q += blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < W; k++) s = s + wsum[k][q];
        partial[(size_t)blockIdx.x * np + q] += s;  // rows zeroed before the first iteration
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

// The x layout's segment and recombination lists from the paired layout's: a vertex row
// F_i + j of part i becomes X_i + j, a partial row Z2 + q becomes Z1 + q.
__device__ __forceinline__ int kg_map_row(const KgRowMap& m, int r) {
    if (r > m.Z2) return m.Z1 + (r - m.Z2);
    int i = 0;
    for (int k = 1; k < m.n; k++)
        if (r >= m.F[k]) i = k;  // parts in increasing F
    return m.X[i] + (r - m.F[i]);
}

__global__ void k_kg_relayout(const Chunk* __restrict__ ch, int nch, const Combine* __restrict__ cb, int ncb,
                              KgRowMap m, Chunk* __restrict__ ch1, Combine* __restrict__ cb1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nch) {
        Chunk c = ch[i];
        c.v = kg_map_row(m, c.v);
        c.out = kg_map_row(m, c.out);
        ch1[i] = c;
    } else if (i < nch + ncb) {
        Combine c = cb[i - nch];
        c.v = kg_map_row(m, c.v);
        c.first = kg_map_row(m, c.first);
        cb1[i - nch] = c;
    }
}

}  // namespace

constexpr int kKparamBlocks = 512;   // k_kparam_partial's grid
constexpr int kKgGradBlocks = 2048;  // at most this many blocks of k_kg_grad

size_t kparam_partial_doubles(int d) { return (size_t)std::max(kKparamBlocks, kKgGradBlocks) * d * d; }

// The reference's constants (permutohedral.cpp:620-628): double, stored as float.
static GradScale grad_scale(int d) {
    GradScale gs{};
    const float inv_std_dev = (float)(std::sqrt(2.0 / 3.0) * (d + 1));
    for (int i = 0; i < d; i++) gs.sf[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    gs.alpha = 1.0f / (1 + std::pow(2.0f, (float)-d)) / (float)(d + 1);
    return gs;
}

void launch_kg_prepare(const float* Qt, const float* b, const float* norm, int N, int L, const SliceTable& st_in,
                       int ntype, float* lbl, float* P, hipStream_t s) {
    if (!N) return;
    SliceTable st = st_in;
    const int cwf = assign_compat_lds(st, L);
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH_CWL(sh, cwf, k_kg_prepare, blocks_for(N, sh.items), sizeof(float) * (sh.block + cwf), s, Qt, b,
                      norm, N, L, sh.S, sh.per, st, ntype, lbl, P);
}

void launch_kg_x(const float* b, const float* lbl, const float* norm, int N, int L, const SliceTable& st, int ntype,
                 float* x, hipStream_t s) {
    if (!N) return;
    const LaunchShape sh = shape_for(L);
    DCRF_DISPATCH(sh, k_kg_x, blocks_for(N, sh.items), 0, s, b, lbl, norm, N, L, sh.S, sh.per, st, ntype, x);
}

bool kg_grad_fused(int L, int d) { return d >= 1 && d <= 6 && shape_for(L).wave; }

int kg_grad_blocks(int N, int L) {
    const int need = std::max(1, (int)blocks_for(N, shape_for(L).items));
    const int rounds = (need + kKgGradBlocks - 1) / kKgGradBlocks;
    return (need + rounds - 1) / rounds;  // nb * rounds >= need, nb <= kKgGradBlocks
}

void launch_kg_grad(const int* vtx, const int* rank, const float* fwd, const float* rev, const float* P,
                    const float* xr, const float* ones, const float* x, const float* f, const int* pix_base,
                    const int* pad_base, int n_inst, int N, int L, int d, bool full, double* partial,
                    hipStream_t s) {
    const int np = full ? d * d : d;
    if (!N || !np) return;
    if (!kg_grad_fused(L, d)) throw Error{DCRF_ERR_STATE, "kernel gradient: no lane-group kernel for this shape"};
    const LaunchShape sh = shape_for(L);
    const int need = (int)blocks_for(N, sh.items);
    const int rounds = (need + kKgGradBlocks - 1) / kKgGradBlocks;
    const int nb = kg_grad_blocks(N, L);
    const KgGrad a{fwd, rev, P, xr, ones, x, f, pix_base, pad_base, n_inst, full ? 1 : 0};
    const GradScale gs = grad_scale(d);
#define DCRF_KGG(RR, DD)                                                                                        \
    case RR * 8 + DD:                                                                                           \
        k_kg_grad<RR, DD><<<nb, kWaveBlock, 0, s>>>(vtx, rank, a, N, L, sh.S, sh.per, rounds, gs, partial); \
        break;
    switch (sh.R * 8 + d) {
        DCRF_KGG(1, 1) DCRF_KGG(1, 2) DCRF_KGG(1, 3) DCRF_KGG(1, 4) DCRF_KGG(1, 5) DCRF_KGG(1, 6)
        DCRF_KGG(2, 1) DCRF_KGG(2, 2) DCRF_KGG(2, 3) DCRF_KGG(2, 4) DCRF_KGG(2, 5) DCRF_KGG(2, 6)
        DCRF_KGG(3, 1) DCRF_KGG(3, 2) DCRF_KGG(3, 3) DCRF_KGG(3, 4) DCRF_KGG(3, 5) DCRF_KGG(3, 6)
        DCRF_KGG(4, 1) DCRF_KGG(4, 2) DCRF_KGG(4, 3) DCRF_KGG(4, 4) DCRF_KGG(4, 5) DCRF_KGG(4, 6)
        default:
            throw Error{DCRF_ERR_STATE, "kernel gradient: no lane-group kernel for this shape"};
    }
#undef DCRF_KGG
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_kparam_final(const double* partial, int rows, int np, double* acc, hipStream_t s) {
    if (rows <= 0 || np <= 0) return;
    k_kparam_final<<<np, kThreads, 0, s>>>(partial, rows, np, acc);
    DCRF_HIP_CHECK(hipGetLastError());
}

void launch_slice_grad(const float* vals, int ldv, int ks, const int* vtx, const int* rank, const float* c, int ldc,
                       int N, int L, int d, int accumulate, float* df, hipStream_t s) {
    if (!N) return;
    const GradScale gs = grad_scale(d);
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

void launch_kg_relayout(const Chunk* ch, int nch, const Combine* cb, int ncb, const KgRowMap& m, Chunk* ch1,
                        Combine* cb1, hipStream_t s) {
    const int n = nch + ncb;
    if (n <= 0) return;
    k_kg_relayout<<<blocks_for(n), kThreads, 0, s>>>(ch, nch, cb, ncb, m, ch1, cb1);
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
