// Internal declarations shared by the HIP translation units of libdcrf_hip.so.
//
// Device layout of a model (N variables over all instances, L labels):
//   Q, unary                [N][L] float, label fastest (the reference's Eigen storage)
// Per pairwise term k with feature dimension d (term-local vertex ids 0..V_k-1):
//   vtx, bary, rank         [N][d+1]   lattice vertex / barycentric weight / rank per
//                                      (pixel, remainder)   (Permutohedral offset_ ...)
//   nbr                     [d+1][V][2] blur neighbours, -1 absent
//   csr_ptr, csr_e          vertex -> its (pixel, remainder) entries, pixel ascending
//   chunks                  the splat's segments of at most `chunk` entries
// Global filter layout (built once the terms are fixed): the vertices of all terms
// are concatenated in order of decreasing d, followed by one all-zero row (the
// "missing neighbour" row) and the partial-sum rows of multi-chunk vertices.  Blur
// pass s then touches a prefix of the vertex space: the terms with d >= s.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dcrf_hip.h"

namespace dcrf {

// ---------------------------------------------------------------- errors
struct Error {
    int code;
    std::string msg;
};
void set_error(int code, const std::string& msg);
#define DCRF_HIP_CHECK(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::dcrf::Error{e_ == hipErrorOutOfMemory ? DCRF_ERR_MEMORY : DCRF_ERR_DEVICE, \
                                std::string(#expr) + ": " + hipGetErrorString(e_)};      \
    } while (0)

// ---------------------------------------------------------------- device buffers
// Caching pool over hipMalloc (pool.hip): a freed block is reused after the next device
// synchronisation the pool performs, so that no stream can still be using it.
void* pool_alloc(size_t bytes);
void pool_free(void* p);
size_t pool_release_all();

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) pool_free(p);
        p = nullptr;
        n = 0;
    }
    // (Re)allocate to hold at least `count` elements; contents are not preserved.
    void reserve(size_t count) {
        if (count <= n && p) return;
        release();
        size_t bytes = sizeof(T) * (count ? count : 1);
        p = static_cast<T*>(pool_alloc(bytes));
        n = count;
    }
    T* get() const { return p; }
};

// ---------------------------------------------------------------- host transfers
// Copies between pageable host memory and the device; those of kStagedMin bytes and
// more go through the process's page-locked staging area in chunks moved by several
// host threads.  Both return when the copy is done (after the work queued on s).
constexpr size_t kStagedMin = (size_t)8 << 20;
void copy_to_host(void* dst, const void* src, size_t n, hipStream_t s);
void copy_to_device(void* dst, const void* src, size_t n, hipStream_t s);

// ---------------------------------------------------------------- lattice of one term
struct Chunk {
    int v;    // vertex (term-local, later global)
    int beg;  // first entry
    int end;  // one past the last entry
    int out;  // output row: >= 0 vertex row; < 0 partial row -(r+1) (term-local)
};
struct Combine {
    int v;      // vertex
    int first;  // first partial row
    int count;  // number of partial rows
    int pad;    // one-launch lists: 0 for a whole vertex, else its number of pieces
};

struct TermLattice {
    int d = 0;
    int V = 0;                    // vertices over all instances
    std::vector<int> inst_vbase;  // first vertex of each instance (size n_inst+1)
    int overflow = 0;
    DevBuf<int> vtx;     // [N][d+1]
    DevBuf<float> bary;  // [N][d+1]
    DevBuf<int> rank;    // [N][d+1]
    DevBuf<int> nbr;     // [(d+1)][V][2]
    DevBuf<int16_t> keys;  // [V][d]
    DevBuf<int> csr_e;     // [N*(d+1)] entry index p*(d+1)+r, sorted by (tile of p, vertex, p)
    int n_tiles = 0;
    std::vector<int> run_beg;   // runs of equal (tile, vertex) in csr_e order (size n_runs + 1)
    std::vector<int> run_v, run_t;
    std::vector<int> tile_ptr;  // first csr_e position of each tile (size n_tiles + 1)
    // splat segments, term-local, listed tile by tile: out >= 0 vertex row, < 0 partial
    // row -(r+1); chunk_tile_ptr[t] = first segment of tile t
    std::vector<Chunk> chunks;
    std::vector<int> chunk_tile_ptr;
    std::vector<Combine> combines;  // vertices with more than one segment
    int n_partials = 0;
    int chunk_len = 0;
};

// The lattice filters of the kernel-parameter gradient (the lattice_.compute and
// Permutohedral::gradient calls of DenseKernel::featureGradient, pairwise.cpp:88-114,
// without normaliser) for up to DCRF_MAX_TERMS / 2 terms at once, in vertex rows of
// their own.  The paired inputs [a | c] (2L floats per pixel) of all the plan's terms are
// splatted by one launch and blurred forward and in reverse by shared launches: rows
// [F, F + V) and [R, R + V) hold a term's forward and reverse chain, terms with more
// blur passes first (each pass then covers a prefix of the rows), then the zero row Z2
// and the splat's partial rows.  The normaliser inputs x use a second layout with the
// forward rows only, [X, X + V), zero row Z1, partial rows.  The splat entries of the
// plan's i-th term read input row i*N + p for pixel p (inputs stacked term-major).
struct KgPart {
    int term = 0;            // index into the model's terms
    int V = 0;               // vertices of its lattice
    int F = 0, R = 0, X = 0;  // first rows of the forward, reverse and x chains
    int passes = 0;          // blur passes of the lattice
    std::vector<int> steps;  // lattice steps of each pass
    DevBuf<float> sw;        // [N][d+1] slice weights bary / (1 + 2^-d)
    const float* ones_rev = nullptr;  // [V] the all-ones input splatted and blurred in reverse
                                      // order (rows of KgPlan::ones)
};
struct KgPlan {
    std::vector<KgPart> parts;
    int gen = -1;  // the model's lattice generation the plan was built for
    int Z2 = 0, Z1 = 0, passes = 0, nch = 0, ncb = 0;
    DevBuf<int2> ent;          // (stacked input row, barycentric weight bits) in segment order
    DevBuf<Chunk> ch2, ch1;    // segments with output rows of the [a | c] and x layouts
    DevBuf<Combine> cb2, cb1;  // vertices with several segments, in the two layouts
    std::vector<int> vp2, vp1;  // per pass: one past the last row it writes
    DevBuf<int> tab;            // the pass tables of all chains
    std::vector<std::vector<int>> tab_off;  // per pass: forward, reverse, x table of each part
    DevBuf<float> S2, X2[2], S1, B1;  // splat and ping-pong rows of the two layouts
    DevBuf<float> ones[3];            // the all-ones input's splat and blur rows (one value each)
};

struct Term {
    int ktype = DCRF_DIAG_KERNEL, ntype = DCRF_NORMALIZE_SYMMETRIC, ckind = DCRF_COMPAT_POTTS;
    std::vector<float> compat;  // Potts [1], Diagonal [L], Matrix [L*L] symmetrised
    DevBuf<float> d_compat;
    TermLattice lat;
    DevBuf<float> norm;  // [N]
    DevBuf<float> feat0;         // [npad][d] features as given (DenseKernel::f_)
    std::vector<float> kparams;  // DIAG [d] scales, FULL [d*d] column-major (parameters_)
    int num_kparams() const {
        return ktype == DCRF_CONST_KERNEL ? 0 : (ktype == DCRF_DIAG_KERNEL ? lat.d : lat.d * lat.d);
    }
};

// Instance geometry of a model (a plain model is one instance).
struct Geometry {
    int n_inst = 1;
    std::vector<int> N;         // per instance
    std::vector<int> W, H;      // per instance (2-D models)
    std::vector<int> pix_base;  // size n+1
    std::vector<int> pad_base;  // size n+1, padded point bases (ceil4 per instance)
    bool is2d = false;
    int N_total() const { return pix_base.back(); }
    int npad_total() const { return pad_base.back(); }
};

// Device copies of the geometry tables.
struct GeometryDev {
    DevBuf<int> pix_base, pad_base, W;
};

// Splat tiles: pixels are grouped into tiles (T x T squares of a 2-D instance, runs of
// T*T consecutive pixels otherwise, never crossing instances); the contributors of a
// heavy vertex are cut at tile boundaries and all segments are processed tile by tile,
// so the input rows of one tile are fetched once and re-read from cache.
struct TileMap {
    int side = 0;               // T
    std::vector<int> base;      // first tile of each instance (size n_inst + 1)
    DevBuf<int> base_dev;
    int n_tiles() const { return base.empty() ? 0 : base.back(); }
};
void make_tile_map(const Geometry& g, int side, TileMap& tm);
// Default tile side (128: measured fastest at 1024x768 with 21 labels, where the
// bilateral lattice's vertices span hundreds of pixels and smaller tiles multiply the
// partial rows).
int default_tile_side(int L);

// Tile geometry as the kernels see it.
struct TileGeom {
    const int* tile_base;  // [n_inst + 1]
    const int* pix_base;   // [n_inst + 1]
    const int* W;          // 2-D: width of each instance
    int n_inst, is2d, side;
};

// ---------------------------------------------------------------- lattice build
// Builds the lattice of one term from padded point-major features feat [npad][d]
// (pad points zero).  Fills lat.{vtx,bary,rank,nbr,keys,csr_ptr,csr_e,V,inst_vbase}.
void build_lattice(const Geometry& g, const GeometryDev& gd, const TileMap& tm, const float* feat, int d,
                   TermLattice& lat, float hash_load, hipStream_t s);
// Splat segments (host): the runs of equal (tile, vertex) cut into pieces of at most
// `chunk` entries, listed tile by tile.
void build_chunks(TermLattice& lat, int chunk);
// Feature generators into padded point-major layout [npad][d].
void features_from_dxN(const Geometry& g, const GeometryDev& gd, const float* feat_dxN, int d,
                       float* out, hipStream_t s);
void features_gaussian_2d(const Geometry& g, const GeometryDev& gd, float sx, float sy,
                          float* out, hipStream_t s);
void features_bilateral_2d(const Geometry& g, const GeometryDev& gd, float sx, float sy, float sr,
                           float sg, float sb, const uint8_t* rgb, float* out, hipStream_t s);

// ---------------------------------------------------------------- filter kernels
struct BlurTerm {
    int vbase;       // first global vertex
    int V;           // vertices
    int d;           // feature dimension (axes 0..d)
    const int2* nbr; // [(d+1)][V] global neighbour rows (zero row for absent)
};
struct BlurTable {
    int nterms;
    BlurTerm t[DCRF_MAX_TERMS];
};

// A pass of K = 1..3 consecutive blur steps of one term: per vertex v the 3^K rows of
// the step tree (v and its two neighbours along the pass's last axis, each expanded
// along the axis before, ...), zero row for an absent vertex, at tab[v * 3^K ..].
struct PassTerm {
    int vbase;        // first global vertex
    int V;
    const int* tab;   // [V][3^steps]
    int steps;        // K
};
struct PassTable {
    int nterms;
    PassTerm t[DCRF_MAX_TERMS];
    int zero_row = -1;  // the tables' row for absent vertices (all zeros), or -1
};

// One term as seen by the per-pixel kernels (slice + compatibility).
struct SliceTerm {
    const int* vtx;     // [N][d+1] global vertex rows
    const float* sw;    // [N][d+1] slice weights: bary * alpha * post-normaliser
    const float* vals;  // vertex rows after the last blur pass
    int d1;
    int ckind;          // DCRF_COMPAT_*, or -1 for none
    const float* cw;    // device: Potts [1], Diagonal [L], Matrix [L*L] (symmetric)
    const float* kw;    // device: the term's weight (one float), null = 1
    float* mout;        // optional output of the raw message K~Q [N][L] (training cache)
    int ldv = 0;        // row stride of vals (floats)
    int cw_lds = -1;    // Matrix: offset (floats) of the LDS copy of cw in the kernel's compat
                        // region, or -1 to read cw from global memory (set by the launchers)
    int ldw = 0;        // row stride (floats) of that LDS copy: L, or L rounded up to 4
    // Adjoint kernels: the Potts / Diagonal statistic sum_p (K~^T b)[p] .* Q[p] of this
    // term, one row of L float sums per block at gstat[block * L] (null: none), summed
    // through LDS slot gslot of the kernel's statistic region.
    float* gstat = nullptr;
    int gslot = -1;
};
struct SliceTable {
    int nterms;
    SliceTerm t[DCRF_MAX_TERMS];
    int nstat = 0;  // terms with a statistic (gslot 0 .. nstat-1)
};
// Give each Matrix term an LDS copy of its L x L compatibility (cw_lds) while they fit
// kCompatLdsFloats in total; returns the floats of the compat LDS region.  pad: rows of
// (L + 3) & ~3 floats (16-byte aligned, for vector reads), else of L.
constexpr int kCompatLdsFloats = 8192;
int assign_compat_lds(SliceTable& st, int L, bool pad = false);

// Lane-group geometry of the per-item kernels (see dcrf_device.h).
struct LaunchShape {
    int block;  // threads per block
    int per;    // groups per wave (wave mode) or per block (block mode)
    bool wave;  // S <= 64: groups live inside one wave
    int items;  // items per block
    int R;      // labels per lane (1..4)
    int S;      // lanes per item, ceil(L / R)
};
int choose_labels_per_lane(int L);
LaunchShape shape_for(int L);
LaunchShape shape_for_r(int L, int R);  // the shape for R labels per lane

// Splat: vals[out] = sum over a segment's entries i of w_i * in[p_i], in entry order;
// ent[i] = (p_i, bits of w_i).  Global form: p_i is the pixel (row of `in`).
void launch_splat(const Chunk* chunks, int n_chunks, const int2* ent, const float* in, int ldi, float* vals,
                  int ldv, int L, hipStream_t s);
// Tiled form: one block per tile stages the tile's input rows, entries and segments in
// LDS; p_i is then the index of the entry's pixel inside its tile (row-major in the
// tile).  The entries of tile t are [tile_e[t], tile_e[t+1]), its segments
// [tile_seg[t], tile_seg[t+1]).
struct TiledSplat {
    const Chunk* chunks;
    const int* tile_seg;
    const int* tile_e;
    int n_tiles;
    TileGeom geom;
    int e_max, s_max;  // most entries / segments of one tile
};
// LDS bytes of the tiled splat (0 if it exceeds lds_cap, the LDS one workgroup may
// allocate, or L needs the block-mode lane groups: use the global form).
size_t tiled_splat_lds(int L, int side, int e_max, int s_max, size_t lds_cap);
void launch_splat_tiled(const TiledSplat& ts, const int2* ent, const float* in, float* vals, int ldv, int L,
                        size_t lds_cap, hipStream_t s);
void launch_combine(const Combine* c, int n, float* vals, int ldv, int L, hipStream_t s);
// The one-launch list (wave-mode shapes): light vertices one item each, every heavy
// vertex a whole wave of pieces (Combine::pad).
void launch_combine_waves(const Combine* c, int n, float* vals, int ldv, int L, hipStream_t s);
// One blur step over the vertices [v0, Vp) (a range of the terms with d >= step).
void launch_blur(const float* src, float* dst, int v0, int Vp, int ldv, int L, const BlurTable& bt, int step,
                 int reverse, hipStream_t s);
// One multi-step pass over the vertices [v0, Vp): bit-identical to its blur steps.
void launch_blur_pass(const float* src, float* dst, int v0, int Vp, int ldv, int L, const PassTable& pt,
                      hipStream_t s, int R = 0);
// The table of a K-step pass (axes[0] first) of one term from its global neighbour
// rows nbr [(d+1)][V].
void launch_pass_table(const int2* nbr, int V, int vbase, int zero_row, int K, const int* axes, int* tab,
                       hipStream_t s);
// Mean-field update: E = -U - sum_k compat_k(slice_k); Q = softmax(E).  U may be null.
// tmp1/tmp2 optional outputs (energy, last term's message).
void launch_meanfield(const float* U, float* Q, int N, int L, const SliceTable& st, float* tmp1,
                      float* tmp2, hipStream_t s);
// The same update for models of one term with d+1 in {3, 6} or two terms with
// (d+1) = (3, 6) / (6, 3) (fused_meanfield_ok): slc[p] holds the (vertex, weight bits)
// pairs of every term of pixel p, term 0 first (launch_interleave_slice builds it).
bool fused_meanfield_ok(int L, int nterms, int d1_0, int d1_1);
// Run form of the fused update: pixels grouped into runs (TW x TH rectangles of 2-D
// instances, runs of TW*TH consecutive pixels otherwise) and visited in run-major
// order; per run the distinct vertex rows its slice pairs read (tagged (row << 1) |
// buffer: vals v0 or v1), and the pairs with the row replaced by its index in the run's
// list.  The kernel stages a run's rows in LDS.
struct RunSliceDev {
    int P, n_runs, rows_lds;  // rows_lds: LDS floats for the largest run's rows (even)
    const int* base;          // [n_runs + 1] first distinct row of each run
    const int* rows;          // tagged rows
    const int* pix_base;      // [n_runs + 1] first position of each run in run-major order
    const int* rpix;          // [N] pixel at each run-major position
    const int2* loc;          // [N][DS] by run-major position: (index in the run's list, weight bits)
    const float* v0;
    const float* v1;
};
struct RunSlice {
    int P = 0, n_runs = 0, u_max = 0;
    DevBuf<int> base, rows, pix_base, rpix;
    DevBuf<int2> loc;
};
// par_mask bit j: entry j of a pixel's slice pairs reads the second value buffer.
void build_run_slice(const Geometry& g, const GeometryDev& gd, const int2* slc, int ds, unsigned par_mask, int TW,
                     int TH, RunSlice& rs, hipStream_t s);
size_t meanfield_runs_lds(int L, int ds, int P, int rows_lds, int cwf);
void launch_meanfield_runs(const float* U, float* Q, int N, int L, const SliceTable& st, const RunSliceDev& rs,
                           hipStream_t s);
// bw > 0: Q is written in 8 x 8 pixel blocks of an image bw pixels wide (q_block_pos,
// rows of ldq floats), for the splat over block-position entries (launch_entry_block).
void launch_meanfield_fused(const float* U, float* Q, int ldq, int N, int L, const SliceTable& st,
                            const int2* slc, hipStream_t s, int bw = 0);
void launch_entry_block(const int2* ent, size_t E, int W, int2* out, hipStream_t s);
// rows of a Q held in 8 x 8 pixel blocks for an image of W x H pixels
inline size_t q_block_rows(int W, int H) { return (size_t)((W + 7) / 8) * ((H + 7) / 8) * 64; }
void launch_interleave_slice(const int* vtx, const float* sw, int N, int d1, int ds, int off, int2* slc,
                             hipStream_t s);
// out = compat_k(slice_k) for term index `which` of st.
void launch_slice_term(float* out, int N, int L, const SliceTable& st, int which, hipStream_t s);
// Q0 = softmax(-U).
void launch_softmax_neg(const float* U, float* Q, int N, int L, hipStream_t s);
// KL divergence partial sums (one double per block).
int launch_kl(const float* Q, const float* U, int N, int L, const SliceTable& st, double* partials,
              hipStream_t s);
// norm = f(K*1) per the normalisation type, from the sliced values (L = 1).
void launch_norm_finish(float* norm, int N, int ntype, hipStream_t s);
void launch_fill(float* p, float v, size_t n, hipStream_t s);
// out[c][r] = in[r][c] (in is rows x cols row-major).
void launch_transpose(const float* in, float* out, int rows, int cols, hipStream_t s);
// out[p] = the first label of the largest Q[p][l] (DenseCRF::currentMap), Q row stride ldq.
void launch_argmax(const float* Q, int ldq, int N, int L, int16_t* out, hipStream_t s);
// ---------------------------------------------------------------- adjoint kernels
// b = sumAndNormalize(Q .* dQ, Q); dU = b.
void launch_sn_first(const float* Q, const float* dQ, float* b, float* dU, int N, int L, hipStream_t s);
// b <- sumAndNormalize(Q .* sum_k w_k C_k^T slice^T_k, Q) (b is written, not read); dU += b.
// Terms with st.t[k].gstat also get their statistic rows (see SliceTerm).
void launch_adjoint(const float* Q, float* b, float* dU, int N, int L, const SliceTable& st, hipStream_t s);
// The same over interleaved slice pairs (those of launch_meanfield_fused, with the
// transposed filter's slice weights); models that fused_meanfield_ok accepts.
void launch_adjoint_fused(const float* Q, float* b, float* dU, int N, int L, const SliceTable& st, const int2* slc,
                          hipStream_t s);
// Blocks of one adjoint launch over N pixels (the rows of its statistic output).
int adjoint_blocks(int N, int L);
// acc += sum_{t < iters} sum_p b_t[p] M_t[p]^T (L x L), with b_t = b + t * bstride and
// M_t = M + t * mstride, on the matrix cores; one launch and one reduction for all
// iterations.  partial holds grad_partial_floats(N, L, iters) floats, mid
// rows_mid_doubles(L * L) doubles.
size_t grad_partial_floats(int N, int L, int iters);
void launch_grad_stat(const float* b, size_t bstride, const float* M, size_t mstride, int iters, int N, int L,
                      float* partial, double* mid, double* acc, hipStream_t s);
// acc[j] += sum_{r < rows} part[r * cols + j] for j < cols, in double, in a fixed order
// (two kernels; mid holds rows_mid_doubles(cols) doubles).
size_t rows_mid_doubles(int cols);
void launch_sum_rows(const float* part, long long rows, int cols, double* mid, double* acc, hipStream_t s);
void launch_fill_double(double* p, size_t n, hipStream_t s);
void launch_compat_finish(const double* G, int ckind, int L, const float* cw, const float* kw, float* dcompat,
                          float* dkw, hipStream_t s);
void launch_set_compat(const float* raw, int ckind, int L, float* cw, hipStream_t s);

// ---------------------------------------------------------------- kernel-parameter gradient
// lbl (optional) = C (w Q_t) of the term in st.t[0] (its compatibility and weight), and
// P [N][2L] = [a | c], a = b (times n for SYMMETRIC / AFTER), c = lbl (times n for
// SYMMETRIC / BEFORE).
void launch_kg_prepare(const float* Qt, const float* b, const float* norm, int N, int L, const SliceTable& st,
                       int ntype, float* lbl, float* P, hipStream_t s);
// The normaliser-derivative input x of featureGradient for ntype, with fa sliced by
// st.t[0] and fb by st.t[1] (a term with null vals: not needed).
void launch_kg_x(const float* b, const float* lbl, const float* norm, int N, int L, const SliceTable& st, int ntype,
                 float* x, hipStream_t s);
// Whether launch_kg_grad has a kernel for L labels and feature dimension d.
bool kg_grad_fused(int L, int d);
// The parameter gradient of fg = -G(x, 1) + G(a, c) (xr null: fg = G(a, c)) against
// the padded feature table f, in one pass over the points: each of the
// kg_grad_blocks(N, L) blocks adds its sums to its row of np = d (or d*d) doubles in
// partial, rows that the caller zeroes before the first reverse iteration; then
// launch_kparam_final(partial, kg_grad_blocks(N, L), np, acc) adds them to acc.
int kg_grad_blocks(int N, int L);
void launch_kg_grad(const int* vtx, const int* rank, const float* fwd, const float* rev, const float* P,
                    const float* xr, const float* ones, const float* x, const float* f, const int* pix_base,
                    const int* pad_base, int n_inst, int N, int L, int d, bool full, double* partial,
                    hipStream_t s);
// acc[q] += sum over rows b < rows of partial[b * np + q], for q < np (fixed order).
void launch_kparam_final(const double* partial, int rows, int np, double* acc, hipStream_t s);
// One direction of the lattice gradient: df (+)= slicing gradient of the blurred rows
// vals (row stride ldv, label stride ks) paired with c [N][L] (row stride ldc; null:
// ones).
void launch_slice_grad(const float* vals, int ldv, int ks, const int* vtx, const int* rank, const float* c, int ldc,
                       int N, int L, int d, int accumulate, float* df, hipStream_t s);
// acc (d or d*d doubles) += parameter gradient of fg = (r ? -r + g : g) against the
// padded feature table f; partial holds kparam_partial_doubles(d).
size_t kparam_partial_doubles(int d);
void launch_kparam_reduce(const float* r, const float* g, const float* f, const int* pix_base, const int* pad_base,
                          int n_inst, int N, int d, bool full, double* partial, double* acc, hipStream_t s);
// The row numbering of a kernel-gradient plan's x layout (KgPlan): part i's vertex rows
// start at F[i] in the paired layout and at X[i] in the x layout; partial rows follow
// the zero rows Z2 and Z1.
struct KgRowMap {
    int n, Z2, Z1;
    int F[DCRF_MAX_TERMS / 2], X[DCRF_MAX_TERMS / 2];
};
// ch1 / cb1 = ch / cb with their rows in the x layout.
void launch_kg_relayout(const Chunk* ch, int nch, const Combine* cb, int ncb, const KgRowMap& m, Chunk* ch1,
                        Combine* cb1, hipStream_t s);
// out[i] = (float)acc[i].
void launch_kparam_store(const double* acc, int n, float* out, hipStream_t s);
// Scaled features diag(P) f or P f (P column-major) over the padded table.
void launch_scale_features(const float* f, int npad, int d, bool full, const float* P, float* out, hipStream_t s);

// ---------------------------------------------------------------- logistic unary
// U [N][M] = (Lm [M][F] fFN [F][N])^T on the matrix cores.
void launch_logistic_unary(const float* fFN, const float* Lm, int N, int F, int M, float* U, hipStream_t s);
// Unary helpers on the device, output [N][L].
void launch_unary_from_softmax(const float* sm, int L, int N, double scale, bool use_scale, double clip,
                               bool use_clip, float* U, hipStream_t s);
void launch_unary_from_labels(const int* labels, int N, int L, float p_energy, float n_energy, float u_energy,
                              bool zero_unsure, float* U, hipStream_t s);
// out [LA][LB] (double) = sum_p A[p][a] B[p][c], A [N][LA], B [N][LB]; partial holds
// gram_partial_floats(LA, LB, N) floats.
size_t gram_partial_floats(int LA, int LB, int N);
// ---------------------------------------------------------------- objectives
// ObjectiveFunction::evaluate on the device (DCRF_OBJECTIVE_*): dQ [N][L] written whole;
// partial holds objective_partials(kind, N, L) doubles: for LogLikelihood / Hamming the
// per-block sums of r (the first objective_partials entries, added in order), for IoU
// the block statistics followed by the final [L][2] (in_l, un_l) at offset
// objective_partials - 2 L.  w: Hamming's class weights on the device (one per label
// that occurs in gt).
int objective_partials(int kind, int N, int L);
void launch_objective(int kind, const float* Q, const int16_t* gt, int N, int L, float param, const float* w,
                      float* dQ, double* partial, hipStream_t s);
void launch_gram(const float* A, int LA, const float* B, int LB, int N, float* partial, double* out, hipStream_t s);

// Global layout remaps.
void launch_remap_nbr(const int* src, int2* dst, int n, int vbase, int zero_row, hipStream_t s);
// Entry arrays of one term: sorted entry j (csr_e[j] = p*(d+1)+r) goes to position
// base[t] + j, t the tile of j (tile_ptr: first j of each tile; base null: position j);
// ent = (the pixel p, or its index inside its tile when `local` is given, plus pix_off;
// bits of the barycentric weight, times norm[p] when use_norm).
void launch_entry_gather(const int* csr_e, int E, int d1, const float* bary, const float* norm, int use_norm,
                         const int* tile_ptr, const int* base, int n_tiles, const TileGeom* local, int2* ent_out,
                         hipStream_t s, int pix_off = 0);
void launch_slice_weights(const int* vtx, const float* bary, const float* norm, int use_norm,
                          float alpha, int N, int d1, int vbase, int* gvtx, float* sw,
                          hipStream_t s);

}  // namespace dcrf
