// Parallel permutohedral-lattice build for gfx950.
//
// Replaces Permutohedral::init (SSE2 form, pydensecrf/densecrf/src/permutohedral.cpp:
// 140-321).  Stages, each one kernel over all instances of a model:
//   1. features     -> padded point-major [npad][d] (pad lanes of every instance are 0.0,
//                      like the reference's 4-wide blocks, :191-196)
//   2. embed        -> per (point, remainder): lattice key (d int16 coordinates plus the
//                      instance id), barycentric weight and rank; the float op order of
//                      :199-263 is kept (contraction off) so results are bit-identical
//   3. hash insert  -> lock-free open addressing; each slot keeps the smallest entry index
//                      of its key (atomicMin), i.e. the key's first occurrence
//   4. number       -> scan over "entry is its key's first occurrence" flags: vertex ids in
//                      first-occurrence order, which is the reference hash table's
//                      insertion order (:103-111, :272), so offsets match bit for bit
//   5. neighbours   -> per (vertex, axis) lookups of key +- (d+1)e_j style moves (:305-318)
//   6. CSR          -> stable radix sort of entries by vertex: each vertex's contributors
//                      in ascending pixel order, for the atomic-free splat
#include "dcrf_internal.h"

#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>

#pragma clang fp contract(off)

namespace dcrf {
namespace {

constexpr int kEmpty = 0x7fffffff;
constexpr int kThreads = 256;

inline unsigned blocks_for(size_t n, int t = kThreads) { return (unsigned)((n + t - 1) / t); }

__device__ inline int find_instance(const int* base, int n_inst, int q) {
    int lo = 0, hi = n_inst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------------ features
__global__ void k_feat_dxN(const float* __restrict__ f, int d, int N_total, const int* pix_base,
                           const int* pad_base, int n_inst, int npad, float* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npad) return;
    const int i = n_inst > 1 ? find_instance(pad_base, n_inst, q) : 0;
    const int local = q - pad_base[i];
    const int Ni = pix_base[i + 1] - pix_base[i];
    float* o = out + (size_t)q * d;
    if (local < Ni) {
        const int p = pix_base[i] + local;
        for (int j = 0; j < d; j++) o[j] = f[(size_t)j * N_total + p];
    } else {
        for (int j = 0; j < d; j++) o[j] = 0.0f;
    }
}

// DenseCRF2D::addPairwiseGaussian / addPairwiseBilateral features (densecrf.cpp:61-81):
// (x/sx, y/sy[, r/sr, g/sg, b/sb]) with IEEE float division.
__global__ void k_feat_2d(const int* pix_base, const int* pad_base, const int* Wd, int n_inst,
                          int npad, float sx, float sy, const uint8_t* __restrict__ rgb, float sr,
                          float sg, float sb, float* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npad) return;
    const int i = n_inst > 1 ? find_instance(pad_base, n_inst, q) : 0;
    const int local = q - pad_base[i];
    const int Ni = pix_base[i + 1] - pix_base[i];
    const int d = rgb ? 5 : 2;
    float* o = out + (size_t)q * d;
    if (local < Ni) {
        const int W = Wd[i];
        const int x = local % W, y = local / W;
        o[0] = (float)x / sx;
        o[1] = (float)y / sy;
        if (rgb) {
            const size_t p = (size_t)pix_base[i] + local;
            o[2] = (float)rgb[p * 3 + 0] / sr;
            o[3] = (float)rgb[p * 3 + 1] / sg;
            o[4] = (float)rgb[p * 3 + 2] / sb;
        }
    } else {
        for (int j = 0; j < d; j++) o[j] = 0.0f;
    }
}

// ------------------------------------------------------------------ keys
template <int D>
struct Key {
    static constexpr int W = (D + 4) / 4;  // d int16 coordinates + the instance id
    uint64_t w[W];
};

template <int D>
__device__ inline uint32_t key_hash(const Key<D>& k) {
    uint64_t h = 0x243F6A8885A308D3ull;
#pragma unroll
    for (int i = 0; i < Key<D>::W; i++) {
        h ^= k.w[i];
        h *= 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
    }
    return (uint32_t)(h ^ (h >> 17));
}

template <int D>
__device__ inline bool key_eq(const Key<D>& a, const Key<D>& b) {
    bool eq = true;
#pragma unroll
    for (int i = 0; i < Key<D>::W; i++) eq = eq && (a.w[i] == b.w[i]);
    return eq;
}

template <int D>
__device__ inline void key_set(Key<D>& k, int i, int v) {
    const uint64_t m = (uint64_t)0xffff << ((i & 3) * 16);
    k.w[i >> 2] = (k.w[i >> 2] & ~m) | ((uint64_t)(uint16_t)v << ((i & 3) * 16));
}

template <int D>
__device__ inline int key_get(const Key<D>& k, int i) {
    return (int)(int16_t)(uint16_t)(k.w[i >> 2] >> ((i & 3) * 16));
}

struct ScaleArg {
    float s[DCRF_MAX_DIM];
    float inv_d1;
};

// Permutohedral::init per point (permutohedral.cpp:199-276); see the file header.
template <int D>
__global__ void __launch_bounds__(kThreads) k_embed(const float* __restrict__ feat, int npad,
                                                    const int* pix_base, const int* pad_base,
                                                    int n_inst, ScaleArg sc, Key<D>* __restrict__ keys,
                                                    float* __restrict__ bary, int* __restrict__ rank,
                                                    int* __restrict__ overflow) {
    constexpr int D1 = D + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npad) return;
    const int inst = n_inst > 1 ? find_instance(pad_base, n_inst, q) : 0;
    const int local = q - pad_base[inst];
    const bool real = local < pix_base[inst + 1] - pix_base[inst];
    const int p = pix_base[inst] + local;

    float f[D];
#pragma unroll
    for (int j = 0; j < D; j++) f[j] = feat[(size_t)q * D + j];
    // elevate (:199-207)
    float elev[D1];
    float sm = 0.0f;
#pragma unroll
    for (int j = D; j > 0; j--) {
        const float cf = f[j - 1] * sc.s[j - 1];
        const float jc = (float)j * cf;
        elev[j] = sm - jc;
        sm = sm + cf;
    }
    elev[0] = sm;
    // closest remainder-0 point, round half to even (:209-221)
    const float fd1 = (float)D1;
    float rem0[D1];
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i <= D; i++) {
        float v = sc.inv_d1 * elev[i];
        v = rintf(v);
        rem0[i] = v * fd1;
        sum = sum + v;
    }
    // rank (:223-233)
    float rk[D1];
#pragma unroll
    for (int i = 0; i <= D; i++) rk[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const float di = elev[i] - rem0[i];
#pragma unroll
        for (int j = i + 1; j <= D; j++) {
            const float dj = elev[j] - rem0[j];
            const float c = di < dj ? 1.0f : 0.0f;
            rk[i] = rk[i] + c;
            rk[j] = rk[j] + (1.0f - c);
        }
    }
    // back onto the plane (:235-242)
#pragma unroll
    for (int i = 0; i <= D; i++) {
        rk[i] = rk[i] + sum;
        const float add = rk[i] < 0.0f ? fd1 : 0.0f;
        const float sub = rk[i] >= fd1 ? fd1 : 0.0f;
        rk[i] = rk[i] + (add - sub);
        rem0[i] = rem0[i] + (add - sub);
    }
    // barycentric (:244-263).  Position D-rank_i receives +v_i and position D-rank_i+1
    // receives -v_i; with ranks a permutation each slot gets at most one of each, and
    // (0 + a) - b == (0 - b) + a in float, so slot r holds v_{rank=D-r} - v_{rank=D-r+1}.
    float vr[D1];
#pragma unroll
    for (int r = 0; r <= D; r++) vr[r] = 0.0f;
#pragma unroll
    for (int i = 0; i <= D; i++) {
        const float e = elev[i] - rem0[i];
        const float v = e * sc.inv_d1;
        const int ri = (int)rk[i];
#pragma unroll
        for (int r = 0; r <= D; r++) vr[r] = (ri == r) ? v : vr[r];
    }
    float bc[D1];
#pragma unroll
    for (int r = 1; r <= D; r++) bc[r] = vr[D - r] - vr[D - r + 1];
    {
        const float last = 0.0f - vr[0];  // slot D+1 holds -v_{rank=0}
        bc[0] = vr[D] + (1.0f + last);
    }
    // vertices of the enclosing simplex (:266-276):
    // key_i(r) = rem0_i + canonical[r][rank_i], canonical[r][k] = r (k <= D-r), r-D-1 (else)
    bool ovf = false;
#pragma unroll
    for (int r = 0; r <= D; r++) {
        Key<D> k;
#pragma unroll
        for (int w = 0; w < Key<D>::W; w++) k.w[w] = 0;
#pragma unroll
        for (int i = 0; i < D; i++) {
            const int ri = (int)rk[i];
            const float c = (float)(ri <= D - r ? r : r - D1);
            const float kv = rem0[i] + c;
            const int ki = (int)kv;
            ovf = ovf || ki > 32767 || ki < -32768;
            key_set<D>(k, i, ki);
        }
        key_set<D>(k, D, inst);
        keys[(size_t)q * D1 + r] = k;
        if (real) {
            bary[(size_t)p * D1 + r] = bc[r];
            rank[(size_t)p * D1 + r] = (int)rk[r];
        }
    }
    if (ovf) overflow[threadIdx.x & 63] = 1;
}

template <int D>
__global__ void __launch_bounds__(kThreads) k_hash_insert(const Key<D>* __restrict__ keys, int E,
                                                          int* table, uint32_t mask,
                                                          int* __restrict__ slot) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const Key<D> k = keys[e];
    uint32_t h = key_hash<D>(k) & mask;
    for (;;) {
        // a workgroup-scope read may return a stale value from the CU's cache, which is
        // safe: a slot only goes from empty to entries of one key, whose value then only
        // falls, so a stale non-empty value names the right key and is >= the current one
        // (the e < cur test below stays exact), and a stale empty value ends in the CAS,
        // which reads the slot at device scope
        int cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == kEmpty) {
            const int old = atomicCAS(&table[h], kEmpty, e);
            if (old == kEmpty) {
                slot[e] = (int)h;
                return;
            }
            cur = old;
        }
        if (key_eq<D>(keys[cur], k)) {
            // the slot keeps the smallest entry of its key (first occurrence); the value
            // only falls, so an entry above the one just read cannot lower it: most
            // entries of a popular vertex skip the atomic
            if (e < cur) atomicMin(&table[h], e);
            slot[e] = (int)h;
            return;
        }
        h = (h + 1) & mask;
    }
}

template <int D>
__device__ inline int hash_find(const Key<D>* __restrict__ keys, const int* __restrict__ table,
                                uint32_t mask, const Key<D>& k) {
    uint32_t h = key_hash<D>(k) & mask;
    for (;;) {
        const int cur = table[h];
        if (cur == kEmpty) return -1;
        if (key_eq<D>(keys[cur], k)) return cur;
        h = (h + 1) & mask;
    }
}

__global__ void k_mark_first(const int* __restrict__ table, const int* __restrict__ slot, int E,
                             int* __restrict__ flag) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    flag[e] = table[slot[e]] == e ? 1 : 0;
}

__global__ void k_assign(const int* __restrict__ table, const int* __restrict__ slot,
                         const int* __restrict__ vid, int E, int d1, const int* pix_base,
                         const int* pad_base, int n_inst, int* __restrict__ vtx,
                         int* __restrict__ rep) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int first = table[slot[e]];
    const int v = vid[first];
    if (first == e) rep[v] = e;
    const int q = e / d1, r = e - q * d1;
    const int inst = n_inst > 1 ? find_instance(pad_base, n_inst, q) : 0;
    const int local = q - pad_base[inst];
    if (local < pix_base[inst + 1] - pix_base[inst]) vtx[(size_t)(pix_base[inst] + local) * d1 + r] = v;
}

// Blur neighbours (permutohedral.cpp:296-318): along axis j, n1 = key - 1 with
// n1[j] = key[j] + d and n2 = key + 1 with n2[j] = key[j] - d; axis d moves every
// stored coordinate.  Also exports the vertex keys.
template <int D>
__global__ void __launch_bounds__(kThreads) k_neighbors(const Key<D>* __restrict__ keys,
                                                        const int* __restrict__ table, uint32_t mask,
                                                        const int* __restrict__ vid,
                                                        const int* __restrict__ rep, int V,
                                                        int* __restrict__ nbr,
                                                        int16_t* __restrict__ keys_out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)V * (D + 1)) return;
    const int j = (int)(t / V), v = (int)(t - (long long)j * V);
    const Key<D> k = keys[rep[v]];
    Key<D> n1 = k, n2 = k;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const int c = key_get<D>(k, i);
        key_set<D>(n1, i, i == j ? c + D : c - 1);
        key_set<D>(n2, i, i == j ? c - D : c + 1);
        if (j == 0) keys_out[(size_t)v * D + i] = (int16_t)c;
    }
    const int f1 = hash_find<D>(keys, table, mask, n1);
    const int f2 = hash_find<D>(keys, table, mask, n2);
    nbr[2 * (size_t)t] = f1 >= 0 ? vid[f1] : -1;
    nbr[2 * (size_t)t + 1] = f2 >= 0 ? vid[f2] : -1;
}

__global__ void k_fill_int(int* p, int v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Sort key of a real entry: (tile of its pixel << vbits) | vertex.
__global__ void k_entry_keys(const int* __restrict__ vtx, int Er, int d1, const int* pix_base, const int* Wd,
                             const int* tile_base, int n_inst, int is2d, int side, int vbits,
                             unsigned long long* __restrict__ key) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Er) return;
    const int p = e / d1;
    const int i = n_inst > 1 ? find_instance(pix_base, n_inst, p) : 0;
    const int local = p - pix_base[i];
    int tile;
    if (is2d) {
        const int W = Wd[i];
        const int ntx = (W + side - 1) / side;
        tile = tile_base[i] + (local / W / side) * ntx + (local % W) / side;
    } else {
        tile = tile_base[i] + local / (side * side);
    }
    key[e] = ((unsigned long long)(unsigned)tile << vbits) | (unsigned)vtx[e];
}

// Starts of the runs of equal keys in a sorted array.
__global__ void k_run_flags(const unsigned long long* __restrict__ key, int n, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ void k_run_compact(const unsigned long long* __restrict__ key, const int* __restrict__ flag,
                              const int* __restrict__ rid, int n, int* __restrict__ beg,
                              unsigned long long* __restrict__ rkey) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    beg[rid[i]] = i;
    rkey[rid[i]] = key[i];
}

__global__ void k_gather_int(const int* __restrict__ src, const int* __restrict__ idx, int n,
                             int* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

uint32_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    if (p > (1ull << 31)) throw Error{DCRF_ERR_VALUE, "lattice hash table would exceed 2^31 slots"};
    return (uint32_t)p;
}

template <int D>
void build_lattice_d(const Geometry& g, const GeometryDev& gd, const TileMap& tm, const float* feat,
                     TermLattice& lat, float hash_load, hipStream_t s) {
    constexpr int D1 = D + 1;
    const int npad = g.npad_total();
    const int N = g.N_total();
    const long long E_ll = (long long)npad * D1;
    if (E_ll > 0x7ffffff0ll) throw Error{DCRF_ERR_VALUE, "too many lattice entries for one term"};
    const int E = (int)E_ll;
    const int Er = N * D1;
    lat.d = D;

    // scale factors exactly as permutohedral.cpp:176-179 (double, stored as float)
    ScaleArg sc{};
    const float inv_std_dev = (float)(std::sqrt(2.0 / 3.0) * D1);
    for (int i = 0; i < D; i++) sc.s[i] = (float)(1.0 / std::sqrt((double)((i + 2) * (i + 1))) * inv_std_dev);
    sc.inv_d1 = 1.0f / (float)D1;

    DevBuf<Key<D>> keys;
    keys.reserve(E);
    lat.vtx.reserve((size_t)Er);
    lat.bary.reserve((size_t)Er);
    lat.rank.reserve((size_t)Er);
    DevBuf<int> ovf;
    ovf.reserve(64);
    k_fill_int<<<1, 64, 0, s>>>(ovf.get(), 0, 64);
    DCRF_HIP_CHECK(hipGetLastError());
    k_embed<D><<<blocks_for(npad), kThreads, 0, s>>>(feat, npad, gd.pix_base.get(), gd.pad_base.get(),
                                                     g.n_inst, sc, keys.get(), lat.bary.get(),
                                                     lat.rank.get(), ovf.get());
    DCRF_HIP_CHECK(hipGetLastError());

    if (!(hash_load > 0.0f && hash_load <= 0.5f)) hash_load = 0.5f;
    const uint32_t cap = next_pow2((uint64_t)std::ceil((double)E / hash_load));
    DevBuf<int> table, slot, flag, vid;
    table.reserve(cap);
    slot.reserve(E);
    flag.reserve(E);
    vid.reserve(E);
    k_fill_int<<<blocks_for(cap), kThreads, 0, s>>>(table.get(), kEmpty, cap);
    DCRF_HIP_CHECK(hipGetLastError());
    k_hash_insert<D><<<blocks_for(E), kThreads, 0, s>>>(keys.get(), E, table.get(), cap - 1, slot.get());
    DCRF_HIP_CHECK(hipGetLastError());
    k_mark_first<<<blocks_for(E), kThreads, 0, s>>>(table.get(), slot.get(), E, flag.get());
    DCRF_HIP_CHECK(hipGetLastError());
    DevBuf<char> scan_tmp;  // lives until the synchronisation below
    {
        size_t tmp_bytes = 0;
        DCRF_HIP_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, flag.get(), vid.get(), 0, (size_t)E,
                                               rocprim::plus<int>(), s));
        scan_tmp.reserve(tmp_bytes);
        DCRF_HIP_CHECK(rocprim::exclusive_scan(scan_tmp.get(), tmp_bytes, flag.get(), vid.get(), 0, (size_t)E,
                                               rocprim::plus<int>(), s));
    }
    int last[2];
    DCRF_HIP_CHECK(hipMemcpyAsync(&last[0], vid.get() + (E - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP_CHECK(hipMemcpyAsync(&last[1], flag.get() + (E - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    int ovf_h[64];
    DCRF_HIP_CHECK(hipMemcpyAsync(ovf_h, ovf.get(), sizeof(ovf_h), hipMemcpyDeviceToHost, s));
    DCRF_HIP_CHECK(hipStreamSynchronize(s));
    const int V = last[0] + last[1];
    lat.V = V;
    lat.overflow = 0;
    for (int i = 0; i < 64; i++) lat.overflow |= ovf_h[i];

    DevBuf<int> rep;
    rep.reserve(V);
    k_assign<<<blocks_for(E), kThreads, 0, s>>>(table.get(), slot.get(), vid.get(), E, D1, gd.pix_base.get(),
                                                gd.pad_base.get(), g.n_inst, lat.vtx.get(), rep.get());
    DCRF_HIP_CHECK(hipGetLastError());
    lat.nbr.reserve((size_t)2 * D1 * V);
    lat.keys.reserve((size_t)V * D);
    k_neighbors<D><<<blocks_for((size_t)V * D1), kThreads, 0, s>>>(keys.get(), table.get(), cap - 1, vid.get(),
                                                                   rep.get(), V, lat.nbr.get(), lat.keys.get());
    DCRF_HIP_CHECK(hipGetLastError());

    // first vertex of every instance: the first entry of an instance is always new (one
    // instance: vertex 0, no device round trip)
    if (g.n_inst == 1) {
        lat.inst_vbase.assign({0, V});
    } else {
        std::vector<int> idx(g.n_inst);
        for (int i = 0; i < g.n_inst; i++) idx[i] = g.pad_base[i] * D1;
        DevBuf<int> didx, dout;
        didx.reserve(g.n_inst);
        dout.reserve(g.n_inst);
        DCRF_HIP_CHECK(hipMemcpyAsync(didx.get(), idx.data(), sizeof(int) * g.n_inst, hipMemcpyHostToDevice, s));
        k_gather_int<<<blocks_for(g.n_inst), kThreads, 0, s>>>(vid.get(), didx.get(), g.n_inst, dout.get());
        DCRF_HIP_CHECK(hipGetLastError());
        lat.inst_vbase.assign(g.n_inst + 1, V);
        DCRF_HIP_CHECK(hipMemcpyAsync(lat.inst_vbase.data(), dout.get(), sizeof(int) * g.n_inst,
                                      hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
    }

    // Entries sorted by (tile of the pixel, vertex), stably (pixel order within a key),
    // and the runs of equal (tile, vertex): the splat segments are cut from the runs.
    lat.csr_e.reserve((size_t)Er);
    lat.n_tiles = tm.n_tiles();
    {
        int vbits = 1;
        while ((1ll << vbits) <= (long long)V) vbits++;
        int tbits = 1;
        while ((1ll << tbits) <= (long long)lat.n_tiles) tbits++;
        if (vbits + tbits > 63) throw Error{DCRF_ERR_VALUE, "too many lattice vertices and tiles"};
        DevBuf<int> iota, flag2, rid;
        DevBuf<unsigned long long> key, sorted_key;
        iota.reserve(Er);
        key.reserve(Er);
        sorted_key.reserve(Er);
        k_iota<<<blocks_for(Er), kThreads, 0, s>>>(iota.get(), Er);
        DCRF_HIP_CHECK(hipGetLastError());
        k_entry_keys<<<blocks_for(Er), kThreads, 0, s>>>(lat.vtx.get(), Er, D1, gd.pix_base.get(), gd.W.get(),
                                                         tm.base_dev.get(), g.n_inst, g.is2d ? 1 : 0, tm.side,
                                                         vbits, key.get());
        DCRF_HIP_CHECK(hipGetLastError());
        size_t tmp_bytes = 0;
        DCRF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, key.get(), sorted_key.get(), iota.get(),
                                                 lat.csr_e.get(), (size_t)Er, 0, vbits + tbits, s));
        DevBuf<char> tmp;
        tmp.reserve(tmp_bytes);
        DCRF_HIP_CHECK(rocprim::radix_sort_pairs(tmp.get(), tmp_bytes, key.get(), sorted_key.get(), iota.get(),
                                                 lat.csr_e.get(), (size_t)Er, 0, vbits + tbits, s));
        flag2.reserve(Er);
        rid.reserve(Er);
        k_run_flags<<<blocks_for(Er), kThreads, 0, s>>>(sorted_key.get(), Er, flag2.get());
        DCRF_HIP_CHECK(hipGetLastError());
        size_t scan_bytes = 0;
        DCRF_HIP_CHECK(rocprim::exclusive_scan(nullptr, scan_bytes, flag2.get(), rid.get(), 0, (size_t)Er,
                                               rocprim::plus<int>(), s));
        DevBuf<char> stmp;
        stmp.reserve(scan_bytes);
        DCRF_HIP_CHECK(rocprim::exclusive_scan(stmp.get(), scan_bytes, flag2.get(), rid.get(), 0, (size_t)Er,
                                               rocprim::plus<int>(), s));
        int tail[2];
        DCRF_HIP_CHECK(hipMemcpyAsync(&tail[0], rid.get() + (Er - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipMemcpyAsync(&tail[1], flag2.get() + (Er - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        const int n_runs = tail[0] + tail[1];
        DevBuf<int> rbeg;
        DevBuf<unsigned long long> rkey;
        rbeg.reserve(n_runs);
        rkey.reserve(n_runs);
        k_run_compact<<<blocks_for(Er), kThreads, 0, s>>>(sorted_key.get(), flag2.get(), rid.get(), Er, rbeg.get(),
                                                          rkey.get());
        DCRF_HIP_CHECK(hipGetLastError());
        lat.run_beg.resize(n_runs + 1);
        std::vector<unsigned long long> hk(n_runs);
        copy_to_host(lat.run_beg.data(), rbeg.get(), sizeof(int) * n_runs, s);
        copy_to_host(hk.data(), rkey.get(), sizeof(unsigned long long) * n_runs, s);
        lat.run_beg[n_runs] = Er;
        lat.run_v.resize(n_runs);
        lat.run_t.resize(n_runs);
        const unsigned long long vmask = (1ull << vbits) - 1;
        for (int r = 0; r < n_runs; r++) {
            lat.run_v[r] = (int)(hk[r] & vmask);
            lat.run_t[r] = (int)(hk[r] >> vbits);
        }
        // first sorted entry of every tile
        lat.tile_ptr.assign(lat.n_tiles + 1, Er);
        for (int r = n_runs - 1; r >= 0; r--) lat.tile_ptr[lat.run_t[r]] = lat.run_beg[r];
        for (int t = lat.n_tiles - 1; t >= 0; t--) lat.tile_ptr[t] = std::min(lat.tile_ptr[t], lat.tile_ptr[t + 1]);
    }
}

// Run of pixel p: a TW x TH rectangle of its 2-D instance (row-major over the image), or
// a run of TW*TH consecutive pixels of other instances.
__device__ inline int run_of(int p, const int* pix_base, const int* Wd, const int* run_base, int n_inst, int is2d,
                             int TW, int TH) {
    const int i = n_inst > 1 ? find_instance(pix_base, n_inst, p) : 0;
    const int local = p - pix_base[i];
    if (is2d) {
        const int W = Wd[i];
        const int y = local / W, x = local - y * W;
        const int ntx = (W + TW - 1) / TW;
        return run_base[i] + (y / TH) * ntx + x / TW;
    }
    return run_base[i] + local / (TW * TH);
}

// Pixel keys (run << 32) | p: sorted, they give the run-major pixel order.
__global__ void k_rs_pixkeys(int N, const int* pix_base, const int* Wd, const int* run_base, int n_inst, int is2d,
                             int TW, int TH, unsigned long long* __restrict__ key) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    key[p] = ((unsigned long long)(unsigned)run_of(p, pix_base, Wd, run_base, n_inst, is2d, TW, TH) << 32) | (unsigned)p;
}

// rpix[k] = the k-th pixel in run-major order, pos[p] = k, pbase[run] = first k of the run.
__global__ void k_rs_order(const unsigned long long* __restrict__ skey, int N, int* __restrict__ rpix,
                           int* __restrict__ pos, int* __restrict__ pbase) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int p = (int)(unsigned)(skey[k] & 0xffffffffu);
    const int run = (int)(skey[k] >> 32);
    rpix[k] = p;
    pos[p] = k;
    if (k == 0 || (int)(skey[k - 1] >> 32) != run) pbase[run] = k;
}

// Run slice: key of slice pair e = p*ds + j: (run of p << 32) | ((vertex << 1) | buffer).
__global__ void k_rs_keys(const int2* __restrict__ slc, int n, int ds, unsigned par_mask, const int* pix_base,
                          const int* Wd, const int* run_base, int n_inst, int is2d, int TW, int TH,
                          unsigned long long* __restrict__ key) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int p = e / ds, j = e - p * ds;
    const unsigned tag = ((unsigned)slc[e].x << 1) | ((par_mask >> j) & 1u);
    key[e] = ((unsigned long long)(unsigned)run_of(p, pix_base, Wd, run_base, n_inst, is2d, TW, TH) << 32) | tag;
}

__global__ void k_rs_compact(const unsigned long long* __restrict__ key, const int* __restrict__ flag,
                             const int* __restrict__ uid, int n, int* __restrict__ rows, int* __restrict__ base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    rows[uid[i]] = (int)(unsigned)(key[i] & 0xffffffffu);
    if (i == 0 || (key[i] >> 32) != (key[i - 1] >> 32)) base[(int)(key[i] >> 32)] = uid[i];
}

// uid is the exclusive scan of the key-start flags: the distinct-key index of sorted
// entry i is uid[i] + flag[i] - 1.
__global__ void k_rs_loc(const unsigned long long* __restrict__ key, const int* __restrict__ perm,
                         const int* __restrict__ uid, const int* __restrict__ flag, const int* __restrict__ base,
                         const int2* __restrict__ slc, const int* __restrict__ pos, int ds, int n,
                         int2* __restrict__ loc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = perm[i];
    const int p = e / ds, j = e - p * ds;
    loc[(size_t)pos[p] * ds + j] = make_int2(uid[i] + flag[i] - 1 - base[(int)(key[i] >> 32)], slc[e].y);
}

}  // namespace

void build_run_slice(const Geometry& g, const GeometryDev& gd, const int2* slc, int ds, unsigned par_mask, int TW,
                     int TH, RunSlice& rs, hipStream_t s) {
    const int N = g.N_total();
    const long long n_ll = (long long)N * ds;
    if (n_ll > 0x7ffffff0ll) throw Error{DCRF_ERR_VALUE, "too many slice pairs"};
    const int n = (int)n_ll;
    rs.P = TW * TH;
    std::vector<int> rb(1, 0);
    for (int i = 0; i < g.n_inst; i++) {
        long long nr;
        if (g.is2d)
            nr = (long long)((g.W[i] + TW - 1) / TW) * ((g.H[i] + TH - 1) / TH);
        else
            nr = ((long long)g.N[i] + rs.P - 1) / rs.P;
        rb.push_back((int)(rb.back() + nr));
    }
    rs.n_runs = rb.back();
    int rbits = 1;
    while ((1ll << rbits) <= (long long)rs.n_runs) rbits++;
    DevBuf<int> run_base, pos;
    run_base.reserve(rb.size());
    DCRF_HIP_CHECK(hipMemcpyAsync(run_base.get(), rb.data(), sizeof(int) * rb.size(), hipMemcpyHostToDevice, s));
    const int* Wd = g.is2d ? gd.W.get() : nullptr;
    const int is2d = g.is2d ? 1 : 0;
    // run-major pixel order
    rs.rpix.reserve(N);
    rs.pix_base.reserve(rs.n_runs + 1);
    pos.reserve(N);
    {
        DevBuf<unsigned long long> pk, spk;
        pk.reserve(N);
        spk.reserve(N);
        k_rs_pixkeys<<<blocks_for(N), kThreads, 0, s>>>(N, gd.pix_base.get(), Wd, run_base.get(), g.n_inst, is2d, TW,
                                                        TH, pk.get());
        DCRF_HIP_CHECK(hipGetLastError());
        size_t tb = 0;
        DCRF_HIP_CHECK(rocprim::radix_sort_keys(nullptr, tb, pk.get(), spk.get(), (size_t)N, 0, 32 + rbits, s));
        DevBuf<char> tmp;
        tmp.reserve(tb);
        DCRF_HIP_CHECK(rocprim::radix_sort_keys(tmp.get(), tb, pk.get(), spk.get(), (size_t)N, 0, 32 + rbits, s));
        k_rs_order<<<blocks_for(N), kThreads, 0, s>>>(spk.get(), N, rs.rpix.get(), pos.get(), rs.pix_base.get());
        DCRF_HIP_CHECK(hipGetLastError());
        DCRF_HIP_CHECK(hipMemcpyAsync(rs.pix_base.get() + rs.n_runs, &N, sizeof(int), hipMemcpyHostToDevice, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
    }
    DevBuf<unsigned long long> key, skey;
    DevBuf<int> iota, perm, flag, uid;
    key.reserve(n);
    skey.reserve(n);
    iota.reserve(n);
    perm.reserve(n);
    flag.reserve(n);
    uid.reserve(n);
    k_rs_keys<<<blocks_for(n), kThreads, 0, s>>>(slc, n, ds, par_mask, gd.pix_base.get(), Wd, run_base.get(),
                                                  g.n_inst, is2d, TW, TH, key.get());
    DCRF_HIP_CHECK(hipGetLastError());
    k_iota<<<blocks_for(n), kThreads, 0, s>>>(iota.get(), n);
    DCRF_HIP_CHECK(hipGetLastError());
    size_t tb = 0;
    DCRF_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, tb, key.get(), skey.get(), iota.get(), perm.get(), (size_t)n, 0,
                                             32 + rbits, s));
    DevBuf<char> tmp;
    tmp.reserve(tb);
    DCRF_HIP_CHECK(rocprim::radix_sort_pairs(tmp.get(), tb, key.get(), skey.get(), iota.get(), perm.get(), (size_t)n, 0,
                                             32 + rbits, s));
    k_run_flags<<<blocks_for(n), kThreads, 0, s>>>(skey.get(), n, flag.get());
    DCRF_HIP_CHECK(hipGetLastError());
    size_t sb = 0;
    DCRF_HIP_CHECK(rocprim::exclusive_scan(nullptr, sb, flag.get(), uid.get(), 0, (size_t)n, rocprim::plus<int>(), s));
    DevBuf<char> stmp;
    stmp.reserve(sb);
    DCRF_HIP_CHECK(rocprim::exclusive_scan(stmp.get(), sb, flag.get(), uid.get(), 0, (size_t)n, rocprim::plus<int>(), s));
    int tail[2];
    DCRF_HIP_CHECK(hipMemcpyAsync(&tail[0], uid.get() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP_CHECK(hipMemcpyAsync(&tail[1], flag.get() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    DCRF_HIP_CHECK(hipStreamSynchronize(s));
    const int total = tail[0] + tail[1];
    rs.rows.reserve(total);
    rs.base.reserve(rs.n_runs + 1);
    k_rs_compact<<<blocks_for(n), kThreads, 0, s>>>(skey.get(), flag.get(), uid.get(), n, rs.rows.get(), rs.base.get());
    DCRF_HIP_CHECK(hipGetLastError());
    DCRF_HIP_CHECK(hipMemcpyAsync(rs.base.get() + rs.n_runs, &total, sizeof(int), hipMemcpyHostToDevice, s));
    rs.loc.reserve(n);
    k_rs_loc<<<blocks_for(n), kThreads, 0, s>>>(skey.get(), perm.get(), uid.get(), flag.get(), rs.base.get(), slc,
                                                  pos.get(), ds, n, rs.loc.get());
    DCRF_HIP_CHECK(hipGetLastError());
    std::vector<int> hb(rs.n_runs + 1);
    DCRF_HIP_CHECK(hipMemcpyAsync(hb.data(), rs.base.get(), sizeof(int) * hb.size(), hipMemcpyDeviceToHost, s));
    DCRF_HIP_CHECK(hipStreamSynchronize(s));
    rs.u_max = 0;
    for (int r = 0; r < rs.n_runs; r++) rs.u_max = std::max(rs.u_max, hb[r + 1] - hb[r]);
}

int default_tile_side(int L) {
    (void)L;
    return 128;
}

void make_tile_map(const Geometry& g, int side, TileMap& tm) {
    tm.side = side;
    tm.base.assign(1, 0);
    for (int i = 0; i < g.n_inst; i++) {
        long long nt;
        if (g.is2d)
            nt = (long long)((g.W[i] + side - 1) / side) * ((g.H[i] + side - 1) / side);
        else
            nt = ((long long)g.N[i] + (long long)side * side - 1) / ((long long)side * side);
        tm.base.push_back((int)(tm.base.back() + nt));
    }
    tm.base_dev.reserve(tm.base.size());
    DCRF_HIP_CHECK(hipMemcpy(tm.base_dev.get(), tm.base.data(), sizeof(int) * tm.base.size(), hipMemcpyHostToDevice));
}

void build_lattice(const Geometry& g, const GeometryDev& gd, const TileMap& tm, const float* feat, int d,
                   TermLattice& lat, float hash_load, hipStream_t s) {
    switch (d) {
#define DCRF_CASE(D) \
    case D: build_lattice_d<D>(g, gd, tm, feat, lat, hash_load, s); break;
        DCRF_CASE(1) DCRF_CASE(2) DCRF_CASE(3) DCRF_CASE(4) DCRF_CASE(5) DCRF_CASE(6) DCRF_CASE(7)
        DCRF_CASE(8) DCRF_CASE(9) DCRF_CASE(10) DCRF_CASE(11) DCRF_CASE(12) DCRF_CASE(13)
        DCRF_CASE(14) DCRF_CASE(15)
#undef DCRF_CASE
        default:
            throw Error{DCRF_ERR_VALUE, "feature dimension must be between 1 and " +
                                            std::to_string(DCRF_MAX_DIM)};
    }
}

void build_chunks(TermLattice& lat, int chunk) {
    const int V = lat.V, n_runs = (int)lat.run_v.size(), T = lat.n_tiles;
    std::vector<Chunk>& ch = lat.chunks;
    ch.clear();
    ch.reserve(n_runs + n_runs / 4 + 16);
    std::vector<int> seg_tile;
    seg_tile.reserve(ch.capacity());
    // vertices without real contributors (created by padding lanes only) still need
    // their row zeroed every iteration: an empty segment each, listed with tile 0
    std::vector<char> seen(V, 0);
    for (int r = 0; r < n_runs; r++) seen[lat.run_v[r]] = 1;
    for (int v = 0; v < V; v++)
        if (!seen[v]) {
            ch.push_back(Chunk{v, 0, 0, v});
            seg_tile.push_back(0);
        }
    std::vector<int> cnt(V, 0);
    const size_t first_real = ch.size();
    for (int r = 0; r < n_runs; r++) {
        const int b = lat.run_beg[r], e = lat.run_beg[r + 1], v = lat.run_v[r];
        for (int i = b; i < e; i += chunk) {
            ch.push_back(Chunk{v, i, std::min(e, i + chunk), 0});
            seg_tile.push_back(lat.run_t[r]);
            cnt[v]++;
        }
    }
    // vertices with one segment write their row directly; the others get partial rows,
    // contiguous per vertex in segment (= tile) order, recombined by the combine pass
    std::vector<int> pfirst(V, -1);
    lat.combines.clear();
    int prow = 0;
    for (int v = 0; v < V; v++)
        if (cnt[v] > 1) {
            pfirst[v] = prow;
            lat.combines.push_back(Combine{v, prow, cnt[v], 0});
            prow += cnt[v];
        }
    std::vector<int> used(V, 0);
    for (size_t c = first_real; c < ch.size(); c++) {
        const int v = ch[c].v;
        ch[c].out = pfirst[v] < 0 ? v : -(pfirst[v] + used[v]++) - 1;
    }
    lat.chunk_tile_ptr.assign(T + 1, (int)ch.size());
    for (int c = (int)ch.size() - 1; c >= 0; c--) lat.chunk_tile_ptr[seg_tile[c]] = c;
    for (int t = T - 1; t >= 0; t--) lat.chunk_tile_ptr[t] = std::min(lat.chunk_tile_ptr[t], lat.chunk_tile_ptr[t + 1]);
    lat.n_partials = prow;
    lat.chunk_len = chunk;
}

void features_from_dxN(const Geometry& g, const GeometryDev& gd, const float* feat_dxN, int d, float* out,
                       hipStream_t s) {
    const int npad = g.npad_total();
    k_feat_dxN<<<blocks_for(npad), kThreads, 0, s>>>(feat_dxN, d, g.N_total(), gd.pix_base.get(),
                                                     gd.pad_base.get(), g.n_inst, npad, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void features_gaussian_2d(const Geometry& g, const GeometryDev& gd, float sx, float sy, float* out,
                          hipStream_t s) {
    const int npad = g.npad_total();
    k_feat_2d<<<blocks_for(npad), kThreads, 0, s>>>(gd.pix_base.get(), gd.pad_base.get(), gd.W.get(), g.n_inst,
                                                    npad, sx, sy, nullptr, 1.f, 1.f, 1.f, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

void features_bilateral_2d(const Geometry& g, const GeometryDev& gd, float sx, float sy, float sr, float sg,
                           float sb, const uint8_t* rgb, float* out, hipStream_t s) {
    const int npad = g.npad_total();
    k_feat_2d<<<blocks_for(npad), kThreads, 0, s>>>(gd.pix_base.get(), gd.pad_base.get(), gd.W.get(), g.n_inst,
                                                    npad, sx, sy, rgb, sr, sg, sb, out);
    DCRF_HIP_CHECK(hipGetLastError());
}

}  // namespace dcrf
