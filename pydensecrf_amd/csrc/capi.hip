// Host orchestration and the extern "C" boundary (include/dcrf_hip.h).
//
// A dcrf_model mirrors the reference's DenseCRF (pydensecrf/densecrf/src/densecrf.cpp):
// it owns the unary, the list of pairwise terms (each a lattice + normaliser + label
// compatibility, pairwise.cpp:116-122 / labelcompatibility.cpp) and runs mean-field
// inference.  Construction touches no device; the first call that needs the GPU
// creates the model's stream on the current HIP device.  No C++ exception crosses the
// boundary: every entry point returns a status and records a message.
#include "dcrf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <numeric>

using namespace dcrf;

namespace dcrf {
namespace {
thread_local std::string g_last_error;
}
void set_error(int code, const std::string& msg) {
    (void)code;
    g_last_error = msg;
}
}  // namespace dcrf

namespace {

constexpr int kDefaultChunk = 64;

struct TermLayout {
    DevBuf<int2> nbr;    // [(d+1)][V] global rows
    DevBuf<int> vtx;     // [N][d+1] global rows
    DevBuf<float> swF;   // slice weights, forward
    DevBuf<float> swT;   // slice weights, transpose (unused when symmetric)
};

struct PendingTiming {
    hipEvent_t start, stop;
    double bytes;
};

}  // namespace

struct dcrf_model {
    Geometry geo;
    GeometryDev geod;
    int L = 0;
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    std::vector<float> unary_host;  // (L, N) row-major as given
    bool has_unary = false, unary_dirty = false;
    DevBuf<float> unary;  // [N][L]
    std::vector<std::unique_ptr<Term>> terms;
    DevBuf<float> Q, work, tmp1, tmp2, stage;
    float hash_load = 0.5f;
    int splat_chunk = 0;  // 0: default

    // global filter layout
    bool layout_valid = false;
    int V_total = 0, rows0 = 0, E_total = 0, C_total = 0, CB_total = 0, max_d = 0;
    std::vector<int> order, vbase;
    std::vector<int> prefixV;
    DevBuf<float> vals[2];
    DevBuf<Chunk> g_chunks;
    DevBuf<Combine> g_comb;
    DevBuf<int> g_pt;
    DevBuf<float> g_wF, g_wT;
    bool all_symmetric = true;
    std::vector<std::unique_ptr<TermLayout>> tl;

    // timing
    bool timing = false;
    long long blur_launches = 0;
    double blur_ms = 0.0, blur_bytes = 0.0;
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> free_events;

    ~dcrf_model() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (auto e : free_events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int N() const { return geo.N_total(); }
    int chunk() const { return splat_chunk > 0 ? splat_chunk : kDefaultChunk; }
};

namespace {

template <typename F>
int guard(F&& f) {
    try {
        f();
        set_error(DCRF_OK, "");
        return DCRF_OK;
    } catch (const Error& e) {
        set_error(e.code, e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error(DCRF_ERR_MEMORY, "host allocation failed");
        return DCRF_ERR_MEMORY;
    } catch (const std::exception& e) {
        set_error(DCRF_ERR_DEVICE, e.what());
        return DCRF_ERR_DEVICE;
    }
}

void require(bool cond, const std::string& msg, int code = DCRF_ERR_VALUE) {
    if (!cond) throw Error{code, msg};
}

hipEvent_t take_event(dcrf_model& m) {
    if (!m.free_events.empty()) {
        hipEvent_t e = m.free_events.back();
        m.free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    DCRF_HIP_CHECK(hipEventCreate(&e));
    return e;
}

void collect_timing(dcrf_model& m) {
    for (auto& p : m.pending) {
        DCRF_HIP_CHECK(hipEventSynchronize(p.stop));
        float ms = 0.f;
        DCRF_HIP_CHECK(hipEventElapsedTime(&ms, p.start, p.stop));
        m.blur_ms += ms;
        m.blur_bytes += p.bytes;
        m.blur_launches++;
        m.free_events.push_back(p.start);
        m.free_events.push_back(p.stop);
    }
    m.pending.clear();
}

void ensure_device(dcrf_model& m) {
    if (m.dev_ready) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error{DCRF_ERR_DEVICE, "no HIP device is available"};
    DCRF_HIP_CHECK(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
    const Geometry& g = m.geo;
    m.geod.pix_base.reserve(g.pix_base.size());
    m.geod.pad_base.reserve(g.pad_base.size());
    m.geod.W.reserve(g.n_inst);
    DCRF_HIP_CHECK(hipMemcpy(m.geod.pix_base.get(), g.pix_base.data(), sizeof(int) * g.pix_base.size(),
                             hipMemcpyHostToDevice));
    DCRF_HIP_CHECK(hipMemcpy(m.geod.pad_base.get(), g.pad_base.data(), sizeof(int) * g.pad_base.size(),
                             hipMemcpyHostToDevice));
    std::vector<int> W = g.is2d ? g.W : std::vector<int>(g.n_inst, 1);
    DCRF_HIP_CHECK(hipMemcpy(m.geod.W.get(), W.data(), sizeof(int) * g.n_inst, hipMemcpyHostToDevice));
    m.dev_ready = true;
}

void ensure_unary(dcrf_model& m) {
    if (!m.has_unary || !m.unary_dirty) return;
    const size_t n = (size_t)m.N() * m.L;
    m.stage.reserve(n);
    m.unary.reserve(n);
    DCRF_HIP_CHECK(hipMemcpyAsync(m.stage.get(), m.unary_host.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                                  m.stream));
    launch_transpose(m.stage.get(), m.unary.get(), m.L, m.N(), m.stream);
    DCRF_HIP_CHECK(hipStreamSynchronize(m.stream));
    m.unary_dirty = false;
}

inline bool pre_norm(int ntype, bool transpose) {
    return ntype == DCRF_NORMALIZE_SYMMETRIC || (ntype == DCRF_NORMALIZE_BEFORE && !transpose) ||
           (ntype == DCRF_NORMALIZE_AFTER && transpose);
}
inline bool post_norm(int ntype, bool transpose) {
    return ntype == DCRF_NORMALIZE_SYMMETRIC || (ntype == DCRF_NORMALIZE_BEFORE && transpose) ||
           (ntype == DCRF_NORMALIZE_AFTER && !transpose);
}
// Alpha of the slice (permutohedral.cpp:571): 1 / (1 + 2^-d) in float.
inline float slice_alpha(int d) { return 1.0f / (1.0f + std::pow(2.0f, (float)-d)); }

// norm = K*1 for one term (pairwise.cpp:44), then the normalisation of :46-61.
void compute_norm(dcrf_model& m, Term& t) {
    TermLattice& lat = t.lat;
    const int N = m.N(), d1 = lat.d + 1, V = lat.V, E = N * d1;
    hipStream_t s = m.stream;
    DevBuf<int> pt;
    DevBuf<float> w, sw, vals0, vals1, ones;
    DevBuf<Chunk> ch;
    DevBuf<Combine> cb;
    DevBuf<int2> nbr;
    pt.reserve(E);
    w.reserve(E);
    sw.reserve(E);
    launch_csr_gather(lat.csr_e.get(), E, d1, lat.bary.get(), nullptr, 0, pt.get(), w.get(), s);
    const int rows = V + 1 + lat.n_partials;
    vals0.reserve(rows);
    vals1.reserve(V + 1);
    launch_fill(vals0.get(), 0.f, rows, s);
    launch_fill(vals1.get(), 0.f, V + 1, s);
    ch.reserve(lat.n_chunks);
    cb.reserve(lat.n_combines);
    launch_remap_chunks(lat.chunks.get(), ch.get(), lat.n_chunks, 0, 0, V + 1, s);
    launch_remap_combines(lat.combines.get(), cb.get(), lat.n_combines, 0, V + 1, s);
    nbr.reserve((size_t)d1 * V);
    launch_remap_nbr(lat.nbr.get(), nbr.get(), d1 * V, 0, V, s);
    ones.reserve(N);
    launch_fill(ones.get(), 1.f, N, s);
    launch_splat(ch.get(), lat.n_chunks, pt.get(), w.get(), ones.get(), vals0.get(), 1, s);
    launch_combine(cb.get(), lat.n_combines, vals0.get(), 1, s);
    BlurTable bt{};
    bt.nterms = 1;
    bt.t[0] = BlurTerm{0, V, lat.d, nbr.get()};
    float* bufs[2] = {vals0.get(), vals1.get()};
    for (int step = 0; step <= lat.d; step++) launch_blur(bufs[step & 1], bufs[(step + 1) & 1], V, 1, bt, step, 0, s);
    DevBuf<int> gv;
    gv.reserve(E);
    launch_slice_weights(lat.vtx.get(), lat.bary.get(), nullptr, 0, slice_alpha(lat.d), N, d1, 0, gv.get(), sw.get(), s);
    SliceTable st{};
    st.nterms = 1;
    st.t[0] = SliceTerm{gv.get(), sw.get(), bufs[d1 & 1], d1, -1, 0.f, nullptr, 1.0f};
    t.norm.reserve(N);
    launch_slice_term(t.norm.get(), N, 1, st, 0, s);
    if (t.ntype == DCRF_NO_NORMALIZATION) {
        // mean_norm = N / sum(norm) per instance, in float (pairwise.cpp:46-53)
        std::vector<float> h(N);
        DCRF_HIP_CHECK(hipMemcpyAsync(h.data(), t.norm.get(), sizeof(float) * N, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        for (int i = 0; i < m.geo.n_inst; i++) {
            const int b = m.geo.pix_base[i], e = m.geo.pix_base[i + 1];
            float mean = 0.f;
            for (int p = b; p < e; p++) mean += h[p];
            mean = (float)(e - b) / mean;
            for (int p = b; p < e; p++) h[p] = mean;
        }
        DCRF_HIP_CHECK(hipMemcpyAsync(t.norm.get(), h.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    } else {
        launch_norm_finish(t.norm.get(), N, t.ntype, s);
    }
    DCRF_HIP_CHECK(hipStreamSynchronize(s));
}

void add_term(dcrf_model& m, int d, int ckind, const float* compat, int ktype, int ntype,
              const std::function<void(float*)>& make_features) {
    require((int)m.terms.size() < DCRF_MAX_TERMS, "at most " + std::to_string(DCRF_MAX_TERMS) + " pairwise terms");
    require(d >= 1 && d <= DCRF_MAX_DIM, "feature dimension must be between 1 and " + std::to_string(DCRF_MAX_DIM));
    require(ckind >= DCRF_COMPAT_POTTS && ckind <= DCRF_COMPAT_MATRIX, "unknown label compatibility kind");
    require(ktype >= DCRF_CONST_KERNEL && ktype <= DCRF_FULL_KERNEL, "unknown kernel type");
    require(ntype >= DCRF_NO_NORMALIZATION && ntype <= DCRF_NORMALIZE_SYMMETRIC, "unknown normalization type");
    require(compat != nullptr, "compatibility parameters missing");
    ensure_device(m);
    auto t = std::make_unique<Term>();
    t->ktype = ktype;
    t->ntype = ntype;
    t->ckind = ckind;
    const int L = m.L;
    if (ckind == DCRF_COMPAT_POTTS) {
        t->compat.assign(compat, compat + 1);
    } else if (ckind == DCRF_COMPAT_DIAGONAL) {
        t->compat.assign(compat, compat + L);
    } else {
        // MatrixCompatibility stores 0.5*(m + m^T) (labelcompatibility.cpp:79)
        t->compat.resize((size_t)L * L);
        for (int a = 0; a < L; a++)
            for (int b = 0; b < L; b++) {
                const float s = compat[a * L + b] + compat[b * L + a];
                t->compat[(size_t)a * L + b] = 0.5f * s;
            }
    }
    if (ckind != DCRF_COMPAT_POTTS) {
        t->d_compat.reserve(t->compat.size());
        DCRF_HIP_CHECK(hipMemcpy(t->d_compat.get(), t->compat.data(), sizeof(float) * t->compat.size(),
                                 hipMemcpyHostToDevice));
    }
    DevBuf<float> feat;
    feat.reserve((size_t)m.geo.npad_total() * d);
    make_features(feat.get());
    build_lattice(m.geo, m.geod, feat.get(), d, t->lat, m.hash_load, m.stream);
    if (t->lat.overflow)
        throw Error{DCRF_ERR_OVERFLOW,
                    "a lattice coordinate exceeds the int16 range of the lattice keys; scale the features down"};
    build_chunks(t->lat, m.chunk(), m.stream);
    compute_norm(m, *t);
    m.terms.push_back(std::move(t));
    m.layout_valid = false;
}

void ensure_layout(dcrf_model& m) {
    if (m.layout_valid) return;
    hipStream_t s = m.stream;
    const int T = (int)m.terms.size();
    const int N = m.N(), L = m.L;
    for (auto& t : m.terms)
        if (t->lat.chunk_len != m.chunk()) build_chunks(t->lat, m.chunk(), s);
    m.order.resize(T);
    std::iota(m.order.begin(), m.order.end(), 0);
    std::stable_sort(m.order.begin(), m.order.end(),
                     [&](int a, int b) { return m.terms[a]->lat.d > m.terms[b]->lat.d; });
    m.vbase.assign(T, 0);
    int V = 0, P = 0, E = 0, C = 0, CB = 0;
    m.max_d = -1;
    m.all_symmetric = true;
    for (int k : m.order) {
        const TermLattice& lat = m.terms[k]->lat;
        m.vbase[k] = V;
        V += lat.V;
        P += lat.n_partials;
        E += N * (lat.d + 1);
        C += lat.n_chunks;
        CB += lat.n_combines;
        m.max_d = std::max(m.max_d, lat.d);
        if (m.terms[k]->ntype != DCRF_NORMALIZE_SYMMETRIC && m.terms[k]->ntype != DCRF_NO_NORMALIZATION)
            m.all_symmetric = false;
    }
    m.V_total = V;
    m.rows0 = V + 1 + P;
    m.E_total = E;
    m.C_total = C;
    m.CB_total = CB;
    m.vals[0].reserve((size_t)m.rows0 * L);
    m.vals[1].reserve((size_t)(V + 1) * L);
    launch_fill(m.vals[0].get(), 0.f, (size_t)m.rows0 * L, s);
    launch_fill(m.vals[1].get(), 0.f, (size_t)(V + 1) * L, s);
    m.g_chunks.reserve(C);
    m.g_comb.reserve(CB);
    m.g_pt.reserve(E);
    m.g_wF.reserve(E);
    if (!m.all_symmetric) m.g_wT.reserve(E);
    m.tl.clear();
    m.tl.resize(T);
    int ebase = 0, cbase = 0, cbbase = 0, pbase = V + 1;
    for (int k : m.order) {
        Term& t = *m.terms[k];
        const TermLattice& lat = t.lat;
        const int d1 = lat.d + 1, Ek = N * d1;
        launch_csr_gather(lat.csr_e.get(), Ek, d1, lat.bary.get(), t.norm.get(), pre_norm(t.ntype, false),
                          m.g_pt.get() + ebase, m.g_wF.get() + ebase, s);
        if (!m.all_symmetric)
            launch_csr_gather(lat.csr_e.get(), Ek, d1, lat.bary.get(), t.norm.get(), pre_norm(t.ntype, true),
                              m.g_pt.get() + ebase, m.g_wT.get() + ebase, s);
        launch_remap_chunks(lat.chunks.get(), m.g_chunks.get() + cbase, lat.n_chunks, m.vbase[k], ebase, pbase, s);
        launch_remap_combines(lat.combines.get(), m.g_comb.get() + cbbase, lat.n_combines, m.vbase[k], pbase, s);
        auto lay = std::make_unique<TermLayout>();
        lay->nbr.reserve((size_t)d1 * lat.V);
        launch_remap_nbr(lat.nbr.get(), lay->nbr.get(), d1 * lat.V, m.vbase[k], V, s);
        lay->vtx.reserve(Ek);
        lay->swF.reserve(Ek);
        const float alpha = slice_alpha(lat.d);
        launch_slice_weights(lat.vtx.get(), lat.bary.get(), t.norm.get(), post_norm(t.ntype, false), alpha, N, d1,
                             m.vbase[k], lay->vtx.get(), lay->swF.get(), s);
        if (!m.all_symmetric) {
            lay->swT.reserve(Ek);
            launch_slice_weights(lat.vtx.get(), lat.bary.get(), t.norm.get(), post_norm(t.ntype, true), alpha, N, d1,
                                 m.vbase[k], lay->vtx.get(), lay->swT.get(), s);
        }
        m.tl[k] = std::move(lay);
        ebase += Ek;
        cbase += lat.n_chunks;
        cbbase += lat.n_combines;
        pbase += lat.n_partials;
    }
    m.prefixV.assign(std::max(m.max_d + 1, 0), 0);
    for (int st = 0; st <= m.max_d; st++)
        for (int k = 0; k < T; k++)
            if (m.terms[k]->lat.d >= st) m.prefixV[st] += m.terms[k]->lat.V;
    DCRF_HIP_CHECK(hipStreamSynchronize(s));
    m.layout_valid = true;
}

BlurTable blur_table(const dcrf_model& m) {
    BlurTable bt{};
    bt.nterms = (int)m.order.size();
    for (int i = 0; i < bt.nterms; i++) {
        const int k = m.order[i];
        bt.t[i] = BlurTerm{m.vbase[k], m.terms[k]->lat.V, m.terms[k]->lat.d, m.tl[k]->nbr.get()};
    }
    return bt;
}

SliceTable slice_table(const dcrf_model& m, bool transpose) {
    SliceTable st{};
    st.nterms = (int)m.terms.size();
    for (int k = 0; k < st.nterms; k++) {
        const Term& t = *m.terms[k];
        const TermLayout& lay = *m.tl[k];
        const float* sw = (transpose && !m.all_symmetric) ? lay.swT.get() : lay.swF.get();
        st.t[k] = SliceTerm{lay.vtx.get(), sw, m.vals[(t.lat.d + 1) & 1].get(), t.lat.d + 1, t.ckind,
                            t.compat.empty() ? 0.f : t.compat[0],
                            t.ckind == DCRF_COMPAT_POTTS ? nullptr : t.d_compat.get(), 1.0f};
    }
    return st;
}

// Splat, combine and blur every term's lattice: afterwards term k's filtered rows are
// in vals[(d_k+1) & 1].  `transpose` selects the transposed filter (reverse blur
// order and the transpose's normaliser placement, pairwise.cpp:63-80).
void run_filter(dcrf_model& m, const float* in, bool transpose, hipStream_t s) {
    const int L = m.L;
    const float* w = (transpose && !m.all_symmetric) ? m.g_wT.get() : m.g_wF.get();
    launch_splat(m.g_chunks.get(), m.C_total, m.g_pt.get(), w, in, m.vals[0].get(), L, s);
    launch_combine(m.g_comb.get(), m.CB_total, m.vals[0].get(), L, s);
    const BlurTable bt = blur_table(m);
    for (int step = 0; step <= m.max_d; step++) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (m.timing) {
            e0 = take_event(m);
            e1 = take_event(m);
            DCRF_HIP_CHECK(hipEventRecord(e0, s));
        }
        launch_blur(m.vals[step & 1].get(), m.vals[(step + 1) & 1].get(), m.prefixV[step], L, bt, step,
                    transpose ? 1 : 0, s);
        if (m.timing) {
            DCRF_HIP_CHECK(hipEventRecord(e1, s));
            m.pending.push_back(PendingTiming{e0, e1, (double)m.prefixV[step] * (2.0 * L * 4 + 8)});
        }
    }
}

void prepare(dcrf_model& m) {
    ensure_device(m);
    ensure_unary(m);
    ensure_layout(m);
    const size_t n = (size_t)m.N() * m.L;
    m.Q.reserve(n);
}

void run_inference(dcrf_model& m, int n_iter, const float* U, float* Q, hipStream_t s) {
    launch_softmax_neg(U, Q, m.N(), m.L, s);
    const SliceTable st = slice_table(m, false);
    for (int it = 0; it < n_iter; it++) {
        run_filter(m, Q, false, s);
        launch_meanfield(U, Q, m.N(), m.L, st, nullptr, nullptr, s);
    }
}

dcrf_model* new_model(const std::vector<int>& N, const std::vector<int>* W, const std::vector<int>* H, int L) {
    require(!N.empty(), "a model needs at least one instance");
    require(L >= 1 && L <= DCRF_MAX_LABELS, "number of labels must be between 1 and " +
                                                std::to_string(DCRF_MAX_LABELS));
    auto m = std::make_unique<dcrf_model>();
    m->L = L;
    Geometry& g = m->geo;
    g.n_inst = (int)N.size();
    g.N = N;
    g.pix_base.assign(1, 0);
    g.pad_base.assign(1, 0);
    long long pix = 0, pad = 0;
    for (int n : N) {
        require(n >= 1, "number of variables must be positive");
        pix += n;
        pad += (n + 3) / 4 * 4;
        require(pix * L < (1ll << 31) && pad * (DCRF_MAX_DIM + 1) < (1ll << 31),
                "model too large for 32-bit indexing");
        g.pix_base.push_back((int)pix);
        g.pad_base.push_back((int)pad);
    }
    if (W) {
        g.is2d = true;
        g.W = *W;
        g.H = *H;
    }
    return m.release();
}

}  // namespace

// =====================================================================================
extern "C" {

const char* dcrf_last_error(void) { return g_last_error.c_str(); }

int dcrf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int dcrf_create(int N, int L, dcrf_model** out) {
    return guard([&] {
        require(out != nullptr, "null output pointer");
        *out = new_model({N}, nullptr, nullptr, L);
    });
}

int dcrf_create_2d(int W, int H, int L, dcrf_model** out) {
    return guard([&] {
        require(out != nullptr, "null output pointer");
        require(W >= 1 && H >= 1, "width and height must be positive");
        require((long long)W * H < (1ll << 31), "image too large");
        std::vector<int> w{W}, h{H};
        *out = new_model({W * H}, &w, &h, L);
    });
}

int dcrf_create_batch(int n, const int* N, int L, dcrf_model** out) {
    return guard([&] {
        require(out != nullptr && N != nullptr && n >= 1, "bad batch description");
        *out = new_model(std::vector<int>(N, N + n), nullptr, nullptr, L);
    });
}

int dcrf_create_batch_2d(int n, const int* W, const int* H, int L, dcrf_model** out) {
    return guard([&] {
        require(out != nullptr && W != nullptr && H != nullptr && n >= 1, "bad batch description");
        std::vector<int> w(W, W + n), h(H, H + n), N(n);
        for (int i = 0; i < n; i++) {
            require(w[i] >= 1 && h[i] >= 1 && (long long)w[i] * h[i] < (1ll << 31), "bad image size");
            N[i] = w[i] * h[i];
        }
        *out = new_model(N, &w, &h, L);
    });
}

void dcrf_destroy(dcrf_model* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    delete m;
}

int dcrf_set_unary(dcrf_model* m, const float* U) {
    return guard([&] {
        require(m && U, "null argument");
        const size_t n = (size_t)m->N() * m->L;
        m->unary_host.assign(U, U + n);
        m->has_unary = true;
        m->unary_dirty = true;
    });
}

int dcrf_set_unary_instance(dcrf_model* m, int i, const float* U) {
    return guard([&] {
        require(m && U, "null argument");
        require(i >= 0 && i < m->geo.n_inst, "instance index out of range");
        const int N = m->N(), L = m->L, b = m->geo.pix_base[i], Ni = m->geo.N[i];
        if (m->unary_host.size() != (size_t)N * L) m->unary_host.assign((size_t)N * L, 0.f);
        for (int l = 0; l < L; l++)
            std::memcpy(m->unary_host.data() + (size_t)l * N + b, U + (size_t)l * Ni, sizeof(float) * Ni);
        m->has_unary = true;
        m->unary_dirty = true;
    });
}

int dcrf_add_pairwise(dcrf_model* m, const float* feat, int d, int ckind, const float* compat, int ktype,
                      int ntype) {
    return guard([&] {
        require(m && feat, "null argument");
        require(d >= 1 && d <= DCRF_MAX_DIM, "feature dimension must be between 1 and " + std::to_string(DCRF_MAX_DIM));
        const size_t n = (size_t)d * m->N();
        add_term(*m, d, ckind, compat, ktype, ntype, [&](float* out) {
            DevBuf<float> raw;
            raw.reserve(n);
            DCRF_HIP_CHECK(hipMemcpy(raw.get(), feat, sizeof(float) * n, hipMemcpyHostToDevice));
            features_from_dxN(m->geo, m->geod, raw.get(), d, out, m->stream);
            DCRF_HIP_CHECK(hipStreamSynchronize(m->stream));
        });
    });
}

int dcrf_add_gaussian_2d(dcrf_model* m, float sx, float sy, int ckind, const float* compat, int ktype, int ntype) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(m->geo.is2d, "addPairwiseGaussian needs a 2-D model", DCRF_ERR_STATE);
        add_term(*m, 2, ckind, compat, ktype, ntype,
                 [&](float* out) { features_gaussian_2d(m->geo, m->geod, sx, sy, out, m->stream); });
    });
}

int dcrf_add_bilateral_2d(dcrf_model* m, float sx, float sy, float sr, float sg, float sb, const uint8_t* rgb,
                          int ckind, const float* compat, int ktype, int ntype) {
    return guard([&] {
        require(m && rgb, "null argument");
        require(m->geo.is2d, "addPairwiseBilateral needs a 2-D model", DCRF_ERR_STATE);
        const size_t n = (size_t)m->N() * 3;
        add_term(*m, 5, ckind, compat, ktype, ntype, [&](float* out) {
            DevBuf<uint8_t> img;
            img.reserve(n);
            DCRF_HIP_CHECK(hipMemcpy(img.get(), rgb, n, hipMemcpyHostToDevice));
            features_bilateral_2d(m->geo, m->geod, sx, sy, sr, sg, sb, img.get(), out, m->stream);
            DCRF_HIP_CHECK(hipStreamSynchronize(m->stream));
        });
    });
}

int dcrf_inference(dcrf_model* m, int n_iter, float* Q) {
    return guard([&] {
        require(m && Q, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        prepare(*m);
        const float* U = m->has_unary ? m->unary.get() : nullptr;
        run_inference(*m, n_iter, U, m->Q.get(), m->stream);
        DCRF_HIP_CHECK(hipMemcpyAsync(Q, m->Q.get(), sizeof(float) * (size_t)m->N() * m->L, hipMemcpyDeviceToHost,
                                      m->stream));
        DCRF_HIP_CHECK(hipStreamSynchronize(m->stream));
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_inference_device(dcrf_model* m, int n_iter, const float* U_dev, float* Q_dev, void* stream) {
    return guard([&] {
        require(m && Q_dev, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        prepare(*m);
        hipStream_t s = (hipStream_t)stream;  // 0 is the null stream, as in HIP
        const float* U = U_dev ? U_dev : (m->has_unary ? m->unary.get() : nullptr);
        run_inference(*m, n_iter, U, Q_dev, s);
    });
}

int dcrf_start_inference(dcrf_model* m, float* Q) {
    return guard([&] {
        require(m && Q, "null argument");
        const size_t n = (size_t)m->N() * m->L;
        if (!m->has_unary) {  // densecrf.cpp:180-184: zeros without a unary
            std::memset(Q, 0, sizeof(float) * n);
            return;
        }
        ensure_device(*m);
        ensure_unary(*m);
        m->Q.reserve(n);
        launch_softmax_neg(m->unary.get(), m->Q.get(), m->N(), m->L, m->stream);
        DCRF_HIP_CHECK(hipMemcpyAsync(Q, m->Q.get(), sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
        DCRF_HIP_CHECK(hipStreamSynchronize(m->stream));
    });
}

int dcrf_step_inference(dcrf_model* m, float* Q, float* tmp1, float* tmp2) {
    return guard([&] {
        require(m && Q, "null argument");
        prepare(*m);
        const size_t n = (size_t)m->N() * m->L;
        m->tmp1.reserve(n);
        m->tmp2.reserve(n);
        hipStream_t s = m->stream;
        DCRF_HIP_CHECK(hipMemcpyAsync(m->Q.get(), Q, sizeof(float) * n, hipMemcpyHostToDevice, s));
        if (m->terms.empty()) launch_fill(m->tmp2.get(), 0.f, n, s);
        run_filter(*m, m->Q.get(), false, s);
        launch_meanfield(m->has_unary ? m->unary.get() : nullptr, m->Q.get(), m->N(), m->L, slice_table(*m, false),
                         m->tmp1.get(), m->terms.empty() ? nullptr : m->tmp2.get(), s);
        DCRF_HIP_CHECK(hipMemcpyAsync(Q, m->Q.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        if (tmp1) DCRF_HIP_CHECK(hipMemcpyAsync(tmp1, m->tmp1.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        if (tmp2) DCRF_HIP_CHECK(hipMemcpyAsync(tmp2, m->tmp2.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_kl_divergence(dcrf_model* m, const float* Q, double* kl) {
    return guard([&] {
        require(m && Q && kl, "null argument");
        prepare(*m);
        const size_t n = (size_t)m->N() * m->L;
        m->work.reserve(n);
        hipStream_t s = m->stream;
        DCRF_HIP_CHECK(hipMemcpyAsync(m->work.get(), Q, sizeof(float) * n, hipMemcpyHostToDevice, s));
        run_filter(*m, m->work.get(), false, s);
        const LaunchShape sh = shape_for(m->L);
        const int nb = (m->N() + sh.per - 1) / sh.per;
        DevBuf<double> part;
        part.reserve(nb);
        launch_kl(m->work.get(), m->has_unary ? m->unary.get() : nullptr, m->N(), m->L, slice_table(*m, false),
                  part.get(), s);
        std::vector<double> h(nb);
        DCRF_HIP_CHECK(hipMemcpyAsync(h.data(), part.get(), sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        double acc = 0.0;
        for (double v : h) acc += v;
        *kl = acc;
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_apply_pairwise(dcrf_model* m, int term, const float* Q, float* out, int transpose) {
    return guard([&] {
        require(m && Q && out, "null argument");
        require(term >= 0 && term < (int)m->terms.size(), "pairwise term index out of range");
        prepare(*m);
        const size_t n = (size_t)m->N() * m->L;
        m->work.reserve(n);
        m->tmp1.reserve(n);
        hipStream_t s = m->stream;
        DCRF_HIP_CHECK(hipMemcpyAsync(m->work.get(), Q, sizeof(float) * n, hipMemcpyHostToDevice, s));
        run_filter(*m, m->work.get(), transpose != 0, s);
        launch_slice_term(m->tmp1.get(), m->N(), m->L, slice_table(*m, transpose != 0), term, s);
        DCRF_HIP_CHECK(hipMemcpyAsync(out, m->tmp1.get(), sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_num_variables(const dcrf_model* m) { return m ? m->N() : 0; }
int dcrf_num_labels(const dcrf_model* m) { return m ? m->L : 0; }
int dcrf_num_terms(const dcrf_model* m) { return m ? (int)m->terms.size() : 0; }
int dcrf_num_instances(const dcrf_model* m) { return m ? m->geo.n_inst : 0; }

int dcrf_term_info(const dcrf_model* m, int term, int* d, int* V) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(term >= 0 && term < (int)m->terms.size(), "pairwise term index out of range");
        if (d) *d = m->terms[term]->lat.d;
        if (V) *V = m->terms[term]->lat.V;
    });
}

int dcrf_term_instance_vertices(const dcrf_model* m, int term, int inst, int* V) {
    return guard([&] {
        require(m && V, "null argument");
        require(term >= 0 && term < (int)m->terms.size(), "pairwise term index out of range");
        require(inst >= 0 && inst < m->geo.n_inst, "instance index out of range");
        const auto& vb = m->terms[term]->lat.inst_vbase;
        *V = vb[inst + 1] - vb[inst];
    });
}

int dcrf_get_lattice(const dcrf_model* m, int term, int inst, int* offset, int* rank, float* bary, int* nbr,
                     int16_t* keys, float* norm) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(term >= 0 && term < (int)m->terms.size(), "pairwise term index out of range");
        require(inst >= 0 && inst < m->geo.n_inst, "instance index out of range");
        const Term& t = *m->terms[term];
        const TermLattice& lat = t.lat;
        const int d1 = lat.d + 1;
        const int pb = m->geo.pix_base[inst], Ni = m->geo.N[inst];
        const int vb = lat.inst_vbase[inst], Vi = lat.inst_vbase[inst + 1] - vb;
        const size_t ne = (size_t)Ni * d1;
        if (offset) {
            DCRF_HIP_CHECK(hipMemcpy(offset, lat.vtx.get() + (size_t)pb * d1, sizeof(int) * ne, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < ne; i++) offset[i] -= vb;
        }
        if (rank) DCRF_HIP_CHECK(hipMemcpy(rank, lat.rank.get() + (size_t)pb * d1, sizeof(int) * ne, hipMemcpyDeviceToHost));
        if (bary) DCRF_HIP_CHECK(hipMemcpy(bary, lat.bary.get() + (size_t)pb * d1, sizeof(float) * ne, hipMemcpyDeviceToHost));
        if (nbr) {
            std::vector<int> all((size_t)2 * d1 * lat.V);
            DCRF_HIP_CHECK(hipMemcpy(all.data(), lat.nbr.get(), sizeof(int) * all.size(), hipMemcpyDeviceToHost));
            for (int j = 0; j < d1; j++)
                for (int v = 0; v < Vi; v++)
                    for (int c = 0; c < 2; c++) {
                        const int x = all[2 * ((size_t)j * lat.V + vb + v) + c];
                        nbr[2 * ((size_t)j * Vi + v) + c] = x < 0 ? -1 : x - vb;
                    }
        }
        if (keys)
            DCRF_HIP_CHECK(hipMemcpy(keys, lat.keys.get() + (size_t)vb * lat.d, sizeof(int16_t) * (size_t)Vi * lat.d,
                                     hipMemcpyDeviceToHost));
        if (norm) DCRF_HIP_CHECK(hipMemcpy(norm, t.norm.get() + pb, sizeof(float) * Ni, hipMemcpyDeviceToHost));
    });
}

int dcrf_set_splat_chunk(dcrf_model* m, int chunk) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(chunk >= 0 && chunk <= (1 << 20), "chunk length out of range");
        if (chunk != m->splat_chunk) {
            m->splat_chunk = chunk;
            m->layout_valid = false;
        }
    });
}

int dcrf_enable_timing(dcrf_model* m, int on) {
    return guard([&] {
        require(m != nullptr, "null model");
        if (m->timing && !on) collect_timing(*m);
        m->timing = on != 0;
        m->blur_launches = 0;
        m->blur_ms = 0.0;
        m->blur_bytes = 0.0;
    });
}

int dcrf_blur_stats(dcrf_model* m, long long* launches, double* ms, double* bytes) {
    return guard([&] {
        require(m != nullptr, "null model");
        collect_timing(*m);
        if (launches) *launches = m->blur_launches;
        if (ms) *ms = m->blur_ms;
        if (bytes) *bytes = m->blur_bytes;
    });
}

int dcrf_iteration_bytes(const dcrf_model* m, double* bytes) {
    return guard([&] {
        require(m && bytes, "null argument");
        const double N = m->N(), L = m->L;
        double b = 3.0 * N * L * 4.0;
        for (const auto& t : m->terms) {
            const double d1 = t->lat.d + 1, V = t->lat.V;
            b += 2.0 * N * d1 * 8.0 + 2.0 * N * 4.0 + V * L * 4.0 * (2.0 + 2.0 * d1) + d1 * V * 8.0;
        }
        *bytes = b;
    });
}

}  // extern "C"
