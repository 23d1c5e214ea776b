// Host orchestration and the extern "C" boundary (include/dcrf_hip.h).
//
// A dcrf_model mirrors the reference's DenseCRF (pydensecrf/densecrf/src/densecrf.cpp):
// it owns the unary, the list of pairwise terms (each a lattice + normaliser + label
// compatibility, pairwise.cpp:116-122 / labelcompatibility.cpp) and runs mean-field
// inference.  Construction touches no device; the first call that needs the GPU
// creates the model's stream on the current HIP device.  No C++ exception crosses the
// boundary: every entry point returns a status and records a message.
#include "dcrf_internal.h"

#include <sys/mman.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <thread>

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

constexpr int kDefaultChunk = 32;
constexpr int kCombineSplit = 16;  // longest serial chain of partial rows in one combine

struct TermLayout {
    DevBuf<int2> nbr;    // [(d+1)][V] global rows
    // multi-step blur passes: steps per pass, table offsets (ints) per pass, tables
    // forward / reverse
    std::vector<int> pass_steps, pass_off;
    DevBuf<int> passF, passT;
    DevBuf<int> vtx;     // [N][d+1] global rows
    DevBuf<float> swF;   // slice weights, forward
    DevBuf<float> swT;   // slice weights, transpose (unused when symmetric)
};

// Kernel classes timed by dcrf_enable_timing.
enum { KC_SPLAT = 0, KC_COMBINE = 1, KC_BLUR = 2, KC_MEANFIELD = 3, KC_COUNT = 4 };

struct PendingTiming {
    hipEvent_t start, stop;
    int cls;
    double bytes;
};

}  // namespace

namespace dcrf {

// ---------------------------------------------------------------- host transfers
// Large copies between pageable host memory (the caller's arrays) and the device go
// through one page-locked staging area per process (hipHostMalloc, grown on demand,
// shared by all models under a lock, kept until dcrf_release_cached_memory) in chunks:
// the DMA engine moves some chunks while host threads copy others in or out, where the
// runtime would stage pageable memory through one host thread.
constexpr size_t kStageMax = (size_t)256 << 20;  // larger copies: windows of this size
constexpr int kCopyThreads = 8;

struct HostStage {
    std::mutex mu;
    char* p = nullptr;
    size_t cap = 0;
};
HostStage& host_stage() {
    static HostStage* h = new HostStage;  // never destroyed: no runtime call at process exit
    return *h;
}

// Threads for a host copy of n bytes: about one per 2 MiB, at most kCopyThreads and the
// hardware threads.
int copy_threads(size_t n) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max<size_t>(1, std::min({(size_t)kCopyThreads, hw, n >> 21}));
}

// Chunk t of n bytes split T ways, 4 KiB aligned: [off(t), off(t + 1)).
inline size_t chunk_off(size_t n, int T, int t) { return t >= T ? n : (n / T * t) & ~(size_t)4095; }

// Chunks of one staged window of n bytes moved by T threads: about 2 MiB each, at least
// one and at most 8 per thread (thread t takes chunks t, t + T, ...).  The host copy
// left after the window's last DMA (downloads), or made before its first (uploads), is
// then one small chunk rather than a thread's whole share.
inline int window_chunks(size_t n, int T) { return (int)std::clamp<size_t>(n >> 21, (size_t)T, (size_t)8 * T); }

// work(t) for t < T: t = 0 on this thread, the others on threads of their own where the
// system starts them (else here as well).  work must not throw.
template <typename F>
void run_parallel(int T, const F& work) {
    std::vector<std::thread> th;
    int t = 1;
    try {
        for (; t < T; t++) th.emplace_back(work, t);
    } catch (...) {
    }
    work(0);
    for (int u = t; u < T; u++) work(u);
    for (auto& x : th) x.join();
}

// dst = src (host memory, n bytes) with copy_threads(n) threads.
void host_copy(void* dst, const void* src, size_t n) {
    const int T = copy_threads(n);
    run_parallel(T, [&](int t) {
        const size_t a = chunk_off(n, T, t), b = chunk_off(n, T, t + 1);
        std::memcpy((char*)dst + a, (const char*)src + a, b - a);
    });
}

// rows rows of cols floats from src (row stride ss) to dst (row stride ds), rows shared
// out among copy_threads threads (row t, t + T, ...).
void copy_rows(float* dst, size_t ds, const float* src, size_t ss, int rows, size_t cols) {
    const int T = std::min(copy_threads((size_t)rows * cols * sizeof(float)), std::max(rows, 1));
    run_parallel(T, [&](int t) {
        for (int r = t; r < rows; r += T) std::memcpy(dst + r * ds, src + r * ss, sizeof(float) * cols);
    });
}

// The staging area of at least n bytes, or null where the runtime cannot give that much
// page-locked memory (the caller then copies without staging).
char* stage_reserve(HostStage& H, size_t n) {
    if (n > H.cap) {
        if (H.p) DCRF_HIP_CHECK(hipHostFree(H.p));
        H.p = nullptr;
        H.cap = 0;
        if (hipHostMalloc((void**)&H.p, n, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();  // clear the failure: the copy goes unstaged
            H.p = nullptr;
            return nullptr;
        }
        H.cap = n;
    }
    return H.p;
}

// Events of one staged copy, destroyed on every exit.
struct ChunkEvents {
    std::vector<hipEvent_t> ev;
    explicit ChunkEvents(int T) : ev(T, nullptr) {
        for (auto& e : ev) DCRF_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    ~ChunkEvents() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void copy_to_host_window(char* stage, void* dst, const void* src, size_t n, hipStream_t s);
void copy_to_device_window(char* stage, void* dst, const void* src, size_t n, hipStream_t s);

// dst (host) = src (device), n bytes, after the work queued on s; returns when done.
// Copies above kStageMax go through the staging area one window at a time.
void copy_to_host(void* dst, const void* src, size_t n, hipStream_t s) {
    if (n < kStagedMin) {
        DCRF_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    HostStage& H = host_stage();
    std::lock_guard<std::mutex> lk(H.mu);
    char* stage = stage_reserve(H, std::min(n, kStageMax));
    if (!stage) {
        DCRF_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    for (size_t o = 0; o < n; o += kStageMax)
        copy_to_host_window(stage, (char*)dst + o, (const char*)src + o, std::min(kStageMax, n - o), s);
}

// dst (device) = src (host), n bytes, queued on s; returns when the copy is done.
void copy_to_device(void* dst, const void* src, size_t n, hipStream_t s) {
    if (n < kStagedMin) {
        DCRF_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    HostStage& H = host_stage();
    std::lock_guard<std::mutex> lk(H.mu);
    char* stage = stage_reserve(H, std::min(n, kStageMax));
    if (!stage) {
        DCRF_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s));
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    for (size_t o = 0; o < n; o += kStageMax)
        copy_to_device_window(stage, (char*)dst + o, (const char*)src + o, std::min(kStageMax, n - o), s);
}

// One window (n <= the staging area's size) of copy_to_host: every chunk's DMA is
// queued at once, in order, and each chunk is copied out as soon as its event fires.
void copy_to_host_window(char* stage, void* dst, const void* src, size_t n, hipStream_t s) {
    const int T = copy_threads(n), C = window_chunks(n, T);
    ChunkEvents E(C);
    for (int c = 0; c < C; c++) {  // all chunks queued before any thread starts
        const size_t a = chunk_off(n, C, c), b = chunk_off(n, C, c + 1);
        DCRF_HIP_CHECK(hipMemcpyAsync(stage + a, (const char*)src + a, b - a, hipMemcpyDeviceToHost, s));
        DCRF_HIP_CHECK(hipEventRecord(E.ev[c], s));
    }
    std::vector<hipError_t> err(C, hipSuccess);
    run_parallel(T, [&](int t) {  // each chunk leaves the staging area as soon as it arrived
        for (int c = t; c < C; c += T) {
            err[c] = hipEventSynchronize(E.ev[c]);
            if (err[c] != hipSuccess) return;
            const size_t a = chunk_off(n, C, c), b = chunk_off(n, C, c + 1);
            std::memcpy((char*)dst + a, stage + a, b - a);
        }
    });
    for (hipError_t e : err) DCRF_HIP_CHECK(e);
}

// One window of copy_to_device: each thread stages its chunks in turn and queues each
// chunk's DMA as soon as it is staged; returns when the window's copies are done (the
// staging area is free again).
void copy_to_device_window(char* stage, void* dst, const void* src, size_t n, hipStream_t s) {
    const int T = copy_threads(n), C = window_chunks(n, T);
    std::vector<hipError_t> err(C, hipSuccess);
    run_parallel(T, [&](int t) {
        for (int c = t; c < C; c += T) {
            const size_t a = chunk_off(n, C, c), b = chunk_off(n, C, c + 1);
            std::memcpy(stage + a, (const char*)src + a, b - a);
            err[c] = hipMemcpyAsync((char*)dst + a, stage + a, b - a, hipMemcpyHostToDevice, s);
            if (err[c] != hipSuccess) return;
        }
    });
    const hipError_t sync = hipStreamSynchronize(s);  // the staging area is free again
    for (hipError_t e : err) DCRF_HIP_CHECK(e);
    DCRF_HIP_CHECK(sync);
}

// Frees the staging area; returns its bytes.
size_t release_host_stage() {
    HostStage& H = host_stage();
    std::lock_guard<std::mutex> lk(H.mu);
    const size_t b = H.cap;
    if (H.p) DCRF_HIP_CHECK(hipHostFree(H.p));
    H.p = nullptr;
    H.cap = 0;
    return b;
}

}  // namespace dcrf

// Host floats left uninitialised when allocated (a copy fills them).  Blocks of 4 MiB
// and more are 2 MiB aligned and advised for transparent huge pages: a 66 MB unary then
// faults in ~33 pages on its first copy instead of ~16,000.
struct HostFloats {
    struct Free {
        void operator()(float* q) const { std::free(q); }
    };
    std::unique_ptr<float, Free> p;
    size_t n = 0;
    void resize(size_t count) {
        if (count == n) return;
        p.reset();
        n = 0;
        if (!count) return;
        const size_t bytes = sizeof(float) * count;
        void* q = nullptr;
        if (bytes >= ((size_t)4 << 20)) {
            if (posix_memalign(&q, (size_t)2 << 20, bytes) != 0) q = nullptr;
            if (q) (void)madvise(q, bytes, MADV_HUGEPAGE);  // advice only
        } else {
            q = std::malloc(bytes);
        }
        if (!q) throw std::bad_alloc();
        p.reset(static_cast<float*>(q));
        n = count;
    }
    float* data() { return p.get(); }
    const float* data() const { return p.get(); }
    size_t size() const { return n; }
};

struct dcrf_model {
    Geometry geo;
    GeometryDev geod;
    int L = 0;
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    HostFloats unary_host;  // (L, N) row-major as given
    std::vector<char> unary_unset;  // per instance: its columns of unary_host hold no data yet
    bool has_unary = false, unary_dirty = false;
    DevBuf<float> unary;  // [N][L]
    // LogisticUnaryEnergy (U = L f): parameters L [L][F] row-major, features [N][F]
    bool logistic = false;
    int lF = 0;
    std::vector<float> lL;
    DevBuf<float> lL_dev, lf, lfFN;  // features [N][F] and as given [F][N]
    std::vector<float> ugrad_param;  // its parameter gradient from the last dcrf_gradient
    std::vector<std::unique_ptr<Term>> terms;
    DevBuf<float> Q, work, tmp1, tmp2, stage;
    DevBuf<float> Qp;  // inference iterate with the vertex-row stride (see run_inference)
    int ldv = 0;       // row stride (floats) of the vertex rows: L rounded to a cache line
    float hash_load = 0.5f;
    int splat_chunk = 0;  // 0: default
    int tile_side = 0;  // 0: default_tile_side(L)
    TileMap tiles;
    bool tiles_ready = false;

    // global filter layout
    bool layout_valid = false;
    int V_total = 0, rows0 = 0, E_total = 0, C_total = 0, CB_total = 0, CB2_total = 0, max_d = 0;
    std::vector<int> order, vbase;
    std::vector<int> prefixV;
    DevBuf<float> vals[2];
    DevBuf<Chunk> g_chunks;
    DevBuf<Combine> g_comb, g_comb2;  // combine pieces, then the second level of split vertices
    // host scratch of the layout's in-wave runs, kept between layout builds (a kernel-
    // parameter change rebuilds the layout; fresh vectors cost their page faults each time)
    std::vector<Chunk> gch_scratch;
    std::vector<int> runs_scratch, first_scratch;
    bool comb_waves = false;          // g_comb is a one-launch list (launch_combine_waves)
    DevBuf<int> g_tile_seg, g_tile_e;  // first segment / entry of each tile
    bool tiled = false;                // splat through LDS tiles (else the global form)
    size_t lds_cap = 0;                // LDS bytes one workgroup may allocate
    TiledSplat tsplat{};
    DevBuf<int2> g_entF, g_entT;  // splat entries (pixel, weight bits), forward / transposed
    // opt-in: Q between inference iterations in 8 x 8 pixel blocks (2-D single images),
    // read by the splat through entries holding block positions
    bool qblocks = false, qblocks_ready = false;
    // opt-in: blur passes of fewer than blur_small_v vertices with blur_r labels per lane
    int blur_r = 0, blur_small_v = 0;
    DevBuf<int2> g_entB;
    DevBuf<float> Qb;
    bool fused = false;           // inference through k_meanfield_fused
    DevBuf<int2> g_slc;           // its per-pixel slice pairs of every term
    DevBuf<int2> g_slcT;          // the same with the transposed filter's slice weights (adjoint;
    bool slcT_valid = false;      // built on first use when they differ from the forward's)
    bool runs = false;            // ... or through k_meanfield_runs (rows staged per pixel run)
    int mf_mode = 0;              // dcrf_set_meanfield_kernel: 0 auto (fused), 1 generic, 2 fused, 3 runs
    bool pairs = true;            // blur up to three steps per pass (dcrf_set_blur_pairs)
    int blur_steps = 2;           // most steps per multi-step pass, 2 or 3 (dcrf_set_blur_steps)
    int n_passes = 0;             // launches per filter with multi-step passes
    std::vector<int> prefixP;     // vertices of the terms with more than p passes
    int run_rows = 16;            // pixel rows of a run rectangle
    RunSlice rslice;
    int rows_lds = 0;
    bool all_symmetric = true;
    std::vector<std::unique_ptr<TermLayout>> tl;

    // term weights (device, [T]); null unless set
    DevBuf<float> kw;
    bool kw_set = false;
    // training checkpoints: Q_0..Q_n and the raw messages K~_k Q_t
    int train_n = -1;
    const float* train_kw = nullptr;
    DevBuf<float> qcache, mcache, bbuf, partial;
    DevBuf<double> rmid;  // first-stage sums of launch_sum_rows
    DevBuf<double> gacc, gacc2;
    // objective on the device: ground truth, Hamming's class weights, partial sums
    DevBuf<int16_t> gt_dev;
    DevBuf<float> obj_w;
    DevBuf<double> obj_partial;
    // kernel-parameter gradient: the filter plans (built on first use for the lattices of
    // generation lat_gen, which every lattice change advances), 4 [N][L] work slabs per
    // term of a plan (lbl, [a | c], x), 2 [N][d] feature gradients (one-thread-per-point
    // form), per-block partials and the accumulated gradient of all terms
    std::vector<std::unique_ptr<KgPlan>> kgplans;
    int lat_gen = 0;
    DevBuf<float> kgbuf, kgdf;
    DevBuf<int16_t> labels;  // dcrf_map / dcrf_current_map
    DevBuf<double> kpartial, kacc;

    // captured inference (dcrf_capture_inference)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graph_valid = false;

    // timing
    bool timing = false;
    long long kc_launches[KC_COUNT] = {};
    double kc_ms[KC_COUNT] = {}, kc_bytes[KC_COUNT] = {};
    std::vector<PendingTiming> pending;
    std::vector<hipEvent_t> free_events;

    ~dcrf_model() {
        for (auto& p : pending) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (auto e : free_events) (void)hipEventDestroy(e);
        drop_graph();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // what a captured graph reads: the model's own unary buffer (nullptr: the caller's
    // unary, or none) and the term-weight buffer (nullptr: no weights)
    bool graph_own_unary = false;
    const float* graph_unary = nullptr;
    const float* graph_kw = nullptr;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_valid = false;
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
        m.kc_ms[p.cls] += ms;
        m.kc_bytes[p.cls] += p.bytes;
        m.kc_launches[p.cls]++;
        m.free_events.push_back(p.start);
        m.free_events.push_back(p.stop);
    }
    m.pending.clear();
}

// Brackets one launch with events when timing is on.
struct Timed {
    dcrf_model& m;
    hipStream_t s;
    int cls;
    double bytes;
    hipEvent_t e0 = nullptr;
    Timed(dcrf_model& m_, hipStream_t s_, int cls_, double bytes_) : m(m_), s(s_), cls(cls_), bytes(bytes_) {
        if (m.timing) {
            e0 = take_event(m);
            DCRF_HIP_CHECK(hipEventRecord(e0, s));
        }
    }
    void done() {
        if (!e0) return;
        hipEvent_t e1 = take_event(m);
        DCRF_HIP_CHECK(hipEventRecord(e1, s));
        m.pending.push_back(PendingTiming{e0, e1, cls, bytes});
        e0 = nullptr;
    }
};

void ensure_device(dcrf_model& m) {
    if (m.dev_ready) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        throw Error{DCRF_ERR_DEVICE, "no HIP device is available"};
    DCRF_HIP_CHECK(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
    int dev = 0, lds = 0;
    DCRF_HIP_CHECK(hipGetDevice(&dev));
    DCRF_HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    m.lds_cap = (size_t)std::max(lds, 0);
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

// dst = src on the device (large lists through the staging area; then the copy is
// complete on return, else it is queued on s).
template <typename T>
void upload(DevBuf<T>& dst, const std::vector<T>& src, hipStream_t s) {
    dst.reserve(src.size());
    if (src.empty()) return;
    const size_t bytes = sizeof(T) * src.size();
    if (bytes >= kStagedMin)
        copy_to_device(dst.get(), src.data(), bytes, s);
    else
        DCRF_HIP_CHECK(hipMemcpyAsync(dst.get(), src.data(), bytes, hipMemcpyHostToDevice, s));
}

void ensure_unary(dcrf_model& m) {
    if (!m.has_unary || !m.unary_dirty) return;
    const size_t n = (size_t)m.N() * m.L;
    if (m.logistic) {
        // LogisticUnaryEnergy::get (unary.cpp:50-52) on the matrix cores
        m.unary.reserve(n);
        upload(m.lL_dev, m.lL, m.stream);
        launch_logistic_unary(m.lfFN.get(), m.lL_dev.get(), m.N(), m.lF, m.L, m.unary.get(), m.stream);
        DCRF_HIP_CHECK(hipStreamSynchronize(m.stream));
        m.unary_dirty = false;
        return;
    }
    m.stage.reserve(n);
    m.unary.reserve(n);
    for (int i = 0; i < (int)m.unary_unset.size(); i++)
        if (m.unary_unset[i]) {  // an instance never given a unary: zeros, as the first set left them
            const int b = m.geo.pix_base[i], Ni = m.geo.N[i];
            for (int l = 0; l < m.L; l++) std::fill_n(m.unary_host.data() + (size_t)l * m.N() + b, Ni, 0.f);
            m.unary_unset[i] = 0;
        }
    copy_to_device(m.stage.get(), m.unary_host.data(), sizeof(float) * n, m.stream);
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

TileGeom tile_geom(const dcrf_model& m) {
    return TileGeom{m.tiles.base_dev.get(), m.geod.pix_base.get(), m.geo.is2d ? m.geod.W.get() : nullptr,
                    m.geo.n_inst, m.geo.is2d ? 1 : 0, m.tiles.side};
}

// Most entries and segments of one tile, from first-of-tile tables.
int max_run(const std::vector<int>& first) {
    int mx = 0;
    for (size_t t = 0; t + 1 < first.size(); t++) mx = std::max(mx, first[t + 1] - first[t]);
    return mx;
}

// The splat's wave mode loads a wave's entries as one range: consecutive segments must
// adjoin (segment c+1 begins where segment c ends), which the tile-major layout gives.
void require_adjoining(const std::vector<Chunk>& ch) {
    for (size_t c = 1; c < ch.size(); c++)
        if (ch[c].beg != ch[c - 1].end) throw Error{DCRF_ERR_STATE, "splat layout: segments do not adjoin"};
}

// norm = K*1 for one term (pairwise.cpp:44), then the normalisation of :46-61.
void compute_norm(dcrf_model& m, Term& t) {
    TermLattice& lat = t.lat;
    const int N = m.N(), d1 = lat.d + 1, V = lat.V, E = N * d1;
    hipStream_t s = m.stream;
    DevBuf<int> tseg, te;
    DevBuf<int2> ent;
    DevBuf<float> sw, vals0, vals1, ones;
    DevBuf<Chunk> ch;
    DevBuf<Combine> cb;
    DevBuf<int2> nbr;
    ent.reserve(E);
    sw.reserve(E);
    // one term: entry positions are the sorted positions; rows V+1.. hold the partials
    std::vector<Chunk> hch = lat.chunks;
    for (auto& c : hch) c.out = c.out >= 0 ? c.out : V + 1 + (-c.out - 1);
    require_adjoining(hch);
    std::vector<Combine> hcb = lat.combines;
    for (auto& c : hcb) c.first += V + 1;
    upload(ch, hch, s);
    upload(cb, hcb, s);
    upload(tseg, lat.chunk_tile_ptr, s);
    upload(te, lat.tile_ptr, s);
    const TileGeom tg = tile_geom(m);
    const TiledSplat ts{ch.get(), tseg.get(), te.get(), lat.n_tiles, tg, max_run(lat.tile_ptr),
                        max_run(lat.chunk_tile_ptr)};
    const bool tiled = tiled_splat_lds(1, tg.side, ts.e_max, ts.s_max, m.lds_cap) > 0;
    launch_entry_gather(lat.csr_e.get(), E, d1, lat.bary.get(), nullptr, 0, nullptr, nullptr, 0, tiled ? &tg : nullptr,
                        ent.get(), s);
    const int rows = V + 1 + lat.n_partials;
    vals0.reserve(rows);
    vals1.reserve(V + 1);
    launch_fill(vals0.get(), 0.f, rows, s);
    launch_fill(vals1.get(), 0.f, V + 1, s);
    nbr.reserve((size_t)d1 * V);
    launch_remap_nbr(lat.nbr.get(), nbr.get(), d1 * V, 0, V, s);
    ones.reserve(N);
    launch_fill(ones.get(), 1.f, N, s);
    if (tiled)
        launch_splat_tiled(ts, ent.get(), ones.get(), vals0.get(), 1, 1, m.lds_cap, s);
    else
        launch_splat(ch.get(), (int)hch.size(), ent.get(), ones.get(), 1, vals0.get(), 1, 1, s);
    launch_combine(cb.get(), (int)hcb.size(), vals0.get(), 1, 1, s);
    BlurTable bt{};
    bt.nterms = 1;
    bt.t[0] = BlurTerm{0, V, lat.d, nbr.get()};
    float* bufs[2] = {vals0.get(), vals1.get()};
    for (int step = 0; step <= lat.d; step++)
        launch_blur(bufs[step & 1], bufs[(step + 1) & 1], 0, V, 1, 1, bt, step, 0, s);
    DevBuf<int> gv;
    gv.reserve(E);
    launch_slice_weights(lat.vtx.get(), lat.bary.get(), nullptr, 0, slice_alpha(lat.d), N, d1, 0, gv.get(), sw.get(), s);
    SliceTable st{};
    st.nterms = 1;
    st.t[0] = SliceTerm{gv.get(), sw.get(), bufs[d1 & 1], d1, -1, nullptr, nullptr, nullptr, 1};
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

// Multi-step blur schedule: a lattice runs its d+1 steps in ceil((d+1)/K) passes of at
// most K steps, spread as evenly as possible (longer passes first); the number of passes
// falls with d, so in the global filter pass p covers a prefix of the vertex space.
// K = 2 by default: passes of three steps (27 rows per vertex) for every lattice
// measured slower at 1024x768x21 (49 us against 28 us of blur per iteration), as the
// 27 gathers per vertex of the large Gaussian lattice cost more than the launch saved.
constexpr int kMaxPassSteps = 2;
std::vector<int> pass_schedule(int steps, int K = kMaxPassSteps) {
    const int n = (steps + K - 1) / K;
    std::vector<int> out(n, steps / n);
    for (int i = 0; i < steps % n; i++) out[i]++;
    return out;
}

// Most steps per pass of every term (dcrf_set_blur_steps: 2, the default, or 3).  Each
// pass is one launch over the terms that still have passes, and a short launch costs
// ~5 us whatever its work (a pass over the 36 k vertices of the 1024x768 bilateral
// lattice alone: 5.1-5.5 us), but three-step passes for that lattice only (3 + 3 beside
// the Gaussian lattice's 2 + 1: two launches instead of three) measured 2 x 25.8 us of
// blur per iteration against 3 x 12.9 us: its 27-row trees, folded one 9-row subtree
// after another, make each launch's tail three memory round trips long.
std::vector<int> pass_step_limits(const dcrf_model& m) {
    return std::vector<int>(m.terms.size(), m.blur_steps);
}

// Pass tables of one lattice (forward axes 0..d, reverse d..0) over its neighbour rows
// nbr [(d+1)][V] (global rows from vbase, zero row zero_row), passes of at most K steps.
void build_pass_tables(const int2* nbr, int V, int d, int vbase, int zero_row, std::vector<int>& steps,
                       std::vector<int>& off, DevBuf<int>& fwd, DevBuf<int>& rev, hipStream_t s,
                       int K = kMaxPassSteps) {
    steps = pass_schedule(d + 1, K);
    off.assign(1, 0);
    for (int q : steps) off.push_back(off.back() + V * (q == 1 ? 3 : q == 2 ? 9 : 27));
    fwd.reserve(off.back());
    rev.reserve(off.back());
    int first = 0;
    for (size_t p = 0; p < steps.size(); p++) {
        const int q = steps[p];
        int fa[3], ta[3];
        for (int j = 0; j < q; j++) {
            fa[j] = first + j;
            ta[j] = d - first - j;
        }
        launch_pass_table(nbr, V, vbase, zero_row, q, fa, fwd.get() + off[p], s);
        launch_pass_table(nbr, V, vbase, zero_row, q, ta, rev.get() + off[p], s);
        first += q;
    }
}

// ---------------------------------------------------------------- kernel-parameter gradient
// The blur passes of one layout (x: the x chains, else the forward and reverse chains
// of the paired inputs) from the splat rows src through the ping-pong pair buf: pass p
// runs the chains of the parts with more than p passes over rows [0, vp[p]).  Part i's
// result is in buf[(passes - 1) & 1].
void kg_blur(const KgPlan& P, bool x, const float* src, float* const buf[2], int Lc, hipStream_t s) {
    const float* in = src;
    for (int p = 0; p < P.passes; p++) {
        PassTable pt{};
        pt.nterms = 0;
        for (size_t i = 0; i < P.parts.size() && P.parts[i].passes > p; i++) {
            const KgPart& part = P.parts[i];
            const int q = part.steps[p];
            const int* tab = P.tab.get();
            const std::vector<int>& off = P.tab_off[p];
            if (x) {
                pt.t[pt.nterms++] = PassTerm{part.X, part.V, tab + off[3 * i + 2], q};
            } else {
                pt.t[pt.nterms++] = PassTerm{part.F, part.V, tab + off[3 * i], q};
                pt.t[pt.nterms++] = PassTerm{part.R, part.V, tab + off[3 * i + 1], q};
            }
        }
        pt.zero_row = x ? P.Z1 : P.Z2;
        launch_blur_pass(in, buf[p & 1], 0, x ? P.vp1[p] : P.vp2[p], Lc, Lc, pt, s);
        in = buf[p & 1];
    }
}

// Fill plan P for the model terms `terms` (at most DCRF_MAX_TERMS / 2; P's device
// buffers are reused where they are large enough): the layouts of KgPlan, the merged
// segment and entry lists (the x layout's lists derived on the device), the pass tables
// of every chain, and per term the slice weights and the all-ones input splatted and
// blurred in reverse order (one value per vertex, through the plan's own launches: the
// same segments, sums and pass order as a filter of that term alone).
void build_kg_plan(dcrf_model& m, std::vector<int> terms, KgPlan& P) {
    hipStream_t s = m.stream;
    const int N = m.N();
    const size_t L = m.L;
    auto steps_of = [&](int k) { return pass_schedule(m.terms[k]->lat.d + 1, kMaxPassSteps); };
    std::stable_sort(terms.begin(), terms.end(),
                     [&](int a, int b) { return steps_of(a).size() > steps_of(b).size(); });
    std::vector<KgPart> old = std::move(P.parts);
    P.parts.clear();
    P.passes = 0;
    int row = 0, xrow = 0, parts_total = 0;
    size_t E = 0;
    for (size_t i = 0; i < terms.size(); i++) {
        const int k = terms[i];
        const TermLattice& lat = m.terms[k]->lat;
        KgPart part;
        if (i < old.size()) part.sw = std::move(old[i].sw);
        part.term = k;
        part.V = lat.V;
        part.steps = steps_of(k);
        part.passes = (int)part.steps.size();
        part.F = row;
        part.R = row + lat.V;
        part.X = xrow;
        row += 2 * lat.V;
        xrow += lat.V;
        parts_total += lat.n_partials;
        E += (size_t)N * (lat.d + 1);
        P.passes = std::max(P.passes, part.passes);
        P.parts.push_back(std::move(part));
    }
    P.Z2 = row;
    P.Z1 = xrow;
    // merged segments: entries of the i-th term from e0, its partial rows from p0 after
    // the zero row
    std::vector<Chunk> ch2;
    std::vector<Combine> cb2;
    P.ent.reserve(E);
    size_t e0 = 0;
    int p0 = 0;
    for (size_t i = 0; i < P.parts.size(); i++) {
        const KgPart& part = P.parts[i];
        const TermLattice& lat = m.terms[part.term]->lat;
        const int d1 = lat.d + 1;
        launch_entry_gather(lat.csr_e.get(), N * d1, d1, lat.bary.get(), nullptr, 0, nullptr, nullptr, 0, nullptr,
                            P.ent.get() + e0, s, (int)i * N);
        for (const Chunk& c : lat.chunks)
            ch2.push_back(Chunk{part.F + c.v, c.beg + (int)e0, c.end + (int)e0,
                                c.out >= 0 ? part.F + c.out : P.Z2 + 1 + p0 + (-c.out - 1)});
        for (const Combine& c : lat.combines)
            cb2.push_back(Combine{part.F + c.v, P.Z2 + 1 + p0 + c.first, c.count, c.pad});
        e0 += (size_t)N * d1;
        p0 += lat.n_partials;
    }
    require_adjoining(ch2);
    if (!ch2.empty() && (ch2.front().beg != 0 || (size_t)ch2.back().end != E))
        throw Error{DCRF_ERR_STATE, "kernel gradient: segments do not cover the entries"};
    upload(P.ch2, ch2, s);
    upload(P.cb2, cb2, s);
    P.nch = (int)ch2.size();
    P.ncb = (int)cb2.size();
    KgRowMap rm{};
    rm.n = (int)P.parts.size();
    rm.Z2 = P.Z2;
    rm.Z1 = P.Z1;
    for (size_t i = 0; i < P.parts.size(); i++) {
        rm.F[i] = P.parts[i].F;
        rm.X[i] = P.parts[i].X;
    }
    P.ch1.reserve(P.nch);
    P.cb1.reserve(P.ncb);
    launch_kg_relayout(P.ch2.get(), P.nch, P.cb2.get(), P.ncb, rm, P.ch1.get(), P.cb1.get(), s);
    // pass tables: pass p of term i expands the axes of its p-th pass; the forward chain
    // reads rows from F, the reverse chain from F in its first pass (the splat rows) and
    // from R after, the x chain from X
    size_t tab_total = 0;
    for (const KgPart& part : P.parts)
        for (int q : part.steps) tab_total += (size_t)3 * part.V * (q == 1 ? 3 : 9);
    P.tab.reserve(std::max<size_t>(tab_total, 1));
    P.tab_off.assign(P.passes, std::vector<int>(3 * P.parts.size(), -1));
    P.vp2.assign(P.passes, 0);
    P.vp1.assign(P.passes, 0);
    size_t toff = 0;
    for (size_t i = 0; i < P.parts.size(); i++) {
        KgPart& part = P.parts[i];
        const TermLattice& lat = m.terms[part.term]->lat;
        const int V = lat.V, d = lat.d, d1 = d + 1;
        DevBuf<int2> nF, nR, nX;  // the neighbour rows in each chain's numbering
        nF.reserve((size_t)d1 * V);
        nR.reserve((size_t)d1 * V);
        nX.reserve((size_t)d1 * V);
        launch_remap_nbr(lat.nbr.get(), nF.get(), d1 * V, part.F, P.Z2, s);
        launch_remap_nbr(lat.nbr.get(), nR.get(), d1 * V, part.R, P.Z2, s);
        launch_remap_nbr(lat.nbr.get(), nX.get(), d1 * V, part.X, P.Z1, s);
        int first = 0;
        for (int p = 0; p < part.passes; p++) {
            const int q = part.steps[p];
            int fa[3], ta[3];
            for (int j = 0; j < q; j++) {
                fa[j] = first + j;
                ta[j] = d - first - j;
            }
            const size_t n = (size_t)V * (q == 1 ? 3 : 9);
            std::vector<int>& off = P.tab_off[p];
            off[3 * i] = (int)toff;
            launch_pass_table(nF.get(), V, part.F, P.Z2, q, fa, P.tab.get() + toff, s);
            toff += n;
            off[3 * i + 1] = (int)toff;
            if (p == 0)
                launch_pass_table(nF.get(), V, part.F, P.Z2, q, ta, P.tab.get() + toff, s);
            else
                launch_pass_table(nR.get(), V, part.R, P.Z2, q, ta, P.tab.get() + toff, s);
            toff += n;
            off[3 * i + 2] = (int)toff;
            launch_pass_table(nX.get(), V, part.X, P.Z1, q, fa, P.tab.get() + toff, s);
            toff += n;
            P.vp2[p] = part.R + V;  // the parts are in decreasing pass count
            P.vp1[p] = part.X + V;
            first += q;
        }
        DevBuf<int> gv;
        gv.reserve((size_t)N * d1);
        part.sw.reserve((size_t)N * d1);
        launch_slice_weights(lat.vtx.get(), lat.bary.get(), nullptr, 0, slice_alpha(d), N, d1, 0, gv.get(),
                             part.sw.get(), s);
    }
    // vertex rows: the zero rows stay zero (splats write vertex and partial rows, the
    // blur passes rows below the zero row)
    auto zeroed = [&](DevBuf<float>& b, size_t n) {
        b.reserve(n);
        launch_fill(b.get(), 0.f, n, s);
    };
    const size_t rows2 = (size_t)P.Z2 + 1 + parts_total;
    zeroed(P.S2, rows2 * 2 * L);
    zeroed(P.X2[0], (size_t)(P.Z2 + 1) * 2 * L);
    zeroed(P.X2[1], (size_t)(P.Z2 + 1) * 2 * L);
    zeroed(P.S1, ((size_t)P.Z1 + 1 + parts_total) * L);
    zeroed(P.B1, (size_t)(P.Z1 + 1) * L);
    // the all-ones input of every term (stacked, one value per pixel) through the plan's
    // splat and both chains; the reverse rows are the terms' K^T 1
    zeroed(P.ones[0], rows2);
    zeroed(P.ones[1], (size_t)P.Z2 + 1);
    zeroed(P.ones[2], (size_t)P.Z2 + 1);
    DevBuf<float> in1;
    in1.reserve((size_t)P.parts.size() * N);
    launch_fill(in1.get(), 1.f, (size_t)P.parts.size() * N, s);
    launch_splat(P.ch2.get(), P.nch, P.ent.get(), in1.get(), 1, P.ones[0].get(), 1, 1, s);
    launch_combine(P.cb2.get(), P.ncb, P.ones[0].get(), 1, 1, s);
    float* const obuf[2] = {P.ones[1].get(), P.ones[2].get()};
    kg_blur(P, false, P.ones[0].get(), obuf, 1, s);
    for (KgPart& part : P.parts) part.ones_rev = obuf[(part.passes - 1) & 1] + part.R;
    P.gen = m.lat_gen;
    DCRF_HIP_CHECK(hipStreamSynchronize(s));  // the uploads from host lists are complete
}

// The plans of the terms with kernel parameters, rebuilt (in place) when a lattice
// changed.  A plan takes at most DCRF_MAX_TERMS / 2 terms (a pass launch holds two chains
// per term), and its stacked input rows, entries and vertex rows stay within int.
void ensure_kg_plans(dcrf_model& m) {
    if (!m.kgplans.empty() && m.kgplans.front()->gen == m.lat_gen) return;
    const long long N = m.N(), lim = std::numeric_limits<int>::max();
    std::vector<std::vector<int>> groups;
    std::vector<int> g;
    long long e = 0, rows = 0;
    for (int k = 0; k < (int)m.terms.size(); k++) {
        const Term& t = *m.terms[k];
        if (!t.num_kparams()) continue;
        const long long ek = N * (t.lat.d + 1), rk = 2LL * t.lat.V + t.lat.n_partials;
        if (!g.empty() && ((int)g.size() == DCRF_MAX_TERMS / 2 || e + ek > lim || rows + rk + 1 > lim ||
                           (long long)(g.size() + 1) * N > lim)) {
            groups.push_back(g);
            g.clear();
            e = rows = 0;
        }
        g.push_back(k);
        e += ek;
        rows += rk;
    }
    if (!g.empty()) groups.push_back(g);
    m.kgplans.resize(groups.size());
    for (size_t i = 0; i < groups.size(); i++) {
        if (!m.kgplans[i]) m.kgplans[i] = std::make_unique<KgPlan>();
        build_kg_plan(m, groups[i], *m.kgplans[i]);
    }
}

// acc_k += kernelGradient(b, Q_t) for every term k of plan P (pairwise.cpp:202-207,
// 152-163, 88-114): fg = -G(x, 1) + G(a, c) with a, c and x per the normalisation type.
// Per term one pass forms lbl = C (w Q_t) and the inputs a and c; the [a | c] of all the
// plan's terms are splatted together and blurred in both orders: the forward rows give
// K a (G(a, c), direction 0) and K c, the reverse rows K^T a and K^T c (G(a, c),
// direction 1).  x (from the slices K^T a and K c in one pass per term) needs one more
// splat and forward blur, and G(x, 1)'s direction 1 is the term's reverse-blurred ones.
// The four slicing gradients and the parameter products then run in one pass per term
// (its partial rows at partial + k * pstride, summed after the last iteration) where the
// lane-group kernel covers the shape, else per direction with a reduction into acc.
void kernel_grad_group(dcrf_model& m, const KgPlan& P, const float* b, const float* Qt, const float* kw,
                       const std::vector<int>& koff, size_t pstride, hipStream_t s) {
    const int N = m.N(), L = m.L, L2 = 2 * L;
    const size_t NL = (size_t)N * L;
    const int T = (int)P.parts.size();
    float* lbl = m.kgbuf.get();                // [T][N][L]
    float* pk = lbl + (size_t)T * NL;          // [T][N][2L]
    float* xb = pk + (size_t)2 * T * NL;       // [T][N][L]
    bool any_x = false;
    for (int i = 0; i < T; i++) {
        const int k = P.parts[i].term;
        const Term& t = *m.terms[k];
        const bool has_x = t.ntype != DCRF_NO_NORMALIZATION;
        const bool need_fa = has_x && t.ntype != DCRF_NORMALIZE_AFTER;  // x reads fa and lbl
        any_x = any_x || has_x;
        SliceTable st{};
        st.nterms = 1;
        st.t[0] = SliceTerm{nullptr, nullptr, nullptr, 1, t.ckind, t.d_compat.get(), kw ? kw + k : nullptr, nullptr, L};
        launch_kg_prepare(Qt, b, t.norm.get(), N, L, st, t.ntype, need_fa ? lbl + i * NL : nullptr,
                          pk + (size_t)i * 2 * NL, s);
    }
    launch_splat(P.ch2.get(), P.nch, P.ent.get(), pk, L2, P.S2.get(), L2, L2, s);
    launch_combine(P.cb2.get(), P.ncb, P.S2.get(), L2, L2, s);
    float* const pbuf[2] = {P.X2[0].get(), P.X2[1].get()};
    kg_blur(P, false, P.S2.get(), pbuf, L2, s);
    float* const xbuf[2] = {P.B1.get(), P.S1.get()};  // pass 0 reads the splat rows S1
    auto fwd_of = [&](int i) { return pbuf[(P.parts[i].passes - 1) & 1] + (size_t)P.parts[i].F * L2; };
    auto rev_of = [&](int i) { return pbuf[(P.parts[i].passes - 1) & 1] + (size_t)P.parts[i].R * L2; };
    if (any_x) {
        for (int i = 0; i < T; i++) {
            const Term& t = *m.terms[P.parts[i].term];
            float* x = xb + (size_t)i * NL;
            if (t.ntype == DCRF_NO_NORMALIZATION) {  // splatted with the others, never read
                launch_fill(x, 0.f, NL, s);
                continue;
            }
            const bool need_fa = t.ntype != DCRF_NORMALIZE_AFTER, need_fb = t.ntype != DCRF_NORMALIZE_BEFORE;
            SliceTable st{};
            st.nterms = 2;
            const int* vtx = t.lat.vtx.get();
            const float* sw = P.parts[i].sw.get();
            const int d1 = t.lat.d + 1;
            st.t[0] = SliceTerm{vtx, sw, need_fa ? rev_of(i) : nullptr, d1, -1, nullptr, nullptr, nullptr, L2};  // K^T a
            st.t[1] = SliceTerm{vtx, sw, need_fb ? fwd_of(i) + L : nullptr, d1, -1, nullptr, nullptr, nullptr, L2};  // K c
            launch_kg_x(b, lbl + (size_t)i * NL, t.norm.get(), N, L, st, t.ntype, x, s);
        }
        launch_splat(P.ch1.get(), P.nch, P.ent.get(), xb, L, P.S1.get(), L, L, s);
        launch_combine(P.cb1.get(), P.ncb, P.S1.get(), L, L, s);
        kg_blur(P, true, P.S1.get(), xbuf, L, s);
    }
    for (int i = 0; i < T; i++) {
        const KgPart& part = P.parts[i];
        const int k = part.term;
        const Term& t = *m.terms[k];
        const int d = t.lat.d;
        const bool has_x = t.ntype != DCRF_NO_NORMALIZATION;
        const float* fwd = fwd_of(i);
        const float* rev = rev_of(i);
        const float* xr = has_x ? xbuf[(part.passes - 1) & 1] + (size_t)part.X * L : nullptr;
        const float* pki = pk + (size_t)i * 2 * NL;
        const float* x = xb + (size_t)i * NL;
        const int* vtx = t.lat.vtx.get();
        const int* rank = t.lat.rank.get();
        const bool full = t.ktype == DCRF_FULL_KERNEL;
        double* partial = m.kpartial.get() + (size_t)k * pstride;
        if (kg_grad_fused(L, d)) {
            launch_kg_grad(vtx, rank, fwd, rev, pki, xr, part.ones_rev, x, t.feat0.get(), m.geod.pix_base.get(),
                           m.geod.pad_base.get(), m.geo.n_inst, N, L, d, full, partial, s);
            continue;
        }
        // more labels than a wave's lanes hold, or d > 6: one thread per point and direction
        float* dfr = m.kgdf.get();
        float* dfg = dfr + (size_t)N * d;
        if (has_x) {
            launch_slice_grad(xr, L, 1, vtx, rank, nullptr, 0, N, L, d, 0, dfr, s);
            launch_slice_grad(part.ones_rev, 1, 0, vtx, rank, x, L, N, L, d, 1, dfr, s);
        }
        launch_slice_grad(fwd, L2, 1, vtx, rank, pki + L, L2, N, L, d, 0, dfg, s);
        launch_slice_grad(rev + L, L2, 1, vtx, rank, pki, L2, N, L, d, 1, dfg, s);
        launch_kparam_reduce(has_x ? dfr : nullptr, dfg, t.feat0.get(), m.geod.pix_base.get(), m.geod.pad_base.get(),
                             m.geo.n_inst, N, d, full, partial, m.kacc.get() + koff[k], s);
    }
}

int num_kernel_params(const dcrf_model& m) {
    int n = 0;
    for (const auto& t : m.terms) n += t->num_kparams();
    return n;
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
    t->d_compat.reserve(t->compat.size());
    DCRF_HIP_CHECK(hipMemcpy(t->d_compat.get(), t->compat.data(), sizeof(float) * t->compat.size(),
                             hipMemcpyHostToDevice));
    DevBuf<float> feat;
    feat.reserve((size_t)m.geo.npad_total() * d);
    make_features(feat.get());
    if (!m.tiles_ready) {
        make_tile_map(m.geo, m.tile_side ? m.tile_side : default_tile_side(m.L), m.tiles);
        m.tiles_ready = true;
    }
    build_lattice(m.geo, m.geod, m.tiles, feat.get(), d, t->lat, m.hash_load, m.stream);
    if (t->lat.overflow)
        throw Error{DCRF_ERR_OVERFLOW,
                    "a lattice coordinate exceeds the int16 range of the lattice keys; scale the features down"};
    build_chunks(t->lat, m.chunk());
    compute_norm(m, *t);
    // DenseKernel keeps f_ and starts from unit scales / the identity (pairwise.cpp:116-122)
    t->feat0 = std::move(feat);
    t->kparams.assign(ktype == DCRF_FULL_KERNEL ? (size_t)d * d : (size_t)d, 0.0f);
    for (int j = 0; j < d; j++) t->kparams[ktype == DCRF_FULL_KERNEL ? j * d + j : j] = 1.0f;
    m.terms.push_back(std::move(t));
    m.layout_valid = false;
    m.lat_gen++;
    m.kw_set = false;  // term weights are per term: set them again after adding terms
    m.train_n = -1;
}

SliceTable slice_table(const dcrf_model& m, bool transpose, const float* kw = nullptr, float* mout = nullptr);

// The vertex buffer holding term k's filtered rows after the global filter: one buffer
// switch per blur step, or per multi-step pass.
int result_parity(const dcrf_model& m, int k) {
    return (m.pairs ? (int)m.tl[k]->pass_steps.size() : m.terms[k]->lat.d + 1) & 1;
}

// Launches per filter with multi-step passes (after the terms' pass tables exist).
int blur_launches(const dcrf_model& m) {
    int P = 0;
    for (const auto& tl : m.tl) P = std::max(P, (int)tl->pass_steps.size());
    return P;
}

void ensure_layout(dcrf_model& m) {
    if (m.layout_valid) return;
    m.drop_graph();
    m.qblocks_ready = false;
    hipStream_t s = m.stream;
    const int T = (int)m.terms.size();
    const int N = m.N(), L = m.L;
    for (auto& t : m.terms)
        if (t->lat.chunk_len != m.chunk()) build_chunks(t->lat, m.chunk());
    // terms in layout order: more blur passes first (then higher d), so that the terms
    // of pass p are a prefix of the vertex space
    const std::vector<int> step_limit = pass_step_limits(m);
    std::vector<int> npass(T);
    for (int k = 0; k < T; k++) npass[k] = (int)pass_schedule(m.terms[k]->lat.d + 1, step_limit[k]).size();
    m.order.resize(T);
    std::iota(m.order.begin(), m.order.end(), 0);
    std::stable_sort(m.order.begin(), m.order.end(), [&](int a, int b) {
        if (npass[a] != npass[b]) return npass[a] > npass[b];
        return m.terms[a]->lat.d > m.terms[b]->lat.d;
    });
    m.vbase.assign(T, 0);
    std::vector<int> pbase(T, 0);
    int V = 0, P = 0, E = 0;
    m.max_d = -1;
    m.all_symmetric = true;
    for (int k : m.order) {
        const TermLattice& lat = m.terms[k]->lat;
        m.vbase[k] = V;
        V += lat.V;
        E += N * (lat.d + 1);
        m.max_d = std::max(m.max_d, lat.d);
        if (m.terms[k]->ntype != DCRF_NORMALIZE_SYMMETRIC && m.terms[k]->ntype != DCRF_NO_NORMALIZATION)
            m.all_symmetric = false;
    }
    for (int k : m.order) {
        pbase[k] = V + 1 + P;
        P += m.terms[k]->lat.n_partials;
    }
    m.V_total = V;
    m.rows0 = V + 1 + P;
    m.E_total = E;
    // Segments and entries, tile by tile, the terms in layout order inside a tile: tile
    // t's entries of term k (sorted positions tile_ptr[t] ..) move to ebase[k][t] + j.
    const int NT = m.tiles.n_tiles();
    const LaunchShape csh = shape_for(L);
    // the LDS-tiled splat (small tiles only) if the largest tile's entries and segments fit
    int e_max = 0, s_max = 0;
    for (int t = 0; t < NT; t++) {
        int et = 0, st = 0;
        for (int k : m.order) {
            const TermLattice& lat = m.terms[k]->lat;
            et += lat.tile_ptr[t + 1] - lat.tile_ptr[t];
            st += lat.chunk_tile_ptr[t + 1] - lat.chunk_tile_ptr[t];
        }
        e_max = std::max(e_max, et);
        s_max = std::max(s_max, st);
    }
    const TileGeom tg = tile_geom(m);
    m.tiled = tiled_splat_lds(L, tg.side, e_max, s_max, m.lds_cap) > 0;
    // In-wave runs (wave-mode shapes, global splat): a vertex's segments that sit side by
    // side in one wave are added inside the wave (k_splat folds each such run into its
    // first lane group); the later segments of a run get out = -1 and write nothing, and
    // the vertex gets one partial row per run, or its final row when it has one run.  At
    // 1024x768x21 this cuts the partial rows from 188 k to 72 k and the vertices with more
    // than 16 of them from 2,581 to 52.  The runs are part of the summation order, so
    // they must not depend on where a segment sits in the whole list: every (tile, term)
    // group starts a wave (empty fillers, v = out = -1, pad the group before), and the
    // empty segments of vertices that only padding lanes touch go to waves of their own
    // at the end.  An instance of a batch then gets the bits of the same image alone.
    // Block-mode shapes and the tiled splat keep one row per segment.
    const bool runs = csh.wave && !m.tiled;
    const int per = csh.per;
    std::vector<Chunk>& gch = m.gch_scratch;  // (capacity kept between layout builds)
    gch.clear();
    std::vector<int>& nruns = m.runs_scratch;
    std::vector<int>& pfirst = m.first_scratch;
    if (runs) {
        nruns.assign(V + 1, 0);
        pfirst.assign(V + 1, -1);
    }
    int slot = 0, prev = -1;  // position inside the current wave, the vertex before it
    auto put = [&](const Chunk& x) {  // marks run heads (out = 1) and continuations (out = -1)
        const bool h = x.v >= 0 && (slot == 0 || prev != x.v);
        gch.push_back(Chunk{x.v, x.beg, x.end, h ? 1 : -1});
        if (h) nruns[x.v]++;
        prev = x.v;
        if (++slot == per) slot = 0;
    };
    std::vector<Combine> gcb;
    std::vector<int> gts(NT + 1), gte(NT + 1);
    std::vector<std::vector<int>> ebase(T, std::vector<int>(NT));
    std::vector<int> empty_segs;  // vertices of the empty segments (padding lanes only)
    int e = 0;
    for (int t = 0; t < NT; t++) {
        gts[t] = (int)gch.size();
        gte[t] = e;
        for (int k : m.order) {
            const TermLattice& lat = m.terms[k]->lat;
            const int eb = e - lat.tile_ptr[t];
            ebase[k][t] = eb;
            for (int c = lat.chunk_tile_ptr[t]; c < lat.chunk_tile_ptr[t + 1]; c++) {
                Chunk x = lat.chunks[c];
                x.v += m.vbase[k];
                x.beg += eb;
                x.end += eb;
                if (runs) {
                    if (x.beg == x.end)
                        empty_segs.push_back(x.v);
                    else
                        put(x);
                    continue;
                }
                x.out = x.out >= 0 ? x.out + m.vbase[k] : pbase[k] + (-x.out - 1);
                gch.push_back(x);
            }
            e += lat.tile_ptr[t + 1] - lat.tile_ptr[t];
            if (runs)
                while (slot) put(Chunk{-1, e, e, 0});
        }
    }
    if (e != E) throw Error{DCRF_ERR_STATE, "splat layout: entry count mismatch"};
    if (runs) {
        for (int v : empty_segs) put(Chunk{v, e, e, 0});
        while (slot) put(Chunk{-1, e, e, 0});
    }
    gts[NT] = (int)gch.size();
    gte[NT] = e;
    require_adjoining(gch);
    upload(m.g_tile_seg, gts, s);
    upload(m.g_tile_e, gte, s);
    std::vector<Combine> combs;  // (vertex, first partial row, rows) of every split vertex
    if (runs) {  // outputs: partial rows contiguous per vertex, in the order of its first run
        int prow = V + 1;
        for (Chunk& x : gch) {
            if (x.out < 0) continue;  // a filler or a run's continuation
            const int v = x.v;
            if (nruns[v] == 1) {
                x.out = v;
                continue;
            }
            if (pfirst[v] < 0) {
                pfirst[v] = prow;
                combs.push_back(Combine{v, prow, nruns[v], 0});
                prow += nruns[v];
            }
            x.out = pfirst[v]++;  // pfirst[v] walks the vertex's partial rows
        }
        m.rows0 = prow;
    } else {
        for (int k : m.order)
            for (Combine c : m.terms[k]->lat.combines) {
                c.v += m.vbase[k];
                c.first += pbase[k];
                combs.push_back(c);
            }
    }
    // Vertices split into more than kCombineSplit partial rows must not leave one lane
    // group a long serial chain (before the in-wave runs, the heaviest bilateral vertices
    // at 1024x768 had ~245 partial rows).  Wave-mode shapes recombine in one launch
    // (launch_combine_waves): such a vertex fills a whole wave, its rows cut into up to
    // `per` pieces summed side by side and then added in piece order.  Block-mode shapes
    // recombine in two levels: pieces of about kCombineSplit rows into second-level rows
    // (appended after the partial rows), then those into the vertex row (a second launch).
    std::vector<Combine> gcb2, heavy;
    m.comb_waves = csh.wave;
    for (const Combine& c : combs) {
        if (c.count <= kCombineSplit) {
            gcb.push_back(c);
            continue;
        }
        if (m.comb_waves) {
            const int pieces = std::min(csh.per, (c.count + kCombineSplit - 1) / kCombineSplit);
            for (int j = 0, off = 0; j < csh.per; j++) {
                const int n = j < pieces ? (c.count - off) / (pieces - j) : 0;
                heavy.push_back(Combine{c.v, c.first + off, n, pieces});
                off += n;
            }
            continue;
        }
        const int pieces = (c.count + kCombineSplit - 1) / kCombineSplit;
        const int row2 = m.rows0;
        m.rows0 += pieces;
        for (int j = 0, off = 0; j < pieces; j++) {
            const int n = (c.count - off) / (pieces - j);
            gcb.push_back(Combine{row2 + j, c.first + off, n, 0});
            off += n;
        }
        gcb2.push_back(Combine{c.v, row2, pieces, 0});
    }
    if (!heavy.empty()) {  // light items padded to whole waves, then one wave per heavy vertex
        while (gcb.size() % csh.per) gcb.push_back(Combine{0, 0, 0, 0});
        gcb.insert(gcb.end(), heavy.begin(), heavy.end());
    }
    m.C_total = (int)gch.size();
    m.CB_total = (int)gcb.size();
    m.CB2_total = (int)gcb2.size();
    upload(m.g_chunks, gch, s);
    upload(m.g_comb, gcb, s);
    upload(m.g_comb2, gcb2, s);
    m.tsplat = TiledSplat{m.g_chunks.get(), m.g_tile_seg.get(), m.g_tile_e.get(), NT, tg, e_max, s_max};

    // vertex rows packed (stride L): rows padded to start cache lines (32 floats at
    // L = 21, for Q between iterations too) measured 8 % slower per step at 1024x768x21
    m.ldv = L;
    m.vals[0].reserve((size_t)m.rows0 * m.ldv);
    m.vals[1].reserve((size_t)(V + 1) * m.ldv);
    launch_fill(m.vals[0].get(), 0.f, (size_t)m.rows0 * m.ldv, s);
    launch_fill(m.vals[1].get(), 0.f, (size_t)(V + 1) * m.ldv, s);
    m.g_entF.reserve(E);
    if (!m.all_symmetric) m.g_entT.reserve(E);
    m.tl.clear();
    m.tl.resize(T);
    std::vector<DevBuf<int>> staging(2 * T);  // live until the synchronisation below
    for (int i = 0; i < T; i++) {
        const int k = m.order[i];
        Term& t = *m.terms[k];
        const TermLattice& lat = t.lat;
        const int d1 = lat.d + 1, Ek = N * d1;
        DevBuf<int>& tptr = staging[2 * i];
        DevBuf<int>& eb = staging[2 * i + 1];
        upload(tptr, lat.tile_ptr, s);
        upload(eb, ebase[k], s);
        const TileGeom* local = m.tiled ? &tg : nullptr;
        launch_entry_gather(lat.csr_e.get(), Ek, d1, lat.bary.get(), t.norm.get(), pre_norm(t.ntype, false),
                            tptr.get(), eb.get(), NT, local, m.g_entF.get(), s);
        if (!m.all_symmetric)
            launch_entry_gather(lat.csr_e.get(), Ek, d1, lat.bary.get(), t.norm.get(), pre_norm(t.ntype, true),
                                tptr.get(), eb.get(), NT, local, m.g_entT.get(), s);
        auto lay = std::make_unique<TermLayout>();
        lay->nbr.reserve((size_t)d1 * lat.V);
        launch_remap_nbr(lat.nbr.get(), lay->nbr.get(), d1 * lat.V, m.vbase[k], V, s);
        if (m.pairs)  // forward passes walk axes 0..d, the transpose's d..0 (permutohedral.cpp:556)
            build_pass_tables(lay->nbr.get(), lat.V, lat.d, m.vbase[k], V, lay->pass_steps, lay->pass_off,
                              lay->passF, lay->passT, s, step_limit[k]);
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
    }
    // interleaved slice pairs for the fused mean-field kernel
    m.runs = false;
    m.slcT_valid = false;
    m.fused = m.mf_mode != 1 &&
              fused_meanfield_ok(L, T, T > 0 ? m.terms[0]->lat.d + 1 : 0, T > 1 ? m.terms[1]->lat.d + 1 : 0);
    if (m.fused) {
        int ds = 0;
        for (int k = 0; k < T; k++) ds += m.terms[k]->lat.d + 1;
        m.g_slc.reserve((size_t)N * ds);
        int off = 0;
        for (int k = 0; k < T; k++) {
            const int d1 = m.terms[k]->lat.d + 1;
            launch_interleave_slice(m.tl[k]->vtx.get(), m.tl[k]->swF.get(), N, d1, ds, off, m.g_slc.get(), s);
            off += d1;
        }
        // pixel runs (on request): entry j reads vals[result_parity] of its term
        m.runs = false;
        if (m.mf_mode == 3 && m.V_total < (1 << 30)) {
            unsigned par = 0;
            int j = 0;
            for (int k = 0; k < T; k++) {
                const int d1 = m.terms[k]->lat.d + 1;
                for (int q = 0; q < d1; q++, j++)
                    if (result_parity(m, k)) par |= 1u << j;
            }
            // rectangles of 2 waves' pixels per row x 16 rows (18 x 16 at 21 labels)
            // rectangles two waves of pixels wide, at most run_rows high and at most
            // 24 KiB of slice pairs
            const LaunchShape shp = shape_for(L);
            const int TW = 2 * shp.per;
            const int TH = std::max(1, std::min(m.run_rows, (24 * 1024) / (TW * ds * 8)));
            build_run_slice(m.geo, m.geod, m.g_slc.get(), ds, par, TW, TH, m.rslice, s);
            m.rows_lds = (m.rslice.u_max * L + 1) & ~1;
            SliceTable probe = slice_table(m, false);
            const int cwf = assign_compat_lds(probe, L);
            const size_t lds = meanfield_runs_lds(L, ds, m.rslice.P, m.rows_lds, cwf);
            m.runs = m.mf_mode == 3 && lds <= 64 * 1024 && lds <= m.lds_cap;
        }
    }
    m.prefixV.assign(std::max(m.max_d + 1, 0), 0);
    for (int st = 0; st <= m.max_d; st++)
        for (int k = 0; k < T; k++)
            if (m.terms[k]->lat.d >= st) m.prefixV[st] += m.terms[k]->lat.V;
    m.n_passes = m.pairs ? blur_launches(m) : 0;
    m.prefixP.assign(m.n_passes, 0);
    for (int p = 0; p < m.n_passes; p++)
        for (int k = 0; k < T; k++)
            if ((int)m.tl[k]->pass_steps.size() > p) m.prefixP[p] += m.terms[k]->lat.V;
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

// kw: device [T] term weights or null; mout: device [T][N][L] raw-message output or null.
SliceTable slice_table(const dcrf_model& m, bool transpose, const float* kw, float* mout) {
    SliceTable st{};
    st.nterms = (int)m.terms.size();
    const size_t NL = (size_t)m.N() * m.L;
    for (int k = 0; k < st.nterms; k++) {
        const Term& t = *m.terms[k];
        const TermLayout& lay = *m.tl[k];
        const float* sw = (transpose && !m.all_symmetric) ? lay.swT.get() : lay.swF.get();
        st.t[k] = SliceTerm{lay.vtx.get(), sw, m.vals[result_parity(m, k)].get(), t.lat.d + 1, t.ckind,
                            t.d_compat.get(), kw ? kw + k : nullptr, mout ? mout + (size_t)k * NL : nullptr, m.ldv};
    }
    return st;
}

int compat_raw_len(int ckind, int L) {
    return ckind == DCRF_COMPAT_POTTS ? 1 : (ckind == DCRF_COMPAT_DIAGONAL ? L : L * L);
}
int compat_packed_len(int ckind, int L) {
    return ckind == DCRF_COMPAT_MATRIX ? L * (L + 1) / 2 : compat_raw_len(ckind, L);
}
int stat_len(int ckind, int L) { return ckind == DCRF_COMPAT_MATRIX ? L * L : L; }

// Splat, combine and blur every term's lattice: afterwards term k's filtered rows are
// in vals[result_parity(m, d_k)].  `transpose` selects the transposed filter (reverse blur
// order and the transpose's normaliser placement, pairwise.cpp:63-80).
// Algorithmic bytes of the splat (input rows once; per term the (pixel, weight) pairs,
// the normaliser and the vertex rows written) and of the per-pixel update (unary read,
// Q written; per term the slice pairs, the normaliser and the vertex rows read).  With
// the blur steps' V (2 L 4 + 8) they add up to dcrf_iteration_bytes.
double splat_bytes(const dcrf_model& m) {
    const double N = m.N(), L = m.L;
    double b = N * L * 4.0;
    for (const auto& t : m.terms) b += N * (t->lat.d + 1) * 8.0 + N * 4.0 + (double)t->lat.V * L * 4.0;
    return b;
}
double meanfield_bytes(const dcrf_model& m) { return splat_bytes(m) + m.N() * (double)m.L * 4.0; }

// in: the [N] input rows with row stride ldi (floats).
void run_filter(dcrf_model& m, const float* in, int ldi, bool transpose, hipStream_t s,
                const int2* ent_in = nullptr) {
    const int L = m.L;
    const int2* ent = ent_in ? ent_in : ((transpose && !m.all_symmetric) ? m.g_entT.get() : m.g_entF.get());
    {
        Timed t(m, s, KC_SPLAT, splat_bytes(m));
        if (m.tiled)
            launch_splat_tiled(m.tsplat, ent, in, m.vals[0].get(), m.ldv, L, m.lds_cap, s);  // ldi == L
        else
            launch_splat(m.g_chunks.get(), m.C_total, ent, in, ldi, m.vals[0].get(), m.ldv, L, s);
        t.done();
    }
    if (m.CB_total) {
        Timed t(m, s, KC_COMBINE, 0.0);
        if (m.comb_waves)
            launch_combine_waves(m.g_comb.get(), m.CB_total, m.vals[0].get(), m.ldv, L, s);
        else
            launch_combine(m.g_comb.get(), m.CB_total, m.vals[0].get(), m.ldv, L, s);
        if (m.CB2_total) launch_combine(m.g_comb2.get(), m.CB2_total, m.vals[0].get(), m.ldv, L, s);
        t.done();
    }
    if (m.pairs) {
        // up to three blur steps per pass; terms in layout order, pass p covers the
        // terms with more than p passes (a prefix of the vertex space)
        for (int p = 0; p < m.n_passes; p++) {
            PassTable pt{};
            pt.nterms = (int)m.order.size();
            pt.zero_row = m.V_total;  // zero in both vertex buffers
            double bytes = 0.0;
            for (int i = 0; i < pt.nterms; i++) {
                const int k = m.order[i];
                const TermLattice& lat = m.terms[k]->lat;
                const TermLayout& lay = *m.tl[k];
                const bool on = (int)lay.pass_steps.size() > p;
                const int steps = on ? lay.pass_steps[p] : 1;
                pt.t[i] = PassTerm{m.vbase[k], lat.V,
                                   on ? (transpose ? lay.passT.get() : lay.passF.get()) + lay.pass_off[p] : nullptr,
                                   steps};
                if (on) bytes += (double)lat.V * steps * (2.0 * L * 4 + 8);
            }
            Timed t(m, s, KC_BLUR, bytes);
            const int pr = m.blur_r > 0 && (m.blur_small_v <= 0 || m.prefixP[p] < m.blur_small_v) ? m.blur_r : 0;
            launch_blur_pass(m.vals[p & 1].get(), m.vals[(p + 1) & 1].get(), 0, m.prefixP[p], m.ldv, L, pt, s, pr);
            t.done();
        }
        return;
    }
    const BlurTable bt = blur_table(m);
    for (int step = 0; step <= m.max_d; step++) {
        Timed t(m, s, KC_BLUR, (double)m.prefixV[step] * (2.0 * L * 4 + 8));
        launch_blur(m.vals[step & 1].get(), m.vals[(step + 1) & 1].get(), 0, m.prefixV[step], m.ldv, L, bt, step,
                    transpose ? 1 : 0, s);
        t.done();
    }
}

void prepare(dcrf_model& m) {
    ensure_device(m);
    ensure_unary(m);
    ensure_layout(m);
    const size_t n = (size_t)m.N() * m.L;
    m.Q.reserve(n);
    if (m.fused && m.ldv != m.L) m.Qp.reserve((size_t)m.N() * m.ldv);  // before any graph capture
    if (m.qblocks && !m.qblocks_ready && m.fused && !m.runs && !m.tiled && m.ldv == m.L && m.geo.is2d &&
        m.geo.n_inst == 1) {
        m.g_entB.reserve(m.E_total);
        launch_entry_block(m.g_entF.get(), (size_t)m.E_total, m.geo.W[0], m.g_entB.get(), m.stream);
        m.Qb.reserve(q_block_rows(m.geo.W[0], m.geo.H[0]) * m.L);
        DCRF_HIP_CHECK(hipStreamSynchronize(m.stream));
        m.qblocks_ready = true;
    }
}

// With the fused kernel (and the global splat), the iterations between the first and
// the last keep Q in m.Qp with the vertex-row stride: every row the splat gathers then
// starts a cache line.  The first splat reads the caller's Q, the last update writes it.
void run_inference(dcrf_model& m, int n_iter, const float* U, float* Q, hipStream_t s, bool allow_pad = true) {
    const int L = m.L, N = m.N();
    launch_softmax_neg(U, Q, N, L, s);
    const SliceTable st = slice_table(m, false, m.kw_set ? m.kw.get() : nullptr);
    const bool pad = allow_pad && m.fused && !m.runs && !m.tiled && n_iter > 1 && m.ldv != L && m.Qp.get();
    // Q in 8 x 8 blocks between iterations: the first splat reads the caller's Q, the
    // later ones m.Qb through the block-position entries; the last update writes Q
    const bool blk = !pad && m.qblocks_ready && n_iter > 1 && m.fused && !m.runs && !m.tiled;
    for (int it = 0; it < n_iter; it++) {
        if (blk) {
            run_filter(m, it > 0 ? m.Qb.get() : Q, L, false, s, it > 0 ? m.g_entB.get() : nullptr);
            Timed t(m, s, KC_MEANFIELD, meanfield_bytes(m));
            const bool last = it + 1 == n_iter;
            launch_meanfield_fused(U, last ? Q : m.Qb.get(), L, N, L, st, m.g_slc.get(), s, last ? 0 : m.geo.W[0]);
            t.done();
            continue;
        }
        const bool in_p = pad && it > 0, out_p = pad && it + 1 < n_iter;
        run_filter(m, in_p ? m.Qp.get() : Q, in_p ? m.ldv : L, false, s);
        Timed t(m, s, KC_MEANFIELD, meanfield_bytes(m));
        if (m.runs)
            launch_meanfield_runs(U, Q, m.N(), m.L, st,
                                  RunSliceDev{m.rslice.P, m.rslice.n_runs, m.rows_lds, m.rslice.base.get(),
                                              m.rslice.rows.get(), m.rslice.pix_base.get(), m.rslice.rpix.get(),
                                              m.rslice.loc.get(), m.vals[0].get(), m.vals[1].get()},
                                  s);
        else if (m.fused)
            launch_meanfield_fused(U, out_p ? m.Qp.get() : Q, out_p ? m.ldv : L, N, L, st, m.g_slc.get(), s);
        else
            launch_meanfield(U, Q, m.N(), m.L, st, nullptr, nullptr, s);
        t.done();
    }
}

int matrix_terms(const dcrf_model& m) {
    int nm = 0;
    for (const auto& t : m.terms) nm += t->ckind == DCRF_COMPAT_MATRIX;
    return nm;
}

// Forward pass that keeps Q_0..Q_n and the Matrix terms' raw messages for the adjoint.
void forward_train(dcrf_model& m, int n_iter, const float* U, const float* compat_raw, const float* kw,
                   hipStream_t s) {
    const int N = m.N(), L = m.L, T = (int)m.terms.size();
    const size_t NL = (size_t)N * L;
    if (compat_raw) {
        size_t off = 0;
        for (auto& t : m.terms) {
            launch_set_compat(compat_raw + off, t->ckind, L, t->d_compat.get(), s);
            off += compat_raw_len(t->ckind, L);
        }
    }
    m.qcache.reserve((size_t)(n_iter + 1) * NL);
    // raw messages of the Matrix terms only (their statistic is a full L x L product
    // taken after the adjoint loop); the others' statistic comes from the adjoint kernels
    const int nm = matrix_terms(m);
    m.mcache.reserve(std::max<size_t>((size_t)n_iter * nm * NL, 1));
    launch_softmax_neg(U, m.qcache.get(), N, L, s);
    for (int it = 0; it < n_iter; it++) {
        const float* Qt = m.qcache.get() + (size_t)it * NL;
        float* Qn = m.qcache.get() + (size_t)(it + 1) * NL;
        run_filter(m, Qt, L, false, s);
        SliceTable st = slice_table(m, false, kw);
        for (int k = 0, j = 0; k < T; k++)
            if (m.terms[k]->ckind == DCRF_COMPAT_MATRIX)
                st.t[k].mout = m.mcache.get() + ((size_t)it * nm + j++) * NL;
        if (m.fused)  // same arithmetic as k_meanfield, gathers of all entries up front
            launch_meanfield_fused(U, Qn, L, N, L, st, m.g_slc.get(), s);
        else
            launch_meanfield(U, Qn, N, L, st, nullptr, nullptr, s);
    }
    m.train_n = n_iter;
    m.train_kw = kw;
}

// Adjoint of forward_train.  dQ: dL/dQ_n.  Outputs dU (always), per-term statistics in
// m.gacc, and optionally raw compatibility / term-weight gradients on the device.
// kernel: also accumulate the kernel-parameter gradient of every term into m.kacc.
// The interleaved slice pairs of the adjoint's fused kernel: the forward's when every
// term's transposed slice weights equal its forward ones (SYMMETRIC / NO_NORMALIZATION),
// else pairs with the transposed weights, built once per layout.
const int2* adjoint_pairs(dcrf_model& m, hipStream_t s) {
    if (m.all_symmetric) return m.g_slc.get();
    if (!m.slcT_valid) {
        const int T = (int)m.terms.size(), N = m.N();
        int ds = 0;
        for (int k = 0; k < T; k++) ds += m.terms[k]->lat.d + 1;
        m.g_slcT.reserve((size_t)N * ds);
        int off = 0;
        for (int k = 0; k < T; k++) {
            const int d1 = m.terms[k]->lat.d + 1;
            launch_interleave_slice(m.tl[k]->vtx.get(), m.tl[k]->swT.get(), N, d1, ds, off, m.g_slcT.get(), s);
            off += d1;
        }
        m.slcT_valid = true;
    }
    return m.g_slcT.get();
}

void backward(dcrf_model& m, const float* dQ, float* dU, float* dcompat, float* dkw, hipStream_t s,
              bool kernel = false) {
    if (m.train_n < 0) throw Error{DCRF_ERR_STATE, "backward needs a preceding training forward"};
    const int N = m.N(), L = m.L, T = (int)m.terms.size(), n = m.train_n;
    const size_t NL = (size_t)N * L;
    // b_t, the adjoint entering reverse iteration t - 1, in slot t (t = n .. 0).  The
    // Matrix terms' statistic reads every b_t after the loop, so with such a term every
    // slot stays; otherwise two slots alternate.
    const int nm = matrix_terms(m);
    m.bbuf.reserve((size_t)(nm ? n + 1 : std::min(n + 1, 2)) * NL);
    auto bslot = [&](int t) { return m.bbuf.get() + (size_t)(nm ? t : (t & 1)) * NL; };
    const float* Qc = m.qcache.get();
    launch_sn_first(Qc + (size_t)n * NL, dQ, bslot(n), dU, N, L, s);
    std::vector<size_t> goff(T + 1, 0);
    for (int k = 0; k < T; k++) goff[k + 1] = goff[k] + stat_len(m.terms[k]->ckind, L);
    m.gacc.reserve(std::max<size_t>(goff[T], 1));
    launch_fill_double(m.gacc.get(), goff[T], s);
    // the other terms' statistic rows: term j (in term order among them), iteration it
    // at partial + (j * n + it) * nblk * L, one row of L sums per adjoint block
    const int nd = T - nm;
    const size_t nblk = (size_t)adjoint_blocks(N, L);
    size_t pf = std::max<size_t>((size_t)nd * n * nblk * L, 1);
    if (nm) pf = std::max(pf, grad_partial_floats(N, L, n));
    m.partial.reserve(pf);
    m.rmid.reserve(rows_mid_doubles(nm ? L * L : L));
    std::vector<int> koff(T + 1, 0);
    for (int k = 0; k < T; k++) koff[k + 1] = koff[k] + m.terms[k]->num_kparams();
    kernel = kernel && koff[T] > 0;
    size_t pstride = 0;  // per term: its partial rows in kpartial
    if (kernel) {
        ensure_kg_plans(m);
        int dmax = 1, dslow = 0;  // dslow: largest d of a term without the one-pass gradient
        for (auto& t : m.terms) {
            dmax = std::max(dmax, t->lat.d);
            if (t->num_kparams() && !kg_grad_fused(L, t->lat.d)) dslow = std::max(dslow, t->lat.d);
        }
        size_t tmax = 1;
        for (const auto& P : m.kgplans) tmax = std::max(tmax, P->parts.size());
        m.kgbuf.reserve(4 * tmax * NL);
        if (dslow) m.kgdf.reserve((size_t)2 * N * dslow);
        pstride = kparam_partial_doubles(dmax);
        m.kpartial.reserve(pstride * T);
        m.kacc.reserve(koff[T]);
        launch_fill_double(m.kacc.get(), koff[T], s);
        for (int k = 0; k < T; k++)  // the one-pass kernel adds every iteration to its rows
            if (m.terms[k]->num_kparams() && kg_grad_fused(L, m.terms[k]->lat.d))
                launch_fill_double(m.kpartial.get() + (size_t)k * pstride,
                                   (size_t)kg_grad_blocks(N, L) * m.terms[k]->num_kparams(), s);
    }
    for (int it = n - 1; it >= 0; it--) {
        float* b = bslot(it + 1);
        if (kernel) {
            // DenseCRF::gradient takes the kernel gradient with b before its update
            // (densecrf.cpp:280-284)
            for (const auto& P : m.kgplans)
                kernel_grad_group(m, *P, b, Qc + (size_t)it * NL, m.train_kw, koff, pstride, s);
        }
        run_filter(m, b, L, true, s);
        SliceTable st = slice_table(m, true, m.train_kw);
        for (int k = 0, j = 0; k < T; k++)
            if (m.terms[k]->ckind != DCRF_COMPAT_MATRIX) {
                st.t[k].gstat = m.partial.get() + ((size_t)j * n + it) * nblk * L;
                st.t[k].gslot = j++;
                st.nstat = j;
            }
        if (m.fused)
            launch_adjoint_fused(Qc + (size_t)it * NL, bslot(it), dU, N, L, st, adjoint_pairs(m, s), s);
        else
            launch_adjoint(Qc + (size_t)it * NL, bslot(it), dU, N, L, st, s);
    }
    if (kernel)  // the one-pass kernel gradient's partial rows hold all iterations
        for (int k = 0; k < T; k++)
            if (m.terms[k]->num_kparams() && kg_grad_fused(L, m.terms[k]->lat.d))
                launch_kparam_final(m.kpartial.get() + (size_t)k * pstride, kg_grad_blocks(N, L),
                                    m.terms[k]->num_kparams(), m.kacc.get() + koff[k], s);
    // G_k = sum_t b_{t+1} (K~_k Q_t)^T over all iterations: the statistic rows of the
    // Potts / Diagonal terms first (they share the partial buffer with the Matrix
    // terms' products), then for Matrix terms b_{t+1} in slot t + 1 against the messages
    // of iteration t
    for (int k = 0, j = 0; k < T; k++)
        if (m.terms[k]->ckind != DCRF_COMPAT_MATRIX) {
            launch_sum_rows(m.partial.get() + (size_t)j * n * nblk * L, (long long)n * nblk, L, m.rmid.get(),
                            m.gacc.get() + goff[k], s);
            j++;
        }
    for (int k = 0, j = 0; k < T; k++)
        if (m.terms[k]->ckind == DCRF_COMPAT_MATRIX) {
            launch_grad_stat(bslot(1), NL, m.mcache.get() + (size_t)j * NL, (size_t)nm * NL, n, N, L,
                             m.partial.get(), m.rmid.get(), m.gacc.get() + goff[k], s);
            j++;
        }
    if (dcompat || dkw) {
        size_t coff = 0;
        for (int k = 0; k < T; k++) {
            const Term& t = *m.terms[k];
            launch_compat_finish(m.gacc.get() + goff[k], t.ckind, L, t.d_compat.get(),
                                 m.train_kw ? m.train_kw + k : nullptr, dcompat ? dcompat + coff : nullptr,
                                 dkw ? dkw + k : nullptr, s);
            coff += compat_raw_len(t.ckind, L);
        }
    }
}

// Hamming's class weights (objective.cpp:51-63): the given ones, or from the class
// counts c of gt as (c / sum c)^-pow normalised by sum c w.  Checks that every label of
// gt inside [0, L) has a weight.
std::vector<float> hamming_weights(const int16_t* gt, int N, int L, float pow_, const float* cw_in, int ncw) {
    std::vector<float> w;
    if (cw_in) {
        w.assign(cw_in, cw_in + ncw);
    } else {
        int M = 0;
        for (int i = 0; i < N; i++) M = std::max(M, gt[i] + 1);
        std::vector<float> cnt(M, 0.f);
        for (int i = 0; i < N; i++)
            if (gt[i] >= 0) cnt[gt[i]] += 1.0f;
        float total = 0.f;
        for (float c : cnt) total += c;
        w.resize(M);
        float z = 0.f;
        for (int l = 0; l < M; l++) {
            w[l] = std::pow(cnt[l] / total, -pow_);
            z += cnt[l] * w[l];
        }
        for (int l = 0; l < M; l++) w[l] = w[l] / z;
    }
    for (int i = 0; i < N; i++) {
        const int g = gt[i];
        if (g >= 0 && g < L) require(g < (int)w.size(), "a ground-truth label has no class weight");
    }
    if (w.empty()) w.assign(1, 0.f);  // no labelled pixel: the kernel reads no weight
    return w;
}

// The objective's value from launch_objective's partial sums (OP: its partial array):
// LogLikelihood / Hamming add the block sums in order; IoU is sum_l in_l / un_l / L.
double objective_value(int kind, const std::vector<double>& OP, int L) {
    double r = 0.0;
    if (kind == DCRF_OBJECTIVE_IOU) {
        const double* st = OP.data() + OP.size() - 2 * (size_t)L;
        for (int l = 0; l < L; l++) r += st[2 * l] / st[2 * l + 1];
        return r / L;
    }
    for (double v : OP) r += v;
    return r;
}

// ---- allocation-free inference (dcrf_inference_device_noalloc)
// The filter's vertex rows (splat output, partial rows, both blur buffers) in one static
// device array per GPU instead of the model's buffers.  Host code must not use the
// array's name as an address: the address comes from hipGetSymbolAddress, once per
// device, and is kept in scratch_dev.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];
float* scratch_dev[kMaxDevices] = {};
std::mutex scratch_mu;

float* scratch_for_current_device() {
    int dev = 0;
    DCRF_HIP_CHECK(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, "device index out of range for the allocation-free inference");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        DCRF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "the scratch array has no device address");
        scratch_dev[dev] = static_cast<float*>(p);
    }
    return scratch_dev[dev];
}

// Points the model's two vertex-row buffers at other memory for its lifetime (nothing
// may reserve or release them meanwhile), then gives the model its own back.
struct BorrowedRows {
    dcrf_model& m;
    float* own_p[2];
    size_t own_n[2];
    BorrowedRows(dcrf_model& m_, float* r0, size_t n0, float* r1, size_t n1) : m(m_) {
        for (int i = 0; i < 2; i++) {
            own_p[i] = m.vals[i].p;
            own_n[i] = m.vals[i].n;
        }
        m.vals[0].p = r0;
        m.vals[0].n = n0;
        m.vals[1].p = r1;
        m.vals[1].n = n1;
    }
    ~BorrowedRows() {
        for (int i = 0; i < 2; i++) {
            m.vals[i].p = own_p[i];
            m.vals[i].n = own_n[i];
        }
    }
    BorrowedRows(const BorrowedRows&) = delete;
    BorrowedRows& operator=(const BorrowedRows&) = delete;
};

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

int dcrf_release_cached_memory(size_t* bytes) {
    return guard([&] {
        const size_t b = pool_release_all() + release_host_stage();
        if (bytes) *bytes = b;
    });
}

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
        m->unary_host.resize(n);
        host_copy(m->unary_host.data(), U, sizeof(float) * n);
        m->unary_unset.clear();
        m->has_unary = true;
        m->unary_dirty = true;
        m->logistic = false;
    });
}

int dcrf_set_logistic_unary(dcrf_model* m, const float* Lmat, int F, const float* feat) {
    return guard([&] {
        require(m && Lmat && feat, "null argument");
        require(F >= 1, "the logistic unary needs at least one feature");
        const int N = m->N(), L = m->L;
        require((long long)N * F < (1ll << 31), "logistic unary features too large for 32-bit indexing");
        ensure_device(*m);
        hipStream_t s = m->stream;
        m->lL.assign(Lmat, Lmat + (size_t)L * F);
        // features (F, N) row-major, kept as given and as [N][F]
        m->lfFN.reserve((size_t)N * F);
        m->lf.reserve((size_t)N * F);
        DCRF_HIP_CHECK(hipMemcpyAsync(m->lfFN.get(), feat, sizeof(float) * N * F, hipMemcpyHostToDevice, s));
        launch_transpose(m->lfFN.get(), m->lf.get(), F, N, s);
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        m->lF = F;
        m->logistic = true;
        m->has_unary = true;
        m->unary_dirty = true;
        m->ugrad_param.clear();
    });
}

int dcrf_num_unary_params(const dcrf_model* m) { return m && m->logistic ? m->L * m->lF : 0; }

int dcrf_get_unary_params(const dcrf_model* m, float* params) {
    return guard([&] {
        require(m && params, "null argument");
        // LogisticUnaryEnergy::parameters (unary.cpp:53-57): L packed column-major
        const int L = m->L, F = m->lF;
        if (!m->logistic) return;
        for (int c = 0; c < F; c++)
            for (int r = 0; r < L; r++) params[(size_t)c * L + r] = m->lL[(size_t)r * F + c];
    });
}

int dcrf_set_unary_params(dcrf_model* m, const float* params) {
    return guard([&] {
        require(m && params, "null argument");
        require(m->logistic, "the unary has no parameters (only a logistic unary does)");
        const int L = m->L, F = m->lF;
        for (int c = 0; c < F; c++)
            for (int r = 0; r < L; r++) m->lL[(size_t)r * F + c] = params[(size_t)c * L + r];
        m->unary_dirty = true;
        m->train_n = -1;
    });
}

int dcrf_unary_param_gradient(const dcrf_model* m, float* grad) {
    return guard([&] {
        require(m && grad, "null argument");
        require(m->logistic, "the unary has no parameters (only a logistic unary does)");
        require(!m->ugrad_param.empty(), "no dcrf_gradient call since the logistic unary was set");
        std::copy(m->ugrad_param.begin(), m->ugrad_param.end(), grad);
    });
}

int dcrf_set_unary_instance(dcrf_model* m, int i, const float* U) {
    return guard([&] {
        require(m && U, "null argument");
        require(!m->logistic, "the model has a logistic unary: set a full unary with dcrf_set_unary first");
        require(i >= 0 && i < m->geo.n_inst, "instance index out of range");
        const int N = m->N(), L = m->L, b = m->geo.pix_base[i], Ni = m->geo.N[i];
        if (m->unary_host.size() != (size_t)N * L) {  // instances not set are zero at the upload
            m->unary_host.resize((size_t)N * L);
            m->unary_unset.assign(m->geo.n_inst, 1);
        }
        copy_rows(m->unary_host.data() + b, N, U, Ni, L, Ni);
        if (!m->unary_unset.empty()) m->unary_unset[i] = 0;
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
        copy_to_host(Q, m->Q.get(), sizeof(float) * (size_t)m->N() * m->L, m->stream);
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_map(dcrf_model* m, int n_iter, int16_t* labels) {
    return guard([&] {
        require(m && labels, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        prepare(*m);
        const float* U = m->has_unary ? m->unary.get() : nullptr;
        run_inference(*m, n_iter, U, m->Q.get(), m->stream);
        m->labels.reserve(m->N());
        launch_argmax(m->Q.get(), m->L, m->N(), m->L, m->labels.get(), m->stream);
        copy_to_host(labels, m->labels.get(), sizeof(int16_t) * m->N(), m->stream);
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_current_map(dcrf_model* m, const float* Q, int16_t* labels) {
    return guard([&] {
        require(m && Q && labels, "null argument");
        ensure_device(*m);
        const size_t n = (size_t)m->N() * m->L;
        m->work.reserve(n);
        m->labels.reserve(m->N());
        copy_to_device(m->work.get(), Q, sizeof(float) * n, m->stream);
        launch_argmax(m->work.get(), m->L, m->N(), m->L, m->labels.get(), m->stream);
        copy_to_host(labels, m->labels.get(), sizeof(int16_t) * m->N(), m->stream);
    });
}

int dcrf_current_map_device(dcrf_model* m, const float* Q_dev, int16_t* labels_dev, void* stream) {
    return guard([&] {
        require(m && Q_dev && labels_dev, "null argument");
        ensure_device(*m);
        launch_argmax(Q_dev, m->L, m->N(), m->L, labels_dev, (hipStream_t)stream);
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

int dcrf_capture_inference(dcrf_model* m, int n_iter, const float* U_dev, float* Q_dev) {
    return guard([&] {
        require(m && Q_dev, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        require(!m->timing, "disable kernel timing before capturing a graph", DCRF_ERR_STATE);
        prepare(*m);
        m->drop_graph();
        const float* U = U_dev ? U_dev : (m->has_unary ? m->unary.get() : nullptr);
        DCRF_HIP_CHECK(hipStreamSynchronize(m->stream));
        DCRF_HIP_CHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        try {
            run_inference(*m, n_iter, U, Q_dev, m->stream);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(m->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        DCRF_HIP_CHECK(hipStreamEndCapture(m->stream, &m->graph));
        DCRF_HIP_CHECK(hipGraphInstantiate(&m->graph_exec, m->graph, nullptr, nullptr, 0));
        m->graph_own_unary = U_dev == nullptr;
        m->graph_unary = U_dev ? nullptr : U;
        m->graph_kw = m->kw_set ? m->kw.get() : nullptr;
        m->graph_valid = true;
    });
}

int dcrf_launch_graph(dcrf_model* m, void* stream) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(m->graph_valid && m->layout_valid, "no captured inference (or the model changed since)",
                DCRF_ERR_STATE);
        // The graph reads the unary and term-weight buffers it was captured with: new
        // values in the same buffers are picked up (a unary set since the capture is
        // uploaded first), a buffer that appeared, went or moved needs a new capture.
        if (m->graph_own_unary) {
            require(m->has_unary == (m->graph_unary != nullptr),
                    "the unary was set or removed since the capture: capture again", DCRF_ERR_STATE);
            if (m->has_unary) {
                ensure_unary(*m);
                require(m->unary.get() == m->graph_unary, "the unary buffer moved since the capture: capture again",
                        DCRF_ERR_STATE);
            }
        }
        require((m->kw_set ? m->kw.get() : nullptr) == m->graph_kw,
                "term weights were set or removed since the capture: capture again", DCRF_ERR_STATE);
        DCRF_HIP_CHECK(hipGraphLaunch(m->graph_exec, (hipStream_t)stream));
    });
}

int dcrf_inference_device_noalloc(dcrf_model* m, int n_iter, const float* U_dev, float* Q_dev, void* stream) {
    return guard([&] {
        require(m && Q_dev, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        require(m->dev_ready && m->layout_valid,
                "the model's layout is not built: run dcrf_inference_device (or dcrf_inference) on it once first",
                DCRF_ERR_STATE);
        require(!m->timing, "disable kernel timing for the allocation-free inference", DCRF_ERR_STATE);
        require(!m->tiled && !m->runs,
                "the allocation-free inference runs the default splat and mean-field kernels only", DCRF_ERR_STATE);
        require(U_dev || !m->has_unary || !m->unary_dirty,
                "the unary changed since the last call: upload it with dcrf_inference_device first, or pass U_dev",
                DCRF_ERR_STATE);
        const size_t n0 = (size_t)m->rows0 * m->ldv, n1 = (size_t)(m->V_total + 1) * m->ldv;
        const size_t off1 = (n0 + 63) / 64 * 64;  // second buffer 256-byte aligned
        const size_t need = off1 + n1;
        if (need > kNoAllocScratchBytes / sizeof(float))
            throw Error{DCRF_ERR_VALUE, "the model needs " + std::to_string(need * sizeof(float)) +
                                            " bytes of vertex rows; the allocation-free inference holds at most " +
                                            std::to_string(kNoAllocScratchBytes) + " (use dcrf_inference_device)"};
        float* base = scratch_for_current_device();
        hipStream_t s = (hipStream_t)stream;
        // the rows start at zero, as the layout leaves the model's own buffers (the
        // missing-neighbour row is read as zero); the array keeps the last call's rows
        launch_fill(base, 0.f, need, s);
        BorrowedRows rows(*m, base, n0, base + off1, n1);
        const float* U = U_dev ? U_dev : (m->has_unary ? m->unary.get() : nullptr);
        run_inference(*m, n_iter, U, Q_dev, s, false);
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
        copy_to_host(Q, m->Q.get(), sizeof(float) * n, m->stream);
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
        copy_to_device(m->Q.get(), Q, sizeof(float) * n, s);
        if (m->terms.empty()) launch_fill(m->tmp2.get(), 0.f, n, s);
        run_filter(*m, m->Q.get(), m->L, false, s);
        launch_meanfield(m->has_unary ? m->unary.get() : nullptr, m->Q.get(), m->N(), m->L,
                         slice_table(*m, false, m->kw_set ? m->kw.get() : nullptr), m->tmp1.get(), m->terms.empty() ? nullptr : m->tmp2.get(), s);
        copy_to_host(Q, m->Q.get(), sizeof(float) * n, s);
        if (tmp1) copy_to_host(tmp1, m->tmp1.get(), sizeof(float) * n, s);
        if (tmp2) copy_to_host(tmp2, m->tmp2.get(), sizeof(float) * n, s);
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
        copy_to_device(m->work.get(), Q, sizeof(float) * n, s);
        run_filter(*m, m->work.get(), m->L, false, s);
        const LaunchShape sh = shape_for(m->L);
        const int nb = (m->N() + sh.items - 1) / sh.items;
        DevBuf<double> part;
        part.reserve(nb);
        launch_kl(m->work.get(), m->has_unary ? m->unary.get() : nullptr, m->N(), m->L,
                  slice_table(*m, false, m->kw_set ? m->kw.get() : nullptr), part.get(), s);
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
        copy_to_device(m->work.get(), Q, sizeof(float) * n, s);
        run_filter(*m, m->work.get(), m->L, transpose != 0, s);
        launch_slice_term(m->tmp1.get(), m->N(), m->L,
                          slice_table(*m, transpose != 0, m->kw_set ? m->kw.get() : nullptr), term, s);
        copy_to_host(out, m->tmp1.get(), sizeof(float) * n, s);
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_unary_from_softmax_device(const float* sm, int L, int N, double scale, double clip, float* U, void* stream) {
    return guard([&] {
        require(sm && U, "null argument");
        require(L >= 1 && N >= 0, "bad shape");
        require(!(scale > 1.0), "scale needs to be in (0, 1]");
        launch_unary_from_softmax(sm, L, N, scale, scale > 0.0, clip, clip > 0.0, U, (hipStream_t)stream);
    });
}

int dcrf_unary_from_labels_device(const int32_t* labels, int N, int L, double gt_prob, int zero_unsure, float* U,
                                  void* stream) {
    return guard([&] {
        require(labels && U, "null argument");
        require(gt_prob > 0.0 && gt_prob < 1.0, "gt_prob must be in (0, 1)");
        require(L >= 2, "at least two labels");
        // the energies as numpy forms them: float64 logs stored to float32
        const double n_e = -std::log((1.0 - gt_prob) / (L - 1));
        const double p_e = -std::log(gt_prob);
        const double u_e = -std::log(1.0 / L);
        launch_unary_from_labels(labels, N, L, (float)p_e, (float)n_e, (float)u_e, zero_unsure != 0, U,
                                 (hipStream_t)stream);
    });
}

int dcrf_unary_energy(dcrf_model* m, const int16_t* labels, float* out) {
    return guard([&] {
        require(m && labels && out, "null argument");
        const int N = m->N(), L = m->L;
        // DenseCRF::unaryEnergy (densecrf.cpp:141-152): U(l_i, i), 0 for a label outside
        // [0, L) or without a unary
        std::fill(out, out + N, 0.0f);
        if (!m->has_unary) return;
        ensure_device(*m);
        ensure_unary(*m);
        std::vector<float> U((size_t)N * L);
        DCRF_HIP_CHECK(hipMemcpy(U.data(), m->unary.get(), sizeof(float) * U.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++)
            if (labels[i] >= 0 && labels[i] < L) out[i] = U[(size_t)i * L + labels[i]];
    });
}

int dcrf_pairwise_energy(dcrf_model* m, const int16_t* labels, int term, float* out) {
    return guard([&] {
        require(m && labels && out, "null argument");
        const int T = (int)m->terms.size();
        require(term == -1 || (term >= 0 && term < T), "pairwise term index out of range");
        const int N = m->N(), L = m->L;
        const size_t n = (size_t)N * L;
        std::fill(out, out + N, 0.0f);
        if (!T) return;
        // DenseCRF::pairwiseEnergy (densecrf.cpp:153-177): the one-hot labelling Q
        // through the term's message, -0.5 Q'(l_i, i); term -1 sums the terms in order
        prepare(*m);
        std::vector<float> Q(n, 0.0f), x(n);
        for (int i = 0; i < N; i++)
            if (labels[i] >= 0 && labels[i] < L) Q[(size_t)i * L + labels[i]] = 1.0f;
        m->work.reserve(n);
        m->tmp1.reserve(n);
        hipStream_t s = m->stream;
        copy_to_device(m->work.get(), Q.data(), sizeof(float) * n, s);
        run_filter(*m, m->work.get(), L, false, s);
        const SliceTable st = slice_table(*m, false, m->kw_set ? m->kw.get() : nullptr);
        for (int k = term < 0 ? 0 : term; k < (term < 0 ? T : term + 1); k++) {
            launch_slice_term(m->tmp1.get(), N, L, st, k, s);
            copy_to_host(x.data(), m->tmp1.get(), sizeof(float) * n, s);
            for (int i = 0; i < N; i++) {
                const float e = labels[i] >= 0 && labels[i] < L ? -0.5f * x[(size_t)i * L + labels[i]] : 0.0f;
                out[i] = out[i] + e;
            }
        }
    });
}

int dcrf_set_compat(dcrf_model* m, int term, const float* params) {
    return guard([&] {
        require(m && params, "null argument");
        require(term >= 0 && term < (int)m->terms.size(), "pairwise term index out of range");
        Term& t = *m->terms[term];
        const int L = m->L;
        if (t.ckind == DCRF_COMPAT_MATRIX) {
            for (int a = 0; a < L; a++)
                for (int b = 0; b < L; b++) {
                    const float sum = params[a * L + b] + params[b * L + a];
                    t.compat[(size_t)a * L + b] = 0.5f * sum;
                }
        } else {
            t.compat.assign(params, params + compat_raw_len(t.ckind, L));
        }
        ensure_device(*m);
        DCRF_HIP_CHECK(hipMemcpy(t.d_compat.get(), t.compat.data(), sizeof(float) * t.compat.size(),
                                 hipMemcpyHostToDevice));
    });
}

int dcrf_set_kernel_weights(dcrf_model* m, const float* w) {
    return guard([&] {
        require(m != nullptr, "null model");
        const int T = (int)m->terms.size();
        if (!w) {
            m->kw_set = false;
            return;
        }
        ensure_device(*m);
        m->kw.reserve(std::max(T, 1));
        DCRF_HIP_CHECK(hipMemcpy(m->kw.get(), w, sizeof(float) * T, hipMemcpyHostToDevice));
        m->kw_set = true;
    });
}

int dcrf_num_compat_params(const dcrf_model* m) {
    if (!m) return 0;
    int n = 0;
    for (const auto& t : m->terms) n += compat_packed_len(t->ckind, m->L);
    return n;
}

int dcrf_get_compat_params(const dcrf_model* m, float* p) {
    return guard([&] {
        require(m && p, "null argument");
        const int L = m->L;
        for (const auto& tp : m->terms) {
            std::vector<float> c(tp->compat.size());
            if (m->dev_ready)
                DCRF_HIP_CHECK(hipMemcpy(c.data(), tp->d_compat.get(), sizeof(float) * c.size(), hipMemcpyDeviceToHost));
            else
                c = tp->compat;
            if (tp->ckind == DCRF_COMPAT_MATRIX) {
                for (int i = 0; i < L; i++)
                    for (int j = i; j < L; j++) *p++ = c[(size_t)i * L + j];
            } else {
                for (float v : c) *p++ = v;
            }
        }
    });
}

int dcrf_set_compat_params(dcrf_model* m, const float* p) {
    return guard([&] {
        require(m && p, "null argument");
        ensure_device(*m);
        const int L = m->L;
        // setLabelCompatibilityParameters (densecrf.cpp:329-341): Potts w, Diagonal w[L],
        // Matrix upper triangle w(i, j>=i) written to both w(i,j) and w(j,i)
        // (labelcompatibility.cpp:95-100)
        for (auto& tp : m->terms) {
            Term& t = *tp;
            if (t.ckind == DCRF_COMPAT_MATRIX) {
                for (int i = 0; i < L; i++)
                    for (int j = i; j < L; j++) {
                        const float v = *p++;
                        t.compat[(size_t)i * L + j] = v;
                        t.compat[(size_t)j * L + i] = v;
                    }
            } else {
                for (float& v : t.compat) v = *p++;
            }
            DCRF_HIP_CHECK(hipMemcpy(t.d_compat.get(), t.compat.data(), sizeof(float) * t.compat.size(),
                                     hipMemcpyHostToDevice));
        }
    });
}

int dcrf_num_kernel_params(const dcrf_model* m) { return m ? num_kernel_params(*m) : 0; }

int dcrf_get_kernel_params(const dcrf_model* m, float* params) {
    return guard([&] {
        require(m && params, "null argument");
        for (const auto& t : m->terms) {
            std::copy(t->kparams.begin(), t->kparams.begin() + t->num_kparams(), params);
            params += t->num_kparams();
        }
    });
}

int dcrf_set_kernel_params(dcrf_model* m, const float* params) {
    return guard([&] {
        require(m && params, "null argument");
        const int np = num_kernel_params(*m);
        for (int i = 0; i < np; i++) require(std::isfinite(params[i]), "kernel parameters must be finite");
        ensure_device(*m);
        hipStream_t s = m->stream;
        // new lattices and normalisers of every scaled term first, then commit them all
        struct Rebuilt {
            Term* t;
            Term tmp;
            std::vector<float> kp;
        };
        std::vector<std::unique_ptr<Rebuilt>> done;
        const float* p = params;
        const int npad = m->geo.npad_total();
        for (auto& tp : m->terms) {
            Term& t = *tp;
            const int n = t.num_kparams();
            if (!n) continue;
            auto r = std::make_unique<Rebuilt>();
            r->t = &t;
            r->kp.assign(p, p + n);
            p += n;
            const int d = t.lat.d;
            DevBuf<float> P, feat;
            upload(P, r->kp, s);
            feat.reserve((size_t)npad * d);
            launch_scale_features(t.feat0.get(), npad, d, t.ktype == DCRF_FULL_KERNEL, P.get(), feat.get(), s);
            r->tmp.ntype = t.ntype;
            build_lattice(m->geo, m->geod, m->tiles, feat.get(), d, r->tmp.lat, m->hash_load, s);
            if (r->tmp.lat.overflow)
                throw Error{DCRF_ERR_OVERFLOW, "a scaled lattice coordinate exceeds the int16 range of the lattice "
                                               "keys; scale the kernel parameters down"};
            build_chunks(r->tmp.lat, m->chunk());
            compute_norm(*m, r->tmp);
            done.push_back(std::move(r));
        }
        for (auto& r : done) {
            r->t->lat = std::move(r->tmp.lat);
            r->t->norm = std::move(r->tmp.norm);
            r->t->kparams = std::move(r->kp);
        }
        if (!done.empty()) {
            m->layout_valid = false;
            m->lat_gen++;
            m->train_n = -1;
            m->drop_graph();
        }
    });
}

int dcrf_gradient(dcrf_model* m, int n_iter, const int16_t* gt, float robust, double* objective, float* unary_grad,
                  float* compat_grad, float* kweight_grad, float* kernel_grad) {
    return dcrf_gradient_objective(m, n_iter, DCRF_OBJECTIVE_LOG_LIKELIHOOD, gt, robust, nullptr, 0, objective,
                                   unary_grad, compat_grad, kweight_grad, kernel_grad);
}

int dcrf_gradient_objective(dcrf_model* m, int n_iter, int objective_kind, const int16_t* gt, float param,
                            const float* class_weight, int n_class_weight, double* objective, float* unary_grad,
                            float* compat_grad, float* kweight_grad, float* kernel_grad) {
    return guard([&] {
        require(m && gt, "null argument");
        require(n_iter >= 0, "number of iterations must be non-negative");
        require(objective_kind >= DCRF_OBJECTIVE_LOG_LIKELIHOOD && objective_kind <= DCRF_OBJECTIVE_IOU,
                "unknown objective");
        require(!class_weight || n_class_weight >= 0, "negative class-weight count");
        prepare(*m);
        const int N = m->N(), L = m->L, T = (int)m->terms.size();
        const size_t NL = (size_t)N * L;
        hipStream_t s = m->stream;
        const float* U = m->has_unary ? m->unary.get() : nullptr;
        const float* kw = m->kw_set ? m->kw.get() : nullptr;
        // Hamming's class weights come from the host's ground truth (objective.cpp:51-63)
        std::vector<float> hw;
        if (objective_kind == DCRF_OBJECTIVE_HAMMING) hw = hamming_weights(gt, N, L, param, class_weight,
                                                                              n_class_weight);
        forward_train(*m, n_iter, U, nullptr, kw, s);
        m->work.reserve(NL);
        m->tmp1.reserve(NL);
        m->gt_dev.reserve(N);
        DCRF_HIP_CHECK(hipMemcpyAsync(m->gt_dev.get(), gt, sizeof(int16_t) * N, hipMemcpyHostToDevice, s));
        if (!hw.empty()) {
            m->obj_w.reserve(hw.size());
            DCRF_HIP_CHECK(hipMemcpyAsync(m->obj_w.get(), hw.data(), sizeof(float) * hw.size(),
                                          hipMemcpyHostToDevice, s));
        }
        const int npart = objective_partials(objective_kind, N, L);
        m->obj_partial.reserve(npart);
        launch_objective(objective_kind, m->qcache.get() + (size_t)n_iter * NL, m->gt_dev.get(), N, L, param,
                         hw.empty() ? nullptr : m->obj_w.get(), m->work.get(), m->obj_partial.get(), s);
        std::vector<double> OP(npart);
        DCRF_HIP_CHECK(hipMemcpyAsync(OP.data(), m->obj_partial.get(), sizeof(double) * npart,
                                      hipMemcpyDeviceToHost, s));
        const int nkp = num_kernel_params(*m);
        backward(*m, m->work.get(), m->tmp1.get(), nullptr, nullptr, s, kernel_grad && nkp > 0);
        std::vector<double> KG(nkp, 0.0);
        if (kernel_grad && nkp && n_iter > 0)
            DCRF_HIP_CHECK(hipMemcpyAsync(KG.data(), m->kacc.get(), sizeof(double) * nkp, hipMemcpyDeviceToHost, s));
        if (unary_grad)
            copy_to_host(unary_grad, m->tmp1.get(), sizeof(float) * NL, s);
        std::vector<double> UG;
        if (m->logistic) {
            // LogisticUnaryEnergy::gradient summed over the iterations (densecrf.cpp:261-262,
            // 293-294; unary.cpp:64-68): sum_t b_t f^T = (dL/dU) f^T, one gram product
            const int F = m->lF;
            m->partial.reserve(std::max(m->partial.n, gram_partial_floats(L, F, N)));
            m->gacc2.reserve((size_t)L * F);
            launch_gram(m->tmp1.get(), L, m->lf.get(), F, N, m->partial.get(), m->gacc2.get(), s);
            UG.resize((size_t)L * F);
            DCRF_HIP_CHECK(hipMemcpyAsync(UG.data(), m->gacc2.get(), sizeof(double) * L * F, hipMemcpyDeviceToHost, s));
        }
        size_t glen = 0;
        for (auto& t : m->terms) glen += stat_len(t->ckind, L);
        std::vector<double> G(glen);
        DCRF_HIP_CHECK(hipMemcpyAsync(G.data(), m->gacc.get(), sizeof(double) * glen, hipMemcpyDeviceToHost, s));
        std::vector<float> w(T, 1.0f);
        if (kw) DCRF_HIP_CHECK(hipMemcpyAsync(w.data(), kw, sizeof(float) * T, hipMemcpyDeviceToHost, s));
        // the compatibilities the forward used: the device copies (a training forward with
        // raw compatibilities, dcrf_forward_device, writes them there only)
        std::vector<std::vector<float>> cw(T);
        for (int k = 0; k < T; k++) {
            const Term& t = *m->terms[k];
            cw[k].resize(compat_raw_len(t.ckind, L));
            DCRF_HIP_CHECK(hipMemcpyAsync(cw[k].data(), t.d_compat.get(), sizeof(float) * cw[k].size(),
                                          hipMemcpyDeviceToHost, s));
        }
        DCRF_HIP_CHECK(hipStreamSynchronize(s));
        if (objective) *objective = objective_value(objective_kind, OP, L);
        // pack like DenseCRF::gradient (labelcompatibility.cpp:57-61, 76-78, 101-108)
        size_t go = 0;
        float* cp = compat_grad;
        for (int k = 0; k < T; k++) {
            const Term& t = *m->terms[k];
            const double* g = G.data() + go;
            double kg = 0.0;
            if (t.ckind == DCRF_COMPAT_POTTS) {
                double sd = 0.0;
                for (int a = 0; a < L; a++) sd += g[a];
                if (cp) *cp++ = (float)(-w[k] * sd);
                kg = -(double)cw[k][0] * sd;
            } else if (t.ckind == DCRF_COMPAT_DIAGONAL) {
                for (int a = 0; a < L; a++) {
                    if (cp) *cp++ = (float)(w[k] * g[a]);
                    kg += (double)cw[k][a] * g[a];
                }
            } else {
                for (int i = 0; i < L; i++)
                    for (int j = i; j < L; j++) {
                        const double v = i == j ? g[i * L + i] : g[i * L + j] + g[j * L + i];
                        if (cp) *cp++ = (float)(w[k] * v);
                    }
                for (int j = 0; j < L * L; j++) kg += (double)cw[k][j] * g[j];
            }
            if (kweight_grad) kweight_grad[k] = (float)kg;
            go += stat_len(t.ckind, L);
        }
        if (kernel_grad)
            for (int i = 0; i < nkp; i++) kernel_grad[i] = (float)KG[i];
        if (m->logistic) {
            const int F = m->lF;
            m->ugrad_param.resize((size_t)L * F);
            for (int c = 0; c < F; c++)
                for (int r = 0; r < L; r++) m->ugrad_param[(size_t)c * L + r] = (float)UG[(size_t)r * F + c];
        }
        if (m->timing) collect_timing(*m);
    });
}

int dcrf_forward_device(dcrf_model* m, int n_iter, const float* U_dev, const float* compat_dev, const float* kw_dev,
                        float* Q_dev, void* stream) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(n_iter >= 0, "number of iterations must be non-negative");
        prepare(*m);
        hipStream_t s = (hipStream_t)stream;
        const float* U = U_dev ? U_dev : (m->has_unary ? m->unary.get() : nullptr);
        const float* kw = kw_dev ? kw_dev : (m->kw_set ? m->kw.get() : nullptr);
        forward_train(*m, n_iter, U, compat_dev, kw, s);
        if (Q_dev) {
            const size_t NL = (size_t)m->N() * m->L;
            DCRF_HIP_CHECK(hipMemcpyAsync(Q_dev, m->qcache.get() + (size_t)n_iter * NL, sizeof(float) * NL,
                                          hipMemcpyDeviceToDevice, s));
        }
    });
}

int dcrf_backward_device(dcrf_model* m, const float* dQ_dev, float* dU_dev, float* dcompat_dev, float* dkw_dev,
                         float* dkernel_dev, void* stream) {
    return guard([&] {
        require(m && dQ_dev && dU_dev, "null argument");
        hipStream_t s = (hipStream_t)stream;
        const int nkp = num_kernel_params(*m);
        const bool kernel = dkernel_dev && nkp > 0;
        backward(*m, dQ_dev, dU_dev, dcompat_dev, dkw_dev, s, kernel);
        if (kernel) {
            if (m->train_n > 0)
                launch_kparam_store(m->kacc.get(), nkp, dkernel_dev, s);
            else
                launch_fill(dkernel_dev, 0.f, nkp, s);
        }
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

int dcrf_set_splat_tile(dcrf_model* m, int side) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(side >= 0 && side <= 4096, "tile side must be in [0, 4096]");
        require(m->terms.empty(), "set the splat tile before adding pairwise terms", DCRF_ERR_STATE);
        m->tile_side = side;
        m->tiles_ready = false;
    });
}

int dcrf_set_meanfield_kernel(dcrf_model* m, int mode, int run_rows) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(mode >= 0 && mode <= 3, "mean-field kernel mode must be 0 (auto), 1, 2 or 3");
        require(run_rows >= 0 && run_rows <= 64, "run rows must be in [0, 64]");
        const int rr = run_rows ? run_rows : 16;
        if (mode != m->mf_mode || rr != m->run_rows) {
            m->mf_mode = mode;
            m->run_rows = rr;
            m->layout_valid = false;
        }
    });
}

int dcrf_set_blur_lanes(dcrf_model* m, int labels_per_lane, int below_vertices) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(labels_per_lane >= 0 && labels_per_lane <= 4, "labels per lane must be 0 (default) or 1 to 4");
        require(labels_per_lane == 0 || m->L % labels_per_lane == 0, "labels per lane must divide the label count");
        require(labels_per_lane == 0 || shape_for_r(m->L, labels_per_lane).wave,
                "too few labels per lane for a lane group within one wave");
        m->blur_r = labels_per_lane;
        m->blur_small_v = below_vertices;
        m->drop_graph();
    });
}

int dcrf_set_q_blocks(dcrf_model* m, int on) {
    return guard([&] {
        require(m != nullptr, "null model");
        if (m->qblocks != (on != 0)) {
            m->qblocks = on != 0;
            m->qblocks_ready = false;
            m->drop_graph();
        }
    });
}

int dcrf_set_blur_steps(dcrf_model* m, int steps) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(steps == 2 || steps == 3, "blur steps per pass: 2 or 3");
        if (m->blur_steps != steps) {
            m->blur_steps = steps;
            m->layout_valid = false;
            m->train_n = -1;
        }
    });
}

int dcrf_set_blur_pairs(dcrf_model* m, int on) {
    return guard([&] {
        require(m != nullptr, "null model");
        if (m->pairs != (on != 0)) {
            m->pairs = on != 0;
            m->layout_valid = false;
            m->train_n = -1;
        }
    });
}

int dcrf_set_hash_load(dcrf_model* m, float load) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(load > 0.0f && load <= 0.5f, "hash load factor must be in (0, 0.5]");
        m->hash_load = load;
    });
}

int dcrf_enable_timing(dcrf_model* m, int on) {
    return guard([&] {
        require(m != nullptr, "null model");
        if (m->timing && !on) collect_timing(*m);
        m->timing = on != 0;
        for (int c = 0; c < KC_COUNT; c++) {
            m->kc_launches[c] = 0;
            m->kc_ms[c] = 0.0;
            m->kc_bytes[c] = 0.0;
        }
    });
}

int dcrf_kernel_stats(dcrf_model* m, int kernel, long long* launches, double* ms, double* bytes) {
    return guard([&] {
        require(m != nullptr, "null model");
        require(kernel >= 0 && kernel < KC_COUNT, "unknown kernel class");
        collect_timing(*m);
        if (launches) *launches = m->kc_launches[kernel];
        if (ms) *ms = m->kc_ms[kernel];
        if (bytes) *bytes = m->kc_bytes[kernel];
    });
}

int dcrf_splat_layout(dcrf_model* m, int* tiled, int* tile_side, int* n_tiles, int* segments, int* partials,
                      int* meanfield) {
    return guard([&] {
        require(m && tiled && tile_side && n_tiles && segments && partials && meanfield, "null argument");
        ensure_device(*m);
        ensure_layout(*m);
        *tiled = m->tiled ? 1 : 0;
        *tile_side = m->tiles.side;
        *n_tiles = m->tiles.n_tiles();
        *segments = m->C_total;
        *partials = m->rows0 - m->V_total - 1;
        *meanfield = m->runs ? 2 : (m->fused ? 1 : 0);
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
