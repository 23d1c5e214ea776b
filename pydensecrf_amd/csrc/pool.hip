// Device memory for the library's buffers (DevBuf): a caching pool over hipMalloc.
//
// Building a model allocates and frees some hundred temporaries (hash tables, sort
// buffers, segment lists, staging), and every hipFree waits for the device and unmaps
// memory: at 1024x768 the two lattice builds and the layout spent ~19 of their ~21 ms in
// hipMalloc / hipFree.  Freed blocks are kept instead and handed out again:
//   - a freed block first waits in a pending list: the pool does not know which stream
//     last used it, so it becomes reusable only after a hipDeviceSynchronize that the
//     pool itself performs (when an allocation finds no free block but pending ones);
//   - sizes are rounded up (256 B below 1 MiB, 1 MiB above) and a request takes the
//     smallest free block of at least its size and at most twice it;
//   - above kCacheLimit cached bytes, and when hipMalloc runs out of memory, the cached
//     blocks of the device are returned to the runtime (after a device synchronisation);
//     dcrf_release_cached_memory does the same on request.
// Blocks are per device (the current device at allocation).  The pool itself is never
// destroyed: its blocks go back with the process.
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "dcrf_internal.h"

namespace dcrf {
namespace {

constexpr size_t kCacheLimit = (size_t)4 << 30;

struct Block {
    size_t size;
    int dev;
};

struct Pool {
    std::mutex mu;
    std::unordered_map<void*, Block> live;
    std::map<int, std::multimap<size_t, void*>> reusable;
    std::map<int, std::vector<std::pair<void*, size_t>>> pending;
    size_t cached = 0;
};

Pool& pool() {
    static Pool* p = new Pool;  // never destroyed (the runtime may be gone at exit)
    return *p;
}

size_t round_size(size_t b) {
    if (b == 0) b = 1;
    if (b <= ((size_t)1 << 20)) return (b + 255) & ~(size_t)255;
    return (b + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
}

// Move the device's pending blocks to the reusable ones; the caller holds the lock and
// has `dev` current.
void settle(Pool& P, int dev) {
    auto& pend = P.pending[dev];
    if (pend.empty()) return;
    (void)hipDeviceSynchronize();
    auto& fl = P.reusable[dev];
    for (auto& b : pend) fl.emplace(b.second, b.first);
    pend.clear();
}

// Return every cached block of `dev` to the runtime; the caller holds the lock and has
// `dev` current.
size_t release(Pool& P, int dev) {
    settle(P, dev);
    size_t bytes = 0;
    auto& fl = P.reusable[dev];
    for (auto& b : fl) {
        (void)hipFree(b.second);
        bytes += b.first;
    }
    fl.clear();
    P.cached -= bytes;
    return bytes;
}

void* take(Pool& P, int dev, size_t s) {
    auto& fl = P.reusable[dev];
    auto it = fl.lower_bound(s);
    if (it == fl.end() || it->first > 2 * s) return nullptr;
    void* p = it->second;
    const size_t size = it->first;
    fl.erase(it);
    P.cached -= size;
    P.live[p] = Block{size, dev};
    return p;
}

}  // namespace

void* pool_alloc(size_t bytes) {
    Pool& P = pool();
    int dev = 0;
    DCRF_HIP_CHECK(hipGetDevice(&dev));
    const size_t s = round_size(bytes);
    std::lock_guard<std::mutex> lk(P.mu);
    if (void* p = take(P, dev, s)) return p;
    if (!P.pending[dev].empty()) {
        settle(P, dev);
        if (void* p = take(P, dev, s)) return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, s);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        release(P, dev);
        e = hipMalloc(&p, s);
    }
    if (e != hipSuccess)
        throw Error{e == hipErrorOutOfMemory ? DCRF_ERR_MEMORY : DCRF_ERR_DEVICE,
                    std::string("hipMalloc: ") + hipGetErrorString(e)};
    P.live[p] = Block{s, dev};
    return p;
}

void pool_free(void* p) {
    if (!p) return;
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return;  // not ours
    const Block b = it->second;
    P.live.erase(it);
    P.pending[b.dev].emplace_back(p, b.size);
    P.cached += b.size;
    if (P.cached > kCacheLimit) {
        int cur = 0;
        if (hipGetDevice(&cur) == hipSuccess && hipSetDevice(b.dev) == hipSuccess) {
            release(P, b.dev);
            (void)hipSetDevice(cur);
        }
    }
}

size_t pool_release_all() {
    Pool& P = pool();
    std::lock_guard<std::mutex> lk(P.mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return 0;
    size_t bytes = 0;
    std::vector<int> devs;
    for (auto& kv : P.reusable) devs.push_back(kv.first);
    for (auto& kv : P.pending)
        if (!kv.second.empty()) devs.push_back(kv.first);
    for (int d : devs)
        if (hipSetDevice(d) == hipSuccess) bytes += release(P, d);
    (void)hipSetDevice(cur);
    return bytes;
}

}  // namespace dcrf
