// host.cpp -- the C ABI (include/genoaligner/genoaligner_c.h) over the gfx950
// kernels: request validation, batch packing, device buffers on a per-call
// stream, launches, and result decoding with the library-side CIGAR checks.
//
// Batch and result rules restate the reference API (src/api/api.cpp):
//   WFA   validation :245-266, min smax :271-272, any_cigar :273-274 and the
//         mixed-batch rule that every resolved pair then gets a CIGAR :456-469,
//         decode + validate :71-106
//   SW    validation :600-616 with the scheme rules :515-536, buffer sizing
//         :636-647, result fill :542-580 and the score-only fill :829-838
// The error strings are the reference's, so callers that log or match them see
// the same text.
#include "ga_device.hpp"
#include "../../include/genoaligner/genoaligner_c.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using ga::kSeqPad;

// ---------------------------------------------------------------------------
// device plumbing
// ---------------------------------------------------------------------------
std::once_flag g_pool_once;

void init_pool()
{
    // Keep freed stream-ordered allocations cached in the device's default pool
    // instead of returning them to the driver at every synchronisation.
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
}

// A small pool of streams (default priority), reused across calls.
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams;

hipStream_t acquire_stream()
{
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (!g_streams.empty()) {
            hipStream_t s = g_streams.back();
            g_streams.pop_back();
            return s;
        }
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
    return s;
}

void release_stream(hipStream_t s)
{
    if (!s) return;
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_streams.push_back(s);
}

// ---------------------------------------------------------------------------
// Workspace blocks.  Large per-call workspaces (wavefront stacks and direction
// tables, up to tens of GB) come from a process-wide cache of device blocks,
// not from the stream-ordered pool: allocating and freeing blocks of that size
// through the pool took 0.3-2 s now and then (the runtime maps and unmaps
// their memory) against 15 ms of kernels per call.  A block handed back
// carries an event recorded on the returning stream after its last use, and
// the next user's stream waits on that event before touching the block.
// Small requests, and every request on a capturing stream, stay stream-ordered
// allocations (a graph records those as its own allocation nodes).  Idle
// blocks of a device are capped at 3/8 of its memory (the oldest go first);
// ga_trim_workspace() frees them all.
// ---------------------------------------------------------------------------
constexpr size_t kWsCacheMin = size_t(32) << 20;

struct WsBlock {
    void* p;
    size_t bytes;
    int dev;
    hipEvent_t done;
};
std::mutex g_ws_mu;
std::vector<WsBlock> g_ws_idle;   // handed back, oldest first
std::vector<WsBlock> g_ws_busy;   // handed out

// sizes rounded up to a granularity of 1/16..1/32 of the request, so requests
// that differ a little reuse one block
size_t ws_round(size_t b)
{
    size_t g = size_t(2) << 20;
    while (g * 16 < b) g <<= 1;
    return (b + g - 1) / g * g;
}

// Free idle blocks: those of `dev` (every device if dev < 0) until the idle
// bytes of that device are at most `keep`.
void ws_trim(int dev, size_t keep)
{
    std::vector<WsBlock> drop;
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        std::vector<size_t> idle_bytes;
        for (const WsBlock& b : g_ws_idle) {
            if (size_t(b.dev) >= idle_bytes.size()) idle_bytes.resize(size_t(b.dev) + 1, 0);
            idle_bytes[size_t(b.dev)] += b.bytes;
        }
        std::vector<WsBlock> kept;
        for (const WsBlock& b : g_ws_idle) {
            size_t& held = idle_bytes[size_t(b.dev)];
            if ((dev < 0 || b.dev == dev) && held > keep) {
                held -= b.bytes;
                drop.push_back(b);
            } else {
                kept.push_back(b);
            }
        }
        g_ws_idle.swap(kept);
    }
    if (drop.empty()) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (const WsBlock& b : drop) {
        if (b.dev != cur) (void)hipSetDevice(b.dev);
        (void)hipEventSynchronize(b.done);
        (void)hipEventDestroy(b.done);
        (void)hipFree(b.p);
        if (b.dev != cur) (void)hipSetDevice(cur);
    }
}

size_t ws_idle_cap(int dev)
{
    static std::atomic<size_t> cap[64];
    if (dev < 0 || dev >= 64) return 0;
    size_t c = cap[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        size_t free_b = 0, total = 0;
        if (hipMemGetInfo(&free_b, &total) != hipSuccess) return 0;
        c = total / 8 * 3;
        cap[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

// Drop-in for hipMallocAsync(p, bytes, st) on the workspace paths.
hipError_t ws_alloc(void** p, size_t bytes, hipStream_t st)
{
    *p = nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (bytes < kWsCacheMin || hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return hipMallocAsync(p, bytes, st);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        // best fit among idle blocks of this device, at most twice the request
        std::lock_guard<std::mutex> lk(g_ws_mu);
        size_t best = g_ws_idle.size();
        for (size_t i = 0; i < g_ws_idle.size(); ++i) {
            const WsBlock& b = g_ws_idle[i];
            if (b.dev == dev && b.bytes >= bytes && b.bytes / 2 <= bytes &&
                (best == g_ws_idle.size() || b.bytes < g_ws_idle[best].bytes))
                best = i;
        }
        if (best < g_ws_idle.size()) {
            const WsBlock b = g_ws_idle[best];
            e = hipStreamWaitEvent(st, b.done, 0);
            if (e != hipSuccess) return e;
            g_ws_idle.erase(g_ws_idle.begin() + std::ptrdiff_t(best));
            g_ws_busy.push_back(b);
            *p = b.p;
            return hipSuccess;
        }
    }
    WsBlock b{nullptr, ws_round(bytes), dev, nullptr};
    e = hipMalloc(&b.p, b.bytes);
    if (e != hipSuccess) {
        // out of memory: give the idle blocks of this device back and retry once
        (void)hipGetLastError();
        ws_trim(dev, 0);
        e = hipMalloc(&b.p, b.bytes);
        if (e != hipSuccess) { (void)hipGetLastError(); return e; }
    }
    e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
    if (e != hipSuccess) { (void)hipFree(b.p); return e; }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws_busy.push_back(b);
    *p = b.p;
    return hipSuccess;
}

// Drop-in for hipFreeAsync(p, st): a cached block goes back to the idle list
// once an event marks the end of its use on `st`.
hipError_t ws_free(void* p, hipStream_t st)
{
    if (!p) return hipSuccess;
    WsBlock b{};
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        auto it = std::find_if(g_ws_busy.begin(), g_ws_busy.end(), [p](const WsBlock& x) { return x.p == p; });
        if (it == g_ws_busy.end()) return hipFreeAsync(p, st);
        b = *it;
        g_ws_busy.erase(it);
    }
    hipError_t e = hipEventRecord(b.done, st);
    if (e != hipSuccess) {
        // no event to order the next use by: wait for the stream here
        (void)hipStreamSynchronize(st);
        (void)hipEventDestroy(b.done);
        if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipFree(b.p);
            return e;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        g_ws_idle.push_back(b);
    }
    ws_trim(b.dev, ws_idle_cap(b.dev));
    return e;
}

// Stream-ordered device allocations freed together at the end of a call.
struct DevArena {
    hipStream_t st = nullptr;
    std::vector<void*> ptrs;
    bool ok = true;
    template <class T> T* alloc(size_t count)
    {
        void* p = nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 64);
        if (ws_alloc(&p, bytes, st) != hipSuccess) { ok = false; return nullptr; }
        ptrs.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T> T* upload(const std::vector<T>& v)
    {
        T* d = alloc<T>(v.size());
        if (d && !v.empty() &&
            hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
            ok = false;
        return d;
    }
    // bytes from host memory (pinned, so the copy stays asynchronous)
    uint8_t* upload_raw(const uint8_t* src, size_t bytes)
    {
        uint8_t* d = alloc<uint8_t>(bytes);
        if (d && bytes > 0 && hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess) ok = false;
        return d;
    }
    void release()
    {
        for (void* p : ptrs) (void)ws_free(p, st);
        ptrs.clear();
    }
};

// Packed byte buffer: kSeqPad bytes before the first and after the last
// sequence, sequences back to back, int64 offsets.
struct Packed {
    std::vector<uint8_t> bytes;
    std::vector<int64_t> off;
    // All n sequences at once (get(k) -> {bytes, length}): offsets by a prefix
    // sum, then the copies on the host worker pool; codes (optional) receives
    // map[byte] for every byte of the buffer, padding included.
    template <class Get> void build(int64_t n, Get get, const uint8_t* map = nullptr,
                                    std::vector<uint8_t>* codes = nullptr);
    // The same layout into caller memory: layout() sets the offsets and
    // returns the buffer size, fill() writes the pads and the sequences (and,
    // with map, the codes into code_dst).
    template <class Get> int64_t layout(int64_t n, Get get);
    template <class Get> void fill(int64_t n, Get get, uint8_t* dst, int64_t size, const uint8_t* map = nullptr,
                                   uint8_t* code_dst = nullptr) const;
    // Sequences [k0, k1) of the same layout (and the pads at its ends), and
    // the byte range they fill: a piece of the buffer to upload on its own.
    template <class Get> void fill_range(int64_t k0, int64_t k1, int64_t n, Get get, uint8_t* dst,
                                         int64_t size) const;
    std::pair<int64_t, int64_t> piece(int64_t k0, int64_t k1, int64_t n, int64_t size) const
    {
        return {k0 == 0 ? 0 : off[size_t(k0)], k1 == n ? size : off[size_t(k1)]};
    }
};

// A process-wide pool of host worker threads (half the host threads, at most
// 8 with the caller), shared by all callers: jobs queue in order, woken
// workers take chunks of the front job, and a caller works on its own job
// until all chunks are taken, then sleeps until the chunks still running are
// done.  Half, and sleeping rather than spinning, because the batch calls run
// beside the serving loop's threads and the GPU runtime's: with a worker per
// CPU and spinning callers, a caller could wait a whole scheduler time slice
// for a preempted worker's chunk.  The pool is never torn down (its threads
// end with the process), so no static destructor joins threads at exit.
class WorkPool {
public:
    static WorkPool& get()
    {
        static WorkPool* pool = new WorkPool();
        return *pool;
    }
    int workers() const { return int(threads_.size()); }
    void run(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& fn)
    {
        auto job = std::make_shared<Job>();
        job->fn = &fn;
        job->n = n;
        job->grain = grain;
        {
            std::lock_guard<std::mutex> lk(mu_);
            queue_.push_back(job);
        }
        const int64_t chunks = (n + grain - 1) / grain;
        const int wake = int(std::min<int64_t>(chunks - 1, int64_t(threads_.size())));
        for (int k = 0; k < wake; ++k) cv_.notify_one();
        work(*job);
        {
            std::unique_lock<std::mutex> lk(job->m);
            job->cv.wait(lk, [&] { return job->done.load() >= n; });
        }
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = queue_.begin(); it != queue_.end(); ++it)
            if (*it == job) { queue_.erase(it); break; }
    }

private:
    struct Job {
        const std::function<void(int64_t, int64_t)>* fn = nullptr;
        int64_t n = 0, grain = 1;
        std::atomic<int64_t> next{0}, done{0};
        std::mutex m;
        std::condition_variable cv;
    };
    WorkPool()
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int nt = int(std::min<unsigned>(hw ? hw : 1, 16u)) / 2 - 1;
        for (int t = 0; t < nt; ++t) threads_.emplace_back([this] { loop(); });
    }
    static void work(Job& j)
    {
        for (;;) {
            const int64_t i0 = j.next.fetch_add(j.grain);
            if (i0 >= j.n) break;
            const int64_t i1 = std::min(j.n, i0 + j.grain);
            (*j.fn)(i0, i1);
            if (j.done.fetch_add(i1 - i0) + (i1 - i0) >= j.n) {
                std::lock_guard<std::mutex> lk(j.m);    // the last chunk wakes the caller
                j.cv.notify_all();
            }
        }
    }
    void loop()
    {
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return !queue_.empty(); });
                job = queue_.front();
                if (job->next.load() >= job->n) {   // every chunk taken: drop it, look again
                    queue_.pop_front();
                    continue;
                }
            }
            work(*job);
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<Job>> queue_;
    std::vector<std::thread> threads_;
};

// Run fn(i) for i in [0, n) on the pool, `grain` indices per chunk.
template <class F> void parallel_for(int64_t n, F&& fn, int64_t grain = 2048)
{
    if (n <= 0) return;
    if (n <= grain || WorkPool::get().workers() == 0) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    const std::function<void(int64_t, int64_t)> chunk = [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) fn(i);
    };
    WorkPool::get().run(n, grain, chunk);
}

template <class Get>
void Packed::build(int64_t n, Get get, const uint8_t* map, std::vector<uint8_t>* codes)
{
    off.resize(size_t(n));
    int64_t pos = kSeqPad;
    for (int64_t k = 0; k < n; ++k) {
        off[size_t(k)] = pos;
        const int len = get(k).second;
        pos += len > 0 ? len : 0;
    }
    bytes.assign(size_t(pos + kSeqPad), 0);
    if (codes) codes->assign(bytes.size(), map[0]);
    parallel_for(n, [&](int64_t k) {
        const auto sq = get(k);
        if (sq.second <= 0) return;
        uint8_t* dst = bytes.data() + off[size_t(k)];
        std::memcpy(dst, sq.first, size_t(sq.second));
        if (codes) {
            uint8_t* c = codes->data() + off[size_t(k)];
            for (int x = 0; x < sq.second; ++x) c[x] = map[dst[x]];
        }
    }, 16);
}

template <class Get>
int64_t Packed::layout(int64_t n, Get get)
{
    off.resize(size_t(n));
    int64_t pos = kSeqPad;
    for (int64_t k = 0; k < n; ++k) {
        off[size_t(k)] = pos;
        const int len = get(k).second;
        pos += len > 0 ? len : 0;
    }
    return pos + kSeqPad;
}

template <class Get>
void Packed::fill_range(int64_t k0, int64_t k1, int64_t n, Get get, uint8_t* dst, int64_t size) const
{
    if (k0 == 0) std::memset(dst, 0, size_t(kSeqPad));
    if (k1 == n) std::memset(dst + size - kSeqPad, 0, size_t(kSeqPad));
    parallel_for(k1 - k0, [&](int64_t q) {
        const auto sq = get(k0 + q);
        if (sq.second > 0) std::memcpy(dst + off[size_t(k0 + q)], sq.first, size_t(sq.second));
    }, 16);
}

template <class Get>
void Packed::fill(int64_t n, Get get, uint8_t* dst, int64_t size, const uint8_t* map, uint8_t* code_dst) const
{
    // sequences sit back to back: only the head and tail pads need clearing
    std::memset(dst, 0, size_t(kSeqPad));
    std::memset(dst + size - kSeqPad, 0, size_t(kSeqPad));
    if (code_dst) {
        std::memset(code_dst, map[0], size_t(kSeqPad));
        std::memset(code_dst + size - kSeqPad, map[0], size_t(kSeqPad));
    }
    parallel_for(n, [&](int64_t k) {
        const auto sq = get(k);
        if (sq.second <= 0) return;
        uint8_t* d = dst + off[size_t(k)];
        std::memcpy(d, sq.first, size_t(sq.second));
        if (code_dst) {
            uint8_t* c = code_dst + off[size_t(k)];
            for (int x = 0; x < sq.second; ++x) c[x] = map[d[x]];
        }
    }, 16);
}

// ---------------------------------------------------------------------------
// Pinned host staging for the host entry points' copies.  The runtime copies
// pageable memory through small bounce buffers, and a device-to-host copy
// into pageable memory also waits for the stream; from pinned memory
// (hipHostMalloc) copies run at DMA speed and stay asynchronous.  Blocks are
// cached process-wide (best fit, at most twice the request) and kept.
// ---------------------------------------------------------------------------
struct PinnedBlock {
    uint8_t* p = nullptr;
    size_t bytes = 0;
};
std::mutex g_pin_mu;
std::vector<PinnedBlock> g_pin_idle;

// Blocks held for one call; given back when it ends.  get() returns nullptr
// when pinned memory cannot be had (callers then copy from pageable memory).
struct PinnedHold {
    std::vector<PinnedBlock> held;
    PinnedHold() = default;
    PinnedHold(const PinnedHold&) = delete;              // the blocks go back once
    PinnedHold& operator=(const PinnedHold&) = delete;
    uint8_t* get(size_t bytes)
    {
        bytes = std::max<size_t>(bytes, 64);
        {
            std::lock_guard<std::mutex> lk(g_pin_mu);
            size_t best = g_pin_idle.size();
            for (size_t i = 0; i < g_pin_idle.size(); ++i) {
                const PinnedBlock& b = g_pin_idle[i];
                if (b.bytes >= bytes && b.bytes / 2 <= bytes &&
                    (best == g_pin_idle.size() || b.bytes < g_pin_idle[best].bytes))
                    best = i;
            }
            if (best < g_pin_idle.size()) {
                held.push_back(g_pin_idle[best]);
                g_pin_idle.erase(g_pin_idle.begin() + std::ptrdiff_t(best));
                return held.back().p;
            }
        }
        PinnedBlock b;
        b.bytes = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
        void* p = nullptr;
        if (hipHostMalloc(&p, b.bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        b.p = static_cast<uint8_t*>(p);
        held.push_back(b);
        return b.p;
    }
    ~PinnedHold()
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        for (const PinnedBlock& b : held) g_pin_idle.push_back(b);
    }
};

// ---------------------------------------------------------------------------
// optional kernel timing (per calling thread)
// ---------------------------------------------------------------------------
struct KernelTimer {
    bool enabled = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
    size_t used = 0;
    void reset() { used = 0; }
    void begin(hipStream_t st)
    {
        if (!enabled) return;
        if (used == pairs.size()) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
            pairs.emplace_back(a, b);
        }
        (void)hipEventRecord(pairs[used].first, st);
    }
    void end(hipStream_t st)
    {
        if (!enabled || used >= pairs.size()) return;
        (void)hipEventRecord(pairs[used].second, st);
        ++used;
    }
    float total_ms()
    {
        if (used == 0) return -1.0f;
        float sum = 0.0f;
        for (size_t i = 0; i < used; ++i) {
            if (hipEventSynchronize(pairs[i].second) != hipSuccess) return -1.0f;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pairs[i].first, pairs[i].second) != hipSuccess) return -1.0f;
            sum += ms;
        }
        return sum;
    }
};
thread_local KernelTimer t_timer;

// Brackets the SW forward kernel inside sw_launch (the walk is not timed).
struct TimerHooks : ga::LaunchHooks {
    void before(hipStream_t st) override { t_timer.begin(st); }
    void after(hipStream_t st) override { t_timer.end(st); }
};
thread_local TimerHooks t_timer_hooks;

// ---------------------------------------------------------------------------
// the library-side SW CIGAR check (api.cpp:114-156) on the host, for
// ga_sw_validate_cigar (the batches run their checks on the device)
// ---------------------------------------------------------------------------
// End of the run of cg[c] in cg[0 .. len): eight bytes at a time against the
// run's byte broadcast, then the first differing byte.
int cigar_run_end(const char* cg, int c, int len)
{
    const uint64_t rep = uint64_t(uint8_t(cg[c])) * 0x0101010101010101ull;
    int e = c + 1;
    while (e + 8 <= len) {
        uint64_t w;
        std::memcpy(&w, cg + e, 8);
        const uint64_t x = w ^ rep;
        if (x) return e + (__builtin_ctzll(x) >> 3);
        e += 8;
    }
    while (e < len && cg[e] == cg[c]) ++e;
    return e;
}

// The same checks as the reference's SW validator (api.cpp:114-156), a run
// of equal ops at a time: every op inside the span [start, end] and the
// sequences, M on equal and X on unequal bytes, the walk ending at (end_i,
// end_j), and the affine re-score (a gap run of L costs go + (L-1) ge) equal
// to `score`.  Positions only grow along a run, so a run's ops lie in range
// exactly up to the first that does not, and its byte tests are one
// branch-free loop.
void sw_validate(const char* cg, int len, const char* text, int text_len, int start_i, int end_i,
                 const char* pattern, int pattern_len, int start_j, int end_j,
                 const ga_sw_scoring& sc, int score, int32_t* rescore_ok, int32_t* wellformed_ok)
{
    int i = start_i, j = start_j, s = 0, run = 0;    // run: open gap, > 0 inserted, < 0 deleted
    bool wf = true;
    auto close_gap = [&] {
        if (run != 0) {
            const int L = run > 0 ? run : -run;
            s -= sc.gap_open + (L - 1) * sc.gap_extend;
            run = 0;
        }
    };
    const int lim_i = std::min(end_i + 1, text_len), lim_j = std::min(end_j + 1, pattern_len);
    for (int c = 0; c < len;) {
        const int e = cigar_run_end(cg, c, len);
        const int L = e - c;
        switch (cg[c]) {
            case 'M':
            case 'X': {
                close_gap();
                const int in = std::max(0, std::min(L, std::min(lim_i - i, lim_j - j)));
                if (in < L) wf = false;
                const uint8_t* a = reinterpret_cast<const uint8_t*>(text) + i;
                const uint8_t* b = reinterpret_cast<const uint8_t*>(pattern) + j;
                uint8_t bad = 0;
                if (cg[c] == 'M') {
                    for (int t = 0; t < in; ++t) bad |= uint8_t(a[t] ^ b[t]);
                    s += L * sc.match;
                } else {
                    for (int t = 0; t < in; ++t) bad |= uint8_t(a[t] == b[t]);
                    s += L * sc.mismatch;
                }
                if (bad) wf = false;
                i += L; j += L;
                break;
            }
            case 'I':
                if (run < 0) close_gap();
                if (i + L - 1 > end_i || i + L - 1 >= text_len) wf = false;
                i += L; run += L;
                break;
            case 'D':
                if (run > 0) close_gap();
                if (j + L - 1 > end_j || j + L - 1 >= pattern_len) wf = false;
                j += L; run -= L;
                break;
            default: wf = false; break;
        }
        c = e;
    }
    close_gap();
    if (i != end_i + 1 || j != end_j + 1) wf = false;
    *wellformed_ok = wf;
    *rescore_ok = wf && (s == score);
}

// ---------------------------------------------------------------------------
// WFA batch on the device
// ---------------------------------------------------------------------------
// Size classes of the persistent WFA kernel, by a pair's smax: each class is
// one launch whose LDS wavefronts and per-workgroup stacks are sized for its
// bound; the stacks of one class launch stay under kWfaStackBudget.
constexpr int kWfaClassBound[] = {ga::kWfaClassBound0, ga::kWfaClassBound1, ga::kWfaClassBound2, GA_MAX_SMAX_EXT};
constexpr int kWfaNumClasses = 4;
// 12 GiB: ~1,460 groups of the largest class (8.4 MB stacks each).  The 200k
// mixed batch runs as fast as with 24 GiB and slows below 8 GiB (30.7 ms at 8
// to 24, 34.7 ms at 6, 47 ms at 4); batches whose largest class holds few
// pairs gain from the lower cap (the routing batch 15.6 -> 14.9 ms), since
// the other classes, on their own streams, get more of the chip.
constexpr int64_t kWfaStackBudget = int64_t(12) << 30;   // bytes of stacks per class launch
// Per-thread cap on the workspace of one batch call (0: the defaults above);
// a serving loop gives each of its concurrent slots a share of its budget.
thread_local int64_t t_workspace_budget = 0;
int64_t wfa_stack_budget()
{
    return t_workspace_budget > 0 ? std::max<int64_t>(t_workspace_budget / 4, 1 << 20) : kWfaStackBudget;
}
// Launches with several waves per workgroup and CIGARs defer their walks: each
// workgroup's stack region holds up to kWfaDeferMult worst-case stacks (as the
// budget allows), so several pairs' walks run together on all its waves.
constexpr int64_t kWfaDeferBudget = int64_t(32) << 30;   // bytes of stack regions per class launch
constexpr int64_t kWfaDeferMult = 4;

// The stack bytes a class launch may take with deferred walks: the default
// deferral budget unless a thread budget (serving) or a caller's region caps it.
int64_t wfa_defer_cap(int64_t stack_cap, bool hard_cap)
{
    return (hard_cap || t_workspace_budget > 0) ? stack_cap : std::max(stack_cap, kWfaDeferBudget);
}

// Workgroups and stack region of a class launch (sets n_groups, ws_stride):
// workgroups capped so one worst-case stack each fits stack_cap, regions of up
// to kWfaDeferMult stacks within defer_cap.  Returns the region bytes per
// workgroup.
int64_t wfa_size_stacks(ga::WfaQueueLaunch& L, bool trace, int64_t stack_cap, int64_t defer_cap)
{
    int groups = ga::wfa_queue_groups(L);
    const int64_t one = int64_t(ga::wfa_queue_ws_bytes(L.smax_cls, 1, L.wide));
    int64_t mult = 1;
    if (trace) {
        groups = int(std::max<int64_t>(1, std::min<int64_t>(groups, stack_cap / one)));
        if (L.waves > 1)
            mult = std::max<int64_t>(1, std::min<int64_t>(kWfaDeferMult, defer_cap / (int64_t(groups) * one)));
    }
    L.n_groups = groups;
    L.ws_stride = one * mult / (L.wide ? 4 : 2);
    return one * mult;
}

struct WfaPlan {
    std::vector<int32_t> order;                 // pair ids, class by class (empty: identity)
    std::vector<int> bound;                     // per class: smax bound
    std::vector<int> max_plen, max_tlen;        // per class: longest sequences
    std::vector<int64_t> first, count;          // per class: slice of `order`
};

// Classes by smax; inside a class the longest pairs are claimed first (a
// counting sort on (pattern_len + text_len) / 32, so planning stays O(n), over
// the bins the batch occupies only).  `ids` (optional) restricts the plan to
// those pairs.  A batch of one class whose pairs share one length bin keeps
// its own order: `order` stays empty and the launch claims pairs 0 .. n-1.
WfaPlan plan_wfa(int64_t n, const int32_t* smax, const int32_t* plen, const int32_t* tlen,
                 const std::vector<int32_t>* ids = nullptr)
{
    constexpr int kKeys = 1 << 14;
    WfaPlan P;
    const int64_t cnt = ids ? int64_t(ids->size()) : n;
    auto pair_at = [&](int64_t x) { return ids ? int64_t((*ids)[size_t(x)]) : x; };
    std::vector<uint16_t> key(static_cast<size_t>(cnt));
    std::vector<uint8_t> cls(static_cast<size_t>(cnt));
    int cmax[kWfaNumClasses] = {0, 0, 0, 0}, pmax[kWfaNumClasses] = {0, 0, 0, 0}, tmax[kWfaNumClasses] = {0, 0, 0, 0};
    int kmin[kWfaNumClasses] = {kKeys, kKeys, kKeys, kKeys}, kmax[kWfaNumClasses] = {-1, -1, -1, -1};
    int64_t ccount[kWfaNumClasses] = {0, 0, 0, 0};
    for (int64_t x = 0; x < cnt; ++x) {
        const int64_t i = pair_at(x);
        int c = 0;
        while (c < kWfaNumClasses - 1 && smax[i] > kWfaClassBound[c]) ++c;
        const int64_t l = (int64_t(std::max(plen[i], 0)) + std::max(tlen[i], 0)) >> 5;
        const int k = int(kKeys - 1 - std::min<int64_t>(l, kKeys - 1));      // longest first
        cls[size_t(x)] = uint8_t(c);
        key[size_t(x)] = uint16_t(k);
        cmax[c] = std::max(cmax[c], int(smax[i]));
        pmax[c] = std::max(pmax[c], int(plen[i]));
        tmax[c] = std::max(tmax[c], int(tlen[i]));
        kmin[c] = std::min(kmin[c], k);
        kmax[c] = std::max(kmax[c], k);
        ++ccount[c];
    }
    int used = 0;
    bool one_bin = true;
    std::vector<int64_t> hist[kWfaNumClasses];
    int64_t at = 0;
    for (int c = 0; c < kWfaNumClasses; ++c) {
        if (ccount[c] == 0) continue;
        ++used;
        one_bin = one_bin && kmin[c] == kmax[c];
        P.bound.push_back(std::max(cmax[c], 0));
        P.max_plen.push_back(pmax[c]);
        P.max_tlen.push_back(tmax[c]);
        P.first.push_back(at);
        P.count.push_back(ccount[c]);
        hist[c].assign(size_t(kmax[c] - kmin[c] + 1), 0);
        at += ccount[c];
    }
    if (used == 1 && one_bin && !ids) return P;             // identity order
    for (int64_t x = 0; x < cnt; ++x) {
        const int c = cls[size_t(x)];
        ++hist[c][size_t(key[size_t(x)] - kmin[c])];
    }
    at = 0;
    for (int c = 0; c < kWfaNumClasses; ++c)                // exclusive scan -> slots
        for (auto& h : hist[c]) {
            const int64_t v = h;
            h = at;
            at += v;
        }
    P.order.resize(static_cast<size_t>(cnt));
    for (int64_t x = 0; x < cnt; ++x) {
        const int c = cls[size_t(x)];
        P.order[size_t(hist[c][size_t(key[size_t(x)] - kmin[c])]++)] = int32_t(pair_at(x));
    }
    return P;
}

// Wavefronts per workgroup for a class: wide wavefronts split over more waves.
// Wavefronts per workgroup of each size class: the autotuner's choice when
// one is set (process-wide), else the default (wide wavefronts over more waves).
std::atomic<int> g_tuned_waves[4] = {{0}, {0}, {0}, {0}};
int wfa_class_of(int bound)
{
    int c = 0;
    while (c < 3 && bound > kWfaClassBound[c]) ++c;
    return c;
}
// Waves per workgroup of a size class: the tuned value, else by class.  The
// largest class takes 4 waves when it holds many pairs (throughput: more pairs
// in flight per CU; measured 30 ms against 38 ms with 8 on the 200k mixed
// batch) and 8 when it holds few or an unknown number, `count` < 0
// (device-planned batches: the router's passes leave few pairs there), where
// the batch's tail is a handful of long pairs and latency decides.
int wfa_class_waves(int bound, int64_t count = -1)
{
    const int t = g_tuned_waves[wfa_class_of(bound)].load(std::memory_order_relaxed);
    if (t == 1 || t == 2 || t == 4 || t == 8) return t;
    if (bound <= 64) return 1;
    if (bound <= 256) return 2;
    if (bound <= 1024) return 4;
    return count >= int64_t(ga::device_cus()) * 24 ? 4 : 8;
}
constexpr size_t kWfaMaxStagedLds = 64 * 1024;

const char* enqueue_wfa(const ga::WfaQueueLaunch& base, const WfaPlan& plan, bool trace,
                        hipStream_t st, KernelTimer* timer)
{
    const size_t nc = plan.bound.size();
    if (nc == 0) return nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_next = nullptr;
    auto drop_queue = [&] {
        if (d_order) (void)ws_free(d_order, st);
        if (d_next) (void)ws_free(d_next, st);
    };
    const bool identity = plan.order.empty();
    if ((!identity && ws_alloc(reinterpret_cast<void**>(&d_order), plan.order.size() * 4 + 64, st) != hipSuccess) ||
        ws_alloc(reinterpret_cast<void**>(&d_next), nc * 4 + 64, st) != hipSuccess) {
        drop_queue();
        return "hipMalloc(WFA queue)";
    }
    if (!identity && hipMemcpyAsync(d_order, plan.order.data(), plan.order.size() * 4, hipMemcpyHostToDevice,
                                    st) != hipSuccess) {
        drop_queue();
        return "hipMemcpy(WFA queue)";
    }
    // set-up first (launch shapes, stacks), so the timed region holds only GPU work
    const char* err = nullptr;
    std::vector<ga::WfaQueueLaunch> Ls(nc);
    for (size_t c = 0; c < nc && !err; ++c) {
        ga::WfaQueueLaunch& L = Ls[c];
        L = base;
        L.order = identity ? nullptr : d_order + plan.first[c];
        L.n = int32_t(plan.count[c]);
        L.next = d_next + c;
        L.counter_zeroed = true;                       // zeroed above, before the timed region
        L.smax_cls = plan.bound[c];
        L.max_plen = plan.max_plen[c];
        L.max_tlen = plan.max_tlen[c];
        L.waves = wfa_class_waves(L.smax_cls, plan.count[c]);
        L.staged = true;
        if (ga::wfa_lds_bytes(L) > kWfaMaxStagedLds) L.staged = false;
        const int64_t per_group = wfa_size_stacks(L, trace, wfa_stack_budget(), wfa_defer_cap(wfa_stack_budget(), false));
        L.ws = nullptr;
        if (trace && ws_alloc(&L.ws, size_t(per_group) * size_t(L.n_groups), st) != hipSuccess)
            err = "hipMalloc(wavefront stacks)";
    }
    // claim counters, unless every class is scheduled statically throughout
    bool dynamic = false;
    for (size_t c = 0; c < nc; ++c) dynamic = dynamic || !ga::wfa_fully_static(Ls[c]);
    if (!err && dynamic && ga::zero_i32_launch(d_next, int(nc), st) != hipSuccess) err = "WFA counter launch";
    if (timer && !err) timer->begin(st);
    hipEvent_t fork = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joins;
    if (!err && nc > 1 && (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess ||
                           hipEventRecord(fork, st) != hipSuccess))
        err = "hipEventRecord";
    for (size_t c = 0; c < nc && !err; ++c) {
        hipStream_t cs = st;
        if (c > 0) {
            cs = acquire_stream();
            if (!cs || hipStreamWaitEvent(cs, fork, 0) != hipSuccess) { err = "hipStreamWaitEvent"; break; }
            side.push_back(cs);
        }
        if (ga::wfa_queue_launch(Ls[c], cs) != hipSuccess) err = "WFA kernel launch";
        if (c > 0) {
            hipEvent_t j = nullptr;
            if (hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess &&
                hipEventRecord(j, cs) == hipSuccess && hipStreamWaitEvent(st, j, 0) == hipSuccess)
                joins.push_back(j);
            else if (!err)
                err = "hipEventRecord";
        }
    }
    if (timer && !err) timer->end(st);
    for (hipEvent_t j : joins) (void)hipEventDestroy(j);
    if (fork) (void)hipEventDestroy(fork);
    for (hipStream_t cs : side) release_stream(cs);
    for (auto& L : Ls)
        if (L.ws) (void)ws_free(L.ws, st);
    drop_queue();
    return err;
}

// Launch shapes and stacks of the four size classes of a device-planned batch
// (class sizes unknown on the host: every class that can hold a pair is
// launched over a full grid, and a class whose plan count is 0 exits at once).
struct WfaDevClasses {
    std::vector<ga::WfaQueueLaunch> L;
    int32_t* order = nullptr;
    int32_t* next = nullptr;
    ga::WfaDevPlan plan{};
    bool planless = false;      // one class holds every pair: no plan kernels
};

// `stack_cap`: bytes of wavefront stacks one class launch may hold (its
// workgroup count is capped to fit); `stack_alloc` (default: `alloc`) gives
// each class its stacks.
const char* setup_wfa_dev_classes(const ga::WfaQueueLaunch& base, int64_t n, int smax_max, int max_len,
                                  bool trace, WfaDevClasses& C, const std::function<void*(size_t)>& alloc,
                                  int64_t stack_cap = -1,
                                  const std::function<void*(size_t)>* stack_alloc = nullptr)
{
    const bool hard_cap = stack_cap >= 0;            // a caller's region: no deferral beyond it
    if (stack_cap < 0) stack_cap = wfa_stack_budget();
    if (!stack_alloc) stack_alloc = &alloc;
    C.order = static_cast<int32_t*>(alloc(size_t(n) * 4 + 64));
    C.next = static_cast<int32_t*>(alloc(64));
    C.plan.hist = static_cast<int32_t*>(alloc(size_t(ga::kWfaClasses) * ga::kPlanKeys * 4));
    C.plan.cls = static_cast<int32_t*>(alloc(64));
    if (!C.order || !C.next || !C.plan.hist || !C.plan.cls) return "hipMalloc(WFA plan)";
    C.plan.smax = base.smax; C.plan.plen = base.plen; C.plan.tlen = base.tlen;
    C.plan.n = int32_t(n);
    C.plan.order = C.order;
    C.plan.next = C.next;
    // every smax within the smallest class: one launch over the pairs in
    // index order, no plan (the claim order only balances long pairs)
    C.planless = smax_max <= kWfaClassBound[0];
    for (int c = 0; c < ga::kWfaClasses; ++c) {
        if (c > 0 && smax_max <= kWfaClassBound[c - 1]) break;        // no pair can fall in class c
        ga::WfaQueueLaunch L = base;
        L.smax_cls = c < ga::kWfaClasses - 1 ? std::min(kWfaClassBound[c], smax_max) : smax_max;
        L.max_plen = max_len;
        L.max_tlen = max_len;
        L.waves = wfa_class_waves(L.smax_cls);
        L.staged = true;
        if (ga::wfa_lds_bytes(L) > kWfaMaxStagedLds) L.staged = false;
        L.n = int32_t(n);
        L.order = C.planless ? nullptr : C.order;
        L.next = C.next + c;
        L.dev_class = C.planless ? nullptr : C.plan.cls + 2 * c;
        const int64_t per_group = wfa_size_stacks(L, trace, stack_cap, wfa_defer_cap(stack_cap, hard_cap));
        L.ws = trace ? (*stack_alloc)(size_t(per_group) * size_t(L.n_groups)) : nullptr;
        if (trace && !L.ws) return "hipMalloc(wavefront stacks)";
        C.L.push_back(L);
    }
    return nullptr;
}

// Plan on the device, then the class launches on forked streams joined back
// into `st`; `side` holds one stream per class beyond the first (so the same
// streams can serve a graph capture).  No host synchronisation.
// Classes from `detach` on are not joined: their completion is recorded in
// ev[c] for the caller to wait on.
const char* enqueue_wfa_dev(const WfaDevClasses& C, hipStream_t st, const std::vector<hipStream_t>& side,
                            std::vector<hipEvent_t>& ev, size_t detach = SIZE_MAX)
{
    if (!C.planless && ga::wfa_plan_launch(C.plan, st) != hipSuccess) return "WFA plan launch";
    const size_t nc = C.L.size();
    // ev[0] forks, ev[1..] join
    if (nc > 1 && hipEventRecord(ev[0], st) != hipSuccess) return "hipEventRecord";
    for (size_t c = 0; c < nc; ++c) {
        hipStream_t cs = c == 0 ? st : side[c - 1];
        if (c > 0 && hipStreamWaitEvent(cs, ev[0], 0) != hipSuccess) return "hipStreamWaitEvent";
        if (ga::wfa_queue_launch(C.L[c], cs) != hipSuccess) return "WFA kernel launch";
        if (c > 0 && hipEventRecord(ev[c], cs) != hipSuccess) return "hipEventRecord";
        if (c > 0 && c < detach && hipStreamWaitEvent(st, ev[c], 0) != hipSuccess) return "hipStreamWaitEvent";
    }
    return nullptr;
}

struct WfaOut {
    std::vector<int32_t> score, cig_len;
    std::vector<uint8_t> vflags;     // the device's CIGAR checks: rescore_ok | wellformed_ok << 1
    std::vector<int64_t> slot_off;   // slot of pair i in the arena
    const char* cig_src = nullptr;   // where the CIGAR slots came back (pinned staging or the arena)
    PinnedHold pin;
};

// Runs every pair with its own smax; trace[i] == 0 means score only (no CIGAR
// for that pair).  The CIGAR bytes land in `arena` (slot layout).
const char* run_wfa(const ga_align_req* reqs, int64_t n, const std::vector<int32_t>& smax,
                    const std::vector<uint8_t>& trace, char* arena, WfaOut& out)
{
    std::call_once(g_pool_once, init_pool);
    out.slot_off.resize(static_cast<size_t>(n));
    int64_t slot = 0;
    std::vector<int32_t> plen(static_cast<size_t>(n)), tlen(static_cast<size_t>(n));
    bool any_trace = false;
    int32_t max_text = 0;
    for (int64_t i = 0; i < n; ++i) {
        const ga_align_req& r = reqs[i];
        max_text = std::max(max_text, r.text_len);
        plen[size_t(i)] = r.pattern_len;
        tlen[size_t(i)] = r.text_len;
        out.slot_off[size_t(i)] = slot;
        slot += std::max(r.pattern_len, 0) + std::max(r.text_len, 0);
        any_trace |= trace[size_t(i)] != 0;
    }
    // both sequence sets packed straight into pinned staging where it can be had
    Packed P, T;
    const auto pget = [&](int64_t i) { return std::make_pair(reqs[i].pattern, reqs[i].pattern_len); };
    const auto tget = [&](int64_t i) { return std::make_pair(reqs[i].text, reqs[i].text_len); };
    const int64_t psz = P.layout(n, pget), tsz = T.layout(n, tget);
    uint8_t* ppin = out.pin.get(size_t(psz));
    uint8_t* tpin = ppin ? out.pin.get(size_t(tsz)) : nullptr;
    const bool pinned = ppin && tpin;
    if (pinned) {
        P.fill(n, pget, ppin, psz);
        T.fill(n, tget, tpin, tsz);
    } else {
        P.build(n, pget);
        T.build(n, tget);
    }
    const WfaPlan plan = plan_wfa(n, smax.data(), plen.data(), tlen.data());

    hipStream_t st = acquire_stream();
    if (!st) return "hipStreamCreate";
    DevArena D;
    D.st = st;
    const char* err = nullptr;
    ga::WfaQueueLaunch L{};
    L.pat = pinned ? D.upload_raw(ppin, size_t(psz)) : D.upload(P.bytes);
    L.poff = D.upload(P.off); L.plen = D.upload(plen);
    L.txt = pinned ? D.upload_raw(tpin, size_t(tsz)) : D.upload(T.bytes);
    L.toff = D.upload(T.off); L.tlen = D.upload(tlen);
    L.smax = D.upload(smax);
    L.wide = max_text > 32766;
    L.score = D.alloc<int32_t>(size_t(n));
    if (any_trace) {
        L.want = D.upload(trace);
        L.cigar = D.alloc<uint8_t>(size_t(slot));
        L.cig_off = D.upload(out.slot_off);
        L.cig_len = D.alloc<int32_t>(size_t(n));
    }
    if (!D.ok) err = "hipMalloc/hipMemcpy (WFA buffers)";
    t_timer.reset();
    float rq = 0.0f;
    const float* rp = ga::router_params(&rq);
    if (!err && rp) {
        // a loaded network chooses each pair's first bound and size class;
        // every pair it sends too low is re-run with its own smax, so the
        // results are the plain path's pair for pair
        ga_wfa_dev_batch b{};
        b.pattern = L.pat; b.pattern_off = L.poff; b.pattern_len = L.plen;
        b.text = L.txt; b.text_off = L.toff; b.text_len = L.tlen;
        b.smax = L.smax; b.n_pairs = n;
        b.smax_max = *std::max_element(smax.begin(), smax.end());
        b.max_len = std::max(*std::max_element(plen.begin(), plen.end()), max_text);
        b.score = L.score;
        if (any_trace) { b.cigar = L.cigar; b.cigar_off = L.cig_off; b.cigar_len = L.cig_len; b.want_cigar = L.want; }
        err = ga::wfa_routed_run(&b, rp, rq, nullptr, st);
    } else if (!err) {
        err = enqueue_wfa(L, plan, any_trace, st, &t_timer);
    }
    // the CIGAR checks on the device, once every CIGAR is written
    uint8_t* d_vf = nullptr;
    if (!err && any_trace) {
        d_vf = D.alloc<uint8_t>(size_t(n));
        ga::SwValidate V{};
        V.txt = L.txt; V.toff = L.toff; V.tlen = L.tlen;
        V.pat = L.pat; V.poff = L.poff; V.plen = L.plen;
        V.cigar = L.cigar; V.cig_off = L.cig_off;
        V.score = L.score; V.stride = 1; V.len = L.cig_len; V.len_stride = 1;
        V.walk = L.want; V.flags = d_vf; V.n = n; V.edit = 1;
        if (!D.ok || ga::sw_validate_launch(V, st) != hipSuccess) err = "WFA validate launch";
    }
    // scores, CIGAR lengths, check flags and slots back into pinned staging
    // (pageable fallback: the vectors and the caller's arena)
    out.score.assign(static_cast<size_t>(n), -1);
    out.cig_len.assign(static_cast<size_t>(n), 0);
    if (d_vf) out.vflags.assign(static_cast<size_t>(n), 0);
    int32_t* sc_h = reinterpret_cast<int32_t*>(out.pin.get(size_t(n) * 8));
    int32_t* cl_h = sc_h ? sc_h + n : nullptr;
    char* cdst = arena;
    if (any_trace && slot > 0)
        if (uint8_t* c = out.pin.get(size_t(slot))) cdst = reinterpret_cast<char*>(c);
    if (!err) {
        if (hipMemcpyAsync(sc_h ? sc_h : out.score.data(), L.score, size_t(n) * 4, hipMemcpyDeviceToHost, st) !=
            hipSuccess)
            err = "hipMemcpy(results)";
        if (!err && any_trace) {
            if (hipMemcpyAsync(cl_h ? cl_h : out.cig_len.data(), L.cig_len, size_t(n) * 4, hipMemcpyDeviceToHost,
                               st) != hipSuccess ||
                (d_vf && hipMemcpyAsync(out.vflags.data(), d_vf, size_t(n), hipMemcpyDeviceToHost, st) != hipSuccess) ||
                (slot > 0 && hipMemcpyAsync(cdst, L.cigar, size_t(slot), hipMemcpyDeviceToHost, st) != hipSuccess))
                err = "hipMemcpy(results)";
        }
    }
    D.release();
    const hipError_t se = hipStreamSynchronize(st);
    if (!err && se != hipSuccess) err = hipGetErrorString(se);
    release_stream(st);
    out.cig_src = cdst;
    if (!err && sc_h) {
        std::copy(sc_h, sc_h + n, out.score.begin());
        if (any_trace) std::copy(cl_h, cl_h + n, out.cig_len.begin());
    }
    return err;
}

void reset_wfa(ga_align_res* res, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) res[i] = ga_align_res{-1, 0, 0, 0, 0, 0};
}

int wfa_common(const ga_align_req* reqs, int64_t n, ga_align_res* res, char* arena,
               int64_t arena_bytes, int32_t* resolved_count, int32_t* unresolved_count,
               const char** error, bool ext)
{
    if (error) *error = nullptr;
    if (resolved_count) *resolved_count = 0;
    if (unresolved_count) *unresolved_count = 0;
    if (n <= 0) return GA_OK;
    reset_wfa(res, n);
    auto refuse = [&](const char* msg) {
        if (error) *error = msg;
        reset_wfa(res, n);
        return GA_INVALID_ARGUMENT;
    };
    if (!reqs || !res) return refuse("null request or result array");
    const int max_smax = ext ? GA_MAX_SMAX_EXT : GA_MAX_SMAX;
    for (int64_t i = 0; i < n; ++i) {
        const ga_align_req& r = reqs[i];
        if ((r.pattern_len > 0 && r.pattern == nullptr) || (r.text_len > 0 && r.text == nullptr))
            return refuse("null pointer with non-zero length");
        if (r.pattern_len < 0 || r.text_len < 0) return refuse("negative length");
        if (r.smax < 0 || r.smax > max_smax)
            return refuse(ext ? "smax out of range [0, 4000]" : "smax out of range [0, 511]");
    }
    const int64_t need = ga_cigar_arena_bytes_wfa(reqs, n);
    if (need > 0 && (arena == nullptr || arena_bytes < need))
        return refuse("CIGAR arena smaller than the batch needs (ga_cigar_arena_bytes_wfa)");

    std::vector<int32_t> smax(static_cast<size_t>(n));
    std::vector<uint8_t> trace(static_cast<size_t>(n));
    bool any_cigar = false;
    int32_t smin = reqs[0].smax;
    for (int64_t i = 0; i < n; ++i) {
        any_cigar |= reqs[i].with_cigar != 0;
        smin = std::min(smin, reqs[i].smax);
    }
    for (int64_t i = 0; i < n; ++i) {
        if (ext) {
            smax[size_t(i)] = reqs[i].smax;
            trace[size_t(i)] = reqs[i].with_cigar != 0;
        } else {
            smax[size_t(i)] = smin;
            trace[size_t(i)] = any_cigar;
        }
    }
    WfaOut out;
    const char* err = run_wfa(reqs, n, smax, trace, arena, out);
    if (err) {
        if (error) *error = err;
        reset_wfa(res, n);
        return GA_DEVICE_ERROR;
    }
    parallel_for(n, [&](int64_t i) {
        const ga_align_req& r = reqs[i];
        ga_align_res& o = res[i];
        o.score = out.score[size_t(i)];
        o.resolved = o.score >= 0;
        const int len = out.cig_len[size_t(i)];
        const int slot_len = std::max(r.pattern_len, 0) + std::max(r.text_len, 0);
        o.cigar_off = out.slot_off[size_t(i)] + slot_len - len;
        if (o.resolved && trace[size_t(i)] && len > 0) {
            o.cigar_len = len;
            if (out.cig_src != arena) std::memcpy(arena + o.cigar_off, out.cig_src + o.cigar_off, size_t(len));
            // the device's checks (every CIGAR batch runs them; without their
            // flags both stay 0)
            if (!out.vflags.empty()) {
                o.rescore_ok = out.vflags[size_t(i)] & 1;
                o.wellformed_ok = (out.vflags[size_t(i)] >> 1) & 1;
            }
        } else {
            o.cigar_len = 0;
            o.cigar_off = out.slot_off[size_t(i)];
        }
    });
    int32_t rc = 0;
    for (int64_t i = 0; i < n; ++i) rc += res[i].resolved ? 1 : 0;
    if (resolved_count) *resolved_count = rc;
    if (unresolved_count) *unresolved_count = int32_t(n) - rc;
    return GA_OK;
}

// ---------------------------------------------------------------------------
// SW batch on the device
// ---------------------------------------------------------------------------
constexpr int64_t kSwDirBudgetDefault = int64_t(16) << 30;   // direction nibbles per chunk
int64_t sw_dir_budget()
{
    return t_workspace_budget > 0 ? std::max<int64_t>(t_workspace_budget / 2, 1 << 20) : kSwDirBudgetDefault;
}
constexpr int kLdsBndMaxRows = 8192;                     // 2 ints per row in LDS
constexpr int kLdsBndMaxRowsMatrix = 7168;               // leaves 8 KB for the score profile

// Substitution-matrix scheme on the host: residue code of every byte (the
// alphabet's letters 0 .. k-1, any other byte kSwUnknownCode) and the 32 x 32
// table over codes (rows: text residue, columns: pattern residue).
constexpr int kSwUnknownCode = 31;
struct SwMatrixHost {
    uint8_t code[256];
    int32_t sub[32 * 32];
    int32_t gap_open, gap_extend;
    int score(uint8_t t, uint8_t p) const { return sub[code[t] * 32 + code[p]]; }
    bool fits8() const
    {
        for (int x = 0; x < 32 * 32; ++x)
            if (sub[x] < -128 || sub[x] > 127) return false;
        return true;
    }
};

const char* sw_scheme_error(const ga_sw_scoring& s, bool with_cigar)
{
    if (s.match <= 0) return "SW scoring: match must be > 0 (the score is 0 for every input otherwise)";
    if (s.mismatch >= s.match) return "SW scoring: mismatch must be < match";
    if (s.gap_open <= 0 || s.gap_extend <= 0) return "SW scoring: gap costs must be > 0";
    if (with_cigar && s.gap_extend > s.gap_open)
        return "SW scoring: gap_extend > gap_open with with_cigar cannot re-score "
               "to its own score (api.hpp; RESULTADO_H9_SW2_TRACE.md). Score-only "
               "requests may use it";
    return nullptr;
}

// Enqueue the packed 16-bit path for one chunk of waves: the packed kernel
// (forward and walk), then the int32 kernel (forward + walk) over the pairs it
// flagged.  A flagged pair reuses its wave's direction region, which still holds
// the partner half's packed flags until the packed walk has run -- hence this
// order.  All on `st`, no host synchronisation.
hipError_t enqueue_sw16_chunk(const ga::Sw16Launch& P16, bool trace, hipStream_t st,
                              ga::LaunchHooks* hooks, const SwMatrixHost* mat = nullptr)
{
    // (the packed kernels' trace variants walk their pairs before the re-run
    // kernel reuses the direction regions)
    hipError_t e = mat ? ga::sw16m_launch(P16, trace, st, hooks) : ga::sw16_launch(P16, trace, st, hooks);
    if (e != hipSuccess) return e;
    ga::SwLaunch R{};
    R.txt = P16.txt; R.toff = P16.toff; R.tlen = P16.tlen;
    R.pat = P16.pat; R.poff = P16.poff; R.plen = P16.plen;
    R.n = 2 * P16.n_waves; R.max_m = P16.max_m;
    R.match = P16.match; R.mismatch = P16.mismatch; R.go = P16.go; R.ge = P16.ge;
    R.res = P16.res; R.dirs = P16.dirs; R.dir_off = nullptr; R.dir_stride = P16.dir_stride;
    R.walk = P16.walk; R.cigar = P16.cigar; R.cig_off = P16.cig_off; R.bnd = nullptr;
    R.first_pair = P16.first_pair; R.rerun = P16.rerun;
    R.pair_a = P16.pair_a; R.pair_b = P16.pair_b; R.n_total = P16.n_total;
    if (mat) { R.tcode = P16.tcode; R.pcode = P16.pcode; R.sub = P16.sub; R.sub8 = mat->fits8(); }
    return ga::sw_launch(R, trace, st);
}

// Device SW over pairs given as pointers; res8 receives 8 ints per pair (see
// ga::SwLaunch).  walk[i] selects pairs whose span + CIGAR are produced.
// The CIGAR slots come back into pinned staging from `pin` (or, without
// pinned memory, into arena): *cig_src tells the caller where they are.
const char* run_sw(const ga_sw_req* reqs, int64_t n, const std::vector<uint8_t>& skip,
                   bool trace, const std::vector<uint8_t>& walk, char* arena,
                   std::vector<int32_t>& res8, std::vector<int64_t>& slot_off,
                   PinnedHold& pin, const char** cig_src, const SwMatrixHost* mat = nullptr,
                   std::vector<uint8_t>* vflags = nullptr, std::vector<int64_t>* csrc_off = nullptr)
{
    *cig_src = arena;
    std::call_once(g_pool_once, init_pool);
    t_timer.reset();
    // Order on the device: pairs for the int32 kernel first (texts longer than
    // the packed kernel takes, and empty or skipped pairs), then the packed
    // kernel's pairs sorted by shape so each wave's two pairs are alike.
    std::vector<int64_t> order;
    order.reserve(size_t(n));
    std::vector<int64_t> packed;
    slot_off.resize(static_cast<size_t>(n));
    int64_t slot = 0;
    // substitution-matrix pairs take the packed kernel when every biased
    // score fits a byte
    const int mbias = mat ? ga::sw16m_bias(mat->sub, mat->gap_open) : -1;
    for (int64_t i = 0; i < n; ++i) {
        const ga_sw_req& r = reqs[i];
        slot_off[size_t(i)] = slot;
        slot += std::max(r.text_len, 0) + std::max(r.pattern_len, 0);
        const bool sk = skip[size_t(i)] != 0;
        const bool fits = mat ? (mbias >= 0 && r.text_len <= ga::kSw16mMaxRows)
                              : (r.text_len <= ga::kSw16MaxRows && (!trace || r.pattern_len <= ga::kSw16MaxCols));
        if (!sk && r.text_len > 0 && r.pattern_len > 0 && fits)
            packed.push_back(i);
        else
            order.push_back(i);
    }
    const int64_t n32 = int64_t(order.size());
    if (mat) {
        // a workgroup computes max(rows) x max(strips) of its two pairs: pair
        // pairs with the same strip count and close row counts, largest first
        std::stable_sort(packed.begin(), packed.end(), [&](int64_t a, int64_t b) {
            const int sa = (reqs[a].pattern_len + 63) >> 6, sb = (reqs[b].pattern_len + 63) >> 6;
            if (sa != sb) return sa > sb;
            return reqs[a].text_len > reqs[b].text_len;
        });
        // then the workgroups (consecutive pairs of that order; an odd last
        // pair stays last, alone) by their cost, strips x (rows + 63), largest
        // first: the longest start first and the short ones fill in behind
        const size_t nd = packed.size() / 2;
        std::vector<std::pair<int64_t, size_t>> duo(nd);
        for (size_t d = 0; d < nd; ++d) {
            const int64_t rows = std::max(reqs[packed[2 * d]].text_len, reqs[packed[2 * d + 1]].text_len);
            duo[d] = {int64_t((reqs[packed[2 * d]].pattern_len + 63) >> 6) * (rows + 63), d};
        }
        std::stable_sort(duo.begin(), duo.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        std::vector<int64_t> by_cost;
        by_cost.reserve(packed.size());
        for (const auto& d : duo) {
            by_cost.push_back(packed[2 * d.second]);
            by_cost.push_back(packed[2 * d.second + 1]);
        }
        if (packed.size() & 1) by_cost.push_back(packed.back());
        packed.swap(by_cost);
    } else {
        std::stable_sort(packed.begin(), packed.end(), [&](int64_t a, int64_t b) {
            if (reqs[a].text_len != reqs[b].text_len) return reqs[a].text_len > reqs[b].text_len;
            return reqs[a].pattern_len > reqs[b].pattern_len;
        });
    }
    order.insert(order.end(), packed.begin(), packed.end());

    std::vector<int32_t> tl(static_cast<size_t>(n)), pl(static_cast<size_t>(n));
    std::vector<int64_t> cig_off(static_cast<size_t>(n));
    std::vector<uint8_t> walk_p(static_cast<size_t>(n));
    int32_t max_m32 = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = order[size_t(k)];
        const ga_sw_req& r = reqs[i];
        const bool sk = skip[size_t(i)] != 0;
        tl[size_t(k)] = sk ? 0 : r.text_len;
        pl[size_t(k)] = sk ? 0 : r.pattern_len;
        cig_off[size_t(k)] = slot_off[size_t(i)];
        walk_p[size_t(k)] = walk[size_t(i)];
        if (k < n32) max_m32 = std::max(max_m32, tl[size_t(k)]);
    }
    // the sequences in launch order, and (matrix scoring) their residue codes,
    // laid out like the bytes (the padding maps to the code of byte 0); packed
    // straight into pinned staging where it can be had
    Packed T, P;
    std::vector<uint8_t> tcode, pcode;
    const auto tget = [&](int64_t k) { return std::make_pair(reqs[order[size_t(k)]].text, tl[size_t(k)]); };
    const auto pget = [&](int64_t k) { return std::make_pair(reqs[order[size_t(k)]].pattern, pl[size_t(k)]); };
    const int64_t tsz = T.layout(n, tget), psz = P.layout(n, pget);
    const int cm = mat ? 2 : 1;                      // bytes (and codes)
    uint8_t* tpin = pin.get(size_t(tsz) * cm);
    uint8_t* ppin = tpin ? pin.get(size_t(psz) * cm) : nullptr;
    // large DNA batches are packed and uploaded in pieces below
    const bool pipe = tpin && ppin && !mat && n >= 4096;
    if (tpin && ppin) {
        if (!pipe) {
            T.fill(n, tget, tpin, tsz, mat ? mat->code : nullptr, mat ? tpin + tsz : nullptr);
            P.fill(n, pget, ppin, psz, mat ? mat->code : nullptr, mat ? ppin + psz : nullptr);
        }
    } else {
        T.build(n, tget, mat ? mat->code : nullptr, mat ? &tcode : nullptr);
        P.build(n, pget, mat ? mat->code : nullptr, mat ? &pcode : nullptr);
    }
    const bool pinned = tpin && ppin;

    hipStream_t st = acquire_stream();
    if (!st) return "hipStreamCreate";
    DevArena D;
    D.st = st;
    const char* err = nullptr;
    uint8_t* d_txt = nullptr;
    uint8_t* d_pat = nullptr;
    if (pipe) {
        // four pieces: each piece's copy runs while the host packs the next
        d_txt = D.alloc<uint8_t>(size_t(tsz));
        d_pat = D.alloc<uint8_t>(size_t(psz));
        for (int q = 0; q < 4 && D.ok; ++q) {
            const int64_t k0 = n * q / 4, k1 = n * (q + 1) / 4;
            T.fill_range(k0, k1, n, tget, tpin, tsz);
            P.fill_range(k0, k1, n, pget, ppin, psz);
            const auto tb = T.piece(k0, k1, n, tsz), pb = P.piece(k0, k1, n, psz);
            if (hipMemcpyAsync(d_txt + tb.first, tpin + tb.first, size_t(tb.second - tb.first),
                               hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_pat + pb.first, ppin + pb.first, size_t(pb.second - pb.first),
                               hipMemcpyHostToDevice, st) != hipSuccess)
                D.ok = false;
        }
    } else {
        d_txt = pinned ? D.upload_raw(tpin, size_t(tsz)) : D.upload(T.bytes);
        d_pat = pinned ? D.upload_raw(ppin, size_t(psz)) : D.upload(P.bytes);
    }
    const uint8_t* d_tcode = !mat ? nullptr : pinned ? D.upload_raw(tpin + tsz, size_t(tsz)) : D.upload(tcode);
    const uint8_t* d_pcode = !mat ? nullptr : pinned ? D.upload_raw(ppin + psz, size_t(psz)) : D.upload(pcode);
    std::vector<int32_t> subv;
    if (mat) subv.assign(mat->sub, mat->sub + 32 * 32);
    const int32_t* d_sub = mat ? D.upload(subv) : nullptr;
    auto* d_toff = D.upload(T.off);
    auto* d_poff = D.upload(P.off);
    auto* d_tl = D.upload(tl);
    auto* d_pl = D.upload(pl);
    auto* d_res = D.alloc<int32_t>(size_t(n) * 8);
    auto* d_rerun = D.alloc<uint8_t>(size_t(n));
    uint8_t* d_cig = nullptr;
    int64_t* d_slot = nullptr;
    uint8_t* d_walk = nullptr;
    if (trace) {
        d_cig = D.alloc<uint8_t>(size_t(slot));
        d_slot = D.upload(cig_off);
        d_walk = D.upload(walk_p);
    }
    const ga_sw_scoring sc = reqs[0].scoring;

    // ---- int32 kernel: pairs [0, n32) -------------------------------------
    if (n32 > 0 && D.ok) {
        std::vector<std::pair<int64_t, int64_t>> chunks;
        int64_t* d_diroff = nullptr;
        uint32_t* d_dirs = nullptr;
        if (trace) {
            int64_t a = 0, words = 0, max_words = 0;
            std::vector<int64_t> local(static_cast<size_t>(n32));
            for (int64_t k = 0; k < n32; ++k) {
                const int64_t w = ga::sw_dir_words(tl[size_t(k)], pl[size_t(k)]);
                if (k > a && (words + w) * 4 > sw_dir_budget()) {
                    chunks.emplace_back(a, k);
                    max_words = std::max(max_words, words);
                    a = k; words = 0;
                }
                local[size_t(k)] = words;
                words += w;
            }
            chunks.emplace_back(a, n32);
            max_words = std::max(max_words, words);
            d_diroff = D.upload(local);
            d_dirs = D.alloc<uint32_t>(size_t(max_words));
        } else {
            chunks.emplace_back(0, n32);
        }
        const bool lds_bnd = max_m32 <= (mat ? kLdsBndMaxRowsMatrix : kLdsBndMaxRows);
        int32_t* d_bnd = nullptr;
        const int64_t bnd_chunk = int64_t(1) << 16;
        if (!lds_bnd) d_bnd = D.alloc<int32_t>(size_t(std::min<int64_t>(n32, bnd_chunk)) * 2 * size_t(max_m32));
        if (!D.ok) err = "hipMalloc(dir table) -- batch workspace exceeds device memory; split the batch";
        for (auto& ch : chunks) {
            if (err) break;
            for (int64_t a = ch.first; a < ch.second && !err;) {
                const int64_t len = std::min(ch.second - a, lds_bnd ? (int64_t(1) << 30) : bnd_chunk);
                ga::SwLaunch L{};
                L.txt = d_txt; L.toff = d_toff; L.tlen = d_tl;
                L.pat = d_pat; L.poff = d_poff; L.plen = d_pl;
                L.n = int32_t(len); L.max_m = max_m32;
                L.match = sc.match; L.mismatch = sc.mismatch; L.go = sc.gap_open; L.ge = sc.gap_extend;
                L.res = d_res; L.dirs = d_dirs; L.dir_off = d_diroff; L.walk = d_walk;
                L.cigar = d_cig; L.cig_off = d_slot; L.bnd = d_bnd; L.first_pair = int32_t(a);
                L.n_total = int32_t(n);
                L.tcode = d_tcode; L.pcode = d_pcode; L.sub = d_sub;
                if (mat) { L.go = mat->gap_open; L.ge = mat->gap_extend; L.sub8 = mat->fits8(); }
                if (ga::sw_launch(L, trace, st, &t_timer_hooks) != hipSuccess) err = "SW kernel launch";
                a += len;
            }
        }
    }

    // ---- packed kernel: pairs [n32, n), two per wave ------------------------
    const int64_t npk = n - n32;
    if (npk > 0 && !err) {
        const int64_t n_waves = (npk + 1) / 2;
        // direction words per wave (DNA) or per workgroup (matrix kernel)
        const auto dir_words = [&](int m_, int n_) {
            return mat ? ga::sw16m_dir_words_per_group(m_, n_) : ga::sw16_dir_words_per_wave(m_, n_);
        };
        // Chunks of waves: a chunk's direction table stays within the
        // direction budget (16 GiB by default; with a caller's workspace
        // budget, half of the direction share, as two chunks may be in
        // flight, one per stream), and
        // (DNA) a wave whose LDS need is below half of its chunk's starts a
        // chunk of its own -- a launch sizes every wave's LDS for its largest
        // pair, and waves come longest text first, so short pairs would run
        // at a long pair's occupancy.
        struct WChunk { int64_t w0, w1; int32_t cm, cn; double steps; };
        std::vector<WChunk> wchunks;
        {
            const int64_t budget = t_workspace_budget > 0 ? std::max<int64_t>(sw_dir_budget() / 2, 1 << 20)
                                                          : sw_dir_budget();
            int64_t w0 = 0, mm = 0, mn = 0;
            double steps = 0.0;
            for (int64_t w = 0; w < n_waves; ++w) {
                int64_t wm = 0, wn = 0;
                for (int h = 0; h < 2; ++h) {
                    const int64_t k = n32 + 2 * w + h;
                    if (k < n) { wm = std::max<int64_t>(wm, tl[size_t(k)]); wn = std::max<int64_t>(wn, pl[size_t(k)]); }
                }
                const int64_t nm = std::max(mm, wm), nn = std::max(mn, wn);
                const int64_t stride = trace ? dir_words(int(nm), int(nn)) : 0;
                const bool lds_split = !mat && w > w0 &&
                                       2 * ga::sw16_lds_need(int(wm), int(wn), trace) <
                                           ga::sw16_lds_need(int(mm), int(mn), trace);
                if (w > w0 && (stride * (w - w0 + 1) * 4 > budget || lds_split)) {
                    wchunks.push_back({w0, w, int32_t(mm), int32_t(mn), steps});
                    w0 = w; mm = wm; mn = wn; steps = 0.0;
                } else {
                    mm = nm; mn = nn;
                }
                steps += double(wm + 64) * double((wn + 63) / 64);
            }
            wchunks.push_back({w0, n_waves, int32_t(mm), int32_t(mn), steps});
        }
        // Several chunks run on two streams.  A chunk is latency-bound when
        // its longest wave (a wave runs its strips one after another) outlasts
        // its wave-steps spread over the device's 1024 SIMDs: measured, a lone
        // wave's step takes about 72 ns and a full device retires a wave-step
        // per SIMD about every 44 ns.  Latency-bound chunks (the few longest
        // texts) take `st`, the others the second stream, so the device is not
        // left idle behind a few long pairs; chunks all of one kind alternate,
        // which overlaps each chunk's tail with the next.
        const int nsq = wchunks.size() > 1 ? 2 : 1;
        std::vector<int> ch_stream(wchunks.size(), 0);
        int64_t dir_need[2] = {0, 0}, sel_need[2] = {0, 0}, cw_need[2] = {0, 0};
        {
            std::vector<uint8_t> lat_bound(wchunks.size());
            size_t n_lat = 0;
            for (size_t c = 0; c < wchunks.size(); ++c) {
                const WChunk& ch = wchunks[c];
                const double lone = 72.0 * double(ch.cm + 64) * double((ch.cn + 63) / 64);
                const double full = 44.0 * ch.steps / 1024.0;
                lat_bound[c] = lone > full;
                n_lat += lat_bound[c];
            }
            const bool mixed = n_lat > 0 && n_lat < wchunks.size();
            for (size_t c = 0; c < wchunks.size(); ++c) {
                const WChunk& ch = wchunks[c];
                const int q = nsq == 1 ? 0 : mixed ? (lat_bound[c] ? 0 : 1) : int(c & 1);
                ch_stream[c] = q;
                const int64_t cw = ch.w1 - ch.w0;
                if (trace) dir_need[q] = std::max(dir_need[q], dir_words(ch.cm, ch.cn) * cw);
                sel_need[q] = std::max(sel_need[q], cw * ga::sw16_sel_stride(ch.cm));
                cw_need[q] = std::max(cw_need[q], cw);
            }
        }
        uint32_t* d_dirs16[2] = {nullptr, nullptr};
        uint32_t* d_sel[2] = {nullptr, nullptr};
        int32_t* d_blist[2] = {nullptr, nullptr};        // the DNA kernel's literal-byte wave lists
        for (int q = 0; q < nsq; ++q) {
            if (trace) d_dirs16[q] = D.alloc<uint32_t>(size_t(dir_need[q]));
            d_sel[q] = D.alloc<uint32_t>(size_t(sel_need[q]));
            if (!mat) d_blist[q] = D.alloc<int32_t>(size_t(cw_need[q] + 1));
        }
        if (!D.ok) err = "hipMalloc(dir table) -- batch workspace exceeds device memory; split the batch";
        // the second stream starts behind the uploads and `st` waits for it
        // before the results are copied back; with two streams the kernel
        // timer takes one interval around both
        hipStream_t sq[2] = {st, nullptr};
        hipEvent_t fork = nullptr, join = nullptr;
        if (!err && nsq > 1) {
            sq[1] = acquire_stream();
            if (!sq[1] || hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess ||
                hipEventRecord(fork, st) != hipSuccess || hipStreamWaitEvent(sq[1], fork, 0) != hipSuccess)
                err = "hipStreamWaitEvent";
        }
        ga::LaunchHooks* hooks = nsq > 1 ? nullptr : &t_timer_hooks;
        if (!err && nsq > 1) t_timer.begin(st);
        for (size_t c = 0; c < wchunks.size() && !err; ++c) {
            const WChunk& ch = wchunks[c];
            const int q = ch_stream[c];
            ga::Sw16Launch L{};
            L.txt = d_txt; L.toff = d_toff; L.tlen = d_tl;
            L.pat = d_pat; L.poff = d_poff; L.plen = d_pl;
            L.first_pair = int32_t(n32 + 2 * ch.w0); L.n_total = int32_t(n);
            L.n_waves = int32_t(ch.w1 - ch.w0); L.max_m = ch.cm; L.max_n = ch.cn;
            L.match = sc.match; L.mismatch = sc.mismatch; L.go = sc.gap_open; L.ge = sc.gap_extend;
            L.res = d_res; L.rerun = d_rerun;
            L.dirs = d_dirs16[q]; L.dir_stride = trace ? dir_words(ch.cm, ch.cn) : 0;
            L.selbuf = d_sel[q];
            L.bytes_list = d_blist[q];
            L.walk = d_walk; L.cigar = d_cig; L.cig_off = d_slot;
            if (mat) {
                L.go = mat->gap_open; L.ge = mat->gap_extend;
                L.tcode = d_tcode; L.pcode = d_pcode; L.sub = d_sub; L.bias = mbias;
            }
            if (enqueue_sw16_chunk(L, trace, sq[q], hooks, mat) != hipSuccess) err = "SW kernel launch";
        }
        if (sq[1]) {
            // joined even after an error, so `st` never runs ahead of it
            if ((hipEventRecord(join, sq[1]) != hipSuccess || hipStreamWaitEvent(st, join, 0) != hipSuccess) && !err)
                err = "hipStreamWaitEvent";
            if (err) (void)hipStreamSynchronize(sq[1]);
        }
        if (nsq > 1) t_timer.end(st);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (sq[1]) release_stream(sq[1]);
    }

    // the CIGAR checks on the device (trace batches), once every walk
    // has written its CIGAR
    uint8_t* d_vf = nullptr;
    if (vflags && trace && !err) {
        d_vf = D.alloc<uint8_t>(size_t(n));
        ga::SwValidate V{};
        V.txt = d_txt; V.toff = d_toff; V.tlen = d_tl;
        V.pat = d_pat; V.poff = d_poff; V.plen = d_pl;
        V.cigar = d_cig; V.cig_off = d_slot;
        V.score = d_res; V.ei = d_res + 1; V.ej = d_res + 2; V.si = d_res + 5; V.sj = d_res + 6; V.stride = 8;
        V.len = d_res + 7; V.len_stride = 8;
        V.walk = d_walk; V.flags = d_vf; V.n = n;
        V.match = sc.match; V.mismatch = sc.mismatch; V.go = sc.gap_open; V.ge = sc.gap_extend;
        if (mat) {
            V.go = mat->gap_open; V.ge = mat->gap_extend;
            V.tcode = d_tcode; V.pcode = d_pcode; V.sub = d_sub;
        }
        if (!D.ok || ga::sw_validate_launch(V, st) != hipSuccess) err = "SW validate launch";
    }
    // results, flags and CIGAR slots back into pinned staging (pageable
    // fallback: the caller's arena and vectors)
    std::vector<int32_t> r8v;
    int32_t* r8p = reinterpret_cast<int32_t*>(pin.get(size_t(n) * 32));
    if (!r8p) { r8v.assign(static_cast<size_t>(n) * 8, -1); r8p = r8v.data(); }
    std::vector<uint8_t> vfv;
    uint8_t* vfp = d_vf ? pin.get(size_t(n)) : nullptr;
    if (d_vf && !vfp) { vfv.assign(static_cast<size_t>(n), 0); vfp = vfv.data(); }
    char* cdst = arena;
    // (with the device's checks) the CIGARs come back packed: the slots
    // are sized m + n per pair, the CIGARs a fraction of that
    const bool compact = d_vf && csrc_off && slot > 0;
    std::vector<int64_t> coff;                            // compact offsets, launch order
    if (!compact && trace && slot > 0)
        if (uint8_t* c = pin.get(size_t(slot))) cdst = reinterpret_cast<char*>(c);
    if (!err) {
        if (hipMemcpyAsync(r8p, d_res, size_t(n) * 32, hipMemcpyDeviceToHost, st) != hipSuccess ||
            (d_vf && hipMemcpyAsync(vfp, d_vf, size_t(n), hipMemcpyDeviceToHost, st) != hipSuccess) ||
            (!compact && trace && slot > 0 &&
             hipMemcpyAsync(cdst, d_cig, size_t(slot), hipMemcpyDeviceToHost, st) != hipSuccess))
            err = "hipMemcpy(results)";
    }
    if (compact && !err) {
        // the lengths from the results' host copy, the offsets uploaded, one
        // kernel packs the CIGARs, one short copy brings them back
        const hipError_t e1 = hipStreamSynchronize(st);
        if (e1 != hipSuccess) err = hipGetErrorString(e1);
        if (!err) {
            coff.resize(size_t(n) + 1);
            int64_t tot = 0;
            for (int64_t k = 0; k < n; ++k) {
                coff[size_t(k)] = tot;
                const int32_t* q = r8p + k * 8;
                if (walk_p[size_t(k)] && q[0] > 0 && q[7] > 0) tot += q[7];
            }
            coff[size_t(n)] = tot;
            uint8_t* cc = tot > 0 ? pin.get(size_t(tot)) : nullptr;
            if (tot > 0 && !cc) {
                // no pinned memory: the slots straight into the caller's arena
                coff.clear();
                if (hipMemcpyAsync(arena, d_cig, size_t(slot), hipMemcpyDeviceToHost, st) != hipSuccess)
                    err = "hipMemcpy(CIGARs)";
            } else if (tot > 0) {
                const int64_t* d_coff = D.upload(coff);
                uint8_t* d_cc = D.alloc<uint8_t>(size_t(tot));
                if (!D.ok || ga::sw_cigar_compact_launch(d_cig, d_slot, d_tl, d_pl, d_coff, n, d_cc, st) != hipSuccess ||
                    hipMemcpyAsync(cc, d_cc, size_t(tot), hipMemcpyDeviceToHost, st) != hipSuccess)
                    err = "hipMemcpy(CIGARs)";
                cdst = reinterpret_cast<char*>(cc);
            }
        }
    }
    D.release();
    const hipError_t se = hipStreamSynchronize(st);
    if (!err && se != hipSuccess) err = hipGetErrorString(se);
    release_stream(st);
    *cig_src = cdst;
    res8.resize(static_cast<size_t>(n) * 8);
    if (d_vf) vflags->resize(static_cast<size_t>(n));
    if (csrc_off) csrc_off->assign(coff.empty() ? 0 : static_cast<size_t>(n), 0);
    parallel_for(n, [&](int64_t k) {
        std::copy(r8p + k * 8, r8p + k * 8 + 8, res8.begin() + order[size_t(k)] * 8);
        if (d_vf) (*vflags)[size_t(order[size_t(k)])] = vfp[k];
        if (!coff.empty()) (*csrc_off)[size_t(order[size_t(k)])] = coff[size_t(k)];
    });
    if (vflags && !d_vf) vflags->clear();
    return err;
}

void reset_sw(ga_sw_res* res, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) res[i] = ga_sw_res{-1, 0, 0, 0, -1, -1, -1, -1, 0, 0, 0};
}

// Upper distance bound of each router bucket; the last bucket means "above
// the pair's smax" (unresolved).
constexpr int32_t kBucketBound[16] = {0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048, INT32_MAX};

const char* check_dev_batch(const ga_wfa_dev_batch* b)
{
    if (!b) return "null batch";
    if (b->smax_max < 0 || b->smax_max > GA_MAX_SMAX_EXT) return "smax_max out of range [0, 4000]";
    if (b->n_pairs > INT32_MAX) return "too many pairs";
    return nullptr;
}

}  // namespace

// A device-planned WFA batch whose every scratch buffer is carved from one
// caller-given device region of plan_bytes + stack_bytes (no allocation):
// the claim order, plan histogram, class slices and claim counters from the
// first plan_bytes, and the wavefront stacks of each class from the rest,
// which every class reuses in turn -- the classes run one after another on
// `st` (no forked streams, no events).  The caller has checked the batch and
// that it fits.
const char* ga::wfa_device_in_region(const ga_wfa_dev_batch* b, hipStream_t st, uint8_t* region,
                                     size_t plan_bytes, size_t stack_bytes)
{
    ga::WfaQueueLaunch L{};
    L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
    L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
    L.smax = b->smax; L.want = b->want_cigar;
    L.score = b->score; L.cigar = b->cigar; L.cig_off = b->cigar_off; L.cig_len = b->cigar_len;
    L.wide = b->max_len > 32766;
    size_t used = 0;
    auto carve = [&](size_t bytes) -> void* {
        const size_t at = (used + 255) & ~size_t(255);
        if (at + bytes > plan_bytes) return nullptr;
        used = at + bytes;
        return region + at;
    };
    std::function<void*(size_t)> alloc = carve;
    std::function<void*(size_t)> stacks = [&](size_t bytes) -> void* {
        return bytes <= stack_bytes ? region + plan_bytes : nullptr;
    };
    WfaDevClasses C;
    if (const char* e = setup_wfa_dev_classes(L, b->n_pairs, b->smax_max, b->max_len, b->cigar != nullptr, C,
                                              alloc, int64_t(stack_bytes), &stacks))
        return e;
    if (ga::wfa_plan_launch(C.plan, st) != hipSuccess) return "WFA plan launch";
    for (const auto& cl : C.L)
        if (ga::wfa_queue_launch(cl, st) != hipSuccess) return "WFA kernel launch";
    return nullptr;
}

void ga::host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& chunk)
{
    if (n <= 0) return;
    if (n <= grain || WorkPool::get().workers() == 0) {
        chunk(0, n);
        return;
    }
    WorkPool::get().run(n, grain, chunk);
}

extern "C" {

int64_t ga_cigar_arena_bytes_wfa(const ga_align_req* reqs, int64_t n)
{
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) s += std::max(reqs[i].pattern_len, 0) + std::max(reqs[i].text_len, 0);
    return s;
}

int64_t ga_cigar_arena_bytes_sw(const ga_sw_req* reqs, int64_t n)
{
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) s += std::max(reqs[i].pattern_len, 0) + std::max(reqs[i].text_len, 0);
    return s;
}

int ga_align_batch(const ga_align_req* reqs, int64_t n, ga_align_res* res, char* arena,
                   int64_t arena_bytes, int32_t* resolved_count, int32_t* unresolved_count,
                   const char** error)
{
    return wfa_common(reqs, n, res, arena, arena_bytes, resolved_count, unresolved_count, error, false);
}

int ga_align_batch_ext(const ga_align_req* reqs, int64_t n, ga_align_res* res, char* arena,
                       int64_t arena_bytes, int32_t* resolved_count, int32_t* unresolved_count,
                       const char** error)
{
    return wfa_common(reqs, n, res, arena, arena_bytes, resolved_count, unresolved_count, error, true);
}

static int sw_common(const ga_sw_req* reqs, int64_t n, ga_sw_res* res, char* arena, int64_t arena_bytes,
                     int32_t* resolved_count, int32_t* unresolved_count, const char** error,
                     const SwMatrixHost* mat)
{
    if (error) *error = nullptr;
    if (resolved_count) *resolved_count = 0;
    if (unresolved_count) *unresolved_count = 0;
    if (n <= 0) return GA_OK;
    reset_sw(res, n);
    auto refuse = [&](const char* msg) {
        if (error) *error = msg;
        reset_sw(res, n);
        return GA_INVALID_ARGUMENT;
    };
    if (!reqs || !res) return refuse("null request or result array");
    const ga_sw_scoring s0 = reqs[0].scoring;
    for (int64_t i = 0; i < n; ++i) {
        const ga_sw_req& r = reqs[i];
        if ((r.text_len > 0 && r.text == nullptr) || (r.pattern_len > 0 && r.pattern == nullptr))
            return refuse("null pointer with non-zero length");
        if (r.text_len < 0 || r.pattern_len < 0) return refuse("negative length");
        if (mat) {
            if (r.with_cigar && mat->gap_extend > mat->gap_open)
                return refuse("SW matrix scoring: gap_extend > gap_open with with_cigar cannot re-score "
                              "to its own score; score-only requests may use it");
            continue;                             // the matrix scheme replaces r.scoring
        }
        if (const char* e = sw_scheme_error(r.scoring, r.with_cigar != 0)) return refuse(e);
        if (r.scoring.match != s0.match || r.scoring.mismatch != s0.mismatch ||
            r.scoring.gap_open != s0.gap_open || r.scoring.gap_extend != s0.gap_extend)
            return refuse("all requests in an SW batch must share one scoring scheme");
    }
    bool any_cigar = false;
    int64_t max_cells = 0;
    for (int64_t i = 0; i < n; ++i) {
        any_cigar |= reqs[i].with_cigar != 0;
        max_cells = std::max(max_cells, (int64_t(reqs[i].text_len) + 1) * (int64_t(reqs[i].pattern_len) + 1));
    }
    const int64_t need = ga_cigar_arena_bytes_sw(reqs, n);
    if (any_cigar && need > 0 && (arena == nullptr || arena_bytes < need))
        return refuse("CIGAR arena smaller than the batch needs (ga_cigar_arena_bytes_sw)");
    // api.cpp:644-645 and sw_kernel_impl.hip:276-280: a pair whose matrix exceeds
    // the batch's direction stride is too_large.
    const int64_t dir_stride = std::min<int64_t>(max_cells, GA_SW_MAX_TRACE_CELLS);
    std::vector<uint8_t> skip(static_cast<size_t>(n), 0), walk(static_cast<size_t>(n), 0);
    for (int64_t i = 0; i < n; ++i) {
        const ga_sw_req& r = reqs[i];
        if (any_cigar && r.text_len > 0 && r.pattern_len > 0 &&
            (int64_t(r.text_len) + 1) * (int64_t(r.pattern_len) + 1) > dir_stride)
            skip[size_t(i)] = 1;
        walk[size_t(i)] = r.with_cigar != 0;
    }
    std::vector<int32_t> r8;
    std::vector<int64_t> slot_off;
    PinnedHold pin;
    const char* cig_src = arena;
    std::vector<uint8_t> vflags;                  // the device's CIGAR checks (DNA batches)
    std::vector<int64_t> csrc;                    // where each CIGAR sits in cig_src (packed), if so
    const char* err = run_sw(reqs, n, skip, any_cigar, walk, arena, r8, slot_off, pin, &cig_src, mat, &vflags,
                             &csrc);
    const bool dev_checks = !vflags.empty();
    if (err) {
        if (error) *error = err;
        reset_sw(res, n);
        return GA_DEVICE_ERROR;
    }
    parallel_for(n, [&](int64_t i) {
        const ga_sw_req& r = reqs[i];
        ga_sw_res& o = res[i];
        const int32_t* q = &r8[size_t(i) * 8];
        o.cigar_off = slot_off[size_t(i)];
        if (skip[size_t(i)]) { o.too_large = 1; return; }
        if (q[0] < 0) return;
        o.score = q[0];
        o.resolved = 1;
        if (!any_cigar) {                         // score-only batch (api.cpp:829-838)
            o.end_i = q[3]; o.end_j = q[4];
            if (o.score == 0) o.rescore_ok = o.wellformed_ok = 1;
            return;
        }
        o.end_i = q[1]; o.end_j = q[2];           // trace fill (api.cpp:542-580)
        if (o.score == 0) { o.rescore_ok = o.wellformed_ok = 1; return; }
        if (!r.with_cigar) return;
        o.start_i = q[5]; o.start_j = q[6];
        const int len = q[7];
        if (len > 0) {
            const int slot_len = std::max(r.text_len, 0) + std::max(r.pattern_len, 0);
            o.cigar_off = slot_off[size_t(i)] + slot_len - len;
            o.cigar_len = len;
            if (cig_src != arena)
                std::memcpy(arena + o.cigar_off, cig_src + (csrc.empty() ? o.cigar_off : csrc[size_t(i)]), size_t(len));
            // the device's checks (every CIGAR batch runs them; without
            // their flags both stay 0)
            if (dev_checks) {
                o.rescore_ok = vflags[size_t(i)] & 1;
                o.wellformed_ok = (vflags[size_t(i)] >> 1) & 1;
            }
        }
    });
    int32_t rc = 0;
    for (int64_t i = 0; i < n; ++i) rc += res[i].resolved ? 1 : 0;
    if (resolved_count) *resolved_count = rc;
    if (unresolved_count) *unresolved_count = int32_t(n) - rc;
    return GA_OK;
}

int ga_sw_validate_cigar(const char* cigar, int32_t cigar_len, const char* text, int32_t text_len,
                         int32_t start_i, int32_t end_i, const char* pattern, int32_t pattern_len,
                         int32_t start_j, int32_t end_j, const ga_sw_scoring* scoring, int32_t score,
                         int32_t* rescore_ok, int32_t* wellformed_ok)
{
    if (!scoring || !rescore_ok || !wellformed_ok || cigar_len < 0 || text_len < 0 || pattern_len < 0 ||
        (cigar_len > 0 && !cigar) || (text_len > 0 && !text) || (pattern_len > 0 && !pattern))
        return GA_INVALID_ARGUMENT;
    sw_validate(cigar, cigar_len, text, text_len, start_i, end_i, pattern, pattern_len, start_j, end_j, *scoring,
                score, rescore_ok, wellformed_ok);
    return GA_OK;
}

int ga_align_sw_batch(const ga_sw_req* reqs, int64_t n, ga_sw_res* res, char* arena,
                      int64_t arena_bytes, int32_t* resolved_count, int32_t* unresolved_count,
                      const char** error)
{
    return sw_common(reqs, n, res, arena, arena_bytes, resolved_count, unresolved_count, error, nullptr);
}

int ga_align_sw_matrix_batch(const ga_sw_req* reqs, int64_t n, const ga_sw_matrix* matrix, ga_sw_res* res,
                             char* arena, int64_t arena_bytes, int32_t* resolved_count,
                             int32_t* unresolved_count, const char** error)
{
    if (error) *error = nullptr;
    if (n > 0 && res) reset_sw(res, n);
    if (!matrix) { if (error) *error = "null substitution matrix"; return GA_INVALID_ARGUMENT; }
    SwMatrixHost mt{};
    const int k = int(strnlen(matrix->alphabet, sizeof(matrix->alphabet)));
    if (k < 1 || k > 31) { if (error) *error = "SW matrix: alphabet must hold 1..31 letters"; return GA_INVALID_ARGUMENT; }
    if (matrix->gap_open < 0 || matrix->gap_extend < 0) {
        if (error) *error = "SW matrix: gap costs must be >= 0";
        return GA_INVALID_ARGUMENT;
    }
    std::memset(mt.code, kSwUnknownCode, sizeof(mt.code));
    for (int a = 0; a < k; ++a) {
        const uint8_t c = uint8_t(matrix->alphabet[a]);
        if (mt.code[c] != kSwUnknownCode) { if (error) *error = "SW matrix: repeated alphabet letter"; return GA_INVALID_ARGUMENT; }
        mt.code[c] = uint8_t(a);
    }
    for (int a = 0; a < 32; ++a)
        for (int b = 0; b < 32; ++b)
            mt.sub[a * 32 + b] = (a < k && b < k) ? int32_t(matrix->score[a][b]) : matrix->unknown;
    mt.gap_open = matrix->gap_open;
    mt.gap_extend = matrix->gap_extend;
    return sw_common(reqs, n, res, arena, arena_bytes, resolved_count, unresolved_count, error, &mt);
}

void ga_sw_matrix_blosum62(ga_sw_matrix* out, int32_t gap_open, int32_t gap_extend)
{
    // BLOSUM62 (Henikoff & Henikoff 1992) as the reference holds it
    // (src/msa/msa_ref.cpp:116-137), rows and columns in the order below; the
    // reference's row M, column H reads -1 where the published matrix has -2,
    // and that value is kept so results match the reference's table.
    static const char kOrder[] = "ARNDCQEGHILKMFPSTWYV";
    static const int8_t kB62[20][20] = {
        { 4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0},
        {-1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3},
        {-2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3},
        {-2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3},
        { 0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1},
        {-1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2},
        {-1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2},
        { 0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3},
        {-2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3},
        {-1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3},
        {-1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1},
        {-1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2},
        {-1,-1,-2,-3,-1, 0,-2,-3,-1, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1},
        {-2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1},
        {-1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2},
        { 1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2},
        { 0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0},
        {-3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3},
        {-2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1},
        { 0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4},
    };
    if (!out) return;
    std::memset(out, 0, sizeof(*out));
    std::memcpy(out->alphabet, kOrder, 20);
    for (int a = 0; a < 20; ++a)
        for (int b = 0; b < 20; ++b) out->score[a][b] = kB62[a][b];
    out->unknown = -1;
    out->gap_open = gap_open;
    out->gap_extend = gap_extend;
}

void ga_set_kernel_timing(int enabled) { t_timer.enabled = enabled != 0; }

float ga_last_kernel_ms(void) { return t_timer.total_ms(); }

const char* ga_backend_name(void) { return "rocm"; }

int ga_device_available(void)
{
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
}

const char* ga_device_name(void)
{
    static const std::string name = [] {
        hipDeviceProp_t p{};
        if (hipGetDeviceProperties(&p, 0) != hipSuccess) return std::string("none");
        return std::string(p.name);
    }();
    return name.c_str();
}

// ---- device-resident batches ------------------------------------------------
int ga_wfa_device(const ga_wfa_dev_batch* b, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (const char* e = check_dev_batch(b)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (b->n_pairs <= 0) return GA_OK;
    std::call_once(g_pool_once, init_pool);
    hipStream_t st = static_cast<hipStream_t>(stream);
    ga::WfaQueueLaunch L{};
    L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
    L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
    L.smax = b->smax; L.want = b->want_cigar;
    L.score = b->score; L.cigar = b->cigar; L.cig_off = b->cigar_off; L.cig_len = b->cigar_len;
    L.wide = b->max_len > 32766;
    t_timer.reset();
    const char* err = nullptr;
    if (b->host_smax && b->host_pattern_len && b->host_text_len) {
        // host plan: exact class maxima, sizes known before launch; a batch
        // whose every smax is in the smallest class needs no pass over its
        // pairs (one class in index order, bounded by smax_max and max_len)
        WfaPlan plan;
        if (b->smax_max <= kWfaClassBound[0]) {
            plan.bound = {b->smax_max};
            plan.max_plen = {b->max_len};
            plan.max_tlen = {b->max_len};
            plan.first = {0};
            plan.count = {b->n_pairs};
        } else {
            plan = plan_wfa(b->n_pairs, b->host_smax, b->host_pattern_len, b->host_text_len);
        }
        err = enqueue_wfa(L, plan, b->cigar != nullptr, st, &t_timer);
    } else {
        // device plan: classes and claim order computed on the GPU
        std::vector<void*> held;
        auto alloc = [&](size_t bytes) -> void* {
            void* p = nullptr;
            if (ws_alloc(&p, bytes, st) != hipSuccess) return nullptr;
            held.push_back(p);
            return p;
        };
        WfaDevClasses C;
        err = setup_wfa_dev_classes(L, b->n_pairs, b->smax_max, b->max_len, b->cigar != nullptr, C, alloc);
        std::vector<hipStream_t> side;
        std::vector<hipEvent_t> ev(ga::kWfaClasses, nullptr);
        for (size_t c = 1; !err && c < C.L.size(); ++c) {
            hipStream_t cs = acquire_stream();
            if (!cs) err = "hipStreamCreate";
            else side.push_back(cs);
        }
        for (auto& e : ev)
            if (!err && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) err = "hipEventCreate";
        if (!err) {
            t_timer.begin(st);
            err = enqueue_wfa_dev(C, st, side, ev);
            t_timer.end(st);
        }
        for (void* p : held) (void)ws_free(p, st);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t cs : side) release_stream(cs);
    }
    if (err) {
        if (error) *error = err;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

// ------------------------------------------------------------------ graphs
struct ga_wfa_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    std::vector<void*> bufs;
};

void ga_wfa_graph_destroy(ga_wfa_graph* g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto e : g->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto s : g->side)
        if (s) (void)hipStreamDestroy(s);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    for (void* p : g->bufs) (void)hipFree(p);
    delete g;
}

int ga_wfa_graph_create(const ga_wfa_dev_batch* b, ga_wfa_graph** out, const char** error)
{
    if (error) *error = nullptr;
    if (!out) { if (error) *error = "null output"; return GA_INVALID_ARGUMENT; }
    *out = nullptr;
    if (const char* e = check_dev_batch(b)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (b->n_pairs <= 0) { if (error) *error = "empty batch"; return GA_INVALID_ARGUMENT; }
    auto* g = new ga_wfa_graph;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        g->bufs.push_back(p);
        return p;
    };
    ga::WfaQueueLaunch L{};
    L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
    L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
    L.smax = b->smax; L.want = b->want_cigar;
    L.score = b->score; L.cigar = b->cigar; L.cig_off = b->cigar_off; L.cig_len = b->cigar_len;
    L.wide = b->max_len > 32766;
    WfaDevClasses C;
    const char* err = setup_wfa_dev_classes(L, b->n_pairs, b->smax_max, b->max_len, b->cigar != nullptr, C, alloc);
    if (!err && hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking) != hipSuccess) err = "hipStreamCreate";
    for (size_t c = 1; !err && c < C.L.size(); ++c) {
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) err = "hipStreamCreate";
        else g->side.push_back(cs);
    }
    g->ev.assign(ga::kWfaClasses, nullptr);
    for (auto& e : g->ev)
        if (!err && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) err = "hipEventCreate";
    if (!err) {
        if (hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            err = "hipStreamBeginCapture";
        } else {
            err = enqueue_wfa_dev(C, g->cap, g->side, g->ev);
            const hipError_t ce = hipStreamEndCapture(g->cap, &g->graph);
            if (!err && ce != hipSuccess) err = "hipStreamEndCapture";
        }
    }
    if (!err && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        err = "hipGraphInstantiate";
    if (err) {
        ga_wfa_graph_destroy(g);
        if (error) *error = err;
        return GA_DEVICE_ERROR;
    }
    *out = g;
    return GA_OK;
}

int ga_wfa_graph_launch(ga_wfa_graph* g, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (!g || !g->exec) { if (error) *error = "null graph"; return GA_INVALID_ARGUMENT; }
    if (hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "hipGraphLaunch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_sw_validate_device(const ga_sw_dev_batch* b, uint8_t* flags, void* stream, const char** error)
{
    if (error) *error = nullptr;
    auto refuse = [&](const char* m) { if (error) *error = m; return GA_INVALID_ARGUMENT; };
    if (!b) return refuse("null batch");
    if (b->n_pairs <= 0) return GA_OK;
    if (!b->with_cigar) return refuse("the batch holds no CIGARs (with_cigar 0)");
    if (const char* e = sw_scheme_error(b->scoring, true)) return refuse(e);
    if (!flags || !b->out || !b->cigar || !b->cigar_off || !b->cigar_len || !b->text || !b->text_off ||
        !b->text_len || !b->pattern || !b->pattern_off || !b->pattern_len)
        return refuse("null device pointer");
    ga::SwValidate V{};
    V.txt = b->text; V.toff = b->text_off; V.tlen = b->text_len;
    V.pat = b->pattern; V.poff = b->pattern_off; V.plen = b->pattern_len;
    V.cigar = b->cigar; V.cig_off = b->cigar_off;
    V.score = b->out; V.si = b->out + 1; V.sj = b->out + 2; V.ei = b->out + 3; V.ej = b->out + 4; V.stride = 5;
    V.len = b->cigar_len; V.len_stride = 1;
    V.walk = nullptr; V.flags = flags; V.n = b->n_pairs;
    V.match = b->scoring.match; V.mismatch = b->scoring.mismatch;
    V.go = b->scoring.gap_open; V.ge = b->scoring.gap_extend;
    if (ga::sw_validate_launch(V, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "SW validate launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_sw_device(const ga_sw_dev_batch* b, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (!b) { if (error) *error = "null batch"; return GA_INVALID_ARGUMENT; }
    if (b->n_pairs <= 0) return GA_OK;
    if (b->n_pairs > INT32_MAX) { if (error) *error = "too many pairs"; return GA_INVALID_ARGUMENT; }
    if (const char* e = sw_scheme_error(b->scoring, b->with_cigar != 0)) {
        if (error) *error = e;
        return GA_INVALID_ARGUMENT;
    }
    const bool trace = b->with_cigar != 0;
    if (!b->out || (trace && (!b->cigar || !b->cigar_off))) {
        if (error) *error = "null output buffer";
        return GA_INVALID_ARGUMENT;
    }
    std::call_once(g_pool_once, init_pool);
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<void*> owned;
    auto dalloc = [&](void** p, size_t bytes) {
        if (ws_alloc(p, std::max<size_t>(bytes, 64), st) != hipSuccess) return false;
        owned.push_back(*p);
        return true;
    };
    auto release = [&] { for (void* p : owned) (void)ws_free(p, st); };
    int32_t* d_res = nullptr;
    if (b->max_text_len <= ga::kSw16MaxRows && (!trace || b->max_pattern_len <= ga::kSw16MaxCols)) {
        // packed path: pairs (2w, 2w+1) share a wave
        const int64_t n_waves = (b->n_pairs + 1) / 2;
        const int64_t stride = trace ? ga::sw16_dir_words_per_wave(b->max_text_len, b->max_pattern_len) : 0;
        int64_t cw = n_waves;
        if (trace && stride > 0) cw = std::max<int64_t>(1, std::min<int64_t>(cw, sw_dir_budget() / (stride * 4)));
        uint8_t* d_rerun = nullptr;
        uint32_t* d_dirs = nullptr;
        bool ok = dalloc(reinterpret_cast<void**>(&d_res), size_t(b->n_pairs) * 32) &&
                  dalloc(reinterpret_cast<void**>(&d_rerun), size_t(b->n_pairs));
        if (ok && trace) ok = dalloc(reinterpret_cast<void**>(&d_dirs), size_t(cw * stride) * 4);
        uint32_t* d_sel = nullptr;
        if (ok) ok = dalloc(reinterpret_cast<void**>(&d_sel), size_t(cw * ga::sw16_sel_stride(b->max_text_len)) * 4);
        int32_t* d_blist = nullptr;                        // the literal-byte pass's wave list
        if (ok) ok = dalloc(reinterpret_cast<void**>(&d_blist), size_t(cw + 1) * 4);
        int rc = GA_OK;
        if (!ok) { if (error) *error = "hipMalloc(SW workspace)"; rc = GA_DEVICE_ERROR; }
        t_timer.reset();
        for (int64_t w0 = 0; w0 < n_waves && rc == GA_OK; w0 += cw) {
            ga::Sw16Launch L{};
            L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
            L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
            L.first_pair = int32_t(2 * w0); L.n_total = int32_t(b->n_pairs);
            L.n_waves = int32_t(std::min(cw, n_waves - w0)); L.max_m = b->max_text_len;
            L.max_n = b->max_pattern_len;
            L.match = b->scoring.match; L.mismatch = b->scoring.mismatch;
            L.go = b->scoring.gap_open; L.ge = b->scoring.gap_extend;
            L.res = d_res; L.rerun = d_rerun; L.dirs = d_dirs; L.dir_stride = stride; L.selbuf = d_sel;
            L.bytes_list = d_blist;
            L.walk = nullptr; L.cigar = b->cigar; L.cig_off = b->cigar_off;
            if (enqueue_sw16_chunk(L, trace, st, &t_timer_hooks) != hipSuccess) {
                if (error) *error = "SW kernel launch";
                rc = GA_DEVICE_ERROR;
            }
        }
        if (rc == GA_OK && ga::sw_gather(d_res, b->n_pairs, trace, b->out, b->cigar_len, st) != hipSuccess) {
            if (error) *error = "SW gather launch";
            rc = GA_DEVICE_ERROR;
        }
        release();
        return rc;
    }
    // Every pair gets a direction table of the batch's largest shape.
    const int64_t per = ga::sw_dir_words(b->max_text_len, b->max_pattern_len);
    const bool lds_bnd = b->max_text_len <= kLdsBndMaxRows;
    int64_t chunk = b->n_pairs;
    if (trace) chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, sw_dir_budget() / std::max<int64_t>(per * 4, 1)));
    if (!lds_bnd) chunk = std::min<int64_t>(chunk, int64_t(1) << 16);
    uint32_t* d_dirs = nullptr;
    int32_t* d_bnd = nullptr;
    bool ok = dalloc(reinterpret_cast<void**>(&d_res), size_t(b->n_pairs) * 32);
    if (ok && trace) ok = dalloc(reinterpret_cast<void**>(&d_dirs), size_t(chunk * per) * 4);
    if (ok && !lds_bnd)
        ok = dalloc(reinterpret_cast<void**>(&d_bnd), size_t(chunk) * 2 * size_t(b->max_text_len) * 4);
    int rc = GA_OK;
    if (!ok) { if (error) *error = "hipMalloc(SW workspace)"; rc = GA_DEVICE_ERROR; }
    t_timer.reset();
    for (int64_t a = 0; a < b->n_pairs && rc == GA_OK; a += chunk) {
        ga::SwLaunch L{};
        L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
        L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
        L.n = int32_t(std::min(chunk, b->n_pairs - a)); L.max_m = b->max_text_len;
        L.match = b->scoring.match; L.mismatch = b->scoring.mismatch;
        L.go = b->scoring.gap_open; L.ge = b->scoring.gap_extend;
        L.res = d_res; L.dirs = d_dirs; L.dir_off = nullptr; L.dir_stride = per;
        L.walk = nullptr; L.cigar = b->cigar; L.cig_off = b->cigar_off; L.bnd = d_bnd;
        L.first_pair = int32_t(a); L.n_total = int32_t(b->n_pairs);
        if (ga::sw_launch(L, trace, st, &t_timer_hooks) != hipSuccess) {
            if (error) *error = "SW kernel launch";
            rc = GA_DEVICE_ERROR;
        }
    }
    if (rc == GA_OK && ga::sw_gather(d_res, b->n_pairs, trace, b->out, b->cigar_len, st) != hipSuccess) {
        if (error) *error = "SW gather launch";
        rc = GA_DEVICE_ERROR;
    }
    release();
    return rc;
}

// ------------------------------------------------------------------ router
int ga_sketch_device(const ga_wfa_dev_batch* b, float* feat, int32_t* lower_bound, void* stream,
                     const char** error)
{
    if (error) *error = nullptr;
    if (const char* e = check_dev_batch(b)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (b->n_pairs <= 0) return GA_OK;
    if (!feat || !lower_bound) { if (error) *error = "null output"; return GA_INVALID_ARGUMENT; }
    ga::SketchLaunch L{b->pattern, b->pattern_off, b->pattern_len, b->text, b->text_off, b->text_len,
                       b->n_pairs, feat, lower_bound};
    if (ga::sketch_launch(L, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "sketch kernel launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_mlp_forward_device(const float* params, const float* feat, int64_t n, float* h1, float* h2,
                          float* logits, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (n < 0 || !params || (n > 0 && (!feat || !h1 || !h2 || !logits))) {
        if (error) *error = "invalid MLP arguments";
        return GA_INVALID_ARGUMENT;
    }
    if (ga::mlp_forward(params, feat, n, h1, h2, logits, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "MLP forward launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_mlp_logits_device(const float* params, const float* feat, int64_t n, float* logits, void* stream,
                         const char** error)
{
    if (error) *error = nullptr;
    if (n < 0 || !params || (n > 0 && (!feat || !logits))) {
        if (error) *error = "invalid MLP arguments";
        return GA_INVALID_ARGUMENT;
    }
    if (reinterpret_cast<uintptr_t>(params) & 15u) {
        if (error) *error = "MLP parameters not 16-byte aligned";
        return GA_INVALID_ARGUMENT;
    }
    if (ga::mlp_logits(params, feat, n, logits, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "MLP logits launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_mlp_backward_device(const float* params, const float* feat, const float* h1, const float* h2,
                           const float* logits, const int32_t* label, int64_t n, float* grads, float* loss,
                           float* work, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (n <= 0 || !params || !feat || !h1 || !h2 || !logits || !label || !grads || !loss || !work) {
        if (error) *error = "invalid MLP arguments";
        return GA_INVALID_ARGUMENT;
    }
    float* dL = work;
    float* dH2 = dL + n * ga::kBuckets;
    float* dH1 = dH2 + n * ga::kHidden;
    if (ga::mlp_backward(params, feat, h1, h2, logits, label, n, grads, loss, dL, dH2, dH1,
                         static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "MLP backward launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_mlp_sgd_device(float* params, const float* grads, float* momentum, float lr, float mu, void* stream,
                      const char** error)
{
    if (error) *error = nullptr;
    if (!params || !grads || !momentum) { if (error) *error = "null MLP buffer"; return GA_INVALID_ARGUMENT; }
    if (ga::mlp_sgd(params, grads, momentum, lr, mu, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "SGD launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_mlp_route_device(const float* logits, int64_t n, float q, int32_t* bucket, void* stream,
                        const char** error)
{
    if (error) *error = nullptr;
    if (n < 0 || (n > 0 && (!logits || !bucket))) { if (error) *error = "invalid route arguments"; return GA_INVALID_ARGUMENT; }
    if (ga::mlp_route(logits, n, q, bucket, static_cast<hipStream_t>(stream)) != hipSuccess) {
        if (error) *error = "route launch";
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int32_t ga_route_bucket_bound(int32_t bucket)
{
    return (bucket >= 0 && bucket < ga::kBuckets) ? kBucketBound[bucket] : -1;
}

// Stack bytes per class launch of the routed path's second pass.
constexpr int64_t kRerunStackCap = int64_t(4) << 30;

// The routed WFA batch (sketch, network, routed pass, plain re-run) enqueued
// on st; with stats non-null the stream is waited for once at the end.
extern "C++" const char* ga::wfa_routed_run(const ga_wfa_dev_batch* b, const float* params, float q, ga_route_stats* stats,
                               hipStream_t st)
{
    const int64_t n = b->n_pairs;
    if (n <= 0) return nullptr;
    std::call_once(g_pool_once, init_pool);
    t_timer.reset();
    std::vector<void*> held;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (ws_alloc(&p, bytes, st) != hipSuccess) return nullptr;
        held.push_back(p);
        return p;
    };
    auto release = [&](size_t from) {
        for (size_t k = from; k < held.size(); ++k) (void)ws_free(held[k], st);
        held.resize(from);
    };
    const char* err = nullptr;
    float* feat = static_cast<float*>(alloc(size_t(n) * ga::kSketchDim * 4));
    int32_t* lb = static_cast<int32_t*>(alloc(size_t(n) * 4));
    float* logits = static_cast<float*>(alloc(size_t(n) * ga::kBuckets * 4));
    // the fused forward reads the weights as float4: parameters off a 16-byte
    // boundary take the layer-by-layer forward (activations in scratch)
    const bool fused = (reinterpret_cast<uintptr_t>(params) & 15u) == 0;
    float* h1 = fused ? nullptr : static_cast<float*>(alloc(size_t(n) * ga::kHidden * 4));
    float* h2 = fused ? nullptr : static_cast<float*>(alloc(size_t(n) * ga::kHidden * 4));
    int32_t* bucket = static_cast<int32_t*>(alloc(size_t(n) * 4));
    int32_t* sm1 = static_cast<int32_t*>(alloc(size_t(n) * 4));
    int32_t* sm2 = static_cast<int32_t*>(alloc(size_t(n) * 4));
    auto* dstats = static_cast<unsigned long long*>(alloc(64));
    if (!feat || !lb || !logits || !bucket || !sm1 || !sm2 || !dstats || (!fused && (!h1 || !h2)))
        err = "hipMalloc(router buffers)";
    ga::SketchLaunch S{b->pattern, b->pattern_off, b->pattern_len, b->text, b->text_off, b->text_len, n, feat, lb};
    if (!err && (ga::zero_u64_launch(dstats, 4, st) != hipSuccess || ga::sketch_launch(S, st) != hipSuccess ||
                 (fused ? ga::mlp_logits(params, feat, n, logits, st)
                        : ga::mlp_forward(params, feat, n, h1, h2, logits, st)) != hipSuccess ||
                 ga::mlp_route(logits, n, q, bucket, st) != hipSuccess ||
                 ga::route_first_pass_launch(bucket, lb, b->smax, b->pattern_len, b->text_len, int(n), sm1, b->score,
                                             b->cigar ? b->cigar_len : nullptr, dstats, st) != hipSuccess))
        err = "router kernel launch";
    ga::WfaQueueLaunch L{};
    L.pat = b->pattern; L.poff = b->pattern_off; L.plen = b->pattern_len;
    L.txt = b->text; L.toff = b->text_off; L.tlen = b->text_len;
    L.want = b->want_cigar;
    L.score = b->score; L.cigar = b->cigar; L.cig_off = b->cigar_off; L.cig_len = b->cigar_len;
    L.wide = b->max_len > 32766;
    // each pass forks its classes onto streams of its own
    std::vector<hipStream_t> side[2];
    std::vector<hipEvent_t> ev[2] = {std::vector<hipEvent_t>(ga::kWfaClasses, nullptr),
                                     std::vector<hipEvent_t>(ga::kWfaClasses, nullptr)};
    for (int p = 0; p < 2; ++p) {
        for (int c = 1; !err && c < ga::kWfaClasses; ++c) {
            hipStream_t cs = acquire_stream();
            if (!cs) err = "hipStreamCreate";
            else side[p].push_back(cs);
        }
        for (auto& e : ev[p])
            if (!err && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) err = "hipEventCreate";
    }
    // Two device-planned passes, no host synchronisation: the routed bounds,
    // then the pairs not resolved under them with their own smax.  A routed
    // bound in the largest class is the pair's own smax (route_first_pass), so
    // no pair of that class needs a second try, and the second pass waits only
    // for the first pass's smaller classes: it runs beside the largest class's
    // tail.  Both passes' buffers stay held until the end.
    size_t detached = SIZE_MAX;
    if (!err) t_timer.begin(st);                          // both passes: one interval
    for (int pass = 0; pass < 2 && !err; ++pass) {
        if (pass == 1 && ga::route_second_pass_launch(sm1, b->smax, b->score, int(n), sm2, dstats, st) != hipSuccess) {
            err = "router kernel launch";
            break;
        }
        L.smax = pass == 0 ? sm1 : sm2;
        WfaDevClasses C;
        // the second pass holds its stacks beside the first pass's largest
        // class and runs the few pairs the router sent too low: a small cap
        err = setup_wfa_dev_classes(L, n, b->smax_max, b->max_len, b->cigar != nullptr, C, alloc,
                                    pass == 0 ? -1 : kRerunStackCap);
        if (!err) {
            const size_t detach = pass == 0 ? size_t(ga::kWfaClasses - 1) : SIZE_MAX;
            err = enqueue_wfa_dev(C, st, side[pass], ev[pass], detach);
            if (pass == 0 && C.L.size() > detach) detached = detach;
        }
    }
    // join the first pass's largest class
    if (!err && detached != SIZE_MAX && hipStreamWaitEvent(st, ev[0][detached], 0) != hipSuccess)
        err = "hipStreamWaitEvent";
    t_timer.end(st);
    unsigned long long hs[4] = {0, 0, 0, 0};
    if (!err && stats && (hipMemcpyAsync(hs, dstats, sizeof(hs), hipMemcpyDeviceToHost, st) != hipSuccess ||
                          hipStreamSynchronize(st) != hipSuccess))
        err = "hipMemcpy(router stats)";
    release(0);
    for (auto& evp : ev)
        for (auto& e : evp)
            if (e) (void)hipEventDestroy(e);
    for (auto& sp : side)
        for (hipStream_t cs : sp) release_stream(cs);
    if (!err && stats) {
        stats->shortcut = int64_t(hs[0]);
        stats->routed = int64_t(hs[1]);
        stats->plain = int64_t(hs[2]);
        stats->rerun = int64_t(hs[3]);
    }
    return err;
}

int ga_wfa_routed_device(const ga_wfa_dev_batch* b, const float* params, float q, ga_route_stats* stats,
                         void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (stats) *stats = ga_route_stats{};
    if (const char* e = check_dev_batch(b)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (!params) { if (error) *error = "null MLP parameters"; return GA_INVALID_ARGUMENT; }
    if (const char* e = ga::wfa_routed_run(b, params, q, stats, static_cast<hipStream_t>(stream))) {
        if (error) *error = e;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

// ------------------------------------------------------ the library's router
namespace {
std::atomic<const float*> g_router_params{nullptr};
std::atomic<float> g_router_q{0.99f};
}  // namespace

int ga_router_set(const float* params, float q, const char** error)
{
    if (error) *error = nullptr;
    if (!params) { g_router_params.store(nullptr); return GA_OK; }
    if (!(q <= 1.0f)) { if (error) *error = "routing quantile above 1"; return GA_INVALID_ARGUMENT; }
    // one device copy per call, kept until the process ends: batches already
    // running keep reading the network they started with
    float* d = nullptr;
    if (hipMalloc(&d, size_t(ga::kMlpParams) * 4) != hipSuccess) {
        if (error) *error = "hipMalloc(router parameters)";
        return GA_DEVICE_ERROR;
    }
    if (hipMemcpy(d, params, size_t(ga::kMlpParams) * 4, hipMemcpyHostToDevice) != hipSuccess) {
        if (error) *error = "hipMemcpy(router parameters)";
        return GA_DEVICE_ERROR;
    }
    g_router_q.store(q);
    g_router_params.store(d);
    return GA_OK;
}

int ga_router_enabled(void) { return g_router_params.load() != nullptr ? 1 : 0; }

extern "C++" const float* ga::router_params(float* q)
{
    const float* p = g_router_params.load();
    if (q) *q = g_router_q.load();
    return p;
}

void ga_set_thread_workspace_budget(int64_t bytes) { t_workspace_budget = bytes > 0 ? bytes : 0; }

void ga_trim_workspace(void) { ws_trim(-1, 0); }

// ------------------------------------------------------------------ autotuner
void ga_wfa_set_tuning(const ga_wfa_tuning* t)
{
    for (int c = 0; c < 4; ++c) g_tuned_waves[c].store(t ? t->waves[c] : 0, std::memory_order_relaxed);
}

void ga_wfa_get_tuning(ga_wfa_tuning* t)
{
    if (!t) return;
    for (int c = 0; c < 4; ++c) t->waves[c] = wfa_class_waves(c == 0 ? 64 : c == 1 ? 256 : c == 2 ? 1024 : 2048);
    for (int c = 0; c < 4; ++c) t->class_ms[c] = -1.0f;
}

int ga_wfa_autotune(const ga_wfa_dev_batch* b, int reps, ga_wfa_tuning* out, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (const char* e = check_dev_batch(b)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (!out) { if (error) *error = "null output"; return GA_INVALID_ARGUMENT; }
    if (!b->host_smax || !b->host_pattern_len || !b->host_text_len) {
        if (error) *error = "autotuning needs the host copies of smax and the lengths";
        return GA_INVALID_ARGUMENT;
    }
    if (b->n_pairs <= 0) { if (error) *error = "empty batch"; return GA_INVALID_ARGUMENT; }
    std::call_once(g_pool_once, init_pool);
    hipStream_t st = static_cast<hipStream_t>(stream);
    reps = std::max(1, std::min(reps, 16));
    const WfaPlan plan = plan_wfa(b->n_pairs, b->host_smax, b->host_pattern_len, b->host_text_len);
    ga_wfa_tuning t{};
    ga_wfa_get_tuning(&t);
    ga::WfaQueueLaunch base{};
    base.pat = b->pattern; base.poff = b->pattern_off; base.plen = b->pattern_len;
    base.txt = b->text; base.toff = b->text_off; base.tlen = b->text_len;
    base.smax = b->smax; base.want = b->want_cigar;
    base.score = b->score; base.cigar = b->cigar; base.cig_off = b->cigar_off; base.cig_len = b->cigar_len;
    base.wide = b->max_len > 32766;
    const char* err = nullptr;
    int32_t* d_order = nullptr;
    int32_t* d_next = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ws_alloc(reinterpret_cast<void**>(&d_order), plan.order.size() * 4 + 64, st) != hipSuccess ||
        ws_alloc(reinterpret_cast<void**>(&d_next), 64, st) != hipSuccess ||
        hipMemcpyAsync(d_order, plan.order.data(), plan.order.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        err = "autotune set-up";
    // each class on its own: time every candidate, keep the fastest
    for (size_t c = 0; c < plan.bound.size() && !err; ++c) {
        const int cls = wfa_class_of(plan.bound[c]);
        float best = -1.0f;
        int best_w = t.waves[cls];
        for (int w : {1, 2, 4, 8}) {
            ga::WfaQueueLaunch L = base;
            L.order = plan.order.empty() ? nullptr : d_order + plan.first[c];
            L.n = int32_t(plan.count[c]);
            L.next = d_next;
            L.smax_cls = plan.bound[c];
            L.max_plen = plan.max_plen[c];
            L.max_tlen = plan.max_tlen[c];
            L.waves = w;
            L.staged = ga::wfa_lds_bytes(L) <= kWfaMaxStagedLds;
            const bool trace = b->cigar != nullptr;
            const int64_t per_group = wfa_size_stacks(L, trace, wfa_stack_budget(), wfa_defer_cap(wfa_stack_budget(), false));
            L.ws = nullptr;
            if (trace && ws_alloc(&L.ws, size_t(per_group) * size_t(L.n_groups), st) != hipSuccess) {
                err = "hipMalloc(wavefront stacks)";
                break;
            }
            float tot = 0.0f;
            for (int r = 0; r <= reps && !err; ++r) {         // first run warms up
                if (hipEventRecord(e0, st) != hipSuccess || ga::wfa_queue_launch(L, st) != hipSuccess ||
                    hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    err = "autotune launch";
                    break;
                }
                float ms = 0.0f;
                if (r > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) tot += ms;
            }
            if (L.ws) (void)ws_free(L.ws, st);
            if (err) break;
            const float avg = tot / float(reps);
            if (best < 0.0f || avg < best) { best = avg; best_w = w; }
        }
        t.waves[cls] = best_w;
        t.class_ms[cls] = best;
    }
    if (d_order) (void)ws_free(d_order, st);
    if (d_next) (void)ws_free(d_next, st);
    (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err) { if (error) *error = err; return GA_DEVICE_ERROR; }
    *out = t;
    return GA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// MSA: profile-profile Gotoh batches (the reference's msa_pp_trace_kernel
// launches, src/msa/msa_gpu.cpp, over genomsa::align_profiles semantics)
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kMsaDirBudget = int64_t(8) << 30;     // direction bytes per launch

const char* msa_params_error(const ga_msa_params* p)
{
    if (!p) return "null MSA parameters";
    if (p->alpha != 4 && p->alpha != 20) return "MSA alphabet must be 4 (DNA) or 20 (protein)";
    return nullptr;
}

ga::MsaParamsDev msa_dev_params(const ga_msa_params& p)
{
    return ga::MsaParamsDev{p.match, p.ts, p.tv, p.gap_open, p.gap_extend, p.free_end_gaps != 0 ? 1 : 0,
                            p.psgp != 0 ? 1 : 0, p.psgp_scale, p.psgp_min_open, p.psgp_min_ext, p.alpha};
}

// Substitution scores as the reference's col_score reads them
// (msa_ref.cpp:1023-1028 for DNA: match, transitions A<->G and C<->T, other
// transversions; the matrix otherwise).
std::vector<float> msa_sub_table(const ga_msa_params& p)
{
    const int al = p.alpha;
    std::vector<float> t(size_t(al) * al);
    for (int a = 0; a < al; ++a)
        for (int b = 0; b < al; ++b) {
            float v;
            if (al == 4) {
                const bool ts = (a == 0 && b == 2) || (a == 2 && b == 0) || (a == 1 && b == 3) || (a == 3 && b == 1);
                v = a == b ? p.match : (ts ? p.ts : p.tv);
            } else {
                v = p.sub[a * 20 + b];
            }
            t[size_t(a) * al + b] = v;
        }
    return t;
}

// Sparse columns of a device profile table, then the pairs in launches of at
// most kMsaDirBudget direction bytes.  hlen / ha / hb: host copies of the
// profile lengths and pair indices.
const char* enqueue_msa(const ga_msa_params& p, const float* cols, const float* occ, int64_t n_cols,
                        const int64_t* col_off, const int32_t* len, const int32_t* a_idx, const int32_t* b_idx,
                        int64_t n, const int32_t* hlen, const int32_t* ha, const int32_t* hb, float* score,
                        int32_t* span, uint8_t* cigar, const int64_t* cig_off, int32_t* cig_len, hipStream_t st)
{
    DevArena D;
    D.st = st;
    const int al = p.alpha;
    ga::MsaSparseLaunch S{};
    S.cols = cols; S.occ = occ; S.n_cols = n_cols; S.P = msa_dev_params(p);
    S.nz_cnt = D.alloc<uint8_t>(size_t(std::max<int64_t>(n_cols, 1)));
    S.nz_sym = D.alloc<uint8_t>(size_t(std::max<int64_t>(n_cols, 1)) * al);
    S.nz_frq = D.alloc<float>(size_t(std::max<int64_t>(n_cols, 1)) * al);
    S.gopen = D.alloc<float>(size_t(std::max<int64_t>(n_cols, 1)));
    S.gext = D.alloc<float>(size_t(std::max<int64_t>(n_cols, 1)));
    const float* d_sub = D.upload(msa_sub_table(p));
    // per pair: direction words and boundary floats relative to its launch
    std::vector<int64_t> dir_off(static_cast<size_t>(n)), bnd_off(static_cast<size_t>(n));
    std::vector<int64_t> first{0};
    std::vector<int64_t> dwords, bfloats;
    int64_t dw = 0, bf = 0;
    for (int64_t q = 0; q < n; ++q) {
        const int M = std::max(hlen[ha[q]], 0), N = std::max(hlen[hb[q]], 0);
        const int64_t w = ga::msa_dir_words(M, N);
        if (dw > 0 && (dw + w) * 4 > kMsaDirBudget) {
            first.push_back(q);
            dwords.push_back(dw);
            bfloats.push_back(bf);
            dw = bf = 0;
        }
        dir_off[size_t(q)] = dw;
        bnd_off[size_t(q)] = bf;
        dw += w;
        bf += 3 * (int64_t(M) + 1);
    }
    first.push_back(n);
    dwords.push_back(dw);
    bfloats.push_back(bf);
    const int64_t* d_dir_off = D.upload(dir_off);
    const int64_t* d_bnd_off = D.upload(bnd_off);
    if (!D.ok) { D.release(); return "hipMalloc/hipMemcpy (MSA buffers)"; }
    const char* err = nullptr;
    if (ga::msa_sparse_launch(S, st) != hipSuccess) err = "MSA sparse kernel launch";
    ga::MsaLaunch L{};
    L.nz_cnt = S.nz_cnt; L.nz_sym = S.nz_sym; L.nz_frq = S.nz_frq; L.gopen = S.gopen; L.gext = S.gext;
    L.n_cols = n_cols; L.col_off = col_off; L.len = len; L.a_idx = a_idx; L.b_idx = b_idx;
    L.sub = d_sub; L.P = S.P;
    L.dir_off = d_dir_off; L.bnd_off = d_bnd_off;
    L.score = score; L.span = span; L.cigar = cigar; L.cig_off = cig_off; L.cig_len = cig_len;
    for (size_t c = 0; c + 1 < first.size() && !err; ++c) {
        void* dirs = nullptr;
        void* bnd = nullptr;
        if (ws_alloc(&dirs, size_t(std::max<int64_t>(dwords[c], 16)) * 4, st) != hipSuccess ||
            ws_alloc(&bnd, size_t(std::max<int64_t>(bfloats[c], 16)) * 4, st) != hipSuccess) {
            if (dirs) (void)ws_free(dirs, st);
            err = "hipMalloc(MSA direction bytes)";
            break;
        }
        L.dirs = static_cast<uint32_t*>(dirs);
        L.bnd = static_cast<float*>(bnd);
        L.first_pair = int32_t(first[c]);
        if (ga::msa_pp_launch(L, int(first[c + 1] - first[c]), st, &t_timer_hooks) != hipSuccess)
            err = "MSA kernel launch";
        (void)ws_free(dirs, st);
        (void)ws_free(bnd, st);
    }
    D.release();
    return err;
}

}  // namespace

extern "C" {

int ga_msa_profile_from_rows(const char* rows, int32_t nrows, int32_t L, int32_t alpha, float* cols, float* occ,
                             float* gap)
{
    // msa_ref.cpp:21-106 (symbol fractions) and :773-794 (counts over rows)
    if ((alpha != 4 && alpha != 20) || nrows <= 0 || L < 0 || (L > 0 && (!rows || !cols || !occ)))
        return GA_INVALID_ARGUMENT;
    auto counts = [alpha](char c, float* cnt) -> bool {
        for (int s = 0; s < alpha; ++s) cnt[s] = 0.f;
        if (alpha == 4) {
            int mask = 0;
            switch (c) {
            case 'A': case 'a': mask = 1; break;
            case 'C': case 'c': mask = 2; break;
            case 'G': case 'g': mask = 4; break;
            case 'T': case 't': case 'U': case 'u': mask = 8; break;
            case 'R': mask = 5; break;
            case 'Y': mask = 10; break;
            case 'M': mask = 3; break;
            case 'K': mask = 12; break;
            case 'S': mask = 6; break;
            case 'W': mask = 9; break;
            case 'H': mask = 11; break;
            case 'B': mask = 14; break;
            case 'V': mask = 7; break;
            case 'D': mask = 13; break;
            case 'N': case 'n': case 'X': mask = 15; break;
            default: return false;
            }
            const float f = 1.0f / float(__builtin_popcount(unsigned(mask)));
            for (int s = 0; s < 4; ++s)
                if (mask & (1 << s)) cnt[s] = f;
            return true;
        }
        switch (c) {
        case 'B': cnt[2] = cnt[3] = 0.5f; return true;
        case 'Z': cnt[5] = cnt[6] = 0.5f; return true;
        case 'J': cnt[9] = cnt[10] = 0.5f; return true;
        case 'U': case 'u': cnt[4] = 1.f; return true;
        case 'X': case 'x': case 'O': case 'o':
            for (int s = 0; s < 20; ++s) cnt[s] = 1.f / 20.f;
            return true;
        default: break;
        }
        static const char aa[] = "ARNDCQEGHILKMFPSTWYV";
        const char* q = static_cast<const char*>(std::memchr(aa, c >= 'a' && c <= 'z' ? c - 32 : c, 20));
        if (!q || c == 0) return false;
        cnt[q - aa] = 1.f;
        return true;
    };
    for (int i = 0; i < L; ++i) {
        for (int s = 0; s < alpha; ++s) cols[size_t(i) * alpha + s] = 0.f;
        occ[i] = 0.f;
        if (gap) gap[i] = 0.f;
    }
    float cnt[20];
    for (int r = 0; r < nrows; ++r)
        for (int i = 0; i < L; ++i)
            if (counts(rows[size_t(r) * L + i], cnt)) {
                for (int s = 0; s < alpha; ++s) cols[size_t(i) * alpha + s] += cnt[s];
                occ[i] += 1.f;
            } else if (gap) {
                gap[i] += 1.f;
            }
    for (int i = 0; i < L; ++i) {
        for (int s = 0; s < alpha; ++s) cols[size_t(i) * alpha + s] /= float(nrows);
        occ[i] /= float(nrows);
        if (gap) gap[i] /= float(nrows);
    }
    return GA_OK;
}

int ga_msa_align_profiles(const ga_msa_profile* A, const ga_msa_profile* B, int64_t n, const ga_msa_params* p,
                          ga_msa_pp_res* res, char* arena, int64_t arena_bytes, const char** error)
{
    if (error) *error = nullptr;
    auto refuse = [&](const char* m) {
        if (error) *error = m;
        return GA_INVALID_ARGUMENT;
    };
    if (n <= 0) return GA_OK;
    if (const char* e = msa_params_error(p)) return refuse(e);
    if (!A || !B || !res) return refuse("null profile or result array");
    if (n > INT32_MAX) return refuse("too many pairs");
    for (int64_t q = 0; q < n; ++q) res[q] = ga_msa_pp_res{0.f, 0, 0, 0, 0, 0, 0};
    int64_t need = 0;
    for (int64_t q = 0; q < n; ++q) {
        for (const ga_msa_profile* f : {A + q, B + q}) {
            if (f->len < 0) return refuse("negative profile length");
            if (f->len > 0 && (!f->cols || !f->occ)) return refuse("null profile columns with a positive length");
        }
        need += int64_t(A[q].len) + B[q].len;
    }
    if (need > 0 && (!arena || arena_bytes < need)) return refuse("CIGAR arena smaller than the batch needs");
    std::call_once(g_pool_once, init_pool);
    const int al = p->alpha;
    // column table: each distinct profile once (a profile shared by several
    // pairs, one bait against many, is uploaded once)
    std::vector<const ga_msa_profile*> prof;
    std::vector<int32_t> ha(static_cast<size_t>(n)), hb(static_cast<size_t>(n));
    {
        std::unordered_map<const float*, int32_t> seen;
        auto id = [&](const ga_msa_profile* f) -> int32_t {
            if (f->len > 0) {
                auto it = seen.find(f->cols);
                if (it != seen.end() && prof[size_t(it->second)]->occ == f->occ &&
                    prof[size_t(it->second)]->len == f->len)
                    return it->second;
            }
            prof.push_back(f);
            const int32_t k = int32_t(prof.size() - 1);
            if (f->len > 0) seen.emplace(f->cols, k);
            return k;
        };
        for (int64_t q = 0; q < n; ++q) {
            ha[size_t(q)] = id(A + q);
            hb[size_t(q)] = id(B + q);
        }
    }
    std::vector<int32_t> hlen(prof.size());
    std::vector<int64_t> col_off(prof.size());
    int64_t n_cols = 0;
    for (size_t k = 0; k < prof.size(); ++k) {
        hlen[k] = prof[k]->len;
        col_off[k] = n_cols;
        n_cols += prof[k]->len;
    }
    std::vector<float> cols(size_t(std::max<int64_t>(n_cols, 1)) * al), occ(size_t(std::max<int64_t>(n_cols, 1)));
    for (size_t k = 0; k < prof.size(); ++k)
        if (prof[k]->len > 0) {
            std::memcpy(cols.data() + size_t(col_off[k]) * al, prof[k]->cols, size_t(prof[k]->len) * al * 4);
            std::memcpy(occ.data() + col_off[k], prof[k]->occ, size_t(prof[k]->len) * 4);
        }
    std::vector<int64_t> cig_off(static_cast<size_t>(n));
    int64_t slot = 0;
    for (int64_t q = 0; q < n; ++q) {
        cig_off[size_t(q)] = slot;
        slot += int64_t(A[q].len) + B[q].len;
    }
    hipStream_t st = acquire_stream();
    if (!st) { if (error) *error = "hipStreamCreate"; return GA_DEVICE_ERROR; }
    DevArena D;
    D.st = st;
    const float* d_cols = D.upload(cols);
    const float* d_occ = D.upload(occ);
    const int64_t* d_col_off = D.upload(col_off);
    const int32_t* d_len = D.upload(hlen);
    const int32_t* d_a = D.upload(ha);
    const int32_t* d_b = D.upload(hb);
    const int64_t* d_cig_off = D.upload(cig_off);
    float* d_score = D.alloc<float>(size_t(n));
    int32_t* d_span = D.alloc<int32_t>(size_t(n) * 4);
    int32_t* d_clen = D.alloc<int32_t>(size_t(n));
    uint8_t* d_cig = D.alloc<uint8_t>(size_t(std::max<int64_t>(slot, 1)));
    const char* err = D.ok ? nullptr : "hipMalloc/hipMemcpy (MSA buffers)";
    t_timer.reset();
    if (!err)
        err = enqueue_msa(*p, d_cols, d_occ, n_cols, d_col_off, d_len, d_a, d_b, n, hlen.data(), ha.data(), hb.data(),
                          d_score, d_span, d_cig, d_cig_off, d_clen, st);
    std::vector<float> score(static_cast<size_t>(n));
    std::vector<int32_t> span(static_cast<size_t>(n) * 4), clen(static_cast<size_t>(n));
    if (!err && (hipMemcpyAsync(score.data(), d_score, size_t(n) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipMemcpyAsync(span.data(), d_span, size_t(n) * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipMemcpyAsync(clen.data(), d_clen, size_t(n) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 (slot > 0 && hipMemcpyAsync(arena, d_cig, size_t(slot), hipMemcpyDeviceToHost, st) != hipSuccess)))
        err = "hipMemcpy(MSA results)";
    D.release();
    const hipError_t se = hipStreamSynchronize(st);
    if (!err && se != hipSuccess) err = hipGetErrorString(se);
    release_stream(st);
    if (err) {
        if (error) *error = err;
        for (int64_t q = 0; q < n; ++q) res[q] = ga_msa_pp_res{0.f, 0, 0, 0, 0, 0, 0};
        return GA_DEVICE_ERROR;
    }
    for (int64_t q = 0; q < n; ++q) {
        ga_msa_pp_res& r = res[q];
        r.score = score[size_t(q)];
        r.ai = span[size_t(q) * 4]; r.aj = span[size_t(q) * 4 + 1];
        r.bi = span[size_t(q) * 4 + 2]; r.bj = span[size_t(q) * 4 + 3];
        r.cigar_len = clen[size_t(q)];
        r.cigar_off = cig_off[size_t(q)] + int64_t(A[q].len) + B[q].len - r.cigar_len;
    }
    return GA_OK;
}

int ga_msa_pp_device(const ga_msa_dev_batch* b, const ga_msa_params* p, void* stream, const char** error)
{
    if (error) *error = nullptr;
    if (const char* e = msa_params_error(p)) { if (error) *error = e; return GA_INVALID_ARGUMENT; }
    if (!b) { if (error) *error = "null batch"; return GA_INVALID_ARGUMENT; }
    if (b->n_pairs <= 0) return GA_OK;
    if (b->n_pairs > INT32_MAX) { if (error) *error = "too many pairs"; return GA_INVALID_ARGUMENT; }
    if (!b->host_len || !b->host_a_idx || !b->host_b_idx) {
        if (error) *error = "ga_msa_pp_device needs the host copies of the lengths and pair indices";
        return GA_INVALID_ARGUMENT;
    }
    for (int32_t k = 0; k < b->n_profiles; ++k)
        if (b->host_len[k] < 0) { if (error) *error = "negative profile length"; return GA_INVALID_ARGUMENT; }
    for (int64_t q = 0; q < b->n_pairs; ++q)
        if (b->host_a_idx[q] < 0 || b->host_a_idx[q] >= b->n_profiles || b->host_b_idx[q] < 0 ||
            b->host_b_idx[q] >= b->n_profiles) {
            if (error) *error = "pair index outside the profile table";
            return GA_INVALID_ARGUMENT;
        }
    std::call_once(g_pool_once, init_pool);
    t_timer.reset();
    const char* err = enqueue_msa(*p, b->cols, b->occ, b->n_cols, b->col_off, b->len, b->a_idx, b->b_idx, b->n_pairs,
                                  b->host_len, b->host_a_idx, b->host_b_idx, b->score, b->span, b->cigar,
                                  b->cigar_off, b->cigar_len, static_cast<hipStream_t>(stream));
    if (err) {
        if (error) *error = err;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// guide tree: k-mer distances and neighbour joining
// ---------------------------------------------------------------------------
extern "C" {

int ga_kmer_distances(const char* seqs, const int64_t* off, const int32_t* len, int32_t n, int32_t k, int32_t alpha,
                      float* D, const char** error)
{
    if (error) *error = nullptr;
    auto refuse = [&](const char* m) {
        if (error) *error = m;
        return GA_INVALID_ARGUMENT;
    };
    if (n <= 1) return GA_OK;
    if (alpha != 4 && alpha != 20) return refuse("k-mer alphabet must be 4 (DNA) or 20 (protein)");
    if (n > 65535) return refuse("more than 65,535 sequences");
    if (!off || !len || !D) return refuse("null sequence or distance array");
    if (k < 1) return refuse("k must be >= 1");
    double codes = 1;
    for (int q = 0; q < k; ++q) codes *= alpha;
    if (codes > double(1 << 24)) return refuse("alpha^k above 2^24 k-mer codes");
    int64_t total = 0;
    int32_t max_len = 0;
    for (int32_t s = 0; s < n; ++s) {
        if (len[s] < 0) return refuse("negative sequence length");
        if (len[s] > 0 && !seqs) return refuse("null sequence with a positive length");
        total = std::max<int64_t>(total, off[s] + len[s]);
        max_len = std::max(max_len, len[s]);
    }
    std::call_once(g_pool_once, init_pool);
    const int64_t words = (int64_t(codes) + 31) / 32;
    hipStream_t st = acquire_stream();
    if (!st) { if (error) *error = "hipStreamCreate"; return GA_DEVICE_ERROR; }
    DevArena Da;
    Da.st = st;
    std::vector<uint8_t> bytes(static_cast<size_t>(std::max<int64_t>(total, 1)));
    if (total > 0) std::memcpy(bytes.data(), seqs, size_t(total));
    ga::KmerLaunch K{};
    K.seq = Da.upload(bytes);
    K.off = Da.upload(std::vector<int64_t>(off, off + n));
    K.len = Da.upload(std::vector<int32_t>(len, len + n));
    K.n = n; K.k = k; K.alpha = alpha; K.words = words;
    K.bits = Da.alloc<uint32_t>(size_t(n) * size_t(words));
    K.count = Da.alloc<int32_t>(size_t(n));
    const int64_t np = int64_t(n) * (n - 1) / 2;
    K.dist = Da.alloc<float>(size_t(np));
    const char* err = Da.ok ? nullptr : "hipMalloc/hipMemcpy (k-mer buffers)";
    t_timer.reset();
    if (!err && int64_t(n) * words > INT32_MAX) err = "k-mer bitsets larger than 2^31 words";
    if (!err && ga::zero_i32_launch(reinterpret_cast<int32_t*>(K.bits), int(int64_t(n) * words), st) != hipSuccess)
        err = "k-mer kernel launch";
    if (!err && ga::kmer_distances_launch(K, max_len, st, &t_timer_hooks) != hipSuccess) err = "k-mer kernel launch";
    if (!err && hipMemcpyAsync(D, K.dist, size_t(np) * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        err = "hipMemcpy(distances)";
    Da.release();
    const hipError_t se = hipStreamSynchronize(st);
    if (!err && se != hipSuccess) err = hipGetErrorString(se);
    release_stream(st);
    if (err) {
        if (error) *error = err;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

int ga_nj_tree(const float* D, int32_t n, int32_t* children, const char** error)
{
    if (error) *error = nullptr;
    if (n <= 1) return GA_OK;
    if (!D || !children) { if (error) *error = "null distance or tree array"; return GA_INVALID_ARGUMENT; }
    if (n > 65535) { if (error) *error = "more than 65,535 sequences"; return GA_INVALID_ARGUMENT; }
    std::call_once(g_pool_once, init_pool);
    hipStream_t st = acquire_stream();
    if (!st) { if (error) *error = "hipStreamCreate"; return GA_DEVICE_ERROR; }
    DevArena Da;
    Da.st = st;
    const int64_t np = int64_t(n) * (n - 1) / 2;
    ga::NjLaunch J{};
    J.n = n;
    J.packed = Da.upload(std::vector<float>(D, D + np));
    J.d = Da.alloc<double>(size_t(n) * size_t(n));
    J.r = Da.alloc<double>(size_t(n));
    J.part = Da.alloc<ga::NjBest>(size_t(n));
    J.alive0 = Da.alloc<int32_t>(size_t(n));
    J.alive1 = Da.alloc<int32_t>(size_t(n));
    J.slot_node = Da.alloc<int32_t>(size_t(n));
    J.children = Da.alloc<int32_t>(2 * size_t(n));
    const char* err = Da.ok ? nullptr : "hipMalloc/hipMemcpy (NJ buffers)";
    t_timer.reset();
    if (!err && ga::nj_launch(J, st, &t_timer_hooks) != hipSuccess) err = "NJ kernel launch";
    if (!err && hipMemcpyAsync(children, J.children, size_t(n - 1) * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        err = "hipMemcpy(tree)";
    Da.release();
    const hipError_t se = hipStreamSynchronize(st);
    if (!err && se != hipSuccess) err = hipGetErrorString(se);
    release_stream(st);
    if (err) {
        if (error) *error = err;
        return GA_DEVICE_ERROR;
    }
    return GA_OK;
}

}  // extern "C"
