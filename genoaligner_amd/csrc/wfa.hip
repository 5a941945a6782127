// wfa.hip -- gfx950 kernels for bounded unit-cost edit distance (WFA) with an
// on-device traceback that writes the CIGAR straight into the caller's arena.
//
// Algorithm (reference: include/genoaligner/backend/wfa_kernel.hip):
//   offsets are text coordinates h, pattern coordinate v = h - k (:103-110);
//   compute-next  max(del = W[s-1][k+1], max(ins = W[s-1][k-1], mism = W[s-1][k]) + 1),
//                 NULL when unreachable, h > n or v > m (:119-127);
//   extension     greedy run of equal bytes along the diagonal (:130-136);
//   termination   W[s][n-m] >= n (:552-555);
//   backtrace     mismatch (W[s-1][k]+1) >=, then insertion (W[s-1][k-1]+1) >,
//                 then deletion (W[s-1][k+1]) >; runs of M between; leading
//                 matches, then D, then I (:338-437).
//
// Mapping to CDNA4: a persistent workgroup of 1, 2, 4 or 8 waves (by size
// class) aligns one pair at a time and claims the next from a queue.  Thread t
// owns diagonals lo + t, lo + t + 64 * waves, ... of wavefront s, lo .. hi
// being the diagonals budget pruning keeps.  The two live wavefronts sit in
// LDS (ping-pong, smax + 5 slots each: the live range at a per-score origin
// and two NULL guard slots per side); in the trace variant every wavefront is
// also stored to an HBM stack (row s at offset s*s, 2s+1 entries, coalesced
// across lanes) that the traceback reads back.  The extension compares 16
// bases per probe over 2-bit copies of both sequences staged in LDS, or 8
// bytes per step over the raw bytes of a pair holding any byte other than
// A/C/G/T, so every byte value keeps its literal meaning (N matches only N,
// case is significant).  A traceback runs in lockstep on one wave (its state
// is uniform) and writes each run of M's cooperatively, 64 bytes per store
// instruction; with several waves per workgroup the walks are deferred and
// shared out one pair per wave.
#include "ga_device.hpp"

#include <atomic>

#include <mutex>
#include <vector>

namespace ga {
namespace {

// 8 bytes starting at any byte address.  Reads the two aligned words that
// contain them, so up to 15 bytes past `p` must be readable (sequence buffers
// carry kSeqPad bytes of padding on both sides).
__device__ __forceinline__ uint64_t load8(const uint8_t* p)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint64_t* w = reinterpret_cast<const uint64_t*>(a & ~uintptr_t(7));
    const unsigned sh = unsigned(a & 7u) * 8u;
    const uint64_t lo = w[0];
    const uint64_t hi = w[1];
    return sh ? ((lo >> sh) | (hi << (64u - sh))) : lo;
}

// Length of the run of equal bytes P[0..] == T[0..], capped at lim.
__device__ __forceinline__ int match_run(const uint8_t* P, const uint8_t* T, int lim)
{
    int run = 0;
    while (run < lim) {
        const uint64_t x = load8(P + run) ^ load8(T + run);
        if (x) { run += __builtin_ctzll(x) >> 3; break; }
        run += 8;
    }
    return run < lim ? run : lim;
}

// ---------------------------------------------------------------------------
// Persistent, queue-fed kernel.  A workgroup of WAVES wavefronts aligns one
// pair at a time and claims the next from an atomic counter, so a class of
// pairs of very different cost (lengths 100 bp .. 10 kbp, distances up to
// 2048) keeps every CU busy until the queue drains; the host orders each class
// longest first.  Per pair the workgroup
//   * stages both sequences into LDS (STAGED; the extension then never touches
//     global memory), else extends from the packed global buffers;
//   * keeps the two live wavefronts in LDS as Off (16-bit when every text is
//     shorter than 32767 bytes), thread t owning diagonals -s + t, -s + t +
//     64*WAVES, ...;
//   * stores every wavefront to the workgroup's HBM stack (trace pairs only):
//     row s at offset s*s, 2s+1 entries, reused from pair to pair;
//   * runs the traceback on wave 0 (uniform state, 64-wide CIGAR stores).
// An offset is reachable iff it is >= 0: real offsets are, and everything
// derived from the unreachable sentinel (INT16_MIN or INT32_MIN/2, plus at
// most one) is negative; this is the reference's `> NEG + 1024` test on the
// values the recurrence can produce.
template <typename Off> struct OffNull;
template <> struct OffNull<int16_t> { static constexpr int v = INT16_MIN; };
template <> struct OffNull<int32_t> { static constexpr int v = kWfaNeg; };

template <typename Off>
__device__ __forceinline__ int ws_get(const Off* ws, int s, int k)
{
    if (k < -s || k > s) return OffNull<Off>::v;
    return ws[int64_t(s) * s + k + s];
}

// 2-bit packed sequences in LDS: base i is bits 2(i mod 16) .. +1 of word
// i / 16, code (byte >> 1) & 3 (A 0, C 1, T 2, G 3).  Used only for pairs whose
// bytes are all A, C, G or T, where code equality is byte equality.
__device__ __forceinline__ uint32_t base_code(uint32_t b) { return (b >> 1) & 3u; }
__device__ __forceinline__ bool is_acgt(uint32_t b) { return b == 'A' || b == 'C' || b == 'G' || b == 'T'; }


// Run of equal bases P[v..] == T[h..], capped at lim (lim >= 0; bases v and h
// must lie inside the staged words even when lim == 0).  Bases i .. i+15 of a
// staged sequence are two aligned words joined by one bit funnel shift
// (v_alignbit_b32).
__device__ __forceinline__ int lds_match_run(const uint32_t* P, int v, const uint32_t* T, int h, int lim)
{
    // word pointers and bit shifts fixed for the run; the first probe is
    // straight-line code (most runs end in it), the loop only continues runs
    const uint32_t* p = P + (v >> 4);
    const uint32_t* t = T + (h >> 4);
    const uint32_t sp = (uint32_t(v) & 15u) * 2u, st = (uint32_t(h) & 15u) * 2u;
    // the four words are read before either is used: one LDS round trip
    const uint32_t p0 = p[0], p1 = p[1], t0 = t[0], t1 = t[1];
    __builtin_amdgcn_sched_barrier(0);
    uint32_t x = __builtin_amdgcn_alignbit(p1, p0, sp) ^ __builtin_amdgcn_alignbit(t1, t0, st);
    // equal bases before the first differing one: (lowest set bit) / 2, where
    // v_ffbl_b32 gives 0xffffffff for x == 0, so the capped run exceeds 16
    // exactly when all 16 bases match and the run may go on
    uint32_t low;
    asm("v_ffbl_b32 %0, %1" : "=v"(low) : "v"(x));
    uint32_t run = low >> 1;
    run = run < uint32_t(lim) ? run : uint32_t(lim);
    if (run > 16) {
        uint32_t r = 16;
        for (;;) {
            ++p;
            ++t;
            x = __builtin_amdgcn_alignbit(p[1], p[0], sp) ^ __builtin_amdgcn_alignbit(t[1], t[0], st);
            if (x != 0 || r + 16 >= uint32_t(lim)) break;
            r += 16;
        }
        run = r + (uint32_t(__builtin_ctzg(x, 32)) >> 1);
        run = run < uint32_t(lim) ? run : uint32_t(lim);
    }
    return int(run);
}

// Stage one sequence as 2-bit words (16 bases per thread per step, two spare
// words past the end); returns false (for this thread) on any non-ACGT byte.
template <int NT>
__device__ __forceinline__ bool stage2(uint32_t* dst, const uint8_t* src, int len, int tid)
{
    bool ok = true;
    const int words = ((len + 15) >> 4) + 2;
    for (int x = tid; x < words; x += NT) {
        uint32_t w = 0;
        const int base = 16 * x;
        if (base < len) {
            const uint64_t a = load8(src + base);
            const uint64_t b = load8(src + base + 8);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t c = uint32_t(((j < 8 ? a : b) >> (8 * (j & 7))) & 0xffu);
                if (base + j < len) {
                    ok = ok && is_acgt(c);
                    w |= base_code(c) << (2 * j);
                }
            }
        }
        dst[x] = w;
    }
    return ok;
}

// LDS layout of a workgroup: [claim slots: 4 ints][pattern, 2-bit][text, 2-bit]
// [wavefront A][wavefront B][traceback window][deferred-walk queue], each
// region 16-byte aligned (the sequences first: the pattern's base then fits
// the offset field of the extension's LDS reads, and no add per diagonal is
// spent on it).  Budget pruning keeps at most smax + 1 diagonals live (|k| <=
// s and |k - ak| <= smax - s), so a wavefront buffer holds smax + 5 slots:
// the live range at slot k - lo + 2 and two guard slots on each side.  With
// deferred walks the two buffers' region is at least one walk window per
// wave (the walks use it between pairs).
__host__ __device__ inline int wfa_wf_bytes(int smax_cls, int off_bytes)
{
    return ((smax_cls + 5) * off_bytes + 15) & ~15;
}
__host__ __device__ inline int wfa_seq_bytes(int len) { return ((((len > 0 ? len : 0) + 15) / 16 + 2) * 4 + 15) & ~15; }

// Traceback window: kTbRows rows of the stack below the current score, at
// diagonals k0 - kTbRows .. k0 + kTbRows, loaded with one batch of coalesced
// reads and walked out of LDS for up to kTbRows steps.
constexpr int kTbRows = 16;
constexpr int kTbCols = 2 * kTbRows + 1;
constexpr int kTbBatch = 5;          // window loads a lane has in flight at once
__host__ __device__ inline int wfa_win_bytes(int off_bytes) { return (kTbRows * kTbCols * off_bytes + 15) & ~15; }
__host__ __device__ inline int wfa_region_bytes(int smax_cls, int off_bytes, bool defer, int waves)
{
    const int two = 2 * wfa_wf_bytes(smax_cls, off_bytes);
    const int wins = defer ? waves * wfa_win_bytes(off_bytes) : 0;
    return two > wins ? two : wins;
}

// Forward pass over one pair: wavefronts s = 1 .. smax until W[s][n-m] >= n.
// Returns the score or -1.  PACKED: extend over the 2-bit LDS copies, else over
// the byte sequences in global memory.
template <bool TRACE, typename Off, bool PACKED, int NT>
__device__ __forceinline__ int wfa_forward(const uint32_t* sP, const uint32_t* sT, const uint8_t* P,
                                           const uint8_t* T, int m, int n, int smax,
                                           Off* A, Off* B, Off* ws, bool trace, int tid)
{
    // Wavefront s, diagonals lo .. hi, sits at slots k - lo + 2 of its buffer,
    // and slots 0, 1 (k = lo - 2, lo - 1) and hi - lo + 3, hi - lo + 4 (k =
    // hi + 1, hi + 2) hold NULL: everything the next score reads (the range
    // moves by at most one per side per score) was written by this one.
    // Score 0 is diagonal 0 alone.
    if (tid < 5) {
        Off v = Off(OffNull<Off>::v);
        if (tid == 2) {
            const int o0 = PACKED ? lds_match_run(sP, 0, sT, 0, m < n ? m : n) : match_run(P, T, m < n ? m : n);
            v = Off(o0);
            if (trace) ws[0] = v;
        }
        A[tid] = v;
    }
    __syncthreads();
    const int ak = n - m;
    if (ak == 0 && int(A[2]) >= n) return 0;
    Off* prev = A;
    Off* cur = B;
    int plo = 0;                                               // lo of the previous score
    for (int s = 1; s <= smax; ++s) {
        Off* wrow = TRACE ? ws + int64_t(s) * s : nullptr;   // row s of the stack (wave-uniform)
        // Budget pruning: every step changes the diagonal by at most one, so
        // from score s only diagonals with |k - ak| <= smax - s can still
        // reach the end diagonal ak by smax; the others are not computed.  A
        // kept diagonal reads only kept diagonals of the previous score (its
        // neighbours are within smax - s + 1 of ak), so kept offsets, the end
        // test and the backtrace (whose candidates are kept too) are exactly
        // those of the full wavefront.  A pair that runs to smax computes
        // about half the cells.
        const int bud = smax - s;
        const int lo = -s > ak - bud ? -s : ak - bud;
        const int hi = s < ak + bud ? s : ak + bud;
        const Off* pk = prev + 2 - plo;                        // pk[k] = W[s-1][k]
        Off* ck = cur + 2 - lo;                                // ck[k] = W[s][k]
        for (int k = lo + tid; k <= hi; k += NT) {
            const int ins = pk[k - 1];
            const int mis = pk[k];
            const int del = pk[k + 1];
            const int mid = ins > mis ? ins : mis;
            int mx = del > mid + 1 ? del : mid + 1;
            const int v = mx - k;
            const bool ok = mx >= 0 && mx <= n && v <= m;
            if (PACKED) {
                // invalid diagonals probe (0, 0) with lim 0 and are dropped
                const int lim = ok ? ((m - v) < (n - mx) ? (m - v) : (n - mx)) : 0;
                const int run = lds_match_run(sP, ok ? v : 0, sT, ok ? mx : 0, lim);
                mx = ok ? mx + run : OffNull<Off>::v;
            } else if (ok) {
                const int lim = (m - v) < (n - mx) ? (m - v) : (n - mx);
                mx += match_run(P + v, T + mx, lim);
            } else {
                mx = OffNull<Off>::v;
            }
            ck[k] = Off(mx);
            // (a 32-bit byte offset from the wave-uniform row base: one
            // scalar-base store, no per-lane 64-bit address)
            if (TRACE && trace)
                *reinterpret_cast<Off*>(reinterpret_cast<char*>(wrow) + uint32_t(k + s) * uint32_t(sizeof(Off))) = Off(mx);
        }
        if (tid >= NT - 4) ck[tid < NT - 2 ? lo - (NT - tid) + 2 : hi + (tid - (NT - 3))] = Off(OffNull<Off>::v);
        __syncthreads();
        if (ak >= lo && ak <= hi && int(ck[ak]) >= n) return s;
        Off* t = prev; prev = cur; cur = t;
        plo = lo;
    }
    return -1;
}

// The reference's traceback (wfa_kernel.hip:338-437) of one pair on one whole
// wave (wave-uniform state): from the end of diagonal n - m at `score`, over
// the pair's stack (row s at entry s^2, diagonal k at s + k), emitting the
// CIGAR right-aligned into cg[0 .. m + n).  Returns its length.  win:
// kTbRows x kTbCols entries of LDS for the wave's windows.
template <typename Off>
__device__ int wfa_walk(const Off* ws, Off* win, uint8_t* const cg, int m, int n, int score, int lane)
{
    const int ak = n - m;
    int pos = m + n;
    auto emit = [&](uint8_t c, int cnt) {
        if (cnt > pos) cnt = pos;
        if (cnt <= 0) return;
        for (int x = lane; x < cnt; x += 64) cg[pos - 1 - x] = c;
        pos -= cnt;
    };
    if (score == 0) {
        emit('M', n);
    } else {
        int sc = score, k = ak;
        int offset = ws_get(ws, score, k);
        int h = offset, v = offset - k;
        emit('D', v - m);
        emit('I', h - n);
        if (v > m) v = m;
        if (h > n) h = n;
        int top = 0, k0 = 0, d = kTbRows;
        while (v > 0 && h > 0 && sc > 0) {
            if (d == kTbRows) {
                // rows sc-1 .. sc-kTbRows, diagonals k-kTbRows .. k+kTbRows
                top = sc;
                k0 = k;
                d = 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                // kTbBatch unconditional loads in flight at a time (an
                // entry outside the stack reads entry 0 and is replaced
                // by NULL), not one load per round trip
#pragma unroll 1
                for (int e0 = lane; e0 < kTbRows * kTbCols; e0 += 64 * kTbBatch) {
                    Off x[kTbBatch];
#pragma unroll
                    for (int u = 0; u < kTbBatch; ++u) {
                        const int e = e0 + 64 * u;
                        const int r = top - 1 - e / kTbCols;
                        const int c = k0 - kTbRows + e % kTbCols;
                        const bool ok = e < kTbRows * kTbCols && r >= 0 && c >= -r && c <= r;
                        x[u] = ws[ok ? uint32_t(r * r + c + r) : 0u];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < kTbBatch; ++u) {
                        const int e = e0 + 64 * u;
                        const int r = top - 1 - e / kTbCols;
                        const int c = k0 - kTbRows + e % kTbCols;
                        const bool ok = r >= 0 && c >= -r && c <= r;
                        if (e < kTbRows * kTbCols) win[e] = ok ? x[u] : Off(OffNull<Off>::v);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            const Off* wr = win + d * kTbCols + (k - k0 + kTbRows);
            const int misms = int(wr[0]) + 1;
            const int ins = int(wr[-1]) + 1;
            const int del = int(wr[1]);
            int best = kWfaNeg, op = 0;
            if (misms >= 0 && misms >= best) { best = misms; op = 1; }
            if (ins >= 0 && ins > best) { best = ins; op = 2; }
            if (del >= 0 && del > best) { best = del; op = 3; }
            if (op == 0) break;
            // the run of matches down to the predecessor, then the edit, in
            // one masked store loop (two when the slot would overflow, which
            // consistent stacks never do)
            const int run = offset - best > 0 ? offset - best : 0;
            const uint8_t oc = op == 1 ? uint8_t('X') : op == 2 ? uint8_t('I') : uint8_t('D');
            if (run + 1 <= pos) {
                for (int x = lane; x <= run; x += 64) cg[pos - 1 - x] = x < run ? uint8_t('M') : oc;
                pos -= run + 1;
            } else {
                emit('M', run);
                emit(oc, 1);
            }
            offset = best - (op != 3 ? 1 : 0);
            k += (op == 3 ? 1 : 0) - (op == 2 ? 1 : 0);
            v = offset - k;
            h = offset;
            --sc;
            ++d;
        }
        if (v > 0 && h > 0) {
            const int nm = v < h ? v : h;
            emit('M', nm);
            v -= nm;
            h -= nm;
        }
        emit('D', v);
        emit('I', h);
    }
    return m + n - pos;
}

// Eight waves per SIMD: <= 64 VGPRs and <= 96 SGPRs, so four 8-wave
// workgroups (or 32 single waves) fit a CU when LDS allows.
template <bool TRACE, typename Off, bool STAGED, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(8, 8)))
void wfa_kernel(WfaQueueLaunch L)
{
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr int NT = 64 * WAVES;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int S = L.smax_cls;
    const int wfb = wfa_wf_bytes(S, int(sizeof(Off)));
    int* claim = reinterpret_cast<int*>(smem);
    const int pbytes = STAGED ? wfa_seq_bytes(L.max_plen) : 0;
    const int tbytes = STAGED ? wfa_seq_bytes(L.max_tlen) : 0;
    uint32_t* sP = reinterpret_cast<uint32_t*>(smem + 16);
    uint32_t* sT = reinterpret_cast<uint32_t*>(smem + 16 + pbytes);
    uint8_t* wfs = smem + 16 + pbytes + tbytes;
    Off* A = reinterpret_cast<Off*>(wfs);
    Off* B = reinterpret_cast<Off*>(wfs + wfb);
    // the workgroup's stack region: ws_stride entries (at least one worst-case
    // stack of (S + 1)^2 entries; more when walks are deferred)
    const int64_t stride = L.ws_stride > 0 ? L.ws_stride : (int64_t(S) + 1) * (int64_t(S) + 1);
    Off* ws = TRACE ? static_cast<Off*>(L.ws) + int64_t(blockIdx.x) * stride : nullptr;
    // Deferred walks (several waves per workgroup): each traced pair's stack is
    // left in the region and its walk queued in LDS; when the region cannot
    // hold the next pair's worst case (or the queue is full) all waves walk
    // the queued pairs, wave w taking pairs w, w + WAVES, ..., each with a
    // window in the wavefront buffers (free between pairs).  So no wave waits
    // while one walks.
    const int winb = wfa_win_bytes(int(sizeof(Off)));
    constexpr bool defer = TRACE && WAVES > 1;
    const int region = wfa_region_bytes(S, int(sizeof(Off)), defer, WAVES);
    Off* win = reinterpret_cast<Off*>(wfs + region);
    int* pend = reinterpret_cast<int*>(wfs + region + winb);
    int npend = 0;
    int64_t used = 0;
    auto walk_pending = [&]() {
        Off* wwin = reinterpret_cast<Off*>(wfs + wave * winb);
        for (int i = wave; i < npend; i += WAVES) {
            const int pid = pend[3 * i];
            const int64_t off = int64_t(uint32_t(pend[3 * i + 1]));
            const int sc = pend[3 * i + 2];
            const int len = wfa_walk<Off>(ws + off, wwin, L.cigar + L.cig_off[pid], L.plen[pid], L.tlen[pid], sc,
                                          lane);
            if (lane == 0) L.cig_len[pid] = len;
        }
        npend = 0;
        used = 0;
    };
    const int G = int(gridDim.x);
    // class size and the start of its slice of the claim order: from the launch,
    // or (device-planned batches) from the plan the device wrote
    int n_cls = L.n;
    const int32_t* order = L.order;
    if (L.dev_class) {
        n_cls = __builtin_amdgcn_readfirstlane(L.dev_class[1]);
        order = L.order + __builtin_amdgcn_readfirstlane(L.dev_class[0]);
    }
    // Static rounds first: in round r workgroup g takes pair r*G + g; the
    // counter then starts at rounds*G.  Claims on the one counter serialize
    // (≈ 12 ns each on this device), which caps a launch of short pairs of the
    // smallest class (smax <= 64, sequences <= 1024): there a class with at
    // most 8 pairs per workgroup is scheduled statically throughout and a
    // larger one keeps only its last two rounds' worth of pairs for the
    // counter.  Other launches take one static round: their pairs' costs differ
    // by orders of magnitude (lengths up to the launch maximum) and a claim
    // costs little next to a pair, so dynamic claims balance them.
    const bool short_pairs = S <= 64 && L.max_plen <= 1024 && L.max_tlen <= 1024;
    const int rounds = !short_pairs ? 1 : n_cls <= 8 * G ? (n_cls + G - 1) / G : n_cls / G - 2;
    for (int it = 0;; ++it) {
        // The dynamic claim has no branch on the lane id: wave 0 takes part in
        // one atomic adding (lane == 0) and lane 0's result is read into a
        // scalar register.  (With `if (lane == 0) q = atomicAdd(..)` and a
        // shuffle, the compiler may thread the loop-invariant branch into the
        // loop and leave lanes 1..63 looping on a stale claim while lane 0 is
        // masked off.)
        int q;
        if (it < rounds) {
            q = int(blockIdx.x) + it * G;
        } else if (n_cls <= rounds * G) {
            break;
        } else if (WAVES == 1) {
            q = rounds * G + __builtin_amdgcn_readfirstlane(atomicAdd(L.next, lane == 0 ? 1 : 0));
        } else {
            if (wave == 0) claim[it & 1] = __builtin_amdgcn_readfirstlane(atomicAdd(L.next, lane == 0 ? 1 : 0));
            __syncthreads();
            q = rounds * G + __builtin_amdgcn_readfirstlane(claim[it & 1]);
        }
        if (q >= n_cls) break;                    // every wave ends here once the queue is empty
        const int pid = order ? order[q] : q;

        // smax < 0: the pair is not in this pass (a launch without a plan runs
        // every pair in index order; the plans leave such pairs out)
        if (L.smax[pid] < 0) continue;
        const int m = L.plen[pid];
        const int n = L.tlen[pid];
        const int smax = L.smax[pid] < S ? L.smax[pid] : S;
        const uint8_t* P = L.pat + L.poff[pid];
        const uint8_t* T = L.txt + L.toff[pid];
        const bool trace = TRACE && (L.want == nullptr || L.want[pid] != 0);
        if (m <= 0 || n <= 0) {
            const int mm = m > 0 ? m : 0, nn = n > 0 ? n : 0;
            if (tid == 0) L.score[pid] = (m <= 0) ? nn : mm;
            if (trace) {
                uint8_t* cg = L.cigar + L.cig_off[pid];
                const uint8_t c = (m <= 0) ? uint8_t('I') : uint8_t('D');
                for (int x = tid; x < mm + nn; x += NT) cg[x] = c;
                if (tid == 0) L.cig_len[pid] = mm + nn;
            } else if (TRACE && L.cig_len && tid == 0) {
                L.cig_len[pid] = 0;
            }
            continue;
        }
        // the end diagonal n - m is first reached at score |n - m|: a pair whose
        // length difference exceeds its bound is unresolved without a pass
        if ((n > m ? n - m : m - n) > smax) {
            if (tid == 0) {
                L.score[pid] = -1;
                if (TRACE && L.cig_len) L.cig_len[pid] = 0;
            }
            continue;
        }
        Off* wsp = ws;
        if (defer && trace) {
            // room for this pair's worst case: (smax + 1)^2 entries
            const int64_t need = ((int64_t(smax) + 1) * (smax + 1) + 1) & ~int64_t(1);
            if (used + need > stride || npend == kWfaPend) {
                walk_pending();
                __syncthreads();                   // the walks' windows live in A and B
            }
            wsp = ws + used;
        }
        // STAGED: 2-bit copies in LDS unless a byte other than A/C/G/T occurs
        bool packed = false;
        if (STAGED) {
            const bool ok = stage2<NT>(sP, P, m, tid) & stage2<NT>(sT, T, n, tid);
            packed = __syncthreads_and(ok);
        } else {
            __syncthreads();
        }
        const int ak = n - m;
        const int score = packed ? wfa_forward<TRACE, Off, true, NT>(sP, sT, P, T, m, n, smax, A, B, wsp, trace, tid)
                                 : wfa_forward<TRACE, Off, false, NT>(sP, sT, P, T, m, n, smax, A, B, wsp, trace, tid);
        if (tid == 0) L.score[pid] = score;
        if (TRACE && wave == 0) {
            if (!trace || score < 0) {
                if (lane == 0 && L.cig_len) L.cig_len[pid] = 0;
            } else if (defer && score > 0) {
                // walked later with other pairs of this workgroup (all waves)
                if (lane == 0) {
                    pend[3 * npend] = pid;
                    pend[3 * npend + 1] = int(uint32_t(used));
                    pend[3 * npend + 2] = score;
                }
            } else {
                const int len = wfa_walk<Off>(wsp, win, L.cigar + L.cig_off[pid], m, n, score, lane);
                if (lane == 0) L.cig_len[pid] = len;
            }
        }
        if (defer && trace && score > 0) {
            // every wave keeps the same region bookkeeping (uniform values)
            ++npend;
            used += ((int64_t(score) + 1) * (score + 1) + 1) & ~int64_t(1);
        }
        // the next pair rewrites the LDS wavefronts, sequences and the stack
        __syncthreads();
    }
    if (defer) walk_pending();
}

__global__ void zero_i32_kernel(int32_t* p, int n)
{
    for (int i = int(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += int(gridDim.x * blockDim.x)) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Device-side planning: size class by smax, claim order by length (longest
// first) as a counting sort over kPlanKeys length bins of 32 bytes per class:
// histogram, one-workgroup exclusive scan, scatter.  Order inside a bin is
// whatever the atomics give; results do not depend on it.  A pair whose smax
// is negative is left out of the plan (the router's passes use this).
__device__ __forceinline__ int plan_class(int smax)
{
    return smax <= kWfaClassBound0 ? 0 : smax <= kWfaClassBound1 ? 1 : smax <= kWfaClassBound2 ? 2 : 3;
}
__device__ __forceinline__ int plan_key(int m, int n)
{
    const int l = ((m > 0 ? m : 0) + (n > 0 ? n : 0)) >> 5;
    return kPlanKeys - 1 - (l < kPlanKeys - 1 ? l : kPlanKeys - 1);
}

// Large batches: each workgroup counts kPlanChunk pairs into an LDS histogram
// and adds its non-zero bins to the global one (a global atomic per pair
// serialised on the few bins most pairs share).
constexpr int kPlanChunk = 4096;
constexpr int kPlanBins = kWfaClasses * kPlanKeys;

__device__ __forceinline__ int plan_bin(const WfaDevPlan& P, int i)
{
    return plan_class(P.smax[i]) * kPlanKeys + plan_key(P.plen[i], P.tlen[i]);
}

__global__ __launch_bounds__(256) void plan_hist_kernel(WfaDevPlan P)
{
    __shared__ int hist[kPlanBins];
    const int t = threadIdx.x;
    for (int b = t; b < kPlanBins; b += 256) hist[b] = 0;
    __syncthreads();
    const int i0 = int(blockIdx.x) * kPlanChunk;
    const int i1 = i0 + kPlanChunk < P.n ? i0 + kPlanChunk : P.n;
    for (int i = i0 + t; i < i1; i += 256)
        if (P.smax[i] >= 0) atomicAdd(&hist[plan_bin(P, i)], 1);   // smax < 0: not in this plan
    __syncthreads();
    for (int b = t; b < kPlanBins; b += 256)
        if (hist[b]) atomicAdd(&P.hist[b], hist[b]);
}

// hist (kWfaClasses * kPlanKeys counts) -> exclusive offsets; cls[2c] = first
// pair of class c in the order, cls[2c+1] = its size
__global__ __launch_bounds__(1024) void plan_scan_kernel(WfaDevPlan P)
{
    __shared__ int part[1024];
    constexpr int kPer = kWfaClasses * kPlanKeys / 1024;
    const int t = threadIdx.x;
    int v[kPer];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) { v[q] = P.hist[t * kPer + q]; sum += v[q]; }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {             // inclusive Hillis-Steele scan
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int run = part[t] - sum;                                // exclusive prefix of this thread
#pragma unroll
    for (int q = 0; q < kPer; ++q) { P.hist[t * kPer + q] = run; run += v[q]; }
    // class c spans bins [c*K, (c+1)*K): its first offset is that of bin c*K
    __syncthreads();
    if (t < kWfaClasses) {
        const int first = P.hist[t * kPlanKeys];
        const int end = t + 1 < kWfaClasses ? P.hist[(t + 1) * kPlanKeys] : part[1023];
        P.cls[2 * t] = first;
        P.cls[2 * t + 1] = end - first;
    }
}

// The whole plan in one workgroup, histogram in LDS (batches up to
// kPlanSmallMax pairs): one launch instead of four.  Also zeroes the classes'
// claim counters.
__global__ __launch_bounds__(1024) void plan_small_kernel(WfaDevPlan P)
{
    __shared__ int hist[kWfaClasses * kPlanKeys];
    __shared__ int part[1024];
    constexpr int kPer = kWfaClasses * kPlanKeys / 1024;
    const int t = threadIdx.x;
    for (int b = t; b < kWfaClasses * kPlanKeys; b += 1024) hist[b] = 0;
    if (t < kWfaClasses) P.next[t] = 0;
    __syncthreads();
    for (int i = t; i < P.n; i += 1024)
        if (P.smax[i] >= 0) atomicAdd(&hist[plan_class(P.smax[i]) * kPlanKeys + plan_key(P.plen[i], P.tlen[i])], 1);
    __syncthreads();
    int v[kPer];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) { v[q] = hist[t * kPer + q]; sum += v[q]; }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int run = part[t] - sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q) { hist[t * kPer + q] = run; run += v[q]; }
    __syncthreads();
    if (t < kWfaClasses) {
        const int first = hist[t * kPlanKeys];
        const int end = t + 1 < kWfaClasses ? hist[(t + 1) * kPlanKeys] : part[1023];
        P.cls[2 * t] = first;
        P.cls[2 * t + 1] = end - first;
    }
    __syncthreads();
    for (int i = t; i < P.n; i += 1024) {
        if (P.smax[i] < 0) continue;
        const int pos = atomicAdd(&hist[plan_class(P.smax[i]) * kPlanKeys + plan_key(P.plen[i], P.tlen[i])], 1);
        P.order[pos] = i;
    }
}

// The same chunks again: each workgroup reserves one range per non-zero bin
// with a global atomic, then places its pairs in it through LDS counters.
__global__ __launch_bounds__(256) void plan_scatter_kernel(WfaDevPlan P)
{
    __shared__ int hist[kPlanBins];
    const int t = threadIdx.x;
    for (int b = t; b < kPlanBins; b += 256) hist[b] = 0;
    __syncthreads();
    const int i0 = int(blockIdx.x) * kPlanChunk;
    const int i1 = i0 + kPlanChunk < P.n ? i0 + kPlanChunk : P.n;
    for (int i = i0 + t; i < i1; i += 256)
        if (P.smax[i] >= 0) atomicAdd(&hist[plan_bin(P, i)], 1);
    __syncthreads();
    for (int b = t; b < kPlanBins; b += 256)
        if (hist[b]) hist[b] = atomicAdd(&P.hist[b], hist[b]);
    __syncthreads();
    for (int i = i0 + t; i < i1; i += 256)
        if (P.smax[i] >= 0) P.order[atomicAdd(&hist[plan_bin(P, i)], 1)] = i;
}

}  // namespace

}  // namespace ga

namespace ga {

int device_cus()
{
    // per device, looked up once (threads racing on the first lookup store the
    // same value)
    static std::atomic<int> cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int c = cus[dev].load(std::memory_order_relaxed);
    if (c == 0) {
        hipDeviceProp_t p{};
        c = (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) ? p.multiProcessorCount
                                                                                        : 256;
        cus[dev].store(c, std::memory_order_relaxed);
    }
    return c;
}

namespace {

template <bool TRACE, typename Off, bool STAGED, int WAVES>
hipError_t launch_one(const WfaQueueLaunch& L, size_t lds, hipStream_t st)
{
    auto k = wfa_kernel<TRACE, Off, STAGED, WAVES>;
    hipLaunchKernelGGL(k, dim3(unsigned(L.n_groups)), dim3(64 * WAVES), lds, st, L);
    return hipGetLastError();
}

template <bool TRACE, typename Off, bool STAGED, int WAVES>
int resident_groups(size_t lds)
{
    int nb = 0;
    auto k = wfa_kernel<TRACE, Off, STAGED, WAVES>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 64 * WAVES, lds) != hipSuccess || nb < 1) nb = 1;
    return nb;
}

template <bool TRACE, typename Off, bool STAGED>
hipError_t dispatch_waves(const WfaQueueLaunch& L, size_t lds, hipStream_t st, int* groups)
{
    switch (L.waves) {
    case 1: if (groups) { *groups = resident_groups<TRACE, Off, STAGED, 1>(lds); return hipSuccess; }
            return launch_one<TRACE, Off, STAGED, 1>(L, lds, st);
    case 2: if (groups) { *groups = resident_groups<TRACE, Off, STAGED, 2>(lds); return hipSuccess; }
            return launch_one<TRACE, Off, STAGED, 2>(L, lds, st);
    case 4: if (groups) { *groups = resident_groups<TRACE, Off, STAGED, 4>(lds); return hipSuccess; }
            return launch_one<TRACE, Off, STAGED, 4>(L, lds, st);
    default: if (groups) { *groups = resident_groups<TRACE, Off, STAGED, 8>(lds); return hipSuccess; }
            return launch_one<TRACE, Off, STAGED, 8>(L, lds, st);
    }
}

// groups != null: report resident workgroups per CU instead of launching
hipError_t dispatch(const WfaQueueLaunch& L, size_t lds, hipStream_t st, int* groups)
{
    const bool tr = L.cigar != nullptr;
    if (L.wide) {
        if (tr) return L.staged ? dispatch_waves<true, int32_t, true>(L, lds, st, groups)
                                : dispatch_waves<true, int32_t, false>(L, lds, st, groups);
        return L.staged ? dispatch_waves<false, int32_t, true>(L, lds, st, groups)
                        : dispatch_waves<false, int32_t, false>(L, lds, st, groups);
    }
    if (tr) return L.staged ? dispatch_waves<true, int16_t, true>(L, lds, st, groups)
                            : dispatch_waves<true, int16_t, false>(L, lds, st, groups);
    return L.staged ? dispatch_waves<false, int16_t, true>(L, lds, st, groups)
                    : dispatch_waves<false, int16_t, false>(L, lds, st, groups);
}

}  // namespace

size_t wfa_lds_bytes(const WfaQueueLaunch& L)
{
    const bool defer = L.cigar && L.waves > 1;
    size_t b = 16 + size_t(wfa_region_bytes(L.smax_cls, L.wide ? 4 : 2, defer, L.waves)) +
               size_t(wfa_win_bytes(L.wide ? 4 : 2));
    if (defer) b += size_t(kWfaPend) * 12;                          // the deferred-walk queue
    if (L.staged) b += size_t(wfa_seq_bytes(L.max_plen)) + size_t(wfa_seq_bytes(L.max_tlen));
    return b;
}

int wfa_queue_groups(const WfaQueueLaunch& L)
{
    // resident workgroups per CU, cached per (device, variant, LDS bytes)
    struct Entry { int dev, variant; size_t lds; int per_cu; };
    static std::mutex mu;
    static std::vector<Entry> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int variant = (L.cigar ? 1 : 0) | (L.wide ? 2 : 0) | (L.staged ? 4 : 0) | (L.waves << 3);
    const size_t lds = wfa_lds_bytes(L);
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const Entry& e : cache)
            if (e.dev == dev && e.variant == variant && e.lds == lds) { per_cu = e.per_cu; break; }
    }
    if (per_cu == 0) {
        per_cu = 1;
        (void)dispatch(L, lds, nullptr, &per_cu);
        std::lock_guard<std::mutex> lk(mu);
        cache.push_back({dev, variant, lds, per_cu});
    }
    const int64_t g = int64_t(device_cus()) * per_cu;
    return int(g < L.n ? g : L.n);
}

size_t wfa_queue_ws_bytes(int smax_cls, int n_groups, bool wide)
{
    return size_t(n_groups) * size_t(smax_cls + 1) * size_t(smax_cls + 1) * (wide ? 4u : 2u);
}

hipError_t wfa_queue_launch(const WfaQueueLaunch& L, hipStream_t st)
{
    if (L.n <= 0 || L.n_groups <= 0) return hipSuccess;
    // the claim counter is zeroed by a kernel, not a memset node, so this
    // submission can be captured into a graph and replayed (device-planned
    // classes: the plan kernels zero it)
    if (!L.dev_class && !L.counter_zeroed && !wfa_fully_static(L)) {
        hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, st, L.next, 1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return dispatch(L, wfa_lds_bytes(L), st, nullptr);
}

hipError_t zero_i32_launch(int32_t* p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(unsigned((n + 255) / 256 < 64 ? (n + 255) / 256 : 64)), dim3(256), 0,
                       st, p, n);
    return hipGetLastError();
}

hipError_t wfa_plan_launch(const WfaDevPlan& P, hipStream_t st)
{
    if (P.n <= 0) return hipSuccess;
    if (P.n <= kPlanSmallMax) {
        hipLaunchKernelGGL(plan_small_kernel, dim3(1), dim3(1024), 0, st, P);
        return hipGetLastError();
    }
    const unsigned nb = unsigned((P.n + kPlanChunk - 1) / kPlanChunk);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, st, P.next, kWfaClasses);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(16), dim3(256), 0, st, P.hist, kWfaClasses * kPlanKeys);
    hipLaunchKernelGGL(plan_hist_kernel, dim3(nb), dim3(256), 0, st, P);
    hipLaunchKernelGGL(plan_scan_kernel, dim3(1), dim3(1024), 0, st, P);
    hipLaunchKernelGGL(plan_scatter_kernel, dim3(nb), dim3(256), 0, st, P);
    return hipGetLastError();
}

}  // namespace ga
