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
// Mapping to CDNA4: one 64-lane wavefront per pair, one workgroup of exactly one
// wave, so __syncthreads() costs a single-wave barrier.  Lane q owns diagonals
// k = -s + q, -s + q + 64, ... of wavefront s.  The two live wavefronts sit in
// LDS (ping-pong, 2*smax+3 slots each, as the reference lays them out); in the
// trace variant every wavefront is also stored to a per-pair HBM stack (row s
// at offset s*s, 2s+1 ints, coalesced across lanes) that the traceback reads
// back.  The extension compares 8 bytes per step with a 64-bit XOR and a
// count-trailing-zeros, over raw bytes, so every byte value keeps its literal
// meaning (N matches only N, case is significant).  The traceback runs in
// lockstep on all 64 lanes (its state is uniform) and writes each run of M's
// cooperatively, 64 bytes per store instruction.
#include "ga_device.hpp"

namespace ga {
namespace {

__device__ __forceinline__ bool reachable(int off) { return off > kWfaNeg + 1024; }

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
// Persistent, queue-fed kernel.  One wave aligns one pair at a time and claims
// the next from an atomic counter, so a class of pairs of very different cost
// (lengths 100 bp .. 10 kbp, distances up to 2048) keeps every wave busy until
// the queue drains; the host orders each class longest first.  The wavefront
// stack is per wave, reused from pair to pair: row s at offset s*s, 2s+1
// entries, 16-bit offsets when every text is shorter than 32767 bytes (the
// offset of a reachable diagonal never exceeds the text length), else 32-bit.
template <typename Off> struct OffNull;
template <> struct OffNull<int16_t> { static constexpr int16_t v = INT16_MIN; };
template <> struct OffNull<int32_t> { static constexpr int32_t v = kWfaNeg; };

template <typename Off>
__device__ __forceinline__ int ws_get(const Off* ws, int s, int k)
{
    if (k < -s || k > s) return kWfaNeg;
    const int v = ws[int64_t(s) * s + k + s];
    return v == OffNull<Off>::v ? kWfaNeg : v;
}

template <bool TRACE, typename Off>
__device__ void wfa_pair_q(const WfaQueueLaunch& L, int pid, int lane, int* lds, Off* ws)
{
    const int m = L.plen[pid];
    const int n = L.tlen[pid];
    const int smax = L.smax[pid] < L.smax_cls ? L.smax[pid] : L.smax_cls;
    const uint8_t* P = L.pat + L.poff[pid];
    const uint8_t* T = L.txt + L.toff[pid];
    const bool trace = TRACE && (L.want == nullptr || L.want[pid] != 0);
    if (m <= 0 || n <= 0) {
        const int mm = m > 0 ? m : 0, nn = n > 0 ? n : 0;
        if (lane == 0) L.score[pid] = (m <= 0) ? nn : mm;
        if (trace) {
            uint8_t* cg = L.cigar + L.cig_off[pid];
            const uint8_t c = (m <= 0) ? uint8_t('I') : uint8_t('D');
            for (int q = lane; q < mm + nn; q += 64) cg[q] = c;
            if (lane == 0) L.cig_len[pid] = mm + nn;
        } else if (TRACE && L.cig_len && lane == 0) {
            L.cig_len[pid] = 0;
        }
        return;
    }
    const int S = L.smax_cls;
    const int W = 2 * S + 3;
    int* A = lds;
    int* B = lds + W;
    for (int q = lane; q < 2 * W; q += 64) lds[q] = kWfaNeg;
    const int o0 = match_run(P, T, m < n ? m : n);
    __syncthreads();
    if (lane == 0) {
        A[S + 1] = o0;
        if (trace) ws[0] = Off(o0);
    }
    __syncthreads();
    const int ak = n - m;
    int score = -1;
    if (ak == 0 && o0 >= n) {
        score = 0;
    } else {
        int* prev = A;
        int* cur = B;
        for (int s = 1; s <= smax; ++s) {
            for (int base = -s; base <= s; base += 64) {
                const int k = base + lane;
                if (k <= s) {
                    const int ins = prev[k + S];
                    const int mis = prev[k + S + 1];
                    const int del = prev[k + S + 2];
                    const int mid = ins > mis ? ins : mis;
                    int mx = del > mid + 1 ? del : mid + 1;
                    if (!reachable(mx) || mx > n || mx - k > m) {
                        mx = kWfaNeg;
                    } else {
                        const int v = mx - k;
                        const int lim = (m - v) < (n - mx) ? (m - v) : (n - mx);
                        mx += match_run(P + v, T + mx, lim);
                    }
                    cur[k + S + 1] = mx;
                    if (trace) ws[int64_t(s) * s + k + s] = reachable(mx) ? Off(mx) : OffNull<Off>::v;
                }
            }
            __syncthreads();
            if (ak >= -s && ak <= s) {
                const int off = cur[ak + S + 1];
                if (reachable(off) && off >= n) { score = s; break; }
            }
            int* t = prev; prev = cur; cur = t;
        }
    }
    if (lane == 0) L.score[pid] = score;
    if (!TRACE) return;
    if (!trace || score < 0) {
        if (lane == 0 && L.cig_len) L.cig_len[pid] = 0;
        __syncthreads();
        return;
    }
    __syncthreads();
    uint8_t* cg = L.cigar + L.cig_off[pid];
    int pos = m + n;
    auto emit = [&](uint8_t c, int cnt) {
        if (cnt > pos) cnt = pos;
        if (cnt <= 0) return;
        for (int q = lane; q < cnt; q += 64) cg[pos - 1 - q] = c;
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
        while (v > 0 && h > 0 && sc > 0) {
            const int sp = sc - 1;
            const int misms = ws_get(ws, sp, k) + 1;
            const int ins = ws_get(ws, sp, k - 1) + 1;
            const int del = ws_get(ws, sp, k + 1);
            int best = kWfaNeg, op = 0;
            if (reachable(misms) && misms >= best) { best = misms; op = 1; }
            if (reachable(ins) && ins > best) { best = ins; op = 2; }
            if (reachable(del) && del > best) { best = del; op = 3; }
            if (op == 0) break;
            emit('M', offset - best);
            offset = best;
            if (op == 1)      { emit('X', 1); --offset; }
            else if (op == 2) { emit('I', 1); --k; --offset; }
            else              { emit('D', 1); ++k; }
            v = offset - k;
            h = offset;
            --sc;
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
    if (lane == 0) L.cig_len[pid] = m + n - pos;
    __syncthreads();
}

__global__ void zero_i32_kernel(int32_t* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

template <bool TRACE, typename Off>
__global__ __launch_bounds__(64) void wfa_queue_kernel(WfaQueueLaunch L)
{
    extern __shared__ int lds[];
    const int lane = threadIdx.x;
    Off* ws = TRACE ? static_cast<Off*>(L.ws) + int64_t(blockIdx.x) * (int64_t(L.smax_cls) + 1) * (int64_t(L.smax_cls) + 1)
                    : nullptr;
    for (;;) {
        // Claim without a branch on the lane id: every lane takes part in one
        // atomic adding (lane == 0), and lane 0's result is read into a scalar
        // register.  (With `if (lane == 0) q = atomicAdd(..)` and a shuffle,
        // the compiler may thread the loop-invariant branch into the loop and
        // leave lanes 1..63 looping on a stale claim while lane 0 is masked off.)
        const int mine = atomicAdd(L.next, lane == 0 ? 1 : 0);
        const int q = __builtin_amdgcn_readfirstlane(mine);
        if (q >= L.n) break;                      // every wave ends here once the queue is empty
        const int pid = L.order ? L.order[q] : q;
        wfa_pair_q<TRACE, Off>(L, pid, lane, lds, ws);
    }
}

}  // namespace

}  // namespace ga

namespace ga {

int wfa_queue_waves(int smax_cls, int n)
{
    // LDS per wave: two wavefronts of 2*smax+3 ints; at most 16 waves per CU.
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p{};
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    const size_t lds = size_t(2) * size_t(2 * smax_cls + 3) * sizeof(int);
    int per_cu = int((160 * 1024) / (lds > 0 ? lds : 1));
    if (per_cu > 16) per_cu = 16;
    if (per_cu < 1) per_cu = 1;
    const int waves = cus * per_cu;
    return n < waves ? n : waves;
}

size_t wfa_queue_ws_bytes(int smax_cls, int n_waves, bool wide)
{
    return size_t(n_waves) * size_t(smax_cls + 1) * size_t(smax_cls + 1) * (wide ? 4u : 2u);
}

hipError_t wfa_queue_launch(const WfaQueueLaunch& L, hipStream_t st)
{
    if (L.n <= 0 || L.n_waves <= 0) return hipSuccess;
    const size_t lds = size_t(2) * size_t(2 * L.smax_cls + 3) * sizeof(int);
    // the claim counter is zeroed by a kernel, not a memset node, so this
    // submission can be captured into a graph and replayed
    hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, st, L.next, 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid(unsigned(L.n_waves)), block(64);
    if (!L.cigar)     hipLaunchKernelGGL((wfa_queue_kernel<false, int16_t>), grid, block, lds, st, L);
    else if (L.wide)  hipLaunchKernelGGL((wfa_queue_kernel<true, int32_t>), grid, block, lds, st, L);
    else              hipLaunchKernelGGL((wfa_queue_kernel<true, int16_t>), grid, block, lds, st, L);
    return hipGetLastError();
}

}  // namespace ga
