// sw.hip -- gfx950 kernels for local alignment with affine gaps (Gotoh
// Smith-Waterman): score, end coordinate, and (trace variant) a 4-bit direction
// code per cell in HBM that a walk kernel turns into span + CIGAR.
//
// Recurrence and tie rules (reference: include/genoaligner/backend/
// sw_kernel_impl.hip:292-322, sw_kernel.hip:76-91), text = row axis i,
// pattern = column axis j, both 1-based in the matrix:
//   E[i][j] = max(H[i][j-1] - go, E[i][j-1] - ge)   extend only if strictly greater
//   F[i][j] = max(H[i-1][j] - go, F[i-1][j] - ge)   extend only if strictly greater
//   H[i][j] = diag + s, then E if E > H, then F if F > H; H <= 0 -> 0 (STOP)
//   boundaries H = E = F = 0 on row 0 and column 0 (:286-296)
//   best = first strict maximum in row-major order (:320)
// The direction code is the reference's byte (hsrc | EEXT 4 | FEXT 8) kept in
// a nibble, and the walk is the reference's three-state walk (:198-242).
//
// Mapping to CDNA4: one wave64 per pair.  The pattern is cut into strips of 64
// columns; in a strip lane l owns column j = 64*strip + l + 1 and the text is
// streamed through the wave with a one-step skew per lane, so at step t lane l
// computes row i = t - l + 1.  Everything a cell needs from its left neighbour
// (H and E of the same row, and H of the row above) is what lane l-1 held one
// and two steps earlier, moved by one cross-lane shift per step; F and H of the
// row above stay in the lane's own registers.  Lane 0 takes its left column
// from a boundary buffer that lane 63 of the previous strip filled (in LDS when
// it fits).  Each lane packs 8 direction nibbles per 32-bit word and stores one
// word every 8 steps; a wave's 64 words form one contiguous 256-byte store.
//
// End coordinates.  Each lane keeps the first strict maximum of its own columns
// in (row, strip) order.  That per-lane record is exactly the reference score
// kernel's per-lane record at warp width 64 (its lane (j-1) % 64 owns the same
// columns), so the score-only end follows its xor-butterfly argmax
// (sw_kernel_impl.hip:146,161-167); the trace end is the row-major minimum over
// lanes holding the maximum.
#include "ga_device.hpp"

#include <algorithm>
#include <type_traits>

namespace ga {
namespace {

// Re-run mode: the pair held by slot q of a packed launch (-1 when empty).
__device__ __forceinline__ int slot_pair(const SwLaunch& L, int q)
{
    const int w = q >> 1;
    int pid;
    if (L.pair_a) pid = (q & 1) ? L.pair_b[w] : L.pair_a[w];
    else          pid = L.first_pair + q;
    return pid < L.n_total ? pid : -1;
}

// First direction word of a pair: explicit offsets, or its packed slot's
// region in re-run mode, or a fixed stride per block.
__device__ __forceinline__ int64_t dir_base(const SwLaunch& L, int pid, int q)
{
    if (L.dir_off) return L.dir_off[pid];
    if (L.rerun) return int64_t(q >> 1) * L.dir_stride + int64_t(q & 1) * (L.dir_stride / 2);
    return int64_t(q) * L.dir_stride;
}

__device__ __forceinline__ uint64_t load8(const uint8_t* p)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint64_t* w = reinterpret_cast<const uint64_t*>(a & ~uintptr_t(7));
    const unsigned sh = unsigned(a & 7u) * 8u;
    const uint64_t lo = w[0];
    const uint64_t hi = w[1];
    return sh ? ((lo >> sh) | (hi << (64u - sh))) : lo;
}

constexpr int kSwProfileInts = 32 * 64;

// One lane's running state in a strip.
struct SwLane {
    int h, e, f;        // this lane's last row
    int hl_prev;        // H[i-1][j-1]
    int bHv, bIv;       // first strict maximum of the lane's column in this strip
};

// Eight steps t0 .. t0+7 of a strip (rows t0 - lane .. t0 - lane + 7).  Lane 0
// takes its left column from bHr/bEr (wave-uniform LDS or global reads), lane
// 63 hands its column to the next strip through bHw/bEw.  Returns the 8
// direction nibbles of the lane (TRACE).  FULL: every lane's 8 rows exist and
// every lane's column exists (t0 >= 63, t0 + 8 <= m, a full strip), so no
// row needs the activity test, and lane 63 hands its 8 rows over after them.
template <bool TRACE, bool MATRIX, class PT = int, bool FULL = false>
__device__ __forceinline__ uint32_t sw_chunk(SwLane& S, int t0, int lane, int m, bool jvalid, uint8_t pc,
                                             uint64_t tcur, const int* bHr, const int* bEr, int* bHw,
                                             int* bEw, const PT* prof, int go, int ge, int sm, int sx)
{
    uint32_t acc = 0;
    int hvs[8], ees[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int i = t0 + u - lane;         // 0-based row of this step
        // Left neighbour's H and E, one step old, by a whole-wave shift
        // right (DPP wave_shr:1); lane 0 receives the strip boundary
        // (row t0+u), read at a wave-uniform address.
        const int i0 = t0 + u;
        const int bh = i0 < m ? bHr[i0] : 0;
        const int be = i0 < m ? bEr[i0] : 0;
        const int hl = __builtin_amdgcn_update_dpp(bh, S.h, 0x138, 0xf, 0xf, false);
        const int el = __builtin_amdgcn_update_dpp(be, S.e, 0x138, 0xf, 0xf, false);
        const int diag = S.hl_prev;
        S.hl_prev = hl;
        const bool active = FULL || ((i >= 0) && (i < m) && jvalid);
        if (active) {
            const uint8_t tc = uint8_t(tcur >> (8 * u));
            int ee = hl - go;
            uint32_t code = 0;
            { const int e2 = el - ge; if (e2 > ee) { ee = e2; code |= 4u; } }
            int ff = S.h - go;
            { const int f2 = S.f - ge; if (f2 > ff) { ff = f2; code |= 8u; } }
            int hv = diag + (MATRIX ? int(prof[(tc & 31) * 64 + lane]) : (tc == pc ? sm : sx));
            uint32_t hsrc = 1u;
            if (ee > hv) { hv = ee; hsrc = 2u; }
            if (ff > hv) { hv = ff; hsrc = 3u; }
            if (hv <= 0) { hv = 0; hsrc = 0u; }
            S.h = hv; S.e = ee; S.f = ff;
            if (hv > S.bHv) { S.bHv = hv; S.bIv = i; }
            if (TRACE) acc |= (code | hsrc) << (4 * u);
            if (FULL) { hvs[u] = hv; ees[u] = ee; }
            else if (lane == 63) { bHw[i] = hv; bEw[i] = ee; }
        }
    }
    if (FULL && lane == 63) {
#pragma unroll
        for (int u = 0; u < 8; ++u) { bHw[t0 + u - 63] = hvs[u]; bEw[t0 + u - 63] = ees[u]; }
    }
    return acc;
}

// A strip's first maximum folded into the lane's record (strips in order:
// a later strip wins only with a greater value or the same value on an
// earlier row).
__device__ __forceinline__ void sw_fold(const SwLane& S, int j, int& gH, int& gI, int& gJ)
{
    if (S.bHv > gH || (S.bHv == gH && S.bHv > 0 && S.bIv < gI)) { gH = S.bHv; gI = S.bIv; gJ = j; }
}

// Both end cells from the per-lane records, and the result row of the pair.
__device__ __forceinline__ void sw_finish(int gH, int gI, int gJ, int lane, int32_t* res)
{
    // Score-only end: the reference's xor-butterfly argmax at width 64.
    int lH = gH, lI = gI, lJ = gJ;
    // Trace end: the row-major first cell among lanes holding the maximum.
    int rH = gH, rI = gI, rJ = gJ;
    for (int off = 32; off > 0; off >>= 1) {
        const int ob = __shfl_xor(lH, off), oi = __shfl_xor(lI, off), oj = __shfl_xor(lJ, off);
        if (ob > lH) { lH = ob; lI = oi; lJ = oj; }
        const int rb = __shfl_xor(rH, off), ri = __shfl_xor(rI, off), rj = __shfl_xor(rJ, off);
        if (rb > rH || (rb == rH && rb > 0 && (ri < rI || (ri == rI && rj < rJ)))) {
            rH = rb; rI = ri; rJ = rj;
        }
    }
    if (lane == 0) {
        res[0] = lH;          // lH == rH: both are the maximum
        res[1] = rI; res[2] = rJ;
        res[3] = lI; res[4] = lJ;
        res[5] = -1; res[6] = -1; res[7] = 0;
    }
}

template <bool TRACE, bool LDS_BND, bool MATRIX>
__global__ __launch_bounds__(64) void sw_kernel(SwLaunch L)
{
    extern __shared__ int lds[];
    const int slot = blockIdx.x;
    int pid = L.first_pair + slot;
    if (L.rerun) {
        pid = slot_pair(L, slot);
        if (pid < 0 || !L.rerun[pid]) return;
    }
    const int lane = threadIdx.x;
    const int m = L.tlen[pid];
    const int n = L.plen[pid];
    int32_t* res = L.res + int64_t(pid) * 8;
    if (m <= 0 || n <= 0) {
        if (lane == 0) {
            res[0] = 0;
            for (int q = 1; q < 7; ++q) res[q] = -1;
            res[7] = 0;
        }
        return;
    }
    const uint8_t* T = L.txt + L.toff[pid];
    const uint8_t* P = L.pat + L.poff[pid];
    int* bH = LDS_BND ? lds + (MATRIX ? kSwProfileInts : 0) : L.bnd + int64_t(slot) * 2 * L.max_m;
    // MATRIX: each lane's column of the score table for its pattern residue,
    // prof[code * 64 + lane] (one LDS read per cell, only by the owning lane)
    int* prof = lds;
    int* bE = bH + L.max_m;
    for (int q = lane; q < m; q += 64) { bH[q] = 0; bE[q] = 0; }
    __syncthreads();

    const int go = L.go, ge = L.ge, sm = L.match, sx = L.mismatch;
    const int nstrips = (n + 63) >> 6;
    const int nchunks = (m + 63 + 7) >> 3;          // 8-step chunks per strip
    uint32_t* dirs = nullptr;
    if (TRACE) dirs = L.dirs + dir_base(L, pid, slot);
    int gH = 0, gI = -1, gJ = -1;
    // text bytes (or residue codes) for rows t - lane .. t - lane + 7
    const uint8_t* TS = MATRIX ? L.tcode + L.toff[pid] : T;

    for (int s = 0; s < nstrips; ++s) {
        const int j = (s << 6) + lane;               // 0-based column of this lane
        const bool jvalid = j < n;
        const bool full_strip = (s << 6) + 64 <= n;
        const uint8_t pc = jvalid ? (MATRIX ? L.pcode[L.poff[pid] + j] : P[j]) : uint8_t(0);
        if (MATRIX)
            for (int c = 0; c < 32; ++c) prof[c * 64 + lane] = L.sub[c * 32 + pc];
        SwLane S{0, 0, 0, 0, 0, -1};
        uint64_t tcur = load8(TS - lane);            // prefetched a chunk ahead
        for (int c = 0; c < nchunks; ++c) {
            const int t0 = c << 3;
            const uint64_t tnext = load8(TS + t0 + 8 - lane);
            const uint32_t acc = (t0 >= 63 && t0 + 8 <= m && full_strip)
                ? sw_chunk<TRACE, MATRIX, int, true>(S, t0, lane, m, jvalid, pc, tcur, bH, bE, bH, bE,
                                                     prof, go, ge, sm, sx)
                : sw_chunk<TRACE, MATRIX, int, false>(S, t0, lane, m, jvalid, pc, tcur, bH, bE, bH, bE,
                                                      prof, go, ge, sm, sx);
            if (TRACE) dirs[(int64_t(s) * nchunks + c) * 64 + lane] = acc;
            tcur = tnext;
        }
        sw_fold(S, j, gH, gI, gJ);
        __syncthreads();
    }
    sw_finish(gH, gI, gJ, lane, res);
}

// Substitution-matrix pairs with several waves per pair.  The strips of one
// pair run as a pipeline over the kMwWaves waves of a workgroup: wave w takes
// strips w, w + kMwWaves, ..., strip s starting kMwLag 8-row chunks after
// strip s - 1 (lane 63 of strip s - 1 hands row r over at step r + 63, lane 0
// of strip s needs it at step r) and no earlier than its wave is free; one
// barrier per chunk keeps the waves in step.  Strip s reads the boundary
// column buffer (s + 1) & 1 and writes buffer s & 1: strip s + 2 overwrites a
// row of that buffer only after strip s + 1 has read it.  Each strip computes
// exactly what the one-wave kernel computes for it, so results, direction
// words and end cells are the same; the per-lane first-maximum records of the
// waves are merged by the same order (value, then row, then column).
constexpr int kMwWaves = 4;
constexpr int kMwLag = 9;           // > 63 / 8 + 1 chunks
constexpr size_t kSwMwMaxLds = 96 * 1024;

// P8: the profiles hold the scores as int8 (a table whose entries all fit),
// a quarter of the LDS, so more workgroups fit a CU (BLOSUM62 with 2,000-row
// texts: 43 KB instead of 67 KB, three workgroups instead of two).
__host__ __device__ inline size_t sw_mw_lds_bytes(int max_m, bool p8)
{
    return size_t(kMwWaves) * kSwProfileInts * (p8 ? 1 : 4) + size_t(4) * size_t(max_m) * 4 +
           size_t(kMwWaves) * 64 * 12;
}

template <bool TRACE, bool P8>
__global__ __launch_bounds__(64 * kMwWaves) void sw_matrix_mw_kernel(SwLaunch L)
{
    using PT = typename std::conditional<P8, int8_t, int>::type;
    extern __shared__ int lds[];
    const int slot = blockIdx.x;
    const int pid = L.first_pair + slot;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int m = L.tlen[pid];
    const int n = L.plen[pid];
    int32_t* res = L.res + int64_t(pid) * 8;
    if (m <= 0 || n <= 0) {                         // the whole workgroup leaves
        if (tid == 0) {
            res[0] = 0;
            for (int q = 1; q < 7; ++q) res[q] = -1;
            res[7] = 0;
        }
        return;
    }
    PT* prof = reinterpret_cast<PT*>(lds) + w * kSwProfileInts;
    int* bnd = reinterpret_cast<int*>(reinterpret_cast<PT*>(lds) + kMwWaves * kSwProfileInts);   // [buffer 0, 1][H, E][max_m]
    int* rec = bnd + 4 * L.max_m;                   // [wave][lane][H, I, J]
    const int mm = L.max_m;
    for (int q = tid; q < m; q += 64 * kMwWaves) { bnd[2 * mm + q] = 0; bnd[3 * mm + q] = 0; }
    __syncthreads();

    const int go = L.go, ge = L.ge;
    const int nstrips = (n + 63) >> 6;
    const int nchunks = (m + 63 + 7) >> 3;
    const int period = nchunks > kMwWaves * kMwLag ? nchunks : kMwWaves * kMwLag;
    const int last = nstrips - 1;
    const int total = (last / kMwWaves) * period + (last % kMwWaves) * kMwLag + nchunks;
    uint32_t* dirs = nullptr;
    if (TRACE) dirs = L.dirs + dir_base(L, pid, slot);
    const uint8_t* TS = L.tcode + L.toff[pid];
    int gH = 0, gI = -1, gJ = -1;
    SwLane S{0, 0, 0, 0, 0, -1};
    int j = 0;
    bool jvalid = false;
    uint64_t tcur = 0;
    for (int G = 0; G < total; ++G) {
        // this wave's strip and chunk at pipeline step G (wave-uniform)
        int s = -1, c = 0;
        if (G >= w * kMwLag) {
            const int r = (G - w * kMwLag) / period;
            const int cc = G - w * kMwLag - r * period;
            if (cc < nchunks && r * kMwWaves + w < nstrips) { s = r * kMwWaves + w; c = cc; }
        }
        if (s >= 0) {
            if (c == 0) {
                j = (s << 6) + lane;
                jvalid = j < n;
                const uint8_t pc = jvalid ? L.pcode[L.poff[pid] + j] : uint8_t(0);
                for (int k = 0; k < 32; ++k) prof[k * 64 + lane] = PT(L.sub[k * 32 + pc]);
                S = SwLane{0, 0, 0, 0, 0, -1};
                tcur = load8(TS - lane);
            }
            const int t0 = c << 3;
            const uint64_t tnext = load8(TS + t0 + 8 - lane);
            const int* bHr = bnd + ((s + 1) & 1) * 2 * mm;
            int* bHw = bnd + (s & 1) * 2 * mm;
            const uint32_t acc = (t0 >= 63 && t0 + 8 <= m && (s << 6) + 64 <= n)
                ? sw_chunk<TRACE, true, PT, true>(S, t0, lane, m, jvalid, 0, tcur, bHr, bHr + mm, bHw,
                                                  bHw + mm, prof, go, ge, 0, 0)
                : sw_chunk<TRACE, true, PT, false>(S, t0, lane, m, jvalid, 0, tcur, bHr, bHr + mm, bHw,
                                                   bHw + mm, prof, go, ge, 0, 0);
            if (TRACE) dirs[(int64_t(s) * nchunks + c) * 64 + lane] = acc;
            tcur = tnext;
            if (c == nchunks - 1) sw_fold(S, j, gH, gI, gJ);
        }
        __syncthreads();
    }
    // merge the waves' per-lane records (each is the first maximum over its
    // own strips; the order value, row, column makes the merge exact)
    rec[(w * 64 + lane) * 3 + 0] = gH;
    rec[(w * 64 + lane) * 3 + 1] = gI;
    rec[(w * 64 + lane) * 3 + 2] = gJ;
    __syncthreads();
    if (w != 0) return;
    for (int v = 1; v < kMwWaves; ++v) {
        const int h = rec[(v * 64 + lane) * 3 + 0];
        const int i = rec[(v * 64 + lane) * 3 + 1];
        const int jj = rec[(v * 64 + lane) * 3 + 2];
        if (h > gH || (h == gH && h > 0 && (i < gI || (i == gI && jj < gJ)))) { gH = h; gI = i; gJ = jj; }
    }
    sw_finish(gH, gI, gJ, lane, res);
}

// One thread per pair: the reference's walk over the nibble table, writing the
// CIGAR right-aligned into the pair's slot of text_len + pattern_len bytes.
__global__ __launch_bounds__(64) void sw_walk_kernel(SwLaunch L)
{
    const int q = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (q >= L.n) return;
    int pid = L.first_pair + q;
    if (L.rerun) {
        pid = slot_pair(L, q);
        if (pid < 0 || !L.rerun[pid]) return;
    }
    int32_t* res = L.res + int64_t(pid) * 8;
    const int score = res[0];
    if (score <= 0 || (L.walk && !L.walk[pid])) return;
    const int m = L.tlen[pid], n = L.plen[pid];
    const uint8_t* T = L.txt + L.toff[pid];
    const uint8_t* P = L.pat + L.poff[pid];
    const uint32_t* dirs = L.dirs + dir_base(L, pid, q);
    const int nchunks = (m + 63 + 7) >> 3;
    uint8_t* cg = L.cigar + L.cig_off[pid];
    int pos = m + n;
    int i = res[1] + 1, j = res[2] + 1, state = 0;
    while (i > 0 && j > 0 && pos > 0) {
        const int j0 = j - 1, sidx = j0 >> 6, l = j0 & 63, t = (i - 1) + l;
        const uint32_t w = dirs[(int64_t(sidx) * nchunks + (t >> 3)) * 64 + l];
        const uint32_t b = (w >> (4 * (t & 7))) & 15u;
        if (state == 0) {
            const uint32_t hsrc = b & 3u;
            if (hsrc == 0u) break;
            if (hsrc == 1u) {
                cg[--pos] = (T[i - 1] == P[j - 1]) ? 'M' : 'X';
                --i; --j;
            } else if (hsrc == 2u) {
                cg[--pos] = 'D';
                state = (b & 4u) ? 1 : 0;
                --j;
            } else {
                cg[--pos] = 'I';
                state = (b & 8u) ? 2 : 0;
                --i;
            }
        } else if (state == 1) {
            cg[--pos] = 'D';
            --j;
            if (!(b & 4u)) state = 0;
        } else {
            cg[--pos] = 'I';
            --i;
            if (!(b & 8u)) state = 0;
        }
    }
    res[5] = i;
    res[6] = j;
    res[7] = m + n - pos;
}

__global__ __launch_bounds__(256) void sw_gather_kernel(const int32_t* res8, int64_t n, int trace,
                                                       int32_t* out5, int32_t* cig_len)
{
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const int32_t* q = res8 + i * 8;
    int32_t* o = out5 + i * 5;
    o[0] = q[0];
    if (trace) { o[1] = q[5]; o[2] = q[6]; o[3] = q[1]; o[4] = q[2]; }
    else       { o[1] = -1;   o[2] = -1;   o[3] = q[3]; o[4] = q[4]; }
    if (cig_len) cig_len[i] = trace ? q[7] : 0;
}

// One wave per pair, 64 CIGAR ops per step.  Lane l takes op c0 + l; its
// cell (i, j) is the carried position plus the text and pattern advances of
// the ops before it in the step (ballot + mbcnt), so every op's range and
// byte tests run at once.  The re-score counts the step's M, X and gap ops
// and its gap-run starts (a gap op whose predecessor is another op): a run
// of L gap ops costs go + (L-1) ge = L ge + (go - ge).  The host validator
// keeps a run open across a foreign byte ("IZI" is one run there, two here);
// such a CIGAR is not well formed either way, so both flags agree.
__global__ __launch_bounds__(256) void sw_validate_kernel(SwValidate V)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int64_t nwaves = int64_t(gridDim.x) * 4;
    for (int64_t k = wave; k < V.n; k += nwaves) {
        const int64_t q = k * V.stride;
        const int score = V.score[q], len = V.len[k * V.len_stride];
        if (score < (V.edit ? 0 : 1) || len <= 0 || (V.walk && !V.walk[k])) {
            if (lane == 0) V.flags[k] = 0;
            continue;
        }
        const int m = V.tlen[k], n = V.plen[k];
        if (len > m + n) {                                // no such slot: not well formed
            if (lane == 0) V.flags[k] = 0;
            continue;
        }
        const int end_i = V.edit ? m - 1 : V.ei[q], end_j = V.edit ? n - 1 : V.ej[q];
        const uint8_t* cg = V.cigar + V.cig_off[k] + (m + n - len);
        const uint8_t* T = V.txt + V.toff[k];
        const uint8_t* P = V.pat + V.poff[k];
        const uint8_t* TC = V.sub ? V.tcode + V.toff[k] : nullptr;
        const uint8_t* PC = V.sub ? V.pcode + V.poff[k] : nullptr;
        const int lim_i = end_i + 1 < m ? end_i + 1 : m, lim_j = end_j + 1 < n ? end_j + 1 : n;
        int i0 = V.edit ? 0 : V.si[q], j0 = V.edit ? 0 : V.sj[q];
        int s = 0;
        bool bad = false;
        uint8_t prev = 0;                                 // op before the step
        for (int c0 = 0; c0 < len; c0 += 64) {
            const bool live = c0 + lane < len;
            const uint8_t op = live ? cg[c0 + lane] : uint8_t(0);
            const bool isM = op == 'M', isX = op == 'X', isI = op == 'I', isD = op == 'D';
            const uint64_t bi = __ballot(isM || isX || isI), bj = __ballot(isM || isX || isD);
            const int di = int(__builtin_amdgcn_mbcnt_hi(uint32_t(bi >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bi), 0u)));
            const int dj = int(__builtin_amdgcn_mbcnt_hi(uint32_t(bj >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bj), 0u)));
            const int i = i0 + di, j = j0 + dj;
            bool b = live && !(isM || isX || isI || isD);
            int sub_s = 0;                                // (matrix) this op's table score
            if (isM || isX) {
                if (i >= 0 && j >= 0 && i < lim_i && j < lim_j) {
                    b |= isM ? T[i] != P[j] : T[i] == P[j];
                    if (V.sub) sub_s = V.sub[int(TC[i]) * 32 + int(PC[j])];
                } else {
                    b = true;                             // (a negative start is no SW result)
                }
            }
            if (isI && !V.edit) b |= i >= lim_i;          // (edit: only the ends are checked)
            if (isD && !V.edit) b |= j >= lim_j;
            bad |= __any(b) != 0;
            // predecessor op: the lane below, or the last op of the step before
            const uint8_t before = uint8_t(__shfl_up(int(op), 1));
            const uint8_t pr = lane == 0 ? prev : before;
            const bool start = (isI || isD) && pr != op;
            if (V.edit) {
                s += int(__popcll(__ballot(isX || isI || isD)));
            } else if (V.sub) {
                for (int off = 32; off > 0; off >>= 1) sub_s += __shfl_xor(sub_s, off);
                s += sub_s;
            } else {
                s += int(__popcll(__ballot(isM))) * V.match + int(__popcll(__ballot(isX))) * V.mismatch;
            }
            if (!V.edit)
                s -= int(__popcll(__ballot(isI || isD))) * V.ge + int(__popcll(__ballot(start))) * (V.go - V.ge);
            i0 += int(__popcll(bi));
            j0 += int(__popcll(bj));
            prev = uint8_t(__shfl(int(op), 63));
        }
        const bool wf = !bad && i0 == end_i + 1 && j0 == end_j + 1;
        const bool rs = s == score && (wf || V.edit);
        if (lane == 0) V.flags[k] = uint8_t((rs ? 1 : 0) | (wf ? 2 : 0));
    }
}

// One wave per pair, one byte per lane per step.
__global__ __launch_bounds__(256) void sw_cigar_compact_kernel(const uint8_t* cigar, const int64_t* slot_off,
                                                               const int32_t* tlen, const int32_t* plen,
                                                               const int64_t* coff, int64_t n, uint8_t* out)
{
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6;
    const int64_t nwaves = int64_t(gridDim.x) * 4;
    for (int64_t k = wave; k < n; k += nwaves) {
        const int64_t len = coff[k + 1] - coff[k];
        if (len <= 0) continue;
        const uint8_t* src = cigar + slot_off[k] + tlen[k] + plen[k] - len;
        uint8_t* dst = out + coff[k];
        for (int64_t q = lane; q < len; q += 64) dst[q] = src[q];
    }
}

}  // namespace

hipError_t sw_cigar_compact_launch(const uint8_t* cigar, const int64_t* slot_off, const int32_t* tlen,
                                   const int32_t* plen, const int64_t* coff, int64_t n, uint8_t* out,
                                   hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 3) / 4, 8192);
    hipLaunchKernelGGL(sw_cigar_compact_kernel, dim3(unsigned(blocks)), dim3(256), 0, st, cigar, slot_off, tlen,
                       plen, coff, n, out);
    return hipGetLastError();
}

hipError_t sw_validate_launch(const SwValidate& V, hipStream_t st)
{
    if (V.n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((V.n + 3) / 4, 8192);
    hipLaunchKernelGGL(sw_validate_kernel, dim3(unsigned(blocks)), dim3(256), 0, st, V);
    return hipGetLastError();
}

hipError_t sw_gather(const int32_t* res8, int64_t n, bool trace, int32_t* out5, int32_t* cig_len,
                     hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sw_gather_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, res8, n,
                       int(trace), out5, cig_len);
    return hipGetLastError();
}

namespace {
template <bool TRACE, bool MATRIX>
void launch_forward(const SwLaunch& L, bool in_lds, size_t lds, hipStream_t st)
{
    const dim3 grid(unsigned(L.n)), block(64);
    if (in_lds) hipLaunchKernelGGL((sw_kernel<TRACE, true, MATRIX>), grid, block, lds, st, L);
    else        hipLaunchKernelGGL((sw_kernel<TRACE, false, MATRIX>), grid, block, lds, st, L);
}
}  // namespace

hipError_t sw_launch(const SwLaunch& L, bool trace, hipStream_t st, LaunchHooks* hooks)
{
    if (L.n <= 0) return hipSuccess;
    const bool matrix = L.sub != nullptr;
    const size_t bnd_bytes = size_t(2) * size_t(L.max_m) * sizeof(int);
    const bool in_lds = L.bnd == nullptr;
    const size_t lds = (in_lds ? bnd_bytes : 0) + (matrix ? size_t(kSwProfileInts) * sizeof(int) : 0);
    // substitution-matrix batches (no re-run slots) take the multi-wave
    // pipeline while both boundary buffers fit the LDS next to the profiles
    const size_t mw_lds = sw_mw_lds_bytes(L.max_m, L.sub8);
    const bool mw = matrix && !L.rerun && mw_lds <= kSwMwMaxLds;
    if (hooks) hooks->before(st);
    if (mw) {
        const dim3 grid(unsigned(L.n)), block(64 * kMwWaves);
        if (L.sub8) {
            if (trace) hipLaunchKernelGGL((sw_matrix_mw_kernel<true, true>), grid, block, mw_lds, st, L);
            else       hipLaunchKernelGGL((sw_matrix_mw_kernel<false, true>), grid, block, mw_lds, st, L);
        } else {
            if (trace) hipLaunchKernelGGL((sw_matrix_mw_kernel<true, false>), grid, block, mw_lds, st, L);
            else       hipLaunchKernelGGL((sw_matrix_mw_kernel<false, false>), grid, block, mw_lds, st, L);
        }
        if (hooks) hooks->after(st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || !trace) return e;
        const dim3 wgrid(unsigned((L.n + 63) / 64)), wblock(64);
        hipLaunchKernelGGL(sw_walk_kernel, wgrid, wblock, 0, st, L);
        return hipGetLastError();
    }
    if (trace) {
        if (matrix) launch_forward<true, true>(L, in_lds, lds, st);
        else        launch_forward<true, false>(L, in_lds, lds, st);
        if (hooks) hooks->after(st);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const dim3 wgrid(unsigned((L.n + 63) / 64)), block(64);
        hipLaunchKernelGGL(sw_walk_kernel, wgrid, block, 0, st, L);
    } else {
        if (matrix) launch_forward<false, true>(L, in_lds, lds, st);
        else        launch_forward<false, false>(L, in_lds, lds, st);
        if (hooks) hooks->after(st);
    }
    return hipGetLastError();
}

}  // namespace ga
