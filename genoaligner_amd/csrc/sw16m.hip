// sw16m.hip -- packed 16-bit Smith-Waterman with a substitution matrix
// (protein, BLOSUM62 and the like) for gfx950.
//
// Same recurrence, tie rules, end cells and direction semantics as sw.hip's
// int32 kernels (reference: include/genoaligner/backend/sw_kernel_impl.hip:
// 22-168, 252-336, walk :198-242), computed with the packed arithmetic of
// sw16.hip (strip_group in sw16_ops.hpp): two pairs per workgroup, pair A in
// the low and pair B in the high half of every 32-bit register, unsigned
// 16-bit values with E/F biased by gap_open and scores biased by
// bias = max(gap_open, -min score) so that every value is >= 0.  A pair whose
// maximum reaches 32767 - 255 - bias is flagged and re-run by the int32
// kernel (the biased sums could leave 16 bits).
//
// Scores.  Lane l of strip s owns pattern column j = 64 s + l of both pairs.
// Per wave, the LDS holds two byte tables, u_A[c][l] and u_B[c][l] =
// score(text residue c, the lane's pattern residue) + bias for c = 0..31,
// and a row c = 32 of zeros for rows past a pair's end (u = 0 means
// s = -bias <= -1: such cells stay below the pair's maximum, as in sw16.hip).
// A text row's selector holds 64 c_A in its low and the B table's offset
// plus 64 c_B in its high 16 bits; a lane adds its own byte address to each
// half (two adds), reads both bytes (two LDS reads) and merges them into the
// packed u with one v_perm.
//
// Parallelism.  A protein batch has few pairs of many strips, so the strips
// of a workgroup's two pairs run as a pipeline over its kWaves waves: wave w
// takes strips w, w + kWaves, ...; strip s starts kLag 8-row chunks after
// strip s - 1 (lane 63 of strip s - 1 hands row r over at step r + 63, lane 0
// of strip s reads it one chunk ahead of step r) and no earlier than its
// wave is free; one barrier per chunk keeps the waves in step.  Strip s reads
// the boundary column buffer (s + 1) & 1 and writes buffer s & 1; strip s + 2
// overwrites a row of that buffer only long after strip s + 1 has read it.
// Each strip computes exactly the cells, flags and first maxima of the
// one-wave kernel; the waves' per-lane records merge by (value, row, column),
// the order in which the one-wave kernel folds strips.
//
// Trace: the 4 flags per cell and pair of sw16.hip's non-HMOD layout, one
// word per lane and 4 rows; a walk kernel (one thread per pair) then runs the
// reference's three-state walk over them, carrying the cell value to find
// STOP, before any int32 re-run reuses the workgroup's direction region.
#include "sw16_ops.hpp"

namespace ga {
namespace {

constexpr int kWaves = 4;                 // waves per workgroup
constexpr int kLag = 10;                  // 8-row chunks between consecutive strips
constexpr int kZeroCode = 32;             // table row of u = 0
constexpr int kTabBytes = 33 * 64;        // one half's table: 33 rows x 64 lanes
constexpr int kProfBytes = kWaves * 2 * kTabBytes;
constexpr uint32_t kSelZero = uint32_t(kZeroCode * 64) | (uint32_t(kTabBytes + kZeroCode * 64) << 16);

// Boundary rows in use: the text rows rounded up to whole chunks of 8.
__host__ __device__ inline int bnd_rows(int m)
{
    const int r = (m + 7) & ~7;
    return r > 8 ? r : 8;
}

__host__ __device__ inline size_t lds_bytes(int max_m)
{
    return size_t(kProfBytes) + size_t(2) * size_t(bnd_rows(max_m)) * 8;
}

// The workgroup's two pairs and the shape they are computed over.
struct Duo {
    int ia, ib;                 // pair ids (ib -1: none)
    int mA, nA, mB, nB;         // lengths as given
    int mAp, nAp, mBp, nBp;     // 0 for a pair with an empty side
    int M, N;                   // rows and columns computed
    int nq;                     // flag words per lane column and strip
};

__device__ __forceinline__ Duo duo_of(const Sw16Launch& L, int g)
{
    Duo d;
    d.ia = L.pair_a ? L.pair_a[g] : L.first_pair + 2 * g;
    d.ib = L.pair_b ? L.pair_b[g] : L.first_pair + 2 * g + 1;
    if (d.ib >= L.n_total) d.ib = -1;
    d.mA = L.tlen[d.ia]; d.nA = L.plen[d.ia];
    d.mB = d.ib >= 0 ? L.tlen[d.ib] : 0; d.nB = d.ib >= 0 ? L.plen[d.ib] : 0;
    const bool la = d.mA > 0 && d.nA > 0, lb = d.mB > 0 && d.nB > 0;
    d.mAp = la ? d.mA : 0; d.nAp = la ? d.nA : 0;
    d.mBp = lb ? d.mB : 0; d.nBp = lb ? d.nB : 0;
    d.M = d.mAp > d.mBp ? d.mAp : d.mBp;
    d.N = d.nAp > d.nBp ? d.nAp : d.nBp;
    d.nq = ((d.M + 63 + 7) & ~7) >> 2;
    return d;
}

template <bool TRACE, bool GK>
__global__ __launch_bounds__(64 * kWaves) void sw16m_kernel(Sw16Launch L)
{
    extern __shared__ __align__(16) uint32_t lds16[];
    const int g = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const Duo D = duo_of(L, g);
    const int M = D.M, N = D.N;

    if (M == 0) {                              // both pairs have an empty side
        if (tid < 2) {
            const int pid = tid == 0 ? D.ia : D.ib;
            if (pid >= 0) {
                int32_t* res = L.res + int64_t(pid) * 8;
                res[0] = 0;
                for (int q = 1; q < 7; ++q) res[q] = -1;
                res[7] = 0;
                L.rerun[pid] = 0;
            }
        }
        return;
    }

    uint8_t* prof = reinterpret_cast<uint8_t*>(lds16) + w * 2 * kTabBytes;
    uint2* bnd = reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(lds16) + kProfBytes);
    const int Mr = bnd_rows(M);
    const int go = L.go, ge = L.ge;
    const uint32_t go2 = uint32_t(go) * 0x10001u, ge2 = uint32_t(ge) * 0x10001u;
    const int bias = L.bias;
    const uint32_t b2 = uint32_t(bias) * 0x10001u, gk2 = uint32_t(bias - go) * 0x10001u;

    // Text selectors of both pairs (guard rows around them) and both boundary
    // buffers at the zero boundary (H = 0, E = -go).
    const uint8_t* TCA = L.tcode + L.toff[D.ia];
    const uint8_t* TCB = D.ib >= 0 ? L.tcode + L.toff[D.ib] : TCA;
    uint32_t* sel = L.selbuf + int64_t(g) * sw16_sel_stride(L.max_m) + 64;
    for (int i = tid - 64; i < M + 80; i += 64 * kWaves) {
        uint32_t sv = kSelZero;
        if (i >= 0 && i < M) {
            const uint32_t ca = i < D.mAp ? uint32_t(TCA[i]) : uint32_t(kZeroCode);
            const uint32_t cb = i < D.mBp ? uint32_t(TCB[i]) : uint32_t(kZeroCode);
            sv = (ca * 64u) | ((uint32_t(kTabBytes) + cb * 64u) << 16);
        }
        sel[i] = sv;
    }
    for (int i = tid; i < 2 * Mr; i += 64 * kWaves) bnd[i] = make_uint2(0u, go2);
    __syncthreads();

    const int nstrips = (N + 63) >> 6;
    const int nchunks = (M + 63 + 7) >> 3;
    const int nq = nchunks * 2;
    const int period = nchunks > kWaves * kLag ? nchunks : kWaves * kLag;
    const int last = nstrips - 1;
    const int total = (last / kWaves) * period + (last % kWaves) * kLag + nchunks;
    const int nmin = D.nAp < D.nBp ? D.nAp : D.nBp;
    uint32_t* dirs = TRACE ? L.dirs + int64_t(g) * L.dir_stride : nullptr;
    const uint8_t* PCA = L.pcode + L.poff[D.ia];
    const uint8_t* PCB = D.ib >= 0 ? L.pcode + L.poff[D.ib] : PCA;
    const uint32_t* sp = sel - lane;           // lane l reads row t - l
    const uint8_t* pb = prof + lane;           // the lane's byte of every table row

    int gH[2] = {0, 0}, gI[2] = {-1, -1}, gJ[2] = {-1, -1};
    uint32_t thrm1 = 0;                        // per half: this wave's best so far - 1
    Strip S{};
    int j = 0;
    bool fast = false;
    // Prefetch sets, indexed by the parity of the group (4 rows): the group
    // about to run has its u and boundary rows; the selectors run two groups
    // ahead.  The sets rotate with the group index, so no register copy waits
    // for a load in flight.
    uint32_t sv[2][4], uu[2][4];
    uint2 bv[2][4];
    const int last_t = nchunks * 8 - 4;        // first row of the strip's last group
    auto ld_sel = [&](int t, uint32_t (&s4)[4]) {
        const int tt = t < last_t ? t : last_t;
#pragma unroll
        for (int u = 0; u < 4; ++u) s4[u] = sp[tt + u];
    };
    auto ld_u = [&](const uint32_t (&s4)[4], uint32_t (&u4)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
            u4[u] = __builtin_amdgcn_perm(uint32_t(pb[s4[u] >> 16]), uint32_t(pb[s4[u] & 0xffffu]), 0x0c040c00u);
    };
    auto ld_bnd = [&](const uint2* br, int t, uint2 (&b4)[4]) {
        const uint2* b = br + (t < Mr - 4 ? t : Mr - 4);
#pragma unroll
        for (int u = 0; u < 4; ++u) b4[u] = b[u];
    };

    for (int G = 0; G < total; ++G) {
        // this wave's strip and chunk at pipeline step G (wave-uniform)
        int s = -1, c = 0;
        if (G >= w * kLag) {
            const int r = (G - w * kLag) / period;
            const int cc = G - w * kLag - r * period;
            if (cc < nchunks && r * kWaves + w < nstrips) { s = r * kWaves + w; c = cc; }
        }
        if (s >= 0) {
            const uint2* br = bnd + ((s + 1) & 1) * Mr;
            uint2* bw = bnd + (s & 1) * Mr;
            if (c == 0) {
                j = (s << 6) + lane;
                const uint32_t pa = j < D.nAp ? PCA[j] : 0u, pbc = j < D.nBp ? PCB[j] : 0u;
                for (int code = 0; code < 32; ++code) {
                    prof[code * 64 + lane] = uint8_t(j < D.nAp ? L.sub[code * 32 + pa] + bias : 0);
                    prof[kTabBytes + code * 64 + lane] = uint8_t(j < D.nBp ? L.sub[code * 32 + pbc] + bias : 0);
                }
                prof[kZeroCode * 64 + lane] = 0;
                prof[kTabBytes + kZeroCode * 64 + lane] = 0;
                S.H = 0; S.Eh = go2; S.Fh = go2; S.diag = 0;
                S.bestH = thrm1; S.bestT = 0xffffffffu;   // record at the bound (see strip_group)
                S.colmask = (j < D.nAp ? 0xffffu : 0u) | (j < D.nBp ? 0xffff0000u : 0u);
                fast = (s << 6) + 64 <= nmin;
                ld_sel(0, sv[0]);
                ld_sel(4, sv[1]);
                ld_u(sv[0], uu[0]);
                ld_bnd(br, 0, bv[0]);
            }
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int t0 = c * 8 + 4 * p;
                // prefetch: selectors of group +2, u and boundary rows of group +1
                ld_sel(t0 + 8, sv[p]);
                ld_u(sv[p ^ 1], uu[p ^ 1]);
                ld_bnd(br, t0 + 4, bv[p ^ 1]);
                uint32_t acc = 0, hout[4], eout[4];
                const uint32_t (&u4)[4] = uu[p];
                const auto mat_u = [&](int u) { return u4[u]; };
                if (fast)
                    strip_group<TRACE, false, false, GK>(S, t0, lane, mat_u, bv[p], b2, gk2, ge2, acc, hout, eout, thrm1);
                else
                    strip_group<TRACE, false, true, true>(S, t0, lane, mat_u, bv[p], b2, gk2, ge2, acc, hout, eout, thrm1);
                if (TRACE) dirs[(int64_t(s) * nq + (t0 >> 2)) * 64 + lane] = acc;
                // lane 63 hands rows t0-63 .. t0-60 to strip s + 1
                if (lane == 63) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = t0 + u - 63;
                        if (r >= 0 && r < Mr) bw[r] = make_uint2(hout[u], eout[u]);
                    }
                }
            }
            if (c == nchunks - 1) {
                // fold this strip's first maxima (per half) into the running records
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int bh = int((S.bestH >> (16 * h)) & 0xffffu);
                    const int bt = int((S.bestT >> (16 * h)) & 0xffffu);
                    const int bi = bh > 0 ? bt - lane : -1;
                    if (bh > gH[h] || (bh == gH[h] && bh > 0 && bi < gI[h])) { gH[h] = bh; gI[h] = bi; gJ[h] = j; }
                }
            }
        }
        __syncthreads();
    }

    // merge the waves' per-lane records (LDS of the tables, free now)
    int* rec = reinterpret_cast<int*>(lds16);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int* q = rec + ((h * kWaves + w) * 64 + lane) * 3;
        q[0] = gH[h]; q[1] = gI[h]; q[2] = gJ[h];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        for (int v = 1; v < kWaves; ++v) {
            const int* q = rec + ((h * kWaves + v) * 64 + lane) * 3;
            const int hh = q[0], ii = q[1], jj = q[2];
            if (hh > gH[h] || (hh == gH[h] && hh > 0 && (ii < gI[h] || (ii == gI[h] && jj < gJ[h])))) {
                gH[h] = hh; gI[h] = ii; gJ[h] = jj;
            }
        }
    }
    const int limit = 32767 - 255 - bias;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int pid = h == 0 ? D.ia : D.ib;
        if (pid < 0) continue;
        // score-only end: the reference's xor-butterfly argmax at width 64;
        // trace end: the row-major first cell among lanes holding the maximum
        int lH = gH[h], lI = gI[h], lJ = gJ[h];
        int rH = gH[h], rI = gI[h], rJ = gJ[h];
        for (int off = 32; off > 0; off >>= 1) {
            const int ob = __shfl_xor(lH, off), oi = __shfl_xor(lI, off), oj = __shfl_xor(lJ, off);
            if (ob > lH) { lH = ob; lI = oi; lJ = oj; }
            const int rb = __shfl_xor(rH, off), ri = __shfl_xor(rI, off), rj = __shfl_xor(rJ, off);
            if (rb > rH || (rb == rH && rb > 0 && (ri < rI || (ri == rI && rj < rJ)))) {
                rH = rb; rI = ri; rJ = rj;
            }
        }
        if (lane == 0) {
            int32_t* res = L.res + int64_t(pid) * 8;
            const int m = h == 0 ? D.mA : D.mB, n = h == 0 ? D.nA : D.nB;
            if (m <= 0 || n <= 0) {
                res[0] = 0;
                for (int q = 1; q < 7; ++q) res[q] = -1;
                res[7] = 0;
                L.rerun[pid] = 0;
            } else {
                const bool rerun = lH >= limit;
                res[0] = rerun ? kSwRerun : lH;
                res[1] = rI; res[2] = rJ; res[3] = lI; res[4] = lJ;
                res[5] = -1; res[6] = -1; res[7] = 0;
                L.rerun[pid] = rerun ? 1 : 0;
            }
        }
    }
}

// One thread per pair: the reference's three-state walk (sw_kernel_impl.hip:
// 198-242) over the packed flags, from the trace end cell with the cell's
// value carried (score; minus s on a diagonal step, plus ge on an extension,
// plus go on an opening), stopping where that value is 0 in the H state.
// Writes the CIGAR right-aligned into the pair's slot of m + n bytes.
__global__ __launch_bounds__(64) void sw16m_walk_kernel(Sw16Launch L)
{
    const int q = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (q >= 2 * L.n_waves) return;
    const int g = q >> 1, h = q & 1;
    const Duo D = duo_of(L, g);
    const int pid = h == 0 ? D.ia : D.ib;
    if (pid < 0 || (L.walk && !L.walk[pid])) return;
    int32_t* res = L.res + int64_t(pid) * 8;
    const int score = res[0];
    if (score <= 0) return;                    // nothing to walk, or re-run (kSwRerun)
    const int m = h == 0 ? D.mA : D.mB, n = h == 0 ? D.nA : D.nB;
    const uint8_t* T = L.txt + L.toff[pid];
    const uint8_t* P = L.pat + L.poff[pid];
    const uint8_t* TC = L.tcode + L.toff[pid];
    const uint8_t* PC = L.pcode + L.poff[pid];
    const uint32_t* dirs = L.dirs + int64_t(g) * L.dir_stride;
    uint8_t* cg = L.cigar + L.cig_off[pid];
    const int nq = D.nq;
    const int go = L.go, ge = L.ge;
    int pos = m + n, state = 0, v = score;
    int i = res[1] + 1, j = res[2] + 1;
    while (i > 0 && j > 0 && pos > 0 && !(state == 0 && v == 0)) {
        const int l = (j - 1) & 63, s = (j - 1) >> 6, t = i - 1 + l;
        const uint32_t word = dirs[(int64_t(s) * nq + (t >> 2)) * 64 + l];
        const int b = 8 * h + (t & 3);
        if (state == 0 && !((word >> (b + 4)) & 1u)) {
            cg[--pos] = T[i - 1] == P[j - 1] ? 'M' : 'X';
            v -= L.sub[int(TC[i - 1]) * 32 + int(PC[j - 1])];
            --i; --j;
        } else if (state == 1 || (state == 0 && !((word >> (b + 20)) & 1u))) {
            cg[--pos] = 'D';
            const bool ext = ((word >> b) & 1u) != 0;
            state = ext ? 1 : 0;
            v += ext ? ge : go;
            --j;
        } else {
            cg[--pos] = 'I';
            const bool ext = ((word >> (b + 16)) & 1u) != 0;
            state = ext ? 2 : 0;
            v += ext ? ge : go;
            --i;
        }
    }
    res[5] = i;
    res[6] = j;
    res[7] = m + n - pos;
}

}  // namespace

int sw16m_bias(const int32_t* sub, int gap_open)
{
    int lo = 0, hi = 0;
    for (int x = 0; x < 32 * 32; ++x) {
        lo = sub[x] < lo ? sub[x] : lo;
        hi = sub[x] > hi ? sub[x] : hi;
    }
    const int bias = gap_open > -lo ? gap_open : -lo;
    if (gap_open < 0 || bias + hi > 255 || bias > 255) return -1;
    return bias;
}

hipError_t sw16m_launch(const Sw16Launch& L, bool trace, hipStream_t st, LaunchHooks* hooks)
{
    if (L.n_waves <= 0) return hipSuccess;
    const size_t lds = lds_bytes(L.max_m);
    const dim3 grid(unsigned(L.n_waves)), block(64 * kWaves);
    const bool gk = L.bias > L.go;                 // -min score > gap_open: H needs bias - go
    if (hooks) hooks->before(st);
    if (gk) {
        if (trace) hipLaunchKernelGGL((sw16m_kernel<true, true>), grid, block, lds, st, L);
        else       hipLaunchKernelGGL((sw16m_kernel<false, true>), grid, block, lds, st, L);
    } else {
        if (trace) hipLaunchKernelGGL((sw16m_kernel<true, false>), grid, block, lds, st, L);
        else       hipLaunchKernelGGL((sw16m_kernel<false, false>), grid, block, lds, st, L);
    }
    if (hooks) hooks->after(st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !trace) return e;
    if (hooks) hooks->before(st);
    hipLaunchKernelGGL(sw16m_walk_kernel, dim3(unsigned((2 * L.n_waves + 63) / 64)), dim3(64), 0, st, L);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

}  // namespace ga
