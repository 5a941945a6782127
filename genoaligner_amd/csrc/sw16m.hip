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
// Per wave, the LDS holds two byte tables, u_A and u_B = score(text residue
// c, the lane's pattern residue) + bias for c = 0..31, and codes 32..35 of
// zeros for rows past a pair's end (u = 0 means s = -bias <= -1: such cells
// stay below the pair's maximum, as in sw16.hip).  Code c of lane l is byte
// c % 4 of word (c / 4) * 64 + l, so the 64 lanes of a lookup read 64 words
// in 64 banks.  A text row's selector holds tab_off(c_A) in its low and the
// B table's offset plus tab_off(c_B) in its high 16 bits; a lane adds its own
// byte address to each half (two adds), reads both bytes (two LDS reads) and
// merges them into the packed u with one v_perm.
//
// Parallelism.  A protein batch has few pairs of many strips, so the strips
// of a workgroup's two pairs run as a pipeline over its WAVES waves (4, or 8
// when the launch's pairs have 8 strips or more, which halves the longest
// workgroup's run): wave w takes strips w, w + WAVES, ...; strip s starts kLag
// chunks of kChunk rows after
// strip s - 1 (lane 63 of strip s - 1 hands row r over at step r + 63, lane 0
// of strip s reads it one chunk ahead of step r) and no earlier than its
// wave is free; one barrier per chunk keeps the waves in step.  One boundary
// column buffer serves every strip: row r is written by strip s at its step
// r + 63, read by strip s + 1 one group ahead of its step r (kLag chunks
// after strip s started, so at least two barriers after that write) and
// overwritten by strip s + 1 only at its step r + 63, after that read.
// (Reads clamped to the last rows, in the drain past row M, may see the
// strip's own rows; they feed only drained cells.)
// Each strip computes exactly the cells, flags and first maxima of the
// one-wave kernel; the waves' per-lane records merge by (value, row, column),
// the order in which the one-wave kernel folds strips.
//
// Trace: the 4 flags per cell and pair of sw16.hip's non-HMOD layout, one
// word per lane and 4 rows; after the strips, wave 0 runs the reference's
// three-state walk of both pairs over them (walk_duo), carrying the cell
// value to find STOP, before any int32 re-run reuses the workgroup's
// direction region.
#include "sw16_ops.hpp"

#include <type_traits>

namespace ga {
namespace {

constexpr int kChunk = 16;               // text rows per pipeline step (one barrier each)
// Pipeline steps between consecutive strips: the rows a strip reads in a step
// (and the group it prefetches) were handed over in an earlier step.
constexpr int kLag = 2 + 66 / kChunk;
constexpr int kZeroCode = 32;             // table row of u = 0
// One half's table: codes 0..35 in 9 rows of 64 lane words, 4 codes per word,
// so the lanes of a lookup read 64 different words in 64 different banks
// whatever codes they hold.
constexpr int kTabBytes = 9 * 64 * 4;
__host__ __device__ constexpr int prof_bytes(int waves) { return waves * 2 * kTabBytes; }
__host__ __device__ constexpr uint32_t tab_off(uint32_t c) { return (c >> 2) * 256u + (c & 3u); }
// merge records (in the tables' LDS)
__host__ __device__ constexpr int rec_bytes(int waves) { return 2 * waves * 64 * 3 * 4; }
constexpr uint32_t kSelZero = tab_off(kZeroCode) | ((uint32_t(kTabBytes) + tab_off(kZeroCode)) << 16);

// Boundary rows in use: the text rows rounded up to whole chunks of 8.
__host__ __device__ inline int bnd_rows(int m)
{
    const int r = (m + 7) & ~7;
    return r > 8 ? r : 8;
}

__host__ __device__ inline size_t lds_bytes(int max_m, int waves)
{
    return size_t(prof_bytes(waves)) + size_t(bnd_rows(max_m)) * 8;
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
    d.nq = ((d.M + 63 + kChunk - 1) / kChunk) * (kChunk / 4);
    return d;
}

// The reference's three-state walk (sw_kernel_impl.hip:198-242) over the
// packed flags, run by wave 0 of each workgroup of the forward kernel once its
// strips are done (other workgroups' strips hide its latency): lane 0 walks
// pair A and lane 32 pair B, from the trace end cell with the cell's value
// carried (score; minus s on a diagonal step, plus ge on an extension, plus
// go on an opening) and stopping where that value is 0 in the H state.  When
// a walk leaves its window, its half-wave loads the kWinC columns x kWinR rows
// up and left of the current cell in one coalesced round: lane c of the half
// the flag words of column j - c, the text byte and residue code of row i - c
// and the pattern byte and code of column j - c, so the steps inside the
// window read only LDS.  The CIGAR goes right-aligned into the pair's slot of
// m + n bytes.  Flag layout (non-HMOD, sw16.hip): cell (i, j) of half h sits
// in word (s * nq + (t >> 2)) * 64 + l, s and l the strip and lane of column
// j - 1 and t = i - 1 + l; with b = 8h + (t & 3), bit b is E extended, b + 4
// not-from-the-diagonal, b + 16 F extended and b + 20 F over E.
struct WalkLds {
    uint32_t win[2][kWinC * kWinQ];
    uint8_t wt[2][kWinR], wtc[2][kWinR], wp[2][kWinC], wpc[2][kWinC];
    int16_t sub_s[32 * 32];
};
static_assert(rec_bytes(4) + sizeof(WalkLds) <= size_t(prof_bytes(4)), "the walk's LDS lies inside the tables' region");

// Called by one whole wave after the forward pass of its workgroup's pairs.
// score / ei / ej: the half's score (> 0 to walk) and 0-based trace end cell.
__device__ void walk_duo(const Sw16Launch& L, const Duo& D, int g, int lane, int score, int ei, int ej,
                         bool walk_it, WalkLds& S)
{
    auto& win = S.win;
    auto& wt = S.wt;
    auto& wtc = S.wtc;
    auto& wp = S.wp;
    auto& wpc = S.wpc;
    auto& sub_s = S.sub_s;
    const int hw = lane >> 5, c = lane & 31;
    for (int q = lane; q < 32 * 32; q += 64) sub_s[q] = int16_t(L.sub[q]);
    __builtin_amdgcn_wave_barrier();
    const int pid = hw == 0 ? D.ia : D.ib;
    const int m = hw == 0 ? D.mA : D.mB, n = hw == 0 ? D.nA : D.nB;
    int i = ei + 1, j = ej + 1;
    const bool walker = c == 0 && walk_it;
    const int ps = pid >= 0 ? pid : D.ia;
    const uint8_t* T = L.txt + L.toff[ps];
    const uint8_t* P = L.pat + L.poff[ps];
    const uint8_t* TC = L.tcode + L.toff[ps];
    const uint8_t* PC = L.pcode + L.poff[ps];
    uint8_t* cg = L.cigar + (walker ? L.cig_off[pid] : 0);
    const uint32_t* dirs = L.dirs + int64_t(g) * L.dir_stride;
    const int nq = D.nq;
    const int go = L.go, ge = L.ge;
    uint32_t* w = win[hw];
    int pos = m + n, state = 0, v = score;
    bool active = walker, have = false;
    int iw = 0, jw = 0;                           // window anchor (walker lanes)
    while (__any(active)) {
        if (active && (!(i > 0 && j > 0 && pos > 0) || (state == 0 && v == 0))) active = false;
        const bool need = active && !(have && iw - i < kWinR && jw - j < kWinC);
        const int need_h = __shfl(int(need), hw * 32);
        const int ih = __shfl(i, hw * 32), jh = __shfl(j, hw * 32);
        if (need_h) {
            const int jj = jh - c;                // 1-based column of this lane
            const int ii = ih - c;                // 1-based row of this lane
            // every load issued before any is waited for: out-of-range words
            // and bytes read index 0 and are discarded
            const int jc = jj >= 1 ? jj : 1, ic = ii >= 1 ? ii : 1;
            const int l = (jc - 1) & 63, s = (jc - 1) >> 6;
            const int q0 = (ih - kWinR + l) >> 2;       // floor
            uint32_t val[kWinQ];
#pragma unroll
            for (int k = 0; k < kWinQ; ++k) {
                const int q = q0 + k;
                val[k] = dirs[(int64_t(s) * nq + (q >= 0 && q < nq ? q : 0)) * 64 + l];
            }
            const uint8_t pv = P[jc - 1], pcv = PC[jc - 1], tv = T[ic - 1], tcv = TC[ic - 1];
            if (jj >= 1) {
#pragma unroll
                for (int k = 0; k < kWinQ; ++k) {
                    const int q = q0 + k;
                    w[c * kWinQ + k] = (q >= 0 && q < nq) ? val[k] : 0u;
                }
                wp[hw][c] = pv;
                wpc[hw][c] = pcv;
            }
            if (ii >= 1) {
                wt[hw][c] = tv;
                wtc[hw][c] = tcv;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (need) { iw = i; jw = j; have = true; }
        if (!active) continue;
        // at most kWinR + kWinC - 1 steps stay inside the window
        while (i > 0 && j > 0 && pos > 0 && !(state == 0 && v == 0) && iw - i < kWinR && jw - j < kWinC) {
            const int l = (j - 1) & 63, t = i - 1 + l;
            const uint32_t word = w[(jw - j) * kWinQ + ((t >> 2) - ((iw - kWinR + l) >> 2))];
            const int b = 8 * hw + (t & 3);
            if (state == 0 && !((word >> (b + 4)) & 1u)) {
                const int ri = iw - i, cj = jw - j;
                cg[--pos] = wt[hw][ri] == wp[hw][cj] ? 'M' : 'X';
                v -= sub_s[int(wtc[hw][ri]) * 32 + int(wpc[hw][cj])];
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
    }
    if (walker) {
        int32_t* res = L.res + int64_t(pid) * 8;
        res[5] = i;
        res[6] = j;
        res[7] = m + n - pos;
    }
}

template <bool TRACE, bool GK, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void sw16m_kernel(Sw16Launch L)
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
    uint2* bnd = reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(lds16) + prof_bytes(WAVES));
    const int Mr = bnd_rows(M);
    const int go = L.go, ge = L.ge;
    const uint32_t go2 = uint32_t(go) * 0x10001u, ge2 = uint32_t(ge) * 0x10001u;
    const int bias = L.bias;
    const uint32_t b2 = uint32_t(bias) * 0x10001u, gk2 = uint32_t(bias - go) * 0x10001u;

    // Text selectors of both pairs (guard rows around them) and the boundary
    // buffer at the zero boundary (H = 0, E = -go).
    const uint8_t* TCA = L.tcode + L.toff[D.ia];
    const uint8_t* TCB = D.ib >= 0 ? L.tcode + L.toff[D.ib] : TCA;
    uint32_t* sel = L.selbuf + int64_t(g) * sw16_sel_stride(L.max_m) + 64;
    for (int i = tid - 64; i < M + 80; i += 64 * WAVES) {
        uint32_t sv = kSelZero;
        if (i >= 0 && i < M) {
            const uint32_t ca = i < D.mAp ? uint32_t(TCA[i]) : uint32_t(kZeroCode);
            const uint32_t cb = i < D.mBp ? uint32_t(TCB[i]) : uint32_t(kZeroCode);
            sv = tab_off(ca) | ((uint32_t(kTabBytes) + tab_off(cb)) << 16);
        }
        sel[i] = sv;
    }
    for (int i = tid; i < Mr; i += 64 * WAVES) bnd[i] = make_uint2(0u, go2);
    __syncthreads();

    const int nstrips = (N + 63) >> 6;
    const int nchunks = (M + 63 + kChunk - 1) / kChunk;
    const int nq = nchunks * (kChunk / 4);
    const int period = nchunks > WAVES * kLag ? nchunks : WAVES * kLag;
    const int last = nstrips - 1;
    const int total = (last / WAVES) * period + (last % WAVES) * kLag + nchunks;
    const int nmin = D.nAp < D.nBp ? D.nAp : D.nBp;
    uint32_t* dirs = TRACE ? L.dirs + int64_t(g) * L.dir_stride : nullptr;
    const uint8_t* PCA = L.pcode + L.poff[D.ia];
    const uint8_t* PCB = D.ib >= 0 ? L.pcode + L.poff[D.ib] : PCA;
    const uint32_t* sp = sel - lane;           // lane l reads row t - l
    const uint8_t* pb = prof + lane * 4;       // the lane's word of every table row

    int gH[2] = {0, 0}, gI[2] = {-1, -1}, gJ[2] = {-1, -1};
    uint32_t thrm1 = 0;                        // per half: this wave's best so far - 1
    const int last_t = nchunks * kChunk - 4;   // first row of the strip's last group
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

    // One strip s, nchunks pipeline steps (one barrier each).  FAST: every
    // column of the strip exists in both pairs (no column masks).
    auto run_strip = [&](auto fast_tag, int s) {
        constexpr bool FAST = decltype(fast_tag)::value;
        const uint2* br = bnd;
        uint2* bw = bnd;
        const int j = (s << 6) + lane;
        const uint32_t pa = j < D.nAp ? PCA[j] : 0u, pbc = j < D.nBp ? PCB[j] : 0u;
        // the lane's two score columns: 64 table reads, 16 in flight at a
        // time, written as words of 4 codes; codes 32..35 (the zero row) are 0
        uint32_t* pw = reinterpret_cast<uint32_t*>(prof) + lane;
#pragma unroll
        for (int c0 = 0; c0 < 32; c0 += 8) {
            int32_t sa[8], sb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                sa[u] = L.sub[(c0 + u) * 32 + pa];
                sb[u] = L.sub[(c0 + u) * 32 + pbc];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                uint32_t wa = 0, wb = 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    wa |= (uint32_t(sa[4 * q + u] + bias) & 0xffu) << (8 * u);
                    wb |= (uint32_t(sb[4 * q + u] + bias) & 0xffu) << (8 * u);
                }
                pw[(c0 / 4 + q) * 64] = j < D.nAp ? wa : 0u;
                pw[kTabBytes / 4 + (c0 / 4 + q) * 64] = j < D.nBp ? wb : 0u;
            }
        }
        pw[8 * 64] = 0u;
        pw[kTabBytes / 4 + 8 * 64] = 0u;
        Strip S{};
        S.H = 0; S.Eh = go2; S.Fh = go2; S.diag = 0;
        S.bestH = thrm1; S.bestT = 0xffffffffu;   // record at the bound (see strip_group)
        S.colmask = (j < D.nAp ? 0xffffu : 0u) | (j < D.nBp ? 0xffff0000u : 0u);
        // Prefetch sets, indexed by the parity of the group (4 rows): the
        // group about to run has its u and boundary rows; the selectors run
        // two groups ahead.  The sets rotate with the group index, so no
        // register copy waits for a load in flight.
        uint32_t sv[2][4], uu[2][4];
        uint2 bv[2][4];
        ld_sel(0, sv[0]);
        ld_sel(4, sv[1]);
        ld_u(sv[0], uu[0]);
        ld_bnd(br, 0, bv[0]);
        for (int c = 0; c < nchunks; ++c) {
#pragma unroll
            for (int k = 0; k < kChunk / 4; ++k) {
                const int p = k & 1;
                const int t0 = c * kChunk + 4 * k;
                // prefetch: selectors of group +2, u and boundary rows of group +1
                ld_sel(t0 + 8, sv[p]);
                ld_u(sv[p ^ 1], uu[p ^ 1]);
                ld_bnd(br, t0 + 4, bv[p ^ 1]);
                uint32_t acc = 0, hout[4], eout[4];
                const uint32_t (&u4)[4] = uu[p];
                const auto mat_u = [&](int u) { return u4[u]; };
                strip_group<TRACE, kLayFlags, !FAST, FAST ? GK : true>(S, t0, lane, mat_u, bv[p], b2, gk2, ge2, acc,
                                                                  hout, eout, thrm1);
                if (TRACE) dirs[(int64_t(s) * nq + (t0 >> 2)) * 64 + lane] = acc;
                // lane 63 hands rows t0-63 .. t0-60 to strip s + 1; rows
                // outside 0 .. Mr-1 are dropped (rows M .. Mr-1 only reach
                // drained cells)
                if (lane == 63) {
                    if (t0 >= 64 && t0 - 60 < Mr) {
                        uint2* b = bw + (t0 - 63);
#pragma unroll
                        for (int u = 0; u < 4; ++u) b[u] = make_uint2(hout[u], eout[u]);
                    } else {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int r = t0 + u - 63;
                            if (r >= 0 && r < Mr) bw[r] = make_uint2(hout[u], eout[u]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        // fold this strip's first maxima (per half) into the running records
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int bh = int((S.bestH >> (16 * h)) & 0xffffu);
            const int bt = int((S.bestT >> (16 * h)) & 0xffffu);
            const int bi = bh > 0 ? bt - lane : -1;
            if (bh > gH[h] || (bh == gH[h] && bh > 0 && bi < gI[h])) { gH[h] = bh; gI[h] = bi; gJ[h] = j; }
        }
    };

    // Wave w runs strips w, w + WAVES, ...; strip s occupies pipeline steps
    // (s / WAVES) * period + w * kLag onwards.  Every wave passes the same
    // number of barriers (total), idle steps included.
    int G = 0;
    for (int s = w; s < nstrips; s += WAVES) {
        const int start = (s / WAVES) * period + w * kLag;
        for (; G < start; ++G) __syncthreads();
        if ((s << 6) + 64 <= nmin) run_strip(std::true_type{}, s);
        else                       run_strip(std::false_type{}, s);
        G += nchunks;
    }
    for (; G < total; ++G) __syncthreads();

    // merge the waves' per-lane records (LDS of the tables, free now)
    int* rec = reinterpret_cast<int*>(lds16);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int* q = rec + ((h * WAVES + w) * 64 + lane) * 3;
        q[0] = gH[h]; q[1] = gI[h]; q[2] = gJ[h];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        for (int v = 1; v < WAVES; ++v) {
            const int* q = rec + ((h * WAVES + v) * 64 + lane) * 3;
            const int hh = q[0], ii = q[1], jj = q[2];
            if (hh > gH[h] || (hh == gH[h] && hh > 0 && (ii < gI[h] || (ii == gI[h] && jj < gJ[h])))) {
                gH[h] = hh; gI[h] = ii; gJ[h] = jj;
            }
        }
    }
    const int limit = 32767 - 255 - bias;
    int wS[2] = {0, 0}, wI[2] = {-1, -1}, wJ[2] = {-1, -1};
    bool wOk[2] = {false, false};
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
        const int m = h == 0 ? D.mA : D.mB, n = h == 0 ? D.nA : D.nB;
        wS[h] = lH; wI[h] = rI; wJ[h] = rJ;
        wOk[h] = m > 0 && n > 0 && lH > 0 && lH < limit && (!L.walk || L.walk[pid]);
    }
    if (!TRACE) return;
    // the walks (the re-run kernel walks the pairs it re-runs); the tables'
    // LDS behind the merge records holds the walk's windows
    WalkLds& WS = *reinterpret_cast<WalkLds*>(reinterpret_cast<uint8_t*>(lds16) + rec_bytes(WAVES));
    const int hw = lane >> 5;
    walk_duo(L, D, g, lane, hw == 0 ? wS[0] : wS[1], hw == 0 ? wI[0] : wI[1], hw == 0 ? wJ[0] : wJ[1],
             hw == 0 ? wOk[0] : wOk[1], WS);
}

}  // namespace

int64_t sw16m_dir_words_per_group(int max_m, int max_n)
{
    if (max_m <= 0 || max_n <= 0) return 0;
    const int64_t strips = (max_n + 63) / 64;
    const int64_t own = strips * ((int64_t(max_m) + 63 + kChunk - 1) / kChunk) * (kChunk / 4) * 64;
    const int64_t alt = 2 * strips * int64_t((max_m + 63 + 7) / 8) * 64;
    return own > alt ? own : alt;
}

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
    const bool gk = L.bias > L.go;                 // -min score > gap_open: H needs bias - go
    const bool wide = (L.max_n + 63) / 64 >= 8;    // 8 strips or more: 8 waves
    const int waves = wide ? 8 : 4;
    const size_t lds = lds_bytes(L.max_m, waves);
    const dim3 grid(unsigned(L.n_waves)), block(64 * waves);
    if (hooks) hooks->before(st);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, lds, st, L); };
    if (wide) {
        if (gk) trace ? go(sw16m_kernel<true, true, 8>) : go(sw16m_kernel<false, true, 8>);
        else    trace ? go(sw16m_kernel<true, false, 8>) : go(sw16m_kernel<false, false, 8>);
    } else {
        if (gk) trace ? go(sw16m_kernel<true, true, 4>) : go(sw16m_kernel<false, true, 4>);
        else    trace ? go(sw16m_kernel<true, false, 4>) : go(sw16m_kernel<false, false, 4>);
    }
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

}  // namespace ga
