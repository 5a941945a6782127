// sw16.hip -- packed 16-bit Smith-Waterman for gfx950: two pairs per wave64,
// pair A in the low and pair B in the high half of every 32-bit register, so
// each VALU instruction advances two DP cells.
//
// Same recurrence, tie rules and end-cell rules as sw.hip (reference:
// include/genoaligner/backend/sw_kernel_impl.hip:22-168, :252-336); same frame:
// lane l owns pattern column j = 64*strip + l, the text streams through the wave
// with a one-row skew per lane, neighbour values move by DPP wave_shr:1.
//
// Arithmetic.  All values are unsigned 16-bit and never negative:
//   Eh = E + go,  Fh = F + go   (>= 0 wherever they can matter)
//   Eh[i][j] = max(H[i][j-1], Eh[i][j-1] -sat ge)
//   Fh[i][j] = max(H[i-1][j], Fh[i-1][j] -sat ge)
//   H[i][j]  = max((diag + u) -sat m0, max(Eh, Fh) -sat go)
// where -sat saturates at 0, m0 = max(0, -mismatch) and u = s + m0 >= 0 is the
// biased substitution score.  Negative E/F values never reach H nor a cell on a
// traceback path, so clamping them at -go changes no output.  Scores are
// limited to < 32768 - 255 - go; a pair that reaches that bound is flagged and
// re-run by the int32 kernel.
//
// Scoring (DNA over exactly the bytes A, C, G, T): each lane keeps, per half,
// a 4-byte table of u for its pattern byte against A/C/G/T; the text streams as
// per-row v_perm selectors (byte 0: code of pair A's text byte, byte 2: 4 + code
// of pair B's), so one v_perm_b32 yields both u values zero-extended.  A pair
// containing any other byte is flagged (literal byte semantics: N matches only
// N, case is significant) and re-run by the literal-byte pass of the same
// kernel (sw16_bytes_kernel: u by a packed compare of raw bytes) when the
// scheme's mismatch is negative, else by the int32 kernel.
//
// Traceback record (trace variant): 4 flags per cell and pair, taken from sign
// bits of packed differences --
//   nd  H did not come from the diagonal (max(Eh,Fh) - go > diag + s)
//   fe  F beat E                           (Fh > Eh)
//   ee  E extended                          (Eh_left -sat ge > H_left)
//   fx  F extended                          (Fh_up -sat ge > H_up)
// gathered with two v_perm_b32 and accumulated 4 steps per dword.  STOP needs
// no flag: the walk carries the exact H/E/F value of its current cell (score at
// the end cell; minus s on a diagonal step, plus ge on an extension, plus go on
// an opening) and stops where that value is 0, as the reference's value walk
// does (tests/sw/test_sw_trace.cpp:68-136).
#include "sw16_ops.hpp"

namespace ga {
namespace {

__device__ __forceinline__ int dna_code(uint8_t c)
{
    return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1;
}

// Score-table dword for one pattern byte: byte k = (q == "ACGT"[k] ? match : mismatch) + bias.
__device__ __forceinline__ uint32_t u_table(uint8_t q, int match, int mismatch, int bias)
{
    const uint32_t um = uint32_t(match + bias) & 0xffu, ux = uint32_t(mismatch + bias) & 0xffu;
    uint32_t t = 0;
    const uint8_t acgt[4] = {'A', 'C', 'G', 'T'};
#pragma unroll
    for (int k = 0; k < 4; ++k) t |= (q == acgt[k] ? um : ux) << (8 * k);
    return t;
}

constexpr uint32_t kSelNone = 0x0c0c0c0cu;   // v_perm selector: both halves u = 0

// Flag-word window of a walk (walk16): kWinC columns ending at the anchor
// column (one per lane of the walk's half-wave), and per column the kWinQ
// words that hold the kWinR rows ending at the anchor row.
constexpr size_t kWinBytes = size_t(2) * kWinC * kWinQ * 4;   // both walks of a wave

// Offset of the walk windows behind the trace staging area (2 CIGAR slots and
// both pairs' sequences: at most 4 * (max_m + max_n) bytes).
__host__ __device__ inline size_t sw16_win_offset(int max_m, int max_n)
{
    return (size_t(4) * size_t(max_m + max_n) + 15) & ~size_t(15);
}

// Rows of the LDS boundary column: the text rows rounded up to whole groups of
// 4 (at least one group).  8 bytes per row, so a 1024-row text takes 8 KB and
// 20 waves fit a CU's 160 KB.
__host__ __device__ inline int sw16_bnd_rows(int m)
{
    const int r = (m + 3) & ~3;
    return r > 4 ? r : 4;
}

// LDS layout: the boundary column (H and Eh of both pairs per text row).  Once
// the strips are done (trace variant) the same bytes hold the staging area and
// the windows.
__host__ __device__ inline size_t sw16_lds_bytes(int max_m, int max_n)
{
    const size_t bnd = size_t(sw16_bnd_rows(max_m)) * 8;
    const size_t stage = max_n > 0 ? sw16_win_offset(max_m, max_n) + kWinBytes : 0;
    return bnd > stage ? bnd : stage;
}

// The reference's three-state walk (sw_kernel_impl.hip:198-242) driven by the 4
// flags and the carried cell value, from the 1-based end cell (i, j) with H =
// score; writes the CIGAR right-aligned into cg[0 .. m+n) and returns its
// length; *si, *sj receive the 0-based start.  Both walks of a wave run at once
// (pair A on lane 0, pair B on lane 32; `walker` marks those lanes when their
// pair is walked).  The flag words come from LDS windows: when a walk leaves its
// window, its half-wave loads the kWinC x kWinR cells up and left of the
// current cell (lane c of the half loads column j - c, coalesced across the
// half), so a walk waits for one round of loads per window instead of one
// every few steps.  Every lane of the wave calls it; lanes other than the
// walkers return 0.  Flag word layout: cell (i, j) of half h sits in word
// (s * nq + (t >> 2)) * 64 + l, with s, l the strip and lane of column j - 1
// and t = i - 1 + l; with r = t & 3 and b = 8h + r,
//   !HMOD: bit b + 4 is nd, b + 20 fe, b ee and b + 16 fx;
//   HMOD:  bit b is nd, b + 16 fe, and bits 16h + 4 + 2(r & 1) + 8(r >> 1) and
//          the next hold H mod 4.  A gap's open-or-extend choice then follows
//          from the cell it came from: in E at (i, j) with value v, the gap
//          opened iff H(i, j-1) = v + go, and otherwise (extension, strictly
//          better) v + ge <= H(i, j-1) < v + go, so H mod 4 decides it when
//          go - ge <= 3 (F likewise with H(i-1, j)).  The walk reads that cell's
//          word too, so the window keeps a margin of one row and one column
//          (a window's words start at row iw - kWinR + 1 or up to 3 rows
//          lower, depending on the word alignment).
template <bool HMOD, class CP, class SP>
__device__ int walk16(const uint32_t* dirs, int nq, uint32_t* win, bool walker, int lane, SP T, SP P,
                      int m, int n, int i, int j, int score, int match, int mismatch, int go, int ge,
                      CP cg, int* si, int* sj)
{
    // rows and columns a step may use: with HMOD a gap step also reads the
    // cell above or to the left, which must be inside the window too
    constexpr int kRows = HMOD ? kWinR - 1 : kWinR;
    constexpr int kCols = HMOD ? kWinC - 1 : kWinC;
    const int hw = lane >> 5, c = lane & 31;
    uint32_t* w = win + hw * (kWinC * kWinQ);
    int pos = m + n, state = 0, v = score;
    bool active = walker;
    bool have = false;
    int iw = 0, jw = 0;                            // window anchor (walker lanes)
    // the window word of cell (ii, jj) (inside the window)
    auto word_at = [&](int ii, int jj) {
        const int l = (jj - 1) & 63, t = ii - 1 + l;
        return w[(jw - jj) * kWinQ + ((t >> 2) - ((iw - kWinR + l) >> 2))];
    };
    auto hmod_at = [&](int ii, int jj) {
        const int l = (jj - 1) & 63, r = (ii - 1 + l) & 3;
        return int((word_at(ii, jj) >> (16 * hw + 4 + 2 * (r & 1) + 8 * (r >> 1))) & 3u);
    };
    while (__any(active)) {
        if (active && (!(i > 0 && j > 0 && pos > 0) || (state == 0 && v == 0))) active = false;
        const bool need = active && !(have && iw - i < kRows && jw - j < kCols);
        const int need_h = __shfl(int(need), hw * 32);
        const int ih = __shfl(i, hw * 32), jh = __shfl(j, hw * 32);
        if (need_h) {
            const int jj = jh - c;                 // 1-based column of this lane
            if (jj >= 1) {
                const int l = (jj - 1) & 63, s = (jj - 1) >> 6;
                const int q0 = (ih - kWinR + l) >> 2;   // floor
                // every load issued before any is waited for (words outside
                // the table read word 0 and are replaced by 0)
                uint32_t val[kWinQ];
#pragma unroll
                for (int k = 0; k < kWinQ; ++k) {
                    const int q = q0 + k;
                    val[k] = dirs[(int64_t(s) * nq + (q >= 0 && q < nq ? q : 0)) * 64 + l];
                }
#pragma unroll
                for (int k = 0; k < kWinQ; ++k) {
                    const int q = q0 + k;
                    w[c * kWinQ + k] = (q >= 0 && q < nq) ? val[k] : 0u;
                }
            }
        }
        if (need) { iw = i; jw = j; have = true; }
        if (!active) continue;
        // at most kRows + kCols - 1 steps stay inside the window
        while (i > 0 && j > 0 && pos > 0 && !(state == 0 && v == 0) && iw - i < kRows && jw - j < kCols) {
            const uint32_t word = word_at(i, j);
            const int b = 8 * hw + ((((j - 1) & 63) + i - 1) & 3);
            const int bnd = HMOD ? b : b + 4, bfe = HMOD ? b + 16 : b + 20;
            if (state == 0 && !((word >> bnd) & 1u)) {
                const bool eq = T[i - 1] == P[j - 1];
                cg[--pos] = eq ? 'M' : 'X';
                v -= eq ? match : mismatch;
                --i; --j;
            } else if (state == 1 || (state == 0 && !((word >> bfe) & 1u))) {
                cg[--pos] = 'D';
                const bool ext = HMOD ? ((hmod_at(i, j - 1) - v - go) & 3) != 0 : ((word >> b) & 1u) != 0;
                state = ext ? 1 : 0;
                v += ext ? ge : go;
                --j;
            } else {
                cg[--pos] = 'I';
                const bool ext = HMOD ? ((hmod_at(i - 1, j) - v - go) & 3) != 0 : ((word >> (b + 16)) & 1u) != 0;
                state = ext ? 2 : 0;
                v += ext ? ge : go;
                --i;
            }
        }
    }
    *si = i;
    *sj = j;
    return walker ? m + n - pos : 0;
}

// The walk over the kLayHmod8 record (H mod 8 and fe per cell), in runs: the
// half-wave of a pair takes up to 31 steps of its current direction per round,
// lane c testing step c, and a ballot finds the first step that leaves the
// run.  Only the cell's value v and the record are needed:
//   H at (i, j), v > 0: the diagonal is taken iff H(i-1, j-1) + s = v (the
//     reference prefers the diagonal on ties).  Where it is not taken,
//     v - s - H(i-1, j-1) lies in (0, match - mismatch]: a path that reaches
//     (i, j) through a gap in row i (or column j) enters that row at some
//     column c' <= j - 1 from row i - 1, and the same path with the gap moved
//     to row i - 1 reaches (i-1, j-1) at a cost at most s(c') <= match lower.
//     So for match - mismatch <= 7, H mod 8 of the diagonal cell decides.
//     Otherwise fe (F > E, E winning ties) chooses the gap.  Along a diagonal
//     run, lane c's value is v minus the scores of steps 0 .. c-1: a ballot of
//     the matches and a count of the bits below the lane.
//   E at (i, j) with value v (a D): the gap opened iff H(i, j-1) = v + go, and
//     otherwise (extension, strictly better) v + ge <= H(i, j-1) < v + go; for
//     go - ge <= 7, H mod 8 decides.  Lane c tests the cell left of step c,
//     whose value is v + c ge if steps 0 .. c-1 all extended.  F alike with
//     the cell above (an I).
// The window is loaded as in walk16 (lane c of the half loads column jw - c,
// rows iw - 31 .. iw at least) whenever fewer than kRunMin steps of the
// current direction are left inside it.  Every lane of a walked half takes
// part; `walk_h` marks them.  Returns the CIGAR length in every lane of the
// half (0 where the half is not walked).
constexpr int kRunMin = 8;

template <class CP, class SP>
__device__ int walk16r(const uint32_t* dirs, int nq, uint32_t* win, bool walk_h, int lane, SP T, SP P,
                       int m, int n, int i, int j, int score, int match, int mismatch, int go, int ge,
                       CP cg, int* si, int* sj)
{
    const int hw = lane >> 5, c = lane & 31;
    uint32_t* w = win + hw * (kWinC * kWinQ);
    int pos = m + n, st = 0, v = score;            // st: 0 H, 1 E (D), 2 F (I)
    bool active = walk_h;
    bool have = false;
    int iw = 0, jw = 0;                            // window anchor (uniform per half)
    // the record nibble of cell (ii, jj), inside the window
    auto nibble = [&](int ii, int jj) {
        const int l = (jj - 1) & 63, t = ii - 1 + l, r = t & 3;
        const uint32_t wd = w[(jw - jj) * kWinQ + ((t >> 2) - ((iw - kWinR + l) >> 2))];
        return (wd >> (16 * (r & 1) + 8 * hw + 4 * (r >> 1))) & 0xfu;
    };
    while (__any(active)) {
        if (active && (i <= 0 || j <= 0 || pos <= 0 || (st == 0 && v <= 0))) active = false;
        if (!active) continue;
        // steps of the current direction left inside the window (cells up to
        // row iw - 31 and column jw - 31); a diagonal step also reads the
        // cell beyond it
        int R = 0;
        if (have) {
            const int ri = i - iw + 31, rj = j - jw + 31;
            R = st == 0 ? min(ri, rj) : st == 1 ? rj : ri;
            R = min(R, 31);
        }
        if (R < kRunMin) {
            const int jc = j - c;                  // 1-based column of this lane
            if (jc >= 1) {
                const int l = (jc - 1) & 63, s = (jc - 1) >> 6;
                const int q0 = (i - kWinR + l) >> 2;    // floor
                uint32_t val[kWinQ];
#pragma unroll
                for (int k = 0; k < kWinQ; ++k) {
                    const int q = q0 + k;
                    val[k] = dirs[(int64_t(s) * nq + (q >= 0 && q < nq ? q : 0)) * 64 + l];
                }
#pragma unroll
                for (int k = 0; k < kWinQ; ++k) {
                    const int q = q0 + k;
                    w[c * kWinQ + k] = (q >= 0 && q < nq) ? val[k] : 0u;
                }
            }
            iw = i; jw = j; have = true;
            R = 31;
        }
        const int k = st == 0 ? c : c + 1;
        const int ii = st == 1 ? i : i - k;
        const int jj = st == 2 ? j : j - k;
        const bool inb = ii >= 1 && jj >= 1;
        // a diagonal round reads cells 0 .. R (cell c + 1 is the diagonal
        // cell of step c), a gap round cells 0 .. R - 1
        const uint32_t nib = (inb && (st == 0 ? c <= R : c < R)) ? nibble(ii, jj) : 0u;
        const int hm = int(nib & 7u);
        bool brk;
        char ch;
        uint32_t eh = 0;
        if (st == 0) {
            // H mod 8 of the diagonal cell: the next lane's cell
            const int hp = __builtin_amdgcn_update_dpp(0, hm, 0x130, 0xf, 0xf, false);
            const bool eq = inb && T[ii - 1] == P[jj - 1];
            const uint64_t eb = __ballot(eq);
            eh = hw ? uint32_t(eb >> 32) : uint32_t(eb);
            const int nb = __popc(eh & ((1u << c) - 1u));
            const int vc = v - match * nb - mismatch * (c - nb);
            const int sc = eq ? match : mismatch;
            brk = c >= R || !inb || vc <= 0 || ((hp + sc - vc) & 7) != 0;
            ch = eq ? 'M' : 'X';
        } else {
            brk = c >= R || ((hm - v - c * ge - go) & 7) == 0;
            ch = st == 1 ? 'D' : 'I';
        }
        const uint64_t bb = __ballot(brk);
        const uint32_t bh = hw ? uint32_t(bb >> 32) : uint32_t(bb);
        const int L = __builtin_ctz(bh);            // lane 31 always breaks (R <= 31)
        const int cnt = st == 0 ? L : (L < R ? L + 1 : R);
        if (c < cnt && pos - 1 - c >= 0) cg[pos - 1 - c] = ch;
        pos -= cnt;
        if (st == 0) {
            const int nb = __popc(eh & ((1u << L) - 1u));
            const uint32_t nibL = uint32_t(__shfl(int(nib), 32 * hw + L));
            v -= match * nb + mismatch * (L - nb);
            i -= L;
            j -= L;
            // a break inside the reach: the border, H = 0 (both end the walk
            // at the loop head) or a gap, chosen by fe
            if (L < R) st = (nibL & 8u) ? 2 : 1;
            if (L < R && (i <= 0 || j <= 0 || v <= 0)) st = 0;
        } else {
            const bool opened = L < R;
            v += opened ? L * ge + go : R * ge;
            if (st == 1) j -= cnt; else i -= cnt;
            if (opened) st = 0;
        }
    }
    *si = i;
    *sj = j;
    return walk_h ? m + n - pos : 0;
}

// The score bias b = max(gap_open, -mismatch, 0) exceeds gap_open, so H needs
// the term b - go (same test as the kernel's bias).
inline bool sw16_gk(const Sw16Launch& L)
{
    const int m0 = L.mismatch < 0 ? -L.mismatch : 0;
    return m0 > L.go;
}

// The trace record layout: H mod 8 and fe when match - mismatch <= 7 and
// 0 <= go - ge <= 7 (walk16r explains why that suffices), else H mod 4 in place
// of the gap flags when 0 <= go - ge <= 3, else the four flags.
inline int sw16_layout(const Sw16Launch& L)
{
    if (L.go >= L.ge && L.go - L.ge <= 7 && L.match - L.mismatch <= 7) return kLayHmod8;
    if (L.go >= L.ge && L.go - L.ge <= 3) return kLayHmod4;
    return kLayFlags;
}

// At most 96 VGPRs (5 waves per SIMD) for the trace variant.
template <bool B> struct BoolC { static constexpr bool value = B; };
constexpr BoolC<true> kTrue{};
constexpr BoolC<false> kFalse{};

// The pass over wave w's two pairs.
// GK: gap_open < -mismatch, where H needs the extra bias term (sw16_gk).
// LAY (trace only): the trace record layout (sw16_layout).
// BYTES: the literal-byte pass: only the wave's pairs the first pass flagged
// (rerun == 2) are run, the other half is left as it is; text rows stream as
// raw bytes (0x100 past a pair's end, equal to no byte), each lane keeps its
// raw pattern bytes, and u = match or mismatch comes from a packed compare
// (three instructions where the A/C/G/T tables take one perm).  Rows past a
// pair's end then score the mismatch instead of u = 0: the drained-cell
// argument (s <= -1) needs mismatch < 0, which the launch requires.
template <bool TRACE, bool GK, int LAY, bool BYTES>
__device__ __forceinline__ void sw16_wave(const Sw16Launch& L, const int w)
{
    constexpr int kDepth = 3;
    extern __shared__ __align__(16) uint32_t lds16[];
    const int lane = threadIdx.x;
    const int ia = L.pair_a ? L.pair_a[w] : L.first_pair + 2 * w;
    int ib = L.pair_b ? L.pair_b[w] : L.first_pair + 2 * w + 1;
    if (ib >= L.n_total) ib = -1;
    bool liveA = true, liveB = ib >= 0;
    if (BYTES) {
        liveA = L.rerun[ia] == 2;
        liveB = ib >= 0 && L.rerun[ib] == 2;
    }
    const int mA = liveA ? L.tlen[ia] : 0, nA = liveA ? L.plen[ia] : 0;
    const int mB = liveB ? L.tlen[ib] : 0, nB = liveB ? L.plen[ib] : 0;
    const int mAp = mA > 0 && nA > 0 ? mA : 0, mBp = mB > 0 && nB > 0 ? mB : 0;
    const int M = mAp > mBp ? mAp : mBp;
    const int N = nA > nB ? nA : nB;
    const uint8_t* TA = L.txt + L.toff[ia];
    const uint8_t* PA = L.pat + L.poff[ia];
    const uint8_t* TB = ib >= 0 ? L.txt + L.toff[ib] : TA;
    const uint8_t* PB = ib >= 0 ? L.pat + L.poff[ib] : PA;

    // Per-row selectors (scratch in HBM/L2, guard rows on both sides) and, in
    // LDS, the strip boundary column (H and Eh of both pairs per text row).
    uint32_t* sel = L.selbuf + int64_t(w) * sw16_sel_stride(L.max_m) + 64;
    uint2* bnd = reinterpret_cast<uint2*>(lds16);
    const int go = L.go, ge = L.ge;
    const uint32_t go2 = uint32_t(go) * 0x10001u, ge2 = uint32_t(ge) * 0x10001u;
    const int m0 = L.mismatch < 0 ? -L.mismatch : 0;
    const int bias = go > m0 ? go : m0;            // u = s + bias >= 0
    const uint32_t b2 = uint32_t(bias) * 0x10001u, gk2 = uint32_t(bias - go) * 0x10001u;
    const int Mr = sw16_bnd_rows(M);             // boundary rows in use (<= the LDS rows)

    bool badA = false, badB = false;
    for (int i = lane - 64; i < M + 80; i += 64) {
        uint32_t sv = BYTES ? 0x01000100u : kSelNone;
        if (i >= 0 && i < M) {
            if (BYTES) {
                sv = (i < mAp ? uint32_t(TA[i]) : 0x100u) | ((i < mBp ? uint32_t(TB[i]) : 0x100u) << 16);
            } else {
                uint32_t lo = 12u, hi = 12u;
                if (i < mAp) { const int c = dna_code(TA[i]); badA |= c < 0; lo = uint32_t(c < 0 ? 12 : c); }
                if (i < mBp) { const int c = dna_code(TB[i]); badB |= c < 0; hi = uint32_t(c < 0 ? 12 : 4 + c); }
                sv = lo | (12u << 8) | (hi << 16) | (12u << 24);
            }
        }
        sel[i] = sv;
    }
    for (int i = lane; i < Mr; i += 64) bnd[i] = make_uint2(0u, go2);
    if (!BYTES) {
        for (int j = lane; j < N; j += 64) {
            if (j < nA && mAp > 0) badA |= dna_code(PA[j]) < 0;
            if (j < nB && mBp > 0) badB |= dna_code(PB[j]) < 0;
        }
        badA = __any(badA);
        badB = __any(badB);
    }
    // bytes: u = um + z (ux - um) per half, z = min(text byte ^ pattern byte, 1)
    const uint32_t um2 = uint32_t(L.match + bias) * 0x10001u;
    const uint32_t d2 = (uint32_t(L.mismatch - L.match) & 0xffffu) * 0x10001u;
    __syncthreads();

    const int nstrips = (N + 63) >> 6;
    const int nsteps = (M + 63 + 3) & ~3;
    const int nq = nsteps >> 2;
    const int nmin = nA < nB ? nA : nB;
    uint32_t* dirs = TRACE ? L.dirs + int64_t(w) * L.dir_stride : nullptr;
    // buffer resource over this wave's trace words (stores beyond them are
    // dropped by the range check)
    const __amdgpu_buffer_rsrc_t drs =
        __builtin_amdgcn_make_buffer_rsrc(TRACE ? static_cast<void*>(dirs) : static_cast<void*>(L.res), 0, TRACE ? int(L.dir_stride * 4) : 0, 0x00020000);
    int gH[2] = {0, 0}, gI[2] = {-1, -1}, gJ[2] = {-1, -1};
    uint32_t thrm1 = 0;                           // per half: the pair's best so far - 1 (uniform)

    for (int s = 0; s < nstrips; ++s) {
        const int j = (s << 6) + lane;
        Strip S;
        if (BYTES) {
            // the lane's raw pattern bytes (0x200 past a pattern's end: a masked column)
            S.tabA = (j < nA ? uint32_t(PA[j]) : 0x200u) | ((j < nB ? uint32_t(PB[j]) : 0x200u) << 16);
            S.tabB = 0;
        } else {
            S.tabA = u_table(j < nA ? PA[j] : uint8_t(0), L.match, L.mismatch, bias);
            S.tabB = u_table(j < nB ? PB[j] : uint8_t(0), L.match, L.mismatch, bias);
        }
        S.colmask = (j < nA ? 0xffffu : 0u) | (j < nB ? 0xffff0000u : 0u);
        S.H = 0; S.Eh = go2; S.Fh = go2; S.diag = 0;
        S.bestH = thrm1; S.bestT = 0xffffffffu;     // record at the bound (see strip_group)
        const bool full = (s << 6) + 64 <= nmin;  // every column of the strip exists in both pairs

        // Selectors and boundary rows of group t0 are read kDepth groups ahead,
        // into one of kDepth register sets that rotate with the group index, so
        // no register copy waits for a load in flight.  Every lane reads the
        // group's boundary rows (one LDS address, a broadcast); only lane 0
        // keeps them (shr1).  Groups past the last one repeat it: those rows
        // only reach drained cells.  Guard rows keep every selector read inside
        // its buffer.
        const uint32_t* sp = sel - lane;
        auto ldg = [&](int t, uint32_t (&sv)[4], uint2 (&bv)[4]) {
            const uint2* b = bnd + min(t, Mr - 4);
#pragma unroll
            for (int u = 0; u < 4; ++u) { sv[u] = sp[t + u]; bv[u] = b[u]; }
        };
        uint32_t sv[kDepth][4];
        uint2 bv[kDepth][4];
#pragma unroll
        for (int k = 0; k < kDepth; ++k) ldg(4 * k, sv[k], bv[k]);
        // Wait here for these sets: each group then waits only for what the
        // loop itself issued (with the stores of the trace variant between the
        // loads, the wait counts in the loop would otherwise be set by this
        // prologue and hold up every iteration until an older store is done).
#pragma unroll
        for (int k = 0; k < kDepth; ++k) asm volatile("" ::"v"(sv[k][0]));
        // A strip whose columns all exist in both pairs needs no mask at all.
        const int t_fast = full ? nsteps : 0;
        // For lane 63, the byte offset of boundary row t0 - 63 in the LDS
        // column: kept in a VGPR that advances with the groups (an offset the
        // compiler sees as uniform lives in an SGPR and costs a move into a
        // VGPR before every store).
        int boff = -63 * 8;
        asm volatile("" : "+v"(boff));
        auto group = [&](auto fast, int t0, const uint32_t (&sv)[4], const uint2 (&bv)[4]) {
            uint32_t acc = 0, hout[4], eout[4];
            // both pairs' u: from the lane's tables by the row's v_perm selector,
            // or (BYTES) by comparing the row's text bytes with the lane's
            // pattern bytes
            const auto dna_u = [&](int u) {
                if (BYTES) {
                    const uint32_t z = as_u(__builtin_elementwise_min(as_v(sv[u] ^ S.tabA), as_v(0x00010001u)));
                    return as_u(as_v(z) * as_v(d2) + as_v(um2));
                }
                return __builtin_amdgcn_perm(S.tabB, S.tabA, sv[u]);
            };
            if (decltype(fast)::value || t0 < t_fast)
                strip_group<TRACE, LAY, false, GK>(S, t0, lane, dna_u, bv, b2, gk2, ge2, acc, hout, eout, thrm1);
            else
                strip_group<TRACE, LAY, true, true>(S, t0, lane, dna_u, bv, b2, gk2, ge2, acc, hout, eout, thrm1);
            // the group's trace word: lane offset in a VGPR, group offset in an
            // SGPR (no per-group address arithmetic)
            if (TRACE)
                __builtin_amdgcn_raw_buffer_store_b32(acc, drs, lane * 4, (s * nq + (t0 >> 2)) * 256, 0);
            // lane 63 hands rows t0-63 .. t0-60 to the next strip (lane 0 has
            // read each row before it is written); rows outside 0 .. Mr-1 are
            // dropped, and rows M .. Mr-1 only reach drained cells
            uint2* b = reinterpret_cast<uint2*>(reinterpret_cast<char*>(bnd) + (boff + t0 * 8));
            if (lane == 63) {
                if (t0 >= 64 && t0 - 60 < Mr) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) b[u] = make_uint2(hout[u], eout[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int r = t0 + u - 63;
                        if (r >= 0 && r < Mr) b[u] = make_uint2(hout[u], eout[u]);
                    }
                }
            }
        };
        constexpr int kSpan = 4 * kDepth;
        int t0 = 0;
        for (; t0 + kSpan <= t_fast; t0 += kSpan) {
#pragma unroll
            for (int k = 0; k < kDepth; ++k) {
                group(kTrue, t0 + 4 * k, sv[k], bv[k]);
                ldg(t0 + 4 * k + kSpan, sv[k], bv[k]);
            }
        }
        for (; t0 < nsteps; t0 += kSpan) {
#pragma unroll
            for (int k = 0; k < kDepth; ++k) {
                if (t0 + 4 * k >= nsteps) break;
                group(kFalse, t0 + 4 * k, sv[k], bv[k]);
                ldg(t0 + 4 * k + kSpan, sv[k], bv[k]);
            }
        }
        // fold this strip's first maxima (per half) into the running records
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int bh = int((S.bestH >> (16 * h)) & 0xffffu);
            const int bt = int((S.bestT >> (16 * h)) & 0xffffu);
            const int bi = bh > 0 ? bt - lane : -1;
            if (bh > gH[h] || (bh == gH[h] && bh > 0 && bi < gI[h])) { gH[h] = bh; gI[h] = bi; gJ[h] = j; }
        }
        __syncthreads();
    }

    const int limit = 32767 - 255 - bias;
    int endI[2] = {-1, -1}, endJ[2] = {-1, -1}, endS[2] = {0, 0};
    bool redo[2] = {false, false};
    // literal-byte pass for pairs holding another byte (the launch provides
    // the list and a scheme with mismatch < 0); else the int32 kernel
    const bool to_bytes = !BYTES && L.bytes_list != nullptr && L.mismatch < 0;
    bool listed = false;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int pid = h == 0 ? ia : ib;
        if (pid < 0 || !(h == 0 ? liveA : liveB)) continue;
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
        endI[h] = rI; endJ[h] = rJ; endS[h] = lH;
        const bool bad = h == 0 ? badA : badB;
        redo[h] = bad || lH >= limit;
        const uint8_t flag = !redo[h] ? 0 : (bad && to_bytes) ? 2 : 1;
        listed |= flag == 2;
        if (lane == 0) {
            int32_t* res = L.res + int64_t(pid) * 8;
            const bool rerun = redo[h];
            const int m = h == 0 ? mA : mB, n = h == 0 ? nA : nB;
            if (m <= 0 || n <= 0) {
                res[0] = 0;
                for (int q = 1; q < 7; ++q) res[q] = -1;
                res[7] = 0;
                L.rerun[pid] = 0;
            } else {
                res[0] = rerun ? kSwRerun : lH;
                res[1] = rI; res[2] = rJ; res[3] = lI; res[4] = lJ;
                res[5] = -1; res[6] = -1; res[7] = 0;
                L.rerun[pid] = flag;
            }
        }
    }
    if (listed && lane == 0) L.bytes_list[1 + atomicAdd(&L.bytes_list[0], 1)] = w;
    if (!TRACE) return;
    // The walks, fused: lane 0 walks pair A, lane 32 pair B, while the wave's
    // slot would otherwise sit idle; their load latency overlaps the forward
    // work of the other waves on the SIMD.  Both pairs' sequences and CIGARs are
    // staged in the LDS (free once the strips are done): the walk's only global
    // accesses are the flag words, and the CIGARs leave in coalesced stores.
    __syncthreads();
    uint8_t* st8 = reinterpret_cast<uint8_t*>(lds16);
    const int slotA = mA + nA, slotB = mB + nB;
    uint8_t* cgA = st8;
    uint8_t* cgB = cgA + slotA;
    uint8_t* sTA = cgB + slotB;
    uint8_t* sPA = sTA + mA;
    uint8_t* sTB = sPA + nA;
    uint8_t* sPB = sTB + mB;
    for (int q = lane; q < mA; q += 64) sTA[q] = TA[q];
    for (int q = lane; q < nA; q += 64) sPA[q] = PA[q];
    for (int q = lane; q < mB; q += 64) sTB[q] = TB[q];
    for (int q = lane; q < nB; q += 64) sPB[q] = PB[q];
    __syncthreads();
    uint32_t* win = reinterpret_cast<uint32_t*>(st8 + sw16_win_offset(L.max_m, L.max_n));
    const int hw = lane >> 5;
    const int pidw = hw == 0 ? ia : ib;
    const bool walker = (lane & 31) == 0 && pidw >= 0 && !redo[hw] && endS[hw] > 0 &&
                        (!L.walk || L.walk[pidw]);
    int si = -1, sj = -1;
    int len_h;
    if (LAY == kLayHmod8) {
        const bool walk_h = pidw >= 0 && !redo[hw] && endS[hw] > 0 && (!L.walk || L.walk[pidw]);
        len_h = walk16r(dirs, nq, win, walk_h, lane, hw == 0 ? sTA : sTB, hw == 0 ? sPA : sPB,
                        hw == 0 ? mA : mB, hw == 0 ? nA : nB, endI[hw] + 1, endJ[hw] + 1,
                        endS[hw], L.match, L.mismatch, go, ge, hw == 0 ? cgA : cgB, &si, &sj);
    } else {
        len_h = walk16<LAY == kLayHmod4>(dirs, nq, win, walker, lane, hw == 0 ? sTA : sTB,
                                         hw == 0 ? sPA : sPB, hw == 0 ? mA : mB, hw == 0 ? nA : nB,
                                         endI[hw] + 1, endJ[hw] + 1, endS[hw], L.match, L.mismatch,
                                         go, ge, hw == 0 ? cgA : cgB, &si, &sj);
    }
    if (walker) {
        int32_t* res = L.res + int64_t(pidw) * 8;
        res[5] = si;
        res[6] = sj;
        res[7] = len_h;
    }
    __syncthreads();
    const int lenA = __shfl(len_h, 0), lenB = __shfl(len_h, 32);
    if (lenA > 0) {
        uint8_t* dst = L.cigar + L.cig_off[ia] + (slotA - lenA);
        for (int q = lane; q < lenA; q += 64) dst[q] = cgA[slotA - lenA + q];
    }
    if (lenB > 0) {
        uint8_t* dst = L.cigar + L.cig_off[ib] + (slotB - lenB);
        for (int q = lane; q < lenB; q += 64) dst[q] = cgB[slotB - lenB + q];
    }
}

template <bool TRACE, bool GK, int LAY>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
void sw16_kernel(Sw16Launch L)
{
    sw16_wave<TRACE, GK, LAY, false>(L, int(blockIdx.x));
}

// The literal-byte pass over the waves the first pass listed: a small grid
// takes them in turn (most launches list none, and a full grid of waves that
// exit at once would still cost its dispatch).
template <bool TRACE, bool GK, int LAY>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
void sw16_bytes_kernel(Sw16Launch L)
{
    const int n = __builtin_amdgcn_readfirstlane(L.bytes_list[0]);
    for (int k = int(blockIdx.x); k < n; k += int(gridDim.x)) {
        sw16_wave<TRACE, GK, LAY, true>(L, __builtin_amdgcn_readfirstlane(L.bytes_list[1 + k]));
        __syncthreads();                         // the LDS is the next wave's
    }
}

__global__ void sw16_zero_kernel(int32_t* p) { if (threadIdx.x == 0) p[0] = 0; }

template <bool GK, int LAY>
void sw16_launch_t(const Sw16Launch& L, bool trace, unsigned bytes_grid, const dim3& grid, size_t lds,
                   hipStream_t st)
{
    const dim3 block(64);
    if (lds > 65536) {
        // above 64 KB of dynamic LDS the kernels are told first (the longest
        // texts with long patterns in a CIGAR batch)
        const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
        const void* k[4] = {reinterpret_cast<const void*>(&sw16_kernel<true, GK, LAY>),
                            reinterpret_cast<const void*>(&sw16_kernel<false, GK, kLayFlags>),
                            reinterpret_cast<const void*>(&sw16_bytes_kernel<true, GK, LAY>),
                            reinterpret_cast<const void*>(&sw16_bytes_kernel<false, GK, kLayFlags>)};
        for (const void* f : k) (void)hipFuncSetAttribute(f, a, int(lds));
    }
    if (trace) hipLaunchKernelGGL((sw16_kernel<true, GK, LAY>), grid, block, lds, st, L);
    else       hipLaunchKernelGGL((sw16_kernel<false, GK, kLayFlags>), grid, block, lds, st, L);
    if (bytes_grid == 0) return;
    if (trace) hipLaunchKernelGGL((sw16_bytes_kernel<true, GK, LAY>), dim3(bytes_grid), block, lds, st, L);
    else       hipLaunchKernelGGL((sw16_bytes_kernel<false, GK, kLayFlags>), dim3(bytes_grid), block, lds, st, L);
}

}  // namespace

int64_t sw16_lds_need(int max_m, int max_n, bool trace)
{
    return int64_t(sw16_lds_bytes(max_m, trace ? max_n : 0));
}

hipError_t sw16_launch(const Sw16Launch& L, bool trace, hipStream_t st, LaunchHooks* hooks)
{
    if (L.n_waves <= 0) return hipSuccess;
    const size_t lds = sw16_lds_bytes(L.max_m, trace ? L.max_n : 0);
    const dim3 grid(unsigned(L.n_waves));
    if (hooks) hooks->before(st);
    const int lay = sw16_layout(L);
    // the literal-byte pass runs when the launch gives its list and the scheme
    // allows it; its counter is cleared first by a kernel (a memset node is
    // not used, so the sequence stays safe to capture)
    const bool bytes = L.bytes_list != nullptr && L.mismatch < 0;
    const unsigned bytes_grid = bytes ? unsigned(std::min(L.n_waves, 1024)) : 0u;
    if (bytes) hipLaunchKernelGGL(sw16_zero_kernel, dim3(1), dim3(64), 0, st, L.bytes_list);
    const bool gk = sw16_gk(L);
    if (gk && lay == kLayHmod8)       sw16_launch_t<true, kLayHmod8>(L, trace, bytes_grid, grid, lds, st);
    else if (gk && lay == kLayHmod4)  sw16_launch_t<true, kLayHmod4>(L, trace, bytes_grid, grid, lds, st);
    else if (gk)                      sw16_launch_t<true, kLayFlags>(L, trace, bytes_grid, grid, lds, st);
    else if (lay == kLayHmod8)        sw16_launch_t<false, kLayHmod8>(L, trace, bytes_grid, grid, lds, st);
    else if (lay == kLayHmod4)        sw16_launch_t<false, kLayHmod4>(L, trace, bytes_grid, grid, lds, st);
    else                              sw16_launch_t<false, kLayFlags>(L, trace, bytes_grid, grid, lds, st);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

}  // namespace ga
