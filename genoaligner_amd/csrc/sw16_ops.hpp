// sw16_ops.hpp -- packed 16-bit Smith-Waterman arithmetic shared by the DNA
// kernel (sw16.hip: scores from per-lane v_perm tables) and the
// substitution-matrix kernel (sw16m.hip: scores from per-lane LDS tables).
// Two pairs per wave64, pair A in the low and pair B in the high half of every
// 32-bit register; the recurrence, the trace flags and the first-maximum
// bookkeeping of one group of four text rows are strip_group below.  The
// arithmetic and its invariants are described at the top of sw16.hip.
#pragma once

#include "ga_device.hpp"

namespace ga {
namespace {

typedef short v2i16 __attribute__((ext_vector_type(2)));
typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t as_u(v2u16 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ v2u16 as_v(uint32_t x) { return __builtin_bit_cast(v2u16, x); }

// per-half unsigned max, saturating subtract and add; signed difference
__device__ __forceinline__ uint32_t pmax(uint32_t a, uint32_t b)
{
    return as_u(__builtin_elementwise_max(as_v(a), as_v(b)));
}
__device__ __forceinline__ uint32_t psubs(uint32_t a, uint32_t b)
{
    return as_u(__builtin_elementwise_sub_sat(as_v(a), as_v(b)));
}
// saturating at 65535, so a pair whose scores outgrow 16 bits keeps a maximum
// above the re-run threshold instead of wrapping below it
__device__ __forceinline__ uint32_t padd(uint32_t a, uint32_t b)
{
    return as_u(__builtin_elementwise_add_sat(as_v(a), as_v(b)));
}
__device__ __forceinline__ uint32_t pdiff(uint32_t a, uint32_t b)   // a - b per half, signed
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2i16, a) - __builtin_bit_cast(v2i16, b));
}
// lane l <- lane l-1 (DPP wave_shr:1); lane 0, which has no source lane, keeps
// `old` (bound_ctrl off), the boundary value: one v_mov_b32_dpp into the
// register the boundary row was read into.
__device__ __forceinline__ uint32_t shr1(uint32_t old, uint32_t v)
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(v), 0x138, 0xf, 0xf, false));
}

// (m & a) | (~m & b) as one v_bfi_b32 (written out, since the compiler turns
// the C form with constant masks into separate and/or instructions)
__device__ __forceinline__ uint32_t bsel(uint32_t m, uint32_t a, uint32_t b)
{
    uint32_t d;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(d) : "s"(m), "v"(a), "v"(b));
    return d;
}

// Packed u16 maximum over the wave, as a uniform value: row shifts leave each
// row's maximum in its lane 15, the row broadcasts 15 and 31 carry them up to
// lane 63 (lanes without a source read 0, the identity).
__device__ __forceinline__ uint32_t wave_pmax(uint32_t v)
{
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, true)));
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, true)));
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, true)));
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, true)));
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false)));
    v = pmax(v, uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false)));
    return uint32_t(__builtin_amdgcn_readlane(int(v), 63));
}

// Flag-word window of a traceback walk: kWinC columns ending at the anchor
// column (one per lane of the walk's half-wave), and per column the kWinQ
// words that hold the kWinR rows ending at the anchor row.
constexpr int kWinR = 32, kWinC = 32, kWinQ = kWinR / 4 + 1;

// Per-lane state of one strip.
struct Strip {
    uint32_t H, Eh, Fh, diag;      // last row's H, Eh, Fh; H of the left column one row up
    uint32_t bestH, bestT;         // per half: first maximum of this column (at least thr - 1) and its step
    uint32_t tabA, tabB;           // u tables of this lane's pattern bytes
    uint32_t colmask;              // 0xffff per half whose column exists
};

// Trace record layouts (4 bits per cell and pair, 4 rows per dword):
//   kLayFlags  nd, fe, ee, fx as flags
//   kLayHmod4  nd, fe as flags, H mod 4 in place of ee, fx
//   kLayHmod8  H mod 8 and fe only (the walk derives nd from H mod 8)
constexpr int kLayFlags = 0, kLayHmod4 = 1, kLayHmod8 = 2;

// Four steps t0 .. t0+3 of a strip.  score(u) gives the biased scores u of
// both pairs for step t0 + u (one per half), bv the group's boundary rows
// (prefetched by the caller).  COLS: columns past a pair's end are masked out
// of the maxima (partial strips only).  LAY: the trace record layout.
//
// Rows past a pair's end (the drain, and the shorter pair of a wave) need no
// mask: their selectors give u = 0, i.e. s = -bias <= -1, and gap costs are at
// least 1, so each such cell's H, E and F are below a value of an earlier
// cell: by induction H <= max(0, R - 1), R being the pair's largest real H.
// They never reach the pair's maximum, and real cells never depend on them.
template <bool TRACE, int LAY, bool COLS, bool GK, class Score>
__device__ __forceinline__ void strip_group(Strip& S, int t0, int lane, Score score,
                                            const uint2 (&bv)[4], uint32_t b2, uint32_t gk2,
                                            uint32_t ge2, uint32_t& acc, uint32_t (&hout)[4],
                                            uint32_t (&eout)[4], uint32_t& thrm1)
{
    uint32_t Hb[4], Q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const uint32_t Hl = shr1(bv[u].x, S.H);
        const uint32_t El = shr1(bv[u].y, S.Eh);
        const uint32_t uu = score(u);
        const uint32_t xe = psubs(El, ge2);
        const uint32_t En = pmax(Hl, xe);
        const uint32_t xf = psubs(S.Fh, ge2);
        const uint32_t Fn = pmax(S.H, xf);
        // H = max(0, diag + s, max(E, F)) with u = s + b and Eh, Fh = E, F + go:
        //   = sat(max(diag + u, max(Eh, Fh) + (b - go)) - b), and b = go in the
        // usual case gap_open >= -mismatch, where the middle add vanishes.
        const uint32_t A1 = padd(S.diag, uu);
        const uint32_t G = pmax(En, Fn);
        const uint32_t A2 = GK ? padd(G, gk2) : G;
        const uint32_t Hn = psubs(pmax(A1, A2), b2);
        uint32_t hb = Hn;
        if (COLS) hb &= S.colmask;
        Hb[u] = hb;
        if (TRACE && LAY == kLayHmod8) {
            Q[u] = pdiff(En, Fn);                 // sign bit of each half: fe
        } else if (TRACE) {
            // nd: A2 > A1 (both carry the same bias, so the comparison is exact).
            // v_perm selectors 8..11 replicate the sign bit of a half into a
            // whole byte (8: S1 low half, 9: S1 high, 10: S0 low, 11: S0 high),
            // so each perm yields 0xff/0x00 byte masks for two flags of both
            // pairs; Q[u] keeps nd/fe in the high and ee/fx in the low nibble.
            const uint32_t P1 = __builtin_amdgcn_perm(pdiff(En, Fn), pdiff(A1, A2), 0x0b0a0908u);
            if (LAY == kLayHmod4) {
                Q[u] = P1;
            } else {
                const uint32_t P2 = __builtin_amdgcn_perm(pdiff(S.H, xf), pdiff(Hl, xe), 0x0b0a0908u);
                Q[u] = bsel(0xf0f0f0f0u, P1, P2);
            }
        }
        hout[u] = Hn;
        eout[u] = En;
        S.diag = Hl;
        S.H = Hn;
        S.Eh = En;
        S.Fh = Fn;
    }
    // Row u of the group goes to bit 4+u (nd, fe) and bit u (ee, fx) of each
    // byte: bits alternate between rows 0/1 and 2/3, then pairs of bits
    // between the two.  kLayHmod4: the flags give bits u and 4+u copies of nd, fe;
    // the high nibbles then take H mod 4 of each row (one shift and one
    // bit-select per row, where the two gap flags took two differences, a
    // perm and a select).
    // kLayHmod8: row r of the group and half h take the nibble at bit
    // 16 (r & 1) + 8 h + 4 (r >> 1): H mod 8 in its low 3 bits, fe in bit 3.
    // Two perms gather the low bytes of H (rows 0/1, rows 2/3), two more the
    // high bytes of the differences (their sign bits: fe); the fe bytes of
    // rows 0/1 are gathered one byte rotated so that one rotate by 4 bits
    // puts them at bit 3 of the right byte, next to H mod 8 of rows 2/3.
    if (TRACE && LAY == kLayHmod8) {
        const uint32_t h01 = __builtin_amdgcn_perm(hout[1], hout[0], 0x06040200u);
        const uint32_t h23 = __builtin_amdgcn_perm(hout[3], hout[2], 0x06040200u);
        const uint32_t d23 = __builtin_amdgcn_perm(Q[3], Q[2], 0x07050301u);
        const uint32_t d01 = __builtin_amdgcn_perm(Q[1], Q[0], 0x01070503u);
        const uint32_t lo = bsel(0x07070707u, h01, d23);
        const uint32_t y = bsel(0x07070707u, h23, d01);
        acc = bsel(0x87878787u, lo, __builtin_amdgcn_alignbit(y, y, 28));
    } else if (TRACE) {
        acc = bsel(0x33333333u, bsel(0x55555555u, Q[0], Q[1]), bsel(0x55555555u, Q[2], Q[3]));
        if (LAY == kLayHmod4) {
            acc = bsel(0x00030003u << 4, hout[0] << 4, acc);
            acc = bsel(0x00030003u << 6, hout[1] << 6, acc);
            acc = bsel(0x00030003u << 12, hout[2] << 12, acc);
            acc = bsel(0x00030003u << 14, hout[3] << 14, acc);
        }
    }
    // First maxima.  Only a cell at or above the pair's best so far can be the
    // end cell, so a lane's column record needs the exact per-step update only
    // when its H exceeds max(record, thr - 1), thr being the best over the wave
    // (per half, uniform; thrm1 = thr - 1).  The record is kept at that bound
    // (raised to it with a stale step where the column's own maximum is lower:
    // such a record is below the pair's maximum and loses the final
    // comparison), so one comparison with it suffices.  Without the bound
    // nearly every group of a strip had some column reaching a new maximum.
    const uint32_t nb = pmax(pmax(S.bestH, Hb[0]), pmax(pmax(Hb[1], Hb[2]), Hb[3]));
    if (__any(nb != S.bestH)) {
        uint32_t b = S.bestH;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // 0xffff per half where Hb[u] > b: v_perm selectors 8 and 9
            // replicate the sign bits of the low and high half of b - Hb[u]
            const uint32_t up = __builtin_amdgcn_perm(0u, pdiff(b, Hb[u]), 0x09090808u);
            b = pmax(b, Hb[u]);
            S.bestT = (up & (uint32_t(t0 + u) * 0x10001u)) | (~up & S.bestT);
        }
        thrm1 = psubs(wave_pmax(b), 0x00010001u);
        S.bestH = pmax(b, thrm1);
    }
}

}  // namespace
}  // namespace ga
