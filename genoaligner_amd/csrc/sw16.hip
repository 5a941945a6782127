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
// containing any other byte is flagged and re-run by the int32 kernel (literal
// byte semantics: N matches only N, case is significant).
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
__device__ __forceinline__ uint32_t psign_mask(uint32_t d)          // 0xffff where half < 0
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(v2i16, d) >> short(15));
}
__device__ __forceinline__ uint32_t shr1(uint32_t old, uint32_t v)  // lane l <- lane l-1; lane 0 <- old
{
    return uint32_t(__builtin_amdgcn_update_dpp(int(old), int(v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ int dna_code(uint8_t c)
{
    return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1;
}

// Score-table dword for one pattern byte: byte k = (q == "ACGT"[k] ? match : mismatch) + m0.
__device__ __forceinline__ uint32_t u_table(uint8_t q, int match, int mismatch, int m0)
{
    const uint32_t um = uint32_t(match + m0) & 0xffu, ux = uint32_t(mismatch + m0) & 0xffu;
    uint32_t t = 0;
    const uint8_t acgt[4] = {'A', 'C', 'G', 'T'};
#pragma unroll
    for (int k = 0; k < 4; ++k) t |= (q == acgt[k] ? um : ux) << (8 * k);
    return t;
}

constexpr uint32_t kSelNone = 0x0c0c0c0cu;   // v_perm selector: both halves u = 0

template <bool TRACE>
__global__ __launch_bounds__(64) void sw16_kernel(Sw16Launch L)
{
    extern __shared__ uint32_t lds16[];
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int ia = L.pair_a ? L.pair_a[w] : L.first_pair + 2 * w;
    int ib = L.pair_b ? L.pair_b[w] : L.first_pair + 2 * w + 1;
    if (ib >= L.n_total) ib = -1;
    const int mA = L.tlen[ia], nA = L.plen[ia];
    const int mB = ib >= 0 ? L.tlen[ib] : 0, nB = ib >= 0 ? L.plen[ib] : 0;
    const int mAp = mA > 0 && nA > 0 ? mA : 0, mBp = mB > 0 && nB > 0 ? mB : 0;
    const int M = mAp > mBp ? mAp : mBp;
    const int N = nA > nB ? nA : nB;
    const uint8_t* TA = L.txt + L.toff[ia];
    const uint8_t* PA = L.pat + L.poff[ia];
    const uint8_t* TB = ib >= 0 ? L.txt + L.toff[ib] : TA;
    const uint8_t* PB = ib >= 0 ? L.pat + L.poff[ib] : PA;

    // LDS: per-row selectors with 64 guard rows on each side, then the strip
    // boundary column (H and Eh of both pairs, one uint2 per text row).
    uint32_t* sel = lds16 + 64;
    uint2* bnd = reinterpret_cast<uint2*>(lds16 + (L.max_m + 136));
    const int go = L.go, ge = L.ge;
    const uint32_t go2 = uint32_t(go) * 0x10001u, ge2 = uint32_t(ge) * 0x10001u;
    const int m0 = L.mismatch < 0 ? -L.mismatch : 0;
    const uint32_t m02 = uint32_t(m0) * 0x10001u;

    bool badA = false, badB = false;
    for (int i = lane - 64; i < M + 72; i += 64) {
        uint32_t s = kSelNone;
        if (i >= 0 && i < M) {
            uint32_t lo = 12u, hi = 12u;
            if (i < mAp) { const int c = dna_code(TA[i]); badA |= c < 0; lo = uint32_t(c < 0 ? 12 : c); }
            if (i < mBp) { const int c = dna_code(TB[i]); badB |= c < 0; hi = uint32_t(c < 0 ? 12 : 4 + c); }
            s = lo | (12u << 8) | (hi << 16) | (12u << 24);
        }
        sel[i] = s;
    }
    for (int i = lane; i < M; i += 64) bnd[i] = make_uint2(0u, go2);
    for (int j = lane; j < N; j += 64) {
        if (j < nA && mAp > 0) badA |= dna_code(PA[j]) < 0;
        if (j < nB && mBp > 0) badB |= dna_code(PB[j]) < 0;
    }
    badA = __any(badA);
    badB = __any(badB);
    __syncthreads();

    if (TRACE && lane == 0) L.wave_m[w] = M;
    const int nstrips = (N + 63) >> 6;
    const int nsteps = (M + 63 + 3) & ~3;
    const int nq = nsteps >> 2;
    uint32_t* dirs = TRACE ? L.dirs + int64_t(w) * L.dir_stride : nullptr;
    int gH[2] = {0, 0}, gI[2] = {-1, -1}, gJ[2] = {-1, -1};

    for (int s = 0; s < nstrips; ++s) {
        const int j = (s << 6) + lane;
        const uint8_t qa = j < nA ? PA[j] : uint8_t(0);
        const uint8_t qb = j < nB ? PB[j] : uint8_t(0);
        const uint32_t tabA = u_table(qa, L.match, L.mismatch, m0);
        const uint32_t tabB = u_table(qb, L.match, L.mismatch, m0);
        const uint32_t colmask = (j < nA ? 0xffffu : 0u) | (j < nB ? 0xffff0000u : 0u);
        uint32_t H = 0, Eh = go2, Fh = go2, diag = 0;
        uint32_t bestH = 0, bestT = 0xffffffffu;
        for (int t0 = 0; t0 < nsteps; t0 += 4) {
            uint32_t acc = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u;
                const uint2 b = t < M ? bnd[t] : make_uint2(0u, go2);
                const uint32_t Hl = shr1(b.x, H);
                const uint32_t El = shr1(b.y, Eh);
                const uint32_t sv = sel[t - lane];
                const uint32_t uu = __builtin_amdgcn_perm(tabB, tabA, sv);
                const uint32_t xe = psubs(El, ge2);
                const uint32_t En = pmax(Hl, xe);
                const uint32_t xf = psubs(Fh, ge2);
                const uint32_t Fn = pmax(H, xf);
                const uint32_t A1 = psubs(padd(diag, uu), m02);
                const uint32_t G = pmax(En, Fn);
                const uint32_t A2 = psubs(G, go2);
                const uint32_t Hn = pmax(A1, A2);
                // live cells only: column in range, row in range (rows < 0 hold 0)
                uint32_t live = colmask;
                if (t >= M) {
                    const int i = t - lane;
                    live &= (i < mAp ? 0xffffu : 0u) | (i < mBp ? 0xffff0000u : 0u);
                } else if (t >= (mAp < mBp ? mAp : mBp)) {
                    const int i = t - lane;
                    live &= (i < mAp ? 0xffffu : 0u) | (i < mBp ? 0xffff0000u : 0u);
                }
                const uint32_t Hb = Hn & live;
                const uint32_t up = psign_mask(pdiff(bestH, Hb));
                bestH = pmax(bestH, Hb);
                bestT = (up & (uint32_t(t) * 0x10001u)) | (~up & bestT);
                if (TRACE) {
                    const uint32_t f_nd = pdiff(A1, A2);
                    const uint32_t f_fe = pdiff(En, Fn);
                    const uint32_t f_ee = pdiff(Hl, xe);
                    const uint32_t f_fx = pdiff(H, xf);
                    const uint32_t P1 = __builtin_amdgcn_perm(f_fe, f_nd, 0x07050301u);
                    const uint32_t P2 = __builtin_amdgcn_perm(f_fx, f_ee, 0x07050301u);
                    const uint32_t Z = (P1 & 0x80808080u) | ((P2 >> 4) & 0x08080808u);
                    acc = (u == 0) ? Z : ((acc >> 1) | Z);
                }
                if (lane == 63 && t >= 63 && t - 63 < M) bnd[t - 63] = make_uint2(Hn, En);
                diag = Hl;
                H = Hn;
                Eh = En;
                Fh = Fn;
            }
            if (TRACE) dirs[(int64_t(s) * nq + (t0 >> 2)) * 64 + lane] = acc;
        }
        // fold this strip's first maxima (per half) into the running records
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int bh = int((bestH >> (16 * h)) & 0xffffu);
            const int bt = int((bestT >> (16 * h)) & 0xffffu);
            const int bi = bh > 0 ? bt - lane : -1;
            if (bh > gH[h] || (bh == gH[h] && bh > 0 && bi < gI[h])) { gH[h] = bh; gI[h] = bi; gJ[h] = j; }
        }
        __syncthreads();
    }

    const int limit = 32767 - 255 - go;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int pid = h == 0 ? ia : ib;
        if (pid < 0) continue;
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
            const bool rerun = (h == 0 ? badA : badB) || lH >= limit;
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
                L.rerun[pid] = rerun ? 1 : 0;
            }
        }
    }
}

__device__ __forceinline__ uint32_t flag16(const uint32_t* dirs, int nq, int i, int j, int h, int f)
{
    const int j0 = j - 1, s = j0 >> 6, l = j0 & 63, t = (i - 1) + l;
    const uint32_t w = dirs[(int64_t(s) * nq + (t >> 2)) * 64 + l];
    const int byte = (f == 0 || f == 2) ? h : 2 + h;
    const int bit = (f < 2) ? 4 + (t & 3) : (t & 3);
    return (w >> (8 * byte + bit)) & 1u;
}

// One thread per pair: the reference's three-state walk driven by the 4 flags
// and the carried cell value.
__global__ __launch_bounds__(64) void sw16_walk_kernel(Sw16Launch L)
{
    const int q = int(blockIdx.x) * 64 + int(threadIdx.x);
    if (q >= 2 * L.n_waves) return;
    const int w = q >> 1, h = q & 1;
    int pid = L.pair_a ? (h ? L.pair_b[w] : L.pair_a[w]) : L.first_pair + q;
    if (pid >= L.n_total || pid < 0) return;
    if (L.rerun[pid]) return;
    int32_t* res = L.res + int64_t(pid) * 8;
    const int score = res[0];
    if (score <= 0 || (L.walk && !L.walk[pid])) return;
    const int m = L.tlen[pid], n = L.plen[pid];
    const uint8_t* T = L.txt + L.toff[pid];
    const uint8_t* P = L.pat + L.poff[pid];
    const int Mw = L.wave_m[w];
    const int nq = ((Mw + 63 + 3) & ~3) >> 2;
    const uint32_t* dirs = L.dirs + int64_t(w) * L.dir_stride;
    uint8_t* cg = L.cigar + L.cig_off[pid];
    int pos = m + n;
    int i = res[1] + 1, j = res[2] + 1, state = 0, v = score;
    while (i > 0 && j > 0 && pos > 0) {
        if (state == 0) {
            if (v == 0) break;
            if (!flag16(dirs, nq, i, j, h, 0)) {
                const bool eq = T[i - 1] == P[j - 1];
                cg[--pos] = eq ? 'M' : 'X';
                v -= eq ? L.match : L.mismatch;
                --i; --j;
            } else if (!flag16(dirs, nq, i, j, h, 1)) {
                cg[--pos] = 'D';
                const bool ext = flag16(dirs, nq, i, j, h, 2);
                state = ext ? 1 : 0;
                v += ext ? L.ge : L.go;
                --j;
            } else {
                cg[--pos] = 'I';
                const bool ext = flag16(dirs, nq, i, j, h, 3);
                state = ext ? 2 : 0;
                v += ext ? L.ge : L.go;
                --i;
            }
        } else if (state == 1) {
            cg[--pos] = 'D';
            const bool ext = flag16(dirs, nq, i, j, h, 2);
            if (!ext) state = 0;
            v += ext ? L.ge : L.go;
            --j;
        } else {
            cg[--pos] = 'I';
            const bool ext = flag16(dirs, nq, i, j, h, 3);
            if (!ext) state = 0;
            v += ext ? L.ge : L.go;
            --i;
        }
    }
    res[5] = i;
    res[6] = j;
    res[7] = m + n - pos;
}

}  // namespace

hipError_t sw16_launch(const Sw16Launch& L, bool trace, hipStream_t st, LaunchHooks* hooks)
{
    if (L.n_waves <= 0) return hipSuccess;
    const size_t lds = size_t(L.max_m + 136) * 4 + size_t(L.max_m) * 8;
    const dim3 grid(unsigned(L.n_waves)), block(64);
    if (hooks) hooks->before(st);
    if (trace) hipLaunchKernelGGL(sw16_kernel<true>, grid, block, lds, st, L);
    else       hipLaunchKernelGGL(sw16_kernel<false>, grid, block, lds, st, L);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

hipError_t sw16_walk_launch(const Sw16Launch& L, hipStream_t st)
{
    if (L.n_waves <= 0) return hipSuccess;
    hipLaunchKernelGGL(sw16_walk_kernel, dim3(unsigned((2 * L.n_waves + 63) / 64)), dim3(64), 0, st, L);
    return hipGetLastError();
}

}  // namespace ga
