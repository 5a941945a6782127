// msa.hip -- gfx950 kernels for the profile-profile Gotoh DP of the
// reference's progressive MSA (include/genoaligner/backend/
// msa_pp_kernel_impl.hip:35-209, semantics fixed by src/msa/msa_ref.cpp:
// 1063-1164):
//   M[i][j]  = col(A_i-1, B_j-1) + max(M, Ix, Iy)[i-1][j-1]   ties M > Ix > Iy
//   Ix[i][j] = max(M[i-1][j] - open(occA_i-1), Ix[i-1][j] - ext(occA_i-1))
//   Iy[i][j] = max(M[i][j-1] - open(occB_j-1), Iy[i][j-1] - ext(occB_j-1))
//   col(a, b) = sum over symbols x of a, then y of b, of fa_x * fb_y * S(x, y)
// in float, every operation in the reference's order and unfused (the
// library is built with -ffp-contract=off).  free_end_gaps: row 0 and column 0
// are 0 and the end is the best cell of the last column (rows ascending, M
// then Ix) or the last row (columns ascending, M then Iy), first maximum
// winning; otherwise the end is M[M][N].  A direction byte per cell (M source
// in bits 0-1, Ix extended 4, Iy extended 8) drives the reference's
// three-state walk.
//
// Mapping to CDNA4: one wave64 per pair.  Profile B's columns are cut into
// strips of 64, lane l owning column j = 64 s + l + 1; A's columns stream
// through the wave as rows with a one-step skew per lane (at step t lane l
// computes row t - l).  A cell's left neighbour (row i, column j-1) and its
// diagonal (row i-1, column j-1) are what lane l-1 produced one and two steps
// earlier, moved by one DPP wave_shr:1 per value and step; the row above stays
// in the lane's registers.  Lane 0 takes its left column from the strip
// boundary buffer that lane 63 of the previous strip filled (or, in strip 0,
// from the boundary formulas).  Profile columns come in sparse form (the
// non-zero symbol fractions in ascending symbol order, from msa_sparse_kernel):
// a zero fraction adds exactly +0 to a sum that is never -0, so skipping it
// leaves every bit of the reference's sum unchanged.  Direction bytes are
// stored in the computation's own skewed order, four steps per 32-bit word, a
// wave's 64 words one contiguous 256-byte store.
#include "ga_device.hpp"

#include <type_traits>

namespace ga {
namespace {

constexpr float kMsaNeg = -1e30f;

__device__ __forceinline__ float dpp_shr1(float old, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf,
                                                      false));
}

// Position-specific gap penalties (include/genoaligner/msa/msa.hpp:123-134).
__device__ __forceinline__ float psgp_open(float occ, const MsaParamsDev& P)
{
    if (!P.psgp || occ >= 1.0f) return P.go * occ;
    const float p = P.psgp_scale * P.go * occ;
    const float fl = P.psgp_min_open * P.go;
    return p > fl ? p : fl;
}
__device__ __forceinline__ float psgp_ext(float occ, const MsaParamsDev& P)
{
    if (!P.psgp || occ >= 1.0f) return P.ge;
    const float p = P.ge * occ;
    const float fl = P.psgp_min_ext * P.ge;
    return p > fl ? p : fl;
}

// One thread per profile column: its non-zero fractions in symbol order and
// its gap-open / gap-extend costs.
__global__ void msa_sparse_kernel(MsaSparseLaunch S)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= S.n_cols) return;
    const float* f = S.cols + c * S.P.alpha;
    int k = 0;
    for (int x = 0; x < S.P.alpha; ++x) {
        const float v = f[x];
        if (v != 0.f) {
            S.nz_sym[int64_t(k) * S.n_cols + c] = uint8_t(x);
            S.nz_frq[int64_t(k) * S.n_cols + c] = v;
            ++k;
        }
    }
    S.nz_cnt[c] = uint8_t(k);
    S.gopen[c] = psgp_open(S.occ[c], S.P);
    S.gext[c] = psgp_ext(S.occ[c], S.P);
}

struct Best {
    float v;
    int key;        // position in the reference's end-point scan order
};
__device__ __forceinline__ void offer(Best& b, float v, int key)
{
    if (v > b.v || (v == b.v && key < b.key)) { b.v = v; b.key = key; }
}

// The walk's window over the skewed direction words: one strip, 16 word
// rows (64 steps) of all 64 lanes, loaded by the whole wave when the walk
// leaves it.  The walk itself runs on every lane with wave-uniform state.
constexpr int kWinRows = 16;

struct WalkWin {
    uint32_t* w;         // kWinRows x 64 words (LDS)
    int s = -1, lo = 0;  // strip and first word row held
};

__device__ __forceinline__ uint32_t dir_at(const uint32_t* d, int tq, WalkWin& W, int lane, int i, int j)
{
    const int s = (j - 1) >> 6, l = (j - 1) & 63, t = i + l, wr = t >> 2;
    if (s != W.s || wr < W.lo || wr >= W.lo + kWinRows) {
        W.s = s;
        W.lo = wr - (kWinRows - 1) > 0 ? wr - (kWinRows - 1) : 0;
        __syncthreads();
        // every load issued before any is waited for (rows past the table
        // read its last row; the walk never reads them)
        const uint32_t* ds = d + int64_t(s) * tq * 64 + lane;
#pragma unroll
        for (int q0 = 0; q0 < kWinRows; q0 += 8) {
            uint32_t val[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int r = W.lo + q0 + q < tq ? W.lo + q0 + q : tq - 1;
                val[q] = ds[uint32_t(r) * 64u];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) W.w[(q0 + q) * 64 + lane] = val[q];
        }
        __syncthreads();
    }
    const uint32_t word = W.w[(wr - W.lo) * 64 + l];
    return (uint32_t(__builtin_amdgcn_readfirstlane(int(word))) >> (8 * (t & 3))) & 0xffu;
}

// Three-state walk from the end cell (msa_ref.cpp:1131-1156); emits the
// column ops right-aligned in the pair's CIGAR slot (lane 0 stores).
__device__ void msa_walk(const uint32_t* d, int tq, WalkWin& W, int lane, int ei, int ej, char est,
                         uint8_t* slot_end, int* oi, int* oj, int* used_out)
{
    int i = ei, j = ej, used = 0;
    char st = est;
    while (i > 0 || j > 0) {
        if (st == 'M') {
            if (i == 0 || j == 0) break;
            if (lane == 0) slot_end[-1 - used] = 'M';
            ++used;
            const uint32_t h = dir_at(d, tq, W, lane, i, j) & 3u;
            --i; --j;
            st = h == 0 ? 'M' : (h == 1 ? 'I' : 'D');
        } else if (st == 'I') {
            if (i == 0) break;
            if (lane == 0) slot_end[-1 - used] = 'I';
            ++used;
            if (j > 0) st = (dir_at(d, tq, W, lane, i, j) & 4u) ? 'I' : 'M';
            --i;
        } else {
            if (j == 0) break;
            if (lane == 0) slot_end[-1 - used] = 'D';
            ++used;
            if (i > 0) st = (dir_at(d, tq, W, lane, i, j) & 8u) ? 'D' : 'M';
            --j;
        }
    }
    *oi = i; *oj = j; *used_out = used;
}

__device__ __forceinline__ float boundary_gap(float open, int k, float ge)
{
    return -(open + float(k - 1) * ge);
}

// A pair with an empty profile: only boundary cells exist (row 0 / column 0).
__device__ void msa_degenerate(const MsaLaunch& L, int pid, int64_t ca, int M, int64_t cb, int N)
{
    const MsaParamsDev& P = L.P;
    const bool fr = P.free_end != 0;
    auto Mv = [&](int i, int j) -> float { return (i == 0 && j == 0) ? 0.f : (fr ? 0.f : kMsaNeg); };
    auto Ixv = [&](int i) -> float { return i == 0 ? kMsaNeg : (fr ? 0.f : boundary_gap(L.gopen[ca + i - 1], i, P.ge)); };
    auto Iyv = [&](int j) -> float { return j == 0 ? kMsaNeg : (fr ? 0.f : boundary_gap(L.gopen[cb + j - 1], j, P.ge)); };
    float best = Mv(M, N);
    int ei = M, ej = N;
    char est = 'M';
    if (fr) {
        best = kMsaNeg;
        // column N (only column 0 or row 0 cells exist here)
        for (int i = 0; i <= M; ++i) {
            const float vm = N == 0 ? Mv(i, 0) : Mv(0, N);
            const float vx = N == 0 ? Ixv(i) : kMsaNeg;
            if (N != 0 && i > 0) break;                  // M == 0 here: only row 0
            if (vm > best) { best = vm; ei = i; ej = N; est = 'M'; }
            if (vx > best) { best = vx; ei = i; ej = N; est = 'I'; }
        }
        for (int j = 0; j <= N; ++j) {
            const float vm = M == 0 ? Mv(0, j) : Mv(M, 0);
            const float vy = M == 0 ? Iyv(j) : kMsaNeg;
            if (M != 0 && j > 0) break;                  // N == 0 here: only column 0
            if (vm > best) { best = vm; ei = M; ej = j; est = 'M'; }
            if (vy > best) { best = vy; ei = M; ej = j; est = 'D'; }
        }
    }
    uint8_t* slot_end = L.cigar + L.cig_off[pid] + M + N;
    int i = ei, j = ej, used = 0;
    char st = est;
    while (i > 0 || j > 0) {                             // no interior cell: no direction byte is read
        if (st == 'M') break;
        if (st == 'I') {
            if (i == 0) break;
            slot_end[-1 - used++] = 'I';
            --i;
        } else {
            if (j == 0) break;
            slot_end[-1 - used++] = 'D';
            --j;
        }
    }
    L.score[pid] = best;
    int32_t* sp = L.span + 4 * int64_t(pid);
    sp[0] = i; sp[1] = j; sp[2] = ei; sp[3] = ej;
    L.cig_len[pid] = used;
}

// Rows of A (its columns, with their gap costs) and of the strip boundary
// column come through an LDS ring of 128 rows: at the start of each 64-step
// block the wave writes the rows fetched during the previous block into the
// ring and issues the global loads for the next, so no step waits on HBM.
// At step t of block t0 the lanes need A rows t - 63 .. t <= t0 + 64 and
// lane 0 the boundary row t: rows t0 - 63 .. t0 + 64 are in the ring.
constexpr int kRing = 128;

template <int K>
struct RowFetch {
    int cnt;
    float open, ext;
    uint32_t sym[(K + 3) / 4];
    float frq[K];
    float bm, bx, by;
};

template <int K>
__device__ __forceinline__ void fetch_rows(RowFetch<K>& f, const MsaLaunch& L, int64_t ca, int M, int t0, int lane,
                                           bool strip0, const float* bM, const float* bX, const float* bY)
{
    const int r = t0 + 1 + lane;                      // 1-based row of A
    const int64_t C = L.n_cols;
    f.cnt = 0; f.open = 0.f; f.ext = 0.f;
    for (int q = 0; q < (K + 3) / 4; ++q) f.sym[q] = 0;
    if (r <= M) {
        const int64_t c = ca + r - 1;
        f.cnt = L.nz_cnt[c];
        f.open = L.gopen[c];
        f.ext = L.gext[c];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            f.sym[k >> 2] |= uint32_t(L.nz_sym[int64_t(k) * C + c]) << (8 * (k & 3));
            f.frq[k] = L.nz_frq[int64_t(k) * C + c];
        }
    }
    const int rb = t0 + lane;                          // boundary row
    if (!strip0 && rb <= M) { f.bm = bM[rb]; f.bx = bX[rb]; f.by = bY[rb]; }
}

template <int K>
__global__ __launch_bounds__(64) void msa_pp_kernel(MsaLaunch L)
{
    constexpr bool kRegB = K <= 4;
    __shared__ float sub_s[K * K];
    __shared__ float bfrq[kRegB ? 1 : K * 64];
    __shared__ uint8_t bsym[kRegB ? 1 : K * 64];
    __shared__ float r_open[kRing], r_ext[kRing];
    __shared__ int r_cnt[kRing];
    __shared__ uint8_t r_sym[kRing * K];
    __shared__ float r_frq[kRing * K];
    __shared__ float r_bm[kRing], r_bx[kRing], r_by[kRing];
    __shared__ uint32_t win[kWinRows * 64];
    __shared__ float gsum[K * 64];
    __shared__ Best red[64];
    __shared__ float mmn;
    const int pid = L.first_pair + int(blockIdx.x);
    const int lane = threadIdx.x;
    const MsaParamsDev P = L.P;
    for (int q = lane; q < K * K; q += 64) sub_s[q] = L.sub[q];
    const int a = L.a_idx[pid], b = L.b_idx[pid];
    const int M = L.len[a], N = L.len[b];
    const int64_t ca = L.col_off[a], cb = L.col_off[b];
    __syncthreads();
    if (M <= 0 || N <= 0) {
        if (lane == 0) msa_degenerate(L, pid, ca, M > 0 ? M : 0, cb, N > 0 ? N : 0);
        return;
    }
    const bool fr = P.free_end != 0;
    const float ge = P.ge;
    const int64_t C = L.n_cols;
    uint32_t* dirs = L.dirs + L.dir_off[pid];
    float* bM = L.bnd + L.bnd_off[pid];
    float* bX = bM + (M + 1);
    float* bY = bX + (M + 1);
    const int nstrips = (N + 63) >> 6;
    const int T = M + 64;                               // steps t = 0 .. M + 63
    const int tq = (T + 3) >> 2;
    Best best{-3.0e38f, 0x7fffffff};
    if (lane == 0 && fr) offer(best, 0.f, 2 * (M + 1));   // row M, column 0: M[M][0]
    for (int s = 0; s < nstrips; ++s) {
        const int j = (s << 6) + lane + 1;
        const bool jv = j <= N;
        const int64_t cB = cb + j - 1;
        const int nb = jv ? int(L.nz_cnt[cB]) : 0;
        const float openA = jv ? L.gopen[cB] : 0.f, extA = jv ? L.gext[cB] : 0.f;
        float prevM = fr ? 0.f : kMsaNeg, prevIx = kMsaNeg;          // M[0][j], Ix[0][j]
        const float row0Iy = fr ? 0.f : boundary_gap(openA, j, ge);
        float outM = kMsaNeg, outIx = kMsaNeg, outIy = kMsaNeg;
        float dM = kMsaNeg, dIx = kMsaNeg, dIy = kMsaNeg;
        uint32_t acc = 0;
        const bool full_strip = (s << 6) + 64 <= N;
        const bool coln = (s << 6) + 64 == N;               // lane 63 holds column N
        // the lane's column of B: in registers for DNA (at most 4 symbols),
        // in the workgroup's LDS for larger alphabets
        float fbv[kRegB ? K : 1];
        int ybv[kRegB ? K : 1];
        if constexpr (kRegB) {
#pragma unroll
            for (int kb = 0; kb < K; ++kb) {
                fbv[kb] = kb < nb ? L.nz_frq[int64_t(kb) * C + cB] : 0.f;
                ybv[kb] = kb < nb ? int(L.nz_sym[int64_t(kb) * C + cB]) : 0;
            }
        } else {
            for (int k = 0; k < nb; ++k) {
                bsym[k * 64 + lane] = L.nz_sym[int64_t(k) * C + cB];
                bfrq[k * 64 + lane] = L.nz_frq[int64_t(k) * C + cB];
            }
        }
        // acc += fa * fb * S[x][y] over the lane's B symbols, in the
        // reference's term order
        auto add_col = [&](float& acc, int x, float fa) {
            if constexpr (kRegB) {
#pragma unroll
                for (int kb = 0; kb < K; ++kb)
                    if (kb < nb) acc += fa * fbv[kb] * sub_s[x * K + ybv[kb]];
            } else {
                for (int kb = 0; kb < nb; ++kb)
                    acc += fa * bfrq[kb * 64 + lane] * sub_s[x * K + bsym[kb * 64 + lane]];
            }
        };
        // G[x] = the substitution sum of a one-symbol A column x (fraction
        // 1) against the lane's column of B
        for (int x = 0; x < K; ++x) {
            float g = 0.f;
            add_col(g, x, 1.0f);
            gsum[x * 64 + lane] = g;
        }
        RowFetch<K> pf;
        fetch_rows<K>(pf, L, ca, M, 0, lane, s == 0, bM, bX, bY);
        for (int t0 = 0; t0 < T; t0 += 64) {
            __syncthreads();                             // the previous block's ring reads are done
            {
                const int r = t0 + 1 + lane, q = r & (kRing - 1);
                r_cnt[q] = pf.cnt; r_open[q] = pf.open; r_ext[q] = pf.ext;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    r_sym[k * kRing + q] = uint8_t(pf.sym[k >> 2] >> (8 * (k & 3)));
                    r_frq[k * kRing + q] = pf.frq[k];
                }
                const int qb = (t0 + lane) & (kRing - 1);
                r_bm[qb] = pf.bm; r_bx[qb] = pf.bx; r_by[qb] = pf.by;
            }
            __syncthreads();
            if (t0 + 64 < T) fetch_rows<K>(pf, L, ca, M, t0 + 64, lane, s == 0, bM, bX, bY);
            const int tend = t0 + 64 < T ? t0 + 64 : T;
            // One step.  FULL: every lane's row exists and is below the
            // last row (63 < t0, t0 + 63 < M) and every lane's column
            // exists, none of them column N (no activity tests, no end-point
            // bookkeeping).  S0: strip 0 (the left column is the boundary).
            auto step = [&](int t, auto full_tag, auto s0_tag) {
                constexpr bool FULL = decltype(full_tag)::value;
                constexpr bool S0 = decltype(s0_tag)::value;
                const int i = t - lane;
                // lane 0's left column at row t: the column-0 boundary or
                // the previous strip's last column (wave-uniform values)
                float lbM = kMsaNeg, lbX = kMsaNeg, lbY = kMsaNeg;
                if (FULL || t <= M) {
                    const int qt = t & (kRing - 1);
                    if (S0) {
                        lbM = (!FULL && t == 0) ? 0.f : (fr ? 0.f : kMsaNeg);
                        lbX = (!FULL && t == 0) ? kMsaNeg : (fr ? 0.f : boundary_gap(r_open[qt], t, ge));
                    } else {
                        lbM = r_bm[qt]; lbX = r_bx[qt]; lbY = r_by[qt];
                    }
                }
                const float lM = dpp_shr1(lbM, outM);
                const float lX = dpp_shr1(lbX, outIx);
                const float lY = dpp_shr1(lbY, outIy);
                uint32_t code = 0;
                if (!FULL && i == 0) {
                    outM = prevM; outIx = prevIx; outIy = row0Iy;
                    if (jv && j == N) {                  // column N, row 0
                        offer(best, prevM, 0);
                        offer(best, prevIx, 1);
                    }
                } else if (FULL || (i >= 1 && i <= M && jv)) {
                    const int qi = i & (kRing - 1);
                    const float openB = r_open[qi], extB = r_ext[qi];
                    float sc = 0.f;
                    const int na = r_cnt[qi];
                    const int x0 = r_sym[qi];
                    const float f0 = r_frq[qi];
                    // A column of one symbol at fraction 1: fa * fb == fb
                    // exactly, so the sum is the lane's precomputed G[x]
                    if (__all(na == 1 && f0 == 1.0f)) {
                        sc = gsum[x0 * 64 + lane];
                    } else {
                        for (int ka = 0; ka < na; ++ka)
                            add_col(sc, r_sym[ka * kRing + qi], r_frq[ka * kRing + qi]);
                    }
                    const float oIx = prevM - openB, eIx = prevIx - extB;
                    const float nIx = oIx >= eIx ? oIx : eIx;
                    const float oIy = lM - openA, eIy = lY - extA;
                    const float cIy = oIy >= eIy ? oIy : eIy;
                    float mx = dM;
                    uint32_t h = 0;
                    if (dIx > mx) { mx = dIx; h = 1; }
                    if (dIy > mx) { mx = dIy; h = 2; }
                    const float cm = sc + mx;
                    code = h | (eIx > oIx ? 4u : 0u) | (eIy > oIy ? 8u : 0u);
                    outM = cm; outIx = nIx; outIy = cIy;
                    prevM = cm; prevIx = nIx;
                    if (!FULL) {
                        if (fr) {
                            if (j == N) { offer(best, cm, 2 * i); offer(best, nIx, 2 * i + 1); }
                            if (i == M) { offer(best, cm, 2 * (M + 1) + 2 * j); offer(best, cIy, 2 * (M + 1) + 2 * j + 1); }
                        } else if (i == M && j == N) {
                            mmn = cm;
                        }
                    }
                }
                dM = lM; dIx = lX; dIy = lY;
                if (lane == 63 && (FULL || (i >= 0 && i <= M))) { bM[i] = outM; bX[i] = outIx; bY[i] = outIy; }
                acc |= code << (8 * (t & 3));
                if (FULL ? (t & 3) == 3 : ((t & 3) == 3 || t == T - 1)) {
                    dirs[(int64_t(s) * tq + (t >> 2)) * 64 + lane] = acc;
                    acc = 0;
                }
            };
            using TT = std::true_type;
            using FF = std::false_type;
            if (full_strip && !coln && t0 >= 64 && t0 + 63 < M) {
                auto run4 = [&](auto s0_tag) {
                    for (int t = t0; t < tend; t += 4) {
                        step(t, TT{}, s0_tag);
                        step(t + 1, TT{}, s0_tag);
                        step(t + 2, TT{}, s0_tag);
                        step(t + 3, TT{}, s0_tag);
                    }
                };
                if (s == 0) run4(TT{}); else run4(FF{});
            } else if (s == 0) {
                for (int t = t0; t < tend; ++t) step(t, FF{}, TT{});
            } else {
                for (int t = t0; t < tend; ++t) step(t, FF{}, FF{});
            }
        }
        __syncthreads();
    }
    // first maximum of the end-point scan: the greatest value, then the
    // earliest position in the scan order
    red[lane] = best;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) {
        if (lane < off) {
            const Best o = red[lane + off];
            Best m = red[lane];
            offer(m, o.v, o.key);
            red[lane] = m;
        }
        __syncthreads();
    }
    __threadfence();
    int ei = M, ej = N;
    char est = 'M';
    float sc;
    if (fr) {
        const Best w = red[0];
        sc = w.v;
        if (w.key < 2 * (M + 1)) {
            ei = w.key >> 1; ej = N; est = (w.key & 1) ? 'I' : 'M';
        } else {
            const int q = w.key - 2 * (M + 1);
            ei = M; ej = q >> 1; est = (q & 1) ? 'D' : 'M';
        }
    } else {
        sc = mmn;
    }
    WalkWin W;
    W.w = win;
    int oi, oj, used;
    msa_walk(dirs, tq, W, lane, ei, ej, est, L.cigar + L.cig_off[pid] + M + N, &oi, &oj, &used);
    if (lane == 0) {
        L.score[pid] = sc;
        int32_t* sp = L.span + 4 * int64_t(pid);
        sp[0] = oi; sp[1] = oj; sp[2] = ei; sp[3] = ej;
        L.cig_len[pid] = used;
    }
}

}  // namespace

int64_t msa_dir_words(int M, int N)
{
    if (M <= 0 || N <= 0) return 0;
    return int64_t((N + 63) >> 6) * int64_t((M + 64 + 3) >> 2) * 64;
}

hipError_t msa_sparse_launch(const MsaSparseLaunch& S, hipStream_t st)
{
    if (S.n_cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(msa_sparse_kernel, dim3(unsigned((S.n_cols + 255) / 256)), dim3(256), 0, st, S);
    return hipGetLastError();
}

hipError_t msa_pp_launch(const MsaLaunch& L, int n_pairs, hipStream_t st, LaunchHooks* hooks)
{
    if (n_pairs <= 0) return hipSuccess;
    if (hooks) hooks->before(st);
    if (L.P.alpha == 4) hipLaunchKernelGGL(msa_pp_kernel<4>, dim3(unsigned(n_pairs)), dim3(64), 0, st, L);
    else hipLaunchKernelGGL(msa_pp_kernel<20>, dim3(unsigned(n_pairs)), dim3(64), 0, st, L);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

}  // namespace ga
