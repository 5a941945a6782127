// fasta.hip -- FASTA ingest on the device: a whole file in HBM becomes its
// records' sequences back to back (the byte layout the batch kernels read),
// with the host reader's rules (fasta.cpp; reference include/genoaligner/io/
// fasta.hpp:50-175).
//
// Every byte is classified from its line: a line starts at byte 0, after a
// \n, and after a \r not followed by \n; a line whose first byte is '>' is a
// header; \n and \r are never sequence.  A byte is kept iff its line starts
// at or after the first header and is not a header, and the byte is not a
// line break.  So a byte's fate depends only on the line start in force and
// whether that line is a header: the state carried from tile to tile is one
// number, (line start << 1) | header, combined by max.
//
// A tile is 512 threads x 32 bytes (16 KiB).  Two passes over the file's
// tiles and one small scan between them:
//   A  per tile: the last line state, the header count, the first header, the
//      error candidates (a header with an empty id; a line before the first
//      header starting with a byte other than '>', space, tab or a line
//      break), and the kept bytes under either entry state: bytes before the
//      tile's first line start (kept iff the line in force is a sequence
//      line), bytes after it (decided inside the tile), and bytes after the
//      tile's first header (for the tile holding the file's first header);
//   S  three small launches over groups of 1024 tiles: the file's first
//      header and error candidates, the line-state carry into each tile,
//      headers and kept bytes before each tile, the record and byte totals,
//      and whether the file is well formed (the kernels below do nothing
//      otherwise);
//   E  per tile: scatter of the kept bytes, each record's start, a clear of
//      the record's A/C/G/T flag at any other kept byte, and (optional) the
//      2-bit codes of the 16-byte output words inside the tile's output;
//   F  per record: length and the empty-record check;
//   G  (optional) the 2-bit words cut by tile seams.
// Each thread reads its 32 bytes with two 128-bit loads (when the file is
// 16-byte aligned) plus the byte before and the byte after them.  The host
// waits once, at the end; a malformed file is only flagged on the device and
// the host reader words the message.
#include "../../include/genoaligner/genoaligner_c.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

bool ga_fasta_host_message(const char* data, int64_t len, std::string& err);

namespace {

constexpr int kFaThreads = 512;
constexpr int kFaWaves = kFaThreads / 64;
constexpr int kFaBytes = 32;                          // per thread
constexpr int64_t kFaTile = int64_t(kFaThreads) * kFaBytes;
constexpr int kScanThreads = 1024;

struct FaInfo {
    int64_t first_hdr;      // first header line start (len if none)
    int64_t n_records;
    int64_t total;          // kept bytes
    int64_t bad_pre;        // first offending line start (len if none)
    int64_t bad_id;         // first header with an empty id (len if none)
    int32_t bad_rec;        // 1: some record has no sequence byte; 2: one is longer than INT32_MAX
    int32_t ok;             // well formed so far, and n_records <= max_records
};

// Per-tile values of pass A (ntiles each); the scan rewrites three of them.
struct FaTiles {
    int64_t* state;         // last (line start << 1 | header) in the tile, -1 if none -> carry into the tile
    int64_t* nhdr;          // headers -> headers before the tile
    int64_t* first_gt;      // first header in the tile (len if none)
    int64_t* bad_pre;
    int64_t* bad_id;
    int64_t* kpre;          // non-break bytes before the tile's first line start
    int64_t* kloc;          // kept bytes after it -> kept bytes before the tile
    int64_t* kfh;           // kept bytes after the tile's first header
};

__device__ __forceinline__ bool fa_base(uint8_t c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

// One thread's 32 bytes, and its neighbours: prev = the byte before (a line
// break at the file start, so byte 0 starts a line), next = the byte after
// ('\n' past the end).  Bytes past the end read 0 and are never used.
struct FaBytes {
    uint8_t b[kFaBytes];
    uint8_t prev, next;
};
template <bool AL>
__device__ __forceinline__ FaBytes fa_load(const uint8_t* d, int64_t len, int64_t base)
{
    FaBytes r;
    if (AL && base + kFaBytes <= len) {
        const uint4 v0 = *reinterpret_cast<const uint4*>(d + base);
        const uint4 v1 = *reinterpret_cast<const uint4*>(d + base + 16);
        const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
        for (int q = 0; q < kFaBytes; ++q) r.b[q] = uint8_t(w[q >> 2] >> (8 * (q & 3)));
    } else {
#pragma unroll
        for (int q = 0; q < kFaBytes; ++q) r.b[q] = base + q < len ? d[base + q] : uint8_t(0);
    }
    r.prev = base > 0 && base - 1 < len ? d[base - 1] : uint8_t('\n');
    r.next = base + kFaBytes < len ? d[base + kFaBytes] : uint8_t('\n');
    return r;
}
// Bit k of the result: byte k of x equals the byte in c4 (c * 0x01010101).
// Exact per byte (no borrow between bytes): the high bit of each byte of
// ((t & 0x7f..) + 0x7f..) | t is set iff that byte of t is nonzero.
__device__ __forceinline__ uint32_t fa_eq4(uint32_t x, uint32_t c4)
{
    const uint32_t t = x ^ c4;
    const uint32_t z = ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t) & 0x80808080u;
    return ((z >> 7) * 0x00204081u) >> 21 & 0xfu;     // bits 0, 8, 16, 24 -> 0..3
}
// 2-bit codes of the 4 bytes of w as 8 bits, byte k at bits 2k: A 0, C 1,
// G 2, T 3 (((c >> 1) ^ (c >> 2)) & 3 of the ASCII byte), any other byte 0.
__device__ __forceinline__ uint32_t fa_code8(uint32_t w)
{
    const uint32_t m4 = fa_eq4(w, 0x41414141u) | fa_eq4(w, 0x43434343u) | fa_eq4(w, 0x47474747u) |
                        fa_eq4(w, 0x54545454u);
    const uint32_t keep = ((m4 * 0x00204081u) & 0x01010101u) * 0xffu;    // bit k -> byte k
    uint32_t c = ((w >> 1) ^ (w >> 2)) & 0x03030303u & keep;
    c = (c | (c >> 6)) & 0x000f000fu;
    return (c | (c >> 12)) & 0xffu;
}

__device__ __forceinline__ uint32_t fa_mask(const uint32_t (&w)[kFaBytes / 4], uint8_t c)
{
    const uint32_t c4 = uint32_t(c) * 0x01010101u;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < kFaBytes / 4; ++k) m |= fa_eq4(w[k], c4) << (4 * k);
    return m;
}
__device__ __forceinline__ uint32_t fa_from(int q) { return q >= 32 ? 0u : ~0u << q; }     // bits q..31
__device__ __forceinline__ uint32_t fa_below(int q) { return q >= 32 ? ~0u : (1u << q) - 1u; }  // bits 0..q-1

__device__ __forceinline__ uint32_t fa_mask_bytes(const FaBytes& r, uint8_t c)
{
    uint32_t w[kFaBytes / 4];
#pragma unroll
    for (int k = 0; k < kFaBytes / 4; ++k)
        w[k] = uint32_t(r.b[4 * k]) | uint32_t(r.b[4 * k + 1]) << 8 | uint32_t(r.b[4 * k + 2]) << 16 |
               uint32_t(r.b[4 * k + 3]) << 24;
    return fa_mask(w, c);
}

// The classes of one thread's 32 bytes as bit masks (bit q = byte base + q).
struct FaMasks {
    uint32_t valid;         // inside the file
    uint32_t eol;           // \n or \r
    uint32_t ls;            // starts a line
    uint32_t hs;            // starts a header line
    uint32_t hl;            // inside a header line that starts in these bytes
    uint32_t other;         // not ' ', '\t', '\n', '\r' or '>'
    uint32_t blank;         // ' ', '\t', '\n' or '\r'
};
template <bool AL>
__device__ __forceinline__ FaMasks fa_masks(const uint8_t* d, int64_t len, int64_t base, FaBytes& r)
{
    r = fa_load<AL>(d, len, base);
    uint32_t w[kFaBytes / 4];
#pragma unroll
    for (int k = 0; k < kFaBytes / 4; ++k)
        w[k] = uint32_t(r.b[4 * k]) | uint32_t(r.b[4 * k + 1]) << 8 | uint32_t(r.b[4 * k + 2]) << 16 |
               uint32_t(r.b[4 * k + 3]) << 24;
    FaMasks m;
    const int64_t nv = len - base;
    m.valid = nv >= kFaBytes ? ~0u : nv <= 0 ? 0u : fa_below(int(nv));
    const uint32_t nl = fa_mask(w, '\n'), cr = fa_mask(w, '\r'), gt = fa_mask(w, '>');
    const uint32_t sp = fa_mask(w, ' '), tab = fa_mask(w, '\t');
    m.eol = (nl | cr) & m.valid;
    // byte q starts a line iff byte q-1 is \n, or is \r and byte q is not \n
    const uint32_t pnl = nl << 1 | (r.prev == '\n' ? 1u : 0u);
    const uint32_t pcr = cr << 1 | (r.prev == '\r' ? 1u : 0u);
    m.ls = (pnl | (pcr & ~nl)) & m.valid;
    m.hs = m.ls & gt;
    m.other = ~(nl | cr | gt | sp | tab) & m.valid;
    m.blank = (nl | cr | sp | tab) & m.valid;
    m.hl = 0;
    for (uint32_t rem = m.hs; rem;) {                 // each header line runs to the next line start
        const int s = __builtin_ctz(rem);
        rem &= rem - 1;
        const uint32_t after = m.ls & fa_from(s + 1);
        const int e = after ? __builtin_ctz(after) : 32;
        m.hl |= fa_from(s) & fa_below(e);
    }
    return m;
}

struct OpSum { template <class V> __device__ V operator()(V a, V b) const { return a + b; } };
struct OpMax { template <class V> __device__ V operator()(V a, V b) const { return a > b ? a : b; } };
struct OpMin { template <class V> __device__ V operator()(V a, V b) const { return a < b ? a : b; } };

template <class V, class Op>
__device__ __forceinline__ V wave_reduce(V v, Op op)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = op(v, V(__shfl_xor(v, off, 64)));
    return v;
}
template <class V, class Op>
__device__ __forceinline__ V wave_inscan(V v, Op op)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const V x = __shfl_up(v, off, 64);
        if (lane >= off) v = op(v, x);
    }
    return v;
}

// Exclusive scan over the workgroup (at most 16 waves); *all gets the
// reduction.  wtmp: 16 entries of LDS owned by this call site.
template <class V> struct FaSame { using type = V; };   // keeps `all` out of deduction (nullptr)
template <class V, class Op>
__device__ V block_exscan(V v, V id, Op op, V* wtmp, typename FaSame<V>::type* all)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const V inc = wave_inscan(v, op);
    if (lane == 63) wtmp[w] = inc;
    __syncthreads();
    V pre = id, tot = id;
    for (int k = 0; k < nw; ++k) {
        if (k < w) pre = op(pre, wtmp[k]);
        tot = op(tot, wtmp[k]);
    }
    V ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = id;
    if (all) *all = tot;
    __syncthreads();
    return op(pre, ex);
}
template <class V, class Op>
__device__ V block_reduce(V v, V id, Op op, V* wtmp)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_reduce(v, op);
    if (lane == 0) wtmp[w] = v;
    __syncthreads();
    V tot = id;
    for (int k = 0; k < nw; ++k) tot = op(tot, wtmp[k]);
    __syncthreads();
    return tot;
}

// A: per-tile values.  Positions inside the thread block are tile-relative
// (< 16 KiB), so every reduction runs on 32-bit values, and the four counts
// travel as two 16-bit pairs (each at most 16,384).
template <bool AL>
__global__ __launch_bounds__(kFaThreads) void fa_tiles_a(const uint8_t* d, int64_t len, FaTiles T)
{
    __shared__ int32_t wt[2][16];
    __shared__ int32_t part[4][kFaWaves];
    const int64_t tile0 = int64_t(blockIdx.x) * kFaTile;
    const int rel = int(threadIdx.x) * kFaBytes;      // this thread's first byte, from the tile start
    const int64_t base = tile0 + rel;
    FaBytes r;
    const FaMasks m = fa_masks<AL>(d, len, base, r);
    const int fls = m.ls ? __builtin_ctz(m.ls) : 32;
    const int fhs = m.hs ? __builtin_ctz(m.hs) : 32;
    const uint32_t seqb = m.valid & ~m.eol;           // bytes that are sequence if their line is
    const int kpre = __builtin_popcount(seqb & fa_below(fls));
    const int kloc = __builtin_popcount(seqb & fa_from(fls) & ~m.hl);
    const int kfh = __builtin_popcount(seqb & fa_from(fhs) & ~m.hl);
    int32_t last = -1;                                // (tile-relative line start << 1) | header
    if (m.ls) {
        const int q = 31 - __builtin_clz(m.ls);
        last = 2 * (rel + q) + int((m.hs >> q) & 1);
    }
    const int32_t first_gt = m.hs ? rel + fhs : int32_t(kFaTile);
    const uint32_t bp = m.ls & m.other;               // a line that must not come before the first header
    const int32_t pre = bp ? rel + __builtin_ctz(bp) : int32_t(kFaTile);
    // a header whose next byte is a blank, a line break or the end of the file has an empty id
    const bool next_blank = base + kFaBytes >= len || r.next == ' ' || r.next == '\t' || r.next == '\n' ||
                            r.next == '\r';
    const uint32_t blank_next = ((m.blank | ~m.valid) >> 1) | (next_blank ? 0x80000000u : 0u);
    const uint32_t bi_m = m.hs & blank_next;
    const int32_t bid = bi_m ? rel + __builtin_ctz(bi_m) : int32_t(kFaTile);
    // the line in force at this thread's first byte, from the tile's earlier threads (-1: none)
    int32_t tile_last;
    const int32_t e = block_exscan(last, int32_t(-1), OpMax(), wt[0], &tile_last);
    const int32_t g = block_reduce(first_gt, int32_t(kFaTile), OpMin(), wt[1]);
    const bool e_seq = e >= 0 && !(e & 1);
    const int32_t c_pre = e < 0 ? kpre : 0;
    const int32_t c_loc = kloc + (e_seq ? kpre : 0);
    const int32_t c_fh = (g < kFaTile && first_gt == g) ? kfh : (g < rel ? c_loc : 0);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t r_a = wave_reduce(int32_t(__builtin_popcount(m.hs)) | c_fh << 16, OpSum());
    const int32_t r_b = wave_reduce(c_pre | c_loc << 16, OpSum());
    const int32_t r_pre = wave_reduce(pre, OpMin());
    const int32_t r_bid = wave_reduce(bid, OpMin());
    if (lane == 0) {
        part[0][w] = r_a;
        part[1][w] = r_b;
        part[2][w] = r_pre;
        part[3][w] = r_bid;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s_nh = 0, s_cfh = 0, s_cpre = 0, s_cloc = 0;
        int32_t s_pre = kFaTile, s_bid = kFaTile;
        for (int k = 0; k < kFaWaves; ++k) {
            s_nh += part[0][k] & 0xffff;
            s_cfh += part[0][k] >> 16;
            s_cpre += part[1][k] & 0xffff;
            s_cloc += part[1][k] >> 16;
            s_pre = part[2][k] < s_pre ? part[2][k] : s_pre;
            s_bid = part[3][k] < s_bid ? part[3][k] : s_bid;
        }
        const int64_t t = blockIdx.x;
        T.state[t] = tile_last < 0 ? int64_t(-1) : 2 * tile0 + tile_last;
        T.nhdr[t] = s_nh;
        T.first_gt[t] = g < kFaTile ? tile0 + g : len;
        T.bad_pre[t] = s_pre < kFaTile ? tile0 + s_pre : len;
        T.bad_id[t] = s_bid < kFaTile ? tile0 + s_bid : len;
        T.kpre[t] = s_cpre;
        T.kloc[t] = s_cloc;
        T.kfh[t] = s_cfh;
    }
}

// S, in three launches over groups of kScanThreads tiles (one thread a tile):
//   S1  per group: first header, error candidates, last line state, headers;
//   S2  per group: the carries into the group from the earlier groups'
//       aggregates, then per tile the line-state carry (T.state), headers
//       before it (T.nhdr) and its kept bytes under its entry state (into
//       T.kloc); the group's kept sum; group 0 also fills *info but for the
//       totals;
//   S3  per group: kept bytes before each tile (T.kloc); group 0 fills the
//       byte total and the well-formed flag.
// A group reads at most all groups' aggregates (a few hundred for a GB file).
struct FaGroups {
    int64_t* first_gt;
    int64_t* bad_pre;
    int64_t* bad_id;
    int64_t* state;
    int64_t* nhdr;
    int64_t* kept;
};

__global__ __launch_bounds__(kScanThreads) void fa_scan_up(FaTiles T, int64_t n, int64_t len, FaGroups G)
{
    __shared__ int64_t wt[5][16];
    const int64_t k = int64_t(blockIdx.x) * kScanThreads + threadIdx.x;
    const bool in = k < n;
    const int64_t fg = block_reduce(in ? T.first_gt[k] : len, len, OpMin(), wt[0]);
    const int64_t bp = block_reduce(in ? T.bad_pre[k] : len, len, OpMin(), wt[1]);
    const int64_t bi = block_reduce(in ? T.bad_id[k] : len, len, OpMin(), wt[2]);
    const int64_t st = block_reduce(in ? T.state[k] : int64_t(-1), int64_t(-1), OpMax(), wt[3]);
    const int64_t nh = block_reduce(in ? T.nhdr[k] : int64_t(0), int64_t(0), OpSum(), wt[4]);
    if (threadIdx.x == 0) {
        G.first_gt[blockIdx.x] = fg;
        G.bad_pre[blockIdx.x] = bp;
        G.bad_id[blockIdx.x] = bi;
        G.state[blockIdx.x] = st;
        G.nhdr[blockIdx.x] = nh;
    }
}

__global__ __launch_bounds__(kScanThreads) void fa_scan_mid(FaTiles T, int64_t n, int64_t len, FaGroups G,
                                                            int64_t ngroups, FaInfo* info)
{
    __shared__ int64_t wt[8][16];
    const int64_t gi = blockIdx.x;
    // aggregates of all groups (first header, errors, records) and of the earlier ones (carries)
    int64_t fg = len, bp = len, bi = len, nrec = 0, cst = -1, chd = 0;
    for (int64_t q = threadIdx.x; q < ngroups; q += kScanThreads) {
        fg = G.first_gt[q] < fg ? G.first_gt[q] : fg;
        bp = G.bad_pre[q] < bp ? G.bad_pre[q] : bp;
        bi = G.bad_id[q] < bi ? G.bad_id[q] : bi;
        nrec += G.nhdr[q];
        if (q < gi) {
            cst = G.state[q] > cst ? G.state[q] : cst;
            chd += G.nhdr[q];
        }
    }
    const int64_t first_hdr = block_reduce(fg, len, OpMin(), wt[0]);
    const int64_t bad_pre = block_reduce(bp, len, OpMin(), wt[1]);
    const int64_t bad_id = block_reduce(bi, len, OpMin(), wt[2]);
    const int64_t n_rec = block_reduce(nrec, int64_t(0), OpSum(), wt[3]);
    const int64_t g_st = block_reduce(cst, int64_t(-1), OpMax(), wt[4]);
    const int64_t g_hd = block_reduce(chd, int64_t(0), OpSum(), wt[5]);
    const int64_t k = gi * kScanThreads + threadIdx.x;
    const bool in = k < n;
    const int64_t s = in ? T.state[k] : int64_t(-1), h = in ? T.nhdr[k] : int64_t(0);
    int64_t e_st = block_exscan(s, int64_t(-1), OpMax(), wt[6], nullptr);
    e_st = e_st > g_st ? e_st : g_st;
    const int64_t e_hd = g_hd + block_exscan(h, int64_t(0), OpSum(), wt[7], nullptr);
    const int64_t ft = first_hdr < len ? first_hdr / kFaTile : n;
    int64_t kk = 0;
    if (in) {
        const bool seq_entry = e_st >= 0 && !(e_st & 1);
        kk = k < ft ? 0 : k == ft ? T.kfh[k] : T.kloc[k] + (seq_entry ? T.kpre[k] : 0);
        T.state[k] = e_st;
        T.nhdr[k] = e_hd;
        T.kloc[k] = kk;
    }
    const int64_t ks = block_reduce(kk, int64_t(0), OpSum(), wt[0]);
    if (threadIdx.x == 0) {
        G.kept[gi] = ks;
        if (gi == 0) {
            info->first_hdr = first_hdr;
            info->n_records = n_rec;
            info->bad_pre = bad_pre;
            info->bad_id = bad_id;
            info->bad_rec = 0;
        }
    }
}

__global__ __launch_bounds__(kScanThreads) void fa_scan_down(FaTiles T, int64_t n, int64_t len, FaGroups G,
                                                             int64_t ngroups, int64_t max_records, FaInfo* info)
{
    __shared__ int64_t wt[3][16];
    const int64_t gi = blockIdx.x;
    int64_t before = 0, all = 0;
    for (int64_t q = threadIdx.x; q < ngroups; q += kScanThreads) {
        all += G.kept[q];
        before += q < gi ? G.kept[q] : 0;
    }
    before = block_reduce(before, int64_t(0), OpSum(), wt[0]);
    all = block_reduce(all, int64_t(0), OpSum(), wt[1]);
    const int64_t k = gi * kScanThreads + threadIdx.x;
    const bool in = k < n;
    const int64_t ex = before + block_exscan(in ? T.kloc[k] : int64_t(0), int64_t(0), OpSum(), wt[2], nullptr);
    if (in) T.kloc[k] = ex;
    if (gi == 0 && threadIdx.x == 0) {
        const int64_t n_rec = info->n_records;
        info->total = all;
        info->ok = n_rec > 0 && info->bad_pre >= info->first_hdr && info->bad_id >= len && n_rec <= max_records
                       ? 1 : 0;
    }
}

__global__ __launch_bounds__(kFaThreads) void fa_flags_init(uint8_t* acgt, int64_t cap, const FaInfo* info)
{
    const int64_t k = int64_t(blockIdx.x) * kFaThreads + threadIdx.x;
    if (info->ok && k < info->n_records && k < cap) acgt[k] = 1;
}

// E: scatter, record starts, and the A/C/G/T flags (acgt[] holds 1 for every
// record on entry; a kept byte other than A, C, G, T clears its record's).
// The tile's kept bytes are gathered in LDS at (output address mod 16) + their
// offset in the tile, so that 16-byte-aligned output runs are 16-byte-aligned
// in LDS, then written with 128-bit stores (bytes at the two ends).
template <bool AL>
__global__ __launch_bounds__(kFaThreads) void fa_tiles_e(const uint8_t* d, int64_t len, FaTiles T,
                                                        const FaInfo* info, uint8_t* seq, int64_t* seq_off,
                                                        uint8_t* acgt, uint32_t* pack, int64_t max_records)
{
    __shared__ int32_t wt[2][16];
    __shared__ __align__(16) uint8_t stage[kFaTile + 16];
    if (!info->ok) return;                            // the same for the whole workgroup
    const int64_t first_hdr = info->first_hdr;
    const int64_t tile0 = int64_t(blockIdx.x) * kFaTile;
    const int rel = int(threadIdx.x) * kFaBytes;
    const int64_t base = tile0 + rel;
    FaBytes r;
    const FaMasks m = fa_masks<AL>(d, len, base, r);
    int32_t last = -1;                                // (tile-relative line start << 1) | header
    if (m.ls) {
        const int q = 31 - __builtin_clz(m.ls);
        last = 2 * (rel + q) + int((m.hs >> q) & 1);
    }
    const int32_t ex = block_exscan(last, int32_t(-1), OpMax(), wt[0], nullptr);
    const int64_t carry = T.state[blockIdx.x];
    // the line in force at this thread's first byte
    const int64_t entry = ex >= 0 ? 2 * tile0 + ex : carry;
    const int fls = m.ls ? __builtin_ctz(m.ls) : 32;
    // bytes before the first line start here: kept iff the entry line is a sequence line after the first header
    const bool entry_keep = entry >= 0 && !(entry & 1) && (entry >> 1) >= first_hdr;
    // bytes of lines starting here: kept iff the line is not a header and starts at or after the first header
    const int64_t cut = first_hdr - base;
    const uint32_t after_first = cut <= 0 ? ~0u : cut >= 32 ? 0u : fa_from(int(cut));
    const uint32_t keep = m.valid & ~m.eol & ~m.hl &
                          ((entry_keep ? fa_below(fls) : 0u) | (fa_from(fls) & after_first));
    // kept bytes and headers before this thread in one scan: headers | kept << 16 (each <= 16,384)
    const int32_t packed = int32_t(__builtin_popcount(m.hs)) | int32_t(__builtin_popcount(keep)) << 16;
    int32_t tile_sum;
    const int32_t pex = block_exscan(packed, int32_t(0), OpSum(), wt[1], &tile_sum);
    const int64_t ob = T.kloc[blockIdx.x];           // output offset of the tile's first kept byte
    const int sh = int(ob & 15);
    int lp = pex >> 16;                               // offset in the tile of this thread's first kept byte
    int64_t rec = T.nhdr[blockIdx.x] + (pex & 0xffff);
    // record starts, and the A/C/G/T flag per record segment of these bytes
    const uint32_t a_ = fa_mask_bytes(r, 'A') | fa_mask_bytes(r, 'C') | fa_mask_bytes(r, 'G') |
                        fa_mask_bytes(r, 'T');
    const uint32_t odd = keep & ~a_;                  // kept bytes that are not A, C, G or T
    uint32_t seg = fa_below(m.hs ? __builtin_ctz(m.hs) : 32);
    if ((odd & seg) && rec >= 1 && rec - 1 < max_records) acgt[rec - 1] = 0;
    for (uint32_t rem = m.hs; rem;) {
        const int s = __builtin_ctz(rem);
        rem &= rem - 1;
        if (rec < max_records) seq_off[rec] = ob + lp + __builtin_popcount(keep & fa_below(s));
        seg = fa_from(s) & fa_below(rem ? __builtin_ctz(rem) : 32);
        if ((odd & seg) && rec < max_records) acgt[rec] = 0;
        ++rec;
    }
#pragma unroll
    for (int q = 0; q < kFaBytes; ++q)
        if (keep & (1u << q)) stage[sh + lp++] = r.b[q];
    __syncthreads();
    const int nk = tile_sum >> 16;
    if (nk == 0) return;
    // output [ob, ob + nk): bytes up to the first 16-byte boundary, 128-bit words, bytes after the last
    const int64_t a0 = (ob + 15) & ~int64_t(15), a1 = (ob + nk) & ~int64_t(15);
    if (a0 >= a1) {
        for (int j = threadIdx.x; j < nk; j += kFaThreads) seq[ob + j] = stage[sh + j];
        return;
    }
    const int head = int(a0 - ob), words = int((a1 - a0) >> 4), tail = int(ob + nk - a1);
    for (int j = threadIdx.x; j < head; j += kFaThreads) seq[ob + j] = stage[sh + j];
    for (int j = threadIdx.x; j < words; j += kFaThreads) {
        const uint4 v = *reinterpret_cast<const uint4*>(stage + sh + head + 16 * j);
        *reinterpret_cast<uint4*>(seq + a0 + 16 * int64_t(j)) = v;
        if (pack)   // the 2-bit word of these 16 output bytes (words cut by a tile seam: fa_pack2_edges)
            pack[(a0 >> 4) + j] = fa_code8(v.x) | fa_code8(v.y) << 8 | fa_code8(v.z) << 16 | fa_code8(v.w) << 24;
    }
    for (int j = threadIdx.x; j < tail; j += kFaThreads) seq[a1 + j] = stage[sh + head + 16 * words + j];
}

// F: one thread per record: length and the empty-record check.
__global__ __launch_bounds__(kFaThreads) void fa_records(const int64_t* seq_off, int64_t cap, int32_t* seq_len,
                                                        uint8_t* acgt, FaInfo* info)
{
    const int64_t k = int64_t(blockIdx.x) * kFaThreads + threadIdx.x;
    if (!info->ok) return;
    const int64_t n = info->n_records;
    if (k >= n || k >= cap) return;
    const int64_t o = seq_off[k];
    const int64_t e = k + 1 < n ? seq_off[k + 1] : info->total;
    seq_len[k] = e - o > INT32_MAX ? INT32_MAX : int32_t(e - o);
    if (e == o) {
        acgt[k] = 0;
        info->bad_rec = 1;
    } else if (e - o > INT32_MAX) {
        info->bad_rec = 2;
    }
}

// G: the 2-bit words (16 bases to a 32-bit word, base p at bits 2(p mod 16)
// of word p/16) that pass E could not write whole: the word holding a tile's
// first output byte and the one holding its last, when a tile seam falls
// inside them.  One thread a tile, from the finished sequence stream; tiles
// sharing a word write the same value.
__global__ __launch_bounds__(kFaThreads) void fa_pack2_edges(const uint8_t* seq, FaTiles T, int64_t n,
                                                            const FaInfo* info, uint32_t* pack)
{
    if (!info->ok) return;
    const int64_t t = int64_t(blockIdx.x) * kFaThreads + threadIdx.x;
    if (t >= n) return;
    const int64_t total = info->total;
    const int64_t ob = T.kloc[t], oe = t + 1 < n ? T.kloc[t + 1] : total;
    if (oe <= ob) return;
    const int64_t ends[2] = {(ob & 15) ? ob >> 4 : -1, (oe & 15) ? (oe - 1) >> 4 : -1};
    for (int e = 0; e < 2; ++e) {
        const int64_t w = ends[e];
        if (w < 0) continue;
        uint32_t word = 0;
        for (int q = 0; q < 16 && 16 * w + q < total; ++q) {
            const uint32_t c = seq[16 * w + q];
            word |= (fa_base(uint8_t(c)) ? ((c >> 1) ^ (c >> 2)) & 3u : 0u) << (2 * q);
        }
        pack[w] = word;
    }
}

}  // namespace

extern "C" int ga_fasta_device(const uint8_t* data, int64_t len, const ga_fasta_dev_out* out, int64_t* n_records,
                               int64_t* total, void* stream, const char** error)
{
    static thread_local std::string msg;
    if (error) *error = nullptr;
    if (n_records) *n_records = 0;
    if (total) *total = 0;
    if (!out || !out->seq || !out->seq_off || !out->seq_len || !out->acgt || (!data && len > 0) || len < 0 ||
        out->max_records < 0) {
        if (error) *error = "null argument";
        return GA_INVALID_ARGUMENT;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    auto fail_dev = [&](const char* what) {
        msg = what;
        if (error) *error = msg.c_str();
        return GA_DEVICE_ERROR;
    };
    const int64_t ntiles = len > 0 ? (len + kFaTile - 1) / kFaTile : 0;
    // scratch: 8 per-tile arrays, 6 per-group arrays and the info block, one allocation
    const int64_t nt = ntiles > 0 ? ntiles : 1;
    const int64_t ngroups = (nt + kScanThreads - 1) / kScanThreads;
    void* scratch = nullptr;
    if (hipMallocAsync(&scratch, size_t(8 * nt + 6 * ngroups) * sizeof(int64_t) + sizeof(FaInfo), st) !=
        hipSuccess)
        return fail_dev("hipMalloc(FASTA scratch)");
    int64_t* a = static_cast<int64_t*>(scratch);
    const FaTiles T{a, a + nt, a + 2 * nt, a + 3 * nt, a + 4 * nt, a + 5 * nt, a + 6 * nt, a + 7 * nt};
    int64_t* gp = a + 8 * nt;
    const FaGroups G{gp, gp + ngroups, gp + 2 * ngroups, gp + 3 * ngroups, gp + 4 * ngroups, gp + 5 * ngroups};
    FaInfo* info = reinterpret_cast<FaInfo*>(gp + 6 * ngroups);
    FaInfo h{};
    hipError_t e = hipSuccess;
    const bool al = (reinterpret_cast<uintptr_t>(data) & 15) == 0;
    if (ntiles > 0) {
        const dim3 g{unsigned(ntiles), 1, 1}, b{unsigned(kFaThreads), 1, 1};
        if (al) hipLaunchKernelGGL(fa_tiles_a<true>, g, b, 0, st, data, len, T);
        else    hipLaunchKernelGGL(fa_tiles_a<false>, g, b, 0, st, data, len, T);
        const dim3 gg{unsigned(ngroups), 1, 1}, sb{unsigned(kScanThreads), 1, 1};
        hipLaunchKernelGGL(fa_scan_up, gg, sb, 0, st, T, ntiles, len, G);
        hipLaunchKernelGGL(fa_scan_mid, gg, sb, 0, st, T, ntiles, len, G, ngroups, info);
        hipLaunchKernelGGL(fa_scan_down, gg, sb, 0, st, T, ntiles, len, G, ngroups, out->max_records, info);
        const int64_t cap = out->max_records < len ? out->max_records : len;   // records <= bytes
        if (cap > 0) {
            const dim3 rg{unsigned((cap + kFaThreads - 1) / kFaThreads), 1, 1};
            hipLaunchKernelGGL(fa_flags_init, rg, b, 0, st, out->acgt, cap, info);
            if (al)
                hipLaunchKernelGGL(fa_tiles_e<true>, g, b, 0, st, data, len, T, info, out->seq, out->seq_off,
                                   out->acgt, out->pack2, out->max_records);
            else
                hipLaunchKernelGGL(fa_tiles_e<false>, g, b, 0, st, data, len, T, info, out->seq, out->seq_off,
                                   out->acgt, out->pack2, out->max_records);
            hipLaunchKernelGGL(fa_records, rg, b, 0, st, out->seq_off, cap, out->seq_len, out->acgt, info);
            if (out->pack2) {
                const dim3 pg{unsigned((ntiles + kFaThreads - 1) / kFaThreads), 1, 1};
                hipLaunchKernelGGL(fa_pack2_edges, pg, b, 0, st, out->seq, T, ntiles, info, out->pack2);
            }
        }
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h, info, sizeof(FaInfo), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFreeAsync(scratch, st);
    if (e != hipSuccess) return fail_dev("FASTA kernels");
    const bool malformed = ntiles == 0 || h.n_records == 0 || h.bad_pre < h.first_hdr || h.bad_id < len ||
                           (h.ok && h.bad_rec != 0);
    if (malformed) {
        // word the error as the host reader does, on a copy of the bytes
        std::vector<char> host(static_cast<size_t>(len));
        if (len > 0 && (hipMemcpy(host.data(), data, size_t(len), hipMemcpyDeviceToHost) != hipSuccess))
            return fail_dev("hipMemcpy(FASTA bytes)");
        std::string err;
        if (ga_fasta_host_message(host.data(), len, err))
            err = h.bad_rec == 2 ? "record longer than 2147483647 bytes" : "malformed FASTA";
        msg = err;
        if (error) *error = msg.c_str();
        return GA_INVALID_ARGUMENT;
    }
    if (h.n_records > out->max_records) {
        // *n_records reports the count needed; nothing was written
        if (n_records) *n_records = h.n_records;
        msg = "more records than max_records";
        if (error) *error = msg.c_str();
        return GA_INVALID_ARGUMENT;
    }
    if (n_records) *n_records = h.n_records;
    if (total) *total = h.total;
    return GA_OK;
}
