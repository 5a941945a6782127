// ga_device.hpp -- launch-side declarations shared by the gfx950 kernels and the
// host layer.  Everything here is internal to libgenoaligner_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include <functional>

struct ga_wfa_dev_batch;
struct ga_route_stats;

namespace ga {

// Packed sequence buffers carry this many readable bytes before their first
// and after their last sequence: the extension and streaming loops read aligned
// 8-byte words up to 72 bytes outside the sequence they work on.
constexpr int kSeqPad = 128;

// Sentinel of an unreachable wavefront offset and its threshold test
// (reference: include/genoaligner/backend/wfa_kernel.hip:103,110).
constexpr int kWfaNeg = INT32_MIN / 2;

// Persistent WFA over one size class: n_groups workgroups of `waves`
// wavefronts claim pairs order[0..n) from an atomic counter (zeroed by the
// launcher) and align them one at a time; want[pid] selects score + CIGAR (else
// score only).  Each workgroup owns a wavefront stack of (smax_cls+1)^2
// offsets, 16-bit unless `wide` (needed when a text is 32767 bytes or longer).
// `staged`: both sequences of a pair are copied into LDS (sized by max_plen,
// max_tlen) and extended there.
struct WfaQueueLaunch {
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;
    const int32_t* smax;
    const uint8_t* want;     // null: every pair gets a CIGAR when cigar != null
    const int32_t* order;    // claim order (pair ids); null: pairs 0 .. n-1
    int32_t n;
    int32_t* next;           // claim counter, 0 at launch
    int32_t smax_cls;        // >= smax of every pair in the class (LDS + stack sizing)
    int32_t max_plen, max_tlen;
    int32_t waves;           // wavefronts per workgroup: 1, 2, 4 or 8
    int32_t n_groups;
    int32_t* score;
    uint8_t* cigar; const int64_t* cig_off; int32_t* cig_len;
    void* ws;                // n_groups * ws_stride entries of int16 (int32 if wide)
    int64_t ws_stride;       // entries per workgroup, >= (smax_cls+1)^2 (0: exactly that)
    bool wide;
    bool staged;
    const int32_t* dev_class; // device-planned: {first, count} of this class (then n, order ignored / offset)
    bool counter_zeroed;      // host-planned: `next` already zeroed on the stream (else the launcher zeroes it)
};

// Deferred walks: pairs a workgroup queues before walking them together.
constexpr int kWfaPend = 32;

// Size classes by smax (device planning uses the same bounds as the host plan).
constexpr int kWfaClasses = 4;
constexpr int kWfaClassBound0 = 64, kWfaClassBound1 = 256, kWfaClassBound2 = 1024;
constexpr int kPlanKeys = 1024;              // length bins of 32 bytes per class

struct WfaDevPlan {
    const int32_t* smax; const int32_t* plen; const int32_t* tlen;
    int32_t n;
    int32_t* hist;           // kWfaClasses * kPlanKeys
    int32_t* cls;            // 2 * kWfaClasses: {first, count} per class
    int32_t* order;          // n
    int32_t* next;           // kWfaClasses claim counters, zeroed by the plan
};
constexpr int kPlanSmallMax = 65536;         // single-workgroup plan up to this many pairs
hipError_t wfa_plan_launch(const WfaDevPlan& P, hipStream_t st);

size_t wfa_lds_bytes(const WfaQueueLaunch& L);     // dynamic LDS of one workgroup
int wfa_queue_groups(const WfaQueueLaunch& L);     // resident workgroups (<= n)
// True when a launch is scheduled statically throughout (short pairs of the
// smallest class, at most 8 per workgroup): it never reads its claim counter.
inline bool wfa_fully_static(const WfaQueueLaunch& L)
{
    return L.smax_cls <= kWfaClassBound0 && L.max_plen <= 1024 && L.max_tlen <= 1024 && !L.dev_class &&
           int64_t(L.n) <= int64_t(8) * L.n_groups;
}
size_t wfa_queue_ws_bytes(int smax_cls, int n_groups, bool wide);
hipError_t wfa_queue_launch(const WfaQueueLaunch& L, hipStream_t st);
// p[0 .. n) = 0 by a kernel (graph-capturable, unlike a memset node)
hipError_t zero_i32_launch(int32_t* p, int n, hipStream_t st);

// ---------------------------------------------------------------- router
constexpr int kSketchDim = 64;
constexpr int kHidden = 128;
constexpr int kBuckets = 16;
// parameter offsets (floats) of the MLP 64 -> 128 -> 128 -> 16; W row-major [in][out]
struct MlpLayout {
    static constexpr int w1 = 0;
    static constexpr int b1 = w1 + kSketchDim * kHidden;
    static constexpr int w2 = b1 + kHidden;
    static constexpr int b2 = w2 + kHidden * kHidden;
    static constexpr int w3 = b2 + kHidden;
    static constexpr int b3 = w3 + kHidden * kBuckets;
    static constexpr int total = b3 + kBuckets;
};
constexpr int kMlpParams = MlpLayout::total;     // 26,896

struct SketchLaunch {
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;
    int64_t n;
    float* feat;             // n * kSketchDim
    int32_t* lower_bound;    // n
};
// Compute units of the current device (cached per device).
int device_cus();
hipError_t sketch_launch(const SketchLaunch& L, hipStream_t st);
hipError_t mlp_forward(const float* params, const float* X, int64_t N, float* H1, float* H2, float* logits,
                       hipStream_t st);
// logits only, layers fused (inference; params 16-byte aligned)
hipError_t mlp_logits(const float* params, const float* X, int64_t N, float* logits, hipStream_t st);
hipError_t mlp_backward(const float* params, const float* X, const float* H1, const float* H2, const float* logits,
                        const int32_t* label, int64_t N, float* grads, float* loss, float* dL, float* dH2,
                        float* dH1, hipStream_t st);
hipError_t mlp_sgd(float* params, const float* grads, float* mom, float lr, float mu, hipStream_t st);
hipError_t mlp_route(const float* logits, int64_t N, float q, int32_t* bucket, hipStream_t st);
hipError_t route_first_pass_launch(const int32_t* bucket, const int32_t* lb, const int32_t* smax, const int32_t* plen,
                                   const int32_t* tlen, int n, int32_t* sm1, int32_t* score, int32_t* cig_len,
                                   unsigned long long* stats, hipStream_t st);
hipError_t route_second_pass_launch(const int32_t* sm1, const int32_t* smax, const int32_t* score, int n, int32_t* sm2,
                                    unsigned long long* stats, hipStream_t st);
hipError_t zero_u64_launch(unsigned long long* p, int n, hipStream_t st);

struct SwLaunch {
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;   // rows
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;   // columns
    int32_t n;              // blocks of the forward kernel / threads of the walk
    int32_t max_m;          // max text_len in the launch
    int32_t match, mismatch, go, ge;
    // out: 8 ints per pair: score, end_i, end_j (row-major first), lane_end_i,
    // lane_end_j (wave64 lane rule), start_i, start_j, cigar_len
    int32_t* res;
    uint32_t* dirs;         // direction nibbles (trace)
    const int64_t* dir_off; // per pair word offset into dirs, or null: then pair
    int64_t dir_stride;     //   first_pair + q starts at word q * dir_stride
    const uint8_t* walk;    // per pair: 1 = emit span + CIGAR (trace); null = all
    uint8_t* cigar; const int64_t* cig_off;
    int32_t* bnd;           // global strip-boundary scratch (max_m*2 ints per pair) or null
    int32_t first_pair;
    // Re-run mode: block/thread q serves slot (wave q/2, half q%2) of a packed
    // launch (pairs from pair_a/pair_b, or first_pair + q when those are null),
    // only when rerun[pid] is set; its direction words start at
    // dirs + (q/2) * dir_stride + (q%2) * (dir_stride/2).
    const uint8_t* rerun;
    const int32_t* pair_a; const int32_t* pair_b;
    int32_t n_total;
    // Substitution-matrix scoring (null: match / mismatch): residue codes
    // 0..31 of the text and pattern, laid out like txt / pat, and the 32 x 32
    // score table sub[text code * 32 + pattern code].
    const uint8_t* tcode; const uint8_t* pcode; const int32_t* sub;
    bool sub8;              // every sub entry fits int8 (8-bit LDS profiles)
};

// Marks, in res[0] of a pair, that the packed 16-bit kernel left it to the int32 kernel.
constexpr int kSwRerun = -2;

struct Sw16Launch {
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;
    const int32_t* pair_a; const int32_t* pair_b;   // per wave; null: first_pair + 2w, + 1
    int32_t first_pair;
    int32_t n_total;        // pairs in the arrays (a half whose pair id >= n_total is empty)
    int32_t n_waves;
    int32_t max_m;          // LDS sizing: max text_len over the launch's pairs
    int32_t max_n;          //   and max pattern_len (trace: staging for the walk)
    int32_t match, mismatch, go, ge;
    int32_t* res;           // 8 ints per pair, as SwLaunch
    uint8_t* rerun;         // out: 1 = needs the int32 kernel (2, inside a launch: the literal-byte pass)
    // Optional (DNA launches): n_waves + 1 ints of scratch; with it, pairs holding
    // a byte other than A/C/G/T get a packed literal-byte pass instead of the
    // int32 kernel (schemes with mismatch < 0)
    int32_t* bytes_list;
    uint32_t* dirs; int64_t dir_stride;             // words per wave (trace)
    uint32_t* selbuf;       // per-wave text selector rows (scratch), sw16_sel_stride dwords per wave
    const uint8_t* walk;    // per pair: emit span + CIGAR; null = all
    uint8_t* cigar; const int64_t* cig_off;
    // substitution-matrix launches (sw16m_launch): residue codes 0..31 laid
    // out like the bytes, the 32 x 32 score table (row: text residue), and the
    // bias the kernel adds to every score (every biased score fits 0..255)
    const uint8_t* tcode; const uint8_t* pcode; const int32_t* sub;
    int32_t bias;
};

// Largest text length the packed kernel takes (LDS: 8 bytes per text row,
// 64 KB at this length).
constexpr int kSw16MaxRows = 8192;
// Largest pattern length the packed kernel takes in a CIGAR batch (its walk
// stages both pairs' sequences and CIGARs, 4 bytes per row and column, in LDS:
// with kSw16MaxRows at most 66.3 KB per wave, of the 160 KB a workgroup may
// hold).
constexpr int kSw16MaxCols = 8192;

// Selector scratch per wave: 64 guard rows before and 80 after max_m rows
// (reads reach 63 rows before row 0 and 77 rows past the last row).
__host__ __device__ inline int64_t sw16_sel_stride(int max_m) { return (int64_t(max_m) + 144 + 3) & ~int64_t(3); }

inline int64_t sw16_dir_words_per_wave(int max_m, int max_n)
{
    if (max_m <= 0 || max_n <= 0) return 0;
    const int64_t strips = (max_n + 63) / 64;
    const int64_t nq = ((max_m + 63 + 3) & ~3) / 4;
    const int64_t own = strips * nq * 64;
    // room for the int32 layout of both halves when they are re-run
    const int64_t alt = 2 * strips * int64_t((max_m + 63 + 7) / 8) * 64;
    return own > alt ? own : alt;
}

// Words of direction nibbles one pair needs for the trace path.
inline int64_t sw_dir_words(int m, int n)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t strips = (n + 63) / 64;
    const int64_t wps = (m + 63 + 7) / 8;
    return strips * wps * 64;
}

// Optional callbacks around a launcher's dominant kernel (timing).
struct LaunchHooks {
    virtual void before(hipStream_t) {}
    virtual void after(hipStream_t) {}
    virtual ~LaunchHooks() = default;
};

hipError_t sw_launch(const SwLaunch& L, bool trace, hipStream_t st, LaunchHooks* hooks = nullptr);

// Packed kernel over L.n_waves waves (two pairs each), then -- separately --
// its walk over 2 * n_waves slots.
hipError_t sw16_launch(const Sw16Launch& L, bool trace, hipStream_t st, LaunchHooks* hooks = nullptr);
// LDS bytes per wave of a packed DNA launch over pairs up to these lengths.
int64_t sw16_lds_need(int max_m, int max_n, bool trace);

// Substitution-matrix pairs, two per workgroup (packed 16-bit), their strips
// pipelined over the workgroup's waves; n_waves counts workgroups.  The trace
// variant then walks every pair not flagged for the int32 re-run.
constexpr int kSw16mMaxRows = 4096;
int sw16m_bias(const int32_t* sub32x32, int gap_open);   // -1: some biased score leaves 0..255
// Direction words per workgroup of sw16m_launch (its own layout, or the int32
// layout of both pairs when they are re-run, whichever is larger).
int64_t sw16m_dir_words_per_group(int max_m, int max_n);
hipError_t sw16m_launch(const Sw16Launch& L, bool trace, hipStream_t st, LaunchHooks* hooks = nullptr);

// CIGAR checks on the device (the reference's sw_validate, api.cpp:114-156,
// one wave per pair): pair k's fields are read at score[k * stride] etc.,
// its CIGAR at cigar + cig_off[k] + text_len[k] + pattern_len[k] - len (the
// right-aligned slot every SW path writes).  flags[k] = rescore_ok |
// wellformed_ok << 1 for pairs with score > 0, len > 0 and walk[k] (walk may
// be null: all), 0 for the others.
struct SwValidate {
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;
    const uint8_t* cigar; const int64_t* cig_off;
    const int32_t* score; const int32_t* si; const int32_t* sj; const int32_t* ei; const int32_t* ej;
    int stride;                                   // ints between one pair's fields and the next's
    const int32_t* len; int len_stride;
    const uint8_t* walk;
    uint8_t* flags;
    int64_t n;
    int match, mismatch, go, ge;
    // substitution-matrix batches: M and X score sub[tcode * 32 + pcode]
    // (residue codes laid out like the bytes) instead of match / mismatch
    const uint8_t* tcode; const uint8_t* pcode; const int32_t* sub;
    // global edit CIGARs (the WFA check, api.cpp:84-106): from (0, 0) to the
    // ends of both sequences, the score = X + I + D ops, rescore_ok on the
    // score alone; si/sj/ei/ej unused, pairs with score >= 0 checked
    int edit;
};
hipError_t sw_validate_launch(const SwValidate& V, hipStream_t st);
// Pair k's CIGAR (the coff[k+1] - coff[k] bytes ending its slot at
// cigar + slot_off[k] + tlen[k] + plen[k]) copied to out + coff[k]: the
// CIGARs back to back for a short copy to the host.
hipError_t sw_cigar_compact_launch(const uint8_t* cigar, const int64_t* slot_off, const int32_t* tlen,
                                   const int32_t* plen, const int64_t* coff, int64_t n, uint8_t* out,
                                   hipStream_t st);

// Reduce the 8-int per-pair records to the device-resident result layout:
// out5 = score, start_i, start_j, end_i, end_j (trace: row-major end and span;
// score-only: the wave64 lane-rule end), cig_len = CIGAR length (may be null).
hipError_t sw_gather(const int32_t* res8, int64_t n, bool trace, int32_t* out5,
                     int32_t* cig_len, hipStream_t st);

// ga_wfa_device with device planning and every scratch buffer carved from
// one device region (plan_bytes for the claim order and plan, stack_bytes of
// wavefront stacks reused by each size class in turn); the classes run one
// after another on `st`.  Allocates nothing.  Returns an error string or null.
// The routed WFA batch (router kernels, routed pass, plain re-run) on st.
const char* wfa_routed_run(const ga_wfa_dev_batch* b, const float* params, float q, ga_route_stats* stats,
                           hipStream_t st);
// The network loaded by ga_router_set (device pointer, null when none) and its quantile.
const float* router_params(float* q);
const char* wfa_device_in_region(const ga_wfa_dev_batch* b, hipStream_t st, uint8_t* region, size_t plan_bytes,
                                 size_t stack_bytes);

// ---------------------------------------------------------------- MSA
// Profile-profile Gotoh (msa.hip).  Profiles live in one column table (all
// profiles back to back, profile k at columns col_off[k] .. + len[k]); the
// sparse form holds, per column c, nz_cnt[c] non-zero symbol fractions in
// symbol order at nz_sym / nz_frq[k * n_cols + c] and the column's gap costs.
struct MsaParamsDev {
    float match, ts, tv, go, ge;
    int free_end, psgp;
    float psgp_scale, psgp_min_open, psgp_min_ext;
    int alpha;                   // 4 or 20
};
struct MsaSparseLaunch {
    const float* cols;           // n_cols x alpha fractions
    const float* occ;            // n_cols
    int64_t n_cols;
    MsaParamsDev P;
    uint8_t* nz_cnt; uint8_t* nz_sym; float* nz_frq;
    float* gopen; float* gext;
};
struct MsaLaunch {
    const uint8_t* nz_cnt; const uint8_t* nz_sym; const float* nz_frq;
    const float* gopen; const float* gext;
    int64_t n_cols;
    const int64_t* col_off; const int32_t* len;        // per profile
    const int32_t* a_idx; const int32_t* b_idx;        // per pair
    int32_t first_pair;
    const float* sub;            // alpha x alpha substitution scores (device)
    MsaParamsDev P;
    uint32_t* dirs; const int64_t* dir_off;            // msa_dir_words(M, N) words per pair
    float* bnd; const int64_t* bnd_off;                // 3 (M + 1) floats per pair
    float* score; int32_t* span;                       // span: ai, aj, bi, bj
    uint8_t* cigar; const int64_t* cig_off; int32_t* cig_len;   // slot of M + N bytes, CIGAR right-aligned
};
int64_t msa_dir_words(int M, int N);
hipError_t msa_sparse_launch(const MsaSparseLaunch& S, hipStream_t st);
hipError_t msa_pp_launch(const MsaLaunch& L, int n_pairs, hipStream_t st, LaunchHooks* hooks = nullptr);

// k-mer distances and neighbour joining (nj.hip).
struct KmerLaunch {
    const uint8_t* seq; const int64_t* off; const int32_t* len;
    int32_t n, k, alpha;
    int64_t words;               // bitset words per sequence: ceil(alpha^k / 32)
    uint32_t* bits;              // n x words, zeroed by the caller
    int32_t* count;              // n: |K_i|
    float* dist;                 // n (n - 1) / 2 packed upper triangle
};
hipError_t kmer_distances_launch(const KmerLaunch& K, int max_len, hipStream_t st, LaunchHooks* hooks = nullptr);

struct NjBest {                  // a merge candidate; a < 0: none
    double q;
    int32_t a, b;
};
struct NjLaunch {
    int32_t n;
    const float* packed;         // n (n - 1) / 2 distances
    double* d;                   // n x n
    double* r;                   // n
    NjBest* part;                // n per-row first minima
    int32_t* alive0; int32_t* alive1; int32_t* slot_node;   // n each
    int32_t* children;           // 2 (n - 1): children of internal nodes n .. 2n-2
};
hipError_t nj_launch(const NjLaunch& J, hipStream_t st, LaunchHooks* hooks = nullptr);

// Run chunk(k0, k1) over [0, n) on the host work pool, about `grain` indices a
// chunk; returns when every chunk has run.
void host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& chunk);

}  // namespace ga
