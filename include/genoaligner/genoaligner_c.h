/*
 * genoaligner_c.h -- C ABI of the MI355X (gfx950) pairwise aligner.
 *
 * Plain C, POD structs, plain pointers and sizes.  Every entry point below
 * either replaces one function of the reference's public C++ API
 * (include/genoaligner/api.hpp, implemented in src/api/api.cpp) with the same
 * argument meaning, result rules and error behaviour, or is an explicitly
 * marked extension with no reference counterpart.  include/genoaligner/api.hpp
 * in this repository re-exposes the reference's C++ surface on top of this ABI.
 *
 * Conventions shared by all entry points
 *   - Status codes mirror BatchResult::Status / SWBatchResult::Status
 *     (api.hpp:158, :306): GA_OK, GA_INVALID_ARGUMENT, GA_DEVICE_ERROR.
 *     `*error` receives a static string, non-NULL iff the status is not GA_OK.
 *   - A batch is refused as a whole, before any device work, on any invalid
 *     request; results are then reset to their defaults.
 *   - CIGARs are byte strings over {M,X,I,D}; I consumes text, D consumes
 *     pattern.  They are written into a caller-owned arena: pair i's CIGAR is
 *     cigar_arena[res[i].cigar_off .. + res[i].cigar_len).  The arena must hold
 *     at least sum(pattern_len + text_len) bytes over the batch
 *     (ga_cigar_arena_bytes_wfa/_sw compute it); a smaller arena is refused with
 *     GA_INVALID_ARGUMENT before any device work.
 *   - Calls are synchronous and thread-safe for disjoint inputs (api.hpp:91-111):
 *     each call runs on its own HIP stream; the input bytes must stay alive and
 *     unmodified until the call returns.
 */
#ifndef GENOALIGNER_C_H
#define GENOALIGNER_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GA_OK = 0, GA_INVALID_ARGUMENT = 1, GA_DEVICE_ERROR = 2 };

/* api.cpp:227 -- the reference contract's smax bound for ga_align_batch. */
#define GA_MAX_SMAX 511
/* Extension bound for ga_align_batch_ext (per-pair smax). */
#define GA_MAX_SMAX_EXT 4000
/* api.hpp:256 -- largest SW matrix ((text_len+1)*(pattern_len+1)) traced. */
#define GA_SW_MAX_TRACE_CELLS (1 << 26)

/* ---- WFA edit distance ---------------------------------------------------- */

/* AlignRequest, api.hpp:81-88 (same fields, same defaults: smax 64, with_cigar 1). */
typedef struct ga_align_req {
    const char* pattern;
    int32_t     pattern_len;
    const char* text;
    int32_t     text_len;
    int32_t     smax;
    int32_t     with_cigar;
} ga_align_req;

/* AlignResult, api.hpp:64-76. */
typedef struct ga_align_res {
    int32_t score;          /* edit distance; -1 = unresolved (distance > smax) */
    int32_t resolved;
    int64_t cigar_off;      /* into the caller's arena */
    int32_t cigar_len;      /* 0 when unresolved */
    int32_t rescore_ok;
    int32_t wellformed_ok;
} ga_align_res;

/* Replaces genoaligner::align_batch (api.hpp:164, api.cpp:234-497).
 * One smax for the batch: the minimum over the requests, each of which must lie
 * in [0, GA_MAX_SMAX].  If any request has with_cigar, every resolved pair gets
 * a CIGAR and both validators (api.cpp:380-470); otherwise scores only. */
int ga_align_batch(const ga_align_req* reqs, int64_t n, ga_align_res* res,
                   char* cigar_arena, int64_t arena_bytes,
                   int32_t* resolved_count, int32_t* unresolved_count,
                   const char** error);

/* EXTENSION (no reference counterpart): per-pair smax in [0, GA_MAX_SMAX_EXT]
 * and per-pair with_cigar (score-only pairs get no CIGAR).  Otherwise the same
 * per-pair results as ga_align_batch with that pair's smax. */
int ga_align_batch_ext(const ga_align_req* reqs, int64_t n, ga_align_res* res,
                       char* cigar_arena, int64_t arena_bytes,
                       int32_t* resolved_count, int32_t* unresolved_count,
                       const char** error);

/* ---- Smith-Waterman, affine gaps ------------------------------------------ */

/* SWScoring, api.hpp:246-251 (defaults 2, -3, 5, 2). */
typedef struct ga_sw_scoring {
    int32_t match;
    int32_t mismatch;
    int32_t gap_open;
    int32_t gap_extend;
} ga_sw_scoring;

/* SWRequest, api.hpp:261-268: text is the row axis, pattern the column axis. */
typedef struct ga_sw_req {
    const char*   text;
    int32_t       text_len;
    const char*   pattern;
    int32_t       pattern_len;
    ga_sw_scoring scoring;
    int32_t       with_cigar;
} ga_sw_req;

/* SWAlignResult, api.hpp:273-294. */
typedef struct ga_sw_res {
    int32_t score;
    int32_t resolved;
    int64_t cigar_off;
    int32_t cigar_len;
    int32_t start_i, start_j, end_i, end_j;
    int32_t rescore_ok;
    int32_t wellformed_ok;
    int32_t too_large;
} ga_sw_res;

/* Replaces genoaligner::align_sw_batch (api.hpp:312, api.cpp:584-847). */
int ga_align_sw_batch(const ga_sw_req* reqs, int64_t n, ga_sw_res* res,
                      char* cigar_arena, int64_t arena_bytes,
                      int32_t* resolved_count, int32_t* unresolved_count,
                      const char** error);

/* ---- device queries (api.hpp:485-487) -------------------------------------- */
const char* ga_backend_name(void);
const char* ga_device_name(void);
int         ga_device_available(void);

/* Bytes of CIGAR arena a batch needs: sum over pairs of pattern_len + text_len
 * (negative lengths count as 0). */
/* ---- substitution-matrix Smith-Waterman (EXTENSION: no reference counterpart;
 * the reference's SW takes match/mismatch only) ------------------------------
 * Same local affine alignment, tie rules, CIGAR, too_large and score-only end
 * rules as ga_align_sw_batch, with the substitution score of text residue a
 * against pattern residue b read from score[code(a)][code(b)], code = index of
 * the byte in `alphabet`; a pair involving a byte outside the alphabet scores
 * `unknown`.  M / X in the CIGAR still mean equal / unequal bytes.  A gap of
 * length L costs gap_open + (L - 1) * gap_extend (so BLAST's 11/1 is
 * gap_open 12, gap_extend 1 here).  The requests' `scoring` fields are
 * ignored.  rescore_ok re-scores the CIGAR with the matrix. */
typedef struct ga_sw_matrix {
    char    alphabet[32];      /* 1..31 distinct letters, NUL-terminated if shorter */
    int8_t  score[32][32];
    int32_t unknown;
    int32_t gap_open, gap_extend;
} ga_sw_matrix;

int ga_align_sw_matrix_batch(const ga_sw_req* reqs, int64_t n, const ga_sw_matrix* matrix, ga_sw_res* res,
                             char* arena, int64_t arena_bytes, int32_t* resolved_count,
                             int32_t* unresolved_count, const char** error);

/* BLOSUM62 over ARNDCQEGHILKMFPSTWYV as the reference's table holds it
 * (src/msa/msa_ref.cpp:116-137; its row M, column H is -1, the published
 * matrix has -2), unknown residues -1, with the given gaps. */
void ga_sw_matrix_blosum62(ga_sw_matrix* out, int32_t gap_open, int32_t gap_extend);

int64_t ga_cigar_arena_bytes_wfa(const ga_align_req* reqs, int64_t n);
int64_t ga_cigar_arena_bytes_sw(const ga_sw_req* reqs, int64_t n);

/* The CIGAR check ga_align_sw_batch applies to each result (the reference's
 * sw_validate, src/api/api.cpp:114-156): wellformed_ok when every op stays in
 * [start, end] and inside both sequences, M pairs equal and X unequal bytes,
 * and the walk ends at (end_i, end_j); rescore_ok when also the affine
 * re-score (a gap run of L costs gap_open + (L-1) gap_extend) equals score.
 * EXTENSION (static in the reference), for callers that check CIGARs of their
 * own.  GA_INVALID_ARGUMENT for a null pointer with a non-zero length. */
int ga_sw_validate_cigar(const char* cigar, int32_t cigar_len, const char* text, int32_t text_len,
                         int32_t start_i, int32_t end_i, const char* pattern, int32_t pattern_len,
                         int32_t start_j, int32_t end_j, const ga_sw_scoring* scoring, int32_t score,
                         int32_t* rescore_ok, int32_t* wellformed_ok);

/* ---- device-resident batches (EXTENSION: no reference counterpart) ---------
 * For callers whose sequences already sit in device memory (the benchmark,
 * the request loop).  All pointers are device pointers.  Sequences are packed
 * byte arrays addressed by int64 offsets; the packed arrays must be readable
 * for 128 bytes before their first and after their last sequence.  `cigar_off[i]` gives pair i's slot of
 * pattern_len + text_len bytes in `cigar`; the CIGAR is right-aligned in its
 * slot: it occupies cigar[cigar_off[i] + slot - cigar_len[i] .. + slot).
 * `stream` is a hipStream_t (NULL = the null stream).  Work is enqueued
 * asynchronously; results are valid once the stream is synchronised. */
typedef struct ga_wfa_dev_batch {
    const uint8_t* pattern; const int64_t* pattern_off; const int32_t* pattern_len;
    const uint8_t* text;    const int64_t* text_off;    const int32_t* text_len;
    const int32_t* smax;            /* per pair */
    int64_t        n_pairs;
    int32_t        smax_max;        /* >= every smax[i]; <= GA_MAX_SMAX_EXT */
    int32_t        max_len;         /* >= every pattern_len and text_len */
    int32_t*       score;           /* out */
    uint8_t*       cigar;           /* NULL = score only */
    const int64_t* cigar_off;
    int32_t*       cigar_len;       /* out */
    /* Optional scheduling metadata (may all be NULL):
     *   want_cigar    device, per pair: 0 = score only for that pair
     *   host_smax,    host copies of smax, pattern_len and text_len: with them
     *   host_pattern_len,  the batch is split into smax size classes, each run
     *   host_text_len      longest pair first, classes concurrently. */
    const uint8_t* want_cigar;
    const int32_t* host_smax;
    const int32_t* host_pattern_len;
    const int32_t* host_text_len;
} ga_wfa_dev_batch;

int ga_wfa_device(const ga_wfa_dev_batch* b, void* stream, const char** error);

/* EXTENSION with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): ga_wfa_device with device planning (the
 * host copies in `b` are ignored) that calls no hipMalloc / hipFree and no
 * allocating library routine.  Every scratch buffer (claim order, plan,
 * claim counters, wavefront stacks) lives in one static device array of
 * GA_WFA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the size
 * classes run one after another on `stream`, each with as many workgroups as
 * its stacks fit in the array.  Same scores and CIGARs as ga_wfa_device.
 * Refused with GA_INVALID_ARGUMENT: more than GA_WFA_NOALLOC_MAX_PAIRS pairs;
 * with CIGAR, a (smax_max + 1)^2 stack of 16-bit offsets (32-bit when max_len
 * >= 32767) larger than the array's stack part (any smax_max <= 4000 with
 * 16-bit offsets; up to 3962 with 32-bit ones).  Calls on one GPU share its
 * array: calls that overlap on one GPU race. */
#define GA_WFA_NOALLOC_SCRATCH_BYTES (64 << 20)
#define GA_WFA_NOALLOC_MAX_PAIRS (1 << 20)
int ga_wfa_device_noalloc(const ga_wfa_dev_batch* b, void* stream, const char** error);

/* Without the host copies, ga_wfa_device plans on the device: a histogram /
 * scan / scatter over (size class, length) gives each class its slice of a
 * longest-first claim order, and every class that can hold a pair is launched
 * over a full persistent grid, with no host synchronisation.
 *
 * Graph submission (EXTENSION): ga_wfa_graph_create captures that whole
 * device-planned submission of batch `b` (its buffers, sizes and bounds are
 * baked in) into a HIP graph; ga_wfa_graph_launch replays it on a stream, so a
 * caller that refills the same buffers resubmits with one call.  Workspace is
 * held by the graph until ga_wfa_graph_destroy. */
/* Autotuner (EXTENSION): wavefronts per workgroup (1, 2, 4 or 8) for each of
 * the four WFA size classes (smax <= 64, <= 256, <= 1024, above).
 * ga_wfa_autotune times every candidate on the pairs of batch `b` (which needs
 * the host copies; its outputs are written, and equal for every candidate)
 * and returns the fastest per class with its time; ga_wfa_set_tuning makes a
 * tuning current for the process (NULL restores the defaults). */
typedef struct ga_wfa_tuning {
    int32_t waves[4];
    float   class_ms[4];     /* autotune: time of the chosen candidate, -1 when the class was empty */
} ga_wfa_tuning;
int  ga_wfa_autotune(const ga_wfa_dev_batch* b, int reps, ga_wfa_tuning* out, void* stream, const char** error);
void ga_wfa_set_tuning(const ga_wfa_tuning* t);
void ga_wfa_get_tuning(ga_wfa_tuning* t);

typedef struct ga_wfa_graph ga_wfa_graph;
int  ga_wfa_graph_create(const ga_wfa_dev_batch* b, ga_wfa_graph** out, const char** error);
int  ga_wfa_graph_launch(ga_wfa_graph* g, void* stream, const char** error);
void ga_wfa_graph_destroy(ga_wfa_graph* g);

/* ---- learned router (EXTENSION: no reference counterpart) ------------------
 * Sketch of a pair: GA_SKETCH_DIM floats of 4-mer / length features and the
 * exact lower bound max(|m - n|, ceil(||Q4(P) - Q4(T)||_1 / 8)) on its edit
 * distance.  The MLP (GA_SKETCH_DIM -> 128 -> 128 -> GA_ROUTE_BUCKETS, ReLU,
 * FP32) keeps its GA_MLP_PARAMS parameters in one device array:
 * W1[64][128], b1[128], W2[128][128], b2[128], W3[128][16], b3[16]. */
#define GA_SKETCH_DIM 64
#define GA_MLP_HIDDEN 128
#define GA_ROUTE_BUCKETS 16
#define GA_MLP_PARAMS 26896

int ga_sketch_device(const ga_wfa_dev_batch* b, float* feat, int32_t* lower_bound, void* stream,
                     const char** error);
/* h1, h2: n x 128 activations kept for the backward pass; logits n x 16 */
int ga_mlp_forward_device(const float* params, const float* feat, int64_t n, float* h1, float* h2,
                          float* logits, void* stream, const char** error);
/* Inference only: logits (n x 16) with the three layers fused and no
 * activations kept (the routed path's forward; sums by fused multiply-adds, so
 * the logits may differ from ga_mlp_forward_device's in the last bits).  No
 * counterpart in the reference: the learned router has none.  params must be
 * 16-byte aligned. */
int ga_mlp_logits_device(const float* params, const float* feat, int64_t n, float* logits, void* stream,
                         const char** error);
/* mean softmax cross-entropy over the n samples: grads (GA_MLP_PARAMS) and the
 * per-sample loss; work holds n * (16 + 128 + 128) floats */
int ga_mlp_backward_device(const float* params, const float* feat, const float* h1, const float* h2,
                           const float* logits, const int32_t* label, int64_t n, float* grads, float* loss,
                           float* work, void* stream, const char** error);
/* momentum <- mu * momentum + grads; params <- params - lr * momentum */
int ga_mlp_sgd_device(float* params, const float* grads, float* momentum, float lr, float mu, void* stream,
                      const char** error);
/* bucket = smallest bucket whose cumulative softmax probability reaches q
 * (q <= 0: arg-max) */
int ga_mlp_route_device(const float* logits, int64_t n, float q, int32_t* bucket, void* stream,
                        const char** error);
/* distance bound of a bucket (the last bucket: above smax, INT32_MAX) */
int32_t ga_route_bucket_bound(int32_t bucket);

typedef struct ga_route_stats {
    int64_t shortcut;   /* unresolved by the exact lower bound, not aligned */
    int64_t routed;     /* first tried under a bound below their smax */
    int64_t plain;      /* first tried under their own smax */
    int64_t rerun;      /* routed pairs not resolved under the bound: re-run with their smax */
} ga_route_stats;

/* ga_wfa_device with routing: same outputs pair for pair (the router only
 * chooses the first bound and size class; every pair not resolved under a
 * routed bound is re-run with its own smax).  Sketch, network, both passes and
 * their plans run on the device with no host synchronisation; with `stats`
 * non-NULL the call waits for the stream once at the end to report them. */
int ga_wfa_routed_device(const ga_wfa_dev_batch* b, const float* params, float q, ga_route_stats* stats,
                         void* stream, const char** error);

/* The library's router (EXTENSION; no reference counterpart).  With a network
 * loaded (GA_MLP_PARAMS host floats, copied to the device; q the routing
 * quantile as for ga_wfa_routed_device), ga_align_batch and
 * ga_align_batch_ext run their WFA batches through the routed path: same
 * outputs pair for pair, the network choosing each pair's first bound and
 * size class.  params NULL unloads it.  Every call keeps its own device copy
 * until the process ends (batches already running keep the network they
 * started with). */
int ga_router_set(const float* params, float q, const char** error);
int ga_router_enabled(void);

typedef struct ga_sw_dev_batch {
    const uint8_t* text;    const int64_t* text_off;    const int32_t* text_len;
    const uint8_t* pattern; const int64_t* pattern_off; const int32_t* pattern_len;
    int64_t        n_pairs;
    int32_t        max_text_len, max_pattern_len;
    ga_sw_scoring  scoring;
    int32_t        with_cigar;      /* 1: trace path (row-major end, span, CIGAR) */
    int32_t*       out;             /* out: 5 ints per pair: score, start_i, start_j, end_i, end_j */
    uint8_t*       cigar;           /* with_cigar only */
    const int64_t* cigar_off;
    int32_t*       cigar_len;       /* out */
} ga_sw_dev_batch;

int ga_sw_device(const ga_sw_dev_batch* b, void* stream, const char** error);
/* ga_sw_validate_cigar on the device for a batch ga_sw_device has filled
 * (with_cigar 1): pair k's CIGAR is the cigar_len[k] bytes ending its slot
 * (cigar + cigar_off[k] + text_len[k] + pattern_len[k]), checked against
 * out[5k .. 5k+5) and the sequences; flags[k] (device, n_pairs bytes) =
 * rescore_ok | wellformed_ok << 1 for pairs with score > 0 and a CIGAR, 0
 * for the others.  The checks ga_align_sw_batch (and, with the table,
 * ga_align_sw_matrix_batch) runs on its own results.
 * Enqueued on `stream`.  EXTENSION: no reference counterpart. */
int ga_sw_validate_device(const ga_sw_dev_batch* b, uint8_t* flags, void* stream, const char** error);

/* ---- kernel timing (EXTENSION: measurement support) ------------------------
 * When enabled (per calling thread), the batch entry points (ga_align_batch,
 * _ext, ga_align_sw_batch, ga_align_sw_matrix_batch) and the device ones
 * (ga_wfa_device, ga_wfa_routed_device, ga_sw_device) record HIP events on
 * their stream around their dominant kernels (the WFA class launches; the SW
 * forward kernels) and ga_last_kernel_ms() returns the summed duration of
 * those launches in the calling thread's last such call, waiting for it.
 * Returns a negative value when nothing was recorded. */
void  ga_set_kernel_timing(int enabled);
float ga_last_kernel_ms(void);

/* Caps the device workspace (wavefront stacks, direction tables) that batch
 * calls made from the calling thread hold at once; 0 restores the defaults.
 * Larger batches then run in more chunks.  (EXTENSION: for serving loops that
 * run several batches concurrently under one memory budget.) */
void ga_set_thread_workspace_budget(int64_t bytes);

/* Frees the idle device blocks that batch calls keep for reuse (large
 * workspaces: wavefront stacks, direction tables).  Calls still running keep
 * theirs.  Idle blocks are otherwise capped at 3/8 of each device's memory.
 * (EXTENSION: no counterpart in the reference, whose workspaces are host
 * memory.) */
void ga_trim_workspace(void);

/* ---- FASTA (reference: include/genoaligner/io/fasta.hpp:50-175) ------------
 * ga_fasta_parse / ga_fasta_read_file replace io::read_fasta /
 * io::read_fasta_file (fasta.hpp:47-53) with the same rules: raw bytes kept
 * (case, N, IUPAC codes), lines joined, \n, \r\n and a lone \r end a line,
 * blank lines skipped, the first header token is the id and the rest the
 * description.  A malformed file returns GA_INVALID_ARGUMENT and *error holds
 * the reference's message, e.g. "record has a header but no sequence (line 3)"
 * (valid until the next FASTA call on the thread).  The records stay valid
 * until ga_fasta_free. */
typedef struct ga_fasta ga_fasta;
int     ga_fasta_parse(const char* data, int64_t len, ga_fasta** out, const char** error);
int     ga_fasta_read_file(const char* path, ga_fasta** out, const char** error);
int64_t ga_fasta_count(const ga_fasta* f);
int     ga_fasta_record(const ga_fasta* f, int64_t k, const char** id, int64_t* id_len, const char** desc,
                        int64_t* desc_len, const char** seq, int64_t* seq_len);
void    ga_fasta_free(ga_fasta* f);
/* All records at once (EXTENSION: bulk access for bindings): ga_fasta_sizes
 * gives the total id, description and sequence bytes; ga_fasta_export writes
 * them back to back, record k at id_off[k] .. id_off[k+1] (and likewise), the
 * offset arrays holding count+1 entries.  Any pointer may be NULL. */
void    ga_fasta_sizes(const ga_fasta* f, int64_t* id_bytes, int64_t* desc_bytes, int64_t* seq_bytes);
int     ga_fasta_export(const ga_fasta* f, char* ids, int64_t* id_off, char* descs, int64_t* desc_off,
                        char* seqs, int64_t* seq_off);

/* FASTA ingest on the device (EXTENSION: no reference counterpart, whose reader
 * runs on the host only).  `data` holds a whole FASTA file
 * in device memory; its records' sequences are written back to back into
 * `seq` (same rules and bytes as ga_fasta_parse), record k at seq_off[k] with
 * seq_len[k] bytes, acgt[k] = 1 iff every byte is A, C, G or T (the 2-bit
 * packable records; the others keep their literal bytes).  When pack2 is not
 * NULL it receives the 2-bit codes of all of seq, 16 bases to a word, byte p of
 * seq at bits 2*(p%16) of pack2[p/16], A 0, C 1, G 2, T 3 (any other byte 0;
 * its record has acgt 0).  Buffers: seq >= len bytes, pack2 >= (len+15)/16
 * words, the per-record arrays >= max_records entries (the number of '>' bytes
 * bounds the records).  *n_records and *total (bytes in seq) return on the
 * host; the call waits for the stream.  A malformed file returns
 * GA_INVALID_ARGUMENT with ga_fasta_parse's message.  More than max_records
 * records returns GA_INVALID_ARGUMENT with the count needed in *n_records
 * (nothing written; call again with larger arrays); a record longer than
 * INT32_MAX bytes also returns GA_INVALID_ARGUMENT. */
typedef struct ga_fasta_dev_out {
    uint8_t*  seq;
    int64_t*  seq_off;
    int32_t*  seq_len;
    uint8_t*  acgt;
    uint32_t* pack2;      /* optional */
    int64_t   max_records;
} ga_fasta_dev_out;
int ga_fasta_device(const uint8_t* data, int64_t len, const ga_fasta_dev_out* out, int64_t* n_records,
                    int64_t* total, void* stream, const char** error);

/* ---- gene QC walk (reference: tools/gene_qc.cpp:146-217) ------------------
 * The per-record analysis of the reference's gene_qc tool, over the results
 * and CIGAR arena of a ga_align_sw_batch call whose requests align record k
 * (text) against the bait (pattern): text coverage of the span, indel runs
 * with length % 3 != 0, the net indel length mod 3 (-1, 0, +1), in-frame
 * stops of the bait's codons in `frame` (code 1 standard, 2 vertebrate
 * mitochondrial, any other TAA/TAG only), the span's terminal stop reported
 * apart.  Records with
 * resolved == 0 or score <= 0 get an all-zero row.  text/text_off/text_len
 * are the records' bytes as given to the batch.  Runs on the host. */
typedef struct ga_qc_row {
    double  cov_text;
    int32_t fs_runs;
    int32_t fs_net;
    int32_t stops;
    int32_t stop_end;
} ga_qc_row;
int ga_gene_qc_walk(const ga_sw_res* res, int64_t n, const uint8_t* cigar_arena, const uint8_t* text,
                    const int64_t* text_off, const int32_t* text_len, int32_t frame, int32_t code,
                    ga_qc_row* rows);
/* The tool's TSV (reference: tools/gene_qc.cpp:168, 239-262): the header and
 * one row per record -- id (record k's bytes ids[id_off[k] .. id_off[k+1]),
 * up to a NUL byte), text length, score, si, sj, ei, ej, cov_text (%.3f),
 * fs_runs, fs_net, stops, stop_end, flag (NO_ALIGN when unresolved or
 * score <= min_score; else LOW_COV when cov_text < min_cov, +FS=net when
 * fs_net != 0, +STOP=n when stops > 0, or OK) -- written into out when
 * out_cap holds it.  Returns the TSV's length in bytes (nothing written when
 * out is NULL or out_cap is smaller), or -1 for invalid arguments.  counts
 * (6) receives the summary line's OK, NO_ALIGN, LOW_COV, FS, STOP and invalid
 * counts (gene_qc.cpp:255-266).  EXTENSION (the reference tool formats inside
 * its main).  Runs on the host's worker threads. */
int64_t ga_gene_qc_tsv(const ga_sw_res* res, const ga_qc_row* rows, int64_t n, const char* ids,
                       const int64_t* id_off, const int32_t* text_len, double min_cov, int32_t min_score,
                       char* out, int64_t out_cap, int64_t* counts);

/* ---- MSA building blocks (reference: src/msa/msa_ref.cpp, the genomsa
 * engine; include/genoaligner/backend/msa_pp_kernel_impl.hip,
 * nj_kernel_impl.hip) ----------------------------------------------------- */

/* genomsa::Params fields of the profile DP (include/genoaligner/msa/msa.hpp:
 * 36-60): alpha 4 (IUPAC DNA: match / ts / tv scores) or 20 (amino acids:
 * sub[a * 20 + b], BLOSUM order ARNDCQEGHILKMFPSTWYV). */
typedef struct ga_msa_params {
    float   match, ts, tv;
    float   gap_open, gap_extend;
    int32_t free_end_gaps;
    int32_t psgp;
    float   psgp_scale, psgp_min_open, psgp_min_ext;
    int32_t alpha;
    float   sub[20 * 20];
} ga_msa_params;

/* A profile on the host: len columns of alpha symbol fractions
 * (cols[c * alpha + s]) and their occupancy (non-gap fraction). */
typedef struct ga_msa_profile {
    const float* cols;
    const float* occ;
    int32_t      len;
} ga_msa_profile;

/* genomsa::AlignResult (msa.hpp:339-344): score, aligned span (first
 * consumed column of A / B, ends exclusive) and the column CIGAR over
 * {M, I (A column only), D (B column only)} at arena[cigar_off ..). */
typedef struct ga_msa_pp_res {
    float   score;
    int32_t ai, aj, bi, bj;
    int64_t cigar_off;
    int32_t cigar_len;
} ga_msa_pp_res;

/* Replaces genomsa::align_profiles (msa.hpp:345, msa_ref.cpp:1166-1169) for
 * a batch of profile pairs -- the reference's msa_pp_trace_kernel launch
 * (msa_pp_kernel_impl.hip:35-209) -- with the same float results bit for
 * bit.  The arena needs sum(A.len + B.len) bytes.  Refused with
 * GA_INVALID_ARGUMENT: alpha other than 4 or 20, a negative length, a null
 * array with a positive length, a small arena. */
int ga_msa_align_profiles(const ga_msa_profile* A, const ga_msa_profile* B, int64_t n, const ga_msa_params* p,
                          ga_msa_pp_res* res, char* cigar_arena, int64_t arena_bytes, const char** error);

/* genomsa::profile_from_seq / profile_update_counts (msa.hpp:322-325,
 * msa_ref.cpp:751-794): the profile of nrows aligned rows of L bytes (row r
 * at rows + r * L); IUPAC letters (alpha 4) and B / Z / J / X / O / U (alpha
 * 20) count fractionally, '-' and unknown letters as gap.  gap (optional)
 * receives the gap fraction the reference keeps in its slot `alpha`. */
int ga_msa_profile_from_rows(const char* rows, int32_t nrows, int32_t L, int32_t alpha, float* cols, float* occ,
                             float* gap);

/* Device-resident batch (EXTENSION: no reference counterpart): profiles in
 * one column table on the device (profile k at columns col_off[k] ..
 * + len[k]), pair p aligns a_idx[p] against b_idx[p]; results to device
 * arrays (span: ai, aj, bi, bj per pair; CIGAR right-aligned in a slot of
 * len(A) + len(B) bytes at cigar_off[p]).  host_len / host_a_idx / host_b_idx
 * are host copies the library sizes its workspace with. */
typedef struct ga_msa_dev_batch {
    const float*   cols;          /* n_cols x alpha */
    const float*   occ;
    int64_t        n_cols;
    const int64_t* col_off;
    const int32_t* len;
    int32_t        n_profiles;
    const int32_t* a_idx;
    const int32_t* b_idx;
    int64_t        n_pairs;
    float*         score;
    int32_t*       span;
    uint8_t*       cigar;
    const int64_t* cigar_off;
    int32_t*       cigar_len;
    const int32_t* host_len;
    const int32_t* host_a_idx;
    const int32_t* host_b_idx;
} ga_msa_dev_batch;
int ga_msa_pp_device(const ga_msa_dev_batch* b, const ga_msa_params* p, void* stream, const char** error);

/* Replaces genomsa::kmer_distances / kmer_distances_mt (msa.hpp:329-333,
 * msa_ref.cpp:802-898): the packed upper triangle (row i, then j > i, index
 * i*n - i*(i+1)/2 + j - i - 1) of D = 1 - |K_i & K_j| / min(|K_i|, |K_j|)
 * (1 when the smaller set is empty) over the sets of k-mers of unambiguous
 * letters, sequence s at seqs[off[s] .. + len[s]).  The same floats bit for
 * bit.  Limits: alpha 4 or 20, alpha^k <= 2^24, n <= 65,535. */
int ga_kmer_distances(const char* seqs, const int64_t* off, const int32_t* len, int32_t n, int32_t k,
                      int32_t alpha, float* D, const char** error);

/* Replaces genomsa::nj_tree / nj_tree_gpu (msa.hpp:341-357, msa_ref.cpp:
 * 904-947, nj_kernel_impl.hip:84-187): deterministic neighbour joining over
 * the packed distances, the same tree: children[2*(u-n)] and
 * children[2*(u-n)+1] are the child node ids of internal node u, n <= u <=
 * 2n-2 (leaves 0 .. n-1, root 2n-2).  n <= 65,535; n < 2 writes nothing. */
int ga_nj_tree(const float* D, int32_t n, int32_t* children, const char** error);

#ifdef __cplusplus
}
#endif
#endif /* GENOALIGNER_C_H */
