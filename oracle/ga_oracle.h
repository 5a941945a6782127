/*
 * ga_oracle.h -- CPU restatement of genoaligner's pairwise alignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (genoaligner_amd) never links or calls it and fails loudly when its
 * HIP library is missing.
 *
 * Every function restates a named piece of the reference (paths relative to the
 * reference repository root) in plain C99, with the same integer arithmetic,
 * sentinels and tie rules, so that results are bit-identical:
 *
 *   WFA edit distance + CIGAR  include/genoaligner/backend/wfa_kernel.hip:103-145
 *                              (offsets, compute, extend), :467-592 (forward pass
 *                              retaining every wavefront), :297-437 (backtrace)
 *   Levenshtein DP             src/reference/edit_distance_cpu.hpp:19-39
 *   WFA CIGAR validation       src/api/api.cpp:84-106
 *   SW trace (Gotoh + walk)    include/genoaligner/backend/sw_kernel_impl.hip:198-336
 *   SW score-only end rule     include/genoaligner/backend/sw_kernel_impl.hip:22-168
 *                              (per-lane first strict max, xor-butterfly argmax)
 *   SW value walk              tests/sw/test_sw_trace.cpp:68-136
 *   SW CIGAR validation        src/api/api.cpp:114-156
 *
 * Parity pins (what checks each function against something other than itself)
 * are listed in tests/test_oracle.py; every function here has at least one.
 *
 * CIGARs are returned as forward strings over {M,X,I,D} (I consumes text, D
 * consumes pattern), exactly as src/api/api.cpp:71-78 decodes them.
 */
#ifndef GA_ORACLE_H
#define GA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- WFA (unit-cost edit distance bounded by smax) ------------------------ */

/* Score + CIGAR for one pair, following wfa_trace_kernel.  `smax` may be any
 * value >= 0 (the reference API caps it at 511; the oracle does not).
 * Returns 0 on success, -1 on allocation failure.
 *   *score      edit distance, or -1 when unresolved within smax
 *   cigar       forward CIGAR chars (capacity >= m + n); may be NULL
 *   *cigar_len  number of chars written (0 when unresolved or both empty)   */
int ga_oracle_wfa(const char* pattern, int m, const char* text, int n, int smax,
                  int* score, char* cigar, int* cigar_len);

/* Score only (wfa_score_kernel_flat: same recurrence, same termination). */
int ga_oracle_wfa_score(const char* pattern, int m, const char* text, int n, int smax);

/* Classic O(mn) Levenshtein distance (edit_distance_cpu). */
int ga_oracle_levenshtein(const char* pattern, int m, const char* text, int n);

/* validate(): rescore_ok = (#X+#I+#D == score); wellformed_ok = consumes exactly
 * m and n, M only on equal bytes, X only on unequal bytes. */
void ga_oracle_wfa_validate(const char* cigar, int len, int score,
                            const char* pattern, int m, const char* text, int n,
                            int* rescore_ok, int* wellformed_ok);

/* ---- Smith-Waterman, affine gaps -------------------------------------------
 * `sub` is a 256x256 row-major substitution table indexed [text byte][pattern
 * byte]; ga_oracle_sw_table_mm() fills it for the reference's match/mismatch
 * scheme (s = a == b ? match : mismatch, sw_kernel_impl.hip:308).            */
void ga_oracle_sw_table_mm(int match, int mismatch, int32_t* sub);

/* sw_trace_kernel + sw_trace_walk for one pair.  text is the row axis (m),
 * pattern the column axis (n).  Returns 0 on success, -1 on allocation failure.
 * Outputs follow SWTraceResult: score 0 -> coords all -1 and empty CIGAR.   */
int ga_oracle_sw_trace(const char* text, int m, const char* pattern, int n,
                       const int32_t* sub, int gap_open, int gap_extend,
                       int* score, int* start_i, int* start_j, int* end_i, int* end_j,
                       char* cigar, int* cigar_len);

/* sw_score_kernel's (score, end_i, end_j) for one pair at warp width `wsize`
 * (64 on AMD).  The end coordinate follows the kernel's per-lane first strict
 * maximum and its xor-butterfly argmax, which keeps a lane's own value on ties.
 * Empty side -> (0, -1, -1). */
void ga_oracle_sw_score_lane(const char* text, int m, const char* pattern, int n,
                             const int32_t* sub, int gap_open, int gap_extend, int wsize,
                             int* score, int* end_i, int* end_j);

/* Independent value walk over full H/E/F matrices (test_sw_trace.cpp:68-136). */
int ga_oracle_sw_value_walk(const char* text, int m, const char* pattern, int n,
                            const int32_t* sub, int gap_open, int gap_extend,
                            int* score, int* start_i, int* start_j, int* end_i, int* end_j,
                            char* cigar, int* cigar_len);

/* sw_validate(): affine re-score over the span and well-formedness. */
void ga_oracle_sw_validate(const char* cigar, int len,
                           const char* text, int text_len, int start_i, int end_i,
                           const char* pattern, int pattern_len, int start_j, int end_j,
                           const int32_t* sub, int gap_open, int gap_extend,
                           int score, int* rescore_ok, int* wellformed_ok);

/* ---- Batch helpers used as the timed CPU baseline --------------------------
 * Run the per-pair functions over a packed batch on `threads` POSIX threads.
 * Sequences are concatenated in `pat`/`txt` with int64 offsets.  Outputs are
 * per pair; CIGARs are written into `cigar` at cig_off[i] (capacity m+n).   */
int ga_oracle_wfa_batch(const char* pat, const int64_t* pat_off, const int32_t* pat_len,
                        const char* txt, const int64_t* txt_off, const int32_t* txt_len,
                        const int32_t* smax, const uint8_t* want_cigar, int64_t n_pairs,
                        int32_t* score, char* cigar, const int64_t* cig_off, int32_t* cig_len,
                        int threads);

int ga_oracle_sw_batch(const char* txt, const int64_t* txt_off, const int32_t* txt_len,
                       const char* pat, const int64_t* pat_off, const int32_t* pat_len,
                       int64_t n_pairs, const int32_t* sub, int gap_open, int gap_extend,
                       int with_cigar, int32_t* out5 /* score,si,sj,ei,ej per pair */,
                       char* cigar, const int64_t* cig_off, int32_t* cig_len, int threads);

#ifdef __cplusplus
}
#endif
#endif /* GA_ORACLE_H */
