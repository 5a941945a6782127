/* ga_msa_oracle.h -- CPU restatement of the reference's MSA building blocks
 * (src/msa/msa_ref.cpp) used as the checker for the profile-profile Gotoh
 * kernel and the k-mer distance / neighbour-joining kernels.
 * TEST INFRASTRUCTURE ONLY: nothing in the product links or calls it. */
#ifndef GA_MSA_ORACLE_H
#define GA_MSA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* genomsa::Params fields the DP reads (msa.hpp:36-60). */
typedef struct ga_oracle_msa_params {
    float match, ts, tv;          /* alpha == 4 substitution scores */
    float gap_open, gap_extend;
    int   free_end_gaps;
    int   psgp;
    float psgp_scale, psgp_min_open, psgp_min_ext;
    int   alpha;                  /* 4 (IUPAC DNA) or 20 (amino acids) */
    const float* sub;             /* alpha x alpha, row-major, alpha > 4 */
} ga_oracle_msa_params;

/* profile_update_counts over `nrows` aligned rows of L bytes each (row r at
 * rows + r*L): cols[c*alpha + s] = fraction of symbol s in column c,
 * occ[c] = non-gap fraction. */
int ga_oracle_msa_profile(const char* rows, int nrows, int L, int alpha, float* cols, float* occ);

/* align_profiles (semiglobal / global Gotoh, profile columns).  cols are
 * L x alpha as above.  span = {ai, aj, bi, bj}; cigar receives M/I/D (I = A
 * column only, D = B column only), capacity M + N. */
int ga_oracle_msa_align(const float* colsA, const float* occA, int M, const float* colsB, const float* occB,
                        int N, const ga_oracle_msa_params* P, float* score, int32_t* span, char* cigar,
                        int32_t* cigar_len);

/* Many pairs on `threads` threads: pair p aligns profile a_idx[p] against
 * b_idx[p] of a profile table (cols at col_off[k]*alpha, occ at col_off[k],
 * len[k] columns); cigar at cig_off[p] with capacity len(A)+len(B). */
int ga_oracle_msa_align_batch(const float* cols, const float* occ, const int64_t* col_off, const int32_t* len,
                              const int32_t* a_idx, const int32_t* b_idx, int64_t n_pairs,
                              const ga_oracle_msa_params* P, float* score, int32_t* span, char* cigar,
                              const int64_t* cig_off, int32_t* cig_len, int threads);

/* kmer_distances: D[idx(i,j)] (packed upper triangle, row i then j > i) =
 * 1 - |K_i & K_j| / min(|K_i|, |K_j|) (1 when the smaller set is empty). */
int ga_oracle_kmer_distances(const char* seqs, const int64_t* off, const int32_t* len, int n, int k, int alpha,
                             float* D);

/* nj_tree over the packed distances: children[2*(u-n)], children[2*(u-n)+1]
 * = the child node ids of internal node u (n <= u <= 2n-2; the root is
 * 2n-2).  n >= 2. */
int ga_oracle_nj_tree(const float* D, int n, int32_t* children);

#ifdef __cplusplus
}
#endif
#endif
