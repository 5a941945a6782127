/* ga_msa_oracle.c -- CPU restatement of the reference's MSA building blocks
 * (src/msa/msa_ref.cpp), the checker of the profile-profile Gotoh kernel and
 * of the k-mer distance and neighbour-joining kernels.
 * TEST INFRASTRUCTURE ONLY.  Built with -ffp-contract=off: every float and
 * double operation below is evaluated in the reference's order, unfused.
 *
 *   sym_counts / iupac_counts / aa_index   msa_ref.cpp:21-106
 *   sym_index                              msa_ref.cpp:109-114
 *   profile_update_counts                  msa_ref.cpp:773-794
 *   psgp_open / psgp_ext                   include/genoaligner/msa/msa.hpp:123-134
 *   sub_score / col_score                  msa_ref.cpp:1023-1047
 *   align_profiles_diag (no band)          msa_ref.cpp:1063-1164
 *   kmer_distances                         msa_ref.cpp:802-841
 *   nj_tree                                msa_ref.cpp:904-947
 */
#include "ga_msa_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MSA_NEG (-1e30f)

/* ---------------------------------------------------------------- symbols */
static int iupac_counts(char c, float cnt[4])
{
    int mask = 0;
    cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0.f;
    switch (c) {
    case 'A': case 'a': mask = 1; break;
    case 'C': case 'c': mask = 2; break;
    case 'G': case 'g': mask = 4; break;
    case 'T': case 't': case 'U': case 'u': mask = 8; break;
    case 'R': mask = 1 | 4; break;
    case 'Y': mask = 2 | 8; break;
    case 'M': mask = 1 | 2; break;
    case 'K': mask = 4 | 8; break;
    case 'S': mask = 2 | 4; break;
    case 'W': mask = 1 | 8; break;
    case 'H': mask = 1 | 2 | 8; break;
    case 'B': mask = 2 | 4 | 8; break;
    case 'V': mask = 1 | 2 | 4; break;
    case 'D': mask = 1 | 4 | 8; break;
    case 'N': case 'n': case 'X': mask = 15; break;
    default: return 0;
    }
    const float f = 1.0f / (float)__builtin_popcount((unsigned)mask);
    for (int b = 0; b < 4; ++b)
        if (mask & (1 << b)) cnt[b] = f;
    return 1;
}

static int aa_index(char c)
{
    static const char up[] = "ARNDCQEGHILKMFPSTWYV";
    for (int i = 0; i < 20; ++i)
        if (c == up[i] || c == up[i] + 32) return i;
    return -1;
}

static int sym_counts(char c, float* cnt, int alpha)
{
    if (alpha == 4) return iupac_counts(c, cnt);
    for (int i = 0; i < alpha; ++i) cnt[i] = 0.f;
    switch (c) {
    case 'B': cnt[2] = cnt[3] = 0.5f; return 1;
    case 'Z': cnt[5] = cnt[6] = 0.5f; return 1;
    case 'J': cnt[9] = cnt[10] = 0.5f; return 1;
    case 'U': case 'u': cnt[4] = 1.f; return 1;
    case 'X': case 'x': case 'O': case 'o':
        for (int i = 0; i < alpha; ++i) cnt[i] = 1.f / (float)alpha;
        return 1;
    default: break;
    }
    const int s = aa_index(c);
    if (s < 0) return 0;
    cnt[s] = 1.f;
    return 1;
}

static int sym_index(char c, int alpha)
{
    float cnt[20];
    if (!sym_counts(c, cnt, alpha)) return -1;
    for (int i = 0; i < alpha; ++i)
        if (cnt[i] == 1.f) return i;
    return -1;
}

int ga_oracle_msa_profile(const char* rows, int nrows, int L, int alpha, float* cols, float* occ)
{
    if (alpha != 4 && alpha != 20) return -1;
    for (int i = 0; i < L; ++i) {
        for (int b = 0; b < alpha; ++b) cols[(size_t)i * alpha + b] = 0.f;
        occ[i] = 0.f;
    }
    for (int r = 0; r < nrows; ++r)
        for (int i = 0; i < L; ++i) {
            float cnt[20];
            if (sym_counts(rows[(size_t)r * L + i], cnt, alpha)) {
                for (int b = 0; b < alpha; ++b) cols[(size_t)i * alpha + b] += cnt[b];
                occ[i] += 1.f;
            }
        }
    for (int i = 0; i < L; ++i) {
        for (int b = 0; b < alpha; ++b) cols[(size_t)i * alpha + b] /= (float)nrows;
        occ[i] /= (float)nrows;
    }
    return 0;
}

/* -------------------------------------------------------------- scoring */
static float psgp_open(float occ, const ga_oracle_msa_params* P)
{
    if (!P->psgp || occ >= 1.0f) return P->gap_open * occ;
    const float p = P->psgp_scale * P->gap_open * occ;
    const float fl = P->psgp_min_open * P->gap_open;
    return p > fl ? p : fl;
}

static float psgp_ext(float occ, const ga_oracle_msa_params* P)
{
    if (!P->psgp || occ >= 1.0f) return P->gap_extend;
    const float p = P->gap_extend * occ;
    const float fl = P->psgp_min_ext * P->gap_extend;
    return p > fl ? p : fl;
}

static float sub_score(int a, int b, const ga_oracle_msa_params* P)
{
    if (a == b) return P->match;
    const int ts = (a == 0 && b == 2) || (a == 2 && b == 0) || (a == 1 && b == 3) || (a == 3 && b == 1);
    return ts ? P->ts : P->tv;
}

static float col_score(const float* ca, const float* cb, const ga_oracle_msa_params* P)
{
    const int al = P->alpha;
    float s = 0.f;
    for (int a = 0; a < al; ++a) {
        const float fa = ca[a];
        if (fa == 0.f) continue;
        for (int b = 0; b < al; ++b) {
            const float fb = cb[b];
            if (fb == 0.f) continue;
            s += fa * fb * (al == 4 ? sub_score(a, b, P) : P->sub[a * al + b]);
        }
    }
    return s;
}

/* ------------------------------------------------------ semiglobal Gotoh */
int ga_oracle_msa_align(const float* colsA, const float* occA, int M, const float* colsB, const float* occB,
                        int N, const ga_oracle_msa_params* P, float* score, int32_t* span, char* cigar,
                        int32_t* cigar_len)
{
    const int al = P->alpha;
    const size_t W = (size_t)N + 1, cells = ((size_t)M + 1) * W;
    float* mM = (float*)malloc(cells * sizeof(float));
    float* mIx = (float*)malloc(cells * sizeof(float));
    float* mIy = (float*)malloc(cells * sizeof(float));
    unsigned char* dir = (unsigned char*)calloc(cells, 1);
    if (!mM || !mIx || !mIy || !dir) {
        free(mM); free(mIx); free(mIy); free(dir);
        return -1;
    }
    for (size_t c = 0; c < cells; ++c) mM[c] = mIx[c] = mIy[c] = MSA_NEG;
    mM[0] = 0.f;
    for (int i = 1; i <= M; ++i) {
        mIx[i * W] = P->free_end_gaps ? 0.f : -(psgp_open(occA[i - 1], P) + (float)(i - 1) * P->gap_extend);
        mM[i * W] = P->free_end_gaps ? 0.f : MSA_NEG;
    }
    for (int j = 1; j <= N; ++j) {
        mIy[j] = P->free_end_gaps ? 0.f : -(psgp_open(occB[j - 1], P) + (float)(j - 1) * P->gap_extend);
        mM[j] = P->free_end_gaps ? 0.f : MSA_NEG;
    }
    for (int i = 1; i <= M; ++i) {
        for (int j = 1; j <= N; ++j) {
            const size_t c = (size_t)i * W + (size_t)j;
            const float s = col_score(colsA + (size_t)(i - 1) * al, colsB + (size_t)(j - 1) * al, P);
            const float openB = psgp_open(occA[i - 1], P), extB = psgp_ext(occA[i - 1], P);
            const float openA = psgp_open(occB[j - 1], P), extA = psgp_ext(occB[j - 1], P);
            const float oIx = mM[c - W] - openB, eIx = mIx[c - W] - extB;
            const float oIy = mM[c - 1] - openA, eIy = mIy[c - 1] - extA;
            mIx[c] = oIx < eIx ? eIx : oIx;           /* std::max(oIx, eIx) */
            mIy[c] = oIy < eIy ? eIy : oIy;
            float mx = mM[c - W - 1];
            unsigned char h = 0;
            if (mIx[c - W - 1] > mx) { mx = mIx[c - W - 1]; h = 1; }
            if (mIy[c - W - 1] > mx) { mx = mIy[c - W - 1]; h = 2; }
            mM[c] = s + mx;
            dir[c] = (unsigned char)(h | (eIx > oIx ? 4 : 0) | (eIy > oIy ? 8 : 0));
        }
    }
    int ei = M, ej = N;
    float best = mM[(size_t)M * W + (size_t)N];
    char est = 'M';
    if (P->free_end_gaps) {
        best = MSA_NEG;
        for (int i = 0; i <= M; ++i) {
            const size_t c = (size_t)i * W + (size_t)N;
            if (mM[c] > best) { best = mM[c]; ei = i; ej = N; est = 'M'; }
            if (mIx[c] > best) { best = mIx[c]; ei = i; ej = N; est = 'I'; }
        }
        for (int j = 0; j <= N; ++j) {
            const size_t c = (size_t)M * W + (size_t)j;
            if (mM[c] > best) { best = mM[c]; ei = M; ej = j; est = 'M'; }
            if (mIy[c] > best) { best = mIy[c]; ei = M; ej = j; est = 'D'; }
        }
    }
    int i = ei, j = ej, used = 0;
    char st = est;
    while (i > 0 || j > 0) {
        const size_t c = (size_t)i * W + (size_t)j;
        if (st == 'M') {
            if (i == 0 || j == 0) break;
            cigar[used++] = 'M';
            const int h = dir[c] & 3;
            --i; --j;
            st = h == 0 ? 'M' : (h == 1 ? 'I' : 'D');
        } else if (st == 'I') {
            if (i == 0) break;
            cigar[used++] = 'I';
            if (j > 0) st = (dir[c] & 4) ? 'I' : 'M';
            --i;
        } else {
            if (j == 0) break;
            cigar[used++] = 'D';
            if (i > 0) st = (dir[c] & 8) ? 'D' : 'M';
            --j;
        }
    }
    for (int a = 0, b = used - 1; a < b; ++a, --b) {
        const char t = cigar[a];
        cigar[a] = cigar[b];
        cigar[b] = t;
    }
    *score = best;
    span[0] = i; span[1] = j; span[2] = ei; span[3] = ej;
    *cigar_len = used;
    free(mM); free(mIx); free(mIy); free(dir);
    return 0;
}

typedef struct {
    const float *cols, *occ;
    const int64_t* col_off;
    const int32_t *len, *a_idx, *b_idx;
    int64_t n;
    const ga_oracle_msa_params* P;
    float* score;
    int32_t* span;
    char* cigar;
    const int64_t* cig_off;
    int32_t* cig_len;
    int64_t next;
    pthread_mutex_t mu;
    int err;
} msa_job;

static void* msa_worker(void* arg)
{
    msa_job* J = (msa_job*)arg;
    const int al = J->P->alpha;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const int64_t p = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (p >= J->n) break;
        const int a = J->a_idx[p], b = J->b_idx[p];
        if (ga_oracle_msa_align(J->cols + J->col_off[a] * al, J->occ + J->col_off[a], J->len[a],
                                J->cols + J->col_off[b] * al, J->occ + J->col_off[b], J->len[b], J->P,
                                J->score + p, J->span + 4 * p, J->cigar + J->cig_off[p], J->cig_len + p) != 0)
            J->err = 1;
    }
    return NULL;
}

int ga_oracle_msa_align_batch(const float* cols, const float* occ, const int64_t* col_off, const int32_t* len,
                              const int32_t* a_idx, const int32_t* b_idx, int64_t n_pairs,
                              const ga_oracle_msa_params* P, float* score, int32_t* span, char* cigar,
                              const int64_t* cig_off, int32_t* cig_len, int threads)
{
    msa_job J = {cols, occ, col_off, len, a_idx, b_idx, n_pairs, P, score, span, cigar, cig_off, cig_len, 0,
                 PTHREAD_MUTEX_INITIALIZER, 0};
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, msa_worker, &J);
    msa_worker(&J);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    return J.err ? -1 : 0;
}

/* ------------------------------------------------------------ distances */
static int cmp_u64(const void* x, const void* y)
{
    const uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
    return a < b ? -1 : a > b;
}

int ga_oracle_kmer_distances(const char* seqs, const int64_t* off, const int32_t* len, int n, int k, int alpha,
                             float* D)
{
    uint64_t kbase = 1;
    for (int i = 0; i < k; ++i) kbase *= (uint64_t)alpha;
    uint64_t** set = (uint64_t**)calloc((size_t)n, sizeof(uint64_t*));
    int64_t* cnt = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    if (!set || !cnt) return -1;
    for (int s = 0; s < n; ++s) {
        set[s] = (uint64_t*)malloc(((size_t)(len[s] > 0 ? len[s] : 0) + 1) * sizeof(uint64_t));
        const char* q = seqs + off[s];
        uint64_t h = 0;
        int run = 0;
        int64_t c = 0;
        for (int i = 0; i < len[s]; ++i) {
            const int b = sym_index(q[i], alpha);
            if (b < 0) { run = 0; h = 0; continue; }
            h = (h * (uint64_t)alpha + (uint64_t)b) % kbase;
            if (++run >= k) set[s][c++] = h;
        }
        qsort(set[s], (size_t)c, sizeof(uint64_t), cmp_u64);
        int64_t u = 0;
        for (int64_t i = 0; i < c; ++i)
            if (u == 0 || set[s][u - 1] != set[s][i]) set[s][u++] = set[s][i];
        cnt[s] = u;
    }
    size_t at = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            int64_t x = 0, y = 0, inter = 0;
            while (x < cnt[i] && y < cnt[j]) {
                if (set[i][x] < set[j][y]) ++x;
                else if (set[i][x] > set[j][y]) ++y;
                else { ++inter; ++x; ++y; }
            }
            const int64_t mn = cnt[i] < cnt[j] ? cnt[i] : cnt[j];
            D[at++] = mn ? 1.0f - (float)inter / (float)mn : 1.0f;
        }
    for (int s = 0; s < n; ++s) free(set[s]);
    free(set);
    free(cnt);
    return 0;
}

/* ---------------------------------------------------------------- NJ */
int ga_oracle_nj_tree(const float* Dp, int n, int32_t* children)
{
    if (n < 2) return -1;
    const int NN = 2 * n - 1;
    double* d = (double*)calloc((size_t)NN * NN, sizeof(double));
    int* alive = (int*)malloc((size_t)n * sizeof(int));
    int* rest = (int*)malloc((size_t)n * sizeof(int));
    double* r = (double*)malloc((size_t)n * sizeof(double));
    if (!d || !alive || !rest || !r) return -1;
    size_t at = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j, ++at)
            d[(size_t)i * NN + j] = d[(size_t)j * NN + i] = Dp[at];
    int m = n, u = n;
    for (int i = 0; i < n; ++i) alive[i] = i;
    while (m > 2) {
        for (int a = 0; a < m; ++a) {
            r[a] = 0;
            for (int b = 0; b < m; ++b)
                if (a != b) r[a] += d[(size_t)alive[a] * NN + alive[b]];
        }
        double best = 0;
        int bi = -1, bj = -1;
        for (int a = 0; a < m; ++a)
            for (int b = a + 1; b < m; ++b) {
                const double q = (m - 2) * d[(size_t)alive[a] * NN + alive[b]] - r[a] - r[b];
                if (bi < 0 || q < best) { best = q; bi = a; bj = b; }
            }
        const int xi = alive[bi], xj = alive[bj];
        children[2 * (u - n)] = xi;
        children[2 * (u - n) + 1] = xj;
        int nr = 0;
        for (int a = 0; a < m; ++a)
            if (a != bi && a != bj) rest[nr++] = alive[a];
        for (int t = 0; t < nr; ++t) {
            const int v = rest[t];
            d[(size_t)u * NN + v] = d[(size_t)v * NN + u] =
                0.5 * (d[(size_t)xi * NN + v] + d[(size_t)xj * NN + v] - d[(size_t)xi * NN + xj]);
        }
        for (int t = 0; t < nr; ++t) alive[t] = rest[t];
        alive[nr] = u;
        m = nr + 1;
        ++u;
    }
    children[2 * (u - n)] = alive[0];
    children[2 * (u - n) + 1] = alive[1];
    free(d); free(alive); free(rest); free(r);
    return 0;
}
