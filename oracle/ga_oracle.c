/*
 * ga_oracle.c -- CPU restatement of genoaligner's pairwise alignment path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See ga_oracle.h for the contract and
 * the reference lines each function restates.  Plain C99 + POSIX threads.
 */
#define _POSIX_C_SOURCE 200809L
#include "ga_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ========================================================================== */
/* WFA                                                                         */
/* ========================================================================== */

/* wfa_kernel.hip:103 -- WAVEFRONT_OFFSET_NULL; :110 -- threshold reachability. */
#define WFA_NEG (INT32_MIN / 2)
static int wfa_reachable(int off) { return off > WFA_NEG + 1024; }

/* wfa_kernel.hip:119-127 -- compute-next for one diagonal. */
static int wfa_compute(int prev_ins, int prev_del, int prev_mism, int k, int m, int n)
{
    const int mid = prev_ins > prev_mism ? prev_ins : prev_mism;
    int mx = (prev_del > mid + 1) ? prev_del : mid + 1;
    if (!wfa_reachable(mx)) return WFA_NEG;
    if (mx > n) return WFA_NEG;          /* h = offset      */
    if (mx - k > m) return WFA_NEG;      /* v = offset - k  */
    return mx;
}

/* wfa_kernel.hip:130-136 -- greedy byte-wise extension along diagonal k. */
static int wfa_extend(const char* pattern, const char* text, int k, int offset, int m, int n)
{
    if (!wfa_reachable(offset)) return offset;
    int v = offset - k, h = offset;
    while (v < m && h < n && pattern[v] == text[h]) { ++v; ++h; ++offset; }
    return offset;
}

/* The forward pass keeps every wavefront (wfa_kernel.hip:528-558).  Row s holds
 * diagonals k in [-s, s] at W[s*s + k + s]; reads outside that range are the
 * sentinel, exactly as the reference's NEG-initialised slots and its M() lambda
 * (wfa_kernel.hip:360-364) behave. */
typedef struct { int* w; long cap_rows; } WfStack;

static int wf_get(const WfStack* st, int s, int k)
{
    if (s < 0 || k < -s || k > s) return WFA_NEG;
    return st->w[(long)s * s + k + s];
}

static int wf_reserve(WfStack* st, int rows)
{
    if (rows <= st->cap_rows) return 0;
    long nr = st->cap_rows ? st->cap_rows : 16;
    while (nr < rows) nr *= 2;
    int* nw = (int*)realloc(st->w, (size_t)nr * (size_t)nr * sizeof(int));
    if (!nw) return -1;
    st->w = nw; st->cap_rows = nr;
    return 0;
}

/* Forward pass.  Returns the score (>= 0) or -1 when unresolved within smax;
 * *rows_out = number of wavefronts kept.  m > 0 and n > 0 here. */
static int wfa_forward(const char* pattern, int m, const char* text, int n, int smax,
                       WfStack* st, int* err)
{
    *err = 0;
    if (wf_reserve(st, 1)) { *err = 1; return -1; }
    int offset = 0;
    while (offset < m && offset < n && pattern[offset] == text[offset]) ++offset;
    st->w[0] = offset;
    const int alignment_k = n - m;
    if (alignment_k == 0 && offset >= n) return 0;
    for (int s = 1; s <= smax; ++s) {
        if (wf_reserve(st, s + 1)) { *err = 1; return -1; }
        for (int k = -s; k <= s; ++k) {
            const int ins  = wf_get(st, s - 1, k - 1);
            const int del  = wf_get(st, s - 1, k + 1);
            const int mism = wf_get(st, s - 1, k);
            const int mx = wfa_compute(ins, del, mism, k, m, n);
            st->w[(long)s * s + k + s] = wfa_extend(pattern, text, k, mx, m, n);
        }
        if (alignment_k >= -s && alignment_k <= s) {
            const int off = wf_get(st, s, alignment_k);
            if (wfa_reachable(off) && off >= n) return s;
        }
    }
    return -1;
}

/* Reversed op buffer, as the reference's cigar_rev (wfa_kernel.hip:297-302). */
typedef struct { char* rev; int used; } RevBuf;
static void rev_push(RevBuf* b, char op) { b->rev[b->used++] = op; }

/* wfa_kernel.hip:338-437 -- the backward walk, with its tie priority
 * (mismatch >=, then insertion >, then deletion >). */
static void wfa_traceback(const char* pattern, const char* text, int m, int n,
                          int score_total, const WfStack* st, RevBuf* out)
{
    const int alignment_k = n - m;
    int score = score_total, k = alignment_k;
    int offset = wf_get(st, score_total, k);
    if (offset == WFA_NEG) return;
    int h = offset, v = offset - k;
    /* trailing tail (wfa_kernel.hip:309-313) */
    for (int i = v - m; i > 0; --i) rev_push(out, 'D');
    for (int i = h - n; i > 0; --i) rev_push(out, 'I');
    if (v > m) v = m;
    if (h > n) h = n;
    while (v > 0 && h > 0 && score > 0) {
        const int sp = score - 1;
        const int pm = wf_get(st, sp, k), pi = wf_get(st, sp, k - 1), pd = wf_get(st, sp, k + 1);
        const int misms = pm + 1, ins = pi + 1, del = pd;
        int best = WFA_NEG, op = 0;
        if (wfa_reachable(misms) && misms >= best) { best = misms; op = 1; }
        if (wfa_reachable(ins)   && ins   >  best) { best = ins;   op = 2; }
        if (wfa_reachable(del)   && del   >  best) { best = del;   op = 3; }
        if (op == 0) break;
        for (int i = offset - best; i > 0; --i) rev_push(out, 'M');
        offset = best;
        if (op == 1)      { rev_push(out, 'X'); --offset; }
        else if (op == 2) { rev_push(out, 'I'); --k; --offset; }
        else              { rev_push(out, 'D'); ++k; }
        v = offset - k; h = offset;
        --score;
    }
    if (v > 0 && h > 0) {
        const int nm = v < h ? v : h;
        for (int i = 0; i < nm; ++i) rev_push(out, 'M');
        v -= nm; h -= nm;
    }
    while (v > 0) { rev_push(out, 'D'); --v; }
    while (h > 0) { rev_push(out, 'I'); --h; }
}

static void emit_forward(const RevBuf* b, char* cigar, int* cigar_len)
{
    if (cigar) for (int j = 0; j < b->used; ++j) cigar[j] = b->rev[b->used - 1 - j];
    *cigar_len = b->used;
}

int ga_oracle_wfa(const char* pattern, int m, const char* text, int n, int smax,
                  int* score, char* cigar, int* cigar_len)
{
    *cigar_len = 0;
    /* degenerate pair: score is the other length regardless of smax, CIGAR all
     * I (empty pattern) or all D (empty text) -- wfa_kernel.hip:480-502 */
    if (m <= 0 || n <= 0) {
        *score = (m <= 0) ? (n > 0 ? n : 0) : m;
        if (cigar) {
            int c = 0;
            for (int i = 0; i < n; ++i) cigar[c++] = 'I';
            for (int i = 0; i < m; ++i) cigar[c++] = 'D';
        }
        *cigar_len = (m > 0 ? m : 0) + (n > 0 ? n : 0);
        return 0;
    }
    WfStack st = {NULL, 0};
    int err = 0;
    const int s = wfa_forward(pattern, m, text, n, smax, &st, &err);
    if (err) { free(st.w); return -1; }
    *score = s;
    if (s < 0) { free(st.w); return 0; }
    RevBuf rb;
    rb.rev = (char*)malloc((size_t)m + (size_t)n + 4);
    rb.used = 0;
    if (!rb.rev) { free(st.w); return -1; }
    if (s == 0) {
        for (int i = 0; i < n; ++i) rev_push(&rb, 'M');   /* wfa_kernel.hip:563-574 */
    } else {
        wfa_traceback(pattern, text, m, n, s, &st, &rb);
    }
    emit_forward(&rb, cigar, cigar_len);
    free(rb.rev);
    free(st.w);
    return 0;
}

int ga_oracle_wfa_score(const char* pattern, int m, const char* text, int n, int smax)
{
    if (m <= 0 || n <= 0) return (m <= 0) ? (n > 0 ? n : 0) : m;   /* wfa_score_flat.hip:73-77 */
    WfStack st = {NULL, 0};
    int err = 0;
    const int s = wfa_forward(pattern, m, text, n, smax, &st, &err);
    free(st.w);
    return err ? -2 : s;
}

/* edit_distance_cpu.hpp:19-39 */
int ga_oracle_levenshtein(const char* pattern, int m, const char* text, int n)
{
    if (m == 0) return n;
    if (n == 0) return m;
    int* prev = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    int* cur  = (int*)malloc(sizeof(int) * (size_t)(n + 1));
    if (!prev || !cur) { free(prev); free(cur); return -2; }
    for (int j = 0; j <= n; ++j) prev[j] = j;
    for (int i = 1; i <= m; ++i) {
        cur[0] = i;
        for (int j = 1; j <= n; ++j) {
            const int sub = prev[j - 1] + (pattern[i - 1] == text[j - 1] ? 0 : 1);
            const int del = prev[j] + 1;
            const int ins = cur[j - 1] + 1;
            int v = sub < del ? sub : del;
            cur[j] = v < ins ? v : ins;
        }
        int* t = prev; prev = cur; cur = t;
    }
    const int r = prev[n];
    free(prev); free(cur);
    return r;
}

/* api.cpp:84-106 */
void ga_oracle_wfa_validate(const char* cigar, int len, int score,
                            const char* pattern, int m, const char* text, int n,
                            int* rescore_ok, int* wellformed_ok)
{
    int s = 0, pi = 0, ti = 0, wf = 1;
    for (int c = 0; c < len; ++c) {
        switch (cigar[c]) {
            case 'M':
                if (pi >= m || ti >= n || pattern[pi] != text[ti]) wf = 0;
                ++pi; ++ti; break;
            case 'X':
                if (pi >= m || ti >= n || pattern[pi] == text[ti]) wf = 0;
                ++pi; ++ti; ++s; break;
            case 'I': ++ti; ++s; break;
            case 'D': ++pi; ++s; break;
            default: wf = 0; break;
        }
    }
    if (pi != m || ti != n) wf = 0;
    *wellformed_ok = wf;
    *rescore_ok = (s == score);
}

/* ========================================================================== */
/* Smith-Waterman                                                              */
/* ========================================================================== */

void ga_oracle_sw_table_mm(int match, int mismatch, int32_t* sub)
{
    for (int a = 0; a < 256; ++a)
        for (int b = 0; b < 256; ++b) sub[a * 256 + b] = (a == b) ? match : mismatch;
}

#define SUB(t, p) sub[(int)(unsigned char)(t) * 256 + (int)(unsigned char)(p)]

/* direction byte fields, sw_kernel.hip:218-223 */
enum { DIR_STOP = 0, DIR_DIAG = 1, DIR_E = 2, DIR_F = 3, DIR_EEXT = 4, DIR_FEXT = 8 };

int ga_oracle_sw_trace(const char* text, int m, const char* pattern, int n,
                       const int32_t* sub, int gap_open, int gap_extend,
                       int* score, int* start_i, int* start_j, int* end_i, int* end_j,
                       char* cigar, int* cigar_len)
{
    *score = 0; *start_i = *start_j = *end_i = *end_j = -1; *cigar_len = 0;
    if (m <= 0 || n <= 0) return 0;                         /* :270-273 */
    const int rs = n + 1;
    uint8_t* d = (uint8_t*)malloc((size_t)(m + 1) * (size_t)rs);
    int* Hprev = (int*)malloc(sizeof(int) * (size_t)rs);
    int* Farr  = (int*)malloc(sizeof(int) * (size_t)rs);
    if (!d || !Hprev || !Farr) { free(d); free(Hprev); free(Farr); return -1; }
    for (int j = 0; j <= n; ++j) { Hprev[j] = 0; Farr[j] = 0; d[j] = DIR_STOP; }   /* :286-287 */
    for (int i = 0; i <= m; ++i) d[(size_t)i * rs] = DIR_STOP;
    const int go = gap_open, ge = gap_extend;
    int best = 0, bi = -1, bj = -1;
    for (int i = 1; i <= m; ++i) {                           /* :292-322 */
        const char a = text[i - 1];
        int hleft = 0, e_row = 0, diag = 0;
        uint8_t* drow = d + (size_t)i * rs;
        for (int j = 1; j <= n; ++j) {
            int e = hleft - go; uint8_t esrc = 0;
            { const int e2 = e_row - ge; if (e2 > e) { e = e2; esrc = DIR_EEXT; } }
            int f = Hprev[j] - go; uint8_t fsrc = 0;
            { const int f2 = Farr[j] - ge; if (f2 > f) { f = f2; fsrc = DIR_FEXT; } }
            int h = diag + SUB(a, pattern[j - 1]);
            uint8_t hsrc = DIR_DIAG;
            if (e > h) { h = e; hsrc = DIR_E; }
            if (f > h) { h = f; hsrc = DIR_F; }
            if (h <= 0) { h = 0; hsrc = DIR_STOP; }
            drow[j] = (uint8_t)(hsrc | esrc | fsrc);
            diag = Hprev[j]; Hprev[j] = h; Farr[j] = f; hleft = h; e_row = e;
            if (h > best) { best = h; bi = i; bj = j; }
        }
    }
    *score = best;
    if (best > 0) {
        /* sw_trace_walk, sw_kernel_impl.hip:198-242 */
        char* rev = (char*)malloc((size_t)m + (size_t)n + 4);
        if (!rev) { free(d); free(Hprev); free(Farr); return -1; }
        int used = 0, i = bi, j = bj, state = 0;
        while (i > 0 && j > 0) {
            const uint8_t b = d[(size_t)i * rs + j];
            if (state == 0) {
                const uint8_t hsrc = b & 3;
                if (hsrc == DIR_STOP) break;
                if (hsrc == DIR_DIAG) {
                    rev[used++] = (text[i - 1] == pattern[j - 1]) ? 'M' : 'X';
                    --i; --j;
                } else if (hsrc == DIR_E) {
                    rev[used++] = 'D';
                    state = (b & DIR_EEXT) ? 1 : 0;
                    --j;
                } else {
                    rev[used++] = 'I';
                    state = (b & DIR_FEXT) ? 2 : 0;
                    --i;
                }
            } else if (state == 1) {
                rev[used++] = 'D';
                --j;
                if (!(b & DIR_EEXT)) state = 0;
            } else {
                rev[used++] = 'I';
                --i;
                if (!(b & DIR_FEXT)) state = 0;
            }
        }
        *start_i = i; *start_j = j; *end_i = bi - 1; *end_j = bj - 1;
        if (cigar) for (int c = 0; c < used; ++c) cigar[c] = rev[used - 1 - c];
        *cigar_len = used;
        free(rev);
    }
    free(d); free(Hprev); free(Farr);
    return 0;
}

void ga_oracle_sw_score_lane(const char* text, int m, const char* pattern, int n,
                             const int32_t* sub, int gap_open, int gap_extend, int wsize,
                             int* score, int* end_i, int* end_j)
{
    *score = 0; *end_i = -1; *end_j = -1;
    if (m <= 0 || n <= 0) return;                            /* sw_kernel_impl.hip:47-50 */
    if (wsize <= 0 || wsize > 1024 || (wsize & (wsize - 1))) { *score = -2; return; }
    int* Hprev = (int*)calloc((size_t)n + 1, sizeof(int));
    int* Farr  = (int*)calloc((size_t)n + 1, sizeof(int));
    int* lb = (int*)calloc((size_t)wsize * 3, sizeof(int));
    if (!Hprev || !Farr || !lb) { free(Hprev); free(Farr); free(lb); *score = -2; return; }
    for (int l = 0; l < wsize; ++l) { lb[3 * l] = 0; lb[3 * l + 1] = -1; lb[3 * l + 2] = -1; }
    const int go = gap_open, ge = gap_extend;
    /* The cell values are the Gotoh recurrence (H is independent of the E
     * boundary convention, which only moves non-positive E values).  Each lane
     * l owns columns j with (j-1) % wsize == l and walks them row by row, so its
     * first strict maximum is the first in row-major order among its columns
     * (sw_kernel_impl.hip:86-147). */
    for (int i = 1; i <= m; ++i) {
        const char a = text[i - 1];
        int hleft = 0, e_row = 0, diag = 0;
        for (int j = 1; j <= n; ++j) {
            int e = hleft - go;
            { const int e2 = e_row - ge; if (e2 > e) e = e2; }
            int f = Hprev[j] - go;
            { const int f2 = Farr[j] - ge; if (f2 > f) f = f2; }
            int h = diag + SUB(a, pattern[j - 1]);
            if (e > h) h = e;
            if (f > h) h = f;
            if (h < 0) h = 0;
            diag = Hprev[j]; Hprev[j] = h; Farr[j] = f; hleft = h; e_row = e;
            const int l = (j - 1) % wsize;
            if (h > lb[3 * l]) { lb[3 * l] = h; lb[3 * l + 1] = i - 1; lb[3 * l + 2] = j - 1; }
        }
    }
    /* xor-butterfly argmax, every lane updating from its partner's previous
     * value; a partner wins only when strictly greater (sw_kernel_impl.hip:161-167). */
    int* nb = (int*)malloc(sizeof(int) * (size_t)wsize * 3);
    if (!nb) { free(Hprev); free(Farr); free(lb); *score = -2; return; }
    for (int off = wsize / 2; off > 0; off >>= 1) {
        for (int l = 0; l < wsize; ++l) {
            const int o = l ^ off;
            if (lb[3 * o] > lb[3 * l]) {
                nb[3 * l] = lb[3 * o]; nb[3 * l + 1] = lb[3 * o + 1]; nb[3 * l + 2] = lb[3 * o + 2];
            } else {
                nb[3 * l] = lb[3 * l]; nb[3 * l + 1] = lb[3 * l + 1]; nb[3 * l + 2] = lb[3 * l + 2];
            }
        }
        memcpy(lb, nb, sizeof(int) * (size_t)wsize * 3);
    }
    *score = lb[0]; *end_i = lb[1]; *end_j = lb[2];
    free(nb); free(Hprev); free(Farr); free(lb);
}

/* tests/sw/test_sw_trace.cpp:68-136 -- full matrices and a value-comparison walk. */
int ga_oracle_sw_value_walk(const char* text, int m, const char* pattern, int n,
                            const int32_t* sub, int gap_open, int gap_extend,
                            int* score, int* start_i, int* start_j, int* end_i, int* end_j,
                            char* cigar, int* cigar_len)
{
    *score = 0; *start_i = *start_j = *end_i = *end_j = -1; *cigar_len = 0;
    if (m <= 0 || n <= 0) return 0;
    const size_t W = (size_t)n + 1, cells = ((size_t)m + 1) * W;
    int* H = (int*)calloc(cells, sizeof(int));
    int* E = (int*)calloc(cells, sizeof(int));
    int* F = (int*)calloc(cells, sizeof(int));
    if (!H || !E || !F) { free(H); free(E); free(F); return -1; }
#define AT(M, i, j) M[(size_t)(i) * W + (size_t)(j)]
    int best = 0, bi = 0, bj = 0;
    for (int i = 1; i <= m; ++i)
        for (int j = 1; j <= n; ++j) {
            const int s = SUB(text[i - 1], pattern[j - 1]);
            int e = AT(H, i, j - 1) - gap_open;
            { const int t = AT(E, i, j - 1) - gap_extend; if (t > e) e = t; }
            int f = AT(H, i - 1, j) - gap_open;
            { const int t = AT(F, i - 1, j) - gap_extend; if (t > f) f = t; }
            int h = AT(H, i - 1, j - 1) + s;
            if (e > h) h = e;
            if (f > h) h = f;
            if (h < 0) h = 0;
            AT(H, i, j) = h; AT(E, i, j) = e; AT(F, i, j) = f;
            if (h > best) { best = h; bi = i; bj = j; }
        }
    *score = best;
    if (best > 0) {
        char* rev = (char*)malloc((size_t)m + (size_t)n + 4);
        if (!rev) { free(H); free(E); free(F); return -1; }
        int used = 0, i = bi, j = bj, state = 0;
        while (i > 0 && j > 0) {
            if (state == 0) {
                const int h = AT(H, i, j);
                if (h == 0) break;
                const int s = SUB(text[i - 1], pattern[j - 1]);
                if (h == AT(H, i - 1, j - 1) + s) {
                    rev[used++] = (text[i - 1] == pattern[j - 1]) ? 'M' : 'X';
                    --i; --j;
                } else if (h == AT(E, i, j)) {
                    rev[used++] = 'D';
                    const int ext = (AT(E, i, j - 1) - gap_extend) > (AT(H, i, j - 1) - gap_open);
                    --j; state = ext ? 1 : 0;
                } else {
                    rev[used++] = 'I';
                    const int ext = (AT(F, i - 1, j) - gap_extend) > (AT(H, i - 1, j) - gap_open);
                    --i; state = ext ? 2 : 0;
                }
            } else if (state == 1) {
                rev[used++] = 'D';
                const int ext = (AT(E, i, j - 1) - gap_extend) > (AT(H, i, j - 1) - gap_open);
                --j; if (!ext) state = 0;
            } else {
                rev[used++] = 'I';
                const int ext = (AT(F, i - 1, j) - gap_extend) > (AT(H, i - 1, j) - gap_open);
                --i; if (!ext) state = 0;
            }
        }
        *start_i = i; *start_j = j; *end_i = bi - 1; *end_j = bj - 1;
        if (cigar) for (int c = 0; c < used; ++c) cigar[c] = rev[used - 1 - c];
        *cigar_len = used;
        free(rev);
    }
#undef AT
    free(H); free(E); free(F);
    return 0;
}

/* api.cpp:114-156 */
void ga_oracle_sw_validate(const char* cigar, int len,
                           const char* text, int text_len, int start_i, int end_i,
                           const char* pattern, int pattern_len, int start_j, int end_j,
                           const int32_t* sub, int gap_open, int gap_extend,
                           int score, int* rescore_ok, int* wellformed_ok)
{
    int i = start_i, j = start_j, s = 0, run = 0, wf = 1;
#define CLOSE_GAP() do { if (run != 0) { const int L = run > 0 ? run : -run; \
        s -= gap_open + (L - 1) * gap_extend; run = 0; } } while (0)
    for (int c = 0; c < len; ++c) {
        switch (cigar[c]) {
            case 'M':
                CLOSE_GAP();
                if (i > end_i || j > end_j || i >= text_len || j >= pattern_len ||
                    text[i] != pattern[j]) wf = 0;
                if (wf) s += SUB(text[i], pattern[j]);
                ++i; ++j; break;
            case 'X':
                CLOSE_GAP();
                if (i > end_i || j > end_j || i >= text_len || j >= pattern_len ||
                    text[i] == pattern[j]) wf = 0;
                if (wf) s += SUB(text[i], pattern[j]);
                ++i; ++j; break;
            case 'I':
                if (run < 0) CLOSE_GAP();
                if (i > end_i || i >= text_len) wf = 0;
                ++i; ++run; break;
            case 'D':
                if (run > 0) CLOSE_GAP();
                if (j > end_j || j >= pattern_len) wf = 0;
                ++j; --run; break;
            default: wf = 0; break;
        }
    }
    CLOSE_GAP();
#undef CLOSE_GAP
    if (i != end_i + 1 || j != end_j + 1) wf = 0;
    *wellformed_ok = wf;
    *rescore_ok = wf && (s == score);
}

/* ========================================================================== */
/* Threaded batch drivers (timed CPU baseline)                                 */
/* ========================================================================== */

typedef struct {
    int kind;                    /* 0 = WFA, 1 = SW */
    const char *a, *b;
    const int64_t *a_off, *b_off;
    const int32_t *a_len, *b_len;
    const int32_t* smax;
    const uint8_t* want_cigar;
    int64_t n;
    const int32_t* sub;
    int go, ge, with_cigar;
    int32_t* out;
    char* cigar;
    const int64_t* cig_off;
    int32_t* cig_len;
    int64_t next;                /* work counter (atomic) */
    int failed;
} Job;

static void* worker(void* arg)
{
    Job* J = (Job*)arg;
    for (;;) {
        const int64_t i0 = __atomic_fetch_add(&J->next, 64, __ATOMIC_RELAXED);
        if (i0 >= J->n) break;
        const int64_t i1 = i0 + 64 < J->n ? i0 + 64 : J->n;
        for (int64_t i = i0; i < i1; ++i) {
            const char* A = J->a + J->a_off[i];
            const char* B = J->b + J->b_off[i];
            char* cg = J->cigar ? J->cigar + J->cig_off[i] : NULL;
            int len = 0;
            if (J->kind == 0) {
                int sc = -1;
                const int want = J->want_cigar ? J->want_cigar[i] : 1;
                if (want) {
                    if (ga_oracle_wfa(A, J->a_len[i], B, J->b_len[i], J->smax[i], &sc, cg, &len))
                        __atomic_store_n(&J->failed, 1, __ATOMIC_RELAXED);
                } else {
                    sc = ga_oracle_wfa_score(A, J->a_len[i], B, J->b_len[i], J->smax[i]);
                }
                J->out[i] = sc;
                if (J->cig_len) J->cig_len[i] = len;
            } else {
                int32_t* o = J->out + 5 * i;
                int sc, si, sj, ei, ej;
                if (J->with_cigar) {
                    if (ga_oracle_sw_trace(A, J->a_len[i], B, J->b_len[i], J->sub, J->go, J->ge,
                                           &sc, &si, &sj, &ei, &ej, cg, &len))
                        __atomic_store_n(&J->failed, 1, __ATOMIC_RELAXED);
                } else {
                    ga_oracle_sw_score_lane(A, J->a_len[i], B, J->b_len[i], J->sub, J->go, J->ge,
                                            64, &sc, &ei, &ej);
                    si = sj = -1;
                }
                o[0] = sc; o[1] = si; o[2] = sj; o[3] = ei; o[4] = ej;
                if (J->cig_len) J->cig_len[i] = len;
            }
        }
    }
    return NULL;
}

static int run_job(Job* J, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 0; t < threads; ++t) {
        if (pthread_create(&th[t], NULL, worker, J) != 0) break;
        ++started;
    }
    if (started == 0) worker(J);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    return J->failed ? -1 : 0;
}

int ga_oracle_wfa_batch(const char* pat, const int64_t* pat_off, const int32_t* pat_len,
                        const char* txt, const int64_t* txt_off, const int32_t* txt_len,
                        const int32_t* smax, const uint8_t* want_cigar, int64_t n_pairs,
                        int32_t* score, char* cigar, const int64_t* cig_off, int32_t* cig_len,
                        int threads)
{
    Job J;
    memset(&J, 0, sizeof J);
    J.kind = 0; J.a = pat; J.a_off = pat_off; J.a_len = pat_len;
    J.b = txt; J.b_off = txt_off; J.b_len = txt_len; J.smax = smax;
    J.want_cigar = want_cigar; J.n = n_pairs; J.out = score;
    J.cigar = cigar; J.cig_off = cig_off; J.cig_len = cig_len;
    return run_job(&J, threads);
}

int ga_oracle_sw_batch(const char* txt, const int64_t* txt_off, const int32_t* txt_len,
                       const char* pat, const int64_t* pat_off, const int32_t* pat_len,
                       int64_t n_pairs, const int32_t* sub, int gap_open, int gap_extend,
                       int with_cigar, int32_t* out5,
                       char* cigar, const int64_t* cig_off, int32_t* cig_len, int threads)
{
    Job J;
    memset(&J, 0, sizeof J);
    J.kind = 1; J.a = txt; J.a_off = txt_off; J.a_len = txt_len;
    J.b = pat; J.b_off = pat_off; J.b_len = pat_len; J.n = n_pairs;
    J.sub = sub; J.go = gap_open; J.ge = gap_extend; J.with_cigar = with_cigar;
    J.out = out5; J.cigar = cigar; J.cig_off = cig_off; J.cig_len = cig_len;
    return run_job(&J, threads);
}
