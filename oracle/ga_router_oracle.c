/* ga_router_oracle.c -- see ga_router_oracle.h.  TEST INFRASTRUCTURE ONLY. */
#include "ga_router_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static int code4(unsigned char b)
{
    return b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : b == 'T' ? 3 : -1;
}

static void count4(const char* s, int len, int* hist)
{
    for (int i = 0; i + 3 < len; ++i) {
        const int a = code4((unsigned char)s[i]), b = code4((unsigned char)s[i + 1]);
        const int c = code4((unsigned char)s[i + 2]), d = code4((unsigned char)s[i + 3]);
        if ((a | b | c | d) >= 0) ++hist[(a << 6) | (b << 4) | (c << 2) | d];
    }
}

void ga_oracle_sketch(const char* pattern, int m, const char* text, int n, float* feat, int32_t* lower_bound)
{
    int hp[256], ht[256];
    memset(hp, 0, sizeof hp);
    memset(ht, 0, sizeof ht);
    count4(pattern, m, hp);
    count4(text, n, ht);
    int l1 = 0;
    for (int b = 0; b < 256; ++b) l1 += abs(hp[b] - ht[b]);
    const int mm = m > 0 ? m : 0, nn = n > 0 ? n : 0;
    const int mx = (mm > nn ? mm : nn) > 0 ? (mm > nn ? mm : nn) : 1;
    const int dl = mm > nn ? mm - nn : nn - mm;
    const float fm = (float)mx;
    feat[0] = (float)mm / 10000.0f;
    feat[1] = (float)nn / 10000.0f;
    feat[2] = (float)dl / fm;
    feat[3] = (float)l1 / (float)(2 * mx);
    for (int j = 0; j < 60; ++j) {
        int g = 0;
        for (int b = j; b < 256; b += 60) g += abs(hp[b] - ht[b]);
        feat[4 + j] = (float)(g * 30) / fm;
    }
    const int qb = (l1 + 7) / 8;
    *lower_bound = dl > qb ? dl : qb;
}

enum { W1 = 0, B1 = W1 + GA_OR_DIM * GA_OR_HID, W2 = B1 + GA_OR_HID, B2 = W2 + GA_OR_HID * GA_OR_HID,
       W3 = B2 + GA_OR_HID, B3 = W3 + GA_OR_HID * GA_OR_OUT };

static void dense(const float* x, const float* w, const float* b, float* y, int64_t n, int in, int out, int relu)
{
    for (int64_t r = 0; r < n; ++r)
        for (int o = 0; o < out; ++o) {
            float acc = 0.0f;
            for (int i = 0; i < in; ++i) {
                const float p = x[r * in + i] * w[i * out + o];
                acc = acc + p;
            }
            acc = acc + b[o];
            y[r * out + o] = relu ? (acc > 0.0f ? acc : 0.0f) : acc;
        }
}

void ga_oracle_mlp_forward(const float* params, const float* x, int64_t n, float* h1, float* h2, float* logits)
{
    dense(x, params + W1, params + B1, h1, n, GA_OR_DIM, GA_OR_HID, 1);
    dense(h1, params + W2, params + B2, h2, n, GA_OR_HID, GA_OR_HID, 1);
    dense(h2, params + W3, params + B3, logits, n, GA_OR_HID, GA_OR_OUT, 0);
}

/* dW[i][o] = sum_r a[r][i] d[r][o]; db[o] = sum_r d[r][o] */
static void weight_grad(const float* a, const float* d, float* dw, float* db, int64_t n, int in, int out)
{
    for (int i = 0; i < in; ++i)
        for (int o = 0; o < out; ++o) {
            float acc = 0.0f;
            for (int64_t r = 0; r < n; ++r) { const float p = a[r * in + i] * d[r * out + o]; acc = acc + p; }
            dw[i * out + o] = acc;
        }
    for (int o = 0; o < out; ++o) {
        float acc = 0.0f;
        for (int64_t r = 0; r < n; ++r) acc = acc + d[r * out + o];
        db[o] = acc;
    }
}

/* dx[r][i] = (sum_o d[r][o] w[i][o]) * (a[r][i] > 0) */
static void data_grad(const float* d, const float* w, const float* a, float* dx, int64_t n, int in, int out)
{
    for (int64_t r = 0; r < n; ++r)
        for (int i = 0; i < in; ++i) {
            float acc = 0.0f;
            for (int o = 0; o < out; ++o) { const float p = d[r * out + o] * w[i * out + o]; acc = acc + p; }
            dx[r * in + i] = a[r * in + i] > 0.0f ? acc : 0.0f;
        }
}

void ga_oracle_mlp_backward(const float* params, const float* x, const float* h1, const float* h2,
                            const float* logits, const int32_t* label, int64_t n, float* grads, float* loss)
{
    float* dl = (float*)malloc(sizeof(float) * (size_t)n * GA_OR_OUT);
    float* d2 = (float*)malloc(sizeof(float) * (size_t)n * GA_OR_HID);
    float* d1 = (float*)malloc(sizeof(float) * (size_t)n * GA_OR_HID);
    const float inv_n = 1.0f / (float)n;
    for (int64_t r = 0; r < n; ++r) {
        const float* z = logits + r * GA_OR_OUT;
        float mx = z[0];
        for (int c = 1; c < GA_OR_OUT; ++c) mx = z[c] > mx ? z[c] : mx;
        float e[GA_OR_OUT], sum = 0.0f;
        for (int c = 0; c < GA_OR_OUT; ++c) { e[c] = expf(z[c] - mx); sum = sum + e[c]; }
        const int y = label[r];
        for (int c = 0; c < GA_OR_OUT; ++c) dl[r * GA_OR_OUT + c] = (e[c] / sum - (c == y ? 1.0f : 0.0f)) * inv_n;
        loss[r] = logf(sum) - (z[y] - mx);
    }
    weight_grad(h2, dl, grads + W3, grads + B3, n, GA_OR_HID, GA_OR_OUT);
    data_grad(dl, params + W3, h2, d2, n, GA_OR_HID, GA_OR_OUT);
    weight_grad(h1, d2, grads + W2, grads + B2, n, GA_OR_HID, GA_OR_HID);
    data_grad(d2, params + W2, h1, d1, n, GA_OR_HID, GA_OR_HID);
    weight_grad(x, d1, grads + W1, grads + B1, n, GA_OR_DIM, GA_OR_HID);
    free(dl);
    free(d2);
    free(d1);
}
