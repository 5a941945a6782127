/*
 * ga_router_oracle.h -- CPU restatement of the learned router's arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; only tests/ and bench.py's baseline
 * leg load it).  The router has no counterpart in the reference: these
 * functions restate this build's own definitions so that the GPU kernels can be
 * compared with them, and they are pinned in tests/test_router_oracle.py by
 * independent means (a numpy recount of 4-mers, the invariant "lower bound <=
 * Levenshtein distance", a float64 numpy forward pass and finite differences of
 * the loss).
 *
 * Sketch of a pair (P = pattern, length m; T = text, length n):
 *   cP[b], cT[b]   counts of the 4-mers b (A0 C1 G2 T3, first base most
 *                  significant) over windows of A/C/G/T bytes only
 *   L1             sum_b |cP[b] - cT[b]|
 *   lower bound    max(|m - n|, ceil(L1 / 8))     (q-gram lemma, q = 4)
 *   features       f0 = m/10000, f1 = n/10000, f2 = |m-n|/M, f3 = L1/(2M),
 *                  f[4+j] = 30 * G[j] / M for j < 60, G[j] = sum over b with
 *                  b mod 60 == j of |cP[b] - cT[b]|;  M = max(m, n, 1)
 * MLP: 64 -> 128 (ReLU) -> 128 (ReLU) -> 16, parameters packed as
 *   W1[64][128], b1[128], W2[128][128], b2[128], W3[128][16], b3[16]
 * with each output accumulated over the input index in ascending order
 * (separate multiply and add) and the bias added last; loss = mean softmax
 * cross-entropy.
 */
#ifndef GA_ROUTER_ORACLE_H
#define GA_ROUTER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GA_OR_DIM 64
#define GA_OR_HID 128
#define GA_OR_OUT 16
#define GA_OR_PARAMS 26896

void ga_oracle_sketch(const char* pattern, int m, const char* text, int n, float* feat, int32_t* lower_bound);

void ga_oracle_mlp_forward(const float* params, const float* x, int64_t n, float* h1, float* h2, float* logits);

/* grads (GA_OR_PARAMS) of the mean cross-entropy; loss[r] per sample */
void ga_oracle_mlp_backward(const float* params, const float* x, const float* h1, const float* h2,
                            const float* logits, const int32_t* label, int64_t n, float* grads, float* loss);

#ifdef __cplusplus
}
#endif
#endif
