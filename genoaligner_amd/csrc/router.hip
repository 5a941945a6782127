// router.hip -- gfx950 kernels of the learned router: the 4-mer/length sketch
// of a pair, and a small FP32 MLP (64 -> 128 -> 128 -> 16, ReLU) with its
// forward pass, backward pass, SGD-with-momentum update and the routing rule.
//
// Nothing here changes an alignment result: the router only picks the smax a
// pair is first tried with (and so its kernel size class); a pair that is not
// resolved under that bound is re-run with its own smax by the host layer.
// The one exact decision taken here is the q-gram lower bound: every edit
// changes at most q q-grams of one sequence, so ||Q4(a) - Q4(b)||_1 <= 8 ED
// (and |m - n| <= ED); a pair whose bound exceeds its smax is unresolved.
//
// The MLP is a tiny dense contraction (26,896 parameters); it runs on FP32
// VALU with LDS-staged input rows.  Accumulation order is fixed (input index
// ascending, separate multiply and add) so a CPU restatement reproduces it.
#include "ga_device.hpp"

namespace ga {
namespace {

// 8 bytes starting at any byte address (two aligned words; sequences carry
// 128 readable pad bytes on both sides).
__device__ __forceinline__ uint64_t load8u(const uint8_t* p)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uint64_t* w = reinterpret_cast<const uint64_t*>(a & ~uintptr_t(7));
    const unsigned sh = unsigned(a & 7u) * 8u;
    const uint64_t lo = w[0];
    const uint64_t hi = w[1];
    return sh ? ((lo >> sh) | (hi << (64u - sh))) : lo;
}

// 4-mer counts of one sequence into an LDS histogram (windows holding a byte
// other than A/C/G/T are skipped).  Lane l rolls a 4-mer code over windows
// [l*chunk, (l+1)*chunk), reading each base once, 8 bytes per load.  Codes
// come four bytes at a time: ((b >> 1) ^ (b >> 2)) & 3 is A 0, C 1, G 2, T 3
// (bits 1..3 of the ASCII letters), and a byte is one of them iff the code's
// letter (a byte select from "ACGT") equals it.
__device__ void count4(const uint8_t* s, int len, int* hist, int lane)
{
    const int nwin = len - 3;
    if (nwin <= 0) return;
    const int chunk = (nwin + 63) / 64;
    const int w0 = lane * chunk;
    const int w1 = w0 + chunk < nwin ? w0 + chunk : nwin;
    if (w0 >= w1) return;
    // bases w0 .. w1 + 2 feed windows w0 .. w1 - 1
    uint32_t code = 0;                       // rolling 8-bit code
    int valid = 0;                           // trailing run of A/C/G/T bases
    for (int b = w0; b < w1 + 3; b += 8) {
        const uint64_t x = load8u(s + b);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t w = uint32_t(x >> (32 * h));
            const uint32_t c = ((w >> 1) ^ (w >> 2)) & 0x03030303u;
            const uint32_t bad = __builtin_amdgcn_perm(0u, 0x54474341u, c) ^ w;   // "ACGT"
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int pos = b + 4 * h + q;
                if (pos < w1 + 3) {
                    code = ((code << 2) | ((c >> (8 * q)) & 3u)) & 0xffu;
                    valid = ((bad >> (8 * q)) & 0xffu) ? 0 : valid + 1;
                    if (pos >= w0 + 3 && valid >= 4) atomicAdd(&hist[code], 1);
                }
            }
        }
    }
}

__device__ __forceinline__ int wave_sum(int v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One wave per pair, four pairs per workgroup (each wave with its own LDS
// histograms; a launch of one-wave workgroups was bound by workgroup dispatch
// at ~5 ns each): features (kSketchDim floats) and the exact lower bound
// max(|m - n|, ceil(L1 / 8)) on the edit distance.
constexpr int kSketchWaves = 4;
__global__ __launch_bounds__(64 * kSketchWaves) void sketch_kernel(SketchLaunch L)
{
    __shared__ int hist[kSketchWaves][2][256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t pid = int64_t(blockIdx.x) * kSketchWaves + wave;
    if (pid >= L.n) return;                      // whole waves only; no workgroup barrier follows
    int* hp = hist[wave][0];
    int* ht = hist[wave][1];
    // the histograms are this wave's own: a wave-scope fence and barrier order
    // the clearing before the counting and the counting before the reads
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    for (int b = lane; b < 256; b += 64) { hp[b] = 0; ht[b] = 0; }
    wave_sync();
    const int m = L.plen[pid], n = L.tlen[pid];
    count4(L.pat + L.poff[pid], m, hp, lane);
    count4(L.txt + L.toff[pid], n, ht, lane);
    wave_sync();
    int part = 0;
    for (int b = lane; b < 256; b += 64) part += abs(hp[b] - ht[b]);
    const int l1 = wave_sum(part);
    int grp = 0;                                    // group j = bins b with b % 60 == j
    if (lane < 60)
        for (int b = lane; b < 256; b += 60) grp += abs(hp[b] - ht[b]);
    const int mm = m > 0 ? m : 0, nn = n > 0 ? n : 0;
    const int mx = (mm > nn ? mm : nn) > 0 ? (mm > nn ? mm : nn) : 1;
    const int dl = mm > nn ? mm - nn : nn - mm;
    const float fm = float(mx);
    float* f = L.feat + pid * kSketchDim;
    if (lane == 0) {
        f[0] = float(mm) / 10000.0f;
        f[1] = float(nn) / 10000.0f;
        f[2] = float(dl) / fm;
        f[3] = float(l1) / float(2 * mx);
        const int qb = (l1 + 7) / 8;
        L.lower_bound[pid] = dl > qb ? dl : qb;
    }
    if (lane < 60) f[4 + lane] = float(grp * 30) / fm;
}

// ---------------------------------------------------------------- MLP
// Y[r, o] = act(b[o] + sum_i X[r, i] W[i, o]); a workgroup of 256 threads
// covers kRows rows, staged in LDS; W (at most 64 KB) is read through the
// caches, 64 consecutive outputs per wave.
constexpr int kRows = 32;

template <bool RELU>
__global__ __launch_bounds__(256) void dense_fwd(const float* X, const float* W, const float* bias,
                                                float* Y, int64_t N, int I, int O)
{
    extern __shared__ float sm[];
    float* sX = sm;                  // kRows * I
    const int t = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * kRows;
    for (int x = t; x < kRows * I; x += 256) {
        const int64_t r = r0 + x / I;
        sX[x] = r < N ? X[r * I + x % I] : 0.0f;
    }
    __syncthreads();
    // a thread takes 4 rows of one output: each weight read serves 4 rows
    // (the weights come through L2; 64 KB does not fit L1), and each row's sum
    // runs over i ascending, multiply then add, as a one-row loop would
    constexpr int kRt = 4;
    for (int x = t; x < (kRows / kRt) * O; x += 256) {
        const int rg = x / O, o = x % O;
        const float* xr = sX + rg * kRt * I;
        float acc[kRt] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int i = 0; i < I; ++i) {
            const float w = W[i * O + o];
#pragma unroll
            for (int q = 0; q < kRt; ++q) acc[q] += xr[q * I + i] * w;
        }
#pragma unroll
        for (int q = 0; q < kRt; ++q) {
            const int64_t r = r0 + rg * kRt + q;
            if (r >= N) break;
            const float v = acc[q] + bias[o];
            Y[r * O + o] = RELU ? (v > 0.0f ? v : 0.0f) : v;
        }
    }
}

// Inference only (the routed path): the three layers fused, logits out.  A
// workgroup takes kLRows rows; the layer inputs stay in LDS, transposed
// (column i of the tile is kLRows consecutive floats), so a thread's 4 rows
// of one input come in one read; each thread of a hidden layer computes a
// 4-row x 4-output tile, one weight read serving 16 products.  Each sum runs
// over the inputs in ascending order with fused multiply-adds, then adds the
// bias (the training forward above multiplies and adds separately: logits
// may differ from it in the last bits, which changes no alignment result).
constexpr int kLRows = 32;

__global__ __launch_bounds__(256) void mlp_logits_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ X, float* logits, int64_t N)
{
    __shared__ __align__(16) float xt[kSketchDim * kLRows];
    __shared__ __align__(16) float h1t[kHidden * kLRows];
    __shared__ __align__(16) float h2t[kHidden * kLRows];
    const MlpLayout P;
    const int t = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * kLRows;
    for (int x = t; x < kLRows * kSketchDim; x += 256) {
        const int r = x / kSketchDim, i = x % kSketchDim;
        xt[i * kLRows + r] = r0 + r < N ? X[(r0 + r) * kSketchDim + i] : 0.0f;
    }
    __syncthreads();
    const int rg = (t >> 5) * 4, og = (t & 31) * 4;
    auto hidden = [&](const float* in_t, int I, const float* W, const float* b, float* out_t) {
        float acc[4][4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int o = 0; o < 4; ++o) acc[q][o] = 0.0f;
        for (int i = 0; i < I; ++i) {
            const float4 xv = *reinterpret_cast<const float4*>(in_t + i * kLRows + rg);
            const float4 wv = *reinterpret_cast<const float4*>(W + i * kHidden + og);
            const float xs[4] = {xv.x, xv.y, xv.z, xv.w};
            const float ws[4] = {wv.x, wv.y, wv.z, wv.w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int o = 0; o < 4; ++o) acc[q][o] = __builtin_fmaf(xs[q], ws[o], acc[q][o]);
        }
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const float bo = b[og + o];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float v = acc[q][o] + bo;
                out_t[(og + o) * kLRows + rg + q] = v > 0.0f ? v : 0.0f;
            }
        }
    };
    hidden(xt, kSketchDim, params + P.w1, params + P.b1, h1t);
    __syncthreads();
    hidden(h1t, kHidden, params + P.w2, params + P.b2, h2t);
    __syncthreads();
    // output layer: row t / 8, outputs 2 (t % 8) and the next
    const int r = t >> 3, o = (t & 7) * 2;
    float a0 = 0.0f, a1 = 0.0f;
    for (int i = 0; i < kHidden; ++i) {
        const float x = h2t[i * kLRows + r];
        const float2 wv = *reinterpret_cast<const float2*>(params + P.w3 + i * kBuckets + o);
        a0 = __builtin_fmaf(x, wv.x, a0);
        a1 = __builtin_fmaf(x, wv.y, a1);
    }
    if (r0 + r < N) {
        logits[(r0 + r) * kBuckets + o] = a0 + params[P.b3 + o];
        logits[(r0 + r) * kBuckets + o + 1] = a1 + params[P.b3 + o + 1];
    }
}

// dX[r, i] = (sum_o dY[r, o] W[i, o]) * (A[r, i] > 0), A the layer input
// after its ReLU.
__global__ __launch_bounds__(256) void dense_bwd_data(const float* dY, const float* W, const float* A,
                                                     float* dX, int64_t N, int I, int O)
{
    extern __shared__ float sm[];
    float* sD = sm;                  // kRows * O
    const int t = threadIdx.x;
    const int64_t r0 = int64_t(blockIdx.x) * kRows;
    for (int x = t; x < kRows * O; x += 256) {
        const int64_t r = r0 + x / O;
        sD[x] = r < N ? dY[r * O + x % O] : 0.0f;
    }
    __syncthreads();
    for (int x = t; x < kRows * I; x += 256) {
        const int rr = x / I, i = x % I;
        const int64_t r = r0 + rr;
        if (r >= N) continue;
        float acc = 0.0f;
        for (int o = 0; o < O; ++o) acc += sD[rr * O + o] * W[i * O + o];
        dX[r * I + i] = A[r * I + i] > 0.0f ? acc : 0.0f;
    }
}

// dW[i, o] = sum_r X[r, i] dY[r, o]; db[o] = sum_r dY[r, o] (r ascending).
// One thread per weight; the threads of row i == I compute the bias.
__global__ __launch_bounds__(256) void dense_bwd_weight(const float* X, const float* dY, float* dW,
                                                       float* db, int64_t N, int I, int O)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= (I + 1) * O) return;
    const int i = x / O, o = x % O;
    float acc = 0.0f;
    if (i < I) {
        for (int64_t r = 0; r < N; ++r) acc += X[r * I + i] * dY[r * O + o];
        dW[i * O + o] = acc;
    } else {
        for (int64_t r = 0; r < N; ++r) acc += dY[r * O + o];
        db[o] = acc;
    }
}

// Softmax cross-entropy over kBuckets logits: loss[r] = -log p[label], and
// dL[r, c] = (p[c] - [c == label]) / N (mean over the batch).
__global__ __launch_bounds__(256) void softmax_xent(const float* logits, const int32_t* label, float* dL,
                                                   float* loss, int64_t N)
{
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= N) return;
    const float* z = logits + r * kBuckets;
    float mx = z[0];
    for (int c = 1; c < kBuckets; ++c) mx = z[c] > mx ? z[c] : mx;
    float e[kBuckets];
    float sum = 0.0f;
    for (int c = 0; c < kBuckets; ++c) { e[c] = expf(z[c] - mx); sum += e[c]; }
    const int y = label[r];
    const float inv_n = 1.0f / float(N);
    for (int c = 0; c < kBuckets; ++c) dL[r * kBuckets + c] = (e[c] / sum - (c == y ? 1.0f : 0.0f)) * inv_n;
    loss[r] = logf(sum) - (z[y] - mx);
}

__global__ __launch_bounds__(256) void sgd_momentum(float* p, const float* g, float* v, int n, float lr, float mu)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= n) return;
    const float vv = mu * v[x] + g[x];
    v[x] = vv;
    p[x] = p[x] - lr * vv;
}

// Routing rule: the smallest bucket whose cumulative softmax probability
// reaches q (q = 0: the arg-max bucket).
__global__ __launch_bounds__(256) void route_buckets(const float* logits, int64_t N, float q, int32_t* bucket)
{
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= N) return;
    const float* z = logits + r * kBuckets;
    float mx = z[0];
    int am = 0;
    for (int c = 1; c < kBuckets; ++c)
        if (z[c] > mx) { mx = z[c]; am = c; }
    if (q <= 0.0f) { bucket[r] = am; return; }
    float e[kBuckets];
    float sum = 0.0f;
    for (int c = 0; c < kBuckets; ++c) { e[c] = expf(z[c] - mx); sum += e[c]; }
    float cum = 0.0f;
    int b = kBuckets - 1;
    for (int c = 0; c < kBuckets; ++c) {
        cum += e[c];
        if (cum >= q * sum) { b = c; break; }
    }
    bucket[r] = b;
}

__global__ void zero_u64(unsigned long long* p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

__constant__ int32_t kBound[kBuckets] = {0, 1, 2, 4, 8, 16, 32, 64, 128, 256, 384, 512, 768, 1024, 2048, 0x7fffffff};

// Adds, for each k, the number of threads of the (256-thread) workgroup with
// pred[k] to ctr[k]: wave counts gathered in LDS, one atomic per workgroup and
// counter (atomics from every wave on the same few counters serialise).  Every
// thread of the workgroup calls it.
template <int K>
__device__ __forceinline__ void block_count(const bool (&pred)[K], unsigned long long* ctr)
{
    __shared__ unsigned int part[4][K];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const unsigned long long m = __ballot(pred[k]);
        if ((threadIdx.x & 63) == 0) part[w][k] = unsigned(__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const unsigned sum = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
        if (sum) atomicAdd(&ctr[threadIdx.x], (unsigned long long)sum);
    }
}

// First pass of a routed batch: each pair's bound = its predicted bucket's
// bound (never below the exact lower bound, never above its own smax); a pair
// whose exact lower bound exceeds its smax is unresolved at once (score -1, no
// CIGAR) and left out (sm1 = -1).  stats[0..2] += shortcut, routed, plain.
__global__ __launch_bounds__(256) void route_first_pass(const int32_t* bucket, const int32_t* lb, const int32_t* smax,
                                                       const int32_t* plen, const int32_t* tlen, int n, int32_t* sm1,
                                                       int32_t* score, int32_t* cig_len, unsigned long long* stats)
{
    const int i = int(blockIdx.x * 256 + threadIdx.x);
    int kind = -1;                                          // 0 shortcut, 1 routed, 2 plain
    if (i < n) {
        const int s = smax[i];
        const bool degenerate = plen[i] <= 0 || tlen[i] <= 0;   // scores the other length whatever its smax
        if (!degenerate && lb[i] > s) {
            sm1[i] = -1;
            score[i] = -1;
            if (cig_len) cig_len[i] = 0;
            kind = 0;
        } else {
            int b = bucket[i];
            b = b < 0 ? 0 : (b >= kBuckets ? kBuckets - 1 : b);
            int v = kBound[b] < s ? kBound[b] : s;
            const int l = lb[i] < s ? lb[i] : s;
            v = v > l ? v : l;
            // a bound in the largest size class is the pair's own smax: the
            // second pass then only re-runs pairs of the smaller classes
            if (v > kWfaClassBound2) v = s;
            sm1[i] = v;
            kind = v < s ? 1 : 2;
        }
    }
    const bool pred[3] = {kind == 0, kind == 1, kind == 2};
    block_count(pred, stats);
}

// Second pass: the routed pairs not resolved under their first bound, with
// their own smax; everything else is left out.  stats[3] += re-runs.
__global__ __launch_bounds__(256) void route_second_pass(const int32_t* sm1, const int32_t* smax, const int32_t* score,
                                                        int n, int32_t* sm2, unsigned long long* stats)
{
    const int i = int(blockIdx.x * 256 + threadIdx.x);
    bool again = false;
    if (i < n) {
        again = sm1[i] >= 0 && sm1[i] < smax[i] && score[i] < 0;
        sm2[i] = again ? smax[i] : -1;
    }
    const bool pred[1] = {again};
    block_count(pred, stats + 3);
}

inline unsigned blocks(int64_t n, int per) { return unsigned((n + per - 1) / per); }

}  // namespace

hipError_t sketch_launch(const SketchLaunch& L, hipStream_t st)
{
    if (L.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sketch_kernel, dim3(unsigned((L.n + kSketchWaves - 1) / kSketchWaves)),
                       dim3(64 * kSketchWaves), 0, st, L);
    return hipGetLastError();
}

hipError_t mlp_forward(const float* params, const float* X, int64_t N, float* H1, float* H2, float* logits,
                       hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    const MlpLayout P;
    const unsigned g = blocks(N, kRows);
    hipLaunchKernelGGL((dense_fwd<true>), dim3(g), dim3(256), kRows * kSketchDim * 4, st,
                       X, params + P.w1, params + P.b1, H1, N, kSketchDim, kHidden);
    hipLaunchKernelGGL((dense_fwd<true>), dim3(g), dim3(256), kRows * kHidden * 4, st,
                       H1, params + P.w2, params + P.b2, H2, N, kHidden, kHidden);
    hipLaunchKernelGGL((dense_fwd<false>), dim3(g), dim3(256), kRows * kHidden * 4, st,
                       H2, params + P.w3, params + P.b3, logits, N, kHidden, kBuckets);
    return hipGetLastError();
}

hipError_t mlp_logits(const float* params, const float* X, int64_t N, float* logits, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    // the float4 weight reads need 16-byte aligned parameters
    if (reinterpret_cast<uintptr_t>(params) & 15u) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_logits_kernel, dim3(blocks(N, kLRows)), dim3(256), 0, st, params, X, logits, N);
    return hipGetLastError();
}

hipError_t mlp_backward(const float* params, const float* X, const float* H1, const float* H2, const float* logits,
                        const int32_t* label, int64_t N, float* grads, float* loss, float* dL, float* dH2,
                        float* dH1, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    const MlpLayout P;
    const unsigned g = blocks(N, kRows);
    hipLaunchKernelGGL(softmax_xent, dim3(blocks(N, 256)), dim3(256), 0, st, logits, label, dL, loss, N);
    hipLaunchKernelGGL(dense_bwd_weight, dim3(blocks((kHidden + 1) * kBuckets, 256)), dim3(256), 0, st,
                       H2, dL, grads + P.w3, grads + P.b3, N, kHidden, kBuckets);
    hipLaunchKernelGGL(dense_bwd_data, dim3(g), dim3(256), kRows * kBuckets * 4, st,
                       dL, params + P.w3, H2, dH2, N, kHidden, kBuckets);
    hipLaunchKernelGGL(dense_bwd_weight, dim3(blocks((kHidden + 1) * kHidden, 256)), dim3(256), 0, st,
                       H1, dH2, grads + P.w2, grads + P.b2, N, kHidden, kHidden);
    hipLaunchKernelGGL(dense_bwd_data, dim3(g), dim3(256), kRows * kHidden * 4, st,
                       dH2, params + P.w2, H1, dH1, N, kHidden, kHidden);
    hipLaunchKernelGGL(dense_bwd_weight, dim3(blocks((kSketchDim + 1) * kHidden, 256)), dim3(256), 0, st,
                       X, dH1, grads + P.w1, grads + P.b1, N, kSketchDim, kHidden);
    return hipGetLastError();
}

hipError_t mlp_sgd(float* params, const float* grads, float* mom, float lr, float mu, hipStream_t st)
{
    hipLaunchKernelGGL(sgd_momentum, dim3(blocks(kMlpParams, 256)), dim3(256), 0, st, params, grads, mom,
                       int(kMlpParams), lr, mu);
    return hipGetLastError();
}

hipError_t mlp_route(const float* logits, int64_t N, float q, int32_t* bucket, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(route_buckets, dim3(blocks(N, 256)), dim3(256), 0, st, logits, N, q, bucket);
    return hipGetLastError();
}

hipError_t route_first_pass_launch(const int32_t* bucket, const int32_t* lb, const int32_t* smax, const int32_t* plen,
                                   const int32_t* tlen, int n, int32_t* sm1, int32_t* score, int32_t* cig_len,
                                   unsigned long long* stats, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(route_first_pass, dim3(blocks(n, 256)), dim3(256), 0, st, bucket, lb, smax, plen, tlen, n,
                       sm1, score, cig_len, stats);
    return hipGetLastError();
}

hipError_t route_second_pass_launch(const int32_t* sm1, const int32_t* smax, const int32_t* score, int n, int32_t* sm2,
                                    unsigned long long* stats, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(route_second_pass, dim3(blocks(n, 256)), dim3(256), 0, st, sm1, smax, score, n, sm2, stats);
    return hipGetLastError();
}

hipError_t zero_u64_launch(unsigned long long* p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace ga
