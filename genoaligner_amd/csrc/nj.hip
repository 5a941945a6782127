// nj.hip -- gfx950 kernels for the guide-tree step of the reference's MSA
// engine: fragment-corrected k-mer distances (src/msa/msa_ref.cpp:802-841)
// and exact neighbour joining (msa_ref.cpp:904-947; the reference's device
// version include/genoaligner/backend/nj_kernel_impl.hip:84-187).
//
// k-mer distances.  A sequence's k-mer set is a bitset over the alpha^k
// codes: one thread per position sets the bit of the k-mer ending there when
// its k symbols are all unambiguous (the reference's rolling hash mod alpha^k
// is exactly that k-mer's base-alpha value).  Pair (i, j) then needs
// |K_i & K_j| = popcount(B_i & B_j): 16 x 16 pair tiles, the two blocks of
// bitsets staged through LDS 64 words at a time.  D = 1 - inter / min(|K_i|,
// |K_j|) with the quotient rounded once to float (computed in double, whose
// 53 bits make the double-then-float rounding equal the direct float
// division), 1 when the smaller set is empty.
//
// Neighbour joining.  d is n x n double, slot i holding leaf i; a merge
// writes the new node into its first child's slot.  One round: row sums in
// the live order, one thread per row, added strictly in that order (the
// reference's sum order, so every bit agrees); per row the first minimum of
// q = (m - 2) d(a, b) - r(a) - r(b) over b > a (lexicographic on (q, a, b),
// the reference's scan order); one workgroup reduces the rows, records the
// merge, writes the new node's distances 0.5 (d(i, v) + d(j, v) - d(i, j))
// and compacts the live list with the new node last.  All arithmetic is
// unfused (-ffp-contract=off).
#include "ga_device.hpp"

namespace ga {
namespace {

__device__ __forceinline__ int sym_index(uint8_t c, int alpha)
{
    if (alpha == 4) {
        switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': case 'U': case 'u': return 3;
        default: return -1;
        }
    }
    switch (c) {
    case 'A': case 'a': return 0;
    case 'R': case 'r': return 1;
    case 'N': case 'n': return 2;
    case 'D': case 'd': return 3;
    case 'C': case 'c': case 'U': case 'u': return 4;
    case 'Q': case 'q': return 5;
    case 'E': case 'e': return 6;
    case 'G': case 'g': return 7;
    case 'H': case 'h': return 8;
    case 'I': case 'i': return 9;
    case 'L': case 'l': return 10;
    case 'K': case 'k': return 11;
    case 'M': case 'm': return 12;
    case 'F': case 'f': return 13;
    case 'P': case 'p': return 14;
    case 'S': case 's': return 15;
    case 'T': case 't': return 16;
    case 'W': case 'w': return 17;
    case 'Y': case 'y': return 18;
    case 'V': case 'v': return 19;
    default: return -1;
    }
}

__global__ void kmer_bits_kernel(KmerLaunch K)
{
    const int s = int(blockIdx.y);
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;     // position of the k-mer's last symbol
    const int L = K.len[s];
    if (p >= L || p + 1 < K.k) return;
    const uint8_t* q = K.seq + K.off[s];
    uint64_t h = 0;
    for (int u = int(p) - K.k + 1; u <= int(p); ++u) {
        const int b = sym_index(q[u], K.alpha);
        if (b < 0) return;
        h = h * uint64_t(K.alpha) + uint64_t(b);
    }
    atomicOr(K.bits + int64_t(s) * K.words + int64_t(h >> 5), 1u << (h & 31));
}

__global__ void kmer_count_kernel(KmerLaunch K)
{
    const int s = int(blockIdx.x);
    int c = 0;
    for (int64_t w = threadIdx.x; w < K.words; w += blockDim.x) c += __popc(K.bits[int64_t(s) * K.words + w]);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) K.count[s] = part[0] + part[1] + part[2] + part[3];
}

// 16 x 16 pairs per workgroup (blocks bi <= bj), 64 words per LDS stage.
__global__ __launch_bounds__(256) void kmer_pairs_kernel(KmerLaunch K, int nblk)
{
    __shared__ uint32_t si[16][65], sj[16][65];
    // tile index -> (bi, bj), bi <= bj, row by row
    int t = int(blockIdx.x), bi = 0;
    while (t >= nblk - bi) { t -= nblk - bi; ++bi; }
    const int bj = bi + t;
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15;
    const int i = bi * 16 + ti, j = bj * 16 + tj;
    int inter = 0;
    for (int64_t w0 = 0; w0 < K.words; w0 += 64) {
        for (int q = threadIdx.x; q < 16 * 64; q += 256) {
            const int r = q >> 6, c = q & 63;
            const int64_t w = w0 + c;
            const int ii = bi * 16 + r, jj = bj * 16 + r;
            si[r][c] = (ii < K.n && w < K.words) ? K.bits[int64_t(ii) * K.words + w] : 0u;
            sj[r][c] = (jj < K.n && w < K.words) ? K.bits[int64_t(jj) * K.words + w] : 0u;
        }
        __syncthreads();
#pragma unroll 8
        for (int c = 0; c < 64; ++c) inter += __popc(si[ti][c] & sj[tj][c]);
        __syncthreads();
    }
    if (i < K.n && j < K.n && i < j) {
        const int ci = K.count[i], cj = K.count[j];
        const int mn = ci < cj ? ci : cj;
        const float d = mn ? 1.0f - float(double(inter) / double(mn)) : 1.0f;
        K.dist[int64_t(i) * K.n - int64_t(i) * (i + 1) / 2 + (j - i - 1)] = d;
    }
}

// ---------------------------------------------------------------- NJ
using NjBestD = NjBest;
__device__ __forceinline__ bool nj_before(const NjBestD& x, const NjBestD& y)
{
    if (y.a < 0) return x.a >= 0;
    if (x.a < 0) return false;
    if (x.q != y.q) return x.q < y.q;
    if (x.a != y.a) return x.a < y.a;
    return x.b < y.b;
}

__global__ void nj_init_kernel(NjLaunch J)
{
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int n = J.n;
    if (e >= int64_t(n) * n) return;
    const int i = int(e / n), j = int(e % n);
    double v = 0.0;
    if (i != j) {
        const int a = i < j ? i : j, b = i < j ? j : i;
        v = double(J.packed[int64_t(a) * n - int64_t(a) * (a + 1) / 2 + (b - a - 1)]);
    }
    J.d[e] = v;
    if (e < n) {
        J.alive0[e] = int(e);
        J.slot_node[e] = int(e);
    }
}

__global__ __launch_bounds__(64) void nj_rowsum_kernel(NjLaunch J, const int* __restrict__ alive, int m)
{
    __shared__ int sa[64];
    const int a = int(blockIdx.x) * 64 + int(threadIdx.x);
    const int n = J.n;
    const int col = a < m ? alive[a] : 0;
    double s = 0.0;
    for (int b0 = 0; b0 < m; b0 += 64) {
        __syncthreads();
        if (b0 + int(threadIdx.x) < m) sa[threadIdx.x] = alive[b0 + threadIdx.x];
        __syncthreads();
        const int tile = m - b0 < 64 ? m - b0 : 64;
        if (a < m) {
            int k = 0;
            for (; k + 8 <= tile; k += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = J.d[int64_t(sa[k + u]) * n + col];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (b0 + k + u != a) s += v[u];
            }
            for (; k < tile; ++k)
                if (b0 + k != a) s += J.d[int64_t(sa[k]) * n + col];
        }
    }
    if (a < m) J.r[a] = s;
}

__device__ void nj_block_min(NjBestD* sh)
{
    for (int s = int(blockDim.x) / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s && nj_before(sh[threadIdx.x + s], sh[threadIdx.x])) sh[threadIdx.x] = sh[threadIdx.x + s];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void nj_argmin_kernel(NjLaunch J, const int* __restrict__ alive, int m)
{
    __shared__ NjBestD sh[256];
    const int a = int(blockIdx.x);
    NjBestD best{0.0, -1, -1};
    const double mm = double(m - 2);
    const double ra = J.r[a];
    const double* row = J.d + int64_t(alive[a]) * J.n;
    for (int b = a + 1 + int(threadIdx.x); b < m; b += int(blockDim.x)) {
        const double q = mm * row[alive[b]] - ra - J.r[b];
        const NjBestD c{q, a, b};
        if (nj_before(c, best)) best = c;
    }
    sh[threadIdx.x] = best;
    __syncthreads();
    nj_block_min(sh);
    if (threadIdx.x == 0) J.part[a] = sh[0];
}

__global__ __launch_bounds__(256) void nj_merge_kernel(NjLaunch J, const int* __restrict__ alive_in,
                                                       int* __restrict__ alive_out, int m, int round)
{
    __shared__ NjBestD sh[256];
    NjBestD best{0.0, -1, -1};
    for (int a = int(threadIdx.x); a < m - 1; a += int(blockDim.x))
        if (nj_before(J.part[a], best)) best = J.part[a];
    sh[threadIdx.x] = best;
    __syncthreads();
    nj_block_min(sh);
    const int bi = sh[0].a, bj = sh[0].b;
    const int n = J.n;
    const int si = alive_in[bi], sj = alive_in[bj];
    double* ri = J.d + int64_t(si) * n;
    const double* rj = J.d + int64_t(sj) * n;
    const double dij = ri[sj];
    if (threadIdx.x == 0) {
        J.children[2 * round] = J.slot_node[si];
        J.children[2 * round + 1] = J.slot_node[sj];
    }
    __syncthreads();                                   // every thread has read slot_node[si] / dij's row
    for (int k = int(threadIdx.x); k < m; k += int(blockDim.x)) {
        if (k == bi || k == bj) continue;
        const int v = alive_in[k];
        const double nv = 0.5 * (ri[v] + rj[v] - dij);
        ri[v] = nv;
        J.d[int64_t(v) * n + si] = nv;
        alive_out[k < bi ? k : (k < bj ? k - 1 : k - 2)] = v;
    }
    if (threadIdx.x == 0) {
        alive_out[m - 2] = si;
        J.slot_node[si] = n + round;
    }
}

__global__ void nj_root_kernel(NjLaunch J, const int* __restrict__ alive, int round)
{
    if (threadIdx.x == 0) {
        J.children[2 * round] = J.slot_node[alive[0]];
        J.children[2 * round + 1] = J.slot_node[alive[1]];
    }
}

}  // namespace

hipError_t kmer_distances_launch(const KmerLaunch& K, int max_len, hipStream_t st, LaunchHooks* hooks)
{
    if (K.n <= 0) return hipSuccess;
    if (hooks) hooks->before(st);
    if (max_len > 0 && max_len >= K.k)
        hipLaunchKernelGGL(kmer_bits_kernel, dim3(unsigned((max_len + 255) / 256), unsigned(K.n)), dim3(256), 0, st,
                           K);
    hipLaunchKernelGGL(kmer_count_kernel, dim3(unsigned(K.n)), dim3(256), 0, st, K);
    const int nblk = (K.n + 15) / 16;
    if (K.n > 1)
        hipLaunchKernelGGL(kmer_pairs_kernel, dim3(unsigned(nblk * (nblk + 1) / 2)), dim3(256), 0, st, K, nblk);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

hipError_t nj_launch(const NjLaunch& J, hipStream_t st, LaunchHooks* hooks)
{
    const int n = J.n;
    if (n < 2) return hipSuccess;
    if (hooks) hooks->before(st);
    const int64_t nn = int64_t(n) * n;
    hipLaunchKernelGGL(nj_init_kernel, dim3(unsigned((nn + 255) / 256)), dim3(256), 0, st, J);
    int* buf[2] = {J.alive0, J.alive1};
    int round = 0;
    for (int m = n; m > 2; --m, ++round) {
        const int* in = buf[round & 1];
        int* out = buf[(round + 1) & 1];
        hipLaunchKernelGGL(nj_rowsum_kernel, dim3(unsigned((m + 63) / 64)), dim3(64), 0, st, J, in, m);
        hipLaunchKernelGGL(nj_argmin_kernel, dim3(unsigned(m - 1)), dim3(256), 0, st, J, in, m);
        hipLaunchKernelGGL(nj_merge_kernel, dim3(1), dim3(256), 0, st, J, in, out, m, round);
    }
    hipLaunchKernelGGL(nj_root_kernel, dim3(1), dim3(64), 0, st, J, buf[round & 1], round);
    if (hooks) hooks->after(st);
    return hipGetLastError();
}

}  // namespace ga
