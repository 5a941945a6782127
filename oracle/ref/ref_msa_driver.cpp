// ref_msa_driver.cpp -- C entry points over the reference's MSA building
// blocks (src/msa/msa_ref.cpp: profile_update_counts, align_profiles,
// kmer_distances, nj_tree) and its profile-profile kernel body
// (include/genoaligner/backend/msa_pp_kernel_impl.hip, run one thread per
// pair on the CPU through the reference's HIP CPU shim, as its
// tests/parity/msa_pp_parity.cpp runs it).
//
// TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles this file together
// with the reference's msa_ref.cpp, where it lies, into
// oracle/_ref/libga_ref_msa.so.  Nothing here is copied from the reference:
// it converts plain arrays to the reference's types, calls its functions and
// copies the results out.  tests/golden/make_golden.py writes the MSA golden
// vectors from it and tests/test_msa_oracle.py checks the C restatement
// (oracle/ga_msa_oracle.c) against it.
#include <genoaligner/msa/msa.hpp>
#include <hip/hip_runtime.h>
#include <genoaligner/backend/msa_pp_kernel_impl.hip>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

// launch globals of the shim (this translation unit owns them)
thread_local uint3 threadIdx{0, 0, 0};
uint3 blockIdx{0, 0, 0};
dim3 blockDim{1, 1, 1};
dim3 gridDim{1, 1, 1};

namespace {

struct Prm {                     // the layout of ga_oracle_msa_params
    float match, ts, tv;
    float gap_open, gap_extend;
    int free_end_gaps;
    int psgp;
    float psgp_scale, psgp_min_open, psgp_min_ext;
    int alpha;
    const float* sub;
};

genomsa::Params params(const Prm* p)
{
    genomsa::Params P;
    P.match = p->match; P.ts = p->ts; P.tv = p->tv;
    P.gap_open = p->gap_open; P.gap_extend = p->gap_extend;
    P.free_end_gaps = p->free_end_gaps != 0;
    P.psgp = p->psgp != 0;
    P.psgp_scale = p->psgp_scale; P.psgp_min_open = p->psgp_min_open; P.psgp_min_ext = p->psgp_min_ext;
    P.alpha = p->alpha;
    if (p->alpha > 4 && p->sub)
        for (int i = 0; i < p->alpha * p->alpha; ++i) P.sub[size_t(i)] = p->sub[i];
    return P;
}

genomsa::Profile profile(const char* rows, int nrows, int L, int alpha)
{
    genomsa::Profile p;
    p.alpha = alpha;
    p.nseq = nrows;
    for (int r = 0; r < nrows; ++r) {
        p.rows.emplace_back(rows + size_t(r) * size_t(L), size_t(L));
        p.ids.push_back(r);
    }
    genomsa::profile_update_counts(p);
    return p;
}

}  // namespace

extern "C" {

int ga_ref_msa_profile(const char* rows, int nrows, int L, int alpha, float* cols, float* occ)
{
    const genomsa::Profile p = profile(rows, nrows, L, alpha);
    for (int i = 0; i < L; ++i) {
        for (int b = 0; b < alpha; ++b) cols[size_t(i) * alpha + b] = p.cols[size_t(i)][size_t(b)];
        occ[i] = p.occ[size_t(i)];
    }
    return 0;
}

// align_profiles over two profiles given by their aligned rows
int ga_ref_msa_align(const char* rowsA, int nA, int LA, const char* rowsB, int nB, int LB, const Prm* prm,
                     float* score, int32_t* span, char* cigar, int32_t* cigar_len)
{
    const genomsa::Params P = params(prm);
    const genomsa::Profile A = profile(rowsA, nA, LA, prm->alpha), B = profile(rowsB, nB, LB, prm->alpha);
    const genomsa::AlignResult r = genomsa::align_profiles(A, B, P);
    *score = r.score;
    span[0] = r.ai; span[1] = r.aj; span[2] = r.bi; span[3] = r.bj;
    std::memcpy(cigar, r.cigar.data(), r.cigar.size());
    *cigar_len = int32_t(r.cigar.size());
    return 0;
}

// the shipped kernel body, one thread per pair, on the same two profiles
int ga_ref_msa_pp_kernel(const char* rowsA, int nA, int LA, const char* rowsB, int nB, int LB, const Prm* prm,
                         float* score, int32_t* span, char* cigar, int32_t* cigar_len)
{
    using namespace genoaligner;
    const genomsa::Params P = params(prm);
    const genomsa::Profile A = profile(rowsA, nA, LA, prm->alpha), B = profile(rowsB, nB, LB, prm->alpha);
    std::vector<float> ca[MSA_MAX_SYMS], cb[MSA_MAX_SYMS];
    std::vector<float> oa = A.occ, ob = B.occ;
    MsaPPPair pp{};
    for (int b = 0; b < P.alpha; ++b) {
        for (const auto& c : A.cols) ca[b].push_back(c[size_t(b)]);
        for (const auto& c : B.cols) cb[b].push_back(c[size_t(b)]);
        pp.A.cols[b] = ca[b].data();
        pp.B.cols[b] = cb[b].data();
    }
    pp.A.occ = oa.data(); pp.A.len = A.ncols();
    pp.B.occ = ob.data(); pp.B.len = B.ncols();
    MsaPPParams k{P.match, P.ts, P.tv, P.gap_open, P.gap_extend, P.free_end_gaps ? 1 : 0, P.psgp ? 1 : 0,
                  P.psgp_scale, P.psgp_min_open, P.psgp_min_ext, P.alpha, {}, nullptr};
    if (P.alpha > 4 && P.alpha <= MSA_SUB_INLINE_SYMS) std::memcpy(k.sub, P.sub.data(), sizeof(k.sub));
    const size_t dsz = size_t(A.ncols() + 1) * size_t(B.ncols() + 1);
    const size_t ssz = size_t(3) * size_t(B.ncols() + 1) + 2 * size_t(A.ncols() + 1);
    const int cap = A.ncols() + B.ncols() + 4;
    std::vector<uint8_t> dirs(dsz), cg(static_cast<size_t>(cap));
    std::vector<float> scr(ssz);
    int meta[2] = {0, 0};
    const size_t zero[1] = {0}, avd[1] = {dsz}, avs[1] = {ssz};
    MsaPPResult res{};
    blockIdx = uint3{0, 0, 0}; threadIdx = uint3{0, 0, 0};
    blockDim = dim3{1, 1, 1}; gridDim = dim3{1, 1, 1};
    msa_pp_trace_kernel(&pp, k, &res, dirs.data(), zero, avd, cg.data(), zero, meta, scr.data(), zero, avs, 1, cap);
    static const char ops[] = "MID";
    for (int q = 0; q < meta[1]; ++q) cigar[q] = ops[cg[size_t(meta[1] - 1 - q)]];
    *cigar_len = meta[1];
    *score = res.score;
    span[0] = res.ai; span[1] = res.aj; span[2] = res.bi; span[3] = res.bj;
    return 0;
}

// kmer_distances over n sequences packed back to back
int ga_ref_kmer_distances(const char* seqs, const int64_t* off, const int32_t* len, int n, int k, int alpha,
                          float* D)
{
    std::vector<std::string> s;
    for (int i = 0; i < n; ++i) s.emplace_back(seqs + off[i], size_t(len[i]));
    const std::vector<float> d = genomsa::kmer_distances(s, k, alpha);
    std::memcpy(D, d.data(), d.size() * sizeof(float));
    return 0;
}

int ga_ref_nj_tree(const float* D, int n, int32_t* children)
{
    const std::vector<float> d(D, D + size_t(n) * size_t(n - 1) / 2);
    const genomsa::Tree t = genomsa::nj_tree(d, n);
    for (int u = n; u < int(t.nodes.size()); ++u) {
        children[2 * (u - n)] = t.nodes[size_t(u)].left;
        children[2 * (u - n) + 1] = t.nodes[size_t(u)].right;
    }
    return t.root;
}

}  // extern "C"
