// genoaligner -- MSA building blocks, MI355X (gfx950) build.
//
// The reference's genomsa surface for the steps that run on the GPU
// (reference: include/genoaligner/msa/msa.hpp), kept name for name and
// default for default so a caller of these functions recompiles unchanged:
//   Params (the fields the DP and the guide tree read), protein_params,
//   Profile, profile_from_seq, profile_update_counts, merge_profiles,
//   AlignResult, align_profiles (msa.hpp:339-346),
//   kmer_distances, kmer_distances_mt (msa.hpp:329-333),
//   Node, Tree, nj_tree, nj_tree_mt, nj_tree_gpu (msa.hpp:336-357).
// Header-only over the C ABI in genoaligner_c.h, which does the work:
// align_profiles, kmer_distances and nj_tree run on the GPU and return the
// reference's results bit for bit.  Extension: align_profiles_batch (many
// independent pairs in one launch, as the reference's level-batched driver
// src/msa/msa_gpu.cpp launches them).  Failures throw std::runtime_error
// (the reference's functions have no error channel); nj_tree_gpu reports
// through `err` and returns an empty Tree, as the reference's does.
#ifndef GENOALIGNER_MSA_MSA_HPP
#define GENOALIGNER_MSA_MSA_HPP

#include "../genoaligner_c.h"

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace genomsa {

constexpr int MSA_MAX_SYMS = 65;

struct Params {
    int   kmer_k = 5;
    int   alpha = 4;                 // 4 = DNA (IUPAC), 20 = protein
    float match = 2.0f;
    float ts = -1.0f;
    float tv = -2.0f;
    std::array<float, MSA_MAX_SYMS * MSA_MAX_SYMS> sub{};   // row-major alpha x alpha (alpha 20)
    float gap_open = 3.0f;
    float gap_extend = 1.0f;
    bool  free_end_gaps = true;
    bool  psgp = true;
    float psgp_scale = 0.5f;
    float psgp_min_open = 0.1f;
    float psgp_min_ext = 0.2f;
};

// BLOSUM62 in ARNDCQEGHILKMFPSTWYV order, as the reference's MSA engine holds
// it (src/msa/msa_ref.cpp:117-138).
inline Params protein_params()
{
    static const signed char b62[20][20] = {
        {4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0},
        {-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3},
        {-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3},
        {-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3},
        {0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1},
        {-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2},
        {-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2},
        {0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3},
        {-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3},
        {-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3},
        {-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1},
        {-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2},
        {-1, -1, -2, -3, -1, 0, -2, -3, -1, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1},
        {-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1},
        {-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2},
        {1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2},
        {0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0},
        {-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3},
        {-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1},
        {0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4}};
    Params P;
    P.alpha = 20;
    P.kmer_k = 2;
    P.gap_open = 11.0f;
    P.gap_extend = 1.0f;
    P.free_end_gaps = false;
    for (int a = 0; a < 20; ++a)
        for (int b = 0; b < 20; ++b) P.sub[size_t(a * 20 + b)] = float(b62[a][b]);
    return P;
}

// cols[c][s]: fraction of symbol s in column c, cols[c][alpha] the gap
// fraction; occ[c] the non-gap fraction.
struct Profile {
    std::vector<std::array<float, MSA_MAX_SYMS + 1>> cols;
    std::vector<float> occ;
    std::vector<std::string> rows;
    std::vector<int> ids;
    int nseq = 0;
    int alpha = 4;
    int ncols() const { return int(cols.size()); }
};

inline void profile_update_counts(Profile& p)
{
    const int L = p.rows.empty() ? 0 : int(p.rows[0].size());
    std::string flat;
    for (const auto& r : p.rows) {
        if (int(r.size()) != L) throw std::runtime_error("genomsa: profile rows differ in length");
        flat += r;
    }
    std::vector<float> c(static_cast<size_t>(L) * size_t(p.alpha)), occ(static_cast<size_t>(L)),
        gap(static_cast<size_t>(L));
    if (L > 0 && ga_msa_profile_from_rows(flat.data(), int32_t(p.rows.size()), L, p.alpha, c.data(), occ.data(),
                                          gap.data()) != GA_OK)
        throw std::runtime_error("genomsa: invalid profile (alphabet must be 4 or 20)");
    p.nseq = int(p.rows.size());
    p.cols.assign(size_t(L), {});
    p.occ = occ;
    for (int i = 0; i < L; ++i) {
        for (int s = 0; s < p.alpha; ++s) p.cols[size_t(i)][size_t(s)] = c[size_t(i) * p.alpha + s];
        p.cols[size_t(i)][size_t(p.alpha)] = gap[size_t(i)];
    }
}

inline Profile profile_from_seq(const std::string& seq, int alpha = 4)
{
    Profile p;
    p.alpha = alpha;
    p.rows.push_back(seq);
    p.ids.push_back(0);
    profile_update_counts(p);
    return p;
}

struct AlignResult {
    float score = 0;
    std::string cigar;          // 'M', 'I' (A column only), 'D' (B column only)
    int ai = 0, aj = 0;
    int bi = 0, bj = 0;
};

inline ga_msa_params c_params(const Params& P)
{
    ga_msa_params c{};
    c.match = P.match; c.ts = P.ts; c.tv = P.tv;
    c.gap_open = P.gap_open; c.gap_extend = P.gap_extend;
    c.free_end_gaps = P.free_end_gaps ? 1 : 0;
    c.psgp = P.psgp ? 1 : 0;
    c.psgp_scale = P.psgp_scale; c.psgp_min_open = P.psgp_min_open; c.psgp_min_ext = P.psgp_min_ext;
    c.alpha = P.alpha;
    if (P.alpha == 20)
        for (int i = 0; i < 400; ++i) c.sub[i] = P.sub[size_t(i)];
    return c;
}

// Extension: many independent pairs in one launch.
inline std::vector<AlignResult> align_profiles_batch(const std::vector<std::pair<const Profile*, const Profile*>>& pairs,
                                                     const Params& P)
{
    const size_t n = pairs.size();
    std::vector<AlignResult> out(n);
    if (n == 0) return out;
    std::vector<std::vector<float>> keep;
    std::vector<ga_msa_profile> A(n), B(n);
    int64_t need = 0;
    auto view = [&](const Profile* p) {
        std::vector<float> c(static_cast<size_t>(p->ncols()) * size_t(P.alpha));
        for (int i = 0; i < p->ncols(); ++i)
            for (int s = 0; s < P.alpha; ++s) c[size_t(i) * P.alpha + s] = p->cols[size_t(i)][size_t(s)];
        keep.push_back(std::move(c));
        return ga_msa_profile{keep.back().data(), p->occ.data(), int32_t(p->ncols())};
    };
    keep.reserve(2 * n);
    for (size_t q = 0; q < n; ++q) {
        A[q] = view(pairs[q].first);
        B[q] = view(pairs[q].second);
        need += A[q].len + B[q].len;
    }
    std::vector<ga_msa_pp_res> res(n);
    std::vector<char> arena(static_cast<size_t>(need > 0 ? need : 1));
    const ga_msa_params c = c_params(P);
    const char* err = nullptr;
    if (ga_msa_align_profiles(A.data(), B.data(), int64_t(n), &c, res.data(), arena.data(), need, &err) != GA_OK)
        throw std::runtime_error(std::string("genomsa::align_profiles: ") + (err ? err : "failed"));
    for (size_t q = 0; q < n; ++q) {
        out[q].score = res[q].score;
        out[q].cigar.assign(arena.data() + res[q].cigar_off, size_t(res[q].cigar_len));
        out[q].ai = res[q].ai; out[q].aj = res[q].aj; out[q].bi = res[q].bi; out[q].bj = res[q].bj;
    }
    return out;
}

inline AlignResult align_profiles(const Profile& A, const Profile& B, const Params& P)
{
    return align_profiles_batch({{&A, &B}}, P)[0];
}

// Interleave two profiles' rows by their alignment, the columns outside the
// span as gap blocks (msa_ref.cpp:1177-1220), and count the result.
inline Profile merge_profiles(const Profile& A, const Profile& B, const AlignResult& aln)
{
    Profile out;
    out.alpha = A.alpha;
    const int la = A.ncols(), lb = B.ncols();
    for (const auto& row : A.rows) {
        std::string s = row.substr(0, size_t(aln.ai)) + std::string(size_t(aln.aj), '-');
        int ca = aln.ai;
        for (char op : aln.cigar) s.push_back(op == 'D' ? '-' : row[size_t(ca++)]);
        s += row.substr(size_t(ca)) + std::string(size_t(lb - aln.bj), '-');
        out.rows.push_back(s);
    }
    for (const auto& row : B.rows) {
        std::string s = std::string(size_t(aln.ai), '-') + row.substr(0, size_t(aln.aj));
        int cb = aln.aj;
        for (char op : aln.cigar) s.push_back(op == 'I' ? '-' : row[size_t(cb++)]);
        s += std::string(size_t(la - aln.bi), '-') + row.substr(size_t(cb));
        out.rows.push_back(s);
    }
    out.ids = A.ids;
    out.ids.insert(out.ids.end(), B.ids.begin(), B.ids.end());
    profile_update_counts(out);
    return out;
}

inline std::vector<float> kmer_distances(const std::vector<std::string>& seqs, int k, int alpha = 4)
{
    const int n = int(seqs.size());
    std::vector<float> D(n > 1 ? size_t(n) * size_t(n - 1) / 2 : 0);
    if (n < 2) return D;
    std::string flat;
    std::vector<int64_t> off(static_cast<size_t>(n));
    std::vector<int32_t> len(static_cast<size_t>(n));
    for (int s = 0; s < n; ++s) {
        off[size_t(s)] = int64_t(flat.size());
        len[size_t(s)] = int32_t(seqs[size_t(s)].size());
        flat += seqs[size_t(s)];
    }
    const char* err = nullptr;
    if (ga_kmer_distances(flat.data(), off.data(), len.data(), n, k, alpha, D.data(), &err) != GA_OK)
        throw std::runtime_error(std::string("genomsa::kmer_distances: ") + (err ? err : "failed"));
    return D;
}
inline std::vector<float> kmer_distances_mt(const std::vector<std::string>& seqs, int k, int /*threads*/,
                                            int alpha = 4)
{
    return kmer_distances(seqs, k, alpha);
}

struct Node { int left = -1, right = -1; };
struct Tree {
    std::vector<Node> nodes;
    int root = -1;
};

inline Tree nj_tree_gpu(const std::vector<float>& dist_packed, int n, std::string& err)
{
    Tree t;
    err.clear();
    if (n == 1) { t.nodes.resize(1); t.root = 0; return t; }
    std::vector<int32_t> ch(static_cast<size_t>(n > 1 ? 2 * (n - 1) : 0));
    const char* e = nullptr;
    if (n < 2 || dist_packed.size() != size_t(n) * size_t(n - 1) / 2 ||
        ga_nj_tree(dist_packed.data(), n, ch.data(), &e) != GA_OK) {
        err = e ? e : "genomsa::nj_tree: distance array is not n (n - 1) / 2 long";
        return Tree{};
    }
    t.nodes.resize(size_t(n));
    for (int u = 0; u < n - 1; ++u) t.nodes.push_back(Node{ch[size_t(2 * u)], ch[size_t(2 * u + 1)]});
    t.root = 2 * n - 2;
    return t;
}
inline Tree nj_tree(const std::vector<float>& dist_packed, int n)
{
    std::string err;
    Tree t = nj_tree_gpu(dist_packed, n, err);
    if (!err.empty()) throw std::runtime_error(err);
    return t;
}
inline Tree nj_tree_mt(const std::vector<float>& dist_packed, int n, int /*threads*/)
{
    return nj_tree(dist_packed, n);
}

}  // namespace genomsa

#endif  // GENOALIGNER_MSA_MSA_HPP
