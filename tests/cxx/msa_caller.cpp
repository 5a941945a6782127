// A C++ caller of include/genoaligner/msa/msa.hpp (the reference's genomsa
// surface): every golden vector of GOLDEN_DIR/msa_pp.tsv through
// profile_update_counts + align_profiles (and the batch extension), and of
// kmer_nj.tsv through kmer_distances + nj_tree, compared bit for bit; then a
// small progressive merge (merge_profiles over the guide tree's order).
//
//   msa_caller GOLDEN_DIR        exit 0 iff every vector matched
#include "genoaligner/msa/msa.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::vector<std::string> split(const std::string& s, char c)
{
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string x;
    while (std::getline(ss, x, c)) out.push_back(x);
    if (!s.empty() && s.back() == c) out.push_back("");
    return out;
}

float f32(const std::string& h)
{
    const uint32_t u = uint32_t(std::strtoul(h.c_str(), nullptr, 16));
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

genomsa::Params params(const std::string& tag)
{
    genomsa::Params P;
    if (tag == "legacy") P.psgp = false;
    if (tag == "global") P.free_end_gaps = false;
    if (tag == "protein") P = genomsa::protein_params();
    if (tag == "protein_semi") {
        P = genomsa::protein_params();
        P.free_end_gaps = true;
    }
    return P;
}

genomsa::Profile prof(const std::vector<std::string>& rows, int alpha)
{
    genomsa::Profile p;
    p.alpha = alpha;
    p.rows = rows;
    genomsa::profile_update_counts(p);
    return p;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    if (!ga_device_available()) {
        std::printf("no device\n");
        return 3;
    }
    const std::string dir = argv[1];
    int checked = 0, bad = 0;
    std::ifstream in(dir + "/msa_pp.tsv");
    std::string line;
    std::vector<genomsa::Profile> As, Bs;
    std::vector<genomsa::AlignResult> want;
    std::vector<std::string> tags;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        const auto f = split(line, '\t');
        const genomsa::Params P = params(f[0]);
        const genomsa::Profile A = prof(split(f[1], ','), P.alpha), B = prof(split(f[2], ','), P.alpha);
        genomsa::AlignResult w;
        w.score = f32(f[3]);
        w.ai = std::atoi(f[4].c_str()); w.aj = std::atoi(f[5].c_str());
        w.bi = std::atoi(f[6].c_str()); w.bj = std::atoi(f[7].c_str());
        w.cigar = f[8] == "-" ? "" : f[8];
        const genomsa::AlignResult g = genomsa::align_profiles(A, B, P);
        ++checked;
        if (std::memcmp(&g.score, &w.score, 4) != 0 || g.cigar != w.cigar || g.ai != w.ai || g.aj != w.aj ||
            g.bi != w.bi || g.bj != w.bj) {
            if (bad < 10) std::printf("MISMATCH msa row %d (%s)\n", checked, f[0].c_str());
            ++bad;
        }
        if (f[0] == "default") { As.push_back(A); Bs.push_back(B); want.push_back(w); }
    }
    // the same default-parameter pairs in one batch
    std::vector<std::pair<const genomsa::Profile*, const genomsa::Profile*>> pairs;
    for (size_t q = 0; q < As.size(); ++q) pairs.emplace_back(&As[q], &Bs[q]);
    const auto got = genomsa::align_profiles_batch(pairs, genomsa::Params{});
    for (size_t q = 0; q < got.size(); ++q) {
        ++checked;
        if (std::memcmp(&got[q].score, &want[q].score, 4) != 0 || got[q].cigar != want[q].cigar) ++bad;
    }
    std::ifstream kin(dir + "/kmer_nj.tsv");
    while (std::getline(kin, line)) {
        if (line.empty() || line[0] == '#') continue;
        const auto f = split(line, '\t');
        const int k = std::atoi(f[0].c_str()), alpha = std::atoi(f[1].c_str());
        std::vector<std::string> seqs;
        for (auto& s : split(f[2], ',')) seqs.push_back(s == "-" ? "" : s);
        const std::vector<float> D = genomsa::kmer_distances(seqs, k, alpha);
        const auto dh = split(f[3], ',');
        ++checked;
        bool ok = D.size() == dh.size();
        for (size_t q = 0; ok && q < D.size(); ++q) {
            const float w = f32(dh[q]);
            ok = std::memcmp(&D[q], &w, 4) == 0;
        }
        const genomsa::Tree t = genomsa::nj_tree(D, int(seqs.size()));
        const auto ch = split(f[4], ',');
        const int n = int(seqs.size());
        ok = ok && int(t.nodes.size()) == 2 * n - 1 && t.root == 2 * n - 2;
        for (int u = 0; ok && u < n - 1; ++u) {
            const auto pr = split(ch[size_t(u)], ':');
            ok = t.nodes[size_t(n + u)].left == std::atoi(pr[0].c_str()) &&
                 t.nodes[size_t(n + u)].right == std::atoi(pr[1].c_str());
        }
        if (!ok) {
            if (bad < 10) std::printf("MISMATCH kmer/nj row (k %d, %d sequences)\n", k, n);
            ++bad;
        }
    }
    // a progressive merge of four sequences over their guide tree
    const std::vector<std::string> seqs = {"ACGTACGTTTGACCA", "ACGTACGTTGACCA", "ACGAACGTTTGACGA", "TTACGTACGTTTGA"};
    const genomsa::Tree t = genomsa::nj_tree(genomsa::kmer_distances(seqs, 3), 4);
    std::vector<genomsa::Profile> node(t.nodes.size());
    for (int i = 0; i < 4; ++i) node[size_t(i)] = genomsa::profile_from_seq(seqs[size_t(i)]);
    for (size_t u = 4; u < t.nodes.size(); ++u) {
        const auto& L = node[size_t(t.nodes[u].left)];
        const auto& R = node[size_t(t.nodes[u].right)];
        node[u] = genomsa::merge_profiles(L, R, genomsa::align_profiles(L, R, genomsa::Params{}));
    }
    for (const auto& r : node[size_t(t.root)].rows) std::printf("MSA %s\n", r.c_str());
    std::printf("GOLDEN checked %d bad %d\n", checked, bad);
    return bad == 0 ? 0 : 1;
}
