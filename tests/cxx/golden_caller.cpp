// A C++ caller of the drop-in surface (include/genoaligner/api.hpp), in the
// shape of the reference's own callers (examples/align_fasta.cpp:55-68,
// examples/align_sw.cpp:61-74): vectors of requests holding pointers,
// align_batch / align_sw_batch / align / align_sw, status checked first.
//
//   golden_caller GOLDEN_DIR
//
// Part 1 replays every golden vector of GOLDEN_DIR (wfa_trace.tsv,
// sw_trace.tsv, sw_score_w64.tsv: written from the reference's own kernels)
// through the C++ surface and counts the results that differ.  Part 2 runs the
// two examples' request shapes on GOLDEN_DIR/mtdna_human.fa and prints every
// result ("EX ..." lines) for the caller's test to check against the oracle.
// Exit status 0 iff every golden vector matched.
#include "genoaligner/api.hpp"
#include "genoaligner/io/fasta.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

using namespace genoaligner;

namespace {

std::vector<std::vector<std::string>> read_tsv(const std::string& path)
{
    std::vector<std::vector<std::string>> rows;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::vector<std::string> f;
        std::stringstream ss(line);
        std::string x;
        while (std::getline(ss, x, '\t')) f.push_back(x == "-" ? std::string() : x);
        rows.push_back(f);
    }
    return rows;
}

int bad = 0, checked = 0;

void expect(bool ok, const char* what, size_t row)
{
    ++checked;
    if (!ok) {
        if (bad < 20) std::printf("MISMATCH %s row %zu\n", what, row);
        ++bad;
    }
}

void golden_wfa(const std::string& dir)
{
    // smax, pattern, text, score, cigar, flat score
    const auto rows = read_tsv(dir + "/wfa_trace.tsv");
    std::map<int, std::vector<size_t>> by_smax;
    for (size_t i = 0; i < rows.size(); ++i) by_smax[std::atoi(rows[i][0].c_str())].push_back(i);
    for (const auto& kv : by_smax) {
        for (int with_cigar = 0; with_cigar < 2; ++with_cigar) {
            std::vector<AlignRequest> reqs;
            for (size_t i : kv.second) {
                AlignRequest r;
                r.pattern = rows[i][1].data();
                r.pattern_len = int(rows[i][1].size());
                r.text = rows[i][2].data();
                r.text_len = int(rows[i][2].size());
                r.smax = kv.first;
                r.with_cigar = with_cigar != 0;
                reqs.push_back(r);
            }
            const BatchResult b = align_batch(reqs);
            if (!b.ok()) {
                std::printf("align_batch failed: %s\n", b.error ? b.error : "?");
                ++bad;
                continue;
            }
            int resolved = 0;
            for (size_t k = 0; k < kv.second.size(); ++k) {
                const size_t i = kv.second[k];
                const AlignResult& r = b.results[k];
                const int want = std::atoi(rows[i][3].c_str());
                resolved += want >= 0;
                expect(r.score == want && r.resolved == (want >= 0), "wfa score", i);
                if (with_cigar) {
                    expect(r.cigar == rows[i][4], "wfa cigar", i);
                    if (!r.cigar.empty()) expect(r.rescore_ok && r.wellformed_ok, "wfa validators", i);
                } else {
                    expect(r.cigar.empty(), "wfa score-only cigar", i);
                }
            }
            expect(b.resolved_count == resolved && b.unresolved_count == int(kv.second.size()) - resolved,
                   "wfa counts", kv.second[0]);
        }
    }
    // the one-pair entry point on the first rows
    for (size_t i = 0; i < rows.size() && i < 40; ++i) {
        AlignRequest r;
        r.pattern = rows[i][1].data();
        r.pattern_len = int(rows[i][1].size());
        r.text = rows[i][2].data();
        r.text_len = int(rows[i][2].size());
        r.smax = std::atoi(rows[i][0].c_str());
        const AlignResult a = align(r);
        expect(a.score == std::atoi(rows[i][3].c_str()) && a.cigar == rows[i][4], "align", i);
    }
}

void golden_sw(const std::string& dir)
{
    // match, mismatch, gap_open, gap_extend, text, pattern, score, start_i, start_j, end_i, end_j, cigar
    const auto rows = read_tsv(dir + "/sw_trace.tsv");
    std::map<std::tuple<int, int, int, int>, std::vector<size_t>> by_scheme;
    for (size_t i = 0; i < rows.size(); ++i)
        by_scheme[std::make_tuple(std::atoi(rows[i][0].c_str()), std::atoi(rows[i][1].c_str()),
                                  std::atoi(rows[i][2].c_str()), std::atoi(rows[i][3].c_str()))]
            .push_back(i);
    for (const auto& kv : by_scheme) {
        std::vector<SWRequest> reqs;
        for (size_t i : kv.second) {
            SWRequest r;
            r.text = rows[i][4].data();
            r.text_len = int(rows[i][4].size());
            r.pattern = rows[i][5].data();
            r.pattern_len = int(rows[i][5].size());
            r.scoring = {std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                         std::get<3>(kv.first)};
            r.with_cigar = true;
            reqs.push_back(r);
        }
        const SWBatchResult b = align_sw_batch(reqs);
        if (!b.ok()) {
            std::printf("align_sw_batch failed: %s\n", b.error ? b.error : "?");
            ++bad;
            continue;
        }
        for (size_t k = 0; k < kv.second.size(); ++k) {
            const size_t i = kv.second[k];
            const SWAlignResult& r = b.results[k];
            expect(r.score == std::atoi(rows[i][6].c_str()) && r.start_i == std::atoi(rows[i][7].c_str()) &&
                       r.start_j == std::atoi(rows[i][8].c_str()) && r.end_i == std::atoi(rows[i][9].c_str()) &&
                       r.end_j == std::atoi(rows[i][10].c_str()) && r.cigar == rows[i][11],
                   "sw trace", i);
            expect(r.resolved && r.rescore_ok && r.wellformed_ok && !r.too_large, "sw flags", i);
        }
        const SWAlignResult one = align_sw(reqs[0]);
        expect(one.score == b.results[0].score && one.cigar == b.results[0].cigar, "align_sw", kv.second[0]);
    }
    // score-only: score and the end cell by the wave64 lane rule
    const auto srows = read_tsv(dir + "/sw_score_w64.tsv");
    for (size_t i = 0; i < srows.size(); ++i) {
        SWRequest r;
        r.text = srows[i][4].data();
        r.text_len = int(srows[i][4].size());
        r.pattern = srows[i][5].data();
        r.pattern_len = int(srows[i][5].size());
        r.scoring = {std::atoi(srows[i][0].c_str()), std::atoi(srows[i][1].c_str()),
                     std::atoi(srows[i][2].c_str()), std::atoi(srows[i][3].c_str())};
        r.with_cigar = false;
        const SWBatchResult b = align_sw_batch(std::vector<SWRequest>{r});
        expect(b.ok() && b.results[0].score == std::atoi(srows[i][6].c_str()) &&
                   b.results[0].end_i == std::atoi(srows[i][7].c_str()) &&
                   b.results[0].end_j == std::atoi(srows[i][8].c_str()) && b.results[0].cigar.empty(),
               "sw score-only", i);
    }
}

// examples/align_fasta.cpp:44-68 and examples/align_sw.cpp:48-74, on the
// human mtDNA record
int examples(const std::string& dir)
{
    std::vector<io::FastaRecord> recs;
    std::string err;
    if (!io::read_fasta_file(dir + "/mtdna_human.fa", &recs, &err) || recs.empty()) {
        std::printf("cannot read mtdna_human.fa: %s\n", err.c_str());
        return 1;
    }
    const std::string& seq = recs[0].sequence;

    struct Window { std::string pattern, text; };
    std::vector<Window> wins;
    for (int i = 0; i < 3; ++i) {
        Window w;
        w.pattern = seq.substr(size_t(1000 + i * 700), 200);
        w.text = w.pattern;
        w.text[size_t(20 + i * 30)] = w.text[size_t(20 + i * 30)] == 'A' ? 'C' : 'A';
        wins.push_back(w);
    }
    std::vector<AlignRequest> reqs;
    for (const Window& w : wins) {
        AlignRequest r;
        r.pattern = w.pattern.data();
        r.pattern_len = int(w.pattern.size());
        r.text = w.text.data();
        r.text_len = int(w.text.size());
        r.smax = 64;
        r.with_cigar = true;
        reqs.push_back(r);
    }
    const BatchResult b = align_batch(reqs);
    if (!b.ok()) {
        std::printf("align_batch failed: %s\n", b.error ? b.error : "?");
        return 1;
    }
    for (size_t i = 0; i < b.results.size(); ++i)
        std::printf("EX wfa %zu %s %s %d %d %s %d %d\n", i, wins[i].pattern.c_str(), wins[i].text.c_str(),
                    b.results[i].score, int(b.results[i].resolved), b.results[i].cigar.c_str(),
                    int(b.results[i].rescore_ok), int(b.results[i].wellformed_ok));

    std::vector<Window> swins;
    for (int i = 0; i < 3; ++i) {
        Window w;
        const std::string core = seq.substr(size_t(400 + i * 200), 80);
        w.pattern = std::string("TTTTTTTTTT") + core + std::string("TTTTTTTTTT");
        w.text = std::string("AAAAAAAAAA") + core + std::string("AAAAAAAAAA");
        w.text[size_t(15 + i * 7)] = 'N';
        swins.push_back(w);
    }
    std::vector<SWRequest> sreqs;
    for (const Window& w : swins) {
        SWRequest r;
        r.pattern = w.pattern.data();
        r.pattern_len = int(w.pattern.size());
        r.text = w.text.data();
        r.text_len = int(w.text.size());
        r.scoring = {2, -3, 3, 1};
        r.with_cigar = true;
        sreqs.push_back(r);
    }
    const SWBatchResult sb = align_sw_batch(sreqs);
    if (!sb.ok()) {
        std::printf("align_sw_batch failed: %s\n", sb.error ? sb.error : "?");
        return 1;
    }
    for (size_t i = 0; i < sb.results.size(); ++i) {
        const SWAlignResult& r = sb.results[i];
        std::printf("EX sw %zu %s %s %d %d %d %d %d %s %d %d\n", i, swins[i].text.c_str(),
                    swins[i].pattern.c_str(), r.score, r.start_i, r.start_j, r.end_i, r.end_j, r.cigar.c_str(),
                    int(r.rescore_ok), int(r.wellformed_ok));
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: golden_caller GOLDEN_DIR\n");
        return 2;
    }
    if (!device_available()) {
        std::printf("no device\n");
        return 3;
    }
    std::printf("backend %s device %s\n", backend_name(), device_name());
    const std::string dir = argv[1];
    golden_wfa(dir);
    golden_sw(dir);
    const int ex = examples(dir);
    std::printf("GOLDEN checked %d bad %d\n", checked, bad);
    return bad == 0 && ex == 0 ? 0 : 1;
}
