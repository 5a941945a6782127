// gene_qc.cpp -- the per-record CIGAR walk of the gene QC task (reference
// tools/gene_qc.cpp:146-217), on the host over the CIGAR arena that
// ga_align_sw_batch fills, one record per task of the host work pool.
//
// For a record aligned with score > 0 (text = the record, pattern = the bait
// read in `frame`): the text coverage of the span; codons of the bait
// (positions j with (j - frame) mod 3 == 0, 1, 2) whose three columns pair
// contiguous text bases through M/X ops, and the stops among them (code 1:
// TAA TAG TGA; code 2, vertebrate mitochondrial: TAA TAG AGA AGG; any other
// code: TAA TAG); a stop in
// the span's last complete codon is the gene's own terminator (stop_end); the
// indel runs whose length is not a multiple of 3, and the net indel length
// mod 3 as -1, 0 or +1.
#include "../../include/genoaligner/genoaligner_c.h"

#include "ga_device.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

bool is_stop(uint8_t a, uint8_t b, uint8_t c, int code)
{
    if (a == 'T' && b == 'A' && (c == 'A' || c == 'G')) return true;                 // TAA TAG
    if (code == 1 && a == 'T' && b == 'G' && c == 'A') return true;                  // TGA
    if (code == 2 && a == 'A' && b == 'G' && (c == 'A' || c == 'G')) return true;    // AGA AGG
    return false;
}

void walk_one(const ga_sw_res& r, const uint8_t* cig, const uint8_t* t, int32_t tlen, int frame, int code,
              ga_qc_row& row)
{
    row = ga_qc_row{0.0, 0, 0, 0, 0};
    if (!r.resolved || r.score <= 0) return;
    row.cov_text = double(r.end_i - r.start_i + 1) / double(tlen);
    int64_t i = r.start_i, j = r.start_j;
    uint8_t codon[3] = {0, 0, 0};
    int fill = -1;                      // bait codon being built (-1: none)
    int64_t i0 = -1;                    // text index of its first base
    int64_t last_codon_j = -1, stop_at = -1;
    int stops = 0;
    for (int32_t p = 0; p < r.cigar_len; ++p) {
        const uint8_t op = cig[p];
        if (op == 'M' || op == 'X') {
            const int pos = int(((j - frame) % 3 + 3) % 3);
            if (pos == 0) {
                fill = 0;
                i0 = i;
            }
            if (fill == pos) codon[fill++] = i < tlen ? t[i] : uint8_t(0);
            else fill = -1;
            if (fill == 3) {
                if (i == i0 + 2) {      // three contiguous text bases
                    last_codon_j = j - 2;
                    if (is_stop(codon[0], codon[1], codon[2], code)) {
                        ++stops;
                        stop_at = j - 2;
                    }
                }
                fill = -1;
            }
            ++i;
            ++j;
        } else if (op == 'I') {
            ++i;
        } else if (op == 'D') {
            fill = -1;
            ++j;
        }
    }
    int runs = 0;
    int64_t ins = 0, del = 0;
    for (int32_t p = 0; p < r.cigar_len;) {
        const uint8_t o = cig[p];
        if (o != 'I' && o != 'D') {
            ++p;
            continue;
        }
        int32_t q = p;
        while (q < r.cigar_len && cig[q] == o) ++q;
        if ((q - p) % 3 != 0) ++runs;
        (o == 'I' ? ins : del) += q - p;
        p = q;
    }
    int net = int(((ins - del) % 3 + 3) % 3);
    if (net > 1) net -= 3;
    row.fs_runs = runs;
    row.fs_net = net;
    if (stops > 0 && stop_at == last_codon_j) {
        row.stop_end = 1;
        --stops;
    }
    row.stops = stops;
}

}  // namespace

extern "C" int ga_gene_qc_walk(const ga_sw_res* res, int64_t n, const uint8_t* cigar_arena, const uint8_t* text,
                               const int64_t* text_off, const int32_t* text_len, int32_t frame, int32_t code,
                               ga_qc_row* rows)
{
    if (n < 0 || (n > 0 && (!res || !text || !text_off || !text_len || !rows)) || frame < 0 || frame > 2)
        return GA_INVALID_ARGUMENT;
    for (int64_t k = 0; k < n; ++k) {
        const ga_sw_res& r = res[k];
        if (r.resolved && r.score > 0) {
            if (!cigar_arena && r.cigar_len > 0) return GA_INVALID_ARGUMENT;
            if (r.start_i < 0 || r.end_i >= text_len[k] || r.start_i > r.end_i || r.start_j < 0)
                return GA_INVALID_ARGUMENT;
        }
    }
    ga::host_parallel_for(n, 256, [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k)
            walk_one(res[k], cigar_arena ? cigar_arena + res[k].cigar_off : nullptr, text + text_off[k],
                     text_len[k], frame, code, rows[k]);
    });
    return GA_OK;
}

namespace {

// One TSV row (tools/gene_qc.cpp:239-262): the flag from the score, the
// coverage and the walk's counts, then the line as the tool's fprintf writes
// it (the id up to a NUL byte, as %s stops there).  Returns the flag's class
// bits: 1 OK, 2 NO_ALIGN, 4 LOW_COV, 8 FS, 16 STOP.
int qc_line(std::string& s, const ga_sw_res& r, const ga_qc_row& q, const char* id, int64_t id_len,
            int32_t tlen, double min_cov, int32_t min_score)
{
    char flag[64];
    int cls = 0;
    if (!r.resolved || r.score <= min_score) {
        std::snprintf(flag, sizeof flag, "NO_ALIGN");
        cls = 2;
    } else {
        int p = 0;
        if (q.cov_text < min_cov) { p += std::snprintf(flag + p, sizeof flag - size_t(p), "LOW_COV"); cls |= 4; }
        if (q.fs_net != 0) {
            p += std::snprintf(flag + p, sizeof flag - size_t(p), "%sFS=%d", p ? "+" : "", q.fs_net);
            cls |= 8;
        }
        if (q.stops > 0) {
            p += std::snprintf(flag + p, sizeof flag - size_t(p), "%sSTOP=%d", p ? "+" : "", q.stops);
            cls |= 16;
        }
        if (p == 0) { std::snprintf(flag, sizeof flag, "OK"); cls = 1; }
    }
    const void* nul = std::memchr(id, 0, size_t(id_len));
    if (nul) id_len = static_cast<const char*>(nul) - id;
    s.append(id, size_t(id_len));
    char buf[256];
    const int w = std::snprintf(buf, sizeof buf, "\t%d\t%d\t%d\t%d\t%d\t%d\t%.3f\t%d\t%d\t%d\t%d\t%s\n", tlen,
                                r.score, r.start_i, r.start_j, r.end_i, r.end_j, q.cov_text, q.fs_runs, q.fs_net,
                                q.stops, q.stop_end, flag);
    s.append(buf, size_t(w));
    return cls;
}

}  // namespace

extern "C" int64_t ga_gene_qc_tsv(const ga_sw_res* res, const ga_qc_row* rows, int64_t n, const char* ids,
                                  const int64_t* id_off, const int32_t* text_len, double min_cov,
                                  int32_t min_score, char* out, int64_t out_cap, int64_t* counts)
{
    if (n < 0 || (n > 0 && (!res || !rows || !ids || !id_off || !text_len)) || !counts) return -1;
    static const char kHeader[] = "acc\tlen\tscore\tsi\tsj\tei\tej\tcov_text\tfs_runs\tfs_net\tstops\tstop_end\tflag\n";
    constexpr int64_t kBlock = 512;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    std::vector<std::string> part(static_cast<size_t>(nb));
    std::vector<std::array<int64_t, 6>> cnt(static_cast<size_t>(nb));
    ga::host_parallel_for(nb, 1, [&](int64_t b0, int64_t b1) {
        for (int64_t b = b0; b < b1; ++b) {
            std::string& s = part[size_t(b)];
            std::array<int64_t, 6>& c = cnt[size_t(b)];
            c.fill(0);
            s.reserve(size_t(kBlock) * 96);
            for (int64_t k = b * kBlock; k < std::min(n, (b + 1) * kBlock); ++k) {
                const ga_sw_res& r = res[k];
                const int cls = qc_line(s, r, rows[k], ids + id_off[k], id_off[k + 1] - id_off[k], text_len[k],
                                        min_cov, min_score);
                c[0] += cls == 1;
                c[1] += (cls & 2) != 0;
                c[2] += (cls & 4) != 0;
                c[3] += (cls & 8) != 0;
                c[4] += (cls & 16) != 0;
                c[5] += r.resolved && r.score > 0 && (!r.rescore_ok || !r.wellformed_ok);
            }
        }
    });
    int64_t total = int64_t(sizeof kHeader) - 1;
    for (int f = 0; f < 6; ++f) counts[f] = 0;
    for (int64_t b = 0; b < nb; ++b) {
        total += int64_t(part[size_t(b)].size());
        for (int f = 0; f < 6; ++f) counts[f] += cnt[size_t(b)][size_t(f)];
    }
    if (!out || out_cap < total) return total;
    std::memcpy(out, kHeader, sizeof kHeader - 1);
    int64_t pos = int64_t(sizeof kHeader) - 1;
    for (int64_t b = 0; b < nb; ++b) {
        std::memcpy(out + pos, part[size_t(b)].data(), part[size_t(b)].size());
        pos += int64_t(part[size_t(b)].size());
    }
    return total;
}
