// fasta.cpp -- FASTA reading on the host, with the reference reader's rules
// (include/genoaligner/io/fasta.hpp:50-175), behind the C ABI:
//   * records start at a line whose first byte is '>'; the header runs to the
//     end of that line, trailing spaces and tabs stripped; the id is its first
//     token (up to a space or tab), the description the rest after the
//     separating blanks;
//   * sequence lines are taken byte for byte (case, N, IUPAC codes, spaces and
//     a '>' inside a line all kept), line breaks removed; \n, \r\n and a lone
//     \r all end a line; lines holding only a line break are skipped;
//   * before the first header a line starting with a space, tab or line break
//     is skipped whole; any other first byte is an error;
//   * errors, with the reference's wording and line numbers: "expected '>' at
//     start of a record, found 'c'", "empty identifier in header", "record has
//     a header but no sequence" (line of that header), "no records found".
#include "../../include/genoaligner/genoaligner_c.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

struct ga_fasta {
    std::vector<std::string> id, desc, seq;
};

namespace {

thread_local std::string t_fasta_error;   // *error of the FASTA calls on this thread

// First \n or \r at or after i (total if none): memchr for each, the \r
// search bounded by the \n found.
int64_t line_end(const char* data, int64_t i, int64_t total)
{
    const void* nl = std::memchr(data + i, '\n', size_t(total - i));
    const int64_t e = nl ? static_cast<const char*>(nl) - data : total;
    const void* cr = std::memchr(data + i, '\r', size_t(e - i));
    return cr ? static_cast<const char*>(cr) - data : e;
}

bool parse_fasta(const char* data, int64_t total, ga_fasta& F, std::string& err)
{
    int64_t i = 0;
    long line = 1;
    auto fail = [&](const std::string& why, long at) {
        err = why + " (line " + std::to_string(at) + ")";
        return false;
    };
    auto consume_eol = [&] {
        i = line_end(data, i, total);
        if (i >= total) return;
        if (data[i] == '\r' && i + 1 < total && data[i + 1] == '\n') ++i;
        ++i;
        ++line;
    };
    bool in_record = false, have_seq = false;
    long record_line = 0;
    while (i < total) {
        const char c = data[i];
        if (!in_record) {
            if (c == '\n' || c == '\r' || c == ' ' || c == '\t') { consume_eol(); continue; }
            if (c != '>') return fail("expected '>' at start of a record, found '" + std::string(1, c) + "'", line);
            ++i;
            const int64_t h0 = i;
            i = line_end(data, i, total);
            int64_t h1 = i;
            while (h1 > h0 && (data[h1 - 1] == ' ' || data[h1 - 1] == '\t')) --h1;
            int64_t sp = h0;
            while (sp < h1 && data[sp] != ' ' && data[sp] != '\t') ++sp;
            if (sp == h0) return fail("empty identifier in header", line);
            F.id.emplace_back(data + h0, size_t(sp - h0));
            int64_t d0 = sp;
            while (d0 < h1 && (data[d0] == ' ' || data[d0] == '\t')) ++d0;
            F.desc.emplace_back(data + d0, size_t(h1 - d0));
            F.seq.emplace_back();
            in_record = true;
            have_seq = false;
            record_line = line;
            consume_eol();
            continue;
        }
        if (c == '>') {
            if (!have_seq) return fail("record has a header but no sequence", record_line);
            in_record = false;
            continue;
        }
        if (c == '\n' || c == '\r') { consume_eol(); continue; }
        const int64_t s0 = i;
        i = line_end(data, i, total);
        if (i > s0) {
            F.seq.back().append(data + s0, size_t(i - s0));
            have_seq = true;
        }
    }
    if (in_record && !have_seq) return fail("record has a header but no sequence", record_line);
    if (F.id.empty()) return fail("no records found", line);
    return true;
}

}  // namespace

// Also used by the device ingest to word its errors.
bool ga_fasta_host_message(const char* data, int64_t len, std::string& err)
{
    ga_fasta F;
    return parse_fasta(data, len, F, err);
}

extern "C" {

int ga_fasta_parse(const char* data, int64_t len, ga_fasta** out, const char** error)
{
    if (error) *error = nullptr;
    if (!out || (!data && len > 0) || len < 0) {
        if (error) *error = "null argument";
        return GA_INVALID_ARGUMENT;
    }
    *out = nullptr;
    ga_fasta* F = new ga_fasta();
    if (!parse_fasta(data, len, *F, t_fasta_error)) {
        delete F;
        if (error) *error = t_fasta_error.c_str();
        return GA_INVALID_ARGUMENT;
    }
    *out = F;
    return GA_OK;
}

int ga_fasta_read_file(const char* path, ga_fasta** out, const char** error)
{
    if (error) *error = nullptr;
    if (!path || !out) {
        if (error) *error = "null argument";
        return GA_INVALID_ARGUMENT;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        t_fasta_error = std::string("cannot open '") + path + "'";
        if (error) *error = t_fasta_error.c_str();
        return GA_INVALID_ARGUMENT;
    }
    std::string data;
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, k);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
        t_fasta_error = "read error (line 0)";
        if (error) *error = t_fasta_error.c_str();
        return GA_INVALID_ARGUMENT;
    }
    return ga_fasta_parse(data.data(), int64_t(data.size()), out, error);
}

int64_t ga_fasta_count(const ga_fasta* f) { return f ? int64_t(f->id.size()) : 0; }

int ga_fasta_record(const ga_fasta* f, int64_t k, const char** id, int64_t* id_len, const char** desc,
                    int64_t* desc_len, const char** seq, int64_t* seq_len)
{
    if (!f || k < 0 || k >= int64_t(f->id.size())) return GA_INVALID_ARGUMENT;
    const size_t q = size_t(k);
    if (id) *id = f->id[q].data();
    if (id_len) *id_len = int64_t(f->id[q].size());
    if (desc) *desc = f->desc[q].data();
    if (desc_len) *desc_len = int64_t(f->desc[q].size());
    if (seq) *seq = f->seq[q].data();
    if (seq_len) *seq_len = int64_t(f->seq[q].size());
    return GA_OK;
}

void ga_fasta_sizes(const ga_fasta* f, int64_t* id_bytes, int64_t* desc_bytes, int64_t* seq_bytes)
{
    int64_t a = 0, b = 0, c = 0;
    if (f)
        for (size_t k = 0; k < f->id.size(); ++k) {
            a += int64_t(f->id[k].size());
            b += int64_t(f->desc[k].size());
            c += int64_t(f->seq[k].size());
        }
    if (id_bytes) *id_bytes = a;
    if (desc_bytes) *desc_bytes = b;
    if (seq_bytes) *seq_bytes = c;
}

int ga_fasta_export(const ga_fasta* f, char* ids, int64_t* id_off, char* descs, int64_t* desc_off, char* seqs,
                    int64_t* seq_off)
{
    if (!f) return GA_INVALID_ARGUMENT;
    int64_t a = 0, b = 0, c = 0;
    const size_t n = f->id.size();
    for (size_t k = 0; k < n; ++k) {
        if (id_off) id_off[k] = a;
        if (desc_off) desc_off[k] = b;
        if (seq_off) seq_off[k] = c;
        if (ids) std::memcpy(ids + a, f->id[k].data(), f->id[k].size());
        if (descs) std::memcpy(descs + b, f->desc[k].data(), f->desc[k].size());
        if (seqs) std::memcpy(seqs + c, f->seq[k].data(), f->seq[k].size());
        a += int64_t(f->id[k].size());
        b += int64_t(f->desc[k].size());
        c += int64_t(f->seq[k].size());
    }
    if (id_off) id_off[n] = a;
    if (desc_off) desc_off[n] = b;
    if (seq_off) seq_off[n] = c;
    return GA_OK;
}

void ga_fasta_free(ga_fasta* f) { delete f; }

}  // extern "C"
