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

#include "ga_device.hpp"

#include <sys/mman.h>
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// Records back to back in three buffers; record k's id is
// ids[id_off[k] .. id_off[k+1]), likewise its description and sequence.
struct ga_fasta {
    std::string ids, descs, seqs;
    std::vector<int64_t> id_off{0}, desc_off{0}, seq_off{0};
    int64_t count() const { return int64_t(id_off.size()) - 1; }
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

// The reader's rules over data[0 .. total), records appended to F.  A
// segment of a parallel parse may hold no record (need_record false).
bool parse_fasta(const char* data, int64_t total, ga_fasta& F, std::string& err, bool need_record = true)
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
    const int64_t first = F.count();
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
            if (F.count() > first) F.seq_off.push_back(int64_t(F.seqs.size()));   // close the previous record
            F.ids.append(data + h0, size_t(sp - h0));
            F.id_off.push_back(int64_t(F.ids.size()));
            int64_t d0 = sp;
            while (d0 < h1 && (data[d0] == ' ' || data[d0] == '\t')) ++d0;
            F.descs.append(data + d0, size_t(h1 - d0));
            F.desc_off.push_back(int64_t(F.descs.size()));
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
            F.seqs.append(data + s0, size_t(i - s0));
            have_seq = true;
        }
    }
    if (in_record && !have_seq) return fail("record has a header but no sequence", record_line);
    if (F.count() > first) F.seq_off.push_back(int64_t(F.seqs.size()));
    if (need_record && F.count() == 0) return fail("no records found", line);
    return true;
}

// Files this large are parsed in segments on the host's worker threads.
constexpr int64_t kParallelBytes = int64_t(4) << 20;

// A record starts at every '>' that begins a line (byte 0, or after \n or
// \r: a lone \r ends a line too), and nowhere else: before the first record
// only lines starting with a blank or a line break are skipped, and such a
// line is skipped whole.  So segments cut at those bytes parse alone under
// the same rules, each into its own buffers, and are joined in order.  Any
// error sends the caller back to the sequential parse, whose message and line
// number are the reference's.
bool parse_fasta_parallel(const char* data, int64_t total, ga_fasta& F)
{
    const int64_t nseg = std::min<int64_t>(64, std::max<int64_t>(2, total / (int64_t(1) << 20)));
    std::vector<int64_t> cut(size_t(nseg) + 1, total);
    cut[0] = 0;
    ga::host_parallel_for(nseg - 1, 1, [&](int64_t a, int64_t b) {
        for (int64_t s = a + 1; s < b + 1; ++s) {
            // first record start at or after the segment's nominal start
            int64_t p = total * s / nseg;
            while (p < total) {
                const void* q = std::memchr(data + p, '>', size_t(total - p));
                if (!q) { p = total; break; }
                p = static_cast<const char*>(q) - data;
                if (p == 0 || data[p - 1] == '\n' || data[p - 1] == '\r') break;
                ++p;
            }
            cut[size_t(s)] = p;
        }
    });
    for (int64_t s = 1; s <= nseg; ++s) cut[size_t(s)] = std::max(cut[size_t(s)], cut[size_t(s) - 1]);
    std::vector<ga_fasta> part(static_cast<size_t>(nseg));
    std::vector<char> ok(static_cast<size_t>(nseg), 1);
    ga::host_parallel_for(nseg, 1, [&](int64_t a, int64_t b) {
        std::string e;
        for (int64_t s = a; s < b; ++s) {
            const int64_t len = cut[size_t(s) + 1] - cut[size_t(s)];
            if (len == 0) continue;
            ok[size_t(s)] = parse_fasta(data + cut[size_t(s)], len, part[size_t(s)], e, false);
        }
    });
    for (int64_t s = 0; s < nseg; ++s)
        if (!ok[size_t(s)]) return false;
    // join: sizes, then the copies
    std::vector<int64_t> n0(size_t(nseg) + 1, 0), i0(n0), d0(n0), q0(n0);
    for (int64_t s = 0; s < nseg; ++s) {
        const ga_fasta& P = part[size_t(s)];
        n0[size_t(s) + 1] = n0[size_t(s)] + P.count();
        i0[size_t(s) + 1] = i0[size_t(s)] + int64_t(P.ids.size());
        d0[size_t(s) + 1] = d0[size_t(s)] + int64_t(P.descs.size());
        q0[size_t(s) + 1] = q0[size_t(s)] + int64_t(P.seqs.size());
    }
    const size_t n = size_t(n0[size_t(nseg)]);
    F.ids.resize(size_t(i0[size_t(nseg)]));
    F.descs.resize(size_t(d0[size_t(nseg)]));
    F.seqs.resize(size_t(q0[size_t(nseg)]));
    F.id_off.assign(n + 1, 0);
    F.desc_off.assign(n + 1, 0);
    F.seq_off.assign(n + 1, 0);
    ga::host_parallel_for(nseg, 1, [&](int64_t a, int64_t b) {
        for (int64_t s = a; s < b; ++s) {
            const ga_fasta& P = part[size_t(s)];
            std::memcpy(&F.ids[size_t(i0[size_t(s)])], P.ids.data(), P.ids.size());
            std::memcpy(&F.descs[size_t(d0[size_t(s)])], P.descs.data(), P.descs.size());
            std::memcpy(&F.seqs[size_t(q0[size_t(s)])], P.seqs.data(), P.seqs.size());
            for (int64_t k = 1; k <= P.count(); ++k) {
                const size_t o = size_t(n0[size_t(s)] + k);
                F.id_off[o] = i0[size_t(s)] + P.id_off[size_t(k)];
                F.desc_off[o] = d0[size_t(s)] + P.desc_off[size_t(k)];
                F.seq_off[o] = q0[size_t(s)] + P.seq_off[size_t(k)];
            }
        }
    });
    return n > 0;                                 // none: the sequential parse words it
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
    if (len >= kParallelBytes && parse_fasta_parallel(data, len, *F)) {
        *out = F;
        return GA_OK;
    }
    *F = ga_fasta();
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
    // a regular file is parsed from a read-only mapping (no copy); anything
    // else is read whole, in one read when its size is known
    struct stat st {};
    if (fstat(fileno(f), &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
        void* m = mmap(nullptr, size_t(st.st_size), PROT_READ, MAP_PRIVATE | MAP_POPULATE, fileno(f), 0);
        if (m != MAP_FAILED) {
            std::fclose(f);
            const int rc = ga_fasta_parse(static_cast<const char*>(m), int64_t(st.st_size), out, error);
            munmap(m, size_t(st.st_size));
            return rc;
        }
    }
    std::string data;
    if (S_ISREG(st.st_mode) && st.st_size > 0) data.resize(size_t(st.st_size));
    size_t have = 0;
    for (;;) {
        if (have == data.size()) data.resize(std::max<size_t>(data.size() * 2, 65536));
        const size_t k = std::fread(&data[have], 1, data.size() - have, f);
        have += k;
        if (k == 0) break;
    }
    data.resize(have);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
        t_fasta_error = "read error (line 0)";
        if (error) *error = t_fasta_error.c_str();
        return GA_INVALID_ARGUMENT;
    }
    return ga_fasta_parse(data.data(), int64_t(data.size()), out, error);
}

int64_t ga_fasta_count(const ga_fasta* f) { return f ? f->count() : 0; }

int ga_fasta_record(const ga_fasta* f, int64_t k, const char** id, int64_t* id_len, const char** desc,
                    int64_t* desc_len, const char** seq, int64_t* seq_len)
{
    if (!f || k < 0 || k >= f->count()) return GA_INVALID_ARGUMENT;
    const size_t q = size_t(k);
    if (id) *id = f->ids.data() + f->id_off[q];
    if (id_len) *id_len = f->id_off[q + 1] - f->id_off[q];
    if (desc) *desc = f->descs.data() + f->desc_off[q];
    if (desc_len) *desc_len = f->desc_off[q + 1] - f->desc_off[q];
    if (seq) *seq = f->seqs.data() + f->seq_off[q];
    if (seq_len) *seq_len = f->seq_off[q + 1] - f->seq_off[q];
    return GA_OK;
}

void ga_fasta_sizes(const ga_fasta* f, int64_t* id_bytes, int64_t* desc_bytes, int64_t* seq_bytes)
{
    if (id_bytes) *id_bytes = f ? int64_t(f->ids.size()) : 0;
    if (desc_bytes) *desc_bytes = f ? int64_t(f->descs.size()) : 0;
    if (seq_bytes) *seq_bytes = f ? int64_t(f->seqs.size()) : 0;
}

int ga_fasta_export(const ga_fasta* f, char* ids, int64_t* id_off, char* descs, int64_t* desc_off, char* seqs,
                    int64_t* seq_off)
{
    if (!f) return GA_INVALID_ARGUMENT;
    const size_t n1 = f->id_off.size();
    if (id_off) std::memcpy(id_off, f->id_off.data(), n1 * sizeof(int64_t));
    if (desc_off) std::memcpy(desc_off, f->desc_off.data(), n1 * sizeof(int64_t));
    if (seq_off) std::memcpy(seq_off, f->seq_off.data(), n1 * sizeof(int64_t));
    if (ids) std::memcpy(ids, f->ids.data(), f->ids.size());
    if (descs) std::memcpy(descs, f->descs.data(), f->descs.size());
    if (seqs) std::memcpy(seqs, f->seqs.data(), f->seqs.size());
    return GA_OK;
}

void ga_fasta_free(ga_fasta* f) { delete f; }

}  // extern "C"
