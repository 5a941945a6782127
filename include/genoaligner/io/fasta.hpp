// genoaligner::io FASTA reading for C++ callers: the reference's
// include/genoaligner/io/fasta.hpp:38-55 surface (FastaRecord, read_fasta,
// read_fasta_file), header-only over the C ABI's reader (ga_fasta_parse), so
// the records and the error messages are the C ABI's, which follow the
// reference reader's rules and wording.
//
// One difference on failure: the reference may leave the records it read
// before the error appended to *out (its tests tell callers to discard them);
// these functions append nothing when they return false.
#ifndef GENOALIGNER_AMD_IO_FASTA_HPP
#define GENOALIGNER_AMD_IO_FASTA_HPP

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../genoaligner_c.h"

namespace genoaligner {
namespace io {

struct FastaRecord {
    std::string id;           // first whitespace-delimited token of the header
    std::string description;  // the rest of the header line (may be empty)
    std::string sequence;     // raw bytes, line breaks removed, case kept
};

// Every record of an open FILE*, appended to *out (fasta.hpp:50, :59-157).
inline bool read_fasta(FILE* in, std::vector<FastaRecord>* out, std::string* error)
{
    if (!in) { if (error) *error = "null FILE*"; return false; }
    if (!out) { if (error) *error = "null output vector"; return false; }
    std::string data;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), in)) > 0) data.append(buf, n);
    if (std::ferror(in)) { if (error) *error = "read error (line 0)"; return false; }
    ga_fasta* f = nullptr;
    const char* e = nullptr;
    if (ga_fasta_parse(data.data(), int64_t(data.size()), &f, &e) != GA_OK) {
        // the reference tests "no records" on the caller's vector after appending
        if (e && std::strncmp(e, "no records found", 16) == 0 && !out->empty()) return true;
        if (error) *error = e ? e : "";
        return false;
    }
    const int64_t count = ga_fasta_count(f);
    out->reserve(out->size() + size_t(count));
    for (int64_t k = 0; k < count; ++k) {
        const char *id, *desc, *seq;
        int64_t il, dl, sl;
        ga_fasta_record(f, k, &id, &il, &desc, &dl, &seq, &sl);
        out->push_back(FastaRecord{std::string(id, size_t(il)), std::string(desc, size_t(dl)),
                                   std::string(seq, size_t(sl))});
    }
    ga_fasta_free(f);
    return true;
}

// Open a path and read it (fasta.hpp:54-55, :159-170).
inline bool read_fasta_file(const std::string& path, std::vector<FastaRecord>* out, std::string* error)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        if (error) *error = "cannot open '" + path + "'";
        return false;
    }
    const bool ok = read_fasta(f, out, error);
    std::fclose(f);
    return ok;
}

}  // namespace io
}  // namespace genoaligner

#endif  // GENOALIGNER_AMD_IO_FASTA_HPP
