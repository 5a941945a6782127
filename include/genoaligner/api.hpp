// genoaligner -- pairwise alignment API, MI355X (gfx950) build.
//
// The C++ surface of the reference's pairwise path (reference:
// include/genoaligner/api.hpp:61-313 and :485-487), kept field-for-field and
// default-for-default so existing callers recompile unchanged.  It is a thin
// wrapper over the C ABI in genoaligner_c.h, which does all the work on the GPU.
//
// Contract highlights (the reference's, unchanged):
//   * align_batch: one smax per batch, the MINIMUM over the requests, each in
//     [0, 511]; invalid requests refuse the whole batch (invalid_argument);
//     unresolved (distance > smax) is score -1, not an error; if any request
//     asks for a CIGAR every resolved pair gets one, checked by re-scoring and
//     well-formedness.
//   * align_sw_batch: one scoring scheme per batch; gap_extend > gap_open is
//     refused with_cigar; score 0 resolves with coordinates -1; a pair whose
//     matrix exceeds SW_MAX_TRACE_CELLS reports too_large.
//   * Pointers are stored, not copied: inputs must outlive the call.
//
// Deliberate differences, all strictly more capable:
//   * no shared-memory size refusals (the reference refuses with_cigar at large
//     smax and score-only SW at long patterns on shared-memory grounds);
//   * align_batch_ext: per-request smax up to 4000 and per-request with_cigar;
//   * align_sw_matrix_batch: substitution-matrix (protein) Smith-Waterman.
#ifndef GENOALIGNER_API_HPP
#define GENOALIGNER_API_HPP

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace genoaligner {

struct AlignResult {
    int         score    = -1;
    bool        resolved = false;
    std::string cigar;
    bool        rescore_ok    = false;
    bool        wellformed_ok = false;
};

struct AlignRequest {
    const char* pattern     = nullptr;
    int         pattern_len = 0;
    const char* text        = nullptr;
    int         text_len    = 0;
    int         smax        = 64;
    bool        with_cigar  = true;
};

AlignResult align(const AlignRequest& req);

struct BatchResult {
    std::vector<AlignResult> results;
    int resolved_count   = 0;
    int unresolved_count = 0;
    enum class Status { ok, invalid_argument, device_error };
    Status status = Status::ok;
    const char* error = nullptr;
    bool ok() const { return status == Status::ok; }
};
BatchResult align_batch(const std::vector<AlignRequest>& reqs);

// Extension: per-request smax (<= 4000) and per-request with_cigar.
BatchResult align_batch_ext(const std::vector<AlignRequest>& reqs);

struct SWScoring {
    int match      = 2;
    int mismatch   = -3;
    int gap_open   = 5;
    int gap_extend = 2;
};

static constexpr int SW_MAX_TRACE_CELLS = 1 << 26;

struct SWRequest {
    const char* text        = nullptr;
    int         text_len    = 0;
    const char* pattern     = nullptr;
    int         pattern_len = 0;
    SWScoring   scoring;
    bool        with_cigar  = true;
};

struct SWAlignResult {
    int         score    = -1;
    bool        resolved = false;
    std::string cigar;
    int         start_i = -1;
    int         start_j = -1;
    int         end_i   = -1;
    int         end_j   = -1;
    bool        rescore_ok    = false;
    bool        wellformed_ok = false;
    bool        too_large = false;
};

struct SWBatchResult {
    std::vector<SWAlignResult> results;
    int resolved_count   = 0;
    int unresolved_count = 0;
    enum class Status { ok, invalid_argument, device_error };
    Status status = Status::ok;
    const char* error = nullptr;
    bool ok() const { return status == Status::ok; }
};
SWBatchResult align_sw_batch(const std::vector<SWRequest>& reqs);
SWAlignResult align_sw(const SWRequest& req);

// Extension: Smith-Waterman under a substitution matrix (protein).  Same rules
// as align_sw_batch; the requests' `scoring` is replaced by the scheme below:
// score[a][b] for text residue alphabet[a] against pattern residue alphabet[b],
// `unknown` for any pair with a byte outside the alphabet, and a gap of length
// L costing gap_open + (L - 1) * gap_extend.
struct SWMatrixScoring {
    std::string alphabet;                    // 1..31 distinct letters
    std::vector<std::vector<int>> score;     // alphabet.size() x alphabet.size(), int8 range
    int unknown    = -1;
    int gap_open   = 11;
    int gap_extend = 1;
};
SWMatrixScoring blosum62(int gap_open = 11, int gap_extend = 1);
SWBatchResult align_sw_matrix_batch(const std::vector<SWRequest>& reqs, const SWMatrixScoring& scheme);

const char* backend_name();
const char* device_name();
bool        device_available();

}  // namespace genoaligner

#endif  // GENOALIGNER_API_HPP
