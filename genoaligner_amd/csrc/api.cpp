// api.cpp -- the reference's C++ surface (include/genoaligner/api.hpp) over
// the C ABI.  Marshalling only: every rule lives behind ga_align_batch /
// ga_align_sw_batch.
#include <algorithm>
#include <cstring>
#include "../../include/genoaligner/api.hpp"
#include "../../include/genoaligner/genoaligner_c.h"

namespace genoaligner {
namespace {

template <class S> S status_of(int rc)
{
    return rc == GA_OK ? S::ok : rc == GA_INVALID_ARGUMENT ? S::invalid_argument : S::device_error;
}

BatchResult wfa_call(const std::vector<AlignRequest>& reqs, bool ext)
{
    BatchResult out;
    out.results.resize(reqs.size());
    if (reqs.empty()) return out;
    std::vector<ga_align_req> cr(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) {
        const AlignRequest& r = reqs[i];
        cr[i] = ga_align_req{r.pattern, r.pattern_len, r.text, r.text_len, r.smax, r.with_cigar ? 1 : 0};
    }
    std::vector<ga_align_res> res(reqs.size());
    const int64_t need = ga_cigar_arena_bytes_wfa(cr.data(), int64_t(cr.size()));
    std::string arena(size_t(need > 0 ? need : 1), '\0');
    int32_t rc_ok = 0, rc_un = 0;
    const char* err = nullptr;
    const int rc = (ext ? ga_align_batch_ext : ga_align_batch)(cr.data(), int64_t(cr.size()), res.data(),
                                                               &arena[0], need, &rc_ok, &rc_un, &err);
    out.status = status_of<BatchResult::Status>(rc);
    out.error = err;
    if (rc != GA_OK) return out;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const ga_align_res& r = res[i];
        AlignResult& o = out.results[i];
        o.score = r.score;
        o.resolved = r.resolved != 0;
        if (r.cigar_len > 0) o.cigar.assign(arena.data() + r.cigar_off, size_t(r.cigar_len));
        o.rescore_ok = r.rescore_ok != 0;
        o.wellformed_ok = r.wellformed_ok != 0;
    }
    out.resolved_count = rc_ok;
    out.unresolved_count = rc_un;
    return out;
}

}  // namespace

BatchResult align_batch(const std::vector<AlignRequest>& reqs) { return wfa_call(reqs, false); }
BatchResult align_batch_ext(const std::vector<AlignRequest>& reqs) { return wfa_call(reqs, true); }

AlignResult align(const AlignRequest& req)
{
    BatchResult b = align_batch(std::vector<AlignRequest>{req});
    return b.results.empty() ? AlignResult{} : b.results[0];
}

namespace {
// One SW batch through the C ABI, with (matrix != null) or without a
// substitution matrix.
SWBatchResult sw_call(const std::vector<SWRequest>& reqs, const ga_sw_matrix* matrix)
{
    SWBatchResult out;
    out.results.resize(reqs.size());
    if (reqs.empty()) return out;
    std::vector<ga_sw_req> cr(reqs.size());
    for (size_t i = 0; i < reqs.size(); ++i) {
        const SWRequest& r = reqs[i];
        cr[i] = ga_sw_req{r.text, r.text_len, r.pattern, r.pattern_len,
                          ga_sw_scoring{r.scoring.match, r.scoring.mismatch, r.scoring.gap_open,
                                        r.scoring.gap_extend},
                          r.with_cigar ? 1 : 0};
    }
    std::vector<ga_sw_res> res(reqs.size());
    const int64_t need = ga_cigar_arena_bytes_sw(cr.data(), int64_t(cr.size()));
    std::string arena(size_t(need > 0 ? need : 1), '\0');
    int32_t rc_ok = 0, rc_un = 0;
    const char* err = nullptr;
    const int rc = matrix ? ga_align_sw_matrix_batch(cr.data(), int64_t(cr.size()), matrix, res.data(), &arena[0],
                                                     need, &rc_ok, &rc_un, &err)
                          : ga_align_sw_batch(cr.data(), int64_t(cr.size()), res.data(), &arena[0], need,
                                              &rc_ok, &rc_un, &err);
    out.status = status_of<SWBatchResult::Status>(rc);
    out.error = err;
    if (rc != GA_OK) return out;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const ga_sw_res& r = res[i];
        SWAlignResult& o = out.results[i];
        o.score = r.score;
        o.resolved = r.resolved != 0;
        if (r.cigar_len > 0) o.cigar.assign(arena.data() + r.cigar_off, size_t(r.cigar_len));
        o.start_i = r.start_i; o.start_j = r.start_j;
        o.end_i = r.end_i; o.end_j = r.end_j;
        o.rescore_ok = r.rescore_ok != 0;
        o.wellformed_ok = r.wellformed_ok != 0;
        o.too_large = r.too_large != 0;
    }
    out.resolved_count = rc_ok;
    out.unresolved_count = rc_un;
    return out;
}
}  // namespace

SWBatchResult align_sw_batch(const std::vector<SWRequest>& reqs) { return sw_call(reqs, nullptr); }

SWMatrixScoring blosum62(int gap_open, int gap_extend)
{
    ga_sw_matrix m;
    ga_sw_matrix_blosum62(&m, gap_open, gap_extend);
    SWMatrixScoring s;
    s.alphabet.assign(m.alphabet, strnlen(m.alphabet, sizeof(m.alphabet)));
    const size_t k = s.alphabet.size();
    s.score.assign(k, std::vector<int>(k, 0));
    for (size_t a = 0; a < k; ++a)
        for (size_t b = 0; b < k; ++b) s.score[a][b] = m.score[a][b];
    s.unknown = m.unknown;
    s.gap_open = m.gap_open;
    s.gap_extend = m.gap_extend;
    return s;
}

SWBatchResult align_sw_matrix_batch(const std::vector<SWRequest>& reqs, const SWMatrixScoring& scheme)
{
    if (scheme.alphabet.size() > 31) {              // the C struct holds at most 31 letters
        SWBatchResult out;
        out.results.resize(reqs.size());
        out.status = SWBatchResult::Status::invalid_argument;
        out.error = "SW matrix: alphabet must hold 1..31 letters";
        return out;
    }
    ga_sw_matrix m;
    std::memset(&m, 0, sizeof(m));
    const size_t k = std::min(scheme.alphabet.size(), sizeof(m.alphabet));
    std::memcpy(m.alphabet, scheme.alphabet.data(), k);
    for (size_t a = 0; a < k && a < scheme.score.size(); ++a)
        for (size_t b = 0; b < k && b < scheme.score[a].size(); ++b)
            m.score[a][b] = int8_t(scheme.score[a][b]);
    m.unknown = scheme.unknown;
    m.gap_open = scheme.gap_open;
    m.gap_extend = scheme.gap_extend;
    return sw_call(reqs, &m);
}

SWAlignResult align_sw(const SWRequest& req)
{
    SWBatchResult b = align_sw_batch(std::vector<SWRequest>{req});
    return b.results.empty() ? SWAlignResult{} : b.results[0];
}

const char* backend_name() { return ga_backend_name(); }
const char* device_name() { return ga_device_name(); }
bool device_available() { return ga_device_available() != 0; }

}  // namespace genoaligner
