// ref_driver.cpp -- C entry points over the reference's own kernel bodies,
// compiled for the CPU with the reference's own HIP CPU shim
// (tests/parity/hip_cpu_shim/hip/hip_runtime.h in the reference tree).
//
// TEST INFRASTRUCTURE ONLY.  oracle/build_ref.sh compiles this file, which
// includes the reference sources where they lie (src/api/api.cpp and the kernel
// headers it includes) into oracle/_ref/libga_ref.so.  Nothing here is copied
// from the reference: this file only sets the shim's launch globals, sizes the
// buffers the way the reference's launch sites do, calls the kernels, and
// decodes the reversed op arrays into strings.  It is used to write golden
// vectors (tests/golden/make_golden.py) and to cross-check the C restatement
// in oracle/ga_oracle.c.
//
// Launch shapes follow the reference's own CPU harnesses:
//   wfa_trace_kernel      blockDim 1, one call per pair (tests/parity/wfa_parity.cpp:290-300)
//   wfa_score_kernel_flat blockDim 1, smem 2*(2*smax+3) ints (bench/h7_flat_parity.cpp:68-78)
//   sw_trace_kernel       blockDim 1, one call per pair (src/api/api.cpp:668-674)
//   sw_score_kernel       shim::run_block(4*warpSize) per block (src/api/api.cpp:677-689)
//   align_sw_batch        the API itself, shim branch (src/api/api.cpp:649-700)

// The reference API translation unit is included (by path, from the reference
// tree) so that its kernels -- non-inline __global__ functions defined in
// headers -- are compiled exactly once, together with the entry points below.
#include "src/api/api.cpp"
#include "src/reference/edit_distance_cpu.hpp"
#include "include/genoaligner/io/fasta.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace genoaligner;

static char op_char(int op)
{
    return op == 0 ? 'M' : op == 1 ? 'X' : op == 2 ? 'I' : op == 3 ? 'D' : '?';
}

static int decode_rev(const int* rev, int used, char* out)
{
    for (int j = 0; j < used; ++j) out[j] = op_char(rev[used - 1 - j]);
    return used;
}

extern "C" {

// wfa_trace_kernel for one pair.  cigar needs capacity m+n+2.
// Returns 0; *score = -1 when unresolved.  *emitted/*trunc mirror cigar_meta.
int ga_ref_wfa_trace(const char* pattern, int m, const char* text, int n, int smax,
                     int* score, char* cigar, int* cigar_len, int* emitted, int* trunc)
{
    const int wf_stride = 2 * smax + 3;
    const int wf_alloc_max = (smax + 1) * wf_stride;
    const int cigar_cap = m + n + 2;
    std::vector<int> ws((size_t)wf_alloc_max, 0);
    std::vector<int> cg((size_t)cigar_cap, 0);
    std::vector<int> meta(4, 0);
    PairView pv{text, n, pattern, m, smax};
    int sc = -1;
    blockDim = dim3{1, 1, 1};
    threadIdx = uint3{0, 0, 0};
    blockIdx = uint3{0, 0, 0};
    wfa_trace_kernel(&pv, &sc, cg.data(), meta.data(), ws.data(), smax, wf_stride,
                     wf_alloc_max, wf_alloc_max * (int)sizeof(int), cigar_cap);
    *score = sc;
    *emitted = meta[3];
    *trunc = meta[2];
    *cigar_len = 0;
    if (meta[1] > 0 && cigar) *cigar_len = decode_rev(cg.data(), meta[1], cigar);
    return 0;
}

int ga_ref_wfa_score_flat(const char* pattern, int m, const char* text, int n, int smax)
{
    int sc = -12345;
    shim::smem_vec().assign((size_t)2 * (2 * smax + 3), WFA_NEG);
    PairView pv{text, n, pattern, m, smax};
    blockDim = dim3{1, 1, 1};
    threadIdx = uint3{0, 0, 0};
    blockIdx = uint3{0, 0, 0};
    wfa_score_kernel_flat(&pv, &sc, smax);
    return sc;
}

int ga_ref_edit_distance(const char* pattern, int m, const char* text, int n)
{
    return edit_distance_cpu(pattern, m, text, n);
}

// sw_trace_kernel for one pair.  out5 = {score, start_i, start_j, end_i, end_j}.
int ga_ref_sw_trace(const char* text, int m, const char* pattern, int n,
                    int match, int mismatch, int gap_open, int gap_extend,
                    int* out5, char* cigar, int* cigar_len)
{
    SWPairView pv{text, m, pattern, n};
    SWParams prm{match, mismatch, gap_open, gap_extend};
    const size_t dir_stride = ((size_t)m + 1) * ((size_t)n + 1);
    const size_t scratch_stride = 2 * ((size_t)n + 1);
    const int cigar_cap = m + n + 4;
    std::vector<uint8_t> dirs(dir_stride ? dir_stride : 1, 0);
    std::vector<int> scratch(scratch_stride, 0), cg((size_t)cigar_cap, 0), meta(4, 0);
    SWTraceResult tr{-1, -1, -1, -1, -1};
    blockDim = dim3{1, 1, 1};
    threadIdx = uint3{0, 0, 0};
    blockIdx = uint3{0, 0, 0};
    sw_trace_kernel(&pv, prm, &tr, cg.data(), meta.data(), dirs.data(), dir_stride,
                    scratch.data(), scratch_stride, 1, cigar_cap);
    out5[0] = tr.score; out5[1] = tr.start_i; out5[2] = tr.start_j;
    out5[3] = tr.end_i; out5[4] = tr.end_j;
    *cigar_len = 0;
    if (meta[1] > 0 && !meta[2] && cigar) *cigar_len = decode_rev(cg.data(), meta[1], cigar);
    return 0;
}

// sw_score_kernel under warp emulation at width `wsize` for n_pairs pairs laid
// out as packed strings with offsets.  out3 = {score, end_i, end_j} per pair.
int ga_ref_sw_score(const char* text, const long* text_off, const int* text_len,
                    const char* pattern, const long* pat_off, const int* pat_len,
                    int n_pairs, int match, int mismatch, int gap_open, int gap_extend,
                    int wsize, int* out3)
{
    std::vector<SWPairView> pv((size_t)n_pairs);
    int max_n = 0;
    for (int i = 0; i < n_pairs; ++i) {
        pv[(size_t)i] = SWPairView{text + text_off[i], text_len[i], pattern + pat_off[i], pat_len[i]};
        if (pat_len[i] > max_n) max_n = pat_len[i];
    }
    SWParams prm{match, mismatch, gap_open, gap_extend};
    std::vector<SWResult> res((size_t)n_pairs, SWResult{-1, -1, -1});
    warpSize = wsize;
    const int nthreads = SW_WARPS_PER_BLOCK * wsize;
    const int ints_per_w = sw_smem_ints_per_warp(max_n);
    shim::smem_vec().assign((size_t)SW_WARPS_PER_BLOCK * (size_t)ints_per_w, 0);
    blockDim = dim3{(unsigned)nthreads, 1, 1};
    const int nblocks = (n_pairs + SW_WARPS_PER_BLOCK - 1) / SW_WARPS_PER_BLOCK;
    for (int b = 0; b < nblocks; ++b) {
        blockIdx = uint3{(unsigned)b, 0, 0};
        shim::run_block(nthreads, [&] {
            sw_score_kernel(pv.data(), prm, res.data(), n_pairs, ints_per_w);
        });
    }
    blockDim = dim3{1, 1, 1};
    blockIdx = uint3{0, 0, 0};
    warpSize = 32;
    for (int i = 0; i < n_pairs; ++i) {
        out3[3 * i] = res[(size_t)i].score;
        out3[3 * i + 1] = res[(size_t)i].end_i;
        out3[3 * i + 2] = res[(size_t)i].end_j;
    }
    return 0;
}

// The reference API's SW batch (shim branch).  Per pair: out9 = {score, resolved,
// start_i, start_j, end_i, end_j, rescore_ok, wellformed_ok, too_large}; CIGARs
// are written at cig_off[i] (capacity text_len+pattern_len) with lengths in
// cig_len.  Returns the status (0 ok, 1 invalid_argument, 2 device_error).
int ga_ref_align_sw_batch(const char* text, const long* text_off, const int* text_len,
                          const char* pattern, const long* pat_off, const int* pat_len,
                          const unsigned char* null_text, const unsigned char* null_pat,
                          const int* scoring4, const unsigned char* with_cigar, int n_pairs,
                          int* out9, char* cigar, const long* cig_off, int* cig_len)
{
    std::vector<SWRequest> reqs((size_t)n_pairs);
    for (int i = 0; i < n_pairs; ++i) {
        SWRequest& r = reqs[(size_t)i];
        r.text = null_text && null_text[i] ? nullptr : text + text_off[i];
        r.text_len = text_len[i];
        r.pattern = null_pat && null_pat[i] ? nullptr : pattern + pat_off[i];
        r.pattern_len = pat_len[i];
        r.scoring = SWScoring{scoring4[4 * i], scoring4[4 * i + 1], scoring4[4 * i + 2],
                              scoring4[4 * i + 3]};
        r.with_cigar = with_cigar[i] != 0;
    }
    warpSize = 64;
    SWBatchResult b = align_sw_batch(reqs);
    warpSize = 32;
    for (int i = 0; i < n_pairs && i < (int)b.results.size(); ++i) {
        const SWAlignResult& r = b.results[(size_t)i];
        int* o = out9 + 9 * i;
        o[0] = r.score; o[1] = r.resolved; o[2] = r.start_i; o[3] = r.start_j;
        o[4] = r.end_i; o[5] = r.end_j; o[6] = r.rescore_ok; o[7] = r.wellformed_ok;
        o[8] = r.too_large;
        cig_len[i] = (int)r.cigar.size();
        if (cigar && !r.cigar.empty()) std::memcpy(cigar + cig_off[i], r.cigar.data(), r.cigar.size());
    }
    return b.status == SWBatchResult::Status::ok ? 0
         : b.status == SWBatchResult::Status::invalid_argument ? 1 : 2;
}

// The reference API's WFA batch validation (its shim branch reports every pair
// unresolved, so only the status is meaningful).  Returns 0 ok, 1 invalid_argument.
int ga_ref_align_batch_status(const int* pat_len, const int* text_len,
                              const unsigned char* null_pat, const unsigned char* null_text,
                              const int* smax, int n_pairs)
{
    static const char dummy[1] = {0};
    std::vector<std::string> store;
    std::vector<AlignRequest> reqs((size_t)n_pairs);
    for (int i = 0; i < n_pairs; ++i) {
        AlignRequest& r = reqs[(size_t)i];
        store.emplace_back((size_t)(pat_len[i] > 0 ? pat_len[i] : 0) +
                           (size_t)(text_len[i] > 0 ? text_len[i] : 0) + 1, 'A');
        r.pattern = null_pat[i] ? nullptr : store.back().data();
        r.pattern_len = pat_len[i];
        r.text = null_text[i] ? nullptr : store.back().data();
        r.text_len = text_len[i];
        r.smax = smax[i];
    }
    (void)dummy;
    BatchResult b = align_batch(reqs);
    return b.status == BatchResult::Status::ok ? 0
         : b.status == BatchResult::Status::invalid_argument ? 1 : 2;
}

// io::read_fasta over an in-memory file (written to a tmpfile, as a caller's
// file would be read).  Returns the record count, or -1 with the reader's
// message in err.  The records are kept and copied out by ga_ref_fasta_blob as
// [int64 id_len, desc_len, seq_len, id bytes, desc bytes, seq bytes] per record.
static std::string g_fasta_blob;

long long ga_ref_fasta_parse(const char* data, long long len, long long* blob_len, char* err, int err_cap)
{
    g_fasta_blob.clear();
    *blob_len = 0;
    FILE* f = std::tmpfile();
    if (!f) return -2;
    if (len > 0 && std::fwrite(data, 1, size_t(len), f) != size_t(len)) { std::fclose(f); return -2; }
    std::rewind(f);
    std::vector<io::FastaRecord> recs;
    std::string e;
    const bool ok = io::read_fasta(f, &recs, &e);
    std::fclose(f);
    if (!ok) {
        std::snprintf(err, size_t(err_cap), "%s", e.c_str());
        return -1;
    }
    for (const io::FastaRecord& r : recs) {
        const long long l[3] = {(long long)r.id.size(), (long long)r.description.size(),
                                (long long)r.sequence.size()};
        g_fasta_blob.append(reinterpret_cast<const char*>(l), sizeof(l));
        g_fasta_blob += r.id;
        g_fasta_blob += r.description;
        g_fasta_blob += r.sequence;
    }
    *blob_len = (long long)g_fasta_blob.size();
    return (long long)recs.size();
}

void ga_ref_fasta_blob(char* out) { std::memcpy(out, g_fasta_blob.data(), g_fasta_blob.size()); }

// io::read_fasta_file: returns the record count or -1 with the message in err.
long long ga_ref_fasta_file(const char* path, char* err, int err_cap)
{
    std::vector<io::FastaRecord> recs;
    std::string e;
    if (!io::read_fasta_file(path, &recs, &e)) {
        std::snprintf(err, size_t(err_cap), "%s", e.c_str());
        return -1;
    }
    return (long long)recs.size();
}

}  // extern "C"
