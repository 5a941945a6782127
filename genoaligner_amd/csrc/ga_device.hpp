// ga_device.hpp -- launch-side declarations shared by the gfx950 kernels and the
// host layer.  Everything here is internal to libgenoaligner_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ga {

// Packed sequence buffers carry this many readable bytes before their first
// and after their last sequence: the extension and streaming loops read aligned
// 8-byte words up to 72 bytes outside the sequence they work on.
constexpr int kSeqPad = 128;

// Sentinel of an unreachable wavefront offset and its threshold test
// (reference: include/genoaligner/backend/wfa_kernel.hip:103,110).
constexpr int kWfaNeg = INT32_MIN / 2;

struct WfaLaunch {
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;
    const int32_t* smax;
    int32_t n;
    int32_t smax_max;
    int32_t* score;
    // trace only
    uint8_t* cigar; const int64_t* cig_off; int32_t* cig_len;
    int32_t* ws;            // wavefront stack, ws_stride ints per pair slot
    int64_t ws_stride;
    int32_t first_pair;     // pairs [first_pair, first_pair + n) of the arrays
};

// Enqueue the WFA kernel for pairs [L.first_pair, L.first_pair + L.n).
hipError_t wfa_launch(const WfaLaunch& L, bool trace, hipStream_t st);

struct SwLaunch {
    const uint8_t* txt; const int64_t* toff; const int32_t* tlen;   // rows
    const uint8_t* pat; const int64_t* poff; const int32_t* plen;   // columns
    int32_t n;
    int32_t max_m;          // max text_len in the launch
    int32_t match, mismatch, go, ge;
    // out: 8 ints per pair: score, end_i, end_j (row-major first), lane_end_i,
    // lane_end_j (wave64 lane rule), start_i, start_j, cigar_len
    int32_t* res;
    uint32_t* dirs;         // direction nibbles (trace)
    const int64_t* dir_off; // per pair word offset into dirs, or null: then pair
    int64_t dir_stride;     //   first_pair + q starts at word q * dir_stride
    const uint8_t* walk;    // per pair: 1 = emit span + CIGAR (trace); null = all
    uint8_t* cigar; const int64_t* cig_off;
    int32_t* bnd;           // global strip-boundary scratch (max_m*2 ints per pair) or null
    int32_t first_pair;
};

// Words of direction nibbles one pair needs for the trace path.
inline int64_t sw_dir_words(int m, int n)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t strips = (n + 63) / 64;
    const int64_t wps = (m + 63 + 7) / 8;
    return strips * wps * 64;
}

// Optional callbacks around a launcher's dominant kernel (timing).
struct LaunchHooks {
    virtual void before(hipStream_t) {}
    virtual void after(hipStream_t) {}
    virtual ~LaunchHooks() = default;
};

hipError_t sw_launch(const SwLaunch& L, bool trace, hipStream_t st, LaunchHooks* hooks = nullptr);

// Reduce the 8-int per-pair records to the device-resident result layout:
// out5 = score, start_i, start_j, end_i, end_j (trace: row-major end and span;
// score-only: the wave64 lane-rule end), cig_len = CIGAR length (may be null).
hipError_t sw_gather(const int32_t* res8, int64_t n, bool trace, int32_t* out5,
                     int32_t* cig_len, hipStream_t st);

}  // namespace ga
