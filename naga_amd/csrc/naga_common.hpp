// Common definitions for the MI355X nonlocal-LBM library: status codes,
// error text, HIP call checking, host/device qualifiers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define NAGA_HD __host__ __device__
#define NAGA_DEV __device__
#else
#define NAGA_HD
#define NAGA_DEV
#endif

namespace naga_amd {

// Status codes of the C-ABI (include/naga_amd.h).
enum Status : int {
    kOk = 0,
    kInvalidArgument = 1,
    kOutOfMemory = 2,
    kDeviceError = 3,
    kNumericalError = 4,
    kNotSupported = 5,
};

// Exceptions carry a status; the C-ABI layer turns them into codes plus
// thread-local error text.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string& msg) {
    throw Error(status, msg);
}

inline void require(bool ok, const std::string& msg) {
    if (!ok) fail(kInvalidArgument, msg);
}

// Divergence records of a node with tile unions, packed: per group of 8
// records, six 16-byte units of weights (record r's x, y, z at floats 3r ..
// 3r + 2) and one unit of eight 16-bit union positions (record 2m in the
// low half of dword m, 2m + 1 in the high half); then one zero unit, which
// makes the row length odd in units so the records of consecutive nodes
// fall in distinct LDS banks.  k must be a multiple of 8.
constexpr int packed_div_units(int k) { return 7 * (k / 8) + 1; }
// Filter records of a node with tile unions, packed the same way: per group
// of 8 records two units of weights and one unit of 16-bit positions, then
// one zero unit.
constexpr int packed_filter_units(int k) { return 3 * (k / 8) + 1; }
// Record rows of a tile whose nodes use ng groups of 8 records (per-tile
// stencil widths): the groups' units plus one or two zero units, an odd
// number of 16-byte units per node (distinct LDS banks for the nodes of a
// wavefront); equal to the fixed-width rows for ng = k / 8 with k in
// {16, 32, 48}.
constexpr int ragged_div_units(int ng) { return 7 * ng + ((ng & 1) ? 2 : 1); }
constexpr int ragged_filter_units(int ng) { return 3 * ng + ((ng & 1) ? 2 : 1); }

#if defined(__HIPCC__)
inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    int status = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                     ? kOutOfMemory
                     : kDeviceError;
    fail(status, std::string(what) + ": " + hipGetErrorString(e));
}
#define NAGA_HIP(call) ::naga_amd::hip_check((call), #call)
#endif

}  // namespace naga_amd
