// Lattice descriptions for D2Q9 and D3Q27.
//
// The velocity ordering (alpha index) is part of the public interface: the
// distributions exchanged through the C-ABI are indexed as in the reference
// (detail/d2q9_lattice.cuh:44-70, detail/d3q27_lattice.cuh:45-65).  Weights
// follow from the velocity norm; the bounce-back partner is the opposite
// velocity.  Velocity components are packed two bits per alpha so that a
// lane-dependent alpha decodes with a shift instead of a memory lookup.
#pragma once

#include "naga_common.hpp"

namespace naga_amd {

template <int Q>
struct Lattice;

namespace detail {
struct Vel3 {
    int x, y, z;
};
constexpr Vel3 kVel27[27] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},   {0, 0, -1},  {1, 1, 0},   {-1, 1, 0},  {1, -1, 0},
    {-1, -1, 0}, {1, 0, 1},   {-1, 0, 1},  {1, 0, -1},  {-1, 0, -1},
    {0, 1, 1},   {0, -1, 1},  {0, 1, -1},  {0, -1, -1}, {1, 1, 1},
    {-1, 1, 1},  {1, -1, 1},  {-1, -1, 1}, {1, 1, -1},  {-1, 1, -1},
    {1, -1, -1}, {-1, -1, -1}};
constexpr Vel3 kVel9[9] = {{0, 0, 0},  {-1, 1, 0}, {-1, 0, 0},
                           {-1, -1, 0}, {0, -1, 0}, {1, -1, 0},
                           {1, 0, 0},  {1, 1, 0},  {0, 1, 0}};

template <int Q>
constexpr uint64_t pack(int d) {
    uint64_t p = 0;
    for (int a = 0; a < Q; ++a) {
        const Vel3& v = Q == 27 ? kVel27[a] : kVel9[a];
        int c = d == 0 ? v.x : d == 1 ? v.y : v.z;
        p |= uint64_t(c + 1) << (2 * a);
    }
    return p;
}
}  // namespace detail

template <int Q_>
struct Lattice {
    static_assert(Q_ == 9 || Q_ == 27, "D2Q9 and D3Q27 only");
    static constexpr int Q = Q_;
    static constexpr int D = Q_ == 9 ? 2 : 3;
    // AoS row stride (floats) of per-node distribution rows in HBM.  D3Q27
    // rows are padded to one 128-byte line so a neighbour gather touches a
    // single line.
    static constexpr int RS = Q_ == 27 ? 32 : 9;
    // lanes of a wavefront that serve one node in lane-over-alpha kernels
    static constexpr int LPN = Q_ == 27 ? 32 : 9;
    static constexpr int NPW = 64 / LPN;  // nodes per wavefront

    static constexpr uint64_t kPack[3] = {detail::pack<Q_>(0), detail::pack<Q_>(1),
                                          detail::pack<Q_>(2)};

    NAGA_HD static constexpr int c(int a, int d) {
        return int((kPack[d] >> (2 * a)) & 3u) - 1;
    }
    NAGA_HD static constexpr int norm2(int a) {
        return c(a, 0) * c(a, 0) + c(a, 1) * c(a, 1) + (D == 3 ? c(a, 2) * c(a, 2) : 0);
    }
    NAGA_HD static constexpr double w(int a) {
        return D == 3 ? (norm2(a) == 0 ? 8.0 / 27.0
                         : norm2(a) == 1 ? 2.0 / 27.0
                         : norm2(a) == 2 ? 1.0 / 54.0
                                         : 1.0 / 216.0)
                      : (norm2(a) == 0 ? 4.0 / 9.0
                         : norm2(a) == 1 ? 1.0 / 9.0
                                         : 1.0 / 36.0);
    }
    // opposite velocity
    NAGA_HD static constexpr int bb(int a) {
        for (int b = 0; b < Q_; ++b) {
            bool opp = true;
            for (int d = 0; d < D; ++d) opp = opp && c(b, d) == -c(a, d);
            if (opp) return b;
        }
        return -1;
    }
};

// lattice speed of sound, lattices.cuh:46-49
constexpr double kCsLat = 0.57735026918962576450914878050196;

}  // namespace naga_amd
