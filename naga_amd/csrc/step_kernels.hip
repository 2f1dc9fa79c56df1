// Stepping kernels (see step_kernels.hpp for the layout and schedule).
#include <cstdlib>

#include "step_kernels.hpp"

#ifndef NAGA_RECORD_POLICY
#define NAGA_RECORD_POLICY 2
#endif

namespace naga_amd {

namespace {

constexpr int kBlock = 256;  // 4 wavefronts
#ifndef NAGA_P1_BLOCK
#define NAGA_P1_BLOCK 256
#endif
// Collision block: one node per thread, rows transposed through LDS
// ((RS + 1) floats per node), so the block size sets the LDS per block.
constexpr int kP1Block = NAGA_P1_BLOCK;

// Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD works
// on one contiguous range of tiles (neighbouring nodes share its L2).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
    const int64_t q = nb / 8, r = nb % 8;
    const int64_t x = b % 8, l = b / 8;
    return x * q + (x < r ? x : r) + l;
}

// Block position within its scene: the launch's own block for a single
// scene, else the scene owning the block (last s with bstart[s] <= block)
// whose descriptor replaces G and S.
struct BlockCtx {
    int64_t b, nb;
    int scene;
};

__device__ __forceinline__ BlockCtx locate(const BatchArgs& B, StepGeom& G, StepBuffers& S) {
    if (!B.scenes) return {int64_t(blockIdx.x), int64_t(gridDim.x), 0};
    int lo = 0, hi = B.nscenes;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (B.bstart[mid] <= int32_t(blockIdx.x)) lo = mid;
        else hi = mid;
    }
    G = B.scenes[lo].G;
    S = B.scenes[lo].S;
    return {int64_t(blockIdx.x) - B.bstart[lo], int64_t(B.bstart[lo + 1] - B.bstart[lo]), lo};
}

template <int Q>
struct LaneMap {
    using L = Lattice<Q>;
    int64_t node;  // node slot (tile-relative index into the node range)
    int a;         // alpha served by this lane (clamped to Q-1 when idle)
    bool active;   // lane serves a real (node, alpha)
    __device__ LaneMap(int64_t tile) {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int sub = lane / L::LPN;
        const int al = lane - sub * L::LPN;
        node = (tile * (kBlock / 64) + wave) * L::NPW + (sub < L::NPW ? sub : 0);
        active = sub < L::NPW && al < Q;
        a = al < Q ? al : Q - 1;
    }
};

template <int Q>
__device__ __forceinline__ float lane_c(int a, int d) {
    return float(Lattice<Q>::c(a, d));
}

// ---------------------------------------------------------------------------
// p1: macroscopic values and the FP64 collision.  One thread per node; the
// block's rows are moved through LDS so global loads and stores are
// coalesced (rows padded by one float in LDS against bank conflicts).
// Writes s1 (into A): g + collision for fluid nodes, the collision
// increment alone for boundary nodes (p3 adds the interpolated state); the
// unitless (rho - 1, u) per node for the PML pass; the reference's
// macroscopic state (rho, u with boundary velocities zeroed).
// Reference: simulation_engine.cuh:366-433 (macroscopic), :469-531
// (collision), :876-877 (boundary velocity reset).
// ---------------------------------------------------------------------------
// A node's deviation row as contiguous floats.
struct FlatRow {
    float* p;
    __device__ float get(int a) const { return p[a]; }
    __device__ void set(int a, float v) const { p[a] = v; }
};

// Macroscopic values, density source and FP64 collision of node i, its
// deviation row r[0..Q) updated in place.
template <int Q, class Row>
__device__ __forceinline__ float4 p1_node(const StepGeom& G, const StepBuffers& S,
                                          const CollisionParams& cp, int64_t i, Row r) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS;
    double f[Q];
    double rt = 0.0, mom[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const double gv = double(r.get(a));
        rt += gv;
#pragma unroll
        for (int d = 0; d < D; ++d) mom[d] += gv * double(L::c(a, d));
        f[a] = L::w(a) + gv;
    }
    double u[3] = {0.0, 0.0, 0.0};
    const double inv_rho = 1.0 / (1.0 + rt);
#pragma unroll
    for (int d = 0; d < D; ++d) u[d] = mom[d] * inv_rho;
    // density source (simulation_engine.cuh:1057-1107, once per node):
    // f += w s / rho0, rho <- (rho + (rho + s)) / 2, u unchanged
    const double ws = S.src ? double(S.src[i]) / double(G.rho0) : 0.0;
    if (ws != 0.0) {
#pragma unroll
        for (int a = 0; a < Q; ++a) f[a] += L::w(a) * ws;
        rt += 0.5 * ws;
    }
    const double rho = 1.0 + rt;
    const bool bnd = i >= G.nfluid;
    S.rho[i] = float(rho * double(G.rho0));
#pragma unroll
    for (int d = 0; d < D; ++d) S.u[d * G.N + i] = bnd ? 0.0f : float(u[d] * double(G.vel_scale));
    const float4 ru = make_float4(float(rt), float(u[0]), float(u[1]), float(u[2]));
    double df[Q];
    collide<Q>(f, rho, u, cp, df);
    const bool src_node = !bnd && S.bsrc[i];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const double v = bnd ? df[a] : double(r.get(a)) + L::w(a) * ws + df[a];
        r.set(a, float(v));
        if (src_node) S.coll[i * RS + a] = float(df[a]);
    }
    return ru;
}

template <int Q>
__global__ __launch_bounds__(kP1Block) void k_p1(StepGeom G, StepBuffers S,
                                                 CollisionParams cp, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int RS = L::RS, LS = RS + 1;
    __shared__ float rows[kP1Block * LS];
    const BlockCtx bc = locate(B, G, S);
    if (B.scenes) cp = B.scenes[bc.scene].cp;
    const int64_t node0 = bc.b * kP1Block;
    const int64_t left = G.N - node0;
    const int cnt = left < kP1Block ? int(left) : kP1Block;
    const float* src = S.g + node0 * RS;
    for (int e = threadIdx.x; e < cnt * RS; e += kP1Block) rows[(e / RS) * LS + e % RS] = src[e];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < cnt) S.ru[node0 + t] = p1_node<Q>(G, S, cp, node0 + t, FlatRow{rows + t * LS});
    __syncthreads();
    float* dst = S.A + node0 * RS;
    for (int e = threadIdx.x; e < cnt * RS; e += kP1Block) dst[e] = rows[(e / RS) * LS + e % RS];
}

// ---------------------------------------------------------------------------
// PML local terms (pml_absorption.inl:551-631) over the nodes whose Q values
// are used (the layer, plus nodes [0, num_layer) when the layer field reads
// Q at j - num_bulk): Q1 = (feq - w + E) dt/2, Q2 = (Q1 + Q1_prev) dt/2,
// E <- feq - w, and on layer nodes s1 -= sigma (3 (feq - w) + 3 sigma Q1 +
// sigma^2 Q2).  Lane over alpha; each block also emits its partial column
// sums of Q1 and Q2 over layer nodes for the layer means (fixed partition,
// fixed order: deterministic).
// ---------------------------------------------------------------------------
constexpr int kPmlRounds = 8;

// feq_a - w_a (equilibrium_distribution.cuh:52-63, cancellation-free) at
// rho = 1 + rt and velocity u, ru = (rt, u).
template <int Q>
__device__ __forceinline__ double feq_dev(const float4& ru, int a) {
    using L = Lattice<Q>;
    constexpr double cs2 = kCsLat * kCsLat;
    constexpr double inv_cs2 = 1.0 / cs2, inv_2cs4 = 1.0 / (2.0 * cs2 * cs2);
    const double rt = double(ru.x), rho = 1.0 + rt;
    const double u0 = ru.y, u1 = ru.z, u2 = ru.w;
    const double uc = L::D == 3 ? u0 * double(L::c(a, 0)) + u1 * double(L::c(a, 1)) +
                                      u2 * double(L::c(a, 2))
                                : u0 * double(L::c(a, 0)) + u1 * double(L::c(a, 1));
    const double uu = L::D == 3 ? u2 * u2 + (u1 * u1 + u0 * u0) : u1 * u1 + u0 * u0;
    return L::w(a) * (rt + rho * (uc * inv_cs2 + (uc * uc) * inv_2cs4 - uu));
}

// The same in FP32 (the fused collision pass): the terms are deviations of
// order 1e-3 whose FP32 rounding is that of the FP32 rows they update.
template <int Q>
__device__ __forceinline__ float feq_dev32(const float4& ru, int a) {
    using L = Lattice<Q>;
    constexpr float cs2 = float(kCsLat * kCsLat);
    constexpr float inv_cs2 = float(1.0 / (kCsLat * kCsLat));
    constexpr float inv_2cs4 = float(1.0 / (2.0 * kCsLat * kCsLat * kCsLat * kCsLat));
    (void)cs2;
    const float rt = ru.x, rho = 1.0f + rt;
    const float uc = L::D == 3 ? ru.y * float(L::c(a, 0)) + ru.z * float(L::c(a, 1)) +
                                     ru.w * float(L::c(a, 2))
                               : ru.y * float(L::c(a, 0)) + ru.z * float(L::c(a, 1));
    const float uu = L::D == 3 ? ru.w * ru.w + (ru.z * ru.z + ru.y * ru.y)
                               : ru.z * ru.z + ru.y * ru.y;
    return float(L::w(a)) * (rt + rho * (uc * inv_cs2 + (uc * uc) * inv_2cs4 - uu));
}

template <int Q>
__global__ __launch_bounds__(kBlock) void k_pml_local(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int RS = L::RS;
    constexpr int NPB = kBlock / 64 * L::NPW;
    __shared__ double red[NPB][2][Q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / L::LPN, al = lane - sub * L::LPN;
    const bool active = sub < L::NPW && al < Q;
    const int slot = wave * L::NPW + (sub < L::NPW ? sub : 0);
    const int a = al < Q ? al : Q - 1;
    const BlockCtx bc = locate(B, G, S);
    const double hdt = double(G.dt) / 2.0;
    double s1 = 0.0, s2 = 0.0;
    const int64_t base = bc.b * (NPB * kPmlRounds);
    for (int rd = 0; rd < kPmlRounds; ++rd) {
        const int64_t item = base + rd * NPB + slot;
        if (!active || item >= S.n_rq) continue;
        const int64_t i = S.rq_list[item];
        const float4 ru = S.ru[i];
        float4* prev = reinterpret_cast<float4*>(S.E) + i;
        // feq - w now and at the previous step's macroscopic values
        const double fte = feq_dev<Q>(ru, a);
        const double ftp = feq_dev<Q>(*prev, a);
        const int64_t o = i * RS + a;
        const double q1 = (fte + ftp) * hdt;
        const double q2 = (q1 + double(S.Qp[o])) * hdt;
        // the node's lanes (one wavefront) have read prev: the store waits
        // for q1, which needs the loaded value
        float4 keep = ru;
        asm volatile("" : "+v"(keep.x) : "v"(q1));
        if (al == 0) *prev = keep;
        S.Qp[o] = float(q1);
        const bool layer = i >= G.nbulk && i < G.nfluid;
        if (G.pml_ref ? i < G.nlayer : layer) S.Q2v[o] = float(q2);  // rows the layer field reads
        if (layer) {
            const double sig = double(S.sigma[i - G.nbulk]);
            S.A[o] = float(double(S.A[o]) - sig * (3.0 * fte + 3.0 * sig * q1 + sig * sig * q2));
            s1 += q1;
            s2 += q2;
        }
    }
    if (active) {  // every (slot, alpha) pair has exactly one active lane
        red[slot][0][a] = s1;
        red[slot][1][a] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 2 * Q) {
        const int which = threadIdx.x / Q, aa = threadIdx.x % Q;
        double t = 0.0;
        for (int sl = 0; sl < NPB; ++sl) t += red[sl][which][aa];
        S.partial[int64_t(threadIdx.x) * S.nparts + bc.b] = t;
    }
}

// layer means (pml_absorption.inl:775-785, :816-826): one block per column
template <int Q>
__global__ __launch_bounds__(256) void k_means(StepGeom G, StepBuffers S, BatchArgs B) {
    __shared__ double red[256];
    const BlockCtx bc = locate(B, G, S);
    const int col = int(bc.b);  // 0..2Q-1
    double s = 0.0;
#pragma unroll 8
    for (int p = S.part_lo + threadIdx.x; p < S.nparts; p += 256)
        s += S.partial[int64_t(col) * S.nparts + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        S.means[(col / Q) * means_stride(Q) + col % Q] =
            G.nlayer > 0 ? float(red[0] / double(G.nlayer)) : 0.0f;
}

// PML auxiliary fields on layer rows (pml_absorption.inl:440-453, :500-513):
// Gf_a(j) = -2 sigma_j (Q1_a(p) - m1_a) + sigma_j^2 (Q2_a(p) - m2_a),
// p = j - num_bulk (reference indexing) or j.  Rows outside the layer stay
// zero, so the divergence pass needs no branch.
template <int Q>
__global__ __launch_bounds__(kBlock) void k_pml_field(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int RS = L::RS;
    const BlockCtx bc = locate(B, G, S);
    const LaneMap<Q> m(xcd_tile(bc.b, bc.nb));
    if (m.node >= G.nlayer || !m.active) return;
    const int64_t j = G.nbulk + m.node;
    const int64_t p = G.pml_ref ? j - G.nbulk : j;
    const int a = m.a;
    const float sig = S.sigma[m.node];
    const float q1 = S.Qp[p * RS + a], q2 = S.Q2v[p * RS + a];
    S.Gf[j * RS + a] =
        -2.0f * sig * (q1 - S.means[a]) + sig * sig * (q2 - S.means[means_stride(Q) + a]);
}

// ---------------------------------------------------------------------------
// Stencil records staged in LDS.  A block serves NPB node slots; the records
// of its nodes are copied once into LDS (coalesced), then every lane of a
// node reads them with broadcast LDS reads while its gathers go to global
// memory in unrolled batches, so many independent loads are in flight per
// wavefront.
// ---------------------------------------------------------------------------
template <int Q>
struct Tile {
    static constexpr int NPB = kBlock / 64 * Lattice<Q>::NPW;  // node slots
};

// Copy the records of nodes [node0, node0 + cnt) (contiguous slab rows of
// K*REC floats) into lds.
template <int REC>
__device__ __forceinline__ void stage_contiguous(const float* slab, int64_t node0, int cnt,
                                                 int K, float* lds) {
    const float* src = slab + node0 * int64_t(K) * REC;
    const int tot = cnt * K * REC;
    if (REC == 4) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
        float4* d4 = reinterpret_cast<float4*>(lds);
        for (int t = threadIdx.x; t < tot / 4; t += kBlock) d4[t] = s4[t];
    } else if (REC == 2) {
        const float2* s2 = reinterpret_cast<const float2*>(src);
        float2* d2 = reinterpret_cast<float2*>(lds);
        for (int t = threadIdx.x; t < tot / 2; t += kBlock) d2[t] = s2[t];
    } else {
        for (int t = threadIdx.x; t < tot; t += kBlock) lds[t] = src[t];
    }
}

// Copy the records of one node (lanes of the node share the work).
template <int Q, int REC>
__device__ __forceinline__ void stage_node(const float* slab, int64_t node, bool valid, int K,
                                           int a_lane, float* lds) {
    if (!valid) return;
    const float* src = slab + node * int64_t(K) * REC;
    for (int r = a_lane; r < K; r += Lattice<Q>::LPN)
#pragma unroll
        for (int c = 0; c < REC; ++c) lds[r * REC + c] = src[r * REC + c];
}

template <int Q>
struct Lanes {
    using L = Lattice<Q>;
    int slot, a, al;
    bool active, in_wave;
    __device__ Lanes() {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int sub = lane / L::LPN;
        al = lane - sub * L::LPN;
        in_wave = sub < L::NPW;
        active = in_wave && al < Q;
        a = al < Q ? al : Q - 1;
        slot = wave * L::NPW + (sub < L::NPW ? sub : 0);
    }
};

// sum_s (c_a . W_s) fin(idx_s)_a over the staged records rp (REC = 1 + D),
// batches of 8 neighbours.
template <int Q, int KT, int BATCH = 8>
__device__ __forceinline__ float stencil_sum(const float* rp, int K, const float* fin, int a) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS, REC = 1 + D, KMAX = KT ? KT : 64;
    const float cx = lane_c<Q>(a, 0), cy = lane_c<Q>(a, 1),
                cz = D == 3 ? lane_c<Q>(a, 2) : 0.0f;
    float acc = 0.0f;
#pragma unroll
    for (int s0 = 0; s0 < KMAX; s0 += BATCH) {
        if (!KT && s0 >= K) break;
        float v[BATCH], gv[BATCH];
        uint32_t j[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int s = s0 + u;
            const bool in = KT ? (s < KT) : (s < K);
            const float* r = rp + (in ? s : 0) * REC;
            j[u] = __float_as_uint(r[0]);
            float w = cx * r[1] + cy * r[2];
            if (D == 3) w += cz * r[3];
            v[u] = in ? w : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < BATCH; ++u) gv[u] = fin[int64_t(j[u]) * RS + a];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) acc += v[u] * gv[u];
    }
    return acc;
}

// ---------------------------------------------------------------------------
// p2: divergence of the PML auxiliary fields, added to s1
// (pml_absorption.inl:787-851), over the fluid nodes whose stencil reaches
// the layer; the field rows are zero outside the layer.
// ---------------------------------------------------------------------------
template <int Q, int KT>
__global__ __launch_bounds__(kBlock) void k_p2(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS, REC = 1 + D, KMAX = KT ? KT : 64;
    constexpr int NPB = Tile<Q>::NPB;
    __shared__ float srec[NPB * KMAX * REC];
    const BlockCtx bc = locate(B, G, S);
    const int K = KT ? KT : G.k;
    const Lanes<Q> ln;
    const int64_t item = xcd_tile(bc.b, bc.nb) * NPB + ln.slot;
    const bool valid = item < S.n_pml;
    const int64_t i = valid ? S.pml_list[item] : 0;
    float* rp = srec + ln.slot * K * REC;
    stage_node<Q, REC>(S.slab, i, valid && ln.in_wave, K, ln.al, rp);
    __syncthreads();
    if (!valid) return;
    const float acc = stencil_sum<Q, KT>(rp, K, S.Gf, ln.a);
    if (ln.active) S.A[i * RS + ln.a] += acc;
}

// ---------------------------------------------------------------------------
// boundary interpolation (simulation_engine.cuh:434-467 ->
// radial_point_method.cuh:235-250) in deviation form: g_b = sum_s w_s g_src.
// accumulate = true (p3): sources are s1 minus their collision increment
// (the reference interpolates before colliding) and the result is added to
// the row, which holds the boundary node's own collision increment;
// false (p5) overwrites.
// ---------------------------------------------------------------------------
template <int Q, bool ACCUMULATE>
__global__ __launch_bounds__(kBlock) void k_binterp(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int RS = L::RS;
    const BlockCtx bc = locate(B, G, S);
    float* field = ACCUMULATE ? S.A : S.g;
    const float* bslab = S.bslab;
    const float* coll = S.coll;
    const LaneMap<Q> m(xcd_tile(bc.b, bc.nb));
    const int64_t nb = G.N - G.nfluid;
    if (m.node >= nb) return;
    const int a = m.a;
    const float* rec = bslab + m.node * int64_t(G.kb) * 2;
    float acc = 0.0f;
    for (int s = 0; s < G.kb; ++s) {
        const int64_t j = int64_t(__float_as_uint(rec[2 * s]));
        const float v = ACCUMULATE ? field[j * RS + a] - coll[j * RS + a] : field[j * RS + a];
        acc += rec[2 * s + 1] * v;
    }
    const int64_t b = G.nfluid + m.node;
    if (m.active) field[b * RS + a] = ACCUMULATE ? field[b * RS + a] + acc : acc;
}

// ---------------------------------------------------------------------------
// p4: selective filter on fluid nodes (simulation_engine.cuh:1233-1323, with
// the weights normalised once at setup) and bounce-back on boundary nodes
// (:789-944).  Reads s1 (A), writes the new state g.
// ---------------------------------------------------------------------------
template <int Q, int KT>
__global__ __launch_bounds__(kBlock) void k_p4(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS, KMAX = KT ? KT : 64;
    constexpr int NPB = Tile<Q>::NPB;
    __shared__ float srec[NPB * KMAX * 2];
    const BlockCtx bc = locate(B, G, S);
    const int K = KT ? KT : G.k;
    const int64_t node0 = xcd_tile(bc.b, bc.nb) * NPB;
    const int64_t nf = G.nfluid - node0;
    stage_contiguous<2>(S.fslab, node0, nf <= 0 ? 0 : (nf < NPB ? int(nf) : NPB), K, srec);
    __syncthreads();
    const Lanes<Q> ln;
    const int64_t i = node0 + ln.slot;
    if (i >= G.N) return;
    const int a = ln.a;
    const float* s1 = S.A;
    float out;
    if (i < G.nfluid) {
        const float si = s1[i * RS + a];
        const float* rp = srec + ln.slot * K * 2;
        float sum = 0.0f;
#pragma unroll
        for (int s0 = 0; s0 < KMAX; s0 += 8) {
            if (!KT && s0 >= K) break;
            float w[8], gv[8];
            uint32_t j[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int s = s0 + u;
                const bool in = KT ? (s < KT) : (s < K);
                j[u] = __float_as_uint(rp[(in ? s : 0) * 2]);
                w[u] = in ? rp[s * 2 + 1] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) gv[u] = s1[int64_t(j[u]) * RS + a];
#pragma unroll
            for (int u = 0; u < 8; ++u) sum += w[u] * (gv[u] - si);
        }
        out = si + G.filter * sum;
    } else {
        const int64_t b = i - G.nfluid;
        float cn = 0.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) cn += lane_c<Q>(a, d) * S.normals[b * D + d];
        const int src = cn > 0.0f ? L::bb(a) : a;
        out = s1[i * RS + src];
    }
    if (ln.active) S.g[i * RS + a] = out;
}

// ---------------------------------------------------------------------------
// Classic RK4 without an accumulator buffer: the stage outputs are kept,
//   y1 = y + dt/2 k1 (A), y2 = y + dt/2 k2 (B), y3 = y + dt k3 (C),
// and stage 4 forms y + ((y1 - y) + 2 (y2 - y) + (y3 - y)) / 3 + dt/6 k4,
// which is y + dt/6 (k1 + 2 k2 + 2 k3 + k4) (rk4_solver.cuh:48-171).  C is
// the buffer S.acc.  Per node this moves 11 rows per step instead of 14.
// ---------------------------------------------------------------------------
__device__ __forceinline__ const float* stage_in(const StepBuffers& S, int stage) {
    return stage == 1 ? S.g : stage == 2 ? S.A : stage == 3 ? S.B : S.acc;
}
__device__ __forceinline__ float* stage_out(const StepBuffers& S, int stage) {
    return stage == 1 ? S.A : stage == 2 ? S.B : stage == 3 ? S.acc : S.g;
}
template <int STAGE>
__device__ __forceinline__ float rk4_out(float y, float k, float dt, float y1, float y2,
                                         float y3) {
    if (STAGE == 1 || STAGE == 2) return y + (dt * 0.5f) * k;
    if (STAGE == 3) return y + dt * k;
    return y + (((y1 - y) + 2.0f * (y2 - y)) + (y3 - y)) * (1.0f / 3.0f) + (dt * (1.0f / 6.0f)) * k;
}
// Stage 4 in two parts with the arithmetic of rk4_out<4>: the part without
// the rate, y + ((y1 - y) + 2 (y2 - y) + (y3 - y)) / 3, formed before the
// gather (four live values instead of sixteen), then + dt/6 k.
__device__ __forceinline__ float rk4_pre(float y, float y1, float y2, float y3) {
    return y + (((y1 - y) + 2.0f * (y2 - y)) + (y3 - y)) * (1.0f / 3.0f);
}
__device__ __forceinline__ float4 rk4_pre4(const float4& y, const float4& y1, const float4& y2,
                                           const float4& y3) {
    return make_float4(rk4_pre(y.x, y1.x, y2.x, y3.x), rk4_pre(y.y, y1.y, y2.y, y3.y),
                       rk4_pre(y.z, y1.z, y2.z, y3.z), rk4_pre(y.w, y1.w, y2.w, y3.w));
}
__device__ __forceinline__ float4 rk4_post4(const float4& pre, const float4& k, float dt) {
    const float c = dt * (1.0f / 6.0f);
    return make_float4(pre.x + c * k.x, pre.y + c * k.y, pre.z + c * k.z, pre.w + c * k.w);
}

template <int STAGE>
__device__ __forceinline__ float4 rk4_out4(const float4& y, const float4& k, float dt,
                                           const float4& y1, const float4& y2, const float4& y3) {
    return make_float4(rk4_out<STAGE>(y.x, k.x, dt, y1.x, y2.x, y3.x),
                       rk4_out<STAGE>(y.y, k.y, dt, y1.y, y2.y, y3.y),
                       rk4_out<STAGE>(y.z, k.z, dt, y1.z, y2.z, y3.z),
                       rk4_out<STAGE>(y.w, k.w, dt, y1.w, y2.w, y3.w));
}

// ---------------------------------------------------------------------------
// RK4 stage (rk4_solver.cuh:48-171 with pde_right_hand_side.cuh:82-176):
// rate_alpha(i) = -sum_s (c_alpha . W(s,i)) g_in(idx(s,i))_alpha on fluid
// nodes (rates are zero on boundary nodes, so their state is constant
// through the stages).
//   stage 1: A = g + dt/2 r            (and A = B = C = g on boundary rows)
//   stage 2: B = g + dt/2 r             (gathers A)
//   stage 3: C = g + dt r               (gathers B)
//   stage 4: g = g + ((A-g) + 2(B-g) + (C-g))/3 + dt/6 r   (gathers C; in
//            place: each node reads only its own g row here)
// ---------------------------------------------------------------------------
template <int Q, int STAGE, int KT, int BATCH>
__global__ __launch_bounds__(kBlock) void k_stage(StepGeom G, StepBuffers S, BatchArgs B) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS, REC = 1 + D, KMAX = KT ? KT : 64;
    constexpr int NPB = Tile<Q>::NPB;
    __shared__ float srec[NPB * KMAX * REC];
    const BlockCtx bc = locate(B, G, S);
    const int K = KT ? KT : G.k;
    const int64_t node0 = xcd_tile(bc.b, bc.nb) * NPB;
    const Lanes<Q> ln;
    const int64_t i = node0 + ln.slot;
    const int a = ln.a;
    const int64_t o = i * RS + a;
    const float* gin = stage_in(S, STAGE);
    float* gout = stage_out(S, STAGE);
    // own rows first: their HBM latency overlaps the record staging
    const bool fluid = i < G.nfluid;
    const float g0 = (i < G.N) ? S.g[o] : 0.0f;
    float y1 = 0.0f, y2 = 0.0f, y3 = 0.0f;
    if (STAGE == 4 && fluid) {
        y1 = S.A[o];
        y2 = S.B[o];
        y3 = S.acc[o];
    }
    const int64_t nf = G.nfluid - node0;
    stage_contiguous<REC>(S.slab, node0, nf <= 0 ? 0 : (nf < NPB ? int(nf) : NPB), K, srec);
    __syncthreads();
    if (STAGE == 1 && !fluid && i < G.N) {
        if (ln.active) {
            S.A[o] = g0;
            S.B[o] = g0;
            S.acc[o] = g0;
        }
        return;
    }
    if (!fluid) return;
    const float rate = -stencil_sum<Q, KT, BATCH>(srec + ln.slot * K * REC, K, gin, a);
    if (!ln.active) return;
    gout[o] = rk4_out<STAGE>(g0, rate, G.dt, y1, y2, y3);
}

// ---------------------------------------------------------------------------
// D3Q27 kernels with four distributions per lane.  A node's 128-byte row is
// served by 8 lanes, each holding alpha = 4q .. 4q+3 (q = lane % 8; the last
// five slots of the row are padding with zero lattice velocity), so a
// wavefront covers 8 nodes and a block one tile of 32 consecutive nodes.
// Neighbour rows are gathered from HBM/L2 with one 16-byte load per lane,
// or, with tile unions enabled, the distinct stencil neighbours of a tile
// (about 220 rows for 1024 stencil entries on jittered grids) are loaded
// once into LDS and read from there.  Rows in HBM are addressed through
// buffer resources with 32-bit byte offsets (row j at j * 128), which
// bounds the node count by 2^25; loads past the end of a buffer return
// zero instead of faulting.
// ---------------------------------------------------------------------------
constexpr int kV4Lanes = 8;                   // lanes per node

__device__ __forceinline__ double shfl_xor_f64(double v, int m) {
    const long long b = __double_as_longlong(v);
    const int lo = __shfl_xor(int(b & 0xffffffffll), m), hi = __shfl_xor(int(b >> 32), m);
    return __longlong_as_double((long long)(uint32_t(lo)) | ((long long)hi << 32));
}
constexpr int kV4Nodes = kBlock / kV4Lanes;  // nodes per 256-thread block
constexpr int kV4DefaultNodes = 32;          // nodes per block of the stencil kernels

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, uint64_t bytes) {
    const uint32_t n = bytes > 0xffffffffull ? 0xffffffffu : uint32_t(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), short(0), int(n), 0x00020000);
}

template <int AUX = 0>
__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]),
                       __uint_as_float(v[2]), __uint_as_float(v[3]));
}

__device__ __forceinline__ void st4(float* base, uint32_t off, float4 v) {
    *reinterpret_cast<float4*>(reinterpret_cast<char*>(base) + off) = v;
}

__device__ __forceinline__ float get(const float4& v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}

__device__ __forceinline__ void set(float4& v, int e, float x) {
    if (e == 0) v.x = x;
    else if (e == 1) v.y = x;
    else if (e == 2) v.z = x;
    else v.w = x;
}

// lattice velocity component of slot alpha (zero on padding slots)
__device__ __forceinline__ float c27(int a, int d) {
    return a < 27 ? float(Lattice<27>::c(a, d)) : 0.0f;
}

// T_d[e] = sum_s W_d(s) g_s[4q + e] over the staged records rp
// ({index field, W_x, W_y, W_z} per neighbour), the neighbour rows fetched
// by fetch(index field); returns -sum_d c_d T_d per slot.
template <int K, int BATCH, class Fetch>
__device__ __forceinline__ float4 rate_fetch(const float4* rp, Fetch fetch, int q) {
    float tx[4] = {0.f, 0.f, 0.f, 0.f}, ty[4] = {0.f, 0.f, 0.f, 0.f},
          tz[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s0 = 0; s0 < K; s0 += BATCH) {
        float4 rec[BATCH], gv[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) rec[u] = rp[s0 + u];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) gv[u] = fetch(__float_as_uint(rec[u].x));
#pragma unroll
        for (int u = 0; u < BATCH; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v = get(gv[u], e);
                tx[e] += rec[u].y * v;
                ty[e] += rec[u].z * v;
                tz[e] += rec[u].w * v;
            }
    }
    float4 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 4 * q + e;
        set(r, e, -(c27(a, 0) * tx[e] + c27(a, 1) * ty[e] + c27(a, 2) * tz[e]));
    }
    return r;
}

// rate_fetch over packed records (packed_div_units layout): per group of 8
// records six weight units and one unit of 16-bit union positions.  The
// same products in the same order as rate_fetch, so the two agree bit for
// bit.
#ifndef NAGA_PACK_SUB
#define NAGA_PACK_SUB 4
#endif
template <int K, class Fetch>
__device__ __forceinline__ float4 rate_fetch_packed(const float4* rp, Fetch fetch, int q,
                                                    int ng = K / 8) {
    static_assert(K % 8 == 0, "packed records come in groups of 8");
    constexpr int H = NAGA_PACK_SUB;  // records gathered per sub-batch: 8 or 4
    static_assert(H == 8 || H == 4, "sub-batch of 8 or 4 records");
    float tx[4] = {0.f, 0.f, 0.f, 0.f}, ty[4] = {0.f, 0.f, 0.f, 0.f},
          tz[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int g = 0; g < K / 8; ++g) {
        if (g >= ng) break;  // ragged rows: the tile's ng groups
        const float4 iu = rp[7 * g + 6];
#pragma unroll
        for (int h = 0; h < 8; h += H) {
            // records h .. h + H - 1: weights in units 3h/4 .. 3(h + H)/4 - 1
            float4 wu[3 * H / 4], gv[H];
#pragma unroll
            for (int r = 0; r < H; ++r) {
                const uint32_t word = __float_as_uint(get(iu, (h + r) >> 1));
                gv[r] = fetch(((h + r) & 1) ? word >> 16 : word & 0xffffu);
            }
#pragma unroll
            for (int w = 0; w < 3 * H / 4; ++w) wu[w] = rp[7 * g + 3 * h / 4 + w];
#pragma unroll
            for (int r = 0; r < H; ++r) {
                const float wx = get(wu[(3 * r) >> 2], (3 * r) & 3);
                const float wy = get(wu[(3 * r + 1) >> 2], (3 * r + 1) & 3);
                const float wz = get(wu[(3 * r + 2) >> 2], (3 * r + 2) & 3);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = get(gv[r], e);
                    tx[e] += wx * v;
                    ty[e] += wy * v;
                    tz[e] += wz * v;
                }
            }
        }
    }
    float4 r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 4 * q + e;
        set(r, e, -(c27(a, 0) * tx[e] + c27(a, 1) * ty[e] + c27(a, 2) * tz[e]));
    }
    return r;
}

// LDS DMA (gfx950 buffer_load ... lds): one wave instruction copies 64
// 16-byte pieces, lane l's piece landing at lds_base + 16 l whatever its
// source offset, with no VGPR round trip; the base must be wave-uniform.
// Offsets past the buffer resource's size read as zero.
typedef __attribute__((address_space(3))) void lds_void_t;
template <int AUX = 0>
__device__ __forceinline__ void dma16(__amdgpu_buffer_rsrc_t src, uint32_t off, float4* lds_base) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(src, (lds_void_t*)(lds_base), 16, off, 0, 0, AUX);
}

// Cache policy of the record streams (read once per pass): non-temporal,
// so they do not displace the state rows in L2 and the last-level cache.
constexpr int kStreamPolicy = NAGA_RECORD_POLICY;

// Copy n16 <= NMAX contiguous 16-byte pieces of src (from offset 0) to dst
// with the NT threads of the block.
template <int NMAX, int NT>
__device__ __forceinline__ void dma_contiguous(__amdgpu_buffer_rsrc_t src, float4* dst, int n16) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = NT / 64;
    constexpr int per_wave = (NMAX + 64 * nw - 1) / (64 * nw);
#pragma unroll
    for (int t = 0; t < per_wave; ++t) {
        const int c = wave + t * nw;
        if (c * 64 + lane < n16)
            dma16<kStreamPolicy>(src, uint32_t(c * 64 + lane) * 16u, dst + c * 64);
    }
}

// The union range [u0, u0 + U) of tile t, loaded through the scalar cache
// (wave-uniform values in SGPRs, off the vector memory counter).
__device__ __forceinline__ void union_range(const int64_t* uoff, int64_t t, int64_t& u0, int& U) {
    typedef const __attribute__((address_space(4))) int64_t const_i64;
    const_i64* p = (const_i64*)(uoff);
    u0 = p[t];
    U = int(p[t + 1] - u0);
}

// Stage the first Uc union rows of a tile into LDS (urows[r * 8 + q] =
// quarter q of row ul[r]): eight rows per wave instruction, NT threads per
// block.  union_indices loads every index a wave needs (past Uc they read
// as zero through the buffer resource, so the loads are unconditional);
// union_copy then issues all the copies, so a wave waits for one index
// round trip and keeps its copies in flight together.
template <int NT>
struct UnionIdx {
    static constexpr int kChunks = kUnionRowsMax / (8 * (NT / 64));  // 8-row chunks per wave
    uint32_t j[kChunks];
};
template <int NT>
__device__ __forceinline__ UnionIdx<NT> union_indices(const uint32_t* ul, int Uc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = NT / 64;
    const __amdgpu_buffer_rsrc_t ur = rsrc(ul, uint64_t(Uc) * 4u);
    UnionIdx<NT> x;
#pragma unroll
    for (int t = 0; t < UnionIdx<NT>::kChunks; ++t) {
        const int r = (wave + t * nw) * 8 + (lane >> 3);
        x.j[t] = __builtin_amdgcn_raw_buffer_load_b32(ur, uint32_t(r) * 4u, 0, 0);
    }
    return x;
}
template <int NT>
__device__ __forceinline__ void union_copy(float4* urows, const UnionIdx<NT>& x, int Uc,
                                           __amdgpu_buffer_rsrc_t src) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int nw = NT / 64;
    const uint32_t q16 = uint32_t(lane & 7) * 16u;
#pragma unroll
    for (int t = 0; t < UnionIdx<NT>::kChunks; ++t) {
        const int c = wave + t * nw;
        if (c * 8 < Uc && c * 8 + (lane >> 3) < Uc)
            dma16(src, x.j[t] * 128u + q16, urows + c * 64);
    }
}

// Wait for this wave's LDS DMA copies, then for the whole block.
__device__ __forceinline__ void dma_fence() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// The work of one block of a union kernel.  Tile entry e >= 0 is tile e
// (only its first half if the tile was split at setup); e < 0 is the
// second half of split tile split_tiles[-1 - e].  The block serves node
// slots [lo, hi) of the tile's cnt fluid nodes (first: it also serves the
// tile's non-fluid nodes) and stages the union rows [u0, u0 + U) of the
// list, in which those nodes' records hold positions.
struct UnionWork {
    int64_t tile, u0;
    int lo, hi, U;
    bool first;
};

template <int NPB, bool TU>
__device__ __forceinline__ UnionWork union_work(const StepGeom& G, const StepBuffers& S,
                                                int64_t e) {
    typedef const __attribute__((address_space(4))) int32_t const_i32;
    UnionWork w;
    w.first = e >= 0;
    w.tile = e >= 0 ? e : ((const_i32*)S.split_tiles)[-1 - e];
    const int64_t nf = G.nfluid - w.tile * NPB;
    const int cnt = nf <= 0 ? 0 : (nf < NPB ? int(nf) : NPB);
    w.lo = 0;
    w.hi = cnt;
    w.u0 = 0;
    w.U = 0;
    if (TU && cnt > 0) {
        union_range(S.uoff, w.tile, w.u0, w.U);
        if (w.U > S.ucap && S.usplit) {
            const int A = ((const_i32*)S.usplit)[w.tile];
            if (A > 0 && e >= 0) {
                w.hi = NPB / 2;
                w.U = A;
            } else if (A > 0) {
                w.lo = NPB / 2;
                w.u0 += A;
                w.U -= A;
            }
        }
    }
    return w;
}

// Record rows of a union block's tile: groups of 8 per node, units per node
// and the tile's first unit (fixed width: K / 8, KP, tile * NPB * KP;
// ragged: from the per-tile tables, which have one entry per tile of fluid
// nodes; the tiles past them, of boundary nodes only, hold no records and
// read nothing).
struct TileRecs {
    int ng, kp;
    int64_t base;
};
template <bool RG, int K, int KP, int NPB>
__device__ __forceinline__ TileRecs tile_recs(const int32_t* tile_ng, const int64_t* off,
                                              int64_t tile, bool fluid_tile, bool filter) {
    if (!RG) return {K / 8, KP, tile * NPB * KP};
    if (!fluid_tile) return {0, 0, 0};
    typedef const __attribute__((address_space(4))) int32_t const_i32;
    typedef const __attribute__((address_space(4))) int64_t const_i64;
    const int ng = ((const_i32*)tile_ng)[tile];
    return {ng, filter ? ragged_filter_units(ng) : ragged_div_units(ng), ((const_i64*)off)[tile]};
}

// Tile entry of block b of nb of a kernel over the tiles: the regular
// tiles first, then (with tile unions) the second halves of split tiles.
template <bool TU>
__device__ __forceinline__ int64_t union_entry(const StepBuffers& S, int64_t b, int64_t nb) {
    const int64_t t = xcd_tile(b, nb), nreg = TU ? nb - S.n_split : nb;
    return t < nreg ? t : -1 - (t - nreg);
}

// Collision for one D3Q27 scene: one wave per block of 64 nodes, whose
// rows arrive in LDS by DMA (no VGPR round trip), are collided in place
// and stored as coalesced 16-byte pieces.  Piece j of chunk node n sits at
// slot j * 64 + (n ^ j): the DMA writes the wave's 64 slots of one piece
// contiguously, and both a lane's reads of its own row and the store of
// the chunk are free of bank conflicts.  Three waves per SIMD (168 VGPRs)
// measured faster than the VGPR-staged kernel at any occupancy.
constexpr int kP1Chunk = 64;
#ifndef NAGA_P1_DMA
#define NAGA_P1_DMA 1
#endif

// The DMA collision runs with the four-per-lane D3Q27 kernels (their
// 2^25-node bound keeps its 32-bit row offsets exact); the flag is uniform
// over a batch group.
bool p1_uses_dma(int Q, bool unions) { return NAGA_P1_DMA && Q == 27 && unions; }

// The DMA collision carries the PML local terms (k_p1_pml) unless
// NAGA_P1_PML=0 selects the separate PML pass (read once per process).
bool p1_fuses_pml(int Q, bool unions) {
    static const bool on = [] {
        const char* e = std::getenv("NAGA_P1_PML");
        return !(e && e[0] == '0');
    }();
    return on && p1_uses_dma(Q, unions);
}

struct XorRow {
    float* buf;  // pieces of the chunk, swizzled as above
    int t;       // node within the chunk
    __device__ float& at(int a) const {
        const int j = a >> 2;
        return buf[(j * kP1Chunk + (t ^ j)) * 4 + (a & 3)];
    }
    __device__ float get(int a) const { return at(a); }
    __device__ void set(int a, float v) const { at(a) = v; }
};

__global__ __launch_bounds__(kP1Chunk) __attribute__((amdgpu_waves_per_eu(3))) void k_p1_dma(
    StepGeom G, StepBuffers S, CollisionParams cp, BatchArgs B) {
    __shared__ float4 buf[8 * kP1Chunk];
    const int t = threadIdx.x;
    const BlockCtx bc = locate(B, G, S);
    if (B.scenes) cp = B.scenes[bc.scene].cp;
    const __amdgpu_buffer_rsrc_t gsrc = rsrc(S.g, uint64_t(G.N) * 128u);
    const int64_t node0 = bc.b * kP1Chunk;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        dma16(gsrc, uint32_t(node0 + (t ^ j)) * 128u + uint32_t(j) * 16u, buf + j * kP1Chunk);
    dma_fence();
    if (node0 + t < G.N) S.ru[node0 + t] = p1_node<27>(G, S, cp, node0 + t, XorRow{(float*)buf, t});
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int p = r * kP1Chunk + t, n = p >> 3, j = p & 7;
        if (node0 + n < G.N)
            st4(S.A, uint32_t(node0 + n) * 128u + uint32_t(j) * 16u, buf[j * kP1Chunk + (n ^ j)]);
    }
}

// Collision and PML local terms of one D3Q27 scene in one pass, one wave
// per block of 64 nodes (as k_p1_dma, whose independent one-wave blocks
// overlap each other's memory and FP64 work): the wave collides its rows in
// LDS, then its 64 lanes, eight per node (four distributions each), apply
// the PML local terms of k_pml_local_v4 in FP32 to the block's nodes whose
// Q values are maintained, subtracting the layer damping from the collided
// rows before they are stored.  Blocks that hold layer nodes write their
// partial layer sums (column-major, part = block).  Saves the PML pass's
// second read and write of the layer rows, its read of (rho - 1, u) and one
// launch.
constexpr int kP1Fused = kP1Chunk;  // nodes per fused block
#ifndef NAGA_P1_PREFETCH
#define NAGA_P1_PREFETCH 0
#endif

__device__ __forceinline__ bool pml_maintained(const StepGeom& G, int64_t i) {
    return (i >= G.nbulk && i < G.nfluid) || (G.pml_ref && i < G.nlayer && i < G.nfluid);
}

__global__ __launch_bounds__(kP1Fused) __attribute__((amdgpu_waves_per_eu(3))) void k_p1_pml(
    StepGeom G, StepBuffers S, CollisionParams cp, BatchArgs B) {
    __shared__ float4 buf[8 * kP1Chunk];  // collided rows, XorRow layout
    __shared__ float4 rul[kP1Chunk];      // (rho - 1, u) per node
    __shared__ float4 prl[kP1Chunk];      // the previous step's (rho - 1, u) per node
    const int t = threadIdx.x;
    const BlockCtx bc = locate(B, G, S);
    if (B.scenes) cp = B.scenes[bc.scene].cp;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const __amdgpu_buffer_rsrc_t gsrc = rsrc(S.g, rows);
    const int64_t node0 = bc.b * kP1Chunk;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        dma16(gsrc, uint32_t(node0 + (t ^ j)) * 128u + uint32_t(j) * 16u, buf + j * kP1Chunk);
    if (G.nlayer > 0)  // read by the PML terms below, after the collision
        dma16(rsrc(S.E, uint64_t(G.N) * 16u), uint32_t(node0 + t) * 16u, prl);
    dma_fence();
    if (node0 + t < G.N) rul[t] = p1_node<27>(G, S, cp, node0 + t, XorRow{(float*)buf, t});
    __syncthreads();
    const int64_t end = node0 + kP1Chunk;
    const bool any_pml = G.nlayer > 0 && node0 < G.nfluid &&
                         (end > G.nbulk || (G.pml_ref && node0 < G.nlayer));
    const bool layer_block = G.nlayer > 0 && node0 < G.nfluid && end > G.nbulk;
    const int slot = t / kV4Lanes, q = t % kV4Lanes;
    double s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
    if (any_pml) {
        const float hdt = G.dt * 0.5f;
        const __amdgpu_buffer_rsrc_t rQp = rsrc(S.Qp, rows);
        constexpr int kRounds = kP1Chunk / 8;
#if NAGA_P1_PREFETCH
        // every round's loads first: one memory round trip for the phase
        float4 qp[kRounds];
        float sg[kRounds];
#pragma unroll
        for (int rd = 0; rd < kRounds; ++rd) {
            const int64_t i = node0 + rd * 8 + slot;
            const bool layer = i >= G.nbulk && i < G.nfluid;
            qp[rd] = ld4(rQp, uint32_t(i) * 128u + uint32_t(q) * 16u);  // past N: zero
            sg[rd] = layer ? S.sigma[i - G.nbulk] : 0.0f;
        }
#pragma unroll
#else
#pragma unroll 2
#endif
        for (int rd = 0; rd < kRounds; ++rd) {
            const int n = rd * 8 + slot;
            const int64_t i = node0 + n;
            if (i >= G.N || !pml_maintained(G, i)) continue;
            const float4 prv = prl[n];  // the previous step's (rho - 1, u)
            const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
            const bool layer = i >= G.nbulk && i < G.nfluid;
#if NAGA_P1_PREFETCH
            const float4 p4 = qp[rd];
            const float sig = sg[rd];
#else
            const float4 p4 = ld4(rQp, own);
            const float sig = layer ? S.sigma[i - G.nbulk] : 0.0f;
#endif
            const float4 ru = rul[n];
            float4& a4 = buf[q * kP1Chunk + (n ^ q)];
            float4 av = a4;
            float4 po, qo;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int a = 4 * q + e;
                if (a >= 27) {  // padding slots stay zero
                    set(po, e, 0.0f);
                    set(qo, e, 0.0f);
                    continue;
                }
                const float fte = feq_dev32<27>(ru, a);
                const float q1 = (fte + feq_dev32<27>(prv, a)) * hdt;
                const float q2 = (q1 + get(p4, e)) * hdt;
                set(po, e, q1);
                set(qo, e, q2);
                if (layer) {
                    set(av, e, get(av, e) - sig * (3.0f * fte + 3.0f * sig * q1 + sig * sig * q2));
                    s1[e] += double(q1);
                    s2[e] += double(q2);
                }
            }
            // the previous values were staged in LDS before the collision
            if (q == 0) reinterpret_cast<float4*>(S.E)[i] = ru;
            st4(S.Qp, own, po);
            if (G.pml_ref ? i < G.nlayer : layer) st4(S.Q2v, own, qo);  // rows the layer field reads
            if (layer) a4 = av;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int p = r * kP1Chunk + t, n = p >> 3, j = p & 7;
        if (node0 + n < G.N)
            st4(S.A, uint32_t(node0 + n) * 128u + uint32_t(j) * 16u, buf[j * kP1Chunk + (n ^ j)]);
    }
    if (!layer_block) return;
    // sums over the block's eight node slots (lanes q, q + 8, ..., q + 56)
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int o = 8; o < 64; o <<= 1) {
            s1[e] += shfl_xor_f64(s1[e], o);
            s2[e] += shfl_xor_f64(s2[e], o);
        }
    if (slot == 0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int a = 4 * q + e;
            if (a < 27) {
                S.partial[int64_t(a) * S.nparts + bc.b] = s1[e];
                S.partial[int64_t(27 + a) * S.nparts + bc.b] = s2[e];
            }
        }
    }
}

// Waves per SIMD the stage kernels are compiled for: three 256-thread
// blocks per CU with 32-node tiles, two 512-thread blocks with 64.
template <int NPB>
constexpr int stage_waves() { return NPB >= 64 ? 4 : 3; }

// RK4 stage, D3Q27, four per lane, neighbour rows read from the tile's
// union staged in LDS (unions larger than ucap rows read through the list).
template <int STAGE, int K, int BATCH, bool TU, int NPB, bool RG = false>
__global__ __launch_bounds__(NPB * kV4Lanes) __attribute__((amdgpu_waves_per_eu(stage_waves<NPB>()))) void k_stage_u(StepGeom G, StepBuffers S,
                                                            BatchArgs B) {
    constexpr int NT = NPB * kV4Lanes;
    // record rows padded by one entry (or packed, with tile unions): the
    // eight nodes of a wavefront read their records from distinct LDS banks
    constexpr int KP = TU ? packed_div_units(K) : K + 1;
    __shared__ float4 srec[NPB * KP];
    extern __shared__ float4 urows[];
    const BlockCtx bc = locate(B, G, S);
    const UnionWork w = union_work<NPB, TU>(G, S, union_entry<TU>(S, bc.b, bc.nb));
    const int64_t node0 = w.tile * NPB;
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t i = node0 + slot;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
    const float* gin_p = stage_in(S, STAGE);
    float* gout = stage_out(S, STAGE);
    const bool fluid = i < G.nfluid;
    const bool mine = slot >= w.lo && slot < w.hi;  // a fluid node this block serves
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // stage 4 reads the node's own row of its gather input (C): it is in the
    // tile's union (record 0 is the node itself), so it comes from LDS
    // instead of a second read from HBM.  (Stage 1 keeps its early own-row
    // load of g: issued before the union index round trip, it brings the
    // tile's own rows in ahead of the copies; from LDS the stage measured
    // 7 us slower.)
    const bool own_lds = TU && STAGE == 4 && S.self_first && mine && w.U <= S.ucap;
    const float4 g0 = ld4(rsrc(S.g, rows), own);
    const __amdgpu_buffer_rsrc_t gin = rsrc(gin_p, rows);
    const int U = w.U;
    const uint32_t* ul = TU ? S.ulist + w.u0 : nullptr;  // TU: tile unions built
    // records: rows padded by one record (distinct banks per node), copied
    // through VGPRs, or with tile unions by LDS DMA from packed rows
    const TileRecs tr =
        tile_recs<RG, K, KP, NPB>(S.tile_ng, S.rec_off, w.tile, w.hi > w.lo, false);
    const int kp = tr.kp;
    if (w.hi > w.lo) {
        if (!TU) {
            const float4* src = reinterpret_cast<const float4*>(S.slab) + node0 * K;
            for (int t = threadIdx.x; t < w.hi * K; t += NT) srec[t + t / K] = src[t];
        } else {
            // order: union indices, record copies, then the row copies (which
            // wait for the indices only)
            const int Uc = U < S.ucap ? U : S.ucap;
            const UnionIdx<NT> ux = union_indices<NT>(ul, Uc);  // unconditional: no wait here
            __builtin_amdgcn_sched_barrier(0);
            const int n = w.hi - w.lo;
            dma_contiguous<NPB * KP, NT>(
                rsrc(S.slab + (tr.base + int64_t(w.lo) * kp) * 4, uint64_t(n) * kp * 16u),
                srec + w.lo * kp, n * kp);
            __builtin_amdgcn_sched_barrier(0);
            union_copy<NT>(urows, ux, Uc, gin);
        }
    }
    // stage 4: the stage outputs, read for the last time (non-temporal),
    // are loaded once the copies are issued (no registers held through
    // them) and folded before the gather
    float4 y1 = z4, y2 = z4, y3 = z4;
    if (STAGE == 4) {
        y1 = ld4<kStreamPolicy>(rsrc(S.A, rows), own);
        y2 = ld4<kStreamPolicy>(rsrc(S.B, rows), own);
        if (!own_lds) y3 = ld4<kStreamPolicy>(rsrc(S.acc, rows), own);
    }
    if (TU) dma_fence();
    else __syncthreads();
    if (own_lds) {
        const uint32_t p0 = __float_as_uint(srec[slot * kp + 6].x) & 0xffffu;
        y3 = urows[p0 * kV4Lanes + q];
    }
    if (STAGE == 1 && w.first && !fluid && i < G.N) {
        st4(S.A, own, g0);
        st4(S.B, own, g0);
        st4(S.acc, own, g0);
        return;
    }
    if (!mine) return;
    float4 pre = STAGE == 4 ? rk4_pre4(g0, y1, y2, y3) : g0;
    // keep the fold here (the compiler would otherwise sink it to its use
    // and hold the sixteen inputs through the gather)
    if (STAGE == 4) asm volatile("" : "+v"(pre.x), "+v"(pre.y), "+v"(pre.z), "+v"(pre.w));
    float4 r;
    if (!ul)  // records hold node indices
        r = rate_fetch<K, BATCH>(
            srec + slot * kp,
            [&](uint32_t j) { return ld4(gin, j * 128u + uint32_t(q) * 16u); }, q);
    else if (U <= S.ucap)
        r = rate_fetch_packed<K>(srec + slot * kp,
                                 [&](uint32_t j) { return urows[j * kV4Lanes + q]; }, q, tr.ng);
    else  // union over the cap: every row read from HBM through the list
        r = rate_fetch_packed<K>(srec + slot * kp,
                                 [&](uint32_t j) {
                                     return ld4(gin, ul[j] * 128u + uint32_t(q) * 16u);
                                 },
                                 q, tr.ng);
    st4(gout, own, STAGE == 4 ? rk4_post4(pre, r, G.dt) : rk4_out4<STAGE>(g0, r, G.dt, z4, z4, z4));
}

// ---------------------------------------------------------------------------
// D2Q9 stencil passes with three distributions per lane: a node's 36-byte
// row is three 12-byte pieces, so one wavefront gathers 21 neighbour rows
// per instruction (the lane-over-alpha kernels gather 7), and the records
// are read by 3 lanes per node instead of 9.  The records of the block's
// nodes are staged in LDS with one pad float per node row (distinct banks
// for the 21 nodes of a wavefront).  Same products in the same order as the
// lane-over-alpha kernels.
// ---------------------------------------------------------------------------
constexpr int kQ9Lanes = 3;                          // lanes per node
constexpr int kQ9Npw = 21;                           // nodes per wavefront
constexpr int kQ9Npb = kQ9Npw * (kBlock / 64);       // nodes per block (84)
static_assert(kQ9Npb == kQ9TileNodes, "D2Q9 union tiles are the three-per-lane blocks");

struct Q9Lane {
    int slot, q;   // node slot in the block, piece (distributions 3q .. 3q + 2)
    bool active;   // lane 63 of a wavefront serves no node
    __device__ Q9Lane() {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int sub = lane / kQ9Lanes;
        active = sub < kQ9Npw;
        q = active ? lane - sub * kQ9Lanes : 0;
        slot = wave * kQ9Npw + (active ? sub : 0);
    }
};

typedef float f3v __attribute__((ext_vector_type(3)));
__device__ __forceinline__ f3v ld3(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
    f3v o;
    o.x = __uint_as_float(v[0]);
    o.y = __uint_as_float(v[1]);
    o.z = __uint_as_float(v[2]);
    return o;
}
__device__ __forceinline__ void st3(__amdgpu_buffer_rsrc_t r, uint32_t off, f3v x) {
    typedef unsigned u3v __attribute__((ext_vector_type(3)));
    u3v v;
    v.x = __float_as_uint(x.x);
    v.y = __float_as_uint(x.y);
    v.z = __float_as_uint(x.z);
    __builtin_amdgcn_raw_buffer_store_b96(v, r, off, 0, 0);
}

// records of nodes [node0, node0 + cnt) (rows of RW floats) into lds rows
// of RW + 1 floats by LDS DMA: the block's waves copy 64-float chunks of
// the padded layout, lane l of chunk c filling LDS float 64 c + l with its
// source float (a pad slot takes the tile's first float; lanes past the
// last row copy nothing).  Wait with dma_fence.
__device__ __forceinline__ void stage_padded(const float* slab, int64_t node0, int cnt, int RW,
                                             float* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int RP = RW + 1, tot = cnt * RP;
    const __amdgpu_buffer_rsrc_t src = rsrc(slab + node0 * int64_t(RW), uint64_t(cnt) * RW * 4u);
    const int nchunks = (tot + 63) / 64;
    // t / RP by a float reciprocal: t < 2^15 and RP <= 193 keep (t + 1/2) / RP
    // at least 1/(2 RP) from an integer, far above the rounding error
    const float rinv = 1.0f / float(RP);
    for (int c = wave; c < nchunks; c += kBlock / 64) {
        const int t = c * 64 + lane;
        const int node = int((float(t) + 0.5f) * rinv), off = t - node * RP;
        const uint32_t e = off < RW ? uint32_t(node * RW + off) : 0u;
        if (t < tot)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(src, (lds_void_t*)(lds + c * 64), 4, e * 4u,
                                                     0, 0, kStreamPolicy);
    }
}

__device__ __forceinline__ float c9(int a, int d) { return float(Lattice<9>::c(a, d)); }

// sum_s (c_a . W_s) fin(idx_s)_a for the lane's three distributions, the
// neighbour pieces fetched by fetch(index field) (a node index, or a
// position in the tile's union)
template <int BATCH, class Fetch>
__device__ __forceinline__ f3v stencil_sum3(const float* rp, int K, int q, Fetch fetch) {
    float cx[3], cy[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        cx[e] = c9(3 * q + e, 0);
        cy[e] = c9(3 * q + e, 1);
    }
    f3v acc = {0.0f, 0.0f, 0.0f};
    for (int s0 = 0; s0 < K; s0 += BATCH) {
        float wx[BATCH], wy[BATCH];
        f3v gv[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            const int s = s0 + u;
            const bool in = s < K;
            const float* r = rp + (in ? s : 0) * 3;
            const uint32_t j = __float_as_uint(r[0]);
            wx[u] = in ? r[1] : 0.0f;
            wy[u] = in ? r[2] : 0.0f;
            gv[u] = fetch(j);
        }
#pragma unroll
        for (int u = 0; u < BATCH; ++u) {
            acc.x += (cx[0] * wx[u] + cy[0] * wy[u]) * gv[u].x;
            acc.y += (cx[1] * wx[u] + cy[1] * wy[u]) * gv[u].y;
            acc.z += (cx[2] * wx[u] + cy[2] * wy[u]) * gv[u].z;
        }
    }
    return acc;
}

// D2Q9 tile unions: the distinct stencil neighbours of the 84 fluid nodes
// of a block's tile, their rows copied to LDS after the block's records
// (urows[9 r + 3 q + e] = distribution 3 q + e of row ul[r]) so every
// gather of the stencil sums reads LDS.  The block's threads copy the
// 3 Uc 12-byte pieces: every index load first (past Uc they read zero
// through the buffer resource), then every row load, then the stores.
constexpr int kQ9CopyIt = (3 * kQ9UnionRowsMax + kBlock - 1) / kBlock;
struct Q9UnionIdx {
    uint32_t j[kQ9CopyIt];
};
__device__ __forceinline__ Q9UnionIdx q9_union_indices(const uint32_t* ul, int Uc) {
    const __amdgpu_buffer_rsrc_t ur = rsrc(ul, uint64_t(Uc) * 4u);
    const int nit = (3 * Uc + kBlock - 1) / kBlock;
    Q9UnionIdx x;
#pragma unroll
    for (int it = 0; it < kQ9CopyIt; ++it) {
        const uint32_t t = threadIdx.x + it * kBlock;
        x.j[it] = it < nit ? __builtin_amdgcn_raw_buffer_load_b32(ur, (t / 3u) * 4u, 0, 0) : 0u;
    }
    return x;
}
__device__ __forceinline__ void q9_union_rows(float* urows, const Q9UnionIdx& x, int Uc,
                                              __amdgpu_buffer_rsrc_t src) {
    const int nit = (3 * Uc + kBlock - 1) / kBlock;
    f3v v[kQ9CopyIt];
#pragma unroll
    for (int it = 0; it < kQ9CopyIt; ++it) {
        const uint32_t t = threadIdx.x + it * kBlock;
        if (it < nit) v[it] = ld3(src, x.j[it] * 36u + (t % 3u) * 12u);
    }
#pragma unroll
    for (int it = 0; it < kQ9CopyIt; ++it) {
        const int t = int(threadIdx.x) + it * kBlock;
        if (it < nit && t < 3 * Uc) {
            urows[3 * t] = v[it].x;
            urows[3 * t + 1] = v[it].y;
            urows[3 * t + 2] = v[it].z;
        }
    }
}
__device__ __forceinline__ f3v q9_lds_row(const float* urows, uint32_t p, int q) {
    const float* r = urows + p * 9u + uint32_t(q) * 3u;
    return f3v{r[0], r[1], r[2]};
}

template <int STAGE, int BATCH, bool TU = false>
__global__ __launch_bounds__(kBlock) void k_stage3(StepGeom G, StepBuffers S, BatchArgs B) {
    extern __shared__ float srec3[];
    const BlockCtx bc = locate(B, G, S);
    const int K = G.k, RW = 3 * K;
    const int64_t tile = xcd_tile(bc.b, bc.nb);
    const int64_t node0 = tile * kQ9Npb;
    const Q9Lane ln;
    const int64_t i = node0 + ln.slot;
    const uint64_t rows = uint64_t(G.N) * 36u;
    const uint32_t own = uint32_t(i) * 36u + uint32_t(ln.q) * 12u;
    const __amdgpu_buffer_rsrc_t rg = rsrc(S.g, rows);
    const __amdgpu_buffer_rsrc_t gin = rsrc(stage_in(S, STAGE), rows);
    // own rows first: their latency overlaps the record staging
    const bool fluid = i < G.nfluid;
    const f3v g0 = ld3(rg, own);  // past N: zero
    f3v y1 = {0.f, 0.f, 0.f}, y2 = y1, y3 = y1;
    if (STAGE == 4 && fluid) {
        y1 = ld3(rsrc(S.A, rows), own);
        y2 = ld3(rsrc(S.B, rows), own);
        y3 = ld3(rsrc(S.acc, rows), own);
    }
    const int64_t nf = G.nfluid - node0;
    const int cnt = nf <= 0 ? 0 : (nf < kQ9Npb ? int(nf) : kQ9Npb);
    // with tile unions: the tile's union range (tiles of fluid nodes only)
    int64_t u0 = 0;
    int U = 0;
    if (TU && cnt > 0) union_range(S.uoff, tile, u0, U);
    const int Uc = U < S.ucap ? U : S.ucap;
    float* urows = srec3 + kQ9Npb * (RW + 1);
    Q9UnionIdx ux;
    if (TU) ux = q9_union_indices(S.ulist + u0, Uc);
    stage_padded(TU ? S.uslab : S.slab, node0, cnt, RW, srec3);
    if (TU) q9_union_rows(urows, ux, Uc, gin);
    dma_fence();
    if (!ln.active || i >= G.N) return;
    if (STAGE == 1 && !fluid) {
        st3(rsrc(S.A, rows), own, g0);
        st3(rsrc(S.B, rows), own, g0);
        st3(rsrc(S.acc, rows), own, g0);
        return;
    }
    if (!fluid) return;
    // stage 4: the part without the rate, formed before the gather (three
    // live values instead of twelve; the arithmetic of rk4_out<4>)
    f3v pre = g0;
    if (STAGE == 4) {
        pre.x = rk4_pre(g0.x, y1.x, y2.x, y3.x);
        pre.y = rk4_pre(g0.y, y1.y, y2.y, y3.y);
        pre.z = rk4_pre(g0.z, y1.z, y2.z, y3.z);
        asm volatile("" : "+v"(pre.x), "+v"(pre.y), "+v"(pre.z));
    }
    const float* rp = srec3 + ln.slot * (RW + 1);
    const uint32_t lane_off = uint32_t(ln.q) * 12u;
    f3v acc;
    if (!TU) {
        acc = stencil_sum3<BATCH>(rp, K, ln.q,
                                  [&](uint32_t j) { return ld3(gin, j * 36u + lane_off); });
    } else if (U <= S.ucap) {
        acc = stencil_sum3<BATCH>(rp, K, ln.q,
                                  [&](uint32_t p) { return q9_lds_row(urows, p, ln.q); });
    } else {  // union over the cap: rows read from HBM through the list
        const uint32_t* ul = S.ulist + u0;
        acc = stencil_sum3<BATCH>(rp, K, ln.q,
                                  [&](uint32_t p) { return ld3(gin, ul[p] * 36u + lane_off); });
    }
    f3v out;
    if (STAGE == 4) {
        const float c = G.dt * (1.0f / 6.0f);
        out.x = pre.x + c * -acc.x;
        out.y = pre.y + c * -acc.y;
        out.z = pre.z + c * -acc.z;
    } else {
        out.x = rk4_out<STAGE>(g0.x, -acc.x, G.dt, y1.x, y2.x, y3.x);
        out.y = rk4_out<STAGE>(g0.y, -acc.y, G.dt, y1.y, y2.y, y3.y);
        out.z = rk4_out<STAGE>(g0.z, -acc.z, G.dt, y1.z, y2.z, y3.z);
    }
    st3(rsrc(stage_out(S, STAGE), rows), own, out);
}

// p2 for D2Q9, three per lane: the PML divergence over the listed nodes
// (the arithmetic of k_p2), each node's records staged by its three lanes.
__global__ __launch_bounds__(kBlock) void k_p2_3(StepGeom G, StepBuffers S, BatchArgs B) {
    extern __shared__ float srec3[];
    const BlockCtx bc = locate(B, G, S);
    const int K = G.k, RW = 3 * K;
    const Q9Lane ln;
    const int64_t item0 = xcd_tile(bc.b, bc.nb) * kQ9Npb, item = item0 + ln.slot;
    const bool valid = ln.active && item < S.n_pml;
    const int64_t i = valid ? S.pml_list[item] : 0;
    float* rp = srec3 + ln.slot * (RW + 1);
    const int64_t left = S.n_pml - item0;
    const int cnt = left < kQ9Npb ? int(left) : kQ9Npb;
    if (uint64_t(G.nfluid) * uint64_t(RW) * 4u < (uint64_t(1) << 32)) {
        // the listed nodes' records by LDS DMA (32-bit offsets into the slab)
        __shared__ uint32_t rows_of[kQ9Npb];
        if (int(threadIdx.x) < cnt) rows_of[threadIdx.x] = uint32_t(S.pml_list[item0 + threadIdx.x]);
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int RP = RW + 1, tot = cnt * RP;
        const __amdgpu_buffer_rsrc_t src = rsrc(S.slab, uint64_t(G.nfluid) * RW * 4u);
        const float rinv = 1.0f / float(RP);  // as in stage_padded
        for (int c = wave; c < (tot + 63) / 64; c += kBlock / 64) {
            const int t = c * 64 + lane;
            const int node = int((float(t) + 0.5f) * rinv), off = t - node * RP;
            const uint32_t e =
                (t < tot && off < RW) ? rows_of[node < cnt ? node : 0] * uint32_t(RW) + uint32_t(off)
                                      : 0u;
            if (t < tot)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(src, (lds_void_t*)(srec3 + c * 64), 4,
                                                         e * 4u, 0, 0, kStreamPolicy);
        }
        dma_fence();
    } else {
        if (valid) {
            const float* src = S.slab + i * int64_t(RW);
            for (int t = ln.q; t < RW; t += kQ9Lanes) rp[t] = src[t];
        }
        __syncthreads();
    }
    if (!valid) return;
    const uint64_t rows = uint64_t(G.N) * 36u;
    const __amdgpu_buffer_rsrc_t gf = rsrc(S.Gf, rows);
    const uint32_t lane_off = uint32_t(ln.q) * 12u;
    const f3v acc =
        stencil_sum3<8>(rp, K, ln.q, [&](uint32_t j) { return ld3(gf, j * 36u + lane_off); });
    const __amdgpu_buffer_rsrc_t ra = rsrc(S.A, rows);
    const uint32_t own = uint32_t(i) * 36u + uint32_t(ln.q) * 12u;
    f3v a = ld3(ra, own);
    a.x += acc.x;
    a.y += acc.y;
    a.z += acc.z;
    st3(ra, own, a);
}

// Boundary interpolation for D2Q9, three per lane (the arithmetic of
// k_binterp): accumulate = true (p3) interpolates s1 minus the sources'
// collision increments onto the row's own increment, false (p5) overwrites.
template <bool ACCUMULATE>
__global__ __launch_bounds__(kBlock) void k_binterp3(StepGeom G, StepBuffers S, BatchArgs B) {
    const BlockCtx bc = locate(B, G, S);
    const Q9Lane ln;
    const int64_t nbnd = G.N - G.nfluid;
    const int64_t m = xcd_tile(bc.b, bc.nb) * kQ9Npb + ln.slot;
    if (!ln.active || m >= nbnd) return;
    float* field = ACCUMULATE ? S.A : S.g;
    const uint64_t rows = uint64_t(G.N) * 36u;
    const __amdgpu_buffer_rsrc_t f = rsrc(field, rows), cl = rsrc(S.coll, rows);
    const float* rec = S.bslab + m * int64_t(G.kb) * 2;
    const uint32_t lane_off = uint32_t(ln.q) * 12u;
    const uint32_t own = uint32_t(G.nfluid + m) * 36u + lane_off;
    const f3v b = ACCUMULATE ? ld3(f, own) : f3v{0.0f, 0.0f, 0.0f};
    f3v acc = {0.0f, 0.0f, 0.0f};
    for (int s0 = 0; s0 < G.kb; s0 += 8) {
        float w[8];
        f3v v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int s = s0 + u;
            const bool in = s < G.kb;
            const uint32_t j = __float_as_uint(rec[2 * (in ? s : 0)]);
            w[u] = in ? rec[2 * s + 1] : 0.0f;
            v[u] = ld3(f, j * 36u + lane_off);
            if (ACCUMULATE) {
                const f3v c = ld3(cl, j * 36u + lane_off);
                v[u].x = v[u].x - c.x;
                v[u].y = v[u].y - c.y;
                v[u].z = v[u].z - c.z;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            acc.x += w[u] * v[u].x;
            acc.y += w[u] * v[u].y;
            acc.z += w[u] * v[u].z;
        }
    }
    f3v out = acc;
    if (ACCUMULATE) {
        out.x = b.x + acc.x;
        out.y = b.y + acc.y;
        out.z = b.z + acc.z;
    }
    st3(f, own, out);
}

// p4 for D2Q9, three per lane: selective filter on fluid nodes, bounce-back
// on boundary nodes (the arithmetic of k_p4).
template <bool TU = false>
__global__ __launch_bounds__(kBlock) void k_p4_3(StepGeom G, StepBuffers S, BatchArgs B) {
    extern __shared__ float srec3[];
    const BlockCtx bc = locate(B, G, S);
    const int K = G.k, RW = 2 * K;
    const int64_t tile = xcd_tile(bc.b, bc.nb);
    const int64_t node0 = tile * kQ9Npb;
    const Q9Lane ln;
    const int64_t i = node0 + ln.slot;
    const uint64_t rows = uint64_t(G.N) * 36u;
    const uint32_t own = uint32_t(i) * 36u + uint32_t(ln.q) * 12u;
    const __amdgpu_buffer_rsrc_t r1 = rsrc(S.A, rows);
    const f3v si = ld3(r1, own);
    const int64_t nf = G.nfluid - node0;
    const int cnt = nf <= 0 ? 0 : (nf < kQ9Npb ? int(nf) : kQ9Npb);
    int64_t u0 = 0;
    int U = 0;
    if (TU && cnt > 0) union_range(S.uoff, tile, u0, U);
    const int Uc = U < S.ucap ? U : S.ucap;
    float* urows = srec3 + kQ9Npb * (RW + 1);
    Q9UnionIdx ux;
    if (TU) ux = q9_union_indices(S.ulist + u0, Uc);
    stage_padded(TU ? S.ufslab : S.fslab, node0, cnt, RW, srec3);
    if (TU) q9_union_rows(urows, ux, Uc, r1);
    dma_fence();
    if (!ln.active || i >= G.N) return;
    f3v out;
    if (i < G.nfluid) {
        const float* rp = srec3 + ln.slot * (RW + 1);
        const uint32_t lane_off = uint32_t(ln.q) * 12u;
        const uint32_t* ul = S.ulist + u0;
        const bool lds = TU && U <= S.ucap;
        f3v sum = {0.0f, 0.0f, 0.0f};
        for (int s0 = 0; s0 < K; s0 += 8) {
            float w[8];
            f3v gv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int s = s0 + u;
                const bool in = s < K;
                const uint32_t j = __float_as_uint(rp[(in ? s : 0) * 2]);
                w[u] = in ? rp[s * 2 + 1] : 0.0f;
                if (!TU) gv[u] = ld3(r1, j * 36u + lane_off);
                else if (lds) gv[u] = q9_lds_row(urows, j, ln.q);
                else gv[u] = ld3(r1, ul[j] * 36u + lane_off);  // over the cap: through the list
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                sum.x += w[u] * (gv[u].x - si.x);
                sum.y += w[u] * (gv[u].y - si.y);
                sum.z += w[u] * (gv[u].z - si.z);
            }
        }
        out.x = si.x + G.filter * sum.x;
        out.y = si.y + G.filter * sum.y;
        out.z = si.z + G.filter * sum.z;
    } else {
        const int64_t b = i - G.nfluid;
        const float n0 = S.normals[b * 2], n1 = S.normals[b * 2 + 1];
        const float* row = S.A + i * 9;
        float o3[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int a = 3 * ln.q + e;
            float cn = 0.0f;
            cn += c9(a, 0) * n0;
            cn += c9(a, 1) * n1;
            const int src = cn > 0.0f ? Lattice<9>::bb(a) : a;
            o3[e] = row[src];
        }
        out.x = o3[0];
        out.y = o3[1];
        out.z = o3[2];
    }
    st3(rsrc(S.g, rows), own, out);
}

// sum_s w_s (g_s - g_i) over the staged filter records rp ({index field, w}),
// rows fetched by fetch(index field); four slots per lane.
template <int K, int BATCH, class Fetch>
__device__ __forceinline__ float4 filter_sum(const float2* rp, const float4& si, Fetch fetch) {
    float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s0 = 0; s0 < K; s0 += BATCH) {
        float2 rec[BATCH];
        float4 gv[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) rec[u] = rp[s0 + u];
#pragma unroll
        for (int u = 0; u < BATCH; ++u) gv[u] = fetch(__float_as_uint(rec[u].x));
#pragma unroll
        for (int u = 0; u < BATCH; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) sum[e] += rec[u].y * (get(gv[u], e) - get(si, e));
    }
    return make_float4(sum[0], sum[1], sum[2], sum[3]);
}

// filter_sum over packed filter records (packed_filter_units layout): the
// same products in the same order, so the two agree bit for bit.
#ifndef NAGA_FILTER_SUB
#define NAGA_FILTER_SUB 4
#endif
template <int K, class Fetch>
__device__ __forceinline__ float4 filter_sum_packed(const float4* rp, const float4& si,
                                                    Fetch fetch, int ng = K / 8) {
    static_assert(K % 8 == 0, "packed records come in groups of 8");
    constexpr int H = NAGA_FILTER_SUB;  // records gathered per sub-batch: 8 or 4
    static_assert(H == 8 || H == 4, "sub-batch of 8 or 4 records");
    float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int g = 0; g < K / 8; ++g) {
        if (g >= ng) break;  // ragged rows: the tile's ng groups
        const float4 iu = rp[3 * g + 2];
#pragma unroll
        for (int h = 0; h < 8; h += H) {  // records h .. h + H - 1, weights in units h/4 ..
            float4 gv[H];
#pragma unroll
            for (int r = 0; r < H; ++r) {
                const uint32_t word = __float_as_uint(get(iu, (h + r) >> 1));
                gv[r] = fetch(((h + r) & 1) ? word >> 16 : word & 0xffffu);
            }
            float4 wu[H / 4];
#pragma unroll
            for (int w = 0; w < H / 4; ++w) wu[w] = rp[3 * g + h / 4 + w];
#pragma unroll
            for (int r = 0; r < H; ++r)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    sum[e] += get(wu[r >> 2], r & 3) * (get(gv[r], e) - get(si, e));
        }
    }
    return make_float4(sum[0], sum[1], sum[2], sum[3]);
}

// p4 for D3Q27 with the tile unions: selective filter on fluid nodes,
// bounce-back on boundary nodes.
template <int K, int BATCH, bool TU, int NPB, bool RG = false>
__global__ __launch_bounds__(NPB * kV4Lanes) void k_p4_u(StepGeom G, StepBuffers S, BatchArgs B) {
    constexpr int NT = NPB * kV4Lanes;
    // record rows with distinct banks per node: K + 1 records through
    // VGPRs, or packed rows of KF units by LDS DMA
    constexpr int KP = K + 1, KF = packed_filter_units(K);
    constexpr int SU = TU ? NPB * KF : (NPB * KP + 1) / 2;  // 16-byte units
    __shared__ float4 srec4[SU];
    float2* srec = reinterpret_cast<float2*>(srec4);
    extern __shared__ float4 urows[];
    const BlockCtx bc = locate(B, G, S);
    const UnionWork w = union_work<NPB, TU>(G, S, union_entry<TU>(S, bc.b, bc.nb));
    const int64_t node0 = w.tile * NPB;
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t i = node0 + slot;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
    const __amdgpu_buffer_rsrc_t s1 = rsrc(S.A, rows);
    const float4 si = ld4(s1, own);
    const int U = w.U;
    const uint32_t* ul = TU ? S.ulist + w.u0 : nullptr;  // TU: tile unions built
    const TileRecs tr =
        tile_recs<RG, K, KF, NPB>(S.tile_ng, S.frec_off, w.tile, w.hi > w.lo, true);
    const int kf = tr.kp;
    if (w.hi > w.lo) {
        if (!TU) {
            const float2* src = reinterpret_cast<const float2*>(S.fslab) + node0 * K;
            for (int t = threadIdx.x; t < w.hi * K; t += NT) srec[t + t / K] = src[t];
        } else {
            const int Uc = U < S.ucap_f ? U : S.ucap_f;
            const UnionIdx<NT> ux = union_indices<NT>(ul, Uc);
            __builtin_amdgcn_sched_barrier(0);
            const int n = w.hi - w.lo;
            dma_contiguous<NPB * KF, NT>(
                rsrc(S.fslab + (tr.base + int64_t(w.lo) * kf) * 4, uint64_t(n) * kf * 16u),
                srec4 + w.lo * kf, n * kf);
            __builtin_amdgcn_sched_barrier(0);
            union_copy<NT>(urows, ux, Uc, s1);
        }
    }
    if (TU) dma_fence();
    else __syncthreads();
    if (i >= G.N) return;
    const bool fluid = i < G.nfluid;
    // fluid nodes: those this block serves; other nodes: the tile's block
    if (fluid ? (slot < w.lo || slot >= w.hi) : !w.first) return;
    float4 out;
    if (fluid) {
        const uint32_t lane_off = uint32_t(q) * 16u;
        float4 sum;
        if (!ul)
            sum = filter_sum<K, BATCH>(srec + slot * KP, si,
                                       [&](uint32_t j) { return ld4(s1, j * 128u + lane_off); });
        else if (U <= S.ucap_f)
            sum = filter_sum_packed<K>(srec4 + slot * kf, si,
                                       [&](uint32_t j) { return urows[j * kV4Lanes + q]; }, tr.ng);
        else  // union over the cap: every row read from HBM through the list
            sum = filter_sum_packed<K>(srec4 + slot * kf, si, [&](uint32_t j) {
                return ld4(s1, ul[j] * 128u + lane_off);
            }, tr.ng);
#pragma unroll
        for (int e = 0; e < 4; ++e) set(out, e, get(si, e) + G.filter * get(sum, e));
    } else {
        const int64_t b = i - G.nfluid;
        const float n0 = S.normals[b * 3], n1 = S.normals[b * 3 + 1], n2 = S.normals[b * 3 + 2];
        const float* row = S.A + i * 32;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int a = 4 * q + e;
            const float cn = c27(a, 0) * n0 + c27(a, 1) * n1 + c27(a, 2) * n2;
            const int src = (a < 27 && cn > 0.0f) ? Lattice<27>::bb(a) : a;
            set(out, e, src == a ? get(si, e) : row[src]);
        }
    }
    st4(S.g, own, out);
}

// p2 for D3Q27 with four distributions per lane (PML divergence over the
// listed nodes, added to s1), records holding node indices.
template <int K, int BATCH, int NPB>
__global__ __launch_bounds__(NPB * kV4Lanes) void k_p2_v4(StepGeom G, StepBuffers S, BatchArgs B) {
    constexpr int KP = K + 1;  // padded record rows (distinct banks per node)
    __shared__ float4 srec[NPB * KP];
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const BlockCtx bc = locate(B, G, S);
    const int64_t item = xcd_tile(bc.b, bc.nb) * NPB + slot;
    const bool valid = item < S.n_pml;
    const int64_t i = valid ? S.pml_list[item] : 0;
    float4* rp = srec + slot * KP;
    if (valid) {
        const float4* src = reinterpret_cast<const float4*>(S.slab) + i * K;
        for (int r = q; r < K; r += kV4Lanes) rp[r] = src[r];
    }
    __syncthreads();
    if (!valid) return;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const __amdgpu_buffer_rsrc_t gf = rsrc(S.Gf, rows);
    const uint32_t lane_off = uint32_t(q) * 16u;
    const float4 r =
        rate_fetch<K, BATCH>(rp, [&](uint32_t j) { return ld4(gf, j * 128u + lane_off); }, q);
    // the divergence of F = c Gf is the negated rate of Gf
    const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
    const float4 s = ld4(rsrc(S.A, rows), own);
    st4(S.A, own, make_float4(s.x - r.x, s.y - r.y, s.z - r.z, s.w - r.w));
}

// p2 for D3Q27 with tile unions: the PML divergence over every fluid node
// of each listed tile (tiles holding a node with a layer neighbour; for the
// other nodes of such a tile every neighbour's auxiliary row is zero and
// the added divergence is an exact zero), the tile's union rows of Gf and
// its records staged by LDS DMA.
template <int K, int BATCH, int NPB, bool RG = false>
__global__ __launch_bounds__(NPB * kV4Lanes) void k_p2_t(StepGeom G, StepBuffers S, BatchArgs B) {
    constexpr int KP = packed_div_units(K);  // packed record rows
    constexpr int NT = NPB * kV4Lanes;
    __shared__ float4 srec[NPB * KP];
    extern __shared__ float4 urows[];
    const BlockCtx bc = locate(B, G, S);
    // this pass's unions: the layer neighbours and one zero row
    S.slab = S.p2slab;
    S.ulist = S.p2ulist;
    S.uoff = S.p2uoff;
    S.usplit = S.p2usplit;
    S.split_tiles = S.p2split_tiles;
    S.n_split = S.p2n_split;
    typedef const __attribute__((address_space(4))) int32_t const_i32;
    const UnionWork w =
        union_work<NPB, true>(G, S, ((const_i32*)S.ptiles)[xcd_tile(bc.b, bc.nb)]);
    const int64_t node0 = w.tile * NPB;
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t i = node0 + slot;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
    const __amdgpu_buffer_rsrc_t gf = rsrc(S.Gf, rows);
    const __amdgpu_buffer_rsrc_t ar = rsrc(S.A, rows);
    const int U = w.U;
    const uint32_t* ul = S.ulist + w.u0;
    const int Uc = U < S.ucap ? U : S.ucap;
    const UnionIdx<NT> ux = union_indices<NT>(ul, Uc);
    __builtin_amdgcn_sched_barrier(0);
    const int n = w.hi - w.lo;
    const TileRecs tr =
        tile_recs<RG, K, KP, NPB>(S.tile_ng, S.rec_off, w.tile, w.hi > w.lo, false);
    const int kp = tr.kp;
    dma_contiguous<NPB * KP, NT>(rsrc(S.slab + (tr.base + int64_t(w.lo) * kp) * 4, uint64_t(n) * kp * 16u),
                                 srec + w.lo * kp, n * kp);
    __builtin_amdgcn_sched_barrier(0);
    union_copy<NT>(urows, ux, Uc, gf);
    const float4 s = ld4(ar, own);
    dma_fence();
    if (slot < w.lo || slot >= w.hi) return;
    const float4* rp = srec + slot * kp;
    const uint32_t lane_off = uint32_t(q) * 16u;
    float4 r;
    if (U <= S.ucap)
        r = rate_fetch_packed<K>(rp, [&](uint32_t j) { return urows[j * kV4Lanes + q]; }, q,
                                 tr.ng);
    else  // union over the cap: every row read from HBM through the list
        r = rate_fetch_packed<K>(rp, [&](uint32_t j) { return ld4(gf, ul[j] * 128u + lane_off); },
                                 q, tr.ng);
    // the divergence of F = c Gf is the negated rate of Gf
    st4(S.A, own, make_float4(s.x - r.x, s.y - r.y, s.z - r.z, s.w - r.w));
}

// PML auxiliary field rows for D3Q27, four distributions per lane (the
// arithmetic of k_pml_field; padding slots stay zero).
__global__ __launch_bounds__(kBlock) void k_pml_field_v4(StepGeom G, StepBuffers S, BatchArgs B) {
    const BlockCtx bc = locate(B, G, S);
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t m = xcd_tile(bc.b, bc.nb) * kV4Nodes + slot;
    if (m >= G.nlayer) return;
    const int64_t j = G.nbulk + m;
    const int64_t p = G.pml_ref ? j - G.nbulk : j;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const uint32_t lane_off = uint32_t(q) * 16u;
    const float sig = S.sigma[m];
    const float4 q1 = ld4(rsrc(S.Qp, rows), uint32_t(p) * 128u + lane_off);
    const float4 q2 = ld4(rsrc(S.Q2v, rows), uint32_t(p) * 128u + lane_off);
    const float4 m1 = reinterpret_cast<const float4*>(S.means)[q];
    const float4 m2 = reinterpret_cast<const float4*>(S.means + means_stride(27))[q];
    float4 out;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = 4 * q + e;
        const float v = a < 27 ? -2.0f * sig * (get(q1, e) - get(m1, e)) +
                                     sig * sig * (get(q2, e) - get(m2, e))
                               : 0.0f;
        set(out, e, v);
    }
    st4(S.Gf, uint32_t(j) * 128u + lane_off, out);
}

// PML local terms for D3Q27, four distributions per lane (same arithmetic
// as k_pml_local): 32 node slots per round, kPmlRounds rounds per block,
// one row of partial layer sums per block.
__global__ __launch_bounds__(kBlock) void k_pml_local_v4(StepGeom G, StepBuffers S, BatchArgs B) {
    __shared__ double red[kV4Nodes][2][32];
    const BlockCtx bc = locate(B, G, S);
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const double hdt = double(G.dt) / 2.0;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const __amdgpu_buffer_rsrc_t rQp = rsrc(S.Qp, rows), rA = rsrc(S.A, rows);
    double s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t base = bc.b * (kV4Nodes * kPmlRounds);
    for (int rd = 0; rd < kPmlRounds; ++rd) {
        const int64_t item = base + rd * kV4Nodes + slot;
        if (item >= S.n_rq) continue;
        const int64_t i = S.rq_list[item];
        const float4 ru = S.ru[i];
        float4* prev = reinterpret_cast<float4*>(S.E) + i;
        const float4 pr = *prev;  // the previous step's (rho - 1, u)
        const uint32_t own = uint32_t(i) * 128u + uint32_t(q) * 16u;
        const float4 p4 = ld4(rQp, own);
        const bool layer = i >= G.nbulk && i < G.nfluid;
        const double sig = layer ? double(S.sigma[i - G.nbulk]) : 0.0;
        float4 a4 = layer ? ld4(rA, own) : make_float4(0.f, 0.f, 0.f, 0.f);
        float4 po, qo;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int a = 4 * q + e;
            if (a >= 27) {  // padding slots stay zero
                set(po, e, 0.0f);
                set(qo, e, 0.0f);
                continue;
            }
            // feq - w now and at the previous step's macroscopic values
            const double fte = feq_dev<27>(ru, a);
            const double q1 = (fte + feq_dev<27>(pr, a)) * hdt;
            const double q2 = (q1 + double(get(p4, e))) * hdt;
            set(po, e, float(q1));
            set(qo, e, float(q2));
            if (layer) {
                set(a4, e, float(double(get(a4, e)) - sig * (3.0 * fte + 3.0 * sig * q1 + sig * sig * q2)));
                s1[e] += q1;
                s2[e] += q2;
            }
        }
        // the node's eight lanes (one wavefront) have read prev: the store
        // waits for the products of the loaded value
        float4 keep = ru;
        asm volatile("" : "+v"(keep.x) : "v"(get(po, 0)));
        if (q == 0) *prev = keep;
        st4(S.Qp, own, po);
        if (G.pml_ref ? i < G.nlayer : layer) st4(S.Q2v, own, qo);  // rows the layer field reads
        if (layer) st4(S.A, own, a4);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        red[slot][0][4 * q + e] = s1[e];
        red[slot][1][4 * q + e] = s2[e];
    }
    __syncthreads();
    if (threadIdx.x < 2 * 27) {
        const int which = threadIdx.x / 27, aa = threadIdx.x % 27;
        double t = 0.0;
        for (int sl = 0; sl < kV4Nodes; ++sl) t += red[sl][which][aa];
        S.partial[int64_t(threadIdx.x) * S.nparts + bc.b] = t;
    }
}

// Boundary interpolation for D3Q27, four distributions per lane (same
// arithmetic as k_binterp), records read directly by the node's lanes.
template <bool ACCUMULATE>
__global__ __launch_bounds__(kBlock) void k_binterp_v4(StepGeom G, StepBuffers S, BatchArgs B) {
    const BlockCtx bc = locate(B, G, S);
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t nbnd = G.N - G.nfluid;
    const int64_t m = xcd_tile(bc.b, bc.nb) * kV4Nodes + slot;
    if (m >= nbnd) return;
    float* field = ACCUMULATE ? S.A : S.g;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const __amdgpu_buffer_rsrc_t f = rsrc(field, rows), cl = rsrc(S.coll, rows);
    const float2* rp = reinterpret_cast<const float2*>(S.bslab) + m * G.kb;
    const uint32_t lane_off = uint32_t(q) * 16u;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    constexpr int BT = 8;
    for (int s0 = 0; s0 < G.kb; s0 += BT) {
        float2 rec[BT];
        float4 v[BT];
#pragma unroll
        for (int u = 0; u < BT; ++u)
            rec[u] = s0 + u < G.kb ? rp[s0 + u] : make_float2(__uint_as_float(0u), 0.0f);
#pragma unroll
        for (int u = 0; u < BT; ++u) {
            const uint32_t off = __float_as_uint(rec[u].x) * 128u + lane_off;
            v[u] = ld4(f, off);
            if (ACCUMULATE) {
                const float4 c = ld4(cl, off);
                v[u] = make_float4(v[u].x - c.x, v[u].y - c.y, v[u].z - c.z, v[u].w - c.w);
            }
        }
#pragma unroll
        for (int u = 0; u < BT; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] += rec[u].y * get(v[u], e);
    }
    const uint32_t own = uint32_t(G.nfluid + m) * 128u + lane_off;
    float4 out = make_float4(acc[0], acc[1], acc[2], acc[3]);
    if (ACCUMULATE) {
        const float4 b = ld4(f, own);
        out = make_float4(b.x + out.x, b.y + out.y, b.z + out.z, b.w + out.w);
    }
    st4(field, own, out);
}

// Boundary interpolation for D3Q27 with the sources of each 32-node tile
// staged in LDS (the field rows, and for the first pass the collision
// increments too) by LDS DMA: about 6 rows per node instead of 32 (64)
// gathers.  Same records, products and order as k_binterp_v4.
static_assert(kBTileNodes == kV4Nodes, "one boundary tile per block");
// LDS of k_binterp_u: the tile's records (kb <= 63: 32 x 63 {position, w})
// after the field rows and, for the first pass, the increments.
constexpr int kBRecUnits = (kBTileNodes * 63 * 8 + 15) / 16;
template <bool ACCUMULATE>
constexpr size_t binterp_lds() {
    return size_t((ACCUMULATE ? 2 : 1) * kBUnionCap * kV4Lanes + kBRecUnits) * 16u;
}

template <bool ACCUMULATE>
__global__ __launch_bounds__(kBlock) void k_binterp_u(StepGeom G, StepBuffers S, BatchArgs B) {
    extern __shared__ float4 urows[];  // field rows, (ACCUMULATE) increments, records
    float4* recs = urows + (ACCUMULATE ? 2 : 1) * kBUnionCap * kV4Lanes;
    const BlockCtx bc = locate(B, G, S);
    const int slot = threadIdx.x / kV4Lanes, q = threadIdx.x % kV4Lanes;
    const int64_t nbnd = G.N - G.nfluid;
    const int64_t tile = xcd_tile(bc.b, bc.nb);
    const int64_t m0 = tile * kV4Nodes, m = m0 + slot;
    float* field = ACCUMULATE ? S.A : S.g;
    const uint64_t rows = uint64_t(G.N) * 128u;
    const __amdgpu_buffer_rsrc_t f = rsrc(field, rows), cl = rsrc(S.coll, rows);
    int64_t u0;
    int U;
    union_range(S.buoff, tile, u0, U);
    const uint32_t* ul = S.bulist + u0;
    const bool lds = U <= kBUnionCap;
    const UnionIdx<kBlock> ux = union_indices<kBlock>(ul, lds ? U : 0);
    __builtin_amdgcn_sched_barrier(0);
    // the tile's records: cnt * kb pairs from the tile's start (the last
    // tile's piece past the slab reads as zero)
    const int64_t left = nbnd - m0;
    const int cnt = left < kV4Nodes ? int(left) : kV4Nodes;
    dma_contiguous<kBRecUnits, kBlock>(
        rsrc(S.bslab + m0 * G.kb * 2, uint64_t(nbnd - m0) * G.kb * 8u), recs,
        (cnt * G.kb * 8 + 15) / 16);
    __builtin_amdgcn_sched_barrier(0);
    if (lds) {
        union_copy<kBlock>(urows, ux, U, f);
        if (ACCUMULATE) union_copy<kBlock>(urows + kBUnionCap * kV4Lanes, ux, U, cl);
    }
    const uint32_t own = uint32_t(G.nfluid + m) * 128u + uint32_t(q) * 16u;
    // the node's own row (the collision increment it accumulates onto),
    // loaded with the copies in flight (past the rows: zero, never stored)
    const float4 b = ACCUMULATE ? ld4(f, own) : make_float4(0.f, 0.f, 0.f, 0.f);
    dma_fence();
    if (m >= nbnd) return;
    const float2* rp = reinterpret_cast<const float2*>(recs) + slot * G.kb;
    const uint32_t lane_off = uint32_t(q) * 16u;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    constexpr int BT = 8;
    for (int s0 = 0; s0 < G.kb; s0 += BT) {
        float2 rec[BT];
        float4 v[BT];
#pragma unroll
        for (int u = 0; u < BT; ++u)
            rec[u] = s0 + u < G.kb ? rp[s0 + u] : make_float2(__uint_as_float(0u), 0.0f);
#pragma unroll
        for (int u = 0; u < BT; ++u) {
            const uint32_t j = __float_as_uint(rec[u].x);
            if (lds) {
                v[u] = urows[j * kV4Lanes + q];
                if (ACCUMULATE) {
                    const float4 c = urows[(kBUnionCap + j) * kV4Lanes + q];
                    v[u] = make_float4(v[u].x - c.x, v[u].y - c.y, v[u].z - c.z, v[u].w - c.w);
                }
            } else {  // sources over the cap: rows read from HBM through the list
                const uint32_t off = ul[j] * 128u + lane_off;
                v[u] = ld4(f, off);
                if (ACCUMULATE) {
                    const float4 c = ld4(cl, off);
                    v[u] = make_float4(v[u].x - c.x, v[u].y - c.y, v[u].z - c.z, v[u].w - c.w);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < BT; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] += rec[u].y * get(v[u], e);
    }
    float4 out = make_float4(acc[0], acc[1], acc[2], acc[3]);
    if (ACCUMULATE) out = make_float4(b.x + out.x, b.y + out.y, b.z + out.z, b.w + out.w);
    st4(field, own, out);
}

__global__ __launch_bounds__(kBlock) void k_source_field(const int32_t* nodes,
                                                       const int32_t* off, const int32_t* ids,
                                                       const float* prof, int64_t n,
                                                       SourceAmps amps, float* field) {
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int e = off[i]; e < off[i + 1]; ++e) s += prof[e] * amps.a[ids[e]];
    field[nodes[i]] = s;
}

template <int Q>
__global__ __launch_bounds__(kBlock) void k_probe(int64_t n, int k, const uint32_t* idx,
                                                  const float* w, const float* held,
                                                  const float* g, double rho0, double* out) {
    using L = Lattice<Q>;
    const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (p >= n) return;
    double acc = 0.0;
    for (int s = 0; s < k; ++s) {
        const int64_t j = idx[p * k + s];
        double d;
        if (held) {
            d = double(held[j]) - rho0;
        } else {
            double t = 0.0;
#pragma unroll
            for (int a = 0; a < Q; ++a) t += double(g[j * L::RS + a]);
            d = rho0 * t;
        }
        acc += double(w[p * k + s]) * d;
    }
    out[p] = rho0 + acc;
}

template <int Q>
__global__ __launch_bounds__(kBlock) void k_macro(StepGeom G, const float* g,
                                                  float* rho, double* drho,
                                                  float* u) {
    using L = Lattice<Q>;
    constexpr int D = L::D, RS = L::RS;
    const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= G.N) return;
    double rt = 0.0, mom[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const double v = double(g[i * RS + a]);
        rt += v;
#pragma unroll
        for (int d = 0; d < D; ++d) mom[d] += v * double(L::c(a, d));
    }
    const double r = 1.0 + rt;
    if (rho) rho[i] = float(r * double(G.rho0));
    if (drho) drho[i] = rt * double(G.rho0);
    if (u)
#pragma unroll
        for (int d = 0; d < D; ++d) u[d * G.N + i] = float(mom[d] / r * double(G.vel_scale));
}

template <int Q>
int64_t lane_blocks(int64_t nodes) {
    constexpr int npb = kBlock / 64 * Lattice<Q>::NPW;
    return (nodes + npb - 1) / npb;
}

inline void check_launch(const char* what) {
    hip_check(hipGetLastError(), what);
}

}  // namespace

int stage_batch() {
    static const int b = [] {
        const char* e = std::getenv("NAGA_STAGE_BATCH");
        const int v = e ? std::atoi(e) : 16;
        return (v == 8 || v == 16 || v == 32) ? v : 16;
    }();
    return b;
}

// The four-per-lane D3Q27 kernels need k in {16, 32, 48} and 32-bit row offsets.
// NAGA_V4=0 selects the one-per-lane kernels.
bool step_uses_unions(int Q, int k, int64_t N) {
    static const bool on = [] {
        const char* e = std::getenv("NAGA_V4");
        return !(e && e[0] == '0');
    }();
    return Q == 27 && on && (k == 48 || k == 32 || k == 16) && N < (int64_t(1) << 25);
}

// D2Q9 scenes run the three-per-lane stencil kernels (12-byte pieces with
// 32-bit buffer offsets: fewer than 2^26 nodes) unless NAGA_Q9_THREE=0.
bool q9_three_fits(int64_t N) { return N < (int64_t(1) << 26); }
bool q9_three(int Q, int64_t N) {
    static const bool on = [] {
        const char* e = std::getenv("NAGA_Q9_THREE");
        return !(e && e[0] == '0');
    }();
    return on && Q == 9 && q9_three_fits(N);
}

// Nodes per block of the D3Q27 stencil kernels: the union tile size with
// tile unions, else 32, 64 or 128 (records of all of them staged in at
// most 64 KB of LDS; NAGA_V4_NPB overrides the default 32).
int v4_nodes_per_block(int k, bool tiles) {
    static const int env = [] {
        const char* e = std::getenv("NAGA_V4_NPB");
        const int v = e ? std::atoi(e) : kV4DefaultNodes;
        return (v == 32 || v == 64 || v == 128) ? v : kV4DefaultNodes;
    }();
    if (tiles) return union_tile_nodes(k);
    int v = env;
    while (v > kV4Nodes && v * (k + 1) * 16 > 65536) v /= 2;
    return v;
}

int union_tile_nodes(int k) {
    static const bool want32 = [] {
        const char* e = std::getenv("NAGA_TILE");
        return e && std::atoi(e) == 32;
    }();
    return (k <= 32 && !want32) ? 2 * kV4Nodes : kV4Nodes;
}

int union_cap_default(int k) {
    if (union_tile_nodes(k) != kV4Nodes) return 400;  // 64-node tiles, two blocks per CU
    return k <= 32 ? 296 : 440;  // 32-node tiles: three blocks per CU, two for k = 48
}

namespace {

// Passes of one step, in launch order.
enum Pass { kP1, kPl, kPm, kPf, kP2, kP3, kP4, kS1, kS2, kS3, kS4, kP5 };

// How a pass is launched: kernel template choices shared by every scene of
// the launch, and the batch table (or none).
struct PassCfg {
    int ktmpl;     // K template: 32, 16 or 0 (runtime k)
    bool unions;   // D3Q27 four-per-lane kernels with tile unions
    int ucap;      // LDS union rows (dynamic shared memory / 128 B)
    int ucap_f;    // the same for the filter pass
    BatchArgs B;
    int gbatch;    // gather batch of the one-per-lane stage kernels (0: default)
    bool tiles;    // tile unions built (slab index fields are union positions)
    int kmax;      // largest stencil width of the launch's scenes (LDS of the D2Q9 kernels)
    bool ragged;   // per-tile record widths (StepBuffers::tile_ng)
    bool three;    // D2Q9 stencil passes with three distributions per lane
};

template <int Q, int STAGE, int KT>
void launch_stage_scalar(int64_t nb, const StepGeom& G, const StepBuffers& S, BatchArgs B,
                         int gbatch, hipStream_t st) {
    const int b = gbatch > 0 ? gbatch : (KT ? stage_batch() : 8);
    if (b == 32)
        k_stage<Q, STAGE, KT, 32><<<nb, kBlock, 0, st>>>(G, S, B);
    else if (b == 16)
        k_stage<Q, STAGE, KT, 16><<<nb, kBlock, 0, st>>>(G, S, B);
    else
        k_stage<Q, STAGE, KT, 8><<<nb, kBlock, 0, st>>>(G, S, B);
}

// The D3Q27 four-per-lane kernels of the stencil passes for one stencil
// width K (16, 32 or 48) and NPB nodes per block.
template <int K, int NPB>
void launch_v4_npb(int pass, int64_t nb, const StepGeom& G, const StepBuffers& S,
                   const PassCfg& c, hipStream_t st) {
    constexpr int threads = NPB * kV4Lanes;
    const size_t lds = size_t(pass == kP4 ? c.ucap_f : c.ucap) * 128u;
    if (c.tiles && c.ragged) {
        switch (pass) {
            case kP2: k_p2_t<K, 8, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kP4: k_p4_u<K, 8, true, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS1: k_stage_u<1, K, 8, true, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS2: k_stage_u<2, K, 8, true, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS3: k_stage_u<3, K, 8, true, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS4: k_stage_u<4, K, 8, true, NPB, true><<<nb, threads, lds, st>>>(G, S, c.B); return;
            default: fail(kInvalidArgument, "not a stencil pass");
        }
    }
    if (c.tiles) {
        switch (pass) {
            case kP2: k_p2_t<K, 8, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kP4: k_p4_u<K, 8, true, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS1: k_stage_u<1, K, 8, true, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS2: k_stage_u<2, K, 8, true, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS3: k_stage_u<3, K, 8, true, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            case kS4: k_stage_u<4, K, 8, true, NPB><<<nb, threads, lds, st>>>(G, S, c.B); return;
            default: fail(kInvalidArgument, "not a stencil pass");
        }
    }
    switch (pass) {
        case kP2: k_p2_v4<K, 8, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        case kP4: k_p4_u<K, 8, false, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        case kS1: k_stage_u<1, K, 8, false, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        case kS2: k_stage_u<2, K, 8, false, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        case kS3: k_stage_u<3, K, 8, false, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        case kS4: k_stage_u<4, K, 8, false, NPB><<<nb, threads, 0, st>>>(G, S, c.B); return;
        default: fail(kInvalidArgument, "not a stencil pass");
    }
}

template <int K>
void launch_v4(int pass, int64_t nb, const StepGeom& G, const StepBuffers& S, const PassCfg& c,
               hipStream_t st) {
    const int npb = v4_nodes_per_block(K, c.tiles);
    if constexpr ((K + 1) * 128 * 16 <= 65536)
        if (npb == 128) return launch_v4_npb<K, 128>(pass, nb, G, S, c, st);
    if constexpr ((K + 1) * 64 * 16 <= 65536)
        if (npb == 64) return launch_v4_npb<K, 64>(pass, nb, G, S, c, st);
    launch_v4_npb<K, kV4Nodes>(pass, nb, G, S, c, st);
}

template <int Q, int STAGE>
void launch_stage_pass(int64_t nb, const StepGeom& G, const StepBuffers& S, const PassCfg& c,
                       hipStream_t st) {
    if (c.ktmpl == 32) {
        launch_stage_scalar<Q, STAGE, 32>(nb, G, S, c.B, c.gbatch, st);
    } else if (c.ktmpl == 16) {
        launch_stage_scalar<Q, STAGE, 16>(nb, G, S, c.B, c.gbatch, st);
    } else {
        launch_stage_scalar<Q, STAGE, 0>(nb, G, S, c.B, c.gbatch, st);
    }
}

template <int Q>
void launch_pass(int pass, int64_t nb, const StepGeom& G, const StepBuffers& S,
                 const CollisionParams& cp, const PassCfg& c, hipStream_t st) {
    if (nb <= 0) return;
    const bool v4 = Q == 27 && c.unions;
    if (v4 && (pass == kP2 || pass == kP4 || (pass >= kS1 && pass <= kS4))) {
        if (c.ktmpl == 48) launch_v4<48>(pass, nb, G, S, c, st);
        else if (c.ktmpl == 32) launch_v4<32>(pass, nb, G, S, c, st);
        else launch_v4<16>(pass, nb, G, S, c, st);
        check_launch("step pass");
        return;
    }
    if (Q == 9 && c.three &&
        (pass == kP2 || pass == kP3 || pass == kP4 || pass == kP5 ||
         (pass >= kS1 && pass <= kS4))) {
        // records of the block's nodes (one pad float per node row), then
        // with tile unions the union rows (36 bytes each)
        const size_t recs = size_t(kQ9Npb) * ((pass == kP4 ? 2 : 3) * c.kmax + 1) * 4u;
        const size_t lds = recs + (c.tiles ? size_t(c.ucap) * 36u : 0u);
        const int b = c.gbatch > 0 ? c.gbatch : 8;
        switch (pass) {
            case kP2: k_p2_3<<<nb, kBlock, recs, st>>>(G, S, c.B); break;
            case kP3: k_binterp3<true><<<nb, kBlock, 0, st>>>(G, S, c.B); break;
            case kP5: k_binterp3<false><<<nb, kBlock, 0, st>>>(G, S, c.B); break;
            case kP4:
                if (c.tiles) k_p4_3<true><<<nb, kBlock, lds, st>>>(G, S, c.B);
                else k_p4_3<false><<<nb, kBlock, lds, st>>>(G, S, c.B);
                break;
#define NAGA_S3(STG)                                                                          \
    case kS##STG:                                                                             \
        if (c.tiles) {                                                                        \
            if (b == 32) k_stage3<STG, 32, true><<<nb, kBlock, lds, st>>>(G, S, c.B);         \
            else if (b == 16) k_stage3<STG, 16, true><<<nb, kBlock, lds, st>>>(G, S, c.B);    \
            else k_stage3<STG, 8, true><<<nb, kBlock, lds, st>>>(G, S, c.B);                  \
        } else {                                                                              \
            if (b == 32) k_stage3<STG, 32><<<nb, kBlock, lds, st>>>(G, S, c.B);               \
            else if (b == 16) k_stage3<STG, 16><<<nb, kBlock, lds, st>>>(G, S, c.B);          \
            else k_stage3<STG, 8><<<nb, kBlock, lds, st>>>(G, S, c.B);                        \
        }                                                                                     \
        break;
            NAGA_S3(1) NAGA_S3(2) NAGA_S3(3) NAGA_S3(4)
#undef NAGA_S3
            default: break;
        }
        check_launch("step pass");
        return;
    }
    switch (pass) {
        case kP1:
            if (p1_fuses_pml(Q, c.unions)) {
                k_p1_pml<<<nb, kP1Fused, 0, st>>>(G, S, cp, c.B);
            } else if (p1_uses_dma(Q, c.unions)) {
                k_p1_dma<<<nb, kP1Chunk, 0, st>>>(G, S, cp, c.B);
            } else {
                k_p1<Q><<<nb, kP1Block, 0, st>>>(G, S, cp, c.B);
            }
            break;
        case kPl:
            if (v4) k_pml_local_v4<<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_pml_local<Q><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
        case kPm: k_means<Q><<<nb, 256, 0, st>>>(G, S, c.B); break;
        case kPf:
            if (v4) k_pml_field_v4<<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_pml_field<Q><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
        case kP2:
            if (c.ktmpl == 32) k_p2<Q, 32><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else if (c.ktmpl == 16) k_p2<Q, 16><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_p2<Q, 0><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
        case kP3:
            if (v4 && c.tiles)
                k_binterp_u<true><<<nb, kBlock, binterp_lds<true>(), st>>>(G, S, c.B);
            else if (v4) k_binterp_v4<true><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_binterp<Q, true><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
        case kP4:
            if (c.ktmpl == 32) k_p4<Q, 32><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else if (c.ktmpl == 16) k_p4<Q, 16><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_p4<Q, 0><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
        case kS1: launch_stage_pass<Q, 1>(nb, G, S, c, st); break;
        case kS2: launch_stage_pass<Q, 2>(nb, G, S, c, st); break;
        case kS3: launch_stage_pass<Q, 3>(nb, G, S, c, st); break;
        case kS4: launch_stage_pass<Q, 4>(nb, G, S, c, st); break;
        default:
            if (Q == 27 && c.unions && c.tiles)
                k_binterp_u<false><<<nb, kBlock, binterp_lds<false>(), st>>>(G, S, c.B);
            else if (Q == 27 && c.unions) k_binterp_v4<false><<<nb, kBlock, 0, st>>>(G, S, c.B);
            else k_binterp<Q, false><<<nb, kBlock, 0, st>>>(G, S, c.B);
            break;
    }
    check_launch("step pass");
}

int ktemplate(int k) { return (k == 48 || k == 32 || k == 16) ? k : 0; }

}  // namespace

int64_t pml_parts(int Q, int k, int64_t N, int64_t nfluid, int64_t n_rq) {
    if (p1_fuses_pml(Q, step_uses_unions(Q, k, N))) return (nfluid + kP1Fused - 1) / kP1Fused;
    const int64_t per = step_uses_unions(Q, k, N)
                            ? int64_t(kV4Nodes) * kPmlRounds
                            : int64_t(kBlock / 64) * (Q == 27 ? Lattice<27>::NPW : Lattice<9>::NPW) *
                                  kPmlRounds;
    return (n_rq + per - 1) / per;
}

int64_t pml_part_lo(int Q, int k, int64_t N, int64_t nbulk) {
    return p1_fuses_pml(Q, step_uses_unions(Q, k, N)) ? nbulk / kP1Fused : 0;
}

void step_block_counts(int Q, const StepGeom& G, const StepBuffers& S, int64_t* nb, int three) {
    const bool un = step_uses_unions(Q, G.k, G.N);
    const int npb = kBlock / 64 * (Q == 27 ? Lattice<27>::NPW : Lattice<9>::NPW);
    auto lanes = [&](int64_t n) { return (n + npb - 1) / npb; };
    const int64_t tn = un ? v4_nodes_per_block(G.k, S.ulist != nullptr) : kV4Nodes;
    auto tiles = [&](int64_t n) { return (n + tn - 1) / tn; };
    auto tiles32 = [&](int64_t n) { return (n + kV4Nodes - 1) / kV4Nodes; };
    const bool pml = G.nlayer > 0 && S.n_rq > 0;
    const int64_t nbnd = G.N - G.nfluid;
    const bool fused = p1_fuses_pml(Q, un);
    nb[kP1] = fused ? (G.N + kP1Fused - 1) / kP1Fused
                    : p1_uses_dma(Q, un) ? (G.N + kP1Chunk - 1) / kP1Chunk
                                         : (G.N + kP1Block - 1) / kP1Block;
    nb[kPl] = (pml && !fused) ? S.nparts : 0;
    nb[kPm] = pml ? 2 * Q : 0;
    nb[kPf] = pml ? (un ? tiles32(G.nlayer) : lanes(G.nlayer)) : 0;
    nb[kP2] = (pml && S.n_pml > 0)
                  ? (un ? (S.ulist ? S.n_ptiles : tiles(S.n_pml)) : lanes(S.n_pml))
                  : 0;
    nb[kP3] = nbnd > 0 ? (un ? tiles32(nbnd) : lanes(nbnd)) : 0;
    // with tile unions, one more block per split tile (its second half)
    const int64_t ns = (un && S.ulist) ? S.n_split : 0;
    nb[kP4] = un ? tiles(G.N) + ns : lanes(G.N);
    nb[kS1] = un ? tiles(G.N) + ns : lanes(G.N);
    nb[kS2] = nb[kS3] = nb[kS4] = un ? tiles(G.nfluid) + ns : lanes(G.nfluid);
    if (three < 0 ? q9_three(Q, G.N) : (Q == 9 && three > 0)) {
        auto t84 = [](int64_t n) { return (n + kQ9Npb - 1) / kQ9Npb; };
        nb[kP4] = nb[kS1] = t84(G.N);
        nb[kS2] = nb[kS3] = nb[kS4] = t84(G.nfluid);
        nb[kP2] = (pml && S.n_pml > 0) ? t84(S.n_pml) : 0;
        nb[kP3] = nb[kP5] = nbnd > 0 ? t84(nbnd) : 0;
    }
    nb[kP5] = nb[kP3];
}

template <int Q>
void launch_rk4_stage(int stage, const StepGeom& G, const StepBuffers& S, hipStream_t st) {
    int64_t nb[kStepKernels];
    step_block_counts(Q, G, S, nb);
    const PassCfg c{ktemplate(G.k), step_uses_unions(Q, G.k, G.N), S.ucap, S.ucap_f, BatchArgs{}, 0,
                    S.ulist != nullptr, G.k, S.tile_ng != nullptr, q9_three(Q, G.N)};
    launch_pass<Q>(kS1 + stage - 1, nb[kS1 + stage - 1], G, S, CollisionParams{}, c, st);
}

template <int Q>
void launch_step(const StepGeom& G, const StepBuffers& S, const CollisionParams& cp,
                 hipStream_t st, hipEvent_t* ev) {
    int64_t nb[kStepKernels];
    step_block_counts(Q, G, S, nb);
    const PassCfg c{ktemplate(G.k), step_uses_unions(Q, G.k, G.N), S.ucap, S.ucap_f, BatchArgs{}, 0,
                    S.ulist != nullptr, G.k, S.tile_ng != nullptr, q9_three(Q, G.N)};
    for (int p = 0; p < kStepKernels; ++p) {
        if (ev) hip_check(hipEventRecord(ev[p], st), "event");
        launch_pass<Q>(p, nb[p], G, S, cp, c, st);
    }
    if (ev) hip_check(hipEventRecord(ev[kStepKernels], st), "event");
}

template <int Q>
void launch_step_batch(const BatchPlan& plan, hipStream_t st) {
    const StepGeom G{};
    const StepBuffers S{};
    for (int p = 0; p < kStepKernels; ++p) {
        const PassCfg c{plan.k_template, plan.unions, plan.ucap, plan.ucap_f,
                        BatchArgs{plan.dev_scenes, plan.dev_bstart + p * (plan.nscenes + 1),
                                  plan.nscenes},
                        plan.gather_batch, plan.tiles, plan.kmax, plan.ragged, plan.three};
        launch_pass<Q>(p, plan.blocks[p], G, S, CollisionParams{}, c, st);
    }
}

void launch_probe(int Q, int64_t n, int k, const uint32_t* idx, const float* w,
                  const float* held, const float* g, double rho0, double* out,
                  hipStream_t st) {
    if (n == 0) return;
    const int64_t nb = (n + kBlock - 1) / kBlock;
    if (Q == 27)
        k_probe<27><<<nb, kBlock, 0, st>>>(n, k, idx, w, held, g, rho0, out);
    else
        k_probe<9><<<nb, kBlock, 0, st>>>(n, k, idx, w, held, g, rho0, out);
    check_launch("probe");
}

void launch_source_field(const int32_t* nodes, const int32_t* off, const int32_t* ids,
                         const float* prof, int64_t n, const SourceAmps& amps, float* field,
                         hipStream_t st) {
    if (n == 0) return;
    k_source_field<<<(n + kBlock - 1) / kBlock, kBlock, 0, st>>>(nodes, off, ids, prof, n, amps,
                                                                 field);
    check_launch("density source");
}

template <int Q>
void launch_macroscopic(const StepGeom& G, const float* g, float* rho,
                        double* drho, float* u, hipStream_t st) {
    k_macro<Q><<<(G.N + kBlock - 1) / kBlock, kBlock, 0, st>>>(G, g, rho, drho, u);
    check_launch("macroscopic");
}

template void launch_step<9>(const StepGeom&, const StepBuffers&, const CollisionParams&, hipStream_t, hipEvent_t*);
template void launch_step<27>(const StepGeom&, const StepBuffers&, const CollisionParams&, hipStream_t, hipEvent_t*);
template void launch_rk4_stage<9>(int, const StepGeom&, const StepBuffers&, hipStream_t);
template void launch_step_batch<9>(const BatchPlan&, hipStream_t);
template void launch_step_batch<27>(const BatchPlan&, hipStream_t);
template void launch_rk4_stage<27>(int, const StepGeom&, const StepBuffers&, hipStream_t);
template void launch_macroscopic<9>(const StepGeom&, const float*, float*, double*, float*, hipStream_t);
template void launch_macroscopic<27>(const StepGeom&, const float*, float*, double*, float*, hipStream_t);

}  // namespace naga_amd
