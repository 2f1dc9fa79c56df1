// Setup kernels (see setup_kernels.hpp).
#include <cmath>
#include <vector>

#include "setup_kernels.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace naga_amd {

namespace {

constexpr uint64_t kMaxKey = ~0ull;

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const int lo = __shfl(int(uint32_t(v)), src);
    const int hi = __shfl(int(uint32_t(v >> 32)), src);
    return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    const int lo = __shfl_xor(int(uint32_t(v)), m);
    const int hi = __shfl_xor(int(uint32_t(v >> 32)), m);
    return (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo);
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
    return __longlong_as_double(
        (long long)shfl_xor64(uint64_t(__double_as_longlong(v)), m));
}

__device__ __forceinline__ float shfl_xor_d(float v, int m) { return __shfl_xor(v, m); }

// value of lane `src` (wave-uniform) for every lane
__device__ __forceinline__ float bcast_d(float v, int src) {
    const int s = __builtin_amdgcn_readfirstlane(src);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), s));
}
__device__ __forceinline__ double bcast_d(double v, int src) {
    const int s = __builtin_amdgcn_readfirstlane(src);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(uint32_t(b)), s);
    const int hi = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(b) >> 32)), s);
    return __longlong_as_double((long long)((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}

// FP32 squared distance, fixed order, no contraction (the kNN key).
__device__ __forceinline__ float dist2_f32(int D, const float* a, const float* b) {
#pragma clang fp contract(off)
    float s = (a[0] - b[0]) * (a[0] - b[0]);
    for (int d = 1; d < D; ++d) {
        const float t = (a[d] - b[d]) * (a[d] - b[d]);
        s = t + s;
    }
    return s;
}

// bitonic sort of one key per lane across the wavefront, ascending
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kk = 2; kk <= 64; kk <<= 1)
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1) {
            const uint64_t p = shfl_xor64(v, j);
            const bool up = (lane & kk) == 0, lower = (lane & j) == 0;
            v = (lower == up) ? (p < v ? p : v) : (p > v ? p : v);
        }
    return v;
}

// sort a bitonic sequence ascending
__device__ __forceinline__ uint64_t wave_merge64(uint64_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 32; j > 0; j >>= 1) {
        const uint64_t p = shfl_xor64(v, j);
        const bool lower = (lane & j) == 0;
        v = lower ? (p < v ? p : v) : (p > v ? p : v);
    }
    return v;
}

// Merge the candidates of one 64-wide batch into the running top list
// (lane t holds the t-th smallest key).
__device__ __forceinline__ void knn_merge(uint64_t key, uint64_t& top, uint64_t& kth, int k) {
    if (!__any(key < kth)) return;
    const int lane = threadIdx.x & 63;
    const uint64_t sorted = wave_sort64(key);
    const uint64_t rev = shfl64(sorted, 63 - lane);
    top = wave_merge64(top < rev ? top : rev);
    kth = shfl64(top, k - 1);
}

__global__ __launch_bounds__(256) void k_knn(const CellGrid* gp, const float* ref,
                                             const uint32_t* sorted,
                                             const uint32_t* start,
                                             const float* query, int64_t n_q,
                                             int k, float* d2o, uint32_t* idxo) {
    const int lane = threadIdx.x & 63;
    const int64_t q = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (q >= n_q) return;
    const CellGrid G = *gp;
    const int D = G.D;
    float x[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < D; ++d) x[d] = query[q * D + d];
    int64_t qc[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        int64_t c = int64_t(floor((double(x[d]) - G.lo[d]) / G.w));
        qc[d] = c < 0 ? 0 : c >= G.dims[d] ? G.dims[d] - 1 : c;
    }
    uint64_t top = kMaxKey, kth = kMaxKey;
    for (int64_t r = 0;; ++r) {
        const int64_t zr = D == 3 ? r : 0;
        for (int64_t dz = -zr; dz <= zr; ++dz) {
            const int64_t cz = qc[2] + dz;
            if (cz < 0 || cz >= G.dims[2]) continue;
            for (int64_t dy = -r; dy <= r; ++dy) {
                const int64_t cy = qc[1] + dy;
                if (cy < 0 || cy >= G.dims[1]) continue;
                const bool face = (dy == r || dy == -r) || (D == 3 && (dz == r || dz == -r));
                const int64_t row = G.dims[0] * (cy + G.dims[1] * cz);
                int64_t seg[2][2];
                int nseg = 0;
                if (face) {
                    int64_t a = qc[0] - r, b = qc[0] + r;
                    if (a < 0) a = 0;
                    if (b >= G.dims[0]) b = G.dims[0] - 1;
                    seg[nseg][0] = a;
                    seg[nseg++][1] = b;
                } else {
                    if (qc[0] - r >= 0) {
                        seg[nseg][0] = qc[0] - r;
                        seg[nseg++][1] = qc[0] - r;
                    }
                    if (qc[0] + r < G.dims[0]) {
                        seg[nseg][0] = qc[0] + r;
                        seg[nseg++][1] = qc[0] + r;
                    }
                }
                for (int t = 0; t < nseg; ++t) {
                    const int64_t p0 = start[row + seg[t][0]];
                    const int64_t p1 = start[row + seg[t][1] + 1];  // cells are row-major in x
                    for (int64_t base = p0; base < p1; base += 64) {
                        const int64_t p = base + lane;
                        uint64_t key = kMaxKey;
                        if (p < p1) {
                            const uint32_t j = sorted[p];
                            const float dd = dist2_f32(D, x, ref + int64_t(j) * D);
                            key = (uint64_t(__float_as_uint(dd)) << 32) | j;
                        }
                        knn_merge(key, top, kth, k);
                    }
                }
            }
        }
        double bound = 1e300;
        bool all = true;
        for (int d = 0; d < D; ++d) {
            if (qc[d] - r > 0) {
                const double t = double(x[d]) - (G.lo[d] + double(qc[d] - r) * G.w);
                bound = t < bound ? t : bound;
                all = false;
            }
            if (qc[d] + r < G.dims[d] - 1) {
                const double t = (G.lo[d] + double(qc[d] + r + 1) * G.w) - double(x[d]);
                bound = t < bound ? t : bound;
                all = false;
            }
        }
        if (all) break;
        if (kth != kMaxKey && bound > 0.0) {
            const double kd = double(__uint_as_float(uint32_t(kth >> 32)));
            if (kd < bound * bound * (1.0 - 1e-5)) break;
        }
    }
    if (lane < k) {
        d2o[q * k + lane] = __uint_as_float(uint32_t(top >> 32));
        idxo[q * k + lane] = uint32_t(top);
    }
}

// ---------------------------------------------------------------------------
// Wavefront LU with partial pivoting (Golub & Van Loan, Alg. 3.4.1).  Lane r
// holds row r of A (row[c] = A[r][c]) and its right-hand sides.  Row swaps
// are tracked as positions instead of moving data; ties in |pivot| go to the
// smallest current position, as a row-swapping LU scanning downwards picks.
// On return the lane at position p holds unknown p in rhs[].
// Returns the lane's position, or -1 for a zero pivot (wave-uniform).
// ---------------------------------------------------------------------------
// The rows live in LDS (lu[lane * LD + c], LD = lu_stride(NMAX), filled by
// the caller): the steps run as loops over the system's own size, whose
// indices are not compile-time constants (a register array indexed at run
// time would be placed in scratch memory; one unrolled for every size is
// too large to build).  The odd stride spreads the lanes' rows over banks.
// Only lanes below NMAX >= n own a row (lu holds NMAX rows); the others
// read row 0 where a value is needed but never used, and write nothing.
constexpr int lu_stride(int nmax) { return nmax + 1; }
template <int NMAX, class T>
__device__ __forceinline__ T* lu_row(T* lu) {
    const int lane = threadIdx.x & 63;
    return lu + (lane < NMAX ? lane : 0) * lu_stride(NMAX);
}

template <int NMAX, int NR, class T>
__device__ int wave_lu_solve(T* lu, T (&rhs)[NR], int n) {
    constexpr int LD = lu_stride(NMAX);
    const int lane = threadIdx.x & 63;
    T* r = lu_row<NMAX>(lu);
    __syncthreads();  // the rows are in place
    int pos = lane;
    bool done = lane >= n;
    bool singular = false;
    for (int j = 0; j < n; ++j) {
        T v = done ? T(-1) : fabs(r[j]);
        int bp = done ? 1 << 20 : pos, bl = lane;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const T ov = shfl_xor_d(v, o);
            const int op = __shfl_xor(bp, o), ol = __shfl_xor(bl, o);
            if (ov > v || (ov == v && op < bp)) {
                v = ov;
                bp = op;
                bl = ol;
            }
        }
        const int p = __builtin_amdgcn_readfirstlane(bl);
        if (!(v > T(0))) singular = true;
        const int pos_p = __builtin_amdgcn_readlane(pos, p);
        if (pos == j) pos = pos_p;
        if (lane == p) pos = j;
        const T* pr = lu + p * LD;  // the pivot row (every lane reads the same words)
        const T piv = pr[j];
        T prhs[NR];
#pragma unroll
        for (int m = 0; m < NR; ++m) prhs[m] = bcast_d(rhs[m], p);
        const bool upd = !done && lane != p;
        const T l = upd ? r[j] / piv : T(0);
        if (upd) {
            // columns n and above are zero in every row: nothing to update
            for (int c = j + 1; c < n; ++c) r[c] -= l * pr[c];
            r[j] = l;
#pragma unroll
            for (int m = 0; m < NR; ++m) rhs[m] -= l * prhs[m];
        }
        if (lane == p) done = true;
        __syncthreads();  // this step's rows before the next pivot is read
    }
    if (singular) return -1;
    // back substitution, unknown j lives in the lane at position j
    T s[NR];
#pragma unroll
    for (int m = 0; m < NR; ++m) s[m] = rhs[m];
    for (int j = n - 1; j >= 0; --j) {
        const unsigned long long mask = __ballot(pos == j && lane < n);
        const int lj = __builtin_amdgcn_readfirstlane(int(__ffsll((long long)mask) - 1));
        const T ujj = lu[lj * LD + j];
        const T rj = r[j];
#pragma unroll
        for (int m = 0; m < NR; ++m) {
            const T xj = bcast_d(s[m], lj) / ujj;
            if (pos == j) rhs[m] = xj;
            if (pos < j && lane < n) s[m] -= rj * xj;
        }
    }
    return lane < n ? pos : 64;
}

// unscaled quadrature points, [0]: 2D (16 used), [1]: 3D (64); both
// tables are written once per device and never change afterwards, so
// scenes of either dimension can be set up concurrently
__constant__ double c_quad_unit[2][64][3];
__constant__ double c_quad_rw[4];        // radial Gauss weights

__device__ __forceinline__ double dist2_f64(int D, const double* a, const double* b) {
    double s = (a[0] - b[0]) * (a[0] - b[0]);
    for (int d = 1; d < D; ++d) s = (a[d] - b[d]) * (a[d] - b[d]) + s;
    return s;
}

__device__ __forceinline__ double rbf_eval(int rbf, double r2, double spacing) {
    if (rbf == 1) return r2 * sqrt(r2);
    const double c = spacing * 0.1;
    return sqrt(r2 + c * c);
}

// Row r of the augmented RBF matrix [[Phi, P], [P^T, 0]] in raw coordinates
// (interpolation/detail/radial_point_method.cuh:349-398).
template <int NMAX, class T>
__device__ void g0_row(int D, const double (*xs)[3], int k, int rbf,
                       double spacing, int r, T* row) {
    const int n = k + D + 1;
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
        double v = 0.0;
        if (r < n && c < n) {
            if (r < k && c < k)
                v = rbf_eval(rbf, dist2_f64(D, xs[r], xs[c]), spacing);
            else if (r < k)
                v = (c == k) ? 1.0 : xs[r][c - k - 1];
            else if (c < k)
                v = (r == k) ? 1.0 : xs[c][r - k - 1];
        }
        if (r < NMAX) row[c] = T(v);
    }
}

// entry r of G_x (:400-433)
__device__ __forceinline__ double gx_entry(int D, const double (*xs)[3], int k,
                                           int rbf, double spacing, int r,
                                           const double* xq) {
    if (r < k) return rbf_eval(rbf, dist2_f64(D, xs[r], xq), spacing);
    if (r == k) return 1.0;
    return xq[r - k - 1];
}

// antisymmetric w4 B-spline kernel times the quadrature weight
// (nonlocal_calculus/detail/divergence.inl:55-133, :167-201)
__device__ void quad_term(int D, double delta, const double* xi, const double* xq,
                          int q, double* aw) {
    double dx[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) dx[d] = xq[d] - xi[d];
    const double rxy = sqrt(dx[1] * dx[1] + dx[0] * dx[0]);
    const double r = D == 3 ? sqrt(dx[2] * dx[2] + (dx[1] * dx[1] + dx[0] * dx[0])) : rxy;
    double g;
    if (2.0 * r / delta < 1.0)
        g = r * (2.0 * delta - 3.0 * r);
    else if (2.0 * r / delta < 2.0)
        g = (delta - r) * (delta - r);
    else
        g = 0.0;
    double qw = c_quad_rw[q % 4] * delta / 2.0;
    qw *= 2.0 * M_PI / 4.0;
    if (D == 3) qw *= M_PI / 4.0;
    if (D == 2) {
        const double C = (240.0 / 7.0) / (M_PI * pow(delta, 5.0));
        qw *= r;
        aw[0] = C * g * (dx[0] / r) * qw;
        aw[1] = C * g * (dx[1] / r) * qw;
    } else {
        const double C = 48.0 / (M_PI * pow(delta, 6.0));
        const double sp = rxy / r, cp = dx[2] / r;
        const double ct = rxy != 0.0 ? dx[0] / rxy : 0.0;
        const double st = rxy != 0.0 ? dx[1] / rxy : 0.0;
        qw *= r * rxy;
        aw[0] = C * g * sp * ct * qw;
        aw[1] = C * g * sp * st * qw;
        aw[2] = C * g * cp * qw;
    }
}

// One wavefront per query node: assemble G0, the D quadrature right-hand
// sides b_d = sum_q alpha_d(q) w(q) G_x(q), solve G0 y_d = b_d in FP64 and
// write W(d, s) = y_d[s] (+ sum_q alpha_d w for s = 0).  By linearity this
// equals the reference's per-quadrature-point RPIM weights summed against
// the kernel (detail/divergence.inl:203-215) with D solves instead of Nq.
template <int NMAX, int D, class T>
__global__ __launch_bounds__(64) void k_divw(const float* src, const float* query,
                                             int64_t n_q, const uint32_t* idx,
                                             const float* d2, int kmax, int rbf,
                                             const double* hbar, const uint8_t* kvar,
                                             float* W, double* W64, int* info) {
    __shared__ double xs[64][3];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    if (i >= n_q) return;
    // per-node stencil width k <= kmax (entries past k get zero weights)
    const int k = kvar ? int(kvar[i]) : kmax;
    const int n = k + D + 1;
    if (lane >= k && lane < kmax)
        for (int d = 0; d < D; ++d) {
            W[(i * kmax + lane) * D + d] = 0.0f;
            if (W64) W64[(i * kmax + lane) * D + d] = 0.0;
        }
    if (lane < k) {
        const int64_t j = idx[i * kmax + lane];
        for (int d = 0; d < D; ++d) xs[lane][d] = double(src[j * D + d]);
    }
    __syncthreads();
    double xi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) xi[d] = double(query[i * D + d]);
    const double delta = 0.05 * sqrt(double(d2[i * kmax + 1]));
    const double spacing = hbar[i / 16384];
    __shared__ T lu[NMAX * lu_stride(NMAX)];
    g0_row<NMAX>(D, xs, k, rbf, spacing, lane, lu_row<NMAX>(lu));
    double b64[D], S[D];
    for (int d = 0; d < D; ++d) b64[d] = S[d] = 0.0;
    const int Nq = D == 2 ? 16 : 64;
    for (int q = 0; q < Nq; ++q) {
        double xq[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < D; ++d) xq[d] = c_quad_unit[D - 2][q][d] * delta + xi[d];
        double aw[3];
        quad_term(D, delta, xi, xq, q, aw);
        const double gx = lane < n ? gx_entry(D, xs, k, rbf, spacing, lane, xq) : 0.0;
        for (int d = 0; d < D; ++d) {
            b64[d] += aw[d] * gx;
            S[d] += aw[d];
        }
    }
    T b[D];
    for (int d = 0; d < D; ++d) b[d] = T(b64[d]);
    const int pos = wave_lu_solve<NMAX, D, T>(lu, b, n);
    if (pos < 0) {
        if (lane == 0) info[i] = 1;
        return;
    }
    if (lane == 0) info[i] = 0;
    if (pos < k) {
        for (int d = 0; d < D; ++d) {
            const double v = double(b[d]) + (pos == 0 ? S[d] : 0.0);
            W[(i * kmax + pos) * D + d] = float(v);
            if (W64) W64[(i * kmax + pos) * D + d] = v;
        }
    }
}

template <int NMAX, int D, class T>
__global__ __launch_bounds__(64) void k_rpim(const float* src, const float* query,
                                             int64_t n_q, const uint32_t* idx,
                                             int k, int rbf, double spacing,
                                             float* w, double* w64, int* info, int group) {
    __shared__ double xs[64][3];
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    if (i >= n_q) return;
    const int n = k + D + 1;
    if (lane < k) {  // the queries of a group share their group's stencil
        const int64_t j = idx[(i / group) * k + lane];
        for (int d = 0; d < D; ++d) xs[lane][d] = double(src[j * D + d]);
    }
    __syncthreads();
    double xq[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) xq[d] = double(query[i * D + d]);
    __shared__ T lu[NMAX * lu_stride(NMAX)];
    g0_row<NMAX>(D, xs, k, rbf, spacing, lane, lu_row<NMAX>(lu));
    T b[1] = {T(lane < n ? gx_entry(D, xs, k, rbf, spacing, lane, xq) : 0.0)};
    const int pos = wave_lu_solve<NMAX, 1, T>(lu, b, n);
    if (pos < 0) {
        if (lane == 0) info[i] = 1;
        return;
    }
    if (lane == 0) info[i] = 0;
    if (pos < k) {
        w[i * k + pos] = float(b[0]);
        if (w64) w64[i * k + pos] = double(b[0]);
    }
}

template <int NMAX>
__global__ __launch_bounds__(64) void k_batched_lu(int n, int nrhs, int64_t batch,
                                                   double* A, double* B, int* info) {
    const int lane = threadIdx.x;
    const int64_t b = blockIdx.x;
    if (b >= batch) return;
    const double* Ab = A + b * int64_t(n) * n;
    double* Bb = B + b * int64_t(n) * nrhs;
    __shared__ double lu[NMAX * lu_stride(NMAX)];
    double* r = lu_row<NMAX>(lu);
    // right-hand sides are processed four at a time, each group against a
    // fresh copy of the matrix
    for (int m0 = 0; m0 < nrhs; m0 += 4) {
        __syncthreads();  // the previous group's factorisation is done with lu
#pragma unroll 4
        for (int c = 0; c < NMAX; ++c)
            if (lane < NMAX) r[c] = (lane < n && c < n) ? Ab[lane + int64_t(c) * n] : 0.0;
        double rhs[4];
        for (int m = 0; m < 4; ++m)
            rhs[m] = (lane < n && m0 + m < nrhs) ? Bb[lane + int64_t(m0 + m) * n] : 0.0;
        const int pos = wave_lu_solve<NMAX, 4, double>(lu, rhs, n);
        if (pos < 0) {
            if (lane == 0) info[b] = 1;
            return;
        }
        if (pos < n)
            for (int m = 0; m < 4 && m0 + m < nrhs; ++m) Bb[pos + int64_t(m0 + m) * n] = rhs[m];
    }
    if (lane == 0) info[b] = 0;
}

__global__ void k_filter(int D, const float* pts, int64_t n, const uint32_t* idx,
                         int kmax, double sg, const uint8_t* kvar, float* wf) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = kvar ? int(kvar[i]) : kmax;
    double xi[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < D; ++d) xi[d] = double(pts[i * D + d]);
    double sum = 0.0;
    for (int s = 0; s < k; ++s) {
        double xj[3] = {0.0, 0.0, 0.0};
        const int64_t j = idx[i * kmax + s];
        for (int d = 0; d < D; ++d) xj[d] = double(pts[j * D + d]);
        const double r = sqrt(dist2_f64(D, xi, xj)) / sg;
        sum += exp(-0.5 * r * r);
    }
    for (int s = 0; s < kmax; ++s) {
        double wv = 0.0;
        if (s < k) {
            double xj[3] = {0.0, 0.0, 0.0};
            const int64_t j = idx[i * kmax + s];
            for (int d = 0; d < D; ++d) xj[d] = double(pts[j * D + d]);
            const double r = sqrt(dist2_f64(D, xi, xj)) / sg;
            wv = exp(-0.5 * r * r) / sum;
        }
        wf[i * kmax + s] = float(wv);
    }
}

__global__ void k_make_slab(int64_t n, int k, int nc, const uint32_t* idx,
                            const float* w, float* slab) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    float* r = slab + t * (1 + nc);
    r[0] = __uint_as_float(idx[t]);
    for (int c = 0; c < nc; ++c) r[1 + c] = w[t * nc + c];
}

// One thread per (node, 32-bit word) of the packed divergence records.
__global__ void k_pack_div_records(int64_t n, int k, const float* src, uint32_t* dst) {
    const int units = packed_div_units(k);
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * units * 4) return;
    const int64_t i = t / (int64_t(units) * 4);
    const int word = int(t - i * units * 4);
    const int u = word / 4, g = u / 7, w = u % 7;
    const float* rec = src + (i * k + 8 * g) * 4;  // the group's 8 records
    uint32_t v = 0;
    if (u == units - 1) {
        v = 0;  // padding unit
    } else if (w < 6) {
        const int f = 4 * w + word % 4;  // float f of the group's 24 weights
        v = __float_as_uint(rec[(f / 3) * 4 + 1 + f % 3]);
    } else {
        const int m = word % 4;  // records 2m, 2m + 1
        const uint32_t lo = __float_as_uint(rec[(2 * m) * 4]);
        const uint32_t hi = __float_as_uint(rec[(2 * m + 1) * 4]);
        v = (lo & 0xffffu) | (hi << 16);
    }
    dst[t] = v;
}

// One thread per (node, 32-bit word) of the packed filter records.
__global__ void k_pack_filter_records(int64_t n, int k, const float* src, uint32_t* dst) {
    const int units = packed_filter_units(k);
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * units * 4) return;
    const int64_t i = t / (int64_t(units) * 4);
    const int word = int(t - i * units * 4);
    const int u = word / 4, g = u / 3, w = u % 3;
    const float* rec = src + (i * k + 8 * g) * 2;  // the group's 8 {position, weight}
    uint32_t v = 0;
    if (u == units - 1) {
        v = 0;  // padding unit
    } else if (w < 2) {
        v = __float_as_uint(rec[(4 * w + word % 4) * 2 + 1]);
    } else {
        const int m = word % 4;  // records 2m, 2m + 1
        const uint32_t lo = __float_as_uint(rec[(2 * m) * 2]);
        const uint32_t hi = __float_as_uint(rec[(2 * m + 1) * 2]);
        v = (lo & 0xffffu) | (hi << 16);
    }
    dst[t] = v;
}

// One block per tile: the tile's k * tile_nodes (<= 4096) neighbour indices
// are sorted in LDS (bitonic), flagged where they differ from their
// predecessor and compacted by a block-wide scan of the flags.
constexpr int kUnionMax = 4096;

__device__ __forceinline__ int block_exclusive_scan(int* v, int n, int* tmp) {
    // v[0..n) in LDS, n <= kUnionMax, blockDim.x == 256: per-thread chunks of
    // n / 256 (n is a power of two >= 256 here), then a scan of the chunk sums.
    const int per = n / 256, t = threadIdx.x;
    int sum = 0;
    for (int e = 0; e < per; ++e) sum += v[t * per + e];
    tmp[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int add = t >= off ? tmp[t - off] : 0;
        __syncthreads();
        tmp[t] += add;
        __syncthreads();
    }
    int run = tmp[t] - sum;  // exclusive prefix of this chunk
    for (int e = 0; e < per; ++e) {
        const int f = v[t * per + e];
        v[t * per + e] = run;
        run += f;
    }
    __syncthreads();
    const int total = tmp[255];
    __syncthreads();
    return total;
}

// Sorted distinct keys of src[0..M) (M <= kUnionMax) into uniq[0..U) by
// the whole block; returns U.
__device__ int block_union(const uint32_t* src, int M, uint32_t* keys, uint32_t* uniq,
                           int* rank, int* tmp) {
    int P = 256;
    while (P < M) P <<= 1;
    for (int e = threadIdx.x; e < P; e += 256) keys[e] = e < M ? src[e] : 0xffffffffu;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int e = threadIdx.x; e < P / 2; e += 256) {
                const int i = 2 * stride * (e / stride) + (e % stride), j = i + stride;
                const bool up = (i & size) == 0;
                const uint32_t a = keys[i], b = keys[j];
                if ((a > b) == up) {
                    keys[i] = b;
                    keys[j] = a;
                }
            }
            __syncthreads();
        }
    for (int e = threadIdx.x; e < P; e += 256)
        rank[e] = (e < M && (e == 0 || keys[e] != keys[e - 1])) ? 1 : 0;
    __syncthreads();
    const int U = block_exclusive_scan(rank, P, tmp);
    // compact the distinct keys (rank is the exclusive scan of the flags)
    for (int e = threadIdx.x; e < M; e += 256)
        if (e == 0 || keys[e] != keys[e - 1]) uniq[rank[e]] = keys[e];
    __syncthreads();
    return U;
}

// Writes uniq[0..U) to out and, for the M stencil entries src[0..M), the
// entry's position in uniq into the index field of the slab records.
__device__ void write_union(const uint32_t* src, int M, const uint32_t* uniq, int U,
                            uint32_t* out, float* slab_a, int nca, float* slab_b, int ncb) {
    for (int e = threadIdx.x; e < U; e += 256) out[e] = uniq[e];
    for (int r = threadIdx.x; r < M; r += 256) {
        const uint32_t j = src[r];
        int lo = 0, hi = U;  // first position with uniq[pos] >= j
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uniq[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        const float slot = __uint_as_float(uint32_t(lo));
        slab_a[r * (1 + nca)] = slot;
        if (slab_b) slab_b[r * (1 + ncb)] = slot;
    }
    __syncthreads();
}

// One block per tile: the tile's k * tile_nodes (<= 4096) neighbour indices
// are sorted in LDS (bitonic), flagged where they differ from their
// predecessor and compacted by a block-wide scan of the flags.  A tile
// whose union exceeds cap (> 0) while each half-tile's union fits is
// split: its list is the first half's union followed by the second
// half's, each half's records hold positions in its own part, and
// usplit[t] is the first part's length (0 for tiles not split).
// Counting pass (ulist == nullptr): counts[t] = list length, usplit[t].
__global__ __launch_bounds__(256) void k_tile_union(int64_t nfluid, int k, int tile_nodes,
                                                    const uint32_t* idx, int cap,
                                                    int32_t* counts, int32_t* usplit,
                                                    const int64_t* uoff,
                                                    uint32_t* ulist, float* slab_a, int nca,
                                                    float* slab_b, int ncb) {
    __shared__ uint32_t keys[kUnionMax];
    __shared__ uint32_t uniq[kUnionMax];
    __shared__ int rank[kUnionMax];
    __shared__ int tmp[256];
    const int64_t node0 = int64_t(blockIdx.x) * tile_nodes;
    const int64_t left = nfluid - node0;
    const int cnt = left < tile_nodes ? int(left) : tile_nodes;
    const int M = cnt * k;
    const int h = tile_nodes / 2;  // nodes of the first half
    const uint32_t* src = idx + node0 * k;
    if (!ulist) {
        const int U = block_union(src, M, keys, uniq, rank, tmp);
        int total = U, first = 0;
        if (cap > 0 && U > cap && cnt > h) {
            const int UA = block_union(src, h * k, keys, uniq, rank, tmp);
            const int UB = block_union(src + h * k, (cnt - h) * k, keys, uniq, rank, tmp);
            if (UA <= cap && UB <= cap) {
                total = UA + UB;
                first = UA;
            }
        }
        if (threadIdx.x == 0) {
            counts[blockIdx.x] = total;
            if (usplit) usplit[blockIdx.x] = first;
        }
        return;
    }
    uint32_t* out = ulist + uoff[blockIdx.x];
    float* sa = slab_a + node0 * k * (1 + nca);
    float* sb = slab_b ? slab_b + node0 * k * (1 + ncb) : nullptr;
    const int first = usplit ? usplit[blockIdx.x] : 0;
    if (first == 0) {
        const int U = block_union(src, M, keys, uniq, rank, tmp);
        write_union(src, M, uniq, U, out, sa, nca, sb, ncb);
        return;
    }
    const int UA = block_union(src, h * k, keys, uniq, rank, tmp);
    write_union(src, h * k, uniq, UA, out, sa, nca, sb, ncb);
    const int MB = (cnt - h) * k;
    const int UB = block_union(src + h * k, MB, keys, uniq, rank, tmp);
    write_union(src + h * k, MB, uniq, UB, out + UA, sa + int64_t(h) * k * (1 + nca), nca,
                sb ? sb + int64_t(h) * k * (1 + ncb) : nullptr, ncb);
}

// Ragged record rows: node i of tile t = i / T keeps the first upg *
// ng[t] units of its full row (KP units) and pads to `ragged(ng[t])` units
// with zeros; tile t's rows start at unit off[t].
__global__ void k_ragged_records(int64_t n, int T, int KP, int upg, const int32_t* ng,
                                 const int64_t* off, const float4* src, float4* dst) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * KP) return;
    const int64_t i = e / KP;
    const int u = int(e - i * KP);
    const int64_t t = i / T;
    const int g = ng[t];
    const int kpt = upg == 7 ? ragged_div_units(g) : ragged_filter_units(g);
    if (u >= kpt) return;
    dst[off[t] + (i - t * T) * kpt + u] = u < upg * g ? src[e] : make_float4(0.f, 0.f, 0.f, 0.f);
}

// flag stays 1 if idx[i * k] == i for every i < n, else becomes 0.
__global__ void k_self_first(int64_t n, int k, const uint32_t* idx, int* flag) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && idx[i * k] != uint32_t(i)) atomicAnd(flag, 0);
}

// Stencil entries s >= kvar[i] (zero weights) of node i point at i itself.
__global__ void k_self_past_k(int64_t n, int k, const uint8_t* kvar, uint32_t* idx) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * k) return;
    const int64_t i = e / k;
    if (int(e - i * k) >= int(kvar[i])) idx[e] = uint32_t(i);
}

// out[e] = idx[e] if it is in [lo, hi), else zero_row.
__global__ void k_layer_only(int64_t n, const uint32_t* idx, uint32_t lo, uint32_t hi,
                             uint32_t zero_row, uint32_t* out) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t j = idx[e];
    out[e] = (j >= lo && j < hi) ? j : zero_row;
}

__global__ void k_flag(int64_t n, int k, const uint32_t* idx, int64_t lo,
                       int64_t hi, uint8_t* flag) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t f = 0;
    for (int s = 0; s < k; ++s) {
        const int64_t j = idx[i * k + s];
        f |= (j >= lo && j < hi) ? 1 : 0;
    }
    flag[i] = f;
}

__global__ void k_div_apply(int D, int64_t n, int k, const uint32_t* idx,
                            const float* W, const float* F, float off, float* out) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    for (int s = 0; s < k; ++s) {
        const int64_t j = idx[i * k + s];
        float v = 0.0f;
        for (int d = 0; d < D; ++d) v += W[(i * k + s) * D + d] * (F[j * D + d] - off);
        acc += v;
    }
    out[i] = acc;
}

// ---------------------------------------------------------------------------
// Cell binning on the device (segmentation/detail/nd_cubic_segmentation.inl:
// 38-273: bounds, partition sizes, offsets by a scan, indices by a scatter).
// The order of indices within a cell follows the scatter's atomics, as in
// the reference; the kNN over the cells is exact whatever that order.
// ---------------------------------------------------------------------------
// Per-block bounds of the points (grid-stride), partial[b][0..3) minima and
// [3..6) maxima.
__global__ __launch_bounds__(256) void k_bounds(const float* pts, int64_t n, int D,
                                                double* partial) {
    __shared__ float lo[3][256], hi[3][256];
    float l[3] = {3.0e38f, 3.0e38f, 3.0e38f}, h[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x; j < n;
         j += int64_t(gridDim.x) * 256)
        for (int d = 0; d < D; ++d) {
            const float v = pts[j * D + d];
            l[d] = fminf(l[d], v);
            h[d] = fmaxf(h[d], v);
        }
    for (int d = 0; d < 3; ++d) {
        lo[d][threadIdx.x] = l[d];
        hi[d][threadIdx.x] = h[d];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                lo[d][threadIdx.x] = fminf(lo[d][threadIdx.x], lo[d][threadIdx.x + w]);
                hi[d][threadIdx.x] = fmaxf(hi[d][threadIdx.x], hi[d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        partial[blockIdx.x * 6 + threadIdx.x] = double(lo[threadIdx.x][0]);
        partial[blockIdx.x * 6 + 3 + threadIdx.x] = double(hi[threadIdx.x][0]);
    }
}

// The grid from the per-block bounds (nd_cubic_segmentation.cuh:241-279).
__global__ __launch_bounds__(256) void k_grid(const double* partial, int D, int64_t n,
                                              double per_cell, double width, int64_t cap,
                                              CellGrid* g) {
    __shared__ double lo[3][kBoundBlocks], hi[3][kBoundBlocks];
    const int t = threadIdx.x;
    for (int d = 0; d < 3; ++d) {
        lo[d][t] = partial[t * 6 + d];
        hi[d][t] = partial[t * 6 + 3 + d];
    }
    __syncthreads();
    for (int w = kBoundBlocks / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int d = 0; d < 3; ++d) {
                lo[d][t] = fmin(lo[d][t], lo[d][t + w]);
                hi[d][t] = fmax(hi[d][t], hi[d][t + w]);
            }
        __syncthreads();
    }
    if (t != 0) return;
    CellGrid G{};
    G.D = D;
    double ext[3] = {1.0, 1.0, 1.0}, vol = 1.0;
    for (int d = 0; d < 3; ++d) {
        G.lo[d] = d < D ? lo[d][0] : 0.0;
        if (d < D) {
            ext[d] = hi[d][0] - lo[d][0];
            if (ext[d] < 1e-3) ext[d] = 1e-3;  // the reference's tolerance
            vol *= ext[d];
        }
    }
    double w = width > 0.0 ? width : pow(per_cell * vol / double(n), 1.0 / double(D));
    int64_t nc = 1;
    for (int it = 0; it < 4096; ++it) {
        nc = 1;
        for (int d = 0; d < 3; ++d) {
            double m = d < D ? ceil(ext[d] / w) : 1.0;
            m = m < 1.0 ? 1.0 : (m > 1048576.0 ? 1048576.0 : m);
            G.dims[d] = int64_t(m);
            nc *= G.dims[d];
        }
        if (nc <= cap) break;
        w *= 1.25;
    }
    if (nc > cap) {  // unreachable for finite input: one cell then
        G.dims[0] = G.dims[1] = G.dims[2] = 1;
        nc = 1;
        w = 1e300;
    }
    G.w = w;
    G.ncell = nc;
    *g = G;
}

// Cell of a point: floor((x - lo) / w) per axis, clamped to the grid;
// linear index with x fastest (the kNN's cell rows).
__device__ __forceinline__ int64_t cell_of(const CellGrid& G, const float* p) {
    int64_t c = 0, mul = 1;
    for (int d = 0; d < G.D; ++d) {
        int64_t ci = int64_t(floor((double(p[d]) - G.lo[d]) / G.w));
        ci = ci < 0 ? 0 : ci >= G.dims[d] ? G.dims[d] - 1 : ci;
        c += ci * mul;
        mul *= G.dims[d];
    }
    return c;
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_count_cells(const float* pts, int64_t n,
                                                     const CellGrid* gp, uint32_t* counts) {
    const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const CellGrid G = *gp;
    atomicAdd(&counts[cell_of(G, pts + j * G.D)], 1u);
}

// Exclusive scan of in[0..M) into out, block b covering [b * 2048, ..); its
// total goes to bsum[b].
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t* tmp, uint32_t& total) {
    const int t = threadIdx.x;
    tmp[t] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t a = t >= off ? tmp[t - off] : 0u;
        __syncthreads();
        tmp[t] += a;
        __syncthreads();
    }
    total = tmp[255];
    const uint32_t incl = tmp[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_local(const uint32_t* in, int64_t M, uint32_t* out,
                                                    uint32_t* bsum) {
    __shared__ uint32_t tmp[256];
    const int64_t base = int64_t(blockIdx.x) * kScanChunk + threadIdx.x * 8;
    uint32_t v[8], s = 0;
    for (int e = 0; e < 8; ++e) {
        v[e] = base + e < M ? in[base + e] : 0u;
        s += v[e];
    }
    uint32_t total;
    uint32_t run = block_scan_256(s, tmp, total);
    for (int e = 0; e < 8; ++e) {
        if (base + e < M) out[base + e] = run;
        run += v[e];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_sums(uint32_t* bsum, int64_t nb) {
    __shared__ uint32_t tmp[256];
    uint32_t carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += 256) {
        const int64_t i = c0 + threadIdx.x;
        const uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_scan_256(v, tmp, total);
        if (i < nb) bsum[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* out, int64_t M, const uint32_t* bsum) {
    const int64_t base = int64_t(blockIdx.x) * kScanChunk;
    const uint32_t add = bsum[blockIdx.x];
    for (int e = threadIdx.x; e < kScanChunk; e += 256)
        if (base + e < M) out[base + e] += add;
}

__global__ __launch_bounds__(256) void k_scatter_cells(const float* pts, int64_t n,
                                                       const CellGrid* gp, const uint32_t* start,
                                                       uint32_t* fill, uint32_t* sorted) {
    const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (j >= n) return;
    const CellGrid G = *gp;
    const int64_t c = cell_of(G, pts + j * G.D);
    sorted[start[c] + atomicAdd(&fill[c], 1u)] = uint32_t(j);
}

int pick_nmax(int n) {
    const int opts[] = {20, 28, 36, 44, 52, 64};
    for (int o : opts)
        if (n <= o) return o;
    return -1;
}

void fill_quadrature(int D, double (&tab)[64][3]) {
    static const double xr[4] = {0.339981, -0.339981, 0.861136, -0.861136};
    const int Nq = D == 2 ? 16 : 64;
    for (int q = 0; q < 64; ++q)
        for (int d = 0; d < 3; ++d) tab[q][d] = 0.0;
    for (int q = 0; q < Nq; ++q) {
        // nonlocal_calculus/detail/quadrature_points.cuh:74-91, :127-152
        const double r = (xr[q % 4] + 1.0) / 2.0;
        const double ts = 3.0 * M_PI / 2.0 / 3.0;
        if (D == 2) {
            const double th = (q / 4) * ts;
            tab[q][0] = r * cos(th);
            tab[q][1] = r * sin(th);
        } else {
            const double th = ((q / 4) % 4) * ts;
            const double ph = (q / 16) * (ts / 2.0) + M_PI / 4.0 / 2.0;
            tab[q][0] = r * sin(ph) * cos(th);
            tab[q][1] = r * sin(ph) * sin(th);
            tab[q][2] = r * cos(ph);
        }
    }
}

// Uploads both quadrature tables once per device (thread-safe).
void upload_quadrature(int /*D*/) {
    static std::mutex mu;
    static std::vector<int> done;  // devices already holding the tables
    int dev = 0;
    NAGA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (std::find(done.begin(), done.end(), dev) != done.end()) return;
    static const double wr[4] = {0.652145, 0.652145, 0.347855, 0.347855};
    double tab[2][64][3];
    fill_quadrature(2, tab[0]);
    fill_quadrature(3, tab[1]);
    NAGA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_quad_unit), tab, sizeof(tab)));
    NAGA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_quad_rw), wr, sizeof(wr)));
    done.push_back(dev);
}

template <int D, class T>
void launch_divw_d(int nmax, const float* src, const float* query, int64_t n_q,
                   const uint32_t* idx, const float* d2, int k, int rbf,
                   const double* hbar, const uint8_t* kvar, float* W, double* W64, int* info,
                   hipStream_t st) {
#define NAGA_DIVW(NM)                                                                 \
    case NM:                                                                          \
        k_divw<NM, D, T><<<n_q, 64, 0, st>>>(src, query, n_q, idx, d2, k, rbf, hbar, \
                                              kvar, W, W64, info);                    \
        break;
    switch (nmax) {
        NAGA_DIVW(20) NAGA_DIVW(28) NAGA_DIVW(36) NAGA_DIVW(44) NAGA_DIVW(52) NAGA_DIVW(64)
        default: fail(kInvalidArgument, "stencil too wide");
    }
#undef NAGA_DIVW
}

template <int D, class T>
void launch_rpim_d(int nmax, const float* src, const float* query, int64_t n_q,
                   const uint32_t* idx, int k, int rbf, double spacing, float* w,
                   double* w64, int* info, hipStream_t st, int group) {
#define NAGA_RPIM(NM)                                                                  \
    case NM:                                                                           \
        k_rpim<NM, D, T><<<n_q, 64, 0, st>>>(src, query, n_q, idx, k, rbf, spacing, w, \
                                              w64, info, group);                       \
        break;
    switch (nmax) {
        NAGA_RPIM(20) NAGA_RPIM(28) NAGA_RPIM(36) NAGA_RPIM(44) NAGA_RPIM(52) NAGA_RPIM(64)
        default: fail(kInvalidArgument, "stencil too wide");
    }
#undef NAGA_RPIM
}

}  // namespace

size_t cell_bins_bytes(int64_t n) { return bin_layout(n).total; }

CellBins launch_bin_points(int D, const float* pts, int64_t n, double per_cell, double width,
                           void* scratch_mem, size_t bytes, hipStream_t st) {
    require(D >= 1 && D <= 3, "dimension must be 1, 2 or 3");
    require(n > 0, "no reference points");
    require(n < (int64_t(1) << 31), "at most 2^31 - 1 points");
    require(per_cell > 0.0 || width > 0.0, "cell size must be positive");
    const BinLayout L = bin_layout(n);
    require(scratch_mem != nullptr && bytes >= L.total, "cell binning scratch too small");
    char* base = static_cast<char*>(scratch_mem);
    CellBins b{};
    b.grid = reinterpret_cast<CellGrid*>(base + L.grid);
    b.start = reinterpret_cast<uint32_t*>(base + L.start);
    b.counts = reinterpret_cast<uint32_t*>(base + L.counts);
    b.sorted = reinterpret_cast<uint32_t*>(base + L.sorted);
    b.cap = L.cap;
    b.n = n;
    double* partial = reinterpret_cast<double*>(base + L.partial);
    uint32_t* bsum = reinterpret_cast<uint32_t*>(base + L.bsum);
    const int64_t M = L.cap + 1;
    const int64_t nbM = (M + 255) / 256, nbn = (n + 255) / 256;
    k_bounds<<<kBoundBlocks, 256, 0, st>>>(pts, n, D, partial);
    k_grid<<<1, kBoundBlocks, 0, st>>>(partial, D, n, per_cell > 0.0 ? per_cell : 0.0,
                                       per_cell > 0.0 ? 0.0 : width, L.cap, b.grid);
    k_fill_u32<<<nbM, 256, 0, st>>>(b.counts, M, 0u);
    k_count_cells<<<nbn, 256, 0, st>>>(pts, n, b.grid, b.counts);
    k_scan_local<<<L.nblocks, 256, 0, st>>>(b.counts, M, b.start, bsum);
    k_scan_sums<<<1, 256, 0, st>>>(bsum, L.nblocks);
    k_scan_add<<<L.nblocks, 256, 0, st>>>(b.start, M, bsum);
    // the scatter counts each cell up from zero again, leaving the sizes
    k_fill_u32<<<nbM, 256, 0, st>>>(b.counts, M, 0u);
    k_scatter_cells<<<nbn, 256, 0, st>>>(pts, n, b.grid, b.start, b.counts, b.sorted);
    hip_check(hipGetLastError(), "cell binning");
    return b;
}

DeviceCells bin_points_device(int D, const float* pts, int64_t n, hipStream_t st,
                              double per_cell, double width) {
    DeviceCells c;
    const size_t bytes = cell_bins_bytes(n);
    c.mem.allocate(bytes);
    c.bins = launch_bin_points(D, pts, n, per_cell, width, c.mem.get(), bytes, st);
    return c;
}

void launch_knn(const CellBins& bins, const float* ref, const float* query, int64_t n_q, int k,
                float* d2_out, uint32_t* idx_out, hipStream_t st) {
    require(k >= 1 && k <= 64, "k must be in [1, 64]");
    require(bins.n >= k, "fewer reference points than neighbours requested");
    if (n_q == 0) return;
    const int64_t blocks = (n_q * 64 + 255) / 256;
    k_knn<<<blocks, 256, 0, st>>>(bins.grid, ref, bins.sorted, bins.start, query, n_q, k, d2_out,
                                  idx_out);
    hip_check(hipGetLastError(), "knn");
}

void launch_batched_lu_solve(int n, int nrhs, int64_t batch, double* A, double* B,
                             int* info, hipStream_t st) {
    const int nmax = pick_nmax(n);
    require(nmax > 0 && n >= 1, "n must be in [1, 64]");
    require(nrhs >= 1, "nrhs must be >= 1");
    if (batch == 0) return;
#define NAGA_LU(NM) \
    case NM: k_batched_lu<NM><<<batch, 64, 0, st>>>(n, nrhs, batch, A, B, info); break;
    switch (nmax) { NAGA_LU(20) NAGA_LU(28) NAGA_LU(36) NAGA_LU(44) NAGA_LU(52) NAGA_LU(64) }
#undef NAGA_LU
    hip_check(hipGetLastError(), "batched lu");
}

void launch_divergence_weights(int D, const float* src, const float* query,
                               int64_t n_q, const uint32_t* idx, const float* d2,
                               int k, int rbf, const double* hbar, float* W,
                               double* W64, int* info, hipStream_t st, const uint8_t* kvar,
                               int solve_bits) {
    const int nmax = pick_nmax(k + D + 1);
    require(nmax > 0, "stencil too wide (k + D + 1 must be <= 64)");
    require(k >= 2, "k must be >= 2");
    require(solve_bits == 64 || solve_bits == 32, "solve precision must be 64 or 32 bits");
    if (n_q == 0) return;
    upload_quadrature(D);
    const bool f32 = solve_bits == 32;
    if (D == 2 && f32)
        launch_divw_d<2, float>(nmax, src, query, n_q, idx, d2, k, rbf, hbar, kvar, W, W64, info,
                                st);
    else if (D == 2)
        launch_divw_d<2, double>(nmax, src, query, n_q, idx, d2, k, rbf, hbar, kvar, W, W64,
                                 info, st);
    else if (f32)
        launch_divw_d<3, float>(nmax, src, query, n_q, idx, d2, k, rbf, hbar, kvar, W, W64, info,
                                st);
    else
        launch_divw_d<3, double>(nmax, src, query, n_q, idx, d2, k, rbf, hbar, kvar, W, W64,
                                 info, st);
    hip_check(hipGetLastError(), "divergence weights");
}

void launch_rpim_weights(int D, const float* src, const float* query, int64_t n_q,
                         const uint32_t* idx, int k, int rbf, double spacing,
                         float* w, double* w64, int* info, hipStream_t st, int solve_bits,
                         int group) {
    const int nmax = pick_nmax(k + D + 1);
    require(nmax > 0, "stencil too wide (k + D + 1 must be <= 64)");
    require(group >= 1, "group size must be >= 1");
    require(solve_bits == 64 || solve_bits == 32, "solve precision must be 64 or 32 bits");
    if (n_q == 0) return;
    const bool f32 = solve_bits == 32;
    if (D == 2 && f32)
        launch_rpim_d<2, float>(nmax, src, query, n_q, idx, k, rbf, spacing, w, w64, info, st, group);
    else if (D == 2)
        launch_rpim_d<2, double>(nmax, src, query, n_q, idx, k, rbf, spacing, w, w64, info, st, group);
    else if (f32)
        launch_rpim_d<3, float>(nmax, src, query, n_q, idx, k, rbf, spacing, w, w64, info, st, group);
    else
        launch_rpim_d<3, double>(nmax, src, query, n_q, idx, k, rbf, spacing, w, w64, info, st, group);
    hip_check(hipGetLastError(), "rpim weights");
}

void launch_filter_weights(int D, const float* pts, int64_t n, const uint32_t* idx,
                           int k, double h, float* wf, hipStream_t st, const uint8_t* kvar) {
    if (n == 0) return;
    k_filter<<<(n + 255) / 256, 256, 0, st>>>(D, pts, n, idx, k, h / 2.5, kvar, wf);
    hip_check(hipGetLastError(), "filter weights");
}

void launch_make_slab(int64_t n, int k, int nc, const uint32_t* idx, const float* w,
                      float* slab, hipStream_t st) {
    if (n == 0) return;
    k_make_slab<<<(n * k + 255) / 256, 256, 0, st>>>(n, k, nc, idx, w, slab);
    hip_check(hipGetLastError(), "slab");
}

void launch_pack_div_records(int64_t n, int k, const float* src, float* dst, hipStream_t st) {
    require(k % 8 == 0 && k <= 64, "packed divergence records need k a multiple of 8");
    if (n == 0) return;
    const int64_t tot = n * packed_div_units(k) * 4;
    k_pack_div_records<<<(tot + 255) / 256, 256, 0, st>>>(n, k, src,
                                                          reinterpret_cast<uint32_t*>(dst));
    hip_check(hipGetLastError(), "pack divergence records");
}

void launch_pack_filter_records(int64_t n, int k, const float* src, float* dst,
                                hipStream_t st) {
    require(k % 8 == 0 && k <= 64, "packed filter records need k a multiple of 8");
    if (n == 0) return;
    const int64_t tot = n * packed_filter_units(k) * 4;
    k_pack_filter_records<<<(tot + 255) / 256, 256, 0, st>>>(n, k, src,
                                                             reinterpret_cast<uint32_t*>(dst));
    hip_check(hipGetLastError(), "pack filter records");
}

void launch_ragged_records(int64_t n, int T, int KP, int upg, const int32_t* ng,
                           const int64_t* off, const float* src, float* dst, hipStream_t st) {
    if (n == 0) return;
    const int64_t tot = n * KP;
    k_ragged_records<<<(tot + 255) / 256, 256, 0, st>>>(n, T, KP, upg, ng, off,
                                                        reinterpret_cast<const float4*>(src),
                                                        reinterpret_cast<float4*>(dst));
    hip_check(hipGetLastError(), "ragged records");
}

bool stencils_self_first(int64_t n, int k, const uint32_t* idx, hipStream_t st) {
    if (n == 0) return true;
    DeviceBuffer<int> flag(1);
    const int one = 1;
    flag.upload(&one, 1, st);
    k_self_first<<<(n + 255) / 256, 256, 0, st>>>(n, k, idx, flag.get());
    hip_check(hipGetLastError(), "stencil check");
    int out = 0;
    flag.download(&out, 1, st);
    NAGA_HIP(hipStreamSynchronize(st));
    return out == 1;
}

void launch_self_past_k(int64_t n, int k, const uint8_t* kvar, uint32_t* idx, hipStream_t st) {
    if (n == 0) return;
    const int64_t tot = n * k;
    k_self_past_k<<<(tot + 255) / 256, 256, 0, st>>>(n, k, kvar, idx);
    hip_check(hipGetLastError(), "stencil padding");
}

void launch_tile_unions(int64_t nfluid, int k, int tile_nodes, const uint32_t* idx, int cap,
                        int32_t* counts, int32_t* usplit, const int64_t* uoff, uint32_t* ulist,
                        float* slab_a, int ncomp_a, float* slab_b, int ncomp_b, hipStream_t st) {
    require(tile_nodes > 0 && tile_nodes % 2 == 0 && k * tile_nodes <= kUnionMax,
            "stencil too wide for tile unions");
    const int64_t ntiles = (nfluid + tile_nodes - 1) / tile_nodes;
    if (ntiles == 0) return;
    k_tile_union<<<ntiles, 256, 0, st>>>(nfluid, k, tile_nodes, idx, cap, counts, usplit, uoff,
                                          ulist, slab_a, ncomp_a, slab_b, ncomp_b);
    hip_check(hipGetLastError(), "tile unions");
}

void launch_layer_only(int64_t n, const uint32_t* idx, uint32_t lo, uint32_t hi,
                       uint32_t zero_row, uint32_t* out, hipStream_t st) {
    if (n == 0) return;
    k_layer_only<<<(n + 255) / 256, 256, 0, st>>>(n, idx, lo, hi, zero_row, out);
    hip_check(hipGetLastError(), "layer-only stencil");
}

void launch_flag_neighbours(int64_t n, int k, const uint32_t* idx, int64_t lo,
                            int64_t hi, uint8_t* flag, hipStream_t st) {
    if (n == 0) return;
    k_flag<<<(n + 255) / 256, 256, 0, st>>>(n, k, idx, lo, hi, flag);
    hip_check(hipGetLastError(), "flag");
}

void launch_divergence_apply(int D, int64_t n, int k, const uint32_t* idx,
                             const float* W, const float* F, float off, float* out,
                             hipStream_t st) {
    if (n == 0) return;
    k_div_apply<<<(n + 255) / 256, 256, 0, st>>>(D, n, k, idx, W, F, off, out);
    hip_check(hipGetLastError(), "divergence apply");
}

}  // namespace naga_amd
