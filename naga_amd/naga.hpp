// naga.hpp -- C++ facade over the C-ABI (include/naga_amd.h) with the
// reference's naga:: names, so code written against NAGA's nonlocal LBM
// engine builds against this library by swapping the include.
//
// Reference surface mirrored (paths under naga/):
//   fluids/nonlocal_lattice_boltzmann/simulation_engine.cuh:39-121
//   fluids/nonlocal_lattice_boltzmann/simulation_nodes.cuh:41-66
//   fluids/nonlocal_lattice_boltzmann/simulation_variables.cuh:40-60
//   fluids/nonlocal_lattice_boltzmann/node_provider.cuh:41-50
//   fluids/nonlocal_lattice_boltzmann/node_providers/uniform_grid_provider.cuh:65-161
//   segmentation/nearest_neighbors.cuh:95-477
//   linalg/batched_matrix_inverse.cuh:325-397
//   interpolation/radial_point_method.cuh:118-139
//   nonlocal_calculus/divergence.cuh:53-212
//
// Differences a caller sees: arrays are host std::vector (first index
// fastest as in Scalix: points are [i][d]); device state stays inside the
// engine and solution() copies it out; density sources and observers are
// not device lambdas (the initial pulse is set with set_initial_pulse).
// Errors arrive as std::invalid_argument / std::runtime_error with the
// reference's "naga::" prefixes.
#pragma once

#include <array>
#include <cstdint>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "naga_amd.h"

namespace naga {

namespace abi {
inline void check(int status, const char* where) {
    if (status == NAGA_OK) return;
    const std::string msg = std::string(where) + naga_last_error();
    if (status == NAGA_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace abi

namespace fluids::nonlocal_lbm {

template <class T>
struct d2q9_lattice {};
template <class T>
struct d3q27_lattice {};

template <class Lattice>
struct lattice_traits;
template <class T>
struct lattice_traits<d2q9_lattice<T>> {
    using value_type = T;
    static constexpr unsigned size = 9;
    static constexpr unsigned dimensions = 2;
    static constexpr T lattice_speed_of_sound = T(0.57735026918962576450914878050196);
};
template <class T>
struct lattice_traits<d3q27_lattice<T>> {
    using value_type = T;
    static constexpr unsigned size = 27;
    static constexpr unsigned dimensions = 3;
    static constexpr T lattice_speed_of_sound = T(0.57735026918962576450914878050196);
};

template <class T>
struct simulation_nodes {
    std::vector<T> points;            // [i][d], order: bulk, layer, boundary
    std::vector<T> boundary_normals;  // [b][d], inward unit normals
    std::vector<T> layer_absorption;  // [num_layer_points]
    size_t num_bulk_points{};
    size_t num_layer_points{};
    size_t num_ghost_nodes{0};
    size_t num_boundary_points{};
    T nodal_spacing{};
};

template <class T>
struct problem_parameters {
    T fluid_viscosity{};
    T time_step{};
    T speed_of_sound{};
    T nominal_density{};
    T lattice_viscosity{};
    T lattice_time_step{};
    T lattice_characteristic_frequency{};
};

template <class Lattice>
struct state_variables {
    using value_type = typename lattice_traits<Lattice>::value_type;
    static constexpr unsigned lattice_size = lattice_traits<Lattice>::size;
    std::array<std::vector<value_type>, lattice_size> lattice_distributions{};
    struct {
        std::vector<value_type> fluid_velocity;  // [d][i]
        std::vector<value_type> fluid_density;   // [i]
    } macroscopic_values;
};

template <class Lattice>
class node_provider {
  public:
    using value_type = typename lattice_traits<Lattice>::value_type;
    virtual simulation_nodes<value_type> get() const = 0;
    virtual ~node_provider() = default;
};

// Uniform grid over [min, max] with spacing h, absorbing layer of the given
// thickness and peak absorption, interior nodes jittered by jitter * h with
// a seeded SplitMix64 stream (jitter 0 gives the reference's plain grid).
template <class Lattice>
class uniform_grid_provider : public node_provider<Lattice> {
  public:
    using value_type = typename lattice_traits<Lattice>::value_type;
    static constexpr unsigned dimensions = lattice_traits<Lattice>::dimensions;

    uniform_grid_provider(const std::array<double, dimensions>& min,
                          const std::array<double, dimensions>& max, double grid_spacing,
                          double absorption_layer_thickness = 0.0,
                          double absorption_coefficient = 0.0, double jitter = 0.0,
                          uint64_t seed = 0x4E414741ull)
        : min_(min), max_(max), h_(grid_spacing), layer_(absorption_layer_thickness),
          absorption_(absorption_coefficient), jitter_(jitter), seed_(seed) {}

    simulation_nodes<value_type> get() const override {
        const char* where = "naga::nonlocal_lbm::uniform_grid_provider::get: ";
        const int64_t n = naga_uniform_grid_size(int(dimensions), min_.data(), max_.data(), h_);
        if (n <= 0) abi::check(NAGA_INVALID_ARGUMENT, where);
        const size_t nn = size_t(n);
        std::vector<float> pts(nn * dimensions), nrm(nn * dimensions), sig(nn);
        int64_t counts[3];
        abi::check(naga_uniform_grid(int(dimensions), min_.data(), max_.data(), h_, jitter_,
                                     seed_, layer_, absorption_, pts.data(), nrm.data(),
                                     sig.data(), counts),
                   where);
        simulation_nodes<value_type> out;
        out.num_bulk_points = size_t(counts[0]);
        out.num_layer_points = size_t(counts[1]);
        out.num_boundary_points = size_t(counts[2]);
        out.points.assign(pts.begin(), pts.end());
        out.boundary_normals.assign(nrm.begin(), nrm.begin() + counts[2] * dimensions);
        out.layer_absorption.assign(sig.begin(), sig.begin() + counts[1]);
        out.nodal_spacing = value_type(h_);
        return out;
    }

  private:
    std::array<double, dimensions> min_, max_;
    double h_, layer_, absorption_, jitter_;
    uint64_t seed_;
};

// Stencil settings the reference hard-codes in init_domain
// (detail/simulation_engine.cuh:169-258); the defaults here are its values
// except the divergence RBF (r^3 instead of the multiquadric).
struct stencil_config {
    int support_size = 32;
    int boundary_support_size = 32;
    int rbf = NAGA_RBF_POLYHARMONIC3;
    int boundary_rbf = NAGA_RBF_MULTIQUADRIC;
    bool reference_pml_indexing = true;
    int min_support_size = 0;  // > 0: per-node widths in [min, support_size] by density
    double filter_strength = 0.0;  // 0: the reference's 0.75 per step; < 0: no filter
    int solve_precision = 64;      // 32: FP32 weight solves, as the reference
};

// The reference's own stencil settings (reference-parity mode): k = 32
// multiquadric RBFs for the divergence and the boundary interpolator
// (interpolation/radial_point_method.cuh:44-66, 88-90) with FP32 solves
// (linalg/batched_matrix_inverse.cuh:229-250).
inline stencil_config reference_stencil_config() {
    stencil_config s;
    s.rbf = NAGA_RBF_MULTIQUADRIC;
    s.boundary_rbf = NAGA_RBF_MULTIQUADRIC;
    s.solve_precision = 32;
    return s;
}

// Declarative density source (naga_engine_add_density_source): while
// t_start <= t < t_end, nodes within radius of center receive
// amplitude * cos(2 pi frequency t + phase) * P(r / radius) per step.
struct density_source_spec {
    std::array<double, 3> center{};
    double radius = 0.0, amplitude = 0.0, frequency = 0.0, phase = 0.0;
    double t_start = 0.0, t_end = 1e300;
};

template <class Lattice>
class simulation_engine {
  public:
    using value_type = typename lattice_traits<Lattice>::value_type;
    static constexpr unsigned lattice_size = lattice_traits<Lattice>::size;
    static constexpr unsigned dimensions = lattice_traits<Lattice>::dimensions;
    using lattice_type = Lattice;
    using simulation_domain_t = simulation_nodes<value_type>;
    using problem_parameters_t = ::naga::fluids::nonlocal_lbm::problem_parameters<value_type>;
    using solution_t = state_variables<Lattice>;

    explicit simulation_engine(stencil_config stencil = {}) : stencil_(stencil) {}

    void set_problem_parameters(value_type fluid_viscosity, value_type nominal_density,
                                value_type time_step, value_type speed_of_sound,
                                value_type characteristic_frequency = 0) {
        const value_type ratio = lattice_traits<Lattice>::lattice_speed_of_sound / speed_of_sound;
        params_.fluid_viscosity = fluid_viscosity;
        params_.nominal_density = nominal_density;
        params_.time_step = time_step;
        params_.speed_of_sound = speed_of_sound;
        params_.lattice_viscosity = fluid_viscosity / ratio;
        params_.lattice_time_step = time_step / ratio;
        params_.lattice_characteristic_frequency = characteristic_frequency * ratio;
        have_params_ = true;
    }

    void init_domain(const simulation_nodes<value_type>& domain) {
        const char* where = "naga::nonlocal_lbm::simulation_engine::init_domain: ";
        if (!have_params_)
            abi::check(NAGA_INVALID_ARGUMENT, where);
        if (domain.num_ghost_nodes != 0)
            throw std::invalid_argument(std::string(where) + "ghost nodes are not supported");
        naga_scene_config cfg{};
        cfg.lattice = int32_t(lattice_size);
        cfg.k = stencil_.support_size;
        cfg.k_boundary = stencil_.boundary_support_size;
        cfg.rbf = stencil_.rbf;
        cfg.rbf_boundary = stencil_.boundary_rbf;
        cfg.pml_ref_index = stencil_.reference_pml_indexing ? 1 : 0;
        cfg.k_min = stencil_.min_support_size;
        cfg.nu = double(params_.fluid_viscosity);
        cfg.rho0 = double(params_.nominal_density);
        cfg.dt = double(params_.time_step);
        cfg.c = double(params_.speed_of_sound);
        cfg.h = double(domain.nodal_spacing);
        cfg.filter_strength = stencil_.filter_strength;
        cfg.solve_precision = stencil_.solve_precision;
        naga_engine* e = nullptr;
        abi::check(naga_engine_create(&cfg, &e), where);
        handle_.reset(e, [](naga_engine* p) { naga_engine_destroy(p); });
        const size_t N = domain.num_bulk_points + domain.num_layer_points +
                         domain.num_boundary_points;
        if (domain.points.size() != N * dimensions)
            throw std::invalid_argument(
                std::string(where) +
                "The number of points in the simulation domain does not match the sum of the "
                "number of bulk, layer, and boundary points.");
        std::vector<float> p(domain.points.begin(), domain.points.end());
        std::vector<float> n(domain.boundary_normals.begin(), domain.boundary_normals.end());
        std::vector<float> s(domain.layer_absorption.begin(), domain.layer_absorption.end());
        abi::check(naga_engine_init_domain(handle_.get(), p.data(),
                                           int64_t(domain.num_bulk_points),
                                           int64_t(domain.num_layer_points),
                                           int64_t(domain.num_boundary_points),
                                           n.empty() ? nullptr : n.data(),
                                           s.empty() ? nullptr : s.data()),
                   where);
        domain_ = domain;
        reset();
    }

    void init_domain(const node_provider<Lattice>& nodes) { init_domain(nodes.get()); }

    // The one-shot Gaussian density source of the reference's example
    // (examples/nonlocal_lattice_boltzmann.cu:79-125) as an initial state.
    void set_initial_pulse(const std::array<double, dimensions>& center, double amplitude,
                           double width) {
        abi::check(naga_engine_set_pulse(live(), center.data(), amplitude, width),
                   "naga::nonlocal_lbm::simulation_engine::set_initial_pulse: ");
        stale_ = true;
    }

    void step_forward() { step_forward(1); }
    void step_forward(int nsteps) {
        abi::check(naga_engine_step(live(), nsteps),
                   "naga::nonlocal_lbm::simulation_engine::step_forward: ");
        stale_ = true;
    }

    // f = w everywhere, time 0 (detail/simulation_engine.cuh:1377-1386)
    void reset() {
        const std::array<double, dimensions> c{};
        abi::check(naga_engine_set_pulse(live(), c.data(), 0.0, 1.0),
                   "naga::nonlocal_lbm::simulation_engine::reset: ");
        stale_ = true;
    }

    value_type time() const {
        double t = 0.0;
        abi::check(naga_engine_info(live(), nullptr, &t, nullptr, nullptr),
                   "naga::nonlocal_lbm::simulation_engine::time: ");
        return value_type(t);
    }

    const simulation_domain_t& domain() const { return domain_; }
    const problem_parameters_t& problem_parameters() const { return params_; }
    value_type speed_of_sound() const { return params_.speed_of_sound; }

    // Copies the state out of the engine (the values the reference's
    // solution() holds: macroscopic values from the start of the last step).
    const solution_t& solution() const {
        if (stale_) {
            const size_t N = domain_.points.size() / dimensions;
            std::vector<float> rho(N), u(N * dimensions);
            std::vector<double> f(N * lattice_size);
            abi::check(naga_engine_get_macroscopic(live(), rho.data(), u.data(), nullptr, 0),
                       "naga::nonlocal_lbm::simulation_engine::solution: ");
            abi::check(naga_engine_get_distributions(live(), f.data(), nullptr),
                       "naga::nonlocal_lbm::simulation_engine::solution: ");
            solution_.macroscopic_values.fluid_density.assign(rho.begin(), rho.end());
            solution_.macroscopic_values.fluid_velocity.assign(u.begin(), u.end());
            for (unsigned a = 0; a < lattice_size; ++a)
                solution_.lattice_distributions[a].assign(f.begin() + a * N,
                                                          f.begin() + (a + 1) * N);
            stale_ = false;
        }
        return solution_;
    }

    void register_density_source(const density_source_spec& src) {
        abi::check(naga_engine_add_density_source(live(), src.center.data(), src.radius,
                                                  src.amplitude, src.frequency, src.phase,
                                                  src.t_start, src.t_end),
                   "naga::nonlocal_lbm::simulation_engine::register_density_source: ");
    }

    // State carried between steps as bytes (save_state/load_state of
    // simulation_engine.cuh:108-116, in this library's binary format).
    std::vector<char> save_state() const {
        int64_t n = 0;
        abi::check(naga_engine_state_size(live(), &n),
                   "naga::nonlocal_lbm::simulation_engine::save_state: ");
        std::vector<char> buf(static_cast<size_t>(n));
        abi::check(naga_engine_save_state(live(), buf.data(), n),
                   "naga::nonlocal_lbm::simulation_engine::save_state: ");
        return buf;
    }
    void load_state(const std::vector<char>& buf) {
        abi::check(naga_engine_load_state(live(), buf.data(), int64_t(buf.size())),
                   "naga::nonlocal_lbm::simulation_engine::load_state: ");
        stale_ = true;
    }

    naga_engine* native_handle() const { return handle_.get(); }

  private:
    naga_engine* live() const {
        if (!handle_)
            throw std::runtime_error(
                "naga::nonlocal_lbm::simulation_engine: init_domain must be called first");
        return handle_.get();
    }

    stencil_config stencil_;
    problem_parameters_t params_{};
    bool have_params_ = false;
    simulation_domain_t domain_{};
    std::shared_ptr<naga_engine> handle_;
    mutable solution_t solution_{};
    mutable bool stale_ = true;
};

// density_observer (simulation_observers/density_observer.cuh:47-113):
// the density at one location, interpolated from its 32 nearest nodes
// with multiquadric RPIM weights; update() records the value held by the
// engine (from the start of its last step) and the time of that step.
template <class Lattice>
class density_observer {
  public:
    using value_type = typename lattice_traits<Lattice>::value_type;
    static constexpr unsigned dimensions = lattice_traits<Lattice>::dimensions;

    density_observer(const simulation_engine<Lattice>& engine,
                     const std::array<value_type, dimensions>& location,
                     value_type time_multiplier = 1)
        : engine_(&engine), time_multiplier_(time_multiplier) {
        float p[3] = {0.f, 0.f, 0.f};
        for (unsigned d = 0; d < dimensions; ++d) p[d] = float(location[d]);
        abi::check(naga_engine_add_probes(engine.native_handle(), p, 1, NAGA_RBF_MULTIQUADRIC,
                                          &set_),
                   "naga::nonlocal_lbm::density_observer: ");
    }

    // call after each step_forward()
    void update() {
        double v = 0.0, t = 0.0;
        abi::check(naga_engine_sample_probes(engine_->native_handle(), set_, 0, &v),
                   "naga::nonlocal_lbm::density_observer::update: ");
        abi::check(naga_engine_info(engine_->native_handle(), nullptr, &t, nullptr, nullptr),
                   "naga::nonlocal_lbm::density_observer::update: ");
        const double dt = double(engine_->problem_parameters().time_step);
        density_values_.push_back(value_type(v));
        time_values_.push_back(value_type((t - dt) * time_multiplier_));
    }

    const std::vector<value_type>& density_values() const { return density_values_; }
    const std::vector<value_type>& time_values() const { return time_values_; }

  private:
    const simulation_engine<Lattice>* engine_;
    value_type time_multiplier_;
    int32_t set_ = -1;
    std::vector<value_type> density_values_, time_values_;
};

}  // namespace fluids::nonlocal_lbm

namespace segmentation {
// Exact k nearest neighbours (k smallest (d^2, index)) of each query among
// ref; points are [i][d].  d2 and idx come back [query][k].
template <uint32_t Dimensions>
void batched_nearest_neighbors(uint32_t k, const std::vector<float>& ref,
                               const std::vector<float>& query, std::vector<float>& d2,
                               std::vector<uint32_t>& idx) {
    const int64_t nr = int64_t(ref.size() / Dimensions), nq = int64_t(query.size() / Dimensions);
    d2.resize(size_t(nq) * k);
    idx.resize(size_t(nq) * k);
    abi::check(naga_knn(ref.data(), nr, query.data(), nq, int32_t(Dimensions), int32_t(k),
                        d2.data(), idx.data()),
               "naga::segmentation::batched_nearest_neighbors: ");
}
}  // namespace segmentation

namespace linalg {
// Solves A X = B for each of `batch` column-major n x n systems (FP64,
// partial pivoting); the reference inverts instead, which this replaces for
// its callers.  Returns the per-item info (non-zero: zero pivot).
inline std::vector<int32_t> batched_lu_solve(int32_t n, int64_t batch,
                                             const std::vector<double>& A,
                                             const std::vector<double>& B, int32_t nrhs,
                                             std::vector<double>& X) {
    std::vector<int32_t> info(static_cast<size_t>(batch), 0);
    X.resize(size_t(batch) * n * nrhs);
    abi::check(naga_batched_lu_solve_f64(A.data(), n, batch, B.data(), nrhs, X.data(),
                                         info.data()),
               "naga::linalg::batched_matrix_inverse: ");
    return info;
}

// linalg::matrix<T, sparse_csr> (linalg/matrix.cuh:231-437): device-resident
// CSR matrix with FP32 values.
class sparse_csr_matrix {
  public:
    // create_from_index_stencil(columns, index_stencil, values)
    // (matrix.cuh:300-368): index_stencil and values are [row][k]; row r
    // keeps its k entries in stencil order.
    static sparse_csr_matrix create_from_index_stencil(int64_t columns,
                                                       const std::vector<uint32_t>& index_stencil,
                                                       const std::vector<float>& values,
                                                       int32_t k) {
        if (k <= 0 || index_stencil.size() % size_t(k) != 0 ||
            values.size() != index_stencil.size())
            throw std::invalid_argument(
                "naga::linalg::matrix::create_from_index_stencil: index stencil and values "
                "must be [rows][k]");
        sparse_csr_matrix m;
        m.rows_ = int64_t(index_stencil.size() / size_t(k));
        m.cols_ = columns;
        naga_csr* h = nullptr;
        abi::check(naga_csr_from_stencil(m.rows_, columns, k, index_stencil.data(),
                                         values.data(), &h),
                   "naga::linalg::matrix::create_from_index_stencil: ");
        m.h_.reset(h, [](naga_csr* p) { naga_csr_destroy(p); });
        return m;
    }
    // square form (matrix.cuh:370-380)
    static sparse_csr_matrix create_from_index_stencil(const std::vector<uint32_t>& index_stencil,
                                                       const std::vector<float>& values,
                                                       int32_t k) {
        return create_from_index_stencil(int64_t(index_stencil.size() / size_t(k > 0 ? k : 1)),
                                         index_stencil, values, k);
    }
    int64_t rows() const { return rows_; }
    int64_t columns() const { return cols_; }
    naga_csr* handle() const { return h_.get(); }

  private:
    std::shared_ptr<naga_csr> h_;
    int64_t rows_ = 0, cols_ = 0;
};

// matrix_mult (matrix.cuh:438-629): y = alpha A x + beta y
inline void matrix_mult(float alpha, float beta, const sparse_csr_matrix& A,
                        const std::vector<float>& x, std::vector<float>& y) {
    if (int64_t(x.size()) != A.columns())
        throw std::invalid_argument("naga::linalg::matrix_mult: x has incorrect size");
    y.resize(size_t(A.rows()));
    abi::check(naga_csr_mult(A.handle(), alpha, beta, x.data(), y.data()),
               "naga::linalg::matrix_mult: ");
}
inline void matrix_mult(const sparse_csr_matrix& A, const std::vector<float>& x,
                        std::vector<float>& y) {
    matrix_mult(1.0f, 0.0f, A, x, y);
}
}  // namespace linalg

namespace interpolation {
// radial_point_method<T>::compute_weights with group size 1: w[q][k].
template <uint32_t Dimensions>
std::vector<double> rpim_weights(const std::vector<float>& source,
                                 const std::vector<uint32_t>& indices, uint32_t k,
                                 const std::vector<float>& query, int rbf, double spacing) {
    const int64_t nq = int64_t(query.size() / Dimensions);
    std::vector<double> w(size_t(nq) * k);
    abi::check(naga_rpim_weights(source.data(), int64_t(source.size() / Dimensions),
                                 indices.data(), int32_t(k), query.data(), nq,
                                 int32_t(Dimensions), rbf, spacing, w.data()),
               "naga::interpolation::radial_point_method::compute_weights: ");
    return w;
}
}  // namespace interpolation

namespace nonlocal_calculus {
// divergence_operator<T, D>: create builds the stencil and FP64 weights,
// apply evaluates sum_s sum_d W(d,s,i) (F_d(idx(s,i)) - centering) in FP32.
template <class T, uint32_t Dimensions>
class divergence_operator {
  public:
    static divergence_operator create(const std::vector<float>& domain, uint32_t k = 32,
                                      int rbf = NAGA_RBF_POLYHARMONIC3) {
        divergence_operator op;
        op.k_ = k;
        op.n_ = int64_t(domain.size() / Dimensions);
        op.idx_.resize(size_t(op.n_) * k);
        std::vector<double> W(size_t(op.n_) * k * Dimensions);
        abi::check(naga_divergence_weights(domain.data(), op.n_, domain.data(), op.n_,
                                           int32_t(Dimensions), int32_t(k), rbf,
                                           op.idx_.data(), W.data()),
                   "naga::nonlocal_calculus::divergence_operator::create: ");
        op.W_.assign(W.begin(), W.end());
        return op;
    }

    // field: [i][d] values of F at the domain points; result: [i]
    void apply(const std::vector<float>& field, std::vector<float>& result,
               float centering = 0.0f) const {
        result.resize(size_t(n_));
        abi::check(naga_divergence_apply(idx_.data(), W_.data(), n_, int32_t(k_),
                                         int32_t(Dimensions), field.data(),
                                         int64_t(field.size() / Dimensions), centering,
                                         result.data()),
                   "naga::nonlocal_calculus::divergence_operator::apply: ");
    }

    const std::vector<float>& weights() const { return W_; }
    const std::vector<uint32_t>& support_indices() const { return idx_; }
    int64_t size() const { return n_; }
    uint32_t support_size() const { return k_; }

    // from stencils [i][k] and weights [i][k][d] (the operator's load half)
    static divergence_operator from_weights(std::vector<uint32_t> idx, std::vector<float> W,
                                            uint32_t k) {
        divergence_operator op;
        op.k_ = k;
        op.n_ = k ? int64_t(idx.size() / k) : 0;
        op.idx_ = std::move(idx);
        op.W_ = std::move(W);
        return op;
    }

  private:
    uint32_t k_ = 0;
    int64_t n_ = 0;
    std::vector<uint32_t> idx_;
    std::vector<float> W_;
};

// constant_velocity_field (nonlocal_calculus/advection.cuh:58-85)
template <class T, uint32_t Dimensions>
class constant_velocity_field {
  public:
    template <class Velocity>
    static constant_velocity_field create(Velocity&& v) {
        constant_velocity_field m;
        for (uint32_t d = 0; d < Dimensions; ++d) m.v_[d] = float(v[d]);
        return m;
    }
    const float* data() const { return v_; }

  private:
    float v_[Dimensions] = {};
};

// advection_operator (nonlocal_calculus/advection.cuh:88-524).  Stencils and
// weights live on the device; step_forward runs synchronously and returns a
// ready future so callers written against the reference's asynchronous
// executor keep working.
template <class T, uint32_t Dimensions>
class advection_operator {
  public:
    static advection_operator create(const std::vector<float>& domain, uint32_t k = 32,
                                     int rbf = NAGA_RBF_POLYHARMONIC3) {
        advection_operator op;
        op.n_ = int64_t(domain.size() / Dimensions);
        op.k_ = k;
        naga_advection* h = nullptr;
        abi::check(naga_advection_create(domain.data(), op.n_, int32_t(Dimensions), int32_t(k),
                                         rbf, &h),
                   "naga::nonlocal_calculus::advection_operator::create: ");
        op.h_.reset(h, [](naga_advection* p) { naga_advection_destroy(p); });
        return op;
    }
    static advection_operator create(const divergence_operator<T, Dimensions>& div) {
        advection_operator op;
        op.n_ = div.size();
        op.k_ = div.support_size();
        naga_advection* h = nullptr;
        abi::check(naga_advection_from_weights(div.support_indices().data(),
                                               div.weights().data(), op.n_,
                                               int32_t(Dimensions), int32_t(op.k_), &h),
                   "naga::nonlocal_calculus::advection_operator::create: ");
        op.h_.reset(h, [](naga_advection* p) { naga_advection_destroy(p); });
        return op;
    }

    // constant velocity
    std::future<void> step_forward(const constant_velocity_field<T, Dimensions>& v,
                                   const std::vector<float>& f0, std::vector<float>& f, T dt,
                                   T centering_offset = T(0), size_t boundary_idx_start = 0) {
        return run(v.data(), 0, f0, f, dt, centering_offset, boundary_idx_start);
    }
    // per-node velocity [i][d]
    std::future<void> step_forward(const std::vector<float>& velocity,
                                   const std::vector<float>& f0, std::vector<float>& f, T dt,
                                   T centering_offset = T(0), size_t boundary_idx_start = 0) {
        if (velocity.size() != size_t(n_) * Dimensions)
            throw std::invalid_argument(
                "naga::nonlocal_calculus::advection_operator::velocity field has incorrect size");
        return run(velocity.data(), 1, f0, f, dt, centering_offset, boundary_idx_start);
    }

    divergence_operator<T, Dimensions> divergence_op() const {
        std::vector<uint32_t> idx(size_t(n_) * k_);
        std::vector<float> W(size_t(n_) * k_ * Dimensions);
        abi::check(naga_advection_weights(h_.get(), idx.data(), W.data()),
                   "naga::nonlocal_calculus::advection_operator::divergence_op: ");
        return divergence_operator<T, Dimensions>::from_weights(std::move(idx), std::move(W), k_);
    }

  private:
    std::future<void> run(const float* v, int per_node, const std::vector<float>& f0,
                          std::vector<float>& f, T dt, T off, size_t end) {
        if (int64_t(f0.size()) != n_)
            throw std::invalid_argument(
                "naga::nonlocal_calculus::advection_operator::f0 must have the same number of "
                "elements as the domain for which this operator was created.");
        f = f0;
        abi::check(naga_advection_step(h_.get(), v, per_node, f.data(), double(dt), double(off),
                                       int64_t(end), 1),
                   "naga::nonlocal_calculus::advection_operator::step_forward: ");
        std::promise<void> done;
        done.set_value();
        return done.get_future();
    }

    std::shared_ptr<naga_advection> h_;
    int64_t n_ = 0;
    uint32_t k_ = 0;
};
}  // namespace nonlocal_calculus

}  // namespace naga
