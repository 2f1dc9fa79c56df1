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

#include <algorithm>
#include <array>
#include <cmath>
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

// simulation_observer<Lattice> (fluids/nonlocal_lattice_boltzmann/
// simulation_observer.cuh:41-55): update() is called at the start of every
// step with the time and the state of that moment.
template <class Lattice>
class simulation_observer {
  public:
    using value_type = typename lattice_traits<Lattice>::value_type;
    using simulation_domain_t = simulation_nodes<value_type>;
    using problem_parameters_t = ::naga::fluids::nonlocal_lbm::problem_parameters<value_type>;
    using solution_t = state_variables<Lattice>;
    virtual ~simulation_observer() = default;
    virtual void update(const value_type& time, const simulation_domain_t& domain,
                        const problem_parameters_t& parameters, const solution_t& solution) = 0;
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
    // With observers attached every step first hands them the state at its
    // start (update_observers, detail/simulation_engine.cuh:1332, 1444-1448).
    void step_forward(int nsteps) {
        if (observers_.empty()) {
            abi::check(naga_engine_step(live(), nsteps),
                       "naga::nonlocal_lbm::simulation_engine::step_forward: ");
            stale_ = true;
            return;
        }
        for (int n = 0; n < nsteps; ++n) {
            const solution_t& sol = fresh_solution();
            const value_type t = time();
            for (simulation_observer<Lattice>* o : observers_)
                o->update(t, domain_, params_, sol);
            abi::check(naga_engine_step(live(), 1),
                       "naga::nonlocal_lbm::simulation_engine::step_forward: ");
            stale_ = true;
        }
    }

    // attach_observer / detach_observer (simulation_engine.cuh:90-96)
    void attach_observer(simulation_observer<Lattice>& observer) {
        observers_.push_back(&observer);
    }
    void detach_observer(simulation_observer<Lattice>& observer) {
        observers_.erase(std::remove(observers_.begin(), observers_.end(), &observer),
                         observers_.end());
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
    const solution_t& solution() const { return load_solution(false); }

  private:
    // the solution with fresh macroscopic values of the current state (what
    // the reference's step start computes before updating the observers)
    const solution_t& fresh_solution() const { return load_solution(true); }

    const solution_t& load_solution(bool fresh) const {
        if (stale_ || fresh) {
            const size_t N = domain_.points.size() / dimensions;
            std::vector<float> rho(N), u(N * dimensions);
            std::vector<double> f(N * lattice_size);
            abi::check(naga_engine_get_macroscopic(live(), rho.data(), u.data(), nullptr,
                                                   fresh ? 1 : 0),
                       "naga::nonlocal_lbm::simulation_engine::solution: ");
            abi::check(naga_engine_get_distributions(live(), f.data(), nullptr),
                       "naga::nonlocal_lbm::simulation_engine::solution: ");
            solution_.macroscopic_values.fluid_density.assign(rho.begin(), rho.end());
            solution_.macroscopic_values.fluid_velocity.assign(u.begin(), u.end());
            for (unsigned a = 0; a < lattice_size; ++a)
                solution_.lattice_distributions[a].assign(f.begin() + a * N,
                                                          f.begin() + (a + 1) * N);
            stale_ = fresh;  // a fresh copy is not what solution() holds
        }
        return solution_;
    }

  public:
    // The vtk_observer output of the current state (naga_engine_write_vtp).
    void write_vtp(const std::string& path, bool fresh = true) const {
        abi::check(naga_engine_write_vtp(live(), path.c_str(), fresh ? 1 : 0),
                   "naga::nonlocal_lbm::vtk_observer::update: ");
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

    std::vector<simulation_observer<Lattice>*> observers_;
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
class density_observer : public simulation_observer<Lattice> {
  public:
    using base = simulation_observer<Lattice>;
    using typename base::problem_parameters_t;
    using typename base::simulation_domain_t;
    using typename base::solution_t;
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

    // as an attached observer: the density of the state at the step's start
    void update(const value_type& time, const simulation_domain_t&, const problem_parameters_t&,
                const solution_t&) override {
        double v = 0.0;
        abi::check(naga_engine_sample_probes(engine_->native_handle(), set_, 1, &v),
                   "naga::nonlocal_lbm::density_observer::update: ");
        density_values_.push_back(value_type(v));
        time_values_.push_back(value_type(time * time_multiplier_));
    }

    const std::vector<value_type>& density_values() const { return density_values_; }
    const std::vector<value_type>& time_values() const { return time_values_; }

  private:
    const simulation_engine<Lattice>* engine_;
    value_type time_multiplier_;
    int32_t set_ = -1;
    std::vector<value_type> density_values_, time_values_;
};

// vtk_observer<Lattice> (simulation_observers/vtk_observer.cuh:45-431):
// writes output_directory/result.<frame>.vtp at each update, frame =
// floor(time_multiplier * t * frame_rate) when frame_rate != 0 (frames
// already written are skipped), else ceil(t / dt).
template <class Lattice>
class vtk_observer : public simulation_observer<Lattice> {
  public:
    using base = simulation_observer<Lattice>;
    using typename base::problem_parameters_t;
    using typename base::simulation_domain_t;
    using typename base::solution_t;
    using typename base::value_type;

    vtk_observer(const simulation_engine<Lattice>& engine, std::string output_directory,
                 value_type frame_rate = 0, value_type time_multiplier = 1)
        : engine_(&engine), dir_(std::move(output_directory)), frame_rate_(frame_rate),
          time_multiplier_(time_multiplier) {}

    void update(const value_type& time, const simulation_domain_t&,
                const problem_parameters_t& parameters, const solution_t&) override {
        uint32_t frame;
        if (frame_rate_ != 0) {
            frame = uint32_t(time_multiplier_ * time * frame_rate_);
            if (frame < current_frame_) return;
        } else {
            frame = uint32_t(std::ceil(double(time) / double(parameters.time_step) - 1e-9));
        }
        current_frame_ = frame + 1;
        engine_->write_vtp(dir_ + "/result." + std::to_string(frame) + ".vtp", true);
    }

  private:
    const simulation_engine<Lattice>* engine_;
    std::string dir_;
    value_type frame_rate_, time_multiplier_;
    uint32_t current_frame_ = 0;
};

}  // namespace fluids::nonlocal_lbm

namespace segmentation {
// nd_cubic_segmentation<T, D> (segmentation/nd_cubic_segmentation.cuh:
// 219-461), binned on the device: points are [i][d].
template <class T, uint32_t Dimensions>
class nd_cubic_segmentation {
  public:
    nd_cubic_segmentation(const std::vector<T>& points, uint32_t approx_partition_size)
        : n_(points.size() / Dimensions) {
        create(points, int32_t(approx_partition_size), 0.0);
    }
    nd_cubic_segmentation(const std::vector<T>& points, T partition_width)
        : n_(points.size() / Dimensions) {
        create(points, 0, double(partition_width));
    }
    std::array<uint32_t, Dimensions> shape() const { return shape_; }
    size_t partition_count() const { return count_; }
    size_t point_count() const { return n_; }
    T partition_width() const { return T(width_); }
    // partition sizes, index offsets and the indices of the points by cell
    void partitions(std::vector<uint32_t>& sizes, std::vector<uint32_t>& offsets,
                    std::vector<uint32_t>& indices) const {
        sizes.resize(count_);
        offsets.resize(count_);
        indices.resize(n_);
        abi::check(naga_segmentation_partitions(h_.get(), sizes.data(), offsets.data(),
                                                indices.data()),
                   "naga::segmentation::nd_cubic_segmentation: ");
    }
    naga_segmentation* handle() const { return h_.get(); }

  private:
    void create(const std::vector<T>& points, int32_t approx, double width) {
        std::vector<float> p(points.begin(), points.end());
        naga_segmentation* h = nullptr;
        abi::check(naga_segmentation_create(p.data(), int64_t(n_), int32_t(Dimensions), approx,
                                            width, &h),
                   "naga::segmentation::nd_cubic_segmentation: ");
        h_.reset(h, [](naga_segmentation* s) { naga_segmentation_destroy(s); });
        int64_t shape[3];
        int64_t count = 0;
        abi::check(naga_segmentation_shape(h, shape, nullptr, &width_, &count),
                   "naga::segmentation::nd_cubic_segmentation: ");
        for (uint32_t d = 0; d < Dimensions; ++d) shape_[d] = uint32_t(shape[d]);
        count_ = size_t(count);
    }
    std::shared_ptr<naga_segmentation> h_;
    size_t n_ = 0, count_ = 0;
    double width_ = 0.0;
    std::array<uint32_t, Dimensions> shape_{};
};

// batched_nearest_neighbors(k, query_points, segmented_ref_points)
// (segmentation/nearest_neighbors.cuh:454-477): d2 and idx come back
// [query][k].
template <class T, uint32_t Dimensions>
void batched_nearest_neighbors(uint32_t k, const std::vector<T>& query,
                               const nd_cubic_segmentation<T, Dimensions>& segmented,
                               std::vector<float>& d2, std::vector<uint32_t>& idx) {
    std::vector<float> q(query.begin(), query.end());
    const int64_t nq = int64_t(q.size() / Dimensions);
    d2.resize(size_t(nq) * k);
    idx.resize(size_t(nq) * k);
    abi::check(naga_segmentation_knn(segmented.handle(), q.data(), nq, int32_t(k), d2.data(),
                                     idx.data()),
               "naga::segmentation::batched_nearest_neighbors: ");
}

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

// batched_matrix_inverse(A, A_inv, copy_A) (linalg/batched_matrix_inverse.cuh:
// 325-397): A and A_inv hold `batch` column-major n x n matrices; FP64
// arithmetic for either T.  copy_A is accepted for the signature (A is never
// modified here).  Returns the per-item pivot info (the reference leaves it
// unchecked).
template <class T>
std::vector<int32_t> batched_matrix_inverse(int32_t n, int64_t batch, const std::vector<T>& A,
                                            std::vector<T>& A_inv, bool copy_A = true) {
    (void)copy_A;
    if (A.size() != size_t(batch) * n * n)
        throw std::invalid_argument("naga::linalg::Matrices must be square");
    std::vector<double> a(A.begin(), A.end()), inv(A.size());
    std::vector<int32_t> info(size_t(batch), 0);
    abi::check(naga_batched_matrix_inverse_f64(a.data(), n, batch, inv.data(), info.data()),
               "naga::linalg::");
    A_inv.assign(inv.begin(), inv.end());
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
// radial_point_method<T> (interpolation/radial_point_method.cuh:92-280):
// create_interpolator keeps the weights of each query over its group's
// stencil on the device; interpolate applies them.  Points are [i][d];
// interpolating_indices [group][k]; fields [point][component].
template <class T>
class radial_point_method {
  public:
    static size_t get_scratchpad_size(size_t query_size, uint32_t support_size,
                                      uint32_t dimensions, uint32_t group_size = 1) {
        const int64_t v = naga_rpim_scratchpad_size(int64_t(query_size), int32_t(support_size),
                                                    int32_t(dimensions), int32_t(group_size),
                                                    int32_t(sizeof(T)));
        if (v < 0)
            throw std::invalid_argument(
                "naga::interpolation::radial_point_method::query_size must be a multiple of "
                "group_size");
        return size_t(v);
    }

    template <uint32_t Dimensions>
    static radial_point_method create_interpolator(
        const std::vector<T>& source_points, const std::vector<uint32_t>& interpolating_indices,
        uint32_t support_size, const std::vector<T>& query_points, T approx_particle_spacing,
        uint32_t group_size = 1, int rbf = NAGA_RBF_MULTIQUADRIC) {
        std::vector<float> s(source_points.begin(), source_points.end()),
            q(query_points.begin(), query_points.end());
        radial_point_method r;
        r.n_src_ = int64_t(s.size() / Dimensions);
        r.n_q_ = int64_t(q.size() / Dimensions);
        r.k_ = support_size;
        naga_rpim* h = nullptr;
        abi::check(naga_rpim_create(s.data(), r.n_src_, interpolating_indices.data(),
                                    int32_t(support_size), q.data(), r.n_q_, int32_t(Dimensions),
                                    int32_t(group_size), rbf, double(approx_particle_spacing),
                                    &h),
                   "naga::interpolation::radial_point_method::create_interpolator: ");
        r.h_.reset(h, [](naga_rpim* p) { naga_rpim_destroy(p); });
        return r;
    }

    // weights [query][k] (compute_weights)
    std::vector<float> weights() const {
        std::vector<float> w(size_t(n_q_) * k_);
        abi::check(naga_rpim_get_weights(h_.get(), w.data()),
                   "naga::interpolation::radial_point_method: ");
        return w;
    }

    // field [source][components] -> destination [query][components]
    void interpolate(const std::vector<T>& field, std::vector<T>& destination,
                     uint32_t components = 1, T centering_offset = T(0)) const {
        std::vector<float> f(field.begin(), field.end()), d(size_t(n_q_) * components);
        abi::check(naga_rpim_interpolate(h_.get(), f.data(), int64_t(f.size() / components),
                                         int32_t(components), d.data(),
                                         float(centering_offset)),
                   "naga::interpolation::radial_point_method::interpolate: ");
        destination.assign(d.begin(), d.end());
    }

  private:
    std::shared_ptr<naga_rpim> h_;
    int64_t n_src_ = 0, n_q_ = 0;
    uint32_t k_ = 0;
};

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
        return create(domain, domain, 0, k, rbf);
    }

    // create(domain, query_points, override_query_size)
    // (nonlocal_calculus/divergence.cuh:53-166): stencils over the domain
    // for each query; a non-zero override_query_size sizes the operator for
    // that many queries, the rows past the query count zero.
    static divergence_operator create(const std::vector<float>& domain,
                                      const std::vector<float>& query_points,
                                      size_t override_query_size, uint32_t k = 32,
                                      int rbf = NAGA_RBF_POLYHARMONIC3) {
        const int64_t nd = int64_t(domain.size() / Dimensions);
        const int64_t nq = int64_t(query_points.size() / Dimensions);
        const int64_t size = override_query_size == 0 ? nq : int64_t(override_query_size);
        if (size < nq)
            throw std::invalid_argument(
                "naga::nonlocal_calculus::divergence_operator::override_query_size must be at "
                "least as large as the number of query points.");
        divergence_operator op;
        op.k_ = k;
        op.n_ = size;
        op.idx_.assign(size_t(size) * k, 0u);
        std::vector<double> W(size_t(size) * k * Dimensions, 0.0);
        abi::check(naga_divergence_weights(domain.data(), nd, query_points.data(), nq,
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
