// Drives the naga:: C++ facade (naga_amd/naga.hpp) the way reference code
// drives NAGA's engine.  Modes:
//   errors : checks the error mapping without touching a GPU
//   run    : 2D scene from the uniform grid provider, pulse, N steps; prints
//            "sum_drho <value> max_u <value> time <value>"
//   extras : save/load round trip, a density source and an observer
//   operators : advection operator and CSR matrices
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "naga.hpp"

using namespace naga::fluids::nonlocal_lbm;
using lattice = d2q9_lattice<float>;

static int errors() {
    simulation_engine<lattice> engine;
    simulation_nodes<float> nodes;
    nodes.num_bulk_points = 3;
    nodes.points.resize(2);  // inconsistent on purpose: no parameters set yet
    try {
        engine.init_domain(nodes);
        std::puts("FAIL: init_domain without parameters did not throw");
        return 1;
    } catch (const std::invalid_argument& e) {
        if (std::strncmp(e.what(), "naga::nonlocal_lbm::simulation_engine::init_domain: ", 52)) {
            std::printf("FAIL: unexpected message %s\n", e.what());
            return 1;
        }
    }
    try {
        engine.step_forward();
        std::puts("FAIL: step_forward before init_domain did not throw");
        return 1;
    } catch (const std::runtime_error&) {
    }
    std::puts("errors ok");
    return 0;
}

static int run(int steps) {
    const double h = 0.02;
    uniform_grid_provider<lattice> provider({0.0, 0.0}, {1.0, 1.0}, h, 0.1, 0.5, 0.25);
    simulation_engine<lattice> engine;
    const double cs = 0.57735026918962576;
    const double dt = 0.4 * h;
    // tau = 0.6 on D2Q9: nu_lat = (tau - 1/2) dt / 3 with c = c_s
    engine.set_problem_parameters(float((0.6 - 0.5) * dt / 3.0), 1.0f, float(dt), float(cs));
    engine.init_domain(provider);
    engine.set_initial_pulse({0.5, 0.5}, 1e-3, 0.05);
    for (int n = 0; n < steps; ++n) engine.step_forward();
    const auto& sol = engine.solution();
    double sum = 0.0, umax = 0.0;
    for (float r : sol.macroscopic_values.fluid_density) sum += double(r) - 1.0;
    for (float u : sol.macroscopic_values.fluid_velocity) umax = std::fmax(umax, std::fabs(u));
    std::printf("sum_drho %.9e max_u %.9e time %.9e nodes %zu\n", sum, umax,
                double(engine.time()), sol.macroscopic_values.fluid_density.size());
    return 0;
}

// save_state/load_state round trip, a density source and an observer
static int extras() {
    const double h = 0.025;
    uniform_grid_provider<lattice> provider({0.0, 0.0}, {1.0, 1.0}, h, 0.1, 0.5, 0.25);
    simulation_engine<lattice> engine;
    const double dt = 0.4 * h;
    engine.set_problem_parameters(float((0.6 - 0.5) * dt / 3.0), 1.0f, float(dt),
                                  float(0.57735026918962576));
    engine.init_domain(provider);
    density_source_spec src;
    src.center = {0.5, 0.5, 0.0};
    src.radius = 0.1;
    src.amplitude = 1e-3;
    src.frequency = 5.0;
    engine.register_density_source(src);
    density_observer<lattice> probe(engine, {0.6f, 0.5f});
    for (int n = 0; n < 5; ++n) {
        engine.step_forward();
        probe.update();
    }
    const std::vector<char> state = engine.save_state();
    for (int n = 0; n < 5; ++n) engine.step_forward();
    const std::vector<float> a = engine.solution().macroscopic_values.fluid_density;
    engine.load_state(state);
    for (int n = 0; n < 5; ++n) engine.step_forward();
    const std::vector<float> b = engine.solution().macroscopic_values.fluid_density;
    if (a != b) {
        std::puts("FAIL: resumed run differs");
        return 1;
    }
    if (probe.density_values().size() != 5 || probe.time_values()[1] <= probe.time_values()[0]) {
        std::puts("FAIL: observer");
        return 1;
    }
    double dev = 0.0;
    for (float v : probe.density_values()) dev = std::fmax(dev, std::fabs(v - 1.0));
    std::printf("extras ok max_probe_dev %.3e\n", dev);
    return dev > 0.0 ? 0 : 1;
}

// advection_operator built twice (from points, from its divergence
// operator) gives identical steps; a CSR matrix from the stencil with the
// x-weights reproduces divergence_operator::apply on the field (f, 0)
static int operators() {
    using naga::nonlocal_calculus::advection_operator;
    using naga::nonlocal_calculus::constant_velocity_field;
    using naga::nonlocal_calculus::divergence_operator;
    const int m = 41;
    std::vector<float> pts;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            const float jx = 0.2f * std::sin(float(7 * i + 13 * j)), jy = 0.2f * std::cos(float(5 * i + 3 * j));
            pts.push_back((float(i) + jx) / float(m - 1));
            pts.push_back((float(j) + jy) / float(m - 1));
        }
    const size_t n = pts.size() / 2;
    auto adv = advection_operator<float, 2>::create(pts, 16);
    auto again = advection_operator<float, 2>::create(adv.divergence_op());
    std::vector<float> f0(n), fa, fb;
    for (size_t i = 0; i < n; ++i)
        f0[i] = std::exp(-((pts[2 * i] - 0.4f) * (pts[2 * i] - 0.4f) +
                           (pts[2 * i + 1] - 0.5f) * (pts[2 * i + 1] - 0.5f)) / 0.01f);
    const float v[2] = {0.5f, 0.0f};
    auto vel = constant_velocity_field<float, 2>::create(v);
    adv.step_forward(vel, f0, fa, 0.005f).get();
    again.step_forward(vel, f0, fb, 0.005f).get();
    if (fa != fb) {
        std::puts("FAIL: advection operators differ");
        return 1;
    }
    const auto div = adv.divergence_op();
    const uint32_t k = div.support_size();
    std::vector<float> wx(n * k), field(2 * n, 0.0f), viaop, viacsr;
    for (size_t t = 0; t < n * k; ++t) wx[t] = div.weights()[2 * t];
    for (size_t i = 0; i < n; ++i) field[2 * i] = f0[i];
    div.apply(field, viaop);
    auto A = naga::linalg::sparse_csr_matrix::create_from_index_stencil(div.support_indices(), wx,
                                                                        int32_t(k));
    naga::linalg::matrix_mult(A, f0, viacsr);
    double err = 0.0, scale = 0.0;
    for (size_t i = 0; i < n; ++i) {
        err = std::fmax(err, std::fabs(double(viaop[i]) - viacsr[i]));
        scale = std::fmax(scale, std::fabs(double(viaop[i])));
    }
    double moved = 0.0;
    for (size_t i = 0; i < n; ++i) moved = std::fmax(moved, std::fabs(double(fa[i]) - f0[i]));
    std::printf("operators ok csr_vs_apply %.3e moved %.3e\n", err / scale, moved);
    return err <= 1e-5 * scale && moved > 0.0 ? 0 : 1;
}

// reference-named entry points: segmentation and kNN over it, the batched
// inverse, the RPIM interpolator (linear reproduction), divergence_operator::
// create(domain, query, override_query_size), attached observers and the VTK
// output (argv[2]: a directory for the .vtp files)
static int dropin(const char* dir) {
    using naga::segmentation::nd_cubic_segmentation;
    std::vector<float> pts;
    for (int i = 0; i < 2000; ++i) {
        pts.push_back(std::fmod(0.618034f * float(i), 1.0f));
        pts.push_back(std::fmod(0.754878f * float(i) + 0.1f, 1.0f));
    }
    nd_cubic_segmentation<float, 2> seg(pts, 32u);
    std::vector<float> q(pts.begin(), pts.begin() + 200), d2a, d2b;
    std::vector<uint32_t> ia, ib;
    naga::segmentation::batched_nearest_neighbors<float, 2>(16, q, seg, d2a, ia);
    naga::segmentation::batched_nearest_neighbors<2>(16, pts, q, d2b, ib);
    if (ia != ib || d2a != d2b) {
        std::puts("FAIL: kNN over the segmentation differs");
        return 1;
    }
    std::vector<uint32_t> sizes, offs, idx;
    seg.partitions(sizes, offs, idx);
    size_t total = 0;
    for (uint32_t s : sizes) total += s;
    if (total != 2000 || seg.partition_count() != sizes.size()) {
        std::puts("FAIL: partitions");
        return 1;
    }
    const std::vector<double> A = {2, 1, 0, 1, 3, 1, 0, 1, 4};
    std::vector<double> Ainv;
    naga::linalg::batched_matrix_inverse(3, 1, A, Ainv);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s += A[i + 3 * m] * Ainv[m + 3 * j];
            if (std::fabs(s - (i == j)) > 1e-12) {
                std::puts("FAIL: inverse");
                return 1;
            }
        }
    // linear field interpolated exactly (the P-augmentation)
    auto r = naga::interpolation::radial_point_method<float>::create_interpolator<2>(
        pts, ia, 16, q, 0.02f);
    std::vector<float> field(2000), out;
    for (int i = 0; i < 2000; ++i) field[i] = 0.3f + 2.0f * pts[2 * i] - pts[2 * i + 1];
    r.interpolate(field, out);
    for (size_t i = 0; i < out.size(); ++i)
        if (std::fabs(out[i] - (0.3f + 2.0f * q[2 * i] - q[2 * i + 1])) > 1e-4f) {
            std::puts("FAIL: rpim linear reproduction");
            return 1;
        }
    auto op = naga::nonlocal_calculus::divergence_operator<float, 2>::create(pts, q, 150, 16);
    if (op.size() != 150 || op.weights()[size_t(120) * 16 * 2] != 0.0f) {
        std::puts("FAIL: divergence override size");
        return 1;
    }
    const double h = 0.05;
    uniform_grid_provider<lattice> provider({0.0, 0.0}, {1.0, 1.0}, h, 0.1, 0.5, 0.25);
    simulation_engine<lattice> engine;
    const double dt = 0.4 * h;
    engine.set_problem_parameters(float((0.6 - 0.5) * dt / 3.0), 1.0f, float(dt),
                                  float(0.57735026918962576));
    engine.init_domain(provider);
    engine.set_initial_pulse({0.5, 0.5}, 1e-3, 0.1);
    vtk_observer<lattice> vtk(engine, dir);
    density_observer<lattice> probe(engine, {0.5f, 0.5f});
    engine.attach_observer(vtk);
    engine.attach_observer(probe);
    engine.step_forward(2);
    engine.detach_observer(vtk);
    engine.step_forward(1);
    if (probe.density_values().size() != 3 || std::fabs(probe.density_values()[0] - 1.0f) < 1e-5f) {
        std::puts("FAIL: attached density observer");
        return 1;
    }
    for (int f = 0; f < 2; ++f) {
        const std::string path = std::string(dir) + "/result." + std::to_string(f) + ".vtp";
        FILE* fp = std::fopen(path.c_str(), "rb");
        if (!fp) {
            std::printf("FAIL: missing %s\n", path.c_str());
            return 1;
        }
        std::fclose(fp);
    }
    std::puts("dropin ok");
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc > 1 && !std::strcmp(argv[1], "run")) return run(argc > 2 ? std::atoi(argv[2]) : 10);
        if (argc > 1 && !std::strcmp(argv[1], "extras")) return extras();
        if (argc > 1 && !std::strcmp(argv[1], "operators")) return operators();
        if (argc > 2 && !std::strcmp(argv[1], "dropin")) return dropin(argv[2]);
        return errors();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
}
