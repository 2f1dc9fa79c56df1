// extern "C" entry points (include/naga_amd.h): argument checks, exception
// to status translation, thread-local error text.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/naga_amd.h"
#include "calculus.hpp"
#include "engine.hpp"
#include "operators.hpp"
#include "server.hpp"
#include "setup_kernels.hpp"

using namespace naga_amd;

struct naga_engine {
    std::unique_ptr<Engine> impl;
};

struct naga_batch {
    std::unique_ptr<Batch> impl;
};

struct naga_server {
    std::unique_ptr<Server> impl;
};

struct naga_advection {
    std::unique_ptr<Advection> impl;
};

struct naga_csr {
    std::unique_ptr<Csr> impl;
};

struct naga_segmentation {
    std::unique_ptr<Segmentation> impl;
};

struct naga_rpim {
    std::unique_ptr<Rpim> impl;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return NAGA_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return NAGA_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return NAGA_DEVICE_ERROR;
    }
}

Engine& engine_of(naga_engine* e) {
    require(e != nullptr && e->impl != nullptr, "null engine handle");
    return *e->impl;
}

Batch& batch_of(naga_batch* b) {
    require(b != nullptr && b->impl != nullptr, "null batch handle");
    return *b->impl;
}

Server& server_of(naga_server* s) {
    require(s != nullptr && s->impl != nullptr, "null server handle");
    return *s->impl;
}

// ---- synthetic uniform grid (node_providers/uniform_grid_provider.cuh) ----

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t morton(int D, const int64_t* g) {
    uint64_t key = 0;
    const int bits = D == 2 ? 32 : 21;
    for (int b = 0; b < bits; ++b)
        for (int d = 0; d < D; ++d) key |= ((uint64_t(g[d]) >> b) & 1ull) << (b * D + d);
    return key;
}

// Hilbert-curve key of grid point g (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004: axes to transposed Hilbert index, then
// the transposed bits interleaved most significant first) on the curve of
// 2^21 (3D) or 2^32 (2D) points per side.
template <int D>
uint64_t hilbert_t(const int64_t* g) {
    constexpr int bits = D == 2 ? 32 : 21;
    uint32_t X[3] = {uint32_t(g[0]), uint32_t(g[1]), D == 3 ? uint32_t(g[2]) : 0u};
    // the levels above the coordinates' highest set bit only permute the
    // axes (each swaps X[0] with X[1], then with X[2]): applied at once
    const uint32_t mx = X[0] | X[1] | X[2];
    const int b = mx ? 32 - __builtin_clz(mx) : 1;
    const int lv = bits - b;
    if (D == 2) {
        if (lv & 1) std::swap(X[0], X[1]);
    } else if (lv % 3 == 1) {
        const uint32_t x0 = X[0];
        X[0] = X[2];
        X[2] = X[1];
        X[1] = x0;
    } else if (lv % 3 == 2) {
        const uint32_t x0 = X[0];
        X[0] = X[1];
        X[1] = X[2];
        X[2] = x0;
    }
    for (uint32_t Q = 1u << (b - 1); Q > 1; Q >>= 1) {  // inverse undo
        const uint32_t P = Q - 1;
        for (int i = 0; i < D; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < D; ++i) X[i] ^= X[i - 1];  // Gray encode
    uint32_t t = 0;
    for (uint32_t Q = 1u << (b - 1); Q > 1; Q >>= 1)
        if (X[D - 1] & Q) t ^= Q - 1;
    for (int i = 0; i < D; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int l = b - 1; l >= 0; --l)
        for (int i = 0; i < D; ++i) key = (key << 1) | ((X[i] >> l) & 1u);
    return key;
}
uint64_t hilbert(int D, const int64_t* g) { return D == 2 ? hilbert_t<2>(g) : hilbert_t<3>(g); }

// Sort key of the node providers within each node class: Hilbert order
// (consecutive nodes are grid neighbours, so a tile of consecutive nodes is
// compact and its stencils' halo is shared with the tiles next to it in
// the launch), or Morton order with NAGA_NODE_ORDER=morton (read once per
// process; the oracle orders by the Hilbert key).
uint64_t node_key(int D, const int64_t* g) {
    static const bool use_morton = [] {
        const char* e = std::getenv("NAGA_NODE_ORDER");
        return e && std::string(e) == "morton";
    }();
    return use_morton ? morton(D, g) : hilbert(D, g);
}

// ranges/uniform_grid.cuh:55-65, size per dimension ceil(extent/h) + 1 in FP32
int64_t grid_dims(int D, const double* lo, const double* hi, double h, int64_t* dims) {
    int64_t n = 1;
    for (int d = 0; d < D; ++d) {
        const float ext = float(hi[d]) - float(lo[d]);
        dims[d] = int64_t(std::ceil(ext / float(h))) + 1;
        n *= dims[d];
    }
    return n;
}

}  // namespace

extern "C" {

const char* naga_last_error(void) { return g_last_error.c_str(); }

int naga_lattice_tables(int32_t q, double* c, double* w, int32_t* bb, double* K) {
    return guarded([&] {
        require(q == 9 || q == 27, "lattice must be 9 or 27");
        const int D = q == 9 ? 2 : 3;
        for (int a = 0; a < q; ++a) {
            for (int d = 0; d < D; ++d)
                if (c) c[a * D + d] = q == 9 ? Lattice<9>::c(a, d) : Lattice<27>::c(a, d);
            if (w) w[a] = q == 9 ? Lattice<9>::w(a) : Lattice<27>::w(a);
            if (bb) bb[a] = q == 9 ? Lattice<9>::bb(a) : Lattice<27>::bb(a);
            for (int j = 0; j < q; ++j)
                if (K) K[a * q + j] = collision_matrix_entry(q, a, j);
        }
    });
}

int naga_collide(int32_t q, const double* f, double rho, const double* u, double nu_lat,
                 double dt_lat, double* df) {
    return guarded([&] {
        require(q == 9 || q == 27, "lattice must be 9 or 27");
        require(f && u && df, "null argument");
        const CollisionParams p = make_collision_params(q, nu_lat, dt_lat);
        if (q == 27)
            collide<27>(f, rho, u, p, df);
        else
            collide<9>(f, rho, u, p, df);
    });
}

int naga_engine_create(const naga_scene_config* cfg, naga_engine** out) {
    return guarded([&] {
        require(cfg != nullptr && out != nullptr, "null argument");
        SceneConfig c;
        c.Q = cfg->lattice;
        c.k = cfg->k;
        c.kb = cfg->k_boundary;
        c.rbf = cfg->rbf;
        c.rbf_boundary = cfg->rbf_boundary;
        c.pml_ref = cfg->pml_ref_index;
        c.k_min = cfg->k_min;
        c.nu = cfg->nu;
        c.rho0 = cfg->rho0;
        c.dt = cfg->dt;
        c.c = cfg->c;
        c.h = cfg->h;
        c.filter = cfg->filter_strength;
        require(cfg->solve_precision == 0 || cfg->solve_precision == 32 ||
                    cfg->solve_precision == 64,
                "solve_precision must be 0, 32 or 64");
        c.solve_bits = cfg->solve_precision == 32 ? 32 : 64;
        auto* e = new naga_engine{};
        try {
            e->impl = std::make_unique<Engine>(c);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int naga_engine_destroy(naga_engine* e) {
    return guarded([&] { delete e; });
}

int naga_engine_init_domain(naga_engine* e, const float* points, int64_t n_bulk,
                            int64_t n_layer, int64_t n_boundary, const float* normals,
                            const float* layer_absorption) {
    return guarded([&] {
        engine_of(e).init_domain(points, n_bulk, n_layer, n_boundary, normals,
                                 layer_absorption);
    });
}

int naga_engine_get_stencil_sizes(naga_engine* e, uint8_t* k) {
    return guarded([&] { engine_of(e).get_stencil_sizes(k); });
}

int naga_engine_set_pulse(naga_engine* e, const double* center, double amp, double width) {
    return guarded([&] {
        require(center != nullptr, "null center");
        engine_of(e).set_pulse(center, amp, width);
    });
}

int naga_engine_add_density_source(naga_engine* e, const double* center, double radius,
                                   double amplitude, double frequency, double phase,
                                   double t_start, double t_end) {
    return guarded([&] {
        require(center != nullptr, "null center");
        DensitySource s;
        Engine& eng = engine_of(e);
        for (int d = 0; d < eng.dims(); ++d) s.center[d] = center[d];
        s.radius = radius;
        s.amplitude = amplitude;
        s.frequency = frequency;
        s.phase = phase;
        s.t_start = t_start;
        s.t_end = t_end;
        eng.add_density_source(s);
    });
}

int naga_engine_clear_density_sources(naga_engine* e) {
    return guarded([&] { engine_of(e).clear_density_sources(); });
}

int naga_batch_create(naga_batch** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        auto* b = new naga_batch;
        try {
            b->impl = std::make_unique<Batch>();
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int naga_batch_destroy(naga_batch* b) {
    return guarded([&] { delete b; });
}

int naga_batch_add(naga_batch* b, naga_engine* e) {
    return guarded([&] { batch_of(b).add(&engine_of(e)); });
}

int naga_batch_remove(naga_batch* b, naga_engine* e) {
    return guarded([&] { batch_of(b).remove(&engine_of(e)); });
}

int naga_batch_set_graph(naga_batch* b, int32_t on) {
    return guarded([&] { batch_of(b).set_graph(on != 0); });
}

int naga_batch_autotune(naga_batch* b, int32_t trial_steps, int32_t* gather_batch) {
    return guarded([&] {
        const int gb = batch_of(b).autotune(trial_steps);
        if (gather_batch) *gather_batch = gb;
    });
}

int naga_batch_set_q9_kernels(naga_batch* b, int32_t lanes, int32_t unions) {
    return guarded([&] { batch_of(b).set_q9_kernels(lanes, unions != 0); });
}

int naga_batch_tuning(naga_batch* b, int32_t* gather_batch, int32_t* q9_lanes,
                      int32_t* q9_unions, int32_t* graph) {
    return guarded([&] {
        int gb = 0, lanes = 0, un = 0, gr = 0;
        batch_of(b).tuning(&gb, &lanes, &un, &gr);
        if (gather_batch) *gather_batch = gb;
        if (q9_lanes) *q9_lanes = lanes;
        if (q9_unions) *q9_unions = un;
        if (graph) *graph = gr;
    });
}

int naga_engine_add_probes(naga_engine* e, const float* points, int64_t n, int32_t rbf,
                           int32_t* probe_set) {
    return guarded([&] {
        const int id = engine_of(e).add_probes(points, n, rbf);
        if (probe_set) *probe_set = id;
    });
}

int naga_engine_sample_probes(naga_engine* e, int32_t probe_set, int32_t fresh,
                              double* values) {
    return guarded([&] { engine_of(e).sample_probes(probe_set, fresh != 0, values); });
}

int naga_engine_state_size(naga_engine* e, int64_t* bytes) {
    return guarded([&] {
        require(bytes != nullptr, "null argument");
        *bytes = engine_of(e).state_bytes();
    });
}

int naga_engine_save_state(naga_engine* e, void* buf, int64_t capacity) {
    return guarded([&] { engine_of(e).save_state(buf, capacity); });
}

int naga_engine_load_state(naga_engine* e, const void* buf, int64_t bytes) {
    return guarded([&] { engine_of(e).load_state(buf, bytes); });
}

int naga_batch_step(naga_batch* b, int32_t nsteps) {
    return guarded([&] { batch_of(b).step(nsteps); });
}

int naga_batch_step_timed(naga_batch* b, int32_t nsteps, double* ms) {
    return guarded([&] {
        const double v = batch_of(b).step_timed(nsteps);
        if (ms) *ms = v;
    });
}

int naga_server_create(int64_t hbm_budget_bytes, int32_t max_active, int32_t chunk_steps,
                       int32_t use_graph, int32_t setup_threads, naga_server** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        require(hbm_budget_bytes > 0, "budget must be positive");
        auto* s = new naga_server;
        try {
            s->impl = std::make_unique<Server>(size_t(hbm_budget_bytes), max_active, chunk_steps,
                                               use_graph != 0, setup_threads);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int naga_server_destroy(naga_server* s) {
    return guarded([&] { delete s; });
}

int naga_server_submit(naga_server* s, const naga_request* req, int64_t* id) {
    return guarded([&] {
        require(req != nullptr && id != nullptr, "null argument");
        *id = server_of(s).submit(*req);
    });
}

int naga_server_status(naga_server* s, int64_t id, int32_t* state) {
    return guarded([&] {
        require(state != nullptr, "null argument");
        const int st = server_of(s).status(id);
        require(st >= 0, "unknown request id");
        *state = st;
    });
}

int naga_server_result(naga_server* s, int64_t id, int64_t* n_nodes, float* rho, float* u,
                       double* drho, double* latency_ms) {
    return guarded([&] {
        ServedResult r;
        require(server_of(s).result(id, r), "request is not done (or unknown)");
        if (!r.ok) fail(kNumericalError, "request failed: " + r.error);
        if (n_nodes) *n_nodes = r.nodes;
        if (rho) std::memcpy(rho, r.rho.data(), sizeof(float) * r.rho.size());
        if (u) std::memcpy(u, r.u.data(), sizeof(float) * r.u.size());
        if (drho) std::memcpy(drho, r.drho.data(), sizeof(double) * r.drho.size());
        if (latency_ms) *latency_ms = r.latency_ms;
    });
}

int naga_server_release(naga_server* s, int64_t id) {
    return guarded([&] { server_of(s).release(id); });
}

int naga_server_drain(naga_server* s) {
    return guarded([&] { server_of(s).drain(); });
}

int naga_server_get_stats(naga_server* s, naga_server_stats* out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = server_of(s).stats();
    });
}

int naga_engine_set_distributions(naga_engine* e, const double* f) {
    return guarded([&] {
        require(f != nullptr, "null distributions");
        engine_of(e).set_distributions(f);
    });
}

int naga_engine_step(naga_engine* e, int32_t nsteps) {
    return guarded([&] { engine_of(e).step(nsteps); });
}

int naga_engine_step_timed(naga_engine* e, int32_t nsteps, double* ms) {
    return guarded([&] {
        const double t = engine_of(e).step_timed(nsteps);
        if (ms) *ms = t;
    });
}

int naga_engine_profile_step(naga_engine* e, double* ms) {
    return guarded([&] {
        require(ms != nullptr, "null output");
        engine_of(e).profile_step(ms);
    });
}

int naga_engine_get_macroscopic(naga_engine* e, float* rho, float* u, double* drho,
                                int32_t fresh) {
    return guarded([&] { engine_of(e).get_macroscopic(rho, u, drho, fresh != 0); });
}

int naga_engine_get_distributions(naga_engine* e, double* f, double* g) {
    return guarded([&] { engine_of(e).get_distributions(f, g); });
}

int naga_engine_get_setup(naga_engine* e, uint32_t* idx, float* W, float* filt,
                          uint32_t* bidx, float* bw) {
    return guarded([&] { engine_of(e).get_setup(idx, W, filt, bidx, bw); });
}

int naga_engine_get_pml(naga_engine* e, double* Q1_minus_w, double* Q2) {
    return guarded([&] { engine_of(e).get_pml(Q1_minus_w, Q2); });
}

int naga_engine_info(naga_engine* e, int64_t* n_nodes, double* time, int64_t* frame,
                     double* setup_ms) {
    return guarded([&] {
        Engine& en = engine_of(e);
        if (n_nodes) *n_nodes = en.nodes();
        if (time) *time = en.time();
        if (frame) *frame = en.frame();
        if (setup_ms) *setup_ms = en.setup_ms();
    });
}

int naga_engine_write_vtp(naga_engine* e, const char* path, int32_t fresh) {
    return guarded([&] { engine_of(e).write_vtp(path, fresh != 0); });
}

int naga_engine_setup_phases(naga_engine* e, double* ms) {
    return guarded([&] {
        require(ms != nullptr, "null argument");
        const double* p = engine_of(e).setup_phase_ms();
        for (int i = 0; i < Engine::kSetupPhases; ++i) ms[i] = p[i];
    });
}

int64_t naga_uniform_grid_size(int32_t dims, const double* lo, const double* hi, double h) {
    if ((dims != 2 && dims != 3) || !lo || !hi || !(h > 0)) return -1;
    int64_t d[3];
    return grid_dims(dims, lo, hi, h, d);
}

int naga_uniform_grid(int32_t dims, const double* lo, const double* hi, double h,
                      double jitter_frac, uint64_t seed, double layer_thickness,
                      double peak_absorption, float* points, float* normals, float* sigma,
                      int64_t* counts) {
    return guarded([&] {
        const int D = dims;
        require(D == 2 || D == 3, "dims must be 2 or 3");
        require(lo && hi && points && normals && sigma && counts, "null argument");
        require(h > 0, "grid spacing must be positive");
        int64_t gd[3];
        const int64_t N = grid_dims(D, lo, hi, h, gd);
        double sp[3];
        for (int d = 0; d < D; ++d) {
            require(gd[d] >= 2, "grid needs at least two nodes per dimension");
            sp[d] = (hi[d] - lo[d]) / double(gd[d] - 1);
        }
        // grid points with jittered interior nodes (uniform_grid.cuh:89-97)
        std::vector<float> xyz(N * D);
        std::vector<uint8_t> isb(N);
        uint64_t st = seed;
        int64_t nb = 0;
        for (int64_t g = 0; g < N; ++g) {
            int64_t gi[3], rem = g;
            bool bnd = false;
            for (int d = 0; d < D; ++d) {
                gi[d] = rem % gd[d];
                rem /= gd[d];
                bnd = bnd || gi[d] == 0 || gi[d] == gd[d] - 1;  // uniform_grid.cuh:242-253
            }
            isb[g] = bnd;
            nb += bnd;
            for (int d = 0; d < D; ++d) {
                const double u01 = double(splitmix64(st) >> 11) * 0x1.0p-53;
                double x = lo[d] + double(gi[d]) * sp[d];
                if (!bnd) x += (2.0 * u01 - 1.0) * jitter_frac * h;
                xyz[g * D + d] = float(x);
            }
        }
        const int64_t ni = N - nb;
        std::vector<float> bpts, ipts;
        std::vector<int64_t> bg, ig;
        bpts.reserve(nb * D);
        ipts.reserve(ni * D);
        for (int64_t g = 0; g < N; ++g) {
            auto& dst = isb[g] ? bpts : ipts;
            dst.insert(dst.end(), xyz.begin() + g * D, xyz.begin() + (g + 1) * D);
            (isb[g] ? bg : ig).push_back(g);
        }
        // layer: exact 1-NN squared distance to the boundary nodes on the GPU
        // (uniform_grid_provider.cuh:104-127); layer iff d2 <= L^2 in FP32
        std::vector<float> d2(std::max<int64_t>(ni, 1), 0.0f);
        std::vector<uint32_t> nn(std::max<int64_t>(ni, 1), 0u);
        const bool layered = layer_thickness > 0 && ni > 0 && nb > 0;
        if (layered) {
            hipStream_t s;
            NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            try {
                knn_host(D, bpts.data(), nb, ipts.data(), ni, 1, d2.data(), nn.data(), s);
            } catch (...) {
                (void)hipStreamDestroy(s);
                throw;
            }
            (void)hipStreamDestroy(s);
        }
        const float Lf = float(layer_thickness);
        const float L2 = Lf * Lf;
        auto gridkey = [&](int64_t g) {
            int64_t gi[3], rem = g;
            for (int d = 0; d < D; ++d) {
                gi[d] = rem % gd[d];
                rem /= gd[d];
            }
            return node_key(D, gi);
        };
        std::vector<std::pair<uint64_t, int64_t>> kb, kl, kbd;
        for (int64_t t = 0; t < ni; ++t) {
            auto e = std::make_pair(gridkey(ig[t]), t);
            (layered && d2[t] <= L2 ? kl : kb).push_back(e);
        }
        for (int64_t t = 0; t < nb; ++t) kbd.emplace_back(gridkey(bg[t]), t);
        std::sort(kb.begin(), kb.end());
        std::sort(kl.begin(), kl.end());
        std::sort(kbd.begin(), kbd.end());
        int64_t o = 0;
        for (auto& e : kb) std::memcpy(points + D * o++, &ipts[e.second * D], sizeof(float) * D);
        int64_t t = 0;
        for (auto& e : kl) {
            std::memcpy(points + D * o++, &ipts[e.second * D], sizeof(float) * D);
            // detail/uniform_grid_provider.inl:148-159 in FP32
            const float dist = std::sqrt(d2[e.second]);
            const float r = (Lf - dist) / Lf;
            sigma[t++] = float(peak_absorption) * (r * r);
        }
        t = 0;
        for (auto& e : kbd) {
            std::memcpy(points + D * o++, &bpts[e.second * D], sizeof(float) * D);
            int64_t gi[3], rem = bg[e.second];
            for (int d = 0; d < D; ++d) {
                gi[d] = rem % gd[d];
                rem /= gd[d];
            }
            // uniform_grid.cuh:255-293, negated to point inward
            // (uniform_grid_provider.cuh:89-94)
            float nv[3] = {0.f, 0.f, 0.f};
            for (int d = 0; d < D; ++d) nv[d] = gi[d] == 0 ? -1.f : gi[d] == gd[d] - 1 ? 1.f : 0.f;
            float n2 = nv[0] * nv[0];
            for (int d = 1; d < D; ++d) n2 = nv[d] * nv[d] + n2;
            const float nrm = std::sqrt(n2);
            for (int d = 0; d < D; ++d) normals[t * D + d] = -(nv[d] / (nrm == 0.f ? 1.f : nrm));
            ++t;
        }
        counts[0] = int64_t(kb.size());
        counts[1] = int64_t(kl.size());
        counts[2] = nb;
    });
}

int naga_obstacle_grid(int32_t dims, const double* lo, const double* hi, double h,
                       double jitter_frac, uint64_t seed, int32_t n_obstacles,
                       const double* obstacles, double refine_width, double layer_thickness,
                       double peak_absorption, double immersed_layer_thickness,
                       double immersed_absorption, int64_t capacity, float* points,
                       float* normals, float* sigma, int64_t* counts) {
    return guarded([&] {
        const int D = dims;
        require(D == 2 || D == 3, "dims must be 2 or 3");
        require(immersed_layer_thickness >= 0.0 && immersed_absorption >= 0.0,
                "immersed layer thickness and absorption must be >= 0");
        require(lo && hi && counts, "null argument");
        require(h > 0, "grid spacing must be positive");
        require(n_obstacles >= 0 && (n_obstacles == 0 || obstacles), "bad obstacle list");
        for (int j = 0; j < n_obstacles; ++j)
            require(obstacles[j * (D + 1) + D] > 0, "obstacle radius must be positive");
        int64_t gd[3];
        const int64_t N = grid_dims(D, lo, hi, h, gd);
        double sp[3];
        for (int d = 0; d < D; ++d) {
            require(gd[d] >= 2, "grid needs at least two nodes per dimension");
            sp[d] = (hi[d] - lo[d]) / double(gd[d] - 1);
        }
        auto inside = [&](const double* x, double pad) {
            for (int j = 0; j < n_obstacles; ++j) {
                const double* o = obstacles + j * (D + 1);
                double r2 = 0.0;
                for (int d = 0; d < D; ++d) r2 += (x[d] - o[d]) * (x[d] - o[d]);
                if (std::sqrt(r2) < o[D] + pad) return true;
            }
            return false;
        };
        auto qkey = [&](const double* x) {
            int64_t gi[3];
            for (int d = 0; d < D; ++d) {
                const double t = std::round((x[d] - lo[d]) / sp[d]);
                gi[d] = int64_t(std::min(std::max(t, 0.0), double(gd[d] - 1)));
            }
            return node_key(D, gi);
        };
        struct Node {
            uint64_t key;
            int64_t order;
            float x[3];
            float n[3];
        };
        std::vector<Node> bulk, layer, bnd;
        std::vector<float> sig;
        const double L = layer_thickness, Li = immersed_layer_thickness;
        // Absorbing layers (conforming_point_cloud_provider.cuh:219-280): a
        // fluid node within L of the box faces (distance d) gets
        // sigma = peak_absorption (1 - d/L)^2; one within Li of an obstacle
        // surface and nearer to it than to the faces gets
        // immersed_absorption (1 - d/Li)^2 from its nearest obstacle.
        auto classify = [&](const Node& nd, const double* x) {
            double dbox = 1e300;
            for (int d = 0; d < D; ++d) dbox = std::min({dbox, x[d] - lo[d], hi[d] - x[d]});
            double cur = 1e300, s = 0.0;
            bool in_layer = false;
            if (L > 0 && dbox <= L) {
                cur = dbox;
                const double r = (L - std::max(dbox, 0.0)) / L;
                s = peak_absorption * r * r;
                in_layer = true;
            }
            if (Li > 0 && immersed_absorption > 0) {
                double dob = 1e300;
                for (int j = 0; j < n_obstacles; ++j) {
                    const double* o = obstacles + j * (D + 1);
                    double r2 = 0.0;
                    for (int d = 0; d < D; ++d) r2 += (x[d] - o[d]) * (x[d] - o[d]);
                    dob = std::min(dob, std::sqrt(r2) - o[D]);
                }
                if (dob <= Li && dob < cur) {
                    const double r = (Li - std::max(dob, 0.0)) / Li;
                    s = immersed_absorption * r * r;
                    in_layer = true;
                }
            }
            if (in_layer) {
                layer.push_back(nd);
                sig.push_back(float(s));
            } else {
                bulk.push_back(nd);
            }
        };
        // grid nodes: the uniform grid's stream and jitter
        uint64_t st = seed;
        for (int64_t g = 0; g < N; ++g) {
            int64_t gi[3], rem = g;
            bool onface = false;
            for (int d = 0; d < D; ++d) {
                gi[d] = rem % gd[d];
                rem /= gd[d];
                onface = onface || gi[d] == 0 || gi[d] == gd[d] - 1;
            }
            double x[3] = {0.0, 0.0, 0.0};
            for (int d = 0; d < D; ++d) {
                const double u01 = double(splitmix64(st) >> 11) * 0x1.0p-53;
                x[d] = lo[d] + double(gi[d]) * sp[d];
                if (!onface) x[d] += (2.0 * u01 - 1.0) * jitter_frac * h;
            }
            if (inside(x, 0.5 * h)) continue;
            Node nd{node_key(D, gi), g, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
            for (int d = 0; d < D; ++d) nd.x[d] = float(x[d]);
            if (onface) {
                float nv[3] = {0.f, 0.f, 0.f}, n2 = 0.f;
                for (int d = 0; d < D; ++d) {
                    nv[d] = gi[d] == 0 ? 1.f : gi[d] == gd[d] - 1 ? -1.f : 0.f;
                    n2 += nv[d] * nv[d];
                }
                for (int d = 0; d < D; ++d) nd.n[d] = nv[d] / std::sqrt(n2);
                bnd.push_back(nd);
                continue;
            }
            classify(nd, x);
        }
        // refinement: grid cells whose centre lies within refine_width of an
        // obstacle surface (outside it) get the 2^D - 1 half-offset nodes,
        // jittered by jitter_frac * h/2 from a second stream
        int64_t order = N;
        if (refine_width > 0 && n_obstacles > 0) {
            uint64_t st2 = seed ^ 0x5DEECE66DA3B1F27ull;
            int64_t cd[3] = {1, 1, 1};
            int64_t ncell = 1;
            for (int d = 0; d < D; ++d) {
                cd[d] = gd[d] - 1;
                ncell *= cd[d];
            }
            for (int64_t c = 0; c < ncell; ++c) {
                int64_t ci[3] = {0, 0, 0}, rem = c;
                for (int d = 0; d < D; ++d) {
                    ci[d] = rem % cd[d];
                    rem /= cd[d];
                }
                double ctr[3] = {0.0, 0.0, 0.0};
                for (int d = 0; d < D; ++d) ctr[d] = lo[d] + (double(ci[d]) + 0.5) * sp[d];
                double gap = 1e300;
                for (int j = 0; j < n_obstacles; ++j) {
                    const double* o = obstacles + j * (D + 1);
                    double r2 = 0.0;
                    for (int d = 0; d < D; ++d) r2 += (ctr[d] - o[d]) * (ctr[d] - o[d]);
                    gap = std::min(gap, std::sqrt(r2) - o[D]);
                }
                if (gap < -0.5 * h || gap > refine_width) continue;
                for (int m = 1; m < (1 << D); ++m) {
                    double x[3] = {0.0, 0.0, 0.0};
                    bool ok = true;
                    for (int d = 0; d < D; ++d) {
                        const double u01 = double(splitmix64(st2) >> 11) * 0x1.0p-53;
                        x[d] = lo[d] + (double(ci[d]) + 0.5 * ((m >> d) & 1)) * sp[d] +
                               (2.0 * u01 - 1.0) * jitter_frac * 0.5 * h;
                        ok = ok && x[d] > lo[d] && x[d] < hi[d];
                    }
                    if (!ok || inside(x, 0.25 * h)) continue;
                    Node nd{qkey(x), order++, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
                    for (int d = 0; d < D; ++d) nd.x[d] = float(x[d]);
                    classify(nd, x);
                }
            }
        }
        // obstacle surfaces (at half spacing when refined), normals into the fluid
        const double hs = refine_width > 0 ? 0.5 * h : h;
        for (int j = 0; j < n_obstacles; ++j) {
            const double* o = obstacles + j * (D + 1);
            const double R = o[D];
            std::vector<std::array<double, 3>> dirs;
            if (D == 2) {
                const int M = std::max(8, int(std::ceil(2.0 * M_PI * R / hs)));
                for (int m = 0; m < M; ++m) {
                    const double t = 2.0 * M_PI * (m + 0.5) / M;
                    dirs.push_back({std::cos(t), std::sin(t), 0.0});
                }
            } else {
                const int M = std::max(16, int(std::ceil(4.0 * M_PI * R * R / (hs * hs))));
                const double ga = M_PI * (3.0 - std::sqrt(5.0));
                for (int m = 0; m < M; ++m) {
                    const double z = 1.0 - 2.0 * (m + 0.5) / M, rr = std::sqrt(1.0 - z * z);
                    dirs.push_back({rr * std::cos(ga * m), rr * std::sin(ga * m), z});
                }
            }
            for (const auto& v : dirs) {
                double x[3] = {0.0, 0.0, 0.0};
                bool ok = true;
                for (int d = 0; d < D; ++d) {
                    x[d] = o[d] + R * v[d];
                    ok = ok && x[d] > lo[d] && x[d] < hi[d];
                }
                // points inside another obstacle are not on the fluid boundary
                for (int i = 0; ok && i < n_obstacles; ++i) {
                    if (i == j) continue;
                    const double* p = obstacles + i * (D + 1);
                    double r2 = 0.0;
                    for (int d = 0; d < D; ++d) r2 += (x[d] - p[d]) * (x[d] - p[d]);
                    ok = std::sqrt(r2) >= p[D];
                }
                if (!ok) continue;
                Node nd{qkey(x), order++, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
                for (int d = 0; d < D; ++d) {
                    nd.x[d] = float(x[d]);
                    nd.n[d] = float(v[d]);
                }
                bnd.push_back(nd);
            }
        }
        const int64_t total = int64_t(bulk.size() + layer.size() + bnd.size());
        counts[0] = int64_t(bulk.size());
        counts[1] = int64_t(layer.size());
        counts[2] = int64_t(bnd.size());
        if (!points) return;  // size query
        require(capacity >= total, "capacity below the node count (query the size first)");
        require(normals && sigma, "null argument");
        auto by_key = [](const Node& a, const Node& b) {
            return a.key != b.key ? a.key < b.key : a.order < b.order;
        };
        // sort the layer together with its absorption values
        std::vector<int64_t> lperm(layer.size());
        for (size_t i = 0; i < lperm.size(); ++i) lperm[i] = int64_t(i);
        std::sort(lperm.begin(), lperm.end(),
                  [&](int64_t a, int64_t b) { return by_key(layer[a], layer[b]); });
        std::sort(bulk.begin(), bulk.end(), by_key);
        std::sort(bnd.begin(), bnd.end(), by_key);
        int64_t o = 0;
        for (const Node& nd : bulk) std::memcpy(points + D * o++, nd.x, sizeof(float) * D);
        for (size_t i = 0; i < lperm.size(); ++i) {
            std::memcpy(points + D * o++, layer[lperm[i]].x, sizeof(float) * D);
            sigma[i] = sig[lperm[i]];
        }
        int64_t b = 0;
        for (const Node& nd : bnd) {
            std::memcpy(points + D * o++, nd.x, sizeof(float) * D);
            std::memcpy(normals + D * b++, nd.n, sizeof(float) * D);
        }
    });
}

int naga_knn(const float* ref, int64_t n_ref, const float* query, int64_t n_q, int32_t dims,
             int32_t k, float* d2_out, uint32_t* idx_out) {
    return guarded([&] {
        require(ref && query && d2_out && idx_out, "null argument");
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            knn_host(dims, ref, n_ref, query, n_q, k, d2_out, idx_out, s);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

// ---- naga_knn without allocation ----------------------------------------
// The cell binning's scratch (grid, bounds, counts, starts, scan sums,
// sorted indices) for up to kNoAllocMaxRef reference points, one copy per
// GPU.  Host code never uses the array's name as a device address: the
// device address comes from hipGetSymbolAddress, once per device.
namespace {
constexpr int64_t kNoAllocMaxRef = NAGA_KNN_NOALLOC_MAX_REF;
constexpr size_t kNoAllocBytes = bin_layout(kNoAllocMaxRef).total;
}  // namespace
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
std::array<void*, 64> g_scratch_ptr{};

void* scratch_of_current_device() {
    int dev = 0;
    NAGA_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < int(g_scratch_ptr.size()), "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        NAGA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "no device address for the kNN scratch");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int naga_knn_noalloc(const float* ref, int64_t n_ref, const float* query, int64_t n_q,
                     int32_t dims, int32_t k, float* d2_out, uint32_t* idx_out, void* stream) {
    return guarded([&] {
        require(dims == 2 || dims == 3, "dimension must be 2 or 3");
        require(k >= 1 && k <= 64, "k must be in [1, 64]");
        require(n_ref >= k, "fewer reference points than neighbours requested");
        require(n_ref <= kNoAllocMaxRef,
                "n_ref exceeds the no-allocation limit of " + std::to_string(kNoAllocMaxRef) +
                    " reference points (use naga_knn)");
        require(n_q >= 0, "negative query count");
        require(ref && (n_q == 0 || (query && d2_out && idx_out)), "null argument");
        hipStream_t st = static_cast<hipStream_t>(stream);
        void* mem = scratch_of_current_device();
        const CellBins bins = launch_bin_points(dims, ref, n_ref, 2.0, 0.0, mem, kNoAllocBytes, st);
        launch_knn(bins, ref, query, n_q, k, d2_out, idx_out, st);
    });
}

int naga_batched_matrix_inverse_f64(const double* A, int32_t n, int64_t batch, double* A_inv,
                                    int32_t* info) {
    return guarded([&] {
        require(A && A_inv, "null argument");
        require(n >= 1 && n <= 64, "n must be in [1, 64]");
        require(batch >= 0 && batch < int64_t(INT32_MAX),
                "Batch size must be less than std::numeric_limits<int>::max()");
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            // A X = I per item, by the batched LU solve
            std::vector<double> eye(size_t(batch) * n * n, 0.0);
            for (int64_t b = 0; b < batch; ++b)
                for (int i = 0; i < n; ++i) eye[size_t(b) * n * n + size_t(i) * n + i] = 1.0;
            DeviceBuffer<double> dA(size_t(batch) * n * n), dB(size_t(batch) * n * n);
            DeviceBuffer<int> dinfo(std::max<int64_t>(batch, 1));
            dA.upload(A, dA.size(), s);
            dB.upload(eye.data(), dB.size(), s);
            launch_batched_lu_solve(n, n, batch, dA.get(), dB.get(), dinfo.get(), s);
            dB.download(A_inv, dB.size(), s);
            std::vector<int> hi(batch);
            if (batch) dinfo.download(hi.data(), batch, s);
            NAGA_HIP(hipStreamSynchronize(s));
            for (int64_t b = 0; b < batch; ++b)
                if (info) info[b] = hi[b];
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

int naga_batched_lu_solve_f64(const double* A, int32_t n, int64_t batch, const double* B,
                              int32_t nrhs, double* X, int32_t* info) {
    return guarded([&] {
        require(A && B && X, "null argument");
        require(n >= 1 && n <= 64, "n must be in [1, 64]");
        require(nrhs >= 1 && batch >= 0 && batch < int64_t(INT32_MAX), "bad nrhs or batch");
        if (batch == 0) return;
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            DeviceBuffer<double> dA(size_t(batch) * n * n), dB(size_t(batch) * n * nrhs);
            DeviceBuffer<int> dinfo(batch);
            dA.upload(A, dA.size(), s);
            dB.upload(B, dB.size(), s);
            launch_batched_lu_solve(n, nrhs, batch, dA.get(), dB.get(), dinfo.get(), s);
            dB.download(X, dB.size(), s);
            std::vector<int> hi(batch);
            dinfo.download(hi.data(), batch, s);
            NAGA_HIP(hipStreamSynchronize(s));
            for (int64_t b = 0; b < batch; ++b)
                if (info) info[b] = hi[b];
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

int64_t naga_rpim_scratchpad_size(int64_t query_size, int32_t support_size, int32_t dims,
                                  int32_t group_size, int32_t bytes_per_value) {
    if (query_size < 0 || support_size < 1 || dims < 1 || group_size < 1 ||
        (bytes_per_value != 4 && bytes_per_value != 8) || query_size % group_size != 0) {
        g_last_error = "query_size must be a multiple of group_size";
        return -1;
    }
    // the per-group memory of interpolation/detail/radial_point_method.cuh:
    // 38-90: G0 and its inverse, G_x per query, info and pivots
    const int64_t n = int64_t(support_size) + dims + 1;
    const int64_t per_group = 2 * n * n * bytes_per_value + n * bytes_per_value * group_size +
                              4 * int64_t(group_size) + 4 * n * group_size;
    g_last_error.clear();
    return query_size / group_size * per_group;
}

int naga_rpim_create(const float* src, int64_t n_src, const uint32_t* idx, int32_t k,
                     const float* query, int64_t n_q, int32_t dims, int32_t group_size,
                     int32_t rbf, double spacing, naga_rpim** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = nullptr;
        auto r = std::make_unique<naga_rpim>();
        r->impl = std::make_unique<Rpim>(src, n_src, idx, k, query, n_q, dims, group_size, rbf,
                                         spacing, 64);
        *out = r.release();
    });
}

int naga_rpim_destroy(naga_rpim* r) {
    return guarded([&] { delete r; });
}

int naga_rpim_interpolate(naga_rpim* r, const float* field, int64_t n_field, int32_t m,
                          float* dest, float centering_offset) {
    return guarded([&] {
        require(r != nullptr && r->impl, "null interpolator");
        require(n_field == r->impl->sources(),
                "field array must have the same number of rows as the number of source "
                "points used to construct the interpolator");
        r->impl->interpolate(field, m, dest, centering_offset);
    });
}

int naga_rpim_get_weights(naga_rpim* r, float* w) {
    return guarded([&] {
        require(r != nullptr && r->impl, "null interpolator");
        r->impl->weights(w);
    });
}

int naga_segmentation_create(const float* points, int64_t n, int32_t dims,
                             int32_t approx_partition_size, double partition_width,
                             naga_segmentation** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = nullptr;
        auto s = std::make_unique<naga_segmentation>();
        s->impl = std::make_unique<Segmentation>(points, n, dims, approx_partition_size,
                                                 partition_width);
        *out = s.release();
    });
}

int naga_segmentation_destroy(naga_segmentation* s) {
    return guarded([&] { delete s; });
}

int naga_segmentation_shape(naga_segmentation* s, int64_t* shape, double* lower_bounds,
                            double* partition_width, int64_t* partition_count) {
    return guarded([&] {
        require(s != nullptr && s->impl, "null segmentation");
        const CellGrid& g = s->impl->grid();
        for (int d = 0; d < 3; ++d) {
            if (shape) shape[d] = g.dims[d];
            if (lower_bounds) lower_bounds[d] = g.lo[d];
        }
        if (partition_width) *partition_width = g.w;
        if (partition_count) *partition_count = g.ncell;
    });
}

int naga_segmentation_partitions(naga_segmentation* s, uint32_t* sizes, uint32_t* offsets,
                                 uint32_t* indices) {
    return guarded([&] {
        require(s != nullptr && s->impl, "null segmentation");
        s->impl->partitions(sizes, offsets, indices);
    });
}

int naga_segmentation_knn(naga_segmentation* s, const float* query, int64_t n_q, int32_t k,
                          float* d2_out, uint32_t* idx_out) {
    return guarded([&] {
        require(s != nullptr && s->impl, "null segmentation");
        s->impl->knn(query, n_q, k, d2_out, idx_out);
    });
}

int naga_rk4_stage(int32_t q, int64_t n, int64_t n_fluid, int32_t k, const uint32_t* idx,
                   const float* W, double dt, int32_t stage, const float* y, const float* in,
                   const float* y1, const float* y2, float* out) {
    return guarded([&] {
        require(q == 9 || q == 27, "lattice must be D2Q9 (9) or D3Q27 (27)");
        require(stage >= 1 && stage <= 4, "stage must be 1, 2, 3 or 4");
        require(n >= 1 && n_fluid >= 0 && n_fluid <= n && n < (int64_t(1) << 31),
                "need 0 <= n_fluid <= n < 2^31");
        require(k >= 1 && k <= 64, "k must be in [1, 64]");
        require(y && out && (stage == 1 || in) && (stage != 4 || (y1 && y2)),
                "null argument");
        require(n_fluid == 0 || (idx && W), "null stencil");
        for (int64_t t = 0; t < n_fluid * k; ++t) require(idx[t] < n, "index out of range");
        const int D = q == 9 ? 2 : 3, RS = q == 9 ? 9 : 32;
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            auto rows = [&](const float* src) {  // [n][q] -> [n][RS]
                std::vector<float> r(size_t(n) * RS, 0.0f);
                for (int64_t i = 0; i < n; ++i)
                    std::memcpy(&r[size_t(i) * RS], src + size_t(i) * q, sizeof(float) * q);
                return r;
            };
            DeviceBuffer<float> g(size_t(n) * RS), A(size_t(n) * RS), B(size_t(n) * RS),
                C(size_t(n) * RS);
            const std::vector<float> hy = rows(y);
            // every buffer starts as y: rows past n_fluid keep it (zero rate)
            for (DeviceBuffer<float>* b : {&g, &A, &B, &C}) b->upload(hy.data(), hy.size(), s);
            if (stage >= 2) {
                const std::vector<float> hin = rows(in);
                (stage == 2 ? A : stage == 3 ? B : C).upload(hin.data(), hin.size(), s);
            }
            if (stage == 4) {
                const std::vector<float> h1 = rows(y1), h2 = rows(y2);
                A.upload(h1.data(), h1.size(), s);
                B.upload(h2.data(), h2.size(), s);
            }
            DeviceBuffer<uint32_t> didx(std::max<int64_t>(n_fluid * k, 1));
            DeviceBuffer<float> dW(std::max<int64_t>(n_fluid * k * D, 1));
            DeviceBuffer<float> slab(std::max<int64_t>(n_fluid * k * (1 + D), 1));
            if (n_fluid) {
                didx.upload(idx, size_t(n_fluid) * k, s);
                dW.upload(W, size_t(n_fluid) * k * D, s);
                launch_make_slab(n_fluid, k, D, didx.get(), dW.get(), slab.get(), s);
            }
            StepGeom G{};
            G.N = n;
            G.nbulk = n_fluid;
            G.nlayer = 0;
            G.nfluid = n_fluid;
            G.k = k;
            G.kb = 1;
            G.dt = float(dt);
            G.rho0 = 1.0f;
            G.vel_scale = 1.0f;
            StepBuffers S{};
            S.g = g.get();
            S.A = A.get();
            S.B = B.get();
            S.acc = C.get();
            S.slab = slab.get();
            if (q == 27)
                launch_rk4_stage<27>(stage, G, S, s);
            else
                launch_rk4_stage<9>(stage, G, S, s);
            std::vector<float> res(size_t(n) * RS);
            (stage == 1 ? A : stage == 2 ? B : stage == 3 ? C : g).download(res.data(), res.size(),
                                                                          s);
            NAGA_HIP(hipStreamSynchronize(s));
            for (int64_t i = 0; i < n; ++i)
                std::memcpy(out + size_t(i) * q, &res[size_t(i) * RS], sizeof(float) * q);
        } catch (...) {
            (void)hipStreamDestroy(s);
            throw;
        }
        (void)hipStreamDestroy(s);
    });
}

int naga_rpim_weights(const float* src, int64_t n_src, const uint32_t* idx, int32_t k,
                      const float* query, int64_t n_q, int32_t dims, int32_t rbf,
                      double spacing, double* w) {
    return guarded([&] {
        require(src && idx && query && w, "null argument");
        require(dims == 2 || dims == 3, "dims must be 2 or 3");
        require(k >= 1 && k + dims + 1 <= 64, "k + dims + 1 must be <= 64");
        for (int64_t t = 0; t < n_q * k; ++t) require(idx[t] < n_src, "index out of range");
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DeviceBuffer<float> dsrc(n_src * dims), dq(n_q * dims), dw(n_q * k);
        DeviceBuffer<uint32_t> didx(n_q * k);
        DeviceBuffer<double> dw64(n_q * k);
        DeviceBuffer<int> dinfo(n_q);
        dsrc.upload(src, dsrc.size(), s);
        dq.upload(query, dq.size(), s);
        didx.upload(idx, didx.size(), s);
        launch_rpim_weights(dims, dsrc.get(), dq.get(), n_q, didx.get(), k, rbf, spacing,
                            dw.get(), dw64.get(), dinfo.get(), s);
        dw64.download(w, n_q * k, s);
        std::vector<int> hi(n_q);
        dinfo.download(hi.data(), n_q, s);
        NAGA_HIP(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
        for (int64_t q = 0; q < n_q; ++q)
            if (hi[q]) fail(kNumericalError, "singular RPIM system");
    });
}

int naga_divergence_weights(const float* src, int64_t n_src, const float* query, int64_t n_q,
                            int32_t dims, int32_t k, int32_t rbf, uint32_t* idx_out,
                            double* W_out) {
    return guarded([&] {
        require(src && query && idx_out && W_out, "null argument");
        require(dims == 2 || dims == 3, "dims must be 2 or 3");
        require(k >= 2 && k + dims + 1 <= 64, "k must satisfy 2 <= k, k + dims + 1 <= 64");
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        std::vector<float> d2(n_q * k);
        knn_host(dims, src, n_src, query, n_q, k, d2.data(), idx_out, s);
        const int64_t nbatch = (n_q + 16383) / 16384;
        std::vector<double> hbar(std::max<int64_t>(nbatch, 1), 1.0);
        for (int64_t b = 0; b < nbatch; ++b) {
            const int64_t e = std::min<int64_t>((b + 1) * 16384, n_q);
            double acc = 0.0;
            for (int64_t i = b * 16384; i < e; ++i) acc += std::sqrt(double(d2[i * k + 1]));
            hbar[b] = acc / double(e - b * 16384);
        }
        DeviceBuffer<float> dsrc(n_src * dims), dq(n_q * dims), dd2(n_q * k), dW(n_q * k * dims);
        DeviceBuffer<uint32_t> didx(n_q * k);
        DeviceBuffer<double> dW64(n_q * k * dims);
        DeviceBuffer<int> dinfo(n_q);
        dsrc.upload(src, dsrc.size(), s);
        dq.upload(query, dq.size(), s);
        dd2.upload(d2.data(), d2.size(), s);
        didx.upload(idx_out, didx.size(), s);
        auto dh = to_device(hbar, s);
        launch_divergence_weights(dims, dsrc.get(), dq.get(), n_q, didx.get(), dd2.get(), k,
                                  rbf, dh.get(), dW.get(), dW64.get(), dinfo.get(), s);
        dW64.download(W_out, dW64.size(), s);
        std::vector<int> hi(n_q);
        dinfo.download(hi.data(), n_q, s);
        NAGA_HIP(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
        for (int64_t q = 0; q < n_q; ++q)
            if (hi[q]) fail(kNumericalError, "singular RBF system");
    });
}

int naga_divergence_apply(const uint32_t* idx, const float* W, int64_t n, int32_t k,
                          int32_t dims, const float* F, int64_t n_f, float off, float* out) {
    return guarded([&] {
        require(idx && W && F && out, "null argument");
        require(dims == 2 || dims == 3, "dims must be 2 or 3");
        for (int64_t t = 0; t < n * k; ++t) require(idx[t] < n_f, "index out of range");
        hipStream_t s;
        NAGA_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DeviceBuffer<uint32_t> didx(n * k);
        DeviceBuffer<float> dW(n * k * dims), dF(n_f * dims), dout(n);
        didx.upload(idx, didx.size(), s);
        dW.upload(W, dW.size(), s);
        dF.upload(F, dF.size(), s);
        launch_divergence_apply(dims, n, k, didx.get(), dW.get(), dF.get(), off, dout.get(), s);
        dout.download(out, n, s);
        NAGA_HIP(hipStreamSynchronize(s));
        (void)hipStreamDestroy(s);
    });
}

int naga_advection_create(const float* points, int64_t n, int32_t dims, int32_t k, int32_t rbf,
                          naga_advection** out) {
    return guarded([&] {
        require(points && out, "null argument");
        *out = nullptr;
        auto a = std::make_unique<naga_advection>();
        a->impl = std::make_unique<Advection>(points, n, dims, k, rbf);
        *out = a.release();
    });
}

int naga_advection_from_weights(const uint32_t* idx, const float* W, int64_t n, int32_t dims,
                                int32_t k, naga_advection** out) {
    return guarded([&] {
        require(out, "null argument");
        *out = nullptr;
        auto a = std::make_unique<naga_advection>();
        a->impl = std::make_unique<Advection>(idx, W, n, dims, k);
        *out = a.release();
    });
}

int naga_advection_destroy(naga_advection* a) {
    return guarded([&] { delete a; });
}

int naga_advection_step(naga_advection* a, const float* velocity, int32_t per_node, float* f,
                        double dt, double centering_offset, int64_t boundary_start,
                        int32_t nsteps) {
    return guarded([&] {
        require(a && a->impl, "null advection operator");
        require(boundary_start >= 0 && boundary_start <= a->impl->size(),
                "boundary_start out of range");
        a->impl->step(velocity, per_node != 0, f, dt, centering_offset, boundary_start, nsteps);
    });
}

int naga_advection_weights(naga_advection* a, uint32_t* idx, float* W) {
    return guarded([&] {
        require(a && a->impl, "null advection operator");
        a->impl->weights(idx, W);
    });
}

int naga_csr_create(int64_t rows, int64_t cols, const int64_t* row_offsets,
                    const uint32_t* column_indices, const float* values, naga_csr** out) {
    return guarded([&] {
        require(out, "null argument");
        *out = nullptr;
        auto m = std::make_unique<naga_csr>();
        m->impl = std::make_unique<Csr>(rows, cols, row_offsets, column_indices, values);
        *out = m.release();
    });
}

int naga_csr_from_stencil(int64_t rows, int64_t cols, int32_t k, const uint32_t* idx,
                          const float* values, naga_csr** out) {
    return guarded([&] {
        require(idx && values && out, "null argument");
        require(k >= 1, "stencil width must be >= 1");
        require(rows <= int64_t(INT32_MAX) && cols <= int64_t(INT32_MAX),
                "rows and columns must fit in int");
        *out = nullptr;
        std::vector<int64_t> rp(rows + 1);
        for (int64_t r = 0; r <= rows; ++r) rp[r] = r * k;
        auto m = std::make_unique<naga_csr>();
        m->impl = std::make_unique<Csr>(rows, cols, rp.data(), idx, values);
        *out = m.release();
    });
}

int naga_csr_destroy(naga_csr* m) {
    return guarded([&] { delete m; });
}

int naga_csr_mult(naga_csr* m, float alpha, float beta, const float* x, float* y) {
    return guarded([&] {
        require(m && m->impl, "null matrix");
        m->impl->mult(alpha, beta, x, y);
    });
}

}  // extern "C"
