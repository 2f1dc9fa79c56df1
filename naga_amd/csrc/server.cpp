// Serving loop (server.hpp).
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace naga_amd {

namespace {

// Relaxation time -> lattice viscosity with c = c_s (lattice and physical
// units coincide): D3Q27 relaxes shear at 1/(3 nu + 1/2)/dt
// (detail/d3q27_lattice.cuh:177), D2Q9 at 1/(3 nu + dt/2)
// (detail/d2q9_lattice.cuh:136).
double viscosity_for(int Q, double tau, double dt) {
    return Q == 27 ? (tau - 0.5) / 3.0 : (tau - 0.5) * dt / 3.0;
}

struct Nodes {
    std::vector<float> pts, nrm, sig;
    int64_t counts[3] = {0, 0, 0};
};

void make_nodes(const naga_request& r, Nodes& out) {
    const int D = r.lattice == 9 ? 2 : 3;
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {1.0, 1.0, 1.0};
    auto check = [](int st) {
        if (st != NAGA_OK) fail(Status(st), naga_last_error());
    };
    check(naga_obstacle_grid(D, lo, hi, r.h, r.jitter, r.seed, r.n_obstacles, r.obstacles,
                             r.refine_width, r.layer_thickness, r.peak_absorption,
                             r.immersed_layer_thickness, r.immersed_absorption, 0, nullptr,
                             nullptr, nullptr, out.counts));
    const int64_t n = out.counts[0] + out.counts[1] + out.counts[2];
    out.pts.resize(size_t(n) * D);
    out.nrm.resize(size_t(std::max<int64_t>(out.counts[2], 1)) * D);
    out.sig.resize(size_t(std::max<int64_t>(out.counts[1], 1)));
    check(naga_obstacle_grid(D, lo, hi, r.h, r.jitter, r.seed, r.n_obstacles, r.obstacles,
                             r.refine_width, r.layer_thickness, r.peak_absorption,
                             r.immersed_layer_thickness, r.immersed_absorption, n,
                             out.pts.data(), out.nrm.data(), out.sig.data(), out.counts));
}

// Upper bound of a request's node count without generating the nodes: the
// full grid, every obstacle's surface nodes (at half spacing when refined)
// and, with refinement, 2^D - 1 extra nodes per grid cell of the shell
// around each obstacle.
int64_t node_bound(const naga_request& r) {
    const int D = r.lattice == 9 ? 2 : 3;
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {1.0, 1.0, 1.0};
    const int64_t grid = naga_uniform_grid_size(D, lo, hi, r.h);
    require(grid > 0, "bad grid spacing");
    require(r.refine_width >= 0.0, "refine_width must be >= 0");
    const double hs = r.refine_width > 0 ? 0.5 * r.h : r.h;
    double extra = 0.0;
    for (int j = 0; j < r.n_obstacles; ++j) {
        const double R = r.obstacles[j * (D + 1) + D];
        extra += D == 2 ? std::max(8.0, std::ceil(2.0 * M_PI * R / hs))
                        : std::max(16.0, std::ceil(4.0 * M_PI * R * R / (hs * hs)));
        if (r.refine_width > 0) {
            const double ro = R + r.refine_width + r.h, ri = std::max(R - r.h, 0.0);
            const double vol = D == 2 ? M_PI * (ro * ro - ri * ri)
                                      : 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri);
            const double cells = vol / std::pow(r.h, D) + 4.0 * std::pow(ro / r.h + 2.0, D - 1);
            extra += double((1 << D) - 1) * cells;
        }
    }
    return grid + int64_t(std::ceil(extra));
}

int stencil_k(const naga_request& r) { return r.k > 0 ? r.k : (r.lattice == 9 ? 16 : 32); }

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t i = size_t(std::ceil(p * double(v.size()))) - (p > 0 ? 1 : 0);
    return v[std::min(i, v.size() - 1)];
}

}  // namespace

size_t Server::footprint(int Q, int k, int64_t n) {
    const int D = Q == 9 ? 2 : 3, RS = Q == 9 ? Lattice<9>::RS : Lattice<27>::RS;
    // nine [N][RS] state buffers, stencil products (idx, W, filter, slabs),
    // setup transients, per-node scalars; 25% margin
    const double per = 36.0 * RS + double(k) * (4.0 + 4.0 * D + 4.0 + 4.0 * (1 + D) + 8.0 + 8.0) +
                       96.0;
    return size_t(1.25 * per * double(n)) + (size_t(8) << 20);
}

Server::Server(size_t hbm_budget, int max_active, int chunk_steps, bool graph,
               int setup_threads)
    : budget_(hbm_budget), max_active_(max_active), chunk_(chunk_steps), graph_(graph),
      t0_(std::chrono::steady_clock::now()) {
    require(hbm_budget > 0 && max_active > 0 && chunk_steps > 0 && setup_threads > 0,
            "bad server settings");
    for (int t = 0; t < setup_threads; ++t)
        setup_threads_.emplace_back([this] { setup_loop(); });
    step_thread_ = std::thread([this] { step_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : setup_threads_)
        if (t.joinable()) t.join();
    if (step_thread_.joinable()) step_thread_.join();
}

double Server::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_)
        .count();
}

int64_t Server::submit(const naga_request& req) {
    require(req.lattice == 9 || req.lattice == 27, "lattice must be 9 or 27");
    require(req.steps >= 0, "steps must be >= 0");
    require(req.h > 0 && req.pulse_width > 0, "h and pulse width must be positive");
    require(req.n_obstacles >= 0 && req.n_obstacles <= NAGA_MAX_OBSTACLES, "too many obstacles");
    std::lock_guard<std::mutex> lk(mu_);
    const int64_t id = next_id_++;
    const double t = now_ms();
    if (t_first_submit_ < 0) t_first_submit_ = t;
    queue_.push_back(Pending{id, req, t});
    state_[id] = 0;
    ++submitted_;
    cv_.notify_all();
    return id;
}

int Server::status(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = state_.find(id);
    return it == state_.end() ? -1 : it->second;
}

bool Server::result(int64_t id, ServedResult& out) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = done_.find(id);
    if (it == done_.end()) return false;
    out = it->second;
    return true;
}

void Server::release(int64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    done_.erase(id);
    state_.erase(id);
}

void Server::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return completed_ + failed_ == submitted_; });
}

naga_server_stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    naga_server_stats s{};
    s.submitted = submitted_;
    s.completed = completed_;
    s.failed = failed_;
    s.resident = resident_;
    s.queued = int64_t(queue_.size());
    s.bytes_in_use = int64_t(used_);
    s.p50_ms = percentile(latencies_, 0.50);
    s.p99_ms = percentile(latencies_, 0.99);
    double sum = 0.0;
    for (double v : latencies_) sum += v;
    s.mean_ms = latencies_.empty() ? 0.0 : sum / double(latencies_.size());
    const double span = t_last_done_ - t_first_submit_;
    s.node_updates_per_s = span > 0 ? node_updates_ / (span * 1e-3) : 0.0;
    return s;
}

// Admission and scene setup (one loop per setup thread).  Requests are
// taken in arrival order; the head waits until its footprint fits the
// budget and a slot is free.
void Server::setup_loop() {
    for (;;) {
        Pending p;
        size_t bytes = 0;
        int64_t n = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_) return;
                if (!queue_.empty()) {
                    const naga_request& r = queue_.front().req;
                    try {
                        n = node_bound(r);
                        bytes = footprint(r.lattice, stencil_k(r), n);
                    } catch (const std::exception& e) {
                        Pending bad = queue_.front();
                        queue_.pop_front();
                        ServedResult res;
                        res.error = e.what();
                        done_[bad.id] = std::move(res);
                        state_[bad.id] = 1;
                        ++failed_;
                        cv_.notify_all();
                        continue;
                    }
                    const bool fits = used_ + bytes <= budget_ || resident_ == 0;
                    if (fits && resident_ < max_active_) break;
                }
                cv_.wait(lk);
            }
            p = queue_.front();
            queue_.pop_front();
            used_ += bytes;
            ++resident_;
        }
        Active a;
        a.id = p.id;
        a.t_submit = p.t_submit;
        a.t_admit = now_ms();
        a.bytes = bytes;
        a.steps = a.remaining = p.req.steps;
        try {
            const naga_request& r = p.req;
            Nodes nd;
            make_nodes(r, nd);
            a.nodes = nd.counts[0] + nd.counts[1] + nd.counts[2];
            SceneConfig cfg;
            cfg.Q = r.lattice;
            cfg.k = cfg.kb = stencil_k(r);
            cfg.k_min = r.k_min;
            cfg.rbf = cfg.rbf_boundary = 1;
            cfg.pml_ref = 1;
            cfg.h = r.h;
            // dt = 0.4 of the finest spacing (h / 2 in a refined shell)
            cfg.dt = 0.4 * (r.refine_width > 0 ? 0.5 * r.h : r.h);
            cfg.c = kCsLat;
            cfg.rho0 = 1.0;
            cfg.nu = viscosity_for(r.lattice, r.tau > 0 ? r.tau : 0.6, cfg.dt);
            a.eng = std::make_unique<Engine>(cfg);
            a.eng->init_domain(nd.pts.data(), nd.counts[0], nd.counts[1], nd.counts[2],
                               nd.nrm.data(), nd.sig.data());
            a.eng->set_pulse(r.pulse_center, r.pulse_amplitude, r.pulse_width);
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> lk(mu_);
            ServedResult res;
            res.error = e.what();
            done_[a.id] = std::move(res);
            state_[a.id] = 1;
            used_ -= a.bytes;
            --resident_;
            ++failed_;
            cv_.notify_all();
            continue;
        }
        std::lock_guard<std::mutex> lk(mu_);
        ready_.push_back(std::move(a));
        cv_.notify_all();
    }
}

void Server::finish(Active& a, double now) {
    ServedResult res;
    res.ok = true;
    res.nodes = a.nodes;
    res.steps = a.steps;
    const int D = a.eng->dims();
    res.rho.resize(size_t(a.nodes));
    res.u.resize(size_t(a.nodes) * D);
    res.drho.resize(size_t(a.nodes));
    a.eng->get_macroscopic(res.rho.data(), res.u.data(), res.drho.data(), true);
    a.eng.reset();  // frees the scene's device memory
    const double t = now_ms();
    res.latency_ms = t - a.t_submit;
    res.queue_ms = a.t_admit - a.t_submit;
    (void)now;
    std::lock_guard<std::mutex> lk(mu_);
    done_[a.id] = std::move(res);
    state_[a.id] = 1;
    used_ -= a.bytes;
    --resident_;
    ++completed_;
    node_updates_ += double(a.nodes) * double(a.steps);
    latencies_.push_back(t - a.t_submit);
    t_last_done_ = t;
    cv_.notify_all();
}

// Continuous batching: scenes ready from setup join the batch at chunk
// boundaries; each chunk advances every resident scene by the same number
// of steps, at most chunk_ and at most the smallest remaining count, so
// finished scenes leave as soon as they are done.
void Server::step_loop() {
    Batch batch;
    batch.set_graph(graph_);
    std::vector<Active> active;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !ready_.empty() || !active.empty(); });
            if (stop_ && ready_.empty() && active.empty()) return;
            for (Active& a : ready_) active.push_back(std::move(a));
            ready_.clear();
        }
        // scenes with no steps left (including zero-step requests) finish now
        for (size_t i = 0; i < active.size();) {
            if (active[i].remaining <= 0) {
                if (active[i].eng->batch) batch.remove(active[i].eng.get());
                finish(active[i], now_ms());
                active.erase(active.begin() + i);
            } else {
                if (!active[i].eng->batch) batch.add(active[i].eng.get());
                ++i;
            }
        }
        if (active.empty()) continue;
        int64_t n = chunk_;
        for (const Active& a : active) n = std::min(n, a.remaining);
        try {
            batch.step(int(n));
        } catch (const std::exception& e) {
            // a failed step fails every resident scene of the batch
            for (Active& a : active) {
                if (a.eng->batch) batch.remove(a.eng.get());
                a.eng.reset();
                std::lock_guard<std::mutex> lk(mu_);
                ServedResult res;
                res.error = e.what();
                done_[a.id] = std::move(res);
                state_[a.id] = 1;
                used_ -= a.bytes;
                --resident_;
                ++failed_;
            }
            active.clear();
            cv_.notify_all();
            continue;
        }
        for (Active& a : active) a.remaining -= n;
    }
}

}  // namespace naga_amd
