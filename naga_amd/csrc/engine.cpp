// Engine: setup orchestration and the per-step schedule on one HIP stream.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstring>

#include "setup_kernels.hpp"

namespace naga_amd {

namespace {

template <int Q>
double lat_w(int a) {
    return Lattice<Q>::w(a);
}

double weight(int Q, int a) { return Q == 27 ? lat_w<27>(a) : lat_w<9>(a); }

// Runs the launches of a stream capture; if one of them throws, the
// capture is ended and its partial graph dropped before the error
// propagates, so the stream is usable again.
template <class F>
void capture_guarded(hipStream_t st, F&& body) {
    try {
        body();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
}

// Events recorded on a stream at phase boundaries; read() gives the elapsed
// milliseconds between consecutive marks once the stream has reached them.
struct PhaseMarks {
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    explicit PhaseMarks(hipStream_t s) : st(s) { mark(); }
    ~PhaseMarks() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void mark() {
        hipEvent_t e = nullptr;
        NAGA_HIP(hipEventCreate(&e));
        ev.push_back(e);
        NAGA_HIP(hipEventRecord(e, st));
    }
    void read(double* ms, int n) const {
        for (int i = 0; i < n; ++i) {
            float v = 0.0f;
            if (i + 1 < int(ev.size())) NAGA_HIP(hipEventElapsedTime(&v, ev[i], ev[i + 1]));
            ms[i] = double(v);
        }
    }
};

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};

}  // namespace

void knn_host(int D, const float* ref, int64_t n_ref, const float* query, int64_t n_q,
              int k, float* d2, uint32_t* idx, hipStream_t st) {
    require(D == 2 || D == 3, "dimension must be 2 or 3");
    require(k >= 1 && k <= 64, "k must be in [1, 64]");
    require(n_ref >= k, "fewer reference points than neighbours requested");
    DeviceBuffer<float> dref(n_ref * D), dq(n_q * D), dd2(n_q * k);
    DeviceBuffer<uint32_t> didx(n_q * k);
    dref.upload(ref, n_ref * D, st);
    dq.upload(query, n_q * D, st);
    const DeviceCells cells = bin_points_device(D, dref.get(), n_ref, st);
    launch_knn(cells.bins, dref.get(), dq.get(), n_q, k, dd2.get(), didx.get(), st);
    dd2.download(d2, n_q * k, st);
    didx.download(idx, n_q * k, st);
    NAGA_HIP(hipStreamSynchronize(st));
}

Engine::Engine(const SceneConfig& cfg) : cfg_(cfg) {
    require(cfg.Q == 9 || cfg.Q == 27, "lattice must be D2Q9 (9) or D3Q27 (27)");
    D_ = cfg.Q == 9 ? 2 : 3;
    RS_ = cfg.Q == 9 ? Lattice<9>::RS : Lattice<27>::RS;
    require(cfg.k >= 2 && cfg.k + D_ + 1 <= 64, "k must satisfy 2 <= k and k + D + 1 <= 64");
    require(cfg.k_min >= 0 && (cfg.k_min == 0 || cfg.k_min >= 2), "k_min must be 0 or >= 2");
    require(cfg.kb >= 1 && cfg.kb + D_ + 1 <= 64, "k_boundary must satisfy k_boundary + D + 1 <= 64");
    require(cfg.rbf == 0 || cfg.rbf == 1, "rbf must be 0 (multiquadric) or 1 (r^3)");
    require(cfg.rbf_boundary == 0 || cfg.rbf_boundary == 1, "rbf_boundary must be 0 or 1");
    require(cfg.dt > 0 && cfg.c > 0 && cfg.rho0 > 0 && cfg.h > 0,
            "dt, c, rho0 and h must be positive");
    NAGA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    // set_problem_parameters (detail/simulation_engine.cuh:94-124)
    const double ratio = kCsLat / cfg.c;
    const double dt_lat = cfg.dt / ratio;
    const double nu_lat = cfg.nu / ratio;
    cp_ = make_collision_params(cfg.Q, nu_lat, dt_lat);
    geo_.dt = float(dt_lat);
    geo_.rho0 = float(cfg.rho0);
    geo_.vel_scale = float(cfg.c / kCsLat);
    // selective filter strength (detail/simulation_engine.cuh:1242: 0.75)
    geo_.filter = cfg.filter == 0.0 ? 0.75f : (cfg.filter < 0.0 ? 0.0f : float(cfg.filter));
    geo_.k = cfg.k;
    geo_.kb = cfg.kb;
    geo_.pml_ref = cfg.pml_ref ? 1 : 0;
}

Engine::~Engine() {
    if (batch) batch->remove(this);
    drop_step_graph();
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void Engine::require_domain() const {
    require(ready_, "init_domain must be called first");
}

// D2Q9 tile unions: built when NAGA_Q9_UNIONS=1 (off by default).
static bool q9_unions_wanted(int64_t /*N*/) {
    const char* e = std::getenv("NAGA_Q9_UNIONS");
    return e && e[0] == '1';
}

void Engine::init_domain(const float* pts, int64_t nbulk, int64_t nlayer, int64_t nbnd,
                         const float* normals, const float* sigma) {
    require(batch == nullptr, "remove the engine from its batch before init_domain");
    drop_step_graph();  // the captured step refers to the old buffers
    Timer timer;
    const int D = D_, Q = cfg_.Q, k = cfg_.k, kb = cfg_.kb;
    // shape checks of detail/simulation_engine.cuh:128-167
    require(pts != nullptr, "The number of points in the simulation domain does not match "
                            "the sum of the number of bulk, layer, and boundary points.");
    require(nbulk >= 0 && nlayer >= 0 && nbnd >= 0, "negative node count");
    require(nbnd == 0 || normals != nullptr,
            "The number of boundary normals does not match the number of boundary points.");
    require(nlayer == 0 || sigma != nullptr,
            "The number of layer absorption coefficients does not match the number of layer "
            "points.");
    const int64_t N = nbulk + nlayer + nbnd;
    const int64_t nfluid = nbulk + nlayer;
    require(N >= k, "fewer nodes than the divergence stencil width");
    require(nbnd == 0 || nfluid >= kb, "fewer fluid nodes than the boundary stencil width");
    require(N < (int64_t(1) << 31), "at most 2^31 - 1 nodes per scene");
    N_ = N;
    nbulk_ = nbulk;
    nlayer_ = nlayer;
    nbnd_ = nbnd;
    nfluid_ = nfluid;
    ready_ = false;
    hipStream_t st = stream_;

    // phase marks on the stream: setup_phase_ms_ reports the wall time
    // between consecutive marks
    PhaseMarks marks(st);
    pts_.allocate(N * D);
    pts_.upload(pts, N * D, st);
    hpts_.assign(pts, pts + N * D);
    hnormals_.assign(normals ? normals : pts, normals ? normals + nbnd * D : pts);
    hsigma_.assign(sigma ? sigma : pts, sigma ? sigma + nlayer : pts);
    probes_.clear();
    sources_.clear();
    n_snodes_ = 0;

    // divergence operator over all nodes (init_domain :169-181)
    {
        const DeviceCells cells = bin_points_device(D, pts_.get(), N, st);
        DeviceBuffer<float> d2(N * k);
        idx_.allocate(N * k);
        launch_knn(cells.bins, pts_.get(), pts_.get(), N, k, d2.get(), idx_.get(), st);
        marks.mark();  // binning and kNN done
        // multiquadric shape parameter: mean nearest distance per 16384-query
        // batch (nonlocal_calculus/divergence.cuh:118-160, operators.cuh:96-117)
        const int64_t nbatch = (N + 16383) / 16384;
        std::vector<double> hbar(nbatch, 1.0);
        if (cfg_.rbf == 0) {
            std::vector<float> hd2(N * k);
            d2.download(hd2.data(), N * k, st);
            NAGA_HIP(hipStreamSynchronize(st));
            for (int64_t b = 0; b < nbatch; ++b) {
                const int64_t e = std::min<int64_t>((b + 1) * 16384, N);
                double s = 0.0;
                for (int64_t i = b * 16384; i < e; ++i) s += std::sqrt(double(hd2[i * k + 1]));
                hbar[b] = s / double(e - b * 16384);
            }
        }
        auto dhbar = to_device(hbar, st);
        // per-node stencil widths by local density: k_i = k_min d_ref / d_i
        // clamped to [k_min, k], d_i the mean distance to the four nearest
        // neighbours and d_ref its 90th percentile (the coarse spacing)
        kvar_ = DeviceBuffer<uint8_t>();
        if (cfg_.k_min > 0 && cfg_.k_min < k) {
            require(k >= 5 && k <= 255, "variable stencils need 5 <= k <= 255");
            std::vector<float> hd2(size_t(N) * k);
            d2.download(hd2.data(), hd2.size(), st);
            NAGA_HIP(hipStreamSynchronize(st));
            std::vector<double> dm(N);
            for (int64_t i = 0; i < N; ++i) {
                double a = 0.0;
                for (int s = 1; s <= 4; ++s) a += std::sqrt(double(hd2[size_t(i) * k + s]));
                dm[i] = a / 4.0;
            }
            std::vector<double> sorted_dm(dm);
            const size_t p90 = size_t(0.9 * double(N - 1));
            std::nth_element(sorted_dm.begin(), sorted_dm.begin() + p90, sorted_dm.end());
            const double dref = sorted_dm[p90];
            std::vector<uint8_t> kv(N);
            for (int64_t i = 0; i < N; ++i) {
                const double want = double(cfg_.k_min) * dref / std::max(dm[i], 1e-30);
                kv[i] = uint8_t(std::min<double>(k, std::max<double>(cfg_.k_min, std::round(want))));
            }
            kvar_ = to_device(kv, st);
        }
        W_.allocate(N * k * D);
        DeviceBuffer<int> info(N);
        launch_divergence_weights(D, pts_.get(), pts_.get(), N, idx_.get(), d2.get(), k,
                                  cfg_.rbf, dhbar.get(), W_.get(), nullptr, info.get(), st,
                                  kvar_.get(), cfg_.solve_bits);
        std::vector<int> hinfo(N);
        info.download(hinfo.data(), N, st);
        NAGA_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < N; ++i)
            if (hinfo[i])
                fail(kNumericalError, "singular RBF system while building the divergence "
                                      "operator at node " + std::to_string(i));
    }
    marks.mark();  // divergence weights done
    filt_.allocate(N * k);
    launch_filter_weights(D, pts_.get(), N, idx_.get(), k, cfg_.h, filt_.get(), st, kvar_.get());
    if (kvar_.size() == size_t(N)) launch_self_past_k(N, k, kvar_.get(), idx_.get(), st);
    slab_.allocate(nfluid * k * (1 + D));
    launch_make_slab(nfluid, k, D, idx_.get(), W_.get(), slab_.get(), st);
    fslab_.allocate(nfluid * k * 2);
    launch_make_slab(nfluid, k, 1, idx_.get(), filt_.get(), fslab_.get(), st);
    ucap_ = 0;
    ucap_f_ = 0;
    self_first_ = 0;
    tile_ng_ = DeviceBuffer<int32_t>();
    rec_off_ = DeviceBuffer<int64_t>();
    frec_off_ = DeviceBuffer<int64_t>();
    ulist_ = DeviceBuffer<uint32_t>();
    uoff_ = DeviceBuffer<int64_t>();
    usplit_ = DeviceBuffer<int32_t>();
    split_tiles_ = DeviceBuffer<int32_t>();
    split_of_.clear();
    n_split_ = 0;
    p2ulist_ = DeviceBuffer<uint32_t>();
    p2uoff_ = DeviceBuffer<int64_t>();
    p2usplit_ = DeviceBuffer<int32_t>();
    p2split_tiles_ = DeviceBuffer<int32_t>();
    p2slab_ = DeviceBuffer<float>();
    uslab_ = DeviceBuffer<float>();
    ufslab_ = DeviceBuffer<float>();
    p2split_of_.clear();
    p2n_split_ = 0;
    bulist_ = DeviceBuffer<uint32_t>();
    buoff_ = DeviceBuffer<int64_t>();
    // Tile unions: the gather kernels copy each tile's distinct stencil
    // neighbours into LDS by LDS DMA and read them from there (fewer bytes
    // through the vector L1 than direct gathers: 0.2 against 0.27 ms per
    // RK4 stage on the 1M-node D3Q27 scene); NAGA_UNIONS=0 disables them.
    const char* un_env = std::getenv("NAGA_UNIONS");
    tile_nodes_ = 0;
    if (!(un_env && un_env[0] == '0') && step_uses_unions(Q, k, N) && nfluid > 0) {
        // tile unions; slab index fields become union positions
        std::vector<int32_t> last_counts;  // union rows per tile of the last set built
        const int T = union_tile_nodes(k);
        tile_nodes_ = T;
        const int64_t ntiles = (nfluid + T - 1) / T;
        const char* env = std::getenv("NAGA_UCAP");
        const int cap = env ? std::max(0, std::min(std::atoi(env), kUnionRowsMax))
                            : union_cap_default(k);
        // tiles over the cap are split into two half-tile unions when
        // those fit (NAGA_SPLIT=0: never)
        const char* sp_env = std::getenv("NAGA_SPLIT");
        const bool split = !(sp_env && sp_env[0] == '0');
        // one union set over the stencils idx_in: lists, offsets, split
        // tiles; the index fields of slab_a / slab_b become union positions
        auto build = [&](const uint32_t* idx_in, float* slab_a, int nca, float* slab_b, int ncb,
                         DeviceBuffer<uint32_t>& ulist, DeviceBuffer<int64_t>& uoff,
                         DeviceBuffer<int32_t>& usplit, DeviceBuffer<int32_t>& split_tiles,
                         std::vector<int32_t>& split_of, int64_t& n_split) {
            DeviceBuffer<int32_t> counts(ntiles);
            last_counts.clear();
            usplit.allocate(ntiles);
            launch_tile_unions(nfluid, k, T, idx_in, split ? cap : 0, counts.get(), usplit.get(),
                               nullptr, nullptr, nullptr, 0, nullptr, 0, st);
            std::vector<int32_t> hc(ntiles), hs(ntiles), splits;
            counts.download(hc.data(), ntiles, st);
            usplit.download(hs.data(), ntiles, st);
            NAGA_HIP(hipStreamSynchronize(st));
            std::vector<int64_t> off(ntiles + 1, 0);
            int maxu = 0;
            split_of.assign(ntiles, -1);
            last_counts = hc;
            for (int64_t t = 0; t < ntiles; ++t) {
                off[t + 1] = off[t] + hc[t];
                maxu = std::max(maxu, hc[t]);
                if (hs[t] > 0) {  // split: one more block serves the second half
                    split_of[t] = int32_t(splits.size());
                    splits.push_back(int32_t(t));
                }
            }
            uoff = to_device(off, st);
            ulist.allocate(std::max<int64_t>(off[ntiles], 1));
            launch_tile_unions(nfluid, k, T, idx_in, split ? cap : 0, nullptr, usplit.get(),
                               uoff.get(), ulist.get(), slab_a, nca, slab_b, ncb, st);
            n_split = int64_t(splits.size());
            split_tiles = to_device(splits.empty() ? std::vector<int32_t>{0} : splits, st);
            return maxu;
        };
        // The PML divergence reads Gf, zero outside the layer: its unions
        // hold the layer neighbours and one zero row (every other neighbour
        // mapped to a node outside the layer), about 60% of the full union.
        DeviceBuffer<float> slab2;
        if (nlayer > 0) {
            DeviceBuffer<uint32_t> idx2(nfluid * k);
            const uint32_t zero_row = nbnd > 0 ? uint32_t(nfluid) : 0u;
            launch_layer_only(nfluid * k, idx_.get(), uint32_t(nbulk), uint32_t(nfluid), zero_row,
                              idx2.get(), st);
            slab2.allocate(nfluid * k * (1 + D));
            launch_make_slab(nfluid, k, D, idx_.get(), W_.get(), slab2.get(), st);
            build(idx2.get(), slab2.get(), D, nullptr, 0, p2ulist_, p2uoff_, p2usplit_,
                  p2split_tiles_, p2split_of_, p2n_split_);
            p2slab_.allocate(nfluid * packed_div_units(k) * 4);
            launch_pack_div_records(nfluid, k, slab2.get(), p2slab_.get(), st);
            NAGA_HIP(hipStreamSynchronize(st));
        }
        const int maxu = build(idx_.get(), slab_.get(), D, fslab_.get(), 1, ulist_, uoff_,
                               usplit_, split_tiles_, split_of_, n_split_);
        // the union kernels copy records into LDS as they lie: packed rows
        // (16-bit union positions, an odd number of 16-byte units per node)
        // keep the records of consecutive nodes in distinct LDS banks
        DeviceBuffer<float> ps(nfluid * packed_div_units(k) * 4),
            pf(nfluid * packed_filter_units(k) * 4);
        launch_pack_div_records(nfluid, k, slab_.get(), ps.get(), st);
        launch_pack_filter_records(nfluid, k, fslab_.get(), pf.get(), st);
        NAGA_HIP(hipStreamSynchronize(st));
        slab_ = std::move(ps);
        fslab_ = std::move(pf);
        ucap_ = std::min(maxu, cap);
        // the filter pass's union cap: with 32-node tiles (256-thread
        // blocks) a cap at the 95th percentile of the tiles' unions
        // (multiple of 8) lets a third block share a CU, the tiles above it
        // reading their rows through the list (2.24M-node irregular scene:
        // 3.55 against 3.66 ms per step with the stage cap); with 64-node
        // tiles registers hold the filter to two 512-thread blocks per CU
        // anyway, so it keeps the stage cap (1M-node cube: 144 against
        // 152 us for the pass)
        {
            std::vector<int32_t> c = last_counts;
            int p95 = ucap_;
            if (!c.empty() && T < 64) {
                const size_t at = size_t(0.95 * double(c.size() - 1));
                std::nth_element(c.begin(), c.begin() + at, c.end());
                p95 = (c[at] + 7) / 8 * 8;
            }
            ucap_f_ = std::max(8, std::min(ucap_, p95));
            if (const char* fe = std::getenv("NAGA_UCAP_F"))
                ucap_f_ = std::max(8, std::min(ucap_, std::atoi(fe)));
        }
        // whether every fluid node's first stencil entry is itself (the kNN
        // puts it first unless another node sits at the same point)
        const char* ol_env = std::getenv("NAGA_OWN_LDS");
        self_first_ = (ol_env && ol_env[0] == '0') ? 0
                                                    : int(stencils_self_first(nfluid, k, idx_.get(), st));
        // per-node stencil widths: each tile keeps the groups of 8 records
        // its widest node needs (ragged record rows), so the stencil passes
        // read and sum k_tile records per node instead of k
        const char* rg_env = std::getenv("NAGA_RAGGED");
        if (kvar_.size() == size_t(N) && !(rg_env && rg_env[0] == '0')) {
            std::vector<uint8_t> hk(static_cast<size_t>(N));
            kvar_.download(hk.data(), size_t(N), st);
            NAGA_HIP(hipStreamSynchronize(st));
            std::vector<int32_t> ng(static_cast<size_t>(ntiles));
            std::vector<int64_t> ro(static_cast<size_t>(ntiles) + 1, 0),
                fo(static_cast<size_t>(ntiles) + 1, 0);
            for (int64_t t = 0; t < ntiles; ++t) {
                const int64_t a = t * T, b = std::min<int64_t>(a + T, nfluid);
                int m = 1;
                for (int64_t i = a; i < b; ++i) m = std::max(m, (int(hk[size_t(i)]) + 7) / 8);
                ng[size_t(t)] = std::min(m, k / 8);
                ro[size_t(t) + 1] = ro[size_t(t)] + (b - a) * ragged_div_units(ng[size_t(t)]);
                fo[size_t(t) + 1] = fo[size_t(t)] + (b - a) * ragged_filter_units(ng[size_t(t)]);
            }
            tile_ng_ = to_device(ng, st);
            rec_off_ = to_device(ro, st);
            frec_off_ = to_device(fo, st);
            DeviceBuffer<float> rs(size_t(ro[size_t(ntiles)]) * 4),
                rf(size_t(fo[size_t(ntiles)]) * 4);
            launch_ragged_records(nfluid, T, packed_div_units(k), 7, tile_ng_.get(),
                                  rec_off_.get(), slab_.get(), rs.get(), st);
            launch_ragged_records(nfluid, T, packed_filter_units(k), 3, tile_ng_.get(),
                                  frec_off_.get(), fslab_.get(), rf.get(), st);
            if (p2slab_.size() > 0) {
                DeviceBuffer<float> r2(size_t(ro[size_t(ntiles)]) * 4);
                launch_ragged_records(nfluid, T, packed_div_units(k), 7, tile_ng_.get(),
                                      rec_off_.get(), p2slab_.get(), r2.get(), st);
                NAGA_HIP(hipStreamSynchronize(st));
                p2slab_ = std::move(r2);
            }
            NAGA_HIP(hipStreamSynchronize(st));
            slab_ = std::move(rs);
            fslab_ = std::move(rf);
        }
    } else if (q9_unions_wanted(N) && !(un_env && un_env[0] == '0') && q9_three(Q, N) &&
               nfluid > 0 && k * kQ9TileNodes <= 4096) {
        // D2Q9 tile unions over the three-per-lane kernels' 84-node tiles
        // (no split tiles: a tile whose union is over the cap reads its rows
        // through the list).  Copies of the stage and filter records get
        // union positions in their index fields (uslab, ufslab); the
        // records with node indices stay for the PML divergence (a node
        // list) and for batches that step D2Q9 scenes without unions.
        const int T = kQ9TileNodes;
        const int64_t ntiles = (nfluid + T - 1) / T;
        uslab_.allocate(slab_.size());
        ufslab_.allocate(fslab_.size());
        NAGA_HIP(hipMemcpyAsync(uslab_.get(), slab_.get(), slab_.size() * sizeof(float),
                                hipMemcpyDeviceToDevice, st));
        NAGA_HIP(hipMemcpyAsync(ufslab_.get(), fslab_.get(), fslab_.size() * sizeof(float),
                                hipMemcpyDeviceToDevice, st));
        DeviceBuffer<int32_t> counts(ntiles);
        launch_tile_unions(nfluid, k, T, idx_.get(), 0, counts.get(), nullptr, nullptr, nullptr,
                           nullptr, 0, nullptr, 0, st);
        std::vector<int32_t> hc(static_cast<size_t>(ntiles));
        counts.download(hc.data(), size_t(ntiles), st);
        NAGA_HIP(hipStreamSynchronize(st));
        std::vector<int64_t> off(static_cast<size_t>(ntiles) + 1, 0);
        int maxu = 0;
        for (int64_t t = 0; t < ntiles; ++t) {
            off[size_t(t) + 1] = off[size_t(t)] + hc[size_t(t)];
            maxu = std::max(maxu, hc[size_t(t)]);
        }
        uoff_ = to_device(off, st);
        ulist_.allocate(std::max<int64_t>(off[size_t(ntiles)], 1));
        launch_tile_unions(nfluid, k, T, idx_.get(), 0, nullptr, nullptr, uoff_.get(),
                           ulist_.get(), uslab_.get(), D, ufslab_.get(), 1, st);
        NAGA_HIP(hipStreamSynchronize(st));
        ucap_ = std::min(maxu, kQ9UnionRowsMax);
        ucap_f_ = ucap_;
    }

    marks.mark();  // filter weights, slabs and tile unions done
    // boundary interpolator (init_domain :223-258): kNN of boundary nodes
    // among the bulk and layer nodes, RPIM weights at spacing h
    if (nbnd > 0) {
        const DeviceCells cells = bin_points_device(D, pts_.get(), nfluid, st);
        DeviceBuffer<float> bd2(nbnd * kb);
        bidx_.allocate(nbnd * kb);
        launch_knn(cells.bins, pts_.get(), pts_.get() + nfluid * D, nbnd, kb, bd2.get(),
                   bidx_.get(), st);
        bw_.allocate(nbnd * kb);
        DeviceBuffer<int> info(nbnd);
        launch_rpim_weights(D, pts_.get(), pts_.get() + nfluid * D, nbnd, bidx_.get(), kb,
                            cfg_.rbf_boundary, cfg_.h, bw_.get(), nullptr, info.get(), st,
                            cfg_.solve_bits);
        std::vector<int> hinfo(nbnd);
        info.download(hinfo.data(), nbnd, st);
        NAGA_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < nbnd; ++i)
            if (hinfo[i])
                fail(kNumericalError, "singular RBF system in the boundary interpolator");
        // fluid nodes read by the interpolator keep their collision
        // increment aside for it
        std::vector<uint32_t> hb(nbnd * kb);
        bidx_.download(hb.data(), hb.size(), st);
        NAGA_HIP(hipStreamSynchronize(st));
        std::vector<uint8_t> src(nfluid, 0);
        for (uint32_t j : hb) src[j] = 1;
        bsrc_ = to_device(src, st);
        bslab_.allocate(nbnd * kb * 2);
        launch_make_slab(nbnd, kb, 1, bidx_.get(), bw_.get(), bslab_.get(), st);
        if (tile_nodes_ > 0) {
            // boundary tiles of 32 nodes: their distinct source nodes are
            // staged in LDS by the interpolation kernels; bslab index fields
            // become positions in that list
            const int64_t nbt = (nbnd + kBTileNodes - 1) / kBTileNodes;
            DeviceBuffer<int32_t> counts(nbt);
            launch_tile_unions(nbnd, kb, kBTileNodes, bidx_.get(), 0, counts.get(), nullptr,
                               nullptr, nullptr, nullptr, 0, nullptr, 0, st);
            std::vector<int32_t> hc(nbt);
            counts.download(hc.data(), nbt, st);
            NAGA_HIP(hipStreamSynchronize(st));
            std::vector<int64_t> off(nbt + 1, 0);
            for (int64_t t = 0; t < nbt; ++t) off[t + 1] = off[t] + hc[t];
            buoff_ = to_device(off, st);
            bulist_.allocate(std::max<int64_t>(off[nbt], 1));
            launch_tile_unions(nbnd, kb, kBTileNodes, bidx_.get(), 0, nullptr, nullptr, buoff_.get(),
                               bulist_.get(), bslab_.get(), 1, nullptr, 0, st);
        }
        normals_.allocate(nbnd * D);
        normals_.upload(normals, nbnd * D, st);
    }
    marks.mark();  // boundary interpolator done
    // fluid nodes whose stencil reaches the absorbing layer
    n_pml_ = 0;
    n_ptiles_ = 0;
    ptiles_ = DeviceBuffer<int32_t>();
    if (nlayer > 0) {
        DeviceBuffer<uint8_t> flag(nfluid);
        launch_flag_neighbours(nfluid, k, idx_.get(), nbulk, nfluid, flag.get(), st);
        std::vector<uint8_t> hflag(nfluid);
        flag.download(hflag.data(), nfluid, st);
        NAGA_HIP(hipStreamSynchronize(st));
        std::vector<int32_t> list;
        for (int64_t i = 0; i < nfluid; ++i)
            if (hflag[i]) list.push_back(int32_t(i));
        n_pml_ = int64_t(list.size());
        pml_list_ = to_device(list, st);
        // with tile unions the PML divergence runs over whole tiles: every
        // tile holding a listed node (the others add an exact zero)
        if (ulist_.get() && tile_nodes_ > 0) {
            std::vector<int32_t> tl;
            const int T = tile_nodes_;
            for (int64_t t0 = 0; t0 < nfluid; t0 += T) {
                const int64_t t1 = std::min<int64_t>(t0 + T, nfluid);
                for (int64_t i = t0; i < t1; ++i)
                    if (hflag[i]) {
                        const int64_t t = t0 / T;
                        tl.push_back(int32_t(t));
                        // a split tile's second half: entry -1 - s
                        if (t < int64_t(p2split_of_.size()) && p2split_of_[t] >= 0)
                            tl.push_back(-1 - p2split_of_[t]);
                        break;
                    }
            }
            n_ptiles_ = int64_t(tl.size());
            ptiles_ = to_device(tl.empty() ? std::vector<int32_t>{0} : tl, st);
        }
        sigma_.allocate(nlayer);
        sigma_.upload(sigma, nlayer, st);
    }
    // nodes whose PML Q values are maintained: the layer, plus [0, nlayer)
    // when the layer field reads Q at j - num_bulk
    n_rq_ = 0;
    if (nlayer > 0) {
        std::vector<uint8_t> inq(nfluid, 0);
        for (int64_t i = nbulk; i < nfluid; ++i) inq[i] = 1;
        if (cfg_.pml_ref)
            for (int64_t i = 0; i < nlayer; ++i) inq[i] = 1;
        std::vector<int32_t> rq;
        for (int64_t i = 0; i < nfluid; ++i)
            if (inq[i]) rq.push_back(int32_t(i));
        n_rq_ = int64_t(rq.size());
        rq_list_ = to_device(rq, st);
    }
    // state
    const size_t rows = size_t(N) * RS_;
    if (nbnd == 0) {
        std::vector<uint8_t> none(std::max<int64_t>(nfluid, 1), 0);
        bsrc_ = to_device(none, st);
    }
    for (DeviceBuffer<float>* b : {&g_, &A_, &B_, &acc_, &Qp_, &Q2v_, &coll_, &Gf_}) {
        b->allocate(rows);
        b->zero(st);
    }
    // the previous step's (rho - 1, u) for the PML terms: zero is the rest
    // state, whose equilibrium deviation is zero
    E_.allocate(size_t(N) * 4);
    E_.zero(st);
    rho_.allocate(N);
    u_.allocate(N * D);
    ru_.allocate(N);
    nparts_ = int(std::max<int64_t>(pml_parts(Q, k, N, nfluid, n_rq_), 1));
    part_lo_ = int(std::min<int64_t>(pml_part_lo(Q, k, N, nbulk), nparts_));
    partial_.allocate(size_t(nparts_) * 2 * Q);
    partial_.zero(st);
    means_.allocate(2 * means_stride(Q));
    means_.zero(st);

    geo_.N = N;
    geo_.nbulk = nbulk;
    geo_.nlayer = nlayer;
    geo_.nfluid = nfluid;
    ready_ = true;
    time_ = 0.0;
    frame_ = 0;
    refresh_macroscopic_state();
    marks.mark();  // PML lists and state done
    NAGA_HIP(hipStreamSynchronize(st));
    setup_ms_ = timer.ms();
    marks.read(setup_phase_ms_, kSetupPhases);
}

StepBuffers Engine::buffers() {
    StepBuffers S{};
    S.g = g_.get();
    S.A = A_.get();
    S.B = B_.get();
    S.acc = acc_.get();
    S.E = E_.get();
    S.Qp = Qp_.get();
    S.Q2v = Q2v_.get();
    S.slab = slab_.get();
    S.fslab = fslab_.get();
    S.bslab = bslab_.get();
    S.sigma = sigma_.get();
    S.normals = normals_.get();
    S.coll = coll_.get();
    S.ru = ru_.get();
    S.Gf = Gf_.get();
    S.rq_list = rq_list_.get();
    S.n_rq = n_rq_;
    S.bsrc = bsrc_.get();
    S.pml_list = pml_list_.get();
    S.n_pml = n_pml_;
    S.ptiles = ptiles_.get();
    S.n_ptiles = n_ptiles_;
    S.rho = rho_.get();
    S.u = u_.get();
    S.partial = partial_.get();
    S.means = means_.get();
    S.nparts = nparts_;
    S.part_lo = part_lo_;
    S.src = sources_.empty() ? nullptr : src_field_.get();
    S.ulist = ulist_.get();
    S.uoff = uoff_.get();
    S.usplit = usplit_.get();
    S.split_tiles = split_tiles_.get();
    S.n_split = n_split_;
    S.p2slab = p2slab_.get();
    S.uslab = uslab_.get();
    S.ufslab = ufslab_.get();
    S.p2ulist = p2ulist_.get();
    S.p2uoff = p2uoff_.get();
    S.p2usplit = p2usplit_.get();
    S.p2split_tiles = p2split_tiles_.get();
    S.p2n_split = p2n_split_;
    S.bulist = bulist_.get();
    S.buoff = buoff_.get();
    S.ucap = ucap_;
    S.ucap_f = ucap_f_;
    S.self_first = self_first_;
    S.tile_ng = tile_ng_.size() ? tile_ng_.get() : nullptr;
    S.rec_off = rec_off_.size() ? rec_off_.get() : nullptr;
    S.frec_off = frec_off_.size() ? frec_off_.get() : nullptr;
    return S;
}

// Density sources (detail/simulation_engine.cuh:347-364, 1057-1107): the
// nodes inside any source region, with their (source, profile) entries in
// source order, are listed once; each step a small kernel writes the
// step's source value of every listed node into src_field_, read by p1.
void Engine::add_density_source(const DensitySource& src) {
    require_domain();
    require(int(sources_.size()) < kMaxSources, "at most 16 density sources per engine");
    require(src.radius > 0.0, "density source radius must be positive");
    sources_.push_back(src);
    const int D = D_;
    std::vector<int32_t> nodes, off{0}, ids;
    std::vector<float> prof;
    for (int64_t i = 0; i < N_; ++i) {
        bool any = false;
        for (size_t j = 0; j < sources_.size(); ++j) {
            const DensitySource& S = sources_[j];
            double r2 = 0.0;
            for (int d = 0; d < D; ++d) {
                const double dx = double(hpts_[i * D + d]) - S.center[d];
                r2 += dx * dx;
            }
            const double r = std::sqrt(r2);
            if (r < S.radius) {
                const double x = 2.0 * r / S.radius;
                ids.push_back(int32_t(j));
                prof.push_back(float((1.0 - x * x / 4.0) * std::exp(-x * x)));
                any = true;
            }
        }
        if (any) {
            nodes.push_back(int32_t(i));
            off.push_back(int32_t(ids.size()));
        }
    }
    n_snodes_ = int64_t(nodes.size());
    snodes_ = to_device(nodes, stream_);
    soff_ = to_device(off, stream_);
    sids_ = to_device(ids.empty() ? std::vector<int32_t>{0} : ids, stream_);
    sprof_ = to_device(prof.empty() ? std::vector<float>{0.0f} : prof, stream_);
    if (src_field_.size() != size_t(N_)) src_field_.allocate(N_);
    src_field_.zero(stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    // the batch's scene table carries the source field pointer
    if (batch) batch->invalidate();
}

void Engine::clear_density_sources() {
    require_domain();
    sources_.clear();
    n_snodes_ = 0;
    // the batch's scene table still points at the source field: rebuild it
    // with none (else the last step's source values would be added again)
    if (batch) batch->invalidate();
}

void Engine::refresh_macroscopic_state() {
    if (cfg_.Q == 27)
        launch_macroscopic<27>(geo_, g_.get(), rho_.get(), nullptr, u_.get(), stream_);
    else
        launch_macroscopic<9>(geo_, g_.get(), rho_.get(), nullptr, u_.get(), stream_);
}

SceneDesc Engine::scene_desc() {
    require_domain();
    SceneDesc d{};
    d.G = geo_;
    d.S = buffers();
    d.cp = cp_;
    return d;
}

void Engine::prepare_batched_step(hipStream_t stream) {
    if (sources_.empty()) return;
    SourceAmps amps{};
    for (size_t j = 0; j < sources_.size(); ++j) {
        const DensitySource& D = sources_[j];
        const bool on = time_ >= D.t_start && time_ < D.t_end;
        amps.a[j] = on ? float(D.amplitude * std::cos(2.0 * M_PI * D.frequency * time_ + D.phase))
                       : 0.0f;
    }
    launch_source_field(snodes_.get(), soff_.get(), sids_.get(), sprof_.get(), n_snodes_, amps,
                        src_field_.get(), stream);
}

// density_observer (simulation_observers/density_observer.cuh:62-113):
// kNN of the probe points among all nodes (k = 32) and RPIM weights at the
// nodal spacing; a sample interpolates rho - rho0 and adds rho0.
int Engine::add_probes(const float* points, int64_t n, int rbf) {
    require_domain();
    require(points != nullptr && n > 0, "no probe points");
    require(rbf == 0 || rbf == 1, "rbf must be 0 (multiquadric) or 1 (r^3)");
    auto ps = std::make_unique<ProbeSet>();
    ps->n = n;
    ps->k = int(std::min<int64_t>(32, N_));
    const int k = ps->k;
    std::vector<float> d2(size_t(n) * k);
    std::vector<uint32_t> idx(size_t(n) * k);
    knn_host(D_, hpts_.data(), N_, points, n, k, d2.data(), idx.data(), stream_);
    ps->idx = to_device(idx, stream_);
    DeviceBuffer<float> q(size_t(n) * D_);
    q.upload(points, size_t(n) * D_, stream_);
    ps->w.allocate(size_t(n) * k);
    DeviceBuffer<int> info(n);
    launch_rpim_weights(D_, pts_.get(), q.get(), n, ps->idx.get(), k, rbf, cfg_.h, ps->w.get(),
                        nullptr, info.get(), stream_);
    std::vector<int> hinfo(n);
    info.download(hinfo.data(), n, stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    for (int v : hinfo) require(v == 0, "singular RBF system for a probe point");
    ps->out.allocate(n);
    probes_.push_back(std::move(ps));
    return int(probes_.size()) - 1;
}

void Engine::sample_probes(int set, bool fresh, double* values) {
    require_domain();
    require(set >= 0 && set < int(probes_.size()), "unknown probe set");
    require(values != nullptr, "null output");
    ProbeSet& ps = *probes_[set];
    launch_probe(cfg_.Q, ps.n, ps.k, ps.idx.get(), ps.w.get(), fresh ? nullptr : rho_.get(),
                 g_.get(), cfg_.rho0, ps.out.get(), stream_);
    ps.out.download(values, ps.n, stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
}

EngineState Engine::snapshot() {
    require_domain();
    EngineState st;
    const size_t rows = size_t(N_) * RS_;
    auto copy = [&](DeviceBuffer<float>& dst, const DeviceBuffer<float>& src, size_t n) {
        dst.allocate(n);
        NAGA_HIP(hipMemcpyAsync(dst.get(), src.get(), n * sizeof(float), hipMemcpyDeviceToDevice,
                                stream_));
    };
    copy(st.g, g_, rows);
    copy(st.E, E_, size_t(N_) * 4);
    copy(st.Qp, Qp_, rows);
    copy(st.rho, rho_, size_t(N_));
    copy(st.u, u_, size_t(N_) * D_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    st.time = time_;
    st.frame = frame_;
    return st;
}

void Engine::restore(const EngineState& st) {
    require_domain();
    const size_t rows = size_t(N_) * RS_;
    require(st.g.size() == rows && st.rho.size() == size_t(N_), "state of another scene");
    auto copy = [&](DeviceBuffer<float>& dst, const DeviceBuffer<float>& src) {
        NAGA_HIP(hipMemcpyAsync(dst.get(), src.get(), src.size() * sizeof(float),
                                hipMemcpyDeviceToDevice, stream_));
    };
    copy(g_, st.g);
    copy(E_, st.E);
    copy(Qp_, st.Qp);
    copy(rho_, st.rho);
    copy(u_, st.u);
    NAGA_HIP(hipStreamSynchronize(stream_));
    time_ = st.time;
    frame_ = st.frame;
}

namespace {
struct StateHeader {
    char magic[8];
    int32_t Q, D, RS, pad;
    int64_t N, frame;
    double time;
};
// version 02: the PML history is the previous step's (rho - 1, u) per node
constexpr char kStateMagic[8] = {'N', 'A', 'G', 'A', 'S', 'T', '0', '2'};
}  // namespace

int64_t Engine::state_bytes() const {
    require_domain();
    const int64_t rows = N_ * RS_;
    return int64_t(sizeof(StateHeader)) +
           int64_t(sizeof(float)) * (2 * rows + 4 * N_ + N_ + N_ * D_);
}

void Engine::save_state(void* buf, int64_t capacity) {
    require(buf != nullptr && capacity >= state_bytes(), "state buffer too small");
    StateHeader h{};
    std::memcpy(h.magic, kStateMagic, 8);
    h.Q = cfg_.Q;
    h.D = D_;
    h.RS = RS_;
    h.N = N_;
    h.frame = frame_;
    h.time = time_;
    char* p = static_cast<char*>(buf);
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (DeviceBuffer<float>* b : {&g_, &E_, &Qp_}) {
        b->download(reinterpret_cast<float*>(p), b->size(), stream_);
        p += b->size() * sizeof(float);
    }
    rho_.download(reinterpret_cast<float*>(p), size_t(N_), stream_);
    p += size_t(N_) * sizeof(float);
    u_.download(reinterpret_cast<float*>(p), size_t(N_) * D_, stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
}

void Engine::load_state(const void* buf, int64_t bytes) {
    require(buf != nullptr && bytes >= int64_t(sizeof(StateHeader)), "state buffer too small");
    StateHeader h{};
    std::memcpy(&h, buf, sizeof(h));
    require(std::memcmp(h.magic, kStateMagic, 8) == 0, "not a naga_amd engine state");
    require(h.Q == cfg_.Q && h.N == N_ && h.RS == RS_ && h.D == D_,
            "state was saved from a different scene");
    require(bytes >= state_bytes(), "state buffer truncated");
    const char* p = static_cast<const char*>(buf) + sizeof(h);
    for (DeviceBuffer<float>* b : {&g_, &E_, &Qp_}) {
        b->upload(reinterpret_cast<const float*>(p), b->size(), stream_);
        p += b->size() * sizeof(float);
    }
    rho_.upload(reinterpret_cast<const float*>(p), size_t(N_), stream_);
    p += size_t(N_) * sizeof(float);
    u_.upload(reinterpret_cast<const float*>(p), size_t(N_) * D_, stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    time_ = h.time;
    frame_ = h.frame;
}

void Engine::finish_batched_step() {
    time_ += cfg_.dt;
    ++frame_;
}

void Engine::launch_one_step(hipEvent_t* ev) {
    const StepBuffers S = buffers();
    if (!sources_.empty()) {
        SourceAmps amps{};
        for (size_t j = 0; j < sources_.size(); ++j) {
            const DensitySource& D = sources_[j];
            const bool on = time_ >= D.t_start && time_ < D.t_end;
            amps.a[j] = on ? float(D.amplitude * std::cos(2.0 * M_PI * D.frequency * time_ +
                                                          D.phase))
                           : 0.0f;
        }
        launch_source_field(snodes_.get(), soff_.get(), sids_.get(), sprof_.get(), n_snodes_,
                            amps, src_field_.get(), stream_);
    }
    if (cfg_.Q == 27)
        launch_step<27>(geo_, S, cp_, stream_, ev);
    else
        launch_step<9>(geo_, S, cp_, stream_, ev);
    time_ += cfg_.dt;
    ++frame_;
}

void Engine::drop_step_graph() {
    if (step_exec_) (void)hipGraphExecDestroy(step_exec_);
    if (step_graph_) (void)hipGraphDestroy(step_graph_);
    step_exec_ = nullptr;
    step_graph_ = nullptr;
}

void Engine::launch_steps(int nsteps) {
    static const bool graphs = [] {
        const char* e = std::getenv("NAGA_ENGINE_GRAPH");
        return !(e && e[0] == '0');
    }();
    // density sources change kernel arguments every step: direct launches
    if (!graphs || !sources_.empty() || nsteps < 2) {
        for (int n = 0; n < nsteps; ++n) launch_one_step(nullptr);
        return;
    }
    if (!step_exec_) {
        // the step's kernels take the same arguments every step (buffers,
        // geometry, rates); the clock is advanced on the host per replay
        bool ok = hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed) == hipSuccess;
        if (ok) {
            capture_guarded(stream_, [&] {
                const StepBuffers S = buffers();
                if (cfg_.Q == 27)
                    launch_step<27>(geo_, S, cp_, stream_, nullptr);
                else
                    launch_step<9>(geo_, S, cp_, stream_, nullptr);
            });
            ok = hipStreamEndCapture(stream_, &step_graph_) == hipSuccess && step_graph_;
        }
        ok = ok && hipGraphInstantiate(&step_exec_, step_graph_, nullptr, nullptr, 0) ==
                       hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            drop_step_graph();
            for (int n = 0; n < nsteps; ++n) launch_one_step(nullptr);
            return;
        }
    }
    for (int n = 0; n < nsteps; ++n) {
        NAGA_HIP(hipGraphLaunch(step_exec_, stream_));
        time_ += cfg_.dt;
        ++frame_;
    }
}

void Engine::step(int nsteps) {
    require_domain();
    require(nsteps >= 0, "nsteps must be >= 0");
    launch_steps(nsteps);
    NAGA_HIP(hipStreamSynchronize(stream_));
}

double Engine::step_timed(int nsteps) {
    require_domain();
    hipEvent_t a, b;
    NAGA_HIP(hipEventCreate(&a));
    NAGA_HIP(hipEventCreate(&b));
    NAGA_HIP(hipEventRecord(a, stream_));
    launch_steps(nsteps);
    NAGA_HIP(hipEventRecord(b, stream_));
    NAGA_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    NAGA_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms);
}

void Engine::profile_step(double* ms) {
    require_domain();
    hipEvent_t ev[kStepKernels + 1];
    for (auto& e : ev) NAGA_HIP(hipEventCreate(&e));
    launch_one_step(ev);
    NAGA_HIP(hipEventSynchronize(ev[kStepKernels]));
    for (int t = 0; t < kStepKernels; ++t) {
        float v = 0.0f;
        NAGA_HIP(hipEventElapsedTime(&v, ev[t], ev[t + 1]));
        ms[t] = double(v);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
Batch::Batch() : q9_three_(q9_three(9, 1)) {
    NAGA_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
}

Batch::~Batch() {
    for (Engine* e : scenes_) e->batch = nullptr;
    drop_graph();
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void Batch::drop_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
}

void Batch::add(Engine* e) {
    require(e != nullptr && e->has_domain(), "a batch takes engines with a domain");
    require(e->batch == nullptr, "engine already in a batch");
    scenes_.push_back(e);
    e->batch = this;
    dirty_ = true;
}

void Batch::remove(Engine* e) {
    for (size_t i = 0; i < scenes_.size(); ++i)
        if (scenes_[i] == e) {
            scenes_.erase(scenes_.begin() + i);
            e->batch = nullptr;
            dirty_ = true;
            return;
        }
    fail(kInvalidArgument, "engine is not in the batch");
}

void Batch::set_graph(bool on) {
    use_graph_ = on;
    dirty_ = true;
}

void Batch::set_q9_kernels(int lanes, bool unions) {
    require(lanes == 3 || lanes == 9, "D2Q9 lanes per node must be 3 or 9");
    require(lanes == 3 || !unions, "D2Q9 tile unions need 3 lanes per node");
    q9_three_ = lanes == 3;
    q9_unions_ = unions;
    dirty_ = true;
}

void Batch::tuning(int* gather_batch, int* q9_lanes, int* q9_unions, int* graph) const {
    if (gather_batch) *gather_batch = gather_batch_;
    if (q9_lanes) *q9_lanes = q9_three_ ? 3 : 9;
    if (q9_unions) *q9_unions = q9_unions_ ? 1 : 0;
    if (graph) *graph = use_graph_ ? 1 : 0;
}

int Batch::autotune(int trial_steps) {
    require(trial_steps > 0, "trial_steps must be positive");
    if (scenes_.empty()) return 0;
    std::vector<EngineState> saved;
    saved.reserve(scenes_.size());
    for (Engine* e : scenes_) saved.push_back(e->snapshot());
    double best = 1e300;
    int best_gb = 8;
    bool best_graph = true;
    std::pair<bool, bool> best_q9{q9_three_, q9_unions_};
    bool any_q9 = false, any_q9_unions = false;
    for (Engine* e : scenes_)
        if (e->lattice() == 9) {
            any_q9 = true;
            any_q9_unions = any_q9_unions || e->scene_desc().S.ulist != nullptr;
        }
    // D2Q9 forms: (three per lane, unions); the union form where some
    // scene has its tile unions built
    std::vector<std::pair<bool, bool>> q9_forms{best_q9};
    if (any_q9) q9_forms = {{true, false}, {false, false}};
    if (any_q9_unions) q9_forms.push_back({true, true});
    for (const auto& form : q9_forms)
        for (int gb : {8, 16, 32})
            for (bool graph : {true, false}) {
                gather_batch_ = gb;
                use_graph_ = graph;
                q9_three_ = form.first;
                q9_unions_ = form.second;
                dirty_ = true;
                step(1);  // builds the plan (and captures the graph)
                const double ms = step_timed(trial_steps);
                if (ms < best) {
                    best = ms;
                    best_gb = gb;
                    best_graph = graph;
                    best_q9 = form;
                }
            }
    gather_batch_ = best_gb;
    use_graph_ = best_graph;
    q9_three_ = best_q9.first;
    q9_unions_ = best_q9.second;
    dirty_ = true;
    for (size_t i = 0; i < scenes_.size(); ++i) scenes_[i]->restore(saved[i]);
    return best_gb;
}

// Scenes are grouped by lattice and kernel variant (the four-per-lane D3Q27
// kernels by stencil width; D2Q9 scenes share the runtime-k kernels unless
// all have the same specialised width, and the lanes per node).
void Batch::rebuild() {
    drop_graph();
    groups_.clear();
    any_sources_ = false;
    // groups: lattice, four-per-lane stencil width, and whether the tile
    // unions (and their padded slab layout) were built
    auto three_of = [this](Engine* e) {
        return e->lattice() == 9 && q9_three_ && q9_three_fits(e->nodes());
    };
    // whether the scene's passes read its tile unions: D3Q27 scenes with
    // unions always (their records hold union positions); D2Q9 scenes
    // with unions when the batch runs them three per lane with unions
    auto tiles_of = [&](Engine* e) {
        const bool built = e->scene_desc().S.ulist != nullptr;
        return e->lattice() == 27 ? built : built && three_of(e) && q9_unions_;
    };
    auto key_of = [&](Engine* e) {
        const int Q = e->lattice();
        const bool un = step_uses_unions(Q, e->stencil_k(), e->nodes());
        const bool tiles = tiles_of(e);
        const bool three = three_of(e);
        const bool ragged = e->scene_desc().S.tile_ng != nullptr;
        return Q * 1000 + (un ? e->stencil_k() : 0) + (tiles ? 500 : 0) + (three ? 250 : 0) +
               (ragged ? 125 : 0);
    };
    std::vector<int> keys;
    for (Engine* e : scenes_) {
        const int key = key_of(e);
        size_t g = 0;
        while (g < keys.size() && keys[g] != key) ++g;
        if (g == keys.size()) {
            keys.push_back(key);
            groups_.push_back(std::make_unique<Group>());
            groups_.back()->Q = e->lattice();
            groups_.back()->three = three_of(e);
            groups_.back()->tiles = tiles_of(e);
        }
        groups_[g]->members.push_back(e);
    }
    for (auto& gp : groups_) {
        Group& g = *gp;
        const int n = int(g.members.size());
        std::vector<SceneDesc> descs(n);
        std::vector<int32_t> bstart(size_t(kStepKernels) * (n + 1), 0);
        BatchPlan& P = g.plan;
        P = BatchPlan{};
        P.Q = g.Q;
        P.nscenes = n;
        P.three = g.three;
        int kcommon = g.members[0]->stencil_k();
        for (int s = 0; s < n; ++s) {
            Engine* e = g.members[s];
            descs[s] = e->scene_desc();
            if (e->stencil_k() != kcommon) kcommon = -1;
            P.ucap = std::max(P.ucap, descs[s].S.ucap);
            P.ucap_f = std::max(P.ucap_f, descs[s].S.ucap_f);
            P.kmax = std::max(P.kmax, e->stencil_k());
            any_sources_ = any_sources_ || descs[s].S.src != nullptr;
            int64_t nb[kStepKernels];
            step_block_counts(g.Q, descs[s].G, descs[s].S, nb, g.three ? 1 : 0);
            for (int p = 0; p < kStepKernels; ++p) {
                int32_t* row = bstart.data() + size_t(p) * (n + 1);
                const int64_t next = int64_t(row[s]) + nb[p];
                require(next < (int64_t(1) << 31), "batch too large for one launch");
                row[s + 1] = int32_t(next);
            }
        }
        for (int p = 0; p < kStepKernels; ++p) P.blocks[p] = bstart[size_t(p) * (n + 1) + n];
        P.unions = step_uses_unions(g.Q, g.members[0]->stencil_k(), g.members[0]->nodes());
        P.k_template = (kcommon == 48 || kcommon == 32 || kcommon == 16) ? kcommon : 0;
        P.gather_batch = gather_batch_;
        P.tiles = g.tiles;
        P.ragged = descs[0].S.tile_ng != nullptr;
        for (const SceneDesc& d : descs)
            require(g.Q == 9 ? (!P.tiles || d.S.ulist != nullptr) : (d.S.ulist != nullptr) == P.tiles,
                    "scenes with and without tile unions");
        require(!P.unions || P.k_template != 0, "four-per-lane scenes need a common k");
        g.table = to_device(descs, stream_);
        g.bstart = to_device(bstart, stream_);
        P.dev_scenes = g.table.get();
        P.dev_bstart = g.bstart.get();
    }
    NAGA_HIP(hipStreamSynchronize(stream_));
    dirty_ = false;
}

void Batch::launch_one_step() {
    for (auto& gp : groups_) {
        if (gp->Q == 27)
            launch_step_batch<27>(gp->plan, stream_);
        else
            launch_step_batch<9>(gp->plan, stream_);
    }
}

void Batch::step(int nsteps) {
    require(nsteps >= 0, "nsteps must be >= 0");
    if (scenes_.empty() || nsteps == 0) return;
    if (dirty_) rebuild();
    // steps with density sources change kernel arguments per step: no graph
    bool graph = use_graph_ && !any_sources_;
    if (graph && !exec_) {
        // Relaxed capture: other threads (a serving loop's setup) may
        // allocate and launch on their own streams meanwhile.  A capture
        // that fails anyway falls back to direct launches for this call.
        bool ok = hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed) == hipSuccess;
        if (ok) {
            capture_guarded(stream_, [&] { launch_one_step(); });
            ok = hipStreamEndCapture(stream_, &graph_) == hipSuccess && graph_ != nullptr;
        }
        ok = ok && hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            drop_graph();
            graph = false;
        }
    }
    for (int n = 0; n < nsteps; ++n) {
        if (graph) {
            NAGA_HIP(hipGraphLaunch(exec_, stream_));
        } else {
            for (Engine* e : scenes_) e->prepare_batched_step(stream_);
            launch_one_step();
        }
        for (Engine* e : scenes_) e->finish_batched_step();
    }
    NAGA_HIP(hipStreamSynchronize(stream_));
}

double Batch::step_timed(int nsteps) {
    hipEvent_t a, b;
    NAGA_HIP(hipEventCreate(&a));
    NAGA_HIP(hipEventCreate(&b));
    if (dirty_ && !scenes_.empty()) rebuild();
    NAGA_HIP(hipEventRecord(a, stream_));
    step(nsteps);
    NAGA_HIP(hipEventRecord(b, stream_));
    NAGA_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    NAGA_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return double(ms);
}

void Engine::get_macroscopic(float* rho, float* u, double* drho, bool fresh) {
    require_domain();
    const int64_t N = N_;
    hipStream_t st = stream_;
    if (fresh) {
        DeviceBuffer<float> r(N), uu(N * D_);
        DeviceBuffer<double> dr(N);
        if (cfg_.Q == 27)
            launch_macroscopic<27>(geo_, g_.get(), r.get(), dr.get(), uu.get(), st);
        else
            launch_macroscopic<9>(geo_, g_.get(), r.get(), dr.get(), uu.get(), st);
        if (rho) r.download(rho, N, st);
        if (u) uu.download(u, N * D_, st);
        if (drho) dr.download(drho, N, st);
        NAGA_HIP(hipStreamSynchronize(st));
        return;
    }
    std::vector<float> hr(N);
    rho_.download(hr.data(), N, st);
    if (u) u_.download(u, N * D_, st);
    NAGA_HIP(hipStreamSynchronize(st));
    if (rho) std::memcpy(rho, hr.data(), sizeof(float) * N);
    if (drho)
        for (int64_t i = 0; i < N; ++i) drho[i] = double(hr[i]) - cfg_.rho0;
}

void Engine::get_distributions(double* f, double* g) {
    require_domain();
    const int Q = cfg_.Q;
    std::vector<float> rows(size_t(N_) * RS_);
    g_.download(rows.data(), rows.size(), stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    for (int a = 0; a < Q; ++a) {
        const double w = weight(Q, a);
        for (int64_t i = 0; i < N_; ++i) {
            const double v = double(rows[size_t(i) * RS_ + a]);
            if (g) g[a * N_ + i] = v;
            if (f) f[a * N_ + i] = w + v;
        }
    }
}

void Engine::set_distributions(const double* f) {
    require_domain();
    const int Q = cfg_.Q;
    std::vector<float> rows(size_t(N_) * RS_, 0.0f);
    for (int a = 0; a < Q; ++a) {
        const double w = weight(Q, a);
        for (int64_t i = 0; i < N_; ++i) rows[size_t(i) * RS_ + a] = float(f[a * N_ + i] - w);
    }
    g_.upload(rows.data(), rows.size(), stream_);
    refresh_macroscopic_state();
    NAGA_HIP(hipStreamSynchronize(stream_));
}

void Engine::set_pulse(const double* center, double amp, double width) {
    require_domain();
    require(width > 0, "pulse width must be positive");
    const int Q = cfg_.Q, D = D_;
    std::vector<float> hp(size_t(N_) * D);
    pts_.download(hp.data(), hp.size(), stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    std::vector<float> rows(size_t(N_) * RS_, 0.0f);
    for (int64_t i = 0; i < N_; ++i) {
        double r2 = 0.0;
        for (int d = 0; d < D; ++d) {
            const double t = double(hp[i * D + d]) - center[d];
            r2 += t * t;
        }
        const double p = amp * std::exp(-r2 / (2.0 * width * width));
        for (int a = 0; a < Q; ++a) rows[size_t(i) * RS_ + a] = float(weight(Q, a) * p);
    }
    g_.upload(rows.data(), rows.size(), stream_);
    refresh_macroscopic_state();
    NAGA_HIP(hipStreamSynchronize(stream_));
    time_ = 0.0;
    frame_ = 0;
}

void Engine::get_stencil_sizes(uint8_t* k) {
    require_domain();
    require(k != nullptr, "null output");
    if (kvar_.size() == size_t(N_)) {
        kvar_.download(k, N_, stream_);
        NAGA_HIP(hipStreamSynchronize(stream_));
    } else {
        std::fill(k, k + N_, uint8_t(cfg_.k));
    }
}

void Engine::get_setup(uint32_t* idx, float* W, float* filt, uint32_t* bidx, float* bw) {
    require_domain();
    const int k = cfg_.k, kb = cfg_.kb;
    if (idx) idx_.download(idx, N_ * k, stream_);
    if (W) W_.download(W, N_ * k * D_, stream_);
    if (filt) filt_.download(filt, N_ * k, stream_);
    if (bidx && nbnd_) bidx_.download(bidx, nbnd_ * kb, stream_);
    if (bw && nbnd_) bw_.download(bw, nbnd_ * kb, stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
}

namespace {
// RFC 4648 base64 of n bytes.
void base64_append(std::string& out, const unsigned char* p, size_t n) {
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = (uint32_t(p[i]) << 16) | (uint32_t(p[i + 1]) << 8) | p[i + 2];
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += tab[(v >> 6) & 63];
        out += tab[v & 63];
    }
    if (i < n) {
        uint32_t v = uint32_t(p[i]) << 16;
        if (i + 1 < n) v |= uint32_t(p[i + 1]) << 8;
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += i + 1 < n ? tab[(v >> 6) & 63] : '=';
        out += '=';
    }
}

// One inline binary DataArray: a UInt32 byte count, then the values, both
// base64-encoded in one block.
template <class T>
void vtp_array(std::string& out, const char* type, const char* name, int comps,
               const std::vector<T>& v) {
    out += "        <DataArray type=\"";
    out += type;
    out += "\" Name=\"";
    out += name;
    out += "\" NumberOfComponents=\"" + std::to_string(comps) + "\" format=\"binary\">\n          ";
    std::vector<unsigned char> buf(4 + v.size() * sizeof(T));
    const uint32_t bytes = uint32_t(v.size() * sizeof(T));
    std::memcpy(buf.data(), &bytes, 4);
    if (!v.empty()) std::memcpy(buf.data() + 4, v.data(), v.size() * sizeof(T));
    base64_append(out, buf.data(), buf.size());
    out += "\n        </DataArray>\n";
}
}  // namespace

// The output of vtk_observer::update_all (simulation_observers/
// vtk_observer.cuh:99-247) for this scene: one VTK XML PolyData file, the
// nodes as vertices, point arrays f (f_alpha / w_alpha), density, velocity
// (3 components), absorption, normals (3), type (0 bulk, 1 layer, 2
// boundary); FP32 values, inline binary, uncompressed.
void Engine::write_vtp(const char* path, bool fresh) {
    require_domain();
    require(path != nullptr && path[0] != 0, "empty output path");
    const int Q = cfg_.Q, D = D_;
    const int64_t N = N_;
    std::vector<float> rho(N), u(size_t(N) * D);
    std::vector<double> f(size_t(N) * Q);
    get_macroscopic(rho.data(), u.data(), nullptr, fresh);
    get_distributions(f.data(), nullptr);
    std::vector<float> pts(size_t(N) * 3, 0.0f), fw(size_t(N) * Q), vel(size_t(N) * 3, 0.0f),
        nrm(size_t(N) * 3, 0.0f), absorb(N, 0.0f);
    std::vector<int32_t> type(N, 0), conn(N), offs(N);
    for (int64_t i = 0; i < N; ++i) {
        for (int d = 0; d < D; ++d) {
            pts[size_t(i) * 3 + d] = hpts_[size_t(i) * D + d];
            vel[size_t(i) * 3 + d] = u[size_t(d) * N + i];
        }
        for (int a = 0; a < Q; ++a) fw[size_t(i) * Q + a] = float(f[size_t(a) * N + i] / weight(Q, a));
        if (i >= nfluid_) {
            type[i] = 2;
            for (int d = 0; d < D; ++d)
                nrm[size_t(i) * 3 + d] = hnormals_[size_t(i - nfluid_) * D + d];
        } else if (i >= nbulk_) {
            type[i] = 1;
            absorb[i] = hsigma_[size_t(i - nbulk_)];
        }
        conn[i] = int32_t(i);
        offs[i] = int32_t(i + 1);
    }
    std::string x;
    x.reserve(size_t(N) * (Q + 12) * 6 + 4096);
    x += "<?xml version=\"1.0\"?>\n";
    x += "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\" "
         "header_type=\"UInt32\">\n  <PolyData>\n";
    x += "    <Piece NumberOfPoints=\"" + std::to_string(N) + "\" NumberOfVerts=\"" +
         std::to_string(N) + "\" NumberOfLines=\"0\" NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n";
    x += "      <PointData>\n";
    vtp_array(x, "Float32", "f", Q, fw);
    vtp_array(x, "Float32", "density", 1, rho);
    vtp_array(x, "Float32", "velocity", 3, vel);
    vtp_array(x, "Float32", "absorption", 1, absorb);
    vtp_array(x, "Float32", "normals", 3, nrm);
    vtp_array(x, "Int32", "type", 1, type);
    x += "      </PointData>\n      <Points>\n";
    vtp_array(x, "Float32", "Points", 3, pts);
    x += "      </Points>\n      <Verts>\n";
    vtp_array(x, "Int32", "connectivity", 1, conn);
    vtp_array(x, "Int32", "offsets", 1, offs);
    x += "      </Verts>\n    </Piece>\n  </PolyData>\n</VTKFile>\n";
    FILE* fp = std::fopen(path, "wb");
    if (!fp) fail(kInvalidArgument, std::string("cannot open ") + path + " for writing");
    const size_t wrote = std::fwrite(x.data(), 1, x.size(), fp);
    const int closed = std::fclose(fp);
    if (wrote != x.size() || closed != 0) fail(kDeviceError, std::string("short write to ") + path);
}

// feq_a - w_a at ru = (rho - 1, u), as the PML kernels form it.
template <int Q>
static double feq_dev_host(const float* ru, int a) {
    using L = Lattice<Q>;
    const double cs2 = kCsLat * kCsLat;
    const double rt = double(ru[0]), rho = 1.0 + rt;
    double uc = 0.0, uu = 0.0;
    for (int d = L::D - 1; d >= 0; --d) uu += double(ru[1 + d]) * double(ru[1 + d]);
    for (int d = 0; d < L::D; ++d) uc += double(ru[1 + d]) * double(L::c(a, d));
    return L::w(a) * (rt + rho * (uc / cs2 + (uc * uc) / (2.0 * cs2 * cs2) - uu));
}

void Engine::get_pml(double* E, double* Qp) {
    require_domain();
    const int Q = cfg_.Q;
    std::vector<float> e(size_t(N_) * 4), q(size_t(N_) * RS_);
    E_.download(e.data(), e.size(), stream_);
    Qp_.download(q.data(), q.size(), stream_);
    NAGA_HIP(hipStreamSynchronize(stream_));
    for (int a = 0; a < Q; ++a)
        for (int64_t i = 0; i < N_; ++i) {
            if (E) E[a * N_ + i] = Q == 27 ? feq_dev_host<27>(&e[size_t(i) * 4], a)
                                           : feq_dev_host<9>(&e[size_t(i) * 4], a);
            if (Qp) Qp[a * N_ + i] = double(q[size_t(i) * RS_ + a]);
        }
}

}  // namespace naga_amd
