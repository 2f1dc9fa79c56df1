// One scene of the nonlocal LBM resident on the GPU: setup products, state
// and the step schedule.  Mirrors detail::simulation_engine<Lattice>
// (detail/simulation_engine.cuh:85-1486) behind the C-ABI.
#pragma once

#include <memory>
#include <vector>

#include "device_buffer.hpp"
#include "step_kernels.hpp"

namespace naga_amd {

struct SceneConfig {
    int Q = 27;
    int k = 32;
    int kb = 32;
    int rbf = 1;
    int rbf_boundary = 1;
    int pml_ref = 1;
    int k_min = 0;  // > 0: per-node stencil widths in [k_min, k] by local density
    double nu = 0.0, rho0 = 1.0, dt = 0.0, c = kCsLat, h = 0.01;
    double filter = 0.0;  // 0: the reference's 0.75; > 0: that strength; < 0: off
    int solve_bits = 64;  // precision of the RBF weight solves: 64 or 32
};

// A declarative density source: amplitude * cos(2 pi frequency t + phase)
// times the radial profile (1 - (2r/R)^2 / 4) exp(-(2r/R)^2) within radius
// R of the centre, for t_start <= t < t_end.
class Batch;

// The state of a scene that carries from one step to the next: the
// distributions (as deviations), the PML history, the macroscopic values
// held for solution(), the clock.
struct EngineState {
    DeviceBuffer<float> g, E, Qp, rho, u;
    double time = 0.0;
    int64_t frame = 0;
};

struct DensitySource {
    double center[3] = {0.0, 0.0, 0.0};
    double radius = 0.0, amplitude = 0.0, frequency = 0.0, phase = 0.0;
    double t_start = 0.0, t_end = 0.0;
};

class Engine {
  public:
    explicit Engine(const SceneConfig& cfg);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void init_domain(const float* pts, int64_t nbulk, int64_t nlayer, int64_t nbnd,
                     const float* normals, const float* sigma);
    void set_pulse(const double* center, double amp, double width);
    void set_distributions(const double* f);
    void step(int nsteps);
    double step_timed(int nsteps);
    void profile_step(double* ms);
    void get_macroscopic(float* rho, float* u, double* drho, bool fresh);
    void get_distributions(double* f, double* g);
    void get_setup(uint32_t* idx, float* W, float* filt, uint32_t* bidx, float* bw);
    void get_stencil_sizes(uint8_t* k);
    void get_pml(double* E, double* Qp);
    void add_density_source(const DensitySource& src);
    // density probes: kNN (k = 32, or fewer nodes) among all nodes and
    // RPIM weights at spacing h; returns the probe set's index
    int add_probes(const float* points, int64_t n, int rbf);
    void sample_probes(int set, bool fresh, double* values);
    // device-side copy of the state and its restoration
    EngineState snapshot();
    void restore(const EngineState& st);
    // host serialisation of the state: header + buffers
    int64_t state_bytes() const;
    void save_state(void* buf, int64_t capacity);
    void load_state(const void* buf, int64_t bytes);
    void clear_density_sources();
    // The vtk_observer output of the current scene state (fresh: the state
    // now; else the macroscopic values held from the last step's start).
    void write_vtp(const char* path, bool fresh);

    // Descriptor of this scene for batched launches.
    SceneDesc scene_desc();
    // Per-step host work of a batched step: the density-source field of
    // this step (on `stream`), and the clock advance after it.
    void prepare_batched_step(hipStream_t stream);
    void finish_batched_step();
    bool has_domain() const { return ready_; }
    Batch* batch = nullptr;  // the batch holding this scene, if any
    int stencil_k() const { return cfg_.k; }
    int64_t fluid_nodes() const { return nfluid_; }

    int64_t nodes() const { return N_; }
    double time() const { return time_; }
    int64_t frame() const { return frame_; }
    double setup_ms() const { return setup_ms_; }
    // Setup phases (ms on the engine's stream): binning and kNN, divergence
    // weights, filter weights + slabs + tile unions, boundary interpolator,
    // PML lists and state.
    static constexpr int kSetupPhases = 5;
    const double* setup_phase_ms() const { return setup_phase_ms_; }
    int lattice() const { return cfg_.Q; }
    int dims() const { return D_; }
    hipStream_t stream() const { return stream_; }

  private:
    void require_domain() const;
    void launch_one_step(hipEvent_t* ev);
    // nsteps steps on the stream: one step captured once into a hipGraph and
    // replayed (no density sources; NAGA_ENGINE_GRAPH=0 disables), else
    // direct launches
    void launch_steps(int nsteps);
    void drop_step_graph();
    void refresh_macroscopic_state();
    StepBuffers buffers();
    hipGraph_t step_graph_ = nullptr;
    hipGraphExec_t step_exec_ = nullptr;

    SceneConfig cfg_;
    int D_ = 3, RS_ = 32;
    int64_t N_ = 0, nbulk_ = 0, nlayer_ = 0, nbnd_ = 0, nfluid_ = 0;
    hipStream_t stream_ = nullptr;
    StepGeom geo_{};
    CollisionParams cp_{};
    double time_ = 0.0, setup_ms_ = 0.0;
    double setup_phase_ms_[kSetupPhases] = {};
    int64_t frame_ = 0;
    bool ready_ = false;

    DeviceBuffer<float> pts_, normals_, sigma_;
    DeviceBuffer<float> slab_, fslab_, bslab_;
    DeviceBuffer<uint32_t> idx_, bidx_;
    DeviceBuffer<uint8_t> kvar_;  // per-node stencil widths (empty: all k)
    DeviceBuffer<float> W_, filt_, bw_;
    DeviceBuffer<int32_t> pml_list_;
    DeviceBuffer<int32_t> ptiles_;
    DeviceBuffer<uint8_t> bsrc_;
    DeviceBuffer<float> coll_, Gf_;
    DeviceBuffer<float4> ru_;
    DeviceBuffer<int32_t> rq_list_;
    int64_t n_rq_ = 0;
    DeviceBuffer<float> g_, A_, B_, acc_, E_, Qp_, Q2v_;
    DeviceBuffer<float> rho_, u_;
    DeviceBuffer<double> partial_;
    DeviceBuffer<float> means_;
    int nparts_ = 1;
    int part_lo_ = 0;
    int64_t n_pml_ = 0;
    int64_t n_ptiles_ = 0;
    int tile_nodes_ = 0;  // nodes per union tile (0: no unions)

    DeviceBuffer<uint32_t> ulist_;  // tile unions (four-per-lane D3Q27 kernels)
    DeviceBuffer<int64_t> uoff_;
    DeviceBuffer<int32_t> usplit_;  // per tile: first half-union's length if split, else 0
    DeviceBuffer<int32_t> split_tiles_;  // the split tiles, in order
    std::vector<int32_t> split_of_;      // per tile: its index among the split tiles, or -1
    int64_t n_split_ = 0;
    // the PML divergence's unions (layer neighbours and one zero row) and
    // its packed records holding positions in them
    DeviceBuffer<uint32_t> p2ulist_;
    DeviceBuffer<int64_t> p2uoff_;
    DeviceBuffer<int32_t> p2usplit_, p2split_tiles_;
    std::vector<int32_t> p2split_of_;
    int64_t p2n_split_ = 0;
    DeviceBuffer<float> p2slab_;
    // boundary-interpolation tiles: their distinct source nodes
    DeviceBuffer<uint32_t> bulist_;
    DeviceBuffer<int64_t> buoff_;
    DeviceBuffer<float> uslab_, ufslab_;  // D2Q9 records with union positions
    int ucap_ = 0;
    int ucap_f_ = 0;
    int self_first_ = 0;
    DeviceBuffer<int32_t> tile_ng_;   // ragged record rows: groups of 8 per tile
    DeviceBuffer<int64_t> rec_off_;   // first unit of each tile's divergence (and PML) records
    DeviceBuffer<int64_t> frec_off_;  // first unit of each tile's filter records

    struct ProbeSet {
        int64_t n = 0;
        int k = 0;
        DeviceBuffer<uint32_t> idx;
        DeviceBuffer<float> w;
        DeviceBuffer<double> out;
    };
    std::vector<std::unique_ptr<ProbeSet>> probes_;

    std::vector<float> hpts_;  // host copy of the node coordinates
    std::vector<float> hnormals_, hsigma_;  // host copies of the normals and absorption
    std::vector<DensitySource> sources_;
    DeviceBuffer<int32_t> snodes_, soff_, sids_;
    DeviceBuffer<float> sprof_, src_field_;
    int64_t n_snodes_ = 0;
};

// Many scenes stepped together: every pass of a step is one launch over all
// scenes of a group (scenes sharing lattice and kernel variant), optionally
// replayed from a captured hipGraph.  Scenes stay owned by their engines.
class Batch {
  public:
    Batch();
    ~Batch();
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    void add(Engine* e);
    void remove(Engine* e);
    void step(int nsteps);
    double step_timed(int nsteps);
    void set_graph(bool on);
    // Times trial_steps steps for each gather batch (8, 16, 32), with and
    // without graph replay, and, when the batch holds D2Q9 scenes, for each
    // D2Q9 kernel form: three distributions per lane (3 lanes per node,
    // 84-node tiles) reading the tile unions staged in LDS, the same
    // gathering directly, or one distribution per lane (9 lanes per node,
    // 28-node tiles); keeps the fastest and restores every scene's state
    // (the trials do not advance the scenes).  Returns the chosen gather
    // batch, or 0 if the batch is empty.
    int autotune(int trial_steps);
    // D2Q9 kernel form: lanes per node 3 or 9; unions (lanes 3 only): read
    // the tile unions in LDS where the scene has them.
    void set_q9_kernels(int lanes, bool unions);
    // The current choices: gather batch (0: the kernels' default), D2Q9
    // lanes per node and unions, graph replay.
    void tuning(int* gather_batch, int* q9_lanes, int* q9_unions, int* graph) const;
    int size() const { return int(scenes_.size()); }
    hipStream_t stream() const { return stream_; }
    // A member's scene description changed (density sources added or
    // cleared): the scene tables and the graph are rebuilt before the next
    // step.
    void invalidate() { dirty_ = true; }

  private:
    struct Group {
        int Q = 9;
        bool three = false;  // D2Q9 with three distributions per lane
        bool tiles = false;  // the members' tile unions are read (D3Q27: always with them)
        std::vector<Engine*> members;
        BatchPlan plan;
        DeviceBuffer<SceneDesc> table;
        DeviceBuffer<int32_t> bstart;
    };
    void rebuild();
    void launch_one_step();
    void drop_graph();

    std::vector<Engine*> scenes_;
    std::vector<std::unique_ptr<Group>> groups_;
    hipStream_t stream_ = nullptr;
    bool dirty_ = true;
    bool use_graph_ = true;
    bool any_sources_ = false;
    int gather_batch_ = 0;
    bool q9_three_ = true;   // D2Q9 scenes with three distributions per lane
    bool q9_unions_ = true;  // ... reading their tile unions in LDS
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// Exact kNN of queries among ref (all host arrays), on the GPU.
void knn_host(int D, const float* ref, int64_t n_ref, const float* query,
              int64_t n_q, int k, float* d2, uint32_t* idx, hipStream_t stream);

}  // namespace naga_amd
