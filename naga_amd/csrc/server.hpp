// Serving loop for acoustic-pulse solves with continuous batching.
//
// Requests are admitted while their estimated device footprint fits the
// HBM budget and a slot is free.  Setup threads build each admitted
// request's scene (node generation on the host, kNN and weights on the
// GPU on the scene's own stream) while a stepping thread advances every
// resident scene together through one Batch; scenes join at the next
// chunk boundary and leave as soon as their step count is reached, their
// results copied out and their memory released.  State and stencils stay
// on the device for the whole solve.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/naga_amd.h"
#include "engine.hpp"

namespace naga_amd {

struct ServedResult {
    bool ok = false;
    std::string error;
    int64_t nodes = 0;
    int64_t steps = 0;
    std::vector<float> rho;     // [N]
    std::vector<float> u;       // [D][N]
    std::vector<double> drho;   // [N]
    double latency_ms = 0.0;    // submit to result ready
    double queue_ms = 0.0;      // submit to admission
};

class Server {
  public:
    Server(size_t hbm_budget, int max_active, int chunk_steps, bool graph, int setup_threads);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int64_t submit(const naga_request& req);
    // 0 pending, 1 done (ok or failed), -1 unknown id
    int status(int64_t id);
    // Copies the result out (the server keeps it until released).
    bool result(int64_t id, ServedResult& out);
    void release(int64_t id);
    void drain();  // wait until every submitted request is done
    naga_server_stats stats();

    // device bytes a request is expected to hold while resident
    static size_t footprint(int Q, int k, int64_t n_nodes);

  private:
    struct Pending {
        int64_t id;
        naga_request req;
        double t_submit;
    };
    struct Active {
        int64_t id = 0;
        std::unique_ptr<Engine> eng;
        int64_t remaining = 0, steps = 0, nodes = 0;
        size_t bytes = 0;
        double t_submit = 0.0, t_admit = 0.0;
    };

    void setup_loop();
    void step_loop();
    void finish(Active& a, double now);
    double now_ms() const;

    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Pending> queue_;
    std::vector<Active> ready_;
    std::map<int64_t, ServedResult> done_;
    std::map<int64_t, int> state_;  // 0 pending, 1 done
    size_t budget_, used_ = 0;
    int max_active_, chunk_;
    int resident_ = 0;
    int64_t next_id_ = 1;
    bool stop_ = false;
    bool graph_;
    int64_t submitted_ = 0, completed_ = 0, failed_ = 0;
    double node_updates_ = 0.0, busy_ms_ = 0.0;
    double t_first_submit_ = -1.0, t_last_done_ = 0.0;
    std::vector<double> latencies_;
    std::chrono::steady_clock::time_point t0_;
    std::vector<std::thread> setup_threads_;
    std::thread step_thread_;
};

}  // namespace naga_amd
