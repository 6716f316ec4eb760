// A solve server on one GPU: requests are structured box problems; systems (the matrix and its
// multigrid hierarchy) stay resident in HBM under a byte budget and are shared by every request
// with the same mesh, constraints and material; the requests of one system are solved together
// as the columns of one lockstep batch, new requests taking free columns between slices of
// iterations (continuous batching).
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "sf_handles.hpp"

namespace sf {

struct ServeRequest {
    u64 id = 0;
    sf_workload workload{};
    std::shared_ptr<const Model> model;  // a model request (any mesh); null for a box workload
    std::string key;
    int state = SF_SERVE_QUEUED;
    Error error{SF_OK, ""};
    double submitted = 0, build_seconds = 0, solve_started = 0, finished = 0;
    bool cache_hit = false;
    u64 free_dofs = 0;
    sf_column_progress progress{};
    std::vector<double> displacements;  // 3 per mesh node, fixed components zero
};

struct ServeSystem {
    std::string key;
    sf_static_study *study = nullptr;   // the assembled matrix and the device hierarchy
    std::vector<Face> boundary;         // the mesh's boundary faces in order
    std::vector<Face> sorted_boundary;  // the same, sorted (a model request's traction faces are checked against it)
    std::deque<u64> waiting;            // requests of this system not yet in a batch
    bool ready = false, busy = false;
    u64 bytes = 0;                      // reserved HBM: the system and one full batch
    double last_used = 0;
};

class Server {
   public:
    Server(int device, const sf_server_options &options);
    ~Server();
    u64 submit(const sf_workload &w);
    // A model of any mesh: requests whose models share mesh, fixed nodes, prescribed components and
    // material share one system; their forces and tractions are their own.
    u64 submit_model(const Model &model);
    sf_serve_status status(u64 id);
    // Waits up to `timeout` seconds (negative: without limit) for the request to end.
    sf_serve_status wait(u64 id, double timeout);
    void result(u64 id, double *out, size_t len);
    void release(u64 id);
    sf_server_stats stats();

   private:
    void builder_loop();
    void solver_loop();
    void build(ServeSystem &system, const sf_workload &w, const std::shared_ptr<const Model> &model, sf_context &ctx);
    u64 enqueue(std::unique_ptr<ServeRequest> r);
    void serve(ServeSystem &system, const std::shared_ptr<Context> &ctx);
    void finish_request(ServeRequest &r, int state);
    sf_serve_status describe(const ServeRequest &r) const;
    u64 estimate(const ServeRequest &r) const;
    bool evict_for(u64 bytes, std::vector<sf_static_study *> &doomed);
    ServeRequest &request(u64 id);

    int device_;
    sf_server_options options_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
    u64 next_id_ = 1;
    std::map<u64, std::unique_ptr<ServeRequest>> requests_;
    std::deque<u64> unbuilt_;  // requests whose system is neither resident nor being built
    std::map<std::string, std::unique_ptr<ServeSystem>> systems_;
    std::vector<std::unique_ptr<sf_context>> contexts_;  // outlive every study built on them
    u64 in_use_ = 0, peak_ = 0;
    u64 submitted_ = 0, completed_ = 0, failed_ = 0, builds_ = 0, hits_ = 0, evictions_ = 0;
    std::vector<double> latencies_;
    std::vector<std::thread> threads_;
};

}  // namespace sf
