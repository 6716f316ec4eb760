// The solve server: admission under an HBM budget, resident systems shared by key, and batches
// of requests that admit new columns between slices of lockstep iterations.
#include "sf_serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace sf {

namespace {

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// Everything the matrix and the hierarchy depend on: the mesh, the constraint and the material.
std::string system_key(const sf_workload &w) {
    char buf[640];
    std::snprintf(buf, sizeof buf, "%a %a %a|%a %a %a|%d %d %d|%a %llu|%d %a|%d %a %a %a %a %a|%d|%d %a", w.origin[0], w.origin[1],
                  w.origin[2], w.size[0], w.size[1], w.size[2], w.divisions[0], w.divisions[1], w.divisions[2], w.jitter_fraction,
                  (unsigned long long)w.jitter_seed, w.clamp_axis, w.clamp_value, w.material_kind, w.youngs, w.poisson,
                  w.core_youngs, w.core_poisson, w.shell_depth, w.quadrature_points, w.wall_cells, w.wall_thickness);
    return buf;
}

// A model request's system: its mesh, fixed nodes, prescribed components, material and rule, as
// the bytes of an FNV-1a hash (with the counts, against collisions of unequal sizes).
std::string model_key(const Model &m) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 0x00000100000001b3ull;
        }
    };
    mix(m.mesh.nodes.data(), m.mesh.nodes.size() * sizeof(Vec3));
    mix(m.mesh.tets.data(), m.mesh.tets.size() * sizeof(m.mesh.tets[0]));
    for (u32 f : m.fixed) mix(&f, sizeof f);
    for (const auto &p : m.prescribed) {
        mix(&p.first.first, sizeof(u32));
        mix(&p.first.second, sizeof(u32));
        mix(&p.second, sizeof(double));
    }
    const MaterialField &f = m.field;
    mix(&f.kind, sizeof f.kind);
    for (double v : {f.e, f.nu, f.core_e, f.core_nu, f.shell_depth, f.lo[0], f.lo[1], f.lo[2], f.hi[0], f.hi[1], f.hi[2]}) mix(&v, sizeof v);
    mix(f.transverse, sizeof f.transverse);
    mix(&m.quadrature, sizeof m.quadrature);
    char buf[96];
    std::snprintf(buf, sizeof buf, "model %016llx %zu %zu %zu", (unsigned long long)h, m.mesh.nodes.size(), m.mesh.tets.size(),
                  m.prescribed.size());
    return buf;
}

void check(sf_status s) {
    if (s == SF_OK) return;
    int64_t a = 0, b = 0;
    double v = 0;
    sf_last_error_payload(&a, &b, &v);
    throw Error(s, sf_last_error(), a, b, v);
}

size_t vector_size(int precision) { return precision == SF_PRECISION_SINGLE ? 4 : 8; }

}  // namespace

Server::Server(int device, const sf_server_options &o) : device_(device), options_(o) {
    if (options_.hbm_budget == 0) options_.hbm_budget = 200ull << 30;
    if (options_.max_columns == 0) options_.max_columns = 32;
    if (options_.max_columns > MAX_COLUMNS) throw Error(SF_ERR_INVALID_ARGUMENT, "a batch carries 1 to 64 requests");
    if (options_.slice_iterations == 0) options_.slice_iterations = 8;
    if (options_.tolerance <= 0) options_.tolerance = 1e-8;
    if (options_.builders <= 0) options_.builders = 4;
    if (options_.solvers <= 0) options_.solvers = 8;
    if (options_.precision < 0 || options_.precision > 2) throw Error(SF_ERR_INVALID_ARGUMENT, "unknown precision");
    if (options_.coarse_solve != SF_COARSE_REFERENCE && options_.coarse_solve != SF_COARSE_INVERSE)
        throw Error(SF_ERR_INVALID_ARGUMENT, "unknown coarse solve");
    int count = 0;
    SF_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) throw Error(SF_ERR_INVALID_ARGUMENT, "no such device");
    for (int i = 0; i < options_.builders; ++i) threads_.emplace_back([this] { builder_loop(); });
    for (int i = 0; i < options_.solvers; ++i) threads_.emplace_back([this] { solver_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto &kv : systems_)
        if (kv.second->study) sf_static_study_destroy(kv.second->study);
    systems_.clear();
    contexts_.clear();
}

u64 Server::estimate(const ServeRequest &r) const {
    // A structured box of Tet10: (2 nx + 1)(2 ny + 1)(2 nz + 1) nodes with at most 27 blocks per
    // block row; a model's mesh: its nodes, and at most its 100 block pairs per element. The system
    // holds the assembled f64 blocks twice (row-major and sliced), the hierarchy about 0.6 of that
    // again in f64, and a full batch 14 vectors of every column.
    const sf_workload &w = r.workload;
    double nodes, blocks;
    if (r.model) {
        nodes = (double)r.model->mesh.nodes.size();
        blocks = std::min(100.0 * (double)r.model->mesh.tets.size(), nodes * nodes);
    } else {
        nodes = (2.0 * w.divisions[0] + 1) * (2.0 * w.divisions[1] + 1) * (2.0 * w.divisions[2] + 1);
        blocks = 27 * nodes;
    }
    const double dofs = 3 * nodes;
    const double system = blocks * (72 + 80) * 1.6 + dofs * 8 * 24;
    const double batch = dofs * options_.max_columns * (10 * 8 + 1.4 * 8 * vector_size(options_.precision));
    return (u64)(system + batch);
}

u64 Server::submit(const sf_workload &w) {
    auto r = std::make_unique<ServeRequest>();
    r->workload = w;
    r->key = system_key(w);
    return enqueue(std::move(r));
}

u64 Server::submit_model(const Model &model) {
    auto r = std::make_unique<ServeRequest>();
    r->model = std::make_shared<const Model>(model);
    r->key = model_key(model);
    return enqueue(std::move(r));
}

u64 Server::enqueue(std::unique_ptr<ServeRequest> r) {
    r->submitted = now();
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) throw Error(SF_ERR_INVALID_ARGUMENT, "the server is stopping");
    const u64 id = next_id_++;
    r->id = id;
    ++submitted_;
    auto it = systems_.find(r->key);
    if (it != systems_.end()) {
        r->cache_hit = true;
        ++hits_;
        it->second->waiting.push_back(id);
    } else {
        unbuilt_.push_back(id);
    }
    requests_[id] = std::move(r);
    cv_.notify_all();
    return id;
}

ServeRequest &Server::request(u64 id) {
    auto it = requests_.find(id);
    if (it == requests_.end()) throw Error(SF_ERR_SERVE_UNKNOWN_REQUEST, "no such request", (int64_t)id);
    return *it->second;
}

sf_serve_status Server::describe(const ServeRequest &r) const {
    sf_serve_status s{};
    s.state = r.state;
    s.error = r.error.status;
    s.cache_hit = r.cache_hit ? 1 : 0;
    s.at_rounding_floor = r.progress.at_rounding_floor;
    s.iterations = r.progress.iteration;
    s.relative_residual = r.progress.relative_residual;
    s.build_seconds = r.build_seconds;
    if (r.finished > 0) {
        s.latency_seconds = r.finished - r.submitted;
        if (r.solve_started > 0) s.solve_seconds = r.finished - r.solve_started;
        s.queued_seconds = std::max(0.0, s.latency_seconds - s.solve_seconds - s.build_seconds);
    }
    s.nodes = r.displacements.size() / 3;
    s.free_dofs = r.free_dofs;
    return s;
}

sf_serve_status Server::status(u64 id) {
    std::lock_guard<std::mutex> g(mu_);
    return describe(request(id));
}

sf_serve_status Server::wait(u64 id, double timeout) {
    std::unique_lock<std::mutex> g(mu_);
    ServeRequest *r = &request(id);
    auto done = [&] { return r->state == SF_SERVE_DONE || r->state == SF_SERVE_FAILED; };
    if (timeout < 0) cv_.wait(g, done);
    else cv_.wait_for(g, std::chrono::duration<double>(timeout), done);
    return describe(*r);
}

void Server::result(u64 id, double *out, size_t len) {
    std::lock_guard<std::mutex> g(mu_);
    ServeRequest &r = request(id);
    if (r.state == SF_SERVE_FAILED) throw r.error;
    if (r.state != SF_SERVE_DONE) throw Error(SF_ERR_MODEL_NOT_FINISHED, "the request has not finished", (int64_t)id);
    if (len != r.displacements.size())
        throw Error(SF_ERR_INVALID_ARGUMENT, "the output length is not 3 per node", (int64_t)r.displacements.size(), (int64_t)len);
    std::copy(r.displacements.begin(), r.displacements.end(), out);
}

void Server::release(u64 id) {
    std::lock_guard<std::mutex> g(mu_);
    ServeRequest &r = request(id);
    if (r.state != SF_SERVE_DONE && r.state != SF_SERVE_FAILED)
        throw Error(SF_ERR_INVALID_ARGUMENT, "the request is still open", (int64_t)id);
    requests_.erase(id);
}

sf_server_stats Server::stats() {
    std::lock_guard<std::mutex> g(mu_);
    sf_server_stats s{};
    s.submitted = submitted_;
    s.completed = completed_;
    s.failed = failed_;
    s.builds = builds_;
    s.cache_hits = hits_;
    s.evictions = evictions_;
    s.resident_systems = systems_.size();
    s.hbm_reserved = in_use_;
    s.hbm_peak = peak_;
    if (!latencies_.empty()) {
        std::vector<double> v = latencies_;
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto rank = [&](double q) { return v[std::min(v.size() - 1, (size_t)std::ceil(q * (double)v.size()) - 1)]; };
        s.p50_seconds = rank(0.50);
        s.p99_seconds = rank(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_seconds = sum / (double)v.size();
        s.max_seconds = v.back();
    }
    return s;
}

void Server::finish_request(ServeRequest &r, int state) {
    // called with the lock held
    r.state = state;
    r.finished = now();
    if (state == SF_SERVE_DONE) {
        ++completed_;
        latencies_.push_back(r.finished - r.submitted);
    } else {
        ++failed_;
    }
}

bool Server::evict_for(u64 bytes, std::vector<sf_static_study *> &doomed) {
    // called with the lock held: least recently used idle systems go first
    while (in_use_ + bytes > options_.hbm_budget) {
        ServeSystem *lru = nullptr;
        for (auto &kv : systems_) {
            ServeSystem &s = *kv.second;
            if (s.ready && !s.busy && s.waiting.empty() && (!lru || s.last_used < lru->last_used)) lru = &s;
        }
        if (!lru) return false;
        in_use_ -= lru->bytes;
        doomed.push_back(lru->study);
        ++evictions_;
        systems_.erase(lru->key);
    }
    return true;
}

void Server::build(ServeSystem &system, const sf_workload &w, const std::shared_ptr<const Model> &from, sf_context &ctx) {
    sf_model *model = nullptr;
    if (from) {
        auto m = std::make_unique<sf_model>();
        m->model = *from;
        m->view.mesh = m->model.mesh;
        model = m.release();
    } else {
        check(sf_model_from_workload(&w, &model));
    }
    std::unique_ptr<sf_model, void (*)(sf_model *)> hold(model, sf_model_destroy);
    const sf_pcg_options pcg{options_.tolerance, 0, options_.dot_order, 0};
    sf_multigrid_options mg{3.0, 1500, 12, 3, 30.0, options_.precision, 0, options_.coarse_solve, 0};
    sf_study_builder *b = nullptr;
    check(sf_study_builder_create(&ctx, model, &pcg, 1, &mg, &b));
    sf_build_progress p{};
    sf_status st = sf_study_builder_step(b, ~0ull, nullptr, nullptr, &p);
    if (st != SF_OK) {
        sf_study_builder_destroy(b);
        check(st);
    }
    sf_static_study *study = nullptr;
    st = sf_study_builder_finish(b, &study);  // consumes the builder when it succeeds
    if (st != SF_OK) {
        sf_study_builder_destroy(b);
        check(st);
    }
    system.study = study;
    system.boundary = study->model.mesh.boundary_faces();
    system.sorted_boundary = system.boundary;
    std::sort(system.sorted_boundary.begin(), system.sorted_boundary.end());
}

void Server::builder_loop() {
    std::unique_ptr<sf_context> ctx;
    try {
        ctx = std::make_unique<sf_context>();
        ctx->ctx = std::make_shared<Context>(device_);
    } catch (...) {
        return;
    }
    sf_context *raw = ctx.get();
    {
        std::lock_guard<std::mutex> g(mu_);
        contexts_.push_back(std::move(ctx));
    }
    for (;;) {
        ServeSystem *system = nullptr;
        sf_workload w{};
        std::shared_ptr<const Model> m;
        u64 first = 0;
        std::vector<sf_static_study *> doomed;
        {
            std::unique_lock<std::mutex> g(mu_);
            for (;;) {
                if (stop_) return;
                // the oldest request whose system is not yet resident or being built
                while (!unbuilt_.empty()) {
                    const u64 id = unbuilt_.front();
                    auto it = requests_.find(id);
                    if (it == requests_.end()) {
                        unbuilt_.pop_front();
                        continue;
                    }
                    ServeRequest &r = *it->second;
                    auto s = systems_.find(r.key);
                    if (s != systems_.end()) {  // another builder took this key meanwhile
                        r.cache_hit = true;
                        ++hits_;
                        s->second->waiting.push_back(id);
                        unbuilt_.pop_front();
                        continue;
                    }
                    break;
                }
                if (!unbuilt_.empty()) {
                    ServeRequest &r = *requests_.at(unbuilt_.front());
                    const u64 need = estimate(r);
                    if (need > options_.hbm_budget) {
                        unbuilt_.pop_front();
                        r.error = Error(SF_ERR_SERVE_OVER_BUDGET, "the request needs more HBM than the whole budget", (int64_t)need,
                                        (int64_t)options_.hbm_budget);
                        finish_request(r, SF_SERVE_FAILED);
                        cv_.notify_all();
                        continue;
                    }
                    if (evict_for(need, doomed)) {
                        unbuilt_.pop_front();
                        auto s = std::make_unique<ServeSystem>();
                        s->key = r.key;
                        s->bytes = need;
                        s->waiting.push_back(r.id);
                        r.state = SF_SERVE_BUILDING;
                        system = s.get();
                        systems_[r.key] = std::move(s);
                        in_use_ += need;
                        peak_ = std::max(peak_, in_use_);
                        w = r.workload;
                        m = r.model;
                        first = r.id;
                        break;
                    }
                }
                if (!doomed.empty()) break;  // free the evicted systems first, then look again
                cv_.wait(g);
            }
        }
        for (sf_static_study *s : doomed) sf_static_study_destroy(s);
        if (!system) continue;
        const double t0 = now();
        Error failure{SF_OK, ""};
        try {
            build(*system, w, m, *raw);
        } catch (const Error &e) {
            failure = e;
        } catch (const std::exception &e) {
            failure = Error(SF_ERR_INTERNAL, e.what());
        }
        const double spent = now() - t0;
        std::lock_guard<std::mutex> g(mu_);
        auto r = requests_.find(first);
        if (r != requests_.end()) r->second->build_seconds = spent;
        if (failure.status != SF_OK) {
            // every request of this key fails with the build's error
            for (u64 id : system->waiting) {
                auto it = requests_.find(id);
                if (it == requests_.end()) continue;
                it->second->error = failure;
                finish_request(*it->second, SF_SERVE_FAILED);
            }
            in_use_ -= system->bytes;
            systems_.erase(system->key);
        } else {
            ++builds_;
            system->ready = true;
            system->last_used = now();
        }
        cv_.notify_all();
    }
}

void Server::solver_loop() {
    std::shared_ptr<Context> ctx;
    try {
        ctx = std::make_shared<Context>(device_);
    } catch (...) {
        return;
    }
    for (;;) {
        ServeSystem *system = nullptr;
        {
            std::unique_lock<std::mutex> g(mu_);
            for (;;) {
                if (stop_) return;
                // the system whose oldest waiting request is oldest
                for (auto &kv : systems_) {
                    ServeSystem &s = *kv.second;
                    if (!s.ready || s.busy || s.waiting.empty()) continue;
                    if (!system || s.waiting.front() < system->waiting.front()) system = &s;
                }
                if (system) {
                    system->busy = true;
                    break;
                }
                cv_.wait(g);
            }
        }
        // a device or host failure inside a batch fails the requests it held; anything else
        // thrown (std::bad_alloc from a host vector, say) does the same as an internal error
        // rather than leaving the thread and ending the process
        auto fail_solving = [&](const Error &e) {
            std::lock_guard<std::mutex> g(mu_);
            for (auto &kv : requests_) {
                ServeRequest &r = *kv.second;
                if (r.state == SF_SERVE_SOLVING && r.key == system->key) {
                    r.error = e;
                    finish_request(r, SF_SERVE_FAILED);
                }
            }
        };
        try {
            serve(*system, ctx);
        } catch (const Error &e) {
            fail_solving(e);
        } catch (const std::bad_alloc &) {
            fail_solving(Error(SF_ERR_OUT_OF_MEMORY, "host memory ran out while solving a batch"));
        } catch (const std::exception &e) {
            fail_solving(Error(SF_ERR_INTERNAL, std::string("a batch failed: ") + e.what()));
        } catch (...) {
            fail_solving(Error(SF_ERR_INTERNAL, "a batch failed"));
        }
        std::lock_guard<std::mutex> g(mu_);
        system->busy = false;
        system->last_used = now();
        cv_.notify_all();
    }
}

void Server::serve(ServeSystem &system, const std::shared_ptr<Context> &ctx) {
    sf_static_study &st = *system.study;
    const QuadMesh &mesh = st.model.mesh;
    const FreeProblem &numbering = st.problem;
    const size_t n = st.system.system.size;
    // the right-hand side of a request, or an error that fails it alone: a workload's tractions,
    // or a model's forces and tractions with the system's prescribed components eliminated as
    // the study eliminated them (study.rs:419-474, :567-593)
    auto load_of = [&](const sf_workload &w, const std::shared_ptr<const Model> &m, std::vector<double> &rhs) -> Error {
        try {
            bool any = false;
            if (m) {
                const std::vector<double> load = model_load(*m, &system.sorted_boundary);
                rhs.clear();
                rhs.reserve(3 * numbering.free_count);
                for (size_t node = 0; node < mesh.nodes.size(); ++node)
                    if (numbering.free[node] != FIXED)
                        for (int a = 0; a < 3; ++a) rhs.push_back(load[3 * node + a]);
                for (double v : rhs) any = any || v != 0.0;
                for (const auto &pr : st.prescribed_rhs.prescribed) any = any || pr.second != 0.0;
                apply_prescribed_rhs(st.prescribed_rhs, rhs);
            } else {
                rhs = traction_rhs(mesh, workload_tractions(mesh, system.boundary, w), numbering.free, numbering.free_count);
                for (double v : rhs) any = any || v != 0.0;
            }
            if (!any) return Error(SF_ERR_MODEL_NO_EFFECTIVE_LOAD, "the model has no effective load");
            return Error(SF_OK, "");
        } catch (const Error &e) {
            return e;
        }
    };
    // takes up to `room` waiting requests, failing those whose loads are refused
    auto take = [&](size_t room, std::vector<u64> &ids, std::vector<std::vector<double>> &rhs) {
        struct Picked {
            u64 id;
            sf_workload w;
            std::shared_ptr<const Model> m;
        };
        std::vector<Picked> picked;
        {
            std::lock_guard<std::mutex> g(mu_);
            while (picked.size() < room && !system.waiting.empty()) {
                const u64 id = system.waiting.front();
                system.waiting.pop_front();
                auto it = requests_.find(id);
                if (it == requests_.end()) continue;
                it->second->state = SF_SERVE_SOLVING;
                it->second->solve_started = now();
                picked.push_back({id, it->second->workload, it->second->model});
            }
        }
        for (auto &p : picked) {
            std::vector<double> b;
            Error e = load_of(p.w, p.m, b);
            if (e.status != SF_OK) {
                std::lock_guard<std::mutex> g(mu_);
                ServeRequest &r = *requests_.at(p.id);
                r.error = e;
                finish_request(r, SF_SERVE_FAILED);
                cv_.notify_all();
                continue;
            }
            ids.push_back(p.id);
            rhs.push_back(std::move(b));
        }
    };

    std::vector<u64> ids;
    std::vector<std::vector<double>> rhs;
    size_t width = 1;
    {
        std::lock_guard<std::mutex> g(mu_);
        while (width < system.waiting.size() && width < options_.max_columns) width *= 2;
        width = std::min<size_t>(width, options_.max_columns);
    }
    take(width, ids, rhs);
    if (ids.empty()) return;
    // the free-DOF solution of a request as its nodal displacements
    auto nodal_of = [&](const std::vector<double> &x) {
        std::vector<double> nodal(3 * mesh.nodes.size(), 0.0);
        for (size_t node = 0; node < mesh.nodes.size(); ++node)
            if (numbering.free[node] != FIXED)
                for (int a = 0; a < 3; ++a) nodal[3 * node + a] = x[3 * (size_t)numbering.free[node] + a];
        // a prescribed component is exact (study.rs:233-235); the system's requests share them
        for (const auto &pr : st.model.prescribed) nodal[3 * (size_t)pr.first.first + pr.first.second] = pr.second;
        return nodal;
    };
    // Fewer than four requests: each runs the single-column solve in turn (the same bits as a
    // batch column). On the 300k-DOF block one column alone takes 9.7 ms against 34.6 ms as a
    // one-column batch, two 19.3 ms against 34.1 ms, and four 38.2 ms against 31.1 ms. Requests
    // that arrive meanwhile wait for the next round, where they batch together.
    auto solve_alone = [&](u64 id, const std::vector<double> &b) {
        const sf_pcg_options pcg{options_.tolerance, 0, options_.dot_order, 0};
        sf_column_progress done{};
        std::vector<double> x;
        try {
            DevicePcg single(ctx, st.system.system, st.mg.mg.get(), b, pcg);
            for (;;) {
                const int outcome = single.step(options_.slice_iterations, Cancel{nullptr, nullptr});
                if (outcome != SF_RUNNING) break;
                std::lock_guard<std::mutex> g(mu_);
                if (stop_) return false;
            }
            const sf_progress p = single.progress();
            done.iteration = p.iteration;
            done.relative_residual = p.relative_residual;
            done.at_rounding_floor = p.at_rounding_floor;
            done.outcome = p.outcome;
            x.resize(n);
            single.estimate(x.data());
        } catch (const Error &e) {
            if (e.status == SF_ERR_HIP) throw;  // the device, not the request: the system's failure path
            done.outcome = SF_FAILED;
            done.error = e.status;
            done.error_iteration = (uint64_t)e.a;
            done.error_value = e.value;
        }
        std::vector<double> nodal;
        if (done.outcome == SF_CONVERGED) nodal = nodal_of(x);
        std::lock_guard<std::mutex> g(mu_);
        auto it = requests_.find(id);
        if (it != requests_.end()) {
            ServeRequest &r = *it->second;
            r.progress = done;
            r.free_dofs = n;
            if (done.outcome == SF_CONVERGED) {
                r.displacements = std::move(nodal);
                finish_request(r, SF_SERVE_DONE);
            } else {
                r.error = Error(done.error ? done.error : SF_ERR_INTERNAL, "the solve failed", (int64_t)done.error_iteration, 0,
                                done.error_value);
                finish_request(r, SF_SERVE_FAILED);
            }
        }
        cv_.notify_all();
        return true;
    };
    if (ids.size() < 4) {
        for (size_t k = 0; k < ids.size(); ++k)
            if (!solve_alone(ids[k], rhs[k])) return;
        return;
    }
    const uint32_t M = (uint32_t)width;
    std::vector<double> columns(n * M, 0.0);
    std::vector<u64> slot(M, 0);
    for (size_t k = 0; k < ids.size(); ++k) {
        std::copy(rhs[k].begin(), rhs[k].end(), columns.begin() + k * n);
        slot[k] = ids[k];
    }
    const sf_pcg_options pcg{options_.tolerance, 0, options_.dot_order, 0};
    DevicePcgBatch batch(ctx, st.system.system, st.mg.mg.get(), columns.data(), M, pcg);
    std::vector<sf_column_progress> progress(M);
    std::vector<double> x(n);
    for (;;) {
        batch.step(options_.slice_iterations, Cancel{nullptr, nullptr});
        batch.progress(progress.data());
        bool any_open = false;
        for (uint32_t m = 0; m < M; ++m) {
            if (!slot[m]) continue;
            if (progress[m].outcome == SF_RUNNING) {
                any_open = true;
                continue;
            }
            std::vector<double> nodal;
            if (progress[m].outcome == SF_CONVERGED) {
                batch.estimate_column(m, x.data());
                nodal = nodal_of(x);
            }
            std::lock_guard<std::mutex> g(mu_);
            auto it = requests_.find(slot[m]);
            if (it != requests_.end()) {
                ServeRequest &r = *it->second;
                r.progress = progress[m];
                r.free_dofs = n;
                if (progress[m].outcome == SF_CONVERGED) {
                    r.displacements = std::move(nodal);
                    finish_request(r, SF_SERVE_DONE);
                } else {
                    r.error = Error(progress[m].error ? progress[m].error : SF_ERR_INTERNAL, "the solve failed",
                                    (int64_t)progress[m].error_iteration, 0, progress[m].error_value);
                    finish_request(r, SF_SERVE_FAILED);
                }
            }
            slot[m] = 0;
            cv_.notify_all();
        }
        // continuous batching: waiting requests take the free columns
        std::vector<uint32_t> free_cols;
        for (uint32_t m = 0; m < M; ++m)
            if (!slot[m]) free_cols.push_back(m);
        std::vector<u64> more;
        std::vector<std::vector<double>> more_rhs;
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stop_) return;
        }
        if (!free_cols.empty()) take(free_cols.size(), more, more_rhs);
        if (!more.empty()) {
            std::vector<uint32_t> cols;
            std::vector<const double *> ptrs;
            for (size_t k = 0; k < more.size(); ++k) {
                cols.push_back(free_cols[k]);
                ptrs.push_back(more_rhs[k].data());
                slot[free_cols[k]] = more[k];
            }
            batch.replace_columns(cols, ptrs);
            any_open = true;
        }
        if (!any_open) return;
    }
}

}  // namespace sf
