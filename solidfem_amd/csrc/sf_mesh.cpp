// Host input side of the solve path: meshes (crates/solidfem-mesh), the model's numbering, load
// and rigid-body modes (crates/solidfem-elastic/src/{study,loads,rigid_modes}.rs), the block
// pattern with its assembly map, and the deterministic host thread pool.
#include <algorithm>
#include <string>
#include <cerrno>
#include <cctype>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>

#include "sf_host.hpp"

namespace sf {

// ---- thread pool ----------------------------------------------------------------------------

size_t host_threads() {
    static size_t n = [] {
        // the process's CPU share where the environment states one, else the hardware's threads
        size_t h = std::thread::hardware_concurrency();
        if (const char *omp = std::getenv("OMP_NUM_THREADS")) h = std::min<size_t>(h, (size_t)std::max(1, std::atoi(omp)));
        if (const char *env = std::getenv("SOLIDFEM_HOST_THREADS")) h = (size_t)std::max(1, std::atoi(env));
        return std::max<size_t>(1, std::min<size_t>(h, 64));
    }();
    return n;
}

namespace {

// A persistent pool of host_threads() - 1 workers taking chunks of parallel_for calls from one
// queue. The caller of parallel_for runs chunks too (its own, or any other call's), so calls made
// from several threads at once, or from inside a chunk, always progress. Chunk boundaries depend
// only on n and the thread count, never on which thread runs a chunk.
class HostPool {
   public:
    struct Job {
        const std::function<void(size_t, size_t)> *body = nullptr;
        size_t remaining = 0;  // chunks not yet finished, guarded by the pool's lock
        std::exception_ptr failure;
        std::condition_variable done;
    };
    struct Task {
        Job *job;
        size_t first, last;
    };

    static HostPool &get() {
        // never destroyed: workers may still wait on the queue when the process exits
        static HostPool *pool = new HostPool(host_threads() > 1 ? host_threads() - 1 : 0);
        return *pool;
    }

    void run(Job &job, const std::vector<Task> &tasks) {
        std::unique_lock<std::mutex> lock(m_);
        job.remaining = tasks.size();
        for (size_t i = 1; i < tasks.size(); ++i) queue_.push_back(tasks[i]);
        wake_.notify_all();
        lock.unlock();
        execute(tasks[0]);
        lock.lock();
        while (job.remaining > 0) {
            if (!queue_.empty()) {
                const Task t = queue_.front();
                queue_.pop_front();
                lock.unlock();
                execute(t);
                lock.lock();
            } else {
                job.done.wait(lock);
            }
        }
    }

   private:
    explicit HostPool(size_t workers) {
        for (size_t i = 0; i < workers; ++i)
            std::thread([this] {
                std::unique_lock<std::mutex> lock(m_);
                for (;;) {
                    wake_.wait(lock, [this] { return !queue_.empty(); });
                    const Task t = queue_.front();
                    queue_.pop_front();
                    lock.unlock();
                    execute(t);
                    lock.lock();
                }
            }).detach();
    }

    void execute(const Task &t) {
        std::exception_ptr failure;
        try {
            (*t.job->body)(t.first, t.last);
        } catch (...) {
            failure = std::current_exception();
        }
        std::lock_guard<std::mutex> g(m_);
        if (failure && !t.job->failure) t.job->failure = failure;
        if (--t.job->remaining == 0) t.job->done.notify_all();
    }

    std::mutex m_;
    std::condition_variable wake_;
    std::deque<Task> queue_;
};

}  // namespace

void parallel_for(size_t n, size_t min_chunk, const std::function<void(size_t, size_t)> &body) {
    const size_t threads = host_threads();
    if (n == 0) return;
    const size_t chunks = std::min(threads, (n + min_chunk - 1) / std::max<size_t>(min_chunk, 1));
    if (chunks <= 1) {
        body(0, n);
        return;
    }
    const size_t per = (n + chunks - 1) / chunks;
    HostPool::Job job;
    job.body = &body;
    std::vector<HostPool::Task> tasks;
    for (size_t c = 0; c < chunks; ++c) {
        const size_t first = c * per, last = std::min(n, first + per);
        if (first >= last) break;
        tasks.push_back({&job, first, last});
    }
    HostPool::get().run(job, tasks);
    if (job.failure) std::rethrow_exception(job.failure);
}

// ---- meshes -----------------------------------------------------------------------------------

const int FACE_LOCAL[4][6] = {{1, 2, 3, 5, 9, 8}, {0, 3, 2, 7, 9, 6}, {0, 1, 3, 4, 8, 7}, {0, 2, 1, 6, 5, 4}};
const int EDGE_PAIRS[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

static double signed_volume(const Vec3 &a, const Vec3 &b, const Vec3 &c, const Vec3 &d) {
    // crates/solidfem-mesh/src/tet.rs:9-16
    double u0 = b[0] - a[0], u1 = b[1] - a[1], u2 = b[2] - a[2];
    double v0 = c[0] - a[0], v1 = c[1] - a[1], v2 = c[2] - a[2];
    double w0 = d[0] - a[0], w1 = d[1] - a[1], w2 = d[2] - a[2];
    double det = u0 * (v1 * w2 - v2 * w1) - u1 * (v0 * w2 - v2 * w0) + u2 * (v0 * w1 - v1 * w0);
    return det / 6.0;
}

static std::array<u32, 3> sorted_corners(const std::array<u32, 4> &t, int face) {
    std::array<u32, 3> k = {t[FACE_LOCAL[face][0]], t[FACE_LOCAL[face][1]], t[FACE_LOCAL[face][2]]};
    std::sort(k.begin(), k.end());
    return k;
}

// Sorts `items` by `less` where `lead(item)`, in [0, range), orders items first: a counting sort
// by lead value, then each bucket sorted on the host pool. Wherever items that compare equal are
// identical, the result is std::sort's.
template <typename T, typename Lead, typename Less>
static void lead_sort(std::vector<T> &items, size_t range, Lead lead, Less less) {
    std::vector<size_t> start(range + 1, 0);
    for (const T &x : items) start[(size_t)lead(x) + 1] += 1;
    for (size_t i = 0; i < range; ++i) start[i + 1] += start[i];
    std::vector<T> out(items.size());
    {
        std::vector<size_t> fill(start.begin(), start.end() - 1);
        for (const T &x : items) out[fill[(size_t)lead(x)]++] = x;
    }
    parallel_for(range, 4096, [&](size_t first, size_t last) {
        for (size_t r = first; r < last; ++r) std::sort(out.begin() + (int64_t)start[r], out.begin() + (int64_t)start[r + 1], less);
    });
    items.swap(out);
}

void validate_tet_mesh(const TetMesh &m) {
    // crates/solidfem-mesh/src/tet.rs:59-109
    if (m.tets.empty()) throw Error(SF_ERR_MESH_EMPTY, "the mesh has no tetrahedra");
    for (size_t n = 0; n < m.positions.size(); ++n)
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(m.positions[n][a]))
                throw Error(SF_ERR_MESH_NON_FINITE_POSITION, "node has a NaN or infinite coordinate", (int64_t)n);
    std::vector<char> used(m.positions.size(), 0);
    for (size_t t = 0; t < m.tets.size(); ++t) {
        const auto &tet = m.tets[t];
        for (u32 node : tet) {
            if (node >= m.positions.size())
                throw Error(SF_ERR_MESH_NODE_OUT_OF_RANGE, "tetrahedron names a node that does not exist", (int64_t)t, node);
            used[node] = 1;
        }
        const Vec3 &a = m.positions[tet[0]], &b = m.positions[tet[1]], &c = m.positions[tet[2]], &d = m.positions[tet[3]];
        double volume = signed_volume(a, b, c, d);
        if (std::isnan(volume) || volume <= 0.0)
            throw Error(SF_ERR_MESH_DEGENERATE_TET, "tetrahedron has no positive volume", (int64_t)t, 0, volume);
        auto l2 = [](const Vec3 &p, const Vec3 &q) {
            return (p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]);
        };
        double longest = 0.0;
        for (double e : {l2(a, b), l2(a, c), l2(a, d), l2(b, c), l2(b, d), l2(c, d)}) longest = std::max(longest, e);
        double relative = 6.0 * volume / (longest * std::sqrt(longest));
        if (std::isnan(relative) || relative < 1e-12)
            throw Error(SF_ERR_MESH_FLAT_TET, "tetrahedron is flat to rounding", (int64_t)t, 0, relative);
    }
    for (size_t n = 0; n < used.size(); ++n)
        if (!used[n]) throw Error(SF_ERR_MESH_UNUSED_NODE, "node belongs to no tetrahedron", (int64_t)n);
    std::vector<std::array<u32, 3>> faces(4 * m.tets.size());
    parallel_for(m.tets.size(), 8192, [&](size_t first, size_t last) {
        for (size_t t = first; t < last; ++t)
            for (int f = 0; f < 4; ++f) faces[4 * t + f] = sorted_corners(m.tets[t], f);
    });
    lead_sort(faces, m.positions.size(), [](const std::array<u32, 3> &k) { return k[0]; },
              [](const std::array<u32, 3> &a, const std::array<u32, 3> &b) { return a < b; });
    for (size_t i = 2; i < faces.size(); ++i)
        if (faces[i] == faces[i - 2]) throw Error(SF_ERR_MESH_NON_MANIFOLD_FACE, "a face is shared by more than two tetrahedra");
}

TetMesh make_box_mesh(const double origin[3], const double size[3], const int32_t divisions[3]) {
    // crates/solidfem-mesh/src/box_mesh.rs:22-90: six tetrahedra per cell around the cell's
    // low-to-high diagonal, one per axis order of the walk; odd orders swap two corners.
    static const int WALKS[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    for (int a = 0; a < 3; ++a) {
        if (divisions[a] <= 0) throw Error(SF_ERR_MESH_INVALID_BOX, "each axis needs at least one division");
        if (std::isnan(size[a]) || size[a] <= 0.0) throw Error(SF_ERR_MESH_INVALID_BOX, "each box length must be positive");
    }
    const size_t nx = (size_t)divisions[0] + 1, ny = (size_t)divisions[1] + 1, nz = (size_t)divisions[2] + 1;
    auto line = [](double o, double s, size_t i, size_t d) { return i == d ? o + s : o + s * ((double)i / (double)d); };
    TetMesh m;
    m.positions.resize(nx * ny * nz);
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i)
                m.positions[i + nx * (j + ny * k)] = Vec3{{line(origin[0], size[0], i, nx - 1), line(origin[1], size[1], j, ny - 1),
                                                           line(origin[2], size[2], k, nz - 1)}};
    m.tets.reserve(6 * (nx - 1) * (ny - 1) * (nz - 1));
    for (size_t k = 0; k + 1 < nz; ++k)
        for (size_t j = 0; j + 1 < ny; ++j)
            for (size_t i = 0; i + 1 < nx; ++i)
                for (int w = 0; w < 6; ++w) {
                    size_t at[3] = {i, j, k};
                    u32 v[4];
                    v[0] = (u32)(at[0] + nx * (at[1] + ny * at[2]));
                    for (int s = 0; s < 3; ++s) {
                        at[WALKS[w][s]] += 1;
                        v[s + 1] = (u32)(at[0] + nx * (at[1] + ny * at[2]));
                    }
                    if (w < 3) m.tets.push_back({v[0], v[1], v[2], v[3]});
                    else m.tets.push_back({v[0], v[2], v[1], v[3]});
                }
    return m;
}

static void find_faces(QuadMesh &mesh);

QuadMesh upgrade_to_quadratic(const TetMesh &mesh) {
    // crates/solidfem-mesh/src/quadratic.rs:24-71: one midpoint per distinct edge, numbered
    // after the corners in ascending (min, max) corner order. Edges are packed into u64 keys.
    QuadMesh q;
    q.vertex_count = mesh.positions.size();
    std::vector<u64> edges(mesh.tets.size() * 6);
    parallel_for(mesh.tets.size(), 4096, [&](size_t first, size_t last) {
        for (size_t t = first; t < last; ++t)
            for (int e = 0; e < 6; ++e) {
                u32 x = mesh.tets[t][EDGE_PAIRS[e][0]], y = mesh.tets[t][EDGE_PAIRS[e][1]];
                edges[6 * t + e] = ((u64)std::min(x, y) << 32) | std::max(x, y);
            }
    });
    std::vector<u64> unique = edges;
    lead_sort(unique, mesh.positions.size(), [](u64 e) { return e >> 32; }, [](u64 a, u64 b) { return a < b; });
    unique.erase(std::unique(unique.begin(), unique.end()), unique.end());
    q.nodes.resize(q.vertex_count + unique.size());
    std::copy(mesh.positions.begin(), mesh.positions.end(), q.nodes.begin());
    parallel_for(unique.size(), 8192, [&](size_t first, size_t last) {
        for (size_t e = first; e < last; ++e) {
            const Vec3 &p = mesh.positions[unique[e] >> 32], &r = mesh.positions[unique[e] & 0xffffffffu];
            q.nodes[q.vertex_count + e] = Vec3{{0.5 * (p[0] + r[0]), 0.5 * (p[1] + r[1]), 0.5 * (p[2] + r[2])}};
        }
    });
    q.tets.resize(mesh.tets.size());
    parallel_for(mesh.tets.size(), 4096, [&](size_t first, size_t last) {
        for (size_t t = first; t < last; ++t) {
            for (int c = 0; c < 4; ++c) q.tets[t][c] = mesh.tets[t][c];
            for (int e = 0; e < 6; ++e) {
                size_t at = std::lower_bound(unique.begin(), unique.end(), edges[6 * t + e]) - unique.begin();
                q.tets[t][4 + e] = (u32)(q.vertex_count + at);
            }
        }
    });
    q.corners = mesh.tets;
    find_faces(q);
    return q;
}

static void find_faces(QuadMesh &mesh) {
    // crates/solidfem-mesh/src/quadratic.rs:89-124: faces met once, ascending by sorted corners,
    // then by (tet, face); and the pairs of tetrahedra that meet at a face (quadratic.rs:129-166).
    struct Key {
        std::array<u32, 3> k;
        u32 tet, face;
        bool operator<(const Key &o) const {
            if (k != o.k) return k < o.k;
            if (tet != o.tet) return tet < o.tet;
            return face < o.face;
        }
    };
    std::vector<Key> keys(mesh.corners.size() * 4);
    parallel_for(mesh.corners.size(), 8192, [&](size_t first, size_t last) {
        for (size_t t = first; t < last; ++t)
            for (int f = 0; f < 4; ++f) keys[4 * t + f] = Key{sorted_corners(mesh.corners[t], f), (u32)t, (u32)f};
    });
    lead_sort(keys, mesh.nodes.size(), [](const Key &k) { return k.k[0]; }, [](const Key &a, const Key &b) { return a < b; });
    std::vector<Face> &out = mesh.boundary;
    out.clear();
    mesh.shared_faces.clear();
    for (size_t at = 0; at < keys.size();) {
        size_t end = at + 1;
        while (end < keys.size() && keys[end].k == keys[at].k) ++end;
        if (end == at + 1) {
            Face f;
            f.tet = keys[at].tet;
            f.face = keys[at].face;
            for (int s = 0; s < 6; ++s) f.nodes[s] = mesh.tets[f.tet][FACE_LOCAL[f.face][s]];
            out.push_back(f);
        }
        for (size_t k = at + 1; k < end; ++k) mesh.shared_faces.push_back({keys[k - 1].tet, keys[k].tet});
        at = end;
    }
}

// ---- workload inputs ------------------------------------------------------------------------

static u64 mix64(u64 x) {
    // SplitMix64 finaliser: the counter-based generator of every seeded workload input.
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static double uniform01(u64 bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }
static double gaussian(u64 seed, u64 counter) {
    double u1 = ((double)(mix64(seed + 2 * counter) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    double u2 = uniform01(mix64(seed + 2 * counter + 1));
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

std::vector<Traction> workload_tractions(const QuadMesh &mesh, const std::vector<Face> &boundary, const sf_workload &w) {
    const auto &nodes = mesh.nodes;
    std::vector<Traction> out;
    if (w.load_axis < 0) return out;  // no loaded plane
    if (w.load_axis > 2) throw Error(SF_ERR_MODEL_AXIS_OUT_OF_RANGE, "load axis out of range", w.load_axis);
    int u = (w.load_axis + 1) % 3, v = (w.load_axis + 2) % 3;
    u64 index = 0;
    for (const Face &f : boundary) {
        bool on = true;
        for (int c = 0; c < 3; ++c)
            if (!(std::fabs(nodes[f.nodes[c]][w.load_axis] - w.load_value) < 1e-9)) on = false;
        if (!on) continue;
        std::array<double, 3> t{0, 0, 0};
        if (w.traction_kind == SF_TRACTION_UNIFORM) {
            t = {w.traction[0], w.traction[1], w.traction[2]};
        } else if (w.traction_kind == SF_TRACTION_RANDOM) {
            for (int a = 0; a < 3; ++a)
                t[a] = w.traction_scale * gaussian(w.traction_seed, ((u64)(uint32_t)w.load_case << 40) ^ (index << 2) ^ (u64)a);
        } else {
            const Vec3 &p = nodes[f.nodes[0]], &q = nodes[f.nodes[1]], &r = nodes[f.nodes[2]];
            double cu = (p[u] + q[u] + r[u]) / 3.0 - w.axis_u;
            double cv = (p[v] + q[v] + r[v]) / 3.0 - w.axis_v;
            double radius = std::sqrt(cu * cu + cv * cv);
            t[w.load_axis] = 0.0;
            t[u] = -w.traction_scale * cv / radius;
            t[v] = w.traction_scale * cu / radius;
        }
        out.push_back({f, t});
        ++index;
    }
    return out;
}

namespace {

int longest_axis(const double size[3]) {
    int longest = 0;
    for (int a = 1; a < 3; ++a)
        if (size[a] > size[longest]) longest = a;
    return longest;
}

}  // namespace

TetMesh make_tube_mesh(const sf_workload &w) {
    // demo/wasm/src/part.rs:48-185: the box's grid along the longest axis; across it, `wall_cells`
    // cells through each wall of `wall_thickness` and the rest evenly over the hollow's span, whose
    // cells are removed. Every cell is cut into the same six tetrahedra, so what remains conforms.
    // Vertices are renumbered in order of first use by the kept tetrahedra, corner by corner.
    const int along = longest_axis(w.size), across[2] = {(along + 1) % 3, (along + 2) % 3};
    const int walls = w.wall_cells;
    for (int a = 0; a < 3; ++a)
        if (w.divisions[a] <= 0 || std::isnan(w.size[a]) || w.size[a] <= 0.0)
            throw Error(SF_ERR_MESH_INVALID_BOX, "each axis needs at least one division and a positive length");
    for (int a : across)
        if (w.divisions[a] <= 2 * walls || !(w.wall_thickness > 0.0) || !(2.0 * w.wall_thickness < w.size[a]))
            throw Error(SF_ERR_MESH_INVALID_BOX, "a tube needs a wall thinner than half its side and a hollow of at least one cell");
    std::vector<double> lines[3];
    for (int a = 0; a < 3; ++a) {
        const size_t n = (size_t)w.divisions[a];
        auto &l = lines[a];
        if (a == along) {
            for (size_t i = 0; i <= n; ++i) l.push_back(i == n ? w.origin[a] + w.size[a] : w.origin[a] + w.size[a] * ((double)i / (double)n));
            continue;
        }
        // part.rs:58-82, the span measured from the origin
        const double wall = w.wall_thickness, inner = w.size[a] - 2.0 * wall, step = wall / (double)walls;
        const size_t inner_cells = n - 2 * (size_t)walls;
        for (int i = 0; i <= walls; ++i) l.push_back(w.origin[a] + step * (double)i);
        for (size_t i = 1; i < inner_cells; ++i) l.push_back(w.origin[a] + (wall + inner * (double)i / (double)inner_cells));
        for (int i = 0; i <= walls; ++i) l.push_back(w.origin[a] + (w.size[a] - wall + step * (double)i));
    }
    const size_t n0 = (size_t)w.divisions[0], n1 = (size_t)w.divisions[1], nx = n0 + 1, ny = n1 + 1;
    const double zero[3] = {0.0, 0.0, 0.0}, cells[3] = {(double)w.divisions[0], (double)w.divisions[1], (double)w.divisions[2]};
    const TetMesh unit = make_box_mesh(zero, cells, w.divisions);
    auto hollow = [&](const size_t cell[3]) {
        const size_t j = cell[across[0]], k = cell[across[1]];
        const size_t nj = (size_t)w.divisions[across[0]], nk = (size_t)w.divisions[across[1]];
        return j >= (size_t)walls && j < nj - (size_t)walls && k >= (size_t)walls && k < nk - (size_t)walls;
    };
    TetMesh m;
    std::vector<u32> remap(unit.positions.size(), FIXED);
    for (size_t t = 0; t < unit.tets.size(); ++t) {
        // the box builder emits six tetrahedra per cell, cells x-fastest
        const size_t c = t / 6, cell[3] = {c % n0, (c / n0) % n1, c / (n0 * n1)};
        if (hollow(cell)) continue;
        std::array<u32, 4> tet;
        for (int i = 0; i < 4; ++i) {
            const u32 v = unit.tets[t][(size_t)i];
            if (remap[v] == FIXED) {
                remap[v] = (u32)m.positions.size();
                const size_t at[3] = {v % nx, (v / nx) % ny, v / (nx * ny)};
                m.positions.push_back(Vec3{{lines[0][at[0]], lines[1][at[1]], lines[2][at[2]]}});
            }
            tet[(size_t)i] = remap[v];
        }
        m.tets.push_back(tet);
    }
    return m;
}

TetMesh read_tet_mesh_text(const char *text, size_t length) {
    // tokens separated by white space; the line of each token kept for the refusal
    size_t at = 0, line = 1;
    auto next = [&](std::string &token) -> bool {
        while (at < length && std::isspace((unsigned char)text[at])) {
            if (text[at] == '\n') ++line;
            ++at;
        }
        if (at >= length) return false;
        const size_t from = at;
        while (at < length && !std::isspace((unsigned char)text[at])) ++at;
        token.assign(text + from, at - from);
        return true;
    };
    auto refuse = [&](const char *what) { throw Error(SF_ERR_INVALID_ARGUMENT, std::string("mesh text: ") + what, (int64_t)line); };
    auto count = [&](const char *what) -> uint64_t {
        std::string token;
        if (!next(token)) refuse(what);
        char *end = nullptr;
        errno = 0;
        const unsigned long long v = std::strtoull(token.c_str(), &end, 10);
        if (errno || end == token.c_str() || *end || token[0] == '-' || v >= (1ull << 32)) refuse(what);
        return v;
    };
    auto real = [&](const char *what) -> double {
        std::string token;
        if (!next(token)) refuse(what);
        char *end = nullptr;
        const double v = std::strtod(token.c_str(), &end);
        if (end == token.c_str() || *end) refuse(what);
        return v;
    };
    TetMesh mesh;
    const uint64_t vertices = count("the vertex count is missing or not a count");
    const uint64_t tets = count("the tetrahedron count is missing or not a count");
    mesh.positions.resize(vertices);
    for (auto &p : mesh.positions)
        for (int a = 0; a < 3; ++a) p[a] = real("a vertex coordinate is missing or not a number");
    mesh.tets.resize(tets);
    for (auto &t : mesh.tets)
        for (int a = 0; a < 4; ++a) t[a] = (u32)count("a tetrahedron's vertex index is missing or not an index");
    std::string extra;
    if (next(extra)) refuse("text follows the last tetrahedron");
    validate_tet_mesh(mesh);
    return mesh;
}

Model model_from_workload(const sf_workload &w) {
    if (w.wall_cells < 0) throw Error(SF_ERR_INVALID_ARGUMENT, "wall_cells must not be negative");
    if (w.clamp_axis < -1 || w.clamp_axis > 2) throw Error(SF_ERR_MODEL_AXIS_OUT_OF_RANGE, "clamp axis out of range", w.clamp_axis);
    if (w.load_axis < -1 || w.load_axis > 2) throw Error(SF_ERR_MODEL_AXIS_OUT_OF_RANGE, "load axis out of range", w.load_axis);
    if (w.wall_cells > 0 && w.jitter_fraction > 0.0) throw Error(SF_ERR_INVALID_ARGUMENT, "a tube is not jittered");
    StageClock clock;
    TetMesh linear = w.wall_cells > 0 ? make_tube_mesh(w) : make_box_mesh(w.origin, w.size, w.divisions);
    clock.lap("linear mesh");
    if (w.jitter_fraction > 0.0) {
        // Interior vertices move by (2u - 1) * (fraction * h) per axis, u from (seed ^ (3 v + axis)).
        const size_t nx = (size_t)w.divisions[0] + 1, ny = (size_t)w.divisions[1] + 1;
        double h[3];
        for (int a = 0; a < 3; ++a) h[a] = w.size[a] / (double)w.divisions[a];
        for (size_t v = 0; v < linear.positions.size(); ++v) {
            size_t i = v % nx, j = (v / nx) % ny, k = v / (nx * ny);
            if (i == 0 || j == 0 || k == 0 || i == (size_t)w.divisions[0] || j == (size_t)w.divisions[1] ||
                k == (size_t)w.divisions[2])
                continue;
            for (int a = 0; a < 3; ++a) {
                double u = uniform01(mix64(w.jitter_seed ^ (u64)(3 * v + a)));
                linear.positions[v][a] = linear.positions[v][a] + (2.0 * u - 1.0) * (w.jitter_fraction * h[a]);
            }
        }
    }
    clock.lap("jitter");
    validate_tet_mesh(linear);
    clock.lap("mesh validation");
    Model model;
    model.mesh = upgrade_to_quadratic(linear);
    clock.lap("quadratic upgrade");
    const auto &nodes = model.mesh.nodes;
    if (w.clamp_axis >= 0)
        for (size_t n = 0; n < nodes.size(); ++n)
            if (std::fabs(nodes[n][w.clamp_axis] - w.clamp_value) < 1e-9) model.fixed.insert((u32)n);
    const auto boundary = model.mesh.boundary_faces();
    clock.lap("boundary faces");
    model.tractions = workload_tractions(model.mesh, boundary, w);
    clock.lap("tractions");
    if (w.material_kind != SF_MATERIAL_UNIFORM && w.material_kind != SF_MATERIAL_SHELL_CORE)
        throw Error(SF_ERR_INVALID_ARGUMENT, "a workload's material is uniform or a shell and core");
    model.field.kind = w.material_kind;
    model.field.e = w.youngs;
    model.field.nu = w.poisson;
    model.field.core_e = w.core_youngs;
    model.field.core_nu = w.core_poisson;
    model.field.shell_depth = w.shell_depth;
    const int longest = longest_axis(w.size);
    model.field.transverse[0] = (longest + 1) % 3;
    model.field.transverse[1] = (longest + 2) % 3;
    for (int a = 0; a < 3; ++a) {
        model.field.lo[a] = w.origin[a];
        model.field.hi[a] = w.origin[a] + w.size[a];
    }
    model.quadrature = w.quadrature_points == 14 ? 14 : 4;
    return model;
}

// ---- numbering and load ------------------------------------------------------------------

// Consistent nodal load of a uniform traction over a 6-node face with the 6-point triangle rule
// (crates/solidfem-elastic/src/loads.rs:10-71, crates/solidfem-elastic/src/quadrature.rs:87-94).
static void face_load(const QuadMesh &mesh, const Face &face, const std::array<double, 3> &t, std::vector<double> &load) {
    static const double TA = 0.091576213509771, TB = 0.445948490915965;
    static const double WA = 0.109951743655322, WB = 0.223381589678011;
    static const double RULE[6][3] = {{TA, TA, WA}, {1.0 - 2.0 * TA, TA, WA}, {TA, 1.0 - 2.0 * TA, WA},
                                      {TB, TB, WB}, {1.0 - 2.0 * TB, TB, WB}, {TB, 1.0 - 2.0 * TB, WB}};
    const Vec3 *p[6];
    for (int n = 0; n < 6; ++n) p[n] = &mesh.nodes[face.nodes[n]];
    for (const auto &qp : RULE) {
        const double r = qp[0], s = qp[1], l = 1.0 - r - s;
        const double n[6] = {l * (2.0 * l - 1.0), r * (2.0 * r - 1.0), s * (2.0 * s - 1.0), 4.0 * l * r, 4.0 * r * s, 4.0 * s * l};
        const double dr[6] = {-(4.0 * l - 1.0), 4.0 * r - 1.0, 0.0, 4.0 * (l - r), 4.0 * s, -4.0 * s};
        const double ds[6] = {-(4.0 * l - 1.0), 0.0, 4.0 * s - 1.0, -4.0 * r, 4.0 * r, 4.0 * (l - s)};
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        for (int k = 0; k < 6; ++k)
            for (int x = 0; x < 3; ++x) {
                a[x] += dr[k] * (*p[k])[x];
                b[x] += ds[k] * (*p[k])[x];
            }
        const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
        const double measure = 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2) * qp[2];
        for (int k = 0; k < 6; ++k)
            for (int x = 0; x < 3; ++x) load[3 * (size_t)face.nodes[k] + x] += t[x] * n[k] * measure;
    }
}

// ---- held bodies (crates/solidfem-elastic/src/constraints.rs) -----------------------------

std::vector<std::vector<u32>> face_connected_bodies(const QuadMesh &mesh) {
    // quadratic.rs:129-166: elements sharing a corner triangle join; bodies come out ordered by
    // their lowest element, each body's elements ascending. Any union order gives this partition,
    // so the pairs the face matching of the mesh found are joined in the order found.
    const size_t n = mesh.corners.size();
    std::vector<u32> parent(n);
    std::iota(parent.begin(), parent.end(), 0u);
    auto find = [&](u32 x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    for (const auto &pair : mesh.shared_faces) {
        const u32 r0 = find(pair[0]), r1 = find(pair[1]);
        if (r0 != r1) parent[std::max(r0, r1)] = std::min(r0, r1);
    }
    std::vector<std::vector<u32>> groups;
    std::vector<u32> group_of(n, FIXED);
    for (size_t t = 0; t < n; ++t) {
        const u32 root = find((u32)t);
        if (group_of[root] == FIXED) {
            group_of[root] = (u32)groups.size();
            groups.emplace_back();
        }
        groups[group_of[root]].push_back((u32)t);
    }
    return groups;
}

namespace {

constexpr double COLLINEAR_FRACTION = 1e-9;  // constraints.rs:37-39

double distance(const Vec3 &a, const Vec3 &b) {
    const double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
}

// constraints.rs:256-268: the diagonal of the nodes' bounding box.
double bounding_size(const std::vector<Vec3> &positions, const std::vector<u32> &nodes) {
    Vec3 lo{{INFINITY, INFINITY, INFINITY}}, hi{{-INFINITY, -INFINITY, -INFINITY}};
    for (u32 n : nodes)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], positions[n][a]);
            hi[a] = std::fmax(hi[a], positions[n][a]);
        }
    return distance(lo, hi);
}

// constraints.rs:41-72: -1 when the points hold the body, else the SF_FREEDOM_* left.
int remaining_freedom(const std::vector<Vec3> &points, double size) {
    if (points.empty()) return SF_FREEDOM_FREE;
    const double limit = COLLINEAR_FRACTION * size;
    const Vec3 first = points[0];
    Vec3 far = first;
    double far_distance = 0.0;
    for (const Vec3 &p : points) {
        const double d = distance(first, p);
        if (d > far_distance) {
            far = p;
            far_distance = d;
        }
    }
    if (far_distance <= limit) return SF_FREEDOM_ROTATION_ABOUT_POINT;
    const double axis[3] = {(far[0] - first[0]) / far_distance, (far[1] - first[1]) / far_distance, (far[2] - first[2]) / far_distance};
    for (const Vec3 &p : points) {
        const double v[3] = {p[0] - first[0], p[1] - first[1], p[2] - first[2]};
        const double along = v[0] * axis[0] + v[1] * axis[1] + v[2] * axis[2];
        const double off[3] = {v[0] - along * axis[0], v[1] - along * axis[1], v[2] - along * axis[2]};
        if (std::sqrt(off[0] * off[0] + off[1] * off[1] + off[2] * off[2]) > limit) return -1;
    }
    return SF_FREEDOM_ROTATION_ABOUT_LINE;
}

std::vector<Vec3> held_points(const std::vector<Vec3> &positions, const std::vector<u32> &nodes, const std::vector<char> &held) {
    std::vector<Vec3> out;
    for (u32 n : nodes)
        if (held[n]) out.push_back(positions[n]);
    return out;
}

// constraints.rs:81-135: every body needs three non-collinear held nodes; a body held so holds
// every node it shares with the others, which may anchor them in turn.
void check_constraints(const QuadMesh &mesh, const std::vector<char> &fixed) {
    if (std::find(fixed.begin(), fixed.end(), 1) == fixed.end()) throw Error(SF_ERR_MODEL_NO_FIXED_NODES, "no node is fixed");
    const auto bodies = face_connected_bodies(mesh);
    std::vector<std::vector<u32>> body_nodes(bodies.size());
    std::vector<double> sizes(bodies.size());
    std::vector<u32> stamp(mesh.nodes.size(), FIXED);
    for (size_t b = 0; b < bodies.size(); ++b) {
        auto &nodes = body_nodes[b];
        for (u32 t : bodies[b])
            for (u32 node : mesh.tets[t])
                if (stamp[node] != (u32)b) {
                    stamp[node] = (u32)b;
                    nodes.push_back(node);
                }
        std::sort(nodes.begin(), nodes.end());
        sizes[b] = bounding_size(mesh.nodes, nodes);
    }
    std::vector<char> held = fixed, anchored(bodies.size(), 0);
    for (bool changed = true; changed;) {
        changed = false;
        for (size_t b = 0; b < bodies.size(); ++b) {
            if (anchored[b] || remaining_freedom(held_points(mesh.nodes, body_nodes[b], held), sizes[b]) >= 0) continue;
            anchored[b] = 1;
            changed = true;
            for (u32 n : body_nodes[b]) held[n] = 1;
        }
    }
    for (size_t b = 0; b < bodies.size(); ++b) {
        if (anchored[b]) continue;
        const auto points = held_points(mesh.nodes, body_nodes[b], held);
        const int freedom = remaining_freedom(points, sizes[b]);
        throw Error(SF_ERR_MODEL_BODY_NOT_HELD, "a body is not held", bodies[b][0], (int64_t)points.size(),
                    (double)(freedom < 0 ? SF_FREEDOM_FREE : freedom));
    }
}

// constraints.rs:238-275: the rank of rows of six by elimination with row pivoting per column,
// the pivot the first row of largest magnitude, counted when above COLLINEAR_FRACTION.
size_t pivot_rank(std::vector<std::array<double, 6>> &rows) {
    size_t rank = 0;
    for (int column = 0; column < 6; ++column) {
        if (rank >= rows.size()) break;
        size_t best = rank;
        for (size_t r = rank + 1; r < rows.size(); ++r)
            if (std::fabs(rows[r][column]) > std::fabs(rows[best][column])) best = r;
        if (std::fabs(rows[best][column]) <= COLLINEAR_FRACTION) continue;
        std::swap(rows[rank], rows[best]);
        const auto pivot = rows[rank];
        for (size_t r = rank + 1; r < rows.size(); ++r) {
            const double factor = rows[r][column] / pivot[column];
            if (factor != 0.0)
                for (int k = column; k < 6; ++k) rows[r][k] -= factor * pivot[k];
        }
        rank += 1;
    }
    return rank;
}

// constraints.rs:162-225: with prescribed components, the held components of each body must
// span all six rigid motions (about the body's centroid, scaled by its size).
void check_held_dofs(const QuadMesh &mesh, const std::vector<char> &fixed, const std::map<std::pair<u32, u32>, double> &prescribed) {
    const auto &positions = mesh.nodes;
    std::vector<char> in_body(positions.size(), 0);
    for (const auto &body : face_connected_bodies(mesh)) {
        std::fill(in_body.begin(), in_body.end(), 0);
        std::vector<u32> nodes;
        for (u32 t : body)
            for (u32 n : mesh.tets[t])
                if (!in_body[n]) {
                    in_body[n] = 1;
                    nodes.push_back(n);
                }
        std::sort(nodes.begin(), nodes.end());
        const double size = std::fmax(bounding_size(positions, nodes), 2.2250738585072014e-308);
        double center[3] = {0.0, 0.0, 0.0};
        for (u32 n : nodes)
            for (int a = 0; a < 3; ++a) center[a] += positions[n][a];
        for (double &c : center) c /= (double)nodes.size();
        std::vector<std::array<double, 6>> rows;
        auto add = [&](u32 node, u32 axis) {
            const Vec3 &p = positions[node];
            const double x = (p[0] - center[0]) / size, y = (p[1] - center[1]) / size, z = (p[2] - center[2]) / size;
            if (axis == 0)
                rows.push_back({1.0, 0.0, 0.0, 0.0, z, -y});
            else if (axis == 1)
                rows.push_back({0.0, 1.0, 0.0, -z, 0.0, x});
            else
                rows.push_back({0.0, 0.0, 1.0, y, -x, 0.0});
        };
        for (u32 n : nodes)
            if (fixed[n])
                for (u32 a = 0; a < 3; ++a) add(n, a);
        for (const auto &pr : prescribed) {
            const u32 node = pr.first.first, axis = pr.first.second;
            if (node < positions.size() && in_body[node] && axis < 3) add(node, axis);
        }
        const size_t rank = pivot_rank(rows);
        if (rank < 6) throw Error(SF_ERR_MODEL_DOFS_DO_NOT_HOLD, "the held components leave a rigid motion free", body[0], (int64_t)rank);
    }
}

}  // namespace

std::vector<double> model_load(const Model &model, const std::vector<Face> *sorted_boundary) {
    // crates/solidfem-elastic/src/study.rs:696-735: nodal forces, then tractions, in the order given
    const size_t nn = model.mesh.nodes.size();
    std::vector<double> load(3 * nn, 0.0);
    for (const auto &f : model.nodal_forces) {
        if (f.first >= nn) throw Error(SF_ERR_MODEL_NODE_OUT_OF_RANGE, "force node out of range", f.first);
        for (double c : f.second)
            if (!std::isfinite(c)) throw Error(SF_ERR_MODEL_NON_FINITE_LOAD, "a load is not finite");
        for (int a = 0; a < 3; ++a) load[3 * (size_t)f.first + a] += f.second[a];
    }
    if (!model.tractions.empty()) {
        std::vector<Face> own;
        if (!sorted_boundary) {
            own = model.mesh.boundary_faces();
            std::sort(own.begin(), own.end());
            sorted_boundary = &own;
        }
        for (const auto &t : model.tractions) {
            if (!std::binary_search(sorted_boundary->begin(), sorted_boundary->end(), t.first))
                throw Error(SF_ERR_MODEL_FACE_NOT_ON_BOUNDARY, "the face is not on the boundary", t.first.tet, t.first.face);
            for (double c : t.second)
                if (!std::isfinite(c)) throw Error(SF_ERR_MODEL_NON_FINITE_LOAD, "a traction is not finite");
            face_load(model.mesh, t.first, t.second, load);
        }
    }
    return load;
}

PrescribedRhs prescribed_rhs_program(const HostBsr<double> &a, const std::vector<std::pair<size_t, double>> &prescribed) {
    // what study.rs:567-593 does to the right-hand side, read from the matrix before elimination:
    // a row's off-diagonal entry in an unconstrained column is never changed before that row is
    // eliminated (only eliminating its own row, or its column's, changes it, and a constrained
    // column's term does not reach the right-hand side), and no diagonal entry is ever changed
    PrescribedRhs p;
    p.prescribed = prescribed;
    std::vector<char> constrained(a.dofs(), 0);
    for (const auto &pr : prescribed) constrained[pr.first] = 1;
    p.starts.push_back(0);
    for (const auto &pr : prescribed) {
        const size_t dof = pr.first, brow = dof / 3, row = dof % 3;
        for (int64_t at = a.starts[brow]; at < a.starts[brow + 1]; ++at)
            for (size_t q = 0; q < 3; ++q) {
                const size_t col = (size_t)a.columns[(size_t)at] * 3 + q;
                const double v = a.values[(size_t)at * 9 + row * 3 + q];
                if (col == dof) p.diagonal.push_back(v);
                else if (!constrained[col]) p.entries.push_back({col, v});
            }
        p.starts.push_back((int64_t)p.entries.size());
    }
    return p;
}

void apply_prescribed_rhs(const PrescribedRhs &p, std::vector<double> &rhs) {
    for (size_t k = 0; k < p.prescribed.size(); ++k)
        for (int64_t at = p.starts[k]; at < p.starts[k + 1]; ++at)
            rhs[p.entries[(size_t)at].first] -= p.entries[(size_t)at].second * p.prescribed[k].second;
    for (size_t k = 0; k < p.prescribed.size(); ++k) rhs[p.prescribed[k].first] = p.diagonal[k] * p.prescribed[k].second;
}

FreeProblem number_and_load(const Model &model) {
    const size_t nn = model.mesh.nodes.size();
    // study.rs:419-432: the fixed flags, then the check that every body is held
    std::vector<char> is_fixed(nn, 0);
    for (u32 node : model.fixed) {
        if (node >= nn) throw Error(SF_ERR_MODEL_NODE_OUT_OF_RANGE, "fixed node out of range", node);
        is_fixed[node] = 1;
    }
    StageClock clock;
    if (model.prescribed.empty())
        check_constraints(model.mesh, is_fixed);
    else
        check_held_dofs(model.mesh, is_fixed, model.prescribed);
    clock.lap("  held-body validation");
    const std::vector<double> load = model_load(model);
    clock.lap("  loads");
    FreeProblem p;
    p.free.assign(nn, FIXED);
    for (size_t n = 0; n < nn; ++n)
        if (!model.fixed.count((u32)n)) p.free[n] = (u32)p.free_count++;
    p.rhs.reserve(3 * p.free_count);
    for (size_t n = 0; n < nn; ++n)
        if (p.free[n] != FIXED)
            for (int a = 0; a < 3; ++a) p.rhs.push_back(load[3 * n + a]);
    // study.rs:662-683
    for (const auto &pr : model.prescribed) {
        if (pr.first.second > 2) throw Error(SF_ERR_MODEL_AXIS_OUT_OF_RANGE, "axis out of range", pr.first.second);
        if (!std::isfinite(pr.second)) throw Error(SF_ERR_MODEL_NON_FINITE_LOAD, "a prescribed value is not finite");
        if (pr.first.first >= nn) throw Error(SF_ERR_MODEL_NODE_OUT_OF_RANGE, "prescribed node out of range", pr.first.first);
        u32 index = p.free[pr.first.first];
        if (index == FIXED) throw Error(SF_ERR_MODEL_PRESCRIBED_ON_FIXED, "prescribed node is also fixed", pr.first.first);
        p.prescribed.push_back({3 * (size_t)index + pr.first.second, pr.second});
    }
    std::sort(p.prescribed.begin(), p.prescribed.end(), [](auto &a, auto &b) { return a.first < b.first; });
    bool any = false;
    for (double r : p.rhs) any = any || r != 0.0;
    for (auto &pr : p.prescribed) any = any || pr.second != 0.0;
    if (!any) throw Error(SF_ERR_MODEL_NO_EFFECTIVE_LOAD, "the model has no effective load");
    clock.lap("  numbering");
    return p;
}

void apply_prescribed(HostBsr<double> &a, std::vector<double> &rhs, const std::vector<std::pair<size_t, double>> &prescribed) {
    if (a.rb != 3 || a.cb != 3 || a.values.size() != a.blocks() * 9 || rhs.size() != a.dofs())
        throw Error(SF_ERR_INTERNAL, "prescribed elimination needs the assembled 3x3 blocks and a matching right-hand side");
    const PrescribedRhs program = prescribed_rhs_program(a, prescribed);
    // the slot of block (row, col), which the symmetric pattern holds for every stored mirror
    auto slot = [&](size_t brow, size_t bcol) -> size_t {
        const u32 *first = a.columns.data() + a.starts[brow], *last = a.columns.data() + a.starts[brow + 1];
        const u32 *it = std::lower_bound(first, last, (u32)bcol);
        if (it == last || *it != bcol) throw Error(SF_ERR_INTERNAL, "the stiffness pattern is not symmetric");
        return (size_t)(it - a.columns.data());
    };
    std::vector<std::pair<size_t, double>> row;
    for (const auto &pr : prescribed) {
        const size_t dof = pr.first, brow = dof / 3, p = dof % 3;
        // the row's off-diagonal entries in column order, as they stand before this row's updates,
        // cancelled in the row and in its mirror column
        row.clear();
        for (int64_t at = a.starts[brow]; at < a.starts[brow + 1]; ++at)
            for (size_t q = 0; q < 3; ++q) {
                const size_t col = (size_t)a.columns[(size_t)at] * 3 + q;
                if (col != dof) row.push_back({col, a.values[(size_t)at * 9 + p * 3 + q]});
            }
        for (const auto &e : row) {
            const size_t col = e.first;
            a.values[slot(col / 3, brow) * 9 + (col % 3) * 3 + p] += -e.second;
            a.values[slot(brow, col / 3) * 9 + p * 3 + col % 3] += -e.second;
        }
    }
    apply_prescribed_rhs(program, rhs);
}

std::vector<double> traction_rhs(const QuadMesh &mesh, const std::vector<Traction> &tractions, const std::vector<u32> &free,
                                 size_t free_count) {
    // the traction part of number_and_load, in the same order and arithmetic
    std::vector<double> load(3 * mesh.nodes.size(), 0.0);
    for (const auto &t : tractions) {
        for (double c : t.second)
            if (!std::isfinite(c)) throw Error(SF_ERR_MODEL_NON_FINITE_LOAD, "a traction is not finite");
        face_load(mesh, t.first, t.second, load);
    }
    std::vector<double> rhs;
    rhs.reserve(3 * free_count);
    for (size_t n = 0; n < free.size(); ++n)
        if (free[n] != FIXED)
            for (int a = 0; a < 3; ++a) rhs.push_back(load[3 * n + a]);
    return rhs;
}

std::vector<double> rigid_body_modes(const QuadMesh &mesh, const std::vector<u32> &free, size_t free_count) {
    // crates/solidfem-elastic/src/rigid_modes.rs:36-83: rotations about the free-node centroid,
    // scaled by the largest coordinate offset from it.
    double c[3] = {0, 0, 0};
    for (size_t n = 0; n < free.size(); ++n)
        if (free[n] != FIXED)
            for (int a = 0; a < 3; ++a) c[a] += mesh.nodes[n][a];
    const double divisor = (double)std::max<size_t>(free_count, 1);
    for (double &x : c) x /= divisor;
    double reach = 0.0;
    for (size_t n = 0; n < free.size(); ++n)
        if (free[n] != FIXED)
            for (int a = 0; a < 3; ++a) reach = std::max(reach, std::fabs(mesh.nodes[n][a] - c[a]));
    const double scale = reach > 0.0 ? 1.0 / reach : 1.0;
    std::vector<double> modes(18 * free_count, 0.0);
    parallel_for(free.size(), 16384, [&](size_t first, size_t last) {
        for (size_t n = first; n < last; ++n) {
            if (free[n] == FIXED) continue;
            const double x = (mesh.nodes[n][0] - c[0]) * scale, y = (mesh.nodes[n][1] - c[1]) * scale,
                         z = (mesh.nodes[n][2] - c[2]) * scale;
            double *m = &modes[18 * (size_t)free[n]];
            const double rows[18] = {1, 0, 0, 0, z, -y, 0, 1, 0, -z, 0, x, 0, 0, 1, y, -x, 0};
            std::memcpy(m, rows, sizeof(rows));
        }
    });
    return modes;
}

// ---- pattern ----------------------------------------------------------------------------

AssemblyPattern build_pattern(const QuadMesh &mesh, const std::vector<u32> &free, size_t free_count) {
    StageClock clock;
    // Node -> elements adjacency (ascending elements), by counting sort.
    const size_t ne = mesh.tets.size();
    std::vector<int64_t> adj_start(free_count + 1, 0);
    for (const auto &t : mesh.tets)
        for (u32 node : t)
            if (free[node] != FIXED) adj_start[(size_t)free[node] + 1] += 1;
    for (size_t i = 0; i < free_count; ++i) adj_start[i + 1] += adj_start[i];
    std::vector<u32> adj((size_t)adj_start[free_count]);  // (element << 4) | local
    {
        std::vector<int64_t> fill(adj_start.begin(), adj_start.end() - 1);
        for (size_t e = 0; e < ne; ++e)
            for (int a = 0; a < 10; ++a) {
                u32 f = free[mesh.tets[e][a]];
                if (f != FIXED) adj[(size_t)fill[f]++] = ((u32)e << 4) | (u32)a;
            }
    }
    clock.lap("  adjacency");
    AssemblyPattern out;
    auto &pat = out.pattern;
    pat.rows = pat.cols = free_count;
    pat.rb = pat.cb = 3;
    // A row's distinct free neighbours are marked in a per-task array indexed by free node,
    // stamped with the row so that nothing is cleared between rows; a second array gives each
    // neighbour's place in the row's sorted columns.
    // Pass 1: per row, the number of distinct free neighbours and of contributions.
    std::vector<int64_t> row_len(free_count, 0), row_contrib(free_count, 0);
    parallel_for(free_count, 2048, [&](size_t first, size_t last) {
        std::vector<u32> stamp(free_count, FIXED);
        for (size_t r = first; r < last; ++r) {
            int64_t distinct = 0, all = 0;
            for (int64_t at = adj_start[r]; at < adj_start[r + 1]; ++at)
                for (u32 node : mesh.tets[adj[(size_t)at] >> 4]) {
                    const u32 f = free[node];
                    if (f == FIXED) continue;
                    ++all;
                    if (stamp[f] != (u32)r) {
                        stamp[f] = (u32)r;
                        ++distinct;
                    }
                }
            row_len[r] = distinct;
            row_contrib[r] = all;
        }
    });
    pat.starts.assign(free_count + 1, 0);
    std::vector<int64_t> contrib_row(free_count + 1, 0);
    for (size_t r = 0; r < free_count; ++r) {
        pat.starts[r + 1] = pat.starts[r] + row_len[r];
        contrib_row[r + 1] = contrib_row[r] + row_contrib[r];
    }
    clock.lap("  pattern pass 1");
    const size_t nnzb = (size_t)pat.starts[free_count];
    pat.columns.resize(nnzb);
    out.contribution_starts = HostArray<int64_t>(nnzb + 1);
    out.contributions = HostArray<u32>((size_t)contrib_row[free_count]);
    // Pass 2: the sorted columns, then each block's contributions in ascending element order.
    parallel_for(free_count, 2048, [&](size_t first, size_t last) {
        std::vector<u32> stamp(free_count, FIXED), place(free_count, 0), cols;
        std::vector<int64_t> fill;
        for (size_t r = first; r < last; ++r) {
            cols.clear();
            for (int64_t at = adj_start[r]; at < adj_start[r + 1]; ++at)
                for (u32 node : mesh.tets[adj[(size_t)at] >> 4]) {
                    const u32 f = free[node];
                    if (f != FIXED && stamp[f] != (u32)r) {
                        stamp[f] = (u32)r;
                        cols.push_back(f);
                    }
                }
            std::sort(cols.begin(), cols.end());
            const size_t row0 = (size_t)pat.starts[r];
            for (size_t k = 0; k < cols.size(); ++k) {
                pat.columns[row0 + k] = cols[k];
                place[cols[k]] = (u32)k;
            }
            fill.assign(cols.size(), 0);
            for (int64_t at = adj_start[r]; at < adj_start[r + 1]; ++at)
                for (u32 node : mesh.tets[adj[(size_t)at] >> 4])
                    if (free[node] != FIXED) fill[place[free[node]]] += 1;
            // block-local offsets within this row's contribution range
            int64_t base = contrib_row[r];
            for (size_t k = 0; k < cols.size(); ++k) {
                out.contribution_starts[row0 + k] = base;
                const int64_t c = fill[k];
                fill[k] = base;
                base += c;
            }
            for (int64_t at = adj_start[r]; at < adj_start[r + 1]; ++at) {
                const u32 e = adj[(size_t)at] >> 4, a = adj[(size_t)at] & 15u;
                const auto &t = mesh.tets[e];
                for (u32 b = 0; b < 10; ++b) {
                    const u32 f = free[t[b]];
                    if (f == FIXED) continue;
                    out.contributions[(size_t)fill[place[f]]++] = (e << 7) | (10 * a + b);
                }
            }
        }
    });
    out.contribution_starts[nnzb] = contrib_row[free_count];
    clock.lap("  pattern pass 2");
    return out;
}

PeakStress peak_stress(const double *positions, size_t nodes, const double *nodal_vm, size_t stride, const double *point_values,
                       size_t points, const std::function<bool(const double *)> &inside) {
    PeakStress peak;
    auto offer = [&](bool at_point, size_t index, double value, const double *position) {
        if (inside && !inside(position)) return;
        if (peak.found && (std::isnan(peak.von_mises) || !(std::isnan(value) || value > peak.von_mises))) return;
        peak.found = true;
        peak.at_point = at_point;
        peak.index = index;
        peak.von_mises = value;
        for (int a = 0; a < 3; ++a) peak.position[a] = position[a];
    };
    for (size_t n = 0; n < nodes; ++n) offer(false, n, nodal_vm[n * stride], positions + 3 * n);
    for (size_t p = 0; p < points; ++p) offer(true, p, point_values[p * 10 + 9], point_values + p * 10);
    return peak;
}

namespace {

// check_matches (comparison.rs:199-224): one value per node, and the same whole number of
// integration points for every element (the point layout is element-major by construction)
void check_matches(const ComparedSolve &s, int which) {
    const size_t tets = s.mesh->tets.size();
    const bool per_node = s.nodes == s.mesh->nodes.size();
    const bool covers = tets > 0 && s.point_count > 0 && s.point_count % tets == 0;
    if (!per_node || !covers)
        throw Error(SF_ERR_MODEL_SOLUTION_MESH_MISMATCH, "a solution's arrays do not match its mesh", which);
}

// region_peak (comparison.rs:271-305): refused when no integration point is inside, or the peak
// is not finite or is zero
PeakStress region_peak(const ComparedSolve &s, const std::function<bool(const double *)> &inside, int which) {
    bool any = false;
    for (size_t p = 0; p < s.point_count && !any; ++p) any = !inside || inside(s.points + p * 10);
    if (!any) throw Error(SF_ERR_MODEL_NOTHING_INSIDE_REGION, "no integration point of a mesh is inside the region", which);
    std::vector<double> positions(s.mesh->nodes.size() * 3);
    for (size_t n = 0; n < s.mesh->nodes.size(); ++n)
        for (int a = 0; a < 3; ++a) positions[3 * n + a] = s.mesh->nodes[n][a];
    const PeakStress peak = peak_stress(positions.data(), s.nodes, s.nodal_vm, s.stride, s.points, s.point_count, inside);
    if (!peak.found) throw Error(SF_ERR_MODEL_NOTHING_INSIDE_REGION, "no integration point of a mesh is inside the region", which);
    if (!std::isfinite(peak.von_mises))
        throw Error(SF_ERR_MODEL_NON_FINITE_PEAK, "a peak stress is not a finite number", which, 0, peak.von_mises);
    if (peak.von_mises <= 0.0) throw Error(SF_ERR_MODEL_ZERO_PEAK, "a peak stress is zero", which);
    return peak;
}

// elements_inside (comparison.rs:226-239): elements with at least one point inside, counted once
size_t elements_inside(const ComparedSolve &s, const std::function<bool(const double *)> &inside) {
    const size_t per = s.point_count / s.mesh->tets.size();
    size_t count = 0, last = SIZE_MAX;
    for (size_t p = 0; p < s.point_count; ++p) {
        const size_t tet = p / per;
        if (last != tet && (!inside || inside(s.points + p * 10))) {
            count += 1;
            last = tet;
        }
    }
    return count;
}

// holder_volume (comparison.rs:241-269): a point's own element, or the mean volume of the
// elements holding a node, summed in element order
double holder_volume(const ComparedSolve &s, const PeakStress &peak) {
    const QuadMesh &m = *s.mesh;
    auto volume = [&](const std::array<u32, 10> &t) { return signed_volume(m.nodes[t[0]], m.nodes[t[1]], m.nodes[t[2]], m.nodes[t[3]]); };
    if (peak.at_point) return volume(m.tets[peak.index / (s.point_count / m.tets.size())]);
    double sum = 0.0;
    size_t count = 0;
    for (const auto &t : m.tets)
        if (std::find(t.begin(), t.end(), (u32)peak.index) != t.end()) {
            sum += volume(t);
            count += 1;
        }
    return sum / (double)count;
}

}  // namespace

MeshComparison compare_meshes(const ComparedSolve &coarse, const ComparedSolve &fine,
                              const std::function<bool(const double *)> &inside) {
    // MeshComparison::between, comparison.rs:101-141
    check_matches(coarse, SF_COMPARED_COARSE);
    check_matches(fine, SF_COMPARED_FINE);
    MeshComparison c;
    c.coarse.peak = region_peak(coarse, inside, SF_COMPARED_COARSE);
    c.fine.peak = region_peak(fine, inside, SF_COMPARED_FINE);
    c.coarse.elements = elements_inside(coarse, inside);
    c.fine.elements = elements_inside(fine, inside);
    if (c.fine.elements <= c.coarse.elements)
        throw Error(SF_ERR_MODEL_FINE_MESH_NOT_FINER, "the fine mesh has no more elements in the region than the coarse one",
                    (int64_t)c.coarse.elements, (int64_t)c.fine.elements);
    c.peak_change = (c.fine.peak.von_mises - c.coarse.peak.von_mises) / c.fine.peak.von_mises;
    double squared = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double d = c.fine.peak.position[a] - c.coarse.peak.position[a];
        squared += d * d;
    }
    c.peak_distance = std::sqrt(squared);
    c.coarse.peak_element_volume = holder_volume(coarse, c.coarse.peak);
    c.fine.peak_element_volume = holder_volume(fine, c.fine.peak);
    return c;
}

}  // namespace sf
