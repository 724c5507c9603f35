// pf2_server.hip — serving loop for many independent particle-stage
// simulations on one GPU (C ABI declared in include/pfem2_amd.h).
//
// No counterpart in the reference, which runs one simulation per process.
// Requests are admitted first-come first-served while their device footprint
// fits the server's HBM budget.  The server takes that budget from HIP once,
// at creation, and places every admitted simulation in a free range of it
// (first fit, ranges coalesced on release), so admission and retirement
// never call hipMalloc or hipFree while other simulations run.  Each active
// simulation owns a stream from a pool.  Batched mode (the default of the
// bindings): every step advances all active simulations of one dimension,
// precision and size class together with pf2_batch_step, one launch per
// kernel of the step for all of them, kept a few steps ahead on the group's
// stream; a moving mesh is moved on its simulation's stream after each step.
// The size classes keep small simulations from waiting on large ones: the
// groups' streams run side by side.  Otherwise each
// simulation is kept a few chunks of graph-replayed steps ahead on its own
// stream.  Completion is observed by polling the event of a simulation's
// last step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pfem2_amd.h"

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum ReqState { PENDING = 0, ACTIVE = 1, DONE = 2, FAILED = 3 };

struct Request {
    long long id = 0;
    int state = PENDING;
    // inputs, kept on the host until admission
    int dim = 2, n_cells = 0, n_vertices = 0, n_dofs = 0, degree = 1;
    std::vector<double> X, V;
    std::vector<int> cv, cd;
    pf2_params params{};
    int n_steps = 0;
    // prescribed mesh motion X0 + amp sin(omega k dt + phase) W after step k
    bool moving = false;
    std::vector<double> W;
    double amp = 0, omega = 0, phase = 0;
    unsigned long long bytes = 0;   // footprint reserved in the block (handle + field)
    // while active
    pf2_handle h = nullptr;
    hipStream_t stream = nullptr;
    size_t offset = 0;              // range of the server block
    double* dfield = nullptr;       // V (component-major) refreshed before every step
    int issued = 0;
    std::deque<hipEvent_t> inflight;
    long long n_start = 0, n_end = 0;
    // results
    std::vector<double> V_out;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int err = 0;
    std::string msg;
};

// first-fit ranges of the device block
struct RangeAlloc {
    std::map<size_t, size_t> free_;   // offset -> size
    void init(size_t size) {
        free_.clear();
        free_[0] = size;
    }
    bool take(size_t size, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= size) {
                *off = it->first;
                const size_t rest = it->second - size;
                const size_t at = it->first + size;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void give(size_t off, size_t size) {
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

constexpr size_t kAlign = 1u << 21;   // simulations start on 2 MiB boundaries

size_t round_up(size_t n) { return (n + kAlign - 1) & ~(kAlign - 1); }

}  // namespace

// a batch of the simulations of one dimension and precision, and the events
// of its steps in flight
struct BatchGroup {
    pf2_batch b = nullptr;
    hipStream_t stream = nullptr;
    std::deque<hipEvent_t> inflight;
};
constexpr int kBatchAhead = 3;   // batched steps queued ahead (below the batch's staging ring)
// size classes of batched simulations by cell count: below 16k, below 128k, larger
constexpr int kSizeClasses = 3;
int size_class(int n_cells) { return n_cells < 16384 ? 0 : (n_cells < 131072 ? 1 : 2); }

struct pf2_server_s {
    pf2_server_config cfg{};
    BatchGroup groups[2][2][kSizeClasses];   // [dim - 2][precision][size class]
    char* block = nullptr;
    size_t block_bytes = 0;
    RangeAlloc ranges;
    size_t in_use = 0, peak = 0;
    int max_active_seen = 0;
    long long next_id = 1;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::unordered_map<long long, Request*> all;
    std::vector<hipStream_t> free_streams, streams;
    std::vector<hipEvent_t> free_events, events;
    double t_first = -1, t_last = 0, updates = 0;
    std::vector<double> latencies;
    long long n_failed = 0;
    std::string err;
};

namespace {

thread_local std::string g_serr;

int sfail(int code, const std::string& msg) {
    g_serr = msg;
    return code;
}

hipEvent_t get_event(pf2_server_s* s) {
    if (s->free_events.empty()) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        s->events.push_back(e);
        return e;
    }
    hipEvent_t e = s->free_events.back();
    s->free_events.pop_back();
    return e;
}

hipStream_t get_stream(pf2_server_s* s) {
    if (s->free_streams.empty()) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
        s->streams.push_back(st);
        return st;
    }
    hipStream_t st = s->free_streams.back();
    s->free_streams.pop_back();
    return st;
}

void release(pf2_server_s* s, Request* r) {
    if (r->h) pf2_destroy(r->h);   // synchronises the request's stream
    r->h = nullptr;
    for (hipEvent_t e : r->inflight) s->free_events.push_back(e);
    r->inflight.clear();
    if (r->stream) s->free_streams.push_back(r->stream);
    r->stream = nullptr;
    if (r->bytes) {
        s->ranges.give(r->offset, r->bytes);
        s->in_use -= r->bytes;
    }
    r->dfield = nullptr;
    r->X.clear();
    r->X.shrink_to_fit();
    r->cv.clear();
    r->cv.shrink_to_fit();
    r->cd.clear();
    r->cd.shrink_to_fit();
    r->V.clear();
    r->V.shrink_to_fit();
    r->W.clear();
    r->W.shrink_to_fit();
}

void mark_failed(pf2_server_s* s, Request* r, int code, const std::string& msg) {
    r->err = code;
    r->msg = msg;
    r->state = FAILED;
    r->t_done = now_s();
    ++s->n_failed;
    release(s, r);
}

pf2_mesh mesh_of(const Request* r) {
    pf2_mesh m{};
    m.dim = r->dim;
    m.n_cells = r->n_cells;
    m.n_vertices = r->n_vertices;
    m.n_dofs = r->n_dofs;
    m.vertices = r->X.data();
    m.cell_vertices = r->cv.data();
    m.cell_dofs = r->cd.data();
    m.degree = r->degree;
    return m;
}

unsigned long long footprint(const Request* r) {
    pf2_mesh m = mesh_of(r);
    unsigned long long b = 0;
    if (pf2_estimate_bytes(&m, &r->params, 0, &b)) return 0;
    return round_up(b + (size_t)r->n_dofs * r->dim * 8 + 256);
}

// create the simulation inside its range, seed the lattice
int admit(pf2_server_s* s, Request* r) {
    size_t off = 0;
    if (!s->ranges.take(r->bytes, &off)) return 1;   // fragmented: try again later
    r->offset = off;
    s->in_use += r->bytes;
    s->peak = std::max(s->peak, s->in_use);
    r->stream = get_stream(s);
    if (!r->stream) {
        mark_failed(s, r, PF2_E_HIP, "stream creation failed");
        return 0;
    }
    char* base = s->block + off;
    const size_t field_bytes = ((size_t)r->n_dofs * r->dim * 8 + 255) & ~(size_t)255;
    r->dfield = (double*)base;
    pf2_mesh m = mesh_of(r);
    int rc = pf2_create_in(&m, &r->params, 0, base + field_bytes, r->bytes - field_bytes, r->stream, &r->h);
    if (rc) {
        mark_failed(s, r, rc, pf2_last_error());
        return 0;
    }
    if (hipMemcpyAsync(r->dfield, r->V.data(), r->V.size() * 8, hipMemcpyHostToDevice, r->stream) != hipSuccess) {
        mark_failed(s, r, PF2_E_HIP, "field upload failed");
        return 0;
    }
    long long n = 0;
    if ((rc = pf2_upload_fields(r->h, r->V.data(), r->V.data())) || (rc = pf2_seed_particles(r->h)) ||
        (rc = pf2_n_global_particles(r->h, &n)) || (rc = pf2_set_graphs(r->h, s->cfg.use_graphs)) ||
        (r->moving && (rc = pf2_set_motion(r->h, r->W.data())))) {
        mark_failed(s, r, rc, pf2_last_error());
        return 0;
    }
    r->n_start = n;
    r->state = ACTIVE;
    r->t_admit = now_s();
    s->active.push_back(r);
    s->max_active_seen = std::max(s->max_active_seen, (int)s->active.size());
    return 0;
}

bool batched(const pf2_server_s* s, const Request* r) {
    return s->cfg.batched && r->degree == 1 && (s->cfg.batch_max_cells <= 0 || r->n_cells <= s->cfg.batch_max_cells);
}

// the simulation's step k (1-based) is done: the mesh move of a moving request
int after_step(Request* r, int k) {
    if (!r->moving) return 0;
    const double t = (double)k * r->params.dt;
    return pf2_move_mesh_scaled(r->h, r->amp * std::sin(r->omega * t + r->phase));
}

// issue steps so that `depth` chunks are queued
int issue(pf2_server_s* s, Request* r) {
    if (batched(s, r)) return 0;   // stepped by issue_batched
    const int chunk = s->cfg.steps_per_chunk > 0 ? s->cfg.steps_per_chunk : 8;
    const int depth = s->cfg.chunks_ahead > 0 ? s->cfg.chunks_ahead : 2;
    while ((int)r->inflight.size() < depth && r->issued < r->n_steps) {
        const int k = std::min(chunk, r->n_steps - r->issued);
        for (int i = 0; i < k; ++i) {
            int rc = pf2_step_with_field(r->h, r->dfield);
            if (!rc) rc = after_step(r, r->issued + i + 1);
            if (rc) {
                mark_failed(s, r, rc, pf2_last_error());
                return 1;
            }
        }
        r->issued += k;
        hipEvent_t e = get_event(s);
        if (!e || hipEventRecord(e, r->stream) != hipSuccess) {
            mark_failed(s, r, PF2_E_HIP, "event record failed");
            return 1;
        }
        r->inflight.push_back(e);
    }
    return 0;
}

// Batched mode: up to kBatchAhead batched steps queued per group; each takes
// every active simulation of the group that has steps left.  A simulation
// whose last step was issued gets an event on its stream (after the batch
// and its mesh move), which progress() polls.
int issue_batched(pf2_server_s* s) {
    for (int di = 0; di < 2; ++di)
        for (int k = 0; k < 2 * kSizeClasses; ++k) {
            const int pr = k / kSizeClasses, sc = k % kSizeClasses;
            BatchGroup& g = s->groups[di][pr][sc];
            while (!g.inflight.empty() && hipEventQuery(g.inflight.front()) == hipSuccess) {
                s->free_events.push_back(g.inflight.front());
                g.inflight.pop_front();
            }
            for (;;) {
                std::vector<Request*> mem;
                for (Request* r : s->active)
                    if (r->state == ACTIVE && batched(s, r) && r->dim == di + 2 && r->params.precision == pr &&
                        size_class(r->n_cells) == sc && r->issued < r->n_steps)
                        mem.push_back(r);
                if (mem.empty() || (int)g.inflight.size() >= kBatchAhead) break;
                if (!g.b) {
                    if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess)
                        return sfail(PF2_E_HIP, "stream creation failed");
                    const int cap = std::max(256, s->cfg.max_active);
                    if (int rc = pf2_batch_create(cap, g.stream, &g.b)) return sfail(rc, pf2_last_error());
                }
                std::vector<pf2_handle> hs;
                std::vector<const double*> fs;
                for (Request* r : mem) {
                    hs.push_back(r->h);
                    fs.push_back(r->dfield);
                }
                if (int rc = pf2_batch_step(g.b, (int)hs.size(), hs.data(), fs.data())) {
                    const std::string msg = pf2_last_error();
                    for (Request* r : mem) mark_failed(s, r, rc, msg);
                    break;
                }
                hipEvent_t ge = get_event(s);
                if (!ge || hipEventRecord(ge, g.stream) != hipSuccess) return sfail(PF2_E_HIP, "event record failed");
                g.inflight.push_back(ge);
                for (Request* r : mem) {
                    r->issued += 1;
                    if (int rc = after_step(r, r->issued)) {
                        mark_failed(s, r, rc, pf2_last_error());
                        continue;
                    }
                    if (r->issued == r->n_steps) {
                        hipEvent_t e = get_event(s);
                        if (!e || hipEventRecord(e, r->stream) != hipSuccess) {
                            mark_failed(s, r, PF2_E_HIP, "event record failed");
                            continue;
                        }
                        r->inflight.push_back(e);
                    }
                }
            }
        }
    return 0;
}

// drop finished chunks; returns true once every step has completed
bool progress(pf2_server_s* s, Request* r) {
    while (!r->inflight.empty()) {
        const hipError_t q = hipEventQuery(r->inflight.front());
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
            mark_failed(s, r, PF2_E_HIP, std::string("step failed: ") + hipGetErrorString(q));
            return false;
        }
        s->free_events.push_back(r->inflight.front());
        r->inflight.pop_front();
    }
    return r->state == ACTIVE && r->issued == r->n_steps && r->inflight.empty();
}

void retire(pf2_server_s* s, Request* r) {
    r->t_done = now_s();
    r->V_out.resize((size_t)r->n_dofs * r->dim);
    int rc = pf2_download_field(r->h, r->V_out.data());
    long long n = 0;
    if (!rc) rc = pf2_n_global_particles(r->h, &n);
    if (rc) {
        mark_failed(s, r, rc, pf2_last_error());
        return;
    }
    r->n_end = n;
    r->state = DONE;
    s->latencies.push_back((r->t_done - r->t_submit) * 1e3);
    s->updates += 0.5 * (double)(r->n_start + r->n_end) * r->n_steps;
    s->t_last = std::max(s->t_last, r->t_done);
    release(s, r);
}

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

const char* pf2_server_last_error(void) { return g_serr.c_str(); }

int pf2_server_create(const pf2_server_config* cfg, pf2_server* out) {
    if (!cfg || !out) return sfail(PF2_E_ARG, "cfg/out is NULL");
    *out = nullptr;
    size_t budget = (size_t)cfg->hbm_budget;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return sfail(PF2_E_HIP, "hipMemGetInfo failed");
        budget = (size_t)(0.8 * (double)fr);
    }
    budget &= ~(kAlign - 1);
    if (budget < kAlign) return sfail(PF2_E_ARG, "HBM budget below 2 MiB");
    pf2_server_s* s = new pf2_server_s();
    s->cfg = *cfg;
    if (hipMalloc((void**)&s->block, budget) != hipSuccess) {
        delete s;
        return sfail(PF2_E_HIP, "could not reserve the HBM budget");
    }
    s->block_bytes = budget;
    s->ranges.init(budget);
    *out = s;
    return 0;
}

int pf2_server_destroy(pf2_server s) {
    if (!s) return 0;
    for (int i = 0; i < 2 * 2 * kSizeClasses; ++i) {
        BatchGroup& g = (&s->groups[0][0][0])[i];
        if (g.b) pf2_batch_destroy(g.b);   // synchronises the group's stream
        if (g.stream) hipStreamDestroy(g.stream);
        for (hipEvent_t e : g.inflight) s->free_events.push_back(e);
    }
    for (Request* r : s->active) release(s, r);
    s->active.clear();
    for (auto& kv : s->all) delete kv.second;
    for (hipStream_t st : s->streams) hipStreamDestroy(st);
    for (hipEvent_t e : s->events) hipEventDestroy(e);
    if (s->block) hipFree(s->block);
    delete s;
    return 0;
}

static int submit_impl(pf2_server_s* s, const pf2_mesh* mesh, const pf2_params* params, const double* V, int n_steps,
                const double* W, double amp, double omega, double phase, long long* request_id) {
    if (!s || !mesh || !params || !V || !request_id) return sfail(PF2_E_ARG, "NULL argument");
    if (n_steps < 0) return sfail(PF2_E_ARG, "n_steps must be >= 0");
    if (mesh->dim != 2 && mesh->dim != 3) return sfail(PF2_E_ARG, "dim must be 2 or 3");
    if (mesh->n_cells < 1 || mesh->n_vertices < 1 || mesh->n_dofs < 1 || !mesh->vertices || !mesh->cell_vertices)
        return sfail(PF2_E_ARG, "empty mesh");
    Request* r = new Request();
    r->dim = mesh->dim;
    r->n_cells = mesh->n_cells;
    r->n_vertices = mesh->n_vertices;
    r->n_dofs = mesh->n_dofs;
    r->degree = mesh->degree == 2 ? 2 : 1;
    if (mesh->degree == 2 && !mesh->cell_dofs) {
        delete r;
        return sfail(PF2_E_ARG, "a degree-2 field needs cell_dofs");
    }
    const size_t nck = (size_t)mesh->n_cells << mesh->dim;
    const size_t ndc = r->degree == 2 ? (size_t)mesh->n_cells * (mesh->dim == 2 ? 9 : 27) : nck;
    r->X.assign(mesh->vertices, mesh->vertices + (size_t)mesh->n_vertices * mesh->dim);
    r->cv.assign(mesh->cell_vertices, mesh->cell_vertices + nck);
    const int* cd = mesh->cell_dofs ? mesh->cell_dofs : mesh->cell_vertices;
    r->cd.assign(cd, cd + ndc);
    r->V.assign(V, V + (size_t)mesh->n_dofs * mesh->dim);
    r->params = *params;
    r->n_steps = n_steps;
    if (W) {
        r->moving = true;
        r->params.keep_velocity_ext = 1;
        r->W.assign(W, W + (size_t)mesh->n_vertices * mesh->dim);
        r->amp = amp;
        r->omega = omega;
        r->phase = phase;
    }
    r->bytes = footprint(r);
    if (r->bytes == 0) {
        delete r;
        return sfail(PF2_E_ARG, std::string("invalid request: ") + pf2_last_error());
    }
    if (r->bytes > s->block_bytes) {
        delete r;
        return sfail(PF2_E_CAPACITY, "request needs more HBM than the server budget");
    }
    r->id = s->next_id++;
    r->t_submit = now_s();
    if (s->t_first < 0) s->t_first = r->t_submit;
    s->all[r->id] = r;
    s->pending.push_back(r);
    *request_id = r->id;
    return 0;
}

int pf2_server_submit(pf2_server s, const pf2_mesh* mesh, const pf2_params* params, const double* V, int n_steps,
                      long long* request_id) {
    return submit_impl(s, mesh, params, V, n_steps, nullptr, 0, 0, 0, request_id);
}

int pf2_server_submit_moving(pf2_server s, const pf2_mesh* mesh, const pf2_params* params, const double* V,
                             int n_steps, const double* W, double amplitude, double omega, double phase,
                             long long* request_id) {
    if (!W) return sfail(PF2_E_ARG, "W is NULL");
    return submit_impl(s, mesh, params, V, n_steps, W, amplitude, omega, phase, request_id);
}

int pf2_server_poll(pf2_server s, int* n_pending, int* n_active) {
    if (!s) return sfail(PF2_E_ARG, "server is NULL");
    // retire finished simulations
    for (size_t i = 0; i < s->active.size();) {
        Request* r = s->active[i];
        if (r->state == ACTIVE && progress(s, r)) retire(s, r);
        if (r->state != ACTIVE) {
            s->active.erase(s->active.begin() + i);
            continue;
        }
        ++i;
    }
    // admit in arrival order while the footprint fits
    const int max_active = s->cfg.max_active > 0 ? s->cfg.max_active : 1 << 30;
    while (!s->pending.empty() && (int)s->active.size() < max_active) {
        Request* r = s->pending.front();
        if (s->in_use + r->bytes > s->block_bytes) break;
        if (admit(s, r)) break;
        s->pending.pop_front();
    }
    // keep the batches and every other active simulation a few steps ahead
    if (s->cfg.batched)
        if (int rc = issue_batched(s)) return rc;
    for (size_t i = 0; i < s->active.size();) {
        Request* r = s->active[i];
        issue(s, r);
        if (r->state != ACTIVE) {
            s->active.erase(s->active.begin() + i);
            continue;
        }
        ++i;
    }
    if (n_pending) *n_pending = (int)s->pending.size();
    if (n_active) *n_active = (int)s->active.size();
    return 0;
}

int pf2_server_drain(pf2_server s) {
    if (!s) return sfail(PF2_E_ARG, "server is NULL");
    int np = 0, na = 0;
    for (;;) {
        if (int rc = pf2_server_poll(s, &np, &na)) return rc;
        if (np == 0 && na == 0) return 0;
        if (na == 0 && np > 0) return sfail(PF2_E_STATE, "pending requests cannot be admitted");
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

int pf2_server_result(pf2_server s, long long request_id, double* V_out, long long* n_particles, double* latency_ms) {
    if (!s) return sfail(PF2_E_ARG, "server is NULL");
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return sfail(PF2_E_ARG, "unknown request id");
    Request* r = it->second;
    if (r->state == FAILED) return sfail(r->err ? r->err : PF2_E_HIP, r->msg);
    if (r->state != DONE) return sfail(PF2_E_STATE, "request not finished");
    if (V_out) std::memcpy(V_out, r->V_out.data(), r->V_out.size() * 8);
    if (n_particles) *n_particles = r->n_end;
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    return 0;
}

int pf2_server_stats(pf2_server s, double* out, int cap) {
    if (!s || !out) return sfail(PF2_E_ARG, "NULL argument");
    const double wall = s->t_first >= 0 && s->t_last > s->t_first ? s->t_last - s->t_first : 0.0;
    double mean = 0.0;
    for (double l : s->latencies) mean += l;
    if (!s->latencies.empty()) mean /= (double)s->latencies.size();
    const double v[10] = {(double)s->latencies.size(),
                          (double)s->n_failed,
                          quantile(s->latencies, 0.50),
                          quantile(s->latencies, 0.99),
                          s->updates,
                          wall,
                          wall > 0 ? s->updates / wall : 0.0,
                          (double)s->peak,
                          (double)s->max_active_seen,
                          mean};
    const int n = std::min(cap, 10);
    for (int i = 0; i < n; ++i) out[i] = v[i];
    return n;
}

}  // extern "C"
