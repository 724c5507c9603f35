// pf2_lib.hip — handle management and the C ABI declared in include/pfem2_amd.h.
//
// Host side of the particle stage: it uploads the mesh once, builds the
// vertex star (pfem2ParticleHandler::initialize_maps,
// TurekBenchmark/pfem2particle.cpp:215-219) and the DoF -> cell table of the
// projection, and orders the kernel launches of one step on the handle's
// stream.  No host synchronisation happens inside a step; counts stay on
// the device and every particle kernel loops over the device-side count.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfem2_amd.h"
#include "pf2_batch.hpp"
#include "pf2_fem.hpp"
#include "pf2_kernels.hpp"
#include "pf2_mixed.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPC(expr)                                                                            \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(PF2_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

// Timing ids of pf2_kernel_times / pf2_kernel_name.  The ids of kernels that
// no longer exist (K_GROUP, K_QPREFIX) stay reserved, so that a caller's
// numbering does not shift; they report no launches.
enum KernelId {
    K_MOVE = 0, K_GROUP, K_CELL_COUNT, K_CELL_WRITE, K_SCAN, K_FINALIZE, K_PARTIALS, K_NODE, K_WALL,
    K_SEED, K_MISC, K_MOVE32, K_CELL_WRITE32, K_PARTIALS32, K_SEARCH, K_SEARCH32, K_QPREFIX,
    K_SEARCH_R1, K_SEARCH_R2, K_SEARCH32_R1, K_SEARCH32_R2, K_CELL_BIG, K_RESORT, K_RESORT32, K_CELL_MAP,
    K_CELL_GATHER, K_CELL_GATHER32, K_RESORT_SEARCH, K_RESORT_SEARCH32, K_CELL_PEND, K_SETTLE, K_NKERNELS
};
const char* kKernelNames[K_NKERNELS] = {"k_move", "k_group", "k_cell_count", "k_cell_write", "scan",
                                        "k_finalize", "k_cell_partials", "k_node", "wall_fixup",
                                        "k_seed", "misc", "k_move32", "k_cell_write32", "k_cell_partials32",
                                        "k_search_resume", "k_search_resume32", "k_queue_prefix",
                                        "k_search_resume_r1", "k_search_resume_r2", "k_search_resume32_r1",
                                        "k_search_resume32_r2", "k_cell_big", "k_resort", "k_resort32",
                                        "k_cell_map", "k_cell_gather", "k_cell_gather32", "k_resort_search",
                                        "k_resort_search32", "k_cell_pend", "k_settle"};

// bytes obtained through dalloc on this thread (a handle's footprint is the
// difference across its creation)
thread_local size_t g_alloc_bytes = 0;

// While a handle is created inside caller-provided device memory, dalloc
// carves 256-byte aligned pieces out of it instead of calling hipMalloc.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
};
thread_local Arena g_arena;

template <typename T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    const size_t bytes = n * sizeof(T);
    if (g_arena.base) {
        const size_t at = (g_arena.used + 255) & ~(size_t)255;
        if (at + bytes > g_arena.size) return fail(PF2_E_CAPACITY, "device memory block too small for the handle");
        *p = (T*)(g_arena.base + at);
        g_arena.used = at + bytes;
    } else {
        hipError_t e = hipMalloc((void**)p, bytes);
        if (e != hipSuccess) return fail(PF2_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    g_alloc_bytes += bytes;
    return 0;
}

struct PBufs {
    int* cell = nullptr;
    uint32_t* id = nullptr;
    double* x = nullptr;
    double* xi = nullptr;
    double* v = nullptr;
    double* vext = nullptr;
    // FP32-mixed storage: one record per particle (xi, v)
    float4* s32a = nullptr;
    float2* s32b = nullptr;   // 3-D only
};

pf2::Rec32 rec32(const PBufs& b) { return pf2::Rec32{b.s32a, b.s32b}; }

}  // namespace

struct pf2_handle_s {
    int dim = 2, nv = 4;
    int deg = 1, nd = 4;   // velocity field degree and its DoFs per cell
    int n_cells = 0, n_vertices = 0, n_dofs = 0;
    pf2_params params{};
    long long cap = 0;
    long long n_est = 0;   // host-side estimate of the particle count (grid sizes only)
    bool pending = false;  // the particles are in a pending layout (k_cell_pend)
    hipStream_t stream = nullptr;
    // host copies
    std::vector<double> X;
    std::vector<int> cv, cd, star_off, star_cell, dof_off, dof_cell, cell_star;   // cell_star: pairs
    std::vector<double> star_dir;
    // device mesh
    double* dX = nullptr;
    int *dcv = nullptr, *dcd = nullptr, *dstar_off = nullptr, *dstar_cell = nullptr;
    double* dstar_dir = nullptr;
    int *ddof_off = nullptr, *ddof_cell = nullptr;
    double* dftol = nullptr;
    double* dcellX = nullptr;
    unsigned char* dvbnd = nullptr;   // per vertex: on a boundary face
    int* dvdof = nullptr;             // per vertex: its velocity DoF (-1: in no cell)
    int2* dcell_star = nullptr;
    std::vector<int> cell_nbr;   // face neighbours, 4 (2-D) or 8 (3-D, 6 used) per cell
    int4* dcell_nbr = nullptr;
    double *dcellV = nullptr, *dcellDV = nullptr;
    // fields (node-major)
    double *dV = nullptr, *dVold = nullptr, *dtmp = nullptr;
    // mixed-precision fields and geometry
    float *dV32 = nullptr, *dVold32 = nullptr;
    float* dgeo32 = nullptr;
    // particles
    PBufs buf[2];
    int cur = 0;
    bool fp64_valid = true;   // FP64 copy of the state is current (always true in the FP64 path)
    int* off[2] = {nullptr, nullptr};
    int cur_off = 0;
    // work
    int *dest = nullptr, *pkey = nullptr, *spos = nullptr;
    int* omap = nullptr;   // regroup output map (k_cell_map -> k_cell_gather)
    unsigned char* dcmoved = nullptr;   // per cell: the last geometry update changed it
    int geo_updates = 0;                // geometry updates since the last re-sort
    int *arr_cnt = nullptr, *arr_slot = nullptr, *ovf_head = nullptr;
    int2* ovf = nullptr;
    int aslots = 0;
    unsigned long long *fs_cnt = nullptr, *fs_off = nullptr;
    unsigned long long* hitmask = nullptr;
    int* nsurv = nullptr;
    double* partial = nullptr;
    int* counters = nullptr;
    int* big = nullptr;
    int* big2 = nullptr;
    int* qcount = nullptr;
    int* qprefix = nullptr;
    long long shard_cap = 0;
    int2* queue[2] = {nullptr, nullptr};
    double* park_vext = nullptr;
    int* park_s = nullptr;
    uint32_t* pcounter = nullptr;
    unsigned long long* stats = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    unsigned long long* scan_tiles = nullptr;   // tile sums of the batched offset scan
    // walls
    int n_walls = 0, n_wall_cells = 0;
    int *w_dof = nullptr, *w_mask = nullptr, *w_fix = nullptr, *w_vwall = nullptr, *w_cells = nullptr;
    double *w_bc = nullptr, *w_val = nullptr;
    // profiling
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    double k_ms[K_NKERNELS] = {};
    long long k_n[K_NKERNELS] = {};
    // HIP graphs of one step, keyed by the buffer parity (cur, cur_off) and
    // by whether the step first refreshes V from a device field
    bool graphs = false;
    hipGraphExec_t gexec[8] = {};
    const double* g_field = nullptr;   // device field baked into the refreshing graphs
    size_t bytes = 0;                  // device memory held by the handle
    bool in_block = false;             // created inside caller memory: nothing of it is freed here
    bool packed = true;                // per-cell copies of V, V - V_old are current
    // prescribed mesh motion X = X0 + a W (handles with keep_velocity_ext)
    double *dX0 = nullptr, *dW = nullptr;
    bool x_host_stale = false;         // the device vertices moved without the host copy
    bool own_stream = true;
    int device = 0;                    // the device current at creation
};

namespace {

using namespace pf2;

int mixed(const pf2_handle_s* h) { return h->params.precision == PF2_FP32_MIXED; }

int flush_events(pf2_handle_s* h) {
    if (h->ev_pending.empty()) return 0;
    HIPC(hipStreamSynchronize(h->stream));
    for (auto& p : h->ev_pending) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        h->k_ms[p.first] += ms;
        h->k_n[p.first] += 1;
        h->ev_free.push_back(p.second);
    }
    h->ev_pending.clear();
    return 0;
}

// brackets one launch with events when profiling is on
struct Timed {
    pf2_handle_s* h;
    int kid;
    std::pair<hipEvent_t, hipEvent_t> ev{};
    bool on = false;
    Timed(pf2_handle_s* h_, int kid_) : h(h_), kid(kid_) {
        if (!h->profiling) return;
        if (h->ev_free.empty()) {
            if (h->ev_pending.size() >= 4096) flush_events(h);
        }
        if (h->ev_free.empty()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
            h->ev_free.push_back({a, b});
        }
        ev = h->ev_free.back();
        h->ev_free.pop_back();
        hipEventRecord(ev.first, h->stream);
        on = true;
    }
    ~Timed() {
        if (!on) return;
        hipEventRecord(ev.second, h->stream);
        h->ev_pending.push_back({kid, ev});
    }
};

// cell kernels: 16 (or 8) cells per 256-thread block, blocks looping over cells
// workgroups of a cell kernel (16 cells per workgroup, grid-stride beyond
// `cap` workgroups).  The census runs with a cap of 8192 (one pass over the
// ~100k cells of the cylinder workload: each wavefront's chain of dependent
// reads runs once), the write and sum passes with 4096 (measured fastest)
int cell_grid(int n_cells, int cap = 4096) {
    int g = (n_cells + 15) / 16;
    return g < cap ? (g < 1 ? 1 : g) : cap;
}

int pack_fields(pf2_handle_s* h);
int ensure_packed(pf2_handle_s* h);

// bound of the per-cell particle count after a census with reseed
long long fuse_bound(const pf2_handle_s* h) {
    long long b = h->params.max_per_part;
    for (int d = 0; d < h->dim; ++d) b *= h->params.quantities[d] + 1;
    return b;
}

// search grid per round: the first round sees most parked particles; later
// rounds see fewer, the last one loops every remaining particle to the end
int round_grid(int qgrid, int r) {
    if (r == 0) return qgrid;
    return r == MOVE_ROUNDS ? qgrid / 2 + 1 : qgrid / 4 + 1;
}

int grid_for(long long n, int block, int cap_blocks = 65535 * 4) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap_blocks) g = cap_blocks;
    return (int)g;
}

template <int DIM>
DevMesh<DIM> dev_mesh(pf2_handle_s* h) {
    DevMesh<DIM> m;
    m.n_cells = h->n_cells;
    m.n_vertices = h->n_vertices;
    m.n_dofs = h->n_dofs;
    m.nd = h->nd;
    m.X = h->dX;
    m.cv = h->dcv;
    m.cd = h->dcd;
    m.star_off = h->dstar_off;
    m.star_cell = h->dstar_cell;
    m.star_dir = h->dstar_dir;
    m.dof_off = h->ddof_off;
    m.dof_cell = h->ddof_cell;
    m.ftol = h->dftol;
    m.cellX = h->dcellX;
    m.cell_star = h->dcell_star;
    m.cell_nbr = h->dcell_nbr;
    m.cellV = h->dcellV;     // null for FP32-mixed handles
    m.cellDV = h->dcellDV;
    m.Vn = h->dV;
    m.cmoved = h->dcmoved;
    return m;
}

DevParticles dev_particles(pf2_handle_s* h, int which) {
    DevParticles p;
    const PBufs& b = h->buf[which];
    p.cell = b.cell;
    p.id = b.id;
    p.x = b.x;
    p.xi = b.xi;
    p.v = b.v;
    p.vext = b.vext;
    p.cap = h->cap;
    return p;
}

DevQueues dev_queues(pf2_handle_s* h) {
    DevQueues q;
    q.data[0] = h->queue[0];
    q.data[1] = h->queue[1];
    q.count = h->qcount;
    q.prefix = h->qprefix;
    q.shard_cap = (int)h->shard_cap;
    q.park_vext = h->park_vext;
    q.park_s = h->park_s;
    return q;
}

DevParams dev_params(pf2_handle_s* h) {
    DevParams p;
    p.h = h->params.dt / h->params.substeps;
    p.substeps = h->params.substeps;
    p.max_per_part = h->params.max_per_part;
    p.nparts_in = 1;
    for (int d = 0; d < 3; ++d) p.q[d] = d < h->dim ? h->params.quantities[d] : 1;
    for (int d = 0; d < 3; ++d) p.hq[d] = 1.0 / p.q[d];
    for (int d = 0; d < h->dim; ++d) p.nparts_in *= p.q[d];
    p.keep_vext = h->params.keep_velocity_ext;
    return p;
}

DevWork dev_work(pf2_handle_s* h) {
    DevWork w;
    w.dest = h->dest;
    w.pkey = h->pkey;
    w.spos = h->spos;
    w.arr_cnt = h->arr_cnt;
    w.arr_slot = h->arr_slot;
    w.aslots = h->aslots;
    w.pend_mark = 0;
    w.ovf = h->ovf;
    w.ovf_head = h->ovf_head;
    w.fs_cnt = h->fs_cnt;
    w.fs_off = h->fs_off;
    w.new_off = h->off[1 - h->cur_off];
    w.hitmask = h->hitmask;
    w.nsurv = h->nsurv;
    w.partial = h->partial;
    w.counters = h->counters;
    w.particle_counter = h->pcounter;
    w.stats = h->stats;
    w.big = h->big;
    w.big2 = h->big2;
    return w;
}

// vertex_to_cell_centers_directions restated: normalise(vertex average - vertex)
void compute_star_dirs(pf2_handle_s* h) {
    const int dim = h->dim;
    h->star_dir.assign((size_t)h->star_off[h->n_vertices] * dim, 0.0);
    for (int v = 0; v < h->n_vertices; ++v)
        for (int e = h->star_off[v]; e < h->star_off[v + 1]; ++e) {
            const int c = h->star_cell[e];
            double ctr[3] = {0, 0, 0};
            for (int k = 0; k < h->nv; ++k) {
                const int gv = h->cv[(size_t)c * h->nv + k];
                for (int d = 0; d < dim; ++d) ctr[d] += h->X[(size_t)gv * dim + d];
            }
            const double inv_n = 1.0 / h->nv;
            double t[3], s = 0.0;
            for (int d = 0; d < dim; ++d) t[d] = ctr[d] * inv_n - h->X[(size_t)v * dim + d];
            for (int d = 0; d < dim; ++d) s += t[d] * t[d];
            const double r = 1.0 / std::sqrt(s);
            for (int d = 0; d < dim; ++d) h->star_dir[(size_t)e * dim + d] = t[d] * r;
        }
}

int upload_star_dirs(pf2_handle_s* h) {
    HIPC(hipMemcpyAsync(h->dstar_dir, h->star_dir.data(), h->star_dir.size() * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
    return 0;
}

int check_params(int dim, const pf2_params* p) {
    if (!p) return fail(PF2_E_ARG, "params is NULL");
    if (!(p->dt > 0.0)) return fail(PF2_E_ARG, "dt must be positive");
    if (p->substeps < 1) return fail(PF2_E_ARG, "substeps must be >= 1");
    if (p->max_per_part < 1) return fail(PF2_E_ARG, "max_per_part must be >= 1");
    long long np = 1;
    for (int d = 0; d < dim; ++d) {
        if (p->quantities[d] < 1) return fail(PF2_E_ARG, "quantities must be >= 1");
        np *= p->quantities[d];
    }
    if (np > 64) return fail(PF2_E_ARG, "at most 64 lattice parts per cell are supported");
    if (p->precision != PF2_FP64 && p->precision != PF2_FP32_MIXED) return fail(PF2_E_ARG, "unknown precision");
    return 0;
}

long long default_capacity(int dim, int n_cells, const pf2_params* p) {
    long long per = p->max_per_part;
    for (int d = 0; d < dim; ++d) per *= (p->quantities[d] + 1);
    long long lattice = 1;
    for (int d = 0; d < dim; ++d) lattice *= p->quantities[d];
    per = std::max(per, lattice);
    return per * (long long)n_cells;
}

// arrival slots per cell: the census row capacity of the lattice, so that a
// cell with more arrivals is one the census hands to k_cell_big anyway
int arrival_slots(int dim, const pf2_params* p) {
    long long parts = 1;
    for (int d = 0; d < dim; ++d) parts *= p->quantities[d];
    return census_small_row((int)parts, p->max_per_part) ? GCAP : GCAP_L;
}

int alloc_particles(pf2_handle_s* h) {
    for (int b = 0; b < 2; ++b) {
        PBufs& pb = h->buf[b];
        if (dalloc(&pb.cell, h->cap) || dalloc(&pb.id, h->cap) || dalloc(&pb.x, h->cap * h->dim) ||
            dalloc(&pb.xi, h->cap * h->dim) || dalloc(&pb.v, h->cap * h->dim))
            return PF2_E_HIP;
        if (h->params.keep_velocity_ext && dalloc(&pb.vext, h->cap * h->dim)) return PF2_E_HIP;
        if (mixed(h) && (dalloc(&pb.s32a, h->cap) || (h->dim == 3 && dalloc(&pb.s32b, h->cap)))) return PF2_E_HIP;
    }
    h->shard_cap = 2 * ((h->cap + QSHARDS - 1) / QSHARDS) + 64;
    const long long qlen = QSHARDS * h->shard_cap + h->cap;
    if (dalloc(&h->dest, h->cap) || dalloc(&h->pkey, h->cap) || dalloc(&h->spos, h->cap) || dalloc(&h->omap, h->cap) ||
        dalloc(&h->ovf, h->cap) || dalloc(&h->queue[0], qlen) || dalloc(&h->queue[1], qlen) ||
        dalloc(&h->park_s, h->cap) || dalloc(&h->park_vext, h->cap * h->dim))
        return PF2_E_HIP;
    return 0;
}

void free_particles(pf2_handle_s* h) {
    if (h->in_block) return;
    for (int b = 0; b < 2; ++b) {
        PBufs& pb = h->buf[b];
        hipFree(pb.cell); hipFree(pb.id); hipFree(pb.x); hipFree(pb.xi); hipFree(pb.v); hipFree(pb.vext);
        hipFree(pb.s32a); hipFree(pb.s32b);
        pb = PBufs();
    }
    hipFree(h->dest); hipFree(h->pkey); hipFree(h->spos); hipFree(h->omap); hipFree(h->ovf);
    hipFree(h->queue[0]); hipFree(h->queue[1]); hipFree(h->park_s); hipFree(h->park_vext);
    h->dest = h->pkey = h->spos = h->omap = h->park_s = nullptr;
    h->ovf = nullptr;
    h->queue[0] = h->queue[1] = nullptr;
    h->park_vext = nullptr;
}

template <class T>
int scan(pf2_handle_s* h, const T* in, T* out, int n) {
    Timed t(h, K_SCAN);
    size_t bytes = h->scan_tmp_bytes;
    HIPC(hipcub::DeviceScan::ExclusiveSum(h->scan_tmp, bytes, in, out, n, h->stream));
    return 0;
}

// the sticky overflow flags (counters[2]) of a synchronised read; after
// either, the handle holds no valid particle state
int flag_error(int flags) {
    if (flags & 1) return fail(PF2_E_CAPACITY, "particle capacity exceeded");
    if (flags & 2) return fail(PF2_E_CAPACITY, "more than 2048 particles stayed in or arrived at one cell");
    return 0;
}

int get_count(pf2_handle_s* h, long long* n) {
    int cnt[3];
    HIPC(hipMemcpyAsync(cnt, h->counters, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    if (int r = flag_error(cnt[2])) return r;
    *n = cnt[0];
    h->n_est = cnt[0];
    return 0;
}

// Regroup after k_move / k_resort / k_keys filled dest and pkey:
// arrivals, census, survivors + reseed into the other buffer, swap.
// fp64_state: regroup the FP64 copy of an FP32-mixed handle's state (the
// re-sort and check_empty_cells run in FP64).
template <int DIM>
int regroup(pf2_handle_s* h, int reseed, bool fuse_sums = false, bool defer_finalize = false,
            bool fp64_state = false, bool pend = false) {
    const int nc = h->n_cells;
    // a regroup consumes a pending layout (its census reads it, its
    // finalize clears the device's record of it)
    if (!pend) h->pending = false;
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    // (a pending regroup's census marks its survivors' destinations)
    w.pend_mark = pend ? 1 : 0;
    const int* old_off = h->off[h->cur_off];
    // rows of GCAP for a lattice of more than 16 parts: the cells that do
    // not fit are listed for a second pass in rows of GCAP_L, and only those
    // that do not fit there go to k_cell_big
    const bool small = census_small_row(pp.nparts_in, pp.max_per_part);
    const bool second = small && pp.nparts_in > 16;
    {
        Timed t(h, K_CELL_COUNT);
        if (small)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_count<DIM, GCAP>), dim3(cell_grid(nc, 8192)), dim3(256), 0,
                               h->stream, nc, old_off, h->buf[h->cur].id, pp, w, reseed);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_count<DIM, GCAP_L>), dim3(cell_grid(nc, 8192)), dim3(256), 0,
                               h->stream, nc, old_off, h->buf[h->cur].id, pp, w, reseed);
        if (second)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_count<DIM, GCAP_L, true>), dim3(256), dim3(256), 0, h->stream,
                               nc, old_off, h->buf[h->cur].id, pp, w, reseed);
    }
    {
        Timed t(h, K_CELL_BIG);
        // 1024 workgroups: a few thousand listed cells take 0.034 ms instead
        // of 0.075 with 256
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_big<DIM>), dim3(1024), dim3(256), 0, h->stream, old_off,
                           h->buf[h->cur].id, pp, w, reseed, second ? 1 : 0);
    }
    if (int r = scan(h, h->fs_cnt, h->fs_off, nc + 1)) return r;
    if (pend) {
        // survivors marked and new particles placed in the current buffer;
        // the next move or a settle regroups them (FP64 handles)
        if (mixed(h) && !fp64_state) {
            Timed t(h, K_CELL_PEND);
            const PBufs& a = h->buf[h->cur];
            if (h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend32<DIM, 2>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, nc, old_off, a.cell, a.id, rec32(a), a.vext, h->dV32, h->cap, pp, w);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend32<DIM>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, nc, old_off, a.cell, a.id, rec32(a), a.vext, h->dV32, h->cap, pp, w);
        } else {
            Timed t(h, K_CELL_PEND);
            DevParticles a = dev_particles(h, h->cur);
            const bool gv = !h->dcellV || !h->packed;
            if (gv && h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend<DIM, true, 2>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, dev_mesh<DIM>(h), old_off, a, pp, w);
            else if (gv)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend<DIM, true>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, dev_mesh<DIM>(h), old_off, a, pp, w);
            else if (h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend<DIM, false, 2>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, dev_mesh<DIM>(h), old_off, a, pp, w);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_pend<DIM, false>), dim3(grid_for(nc, 256)), dim3(256), 0,
                                   h->stream, dev_mesh<DIM>(h), old_off, a, pp, w);
        }
        {
            Timed t(h, K_FINALIZE);
            hipLaunchKernelGGL(k_finalize, dim3(1), dim3(QSHARDS), 0, h->stream, nc, w, h->qcount, 1);
        }
        HIPC(hipGetLastError());
        h->pending = true;
        return 0;
    }
    const bool two_pass = !fuse_sums && h->omap && nc < (1 << 25);
    // the one-pass write of an FP64 handle reads the per-cell copies of V
    if (!two_pass && !mixed(h))
        if (int r = ensure_packed(h)) return r;
    if (two_pass) {
        // per-cell map, then the output positions in order (pf2_kernels.hpp)
        Timed t(h, K_CELL_MAP);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_map<DIM>), dim3(cell_grid(nc)), dim3(256), 0, h->stream, nc,
                           old_off, h->cap, pp, w, h->omap);
    }
    if (mixed(h) && !fp64_state && two_pass) {
        Timed t(h, K_CELL_GATHER32);
        const dim3 g(grid_for(h->n_est, 256 * GU, 65535));   // one pass of GU positions per thread
        const PBufs &a = h->buf[h->cur], &b = h->buf[1 - h->cur];
        if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather32<DIM, 2>), g, dim3(256), 0, h->stream, dev_mesh<DIM>(h),
                               h->dV32, a.id, rec32(a), b.cell, b.id, rec32(b), a.vext, b.vext, h->cap, pp, w, h->omap);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather32<DIM>), g, dim3(256), 0, h->stream, dev_mesh<DIM>(h),
                               h->dV32, a.id, rec32(a), b.cell, b.id, rec32(b), a.vext, b.vext, h->cap, pp, w, h->omap);
    } else if (mixed(h) && !fp64_state) {
        Timed t(h, K_CELL_WRITE32);
        launch_cell_write32<DIM>(h->stream, cell_grid(nc), dev_mesh<DIM>(h), h->dV32, old_off, h->buf[h->cur].id,
                                 rec32(h->buf[h->cur]), h->buf[1 - h->cur].cell, h->buf[1 - h->cur].id,
                                 rec32(h->buf[1 - h->cur]), h->buf[h->cur].vext, h->buf[1 - h->cur].vext, h->cap, pp,
                                 w, fuse_sums, h->partial, h->deg);
    } else if (two_pass) {
        Timed t(h, K_CELL_GATHER);
        const dim3 g(grid_for(h->n_est, 256 * GU, 65535));   // one pass of GU positions per thread
        DevParticles a = dev_particles(h, h->cur), b = dev_particles(h, 1 - h->cur);
        // (an FP32-mixed handle's FP64 regroup, and a regroup after a
        // projection left the per-cell copies stale, gather V through the
        // DoF map: the same values)
        const bool gather_v = !h->dcellV || !h->packed;
        if (gather_v && h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather<DIM, true, 2>), g, dim3(256), 0, h->stream,
                               dev_mesh<DIM>(h), a, b, pp, w, h->omap);
        else if (gather_v)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather<DIM, true>), g, dim3(256), 0, h->stream,
                               dev_mesh<DIM>(h), a, b, pp, w, h->omap);
        else if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather<DIM, false, 2>), g, dim3(256), 0, h->stream,
                               dev_mesh<DIM>(h), a, b, pp, w, h->omap);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_gather<DIM, false>), g, dim3(256), 0, h->stream,
                               dev_mesh<DIM>(h), a, b, pp, w, h->omap);
    } else {
        Timed t(h, K_CELL_WRITE);
        // (an FP32-mixed handle's FP64 regroup gathers V and never fuses)
        if (!h->dcellV && h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, true, false, 2>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
        else if (!h->dcellV)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, true, false>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
        else if (fuse_sums && h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, false, true, 2>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
        else if (fuse_sums)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, false, true>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
        else if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, false, false, 2>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write<DIM, false, false>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), h->dV, old_off, dev_particles(h, h->cur),
                               dev_particles(h, 1 - h->cur), pp, w, h->partial);
    }
    if (!defer_finalize) {
        Timed t(h, K_FINALIZE);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(QSHARDS), 0, h->stream, nc, w, h->qcount, 0);
    }
    HIPC(hipGetLastError());
    h->cur = 1 - h->cur;
    h->cur_off = 1 - h->cur_off;
    return 0;
}

// a pending layout regrouped without reseed: the same particles, in cell
// order (k_settle, then the regroup)
template <int DIM>
int settle_dim(pf2_handle_s* h) {
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    {
        Timed t(h, K_SETTLE);
        const dim3 g(grid_for(h->n_est + 1, 256, 16384));
        if (mixed(h))
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_settle32<DIM>), g, dim3(256), 0, h->stream, h->buf[h->cur].cell,
                               rec32(h->buf[h->cur]), pp, w);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_settle<DIM>), g, dim3(256), 0, h->stream, dev_particles(h, h->cur), pp,
                               w);
    }
    HIPC(hipGetLastError());
    h->pending = false;
    return regroup<DIM>(h, 0);
}

int settle(pf2_handle_s* h) {
    if (!h->pending) return 0;
    return h->dim == 2 ? settle_dim<2>(h) : settle_dim<3>(h);
}

template <int DIM>
int run_move(pf2_handle_s* h, int do_correct, int do_move, bool fuse_sums = false) {
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    DevQueues qu = dev_queues(h);
    // the move: up to 8192 workgroups, about one wavefront-sized chunk per
    // wavefront on the cylinder workload (a smaller grid looping over
    // several chunks was 1% slower); the searches: up to 2048 (FP64) or
    // 4096 (FP32: its rounds take contiguous queue ranges per workgroup,
    // 0.8% faster on the 14.6M-particle moving step, the same on smaller
    // ones)
    const int grid = grid_for(h->n_est, 256, 8192);
    const int qgrid = grid_for(h->n_est / 2 + 1, 256, mixed(h) ? 4096 : 2048);
    if (mixed(h)) {
        h->fp64_valid = false;
        PBufs& b = h->buf[h->cur];
        {
            Timed t(h, K_MOVE32);
            launch_move32<DIM>(h->stream, grid, dev_mesh<DIM>(h), h->dV32, h->dVold32, h->dgeo32, b.cell, rec32(b),
                               b.vext, h->cap, pp, w, qu, do_correct, do_move, h->deg);
        }
        if (do_move)
            for (int r = 0; r <= MOVE_ROUNDS; ++r) {
                Timed t(h, r == 0 ? K_SEARCH32 : (r == 1 ? K_SEARCH32_R1 : K_SEARCH32_R2));
                launch_search_resume32<DIM>(h->stream, round_grid(qgrid, r), dev_mesh<DIM>(h), h->dV32,
                                            h->dgeo32, b.cell, rec32(b), b.vext, h->cap, pp, w, qu, r, h->deg);
            }
    } else {
        {
            Timed t(h, K_MOVE);
            if (h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_move<DIM, 2>), dim3(grid), dim3(256), 0, h->stream,
                                   dev_mesh<DIM>(h), h->dV, h->dVold, dev_particles(h, h->cur), pp, w, qu, do_correct,
                                   do_move);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_move<DIM>), dim3(grid), dim3(256), 0, h->stream,
                                   dev_mesh<DIM>(h), h->dV, h->dVold, dev_particles(h, h->cur), pp, w, qu, do_correct,
                                   do_move);
        }
        if (do_move) {
            // one search round with refilling lanes (queue 0)
            Timed t(h, K_SEARCH);
            if (h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume<DIM, 2>), dim3(qgrid), dim3(256), 0, h->stream,
                                   dev_mesh<DIM>(h), h->dV, dev_particles(h, h->cur), pp, w, qu, 0);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume<DIM>), dim3(qgrid), dim3(256), 0, h->stream,
                                   dev_mesh<DIM>(h), h->dV, dev_particles(h, h->cur), pp, w, qu, 0);
        }
    }
    HIPC(hipGetLastError());
    if (!do_move) return 0;
    // a fused step (projection sums in the cell write) leaves the
    // bookkeeping to the node pass's last workgroup
    return regroup<DIM>(h, 1, fuse_sums, fuse_sums);
}

// FP64 re-sort: own-cell tests (k_resort), then the relocation of the
// particles that failed them in full wavefronts (k_resort_search)
template <int DIM>
int resort64(pf2_handle_s* h, int skip) {
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    DevQueues qu = dev_queues(h);
    {
        Timed t(h, K_RESORT);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort<DIM>), dim3(grid_for(h->n_est, 256, 16384)), dim3(256), 0,
                           h->stream, dev_mesh<DIM>(h), dev_particles(h, h->cur), pp, w, qu, skip);
    }
    {
        Timed t(h, K_RESORT_SEARCH);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort_search<DIM>), dim3(grid_for(h->n_est / 2 + 1, 256, 2048)),
                           dim3(256), 0, h->stream, dev_mesh<DIM>(h), dev_particles(h, h->cur), pp, w, qu);
    }
    HIPC(hipGetLastError());
    return 0;
}

// partial: where the per-cell projection sums go (the handle's own buffer
// unless a caller supplies another of the same size)
template <int DIM>
int run_project(pf2_handle_s* h, bool sums_done = false, double* partial = nullptr) {
    const int nc = h->n_cells;
    DevWork w = dev_work(h);
    if (!partial) partial = h->partial;
    w.partial = partial;
    if (sums_done) {
        // the projection sums came with the regroup (k_cell_write<DIM, false, true>,
        // k_cell_write32<DIM, true>)
    } else if (mixed(h)) {
        Timed t(h, K_PARTIALS32);
        launch_partials32<DIM>(h->stream, cell_grid(nc), nc, h->off[h->cur_off], rec32(h->buf[h->cur]), partial,
                               h->counters, h->deg);
    } else {
        Timed t(h, K_PARTIALS);
        if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_partials<DIM, 2>), dim3(cell_grid(nc)), dim3(256), 0,
                               h->stream, nc, h->off[h->cur_off], dev_particles(h, h->cur), partial, h->counters);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_partials<DIM>), dim3(cell_grid(nc)), dim3(256), 0, h->stream,
                               nc, h->off[h->cur_off], dev_particles(h, h->cur), partial, h->counters);
    }
    {
        Timed t(h, K_NODE);
        if (sums_done) {
            // also the regroup's deferred bookkeeping (k_finalize's work); the
            // regroup has swapped the offset buffers, so its new offsets are
            // now the current ones
            DevWork wf = w;
            wf.new_off = h->off[h->cur_off];
            if (h->deg == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_node<DIM, true, 2>), dim3(grid_for(h->n_dofs, 256) + 1),
                                   dim3(256), 0, h->stream, dev_mesh<DIM>(h), partial, h->dV, h->dVold, h->stats, wf,
                                   h->qcount);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_node<DIM, true>), dim3(grid_for(h->n_dofs, 256) + 1),
                                   dim3(256), 0, h->stream, dev_mesh<DIM>(h), partial, h->dV, h->dVold, h->stats, wf,
                                   h->qcount);
        } else if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_node<DIM, false, 2>), dim3(grid_for(h->n_dofs, 256)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), partial, h->dV, h->dVold, h->stats, w, h->qcount);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_node<DIM, false>), dim3(grid_for(h->n_dofs, 256)), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), partial, h->dV, h->dVold, h->stats, w, h->qcount);
    }
    if (h->n_walls > 0) {
        Timed t(h, K_WALL);
        DevWalls wl;
        wl.n = h->n_walls;
        wl.dof = h->w_dof;
        wl.bc = h->w_bc;
        wl.mask = h->w_mask;
        wl.fix = h->w_fix;
        wl.val = h->w_val;
        wl.vwall = h->w_vwall;
        wl.n_cells = h->n_wall_cells;
        wl.cells = h->w_cells;
        wl.counters = h->counters;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wall_flags<DIM>), dim3(grid_for(h->n_walls, 256)), dim3(256), 0,
                           h->stream, wl, h->dV);
        if (mixed(h)) {
            h->fp64_valid = false;
            launch_wall_particles32<DIM>(h->stream, dev_mesh<DIM>(h), wl, h->off[h->cur_off], rec32(h->buf[h->cur]));
        } else {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wall_particles<DIM>), dim3((h->n_wall_cells + 3) / 4), dim3(256), 0,
                               h->stream, dev_mesh<DIM>(h), wl, h->off[h->cur_off], dev_particles(h, h->cur));
        }
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wall_set<DIM>), dim3(grid_for(h->n_walls, 256)), dim3(256), 0, h->stream,
                           wl, h->dV, h->dVold);
    }
    h->packed = false;   // V and V_old changed: refreshed before their next use
    HIPC(hipGetLastError());
    return 0;
}

int download_field(pf2_handle_s* h, double* V) {
    const int n = h->n_dofs;
    if (h->dim == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_comp_major<2>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream, h->dV,
                           h->dtmp, n);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_comp_major<3>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream, h->dV,
                           h->dtmp, n);
    HIPC(hipGetLastError());
    int flags = 0;
    HIPC(hipMemcpyAsync(V, h->dtmp, sizeof(double) * (size_t)n * h->dim, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(&flags, h->counters + 2, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return flag_error(flags);
}

int upload_field(pf2_handle_s* h, const double* src, double* dst, hipMemcpyKind kind) {
    const int n = h->n_dofs;
    HIPC(hipMemcpyAsync(h->dtmp, src, sizeof(double) * (size_t)n * h->dim, kind, h->stream));
    if (h->dim == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_node_major<2>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream,
                           h->dtmp, dst, n);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_node_major<3>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream,
                           h->dtmp, dst, n);
    HIPC(hipGetLastError());
    // the host source buffer may be reused by the caller after return
    if (kind == hipMemcpyHostToDevice) HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

// per-cell copies of V and V - V_old (FP64 handles: FP64 copies; FP32-mixed
// handles: FP32 copies only); called whenever V or V_old change
int pack_fields(pf2_handle_s* h) {
    h->packed = true;
    const long long n = (long long)h->n_cells * h->nd;
    const int g = grid_for(n, 256, 4096);
    if (h->dim == 2) {
        if (h->dcellV)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pack_fields<2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h), h->dV,
                               h->dVold, h->dcellV, mixed(h) ? nullptr : h->dcellDV);
        if (mixed(h))
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pack_fields32<2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h),
                               h->dV, h->dVold, h->dV32, h->dVold32);
    } else {
        if (h->dcellV)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pack_fields<3>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h), h->dV,
                               h->dVold, h->dcellV, mixed(h) ? nullptr : h->dcellDV);
        if (mixed(h))
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pack_fields32<3>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h),
                               h->dV, h->dVold, h->dV32, h->dVold32);
    }
    HIPC(hipGetLastError());
    return 0;
}

int ensure_packed(pf2_handle_s* h) {
    if (h->packed) return 0;
    Timed t(h, K_MISC);
    return pack_fields(h);
}

// V from a device field (component-major) into the node-major copy and the
// per-cell copies, in one pass
int refresh_field(pf2_handle_s* h, const double* field) {
    const long long n = std::max((long long)h->n_cells * h->nd, (long long)h->n_dofs);
    const int g = grid_for(n, 256, 4096);
    Timed t(h, K_MISC);
    if (h->dim == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_refresh<2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h), field,
                           h->dV, h->dVold, h->dcellV, mixed(h) ? nullptr : h->dcellDV, mixed(h) ? h->dV32 : nullptr, h->dVold32);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_refresh<3>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h), field,
                           h->dV, h->dVold, h->dcellV, mixed(h) ? nullptr : h->dcellDV, mixed(h) ? h->dV32 : nullptr, h->dVold32);
    HIPC(hipGetLastError());
    h->packed = true;
    return 0;
}

// FP64 state <-> FP32-mixed state (downloads and inserts go through FP64)
int to_mixed(pf2_handle_s* h) {
    if (!mixed(h)) return 0;
    h->fp64_valid = true;
    launch_state_to32(h->stream, h->dim, h->buf[h->cur].xi, h->buf[h->cur].v, rec32(h->buf[h->cur]), h->cap,
                      h->counters, h->n_est);
    HIPC(hipGetLastError());
    return 0;
}

int from_mixed(pf2_handle_s* h) {
    if (!mixed(h) || h->fp64_valid) return 0;
    if (h->dim == 2)
        launch_state_from32<2>(h->stream, dev_mesh<2>(h), h->buf[h->cur].cell, rec32(h->buf[h->cur]),
                               h->buf[h->cur].x, h->buf[h->cur].xi, h->buf[h->cur].v, h->cap,
                               h->counters, h->n_est);
    else
        launch_state_from32<3>(h->stream, dev_mesh<3>(h), h->buf[h->cur].cell, rec32(h->buf[h->cur]),
                               h->buf[h->cur].x, h->buf[h->cur].xi, h->buf[h->cur].v, h->cap,
                               h->counters, h->n_est);
    HIPC(hipGetLastError());
    h->fp64_valid = true;
    return 0;
}

// Device bytes pf2_create requests for a mesh of these sizes (the same list
// of allocations as pf2_create and alloc_particles), plus 1 MiB for the scan
// state whose size the scan library fixes at run time.
size_t estimate_bytes(int dim, long long nc, long long nvt, long long nd, long long cap, const pf2_params* p,
                      int deg) {
    const long long nv = 1LL << dim, nck = nc * nv;
    const long long ndc = nc * (deg == 2 ? (dim == 2 ? 9 : 27) : nv);   // velocity DoFs of all cells
    size_t b = 0;
    b += nvt * dim * 8 + nck * 4 + ndc * 4 + (nvt + 1) * 4 + nck * 4 + nck * dim * 8;  // X, cv, cd, star
    b += nc;                                                                             // moved-cell flags
    b += nc * 16 * (dim - 1) + nvt * 5;                                                  // face neighbours, vertex flags
    b += (nd + 1) * 4 + ndc * 4 + nc * 8 + nck * dim * 8 + nck * 8;                    // dof map, ftol, cellX/star
    if (p->precision != PF2_FP32_MIXED) b += ndc * dim * 8 * 2;                          // FP64 cell V / dV copies
    b += nd * dim * 8 * 3;                                                              // V, V_old, staging
    b += (nc + 1) * 4 * 10 + nc * 4 * (1 + arrival_slots(dim, p));                     // per-cell work, arrival slots
    b += nc * 8 + nc * 4 + ndc * (dim + 1) * 8 + 32;                                     // census, partials, counters
    b += (size_t)NQROUNDS * (QSHARDS + 1) * QSTRIDE * 4 + (size_t)NQROUNDS * (QSHARDS + 2) * 4;
    b += 4 + (size_t)STAT_SHARDS * NSTAT * 8 + ((nc + 1 + SCAN_TILE - 1) / SCAN_TILE) * 8;
    b += ndc * dim * 4 * 2 + nc * geo32_floats(dim) * 4;                                 // FP32 fields, geometry
    long long per = 4 + 4 + 3 * dim * 8;
    if (p->keep_velocity_ext) {
        per += dim * 8;
        b += 2 * nvt * dim * 8;    // rest positions and displacement shape of a prescribed motion
    }
    if (p->precision == PF2_FP32_MIXED) per += 2 * dim * 4;
    b += 2 * cap * per;                                                                 // two particle buffers
    const long long shard_cap = 2 * ((cap + QSHARDS - 1) / QSHARDS) + 64;
    const long long qlen = QSHARDS * shard_cap + cap;
    b += cap * 4 * 6 + 2 * qlen * 8 + cap * 4 + cap * dim * 8;                          // work, map, overflow, queues, parking
    return b + (1u << 20);
}

void invalidate_graphs(pf2_handle_s* h) {
    bool any = false;
    for (auto g : h->gexec) any = any || g;
    if (!any) return;
    hipStreamSynchronize(h->stream);
    for (auto& g : h->gexec)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// one step: optional refresh of V from a device field (component-major),
// then correct, move (with reseed) and project -- no host synchronisation
int step_body(pf2_handle_s* h, const double* field) {
    // the copies are refreshed unconditionally, so that a captured graph
    // holds the refresh whatever the state at capture
    if (field) {
        if (int r = refresh_field(h, field)) return r;
    } else {
        Timed t(h, K_MISC);
        if (int r = pack_fields(h)) return r;
    }
    // steps whose per-cell final count is bounded by the LDS rows compute
    // the projection sums while regrouping
    const bool fuse = fuse_bound(h) <= FCAP;
    if (int r = (h->dim == 2 ? run_move<2>(h, 1, 1, fuse) : run_move<3>(h, 1, 1, fuse))) return r;
    return h->dim == 2 ? run_project<2>(h, fuse) : run_project<3>(h, fuse);
}

// the step as a HIP graph: captured once per buffer parity, replayed after.
// The host-side effects of a step (buffer swap, FP32 state flag) are applied
// on every replay exactly as the eager step applies them.
int step_graph(pf2_handle_s* h, const double* field) {
    if (field != h->g_field) {
        for (int k = 4; k < 8; ++k)
            if (h->gexec[k]) {
                hipGraphExecDestroy(h->gexec[k]);
                h->gexec[k] = nullptr;
            }
        h->g_field = field;
    }
    const int key = (field ? 4 : 0) + h->cur * 2 + h->cur_off;
    if (!h->gexec[key]) {
        HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int r = step_body(h, field);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (r) {
            if (g) hipGraphDestroy(g);
            return r;
        }
        HIPC(e);
        const hipError_t ei = hipGraphInstantiate(&h->gexec[key], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPC(ei);
        // the capture already applied the step's host-side effects; undo
        // them so that the replay below applies them once
        h->cur = 1 - h->cur;
        h->cur_off = 1 - h->cur_off;
    }
    HIPC(hipGraphLaunch(h->gexec[key], h->stream));
    // host-side effects of an eager step: buffers swapped, FP32 state ahead
    // of its FP64 copy, per-cell field copies stale after the projection
    h->cur = 1 - h->cur;
    h->cur_off = 1 - h->cur_off;
    if (mixed(h)) h->fp64_valid = false;
    h->packed = false;
    return 0;
}

int run_step(pf2_handle_s* h, const double* field) {
    if (h->graphs && !h->profiling) {
        // (a step graph is captured on a settled layout)
        if (int r = settle(h)) return r;
        return step_graph(h, field);
    }
    return step_body(h, field);
}

// ---------------------------------------------------------------------------
// Batched steps (pf2_batch_*): one launch per kernel of the step for all the
// simulations of a batch.  The table of simulations and their workgroup
// ranges is staged in pinned host memory (a ring of BATCH_STAGES buffers,
// so that a caller may queue that many steps before the host waits) and
// copied to the device before each step.
}  // namespace

constexpr int BATCH_STAGES = 4;

struct pf2_batch_s {
    int max_sims = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    char* dtab = nullptr;
    size_t tab_bytes = 0;
    char* htab[BATCH_STAGES] = {};
    hipEvent_t hev[BATCH_STAGES] = {};         // the staging copy of each buffer has run
    int slot = 0;
    std::vector<hipEvent_t> ev;               // per simulation: its stream's work so far
    hipEvent_t done = nullptr;
};

namespace {

size_t batch_table_bytes(int max_sims) {
    const size_t simb = std::max(sizeof(BatchSim<2>), sizeof(BatchSim<3>));
    return simb * max_sims + (size_t)BK_NKINDS * (max_sims + 1) * sizeof(int);
}

template <int DIM>
void batch_fill(pf2_handle_s* h, const double* field, BatchSim<DIM>& S, int* grid) {
    S.m = dev_mesh<DIM>(h);
    S.pp = dev_params(h);
    S.w = dev_work(h);
    S.qu = dev_queues(h);
    S.A = dev_particles(h, h->cur);
    S.B = dev_particles(h, 1 - h->cur);
    S.RA = rec32(h->buf[h->cur]);
    S.RB = rec32(h->buf[1 - h->cur]);
    S.old_off = h->off[h->cur_off];
    S.field = field;
    S.V = h->dV;
    S.Vold = h->dVold;
    S.cellV = h->dcellV;
    S.cellDV = mixed(h) ? nullptr : h->dcellDV;
    S.V32 = mixed(h) ? h->dV32 : nullptr;
    S.dV32 = h->dVold32;
    S.geo = h->dgeo32;
    S.partial = h->partial;
    S.qcount = h->qcount;
    S.tiles = h->scan_tiles;
    const int nc = h->n_cells;
    S.ntiles = (nc + 1 + SCAN_TILE - 1) / SCAN_TILE;
    // each kernel's grid as the simulation's own step would launch it
    const int qgrid = grid_for(h->n_est / 2 + 1, 256, 2048);
    grid[BK_REFRESH] = grid_for(std::max((long long)nc * h->nd, (long long)h->n_dofs), 256, 4096);
    grid[BK_MOVE] = grid_for(h->n_est, 256, 8192);
    for (int r = 0; r <= MOVE_ROUNDS; ++r) grid[BK_SEARCH0 + r] = round_grid(qgrid, r);
    // (batched steps keep rows of GCAP_L for lattices of more than 16 parts:
    // they have no second census pass)
    const bool small = S.pp.nparts_in <= 16;
    grid[BK_COUNT] = small ? cell_grid(nc, 8192) : 0;
    grid[BK_COUNT_L] = small ? 0 : cell_grid(nc, 8192);
    grid[BK_BIG] = 32;
    grid[BK_TILES] = S.ntiles;
    grid[BK_WRITE] = cell_grid(nc);
    grid[BK_PARTIALS] = cell_grid(nc);
    grid[BK_NODE] = grid_for(h->n_dofs, 256);
}

template <int DIM>
int batch_step(pf2_batch_s* b, int n, const pf2_handle* hs, const double* const* fields) {
    const size_t simb = sizeof(BatchSim<DIM>);
    const int slot = b->slot;
    HIPC(hipEventSynchronize(b->hev[slot]));   // this staging buffer's last copy has run
    char* ht = b->htab[slot];
    BatchSim<DIM>* sims = reinterpret_cast<BatchSim<DIM>*>(ht);
    int* starts = reinterpret_cast<int*>(ht + simb * n);
    int tot[BK_NKINDS] = {};
    for (int i = 0; i < n; ++i) {
        int grid[BK_NKINDS];
        batch_fill<DIM>(hs[i], fields[i], sims[i], grid);
        for (int k = 0; k < BK_NKINDS; ++k) {
            starts[(size_t)k * (n + 1) + i] = tot[k];
            tot[k] += grid[k];
        }
    }
    for (int k = 0; k < BK_NKINDS; ++k) starts[(size_t)k * (n + 1) + n] = tot[k];
    // after every simulation's earlier work
    for (int i = 0; i < n; ++i) {
        HIPC(hipEventRecord(b->ev[i], hs[i]->stream));
        HIPC(hipStreamWaitEvent(b->stream, b->ev[i], 0));
    }
    const size_t bytes = simb * n + (size_t)BK_NKINDS * (n + 1) * sizeof(int);
    HIPC(hipMemcpyAsync(b->dtab, ht, bytes, hipMemcpyHostToDevice, b->stream));
    HIPC(hipEventRecord(b->hev[slot], b->stream));
    b->slot = (slot + 1) % BATCH_STAGES;
    BatchView<DIM> bv{reinterpret_cast<const BatchSim<DIM>*>(b->dtab), reinterpret_cast<const int*>(b->dtab + simb * n),
                      n};
    const bool mx = mixed(hs[0]);
    hipStream_t st = b->stream;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_refresh<DIM>), dim3(tot[BK_REFRESH]), dim3(256), 0, st, bv);
    if (mx) hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_move32<DIM>), dim3(tot[BK_MOVE]), dim3(256), 0, st, bv);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_move<DIM>), dim3(tot[BK_MOVE]), dim3(256), 0, st, bv);
    if (!mx)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_search<DIM>), dim3(tot[BK_SEARCH0]), dim3(256), 0, st, bv, 0);
    else
        for (int r = 0; r <= MOVE_ROUNDS; ++r) {
            const dim3 g(tot[BK_SEARCH0 + r]);
            if (r == MOVE_ROUNDS) hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_search32<DIM, true>), g, dim3(256), 0, st, bv, r);
            else hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_search32<DIM, false>), g, dim3(256), 0, st, bv, r);
        }
    if (tot[BK_COUNT])
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_count<DIM, GCAP>), dim3(tot[BK_COUNT]), dim3(256), 0, st, bv);
    if (tot[BK_COUNT_L])
        hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_count<DIM, GCAP_L>), dim3(tot[BK_COUNT_L]), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_big<DIM>), dim3(tot[BK_BIG]), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_scan_tiles<DIM>), dim3(tot[BK_TILES]), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_scan_tops<DIM>), dim3(n), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_scan_apply<DIM>), dim3(tot[BK_TILES]), dim3(256), 0, st, bv);
    if (mx) hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_write32<DIM>), dim3(tot[BK_WRITE]), dim3(256), 0, st, bv);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_write<DIM>), dim3(tot[BK_WRITE]), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_finalize<DIM>), dim3(n), dim3(QSHARDS), 0, st, bv);
    if (mx) hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_partials32<DIM>), dim3(tot[BK_PARTIALS]), dim3(256), 0, st, bv);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_partials<DIM>), dim3(tot[BK_PARTIALS]), dim3(256), 0, st, bv);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(kb_node<DIM>), dim3(tot[BK_NODE]), dim3(256), 0, st, bv);
    HIPC(hipGetLastError());
    // every simulation's later work after the batch
    HIPC(hipEventRecord(b->done, st));
    for (int i = 0; i < n; ++i) HIPC(hipStreamWaitEvent(hs[i]->stream, b->done, 0));
    // the host-side effects of each simulation's own step
    for (int i = 0; i < n; ++i) {
        pf2_handle_s* h = hs[i];
        h->cur = 1 - h->cur;
        h->cur_off = 1 - h->cur_off;
        if (mixed(h)) h->fp64_valid = false;
        h->packed = false;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Projection sums of pf2_distribute_particle_velocities_to_grid_static: a
// file-scope device array instead of the handle's buffer, for callers that
// may not allocate.  Every GPU holds its own copy at its own address; the
// host learns it once per device with hipGetSymbolAddress and passes that
// address to the kernels (the array's name is not a device address in host
// code).  Calls on one GPU share the array, so such calls must not overlap.
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)PF2_STATIC_SCRATCH_BYTES;
__device__ __align__(256) double scratch[STATIC_SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on the current device (looked up once)
int scratch_address(void** out) {
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(PF2_E_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PF2_E_HIP, "hipGetSymbolAddress returned no address");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// field (node-major) -> caller device memory, component-major
int field_to_device(pf2_handle_s* h, double* V_out) {
    const int n = h->n_dofs;
    if (h->dim == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_comp_major<2>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream, h->dV,
                           V_out, n);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_comp_major<3>), dim3(grid_for(n, 256)), dim3(256), 0, h->stream, h->dV,
                           V_out, n);
    HIPC(hipGetLastError());
    return 0;
}

}  // namespace

// ============================================================================
extern "C" {

const char* pf2_last_error(void) { return g_err.c_str(); }

int pf2_create(const pf2_mesh* mesh, const pf2_params* params, long long capacity, pf2_handle* out) {
    return pf2_create_in(mesh, params, capacity, nullptr, 0, nullptr, out);
}

int pf2_create_in(const pf2_mesh* mesh, const pf2_params* params, long long capacity, void* block,
                  unsigned long long block_bytes, void* stream, pf2_handle* out) {
    struct ArenaScope {
        ArenaScope(void* b, size_t n) { g_arena = Arena{(char*)b, b ? n : 0, 0}; }
        ~ArenaScope() { g_arena = Arena{}; }
    } arena_scope(block, block_bytes);
    if (!mesh || !out) return fail(PF2_E_ARG, "mesh/out is NULL");
    *out = nullptr;
    if (mesh->dim != 2 && mesh->dim != 3) return fail(PF2_E_ARG, "dim must be 2 or 3");
    if (mesh->n_cells < 1 || mesh->n_vertices < 1 || mesh->n_dofs < 1 || !mesh->vertices || !mesh->cell_vertices)
        return fail(PF2_E_ARG, "empty mesh");
    if (int r = check_params(mesh->dim, params)) return r;
    if (mesh->degree != 0 && mesh->degree != 1 && mesh->degree != 2) return fail(PF2_E_ARG, "degree must be 1 or 2");
    if (mesh->degree == 2 && !mesh->cell_dofs) return fail(PF2_E_ARG, "a degree-2 field needs cell_dofs");
    pf2_handle_s* h = new pf2_handle_s();
    h->dim = mesh->dim;
    h->nv = 1 << mesh->dim;
    h->deg = mesh->degree == 2 ? 2 : 1;
    h->nd = h->deg == 2 ? (mesh->dim == 2 ? 9 : 27) : h->nv;
    h->n_cells = mesh->n_cells;
    h->n_vertices = mesh->n_vertices;
    h->n_dofs = mesh->n_dofs;
    h->params = *params;
    h->cap = capacity > 0 ? capacity : default_capacity(h->dim, h->n_cells, params);
    h->aslots = arrival_slots(h->dim, params);
    if (h->cap > 0x7fffffffLL) {
        delete h;
        return fail(PF2_E_ARG, "capacity above 2^31-1 particles");
    }
    const size_t nck = (size_t)h->n_cells * h->nv;
    const size_t ndc = (size_t)h->n_cells * h->nd;   // velocity DoF entries of all cells
    h->X.assign(mesh->vertices, mesh->vertices + (size_t)h->n_vertices * h->dim);
    h->cv.assign(mesh->cell_vertices, mesh->cell_vertices + nck);
    h->cd.assign(mesh->cell_dofs ? mesh->cell_dofs : mesh->cell_vertices,
                 (mesh->cell_dofs ? mesh->cell_dofs : mesh->cell_vertices) + ndc);
    for (size_t i = 0; i < nck; ++i) {
        if (h->cv[i] < 0 || h->cv[i] >= h->n_vertices) {
            delete h;
            return fail(PF2_E_ARG, "cell vertex index out of range");
        }
    }
    for (size_t i = 0; i < ndc; ++i) {
        if (h->cd[i] < 0 || h->cd[i] >= h->n_dofs) {
            delete h;
            return fail(PF2_E_ARG, "cell dof index out of range");
        }
    }
    // vertex star (GridTools::vertex_to_cell_map: std::set => ascending cells)
    h->star_off.assign(h->n_vertices + 1, 0);
    for (size_t i = 0; i < nck; ++i) h->star_off[h->cv[i] + 1]++;
    for (int v = 0; v < h->n_vertices; ++v) h->star_off[v + 1] += h->star_off[v];
    h->star_cell.assign(h->star_off[h->n_vertices], 0);
    {
        std::vector<int> fill(h->star_off.begin(), h->star_off.end() - 1);
        for (int c = 0; c < h->n_cells; ++c)
            for (int k = 0; k < h->nv; ++k) h->star_cell[fill[h->cv[(size_t)c * h->nv + k]]++] = c;
    }
    compute_star_dirs(h);
    h->cell_star.resize(2 * nck);
    for (size_t i = 0; i < nck; ++i) {
        h->cell_star[2 * i] = h->star_off[h->cv[i]];
        h->cell_star[2 * i + 1] = h->star_off[h->cv[i] + 1];
    }
    {
        // neighbour across each face (face 2a + s: the local vertices whose
        // bit a is s): the other cell of the face's first vertex's star that
        // holds all the face's vertices; -1 on the boundary
        const int nv = h->nv, nf = 2 * h->dim, stride = h->dim == 2 ? 4 : 8;
        h->cell_nbr.assign((size_t)stride * h->n_cells, -1);
        for (int c = 0; c < h->n_cells; ++c)
            for (int f = 0; f < nf; ++f) {
                int fvtx[4], nfv = 0;
                for (int k = 0; k < nv; ++k)
                    if (((k >> (f >> 1)) & 1) == (f & 1)) fvtx[nfv++] = h->cv[(size_t)c * nv + k];
                int& out = h->cell_nbr[(size_t)c * stride + f];
                for (int e = h->star_off[fvtx[0]]; e < h->star_off[fvtx[0] + 1] && out < 0; ++e) {
                    const int s = h->star_cell[e];
                    if (s == c) continue;
                    int hits = 0;
                    for (int q = 1; q < nfv; ++q)
                        for (int k = 0; k < nv; ++k)
                            if (h->cv[(size_t)s * nv + k] == fvtx[q]) {
                                ++hits;
                                break;
                            }
                    if (hits == nfv - 1) out = s;
                }
            }
    }
    // per vertex: on a face with no neighbour (face->at_boundary()), and its DoF
    std::vector<unsigned char> vbnd(h->n_vertices, 0);
    std::vector<int> vdof(h->n_vertices, -1);
    {
        const int nv = h->nv, nf = 2 * h->dim, stride = h->dim == 2 ? 4 : 8;
        for (int c = 0; c < h->n_cells; ++c)
            for (int f = 0; f < nf; ++f)
                if (h->cell_nbr[(size_t)c * stride + f] < 0)
                    for (int k = 0; k < nv; ++k)
                        if (((k >> (f >> 1)) & 1) == (f & 1)) vbnd[h->cv[(size_t)c * nv + k]] = 1;
        for (int v = 0; v < h->n_vertices; ++v) {
            if (h->star_off[v] == h->star_off[v + 1]) continue;
            const int c0 = h->star_cell[h->star_off[v]];
            for (int k = 0; k < nv; ++k)
                if (h->cv[(size_t)c0 * nv + k] == v) vdof[v] = h->cd[(size_t)c0 * h->nd + k];
        }
    }
    // DoF -> (cell, local DoF), ascending cells
    h->dof_off.assign(h->n_dofs + 1, 0);
    for (size_t i = 0; i < ndc; ++i) h->dof_off[h->cd[i] + 1]++;
    for (int v = 0; v < h->n_dofs; ++v) h->dof_off[v + 1] += h->dof_off[v];
    h->dof_cell.assign(h->dof_off[h->n_dofs], 0);
    {
        std::vector<int> fill(h->dof_off.begin(), h->dof_off.end() - 1);
        for (int c = 0; c < h->n_cells; ++c)
            for (int k = 0; k < h->nd; ++k) h->dof_cell[fill[h->cd[(size_t)c * h->nd + k]]++] = c * h->nd + k;
    }
    auto bail = [&](int r) {
        pf2_destroy(h);
        return r;
    };
    const size_t bytes0 = g_alloc_bytes;
    h->in_block = block != nullptr;
    if (hipGetDevice(&h->device) != hipSuccess) return bail(fail(PF2_E_HIP, "hipGetDevice failed"));
    if (stream) {
        h->stream = (hipStream_t)stream;
        h->own_stream = false;
    } else if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(PF2_E_HIP, "hipStreamCreate failed"));
    const int nc = h->n_cells, nvt = h->n_vertices, nd = h->n_dofs, dim = h->dim;
    if (dalloc(&h->dX, (size_t)nvt * dim) || dalloc(&h->dcv, nck) || dalloc(&h->dcd, ndc) ||
        dalloc(&h->dstar_off, nvt + 1) || dalloc(&h->dstar_cell, h->star_cell.size()) ||
        dalloc(&h->dstar_dir, h->star_dir.size()) || dalloc(&h->ddof_off, nd + 1) ||
        dalloc(&h->ddof_cell, h->dof_cell.size()) || dalloc(&h->dftol, nc) || dalloc(&h->dcellX, nck * dim) || dalloc(&h->dcell_star, nck) ||
        dalloc(&h->dcell_nbr, (size_t)h->n_cells * (h->dim - 1)) || dalloc(&h->dvbnd, nvt) || dalloc(&h->dcmoved, nc) ||
        dalloc(&h->dvdof, nvt) ||
        (!mixed(h) && (dalloc(&h->dcellV, ndc * dim) || dalloc(&h->dcellDV, ndc * dim))) ||
        dalloc(&h->dV, (size_t)nd * dim) ||
        dalloc(&h->dVold, (size_t)nd * dim) || dalloc(&h->dtmp, (size_t)nd * dim) || dalloc(&h->off[0], nc + 1) ||
        dalloc(&h->off[1], nc + 1) || dalloc(&h->arr_cnt, nc + 1) ||
        dalloc(&h->arr_slot, (size_t)nc * h->aslots) || dalloc(&h->ovf_head, nc) || dalloc(&h->fs_cnt, nc + 1) || dalloc(&h->fs_off, nc + 1) || dalloc(&h->hitmask, nc) || dalloc(&h->nsurv, nc) ||
        dalloc(&h->partial, ndc * (dim + 1)) || dalloc(&h->counters, 8) || dalloc(&h->big, nc + 1) || dalloc(&h->big2, nc + 1) ||
        dalloc(&h->qcount, (size_t)NQROUNDS * (QSHARDS + 1) * QSTRIDE) ||
        dalloc(&h->qprefix, (size_t)NQROUNDS * (QSHARDS + 2)) ||
        dalloc(&h->pcounter, 1) || dalloc(&h->stats, STAT_SHARDS * NSTAT) ||
        dalloc(&h->scan_tiles, (size_t)(nc + 1 + SCAN_TILE - 1) / SCAN_TILE))
        return bail(PF2_E_HIP);
    if (mixed(h) || true) {
        if (dalloc(&h->dV32, ndc * dim) || dalloc(&h->dVold32, ndc * dim) ||
            dalloc(&h->dgeo32, (size_t)nc * geo32_floats(dim)))
            return bail(PF2_E_HIP);
    }
    if (alloc_particles(h)) return bail(PF2_E_HIP);
    if (h->params.keep_velocity_ext &&
        (dalloc(&h->dX0, (size_t)nvt * dim) || dalloc(&h->dW, (size_t)nvt * dim)))
        return bail(PF2_E_HIP);
    auto up = [&](void* d, const void* s, size_t bytes) {
        return hipMemcpyAsync(d, s, bytes, hipMemcpyHostToDevice, h->stream);
    };
    if (up(h->dX, h->X.data(), h->X.size() * 8) || up(h->dcv, h->cv.data(), nck * 4) ||
        up(h->dcd, h->cd.data(), ndc * 4) || up(h->dstar_off, h->star_off.data(), h->star_off.size() * 4) ||
        up(h->dstar_cell, h->star_cell.data(), h->star_cell.size() * 4) ||
        up(h->dstar_dir, h->star_dir.data(), h->star_dir.size() * 8) ||
        up(h->ddof_off, h->dof_off.data(), h->dof_off.size() * 4) ||
        up(h->ddof_cell, h->dof_cell.data(), h->dof_cell.size() * 4) ||
        up(h->dcell_star, h->cell_star.data(), h->cell_star.size() * 4) ||
        up(h->dcell_nbr, h->cell_nbr.data(), h->cell_nbr.size() * 4) || up(h->dvbnd, vbnd.data(), nvt) ||
        up(h->dvdof, vdof.data(), (size_t)nvt * 4))
        return bail(fail(PF2_E_HIP, "mesh upload failed"));
    if (hipMemsetAsync(h->dV, 0, sizeof(double) * nd * dim, h->stream) ||
        hipMemsetAsync(h->dVold, 0, sizeof(double) * nd * dim, h->stream) ||
        hipMemsetAsync(h->off[0], 0, sizeof(int) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->off[1], 0, sizeof(int) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->arr_cnt, 0, sizeof(int) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->ovf_head, 0xff, sizeof(int) * nc, h->stream) ||
        hipMemsetAsync(h->fs_cnt, 0, sizeof(unsigned long long) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->counters, 0, sizeof(int) * 8, h->stream) ||
        hipMemsetAsync(h->big, 0, sizeof(int) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->big2, 0, sizeof(int) * (nc + 1), h->stream) ||
        hipMemsetAsync(h->qcount, 0, sizeof(int) * NQROUNDS * (QSHARDS + 1) * QSTRIDE, h->stream) ||
        hipMemsetAsync(h->pcounter, 0, sizeof(uint32_t), h->stream) ||
        hipMemsetAsync(h->stats, 0, sizeof(unsigned long long) * STAT_SHARDS * NSTAT, h->stream))
        return bail(fail(PF2_E_HIP, "initialisation failed"));
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, h->scan_tmp_bytes, h->fs_cnt, h->fs_off, nc + 1, h->stream) !=
        hipSuccess)
        return bail(fail(PF2_E_HIP, "scan sizing failed"));
    if (dalloc((char**)&h->scan_tmp, h->scan_tmp_bytes)) return bail(PF2_E_HIP);
    if (h->dim == 2) launch_geo<2>(h->stream, dev_mesh<2>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    else launch_geo<3>(h->stream, dev_mesh<3>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    if (pack_fields(h)) return bail(PF2_E_HIP);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(PF2_E_HIP, "initialisation sync failed"));
    h->bytes = g_alloc_bytes - bytes0;
    *out = h;
    return 0;
}

int pf2_destroy(pf2_handle h) {
    if (!h) return 0;
    if (h->stream) hipStreamSynchronize(h->stream);
    invalidate_graphs(h);
    free_particles(h);
    void* ptrs[] = {h->dX, h->dcv, h->dcd, h->dstar_off, h->dstar_cell, h->dstar_dir, h->ddof_off, h->ddof_cell,
                    h->dftol, h->dcellX, h->dcell_star, h->dcell_nbr, h->dvbnd, h->dvdof, h->dcellV, h->dcellDV, h->dV, h->dVold, h->dtmp, h->dV32,
                    h->dX0, h->dW, h->dcmoved,
                    h->dVold32, h->dgeo32, h->off[0], h->off[1], h->arr_cnt, h->arr_slot, h->ovf_head, h->fs_cnt,
                    h->fs_off, h->hitmask, h->nsurv, h->partial, h->counters, h->big, h->big2, h->qcount, h->qprefix,
                    h->pcounter, h->stats, h->scan_tmp, h->scan_tiles};
    if (!h->in_block)
        for (void* p : ptrs)
            if (p) hipFree(p);
    // wall lists are allocated after creation, always with hipMalloc
    void* wptrs[] = {h->w_dof, h->w_mask, h->w_fix, h->w_vwall, h->w_cells, h->w_bc, h->w_val};
    for (void* p : wptrs)
        if (p) hipFree(p);
    for (auto& e : h->ev_free) {
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    for (auto& e : h->ev_pending) {
        hipEventDestroy(e.second.first);
        hipEventDestroy(e.second.second);
    }
    if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int pf2_set_params(pf2_handle h, const pf2_params* params) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    invalidate_graphs(h);
    if (int r = check_params(h->dim, params)) return r;
    if (params->precision != h->params.precision)
        return fail(PF2_E_STATE, "precision is fixed at pf2_create");
    if (params->keep_velocity_ext != h->params.keep_velocity_ext)
        return fail(PF2_E_STATE, "keep_velocity_ext is fixed at pf2_create");
    if (default_capacity(h->dim, h->n_cells, params) > h->cap)
        return fail(PF2_E_CAPACITY, "new lattice / MAX exceed the particle capacity");
    h->params = *params;
    return 0;
}

int pf2_upload_fields(pf2_handle h, const double* V, const double* V_old) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (V)
        if (int r = upload_field(h, V, h->dV, hipMemcpyHostToDevice)) return r;
    if (V_old)
        if (int r = upload_field(h, V_old, h->dVold, hipMemcpyHostToDevice)) return r;
    return pack_fields(h);
}

int pf2_upload_fields_device(pf2_handle h, const double* V, const double* V_old) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (V)
        if (int r = upload_field(h, V, h->dV, hipMemcpyDeviceToDevice)) return r;
    if (V_old)
        if (int r = upload_field(h, V_old, h->dVold, hipMemcpyDeviceToDevice)) return r;
    return pack_fields(h);
}

int pf2_download_field(pf2_handle h, double* V) {
    if (!h || !V) return fail(PF2_E_ARG, "NULL argument");
    return download_field(h, V);
}

int pf2_seed_particles(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    invalidate_graphs(h);
    if (int r = ensure_packed(h)) return r;
    long long n = 0;
    if (int r = get_count(h, &n)) return r;
    if (n != 0) return fail(PF2_E_STATE, "seed_particles on a handle that already holds particles");
    DevParams pp = dev_params(h);
    const long long total = (long long)h->n_cells * pp.nparts_in;
    if (total > h->cap) return fail(PF2_E_CAPACITY, "seeding exceeds the particle capacity");
    uint32_t base = 0;
    HIPC(hipMemcpyAsync(&base, h->pcounter, 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    {
        Timed t(h, K_SEED);
        const int g = grid_for(total, 256);
        if (h->dim == 2 && h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_seed<2, 2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h), h->dV,
                               dev_particles(h, h->cur), pp, base, h->off[h->cur_off]);
        else if (h->dim == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_seed<2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h), h->dV,
                               dev_particles(h, h->cur), pp, base, h->off[h->cur_off]);
        else if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_seed<3, 2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h), h->dV,
                               dev_particles(h, h->cur), pp, base, h->off[h->cur_off]);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_seed<3>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h), h->dV,
                               dev_particles(h, h->cur), pp, base, h->off[h->cur_off]);
    }
    HIPC(hipGetLastError());
    const int cnt = (int)total;
    const uint32_t nb = base + (uint32_t)total;
    HIPC(hipMemcpyAsync(h->counters, &cnt, 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->pcounter, &nb, 4, hipMemcpyHostToDevice, h->stream));
    h->n_est = total;
    if (int r = to_mixed(h)) return r;
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int pf2_download_particles(pf2_handle h, int* cell, uint32_t* id, double* x, double* xi, double* v) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    long long n = 0;
    if (int r = get_count(h, &n)) return r;
    if (int r = from_mixed(h)) return r;
    const PBufs& b = h->buf[h->cur];
    const int dim = h->dim;
    if (cell) HIPC(hipMemcpyAsync(cell, b.cell, n * 4, hipMemcpyDeviceToHost, h->stream));
    if (id) HIPC(hipMemcpyAsync(id, b.id, n * 4, hipMemcpyDeviceToHost, h->stream));
    std::vector<double> tmp((size_t)n * dim);
    double* srcs[3] = {b.x, b.xi, b.v};
    double* dsts[3] = {x, xi, v};
    for (int a = 0; a < 3; ++a) {
        if (!dsts[a]) continue;
        for (int d = 0; d < dim; ++d)
            HIPC(hipMemcpyAsync(tmp.data() + (size_t)d * n, srcs[a] + (size_t)d * h->cap, n * 8,
                                hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        for (long long i = 0; i < n; ++i)
            for (int d = 0; d < dim; ++d) dsts[a][i * dim + d] = tmp[(size_t)d * n + i];
    }
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int pf2_insert_particles(pf2_handle h, int n, const int* cell, const double* xi, const double* v) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    invalidate_graphs(h);
    if (int r = ensure_packed(h)) return r;
    if (n < 0 || (n > 0 && (!cell || !xi))) return fail(PF2_E_ARG, "invalid particle arrays");
    const int dim = h->dim;
    for (int i = 0; i < n; ++i) {
        if (cell[i] < 0 || cell[i] >= h->n_cells) return fail(PF2_E_ARG, "cell index out of range");
        for (int d = 0; d < dim; ++d)
            if (!(xi[(size_t)i * dim + d] >= 0.0 && xi[(size_t)i * dim + d] <= 1.0))
                return fail(PF2_E_ARG, "reference coordinates must lie in [0,1]");
    }
    long long n0 = 0;
    if (int r = get_count(h, &n0)) return r;
    if (n0 + n > h->cap) return fail(PF2_E_CAPACITY, "insert exceeds the particle capacity");
    std::vector<int> c0(n0);
    std::vector<uint32_t> id0(n0);
    std::vector<double> x0((size_t)n0 * dim), xi0((size_t)n0 * dim), v0((size_t)n0 * dim);
    if (n0 > 0)
        if (int r = pf2_download_particles(h, c0.data(), id0.data(), x0.data(), xi0.data(), v0.data())) return r;
    uint32_t base = 0;
    HIPC(hipMemcpyAsync(&base, h->pcounter, 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    const long long N = n0 + n;
    struct Rec { int cell; uint32_t id; long long src; };
    std::vector<Rec> recs(N);
    for (long long i = 0; i < n0; ++i) recs[i] = {c0[i], id0[i], i};
    for (int i = 0; i < n; ++i) recs[n0 + i] = {cell[i], base + (uint32_t)i + 1u, n0 + i};
    std::stable_sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
        return a.cell != b.cell ? a.cell < b.cell : a.id < b.id;
    });
    std::vector<int> hc(N), offh(h->n_cells + 1, 0);
    std::vector<uint32_t> hid(N);
    std::vector<unsigned char> flags(N);
    std::vector<double> hx((size_t)dim * N), hxi((size_t)dim * N), hv((size_t)dim * N);
    for (long long i = 0; i < N; ++i) {
        const Rec& r = recs[i];
        hc[i] = r.cell;
        hid[i] = r.id;
        offh[r.cell + 1]++;
        const bool is_new = r.src >= n0;
        flags[i] = is_new ? (v ? 1 : 3) : 0;   // bit0: map location, bit1: interpolate velocity
        for (int d = 0; d < dim; ++d) {
            if (!is_new) {
                hx[(size_t)d * N + i] = x0[r.src * dim + d];
                hxi[(size_t)d * N + i] = xi0[r.src * dim + d];
                hv[(size_t)d * N + i] = v0[r.src * dim + d];
            } else {
                const long long s = r.src - n0;
                hx[(size_t)d * N + i] = 0.0;
                hxi[(size_t)d * N + i] = xi[s * dim + d];
                hv[(size_t)d * N + i] = v ? v[s * dim + d] : 0.0;
            }
        }
    }
    for (int c = 0; c < h->n_cells; ++c) offh[c + 1] += offh[c];
    const PBufs& b = h->buf[h->cur];
    unsigned char* dflags = nullptr;
    if (dalloc(&dflags, N)) return PF2_E_HIP;
    HIPC(hipMemcpyAsync(b.cell, hc.data(), N * 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(b.id, hid.data(), N * 4, hipMemcpyHostToDevice, h->stream));
    for (int d = 0; d < dim; ++d) {
        HIPC(hipMemcpyAsync(b.x + (size_t)d * h->cap, hx.data() + (size_t)d * N, N * 8, hipMemcpyHostToDevice, h->stream));
        HIPC(hipMemcpyAsync(b.xi + (size_t)d * h->cap, hxi.data() + (size_t)d * N, N * 8, hipMemcpyHostToDevice, h->stream));
        HIPC(hipMemcpyAsync(b.v + (size_t)d * h->cap, hv.data() + (size_t)d * N, N * 8, hipMemcpyHostToDevice, h->stream));
    }
    HIPC(hipMemcpyAsync(dflags, flags.data(), N, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->off[h->cur_off], offh.data(), (h->n_cells + 1) * 4, hipMemcpyHostToDevice, h->stream));
    {
        Timed t(h, K_SEED);
        const int g = grid_for(N, 256);
        const int kv = h->params.keep_velocity_ext;
        if (dim == 2 && h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_init_particles<2, 2>), dim3(g), dim3(256), 0, h->stream,
                               dev_mesh<2>(h), h->dV, dev_particles(h, h->cur), dflags, N, kv);
        else if (dim == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_init_particles<2>), dim3(g), dim3(256), 0, h->stream,
                               dev_mesh<2>(h), h->dV, dev_particles(h, h->cur), dflags, N, kv);
        else if (h->deg == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_init_particles<3, 2>), dim3(g), dim3(256), 0, h->stream,
                               dev_mesh<3>(h), h->dV, dev_particles(h, h->cur), dflags, N, kv);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_init_particles<3>), dim3(g), dim3(256), 0, h->stream,
                               dev_mesh<3>(h), h->dV, dev_particles(h, h->cur), dflags, N, kv);
    }
    HIPC(hipGetLastError());
    const int cnt = (int)N;
    const uint32_t nb = base + (uint32_t)n;
    HIPC(hipMemcpyAsync(h->counters, &cnt, 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->pcounter, &nb, 4, hipMemcpyHostToDevice, h->stream));
    h->n_est = N;
    if (int r = to_mixed(h)) return r;
    HIPC(hipStreamSynchronize(h->stream));
    hipFree(dflags);
    return 0;
}

int pf2_correct_particles_velocities(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = ensure_packed(h)) return r;
    return h->dim == 2 ? run_move<2>(h, 1, 0) : run_move<3>(h, 1, 0);
}

int pf2_move_particles(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = ensure_packed(h)) return r;
    return h->dim == 2 ? run_move<2>(h, 0, 1) : run_move<3>(h, 0, 1);
}

int pf2_distribute_particle_velocities_to_grid(pf2_handle h, double* V_out) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    if (int r = (h->dim == 2 ? run_project<2>(h) : run_project<3>(h))) return r;
    if (V_out) return download_field(h, V_out);
    return 0;
}

int pf2_distribute_particle_velocities_to_grid_static(pf2_handle h, double* V_out_device) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    const size_t need = (size_t)h->n_cells * h->nd * (h->dim + 1) * sizeof(double);
    if (need > STATIC_SCRATCH_BYTES)
        return fail(PF2_E_CAPACITY, "projection sums of " + std::to_string(need) + " bytes exceed the static scratch of " +
                                        std::to_string(STATIC_SCRATCH_BYTES) + " bytes");
    int cur = 0;
    HIPC(hipGetDevice(&cur));
    if (cur != h->device) return fail(PF2_E_STATE, "the handle's device is not the current device");
    void* sp = nullptr;
    if (int r = scratch_address(&sp)) return r;
    // no per-kernel timing here: its events would be created on demand
    const bool prof = h->profiling;
    h->profiling = false;
    // (a pending layout is settled with the handle's own buffers)
    int r = settle(h);
    if (!r) r = h->dim == 2 ? run_project<2>(h, false, (double*)sp) : run_project<3>(h, false, (double*)sp);
    if (!r && V_out_device) r = field_to_device(h, V_out_device);
    h->profiling = prof;
    return r;
}

int pf2_batch_create(int max_sims, void* stream, pf2_batch* out) {
    if (!out || max_sims < 1) return fail(PF2_E_ARG, "invalid batch size");
    *out = nullptr;
    pf2_batch_s* b = new pf2_batch_s();
    b->max_sims = max_sims;
    b->tab_bytes = batch_table_bytes(max_sims);
    auto bail = [&](const char* msg) {
        pf2_batch_destroy(b);
        return fail(PF2_E_HIP, msg);
    };
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail("stream creation failed");
        b->own_stream = true;
    }
    if (hipMalloc((void**)&b->dtab, b->tab_bytes) != hipSuccess) return bail("batch table allocation failed");
    for (auto& t : b->htab)
        if (hipHostMalloc((void**)&t, b->tab_bytes, hipHostMallocDefault) != hipSuccess)
            return bail("batch table allocation failed");
    b->ev.assign(max_sims, nullptr);
    for (auto& e : b->ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return bail("event creation failed");
    for (auto& e : b->hev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess || hipEventRecord(e, b->stream) != hipSuccess)
            return bail("event creation failed");
    if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) return bail("event creation failed");
    *out = b;
    return 0;
}

int pf2_batch_destroy(pf2_batch b) {
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    hipFree(b->dtab);
    for (auto t : b->htab)
        if (t) hipHostFree(t);
    for (auto e : b->ev)
        if (e) hipEventDestroy(e);
    for (auto e : b->hev)
        if (e) hipEventDestroy(e);
    if (b->done) hipEventDestroy(b->done);
    if (b->own_stream && b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int pf2_batch_step(pf2_batch b, int n, const pf2_handle* hs, const double* const* fields) {
    if (!b || !hs || !fields || n < 1) return fail(PF2_E_ARG, "NULL argument");
    if (n > b->max_sims) return fail(PF2_E_ARG, "more simulations than the batch holds");
    for (int i = 0; i < n; ++i)
        if (hs[i])
            if (int r = settle(hs[i])) return r;
    for (int i = 0; i < n; ++i) {
        if (!hs[i] || !fields[i]) return fail(PF2_E_ARG, "NULL handle or field");
        if (hs[i]->dim != hs[0]->dim || mixed(hs[i]) != mixed(hs[0]))
            return fail(PF2_E_ARG, "a batch takes simulations of one dimension and precision");
        if (hs[i]->deg != 1 || hs[i]->n_walls > 0)
            return fail(PF2_E_STATE, "a batch takes degree-1 fields without walls");
        for (int j = 0; j < i; ++j)
            if (hs[j] == hs[i]) return fail(PF2_E_ARG, "a simulation appears twice in the batch");
    }
    return hs[0]->dim == 2 ? batch_step<2>(b, n, hs, fields) : batch_step<3>(b, n, hs, fields);
}

int pf2_step(pf2_handle h, double* V_out) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = run_step(h, nullptr)) return r;
    if (V_out) return download_field(h, V_out);
    return 0;
}

int pf2_step_with_field(pf2_handle h, const double* V_device) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (!V_device) return fail(PF2_E_ARG, "V_device is NULL");
    return run_step(h, V_device);
}

int pf2_set_graphs(pf2_handle h, int on) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    HIPC(hipStreamSynchronize(h->stream));
    invalidate_graphs(h);
    h->graphs = on != 0;
    return 0;
}

int pf2_memory_bytes(pf2_handle h, unsigned long long* bytes) {
    if (!h || !bytes) return fail(PF2_E_ARG, "handle/bytes is NULL");
    *bytes = h->bytes;
    return 0;
}

int pf2_estimate_bytes(const pf2_mesh* mesh, const pf2_params* params, long long capacity,
                       unsigned long long* bytes) {
    if (!mesh || !bytes) return fail(PF2_E_ARG, "mesh/bytes is NULL");
    if (mesh->dim != 2 && mesh->dim != 3) return fail(PF2_E_ARG, "dim must be 2 or 3");
    if (int r = check_params(mesh->dim, params)) return r;
    const long long cap = capacity > 0 ? capacity : default_capacity(mesh->dim, mesh->n_cells, params);
    *bytes = estimate_bytes(mesh->dim, mesh->n_cells, mesh->n_vertices, mesh->n_dofs, cap, params,
                            mesh->degree == 2 ? 2 : 1);
    return 0;
}

int pf2_set_walls(pf2_handle h, int n, const int* vertex, const int* dof, const double* bc, const int* mask) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    invalidate_graphs(h);
    hipFree(h->w_dof); hipFree(h->w_mask); hipFree(h->w_fix); hipFree(h->w_vwall); hipFree(h->w_cells);
    hipFree(h->w_bc); hipFree(h->w_val);
    h->w_dof = h->w_mask = h->w_fix = h->w_vwall = h->w_cells = nullptr;
    h->w_bc = h->w_val = nullptr;
    h->n_walls = h->n_wall_cells = 0;
    if (n <= 0) return 0;
    if (h->deg != 1) return fail(PF2_E_STATE, "the wall fix-up acts on Q1 vertex DoFs (degree-1 fields)");
    if (!vertex || !dof || !bc || !mask) return fail(PF2_E_ARG, "wall arrays are NULL");
    struct W { int v, dof, mask; double bc[3]; };
    std::vector<W> ws(n);
    for (int i = 0; i < n; ++i) {
        if (vertex[i] < 0 || vertex[i] >= h->n_vertices || dof[i] < 0 || dof[i] >= h->n_dofs)
            return fail(PF2_E_ARG, "wall vertex/dof out of range");
        ws[i].v = vertex[i];
        ws[i].dof = dof[i];
        ws[i].mask = mask[i];
        for (int d = 0; d < 3; ++d) ws[i].bc[d] = d < h->dim ? bc[(size_t)i * h->dim + d] : 0.0;
    }
    std::stable_sort(ws.begin(), ws.end(), [](const W& a, const W& b) { return a.v < b.v; });
    for (int i = 1; i < n; ++i)
        if (ws[i].v == ws[i - 1].v) return fail(PF2_E_ARG, "duplicate wall vertex");
    std::vector<int> wd(n), wm(n), vwall(h->n_vertices, -1);
    std::vector<double> wbc((size_t)n * 3);
    std::vector<char> mark(h->n_cells, 0);
    for (int i = 0; i < n; ++i) {
        wd[i] = ws[i].dof;
        wm[i] = ws[i].mask;
        for (int d = 0; d < 3; ++d) wbc[(size_t)i * 3 + d] = ws[i].bc[d];
        vwall[ws[i].v] = i;
        for (int e = h->star_off[ws[i].v]; e < h->star_off[ws[i].v + 1]; ++e) mark[h->star_cell[e]] = 1;
    }
    std::vector<int> cells;
    for (int c = 0; c < h->n_cells; ++c)
        if (mark[c]) cells.push_back(c);
    if (dalloc(&h->w_dof, n) || dalloc(&h->w_mask, n) || dalloc(&h->w_fix, n) || dalloc(&h->w_vwall, h->n_vertices) ||
        dalloc(&h->w_cells, cells.size()) || dalloc(&h->w_bc, (size_t)n * 3) || dalloc(&h->w_val, (size_t)n * 3))
        return PF2_E_HIP;
    HIPC(hipMemcpyAsync(h->w_dof, wd.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->w_mask, wm.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->w_vwall, vwall.data(), h->n_vertices * 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->w_cells, cells.data(), cells.size() * 4, hipMemcpyHostToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->w_bc, wbc.data(), wbc.size() * 8, hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    h->n_walls = n;
    h->n_wall_cells = (int)cells.size();
    return 0;
}

int pf2_sort_particles_into_subdomains_and_cells(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    if (!h->params.keep_velocity_ext)
        return fail(PF2_E_STATE, "relocation without advection needs keep_velocity_ext");
    if (int r = from_mixed(h)) return r;
    // particles of unchanged cells keep their xi when one geometry update separates this re-sort from the last
    const int skip = h->geo_updates == 1;
    h->geo_updates = 0;
    if (int r = h->dim == 2 ? resort64<2>(h, skip) : resort64<3>(h, skip)) return r;
    // regroup the FP64 copy, then convert back
    int r = ensure_packed(h);
    if (!r) r = h->dim == 2 ? regroup<2>(h, 0, false, false, true) : regroup<3>(h, 0, false, false, true);
    if (r) return r;
    long long n;
    if (int q = get_count(h, &n)) return q;
    return to_mixed(h);
}

int pf2_check_empty_cells(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = settle(h)) return r;
    if (int r = from_mixed(h)) return r;
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    {
        Timed t(h, K_MISC);
        if (h->dim == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_keys<2>), dim3(grid_for(h->n_est, 256, 4096)), dim3(256), 0,
                               h->stream, dev_particles(h, h->cur), pp, w);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_keys<3>), dim3(grid_for(h->n_est, 256, 4096)), dim3(256), 0,
                               h->stream, dev_particles(h, h->cur), pp, w);
    }
    HIPC(hipGetLastError());
    int r = ensure_packed(h);
    if (!r) r = h->dim == 2 ? regroup<2>(h, 1, false, false, true) : regroup<3>(h, 1, false, false, true);
    if (r) return r;
    long long n;
    if (int q = get_count(h, &n)) return q;
    return to_mixed(h);
}

int pf2_update_vertices(pf2_handle h, const double* vertices, int recompute_maps) {
    if (!h || !vertices) return fail(PF2_E_ARG, "NULL argument");
    if (int r = settle(h)) return r;
    // positions of FP32-mixed particles are F_K(xi) on the mesh before the move
    if (int r = from_mixed(h)) return r;
    h->X.assign(vertices, vertices + (size_t)h->n_vertices * h->dim);
    HIPC(hipMemcpyAsync(h->dX, h->X.data(), h->X.size() * 8, hipMemcpyHostToDevice, h->stream));
    if (recompute_maps) {
        compute_star_dirs(h);
        if (int r = upload_star_dirs(h)) return r;
    }
    if (h->dim == 2) launch_geo<2>(h->stream, dev_mesh<2>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    else launch_geo<3>(h->stream, dev_mesh<3>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    h->geo_updates += 1;
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

// transform_grid + re-sort + check_empty_cells with one census; the new
// vertices come from the host (vertices) or from X0 + a W on the device
int move_mesh_impl(pf2_handle_s* h, const double* vertices, int recompute_maps, const double* scale) {
    if (!h->params.keep_velocity_ext)
        return fail(PF2_E_STATE, "relocation without advection needs keep_velocity_ext");
    if (int r = settle(h)) return r;
    // the FP64 re-sort reads no velocity field, and its regroup takes the
    // velocities of new particles from V through the DoF map when the
    // per-cell copies are stale (after a projection); the FP32-mixed path
    // regroups its FP32 state from the FP32 copies
    if (mixed(h))
        if (int r = ensure_packed(h)) return r;
    PBufs& b = h->buf[h->cur];
    const int grid = grid_for(h->n_est, 256, 4096);
    if (mixed(h)) {
        // locations on the mesh before the move, in FP64
        Timed t(h, K_MISC);
        if (h->dim == 2)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pos32<2>), dim3(grid), dim3(256), 0, h->stream, dev_mesh<2>(h), b.cell,
                               rec32(b), b.x, h->cap, h->counters);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pos32<3>), dim3(grid), dim3(256), 0, h->stream, dev_mesh<3>(h), b.cell,
                               rec32(b), b.x, h->cap, h->counters);
        HIPC(hipGetLastError());
    }
    if (vertices) {
        h->X.assign(vertices, vertices + (size_t)h->n_vertices * h->dim);
        HIPC(hipMemcpyAsync(h->dX, h->X.data(), h->X.size() * 8, hipMemcpyHostToDevice, h->stream));
        h->x_host_stale = false;
    } else {
        const long long n = (long long)h->n_vertices * h->dim;
        hipLaunchKernelGGL(k_displace, dim3(grid_for(n, 256, 4096)), dim3(256), 0, h->stream, h->dX0, h->dW, *scale,
                           h->dX, n);
        HIPC(hipGetLastError());
        h->x_host_stale = true;
    }
    if (recompute_maps) {
        compute_star_dirs(h);
        if (int r = upload_star_dirs(h)) return r;
    }
    if (h->dim == 2) launch_geo<2>(h->stream, dev_mesh<2>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    else launch_geo<3>(h->stream, dev_mesh<3>(h), mixed(h) ? h->dgeo32 : nullptr, h->dftol, h->dcellX, h->dcmoved);
    h->geo_updates += 1;
    DevParams pp = dev_params(h);
    DevWork w = dev_work(h);
    // particles of unchanged cells keep their xi when this is the only geometry update since the last re-sort
    const int skip = h->geo_updates == 1;
    h->geo_updates = 0;
    if (mixed(h)) {
        h->fp64_valid = false;
        DevQueues qu = dev_queues(h);
        const int rgrid = grid_for(h->n_est, 256, 16384), sgrid = grid_for(h->n_est / 2 + 1, 256, 2048);
        {
            Timed t(h, K_RESORT32);
            if (h->dim == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort32<2>), dim3(rgrid), dim3(256), 0, h->stream,
                                   dev_mesh<2>(h), b.cell, b.id, b.x, rec32(b), h->cap, pp, w, qu, skip);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort32<3>), dim3(rgrid), dim3(256), 0, h->stream,
                                   dev_mesh<3>(h), b.cell, b.id, b.x, rec32(b), h->cap, pp, w, qu, skip);
        }
        {
            Timed t(h, K_RESORT_SEARCH32);
            if (h->dim == 2)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort_search32<2>), dim3(sgrid), dim3(256), 0, h->stream,
                                   dev_mesh<2>(h), b.cell, b.x, b.vext, rec32(b), h->cap, pp, w, qu);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_resort_search32<3>), dim3(sgrid), dim3(256), 0, h->stream,
                                   dev_mesh<3>(h), b.cell, b.x, b.vext, rec32(b), h->cap, pp, w, qu);
        }
    } else if (int r = h->dim == 2 ? resort64<2>(h, skip) : resort64<3>(h, skip)) {
        return r;
    }
    HIPC(hipGetLastError());
    // relocation and reseed in one census: the same survivors, ids and
    // counts as the re-sort's census followed by check_empty_cells'.  A
    // handle stepping without graphs leaves its particles in a pending
    // layout that the next move regroups (k_cell_pend[32]): the re-sort's
    // own regroup write is skipped.  (Room for the new particles after the
    // current ones is checked on the estimate here and on the device.)
    const long long nc = h->n_cells;
    const bool pend = !h->graphs && h->omap && nc < (1 << 25) &&
                      h->n_est + nc * dev_params(h).nparts_in <= h->cap;
    return h->dim == 2 ? regroup<2>(h, 1, false, false, false, pend) : regroup<3>(h, 1, false, false, false, pend);
}

int pf2_move_mesh(pf2_handle h, const double* vertices, int recompute_maps) {
    if (!h || !vertices) return fail(PF2_E_ARG, "NULL argument");
    return move_mesh_impl(h, vertices, recompute_maps, nullptr);
}

int pf2_set_motion(pf2_handle h, const double* W) {
    if (!h || !W) return fail(PF2_E_ARG, "NULL argument");
    if (!h->dX0) return fail(PF2_E_STATE, "a prescribed motion needs keep_velocity_ext");
    const size_t n = (size_t)h->n_vertices * h->dim;
    HIPC(hipMemcpyAsync(h->dX0, h->dX, n * 8, hipMemcpyDeviceToDevice, h->stream));
    HIPC(hipMemcpyAsync(h->dW, W, n * 8, hipMemcpyHostToDevice, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int pf2_move_mesh_scaled(pf2_handle h, double a) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (!h->dX0) return fail(PF2_E_STATE, "a prescribed motion needs keep_velocity_ext");
    return move_mesh_impl(h, nullptr, 0, &a);
}

int pf2_reinterpolate_fields(pf2_handle h, const double* U, int n_extra, double* extra) {
    if (!h || !U) return fail(PF2_E_ARG, "NULL argument");
    if (n_extra < 0 || (n_extra > 0 && !extra)) return fail(PF2_E_ARG, "invalid extra fields");
    if (h->deg != 1) return fail(PF2_E_STATE, "field reinterpolation acts on Q1 fields");
    const size_t nvt = h->n_vertices, nd = h->n_dofs, dim = h->dim, ne = (size_t)n_extra * nd;
    double *dU = nullptr, *dVn = nullptr, *dE = nullptr;
    int r = 0;
    if (dalloc(&dU, nvt * dim) || dalloc(&dVn, nd * dim) || dalloc(&dE, 2 * ne)) r = PF2_E_HIP;
    auto done = [&](int rc) {
        hipStreamSynchronize(h->stream);
        hipFree(dU);
        hipFree(dVn);
        hipFree(dE);
        return rc;
    };
    if (r) return done(r);
    if (hipMemcpyAsync(dU, U, nvt * dim * 8, hipMemcpyHostToDevice, h->stream) ||
        hipMemcpyAsync(dVn, h->dV, nd * dim * 8, hipMemcpyDeviceToDevice, h->stream) ||
        (ne && hipMemcpyAsync(dE, extra, ne * 8, hipMemcpyHostToDevice, h->stream)) ||
        (ne && hipMemcpyAsync(dE + ne, dE, ne * 8, hipMemcpyDeviceToDevice, h->stream)))
        return done(fail(PF2_E_HIP, "reinterpolation upload failed"));
    const int g = grid_for((long long)nvt, 256);
    if (h->dim == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_reinterp<2>), dim3(g), dim3(256), 0, h->stream, dev_mesh<2>(h), h->dvbnd,
                           h->dvdof, dU, h->dV, dVn, dE, dE + ne, n_extra);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_reinterp<3>), dim3(g), dim3(256), 0, h->stream, dev_mesh<3>(h), h->dvbnd,
                           h->dvdof, dU, h->dV, dVn, dE, dE + ne, n_extra);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(h->dV, dVn, nd * dim * 8, hipMemcpyDeviceToDevice, h->stream) ||
        (ne && hipMemcpyAsync(extra, dE + ne, ne * 8, hipMemcpyDeviceToHost, h->stream)))
        return done(fail(PF2_E_HIP, "reinterpolation failed"));
    h->packed = false;   // V changed: the per-cell copies are refreshed before their next use
    return done(0);
}

int pf2_calculate_loads(pf2_handle h, const double* V, const double* P, int n_faces, const int* faces,
                        const pf2_loads_params* lp, double* out5) {
    if (!h || !P || !lp || !out5 || (n_faces > 0 && !faces) || n_faces < 0) return fail(PF2_E_ARG, "NULL argument");
    if (h->dim != 2 || h->deg != 1) return fail(PF2_E_STATE, "surface loads are computed for 2-D Q1 fields");
    for (int f = 0; f < n_faces; ++f)
        if (faces[2 * f] < 0 || faces[2 * f] >= h->n_cells || faces[2 * f + 1] < 0 || faces[2 * f + 1] > 3)
            return fail(PF2_E_ARG, "face (cell, face number) out of range");
    if (lp->dof_a >= h->n_dofs || lp->dof_e >= h->n_dofs) return fail(PF2_E_ARG, "pressure DoF out of range");
    const size_t nd = h->n_dofs;
    double *dV = nullptr, *dP = nullptr, *dT = nullptr, *dO = nullptr;
    int* dF = nullptr;
    int r = 0;
    if (dalloc(&dV, 2 * nd) || dalloc(&dP, nd) || dalloc(&dT, (size_t)8 * n_faces + 8) || dalloc(&dO, 8) ||
        dalloc(&dF, (size_t)2 * n_faces + 2))
        r = PF2_E_HIP;
    auto done = [&](int rc) {
        hipStreamSynchronize(h->stream);
        hipFree(dV);
        hipFree(dP);
        hipFree(dT);
        hipFree(dO);
        hipFree(dF);
        return rc;
    };
    if (r) return done(r);
    if (V) {
        if (hipMemcpyAsync(dV, V, 2 * nd * 8, hipMemcpyHostToDevice, h->stream)) return done(fail(PF2_E_HIP, "upload"));
    } else {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_to_comp_major<2>), dim3(grid_for((long long)nd, 256)), dim3(256), 0,
                           h->stream, h->dV, dV, (int)nd);
    }
    if (hipMemcpyAsync(dP, P, nd * 8, hipMemcpyHostToDevice, h->stream) ||
        (n_faces && hipMemcpyAsync(dF, faces, (size_t)n_faces * 8, hipMemcpyHostToDevice, h->stream)))
        return done(fail(PF2_E_HIP, "upload"));
    if (n_faces)
        hipLaunchKernelGGL(k_loads_terms, dim3(grid_for(2LL * n_faces, 256)), dim3(256), 0, h->stream, dev_mesh<2>(h),
                           dF, n_faces, dV, dP, lp->mu, dT);
    hipLaunchKernelGGL(k_loads_sum, dim3(1), dim3(64), 0, h->stream, dT, 2 * n_faces, dP, lp->rho, lp->u_mean,
                       lp->diam, lp->dof_a, lp->dof_e, dO);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(out5, dO, 5 * 8, hipMemcpyDeviceToHost, h->stream))
        return done(fail(PF2_E_HIP, "loads failed"));
    return done(0);
}

int pf2_initialize_maps(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (h->x_host_stale) {
        HIPC(hipMemcpyAsync(h->X.data(), h->dX, h->X.size() * 8, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
        h->x_host_stale = false;
    }
    compute_star_dirs(h);
    if (int r = upload_star_dirs(h)) return r;
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int pf2_n_global_particles(pf2_handle h, long long* n) {
    if (!h || !n) return fail(PF2_E_ARG, "NULL argument");
    return get_count(h, n);
}

int pf2_cell_counts(pf2_handle h, int* counts) {
    if (!h || !counts) return fail(PF2_E_ARG, "NULL argument");
    if (int r = settle(h)) return r;
    long long n;
    if (int r = get_count(h, &n)) return r;
    std::vector<int> off(h->n_cells + 1);
    HIPC(hipMemcpyAsync(off.data(), h->off[h->cur_off], off.size() * 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int c = 0; c < h->n_cells; ++c) counts[c] = off[c + 1] - off[c];
    return 0;
}

int pf2_particle_counter(pf2_handle h, uint32_t* count) {
    if (!h || !count) return fail(PF2_E_ARG, "NULL argument");
    HIPC(hipMemcpyAsync(count, h->pcounter, 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    return 0;
}

int pf2_stats(pf2_handle h, unsigned long long* out5) {
    if (!h || !out5) return fail(PF2_E_ARG, "NULL argument");
    std::vector<unsigned long long> rows((size_t)STAT_SHARDS * NSTAT);
    HIPC(hipMemcpyAsync(rows.data(), h->stats, rows.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 5; ++k) {
        unsigned long long t = 0;
        for (int r = 0; r < STAT_SHARDS; ++r) t += rows[(size_t)r * NSTAT + k];
        out5[k] = t;
    }
    return 0;
}

int pf2_step_info(pf2_handle h, long long* out, int cap) {
    if (!h || !out) return fail(PF2_E_ARG, "NULL argument");
    std::vector<int> pref((size_t)NQROUNDS * (QSHARDS + 2));
    int cnt[4];
    std::vector<int> offs(2);
    HIPC(hipMemcpyAsync(pref.data(), h->qprefix, pref.size() * 4, hipMemcpyDeviceToHost, h->stream));
    HIPC(hipMemcpyAsync(cnt, h->counters, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
    HIPC(hipStreamSynchronize(h->stream));
    const int arrivals = cnt[1];
    long long vals[6] = {};
    vals[0] = cnt[3];
    vals[1] = cnt[0];
    // queued per round of the last move (the FP64 path searches in one
    // round), then the relocations of the last re-sort after a mesh move
    for (int r = 0; r <= (mixed(h) ? MOVE_ROUNDS : 0); ++r) vals[2 + r] = pref[(size_t)r * (QSHARDS + 2) + QSHARDS + 1];
    vals[4] = pref[(size_t)RESORT_QUEUE * (QSHARDS + 2) + QSHARDS + 1];
    vals[5] = arrivals;
    int k = 0;
    for (; k < 6 && k < cap; ++k) out[k] = vals[k];
    return k;
}

int pf2_synchronize(pf2_handle h) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    long long n;
    return get_count(h, &n);
}

int pf2_set_profiling(pf2_handle h, int on) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = flush_events(h)) return r;
    h->profiling = on != 0;
    for (int i = 0; i < K_NKERNELS; ++i) {
        h->k_ms[i] = 0;
        h->k_n[i] = 0;
    }
    return 0;
}

int pf2_kernel_times(pf2_handle h, double* total_ms, long long* launches, int cap) {
    if (!h) return fail(PF2_E_ARG, "handle is NULL");
    if (int r = flush_events(h)) return r;
    for (int i = 0; i < K_NKERNELS && i < cap; ++i) {
        if (total_ms) total_ms[i] = h->k_ms[i];
        if (launches) launches[i] = h->k_n[i];
    }
    return K_NKERNELS;
}

const char* pf2_kernel_name(int i) { return (i >= 0 && i < K_NKERNELS) ? kKernelNames[i] : ""; }

}  // extern "C"
