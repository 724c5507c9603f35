// pf2_kernels.hpp — device kernels of the particle stage (FP64 path).
//
// Data layout in HBM (one simulation):
//   mesh     X[n_vertices*DIM] (row-major), cell_vertices / cell_dofs [n_cells*NV],
//            vertex star CSR (ascending cells) + normalised centre directions,
//            DoF -> (cell, local vertex) CSR in ascending cell order.
//   fields   V, Vold [n_dofs*DIM] node-major (one DoF's components adjacent).
//   particles structure of arrays, component-major (x[d*cap + i]), sorted by
//            cell, ascending id inside a cell; cell_off[n_cells+1] is the CSR.
// Per time step (pf2_step):
//   k_move         one thread per particle: correction, then substeps of
//                  advection in registers until done or the first failed
//                  own-cell test; such particles are parked in a queue.
//   k_search_resume  MOVE_ROUNDS + 1 rounds over the queue: vertex-star
//                  search in full wavefronts, then the remaining substeps.
//                  A particle that ends in another cell takes a slot in that
//                  cell's arrival row (add_arrival).
//   k_cell_count   one 16-lane group per cell: part counts over stayers +
//                  arrivals, survivors (MAX smallest ids per part), empty
//                  parts to reseed, final count (k_cell_big: larger cells).
//   k_cell_write   one 16-lane group per cell: survivors and new particles
//                  into the other buffer at their final positions.
//   k_cell_partials one group per cell: per-cell projection sums in
//                  ascending id order.
//   k_node         one thread per DoF: sum of the cell partials in ascending
//                  cell order, then division by the weight.
// No floating-point atomics anywhere.  Integer atomics only build lists whose
// order is erased again by the id ordering; counters that many wavefronts
// update are sharded over separate cache lines.
#pragma once

#include "pf2_geom.hpp"
#include <stdint.h>

namespace pf2 {

constexpr int WAVE = 64;

// sum over the wavefront (all lanes must call)
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Counters that many wavefronts update are spread over STAT_SHARDS cache
// lines: atomics on one address serialise in the memory system (about 11 ns
// each on MI355X), so a single counter updated once per wavefront would
// cost hundreds of microseconds per launch.
constexpr int STAT_SHARDS = 256;
constexpr int NSTAT = 8;   // u64 per shard row (64 bytes)

__device__ __forceinline__ int wave_shard() {
    return (int)((blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) & (STAT_SHARDS - 1));
}

__device__ __forceinline__ void stat_add(unsigned long long* stats, int k, unsigned long long v) {
    atomicAdd(&stats[wave_shard() * NSTAT + k], v);
}

// atomicAdd(&cnt[key], 1) for the lanes with pred set, one atomic per distinct
// key in the wavefront; returns the previous value for this lane (all lanes call)
__device__ __forceinline__ int wave_inc_by_key(int* cnt, bool pred, int key) {
    int result = 0;
    unsigned long long todo = __ballot(pred);
    const int lane = threadIdx.x & (WAVE - 1);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int k = __shfl(key, leader);
        const unsigned long long same = __ballot(pred && key == k) & todo;
        int base = 0;
        if (lane == leader) base = atomicAdd(&cnt[k], __popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1) result = base + __popcll(same & ((1ull << lane) - 1));
        todo &= ~same;
    }
    return result;
}

// flush per-lane statistics with one sharded atomic per wavefront (all lanes must call)
__device__ __forceinline__ void wave_stat(unsigned long long* stats, int k, int v) {
    const int t = wave_sum(v);
    if ((threadIdx.x & (WAVE - 1)) == 0 && t) stat_add(stats, k, (unsigned long long)t);
}

// Arrivals (particles that changed cell) of the lanes with pred set: particle
// i takes a slot of cell c's row, so the census finds a cell's arrivals with
// no grouping pass; arrivals past the row go to the cell's overflow list (the
// census then handles the cell in k_cell_big).  Lanes with the same cell
// share one atomic, and the atomics of all cells of the wavefront are issued
// together, so the wavefront waits for one returning atomic.  All lanes of
// the wavefront call.
template <class W>
__device__ __forceinline__ void add_arrival(const W& w, bool pred, int c, int i) {
    const int lane = threadIdx.x & (WAVE - 1);
    unsigned long long todo = __ballot(pred);
    if (todo == 0) return;
    int leader = lane, off = 0, cnt = 0;
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        const int key = __shfl(c, l);
        const unsigned long long same = __ballot(pred && c == key) & todo;
        if ((same >> lane) & 1) {
            leader = l;
            off = __popcll(same & ((1ull << lane) - 1));
        }
        if (lane == l) cnt = __popcll(same);
        todo &= ~same;
    }
    int base = 0;
    if (cnt) base = atomicAdd(&w.arr_cnt[c], cnt);
    base = __shfl(base, leader);
    if (!pred) return;
    const int slot = base + off;
    if (slot < w.aslots) {
        w.arr_slot[(size_t)c * w.aslots + slot] = i;
        return;
    }
    const int e = atomicAdd(&w.counters[4], 1);
    const int next = atomicExch(&w.ovf_head[c], e);
    w.ovf[e] = make_int2(i, next);
}

template <int DIM>
struct DevMesh {
    int n_cells, n_vertices, n_dofs;
    int nd;                // velocity DoFs per cell: 2^DIM (Q1) or 3^DIM (Q2)
    const double* X;
    const int* cv;
    const int* cd;         // per cell: its ND velocity DoFs (ND = QF<DIM, DEG>::ND)
    const int* star_off;
    const int* star_cell;
    const double* star_dir;
    const int* dof_off;
    const int* dof_cell;   // packed: cell * ND + local DoF
    const double* ftol;    // per cell: tolerance of the clearly-outside test (-1: none)
    const double* cellX;   // per cell: its vertex coordinates packed [NV][DIM] (copies of X)
    const int2* cell_star; // per cell and local vertex: [begin, end) of that vertex's star
    const int4* cell_nbr;  // per cell, the neighbours across faces 0..2*DIM-1 (-1 on the boundary or past
                           // the faces): one int4 in 2-D, two in 3-D
    const double* cellV;   // per cell: V at its DoFs [ND][DIM] (refreshed when V changes); null for
                           // FP32-mixed handles, whose particle kernels read FP32 copies
    const double* cellDV;  // per cell: V - V_old at its DoFs [ND][DIM] (FP64 handles)
    const double* Vn;      // V, node-major [n_dofs][DIM]
    const unsigned char* cmoved;   // per cell: 1 if the last geometry update changed a vertex of it
};

struct DevParticles {
    int* cell;
    uint32_t* id;
    double* x;
    double* xi;
    double* v;
    double* vext;   // may be null
    long long cap;
};

struct DevParams {
    double h;           // dt / substeps
    int substeps;
    int max_per_part;
    int q[3];
    double hq[3];       // 1.0 / q[d] (the part width, correctly rounded)
    int nparts_in;      // q0*q1[*q2]
    int keep_vext;
};

struct DevWork {
    int* dest;          // per particle: new cell or -1
    int* pkey;          // per particle: part key of the final position
    int* spos;          // per particle: position among the cell's survivors, -1 = dropped
    int* arr_cnt;       // per cell: arrivals
    int* arr_slot;      // per cell: aslots arrival slots (particle indices)
    int aslots;
    int2* ovf;          // arrivals past a cell's slots: (particle, next entry) lists
    int* ovf_head;      // per cell: first overflow entry or -1
    // per cell (n_cells + 1 for the scan): final count << 32 | spawns, and
    // its exclusive scan (particle offset << 32 | spawn offset): one 64-bit
    // scan gives both offsets (the spawn total stays below 2^32, so the low
    // halves never carry)
    unsigned long long* fs_cnt;
    unsigned long long* fs_off;
    int* new_off;       // n_cells + 1: particle offsets, written from fs_off by the cell writes
    unsigned long long* hitmask;   // per cell: in-range parts that hold a particle
    int* nsurv;         // per cell: survivors
    double* partial;    // per cell: NV * (DIM + 1) projection sums
    int* counters;      // [0] n_particles, [1] arrivals of the last regroup, [2] overflow flags,
                        // [3] count before the last regroup, [4] overflow entries in use,
                        // [5] id counter at the last advection (ids up to it hold an inverse-mapped xi),
                        // [6] slots of a pending layout (0: none; see k_cell_pend), [7] its segment end
    int* big;           // [0] number of cells listed for k_cell_big, then the cells
    int* big2;          // the same for the cells that the second census pass (rows of GCAP_L) lists
    int pend_mark;      // census of a pending regroup: survivors' destinations get DEST_KEPT
    uint32_t* particle_counter;
    unsigned long long* stats;   // STAT_SHARDS rows: lost, deleted, spawned, searches, zero-weight
};

// counters[2] bits 0 and 1 (particle capacity exceeded, a cell past the
// census limit) are sticky: the regroup that set them left the particle
// buffers incomplete and the new offsets beyond the capacity.  Every later
// kernel that reads particles or offsets then returns at once, so steps
// already queued (graph replays, chunks issued ahead) touch no memory out of
// bounds; the host reports PF2_E_CAPACITY at its next synchronisation.
constexpr int DEAD_FLAGS = 3;
__device__ __forceinline__ bool state_dead(const int* counters) { return (counters[2] & DEAD_FLAGS) != 0; }

// vertex coordinates and nodal velocities of cell c, from the packed per-cell
// copies, or (GATHER: FP32-mixed handles, which keep no FP64 velocity copy)
// with the velocities gathered from V through the cell's DoF map (the same
// values).  The choice is a template argument so that the kernels of FP64
// handles carry no second path.
template <int DIM, bool GATHER = false, int DEG = 1>
__device__ __forceinline__ void load_cell(const DevMesh<DIM>& m, int c, double (*P)[DIM], double (*Fc)[DIM]) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    const double* cx = m.cellX + (size_t)c * NV * DIM;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) P[k][d] = cx[k * DIM + d];
    if constexpr (GATHER) {
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const double* vf = m.Vn + (size_t)m.cd[(size_t)c * ND + k] * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d) Fc[k][d] = vf[d];
        }
    } else {
        const double* cf = m.cellV + (size_t)c * ND * DIM;
#pragma unroll
        for (int k = 0; k < ND; ++k)
#pragma unroll
            for (int d = 0; d < DIM; ++d) Fc[k][d] = cf[k * DIM + d];
    }
}

// the same, choosing by whether the handle keeps the FP64 copy (seeding)
template <int DIM, int DEG = 1>
__device__ __forceinline__ void load_cell_any(const DevMesh<DIM>& m, int c, double (*P)[DIM], double (*Fc)[DIM]) {
    if (m.cellV) load_cell<DIM, false, DEG>(m, c, P, Fc);
    else load_cell<DIM, true, DEG>(m, c, P, Fc);
}

// vertex coordinates of cell c (packed per-cell copy: one dependent load level)
template <int DIM>
__device__ __forceinline__ void load_vertices(const DevMesh<DIM>& m, int c, double (*P)[DIM]) {
    constexpr int NV = Q1<DIM>::NV;
    const double* cx = m.cellX + (size_t)c * NV * DIM;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) P[k][d] = cx[k * DIM + d];
}

// vertex coordinates gathered from X through the cell's vertex indices
template <int DIM>
__device__ __forceinline__ void gather_vertices(const DevMesh<DIM>& m, int c, double (*P)[DIM]) {
    constexpr int NV = Q1<DIM>::NV;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int gv = m.cv[(size_t)c * NV + k];
#pragma unroll
        for (int d = 0; d < DIM; ++d) P[k][d] = m.X[(size_t)gv * DIM + d];
    }
}

// u = sum_k N_k * F_k accumulated from zero over the cell's DoFs in order
// (velocity at a reference point)
template <int DIM, int DEG = 1>
__device__ __forceinline__ void interp(const double (*Fc)[DIM], const double* xi, double* u) {
    constexpr int ND = QF<DIM, DEG>::ND;
    double N[ND];
    shape_q<DIM, DEG>(xi, N);
#pragma unroll
    for (int d = 0; d < DIM; ++d) u[d] = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) u[d] += N[k] * Fc[k][d];
}

template <int DIM>
__device__ __forceinline__ int part_key(const DevParams& pp, const double* xi) {
    // bits 0-15: part index with the phantom value q_d (the grouping key);
    // bits 16-23: 1 + ordinal of an in-range part, 0 for a phantom part
    int key = 0, stride = 1, ord = 0;
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double h = pp.hq[d];
        const unsigned num = (unsigned)(xi[d] / h);
        key += (int)num * stride;
        stride *= pp.q[d] + 1;
        in = in && (int)num < pp.q[d];
        ord = ord * pp.q[d] + (int)num;
    }
    return key | ((in ? ord + 1 : 0) << 16);
}

// in-range part ordinal of a packed part key, -1 for a phantom part
__device__ __forceinline__ int part_ordinal(int pkey) { return (pkey >> 16) - 1; }

template <int DIM>
__device__ __forceinline__ void lattice_point(const DevParams& pp, int ordinal, double* xi) {
    int num[3];
    int o = ordinal;
#pragma unroll
    for (int d = DIM - 1; d >= 0; --d) {
        num[d] = o % pp.q[d];
        o /= pp.q[d];
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double h = pp.hq[d];
        xi[d] = (num[d] + 1.0 / 2) * h;
    }
}

// Relocation of a particle whose position left cell c
// (pfem2ParticleHandler::sort_particles_into_subdomains_and_cells, the part
// after the own-cell test, cylinder2D/pfem2particle.cpp:278-386).  The
// reference sorts the star of the selected vertex by decreasing alignment
// (stable for equal keys) and takes the first cell that contains the point;
// that is the containing cell with the largest key, earliest star position on
// ties, which is what the loop below selects.  Returns -1 if no star cell
// contains the point.
struct NoPrefetch {
    __device__ void load(int) {}
};

// nodal velocities of a candidate cell (FP64 path)
template <int DIM, int DEG = 1>
struct PrefetchV64 {
    static constexpr int ND = QF<DIM, DEG>::ND;
    const double* cellV;
    double Vc[ND][DIM];
    __device__ void load(int cc) {
        const double* cf = cellV + (size_t)cc * ND * DIM;
#pragma unroll
        for (int k = 0; k < ND; ++k)
#pragma unroll
            for (int d = 0; d < DIM; ++d) Vc[k][d] = cf[k * DIM + d];
    }
};

// `Q` returns the vertices of the cell found; `pre.load(cell)` is called for
// each candidate before its containment test, so that a caller can have the
// candidate's other data loaded alongside its vertices (the data of the
// last candidate loaded is the found cell's).
// Face through which the step from xold to x most likely left the cell: the
// linearised reference displacement J(xi_old)^-1 (x - xold) added to xi_old,
// the component furthest outside [0, 1] (face 2a + s is xi_a = s), -1 if
// none.  Only a guess: the caller verifies it.
template <int DIM>
__device__ __forceinline__ int exit_face(const double (*P)[DIM], const double* xi_old, const double* xold,
                                         const double* x) {
    constexpr int NV = Q1<DIM>::NV;
    double J[DIM][DIM], Ji[DIM][DIM], dx[DIM], dxi[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) J[i][j] = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            double w = ((k >> j) & 1) ? 1.0 : -1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                if (d != j) w *= ((k >> d) & 1) ? xi_old[d] : 1.0 - xi_old[d];
#pragma unroll
            for (int i = 0; i < DIM; ++i) J[i][j] += w * P[k][i];
        }
    // J^-1 = adj(J) / det: the face is chosen from values scaled by det, so
    // no division (the choice is only a guess; a cell with det <= 0 gets an
    // arbitrary one, which the caller's containment test rejects)
    if constexpr (DIM == 2) {
        Ji[0][0] = J[1][1];
        Ji[0][1] = -J[0][1];
        Ji[1][0] = -J[1][0];
        Ji[1][1] = J[0][0];
    } else {
        Ji[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        Ji[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
        Ji[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
        Ji[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        Ji[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
        Ji[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
        Ji[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        Ji[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
        Ji[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    }
    double dt = 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) dt += J[0][j] * Ji[j][0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) dx[d] = x[d] - xold[d];
    matvec<DIM>(Ji, dx, dxi);
    int f = -1;
    double worst = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const double e = xi_old[d] * dt + dxi[d];   // det * (reference coordinate after the move)
        const double o = e < 0.0 ? -e : e - dt;
        if (o > worst) {
            worst = o;
            f = 2 * d + (e < 0.0 ? 0 : 1);
        }
    }
    return f;
}

// margin in reference coordinates beyond which a point inside one cell is
// inside no other cell under the containment test (the Newton acceptance
// radius is 1e-10 of a cell diameter)
constexpr double NBR_MARGIN = 1e-6;

template <int DIM, class Pre, bool SHORTCUT = true>
__device__ __forceinline__ int relocate(const DevMesh<DIM>& m, int c, const double (*P)[DIM], const double* xi_old,
                        const double* vext, const double* x, double* xi_out, double (*Q)[DIM], Pre& pre) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int CH = 4;   // star entries read together
    int nb[2 * DIM];
    if constexpr (SHORTCUT) {
#pragma unroll
        for (int h = 0; h < DIM - 1; ++h) {
            const int4 q = m.cell_nbr[(size_t)c * (DIM - 1) + h];
            nb[4 * h] = q.x;
            nb[4 * h + 1] = q.y;
            if (4 * h + 2 < 2 * DIM) nb[4 * h + 2] = q.z;
            if (4 * h + 3 < 2 * DIM) nb[4 * h + 3] = q.w;
        }
    }
    double xold[DIM], vn[DIM], t[DIM];
    forward<DIM>(P, xi_old, xold);
    const double rv = 1.0 / vnorm<DIM>(vext);
#pragma unroll
    for (int d = 0; d < DIM; ++d) vn[d] = vext[d] * rv;
#pragma unroll
    for (int d = 0; d < DIM; ++d) t[d] = P[0][d] - xold[d];
    double rn = 1.0 / vnorm<DIM>(t);
#pragma unroll
    for (int d = 0; d < DIM; ++d) t[d] *= rn;
    double best = vdot<DIM>(vn, t);
    int cvk = 0;
    double dir[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) dir[d] = P[0][d];
#pragma unroll
    for (int k = 1; k < NV; ++k) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) t[d] = P[k][d] - xold[d];
        rn = 1.0 / vnorm<DIM>(t);
#pragma unroll
        for (int d = 0; d < DIM; ++d) t[d] *= rn;
        const double a = vdot<DIM>(vn, t);
        if (a > best) {
            cvk = k;
            best = a;
#pragma unroll
            for (int d = 0; d < DIM; ++d) dir[d] = P[k][d];
        }
    }
    // the selected vertex's star range, loaded alongside the shortcut's cell
    const int2 se = m.cell_star[(size_t)c * NV + cvk];
    if constexpr (SHORTCUT) {
        // Face-neighbour shortcut.  The neighbour g across the exit face is in
        // the selected vertex's star when that vertex lies on the face; if g
        // contains x at least NBR_MARGIN inside, no other cell contains x,
        // so g is the first containing cell of the star in any order: the
        // result of the loop below, with the same inverse map.
        const int f = exit_face<DIM>(P, xi_old, xold, x);
        if (f >= 0 && ((cvk >> (f >> 1)) & 1) == (f & 1)) {
            int g = nb[0];
#pragma unroll
            for (int q = 1; q < 2 * DIM; ++q)
                if (f == q) g = nb[q];
            if (g >= 0) {
                double r[DIM];
                load_vertices<DIM>(m, g, Q);
                pre.load(g);
                bool in = inverse_map<DIM>(Q, x, r, m.ftol[g], true);
#pragma unroll
                for (int d = 0; d < DIM; ++d) in = in && r[d] >= NBR_MARGIN && r[d] <= 1.0 - NBR_MARGIN;
                if (in) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi_out[d] = r[d];
                    return g;
                }
            }
        }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) dir[d] = x[d] - dir[d];
    rn = 1.0 / vnorm<DIM>(dir);
#pragma unroll
    for (int d = 0; d < DIM; ++d) dir[d] *= rn;
    const int s0 = se.x, n = se.y - se.x;
    if (n <= 32) {
        // candidates in decreasing key order (ties: star order), first
        // containing cell wins -- the reference's sorted loop, stopping early.
        // Each pass re-reads the star (cache hits, CH entries at a time)
        // instead of holding the keys in registers.
        unsigned tried = 0;
        for (int it = 0; it < n; ++it) {
            int best_e = -1, cc = -1;
            double bk = 0.0;
            for (int e0 = 0; e0 < n; e0 += CH) {
                double key[CH];
                int cell[CH];
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const int e = e0 + u < n ? e0 + u : n - 1;
                    double sd[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) sd[d] = m.star_dir[(size_t)(s0 + e) * DIM + d];
                    key[u] = vdot<DIM>(sd, dir);
                    cell[u] = m.star_cell[s0 + e];
                }
#pragma unroll
                for (int u = 0; u < CH; ++u) {
                    const int e = e0 + u;
                    if (e < n && !((tried >> e) & 1u) && (best_e < 0 || key[u] > bk)) {
                        best_e = e;
                        bk = key[u];
                        cc = cell[u];
                    }
                }
            }
            tried |= 1u << best_e;
            double r[DIM];
            load_vertices<DIM>(m, cc, Q);
            pre.load(cc);
            if (inverse_map<DIM>(Q, x, r, m.ftol[cc], true) && inside_unit<DIM>(r)) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi_out[d] = r[d];
                return cc;
            }
        }
        return -1;
    }
    // stars of more than 32 cells: the containing cell with the largest key
    int bestc = -1;
    double bestkey = 0.0;
    for (int e = s0; e < s0 + n; ++e) {
        double sd[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sd[d] = m.star_dir[(size_t)e * DIM + d];
        const double key = vdot<DIM>(sd, dir);
        if (bestc >= 0 && !(key > bestkey)) continue;
        const int cc = m.star_cell[e];
        double r[DIM];
        load_vertices<DIM>(m, cc, Q);
        if (inverse_map<DIM>(Q, x, r, m.ftol[cc], true) && inside_unit<DIM>(r)) {
            bestc = cc;
            bestkey = key;
#pragma unroll
            for (int d = 0; d < DIM; ++d) xi_out[d] = r[d];
        }
    }
    if (bestc >= 0) {
        load_vertices<DIM>(m, bestc, Q);
        pre.load(bestc);
    }
    return bestc;
}

// ---------------------------------------------------------------------------
// Advection substeps (move_particles, TurekBenchmark/pfem2particle.cpp:813-856).
// Particles are independent within move_particles, so the substeps of one
// particle run back to back in registers: u = sum N_k V_k at the current
// reference point, velocity_ext = u, x += (dt/S) u, then the own-cell test
// (pfem2ParticleHandler::sort_particles_into_subdomains_and_cells,
// TurekBenchmark/pfem2particle.cpp:286-302).  A particle whose own-cell test
// fails is parked (position advanced, old reference point, velocity_ext and
// substep index stored) and queued; the vertex-star searches of all queued
// particles then run the vertex-star search in full wavefronts, which
// continue the remaining substeps.
//   FP64 path: one search round (k_search_resume) whose lanes follow their
//   particle through every further crossing and refill from the queue when
//   it is done (search_body).
//   FP32-mixed path: MOVE_ROUNDS parking rounds (k_search_resume32 parks a
//   particle that leaves its new cell again), then a last round that loops
//   every remaining particle to its end.
#ifndef PF2_MOVE_ROUNDS
#define PF2_MOVE_ROUNDS 1
#endif
constexpr int MOVE_ROUNDS = PF2_MOVE_ROUNDS;
static_assert(MOVE_ROUNDS == 1, "one parking round (two measured slower)");
// the queue segments of the refilling search (one cursor each, on its own
// cache line: a single cursor for the whole grid serialises the wavefronts'
// atomics) and the idle lanes a wavefront gathers before it refills
constexpr int RSEGS = 256;
constexpr int REFILL_MIN = 16;
// the round whose shard counters hold the cursors: after the FP32-mixed
// path's last round, so no search pushes to it; the counters are cleared
// with the other queue counters after each regroup
constexpr int CURSOR_ROUND = MOVE_ROUNDS + 1;
// the queue of the re-sort's relocations after a mesh move (both paths)
constexpr int RESORT_QUEUE = MOVE_ROUNDS + 2;
// queues (rounds) of the counters: the moves' rounds, the cursors, the re-sort
constexpr int NQROUNDS = MOVE_ROUNDS + 3;

// Queues of parked particles, one per round, ping-ponging between two
// buffers.  A queue is QSHARDS shards of shard_cap entries plus an overflow
// region of cap entries; each shard's counter sits on its own 128-byte line
// and a wavefront appends to shard (wavefront id mod QSHARDS) with one atomic.
constexpr int QSHARDS = 256;
constexpr int QSTRIDE = 32;   // ints between shard counters

struct DevQueues {
    int2* data[2];        // ping-pong buffers of (particle, cell): QSHARDS * shard_cap + overflow
    int* count;           // NQROUNDS rounds x (QSHARDS + 1) counters, QSTRIDE apart
    int* prefix;          // per round: QSHARDS + 2 ints (exclusive prefix, overflow start, total)
    int shard_cap;
    double* park_vext;    // DIM x cap: velocity_ext of a parked particle
    int* park_s;          // substep index of a parked particle
};

// Loops of wavefronts over n items (particles or queue entries): the
// wavefront handles items base + lane for base = begin, begin + step, ...
// below end, and appends to queue shard `shard`.  Shards follow the item
// order, so a queue keeps the order of the particles (stored by cell) up to
// the order within a shard, and the searches of one wavefront touch nearby
// cells.
struct WaveLoop {
    int begin, end, step, shard;
};

// The workgroups a kernel body loops over: the hardware grid, or, in a
// batched launch that steps several simulations at once, the simulation's
// own range of the launch's workgroups (index within it and its size).
struct Grid {
    unsigned bid, nblk;
};
__device__ __forceinline__ Grid hw_grid() { return Grid{blockIdx.x, gridDim.x}; }

// Both loops work in 32-bit unsigned arithmetic on wavefront-uniform values
// (grids of at most 2^16 workgroups, n < 2^31), which the compiler keeps in
// scalar registers: a 64-bit division per wavefront costs more vector
// instructions than a substep.

// grid-stride: wavefront k takes chunks k, k + W, ... of WAVE items
__device__ __forceinline__ WaveLoop stride_loop(int n, Grid G = hw_grid()) {
    const unsigned wpb = blockDim.x / WAVE, nw = G.nblk * wpb;
    const unsigned k = G.bid * wpb + (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    WaveLoop L;
    L.begin = (int)(k * WAVE);
    L.end = n;
    L.step = (int)(nw * WAVE);
    L.shard = (int)(k * QSHARDS / nw);
    return L;
}

// one contiguous range of whole workgroup-sized chunks per workgroup, the
// wavefronts of a workgroup interleaved within it
__device__ __forceinline__ WaveLoop range_loop(int n, Grid G = hw_grid()) {
    const unsigned g = G.nblk, b = G.bid, un = (unsigned)n;
    const unsigned per = ((un + g - 1) / g + blockDim.x - 1) / blockDim.x * blockDim.x;
    const unsigned long long b0l = (unsigned long long)b * per;
    const unsigned b0 = b0l < un ? (unsigned)b0l : un;
    const unsigned wpb = blockDim.x / WAVE, wv = (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    WaveLoop L;
    L.begin = (int)(b0 + wv * WAVE);
    L.end = (int)(un - b0 > per ? b0 + per : un);
    L.step = blockDim.x;
    L.shard = (int)((b * wpb + wv) * QSHARDS / (g * wpb));
    return L;
}

__device__ __forceinline__ int* queue_count(const DevQueues& qu, int round, int shard) {
    return qu.count + ((size_t)round * (QSHARDS + 1) + shard) * QSTRIDE;
}

// append `value` to shard `sh` for the lanes with pred set (all lanes of the
// wavefront call)
__device__ __forceinline__ void queue_push(const DevQueues& qu, int round, bool pred, int2 value, int sh) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = threadIdx.x & (WAVE - 1);
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(queue_count(qu, round, sh), __popcll(mask));
    base = __shfl(base, leader);
    const int slot = base + __popcll(mask & ((1ull << lane) - 1));
    const bool spill = pred && slot >= qu.shard_cap;
    int2* data = qu.data[round & 1];
    if (pred && !spill) data[(size_t)sh * qu.shard_cap + slot] = value;
    const unsigned long long smask = __ballot(spill);
    if (smask == 0) return;
    const int l2 = __ffsll((long long)smask) - 1;
    int ob = 0;
    if (lane == l2) ob = atomicAdd(queue_count(qu, round, QSHARDS), __popcll(smask));
    ob = __shfl(ob, l2);
    if (spill) data[(size_t)QSHARDS * qu.shard_cap + ob + __popcll(smask & ((1ull << lane) - 1))] = value;
}

// Exclusive prefix of the valid entries per shard, then the overflow start
// and the total, computed by each workgroup of the search kernel from the
// shard counters (blockDim.x == QSHARDS: one thread per shard; the counters
// are a few cache lines, so the redundant work costs less than a launch).
// Workgroup 0 also stores it for diagnostics.
__device__ __forceinline__ int queue_prefix(const DevQueues& qu, int round, int* s_pref, Grid G = hw_grid()) {
    __shared__ int s_wave[QSHARDS / WAVE];
    const int t = threadIdx.x;
    const int lane = t & (WAVE - 1), wv = t / WAVE;
    int v = *queue_count(qu, round, t);
    v = v < qu.shard_cap ? v : qu.shard_cap;
    int incl = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == WAVE - 1) s_wave[wv] = incl;
    __syncthreads();
    int wbase = 0;
    for (int q = 0; q < wv; ++q) wbase += s_wave[q];
    s_pref[t] = wbase + incl - v;
    if (t == QSHARDS - 1) {
        const int tot = wbase + incl;
        s_pref[QSHARDS] = tot;
        s_pref[QSHARDS + 1] = tot + *queue_count(qu, round, QSHARDS);
    }
    __syncthreads();
    if (G.bid == 0) {
        int* pref = qu.prefix + (size_t)round * (QSHARDS + 2);
        pref[t] = s_pref[t];
        if (t < 2) pref[QSHARDS + t] = s_pref[QSHARDS + t];
    }
    return s_pref[QSHARDS + 1];
}

// particle index of queue entry j
__device__ __forceinline__ int2 queue_get(const DevQueues& qu, int round, const int* s_pref, int j) {
    const int2* data = qu.data[round & 1];
    if (j >= s_pref[QSHARDS]) return data[(size_t)QSHARDS * qu.shard_cap + (j - s_pref[QSHARDS])];
    int lo = 0, hi = QSHARDS - 1;   // last shard with s_pref <= j
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pref[mid] <= j) lo = mid;
        else hi = mid - 1;
    }
    return data[(size_t)lo * qu.shard_cap + (j - s_pref[lo])];
}

// A parked particle whose own-cell test k_move left to the search kernel
// (flag in its parked substep index)
constexpr int PARK_NEWTON = 1 << 30;
// destination mark of a particle that a pending census kept (k_cell_pend):
// destination = cell | DEST_KEPT (cell indices stay below 2^25 there)
constexpr int DEST_KEPT = 1 << 30;

// Runs substeps until done (returns 0) or the own-cell test fails (returns
// 1; s is the substep whose test failed).  pending: the test of substep s at
// the current x comes first, before any new substep.
// CLOSED (2-D, k_move): the own-cell test takes only inverse_map's closed
// form and returns 2 where inverse_map would go on to its clearly-outside
// test and Newton iteration (the point is outside the closed form's unit
// square, or the closed form is degenerate); the search kernel finishes that
// test (pending).  This keeps the Newton code, and its registers, out of
// k_move: 94 instead of 167 registers, 5 instead of 3 wavefronts per SIMD.
template <int DIM, int DEG = 1, bool CLOSED = false>
__device__ __forceinline__ int advance64(const DevMesh<DIM>& m, const double* __restrict__ V, const DevParams& pp,
                                         const DevParticles& A, int i, int c, double* x, double* xi, double* vext,
                                         int& s, const double (*P)[DIM], const double (*Vc)[DIM], double tol,
                                         bool pending = false) {
    for (;; ++s) {
        if (!pending) {
            if (s >= pp.substeps) return 0;
            double u[DIM];
            interp<DIM, DEG>(Vc, xi, u);
#pragma unroll
            for (int d = 0; d < DIM; ++d) vext[d] = u[d];
#pragma unroll
            for (int d = 0; d < DIM; ++d) u[d] *= pp.h;
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = x[d] + u[d];
            if (pp.keep_vext && s == pp.substeps - 1)
#pragma unroll
                for (int d = 0; d < DIM; ++d) A.vext[(size_t)d * A.cap + i] = vext[d];
        }
        pending = false;
        double r[DIM];
        if constexpr (CLOSED && DIM == 2) {
            // inverse_map<2>(P, x, r, tol, false) up to its clearly-outside test
            if (inverse_quadratic(P, x, r)) {
                const double eps = 1e-15;
                if (-eps <= r[1] && r[1] <= 1 + eps && -eps <= r[0] && r[0] <= 1 + eps) {
                    if (!inside_unit<DIM>(r)) return 1;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = r[d];
                    continue;
                }
            }
            return 2;
        } else {
            if (inverse_map<DIM>(P, x, r, tol, false) && inside_unit<DIM>(r)) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[d] = r[d];
            } else {
                return 1;
            }
        }
    }
}

// final bookkeeping of a particle whose substeps are complete (all lanes call)
// (seg_end: particles from it on, new particles of a pending layout, are in
// no cell's segment, so they arrive in their cell)
template <int DIM>
__device__ __forceinline__ void finish_particle(const DevParams& pp, const DevParticles& A, const DevWork& w,
                                                bool valid, int i, int c, const double* x, const double* xi,
                                                bool write_state, int seg_end = 0x7fffffff) {
    add_arrival(w, valid && c >= 0 && (c != A.cell[valid ? i : 0] || i >= seg_end), c, i);
    if (!valid) return;
    if (write_state) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            A.x[(size_t)d * A.cap + i] = x[d];
            A.xi[(size_t)d * A.cap + i] = xi[d];
        }
    }
    w.dest[i] = c;
    if (c >= 0) w.pkey[i] = part_key<DIM>(pp, xi);
}

// park a particle for the search of the next round (all lanes call)
template <int DIM>
__device__ __forceinline__ void park_particle(const DevParticles& A, const DevWork& w, const DevQueues& qu,
                                              int round_out, int shard, bool park, int i, int c,
                                              const double* x, const double* xi, const double* vext, int s) {
    queue_push(qu, round_out, park, make_int2(i, c), shard);
    if (!park) return;
    qu.park_s[i] = s;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        A.x[(size_t)d * A.cap + i] = x[d];
        A.xi[(size_t)d * A.cap + i] = xi[d];
        qu.park_vext[(size_t)d * A.cap + i] = vext[d];
    }
}

// k_move: [correction] + substeps until done or the first failed own-cell test.
//   correct_particles_velocities  TurekBenchmark/pfem2particle.cpp:792-811
template <int DIM, int DEG = 1>
__device__ __forceinline__ void move_body(const Grid& G, const DevMesh<DIM>& m, const double* __restrict__ V,
                                          const double* __restrict__ Vold, const DevParticles& A, const DevParams& pp,
                                          const DevWork& w, const DevQueues& qu, int do_correct, int do_move) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    // a pending layout (k_cell_pend): its slots, the current cell of a kept
    // particle in dest (with DEST_KEPT), new particles from the segment end on
    const int pn = w.counters[6];
    const int n = state_dead(w.counters) ? 0 : (pn > 0 ? pn : w.counters[0]);
    const int seg_end = pn > 0 ? w.counters[7] : 0x7fffffff;
    const int lane = threadIdx.x & (WAVE - 1);
    // every particle that exists now ends this advection with xi from an
    // inverse map (the re-sort after a mesh move relies on it)
    if (do_move && G.bid == 0 && threadIdx.x == 0) w.counters[5] = (int)*w.particle_counter;
    const WaveLoop L = stride_loop(n, G);
    for (int base = L.begin; base < L.end; base += L.step) {
        const int i = base + lane;
        bool valid = i < L.end;
        int c = 0;
        if (valid && pn > 0) {
            const int dk = w.dest[i];
            c = dk & ~DEST_KEPT;
            if (dk < 0 || !(dk & DEST_KEPT)) {
                // not kept by the pending census: gone (no cell, no arrival)
                w.dest[i] = -1;
                valid = false;
            }
        } else if (valid) {
            c = A.cell[i];
        }
        double x[DIM], xi[DIM], vext[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = valid ? A.xi[(size_t)d * A.cap + i] : 0.5;
        if (do_correct && valid) {
            double v[DIM], N[ND];
#pragma unroll
            for (int d = 0; d < DIM; ++d) v[d] = A.v[(size_t)d * A.cap + i];
            shape_q<DIM, DEG>(xi, N);
            const double* dv = m.cellDV + (size_t)c * ND * DIM;
#pragma unroll
            for (int k = 0; k < ND; ++k)
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = v[d] + N[k] * dv[k * DIM + d];
#pragma unroll
            for (int d = 0; d < DIM; ++d) A.v[(size_t)d * A.cap + i] = v[d];
        }
        if (!do_move) continue;
        int st = 0, s = 0;
        if (valid) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = A.x[(size_t)d * A.cap + i];
            double P[NV][DIM], Vc[ND][DIM];
            load_cell<DIM, false, DEG>(m, c, P, Vc);
            st = advance64<DIM, DEG, true>(m, V, pp, A, i, c, x, xi, vext, s, P, Vc, m.ftol[c]);
        }
        park_particle<DIM>(A, w, qu, 0, L.shard, valid && st != 0, i, c, x, xi, vext, st == 2 ? s | PARK_NEWTON : s);
        finish_particle<DIM>(pp, A, w, valid && st == 0, i, c, x, xi, true, seg_end);
    }
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? (DEG == 1 ? 5 : 3) : 2))) k_move(DevMesh<DIM> m, const double* __restrict__ V,
                                              const double* __restrict__ Vold, DevParticles A, DevParams pp,
                                              DevWork w, DevQueues qu, int do_correct, int do_move) {
    move_body<DIM, DEG>(hw_grid(), m, V, Vold, A, pp, w, qu, do_correct, do_move);
}

// k_search_resume: the vertex-star search of the parked particles, then
// their remaining substeps, in one round.  A lane follows its particle
// through every further crossing and, when the particle is done, refills
// from the queue: the wavefront's segment of the queue has a cursor (one
// returning atomic per wavefront and refill, taken once REFILL_MIN lanes are
// idle or all are), so a wavefront does not wait for its longest crossing
// chain before it takes new particles, and no particle is parked twice.
// Every iteration (one search, then the substeps up to the next exit)
// consumes a substep, so each lane ends, and a wavefront leaves when its
// segment is drained and none of its lanes holds a particle.
template <int DIM, int DEG = 1, bool RESORT = false>
__device__ __forceinline__ void search_body(const Grid& G, const DevMesh<DIM>& m, const double* __restrict__ V,
                                            const DevParticles& A, const DevParams& pp, const DevWork& w,
                                            const DevQueues& qu, int round) {
    constexpr int NV = Q1<DIM>::NV;
    __shared__ int s_pref[QSHARDS + 2];
    const int n = queue_prefix(qu, round, s_pref, G);
    // the wavefront's segment of the queue and its cursor: as many segments
    // as the grid has wavefronts, up to RSEGS, so every segment has one
    const unsigned wpb = blockDim.x / WAVE, nwv = G.nblk * wpb;
    const unsigned gw = G.bid * wpb + (unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int nseg = (int)(nwv < (unsigned)RSEGS ? nwv : (unsigned)RSEGS);
    const int seg = (int)(gw % (unsigned)nseg);
    const int s0 = (int)((long long)n * seg / nseg), s1 = (int)((long long)n * (seg + 1) / nseg);
    int* cursor = queue_count(qu, CURSOR_ROUND, seg);
    // new particles of a pending layout (from its segment end on) arrive in their cell
    const int seg_end = !RESORT && w.counters[6] > 0 ? w.counters[7] : 0x7fffffff;
    const int lane = threadIdx.x & (WAVE - 1);
    const unsigned long long below = (1ull << lane) - 1;
    int lost = 0, searches = 0;
    bool have = false, more = true, newton = false;
    int i = 0, c = 0, s = 0;
    double x[DIM], xi[DIM], vext[DIM], P[NV][DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = xi[d] = vext[d] = 0.0;
    for (;;) {
        if (more) {   // wavefront-uniform
            const unsigned long long need = __ballot(!have);
            if (need && (__popcll(need) >= REFILL_MIN || need == ~0ull)) {
                const int k = __popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(cursor, k);
                base = s0 + __shfl(base, leader);
                more = base + k < s1;
                const int j = base + __popcll(need & below);
                if (!have && j < s1) {
                    const int2 e = queue_get(qu, round, s_pref, j);
                    i = e.x;
                    c = e.y;
                    const int ps = qu.park_s[i];
                    s = ps & ~PARK_NEWTON;
                    newton = (ps & PARK_NEWTON) != 0;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        x[d] = A.x[(size_t)d * A.cap + i];
                        xi[d] = A.xi[(size_t)d * A.cap + i];
                        vext[d] = qu.park_vext[(size_t)d * A.cap + i];
                    }
                    load_vertices<DIM>(m, c, P);
                    have = true;
                }
            }
        }
        if (__ballot(have) == 0) break;
        bool done = false;
        if constexpr (RESORT) {
            // re-sort after a mesh move: the own-cell test left by k_resort,
            // then the relocation; no advection
            if (have) {
                double r[DIM];
                bool own = false;
                if (newton) {
                    newton = false;
                    own = !clearly_outside<DIM>(P, x, m.ftol[c]) && inverse_newton<DIM>(P, x, r) &&
                          inside_unit<DIM>(r);
                }
                if (!own) {
                    double Q[NV][DIM];
                    ++searches;
                    NoPrefetch pre;
                    c = relocate<DIM>(m, c, P, xi, vext, x, r, Q, pre);
                    if (c < 0) ++lost;
                }
                if (c >= 0)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = r[d];
                done = true;
            }
        } else if (DIM == 2 && have) {
            // a particle k_move flagged has its own-cell test at x pending:
            // clearly outside, it goes to the search; otherwise the test
            // (inverse_map from its closed form, as in advance64) runs first
            // in advance64, which goes on in the cell if x is inside
            bool own = false;
            if (newton) {
                newton = false;
                own = !clearly_outside<DIM>(P, x, m.ftol[c]);
            }
            double r[DIM];
            PrefetchV64<DIM, DEG> pre{m.cellV};
            int cn = c;
            if (own) {
                pre.load(c);
            } else {
                double Q[NV][DIM];
                ++searches;
                cn = relocate<DIM, PrefetchV64<DIM, DEG>, true>(m, c, P, xi, vext, x, r, Q, pre);
                if (cn >= 0)
#pragma unroll
                    for (int k = 0; k < NV; ++k)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) P[k][d] = Q[k][d];
            }
            if (cn < 0) {
                c = -1;
                ++lost;
                done = true;
            } else {
                if (!own) {
                    c = cn;
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = r[d];
                    ++s;
                }
                done = advance64<DIM, DEG>(m, V, pp, A, i, c, x, xi, vext, s, P, pre.Vc, m.ftol[c], own) == 0;
            }
        } else if (have) {
            // 3-D (k_move finishes its own-cell tests)
            double Q[NV][DIM], r[DIM];
            ++searches;
            PrefetchV64<DIM, DEG> pre{m.cellV};
            // (the face-neighbour shortcut's registers push this kernel into heavy spilling)
            const int cn = relocate<DIM, PrefetchV64<DIM, DEG>, false>(m, c, P, xi, vext, x, r, Q, pre);
            if (cn < 0) {
                c = -1;
                ++lost;
                done = true;
            } else {
                c = cn;
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[d] = r[d];
                ++s;
#pragma unroll
                for (int k = 0; k < NV; ++k)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) P[k][d] = Q[k][d];
                done = advance64<DIM, DEG>(m, V, pp, A, i, c, x, xi, vext, s, P, pre.Vc, m.ftol[c]) == 0;
            }
        }
        finish_particle<DIM>(pp, A, w, done, i, c, x, xi, true, seg_end);
        if (done) have = false;
    }
    wave_stat(w.stats, 0, lost);
    wave_stat(w.stats, 3, searches);
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 3 : 2))) k_search_resume(DevMesh<DIM> m, const double* __restrict__ V,
                                                       DevParticles A, DevParams pp, DevWork w, DevQueues qu,
                                                       int round) {
    search_body<DIM, DEG>(hw_grid(), m, V, A, pp, w, qu, round);
}

// the relocations of the re-sort after a mesh move (queue RESORT_QUEUE)
template <int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 3 : 2))) k_resort_search(DevMesh<DIM> m, DevParticles A, DevParams pp, DevWork w, DevQueues qu) {
    search_body<DIM, 1, true>(hw_grid(), m, nullptr, A, pp, w, qu, RESORT_QUEUE);
}

// candidate j of cell c: stayers of the old segment first, then the
// arrivals in the cell's slots (j < n_st + min(arrivals, aslots))
__device__ __forceinline__ int candidate(const DevWork& w, const int* old_off, int c, int n_st, int j) {
    return j < n_st ? old_off[c] + j : w.arr_slot[(size_t)c * w.aslots + (j - n_st)];
}

// candidates beyond the slots: the cell's overflow list, entry t handled by
// lane t % stride
template <class F>
__device__ __forceinline__ void for_overflow(const DevWork& w, int c, int lane, int stride, F f) {
    int t = 0;
    for (int e = w.ovf_head[c]; e >= 0; ++t) {
        const int2 v = w.ovf[e];
        if (t % stride == lane) f(t, v.x);
        e = v.y;
    }
}


// ---------------------------------------------------------------------------
// Cell kernels.  A cell is handled by a group of GS = 16 lanes, so a
// wavefront works on four cells at once (the per-cell work is a short chain
// of dependent loads; several cells per wavefront hide that latency).
constexpr int GS = 16;
constexpr int GROUPS = 256 / GS;     // groups per 256-thread block
constexpr int GCAP = 64;             // candidates staged in LDS per group (small lattices, census_small_row)
constexpr int GCAP_L = 128;          // the same for larger lattices
// the census row of a lattice: GCAP for up to 16 parts, or when the
// survivors of a full cell (MAX per part) leave 8 of GCAP's entries for
// arrivals; else GCAP_L.  A cell whose candidates do not fit goes to
// the second pass in rows of GCAP_L (and k_cell_big) either way; on the 5x5
// lattice with MAX 2 of the moving-mesh workload the smaller rows (LDS for 8
// workgroups per CU instead of 4) take the census from 0.62 to 0.40 ms per
// step, the second pass and k_cell_big adding ~0.04 ms.
constexpr int SMALL_ROW_SURVIVORS = GCAP - 8;
__host__ __device__ inline bool census_small_row(int nparts, int max_per_part) {
    return nparts <= 16 || (long long)max_per_part * nparts <= SMALL_ROW_SURVIVORS;
}
constexpr int FCAP = 64;             // final particles per cell staged for the fused projection sums
constexpr int WU = 4;                // survivors per lane and pass in the regroup write
constexpr int NKMAX = 272;           // part keys counted per group in the census (prod(q_d + 1))

// put(particle, position) for each survivor of cell c as the census left
// them: listed in id order in the cell's slot row when they fit, else by
// the position the census stored per candidate (the old segment's stayers,
// the slots and the overflow list)
template <class F>
__device__ __forceinline__ void survivors_of(const DevWork& w, const int* old_off, int c, int n_st, int kin,
                                             int nsurv, int gl, F put) {
    if (nsurv <= w.aslots) {
        for (int t = gl; t < nsurv; t += GS) put(w.arr_slot[(size_t)c * w.aslots + t], t);
        return;
    }
    for (int j = gl; j < kin; j += GS) {
        const int idx = candidate(w, old_off, c, n_st, j);
        // arrivals are in this cell's slots because their destination is c
        if (j < n_st && w.dest[idx] != c) continue;
        const int pos = w.spos[idx];
        if (pos >= 0) put(idx, pos);
    }
    for_overflow(w, c, gl, GS, [&](int, int idx) {
        const int pos = w.spos[idx];
        if (pos >= 0) put(idx, pos);
    });
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int load_agent(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum / OR over a 16-lane group
__device__ __forceinline__ int group_sum(int v) {
#pragma unroll
    for (int off = 1; off < GS; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

__device__ __forceinline__ unsigned long long group_or(unsigned long long v) {
#pragma unroll
    for (int off = 1; off < GS; off <<= 1) {
        const unsigned lo = __shfl_xor((unsigned)(v & 0xffffffffu), off);
        const unsigned hi = __shfl_xor((unsigned)(v >> 32), off);
        v |= ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// k_cell_count: per-cell part census (check_cell_for_empty_parts,
// cylinder2D/pfem2particle.cpp:510-561).  Part index per direction is
// (unsigned)(xi_d / (1/q_d)) without clamping (xi = 1 gives the extra index
// q_d).  In every part the MAX particles with the smallest ids survive;
// every in-range part that holds no particle is reseeded.  With reseed = 0
// (relocation only) every particle that found a cell survives.
// A cell's candidates (stayers of its old segment, then arrivals) are staged
// in LDS when there are at most CAP of them (GCAP or GCAP_L), else listed
// for k_cell_big.
template <int DIM, int CAP, bool LIST = false>
__device__ __forceinline__ void cell_count_body(const Grid& G, int n_cells, const int* __restrict__ old_off, const uint32_t* __restrict__ pid, const DevParams& pp, const DevWork& w, int reseed) {
    // Staged rows, compacted to the cell's valid candidates (stayers whose
    // destination is c, then the arrivals): candidate index; part key << 32 |
    // id; id of a survivor, all ones otherwise.  Rows are padded to a
    // multiple of 4 entries with fill values (key all ones, id all ones), so
    // the scans below read 16 bytes at a time, and their strides put the four
    // groups of a wavefront on different LDS banks.
    __shared__ int s_idx[GROUPS][CAP + 4];
    __shared__ __align__(16) unsigned long long s_kid[GROUPS][CAP + 2];
    __shared__ __align__(16) uint32_t s_sid[GROUPS][CAP + 4];
    // per group: the cell's particles per part key, one byte each (four to
    // a word, counted with word atomics; counts stay below CAP <= 128)
    // (rows of GCAP count up to 64 keys: 2-D lattices up to 7 x 7, 3-D up to
    // 3 x 3 x 3; the smaller LDS lets 8 instead of 7 workgroups share a CU)
    constexpr int NKC = CAP == GCAP ? 64 : NKMAX;
    __shared__ unsigned int s_cnt[GROUPS][NKC / 4];
    if (state_dead(w.counters)) return;
    const int g = threadIdx.x / GS, gl = threadIdx.x % GS;
    const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
    int nk = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) nk *= pp.q[d] + 1;
    const bool by_count = nk <= NKC;
    for (int t = gl; t < NKC / 4; t += GS) s_cnt[g][t] = 0;
    int deleted = 0, arrivals = 0;
    // LIST: the cells that the first pass listed (w.big), in rows of CAP;
    // the cells that do not fit either are listed again in w.big2
    const int ncl = LIST ? w.big[0] : n_cells;
    for (int cb = G.bid * GROUPS; cb < ncl; cb += G.nblk * GROUPS) {
        const bool cv = cb + g < ncl;
        const int c = LIST ? (cv ? w.big[1 + cb + g] : 0) : cb + g;
        const int n_st = cv ? old_off[c + 1] - old_off[c] : 0;
        const int na = cv ? w.arr_cnt[c] : 0;
        const int k = n_st + na;
        const bool small = k <= CAP && na <= w.aslots;
        if (!LIST && cv && gl == 0) arrivals += na;
        unsigned long long hit = 0;
        // nv: valid candidates, nst of them stayers (the first nst staged
        // entries, in ascending id order like the cell's segment); the same
        // in all lanes of the group
        int nv = 0, nst = 0, nsurv = 0;
        // the group's lanes in its wavefront's ballot start at bit sh
        const int sh = (threadIdx.x & (WAVE - 1)) & ~(GS - 1);
        if (small) {
            // all loads of the row first (up to CAP / GS candidates per lane),
            // then the compaction: the candidates' dependent loads overlap
            constexpr int SU = CAP / GS;
            int idxs[SU];
            unsigned long long kids[SU];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const int j = u * GS + gl;
                idxs[u] = -1;
                kids[u] = ~0ull;
                if (j < k) {
                    const int cidx = candidate(w, old_off, c, n_st, j);
                    const int dc = j < n_st ? w.dest[cidx] : c;
                    const unsigned long long kid = ((unsigned long long)(uint32_t)w.pkey[cidx] << 32) | pid[cidx];
                    // arrivals are listed for c because their destination is c
                    if (dc == c) {
                        idxs[u] = cidx;
                        kids[u] = kid;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                if (u * GS >= k) break;
                const int j = u * GS + gl;
                const int idx = idxs[u];
                const unsigned long long kid = kids[u];
                const unsigned mask = (unsigned)(__ballot(idx >= 0) >> sh) & 0xffffu;
                const unsigned smask = (unsigned)(__ballot(idx >= 0 && j < n_st) >> sh) & 0xffffu;
                if (idx >= 0) {
                    const int at = nv + __popc(mask & ((1u << gl) - 1u));
                    s_idx[g][at] = idx;
                    s_kid[g][at] = kid;
                    s_sid[g][at] = ~0u;
                }
                nv += __popc(mask);
                nst += __popc(smask);
            }
            if (gl < ((nv + 3) & ~3) - nv) {
                s_kid[g][nv + gl] = ~0ull;
                s_sid[g][nv + gl] = ~0u;
            }
        }
        const int nvr = (nv + 3) & ~3;
        wave_sync();
        // rank = particles of the same part with a smaller id, needed only
        // in a part that holds more than MAX particles: in rows of more than
        // 32 candidates the parts' counts first, then a scan of the row for
        // the particles of such parts (in shorter rows a scan for every
        // candidate costs less than the counting)
        const bool counted = small && reseed && by_count && nv > 32;
        if (counted) {
            for (int j = gl; j < nv; j += GS) {
                const uint32_t key = (uint32_t)(s_kid[g][j] >> 32) & 0xffffu;
                atomicAdd(&s_cnt[g][key >> 2], 1u << (8 * (key & 3)));
            }
            wave_sync();
        }
        if (small)
            for (int j = gl; j < nv; j += GS) {
                const unsigned long long me = s_kid[g][j];
                const uint32_t mk = (uint32_t)(me >> 32), mi = (uint32_t)me;
                const int o = part_ordinal((int)mk);
                if (o >= 0) hit |= 1ull << o;
                // rank (fill entries carry no valid key)
                int rank = 0;
                const uint32_t key = mk & 0xffffu;
                const unsigned cnt = counted ? (s_cnt[g][key >> 2] >> (8 * (key & 3))) & 0xffu : 0xffu;
                if (reseed && cnt > (unsigned)pp.max_per_part) {
                    for (int t = 0; t < nvr; t += 2) {
                        const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(&s_kid[g][t]);
                        rank += ((uint32_t)(v.x >> 32) == mk && (uint32_t)v.x < mi) ? 1 : 0;
                        rank += ((uint32_t)(v.y >> 32) == mk && (uint32_t)v.y < mi) ? 1 : 0;
                    }
                }
                const bool surv = rank < pp.max_per_part || !reseed;
                nsurv += surv ? 1 : 0;
                if (surv) s_sid[g][j] = mi;
            }
        else if (cv && gl == 0) {
            // more than CAP candidates: census by k_cell_big (or by the
            // second pass with larger rows)
            int* list = LIST ? w.big2 : w.big;
            const int slot = atomicAdd(&list[0], 1);
            list[1 + slot] = c;
        }
        wave_sync();
        if (counted)   // back to zero for the next cell (the words hold only this cell's parts)
            for (int j = gl; j < nv; j += GS) s_cnt[g][((uint32_t)(s_kid[g][j] >> 32) & 0xffffu) >> 2] = 0u;
        nsurv = group_sum(nsurv);
        // positions of the survivors in ascending id order (ids are below
        // 2^32 - 1, the fill value): the survivors in that order replace the
        // cell's arrival slots (already staged) when they fit, else each
        // candidate gets its position (-1: dropped)
        const bool listed = nsurv <= w.aslots;
        // positions: with few arrivals, a stayer's is the surviving stayers
        // before it plus the surviving arrivals with a smaller id; otherwise
        // (and for the arrivals) a scan of all entries.  A pending layout
        // needs no positions, only its survivors marked.
        const bool pos_needed = small && !w.pend_mark;
        if (small && w.pend_mark)
            for (int j = gl; j < nv; j += GS)
                if (s_sid[g][j] != ~0u) w.dest[s_idx[g][j]] = c | DEST_KEPT;
        const int jp = pos_needed && nv - nst <= 4 ? nst : 0;
        if (pos_needed && jp > 0) {
            int sbase = 0;
            for (int j0 = 0; j0 < jp; j0 += GS) {
                const int j = j0 + gl;
                const bool act = j < jp;
                const uint32_t id = act ? s_sid[g][j] : ~0u;
                const bool sv = act && id != ~0u;
                const unsigned mask = (unsigned)(__ballot(sv) >> sh) & 0xffffu;
                if (sv) {
                    int pos = sbase + __popc(mask & ((1u << gl) - 1u));
                    for (int t = nst; t < nv; ++t) pos += s_sid[g][t] < id ? 1 : 0;
                    const int idx = s_idx[g][j];
                    if (listed) w.arr_slot[(size_t)c * w.aslots + pos] = idx;
                    else w.spos[idx] = pos;
                } else if (act && !listed) {
                    w.spos[s_idx[g][j]] = -1;
                }
                sbase += __popc(mask);
            }
        }
        if (pos_needed)
            for (int j = jp + gl; j < nv; j += GS) {
                const int idx = s_idx[g][j];
                const uint32_t id = s_sid[g][j];
                if (id == ~0u) {
                    if (!listed) w.spos[idx] = -1;
                    continue;
                }
                int pos = 0;
                for (int t = 0; t < nvr; t += 4) {
                    const uint4 v = *reinterpret_cast<const uint4*>(&s_sid[g][t]);
                    pos += (v.x < id ? 1 : 0) + (v.y < id ? 1 : 0) + (v.z < id ? 1 : 0) + (v.w < id ? 1 : 0);
                }
                if (listed) w.arr_slot[(size_t)c * w.aslots + pos] = idx;
                else w.spos[idx] = pos;
            }
        wave_sync();
        hit = group_or(hit);
        if (cv && small && gl == 0) {
            const int spawns = reseed ? __popcll(all & ~hit) : 0;
            w.hitmask[c] = reseed ? hit : all;
            w.nsurv[c] = nsurv;
            w.fs_cnt[c] = ((unsigned long long)(nsurv + spawns) << 32) | (unsigned)spawns;
            deleted += nv - nsurv;
        }
    }
    deleted = wave_sum(deleted);
    if ((threadIdx.x & (WAVE - 1)) == 0 && deleted) stat_add(w.stats, 1, (unsigned long long)deleted);
    if (!LIST) wave_stat(w.stats, 5, arrivals);
}

template <int DIM, int CAP, bool LIST = false>
__global__ void __launch_bounds__(256) k_cell_count(int n_cells, const int* __restrict__ old_off,
                                                    const uint32_t* __restrict__ pid, DevParams pp, DevWork w,
                                                    int reseed) {
    cell_count_body<DIM, CAP, LIST>(hw_grid(), n_cells, old_off, pid, pp, w, reseed);
}

// The census of k_cell_count for the cells it listed (more than CAP
// candidates): one 256-thread block per cell, candidates staged in LDS.
// A cell with more than BIGCAP candidates sets overflow flag 2 and keeps no
// particle.
constexpr int BIGCAP = 2048;

template <int DIM>
__device__ __forceinline__ void cell_big_body(const Grid& G, const int* __restrict__ old_off, const uint32_t* __restrict__ pid, const DevParams& pp, const DevWork& w, int reseed, const int* __restrict__ list) {
    __shared__ int s_idx[BIGCAP];
    __shared__ __align__(16) unsigned long long s_kid[BIGCAP];
    __shared__ __align__(16) uint32_t s_sid[BIGCAP];
    __shared__ unsigned long long s_hit;
    __shared__ int s_nvalid, s_nsurv;
    if (state_dead(w.counters)) return;
    const int tid = threadIdx.x;
    const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
    const int nbig = list[0];
    int deleted = 0;
    for (int b = G.bid; b < nbig; b += G.nblk) {
        const int c = list[1 + b];
        const int n_st = old_off[c + 1] - old_off[c];
        const int na = w.arr_cnt[c];
        const int k = n_st + na;
        const int kin = n_st + (na < w.aslots ? na : w.aslots);   // candidates in the old segment and slots
        if (k > BIGCAP) {
            for (int j = tid; j < kin; j += blockDim.x) w.spos[candidate(w, old_off, c, n_st, j)] = -1;
            for_overflow(w, c, tid, blockDim.x, [&](int, int idx) { w.spos[idx] = -1; });
            if (tid == 0) {
                atomicOr(&w.counters[2], 2);
                w.hitmask[c] = all;
                w.nsurv[c] = 0;
                w.fs_cnt[c] = 0;
            }
            continue;
        }
        const int kr = (k + 3) & ~3;
        if (tid == 0) {
            s_hit = 0;
            s_nvalid = 0;
            s_nsurv = 0;
        }
        for (int j = tid; j < kr; j += blockDim.x) {
            if (j >= kin && j < k) continue;   // overflow entries, staged below
            int idx = -1;
            unsigned long long kid = ~0ull;
            if (j < k) {
                const int cidx = candidate(w, old_off, c, n_st, j);
                if (j >= n_st || w.dest[cidx] == c) {
                    idx = cidx;
                    kid = ((unsigned long long)(uint32_t)w.pkey[cidx] << 32) | pid[cidx];
                }
            }
            s_idx[j] = idx;
            s_kid[j] = kid;
            s_sid[j] = ~0u;
        }
        for_overflow(w, c, tid, blockDim.x, [&](int t, int cidx) {
            s_idx[kin + t] = cidx;
            s_kid[kin + t] = ((unsigned long long)(uint32_t)w.pkey[cidx] << 32) | pid[cidx];
            s_sid[kin + t] = ~0u;
        });
        __syncthreads();
        unsigned long long hit = 0;
        int nvalid = 0, nsurv = 0;
        for (int j = tid; j < k; j += blockDim.x) {
            if (s_idx[j] < 0) continue;
            ++nvalid;
            const unsigned long long me = s_kid[j];
            const int o = part_ordinal((int)(me >> 32));
            if (o >= 0) hit |= 1ull << o;
            int rank = 0;
            if (reseed) {
                const unsigned long long lo = me & 0xffffffff00000000ull;
                for (int t = 0; t < kr; t += 2) {
                    const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(&s_kid[t]);
                    rank += (v.x >= lo && v.x < me) ? 1 : 0;
                    rank += (v.y >= lo && v.y < me) ? 1 : 0;
                }
            }
            const bool surv = rank < pp.max_per_part || !reseed;
            nsurv += surv ? 1 : 0;
            if (surv) s_sid[j] = (uint32_t)me;
        }
        if (hit) atomicOr(&s_hit, hit);
        if (nvalid) atomicAdd(&s_nvalid, nvalid);
        if (nsurv) atomicAdd(&s_nsurv, nsurv);
        __syncthreads();
        // survivors listed in the cell's (staged) arrival slots when they fit
        // (a pending layout: only marked)
        const bool listed = s_nsurv <= w.aslots;
        for (int j = tid; j < k; j += blockDim.x) {
            const int idx = s_idx[j];
            if (idx < 0) continue;
            const uint32_t id = s_sid[j];
            if (w.pend_mark) {
                if (id != ~0u) w.dest[idx] = c | DEST_KEPT;
                continue;
            }
            if (id == ~0u) {
                if (!listed) w.spos[idx] = -1;
                continue;
            }
            int pos = 0;
            for (int t = 0; t < kr; t += 4) {
                const uint4 v = *reinterpret_cast<const uint4*>(&s_sid[t]);
                pos += (v.x < id ? 1 : 0) + (v.y < id ? 1 : 0) + (v.z < id ? 1 : 0) + (v.w < id ? 1 : 0);
            }
            if (listed) w.arr_slot[(size_t)c * w.aslots + pos] = idx;
            else w.spos[idx] = pos;
        }
        if (tid == 0) {
            const unsigned long long hm = s_hit;
            const int spawns = reseed ? __popcll(all & ~hm) : 0;
            w.hitmask[c] = reseed ? hm : all;
            w.nsurv[c] = s_nsurv;
            w.fs_cnt[c] = ((unsigned long long)(s_nsurv + spawns) << 32) | (unsigned)spawns;
            deleted += s_nvalid - s_nsurv;
        }
        __syncthreads();
    }
    if (tid == 0 && deleted) stat_add(w.stats, 1, (unsigned long long)deleted);
}

// second: the cells listed by the second census pass (w.big2)
template <int DIM>
__global__ void __launch_bounds__(256) k_cell_big(const int* __restrict__ old_off, const uint32_t* __restrict__ pid,
                                                  DevParams pp, DevWork w, int reseed, int second) {
    cell_big_body<DIM>(hw_grid(), old_off, pid, pp, w, reseed, second ? w.big2 : w.big);
}

// ---------------------------------------------------------------------------
// k_cell_write: survivors and reseeded particles into buffer B.
// New particles: reference point at the part centre, location mapped,
// velocity interpolated from V (cylinder2D/pfem2particle.cpp:533-549), ids
// continuing the counter in ascending cell, then part, order.
// FP64 projection sums of one cell over n staged particles in order, the
// expressions of k_cell_partials (accumulator a: vertex a / (DIM + 1),
// component a % (DIM + 1), the last one the weight); NA accumulators a0,
// a0 + GS, ... in one pass (their chains are independent, so they overlap)
template <int DIM, int NA, int DEG = 1>
__device__ __forceinline__ void cell_sums64(const double* __restrict__ sxi, const double* __restrict__ sv, int n,
                                            int a0, double* out) {
    constexpr int NACC = QF<DIM, DEG>::ND * (DIM + 1);
    ShapeOne<DIM, DEG> sh[NA];
    double acc[NA];
    int comp[NA];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = a0 + u * GS;
        const int kk = (a < NACC ? a : 0) / (DIM + 1);
        comp[u] = a - kk * (DIM + 1);
        sh[u].init(kk);
        acc[u] = 0.0;
    }
    for (int q = 0; q < n; ++q) {
        double xi[DIM], v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xi[d] = sxi[q * DIM + d];
            v[d] = sv[q * DIM + d];
        }
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const double nk = sh[u](xi);
            if (comp[u] < DIM) {
                double vc = v[0];
#pragma unroll
                for (int d = 1; d < DIM; ++d)
                    if (d == comp[u]) vc = v[d];
                acc[u] += nk * vc;
            } else {
                acc[u] += nk;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < NA; ++u) out[u] = acc[u];
}

template <int DIM, bool GATHER, bool FUSE, int DEG = 1>
__device__ __forceinline__ void cell_write_one(const DevMesh<DIM>& m, const double* __restrict__ V,
                                               const int* __restrict__ old_off, const DevParticles& A,
                                               const DevParticles& B, const DevParams& pp, const DevWork& w, int c,
                                               int gl, double* __restrict__ sxi, double* __restrict__ sv,
                                               double* __restrict__ partial) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    const int n_st = old_off[c + 1] - old_off[c];
    const int na = w.arr_cnt[c];
    const int kin = n_st + (na < w.aslots ? na : w.aslots);
    const unsigned long long fo = w.fs_off[c];
    const long long out0 = (long long)(fo >> 32);
    const int nsurv = w.nsurv[c];
    const int nfinal = (int)(w.fs_cnt[c] >> 32);
    if (gl == 0) w.new_off[c] = (int)out0;
    if (out0 + nfinal > B.cap) {
        if (gl == 0) atomicOr(&w.counters[2], 1);
        return;
    }
    auto put = [&](int idx, int pos) {
        const long long o = out0 + pos;
        B.cell[o] = c;
        B.id[o] = A.id[idx];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double xi = A.xi[(size_t)d * A.cap + idx], v = A.v[(size_t)d * A.cap + idx];
            B.x[(size_t)d * B.cap + o] = A.x[(size_t)d * A.cap + idx];
            B.xi[(size_t)d * B.cap + o] = xi;
            B.v[(size_t)d * B.cap + o] = v;
            if constexpr (FUSE) {
                sxi[pos * DIM + d] = xi;
                sv[pos * DIM + d] = v;
            }
        }
        if (pp.keep_vext)
#pragma unroll
            for (int d = 0; d < DIM; ++d) B.vext[(size_t)d * B.cap + o] = A.vext[(size_t)d * A.cap + idx];
    };
    if (nsurv <= w.aslots) {
        // survivors listed in the slot row: WU of a lane's survivors per
        // pass, all their loads issued before any store, so that a lane has
        // several independent loads in flight
        for (int t0 = 0; t0 < nsurv; t0 += GS * WU) {
            int src[WU];
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                const int t = t0 + gl + u * GS;
                src[u] = t < nsurv ? w.arr_slot[(size_t)c * w.aslots + t] : -1;
            }
            uint32_t pid[WU];
            double px[WU][DIM], pxi[WU][DIM], pv[WU][DIM], pe[WU][DIM];
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                if (src[u] < 0) continue;
                const int idx = src[u];
                pid[u] = A.id[idx];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    px[u][d] = A.x[(size_t)d * A.cap + idx];
                    pxi[u][d] = A.xi[(size_t)d * A.cap + idx];
                    pv[u][d] = A.v[(size_t)d * A.cap + idx];
                    pe[u][d] = pp.keep_vext ? A.vext[(size_t)d * A.cap + idx] : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < WU; ++u) {
                if (src[u] < 0) continue;
                const int pos = t0 + gl + u * GS;
                const long long o = out0 + pos;
                B.cell[o] = c;
                B.id[o] = pid[u];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    B.x[(size_t)d * B.cap + o] = px[u][d];
                    B.xi[(size_t)d * B.cap + o] = pxi[u][d];
                    B.v[(size_t)d * B.cap + o] = pv[u][d];
                    if constexpr (FUSE) {
                        sxi[pos * DIM + d] = pxi[u][d];
                        sv[pos * DIM + d] = pv[u][d];
                    }
                }
                if (pp.keep_vext)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) B.vext[(size_t)d * B.cap + o] = pe[u][d];
            }
        }
    } else {
        survivors_of(w, old_off, c, n_st, kin, nsurv, gl, put);
    }
    const int spawns = nfinal - nsurv;
    if (spawns > 0) {
        const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
        const unsigned long long empty = all & ~w.hitmask[c];
        const uint32_t id0 = *w.particle_counter + (uint32_t)fo;
        for (int o = gl; o < pp.nparts_in; o += GS) {
            if (!((empty >> o) & 1)) continue;
            const int r = __popcll(empty & ((1ull << o) - 1));
            double xi[DIM], P[NV][DIM], Vc[ND][DIM], x[DIM], v[DIM];
            lattice_point<DIM>(pp, o, xi);
            load_cell<DIM, GATHER, DEG>(m, c, P, Vc);
            forward<DIM>(P, xi, x);
            interp<DIM, DEG>(Vc, xi, v);
            const long long q = out0 + nsurv + r;
            B.cell[q] = c;
            B.id[q] = id0 + (uint32_t)r + 1u;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                B.x[(size_t)d * B.cap + q] = x[d];
                B.xi[(size_t)d * B.cap + q] = xi[d];
                B.v[(size_t)d * B.cap + q] = v[d];
                if constexpr (FUSE) {
                    sxi[(nsurv + r) * DIM + d] = xi[d];
                    sv[(nsurv + r) * DIM + d] = v[d];
                }
            }
            if (pp.keep_vext)
#pragma unroll
                for (int d = 0; d < DIM; ++d) B.vext[(size_t)d * B.cap + q] = 0.0;
        }
    }
    if (gl == 0) {
        w.arr_cnt[c] = 0;
        w.ovf_head[c] = -1;
    }
    if constexpr (FUSE) {
        constexpr int NACC = ND * (DIM + 1);
        wave_sync();
        // sums per lane: Q1 1 in 2-D (12 lanes), 2 in 3-D; Q2 2 in 2-D, 7 in 3-D
        constexpr int NA = (NACC + GS - 1) / GS;
        if (gl < NACC) {
            double out[NA];
            cell_sums64<DIM, NA, DEG>(sxi, sv, nfinal, gl, out);
#pragma unroll
            for (int u = 0; u < NA; ++u)
                if (gl + u * GS < NACC) partial[(size_t)c * NACC + gl + u * GS] = out[u];
        }
        wave_sync();
    }
}

// GATHER: an FP32-mixed handle's FP64 regroup (no FP64 per-cell velocity copy).
// FUSE (final count per cell <= FCAP): the group stages the cell's final
// particles in LDS and writes its projection sums, as k_cell_partials does.
template <int DIM, bool GATHER, bool FUSE, int DEG = 1>
__device__ __forceinline__ void cell_write_body(const Grid& G, const DevMesh<DIM>& m, const double* __restrict__ V, const int* __restrict__ old_off, const DevParticles& A, const DevParticles& B, const DevParams& pp, const DevWork& w, double* __restrict__ partial) {
    __shared__ double s_xi[FUSE ? GROUPS : 1][FUSE ? FCAP * DIM : 1];
    __shared__ double s_v[FUSE ? GROUPS : 1][FUSE ? FCAP * DIM : 1];
    if (state_dead(w.counters)) return;
    const int g = threadIdx.x / GS, gl = threadIdx.x % GS;
    for (int c = G.bid * GROUPS + g; c < m.n_cells; c += G.nblk * GROUPS)
        cell_write_one<DIM, GATHER, FUSE, DEG>(m, V, old_off, A, B, pp, w, c, gl, s_xi[FUSE ? g : 0],
                                               s_v[FUSE ? g : 0], partial);
}

template <int DIM, bool GATHER, bool FUSE, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_write(DevMesh<DIM> m, const double* __restrict__ V,
                                                    const int* __restrict__ old_off, DevParticles A,
                                                    DevParticles B, DevParams pp, DevWork w,
                                                    double* __restrict__ partial) {
    cell_write_body<DIM, GATHER, FUSE, DEG>(hw_grid(), m, V, old_off, A, B, pp, w, partial);
}

// ---------------------------------------------------------------------------
// The regroup write in two passes (regroups without fused projection sums).
// A cell's output segment starts at an arbitrary position, so the stores of
// k_cell_write's 16-lane groups never cover a whole cache line and L2 fills
// every line it writes from memory first (the write volume again in reads).
// k_cell_map does the per-cell work (offsets, survivors, spawn slots,
// resets) and writes one int per output position: the source particle, or
// for a new particle -1 - (cell * 64 + lattice ordinal).  k_cell_gather then
// walks the output positions in order and copies or creates each particle,
// so every wavefront stores whole, consecutive lines.  Cell indices below
// 2^25 (the map's encoding), which the host checks.
template <int DIM>
__device__ __forceinline__ void cell_map_body(const Grid& G, int n_cells, const int* __restrict__ old_off,
                                              long long cap, const DevParams& pp, const DevWork& w,
                                              int* __restrict__ omap) {
    if (state_dead(w.counters)) return;
    const int g = threadIdx.x / GS, gl = threadIdx.x % GS;
    for (int c = G.bid * GROUPS + g; c < n_cells; c += G.nblk * GROUPS) {
        const int n_st = old_off[c + 1] - old_off[c];
        const int na = w.arr_cnt[c];
        const int kin = n_st + (na < w.aslots ? na : w.aslots);
        const unsigned long long fo = w.fs_off[c];
        const long long out0 = (long long)(fo >> 32);
        const int nsurv = w.nsurv[c];
        const int nfinal = (int)(w.fs_cnt[c] >> 32);
        if (gl == 0) w.new_off[c] = (int)out0;
        if (out0 + nfinal > cap) {
            if (gl == 0) atomicOr(&w.counters[2], 1);
            continue;
        }
        if (nsurv <= w.aslots) {
            for (int t = gl; t < nsurv; t += GS) omap[out0 + t] = w.arr_slot[(size_t)c * w.aslots + t];
        } else {
            survivors_of(w, old_off, c, n_st, kin, nsurv, gl, [&](int idx, int pos) { omap[out0 + pos] = idx; });
        }
        if (nfinal > nsurv) {
            const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
            const unsigned long long empty = all & ~w.hitmask[c];
            for (int o = gl; o < pp.nparts_in; o += GS) {
                if (!((empty >> o) & 1)) continue;
                const int r = __popcll(empty & ((1ull << o) - 1));
                omap[out0 + nsurv + r] = -1 - (c * 64 + o);
            }
        }
        if (gl == 0) {
            w.arr_cnt[c] = 0;
            w.ovf_head[c] = -1;
        }
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) k_cell_map(int n_cells, const int* __restrict__ old_off, long long cap,
                                                  DevParams pp, DevWork w, int* __restrict__ omap) {
    cell_map_body<DIM>(hw_grid(), n_cells, old_off, cap, pp, w, omap);
}

constexpr int GU = 4;   // output positions per thread and pass of the regroup gather

// a new particle at output position o: code = cell * 64 + lattice ordinal
template <int DIM, bool GATHER, int DEG>
__device__ __forceinline__ void gather_spawn(const DevMesh<DIM>& m, const DevParticles& B, const DevParams& pp,
                                             const DevWork& w, long long o, int code) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    const int c = code >> 6, lo = code & 63;
    const unsigned long long fo = w.fs_off[c];
    const int r = (int)(o - (long long)(fo >> 32)) - w.nsurv[c];
    double xi[DIM], P[NV][DIM], Vc[ND][DIM], x[DIM], v[DIM];
    lattice_point<DIM>(pp, lo, xi);
    load_cell<DIM, GATHER, DEG>(m, c, P, Vc);
    forward<DIM>(P, xi, x);
    interp<DIM, DEG>(Vc, xi, v);
    B.cell[o] = c;
    B.id[o] = *w.particle_counter + (uint32_t)fo + (uint32_t)r + 1u;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        B.x[(size_t)d * B.cap + o] = x[d];
        B.xi[(size_t)d * B.cap + o] = xi[d];
        B.v[(size_t)d * B.cap + o] = v[d];
        if (pp.keep_vext) B.vext[(size_t)d * B.cap + o] = 0.0;
    }
}

// one output position per lane and slice, GU slices of the workgroup's
// positions per pass, all loads of a pass issued before its stores: a
// survivor's record is copied (its cell is its destination), a new particle
// is created at its lattice point as k_cell_write creates it (id after the
// cell's spawn offset, velocity interpolated from the cell's DoFs,
// velocity_ext zero)
template <int DIM, bool GATHER, int DEG = 1>
__device__ __forceinline__ void cell_gather_body(const Grid& G, const DevMesh<DIM>& m, const DevParticles& A,
                                                 const DevParticles& B, const DevParams& pp, const DevWork& w,
                                                 const int* __restrict__ omap) {
    if (state_dead(w.counters)) return;
    const long long total = (long long)(w.fs_off[m.n_cells] >> 32);
    const long long n = total < B.cap ? total : B.cap;
    const long long bd = blockDim.x, span = (long long)G.nblk * bd * GU;
    for (long long o0 = (long long)G.bid * bd * GU + threadIdx.x; o0 < n; o0 += span) {
        int src[GU];
        bool act[GU];
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            const long long o = o0 + u * bd;
            act[u] = o < n;
            src[u] = act[u] ? omap[o] : -1;
        }
        int pc[GU];
        uint32_t pid[GU];
        double px[GU][DIM], pxi[GU][DIM], pv[GU][DIM], pe[GU][DIM];
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            if (!act[u] || src[u] < 0) continue;
            const int s = src[u];
            pc[u] = w.dest[s];
            pid[u] = A.id[s];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                px[u][d] = A.x[(size_t)d * A.cap + s];
                pxi[u][d] = A.xi[(size_t)d * A.cap + s];
                pv[u][d] = A.v[(size_t)d * A.cap + s];
                pe[u][d] = pp.keep_vext ? A.vext[(size_t)d * A.cap + s] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            if (!act[u]) continue;
            const long long o = o0 + u * bd;
            if (src[u] < 0) {
                gather_spawn<DIM, GATHER, DEG>(m, B, pp, w, o, -1 - src[u]);
                continue;
            }
            B.cell[o] = pc[u];
            B.id[o] = pid[u];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                B.x[(size_t)d * B.cap + o] = px[u][d];
                B.xi[(size_t)d * B.cap + o] = pxi[u][d];
                B.v[(size_t)d * B.cap + o] = pv[u][d];
                if (pp.keep_vext) B.vext[(size_t)d * B.cap + o] = pe[u][d];
            }
        }
    }
}

template <int DIM, bool GATHER, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_gather(DevMesh<DIM> m, DevParticles A, DevParticles B, DevParams pp,
                                                     DevWork w, const int* __restrict__ omap) {
    cell_gather_body<DIM, GATHER, DEG>(hw_grid(), m, A, B, pp, w, omap);
}

// totals after the scans: particle count, id counter, statistics; resets the
// queue counters for the next step (one workgroup of QSHARDS threads: k_finalize,
// or the last workgroup of k_node in a fused step)
__device__ __forceinline__ void finalize_block(int n_cells, const DevWork& w, int* queue_count,
                                               bool keep_pending = false) {
    static_assert(QSHARDS == STAT_SHARDS, "one thread per stat row");
    __shared__ int s_arr;
    const int t = threadIdx.x;
    for (int r = 0; r < NQROUNDS; ++r) {
        queue_count[((size_t)r * (QSHARDS + 1) + t) * QSTRIDE] = 0;
        if (t == 0) queue_count[((size_t)r * (QSHARDS + 1) + QSHARDS) * QSTRIDE] = 0;
    }
    // arrivals of this regroup (diagnostics): the stat rows' slot 5, cleared
    if (t == 0) s_arr = 0;
    __syncthreads();
    const int a = (int)w.stats[(size_t)t * NSTAT + 5];
    w.stats[(size_t)t * NSTAT + 5] = 0;
    if (a) atomicAdd(&s_arr, a);
    __syncthreads();
    if (t != 0) return;
    w.counters[1] = s_arr;
    const unsigned long long fo = w.fs_off[n_cells];
    // a dead state publishes no particles (its total may exceed the capacity)
    const bool dead = state_dead(w.counters);
    const int total = dead ? 0 : (int)(fo >> 32);
    const int spawned = dead ? 0 : (int)(uint32_t)fo;
    w.new_off[n_cells] = total;
    w.stats[2] += (unsigned long long)spawned;
    *w.particle_counter += (uint32_t)spawned;
    w.counters[3] = w.counters[0];   // count before the step (diagnostics)
    w.counters[0] = total;
    w.counters[4] = 0;
    // a regroup consumes a pending layout; k_cell_pend has just set one up
    if (!keep_pending) w.counters[6] = 0;
    w.big[0] = 0;
    w.big2[0] = 0;
}

__global__ void __launch_bounds__(QSHARDS) k_finalize(int n_cells, DevWork w, int* queue_count, int keep_pending) {
    finalize_block(n_cells, w, queue_count, keep_pending != 0);
}

// ---------------------------------------------------------------------------
// Pending layout.  After the census of a re-sort (relocation + reseed), the
// particles are not moved into a new buffer: the census's survivors are
// marked in place (destination = cell | DEST_KEPT), its
// new particles are created after the current segments (position seg_end +
// the cell's spawn offset + rank, ids and values as k_cell_gather creates
// them), and counters[6], [7] record the slots and seg_end.  The next move
// takes a particle's cell from dest, skips the unmarked ones and makes the
// new ones arrive in their cell, so its census regroups the pending layout
// and the re-sort's regroup write is never done.  Any other use of the
// particles first settles the layout (k_settle + a regroup without reseed).
// (The census marks the survivors when DevWork::pend_mark is set.)
template <int DIM, bool GATHER, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_pend(DevMesh<DIM> m, const int* __restrict__ old_off, DevParticles A,
                                                   DevParams pp, DevWork w) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    if (state_dead(w.counters)) return;
    const int nc = m.n_cells;
    const int seg_end = old_off[nc];
    const unsigned long long ftot = w.fs_off[nc];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long long slots = (long long)seg_end + (uint32_t)ftot;
        if (slots > A.cap) atomicOr(&w.counters[2], 1);
        w.counters[6] = slots > A.cap ? 0 : (int)slots;
        w.counters[7] = seg_end;
    }
    // one cell per thread: new particles are few, and most cells only get
    // their resets
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if ((uint32_t)w.fs_cnt[c] != 0u) {   // spawns, in lattice order
            const unsigned long long fo = w.fs_off[c];
            const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
            const unsigned long long empty = all & ~w.hitmask[c];
            double P[NV][DIM], Vc[ND][DIM];
            load_cell<DIM, GATHER, DEG>(m, c, P, Vc);
            int r = 0;
            for (unsigned long long e = empty; e; e &= e - 1, ++r) {
                const int o = __ffsll((long long)e) - 1;
                const long long q = (long long)seg_end + (uint32_t)fo + r;
                if (q >= A.cap) continue;   // (flagged above)
                double xi[DIM], x[DIM], v[DIM];
                lattice_point<DIM>(pp, o, xi);
                forward<DIM>(P, xi, x);
                interp<DIM, DEG>(Vc, xi, v);
                A.cell[q] = c;
                A.id[q] = *w.particle_counter + (uint32_t)fo + (uint32_t)r + 1u;
                w.dest[q] = c | DEST_KEPT;
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    A.x[(size_t)d * A.cap + q] = x[d];
                    A.xi[(size_t)d * A.cap + q] = xi[d];
                    A.v[(size_t)d * A.cap + q] = v[d];
                    if (pp.keep_vext) A.vext[(size_t)d * A.cap + q] = 0.0;
                }
            }
        }
        w.arr_cnt[c] = 0;
        w.ovf_head[c] = -1;
    }
}

// settles a pending layout for a regroup without reseed: every kept or new
// particle keeps its cell (destination) and gets its part key back, and the
// ones outside their cell's segment arrive there; the others are gone
template <int DIM>
__global__ void k_settle(DevParticles A, DevParams pp, DevWork w) {
    const int pn = w.counters[6];
    const int seg_end = w.counters[7];
    const int lane = threadIdx.x & (WAVE - 1);
    for (int base = blockIdx.x * blockDim.x + (threadIdx.x - lane); base < pn; base += gridDim.x * blockDim.x) {
        const int i = base + lane;
        const bool valid = i < pn;
        int c = -1;
        if (valid) {
            const int dk = w.dest[i];
            c = dk >= 0 && (dk & DEST_KEPT) ? dk & ~DEST_KEPT : -1;
            w.dest[i] = c;
            if (c >= 0) {
                double xi[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[d] = A.xi[(size_t)d * A.cap + i];
                w.pkey[i] = part_key<DIM>(pp, xi);
            }
        }
        add_arrival(w, valid && c >= 0 && (c != A.cell[valid ? i : 0] || i >= seg_end), c, i);
    }
}

// ---------------------------------------------------------------------------
// k_cell_partials: per-cell local sums of the projection
// (TurekBenchmark/pfem2particle.cpp:880-893): for each cell vertex k,
// sum_p N_k(xi_p) v_p and sum_p N_k(xi_p), particles in ascending id order.
// A group of PGS lanes per cell stages PGS particles at a time in LDS; lane
// a = k*(DIM+1) + comp of the group then adds its terms in order.
// Q1: one lane per accumulator (12 or 32 lanes per cell); Q2: 27 (2-D, one
// lane each) or 108 accumulators (3-D: a group of 64 lanes, two each)
template <int DIM, int DEG = 1>
struct PartialsShape {
    static constexpr int NACC = QF<DIM, DEG>::ND * (DIM + 1);
    static constexpr int PGS = NACC <= 16 ? 16 : (NACC <= 32 ? 32 : 64);
    static constexpr int NA = (NACC + PGS - 1) / PGS;   // accumulators per lane
    static constexpr int NG = 256 / PGS;
};

template <int DIM, int DEG = 1>
__device__ __forceinline__ void partials_body(const Grid& G, int n_cells, const int* __restrict__ off, const DevParticles& A, double* __restrict__ partial, const int* __restrict__ counters) {
    if (state_dead(counters)) return;
    constexpr int NACC = PartialsShape<DIM, DEG>::NACC;
    constexpr int PGS = PartialsShape<DIM, DEG>::PGS;
    constexpr int NG = PartialsShape<DIM, DEG>::NG;
    constexpr int NA = PartialsShape<DIM, DEG>::NA;
    __shared__ double s_xi[NG][PGS][DIM];
    __shared__ double s_v[NG][PGS][DIM];
    const int g = threadIdx.x / PGS, gl = threadIdx.x % PGS;
    // lane gl holds accumulators gl, gl + PGS, ... (a = k * (DIM + 1) + comp)
    ShapeOne<DIM, DEG> sh[NA];
    int comp[NA];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = gl + u * PGS;
        const int kk = (a < NACC ? a : 0) / (DIM + 1);
        comp[u] = a - kk * (DIM + 1);
        sh[u].init(kk);
    }
    for (int cb = G.bid * NG; cb < n_cells; cb += G.nblk * NG) {
        const int c = cb + g;
        const bool cv = c < n_cells;
        const int b = cv ? off[c] : 0, e = cv ? off[c + 1] : 0;
        int trips = (e - b + PGS - 1) / PGS;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) trips = max(trips, __shfl_xor(trips, o));
        double acc[NA];
#pragma unroll
        for (int u = 0; u < NA; ++u) acc[u] = 0.0;
        for (int t = 0; t < trips; ++t) {
            const int j0 = b + t * PGS;
            const int j = j0 + gl;
            if (j < e) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    s_xi[g][gl][d] = A.xi[(size_t)d * A.cap + j];
                    s_v[g][gl][d] = A.v[(size_t)d * A.cap + j];
                }
            }
            wave_sync();
            const int cnt = (e - j0) < PGS ? (e - j0) : PGS;
            if (gl < NACC)
                for (int q = 0; q < cnt; ++q) {
                    // N_k of this lane's node alone: per direction the
                    // factor of the node's axis, multiplied in shape_q's
                    // order, so the value is shape_q's N[k] (up to the sign
                    // of a zero factor, which no sum here sees: acc starts
                    // at +0)
                    double xi[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = s_xi[g][q][d];
#pragma unroll
                    for (int u = 0; u < NA; ++u) {
                        const double nk = sh[u](xi);
                        if (comp[u] < DIM) {
                            double vc = s_v[g][q][0];
#pragma unroll
                            for (int d = 1; d < DIM; ++d)
                                if (d == comp[u]) vc = s_v[g][q][d];
                            acc[u] += nk * vc;
                        } else {
                            acc[u] += nk;
                        }
                    }
                }
            wave_sync();
        }
        if (cv)
#pragma unroll
            for (int u = 0; u < NA; ++u)
                if (gl + u * PGS < NACC) partial[(size_t)c * NACC + gl + u * PGS] = acc[u];
    }
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_partials(int n_cells, const int* __restrict__ off, DevParticles A,
                                                       double* __restrict__ partial, const int* __restrict__ counters) {
    partials_body<DIM, DEG>(hw_grid(), n_cells, off, A, partial, counters);
}

// k_node: node sums over the cells that contain the DoF (ascending cell
// order, TurekBenchmark/pfem2particle.cpp:895-909) divided by the weight.
// Writes V and V_old (both hold the projected field afterwards).
template <int DIM, bool FIN, int DEG = 1>
__device__ __forceinline__ void node_body(const Grid& G, const DevMesh<DIM>& m, const double* __restrict__ partial, double* __restrict__ V, double* __restrict__ Vold, unsigned long long* stats, const DevWork& w, int* queue_count) {
    constexpr int NV = Q1<DIM>::NV;   // cells per chunk
    constexpr int ND = QF<DIM, DEG>::ND;
    constexpr int NACC = ND * (DIM + 1);
    // FIN (a step whose regroup left its bookkeeping for later): the last
    // workgroup runs it; it depends on the cell write only, as the node sums do
    if (FIN && G.bid == G.nblk - 1) {
        finalize_block(m.n_cells, w, queue_count);
        return;
    }
    if (state_dead(w.counters)) return;
    const int nb = G.nblk - (FIN ? 1 : 0);
    for (int i = G.bid * blockDim.x + threadIdx.x; i < m.n_dofs; i += nb * blockDim.x) {
        double acc[DIM], wsum = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
        // the node's cells in chunks of NV: the chunk's cell indices, then
        // its partials, are loaded together (two dependent levels per chunk
        // instead of two per cell); the sums keep the cell order
        const int e0 = m.dof_off[i], e1 = m.dof_off[i + 1];
        for (int e = e0; e < e1; e += NV) {
            int ck[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) ck[q] = e + q < e1 ? m.dof_cell[e + q] : -1;
            double pv[NV][DIM + 1];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                const int c = ck[q] >= 0 ? ck[q] / ND : 0, k = ck[q] >= 0 ? ck[q] - c * ND : 0;
                const double* p = partial + (size_t)c * NACC + k * (DIM + 1);
#pragma unroll
                for (int d = 0; d <= DIM; ++d) pv[q][d] = ck[q] >= 0 ? p[d] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < NV; ++q)
                if (ck[q] >= 0) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) acc[d] += pv[q][d];
                    wsum += pv[q][DIM];
                }
        }
        if (wsum == 0.0) stat_add(stats, 4, 1ull);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double val = acc[d] / wsum;
            V[(size_t)i * DIM + d] = val;
            Vold[(size_t)i * DIM + d] = val;
        }
    }
}

template <int DIM, bool FIN, int DEG = 1>
__global__ void __launch_bounds__(256) k_node(DevMesh<DIM> m, const double* __restrict__ partial,
                                              double* __restrict__ V, double* __restrict__ Vold,
                                              unsigned long long* stats, DevWork w, int* queue_count) {
    node_body<DIM, FIN, DEG>(hw_grid(), m, partial, V, Vold, stats, w, queue_count);
}

// ---------------------------------------------------------------------------
// wall fix-up (cylinderDyM/pfem2particle.cpp:960-992)
struct DevWalls {
    int n;
    const int* dof;         // per wall node
    const double* bc;       // n * 3
    const int* mask;        // per wall node
    int* fix;               // per wall node: deviation above 1e-5 in an enforced component
    double* val;            // n * 3 projected values before the fix
    const int* vwall;       // per vertex: wall index or -1
    int n_cells;            // cells adjacent to a wall node
    const int* cells;
    const int* counters;    // the handle's counters (dead-state flags)
};

template <int DIM>
__global__ void k_wall_flags(DevWalls wl, const double* __restrict__ V) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < wl.n; i += gridDim.x * blockDim.x) {
        const int dof = wl.dof[i];
        int fix = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double val = V[(size_t)dof * DIM + d];
            wl.val[i * 3 + d] = val;
            if (((wl.mask[i] >> d) & 1) && fabs(val - wl.bc[i * 3 + d]) > 1e-5) fix = 1;
        }
        wl.fix[i] = fix;
    }
}

// one wavefront per wall-adjacent cell; per particle the corrections of the
// cell's wall vertices are applied in ascending vertex order, as the
// reference loops over wall nodes in ascending order
template <int DIM>
__global__ void k_wall_particles(DevMesh<DIM> m, DevWalls wl, const int* __restrict__ off, DevParticles A) {
    constexpr int NV = Q1<DIM>::NV;
    const int lane = threadIdx.x & (WAVE - 1);
    const int t = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (t >= wl.n_cells || state_dead(wl.counters)) return;
    const int c = wl.cells[t];
    int order[NV], widx[NV], nw = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const int gv = m.cv[(size_t)c * NV + k];
        const int wi = wl.vwall[gv];
        if (wi >= 0 && wl.fix[wi]) {
            // insertion by ascending global vertex index
            int p = nw;
            while (p > 0 && m.cv[(size_t)c * NV + order[p - 1]] > gv) {
                order[p] = order[p - 1];
                widx[p] = widx[p - 1];
                --p;
            }
            order[p] = k;
            widx[p] = wi;
            ++nw;
        }
    }
    if (nw == 0) return;
    for (int j = off[c] + lane; j < off[c + 1]; j += WAVE) {
        double xi[DIM], v[DIM], N[NV];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xi[d] = A.xi[(size_t)d * A.cap + j];
            v[d] = A.v[(size_t)d * A.cap + j];
        }
        shape<DIM>(xi, N);
        for (int q = 0; q < nw; ++q) {
            const int wi = widx[q];
            double nk = N[0];
#pragma unroll
            for (int r = 1; r < NV; ++r)
                if (r == order[q]) nk = N[r];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                if ((wl.mask[wi] >> d) & 1) v[d] = v[d] + nk * (wl.bc[wi * 3 + d] - wl.val[wi * 3 + d]);
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) A.v[(size_t)d * A.cap + j] = v[d];
    }
}

template <int DIM>
__global__ void k_wall_set(DevWalls wl, double* __restrict__ V, double* __restrict__ Vold) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < wl.n; i += gridDim.x * blockDim.x) {
        const int dof = wl.dof[i];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            if ((wl.mask[i] >> d) & 1) {
                V[(size_t)dof * DIM + d] = wl.bc[i * 3 + d];
                Vold[(size_t)dof * DIM + d] = wl.bc[i * 3 + d];
            }
    }
}

// ---------------------------------------------------------------------------
// seeding: q0*q1[*q2] particles per cell in lattice order, ids in cell order
// (cylinder2D/pfem2particle.cpp:487-508, 563-580)
template <int DIM, int DEG = 1>
__global__ void k_seed(DevMesh<DIM> m, const double* __restrict__ V, DevParticles A, DevParams pp,
                       uint32_t id_base, int* off) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    const long long total = (long long)m.n_cells * pp.nparts_in;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(i / pp.nparts_in);
        const int o = (int)(i - (long long)c * pp.nparts_in);
        double xi[DIM], P[NV][DIM], Vc[ND][DIM], x[DIM], v[DIM];
        lattice_point<DIM>(pp, o, xi);
        load_cell_any<DIM, DEG>(m, c, P, Vc);
        forward<DIM>(P, xi, x);
        interp<DIM, DEG>(Vc, xi, v);
        A.cell[i] = c;
        A.id[i] = id_base + (uint32_t)i + 1u;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            A.x[(size_t)d * A.cap + i] = x[d];
            A.xi[(size_t)d * A.cap + i] = xi[d];
            A.v[(size_t)d * A.cap + i] = v[d];
            if (pp.keep_vext) A.vext[(size_t)d * A.cap + i] = 0.0;
        }
        if (o == 0) off[c] = (int)i;
        if (i == total - 1) off[m.n_cells] = (int)total;
    }
}

// location from (cell, reference point) and, optionally, velocity from V, for
// particles placed by pf2_insert_particles
template <int DIM, int DEG = 1>
__global__ void k_init_particles(DevMesh<DIM> m, const double* __restrict__ V, DevParticles A,
                                 const unsigned char* __restrict__ need_v, long long n, int keep_vext) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int c = A.cell[i];
        double xi[DIM], P[NV][DIM], Vc[ND][DIM], x[DIM], v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = A.xi[(size_t)d * A.cap + i];
        load_cell_any<DIM, DEG>(m, c, P, Vc);
        forward<DIM>(P, xi, x);
#pragma unroll
        for (int d = 0; d < DIM; ++d) A.x[(size_t)d * A.cap + i] = x[d];
        if (need_v[i]) {
            interp<DIM, DEG>(Vc, xi, v);
#pragma unroll
            for (int d = 0; d < DIM; ++d) A.v[(size_t)d * A.cap + i] = v[d];
        }
        if (keep_vext)
#pragma unroll
            for (int d = 0; d < DIM; ++d) A.vext[(size_t)d * A.cap + i] = 0.0;
    }
}

// re-sort after a mesh move: own-cell test with the stored last advection
// velocity, no advection; a particle that fails it (or, in 2-D, whose test
// needs inverse_map's clearly-outside test and Newton iteration) is parked
// with no substeps left for k_resort_search, which relocates it in full
// wavefronts.  skip_static: the geometry changed once since the last
// re-sort, so a particle whose cell kept its vertices and whose xi came from
// an inverse map in that cell (ids up to the counter of the last advection)
// keeps its xi and cell: the inverse map would return the same bits
template <int DIM>
__global__ void __launch_bounds__(256) k_resort(DevMesh<DIM> m, DevParticles A, DevParams pp, DevWork w,
                                                DevQueues qu, int skip_static) {
    constexpr int NV = Q1<DIM>::NV;
    const int n = state_dead(w.counters) ? 0 : w.counters[0];
    const int lane = threadIdx.x & (WAVE - 1);
    const uint32_t id_exact = (uint32_t)w.counters[5];
    const WaveLoop L = stride_loop(n);
    for (int base = L.begin; base < L.end; base += L.step) {
        const int i = base + lane;
        const bool valid = i < L.end;
        const int c = valid ? A.cell[i] : -1;
        double x[DIM], xi[DIM], vext[DIM], r[DIM];
        int st = 0;   // 0 stays, 1 relocate, 2 own-cell test pending
        if (valid) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) xi[d] = A.xi[(size_t)d * A.cap + i];
            if (!(skip_static && !m.cmoved[c] && A.id[i] <= id_exact)) {
                double P[NV][DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = A.x[(size_t)d * A.cap + i];
                load_vertices<DIM>(m, c, P);
                if constexpr (DIM == 2) {
                    // inverse_map<2>(P, x, r, tol, false) up to its clearly-outside test
                    st = 2;
                    if (inverse_quadratic(P, x, r)) {
                        const double eps = 1e-15;
                        if (-eps <= r[1] && r[1] <= 1 + eps && -eps <= r[0] && r[0] <= 1 + eps)
                            st = inside_unit<DIM>(r) ? 0 : 1;
                    }
                } else {
                    st = inverse_map<DIM>(P, x, r, m.ftol[c], false) && inside_unit<DIM>(r) ? 0 : 1;
                }
                if (st == 0) {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        xi[d] = r[d];
                        A.xi[(size_t)d * A.cap + i] = r[d];
                    }
                } else {
#pragma unroll
                    for (int d = 0; d < DIM; ++d) vext[d] = A.vext ? A.vext[(size_t)d * A.cap + i] : 0.0;
                }
            }
            if (st == 0) {
                w.dest[i] = c;
                w.pkey[i] = part_key<DIM>(pp, xi);
            }
        }
        const int sp = pp.substeps - 1;
        park_particle<DIM>(A, w, qu, RESORT_QUEUE, L.shard, valid && st != 0, i, c, x, xi, vext,
                           st == 2 ? sp | PARK_NEWTON : sp);
    }
}

// dest / key of every particle for a reseed pass without movement
template <int DIM>
__global__ void k_keys(DevParticles A, DevParams pp, DevWork w) {
    const int n = state_dead(w.counters) ? 0 : w.counters[0];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double xi[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = A.xi[(size_t)d * A.cap + i];
        w.dest[i] = A.cell[i];
        w.pkey[i] = part_key<DIM>(pp, xi);
    }
}

// per-cell copies of V and V - V_old at the cell's DoFs (one load level in
// the particle kernels); V - V_old is the same per-node difference the
// correction forms (TurekBenchmark/pfem2particle.cpp:806-807)
// prescribed mesh motion: X = X0 + a W
__global__ void k_displace(const double* __restrict__ X0, const double* __restrict__ W, double a,
                           double* __restrict__ X, long long n) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        X[t] = X0[t] + a * W[t];
}

// V refreshed from a component-major device field: the node-major copy and
// the per-cell copies of V and V - V_old (FP32 copies when V32 is given; the
// FP64 V - V_old, read only by the FP64 correction, when cellDV is given)
template <int DIM>
__device__ __forceinline__ void refresh_body(const Grid& G, const DevMesh<DIM>& m, const double* __restrict__ Vcm, double* __restrict__ V, const double* __restrict__ Vold, double* __restrict__ cellV, double* __restrict__ cellDV, float* __restrict__ V32, float* __restrict__ DV32) {
    const long long nck = (long long)m.n_cells * m.nd, nd = m.n_dofs;
    const long long n = nck > nd ? nck : nd;
    for (long long t = G.bid * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)G.nblk * blockDim.x) {
        if (t < nd)
#pragma unroll
            for (int d = 0; d < DIM; ++d) V[(size_t)t * DIM + d] = Vcm[(size_t)d * nd + t];
        if (t < nck) {
            const int dof = m.cd[t];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const double a = Vcm[(size_t)d * nd + dof], b = Vold[(size_t)dof * DIM + d];
                if (cellV) cellV[(size_t)t * DIM + d] = a;
                if (cellDV) cellDV[(size_t)t * DIM + d] = a - b;
                if (V32) {
                    V32[(size_t)t * DIM + d] = (float)a;
                    DV32[(size_t)t * DIM + d] = (float)(a - b);
                }
            }
        }
    }
}

template <int DIM>
__global__ void k_refresh(DevMesh<DIM> m, const double* __restrict__ Vcm, double* __restrict__ V,
                          const double* __restrict__ Vold, double* __restrict__ cellV, double* __restrict__ cellDV,
                          float* __restrict__ V32, float* __restrict__ DV32) {
    refresh_body<DIM>(hw_grid(), m, Vcm, V, Vold, cellV, cellDV, V32, DV32);
}

template <int DIM>
__global__ void k_pack_fields(DevMesh<DIM> m, const double* __restrict__ V, const double* __restrict__ Vold,
                              double* __restrict__ cellV, double* __restrict__ cellDV) {
    const long long n = (long long)m.n_cells * m.nd;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int dof = m.cd[t];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double a = V[(size_t)dof * DIM + d], b = Vold[(size_t)dof * DIM + d];
            if (cellV) cellV[(size_t)t * DIM + d] = a;
            if (cellDV) cellDV[(size_t)t * DIM + d] = a - b;
        }
    }
}

__global__ void k_fill_int(int* p, int val, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = val;
}

// nodal field layout conversion: component-major <-> node-major
template <int DIM>
__global__ void k_to_node_major(const double* __restrict__ src, double* __restrict__ dst, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int d = 0; d < DIM; ++d) dst[(size_t)i * DIM + d] = src[(size_t)d * n + i];
}

template <int DIM>
__global__ void k_to_comp_major(const double* __restrict__ src, double* __restrict__ dst, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int d = 0; d < DIM; ++d) dst[(size_t)d * n + i] = src[(size_t)i * DIM + d];
}

}  // namespace pf2
