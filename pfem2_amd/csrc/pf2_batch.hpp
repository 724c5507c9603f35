// pf2_batch.hpp — one step of many simulations with one launch per kernel.
//
// A batch lists simulations of the same dimension and precision.  Every
// kernel of the step (field refresh, move, relocation rounds, census, offset
// scan, write, bookkeeping, projection) is launched once for the whole
// batch: the launch's workgroups are split into consecutive ranges, one per
// simulation (sized as that simulation's own launch would be), and each
// workgroup runs the kernel body of its simulation with the range as its
// grid.  A simulation's results are therefore those of its own launches, bit
// for bit: the same bodies, the same grids, the same data.
//
// Device table of a batch of n simulations: BatchSim[n], then for each kernel
// kind the first workgroup of every simulation (n + 1 ints per kind, the
// last one the launch's total).
#pragma once

#include "pf2_mixed.hpp"

namespace pf2 {

enum BatchKind {
    BK_REFRESH = 0, BK_MOVE, BK_SEARCH0, BK_SEARCH1, BK_SEARCH2, BK_COUNT, BK_COUNT_L, BK_BIG, BK_TILES,
    BK_WRITE, BK_PARTIALS, BK_NODE, BK_NKINDS
};
static_assert(MOVE_ROUNDS <= 2, "kinds for up to three relocation rounds");

template <int DIM>
struct BatchSim {
    DevMesh<DIM> m;
    DevParams pp;
    DevWork w;                    // new_off: the other offset buffer
    DevQueues qu;
    DevParticles A, B;            // current and other particle buffer
    Rec32 RA, RB;                 // their FP32 records (FP32-mixed)
    const int* old_off;           // current offsets
    const double* field;          // the step's V, component-major
    double *V, *Vold, *cellV, *cellDV;
    float *V32, *dV32;
    const float* geo;
    double* partial;
    int* qcount;
    unsigned long long* tiles;    // tile sums of the offset scan
    int ntiles;                   // tiles of SCAN_TILE entries of the n_cells + 1 counts
};

constexpr int SCAN_TILE = 1024;   // entries per workgroup of the offset scan (4 per thread)

// the simulation whose range of kind K holds workgroup b: the last one whose
// first workgroup is <= b (simulations with an empty range are skipped)
__device__ __forceinline__ int batch_find(const int* __restrict__ st, int nsim, int b) {
    int lo = 0, hi = nsim - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int DIM>
struct BatchView {
    const BatchSim<DIM>* sims;
    const int* starts;
    int nsim;
    // this workgroup's simulation and its grid for kind K
    __device__ __forceinline__ const BatchSim<DIM>& at(int K, Grid& G) const {
        const int* st = starts + (size_t)K * (nsim + 1);
        const int s = batch_find(st, nsim, (int)blockIdx.x);
        G.bid = blockIdx.x - (unsigned)st[s];
        G.nblk = (unsigned)(st[s + 1] - st[s]);
        return sims[s];
    }
};

template <int DIM>
__global__ void __launch_bounds__(256) kb_refresh(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_REFRESH, G);
    refresh_body<DIM>(G, S.m, S.field, S.V, S.Vold, S.cellV, S.cellDV, S.V32, S.dV32);
}

template <int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 3 : 2))) kb_move(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_MOVE, G);
    move_body<DIM>(G, S.m, S.V, S.Vold, S.A, S.pp, S.w, S.qu, 1, 1);
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_move32(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_MOVE, G);
    move32_body<DIM>(G, S.m, S.V32, S.dV32, S.geo, S.A.cell, S.RA, S.A.vext, S.A.cap, S.pp, S.w, S.qu, 1, 1);
}

template <int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 3 : 2))) kb_search(BatchView<DIM> bv, int round) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_SEARCH0 + round, G);
    search_body<DIM>(G, S.m, S.V, S.A, S.pp, S.w, S.qu, round);
}

template <int DIM, bool LAST>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LAST ? 2 : (DIM == 2 ? 4 : 3)))) kb_search32(BatchView<DIM> bv, int round) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_SEARCH0 + round, G);
    search32_body<DIM, LAST>(G, S.m, S.V32, S.geo, S.A.cell, S.RA, S.A.vext, S.A.cap, S.pp, S.w, S.qu, round);
}

template <int DIM, int CAP>
__global__ void __launch_bounds__(256) kb_count(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(CAP == GCAP ? BK_COUNT : BK_COUNT_L, G);
    cell_count_body<DIM, CAP>(G, S.m.n_cells, S.old_off, S.A.id, S.pp, S.w, 1);
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_big(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_BIG, G);
    cell_big_body<DIM>(G, S.old_off, S.A.id, S.pp, S.w, 1, S.w.big);
}

// ---- exclusive scan of the per-cell counts (fs_cnt -> fs_off), per simulation
// sum of 4 consecutive u64 per thread over the workgroup (all threads call)
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long v, unsigned long long* s_w,
                                                              unsigned long long* total) {
    const int t = threadIdx.x, lane = t & (WAVE - 1), wv = t / WAVE;
    unsigned long long incl = v;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const unsigned long long o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    if (lane == WAVE - 1) s_w[wv] = incl;
    __syncthreads();
    unsigned long long base = 0, tot = 0;
    for (int q = 0; q < (int)(blockDim.x / WAVE); ++q) {
        if (q < wv) base += s_w[q];
        tot += s_w[q];
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// workgroup t of a simulation: the sum of its tile of counts
template <int DIM>
__global__ void __launch_bounds__(256) kb_scan_tiles(BatchView<DIM> bv) {
    __shared__ unsigned long long s_w[256 / WAVE];
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_TILES, G);
    const long long n = (long long)S.m.n_cells + 1;
    const long long i0 = (long long)G.bid * SCAN_TILE + threadIdx.x * 4;
    unsigned long long v = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (i0 + u < n) v += S.w.fs_cnt[i0 + u];
    unsigned long long tot;
    block_excl_scan(v, s_w, &tot);
    if (threadIdx.x == 0) S.tiles[G.bid] = tot;
}

// workgroup s: exclusive scan of simulation s's tile sums, in place
template <int DIM>
__global__ void __launch_bounds__(256) kb_scan_tops(BatchView<DIM> bv) {
    __shared__ unsigned long long s_w[256 / WAVE];
    const BatchSim<DIM>& S = bv.sims[blockIdx.x];
    unsigned long long carry = 0;
    for (int b = 0; b < S.ntiles; b += 256) {
        const int t = b + threadIdx.x;
        const unsigned long long v = t < S.ntiles ? S.tiles[t] : 0;
        unsigned long long tot;
        const unsigned long long ex = block_excl_scan(v, s_w, &tot);
        if (t < S.ntiles) S.tiles[t] = carry + ex;
        carry += tot;
    }
}

// workgroup t of a simulation: the exclusive scan of its tile plus the tile's offset
template <int DIM>
__global__ void __launch_bounds__(256) kb_scan_apply(BatchView<DIM> bv) {
    __shared__ unsigned long long s_w[256 / WAVE];
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_TILES, G);
    const long long n = (long long)S.m.n_cells + 1;
    const long long i0 = (long long)G.bid * SCAN_TILE + threadIdx.x * 4;
    unsigned long long v[4], sum = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u] = i0 + u < n ? S.w.fs_cnt[i0 + u] : 0;
        sum += v[u];
    }
    unsigned long long tot;
    unsigned long long run = S.tiles[G.bid] + block_excl_scan(sum, s_w, &tot);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (i0 + u < n) S.w.fs_off[i0 + u] = run;
        run += v[u];
    }
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_write(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_WRITE, G);
    cell_write_body<DIM, false, false>(G, S.m, S.V, S.old_off, S.A, S.B, S.pp, S.w, S.partial);
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_write32(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_WRITE, G);
    cell_write32_body<DIM, false>(G, S.m, S.V32, S.old_off, S.A.id, S.RA, S.B.cell, S.B.id, S.RB, S.A.vext,
                                  S.B.vext, S.A.cap, S.pp, S.w, S.partial);
}

// workgroup s: simulation s's bookkeeping after the write (k_finalize's work)
template <int DIM>
__global__ void __launch_bounds__(QSHARDS) kb_finalize(BatchView<DIM> bv) {
    const BatchSim<DIM>& S = bv.sims[blockIdx.x];
    finalize_block(S.m.n_cells, S.w, S.qcount);
}

// after the write the other buffer holds the particles and new_off their offsets
template <int DIM>
__global__ void __launch_bounds__(256) kb_partials(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_PARTIALS, G);
    partials_body<DIM>(G, S.m.n_cells, S.w.new_off, S.B, S.partial, S.w.counters);
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_partials32(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_PARTIALS, G);
    partials32_body<DIM>(G, S.m.n_cells, S.w.new_off, S.RB, S.partial, S.w.counters);
}

template <int DIM>
__global__ void __launch_bounds__(256) kb_node(BatchView<DIM> bv) {
    Grid G;
    const BatchSim<DIM>& S = bv.at(BK_NODE, G);
    node_body<DIM, false>(G, S.m, S.partial, S.V, S.Vold, S.w.stats, S.w, S.qcount);
}

}  // namespace pf2
