// Replays the particle part of a reference case driver through the C++
// facade (pfem2_amd.hpp), in the driver's call order
// (cylinder2D/cylinder2D.cc:850, 865-867):
//   seed_particles(quantities)
//   every step: solutionV = FEM result; correct_particles_velocities();
//               move_particles(); distribute_particle_velocities_to_grid()
// The FEM result is a fixed field read from the input file.
//
// usage: driver_replay <in.bin> <out.bin>
//   in.bin : int32 [dim, n_cells, n_vertices, n_dofs, steps, substeps, max_per_part,
//                   q0, q1, q2, precision], float64 dt,
//            float64 vertices[n_vertices*dim], int32 cells[n_cells*2^dim],
//            float64 V[dim*n_dofs] (component-major)
//   out.bin: int64 particles, uint32 particleCount, int64 sum of cell ids over
//            particles, float64 V[dim*n_dofs] (the last projection)
#include <cstdio>
#include <cstdint>
#include <vector>

#include "pfem2_amd.hpp"

template <typename T>
static bool rd(FILE* f, T* p, size_t n) {
    return fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hd[11];
    double dt = 0;
    if (!rd(f, hd, 11) || !rd(f, &dt, 1)) return 2;
    const int dim = hd[0], nc = hd[1], nvt = hd[2], nd = hd[3], steps = hd[4];
    pfem2_amd::Mesh mesh;
    mesh.dim = dim;
    mesh.vertices.resize((size_t)nvt * dim);
    mesh.cell_vertices.resize((size_t)nc << dim);
    mesh.n_dofs = nd;
    std::vector<double> V((size_t)dim * nd);
    if (!rd(f, mesh.vertices.data(), mesh.vertices.size()) || !rd(f, mesh.cell_vertices.data(), mesh.cell_vertices.size()) ||
        !rd(f, V.data(), V.size()))
        return 2;
    fclose(f);
    try {
        pfem2_amd::pfem2Solver solver(mesh, dt, hd[5], hd[6], hd[10]);
        for (int d = 0; d < dim; ++d) solver.solutionV[d].assign(V.begin() + (size_t)d * nd, V.begin() + (size_t)(d + 1) * nd);
        std::vector<unsigned> q(hd + 7, hd + 7 + dim);
        solver.seed_particles(q);
        for (int s = 0; s < steps; ++s) {
            for (int d = 0; d < dim; ++d)
                solver.solutionV[d].assign(V.begin() + (size_t)d * nd, V.begin() + (size_t)(d + 1) * nd);
            solver.correct_particles_velocities();
            solver.move_particles();
            solver.distribute_particle_velocities_to_grid();
        }
        const long long n = solver.particle_handler.n_global_particles();
        long long cell_sum = 0;
        for (const auto& p : solver.particle_handler.particles()) cell_sum += p.tria_position;
        const uint32_t cnt = solver.particleCount();
        FILE* o = fopen(argv[2], "wb");
        if (!o) return 2;
        fwrite(&n, 8, 1, o);
        fwrite(&cnt, 4, 1, o);
        fwrite(&cell_sum, 8, 1, o);
        for (int d = 0; d < dim; ++d) fwrite(solver.solutionV[d].data(), 8, nd, o);
        fclose(o);
    } catch (const pfem2_amd::Error& e) {
        fprintf(stderr, "pfem2_amd error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
