// pfem2_amd.hpp — header-only C++ facade over the C ABI in pfem2_amd.h with
// the class and method names of the reference's particle interface:
//   pfem2Solver           cylinder2D/pfem2particle.h:115-192 (the particle
//                         members; the FEM members stay with the driver)
//   pfem2ParticleHandler  cylinder2D/pfem2particle.h:75-112 and
//                         TurekBenchmark/pfem2particle.h:96-130
// A driver ported from the reference keeps its call sequence: it writes the
// FEM result into `solutionV[d]`, calls correct_particles_velocities(),
// move_particles() and distribute_particle_velocities_to_grid(), and reads
// the projected field back from `solutionV[d]`.  The reference's deal.II
// vectors become std::vector<double> per velocity component.  Failures
// throw pfem2_amd::Error with the library's message (the reference returns
// void and aborts on device errors).
#ifndef PFEM2_AMD_HPP
#define PFEM2_AMD_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "pfem2_amd.h"

namespace pfem2_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != 0) throw Error(rc, pf2_last_error());
}

// Triangulation + DoFHandler as plain arrays (see pfem2_amd.h)
struct Mesh {
    int dim = 2;
    std::vector<double> vertices;     // n_vertices x dim
    std::vector<int> cell_vertices;   // n_cells x 2^dim, deal.II vertex order
    std::vector<int> cell_dofs;       // empty: same as cell_vertices (degree 1)
    int n_dofs = 0;                   // 0: one DoF per vertex
    int degree = 1;                   // velocity field degree (2: cell_dofs has 3^dim DoFs per cell)
    int n_cells() const { return (int)(cell_vertices.size() >> dim); }
    int n_vertices() const { return (int)(vertices.size() / dim); }
};

// One particle as the reference's iteration exposes it
// (pfem2Particle getters, cylinder2D/pfem2particle.h:32-73)
struct Particle {
    int tria_position;
    unsigned id;
    double location[3];
    double reference_location[3];
    double velocity[3];
    const double* get_location() const { return location; }
    const double* get_reference_location() const { return reference_location; }
    double get_velocity_component(int i) const { return velocity[i]; }
    unsigned get_id() const { return id; }
};

class pfem2ParticleHandler {
public:
    void initialize_maps() { check(pf2_initialize_maps(h_)); }
    void sort_particles_into_subdomains_and_cells() { check(pf2_sort_particles_into_subdomains_and_cells(h_)); }
    long long n_global_particles() const {
        long long n = 0;
        check(pf2_n_global_particles(h_, &n));
        return n;
    }
    // the particles in the order of the reference's iteration: by cell,
    // ascending id inside a cell
    std::vector<Particle> particles() const {
        const long long n = n_global_particles();
        std::vector<int> cell(n);
        std::vector<uint32_t> id(n);
        std::vector<double> x(n * dim_), xi(n * dim_), v(n * dim_);
        check(pf2_download_particles(h_, cell.data(), id.data(), x.data(), xi.data(), v.data()));
        std::vector<Particle> out(n);
        for (long long i = 0; i < n; ++i) {
            Particle& p = out[i];
            p.tria_position = cell[i];
            p.id = id[i];
            for (int d = 0; d < 3; ++d) {
                p.location[d] = d < dim_ ? x[i * dim_ + d] : 0.0;
                p.reference_location[d] = d < dim_ ? xi[i * dim_ + d] : 0.0;
                p.velocity[d] = d < dim_ ? v[i * dim_ + d] : 0.0;
            }
        }
        return out;
    }
    std::vector<int> n_particles_in_cells() const {
        std::vector<int> c(n_cells_);
        check(pf2_cell_counts(h_, c.data()));
        return c;
    }
    pf2_handle handle() const { return h_; }

private:
    friend class pfem2Solver;
    pf2_handle h_ = nullptr;
    int dim_ = 2, n_cells_ = 0;
};

class pfem2Solver {
public:
    // time_step, PARTICLES_MOVEMENT_STEPS and MAX_PARTICLES_PER_CELL_PART of
    // the reference (cylinder2D/pfem2particle.h:4-5, 161)
    pfem2Solver(Mesh mesh, double time_step, int particles_movement_steps, int max_particles_per_cell_part,
                int precision = PF2_FP64, bool keep_velocity_ext = false)
        : time_step(time_step), mesh_(std::move(mesh)), substeps_(particles_movement_steps),
          max_(max_particles_per_cell_part), precision_(precision), keep_(keep_velocity_ext) {
        const int nd = mesh_.n_dofs > 0 ? mesh_.n_dofs : mesh_.n_vertices();
        for (int d = 0; d < mesh_.dim; ++d) solutionV[d].assign(nd, 0.0);
    }
    ~pfem2Solver() {
        if (particle_handler.h_) pf2_destroy(particle_handler.h_);
    }
    pfem2Solver(const pfem2Solver&) = delete;
    pfem2Solver& operator=(const pfem2Solver&) = delete;

    // cylinder2D/pfem2particle.cpp:563-580: lattice of quantities[d] points
    // per direction in every cell, velocities interpolated from solutionV
    void seed_particles(const std::vector<unsigned>& quantities) {
        if (particle_handler.h_) throw Error(PF2_E_STATE, "particles were already seeded");
        pf2_params p{};
        p.dt = time_step;
        p.substeps = substeps_;
        p.max_per_part = max_;
        for (int d = 0; d < 3; ++d) p.quantities[d] = d < (int)quantities.size() ? (int)quantities[d] : 1;
        p.precision = precision_;
        p.keep_velocity_ext = keep_ ? 1 : 0;
        pf2_mesh m{};
        m.dim = mesh_.dim;
        m.n_cells = mesh_.n_cells();
        m.n_vertices = mesh_.n_vertices();
        m.n_dofs = (int)solutionV[0].size();
        m.vertices = mesh_.vertices.data();
        m.cell_vertices = mesh_.cell_vertices.data();
        m.cell_dofs = mesh_.cell_dofs.empty() ? nullptr : mesh_.cell_dofs.data();
        m.degree = mesh_.degree;
        check(pf2_create(&m, &p, 0, &particle_handler.h_));
        particle_handler.dim_ = mesh_.dim;
        particle_handler.n_cells_ = m.n_cells;
        const std::vector<double> V = packed();
        check(pf2_upload_fields(particle_handler.h_, V.data(), V.data()));
        check(pf2_seed_particles(particle_handler.h_));
    }
    // cylinder2D/pfem2particle.cpp:582-603; sends the driver's new solutionV
    // (old_solution is resident: the last projected field)
    void correct_particles_velocities() {
        const std::vector<double> V = packed();
        check(pf2_upload_fields(h(), V.data(), nullptr));
        check(pf2_correct_particles_velocities(h()));
    }
    // cylinder2D/pfem2particle.cpp:605-647
    void move_particles() { check(pf2_move_particles(h())); }
    // TurekBenchmark/pfem2particle.cpp:858-918: the projection replaces solutionV
    void distribute_particle_velocities_to_grid() {
        std::vector<double> V(solutionV[0].size() * mesh_.dim);
        check(pf2_distribute_particle_velocities_to_grid(h(), V.data()));
        const size_t nd = solutionV[0].size();
        for (int d = 0; d < mesh_.dim; ++d)
            solutionV[d].assign(V.begin() + d * nd, V.begin() + (d + 1) * nd);
    }
    // cylinderDyM/pfem2particle.cpp:759-763
    void check_empty_cells() { check(pf2_check_empty_cells(h())); }
    // cylinderDyM/pfem2particle.cpp:1252-1273: new vertex positions
    void transform_grid(const std::vector<double>& vertices) {
        mesh_.vertices = vertices;
        check(pf2_update_vertices(h(), vertices.data(), 0));
    }
    // wallsAndBodyDoFs (cylinderDyM/pfem2particle.h:253)
    void set_walls(const std::vector<int>& vertex, const std::vector<int>& dof, const std::vector<double>& bc,
                   const std::vector<int>& mask) {
        check(pf2_set_walls(h(), (int)vertex.size(), vertex.data(), dof.data(), bc.data(), mask.data()));
    }
    unsigned particleCount() const {
        uint32_t c = 0;
        check(pf2_particle_counter(particle_handler.h_, &c));
        return c;
    }

    double time_step;
    std::vector<double> solutionV[3];
    pfem2ParticleHandler particle_handler;

private:
    pf2_handle h() const {
        if (!particle_handler.h_) throw Error(PF2_E_STATE, "seed_particles has not been called");
        return particle_handler.h_;
    }
    std::vector<double> packed() const {
        std::vector<double> V;
        for (int d = 0; d < mesh_.dim; ++d) V.insert(V.end(), solutionV[d].begin(), solutionV[d].end());
        return V;
    }
    Mesh mesh_;
    int substeps_, max_, precision_;
    bool keep_;
};

}  // namespace pfem2_amd

#endif  // PFEM2_AMD_HPP
