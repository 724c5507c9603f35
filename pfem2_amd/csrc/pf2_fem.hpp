// pf2_fem.hpp — nodal routines that the reference's drivers run around the
// particle stage, on the device:
//   k_reinterp      pfem2Solver::reinterpolate_fields
//                   (cylinderDyM/pfem2particle.cpp:1144-1250): nodal fields of
//                   the mesh before a move, evaluated at the displaced vertices
//   k_loads_terms   pfem2Solver::calculate_loads
//   k_loads_sum     (TurekBenchmark/pfem2particle.cpp:920-970): pressure and
//                   viscous force terms per boundary-face Gauss point, then
//                   their sums in face order
// Both follow the reference's expressions in its order (FP64, no
// contraction), so that they equal a host evaluation bit for bit.
#pragma once

#include "pf2_kernels.hpp"

namespace pf2 {

template <int DIM>
__device__ __forceinline__ bool inside_unit_eps(const double* xi, double eps) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        if (!(xi[d] >= -eps && xi[d] <= 1.0 + eps)) return false;
    return true;
}

// One thread per vertex.  A boundary vertex (vbnd) keeps its values.  The
// search starts in the vertex's first star cell (the smallest index, where
// the reference's moving node records its cell), then the other star cells
// in ascending order; a cell holds the displaced vertex when its inverse
// map (no clearly-outside shortcut: the acceptance is widened by 1e-5) lies
// in the unit cell widened by 1e-5.  A failed inverse map leaves the
// reference coordinates as they were: (0, 0[, 0]) in the first cell, which
// counts as inside, as the reference's caught exception leaves its
// localCoords.  Fields: V node-major (DIM per DoF), the extras component-major
// (n_extra x n_dofs).  Writes Vn and En at the vertices' DoFs; they must not
// alias V and E and must hold copies of them beforehand (a DoF no vertex
// carries keeps its value).
template <int DIM>
__global__ void k_reinterp(DevMesh<DIM> m, const unsigned char* __restrict__ vbnd, const int* __restrict__ vdof,
                           const double* __restrict__ U, const double* __restrict__ V, double* __restrict__ Vn,
                           const double* __restrict__ E, double* __restrict__ En, int n_extra) {
    constexpr int NV = Q1<DIM>::NV;
    const long long nd = m.n_dofs;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < m.n_vertices; v += gridDim.x * blockDim.x) {
        const int dof = vdof[v];
        if (dof < 0) continue;   // a vertex of no cell carries no DoF
        int cell = -1;
        double xi[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = 0.0;
        if (!vbnd[v]) {
            double p[DIM], P[NV][DIM], r[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) p[d] = m.X[(size_t)v * DIM + d] + U[(size_t)v * DIM + d];
            const int s0 = m.star_off[v], s1 = m.star_off[v + 1];
            const int c0 = m.star_cell[s0];
            gather_vertices<DIM>(m, c0, P);
            if (inverse_map<DIM>(P, p, r, -1.0, false))
#pragma unroll
                for (int d = 0; d < DIM; ++d) xi[d] = r[d];
            if (inside_unit_eps<DIM>(xi, 1e-5)) {
                cell = c0;
            } else {
                for (int e = s0; e < s1; ++e) {
                    const int c = m.star_cell[e];
                    if (c == c0) continue;
                    gather_vertices<DIM>(m, c, P);
                    if (inverse_map<DIM>(P, p, r, -1.0, false))
#pragma unroll
                        for (int d = 0; d < DIM; ++d) xi[d] = r[d];
                    if (inside_unit_eps<DIM>(xi, 1e-5)) {
                        cell = c;
                        break;
                    }
                }
            }
        }
        if (cell >= 0) {
            double N[NV];
            shape<DIM>(xi, N);
            double acc[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
            for (int k = 0; k < NV; ++k) {
                const int q = m.cd[(size_t)cell * NV + k];
#pragma unroll
                for (int d = 0; d < DIM; ++d) acc[d] += N[k] * V[(size_t)q * DIM + d];
            }
#pragma unroll
            for (int d = 0; d < DIM; ++d) Vn[(size_t)dof * DIM + d] = acc[d] / 1.0;
            for (int x = 0; x < n_extra; ++x) {
                double a = 0.0;
                for (int k = 0; k < NV; ++k) a += N[k] * E[(size_t)x * nd + m.cd[(size_t)cell * NV + k]];
                En[(size_t)x * nd + dof] = a / 1.0;
            }
        } else {
            // the values stay (0 + value, as the reference accumulates them)
#pragma unroll
            for (int d = 0; d < DIM; ++d) Vn[(size_t)dof * DIM + d] = (0.0 + V[(size_t)dof * DIM + d]) / 1.0;
            for (int x = 0; x < n_extra; ++x) En[(size_t)x * nd + dof] = (0.0 + E[(size_t)x * nd + dof]) / 1.0;
        }
    }
}

// Per boundary face f (cell, face number) and Gauss point q of QGauss<1>(2)
// (unit face coordinate 1/2 -+ 1/(2 sqrt 3), weight 1/2), the four force
// terms of calculate_loads: mu dVt/dn n_y JxW, p n_x JxW, mu dVt/dn n_x JxW,
// p n_y JxW (the reference adds the first and subtracts the others).
// FEFaceValues of MappingQ1: face points (0, s), (1, s), (s, 0), (s, 1) for
// faces 0-3; tangent J t_f with t_f = (0,-1), (0,1), (1,0), (-1,0); boundary
// form (t_y, -t_x): its length times the weight is JxW and it divided by its
// length the unit normal; shape gradients J^-T grad_xi N.  Vcm: velocity
// component-major (2 x n_dofs); P: pressure (n_dofs).
__global__ void k_loads_terms(DevMesh<2> m, const int* __restrict__ faces, int n_faces,
                              const double* __restrict__ Vcm, const double* __restrict__ Pf, double mu,
                              double* __restrict__ terms) {
    const long long nd = m.n_dofs;
    const double g = 0.5 / sqrt(3.0);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n_faces; t += gridDim.x * blockDim.x) {
        const int f = t >> 1, q = t & 1;
        const int c = faces[2 * f], fn = faces[2 * f + 1];
        double P[4][2], xi[2], N[4], G[4][2];
        gather_vertices<2>(m, c, P);
        const double s = q == 0 ? 0.5 - g : 0.5 + g;
        xi[0] = fn == 0 ? 0.0 : (fn == 1 ? 1.0 : s);
        xi[1] = fn == 2 ? 0.0 : (fn == 3 ? 1.0 : s);
        shape<2>(xi, N);
        shape_grad<2>(xi, G);
        double J[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) J[i][j] += P[k][i] * G[k][j];
        const double tfx = fn == 2 ? 1.0 : (fn == 3 ? -1.0 : 0.0);
        const double tfy = fn == 0 ? -1.0 : (fn == 1 ? 1.0 : 0.0);
        const double t0 = J[0][0] * tfx + J[0][1] * tfy;
        const double t1 = J[1][0] * tfx + J[1][1] * tfy;
        const double len = sqrt(t1 * t1 + t0 * t0);
        const double nx = t1 / len, ny = -t0 / len;
        const double JxW = len * 0.5;
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double inv = 1.0 / det;
        double pv = 0.0, dvt = 0.0;
        for (int k = 0; k < 4; ++k) {
            const int dof = m.cd[(size_t)c * 4 + k];
            const double gx = (J[1][1] * G[k][0] - J[1][0] * G[k][1]) * inv;
            const double gy = (J[0][0] * G[k][1] - J[0][1] * G[k][0]) * inv;
            pv += Pf[dof] * N[k];
            dvt += (Vcm[dof] * ny - Vcm[nd + dof] * nx) * (gx * nx + gy * ny);
        }
        double* o = terms + (size_t)t * 4;
        o[0] = mu * dvt * ny * JxW;
        o[1] = pv * nx * JxW;
        o[2] = mu * dvt * nx * JxW;
        o[3] = pv * ny * JxW;
    }
}

// the sums in face order and the coefficients 2 F / (rho u_mean^2 diam),
// then the pressure difference P[dof_a] - P[dof_e] (one thread)
__global__ void k_loads_sum(const double* __restrict__ terms, int n_terms, const double* __restrict__ Pf,
                            double rho, double u_mean, double diam, int dof_a, int dof_e, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double Fx_nu = 0.0, Fx_p = 0.0, Fy_nu = 0.0, Fy_p = 0.0;
    for (int t = 0; t < n_terms; ++t) {
        const double* o = terms + (size_t)t * 4;
        Fx_nu += o[0];
        Fx_p -= o[1];
        Fy_nu -= o[2];
        Fy_p -= o[3];
    }
    const double s = rho * u_mean * u_mean * diam;
    out[0] = 2.0 * Fx_nu / s;
    out[1] = 2.0 * Fx_p / s;
    out[2] = 2.0 * Fy_nu / s;
    out[3] = 2.0 * Fy_p / s;
    double pd = 0.0;
    if (dof_a >= 0) pd += Pf[dof_a];
    if (dof_e >= 0) pd -= Pf[dof_e];
    out[4] = pd;
}

}  // namespace pf2
