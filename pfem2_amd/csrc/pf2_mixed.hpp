// pf2_mixed.hpp — FP32-mixed variant of the particle stage.
//
// A particle is stored as (cell, id, xi in FP32, v in FP32): 8 + 4*DIM*2
// bytes instead of 8 + 8*DIM*3 in the FP64 path (there is no absolute
// position; it is F_K(xi) on demand); xi and v sit together in one record.  Within a substep the displacement
// h*u is applied in reference coordinates of the current cell by a short
// FP32 Newton solve of F_K(xi + delta) - F_K(xi) = h*u, written with the
// exact multilinear difference so no absolute coordinates enter the FP32
// arithmetic.  Only a particle whose new reference point leaves the unit
// cell goes through the FP64 path: F_K(xi) + h*u in FP64 and the relocation
// rule of the reference (vertex star, FP64 inverse map).  Projection sums use
// Kahan-compensated FP32 per cell and FP64 node sums.
#pragma once

#include "pf2_kernels.hpp"

namespace pf2 {

inline int geo32_floats(int dim) { return dim == 2 ? 6 : 21; }

// FP32 particle state, one record per particle, so that a scattered access
// touches one cache line (2-D) or two (3-D):
//   2-D  a = {xi0, xi1, v0, v1}
//   3-D  a = {xi0, xi1, xi2, v0}, b = {v1, v2}
struct Rec32 {
    float4* a;
    float2* b;   // 3-D only
};

template <int DIM>
__device__ __forceinline__ void rec_load(const Rec32& r, long long i, float* xi, float* v) {
    const float4 q = r.a[i];
    if constexpr (DIM == 2) {
        xi[0] = q.x;
        xi[1] = q.y;
        v[0] = q.z;
        v[1] = q.w;
    } else {
        const float2 t = r.b[i];
        xi[0] = q.x;
        xi[1] = q.y;
        xi[2] = q.z;
        v[0] = q.w;
        v[1] = t.x;
        v[2] = t.y;
    }
}

template <int DIM>
__device__ __forceinline__ void rec_load_xi(const Rec32& r, long long i, float* xi) {
    if constexpr (DIM == 2) {
        const float2 q = reinterpret_cast<const float2*>(r.a + i)[0];
        xi[0] = q.x;
        xi[1] = q.y;
    } else {
        const float4 q = r.a[i];
        xi[0] = q.x;
        xi[1] = q.y;
        xi[2] = q.z;
    }
}

template <int DIM>
__device__ __forceinline__ void rec_store(const Rec32& r, long long i, const float* xi, const float* v) {
    if constexpr (DIM == 2) {
        r.a[i] = make_float4(xi[0], xi[1], v[0], v[1]);
    } else {
        r.a[i] = make_float4(xi[0], xi[1], xi[2], v[0]);
        r.b[i] = make_float2(v[1], v[2]);
    }
}

template <int DIM>
__device__ __forceinline__ void rec_store_xi(const Rec32& r, long long i, const float* xi) {
    if constexpr (DIM == 2) {
        reinterpret_cast<float2*>(r.a + i)[0] = make_float2(xi[0], xi[1]);
    } else {
        float* p = reinterpret_cast<float*>(r.a + i);
        p[0] = xi[0];
        p[1] = xi[1];
        p[2] = xi[2];
    }
}

template <int DIM>
__device__ __forceinline__ void rec_store_v(const Rec32& r, long long i, const float* v) {
    if constexpr (DIM == 2) {
        reinterpret_cast<float2*>(r.a + i)[1] = make_float2(v[0], v[1]);
    } else {
        reinterpret_cast<float*>(r.a + i)[3] = v[0];
        r.b[i] = make_float2(v[1], v[2]);
    }
}

template <int DIM>
__device__ __forceinline__ void rec_copy(const Rec32& src, long long i, const Rec32& dst, long long o) {
    dst.a[o] = src.a[i];
    if constexpr (DIM == 3) dst.b[o] = src.b[i];
}

// multilinear coefficients of F_K relative to vertex 0, in FP32
//   2-D: F = x0 + a xi0 + b xi1 + c xi0 xi1
//   3-D: F = x0 + a xi0 + b xi1 + c xi2 + d xi0 xi1 + e xi0 xi2 + f xi1 xi2 + g xi0 xi1 xi2
template <int DIM>
__global__ void k_geo32(DevMesh<DIM> m, float* __restrict__ geo) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int NG = DIM == 2 ? 6 : 21;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m.n_cells; c += gridDim.x * blockDim.x) {
        double P[NV][DIM];
        gather_vertices<DIM>(m, c, P);
        float* g = geo + (size_t)c * NG;
        if constexpr (DIM == 2) {
            for (int d = 0; d < 2; ++d) {
                g[0 + d] = (float)(P[1][d] - P[0][d]);
                g[2 + d] = (float)(P[2][d] - P[0][d]);
                g[4 + d] = (float)((P[0][d] - P[1][d]) - (P[2][d] - P[3][d]));
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                const double x0 = P[0][d], x1 = P[1][d], x2 = P[2][d], x3 = P[3][d];
                const double x4 = P[4][d], x5 = P[5][d], x6 = P[6][d], x7 = P[7][d];
                g[0 + d] = (float)(x1 - x0);                                    // a
                g[3 + d] = (float)(x2 - x0);                                    // b
                g[6 + d] = (float)(x4 - x0);                                    // c
                g[9 + d] = (float)((x0 - x1) - (x2 - x3));                      // d (xi0 xi1)
                g[12 + d] = (float)((x0 - x1) - (x4 - x5));                     // e (xi0 xi2)
                g[15 + d] = (float)((x0 - x2) - (x4 - x6));                     // f (xi1 xi2)
                g[18 + d] = (float)(((x1 - x0) - (x3 - x2)) - ((x5 - x4) - (x7 - x6)));   // g
            }
        }
    }
}

// per-cell FP64 tolerance of the clearly-outside test, the packed vertex
// copy, and whether the copy changed (cmoved: any coordinate differs from
// the one it replaces)
template <int DIM>
__global__ void k_ftol(DevMesh<DIM> m, double* __restrict__ ftol, double* __restrict__ cellX,
                       unsigned char* __restrict__ cmoved) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < m.n_cells; c += gridDim.x * blockDim.x) {
        double P[Q1<DIM>::NV][DIM];
        gather_vertices<DIM>(m, c, P);
        ftol[c] = outside_tol<DIM>(P);
        bool moved = false;
#pragma unroll
        for (int k = 0; k < Q1<DIM>::NV; ++k)
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                double* q = cellX + ((size_t)c * Q1<DIM>::NV + k) * DIM + d;
                moved = moved || *q != P[k][d];
                *q = P[k][d];
            }
        cmoved[c] = moved ? 1 : 0;
    }
}

template <int DIM>
void launch_geo(hipStream_t s, DevMesh<DIM> m, float* geo, double* ftol, double* cellX, unsigned char* cmoved) {
    // geo: the FP32 geometry, for FP32-mixed handles only
    if (geo) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_geo32<DIM>), dim3((m.n_cells + 255) / 256), dim3(256), 0, s, m, geo);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ftol<DIM>), dim3((m.n_cells + 255) / 256), dim3(256), 0, s, m, ftol, cellX,
                       cmoved);
}

template <int DIM>
__device__ __forceinline__ void shape32(const float* xi, float* N) {
#pragma unroll
    for (int k = 0; k < Q1<DIM>::NV; ++k) {
        float val = 1.0f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) val *= ((k >> d) & 1) ? xi[d] : (1.0f - xi[d]);
        N[k] = val;
    }
}

// N_k alone, for a lane with a fixed vertex k: factor d is t_d xi_d + s_d
// (xi_d, or 1 - xi_d with one rounding), multiplied in shape32's order, so
// the value is shape32's N[k]
template <int DIM>
__device__ __forceinline__ void shape_one32_coeffs(int k, float* t, float* s) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const bool one = (k >> d) & 1;
        t[d] = one ? 1.0f : -1.0f;
        s[d] = one ? 0.0f : 1.0f;
    }
}

template <int DIM>
__device__ __forceinline__ float shape_one32(const float* t, const float* s, const float* xi) {
    float val = __fmaf_rn(t[0], xi[0], s[0]);
#pragma unroll
    for (int d = 1; d < DIM; ++d) val *= __fmaf_rn(t[d], xi[d], s[d]);
    return val;
}

// The local difference of F_K splits into the Jacobian J(xi) at the current
// reference point and a nonlinear part Q:
//   F(xi + delta) - F(xi) = J(xi) delta + Q(delta)
//   2-D: Q = c d0 d1
//   3-D: Q = d d0 d1 + e d0 d2 + f d1 d2 + g (xi0 d1 d2 + xi1 d0 d2 + xi2 d0 d1 + d0 d1 d2)
//
// delta with F(xi + delta) - F(xi) = dx: delta = J^-1 (dx - Q(delta)) by
// fixed-point iteration from delta = J^-1 dx (contraction ~ |Q'| / |J| ~
// the step in cell units), carried out in reference coordinates (J^-1
// applied once to dx and to Q's coefficient vectors).  Returns false when
// J is not positive or the change of a fourth iteration, which is the
// residual mapped by J^-1, exceeds 1e-6 |J^-1 dx| in the 1-norm (the caller
// then takes the FP64 path).  Three iterations and this bound keep the
// projected field of the FP32 path within ~2e-6 (relative L2) of the FP64
// path on the cylinder workload; two iterations and 1e-5 gave 1.1e-5.
#ifndef PF2_LS_ITERS
#define PF2_LS_ITERS 3
#endif
#ifndef PF2_LS_TOL
#define PF2_LS_TOL 1e-6f
#endif
template <int DIM>
__device__ __forceinline__ bool local_step(const float* g, const float* xi, const float* dx, float* dl) {
    if constexpr (DIM == 2) {
        // 2-D: F(xi + delta) - F(xi) = J delta + c delta0 delta1 with c =
        // (g4, g5), so the same iteration reads delta = e - w delta0 delta1
        // with e = J^-1 dx and w = J^-1 c (three instructions per iteration
        // instead of nine).  The change of a fourth iteration is the
        // residual J^-1 (dx - F(xi + delta) + F(xi)); the step returns that
        // fourth iterate when the change is within 1e-6 |e| (1-norms).  No
        // early exit: with det <= 0 the result is discarded.
        const float J00 = __fmaf_rn(g[4], xi[1], g[0]), J10 = __fmaf_rn(g[5], xi[1], g[1]);
        const float J01 = __fmaf_rn(g[4], xi[0], g[2]), J11 = __fmaf_rn(g[5], xi[0], g[3]);
        const float det = __fmaf_rn(J00, J11, -(J01 * J10));
        const float r = __builtin_amdgcn_rcpf(det);
        const float e0 = __fmaf_rn(J11, dx[0], -(J01 * dx[1])) * r;
        const float e1 = __fmaf_rn(J00, dx[1], -(J10 * dx[0])) * r;
        const float w0 = __fmaf_rn(J11, g[4], -(J01 * g[5])) * r;
        const float w1 = __fmaf_rn(J00, g[5], -(J10 * g[4])) * r;
        float d0 = e0, d1 = e1;
#pragma unroll
        for (int it = 0; it < PF2_LS_ITERS; ++it) {
            const float p = d0 * d1;
            d0 = __fmaf_rn(-w0, p, e0);
            d1 = __fmaf_rn(-w1, p, e1);
        }
        const float p = d0 * d1;
        const float n0 = __fmaf_rn(-w0, p, e0), n1 = __fmaf_rn(-w1, p, e1);
        const float r1n = fabsf(n0 - d0) + fabsf(n1 - d1), e1n = fabsf(e0) + fabsf(e1);
        dl[0] = n0;
        dl[1] = n1;
        return (det > 0.0f) & (r1n <= __fmaf_rn(PF2_LS_TOL, e1n, 1e-30f));
    } else {
        // 3-D: Q(delta) = D d0 d1 + E d0 d2 + F d1 d2 + G d0 d1 d2 with
        // D = d + g xi2, E = e + g xi1, F = f + g xi0, G = g (the xi-dependent
        // part of the trilinear term folded in), so the iteration reads
        // delta = e - W p with e = J^-1 dx, W = J^-1 [D E F G] and
        // p = (d0 d1, d0 d2, d1 d2, d0 d1 d2): 16 instructions per iteration
        // instead of ~45.  Same convergence test and fourth iterate as in 2-D.
        const float s12 = xi[1] * xi[2], s02 = xi[0] * xi[2], s01 = xi[0] * xi[1];
        float J[3][3], M[4][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float a = g[i], b = g[3 + i], c = g[6 + i];
            const float dd = g[9 + i], ee = g[12 + i], ff = g[15 + i], gg = g[18 + i];
            J[i][0] = __fmaf_rn(gg, s12, __fmaf_rn(ee, xi[2], __fmaf_rn(dd, xi[1], a)));
            J[i][1] = __fmaf_rn(gg, s02, __fmaf_rn(ff, xi[2], __fmaf_rn(dd, xi[0], b)));
            J[i][2] = __fmaf_rn(gg, s01, __fmaf_rn(ff, xi[1], __fmaf_rn(ee, xi[0], c)));
            M[0][i] = __fmaf_rn(gg, xi[2], dd);
            M[1][i] = __fmaf_rn(gg, xi[1], ee);
            M[2][i] = __fmaf_rn(gg, xi[0], ff);
            M[3][i] = gg;
        }
        float A[3][3];   // adjugate: J^-1 = A / det
        A[0][0] = __fmaf_rn(J[1][1], J[2][2], -(J[1][2] * J[2][1]));
        A[0][1] = __fmaf_rn(J[0][2], J[2][1], -(J[0][1] * J[2][2]));
        A[0][2] = __fmaf_rn(J[0][1], J[1][2], -(J[0][2] * J[1][1]));
        A[1][0] = __fmaf_rn(J[1][2], J[2][0], -(J[1][0] * J[2][2]));
        A[1][1] = __fmaf_rn(J[0][0], J[2][2], -(J[0][2] * J[2][0]));
        A[1][2] = __fmaf_rn(J[0][2], J[1][0], -(J[0][0] * J[1][2]));
        A[2][0] = __fmaf_rn(J[1][0], J[2][1], -(J[1][1] * J[2][0]));
        A[2][1] = __fmaf_rn(J[0][1], J[2][0], -(J[0][0] * J[2][1]));
        A[2][2] = __fmaf_rn(J[0][0], J[1][1], -(J[0][1] * J[1][0]));
        const float det = __fmaf_rn(J[0][2], A[2][0], __fmaf_rn(J[0][1], A[1][0], J[0][0] * A[0][0]));
        const float r = __builtin_amdgcn_rcpf(det);
        float e[3], W[4][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            e[i] = __fmaf_rn(A[i][2], dx[2], __fmaf_rn(A[i][1], dx[1], A[i][0] * dx[0])) * r;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                W[q][i] = __fmaf_rn(A[i][2], M[q][2], __fmaf_rn(A[i][1], M[q][1], A[i][0] * M[q][0])) * r;
        }
        float d0 = e[0], d1 = e[1], d2 = e[2];
        auto next = [&](float& n0, float& n1, float& n2) {
            const float p01 = d0 * d1, p02 = d0 * d2, p12 = d1 * d2, p012 = p01 * d2;
            float n[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                n[i] = __fmaf_rn(-W[3][i], p012,
                                 __fmaf_rn(-W[2][i], p12, __fmaf_rn(-W[1][i], p02, __fmaf_rn(-W[0][i], p01, e[i]))));
            n0 = n[0];
            n1 = n[1];
            n2 = n[2];
        };
#pragma unroll
        for (int it = 0; it < PF2_LS_ITERS; ++it) next(d0, d1, d2);
        float n0, n1, n2;
        next(n0, n1, n2);
        const float r1n = fabsf(n0 - d0) + fabsf(n1 - d1) + fabsf(n2 - d2);
        const float e1n = fabsf(e[0]) + fabsf(e[1]) + fabsf(e[2]);
        dl[0] = n0;
        dl[1] = n1;
        dl[2] = n2;
        return (det > 0.0f) & (r1n <= __fmaf_rn(PF2_LS_TOL, e1n, 1e-30f));
    }
}

// FP32 coefficients and nodal velocities of cell c (packed per-cell copies,
// QF<DIM, DEG>::ND nodes per cell)
template <int DIM, int DEG = 1>
__device__ __forceinline__ void load_cell32(const float* __restrict__ geo, const float* __restrict__ cellV32, int c,
                                            float* g, float (*Fc)[DIM]) {
    constexpr int ND = QF<DIM, DEG>::ND;
    constexpr int NG = DIM == 2 ? 6 : 21;
#pragma unroll
    for (int q = 0; q < NG; ++q) g[q] = geo[(size_t)c * NG + q];
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) Fc[k][d] = cellV32[((size_t)c * ND + k) * DIM + d];
}

// FP32 coefficients and nodal velocities of a candidate cell
template <int DIM, int DEG = 1>
struct PrefetchCell32 {
    const float* geo;
    const float* V32;
    float g[DIM == 2 ? 6 : 21];
    float Vc[QF<DIM, DEG>::ND][DIM];
    __device__ void load(int cc) { load_cell32<DIM, DEG>(geo, V32, cc, g, Vc); }
};

template <int DIM>
__device__ __forceinline__ int part_key32(const DevParams& pp, const float* xi) {
    double xd[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) xd[d] = (double)xi[d];
    return part_key<DIM>(pp, xd);
}

// multilinear coefficients of the nodal values of a cell (in place):
//   2-D  u = c0 + c1 xi0 + c2 xi1 + c3 xi0 xi1
//   3-D  u = c0 + c1 xi0 + c2 xi1 + c3 xi2 + c4 xi0 xi1 + c5 xi0 xi2 + c6 xi1 xi2 + c7 xi0 xi1 xi2
// so that a substep evaluates u with DIM nested FMA chains instead of the
// shape functions and a sum over the vertices
template <int DIM>
__device__ __forceinline__ void to_coeffs32(float (*c)[DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        const float v0 = c[0][d], v1 = c[1][d], v2 = c[2][d], v3 = c[3][d];
        c[1][d] = v1 - v0;
        c[2][d] = v2 - v0;
        c[3][d] = (v3 - v2) - (v1 - v0);
        if constexpr (DIM == 3) {
            const float v4 = c[4][d], v5 = c[5][d], v6 = c[6][d], v7 = c[7][d];
            c[3][d] = v4 - v0;
            c[4][d] = (v3 - v2) - (v1 - v0);
            c[5][d] = (v5 - v4) - (v1 - v0);
            c[6][d] = (v6 - v4) - (v2 - v0);
            c[7][d] = ((v7 - v6) - (v5 - v4)) - ((v3 - v2) - (v1 - v0));
        }
    }
}

template <int DIM>
__device__ __forceinline__ float eval_coeffs32(const float (*c)[DIM], int d, const float* xi) {
    if constexpr (DIM == 2)
        return __fmaf_rn(__fmaf_rn(c[3][d], xi[0], c[2][d]), xi[1], __fmaf_rn(c[1][d], xi[0], c[0][d]));
    else
        return __fmaf_rn(__fmaf_rn(__fmaf_rn(c[7][d], xi[0], c[6][d]), xi[1], __fmaf_rn(c[5][d], xi[0], c[3][d])),
                         xi[2],
                         __fmaf_rn(__fmaf_rn(c[4][d], xi[0], c[2][d]), xi[1], __fmaf_rn(c[1][d], xi[0], c[0][d])));
}

// A cell's FP32 field of degree DEG as monomial coefficients of its nodal
// values, evaluated with nested FMAs.  DEG 1: to_coeffs32 / eval_coeffs32.
// DEG 2: the nodal values in tensor order (index a_0 + 3 a_1 [+ 9 a_2], axis
// index a: t = 0, 1, 1/2), then along each direction
//   (f(0), f(1), f(1/2)) -> (c0, c1, c2),  f(t) = c0 + c1 t + c2 t^2,
//   c1 = 4 f(1/2) - 3 f(0) - f(1),  c2 = 2 (f(0) + f(1)) - 4 f(1/2),
// so that entry a_0 + 3 a_1 [+ 9 a_2] is the coefficient of
// xi0^a_0 xi1^a_1 [xi2^a_2].
template <int DIM, int DEG>
struct Field32 {
    static constexpr int ND = QF<DIM, DEG>::ND;
    float c[ND][DIM];
    __device__ __forceinline__ void init(const float (*Vc)[DIM]) {
        if constexpr (DEG == 1) {
#pragma unroll
            for (int k = 0; k < ND; ++k)
#pragma unroll
                for (int d = 0; d < DIM; ++d) c[k][d] = Vc[k][d];
            to_coeffs32<DIM>(c);
        } else {
#pragma unroll
            for (int k = 0; k < ND; ++k) {
                int t = 0, st = 1;
#pragma unroll
                for (int e = 0; e < DIM; ++e) {
                    t += q2_axis(DIM, k, e) * st;
                    st *= 3;
                }
#pragma unroll
                for (int d = 0; d < DIM; ++d) c[t][d] = Vc[k][d];
            }
#pragma unroll
            for (int e = 0, st = 1; e < DIM; ++e, st *= 3)
#pragma unroll
                for (int b = 0; b < ND; ++b) {
                    if ((b / st) % 3 != 0) continue;   // first entry of a line along direction e
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        const float f0 = c[b][d], f1 = c[b + st][d], fh = c[b + 2 * st][d];
                        c[b + st][d] = (4.0f * fh - 3.0f * f0) - f1;
                        c[b + 2 * st][d] = 2.0f * (f0 + f1) - 4.0f * fh;
                    }
                }
        }
    }
    // component d at xi
    __device__ __forceinline__ float eval(int d, const float* xi) const {
        if constexpr (DEG == 1) {
            return eval_coeffs32<DIM>(c, d, xi);
        } else {
            auto row = [&](int b) {   // c[b] + c[b+1] xi0 + c[b+2] xi0^2
                return __fmaf_rn(__fmaf_rn(c[b + 2][d], xi[0], c[b + 1][d]), xi[0], c[b][d]);
            };
            auto plane = [&](int b) {   // rows b, b+3, b+6 in xi1
                return __fmaf_rn(__fmaf_rn(row(b + 6), xi[1], row(b + 3)), xi[1], row(b));
            };
            if constexpr (DIM == 2) return plane(0);
            else return __fmaf_rn(__fmaf_rn(plane(18), xi[2], plane(9)), xi[2], plane(0));
        }
    }
};

// N_k of one node k of a degree-DEG cell in FP32 (a lane's node in the
// projection sums): DEG 1 as shape_one32; DEG 2 the product over the
// directions of the 1-D Lagrange factor (c0 t + c1)(c2 t + c3) of the node's
// axis index (lag2_coeffs)
template <int DIM, int DEG>
struct ShapeOne32 {
    float c[DIM][4];
    __device__ __forceinline__ void init(int k) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            if constexpr (DEG == 1) {
                const bool one = (k >> d) & 1;
                c[d][0] = one ? 1.0f : -1.0f;
                c[d][1] = one ? 0.0f : 1.0f;
            } else {
                double q[4];
                lag2_coeffs(q2_axis(DIM, k, d), q);
#pragma unroll
                for (int j = 0; j < 4; ++j) c[d][j] = (float)q[j];
            }
        }
    }
    __device__ __forceinline__ float factor(int d, float t) const {
        if constexpr (DEG == 1) return __fmaf_rn(c[d][0], t, c[d][1]);
        else return __fmaf_rn(c[d][0], t, c[d][1]) * __fmaf_rn(c[d][2], t, c[d][3]);
    }
    __device__ __forceinline__ float operator()(const float* xi) const {
        float v = factor(0, xi[0]);
#pragma unroll
        for (int d = 1; d < DIM; ++d) v *= factor(d, xi[d]);
        return v;
    }
};

// FP32 substeps until done (0) or the reference point leaves the unit cell (1);
// u (the last velocity) is returned for the relocation.  Vc: the cell's nodal
// velocities (QF<DIM, DEG>::ND nodes).
template <int DIM, int DEG = 1>
__device__ __forceinline__ int advance32(const DevParams& pp, float h32, int i, double* __restrict__ pvext,
                                         long long cap, float* xi, float* u, int& s, const float* g,
                                         const float (*Vc)[DIM]) {
    Field32<DIM, DEG> cf;
    cf.init(Vc);
    for (; s < pp.substeps; ++s) {
        float dx[DIM], dl[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) u[d] = cf.eval(d, xi);
#pragma unroll
        for (int d = 0; d < DIM; ++d) dx[d] = u[d] * h32;
        if (pp.keep_vext && s == pp.substeps - 1 && pvext)
#pragma unroll
            for (int d = 0; d < DIM; ++d) pvext[(size_t)d * cap + i] = (double)u[d];
        // one exit per substep (the local step has no early exit)
        bool in = local_step<DIM>(g, xi, dx, dl);
        float e[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            e[d] = xi[d] + dl[d];
            in = in & (e[d] >= 0.0f) & (e[d] <= 1.0f);
        }
        if (!in) return 1;
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = e[d];
    }
    return 0;
}

// (seg_end: particles from it on, new particles of a pending layout, arrive
// in their cell)
template <int DIM>
__device__ __forceinline__ void finish32(const DevParams& pp, const int* __restrict__ pcell, const Rec32& R,
                                         const DevWork& w, bool valid, int i, int c, const float* xi,
                                         int seg_end = 0x7fffffff) {
    add_arrival(w, valid && c >= 0 && (c != pcell[valid ? i : 0] || i >= seg_end), c, i);
    if (!valid) return;
    rec_store_xi<DIM>(R, i, xi);
    w.dest[i] = c;
    if (c >= 0) w.pkey[i] = part_key32<DIM>(pp, xi);
}

template <int DIM>
__device__ __forceinline__ void park32(const Rec32& R, long long cap, const DevWork& w, const DevQueues& qu,
                                       int round_out, int shard, bool park, int i, int c, const float* xi,
                                       const float* u, int s) {
    queue_push(qu, round_out, park, make_int2(i, c), shard);
    if (!park) return;
    qu.park_s[i] = s;
    if constexpr (DIM == 2) {
        // 2-D: reference point and velocity_ext in one 16-byte record (the
        // FP64 parking array has room for it), read back by the search
        reinterpret_cast<float4*>(qu.park_vext)[i] = make_float4(xi[0], xi[1], u[0], u[1]);
    } else {
        rec_store_xi<DIM>(R, i, xi);
#pragma unroll
        for (int d = 0; d < DIM; ++d) qu.park_vext[(size_t)d * cap + i] = (double)u[d];
    }
}

template <int DIM, int DEG = 1>
__device__ __forceinline__ void move32_body(const Grid& G, const DevMesh<DIM>& m, const float* __restrict__ V32, const float* __restrict__ dV32, const float* __restrict__ geo, const int* __restrict__ pcell, const Rec32& R, double* __restrict__ pvext, long long cap, const DevParams& pp, const DevWork& w, const DevQueues& qu, int do_correct, int do_move) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    constexpr int NG = DIM == 2 ? 6 : 21;
    // a pending layout (k_cell_pend32): as in move_body
    const int pn = w.counters[6];
    const int n = state_dead(w.counters) ? 0 : (pn > 0 ? pn : w.counters[0]);
    const int seg_end = pn > 0 ? w.counters[7] : 0x7fffffff;
    const float h32 = (float)pp.h;
    const int lane = threadIdx.x & (WAVE - 1);
    // the particles that exist now end this advection with xi from the
    // local step or an inverse map in their cell (see k_resort32)
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
                w.dest[i] = -1;
                valid = false;
            }
        } else if (valid) {
            c = pcell[i];
        }
        float xi[DIM], v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xi[d] = v[d] = 0.5f;
        if (valid) rec_load<DIM>(R, i, xi, v);
        if (do_correct && valid) {
            if constexpr (DEG == 1) {
                float N[NV];
                shape32<DIM>(xi, N);
#pragma unroll
                for (int k = 0; k < NV; ++k)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) v[d] = __fmaf_rn(N[k], dV32[((size_t)c * NV + k) * DIM + d], v[d]);
            } else {
                // the increment V - V_old at xi from its monomial coefficients
                float dVc[ND][DIM];
#pragma unroll
                for (int k = 0; k < ND; ++k)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) dVc[k][d] = dV32[((size_t)c * ND + k) * DIM + d];
                Field32<DIM, DEG> f;
                f.init(dVc);
#pragma unroll
                for (int d = 0; d < DIM; ++d) v[d] = v[d] + f.eval(d, xi);
            }
            rec_store_v<DIM>(R, i, v);
        }
        if (!do_move) continue;
        int st = 0, s = 0;
        float u[DIM];
        if (valid) {
            float g[NG], Vc[ND][DIM];
            load_cell32<DIM, DEG>(geo, V32, c, g, Vc);
            st = advance32<DIM, DEG>(pp, h32, i, pvext, cap, xi, u, s, g, Vc);
        }
        park32<DIM>(R, cap, w, qu, 0, L.shard, valid && st == 1, i, c, xi, u, s);
        finish32<DIM>(pp, pcell, R, w, valid && st == 0, i, c, xi, seg_end);
    }
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) k_move32(DevMesh<DIM> m, const float* __restrict__ V32,
                                                const float* __restrict__ dV32, const float* __restrict__ geo,
                                                const int* __restrict__ pcell, Rec32 R, double* __restrict__ pvext,
                                                long long cap, DevParams pp, DevWork w, DevQueues qu, int do_correct,
                                                int do_move) {
    move32_body<DIM, DEG>(hw_grid(), m, V32, dV32, geo, pcell, R, pvext, cap, pp, w, qu, do_correct, do_move);
}

// FP64 relocation of queued particles (the reference rule), then the
// remaining FP32 substeps
// (LAST: the last round, which loops every particle to its end.  Its
// particles cross several cells per step and the face-neighbour shortcut
// misses more often there: measured 4-8% slower with it.  It runs few
// wavefronts, so it takes the full register budget (2 per SIMD, no spills).
// The first round is latency-bound with many wavefronts: in 2-D 4 per SIMD
// (128 registers, 8 bytes per lane spilled) runs it 12% faster than 3
// without spills; in 3-D 3 per SIMD (212 bytes spilled) 3% faster than 2.)
template <int DIM, bool LAST, int DEG = 1>
__device__ __forceinline__ void search32_body(const Grid& G, const DevMesh<DIM>& m, const float* __restrict__ V32, const float* __restrict__ geo, const int* __restrict__ pcell, const Rec32& R, double* __restrict__ pvext, long long cap, const DevParams& pp, const DevWork& w, const DevQueues& qu, int round) {
    constexpr int NV = Q1<DIM>::NV;
    constexpr int ND = QF<DIM, DEG>::ND;
    constexpr int NG = DIM == 2 ? 6 : 21;
    __shared__ int s_pref[QSHARDS + 2];
    const int n = queue_prefix(qu, round, s_pref, G);
    const int seg_end = w.counters[6] > 0 ? w.counters[7] : 0x7fffffff;
    constexpr bool last = LAST;
    const float h32 = (float)pp.h;
    const int lane = threadIdx.x & (WAVE - 1);
    const WaveLoop L = range_loop(n, G);
    int lost = 0, searches = 0;
    for (int base = L.begin; base < L.end; base += L.step) {
        const int j = base + lane;
        const bool valid = j < L.end;
        const int2 e = valid ? queue_get(qu, round, s_pref, j) : make_int2(-1, -1);
        const int i = e.x;
        int c = e.y;
        int s = valid ? qu.park_s[i] : 0;
        int st = 0;
        float xi[DIM], u[DIM];
        if (valid) {
            if constexpr (DIM == 2) {
                const float4 pk = reinterpret_cast<const float4*>(qu.park_vext)[i];
                xi[0] = pk.x;
                xi[1] = pk.y;
                u[0] = pk.z;
                u[1] = pk.w;
            } else {
                rec_load_xi<DIM>(R, i, xi);
#pragma unroll
                for (int d = 0; d < DIM; ++d) u[d] = (float)qu.park_vext[(size_t)d * cap + i];
            }
            // 2-D: the found cell's vertices and FP32 data are loaded with the
            // candidate and reused; 3-D reloads them (register budget)
            double P[NV][DIM];
            load_vertices<DIM>(m, c, P);
            for (;;) {
                ++searches;
                double Q[NV][DIM], xid[DIM], x[DIM], vext[DIM], r[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    xid[d] = (double)xi[d];
                    vext[d] = (double)u[d];
                }
                forward<DIM>(P, xid, x);
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = x[d] + (double)(u[d] * h32);
                int cn;
                if constexpr (DIM == 2) {
                    PrefetchCell32<DIM, DEG> pre{geo, V32};
                    cn = relocate<DIM, PrefetchCell32<DIM, DEG>, !LAST>(m, c, P, xid, vext, x, r, Q, pre);
                    if (cn >= 0) {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            const float t = (float)r[d];
                            xi[d] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                        }
                        ++s;
                        st = advance32<DIM, DEG>(pp, h32, i, pvext, cap, xi, u, s, pre.g, pre.Vc);
                    }
                } else {
                    NoPrefetch pre;
                    cn = relocate<DIM>(m, c, P, xid, vext, x, r, Q, pre);
                    if (cn >= 0) {
#pragma unroll
                        for (int d = 0; d < DIM; ++d) {
                            const float t = (float)r[d];
                            xi[d] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                        }
                        float g[NG], Vc[ND][DIM];
                        load_cell32<DIM, DEG>(geo, V32, cn, g, Vc);
                        ++s;
                        st = advance32<DIM, DEG>(pp, h32, i, pvext, cap, xi, u, s, g, Vc);
                    }
                }
                if (cn < 0) {
                    c = -1;
                    ++lost;
                    st = 0;
                    break;
                }
                c = cn;
                if (st == 0 || !last) break;
                if constexpr (DIM == 2) {
#pragma unroll
                    for (int k = 0; k < NV; ++k)
#pragma unroll
                        for (int d = 0; d < DIM; ++d) P[k][d] = Q[k][d];
                } else {
                    load_vertices<DIM>(m, c, P);
                }
            }
        }
        if (!last) park32<DIM>(R, cap, w, qu, round + 1, L.shard, valid && st == 1, i, c, xi, u, s);
        finish32<DIM>(pp, pcell, R, w, valid && st == 0, i, c, xi, seg_end);
    }
    wave_stat(w.stats, 0, lost);
    wave_stat(w.stats, 3, searches);
}

template <int DIM, bool LAST, int DEG = 1>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LAST ? 2 : (DIM == 2 ? 4 : 3)))) k_search_resume32(DevMesh<DIM> m, const float* __restrict__ V32,
                                                         const float* __restrict__ geo, const int* __restrict__ pcell,
                                                         Rec32 R, double* __restrict__ pvext,
                                                         long long cap, DevParams pp, DevWork w, DevQueues qu,
                                                         int round) {
    search32_body<DIM, LAST, DEG>(hw_grid(), m, V32, geo, pcell, R, pvext, cap, pp, w, qu, round);
}

// deg: the velocity field's degree (1 or 2)
template <int DIM>
void launch_move32(hipStream_t s, int grid, DevMesh<DIM> m, const float* V32, const float* dV32, const float* geo,
                   int* pcell, Rec32 R, double* pvext, long long cap, DevParams pp, DevWork w,
                   DevQueues qu, int do_correct, int do_move, int deg) {
    if (deg == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_move32<DIM, 2>), dim3(grid), dim3(256), 0, s, m, V32, dV32, geo, pcell, R,
                           pvext, cap, pp, w, qu, do_correct, do_move);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_move32<DIM>), dim3(grid), dim3(256), 0, s, m, V32, dV32, geo, pcell, R,
                           pvext, cap, pp, w, qu, do_correct, do_move);
}

template <int DIM>
void launch_search_resume32(hipStream_t s, int grid, DevMesh<DIM> m, const float* V32, const float* geo,
                            const int* pcell, Rec32 R, double* pvext, long long cap, DevParams pp, DevWork w,
                            DevQueues qu, int round, int deg) {
    const bool last = round == MOVE_ROUNDS;
    if (deg == 2 && last)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume32<DIM, true, 2>), dim3(grid), dim3(256), 0, s, m, V32, geo,
                           pcell, R, pvext, cap, pp, w, qu, round);
    else if (deg == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume32<DIM, false, 2>), dim3(grid), dim3(256), 0, s, m, V32,
                           geo, pcell, R, pvext, cap, pp, w, qu, round);
    else if (last)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume32<DIM, true>), dim3(grid), dim3(256), 0, s, m, V32, geo,
                           pcell, R, pvext, cap, pp, w, qu, round);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_search_resume32<DIM, false>), dim3(grid), dim3(256), 0, s, m, V32, geo,
                           pcell, R, pvext, cap, pp, w, qu, round);
}

// rows of the fused projection sums (FCAP, pf2_kernels.hpp): a cell's final
// particles (records; in 3-D the second part of each record in a row of its
// own)

// Kahan-compensated FP32 projection sums of one cell over n staged records
// in order (accumulator a < NACC: vertex a / (DIM + 1), component a % (DIM +
// 1), the last one the weight), as k_cell_partials32 computes them
template <int DIM, int NA, int DEG = 1>
__device__ __forceinline__ void cell_sums32(const float4* __restrict__ rec, const float2* __restrict__ recb, int n,
                                            int a0, double* out) {
    // NA accumulators a0, a0 + GS, ... in one pass over the records (their
    // Kahan chains are independent, so they overlap); accumulators past
    // NACC are computed for node 0 and not stored by the caller
    constexpr int NACC = QF<DIM, DEG>::ND * (DIM + 1);
    ShapeOne32<DIM, DEG> sh[NA];
    float sum[NA], cmp[NA];
    int comp[NA];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = a0 + u * GS;
        const int kk = (a < NACC ? a : 0) / (DIM + 1);
        comp[u] = a - kk * (DIM + 1);
        sh[u].init(kk);
        sum[u] = 0.0f;
        cmp[u] = 0.0f;
    }
    for (int q = 0; q < n; ++q) {
        const float4 r = rec[q];
        float xi[DIM], v[DIM];
        if constexpr (DIM == 2) {
            xi[0] = r.x;
            xi[1] = r.y;
            v[0] = r.z;
            v[1] = r.w;
        } else {
            const float2 rb = recb[q];
            xi[0] = r.x;
            xi[1] = r.y;
            xi[2] = r.z;
            v[0] = r.w;
            v[1] = rb.x;
            v[2] = rb.y;
        }
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const float nk = sh[u](xi);
            float vc = 1.0f;   // the weight term (comp == DIM) is N_k itself
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                if (d == comp[u]) vc = v[d];
            const float term = nk * vc;
            const float y = term - cmp[u];
            const float tt = sum[u] + y;
            cmp[u] = (tt - sum[u]) - y;
            sum[u] = tt;
        }
    }
#pragma unroll
    for (int u = 0; u < NA; ++u) out[u] = (double)sum[u] - (double)cmp[u];
}

// FUSE (final count per cell <= FCAP): the group also stages the cell's
// final records in LDS and writes its projection sums, so no separate
// partials pass re-reads the particles (3-D: each lane of the group owns two
// of the 32 sums)
template <int DIM, bool FUSE, int DEG = 1>
__device__ __forceinline__ void cell_write32_one(const DevMesh<DIM>& m, const float* __restrict__ V32c,
                                                 const int* __restrict__ old_off, const uint32_t* __restrict__ aid,
                                                 const Rec32& A, int* __restrict__ bcell, uint32_t* __restrict__ bid,
                                                 const Rec32& B, const double* __restrict__ avext,
                                                 double* __restrict__ bvext, long long cap, const DevParams& pp,
                                                 const DevWork& w, int c, int gl, float4* __restrict__ srec,
                                                 float2* __restrict__ srecb, double* __restrict__ partial) {
    constexpr int ND = QF<DIM, DEG>::ND;
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
        return;
    }
    auto put = [&](int idx, int pos) {
        const long long o = out0 + pos;
        bcell[o] = c;
        bid[o] = aid[idx];
        if constexpr (FUSE) {
            const float4 r = A.a[idx];
            B.a[o] = r;
            srec[pos] = r;
            if constexpr (DIM == 3) {
                const float2 rb = A.b[idx];
                B.b[o] = rb;
                srecb[pos] = rb;
            }
        } else {
            rec_copy<DIM>(A, idx, B, o);
        }
        if (pp.keep_vext)
#pragma unroll
            for (int d = 0; d < DIM; ++d) bvext[(size_t)d * cap + o] = avext[(size_t)d * cap + idx];
    };
    survivors_of(w, old_off, c, n_st, kin, nsurv, gl, put);
    const int spawns = nfinal - nsurv;
    if (spawns > 0) {
        const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
        const unsigned long long empty = all & ~w.hitmask[c];
        const uint32_t id0 = *w.particle_counter + (uint32_t)fo;
        for (int o = gl; o < pp.nparts_in; o += GS) {
            if (!((empty >> o) & 1)) continue;
            const int r = __popcll(empty & ((1ull << o) - 1));
            // velocity from the cell's FP32 copy of V (this handle keeps no
            // FP64 per-cell copy), interpolated in FP64
            double xi[DIM], Vc[ND][DIM], v[DIM];
            lattice_point<DIM>(pp, o, xi);
#pragma unroll
            for (int k = 0; k < ND; ++k)
#pragma unroll
                for (int d = 0; d < DIM; ++d) Vc[k][d] = (double)V32c[((size_t)c * ND + k) * DIM + d];
            interp<DIM, DEG>(Vc, xi, v);
            const long long q = out0 + nsurv + r;
            bcell[q] = c;
            bid[q] = id0 + (uint32_t)r + 1u;
            float xf[DIM], vf[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xf[d] = (float)xi[d];
                vf[d] = (float)v[d];
                if (pp.keep_vext) bvext[(size_t)d * cap + q] = 0.0;
            }
            rec_store<DIM>(B, q, xf, vf);
            if constexpr (FUSE) {
                if constexpr (DIM == 2) {
                    srec[nsurv + r] = make_float4(xf[0], xf[1], vf[0], vf[1]);
                } else {
                    srec[nsurv + r] = make_float4(xf[0], xf[1], xf[2], vf[0]);
                    srecb[nsurv + r] = make_float2(vf[1], vf[2]);
                }
            }
        }
    }
    if (gl == 0) {
        w.arr_cnt[c] = 0;
        w.ovf_head[c] = -1;
    }
    if constexpr (FUSE) {
        constexpr int NACC = PartialsShape<DIM, DEG>::NACC;
        wave_sync();
        // sums per lane: Q1 1 in 2-D (12 lanes), 2 in 3-D; Q2 2 in 2-D, 7 in 3-D
        constexpr int NA = (NACC + GS - 1) / GS;
        if (gl < NACC) {
            double out[NA];
            cell_sums32<DIM, NA, DEG>(srec, srecb, nfinal, gl, out);
#pragma unroll
            for (int u = 0; u < NA; ++u)
                if (gl + u * GS < NACC) partial[(size_t)c * NACC + gl + u * GS] = out[u];
        }
        wave_sync();
    }
}

template <int DIM, bool FUSE, int DEG = 1>
__device__ __forceinline__ void cell_write32_body(const Grid& G, const DevMesh<DIM>& m, const float* __restrict__ V32c, const int* __restrict__ old_off, const uint32_t* __restrict__ aid, const Rec32& A, int* __restrict__ bcell, uint32_t* __restrict__ bid, const Rec32& B, const double* __restrict__ avext, double* __restrict__ bvext, long long cap, const DevParams& pp, const DevWork& w, double* __restrict__ partial) {
    constexpr bool FB = FUSE && DIM == 3;
    __shared__ float4 s_rec[FUSE ? GROUPS : 1][FUSE ? FCAP : 1];
    __shared__ float2 s_recb[FB ? GROUPS : 1][FB ? FCAP : 1];
    if (state_dead(w.counters)) return;
    const int g = threadIdx.x / GS, gl = threadIdx.x % GS;
    for (int c = G.bid * GROUPS + g; c < m.n_cells; c += G.nblk * GROUPS)
        cell_write32_one<DIM, FUSE, DEG>(m, V32c, old_off, aid, A, bcell, bid, B, avext, bvext, cap, pp, w, c, gl,
                                         s_rec[FUSE ? g : 0], s_recb[FB ? g : 0], partial);
}

template <int DIM, bool FUSE, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_write32(DevMesh<DIM> m, const float* __restrict__ V32c,
                                                     const int* __restrict__ old_off, const uint32_t* __restrict__ aid,
                                                     Rec32 A, int* __restrict__ bcell, uint32_t* __restrict__ bid,
                                                     Rec32 B, const double* __restrict__ avext,
                                                     double* __restrict__ bvext, long long cap, DevParams pp,
                                                     DevWork w, double* __restrict__ partial) {
    cell_write32_body<DIM, FUSE, DEG>(hw_grid(), m, V32c, old_off, aid, A, bcell, bid, B, avext, bvext, cap, pp, w,
                                      partial);
}

// The second pass of the two-pass regroup write (k_cell_map, pf2_kernels.hpp)
// for FP32 records: output positions in order (GU slices per pass, loads
// before stores), a survivor's record copied (its cell is its destination),
// a new particle created as k_cell_write32 creates it.  Every wavefront
// stores whole, consecutive lines.
template <int DIM, int DEG>
__device__ __forceinline__ void gather32_spawn(const float* __restrict__ V32c, int* __restrict__ bcell,
                                               uint32_t* __restrict__ bid, const Rec32& B, double* __restrict__ bvext,
                                               long long cap, const DevParams& pp, const DevWork& w, long long o,
                                               int code) {
    constexpr int ND = QF<DIM, DEG>::ND;
    const int c = code >> 6, lo = code & 63;
    const unsigned long long fo = w.fs_off[c];
    const int r = (int)(o - (long long)(fo >> 32)) - w.nsurv[c];
    // velocity from the cell's FP32 copy of V, interpolated in FP64
    double xi[DIM], Vc[ND][DIM], v[DIM];
    lattice_point<DIM>(pp, lo, xi);
#pragma unroll
    for (int k = 0; k < ND; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) Vc[k][d] = (double)V32c[((size_t)c * ND + k) * DIM + d];
    interp<DIM, DEG>(Vc, xi, v);
    float xf[DIM], vf[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        xf[d] = (float)xi[d];
        vf[d] = (float)v[d];
    }
    bcell[o] = c;
    bid[o] = *w.particle_counter + (uint32_t)fo + (uint32_t)r + 1u;
    rec_store<DIM>(B, o, xf, vf);
    if (pp.keep_vext)
#pragma unroll
        for (int d = 0; d < DIM; ++d) bvext[(size_t)d * cap + o] = 0.0;
}

template <int DIM, int DEG = 1>
__device__ __forceinline__ void cell_gather32_body(const Grid& G, const DevMesh<DIM>& m, const float* __restrict__ V32c,
                                                   const uint32_t* __restrict__ aid, const Rec32& A,
                                                   int* __restrict__ bcell, uint32_t* __restrict__ bid, const Rec32& B,
                                                   const double* __restrict__ avext, double* __restrict__ bvext,
                                                   long long cap, const DevParams& pp, const DevWork& w,
                                                   const int* __restrict__ omap) {
    if (state_dead(w.counters)) return;
    const long long total = (long long)(w.fs_off[m.n_cells] >> 32);
    const long long n = total < cap ? total : cap;
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
        float4 ra[GU];
        float2 rb[GU];
        double pe[GU][DIM];
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            if (!act[u] || src[u] < 0) continue;
            const int s = src[u];
            pc[u] = w.dest[s];
            pid[u] = aid[s];
            ra[u] = A.a[s];
            if constexpr (DIM == 3) rb[u] = A.b[s];
#pragma unroll
            for (int d = 0; d < DIM; ++d) pe[u][d] = pp.keep_vext ? avext[(size_t)d * cap + s] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            if (!act[u]) continue;
            const long long o = o0 + u * bd;
            if (src[u] < 0) {
                gather32_spawn<DIM, DEG>(V32c, bcell, bid, B, bvext, cap, pp, w, o, -1 - src[u]);
                continue;
            }
            bcell[o] = pc[u];
            bid[o] = pid[u];
            B.a[o] = ra[u];
            if constexpr (DIM == 3) B.b[o] = rb[u];
            if (pp.keep_vext)
#pragma unroll
                for (int d = 0; d < DIM; ++d) bvext[(size_t)d * cap + o] = pe[u][d];
        }
    }
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_gather32(DevMesh<DIM> m, const float* __restrict__ V32c,
                                                       const uint32_t* __restrict__ aid, Rec32 A,
                                                       int* __restrict__ bcell, uint32_t* __restrict__ bid, Rec32 B,
                                                       const double* __restrict__ avext, double* __restrict__ bvext,
                                                       long long cap, DevParams pp, DevWork w,
                                                       const int* __restrict__ omap) {
    cell_gather32_body<DIM, DEG>(hw_grid(), m, V32c, aid, A, bcell, bid, B, avext, bvext, cap, pp, w, omap);
}

// fuse = the projection sums are computed here (final counts per cell
// bounded by FCAP, i.e. MAX * prod(q_d + 1) <= FCAP and a census with reseed)
template <int DIM>
void launch_cell_write32(hipStream_t s, int grid, DevMesh<DIM> m, const float* V32c, const int* old_off, const uint32_t* aid,
                         Rec32 A, int* bcell, uint32_t* bid, Rec32 B, const double* avext, double* bvext, long long cap,
                         DevParams pp, DevWork w, bool fuse, double* partial, int deg) {
    if (fuse && deg == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write32<DIM, true, 2>), dim3(grid), dim3(256), 0, s, m, V32c,
                           old_off, aid, A, bcell, bid, B, avext, bvext, cap, pp, w, partial);
    else if (deg == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write32<DIM, false, 2>), dim3(grid), dim3(256), 0, s, m, V32c,
                           old_off, aid, A, bcell, bid, B, avext, bvext, cap, pp, w, partial);
    else if (fuse)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write32<DIM, true>), dim3(grid), dim3(256), 0, s, m, V32c, old_off,
                           aid, A, bcell, bid, B, avext, bvext, cap, pp, w, partial);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_write32<DIM, false>), dim3(grid), dim3(256), 0, s, m, V32c, old_off,
                           aid, A, bcell, bid, B, avext, bvext, cap, pp, w, partial);
}

// per-cell projection sums in FP32 with Kahan compensation, ascending id order
// (groups of lanes per cell as in k_cell_partials)
template <int DIM, int DEG = 1>
__device__ __forceinline__ void partials32_body(const Grid& G, int n_cells, const int* __restrict__ off, const Rec32& R, double* __restrict__ partial, const int* __restrict__ counters) {
    if (state_dead(counters)) return;
    constexpr int NACC = PartialsShape<DIM, DEG>::NACC;
    constexpr int PGS = PartialsShape<DIM, DEG>::PGS;
    constexpr int NG = PartialsShape<DIM, DEG>::NG;
    constexpr int NA = PartialsShape<DIM, DEG>::NA;
    __shared__ float s_xi[NG][PGS][DIM];
    __shared__ float s_v[NG][PGS][DIM];
    const int g = threadIdx.x / PGS, gl = threadIdx.x % PGS;
    // lane gl holds accumulators gl, gl + PGS, ... (a = k * (DIM + 1) + comp)
    ShapeOne32<DIM, DEG> sh[NA];
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
        float sum[NA], cmp[NA];
#pragma unroll
        for (int u = 0; u < NA; ++u) sum[u] = cmp[u] = 0.0f;
        for (int t = 0; t < trips; ++t) {
            const int j0 = b + t * PGS;
            const int j = j0 + gl;
            if (j < e) {
                float xi[DIM], v[DIM];
                rec_load<DIM>(R, j, xi, v);
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    s_xi[g][gl][d] = xi[d];
                    s_v[g][gl][d] = v[d];
                }
            }
            wave_sync();
            const int cnt = (e - j0) < PGS ? (e - j0) : PGS;
            if (gl < NACC)
                for (int q = 0; q < cnt; ++q) {
                    float xi[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = s_xi[g][q][d];
#pragma unroll
                    for (int u = 0; u < NA; ++u) {
                        const float nk = sh[u](xi);
                        float vc = 1.0f;   // the weight term (comp == DIM) is N_k itself
#pragma unroll
                        for (int d = 0; d < DIM; ++d)
                            if (d == comp[u]) vc = s_v[g][q][d];
                        const float term = nk * vc;
                        const float y = term - cmp[u];
                        const float tt = sum[u] + y;
                        cmp[u] = (tt - sum[u]) - y;
                        sum[u] = tt;
                    }
                }
            wave_sync();
        }
#pragma unroll
        for (int u = 0; u < NA; ++u)
            if (cv && gl + u * PGS < NACC) partial[(size_t)c * NACC + gl + u * PGS] = (double)sum[u] - (double)cmp[u];
    }
}

template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_partials32(int n_cells, const int* __restrict__ off, Rec32 R,
                                                         double* __restrict__ partial, const int* __restrict__ counters) {
    partials32_body<DIM, DEG>(hw_grid(), n_cells, off, R, partial, counters);
}

template <int DIM>
void launch_partials32(hipStream_t s, int grid, int n_cells, const int* off, Rec32 R, double* partial,
                       const int* counters, int deg = 1) {
    if (deg == 2)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_partials32<DIM, 2>), dim3(grid), dim3(256), 0, s, n_cells, off, R,
                           partial, counters);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cell_partials32<DIM>), dim3(grid), dim3(256), 0, s, n_cells, off, R,
                           partial, counters);
}

template <int DIM>
__global__ void k_wall_particles32(DevMesh<DIM> m, DevWalls wl, const int* __restrict__ off, Rec32 R) {
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
        float xi[DIM], v[DIM], N[NV];
        rec_load<DIM>(R, j, xi, v);
        shape32<DIM>(xi, N);
        for (int q = 0; q < nw; ++q) {
            const int wi = widx[q];
            float nk = N[0];
#pragma unroll
            for (int r = 1; r < NV; ++r)
                if (r == order[q]) nk = N[r];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                if ((wl.mask[wi] >> d) & 1) v[d] = __fmaf_rn(nk, (float)(wl.bc[wi * 3 + d] - wl.val[wi * 3 + d]), v[d]);
        }
        rec_store_v<DIM>(R, j, v);
    }
}

template <int DIM>
void launch_wall_particles32(hipStream_t s, DevMesh<DIM> m, DevWalls wl, const int* off, Rec32 R) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wall_particles32<DIM>), dim3((wl.n_cells + 3) / 4), dim3(256), 0, s, m, wl,
                       off, R);
}

// per-cell FP32 copies of V and V - V_old at the cell's DoFs
template <int DIM>
__global__ void k_pack_fields32(DevMesh<DIM> m, const double* __restrict__ V, const double* __restrict__ Vold,
                                float* __restrict__ cellV32, float* __restrict__ cellDV32) {
    const long long n = (long long)m.n_cells * m.nd;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int dof = m.cd[t];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const double a = V[(size_t)dof * DIM + d], b = Vold[(size_t)dof * DIM + d];
            cellV32[(size_t)t * DIM + d] = (float)a;
            cellDV32[(size_t)t * DIM + d] = (float)(a - b);
        }
    }
}

template <int DIM>
__global__ void k_state_to32(const double* __restrict__ xi, const double* __restrict__ v, Rec32 R, long long cap,
                             const int* __restrict__ counters) {
    const long long n = state_dead(counters) ? 0 : counters[0];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float xf[DIM], vf[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xf[d] = (float)xi[(size_t)d * cap + i];
            vf[d] = (float)v[(size_t)d * cap + i];
        }
        rec_store<DIM>(R, i, xf, vf);
    }
}

inline void launch_state_to32(hipStream_t s, int dim, const double* xi, const double* v, Rec32 R, long long cap,
                              const int* counters, long long n_est) {
    long long g = (n_est + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    if (dim == 2) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_state_to32<2>), dim3((int)g), dim3(256), 0, s, xi, v, R, cap, counters);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_state_to32<3>), dim3((int)g), dim3(256), 0, s, xi, v, R, cap, counters);
}

// FP64 view of the FP32 state: xi, v widened, location mapped from (cell, xi)
template <int DIM>
__global__ void k_state_from32(DevMesh<DIM> m, const int* __restrict__ pcell, Rec32 R, double* __restrict__ x,
                               double* __restrict__ xi, double* __restrict__ v, long long cap,
                               const int* __restrict__ counters) {
    constexpr int NV = Q1<DIM>::NV;
    const long long n = state_dead(counters) ? 0 : counters[0];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double xd[DIM], P[NV][DIM], xx[DIM];
        float xf[DIM], vf[DIM];
        rec_load<DIM>(R, i, xf, vf);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            xd[d] = (double)xf[d];
            xi[(size_t)d * cap + i] = xd[d];
            v[(size_t)d * cap + i] = (double)vf[d];
        }
        load_vertices<DIM>(m, pcell[i], P);
        forward<DIM>(P, xd, xx);
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[(size_t)d * cap + i] = xx[d];
    }
}

template <int DIM>
void launch_state_from32(hipStream_t s, DevMesh<DIM> m, const int* pcell, Rec32 R, double* x, double* xi, double* v,
                         long long cap, const int* counters, long long n_est) {
    long long g = (n_est + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_state_from32<DIM>), dim3((int)g), dim3(256), 0, s, m, pcell, R, x, xi, v,
                       cap, counters);
}

// FP64 location F_K(xi) of every FP32-state particle on the current
// geometry (before the vertices move)
template <int DIM>
__global__ void k_pos32(DevMesh<DIM> m, const int* __restrict__ pcell, Rec32 R, double* __restrict__ x,
                        long long cap, const int* __restrict__ counters) {
    constexpr int NV = Q1<DIM>::NV;
    const long long n = state_dead(counters) ? 0 : counters[0];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float xf[DIM];
        double xd[DIM], P[NV][DIM], xx[DIM];
        rec_load_xi<DIM>(R, i, xf);
#pragma unroll
        for (int d = 0; d < DIM; ++d) xd[d] = (double)xf[d];
        load_vertices<DIM>(m, pcell[i], P);
        forward<DIM>(P, xd, xx);
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[(size_t)d * cap + i] = xx[d];
    }
}

// FP32-mixed re-sort after a mesh move, with the FP64 positions of k_pos32:
// the own-cell test in FP64 (2-D: inverse_map's closed form; the rest of the
// test is left pending, as in the FP64 re-sort), and only the reference
// coordinate of a particle that stays is stored back (clamped FP32).  A
// particle that fails the test is queued for k_resort_search32 (its x,
// reference point and velocity_ext stay where they are).  skip_static as in
// k_resort: a particle of a cell that kept its vertices, whose xi came from
// the local step or an inverse map in that cell, keeps its record.
template <int DIM>
__global__ void __launch_bounds__(256) k_resort32(DevMesh<DIM> m, const int* __restrict__ pcell,
                                                  const uint32_t* __restrict__ pid, const double* __restrict__ px,
                                                  Rec32 R, long long cap, DevParams pp, DevWork w, DevQueues qu,
                                                  int skip_static) {
    constexpr int NV = Q1<DIM>::NV;
    const int n = state_dead(w.counters) ? 0 : w.counters[0];
    const uint32_t id_exact = (uint32_t)w.counters[5];
    const int lane = threadIdx.x & (WAVE - 1);
    const WaveLoop L = stride_loop(n);
    for (int base = L.begin; base < L.end; base += L.step) {
        const int i = base + lane;
        const bool valid = i < L.end;
        const int c = valid ? pcell[i] : 0;
        int st = 0;   // 0 stays, 1 relocate, 2 own-cell test pending
        if (valid) {
            float xf[DIM];
            rec_load_xi<DIM>(R, i, xf);
            if (!(skip_static && !m.cmoved[c] && pid[i] <= id_exact)) {
                double x[DIM], r[DIM], P[NV][DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) x[d] = px[(size_t)d * cap + i];
                load_vertices<DIM>(m, c, P);
                if constexpr (DIM == 2) {
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
                        const float t = (float)r[d];
                        xf[d] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                    }
                    rec_store_xi<DIM>(R, i, xf);
                }
            }
            if (st == 0) {
                w.dest[i] = c;
                w.pkey[i] = part_key32<DIM>(pp, xf);
            } else {
                qu.park_s[i] = st == 2 ? PARK_NEWTON : 0;
            }
        }
        queue_push(qu, RESORT_QUEUE, valid && st != 0, make_int2(i, c), L.shard);
    }
}

// the relocations of the FP32-mixed re-sort (queue RESORT_QUEUE): a pending
// own-cell test first, then the vertex-star rule with the stored
// velocity_ext, in FP64; one search per particle, one queue entry per lane
template <int DIM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIM == 2 ? 3 : 2))) k_resort_search32(DevMesh<DIM> m, const int* __restrict__ pcell, const double* __restrict__ px, const double* __restrict__ pvext, Rec32 R, long long cap, DevParams pp, DevWork w, DevQueues qu) {
    constexpr int NV = Q1<DIM>::NV;
    __shared__ int s_pref[QSHARDS + 2];
    const int n = queue_prefix(qu, RESORT_QUEUE, s_pref);
    const int lane = threadIdx.x & (WAVE - 1);
    const WaveLoop L = range_loop(n);
    int lost = 0, searches = 0;
    for (int base = L.begin; base < L.end; base += L.step) {
        const int j = base + lane;
        const bool valid = j < L.end;
        const int2 e = valid ? queue_get(qu, RESORT_QUEUE, s_pref, j) : make_int2(0, -1);
        const int i = e.x;
        int c = e.y;
        if (valid) {
            double x[DIM], xi[DIM], vext[DIM], r[DIM], P[NV][DIM];
            float xf[DIM];
            rec_load_xi<DIM>(R, i, xf);
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                xi[d] = (double)xf[d];
                x[d] = px[(size_t)d * cap + i];
                vext[d] = pvext ? pvext[(size_t)d * cap + i] : 0.0;
            }
            load_vertices<DIM>(m, c, P);
            const bool own = qu.park_s[i] == PARK_NEWTON && !clearly_outside<DIM>(P, x, m.ftol[c]) &&
                             inverse_newton<DIM>(P, x, r) && inside_unit<DIM>(r);
            if (!own) {
                ++searches;
                double Q[NV][DIM];
                NoPrefetch pre;
                c = relocate<DIM>(m, c, P, xi, vext, x, r, Q, pre);
                if (c < 0) ++lost;
            }
            if (c >= 0) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    const float t = (float)r[d];
                    xf[d] = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                }
                rec_store_xi<DIM>(R, i, xf);
                w.pkey[i] = part_key32<DIM>(pp, xf);
            }
            w.dest[i] = c;
        }
        add_arrival(w, valid && c >= 0 && c != pcell[valid ? i : 0], c, i);
    }
    wave_stat(w.stats, 0, lost);
    wave_stat(w.stats, 3, searches);
}

// the pending layout of an FP32-mixed handle after a re-sort (k_cell_pend):
// new particles created as k_cell_gather32 creates them, after the current
// segments (the census marked the survivors)
template <int DIM, int DEG = 1>
__global__ void __launch_bounds__(256) k_cell_pend32(int n_cells, const int* __restrict__ old_off, int* __restrict__ pcell,
                                                     uint32_t* __restrict__ pid, Rec32 R, double* __restrict__ pvext,
                                                     const float* __restrict__ V32, long long cap, DevParams pp,
                                                     DevWork w) {
    constexpr int ND = QF<DIM, DEG>::ND;
    if (state_dead(w.counters)) return;
    const int nc = n_cells;
    const int seg_end = old_off[nc];
    const unsigned long long ftot = w.fs_off[nc];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const long long slots = (long long)seg_end + (uint32_t)ftot;
        if (slots > cap) atomicOr(&w.counters[2], 1);
        w.counters[6] = slots > cap ? 0 : (int)slots;
        w.counters[7] = seg_end;
    }
    // one cell per thread, as in k_cell_pend
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if ((uint32_t)w.fs_cnt[c] != 0u) {   // spawns, in lattice order
            const unsigned long long fo = w.fs_off[c];
            const unsigned long long all = pp.nparts_in >= 64 ? ~0ull : ((1ull << pp.nparts_in) - 1);
            const unsigned long long empty = all & ~w.hitmask[c];
            double Vc[ND][DIM];
#pragma unroll
            for (int k = 0; k < ND; ++k)
#pragma unroll
                for (int d = 0; d < DIM; ++d) Vc[k][d] = (double)V32[((size_t)c * ND + k) * DIM + d];
            int r = 0;
            for (unsigned long long e = empty; e; e &= e - 1, ++r) {
                const int o = __ffsll((long long)e) - 1;
                const long long q = (long long)seg_end + (uint32_t)fo + r;
                if (q >= cap) continue;   // (flagged above)
                double xi[DIM], v[DIM];
                lattice_point<DIM>(pp, o, xi);
                interp<DIM, DEG>(Vc, xi, v);
                float xf[DIM], vf[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    xf[d] = (float)xi[d];
                    vf[d] = (float)v[d];
                }
                pcell[q] = c;
                pid[q] = *w.particle_counter + (uint32_t)fo + (uint32_t)r + 1u;
                w.dest[q] = c | DEST_KEPT;
                rec_store<DIM>(R, q, xf, vf);
                if (pp.keep_vext)
#pragma unroll
                    for (int d = 0; d < DIM; ++d) pvext[(size_t)d * cap + q] = 0.0;
            }
        }
        w.arr_cnt[c] = 0;
        w.ovf_head[c] = -1;
    }
}

// k_settle for the FP32 state
template <int DIM>
__global__ void k_settle32(const int* __restrict__ pcell, Rec32 R, DevParams pp, DevWork w) {
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
                float xf[DIM];
                rec_load_xi<DIM>(R, i, xf);
                w.pkey[i] = part_key32<DIM>(pp, xf);
            }
        }
        add_arrival(w, valid && c >= 0 && (c != pcell[valid ? i : 0] || i >= seg_end), c, i);
    }
}

}  // namespace pf2
