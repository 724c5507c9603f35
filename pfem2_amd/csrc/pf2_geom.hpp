// pf2_geom.hpp — Q1 geometry of quadrilateral / hexahedral cells on the device.
//
// These are the deal.II primitives the particle stage calls (the library is not
// available to this build, so they are written from their definitions):
//   FE_Q<dim>(1)::shape_value            cylinder2D/pfem2particle.cpp:501,594,625,670
//   MappingQ1::transform_unit_to_real    cylinder2D/pfem2particle.cpp:123,497,536
//   MappingQ1::transform_real_to_unit    cylinder2D/pfem2particle.cpp:228,320
//   GeometryInfo::is_inside_unit_cell    cylinder2D/pfem2particle.cpp:230,321
// Arithmetic contract: every expression is evaluated in the order written,
// without fused multiply-add (the library is compiled with -ffp-contract=off),
// so that a cell, a reference point and a location give the same bits as a
// host computation in IEEE double.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace pf2 {

#define PF2_HD __host__ __device__ __forceinline__

template <int DIM> struct Q1 { static constexpr int NV = 1 << DIM; };

PF2_HD double dmax(double a, double b) { return (a < b) ? b : a; }
PF2_HD double dmin(double a, double b) { return (b < a) ? b : a; }

// N_k(xi) = prod_d L_{bit_d(k)}(xi_d), L0(t) = 1 - t, L1(t) = t
template <int DIM>
PF2_HD void shape(const double* xi, double* N) {
#pragma unroll
    for (int k = 0; k < Q1<DIM>::NV; ++k) {
        double val = 1.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) val *= ((k >> d) & 1) ? xi[d] : (1.0 - xi[d]);
        N[k] = val;
    }
}

// ---------------------------------------------------------------------------
// Velocity fields of degree DEG on the Q1 geometry.  DEG = 1: the vertex
// values above.  DEG = 2 (FE_Q<dim>(2)): 3^dim nodes per cell in deal.II's
// order -- vertices (lexicographic), then lines, then (3-D) faces, then the
// interior.  Node k sits at unit coordinates (t(a_0), t(a_1)[, t(a_2)]) with
// axis index a in {0, 1, 2} for t = 0, 1, 1/2, and
//   N_k(xi) = L_{a_0}(xi_0) L_{a_1}(xi_1) [L_{a_2}(xi_2)]   (left to right)
//   L_0(t) = (1 - t)(1 - 2t),  L_1(t) = t (2t - 1),  L_2(t) = 4t (1 - t)
// each factor evaluated as written (two roundings and one product).
// 2-D lines: 0 x=0, 1 x=1, 2 y=0, 3 y=1.  3-D lines: 0-3 those of the face
// z=0, 4-7 of z=1, 8-11 parallel to z at the vertices 0-3; faces 0-5:
// x=0, x=1, y=0, y=1, z=0, z=1.
template <int DIM, int DEG>
struct QF {
    static constexpr int ND = DEG == 1 ? (1 << DIM) : (DIM == 2 ? 9 : 27);
};

// axis index of node k (of a degree-2 cell) in direction d
PF2_HD int q2_axis(int dim, int k, int d) {
    // 2-D: 9 nodes x 2 axes; 3-D: 27 nodes x 3 axes
    constexpr signed char A2[9][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2}, {2, 0}, {2, 1}, {2, 2}};
    constexpr signed char A3[27][3] = {
        {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},   // vertices
        {0, 2, 0}, {1, 2, 0}, {2, 0, 0}, {2, 1, 0},                                           // lines z = 0
        {0, 2, 1}, {1, 2, 1}, {2, 0, 1}, {2, 1, 1},                                           // lines z = 1
        {0, 0, 2}, {1, 0, 2}, {0, 1, 2}, {1, 1, 2},                                           // lines along z
        {0, 2, 2}, {1, 2, 2}, {2, 0, 2}, {2, 1, 2}, {2, 2, 0}, {2, 2, 1},                     // faces
        {2, 2, 2}};                                                                           // interior
    return dim == 2 ? A2[k][d] : A3[k][d];
}

// the 1-D Lagrange factor of axis index a as p(t) * r(t) with
// p = c[0] t + c[1] and r = c[2] t + c[3]; every product c t is exact
// (c in {+-1, +-2, 4}), so fma gives the expression of the definition
PF2_HD void lag2_coeffs(int a, double* c) {
    c[0] = a == 0 ? -1.0 : (a == 1 ? 1.0 : 4.0);
    c[1] = a == 0 ? 1.0 : 0.0;
    c[2] = a == 0 ? -2.0 : (a == 1 ? 2.0 : -1.0);
    c[3] = a == 0 ? 1.0 : (a == 1 ? -1.0 : 1.0);
}

PF2_HD double lag2(int a, double t) {
    if (a == 0) return (1.0 - t) * (1.0 - 2.0 * t);
    if (a == 1) return t * (2.0 * t - 1.0);
    return (4.0 * t) * (1.0 - t);
}

// N_k(xi) of every node, k ascending
template <int DIM, int DEG>
PF2_HD void shape_q(const double* xi, double* N) {
    if constexpr (DEG == 1) {
        shape<DIM>(xi, N);
    } else {
        double L[DIM][3];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int a = 0; a < 3; ++a) L[d][a] = lag2(a, xi[d]);
#pragma unroll
        for (int k = 0; k < QF<DIM, 2>::ND; ++k) {
            double val = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) val *= L[d][q2_axis(DIM, k, d)];
            N[k] = val;
        }
    }
}

// N_k of one node k chosen at run time (a lane's node in the projection
// sums): per direction the factor coefficients, then the product in
// shape_q's order, so that the value equals shape_q's N[k]
template <int DIM, int DEG>
struct ShapeOne {
    double c[DIM][4];
    __host__ __device__ void init(int k) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            if constexpr (DEG == 1) {
                const bool one = (k >> d) & 1;
                c[d][0] = one ? 1.0 : -1.0;
                c[d][1] = one ? 0.0 : 1.0;
            } else {
                lag2_coeffs(q2_axis(DIM, k, d), c[d]);
            }
        }
    }
    __host__ __device__ double factor(int d, double t) const {
        if constexpr (DEG == 1) return fma(c[d][0], t, c[d][1]);
        else return fma(c[d][0], t, c[d][1]) * fma(c[d][2], t, c[d][3]);
    }
    __host__ __device__ double operator()(const double* xi) const {
        double v = factor(0, xi[0]);
#pragma unroll
        for (int d = 1; d < DIM; ++d) v *= factor(d, xi[d]);
        return v;
    }
};

template <int DIM>
PF2_HD void shape_grad(const double* xi, double (*G)[DIM]) {
#pragma unroll
    for (int k = 0; k < Q1<DIM>::NV; ++k)
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
            double g = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
                const int bit = (k >> d) & 1;
                if (d == j) g *= bit ? 1.0 : -1.0;
                else g *= bit ? xi[d] : (1.0 - xi[d]);
            }
            G[k][j] = g;
        }
}

// x = sum_k P_k * N_k(xi), accumulated from zero in k order
template <int DIM>
PF2_HD void forward(const double (*P)[DIM], const double* xi, double* x) {
    double N[Q1<DIM>::NV];
    shape<DIM>(xi, N);
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = 0.0;
#pragma unroll
    for (int k = 0; k < Q1<DIM>::NV; ++k)
#pragma unroll
        for (int d = 0; d < DIM; ++d) x[d] += P[k][d] * N[k];
}

template <int DIM>
PF2_HD double vnorm(const double* a) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) s += a[d] * a[d];
    return sqrt(s);
}

template <int DIM>
PF2_HD double vdot(const double* a, const double* b) {
    double s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) s += a[d] * b[d];
    return s;
}

// longest diagonal (deal.II cell->diameter() for quads and hexes)
template <int DIM>
PF2_HD double diameter(const double (*P)[DIM]) {
    double best = 0.0, t[DIM];
    constexpr int NP = DIM == 2 ? 2 : 4;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        const int a = (Q1<DIM>::NV - 1) - i, b = i;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t[d] = P[a][d] - P[b][d];
        const double n = vnorm<DIM>(t);
        if (n > best) best = n;
    }
    return best;
}

template <int DIM>
PF2_HD double det(const double (*t)[DIM]) {
    if constexpr (DIM == 2) {
        return t[0][0] * t[1][1] - t[1][0] * t[0][1];
    } else {
        return t[0][0] * t[1][1] * t[2][2] - t[0][0] * t[1][2] * t[2][1]
             - t[1][0] * t[0][1] * t[2][2] + t[1][0] * t[0][2] * t[2][1]
             + t[2][0] * t[0][1] * t[1][2] - t[2][0] * t[0][2] * t[1][1];
    }
}

template <int DIM>
PF2_HD void invert(const double (*t)[DIM], double (*r)[DIM]) {
    const double inv = 1.0 / det<DIM>(t);
    if constexpr (DIM == 2) {
        r[0][0] = t[1][1] * inv;
        r[0][1] = -t[0][1] * inv;
        r[1][0] = -t[1][0] * inv;
        r[1][1] = t[0][0] * inv;
    } else {
        r[0][0] = (t[1][1] * t[2][2] - t[1][2] * t[2][1]) * inv;
        r[0][1] = (t[0][2] * t[2][1] - t[0][1] * t[2][2]) * inv;
        r[0][2] = (t[0][1] * t[1][2] - t[0][2] * t[1][1]) * inv;
        r[1][0] = (t[1][2] * t[2][0] - t[1][0] * t[2][2]) * inv;
        r[1][1] = (t[0][0] * t[2][2] - t[0][2] * t[2][0]) * inv;
        r[1][2] = (t[0][2] * t[1][0] - t[0][0] * t[1][2]) * inv;
        r[2][0] = (t[1][0] * t[2][1] - t[1][1] * t[2][0]) * inv;
        r[2][1] = (t[0][1] * t[2][0] - t[0][0] * t[2][1]) * inv;
        r[2][2] = (t[0][0] * t[1][1] - t[0][1] * t[1][0]) * inv;
    }
}

template <int DIM>
PF2_HD void matvec(const double (*m)[DIM], const double* v, double* r) {
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double s = m[i][0] * v[0];
#pragma unroll
        for (int j = 1; j < DIM; ++j) s += m[i][j] * v[j];
        r[i] = s;
    }
}

// Exact inverse of the bilinear map by the quadratic formula for the second
// reference coordinate (the method of deal.II's 2-D MappingQ1 inverse).
// false where no root is usable.
PF2_HD bool inverse_quadratic(const double (*P)[2], const double* p, double* out) {
    const double x = p[0], y = p[1];
    const double x0 = P[0][0], x1 = P[1][0], x2 = P[2][0], x3 = P[3][0];
    const double y0 = P[0][1], y1 = P[1][1], y2 = P[2][1], y3 = P[3][1];
    const double a = (x1 - x3) * (y0 - y2) - (x0 - x2) * (y1 - y3);
    const double b = -(x0 - x1 - x2 + x3) * y + (x - 2 * x1 + x3) * y0 - (x - 2 * x0 + x2) * y1
                     - (x - x1) * y2 + (x - x0) * y3;
    const double c = (x0 - x1) * y - (x - x1) * y0 + (x - x0) * y1;
    const double disc = b * b - 4 * a * c;
    if (!(disc >= 0.0)) return false;
    const double sq = sqrt(disc);
    double eta1, eta2;
    if (b != 0.0 && fabs(b) == sq) {
        eta1 = -c / b;
        eta2 = eta1;
    } else if (fabs(a) < 1e-8 * fabs(b)) {
        eta1 = 2 * c / (-b - sq);
        eta2 = 2 * c / (-b + sq);
    } else {
        eta1 = (-b - sq) / (2 * a);
        eta2 = (-b + sq) / (2 * a);
    }
    const double eta = (fabs(eta1 - 0.5) < fabs(eta2 - 0.5)) ? eta1 : eta2;
    const double sub0 = -eta * x2 + x0 * (eta - 1);
    const double den0 = eta * x3 - x1 * (eta - 1) + sub0;
    const double max_x = dmax(dmax(fabs(x0), fabs(x1)), dmax(fabs(x2), fabs(x3)));
    if (fabs(den0) > 1e-10 * max_x) {
        out[0] = (x + sub0) / den0;
        out[1] = eta;
        return true;
    }
    const double max_y = dmax(dmax(fabs(y0), fabs(y1)), dmax(fabs(y2), fabs(y3)));
    const double sub1 = -eta * y2 + y0 * (eta - 1);
    const double den1 = eta * y3 - y1 * (eta - 1) + sub1;
    if (fabs(den1) > 1e-10 * max_y) {
        out[0] = (sub1 + y) / den1;
        out[1] = eta;
        return true;
    }
    return false;
}

// Newton iteration with step-halving line search, started from the affine
// least-squares fit of the cell (deal.II MappingQGeneric's inverse).
template <int DIM>
PF2_HD bool inverse_newton(const double (*P)[DIM], const double* p, double* out) {
    constexpr int NV = Q1<DIM>::NV;
    // least-squares affine fit: A[:, j] = sum_k P_k (2 b_kj - 1) / 2^(DIM-1),
    // b = sum_k w_k P_k with w_k = (1 - sum_j (2 b_kj - 1)) / 2^DIM
    double A[DIM][DIM], bv[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        bv[i] = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) A[i][j] = 0.0;
    }
    constexpr double ka = DIM == 2 ? 0.5 : 0.25;
    constexpr double inv_nv = 1.0 / NV;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        int s = 0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += ((k >> j) & 1) ? 1 : -1;
        const double kb = (1 - s) * inv_nv;
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            bv[i] += kb * P[k][i];
#pragma unroll
            for (int j = 0; j < DIM; ++j) A[i][j] += P[k][i] * (((k >> j) & 1) ? ka : -ka);
        }
    }
    double Ai[DIM][DIM], pb[DIM], xi[DIM];
    invert<DIM>(A, Ai);
#pragma unroll
    for (int d = 0; d < DIM; ++d) pb[d] = p[d] - bv[d];
    matvec<DIM>(Ai, pb, xi);

    double pr[DIM], f[DIM];
    forward<DIM>(P, xi, pr);
#pragma unroll
    for (int d = 0; d < DIM; ++d) f[d] = pr[d] - p[d];
    if (vnorm<DIM>(f) < 1e-10 * diameter<DIM>(P)) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) out[d] = xi[d];
        return true;
    }
    int it = 0;
    double wn;
    do {
        double G[NV][DIM], df[DIM][DIM];
        shape_grad<DIM>(xi, G);
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) df[i][j] = 0.0;
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) df[i][j] += P[k][i] * G[k][j];
        if (!(det<DIM>(df) > 0)) return false;
        double dfi[DIM][DIM], delta[DIM];
        invert<DIM>(df, dfi);
        matvec<DIM>(dfi, f, delta);
        double step = 1.0;
        const double fn = vnorm<DIM>(f);
        for (;;) {
            double xt[DIM], prt[DIM], ft[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) xt[d] = xi[d] - step * delta[d];
            forward<DIM>(P, xt, prt);
#pragma unroll
            for (int d = 0; d < DIM; ++d) ft[d] = prt[d] - p[d];
            if (vnorm<DIM>(ft) < fn) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    xi[d] = xt[d];
                    f[d] = ft[d];
                }
                break;
            } else if (step > 0.05) {
                step /= 2;
            } else {
                return false;
            }
        }
        ++it;
        if (it > 20) return false;
        double wf[DIM];
        matvec<DIM>(dfi, f, wf);
        wn = vnorm<DIM>(wf);
    } while (wn > 1e-11);
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = xi[d];
    return true;
}

// Tolerance of the clearly-outside test of a cell: 1e-8 * diameter, or -1
// when the test does not apply (a non-convex quadrilateral).
template <int DIM>
PF2_HD double outside_tol(const double (*P)[DIM]) {
    const double tol = 1e-8 * diameter<DIM>(P);
    if constexpr (DIM == 2) {
        constexpr int cyc[4] = {0, 1, 3, 2};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = cyc[i], b = cyc[(i + 1) & 3], c = cyc[(i + 2) & 3];
            const double e0x = P[b][0] - P[a][0], e0y = P[b][1] - P[a][1];
            const double e1x = P[c][0] - P[b][0], e1y = P[c][1] - P[b][1];
            if (!(e0x * e1y - e0y * e1x > 0.0)) return -1.0;
        }
    }
    return tol;
}

// A point farther than tol outside the cell: outside a convex quadrilateral's
// edges (2-D) or outside the vertex bounding box (3-D).  Neither inverse can
// return a unit-cell point for it (its residual is at least the distance).
template <int DIM>
PF2_HD bool clearly_outside(const double (*P)[DIM], const double* p, double tol) {
    if (!(tol >= 0.0)) return false;
    if constexpr (DIM == 2) {
        constexpr int cyc[4] = {0, 1, 3, 2};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = cyc[i], b = cyc[(i + 1) & 3];
            const double ex = P[b][0] - P[a][0], ey = P[b][1] - P[a][1];
            const double wx = p[0] - P[a][0], wy = p[1] - P[a][1];
            const double cr = ex * wy - ey * wx;
            const double el2 = ex * ex + ey * ey;
            if (cr < 0.0 && cr * cr > tol * tol * el2) return true;
        }
        return false;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double lo = P[0][d], hi = P[0][d];
#pragma unroll
            for (int k = 1; k < 8; ++k) {
                lo = dmin(lo, P[k][d]);
                hi = dmax(hi, P[k][d]);
            }
            if (p[d] < lo - tol || p[d] > hi + tol) return true;
        }
        return false;
    }
}

// MappingQ1::transform_real_to_unit_cell: false = transformation failed.
// filter_first: clearly-outside test before the closed form (star
// candidates); otherwise the closed form runs first and the test only guards
// the Newton fallback (own-cell test).  Both orders give the same answer.
template <int DIM>
PF2_HD bool inverse_map(const double (*P)[DIM], const double* p, double* out, double tol, bool filter_first) {
    if ((filter_first || DIM == 3) && clearly_outside<DIM>(P, p, tol)) return false;
    if constexpr (DIM == 2) {
        double r[2];
        if (inverse_quadratic(P, p, r)) {
            const double eps = 1e-15;
            if (-eps <= r[1] && r[1] <= 1 + eps && -eps <= r[0] && r[0] <= 1 + eps) {
                out[0] = r[0];
                out[1] = r[1];
                return true;
            }
        }
        if (!filter_first && clearly_outside<DIM>(P, p, tol)) return false;
    }
    return inverse_newton<DIM>(P, p, out);
}

template <int DIM>
PF2_HD bool inside_unit(const double* xi) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
        if (!(xi[d] >= 0.0 && xi[d] <= 1.0)) return false;
    return true;
}

}  // namespace pf2
