"""Host FEM stand-in for the 2-D channel driver: one fractional step on Q1.

The reference's case drivers advance the fluid with deal.II between two
particle stages (cylinder2D/cylinder2D.cc:385-731: implicit viscous
prediction of Vx and Vy, pressure Poisson equation, velocity correction,
then old_solution = solution for the next correction).  deal.II is not
available here, so this module restates that step on the same Q1 mesh with
scipy, in a reduced form:

  prediction  (rho M + mu dt K) V* = rho M V_proj      per component,
              Dirichlet walls (0) and inlet (Poiseuille profile)
  pressure    K p = -(rho/dt) (D V*),  (D V*)_i = int N_i div V*,
              dp/dn = 0 on walls and inlet, p = 0 at the outlet
  correction  M V = M V* - (dt/rho) G p,  (G p)_i = int N_i grad p,
              Dirichlet values re-imposed

(the reference's prediction also carries the off-diagonal viscous terms of
the full stress tensor; they vanish for the channel's unidirectional flow).
M, K, D and G are assembled with 2x2 Gauss quadrature on the bilinear cells;
the three systems are factorised once (scipy.sparse.linalg.splu).  The host
step is deterministic, so two runs fed the same projected field give the
same bits.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_G = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])


def _q1_at(xi, eta):
    """Q1 shape values and reference gradients at (xi, eta), deal.II vertex order."""
    N = np.array([(1 - xi) * (1 - eta), xi * (1 - eta), (1 - xi) * eta, xi * eta])
    dN = np.array([[-(1 - eta), -(1 - xi)], [(1 - eta), -xi], [-eta, (1 - xi)], [eta, xi]])
    return N, dN


def assemble(mesh):
    """Consistent mass M, stiffness K and the two gradient operators
    Gx, Gy with (G_d)_{ij} = int N_i dN_j/dx_d (all n_dofs x n_dofs)."""
    P = mesh.vertices[mesh.cells]                      # (nc, 4, 2)
    nc = mesh.n_cells
    Me = np.zeros((nc, 4, 4))
    Ke = np.zeros((nc, 4, 4))
    Ge = np.zeros((2, nc, 4, 4))
    for xi in _G:
        for eta in _G:
            N, dN = _q1_at(xi, eta)
            J = np.einsum("cka,kb->cab", P, dN)             # dx_a / dxi_b
            det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
            Jinv = np.empty_like(J)
            Jinv[:, 0, 0] = J[:, 1, 1] / det
            Jinv[:, 0, 1] = -J[:, 0, 1] / det
            Jinv[:, 1, 0] = -J[:, 1, 0] / det
            Jinv[:, 1, 1] = J[:, 0, 0] / det
            w = det * 0.25                                  # Gauss weights 1/2 x 1/2
            grad = np.einsum("kb,cba->cka", dN, Jinv)      # dN_k/dx_a
            Me += w[:, None, None] * np.outer(N, N)[None]
            Ke += w[:, None, None] * np.einsum("cia,cja->cij", grad, grad)
            for d in range(2):
                Ge[d] += w[:, None, None] * np.einsum("i,cj->cij", N, grad[:, :, d])
    rows = np.repeat(mesh.cells, 4, axis=1).ravel()
    cols = np.tile(mesh.cells, (1, 4)).ravel()
    n = mesh.n_dofs

    def mat(E):
        return sp.csr_matrix((E.ravel(), (rows, cols)), shape=(n, n))

    return mat(Me), mat(Ke), mat(Ge[0]), mat(Ge[1])


def _dirichlet(A, nodes):
    """A with the rows and columns of `nodes` replaced by the identity."""
    A = A.tolil()
    A[nodes, :] = 0.0
    A[:, nodes] = 0.0
    for i in nodes:
        A[i, i] = 1.0
    return A.tocsc()


class ChannelFEM:
    """Fractional-step stand-in for the channel driver (walls at y = 0 and
    y = H, Poiseuille inflow with peak `umax` at x = 0, outlet at x = L)."""

    def __init__(self, mesh, dt, umax=1.0, H=1.0, mu=0.01, rho=1.0):
        self.mesh, self.dt, self.mu, self.rho = mesh, dt, mu, rho
        self.M, self.K, self.Gx, self.Gy = assemble(mesh)
        bv = mesh.boundary_vertices
        self.walls = np.unique(np.concatenate([bv["bottom"], bv["top"]]))
        self.inlet = np.setdiff1d(bv["inlet"], self.walls)
        self.outlet = np.asarray(bv["outlet"])
        y = mesh.vertices[self.inlet, 1]
        self.u_in = 4.0 * umax * y * (H - y) / H ** 2
        self.vel_nodes = np.concatenate([self.walls, self.inlet])
        self.vel_values = [np.concatenate([np.zeros(len(self.walls)), self.u_in]),
                           np.zeros(len(self.vel_nodes))]
        A = rho * self.M + mu * dt * self.K
        self._pred = spla.splu(_dirichlet(A, self.vel_nodes))
        self._pres = spla.splu(_dirichlet(self.K, self.outlet))
        self._mass = spla.splu(_dirichlet(self.M, self.vel_nodes))

    def _impose(self, rhs, d):
        rhs = rhs.copy()
        rhs[self.vel_nodes] = self.vel_values[d]
        return rhs

    def divergence(self, V):
        """(D V)_i = int N_i div V (the Q1 field's divergence tested with N_i);
        the outlet rows, where p is prescribed, are zero."""
        div = self.Gx @ V[0] + self.Gy @ V[1]
        div[self.outlet] = 0.0
        return div

    def step(self, V_proj):
        """New nodal velocity (2, n_dofs) from the projected particle field."""
        Vs = np.empty_like(V_proj)
        for d in range(2):
            Vs[d] = self._pred.solve(self._impose(self.rho * (self.M @ V_proj[d]), d))
        rhs_p = -(self.rho / self.dt) * self.divergence(Vs)
        rhs_p[self.outlet] = 0.0
        p = self._pres.solve(rhs_p)
        V = np.empty_like(V_proj)
        for d, G in enumerate((self.Gx, self.Gy)):
            V[d] = self._mass.solve(self._impose(self.M @ Vs[d] - (self.dt / self.rho) * (G @ p), d))
        self.p = p
        return V
