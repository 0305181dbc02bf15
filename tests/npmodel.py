"""Independent two-lattice model of the same physics, in numpy.

It shares no code or layout with the oracle or the product: populations are
stored naturally (P[a][x] = post-collision population at x moving along c_a),
streaming is np.roll along each axis with halfway bounce-back taken from the
own node's opposite population when the upstream node is solid, and the
collision is the textbook He-Luo incompressible BGK on shifted populations
(f~ = f - w rho_A) with the velocity-shift body force u_eq = j/rho_A + tau g,
plus the Inamuro linear-equilibrium solute g_eq = w c (1 + 3 c.u_eq).

The in-place AA storage of the oracle maps onto this model as follows:
after an even number of steps slot a at x holds P_a(x); after an odd number
slot opp(a) at x holds the streamed, not yet collided F_a(x).
"""
import numpy as np

C19 = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0], [-1, -1, 0], [-1, 1, 0],
    [1, -1, 0], [0, 0, 1], [0, 0, -1], [1, 0, 1], [-1, 0, -1], [-1, 0, 1], [1, 0, -1], [0, 1, 1],
    [0, -1, -1], [0, -1, 1], [0, 1, -1]])


def lattice(dims):
    """velocities and weights from the isotropy conditions"""
    if dims == 2:
        c = C19[:9, :2]
        sq = (c ** 2).sum(1)
        w = np.where(sq == 0, 4 / 9, np.where(sq == 1, 1 / 9, 1 / 36))
    else:
        c = C19
        sq = (c ** 2).sum(1)
        w = np.where(sq == 0, 1 / 3, np.where(sq == 1, 1 / 18, 1 / 36))
    opp = np.array([int(np.where((c == -c[a]).all(1))[0][0]) for a in range(len(c))])
    return c, w, opp


class NpLBM:
    def __init__(self, dims, shape, mask, tau, rho_A=(1.0, 1.0), g=(0, 0, 0), nsubs=1):
        """shape = (nj, ni) or (nk, nj, ni); g = (gx, gy, gz) of the flow."""
        self.dims = dims
        self.c, self.w, self.opp = lattice(dims)
        self.Q = len(self.w)
        self.shape = tuple(shape)
        self.solid = np.asarray(mask).reshape(shape) == 1
        self.tau = tau
        self.rho_A = rho_A
        self.g = np.asarray(g[:dims], dtype=float)
        self.nsubs = nsubs
        # numpy axis of each velocity component: x is the last axis
        self.axis = [dims - 1 - d for d in range(dims)]

    def stream(self, P):
        F = np.empty_like(P)
        for a in range(self.Q):
            shifted = P[a]
            src_solid = self.solid
            for d in range(self.dims):
                if self.c[a, d]:
                    shifted = np.roll(shifted, self.c[a, d], axis=self.axis[d])
                    src_solid = np.roll(src_solid, self.c[a, d], axis=self.axis[d])
            F[a] = np.where(src_solid, P[self.opp[a]], shifted)
        return F

    def moments(self, F):
        drho = F.sum(0)
        j = np.stack([np.tensordot(self.c[:, d], F, axes=(0, 0)) for d in range(self.dims)])
        u = j / self.rho_A[0] + (self.tau[0] * self.g).reshape((-1,) + (1,) * self.dims)
        return drho, u

    def collide_flow(self, F):
        drho, u = self.moments(F)
        cu = np.tensordot(self.c.astype(float), u, axes=(1, 0))
        usq = (u ** 2).sum(0)
        feq = self.w.reshape((-1,) + (1,) * self.dims) * (
            drho + self.rho_A[0] * (3 * cu + 4.5 * cu ** 2 - 1.5 * usq))
        P = F - (F - feq) / self.tau[0]
        return np.where(self.solid, F, P), drho, u

    def collide_solute(self, G, u):
        conc = G.sum(0)
        cu = np.tensordot(self.c.astype(float), u, axes=(1, 0))
        geq = self.w.reshape((-1,) + (1,) * self.dims) * conc * (1 + 3 * cu)
        P = G - (G - geq) / self.tau[1]
        return np.where(self.solid, G, P), conc
