// lbm_tlm.cpp -- FP64 tangent-linear model of the D3Q19 flow + solute
// objective.  TEST INFRASTRUCTURE (the reference the GPU discrete adjoint is compared with);
// never shipped or measured as the product.
//
// The model is written once, generic in its number type, on the natural
// two-lattice form of the same discrete dynamics (post-collision
// populations P_a(x); each step streams them with halfway bounce-back,
// overwrites the unknown populations of the z-face nodes, records the
// objective and collides).  Instantiated with a dual number (value,
// derivative) it propagates one directional derivative exactly (forward
// mode), so dJ.dp for any direction dp can be compared with the adjoint
// gradient.  Instantiated with double it gives J itself.
//
// Physics as in lbm_oracle.c (He-Luo incompressible BGK on shifted
// populations with u_eq = j/rho_A + tau g, Inamuro solute advected by u_eq,
// z-face Zou-He pressure, solute Dirichlet inlet map and zero-gradient
// outlet), periodic in x and y, started from rest with zero solute.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

struct Dual {
  double v, d;
  Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
inline Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
inline Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
inline double val(Dual a) { return a.v; }
inline double der(Dual a) { return a.d; }
inline double val(double a) { return a; }
inline double der(double) { return 0.0; }

const int CX[19] = {0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0};
const int CY[19] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1};
const int CZ[19] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
const int UP[5] = {9, 11, 13, 15, 17};
const int DN[5] = {10, 12, 14, 16, 18};

inline int opp(int a) { return a == 0 ? 0 : ((a & 1) ? a + 1 : a - 1); }
inline double wt(int a) { return a == 0 ? 1. / 3. : ((a < 5 || a == 9 || a == 10) ? 1. / 18. : 1. / 36.); }

struct Setup {
  int ni, nj, nk, nsteps, window, zf;
  int kind;  // 0: solute flux through plane zf; 1: permeability
  const uint8_t *mask;
  double rho_bot, rho_top;
  // observation points: J += w (c(node, step) - target)^2
  int nobs;
  const long *onode;
  const int *ostep;
  const double *otarget, *oweight;
};

template <typename T>
T run(const Setup &S, T tau, T taus, const T *g, const std::vector<T> &cmap) {
  const int ni = S.ni, nj = S.nj, nk = S.nk;
  const long E = (long)ni * nj, N = E * nk;
  std::vector<T> P(19 * N), G(19 * N), F(19 * N), H(19 * N);
  auto solid = [&](long n) { return S.mask[n] == 1; };
  auto idx = [&](int i, int j, int k) {
    i = (i + ni) % ni;
    j = (j + nj) % nj;
    k = (k + nk) % nk;
    return (long)i + (long)ni * j + E * k;
  };
  const T one(1.0), om = one / tau, oms = one / taus;
  T flux(0.0), obs(0.0), jz(0.0);
  long nfluid = 0;
  for (long n = 0; n < N; n++) nfluid += !solid(n);
  for (int t = 1; t <= S.nsteps; t++) {
    // stream with halfway bounce-back
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
          long n = idx(i, j, k);
          if (solid(n)) continue;
          for (int a = 0; a < 19; a++) {
            long m = idx(i - CX[a], j - CY[a], k - CZ[a]);
            F[a * N + n] = solid(m) ? P[opp(a) * N + n] : P[a * N + m];
            H[a * N + n] = solid(m) ? G[opp(a) * N + n] : G[a * N + m];
          }
        }
    // z faces: the unknown populations of the face nodes are overwritten
    for (int top = 0; top < 2; top++) {
      int k = top ? nk - 1 : 0;
      for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
          long n = idx(i, j, k);
          if (solid(n)) continue;
          T f[19], h[19];
          for (int a = 0; a < 19; a++) {
            f[a] = F[a * N + n];
            h[a] = H[a * N + n];
          }
          T P0 = f[0];
          for (int a = 1; a < 9; a++) P0 += f[a];
          T M(0.0);
          for (int q = 0; q < 5; q++) M += f[top ? UP[q] : DN[q]];
          T Nx = ((f[1] + f[5] + f[8]) - (f[2] + f[7] + f[6])) / T(2.0);
          T Ny = ((f[3] + f[5] + f[7]) - (f[4] + f[6] + f[8])) / T(2.0);
          if (!top) {
            T J = T(S.rho_bot - 1.0) - (P0 + T(2.0) * M);
            F[9 * N + n] = f[10] + J / T(3.0);
            F[11 * N + n] = f[12] + J / T(6.0) - Nx;
            F[13 * N + n] = f[14] + J / T(6.0) + Nx;
            F[15 * N + n] = f[16] + J / T(6.0) - Ny;
            F[17 * N + n] = f[18] + J / T(6.0) + Ny;
            T known(0.0);
            for (int a = 0; a < 9; a++) known += h[a];
            for (int q = 0; q < 5; q++) known += h[DN[q]];
            T cs = (cmap[i + (long)ni * j] - known) / T(1. / 6.);
            for (int q = 0; q < 5; q++) H[UP[q] * N + n] = T(wt(UP[q])) * cs;
          } else {
            T J = (P0 + T(2.0) * M) - T(S.rho_top - 1.0);
            F[10 * N + n] = f[9] - J / T(3.0);
            F[12 * N + n] = f[11] - J / T(6.0) + Nx;
            F[14 * N + n] = f[13] - J / T(6.0) - Nx;
            F[16 * N + n] = f[15] - J / T(6.0) + Ny;
            F[18 * N + n] = f[17] - J / T(6.0) - Ny;
            T known(0.0);
            for (int q = 0; q < 5; q++) known += h[UP[q]];
            T cs = known / T(1. / 6.);
            for (int q = 0; q < 5; q++) H[DN[q] * N + n] = T(wt(DN[q])) * cs;
          }
        }
    }
    // observations: the pre-collision concentration at (node, step)
    for (int q = 0; q < S.nobs; q++) {
      if (S.ostep[q] != t) continue;
      const long n = S.onode[q];
      T c = H[n];
      for (int a = 1; a < 19; a++) c += H[a * N + n];
      const T r = c - T(S.otarget[q]);
      obs += T(S.oweight ? S.oweight[q] : 1.0) * r * r;
    }
    // objective: solute flux through plane zf, or the flow's z momentum
    // over all nodes, over the last `window` steps
    if (S.window > 0 && t > S.nsteps - S.window && S.kind == 1) {
      for (long n = 0; n < N; n++) {
        if (solid(n)) continue;
        for (int a = 0; a < 19; a++)
          if (CZ[a]) jz += T((double)CZ[a]) * F[a * N + n];
      }
    }
    if (S.window > 0 && t > S.nsteps - S.window && S.kind == 0) {
      for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
          long n = idx(i, j, S.zf);
          if (solid(n)) continue;
          for (int a = 0; a < 19; a++)
            if (CZ[a]) flux += T((double)CZ[a]) * H[a * N + n];
        }
    }
    // collide
    for (long n = 0; n < N; n++) {
      if (solid(n)) continue;
      T drho(0.0), u[3] = {T(0.0), T(0.0), T(0.0)}, c(0.0);
      for (int a = 0; a < 19; a++) {
        T fa = F[a * N + n];
        drho += fa;
        if (CX[a]) u[0] += T((double)CX[a]) * fa;
        if (CY[a]) u[1] += T((double)CY[a]) * fa;
        if (CZ[a]) u[2] += T((double)CZ[a]) * fa;
        c += H[a * N + n];
      }
      for (int d = 0; d < 3; d++) u[d] = u[d] + tau * g[d];
      T usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      for (int a = 0; a < 19; a++) {
        T cu = T((double)CX[a]) * u[0] + T((double)CY[a]) * u[1] + T((double)CZ[a]) * u[2];
        T w(wt(a));
        T feq = w * (drho + (T(3.0) * cu + T(4.5) * cu * cu - T(1.5) * usq));
        T geq = w * c * (one + T(3.0) * cu);
        P[a * N + n] = F[a * N + n] + om * (feq - F[a * N + n]);
        G[a * N + n] = H[a * N + n] + oms * (geq - H[a * N + n]);
      }
    }
  }
  if (S.window > 0 && S.kind == 1) {
    // permeability k = rho_A nu <u_z,phys> / G, rho_A = 1: nu = (tau - 1/2)/3,
    // u_phys = j/rho_A + g/2 (superficial average over all N nodes), G the
    // driving gradient (rho_bot - rho_top)/(3 (nk - 1)) + rho_A g_z
    const T nu = (tau - T(0.5)) / T(3.0);
    const T G = T((S.rho_bot - S.rho_top) / (3.0 * (nk - 1))) + g[2];
    const T W((double)S.window);
    const T uphys = (jz + W * T((double)nfluid) * g[2] / T(2.0)) / (W * T((double)N));
    return nu * uphys / G + obs;
  }
  if (S.window > 0) return flux / T((double)S.window * (double)E) + obs;
  return obs;
}

}  // namespace

extern "C" {

// J and the directional derivative dJ = grad J . (dg, dtau, dtaus, dcmap)
// for rho_A = 1, started from rest with zero solute.  params[6] =
// {gx, gy, gz, tau, tau_s, unused}; dparams[5] = {dgx, dgy, dgz, dtau,
// dtau_s}; cmap, dcmap: [nj][ni].  J is the flux objective over the last
// `window` steps (none for window 0) plus the misfit of the nobs
// observation points (dense node, step 1..nsteps, target, weight or NULL).
// kind 0: the solute flux through plane zf; kind 1: the permeability
// rho_A nu <u_z,phys> / G averaged over the window (zf unused).
int orc_tlm_objective_kind(int kind, int ni, int nj, int nk, const uint8_t *mask, double rho_bot, double rho_top,
                           int zf, int nsteps, int window, const double *params, const double *cmap, int nobs,
                           const long *onode, const int *ostep, const double *otarget, const double *oweight,
                           const double *dparams, const double *dcmap, double *J, double *dJ) {
  if (ni <= 0 || nj <= 0 || nk <= 2 || !mask || !params || !cmap || window < 0 || window > nsteps) return 1;
  if (nobs < 0 || (window == 0 && nobs == 0) || (nobs && (!onode || !ostep || !otarget))) return 1;
  if (kind != 0 && kind != 1) return 1;
  Setup S{ni, nj, nk, nsteps, window, zf, kind, mask, rho_bot, rho_top, nobs, onode, ostep, otarget, oweight};
  const long E = (long)ni * nj;
  if (!dparams) {
    std::vector<double> cm(cmap, cmap + E);
    double g[3] = {params[0], params[1], params[2]};
    *J = run<double>(S, params[3], params[4], g, cm);
    if (dJ) *dJ = 0.0;
    return 0;
  }
  std::vector<Dual> cm(E);
  for (long m = 0; m < E; m++) cm[m] = Dual(cmap[m], dcmap ? dcmap[m] : 0.0);
  Dual g[3] = {Dual(params[0], dparams[0]), Dual(params[1], dparams[1]), Dual(params[2], dparams[2])};
  Dual r = run<Dual>(S, Dual(params[3], dparams[3]), Dual(params[4], dparams[4]), g, cm);
  *J = val(r);
  if (dJ) *dJ = der(r);
  return 0;
}

int orc_tlm_objective_obs(int ni, int nj, int nk, const uint8_t *mask, double rho_bot, double rho_top, int zf,
                          int nsteps, int window, const double *params, const double *cmap, int nobs,
                          const long *onode, const int *ostep, const double *otarget, const double *oweight,
                          const double *dparams, const double *dcmap, double *J, double *dJ) {
  return orc_tlm_objective_kind(0, ni, nj, nk, mask, rho_bot, rho_top, zf, nsteps, window, params, cmap, nobs,
                                onode, ostep, otarget, oweight, dparams, dcmap, J, dJ);
}

// the flux objective alone
int orc_tlm_objective(int ni, int nj, int nk, const uint8_t *mask, double rho_bot, double rho_top, int zf,
                      int nsteps, int window, const double *params, const double *cmap, const double *dparams,
                      const double *dcmap, double *J, double *dJ) {
  if (window < 1) return 1;
  return orc_tlm_objective_obs(ni, nj, nk, mask, rho_bot, rho_top, zf, nsteps, window, params, cmap, 0, nullptr,
                               nullptr, nullptr, nullptr, dparams, dcmap, J, dJ);
}
}
