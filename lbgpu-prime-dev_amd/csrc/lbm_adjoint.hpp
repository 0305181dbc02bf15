// Discrete adjoint of the D3Q19 flow + solute time step (FP64, periodic
// lattice, z-face boundary conditions).  No reference counterpart.
//
// A time step is  state_t = K_t(B_t(state_{t-1}))  with B_t the z-face
// conditions and K_t the half-step kernel.  K_t gathers every stored slot
// exactly once, collides per node and scatters every slot exactly once, so
// its transpose reads the adjoint at the places K_t wrote, applies the
// transposed collision Jacobian at the forward state, and writes at the
// places K_t read - in place and race-free for the same reason as the
// forward half-step.  B_t overwrote the unknown populations of the face
// nodes with linear combinations of the known ones; its transpose adds the
// unknowns' adjoint to the known ones and clears the unknowns.
//
// Objective: J = 1/(W A) sum over the last W steps of the solute flux
// sum_a c_a,z g_a through plane zf (pre-collision, after the BCs), A =
// LX*LY.  Parameters: body force g, tau, tau_s and the inlet concentration
// map c_b(x, y).
#pragma once

#include "lbm_aux.hpp"

namespace lbm {

struct AdjArgs {
  StepArgs<double, double> A;  // A.F: forward state before the step, after its BCs
  double *L;                   // adjoint state, same tile layout as the populations
  double seed;                 // dJ / d(plane flux of this step); 0 outside the window
  int zf;                      // flux plane
  double *gacc;                // [launched blocks][5]: dJ/dg_x, g_y, g_z, tau, tau_s
};

// slot offset a half-step reads for direction a (the pull) / writes (the push)
template <bool ODD>
__device__ __forceinline__ uint32_t pull_off(const int (&nb)[19], int s, int a, int idx) {
  if (a == 0) return foff<19, 2>(s, 0, idx);
  if (ODD) return nb[a] >= 0 ? foff<19, 2>(s, a, nb[a]) : foff<19, 2>(s, OPP(a), idx);
  return foff<19, 2>(s, OPP(a), idx);
}
template <bool ODD>
__device__ __forceinline__ uint32_t push_off(const int (&nb)[19], int s, int a, int idx) {
  if (a == 0) return foff<19, 2>(s, 0, idx);
  if (ODD) return nb[OPP(a)] >= 0 ? foff<19, 2>(s, OPP(a), nb[OPP(a)]) : foff<19, 2>(s, a, idx);
  return foff<19, 2>(s, a, idx);
}

// Waves per SIMD the register allocation aims at (the even step fits three,
// the odd step with its neighbour ranks two).
#ifndef LBM_ADJ_WAVES_ODD
#define LBM_ADJ_WAVES_ODD 2
#endif
#ifndef LBM_ADJ_WAVES_EVEN
#define LBM_ADJ_WAVES_EVEN 3
#endif

template <bool ODD, bool OFF32>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ODD ? LBM_ADJ_WAVES_ODD : LBM_ADJ_WAVES_EVEN)))
void k_adjoint_step(AdjArgs J) {
  constexpr int Q = 19;
  const int lb = xcd_block(J.A.nblk8);
  const int idx = lb * kBlock + (int)threadIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (idx < J.A.nf) {
    const StepArgs<double, double> &A = J.A;
    int nb[Q];
    nb[0] = idx;
    int i, j, k;
    coords<3>(A.G, (uint32_t)A.G.node[idx], i, j, k);
    if (ODD) {
      int r[18];
      neighbors_rows<3>(A.G, i, j, k, r);
#pragma unroll
      for (int a = 1; a < Q; a++) nb[a] = r[a - 1];
    }
    // The transposed collision needs, of the adjoint, its weighted moments
    //   Lr = sum w lam, L1 = sum w lam c, L2 = sum w lam c c^T,
    //   M0 = sum w mu,  M1 = sum w mu c
    // and, of the forward populations, their moments and the two sums
    // sum lam f, sum mu g; lam and mu themselves are needed once more for
    // the store.  Every per-direction sum of the collision Jacobian is a
    // combination of these (D3Q19 moments up to second order), so no
    // per-direction temporaries are kept.
    double lam[Q], mu[Q];
    double Lr = 0.0, L1[3] = {0.0, 0.0, 0.0}, L2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // xx yy zz xy xz yz
    double M0 = 0.0, M1[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < Q; a++) {
      const pair_t<double> v = *(const pair_t<double> *)elem<OFF32>(J.L, push_off<ODD>(nb, 0, a, idx));
      lam[a] = v.x;
      mu[a] = v.y;
      const double wl = W<19>(a) * v.x, wm = W<19>(a) * v.y;
      Lr += wl;
      M0 += wm;
      if (CX(a) != 0) {
        L1[0] += CX(a) > 0 ? wl : -wl;
        M1[0] += CX(a) > 0 ? wm : -wm;
        L2[0] += wl;
      }
      if (CY(a) != 0) {
        L1[1] += CY(a) > 0 ? wl : -wl;
        M1[1] += CY(a) > 0 ? wm : -wm;
        L2[1] += wl;
      }
      if (CZ(a) != 0) {
        L1[2] += CZ(a) > 0 ? wl : -wl;
        M1[2] += CZ(a) > 0 ? wm : -wm;
        L2[2] += wl;
      }
      if (CX(a) * CY(a) != 0) L2[3] += CX(a) * CY(a) > 0 ? wl : -wl;
      if (CX(a) * CZ(a) != 0) L2[4] += CX(a) * CZ(a) > 0 ? wl : -wl;
      if (CY(a) * CZ(a) != 0) L2[5] += CY(a) * CZ(a) > 0 ? wl : -wl;
    }
    const double rhoA = A.rhoA, tau = A.tau, om = A.inv_tau, oms = A.inv_tau_s;
    double drho = 0.0, u[3] = {0.0, 0.0, 0.0}, c = 0.0, lf = 0.0, mg = 0.0;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      const pair_t<double> v = *(const pair_t<double> *)elem<OFF32>(A.F, pull_off<ODD>(nb, 0, a, idx));
      drho += v.x;
      if (CX(a) != 0) u[0] += CX(a) > 0 ? v.x : -v.x;
      if (CY(a) != 0) u[1] += CY(a) > 0 ? v.x : -v.x;
      if (CZ(a) != 0) u[2] += CZ(a) > 0 ? v.x : -v.x;
      c += v.y;
      lf += lam[a] * v.x;
      mg += mu[a] * v.y;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) u[d] = u[d] / rhoA + tau * A.g[d];
    const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    // solute: g* = (1 - ws) g + ws w c (1 + 3 c.u)
    //   Sg = sum mu w (1 + 3 c.u), dts = sum mu (geq - g)
    const double Sg = M0 + 3.0 * (u[0] * M1[0] + u[1] * M1[1] + u[2] * M1[2]);
    const double dts = c * Sg - mg;
    // flow: f* = (1 - w) f + w feq(drho, u)
    //   Lu = sum lam w rhoA (3 c + 9 (c.u) c - 3 u), dt = sum lam (feq - f)
    double L2u[3];
    L2u[0] = L2[0] * u[0] + L2[3] * u[1] + L2[4] * u[2];
    L2u[1] = L2[3] * u[0] + L2[1] * u[1] + L2[5] * u[2];
    L2u[2] = L2[4] * u[0] + L2[5] * u[1] + L2[2] * u[2];
    double Lu[3];
#pragma unroll
    for (int d = 0; d < 3; d++) Lu[d] = rhoA * (3.0 * L1[d] + 9.0 * L2u[d] - 3.0 * u[d] * Lr);
    const double uL1 = u[0] * L1[0] + u[1] * L1[1] + u[2] * L1[2];
    const double uL2u = u[0] * L2u[0] + u[1] * L2u[1] + u[2] * L2u[2];
    const double dt = drho * Lr + rhoA * (3.0 * uL1 + 4.5 * uL2u - 1.5 * usq * Lr) - lf;
    double G[3];
#pragma unroll
    for (int d = 0; d < 3; d++) G[d] = om * Lu[d] + 3.0 * oms * c * M1[d];
    const double seed = k == J.zf ? J.seed : 0.0;
    double Gr[3];
#pragma unroll
    for (int d = 0; d < 3; d++) Gr[d] = G[d] / rhoA;
#pragma unroll
    for (int b = 0; b < Q; b++) {
      double lb_ = (1.0 - om) * lam[b] + om * Lr;
      if (CX(b) != 0) lb_ += CX(b) > 0 ? Gr[0] : -Gr[0];
      if (CY(b) != 0) lb_ += CY(b) > 0 ? Gr[1] : -Gr[1];
      if (CZ(b) != 0) lb_ += CZ(b) > 0 ? Gr[2] : -Gr[2];
      double mb = (1.0 - oms) * mu[b] + oms * Sg;
      if (CZ(b) != 0) mb += CZ(b) > 0 ? seed : -seed;
      pair_t<double> v;
      v.x = lb_;
      v.y = mb;
      *(pair_t<double> *)elem<OFF32>(J.L, pull_off<ODD>(nb, 0, b, idx)) = v;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) acc[d] = tau * G[d];
    acc[3] = -om * om * dt + A.g[0] * G[0] + A.g[1] * G[1] + A.g[2] * G[2];
    acc[4] = -oms * oms * dts;
  }
  __shared__ double red[5];
  const int kinds[5] = {0, 0, 0, 0, 0};
  block_reduce<5>(acc, kinds, red);
  if (threadIdx.x < 5) J.gacc[(long)lb * 5 + threadIdx.x] += red[threadIdx.x];
}

// Transpose of the z-face conditions of k_bc_zface for one face.
struct AdjBcArgs {
  StepArgs<double, double> A;
  double *L;
  const int *list;
  int n;
  int phase;
};

template <bool ODD>
__device__ __forceinline__ void adj_face(double *L, const int (&nb)[19], int s, int idx, bool top, bool flow,
                                        double &dcb) {
  constexpr int UP[5] = {9, 11, 13, 15, 17};
  constexpr int DN[5] = {10, 12, 14, 16, 18};
  uint32_t loc[19];
#pragma unroll
  for (int a = 0; a < 19; a++) loc[a] = pull_off<ODD>(nb, s, a, idx);
  double lu[5], dl[19];
#pragma unroll
  for (int a = 0; a < 19; a++) dl[a] = 0.0;
#pragma unroll
  for (int q = 0; q < 5; q++) lu[q] = L[loc[top ? DN[q] : UP[q]]];
  if (flow) {
    double lJ, lNx, lNy;
    if (!top) {
      // J = drb - (P + 2M); v9 = f10 + J/3, v11 = f12 + J/6 - Nx, v13 = f14 + J/6 + Nx,
      // v15 = f16 + J/6 - Ny, v17 = f18 + J/6 + Ny
      lJ = lu[0] / 3.0 + (lu[1] + lu[2] + lu[3] + lu[4]) / 6.0;
      lNx = -lu[1] + lu[2];
      lNy = -lu[3] + lu[4];
#pragma unroll
      for (int q = 0; q < 5; q++) dl[DN[q]] += lu[q] - 2.0 * lJ;
#pragma unroll
      for (int a = 0; a < 9; a++) dl[a] -= lJ;
    } else {
      // J = (P + 2M) - drb; v10 = f9 - J/3, v12 = f11 - J/6 + Nx, v14 = f13 - J/6 - Nx,
      // v16 = f15 - J/6 + Ny, v18 = f17 - J/6 - Ny
      lJ = -lu[0] / 3.0 - (lu[1] + lu[2] + lu[3] + lu[4]) / 6.0;
      lNx = lu[1] - lu[2];
      lNy = lu[3] - lu[4];
#pragma unroll
      for (int q = 0; q < 5; q++) dl[UP[q]] += lu[q] + 2.0 * lJ;
#pragma unroll
      for (int a = 0; a < 9; a++) dl[a] += lJ;
    }
    // Nx = ((f1 + f5 + f8) - (f2 + f7 + f6)) / 2, Ny = ((f3 + f5 + f7) - (f4 + f6 + f8)) / 2
    dl[1] += lNx / 2.0;
    dl[5] += lNx / 2.0;
    dl[8] += lNx / 2.0;
    dl[2] -= lNx / 2.0;
    dl[7] -= lNx / 2.0;
    dl[6] -= lNx / 2.0;
    dl[3] += lNy / 2.0;
    dl[5] += lNy / 2.0;
    dl[7] += lNy / 2.0;
    dl[4] -= lNy / 2.0;
    dl[6] -= lNy / 2.0;
    dl[8] -= lNy / 2.0;
  } else {
    // solute: unknown v_u = w_u c*; bottom c* = (c_b - known) / (1/6), top c* = sum(+z) / (1/6)
    double lcs = 0.0;
#pragma unroll
    for (int q = 0; q < 5; q++) lcs += W<19>(top ? DN[q] : UP[q]) * lu[q];
    const double k6 = lcs / (1. / 6.);
    if (!top) {
#pragma unroll
      for (int a = 0; a < 9; a++) dl[a] -= k6;
#pragma unroll
      for (int q = 0; q < 5; q++) dl[DN[q]] -= k6;
      dcb = k6;
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) dl[UP[q]] += k6;
    }
  }
#pragma unroll
  for (int a = 0; a < 19; a++) {
    bool unknown = false;
#pragma unroll
    for (int q = 0; q < 5; q++) unknown |= a == (top ? DN[q] : UP[q]);
    if (unknown)
      L[loc[a]] = 0.0;
    else if (dl[a] != 0.0)
      L[loc[a]] += dl[a];
  }
}

__global__ void k_adjoint_zface(AdjBcArgs B, ZfaceArgs Z, double *cgrad) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.n) return;
  const int idx = B.list[t];
  int nb[19];
  nb[0] = idx;
  double dcb = 0.0;
  if (B.phase == 1) {
    int i, j, k;
    coords<3>(B.A.G, (uint32_t)B.A.G.node[idx], i, j, k);
    int r[18];
    neighbors_rows<3>(B.A.G, i, j, k, r);
#pragma unroll
    for (int a = 1; a < 19; a++) nb[a] = r[a - 1];
    if (Z.flow) adj_face<true>(B.L, nb, 0, idx, Z.top, true, dcb);
    if (Z.solute) adj_face<true>(B.L, nb, 1, idx, Z.top, false, dcb);
  } else {
    if (Z.flow) adj_face<false>(B.L, nb, 0, idx, Z.top, true, dcb);
    if (Z.solute) adj_face<false>(B.L, nb, 1, idx, Z.top, false, dcb);
  }
  if (!Z.top && Z.solute && cgrad) cgrad[B.A.G.node[idx]] += dcb;
}

}  // namespace lbm
