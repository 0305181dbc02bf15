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
// Objectives: J = 1/(W A) sum over the last W steps of the solute flux
// sum_a c_a,z g_a through plane zf (pre-collision, after the BCs), A =
// LX*LY; or the permeability rho_A nu <u_z,phys> / G over the last W steps,
// which is linear in the flow's z momentum sum_a c_a,z f_a of every node
// (plus terms in tau and g_z the host adds).  Parameters: body force g,
// tau, tau_s and the inlet concentration map c_b(x, y).
#pragma once

#include "lbm_aux.hpp"

namespace lbm {

struct AdjArgs {
  StepArgs<double, double> A;  // A.F: forward state before the step, after its BCs
                               // unless A.zb.on: then the face nodes' boundary
                               // values are formed here from their known ones
  double *L;                   // adjoint state, same tile layout as the populations
  const int *amap;             // node block of each launched off-face block, or null
  double seed;                 // dJ / d(plane flux of this step); 0 outside the window
  double kseed;                // dJ / d(z momentum of a node, this step); 0 outside the
                               // window or for the flux objective
  int zf;                      // flux plane
  double *gacc;                // [launched blocks][5]: dJ/dg_x, g_y, g_z, tau, tau_s
  double *cgrad;               // [LY][LX] dJ/dc_bot_map (face-node kernel)
  const double *obs;           // [nf] dJ/dc of this step's observations (pre-collision
                               // solute concentration per node); null on other steps
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

// Threads per workgroup of the adjoint step: one wave.  With one or two
// waves per SIMD (the register counts below) a workgroup of four waves
// holds its CU's slots until its slowest wave ends; one-wave workgroups are
// replaced wave by wave (kernels of a 128^3 sweep: 289 ms against 305 ms
// with 256 threads, 291 ms with 128; capping the workgroups per CU gained
// nothing).  The gradient partials are per node block, so the block size
// fixes the order of their sums.
constexpr int kAdjBlock = 64;

// Waves per SIMD the register allocation aims at.  The even step at two
// waves (about 236 registers): squeezed to three it fits in 150 registers
// only by reusing one register quadruple for all 19 stores, and each store
// then waits for every memory access before it (233 us against 133 us per
// launch at 128^3).  The odd step, with its neighbour ranks, runs at one
// wave (344 registers): squeezed to two (256 registers and 356 B of spills
// per lane) the adjoint kernels of a 128^3 sweep took 313 ms instead of 303.
#ifndef LBM_ADJ_WAVES_ODD
#define LBM_ADJ_WAVES_ODD 1
#endif
#ifndef LBM_ADJ_WAVES_EVEN
#define LBM_ADJ_WAVES_EVEN 2
#endif

// Weighted moments of the adjoint of one node:
//   Lr = sum w lam, L1 = sum w lam c, L2 = sum w lam c c^T (xx yy zz xy xz yz),
//   M0 = sum w mu,  M1 = sum w mu c.
// Every per-direction sum of the transposed collision is a combination of
// these (D3Q19 moments up to second order), so no per-direction
// temporaries are needed.
struct AdjMoments {
  double Lr = 0.0, L1[3] = {0.0, 0.0, 0.0}, L2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double M0 = 0.0, M1[3] = {0.0, 0.0, 0.0};
};

// the adjoint at the push positions (where the forward half-step wrote)
template <bool ODD, bool OFF32>
__device__ __forceinline__ void adj_load(const AdjArgs &J, const int (&nb)[19], int idx, double (&lam)[19],
                                         double (&mu)[19], AdjMoments &m) {
#pragma unroll
  for (int a = 0; a < 19; a++) {
    const pair_t<double> v = ld_stream<!ODD>((const pair_t<double> *)elem<OFF32>(J.L, push_off<ODD>(nb, 0, a, idx)));
    lam[a] = v.x;
    mu[a] = v.y;
    const double wl = W<19>(a) * v.x, wm = W<19>(a) * v.y;
    m.Lr += wl;
    m.M0 += wm;
    if (CX(a) != 0) {
      m.L1[0] += CX(a) > 0 ? wl : -wl;
      m.M1[0] += CX(a) > 0 ? wm : -wm;
      m.L2[0] += wl;
    }
    if (CY(a) != 0) {
      m.L1[1] += CY(a) > 0 ? wl : -wl;
      m.M1[1] += CY(a) > 0 ? wm : -wm;
      m.L2[1] += wl;
    }
    if (CZ(a) != 0) {
      m.L1[2] += CZ(a) > 0 ? wl : -wl;
      m.M1[2] += CZ(a) > 0 ? wm : -wm;
      m.L2[2] += wl;
    }
    if (CX(a) * CY(a) != 0) m.L2[3] += CX(a) * CY(a) > 0 ? wl : -wl;
    if (CX(a) * CZ(a) != 0) m.L2[4] += CX(a) * CZ(a) > 0 ? wl : -wl;
    if (CY(a) * CZ(a) != 0) m.L2[5] += CY(a) * CZ(a) > 0 ? wl : -wl;
  }
}

// what the transposed collision needs of the forward populations of one
// node: their moments and the two sums sum lam f, sum mu g
struct AdjState {
  double drho = 0.0, j[3] = {0.0, 0.0, 0.0}, c = 0.0, lf = 0.0, mg = 0.0;
};

template <int A_>
__device__ __forceinline__ void adj_state_add(AdjState &st, double f, double g, double lam, double mu) {
  st.drho += f;
  if (CX(A_) != 0) st.j[0] += CX(A_) > 0 ? f : -f;
  if (CY(A_) != 0) st.j[1] += CY(A_) > 0 ? f : -f;
  if (CZ(A_) != 0) st.j[2] += CZ(A_) > 0 ? f : -f;
  st.c += g;
  st.lf += lam * f;
  st.mg += mu * g;
}

template <int A_, int END>
__device__ __forceinline__ void adj_state_arrays(AdjState &st, const double (&f)[19], const double (&g)[19],
                                                 const double (&lam)[19], const double (&mu)[19]) {
  if constexpr (A_ < END) {
    adj_state_add<A_>(st, f[A_], g[A_], lam[A_], mu[A_]);
    adj_state_arrays<A_ + 1, END>(st, f, g, lam, mu);
  }
}

// The forward state is only read here: streaming loads in both half-steps
// leave the L2 to the adjoint's read-then-written lines (adjoint kernels of
// a 128^3 sweep: 289 -> 286 ms with them in the odd step too).
template <int A_, int END, bool ODD, bool OFF32>
__device__ __forceinline__ void adj_state_stream(AdjState &st, const AdjArgs &J, const int (&nb)[19], int idx,
                                                 const double (&lam)[19], const double (&mu)[19]) {
  if constexpr (A_ < END) {
    const pair_t<double> v =
        ld_stream<true>((const pair_t<double> *)elem<OFF32>(J.A.F, pull_off<ODD>(nb, 0, A_, idx)));
    adj_state_add<A_>(st, v.x, v.y, lam[A_], mu[A_]);
    adj_state_stream<A_ + 1, END, ODD, OFF32>(st, J, nb, idx, lam, mu);
  }
}

// Transposed collision at one node.  Flow f* = (1 - w) f + w feq(drho, u),
// solute g* = (1 - ws) g + ws w c (1 + 3 c.u), u = j/rho_A + tau g.
// Returns the new adjoint through lb(b) = (1 - w) lam_b + w Lr + c_b.Gr and
// mb(b) = (1 - ws) mu_b + ws Sg (+ the flux seed), and the parameter terms.
struct AdjOut {
  double Lr, Sg, Gr[3], acc[5];
};

__device__ __forceinline__ AdjOut adj_collide(const AdjArgs &J, const AdjMoments &m, const AdjState &st) {
  const StepArgs<double, double> &A = J.A;
  const double rhoA = A.rhoA, tau = A.tau, om = A.inv_tau, oms = A.inv_tau_s;
  double u[3];
#pragma unroll
  for (int d = 0; d < 3; d++) u[d] = st.j[d] / rhoA + tau * A.g[d];
  const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  // solute: Sg = sum mu w (1 + 3 c.u), dts = sum mu (geq - g)
  const double Sg = m.M0 + 3.0 * (u[0] * m.M1[0] + u[1] * m.M1[1] + u[2] * m.M1[2]);
  const double dts = st.c * Sg - st.mg;
  // flow: Lu = sum lam w rhoA (3 c + 9 (c.u) c - 3 u), dt = sum lam (feq - f)
  double L2u[3];
  L2u[0] = m.L2[0] * u[0] + m.L2[3] * u[1] + m.L2[4] * u[2];
  L2u[1] = m.L2[3] * u[0] + m.L2[1] * u[1] + m.L2[5] * u[2];
  L2u[2] = m.L2[4] * u[0] + m.L2[5] * u[1] + m.L2[2] * u[2];
  double Lu[3];
#pragma unroll
  for (int d = 0; d < 3; d++) Lu[d] = rhoA * (3.0 * m.L1[d] + 9.0 * L2u[d] - 3.0 * u[d] * m.Lr);
  const double uL1 = u[0] * m.L1[0] + u[1] * m.L1[1] + u[2] * m.L1[2];
  const double uL2u = u[0] * L2u[0] + u[1] * L2u[1] + u[2] * L2u[2];
  const double dt = st.drho * m.Lr + rhoA * (3.0 * uL1 + 4.5 * uL2u - 1.5 * usq * m.Lr) - st.lf;
  AdjOut o;
  double G[3];
#pragma unroll
  for (int d = 0; d < 3; d++) G[d] = om * Lu[d] + 3.0 * oms * st.c * m.M1[d];
#pragma unroll
  for (int d = 0; d < 3; d++) o.Gr[d] = G[d] / rhoA;
  o.Lr = m.Lr;
  o.Sg = Sg;
#pragma unroll
  for (int d = 0; d < 3; d++) o.acc[d] = tau * G[d];
  o.acc[3] = -om * om * dt + A.g[0] * G[0] + A.g[1] * G[1] + A.g[2] * G[2];
  o.acc[4] = -oms * oms * dts;
  return o;
}

// new adjoint of direction B_ at the pull position; seed: dJ/d(plane flux)
// at the flux plane, oseed: dJ/dc of an observation at this node (c = sum
// of the pre-collision solute populations)
template <int B_>
__device__ __forceinline__ void adj_new(const AdjArgs &J, const AdjOut &o, double seed, double oseed, double lam,
                                        double mu, double &lb, double &mb) {
  const double om = J.A.inv_tau, oms = J.A.inv_tau_s;
  lb = (1.0 - om) * lam + om * o.Lr;
  if (CX(B_) != 0) lb += CX(B_) > 0 ? o.Gr[0] : -o.Gr[0];
  if (CY(B_) != 0) lb += CY(B_) > 0 ? o.Gr[1] : -o.Gr[1];
  if (CZ(B_) != 0) lb += CZ(B_) > 0 ? o.Gr[2] + J.kseed : -o.Gr[2] - J.kseed;
  mb = (1.0 - oms) * mu + oms * o.Sg;
  if (CZ(B_) != 0) mb += CZ(B_) > 0 ? seed : -seed;
  mb += oseed;
}

template <int B_, bool ODD, bool OFF32>
__device__ __forceinline__ void adj_store_all(const AdjArgs &J, const AdjOut &o, double seed, double oseed,
                                              const int (&nb)[19], int idx, const double (&lam)[19],
                                              const double (&mu)[19]) {
  if constexpr (B_ < 19) {
    double nl, nm;
    adj_new<B_>(J, o, seed, oseed, lam[B_], mu[B_], nl, nm);
    pair_t<double> v;
    v.x = nl;
    v.y = nm;
    st_stream<!ODD>((pair_t<double> *)elem<OFF32>(J.L, pull_off<ODD>(nb, 0, B_, idx)), v);
    adj_store_all<B_ + 1, ODD, OFF32>(J, o, seed, oseed, nb, idx, lam, mu);
  }
}

template <int B_>
__device__ __forceinline__ void adj_new_all(const AdjArgs &J, const AdjOut &o, double seed, double oseed,
                                            const double (&lam)[19], const double (&mu)[19], double (&lb)[19],
                                            double (&mb)[19]) {
  if constexpr (B_ < 19) {
    adj_new<B_>(J, o, seed, oseed, lam[B_], mu[B_], lb[B_], mb[B_]);
    adj_new_all<B_ + 1>(J, o, seed, oseed, lam, mu, lb, mb);
  }
}

template <bool ODD>
__device__ __forceinline__ void adj_neighbours(const StepArgs<double, double> &A, int idx, int (&nb)[19], int &k) {
  nb[0] = idx;
  int i, j;
  coords<3>(A.G, (uint32_t)A.G.node[idx], i, j, k);
  if (ODD) {
    int r[18];
    neighbors_rows<3>(A.G, i, j, k, r);
#pragma unroll
    for (int a = 1; a < 19; a++) nb[a] = r[a - 1];
  }
}

__device__ __forceinline__ void adj_block_sum(const AdjArgs &J, long slot, double (&acc)[5]) {
  __shared__ double red[5];
  const int kinds[5] = {0, 0, 0, 0, 0};
  block_reduce<5, kAdjBlock / 64>(acc, kinds, red);
  if (threadIdx.x < 5) J.gacc[slot * 5 + threadIdx.x] += red[threadIdx.x];
}

// transpose of the flow z-face condition on the new adjoint of a face node:
// the unknowns' adjoint goes to the known populations it was formed from
__device__ __forceinline__ void adj_zface_flow(bool top, double (&l)[19]) {
  double lu[5], dl[19];
#pragma unroll
  for (int a = 0; a < 19; a++) dl[a] = 0.0;
  if (top) {
#pragma unroll
    for (int q = 0; q < 5; q++) lu[q] = l[kDn[q]];
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) lu[q] = l[kUp[q]];
  }
  double lJ, lNx, lNy;
  if (!top) {
    // J = drb - (P + 2M); v9 = f10 + J/3, v11 = f12 + J/6 - Nx, v13 = f14 + J/6 + Nx,
    // v15 = f16 + J/6 - Ny, v17 = f18 + J/6 + Ny
    lJ = lu[0] / 3.0 + (lu[1] + lu[2] + lu[3] + lu[4]) / 6.0;
    lNx = -lu[1] + lu[2];
    lNy = -lu[3] + lu[4];
#pragma unroll
    for (int q = 0; q < 5; q++) dl[kDn[q]] += lu[q] - 2.0 * lJ;
#pragma unroll
    for (int a = 0; a < 9; a++) dl[a] -= lJ;
  } else {
    // J = (P + 2M) - drb; v10 = f9 - J/3, v12 = f11 - J/6 + Nx, v14 = f13 - J/6 - Nx,
    // v16 = f15 - J/6 + Ny, v18 = f17 - J/6 - Ny
    lJ = -lu[0] / 3.0 - (lu[1] + lu[2] + lu[3] + lu[4]) / 6.0;
    lNx = lu[1] - lu[2];
    lNy = lu[3] - lu[4];
#pragma unroll
    for (int q = 0; q < 5; q++) dl[kUp[q]] += lu[q] + 2.0 * lJ;
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
#pragma unroll
  for (int a = 0; a < 19; a++) l[a] += dl[a];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const double od = l[kDn[q]], ou = l[kUp[q]];
    l[kDn[q]] = top ? 0.0 : od;
    l[kUp[q]] = top ? ou : 0.0;
  }
}

// transpose of the solute z-face condition; returns dJ/dc_b of the node
__device__ __forceinline__ double adj_zface_solute(bool top, double (&l)[19]) {
  double lcs = 0.0;
  if (top) {
#pragma unroll
    for (int q = 0; q < 5; q++) lcs += W<19>(kDn[q]) * l[kDn[q]];
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) lcs += W<19>(kUp[q]) * l[kUp[q]];
  }
  const double k6 = lcs / (1. / 6.);
  // bottom: the known in-plane and down-movers lose k6, the up-movers are
  // cleared; top: the up-movers gain k6, the down-movers are cleared
  // (value selects only: the elements stay registers)
  const double kb = top ? 0.0 : k6;
#pragma unroll
  for (int a = 0; a < 9; a++) l[a] -= kb;
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const double od = l[kDn[q]], ou = l[kUp[q]];
    l[kDn[q]] = top ? 0.0 : od - k6;
    l[kUp[q]] = top ? ou + k6 : 0.0;
  }
  return top ? 0.0 : k6;
}

// Adjoint of one half-step at the fluid nodes of the two z faces, with the
// face conditions of the forward step formed from the state (as the fused
// forward half-step formed them) and their transpose applied before the
// store.  Thread t < lo_end takes compact node t, the others the top face.
template <bool ODD, bool OFF32>
__device__ __forceinline__ void adj_face_block(const AdjArgs &J, long slot, int fb) {
  const StepArgs<double, double> &A = J.A;
  const int t = fb * kAdjBlock + (int)threadIdx.x;
  const int nlo = A.zb.lo_end, nface = nlo + (A.nf - A.zb.hi_begin);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (t < nface) {
    const int idx = t < nlo ? t : A.zb.hi_begin + (t - nlo);
    const bool top = idx >= A.zb.hi_begin;
    int nb[19], k;
    adj_neighbours<ODD>(A, idx, nb, k);
    double lam[19], mu[19];
    AdjMoments m;
    adj_load<ODD, OFF32>(J, nb, idx, lam, mu, m);
    double f[19], g[19];
#pragma unroll
    for (int a = 0; a < 19; a++) {
      const pair_t<double> v = *(const pair_t<double> *)elem<OFF32>(A.F, pull_off<ODD>(nb, 0, a, idx));
      f[a] = v.x;
      g[a] = v.y;
    }
    zface_fix<double, double, true>(A, idx, f, g);
    AdjState st;
    adj_state_arrays<0, 19>(st, f, g, lam, mu);
    const AdjOut o = adj_collide(J, m, st);
    double lb[19], mb[19];
    adj_new_all<0>(J, o, k == J.zf ? J.seed : 0.0, J.obs ? J.obs[idx] : 0.0, lam, mu, lb, mb);
    // the adjoint runs with pressure on both faces (or none) and a
    // Dirichlet inlet with a zero-gradient outlet (or none)
    if (A.zb.flow[0] != kZNone) adj_zface_flow(top, lb);
    if (A.zb.conc[0] != kCNone) {
      const double dcb = adj_zface_solute(top, mb);
      if (!top && J.cgrad) J.cgrad[A.G.node[idx]] += dcb;
    }
#pragma unroll
    for (int b = 0; b < 19; b++) {
      pair_t<double> v;
      v.x = lb[b];
      v.y = mb[b];
      *(pair_t<double> *)elem<OFF32>(J.L, pull_off<ODD>(nb, 0, b, idx)) = v;
    }
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = o.acc[q];
  }
  adj_block_sum(J, slot, acc);
}

// Adjoint of one half-step.  With A.zb.on the first nfb blocks (a multiple
// of 8, so the XCD mapping of the others is unchanged) take the z-face
// nodes (adj_face_block) and the others every node off the faces; without
// it (nfb = 0) every block takes the nodes of its range.  The node block of
// the others comes from the row-band map J.amap (a permutation of all
// 8 * nblk8 blocks) when there is one, else from the XCD-contiguous remap.
// Gradient partials: slot b of face block b, slot nfb + node block of the
// others (so the block order does not change the sums).
template <bool ODD, bool OFF32>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(ODD ? LBM_ADJ_WAVES_ODD : LBM_ADJ_WAVES_EVEN)))
void k_adjoint_step(AdjArgs J, int nfb) {
  if ((int)blockIdx.x < nfb) {
    adj_face_block<ODD, OFF32>(J, blockIdx.x, blockIdx.x);
    return;
  }
  const int lb = J.amap ? J.amap[blockIdx.x - nfb] : xcd_block(J.A.nblk8, (int)blockIdx.x - nfb);
  const int idx = lb * kAdjBlock + (int)threadIdx.x;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (idx < J.A.nf && !zface_node<19>(J.A, idx)) {
    int nb[19], k;
    adj_neighbours<ODD>(J.A, idx, nb, k);
    double lam[19], mu[19];
    AdjMoments m;
    adj_load<ODD, OFF32>(J, nb, idx, lam, mu, m);
    AdjState st;
    adj_state_stream<0, 19, ODD, OFF32>(st, J, nb, idx, lam, mu);
    const AdjOut o = adj_collide(J, m, st);
    adj_store_all<0, ODD, OFF32>(J, o, k == J.zf ? J.seed : 0.0, J.obs ? J.obs[idx] : 0.0, nb, idx, lam, mu);
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = o.acc[q];
  }
  adj_block_sum(J, nfb + (long)lb, acc);
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
    if (Z.flow != kZNone) adj_face<true>(B.L, nb, 0, idx, Z.top, true, dcb);
    if (Z.conc != kCNone) adj_face<true>(B.L, nb, 1, idx, Z.top, false, dcb);
  } else {
    if (Z.flow != kZNone) adj_face<false>(B.L, nb, 0, idx, Z.top, true, dcb);
    if (Z.conc != kCNone) adj_face<false>(B.L, nb, 1, idx, Z.top, false, dcb);
  }
  if (!Z.top && Z.conc != kCNone && cgrad) cgrad[B.A.G.node[idx]] += dcb;
}

// Observation points of the objective (compact node of each; the
// concentration is the solute mv value a PEST-mode step wrote).
__global__ void k_obs_gather(const double *mv_c, const int *node, int n, double *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = mv_c[node[q]];
}

// put (or, with val = null, clear) this step's dJ/dc at its observation nodes
__global__ void k_obs_seed(double *seed, const int *node, const double *val, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) seed[node[q]] = val ? val[q] : 0.0;
}

}  // namespace lbm
