/*
 * lbm_oracle.c -- CPU oracle (TEST INFRASTRUCTURE, never shipped or measured
 * as the product).  See lbm_oracle.h for the reference functions restated.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  The
 * arithmetic of every population update is written in the same order as the
 * reference kernels so that an implementation that keeps that order without
 * fused multiply-adds reproduces it bit for bit.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   node update (odd/even), gather/scatter, bounce-back, slot swap, layout
 *       - independent two-lattice numpy model (pull streaming with
 *         np.roll, explicit bounce-back, textbook He-Luo equilibrium):
 *         agreement to 1e-14 over random states, masks and step counts;
 *       - rest state stays exactly zero; mass invariant to round-off;
 *       - block-order independence (OpenMP vs serial bitwise).
 *   relaxation / equilibrium (calc_f_tilde_d)
 *       - moments of the post-collision state: M0 = drho, M1 = rho_A u_eq,
 *         M2 = rho_A u u + drho/3 I (closed forms on random inputs).
 *   body force (apply_accel_mv)
 *       - impulse law: sum_x j(t) = sum_x j(0) + t rho_A g N_f in a
 *         periodic all-fluid box; shear-wave decay rate nu=(tau-1/2)/3;
 *         Poiseuille profile in a bounce-back channel.
 *   2-D Zou-He (k_sysbound_pressure_*), zero gradient (k_sysbound_zeroconcgrad_*)
 *       - boundary-node identities: rho = rho_b, u_x = 0 after the BC;
 *         zero-gradient: incoming = w_i c*, with c* from the outgoing set.
 *   3-D z-face Zou-He, solute Dirichlet / zero gradient (no reference code)
 *       - boundary-node identities (rho = rho_b, j_x = j_y = 0; c = c_b);
 *         steady flux equal across planes; 1-D advection-diffusion closed form.
 *   halo pack/unpack (k_bound_*)
 *       - halo mode with periodic self-exchange equals the wrapped run bitwise.
 *   reductions (compute.c)
 *       - brute-force numpy over the same mv arrays.
 *   Walsh partial bounce-back (WALSH_NS_ON path)
 *       - ns = 0 reproduces the plain update bitwise; ns = 1 reverses every
 *         population, so the state repeats exactly every two steps; mass is
 *         conserved for any ns; in a uniform-ns periodic box the total
 *         momentum follows J' = J ((1-ns)(1-w ns) - ns) + (1-ns)^2 rho_A g N.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q19 velocity set in the reference's order (opposite directions adjacent,
 * D2Q9 first): src/lbgpu_prime.h:78-84. */
static const int VX[19] = {0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0};
static const int VY[19] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1};
static const int VZ[19] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};

/* D2Q9 names (src/lbgpu_prime.h:91-100) */
enum { C_ = 0, E_ = 1, W_ = 2, N_ = 3, S_ = 4, NE_ = 5, SW_ = 6, NW_ = 7, SE_ = 8 };

struct orc {
  orc_params p;
  int Q, U, D;
  long N, E;
  long cs[20];
  double *f;            /* nsubs * cs[Q] + halo padding, see fidx() */
  long fbase;           /* offset so that index -E of slot 0 is valid */
  double *mv;           /* [subs][1+D][N] */
  unsigned char *sol;   /* [E | N | E] */
  double wt[19];
  double inv_tau[2];
  int time;
  double flux, fmass;
  double *ns;           /* [N] Walsh partial bounce-back fraction, or NULL */
  /* z-face conditions per face ([0] z = 0, [1] z = nk-1), resolved:
   * flow 0 none, 1 pressure, 2 velocity; solute 0 none, 1 Dirichlet,
   * 2 zero gradient; zfval: rho_b or rho_A u_z; zcval: Dirichlet c */
  int zflow[2], zconc[2];
  double zfval[2], zcval[2];
};

static int g_threads = 0;

int orc_num_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) { g_threads = n; }

static inline double *fslot(const orc *o, int s, int a) {
  return o->f + o->fbase + (long)s * o->cs[o->Q] + o->cs[a];
}

/* cumul_stride: src/latman.c:194-219.  Unbound directions are packed N
 * each; each bound pair is [E halo][N odd][N even][E halo]. */
static void build_strides(orc *o) {
  int a;
  long N = o->N, E = o->E;
  o->cs[0] = 0;
  for (a = 1; a < o->U; a++) o->cs[a] = o->cs[a - 1] + N;
  o->cs[a] = o->cs[a - 1] + N + E;
  a++;
  o->cs[a] = o->cs[a - 1] + N;
  a++;
  for (; a < o->Q; a += 2) {
    o->cs[a] = o->cs[a - 2] + 2 * N + 2 * E;
    o->cs[a + 1] = o->cs[a - 1] + 2 * N + 2 * E;
  }
  o->cs[a] = o->cs[a - 1] + N + E;
}

orc *orc_create(const orc_params *p) {
  orc *o = (orc *)calloc(1, sizeof(orc));
  int a;
  if (!o) return NULL;
  o->p = *p;
  o->D = p->dims;
  o->Q = p->dims == 2 ? 9 : 19;
  o->U = p->dims == 2 ? 3 : 9;
  o->N = (long)p->ni * p->nj * (p->dims == 2 ? 1 : p->nk);
  o->E = p->dims == 2 ? p->ni : (long)p->ni * p->nj;
  if (p->dims == 2) o->p.nk = 1;
  build_strides(o);
  o->fbase = o->E;
  o->f = (double *)calloc((size_t)(p->nsubs * o->cs[o->Q] + 2 * o->E), sizeof(double));
  o->mv = (double *)calloc((size_t)(p->nsubs * (1 + o->D) * o->N), sizeof(double));
  o->sol = (unsigned char *)calloc((size_t)(o->N + 2 * o->E), 1);
  /* weights: src/latman.c:161-192 */
  if (o->D == 2) {
    double W0 = 4. / 9., W1 = 1. / 9., W2 = 1. / 36.;
    o->wt[0] = W0;
    for (a = 1; a < 5; a++) o->wt[a] = W1;
    for (; a < 9; a++) o->wt[a] = W2;
    for (; a < 19; a++) o->wt[a] = 0.;
  } else {
    double W0 = 1. / 3., W1 = 1. / 18., W2 = 1. / 36.;
    o->wt[0] = W0;
    for (a = 1; a < 5; a++) o->wt[a] = W1;
    for (; a < 9; a++) o->wt[a] = W2;
    for (; a < 11; a++) o->wt[a] = W1;
    for (; a < 19; a++) o->wt[a] = W2;
  }
  for (a = 0; a < p->nsubs; a++) o->inv_tau[a] = 1. / p->tau[a];
  for (a = 0; a < 2; a++) {
    int fk = p->zface_flow[a], ck = p->zface_conc[a];
    o->zflow[a] = fk == 2 ? 1 : fk == 3 ? 2 : fk == 1 ? 0 : (p->zface_pressure ? 1 : 0);
    o->zfval[a] = o->zflow[a] == 2 ? p->rho_A[0] * p->zface_uz[a] : (a ? p->rho_top : p->rho_bot);
    o->zconc[a] = p->nsubs != 2 ? 0 : ck == 2 ? 1 : ck == 3 ? 2 : ck == 1 ? 0 : (p->zface_solute ? (a ? 2 : 1) : 0);
    o->zcval[a] = a ? p->c_top : p->c_bot;
    if (p->dims != 3) o->zflow[a] = o->zconc[a] = 0;
  }
  o->time = 0;
  return o;
}

void orc_destroy(orc *o) {
  if (!o) return;
  free(o->f);
  free(o->mv);
  free(o->sol);
  free(o->ns);
  free(o);
}

int orc_set_ns(orc *o, const double *ns) {
  free(o->ns);
  o->ns = NULL;
  if (!ns) return 0;
  o->ns = (double *)malloc(sizeof(double) * o->N);
  if (!o->ns) return 1;
  memcpy(o->ns, ns, sizeof(double) * o->N);
  return 0;
}

void orc_set_time(orc *o, int t) { o->time = t; }
int orc_get_time(const orc *o) { return o->time; }
long orc_raw_size(const orc *o) { return o->cs[o->Q]; }

/* solid halos: periodic copy (src/lbgpu_prime.cu:209-224).  The reference
 * then zeroes the halo planes of the global ends (:230-262); with one rank
 * and a periodic self-exchange the periodic copy is what the exchange
 * delivers, so that is what is kept here. */
void orc_set_solids(orc *o, const uint8_t *mask) {
  long E = o->E, N = o->N;
  memcpy(o->sol + E, mask, (size_t)N);
  memcpy(o->sol + E + N, o->sol + E, (size_t)E);
  memcpy(o->sol, o->sol + N, (size_t)E);
}

static inline int is_solid(const orc *o, long n) { return o->sol[n + o->E] == 1; }

void orc_set_f(orc *o, int s, const double *f1d) {
  for (int a = 0; a < o->Q; a++) memcpy(fslot(o, s, a), f1d + (long)a * o->N, sizeof(double) * o->N);
}

void orc_get_f(const orc *o, int s, double *f1d) {
  for (int a = 0; a < o->Q; a++) memcpy(f1d + (long)a * o->N, fslot(o, s, a), sizeof(double) * o->N);
}

void orc_get_raw(const orc *o, int s, double *out) {
  memcpy(out, o->f + o->fbase + (long)s * o->cs[o->Q], sizeof(double) * o->cs[o->Q]);
}

/* init_problem with u = 0 and initial_condition 0: src/latman.c:1786-1893.
 * Flow (ACCURATE): f~_a = w_a rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2) = 0.
 * Solute: g_a = w_a c (1 + 3 c.u + ...) = w_a c. */
void orc_init_rest(orc *o, const double *c0) {
  long n;
  int a;
  for (a = 0; a < o->Q; a++) {
    double *f0 = fslot(o, 0, a);
    for (n = 0; n < o->N; n++) f0[n] = o->wt[a] * 1.0 * (3. * 0. + 4.5 * 0. * 0. - 1.5 * 0.);
  }
  if (o->p.nsubs == 2) {
    for (a = 0; a < o->Q; a++) {
      double *g = fslot(o, 1, a);
      for (n = 0; n < o->N; n++) {
        double c = c0 ? c0[n] : 0.;
        g[n] = o->wt[a] * c * (1. + 3. * 0. + 4.5 * 0. * 0. - 1.5 * 0.);
      }
    }
  }
}

/* ---- neighbour addressing: get_f1d_d / set_f1d_d (src/lattice.c) ---- */

static inline long nb_index(const orc *o, int i, int j, int k) {
  const orc_params *p = &o->p;
  if (i < 0) i += p->ni;
  if (i == p->ni) i = 0;
  if (o->D == 3 || !p->halo_mode) {
    if (j < 0) j += p->nj;
    if (j == p->nj) j = 0;
  }
  if (!p->halo_mode) {
    if (k < 0) k += p->nk;
    if (k == p->nk) k = 0;
  }
  return (long)i + (long)p->ni * j + (long)p->ni * p->nj * k;
}

/* value of f_a at (i0+di, j0+dj, k0+dk), or the own slot a+da if that
 * neighbour is solid (halfway bounce-back): src/lattice.c:2660-2744 */
static inline double get_f(const orc *o, int s, int i0, int j0, int k0, long n0, int di, int dj,
                           int dk, int a, int da) {
  long n = nb_index(o, i0 + di, j0 + dj, k0 + dk);
  if (!is_solid(o, n)) return fslot(o, s, a)[n];
  return fslot(o, s, a + da)[n0];
}

/* store into slot a+da at the neighbour, or slot a of the own node if the
 * neighbour is solid: src/lattice.c:2765-2848 */
static inline void set_f(orc *o, int s, int i0, int j0, int k0, long n0, int di, int dj, int dk,
                         int a, int da, double v) {
  long n = nb_index(o, i0 + di, j0 + dj, k0 + dk);
  if (!is_solid(o, n))
    fslot(o, s, a + da)[n] = v;
  else
    fslot(o, s, a)[n0] = v;
}

/* gather as the odd half-step does (stream from x - c_a) */
static inline void gather_pull(const orc *o, int s, int i, int j, int k, long n, double *f) {
  f[0] = get_f(o, s, i, j, k, n, 0, 0, 0, 0, 0);
  for (int a = 1; a < o->Q; a += 2) f[a] = get_f(o, s, i, j, k, n, -VX[a], -VY[a], -VZ[a], a, 1);
  for (int a = 2; a < o->Q; a += 2) f[a] = get_f(o, s, i, j, k, n, -VX[a], -VY[a], -VZ[a], a, -1);
}

/* gather as the even half-step does (own node, opposite slots swapped) */
static inline void gather_local(const orc *o, int s, long n, double *f) {
  f[0] = fslot(o, s, 0)[n];
  for (int a = 1; a < o->Q; a += 2) f[a] = fslot(o, s, a + 1)[n];
  for (int a = 2; a < o->Q; a += 2) f[a] = fslot(o, s, a - 1)[n];
}

static inline void set_mv(orc *o, int s, long n, int c, double v) {
  o->mv[(long)s * o->N * (1 + o->D) + (long)c * o->N + n] = v;
}

/* One node of k_stream_collide_stream (odd) or k_collide (even).
 * Adds the plane-flux contribution to *jflux and *jmass. */
static void update_node(orc *o, int odd, int t, int write_mv, int i, int j, int k, double *jflux,
                        double *jmass) {
  const orc_params *p = &o->p;
  const int Q = o->Q, D = o->D;
  long n = (long)i + (long)p->ni * j + (long)p->ni * p->nj * k;
  double f[19], u[3] = {0., 0., 0.}, rho = 0., usq = 0.;
  int s, a, plane = D == 2 ? j : k;

  if (is_solid(o, n)) return;
  for (s = 0; s < p->nsubs; s++) {
    if (s == 1 && !(t >= p->sigma_t_on && t <= p->sigma_t_off)) continue;
    if (odd)
      gather_pull(o, s, i, j, k, n, f);
    else
      gather_local(o, s, n, f);

    if (s == 1) {
      /* solute: concentration only; u is the flow's u_eq (D8) */
      rho = f[0];
      for (a = 1; a < Q; a++) rho += f[a];
      if (!odd) rho *= (1.0 - p->sink * p->tau[1]); /* k_collide.c:96-97 */
      if (plane == p->flux_plane) {
        double jl = 0., m = 0.;
        for (a = 0; a < Q; a++) {
          jl += (double)(D == 2 ? VY[a] : VZ[a]) * f[a];
          m += f[a];
        }
        *jflux += jl;
        *jmass += m;
      }
    } else {
      rho = f[0];
      for (a = 1; a < Q; a++) rho += f[a];
      u[0] = 0.;
      u[1] = 0.;
      u[2] = 0.;
      for (a = 0; a < Q; a++) {
        u[0] += ((double)VX[a]) * f[a];
        u[1] += ((double)VY[a]) * f[a];
        if (D == 3) u[2] += ((double)VZ[a]) * f[a];
      }
      /* Walsh partial bounce-back scales the momentum by 1 - ns
       * (src/k_stream_collide_stream.c:129-135,183-192) */
      if (o->ns)
        for (a = 0; a < D; a++) u[a] *= (1. - o->ns[n]);
      for (a = 0; a < D; a++) u[a] /= p->rho_A[s];
    }
    /* body force / sink: src/k_stream_collide_stream.c:237-258; with ns
     * the force too is scaled by 1 - ns (apply_accel_mv, src/lattice.c:2936-2939) */
    if (s != 1) {
      if (o->ns)
        for (a = 0; a < D; a++) u[a] += (1. - o->ns[n]) * p->tau[s] * p->gaccel[s * 3 + a];
      else
        for (a = 0; a < D; a++) u[a] += p->tau[s] * p->gaccel[s * 3 + a];
    } else {
      rho *= (1.0 - p->sink * p->tau[1]);
    }
    if (write_mv == 2) {
      /* pest_output_flag_c off a frame end: flow u only, solute c only
       * (src/k_stream_collide_stream.c:305-334, src/k_collide.c:313-340) */
      if (s == 0)
        for (a = 0; a < D; a++) set_mv(o, s, n, a + 1, u[a]);
      else
        set_mv(o, s, n, 0, rho);
    } else if (write_mv) {
      set_mv(o, s, n, 0, s == 0 ? rho + p->rho_A[s] : rho);
      for (a = 0; a < D; a++) set_mv(o, s, n, a + 1, u[a]);
    }
    /* relaxation: calc_f_tilde_d, src/lattice.c:2850-2924 */
    if (s != 1) {
      usq = u[0] * u[0] + u[1] * u[1];
      if (D == 3) usq += u[2] * u[2];
    }
    double fpre[19];
    if (o->ns)
      for (a = 0; a < Q; a++) fpre[a] = f[a];
    for (a = 0; a < Q; a++) {
      double vdotu;
      f[a] *= (1. - o->inv_tau[s]);
      vdotu = ((double)VX[a]) * u[0] + ((double)VY[a]) * u[1];
      if (D == 3) vdotu += ((double)VZ[a]) * u[2];
      if (s == 1) {
        f[a] += o->wt[a] * rho * (1. + 3. * vdotu) * o->inv_tau[s];
      } else {
        f[a] += o->wt[a] * rho * o->inv_tau[s];
        f[a] += o->wt[a] * p->rho_A[s] * (3. * vdotu + 4.5 * vdotu * vdotu - 1.5 * usq) * o->inv_tau[s];
      }
    }
    /* Walsh: post-collision (1 - ns) BGK + ns times the opposite pre-collision
     * population (src/k_stream_collide_stream.c:363-383, src/k_collide.c:373-392) */
    if (o->ns) {
      const double ns = o->ns[n];
      f[0] *= 1. - ns;
      f[0] += ns * fpre[0];
      for (a = 1; a < Q; a += 2) {
        f[a] *= 1. - ns;
        f[a] += ns * fpre[a + 1];
        f[a + 1] *= 1. - ns;
        f[a + 1] += ns * fpre[a];
      }
    }
    if (odd) {
      set_f(o, s, i, j, k, n, 0, 0, 0, 0, 0, f[0]);
      for (a = 1; a < Q; a += 2) set_f(o, s, i, j, k, n, VX[a], VY[a], VZ[a], a, 1, f[a]);
      for (a = 2; a < Q; a += 2) set_f(o, s, i, j, k, n, VX[a], VY[a], VZ[a], a, -1, f[a]);
    } else {
      for (a = 0; a < Q; a++) fslot(o, s, a)[n] = f[a];
    }
  }
}

static void run_kernel(orc *o, int odd, int t, int write_mv) {
  const orc_params *p = &o->p;
  int outer = o->D == 2 ? p->nj : p->nk;
  double jf = 0., jm = 0.;
  int nthreads = orc_num_threads();
  (void)nthreads;
#pragma omp parallel for schedule(static) reduction(+ : jf, jm) num_threads(nthreads)
  for (int q = 0; q < outer; q++) {
    if (o->D == 2) {
      for (int i = 0; i < p->ni; i++) update_node(o, odd, t, write_mv, i, q, 0, &jf, &jm);
    } else {
      for (int j = 0; j < p->nj; j++)
        for (int i = 0; i < p->ni; i++) update_node(o, odd, t, write_mv, i, j, q, &jf, &jm);
    }
  }
  o->flux = jf;
  o->fmass = jm;
}

void orc_last_flux(const orc *o, double *flux, double *mass) {
  *flux = o->flux;
  *mass = o->fmass;
}

/* ---- 2-D system boundary conditions (reference) ---- */

/* rho selection with time series and ramp: src/bcs.c:287-319 */
static double bc_rho(const orc *o, int s, int which, int t) {
  const orc_params *p = &o->p;
  double rho, ramp;
  if (which == 0 && p->pressure_n_in[s] == 2 && p->pressure_n_series[s])
    rho = p->pfadd + p->pfmul * p->pressure_n_series[s][t % p->pressure_n_series_len[s]];
  else if (which == 2 && p->pressure_s_in[s] == 2 && p->pressure_s_series[s])
    rho = p->pfadd + p->pfmul * p->pressure_s_series[s][t % p->pressure_s_series_len[s]];
  else if (which == 0 || which == 2)
    rho = p->rho_in;
  else
    rho = p->rho_out;
  if (p->bc_ramp_start >= 0 && p->bc_ramp_stop >= 0) {
    if (t < p->bc_ramp_start) {
      rho = p->rho_A[s];
    } else if (t < p->bc_ramp_stop) {
      ramp = ((double)(t - p->bc_ramp_start)) / ((double)(p->bc_ramp_stop - p->bc_ramp_start));
      rho *= ramp;
      rho += (1. - ramp) * p->rho_A[s];
    }
  }
  return rho;
}

/* u_y of a velocity row (which: 0 n_in, 1 n_out, 2 s_in, 3 s_out) with the
 * pressure rows' ramp, from 0 instead of from rho_A (no reference code: the
 * reference parses the velocity_{n,s} flags and never applies them) */
static double bc_uy(const orc *o, int which, int t) {
  const orc_params *p = &o->p;
  double u = (which == 0 || which == 2) ? p->uy_in : p->uy_out, ramp;
  if (p->bc_ramp_start >= 0 && p->bc_ramp_stop >= 0) {
    if (t < p->bc_ramp_start) {
      u = 0.0;
    } else if (t < p->bc_ramp_stop) {
      ramp = ((double)(t - p->bc_ramp_start)) / ((double)(p->bc_ramp_stop - p->bc_ramp_start));
      u *= ramp;
    }
  }
  return u;
}

/* phase-1 store of a velocity row's unknown a at node (i, j): slot a at
 * the pull source x - c_a, or the own slot opp(a) when that source is solid
 * (the unknown replaces the bounced population; the 3-D faces' rule) */
static void row_write_vel(orc *o, int i, int j, long n, int a, double v) {
  int oa = (a & 1) ? a + 1 : a - 1;
  long src = nb_index(o, i - VX[a], j - VY[a], 0);
  if (!is_solid(o, src))
    fslot(o, 0, a)[src] = v;
  else
    fslot(o, 0, oa)[n] = v;
}

/* k_sysbound_pressure_{n,s}_{1,2}: src/system_boundary_kernels.c:4-451.
 * velocity: the velocity row instead, x = rho_b = rho_A u_y being the row's
 * prescribed y momentum; with x the node's y momentum in both forms
 * (north: x = P + 2M - (rho_b - rho_A) from the mass balance), the three
 * unknowns follow from x the same way.  The reference has no velocity row,
 * so its phase-1 placement rule (set_f1d_d: own slot a when the source is
 * solid) is not followed there: row_write_vel. */
static void bc2d_pressure(orc *o, int s, int north, int phase, double rho_b, int velocity) {
  const orc_params *p = &o->p;
  int j = north ? p->nj - 1 : 0;
  for (int i = 0; i < p->ni; i++) {
    long n = (long)i + (long)p->ni * j;
    double f[9], x;
    if (is_solid(o, n)) continue;
    if (phase == 1)
      gather_pull(o, s, i, j, 0, n, f);
    else
      gather_local(o, s, n, f);
    if (north) {
      x = f[C_] + f[E_] + f[W_] + f[N_] + f[N_] + f[NE_] + f[NE_] + f[NW_] + f[NW_];
      x -= rho_b;
      x += p->rho_A[s];
      if (velocity) x = rho_b;
      f[S_] = f[N_] - (2. / 3.) * x;
      f[SW_] = f[NE_] + (1. / 2.) * (f[E_] - f[W_]) - (1. / 6.) * x;
      f[SE_] = f[NW_] + (1. / 2.) * (f[W_] - f[E_]) - (1. / 6.) * x;
      if (phase == 1 && velocity) {
        row_write_vel(o, i, j, n, S_, f[S_]);
        row_write_vel(o, i, j, n, SE_, f[SE_]);
        row_write_vel(o, i, j, n, SW_, f[SW_]);
      } else if (phase == 1) {
        set_f(o, s, i, j, 0, n, -VX[S_], -VY[S_], 0, S_, 0, f[S_]);
        set_f(o, s, i, j, 0, n, -VX[SE_], -VY[SE_], 0, SE_, 0, f[SE_]);
        set_f(o, s, i, j, 0, n, -VX[SW_], -VY[SW_], 0, SW_, 0, f[SW_]);
      } else {
        fslot(o, s, N_)[n] = f[S_];
        fslot(o, s, NW_)[n] = f[SE_];
        fslot(o, s, NE_)[n] = f[SW_];
      }
    } else {
      x = -f[C_] - f[E_] - f[W_] - f[S_] - f[S_] - f[SE_] - f[SE_] - f[SW_] - f[SW_];
      x += rho_b;
      x -= p->rho_A[s];
      if (velocity) x = rho_b;
      f[N_] = f[S_] + (2. / 3.) * x;
      f[NE_] = f[SW_] + (1. / 2.) * (f[W_] - f[E_]) + (1. / 6.) * x;
      f[NW_] = f[SE_] + (1. / 2.) * (f[E_] - f[W_]) + (1. / 6.) * x;
      if (phase == 1 && velocity) {
        row_write_vel(o, i, j, n, N_, f[N_]);
        row_write_vel(o, i, j, n, NW_, f[NW_]);
        row_write_vel(o, i, j, n, NE_, f[NE_]);
      } else if (phase == 1) {
        set_f(o, s, i, j, 0, n, -VX[N_], -VY[N_], 0, N_, 0, f[N_]);
        set_f(o, s, i, j, 0, n, -VX[NW_], -VY[NW_], 0, NW_, 0, f[NW_]);
        set_f(o, s, i, j, 0, n, -VX[NE_], -VY[NE_], 0, NE_, 0, f[NE_]);
      } else {
        fslot(o, s, S_)[n] = f[N_];
        fslot(o, s, SE_)[n] = f[NW_];
        fslot(o, s, SW_)[n] = f[NE_];
      }
    }
  }
}

/* k_sysbound_zeroconcgrad_{n,s}_{1,2}: src/system_boundary_kernels.c:454-831 */
static void bc2d_zerograd(orc *o, int north, int phase) {
  const orc_params *p = &o->p;
  const int s = 1;
  int j = north ? p->nj - 1 : 0;
  for (int i = 0; i < p->ni; i++) {
    long n = (long)i + (long)p->ni * j;
    double f[9], c;
    if (is_solid(o, n)) continue;
    if (phase == 1)
      gather_pull(o, s, i, j, 0, n, f);
    else
      gather_local(o, s, n, f);
    if (north) {
      c = (f[N_] + f[NE_] + f[NW_]) / (o->wt[S_] + o->wt[SW_] + o->wt[SE_]);
      f[S_] = o->wt[S_] * c;
      f[SW_] = o->wt[SW_] * c;
      f[SE_] = o->wt[SE_] * c;
      if (phase == 1) {
        set_f(o, s, i, j, 0, n, -VX[S_], -VY[S_], 0, S_, 0, f[S_]);
        set_f(o, s, i, j, 0, n, -VX[SE_], -VY[SE_], 0, SE_, 0, f[SE_]);
        set_f(o, s, i, j, 0, n, -VX[SW_], -VY[SW_], 0, SW_, 0, f[SW_]);
      } else {
        fslot(o, s, N_)[n] = f[S_];
        fslot(o, s, NW_)[n] = f[SE_];
        fslot(o, s, NE_)[n] = f[SW_];
      }
    } else {
      c = (f[S_] + f[SW_] + f[SE_]) / (o->wt[N_] + o->wt[NE_] + o->wt[NW_]);
      f[N_] = o->wt[N_] * c;
      f[NE_] = o->wt[NE_] * c;
      f[NW_] = o->wt[NW_] * c;
      if (phase == 1) {
        set_f(o, s, i, j, 0, n, -VX[N_], -VY[N_], 0, N_, 0, f[N_]);
        set_f(o, s, i, j, 0, n, -VX[NW_], -VY[NW_], 0, NW_, 0, f[NW_]);
        set_f(o, s, i, j, 0, n, -VX[NE_], -VY[NE_], 0, NE_, 0, f[NE_]);
      } else {
        fslot(o, s, S_)[n] = f[N_];
        fslot(o, s, SE_)[n] = f[NW_];
        fslot(o, s, SW_)[n] = f[NE_];
      }
    }
  }
}

/* ---- 3-D z-face boundary conditions (no reference counterpart) ----
 *
 * Flow, Zou-He pressure with zero tangential velocity (the D3Q19 form of
 * Hecht & Harting, J. Stat. Mech. 2010, P01018), on shifted populations.
 * With P = sum of the nine in-plane populations, M = sum of the five that
 * leave through the face (known) and rho_b the prescribed density:
 *   bottom (z=0, unknown +z):  J = (rho_b - rho_A) - (P + 2 M)
 *     f_T  = f_B  + J/3
 *     f_TE = f_BW + J/6 - Nx,  f_TW = f_BE + J/6 + Nx
 *     f_TN = f_BS + J/6 - Ny,  f_TS = f_BN + J/6 + Ny
 *   top (z=nk-1, unknown -z):  J = (P + 2 M) - (rho_b - rho_A)
 *     f_B  = f_T  - J/3
 *     f_BW = f_TE - J/6 + Nx,  f_BE = f_TW - J/6 - Nx
 *     f_BS = f_TN - J/6 + Ny,  f_BN = f_TS - J/6 - Ny
 *   Nx = ((f_E + f_NE + f_SE) - (f_W + f_NW + f_SW)) / 2, Ny likewise in y.
 * The shift w_a rho_A cancels in every line because the weights of each
 * pair are equal and the weights of P + 2M sum to one.  J is the node's z
 * momentum sum_a c_a,z f~_a after the condition.
 *
 * Flow, Zou-He velocity (the reference's CPU path has the compressible
 * form, rho = (P + 2M)/(1 + u) with c = rho u, for the top inflow and the
 * bottom outflow, src/bcs.c:766-862): the same five lines with J the
 * prescribed z momentum rho_A u_z -- in the incompressible model the
 * velocity is j / rho_A -- and the density following from the unknowns.
 *
 * Solute, per face: Dirichlet concentration (c_bot or the map at z=0, c_top
 * at z=nk-1) -- the five unknown populations are w_a c*, c* = (c_b - sum of
 * the 14 known) / (sum of their weights = 1/6); zero gradient -- unknown
 * w_a c*, c* = (sum of the five outgoing) / (1/6), the 3-D form of
 * k_sysbound_zeroconcgrad_n.
 *
 * Phase 1 (before the odd half-step) writes each unknown where the pull
 * will read it: slot a at x - c_a, or the own slot opp(a) when x - c_a is
 * solid (so the unknown replaces the bounced population).  Phase 2 (before
 * the even half-step) writes the own slot opp(a). */
static const int UP5[5] = {9, 11, 13, 15, 17};
static const int DN5[5] = {10, 12, 14, 16, 18};

static void zface_write(orc *o, int s, int phase, int i, int j, int k, long n, int a, double v) {
  int oa = (a & 1) ? a + 1 : a - 1;
  if (phase == 1) {
    long src = nb_index(o, i - VX[a], j - VY[a], k - VZ[a]);
    if (!is_solid(o, src))
      fslot(o, s, a)[src] = v;
    else
      fslot(o, s, oa)[n] = v;
  } else {
    fslot(o, s, oa)[n] = v;
  }
}

static void zface_bcs(orc *o, int phase, int t) {
  const orc_params *p = &o->p;
  (void)t;
  for (int top = 0; top < 2; top++) {
    int k = top ? p->nk - 1 : 0;
    for (int j = 0; j < p->nj; j++)
      for (int i = 0; i < p->ni; i++) {
        long n = (long)i + (long)p->ni * j + (long)p->ni * p->nj * k;
        double f[19], g[19];
        if (is_solid(o, n)) continue;
        if (o->zflow[top]) {
          double P, M = 0., J, Nx, Ny, v[5];
          const int *kn = top ? UP5 : DN5;
          if (phase == 1)
            gather_pull(o, 0, i, j, k, n, f);
          else
            gather_local(o, 0, n, f);
          P = f[0];
          for (int a = 1; a < 9; a++) P += f[a];
          for (int q = 0; q < 5; q++) M += f[kn[q]];
          Nx = ((f[1] + f[5] + f[8]) - (f[2] + f[7] + f[6])) / 2.;
          Ny = ((f[3] + f[5] + f[7]) - (f[4] + f[6] + f[8])) / 2.;
          if (!top) {
            J = o->zflow[0] == 2 ? o->zfval[0] : (o->zfval[0] - p->rho_A[0]) - (P + 2. * M);
            v[0] = f[10] + J / 3.;
            v[1] = f[12] + J / 6. - Nx;
            v[2] = f[14] + J / 6. + Nx;
            v[3] = f[16] + J / 6. - Ny;
            v[4] = f[18] + J / 6. + Ny;
            for (int q = 0; q < 5; q++) zface_write(o, 0, phase, i, j, k, n, UP5[q], v[q]);
          } else {
            J = o->zflow[1] == 2 ? o->zfval[1] : (P + 2. * M) - (o->zfval[1] - p->rho_A[0]);
            v[0] = f[9] - J / 3.;
            v[1] = f[11] - J / 6. + Nx;
            v[2] = f[13] - J / 6. - Nx;
            v[3] = f[15] - J / 6. + Ny;
            v[4] = f[17] - J / 6. - Ny;
            for (int q = 0; q < 5; q++) zface_write(o, 0, phase, i, j, k, n, DN5[q], v[q]);
          }
        }
        if (o->zconc[top]) {
          const int dir = o->zconc[top] == 1;
          double cs, known = 0.;
          if (phase == 1)
            gather_pull(o, 1, i, j, k, n, g);
          else
            gather_local(o, 1, n, g);
          if (dir)
            for (int a = 0; a < 9; a++) known += g[a];
          if (!top) {
            double cb = p->c_bot_map ? p->c_bot_map[(long)i + (long)p->ni * j] : o->zcval[0];
            for (int q = 0; q < 5; q++) known += g[DN5[q]];
            cs = dir ? (cb - known) / (1. / 6.) : known / (1. / 6.);
            for (int q = 0; q < 5; q++) zface_write(o, 1, phase, i, j, k, n, UP5[q], o->wt[UP5[q]] * cs);
          } else {
            for (int q = 0; q < 5; q++) known += g[UP5[q]];
            cs = dir ? (o->zcval[1] - known) / (1. / 6.) : known / (1. / 6.);
            for (int q = 0; q < 5; q++) zface_write(o, 1, phase, i, j, k, n, DN5[q], o->wt[DN5[q]] * cs);
          }
        }
      }
  }
}

/* bcs_1 / bcs_2 order: per subs, pressure N in/out, S in/out, then the
 * solute zero-gradient N, S (src/bcs.c:278-555) */
void orc_bcs(orc *o, int phase, int t) {
  const orc_params *p = &o->p;
  if (o->D == 2) {
    for (int s = 0; s < p->nsubs; s++) {
      if (p->pressure_n_in[s]) bc2d_pressure(o, s, 1, phase, bc_rho(o, s, 0, t), 0);
      if (p->pressure_n_out[s]) bc2d_pressure(o, s, 1, phase, bc_rho(o, s, 1, t), 0);
      if (p->pressure_s_in[s]) bc2d_pressure(o, s, 0, phase, bc_rho(o, s, 2, t), 0);
      if (p->pressure_s_out[s]) bc2d_pressure(o, s, 0, phase, bc_rho(o, s, 3, t), 0);
      if (s == 0) {
        if (p->velocity_n_in) bc2d_pressure(o, 0, 1, phase, p->rho_A[0] * bc_uy(o, 0, t), 1);
        if (p->velocity_n_out) bc2d_pressure(o, 0, 1, phase, p->rho_A[0] * bc_uy(o, 1, t), 1);
        if (p->velocity_s_in) bc2d_pressure(o, 0, 0, phase, p->rho_A[0] * bc_uy(o, 2, t), 1);
        if (p->velocity_s_out) bc2d_pressure(o, 0, 0, phase, p->rho_A[0] * bc_uy(o, 3, t), 1);
      }
      if (s == 1 && p->zeroconcgrad_n) bc2d_zerograd(o, 1, phase);
      if (s == 1 && p->zeroconcgrad_s) bc2d_zerograd(o, 0, phase);
    }
  } else if (o->zflow[0] || o->zflow[1] || o->zconc[0] || o->zconc[1]) {
    zface_bcs(o, phase, t);
  }
}

/* ---- halo planes: k_bound_DtH_{1,2} / k_bound_HtD_{1,2} ----
 * message layout [subs][bound pair][E]  (src/boundary_kernels.c:1-139) */
void orc_halo_pack(const orc *o, int phase, double *pos, double *neg) {
  long E = o->E, N = o->N;
  int nb = (o->Q - o->U) / 2;
  for (int s = 0; s < o->p.nsubs; s++)
    for (int q = 0; q < nb; q++) {
      int a = o->U + 2 * q; /* odd: +y / +z */
      double *po = pos + ((long)s * nb + q) * E, *ne = neg + ((long)s * nb + q) * E;
      for (long m = 0; m < E; m++) {
        if (phase == 1) {
          po[m] = fslot(o, s, a)[N - E + m];
          ne[m] = fslot(o, s, a + 1)[m];
        } else {
          ne[m] = fslot(o, s, a)[-E + m];
          po[m] = fslot(o, s, a + 1)[N + m];
        }
      }
    }
}

/* Phase 2 stores a received value only where the neighbour's push
 * actually happened, i.e. where the pushing node across the face is fluid.
 * k_bound_HtD_2 stores every position (src/boundary_kernels.c:106-139),
 * which overwrites the own-slot bounce-back value of a face node whose
 * cross-face upstream node is solid with a stale copy; the masked form
 * keeps the exchanged run identical to the periodic one.  halo_literal
 * selects the reference's form. */
void orc_halo_unpack(orc *o, int phase, const double *pos, const double *neg) {
  long E = o->E, N = o->N;
  int nb = (o->Q - o->U) / 2;
  int ni = o->p.ni, nj = o->D == 3 ? o->p.nj : 1;
  for (int s = 0; s < o->p.nsubs; s++)
    for (int q = 0; q < nb; q++) {
      int a = o->U + 2 * q;
      const double *po = pos + ((long)s * nb + q) * E, *ne = neg + ((long)s * nb + q) * E;
      for (long m = 0; m < E; m++) {
        if (phase == 1) {
          fslot(o, s, a)[-E + m] = po[m];
          fslot(o, s, a + 1)[N + m] = ne[m];
        } else {
          int i = (int)(m % ni), j = (int)(m / ni);
          int iu = (i + VX[a] + ni) % ni, ju = (j + VY[a] + nj) % nj;
          int id = (i - VX[a] + ni) % ni, jd = (j - VY[a] + nj) % nj;
          const int lit = o->p.halo_literal;
          if (lit || o->sol[o->E + N + iu + (long)ni * ju] != 1) fslot(o, s, a)[N - E + m] = ne[m];
          if (lit || o->sol[id + (long)ni * jd] != 1) fslot(o, s, a + 1)[m] = po[m];
        }
      }
    }
}

/* one rank with periodic wrap: what is sent to rank+1 / rank-1 comes back */
void orc_halo_exchange(orc *o, int phase) {
  long len = (long)o->p.nsubs * ((o->Q - o->U) / 2) * o->E;
  double *pos = (double *)malloc(sizeof(double) * len), *neg = (double *)malloc(sizeof(double) * len);
  orc_halo_pack(o, phase, pos, neg);
  orc_halo_unpack(o, phase, pos, neg);
  free(pos);
  free(neg);
}

void orc_step(orc *o, int write_mv) {
  int t = o->time + 1;
  int odd = t & 1;
  if (o->p.halo_mode) orc_halo_exchange(o, odd ? 1 : 2);
  orc_bcs(o, odd ? 1 : 2, t);
  run_kernel(o, odd, t, write_mv);
  o->time = t;
}

void orc_run(orc *o, int nsteps) {
  for (int q = 0; q < nsteps; q++) orc_step(o, q == nsteps - 1);
}

void orc_get_mv(const orc *o, int s, double *rho, double *ux, double *uy, double *uz) {
  long N = o->N;
  const double *b = o->mv + (long)s * N * (1 + o->D);
  if (rho) memcpy(rho, b, sizeof(double) * N);
  if (ux) memcpy(ux, b + N, sizeof(double) * N);
  if (uy) memcpy(uy, b + 2 * N, sizeof(double) * N);
  if (uz) {
    if (o->D == 3)
      memcpy(uz, b + 3 * N, sizeof(double) * N);
    else
      memset(uz, 0, sizeof(double) * N);
  }
}

/* compute_{max,min,ave}_{rho,u}: src/compute.c:1159-1416 */
void orc_reduce_ref(const orc *o, int s, double *out) {
  long N = o->N, n, nn = 0;
  const double *b = o->mv + (long)s * N * (1 + o->D);
  double mx = 0., mn, ave = 0., umax[3] = {0., 0., 0.}, umin[3], uave[3] = {0., 0., 0.};
  for (n = 0; n < N; n++)
    if (!o->sol[n + o->E]) {
      double r = fabs(b[n]);
      if (r > mx) mx = r;
    }
  mn = mx;
  for (n = 0; n < N; n++)
    if (!o->sol[n + o->E]) {
      double r = fabs(b[n]);
      if (r < mn) mn = r;
      ave += b[n];
      nn++;
    }
  for (n = 0; n < N; n++)
    if (!o->sol[n + o->E])
      for (int c = 0; c < o->D; c++) {
        double u = fabs(b[(c + 1) * N + n]);
        if (u > umax[c]) umax[c] = u;
      }
  for (int c = 0; c < 3; c++) umin[c] = umax[c];
  for (n = 0; n < N; n++)
    if (!o->sol[n + o->E])
      for (int c = 0; c < o->D; c++) {
        double u = fabs(b[(c + 1) * N + n]);
        if (u < umin[c]) umin[c] = u;
        uave[c] += b[(c + 1) * N + n];
      }
  if (nn) {
    ave /= nn;
    for (int c = 0; c < 3; c++) uave[c] /= nn;
  }
  out[0] = mx;
  out[1] = mn;
  out[2] = ave;
  for (int c = 0; c < 3; c++) {
    out[3 + c] = umax[c];
    out[6 + c] = umin[c];
    out[9 + c] = uave[c];
  }
}

void orc_reduce_sums(const orc *o, double *out) {
  long N = o->N, n;
  const double *b = o->mv;
  double m = 0., u0 = 0., u1 = 0., u2 = 0., cnt = 0., c = 0.;
  for (n = 0; n < N; n++)
    if (!is_solid(o, n)) {
      m += b[n] - o->p.rho_A[0];
      u0 += b[N + n];
      u1 += b[2 * N + n];
      if (o->D == 3) u2 += b[3 * N + n];
      cnt += 1.;
      if (o->p.nsubs == 2) c += o->mv[N * (1 + o->D) + n];
    }
  out[0] = m;
  out[1] = u0;
  out[2] = u1;
  out[3] = u2;
  out[4] = cnt;
  out[5] = c;
}
