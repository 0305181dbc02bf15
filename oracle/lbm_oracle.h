/*
 * lbm_oracle.h -- CPU oracle for the lbgpu-prime stream-collide time step.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP product path is compared against this
 * library.  Only tests/, __graft_entry__.smoke() and the CPU baseline timing
 * of bench.py may load it; the product library never links,
 * loads or falls back to it.
 *
 * It is a plain-C restatement, in FP64 with -ffp-contract=off, of the GPU
 * kernels of the reference on the reference's own memory layout
 * (cumulative-stride SoA with z/y halo planes):
 *   odd half-step   k_stream_collide_stream   src/k_stream_collide_stream.c:2-442
 *   even half-step  k_collide                 src/k_collide.c:4-434
 *   gather/scatter  get_f1d_d / set_f1d_d     src/lattice.c:2660-2744, 2765-2848
 *   relaxation      calc_f_tilde_d            src/lattice.c:2850-2924
 *   body force      apply_accel_mv            src/lattice.c:2928-2944
 *   2-D Zou-He and zero-gradient BCs          src/system_boundary_kernels.c:4-831
 *   BC host logic   bcs_1 / bcs_2             src/bcs.c:226-557, 1179-1510
 *   halo pack/unpack                          src/boundary_kernels.c:1-139
 *   host reductions compute_{max,min,ave}_*   src/compute.c:1159-1416
 * plus the parts that have no reference GPU code (3-D z-face Zou-He
 * pressure and velocity, whose top/bottom forms the reference's CPU path has
 * in src/bcs.c:592-862; solute Dirichlet / zero-gradient z faces; 2-D Zou-He
 * velocity rows; FP64 mass / velocity / plane-flux reductions), whose
 * definitions are stated in lbm_oracle.c next to the code.
 *
 * Which check pins which function is listed in the header of lbm_oracle.c.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int dims;            /* 2 (D2Q9) or 3 (D3Q19) */
  int ni, nj, nk;      /* LX, LY, LZ (nk = 1 in 2-D) */
  int nsubs;           /* 1 = flow, 2 = flow + Inamuro solute */
  int halo_mode;       /* 0: periodic wrap in every axis (serial build);
                          1: last axis through halo planes filled by a
                             periodic self-exchange (one-rank PARALLEL build) */
  double tau[2];
  double rho_A[2];
  double gaccel[6];    /* [subs*3 + component] */
  double sink;
  int sigma_t_on, sigma_t_off;

  /* reference 2-D system BCs on the N/S (y) faces */
  int pressure_n_in[2], pressure_n_out[2], pressure_s_in[2], pressure_s_out[2];
  int zeroconcgrad_n, zeroconcgrad_s;
  double rho_in, rho_out;
  double pfadd, pfmul;
  int bc_ramp_start, bc_ramp_stop;
  const double *pressure_n_series[2];
  int pressure_n_series_len[2];
  const double *pressure_s_series[2];
  int pressure_s_series_len[2];

  /* 3-D z-face BCs (no reference counterpart) */
  int zface_pressure;  /* Zou-He rho_bot at z=0 and rho_top at z=nk-1 */
  double rho_bot, rho_top;
  int zface_solute;    /* Dirichlet c at z=0, zero gradient at z=nk-1 */
  double c_bot;
  const double *c_bot_map; /* optional [nj][ni] inlet concentration map */

  /* plane (index along the last axis) whose solute moments are summed
   * during each step; -1 = none */
  int flux_plane;

  /* Zou-He velocity rows (2-D, flow): the row's y momentum rho_A u_y is
   * prescribed, u_y = uy_in for *_in and uy_out for *_out, ramped from 0 by
   * bc_ramp_start/stop like the pressure rows' density */
  int velocity_n_in, velocity_n_out, velocity_s_in, velocity_s_out;
  double uy_in, uy_out;
  /* per-face 3-D conditions, [0] z = 0, [1] z = nk-1; 0 keeps what
   * zface_pressure / zface_solute select for the face.  Flow: 1 none,
   * 2 pressure (rho_bot / rho_top), 3 velocity (zface_uz[face]: the face's
   * z momentum rho_A u_z is prescribed).  Solute: 1 none, 2 Dirichlet
   * (c_bot and c_bot_map at z = 0, c_top at z = nk-1), 3 zero gradient. */
  int zface_flow[2];
  double zface_uz[2];
  int zface_conc[2];
  double c_top;
  /* halo_mode 1, phase-2 unpack: 0 stores only where the pushing node
   * across the face is fluid; 1 stores every position (k_bound_HtD_2,
   * src/boundary_kernels.c:106-139) */
  int halo_literal;
} orc_params;

typedef struct orc orc;

orc *orc_create(const orc_params *p);
void orc_destroy(orc *o);
/* Walsh partial bounce-back (Walsh, Burwinkle & Saar 2009; the reference's
 * WALSH_NS_ON path): per node a fraction ns in [0, 1]; NULL turns it off */
int orc_set_ns(orc *o, const double *ns);
int orc_num_threads(void);
void orc_set_threads(int n);

void orc_set_solids(orc *o, const uint8_t *mask); /* N bytes, i fastest */
void orc_set_f(orc *o, int subs, const double *f1d); /* [Q][N] */
void orc_get_f(const orc *o, int subs, double *f1d);
long orc_raw_size(const orc *o);                      /* doubles per subs */
void orc_get_raw(const orc *o, int subs, double *out); /* with halos */
void orc_init_rest(orc *o, const double *c0);         /* c0: [N] or NULL */
void orc_set_time(orc *o, int t);
int orc_get_time(const orc *o);

/* One LBM step: time <- time+1; odd times run stream-collide-stream, even
 * times collide.  write_mv 1 stores the pre-collision moments (frame
 * end); 2 stores only the flow velocity and the solute concentration (the
 * PEST observation switch). */
void orc_step(orc *o, int write_mv);
void orc_run(orc *o, int nsteps);

/* pieces of a step, for tests of each part on its own */
void orc_bcs(orc *o, int phase, int time);
void orc_halo_exchange(orc *o, int phase);
void orc_halo_pack(const orc *o, int phase, double *pos, double *neg);
void orc_halo_unpack(orc *o, int phase, const double *pos, const double *neg);

void orc_get_mv(const orc *o, int subs, double *rho, double *ux, double *uy,
                double *uz);

/* compute.c reductions over non-solid nodes of the mv arrays:
 * out[0..2] = max|rho|, min|rho|, ave rho;
 * out[3..5] = max|u_x|,|u_y|,|u_z|; out[6..8] = min; out[9..11] = ave */
void orc_reduce_ref(const orc *o, int subs, double *out);

/* Build-only FP64 sums over fluid nodes of the mv arrays (subs 0 unless
 * noted): out[0] = sum (rho - rho_A), out[1..3] = sum u_eq components,
 * out[4] = number of fluid nodes, out[5] = sum c (subs 1, 0 if absent). */
void orc_reduce_sums(const orc *o, double *out);

/* Solute first moment sum_a c_a,last g_a and zeroth moment sum_a g_a summed
 * over the fluid nodes of plane p.flux_plane, on the pre-collision
 * populations of the last step (after its halo exchange and BCs). */
void orc_last_flux(const orc *o, double *flux, double *mass);

#ifdef __cplusplus
}
#endif
#endif
