/*
 * lbm_mi355x.h -- C ABI of the MI355X (gfx950) stream-collide library.
 *
 * The reference has no plugin API: its seam is inside main(), where the
 * device globals f_mem_d / mv_mem_d / solids_mem_d (src/lbgpu_prime.cu:17-24),
 * the __constant__ parameter block (src/lattice.c:2222-2268, uploaded in
 * src/latman.c:147-376) and the <<<>>> launches of the time loop
 * (src/lbgpu_prime.cu:369-915) meet the host code.  Each entry point below
 * replaces one piece of that seam; the comment names the piece.
 *
 * Contract (mirrors the reference's error convention, src/lattice.c:3123-3142):
 *   - every function returns 0 on success and nonzero on failure, with the
 *     message in lbm_last_error(); the host maps nonzero to process_exit(1);
 *   - every call is synchronous at the API boundary (one HIP stream inside);
 *   - a handle is not thread-safe.
 * Population arrays at the boundary are FP64, SoA [Q][N] with N = LX*LY*LZ
 * nodes, i fastest, in the reference's direction order (opposite directions
 * adjacent: src/lbgpu_prime.h:78-84) and hold the reference's slot contents
 * (shifted flow populations f - w rho_A; plain solute populations).
 */
#ifndef LBM_MI355X_H
#define LBM_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbm_handle lbm_handle;

/* Population storage.  FP32 stores the shifted populations f - w rho_A;
 * FP16 stores them scaled (store_scale), and the half-steps round each
 * stored FP16 value stochastically (unbiased: up or down with probabilities
 * set by its distance to the two neighbouring FP16 values, from a hash of
 * the node's own state, so results are reproducible and independent of the
 * launch order).  Collision runs in the compute type; reductions are FP64. */
enum lbm_storage { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2 };
enum lbm_compute { LBM_COMPUTE_F64 = 0, LBM_COMPUTE_F32 = 1 };
/* what a step writes to the macroscopic buffer */
enum lbm_macro { LBM_MACRO_NONE = 0, LBM_MACRO_FRAME = 1, LBM_MACRO_PEST = 2 };

/* Parameters: the fields read_params() fills for the hot path
 * (src/params.h:40-661, src/lattice.h:113-300) plus storage precision and
 * the z-face boundary conditions, which have no reference counterpart. */
typedef struct lbm_params {
  int dims;            /* NUM_DIMENSIONS: 2 (D2Q9) or 3 (D3Q19) */
  int ni, nj, nk;      /* LX, LY, LZ (nk = 1 in 2-D) */
  int nsubs;           /* NUM_FLUID_COMPONENTS: 1 flow, 2 flow + solute */
  int halo_mode;       /* 0: periodic wrap in every axis (serial build);
                          1: the last axis crosses the slab faces through
                             halo planes (PARALLEL build, one z-slab) */
  double tau[2];
  double rho_A[2];
  double gaccel[6];    /* [subs*3 + component] */
  double sink;
  int sigma_t_on, sigma_t_off;
  int pressure_n_in[2], pressure_n_out[2], pressure_s_in[2], pressure_s_out[2];
  int zeroconcgrad_n, zeroconcgrad_s;
  double rho_in, rho_out;
  double pfadd, pfmul;
  int bc_ramp_start, bc_ramp_stop;
  const double *pressure_n_series[2];
  int pressure_n_series_len[2];
  const double *pressure_s_series[2];
  int pressure_s_series_len[2];
  int zface_pressure;  /* 3-D Zou-He: rho_bot at z=0, rho_top at z=nk-1 */
  double rho_bot, rho_top;
  int zface_solute;    /* 3-D solute: Dirichlet c at z=0, zero gradient at z=nk-1 */
  double c_bot;
  const double *c_bot_map; /* optional [nj][ni] inlet concentration */
  int flux_plane;      /* plane (last axis) whose solute moments are summed on
                          steps with a macroscopic write; -1 = none */
  int storage;         /* enum lbm_storage */
  int compute;         /* enum lbm_compute */
  double store_scale[2]; /* FP16: stored = (value - shift) * scale; 0 = default
                            (2^14 flow, 2^12 solute: FP16's normal range then
                            holds |f~| < 4 and c up to about 40 with 11
                            significant bits down to 4e-9 and 1.5e-8) */
  double store_shift_c;  /* solute storage shift c0: stored g - w_a c0 (FP32 and
                            FP16 storage; 0 keeps small concentrations at full
                            relative precision) */
  /* Zou-He velocity rows in 2-D (flow only): the reference parses
   * velocity_{n,s}_{in,out} and uy_in / uy_out (src/params.h:154-155,
   * 203-206) but has no kernel for them; here the row's y momentum
   * rho_A u_y is prescribed (uy_in for *_in, uy_out for *_out, ramped from 0
   * like the pressure rows' density).  A row may not carry both a pressure
   * and a velocity condition. */
  int velocity_n_in, velocity_n_out, velocity_s_in, velocity_s_out;
  double uy_in, uy_out;
  /* Per-face 3-D z conditions, [0] the z = 0 face, [1] the z = nk-1 face
   * (the reference's CPU path: pressure_{t,b}_{in,out} and
   * velocity_{t,b}_{in,out}, src/bcs.c:592-862).  0 keeps what
   * zface_pressure / zface_solute select for that face. */
  int zface_flow[2];   /* enum lbm_zface_flow */
  double zface_uz[2];  /* u_z of a LBM_ZFACE_VELOCITY face (positive: +z) */
  int zface_conc[2];   /* enum lbm_zface_conc (solute) */
  double c_top;        /* concentration of a LBM_ZCONC_DIRICHLET z = nk-1 face
                          (the z = 0 face takes c_bot and c_bot_map) */
  /* halo_mode = 1, phase-2 unpack: 0 (default) stores a received value only
   * where the pushing node across the face is fluid, so the exchanged run
   * equals the periodic one; 1 stores every position of the received planes
   * as k_bound_HtD_2 does (src/boundary_kernels.c:106-139), which replaces
   * the own-slot bounce-back value of a face node whose cross-face upstream
   * node is solid with the received (stale) ghost value. */
  int halo_literal;
} lbm_params;

/* flow condition of a z face: pressure takes rho_bot / rho_top, velocity
 * zface_uz[face] (the face's z momentum rho_A u_z is prescribed; the
 * incompressible model's velocity is j / rho_A) */
enum lbm_zface_flow { LBM_ZFACE_DEFAULT = 0, LBM_ZFACE_NONE = 1, LBM_ZFACE_PRESSURE = 2, LBM_ZFACE_VELOCITY = 3 };
/* solute condition of a z face */
enum lbm_zface_conc { LBM_ZCONC_DEFAULT = 0, LBM_ZCONC_NONE = 1, LBM_ZCONC_DIRICHLET = 2, LBM_ZCONC_ZEROGRAD = 3 };

/* reductions over the fluid nodes of the last macroscopic write */
enum lbm_reduction {
  LBM_REDUCE_REF = 0,  /* compute.c set for one subs, 12 values:
                          max|rho|, min|rho|, ave rho, max|u_i| (3),
                          min|u_i| (3), ave u_i (3)  (src/compute.c:1159-1416) */
  LBM_REDUCE_SUMS = 1  /* 6 values: sum(rho - rho_A), sum u_eq (3), number of
                          fluid nodes, sum c (solute) */
};

const char *lbm_last_error(void);

/* construct_lattice: allocation, constants (src/latman.c:91-493).  The
 * kernels address populations with 32-bit byte offsets up to 4 GiB of
 * populations and with 64-bit offsets above; the environment variable
 * LBM_OFFSET64=1, read here, selects the 64-bit variants at every size
 * (same results, so that path can be checked on small lattices). */
int lbm_create(const lbm_params *p, lbm_handle **out);
/* destruct_lattice (src/latman.c:2084-2177) */
int lbm_destroy(lbm_handle *h);

/* solids H2D with halo copies (src/lbgpu_prime.cu:170-257).  mask: N
 * bytes, i fastest; a node is solid when its byte equals 1, as on the
 * reference GPU path (src/lattice.c:2291).  Builds the fluid-node
 * compaction and the per-row fluid bitmasks the kernels derive neighbours
 * from; populations become zero. */
int lbm_set_solids(lbm_handle *h, const uint8_t *mask);

/* Walsh partial bounce-back (the reference's WALSH_NS_ON path, off in its
 * flags: src/k_stream_collide_stream.c:129-135,363-383, src/lattice.c:
 * 2936-2939; ns_mem_d from process_ns / read_ns, src/latman.c:48-80,
 * 700-770): ns [N], a fraction in [0, 1] per node (values at solid nodes are
 * ignored).  Every fluid node then collides to (1 - ns) BGK + ns times its
 * opposite pre-collision populations, with momentum and force scaled by
 * 1 - ns.  The reference's Walsh build computes every node; here nodes whose
 * mask byte is 1 still take halfway bounce-back, so an all-zero mask gives
 * the reference's model.  NULL turns it off; lbm_set_solids clears it. */
int lbm_set_ns(lbm_handle *h, const double *ns);

/* f H2D / D2H per subs (src/lbgpu_prime.cu:317-343, :874-900).  Solid
 * nodes are not stored: download returns 0 there. */
int lbm_upload_f(lbm_handle *h, int subs, const double *f1d);
int lbm_download_f(lbm_handle *h, int subs, double *f1d);
/* lbm_download_f without any allocation (no counterpart in the reference;
 * for callers that cannot allocate device or host memory during a run, such
 * as serving loops that checkpoint the state): the same [Q][N] values, staged
 * in a static device array of the library instead of a buffer allocated per
 * call.  Q x N may not exceed 8,388,608 (64 MiB of FP64: D3Q19 lattices up to
 * 441,505 nodes, D2Q9 up to 932,067); larger lattices are rejected with an
 * error.  Every handle on one GPU shares that GPU's array, and each call
 * rewrites all of the Q x N elements it returns: calls that overlap on one
 * GPU (handles used from several threads at once) race. */
int lbm_download_f_noalloc(lbm_handle *h, int subs, double *f1d);

/* init_problem for u = 0 (src/latman.c:1786-1893): flow f~ = 0, solute
 * g = w c0 (c0: [N] or NULL for 0) */
int lbm_init_rest(lbm_handle *h, const double *c0);

/* time counter of the loop (src/lbgpu_prime.cu:369,601) */
int lbm_set_time(lbm_handle *h, int time);
int lbm_get_time(const lbm_handle *h, int *time);

/* One LBM step at time+1: odd times run the halo exchange (halo_mode),
 * bcs_1 and k_stream_collide_stream (src/lbgpu_prime.cu:382-530); even
 * times the exchange, bcs_2 and k_collide (:612-752).  write_macro is an
 * lbm_macro value: LBM_MACRO_FRAME stores the pre-collision moments
 * (is_end_of_frame_mem_c, src/k_stream_collide_stream.c:262-304);
 * LBM_MACRO_PEST stores only the flow velocity and the solute
 * concentration (pest_output_flag_c, :305-334); other nonzero values act as
 * LBM_MACRO_FRAME. */
int lbm_step(lbm_handle *h, int write_macro);
/* nsteps steps, macroscopic write on the last one if write_macro_last */
int lbm_run(lbm_handle *h, int nsteps, int write_macro_last);

/* The two half-steps on their own, at an explicit time (a1 / a2):
 * k_stream_collide_stream (src/k_stream_collide_stream.c:2-442) and
 * k_collide (src/k_collide.c:4-434), without exchange or BCs. */
int lbm_step_odd(lbm_handle *h, int time, int write_macro);
int lbm_step_even(lbm_handle *h, int time, int write_macro);

/* bcs_1 (phase 1) / bcs_2 (phase 2) at time (src/bcs.c:226-557,1179-1510);
 * in 3-D the z-face conditions */
int lbm_apply_bcs(lbm_handle *h, int phase, int time);

/* k_bound_DtH_{1,2} / k_bound_HtD_{1,2} (src/boundary_kernels.c:1-139):
 * pack the z-face planes into / unpack from DEVICE buffers laid out
 * [subs][bound pair][LX*LY] in the storage type.  Phase-2 unpack stores
 * only where the cross-face pushing node is fluid.
 * lbm_halo_bytes gives the size of one buffer. */
int lbm_halo_bytes(const lbm_handle *h, long *bytes);
int lbm_halo_pack(lbm_handle *h, int phase, void *pos_dev, void *neg_dev);
int lbm_halo_unpack(lbm_handle *h, int phase, const void *pos_dev, const void *neg_dev);
/* one-slab stand-in for process_send_recv_*: pack, then unpack the own
 * buffers as if received from both neighbours (periodic wrap) */
int lbm_halo_periodic(lbm_handle *h, int phase);

/* mv D2H (src/lbgpu_prime.cu:534-583): rho (or c for subs 1) and u_eq per
 * node; solid nodes get 0; uz may be NULL in 2-D */
int lbm_download_macro(lbm_handle *h, int subs, double *rho, double *ux, double *uy, double *uz);
/* mv H2D (src/lbgpu_prime.cu:345-352): sets the macroscopic buffer at the
 * fluid nodes from [N] host arrays; a NULL component is left as it is.
 * Steps with LBM_MACRO_PEST overwrite only part of the buffer. */
int lbm_upload_macro(lbm_handle *h, int subs, const double *rho, const double *ux, const double *uy,
                     const double *uz);

/* mv at single nodes (the per-node transfers of write_PEST_out_data,
 * src/latman.c:1130-1170): for each dense node index in nodes[npts], out
 * receives rho (c for subs 1) and the D velocity components, 1 + D values
 * per node; solid nodes give 0. */
int lbm_sample_macro(lbm_handle *h, int subs, int npts, const long *nodes, double *out);

/* FP64 reductions on the GPU over the last macroscopic write */
int lbm_reduce(lbm_handle *h, int kind, int subs, double *out);
/* solute moments of plane flux_plane on the last step with a macroscopic
 * write (or the last step of an objective run) */
int lbm_last_flux(lbm_handle *h, double *flux, double *mass);

/* Health check of the stored state, the library's form of the reference's
 * failure handling: checkCUDAError -> process_exit(1) after every launch
 * (src/lattice.c:3123-3142) and the negative-population check switched by
 * PUKE_NEGATIVE_DENSITIES ("negative densities generally signify impending
 * doom", src/flags.h:264-270).  One FP64 pass on the GPU over every stored
 * population of the fluid nodes fills *out (which may be NULL) and returns
 * nonzero, with the reason in lbm_last_error(), when
 *   - any population (flow or solute) is NaN or infinite;
 *   - any flow population f_a = f~_a + w_a rho_A or density rho_A + delta-rho
 *     is below zero;
 *   - mass_ref is given and a distribution's mass has drifted from it by more
 *     than mass_tol relative: |mass[s] - mass_ref[s]| > mass_tol * M_s with
 *     M_0 = rho_A N_f + |mass_ref[0]| (the flow's total mass) and
 *     M_1 = |mass_ref[1]|; a NaN entry of mass_ref skips that distribution.
 * Negative solute populations are counted but not an error: advection-
 * diffusion fronts undershoot slightly.  A device error left by an earlier
 * launch is reported too. */
typedef struct lbm_health {
  double mass[2];     /* sum over the fluid nodes of delta-rho (flow) and of c (solute) */
  double min_rho;     /* least flow density rho_A + delta-rho */
  double min_pop[2];  /* least finite population: flow f_a, solute g_a */
  long nonfinite[2];  /* populations that are NaN or infinite */
  long negative[2];   /* populations below zero */
} lbm_health;
int lbm_check(lbm_handle *h, const double *mass_ref, double mass_tol, lbm_health *out);

/* geometry numbers: N, fluid nodes, population slots per direction
 * (fluid nodes and halo ghost entries rounded up to 2048-node tiles) */
int lbm_info(const lbm_handle *h, long *n_nodes, long *n_fluid, long *row_stride);

/* Kernel timing with HIP events on the library's stream.  While enabled
 * every launch is bracketed by events; lbm_profile_read returns the summed
 * milliseconds and launch counts of the LBM_PROFILE_CLASSES kernel classes
 * [0] odd half-step, [1] even half-step, [2] boundary conditions,
 * [3] halo pack/unpack, [4] plane flux, [5] adjoint step (with its
 * boundary transposes; the adjoint's forward replay counts as [0]/[1]),
 * and resets them.  ms and count hold LBM_PROFILE_CLASSES entries. */
#define LBM_PROFILE_CLASSES 6
int lbm_profile_enable(lbm_handle *h, int on);
int lbm_profile_read(lbm_handle *h, double *ms, int *count);

/* Objective of the discrete adjoint (no reference counterpart): runs
 * nsteps steps from the current state like lbm_run, keeping a copy of the
 * state every `every` steps, and returns J, the window term selected by
 * lbm_objective_select over the last `window` steps:
 *   LBM_OBJ_OUTLET_FLUX (default): 1/(window * LX * LY) * sum over the
 *       steps of sum over the fluid nodes of plane flux_plane of
 *       sum_a c_a,z g_a (the solute flux through that plane, pre-collision,
 *       after the BCs);
 *   LBM_OBJ_PERMEABILITY: the Darcy permeability rho_A nu <u_z,phys> / G
 *       averaged over the steps, nu = (tau - 1/2)/3, u_phys = u_eq -
 *       (tau - 1/2) g the flow velocity averaged over all N nodes (solids
 *       count as 0), G = (rho_bot - rho_top)/(3 (LZ - 1)) with the z-face
 *       pressure conditions, plus rho_A g_z (the driving gradient; it must
 *       not be 0).
 * D3Q19 flow + solute, FP64, halo_mode 0, sink 0. */
enum lbm_objective_kind { LBM_OBJ_OUTLET_FLUX = 0, LBM_OBJ_PERMEABILITY = 1 };
int lbm_objective_select(lbm_handle *h, int kind);
int lbm_objective_run(lbm_handle *h, int nsteps, int window, int every, double *J);
/* Observation points of the objective (no reference counterpart; the
 * reference's PEST output samples the same quantity, src/latman.c:1072-1190):
 * node[i] (dense index, a fluid node), step[i] (1..nsteps of the next
 * objective run), target[i] and weight[i] (NULL: all 1) add
 *   sum_i weight[i] * (c(node[i], step[i]) - target[i])^2
 * to J, c being the solute concentration the step writes (pre-collision,
 * after the boundary conditions).  With observations, window may be 0 (no
 * flux term).  n = 0 removes them.  They stay set for later runs. */
int lbm_objective_observations(lbm_handle *h, int n, const long *node, const int *step, const double *target,
                               const double *weight);
/* Gradient of the last objective by the discrete adjoint, replaying each
 * segment between saved states out of place and sweeping it backwards:
 * grad[0..4] = dJ/d(g_x, g_y, g_z, tau, tau_s) of the flow's body force
 * and relaxation times; grad_cmap (optional, [LY][LX]) = dJ/dc_bot_map.
 * Device memory: nsteps/every saved states (objective run) and every + 2
 * states (adjoint run), each the size of the population array; both stay
 * allocated for the next run and are released by lbm_set_solids and
 * lbm_destroy.  An objective run may be swept more than once, except when
 * it kept at least every + 1 states (lbm_objective_keep_states) and left
 * segments to replay: the sweep then replays into the kept buffers, and a
 * second sweep fails until the next objective run. */
int lbm_adjoint_run(lbm_handle *h, double *grad, double *grad_cmap);
/* Device memory the next objective runs may spend on states kept for the
 * adjoint (no reference counterpart; default 0): the run keeps the states
 * of its last whole segments of `every` steps, as many as max_bytes holds
 * (one state is the size of the population array) while the checkpoints of
 * the other segments, the adjoint state, the states of one replayed segment
 * (every + 1; the first kept buffers serve when there are that many) and
 * 1 GiB of device memory stay free, and the adjoint sweep uses them instead
 * of replaying those segments from checkpoints.  The gradient is the same
 * bit for bit.  The kept states stay allocated like the checkpoints. */
int lbm_objective_keep_states(lbm_handle *h, size_t max_bytes);
/* What the last objective run holds for its adjoint (no reference
 * counterpart): the number of its last steps whose states it kept
 * (lbm_objective_keep_states) and the number of segment checkpoints it
 * saved; both 0 before any objective run.  The counts describe the run's
 * buffers: after a sweep that replayed into the kept buffers they no longer
 * hold the run's states (lbm_adjoint_run then refuses a second sweep). */
int lbm_objective_info(const lbm_handle *h, int *kept_steps, int *checkpoints);

/* Device-to-device copy of nbytes (reps times) on a stream of its own:
 * the measured bandwidth in GB/s (bytes read + written per second) is the
 * practical ceiling that the half-steps are compared with. */
int lbm_copy_bandwidth(long nbytes, int reps, double *gbps);

/* Seeded synthetic porous media (xoshiro256** seeded by splitmix64).
 * 2-D: non-overlapping discs, radius U[rmin,rmax], periodic in x, solid
 * rows j=0 and j=nj-1 when walls; discs may not touch the walls.
 * 3-D: overlapping spheres, periodic in x and y, no solid within zbuf
 * planes of either z end.  Filled until the porosity is <= phi.
 * mask gets 1 for solid, 0 for fluid.  Returns the achieved porosity in
 * *phi_out. */
int lbm_geom_discs2d(int ni, int nj, uint64_t seed, double rmin, double rmax, double phi,
                     int walls, uint8_t *mask, double *phi_out);
int lbm_geom_spheres3d(int ni, int nj, int nk, uint64_t seed, double rmin, double rmax,
                       double phi, int zbuf, uint8_t *mask, double *phi_out);

#ifdef __cplusplus
}
#endif
#endif
