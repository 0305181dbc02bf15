/*
 * lbm_io.h -- host-side input-deck and output formats of the reference, as a
 * C ABI over liblbm_mi355x.so.  No GPU is touched by anything declared here.
 *
 * The reference reads a strictly positional params.in, a 24-bit BMP (2-D)
 * or a byte-per-node .raw (3-D) for the solids and optional pressure time
 * series, and writes frames.dat, rho/u*.txt, .dat and 8-bit .raw files per
 * frame under ./out with names from gen_filename.  These functions do the
 * same, so a deck written for the reference runs unchanged on the MI355X
 * library (see the lbgpu_prime driver built beside the library).
 *
 * Every function returns 0 on success and nonzero on failure, with the
 * message in lbm_last_error() (declared in lbm_mi355x.h).
 */
#ifndef LBM_IO_H
#define LBM_IO_H

#include <stddef.h>
#include <stdint.h>

#include "lbm_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The fields read_params() fills (src/params.h:40-661; field meanings in
 * src/lattice.h:113-300), named as there.  `real` is double, the
 * reference's DP_ON build (src/flags.h:73); INTEGER_IC_BOUND is on
 * (src/flags.h:55), so the initial-condition bounds are integers. */
typedef struct lbm_deck {
  int LX, LY, LZ, BX, BY, BZ;
  int length_scale, NumFrames, FrameRate;
  double tau[2];
  double gforce[2][3];
  int end_grav[2];
  int buoyancy, incompressible, simple_diffusion;
  double rho_A[2], rho_B[2];
  double rho_sigma, rho_sigma_in, rho_sigma_out;
  double u_sigma, u_sigma_in, u_sigma_out;
  int sigma_t_on, sigma_t_off, sigma_start, sigma_stop, sigma_btc_rate, sigma_btc_spot;
  int GZL, PressureBC, AllBoundaryPeriodic;
  double sink, pfmul, pfadd;
  double rho_in, rho_out, ux_in, ux_out, uy_in, uy_out, uz_in, uz_out;
  double big_V0, big_V0_solid[2];
  int ns_flag;
  double ns;
  int ic_poisseuille, bc_poisseuille, bc_slip_north, bc_sigma_slip;
  int bc_ramp_start, bc_ramp_stop;
  /* [subs]: top/bottom (z), north/south (y), east/west (x) faces */
  int pressure_t_in[2], pressure_b_in[2], pressure_t_out[2], pressure_b_out[2];
  int velocity_t_in[2], velocity_b_in[2], velocity_t_out[2], velocity_b_out[2];
  int pressure_n_in[2], pressure_s_in[2], pressure_n_out[2], pressure_s_out[2];
  int velocity_n_in[2], velocity_s_in[2], velocity_n_out[2], velocity_s_out[2];
  int pressure_e_in[2], pressure_w_in[2], pressure_e_out[2], pressure_w_out[2];
  int velocity_e_in[2], velocity_w_in[2], velocity_e_out[2], velocity_w_out[2];
  int constcon_t_in, constcon_b_in, constcon_t_out, constcon_b_out;
  int constflx_t_in, constflx_b_in, constflx_t_out, constflx_b_out;
  int constcon_n_in, constcon_s_in, constcon_n_out, constcon_s_out;
  int constflx_n_in, constflx_s_in, constflx_n_out, constflx_s_out;
  int constcon_e_in, constcon_w_in, constcon_e_out, constcon_w_out;
  int constflx_e_in, constflx_w_in, constflx_e_out, constflx_w_out;
  int zeroconcgrad_t, zeroconcgrad_b, zeroconcgrad_n, zeroconcgrad_s;
  int zeroconcgrad_e, zeroconcgrad_w, zeroconcgrad_full;
  int plot_scale_dynamic, use_colormap, initial_condition;
  int x0, y0, z0, r0;
  double cut;
  int x1, x2, y1, y2, z1, z2;
  double rel_x1, rel_x2, rel_y1, rel_y2, rel_z1, rel_z2;
  int dump_rho, dump_u, dump_force, dump_vor;
  /* derived by read_params (src/params.h:380-650) */
  int NumTimeSteps;
  long NumNodes;
} lbm_deck;

/* Number of positional entries of params.in (one value each). */
#define LBM_DECK_ENTRIES 164

/* read_params (src/params.h:40-661) for a build of `dims` dimensions and
 * `nsubs` fluid components (NUM_DIMENSIONS, NUM_FLUID_COMPONENTS; the
 * solute block is read when nsubs == 2, as INAMURO_SIGMA_COMPONENT).
 * Labels are skipped up to the first space unless an entry starts with a
 * digit, '-' or '.' (skip_label, :13-31); each value is scanned with %d or
 * %lf in the reference's order, so an integer written as "1.5" or a real
 * written without a point shifts the following entries exactly as there.
 * Entries the build does not use are scanned and dropped.  The derived
 * defaults follow: LZ = 1 in 2-D, NumTimeSteps = NumFrames * FrameRate,
 * sigma_t_on/off bounds, rho_A[1] = rho_B[0] and rho_B[1] = rho_A[0] for two
 * components, centre/radius/cube defaults for negative x0..z2 and cut.
 * Fields not scanned start at zero.  A file that ends inside a label is an
 * error (the reference loops there for ever). */
int lbm_io_read_deck(const char *path, int dims, int nsubs, lbm_deck *deck);

/* The lbm_params a deck describes for the stream-collide library: tau,
 * rho_A, gforce, sink, sigma_t_on/off, rho_in/out, pfadd/pfmul, the ramp
 * and the boundary flags (storage FP64, collision FP64, halo_mode 0,
 * flux_plane -1; series pointers left NULL):
 *   2-D  pressure_{n,s}_{in,out}[s] and zeroconcgrad_{n,s} as the
 *        reference's GPU path applies them (src/bcs.c:226-557);
 *        velocity_{n,s}_{in,out}[0] become velocity rows with uy_in/uy_out;
 *   3-D  the z faces (t = z = LZ-1, b = z = 0; the reference's GPU build
 *        applies no condition in 3-D, its CPU path has pressure_t_in,
 *        pressure_b_out, velocity_t_in, velocity_b_out, src/bcs.c:592-862):
 *        pressure_{t,b}_{in,out}[0] -> a pressure face with rho_in/rho_out,
 *        velocity_{t,b}_{in,out}[0] -> a velocity face with uz_in/uz_out,
 *        and with the solute constcon_{t,b}_{in,out} -> a Dirichlet face
 *        with rho_sigma_in/rho_sigma_out, zeroconcgrad_{t,b} -> a
 *        zero-gradient face.
 * Every other boundary flag that is set is an error naming it (east/west
 * faces, constant-flux faces, zeroconcgrad_full, second-component flow
 * conditions in 3-D or velocity ones in 2-D, faces of the other dimension,
 * GZL, the slip and Poiseuille switches), as are two flow or two solute
 * conditions on one face: the reference's GPU build ignores them silently,
 * and a deck that sets them expects a condition this build would not
 * apply. */
int lbm_io_deck_params(const lbm_deck *deck, int dims, int nsubs, lbm_params *p);

/* Pressure time series ./in/pressure_{n,s}_in0.in (src/latman.c:536-631):
 * whitespace-separated reals, counted as the reference counts them (a value
 * not followed by anything before the end of the file is not counted).
 * With values == NULL only *count is returned; otherwise at most cap values
 * are stored. */
int lbm_io_read_series(const char *path, double *values, int cap, int *count);

/* read_solids (src/lbio.c:507-565): LX*LY*LZ bytes, one per node, i
 * fastest; the file must hold at least that many.  Byte values are kept
 * as they are (the kernels treat a node as solid when its byte is 1). */
int lbm_io_read_solids_raw(const char *path, int LX, int LY, int LZ, uint8_t *mask);

/* spy_bmp + process_matrix (src/lbio.c:3073-3363, src/latman.c:24-46):
 * uncompressed 24-bit BMP of width LX and height LY, pixel rows read in
 * file order (row j = 0 first) from byte 54 on, rows padded to 4 bytes;
 * a black pixel (0,0,0) is solid (1), anything else fluid (0).  Other bit
 * depths, compression or a size mismatch are errors, as there. */
int lbm_io_read_solids_bmp(const char *path, int LX, int LY, uint8_t *mask);

/* gen_filename (src/lattice.c:371-452) of the GPU build ("cuda_" prefix):
 * <out_path>/cuda_<prefix><size>[_frame%04d][_subs%02d][_proc%04d]<suffix>
 * with size "LXxLY" in 2-D and "LXxLYxLY" in 3-D (the reference prints LY
 * twice); the frame part is left out for frame < 0, the subs part for one
 * component, the proc part for one process. */
int lbm_io_filename(char *out, size_t cap, const char *out_path, int dims, int LX, int LY, int nsubs,
                    int num_procs, int proc_id, const char *prefix, int frame, int subs, const char *suffix);

/* compute_{max,min,ave}_{rho,u} (src/compute.c:1159-1416) on host arrays of
 * N nodes over the nodes whose mask byte is 0 (the host rule of
 * is_not_solid, src/lattice.c:163-194), in node order.  stats[12] =
 * max|rho|, min|rho|, ave rho, max|u_i| (3), min|u_i| (3), ave u_i (3).
 * uz may be NULL (treated as zeros). */
int lbm_io_frame_stats(const double *rho, const double *ux, const double *uy, const double *uz,
                       const uint8_t *mask, long N, double stats[12]);

/* dump_frame_summary (src/lbio.c:140-290): at time 0 the file is created
 * with its two header lines, then one row is appended with the time, the
 * four flux values, min/max/ave u, the max/ave ratios and min/max/ave rho. */
int lbm_io_frames_dat(const char *path, int time, const double stats[12], const double flux[4]);

/* write_rho_txt / write_u_txt (src/lbio.c:2864-2973): " %18.16f" per node,
 * a newline after each row of LX values and after each LX*LY plane. */
int lbm_io_write_txt(const char *path, const double *a, int LX, int LY, int LZ);

/* write_dat (src/lbio.c:2507-2561): N raw doubles, 0 at nodes whose mask
 * byte is nonzero. */
int lbm_io_write_dat(const char *path, const double *a, const uint8_t *mask, long N);

/* write_raw (src/lbio.c:934-986): one byte per node, floor(1 + 254 (a -
 * a_min) / (a_max - a_min)) converted to an 8-bit byte through a 32-bit
 * integer as x86-64 does, 0 at nodes whose mask byte is nonzero.  With
 * uy != NULL the value is the speed sqrt(ux^2 + uy^2 [+ uz^2]) and a_max /
 * a_min are formed from the component extremes (write_raw_u, :987-1060);
 * uz == NULL drops the z term (2-D). */
int lbm_io_write_raw8(const char *path, const double *ux, const double *uy, const double *uz,
                      const uint8_t *mask, long N, const double a_max[3], const double a_min[3]);

/* read_PEST_in_files (src/latman.c:889-1038): observation times from
 * t_path (integers, counted as the pressure series are), and the x and y
 * lattice coordinates of each from x_path and y_path.  Times that decrease
 * are an error ("concentration data not in time order" there).  Points
 * outside 0..LX-1 x 0..LY-1 are dropped; order[] keeps each kept point's
 * position in the files.  With t == NULL only *count (kept points) is
 * returned; otherwise at most cap points are stored. */
int lbm_io_read_observations(const char *t_path, const char *x_path, const char *y_path, int LX, int LY, int cap,
                             int *t, int *x, int *y, int *order, int *count);

/* write_PEST_out_file (src/latman.c:1196-1255) of process 0: under
 * out_dir, conc_data_proc0000.dat and vel_data_proc0000.dat ("%20.17f" per
 * point: the solute concentration and the flow speed |(u_x, u_y)| at the
 * observation) and conc_order_proc0000.dat ("%d": the point's order). */
int lbm_io_write_observations(const char *out_dir, int count, const double *conc, const double *vel,
                              const int *order);

/* main() of the reference (src/lbgpu_prime.cu:92-915) on this library:
 * argv as for the lbgpu_prime executable ([name] [--dims 2|3] [--subs 1|2]
 * [--storage f64|f32|f16] [--compute f64|f32] [--in DIR] [--out DIR]).
 * Reads <in>/params.in (./in_<name>/params.in when a name is given), the
 * solids (<in>/LXxLY.bmp in 2-D, <in>/LXxLYxLZ.raw in 3-D) and the pressure
 * series, starts from init_problem's rest state (initial_condition 0 or 2),
 * writes frame 0, steps NumFrames * FrameRate times and writes each frame
 * (frames.dat, .dat, 3-D .raw, rho/u .txt) under <out>.  With
 * <in>/timestep_file.in present, the observation points of
 * <in>/{timestep,x_coord,y_coord}_file.in are sampled on their steps (the
 * PEST switch) and written at the end (lbm_io_write_observations).
 * Returns 0, or 1 after printing the reason (process_exit(1) there). */
int lbm_deck_main(int argc, char **argv);

/* sizeof(lbm_deck), for bindings that mirror the struct */
size_t lbm_io_deck_bytes(void);

#ifdef __cplusplus
}
#endif

#endif
