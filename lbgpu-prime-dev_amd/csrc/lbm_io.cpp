// Host-side input-deck readers and per-frame writers in the reference's file
// formats (declared in include/lbm_io.h).  Plain C stdio, no GPU calls.
#include <ctype.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "lbm_internal.h"
#include "lbm_io.h"

namespace {

// One positional entry of params.in.  `both` entries are stored for every
// build; `two` entries only when the build has two components (the solute
// component and its boundary flags), otherwise scanned and dropped, as the
// reference's NUM_FLUID_COMPONENTS / INAMURO_SIGMA_COMPONENT branches do.
// `ws` entries use a format with a trailing space, which also swallows the
// whitespace after the value.
enum Kind : char { kInt = 'i', kReal = 'r' };
enum When : char { kBoth = 'b', kTwo = '2' };

struct Entry {
  const char *label;
  Kind kind;
  When when;
  bool ws;
  size_t off;
};

#define EI(lbl, fld) {lbl, kInt, kBoth, false, offsetof(lbm_deck, fld)}
#define EIW(lbl, fld) {lbl, kInt, kBoth, true, offsetof(lbm_deck, fld)}
#define ER(lbl, fld) {lbl, kReal, kBoth, false, offsetof(lbm_deck, fld)}
#define EI2(lbl, fld) {lbl, kInt, kTwo, false, offsetof(lbm_deck, fld)}
#define EIW2(lbl, fld) {lbl, kInt, kTwo, true, offsetof(lbm_deck, fld)}
#define ER2(lbl, fld) {lbl, kReal, kTwo, false, offsetof(lbm_deck, fld)}

// Order and scan formats of read_params (src/params.h:65-375).
const Entry kDeck[] = {
    EI("LX", LX), EI("LY", LY), EI("LZ", LZ), EI("BX", BX), EI("BY", BY), EI("BZ", BZ),
    EI("characteristic_length", length_scale), EI("NumFrames", NumFrames), EI("FrameRate", FrameRate),
    ER("tau[0]", tau[0]), ER("gforce_x[0]", gforce[0][0]), ER("gforce_y[0]", gforce[0][1]),
    ER("gforce_z[0]", gforce[0][2]), EI("end_grav[0]", end_grav[0]),
    ER2("tau[1]", tau[1]), ER2("gforce_x[1]", gforce[1][0]), ER2("gforce_y[1]", gforce[1][1]),
    ER2("gforce_z[1]", gforce[1][2]), EI2("end_grav[1]", end_grav[1]),
    EIW("buoyancy", buoyancy), EIW("incompressible", incompressible), EIW("simple_diffusion", simple_diffusion),
    ER("rho_A", rho_A[0]), ER("rho_B", rho_B[0]),
    ER2("rho_sigma", rho_sigma), ER2("rho_sigma_in", rho_sigma_in), ER2("rho_sigma_out", rho_sigma_out),
    ER2("u_sigma", u_sigma), ER2("u_sigma_in", u_sigma_in), ER2("u_sigma_out", u_sigma_out),
    EIW2("sigma_t_on", sigma_t_on), EIW2("sigma_t_off", sigma_t_off), EIW2("sigma_start", sigma_start),
    EIW2("sigma_stop", sigma_stop), EIW2("sigma_btc_rate", sigma_btc_rate), EIW2("sigma_btc_spot", sigma_btc_spot),
    EIW("GZL", GZL), EIW("PressureBC", PressureBC), EIW("AllBoundaryPeriodic", AllBoundaryPeriodic),
    ER("sink", sink), ER("pfmul", pfmul), ER("pfadd", pfadd), ER("rho_in", rho_in), ER("rho_out", rho_out),
    ER("ux_in", ux_in), ER("ux_out", ux_out), ER("uy_in", uy_in), ER("uy_out", uy_out), ER("uz_in", uz_in),
    ER("uz_out", uz_out), ER("big_V0", big_V0), ER("big_V0_solid[0]", big_V0_solid[0]),
    ER2("big_V0_solid[1]", big_V0_solid[1]),
    EI("ns_flag", ns_flag), ER("ns", ns), EI("ic_poisseuille", ic_poisseuille),
    EI("bc_poisseuille", bc_poisseuille), EI("bc_slip_north", bc_slip_north), EI2("bc_sigma_slip", bc_sigma_slip),
    EI("bc_ramp_start", bc_ramp_start), EI("bc_ramp_stop", bc_ramp_stop),
#define FACE_FLAGS(S, W)                                                                                 \
  W("pressure_t_in[" #S "]", pressure_t_in[S]), W("pressure_b_in[" #S "]", pressure_b_in[S]),            \
      W("pressure_t_out[" #S "]", pressure_t_out[S]), W("pressure_b_out[" #S "]", pressure_b_out[S]),    \
      W("velocity_t_in[" #S "]", velocity_t_in[S]), W("velocity_b_in[" #S "]", velocity_b_in[S]),        \
      W("velocity_t_out[" #S "]", velocity_t_out[S]), W("velocity_b_out[" #S "]", velocity_b_out[S]),    \
      W("pressure_n_in[" #S "]", pressure_n_in[S]), W("pressure_s_in[" #S "]", pressure_s_in[S]),        \
      W("pressure_n_out[" #S "]", pressure_n_out[S]), W("pressure_s_out[" #S "]", pressure_s_out[S]),    \
      W("velocity_n_in[" #S "]", velocity_n_in[S]), W("velocity_s_in[" #S "]", velocity_s_in[S]),        \
      W("velocity_n_out[" #S "]", velocity_n_out[S]), W("velocity_s_out[" #S "]", velocity_s_out[S]),    \
      W("pressure_e_in[" #S "]", pressure_e_in[S]), W("pressure_w_in[" #S "]", pressure_w_in[S]),        \
      W("pressure_e_out[" #S "]", pressure_e_out[S]), W("pressure_w_out[" #S "]", pressure_w_out[S]),    \
      W("velocity_e_in[" #S "]", velocity_e_in[S]), W("velocity_w_in[" #S "]", velocity_w_in[S]),        \
      W("velocity_e_out[" #S "]", velocity_e_out[S]), W("velocity_w_out[" #S "]", velocity_w_out[S])
    FACE_FLAGS(0, EI), FACE_FLAGS(1, EI2),
#undef FACE_FLAGS
#define SOLUTE_FLAGS(F) \
  EI2(#F "_t_in", F##_t_in), EI2(#F "_b_in", F##_b_in), EI2(#F "_t_out", F##_t_out), EI2(#F "_b_out", F##_b_out)
#define SOLUTE_FLAGS_NS(F) \
  EI2(#F "_n_in", F##_n_in), EI2(#F "_s_in", F##_s_in), EI2(#F "_n_out", F##_n_out), EI2(#F "_s_out", F##_s_out)
#define SOLUTE_FLAGS_EW(F) \
  EI2(#F "_e_in", F##_e_in), EI2(#F "_w_in", F##_w_in), EI2(#F "_e_out", F##_e_out), EI2(#F "_w_out", F##_w_out)
    SOLUTE_FLAGS(constcon), SOLUTE_FLAGS(constflx), SOLUTE_FLAGS_NS(constcon), SOLUTE_FLAGS_NS(constflx),
    SOLUTE_FLAGS_EW(constcon), SOLUTE_FLAGS_EW(constflx),
#undef SOLUTE_FLAGS
#undef SOLUTE_FLAGS_NS
#undef SOLUTE_FLAGS_EW
    EI2("zeroconcgrad_t", zeroconcgrad_t), EI2("zeroconcgrad_b", zeroconcgrad_b),
    EI2("zeroconcgrad_n", zeroconcgrad_n), EI2("zeroconcgrad_s", zeroconcgrad_s),
    EI2("zeroconcgrad_e", zeroconcgrad_e), EI2("zeroconcgrad_w", zeroconcgrad_w),
    EI2("zeroconcgrad_full", zeroconcgrad_full),
    EI("plot_scale_dynamic", plot_scale_dynamic), EI("use_colormap", use_colormap),
    EI("initial_condition", initial_condition),
    EI("x0", x0), EI("y0", y0), EI("z0", z0), EI("r0", r0), ER("cut", cut),
    EI("x1", x1), EI("x2", x2), EI("y1", y1), EI("y2", y2), EI("z1", z1), EI("z2", z2),
    ER("rel_x1", rel_x1), ER("rel_x2", rel_x2), ER("rel_y1", rel_y1), ER("rel_y2", rel_y2),
    ER("rel_z1", rel_z1), ER("rel_z2", rel_z2),
    EI("dump_rho", dump_rho), EI("dump_u", dump_u), EI("dump_force", dump_force), EI("dump_vor", dump_vor),
};

#undef EI
#undef EIW
#undef ER
#undef EI2
#undef EIW2
#undef ER2

static_assert(sizeof(kDeck) / sizeof(kDeck[0]) == LBM_DECK_ENTRIES, "params.in has 164 positional entries");

struct File {
  FILE *f;
  explicit File(FILE *p) : f(p) {}
  ~File() {
    if (f) fclose(f);
  }
};

// skip_label (src/params.h:13-31): an entry that starts like a number has
// no label; otherwise everything up to the first space is the label, and
// the spaces after it are skipped.
int skip_label(FILE *in) {
  int c = fgetc(in);
  if (c == EOF) return 1;
  if ((c >= '0' && c <= '9') || c == '-' || c == '.') {
    ungetc(c, in);
    return 0;
  }
  while ((c = fgetc(in)) != ' ')
    if (c == EOF) return 1;
  while ((c = fgetc(in)) == ' ') {
  }
  if (c != EOF) ungetc(c, in);
  return 0;
}

// the floor-then-byte conversion of write_raw: (unsigned char)ROUND(x)
// with ROUND = floor, compiled for x86-64 (cvttsd2si to a 32-bit integer,
// then the low byte; out-of-range and NaN give 0x80000000)
inline uint8_t byte_of(double x) {
  double fl = floor(x);
  int32_t i;
  if (!(fl >= -2147483648.0 && fl < 2147483648.0))
    i = INT32_MIN;
  else
    i = (int32_t)fl;
  return (uint8_t)(i & 0xff);
}

int count_items(FILE *in, const char *fmt, const char *path, int *n);

}  // namespace

extern "C" int lbm_io_read_deck(const char *path, int dims, int nsubs, lbm_deck *d) {
  if (!path || !d) return lbm_fail("null argument");
  if (dims != 2 && dims != 3) return lbm_fail("dims must be 2 or 3");
  if (nsubs != 1 && nsubs != 2) return lbm_fail("nsubs must be 1 or 2");
  File in(fopen(path, "r"));
  if (!in.f) return lbm_fail("fopen(\"%s\",\"r\") = NULL", path);
  memset(d, 0, sizeof *d);
  for (const Entry &e : kDeck) {
    if (skip_label(in.f)) return lbm_fail("%s: end of file before entry \"%s\"", path, e.label);
    const bool keep = e.when == kBoth || nsubs == 2;
    if (e.kind == kInt) {
      int v;
      if (fscanf(in.f, e.ws ? "%d " : "%d", &v) == 1 && keep) memcpy((char *)d + e.off, &v, sizeof v);
    } else {
      double v;
      if (fscanf(in.f, e.ws ? "%lf " : "%lf", &v) == 1 && keep) memcpy((char *)d + e.off, &v, sizeof v);
    }
  }
  // derived values (src/params.h:380-650)
  if (dims == 2 && d->LZ != 1) d->LZ = 1;
  d->NumNodes = (long)d->LX * d->LY * d->LZ;
  d->NumTimeSteps = d->NumFrames * d->FrameRate;
  if (nsubs == 2) {
    if (d->sigma_t_on < 0) d->sigma_t_on = 0;
    if (d->sigma_t_off < 0) d->sigma_t_off = d->NumTimeSteps;
    d->rho_A[1] = d->rho_B[0];
    d->rho_B[1] = d->rho_A[0];
  }
  if (d->x0 < 0) d->x0 = d->LX / 2;
  if (d->y0 < 0) d->y0 = d->LY / 2;
  if (d->z0 < 0) d->z0 = d->LZ / 2;
  if (d->r0 < 0) {
    int m = d->LX < d->LY ? d->LX : d->LY;
    if (d->LZ <= m) m = d->LZ;
    d->r0 = m;
  }
  if (d->cut < 0.) d->cut = 0.3;
  auto bound = [](int &v, double rel, int L, int div) {
    if (v < 0) v = rel < 0. ? L / div : (int)llround(rel * (double)L);
  };
  bound(d->x1, d->rel_x1, d->LX, 4);
  bound(d->x2, d->rel_x2, d->LX, 2);
  bound(d->y1, d->rel_y1, d->LY, 4);
  bound(d->y2, d->rel_y2, d->LY, 2);
  bound(d->z1, d->rel_z1, d->LZ, 4);
  bound(d->z2, d->rel_z2, d->LZ, 2);
  return 0;
}

extern "C" size_t lbm_io_deck_bytes(void) { return sizeof(lbm_deck); }

extern "C" int lbm_io_deck_params(const lbm_deck *d, int dims, int nsubs, lbm_params *p) {
  if (!d || !p) return lbm_fail("null argument");
  if (dims != 2 && dims != 3) return lbm_fail("dims must be 2 or 3");
  if (nsubs != 1 && nsubs != 2) return lbm_fail("nsubs must be 1 or 2");
  memset(p, 0, sizeof *p);
  p->dims = dims;
  p->ni = d->LX;
  p->nj = d->LY;
  p->nk = dims == 3 ? d->LZ : 1;
  p->nsubs = nsubs;
  p->halo_mode = 0;
  for (int s = 0; s < 2; s++) {
    p->tau[s] = s < nsubs ? d->tau[s] : 1.0;
    p->rho_A[s] = d->rho_A[s];
    for (int c = 0; c < 3; c++) p->gaccel[3 * s + c] = s < nsubs ? d->gforce[s][c] : 0.0;
  }
  p->sink = d->sink;
  p->sigma_t_on = d->sigma_t_on;
  p->sigma_t_off = d->sigma_t_off;
  p->zeroconcgrad_n = nsubs == 2 ? d->zeroconcgrad_n : 0;
  p->zeroconcgrad_s = nsubs == 2 ? d->zeroconcgrad_s : 0;
  p->rho_in = d->rho_in;
  p->rho_out = d->rho_out;
  p->pfadd = d->pfadd;
  p->pfmul = d->pfmul;
  p->bc_ramp_start = d->bc_ramp_start;
  p->bc_ramp_stop = d->bc_ramp_stop;
  p->flux_plane = -1;
  p->storage = LBM_STORE_F64;
  p->compute = LBM_COMPUTE_F64;

  // flags no kernel of this build applies
  struct Flag {
    const char *name;
    int v;
  };
  std::vector<Flag> bad;
  auto no = [&](const char *name, int v) {
    if (v) bad.push_back({name, v});
  };
  for (int s = 0; s < 2; s++) {
    const bool two = s == 1;
    auto nm = [&](const char *a, const char *b) { return two ? b : a; };
    no(nm("pressure_e_in[0]", "pressure_e_in[1]"), d->pressure_e_in[s]);
    no(nm("pressure_w_in[0]", "pressure_w_in[1]"), d->pressure_w_in[s]);
    no(nm("pressure_e_out[0]", "pressure_e_out[1]"), d->pressure_e_out[s]);
    no(nm("pressure_w_out[0]", "pressure_w_out[1]"), d->pressure_w_out[s]);
    no(nm("velocity_e_in[0]", "velocity_e_in[1]"), d->velocity_e_in[s]);
    no(nm("velocity_w_in[0]", "velocity_w_in[1]"), d->velocity_w_in[s]);
    no(nm("velocity_e_out[0]", "velocity_e_out[1]"), d->velocity_e_out[s]);
    no(nm("velocity_w_out[0]", "velocity_w_out[1]"), d->velocity_w_out[s]);
    if (dims == 2) {
      no(nm("pressure_t_in[0]", "pressure_t_in[1]"), d->pressure_t_in[s]);
      no(nm("pressure_b_in[0]", "pressure_b_in[1]"), d->pressure_b_in[s]);
      no(nm("pressure_t_out[0]", "pressure_t_out[1]"), d->pressure_t_out[s]);
      no(nm("pressure_b_out[0]", "pressure_b_out[1]"), d->pressure_b_out[s]);
      no(nm("velocity_t_in[0]", "velocity_t_in[1]"), d->velocity_t_in[s]);
      no(nm("velocity_b_in[0]", "velocity_b_in[1]"), d->velocity_b_in[s]);
      no(nm("velocity_t_out[0]", "velocity_t_out[1]"), d->velocity_t_out[s]);
      no(nm("velocity_b_out[0]", "velocity_b_out[1]"), d->velocity_b_out[s]);
    } else {
      no(nm("pressure_n_in[0]", "pressure_n_in[1]"), d->pressure_n_in[s]);
      no(nm("pressure_s_in[0]", "pressure_s_in[1]"), d->pressure_s_in[s]);
      no(nm("pressure_n_out[0]", "pressure_n_out[1]"), d->pressure_n_out[s]);
      no(nm("pressure_s_out[0]", "pressure_s_out[1]"), d->pressure_s_out[s]);
    }
    if (dims == 3 || two) {
      no(nm("velocity_n_in[0]", "velocity_n_in[1]"), d->velocity_n_in[s]);
      no(nm("velocity_s_in[0]", "velocity_s_in[1]"), d->velocity_s_in[s]);
      no(nm("velocity_n_out[0]", "velocity_n_out[1]"), d->velocity_n_out[s]);
      no(nm("velocity_s_out[0]", "velocity_s_out[1]"), d->velocity_s_out[s]);
    }
    if (dims == 3 && two) {
      no("pressure_t_in[1]", d->pressure_t_in[1]);
      no("pressure_b_in[1]", d->pressure_b_in[1]);
      no("pressure_t_out[1]", d->pressure_t_out[1]);
      no("pressure_b_out[1]", d->pressure_b_out[1]);
      no("velocity_t_in[1]", d->velocity_t_in[1]);
      no("velocity_b_in[1]", d->velocity_b_in[1]);
      no("velocity_t_out[1]", d->velocity_t_out[1]);
      no("velocity_b_out[1]", d->velocity_b_out[1]);
    }
  }
  no("constflx_t_in", d->constflx_t_in);
  no("constflx_b_in", d->constflx_b_in);
  no("constflx_t_out", d->constflx_t_out);
  no("constflx_b_out", d->constflx_b_out);
  no("constflx_n_in", d->constflx_n_in);
  no("constflx_s_in", d->constflx_s_in);
  no("constflx_n_out", d->constflx_n_out);
  no("constflx_s_out", d->constflx_s_out);
  no("constflx_e_in", d->constflx_e_in);
  no("constflx_w_in", d->constflx_w_in);
  no("constflx_e_out", d->constflx_e_out);
  no("constflx_w_out", d->constflx_w_out);
  no("constcon_n_in", d->constcon_n_in);
  no("constcon_s_in", d->constcon_s_in);
  no("constcon_n_out", d->constcon_n_out);
  no("constcon_s_out", d->constcon_s_out);
  no("constcon_e_in", d->constcon_e_in);
  no("constcon_w_in", d->constcon_w_in);
  no("constcon_e_out", d->constcon_e_out);
  no("constcon_w_out", d->constcon_w_out);
  no("zeroconcgrad_e", d->zeroconcgrad_e);
  no("zeroconcgrad_w", d->zeroconcgrad_w);
  no("zeroconcgrad_full", d->zeroconcgrad_full);
  if (dims == 2) {
    no("constcon_t_in", d->constcon_t_in);
    no("constcon_b_in", d->constcon_b_in);
    no("constcon_t_out", d->constcon_t_out);
    no("constcon_b_out", d->constcon_b_out);
    no("zeroconcgrad_t", d->zeroconcgrad_t);
    no("zeroconcgrad_b", d->zeroconcgrad_b);
  } else {
    no("zeroconcgrad_n", d->zeroconcgrad_n);
    no("zeroconcgrad_s", d->zeroconcgrad_s);
  }
  no("GZL", d->GZL);
  no("bc_poisseuille", d->bc_poisseuille);
  no("bc_slip_north", d->bc_slip_north);
  no("bc_sigma_slip", d->bc_sigma_slip);
  if (!bad.empty())
    return lbm_fail("%s = %d is set, and this build has no such boundary condition", bad[0].name, bad[0].v);

  if (dims == 2) {
    p->velocity_n_in = d->velocity_n_in[0];
    p->velocity_n_out = d->velocity_n_out[0];
    p->velocity_s_in = d->velocity_s_in[0];
    p->velocity_s_out = d->velocity_s_out[0];
    p->uy_in = d->uy_in;
    p->uy_out = d->uy_out;
    for (int s = 0; s < 2; s++) {
      p->pressure_n_in[s] = d->pressure_n_in[s];
      p->pressure_n_out[s] = d->pressure_n_out[s];
      p->pressure_s_in[s] = d->pressure_s_in[s];
      p->pressure_s_out[s] = d->pressure_s_out[s];
    }
    if ((p->velocity_n_in || p->velocity_n_out) && (p->pressure_n_in[0] || p->pressure_n_out[0]))
      return lbm_fail("the north row has both a pressure and a velocity condition");
    if ((p->velocity_s_in || p->velocity_s_out) && (p->pressure_s_in[0] || p->pressure_s_out[0]))
      return lbm_fail("the south row has both a pressure and a velocity condition");
  } else {
    for (int s = 0; s < 2; s++) {
      p->pressure_n_in[s] = p->pressure_n_out[s] = p->pressure_s_in[s] = p->pressure_s_out[s] = 0;
    }
    p->zeroconcgrad_n = p->zeroconcgrad_s = 0;
    // z faces: [0] bottom (z = 0), [1] top (z = LZ-1)
    const int pin[2] = {d->pressure_b_in[0], d->pressure_t_in[0]};
    const int pout[2] = {d->pressure_b_out[0], d->pressure_t_out[0]};
    const int vin[2] = {d->velocity_b_in[0], d->velocity_t_in[0]};
    const int vout[2] = {d->velocity_b_out[0], d->velocity_t_out[0]};
    const int cin[2] = {d->constcon_b_in, d->constcon_t_in};
    const int cout[2] = {d->constcon_b_out, d->constcon_t_out};
    const int zg[2] = {d->zeroconcgrad_b, d->zeroconcgrad_t};
    static const char *face[2] = {"bottom (b)", "top (t)"};
    for (int f = 0; f < 2; f++) {
      if ((pin[f] != 0) + (pout[f] != 0) + (vin[f] != 0) + (vout[f] != 0) > 1)
        return lbm_fail("the %s face has more than one flow condition", face[f]);
      if ((cin[f] != 0) + (cout[f] != 0) + (zg[f] != 0) > 1)
        return lbm_fail("the %s face has more than one solute condition", face[f]);
      const double rho_b = pin[f] ? d->rho_in : d->rho_out;
      if (pin[f] || pout[f]) {
        p->zface_flow[f] = LBM_ZFACE_PRESSURE;
        if (f) p->rho_top = rho_b;
        else p->rho_bot = rho_b;
      } else if (vin[f] || vout[f]) {
        p->zface_flow[f] = LBM_ZFACE_VELOCITY;
        p->zface_uz[f] = vin[f] ? d->uz_in : d->uz_out;
      } else {
        p->zface_flow[f] = LBM_ZFACE_NONE;
      }
      if (nsubs == 2 && (cin[f] || cout[f])) {
        p->zface_conc[f] = LBM_ZCONC_DIRICHLET;
        const double c = cin[f] ? d->rho_sigma_in : d->rho_sigma_out;
        if (f) p->c_top = c;
        else p->c_bot = c;
      } else if (nsubs == 2 && zg[f]) {
        p->zface_conc[f] = LBM_ZCONC_ZEROGRAD;
      } else {
        p->zface_conc[f] = LBM_ZCONC_NONE;
      }
    }
  }
  return 0;
}

extern "C" int lbm_io_read_series(const char *path, double *values, int cap, int *count) {
  if (!path || !count) return lbm_fail("null argument");
  File in(fopen(path, "r"));
  if (!in.f) return lbm_fail("can't load \"%s\"", path);
  // count as src/latman.c:552-558 does: one scan ahead, count while not at
  // end of file
  int n = 0;
  if (count_items(in.f, "%lf", path, &n)) return 1;
  *count = n;
  if (!values) return 0;
  rewind(in.f);
  for (int q = 0; q < n && q < cap; q++)
    if (fscanf(in.f, "%lf", values + q) != 1) return lbm_fail("%s: value %d unreadable", path, q);
  return 0;
}

extern "C" int lbm_io_read_solids_raw(const char *path, int LX, int LY, int LZ, uint8_t *mask) {
  if (!path || !mask) return lbm_fail("null argument");
  if (LX <= 0 || LY <= 0 || LZ <= 0) return lbm_fail("bad lattice size");
  File in(fopen(path, "rb"));
  if (!in.f) return lbm_fail("Can't open file \"%s\"", path);
  const size_t n = (size_t)LX * LY * LZ;
  size_t got = fread(mask, 1, n, in.f);
  if (got != n) return lbm_fail("%s: can't read image data: read = %zu of %zu", path, got, n);
  return 0;
}

extern "C" int lbm_io_read_solids_bmp(const char *path, int LX, int LY, uint8_t *mask) {
  if (!path || !mask) return lbm_fail("null argument");
  if (LX <= 0 || LY <= 0) return lbm_fail("bad lattice size");
  File in(fopen(path, "rb"));
  if (!in.f) return lbm_fail("File \"%s\" cannot be opened for reading", path);
  unsigned char hdr[54];
  if (fread(hdr, 1, 54, in.f) != 54) return lbm_fail("%s: short BMP header", path);
  if (hdr[0] != 'B' || hdr[1] != 'M') return lbm_fail("%s: can't process this file type", path);
  auto le32 = [&](int o) { return (int32_t)((uint32_t)hdr[o] | (uint32_t)hdr[o + 1] << 8 | (uint32_t)hdr[o + 2] << 16 | (uint32_t)hdr[o + 3] << 24); };
  const int32_t width = le32(18), height = le32(22), compression = le32(30);
  const int bits = (int16_t)((uint16_t)hdr[28] | (uint16_t)hdr[29] << 8);
  if (compression != 0) return lbm_fail("%s: can't handle compression", path);
  if (width != LX) return lbm_fail("%s: LX %d does not match the width %d of the BMP file", path, LX, width);
  if (height != LY) return lbm_fail("%s: LY %d does not match the height %d of the BMP file", path, LY, height);
  if (bits != 24) return lbm_fail("%s: unhandled color depth, BitCount = %d", path, bits);
  // pixel rows follow the 54-byte header directly (the reference does not
  // use bfOffBits); each row holds 3*LX bytes plus padding to 4
  const long bpr = (long)ceil((double)width * (double)bits / 8.);
  const long pad = (4 - bpr % 4) % 4, row = bpr + pad;
  std::vector<unsigned char> buf((size_t)row);
  memset(mask, 0, (size_t)LX * LY);
  for (int j = 0; j < LY; j++) {
    size_t got = fread(buf.data(), 1, (size_t)row, in.f);
    // a short last row keeps the pixels it holds, as the byte loop there does
    for (int i = 0; i < LX && 3L * i + 2 < (long)got; i++) {
      const unsigned char *px = &buf[3 * (size_t)i];
      if (px[0] == 0 && px[1] == 0 && px[2] == 0) mask[(size_t)j * LX + i] = 1;
    }
    if (got < (size_t)row) break;
  }
  return 0;
}

extern "C" int lbm_io_filename(char *out, size_t cap, const char *out_path, int dims, int LX, int LY, int nsubs,
                               int num_procs, int proc_id, const char *prefix, int frame, int subs,
                               const char *suffix) {
  if (!out || !out_path || !prefix || !suffix) return lbm_fail("null argument");
  char size_str[64], frame_str[32], subs_str[32];
  if (dims == 2)
    snprintf(size_str, sizeof size_str, "%dx%d", LX, LY);
  else
    snprintf(size_str, sizeof size_str, "%dx%dx%d", LX, LY, LY);
  if (frame >= 0)
    snprintf(frame_str, sizeof frame_str, "_frame%04d", frame);
  else
    frame_str[0] = '\0';
  if (nsubs == 1)
    subs_str[0] = '\0';
  else
    snprintf(subs_str, sizeof subs_str, "_subs%02d", subs);
  int n;
  if (num_procs > 1)
    n = snprintf(out, cap, "%s/cuda_%s%s%s%s_proc%04d%s", out_path, prefix, size_str, frame_str, subs_str, proc_id,
                 suffix);
  else
    n = snprintf(out, cap, "%s/cuda_%s%s%s%s%s", out_path, prefix, size_str, frame_str, subs_str, suffix);
  if (n < 0 || (size_t)n >= cap) return lbm_fail("file name longer than %zu bytes", cap);
  return 0;
}

extern "C" int lbm_io_frame_stats(const double *rho, const double *ux, const double *uy, const double *uz,
                                  const uint8_t *mask, long N, double st[12]) {
  if (!rho || !ux || !uy || !mask || !st) return lbm_fail("null argument");
  auto z = [&](long n) { return uz ? uz[n] : 0.0; };
  const double *u[3] = {ux, uy, uz};
  auto uc = [&](int c, long n) { return c < 2 ? u[c][n] : z(n); };
  // compute_max_rho / compute_min_rho / compute_ave_rho (src/compute.c:1159-1274)
  double mx = 0., mn, av = 0.;
  long nn = 0;
  for (long n = 0; n < N; n++)
    if (!mask[n]) {
      double r = fabs(rho[n]);
      if (r > mx) mx = r;
    }
  mn = mx;
  for (long n = 0; n < N; n++)
    if (!mask[n]) {
      double r = fabs(rho[n]);
      if (r < mn) mn = r;
    }
  for (long n = 0; n < N; n++)
    if (!mask[n]) {
      av += rho[n];
      nn++;
    }
  if (nn != 0) av = av / nn;
  st[0] = mx;
  st[1] = mn;
  st[2] = av;
  // compute_max_u / compute_min_u / compute_ave_u (:1276-1416)
  double umax[3] = {0., 0., 0.}, umin[3], uav[3] = {0., 0., 0.};
  for (long n = 0; n < N; n++)
    if (!mask[n])
      for (int c = 0; c < 3; c++) {
        double a = fabs(uc(c, n));
        if (a > umax[c]) umax[c] = a;
      }
  for (int c = 0; c < 3; c++) umin[c] = umax[c];
  for (long n = 0; n < N; n++)
    if (!mask[n])
      for (int c = 0; c < 3; c++) {
        double a = fabs(uc(c, n));
        if (a < umin[c]) umin[c] = a;
      }
  nn = 0;
  for (long n = 0; n < N; n++)
    if (!mask[n]) {
      for (int c = 0; c < 3; c++) uav[c] += uc(c, n);
      nn++;
    }
  if (nn != 0)
    for (int c = 0; c < 3; c++) uav[c] = uav[c] / nn;
  for (int c = 0; c < 3; c++) {
    st[3 + c] = umax[c];
    st[6 + c] = umin[c];
    st[9 + c] = uav[c];
  }
  return 0;
}

extern "C" int lbm_io_frames_dat(const char *path, int time, const double st[12], const double flux[4]) {
  if (!path || !st || !flux) return lbm_fail("null argument");
  if (time == 0) {
    File o(fopen(path, "w+"));
    if (!o.f) return lbm_fail("fopen(\"%s\",\"w+\") = NULL", path);
    static const char *cols[] = {"time",     "|j|",      "j_x",      "j_y",      "j_z",      "min_u[x]",
                                 "min_u[y]", "min_u[z]", "max_u[x]", "max_u[y]", "max_u[z]", "ave_u[x]",
                                 "ave_u[y]", "ave_u[z]", "max/ave[x]", "max/ave[y]", "max/ave[z]", "min_rho",
                                 "max_rho",  "ave_rho",  "max/ave"};
    fprintf(o.f, "\n");
    for (const char *c : cols) fprintf(o.f, "%12s ", c);
    fprintf(o.f, "\n");
    for (size_t q = 0; q < sizeof cols / sizeof cols[0]; q++) fprintf(o.f, "------------ ");
    fprintf(o.f, "\n");
  }
  File o(fopen(path, "a+"));
  if (!o.f) return lbm_fail("fopen(\"%s\",\"a+\") = NULL", path);
  const double max_rho = st[0], min_rho = st[1], ave_rho = st[2];
  const double *max_u = st + 3, *min_u = st + 6, *ave_u = st + 9;
  double ratio[3];
  for (int c = 0; c < 3; c++) ratio[c] = ave_u[c] != 0. ? max_u[c] / ave_u[c] : 1.;
  const double rho_ratio = ave_rho != 0. ? max_rho / ave_rho : 1.;
  auto clip = [](double r) { return fabs(r) <= 9999. ? r : INFINITY; };
  // the row format of src/lbio.c:242-263, including the missing blank
  // between j_z and min_u[x]
  fprintf(o.f,
          "%12d %12.7f %12.7f %12.7f %12.7f%12.7f %12.7f %12.7f %12.7f %12.7f %12.7f %12.7f %12.7f %12.7f "
          "%12.7f %12.7f %12.7f %12.7f %12.7f %12.7f %12.7f\n",
          time, flux[0], flux[1], flux[2], flux[3], min_u[0], min_u[1], min_u[2], max_u[0], max_u[1], max_u[2],
          ave_u[0], ave_u[1], ave_u[2], clip(ratio[0]), clip(ratio[1]), clip(ratio[2]), min_rho, max_rho, ave_rho,
          clip(rho_ratio));
  return 0;
}

extern "C" int lbm_io_write_txt(const char *path, const double *a, int LX, int LY, int LZ) {
  if (!path || !a) return lbm_fail("null argument");
  File o(fopen(path, "w+"));
  if (!o.f) return lbm_fail("fopen(\"%s\",\"w+\") = NULL", path);
  long n = 0;
  for (int k = 0; k < LZ; k++) {
    for (int j = 0; j < LY; j++) {
      for (int i = 0; i < LX; i++) fprintf(o.f, " %18.16f", a[n++]);
      fprintf(o.f, "\n");
    }
    fprintf(o.f, "\n");
  }
  return 0;
}

extern "C" int lbm_io_write_dat(const char *path, const double *a, const uint8_t *mask, long N) {
  if (!path || !a || !mask) return lbm_fail("null argument");
  std::vector<double> buf((size_t)N);
  for (long n = 0; n < N; n++) buf[n] = mask[n] ? 0. : a[n];
  File o(fopen(path, "wb"));
  if (!o.f) return lbm_fail("Can't creat file \"%s\"", path);
  if (fwrite(buf.data(), sizeof(double), (size_t)N, o.f) != (size_t)N) return lbm_fail("%s: short write", path);
  return 0;
}

extern "C" int lbm_io_write_raw8(const char *path, const double *ux, const double *uy, const double *uz,
                                 const uint8_t *mask, long N, const double a_max[3], const double a_min[3]) {
  if (!path || !ux || !mask || !a_max || !a_min) return lbm_fail("null argument");
  double hi = a_max[0], lo = a_min[0];
  if (uy) {
    hi = sqrt(a_max[0] * a_max[0] + a_max[1] * a_max[1] + a_max[2] * a_max[2]);
    lo = sqrt(a_min[0] * a_min[0] + a_min[1] * a_min[1] + a_min[2] * a_min[2]);
  }
  std::vector<uint8_t> buf((size_t)N);
  for (long n = 0; n < N; n++) {
    if (mask[n]) {
      buf[n] = 0;
      continue;
    }
    double a = ux[n];
    if (uy) a = sqrt(ux[n] * ux[n] + uy[n] * uy[n] + (uz ? uz[n] * uz[n] : 0));
    buf[n] = byte_of(1. + 254. * (a - lo) / (hi - lo));
  }
  File o(fopen(path, "wb"));
  if (!o.f) return lbm_fail("Can't creat file \"%s\"", path);
  if (fwrite(buf.data(), 1, (size_t)N, o.f) != (size_t)N) return lbm_fail("%s: short write", path);
  return 0;
}

namespace {

// the feof counting idiom of the reference's list readers: one scan ahead,
// count while not at the end of the file
int count_items(FILE *in, const char *fmt, const char *path, int *n) {
  union {
    int i;
    double d;
  } v;
  int c = 0;
  (void)fscanf(in, fmt, &v);
  while (!feof(in)) {
    c++;
    if (fscanf(in, fmt, &v) != 1 && !feof(in)) return lbm_fail("%s: unreadable entry after %d values", path, c);
  }
  *n = c;
  return 0;
}

int read_ints(const char *path, int n, std::vector<int> &out) {
  File in(fopen(path, "r"));
  if (!in.f) return lbm_fail("can't load \"%s\"", path);
  out.resize(n);
  for (int q = 0; q < n; q++)
    if (fscanf(in.f, "%d", &out[q]) != 1) return lbm_fail("%s: %d values expected, %d read", path, n, q);
  return 0;
}

}  // namespace

extern "C" int lbm_io_read_observations(const char *t_path, const char *x_path, const char *y_path, int LX, int LY,
                                        int cap, int *t, int *x, int *y, int *order, int *count) {
  if (!t_path || !x_path || !y_path || !count) return lbm_fail("null argument");
  int n = 0;
  {
    File in(fopen(t_path, "r"));
    if (!in.f) return lbm_fail("can't load \"%s\"", t_path);
    if (count_items(in.f, "%d", t_path, &n)) return 1;
  }
  std::vector<int> tt, xx, yy;
  if (read_ints(t_path, n, tt)) return 1;
  for (int q = 1; q < n; q++)
    if (tt[q] < tt[q - 1]) return lbm_fail("%s: observation times not in time order (entry %d)", t_path, q);
  if (read_ints(x_path, n, xx) || read_ints(y_path, n, yy)) return 1;
  int kept = 0;
  for (int q = 0; q < n; q++) {
    if (xx[q] < 0 || xx[q] > LX - 1 || yy[q] < 0 || yy[q] > LY - 1) continue;
    if (t && kept < cap) {
      t[kept] = tt[q];
      if (x) x[kept] = xx[q];
      if (y) y[kept] = yy[q];
      if (order) order[kept] = q;
    }
    kept++;
  }
  *count = kept;
  return 0;
}

extern "C" int lbm_io_write_observations(const char *out_dir, int count, const double *conc, const double *vel,
                                         const int *order) {
  if (!out_dir || count < 0 || (count && (!conc || !vel || !order))) return lbm_fail("bad argument");
  const char *names[3] = {"conc_data", "vel_data", "conc_order"};
  for (int f = 0; f < 3; f++) {
    char path[1024];
    int len = snprintf(path, sizeof path, "%s/%s_proc%04d.dat", out_dir, names[f], 0);
    if (len < 0 || (size_t)len >= sizeof path) return lbm_fail("file name too long");
    File o(fopen(path, "w+"));
    if (!o.f) return lbm_fail("can't write \"%s\"", path);
    for (int q = 0; q < count; q++) {
      if (f == 0) fprintf(o.f, "%20.17f\n", conc[q]);
      if (f == 1) fprintf(o.f, "%20.17f\n", vel[q]);
      if (f == 2) fprintf(o.f, "%d\n", order[q]);
    }
  }
  return 0;
}
