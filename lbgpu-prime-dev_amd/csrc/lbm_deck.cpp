// lbm_deck_main -- the reference's main() loop (src/lbgpu_prime.cu:92-915)
// over the MI355X library: reads a params.in deck, the solids image and the
// optional pressure series, sets up the rest state of init_problem, steps
// the lattice and writes the per-frame files in the reference's formats.
// The lbgpu_prime executable is this function behind main().
//
//   lbgpu_prime [name] [--dims 2|3] [--subs 1|2] [--storage f64|f32|f16]
//               [--compute f64|f32] [--in DIR] [--out DIR] [--walsh]
//
// name selects ./in_<name>/params.in as argv[1] does in construct_lattice
// (src/latman.c:108-116); solids and series are read from ./in (or --in).
// --dims / --subs stand for the reference's NUM_DIMENSIONS and
// NUM_FLUID_COMPONENTS build flags (defaults 2 and 1, src/flags.h:104,108).
// --walsh stands for a build with WALSH_NS_ON and COMPUTE_ON_SOLIDS
// (src/flags.h:57-63): the kernels then compute every node without
// bounce-back, and the black pixels of the 2-D solids image carry the
// partial bounce-back fraction ns of params.in (process_ns,
// src/latman.c:48-80, 413-416); the host keeps the image's solids for its
// frame statistics and .dat files.
// Any failure prints the reason and returns 1, the status of
// process_exit(1); the lattice is released first.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <exception>
#include <string>
#include <vector>

#include "lbm_io.h"
#include "lbm_mi355x.h"

namespace {

struct Exit {
  std::string msg;
};

[[noreturn]] void die(const char *what) { throw Exit{std::string(what) + ": " + lbm_last_error()}; }

[[noreturn]] void die_msg(const std::string &msg) { throw Exit{msg}; }

struct HandleGuard {
  lbm_handle *h = nullptr;
  ~HandleGuard() {
    if (h) (void)lbm_destroy(h);
  }
};

void check(int rc, const char *what) {
  if (rc) die(what);
}

struct Macro {
  std::vector<double> rho, ux, uy, uz;
  explicit Macro(long N) : rho(N, 0.), ux(N, 0.), uy(N, 0.), uz(N, 0.) {}
};

struct Run {
  int dims = 2, nsubs = 1;
  std::string in_dir = "./in", out_dir = "./out", params_path;
  lbm_deck d{};
  std::vector<uint8_t> mask;
  bool walsh = false;         // a WALSH_NS_ON + COMPUTE_ON_SOLIDS run
  std::vector<Macro> mv;      // what the host holds for output
  std::vector<Macro> init_mv; // init_problem's values (kept at solid nodes)
  long N = 0;

  // a node the kernels never write: an image solid, unless every node is
  // computed
  bool kept(long n) const { return !walsh && mask[n] == 1; }

  std::string name(const char *prefix, int frame, int subs, const char *suffix) const {
    char buf[1024];
    check(lbm_io_filename(buf, sizeof buf, out_dir.c_str(), dims, d.LX, d.LY, nsubs, 1, 0, prefix, frame, subs,
                          suffix),
          "gen_filename");
    return buf;
  }

  // init_problem (src/latman.c:1270-1700): density per initial_condition
  // (0 uniform, 2 cube), zero velocity
  void init_problem() {
    mv.assign(nsubs, Macro(N));
    const int ni = d.LX, nj = d.LY;
    for (int s = 0; s < nsubs; s++) {
      for (long n = 0; n < N; n++) {
        const int i = (int)(n % ni), j = (int)((n / ni) % nj), k = (int)(n / ((long)ni * nj));
        const bool in_cube = i >= d.x1 && i <= d.x2 && j >= d.y1 && j <= d.y2 &&
                             (dims == 2 || (k >= d.z1 && k <= d.z2));
        double r;
        switch (d.initial_condition) {
          case 0:
            r = d.rho_A[s];
            break;
          case 2:
            if (nsubs == 1)
              r = in_cube ? d.rho_B[0] : d.rho_A[0];
            else
              r = s == 0 ? d.rho_A[0] : (in_cube ? d.rho_sigma : 0.);
            break;
          default:
            die_msg("init_problem: unhandled initial_condition = " + std::to_string(d.initial_condition));
        }
        mv[s].rho[n] = r;
      }
    }
    init_mv = mv;
  }

  // output_frame (src/lbio.c:37-133): frame summary, .dat (and in 3-D
  // .raw) files of each component, then the rho and u text files
  void output_frame(int time, int frame) {
    std::vector<std::vector<double>> st(nsubs, std::vector<double>(12));
    const double *uz_or_null = nullptr;
    for (int s = 0; s < nsubs; s++) {
      uz_or_null = mv[s].uz.data();
      check(lbm_io_frame_stats(mv[s].rho.data(), mv[s].ux.data(), mv[s].uy.data(), uz_or_null, mask.data(), N,
                               st[s].data()),
            "compute_*");
    }
    const double flux[4] = {0., 0., 0., 0.};  // compute_flux is a stub there
    for (int s = 0; s < nsubs; s++)
      check(lbm_io_frames_dat(name("frames", -1, s, ".dat").c_str(), time, st[s].data(), flux), "frames.dat");
    for (int s = 0; s < nsubs; s++) {
      const Macro &m = mv[s];
      const double *S = st[s].data();
      if (dims == 3) {
        const double mx1[3] = {S[0], 0, 0}, mn1[3] = {S[1], 0, 0};
        check(lbm_io_write_raw8(name("rho", frame, s, ".raw").c_str(), m.rho.data(), nullptr, nullptr, mask.data(),
                                N, mx1, mn1),
              "write_raw rho");
        check(lbm_io_write_raw8(name("u", frame, s, ".raw").c_str(), m.ux.data(), m.uy.data(), m.uz.data(),
                                mask.data(), N, S + 3, S + 6),
              "write_raw u");
        const char *pre[3] = {"u_x", "u_y", "u_z"};
        const std::vector<double> *comp[3] = {&m.ux, &m.uy, &m.uz};
        for (int c = 0; c < 3; c++) {
          const double mx[3] = {S[3 + c], 0, 0}, mn[3] = {S[6 + c], 0, 0};
          check(lbm_io_write_raw8(name(pre[c], frame, s, ".raw").c_str(), comp[c]->data(), nullptr, nullptr,
                                  mask.data(), N, mx, mn),
                "write_raw u_i");
        }
      }
      check(lbm_io_write_dat(name("rho", frame, s, ".dat").c_str(), m.rho.data(), mask.data(), N), "write_dat");
      check(lbm_io_write_dat(name("u_x", frame, s, ".dat").c_str(), m.ux.data(), mask.data(), N), "write_dat");
      check(lbm_io_write_dat(name("u_y", frame, s, ".dat").c_str(), m.uy.data(), mask.data(), N), "write_dat");
      if (dims == 3)
        check(lbm_io_write_dat(name("u_z", frame, s, ".dat").c_str(), m.uz.data(), mask.data(), N), "write_dat");
    }
    const int LZ = dims == 3 ? d.LZ : 1;
    for (int s = 0; s < nsubs; s++)
      check(lbm_io_write_txt(name("rho", frame, s, ".txt").c_str(), mv[s].rho.data(), d.LX, d.LY, LZ), "rho txt");
    for (int s = 0; s < nsubs; s++) {
      check(lbm_io_write_txt(name("ux", frame, s, ".txt").c_str(), mv[s].ux.data(), d.LX, d.LY, LZ), "u txt");
      check(lbm_io_write_txt(name("uy", frame, s, ".txt").c_str(), mv[s].uy.data(), d.LX, d.LY, LZ), "u txt");
      check(lbm_io_write_txt(name("uz", frame, s, ".txt").c_str(), mv[s].uz.data(), d.LX, d.LY, LZ), "u txt");
    }
  }
};

int parse_int(const char *flag, const char *v) {
  char *end = nullptr;
  long x = strtol(v, &end, 10);
  if (!end || *end) die_msg(std::string("bad value for ") + flag + ": " + v);
  return (int)x;
}

int deck_main(int argc, char **argv) {
  Run R;
  std::string deck_name;
  int storage = LBM_STORE_F64, compute = LBM_COMPUTE_F64;
  for (int a = 1; a < argc; a++) {
    std::string s = argv[a];
    auto need = [&](const char *f) -> const char * {
      if (a + 1 >= argc) die_msg(std::string("missing value for ") + f);
      return argv[++a];
    };
    if (s == "--dims")
      R.dims = parse_int("--dims", need("--dims"));
    else if (s == "--subs")
      R.nsubs = parse_int("--subs", need("--subs"));
    else if (s == "--in")
      R.in_dir = need("--in");
    else if (s == "--out")
      R.out_dir = need("--out");
    else if (s == "--walsh")
      R.walsh = true;
    else if (s == "--storage") {
      std::string v = need("--storage");
      storage = v == "f64" ? LBM_STORE_F64 : v == "f32" ? LBM_STORE_F32 : v == "f16" ? LBM_STORE_F16 : -1;
      if (storage < 0) die_msg("--storage takes f64, f32 or f16");
    } else if (s == "--compute") {
      std::string v = need("--compute");
      compute = v == "f64" ? LBM_COMPUTE_F64 : v == "f32" ? LBM_COMPUTE_F32 : -1;
      if (compute < 0) die_msg("--compute takes f64 or f32");
    } else if (s == "-h" || s == "--help") {
      printf("usage: lbgpu_prime [name] [--dims 2|3] [--subs 1|2] [--storage f64|f32|f16] "
             "[--compute f64|f32] [--in DIR] [--out DIR] [--walsh]\n");
      return 0;
    } else if (!s.empty() && s[0] != '-' && deck_name.empty())
      deck_name = s;
    else
      die_msg("unknown argument " + s);
  }
  if (R.dims != 2 && R.dims != 3) die_msg("--dims must be 2 or 3");
  if (R.nsubs != 1 && R.nsubs != 2) die_msg("--subs must be 1 or 2");
  if (R.walsh && R.dims != 2)
    die_msg("--walsh: the reference sets ns from the 2-D solids image only (src/latman.c:413-416)");

  // construct_lattice (src/latman.c:91-493)
  R.params_path = deck_name.empty() ? R.in_dir + "/params.in" : "./in_" + deck_name + "/params.in";
  printf("Reading params from file \"%s\"\n", R.params_path.c_str());
  check(lbm_io_read_deck(R.params_path.c_str(), R.dims, R.nsubs, &R.d), "read_params");
  lbm_deck &d = R.d;
  const int LZ = R.dims == 3 ? d.LZ : 1;
  R.N = (long)d.LX * d.LY * LZ;
  if (d.LX <= 0 || d.LY <= 0 || LZ <= 0) die_msg("lattice dimensions must be positive");
  // the block-size checks of main (src/lbgpu_prime.cu:113-141)
  if (d.BX <= 0 || d.LX % d.BX != 0)
    die_msg("LX = " + std::to_string(d.LX) + " must be divisible by BX = " + std::to_string(d.BX));
  if (d.BY <= 0 || d.LY % d.BY != 0)
    die_msg("LY = " + std::to_string(d.LY) + " must be divisible by BY = " + std::to_string(d.BY));
  if (R.dims == 3 && (d.BZ <= 0 || d.LZ % d.BZ != 0))
    die_msg("LZ = " + std::to_string(d.LZ) + " must be divisible by BZ = " + std::to_string(d.BZ));

  // solids (src/latman.c:379-428)
  R.mask.assign(R.N, 0);
  char fn[1024];
  if (R.dims == 2) {
    snprintf(fn, sizeof fn, "%s/%dx%d.bmp", R.in_dir.c_str(), d.LX, d.LY);
    FILE *probe = fopen(fn, "rb");
    if (probe) {
      fclose(probe);
      printf("Reading solids from file \"%s\"\n", fn);
      check(lbm_io_read_solids_bmp(fn, d.LX, d.LY, R.mask.data()), "spy_bmp");
    } else {
      // spy_bmp writes a blank image and reads it back: no solids
      printf("File \"%s\" cannot be opened for reading; no solids.\n", fn);
    }
  } else {
    snprintf(fn, sizeof fn, "%s/%dx%dx%d.raw", R.in_dir.c_str(), d.LX, d.LY, d.LZ);
    printf("Reading solids from file \"%s\"\n", fn);
    check(lbm_io_read_solids_raw(fn, d.LX, d.LY, d.LZ, R.mask.data()), "read_solids");
  }

  lbm_params p;
  check(lbm_io_deck_params(&d, R.dims, R.nsubs, &p), "params");
  p.storage = storage;
  p.compute = compute;
  // pressure series (src/latman.c:536-631)
  std::vector<double> series[2];
  const char *sname[2] = {"pressure_n_in0.in", "pressure_s_in0.in"};
  const int want[2] = {d.pressure_n_in[0] == 2, d.pressure_s_in[0] == 2};
  for (int q = 0; q < 2; q++) {
    if (!want[q]) continue;
    std::string path = R.in_dir + "/" + sname[q];
    int cnt = 0;
    check(lbm_io_read_series(path.c_str(), nullptr, 0, &cnt), "pressure series");
    series[q].resize(cnt > 0 ? cnt : 1);
    check(lbm_io_read_series(path.c_str(), series[q].data(), cnt, &cnt), "pressure series");
    if (cnt <= 0) die_msg(path + " holds no values");
    series[q].resize(cnt);
    printf("num %s = %d\n", sname[q], cnt);
  }
  if (want[0]) {
    p.pressure_n_series[0] = series[0].data();
    p.pressure_n_series_len[0] = (int)series[0].size();
  }
  if (want[1]) {
    p.pressure_s_series[0] = series[1].data();
    p.pressure_s_series_len[0] = (int)series[1].size();
  }

  HandleGuard guard;
  check(lbm_create(&p, &guard.h), "construct_lattice");
  lbm_handle *h = guard.h;
  if (R.walsh) {
    // every node is computed; image solids take ns, the others 0
    std::vector<uint8_t> none(R.N, 0);
    std::vector<double> ns(R.N);
    for (long n = 0; n < R.N; n++) ns[n] = R.mask[n] == 1 ? d.ns : 0.0;
    check(lbm_set_solids(h, none.data()), "solids H2D");
    check(lbm_set_ns(h, ns.data()), "ns H2D");
  } else {
    check(lbm_set_solids(h, R.mask.data()), "solids H2D");
  }

  // init_problem, frame 0 from the host values (src/lbgpu_prime.cu:97-101)
  R.init_problem();
  R.output_frame(0, 0);
  check(lbm_init_rest(h, R.nsubs == 2 ? R.mv[1].rho.data() : nullptr), "f H2D");
  for (int s = 0; s < R.nsubs; s++)
    check(lbm_upload_macro(h, s, R.mv[s].rho.data(), R.mv[s].ux.data(), R.mv[s].uy.data(),
                           R.dims == 3 ? R.mv[s].uz.data() : nullptr),
          "mv H2D");

  // PEST observation points (read_PEST_in_files, src/latman.c:889-1038),
  // used when the deck has a timestep file
  const std::string obs_t = R.in_dir + "/timestep_file.in";
  FILE *probe_obs = fopen(obs_t.c_str(), "r");
  const bool pest = probe_obs != nullptr;
  if (probe_obs) fclose(probe_obs);
  std::vector<int> ot, ox, oy, oorder;
  std::vector<double> oconc, ovel;
  if (pest) {
    const std::string xs = R.in_dir + "/x_coord_file.in", ys = R.in_dir + "/y_coord_file.in";
    int cnt = 0;
    check(lbm_io_read_observations(obs_t.c_str(), xs.c_str(), ys.c_str(), d.LX, d.LY, 0, nullptr, nullptr,
                                   nullptr, nullptr, &cnt),
          "read_PEST_in_files");
    ot.resize(cnt);
    ox.resize(cnt);
    oy.resize(cnt);
    oorder.resize(cnt);
    check(lbm_io_read_observations(obs_t.c_str(), xs.c_str(), ys.c_str(), d.LX, d.LY, cnt, ot.data(), ox.data(),
                                   oy.data(), oorder.data(), &cnt),
          "read_PEST_in_files");
    oconc.assign(cnt, 0.0);
    ovel.assign(cnt, 0.0);
    printf("Number of PEST points = %d\n", cnt);
  }
  size_t opos = 0;
  const int D = R.dims;
  std::vector<double> sample(1 + D);

  // time loop (src/lbgpu_prime.cu:369-915): the odd and even half-steps,
  // frame output after the half-step that ends a frame, stop at the final
  // frame
  auto t0 = std::chrono::steady_clock::now();
  double io_s = 0.0;
  int frame = 0, steps = 0;
  for (int time = 1; time <= d.NumTimeSteps; time++) {
    const int end_of_frame = d.FrameRate > 0 && time % d.FrameRate == 0;
    const bool observe = pest && opos < ot.size() && ot[opos] == time;
    const int macro = end_of_frame ? LBM_MACRO_FRAME : observe ? LBM_MACRO_PEST : LBM_MACRO_NONE;
    check(lbm_step(h, macro), time % 2 ? "k_stream_collide_stream" : "k_collide");
    steps++;
    // write_PEST_out_data (src/latman.c:1072-1190): the solute
    // concentration and flow speed at each point observed at this time
    while (pest && opos < ot.size() && ot[opos] == time) {
      const long n = (long)oy[opos] * d.LX + ox[opos];
      double c = 0.0, ux = 0.0, uy = 0.0;
      if (R.kept(n)) {
        if (R.nsubs == 2) c = R.init_mv[1].rho[n];
      } else {
        if (R.nsubs == 2) {
          check(lbm_sample_macro(h, 1, 1, &n, sample.data()), "PEST transfer");
          c = sample[0];
        }
        check(lbm_sample_macro(h, 0, 1, &n, sample.data()), "PEST transfer");
        ux = sample[1];
        uy = sample[2];
      }
      oconc[opos] = c;
      ovel[opos] = sqrt(ux * ux + uy * uy);
      opos++;
    }
    if (!end_of_frame) continue;
    // at every frame: no device error, no NaN/infinite population, no
    // negative flow population (checkCUDAError and PUKE_NEGATIVE_DENSITIES:
    // the run stops with exit status 1)
    check(lbm_check(h, nullptr, 0.0, nullptr), "health check");
    auto a = std::chrono::steady_clock::now();
    ++frame;
    for (int s = 0; s < R.nsubs; s++) {
      Macro &m = R.mv[s];
      check(lbm_download_macro(h, s, m.rho.data(), m.ux.data(), m.uy.data(), R.dims == 3 ? m.uz.data() : nullptr),
            "mv D2H");
      // solid nodes keep what was uploaded at the start (the kernels never
      // write them)
      const Macro &i0 = R.init_mv[s];
      for (long n = 0; n < R.N; n++)
        if (R.kept(n)) {
          m.rho[n] = i0.rho[n];
          m.ux[n] = i0.ux[n];
          m.uy[n] = i0.uy[n];
          if (R.dims == 3) m.uz[n] = i0.uz[n];
        }
    }
    printf("Begin file I/O at time = %d, frame = %d.\n", time, frame);
    R.output_frame(time, frame);
    io_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    if (time % 2 == 1 && frame == d.NumFrames) {
      printf("\n breaking out of time loop on frame %d\n", frame);
      break;
    }
  }
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (pest)
    check(lbm_io_write_observations(R.out_dir.c_str(), (int)ot.size(), oconc.data(), ovel.data(), oorder.data()),
          "write_PEST_out_file");
  const double comp = total - io_s;
  printf("lbgpu_prime: %d steps of %ld nodes in %.3f s (%.3f s stepping, %.3f s file I/O): %.1f MLUPS\n", steps,
         R.N, total, comp, io_s, comp > 0 ? (double)R.N * steps / comp / 1e6 : 0.0);
  guard.h = nullptr;
  check(lbm_destroy(h), "destruct_lattice");
  return 0;
}

}  // namespace

extern "C" int lbm_deck_main(int argc, char **argv) {
  try {
    return deck_main(argc, argv);
  } catch (const Exit &e) {
    fflush(stdout);
    fprintf(stderr, "lbgpu_prime: %s\n", e.msg.c_str());
    return 1;
  } catch (const std::exception &e) {
    fflush(stdout);
    fprintf(stderr, "lbgpu_prime: %s\n", e.what());
    return 1;
  }
}
