// C ABI of the stream-collide library: handle, lattice set-up, time loop
// pieces, transfers, reductions and timing.  See include/lbm_mi355x.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "lbm_adjoint.hpp"
#include "lbm_aux.hpp"
#include "lbm_device.hpp"
#include "lbm_internal.h"
#include "lbm_mi355x.h"

using namespace lbm;

static thread_local std::string g_err;

// Staging memory of lbm_download_f_noalloc: one [Q][N] distribution in
// FP64, at most kScratchDoubles elements (64 MiB: D3Q19 lattices up to
// 441,505 nodes, D2Q9 up to 932,067).  Each GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` only names
// the symbol: its device address on a GPU comes from hipGetSymbolAddress
// with that GPU current (scratch_on_device), and that address is what
// kernels and copies take.
constexpr long kScratchDoubles = 8L << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

extern "C" int lbm_fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

extern "C" const char *lbm_last_error(void) { return g_err.c_str(); }

#define HIP_OK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return lbm_fail("%s:%d %s: %s", __FILE__, __LINE__, #call,      \
                                          hipGetErrorString(e_));                         \
  } while (0)

typedef _Float16 half_t;

namespace {

struct ProfPair {
  hipEvent_t a, b;
  int cls;
};

}  // namespace

struct lbm_handle {
  lbm_params p;
  std::vector<double> series[4];  // n[0], n[1], s[0], s[1]
  std::vector<double> cmap_host;
  int D, Q, U, L;  // dims, directions, unbound directions, last-axis length
  // the z faces' conditions (3-D), resolved from zface_flow / zface_conc,
  // zface_pressure and zface_solute: ZFlow / ZConc per face, and the
  // prescribed values (pressure: rho_b; velocity: rho_A u_z; Dirichlet: c)
  int zflow[2] = {kZNone, kZNone}, zconc[2] = {kCNone, kCNone};
  double zfval[2] = {0., 0.}, zcval[2] = {0., 0.};
  long N, E;
  int nf = 0;
  long tiles = 0;  // kTile-node tiles incl. ghost entries
  int W = 0;       // 64-node words per lattice row
  int nblk8 = 0;  // blocks of kBlock threads / 8
  int anblk8 = 0; // blocks of kAdjBlock threads / 8
  int *bmap = nullptr;  // row-band block map of the odd half-step (its workgroup size)
  int bmap8 = 0;        // its length / 8
  int bmap_bs = 0;      // the node block it was built for
  int lds_cu = 0;       // LDS per CU of the device (bytes)
  int *amap = nullptr;  // row-band block map of the adjoint step (kAdjBlock threads)
  size_t st_bytes = 8;
  void *F = nullptr;
  void *Fout = nullptr;  // where the half-steps write when not F (out-of-place steps)
  // lbm_run's captured chunks of steps, one per starting parity and solute
  // window state (inside the sigma_t_on..sigma_t_off window or not: the
  // launches differ in solute_on and in the fused z-face conditions)
  unsigned gen = 0;  // bumped when the geometry or the Walsh field changes
  hipGraphExec_t chunk[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  unsigned chunk_gen[2][2] = {{0, 0}, {0, 0}};
  void *chunk_F[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  bool chunk_off = false;  // a capture failed: steps are launched one by one
  RowWord *rw = nullptr;
  double *mv = nullptr;
  int *node = nullptr;  // compact -> dense
  std::vector<int> fid; // dense -> compact or -1
  std::vector<int> node_host;
  int *face_lo = nullptr, *face_hi = nullptr, *face_flux = nullptr;
  int n_lo = 0, n_hi = 0, n_flux = 0;
  int zb_ok = 0;  // the z faces are the compact ranges [0, n_lo) and [nf - n_hi, nf)
  double *ns = nullptr;  // [nf] Walsh partial bounce-back fraction, or null
  bool off64 = false;    // 64-bit element offsets at every size (LBM_OFFSET64=1)
  int *plane_lo = nullptr, *plane_hi = nullptr, *tgt_lo = nullptr, *tgt_hi = nullptr;
  void *halo_pos = nullptr, *halo_neg = nullptr;
  double *cmap = nullptr;
  double *red = nullptr;      // reduction partials
  double *flux_dev = nullptr; // [2 + 2 * kFluxBlocks]: result, then block partials
  double flux = 0., fmass = 0.;
  int flux_valid = 0;
  int time = 0;
  hipStream_t stream = nullptr;
  double *scratch = nullptr;  // this device's address of the static staging array
  int profile = 0;
  std::vector<ProfPair> pending;
  std::vector<hipEvent_t> pool;
  double prof_ms[LBM_PROFILE_CLASSES] = {};
  int prof_n[LBM_PROFILE_CLASSES] = {};
  // objective run and its adjoint
  // Buffers stay allocated between calls (the next objective or adjoint
  // run of the same lattice reuses them) and are released with the lattice.
  struct {
    int valid = 0;
    int kind = LBM_OBJ_OUTLET_FLUX;  // the window term (lbm_objective_select)
    int t0 = 0, nsteps = 0, window = 0, every = 0;
    std::vector<void *> ckpt;  // state before steps t0+1, t0+every+1, ...
    size_t keep_bytes = 0;     // lbm_objective_keep_states
    std::vector<void *> kept;  // states (after the BCs) before the last kept steps of the run
    int kept_from = 0;         // the run's first kept step index (0-based; nsteps: none)
    int kept_replay = 0;       // the sweep replays into kept[0..every] (no segment buffers
                               // of its own), which then no longer hold the run's states
    double *jacc = nullptr;
    double J = 0.0;
    double Jwin = 0.0;         // the window term of J
    std::vector<void *> seg;   // adjoint: the states of one replayed segment
    double *L = nullptr;       // adjoint state
    double *gacc = nullptr;    // per-block parameter gradients
    double *cgrad = nullptr;   // inlet-map gradient
    // observation points (lbm_objective_observations), sorted by step
    std::vector<int> o_step, o_node;          // step of the run (1..nsteps), compact node
    std::vector<double> o_target, o_weight;
    int *o_node_d = nullptr;                  // compact nodes, sorted by step
    double *o_c = nullptr;                    // observed concentrations of the last run
    // dJ/dc per (step, node), merged
    std::vector<int> s_step, s_begin;         // per step with observations: first entry
    int *s_node_d = nullptr;
    double *s_val_d = nullptr;
    double *seed = nullptr;                   // [nf] per-node dJ/dc of the step being swept
  } obj;
};

namespace {

constexpr int kRedBlocks = 512;
// largest LX: the row words hold the ranks of a row's nodes next to a word
// as 16-bit offsets from the word's base, which are at most LX
constexpr int kMaxLX = 32767;

hipEvent_t ev_get(lbm_handle *h) {
  if (!h->pool.empty()) {
    hipEvent_t e = h->pool.back();
    h->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct Prof {
  lbm_handle *h;
  int cls;
  hipEvent_t a = nullptr;
  Prof(lbm_handle *h_, int c) : h(h_), cls(c) {
    if (h->profile) {
      a = ev_get(h);
      if (a) (void)hipEventRecord(a, h->stream);
    }
  }
  ~Prof() {
    if (h->profile && a) {
      hipEvent_t b = ev_get(h);
      if (b) {
        (void)hipEventRecord(b, h->stream);
        h->pending.push_back({a, b, cls});
      }
    }
  }
};

template <typename ST, typename CT>
StepArgs<ST, CT> make_args(const lbm_handle *h) {
  StepArgs<ST, CT> A;
  const lbm_params &p = h->p;
  A.F = (ST *)h->F;
  A.Fo = (ST *)(h->Fout ? h->Fout : h->F);
  A.G.rw = h->rw;
  A.G.node = h->node;
  A.G.ni = p.ni;
  A.G.nj = p.nj;
  A.G.nk = h->D == 3 ? p.nk : 1;
  A.G.W = h->W;
  A.G.halo = p.halo_mode;
  A.G.ghost = h->nf;
  A.G.E = (int)h->E;
  A.G.div_ni = make_fastdiv((uint32_t)p.ni);
  A.G.div_nj = make_fastdiv((uint32_t)p.nj);
  A.nf = h->nf;
  A.nblk8 = h->nblk8;
  A.bmap = nullptr;
  A.mv = h->mv;
  A.write_mv = 0;
  A.solute_on = 0;
  A.rhoA = (CT)p.rho_A[0];
  A.tau = (CT)p.tau[0];
  A.inv_tau = (CT)(1. / p.tau[0]);
  for (int d = 0; d < 3; d++) A.g[d] = (CT)p.gaccel[d];
  A.rhoA_s = (CT)p.rho_A[1];
  A.inv_tau_s = (CT)(p.nsubs == 2 ? 1. / p.tau[1] : 1.);
  A.sink_fac = (CT)(1.0 - p.sink * p.tau[1]);
  for (int s = 0; s < 2; s++) {
    double sc = p.store_scale[s];
    if (h->p.storage == LBM_STORE_F16) {
      // FP16 keeps 11 significant bits from 6.1e-5 to 65504: 2^14 places
      // the flow's f~ (|f~| < 4 at the most) and 2^12 the solute's g (c up
      // to about 40) in that range, so small populations keep their
      // relative precision down to 4e-9 (flow) and 1.5e-8 (solute)
      if (sc <= 0) sc = s == 0 ? 16384.0 : 4096.0;
    } else {
      sc = 1.0;
    }
    A.scale[s] = (CT)sc;
    A.inv_scale[s] = (CT)(1.0 / sc);
  }
  A.c0 = (CT)(h->p.storage == LBM_STORE_F64 ? 0.0 : p.store_shift_c);
  A.zb = ZfaceFuse{};
  A.ns = nullptr;
  return A;
}

// Whether the z-face conditions of step t can be formed inside the
// half-step (k_half_step / the face blocks of k_adjoint_step) instead of a separate launch
// storing them first.  Not when the plane flux is taken on a face (it reads
// the stored boundary values) or when the solute condition would write into
// a resting solute (the separate launch does, and later steps see it).
bool zface_any(const lbm_handle *h) {
  return h->zflow[0] != kZNone || h->zflow[1] != kZNone || h->zconc[0] != kCNone || h->zconc[1] != kCNone;
}

bool zface_fused(const lbm_handle *h, int t) {
  const lbm_params &p = h->p;
  if (h->D != 3 || !h->zb_ok || !zface_any(h)) return false;
  const bool sol = h->zconc[0] != kCNone || h->zconc[1] != kCNone;
  if (h->n_flux && (p.flux_plane == 0 || p.flux_plane == p.nk - 1)) return false;
  if (sol && !(t >= p.sigma_t_on && t <= p.sigma_t_off)) return false;
  return true;
}

template <typename ST, typename CT>
void set_zb(const lbm_handle *h, StepArgs<ST, CT> &A) {
  A.zb.on = 1;
  A.zb.lo_end = h->n_lo;
  A.zb.hi_begin = h->nf - h->n_hi;
  for (int f = 0; f < 2; f++) {
    A.zb.flow[f] = h->zflow[f];
    A.zb.conc[f] = h->zconc[f];
    A.zb.fval[f] = h->zfval[f];
    A.zb.cval[f] = h->zcval[f];
  }
  A.zb.cmap = h->cmap;
}

// the ZfaceArgs of one face for the separate boundary launches
ZfaceArgs zface_args(const lbm_handle *h, int top) {
  ZfaceArgs Z;
  Z.top = top;
  Z.flow = h->zflow[top];
  Z.conc = h->zconc[top];
  Z.fval = h->zfval[top];
  Z.cval = h->zcval[top];
  Z.cmap = h->cmap;
  return Z;
}

// Calls fn.template operator()<Q, NS, ST, CT>() for the handle's kind.
template <typename Fn>
int dispatch(const lbm_handle *h, Fn &&fn) {
  const int st = h->p.storage, ct = h->p.compute;
#define LBM_DISPATCH_T(QQ, NN)                                                        \
  if (st == LBM_STORE_F64 && ct == LBM_COMPUTE_F64) return fn.template operator()<QQ, NN, double, double>(); \
  if (st == LBM_STORE_F64 && ct == LBM_COMPUTE_F32) return fn.template operator()<QQ, NN, double, float>();  \
  if (st == LBM_STORE_F32 && ct == LBM_COMPUTE_F64) return fn.template operator()<QQ, NN, float, double>();  \
  if (st == LBM_STORE_F32) return fn.template operator()<QQ, NN, float, float>();                            \
  if (st == LBM_STORE_F16) return fn.template operator()<QQ, NN, half_t, float>();
  if (h->Q == 9 && h->p.nsubs == 1) { LBM_DISPATCH_T(9, 1) }
  if (h->Q == 9 && h->p.nsubs == 2) { LBM_DISPATCH_T(9, 2) }
  if (h->Q == 19 && h->p.nsubs == 1) { LBM_DISPATCH_T(19, 1) }
  if (h->Q == 19 && h->p.nsubs == 2) { LBM_DISPATCH_T(19, 2) }
#undef LBM_DISPATCH_T
  return lbm_fail("unsupported storage/compute combination");
}

inline unsigned blocks_for(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

void free_objective(lbm_handle *h) {
  for (void *q : h->obj.ckpt)
    if (q) (void)hipFree(q);
  h->obj.ckpt.clear();
  for (void *q : h->obj.seg)
    if (q) (void)hipFree(q);
  h->obj.seg.clear();
  for (void *q : h->obj.kept)
    if (q) (void)hipFree(q);
  h->obj.kept.clear();
  void *bufs[] = {h->obj.jacc, h->obj.L, h->obj.gacc, h->obj.cgrad, h->obj.o_node_d, h->obj.o_c,
                  h->obj.s_node_d, h->obj.s_val_d, h->obj.seed};
  for (void *q : bufs)
    if (q) (void)hipFree(q);
  h->obj.jacc = h->obj.L = h->obj.gacc = h->obj.cgrad = h->obj.o_c = h->obj.s_val_d = h->obj.seed = nullptr;
  h->obj.o_node_d = h->obj.s_node_d = nullptr;
  h->obj.o_step.clear();
  h->obj.o_node.clear();
  h->obj.o_target.clear();
  h->obj.o_weight.clear();
  h->obj.s_step.clear();
  h->obj.s_begin.clear();
  h->obj.valid = 0;
}

void drop_chunks(lbm_handle *h) {
  for (int par = 0; par < 2; par++)
    for (int win = 0; win < 2; win++) {
      if (h->chunk[par][win]) (void)hipGraphExecDestroy(h->chunk[par][win]);
      h->chunk[par][win] = nullptr;
    }
  h->gen++;
}

void free_lattice(lbm_handle *h) {
  drop_chunks(h);
  free_objective(h);
  void *ptrs[] = {h->F, h->rw, h->mv, h->node, h->face_lo, h->face_hi, h->face_flux, h->plane_lo,
                  h->plane_hi, h->tgt_lo, h->tgt_hi, h->halo_pos, h->halo_neg, h->ns, h->bmap, h->amap};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  h->F = nullptr;
  h->rw = nullptr;
  h->mv = nullptr;
  h->node = nullptr;
  h->face_lo = h->face_hi = h->face_flux = nullptr;
  h->plane_lo = h->plane_hi = h->tgt_lo = h->tgt_hi = nullptr;
  h->halo_pos = h->halo_neg = nullptr;
  h->ns = nullptr;
  h->bmap = h->amap = nullptr;
  h->bmap8 = h->bmap_bs = 0;
  h->nf = 0;
}

int need_lattice(const lbm_handle *h) {
  if (!h) return lbm_fail("null handle");
  if (!h->F) return lbm_fail("lbm_set_solids has not been called");
  return 0;
}

// the device address of the static staging array on the current device,
// looked up once per device
int scratch_on_device(double **out) {
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return lbm_fail("device %d beyond the library's %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return lbm_fail("no device address for the staging array");
    g_scratch[dev] = (double *)ptr;
  }
  *out = g_scratch[dev];
  return 0;
}

// rho selection of bcs_1 / bcs_2 with time series and ramp (src/bcs.c:287-319)
double bc_rho(const lbm_handle *h, int s, int which, int t) {
  const lbm_params &p = h->p;
  double rho, ramp;
  const std::vector<double> &sn = h->series[s], &ss = h->series[2 + s];
  if (which == 0 && p.pressure_n_in[s] == 2 && !sn.empty())
    rho = p.pfadd + p.pfmul * sn[t % sn.size()];
  else if (which == 2 && p.pressure_s_in[s] == 2 && !ss.empty())
    rho = p.pfadd + p.pfmul * ss[t % ss.size()];
  else if (which == 0 || which == 2)
    rho = p.rho_in;
  else
    rho = p.rho_out;
  if (p.bc_ramp_start >= 0 && p.bc_ramp_stop >= 0) {
    if (t < p.bc_ramp_start) {
      rho = p.rho_A[s];
    } else if (t < p.bc_ramp_stop) {
      ramp = ((double)(t - p.bc_ramp_start)) / ((double)(p.bc_ramp_stop - p.bc_ramp_start));
      rho *= ramp;
      rho += (1. - ramp) * p.rho_A[s];
    }
  }
  return rho;
}

// u_y of a 2-D velocity row (uy_in for the *_in flags, uy_out for *_out),
// with the pressure rows' ramp: 0 before bc_ramp_start, rising linearly to
// the full value at bc_ramp_stop (which = 0 n_in, 1 n_out, 2 s_in, 3 s_out)
double bc_uy(const lbm_handle *h, int which, int t) {
  const lbm_params &p = h->p;
  double u = (which == 0 || which == 2) ? p.uy_in : p.uy_out;
  if (p.bc_ramp_start >= 0 && p.bc_ramp_stop >= 0) {
    if (t < p.bc_ramp_start) {
      u = 0.0;
    } else if (t < p.bc_ramp_stop) {
      const double ramp = ((double)(t - p.bc_ramp_start)) / ((double)(p.bc_ramp_stop - p.bc_ramp_start));
      u *= ramp;
    }
  }
  return u;
}

int upload_ints(int **dst, const std::vector<int> &v) {
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  size_t n = v.empty() ? 1 : v.size();
  HIP_OK(hipMalloc(dst, n * sizeof(int)));
  if (!v.empty()) HIP_OK(hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

// Row-band block map of nb blocks of bs compact nodes, launched as 8 * per
// workgroups: the blocks are ordered by the lattice row (y) of their first
// node (blocks past the last node last) and cut into 8 parts; workgroup
// c + 8 i runs on XCD c (round-robin dispatch) and takes the i-th block of
// part c in node order (-1 past the end of a part).  Each XCD thus sweeps a
// band of rows through all z planes, so the lines a node shares with its z
// neighbours (the partial lines of bounce-back own slots) are reused one
// band-plane later (about 1 MB of traffic at 256^3) instead of one full plane
// later (8 MB, beyond the XCD's 4 MB L2).  The results do not depend on the
// block order: in either AA half-step every node reads and writes only its
// own locations.
int upload_band_map(lbm_handle *h, int bs, long nb, long per, int **dst) {
  const long LX = h->p.ni, LY = h->p.nj;
  std::vector<long> ord(nb);
  for (long b = 0; b < nb; b++) ord[b] = b;
  auto row = [&](long b) { return b * bs < h->nf ? ((long)h->node_host[b * bs] / LX) % LY : LY; };
  std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return row(a) < row(b); });
  std::vector<int> map(8 * per, -1);
  for (int c = 0; c < 8; c++) {
    const long b0 = nb * c / 8, b1 = nb * (c + 1) / 8;
    std::sort(ord.begin() + b0, ord.begin() + b1);
    for (long i = b0; i < b1; i++) map[c + 8 * (i - b0)] = (int)ord[i];
  }
  HIP_OK(hipMalloc(dst, map.size() * sizeof(int)));
  HIP_OK(hipMemcpy(*dst, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

// the odd half-step's workgroup size for the handle's kind, and the LDS a
// half-step workgroup reserves
int odd_block(const lbm_handle *h) {
  return dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    return StepShape<true, NS, ST, CT>::block;
  });
}
template <bool ODD, int NS, typename ST, typename CT>
size_t step_pad(const lbm_handle *h) {
  return step_lds_pad(StepShape<ODD, NS, ST, CT>::per_cu, h->lds_cu);
}

// The 3-D lattice's block maps.  Odd half-step (blocks of odd_block()):
// only when an XCD's band of one plane holds at least 1536 nodes (3 blocks
// of 512, the size the threshold was measured with).  Measured on sphere packs
// of porosity 0.40, odd + even step: 256^3 (6.7 blocks) flow FP64 0.777 ->
// 0.727 ms, flow + solute FP64 1.591 -> 1.497 ms, FP32 0.785 -> 0.741 ms;
// 192^3 (3.9 blocks) 1 to 3 % faster; 128^3 (1.7 blocks) 2 to 4 % slower,
// the even step that follows losing more than the odd step gains.  Adjoint
// step (kAdjBlock, all 8 * anblk8 blocks so that every gradient slot is
// written): at 128^3 0.319 -> 0.305 s of adjoint kernels per sweep.
int build_band_maps(lbm_handle *h) {
  // the odd step's node block for this kind
  const int bso = odd_block(h);
  const long nbo = blocks_for(h->nf, bso);
  if ((double)h->nf / h->p.nk / 8.0 >= 3.0 * kBlockOdd) {
    const long per = (nbo + 7) / 8;
    if (int rc = upload_band_map(h, bso, nbo, per, &h->bmap)) return rc;
    h->bmap8 = (int)per;
    h->bmap_bs = bso;
  }
  return upload_band_map(h, kAdjBlock, 8L * h->anblk8, h->anblk8, &h->amap);
}

// the odd half-step's grid: the row-band map when there is one for this
// workgroup size
template <typename ST, typename CT>
dim3 odd_grid(const lbm_handle *h, StepArgs<ST, CT> &A, int bs) {
  A.nblk8 = (int)((blocks_for(h->nf, bs) + 7) / 8);
  A.bmap = nullptr;
  if (bs == h->bmap_bs && h->bmap) {
    A.bmap = h->bmap;
    A.nblk8 = h->bmap8;
  }
  return dim3(8 * A.nblk8);
}

int launch_half(lbm_handle *h, int odd, int t, int write_mv, bool fuse_bc = false) {
  return dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    if (fuse_bc) set_zb(h, A);
    A.write_mv = write_mv;
    A.solute_on = NS == 2 && t >= h->p.sigma_t_on && t <= h->p.sigma_t_off;
    Prof pr(h, odd ? 0 : 1);
    const bool off32 = !h->off64 && (size_t)h->tiles * NS * Q * kTile * sizeof(ST) <= 0xffffffffull;
    const int bs = odd ? half_step_block<true, NS, ST, CT>() : half_step_block<false, NS, ST, CT>();
    A.nblk8 = (int)((blocks_for(h->nf, bs) + 7) / 8);
    const dim3 grid = odd ? odd_grid(h, A, bs) : dim3(8 * A.nblk8), block(bs);
    const size_t lds = odd ? step_pad<true, NS, ST, CT>(h) : step_pad<false, NS, ST, CT>(h);
    if (h->ns) {
      // Walsh partial bounce-back
      A.ns = h->ns;
      if (odd && off32)
        hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, true, true, true>), grid, block, lds, h->stream, A);
      else if (odd)
        hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, true, false, true>), grid, block, lds, h->stream, A);
      else if (off32)
        hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, false, true, true>), grid, block, lds, h->stream, A);
      else
        hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, false, false, true>), grid, block, lds, h->stream, A);
    } else if (odd && off32)
      hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, true, true>), grid, block, lds, h->stream, A);
    else if (odd)
      hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, true, false>), grid, block, lds, h->stream, A);
    else if (off32)
      hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, false, true>), grid, block, lds, h->stream, A);
    else
      hipLaunchKernelGGL((k_half_step<Q, NS, ST, CT, false, false>), grid, block, lds, h->stream, A);
    HIP_OK(hipGetLastError());
    return 0;
  });
}

int launch_bcs(lbm_handle *h, int phase, int t) {
  const lbm_params &p = h->p;
  return dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    BcArgs<ST, CT> B;
    B.A = make_args<ST, CT>(h);
    B.phase = phase;
    Prof pr(h, 2);
    if constexpr (Q == 9) {
      for (int s = 0; s < p.nsubs; s++) {
        struct { int on, north, which; } pr4[4] = {{p.pressure_n_in[s], 1, 0}, {p.pressure_n_out[s], 1, 1},
                                                   {p.pressure_s_in[s], 0, 2}, {p.pressure_s_out[s], 0, 3}};
        for (auto &c : pr4) {
          if (!c.on) continue;
          B.list = c.north ? h->face_hi : h->face_lo;
          B.n = c.north ? h->n_hi : h->n_lo;
          if (!B.n) continue;
          const double rb = bc_rho(h, s, c.which, t);
          if (s == 0)
            hipLaunchKernelGGL((k_bc2d_pressure<NS, ST, CT, 0>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B,
                               c.north, rb, 0);
          else if constexpr (NS == 2)
            hipLaunchKernelGGL((k_bc2d_pressure<NS, ST, CT, 1>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B,
                               c.north, rb, 0);
        }
        // velocity rows (flow): the prescribed y momentum rho_A u_y
        if (s == 0) {
          struct { int on, north, which; } vr[4] = {{p.velocity_n_in, 1, 0}, {p.velocity_n_out, 1, 1},
                                                    {p.velocity_s_in, 0, 2}, {p.velocity_s_out, 0, 3}};
          for (auto &c : vr) {
            if (!c.on) continue;
            B.list = c.north ? h->face_hi : h->face_lo;
            B.n = c.north ? h->n_hi : h->n_lo;
            if (!B.n) continue;
            const double mom = p.rho_A[0] * bc_uy(h, c.which, t);
            hipLaunchKernelGGL((k_bc2d_pressure<NS, ST, CT, 0>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B,
                               c.north, mom, 1);
          }
        }
        if constexpr (NS == 2) if (s == 1) {
          if (p.zeroconcgrad_n && h->n_hi) {
            B.list = h->face_hi;
            B.n = h->n_hi;
            hipLaunchKernelGGL((k_bc2d_zerograd<ST, CT>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B, 1);
          }
          if (p.zeroconcgrad_s && h->n_lo) {
            B.list = h->face_lo;
            B.n = h->n_lo;
            hipLaunchKernelGGL((k_bc2d_zerograd<ST, CT>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B, 0);
          }
        }
      }
    } else {
      for (int top = 0; top < 2; top++) {
        ZfaceArgs Z = zface_args(h, top);
        if (NS != 2) Z.conc = kCNone;
        if (Z.flow == kZNone && Z.conc == kCNone) continue;
        B.list = top ? h->face_hi : h->face_lo;
        B.n = top ? h->n_hi : h->n_lo;
        if (!B.n) continue;
        hipLaunchKernelGGL((k_bc_zface<NS, ST, CT>), dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B, Z);
      }
    }
    HIP_OK(hipGetLastError());
    return 0;
  });
}

HaloMaps halo_maps(const lbm_handle *h) {
  HaloMaps H;
  H.plane_lo = h->plane_lo;
  H.plane_hi = h->plane_hi;
  H.tgt_lo = h->tgt_lo;
  H.tgt_hi = h->tgt_hi;
  H.E = (int)h->E;
  H.pairs = (h->Q - h->U) / 2;
  H.U = h->U;
  H.ghost = h->nf;
  H.literal = h->p.halo_literal;
  return H;
}

int halo_pack_unpack(lbm_handle *h, int phase, int pack, void *pos, void *neg) {
  if (!h->p.halo_mode) return lbm_fail("halo planes need halo_mode = 1");
  if (phase != 1 && phase != 2) return lbm_fail("phase must be 1 or 2");
  return dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    HaloMaps H = halo_maps(h);
    long total = (long)h->p.nsubs * H.pairs * H.E;
    Prof pr(h, 3);
    if (pack)
      hipLaunchKernelGGL((k_halo_pack<Q, NS, ST, CT>), dim3(blocks_for(total)), dim3(256), 0, h->stream, A, H,
                         h->p.nsubs, phase, (ST *)pos, (ST *)neg);
    else
      hipLaunchKernelGGL((k_halo_unpack<Q, NS, ST, CT>), dim3(blocks_for(total)), dim3(256), 0, h->stream, A, H,
                         h->p.nsubs, phase, (const ST *)pos, (const ST *)neg);
    HIP_OK(hipGetLastError());
    return 0;
  });
}

int launch_flux(lbm_handle *h, int phase, double *acc) {
  return dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    if constexpr (NS == 2) {
      BcArgs<ST, CT> B;
      B.A = make_args<ST, CT>(h);
      B.phase = phase;
      B.list = h->face_flux;
      B.n = h->n_flux;
      Prof pr(h, 4);
      hipLaunchKernelGGL((k_plane_flux<Q, NS, ST, CT>), dim3(kFluxBlocks), dim3(256), 0, h->stream, B,
                         h->flux_dev + 2);
      hipLaunchKernelGGL(k_plane_flux_final, dim3(1), dim3(64), 0, h->stream, h->flux_dev + 2, h->flux_dev, acc);
      HIP_OK(hipGetLastError());
    }
    return 0;
  });
}

int do_step(lbm_handle *h, int write_mv, double *flux_acc = nullptr) {
  int t = h->time + 1;
  int odd = t & 1;
  int phase = odd ? 1 : 2;
  if (h->p.halo_mode) {
    if (halo_pack_unpack(h, phase, 1, h->halo_pos, h->halo_neg)) return 1;
    if (halo_pack_unpack(h, phase, 0, h->halo_pos, h->halo_neg)) return 1;
  }
  const bool fz = zface_fused(h, t);
  if (!fz && launch_bcs(h, phase, t)) return 1;
  int sol_on = h->p.nsubs == 2 && t >= h->p.sigma_t_on && t <= h->p.sigma_t_off;
  // the plane flux is taken on steps with a macroscopic write and inside an objective window
  if (h->n_flux && sol_on && (write_mv || flux_acc)) {
    if (launch_flux(h, phase, flux_acc)) return 1;
    h->flux_valid = 1;
  }
  if (launch_half(h, odd, t, write_mv, fz)) return 1;
  h->time = t;
  return 0;
}

int sync(lbm_handle *h) {
  HIP_OK(hipStreamSynchronize(h->stream));
  return 0;
}

}  // namespace

extern "C" int lbm_create(const lbm_params *p, lbm_handle **out) {
  if (!p || !out) return lbm_fail("lbm_create: null argument");
  *out = nullptr;
  if (p->dims != 2 && p->dims != 3) return lbm_fail("dims must be 2 or 3");
  if (p->ni <= 0 || p->nj <= 0 || (p->dims == 3 && p->nk <= 0)) return lbm_fail("bad lattice size");
  if (p->nsubs != 1 && p->nsubs != 2) return lbm_fail("nsubs must be 1 or 2");
  if (p->halo_literal != 0 && p->halo_literal != 1) return lbm_fail("halo_literal must be 0 or 1");
  for (int s = 0; s < p->nsubs; s++)
    if (!(p->tau[s] > 0.5)) return lbm_fail("tau[%d] = %g must exceed 1/2", s, p->tau[s]);
  if (p->storage < 0 || p->storage > 2 || p->compute < 0 || p->compute > 1)
    return lbm_fail("bad storage/compute");
  if (p->storage == LBM_STORE_F16 && p->compute == LBM_COMPUTE_F64)
    return lbm_fail("FP16 storage runs the FP32 collision");
  long N = (long)p->ni * p->nj * (p->dims == 3 ? p->nk : 1);
  if (N >= (1L << 31) - (1L << 24)) return lbm_fail("lattice too large for 32-bit node indices");
  if (p->ni > kMaxLX) return lbm_fail("LX = %d above %d is not supported", p->ni, kMaxLX);
  if (p->flux_plane >= (p->dims == 2 ? p->nj : p->nk))
    return lbm_fail("flux_plane %d lies outside the lattice", p->flux_plane);
  for (int f = 0; f < 2; f++) {
    if (p->zface_flow[f] < LBM_ZFACE_DEFAULT || p->zface_flow[f] > LBM_ZFACE_VELOCITY)
      return lbm_fail("zface_flow[%d] = %d is not an lbm_zface_flow", f, p->zface_flow[f]);
    if (p->zface_conc[f] < LBM_ZCONC_DEFAULT || p->zface_conc[f] > LBM_ZCONC_ZEROGRAD)
      return lbm_fail("zface_conc[%d] = %d is not an lbm_zface_conc", f, p->zface_conc[f]);
  }
  if (p->dims == 2) {
    if ((p->velocity_n_in || p->velocity_n_out) && (p->pressure_n_in[0] || p->pressure_n_out[0]))
      return lbm_fail("the north row has both a pressure and a velocity condition");
    if ((p->velocity_s_in || p->velocity_s_out) && (p->pressure_s_in[0] || p->pressure_s_out[0]))
      return lbm_fail("the south row has both a pressure and a velocity condition");
  } else if (p->velocity_n_in || p->velocity_n_out || p->velocity_s_in || p->velocity_s_out) {
    return lbm_fail("velocity rows are 2-D conditions (3-D: zface_flow)");
  }
  lbm_handle *h = new lbm_handle();
  h->p = *p;
  if (p->dims == 2) h->p.nk = 1;
  if (p->dims == 3) {
    // per-face z conditions: explicit kinds, else zface_pressure (both
    // faces) and zface_solute (Dirichlet inlet at z = 0, zero gradient at
    // z = nk-1)
    for (int f = 0; f < 2; f++) {
      const int fk = p->zface_flow[f];
      h->zflow[f] = fk == LBM_ZFACE_PRESSURE ? kZPressure
                    : fk == LBM_ZFACE_VELOCITY ? kZVelocity
                    : fk == LBM_ZFACE_NONE     ? kZNone
                    : p->zface_pressure        ? kZPressure
                                               : kZNone;
      h->zfval[f] = h->zflow[f] == kZVelocity ? p->rho_A[0] * p->zface_uz[f] : (f ? p->rho_top : p->rho_bot);
      const int ck = p->zface_conc[f];
      h->zconc[f] = p->nsubs != 2                 ? kCNone
                    : ck == LBM_ZCONC_DIRICHLET ? kCDirichlet
                    : ck == LBM_ZCONC_ZEROGRAD  ? kCZeroGrad
                    : ck == LBM_ZCONC_NONE      ? kCNone
                    : p->zface_solute           ? (f ? kCZeroGrad : kCDirichlet)
                                                : kCNone;
      h->zcval[f] = f ? p->c_top : p->c_bot;
    }
  }
  h->D = p->dims;
  h->Q = p->dims == 2 ? 9 : 19;
  h->U = p->dims == 2 ? 3 : 9;
  h->L = p->dims == 2 ? p->nj : p->nk;
  h->N = N;
  h->E = p->dims == 2 ? p->ni : (long)p->ni * p->nj;
  h->st_bytes = p->storage == LBM_STORE_F64 ? 8 : (p->storage == LBM_STORE_F32 ? 4 : 2);
  // the kernels' 64-bit offset variants normally run only above 4 GiB of
  // populations; LBM_OFFSET64=1 selects them on any lattice
  const char *o64 = getenv("LBM_OFFSET64");
  h->off64 = o64 && o64[0] == '1' && o64[1] == 0;
  for (int s = 0; s < 2; s++) {
    if (p->pressure_n_series[s] && p->pressure_n_series_len[s] > 0)
      h->series[s].assign(p->pressure_n_series[s], p->pressure_n_series[s] + p->pressure_n_series_len[s]);
    if (p->pressure_s_series[s] && p->pressure_s_series_len[s] > 0)
      h->series[2 + s].assign(p->pressure_s_series[s], p->pressure_s_series[s] + p->pressure_s_series_len[s]);
    h->p.pressure_n_series[s] = nullptr;
    h->p.pressure_s_series[s] = nullptr;
  }
  h->p.c_bot_map = nullptr;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return lbm_fail("hipStreamCreate failed");
  }
  if (scratch_on_device(&h->scratch)) {
    const std::string why = g_err;
    lbm_destroy(h);
    return lbm_fail("%s", why.c_str());
  }
  {
    // LDS per CU, for the odd half-step's cap on resident workgroups (none
    // when the device does not report it)
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess)
      h->lds_cu = v;
  }
  if (p->c_bot_map && p->dims == 3) {
    size_t E = (size_t)h->E;
    if (hipMalloc(&h->cmap, E * sizeof(double)) != hipSuccess ||
        hipMemcpy(h->cmap, p->c_bot_map, E * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      lbm_destroy(h);
      return lbm_fail("inlet map upload failed");
    }
  }
  if (hipMalloc(&h->red, sizeof(double) * 12 * kRedBlocks) != hipSuccess ||
      hipMalloc(&h->flux_dev, sizeof(double) * (2 + 2 * kFluxBlocks)) != hipSuccess) {
    lbm_destroy(h);
    return lbm_fail("hipMalloc failed");
  }
  *out = h;
  return 0;
}

extern "C" int lbm_destroy(lbm_handle *h) {
  if (!h) return 0;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  free_lattice(h);
  if (h->cmap) (void)hipFree(h->cmap);
  if (h->red) (void)hipFree(h->red);
  if (h->flux_dev) (void)hipFree(h->flux_dev);
  for (auto &pp : h->pending) {
    (void)hipEventDestroy(pp.a);
    (void)hipEventDestroy(pp.b);
  }
  for (auto e : h->pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

// The body of lbm_set_solids on a handle whose lattice has been released;
// on failure the caller releases whatever it built.
static int build_lattice(lbm_handle *h, const uint8_t *mask) {
  const long N = h->N, E = h->E;
  const int L = h->L;
  h->fid.assign(N, -1);
  h->node_host.clear();
  for (long n = 0; n < N; n++)
    if (mask[n] != 1) {
      h->fid[n] = (int)h->node_host.size();
      h->node_host.push_back((int)n);
    }
  const int nf = (int)h->node_host.size();
  h->nf = nf;
  if (nf == 0) return lbm_fail("lattice has no fluid node");
  long slots = nf + (h->p.halo_mode ? 2 * E : 0);
  h->tiles = (slots + kTile - 1) / kTile;
  long nblk = blocks_for(nf, kBlock);
  h->nblk8 = (int)((nblk + 7) / 8);
  h->anblk8 = (int)((blocks_for(nf, kAdjBlock) + 7) / 8);
  if (h->D == 3)
    if (int rc = build_band_maps(h)) return rc;
  const int Q = h->Q, ns = h->p.nsubs;
  const size_t elems = (size_t)h->tiles * ns * Q * kTile;
  if (elems >= (1ull << 32)) return lbm_fail("lattice too large for 32-bit population offsets");

  HIP_OK(hipMalloc(&h->F, elems * h->st_bytes));
  HIP_OK(hipMemsetAsync(h->F, 0, elems * h->st_bytes, h->stream));
  HIP_OK(hipMalloc(&h->mv, (size_t)ns * (1 + h->D) * nf * sizeof(double)));
  HIP_OK(hipMemsetAsync(h->mv, 0, (size_t)ns * (1 + h->D) * nf * sizeof(double), h->stream));
  if (upload_ints(&h->node, h->node_host)) return 1;
  // fluid bitmask words with their compact base index, one row at a time
  // (LX <= kMaxLX, checked by lbm_create, keeps their 16-bit rank offsets
  // in range)
  const int nx = h->p.ni;
  const long rows = N / nx;
  h->W = (nx + 63) / 64;
  std::vector<RowWord> rw((size_t)rows * h->W);
  fill_row_words(h->fid.data(), rows, nx, rw.data());
  HIP_OK(hipMalloc(&h->rw, rw.size() * sizeof(RowWord)));
  HIP_OK(hipMemcpy(h->rw, rw.data(), rw.size() * sizeof(RowWord), hipMemcpyHostToDevice));

  // faces along the last axis and the flux plane
  std::vector<int> lo, hi, fl;
  for (long m = 0; m < E; m++) {
    if (h->fid[m] >= 0) lo.push_back(h->fid[m]);
    if (h->fid[(L - 1) * E + m] >= 0) hi.push_back(h->fid[(L - 1) * E + m]);
    if (h->p.flux_plane >= 0 && h->p.flux_plane < L && h->fid[(long)h->p.flux_plane * E + m] >= 0)
      fl.push_back(h->fid[(long)h->p.flux_plane * E + m]);
  }
  h->n_lo = (int)lo.size();
  h->n_hi = (int)hi.size();
  h->zb_ok = L >= 2;
  for (size_t q = 0; q < lo.size() && h->zb_ok; q++) h->zb_ok = lo[q] == (int)q;
  for (size_t q = 0; q < hi.size() && h->zb_ok; q++) h->zb_ok = hi[q] == h->nf - h->n_hi + (int)q;
  h->n_flux = h->p.nsubs == 2 ? (int)fl.size() : 0;
  if (upload_ints(&h->face_lo, lo) || upload_ints(&h->face_hi, hi) || upload_ints(&h->face_flux, fl)) return 1;

  if (h->p.halo_mode) {
    const int ni = h->p.ni, nj = h->D == 3 ? h->p.nj : 1, pairs = (Q - h->U) / 2;
    std::vector<int> pl(E), ph(E), tl((size_t)pairs * E), th((size_t)pairs * E);
    for (long m = 0; m < E; m++) {
      pl[m] = h->fid[m];
      ph[m] = h->fid[(L - 1) * E + m];
    }
    for (int q = 0; q < pairs; q++) {
      int a = h->U + 2 * q;
      for (long m = 0; m < E; m++) {
        int i = (int)(m % ni), j = (int)(m / ni);
        int iu = (i + CX(a) + ni) % ni, ju = h->D == 3 ? (j + CY(a) + nj) % nj : 0;
        int id = (i - CX(a) + ni) % ni, jd = h->D == 3 ? (j - CY(a) + nj) % nj : 0;
        bool up_fluid = h->fid[iu + (long)ni * ju] >= 0;               // pusher above the last plane
        bool dn_fluid = h->fid[(L - 1) * E + id + (long)ni * jd] >= 0; // pusher below the first plane
        th[(size_t)q * E + m] = up_fluid ? ph[m] : -1;
        tl[(size_t)q * E + m] = dn_fluid ? pl[m] : -1;
      }
    }
    if (upload_ints(&h->plane_lo, pl) || upload_ints(&h->plane_hi, ph) || upload_ints(&h->tgt_lo, tl) ||
        upload_ints(&h->tgt_hi, th))
      return 1;
    size_t hb = (size_t)h->p.nsubs * pairs * E * h->st_bytes;
    HIP_OK(hipMalloc(&h->halo_pos, hb));
    HIP_OK(hipMalloc(&h->halo_neg, hb));
  }
  HIP_OK(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int lbm_set_solids(lbm_handle *h, const uint8_t *mask) {
  if (!h || !mask) return lbm_fail("lbm_set_solids: null argument");
  HIP_OK(hipStreamSynchronize(h->stream));
  free_lattice(h);
  if (int rc = build_lattice(h, mask)) {
    // no half-built lattice: the handle rejects steps until a later call succeeds
    (void)hipStreamSynchronize(h->stream);
    free_lattice(h);
    return rc;
  }
  return 0;
}

static int transfer_f(lbm_handle *h, int subs, double *f1d, int up) {
  if (need_lattice(h)) return 1;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  if (!f1d) return lbm_fail("null population array");
  const int Q = h->Q, nf = h->nf;
  std::vector<double> st((size_t)Q * nf);
  double *dstage = nullptr;
  HIP_OK(hipMalloc(&dstage, st.size() * sizeof(double)));
  if (up) {
    for (int a = 0; a < Q; a++)
      for (int x = 0; x < nf; x++) st[(size_t)a * nf + x] = f1d[(long)a * h->N + h->node_host[x]];
    HIP_OK(hipMemcpy(dstage, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  int rc = dispatch(h, [&]<int QQ, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    dim3 g(blocks_for(nf)), b(256);
    if (up) {
      if (subs == 0) hipLaunchKernelGGL((k_encode<QQ, NS, ST, CT, 0>), g, b, 0, h->stream, A, dstage);
      else if constexpr (NS == 2) hipLaunchKernelGGL((k_encode<QQ, NS, ST, CT, 1>), g, b, 0, h->stream, A, dstage);
    } else {
      if (subs == 0) hipLaunchKernelGGL((k_decode<QQ, NS, ST, CT, 0>), g, b, 0, h->stream, A, dstage);
      else if constexpr (NS == 2) hipLaunchKernelGGL((k_decode<QQ, NS, ST, CT, 1>), g, b, 0, h->stream, A, dstage);
    }
    HIP_OK(hipGetLastError());
    return 0;
  });
  if (!rc) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = lbm_fail("transfer: %s", hipGetErrorString(e));
  }
  if (!rc && !up) {
    hipError_t e = hipMemcpy(st.data(), dstage, st.size() * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = lbm_fail("transfer: %s", hipGetErrorString(e));
    else {
      memset(f1d, 0, sizeof(double) * Q * h->N);
      for (int a = 0; a < Q; a++)
        for (int x = 0; x < nf; x++) f1d[(long)a * h->N + h->node_host[x]] = st[(size_t)a * nf + x];
    }
  }
  (void)hipFree(dstage);
  return rc;
}

extern "C" int lbm_set_ns(lbm_handle *h, const double *ns) {
  if (need_lattice(h)) return 1;
  HIP_OK(hipStreamSynchronize(h->stream));
  drop_chunks(h);
  if (h->ns) {
    (void)hipFree(h->ns);
    h->ns = nullptr;
  }
  if (!ns) return 0;
  std::vector<double> c((size_t)h->nf);
  for (int x = 0; x < h->nf; x++) {
    const double v = ns[h->node_host[x]];
    if (!(v >= 0.0 && v <= 1.0)) return lbm_fail("ns at node %d is %g, outside [0, 1]", h->node_host[x], v);
    c[x] = v;
  }
  HIP_OK(hipMalloc(&h->ns, c.size() * sizeof(double)));
  HIP_OK(hipMemcpy(h->ns, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int lbm_upload_f(lbm_handle *h, int subs, const double *f1d) {
  return transfer_f(h, subs, const_cast<double *>(f1d), 1);
}

extern "C" int lbm_download_f(lbm_handle *h, int subs, double *f1d) { return transfer_f(h, subs, f1d, 0); }

extern "C" int lbm_download_f_noalloc(lbm_handle *h, int subs, double *f1d) {
  if (need_lattice(h)) return 1;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  if (!f1d) return lbm_fail("null population array");
  const long elems = (long)h->Q * h->N;
  if (elems > kScratchDoubles)
    return lbm_fail("lbm_download_f_noalloc: Q x N = %ld doubles exceed the %ld of its static staging array "
                    "(use lbm_download_f)",
                    elems, kScratchDoubles);
  double *stage = h->scratch;
  int rc = dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    const dim3 g(blocks_for(h->N)), b(256);
    if (subs == 0) hipLaunchKernelGGL((k_decode_dense<Q, NS, ST, CT, 0>), g, b, 0, h->stream, A, stage, (int)h->N);
    else if constexpr (NS == 2)
      hipLaunchKernelGGL((k_decode_dense<Q, NS, ST, CT, 1>), g, b, 0, h->stream, A, stage, (int)h->N);
    HIP_OK(hipGetLastError());
    return 0;
  });
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(f1d, stage, (size_t)elems * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int lbm_init_rest(lbm_handle *h, const double *c0) {
  if (need_lattice(h)) return 1;
  double *dc = nullptr;
  if (c0 && h->p.nsubs == 2) {
    std::vector<double> cc(h->nf);
    for (int x = 0; x < h->nf; x++) cc[x] = c0[h->node_host[x]];
    HIP_OK(hipMalloc(&dc, cc.size() * sizeof(double)));
    HIP_OK(hipMemcpy(dc, cc.data(), cc.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  int rc = dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    hipLaunchKernelGGL((k_init_rest<Q, NS, ST, CT>), dim3(blocks_for(h->nf)), dim3(256), 0, h->stream, A, dc);
    HIP_OK(hipGetLastError());
    return 0;
  });
  if (!rc) rc = sync(h);
  if (dc) (void)hipFree(dc);
  return rc;
}

extern "C" int lbm_set_time(lbm_handle *h, int t) {
  if (!h) return lbm_fail("null handle");
  h->time = t;
  return 0;
}

extern "C" int lbm_get_time(const lbm_handle *h, int *t) {
  if (!h || !t) return lbm_fail("null argument");
  *t = h->time;
  return 0;
}

extern "C" int lbm_step(lbm_handle *h, int write_macro) {
  if (need_lattice(h)) return 1;
  if (do_step(h, write_macro)) return 1;
  return sync(h);
}

namespace {

constexpr int kChunk = 64;  // steps per captured chunk

// Whether the steps t1..t2 launch the same kernels with the same arguments
// as any other stretch of steps of the same parity: no profiling events, no
// halo exchange, 2-D boundary densities without a pressure series or a ramp
// still running, and the solute inside (or outside) its time window
// throughout.
bool chunkable(const lbm_handle *h, int t1, int t2) {
  const lbm_params &p = h->p;
  if (h->profile || p.halo_mode || h->chunk_off) return false;
  if (h->D == 2) {
    for (const auto &s : h->series)
      if (!s.empty()) return false;
    if (p.bc_ramp_start >= 0 && p.bc_ramp_stop >= 0 && t1 < p.bc_ramp_stop) return false;
  }
  if (p.nsubs == 2) {
    const bool in1 = t1 >= p.sigma_t_on && t1 <= p.sigma_t_off;
    const bool in2 = t2 >= p.sigma_t_on && t2 <= p.sigma_t_off;
    if (in1 != in2 || (!in1 && !(t2 < p.sigma_t_on || t1 > p.sigma_t_off))) return false;
  }
  return true;
}

// The HIP graph of kChunk steps from the handle's current time and state
// array, captured from the step launches on first use and kept while the
// geometry, the Walsh field and the state array stay the same; null when
// capture is not possible (the steps are then launched one by one).  One
// graph per starting parity and per solute window state: chunkable() lets
// a run through only when all its steps are on the same side of the
// window, and the graph captured on that side is the one replayed.
hipGraphExec_t chunk_for(lbm_handle *h) {
  const int par = (h->time + 1) & 1;
  const lbm_params &p = h->p;
  const int win = p.nsubs == 2 && h->time + 1 >= p.sigma_t_on && h->time + 1 <= p.sigma_t_off;
  hipGraphExec_t &slot = h->chunk[par][win];
  if (slot && h->chunk_gen[par][win] == h->gen && h->chunk_F[par][win] == h->F) return slot;
  if (slot) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipGraphExecDestroy(slot);
  }
  slot = nullptr;
  const int t_save = h->time;
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    h->chunk_off = true;
    return nullptr;
  }
  int rc = 0;
  for (int i = 0; i < kChunk && !rc; i++) rc = do_step(h, 0);
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  h->time = t_save;
  hipGraphExec_t x = nullptr;
  if (rc || e != hipSuccess || !g || hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    h->chunk_off = true;
    return nullptr;
  }
  (void)hipGraphDestroy(g);
  slot = x;
  h->chunk_gen[par][win] = h->gen;
  h->chunk_F[par][win] = h->F;
  return x;
}

}  // namespace

extern "C" int lbm_run(lbm_handle *h, int nsteps, int write_macro_last) {
  if (need_lattice(h)) return 1;
  int q = 0;
  // the steps before the last in captured chunks where their launches repeat
  if (nsteps - 1 >= kChunk && chunkable(h, h->time + 1, h->time + nsteps - 1)) {
    while (nsteps - 1 - q >= kChunk) {
      hipGraphExec_t g = chunk_for(h);
      if (!g) break;
      HIP_OK(hipGraphLaunch(g, h->stream));
      h->time += kChunk;
      q += kChunk;
    }
  }
  for (; q < nsteps; q++)
    if (do_step(h, q == nsteps - 1 ? write_macro_last : 0)) return 1;
  return sync(h);
}

extern "C" int lbm_step_odd(lbm_handle *h, int t, int write_macro) {
  if (need_lattice(h)) return 1;
  if (launch_half(h, 1, t, write_macro)) return 1;
  return sync(h);
}

extern "C" int lbm_step_even(lbm_handle *h, int t, int write_macro) {
  if (need_lattice(h)) return 1;
  if (launch_half(h, 0, t, write_macro)) return 1;
  return sync(h);
}

extern "C" int lbm_apply_bcs(lbm_handle *h, int phase, int t) {
  if (need_lattice(h)) return 1;
  if (phase != 1 && phase != 2) return lbm_fail("phase must be 1 or 2");
  if (launch_bcs(h, phase, t)) return 1;
  return sync(h);
}

extern "C" int lbm_halo_bytes(const lbm_handle *h, long *bytes) {
  if (!h || !bytes) return lbm_fail("null argument");
  *bytes = (long)h->p.nsubs * ((h->Q - h->U) / 2) * h->E * (long)h->st_bytes;
  return 0;
}

extern "C" int lbm_halo_pack(lbm_handle *h, int phase, void *pos, void *neg) {
  if (need_lattice(h)) return 1;
  if (!pos || !neg) return lbm_fail("null halo buffer");
  if (halo_pack_unpack(h, phase, 1, pos, neg)) return 1;
  return sync(h);
}

extern "C" int lbm_halo_unpack(lbm_handle *h, int phase, const void *pos, const void *neg) {
  if (need_lattice(h)) return 1;
  if (!pos || !neg) return lbm_fail("null halo buffer");
  if (halo_pack_unpack(h, phase, 0, const_cast<void *>(pos), const_cast<void *>(neg))) return 1;
  return sync(h);
}

extern "C" int lbm_halo_periodic(lbm_handle *h, int phase) {
  if (need_lattice(h)) return 1;
  if (halo_pack_unpack(h, phase, 1, h->halo_pos, h->halo_neg)) return 1;
  if (halo_pack_unpack(h, phase, 0, h->halo_pos, h->halo_neg)) return 1;
  return sync(h);
}

extern "C" int lbm_download_macro(lbm_handle *h, int subs, double *rho, double *ux, double *uy, double *uz) {
  if (need_lattice(h)) return 1;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  HIP_OK(hipStreamSynchronize(h->stream));
  const int nf = h->nf, D = h->D;
  std::vector<double> buf((size_t)(1 + D) * nf);
  HIP_OK(hipMemcpy(buf.data(), h->mv + (long)subs * (1 + D) * nf, buf.size() * sizeof(double),
                   hipMemcpyDeviceToHost));
  double *dst[4] = {rho, ux, uy, uz};
  for (int c = 0; c < 4; c++) {
    if (!dst[c]) continue;
    memset(dst[c], 0, sizeof(double) * h->N);
    if (c > D) continue;
    for (int x = 0; x < nf; x++) dst[c][h->node_host[x]] = buf[(size_t)c * nf + x];
  }
  return 0;
}

extern "C" int lbm_sample_macro(lbm_handle *h, int subs, int npts, const long *nodes, double *out) {
  if (need_lattice(h)) return 1;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  if (npts < 0 || (npts && (!nodes || !out))) return lbm_fail("bad sample arguments");
  HIP_OK(hipStreamSynchronize(h->stream));
  const int nf = h->nf, D = h->D;
  const double *mv = h->mv + (long)subs * (1 + D) * nf;
  for (int q = 0; q < npts; q++) {
    if (nodes[q] < 0 || nodes[q] >= h->N) return lbm_fail("sample node %ld outside the lattice", nodes[q]);
    const int x = h->fid[nodes[q]];
    for (int c = 0; c <= D; c++) {
      double *dst = out + (long)q * (1 + D) + c;
      if (x < 0)
        *dst = 0.0;
      else
        HIP_OK(hipMemcpy(dst, mv + (long)c * nf + x, sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  return 0;
}

extern "C" int lbm_upload_macro(lbm_handle *h, int subs, const double *rho, const double *ux, const double *uy,
                                const double *uz) {
  if (need_lattice(h)) return 1;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  HIP_OK(hipStreamSynchronize(h->stream));
  const int nf = h->nf, D = h->D;
  double *mv = h->mv + (long)subs * (1 + D) * nf;
  const double *src[4] = {rho, ux, uy, uz};
  std::vector<double> buf((size_t)nf);
  for (int c = 0; c <= D; c++) {
    if (!src[c]) continue;
    for (int x = 0; x < nf; x++) buf[x] = src[c][h->node_host[x]];
    HIP_OK(hipMemcpy(mv + (long)c * nf, buf.data(), (size_t)nf * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

extern "C" int lbm_reduce(lbm_handle *h, int kind, int subs, double *out) {
  if (need_lattice(h)) return 1;
  if (!out) return lbm_fail("null output");
  if (kind == LBM_REDUCE_SUMS) subs = 0;
  if (subs < 0 || subs >= h->p.nsubs) return lbm_fail("bad subs %d", subs);
  const int D = h->D, nf = h->nf;
  double res[2][12];
  int nsub_red = kind == LBM_REDUCE_SUMS ? h->p.nsubs : 1;
  for (int r = 0; r < nsub_red; r++) {
    int s = kind == LBM_REDUCE_SUMS ? r : subs;
    const double *mv = h->mv + (long)s * (1 + D) * nf;
    double off = (s == 0) ? h->p.rho_A[0] : 0.0;
    if (D == 2)
      hipLaunchKernelGGL(k_reduce_mv<2>, dim3(kRedBlocks), dim3(256), 0, h->stream, mv, nf, off, h->red);
    else
      hipLaunchKernelGGL(k_reduce_mv<3>, dim3(kRedBlocks), dim3(256), 0, h->stream, mv, nf, off, h->red);
    HIP_OK(hipGetLastError());
    std::vector<double> part((size_t)12 * kRedBlocks);
    HIP_OK(hipMemcpyAsync(part.data(), h->red, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_OK(hipStreamSynchronize(h->stream));
    double v[12];
    for (int q = 0; q < 12; q++) v[q] = part[q];
    for (int b = 1; b < kRedBlocks; b++)
      for (int q = 0; q < 12; q++) {
        double x = part[(size_t)b * 12 + q];
        v[q] = q < 4 ? v[q] + x : (q < 8 ? (x > v[q] ? x : v[q]) : (x < v[q] ? x : v[q]));
      }
    for (int q = 0; q < 12; q++) res[r][q] = v[q];
  }
  if (kind == LBM_REDUCE_REF) {
    double *v = res[0];
    double off = subs == 0 ? h->p.rho_A[0] : 0.0;
    out[0] = v[4];
    out[1] = v[8];
    out[2] = off + v[0] / nf;
    for (int c = 0; c < 3; c++) {
      out[3 + c] = c < D ? v[5 + c] : 0.0;
      out[6 + c] = c < D ? v[9 + c] : 0.0;
      out[9 + c] = c < D ? v[1 + c] / nf : 0.0;
    }
    return 0;
  }
  if (kind == LBM_REDUCE_SUMS) {
    out[0] = res[0][0];
    out[1] = res[0][1];
    out[2] = res[0][2];
    out[3] = D == 3 ? res[0][3] : 0.0;
    out[4] = (double)nf;
    out[5] = h->p.nsubs == 2 ? res[1][0] : 0.0;
    return 0;
  }
  return lbm_fail("unknown reduction %d", kind);
}

extern "C" int lbm_last_flux(lbm_handle *h, double *flux, double *mass) {
  if (need_lattice(h)) return 1;
  if (!flux || !mass) return lbm_fail("null output");
  if (!h->flux_valid) return lbm_fail("no plane flux recorded (flux_plane < 0 or no solute)");
  double v[2];
  HIP_OK(hipMemcpyAsync(v, h->flux_dev, sizeof v, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  *flux = v[0];
  *mass = v[1];
  return 0;
}

extern "C" int lbm_check(lbm_handle *h, const double *mass_ref, double mass_tol, lbm_health *out) {
  if (need_lattice(h)) return 1;
  // an error left on the device by an earlier launch is a failure too
  HIP_OK(hipStreamSynchronize(h->stream));
  HIP_OK(hipGetLastError());
  int rc = dispatch(h, [&]<int Q, int NS, typename ST, typename CT>() -> int {
    StepArgs<ST, CT> A = make_args<ST, CT>(h);
    hipLaunchKernelGGL((k_check<Q, NS, ST, CT>), dim3(kRedBlocks), dim3(256), 0, h->stream, A, h->p.rho_A[0],
                       h->red);
    HIP_OK(hipGetLastError());
    return 0;
  });
  if (rc) return rc;
  std::vector<double> part((size_t)kCheckVals * kRedBlocks);
  HIP_OK(hipMemcpyAsync(part.data(), h->red, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  double v[kCheckVals];
  for (int q = 0; q < kCheckVals; q++) v[q] = part[q];
  for (int b = 1; b < kRedBlocks; b++)
    for (int q = 0; q < kCheckVals; q++) {
      const double x = part[(size_t)b * kCheckVals + q];
      v[q] = (q >= 2 && q <= 4) ? fmin(v[q], x) : v[q] + x;
    }
  lbm_health H;
  H.mass[0] = v[0];
  H.mass[1] = h->p.nsubs == 2 ? v[1] : 0.0;
  H.min_rho = v[2];
  H.min_pop[0] = v[3];
  H.min_pop[1] = h->p.nsubs == 2 ? v[4] : 0.0;
  for (int s = 0; s < 2; s++) {
    H.nonfinite[s] = (long)v[5 + s];
    H.negative[s] = (long)v[7 + s];
  }
  if (out) *out = H;
  static const char *name[2] = {"flow", "solute"};
  for (int s = 0; s < h->p.nsubs; s++)
    if (H.nonfinite[s])
      return lbm_fail("health check: %ld %s populations are NaN or infinite", H.nonfinite[s], name[s]);
  if (H.negative[0] || !(H.min_rho > 0.0))
    return lbm_fail("health check: %ld negative flow populations (least %g), least density %g", H.negative[0],
                    H.min_pop[0], H.min_rho);
  if (mass_ref) {
    for (int s = 0; s < h->p.nsubs; s++) {
      if (mass_ref[s] != mass_ref[s]) continue;  // NaN: not checked
      const double M = s == 0 ? h->p.rho_A[0] * (double)h->nf + fabs(mass_ref[0]) : fabs(mass_ref[1]);
      const double d = fabs(H.mass[s] - mass_ref[s]);
      if (!(d <= mass_tol * M))
        return lbm_fail("health check: %s mass %.17g drifted from %.17g (relative %.3g > %g)", name[s], H.mass[s],
                        mass_ref[s], M > 0.0 ? d / M : d, mass_tol);
    }
  }
  return 0;
}

extern "C" int lbm_info(const lbm_handle *h, long *n_nodes, long *n_fluid, long *row_stride) {
  if (!h) return lbm_fail("null handle");
  if (n_nodes) *n_nodes = h->N;
  if (n_fluid) *n_fluid = h->nf;
  if (row_stride) *row_stride = h->tiles * kTile;
  return 0;
}

extern "C" int lbm_profile_enable(lbm_handle *h, int on) {
  if (!h) return lbm_fail("null handle");
  h->profile = on;
  return 0;
}

extern "C" int lbm_profile_read(lbm_handle *h, double *ms, int *count) {
  if (!h) return lbm_fail("null handle");
  HIP_OK(hipStreamSynchronize(h->stream));
  for (auto &pp : h->pending) {
    float t = 0.f;
    HIP_OK(hipEventElapsedTime(&t, pp.a, pp.b));
    h->prof_ms[pp.cls] += t;
    h->prof_n[pp.cls] += 1;
    h->pool.push_back(pp.a);
    h->pool.push_back(pp.b);
  }
  h->pending.clear();
  for (int c = 0; c < LBM_PROFILE_CLASSES; c++) {
    if (ms) ms[c] = h->prof_ms[c];
    if (count) count[c] = h->prof_n[c];
    h->prof_ms[c] = 0;
    h->prof_n[c] = 0;
  }
  return 0;
}

extern "C" int lbm_copy_bandwidth(long nbytes, int reps, double *gbps) {
  if (nbytes < 8 * 256 || reps < 1 || !gbps) return lbm_fail("lbm_copy_bandwidth: bad arguments");
  long n = nbytes / 8;
  double *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.f;
  if (hipMalloc(&a, n * 8) != hipSuccess || hipMalloc(&b, n * 8) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = lbm_fail("lbm_copy_bandwidth: allocation failed");
  } else {
    dim3 g(blocks_for(n)), bl(256);
    (void)hipMemsetAsync(a, 0, n * 8, st);
    hipLaunchKernelGGL(k_copy_probe, g, bl, 0, st, a, b, n);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, g, bl, 0, st, a, b, n);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f)
      rc = lbm_fail("lbm_copy_bandwidth: timing failed");
    else
      *gbps = 2.0 * (double)n * 8.0 * reps / (ms * 1e-3) / 1e9;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ---- objective and discrete adjoint (D3Q19 flow + solute, FP64) ----------

static int adjoint_supported(const lbm_handle *h) {
  const lbm_params &p = h->p;
  if (h->Q != 19 || p.nsubs != 2) return lbm_fail("the adjoint needs D3Q19 flow + solute");
  if (p.storage != LBM_STORE_F64 || p.compute != LBM_COMPUTE_F64) return lbm_fail("the adjoint runs in FP64");
  if (p.halo_mode) return lbm_fail("the adjoint needs the periodic (halo_mode = 0) lattice");
  if (p.sink != 0.0) return lbm_fail("the adjoint needs sink = 0");
  if (h->obj.kind == LBM_OBJ_OUTLET_FLUX && (p.flux_plane < 0 || p.flux_plane >= p.nk))
    return lbm_fail("the objective needs flux_plane");
  if (h->ns) return lbm_fail("the adjoint does not cover partial bounce-back (lbm_set_ns)");
  // the transposed face conditions: pressure on both z faces or on none; a
  // Dirichlet inlet at z = 0 with a zero-gradient outlet, or no solute
  // condition
  if (h->zflow[0] != h->zflow[1] || h->zflow[0] == kZVelocity)
    return lbm_fail("the adjoint needs the z-face pressure condition on both faces or on none");
  if (!((h->zconc[0] == kCDirichlet && h->zconc[1] == kCZeroGrad) ||
        (h->zconc[0] == kCNone && h->zconc[1] == kCNone)))
    return lbm_fail("the adjoint needs a Dirichlet solute inlet at z = 0 with a zero-gradient outlet, or none");
  return 0;
}

static size_t state_bytes(const lbm_handle *h) {
  return (size_t)h->tiles * h->p.nsubs * h->Q * kTile * h->st_bytes;
}

extern "C" int lbm_objective_observations(lbm_handle *h, int n, const long *node, const int *step,
                                          const double *target, const double *weight) {
  if (need_lattice(h)) return 1;
  if (n < 0 || (n && (!node || !step || !target))) return lbm_fail("bad observation arguments");
  std::vector<int> ord(n);
  for (int q = 0; q < n; q++) {
    if (node[q] < 0 || node[q] >= h->N || h->fid[node[q]] < 0)
      return lbm_fail("observation %d: node %ld is not a fluid node of the lattice", q, node[q]);
    if (step[q] < 1) return lbm_fail("observation %d: step %d (steps count from 1)", q, step[q]);
    if (weight && !(weight[q] >= 0.0)) return lbm_fail("observation %d: weight must be >= 0", q);
    ord[q] = q;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return step[a] < step[b]; });
  auto &o = h->obj;
  o.valid = 0;
  o.o_step.assign(n, 0);
  o.o_node.assign(n, 0);
  o.o_target.assign(n, 0.0);
  o.o_weight.assign(n, 1.0);
  for (int q = 0; q < n; q++) {
    o.o_step[q] = step[ord[q]];
    o.o_node[q] = h->fid[node[ord[q]]];
    o.o_target[q] = target[ord[q]];
    if (weight) o.o_weight[q] = weight[ord[q]];
  }
  for (void *p : {(void *)o.o_node_d, (void *)o.o_c}) if (p) (void)hipFree(p);
  o.o_node_d = nullptr;
  o.o_c = nullptr;
  if (n) {
    HIP_OK(hipMalloc(&o.o_node_d, (size_t)n * sizeof(int)));
    HIP_OK(hipMemcpy(o.o_node_d, o.o_node.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&o.o_c, (size_t)n * sizeof(double)));
  }
  return 0;
}

// permeability objective: rho_A nu <u_z,phys> / G over the window, with
// nu = (tau - 1/2)/3, u_phys = u_eq - (tau - 1/2) g (superficial average over
// all N nodes) and G = (rho_bot - rho_top)/(3 (LZ - 1)) (z-face pressure
// conditions) + rho_A g_z, the driving gradient
static double perm_drive(const lbm_handle *h) {
  const lbm_params &p = h->p;
  const double dp = p.zface_pressure ? (p.rho_bot - p.rho_top) / (3.0 * (p.nk - 1)) : 0.0;
  return dp + p.rho_A[0] * p.gaccel[2];
}

extern "C" int lbm_objective_select(lbm_handle *h, int kind) {
  if (!h) return lbm_fail("null handle");
  if (kind != LBM_OBJ_OUTLET_FLUX && kind != LBM_OBJ_PERMEABILITY) return lbm_fail("unknown objective %d", kind);
  h->obj.kind = kind;
  h->obj.valid = 0;
  return 0;
}

extern "C" int lbm_objective_keep_states(lbm_handle *h, size_t max_bytes) {
  if (!h) return lbm_fail("null handle");
  h->obj.keep_bytes = max_bytes;
  return 0;
}

extern "C" int lbm_objective_info(const lbm_handle *h, int *kept_steps, int *checkpoints) {
  if (!h || !kept_steps || !checkpoints) return lbm_fail("null argument");
  *kept_steps = (int)h->obj.kept.size();
  *checkpoints = (int)h->obj.ckpt.size();
  return 0;
}

extern "C" int lbm_objective_run(lbm_handle *h, int nsteps, int window, int every, double *J) {
  if (need_lattice(h) || adjoint_supported(h)) return 1;
  const int nobs = (int)h->obj.o_step.size();
  if (nsteps < 1 || window < 0 || (window == 0 && nobs == 0) || window > nsteps || every < 1 || !J)
    return lbm_fail("bad objective arguments");
  if (nobs && h->obj.o_step.back() > nsteps)
    return lbm_fail("an observation lies after the run's last step (%d > %d)", h->obj.o_step.back(), nsteps);
  const int t0 = h->time;
  if (t0 + 1 < h->p.sigma_t_on || t0 + nsteps > h->p.sigma_t_off)
    return lbm_fail("the solute must be active for the whole objective run");
  const bool perm = h->obj.kind == LBM_OBJ_PERMEABILITY;
  if (perm && window > 0 && !(perm_drive(h) != 0.0))
    return lbm_fail("the permeability needs a driving gradient (z-face pressure difference or g_z)");
  h->obj.valid = 0;
  const size_t sb = state_bytes(h);
  // the last P steps keep their states (P = nsteps - m * every, the fewest
  // segments m left to replay that the budget allows); the m segments
  // before them keep a checkpoint each
  long P = 0;
  if (h->obj.keep_bytes >= sb) {
    const long cap = (long)(h->obj.keep_bytes / sb);
    // The most steps (nsteps - m * every, m = 0, 1, ...) the budget holds
    // whose states fit in device memory together with the checkpoints of the
    // m replayed segments, the adjoint state and, when m > 0, the states of
    // one replayed segment (every + 1, the first kept buffers when there are
    // that many), and 1 GiB besides; memory these buffers already hold
    // counts as available.
    size_t free_b = 0, total_b = 0;
    HIP_OK(hipMemGetInfo(&free_b, &total_b));
    const double held = (double)(h->obj.kept.size() + h->obj.ckpt.size() + h->obj.seg.size()) * (double)sb;
    const double avail = (double)free_b + held - (double)(1ull << 30);
    const long needL = h->obj.L ? 0 : 1;
    for (long m = 0; m * every < nsteps; m++) {
      const long p = nsteps - m * every;
      const long segbuf = m > 0 && p < every + 1 ? every + 1 : 0;
      if (p <= cap && (double)(p + m + segbuf + needL) * (double)sb <= avail) {
        P = p;
        break;
      }
    }
  }
  const int kept_from = nsteps - (int)P;
  const size_t nck = (size_t)((kept_from + every - 1) / every);  // kept_from: nsteps or a multiple of every
  // With every + 1 kept states or more and segments to replay, the sweep
  // replays them into the first kept buffers (it sweeps the kept segments
  // first) instead of segment buffers of its own; the run then supports one
  // sweep.
  const bool kept_replay = nck > 0 && P >= every + 1;
  // release what this run and its sweep will not use before allocating:
  // the sizing above counted it as available
  if (nck == 0 || kept_replay) {
    for (void *q : h->obj.seg) (void)hipFree(q);
    h->obj.seg.clear();
  }
  while ((long)h->obj.kept.size() > P) {
    (void)hipFree(h->obj.kept.back());
    h->obj.kept.pop_back();
  }
  while (h->obj.ckpt.size() > nck) {
    (void)hipFree(h->obj.ckpt.back());
    h->obj.ckpt.pop_back();
  }
  while ((long)h->obj.kept.size() < P) {
    void *c = nullptr;
    HIP_OK(hipMalloc(&c, sb));
    h->obj.kept.push_back(c);
  }
  while (h->obj.ckpt.size() < nck) {
    void *c = nullptr;
    HIP_OK(hipMalloc(&c, sb));
    h->obj.ckpt.push_back(c);
  }
  if (!h->obj.jacc) HIP_OK(hipMalloc(&h->obj.jacc, sizeof(double)));
  HIP_OK(hipMemsetAsync(h->obj.jacc, 0, sizeof(double), h->stream));
  auto &o = h->obj;
  const double *mv_c = h->mv + (long)1 * (1 + h->D) * h->nf;  // solute concentration
  int ob = 0;
  void *const live = h->F;
  // the kept steps run out of place: kept[i] -> kept[i + 1], the last one
  // into the lattice's own array
  struct Restore {
    lbm_handle *h;
    void *F;
    ~Restore() {
      h->F = F;
      h->Fout = nullptr;
    }
  } restore{h, live};
  for (int q = 0; q < nsteps; q++) {
    if (q < kept_from && q % every == 0)
      HIP_OK(hipMemcpyAsync(o.ckpt[q / every], h->F, sb, hipMemcpyDeviceToDevice, h->stream));
    if (q >= kept_from) {
      const int i = q - kept_from;
      if (i == 0) HIP_OK(hipMemcpyAsync(o.kept[0], live, sb, hipMemcpyDeviceToDevice, h->stream));
      h->F = o.kept[i];
      h->Fout = i + 1 < (int)P ? o.kept[i + 1] : live;
    }
    const bool in_window = q >= nsteps - window;
    int oe = ob;
    while (oe < nobs && o.o_step[oe] == q + 1) oe++;
    // an observed step writes the solute concentration (the PEST switch);
    // a window step of the permeability writes the flow velocity
    const int macro = q == nsteps - 1 || (perm && in_window) ? LBM_MACRO_FRAME
                      : oe > ob                               ? LBM_MACRO_PEST
                                                              : LBM_MACRO_NONE;
    if (do_step(h, macro, in_window && !perm ? o.jacc : nullptr)) return 1;
    if (perm && in_window) {
      // sum of u_z over the fluid nodes, added to the running sum
      hipLaunchKernelGGL(k_sum_partial, dim3(kRedBlocks), dim3(256), 0, h->stream, h->mv + (long)3 * h->nf,
                         (long)h->nf, h->red);
      hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, h->stream, h->red, kRedBlocks, o.jacc);
      HIP_OK(hipGetLastError());
    }
    if (oe > ob) {
      hipLaunchKernelGGL(k_obs_gather, dim3(blocks_for(oe - ob)), dim3(256), 0, h->stream, mv_c,
                         o.o_node_d + ob, oe - ob, o.o_c + ob);
      HIP_OK(hipGetLastError());
    }
    ob = oe;
  }
  h->F = live;
  h->Fout = nullptr;
  double acc = 0.0;
  HIP_OK(hipMemcpyAsync(&acc, o.jacc, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  std::vector<double> c(nobs);
  if (nobs) HIP_OK(hipMemcpyAsync(c.data(), o.o_c, (size_t)nobs * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  // observation misfit and its dJ/dc, merged per (step, node)
  double Jobs = 0.0;
  std::vector<int> sn;
  std::vector<double> sv;
  o.s_step.clear();
  o.s_begin.clear();
  for (int q = 0; q < nobs; q++) {
    const double r = c[q] - o.o_target[q];
    Jobs += o.o_weight[q] * r * r;
    const double d = 2.0 * o.o_weight[q] * r;
    if (o.s_step.empty() || o.s_step.back() != o.o_step[q]) {
      o.s_step.push_back(o.o_step[q]);
      o.s_begin.push_back((int)sn.size());
    }
    bool merged = false;
    for (size_t m = o.s_begin.back(); m < sn.size(); m++)
      if (sn[m] == o.o_node[q]) {
        sv[m] += d;
        merged = true;
      }
    if (!merged) {
      sn.push_back(o.o_node[q]);
      sv.push_back(d);
    }
  }
  o.s_begin.push_back((int)sn.size());
  for (void *p : {(void *)o.s_node_d, (void *)o.s_val_d}) if (p) (void)hipFree(p);
  o.s_node_d = nullptr;
  o.s_val_d = nullptr;
  if (!sn.empty()) {
    HIP_OK(hipMalloc(&o.s_node_d, sn.size() * sizeof(int)));
    HIP_OK(hipMalloc(&o.s_val_d, sv.size() * sizeof(double)));
    HIP_OK(hipMemcpy(o.s_node_d, sn.data(), sn.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(o.s_val_d, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  o.t0 = t0;
  o.kept_from = kept_from;
  o.kept_replay = kept_replay ? 1 : 0;
  o.nsteps = nsteps;
  o.window = window;
  o.every = every;
  if (window == 0)
    o.Jwin = 0.0;
  else if (perm) {
    const lbm_params &p = h->p;
    const double nu = (p.tau[0] - 0.5) / 3.0;
    const double su = acc - (double)window * (double)h->nf * (p.tau[0] - 0.5) * p.gaccel[2];
    o.Jwin = p.rho_A[0] * nu * su / ((double)window * (double)h->N * perm_drive(h));
  } else {
    o.Jwin = acc / ((double)window * (double)h->E);
  }
  o.J = o.Jwin + Jobs;
  o.valid = 1;
  *J = o.J;
  return 0;
}

extern "C" int lbm_adjoint_run(lbm_handle *h, double *grad, double *grad_cmap) {
  if (need_lattice(h) || adjoint_supported(h)) return 1;
  if (!grad) return lbm_fail("null gradient output");
  if (!h->obj.valid)
    return lbm_fail("no objective run to differentiate: lbm_objective_run has not been called since the lattice, "
                    "the objective or its observations were set, or since a sweep replayed into its kept states");
  const size_t sb = state_bytes(h);
  const int K = h->obj.every, T = h->obj.nsteps, W_ = h->obj.window, t0 = h->obj.t0;
  const int nseg = (T + K - 1) / K;
  // segments replayed from a checkpoint (kept_from is nsteps or a multiple
  // of K); the others were kept by the objective run
  const int nrep = (h->obj.kept_from + K - 1) / K;
  const int saved_time = h->time;
  void *const saved_F = h->F;  // the replay points h->F at segment buffers
  const int nblk = 8 * h->anblk8;
  int rc = 0;
  auto fail = [&](const char *what, hipError_t e) {
    rc = lbm_fail("adjoint: %s: %s", what, hipGetErrorString(e));
    return rc;
  };
  hipError_t e;
  // workspace: K + 1 segment states (or the first K + 1 kept buffers, free
  // once the kept segments are swept), the adjoint state and the gradient
  // accumulators, kept for the next call
  std::vector<void *> &seg = h->obj.seg;
  const bool kept_replay = h->obj.kept_replay && nrep > 0 && h->obj.kept.size() >= (size_t)K + 1;
  const size_t nsegbuf = nrep > 0 && !kept_replay ? (size_t)K + 1 : 0;
  while (seg.size() > nsegbuf) {
    (void)hipFree(seg.back());
    seg.pop_back();
  }
  while (!rc && seg.size() < nsegbuf) {
    void *q = nullptr;
    if ((e = hipMalloc(&q, sb)) != hipSuccess) fail("segment buffers", e);
    else seg.push_back(q);
  }
  if (!rc && !h->obj.L && (e = hipMalloc(&h->obj.L, sb)) != hipSuccess) fail("adjoint state", e);
  // gradient partials: one slot per block of the step kernel, then one per
  // block of the face-node kernel
  // face blocks of the fused adjoint step: a multiple of 8 (XCD mapping)
  const int nfb = (int)((blocks_for((long)h->n_lo + h->n_hi, kAdjBlock) + 7) / 8 * 8);
  const int nslot = nblk + nfb;
  if (!rc && !h->obj.gacc && (e = hipMalloc(&h->obj.gacc, (size_t)nslot * 5 * sizeof(double))) != hipSuccess)
    fail("gradients", e);
  if (!rc && !h->obj.cgrad && (e = hipMalloc(&h->obj.cgrad, (size_t)h->E * sizeof(double))) != hipSuccess)
    fail("inlet gradient", e);
  if (!rc && !h->obj.s_step.empty() && !h->obj.seed) {
    if ((e = hipMalloc(&h->obj.seed, (size_t)h->nf * sizeof(double))) != hipSuccess) fail("observation seeds", e);
    else (void)hipMemsetAsync(h->obj.seed, 0, (size_t)h->nf * sizeof(double), h->stream);
  }
  double *L = h->obj.L, *gacc = h->obj.gacc, *cgrad = h->obj.cgrad;
  int sidx = (int)h->obj.s_step.size() - 1;  // the sweep visits observed steps backwards
  if (!rc) {
    (void)hipMemsetAsync(L, 0, sb, h->stream);
    (void)hipMemsetAsync(gacc, 0, (size_t)nslot * 5 * sizeof(double), h->stream);
    (void)hipMemsetAsync(cgrad, 0, (size_t)h->E * sizeof(double), h->stream);
  }
  const bool perm = h->obj.kind == LBM_OBJ_PERMEABILITY;
  const double nu = (h->p.tau[0] - 0.5) / 3.0;
  // dJ/d(plane flux) or dJ/d(z momentum of a node) on a window step
  const double seed = perm ? 0.0 : 1.0 / ((double)W_ * (double)h->E);
  const double kseed = perm && W_ > 0 ? nu / ((double)W_ * (double)h->N * perm_drive(h)) : 0.0;
  const bool off32 = !h->off64 && sb <= 0xffffffffull;
  // where the replayed segments' states go
  void *const *rp = kept_replay ? h->obj.kept.data() : seg.data();
  for (int sgi = nseg - 1; sgi >= 0 && !rc; sgi--) {
    const int ts = t0 + sgi * K;                 // state rp[0] is before step ts+1
    const int len = std::min(K, t0 + T - ts);
    // the states before the segment's steps, after their BCs: kept by the
    // objective run, or replayed from the segment's checkpoint
    void *const *sp = sgi >= nrep ? h->obj.kept.data() + ((long)sgi * K - h->obj.kept_from) : rp;
    if (sgi < nrep && (e = hipMemcpyAsync(rp[0], h->obj.ckpt[sgi], sb, hipMemcpyDeviceToDevice, h->stream)) !=
                          hipSuccess) {
      fail("checkpoint restore", e);
      break;
    }
    // replay the segment out of place, keeping every state after its BCs
    for (int q = 1; q <= len && !rc && sgi < nrep; q++) {
      const int t = ts + q;
      h->F = rp[q - 1];
      const bool fz = zface_fused(h, t);
      if (!fz && launch_bcs(h, (t & 1) ? 1 : 2, t)) {
        rc = 1;
        break;
      }
      StepArgs<double, double> A = make_args<double, double>(h);
      if (fz) set_zb(h, A);
      A.Fo = (double *)rp[q];
      A.solute_on = 1;
      const int bs = (t & 1) ? half_step_block<true, 2, double, double>() : half_step_block<false, 2, double, double>();
      const size_t lds = (t & 1) ? step_pad<true, 2, double, double>(h) : step_pad<false, 2, double, double>(h);
      A.nblk8 = (int)((blocks_for(h->nf, bs) + 7) / 8);
      const dim3 grid = (t & 1) ? odd_grid(h, A, bs) : dim3(8 * A.nblk8), block(bs);
      Prof pr(h, (t & 1) ? 0 : 1);
      if (t & 1) {
        if (off32) hipLaunchKernelGGL((k_half_step<19, 2, double, double, true, true>), grid, block, lds, h->stream, A);
        else hipLaunchKernelGGL((k_half_step<19, 2, double, double, true, false>), grid, block, lds, h->stream, A);
      } else {
        if (off32) hipLaunchKernelGGL((k_half_step<19, 2, double, double, false, true>), grid, block, lds, h->stream, A);
        else hipLaunchKernelGGL((k_half_step<19, 2, double, double, false, false>), grid, block, lds, h->stream, A);
      }
      if ((e = hipGetLastError()) != hipSuccess) fail("replay", e);
    }
    // reverse sweep over the segment
    for (int q = len; q >= 1 && !rc; q--) {
      const int t = ts + q;
      const bool odd = t & 1;
      AdjArgs J;
      h->F = sp[q - 1];
      const bool fz = zface_fused(h, t);
      J.A = make_args<double, double>(h);
      J.amap = h->amap;
      J.A.nblk8 = h->anblk8;
      if (fz) set_zb(h, J.A);
      J.L = L;
      J.seed = t > t0 + T - W_ ? seed : 0.0;
      J.kseed = t > t0 + T - W_ ? kseed : 0.0;
      J.zf = h->p.flux_plane;
      J.gacc = gacc;
      J.cgrad = cgrad;
      J.obs = nullptr;
      int s_lo = 0, s_n = 0;
      while (sidx >= 0 && h->obj.s_step[sidx] > t - t0) sidx--;
      if (sidx >= 0 && h->obj.s_step[sidx] == t - t0) {
        s_lo = h->obj.s_begin[sidx];
        s_n = h->obj.s_begin[sidx + 1] - s_lo;
        hipLaunchKernelGGL(k_obs_seed, dim3(blocks_for(s_n)), dim3(256), 0, h->stream, h->obj.seed,
                           h->obj.s_node_d + s_lo, h->obj.s_val_d + s_lo, s_n);
        J.obs = h->obj.seed;
      }
      // clears the seeds of this step once its adjoint kernels are queued
      struct SeedClear {
        lbm_handle *h;
        int lo, n;
        ~SeedClear() {
          if (n)
            hipLaunchKernelGGL(k_obs_seed, dim3(blocks_for(n)), dim3(256), 0, h->stream, h->obj.seed,
                               h->obj.s_node_d + lo, (const double *)nullptr, n);
        }
      } seed_clear{h, s_lo, s_n};
      // with the fused face conditions the first nfb blocks take the face
      // nodes (their conditions formed from the state, transposed before the
      // store); otherwise every node, and the transposes follow
      const int fb = fz ? nfb : 0;
      dim3 grid(nblk + fb), block(kAdjBlock);
      Prof pr(h, 5);
      if (odd) {
        if (off32) hipLaunchKernelGGL((k_adjoint_step<true, true>), grid, block, 0, h->stream, J, fb);
        else hipLaunchKernelGGL((k_adjoint_step<true, false>), grid, block, 0, h->stream, J, fb);
      } else {
        if (off32) hipLaunchKernelGGL((k_adjoint_step<false, true>), grid, block, 0, h->stream, J, fb);
        else hipLaunchKernelGGL((k_adjoint_step<false, false>), grid, block, 0, h->stream, J, fb);
      }
      if ((e = hipGetLastError()) != hipSuccess) {
        fail("adjoint step", e);
        break;
      }
      if (fz) continue;
      if (zface_any(h)) {
        for (int top = 0; top < 2; top++) {
          AdjBcArgs B;
          B.A = J.A;
          B.L = L;
          B.list = top ? h->face_hi : h->face_lo;
          B.n = top ? h->n_hi : h->n_lo;
          B.phase = odd ? 1 : 2;
          if (!B.n) continue;
          const ZfaceArgs Z = zface_args(h, top);
          hipLaunchKernelGGL(k_adjoint_zface, dim3(blocks_for(B.n)), dim3(256), 0, h->stream, B, Z, cgrad);
          if ((e = hipGetLastError()) != hipSuccess) fail("adjoint BC", e);
        }
      }
    }
  }
  h->F = saved_F;
  // the first kept buffers now hold replayed states: this run's kept
  // segments are gone, and another sweep needs another objective run
  if (kept_replay) h->obj.valid = 0;
  if (!rc) {
    std::vector<double> part((size_t)nslot * 5);
    if ((e = hipMemcpyAsync(part.data(), gacc, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess) {
      fail("gradient read-back", e);
    } else {
      for (int q = 0; q < 5; q++) grad[q] = 0.0;
      for (int b = 0; b < nslot; b++)
        for (int q = 0; q < 5; q++) grad[q] += part[(size_t)b * 5 + q];
      if (perm && W_ > 0) {
        // the permeability's own dependence on tau (through nu) and on g_z
        // (u_phys = u_eq - (tau - 1/2) g, and the driving gradient)
        const double Jw = h->obj.Jwin, G = perm_drive(h), rA = h->p.rho_A[0];
        grad[3] += Jw / (3.0 * nu);
        grad[2] += rA * nu * (double)h->nf / (2.0 * G * (double)h->N) - rA * Jw / G;
      }
      if (grad_cmap) {
        if ((e = hipMemcpy(grad_cmap, cgrad, (size_t)h->E * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
          fail("inlet gradient read-back", e);
      }
    }
  }
  (void)hipStreamSynchronize(h->stream);
  h->time = saved_time;
  return rc;
}
