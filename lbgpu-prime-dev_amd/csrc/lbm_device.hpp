// Device code of the stream-collide library (gfx950, wave64).
//
// Storage: populations of fluid nodes only, in tiles of kTile consecutive
// compact nodes (lbm_index.hpp), compact order = dense order (i fastest);
// with a solute, the two populations of a slot are adjacent.
// Slot meaning follows the reference's in-place AA scheme (Bailey et al.,
// ICPP 2009; src/k_stream_collide_stream.c, src/k_collide.c): before an odd
// half-step slot a at x holds the post-collision f_a(x); the odd half-step
// pulls slot a from x - c_a, collides and pushes into slot opp(a) at
// x + c_a; the even half-step reads the own node's opposite slots and writes
// its own slots in order.  Halfway bounce-back: when x - c_a is solid the
// pull reads the own slot opp(a); when x + c_a is solid the push writes the
// own slot a (src/lattice.c:2730-2741, 2835-2844).
//
// All arithmetic is written in the reference's order; the library is built
// with -ffp-contract=off so the FP64 path reproduces the FP64 oracle bit for
// bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "lbm_index.hpp"

namespace lbm {

constexpr int kBlock = 256;
// Workgroup size of the odd (pull/push) half-step of the kinds that run at
// full occupancy (FP64 flow, FP32 storage with the FP32 collision): 8-wave
// workgroups keep more of its 19 gathers per node in flight on a CU (flow
// FP64 at 256^3, porosity 0.40: 0.437 ms at 512 threads against 0.457 ms at
// 256 and 0.491 ms at 1024, contiguous block order).  FP16 storage runs 256
// threads in the row-band order with 256-node blocks (0.242 against 0.267 ms
// at 256^3 in the contiguous order; 512 threads gained 4 % at 256^3 but lost
// 4 % at 192^3).  StepShape below gives every kind's shape.
constexpr int kBlockOdd = 512;
// Launch shape of a half-step kernel instantiation: threads per workgroup,
// and the number of its workgroups one CU may hold (0: as many as the
// registers allow; otherwise each workgroup reserves LDS it does not use,
// lds_per_cu / (per_cu + 1/2) rounded down, so that no more fit).
// With FP64 storage and collision the half-steps run at low occupancy in
// small workgroups: fewer requests in flight per XCD let its L2 serve the
// lines that neighbouring waves share (and the read-then-write of each line)
// before they leave it, and small workgroups retire wave by wave instead of
// waiting for their slowest wave.  Measured at 256^3 (porosity 0.40, odd
// step in the row-band order), odd / even half-step:
//   flow + solute: odd in one-wave workgroups, at most 6 per CU (1.5 waves
//   per SIMD): 0.793 ms against 0.830 with 512 threads (two waves per SIMD,
//   registers allow no more), 0.841 with 256 (three per SIMD), 0.806 / 0.833
//   / 0.814 with at most 8 / 4 / 10 one-wave workgroups; even in two-wave
//   workgroups, at most 2 per CU (one wave per SIMD): 0.642 ms against 0.666
//   with 256 threads uncapped, 0.648 with one-wave workgroups at most 3-6
//   per CU, 0.659 with 256 threads at most 2-3 per CU;
//   flow: odd stays at 512 threads uncapped (0.395 ms; capped or smaller,
//   0.42-0.54); even in two-wave workgroups at most 3 per CU: 0.328 against
//   0.338 ms.
// FP32 storage with the FP64 collision (flow + solute; 141 / 178 registers:
// at most three / two waves per SIMD) takes the small workgroups without a
// cap: odd in one-wave workgroups 0.420 ms against 0.441 with 512 threads,
// even in two-wave workgroups 0.307 against 0.331 with 256 (capped at 1.5
// or 1 wave per SIMD both were slower: 0.46-0.70 / 0.32-0.36 ms).
// FP32 storage with the FP32 collision and FP16 storage keep 512 / 256
// threads uncapped: smaller workgroups or caps made them 1-70 % slower.
template <bool ODD, int NS, typename ST, typename CT>
struct StepShape {
  static constexpr bool f64 = sizeof(ST) == 8 && sizeof(CT) == 8;
  static constexpr bool f32_f64 = NS == 2 && sizeof(ST) == 4 && sizeof(CT) == 8;
  static constexpr int block = ODD ? (NS == 2 && (f64 || f32_f64) ? 64 : (sizeof(ST) >= 4 ? kBlockOdd : kBlock))
                                   : (f64 || f32_f64 ? 128 : kBlock);
  static constexpr int per_cu = !f64 ? 0 : (ODD ? (NS == 2 ? 6 : 0) : (NS == 2 ? 2 : 3));
};
template <bool ODD, int NS, typename ST, typename CT>
constexpr int half_step_block() { return StepShape<ODD, NS, ST, CT>::block; }
// LDS a half-step workgroup reserves to cap the workgroups per CU
inline size_t step_lds_pad(int per_cu, int lds_per_cu) {
  if (per_cu <= 0 || lds_per_cu <= 0) return 0;
  return (size_t)(lds_per_cu / (per_cu + 0.5)) / 2048 * 2048;
}

// D3Q19 velocity set in the reference's order (src/lbgpu_prime.h:78-84):
// opposite directions adjacent, D2Q9 first.
__host__ __device__ constexpr int CX(int a) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0};
  return t[a];
}
__host__ __device__ constexpr int CY(int a) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1};
  return t[a];
}
__host__ __device__ constexpr int CZ(int a) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
  return t[a];
}
__host__ __device__ constexpr int OPP(int a) { return a == 0 ? 0 : ((a & 1) ? a + 1 : a - 1); }

// weights of src/latman.c:161-192
template <int Q>
__host__ __device__ constexpr double W(int a) {
  if (Q == 9) return a == 0 ? 4. / 9. : (a < 5 ? 1. / 9. : 1. / 36.);
  return a == 0 ? 1. / 3. : ((a < 5 || a == 9 || a == 10) ? 1. / 18. : 1. / 36.);
}

// population coding in storage
template <typename ST, typename CT>
struct Codec {
  static __device__ __forceinline__ CT dec(ST v, CT inv_scale, CT shift) {
    // FP32 collision: one fused multiply-add (no FP64 parity to keep)
    if constexpr (sizeof(CT) == 4) return __builtin_fmaf((CT)v, inv_scale, shift);
    else return (CT)v * inv_scale + shift;
  }
  static __device__ __forceinline__ ST enc(CT v, CT scale, CT shift) { return (ST)((v - shift) * scale); }
};
template <>
struct Codec<double, double> {
  static __device__ __forceinline__ double dec(double v, double, double) { return v; }
  static __device__ __forceinline__ double enc(double v, double, double) { return v; }
};
template <>
struct Codec<double, float> {
  static __device__ __forceinline__ float dec(double v, float, float) { return (float)v; }
  static __device__ __forceinline__ double enc(float v, float, float) { return (double)v; }
};

// Population element e of the tile array.  OFF32: the array is smaller
// than 4 GiB, so the byte offset is a 32-bit value added to the uniform
// base (saddr + voffset addressing, half the address registers).
template <bool OFF32, typename T>
__device__ __forceinline__ T *elem(T *F, uint32_t e) {
  if (OFF32) return (T *)((char *)F + (uint32_t)(e * (uint32_t)sizeof(T)));
  return F + e;
}
// Loads and stores of populations.  The even half-step streams whole
// tiles and uses non-temporal accesses; the odd half-step's gathers and
// scatters share partial lines between neighbouring waves, which must stay
// in L2, so it uses ordinary ones.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
  if (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Kinds of the 3-D z-face conditions, one per face ([0] z = 0, [1] z = nk - 1).
// Flow: Zou-He with a prescribed density (pressure) or z velocity; solute:
// Dirichlet concentration or zero gradient.
enum ZFlow : int { kZNone = 0, kZPressure = 1, kZVelocity = 2 };
enum ZConc : int { kCNone = 0, kCDirichlet = 1, kCZeroGrad = 2 };

// 3-D z-face conditions applied by the half-step itself to the populations
// a face node has gathered (the values k_bc_zface would have stored where
// the node reads them).  Compact indices [0, lo_end) are the fluid nodes
// of z = 0 and [hi_begin, nf) those of z = nk - 1 (dense order, z slowest).
struct ZfaceFuse {
  int on;
  int lo_end, hi_begin;
  int flow[2];         // ZFlow per face
  int conc[2];         // ZConc per face
  double fval[2];      // pressure face: rho_b; velocity face: rho_A u_z, its z momentum
  double cval[2];      // Dirichlet concentration
  const double *cmap;  // [nj][ni] Dirichlet concentration of the z = 0 face, or null
};

template <typename ST, typename CT>
struct StepArgs {
  ST *F;          // tiles [nf/kTile][Q][kTile][NS]: populations the half-step reads
  ST *Fo;         // where it writes them (== F in place; another state
                  // buffer when the adjoint replays a segment)
  Geo G;
  int nf;         // fluid nodes
  int nblk8;      // launched blocks / 8
  const int *bmap; // node block of each launched block (-1: none), or null
  double *mv;     // [nsubs][1+D][nf]
  int write_mv;
  int solute_on;
  CT rhoA, tau, inv_tau, g[3];
  CT rhoA_s, inv_tau_s, sink_fac;
  CT scale[2], inv_scale[2], c0;
  ZfaceFuse zb;   // zb.on = 0: no boundary work in the half-step
  const double *ns;  // [nf] Walsh partial bounce-back fraction (WALSH kernels)
};

// Workgroups are dispatched round-robin over the 8 XCDs; map them so that
// each XCD sweeps one contiguous range of fluid nodes (its own L2 keeps the
// partial lines that neighbouring waves share).
__device__ __forceinline__ int xcd_block(int nblk8, int b) { return (b & 7) * nblk8 + (b >> 3); }
__device__ __forceinline__ int xcd_block(int nblk8) { return xcd_block(nblk8, (int)blockIdx.x); }

template <int Q, typename ST, typename CT, int SUBS>
__device__ __forceinline__ CT shift_of(const StepArgs<ST, CT> &A, int a) {
  return SUBS == 1 ? (CT)W<Q>(a) * A.c0 : (CT)0;
}

// nb[a] = compact index of x - c_a (or -1 when solid), a = 1..Q-1
template <int Q>
__device__ __forceinline__ void neighbors(const Geo &G, int x, int (&nb)[Q]) {
  constexpr int D = Q == 9 ? 2 : 3;
  int i, j, k;
  coords<D>(G, (uint32_t)G.node[x], i, j, k);
  int r[Q - 1];
  neighbors_rows<D>(G, i, j, k, r);
  nb[0] = x;
#pragma unroll
  for (int a = 1; a < Q; a++) nb[a] = r[a - 1];
}

// Gather the Q populations one half-step collides.
template <int Q, int NS, typename ST, typename CT, int SUBS, bool ODD, bool OFF32 = false>
__device__ __forceinline__ void gather(const StepArgs<ST, CT> &A, int idx, const int (&nb)[Q],
                                       CT (&f)[Q]) {
#pragma unroll
  for (int a = 0; a < Q; a++) {
    uint32_t off;
    if (a == 0)
      off = foff<Q, NS>(SUBS, 0, idx);
    else if (ODD)
      off = nb[a] >= 0 ? foff<Q, NS>(SUBS, a, nb[a]) : foff<Q, NS>(SUBS, OPP(a), idx);
    else
      off = foff<Q, NS>(SUBS, OPP(a), idx);
    f[a] = Codec<ST, CT>::dec(ld_stream<!ODD>(elem<OFF32>(A.F, off)),
                              A.inv_scale[SUBS], shift_of<Q, ST, CT, SUBS>(A, a));
  }
}

// ---- FP16 stores with stochastic rounding ----------------------------------
// Rounded to nearest, a population that changes little from one step to the
// next rounds the same way at every step and the errors add up (192^3 flow +
// solute, 4000 steps: the density drifted by 1e-4 in nearly closed pockets,
// the concentration behind the front came out up to 4 % low).  The
// half-steps therefore round each stored FP16 value up or down with
// probabilities given by its distance to the two neighbouring FP16 values,
// which makes every rounding unbiased (the errors then add like a random
// walk).  gfx950's v_cvt_sr_f16_f32 does the rounding in one instruction,
// from the top bits of a 32-bit random operand (checked on the device with
// 8192 operands per value: with the operand's top 13 bits uniform the mean
// equals the input to the last bit, subnormal results included, and values
// near 65504 overflow to infinity with the right probability, as round to
// nearest would past 65520; with only its low bits set it truncates).  The
// random operand of a value is the low 13 mantissa bits of another value of
// the same node (its FP32 rounding residue, which the value's own rounding
// does not see) moved to the top, plus a per-node seed hashed from the node
// index and the node's post-collision rest populations: one shift-add per
// value (32-bit multiplies per slot made the FP16 even half-step 10 %
// slower).  The stored values thus depend on the node's state only, not on
// the block order or the launch path (captured graph or step by step), and
// a node whose state does not change keeps storing the same values.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  // integer finaliser "lowbias32" (C. Wellons, hash-prospector)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t sr_seed(int idx, float r0, float r1) {
  const uint32_t b1 = __float_as_uint(r1);
  return mix32((uint32_t)idx * 0x9e3779b9u ^ __float_as_uint(r0) ^ ((b1 << 13) | (b1 >> 19)));
}
// random operand for a value: the low 13 mantissa bits of `other` moved to
// the top, plus the seed
__device__ __forceinline__ uint32_t sr_rand(float other, uint32_t seed) {
  return (__float_as_uint(other) << 19) + seed;
}
typedef _Float16 half2_sr __attribute__((ext_vector_type(2)));
// x and y to FP16 with stochastic rounding, each from the other's residue
__device__ __forceinline__ half2_sr sr_pair(float x, float y, uint32_t seed) {
  half2_sr v = {(_Float16)0, (_Float16)0};
  v = __builtin_amdgcn_cvt_sr_f16_f32(v, x, sr_rand(y, seed), false);
  v = __builtin_amdgcn_cvt_sr_f16_f32(v, y, sr_rand(x, (seed << 16) | (seed >> 16)), true);
  return v;
}
__device__ __forceinline__ _Float16 sr_one(float x, uint32_t r) {
  half2_sr v = {(_Float16)0, (_Float16)0};
  v = __builtin_amdgcn_cvt_sr_f16_f32(v, x, r, false);
  return v.x;
}

template <int Q, int NS, typename ST, typename CT, int SUBS, bool ODD, bool OFF32 = false>
__device__ __forceinline__ void scatter(const StepArgs<ST, CT> &A, int idx, const int (&nb)[Q],
                                        const CT (&f)[Q]) {
  [[maybe_unused]] uint32_t seed = 0;
  if constexpr (sizeof(ST) == 2) seed = sr_seed(idx, (float)f[0], (float)f[1]);
#pragma unroll
  for (int a = 0; a < Q; a++) {
    [[maybe_unused]] constexpr int kPartner[19] = {1, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    uint32_t off;
    if (a == 0)
      off = foff<Q, NS>(SUBS, 0, idx);
    else if (ODD)
      off = nb[OPP(a)] >= 0 ? foff<Q, NS>(SUBS, OPP(a), nb[OPP(a)]) : foff<Q, NS>(SUBS, a, idx);
    else
      off = foff<Q, NS>(SUBS, a, idx);
    if constexpr (sizeof(ST) == 2) {
      // the partner's value: another slot of the node (its opposite)
      const int b = kPartner[a];
      const float y = (float)((f[a] - shift_of<Q, ST, CT, SUBS>(A, a)) * A.scale[SUBS]);
      const float yb = (float)((f[b] - shift_of<Q, ST, CT, SUBS>(A, b)) * A.scale[SUBS]);
      st_stream<!ODD>(elem<OFF32>(A.Fo, off), (ST)sr_one(y, sr_rand(yb, seed)));
    } else {
      st_stream<!ODD>(
          elem<OFF32>(A.Fo, off), Codec<ST, CT>::enc(f[a], A.scale[SUBS], shift_of<Q, ST, CT, SUBS>(A, a)));
    }
  }
}

// Flow and solute together: one 2-wide access per slot moves both
// populations (the pair layout of foff<Q, 2>).
template <typename ST>
using pair_t = ST __attribute__((ext_vector_type(2)));

// NT: streaming loads also in the odd half-step (out-of-place steps)
template <int Q, typename ST, typename CT, bool ODD, bool OFF32, bool NT = false>
__device__ __forceinline__ void gather2(const StepArgs<ST, CT> &A, int idx, const int (&nb)[Q], CT (&f)[Q],
                                        CT (&g)[Q]) {
#pragma unroll
  for (int a = 0; a < Q; a++) {
    uint32_t off;
    if (a == 0)
      off = foff<Q, 2>(0, 0, idx);
    else if (ODD)
      off = nb[a] >= 0 ? foff<Q, 2>(0, a, nb[a]) : foff<Q, 2>(0, OPP(a), idx);
    else
      off = foff<Q, 2>(0, OPP(a), idx);
    const pair_t<ST> v = ld_stream<!ODD || NT>((const pair_t<ST> *)elem<OFF32>(A.F, off));
    f[a] = Codec<ST, CT>::dec(v.x, A.inv_scale[0], (CT)0);
    g[a] = Codec<ST, CT>::dec(v.y, A.inv_scale[1], shift_of<Q, ST, CT, 1>(A, a));
  }
}

template <int Q, typename ST, typename CT, bool ODD, bool OFF32>
__device__ __forceinline__ void scatter2(const StepArgs<ST, CT> &A, int idx, const int (&nb)[Q],
                                         const CT (&f)[Q], const CT (&g)[Q]) {
  [[maybe_unused]] uint32_t seed = 0;
  if constexpr (sizeof(ST) == 2) seed = sr_seed(idx, (float)f[0], (float)g[0]);
#pragma unroll
  for (int a = 0; a < Q; a++) {
    uint32_t off;
    if (a == 0)
      off = foff<Q, 2>(0, 0, idx);
    else if (ODD)
      off = nb[OPP(a)] >= 0 ? foff<Q, 2>(0, OPP(a), nb[OPP(a)]) : foff<Q, 2>(0, a, idx);
    else
      off = foff<Q, 2>(0, a, idx);
    pair_t<ST> v;
    if constexpr (sizeof(ST) == 2) {
      const half2_sr h = sr_pair((float)(f[a] * A.scale[0]),
                                 (float)((g[a] - shift_of<Q, ST, CT, 1>(A, a)) * A.scale[1]), seed);
      v.x = h.x;
      v.y = h.y;
    } else {
      v.x = Codec<ST, CT>::enc(f[a], A.scale[0], (CT)0);
      v.y = Codec<ST, CT>::enc(g[a], A.scale[1], shift_of<Q, ST, CT, 1>(A, a));
    }
    st_stream<!ODD>((pair_t<ST> *)elem<OFF32>(A.Fo, off), v);
  }
}

// acc + c x for a compile-time velocity component c in {-1, 0, 1}.  The
// product by +-1 is exact and adding an exact zero leaves acc unchanged, so
// for finite x the value equals the written acc + (CT)c * x, without the
// multiplications by zero that the compiler must otherwise keep (NaN and
// signed-zero rules).
template <int C, typename CT>
__device__ __forceinline__ CT add_c(CT acc, CT x) {
  if constexpr (C == 0) return acc;
  else if constexpr (C > 0) return acc + x;
  else return acc - x;
}
// c . u over the nonzero components of c, in the written order x, y, z
template <int CXa, int CYa, int CZa, typename CT>
__device__ __forceinline__ CT cdotu(CT u0, CT u1, CT u2) {
  if constexpr (CXa != 0) return add_c<CZa, CT>(add_c<CYa, CT>(CXa > 0 ? u0 : -u0, u1), u2);
  else if constexpr (CYa != 0) return add_c<CZa, CT>(CYa > 0 ? u1 : -u1, u2);
  else if constexpr (CZa != 0) return CZa > 0 ? u2 : -u2;
  else return (CT)0;
}

// Moments, force, macroscopic write and BGK relaxation of one subs, in the
// order of src/k_stream_collide_stream.c:80-390 and
// src/lattice.c:2850-2944.  SUBS 0: He-Luo incompressible flow on shifted
// populations; SUBS 1: Inamuro solute advected by the flow's u_eq.
// WALSH: partial bounce-back (src/k_stream_collide_stream.c:129-135,183-192,
// src/lattice.c:2936-2939): the momentum and the force enter u scaled by
// kns = 1 - ns; the caller mixes the opposite pre-collision populations in.
template <int Q, typename ST, typename CT, int SUBS, bool ODD, bool WALSH = false>
__device__ __forceinline__ void collide(const StepArgs<ST, CT> &A, int idx, CT (&f)[Q], CT (&u)[3],
                                        CT kns = (CT)1) {
  constexpr int D = Q == 9 ? 2 : 3;
  CT rho = f[0];
#pragma unroll
  for (int a = 1; a < Q; a++) rho += f[a];
  const CT rho_pre = rho;  // mass before the sink
  CT jpre[3] = {(CT)0, (CT)0, (CT)0};  // momentum before the collision (FP32 correction)
  if (SUBS == 0) {
    u[0] = (CT)0;
    u[1] = (CT)0;
    u[2] = (CT)0;
    if constexpr (sizeof(CT) == 4) {
      // FP32 collision: the zero terms skipped (same values; 16 % fewer
      // instructions: FP16 odd step 3 % faster)
      [&]<int... A>(std::integer_sequence<int, A...>) {
        ((u[0] = add_c<CX(A), CT>(u[0], f[A]), u[1] = add_c<CY(A), CT>(u[1], f[A]),
          u[2] = D == 3 ? add_c<CZ(A), CT>(u[2], f[A]) : u[2]),
         ...);
      }(std::make_integer_sequence<int, Q>{});
    } else {
      // FP64: as written (the skipped form frees registers, and the flow +
      // solute odd step's second workgroup per CU then slows it by 5 % at
      // 256^3)
#pragma unroll
      for (int a = 0; a < Q; a++) {
        u[0] += ((CT)CX(a)) * f[a];
        u[1] += ((CT)CY(a)) * f[a];
        if (D == 3) u[2] += ((CT)CZ(a)) * f[a];
      }
    }
    if constexpr (sizeof(CT) == 4 && sizeof(ST) >= 4) {
#pragma unroll
      for (int d = 0; d < D; d++) jpre[d] = u[d];
    }
  } else if constexpr (sizeof(CT) == 4 && sizeof(ST) >= 4) {
#pragma unroll
    for (int a = 1; a < Q; a++) {
      jpre[0] += ((CT)CX(a)) * f[a];
      jpre[1] += ((CT)CY(a)) * f[a];
      if (D == 3) jpre[2] += ((CT)CZ(a)) * f[a];
    }
  }
  if (SUBS == 0) {
    if constexpr (WALSH) {
#pragma unroll
      for (int d = 0; d < D; d++) u[d] *= kns;
    }
#pragma unroll
    for (int d = 0; d < D; d++) u[d] /= A.rhoA;
    if constexpr (WALSH) {
#pragma unroll
      for (int d = 0; d < D; d++) u[d] += kns * A.tau * A.g[d];
    } else {
#pragma unroll
      for (int d = 0; d < D; d++) u[d] += A.tau * A.g[d];
    }
  } else {
    if (!ODD) rho *= A.sink_fac;
    rho *= A.sink_fac;
  }
  if (A.write_mv) {
    // write_mv 2 is the PEST observation switch: flow u only, solute c only
    // (src/k_stream_collide_stream.c:305-334)
    double *mv = A.mv + (long)SUBS * (1 + D) * A.nf;
    const bool part = A.write_mv == 2;
    if (!part || SUBS == 1) {
      if (SUBS == 0)
        mv[idx] = (double)rho + (double)A.rhoA;  // in FP64 before the sum (FP32: rho is an offset)
      else
        mv[idx] = (double)rho;
    }
    if (!part || SUBS == 0) {
#pragma unroll
      for (int d = 0; d < D; d++) mv[(long)(d + 1) * A.nf + idx] = (double)u[d];
    }
  }
  CT usq = (CT)0;
  if (SUBS == 0) {
    usq = u[0] * u[0] + u[1] * u[1];
    if (D == 3) usq += u[2] * u[2];
  }
  const CT it = SUBS == 0 ? A.inv_tau : A.inv_tau_s;
  CT vd[Q];
  if constexpr (sizeof(CT) == 4) {
    [&]<int... A>(std::integer_sequence<int, A...>) {
      ((vd[A] = cdotu<CX(A), CY(A), D == 3 ? CZ(A) : 0, CT>(u[0], u[1], u[2])), ...);
    }(std::make_integer_sequence<int, Q>{});
  } else {
#pragma unroll
    for (int a = 0; a < Q; a++) {
      vd[a] = ((CT)CX(a)) * u[0] + ((CT)CY(a)) * u[1];
      if (D == 3) vd[a] += ((CT)CZ(a)) * u[2];
    }
  }
  if constexpr (sizeof(CT) == 4) {
    // FP32 collision, the same update factored per node and fused
    // (FP32 and FP16 storage have no FP64 parity to keep): flow
    //   f (1 - it) + w [it (rho - 1.5 rho_A u.u) + it rho_A v (3 + 4.5 v)],
    // solute  f (1 - it) + w it rho (1 + 3 v),  v = c_a . u
    const CT k1 = (CT)1 - it;
    const CT B = SUBS == 1 ? it * rho : it * (rho - (CT)1.5 * A.rhoA * usq);
    const CT C = it * A.rhoA;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      const CT v = vd[a];
      const CT w = (CT)W<Q>(a);
      CT t;
      if (SUBS == 1) t = B * __builtin_fmaf((CT)3, v, (CT)1);
      else t = __builtin_fmaf(C, v * __builtin_fmaf((CT)4.5, v, (CT)3), B);
      f[a] = __builtin_fmaf(f[a], k1, w * t);
    }
  } else {
#pragma unroll
    for (int a = 0; a < Q; a++) {
      CT fa = f[a] * ((CT)1 - it);
      const CT vdotu = vd[a];
      const CT w = (CT)W<Q>(a);
      if (SUBS == 1) {
        fa += w * rho * ((CT)1 + (CT)3 * vdotu) * it;
      } else {
        fa += w * rho * it;
        fa += w * A.rhoA * ((CT)3 * vdotu + (CT)4.5 * vdotu * vdotu - (CT)1.5 * usq) * it;
      }
      f[a] = fa;
    }
  }
  if constexpr (sizeof(CT) == 4 && sizeof(ST) >= 4) {
    // FP32 collision (FP32 or FP64 storage; FP16 storage, whose rounding
    // is far larger, keeps the plain update and its speed).  The weights
    // rounded to FP32 sum to 1 + 1.5e-8 and
    // give sum_a w_a c_a,x^2 = 1/3 + 2.5e-9, so every collision would add
    // mass and momentum of that relative size: a drift of ~1e-4 in the
    // density and ~5e-5 in the permeability over 4000 steps.  The
    // post-collision moments are set to their exact values instead, with
    // unbiased rounding: the momentum (1 - w) j + w rho_A u (flow) or
    // (1 - w_s) j + w_s c u (solute) through the pair of axis directions
    // of each component, then the mass (flow delta-rho, solute (1 - w_s) c
    // + w_s c after the sink) through the rest population.
    {
      const CT m1 = SUBS == 0 ? A.rhoA : rho;
      CT jt[3];
#pragma unroll
      for (int d = 0; d < D; d++) jt[d] = ((CT)1 - it) * jpre[d] + it * m1 * u[d];
      constexpr int pos[3] = {1, 3, 9}, neg[3] = {2, 4, 10};
#pragma unroll
      for (int d = 0; d < D; d++) {
        CT jc = (CT)0;
#pragma unroll
        for (int a = 1; a < Q; a++) {
          const int c = d == 0 ? CX(a) : (d == 1 ? CY(a) : CZ(a));
          if (c > 0) jc += f[a];
          if (c < 0) jc -= f[a];
        }
        const CT half = (CT)0.5 * (jt[d] - jc);
        f[pos[d]] += half;
        f[neg[d]] -= half;
      }
    }
    const CT mass = SUBS == 0 ? rho_pre : ((CT)1 - it) * rho_pre + it * rho;
    CT others = f[1];
#pragma unroll
    for (int a = 2; a < Q; a++) others += f[a];
    f[0] = mass - others;
  }
}

// ---- 3-D z-face conditions (no reference counterpart) -------------------
// Directions moving up (+z) and down (-z); the unknown populations of a
// bottom-face node are the up-movers, of a top-face node the down-movers.
constexpr int kUp[5] = {9, 11, 13, 15, 17};
constexpr int kDn[5] = {10, 12, 14, 16, 18};

// Zou-He with zero tangential velocity (Hecht & Harting's D3Q19 form): the
// five unknowns of a face node from its known populations and either the
// prescribed density offset val = rho_b - rho_A (pressure: the node's z
// momentum J follows from the mass balance) or the prescribed z momentum
// val = rho_A u_z (velocity: the incompressible model's u = j / rho_A, the
// density then follows from the unknowns).
template <typename CT>
__device__ __forceinline__ void zface_flow_values(const CT (&f)[19], bool top, bool velocity, CT val,
                                                  CT (&v)[5]) {
  CT P = f[0];
#pragma unroll
  for (int a = 1; a < 9; a++) P += f[a];
  // (constant indices in both branches: a select between two elements
  // would index the array at run time and move it to scratch memory)
  CT M = (CT)0;
  if (top) {
#pragma unroll
    for (int q = 0; q < 5; q++) M += f[kUp[q]];
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) M += f[kDn[q]];
  }
  const CT Nx = ((f[1] + f[5] + f[8]) - (f[2] + f[7] + f[6])) / (CT)2;
  const CT Ny = ((f[3] + f[5] + f[7]) - (f[4] + f[6] + f[8])) / (CT)2;
  if (!top) {
    const CT J = velocity ? val : val - (P + (CT)2 * M);
    v[0] = f[10] + J / (CT)3;
    v[1] = f[12] + J / (CT)6 - Nx;
    v[2] = f[14] + J / (CT)6 + Nx;
    v[3] = f[16] + J / (CT)6 - Ny;
    v[4] = f[18] + J / (CT)6 + Ny;
  } else {
    const CT J = velocity ? val : (P + (CT)2 * M) - val;
    v[0] = f[9] - J / (CT)3;
    v[1] = f[11] - J / (CT)6 + Nx;
    v[2] = f[13] - J / (CT)6 - Nx;
    v[3] = f[15] - J / (CT)6 + Ny;
    v[4] = f[17] - J / (CT)6 - Ny;
  }
}

// Solute: the unknowns are w_u c*.  Dirichlet (c = cb): c* = (cb - sum of
// the 14 known populations) / (sum of the unknowns' weights, 1/6).  Zero
// gradient: c* = (sum of the five outgoing populations) / (1/6), the 3-D
// form of k_sysbound_zeroconcgrad_n.  Bottom face: the unknowns move up,
// the outgoing populations down; top face the other way round.
template <typename CT>
__device__ __forceinline__ void zface_solute_values(const CT (&g)[19], bool top, bool dirichlet, CT cb,
                                                    CT (&v)[5]) {
  CT known = (CT)0, cs;
  if (dirichlet) {
#pragma unroll
    for (int a = 0; a < 9; a++) known += g[a];
  }
  if (!top) {
#pragma unroll
    for (int q = 0; q < 5; q++) known += g[kDn[q]];
    cs = dirichlet ? (cb - known) / (CT)(1. / 6.) : known / (CT)(1. / 6.);
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = (CT)W<19>(kUp[q]) * cs;
  } else {
#pragma unroll
    for (int q = 0; q < 5; q++) known += g[kUp[q]];
    cs = dirichlet ? (cb - known) / (CT)(1. / 6.) : known / (CT)(1. / 6.);
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = (CT)W<19>(kDn[q]) * cs;
  }
}

// the prescribed value of a face's flow condition in the compute type:
// rho_b - rho_A formed in FP64 (1.001 rounded to FP32 first would move a
// 1e-3 offset by 5e-5), or the z momentum rho_A u_z formed on the host
template <typename CT>
__device__ __forceinline__ CT zface_flow_val(int kind, double fval, CT rhoA) {
  return kind == kZVelocity ? (CT)fval : (CT)(fval - (double)rhoA);
}

// put the five values into the unknown slots of the gathered array, through
// the storage encoding as a separate boundary launch would
template <typename ST, typename CT, int SUBS>
__device__ __forceinline__ void zface_put(const StepArgs<ST, CT> &A, bool top, const CT (&v)[5], CT (&f)[19]) {
  // value selects only: the elements stay registers
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const CT sd = shift_of<19, ST, CT, SUBS>(A, kDn[q]), su = shift_of<19, ST, CT, SUBS>(A, kUp[q]);
    const CT vd = Codec<ST, CT>::dec(Codec<ST, CT>::enc(v[q], A.scale[SUBS], sd), A.inv_scale[SUBS], sd);
    const CT vu = Codec<ST, CT>::dec(Codec<ST, CT>::enc(v[q], A.scale[SUBS], su), A.inv_scale[SUBS], su);
    const CT od = f[kDn[q]], ou = f[kUp[q]];
    f[kDn[q]] = top ? vd : od;
    f[kUp[q]] = top ? ou : vu;
  }
}

template <int Q, typename ST, typename CT>
__device__ __forceinline__ bool zface_node(const StepArgs<ST, CT> &A, int idx) {
  return Q == 19 && A.zb.on && (idx < A.zb.lo_end || idx >= A.zb.hi_begin);
}

// the z-face conditions of a face node on its gathered flow (and solute)
template <typename ST, typename CT, bool SOLUTE>
__device__ __forceinline__ void zface_fix(const StepArgs<ST, CT> &A, int idx, CT (&f)[19], CT (&g)[19]) {
  const bool top = idx >= A.zb.hi_begin;
  const int fk = top ? A.zb.flow[1] : A.zb.flow[0];
  CT v[5];
  if (fk != kZNone) {
    zface_flow_values<CT>(f, top, fk == kZVelocity, zface_flow_val<CT>(fk, top ? A.zb.fval[1] : A.zb.fval[0], A.rhoA),
                          v);
    zface_put<ST, CT, 0>(A, top, v, f);
  }
  if (SOLUTE) {
    const int ck = top ? A.zb.conc[1] : A.zb.conc[0];
    if (ck != kCNone) {
      CT cb = (CT)(top ? A.zb.cval[1] : A.zb.cval[0]);
      if (!top && A.zb.cmap) cb = (CT)A.zb.cmap[A.G.node[idx]];  // z = 0: dense index = i + ni j
      zface_solute_values<CT>(g, top, ck == kCDirichlet, cb, v);
      zface_put<ST, CT, 1>(A, top, v, g);
    }
  }
}

// Walsh partial bounce-back: post-collision (1 - ns) BGK + ns times the
// opposite pre-collision population (src/k_stream_collide_stream.c:363-383,
// src/k_collide.c:373-392; Walsh, Burwinkle & Saar 2009)
template <int Q, typename CT>
__device__ __forceinline__ void walsh_mix(CT (&f)[Q], const CT (&fp)[Q], CT ns, CT kns) {
  f[0] = f[0] * kns + ns * fp[0];
#pragma unroll
  for (int a = 1; a < Q; a += 2) {
    f[a] = f[a] * kns + ns * fp[a + 1];
    f[a + 1] = f[a + 1] * kns + ns * fp[a];
  }
}

template <int Q, int NS, typename ST, typename CT, bool ODD, bool OFF32, bool WALSH = false>
__global__ __launch_bounds__((half_step_block<ODD, NS, ST, CT>())) void k_half_step(StepArgs<ST, CT> A) {
  // node block: from the row-band map when the launch carries one, else the
  // XCD-contiguous remap
  const int lb = A.bmap ? A.bmap[blockIdx.x] : xcd_block(A.nblk8);
  if (lb < 0) return;
  const int idx = lb * half_step_block<ODD, NS, ST, CT>() + (int)threadIdx.x;
  if (idx >= A.nf) return;
  int nb[Q];
  nb[0] = idx;
  if (ODD) neighbors<Q>(A.G, idx, nb);
  CT f[Q], u[3];
  CT ns = (CT)0, kns = (CT)1;
  if constexpr (WALSH) {
    const double nsd = A.ns[idx];
    ns = (CT)nsd;
    kns = (CT)(1. - nsd);
  }
  if (NS == 2 && A.solute_on) {
    CT g[Q];
    if constexpr (ODD && NS == 2 && sizeof(ST) == 8 && sizeof(CT) == 8) {
      // An out-of-place odd step (the adjoint's forward replay, the kept
      // steps of an objective run) does not write back the lines it reads:
      // streaming loads leave the L2 to the lines it writes (128^3 flow +
      // solute: replay 202 -> 188 ms per sweep, objective run with kept
      // states 188 -> 183 ms; in-place steps unchanged).
      if (A.Fo != A.F) gather2<Q, ST, CT, ODD, OFF32, true>(A, idx, nb, f, g);
      else gather2<Q, ST, CT, ODD, OFF32>(A, idx, nb, f, g);
    } else {
      gather2<Q, ST, CT, ODD, OFF32>(A, idx, nb, f, g);
    }
    if constexpr (Q == 19)
      if (zface_node<Q>(A, idx)) zface_fix<ST, CT, true>(A, idx, f, g);
    if constexpr (WALSH) {
      CT fp[Q], gp[Q];
#pragma unroll
      for (int a = 0; a < Q; a++) {
        fp[a] = f[a];
        gp[a] = g[a];
      }
      collide<Q, ST, CT, 0, ODD, true>(A, idx, f, u, kns);
      collide<Q, ST, CT, 1, ODD, true>(A, idx, g, u, kns);
      walsh_mix<Q, CT>(f, fp, ns, kns);
      walsh_mix<Q, CT>(g, gp, ns, kns);
    } else {
      collide<Q, ST, CT, 0, ODD>(A, idx, f, u);
      collide<Q, ST, CT, 1, ODD>(A, idx, g, u);
    }
    scatter2<Q, ST, CT, ODD, OFF32>(A, idx, nb, f, g);
    return;
  }
  // flow alone (the solute, when present, is outside its active time window)
  gather<Q, NS, ST, CT, 0, ODD, OFF32>(A, idx, nb, f);
  if constexpr (Q == 19)
    if (zface_node<Q>(A, idx)) zface_fix<ST, CT, false>(A, idx, f, f);
  if constexpr (WALSH) {
    CT fp[Q];
#pragma unroll
    for (int a = 0; a < Q; a++) fp[a] = f[a];
    collide<Q, ST, CT, 0, ODD, true>(A, idx, f, u, kns);
    walsh_mix<Q, CT>(f, fp, ns, kns);
  } else {
    collide<Q, ST, CT, 0, ODD>(A, idx, f, u);
  }
  scatter<Q, NS, ST, CT, 0, ODD, OFF32>(A, idx, nb, f);
}

// ---- boundary conditions ------------------------------------------------

template <typename ST, typename CT>
struct BcArgs {
  StepArgs<ST, CT> A;
  const int *list;  // compact indices of the face's fluid nodes
  int n;
  int phase;        // 1: before the odd half-step, 2: before the even one
};

template <int Q, int NS, typename ST, typename CT, int SUBS>
__device__ __forceinline__ void bc_gather(const BcArgs<ST, CT> &B, int idx, int (&nb)[Q], CT (&f)[Q]) {
  nb[0] = idx;
  if (B.phase == 1) {
    neighbors<Q>(B.A.G, idx, nb);
    gather<Q, NS, ST, CT, SUBS, true>(B.A, idx, nb, f);
  } else {
    gather<Q, NS, ST, CT, SUBS, false>(B.A, idx, nb, f);
  }
}

// store the value the coming half-step will collide as f_a:
//   phase 1 -> the slot the pull reads (slot a at x - c_a); when x - c_a is
//   solid the reference writes the own slot a (set_f1d_d, da = 0,
//   src/lattice.c:2835-2844) unless own_opp, which writes the own slot
//   opp(a) that the pull then reads;  phase 2 -> own slot opp(a).
template <int Q, int NS, typename ST, typename CT, int SUBS>
__device__ __forceinline__ void bc_store(const BcArgs<ST, CT> &B, int idx, const int (&nb)[Q], int a,
                                         CT v, bool own_opp) {
  uint32_t off;
  if (B.phase == 1)
    off = nb[a] >= 0 ? foff<Q, NS>(SUBS, a, nb[a]) : foff<Q, NS>(SUBS, own_opp ? OPP(a) : a, idx);
  else
    off = foff<Q, NS>(SUBS, OPP(a), idx);
  B.A.F[off] = Codec<ST, CT>::enc(v, B.A.scale[SUBS], shift_of<Q, ST, CT, SUBS>(B.A, a));
}

// 2-D Zou-He pressure on the N (j = nj-1) or S (j = 0) row:
// k_sysbound_pressure_{n,s}_{1,2}, src/system_boundary_kernels.c:4-451
// The prescribed density arrives in FP64: in the FP32 collision the
// offset rho_b - rho_A is formed in FP64 first (1.001 rounded to FP32 would
// move a 1e-3 offset by 5e-5), in FP64 the reference's order is kept.
// With velocity set (the velocity_{n,s}_{in,out} flags, which the
// reference parses but never applies) the row's y momentum x = rho_A u_y is
// prescribed instead (val = rho_A u_y, formed on the host) and the same
// three unknowns follow from it; having no reference kernel to follow, the
// velocity rows place an unknown whose source is solid in the own slot
// opp(a), where the pull reads it (as the 3-D faces do), instead of the
// reference rows' own slot a (which leaves the bounced value in place and
// hands the unknown to the interior node beside a side wall).
template <int NS, typename ST, typename CT, int SUBS>
__global__ void k_bc2d_pressure(BcArgs<ST, CT> B, int north, double rho_b, int velocity) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.n) return;
  const int idx = B.list[t];
  int nb[9];
  CT f[9];
  bc_gather<9, NS, ST, CT, SUBS>(B, idx, nb, f);
  enum { C_ = 0, E_, W_, N_, S_, NE_, SW_, NW_, SE_ };
  const CT rhoA = SUBS == 0 ? B.A.rhoA : B.A.rhoA_s;
  if (north) {
    CT x = f[C_] + f[E_] + f[W_] + f[N_] + f[N_] + f[NE_] + f[NE_] + f[NW_] + f[NW_];
    if (velocity) {
      x = (CT)rho_b;
    } else if constexpr (sizeof(CT) == 4) {
      x -= (CT)(rho_b - (double)rhoA);
    } else {
      x -= (CT)rho_b;
      x += rhoA;
    }
    CT vS = f[N_] - ((CT)(2. / 3.)) * x;
    CT vSW = f[NE_] + ((CT)(1. / 2.)) * (f[E_] - f[W_]) - ((CT)(1. / 6.)) * x;
    CT vSE = f[NW_] + ((CT)(1. / 2.)) * (f[W_] - f[E_]) - ((CT)(1. / 6.)) * x;
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, S_, vS, velocity);
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, SE_, vSE, velocity);
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, SW_, vSW, velocity);
  } else {
    CT x = -f[C_] - f[E_] - f[W_] - f[S_] - f[S_] - f[SE_] - f[SE_] - f[SW_] - f[SW_];
    if (velocity) {
      x = (CT)rho_b;
    } else if constexpr (sizeof(CT) == 4) {
      x += (CT)(rho_b - (double)rhoA);
    } else {
      x += (CT)rho_b;
      x -= rhoA;
    }
    CT vN = f[S_] + ((CT)(2. / 3.)) * x;
    CT vNE = f[SW_] + ((CT)(1. / 2.)) * (f[W_] - f[E_]) + ((CT)(1. / 6.)) * x;
    CT vNW = f[SE_] + ((CT)(1. / 2.)) * (f[E_] - f[W_]) + ((CT)(1. / 6.)) * x;
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, N_, vN, velocity);
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, NW_, vNW, velocity);
    bc_store<9, NS, ST, CT, SUBS>(B, idx, nb, NE_, vNE, velocity);
  }
}

// 2-D solute zero concentration gradient on the N or S row:
// k_sysbound_zeroconcgrad_{n,s}_{1,2}, src/system_boundary_kernels.c:454-831
template <typename ST, typename CT>
__global__ void k_bc2d_zerograd(BcArgs<ST, CT> B, int north) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.n) return;
  const int idx = B.list[t];
  int nb[9];
  CT f[9];
  bc_gather<9, 2, ST, CT, 1>(B, idx, nb, f);
  enum { C_ = 0, E_, W_, N_, S_, NE_, SW_, NW_, SE_ };
  constexpr double w1 = W<9>(1), w2 = W<9>(5);
  if (north) {
    CT c = (f[N_] + f[NE_] + f[NW_]) / ((CT)(w1 + w2 + w2));
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, S_, (CT)w1 * c, false);
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, SE_, (CT)w2 * c, false);
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, SW_, (CT)w2 * c, false);
  } else {
    CT c = (f[S_] + f[SW_] + f[SE_]) / ((CT)(w1 + w2 + w2));
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, N_, (CT)w1 * c, false);
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, NW_, (CT)w2 * c, false);
    bc_store<9, 2, ST, CT, 1>(B, idx, nb, NE_, (CT)w2 * c, false);
  }
}

// 3-D z faces (the reference's GPU path has none; its CPU path has the
// pressure and velocity forms on the top and bottom faces, src/bcs.c:592-862).
// Flow: Zou-He pressure or velocity with zero tangential velocity in the
// D3Q19 form of Hecht & Harting (J. Stat. Mech. 2010, P01018) on shifted
// populations; solute: Dirichlet concentration or zero gradient.  Unknowns
// replace what streaming would deliver.  One launch per face.
struct ZfaceArgs {
  int top;
  int flow;           // ZFlow of this face
  int conc;           // ZConc of this face (solute)
  double fval;        // pressure: rho_b; velocity: rho_A u_z
  double cval;        // Dirichlet concentration
  const double *cmap; // [nj][ni] Dirichlet concentration of the z = 0 face, or null
};

template <int NS, typename ST, typename CT>
__global__ void k_bc_zface(BcArgs<ST, CT> B, ZfaceArgs Z) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= B.n) return;
  const int idx = B.list[t];
  int nb[19];
  CT f[19], v[5];
  const bool top = Z.top;
  if (Z.flow != kZNone) {
    bc_gather<19, NS, ST, CT, 0>(B, idx, nb, f);
    zface_flow_values<CT>(f, top, Z.flow == kZVelocity, zface_flow_val<CT>(Z.flow, Z.fval, B.A.rhoA), v);
    if (top) {
#pragma unroll
      for (int q = 0; q < 5; q++) bc_store<19, NS, ST, CT, 0>(B, idx, nb, kDn[q], v[q], true);
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) bc_store<19, NS, ST, CT, 0>(B, idx, nb, kUp[q], v[q], true);
    }
  }
  if (NS == 2 && Z.conc != kCNone) {
    bc_gather<19, NS, ST, CT, 1>(B, idx, nb, f);
    CT cb = (CT)Z.cval;
    if (!top && Z.cmap) cb = (CT)Z.cmap[B.A.G.node[idx]];  // z = 0: dense index = i + ni j
    zface_solute_values<CT>(f, top, Z.conc == kCDirichlet, cb, v);
    if (top) {
#pragma unroll
      for (int q = 0; q < 5; q++) bc_store<19, NS, ST, CT, 1>(B, idx, nb, kDn[q], v[q], true);
    } else {
#pragma unroll
      for (int q = 0; q < 5; q++) bc_store<19, NS, ST, CT, 1>(B, idx, nb, kUp[q], v[q], true);
    }
  }
}

}  // namespace lbm
