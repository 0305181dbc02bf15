// Auxiliary device kernels: compaction tables, population upload/download,
// halo planes, plane flux and FP64 reductions.
#pragma once

#include "lbm_device.hpp"

namespace lbm {

// staging [Q][nf] doubles <-> storage rows of one subs
template <int Q, int NS, typename ST, typename CT, int SUBS>
__global__ void k_encode(StepArgs<ST, CT> A, const double *stage) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= A.nf) return;
#pragma unroll
  for (int a = 0; a < Q; a++)
    A.F[foff<Q, NS>(SUBS, a, idx)] =
        Codec<ST, CT>::enc((CT)stage[(long)a * A.nf + idx], A.scale[SUBS], shift_of<Q, ST, CT, SUBS>(A, a));
}

template <int Q, int NS, typename ST, typename CT, int SUBS>
__global__ void k_decode(StepArgs<ST, CT> A, double *stage) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= A.nf) return;
#pragma unroll
  for (int a = 0; a < Q; a++)
    stage[(long)a * A.nf + idx] = (double)Codec<ST, CT>::dec(A.F[foff<Q, NS>(SUBS, a, idx)], A.inv_scale[SUBS],
                                                             shift_of<Q, ST, CT, SUBS>(A, a));
}

// one distribution in the boundary's dense [Q][N] layout: thread n takes
// dense node n, finds its compact index from the row words (-1: solid) and
// writes its Q slots (0 at a solid node), so every element of out is
// written and nothing of an earlier call's contents survives
template <int Q, int NS, typename ST, typename CT, int SUBS>
__global__ void k_decode_dense(StepArgs<ST, CT> A, double *out, int n_nodes) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  const int r = (int)fdiv((uint32_t)n, A.G.div_ni);
  const int idx = rank_of(A.G, n - r * A.G.ni, r);
#pragma unroll
  for (int a = 0; a < Q; a++)
    out[(long)a * n_nodes + n] =
        idx < 0 ? 0.0
                : (double)Codec<ST, CT>::dec(A.F[foff<Q, NS>(SUBS, a, idx)], A.inv_scale[SUBS],
                                             shift_of<Q, ST, CT, SUBS>(A, a));
}

// rest state: f~ = w rho (3 c.u + 4.5 (c.u)^2 - 1.5 u^2) = 0 at u = 0;
// solute g = w c0 (src/latman.c:1786-1893)
template <int Q, int NS, typename ST, typename CT>
__global__ void k_init_rest(StepArgs<ST, CT> A, const double *c0) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= A.nf) return;
#pragma unroll
  for (int a = 0; a < Q; a++) {
    double v = W<Q>(a) * 1.0 * (3. * 0. + 4.5 * 0. * 0. - 1.5 * 0.);
    A.F[foff<Q, NS>(0, a, idx)] = Codec<ST, CT>::enc((CT)v, A.scale[0], (CT)0);
  }
  if (NS == 2) {
    double c = c0 ? c0[idx] : 0.;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      double v = W<Q>(a) * c * (1. + 3. * 0. + 4.5 * 0. * 0. - 1.5 * 0.);
      A.F[foff<Q, NS>(1, a, idx)] = Codec<ST, CT>::enc((CT)v, A.scale[1], shift_of<Q, ST, CT, 1>(A, a));
    }
  }
}

// ---- halo planes: k_bound_DtH_{1,2} / k_bound_HtD_{1,2} ----------------
// message layout [subs][pair q][E], pair q = directions (U+2q, U+2q+1)
struct HaloMaps {
  const int *plane_lo;  // [E] compact index of the first-plane node or -1
  const int *plane_hi;  // [E] last plane
  const int *tgt_lo;    // [pairs][E] phase-2 target in the first plane or -1
  const int *tgt_hi;    // [pairs][E] phase-2 target in the last plane or -1
  int E, pairs, U, ghost;
  int literal;          // phase 2 stores every plane position (k_bound_HtD_2)
};

template <int Q, int NS, typename ST, typename CT>
__global__ void k_halo_pack(StepArgs<ST, CT> A, HaloMaps H, int nsubs, int phase, ST *pos, ST *neg) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)nsubs * H.pairs * H.E;
  if (t >= total) return;
  int m = (int)(t % H.E);
  int q = (int)((t / H.E) % H.pairs);
  int s = (int)(t / ((long)H.E * H.pairs));
  int a = H.U + 2 * q;
  if (phase == 1) {
    int hi = H.plane_hi[m], lo = H.plane_lo[m];
    pos[t] = hi >= 0 ? A.F[foff<Q, NS>(s, a, hi)] : (ST)0;
    neg[t] = lo >= 0 ? A.F[foff<Q, NS>(s, a + 1, lo)] : (ST)0;
  } else {
    neg[t] = A.F[foff<Q, NS>(s, a, H.ghost + m)];
    pos[t] = A.F[foff<Q, NS>(s, a + 1, H.ghost + H.E + m)];
  }
}

template <int Q, int NS, typename ST, typename CT>
__global__ void k_halo_unpack(StepArgs<ST, CT> A, HaloMaps H, int nsubs, int phase, const ST *pos,
                              const ST *neg) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long total = (long)nsubs * H.pairs * H.E;
  if (t >= total) return;
  int m = (int)(t % H.E);
  int q = (int)((t / H.E) % H.pairs);
  int s = (int)(t / ((long)H.E * H.pairs));
  int a = H.U + 2 * q;
  if (phase == 1) {
    A.F[foff<Q, NS>(s, a, H.ghost + m)] = pos[t];
    A.F[foff<Q, NS>(s, a + 1, H.ghost + H.E + m)] = neg[t];
  } else {
    // the masked targets, or every fluid node of the two planes
    const int hi = H.literal ? H.plane_hi[m] : H.tgt_hi[(long)q * H.E + m];
    const int lo = H.literal ? H.plane_lo[m] : H.tgt_lo[(long)q * H.E + m];
    if (hi >= 0) A.F[foff<Q, NS>(s, a, hi)] = neg[t];
    if (lo >= 0) A.F[foff<Q, NS>(s, a + 1, lo)] = pos[t];
  }
}

// ---- FP64 reductions -----------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// kind: 0 sum, 1 max, 2 min.  Reduces K values over the block (NW waves:
// 256 threads unless stated) in a fixed order.
template <int K, int NW = 4>
__device__ void block_reduce(double (&v)[K], const int (&kind)[K], double *out) {
  __shared__ double lds[NW][K];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < K; q++) {
    double x = kind[q] == 0 ? wave_sum(v[q]) : (kind[q] == 1 ? wave_max(v[q]) : wave_min(v[q]));
    if (lane == 0) lds[wv][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    int q = threadIdx.x;
    double x = lds[0][q];
    for (int w = 1; w < NW; w++)
      x = kind[q] == 0 ? x + lds[w][q] : (kind[q] == 1 ? fmax(x, lds[w][q]) : fmin(x, lds[w][q]));
    out[q] = x;
  }
}

// per block: [0] sum(rho - rho_A) (or sum c), [1..3] sum u, [4] max|rho|,
// [5..7] max|u|, [8] min|rho|, [9..11] min|u|
template <int D>
__global__ __launch_bounds__(256) void k_reduce_mv(const double *mv, int nf, double rho_off, double *partial) {
  double v[12];
  constexpr int kinds_init[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
  int kind[12];
#pragma unroll
  for (int q = 0; q < 12; q++) {
    kind[q] = kinds_init[q];
    v[q] = q < 8 ? 0.0 : __builtin_huge_val();
  }
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < nf; idx += (long)gridDim.x * blockDim.x) {
    double r = mv[idx];
    v[0] += r - rho_off;
    v[4] = fmax(v[4], fabs(r));
    v[8] = fmin(v[8], fabs(r));
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double u = d < D ? mv[(long)(d + 1) * nf + idx] : 0.0;
      v[1 + d] += u;
      v[5 + d] = fmax(v[5 + d], fabs(u));
      v[9 + d] = fmin(v[9 + d], fabs(u));
    }
  }
  block_reduce<12>(v, kind, partial + (long)blockIdx.x * 12);
}

// Health of the stored state (lbm_check), over every stored population of
// the fluid nodes, per block: [0] sum of delta-rho, [1] sum of c, [2] least
// density rho_A + delta-rho, [3] least flow population f_a = f~_a + w_a rho_A,
// [4] least solute population g_a, [5] / [6] non-finite flow / solute
// populations, [7] / [8] negative flow / solute populations.  The counts are
// carried as doubles (exact below 2^53).
constexpr int kCheckVals = 9;

template <int Q, int NS, typename ST, typename CT>
__global__ __launch_bounds__(256) void k_check(StepArgs<ST, CT> A, double rho_A, double *partial) {
  double v[kCheckVals];
  const int kind[kCheckVals] = {0, 0, 2, 2, 2, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < kCheckVals; q++) v[q] = kind[q] == 2 ? __builtin_huge_val() : 0.0;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < A.nf; idx += gridDim.x * blockDim.x) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      double m = 0.0;
#pragma unroll
      for (int a = 0; a < Q; a++) {
        const ST raw = A.F[foff<Q, NS>(s, a, idx)];
        const double x = s == 0 ? (double)Codec<ST, CT>::dec(raw, A.inv_scale[0], (CT)0)
                                : (double)Codec<ST, CT>::dec(raw, A.inv_scale[1], shift_of<Q, ST, CT, 1>(A, a));
        const double pop = s == 0 ? x + W<Q>(a) * rho_A : x;
        if (!__builtin_isfinite(x)) {
          v[5 + s] += 1.0;
        } else {
          if (pop < 0.0) v[7 + s] += 1.0;
          v[3 + s] = fmin(v[3 + s], pop);
        }
        m += x;
      }
      v[s] += m;
      if (s == 0 && __builtin_isfinite(m)) v[2] = fmin(v[2], rho_A + m);
    }
  }
  block_reduce<kCheckVals>(v, kind, partial + (long)blockIdx.x * kCheckVals);
}

// solute moments of one plane: sum_a c_a,last g_a and sum_a g_a on the
// populations the coming half-step collides.  Stage 1: kFluxBlocks blocks
// stride over the plane's nodes and leave per-block partial sums; stage 2:
// one block adds them in a fixed order (deterministic) and optionally adds
// the flux to a running sum.
constexpr int kFluxBlocks = 64;

template <int Q, int NS, typename ST, typename CT>
__global__ __launch_bounds__(256) void k_plane_flux(BcArgs<ST, CT> B, double *partial) {
  double v[2] = {0.0, 0.0};
  const int kind[2] = {0, 0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < B.n; t += gridDim.x * blockDim.x) {
    const int idx = B.list[t];
    int nb[Q];
    CT f[Q];
    bc_gather<Q, NS, ST, CT, 1>(B, idx, nb, f);
    double jl = 0.0, m = 0.0;
#pragma unroll
    for (int a = 0; a < Q; a++) {
      jl += (double)(Q == 9 ? CY(a) : CZ(a)) * (double)f[a];
      m += (double)f[a];
    }
    v[0] += jl;
    v[1] += m;
  }
  block_reduce<2>(v, kind, partial + 2 * blockIdx.x);
}

__global__ __launch_bounds__(64) void k_plane_flux_final(const double *partial, double *out, double *acc) {
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < kFluxBlocks; q++) {
      a += partial[2 * q];
      b += partial[2 * q + 1];
    }
    out[0] = a;
    out[1] = b;
    if (acc) acc[0] += a;
  }
}

// acc[0] += sum of x[0..n), in a fixed order: per-block partial sums, then
// one thread adds them in block order
__global__ __launch_bounds__(256) void k_sum_partial(const double *x, long n, double *partial) {
  double v[1] = {0.0};
  const int kind[1] = {0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v[0] += x[i];
  block_reduce<1>(v, kind, partial + blockIdx.x);
}

__global__ __launch_bounds__(64) void k_sum_final(const double *partial, int nb, double *acc) {
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int q = 0; q < nb; q++) a += partial[q];
    acc[0] += a;
  }
}

// plain copy of n doubles: the bandwidth (and HBM byte counter) reference
__global__ __launch_bounds__(256) void k_copy_probe(const double *__restrict__ src, double *__restrict__ dst,
                                                    long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace lbm
