// Index arithmetic shared by the host set-up, the device kernels and the
// CPU index test (tests/index_check.cpp compiles this header with g++).
//
// Populations live in tiles of kTile consecutive fluid nodes:
//   one distribution:  element (slot a, node x) at
//     (x / kTile) * Q*kTile + a * kTile + x % kTile
//   flow + solute:     element (subs s, slot a, node x) at
//     (x / kTile) * 2*Q*kTile + (a * kTile + x % kTile) * 2 + s
// so that one slot of a wave's 64 nodes is one contiguous 512-byte run
// (FP64), consecutive waves continue the same runs, and the flow and solute
// populations of a slot sit side by side (one 2-wide load or store moves
// both).  kTile = 2048 (a tile spans eight workgroups): against 64-node
// tiles (one wave each) it raised the even half-step from 73% to 79% of
// HBM peak at 256^3 and the odd one by 1-2%; 512, 1024 and 4096 measured
// between the two or, for flow + solute, below.
//
// Neighbours are not stored per node.  Each lattice row (fixed j, k) is
// split in 64-node words; for word w of row r, RowWord holds the fluid
// bitmask (bit b set when node 64w+b is fluid) and the compact index of
// the first fluid node at or after 64w.  The compact index of a fluid node
// is then base + popcount(bits below it), read from a few MB that stay in
// cache, instead of 72 bytes of neighbour table per node from HBM.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LBM_HD __host__ __device__ __forceinline__
#else
#define LBM_HD inline
#endif

namespace lbm {

constexpr int kTile = 2048;

// Unsigned 32-bit division by an invariant d >= 1 (Granlund & Montgomery,
// PLDI 1994, Fig. 4.1): q = (t + ((n - t) >> s1)) >> s2, t = mulhi(m, n).
struct FastDiv {
  uint32_t d, m, s1, s2;
};

LBM_HD FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) l++;
  FastDiv f;
  f.d = d;
  f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 0 ? l - 1 : 0;
  return f;
}

LBM_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

LBM_HD uint32_t fdiv(uint32_t n, const FastDiv &f) {
  uint32_t t = mulhi32(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

LBM_HD int popc64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popcll(v);
#else
  return __builtin_popcountll(v);
#endif
}

struct alignas(16) RowWord {
  uint64_t bits;  // fluid bitmask of nodes 64w .. 64w+63 of the row
  int32_t base;   // compact index of the first fluid node at or after 64w
  int16_t dprev;  // rank of the row's node before 64w (periodic) - base, or kSolidRank
  int16_t dnext;  // rank of the row's node after the word (periodic) - base, or kSolidRank
};
constexpr int16_t kSolidRank = -32768;

struct Geo {
  const RowWord *rw;  // [rows][W]
  const int *node;    // compact index -> dense index
  int ni, nj, nk;     // nk = 1 in 2-D
  int W;              // words per row
  int halo;           // last axis through ghost entries
  int ghost;          // first ghost compact index (= number of fluid nodes)
  int E;              // nodes per plane across the last axis
  FastDiv div_ni, div_nj;
};

template <int Q, int NS>
LBM_HD uint32_t foff(int s, int a, uint32_t x) {
  if (NS == 2) return (x / kTile) * (uint32_t)(2 * Q * kTile) + (uint32_t)(a * kTile + (x % kTile)) * 2u + (uint32_t)s;
  return (x / kTile) * (uint32_t)(Q * kTile) + (uint32_t)(a * kTile) + (x % kTile);
}

// fluid test and compact index of node (i, row r): one 16-byte load and
// no branch, so the loads of all neighbours can be in flight together
LBM_HD int rank_of(const Geo &G, int i, int r) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 v = *reinterpret_cast<const uint4 *>(G.rw + ((long)r * G.W + (i >> 6)));
  const uint64_t bits = (uint64_t)v.x | ((uint64_t)v.y << 32);
  const int base = (int)v.z;
#else
  const RowWord &w = G.rw[(long)r * G.W + (i >> 6)];
  const uint64_t bits = w.bits;
  const int base = w.base;
#endif
  const int b = i & 63;
  const int idx = base + popc64(bits & ((1ull << b) - 1ull));
  return ((bits >> b) & 1ull) ? idx : -1;
}

// coordinates of a dense node index: i fastest, then j, then k
template <int D>
LBM_HD void coords(const Geo &G, uint32_t n, int &i, int &j, int &k) {
  uint32_t r = fdiv(n, G.div_ni);
  i = (int)(n - r * (uint32_t)G.ni);
  if (D == 2) {
    j = (int)r;
    k = 0;
  } else {
    uint32_t q = fdiv(r, G.div_nj);
    j = (int)(r - q * (uint32_t)G.nj);
    k = (int)q;
  }
}

// compact index of (i+dx, j+dy, k+dz): periodic in every axis, or across
// the last axis through ghost entries in halo mode; -1 when solid.
// Ghost entries [ghost, ghost+E) stand for the plane before the first one,
// [ghost+E, ghost+2E) for the plane after the last one; their solidity is
// that of the periodic image.  Written with selects only, so the compiler
// can issue the row-word loads of all directions back to back.
LBM_HD int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <int D>
LBM_HD int neighbor(const Geo &G, int i, int j, int k, int dx, int dy, int dz) {
  const int ii = wrap1(i + dx, G.ni);
  int last, lastw, row, p;
  if (D == 2) {
    last = j + dy;
    lastw = wrap1(last, G.nj);
    row = lastw;
    p = ii;
  } else {
    const int jj = wrap1(j + dy, G.nj);
    last = k + dz;
    lastw = wrap1(last, G.nk);
    row = jj + G.nj * lastw;
    p = ii + G.ni * jj;
  }
  const int r = rank_of(G, ii, row);
  const int gh = G.ghost + (last < 0 ? 0 : G.E) + p;
  return (G.halo && last != lastw && r >= 0) ? gh : r;
}

// The same for several directions at once, in two phases: every row-word
// load is issued before any rank is formed, so their latencies overlap.
template <int D, int K>
LBM_HD void neighbors_of(const Geo &G, int i, int j, int k, const int (&dx)[K], const int (&dy)[K],
                         const int (&dz)[K], int (&out)[K]) {
  int col[K], gh[K];
  bool cross[K];
  const RowWord *ptr[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int ii = wrap1(i + dx[q], G.ni);
    int last, lastw, row, p;
    if (D == 2) {
      last = j + dy[q];
      lastw = wrap1(last, G.nj);
      row = lastw;
      p = ii;
    } else {
      const int jj = wrap1(j + dy[q], G.nj);
      last = k + dz[q];
      lastw = wrap1(last, G.nk);
      row = jj + G.nj * lastw;
      p = ii + G.ni * jj;
    }
    col[q] = ii & 63;
    ptr[q] = G.rw + ((long)row * G.W + (ii >> 6));
    cross[q] = G.halo && last != lastw;
    gh[q] = G.ghost + (last < 0 ? 0 : G.E) + p;
  }
  uint64_t bits[K];
  int base[K];
#pragma unroll
  for (int q = 0; q < K; q++) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 v = *reinterpret_cast<const uint4 *>(ptr[q]);
    bits[q] = (uint64_t)v.x | ((uint64_t)v.y << 32);
    base[q] = (int)v.z;
#else
    bits[q] = ptr[q]->bits;
    base[q] = ptr[q]->base;
#endif
  }
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int b = col[q];
    const int r = ((bits[q] >> b) & 1ull) ? base[q] + popc64(bits[q] & ((1ull << b) - 1ull)) : -1;
    out[q] = (cross[q] && r >= 0) ? gh[q] : r;
  }
}

// All neighbours of one node, one row-word load per lattice row: the
// nodes i-1, i, i+1 of a row come from the word holding i, its bitmask and
// the ranks of the row's nodes just outside the word (dprev / dnext, which
// also carry the periodic wrap at the row ends).  Every load is issued
// before any rank is formed.  out[a-1] = compact index of x - c_a.
template <int D>
LBM_HD void neighbors_rows(const Geo &G, int i, int j, int k, int (&out)[D == 2 ? 8 : 18]) {
  // lattice rows (dy, dz) of the nodes x - c_a, in the reference order
  constexpr int R = D == 2 ? 3 : 9;
  constexpr int rdy[9] = {0, -1, 1, 0, 0, -1, 1, 1, -1};
  constexpr int rdz[9] = {0, 0, 0, -1, 1, -1, 1, -1, 1};
  const int b = i & 63, w = i >> 6;
  const int wend = (G.ni - 64 * w) < 64 ? (G.ni - 64 * w) : 64;  // valid bits of the word
  uint64_t bits[R];
  int base[R], dp[R], dn[R], gh[R];
  bool cross[R];
  const RowWord *ptr[R];
#pragma unroll
  for (int q = 0; q < R; q++) {
    int last, lastw, row, p;
    if (D == 2) {
      last = j + rdy[q];
      lastw = wrap1(last, G.nj);
      row = lastw;
      p = 0;
    } else {
      const int jj = wrap1(j + rdy[q], G.nj);
      last = k + rdz[q];
      lastw = wrap1(last, G.nk);
      row = jj + G.nj * lastw;
      p = G.ni * jj;
    }
    ptr[q] = G.rw + ((long)row * G.W + w);
    cross[q] = G.halo && last != lastw;
    gh[q] = G.ghost + (last < 0 ? 0 : G.E) + p;  // + the column below
  }
#pragma unroll
  for (int q = 0; q < R; q++) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 v = *reinterpret_cast<const uint4 *>(ptr[q]);
    bits[q] = (uint64_t)v.x | ((uint64_t)v.y << 32);
    base[q] = (int)v.z;
    dp[q] = (int)(int16_t)(v.w & 0xffffu);
    dn[q] = (int)(int16_t)(v.w >> 16);
#else
    bits[q] = ptr[q]->bits;
    base[q] = ptr[q]->base;
    dp[q] = ptr[q]->dprev;
    dn[q] = ptr[q]->dnext;
#endif
  }
  // rank of column offset dx in row q
  auto rank = [&](int q, int dx) -> int {
    int r;
    if (dx == 0) {
      r = ((bits[q] >> b) & 1ull) ? base[q] + popc64(bits[q] & ((1ull << b) - 1ull)) : -1;
    } else if (dx < 0) {
      const int bb = b - 1;
      const int in = ((bits[q] >> (bb & 63)) & 1ull) ? base[q] + popc64(bits[q] & ((1ull << (bb & 63)) - 1ull)) : -1;
      r = b > 0 ? in : (dp[q] == kSolidRank ? -1 : base[q] + dp[q]);
    } else {
      const int bb = b + 1;
      const int in = ((bits[q] >> (bb & 63)) & 1ull) ? base[q] + popc64(bits[q] & ((1ull << (bb & 63)) - 1ull)) : -1;
      r = bb < wend ? in : (dn[q] == kSolidRank ? -1 : base[q] + dn[q]);
    }
    const int col = wrap1(i + dx, G.ni);
    return (cross[q] && r >= 0) ? gh[q] + col : r;
  };
  if (D == 2) {
    // a: 1 E, 2 W, 3 N, 4 S, 5 NE, 6 SW, 7 NW, 8 SE; x - c_a
    out[0] = rank(0, -1);
    out[1] = rank(0, 1);
    out[2] = rank(1, 0);
    out[3] = rank(2, 0);
    out[4] = rank(1, -1);
    out[5] = rank(2, 1);
    out[6] = rank(1, 1);
    out[7] = rank(2, -1);
  } else {
    out[0] = rank(0, -1);   // 1  (+1, 0, 0)
    out[1] = rank(0, 1);    // 2  (-1, 0, 0)
    out[2] = rank(1, 0);    // 3  (0, +1, 0)
    out[3] = rank(2, 0);    // 4  (0, -1, 0)
    out[4] = rank(1, -1);   // 5  (+1, +1, 0)
    out[5] = rank(2, 1);    // 6  (-1, -1, 0)
    out[6] = rank(1, 1);    // 7  (-1, +1, 0)
    out[7] = rank(2, -1);   // 8  (+1, -1, 0)
    out[8] = rank(3, 0);    // 9  (0, 0, +1)
    out[9] = rank(4, 0);    // 10 (0, 0, -1)
    out[10] = rank(3, -1);  // 11 (+1, 0, +1)
    out[11] = rank(4, 1);   // 12 (-1, 0, -1)
    out[12] = rank(3, 1);   // 13 (-1, 0, +1)
    out[13] = rank(4, -1);  // 14 (+1, 0, -1)
    out[14] = rank(5, 0);   // 15 (0, +1, +1)
    out[15] = rank(6, 0);   // 16 (0, -1, -1)
    out[16] = rank(7, 0);   // 17 (0, -1, +1)
    out[17] = rank(8, 0);   // 18 (0, +1, -1)
  }
}

// Host: fill the row words from the dense -> compact map (fid[n] = -1 for
// solid).  base is the number of fluid nodes before the word in dense
// order, so base(word) + popcount(word) = base(next word) everywhere.
inline void fill_row_words(const int *fid, long rows, int ni, RowWord *rw) {
  const int W = (ni + 63) / 64;
  long count = 0;
  for (long r = 0; r < rows; r++)
    for (int w = 0; w < W; w++) {
      RowWord &x = rw[r * W + w];
      x.bits = 0;
      x.base = (int32_t)count;
      for (int b = 0; b < 64 && 64 * w + b < ni; b++)
        if (fid[r * ni + 64 * w + b] >= 0) {
          x.bits |= 1ull << b;
          count++;
        }
      const int pc = w > 0 ? 64 * w - 1 : ni - 1;
      const int nc = 64 * (w + 1) < ni ? 64 * (w + 1) : 0;
      const int pid = fid[r * ni + pc], nid = fid[r * ni + nc];
      x.dprev = pid < 0 ? kSolidRank : (int16_t)(pid - x.base);
      x.dnext = nid < 0 ? kSolidRank : (int16_t)(nid - x.base);
    }
}

}  // namespace lbm
