// CPU check of the device index arithmetic (lbgpu-prime-dev_amd/csrc/lbm_index.hpp):
// fast division, coordinates, and the bitmask/popcount neighbour ranks
// against a brute-force dense -> compact map, in 2-D and 3-D, periodic and
// halo modes.  Prints "ok" or the first mismatch; exit status 0 on success.
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "lbm_index.hpp"

using namespace lbm;

static const int CXv[19] = {0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0};
static const int CYv[19] = {0, 0, 0, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1};
static const int CZv[19] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};

static int check_div() {
  std::mt19937_64 rng(1);
  for (uint32_t d : {1u, 2u, 3u, 5u, 7u, 64u, 100u, 127u, 128u, 192u, 255u, 256u, 448u, 1000u, 4096u, 65535u}) {
    FastDiv f = make_fastdiv(d);
    for (int t = 0; t < 200000; t++) {
      uint32_t n = t < 100000 ? (uint32_t)t : (uint32_t)(rng() & 0x7fffffffu);
      if (fdiv(n, f) != n / d) {
        printf("fastdiv mismatch n=%u d=%u\n", n, d);
        return 1;
      }
    }
  }
  return 0;
}

template <int D>
static int check_case(int ni, int nj, int nk, int halo, double solid, unsigned seed) {
  const int Q = D == 2 ? 9 : 19;
  const long N = (long)ni * nj * (D == 3 ? nk : 1);
  const long E = D == 2 ? ni : (long)ni * nj;
  const int L = D == 2 ? nj : nk;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> U(0, 1);
  std::vector<int> fid(N), node;
  for (long n = 0; n < N; n++) {
    fid[n] = U(rng) < solid ? -1 : (int)node.size();
    if (fid[n] >= 0) node.push_back((int)n);
  }
  const int nf = (int)node.size();
  const int W = (ni + 63) / 64;
  const long rows = N / ni;
  std::vector<RowWord> rw(rows * W);
  fill_row_words(fid.data(), rows, ni, rw.data());
  Geo G;
  G.rw = rw.data();
  G.node = node.data();
  G.ni = ni;
  G.nj = nj;
  G.nk = D == 3 ? nk : 1;
  G.W = W;
  G.halo = halo;
  G.ghost = nf;
  G.E = (int)E;
  G.div_ni = make_fastdiv(ni);
  G.div_nj = make_fastdiv(nj);
  for (int x = 0; x < nf; x++) {
    long n = node[x];
    int i, j, k;
    coords<D>(G, (uint32_t)n, i, j, k);
    int ri = (int)(n % ni), rj = (int)((n / ni) % nj), rk = D == 2 ? 0 : (int)(n / E);
    if (i != ri || j != rj || k != rk) {
      printf("coords mismatch n=%ld\n", n);
      return 1;
    }
    int rows_nb[18];
    if (D == 2) {
      int t[8];
      neighbors_rows<2>(G, i, j, k, t);
      for (int q = 0; q < 8; q++) rows_nb[q] = t[q];
    } else {
      int t[18];
      neighbors_rows<3>(G, i, j, k, t);
      for (int q = 0; q < 18; q++) rows_nb[q] = t[q];
    }
    for (int a = 1; a < Q; a++) {
      int ii = ((i - CXv[a]) % ni + ni) % ni;
      int jj = j - CYv[a], kk = k - CZv[a];
      int last = D == 2 ? jj : kk;
      if (D == 3) jj = ((jj % nj) + nj) % nj;
      int expect;
      if (last >= 0 && last < L) {
        expect = fid[D == 2 ? ii + (long)ni * last : ii + (long)ni * jj + E * last];
      } else {
        int lw = last < 0 ? last + L : last - L;
        long p = D == 2 ? ii : ii + (long)ni * jj;
        int img = fid[p + E * lw];
        if (!halo) expect = img;
        else expect = img < 0 ? -1 : nf + (last < 0 ? 0 : (int)E) + (int)p;
      }
      int got = neighbor<D>(G, i, j, k, -CXv[a], -CYv[a], -CZv[a]);
      int got2 = rows_nb[a - 1];
      if (got != expect || got2 != expect) {
        printf("neighbour mismatch D=%d halo=%d n=%ld a=%d got %d/%d expect %d\n", D, halo, n, a, got, got2,
               expect);
        return 1;
      }
    }
  }
  // tile offsets are a bijection onto [0, tiles*NS*Q*64)
  const long slots = nf + (halo ? 2 * E : 0);
  const long tiles = (slots + kTile - 1) / kTile;
  std::vector<char> seen(tiles * 2 * Q * kTile, 0);
  for (long x = 0; x < slots; x++)
    for (int s = 0; s < 2; s++)
      for (int a = 0; a < Q; a++) {
        uint32_t o = D == 2 ? foff<9, 2>(s, a, (uint32_t)x) : foff<19, 2>(s, a, (uint32_t)x);
        if (o >= seen.size() || seen[o]) {
          printf("tile offset collision\n");
          return 1;
        }
        seen[o] = 1;
      }
  return 0;
}

int main() {
  if (check_div()) return 1;
  int bad = 0;
  for (int halo = 0; halo < 2; halo++) {
    bad |= check_case<2>(17, 13, 1, halo, 0.3, 1);
    bad |= check_case<2>(128, 9, 1, halo, 0.5, 2);
    bad |= check_case<2>(70, 5, 1, halo, 0.1, 3);
    bad |= check_case<3>(9, 7, 6, halo, 0.3, 4);
    bad |= check_case<3>(64, 5, 4, halo, 0.4, 5);
    bad |= check_case<3>(130, 3, 5, halo, 0.2, 6);
    bad |= check_case<3>(1, 1, 3, halo, 0.0, 7);
    bad |= check_case<3>(65, 4, 5, halo, 0.35, 8);
    bad |= check_case<3>(200, 3, 4, halo, 0.3, 9);
    bad |= check_case<2>(129, 6, 1, halo, 0.3, 10);
    bad |= check_case<3>(3, 2, 5, halo, 0.2, 11);
  }
  if (bad) return 1;
  printf("ok\n");
  return 0;
}
