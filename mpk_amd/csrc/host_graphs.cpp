// host_graphs.cpp -- synthetic matrix generators for the reduced-precision,
// batched and learned-tuning workloads (no reference counterpart: the
// reference only generates stencils, mpk2/gen2.c).
//
// Chung-Lu random graphs (Chung & Lu, "The average distances in random
// graphs with given expected degrees", PNAS 2002): vertex i gets expected
// degree w_i = c (i+1)^(-1/(exponent-1)), edges are drawn by sampling both
// endpoints with probability proportional to w, duplicates and self loops
// dropped, the pattern symmetrised, plus a ring i -- i+1 that keeps the graph
// connected.  The number of sampled edges is (mean_nnz - 3) n / 2, so a row
// holds about mean_nnz entries (diagonal and ring included; duplicate hub
// edges make it slightly fewer).  Values: N(0,1) off the diagonal (Box-Muller
// on the SplitMix64 stream), diagonal = row abs-sum + `diag_shift` (strictly
// diagonally dominant).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace mpkx;

namespace {

struct Rng {  // SplitMix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

extern "C" {

int mpkx_gen_chunglu(int n, double mean_nnz, double exponent, double diag_shift, uint64_t seed,
                     int **ptr_out, int **col_out, double **val_out) {
  clear_error();
  if (n <= 2 || mean_nnz < 3.0 || exponent <= 2.0 || !ptr_out || !col_out || !val_out)
    return fail(MPKX_EINVAL, "gen_chunglu: bad arguments");
  Rng rng{seed};
  // expected degrees w_i = c * (i + i0)^(-1/(exponent-1)), scaled to the mean
  double beta = 1.0 / (exponent - 1.0), i0 = 1.0;
  std::vector<double> w(n), cum(n + 1, 0.0);
  double sum = 0;
  for (int i = 0; i < n; i++) {
    w[i] = std::pow(i + i0, -beta);
    sum += w[i];
  }
  double scale = (mean_nnz - 3.0) * n / sum;
  for (int i = 0; i < n; i++) {
    w[i] *= scale;
    cum[i + 1] = cum[i] + w[i];
  }
  long m = (long)((mean_nnz - 3.0) * n / 2.0);
  std::vector<std::pair<int, int>> edges;
  try {
    edges.reserve((size_t)m * 2);
  } catch (const std::bad_alloc &) {
    return fail(MPKX_ENOMEM, "gen_chunglu: allocation");
  }
  auto pick = [&]() {
    double u = rng.uniform() * cum[n];
    int k = (int)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin()) - 1;
    return std::min(std::max(k, 0), n - 1);
  };
  for (long e = 0; e < m; e++) {
    int a = pick(), b = pick();
    if (a == b) continue;
    edges.push_back({a, b});
    edges.push_back({b, a});
  }
  // a ring keeps every row non-empty and the graph connected
  for (int i = 0; i < n; i++) {
    int j = (i + 1) % n;
    edges.push_back({i, j});
    edges.push_back({j, i});
  }
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  std::vector<int> rp(n + 1, 0), cl;
  std::vector<double> va;
  cl.reserve(edges.size() + n);
  va.reserve(edges.size() + n);
  size_t e = 0;
  for (int i = 0; i < n; i++) {
    size_t d0 = cl.size();
    double abs_sum = 0;
    bool diag_done = false;
    while (e < edges.size() && edges[e].first == i) {
      int j = edges[e].second;
      if (!diag_done && j > i) {
        cl.push_back(i);
        va.push_back(0.0);
        diag_done = true;
      }
      double v = rng.normal();
      cl.push_back(j);
      va.push_back(v);
      abs_sum += std::fabs(v);
      e++;
    }
    if (!diag_done) {
      cl.push_back(i);
      va.push_back(0.0);
    }
    for (size_t t = d0; t < cl.size(); t++)
      if (cl[t] == i) va[t] = abs_sum + diag_shift;
    rp[i + 1] = (int)cl.size();
  }
  int *p = (int *)malloc(sizeof(int) * (n + 1));
  int *c = (int *)malloc(sizeof(int) * (cl.size() + 1));
  double *v = (double *)malloc(sizeof(double) * (va.size() + 1));
  if (!p || !c || !v) {
    free(p), free(c), free(v);
    return fail(MPKX_ENOMEM, "gen_chunglu: allocation");
  }
  memcpy(p, rp.data(), sizeof(int) * (n + 1));
  memcpy(c, cl.data(), sizeof(int) * cl.size());
  memcpy(v, va.data(), sizeof(double) * va.size());
  *ptr_out = p;
  *col_out = c;
  *val_out = v;
  return MPKX_OK;
}

// Banded matrix: row i couples to the `per_row` columns i + offset for
// offsets spread evenly over [-bandwidth, bandwidth] (clipped at the
// borders); N(0,1) off-diagonals, diagonal = abs-sum + diag_shift.
int mpkx_gen_banded(int n, int bandwidth, int per_row, double diag_shift, uint64_t seed, int **ptr_out,
                    int **col_out, double **val_out) {
  clear_error();
  if (n <= 1 || bandwidth < 1 || per_row < 1 || !ptr_out || !col_out || !val_out)
    return fail(MPKX_EINVAL, "gen_banded: bad arguments");
  Rng rng{seed};
  std::vector<int> offs;
  for (int t = 0; t < per_row; t++) {
    long o = -bandwidth + (long)(2L * bandwidth * t) / std::max(1, per_row - 1);
    if (o != 0) offs.push_back((int)o);
  }
  offs.push_back(0);
  std::sort(offs.begin(), offs.end());
  offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
  std::vector<int> rp(n + 1, 0), cl;
  std::vector<double> va;
  for (int i = 0; i < n; i++) {
    size_t d0 = cl.size();
    double abs_sum = 0;
    for (int o : offs) {
      long j = (long)i + o;
      if (j < 0 || j >= n) continue;
      double v = o == 0 ? 0.0 : rng.normal();
      cl.push_back((int)j);
      va.push_back(v);
      abs_sum += std::fabs(v);
    }
    for (size_t t = d0; t < cl.size(); t++)
      if (cl[t] == i) va[t] = abs_sum + diag_shift;
    rp[i + 1] = (int)cl.size();
  }
  int *p = (int *)malloc(sizeof(int) * (n + 1));
  int *c = (int *)malloc(sizeof(int) * (cl.size() + 1));
  double *v = (double *)malloc(sizeof(double) * (va.size() + 1));
  if (!p || !c || !v) {
    free(p), free(c), free(v);
    return fail(MPKX_ENOMEM, "gen_banded: allocation");
  }
  memcpy(p, rp.data(), sizeof(int) * (n + 1));
  memcpy(c, cl.data(), sizeof(int) * cl.size());
  memcpy(v, va.data(), sizeof(double) * va.size());
  *ptr_out = p;
  *col_out = c;
  *val_out = v;
  return MPKX_OK;
}

}  // extern "C"
