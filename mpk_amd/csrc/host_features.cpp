// host_features.cpp -- matrix features for the learned tuning model (no
// reference counterpart).  24 numbers describe the row-length distribution,
// the column spread (bandwidth, locality), the skirt growth of row blocks
// per BFS level, and the right-hand-side block size.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

using namespace mpkx;

namespace {

double lg2(double v) { return std::log2(std::max(v, 1.0)); }

// |ball_d(S)| / |S| for d = 1..depth, S = rows [s0, s0 + len).
void skirt_growth(int n, const int *ptr, const int *col, int s0, int len, int depth, std::vector<int> &dist,
                  std::vector<int> &frontier, double *ratio) {
  std::vector<int> touched;
  frontier.clear();
  for (int i = s0; i < s0 + len; i++) {
    dist[i] = 0;
    frontier.push_back(i);
    touched.push_back(i);
  }
  long count = len;
  std::vector<int> next;
  for (int d = 1; d <= depth; d++) {
    next.clear();
    for (int i : frontier)
      for (int j = ptr[i]; j < ptr[i + 1]; j++) {
        int c = col[j];
        if (dist[c] < 0) {
          dist[c] = d;
          next.push_back(c);
          touched.push_back(c);
        }
      }
    count += (long)next.size();
    ratio[d - 1] += (double)count / len;
    frontier.swap(next);
  }
  for (int i : touched) dist[i] = -1;
  (void)n;
}

}  // namespace

extern "C" {

int mpkx_features(int n, const int *ptr, const int *col, int nrhs, double *feat) {
  clear_error();
  if (n <= 0 || !ptr || !col || nrhs <= 0 || !feat) return fail(MPKX_EINVAL, "features: bad arguments");
  const long nnz = ptr[n];
  std::vector<int> len(n);
  double mean = (double)nnz / n, var = 0;
  long big32 = 0, big256 = 0, nz32 = 0, nz256 = 0;
  double band_sum = 0, band_max = 0;
  long near32 = 0, near4k = 0;
  for (int i = 0; i < n; i++) {
    int d = ptr[i + 1] - ptr[i];
    if (d < 0) return fail(MPKX_EINVAL, "features: decreasing ptr at row %d", i);
    len[i] = d;
    var += (d - mean) * (d - mean);
    if (d > 32) big32++, nz32 += d;
    if (d > 256) big256++, nz256 += d;
    for (int j = ptr[i]; j < ptr[i + 1]; j++) {
      if (col[j] < 0 || col[j] >= n) return fail(MPKX_EINVAL, "features: column out of range in row %d", i);
      double b = std::fabs((double)col[j] - i);
      band_sum += b;
      band_max = std::max(band_max, b);
      near32 += b <= 32;
      near4k += b <= 4096;
    }
  }
  std::vector<int> sorted(len);
  std::sort(sorted.begin(), sorted.end());
  auto q = [&](double f) { return (double)sorted[std::min((long)(f * (n - 1) + 0.5), (long)n - 1)]; };
  const double nz = std::max<long>(nnz, 1);
  double f[24] = {0};
  f[0] = lg2(n);
  f[1] = lg2((double)nnz);
  f[2] = mean;
  f[3] = std::sqrt(var / n) / std::max(mean, 1e-9);
  f[4] = lg2(sorted[n - 1]);
  f[5] = sorted[0];
  f[6] = q(0.5);
  f[7] = q(0.9);
  f[8] = q(0.99);
  f[9] = (double)big32 / n;
  f[10] = (double)big256 / n;
  f[11] = nz32 / nz;
  f[12] = nz256 / nz;
  f[13] = lg2(1.0 + band_sum / nz);
  f[14] = lg2(1.0 + band_max);
  f[15] = near32 / nz;
  f[16] = near4k / nz;
  // skirt growth of 4 row blocks (evenly spaced), depths 1..4, log2 of the mean ratio
  const int blk = std::max(1, std::min(1024, n / 4));
  std::vector<int> dist(n, -1), frontier;
  double ratio[4] = {0, 0, 0, 0};
  for (int t = 0; t < 4; t++) {
    int s0 = (int)((long)t * (n - blk) / 3);
    skirt_growth(n, ptr, col, s0, blk, 4, dist, frontier, ratio);
  }
  for (int d = 0; d < 4; d++) f[17 + d] = std::log2(ratio[d] / 4.0);
  f[21] = lg2(nrhs);
  f[22] = lg2(4.0 * n * nrhs);
  f[23] = lg2(8.0 * nnz + 4.0 * n);
  for (int i = 0; i < 24; i++) feat[i] = f[i];
  return MPKX_OK;
}

}  // extern "C"
