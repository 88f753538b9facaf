// host_levels.cpp -- partitions, levels, skirt and the assignment solver.
//
// The reference computes levels and skirts by iterating a min-plus product
// to a fixpoint once per partition (mpk2/lib.c:183-219, :330-412, :584-615),
// O(P * diameter * nnz).  The same fixpoints are shortest-path distances, so
// here they are computed with one multi-source breadth-first search over
// all partitions at once (levels) and one depth-bounded search per
// partition (skirt); results are identical value for value.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "internal.h"

using namespace mpkx;

namespace {

// Multi-source unit-weight shortest paths restricted to edges inside one
// partition.  dist[] holds the start values (BIG where unseeded) and is
// lowered in place.
void restricted_bfs(int n, const int *ptr, const int *col, const int *part,
                    std::vector<long> &dist, long big) {
  std::vector<std::pair<long, int>> src;
  for (int i = 0; i < n; i++)
    if (dist[i] < big) src.push_back({dist[i], i});
  std::sort(src.begin(), src.end());
  std::vector<std::pair<long, int>> q;
  q.reserve(n);
  size_t qh = 0, si = 0;
  while (qh < q.size() || si < src.size()) {
    std::pair<long, int> u;
    if (si < src.size() && (qh >= q.size() || src[si].first <= q[qh].first))
      u = src[si++];
    else
      u = q[qh++];
    if (u.first > dist[u.second]) continue;  // stale entry
    int i = u.second;
    long nd = u.first + 1;
    for (int j = ptr[i]; j < ptr[i + 1]; j++) {
      int c = col[j];
      if (part[c] == part[i] && nd < dist[c]) {
        dist[c] = nd;
        q.push_back({nd, c});
      }
    }
  }
}

}  // namespace

extern "C" {

int mpkx_levels(int n, const int *ptr, const int *col, const int *part,
                const int *lorg, int *level) {
  clear_error();
  if (n <= 0 || !ptr || !col || !part || !level) return fail(MPKX_EINVAL, "levels: bad arguments");
  long big;
  if (lorg) {
    long mx = lorg[0];
    for (int i = 1; i < n; i++) mx = std::max<long>(mx, lorg[i]);
    big = mx + n + 1;  // update_level's "big enough", mpk2/lib.c:397
  } else {
    big = (long)n + 1;  // comp_level's "big enough", mpk2/lib.c:349
  }
  std::vector<long> dist;
  MPKX_TRY_ALLOC(dist.assign(n, big));
  // Rows adjacent to another partition start from that neighbour's fixed
  // value + 1 (-1 for comp_level, the previous level for update_level).
  for (int i = 0; i < n; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) {
      int c = col[j];
      if (part[c] != part[i]) {
        long v = (lorg ? (long)lorg[c] : -1L) + 1;
        if (v < dist[i]) dist[i] = v;
      }
    }
  restricted_bfs(n, ptr, col, part, dist, big);
  for (int i = 0; i < n; i++) {
    if (dist[i] > INT_MAX) return fail(MPKX_EINVAL, "levels: value overflow");
    level[i] = (int)dist[i];
  }
  return MPKX_OK;
}

int mpkx_level2wcrs(int n, const int *ptr, const int *col, const int *level, int *we) {
  clear_error();
  int mn = level[0];
  for (int i = 1; i < n; i++) mn = std::min(mn, level[i]);
  for (int i = 0; i < n; i++)
    for (int j = ptr[i]; j < ptr[i + 1]; j++) {
      double w = level[i] + level[col[j]] - 2 * mn;
      w = 1e+6 / (w + 1);
      we[j] = w < 1.0 ? 1 : (int)w;
    }
  return MPKX_OK;
}

int mpkx_comp_skirt(int n, const int *ptr, const int *col, const int *part,
                    int npart, int nlevel, const int *lastl, long **sptr_out,
                    int **srow_out, int **sdist_out) {
  clear_error();
  if (n <= 0 || npart <= 0 || nlevel <= 0) return fail(MPKX_EINVAL, "comp_skirt: bad arguments");
  std::vector<std::vector<int>> members(npart);
  for (int i = 0; i < n; i++) {
    if (part[i] < 0 || part[i] >= npart) return fail(MPKX_EINVAL, "comp_skirt: label %d out of range", part[i]);
    members[part[i]].push_back(i);
  }
  long lmin = 0;
  if (lastl)
    for (int i = 0; i < n; i++) lmin = std::min<long>(lmin, lastl[i]);
  // The filter keeps d < nlevel - lastl[j] (or d < nlevel): no kept row is
  // farther than `bound`.
  long bound = (long)nlevel - 1 - lmin;
  std::vector<std::vector<std::pair<int, int>>> per(npart);
  int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nthreads > npart) nthreads = npart;
  std::atomic<long> next{0};
  std::atomic<bool> oom{false};
  auto worker = [&]() {
    std::vector<int> d;
    try {
      d.assign(n, -1);
    } catch (const std::bad_alloc &) {
      oom = true;
      return;
    }
    std::vector<int> touched;
    for (;;) {
      long p = next.fetch_add(1);
      if (p >= npart) break;
      touched.assign(members[p].begin(), members[p].end());
      for (int i : members[p]) d[i] = 0;
      size_t head = 0;
      while (head < touched.size()) {
        int u = touched[head++];
        if (d[u] >= bound) continue;
        for (int j = ptr[u]; j < ptr[u + 1]; j++) {
          int c = col[j];
          if (d[c] < 0) {
            d[c] = d[u] + 1;
            touched.push_back(c);
          }
        }
      }
      auto &out = per[p];
      for (int j : touched) {
        long dj = d[j];
        bool keep = lastl ? (lastl[j] < nlevel - dj) : (dj < nlevel);
        if (keep) out.push_back({j, (int)dj});
      }
      if (bound >= (long)n + 1) {  // unreachable rows carry n+1 (lib.c:603)
        for (int j = 0; j < n; j++)
          if (d[j] < 0) {
            long dj = (long)n + 1;
            bool keep = lastl ? (lastl[j] < nlevel - dj) : (dj < nlevel);
            if (keep) out.push_back({j, (int)dj});
          }
      }
      for (int j : touched) d[j] = -1;
      std::sort(out.begin(), out.end());
    }
  };
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; t++) pool.emplace_back(worker);
  for (auto &th : pool) th.join();
  if (oom) return fail(MPKX_ENOMEM, "comp_skirt: allocation");
  long total = 0;
  for (auto &v : per) total += (long)v.size();
  long *sp = (long *)malloc(sizeof(long) * (npart + 1));
  int *sr = (int *)malloc(sizeof(int) * (total + 1));
  int *sd = (int *)malloc(sizeof(int) * (total + 1));
  if (!sp || !sr || !sd) {
    free(sp), free(sr), free(sd);
    return fail(MPKX_ENOMEM, "comp_skirt: allocation");
  }
  sp[0] = 0;
  for (int p = 0; p < npart; p++) {
    long o = sp[p];
    for (size_t t = 0; t < per[p].size(); t++) {
      sr[o + t] = per[p][t].first;
      sd[o + t] = per[p][t].second;
    }
    sp[p + 1] = o + (long)per[p].size();
  }
  *sptr_out = sp;
  *srow_out = sr;
  *sdist_out = sd;
  return MPKX_OK;
}

// Jonker & Volgenant, "A shortest augmenting path algorithm for dense and
// sparse linear assignment problems", Computing 38 (1987), in the variant
// of mpk2/lapjv.c:141-390: the same phases and the same scan orders, so
// ties resolve to the same permutation.
int mpkx_lapjv(const int *sums, int np, int *perm) {
  clear_error();
  if (np <= 0) return fail(MPKX_EINVAL, "lapjv: npart %d", np);
  const int NONE = -1;
  std::vector<int> cost((size_t)np * np), rowsol(np, NONE), colsol(np, NONE), v(np, 0),
      collist(np), freerows(np), d(np), pred(np);
  int mx = *std::max_element(sums, sums + (size_t)np * np);
  for (size_t t = 0; t < (size_t)np * np; t++) cost[t] = mx - sums[t];
  auto C = [&](int i, int j) { return cost[(size_t)i * np + j]; };
  // Column reduction, last column first.  A row that wins a second
  // column is flagged by storing -(col)-2.
  for (int j = np - 1; j >= 0; j--) {
    collist[j] = j;
    int best = C(0, j), bi = 0;
    for (int i = 1; i < np; i++)
      if (C(i, j) < best) best = C(i, j), bi = i;
    v[j] = best;
    if (rowsol[bi] == NONE) {
      rowsol[bi] = j;
      colsol[j] = bi;
    } else {
      if (rowsol[bi] >= NONE) rowsol[bi] = -rowsol[bi] - 2;
      colsol[j] = NONE;
    }
  }
  // Reduction transfer.
  int nfree = 0;
  for (int i = 0; i < np; i++) {
    if (rowsol[i] == NONE) {
      freerows[nfree++] = i;
    } else if (rowsol[i] < NONE) {
      rowsol[i] = -rowsol[i] - 2;
    } else {
      int j1 = rowsol[i], m = INT_MAX;
      for (int j = 0; j < np; j++)
        if (j != j1 && C(i, j) - v[j] < m) m = C(i, j) - v[j];
      v[j1] -= m;
    }
  }
  // Augmenting row reduction, two passes.
  for (int pass = 0; pass < 2; pass++) {
    int todo = nfree, k = 0;
    nfree = 0;
    while (k < todo) {
      int i = freerows[k++];
      int u1 = C(i, 0) - v[0], u2 = INT_MAX, j1 = 0, j2 = 0;
      for (int j = 1; j < np; j++) {
        int h = C(i, j) - v[j];
        if (h < u2) {
          if (h >= u1) {
            u2 = h;
            j2 = j;
          } else {
            u2 = u1;
            u1 = h;
            j2 = j1;
            j1 = j;
          }
        }
      }
      int i0 = colsol[j1];
      if (u1 < u2)
        v[j1] = v[j1] - u2 + u1;
      else if (i0 > NONE) {
        j1 = j2;
        i0 = colsol[j1];
      }
      if (i0 > NONE) {
        if (u1 < u2)
          freerows[--k] = i0;
        else
          freerows[nfree++] = i0;
      }
      rowsol[i] = j1;
      colsol[j1] = i;
    }
  }
  // Shortest augmenting paths for the remaining free rows.
  int todo = nfree;
  for (int f = 0; f < todo; f++) {
    int i1 = freerows[f];
    int low = 0, up = 0, last = 0, mn = 0, endj = -1;
    for (int j = 0; j < np; j++) {
      d[j] = C(i1, j) - v[j];
      pred[j] = i1;
    }
    while (endj < 0) {
      if (up == low) {
        last = low;
        mn = d[collist[up++]];
        for (int k = up; k < np; k++) {
          int j = collist[k], h = d[j];
          if (h <= mn) {
            if (h < mn) {
              up = low;
              mn = h;
            }
            collist[k] = collist[up];
            collist[up++] = j;
          }
        }
        for (int h = low; h < up && endj < 0; h++)
          if (colsol[collist[h]] == NONE) endj = collist[h];
        if (endj >= 0) break;
      }
      int j1 = collist[low++];
      int i = colsol[j1];
      int u1 = C(i, j1) - v[j1] - mn;
      for (int k = up; k < np; k++) {
        int j = collist[k];
        int h = C(i, j) - v[j] - u1;
        if (h < d[j]) {
          d[j] = h;
          pred[j] = i;
          if (h == mn) {
            if (colsol[j] == NONE) {
              endj = j;
              break;
            }
            collist[k] = collist[up];
            collist[up++] = j;
          }
        }
      }
    }
    for (int k = 0; k < last; k++) {
      int j = collist[k];
      v[j] = v[j] + d[j] - mn;
    }
    int i, j = endj;
    do {
      i = pred[j];
      colsol[j] = i;
      int t = j;
      j = rowsol[i];
      rowsol[i] = t;
    } while (i != i1);
  }
  for (int j = 0; j < np; j++) perm[j] = colsol[j];
  return MPKX_OK;
}

// ---- partitioners ----------------------------------------------------------

int mpkx_partition_grid(int ndim, const int *dims, const int *blocks, const int *shift, int *part) {
  clear_error();
  if (ndim < 1 || ndim > 4) return fail(MPKX_EINVAL, "partition_grid: ndim %d", ndim);
  long n = 1;
  for (int d = 0; d < ndim; d++) {
    if (dims[d] <= 0 || blocks[d] <= 0 || blocks[d] > dims[d])
      return fail(MPKX_EINVAL, "partition_grid: dim %d size %d blocks %d", d, dims[d], blocks[d]);
    n *= dims[d];
  }
  parallel_for(n, [&](long i) {
    long rem = i, stride = n, lab = 0;
    for (int d = 0; d < ndim; d++) {
      stride /= dims[d];
      long c = rem / stride;
      rem %= stride;
      long s = shift ? shift[d] : 0;
      c = ((c + s) % dims[d] + dims[d]) % dims[d];
      lab = lab * blocks[d] + c * blocks[d] / dims[d];
    }
    part[i] = (int)lab;
  });
  return MPKX_OK;
}

namespace {

// Breadth-first order of the rows `rows` (a subgraph given by membership
// mark[] == tag) starting from a pseudo-peripheral row.
void bfs_order(const int *ptr, const int *col, const std::vector<int> &rows,
               std::vector<int> &mark, int tag, std::vector<int> &order,
               std::vector<int> &depth) {
  auto sweep = [&](int start) {
    order.clear();
    order.push_back(start);
    depth[start] = 0;
    mark[start] = tag + 1;
    for (size_t h = 0; h < order.size(); h++) {
      int u = order[h];
      for (int j = ptr[u]; j < ptr[u + 1]; j++) {
        int c = col[j];
        if (mark[c] == tag) {
          mark[c] = tag + 1;
          depth[c] = depth[u] + 1;
          order.push_back(c);
        }
      }
    }
    for (int u : order) mark[u] = tag;
    return order.back();
  };
  int start = rows[0];
  for (int it = 0; it < 2; it++) start = sweep(start);  // move to a far row
  sweep(start);
  if (order.size() < rows.size()) {  // disconnected: append the other components
    for (int u : order) mark[u] = tag + 2;
    for (int r : rows)
      if (mark[r] == tag) {
        std::vector<int> o2;
        o2.push_back(r);
        mark[r] = tag + 2;
        for (size_t h = 0; h < o2.size(); h++) {
          int u = o2[h];
          for (int j = ptr[u]; j < ptr[u + 1]; j++)
            if (mark[col[j]] == tag) {
              mark[col[j]] = tag + 2;
              o2.push_back(col[j]);
            }
        }
        order.insert(order.end(), o2.begin(), o2.end());
      }
    for (int u : order) mark[u] = tag;
  }
}

}  // namespace

int mpkx_partition_graph(int n, const int *ptr, const int *col, int npart, int *part) {
  clear_error();
  if (n <= 0 || npart <= 0 || npart > n) return fail(MPKX_EINVAL, "partition_graph: n %d npart %d", n, npart);
  // Recursive bisection along breadth-first orders: every split cuts the
  // BFS order of the current part at the size ratio of the two halves.
  std::vector<int> mark(n, 0), depth(n, 0), order;
  struct Job {
    std::vector<int> rows;
    int first, count;
  };
  std::vector<Job> stack;
  Job all;
  all.rows.resize(n);
  for (int i = 0; i < n; i++) all.rows[i] = i;
  all.first = 0;
  all.count = npart;
  stack.push_back(std::move(all));
  int tag = 2;
  while (!stack.empty()) {
    Job job = std::move(stack.back());
    stack.pop_back();
    if (job.count == 1) {
      for (int r : job.rows) part[r] = job.first;
      continue;
    }
    tag += 4;
    for (int r : job.rows) mark[r] = tag;
    bfs_order(ptr, col, job.rows, mark, tag, order, depth);
    int c1 = job.count / 2, c2 = job.count - c1;
    size_t cut = (size_t)((double)order.size() * c1 / job.count + 0.5);
    Job a, b;
    a.rows.assign(order.begin(), order.begin() + cut);
    b.rows.assign(order.begin() + cut, order.end());
    for (int r : job.rows) mark[r] = 0;
    a.first = job.first;
    a.count = c1;
    b.first = job.first + c1;
    b.count = c2;
    stack.push_back(std::move(a));
    stack.push_back(std::move(b));
  }
  return MPKX_OK;
}

}  // extern "C"
