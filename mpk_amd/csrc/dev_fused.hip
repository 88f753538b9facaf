// dev_fused.hip -- fused matrix powers kernel ("skirt rounds").
//
// Schedule.  The reference's nphase = 0 schedule (mpk2/driver.c:290-311,
// mpk2/buffers.c:155-179) gives every partition p its skirt: all rows j
// within graph distance d = dist(j, p) < k, and p computes (level l, row j)
// for l <= k - d without any communication.  Here the same construction is
// applied every s levels ("a round"): for levels L0+1 .. L0+w (w <= s) a
// workgroup owning tile p
//   1. gathers level-L0 values of its skirt region (d <= w) from the Krylov
//      block in HBM into LDS -- this replaces the MPI_Alltoallv halo
//      exchange between phases;
//   2. for j = 1..w computes level L0+j of every region row with
//      d <= w - j from the LDS copy of level L0+j-1 (ping-pong buffers),
//      a barrier between levels;
//   3. writes the tile's own rows (d == 0) of every level to HBM.
// Redundant work on the shrinking skirt is traded for halo traffic; s is the
// knob (the reference's "levels per phase").
//
// Data layout per tile, built once on the host:
//   slots   region rows sorted by (distance, row): rows active at relative
//           level j are the prefix cnt[w-j]; the tile's own rows come first.
//   matrix  compute rows (d <= s-1) as a column-major ELL block of W entries
//           per row, [m][w][thread]: row r = m*T + thread lives in thread
//           r % T's registers for the whole launch.  An entry is a 16-bit
//           slot plus a 16-bit value code (dictionary of <= 4 distinct
//           values, held in registers) or a slot with an FP64 value stream.
//   Padding entries are 0xffffffff and skipped.
// Products and sums are rounded separately, in CSR order, so the FP64
// result is bitwise equal to k back-to-back CSR SpMVs (verify_val).
#include <algorithm>
#include <cstring>
#include <memory>
#include <unordered_map>

#include "dev_common.h"

using namespace mpkx;

namespace {

constexpr int HDR = 16;        // ints per tile header
constexpr int MAX_S = 10;      // levels per launch (header holds cnt[0..s])
constexpr uint32_t PAD = 0xffffffffu;
constexpr long LDS_BYTES = 160 * 1024;

template <int T, int MR, int W, bool DICT>
__global__ __launch_bounds__(T) void k_fused(const int *__restrict__ hdr, const int *__restrict__ grow,
                                             const uint32_t *__restrict__ ent,
                                             const double *__restrict__ vals, double d0, double d1,
                                             double d2, double d3, double *vv, int n, int L0, int w,
                                             int ntiles, int maxslots) {
  extern __shared__ double lds[];
  double *buf0 = lds, *buf1 = lds + maxslots;
  const int tile = xcd_remap(blockIdx.x, ntiles);
  const int *h = hdr + (size_t)tile * HDR;
  const int *g = grow + h[0];
  const size_t mat = (size_t)(unsigned)h[1];
  const int *cnt = h + 4;
  const int tid = threadIdx.x;

  // 1. level-L0 values of the region (distance <= w) into LDS
  const int nin = cnt[w];
  const double *src = vv + (size_t)L0 * n;
  for (int t = tid; t < nin; t += T) buf0[t] = src[g[t]];

  // matrix rows of this thread into registers; own-row targets
  const int ncomp = cnt[w - 1], nown = cnt[0];
  uint32_t e[MR][W];
  int own[MR];
#pragma unroll
  for (int m = 0; m < MR; m++) {
    const int r = m * T + tid;
#pragma unroll
    for (int k = 0; k < W; k++) e[m][k] = r < ncomp ? ent[mat + (size_t)(m * W + k) * T + tid] : PAD;
    own[m] = r < nown ? g[r] : -1;
  }
  __syncthreads();

  double *in = buf0, *out = buf1;
  for (int j = 1; j <= w; j++) {
    const int nact = cnt[w - j];
    double *dst = vv + (size_t)(L0 + j) * n;
#pragma unroll
    for (int m = 0; m < MR; m++) {
      const int r = m * T + tid;
      if (r < nact) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < W; k++) {
          const uint32_t x = e[m][k];
          if (x != PAD) {
            double a;
            if (DICT) {
              const uint32_t c = x >> 16;
              a = c == 0 ? d0 : c == 1 ? d1 : c == 2 ? d2 : d3;
            } else {
              a = vals[mat + (size_t)(m * W + k) * T + tid];
            }
            s = __dadd_rn(s, __dmul_rn(a, in[x & 0xffffu]));
          }
        }
        out[r] = s;
        if (own[m] >= 0) dst[own[m]] = s;
      }
    }
    __syncthreads();
    double *t = in;
    in = out;
    out = t;
  }
}

__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

using FusedFn = void (*)(const int *, const int *, const uint32_t *, const double *, double, double,
                         double, double, double *, int, int, int, int, int);

template <int T, int MR, int W>
FusedFn pick_codec(bool dict) {
  return dict ? k_fused<T, MR, W, true> : k_fused<T, MR, W, false>;
}

template <int T, int MR>
FusedFn pick_w(int w, bool dict) {
  switch (w) {
    case 5: return pick_codec<T, MR, 5>(dict);
    case 7: return pick_codec<T, MR, 7>(dict);
    case 9: return pick_codec<T, MR, 9>(dict);
    case 16: return pick_codec<T, MR, 16>(dict);
  }
  return nullptr;
}

FusedFn pick(int T, int MR, int W, bool dict) {
  if (T == 1024) return MR == 4 ? pick_w<1024, 4>(W, dict) : MR == 8 ? pick_w<1024, 8>(W, dict) : nullptr;
  if (T == 512) return MR == 4 ? pick_w<512, 4>(W, dict) : MR == 8 ? pick_w<512, 8>(W, dict) : nullptr;
  return nullptr;
}

}  // namespace

struct mpkx_fused {
  int n = 0, nlevel = 0, s = 0, ntiles = 0, W = 0, MR = 0, T = 0;
  bool dict = false;
  double dv[4] = {0, 0, 0, 0};
  int maxslots = 0;
  size_t lds_bytes = 0;
  long slots = 0, row_levels = 0, matrix_bytes = 0;
  FusedFn fn = nullptr;
  DevArray<int> hdr, grow;
  DevArray<uint32_t> ent;
  DevArray<double> vals;
  GraphCache graph;
};

static void fused_launches(mpkx_fused *f, const double *x, double *vv, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((f->n + 255) / 256)), dim3(256), 0, st, vv, x, (long)f->n);
  for (int L0 = 0; L0 < f->nlevel; L0 += f->s) {
    int w = std::min(f->s, f->nlevel - L0);
    hipLaunchKernelGGL(f->fn, dim3(f->ntiles), dim3(f->T), f->lds_bytes, st, f->hdr.p, f->grow.p,
                       f->ent.p, f->vals.p, f->dv[0], f->dv[1], f->dv[2], f->dv[3], vv, f->n, L0, w,
                       f->ntiles, f->maxslots);
  }
}

extern "C" {

int mpkx_fused_create(mpkx_fused **out, int n, const int *ptr, const int *col, const double *val,
                      const int *part, int npart, int nlevel, int s, int flags) {
  clear_error();
  if (!out || n <= 0 || !ptr || !col || !val || !part || npart <= 0 || nlevel <= 0)
    return fail(MPKX_EINVAL, "fused_create: bad arguments");
  if (s < 1 || s > MAX_S) return fail(MPKX_EINVAL, "fused_create: s=%d outside 1..%d", s, MAX_S);
  s = std::min(s, nlevel);
  int maxdeg = 0;
  for (int i = 0; i < n; i++) maxdeg = std::max(maxdeg, ptr[i + 1] - ptr[i]);
  int W = maxdeg <= 5 ? 5 : maxdeg <= 7 ? 7 : maxdeg <= 9 ? 9 : maxdeg <= 16 ? 16 : 0;
  if (!W) return fail(MPKX_EUNSUPPORTED, "fused_create: rows up to %d entries, at most 16 supported", maxdeg);

  // value codec: up to four distinct values -> register dictionary
  std::vector<double> dict;
  bool use_dict = !(flags & MPKX_FUSED_EXPLICIT);
  if (use_dict) {
    for (long j = 0; j < ptr[n] && use_dict; j++) {
      bool found = false;
      for (double d : dict)
        if (memcmp(&d, &val[j], sizeof(double)) == 0) found = true;
      if (!found) {
        if (dict.size() == 4)
          use_dict = false;
        else
          dict.push_back(val[j]);
      }
    }
  }

  std::vector<std::vector<int>> members(npart);
  for (int i = 0; i < n; i++) {
    if (part[i] < 0 || part[i] >= npart) return fail(MPKX_EINVAL, "fused_create: label %d out of range", part[i]);
    members[part[i]].push_back(i);
  }
  for (int p = 0; p < npart; p++)
    if (members[p].empty()) return fail(MPKX_EINVAL, "fused_create: partition %d is empty", p);

  // Per tile: region = rows within distance s, sorted by (distance, row).
  struct Tile {
    std::vector<int> slots;   // global rows
    std::vector<int> cnt;     // cnt[d] = #slots with distance <= d
    std::vector<int> rows;    // compute-row entries: slot per entry (W per row)
    std::vector<double> vals; // matching values
  };
  std::vector<Tile> tiles(npart);
  std::atomic<int> err{0};
  std::atomic<long> next{0};
  int nthreads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  auto worker = [&]() {
    std::vector<int> dist(n, -1), slot(n, -1);
    for (;;) {
      long p = next.fetch_add(1);
      if (p >= npart || err) break;
      Tile &t = tiles[p];
      std::vector<int> q(members[p]);
      for (int i : q) dist[i] = 0;
      for (size_t h = 0; h < q.size(); h++) {
        int u = q[h];
        if (dist[u] >= s) continue;
        for (int j = ptr[u]; j < ptr[u + 1]; j++)
          if (dist[col[j]] < 0) {
            dist[col[j]] = dist[u] + 1;
            q.push_back(col[j]);
          }
      }
      std::sort(q.begin(), q.end(), [&](int a, int b) { return dist[a] != dist[b] ? dist[a] < dist[b] : a < b; });
      t.slots = q;
      t.cnt.assign(s + 1, 0);
      for (int u : q) t.cnt[dist[u]]++;
      for (int d = 1; d <= s; d++) t.cnt[d] += t.cnt[d - 1];
      for (size_t k = 0; k < q.size(); k++) slot[q[k]] = (int)k;
      int ncomp = t.cnt[s - 1];
      t.rows.assign((size_t)ncomp * W, -1);
      t.vals.assign((size_t)ncomp * W, 0.0);
      for (int r = 0; r < ncomp; r++) {
        int u = q[r];
        for (int j = ptr[u]; j < ptr[u + 1]; j++) {
          t.rows[(size_t)r * W + (j - ptr[u])] = slot[col[j]];
          t.vals[(size_t)r * W + (j - ptr[u])] = val[j];
        }
      }
      for (int u : q) dist[u] = -1, slot[u] = -1;
    }
  };
  {
    std::vector<std::thread> pool;
    for (int i = 0; i < nthreads; i++) pool.emplace_back(worker);
    for (auto &th : pool) th.join();
  }
  int maxslots = 0, maxcomp = 0;
  for (auto &t : tiles) {
    maxslots = std::max(maxslots, (int)t.slots.size());
    maxcomp = std::max(maxcomp, t.cnt[s - 1]);
  }
  if (maxslots > 65535) return fail(MPKX_EUNSUPPORTED, "fused_create: %d region rows exceed 16-bit slots", maxslots);
  size_t lds = (size_t)2 * maxslots * sizeof(double);
  if ((long)lds > LDS_BYTES)
    return fail(MPKX_EUNSUPPORTED, "fused_create: tile region of %d rows needs %zu B LDS (> %ld); use more partitions or a smaller s",
                maxslots, lds, LDS_BYTES);
  int T = 0, MR = 0;
  if (maxcomp <= 512 * 4) T = 512, MR = 4;
  else if (maxcomp <= 1024 * 4) T = 1024, MR = 4;
  else if (maxcomp <= 1024 * 8) T = 1024, MR = 8;
  else return fail(MPKX_EUNSUPPORTED, "fused_create: %d compute rows per tile exceed %d", maxcomp, 8 * 1024);

  auto *f = new (std::nothrow) mpkx_fused;
  if (!f) return fail(MPKX_ENOMEM, "fused_create");
  std::unique_ptr<mpkx_fused> guard(f);
  f->n = n;
  f->nlevel = nlevel;
  f->s = s;
  f->ntiles = npart;
  f->W = W;
  f->MR = MR;
  f->T = T;
  f->dict = use_dict;
  for (size_t i = 0; i < dict.size(); i++) f->dv[i] = dict[i];
  f->maxslots = maxslots;
  f->lds_bytes = lds;
  f->fn = pick(T, MR, W, use_dict);
  if (!f->fn) return fail(MPKX_EUNSUPPORTED, "fused_create: no kernel for T=%d MR=%d W=%d", T, MR, W);

  std::vector<int> hdr((size_t)npart * HDR, 0), grow;
  std::vector<uint32_t> ent;
  std::vector<double> vals;
  long slot_off = 0, mat_off = 0;
  for (int p = 0; p < npart; p++) {
    const Tile &t = tiles[p];
    int ncomp = t.cnt[s - 1];
    int mrt = (ncomp + T - 1) / T;
    int *h = &hdr[(size_t)p * HDR];
    h[0] = (int)slot_off;
    h[1] = (int)mat_off;
    h[2] = (int)t.slots.size();
    h[3] = ncomp;
    for (int d = 0; d <= s; d++) h[4 + d] = t.cnt[d];
    grow.insert(grow.end(), t.slots.begin(), t.slots.end());
    size_t seg = (size_t)mrt * W * T;
    ent.resize(mat_off + seg, PAD);
    if (!use_dict) vals.resize(mat_off + seg, 0.0);
    for (int r = 0; r < ncomp; r++) {
      int m = r / T, th = r % T;
      for (int k = 0; k < W; k++) {
        int sl = t.rows[(size_t)r * W + k];
        if (sl < 0) continue;
        size_t at = mat_off + (size_t)(m * W + k) * T + th;
        uint32_t code = 0;
        if (use_dict) {
          double v = t.vals[(size_t)r * W + k];
          for (size_t c = 0; c < dict.size(); c++)
            if (memcmp(&dict[c], &v, sizeof(double)) == 0) code = (uint32_t)c;
        } else {
          vals[at] = t.vals[(size_t)r * W + k];
        }
        ent[at] = (uint32_t)sl | (code << 16);
      }
    }
    slot_off += (long)t.slots.size();
    mat_off += (long)seg;
    if (mat_off > INT32_MAX || slot_off > INT32_MAX)
      return fail(MPKX_EUNSUPPORTED, "fused_create: plan exceeds 32-bit offsets");
  }
  f->slots = slot_off;
  long rl = 0;
  for (int L0 = 0; L0 < nlevel; L0 += s) {
    int w = std::min(s, nlevel - L0);
    for (auto &t : tiles)
      for (int j = 1; j <= w; j++) rl += t.cnt[w - j];
  }
  f->row_levels = rl;
  f->matrix_bytes = (long)ent.size() * 4 + (long)vals.size() * 8;
  int rc;
  if ((rc = f->hdr.upload(hdr)) || (rc = f->grow.upload(grow)) || (rc = f->ent.upload(ent)) ||
      (rc = f->vals.upload(vals.empty() ? std::vector<double>(1, 0.0) : vals)))
    return rc;
  MPKX_HIP(hipFuncSetAttribute((const void *)f->fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  *out = guard.release();
  return MPKX_OK;
}

void mpkx_fused_destroy(mpkx_fused *f) { delete f; }

int mpkx_fused_exec(mpkx_fused *f, const double *x, double *vv, void *stream, int use_graph) {
  clear_error();
  if (!f || !x || !vv) return fail(MPKX_EINVAL, "fused_exec: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) {
    fused_launches(f, x, vv, st);
    MPKX_HIP(hipGetLastError());
    return MPKX_OK;
  }
  if (!f->graph.matches(x, vv, st, 0)) {
    f->graph.reset();
    hipStream_t cap = st;
    bool own = false;
    if (cap == nullptr) {
      MPKX_HIP(hipStreamCreate(&cap));
      own = true;
    }
    MPKX_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    fused_launches(f, x, vv, cap);
    MPKX_HIP(hipStreamEndCapture(cap, &f->graph.graph));
    MPKX_HIP(hipGraphInstantiate(&f->graph.exec, f->graph.graph, nullptr, nullptr, 0));
    if (own) MPKX_HIP(hipStreamDestroy(cap));
    f->graph.k0 = x;
    f->graph.k1 = vv;
    f->graph.stream = st;
    f->graph.k = 0;
  }
  MPKX_HIP(hipGraphLaunch(f->graph.exec, st));
  return MPKX_OK;
}

int mpkx_fused_round(mpkx_fused *f, double *vv, int round, void *stream) {
  clear_error();
  if (!f || !vv || round < 0 || (long)round * f->s >= f->nlevel) return fail(MPKX_EINVAL, "fused_round: bad arguments");
  int L0 = round * f->s, w = std::min(f->s, f->nlevel - L0);
  hipLaunchKernelGGL(f->fn, dim3(f->ntiles), dim3(f->T), f->lds_bytes, (hipStream_t)stream, f->hdr.p, f->grow.p,
                     f->ent.p, f->vals.p, f->dv[0], f->dv[1], f->dv[2], f->dv[3], vv, f->n, L0, w,
                     f->ntiles, f->maxslots);
  MPKX_HIP(hipGetLastError());
  return MPKX_OK;
}

int mpkx_fused_stats(const mpkx_fused *f, long *max_slots, long *slots, long *row_levels,
                     long *matrix_bytes, int *codec, int *launches) {
  clear_error();
  if (!f) return fail(MPKX_EINVAL, "fused_stats: null");
  if (max_slots) *max_slots = f->maxslots;
  if (slots) *slots = f->slots;
  if (row_levels) *row_levels = f->row_levels;
  if (matrix_bytes) *matrix_bytes = f->matrix_bytes;
  if (codec) *codec = f->dict ? 1 : 0;
  if (launches) *launches = (f->nlevel + f->s - 1) / f->s;
  return MPKX_OK;
}

}  // extern "C"
