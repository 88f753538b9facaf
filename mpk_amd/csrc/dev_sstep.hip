// dev_sstep.hip -- s-step (communication-avoiding) conjugate gradients on
// top of the fused matrix powers kernel: the caller the Krylov block exists
// for (no exec counterpart in the reference; its notebook
// jupyter/s-step-cg.ipynb sketches the method).
//
// CA-CG in the monomial basis (Hoemmen 2010, Carson 2015): every outer
// iteration
//   1. runs two MPKs: P = [p, Ap, .., A^s p] and R = [r, Ar, .., A^(s-1) r]
//      (the plan computes s powers of r; the last one is not used),
//   2. forms the Gram matrix G = Y^T Y of Y = [P, R] (m = 2s+1 columns) in
//      one pass over HBM,
//   3. takes s CG steps on length-m coordinate vectors on the host (the
//      shift B maps column j of a block to column j+1),
//   4. recovers x += Y x', r = Y r', p = Y p' in one pass.
// The operator is A/sigma (sigma = the power of two at or above the largest
// row abs-sum, an exact scaling) applied to b/sigma, which has the same
// solution and keeps the monomial basis from growing.
#include <cmath>
#include <memory>
#include <vector>

#include "dev_common.h"
#include "internal.h"

using namespace mpkx;

namespace {

constexpr int SS_MAX_S = 8;
constexpr int SS_MAXM = 2 * SS_MAX_S + 1;
constexpr int GT = 256, GR = 64;  // threads, rows per Gram tile
constexpr int GRID = 1024;

struct Coef {
  double c[3][SS_MAXM];  // x, r, p
};

__device__ __forceinline__ double yval(const double *vp, int np, const double *vr, long n, int v, long row) {
  return v < np ? vp[(size_t)v * n + row] : vr[(size_t)(v - np) * n + row];
}

// Partial Gram sums: block b sums rows of its tiles for every pair (i <= j).
__global__ __launch_bounds__(GT) void k_gram_partial(const double *__restrict__ vp, int np,
                                                     const double *__restrict__ vr, int nr, long n,
                                                     double *__restrict__ partial) {
  __shared__ double tile[SS_MAXM][GR + 1];
  const int m = np + nr, npairs = m * (m + 1) / 2, t = threadIdx.x;
  int pi = 0, pj = 0;
  if (t < npairs) {  // pair t -> (pi, pj), row-major upper triangle
    int q = t;
    while (q >= m - pi) q -= m - pi, pi++;
    pj = pi + q;
  }
  double acc = 0.0;
  for (long r0 = (long)blockIdx.x * GR; r0 < n; r0 += (long)gridDim.x * GR) {
    for (int idx = t; idx < m * GR; idx += GT) {
      const int v = idx / GR, rr = idx % GR;
      const long row = r0 + rr;
      tile[v][rr] = row < n ? yval(vp, np, vr, n, v, row) : 0.0;
    }
    __syncthreads();
    if (t < npairs)
      for (int rr = 0; rr < GR; rr++) acc = fma(tile[pi][rr], tile[pj][rr], acc);
    __syncthreads();
  }
  if (t < npairs) partial[(size_t)blockIdx.x * npairs + t] = acc;
}

// One workgroup per pair: lane t sums the partials t, t + 256, ..., then a
// pairwise tree in LDS (a fixed order, so the Gram matrix is deterministic).
__global__ __launch_bounds__(GT) void k_gram_finish(const double *__restrict__ partial, int blocks, int npairs,
                                                    double *__restrict__ g) {
  __shared__ double red[GT];
  const int q = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int b = t; b < blocks; b += GT) s += partial[(size_t)b * npairs + q];
  red[t] = s;
  __syncthreads();
  for (int w = GT / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) g[q] = red[0];
}

// x += Y c[0], r = Y c[1], p = Y c[2]
__global__ __launch_bounds__(256) void k_combine3(const double *__restrict__ vp, int np, const double *__restrict__ vr,
                                                  int nr, long n, Coef c, double *__restrict__ x,
                                                  double *__restrict__ r, double *__restrict__ p) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int m = np + nr;
  double sx = 0.0, sr = 0.0, sp = 0.0;
  for (int v = 0; v < m; v++) {
    const double y = yval(vp, np, vr, n, v, i);
    sx = fma(c.c[0][v], y, sx);
    sr = fma(c.c[1][v], y, sr);
    sp = fma(c.c[2][v], y, sp);
  }
  x[i] += sx;
  r[i] = sr;
  p[i] = sp;
}

__global__ void k_start(const double *__restrict__ b, double inv_sigma, long n, double *__restrict__ x,
                        double *__restrict__ r, double *__restrict__ p) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double v = b[i] * inv_sigma;
  x[i] = 0.0;
  r[i] = v;
  p[i] = v;
}

}  // namespace

struct mpkx_sstep {
  int n = 0, s = 0;
  double sigma = 1.0;
  mpkx_fused *mpk = nullptr;
  DevArray<double> vp, vr, r, p, partial, gram;
  ~mpkx_sstep() {
    if (mpk) mpkx_fused_destroy(mpk);
  }
};

extern "C" {

int mpkx_sstep_create(mpkx_sstep **out, int n, const int *ptr, const int *col, const double *val, const int *part,
                      int npart, int s, int levels_per_launch) {
  clear_error();
  if (!out || n <= 0 || !ptr || !col || !val || !part || npart <= 0)
    return fail(MPKX_EINVAL, "sstep_create: bad arguments");
  if (s < 1 || s > SS_MAX_S) return fail(MPKX_EUNSUPPORTED, "sstep_create: s=%d outside 1..%d", s, SS_MAX_S);
  auto *S = new (std::nothrow) mpkx_sstep;
  if (!S) return fail(MPKX_ENOMEM, "sstep_create");
  std::unique_ptr<mpkx_sstep> guard(S);
  S->n = n;
  S->s = s;
  double mx = 0.0;
  for (int i = 0; i < n; i++) {
    double a = 0.0;
    for (int j = ptr[i]; j < ptr[i + 1]; j++) a += std::fabs(val[j]);
    mx = std::max(mx, a);
  }
  int e2 = 0;
  if (mx > 0) {
    double f = std::frexp(mx, &e2);
    if (f == 0.5) e2 -= 1;
  }
  S->sigma = std::ldexp(1.0, e2);
  std::vector<double> scaled(val, val + ptr[n]);
  for (double &v : scaled) v /= S->sigma;
  int rc = mpkx_fused_create(&S->mpk, n, ptr, col, scaled.data(), part, npart, s,
                             std::max(1, std::min(levels_per_launch, s)), MPKX_FUSED_AUTO);
  if (rc) return rc;
  const int m = 2 * s + 1, npairs = m * (m + 1) / 2;
  if ((rc = S->vp.alloc((size_t)(s + 1) * n)) || (rc = S->vr.alloc((size_t)(s + 1) * n)) ||
      (rc = S->r.alloc(n)) || (rc = S->p.alloc(n)) || (rc = S->partial.alloc((size_t)GRID * npairs)) ||
      (rc = S->gram.alloc(npairs)))
    return rc;
  *out = guard.release();
  return MPKX_OK;
}

void mpkx_sstep_destroy(mpkx_sstep *S) { delete S; }

int mpkx_sstep_solve(mpkx_sstep *S, const double *b, double *x, double rtol, int max_outer, int *iters,
                     double *resid, void *stream) {
  clear_error();
  if (!S || !b || !x || !(rtol > 0) || max_outer < 1) return fail(MPKX_EINVAL, "sstep_solve: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int s = S->s, m = 2 * s + 1, npairs = m * (m + 1) / 2, np = s + 1, nr = s;
  const long n = S->n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_start, dim3(nb), dim3(256), 0, st, b, 1.0 / S->sigma, n, x, S->r.p, S->p.p);
  MPKX_HIP(hipGetLastError());
  std::vector<double> gp(npairs), G((size_t)m * m);
  double rr0 = -1.0, rel = 1.0;
  int outer = 0;
  for (; outer < max_outer; outer++) {
    int rc;
    // eager launches: the two calls alternate their (x, vv) pointers, which
    // would re-record a cached graph every time
    if ((rc = mpkx_fused_exec(S->mpk, S->p.p, S->vp.p, st, 0)) || (rc = mpkx_fused_exec(S->mpk, S->r.p, S->vr.p, st, 0)))
      return rc;
    hipLaunchKernelGGL(k_gram_partial, dim3(GRID), dim3(GT), 0, st, S->vp.p, np, S->vr.p, nr, n, S->partial.p);
    hipLaunchKernelGGL(k_gram_finish, dim3(npairs), dim3(GT), 0, st, S->partial.p, GRID, npairs, S->gram.p);
    MPKX_HIP(hipGetLastError());
    MPKX_HIP(hipMemcpyAsync(gp.data(), S->gram.p, sizeof(double) * npairs, hipMemcpyDeviceToHost, st));
    MPKX_HIP(hipStreamSynchronize(st));
    for (int i = 0, q = 0; i < m; i++)
      for (int j = i; j < m; j++, q++) G[(size_t)i * m + j] = G[(size_t)j * m + i] = gp[q];
    // coordinates: p' = e_0, r' = e_{s+1}, x' = 0; B shifts within each block
    std::vector<double> xc(m, 0.0), rc_(m, 0.0), pc(m, 0.0), bp(m), rn(m);
    pc[0] = 1.0;
    rc_[s + 1] = 1.0;
    auto shift = [&](const std::vector<double> &v, std::vector<double> &out) {
      std::fill(out.begin(), out.end(), 0.0);
      for (int j = 0; j < s; j++) out[j + 1] = v[j];           // P block: columns 0..s
      for (int j = s + 1; j < 2 * s; j++) out[j + 1] = v[j];   // R block: columns s+1..2s
    };
    auto gdot = [&](const std::vector<double> &a, const std::vector<double> &c) {
      double t = 0.0;
      for (int i = 0; i < m; i++) {
        double u = 0.0;
        for (int j = 0; j < m; j++) u += G[(size_t)i * m + j] * c[j];
        t += a[i] * u;
      }
      return t;
    };
    double rr = gdot(rc_, rc_);
    if (rr0 < 0) rr0 = rr;
    if (rr <= 0.0 || std::sqrt(rr / rr0) <= rtol) {
      rel = rr0 > 0 ? std::sqrt(std::max(rr, 0.0) / rr0) : 0.0;
      break;
    }
    for (int j = 0; j < s; j++) {
      shift(pc, bp);
      const double pap = gdot(pc, bp);
      if (!(pap > 0.0)) return fail(MPKX_EINVAL, "sstep_solve: p'Ap = %g (matrix not SPD or basis lost)", pap);
      const double alpha = rr / pap;
      for (int i = 0; i < m; i++) {
        xc[i] += alpha * pc[i];
        rn[i] = rc_[i] - alpha * bp[i];
      }
      const double rr1 = gdot(rn, rn);
      const double beta = rr1 / rr;
      for (int i = 0; i < m; i++) pc[i] = rn[i] + beta * pc[i];
      rc_ = rn;
      rr = rr1;
    }
    Coef c{};
    for (int i = 0; i < m; i++) {
      c.c[0][i] = xc[i];
      c.c[1][i] = rc_[i];
      c.c[2][i] = pc[i];
    }
    hipLaunchKernelGGL(k_combine3, dim3(nb), dim3(256), 0, st, S->vp.p, np, S->vr.p, nr, n, c, x, S->r.p, S->p.p);
    MPKX_HIP(hipGetLastError());
    rel = std::sqrt(std::max(rr, 0.0) / rr0);
    if (rel <= rtol) {
      outer++;
      break;
    }
  }
  MPKX_HIP(hipStreamSynchronize(st));
  if (iters) *iters = outer * s;
  if (resid) *resid = rel;
  return MPKX_OK;
}

}  // extern "C"
