// dev_mlp.hip -- the tuning cost model: a 24 -> 64 -> 64 -> 3 MLP (ReLU
// hidden layers, linear heads) trained and evaluated on the GPU (no
// reference counterpart).
//
// Parameter vector (FP32, row-major out x in):
//   W1[64][24] b1[64] W2[64][64] b2[64] W3[3][64] b3[3]
// loss = 1/(2N) sum_n sum_o (y_no - t_no)^2.
//
// One 64-lane wavefront per sample for the forward pass and for the
// back-propagated deltas (hidden width = wavefront width: lane j owns unit
// j); the weight gradients are one lane per parameter summing over the
// samples in a fixed order, so every step is deterministic.  Adam updates
// the parameters in place.
#include <cmath>
#include <memory>
#include <vector>

#include "dev_common.h"
#include "internal.h"

using namespace mpkx;

namespace {

constexpr int NI = 24, NH = 64, NO = 3;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NI, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NO * NH, NPARAM = OFF_B3 + NO;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// h1, h2: [N][64] post-ReLU activations; y: [N][3].
__global__ __launch_bounds__(64) void k_mlp_forward(int N, const float *__restrict__ prm,
                                                    const float *__restrict__ X, float *__restrict__ h1,
                                                    float *__restrict__ h2, float *__restrict__ y) {
  __shared__ float xs[NI], a1[NH];
  const int s = blockIdx.x, j = threadIdx.x;
  if (s >= N) return;
  if (j < NI) xs[j] = X[(size_t)s * NI + j];
  __syncthreads();
  float a = prm[OFF_B1 + j];
  for (int i = 0; i < NI; i++) a = fmaf(prm[OFF_W1 + j * NI + i], xs[i], a);
  a = fmaxf(a, 0.0f);
  a1[j] = a;
  if (h1) h1[(size_t)s * NH + j] = a;
  __syncthreads();
  float b = prm[OFF_B2 + j];
  for (int i = 0; i < NH; i++) b = fmaf(prm[OFF_W2 + j * NH + i], a1[i], b);
  b = fmaxf(b, 0.0f);
  if (h2) h2[(size_t)s * NH + j] = b;
#pragma unroll
  for (int o = 0; o < NO; o++) {
    float t = wave_sum(prm[OFF_W3 + o * NH + j] * b);
    if (j == 0) y[(size_t)s * NO + o] = t + prm[OFF_B3 + o];
  }
}

// dy: [N][3], d2, d1: [N][64]; sl: per-sample loss.
__global__ __launch_bounds__(64) void k_mlp_delta(int N, const float *__restrict__ prm,
                                                  const float *__restrict__ T, const float *__restrict__ y,
                                                  const float *__restrict__ h1, const float *__restrict__ h2,
                                                  float *__restrict__ dy, float *__restrict__ d2,
                                                  float *__restrict__ d1, float *__restrict__ sl) {
  __shared__ float e[NO], g2[NH];
  const int s = blockIdx.x, j = threadIdx.x;
  if (s >= N) return;
  const float invN = 1.0f / (float)N;
  if (j < NO) {
    float r = y[(size_t)s * NO + j] - T[(size_t)s * NO + j];
    e[j] = r * invN;
    dy[(size_t)s * NO + j] = r * invN;
  }
  __syncthreads();
  if (j == 0) {
    float l = 0.0f;
    for (int o = 0; o < NO; o++) {
      float r = y[(size_t)s * NO + o] - T[(size_t)s * NO + o];
      l = fmaf(r, r, l);
    }
    sl[s] = 0.5f * l * invN;
  }
  float a = 0.0f;
#pragma unroll
  for (int o = 0; o < NO; o++) a = fmaf(prm[OFF_W3 + o * NH + j], e[o], a);
  a = h2[(size_t)s * NH + j] > 0.0f ? a : 0.0f;
  g2[j] = a;
  d2[(size_t)s * NH + j] = a;
  __syncthreads();
  float c = 0.0f;
  for (int k = 0; k < NH; k++) c = fmaf(prm[OFF_W2 + k * NH + j], g2[k], c);
  d1[(size_t)s * NH + j] = h1[(size_t)s * NH + j] > 0.0f ? c : 0.0f;
}

// One lane per parameter: sum over samples in index order.
__global__ __launch_bounds__(256) void k_mlp_wgrad(int N, const float *__restrict__ X,
                                                   const float *__restrict__ h1, const float *__restrict__ h2,
                                                   const float *__restrict__ dy, const float *__restrict__ d2,
                                                   const float *__restrict__ d1, float *__restrict__ grad) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NPARAM) return;
  const float *dv, *av;
  int ld_d, ld_a, jd, ja;  // grad = sum_n dv[n*ld_d + jd] * (av ? av[n*ld_a + ja] : 1)
  if (p < OFF_B1) {
    dv = d1, ld_d = NH, jd = p / NI, av = X, ld_a = NI, ja = p % NI;
  } else if (p < OFF_W2) {
    dv = d1, ld_d = NH, jd = p - OFF_B1, av = nullptr, ld_a = 0, ja = 0;
  } else if (p < OFF_B2) {
    int q = p - OFF_W2;
    dv = d2, ld_d = NH, jd = q / NH, av = h1, ld_a = NH, ja = q % NH;
  } else if (p < OFF_W3) {
    dv = d2, ld_d = NH, jd = p - OFF_B2, av = nullptr, ld_a = 0, ja = 0;
  } else if (p < OFF_B3) {
    int q = p - OFF_W3;
    dv = dy, ld_d = NO, jd = q / NH, av = h2, ld_a = NH, ja = q % NH;
  } else {
    dv = dy, ld_d = NO, jd = p - OFF_B3, av = nullptr, ld_a = 0, ja = 0;
  }
  float g = 0.0f;
  if (av)
    for (int n = 0; n < N; n++) g = fmaf(dv[(size_t)n * ld_d + jd], av[(size_t)n * ld_a + ja], g);
  else
    for (int n = 0; n < N; n++) g += dv[(size_t)n * ld_d + jd];
  grad[p] = g;
}

__global__ __launch_bounds__(256) void k_sum_loss(int N, const float *__restrict__ sl, float *__restrict__ out) {
  __shared__ float red[256];
  float a = 0.0f;
  for (int i = threadIdx.x; i < N; i += 256) a += sl[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ prm, const float *__restrict__ grad,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2, float wd) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= NPARAM) return;
  float g = grad[p] + wd * prm[p];
  float mm = b1 * m[p] + (1.0f - b1) * g;
  float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  prm[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

}  // namespace

struct mpkx_mlp {
  int cap = 0;  // samples the activation buffers hold
  int step = 0;
  DevArray<float> prm, grad, m, v, h1, h2, y, dy, d1, d2, sl, loss;
};

static int mlp_reserve(mpkx_mlp *m, int N) {
  if (N <= m->cap) return MPKX_OK;
  int rc;
  size_t h = (size_t)N * NH, o = (size_t)N * NO;
  if ((rc = m->h1.alloc(h)) || (rc = m->h2.alloc(h)) || (rc = m->d1.alloc(h)) || (rc = m->d2.alloc(h)) ||
      (rc = m->y.alloc(o)) || (rc = m->dy.alloc(o)) || (rc = m->sl.alloc(N)))
    return rc;
  m->cap = N;
  return MPKX_OK;
}

extern "C" {

long mpkx_mlp_nparam(void) { return NPARAM; }

int mpkx_mlp_create(mpkx_mlp **out, const float *params) {
  clear_error();
  if (!out || !params) return fail(MPKX_EINVAL, "mlp_create: bad arguments");
  auto *m = new (std::nothrow) mpkx_mlp;
  if (!m) return fail(MPKX_ENOMEM, "mlp_create");
  std::unique_ptr<mpkx_mlp> guard(m);
  std::vector<float> zero(NPARAM, 0.0f);
  int rc;
  if ((rc = m->prm.upload(params, NPARAM)) || (rc = m->grad.upload(zero)) || (rc = m->m.upload(zero)) ||
      (rc = m->v.upload(zero)) || (rc = m->loss.upload(zero.data(), 1)))
    return rc;
  *out = guard.release();
  return MPKX_OK;
}

void mpkx_mlp_destroy(mpkx_mlp *m) { delete m; }

int mpkx_mlp_set_params(mpkx_mlp *m, const float *params) {
  clear_error();
  if (!m || !params) return fail(MPKX_EINVAL, "mlp_set_params: bad arguments");
  MPKX_HIP(hipMemcpy(m->prm.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  std::vector<float> zero(NPARAM, 0.0f);
  MPKX_HIP(hipMemcpy(m->m.p, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  MPKX_HIP(hipMemcpy(m->v.p, zero.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  m->step = 0;
  return MPKX_OK;
}

// Host copies of the parameters and of the last gradient (either may be NULL).
int mpkx_mlp_get(mpkx_mlp *m, float *params, float *grad) {
  clear_error();
  if (!m) return fail(MPKX_EINVAL, "mlp_get: null");
  MPKX_HIP(hipDeviceSynchronize());
  if (params) MPKX_HIP(hipMemcpy(params, m->prm.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (grad) MPKX_HIP(hipMemcpy(grad, m->grad.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return MPKX_OK;
}

// Inference: X device [N][24] -> Y device [N][3].
int mpkx_mlp_forward(mpkx_mlp *m, int N, const float *X, float *Y, void *stream) {
  clear_error();
  if (!m || N <= 0 || !X || !Y) return fail(MPKX_EINVAL, "mlp_forward: bad arguments");
  hipLaunchKernelGGL(k_mlp_forward, dim3(N), dim3(64), 0, (hipStream_t)stream, N, m->prm.p, X, nullptr, nullptr, Y);
  MPKX_HIP(hipGetLastError());
  return MPKX_OK;
}

// Loss and gradient at the current parameters: X [N][24], T [N][3] device;
// *loss (host) receives the loss.
int mpkx_mlp_loss_grad(mpkx_mlp *m, int N, const float *X, const float *T, float *loss, void *stream) {
  clear_error();
  if (!m || N <= 0 || !X || !T) return fail(MPKX_EINVAL, "mlp_loss_grad: bad arguments");
  int rc = mlp_reserve(m, N);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_forward, dim3(N), dim3(64), 0, st, N, m->prm.p, X, m->h1.p, m->h2.p, m->y.p);
  hipLaunchKernelGGL(k_mlp_delta, dim3(N), dim3(64), 0, st, N, m->prm.p, T, m->y.p, m->h1.p, m->h2.p, m->dy.p,
                     m->d2.p, m->d1.p, m->sl.p);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3((NPARAM + 255) / 256), dim3(256), 0, st, N, X, m->h1.p, m->h2.p, m->dy.p,
                     m->d2.p, m->d1.p, m->grad.p);
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, st, N, m->sl.p, m->loss.p);
  MPKX_HIP(hipGetLastError());
  if (loss) {
    MPKX_HIP(hipMemcpyAsync(loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
    MPKX_HIP(hipStreamSynchronize(st));
  }
  return MPKX_OK;
}

// One Adam step (bias-corrected; weight_decay * params added to the
// gradient, i.e. L2 regularisation) with the gradient of the last
// mpkx_mlp_loss_grad call.
int mpkx_mlp_adam(mpkx_mlp *m, float lr, float beta1, float beta2, float eps, float weight_decay, void *stream) {
  clear_error();
  if (!m || !(lr > 0)) return fail(MPKX_EINVAL, "mlp_adam: bad arguments");
  m->step++;
  float c1 = 1.0f / (1.0f - std::pow(beta1, (float)m->step));
  float c2 = 1.0f / (1.0f - std::pow(beta2, (float)m->step));
  hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, (hipStream_t)stream, m->prm.p, m->grad.p,
                     m->m.p, m->v.p, lr, beta1, beta2, eps, c1, c2, weight_decay);
  MPKX_HIP(hipGetLastError());
  return MPKX_OK;
}

}  // extern "C"
