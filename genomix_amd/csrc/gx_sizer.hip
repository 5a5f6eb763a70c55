// Learned table sizing: a small FP32 multilayer perceptron with hand-written
// forward, backward and Adam kernels.
//
// The network maps the sizing features of a read set (a HyperLogLog sketch
// of its first reads plus k, read length, partitions; built in gx_build.hip)
// to log(distinct k-mers / k-mer occurrences) and log(partition skew).  It is
// trained on exact distinct counts that the graph build itself produces; the
// build uses its prediction to size the hash table up front and keeps
// grow-and-rebuild as the fallback when the prediction is too small, so the
// graph output never depends on it.  The reference has no counterpart: its
// external sort spills and its lists grow on demand.
//
// Network: x (din) -> tanh(W1 x + b1) (h) -> tanh(W2 a1 + b2) (h) -> W3 a2 + b3 (dout).
// Parameters, flat, row-major [out][in]: W1, b1, W2, b2, W3, b3.
// Loss: L = 1/(2B) * sum_b sum_o (y_bo - t_bo)^2.
//
// Sizes are tiny (din <= 256, h <= 128, dout <= 8); one block per sample for
// the per-sample passes, one thread per parameter for the batch reductions
// (deterministic sums, no float atomics).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "gx_prims.h"
#include "gx_sizer.h"

namespace gx {

constexpr int kMaxIn = 256, kMaxHidden = 128, kMaxOut = 8;

struct MlpDims {
  int din, h, dout;
  __host__ __device__ size_t w1() const { return 0; }
  __host__ __device__ size_t b1() const { return (size_t)din * h; }
  __host__ __device__ size_t w2() const { return b1() + h; }
  __host__ __device__ size_t b2() const { return w2() + (size_t)h * h; }
  __host__ __device__ size_t w3() const { return b2() + h; }
  __host__ __device__ size_t b3() const { return w3() + (size_t)dout * h; }
  __host__ __device__ size_t total() const { return b3() + dout; }
};

// Forward pass, one block (kMaxHidden threads) per sample.  Keeps the
// activations for the backward pass when a1/a2 are given.
__global__ void __launch_bounds__(kMaxHidden) k_mlp_fwd(MlpDims d, int B, const float* __restrict__ X,
                                                        const float* __restrict__ P, float* __restrict__ A1,
                                                        float* __restrict__ A2, float* __restrict__ Y) {
  __shared__ float x[kMaxIn], a1[kMaxHidden], a2[kMaxHidden];
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= B) return;
  for (int i = j; i < d.din; i += blockDim.x) x[i] = X[(size_t)b * d.din + i];
  __syncthreads();
  if (j < d.h) {
    const float* w = P + d.w1() + (size_t)j * d.din;
    float z = P[d.b1() + j];
    for (int i = 0; i < d.din; i++) z = fmaf(w[i], x[i], z);
    a1[j] = tanhf(z);
    if (A1) A1[(size_t)b * d.h + j] = a1[j];
  }
  __syncthreads();
  if (j < d.h) {
    const float* w = P + d.w2() + (size_t)j * d.h;
    float z = P[d.b2() + j];
    for (int i = 0; i < d.h; i++) z = fmaf(w[i], a1[i], z);
    a2[j] = tanhf(z);
    if (A2) A2[(size_t)b * d.h + j] = a2[j];
  }
  __syncthreads();
  if (j < d.dout) {
    const float* w = P + d.w3() + (size_t)j * d.h;
    float z = P[d.b3() + j];
    for (int i = 0; i < d.h; i++) z = fmaf(w[i], a2[i], z);
    Y[(size_t)b * d.dout + j] = z;
  }
}

// Per-sample deltas: D3 = dL/dz3 = (y - t)/B, D2 = (W3^T D3) * (1 - a2^2),
// D1 = (W2^T D2) * (1 - a1^2); per-sample loss terms.
__global__ void __launch_bounds__(kMaxHidden) k_mlp_delta(MlpDims d, int B, const float* __restrict__ T,
                                                          const float* __restrict__ P,
                                                          const float* __restrict__ A1,
                                                          const float* __restrict__ A2,
                                                          const float* __restrict__ Y, float* __restrict__ D1,
                                                          float* __restrict__ D2, float* __restrict__ D3,
                                                          float* __restrict__ lossb) {
  __shared__ float d3[kMaxOut], d2[kMaxHidden];
  const int b = blockIdx.x, j = threadIdx.x;
  if (b >= B) return;
  if (j == 0) {
    float l = 0.f;
    for (int o = 0; o < d.dout; o++) {
      float e = Y[(size_t)b * d.dout + o] - T[(size_t)b * d.dout + o];
      d3[o] = e / (float)B;
      D3[(size_t)b * d.dout + o] = d3[o];
      l = fmaf(e, e, l);
    }
    lossb[b] = 0.5f * l / (float)B;
  }
  __syncthreads();
  if (j < d.h) {
    float s = 0.f;
    for (int o = 0; o < d.dout; o++) s = fmaf(P[d.w3() + (size_t)o * d.h + j], d3[o], s);
    float a = A2[(size_t)b * d.h + j];
    d2[j] = s * (1.f - a * a);
    D2[(size_t)b * d.h + j] = d2[j];
  }
  __syncthreads();
  if (j < d.h) {
    float s = 0.f;
    for (int o = 0; o < d.h; o++) s = fmaf(P[d.w2() + (size_t)o * d.h + j], d2[o], s);
    float a = A1[(size_t)b * d.h + j];
    D1[(size_t)b * d.h + j] = s * (1.f - a * a);
  }
}

// Gradient of every parameter, summed over the batch in sample order:
// dW[o][i] = sum_b D[b][o] * in[b][i], db[o] = sum_b D[b][o].
__global__ void k_mlp_grad(MlpDims d, int B, const float* __restrict__ X, const float* __restrict__ A1,
                           const float* __restrict__ A2, const float* __restrict__ D1,
                           const float* __restrict__ D2, const float* __restrict__ D3, float* __restrict__ G) {
  const size_t n = d.total();
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const float *Dv, *In;
    int nout, nin, o, i;
    bool bias;
    if (p < d.b1()) { Dv = D1; In = X; nout = d.h; nin = d.din; o = (int)(p / d.din); i = (int)(p % d.din); bias = false; }
    else if (p < d.w2()) { Dv = D1; In = nullptr; nout = d.h; nin = 0; o = (int)(p - d.b1()); i = 0; bias = true; }
    else if (p < d.b2()) { size_t q = p - d.w2(); Dv = D2; In = A1; nout = d.h; nin = d.h; o = (int)(q / d.h); i = (int)(q % d.h); bias = false; }
    else if (p < d.w3()) { Dv = D2; In = nullptr; nout = d.h; nin = 0; o = (int)(p - d.b2()); i = 0; bias = true; }
    else if (p < d.b3()) { size_t q = p - d.w3(); Dv = D3; In = A2; nout = d.dout; nin = d.h; o = (int)(q / d.h); i = (int)(q % d.h); bias = false; }
    else { Dv = D3; In = nullptr; nout = d.dout; nin = 0; o = (int)(p - d.b3()); i = 0; bias = true; }
    float s = 0.f;
    for (int b = 0; b < B; b++) {
      float dv = Dv[(size_t)b * nout + o];
      s = bias ? s + dv : fmaf(dv, In[(size_t)b * nin + i], s);
    }
    G[p] = s;
  }
}

__global__ void k_loss_sum(int B, const float* __restrict__ lossb, float* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += lossb[b];
    *out = s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction; step t starts at 1.
__global__ void k_adam(size_t n, float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                       float* __restrict__ V, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    float g = G[p];
    float m = b1 * M[p] + (1.f - b1) * g;
    float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
  }
}

}  // namespace gx

using namespace gx;

struct gx_mlp {
  MlpDims d{};
  int device = 0;
  hipStream_t st = nullptr;
  std::string err;
  DevBuf P, G, M, V, X, T, A1, A2, Y, D1, D2, D3, lossb, loss;
  int64_t step = 0;
  int bcap = 0;
};

namespace {

template <class F>
int mlp_guard(gx_mlp* m, F&& f) {
  try {
    f();
    return GX_OK;
  } catch (const HipError& e) {
    m->err = e.what();
    return GX_EDEVICE;
  } catch (const std::bad_alloc&) {
    m->err = "out of memory";
    return GX_ENOMEM;
  } catch (const std::exception& e) {
    m->err = e.what();
    return GX_EINVAL;
  }
}

struct MlpArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void ensure_batch(gx_mlp* m, int B) {
  if (B <= 0) throw MlpArgError("batch must be positive");
  if (B <= m->bcap) return;
  const MlpDims& d = m->d;
  m->X.ensure((size_t)B * d.din * 4);
  m->T.ensure((size_t)B * d.dout * 4);
  m->A1.ensure((size_t)B * d.h * 4);
  m->A2.ensure((size_t)B * d.h * 4);
  m->Y.ensure((size_t)B * d.dout * 4);
  m->D1.ensure((size_t)B * d.h * 4);
  m->D2.ensure((size_t)B * d.h * 4);
  m->D3.ensure((size_t)B * d.dout * 4);
  m->lossb.ensure((size_t)B * 4);
  m->bcap = B;
}

// Device arrays of one gradient pass: inputs X [B][din], T [B][dout];
// scratch A1, A2, D1, D2 [B][h], Y, D3 [B][dout], lossb [B] (every element
// written before it is read); outputs G [params] and, when given, loss.
struct GradBufs {
  const float *X, *T;
  float *A1, *A2, *Y, *D1, *D2, *D3, *lossb, *G, *loss;
};

// forward + deltas + gradient + loss sum, enqueued on st
void mlp_gradient_on(const gx_mlp* m, const GradBufs& g, int B, hipStream_t st) {
  const MlpDims& d = m->d;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(B), dim3(kMaxHidden), 0, st, d, B, g.X, m->P.as<float>(), g.A1, g.A2, g.Y);
  GX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_delta, dim3(B), dim3(kMaxHidden), 0, st, d, B, g.T, m->P.as<float>(), g.A1, g.A2, g.Y,
                     g.D1, g.D2, g.D3, g.lossb);
  GX_HIP(hipGetLastError());
  size_t n = d.total();
  hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, B, g.X, g.A1, g.A2, g.D1,
                     g.D2, g.D3, g.G);
  GX_HIP(hipGetLastError());
  if (g.loss) {
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, st, B, g.lossb, g.loss);
    GX_HIP(hipGetLastError());
  }
}

// forward + deltas + gradient into m->G; returns the loss
float mlp_gradient(gx_mlp* m, const float* x, const float* t, int B) {
  const MlpDims& d = m->d;
  ensure_batch(m, B);
  GX_HIP(hipMemcpyAsync(m->X.p, x, (size_t)B * d.din * 4, hipMemcpyHostToDevice, m->st));
  GX_HIP(hipMemcpyAsync(m->T.p, t, (size_t)B * d.dout * 4, hipMemcpyHostToDevice, m->st));
  const GradBufs g{m->X.as<float>(), m->T.as<float>(), m->A1.as<float>(), m->A2.as<float>(), m->Y.as<float>(),
                   m->D1.as<float>(), m->D2.as<float>(), m->D3.as<float>(), m->lossb.as<float>(), m->G.as<float>(),
                   m->loss.as<float>()};
  mlp_gradient_on(m, g, B, m->st);
  float l = 0.f;
  GX_HIP(hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, m->st));
  GX_HIP(hipStreamSynchronize(m->st));
  return l;
}

// Scratch of the allocation-free gradient (gx_mlp_gradient_noalloc): one
// array per device, holding A1, A2, D1, D2, Y, D3 and the per-sample loss of
// a batch of at most kNoAllocBatch samples at the largest network
// dimensions (8.3 MiB).  In host code `scratch` only names the symbol; its
// device address comes from hipGetSymbolAddress, once per device.
constexpr int kNoAllocBatch = 4096;
constexpr size_t kNoAllocFloats =
    (size_t)kNoAllocBatch * (4 * kMaxHidden + 2 * kMaxOut + 1);
static __device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (which is `device`).
float* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw MlpArgError("device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    GX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array", (int)hipErrorInvalidSymbol);
    g_scratch_dev[device] = static_cast<float*>(p);
  }
  return g_scratch_dev[device];
}

}  // namespace

// Device-side inference for the graph build (same stream ordering as the
// caller's stream: the build passes its own).
int gx_mlp_forward_device(gx_mlp* m, const float* dX, int B, float* dY, hipStream_t st) {
  return mlp_guard(m, [&] {
    GX_HIP(hipSetDevice(m->device));
    hipLaunchKernelGGL(k_mlp_fwd, dim3(B), dim3(kMaxHidden), 0, st, m->d, B, dX, m->P.as<float>(), nullptr, nullptr,
                       dY);
    GX_HIP(hipGetLastError());
  });
}

int gx_mlp_device_of(const gx_mlp* m) { return m->device; }
int gx_mlp_dims(const gx_mlp* m, int* din, int* dout) {
  *din = m->d.din;
  *dout = m->d.dout;
  return GX_OK;
}

extern "C" {

int gx_mlp_create(int din, int hidden, int dout, int device, gx_mlp** out) {
  if (!out || din <= 0 || din > kMaxIn || hidden <= 0 || hidden > kMaxHidden || dout <= 0 || dout > kMaxOut)
    return GX_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GX_EDEVICE;
  if (device < 0 || device >= ndev) return GX_EINVAL;
  gx_mlp* m = new (std::nothrow) gx_mlp;
  if (!m) return GX_ENOMEM;
  m->d = MlpDims{din, hidden, dout};
  m->device = device;
  int rc = mlp_guard(m, [&] {
    GX_HIP(hipSetDevice(device));
    GX_HIP(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    size_t n = m->d.total() * 4;
    m->P.ensure(n);
    m->G.ensure(n);
    m->M.ensure(n);
    m->V.ensure(n);
    m->loss.ensure(4);
    GX_HIP(hipMemsetAsync(m->P.p, 0, n, m->st));
    GX_HIP(hipMemsetAsync(m->M.p, 0, n, m->st));
    GX_HIP(hipMemsetAsync(m->V.p, 0, n, m->st));
    GX_HIP(hipStreamSynchronize(m->st));
  });
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return GX_OK;
}

size_t gx_mlp_num_params(const gx_mlp* m) { return m ? m->d.total() : 0; }

int gx_mlp_set_params(gx_mlp* m, const float* p, size_t n) {
  if (!m || !p) return GX_EINVAL;
  return mlp_guard(m, [&] {
    if (n != m->d.total()) throw MlpArgError("parameter count mismatch");
    GX_HIP(hipSetDevice(m->device));
    GX_HIP(hipMemcpyAsync(m->P.p, p, n * 4, hipMemcpyHostToDevice, m->st));
    GX_HIP(hipMemsetAsync(m->M.p, 0, n * 4, m->st));
    GX_HIP(hipMemsetAsync(m->V.p, 0, n * 4, m->st));
    GX_HIP(hipStreamSynchronize(m->st));
    m->step = 0;
  });
}

int gx_mlp_get_params(gx_mlp* m, float* p, size_t n) {
  if (!m || !p) return GX_EINVAL;
  return mlp_guard(m, [&] {
    if (n != m->d.total()) throw MlpArgError("parameter count mismatch");
    GX_HIP(hipSetDevice(m->device));
    GX_HIP(hipMemcpyAsync(p, m->P.p, n * 4, hipMemcpyDeviceToHost, m->st));
    GX_HIP(hipStreamSynchronize(m->st));
  });
}

int gx_mlp_forward(gx_mlp* m, const float* x, int batch, float* y) {
  if (!m || !x || !y) return GX_EINVAL;
  return mlp_guard(m, [&] {
    GX_HIP(hipSetDevice(m->device));
    ensure_batch(m, batch);
    const MlpDims& d = m->d;
    GX_HIP(hipMemcpyAsync(m->X.p, x, (size_t)batch * d.din * 4, hipMemcpyHostToDevice, m->st));
    hipLaunchKernelGGL(k_mlp_fwd, dim3(batch), dim3(kMaxHidden), 0, m->st, d, batch, m->X.as<float>(),
                       m->P.as<float>(), nullptr, nullptr, m->Y.as<float>());
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(y, m->Y.p, (size_t)batch * d.dout * 4, hipMemcpyDeviceToHost, m->st));
    GX_HIP(hipStreamSynchronize(m->st));
  });
}

int gx_mlp_gradient(gx_mlp* m, const float* x, const float* t, int batch, float* grad, float* loss) {
  if (!m || !x || !t) return GX_EINVAL;
  return mlp_guard(m, [&] {
    GX_HIP(hipSetDevice(m->device));
    float l = mlp_gradient(m, x, t, batch);
    if (loss) *loss = l;
    if (grad) {
      GX_HIP(hipMemcpyAsync(grad, m->G.p, m->d.total() * 4, hipMemcpyDeviceToHost, m->st));
      GX_HIP(hipStreamSynchronize(m->st));
    }
  });
}

int gx_mlp_gradient_noalloc(gx_mlp* m, const float* x, const float* t, int batch, float* grad, float* loss,
                            void* stream) {
  if (!m || !x || !t) return GX_EINVAL;
  return mlp_guard(m, [&] {
    if (batch <= 0) throw MlpArgError("batch must be positive");
    if (batch > kNoAllocBatch)
      throw MlpArgError("batch above " + std::to_string(kNoAllocBatch) + " samples, the scratch array's limit");
    GX_HIP(hipSetDevice(m->device));
    float* sp = scratch_on(m->device);
    const MlpDims& d = m->d;
    const size_t bh = (size_t)batch * d.h, bo = (size_t)batch * d.dout;
    GradBufs g{x, t, sp, sp + bh, sp + 4 * bh, sp + 2 * bh, sp + 3 * bh, sp + 4 * bh + bo, sp + 4 * bh + 2 * bo,
               grad ? grad : m->G.as<float>(), loss};
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : m->st;
    mlp_gradient_on(m, g, batch, st);
    if (!stream) GX_HIP(hipStreamSynchronize(st));
  });
}

int gx_mlp_train_step(gx_mlp* m, const float* x, const float* t, int batch, float lr, float* loss) {
  if (!m || !x || !t) return GX_EINVAL;
  return mlp_guard(m, [&] {
    if (!(lr > 0.f)) throw MlpArgError("learning rate must be positive");
    GX_HIP(hipSetDevice(m->device));
    float l = mlp_gradient(m, x, t, batch);
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    size_t n = m->d.total();
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->st, n, m->P.as<float>(),
                       m->G.as<float>(), m->M.as<float>(), m->V.as<float>(), lr, b1, b2, eps, c1, c2);
    GX_HIP(hipGetLastError());
    GX_HIP(hipStreamSynchronize(m->st));
    if (loss) *loss = l;
  });
}

const char* gx_mlp_last_error(const gx_mlp* m) { return m ? m->err.c_str() : "null handle"; }

void gx_mlp_destroy(gx_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->st) {
    (void)hipStreamSynchronize(m->st);
    (void)hipStreamDestroy(m->st);
  }
  delete m;
}

}  // extern "C"
