// fot_mlp.hip -- the learned cost model: a small FP32 MLP
// (in -> hidden -> hidden -> out, ReLU) trained and evaluated on the GPU by
// hand-written kernels.  Every matrix product runs on the matrix cores
// (v_mfma_f32_16x16x4_f32, one 16x16 tile per workgroup, K split across its
// waves); bias, ReLU, the masked MSE gradient and Adam are elementwise
// kernels.  No reference counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "fot_hip.hpp"

using fot::fail;

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]) (ReLU), with A(m,k) =
// A[m*sam + k*sak] and B(k,n) = B[k*sbk + n*sbn]: any transpose through the
// strides.  One 16x16 tile of C per workgroup of kGemmWaves waves; wave w
// takes the K steps w, w + kGemmWaves, ... (four k each), so a long K (the
// weight gradients reduce over the whole batch) is split across the waves,
// and the partial tiles are summed in wave order through LDS.
// MFMA 16x16x4 f32 operand layout: lane l supplies A(i = l%16, k = l/16) and
// B(k = l/16, j = l%16); its four results are C(4*(l/16) + v, l%16).
constexpr uint32_t kGemmWaves = 8;

__global__ __launch_bounds__(64 * kGemmWaves) void gemm_mfma_kernel(
    const float *__restrict__ A, const float *__restrict__ B, float *__restrict__ C,
    const float *__restrict__ bias, uint32_t M, uint32_t N, uint32_t K, uint32_t sam,
    uint32_t sak, uint32_t sbk, uint32_t sbn, uint32_t ldc, int relu, int accumulate) {
  __shared__ v4f part[kGemmWaves][64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t m0 = blockIdx.x * 16u, n0 = blockIdx.y * 16u;
  const uint32_t r = lane & 15u, q = lane >> 4;
  const uint32_t am = m0 + r, bn = n0 + r;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  for (uint32_t k0 = 4u * wave; k0 < K; k0 += 4u * kGemmWaves) {
    const uint32_t k = k0 + q;
    const float a = (am < M && k < K) ? A[size_t(am) * sam + size_t(k) * sak] : 0.f;
    const float b = (bn < N && k < K) ? B[size_t(k) * sbk + size_t(bn) * sbn] : 0.f;
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  v4f sum = part[0][lane];
#pragma unroll
  for (uint32_t w = 1; w < kGemmWaves; ++w) sum += part[w][lane];
  const uint32_t col = n0 + r;
  if (col >= N) return;
  const float bb = bias ? bias[col] : 0.f;
#pragma unroll
  for (uint32_t v = 0; v < 4; ++v) {
    const uint32_t row = m0 + 4u * q + v;
    if (row >= M) continue;
    float x = sum[v] + bb;
    if (relu) x = x > 0.f ? x : 0.f;
    float *c = C + size_t(row) * ldc + col;
    *c = accumulate ? *c + x : x;
  }
}

// dY = 2 (Y - T) mask / count; per-block partial sums of mask (Y - T)^2
__global__ __launch_bounds__(256) void mse_grad_kernel(const float *__restrict__ Y,
                                                       const float *__restrict__ T,
                                                       const float *__restrict__ mask,
                                                       float *__restrict__ dY, uint32_t n,
                                                       float inv_count,
                                                       float *__restrict__ partial) {
  __shared__ float s[256];
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    const float w = mask ? mask[i] : 1.f;
    const float e = (Y[i] - T[i]) * w;
    dY[i] = 2.f * e * inv_count;
    l = e * (Y[i] - T[i]);
  }
  s[threadIdx.x] = l;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// dZ = dH * (Z > 0) in place on dH (Z is the pre-activation's output after
// ReLU: H > 0 exactly where Z > 0)
__global__ __launch_bounds__(256) void relu_grad_kernel(float *__restrict__ dH,
                                                        const float *__restrict__ H, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n && !(H[i] > 0.f)) dH[i] = 0.f;
}

// column sums: db[j] = sum_i dZ[i][j]
__global__ __launch_bounds__(256) void colsum_kernel(const float *__restrict__ dZ,
                                                     float *__restrict__ db, uint32_t rows,
                                                     uint32_t cols) {
  __shared__ float s[256];
  const uint32_t j = blockIdx.x;
  float acc = 0.f;
  for (uint32_t i = threadIdx.x; i < rows; i += 256) acc += dZ[size_t(i) * cols + j];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = s[0];
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                                   float *__restrict__ m, float *__restrict__ v,
                                                   uint32_t n, float lr_t, float b1, float b2,
                                                   float eps, float wd) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi, v[i] = vi;
  p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// C (MxN) = op(A) op(B) through strides; see gemm_mfma_kernel
fot_status gemm(const float *A, const float *B, float *C, const float *bias, uint32_t M,
                uint32_t N, uint32_t K, uint32_t sam, uint32_t sak, uint32_t sbk, uint32_t sbn,
                uint32_t ldc, bool relu, bool accumulate, hipStream_t s) {
  if (M == 0 || N == 0) return FOT_OK;
  hipLaunchKernelGGL(gemm_mfma_kernel, dim3(cdiv(M, 16), cdiv(N, 16)), dim3(64 * kGemmWaves), 0, s,
                     A, B, C,
                     bias, M, N, K, sam, sak, sbk, sbn, ldc, relu ? 1 : 0, accumulate ? 1 : 0);
  FOT_HIP(hipGetLastError());
  return FOT_OK;
}

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

} // namespace

// parameter layout (flat, row-major): W1 [H][I], b1 [H], W2 [H][H], b2 [H],
// W3 [O][H], b3 [O]
struct fot_mlp {
  int device = 0; // the device holding the parameters
  uint32_t I = 0, H = 0, O = 0;
  size_t nparam = 0;
  size_t off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  // activations for up to `cap` rows: X, H1, H2, Y, T, mask, dY, dH2, dH1, partials
  uint32_t cap = 0;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr, *Mk = nullptr;
  float *dY = nullptr, *dH2 = nullptr, *dH1 = nullptr, *part = nullptr;
  uint64_t step = 0;
  std::vector<void *> owned;

  ~fot_mlp() {
    for (void *q : owned) (void)hipFree(q);
  }
  fot_status alloc(float **q, size_t n) {
    void *r = nullptr;
    FOT_HIP(hipMalloc(&r, std::max<size_t>(1, n) * sizeof(float)));
    owned.push_back(r);
    *q = static_cast<float *>(r);
    return FOT_OK;
  }
  fot_status reserve(uint32_t rows) {
    if (rows <= cap) return FOT_OK;
    for (float **q : {&X, &H1, &H2, &Y, &T, &Mk, &dY, &dH2, &dH1, &part}) {
      if (*q) {
        owned.erase(std::find(owned.begin(), owned.end(), static_cast<void *>(*q)));
        (void)hipFree(*q);
        *q = nullptr;
      }
    }
    const size_t r = rows;
    fot_status st;
    if ((st = alloc(&X, r * I)) || (st = alloc(&H1, r * H)) || (st = alloc(&H2, r * H)) ||
        (st = alloc(&Y, r * O)) || (st = alloc(&T, r * O)) || (st = alloc(&Mk, r * O)) ||
        (st = alloc(&dY, r * O)) || (st = alloc(&dH2, r * H)) || (st = alloc(&dH1, r * H)) ||
        (st = alloc(&part, cdiv(uint32_t(r * O), 256) + 1)))
      return st;
    cap = rows;
    return FOT_OK;
  }
  // forward over `rows` rows of x (device) through the hidden activations
  // h1, h2 (rows x H each) into y
  fot_status forward_into(const float *x, float *h1, float *h2, float *y, uint32_t rows,
                          hipStream_t s) const {
    fot_status st;
    if ((st = gemm(x, p + off_w1, h1, p + off_b1, rows, H, I, I, 1, 1, I, H, true, false, s)))
      return st;
    if ((st = gemm(h1, p + off_w2, h2, p + off_b2, rows, H, H, H, 1, 1, H, H, true, false, s)))
      return st;
    return gemm(h2, p + off_w3, y, p + off_b3, rows, O, H, H, 1, 1, H, O, false, false, s);
  }
  fot_status forward(uint32_t rows, hipStream_t s) { return forward_into(X, H1, H2, Y, rows, s); }
  // gradients of the masked MSE into g; returns the loss
  fot_status backward(uint32_t rows, bool masked, float count, float *loss, hipStream_t s) {
    const uint32_t n = rows * O, nb = cdiv(n, 256);
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nb), dim3(256), 0, s, Y, T, masked ? Mk : nullptr,
                       dY, n, 1.f / count, part);
    FOT_HIP(hipGetLastError());
    fot_status st;
    // layer 3: dW3 = dY^T H2, db3 = colsum dY, dH2 = dY W3
    if ((st = gemm(dY, H2, g + off_w3, nullptr, O, H, rows, 1, O, H, 1, H, false, false, s)))
      return st;
    hipLaunchKernelGGL(colsum_kernel, dim3(O), dim3(256), 0, s, dY, g + off_b3, rows, O);
    if ((st = gemm(dY, p + off_w3, dH2, nullptr, rows, H, O, O, 1, H, 1, H, false, false, s)))
      return st;
    hipLaunchKernelGGL(relu_grad_kernel, dim3(cdiv(rows * H, 256)), dim3(256), 0, s, dH2, H2,
                       rows * H);
    // layer 2
    if ((st = gemm(dH2, H1, g + off_w2, nullptr, H, H, rows, 1, H, H, 1, H, false, false, s)))
      return st;
    hipLaunchKernelGGL(colsum_kernel, dim3(H), dim3(256), 0, s, dH2, g + off_b2, rows, H);
    if ((st = gemm(dH2, p + off_w2, dH1, nullptr, rows, H, H, H, 1, H, 1, H, false, false, s)))
      return st;
    hipLaunchKernelGGL(relu_grad_kernel, dim3(cdiv(rows * H, 256)), dim3(256), 0, s, dH1, H1,
                       rows * H);
    // layer 1
    if ((st = gemm(dH1, X, g + off_w1, nullptr, H, I, rows, 1, H, I, 1, I, false, false, s)))
      return st;
    hipLaunchKernelGGL(colsum_kernel, dim3(H), dim3(256), 0, s, dH1, g + off_b1, rows, H);
    FOT_HIP(hipGetLastError());
    if (loss) {
      std::vector<float> h(nb);
      FOT_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(float), hipMemcpyDeviceToHost, s));
      FOT_HIP(hipStreamSynchronize(s));
      double acc = 0;
      for (float x : h) acc += x;
      *loss = float(acc / count);
    }
    return FOT_OK;
  }
  fot_status load_rows(const float *x, const float *t, const float *mask, uint32_t rows) {
    if (fot_status st = reserve(rows)) return st;
    FOT_HIP(hipMemcpy(X, x, sizeof(float) * rows * I, hipMemcpyHostToDevice));
    if (t) FOT_HIP(hipMemcpy(T, t, sizeof(float) * rows * O, hipMemcpyHostToDevice));
    if (mask) FOT_HIP(hipMemcpy(Mk, mask, sizeof(float) * rows * O, hipMemcpyHostToDevice));
    return FOT_OK;
  }
};

static float mask_count(const float *mask, size_t n) {
  if (!mask) return float(n);
  double c = 0;
  for (size_t i = 0; i < n; ++i) c += mask[i];
  return float(std::max(c, 1.0));
}

static fot_status fot_mlp_create_impl(uint32_t in, uint32_t hidden, uint32_t out, uint64_t seed,
                                     fot_mlp **mlp) {
  if (!mlp || in == 0 || hidden == 0 || out == 0 || in > 4096 || hidden > 4096 || out > 4096)
    return fail(FOT_ERR_INVALID_ARGUMENT, "bad MLP shape");
  *mlp = nullptr;
  std::unique_ptr<fot_mlp> M(new fot_mlp);
  FOT_HIP(hipGetDevice(&M->device));
  M->I = in, M->H = hidden, M->O = out;
  M->off_w1 = 0;
  M->off_b1 = M->off_w1 + size_t(hidden) * in;
  M->off_w2 = M->off_b1 + hidden;
  M->off_b2 = M->off_w2 + size_t(hidden) * hidden;
  M->off_w3 = M->off_b2 + hidden;
  M->off_b3 = M->off_w3 + size_t(out) * hidden;
  M->nparam = M->off_b3 + out;
  fot_status st;
  if ((st = M->alloc(&M->p, M->nparam)) || (st = M->alloc(&M->g, M->nparam)) ||
      (st = M->alloc(&M->m, M->nparam)) || (st = M->alloc(&M->v, M->nparam)))
    return st;
  // He-uniform weights (fan in), zero biases
  std::vector<float> h(M->nparam, 0.f);
  SplitMix r{seed};
  auto init = [&](size_t off, uint32_t rows, uint32_t fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (size_t i = 0; i < size_t(rows) * fan_in; ++i) h[off + i] = float((2 * r.uniform() - 1) * a);
  };
  init(M->off_w1, hidden, in);
  init(M->off_w2, hidden, hidden);
  init(M->off_w3, out, hidden);
  FOT_HIP(hipMemcpy(M->p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  FOT_HIP(hipMemset(M->m, 0, M->nparam * sizeof(float)));
  FOT_HIP(hipMemset(M->v, 0, M->nparam * sizeof(float)));
  *mlp = M.release();
  return FOT_OK;
}
extern "C" fot_status fot_mlp_create(uint32_t in, uint32_t hidden, uint32_t out, uint64_t seed,
                                     fot_mlp **mlp) {
  return fot::guarded("fot_mlp_create", [&] {
    return fot_mlp_create_impl(in, hidden, out, seed, mlp);
  });
}

static fot_status fot_mlp_destroy_impl(fot_mlp *mlp) {
  delete mlp;
  return FOT_OK;
}
extern "C" fot_status fot_mlp_destroy(fot_mlp *mlp) {
  return fot::guarded("fot_mlp_destroy", [&] { return fot_mlp_destroy_impl(mlp); });
}

static fot_status fot_mlp_num_params_impl(const fot_mlp *mlp, uint64_t *n) {
  if (!mlp || !n) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *n = mlp->nparam;
  return FOT_OK;
}
extern "C" fot_status fot_mlp_num_params(const fot_mlp *mlp, uint64_t *n) {
  return fot::guarded("fot_mlp_num_params", [&] { return fot_mlp_num_params_impl(mlp, n); });
}

static fot_status fot_mlp_get_params_impl(const fot_mlp *mlp, float *host, uint64_t n) {
  if (!mlp || !host || n != mlp->nparam) return fail(FOT_ERR_INVALID_ARGUMENT, "bad params");
  FOT_HIP(hipMemcpy(host, mlp->p, n * sizeof(float), hipMemcpyDeviceToHost));
  return FOT_OK;
}
extern "C" fot_status fot_mlp_get_params(const fot_mlp *mlp, float *host, uint64_t n) {
  return fot::guarded("fot_mlp_get_params", [&] { return fot_mlp_get_params_impl(mlp, host, n); });
}

static fot_status fot_mlp_set_params_impl(fot_mlp *mlp, const float *host, uint64_t n) {
  if (!mlp || !host || n != mlp->nparam) return fail(FOT_ERR_INVALID_ARGUMENT, "bad params");
  FOT_HIP(hipMemcpy(mlp->p, host, n * sizeof(float), hipMemcpyHostToDevice));
  FOT_HIP(hipMemset(mlp->m, 0, mlp->nparam * sizeof(float)));
  FOT_HIP(hipMemset(mlp->v, 0, mlp->nparam * sizeof(float)));
  mlp->step = 0;
  return FOT_OK;
}
extern "C" fot_status fot_mlp_set_params(fot_mlp *mlp, const float *host, uint64_t n) {
  return fot::guarded("fot_mlp_set_params", [&] { return fot_mlp_set_params_impl(mlp, host, n); });
}

static fot_status fot_mlp_forward_impl(fot_mlp *mlp, const float *x, uint32_t rows, float *y) {
  if (!mlp || !x || !y) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  if (rows == 0) return FOT_OK;
  fot_status st;
  if ((st = mlp->load_rows(x, nullptr, nullptr, rows))) return st;
  if ((st = mlp->forward(rows, nullptr))) return st;
  FOT_HIP(hipMemcpy(y, mlp->Y, sizeof(float) * rows * mlp->O, hipMemcpyDeviceToHost));
  return FOT_OK;
}
extern "C" fot_status fot_mlp_forward(fot_mlp *mlp, const float *x, uint32_t rows, float *y) {
  return fot::guarded("fot_mlp_forward", [&] { return fot_mlp_forward_impl(mlp, x, rows, y); });
}

// ---- inference that allocates nothing ------------------------------------
// The two hidden activations of one fot_mlp_forward_noalloc call live in a
// fixed device array instead of buffers the library allocates: each device
// holds its own copy from first use until the process ends.  Host code never
// uses the array's name as an address (on the host it is a stand-in, not
// device memory); it uses the address hipGetSymbolAddress returns for the
// current device, looked up once per device on the entry point's first call
// there.
constexpr size_t kNoAllocScratchFloats = size_t(1) << 20; // 4 MiB per device
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static float *g_scratch_ptr[kMaxScratchDevices] = {};

static fot_status scratch_of_current_device(float **out) {
  int dev = 0;
  FOT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(FOT_ERR_UNSUPPORTED, "device %d beyond the %d scratch slots", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q)
      return fail(FOT_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float *>(q);
  }
  *out = g_scratch_ptr[dev];
  return FOT_OK;
}

static fot_status fot_mlp_forward_noalloc_impl(const fot_mlp *mlp, const float *x,
                                              uint32_t rows, float *y, void *stream) {
  if (!mlp) return fail(FOT_ERR_INVALID_ARGUMENT, "null MLP");
  if (rows == 0) return FOT_OK;
  if (!x || !y) return fail(FOT_ERR_INVALID_ARGUMENT, "null device buffer");
  if (2 * uint64_t(rows) * mlp->H > kNoAllocScratchFloats)
    return fail(FOT_ERR_INVALID_ARGUMENT,
                "rows %u x hidden %u needs %llu activation floats; the scratch holds %zu "
                "(at most %llu rows)",
                rows, mlp->H, (unsigned long long)(2 * uint64_t(rows) * mlp->H),
                kNoAllocScratchFloats,
                (unsigned long long)(kNoAllocScratchFloats / (2 * uint64_t(mlp->H))));
  int dev = 0;
  FOT_HIP(hipGetDevice(&dev));
  if (dev != mlp->device)
    return fail(FOT_ERR_INVALID_ARGUMENT, "current device %d, the MLP lives on device %d", dev,
                mlp->device);
  float *h = nullptr;
  if (fot_status st = scratch_of_current_device(&h)) return st;
  // every element of both activations is written before it is read: no
  // clearing needed
  return mlp->forward_into(x, h, h + size_t(rows) * mlp->H, y, rows,
                           static_cast<hipStream_t>(stream));
}
extern "C" fot_status fot_mlp_forward_noalloc(const fot_mlp *mlp, const float *x, uint32_t rows,
                                              float *y, void *stream) {
  return fot::guarded("fot_mlp_forward_noalloc", [&] {
    return fot_mlp_forward_noalloc_impl(mlp, x, rows, y, stream);
  });
}

static fot_status fot_mlp_loss_grad_impl(fot_mlp *mlp, const float *x, const float *t,
                                        const float *mask, uint32_t rows, float *loss,
                                        float *grad) {
  if (!mlp || !x || !t || !loss || !grad || rows == 0)
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  fot_status st;
  if ((st = mlp->load_rows(x, t, mask, rows))) return st;
  if ((st = mlp->forward(rows, nullptr))) return st;
  const float count = mask_count(mask, size_t(rows) * mlp->O);
  if ((st = mlp->backward(rows, mask != nullptr, count, loss, nullptr))) return st;
  FOT_HIP(hipMemcpy(grad, mlp->g, mlp->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return FOT_OK;
}
extern "C" fot_status fot_mlp_loss_grad(fot_mlp *mlp, const float *x, const float *t,
                                        const float *mask, uint32_t rows, float *loss,
                                        float *grad) {
  return fot::guarded("fot_mlp_loss_grad", [&] {
    return fot_mlp_loss_grad_impl(mlp, x, t, mask, rows, loss, grad);
  });
}

static fot_status fot_mlp_train_impl(fot_mlp *mlp, const float *x, const float *t,
                                    const float *mask, uint32_t rows, uint32_t epochs, float lr,
                                    float weight_decay, float *loss_per_epoch) {
  if (!mlp || !x || !t || rows == 0) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  fot_status st;
  if ((st = mlp->load_rows(x, t, mask, rows))) return st;
  const float count = mask_count(mask, size_t(rows) * mlp->O);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (uint32_t e = 0; e < epochs; ++e) {
    if ((st = mlp->forward(rows, nullptr))) return st;
    float l = 0;
    if ((st = mlp->backward(rows, mask != nullptr, count, loss_per_epoch ? &l : nullptr, nullptr)))
      return st;
    if (loss_per_epoch) loss_per_epoch[e] = l;
    ++mlp->step;
    const double c1 = 1.0 - std::pow(double(b1), double(mlp->step));
    const double c2 = 1.0 - std::pow(double(b2), double(mlp->step));
    const float lr_t = float(lr * std::sqrt(c2) / c1);
    hipLaunchKernelGGL(adam_kernel, dim3(cdiv(uint32_t(mlp->nparam), 256)), dim3(256), 0, nullptr,
                       mlp->p, mlp->g, mlp->m, mlp->v, uint32_t(mlp->nparam), lr_t, b1, b2, eps,
                       weight_decay);
    FOT_HIP(hipGetLastError());
  }
  FOT_HIP(hipDeviceSynchronize());
  return FOT_OK;
}
extern "C" fot_status fot_mlp_train(fot_mlp *mlp, const float *x, const float *t,
                                    const float *mask, uint32_t rows, uint32_t epochs, float lr,
                                    float weight_decay, float *loss_per_epoch) {
  return fot::guarded("fot_mlp_train", [&] {
    return fot_mlp_train_impl(mlp, x, t, mask, rows, epochs, lr, weight_decay, loss_per_epoch);
  });
}

// The MFMA GEMM on device buffers: C[M][N] = op(A) op(B) with row-major A
// (M x K, or K x M when trans_a) and B (K x N, or N x K when trans_b).
static fot_status fot_gemm_f32_impl(const float *A, const float *B, float *C, uint32_t M,
                                   uint32_t N, uint32_t K, int trans_a, int trans_b,
                                   void *stream) {
  if ((!A || !B || !C) && M && N) return fail(FOT_ERR_INVALID_ARGUMENT, "null matrix");
  const uint32_t sam = trans_a ? 1 : K, sak = trans_a ? M : 1;
  const uint32_t sbk = trans_b ? 1 : N, sbn = trans_b ? K : 1;
  return gemm(A, B, C, nullptr, M, N, K, sam, sak, sbk, sbn, N, false, false,
              static_cast<hipStream_t>(stream));
}
extern "C" fot_status fot_gemm_f32(const float *A, const float *B, float *C, uint32_t M,
                                   uint32_t N, uint32_t K, int trans_a, int trans_b,
                                   void *stream) {
  return fot::guarded("fot_gemm_f32", [&] {
    return fot_gemm_f32_impl(A, B, C, M, N, K, trans_a, trans_b, stream);
  });
}
