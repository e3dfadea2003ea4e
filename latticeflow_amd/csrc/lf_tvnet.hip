// Learned predictor of TV values for the guided divide and conquer: a 3-layer
// 3x3 convolutional net with 16 channels, BF16 operands on the matrix cores
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation), forward and training kernels.
//
// No counterpart in the reference.  The net's output only chooses the initial
// partition and thresholds of the parametric divide and conquer; every result
// is still decided by exact min-cuts (lf_tv_path_guided).
//
// Net: input channels (f_norm, lambda feature), conv3x3 -> 16 + ReLU,
// conv3x3 16 -> 16 + ReLU, conv3x3 16 -> 1, output = f_norm + last conv.
// Zero padding at the image border.  Every layer runs as a 16 x 16 x 16 MFMA
// per tap (channels padded to 16; padded weights stay zero).
//
// Layouts: activations NHWC bf16 with 16 channels (32 bytes per pixel);
// parameters fp32 [W1 | b1 | W2 | b2 | W3 | b3], W[oc][tap][ic] with
// tap = (dy + 1) * 3 + (dx + 1), 16 x 9 x 16 each, biases 16 each.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "../../include/latticeflow_amd.h"
#include "lf_engine.hpp"

namespace lfamd {
namespace {

constexpr int C = 16;
constexpr int TAPS = 9;
constexpr int WSZ = C * TAPS * C;  // one layer's weights
constexpr int LAYER = WSZ + C;     // weights + bias
constexpr int NPARAM = 3 * LAYER;
constexpr int CT = 256;            // threads per block (4 waves)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float x) {
  // round to nearest even (finite values)
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 4 consecutive channels (c0 .. c0+3) of pixel (b, y, x), zero outside the image
__device__ inline bf16x4 load4(const unsigned short* __restrict__ a, int b, int y, int x, int H, int W, int c0) {
  bf16x4 r = {0, 0, 0, 0};
  if (y < 0 || y >= H || x < 0 || x >= W) return r;
  const uint2 u = *reinterpret_cast<const uint2*>(a + ((((long long)b * H + y) * W + x) * C + c0));
  r[0] = (short)(u.x & 0xffffu);
  r[1] = (short)(u.x >> 16);
  r[2] = (short)(u.y & 0xffffu);
  r[3] = (short)(u.y >> 16);
  return r;
}

// ---- parameters -> bf16 copy ---------------------------------------------------
__global__ __launch_bounds__(CT) void k_tvnet_wq(const float* __restrict__ p, unsigned short* __restrict__ q) {
  const int i = blockIdx.x * CT + threadIdx.x;
  if (i < NPARAM) q[i] = f2bf(p[i]);
}

// ---- input: channel 0 = f_norm, channel 1 = lambda feature, others 0 --------
__global__ __launch_bounds__(CT) void k_tvnet_input(const float* __restrict__ f, const float* __restrict__ lam,
                                                    long long hw, int n, unsigned short* __restrict__ x) {
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n * hw) return;
  const int b = (int)(i / hw);
  uint4 z0 = {0, 0, 0, 0}, z1 = {0, 0, 0, 0};
  z0.x = (unsigned)f2bf(f[i]) | ((unsigned)f2bf(lam[b]) << 16);
  uint4* o = reinterpret_cast<uint4*>(x + i * C);
  o[0] = z0;
  o[1] = z1;
}

// ---- forward conv: out = act(W * in + b) ------------------------------------------
// MFMA per tap: A = W[oc][ic] (rows oc), B = in[ic][pixel] (16 pixels of a
// row), D = out[oc][pixel]: lane l holds pixel l % 16, channels 4 (l / 16) + i.
// OUT_F32: only channel 0 is kept, as f32, plus the residual input.
template <bool RELU, bool OUT_F32>
__global__ __launch_bounds__(CT) void k_tvnet_fwd(const unsigned short* __restrict__ in, int n, int H, int W,
                                                  const unsigned short* __restrict__ wq, const float* __restrict__ bias,
                                                  unsigned short* __restrict__ out, float* __restrict__ out32,
                                                  const float* __restrict__ resid) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xblocks = (W + 63) / 64;
  const int row = blockIdx.x / xblocks;  // b * H + y
  if (row >= n * H) return;
  const int b = row / H, y = row % H;
  const int x0 = (blockIdx.x % xblocks) * 64 + wave * 16;
  const int m = lane & 15, g = lane >> 4;
  const int px = x0 + m;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < TAPS; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = (short)wq[(m * TAPS + t) * C + 4 * g + j];
    const bf16x4 bv = load4(in, b, y + dy, px + dx, H, W, 4 * g);
    acc = mfma16(a, bv, acc);
  }
  if (px >= W) return;
  const long long pix = ((long long)b * H + y) * W + px;
  if (OUT_F32) {
    if (g == 0) out32[pix] = acc[0] + bias[0] + resid[pix];
    return;
  }
  unsigned short o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v = acc[i] + bias[4 * g + i];
    if (RELU) v = v > 0.f ? v : 0.f;
    o[i] = f2bf(v);
  }
  uint2 u;
  u.x = (unsigned)o[0] | ((unsigned)o[1] << 16);
  u.y = (unsigned)o[2] | ((unsigned)o[3] << 16);
  *reinterpret_cast<uint2*>(out + pix * C + 4 * g) = u;
}

// ---- loss: mean squared error; gradient of the output layer (channel 0) ----------
__global__ __launch_bounds__(CT) void k_tvnet_loss(const float* __restrict__ pred, const float* __restrict__ target,
                                                   long long P, unsigned short* __restrict__ g3,
                                                   float* __restrict__ loss) {
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  float l = 0.f;
  if (i < P) {
    const float d = pred[i] - target[i];
    l = d * d;
    uint4 z0 = {0, 0, 0, 0}, z1 = {0, 0, 0, 0};
    z0.x = (unsigned)f2bf(2.f * d / (float)P);
    uint4* o = reinterpret_cast<uint4*>(g3 + i * C);
    o[0] = z0;
    o[1] = z1;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_down(l, o, 64);
  __shared__ float s[CT / 64];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (s[0] + s[1] + s[2] + s[3]) / (float)P);
}

// ---- input gradient: gin = (W^T * g) masked by the ReLU of the input layer ---
// gin[p][ic] = sum_t sum_oc W[oc][t][ic] g[p - d_t][oc]; A = W^T (rows ic),
// B = g[oc][pixel] at the mirrored tap, D = gin[ic][pixel].  The result is
// zeroed where the layer's input activation (the previous ReLU output) is 0.
__global__ __launch_bounds__(CT) void k_tvnet_dx(const unsigned short* __restrict__ g, int n, int H, int W,
                                                 const unsigned short* __restrict__ wq,
                                                 const unsigned short* __restrict__ act_in,
                                                 unsigned short* __restrict__ gin) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int xblocks = (W + 63) / 64;
  const int row = blockIdx.x / xblocks;
  if (row >= n * H) return;
  const int b = row / H, y = row % H;
  const int x0 = (blockIdx.x % xblocks) * 64 + wave * 16;
  const int m = lane & 15, gg = lane >> 4;
  const int px = x0 + m;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < TAPS; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = (short)wq[((4 * gg + j) * TAPS + t) * C + m];
    const bf16x4 bv = load4(g, b, y - dy, px - dx, H, W, 4 * gg);
    acc = mfma16(a, bv, acc);
  }
  if (px >= W) return;
  const long long pix = ((long long)b * H + y) * W + px;
  const uint2 ai = *reinterpret_cast<const uint2*>(act_in + pix * C + 4 * gg);
  const unsigned short av[4] = {(unsigned short)(ai.x & 0xffffu), (unsigned short)(ai.x >> 16),
                                (unsigned short)(ai.y & 0xffffu), (unsigned short)(ai.y >> 16)};
  unsigned short o[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = bf2f(av[i]) > 0.f ? f2bf(acc[i]) : (unsigned short)0;
  uint2 u;
  u.x = (unsigned)o[0] | ((unsigned)o[1] << 16);
  u.y = (unsigned)o[2] | ((unsigned)o[3] << 16);
  *reinterpret_cast<uint2*>(gin + pix * C + 4 * gg) = u;
}

// ---- weight gradient: dW[oc][t][ic] = sum_p g[p][oc] x[p + d_t][ic] -----------
// MFMA per tap over strips of 16 pixels: A = g[oc][pixel] (rows oc, K = pixel),
// B = x[pixel + d_t][ic].  Each wave stages its strip of g and the 3 x 18
// pixel halo of x in LDS, accumulates 9 tiles over its strips, and the block
// adds its sums to dW (and the bias gradient, the sum of g) with atomics.
__global__ __launch_bounds__(CT) void k_tvnet_dw(const unsigned short* __restrict__ g,
                                                 const unsigned short* __restrict__ x, int n, int H, int W,
                                                 float* __restrict__ dw, float* __restrict__ db) {
  __shared__ unsigned short sg[4][16 * C];
  __shared__ unsigned short sx[4][3 * 18 * C];
  __shared__ float red[TAPS * C * C];
  __shared__ float redb[C];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, gq = lane >> 4;
  for (int i = threadIdx.x; i < TAPS * C * C; i += CT) red[i] = 0.f;
  if (threadIdx.x < C) redb[threadIdx.x] = 0.f;
  f32x4 acc[TAPS];
#pragma unroll
  for (int t = 0; t < TAPS; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const int xs = (W + 15) / 16;
  const long long strips = (long long)n * H * xs;
  for (long long s = (long long)blockIdx.x * 4 + wave; s < strips; s += (long long)gridDim.x * 4) {
    const int b = (int)(s / ((long long)H * xs));
    const int rem = (int)(s % ((long long)H * xs));
    const int y = rem / xs, x0 = (rem % xs) * 16;
    // g strip: pixel m, channels 4 gq .. +3
    {
      const bf16x4 v = load4(g, b, y, x0 + m, H, W, 4 * gq);
#pragma unroll
      for (int j = 0; j < 4; ++j) sg[wave][m * C + 4 * gq + j] = (unsigned short)v[j];
    }
    // x halo: rows y-1 .. y+1, pixels x0-1 .. x0+16
    for (int e = lane; e < 3 * 18 * 4; e += 64) {
      const int q = e >> 2, c4 = e & 3;
      const int r = q / 18, c = q % 18;
      const bf16x4 v = load4(x, b, y + r - 1, x0 + c - 1, H, W, 4 * c4);
#pragma unroll
      for (int j = 0; j < 4; ++j) sx[wave][q * C + 4 * c4 + j] = (unsigned short)v[j];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // A[oc = m][k = pixel 4 gq + j] = g[pixel][oc]
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a[j] = (short)sg[wave][(4 * gq + j) * C + m];
      bsum += bf2f((unsigned short)a[j]);
    }
#pragma unroll
    for (int t = 0; t < TAPS; ++t) {
      const int dy = t / 3 - 1, dx = t % 3 - 1;
      bf16x4 bv;
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = (short)sx[wave][((dy + 1) * 18 + (4 * gq + j) + dx + 1) * C + m];
      acc[t] = mfma16(a, bv, acc[t]);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  __syncthreads();
  // D[oc = 4 gq + i][ic = m]
#pragma unroll
  for (int t = 0; t < TAPS; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) atomicAdd(&red[((4 * gq + i) * TAPS + t) * C + m], acc[t][i]);
  atomicAdd(&redb[m], bsum);
  __syncthreads();
  for (int i = threadIdx.x; i < TAPS * C * C; i += CT)
    if (red[i] != 0.f) atomicAdd(&dw[i], red[i]);
  if (threadIdx.x < C && redb[threadIdx.x] != 0.f) atomicAdd(&db[threadIdx.x], redb[threadIdx.x]);
}

// ---- Adam (fp32 master parameters) ----------------------------------------------
__global__ __launch_bounds__(CT) void k_tvnet_adam(float* __restrict__ p, const float* __restrict__ grad,
                                                   float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                   float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * CT + threadIdx.x;
  if (i >= NPARAM) return;
  const float gr = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gr;
  const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <typename V>
__global__ __launch_bounds__(CT) void k_tvnet_zero(V* a, long long n) {
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i < n) a[i] = (V)0;
}

inline unsigned nblk(long long n) { return (unsigned)((n + CT - 1) / CT); }

// workspace: wq (bf16 params) | x1 | y1 | y2 | ga | gb (bf16 16 ch) | pred (f32)
struct Work {
  unsigned short *wq, *x1, *y1, *y2, *ga, *gb;
  float* pred;
};

size_t work_bytes(long long P) {
  const size_t act = (size_t)P * C * sizeof(unsigned short);
  return 256 + ((NPARAM * 2 + 255) / 256) * 256 + 5 * ((act + 255) / 256) * 256 + ((P * 4 + 255) / 256) * 256;
}

Work carve(void* base, long long P) {
  char* c = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(base) + 255) / 256 * 256);
  auto take = [&](size_t bytes) {
    char* r = c;
    c += (bytes + 255) / 256 * 256;
    return r;
  };
  Work w;
  const size_t act = (size_t)P * C * sizeof(unsigned short);
  w.wq = reinterpret_cast<unsigned short*>(take(NPARAM * 2));
  w.x1 = reinterpret_cast<unsigned short*>(take(act));
  w.y1 = reinterpret_cast<unsigned short*>(take(act));
  w.y2 = reinterpret_cast<unsigned short*>(take(act));
  w.ga = reinterpret_cast<unsigned short*>(take(act));
  w.gb = reinterpret_cast<unsigned short*>(take(act));
  w.pred = reinterpret_cast<float*>(take((size_t)P * 4));
  return w;
}

void check_shape(int n, int64_t H, int64_t W) {
  if (n <= 0 || H <= 0 || W <= 0) fail(LF_ERR_BAD_ARG, "empty batch");
  if (H > (1 << 20) || W > (1 << 20) || (long long)n * H > (1LL << 30)) fail(LF_ERR_BAD_ARG, "batch too large");
}

// Forward pass into `out` (P floats); uses w.wq, w.x1, w.y1, w.y2.
void forward_to(const float* params, int n, int H, int W, const float* f, const float* lam, const Work& w,
                float* out, hipStream_t s) {
  const long long P = (long long)n * H * W;
  const unsigned rows = (unsigned)((long long)n * H * ((W + 63) / 64));
  hipLaunchKernelGGL(k_tvnet_wq, dim3(nblk(NPARAM)), dim3(CT), 0, s, params, w.wq);
  hipLaunchKernelGGL(k_tvnet_input, dim3(nblk(P)), dim3(CT), 0, s, f, lam, (long long)H * W, n, w.x1);
  hipLaunchKernelGGL((k_tvnet_fwd<true, false>), dim3(rows), dim3(CT), 0, s, w.x1, n, H, W, w.wq, params + WSZ, w.y1,
                     (float*)nullptr, (const float*)nullptr);
  hipLaunchKernelGGL((k_tvnet_fwd<true, false>), dim3(rows), dim3(CT), 0, s, w.y1, n, H, W, w.wq + LAYER,
                     params + LAYER + WSZ, w.y2, (float*)nullptr, (const float*)nullptr);
  hipLaunchKernelGGL((k_tvnet_fwd<false, true>), dim3(rows), dim3(CT), 0, s, w.y2, n, H, W, w.wq + 2 * LAYER,
                     params + 2 * LAYER + WSZ, (unsigned short*)nullptr, out, f);
}

void forward(const float* params, int n, int H, int W, const float* f, const float* lam, Work w, hipStream_t s) {
  forward_to(params, n, H, W, f, lam, w, w.pred, s);
}

// ---- forward pass without allocation ---------------------------------------------
// Scratch of tvnet_forward_noalloc: the bf16 weights and the three activation
// planes of one call live in this file-scope device array; every GPU the
// library runs on holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel: it takes the device address of the
// current GPU's copy from hipGetSymbolAddress (once per device) and uses that.
// Calls that overlap on one GPU share the array and race.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static_assert((size_t)LF_TVNET_NOALLOC_MAX_PIXELS * 3 * C * sizeof(unsigned short) +
                      ((NPARAM * 2 + 255) / 256) * 256 + 3 * 256 <= kNoAllocBytes,
              "the stated pixel limit must fit the scratch array");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current GPU's scratch (looked up on the first call on
// each device, with that device current).
void* scratch_address() {
  int dev = 0;
  LF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(LF_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      fail(LF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

void gradients(const float* params, int n, int H, int W, const float* f, const float* lam, const float* target,
               float* grad, float* loss, Work w, hipStream_t s) {
  const long long P = (long long)n * H * W;
  const unsigned rows = (unsigned)((long long)n * H * ((W + 63) / 64));
  const unsigned dwb = (unsigned)std::min<long long>(1024, std::max<long long>(1, (P / 16 + 63) / 64));
  forward(params, n, H, W, f, lam, w, s);
  hipLaunchKernelGGL(k_tvnet_zero<float>, dim3(nblk(NPARAM)), dim3(CT), 0, s, grad, (long long)NPARAM);
  hipLaunchKernelGGL(k_tvnet_zero<float>, dim3(1), dim3(CT), 0, s, loss, 1LL);
  hipLaunchKernelGGL(k_tvnet_loss, dim3(nblk(P)), dim3(CT), 0, s, w.pred, target, P, w.ga, loss);
  // layer 3: g = ga (channel 0), input y2
  hipLaunchKernelGGL(k_tvnet_dw, dim3(dwb), dim3(CT), 0, s, w.ga, w.y2, n, H, W, grad + 2 * LAYER,
                     grad + 2 * LAYER + WSZ);
  hipLaunchKernelGGL(k_tvnet_dx, dim3(rows), dim3(CT), 0, s, w.ga, n, H, W, w.wq + 2 * LAYER, w.y2, w.gb);
  // layer 2: g = gb, input y1
  hipLaunchKernelGGL(k_tvnet_dw, dim3(dwb), dim3(CT), 0, s, w.gb, w.y1, n, H, W, grad + LAYER, grad + LAYER + WSZ);
  hipLaunchKernelGGL(k_tvnet_dx, dim3(rows), dim3(CT), 0, s, w.gb, n, H, W, w.wq + LAYER, w.y1, w.ga);
  // layer 1: g = ga, input x1
  hipLaunchKernelGGL(k_tvnet_dw, dim3(dwb), dim3(CT), 0, s, w.ga, w.x1, n, H, W, grad, grad + WSZ);
}

}  // namespace

int tvnet_param_count() { return NPARAM; }

size_t tvnet_work_bytes(int n, int64_t H, int64_t W) {
  if (n <= 0 || H <= 0 || W <= 0) return 0;
  return work_bytes((long long)n * H * W);
}

static long long checked_pixels(int n, int64_t H, int64_t W, size_t work_size) {
  check_shape(n, H, W);
  const long long P = (long long)n * H * W;
  if (work_size < work_bytes(P)) fail(LF_ERR_BAD_ARG, "workspace too small (lf_tvnet_work_bytes)");
  return P;
}

void tvnet_forward(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam, float* pred,
                   void* work, size_t work_size, hipStream_t s) {
  const long long P = checked_pixels(n, H, W, work_size);
  Work w = carve(work, P);
  forward(params, n, (int)H, (int)W, f, lam, w, s);
  LF_HIP(hipMemcpyAsync(pred, w.pred, sizeof(float) * P, hipMemcpyDeviceToDevice, s));
  LF_HIP(hipGetLastError());
}

void tvnet_forward_noalloc(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                           float* pred, hipStream_t s) {
  check_shape(n, H, W);
  const long long P = (long long)n * H * W;
  if (P > LF_TVNET_NOALLOC_MAX_PIXELS)
    fail(LF_ERR_BAD_ARG, "n * H * W = " + std::to_string(P) + " pixels exceeds the allocation-free limit of " +
                             std::to_string(LF_TVNET_NOALLOC_MAX_PIXELS) + " (LF_TVNET_NOALLOC_MAX_PIXELS)");
  char* c = reinterpret_cast<char*>(scratch_address());
  const size_t act = ((size_t)P * C * sizeof(unsigned short) + 255) / 256 * 256;
  Work w{};
  w.wq = reinterpret_cast<unsigned short*>(c);
  c += ((NPARAM * 2 + 255) / 256) * 256;
  w.x1 = reinterpret_cast<unsigned short*>(c);
  w.y1 = reinterpret_cast<unsigned short*>(c + act);
  w.y2 = reinterpret_cast<unsigned short*>(c + 2 * act);
  forward_to(params, n, (int)H, (int)W, f, lam, w, pred, s);
  LF_HIP(hipGetLastError());
}

void tvnet_gradients(const float* params, int n, int64_t H, int64_t W, const float* f, const float* lam,
                     const float* target, float* grad, float* loss, void* work, size_t work_size, hipStream_t s) {
  const long long P = checked_pixels(n, H, W, work_size);
  gradients(params, n, (int)H, (int)W, f, lam, target, grad, loss, carve(work, P), s);
  LF_HIP(hipGetLastError());
}

void tvnet_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t s) {
  if (step < 1) fail(LF_ERR_BAD_ARG, "step counts from 1");
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_tvnet_adam, dim3(nblk(NPARAM)), dim3(CT), 0, s, params, grad, m, v, lr, b1, b2, c1, c2, 1e-8f);
  LF_HIP(hipGetLastError());
}

}  // namespace lfamd
