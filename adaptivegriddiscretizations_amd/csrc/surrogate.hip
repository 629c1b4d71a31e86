// surrogate.hip -- kernels of the convolutional surrogate that warm-starts
// the eikonal iteration: 3x3 convolutions on BF16 activations and weights
// with FP32 accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16),
// their two backward passes, 2x2 pooling / upsampling, a masked MSE loss and
// an Adam step.  No counterpart in the reference (its solver has no learned
// component).
//
// Layout: activations NHWC, channels padded to a multiple of 32 (the MFMA
// reduction depth), W a multiple of 16; weights [tap][cout][cin] with
// tap = 3*(dy+1) + (dx+1).  A convolution is an implicit GEMM
//     out[p, co] = sum_{tap, ci} in[p + off(tap), ci] * w[tap, co, ci]
// with M = pixels, N = output channels, K = 9 * input channels.  One wave
// computes a 16-pixel x 32-channel block: per K step of 32 it loads one A
// fragment (16 pixels x 32 channels: 16 contiguous bytes per lane) and two B
// fragments, and issues two MFMAs.  Fragment layout of the 16x16x32 BF16
// MFMA: lane l holds A row l%16 and B column l%16 with reduction indices
// (l/16)*8 .. +7; the accumulator holds D rows (l/16)*4 .. +3 of column l%16.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"

namespace agd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; ++i) z[i] = __bf16(0.f);
  return z;
}

// Grid: one wave per (n, y, 16-pixel segment, 32-channel block); 4 waves
// per workgroup.
__global__ void __launch_bounds__(256) conv3x3_fwd(const __bf16* __restrict__ in,
                                                   const __bf16* __restrict__ w,
                                                   __bf16* __restrict__ out, int N, int H, int W,
                                                   int Cin, int Cout, int relu, int64_t nwaves) {
  const int64_t gw = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (gw >= nwaves) return;  // whole waves exit together
  const int l = threadIdx.x & 63;
  const int cblocks = Cout / 32, segs = W / 16;
  int64_t t = gw;
  const int cb = int(t % cblocks); t /= cblocks;
  const int sx = int(t % segs); t /= segs;
  const int y = int(t % H);
  const int n = int(t / H);
  const int x = sx * 16 + (l & 15);   // this lane's A row (pixel)
  const int kq = (l >> 4) * 8;        // this lane's reduction slice
  const int co0 = cb * 32 + (l & 15);
  f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 9; ++tap) {
    const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
    const bool inb = yy >= 0 && yy < H && xx >= 0 && xx < W;
    const __bf16* ip = in + ((int64_t(n) * H + (inb ? yy : 0)) * W + (inb ? xx : 0)) * Cin + kq;
    const __bf16* wp0 = w + (int64_t(tap) * Cout + co0) * Cin + kq;
    const __bf16* wp1 = wp0 + int64_t(16) * Cin;
    for (int c = 0; c < Cin; c += 32) {
      const bf16x8 a = inb ? *reinterpret_cast<const bf16x8*>(ip + c) : zero8();
      const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(wp0 + c);
      const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(wp1 + c);
      acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
    }
  }
  // D rows (pixels) (l/16)*4 + i, column (channel) l%16
  const int xr = sx * 16 + (l >> 4) * 4;
  __bf16* op = out + ((int64_t(n) * H + y) * W + xr) * Cout + co0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float v0 = acc0[i], v1 = acc1[i];
    if (relu) {
      v0 = v0 > 0.f ? v0 : 0.f;
      v1 = v1 > 0.f ? v1 : 0.f;
    }
    op[int64_t(i) * Cout] = __bf16(v0);
    op[int64_t(i) * Cout + 16] = __bf16(v1);
  }
}

// dW[tap, co, ci] += sum_p gout[p, co] * in[p + off(tap), ci], an implicit
// GEMM with M = output channels, N = input channels, K = pixels.  One
// workgroup (4 waves) per (pixel chunk, tap, 32-co block, 32-ci block); per K
// step of 32 pixels the workgroup stages gout[32 px][32 co] and the
// tap-shifted input [32 px][32 ci] in LDS with 16-byte loads (8 channels of
// one pixel per lane, the pixel's row and column from one 32-bit division),
// stored transposed so that each lane reads its MFMA fragment (8
// consecutive pixels of one channel) with one 16-byte LDS read; wave w
// accumulates the 16x16 block (co half w/2, ci half w%2) in registers and adds
// it to dw once per chunk (FP32 atomics).
constexpr int kWgChunk = 1024;  // pixels per workgroup
__global__ void __launch_bounds__(256) conv3x3_wgrad(const __bf16* __restrict__ in,
                                                     const __bf16* __restrict__ gout,
                                                     float* __restrict__ dw, int N, int H, int W,
                                                     int Cin, int Cout, int nchunks) {
  __shared__ __attribute__((aligned(16))) __bf16 sg[32][40];  // [co][pixel], padded rows
  __shared__ __attribute__((aligned(16))) __bf16 si[32][40];  // [ci][pixel]
  const int nci = Cin / 32, nco = Cout / 32;
  int t = blockIdx.x;
  const int bi = t % nci; t /= nci;
  const int bo = t % nco; t /= nco;
  const int tap = t % 9;
  const int chunk = t / 9;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int npix = N * H * W;
  const int p0 = chunk * kWgChunk;
  const int p1 = min(p0 + kWgChunk, npix);
  const int tid = threadIdx.x;
  const int wv = tid >> 6, l = tid & 63;
  // staging role: lanes 0..127 load gout, 128..255 the shifted input
  const bool ld_g = tid < 128;
  const int r = tid & 127;
  const int px_local = r >> 2, c8 = (r & 3) * 8;
  // MFMA role
  const int coh = wv >> 1, cih = wv & 1;
  const int kq = (l >> 4) * 8;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int p = p0; p < p1; p += 32) {
    const int q = p + px_local;
    bf16x8 v = zero8();
    if (q < p1) {
      if (ld_g) {
        v = *reinterpret_cast<const bf16x8*>(gout + int64_t(q) * Cout + bo * 32 + c8);
      } else {
        const int x = q % W;
        const int rowi = q / W;  // n * H + y
        const int y = rowi % H;
        const int ys = y + dy, xs = x + dx;
        if (ys >= 0 && ys < H && xs >= 0 && xs < W)
          v = *reinterpret_cast<const bf16x8*>(in + (int64_t(rowi + dy) * W + xs) * Cin + bi * 32 + c8);
      }
    }
    __bf16(*dst)[40] = ld_g ? sg : si;
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[c8 + i][px_local] = v[i];
    __syncthreads();
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sg[coh * 16 + (l & 15)][kq]);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&si[cih * 16 + (l & 15)][kq]);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    __syncthreads();
  }
  // D rows (co) (l/16)*4 + i of the wave's block, column (ci) l%16
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int co = bo * 32 + coh * 16 + (l >> 4) * 4 + i;
    const int ci = bi * 32 + cih * 16 + (l & 15);
    atomicAdd(dw + (int64_t(tap) * Cout + co) * Cin + ci, acc[i]);
  }
}

// Weights of the data gradient: wt[8-tap, ci, co] = w[tap, co, ci], so that
// gin = conv3x3_fwd(gout, wt).
__global__ void conv3x3_flip(const __bf16* __restrict__ w, __bf16* __restrict__ wt, int Cin,
                             int Cout) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = int64_t(9) * Cin * Cout;
  if (i >= n) return;
  const int ci = int(i % Cin);
  const int64_t r = i / Cin;
  const int co = int(r % Cout);
  const int tap = int(r / Cout);
  wt[(int64_t(8 - tap) * Cin + ci) * Cout + co] = w[i];
}

__global__ void relu_mask(__bf16* __restrict__ g, const __bf16* __restrict__ x, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && !(float(x[i]) > 0.f)) g[i] = __bf16(0.f);
}

// 2x2 average pooling (in: N, 2H, 2W, C -> out: N, H, W, C) and its adjoint.
__global__ void pool2_fwd(const __bf16* __restrict__ in, __bf16* __restrict__ out, int N, int H,
                          int W, int C) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(N) * H * W * C) return;
  const int c = int(i % C);
  int64_t r = i / C;
  const int x = int(r % W); r /= W;
  const int y = int(r % H);
  const int64_t n = r / H;
  const int64_t W2 = 2 * W;
  const int64_t b = ((n * 2 * H + 2 * y) * W2 + 2 * x) * C + c;
  const float s = float(in[b]) + float(in[b + C]) + float(in[b + W2 * C]) + float(in[b + W2 * C + C]);
  out[i] = __bf16(0.25f * s);
}

__global__ void pool2_bwd(const __bf16* __restrict__ gout, __bf16* __restrict__ gin, int N, int H,
                          int W, int C) {
  // gin: N, 2H, 2W, C
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(N) * 4 * H * W * C) return;
  const int c = int(i % C);
  int64_t r = i / C;
  const int x = int(r % (2 * W)); r /= 2 * W;
  const int y = int(r % (2 * H));
  const int64_t n = r / (2 * H);
  gin[i] = __bf16(0.25f * float(gout[((n * H + y / 2) * W + x / 2) * C + c]));
}

// Nearest 2x upsampling (in: N, H, W, C -> out: N, 2H, 2W, C) and its adjoint.
__global__ void up2_fwd(const __bf16* __restrict__ in, __bf16* __restrict__ out, int N, int H,
                        int W, int C) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(N) * 4 * H * W * C) return;
  const int c = int(i % C);
  int64_t r = i / C;
  const int x = int(r % (2 * W)); r /= 2 * W;
  const int y = int(r % (2 * H));
  const int64_t n = r / (2 * H);
  out[i] = in[((n * H + y / 2) * W + x / 2) * C + c];
}

__global__ void up2_bwd(const __bf16* __restrict__ gout, __bf16* __restrict__ gin, int N, int H,
                        int W, int C) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(N) * H * W * C) return;
  const int c = int(i % C);
  int64_t r = i / C;
  const int x = int(r % W); r /= W;
  const int y = int(r % H);
  const int64_t n = r / H;
  const int64_t W2 = 2 * W;
  const int64_t b = ((n * 2 * H + 2 * y) * W2 + 2 * x) * C + c;
  gin[i] = __bf16(float(gout[b]) + float(gout[b + C]) + float(gout[b + W2 * C]) +
                  float(gout[b + W2 * C + C]));
}

__global__ void add_bf16(const __bf16* __restrict__ a, const __bf16* __restrict__ b,
                         __bf16* __restrict__ out, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __bf16(float(a[i]) + float(b[i]));
}

// Loss sum_p m_p ((pred_p - t_p) / s)^2 / count on channel 0 of pred; writes
// the gradient (channel 0, other channels 0) and adds the loss to *loss.
__global__ void mse_loss(const __bf16* __restrict__ pred, const float* __restrict__ target,
                         const float* __restrict__ mask, __bf16* __restrict__ gpred, int64_t npix,
                         int C, float inv_scale2, float inv_count, float* __restrict__ loss) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float contrib = 0.f;
  if (p < npix) {
    const float m = mask ? mask[p] : 1.f;
    const float d = float(pred[p * C]) - target[p];
    contrib = m * d * d * inv_scale2 * inv_count;
    gpred[p * C] = __bf16(2.f * m * d * inv_scale2 * inv_count);
    for (int c = 1; c < C; ++c) gpred[p * C + c] = __bf16(0.f);
  }
  // wave reduction, one atomic per wave
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, contrib);
}

// Adam step on FP32 master weights; refreshes the BF16 copy.
__global__ void adam_step(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, __bf16* __restrict__ wb, int64_t n, float lr,
                          float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = __bf16(wi);
}

// Warm start: values[y, x] of the fine grid (H x W) = inflate * scale[y, x] *
// (bilinear, cell-centred upsampling of channel 0 of the coarse H/f x W/f
// prediction, clamped below at 0); scale may be null (1).
__global__ void warm_values(const __bf16* __restrict__ pred, int Hc, int Wc, int C, int f,
                            float inflate, const float* __restrict__ scale,
                            float* __restrict__ out, int H, int W) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(H) * W) return;
  const int x = int(i % W), y = int(i / W);
  const float fy = (y + 0.5f) / f - 0.5f, fx = (x + 0.5f) / f - 0.5f;
  int y0 = int(floorf(fy)), x0 = int(floorf(fx));
  const float ty = fy - y0, tx = fx - x0;
  auto at = [&](int yy, int xx) {
    yy = yy < 0 ? 0 : (yy >= Hc ? Hc - 1 : yy);
    xx = xx < 0 ? 0 : (xx >= Wc ? Wc - 1 : xx);
    return float(pred[(int64_t(yy) * Wc + xx) * C]);
  };
  const float v = (1 - ty) * ((1 - tx) * at(y0, x0) + tx * at(y0, x0 + 1)) +
                  ty * ((1 - tx) * at(y0 + 1, x0) + tx * at(y0 + 1, x0 + 1));
  out[i] = inflate * (v > 0.f ? v : 0.f) * (scale ? scale[i] : 1.f);
}

}  // namespace agd

using namespace agd;

namespace {
inline dim3 g1(int64_t n, int b = 256) { return dim3(unsigned((n + b - 1) / b)); }
}

extern "C" {

int agd_nn_conv3x3(const void* in, const void* w, void* out, int32_t N, int32_t H, int32_t W,
                   int32_t Cin, int32_t Cout, int32_t relu, void* stream) {
  AGD_REQUIRE(in && w && out && N >= 1 && H >= 1 && W >= 16, "bad convolution arguments");
  AGD_REQUIRE(W % 16 == 0 && Cin % 32 == 0 && Cout % 32 == 0,
              "convolution needs W % 16 == 0 and channels padded to multiples of 32");
  const int64_t nwaves = int64_t(N) * H * (W / 16) * (Cout / 32);
  hipLaunchKernelGGL(conv3x3_fwd, dim3(unsigned((nwaves + 3) / 4)), dim3(256), 0,
                     as_stream(stream), static_cast<const __bf16*>(in),
                     static_cast<const __bf16*>(w), static_cast<__bf16*>(out), N, H, W, Cin, Cout,
                     relu, nwaves);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_conv3x3_wgrad(const void* in, const void* gout, float* dw, int32_t N, int32_t H,
                         int32_t W, int32_t Cin, int32_t Cout, void* stream) {
  AGD_REQUIRE(in && gout && dw && N >= 1 && H >= 1 && W >= 1, "bad convolution arguments");
  AGD_REQUIRE(Cin % 32 == 0 && Cout % 32 == 0, "channels must be padded to multiples of 32");
  const int64_t npix = int64_t(N) * H * W;
  AGD_REQUIRE(npix < (int64_t(1) << 31), "too many pixels for the weight-gradient kernel");
  const int64_t nchunks = (npix + kWgChunk - 1) / kWgChunk;
  const int64_t nblocks = nchunks * 9 * (Cout / 32) * (Cin / 32);
  AGD_REQUIRE(nblocks < (int64_t(1) << 31), "too many weight-gradient blocks");
  hipLaunchKernelGGL(conv3x3_wgrad, dim3(unsigned(nblocks)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(in), static_cast<const __bf16*>(gout), dw, N, H, W,
                     Cin, Cout, int(nchunks));
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_conv3x3_flip(const void* w, void* wt, int32_t Cin, int32_t Cout, void* stream) {
  AGD_REQUIRE(w && wt && Cin >= 1 && Cout >= 1, "bad weights");
  const int64_t n = int64_t(9) * Cin * Cout;
  hipLaunchKernelGGL(conv3x3_flip, g1(n), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(w), static_cast<__bf16*>(wt), Cin, Cout);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_relu_mask(void* g, const void* x, int64_t n, void* stream) {
  AGD_REQUIRE(g && x && n >= 0, "bad buffers");
  if (n == 0) return AGD_OK;
  hipLaunchKernelGGL(relu_mask, g1(n), dim3(256), 0, as_stream(stream), static_cast<__bf16*>(g),
                     static_cast<const __bf16*>(x), n);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_pool2(const void* in, void* out, int32_t N, int32_t H, int32_t W, int32_t C,
                 int32_t backward, void* stream) {
  // forward: in (N, 2H, 2W, C) -> out (N, H, W, C); backward: gradient of out -> of in
  AGD_REQUIRE(in && out && N >= 1 && H >= 1 && W >= 1 && C >= 1, "bad pooling arguments");
  const int64_t n = int64_t(N) * H * W * C;
  if (backward)
    hipLaunchKernelGGL(pool2_bwd, g1(4 * n), dim3(256), 0, as_stream(stream),
                       static_cast<const __bf16*>(in), static_cast<__bf16*>(out), N, H, W, C);
  else
    hipLaunchKernelGGL(pool2_fwd, g1(n), dim3(256), 0, as_stream(stream),
                       static_cast<const __bf16*>(in), static_cast<__bf16*>(out), N, H, W, C);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_up2(const void* in, void* out, int32_t N, int32_t H, int32_t W, int32_t C,
               int32_t backward, void* stream) {
  // forward: in (N, H, W, C) -> out (N, 2H, 2W, C); backward: gradient of out -> of in
  AGD_REQUIRE(in && out && N >= 1 && H >= 1 && W >= 1 && C >= 1, "bad upsampling arguments");
  const int64_t n = int64_t(N) * H * W * C;
  if (backward)
    hipLaunchKernelGGL(up2_bwd, g1(n), dim3(256), 0, as_stream(stream),
                       static_cast<const __bf16*>(in), static_cast<__bf16*>(out), N, H, W, C);
  else
    hipLaunchKernelGGL(up2_fwd, g1(4 * n), dim3(256), 0, as_stream(stream),
                       static_cast<const __bf16*>(in), static_cast<__bf16*>(out), N, H, W, C);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_add(const void* a, const void* b, void* out, int64_t n, void* stream) {
  AGD_REQUIRE(a && b && out && n >= 0, "bad buffers");
  if (n == 0) return AGD_OK;
  hipLaunchKernelGGL(add_bf16, g1(n), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(a), static_cast<const __bf16*>(b),
                     static_cast<__bf16*>(out), n);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_mse(const void* pred, const float* target, const float* mask, void* gpred, int64_t npix,
               int32_t C, double scale, double count, float* loss, void* stream) {
  AGD_REQUIRE(pred && target && gpred && loss && npix >= 1 && C >= 1 && scale > 0 && count > 0,
              "bad loss arguments");
  hipLaunchKernelGGL(mse_loss, g1(npix), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(pred), target, mask, static_cast<__bf16*>(gpred),
                     npix, C, float(1.0 / (scale * scale)), float(1.0 / count), loss);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_adam(float* w, const float* g, float* m, float* v, void* wb, int64_t n, double lr,
                double b1, double b2, double eps, int32_t step, void* stream) {
  AGD_REQUIRE(w && g && m && v && wb && n >= 0 && step >= 1, "bad optimizer arguments");
  if (n == 0) return AGD_OK;
  const double c1 = 1.0 / (1.0 - std::pow(b1, step)), c2 = 1.0 / (1.0 - std::pow(b2, step));
  hipLaunchKernelGGL(adam_step, g1(n), dim3(256), 0, as_stream(stream), w, g, m, v,
                     static_cast<__bf16*>(wb), n, float(lr), float(b1), float(b2), float(eps),
                     float(c1), float(c2));
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_nn_warm_values(const void* pred, int32_t Hc, int32_t Wc, int32_t C, int32_t factor,
                       double inflate, const float* scale, float* out, int32_t H, int32_t W,
                       void* stream) {
  AGD_REQUIRE(pred && out && Hc >= 1 && Wc >= 1 && C >= 1 && factor >= 1 && H >= 1 && W >= 1,
              "bad warm-start arguments");
  hipLaunchKernelGGL(warm_values, g1(int64_t(H) * W), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(pred), Hc, Wc, C, factor, float(inflate), scale, out,
                     H, W);
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

}  // extern "C"
