// layout.hip -- layout conversion (cell-major <-> tiled), geometry
// preparation (metric -> flattened dual metric in the tiled layout), memory
// helpers and error reporting of the C ABI.
//
// Reference behaviour restated here:
//   block_expand / block_squeeze   FiniteDifferences.py:164-222
//   Riemann: D = ((h m h))^-1, lower-triangular flattening
//            _SetGeometry.py:127-137, riemann.py:38-40,68-73, misc.py:8-15
//   Rander:  [flat(inv(h m h)), h w]  _SetGeometry.py:139, rander.py:61-68
//   Diagonal: (costs h)^-2           _SetGeometry.py:134-135, diagonal.py:32-35
//   TTI:     [l, flat(q), (A h)^-T]  _SetGeometry.py:140-142, Seismic/tti.py:113-125
//   Padding with +inf                _SetGeometry.py:189-190
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "device.hpp"

namespace agd {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

struct LayoutDesc {
  int32_t ndim;
  int32_t shape[3];
  int32_t shape_i[3];
  int32_t shape_o[3];
  int32_t size_i;
  int64_t size_tot;
  int64_t ncell;
};

static int make_layout(int32_t ndim, const int64_t* shape, const int32_t* shape_i, LayoutDesc& L) {
  AGD_REQUIRE(ndim >= 1 && ndim <= 3, "ndim must be 1..3");
  AGD_REQUIRE(shape && shape_i, "null shape");
  L.ndim = ndim;
  L.size_i = 1;
  L.ncell = 1;
  int64_t ntiles = 1;
  for (int k = 0; k < 3; ++k) {
    if (k < ndim) {
      AGD_REQUIRE(shape[k] >= 1 && shape[k] < (1 << 30), "shape out of range");
      AGD_REQUIRE(shape_i[k] >= 1 && shape_i[k] <= 1024, "tile shape out of range");
      L.shape[k] = int32_t(shape[k]);
      L.shape_i[k] = shape_i[k];
      L.shape_o[k] = int32_t(ceil_div(shape[k], shape_i[k]));
    } else {
      L.shape[k] = L.shape_i[k] = L.shape_o[k] = 1;
    }
    L.size_i *= L.shape_i[k];
    L.ncell *= L.shape[k];
    ntiles *= L.shape_o[k];
  }
  L.size_tot = ntiles * L.size_i;
  return AGD_OK;
}

// Tiled index -> cell position (returns false in the padding).
__device__ inline bool tiled_to_cell(const LayoutDesc& L, int64_t n_t, int64_t& cell) {
  int64_t n_o = n_t / L.size_i;
  int n_i = int(n_t - n_o * L.size_i);
  int xo[3], xi[3];
  for (int k = 2; k >= 0; --k) {
    xo[k] = int(n_o % L.shape_o[k]);
    n_o /= L.shape_o[k];
    xi[k] = n_i % L.shape_i[k];
    n_i /= L.shape_i[k];
  }
  cell = 0;
  for (int k = 0; k < 3; ++k) {
    const int x = xo[k] * L.shape_i[k] + xi[k];
    if (x >= L.shape[k]) return false;
    cell = cell * L.shape[k] + x;
  }
  return true;
}

template <class T>
__global__ void expand_kernel(LayoutDesc L, int32_t ncomp, const T* src, int64_t stride, T* dst,
                              T pad) {
  const int64_t n_t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n_t >= L.size_tot) return;
  int64_t cell;
  const bool in = tiled_to_cell(L, n_t, cell);
  for (int c = 0; c < ncomp; ++c) {
    T v = pad;
    if (in) v = stride == 0 ? src[c] : src[c * L.ncell + cell];
    dst[c * L.size_tot + n_t] = v;
  }
}

template <class T>
__global__ void squeeze_kernel(LayoutDesc L, int32_t ncomp, const T* src, T* dst) {
  const int64_t n_t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n_t >= L.size_tot) return;
  int64_t cell;
  if (!tiled_to_cell(L, n_t, cell)) return;
  for (int c = 0; c < ncomp; ++c) dst[c * L.ncell + cell] = src[c * L.size_tot + n_t];
}

// Symmetric 2x2 / 3x3 inverse in double precision (cofactors).
__device__ inline void inv_sym(int d, const double (&a)[3][3], double (&r)[3][3]) {
  if (d == 2) {
    const double det = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    r[0][0] = a[1][1] / det;
    r[1][1] = a[0][0] / det;
    r[0][1] = r[1][0] = -a[0][1] / det;
  } else {
    const double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    const double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    const double c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    const double c12 = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    const double c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    r[0][0] = c00 / det;
    r[0][1] = r[1][0] = c01 / det;
    r[0][2] = r[2][0] = c02 / det;
    r[1][1] = c11 / det;
    r[1][2] = r[2][1] = c12 / det;
    r[2][2] = c22 / det;
  }
}

// General 2x2 / 3x3 inverse in double precision (adjugate / determinant).
__device__ inline void inv_gen(int d, const double (&a)[3][3], double (&r)[3][3]) {
  if (d == 2) {
    const double det = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    r[0][0] = a[1][1] / det;
    r[0][1] = -a[0][1] / det;
    r[1][0] = -a[1][0] / det;
    r[1][1] = a[0][0] / det;
    return;
  }
  double c[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      c[i][j] = a[i1][j1] * a[i2][j2] - a[i1][j2] * a[i2][j1];  // cofactor (i, j)
    }
  const double det = a[0][0] * c[0][0] + a[0][1] * c[0][1] + a[0][2] * c[0][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = c[j][i] / det;
}

template <class S>
__global__ void geometry_kernel(LayoutDesc L, int32_t model, double h0, double h1, double h2,
                                const float* m, const float* w, int64_t stride, S* geom) {
  const int64_t n_t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n_t >= L.size_tot) return;
  const int d = L.ndim;
  const double h[3] = {h0, h1, h2};
  const int sd = d * (d + 1) / 2;
  int64_t cell;
  const bool in = tiled_to_cell(L, n_t, cell);
  const int64_t c = stride == 0 ? 0 : cell;
  const int64_t cs = stride == 0 ? 1 : L.ncell;  // component stride of the source
  if (model == AGD_DIAGONAL) {
    for (int k = 0; k < d; ++k) {
      double v = inf_v<double>();
      if (in) {
        const double ch = double(m[k * cs + c]) * h[k];
        v = 1.0 / (ch * ch);
      }
      geom[k * L.size_tot + n_t] = S(v);
    }
    return;
  }
  if (model == AGD_TTI) {
    // TTI.with_costs(h).flatten(transposed_transformation=True)
    // (Seismic/tti.py:113-125, Metrics/base.py:175-185): [l, flat(q),
    // (A diag(h))^{-T}], A = w (row-major, identity if null).
    const int ng = 5 + d * d;
    if (!in) {
      for (int k = 0; k < ng; ++k) geom[k * L.size_tot + n_t] = inf_v<S>();
      return;
    }
    for (int k = 0; k < 5; ++k) geom[k * L.size_tot + n_t] = S(double(m[k * cs + c]));
    double A[3][3], r[3][3];
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j)
        A[i][j] = (w ? double(w[(i * d + j) * cs + c]) : double(i == j)) * h[j];
    inv_gen(d, A, r);
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j) geom[(5 + i * d + j) * L.size_tot + n_t] = S(r[j][i]);
    return;
  }
  const int ng = model == AGD_RANDER ? sd + d : sd;
  if (!in) {
    for (int k = 0; k < ng; ++k) geom[k * L.size_tot + n_t] = inf_v<S>();
    return;
  }
  double a[3][3], r[3][3];
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) a[i][j] = double(m[(i * d + j) * cs + c]) * h[i] * h[j];
  inv_sym(d, a, r);
  int k = 0;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j <= i; ++j, ++k) geom[k * L.size_tot + n_t] = S(r[i][j]);
  if (model == AGD_RANDER) {
    for (int i = 0; i < d; ++i) geom[(sd + i) * L.size_tot + n_t] = S(double(w[i * cs + c]) * h[i]);
  }
}

}  // namespace agd

using namespace agd;

extern "C" {

const char* agd_last_error(void) { return g_last_error.c_str(); }
int agd_version(void) { return 1; }

int agd_device_count(int32_t* count) {
  int n = 0;
  AGD_HIP_TRY(hipGetDeviceCount(&n));
  if (count) *count = n;
  return AGD_OK;
}

int agd_block_expand(int32_t ndim, const int64_t* shape, const int32_t* shape_i, int32_t ncomp,
                     int32_t elem_size, const void* src, int64_t src_cell_stride, void* dst,
                     double pad, void* stream) {
  LayoutDesc L;
  int st = make_layout(ndim, shape, shape_i, L);
  if (st) return st;
  AGD_REQUIRE(src && dst && ncomp >= 1, "bad buffers");
  AGD_REQUIRE(src_cell_stride == 0 || src_cell_stride == 1, "cell stride must be 0 or 1");
  const dim3 grid(unsigned(ceil_div(L.size_tot, 256))), block(256);
  hipStream_t s = as_stream(stream);
  switch (elem_size) {
    case 1:
      hipLaunchKernelGGL(expand_kernel<uint8_t>, grid, block, 0, s, L, ncomp,
                         static_cast<const uint8_t*>(src), src_cell_stride,
                         static_cast<uint8_t*>(dst), uint8_t(pad));
      break;
    case 4:
      hipLaunchKernelGGL(expand_kernel<float>, grid, block, 0, s, L, ncomp,
                         static_cast<const float*>(src), src_cell_stride,
                         static_cast<float*>(dst), float(pad));
      break;
    case 8:
      hipLaunchKernelGGL(expand_kernel<double>, grid, block, 0, s, L, ncomp,
                         static_cast<const double*>(src), src_cell_stride,
                         static_cast<double*>(dst), pad);
      break;
    default: set_error("invalid argument: elem_size must be 1, 4 or 8"); return AGD_EINVAL;
  }
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_block_squeeze(int32_t ndim, const int64_t* shape, const int32_t* shape_i, int32_t ncomp,
                      int32_t elem_size, const void* src, void* dst, void* stream) {
  LayoutDesc L;
  int st = make_layout(ndim, shape, shape_i, L);
  if (st) return st;
  AGD_REQUIRE(src && dst && ncomp >= 1, "bad buffers");
  const dim3 grid(unsigned(ceil_div(L.size_tot, 256))), block(256);
  hipStream_t s = as_stream(stream);
  switch (elem_size) {
    case 1:
      hipLaunchKernelGGL(squeeze_kernel<uint8_t>, grid, block, 0, s, L, ncomp,
                         static_cast<const uint8_t*>(src), static_cast<uint8_t*>(dst));
      break;
    case 4:
      hipLaunchKernelGGL(squeeze_kernel<float>, grid, block, 0, s, L, ncomp,
                         static_cast<const float*>(src), static_cast<float*>(dst));
      break;
    case 8:
      hipLaunchKernelGGL(squeeze_kernel<double>, grid, block, 0, s, L, ncomp,
                         static_cast<const double*>(src), static_cast<double*>(dst));
      break;
    default: set_error("invalid argument: elem_size must be 1, 4 or 8"); return AGD_EINVAL;
  }
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_prepare_geometry(int32_t model, int32_t ndim, const int64_t* shape, const int32_t* shape_i,
                         int32_t scalar, const double* h, const float* m, const float* w,
                         int64_t cell_stride, void* geom, void* stream) {
  LayoutDesc L;
  int st = make_layout(ndim, shape, shape_i, L);
  if (st) return st;
  AGD_REQUIRE(ndim == 2 || ndim == 3, "ndim must be 2 or 3");
  AGD_REQUIRE(model == AGD_RIEMANN || model == AGD_RANDER || model == AGD_DIAGONAL ||
                  model == AGD_TTI,
              "geometry preparation applies to Diagonal, Riemann, Rander and TTI models");
  AGD_REQUIRE(h && m && geom, "null buffer");
  AGD_REQUIRE(model != AGD_RANDER || w, "Rander model needs the drift w");
  AGD_REQUIRE(cell_stride == 0 || cell_stride == 1, "cell stride must be 0 or 1");
  const dim3 grid(unsigned(ceil_div(L.size_tot, 256))), block(256);
  hipStream_t s = as_stream(stream);
  const double h2 = ndim == 3 ? h[2] : 1.0;
  if (scalar == AGD_F32)
    hipLaunchKernelGGL(geometry_kernel<float>, grid, block, 0, s, L, model, h[0], h[1], h2, m, w,
                       cell_stride, static_cast<float*>(geom));
  else
    hipLaunchKernelGGL(geometry_kernel<double>, grid, block, 0, s, L, model, h[0], h[1], h2, m, w,
                       cell_stride, static_cast<double*>(geom));
  AGD_HIP_TRY(hipGetLastError());
  return AGD_OK;
}

int agd_malloc(void** ptr, int64_t bytes) {
  AGD_REQUIRE(ptr && bytes >= 0, "bad allocation request");
  *ptr = nullptr;
  if (bytes == 0) return AGD_OK;
  hipError_t e = hipMalloc(ptr, size_t(bytes));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return AGD_ENOMEM;
  }
  return AGD_OK;
}

int agd_free(void* ptr) {
  if (ptr) AGD_HIP_TRY(hipFree(ptr));
  return AGD_OK;
}

int agd_memcpy_htod(void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes == 0) return AGD_OK;
  AGD_HIP_TRY(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyHostToDevice, as_stream(stream)));
  AGD_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
  return AGD_OK;
}

int agd_memcpy_dtoh(void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes == 0) return AGD_OK;
  AGD_HIP_TRY(hipMemcpyAsync(dst, src, size_t(bytes), hipMemcpyDeviceToHost, as_stream(stream)));
  AGD_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
  return AGD_OK;
}

int agd_memset(void* dst, int32_t value, int64_t bytes, void* stream) {
  if (bytes == 0) return AGD_OK;
  AGD_HIP_TRY(hipMemsetAsync(dst, value, size_t(bytes), as_stream(stream)));
  return AGD_OK;
}

int agd_stream_synchronize(void* stream) {
  AGD_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
  return AGD_OK;
}

}  // extern "C"
