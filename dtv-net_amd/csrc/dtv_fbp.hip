// FBP front end on gfx950: cosine weighting fused with the filtering of each
// detector row by a spatial kernel (the Ram-Lak taps of the reference's
// ``__conv__``), in one pass over the sinogram.
//
// The reference multiplies by ``cosWeight`` and then runs ``conv2d`` with a
// 1 x (W+1) kernel and padding W/2 over rows of length W
// (StandardGeometry.py:55-65, BeijingGeometry.py:60-63); PyTorch's conv2d is
// a cross-correlation, so
//   out[r][x] = sum_k taps[k] * w[r % wrows][j] * in[r][j],  j = x + k - pad,
// pad = (ntaps - 1) / 2, terms with j outside [0, W) dropped.
//
// One workgroup per row: the weighted row is staged in LDS with zero borders,
// the taps in LDS (padded to a multiple of 8 with zeros); each lane produces
// 8 consecutive outputs with a sliding register window, so every LDS read of
// a sample feeds 8 FMAs and every tap is a broadcast read.
#include <hip/hip_runtime.h>

#include "dtvproj.h"

namespace {

constexpr int kOut = 8;  // outputs per lane

__global__ __launch_bounds__(256) void k_row_filter(const float* __restrict__ in,
                                                    float* __restrict__ out, int rows, int W,
                                                    const float* __restrict__ w, int wrows,
                                                    const float* __restrict__ taps, int ntaps,
                                                    int ntaps8) {
  extern __shared__ float smem[];
  const int pad = (ntaps - 1) / 2;
  const int slen = W + ntaps8 + kOut;  // padded row: s[i] = row[i - pad]
  float* s = smem;
  float* t = smem + slen;
  for (int i = threadIdx.x; i < ntaps8; i += blockDim.x) t[i] = i < ntaps ? taps[i] : 0.f;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const float* ir = in + (size_t)r * W;
    const float* wr = w ? w + (size_t)(r % wrows) * W : nullptr;
    __syncthreads();  // previous row's readers are done with s
    for (int i = threadIdx.x; i < slen; i += blockDim.x) {
      const int j = i - pad;
      float v = 0.f;
      if (j >= 0 && j < W) v = wr ? wr[j] * ir[j] : ir[j];
      s[i] = v;
    }
    __syncthreads();
    for (int x0 = threadIdx.x * kOut; x0 < W; x0 += blockDim.x * kOut) {
      float acc[kOut], win[2 * kOut];
#pragma unroll
      for (int j = 0; j < kOut; ++j) {
        acc[j] = 0.f;
        win[j] = s[x0 + j];
      }
      // taps k .. k+7 against samples x0+k .. x0+k+14
      for (int k = 0; k < ntaps8; k += kOut) {
#pragma unroll
        for (int j = 0; j < kOut; ++j) win[kOut + j] = s[x0 + k + kOut + j];
#pragma unroll
        for (int q = 0; q < kOut; ++q) {
          const float tq = t[k + q];
#pragma unroll
          for (int j = 0; j < kOut; ++j) acc[j] += tq * win[j + q];
        }
#pragma unroll
        for (int j = 0; j < kOut; ++j) win[j] = win[kOut + j];
      }
      float* orow = out + (size_t)r * W;
#pragma unroll
      for (int j = 0; j < kOut; ++j)
        if (x0 + j < W) orow[x0 + j] = acc[j];
    }
  }
}

}  // namespace

extern "C" int dtv_row_filter(const float* input, float* output, int64_t rows, int32_t width,
                              const float* weight, int32_t weight_rows, const float* taps,
                              int32_t n_taps, void* stream) {
  if (!input || !output || !taps) return DTV_ERR_NULL_POINTER;
  if (rows < 1 || width < 1 || n_taps < 1 || n_taps % 2 == 0) return DTV_ERR_INVALID_ARGUMENT;
  if (weight && weight_rows < 1) return DTV_ERR_INVALID_ARGUMENT;
  if (rows >= (1ll << 31)) return DTV_ERR_TOO_LARGE;
  if (input == output) return DTV_ERR_INVALID_ARGUMENT;
  const int ntaps8 = (n_taps + kOut - 1) / kOut * kOut;
  const size_t lds = ((size_t)width + 2 * ntaps8 + kOut) * sizeof(float);
  if (lds > 65536) return DTV_ERR_TOO_LARGE;
  const int lanes = (width + kOut - 1) / kOut;
  const int block = lanes >= 256 ? 256 : ((lanes + 63) / 64) * 64;
  const unsigned grid = (unsigned)(rows < 65536 ? rows : 65536);
  hipLaunchKernelGGL(k_row_filter, dim3(grid), dim3(block), lds, (hipStream_t)stream, input,
                     output, (int)rows, width, weight, weight_rows, taps, n_taps, ntaps8);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}
