// abi_noalloc.hip -- a stability-norm entry point that allocates nothing.
//
// sf_norm_noalloc computes what get_norm_elastic_from_device_ / get_norm_acoustic_from_device_
// compute (the largest |u| of a vector field, or |chi| of a potential, NaN counted as +inf), for a
// field the caller passes as a device pointer, and leaves the result in a device word the caller
// passes. It is for callers that cannot allocate (code under HIP graph capture, serving loops):
// no hipMalloc/hipFree, no host synchronisation, no memset node. The per-block maxima live in the
// file-scope device array `scratch` below (512 KiB per GPU, held until the process ends), so one
// call handles at most SF_NOALLOC_MAX_POINTS points. Calls that overlap on one GPU share that array
// and race; calls on one stream are ordered and need nothing else (every call writes each partial
// it reads).
#include <atomic>

#include "kernels.h"

#define SF_NOALLOC_BLOCK 256
#define SF_NOALLOC_PER_THREAD 8
#define SF_NOALLOC_PARTIALS 131072
#define SF_NOALLOC_MAX_POINTS ((int64_t)SF_NOALLOC_PARTIALS * SF_NOALLOC_BLOCK * SF_NOALLOC_PER_THREAD)
#define SF_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) float scratch[SF_NOALLOC_PARTIALS];

namespace sf {

__device__ __forceinline__ float point_norm(const realw* __restrict__ f, int64_t t, int ncomp) {
  realw v;
  if (ncomp == 3) {
    const realw x = f[3 * t], y = f[3 * t + 1], z = f[3 * t + 2];
    v = sqrtf(x * x + y * y + z * z);
  } else {
    v = fabsf(f[t]);
  }
  return v != v ? INFINITY : v;
}

// block b covers points [b * 2048, (b + 1) * 2048); its maximum goes to partial[b]
__global__ void __launch_bounds__(SF_NOALLOC_BLOCK) k_norm_partials(int64_t npts, int ncomp,
                                                                    const realw* __restrict__ f,
                                                                    float* __restrict__ partial) {
  __shared__ float sm[SF_NOALLOC_BLOCK / 64];
  const int64_t base = (int64_t)blockIdx.x * SF_NOALLOC_BLOCK * SF_NOALLOC_PER_THREAD;
  float m = 0.f;
#pragma unroll
  for (int r = 0; r < SF_NOALLOC_PER_THREAD; r++) {
    const int64_t t = base + (int64_t)r * SF_NOALLOC_BLOCK + threadIdx.x;
    if (t < npts) m = fmaxf(m, point_norm(f, t, ncomp));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = sm[0];
    for (int w = 1; w < SF_NOALLOC_BLOCK / 64; w++) r = fmaxf(r, sm[w]);
    partial[blockIdx.x] = r;
  }
}

// one block folds the nb partial maxima into out[0] (0 when nb = 0)
__global__ void __launch_bounds__(SF_NOALLOC_BLOCK) k_norm_final(int nb, const float* __restrict__ partial,
                                                                 float* __restrict__ out) {
  __shared__ float sm[SF_NOALLOC_BLOCK / 64];
  float m = 0.f;
  for (int b = threadIdx.x; b < nb; b += SF_NOALLOC_BLOCK) m = fmaxf(m, partial[b]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float r = sm[0];
    for (int w = 1; w < SF_NOALLOC_BLOCK / 64; w++) r = fmaxf(r, sm[w]);
    out[0] = r;
  }
}

// device address of `scratch` on each GPU, looked up once per device (each GPU has its own copy
// of the array at its own address; the array's name in host code is only a stand-in)
static std::atomic<float*> g_scratch[SF_NOALLOC_MAX_DEVICES];

static float* scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SF_NOALLOC_MAX_DEVICES) {
    set_error("sf_norm_noalloc: no current HIP device (or device index above 63)");
    return nullptr;
  }
  float* p = g_scratch[dev].load(std::memory_order_acquire);
  if (p) return p;
  void* addr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !addr) {
    set_error("sf_norm_noalloc: hipGetSymbolAddress(scratch) failed");
    return nullptr;
  }
  p = static_cast<float*>(addr);
  g_scratch[dev].store(p, std::memory_order_release);
  return p;
}

}  // namespace sf

// No counterpart in the reference (its get_norm_*_from_device allocate a partial-maximum buffer on
// every call); for callers that cannot allocate. Returns 0, or -1 with the reason in sf_last_error().
extern "C" int sf_norm_noalloc(const realw* field, int64_t npts, int ncomp, realw* out, void* stream) {
  using namespace sf;
  if (ncomp != 1 && ncomp != 3) {
    set_error("sf_norm_noalloc: ncomp must be 1 (potential) or 3 (vector field)");
    return -1;
  }
  if (npts < 0 || npts > SF_NOALLOC_MAX_POINTS) {
    set_error("sf_norm_noalloc: npts must lie in [0, 268435456] (the partial maxima of at most 131072 "
              "blocks of 2048 points fit the library's fixed scratch array)");
    return -1;
  }
  if (!out || (npts > 0 && !field)) {
    set_error("sf_norm_noalloc: null field or output pointer");
    return -1;
  }
  float* part = scratch_on_current_device();
  if (!part) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t per_block = (int64_t)SF_NOALLOC_BLOCK * SF_NOALLOC_PER_THREAD;
  const int nb = (int)((npts + per_block - 1) / per_block);
  if (nb > 0) hipLaunchKernelGGL(k_norm_partials, dim3(nb), dim3(SF_NOALLOC_BLOCK), 0, s, npts, ncomp, field, part);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(SF_NOALLOC_BLOCK), 0, s, nb, part, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return -1;
  }
  return 0;
}
