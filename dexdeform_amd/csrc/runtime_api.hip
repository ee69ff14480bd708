// Device memory, copy and stream helpers with the names and argument order of the
// reference's host API (mpm/csrc/integrator.cu:1987-2072), bound by ctypes in
// mpm/types.py:103-128, plus its compute_grid_lower launcher (integrator.cu:70-82,
// 1619-1625, bound at types.py:130-137). They keep the reference's ctypes layer working
// against this library: its `array` class allocates, uploads, downloads and clears
// through them, and MPMSimulator creates and synchronises its two streams with them.
//
// Differences from the reference, all in error handling: every helper that returns void
// there returns a status here (0, or the hipError_t of the failing call); cuda_alloc
// returns NULL when the allocation fails (the reference prints and returns an
// uninitialised pointer); cuda_stream_create returns NULL on failure.
#include <hip/hip_runtime.h>
#include <climits>
#include <stdio.h>

#include "../../include/dexdeform_mpm.h"

#define DDM_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr int kThreads = 256;

// Lower grid corner of the particles, 10 cells of margin: floor((x - 10 dx) / dx) per
// axis, min-reduced into grid_lower (which the caller initialises, e.g. to INT_MAX).
// Wave-reduce first, then one atomic per axis per wavefront.
__global__ __launch_bounds__(kThreads) void k_compute_grid_lower(const float *__restrict__ x, float dx, float inv_dx,
                                                                 int *__restrict__ grid_lower, int n) {
  int p = blockIdx.x * kThreads + threadIdx.x;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX};
  if (p < n) {
    for (int c = 0; c < 3; ++c) lo[c] = (int)floorf((x[3 * p + c] - dx * 10.f) * inv_dx);
  }
  for (int c = 0; c < 3; ++c) {
    int v = lo[c];
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    lo[c] = v;
  }
  if ((threadIdx.x & 63) == 0 && lo[0] != INT_MAX) {
    atomicMin(grid_lower + 0, lo[0]);
    atomicMin(grid_lower + 1, lo[1]);
    atomicMin(grid_lower + 2, lo[2]);
  }
}

inline int st(hipError_t e) { return e == hipSuccess ? 0 : (int)e; }

}  // namespace

DDM_API void *cuda_alloc(size_t size) {
  void *p = nullptr;
  if (hipMalloc(&p, size) != hipSuccess) return nullptr;
  return p;
}

DDM_API int cuda_free(void *ptr) { return st(hipFree(ptr)); }

DDM_API void print_memory_info(void) {
  int n = 0, id = 0;
  size_t free_b = 0, total_b = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || hipGetDevice(&id) != hipSuccess ||
      hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    printf("GPU memory: query failed\n");
    return;
  }
  printf("GPU %d/%d memory: free=%.3f, total=%.3f\n", id, n, free_b / 1073741824.0, total_b / 1073741824.0);
}

DDM_API hipStream_t cuda_stream_create(void) {
  hipStream_t s = nullptr;
  if (hipStreamCreate(&s) != hipSuccess) return nullptr;  // default priority
  return s;
}

DDM_API int cuda_stream_destroy(hipStream_t stream) { return st(hipStreamDestroy(stream)); }

DDM_API int cuda_stream_sync(hipStream_t stream) { return st(hipStreamSynchronize(stream)); }

DDM_API int cuda_upload(void *device_ptr, void *host_ptr, size_t size) {
  return st(hipMemcpy(device_ptr, host_ptr, size, hipMemcpyHostToDevice));
}

DDM_API int cuda_download(void *host_ptr, void *device_ptr, size_t size) {
  return st(hipMemcpy(host_ptr, device_ptr, size, hipMemcpyDeviceToHost));
}

DDM_API int cuda_copy(void *dst, void *src, size_t size) {
  return st(hipMemcpy(dst, src, size, hipMemcpyDeviceToDevice));
}

DDM_API int cuda_copy2d(void *dst, size_t dpitch, void *src, size_t spitch, size_t width, size_t height) {
  return st(hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToDevice));
}

DDM_API int cuda_zero(void *ptr, size_t size) { return st(hipMemset(ptr, 0, size)); }

DDM_API int cuda_zero_async(void *ptr, size_t size, hipStream_t stream) {
  return st(hipMemsetAsync(ptr, 0, size, stream));
}

DDM_API int cuda_upload_async(void *device_ptr, void *host_ptr, size_t size, hipStream_t stream) {
  return st(hipMemcpyAsync(device_ptr, host_ptr, size, hipMemcpyHostToDevice, stream));
}

DDM_API int cuda_download_async(void *host_ptr, void *device_ptr, size_t size, hipStream_t stream) {
  return st(hipMemcpyAsync(host_ptr, device_ptr, size, hipMemcpyDeviceToHost, stream));
}

DDM_API int cuda_copy_async(void *dst, void *src, size_t size, hipStream_t stream) {
  return st(hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToDevice, stream));
}

DDM_API int compute_grid_lower(float *particle_x, float dx, float inv_dx, int *grid_lower, int dim,
                               hipStream_t stream) {
  if (dim < 0 || !particle_x || !grid_lower) return DDM_EINVAL;
  if (dim == 0) return 0;
  k_compute_grid_lower<<<(dim + kThreads - 1) / kThreads, kThreads, 0, stream>>>(particle_x, dx, inv_dx, grid_lower,
                                                                                dim);
  return st(hipGetLastError());
}
