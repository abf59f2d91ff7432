// Auxiliary device kernels and the HIP runtime glue used by the C host layer.
//
//   K1  twiddle generation (reference twiddle_factors.c:17-131): FP64 on the
//       device, rounded once to the working precision.
//   K7  digit-reversal gather / scatter (reference index_reversals.c:85-94,
//       used by the ordered transforms fft.c:12-71): the index map is the
//       closed form, evaluated per element; no rho table is read.
//   pointwise product (reference convolution.c:20-26) and precision conversion.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <mutex>

#include "pacg.h"
#include "pacg_internal.h"

namespace {

template <typename T>
__global__ void twiddle_kernel(T* __restrict__ out, uint32_t N, uint32_t count, uint32_t step) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  // w_N^e = exp(-2 pi i e / N): sincospi of the exact dyadic 2e/N
  const uint64_t e = ((uint64_t)i * step) % N;
  double s, c;
  sincospi(2.0 * (double)e / (double)N, &s, &c);
  out[2 * i] = (T)c;
  out[2 * i + 1] = (T)(-s);
}

__device__ __forceinline__ uint32_t rev_axis(uint32_t i, int nst, const uint8_t* lg, uint32_t n) {
  // rho(i) = sum_q d_q * n / k_q, d_q the stage-q digit of i
  uint32_t out = 0, sh = 0;
  for (int q = 0; q < nst; ++q) {
    const uint32_t b = lg[q];
    const uint32_t d = (i >> sh) & ((1u << b) - 1u);
    sh += b;
    out += d * (n >> sh);
  }
  return out;
}

template <typename T>
__global__ void permute_kernel(const T* __restrict__ ir, const T* __restrict__ ii, T* __restrict__ orr,
                               T* __restrict__ oi, pacg_perm_desc_t d, uint64_t total, int scatter) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= total) return;
  uint64_t rem = p, acc = 0, mul = 1;
  for (int a = d.nd - 1; a >= 0; --a) {
    const uint32_t i = (uint32_t)(rem % d.dims[a]);
    rem /= d.dims[a];
    acc += (uint64_t)(d.nst[a] ? rev_axis(i, d.nst[a], d.radix_log2[a], d.dims[a]) : i) * mul;
    mul *= d.dims[a];
  }
  if (scatter) {
    orr[acc] = ir[p];
    oi[acc] = ii[p];
  } else {
    orr[p] = ir[acc];
    oi[p] = ii[acc];
  }
}

template <typename T>
__global__ void gather_rho_kernel(const T* __restrict__ ir, const T* __restrict__ ii, T* __restrict__ orr,
                                  T* __restrict__ oi, const int* __restrict__ rho, uint64_t n) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint64_t src = (uint32_t)rho[p];
  orr[p] = ir[src];
  oi[p] = ii[src];
}

template <typename T, typename W>
__global__ void pointwise_kernel(T* __restrict__ zr, T* __restrict__ zi, const W* __restrict__ wr,
                                 const W* __restrict__ wi, uint64_t n, T scale) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const T a = zr[p], b = zi[p], c = (T)wr[p], dd = (T)wi[p];
  zr[p] = (a * c - b * dd) * scale;
  zi[p] = (a * dd + b * c) * scale;
}

template <typename X, typename Y>
__global__ void convert_kernel(const X* __restrict__ xr, const X* __restrict__ xi, Y* __restrict__ yr,
                               Y* __restrict__ yi, uint64_t n, X scale) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  yr[p] = (Y)(xr[p] * scale);
  yi[p] = (Y)(xi[p] * scale);
}

// HBM probe (the GPU counterpart of the reference's memory_bw_test.c:27-67):
// streaming read (reduced so it is not optimised away), write and copy with
// 16-byte accesses over a grid-stride loop.
__global__ void bw_read_kernel(const float4* __restrict__ src, uint64_t n4, float* __restrict__ sink) {
  float acc = 0.f;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
    const float4 v = src[i];
    acc += v.x + v.y + v.z + v.w;
  }
  if (acc == 1234.5f) sink[0] = acc;  // never true for the zero-filled probe buffers
}
__global__ void bw_write_kernel(float4* __restrict__ dst, uint64_t n4) {
  const float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}
__global__ void bw_copy_kernel(float4* __restrict__ dst, const float4* __restrict__ src, uint64_t n4) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

}  // namespace

extern "C" {

int pacg_gen_twiddles(void* out, int prec, uint32_t N, uint32_t count, uint32_t step, void* stream) {
  if (count == 0) return 0;
  if (prec == PACG_F64)
    hipLaunchKernelGGL(twiddle_kernel<double>, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream,
                       (double*)out, N, count, step);
  else
    hipLaunchKernelGGL(twiddle_kernel<float>, dim3(blocks_for(count, 256)), dim3(256), 0, (hipStream_t)stream,
                       (float*)out, N, count, step);
  return (int)hipGetLastError();
}

int pacg_permute(const void* in_re, const void* in_im, void* out_re, void* out_im, int prec,
                 const pacg_perm_desc_t* d, int scatter, void* stream) {
  uint64_t total = 1;
  for (int a = 0; a < d->nd; ++a) total *= d->dims[a];
  if (total == 0) return 0;
  if (prec == PACG_F64)
    hipLaunchKernelGGL(permute_kernel<double>, dim3(blocks_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const double*)in_re, (const double*)in_im, (double*)out_re, (double*)out_im, *d, total,
                       scatter);
  else
    hipLaunchKernelGGL(permute_kernel<float>, dim3(blocks_for(total, 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)in_re, (const float*)in_im, (float*)out_re, (float*)out_im, *d, total, scatter);
  return (int)hipGetLastError();
}

int pacg_gather_rho(const void* in_re, const void* in_im, void* out_re, void* out_im, const int* rho, uint64_t n,
                    int prec, void* stream) {
  if (n == 0) return 0;
  dim3 g(blocks_for(n, 256)), b(256);
  if (prec == PACG_F64)
    hipLaunchKernelGGL(gather_rho_kernel<double>, g, b, 0, (hipStream_t)stream, (const double*)in_re,
                       (const double*)in_im, (double*)out_re, (double*)out_im, rho, n);
  else
    hipLaunchKernelGGL(gather_rho_kernel<float>, g, b, 0, (hipStream_t)stream, (const float*)in_re,
                       (const float*)in_im, (float*)out_re, (float*)out_im, rho, n);
  return (int)hipGetLastError();
}

int pacg_pointwise(void* zre, void* zim, const void* wre, const void* wim, int prec, int wprec, uint64_t n,
                   double scale, void* stream) {
  if (n == 0) return 0;
  dim3 g(blocks_for(n, 256)), b(256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == PACG_F64)
    hipLaunchKernelGGL((pointwise_kernel<double, double>), g, b, 0, s, (double*)zre, (double*)zim,
                       (const double*)wre, (const double*)wim, n, scale);
  else if (wprec == PACG_F64)
    hipLaunchKernelGGL((pointwise_kernel<float, double>), g, b, 0, s, (float*)zre, (float*)zim, (const double*)wre,
                       (const double*)wim, n, (float)scale);
  else
    hipLaunchKernelGGL((pointwise_kernel<float, float>), g, b, 0, s, (float*)zre, (float*)zim, (const float*)wre,
                       (const float*)wim, n, (float)scale);
  return (int)hipGetLastError();
}

int pacg_convert(const void* xre, const void* xim, int xprec, void* yre, void* yim, int yprec, uint64_t n,
                 double scale, void* stream) {
  if (n == 0) return 0;
  dim3 g(blocks_for(n, 256)), b(256);
  hipStream_t s = (hipStream_t)stream;
  if (xprec == PACG_F64 && yprec == PACG_F64)
    hipLaunchKernelGGL((convert_kernel<double, double>), g, b, 0, s, (const double*)xre, (const double*)xim,
                       (double*)yre, (double*)yim, n, scale);
  else if (xprec == PACG_F64)
    hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, s, (const double*)xre, (const double*)xim,
                       (float*)yre, (float*)yim, n, scale);
  else if (yprec == PACG_F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, s, (const float*)xre, (const float*)xim,
                       (double*)yre, (double*)yim, n, (float)scale);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), g, b, 0, s, (const float*)xre, (const float*)xim,
                       (float*)yre, (float*)yim, n, (float)scale);
  return (int)hipGetLastError();
}

int pacg_bw_probe(int mode, void* dst, const void* src, uint64_t bytes, void* stream) {
  const uint64_t n4 = bytes / 16;
  if (n4 == 0) return 0;
  const uint32_t blocks = 256 * 16;
  hipStream_t s = (hipStream_t)stream;
  if (mode == 0)
    hipLaunchKernelGGL(bw_read_kernel, dim3(blocks), dim3(256), 0, s, (const float4*)src, n4, (float*)dst);
  else if (mode == 1)
    hipLaunchKernelGGL(bw_write_kernel, dim3(blocks), dim3(256), 0, s, (float4*)dst, n4);
  else
    hipLaunchKernelGGL(bw_copy_kernel, dim3(blocks), dim3(256), 0, s, (float4*)dst, (const float4*)src, n4);
  return (int)hipGetLastError();
}

// ---- runtime glue ----------------------------------------------------------
int pacg_dev_malloc(void** p, size_t bytes) { return (int)hipMalloc(p, bytes ? bytes : 1); }
int pacg_dev_free(void* p) { return p ? (int)hipFree(p) : 0; }
int pacg_host_malloc(void** p, size_t bytes) { return (int)hipHostMalloc(p, bytes, hipHostMallocDefault); }
int pacg_host_free(void* p) { return p ? (int)hipHostFree(p) : 0; }
int pacg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
}
int pacg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
}
int pacg_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  return (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
}
int pacg_stream_sync(void* stream) { return (int)hipStreamSynchronize((hipStream_t)stream); }
int pacg_stream_create(void** stream) { return (int)hipStreamCreateWithFlags((hipStream_t*)stream, hipStreamNonBlocking); }
int pacg_stream_destroy(void* stream) { return (int)hipStreamDestroy((hipStream_t)stream); }
int pacg_device_sync(void) { return (int)hipDeviceSynchronize(); }
int pacg_get_device(int* dev) { return (int)hipGetDevice(dev); }
int pacg_device_count(int* n) { return (int)hipGetDeviceCount(n); }
int pacg_num_cus(void) {
  static int cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}
int pacg_last_hip_error(void) { return (int)hipGetLastError(); }
const char* pacg_hip_error_string(int code) { return hipGetErrorString((hipError_t)code); }
int pacg_event_create(void** ev) { return (int)hipEventCreate((hipEvent_t*)ev); }
int pacg_event_destroy(void* ev) { return (int)hipEventDestroy((hipEvent_t)ev); }
int pacg_event_record(void* ev, void* stream) { return (int)hipEventRecord((hipEvent_t)ev, (hipStream_t)stream); }
int pacg_event_sync(void* ev) { return (int)hipEventSynchronize((hipEvent_t)ev); }
int pacg_event_elapsed_ms(float* ms, void* start, void* stop) {
  return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}
int pacg_graph_begin(void* stream) {
  return (int)hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal);
}
int pacg_graph_end(void* stream, void** exec) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
  if (e != hipSuccess) return (int)e;
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return (int)e;
  *exec = (void*)x;
  return 0;
}
int pacg_graph_launch(void* exec, void* stream) { return (int)hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream); }
int pacg_graph_destroy(void* exec) { return (int)hipGraphExecDestroy((hipGraphExec_t)exec); }

}  // extern "C"

namespace {
__global__ void empty_kernel() {}
}  // namespace

extern "C" int pacg_launch_floor(int reps, void* stream, float* direct_us, float* graph_us) {
  if (reps < 1 || !direct_us || !graph_us) return (int)hipErrorInvalidValue;
  hipStream_t s = (hipStream_t)stream, own = nullptr;
  hipError_t e = hipSuccess;
  if (!s) {
    if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return (int)e;
    s = own;
  }
  hipEvent_t a = nullptr, b = nullptr;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  float ms = 0.0f;
  if ((e = hipEventCreate(&a)) != hipSuccess || (e = hipEventCreate(&b)) != hipSuccess) goto out;
  for (int i = 0; i < 10; ++i) hipLaunchKernelGGL(empty_kernel, dim3(1), dim3(64), 0, s);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) goto out;
  hipEventRecord(a, s);
  for (int i = 0; i < reps; ++i) hipLaunchKernelGGL(empty_kernel, dim3(1), dim3(64), 0, s);
  hipEventRecord(b, s);
  if ((e = hipEventSynchronize(b)) != hipSuccess || (e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) goto out;
  *direct_us = 1000.0f * ms / (float)reps;
  if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) goto out;
  hipLaunchKernelGGL(empty_kernel, dim3(1), dim3(64), 0, s);
  if ((e = hipStreamEndCapture(s, &g)) != hipSuccess) goto out;
  if ((e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) goto out;
  for (int i = 0; i < 10; ++i) hipGraphLaunch(x, s);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) goto out;
  hipEventRecord(a, s);
  for (int i = 0; i < reps; ++i) hipGraphLaunch(x, s);
  hipEventRecord(b, s);
  if ((e = hipEventSynchronize(b)) != hipSuccess || (e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) goto out;
  *graph_us = 1000.0f * ms / (float)reps;
out:
  if (x) hipGraphExecDestroy(x);
  if (g) hipGraphDestroy(g);
  if (a) hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (own) hipStreamDestroy(own);
  return (int)e;
}

// Scratch of the allocation-free entry point (pacg_convolve_fft_based_noalloc):
// a fixed array in the code object, one copy on every GPU the library runs
// on, held until the process ends.  Its name is a host-side stand-in for a
// device address: host code uses only the address hipGetSymbolAddress gives
// for the current device, resolved once per device and kept here.
static __device__ __align__(256) unsigned char scratch[PACG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

extern "C" int pacg_noalloc_scratch(void** p) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* q = nullptr;
    e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = q;
  }
  *p = g_scratch_ptr[dev];
  return 0;
}
