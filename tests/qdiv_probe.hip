// Test infrastructure only: exposes the shared-denominator quotient of
// storm_amd/csrc/fvt_math.h (recip/qdiv) next to the plain division, and the
// friction factor's pow_one_third, for tests/test_gpu_qdiv.py.  Built by
// __graft_entry__.build() into
// tests/libqdiv_probe.so.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../storm_amd/csrc/fvt_math.h"

using namespace storm;

// out[3i] = n0/d, out[3i+1] = n1/d, out[3i+2] = n2/d through one Recip of d
// (taken with n0); ref[3i..] the plain quotients.
__global__ void k_qdiv(int n, const double* num, const double* den, double* out, double* ref) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = den[i];
  const double a = num[3 * i], b = num[3 * i + 1], c = num[3 * i + 2];
  const Recip r = recip(d, a);
  out[3 * i] = qdiv(r, a, d);
  out[3 * i + 1] = qdiv(r, b, d);
  out[3 * i + 2] = qdiv(r, c, d);
  ref[3 * i] = a / d;
  ref[3 * i + 1] = b / d;
  ref[3 * i + 2] = c / d;
}

extern "C" int qdiv_probe(int n, const double* num, const double* den, double* out, double* ref) {
  double *dn = nullptr, *dd = nullptr, *dout = nullptr, *dref = nullptr;
  if (n <= 0) return 0;
  const size_t b3 = sizeof(double) * 3 * (size_t)n, b1 = sizeof(double) * (size_t)n;
  if (hipMalloc(&dn, b3) != hipSuccess || hipMalloc(&dd, b1) != hipSuccess ||
      hipMalloc(&dout, b3) != hipSuccess || hipMalloc(&dref, b3) != hipSuccess)
    return -1;
  int rc = 0;
  if (hipMemcpy(dn, num, b3, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dd, den, b1, hipMemcpyHostToDevice) != hipSuccess)
    rc = -2;
  if (!rc) {
    hipLaunchKernelGGL(k_qdiv, dim3((n + 255) / 256), dim3(256), 0, 0, n, dn, dd, dout, dref);
    if (hipDeviceSynchronize() != hipSuccess) rc = -3;
  }
  if (!rc && (hipMemcpy(out, dout, b3, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(ref, dref, b3, hipMemcpyDeviceToHost) != hipSuccess))
    rc = -4;
  (void)hipFree(dn); (void)hipFree(dd); (void)hipFree(dout); (void)hipFree(dref);
  return rc;
}

__global__ void k_pow13(int n, const double* h, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = pow_one_third(h[i]);
}

// out[i] = pow_one_third(h[i]) on the device
extern "C" int pow13_probe(int n, const double* h, double* out) {
  double *dh = nullptr, *dout = nullptr;
  if (n <= 0) return 0;
  const size_t b1 = sizeof(double) * (size_t)n;
  if (hipMalloc(&dh, b1) != hipSuccess || hipMalloc(&dout, b1) != hipSuccess) return -1;
  int rc = 0;
  if (hipMemcpy(dh, h, b1, hipMemcpyHostToDevice) != hipSuccess) rc = -2;
  if (!rc) {
    hipLaunchKernelGGL(k_pow13, dim3((n + 255) / 256), dim3(256), 0, 0, n, dh, dout);
    if (hipDeviceSynchronize() != hipSuccess) rc = -3;
  }
  if (!rc && hipMemcpy(out, dout, b1, hipMemcpyDeviceToHost) != hipSuccess) rc = -4;
  (void)hipFree(dh); (void)hipFree(dout);
  return rc;
}
