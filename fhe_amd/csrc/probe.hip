// Issue-rate calibration of the modular products the kernels are built from.
//
// The hot path is bound by 62-bit modular arithmetic built from 32-bit
// multiply-adds, not by HBM, so its roofline is a modular-product rate.
// This kernel measures that ceiling on the device at hand: every thread runs
// eight independent chains of one product form (no loads or stores inside
// the loop), at full occupancy and far more workgroups than compute units.
#include "fhe_amd.h"
#include "kernels.h"

namespace fhe {

namespace {

constexpr int kChains = 8;

template <int FORM>
__global__ __launch_bounds__(kBlock) void modmul_probe_kernel(
    uint64_t *out, int iters, uint64_t q, uint64_t k, uint64_t w,
    uint64_t wq) {
  const uint64_t tid = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint64_t y[kChains];
#pragma unroll
  for (int j = 0; j < kChains; j++)
    y[j] = (tid * 0x9e3779b97f4a7c15ull + uint64_t(j) * 0x632be59bd9b4e019ull) %
           q;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int j = 0; j < kChains; j++) {
      if (FORM == 0)
        y[j] = shoup_mult(y[j], w, wq, q); // lazy butterflies: [0, 2q)
      else
        y[j] = uint64_t(mont_mult_u(int64_t(y[j]), int64_t(w), q, k));
    }
  }
  uint64_t acc = 0;
#pragma unroll
  for (int j = 0; j < kChains; j++)
    acc ^= y[j];
  out[tid] = acc;
}

} // namespace

// Products per second of `form` (0: Shoup, 1: Montgomery) over a grid of
// `blocks` workgroups; synchronizes `stream`.
double measure_modmul_rate(int form, int blocks, int iters, uint64_t q,
                           uint64_t k, uint64_t w, uint64_t wq, uint64_t *out,
                           hipStream_t stream) {
  hipEvent_t start, stop;
  if (hipEventCreate(&start) != hipSuccess)
    return -1;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start);
    return -1;
  }
  auto launch = [&] {
    if (form == 0)
      hipLaunchKernelGGL(modmul_probe_kernel<0>, dim3(blocks), dim3(kBlock), 0,
                         stream, out, iters, q, k, w, wq);
    else
      hipLaunchKernelGGL(modmul_probe_kernel<1>, dim3(blocks), dim3(kBlock), 0,
                         stream, out, iters, q, k, w, wq);
  };
  launch(); // warm-up
  (void)hipEventRecord(start, stream);
  launch();
  (void)hipEventRecord(stop, stream);
  float ms = -1;
  if (hipEventSynchronize(stop) != hipSuccess ||
      hipEventElapsedTime(&ms, start, stop) != hipSuccess)
    ms = -1;
  (void)hipEventDestroy(start);
  (void)hipEventDestroy(stop);
  if (ms <= 0)
    return -1;
  const double products = double(blocks) * kBlock * kChains * double(iters);
  return products / (double(ms) * 1e-3);
}

} // namespace fhe
