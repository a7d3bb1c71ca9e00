// Passes and set-up helpers shared by the periodic and wall-bounded paths.
#pragma once

#include <vector>

#include "context.hpp"
#include "dft.hpp"

namespace dnsmi {

// x pass (c2r, nonlinear term, r2c) over rows (y, z): periodic.hip
// em: EvalMode (EV_RECORD / EV_INC need T = float and the reference buffers)
template <typename T>
int launch_x_pass(Ctx* c, const cx<T>* xin, cx<T>* xout, const double* base, const double* cbase,
                  const double* nbase, int ny, int kid, hipStream_t s, int em);
// z forward over columns (y, kx): periodic.hip
template <typename T>
int launch_z_forward_buf(Ctx* c, const cx<T>* in, cx<T>* out, int ny, int kid, hipStream_t s);

int setup_twiddles(Ctx* c);
int upload_doubles(Ctx* c, double** dst, const std::vector<double>& v);
std::vector<double> real_wavenumbers(int n, double L);
std::vector<double> complex_wavenumbers(int n, double L);
bool supported_length(int P);

// Deterministic block sum: fixed shuffle tree within each wave, then the
// waves in order (thread 0 holds the result).
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
  }
  return s;
}

}  // namespace dnsmi
