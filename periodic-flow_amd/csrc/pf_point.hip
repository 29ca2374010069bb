// Point-wise k-space kernels: stage update, projection, FP64 reductions.
#include "pf_launch.h"

namespace pf {

// steps done: advance the RNG step counter of every active member
__global__ void member_step_kernel(long long* msteps, int n, long long by) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) msteps[i] += by;
}

__global__ void fill_kernel(double* p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Leray projection (math.py:72-100) with optional Nyquist mask and k = 0 mask;
// in may equal out.
template <typename T, int D>
__global__ void __launch_bounds__(kThreads)
project_kernel(StageArgs g, const void* in, void* out, int mask_nyq, int zero_k0) {
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= g.nk) return;
  const Mode<T, D> m = decode<T, D>(g, idx);
  const long long mb = (long long)b * D * g.nk + idx;
  const cplx<T>* src = (const cplx<T>*)in + mb;
  cplx<T>* dst = (cplx<T>*)out + mb;
  cplx<T> v[D];
#pragma unroll
  for (int i = 0; i < D; ++i) v[i] = src[i * g.nk];
  if (zero_k0 < 2) project<T, D>(v, m);
  const bool kill = (mask_nyq && m.nyq) || (zero_k0 && m.k2 == (T)0);
#pragma unroll
  for (int i = 0; i < D; ++i) dst[i * g.nk] = kill ? mk<T>(0, 0) : v[i];
}

static dim3 mode_grid(long long nk, int nmem) {
  return dim3((unsigned)((nk + kThreads - 1) / kThreads), (unsigned)nmem);
}

// Modes per thread: four in 2D when the launch has enough modes to fill
// the GPU several times over (the noise chains of Stokes flow then
// interleave), else one; one in 3D (where four would halve the occupancy).
// cp: cache policy of the state loads and stores (0, or 3: nontemporal).
template <typename T, int CP> static void stage_launch(int dim, const StageArgs& a, hipStream_t s) {
  auto grid = [&](int mpt) {
    return dim3((unsigned)((a.nk + (long long)kThreads * mpt - 1) / ((long long)kThreads * mpt)),
                (unsigned)a.n_members);
  };
  const bool wide = a.nk * a.n_members >= (4ll << 20);
  if (dim == 3) launch_pass_kernel(stage_kernel<T, 3, 1, CP>, grid(1), dim3(kThreads), s, a);
  else if (wide) launch_pass_kernel(stage_kernel<T, 2, 4, CP>, grid(4), dim3(kThreads), s, a);
  else launch_pass_kernel(stage_kernel<T, 2, 1, CP>, grid(1), dim3(kThreads), s, a);
}
template <typename T> int launch_stage(int dim, const StageArgs& a, hipStream_t s, int cp) {
  if (cp == 3) stage_launch<T, 3>(dim, a, s);
  else if (cp == 0) stage_launch<T, 0>(dim, a, s);
  else return -1;
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T>
int launch_project(int dim, const StageArgs& g, const void* in, void* out, int mask_nyq, int zero_k0, int nmem,
                   hipStream_t s) {
  if (dim == 3)
    hipLaunchKernelGGL((project_kernel<T, 3>), mode_grid(g.nk, nmem), dim3(kThreads), 0, s, g, in, out, mask_nyq,
                       zero_k0);
  else
    hipLaunchKernelGGL((project_kernel<T, 2>), mode_grid(g.nk, nmem), dim3(kThreads), 0, s, g, in, out, mask_nyq,
                       zero_k0);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

static unsigned red_blocks(long long nk) {
  long long b = (nk + kThreads - 1) / kThreads;
  if (b > 1024) b = 1024;
  return (unsigned)(b < 1 ? 1 : b);
}

int launch_fill(double* p, long long n, double v, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T> int launch_diag(int dim, const DiagArgs& a, int nmem, hipStream_t s) {
  int rc = launch_fill(a.out, 2LL * nmem, 0.0, s);
  if (rc) return rc;
  dim3 grid(red_blocks(a.nk), (unsigned)nmem);
  if (dim == 3) hipLaunchKernelGGL((diag_kernel<T, 3>), grid, dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL((diag_kernel<T, 2>), grid, dim3(kThreads), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T> int launch_spectrum(int dim, const DiagArgs& a, int nmem, int nshell, hipStream_t s) {
  int rc = launch_fill(a.out, (long long)nshell * nmem, 0.0, s);
  if (rc) return rc;
  dim3 grid(red_blocks(a.nk), (unsigned)nmem);
  if (dim == 3) hipLaunchKernelGGL((spectrum_kernel<T, 3>), grid, dim3(kThreads), 0, s, a, nshell);
  else hipLaunchKernelGGL((spectrum_kernel<T, 2>), grid, dim3(kThreads), 0, s, a, nshell);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T> int launch_band(int dim, const BandArgs& a, int nmem, hipStream_t s) {
  if (dim == 3) hipLaunchKernelGGL((band_kernel<T, 3>), dim3((unsigned)nmem), dim3(kThreads), 0, s, a);
  else hipLaunchKernelGGL((band_kernel<T, 2>), dim3((unsigned)nmem), dim3(kThreads), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T> int launch_sk(int dim, const void* u, double* acc, long long nk, int nmem, hipStream_t s) {
  dim3 grid((unsigned)((nk + kThreads - 1) / kThreads));
  if (dim == 3) hipLaunchKernelGGL((sk_kernel<T, 3>), grid, dim3(kThreads), 0, s, u, acc, nk, nmem);
  else hipLaunchKernelGGL((sk_kernel<T, 2>), grid, dim3(kThreads), 0, s, u, acc, nk, nmem);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

template <typename T>
int launch_draw(void* out, const unsigned long long* seeds, const long long* msteps, long long nk, int nc, int nmem,
                hipStream_t s) {
  dim3 grid((unsigned)((nk + 255) / 256), (unsigned)nmem);
  hipLaunchKernelGGL((draw_kernel<T>), grid, dim3(256), 0, s, out, seeds, msteps, nk, nc);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_member_step(long long* msteps, int n, hipStream_t s, long long by) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(member_step_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, msteps, n, by);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// Members per thread of the register-resident Stokes kernel: enough blocks
// along the members that the launch fills the GPU with about 8 waves per
// SIMD (256 CUs x 32 waves), and no more (each thread keeps its mode's
// wavenumbers and linear operator for all of its members).
template <typename T> int launch_stokes_steps(int dim, const StageArgs& a, int nsteps, double* acc, hipStream_t s) {
  const long long blocks_x = (a.nk + kThreads - 1) / kThreads;
  const long long want = 256LL * 32 * 64 / kThreads;  // blocks for 32 waves on each of 256 CUs
  long long mb = blocks_x * a.n_members / want;
  const int MBv = mb >= 8 ? 8 : (mb >= 4 ? 4 : (mb >= 2 ? 2 : 1));
  const dim3 grid((unsigned)blocks_x, (unsigned)((a.n_members + MBv - 1) / MBv));
  // the noise type as a compile-time choice (1 thermal, 2 correlated)
#define PF_SK(D, M)                                                                                       \
  do {                                                                                                   \
    if (a.noise_type == 0)                                                                               \
      hipLaunchKernelGGL((stokes_steps_kernel<T, D, M, 1>), grid, dim3(kThreads), 0, s, a, nsteps, acc); \
    else                                                                                                 \
      hipLaunchKernelGGL((stokes_steps_kernel<T, D, M, 2>), grid, dim3(kThreads), 0, s, a, nsteps, acc); \
  } while (0)
  if (dim == 3) {
    switch (MBv) { case 8: PF_SK(3, 8); break; case 4: PF_SK(3, 4); break; case 2: PF_SK(3, 2); break; default: PF_SK(3, 1); }
  } else {
    switch (MBv) { case 8: PF_SK(2, 8); break; case 4: PF_SK(2, 4); break; case 2: PF_SK(2, 2); break; default: PF_SK(2, 1); }
  }
#undef PF_SK
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

#define PF_INST(T)                                                                                    \
  template int launch_stage<T>(int, const StageArgs&, hipStream_t, int);                              \
  template int launch_project<T>(int, const StageArgs&, const void*, void*, int, int, int, hipStream_t); \
  template int launch_diag<T>(int, const DiagArgs&, int, hipStream_t);                                \
  template int launch_spectrum<T>(int, const DiagArgs&, int, int, hipStream_t);                       \
  template int launch_band<T>(int, const BandArgs&, int, hipStream_t);                                \
  template int launch_sk<T>(int, const void*, double*, long long, int, hipStream_t);                  \
  template int launch_draw<T>(void*, const unsigned long long*, const long long*, long long, int, int, hipStream_t); \
  template int launch_stokes_steps<T>(int, const StageArgs&, int, double*, hipStream_t);
PF_INST(float)
PF_INST(double)
#undef PF_INST

}  // namespace pf
