// Host-side launchers (one translation unit per kernel family and precision).
#pragma once
#include <hip/hip_ext.h>

#include "pf_kernels.h"

// Transform lengths compiled into the library: N and 3N/2 for N in
// {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 512, 1024} and for N in
// {20, 40, 80, 160, 320}, and N alone (the 2/3 rule) for N in {30, 60, 120,
// 240, 480}.
#define PF_LENGTHS(X) \
  X(8) X(12) X(16) X(18) X(20) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(64) X(72) X(80) X(96) X(120) \
  X(128) X(144) X(160) X(192) X(240) X(256) X(288) X(320) X(384) X(480) X(512) X(768) X(1024) X(1536)

namespace pf {

bool length_supported(int L);

// Events that the next pass kernel launch records itself: set by the pass
// timer around a pass that is one kernel, taken by that launch, so that the
// timing brackets exactly the kernel's execution on the device (events
// recorded as separate stream commands add a few microseconds per launch).
struct PendingEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
PendingEvents& pending_events();

// A pass kernel launch, recording the pending events if there are any.
template <typename F, typename... Args>
inline void launch_pass_kernel(F kernel, dim3 grid, dim3 block, hipStream_t s, Args... args) {
  PendingEvents& pe = pending_events();
  const hipEvent_t a = pe.start, b = pe.stop;
  pe.start = pe.stop = nullptr;
  hipExtLaunchKernelGGL(kernel, grid, block, 0, s, a, b, 0, args...);
}

// Launch variants: V = 0 is the default tile (column passes) or row-pair
// count (row passes); FP32 also builds V = 1 (half), V = 2 (twice) and, for
// column passes, V = 3 (two tiles per block, the second prefetched), V = 4
// (stage kinds: 1024-thread blocks of twice the width), V = 5 (radix plan
// 1), V = 6 (the default tile with nontemporal stores) and V = 7 (nontemporal
// stores and loads); for row passes V = 3 is radix plan 1 and V = 4 the
// default with nontemporal stores.  These are the choices the autotuner
// times.  One translation unit per precision and V.
constexpr int kVariants = 8;
// the radix plan (make_plan) a launch variant uses
constexpr int col_radix_plan(int v) { return v == 5 ? 1 : 0; }
constexpr int row_radix_plan(int v) { return v == 3 ? 1 : 0; }
// the cache policy (cp_store / cp_load) of a launch variant
constexpr int col_cache_policy(int v) { return v == 6 ? 1 : (v == 7 ? 3 : 0); }
constexpr int row_cache_policy(int v) { return v == 4 ? 1 : 0; }
template <typename T, int V> int launch_col_v(int L, int kind, const ColArgs& a, const void* tw, hipStream_t s);
template <typename T, int V> int col_width_q(int L, int kind);
template <typename T, int V> int launch_row_v(int L, int mode, const RowArgs& a, const void* tw, hipStream_t s);
template <typename T, int V> int row_pairs_q(int L, int mode);

template <typename T> constexpr bool has_variants() { return sizeof(T) == 4; }

// Launch one column pass (FFT along a strided axis) of the given ColKind.
template <typename T>
int launch_col(int L, int kind, const ColArgs& a, const void* tw, hipStream_t s, int v) {
  if constexpr (has_variants<T>()) {
    switch (v) {
      case 0: return launch_col_v<T, 0>(L, kind, a, tw, s);
      case 1: return launch_col_v<T, 1>(L, kind, a, tw, s);
      case 2: return launch_col_v<T, 2>(L, kind, a, tw, s);
      case 3: return launch_col_v<T, 3>(L, kind, a, tw, s);
      case 4: return launch_col_v<T, 4>(L, kind, a, tw, s);
      case 5: return launch_col_v<T, 5>(L, kind, a, tw, s);
      case 6: return launch_col_v<T, 6>(L, kind, a, tw, s);
      case 7: return launch_col_v<T, 7>(L, kind, a, tw, s);
      default: return -1;
    }
  } else {
    return v == 0 ? launch_col_v<T, 0>(L, kind, a, tw, s) : -4;
  }
}
// Tile width of a column pass kind at length L and variant v (0: unavailable).
template <typename T> int tile_cols(int L, int kind, int v) {
  if constexpr (has_variants<T>()) {
    switch (v) {
      case 0: return col_width_q<T, 0>(L, kind);
      case 1: return col_width_q<T, 1>(L, kind);
      case 2: return col_width_q<T, 2>(L, kind);
      case 3: return col_width_q<T, 3>(L, kind);
      case 4: return col_width_q<T, 4>(L, kind);
      case 5: return col_width_q<T, 5>(L, kind);
      case 6: return col_width_q<T, 6>(L, kind);
      case 7: return col_width_q<T, 7>(L, kind);
      default: return 0;
    }
  } else {
    return v == 0 ? col_width_q<T, 0>(L, kind) : 0;
  }
}
// Launch one row pass (last axis) of the given RowMode.
template <typename T>
int launch_row(int L, int mode, const RowArgs& a, const void* tw, hipStream_t s, int v) {
  if constexpr (has_variants<T>()) {
    switch (v) {
      case 0: return launch_row_v<T, 0>(L, mode, a, tw, s);
      case 1: return launch_row_v<T, 1>(L, mode, a, tw, s);
      case 2: return launch_row_v<T, 2>(L, mode, a, tw, s);
      case 3: return launch_row_v<T, 3>(L, mode, a, tw, s);
      case 4: return launch_row_v<T, 4>(L, mode, a, tw, s);
      case 5:
      case 6:
      case 7: return -4;  // row passes have no such variants
      default: return -1;
    }
  } else {
    return v == 0 ? launch_row_v<T, 0>(L, mode, a, tw, s) : -4;
  }
}
// Row pairs per block of a row pass at length L and variant v (0: unavailable).
template <typename T> int row_pair_count(int L, int mode, int v) {
  if constexpr (has_variants<T>()) {
    switch (v) {
      case 0: return row_pairs_q<T, 0>(L, mode);
      case 1: return row_pairs_q<T, 1>(L, mode);
      case 2: return row_pairs_q<T, 2>(L, mode);
      case 3: return row_pairs_q<T, 3>(L, mode);
      case 4: return row_pairs_q<T, 4>(L, mode);
      default: return 0;
    }
  } else {
    return v == 0 ? row_pairs_q<T, 0>(L, mode) : 0;
  }
}

// Paired launches of two plans' passes (pf_pair.hip): FP32, both records at
// launch variant v, which pair_variant admits (the default geometry with any
// cache policy), compiled length pairs only; -4 where a pair is not compiled.
bool pair_compiled(int L0, int L1);
bool pair_variant(bool row, int v);
template <typename T>
int launch_col_pair(int kind, int L0, const ColArgs& a0, const void* tw0, int L1, const ColArgs& a1, const void* tw1,
                    hipStream_t s, int v);
template <typename T>
int launch_row_pair(int mode, int L0, const RowArgs& a0, const void* tw0, int L1, const RowArgs& a1, const void* tw1,
                    hipStream_t s, int v);

template <typename T> int launch_stage(int dim, const StageArgs& a, hipStream_t s, int cp = 0);
template <typename T> int launch_diag(int dim, const DiagArgs& a, int n_members, hipStream_t s);
template <typename T> int launch_spectrum(int dim, const DiagArgs& a, int n_members, int nshell, hipStream_t s);
template <typename T> int launch_band(int dim, const BandArgs& a, int n_members, hipStream_t s);
template <typename T> int launch_sk(int dim, const void* u, double* acc, long long nk, int nmem, hipStream_t s);
template <typename T>
int launch_project(int dim, const StageArgs& geo, const void* in, void* out, int mask_nyq, int zero_k0,
                   int n_members, hipStream_t s);
int launch_fill(double* p, long long n, double v, hipStream_t s);
template <typename T>
int launch_draw(void* out, const unsigned long long* seeds, const long long* msteps, long long nk, int nc, int nmem,
                hipStream_t s);
int launch_member_step(long long* msteps, int n, hipStream_t s, long long by = 1);
template <typename T> int launch_stokes_steps(int dim, const StageArgs& a, int nsteps, double* acc, hipStream_t s);

}  // namespace pf
