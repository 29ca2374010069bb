// Paired launches (two plans of different transform lengths in one grid) of
// the 3D nonlinear passes, FP32, at the default geometry with any of its
// cache policies (launch variants 0, 6 and 7 of column passes, 0 and 4 of
// row passes).  Only the length pairs listed in PF_PAIRS are compiled; other
// pairs launch back to back.
#include "pf_launch.h"

// (larger, smaller) transform lengths of ensembles that mix grid sizes:
// 128^3 + 96^3 under the 3/2 rule (192, 144) and the 2/3 rule (128, 96),
// 96^3 + 64^3 under the 3/2 rule (144, 96), 256^3 + 192^3 under 2/3.
#define PF_PAIRS(X) X(192, 144) X(128, 96) X(144, 96) X(256, 192)

namespace pf {

namespace {

template <typename T, int L, int KIND> long long col_blocks(ColArgs& a) {
  constexpr int TC = col_width_v<T, KIND>(L, 0);
  ColGeom& g = a.g;
  g.n_tiles = (g.cols + TC - 1) / TC;
  const long long nfg = (KIND == K_FWD_STAGE3 || KIND == K_FWD_STAGE2) ? 1 : a.n_fields;
  return nfg * g.n_tiles * (long long)g.n_outer * a.n_members;
}

template <typename T, int L0, int L1, int KIND, int CP>
int col_pair_k(ColArgs a0, const void* tw0, ColArgs a1, const void* tw1, hipStream_t s) {
  constexpr int TC0 = col_width_v<T, KIND>(L0, 0), TC1 = col_width_v<T, KIND>(L1, 0);
  if constexpr (TC0 == 0 || TC1 == 0) {
    return -4;
  } else {
    const long long n0 = col_blocks<T, L0, KIND>(a0), n1 = col_blocks<T, L1, KIND>(a1);
    const ColSeg s0{a0, tw0, n0}, s1{a1, tw1, n1};
    const long long split = (s0.nblk + 7) / 8 * 8;
    const long long G = split + s1.nblk;
    if (s0.nblk <= 0 || s1.nblk <= 0 || G > 0x7fffffffLL) return -1;
    hipLaunchKernelGGL((col_pass_pair<T, L0, TC0, L1, TC1, KIND, CP>), dim3((unsigned)G), dim3(kThreads), 0, s, s0,
                       s1, split);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
}

template <typename T, int L0, int L1, int CP>
int col_pair_c(int kind, const ColArgs& a0, const void* tw0, const ColArgs& a1, const void* tw1, hipStream_t s) {
  switch (kind) {
    case K_INV: return col_pair_k<T, L0, L1, K_INV, CP>(a0, tw0, a1, tw1, s);
    case K_FWD: return col_pair_k<T, L0, L1, K_FWD, CP>(a0, tw0, a1, tw1, s);
    case K_FWD_STAGE3: return col_pair_k<T, L0, L1, K_FWD_STAGE3, CP>(a0, tw0, a1, tw1, s);
    case K_INV_UC3: return col_pair_k<T, L0, L1, K_INV_UC3, CP>(a0, tw0, a1, tw1, s);
    case K_INV_CURL3: return col_pair_k<T, L0, L1, K_INV_CURL3, CP>(a0, tw0, a1, tw1, s);
    default: return -4;
  }
}

// v: the launch variant both records use (0, 6 or 7)
template <typename T, int L0, int L1>
int col_pair_l(int kind, const ColArgs& a0, const void* tw0, const ColArgs& a1, const void* tw1, hipStream_t s,
               int v) {
  switch (v) {
    case 0: return col_pair_c<T, L0, L1, col_cache_policy(0)>(kind, a0, tw0, a1, tw1, s);
    case 6: return col_pair_c<T, L0, L1, col_cache_policy(6)>(kind, a0, tw0, a1, tw1, s);
    case 7: return col_pair_c<T, L0, L1, col_cache_policy(7)>(kind, a0, tw0, a1, tw1, s);
    default: return -4;
  }
}

template <typename T, int L0, int L1, int CP>
int row_pair_c(int mode, const RowArgs& a0, const void* tw0, const RowArgs& a1, const void* tw1, hipStream_t s) {
  if (mode != ROW_FUSED3) return -4;
  constexpr int NP0 = row_pairs_v<T, ROW_FUSED3>(L0, 0), NP1 = row_pairs_v<T, ROW_FUSED3>(L1, 0);
  if constexpr (NP0 == 0 || NP1 == 0) {
    return -4;
  } else {
    if (a0.n_rows % 2 || a1.n_rows % 2) return -1;
    RowSeg s0{a0, tw0, (a0.n_rows / 2 + NP0 - 1) / NP0, a0.n_members};
    RowSeg s1{a1, tw1, (a1.n_rows / 2 + NP1 - 1) / NP1, a1.n_members};
    const long long n0 = (long long)s0.gx * s0.gy, n1 = (long long)s1.gx * s1.gy;
    const long long split = (n0 + 7) / 8 * 8;
    if (n0 <= 0 || n1 <= 0 || split + n1 > 0x7fffffffLL) return -1;
    hipLaunchKernelGGL((row_pass_pair<T, L0, NP0, L1, NP1, ROW_FUSED3, CP>), dim3((unsigned)(split + n1)),
                       dim3(kThreads), 0, s, s0, s1, split);
    return hipGetLastError() == hipSuccess ? 0 : -3;
  }
}

// v: the launch variant both records use (0 or 4)
template <typename T, int L0, int L1>
int row_pair_l(int mode, const RowArgs& a0, const void* tw0, const RowArgs& a1, const void* tw1, hipStream_t s,
               int v) {
  switch (v) {
    case 0: return row_pair_c<T, L0, L1, row_cache_policy(0)>(mode, a0, tw0, a1, tw1, s);
    case 4: return row_pair_c<T, L0, L1, row_cache_policy(4)>(mode, a0, tw0, a1, tw1, s);
    default: return -4;
  }
}

}  // namespace

bool pair_variant(bool row, int v) { return v == 0 || (row ? v == 4 : (v == 6 || v == 7)); }

bool pair_compiled(int L0, int L1) {
#define PF_P(a, b) if ((L0 == a && L1 == b) || (L0 == b && L1 == a)) return true;
  PF_PAIRS(PF_P)
#undef PF_P
  return false;
}

template <>
int launch_col_pair<float>(int kind, int L0, const ColArgs& a0, const void* tw0, int L1, const ColArgs& a1,
                           const void* tw1, hipStream_t s, int v) {
#define PF_P(a, b)                                                      \
  if (L0 == a && L1 == b) return col_pair_l<float, a, b>(kind, a0, tw0, a1, tw1, s, v); \
  if (L0 == b && L1 == a) return col_pair_l<float, a, b>(kind, a1, tw1, a0, tw0, s, v);
  PF_PAIRS(PF_P)
#undef PF_P
  return -4;
}

template <>
int launch_row_pair<float>(int mode, int L0, const RowArgs& a0, const void* tw0, int L1, const RowArgs& a1,
                           const void* tw1, hipStream_t s, int v) {
#define PF_P(a, b)                                                      \
  if (L0 == a && L1 == b) return row_pair_l<float, a, b>(mode, a0, tw0, a1, tw1, s, v); \
  if (L0 == b && L1 == a) return row_pair_l<float, a, b>(mode, a1, tw1, a0, tw0, s, v);
  PF_PAIRS(PF_P)
#undef PF_P
  return -4;
}

template <>
int launch_col_pair<double>(int, int, const ColArgs&, const void*, int, const ColArgs&, const void*, hipStream_t,
                            int) {
  return -4;
}
template <>
int launch_row_pair<double>(int, int, const RowArgs&, const void*, int, const RowArgs&, const void*, hipStream_t,
                            int) {
  return -4;
}

}  // namespace pf
