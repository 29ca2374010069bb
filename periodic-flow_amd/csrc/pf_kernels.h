// Kernel templates of the Fourier-Galerkin right-hand side and time step.
//
// HBM layout (C order, complex interleaved, one block per member):
//   state u_hat         (B, d, N0, [N1,] nh)      nh = N_last/2 + 1
//   3D nonlinear passes (per member; rows of A, B, C, D padded to the pitch
//   P >= nh, a multiple of 128 bytes):
//     A  (5, M0, N1, P)    axis-0 inverse of u_hat, omega_y, omega_z         <- col_pass UC3
//     B  (6, M0, M1, P)    axis-1 inverse, velocity and vorticity           <- col_pass CURL3
//     C  (3, M0, M1, P)    c2r, u x omega, r2c, cut     <- row_pass FUSED3
//     D  (3, M0, N1, P)    axis-1 forward, truncated    <- col_pass (planes 128 B apart)
//     R  (3, N0, N1, nh)   axis-0 forward, truncated    <- col_pass (or fused stage)
//   2D: A (3, M0, P) -> C (2, M0, P) -> R (2, N0, nh)
// The axis-0 forward pass enumerates D's N1 rows of nh entries back to back
// and maps each column to D's row (pitch P) and to the state's (pitch nh).
// The padded physical fields never reach HBM: the last axis is taken to
// physical space, multiplied and brought back inside one block.
#pragma once
#include "pf_fft.h"

namespace pf {

constexpr int kThreads = 256;

// ------------------------------------------------------------------- row pass
enum RowMode { ROW_FUSED3 = 0, ROW_FUSED2 = 1, ROW_C2R = 2, ROW_R2C = 3 };

struct RowArgs {
  const void* in;
  void* out;
  int n_rows;                   // rows per member and field (even)
  int N;                        // unpadded length of the last axis
  int nh;                       // N/2 + 1 stored entries
  long long in_row, out_row;    // element strides between rows
  long long in_field, out_field;
  long long in_member, out_member;
  int n_fields;                 // fields handled through grid.y (C2R / R2C)
  int n_members;
  int mask23;                   // 2/3 rule mask on the spectral input
  double scale;                 // forward normalisation
  double out_scale;             // scale of the physical output (C2R)
};

template <int MODE> constexpr int row_slots() {
  return MODE == ROW_FUSED3 ? 6 : (MODE == ROW_FUSED2 ? 3 : 1);
}
template <typename T, int MODE> constexpr int row_pairs(int L) {
  int p = (sizeof(T) == 4 ? 4096 : 2048) / (row_slots<MODE>() * L);
  return p < 1 ? 1 : (p > 8 ? 8 : p);
}
__host__ __device__ constexpr int row_pitch(int L) { return L + L / 16; }
// Row entry e of a slot sits at e + e / row_pad_div(L): one skipped entry
// every 16 for power-of-two lengths (the radix-16/8 strides then fall on
// distinct banks), every 48 where L has a factor 3, whose radix-9/12/24
// first pass writes with stride 9, 12 or 24, none at length 144, and every
// 80 where L has a factor 5 (radix-5 first pass).  Chosen by counting bank
// conflicts of every LDS access of the row pass (64 lanes in groups of 16
// over 32 banks and of 32 over 64) and by timing; at lengths 96, 192, 288
// and 384 this halves them or better, e.g. 2.4 to 0.7 conflict cycles per
// access at 192.  At 144 (plan 9 x 16) no padding measured 2-4% faster than
// one skipped entry every 96 or 32.
__host__ __device__ constexpr int row_pad_div(int L) {
  return L % 5 == 0 ? 80 : (L % 3 ? 16 : (L == 144 ? L : 48));
}

// Hermitian extension of a stored half spectrum to index k of a length-L
// complex transform.  DC keeps its real part only; the stored Nyquist entry
// N/2 keeps its real part (halved when L > N, i.e. 3/2 padding); entries
// beyond N/2 are zero; the 2/3 rule zeroes k > N/3.
template <typename T>
__device__ __forceinline__ cplx<T> hext(const cplx<T>* row, int k, int L, int N, int mask23) {
  const bool mirror = k > L / 2;
  const int kk = mirror ? L - k : k;
  cplx<T> v = mk<T>(0, 0);
  if (kk <= N / 2 && (!mask23 || kk <= N / 3)) {
    v = row[kk];
    if (kk == 0) v.y = 0;
    if (kk == N / 2) { v.y = 0; if (L != N) v.x *= (T)0.5; }
  }
  if (mirror) v.y = -v.y;
  return v;
}

// Z = X + i Y
template <typename T> __device__ __forceinline__ cplx<T> pack(cplx<T> x, cplx<T> y) {
  return mk<T>(x.x - y.y, x.y + y.x);
}

// Row-pair variants of a row pass (the autotuner's choices): V = 0 the
// default count, 1 half of it, 2 twice it, 3 the default count with radix
// plan 1 (make_plan) where L has one, 4 the default count with nontemporal
// stores; 0 where a variant repeats another or does not fit in LDS.
template <typename T, int MODE> constexpr int row_pairs_v(int L, int V) {
  const int np = row_pairs<T, MODE>(L);
  if (V == 3) return has_alt_plan(L) ? np : 0;
  if (V == 4) return np;
  const int r = V == 0 ? np : (V == 1 ? np / 2 : np * 2);
  if (r < 1 || (V != 0 && r == np)) return 0;
  const long long lds = ((long long)r * row_slots<MODE>() * row_pitch(L) + L) * (long long)sizeof(cplx<T>);
  return lds <= 160 * 1024 ? r : 0;
}

// LDS of a row pass block: the twiddle table, then NP pairs of NS slots
template <typename T, int L, int MODE, int NP, int RP = 0> constexpr int row_lds_tw() { return tw_alloc(L, RP); }
template <typename T, int L, int MODE, int NP> constexpr int row_lds_buf() {
  return NP * row_slots<MODE>() * row_pitch(L);
}

// One block of a row pass: block (bx, by) of its launch; tws and buf are the
// block's LDS (row_lds_tw and row_lds_buf entries).
template <typename T, int L, int MODE, int NP, int RP = 0, int CP = 0>
__device__ __forceinline__ void row_body(const RowArgs& a, const cplx<T>* __restrict__ tw, cplx<T>* tws,
                                         cplx<T>* buf, int bx, int by) {
  constexpr int NS = row_slots<MODE>();
  constexpr int LP = row_pitch(L);
  // the table size as a constant (evaluated at run time, the plan's radix
  // array would live in private memory)
  constexpr int TWN = tw_size(L, RP);
  for (int m = threadIdx.x; m < TWN; m += kThreads) tws[m] = tw[m];

  const int npairs = a.n_rows / 2;
  const int pair0 = bx * NP;
  int f = 0, b;
  if (MODE == ROW_C2R || MODE == ROW_R2C) { f = by % a.n_fields; b = by / a.n_fields; }
  else b = by;
  constexpr int PD = row_pad_div(L);
  static_assert(L - 1 + (L - 1) / PD < LP, "row pitch too small for the padding");
  // each row transform within one wavefront for the whole FFT: its passes
  // synchronise the wave only
  constexpr bool WL = wave_local_plan(L, RP);
  auto ix = [](int slot, int e) { return slot * LP + e + e / PD; };
  const int N = a.N;

  // ---- load (every load of a chunk in flight before the LDS stores)
  if constexpr (MODE == ROW_R2C) {
    const T* in = (const T*)a.in + (long long)b * a.in_member + (long long)f * a.in_field;
    constexpr int NIT = (NP * L + kThreads - 1) / kThreads;
    constexpr int CH = NIT < 8 ? NIT : 8;
    for (int base = 0; base < NIT; base += CH) {
      T xr[CH], yr[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int t = threadIdx.x + (base + q) * kThreads;
        xr[q] = yr[q] = (T)0;
        if (base + q >= NIT || t >= NP * L) continue;
        const int p = t / L, x = t % L;
        const int pr = pair0 + p;
        if (pr < npairs) {
          const long long r0 = 2LL * pr;
          xr[q] = in[r0 * a.in_row + x];
          yr[q] = in[(r0 + 1) * a.in_row + x];
        }
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int t = threadIdx.x + (base + q) * kThreads;
        if (base + q >= NIT || t >= NP * L) continue;
        buf[ix(t / L, t % L)] = mk<T>(xr[q], yr[q]);
      }
    }
  } else {
    // stored entries kk = 0..L/2 of each packed pair (X, Y); Z_kk = X + iY and
    // Z_{L-kk} = conj(X) + i conj(Y) (Hermitian extension).  TPS threads
    // own each LDS slot (pair, packed field pair) and step through its
    // entries kk = i, i + TPS, ...: the row pointers are set once per thread.
    const cplx<T>* in = (const cplx<T>*)a.in + (long long)b * a.in_member;
    constexpr int HK = L / 2 + 1;
    constexpr int NSL = NP * NS;
    constexpr int TPS = kThreads / NSL;
    static_assert(TPS >= 1, "more row slots than threads");
    constexpr int NIT = (HK + TPS - 1) / TPS;
    constexpr int CH = NIT < 8 ? NIT : 8;
    const int kmax = a.mask23 ? N / 3 : N / 2;
    const int slot = threadIdx.x / TPS, i0 = threadIdx.x % TPS;
    const bool active = slot < NSL;
    const int p = slot / NS, sl = slot % NS;
    const int pr = pair0 + p;
    const bool rows = active && pr < npairs;
    const cplx<T>* px = in;
    const cplx<T>* py = in;
    if (rows) {
      const long long r0 = 2LL * pr, r1 = r0 + 1;
      if constexpr (MODE == ROW_FUSED3) {
        const int rr = sl / 3, qq = sl % 3;
        const long long r = r0 + rr;
        px = in + (2 * qq) * a.in_field + r * a.in_row;
        py = in + (2 * qq + 1) * a.in_field + r * a.in_row;
      } else if constexpr (MODE == ROW_FUSED2) {
        if (sl == 0) { px = in + r0 * a.in_row; py = in + a.in_field + r0 * a.in_row; }
        else if (sl == 1) { px = in + 2 * a.in_field + r0 * a.in_row; py = in + 2 * a.in_field + r1 * a.in_row; }
        else { px = in + r1 * a.in_row; py = in + a.in_field + r1 * a.in_row; }
      } else {  // ROW_C2R
        const cplx<T>* fin = in + (long long)f * a.in_field;
        px = fin + r0 * a.in_row;
        py = fin + r1 * a.in_row;
      }
    }
    for (int base = 0; base < NIT; base += CH) {
      cplx<T> X[CH], Y[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int kk = i0 + (base + q) * TPS;
        X[q] = Y[q] = mk<T>(0, 0);
        if (!rows || base + q >= NIT || kk >= HK || kk > kmax) continue;
        X[q] = ld_stream<cp_load(CP), T>(px + kk);
        Y[q] = ld_stream<cp_load(CP), T>(py + kk);
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int kk = i0 + (base + q) * TPS;
        if (!active || base + q >= NIT || kk >= HK) continue;
        cplx<T> x = X[q], y = Y[q];
        if (kk == 0) { x.y = 0; y.y = 0; }
        if (kk == N / 2) {
          x.y = 0;
          y.y = 0;
          if (L != N) { x.x *= (T)0.5; y.x *= (T)0.5; }
        }
        buf[ix(slot, kk)] = pack<T>(x, y);
        if (kk > 0 && kk < L / 2) buf[ix(slot, L - kk)] = pack<T>(conj<T>(x), conj<T>(y));
      }
    }
  }
  __syncthreads();

  if constexpr (MODE == ROW_R2C) {
    lds_fft<T, L, -1, NP, kThreads, false, RP, WL>(buf, tws, ix);
  } else {
    lds_fft<T, L, +1, NP * NS, kThreads, false, RP, WL>(buf, tws, ix);
  }

  if constexpr (MODE == ROW_C2R) {
    T* out = (T*)a.out + (long long)b * a.out_member + (long long)f * a.out_field;
    const T os = (T)a.out_scale;
    for (int t = threadIdx.x; t < NP * L; t += kThreads) {
      const int p = t / L, x = t % L;
      const int pr = pair0 + p;
      if (pr >= npairs) continue;
      const long long r0 = 2LL * pr;
      const cplx<T> z = buf[ix(p, x)];
      out[r0 * a.out_row + x] = z.x * os;
      out[(r0 + 1) * a.out_row + x] = z.y * os;
    }
    return;
  }

  if constexpr (MODE == ROW_FUSED3) {
    // u x omega at every physical point of both rows of each pair
    for (int t = threadIdx.x; t < NP * L; t += kThreads) {
      const int p = t / L, x = t % L;
      T c[2][3];
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        const cplx<T> z0 = buf[ix(p * 6 + rr * 3 + 0, x)];
        const cplx<T> z1 = buf[ix(p * 6 + rr * 3 + 1, x)];
        const cplx<T> z2 = buf[ix(p * 6 + rr * 3 + 2, x)];
        const T u0 = z0.x, u1 = z0.y, u2 = z1.x, w0 = z1.y, w1 = z2.x, w2 = z2.y;
        c[rr][0] = u1 * w2 - u2 * w1;
        c[rr][1] = u2 * w0 - u0 * w2;
        c[rr][2] = u0 * w1 - u1 * w0;
      }
      buf[ix(p * 6 + 0, x)] = mk<T>(c[0][0], c[0][1]);
      buf[ix(p * 6 + 3, x)] = mk<T>(c[1][0], c[1][1]);
      buf[ix(p * 6 + 1, x)] = mk<T>(c[0][2], c[1][2]);
    }
    __syncthreads();
    auto fx = [](int ff, int e) {
      const int p = ff / 3, w = ff % 3;
      const int slot = p * 6 + (w == 0 ? 0 : (w == 1 ? 3 : 1));
      return slot * LP + e + e / PD;
    };
    lds_fft<T, L, -1, NP * 3, kThreads, false, RP, WL>(buf, tws, fx);
  } else if constexpr (MODE == ROW_FUSED2) {
    for (int t = threadIdx.x; t < NP * L; t += kThreads) {
      const int p = t / L, x = t % L;
      const cplx<T> za = buf[ix(p * 3 + 0, x)];
      const cplx<T> zw = buf[ix(p * 3 + 1, x)];
      const cplx<T> zb = buf[ix(p * 3 + 2, x)];
      // 2D form of math.py:13-17: (u1 w, -u0 w)
      buf[ix(p * 3 + 0, x)] = mk<T>(za.y * zw.x, -za.x * zw.x);
      buf[ix(p * 3 + 2, x)] = mk<T>(zb.y * zw.y, -zb.x * zw.y);
    }
    __syncthreads();
    auto fx = [](int ff, int e) {
      const int p = ff / 2, w = ff % 2;
      return (p * 3 + 2 * w) * LP + e + e / PD;
    };
    lds_fft<T, L, -1, NP * 2, kThreads, false, RP, WL>(buf, tws, fx);
  }

  // ---- split the packed spectra, truncate, scale, store.  TPO threads own
  // each output slot (pair, output field pair) and step through k.
  constexpr int NFW = MODE == ROW_FUSED3 ? 3 : (MODE == ROW_FUSED2 ? 2 : 1);
  constexpr int NOS = NP * NFW;
  constexpr int TPO = kThreads / NOS;
  constexpr int NKO = (L / 2 + 1 + TPO - 1) / TPO;  // nh <= L/2 + 1
  const int nh = a.nh;
  const T s = (T)a.scale;
  const int osl = threadIdx.x / TPO, k0 = threadIdx.x % TPO;
  const int op = osl / NFW, w = osl % NFW;
  const int opr = pair0 + op;
  if (osl >= NOS || opr >= npairs) return;
  int slot;
  if (MODE == ROW_FUSED3) slot = op * 6 + (w == 0 ? 0 : (w == 1 ? 3 : 1));
  else if (MODE == ROW_FUSED2) slot = op * 3 + 2 * w;
  else slot = op;
  const long long r0 = 2LL * opr, r1 = r0 + 1;
  cplx<T>* out = (cplx<T>*)a.out + (long long)b * a.out_member;
  cplx<T>* oa;
  cplx<T>* ob;
  if constexpr (MODE == ROW_FUSED3) {
    if (w == 0) { oa = out + r0 * a.out_row; ob = out + a.out_field + r0 * a.out_row; }
    else if (w == 1) { oa = out + r1 * a.out_row; ob = out + a.out_field + r1 * a.out_row; }
    else { oa = out + 2 * a.out_field + r0 * a.out_row; ob = out + 2 * a.out_field + r1 * a.out_row; }
  } else if constexpr (MODE == ROW_FUSED2) {
    const long long r = w == 0 ? r0 : r1;
    oa = out + r * a.out_row;
    ob = out + a.out_field + r * a.out_row;
  } else {
    cplx<T>* fo = out + (long long)f * a.out_field;
    oa = fo + r0 * a.out_row;
    ob = fo + r1 * a.out_row;
  }
#pragma unroll
  for (int q = 0; q < NKO; ++q) {
    const int k = k0 + q * TPO;
    if (k >= nh) break;
    const cplx<T> yk = buf[ix(slot, k)];
    const cplx<T> ym = buf[ix(slot, k == 0 ? 0 : L - k)];
    cplx<T> A = mk<T>((T)0.5 * (yk.x + ym.x), (T)0.5 * (yk.y - ym.y));
    cplx<T> B = mk<T>((T)0.5 * (yk.y + ym.y), (T)-0.5 * (yk.x - ym.x));
    if (L != N && k == N / 2) { A = mk<T>(2 * A.x, 0); B = mk<T>(2 * B.x, 0); }
    st_stream<cp_store(CP), T>(oa + k, scl<T>(A, s));
    st_stream<cp_store(CP), T>(ob + k, scl<T>(B, s));
  }
}

template <typename T, int L, int MODE, int NP, int RP = 0, int CP = 0>
__global__ void __launch_bounds__(kThreads) row_pass(RowArgs a, const cplx<T>* __restrict__ tw) {
  __shared__ cplx<T> tws[row_lds_tw<T, L, MODE, NP, RP>()];
  __shared__ cplx<T> buf[row_lds_buf<T, L, MODE, NP>()];
  row_body<T, L, MODE, NP, RP, CP>(a, tw, tws, buf, blockIdx.x, blockIdx.y);
}

// -------------------------------------------------------------- random numbers
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the Box-Muller map.
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64-bit multiply (v_mad_u64_u32) gives both halves
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// u1 in (0, 1], u2 in [0, 1]: g0 = r cos(2 pi u2), g1 = r sin(2 pi u2),
// r = sqrt(-2 ln u1).  FP32 uses the hardware log2 and the hardware sine
// and cosine, whose argument is in revolutions (v_sin_f32(x) = sin(2 pi x)),
// so the angle needs no range reduction; FP64 uses the library functions.
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& g0, float& g1) {
  const float u1 = ((float)a + 0.5f) * 2.3283064365386963e-10f;
  const float u2 = (float)b * 2.3283064365386963e-10f;
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));  // -2 ln 2 log2 u1
  g0 = r * __builtin_amdgcn_cosf(u2);
  g1 = r * __builtin_amdgcn_sinf(u2);
}
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& g0, double& g1) {
  const double u1 = ((double)a + 0.5) * 2.3283064365386963e-10;
  const double u2 = (double)b * 2.3283064365386963e-10;
  const double r = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  g0 = r * cs;
  g1 = r * sn;
}

// two complex normals (words 0-1 and 2-3 through Box-Muller) of one
// Philox4x32-10 call at counter (idx lo, idx hi ^ tag, n lo, n hi), key seed
template <typename T>
__device__ __forceinline__ void rng_normals(unsigned long long seed, long long n, long long idx, uint32_t tag,
                                            cplx<T>& w1, cplx<T>& w2) {
  U4 ctr{(uint32_t)idx, (uint32_t)((unsigned long long)idx >> 32) ^ tag, (uint32_t)n,
         (uint32_t)((unsigned long long)n >> 32)};
  const U4 r = philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  T g0, g1, g2, g3;
  box_muller(r.x, r.y, g0, g1);
  box_muller(r.z, r.w, g2, g3);
  w1 = mk<T>(g0, g1);
  w2 = mk<T>(g2, g3);
}

// raw draws w1[c], w2[c] of mode idx in a step: counter group c at the step
template <typename T>
__device__ __forceinline__ void rng_draw(unsigned long long seed, long long step, long long idx, int c,
                                         cplx<T>& w1, cplx<T>& w2) {
  rng_normals<T>(seed, step, idx, (uint32_t)c << 24, w1, w2);
}

// The single-draw stream (a step whose draws feed one stage): entry e of
// step stp is word pair stp & 1 of counter group e at the step pair stp >> 1,
// so one call serves two consecutive steps (even: the first pair, odd: the
// second).  The top bit of the counter's second word keeps these counters
// apart from rng_draw's.
template <typename T>
__device__ __forceinline__ void rng_pair(unsigned long long seed, long long stp, long long idx, int e,
                                         cplx<T>& even, cplx<T>& odd) {
  rng_normals<T>(seed, stp >> 1, idx, ((uint32_t)e << 24) ^ 0x80000000u, even, odd);
}

// raw draws (B, 2, nc, nk) of one step: the device counterpart of random_fields
template <typename T>
__global__ void __launch_bounds__(256) draw_kernel(void* out, const unsigned long long* seeds,
                                                   const long long* msteps, long long nk, int nc) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= nk) return;
  cplx<T>* o = (cplx<T>*)out + (long long)b * 2 * nc * nk + idx;
  for (int c = 0; c < nc; ++c) {
    cplx<T> w1, w2;
    rng_draw<T>(seeds[b], msteps[b], idx, c, w1, w2);
    o[(long long)c * nk] = w1;
    o[(long long)(nc + c) * nk] = w2;
  }
}

// ------------------------------------------------------------------ stage update
enum Integ { EXPLICIT_PC = 0, IMPLICIT_PC = 1, RK2 = 2, RK3 = 3, RK4 = 4 };

struct StageArgs {
  void* uhat;
  void* u0;
  const void* rhs_nl;          // unprojected nonlinear term or null
  const void* lin_op;          // real (nk) or null: -nu k^2
  const void* corr;            // complex (nk) correlation function
  const void* noise;           // raw draws (B, 2, nc, nk) or null
  const unsigned long long* seeds;
  const double* fcoef;         // (B) forcing coefficient or null
  long long nk;
  int n_members;
  int N0, N1, N2, nh;
  double ks0, ks1, ks2;
  double nu;
  int integ, stage, last_stage, save_u0;
  double ca, cb, dt, dt_sqrt;
  int noise_on, noise_type, use_rng;
  // on-device draws feed one stage of the step only: w1 + a w2 is drawn as
  // sqrt(1 + a^2) w, one complex normal per entry instead of two
  int single_draw;
  double noise_a, noise_mag, noise_norm;
  int mask_nyq;
  double fk2min, fk2max;
  long long step;
  const long long* msteps;     // per-member step counters keying the RNG (null: step)
  // mode 0: fused stage; 1: write the four right-hand-side terms (compute_rhs);
  // 2: stage update from given terms (an integrator's stepping)
  int mode;
  void* t_nl;      // mode 1 out / mode 2 in (projected nonlinear term)
  void* t_lin;     // mode 1 out / mode 2 in-out (implicit_pc overwrites it with L u0)
  void* t_force;
  void* t_noise;
};

template <typename T, int D> struct Mode {
  T k[D];
  T k2;
  bool nyq;
};

// Wavenumbers of the mode at grid indices (i, j, l) (2D: (i, l), j unused).
template <typename T, int D>
__device__ __forceinline__ Mode<T, D> mode_at(const StageArgs& a, int i, int j, int l) {
  Mode<T, D> m;
  if constexpr (D == 3) {
    m.k[0] = (T)(cmode(i, a.N0) * a.ks0);
    m.k[1] = (T)(cmode(j, a.N1) * a.ks1);
    m.k[2] = (T)(l * a.ks2);
    m.nyq = (a.N0 % 2 == 0 && i == a.N0 / 2) || (a.N1 % 2 == 0 && j == a.N1 / 2) ||
            (a.N2 % 2 == 0 && l == a.N2 / 2);
  } else {
    m.k[0] = (T)(cmode(i, a.N0) * a.ks0);
    m.k[1] = (T)(l * a.ks1);
    m.nyq = (a.N0 % 2 == 0 && i == a.N0 / 2) || (a.N1 % 2 == 0 && l == a.N1 / 2);
  }
  m.k2 = 0;
#pragma unroll
  for (int q = 0; q < D; ++q) m.k2 += m.k[q] * m.k[q];
  return m;
}

// Mode of flat index idx = (i * N1 + j) * nh + l (2D: i * nh + l); 32-bit
// division whenever a member's spectrum has fewer than 2^32 modes.
template <typename T, int D>
__device__ __forceinline__ Mode<T, D> decode(const StageArgs& a, long long idx) {
  int i, j = 0, l;
  if (a.nk < (1ll << 32)) {
    const unsigned u = (unsigned)idx, nh = (unsigned)a.nh;
    const unsigned t = u / nh;
    l = (int)(u - t * nh);
    if constexpr (D == 3) { i = (int)(t / (unsigned)a.N1); j = (int)(t - (unsigned)i * (unsigned)a.N1); }
    else i = (int)t;
  } else {
    const long long t = idx / a.nh;
    l = (int)(idx - t * a.nh);
    if constexpr (D == 3) { i = (int)(t / a.N1); j = (int)(t - (long long)i * a.N1); }
    else i = (int)t;
  }
  return mode_at<T, D>(a, i, j, l);
}

template <typename T, int D> __device__ __forceinline__ void project(cplx<T>* v, const Mode<T, D>& m) {
  const T den = m.k2 == (T)0 ? (T)1 : m.k2;
  cplx<T> p = mk<T>(0, 0);
#pragma unroll
  for (int i = 0; i < D; ++i) p = add<T>(p, scl<T>(v[i], m.k[i] / den));
#pragma unroll
  for (int i = 0; i < D; ++i) v[i] = sub<T>(v[i], scl<T>(p, m.k[i]));
}

// The nonlinear, noise and forcing terms of one mode of member b at RNG
// step stp (space_solver.py:240-300): nl projected (math.py:72-100) and
// Nyquist-masked, nz the projected noise scaled by noise_mag, fr the band
// forcing.  nlin: the unprojected nonlinear term when the caller holds it,
// else it is read from a.rhs_nl (or zero without a nonlinear term).  zpre:
// this step's single-draw normals when the caller drew them (rng_pair), else
// they are drawn here.  RES: the register-resident Stokes steps (1 thermal,
// 2 correlated noise), which have no nonlinear term, no injected draws and no
// forcing, take their single-draw normals from zpre and the correlation
// function's value from cpre: the branches they never take are compiled out
// (the operations of the path they take are the same).
template <typename T, int D, int RES = 0>
__device__ __forceinline__ void stage_terms(const StageArgs& a, long long idx, int b, const cplx<T>* nlin,
                                            const cplx<T>* u, const Mode<T, D>& m, long long stp, cplx<T>* nl,
                                            cplx<T>* nz, cplx<T>* fr, const cplx<T>* zpre = nullptr,
                                            const cplx<T>* cpre = nullptr) {
  const long long mb = (long long)b * D * a.nk;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    nl[i] = mk<T>(0, 0);
    nz[i] = mk<T>(0, 0);
    fr[i] = mk<T>(0, 0);
  }
  // The loads of the nonlinear term and of the correlation function are
  // issued first and used after the noise draws, so the draws' arithmetic
  // runs while the loads are in flight.
  const bool has_nl = !RES && (nlin || a.rhs_nl);
  if (has_nl) {
    if (nlin) {
#pragma unroll
      for (int i = 0; i < D; ++i) nl[i] = nlin[i];
    } else {
      const cplx<T>* r = (const cplx<T>*)a.rhs_nl + mb + idx;
#pragma unroll
      for (int i = 0; i < D; ++i) nl[i] = r[i * a.nk];
    }
  }
  const bool thermal = RES ? RES == 1 : a.noise_type == 0;
  cplx<T> C = mk<T>(0, 0);
  if (a.noise_on && !thermal) C = RES ? *cpre : ((const cplx<T>*)a.corr)[idx];

  // noise (space_solver.py:227-265)
  if (a.noise_on) {
    const T fac = m.k2 == (T)0 ? (T)0 : (T)a.noise_norm;
    const T na = (T)a.noise_a;
    const unsigned long long seed = !RES && a.use_rng ? a.seeds[b] : 0ull;
    const int nc = thermal ? D * D : D;
    const cplx<T>* raw = !RES && a.noise ? (const cplx<T>*)a.noise + (long long)b * 2 * nc * a.nk + idx : nullptr;
    cplx<T> w[D * D];
    // entry e of the on-device stream: w1 + a w2, or sqrt(1 + a^2) w with
    // one draw (the single-draw stream) when the step has a single noise stage
    const T sfac = fac * (T)sqrt(1.0 + a.noise_a * a.noise_a);
    auto draw = [&](int e) -> cplx<T> {
      cplx<T> w1, w2;
      if (RES) return scl<T>(zpre[e], sfac);
      if (a.single_draw) {
        if (zpre) return scl<T>(zpre[e], sfac);
        rng_pair<T>(seed, stp, idx, e, w1, w2);
        return scl<T>((stp & 1) ? w2 : w1, sfac);
      }
      rng_draw<T>(seed, stp, idx, e, w1, w2);
      return add<T>(scl<T>(w1, fac), scl<T>(scl<T>(w2, fac), na));
    };
    bool projected = false;  // nz already orthogonal to k
    if (!raw && thermal) {
      // On-device stream, thermal noise.  The reference projects the
      // divergence of the symmetrised tensor, n = P (i W_sym k) (space_solver.py:
      // 240-265 with math.py:72-100).  W_sym has E W_ij W_lm* = s2 (d_il d_jm +
      // d_im d_jl) (entries of variance s2, 2 s2 on the diagonal), so i W_sym k
      // has covariance s2 (k^2 I + k k^T) and n has s2 k^2 P, P = I - k k^T / k^2.
      // i k x z, z a vector of independent entries of variance s2, has
      // covariance s2 (k^2 I - k k^T) = s2 k^2 P as well and is orthogonal to k:
      // the same Gaussian law from D = 3 draws per mode in 3D and one in 2D
      // (z along the third axis) instead of D(D+1)/2, with no projection.
      if constexpr (D == 3) {
        const cplx<T> z0 = draw(0), z1 = draw(1), z2 = draw(2);
        const T k0 = m.k[0], k1 = m.k[1], k2 = m.k[2];
        nz[0] = mul_i<T>(sub<T>(scl<T>(z2, k1), scl<T>(z1, k2)));
        nz[1] = mul_i<T>(sub<T>(scl<T>(z0, k2), scl<T>(z2, k0)));
        nz[2] = mul_i<T>(sub<T>(scl<T>(z1, k0), scl<T>(z0, k1)));
      } else {
        const cplx<T> z = draw(0);
        nz[0] = mul_i<T>(scl<T>(z, m.k[1]));
        nz[1] = mul_i<T>(scl<T>(z, -m.k[0]));
      }
      projected = true;
    } else if (!raw) {
#pragma unroll
      for (int c = 0; c < D; ++c) w[c] = draw(c);
    } else {
#pragma unroll
      for (int c = 0; c < D * D; ++c) {
        if (c < nc) {
          const cplx<T> w1 = raw[(long long)c * a.nk], w2 = raw[(long long)(nc + c) * a.nk];
          w[c] = add<T>(scl<T>(w1, fac), scl<T>(scl<T>(w2, fac), na));
        }
      }
    }
    if (!projected) {
      if (thermal) {
        // math.py / space_solver.py:15-64: i k . (W + W^T) sqrt(1/2)
        const T h = (T)0.70710678118654752440;
#pragma unroll
        for (int i = 0; i < D; ++i) {
          cplx<T> acc = mk<T>(0, 0);
#pragma unroll
          for (int j = 0; j < D; ++j) acc = add<T>(acc, scl<T>(add<T>(w[i * D + j], w[j * D + i]), h * m.k[j]));
          nz[i] = mul_i<T>(acc);
        }
      } else {
#pragma unroll
        for (int i = 0; i < D; ++i) nz[i] = mul<T>(C, w[i]);
      }
      project<T, D>(nz, m);
    }
    const T mag = (T)a.noise_mag;
#pragma unroll
    for (int i = 0; i < D; ++i) nz[i] = scl<T>(nz[i], mag);
  }

  // nonlinear term: Leray projection (math.py:72-100) and optional Nyquist mask
  if (has_nl) {
    project<T, D>(nl, m);
    if (a.mask_nyq && m.nyq) {
#pragma unroll
      for (int i = 0; i < D; ++i) nl[i] = mk<T>(0, 0);
    }
  }

  // constant-power band forcing
  if (!RES && a.fcoef && m.k2 >= (T)a.fk2min && m.k2 <= (T)a.fk2max) {
    const T fc = (T)a.fcoef[b];
#pragma unroll
    for (int i = 0; i < D; ++i) fr[i] = scl<T>(u[i], fc);
  }
}

// The integrator's update of one mode (time_integrator.py:51-318) from its
// terms: u the stage input, uz the step's start, Lk the linear operator.
template <typename T, int D>
__device__ __forceinline__ void stage_combine(const StageArgs& a, const cplx<T>* u, const cplx<T>* uz,
                                              const cplx<T>* nl, const cplx<T>* nz, const cplx<T>* fr, T Lk,
                                              cplx<T>* out) {
  const T dt = (T)a.dt, ds = (T)a.dt_sqrt, ca = (T)a.ca, cb = (T)a.cb;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    cplx<T> r;
    switch (a.integ) {
      case EXPLICIT_PC: {
        const cplx<T> tot = add<T>(add<T>(nl[i], scl<T>(u[i], Lk)), fr[i]);
        r = add<T>(add<T>(uz[i], scl<T>(tot, ca * dt)), scl<T>(nz[i], ds));
        break;
      }
      case IMPLICIT_PC: {
        const cplx<T> lin = a.last_stage ? scl<T>(uz[i], Lk) : scl<T>(u[i], Lk);
        const T inv = (T)1 / ((T)1 - dt / (T)2 * Lk);
        cplx<T> s = add<T>(uz[i], scl<T>(add<T>(nl[i], fr[i]), ca * dt));
        s = add<T>(s, scl<T>(lin, cb * dt));
        s = add<T>(s, scl<T>(nz[i], ds));
        r = scl<T>(s, inv);
        break;
      }
      case RK2: {
        const cplx<T> tot = add<T>(add<T>(nl[i], scl<T>(u[i], Lk)), fr[i]);
        r = add<T>(scl<T>(uz[i], ca), scl<T>(add<T>(u[i], scl<T>(tot, dt)), cb));
        break;
      }
      case RK3: {
        const cplx<T> tot = add<T>(add<T>(nl[i], scl<T>(u[i], Lk)), fr[i]);
        const cplx<T> in = add<T>(add<T>(u[i], scl<T>(tot, dt)), scl<T>(nz[i], ds));
        r = add<T>(scl<T>(uz[i], ca), scl<T>(in, cb));
        break;
      }
      default: {  // RK4
        const cplx<T> tot = add<T>(add<T>(nl[i], scl<T>(u[i], Lk)), fr[i]);
        r = add<T>(uz[i], scl<T>(tot, ca * dt));
        break;
      }
    }
    out[i] = r;
  }
}

// One mode of one member: assemble the stage's terms and update u_hat.
// nlin: the unprojected nonlinear term of this mode when the caller already
// holds it (fused forward pass), else it is read from a.rhs_nl.
// upre / uzpre: u_hat and u0 of this mode already fetched (stage_fetch).
// mpre: the mode's wavenumbers when the caller knows its grid indices.
// CP: cache policy of the u_hat / u0 stores.
template <typename T, int D, int CP = 0>
__device__ __forceinline__ void stage_point(const StageArgs& a, long long idx, int b, const cplx<T>* nlin,
                                            const cplx<T>* upre = nullptr, const cplx<T>* uzpre = nullptr,
                                            const Mode<T, D>* mpre = nullptr) {
  const Mode<T, D> m = mpre ? *mpre : decode<T, D>(a, idx);
  const long long mb = (long long)b * D * a.nk;
  cplx<T>* uh = (cplx<T>*)a.uhat + mb + idx;
  cplx<T>* u0p = (cplx<T>*)a.u0 + mb + idx;

  cplx<T> u[D], uz[D], nl[D], nz[D], fr[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    u[i] = upre ? upre[i] : uh[i * a.nk];
    uz[i] = (a.save_u0 || a.mode == 1) ? u[i] : (uzpre ? uzpre[i] : u0p[i * a.nk]);
  }
  if (a.save_u0 == 1 && a.mode == 0) {
#pragma unroll
    for (int i = 0; i < D; ++i) st_stream<cp_store(CP), T>(u0p + i * a.nk, u[i]);
  }
  const T Lk = a.lin_op ? ((const T*)a.lin_op)[idx] : (T)(-a.nu) * m.k2;

  if (a.mode == 2) {
    // given terms: the caller's arrays, exactly as compute_rhs left them
    const cplx<T>* gn = (const cplx<T>*)a.t_nl + mb + idx;
    cplx<T>* gl = (cplx<T>*)a.t_lin + mb + idx;
    const cplx<T>* gf = (const cplx<T>*)a.t_force + mb + idx;
    const cplx<T>* gz = (const cplx<T>*)a.t_noise + mb + idx;
    const T dt = (T)a.dt, ds = (T)a.dt_sqrt, ca = (T)a.ca, cb = (T)a.cb;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const cplx<T> n_ = gn[i * a.nk], f_ = gf[i * a.nk], z_ = gz[i * a.nk];
      cplx<T> l_ = gl[i * a.nk];
      cplx<T> r;
      switch (a.integ) {
        case EXPLICIT_PC:
          r = add<T>(add<T>(uz[i], scl<T>(add<T>(add<T>(n_, l_), f_), ca * dt)), scl<T>(z_, ds));
          break;
        case IMPLICIT_PC: {
          if (a.last_stage) { l_ = scl<T>(uz[i], Lk); gl[i * a.nk] = l_; }
          const T inv = (T)1 / ((T)1 - dt / (T)2 * Lk);
          cplx<T> s_ = add<T>(uz[i], scl<T>(add<T>(n_, f_), ca * dt));
          s_ = add<T>(s_, scl<T>(l_, cb * dt));
          s_ = add<T>(s_, scl<T>(z_, ds));
          r = scl<T>(s_, inv);
          break;
        }
        case RK2:
          r = add<T>(scl<T>(uz[i], ca), scl<T>(add<T>(u[i], scl<T>(add<T>(add<T>(n_, l_), f_), dt)), cb));
          break;
        case RK3: {
          const cplx<T> in = add<T>(add<T>(u[i], scl<T>(add<T>(add<T>(n_, l_), f_), dt)), scl<T>(z_, ds));
          r = add<T>(scl<T>(uz[i], ca), scl<T>(in, cb));
          break;
        }
        default:
          r = add<T>(uz[i], scl<T>(add<T>(add<T>(n_, l_), f_), ca * dt));
          break;
      }
      uh[i * a.nk] = r;
    }
    return;
  }

  const long long stp = a.msteps ? a.msteps[b] : a.step;
  stage_terms<T, D>(a, idx, b, nlin, u, m, stp, nl, nz, fr);

  if (a.mode == 1) {
    cplx<T>* on = (cplx<T>*)a.t_nl + mb + idx;
    cplx<T>* ol = (cplx<T>*)a.t_lin + mb + idx;
    cplx<T>* of = (cplx<T>*)a.t_force + mb + idx;
    cplx<T>* oz = (cplx<T>*)a.t_noise + mb + idx;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      on[i * a.nk] = nl[i];
      ol[i * a.nk] = scl<T>(u[i], Lk);
      of[i * a.nk] = fr[i];
      oz[i * a.nk] = nz[i];
    }
    return;
  }

  cplx<T> r[D];
  stage_combine<T, D>(a, u, uz, nl, nz, fr, Lk, r);
#pragma unroll
  for (int i = 0; i < D; ++i) st_stream<cp_store(CP), T>(uh + i * a.nk, r[i]);
}

// Loads of stage_point hoisted so that several modes' loads are in flight
// (CP: their cache policy).
template <typename T, int D, int CP = 0>
__device__ __forceinline__ void stage_fetch(const StageArgs& a, long long idx, int b, cplx<T>* u, cplx<T>* uz) {
  const long long mb = (long long)b * D * a.nk + idx;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    u[i] = ld_stream<cp_load(CP), T>((const cplx<T>*)a.uhat + mb + i * a.nk);
    uz[i] = (a.save_u0 || a.mode == 1) ? mk<T>(0, 0) : ld_stream<cp_load(CP), T>((const cplx<T>*)a.u0 + mb + i * a.nk);
  }
}

// MPT modes per thread (block-strided, so each load instruction stays
// coalesced): their u_hat / u0 loads are all issued first and their noise
// chains interleave, which hides both memory and VALU latency.  CP: the
// cache policy of the u_hat / u0 loads and stores.
template <typename T, int D, int MPT, int CP = 0>
__global__ void __launch_bounds__(kThreads) stage_kernel(StageArgs a) {
  const long long base = (long long)blockIdx.x * kThreads * MPT + threadIdx.x;
  const int b = blockIdx.y;
  cplx<T> uu[MPT][D], zz[MPT][D];
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    const long long idx = base + (long long)q * kThreads;
    if (idx < a.nk && a.mode != 2) stage_fetch<T, D, CP>(a, idx, b, uu[q], zz[q]);
  }
#pragma unroll
  for (int q = 0; q < MPT; ++q) {
    const long long idx = base + (long long)q * kThreads;
    if (idx >= a.nk) continue;
    if (a.mode == 2) stage_point<T, D>(a, idx, b, nullptr);
    else stage_point<T, D, CP>(a, idx, b, nullptr, uu[q], zz[q]);
  }
}

// Stokes flow stepped with the state in registers (time_integrator.py:111-129
// run with is_nonlinear = False, space_solver.py:227-265): the step of
// implicit PC whose predictor is dead, u <- (u + dt/2 L u + sqrt(dt/2) n) /
// (1 - dt/2 L), for nsteps steps of members [b0, b0 + MB) of one mode, with
// on-device noise keyed by each member's step counter.  Each step is the
// arithmetic of stage_point for that step (bitwise the same state).  acc
// (nullable) accumulates sum over the members, components and every step of
// |u_hat|^2 in FP64 in registers, added to acc[idx] once at the end.
template <typename T, int D, int MB, int RES>
__global__ void __launch_bounds__(kThreads) stokes_steps_kernel(StageArgs a, int nsteps, double* acc) {
  static_assert(RES == 1 || RES == 2, "thermal (1) or correlated (2) noise");
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= a.nk) return;
  const Mode<T, D> m = decode<T, D>(a, idx);
  const T Lk = a.lin_op ? ((const T*)a.lin_op)[idx] : (T)(-a.nu) * m.k2;
  // the correlation function at this mode (correlated noise)
  const cplx<T> C = RES == 2 && a.noise_on ? ((const cplx<T>*)a.corr)[idx] : mk<T>(0, 0);
  const int b0 = blockIdx.y * MB;
  double s = 0.0;
  for (int bb = 0; bb < MB; ++bb) {
    const int b = b0 + bb;
    if (b >= a.n_members) break;
    cplx<T>* uh = (cplx<T>*)a.uhat + (long long)b * D * a.nk + idx;
    cplx<T> u[D];
#pragma unroll
    for (int i = 0; i < D; ++i) u[i] = uh[i * a.nk];
    const long long st0 = a.msteps[b];
    // the single-draw normals of a step pair, drawn at its first step here
    constexpr int NE = D == 3 ? 3 : 2;
    constexpr int ne = RES == 1 ? (D == 3 ? 3 : 1) : D;
    const unsigned long long seed = a.noise_on ? a.seeds[b] : 0ull;
    cplx<T> ze[NE], zo[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) ze[e] = zo[e] = mk<T>(0, 0);
    for (int t = 0; t < nsteps; ++t) {
      const long long stp = st0 + t;
      if (a.noise_on && (t == 0 || (stp & 1) == 0)) {
#pragma unroll
        for (int e = 0; e < NE; ++e)
          if (e < ne) rng_pair<T>(seed, stp, idx, e, ze[e], zo[e]);
      }
      cplx<T> zs[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) zs[e] = (stp & 1) ? zo[e] : ze[e];
      cplx<T> nl[D], nz[D], fr[D], r[D];
      stage_terms<T, D, RES>(a, idx, b, nullptr, u, m, stp, nl, nz, fr, zs, &C);
      stage_combine<T, D>(a, u, u, nl, nz, fr, Lk, r);
#pragma unroll
      for (int i = 0; i < D; ++i) {
        u[i] = r[i];
        if (acc) s += (double)r[i].x * (double)r[i].x + (double)r[i].y * (double)r[i].y;
      }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) uh[i * a.nk] = u[i];
  }
  if (acc) atomicAdd(acc + idx, s);
}

// ------------------------------------------------------------- reductions (FP64)
template <typename T> __device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

struct DiagArgs {
  const void* uhat;
  double* out;                 // (B, 2): energy, enstrophy
  long long nk;
  int N0, N1, N2, nh;
  double ks0, ks1, ks2;
};

template <typename T, int D>
__global__ void __launch_bounds__(kThreads) diag_kernel(DiagArgs a) {
  __shared__ double sh[kThreads / 64];
  const int b = blockIdx.y;
  double e = 0, z = 0;
  StageArgs g{};
  g.nh = a.nh; g.N0 = a.N0; g.N1 = a.N1; g.N2 = a.N2; g.ks0 = a.ks0; g.ks1 = a.ks1; g.ks2 = a.ks2;
  const cplx<T>* u = (const cplx<T>*)a.uhat + (long long)b * D * a.nk;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < a.nk;
       idx += (long long)gridDim.x * kThreads) {
    const Mode<T, D> m = decode<T, D>(g, idx);
    const int l = (int)(idx % a.nh);
    const int NL = D == 3 ? a.N2 : a.N1;
    const double w = (l == 0 || (NL % 2 == 0 && l == NL / 2)) ? 1.0 : 2.0;
    double ku[D][2];
    double s = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const cplx<T> v = u[i * a.nk + idx];
      ku[i][0] = v.x; ku[i][1] = v.y;
      s += (double)v.x * v.x + (double)v.y * v.y;
    }
    e += w * s;
    double k[D];
#pragma unroll
    for (int i = 0; i < D; ++i) k[i] = (double)m.k[i];
    double zz;
    if constexpr (D == 3) {
      zz = 0;
      for (int c = 0; c < 2; ++c) {
        const double w0 = k[1] * ku[2][c] - k[2] * ku[1][c];
        const double w1 = k[2] * ku[0][c] - k[0] * ku[2][c];
        const double w2 = k[0] * ku[1][c] - k[1] * ku[0][c];
        zz += w0 * w0 + w1 * w1 + w2 * w2;
      }
    } else {
      zz = 0;
      for (int c = 0; c < 2; ++c) {
        const double w0 = k[0] * ku[1][c] - k[1] * ku[0][c];
        zz += w0 * w0;
      }
    }
    z += w * zz;
  }
  e = block_sum<T>(e, sh);
  z = block_sum<T>(z, sh);
  if (threadIdx.x == 0) {
    atomicAdd(&a.out[2 * b + 0], 0.5 * e);
    atomicAdd(&a.out[2 * b + 1], 0.5 * z);
  }
}

// energy in the forcing band -> coefficient P / (2 E_band), one block per member
struct BandArgs {
  const void* uhat;
  double* coef;
  long long nk;
  int N0, N1, N2, nh;
  double ks0, ks1, ks2;
  double kmin2, kmax2, kmax, power;
};

template <typename T, int D>
__global__ void __launch_bounds__(kThreads) band_kernel(BandArgs a) {
  __shared__ double sh[kThreads / 64];
  const int b = blockIdx.x;
  const cplx<T>* u = (const cplx<T>*)a.uhat + (long long)b * D * a.nk;
  const int NL = D == 3 ? a.N2 : a.N1;
  const int NA = a.N0, NB = D == 3 ? a.N1 : 1;
  const int K0 = min((int)floor(a.kmax / a.ks0), NA / 2);
  const int K1 = D == 3 ? min((int)floor(a.kmax / a.ks1), NB / 2) : 0;
  const int K2 = min((int)floor(a.kmax / (D == 3 ? a.ks2 : a.ks1)), a.nh - 1);
  const int lo0 = max(-K0, -(NA / 2)), hi0 = min(K0, (NA - 1) / 2);
  const int lo1 = D == 3 ? max(-K1, -(NB / 2)) : 0, hi1 = D == 3 ? min(K1, (NB - 1) / 2) : 0;
  const int n0 = hi0 - lo0 + 1, n1 = hi1 - lo1 + 1, n2 = K2 + 1;
  double e = 0;
  for (int t = threadIdx.x; t < n0 * n1 * n2; t += kThreads) {
    const int m2 = t % n2, r = t / n2;
    const int m1 = lo1 + r % n1, m0 = lo0 + r / n1;
    const double k0 = m0 * a.ks0;
    double k2;
    long long idx;
    if (D == 3) {
      const double k1 = m1 * a.ks1, kz = m2 * a.ks2;
      k2 = k0 * k0 + k1 * k1 + kz * kz;
      const int i = (m0 + NA) % NA, j = (m1 + NB) % NB;
      idx = ((long long)i * NB + j) * a.nh + m2;
    } else {
      const double k1 = m2 * a.ks1;
      k2 = k0 * k0 + k1 * k1;
      const int i = (m0 + NA) % NA;
      idx = (long long)i * a.nh + m2;
    }
    if (k2 < a.kmin2 || k2 > a.kmax2) continue;
    const double w = (m2 == 0 || (NL % 2 == 0 && m2 == NL / 2)) ? 1.0 : 2.0;
    double s = 0;
    for (int c = 0; c < D; ++c) {
      const cplx<T> v = u[c * a.nk + idx];
      s += (double)v.x * v.x + (double)v.y * v.y;
    }
    e += 0.5 * w * s;
  }
  e = block_sum<T>(e, sh);
  if (threadIdx.x == 0) a.coef[b] = e > 0 ? a.power / (2 * e) : 0.0;
}

// S(k) accumulation: acc[idx] += sum over members and components of |u_hat|^2
template <typename T, int D>
__global__ void __launch_bounds__(kThreads) sk_kernel(const void* uhat, double* acc, long long nk, int nmem) {
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (idx >= nk) return;
  const cplx<T>* u = (const cplx<T>*)uhat;
  double s = 0;
  for (int b = 0; b < nmem; ++b)
    for (int c = 0; c < D; ++c) {
      const cplx<T> v = u[((long long)b * D + c) * nk + idx];
      s += (double)v.x * v.x + (double)v.y * v.y;
    }
  acc[idx] += s;
}

// Shell spectrum: out[b * nshell + round(|k|)] += 1/2 w sum_c |u_hat_c|^2
template <typename T, int D>
__global__ void __launch_bounds__(kThreads) spectrum_kernel(DiagArgs a, int nshell) {
  const int b = blockIdx.y;
  StageArgs g{};
  g.nh = a.nh; g.N0 = a.N0; g.N1 = a.N1; g.N2 = a.N2; g.ks0 = a.ks0; g.ks1 = a.ks1; g.ks2 = a.ks2;
  const cplx<T>* u = (const cplx<T>*)a.uhat + (long long)b * D * a.nk;
  const int NL = D == 3 ? a.N2 : a.N1;
  for (long long idx = (long long)blockIdx.x * kThreads + threadIdx.x; idx < a.nk;
       idx += (long long)gridDim.x * kThreads) {
    const Mode<double, D> m = decode<double, D>(g, idx);
    const int l = (int)(idx % a.nh);
    const double w = (l == 0 || (NL % 2 == 0 && l == NL / 2)) ? 1.0 : 2.0;
    double s = 0;
    for (int c = 0; c < D; ++c) {
      const cplx<T> v = u[c * a.nk + idx];
      s += (double)v.x * v.x + (double)v.y * v.y;
    }
    int sh = (int)rint(sqrt(m.k2));
    if (sh > nshell - 1) sh = nshell - 1;
    atomicAdd(&a.out[(long long)b * nshell + sh], 0.5 * w * s);
  }
}


// ---------------------------------------------------------------- column pass
// FFT along a strided axis: a block loads a tile of L rows x TC contiguous
// columns into LDS (coalesced row segments), transforms every column and
// writes the tile back.  Kinds:
//   K_INV        inverse, rows_in -> L with 3/2 zero padding (Nyquist split)
//   K_INV_UC3    axis 0 of 3D: emits u_hat and omega_y, omega_z of omega_hat = i k x u_hat
//   K_INV_CURL3  axis 1 of 3D from K_INV_UC3's five fields: emits u, omega_x =
//                i (ky w - kz v) formed at the load (ky, kz are diagonal there), omega_y, omega_z
//   K_INV_UC2    axis 0 of 2D: emits u_hat and the scalar vorticity
//   K_FWD        forward, L -> rows_out with truncation (Nyquist sum)
//   K_FWD_STAGE3/2  forward axis 0 of all components of a member, then the
//                stage update of every mode of the tile (no R buffer)
// Under the 2/3 rule ("prune") the inverse passes visit only retained
// modes: rows with |k| > Krow are zero without a load, and the axis-0 pass
// of 3D enumerates only the (j, l) columns with |k_j| <= K1, l <= K2; the
// columns it skips are never read downstream.
enum ColKind {
  K_INV = 0,
  K_FWD = 3,
  K_FWD_STAGE3 = 4,
  K_FWD_STAGE2 = 5,
  K_INV_UC3 = 8,     // u_hat -> velocity (fg 0..2) and omega_y, omega_z (fg 3, 4) of a 3D field
  K_INV_CURL3 = 10,  // K_INV_UC3's fields -> velocity (fg 0..2) and vorticity (fg 3..5)
  K_INV_UC2 = 9,     // u_hat -> velocity (fg 0..1) and scalar vorticity (fg 2) of a 2D field
};

// Column enumeration of a launch: column cc = jj * n2c + l of the tile
// space sits at jj * in_lp + l in the input and jj * out_lp + l in the
// output; columns with l < cload are loaded (the rest read as zero), columns
// with l < cvalid are stored.  Loads and stores may cover whole tiles past
// the meaningful columns where the padded rows have room for them.
struct ColGeom {
  int cols;                     // enumerated columns per matrix
  int n2c, cload, cvalid;
  int n_outer;                  // matrices per field
  int n_tiles;                  // column tiles per matrix (set by the launcher)
  long long in_lp, out_lp;
  long long in_outer, out_outer;
};

struct ColArgs {
  const void* in;
  void* out;
  int rows_in, rows_out;        // rows read / written per matrix
  ColGeom g;
  long long in_row, out_row;    // element strides between rows
  long long in_field, out_field;
  long long in_member, out_member;
  int n_fields;                 // fields per member
  int n_members;
  double scale;
  int gN0, gN1, gN2, gnh;       // unpadded grid (curl kinds)
  double ks0, ks1, ks2;         // wavenumber scales 2 pi / L_i
  int mask23;                   // 2/3 mask while loading (curl kinds)
  int prune;                    // visit retained modes only (2/3 rule)
  int Krow;                     // pruned inverse: rows with |k| > Krow are zero
  int jcompact, K1;             // group jj -> retained axis-1 index |k_j| <= K1 (2/3 rule)
  int f_base;                   // first output field of the launch
  StageArgs st;                 // fused stage kinds
};

template <typename T> constexpr int col_tc(int L) {
  return sizeof(T) == 4 ? (L <= 128 ? 32 : (L <= 384 ? 16 : 8)) : (L <= 128 ? 16 : (L <= 384 ? 8 : 4));
}
// tile width of the fused forward + stage kinds (0: not available)
template <typename T> constexpr int col_tc_stage(int L) {
  return sizeof(T) == 4 ? (L <= 384 ? 8 : (L <= 768 ? 4 : 0)) : (L <= 384 ? 4 : (L <= 768 ? 2 : 0));
}
template <typename T, int KIND> constexpr int col_width(int L) {
  return (KIND == K_FWD_STAGE3 || KIND == K_FWD_STAGE2) ? col_tc_stage<T>(L) : col_tc<T>(L);
}
// Tiles per block: 2 for variant 3 of the plain kinds, whose next tile's
// loads (at most 32 per thread, held in registers) are issued before the
// current tile's transform and stores; otherwise 1.
template <typename T, int KIND> constexpr int col_pipe_v(int L, int V) {
  if (V != 3) return 1;
  const bool kind_ok = KIND == K_INV || KIND == K_FWD;
  const int tc = col_width<T, KIND>(L);
  return kind_ok && tc > 0 && L * tc / kThreads <= 32 ? 2 : 1;
}
// Launch variants (the autotuner's choices): V = 0 the default tile width,
// 1 half of it, 2 twice it, 3 the default width with two tiles per block,
// 4 (3D stage kind) twice the width with 1024-thread blocks, 5 the default
// width with radix plan 1 (make_plan) where L has one, 6 and 7 the default
// width with nontemporal stores (7: and loads); width 0 where a variant does
// not fit in LDS, does not exist or would repeat variant 0.
template <typename T, int KIND> constexpr int col_width_v(int L, int V) {
  const int base = col_width<T, KIND>(L);
  if (base == 0 || V == 0 || V == 6 || V == 7) return base;
  if (V == 5) return has_alt_plan(L) ? base : 0;
  if (V == 3) return col_pipe_v<T, KIND>(L, 3) == 2 ? base : 0;
  // 1024-thread blocks cap a thread at 128 registers: the 2D stage update
  // would spill there, so variant 4 is the 3D stage kind's only
  if (V == 4 && KIND != K_FWD_STAGE3) return 0;
  const int tc = V == 1 ? base / 2 : base * 2;
  if (tc < 1 || tc > 64) return 0;
  const int nf = KIND == K_FWD_STAGE3 ? 3 : (KIND == K_FWD_STAGE2 ? 2 : 1);
  const long long lds = ((long long)L * nf * tc + tw_alloc(L)) * (long long)sizeof(cplx<T>);
  return lds <= 160 * 1024 ? tc : 0;
}
// Threads per block: 1024 for variant 4 of the 3D stage kind (twice the
// default tile width, one block per CU holding the whole LDS), else 256.
template <typename T, int KIND> constexpr int col_threads_v(int L, int V) {
  return (V == 4 && KIND == K_FWD_STAGE3) ? 1024 : kThreads;
}

// Row index of the padded spectrum -> source row and factor (3/2 rule,
// complex axis).  The Nyquist row n/2 is split half-and-half between +-n/2.
__device__ __forceinline__ int pad_src(int e, int n, int L, double& fac) {
  fac = 1.0;
  if (n == L) return e;
  const int h = n / 2;
  if (e < h) return e;
  if (e > L - h) return e - (L - n);
  if (e == h || e == L - h) { fac = 0.5; return h; }
  return -1;
}

// Consecutive block ids are dealt round-robin to the 8 XCDs; give each XCD a
// contiguous run of logical tiles so neighbours share its L2.
__device__ __forceinline__ long long xcd_order(long long bid, long long G) {
  const long long q = G / 8, r = G % 8, x = bid % 8;
  return x * q + (x < r ? x : r) + bid / 8;
}

// a bool as a type, to select a lambda's code path at compile time
template <bool B> struct BoolC { static constexpr bool value = B; };

template <int KIND> constexpr int col_fields() {
  return KIND == K_FWD_STAGE3 ? 3 : (KIND == K_FWD_STAGE2 ? 2 : 1);
}
// LDS of a column pass block: the twiddle table, then L rows of the tile
template <typename T, int L, int TC, int KIND, int RP = 0> constexpr int col_lds_tw() { return tw_alloc(L, RP); }
template <typename T, int L, int TC, int KIND> constexpr int col_lds_buf() { return L * col_fields<KIND>() * TC; }

// One block of a column pass: block bid of the nblk blocks of its launch;
// tws and buf are the block's LDS (col_lds_tw and col_lds_buf entries).
template <typename T, int L, int TC, int KIND, int PIPE, int NT, int RP = 0, int CP = 0>
__device__ __forceinline__ void col_body(const ColArgs& a, const cplx<T>* __restrict__ tw, cplx<T>* tws,
                                         cplx<T>* buf, long long bid, long long nblk) {
  constexpr bool STAGE = KIND == K_FWD_STAGE3 || KIND == K_FWD_STAGE2;
  constexpr bool CURL = KIND == K_INV_CURL3;
  constexpr bool INV = KIND == K_INV || KIND == K_INV_UC3 || KIND == K_INV_UC2 || CURL;
  constexpr bool UC = KIND == K_INV_UC3 || KIND == K_INV_UC2;
  constexpr bool GRID = UC;  // loads need the mode numbers of the unpadded grid
  constexpr int GD = KIND == K_INV_UC3 ? 3 : 2;
  constexpr int NFLD = KIND == K_FWD_STAGE3 ? 3 : (KIND == K_FWD_STAGE2 ? 2 : 1);
  constexpr int W = NFLD * TC;  // LDS columns
  constexpr int NIT = (L * W + NT - 1) / NT;
  // loads kept in flight per thread: 16, or 8 for the FP64 kinds that form
  // curl components (two loads per entry, 4 registers each)
  constexpr int CHM = (sizeof(T) == 8 && (UC || CURL)) ? 8 : 16;
  constexpr int CH = NIT < CHM ? NIT : CHM;
  static_assert(NT % TC == 0, "tile width must divide the block size");
  constexpr int TWN = tw_size(L, RP);  // a constant, as in row_pass
  for (int m = threadIdx.x; m < TWN; m += NT) tws[m] = tw[m];

  const int nfg = STAGE ? 1 : a.n_fields;
  long long lg = xcd_order(bid, nblk);
  const ColGeom& gm = a.g;
  // PIPE consecutive tiles of one (field, matrix, member) per block; the
  // tile index varies fastest, so the blocks an XCD runs together cover
  // neighbouring columns of the same rows; the kinds that form curl
  // components (K_INV_UC3 / UC2, K_INV_CURL3) take the field fastest
  // instead, so that the blocks re-reading a tile's components for the curl
  // run beside the ones that loaded them and find them in the same XCD's L2
  const int ntb = (gm.n_tiles + PIPE - 1) / PIPE;
  int tb, fg;
  if constexpr (UC || CURL) {
    fg = (int)(lg % nfg);
    lg /= nfg;
    tb = (int)(lg % ntb);
    lg /= ntb;
  } else {
    tb = (int)(lg % ntb);
    lg /= ntb;
    fg = (int)(lg % nfg);
    lg /= nfg;
  }
  // The block's role (uniform over the block): a curl component formed from
  // two loaded fields (curlb: the 2D curl, omega_y and omega_z of
  // K_INV_UC3, omega_x of K_INV_CURL3) or one loaded field; cfg: the curl
  // component or the field loaded
  const bool curlb = (UC && fg >= GD) || (CURL && fg == 3);
  const int cfg = KIND == K_INV_UC2 && fg >= 2 ? fg - 2 : (KIND == K_INV_UC3 && fg >= 3 ? fg - 2 : (CURL && fg >= 4 ? fg - 1 : fg));
  const int outer = (int)(lg % gm.n_outer);
  const int b = (int)(lg / gm.n_outer);
  int c0 = tb * PIPE * TC;  // first column of the tile in hand
  const cplx<T>* in = (const cplx<T>*)a.in + (long long)b * a.in_member + (long long)outer * gm.in_outer;

  // enumerated column -> (group j, column l within the group)
  auto cpos = [&](int cc, int& j, int& l) {
    int jj;
    if (gm.cols <= gm.n2c) { jj = 0; l = cc; }          // one group
    else if (gm.n2c == 1) { jj = cc; l = 0; }           // one column per group
    else { jj = cc / gm.n2c; l = cc % gm.n2c; }
    j = (a.jcompact && jj > a.K1) ? jj + a.gN1 - (2 * a.K1 + 1) : jj;
  };
  // LDS column cf -> field, tile column, (j, l), whether it is loaded (ld)
  // and stored (ok)
  struct Slot { int f, c, j, l; bool ld, ok; };
  // the stage kinds' tile starts in group jb0 and spans at most
  // ceil(TC / n2c) + 1 groups (PIPE = 1: c0 is fixed)
  const int jb0 = STAGE ? c0 / gm.n2c : 0;
  auto slot = [&](int cf) {
    Slot r;
    r.f = STAGE ? cf / TC : fg;
    r.c = STAGE ? cf % TC : cf;
    const int cc = c0 + r.c;
    const bool in = cc < gm.cols;
    if constexpr (STAGE) {
      // every column of every group is loaded and updated
      int l = cc - jb0 * gm.n2c, j = jb0;
      while (l >= gm.n2c) { l -= gm.n2c; ++j; }
      r.j = j;
      r.l = l;
      r.ld = r.ok = in;
      return r;
    }
    cpos(in ? cc : 0, r.j, r.l);
    r.ld = in && r.l < gm.cload;
    r.ok = in && r.l < gm.cvalid;
    return r;
  };
  // With W dividing the block size every thread touches one LDS column.
  constexpr bool HOIST = NT % W == 0;
  Slot ts{};
  if constexpr (HOIST) ts = slot(threadIdx.x % W);

  // ---- load: per chunk, issue every load, then combine and store to LDS
  // wavenumbers (ka, kb) of the curl omega_f = i (ka u_ia - kb u_ib) at row
  // e, column (j, l) of the unpadded grid
  auto curl_k = [&](int f, int src, int j, int l, int& ia, int& ib, T& ka, T& kb) {
    const int mi = cmode(src, a.gN0);
    const int mj = GD == 3 ? cmode(j, a.gN1) : 0;
    const T k0 = (T)(mi * a.ks0);
    const T k1 = GD == 3 ? (T)(mj * a.ks1) : (T)(l * a.ks1);
    const T k2 = GD == 3 ? (T)(l * a.ks2) : (T)0;
    if (GD == 2 || f == 2) { ia = 1; ib = 0; ka = k0; kb = k1; }
    else if (f == 0) { ia = 2; ib = 1; ka = k1; kb = k2; }
    else { ia = 0; ib = 2; ka = k2; kb = k0; }
  };
  // K_INV_CURL3's omega_x = i (ky w - kz v) at padded row src (the axis-1
  // mode) and column l (the last-axis mode): inputs w (field 2) and v (1)
  auto vort_k = [&](int src, int l, T& ka, T& kb) {
    ka = (T)(cmode(src, a.rows_in) * a.ks1);
    kb = (T)(l * a.ks2);
  };
  // Loads of this thread's LDS slots base .. base + NQ - 1 of the tile at c0
  // (NQ = the register arrays' length), all issued before any is used.
  // CB: a curl block (a compile-time copy of curlb, so that the
  // unrolled loops hold no per-entry branch between their loads)
  auto issue = [&](int base, auto& ra, auto& rb, auto cb) {
    constexpr bool CB = decltype(cb)::value;
    constexpr int NQ = (int)(sizeof(ra) / sizeof(ra[0]));
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = threadIdx.x + (base + q) * NT;
      ra[q] = mk<T>(0, 0);
      rb[q] = mk<T>(0, 0);
      if (base + q >= NIT || idx >= L * W) continue;
      const int e = idx / W;
      const Slot sl = HOIST ? ts : slot(idx % W);
      if (!sl.ld) continue;
      const int f = sl.f, j = sl.j;
      int l = sl.l;
      const long long col = (long long)j * gm.in_lp + l;
      double fc;
      int src = INV ? pad_src(e, a.rows_in, L, fc) : e;
      if ((KIND == K_INV || CURL) && a.prune && src >= 0 && abs(cmode(src, a.rows_in)) > a.Krow) src = -1;
      if (src < 0) continue;
      if constexpr (CURL) {
        const cplx<T>* p = in + (long long)src * a.in_row + col;
        if constexpr (!CB) {
          ra[q] = ld_stream<cp_load(CP), T>(p + (long long)cfg * a.in_field);
        } else {
          ra[q] = ld_stream<cp_load(CP), T>(p + 2 * a.in_field);
          rb[q] = ld_stream<cp_load(CP), T>(p + a.in_field);
        }
      } else if constexpr (GRID) {
        const int mi = cmode(src, a.gN0);
        const int mj = GD == 3 ? cmode(j, a.gN1) : 0;
        if (a.mask23 && (abs(mi) > a.gN0 / 3 || abs(mj) > a.gN1 / 3 || l > (GD == 3 ? a.gN2 : a.gN1) / 3))
          continue;
        const cplx<T>* p = in + (long long)src * a.in_row + col;
        if constexpr (!CB) {
          ra[q] = ld_stream<cp_load(CP), T>(p + (long long)cfg * a.in_field);
        } else {
          int ia, ib;
          T ka, kb;
          curl_k(cfg, src, j, l, ia, ib, ka, kb);
          ra[q] = ld_stream<cp_load(CP), T>(p + (long long)ia * a.in_field);
          rb[q] = ld_stream<cp_load(CP), T>(p + (long long)ib * a.in_field);
        }
      } else {
        ra[q] = ld_stream<cp_load(CP), T>(in + (long long)f * a.in_field + (long long)src * a.in_row + col);
      }
    }
  };
  // ... combined (curl, 3/2-rule Nyquist halving) and stored to LDS
  auto commit = [&](int base, const auto& ra, const auto& rb, auto cb) {
    constexpr bool CB = decltype(cb)::value;
    constexpr int NQ = (int)(sizeof(ra) / sizeof(ra[0]));
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int idx = threadIdx.x + (base + q) * NT;
      if (base + q >= NIT || idx >= L * W) continue;
      cplx<T> v = ra[q];
      if constexpr (CB) {
        // entries not loaded are zero in ra and rb, so any coefficients serve there
        const int e = idx / W;
        const Slot sl = HOIST ? ts : slot(idx % W);
        double fc;
        int src = pad_src(e, a.rows_in, L, fc);
        if (src < 0) src = 0;
        int ia, ib;
        T ka, kb;
        if constexpr (CURL) vort_k(src, sl.l, ka, kb);
        else curl_k(cfg, src, sl.j, sl.l, ia, ib, ka, kb);
        v = mul_i<T>(mk<T>(ka * ra[q].x - kb * rb[q].x, ka * ra[q].y - kb * rb[q].y));
      }
      if constexpr (INV) {
        // the 3/2 rule halves the Nyquist row it copies to +-n/2
        if (a.rows_in != L) {
          const int e = idx / W, h = a.rows_in / 2;
          if (e == h || e == L - h) v = scl<T>(v, (T)0.5);
        }
      }
      buf[idx] = v;  // idx = e * W + cf
    }
  };
  // The plain kinds with one tile per block skip this staging: their first
  // radix pass reads its inputs straight from global memory (below; 97 -> 94
  // us for the axis-1 forward at 256^3).  The stage kinds keep it (there the
  // direct form measured 3-7% slower at lengths 192 and 144, whose first
  // radix is 3), and so does an inverse whose plan has a single pass (its
  // last pass stores on its own), and so do the FP64 kinds that form curl
  // components where the first radix is 16: two loads of 16 bytes per entry
  // of a radix-16 butterfly held 199 registers (2 waves per SIMD) at 256^3,
  // and staged in chunks of 8 the axis-0 inverse went from 194 to 175 us.
  // With a smaller first radix (6 at length 96) the direct form stays faster
  // (the FP64 64^3 step 5% faster than staged).
  constexpr RadixPlan plD = make_plan(L, RP);
  constexpr bool DIRECT =
      PIPE == 1 && (KIND == K_FWD || ((KIND == K_INV || ((UC || CURL) && (sizeof(T) == 4 || plD.r[0] < 16))) && plD.n >= 2));
  if constexpr (!DIRECT) {
    for (int base = 0; base < NIT; base += CH) {
      cplx<T> ra[CH], rb[CH];
      if ((UC || CURL) && curlb) {
        issue(base, ra, rb, BoolC<true>{});
        commit(base, ra, rb, BoolC<true>{});
      } else {
        issue(base, ra, rb, BoolC<false>{});
        commit(base, ra, rb, BoolC<false>{});
      }
    }
    __syncthreads();
  }

  const T s = (T)a.scale;
  auto addr = [](int ff, int e) { return e * W + ff; };
  for (int it = 0; it < PIPE; ++it) {
  // two-tile blocks: the next tile's loads fly during this tile's transform
  // and stores (the condition is uniform over the block)
  const bool nxt = PIPE > 1 && it + 1 < PIPE && tb * PIPE + it + 1 < gm.n_tiles;
  cplx<T> pa[PIPE > 1 ? NIT : 1], pb[PIPE > 1 ? NIT : 1];
  if constexpr (PIPE > 1) {
    if (nxt) {
      c0 += TC;
      if constexpr (HOIST) ts = slot(threadIdx.x % W);
      issue(0, pa, pb, BoolC<false>{});
      c0 -= TC;
      if constexpr (HOIST) ts = slot(threadIdx.x % W);
    }
  }
  if constexpr (DIRECT) {
    // The first Stockham pass (span 1, no twiddles) with its inputs loaded
    // from global memory: butterfly q = (LDS column q % W, j = q / W) reads
    // rows j + r L/R0 of its column, exactly what the staged tile would
    // have given it, and writes its outputs to rows j R0 + r; then the
    // remaining passes run in LDS.  One LDS round and two barriers fewer.
    constexpr int R0 = plD.r[0];
    constexpr int NB0 = L / R0;
    constexpr int TOT0 = W * NB0;
    constexpr int K0 = (TOT0 + NT - 1) / NT;
    cplx<T> v0[K0][R0];
#pragma unroll
    for (int kk = 0; kk < K0; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if (TOT0 % NT == 0 || q < TOT0) {
        const Slot sl = HOIST ? ts : slot(q % W);
        const int j = q / W;
        const cplx<T>* pcol = in + (long long)sl.j * gm.in_lp + sl.l;
        const cplx<T>* pc = pcol + (long long)sl.f * a.in_field;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          const int e = j + r * NB0;
          cplx<T> v = mk<T>(0, 0);
          if constexpr (UC) {
            // as issue + commit: the 2/3 mask on (i, j, l), then the
            // velocity component, or the curl of two loaded components
            // (zero operands where nothing is loaded)
            double fc;
            const int src = pad_src(e, a.rows_in, L, fc);
            cplx<T> ra = mk<T>(0, 0), rb = mk<T>(0, 0);
            int ia = 0, ib = 0;
            T ka = 0, kb = 0;
            if (curlb) curl_k(cfg, src < 0 ? 0 : src, sl.j, sl.l, ia, ib, ka, kb);
            if (sl.ld && src >= 0) {
              const int mi = cmode(src, a.gN0);
              const int mj = GD == 3 ? cmode(sl.j, a.gN1) : 0;
              const bool cut = a.mask23 && (abs(mi) > a.gN0 / 3 || abs(mj) > a.gN1 / 3 ||
                                            sl.l > (GD == 3 ? a.gN2 : a.gN1) / 3);
              if (!cut) {
                const cplx<T>* p = pcol + (long long)src * a.in_row;
                if (!curlb) {
                  ra = ld_stream<cp_load(CP), T>(p + (long long)cfg * a.in_field);
                } else {
                  ra = ld_stream<cp_load(CP), T>(p + (long long)ia * a.in_field);
                  rb = ld_stream<cp_load(CP), T>(p + (long long)ib * a.in_field);
                }
              }
            }
            v = curlb ? mul_i<T>(mk<T>(ka * ra.x - kb * rb.x, ka * ra.y - kb * rb.y)) : ra;
          } else if constexpr (CURL) {
            // as issue + commit: the 3/2 padding source row and the 2/3
            // rule's zero rows, then one field or the vorticity of two
            double fc;
            int src = pad_src(e, a.rows_in, L, fc);
            if (a.prune && src >= 0 && abs(cmode(src, a.rows_in)) > a.Krow) src = -1;
            if (sl.ld && src >= 0) {
              const cplx<T>* p = pcol + (long long)src * a.in_row;
              if (!curlb) {
                v = ld_stream<cp_load(CP), T>(p + (long long)cfg * a.in_field);
              } else {
                T ka, kb;
                vort_k(src, sl.l, ka, kb);
                const cplx<T> ra = ld_stream<cp_load(CP), T>(p + 2 * a.in_field);
                const cplx<T> rb = ld_stream<cp_load(CP), T>(p + a.in_field);
                v = mul_i<T>(mk<T>(ka * ra.x - kb * rb.x, ka * ra.y - kb * rb.y));
              }
            }
          } else if (sl.ld) {
            if constexpr (INV) {
              // the inverse's row map: 3/2 padding source row (or none),
              // the 2/3 rule's zero rows
              double fc;
              int src = pad_src(e, a.rows_in, L, fc);
              if (a.prune && src >= 0 && abs(cmode(src, a.rows_in)) > a.Krow) src = -1;
              if (src >= 0) v = ld_stream<cp_load(CP), T>(pc + (long long)src * a.in_row);
            } else {
              v = ld_stream<cp_load(CP), T>(pc + (long long)e * a.in_row);
            }
          }
          v0[kk][r] = v;
        }
      }
    }
    constexpr int SG = INV ? +1 : -1;
#pragma unroll
    for (int kk = 0; kk < K0; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if (TOT0 % NT == 0 || q < TOT0) {
        const int cf = q % W, j = q / W;
        if constexpr (INV) {
          // the 3/2 rule halves the Nyquist row it copies to +-n/2
          if (a.rows_in != L) {
            const int h = a.rows_in / 2;
#pragma unroll
            for (int r = 0; r < R0; ++r) {
              const int e = j + r * NB0;
              if (e == h || e == L - h) v0[kk][r] = scl<T>(v0[kk][r], (T)0.5);
            }
          }
        }
        dft<R0, SG, T>(v0[kk]);
#pragma unroll
        for (int r = 0; r < R0; ++r) buf[addr(cf, j * R0 + r)] = v0[kk][r];
      }
    }
    __syncthreads();
    if constexpr (INV) {
      // the middle passes; the last one stores to global memory below
      stockham<T, L, +1, W, NT, 1, R0, true, decltype(addr), plD.n - 1, RP>(buf, tws, addr);
    } else {
      stockham<T, L, -1, W, NT, 1, R0, true, decltype(addr), 99, RP>(buf, tws, addr);
    }
  } else if constexpr (INV) {
    // every pass but the last in LDS; the last one stores to global memory
    lds_fft_head<T, L, +1, W, NT, RP>(buf, tws, addr);
  } else {
    lds_fft<T, L, -1, W, NT, true, RP>(buf, tws, addr);
  }

  if constexpr (INV) {
    // The last Stockham pass, its outputs (natural order: row j + r NS of
    // LDS column q % W for butterfly q = (q % W, j = q / W)) scaled and
    // stored straight to global memory: one LDS round and one barrier
    // fewer than staging the transformed tile.
    cplx<T>* out = (cplx<T>*)a.out + (long long)b * a.out_member + (long long)(a.f_base + fg) * a.out_field +
                   (long long)outer * gm.out_outer;
    constexpr RadixPlan plI = make_plan(L, RP);
    constexpr int PL = plI.n - 1;
    constexpr int RL = plI.r[PL];
    constexpr int NSL = L / RL;  // the last pass's span = its butterflies per transform
    constexpr int TOTL = W * NSL;
    constexpr int KL = (TOTL + NT - 1) / NT;
    constexpr int OFFL = tw_offset(L, PL, RP);
    // W = TC divides the block size: every butterfly of a thread is in its column
    const Slot os = slot(threadIdx.x % TC);
    const long long ocol = (long long)os.j * gm.out_lp + os.l;
    cplx<T> vl[KL][RL];
#pragma unroll
    for (int kk = 0; kk < KL; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if (TOTL % NT == 0 || q < TOTL) {
        const int cf = q % W, j = q / W;
#pragma unroll
        for (int r = 0; r < RL; ++r) vl[kk][r] = buf[addr(cf, j + r * NSL)];
        if constexpr (NSL > 1) {
#pragma unroll
          for (int r = 1; r < RL; ++r) vl[kk][r] = mul<T>(vl[kk][r], twiddle<+1, T>(tws, OFFL + (r - 1) * NSL + j));
        }
        dft<RL, +1, T>(vl[kk]);
      }
    }
#pragma unroll
    for (int kk = 0; kk < KL; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if ((TOTL % NT == 0 || q < TOTL) && os.ok) {
        const int j = q / W;
#pragma unroll
        for (int r = 0; r < RL; ++r) st_stream<cp_store(CP), T>(out + (long long)(j + r * NSL) * a.out_row + ocol, scl<T>(vl[kk][r], s));
      }
    }
  } else {
    const int n = a.rows_out, h = n / 2;
    auto trunc = [&](int o, int cf) -> cplx<T> {
      cplx<T> v;
      if (n == L || o < h) v = buf[o * W + cf];
      else if (o == h) v = add<T>(buf[h * W + cf], buf[(L - h) * W + cf]);
      else v = buf[(o + (L - n)) * W + cf];
      return scl<T>(v, s);
    };
    if constexpr (STAGE) {
      // modes of the tile: fetch u_hat and u0 of SCH modes, then update them.
      // State index of (row o, flat column cc = j nh + l): o * N1 * nh + cc
      // (2D: o * nh + cc).
      constexpr int SCH = 2;
      const long long rowlen = NFLD == 3 ? (long long)a.gN1 * a.gnh : (long long)a.gnh;
      // TC divides the block size: one tile column per thread.  The stage
      // kinds enumerate the state's columns back to back (every one valid)
      const int c = threadIdx.x % TC;
      const bool store = c0 + c < gm.cols;
      const long long mcol = c0 + c;
      // grid indices (j, l) of the column: one division per thread
      const int jt = NFLD == 3 ? (int)((unsigned)mcol / (unsigned)a.gnh) : 0;
      const int lt = (int)mcol - jt * a.gnh;
      for (int idx0 = threadIdx.x; idx0 < n * TC; idx0 += SCH * NT) {
        cplx<T> uu[SCH][NFLD], zz[SCH][NFLD];
#pragma unroll
        for (int q = 0; q < SCH; ++q) {
          const int idx = idx0 + q * NT;
          if (idx < n * TC && store) stage_fetch<T, NFLD, CP>(a.st, (long long)(idx / TC) * rowlen + mcol, b, uu[q], zz[q]);
        }
#pragma unroll
        for (int q = 0; q < SCH; ++q) {
          const int idx = idx0 + q * NT;
          if (idx >= n * TC || !store) continue;
          const int o = idx / TC;
          cplx<T> nl[NFLD];
#pragma unroll
          for (int f = 0; f < NFLD; ++f) nl[f] = trunc(o, f * TC + c);
          const Mode<T, NFLD> md = mode_at<T, NFLD>(a.st, o, jt, lt);
          stage_point<T, NFLD, CP>(a.st, (long long)o * rowlen + mcol, b, nl, uu[q], zz[q], &md);
        }
      }
    } else {
      cplx<T>* out = (cplx<T>*)a.out + (long long)b * a.out_member + (long long)fg * a.out_field +
                     (long long)outer * gm.out_outer;
      const Slot os = slot(threadIdx.x % TC);
      const long long ocol = (long long)os.j * gm.out_lp + os.l;
      for (int idx = threadIdx.x; idx < n * TC; idx += NT) {
        if (!os.ok) break;
        const int o = idx / TC;
        st_stream<cp_store(CP), T>(out + (long long)o * a.out_row + ocol, trunc(o, os.c));
      }
    }
  }
  if (!nxt) break;
  if constexpr (PIPE > 1) {
    {
      __syncthreads();  // every read of this tile's LDS is done
      c0 += TC;
      if constexpr (HOIST) ts = slot(threadIdx.x % W);
      commit(0, pa, pb, BoolC<false>{});
      __syncthreads();
    }
  }
  }
}


// col_body with the cache policy CP, except that K_INV_CURL3's blocks of v
// and w (fields 1, 2) and of omega_x (3, which re-reads v and w) load
// through the caches when CP asks for nontemporal loads: only the blocks of
// u, omega_y and omega_z (0, 4, 5), whose inputs no other block reads, load
// nontemporally.  The field index is the fastest of the block order.
template <typename T, int L, int TC, int KIND, int PIPE, int NT, int RP, int CP>
__device__ __forceinline__ void col_run(const ColArgs& a, const cplx<T>* __restrict__ tw, cplx<T>* tws,
                                        cplx<T>* buf, long long bid, long long nblk) {
  if constexpr (KIND == K_INV_CURL3 && cp_load(CP)) {
    const int fg = (int)(xcd_order(bid, nblk) % a.n_fields);
    if (fg == 0 || fg >= 4) col_body<T, L, TC, KIND, PIPE, NT, RP, CP>(a, tw, tws, buf, bid, nblk);
    else col_body<T, L, TC, KIND, PIPE, NT, RP, CP & 1>(a, tw, tws, buf, bid, nblk);
  } else {
    col_body<T, L, TC, KIND, PIPE, NT, RP, CP>(a, tw, tws, buf, bid, nblk);
  }
}

template <typename T, int L, int TC, int KIND, int PIPE, int NT, int RP = 0, int CP = 0>
__global__ void __launch_bounds__(NT) col_pass(ColArgs a, const cplx<T>* __restrict__ tw) {
  __shared__ cplx<T> tws[col_lds_tw<T, L, TC, KIND, RP>()];
  __shared__ cplx<T> buf[col_lds_buf<T, L, TC, KIND>()];
  col_run<T, L, TC, KIND, PIPE, NT, RP, CP>(a, tw, tws, buf, blockIdx.x, gridDim.x);
}


// ---------------------------------------------------------- paired launches
// Two plans' launches of one pass kind in one grid, for ensembles whose
// members have different transform lengths: blocks [0, split) run segment 0
// at length L0, blocks [split, ...) segment 1 at L1, each exactly as its own
// launch would (the same body, arguments and block numbering).  split is
// segment 0's block count rounded up to a multiple of 8, so that both
// segments keep their XCD-aware tile order.  LDS is sized for the larger.
struct ColSeg {
  ColArgs a;
  const void* tw;
  long long nblk;
};

template <typename T, int L0, int TC0, int L1, int TC1, int KIND, int CP = 0>
__global__ void __launch_bounds__(kThreads) col_pass_pair(ColSeg s0, ColSeg s1, long long split) {
  constexpr int TW0 = col_lds_tw<T, L0, TC0, KIND>(), TW1 = col_lds_tw<T, L1, TC1, KIND>();
  constexpr int BF0 = col_lds_buf<T, L0, TC0, KIND>(), BF1 = col_lds_buf<T, L1, TC1, KIND>();
  __shared__ cplx<T> tws[TW0 > TW1 ? TW0 : TW1];
  __shared__ cplx<T> buf[BF0 > BF1 ? BF0 : BF1];
  const long long bid = blockIdx.x;
  if (bid < split) {
    if (bid < s0.nblk)
      col_run<T, L0, TC0, KIND, 1, kThreads, 0, CP>(s0.a, (const cplx<T>*)s0.tw, tws, buf, bid, s0.nblk);
  } else {
    col_run<T, L1, TC1, KIND, 1, kThreads, 0, CP>(s1.a, (const cplx<T>*)s1.tw, tws, buf, bid - split, s1.nblk);
  }
}

struct RowSeg {
  RowArgs a;
  const void* tw;
  int gx, gy;
};

template <typename T, int L0, int NP0, int L1, int NP1, int MODE, int CP = 0>
__global__ void __launch_bounds__(kThreads) row_pass_pair(RowSeg s0, RowSeg s1, long long split) {
  constexpr int TW0 = row_lds_tw<T, L0, MODE, NP0>(), TW1 = row_lds_tw<T, L1, MODE, NP1>();
  constexpr int BF0 = row_lds_buf<T, L0, MODE, NP0>(), BF1 = row_lds_buf<T, L1, MODE, NP1>();
  __shared__ cplx<T> tws[TW0 > TW1 ? TW0 : TW1];
  __shared__ cplx<T> buf[BF0 > BF1 ? BF0 : BF1];
  const long long bid = blockIdx.x;
  if (bid < split) {
    if (bid < (long long)s0.gx * s0.gy)
      row_body<T, L0, MODE, NP0, 0, CP>(s0.a, (const cplx<T>*)s0.tw, tws, buf, (int)(bid % s0.gx),
                                        (int)(bid / s0.gx));
  } else {
    const long long b1 = bid - split;
    row_body<T, L1, MODE, NP1, 0, CP>(s1.a, (const cplx<T>*)s1.tw, tws, buf, (int)(b1 % s1.gx), (int)(b1 / s1.gx));
  }
}

}  // namespace pf
