// Kernel templates of the Fourier-Galerkin right-hand side and time step.
//
// HBM layout (C order, complex interleaved, one block per member):
//   state u_hat         (B, d, N0, [N1,] nh)      nh = N_last/2 + 1
//   3D nonlinear passes (per member, F = 6 for u_hat and omega_hat):
//     A  (F, M0, N1, nh)   axis-0 inverse, padded       <- col_pass CURL3
//     B  (F, M0, M1, nh)   axis-1 inverse, padded       <- col_pass
//     C  (3, M0, M1, nh)   c2r, u x omega, r2c, cut     <- row_pass FUSED3
//     D  (3, M0, N1, nh)   axis-1 forward, truncated    <- col_pass
//     R  (3, N0, N1, nh)   axis-0 forward, truncated    <- col_pass
//   2D: A (3, M0, nh) -> C (2, M0, nh) -> R (2, N0, nh)
// The padded physical fields never reach HBM: the last axis is taken to
// physical space, multiplied and brought back inside one block.
#pragma once
#include "pf_fft.h"

namespace pf {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- column pass
enum ColMode { COL_PLAIN = 0, COL_CURL3 = 1, COL_CURL2 = 2 };

struct ColArgs {
  const void* in;
  void* out;
  int rows_in, rows_out;        // rows read / written per matrix
  int cols;                     // contiguous columns per matrix
  int n_tiles;                  // column tiles per matrix
  int n_outer;                  // matrices per field
  long long in_row, out_row;    // element strides between rows
  long long in_outer, out_outer;
  long long in_field, out_field;
  long long in_member, out_member;
  int n_fields;                 // output fields per member
  int n_members;
  int inverse;                  // 1: zero-pad rows_in -> L, exp(+i..); 0: exp(-i..), truncate L -> rows_out
  double scale;
  int mode;
  int gN0, gN1, gN2, gnh;       // unpadded grid (curl modes)
  double ks0, ks1, ks2;         // wavenumber scales 2 pi / L_i
  int mask23;                   // 2/3 rule mask while loading (curl modes)
};

template <typename T> constexpr int col_tc(int L) {
  return sizeof(T) == 4 ? (L <= 128 ? 32 : (L <= 384 ? 16 : 8)) : (L <= 128 ? 16 : (L <= 384 ? 8 : 4));
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

template <typename T, int L, int TC>
__global__ void __launch_bounds__(kThreads) col_pass(ColArgs a, const cplx<T>* __restrict__ tw) {
  constexpr int TCP = TC;
  __shared__ cplx<T> tws[L];
  __shared__ cplx<T> buf[L * TCP];
  for (int m = threadIdx.x; m < L; m += kThreads) tws[m] = tw[m];

  const int tile = blockIdx.x % a.n_tiles;
  const int outer = blockIdx.x / a.n_tiles;
  const int f = blockIdx.y % a.n_fields;
  const int b = blockIdx.y / a.n_fields;
  const int c0 = tile * TC;
  const cplx<T>* in = (const cplx<T>*)a.in + (long long)b * a.in_member + (long long)outer * a.in_outer;
  cplx<T>* out = (cplx<T>*)a.out + (long long)b * a.out_member + (long long)f * a.out_field +
                 (long long)outer * a.out_outer;

  for (int idx = threadIdx.x; idx < L * TC; idx += kThreads) {
    const int e = idx / TC, c = idx % TC;
    const int col = c0 + c;
    cplx<T> v = mk<T>(0, 0);
    if (col < a.cols) {
      int src;
      double fac = 1.0;
      if (a.inverse) src = pad_src(e, a.rows_in, L, fac);
      else src = e;
      if (src >= 0) {
        if (a.mode == COL_PLAIN) {
          v = in[(long long)f * a.in_field + (long long)src * a.in_row + col];
        } else if (a.mode == COL_CURL3) {
          const int j = col / a.gnh, l = col % a.gnh;
          const int mi = cmode(src, a.gN0), mj = cmode(j, a.gN1);
          bool keep = true;
          if (a.mask23) keep = abs(mi) <= a.gN0 / 3 && abs(mj) <= a.gN1 / 3 && l <= a.gN2 / 3;
          if (keep) {
            const cplx<T>* p = in + (long long)src * a.in_row + col;
            if (f < 3) {
              v = p[(long long)f * a.in_field];
            } else {
              const T k0 = (T)(mi * a.ks0), k1 = (T)(mj * a.ks1), k2 = (T)(l * a.ks2);
              const cplx<T> u0 = p[0], u1 = p[a.in_field], u2 = p[2 * a.in_field];
              cplx<T> w;
              if (f == 3) w = mk<T>(k1 * u2.x - k2 * u1.x, k1 * u2.y - k2 * u1.y);
              else if (f == 4) w = mk<T>(k2 * u0.x - k0 * u2.x, k2 * u0.y - k0 * u2.y);
              else w = mk<T>(k0 * u1.x - k1 * u0.x, k0 * u1.y - k1 * u0.y);
              v = mul_i<T>(w);
            }
          }
        } else {  // COL_CURL2
          const int l = col;
          const int mi = cmode(src, a.gN0);
          bool keep = true;
          if (a.mask23) keep = abs(mi) <= a.gN0 / 3 && l <= a.gN1 / 3;
          if (keep) {
            const cplx<T>* p = in + (long long)src * a.in_row + col;
            if (f < 2) {
              v = p[(long long)f * a.in_field];
            } else {
              const T k0 = (T)(mi * a.ks0), k1 = (T)(l * a.ks1);
              const cplx<T> u0 = p[0], u1 = p[a.in_field];
              v = mul_i<T>(mk<T>(k0 * u1.x - k1 * u0.x, k0 * u1.y - k1 * u0.y));
            }
          }
        }
        if (fac != 1.0) v = scl<T>(v, (T)fac);
      }
    }
    buf[e * TCP + c] = v;
  }
  __syncthreads();

  auto addr = [](int ff, int e) { return e * TCP + ff; };
  if (a.inverse) lds_fft<T, L, +1, TC, kThreads, true>(buf, tws, addr);
  else lds_fft<T, L, -1, TC, kThreads, true>(buf, tws, addr);

  const T s = (T)a.scale;
  if (a.inverse) {
    for (int idx = threadIdx.x; idx < L * TC; idx += kThreads) {
      const int e = idx / TC, c = idx % TC;
      const int col = c0 + c;
      if (col < a.cols) out[(long long)e * a.out_row + col] = scl<T>(buf[e * TCP + c], s);
    }
  } else {
    const int n = a.rows_out, h = n / 2;
    for (int idx = threadIdx.x; idx < n * TC; idx += kThreads) {
      const int o = idx / TC, c = idx % TC;
      const int col = c0 + c;
      if (col >= a.cols) continue;
      cplx<T> v;
      if (n == L || o < h) v = buf[o * TCP + c];
      else if (o == h) v = add<T>(buf[h * TCP + c], buf[(L - h) * TCP + c]);
      else v = buf[(o + (L - n)) * TCP + c];
      out[(long long)o * a.out_row + col] = scl<T>(v, s);
    }
  }
}

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
__host__ __device__ constexpr int row_pitch(int L) { return L + L / 16 + 1; }

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

template <typename T, int L, int MODE>
__global__ void __launch_bounds__(kThreads) row_pass(RowArgs a, const cplx<T>* __restrict__ tw) {
  constexpr int NS = row_slots<MODE>();
  constexpr int NP = row_pairs<T, MODE>(L);
  constexpr int LP = row_pitch(L);
  __shared__ cplx<T> tws[L];
  __shared__ cplx<T> buf[NP * NS * LP];
  for (int m = threadIdx.x; m < L; m += kThreads) tws[m] = tw[m];

  const int npairs = a.n_rows / 2;
  const int pair0 = blockIdx.x * NP;
  int f = 0, b;
  if (MODE == ROW_C2R || MODE == ROW_R2C) { f = blockIdx.y % a.n_fields; b = blockIdx.y / a.n_fields; }
  else b = blockIdx.y;
  auto ix = [](int slot, int e) { return slot * LP + e + (e >> 4); };
  const int N = a.N;

  // ---- load
  if constexpr (MODE == ROW_R2C) {
    const T* in = (const T*)a.in + (long long)b * a.in_member + (long long)f * a.in_field;
    for (int t = threadIdx.x; t < NP * L; t += kThreads) {
      const int p = t / L, x = t % L;
      const int pr = pair0 + p;
      cplx<T> z = mk<T>(0, 0);
      if (pr < npairs) {
        const long long r0 = 2LL * pr;
        z = mk<T>(in[r0 * a.in_row + x], in[(r0 + 1) * a.in_row + x]);
      }
      buf[ix(p, x)] = z;
    }
  } else {
    const cplx<T>* in = (const cplx<T>*)a.in + (long long)b * a.in_member;
    for (int t = threadIdx.x; t < NP * NS * L; t += kThreads) {
      const int p = t / (NS * L), rem = t % (NS * L);
      const int s = rem / L, k = rem % L;
      const int pr = pair0 + p;
      cplx<T> z = mk<T>(0, 0);
      if (pr < npairs) {
        const long long r0 = 2LL * pr, r1 = r0 + 1;
        cplx<T> X, Y;
        if constexpr (MODE == ROW_FUSED3) {
          const int rr = s / 3, q = s % 3;
          const long long r = r0 + rr;
          X = hext<T>(in + (2 * q) * a.in_field + r * a.in_row, k, L, N, a.mask23);
          Y = hext<T>(in + (2 * q + 1) * a.in_field + r * a.in_row, k, L, N, a.mask23);
        } else if constexpr (MODE == ROW_FUSED2) {
          if (s == 0) {
            X = hext<T>(in + 0 * a.in_field + r0 * a.in_row, k, L, N, a.mask23);
            Y = hext<T>(in + 1 * a.in_field + r0 * a.in_row, k, L, N, a.mask23);
          } else if (s == 1) {
            X = hext<T>(in + 2 * a.in_field + r0 * a.in_row, k, L, N, a.mask23);
            Y = hext<T>(in + 2 * a.in_field + r1 * a.in_row, k, L, N, a.mask23);
          } else {
            X = hext<T>(in + 0 * a.in_field + r1 * a.in_row, k, L, N, a.mask23);
            Y = hext<T>(in + 1 * a.in_field + r1 * a.in_row, k, L, N, a.mask23);
          }
        } else {  // ROW_C2R
          const cplx<T>* fin = in + (long long)f * a.in_field;
          X = hext<T>(fin + r0 * a.in_row, k, L, N, a.mask23);
          Y = hext<T>(fin + r1 * a.in_row, k, L, N, a.mask23);
        }
        z = pack<T>(X, Y);
      }
      buf[ix(p * NS + s, k)] = z;
    }
  }
  __syncthreads();

  if constexpr (MODE == ROW_R2C) {
    lds_fft<T, L, -1, NP, kThreads, false>(buf, tws, ix);
  } else {
    lds_fft<T, L, +1, NP * NS, kThreads, false>(buf, tws, ix);
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
      return slot * LP + e + (e >> 4);
    };
    lds_fft<T, L, -1, NP * 3, kThreads, false>(buf, tws, fx);
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
      return (p * 3 + 2 * w) * LP + e + (e >> 4);
    };
    lds_fft<T, L, -1, NP * 2, kThreads, false>(buf, tws, fx);
  }

  // ---- split the packed spectra, truncate, scale, store
  constexpr int NFW = MODE == ROW_FUSED3 ? 3 : (MODE == ROW_FUSED2 ? 2 : 1);
  const int nh = a.nh;
  const T s = (T)a.scale;
  cplx<T>* out = (cplx<T>*)a.out + (long long)b * a.out_member;
  for (int t = threadIdx.x; t < NP * NFW * nh; t += kThreads) {
    const int p = t / (NFW * nh), rem = t % (NFW * nh);
    const int w = rem / nh, k = rem % nh;
    const int pr = pair0 + p;
    if (pr >= npairs) continue;
    int slot;
    if (MODE == ROW_FUSED3) slot = p * 6 + (w == 0 ? 0 : (w == 1 ? 3 : 1));
    else if (MODE == ROW_FUSED2) slot = p * 3 + 2 * w;
    else slot = p;
    const cplx<T> yk = buf[ix(slot, k)];
    const cplx<T> ym = buf[ix(slot, k == 0 ? 0 : L - k)];
    cplx<T> A = mk<T>((T)0.5 * (yk.x + ym.x), (T)0.5 * (yk.y - ym.y));
    cplx<T> B = mk<T>((T)0.5 * (yk.y + ym.y), (T)-0.5 * (yk.x - ym.x));
    if (L != N && k == N / 2) { A = mk<T>(2 * A.x, 0); B = mk<T>(2 * B.x, 0); }
    A = scl<T>(A, s);
    B = scl<T>(B, s);
    const long long r0 = 2LL * pr, r1 = r0 + 1;
    if constexpr (MODE == ROW_FUSED3) {
      if (w == 0) { out[r0 * a.out_row + k] = A; out[a.out_field + r0 * a.out_row + k] = B; }
      else if (w == 1) { out[r1 * a.out_row + k] = A; out[a.out_field + r1 * a.out_row + k] = B; }
      else { out[2 * a.out_field + r0 * a.out_row + k] = A; out[2 * a.out_field + r1 * a.out_row + k] = B; }
    } else if constexpr (MODE == ROW_FUSED2) {
      const long long r = w == 0 ? r0 : r1;
      out[r * a.out_row + k] = A;
      out[a.out_field + r * a.out_row + k] = B;
    } else {
      cplx<T>* fo = out + (long long)f * a.out_field;
      fo[r0 * a.out_row + k] = A;
      fo[r1 * a.out_row + k] = B;
    }
  }
}

// -------------------------------------------------------------- random numbers
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the Box-Muller map.
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__device__ __forceinline__ void sc_pi(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ __forceinline__ void sc_pi(double x, double* s, double* c) { sincospi(x, s, c); }
__device__ __forceinline__ float ln(float x) { return logf(x); }
__device__ __forceinline__ double ln(double x) { return log(x); }

template <typename T> __device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, T& g0, T& g1) {
  // u1 in (0, 1], u2 in [0, 1)
  const T u1 = ((T)a + (T)0.5) * (T)2.3283064365386963e-10;
  const T u2 = (T)b * (T)2.3283064365386963e-10;
  const T r = sqrt((T)-2 * ln(u1));
  T sn, cs;
  sc_pi((T)2 * u2, &sn, &cs);
  g0 = r * cs;
  g1 = r * sn;
}

// raw draws w1[c], w2[c] of mode idx in a step: four normals per call
template <typename T>
__device__ __forceinline__ void rng_draw(unsigned long long seed, long long step, long long idx, int c,
                                         cplx<T>& w1, cplx<T>& w2) {
  U4 ctr{(uint32_t)idx, (uint32_t)((unsigned long long)idx >> 32) ^ ((uint32_t)c << 24), (uint32_t)step,
         (uint32_t)((unsigned long long)step >> 32)};
  const U4 r = philox(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
  T g0, g1, g2, g3;
  box_muller<T>(r.x, r.y, g0, g1);
  box_muller<T>(r.z, r.w, g2, g3);
  w1 = mk<T>(g0, g1);
  w2 = mk<T>(g2, g3);
}

// raw draws (B, 2, nc, nk) of one step: the device counterpart of random_fields
template <typename T>
__global__ void __launch_bounds__(256) draw_kernel(void* out, const unsigned long long* seeds, long long step,
                                                   long long nk, int nc) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= nk) return;
  cplx<T>* o = (cplx<T>*)out + (long long)b * 2 * nc * nk + idx;
  for (int c = 0; c < nc; ++c) {
    cplx<T> w1, w2;
    rng_draw<T>(seeds[b], step, idx, c, w1, w2);
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
  double noise_a, noise_mag, noise_norm;
  int mask_nyq;
  double fk2min, fk2max;
  long long step;
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

template <typename T, int D>
__device__ __forceinline__ Mode<T, D> decode(const StageArgs& a, long long idx) {
  Mode<T, D> m;
  const int l = (int)(idx % a.nh);
  const long long t = idx / a.nh;
  if constexpr (D == 3) {
    const int j = (int)(t % a.N1), i = (int)(t / a.N1);
    m.k[0] = (T)(cmode(i, a.N0) * a.ks0);
    m.k[1] = (T)(cmode(j, a.N1) * a.ks1);
    m.k[2] = (T)(l * a.ks2);
    m.nyq = (a.N0 % 2 == 0 && i == a.N0 / 2) || (a.N1 % 2 == 0 && j == a.N1 / 2) ||
            (a.N2 % 2 == 0 && l == a.N2 / 2);
  } else {
    const int i = (int)t;
    m.k[0] = (T)(cmode(i, a.N0) * a.ks0);
    m.k[1] = (T)(l * a.ks1);
    m.nyq = (a.N0 % 2 == 0 && i == a.N0 / 2) || (a.N1 % 2 == 0 && l == a.N1 / 2);
  }
  m.k2 = 0;
#pragma unroll
  for (int i = 0; i < D; ++i) m.k2 += m.k[i] * m.k[i];
  return m;
}

template <typename T, int D> __device__ __forceinline__ void project(cplx<T>* v, const Mode<T, D>& m) {
  const T den = m.k2 == (T)0 ? (T)1 : m.k2;
  cplx<T> p = mk<T>(0, 0);
#pragma unroll
  for (int i = 0; i < D; ++i) p = add<T>(p, scl<T>(v[i], m.k[i] / den));
#pragma unroll
  for (int i = 0; i < D; ++i) v[i] = sub<T>(v[i], scl<T>(p, m.k[i]));
}

template <typename T, int D>
__global__ void __launch_bounds__(kThreads) stage_kernel(StageArgs a) {
  const long long idx = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= a.nk) return;
  const Mode<T, D> m = decode<T, D>(a, idx);
  const long long mb = (long long)b * D * a.nk;
  cplx<T>* uh = (cplx<T>*)a.uhat + mb + idx;
  cplx<T>* u0p = (cplx<T>*)a.u0 + mb + idx;

  cplx<T> u[D], uz[D], nl[D], nz[D], fr[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    u[i] = uh[i * a.nk];
    uz[i] = (a.save_u0 || a.mode == 1) ? u[i] : u0p[i * a.nk];
    nl[i] = mk<T>(0, 0);
    nz[i] = mk<T>(0, 0);
    fr[i] = mk<T>(0, 0);
  }
  if (a.save_u0 && a.mode == 0) {
#pragma unroll
    for (int i = 0; i < D; ++i) u0p[i * a.nk] = u[i];
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

  // nonlinear term: Leray projection (math.py:72-100) and optional Nyquist mask
  if (a.rhs_nl) {
    const cplx<T>* r = (const cplx<T>*)a.rhs_nl + mb + idx;
#pragma unroll
    for (int i = 0; i < D; ++i) nl[i] = r[i * a.nk];
    project<T, D>(nl, m);
    if (a.mask_nyq && m.nyq) {
#pragma unroll
      for (int i = 0; i < D; ++i) nl[i] = mk<T>(0, 0);
    }
  }

  // noise (space_solver.py:227-265)
  if (a.noise_on) {
    const T fac = m.k2 == (T)0 ? (T)0 : (T)a.noise_norm;
    const T na = (T)a.noise_a;
    const unsigned long long seed = a.use_rng ? a.seeds[b] : 0ull;
    const int nc = a.noise_type == 0 ? D * D : D;
    const cplx<T>* raw = a.noise ? (const cplx<T>*)a.noise + (long long)b * 2 * nc * a.nk + idx : nullptr;
    cplx<T> w[D * D];
#pragma unroll
    for (int c = 0; c < D * D; ++c) {
      if (c < nc) {
        cplx<T> w1, w2;
        if (raw) { w1 = raw[(long long)c * a.nk]; w2 = raw[(long long)(nc + c) * a.nk]; }
        else rng_draw<T>(seed, a.step, idx, c, w1, w2);
        w[c] = add<T>(scl<T>(w1, fac), scl<T>(scl<T>(w2, fac), na));
      }
    }
    if (a.noise_type == 0) {
      const T h = (T)0.70710678118654752440;
#pragma unroll
      for (int i = 0; i < D; ++i) {
        cplx<T> acc = mk<T>(0, 0);
#pragma unroll
        for (int j = 0; j < D; ++j) acc = add<T>(acc, scl<T>(add<T>(w[i * D + j], w[j * D + i]), h * m.k[j]));
        nz[i] = mul_i<T>(acc);
      }
    } else {
      const cplx<T> C = ((const cplx<T>*)a.corr)[idx];
#pragma unroll
      for (int i = 0; i < D; ++i) nz[i] = mul<T>(C, w[i]);
    }
    project<T, D>(nz, m);
    const T mag = (T)a.noise_mag;
#pragma unroll
    for (int i = 0; i < D; ++i) nz[i] = scl<T>(nz[i], mag);
  }

  // constant-power band forcing
  if (a.fcoef && m.k2 >= (T)a.fk2min && m.k2 <= (T)a.fk2max) {
    const T fc = (T)a.fcoef[b];
#pragma unroll
    for (int i = 0; i < D; ++i) fr[i] = scl<T>(u[i], fc);
  }

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
    uh[i * a.nk] = r;
  }
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

}  // namespace pf
