// Register butterflies and the LDS Stockham engine used by every FFT pass.
//
// A length-L transform (L = 2^a 3^b 5^c) runs as a sequence of radix passes
// (make_plan: radix 5 first, then 3, then 16/8/4/2, or fewer passes with the
// mixed radices 6, 9, 12, 24 where the length is 2^a 3^b).  Pass p with radix R and span Ns (product
// of the earlier radices) reads element j + r L/R, multiplies it by
// W_{Ns R}^{r (j mod Ns)}, does a radix-R DFT in registers and writes it to
// (j / Ns) Ns R + (j mod Ns) + r Ns: the Stockham auto-sort order, so the
// result comes out in natural order with no bit reversal.  Every thread
// reads all of its butterflies of a pass into registers before the block
// barrier and writes them after it, so a pass runs in place in LDS.
#pragma once
#include "pf_common.h"

namespace pf {

struct RadixPlan {
  int n;
  int r[6];
};

// Radix plan rp of length L.  rp = 0, the default: lengths with a factor 3
// whose radix-2/3/4/8/16 plan needs three or more passes use the mixed
// radices 6, 9, 12 and 24 to get by with two (three for 1536), one LDS round
// and barrier pair fewer per transform.  rp = 1: the radix-2/3/4/8/16 plan
// (radix 3 first), an alternative the autotuner times where it differs.
constexpr RadixPlan make_plan(int L, int rp = 0) {
  if (rp == 0) switch (L) {
    case 18: return RadixPlan{2, {9, 2, 0, 0, 0, 0}};
    case 36: return RadixPlan{2, {9, 4, 0, 0, 0, 0}};
    case 72: return RadixPlan{2, {9, 8, 0, 0, 0, 0}};
    case 96: return RadixPlan{2, {6, 16, 0, 0, 0, 0}};
    case 144: return RadixPlan{2, {9, 16, 0, 0, 0, 0}};
    case 192: return RadixPlan{2, {12, 16, 0, 0, 0, 0}};
    case 288: return RadixPlan{2, {12, 24, 0, 0, 0, 0}};
    case 384: return RadixPlan{2, {24, 16, 0, 0, 0, 0}};
    case 1536: return RadixPlan{3, {6, 16, 16, 0, 0, 0}};
    default: break;
  }
  RadixPlan p{0, {0, 0, 0, 0, 0, 0}};
  int a = 0, b = 0, c = 0, n = L;
  while (n % 2 == 0) { n /= 2; ++a; }
  while (n % 3 == 0) { n /= 3; ++b; }
  while (n % 5 == 0) { n /= 5; ++c; }
  for (int i = 0; i < c; ++i) p.r[p.n++] = 5;
  for (int i = 0; i < b; ++i) p.r[p.n++] = 3;
  switch (a) {
    case 1: p.r[p.n++] = 2; break;
    case 2: p.r[p.n++] = 4; break;
    case 3: p.r[p.n++] = 8; break;
    case 4: p.r[p.n++] = 16; break;
    case 5: p.r[p.n++] = 8; p.r[p.n++] = 4; break;
    case 6: p.r[p.n++] = 8; p.r[p.n++] = 8; break;
    case 7: p.r[p.n++] = 16; p.r[p.n++] = 8; break;
    case 8: p.r[p.n++] = 16; p.r[p.n++] = 16; break;
    case 9: p.r[p.n++] = 8; p.r[p.n++] = 8; p.r[p.n++] = 8; break;
    case 10: p.r[p.n++] = 16; p.r[p.n++] = 8; p.r[p.n++] = 8; break;
    case 11: p.r[p.n++] = 16; p.r[p.n++] = 16; p.r[p.n++] = 8; break;
    default: break;
  }
  return p;
}

// whether radix plan 1 of L differs from plan 0
constexpr bool has_alt_plan(int L) { return make_plan(L, 0).n != make_plan(L, 1).n; }

constexpr bool plan_ok(int L) {
  int n = L;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  int a = 0;
  for (int m = L; m % 2 == 0; m /= 2) ++a;
  return n == 1 && a >= 1 && a <= 11 && make_plan(L).n <= 6;
}

// cos(2 pi m / 16) and sin(2 pi m / 16), m = 0..15
constexpr double kCos16[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                               0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
constexpr double kSin16[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                               1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                               0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                               -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};
constexpr double kSqrtHalf = 0.70710678118654752440;
constexpr double kSqrt3Half = 0.86602540378443864676;
// cos and sin of 2 pi / 5 and 4 pi / 5
constexpr double kCos5a = 0.30901699437494742410, kCos5b = -0.80901699437494742410;
constexpr double kSin5a = 0.95105651629515357212, kSin5b = 0.58778525229247312917;

// multiply by W16^m = exp(S 2 pi i m / 16) for a compile-time m
template <int S, int M, typename T> __device__ __forceinline__ cplx<T> w16(cplx<T> a) {
  constexpr int m = M & 15;
  if constexpr (m == 0) return a;
  else if constexpr (m == 4) return mul_si<S, T>(a);
  else if constexpr (m == 8) return mk<T>(-a.x, -a.y);
  else if constexpr (m == 12) return mul_si<-S, T>(a);
  else {
    const T c = (T)kCos16[m], s = (T)(S * kSin16[m]);
    return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

template <int S, typename T> __device__ __forceinline__ void dft2(cplx<T>& a0, cplx<T>& a1) {
  cplx<T> t = a0;
  a0 = add<T>(t, a1);
  a1 = sub<T>(t, a1);
}

template <int S, typename T>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
  cplx<T> t0 = add<T>(a0, a2), t1 = sub<T>(a0, a2);
  cplx<T> t2 = add<T>(a1, a3), t3 = mul_si<S, T>(sub<T>(a1, a3));
  a0 = add<T>(t0, t2);
  a2 = sub<T>(t0, t2);
  a1 = add<T>(t1, t3);
  a3 = sub<T>(t1, t3);
}

template <int S, typename T> __device__ __forceinline__ void dft3(cplx<T>* v) {
  cplx<T> t = add<T>(v[1], v[2]);
  cplx<T> d = scl<T>(mul_si<S, T>(sub<T>(v[1], v[2])), (T)kSqrt3Half);
  cplx<T> m = mk<T>(v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y);
  v[0] = add<T>(v[0], t);
  v[1] = add<T>(m, d);
  v[2] = sub<T>(m, d);
}

// X_k = sum_n x_n w^(n k), w = exp(S 2 pi i / 5): the pairs (x1, x4) and
// (x2, x3) enter X_k and X_(5-k) through their sums (real cosines) and
// differences (imaginary sines)
template <int S, typename T> __device__ __forceinline__ void dft5(cplx<T>* v) {
  const cplx<T> t1 = add<T>(v[1], v[4]), t2 = add<T>(v[2], v[3]);
  const cplx<T> t3 = sub<T>(v[1], v[4]), t4 = sub<T>(v[2], v[3]);
  const T c1 = (T)kCos5a, c2 = (T)kCos5b, s1 = (T)kSin5a, s2 = (T)kSin5b;
  const cplx<T> a1 = mk<T>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  const cplx<T> a2 = mk<T>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  const cplx<T> b1 = mul_si<S, T>(mk<T>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const cplx<T> b2 = mul_si<S, T>(mk<T>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = add<T>(v[0], add<T>(t1, t2));
  v[1] = add<T>(a1, b1);
  v[4] = sub<T>(a1, b1);
  v[2] = add<T>(a2, b2);
  v[3] = sub<T>(a2, b2);
}

template <int S, typename T> __device__ __forceinline__ void dft8(cplx<T>* v) {
  // even / odd split, two radix-4 DFTs, then W8^k on the odd half
  dft4<S, T>(v[0], v[2], v[4], v[6]);
  dft4<S, T>(v[1], v[3], v[5], v[7]);
  cplx<T> o1 = w16<S, 2, T>(v[3]), o2 = w16<S, 4, T>(v[5]), o3 = w16<S, 6, T>(v[7]);
  cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = add<T>(e0, o0); v[4] = sub<T>(e0, o0);
  v[1] = add<T>(e1, o1); v[5] = sub<T>(e1, o1);
  v[2] = add<T>(e2, o2); v[6] = sub<T>(e2, o2);
  v[3] = add<T>(e3, o3); v[7] = sub<T>(e3, o3);
}

template <int S, typename T> __device__ __forceinline__ void dft16(cplx<T>* v) {
  // 4 x 4: Z_r = DFT4(a_r, a_{r+4}, a_{r+8}, a_{r+12}); twiddle W16^{r k'};
  // y_{k' + 4 q} = DFT4 over r.
  dft4<S, T>(v[0], v[4], v[8], v[12]);
  dft4<S, T>(v[1], v[5], v[9], v[13]);
  dft4<S, T>(v[2], v[6], v[10], v[14]);
  dft4<S, T>(v[3], v[7], v[11], v[15]);
  // v[r + 4 k'] holds Z_r[k']
  v[5] = w16<S, 1, T>(v[5]);   v[9] = w16<S, 2, T>(v[9]);   v[13] = w16<S, 3, T>(v[13]);
  v[6] = w16<S, 2, T>(v[6]);   v[10] = w16<S, 4, T>(v[10]); v[14] = w16<S, 6, T>(v[14]);
  v[7] = w16<S, 3, T>(v[7]);   v[11] = w16<S, 6, T>(v[11]); v[15] = w16<S, 9, T>(v[15]);
  dft4<S, T>(v[0], v[1], v[2], v[3]);
  dft4<S, T>(v[4], v[5], v[6], v[7]);
  dft4<S, T>(v[8], v[9], v[10], v[11]);
  dft4<S, T>(v[12], v[13], v[14], v[15]);
  // now v[4 k' + q] holds y_{k' + 4 q}; transpose the 4 x 4 index
  cplx<T> t;
  t = v[1]; v[1] = v[4]; v[4] = t;
  t = v[2]; v[2] = v[8]; v[8] = t;
  t = v[3]; v[3] = v[12]; v[12] = t;
  t = v[6]; v[6] = v[9]; v[9] = t;
  t = v[7]; v[7] = v[13]; v[13] = t;
  t = v[11]; v[11] = v[14]; v[14] = t;
}

// cos(2 pi m / 72) for m = 0..18 (a quarter turn); cos72 / sin72 extend it
// by symmetry to every m.  72 is a multiple of the radices 6, 9, 12, 24.
constexpr double kCos72[19] = {1.0, 0.996194698091745532295, 0.984807753012208059367, 0.96592582628906828675,
                               0.939692620785908384054, 0.906307787036649963243, 0.866025403784438646764,
                               0.819152044288991789684, 0.766044443118978035202, 0.707106781186547524401,
                               0.642787609686539326323, 0.573576436351046096108, 0.5, 0.422618261740699436187,
                               0.342020143325668733044, 0.258819045102520762349, 0.173648177666930348852,
                               0.0871557427476581735581, 0.0};
__host__ __device__ constexpr double cos72(int m) {
  m = ((m % 72) + 72) % 72;
  return m <= 18 ? kCos72[m] : (m <= 36 ? -kCos72[36 - m] : (m <= 54 ? -kCos72[m - 36] : kCos72[72 - m]));
}
__host__ __device__ constexpr double sin72(int m) { return cos72(m - 18); }

// multiply by exp(S 2 pi i m / R) for R dividing 72; m is a constant once the
// caller's loops are unrolled, so the branches and table reads fold away
template <int S, int R, typename T> __device__ __forceinline__ cplx<T> wr(cplx<T> a, int m) {
  static_assert(72 % R == 0, "wr: R must divide 72");
  const int q = ((m % R) + R) % R * (72 / R);
  if (q == 0) return a;
  if (q == 18) return mul_si<S, T>(a);
  if (q == 36) return mk<T>(-a.x, -a.y);
  if (q == 54) return mul_si<-S, T>(a);
  const T c = (T)cos72(q), sn = (T)(S * sin72(q));
  return mk<T>(a.x * c - a.y * sn, a.x * sn + a.y * c);
}

template <int R, int S, typename T> __device__ __forceinline__ void dft(cplx<T>* v);

// DFT of size R1 R2 by one Cooley-Tukey split: input n = R2 n1 + n2, output
// k = k1 + R1 k2; inner DFTs of size R1 over n1, twiddles w_R^(n2 k1), outer
// DFTs of size R2 over n2, results in natural order.
template <int R1, int R2, int S, typename T> __device__ __forceinline__ void dft_ct(cplx<T>* v) {
  constexpr int R = R1 * R2;
  cplx<T> y[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; ++n2) {
    cplx<T> t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[R2 * n1 + n2];
    dft<R1, S, T>(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) y[n2 * R1 + k1] = wr<S, R, T>(t[k1], n2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cplx<T> t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2 * R1 + k1];
    dft<R2, S, T>(t);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
  }
}

template <int R, int S, typename T> __device__ __forceinline__ void dft(cplx<T>* v) {
  if constexpr (R == 2) dft2<S, T>(v[0], v[1]);
  else if constexpr (R == 3) dft3<S, T>(v);
  else if constexpr (R == 5) dft5<S, T>(v);
  else if constexpr (R == 4) dft4<S, T>(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8<S, T>(v);
  else if constexpr (R == 16) dft16<S, T>(v);
  else if constexpr (R == 6) dft_ct<3, 2, S, T>(v);
  else if constexpr (R == 9) dft_ct<3, 3, S, T>(v);
  else if constexpr (R == 12) dft_ct<3, 4, S, T>(v);
  else if constexpr (R == 24) dft_ct<3, 8, S, T>(v);
  else static_assert(R == 2, "unsupported radix");
}

// Twiddle table, one block per pass p >= 1 (pass 0 has span 1 and none):
// entry OFF(p) + (r - 1) Ns + k = exp(-2 pi i r k / (Ns R)).  Lanes that
// take consecutive butterflies read consecutive entries.
constexpr int tw_offset(int L, int P, int rp = 0) {
  RadixPlan pl = make_plan(L, rp);
  int off = 0, ns = 1;
  for (int q = 0; q < P; ++q) {
    if (q > 0) off += (pl.r[q] - 1) * ns;
    ns *= pl.r[q];
  }
  return off;
}
constexpr int tw_size(int L, int rp = 0) { return tw_offset(L, make_plan(L, rp).n, rp); }
constexpr int tw_alloc(int L, int rp = 0) { return tw_size(L, rp) > 0 ? tw_size(L, rp) : 1; }

// conjugated for the inverse
template <int S, typename T> __device__ __forceinline__ cplx<T> twiddle(const cplx<T>* tw, int m) {
  cplx<T> w = tw[m];
  if constexpr (S > 0) w.y = -w.y;
  return w;
}

// Whether every transform of a contiguous-row FFT (consecutive threads on
// consecutive butterflies of one transform) stays within one wavefront from
// pass to pass: every pass of the plan has the same butterfly count L / R,
// and it divides the wavefront, so transform f is always handled by the
// same aligned group of L / R lanes.  Its passes then exchange data through
// LDS with wavefront-level synchronisation instead of block barriers.
constexpr bool wave_local_plan(int L, int rp = 0) {
  const RadixPlan pl = make_plan(L, rp);
  for (int q = 0; q < pl.n; ++q)
    if (pl.r[q] != pl.r[0]) return false;
  const int nb = L / pl.r[0];
  return 64 % nb == 0;
}

// The synchronisation between the read and write halves of a pass and
// between passes: a block barrier, or (WL) a wavefront barrier with
// release / acquire fences at wavefront scope, which orders one wave's LDS
// accesses across lanes.
template <bool WL> __device__ __forceinline__ void fft_sync() {
  if constexpr (WL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham pass and the passes after it (compile-time recursion).
// Addr(f, e) gives the LDS index of element e of transform f.  FFAST maps
// consecutive threads to consecutive transforms (column tiles); otherwise
// to consecutive butterflies of one transform (contiguous rows).
// PEND: stop before pass PEND (the caller runs the remaining passes).
// RP: the radix plan (make_plan).  WL: wavefront-level synchronisation
// inside the transform (contiguous rows of a wave_local_plan only); the
// last pass still ends with a block barrier.
template <typename T, int L, int S, int NF, int NT, int P, int NS, bool FFAST, class Addr, int PEND = 99, int RP = 0,
          bool WL = false>
__device__ __forceinline__ void stockham(cplx<T>* buf, const cplx<T>* tw, const Addr& addr) {
  static_assert(!WL || (!FFAST && wave_local_plan(L, RP) && NT % 64 == 0), "transforms must stay within a wave");
  constexpr RadixPlan pl = make_plan(L, RP);
  if constexpr (P < pl.n && P < PEND) {
    constexpr int R = pl.r[P];
    constexpr int NB = L / R;
    constexpr int TOT = NF * NB;
    constexpr int K = (TOT + NT - 1) / NT;
    cplx<T> v[K][R];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if (TOT % NT == 0 || q < TOT) {
        const int f = FFAST ? q % NF : q / NB;
        const int j = FFAST ? q / NF : q % NB;
        const int k = j % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[kk][r] = buf[addr(f, j + r * NB)];
        if constexpr (NS > 1) {
          constexpr int OFF = tw_offset(L, P, RP);
#pragma unroll
          for (int r = 1; r < R; ++r) v[kk][r] = mul<T>(v[kk][r], twiddle<S, T>(tw, OFF + (r - 1) * NS + k));
        }
        dft<R, S, T>(v[kk]);
      }
    }
    fft_sync<WL>();
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const int q = threadIdx.x + kk * NT;
      if (TOT % NT == 0 || q < TOT) {
        const int f = FFAST ? q % NF : q / NB;
        const int j = FFAST ? q / NF : q % NB;
        const int k = j % NS;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[addr(f, base + r * NS)] = v[kk][r];
      }
    }
    if constexpr (P + 1 < pl.n && P + 1 < PEND) fft_sync<WL>();
    else __syncthreads();
    stockham<T, L, S, NF, NT, P + 1, NS * R, FFAST, Addr, PEND, RP, WL>(buf, tw, addr);
  }
}

// NF independent length-L transforms held in LDS.  Callers synchronise the
// block before the call; the call ends with a block barrier.
template <typename T, int L, int S, int NF, int NT, bool FFAST, int RP = 0, bool WL = false, class Addr>
__device__ __forceinline__ void lds_fft(cplx<T>* buf, const cplx<T>* tw, const Addr& addr) {
  static_assert(plan_ok(L), "FFT length must be 2^a 3^b 5^c with 1 <= a <= 11");
  stockham<T, L, S, NF, NT, 0, 1, FFAST, Addr, 99, RP, WL>(buf, tw, addr);
}

// All passes but the last (the caller runs it); ends with a block barrier
// when any pass ran.
template <typename T, int L, int S, int NF, int NT, int RP = 0, class Addr>
__device__ __forceinline__ void lds_fft_head(cplx<T>* buf, const cplx<T>* tw, const Addr& addr) {
  static_assert(plan_ok(L), "FFT length must be 2^a 3^b 5^c with 1 <= a <= 11");
  stockham<T, L, S, NF, NT, 0, 1, true, Addr, make_plan(L, RP).n - 1, RP>(buf, tw, addr);
}

}  // namespace pf
