// C ABI: plans, workspace, and the per-stage launch sequence.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/periodicflow.h"
#include "pf_launch.h"

using namespace pf;

namespace {

// the last error of the calling thread (a fixed buffer: reporting an error
// allocates nothing)
thread_local char g_err[512];

int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}
int fail(int code, const std::string& msg) { return fail(code, msg.c_str()); }
__attribute__((format(printf, 2, 3))) int failf(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* what) { return failf(PF_EHIP, "%s: %s", what, hipGetErrorString(e)); }

#define PF_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

#define PF_TRY(call)              \
  do {                            \
    int rc_ = (call);             \
    if (rc_) return launch_fail(rc_, #call); \
  } while (0)

int launch_fail(int rc, const char* what) {
  if (rc == -4) return failf(PF_EUNSUPPORTED, "unsupported transform length in %s", what);
  if (rc == -3) return failf(PF_EHIP, "kernel launch failed: %s: %s", what, hipGetErrorString(hipGetLastError()));
  return failf(PF_EINVAL, "invalid launch: %s", what);
}

}  // namespace

struct pf_plan {
  pf_desc d{};
  int dim = 0;
  int N[3] = {1, 1, 1};
  int M[3] = {1, 1, 1};
  int nh = 0;
  long long pitch = 0;           // row pitch of the work buffers (128-byte multiple)
  long long dplane = 0;          // 3D: stride of D's axis-0 planes, N1 P + one 128-byte line
  long long aplane = 0;          // 3D: stride of A's axis-0 planes, N1 P + one 128-byte line
  long long nk = 0, nphys = 0;
  int B = 1;
  bool fp64 = false;
  size_t rsz = 4, csz = 8;
  int nc = 0;
  bool direct23 = false;
  double ks[3] = {1, 1, 1};
  void* u0 = nullptr;
  void* w1 = nullptr;
  void* w2 = nullptr;
  size_t w1_bytes = 0, w2_bytes = 0;
  void* lin_op = nullptr;
  void* corr = nullptr;
  unsigned long long* seeds = nullptr;
  long long* msteps = nullptr;   // per-member RNG step counters (device)
  int nact = 1;                  // members [0, nact) are advanced
  double* fcoef = nullptr;
  // twiddle tables of the plan's distinct transform lengths (at most 2 per
  // axis) and radix plans (make_plan: 0, and 1 where the length has one)
  struct Tw { int L, rp; void* ptr; } tw[12] = {};
  int ntw = 0;
  long long step = 0;
  size_t bytes = 0;
  std::vector<void*> owned;
  // per-pass timing with HIP events (pf_profile_enable)
  bool prof = false;
  uint32_t prof_mask = 0xffffffffu;  // passes that record events (pf_profile_select)
  bool fuse_stage = true;   // axis-0 forward pass fused with the stage update
  int tsel[PF_NPASS] = {};  // launch variant per pass (pf_set_tiles, pf_autotune)
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[PF_NPASS];
  struct LaunchLog* rec = nullptr;  // while set, pass launches are recorded instead of issued
};

pf::PendingEvents& pf::pending_events() {
  static thread_local PendingEvents pe;
  return pe;
}

namespace {

// Times one pass with a pair of events when profiling.  A pass that is one
// kernel launch (kernel = true) hands the events to that launch, which
// records them at the kernel's start and end; otherwise (or when the pass
// launched nothing, e.g. while recording) they are recorded on the stream
// around the pass.
struct PassTimer {
  pf_plan* p;
  int id;
  hipStream_t s;
  bool kernel;
  hipEvent_t a = nullptr, b = nullptr;
  PassTimer(pf_plan* p_, int id_, hipStream_t s_, bool kernel_ = true) : p(p_), id(id_), s(s_), kernel(kernel_) {
    if (p && p->prof && ((p->prof_mask >> id) & 1u) && hipEventCreate(&a) == hipSuccess &&
        hipEventCreate(&b) == hipSuccess) {
      if (kernel) pending_events() = {a, b};
      else (void)hipEventRecord(a, s);
    } else {
      a = b = nullptr;
    }
  }
  ~PassTimer() {
    if (a && b) {
      bool by_kernel = false;
      if (kernel) {
        PendingEvents& pe = pending_events();
        if (pe.start == a) {  // no launch took them
          pe.start = pe.stop = nullptr;
          (void)hipEventRecord(a, s);
        } else {
          by_kernel = true;
        }
      }
      if (!by_kernel) (void)hipEventRecord(b, s);
      p->ev[id].emplace_back(a, b);
    }
  }
};

}  // namespace

// The launches of one stage's nonlinear passes of a plan, recorded (not
// issued) so that several plans' passes can be issued together.
struct LaunchRec {
  bool row;
  int L, kind, v;
  ColArgs ca;
  RowArgs ra;
  const void* tw;
};
struct LaunchLog {
  LaunchRec r[8];
  int n = 0;
  bool overflow = false;
};

namespace {

int dev_alloc(pf_plan* p, void** ptr, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(PF_ENOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
  }
  p->owned.push_back(*ptr);
  p->bytes += bytes;
  return 0;
}

void release(pf_plan* p) {
  for (void* q : p->owned) (void)hipFree(q);
  p->owned.clear();
}

bool supported_length(int L) {
  switch (L) {
#define PF_CASE(n) case n: return true;
    PF_LENGTHS(PF_CASE)
#undef PF_CASE
    default: return false;
  }
}

// Twiddle table of length L under radix plan rp.
const void* tw_of(const pf_plan* p, int L, int rp = 0) {
  for (int i = 0; i < p->ntw; ++i)
    if (p->tw[i].L == L && p->tw[i].rp == rp) return p->tw[i].ptr;
  return nullptr;
}

int make_twiddles(pf_plan* p, int L, int rp = 0) {
  if (rp == 0 && has_alt_plan(L)) {
    int rc = make_twiddles(p, L, 1);
    if (rc) return rc;
  }
  if (tw_of(p, L, rp)) return 0;
  if (p->ntw >= 12) return fail(PF_EINVAL, "too many transform lengths");
  void* dptr = nullptr;
  int rc = dev_alloc(p, &dptr, (size_t)L * p->csz);
  if (rc) return rc;
  // per-pass blocks: entry tw_offset(L, p) + (r - 1) Ns + k = exp(-2 pi i r k / (Ns R))
  const long double twopi = 6.283185307179586476925286766559005768L;
  const RadixPlan pl = make_plan(L, rp);
  std::vector<long double> re, im;
  int ns = 1;
  for (int q = 0; q < pl.n; ++q) {
    const int R = pl.r[q];
    if (q > 0) {
      for (int r = 1; r < R; ++r)
        for (int k = 0; k < ns; ++k) {
          const long double ang = twopi * (long double)(r * k) / (long double)(ns * R);
          re.push_back(cosl(ang));
          im.push_back(-sinl(ang));
        }
    }
    ns *= R;
  }
  if ((int)re.size() != tw_size(L, rp)) return fail(PF_EINVAL, "twiddle table size mismatch");
  if (p->fp64) {
    std::vector<double> h(2 * re.size() + 2, 0.0);
    for (size_t m = 0; m < re.size(); ++m) { h[2 * m] = (double)re[m]; h[2 * m + 1] = (double)im[m]; }
    PF_HIP(hipMemcpy(dptr, h.data(), 2 * re.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h(2 * re.size() + 2, 0.f);
    for (size_t m = 0; m < re.size(); ++m) { h[2 * m] = (float)re[m]; h[2 * m + 1] = (float)im[m]; }
    PF_HIP(hipMemcpy(dptr, h.data(), 2 * re.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  p->tw[p->ntw++] = {L, rp, dptr};
  return 0;
}

// geometry block shared by the point-wise kernels
StageArgs geometry(const pf_plan* p) {
  StageArgs g{};
  g.nk = p->nk;
  g.n_members = p->nact;
  g.nh = p->nh;
  g.N0 = p->N[0];
  g.N1 = p->N[1];
  g.N2 = p->dim == 3 ? p->N[2] : p->N[1];
  g.ks0 = p->ks[0];
  g.ks1 = p->ks[1];
  g.ks2 = p->dim == 3 ? p->ks[2] : 0.0;
  g.nu = p->d.viscosity;
  return g;
}

// v: the launch variant (0 = default; the autotuner's choice per pass)
template <typename T> int row(const pf_plan* p, int L, int mode, int v, const RowArgs& a, hipStream_t s) {
  if (p->rec) {
    LaunchLog& g = *p->rec;
    if (g.n >= 8) { g.overflow = true; return 0; }
    LaunchRec& r = g.r[g.n++];
    r.row = true; r.L = L; r.kind = mode; r.v = v; r.ra = a; r.tw = tw_of(p, L, row_radix_plan(v));
    return 0;
  }
  return launch_row<T>(L, mode, a, tw_of(p, L, row_radix_plan(v)), s, v);
}

// A column pass launch, or its record while the plan records.
template <typename T> int emit_col(pf_plan* p, int L, int kind, const ColArgs& a, hipStream_t s, int v) {
  if (p->rec) {
    LaunchLog& g = *p->rec;
    if (g.n >= 8) { g.overflow = true; return 0; }
    LaunchRec& r = g.r[g.n++];
    r.row = false; r.L = L; r.kind = kind; r.v = v; r.ca = a; r.tw = tw_of(p, L, col_radix_plan(v));
    return 0;
  }
  return launch_col<T>(L, kind, a, tw_of(p, L, col_radix_plan(v)), s, v);
}

ColArgs col_args() {
  ColArgs a{};
  a.g.n_outer = 1;
  a.n_members = 1;
  a.n_fields = 1;
  a.scale = 1.0;
  return a;
}

// a plain column pass over `cols` unit-stride columns of n_outer matrices
template <typename T>
int col_plain(const pf_plan* p, int L, int kind, ColArgs a, int cols, int n_outer, long long in_outer,
              long long out_outer, hipStream_t s) {
  ColGeom& g = a.g;
  g.n2c = g.cload = g.cvalid = g.cols = cols;
  g.n_outer = n_outer;
  g.in_outer = in_outer;
  g.out_outer = out_outer;
  return launch_col<T>(L, kind, a, tw_of(p, L), s, 0);
}

// Loaded / stored column counts of a geometry: the first `nvalid` entries,
// widened to whole tiles where a row holds at least `cap` entries (cap = 0:
// not widened).  Entries past nvalid then read whatever the padded row holds
// and their results land in padding that no later pass reads.
inline int widen(int nvalid, int n2c, int cap) { return cap > 0 ? std::max(nvalid, std::min(n2c, cap)) : nvalid; }

// One strided-axis pass over groups of last-axis entries: `ngroups` groups
// (rows of the axis not transformed) of `nvalid` entries each, group strides
// in_lp / out_lp.  Tiles cover the first `nvalid` entries rounded up to whole
// tiles; loads and stores cover them up to the row capacities cap_in /
// cap_out (see widen).
// flat: the groups' entries are enumerated back to back (a tile may span
// two groups) instead of each group starting a new tile.
template <typename T>
int col_groups(pf_plan* p, int L, int kind, int v, ColArgs a, int ngroups, long long in_lp, long long out_lp,
               int nvalid, hipStream_t s, int cap_in = 0, int cap_out = 0, bool flat = false) {
  const int tc = tile_cols<T>(L, kind, v);
  if (tc <= 0) return -4;
  // the fused stage kinds update every enumerated column: the groups must be
  // enumerated back to back, exactly
  if ((kind == K_FWD_STAGE3 || kind == K_FWD_STAGE2) && !flat)
    return fail(PF_EINVAL, "stage pass needs flat column groups");
  ColGeom& g = a.g;
  g.n2c = flat ? nvalid : (nvalid + tc - 1) / tc * tc;
  g.cload = widen(nvalid, g.n2c, cap_in);
  g.cvalid = widen(nvalid, g.n2c, cap_out);
  g.cols = ngroups * g.n2c;
  g.in_lp = in_lp;
  g.out_lp = out_lp;
  g.n_outer = 1;
  g.in_outer = g.out_outer = 0;
  return emit_col<T>(p, L, kind, a, s, v);
}

// The same over the last-axis columns l of n_outer per-plane matrices (the
// axis-1 passes).
template <typename T>
int col_planes(pf_plan* p, int L, int kind, int v, ColArgs a, int n_outer, long long in_outer, long long out_outer,
               int nvalid, hipStream_t s, int cap_in = 0, int cap_out = 0) {
  const int tc = tile_cols<T>(L, kind, v);
  if (tc <= 0) return -4;
  ColGeom& g = a.g;
  g.n2c = (nvalid + tc - 1) / tc * tc;
  g.cload = widen(nvalid, g.n2c, cap_in);
  g.cvalid = widen(nvalid, g.n2c, cap_out);
  g.cols = g.n2c;
  g.in_lp = g.out_lp = 0;
  g.n_outer = n_outer;
  g.in_outer = in_outer;
  g.out_outer = out_outer;
  return emit_col<T>(p, L, kind, a, s, v);
}

// Inverse passes, the fused last-axis pass and (3D) the axis-1 forward pass
// of the nonlinear term: u_hat -> input of the axis-0 forward pass (w2),
// laid out (B, d, M0, [N1,] P) with the padded row pitch P.
template <typename T> int nonlinear_front(pf_plan* p, const void* uhat, hipStream_t s) {
  const int B = p->nact;
  const long long nh = p->nh, P = p->pitch;
  const bool pr = p->direct23;
  if (p->dim == 3) {
    const int N0 = p->N[0], N1 = p->N[1], N2 = p->N[2];
    const int M0 = p->M[0], M1 = p->M[1], M2 = p->M[2];
    const long long A_f = (long long)M0 * p->aplane, B_f = (long long)M0 * M1 * P;
    const long long D_f = (long long)M0 * p->dplane;  // D: rows of pitch P
    // retained last-axis entries (2/3 rule) or all nh of them (3/2 rule)
    const int lval = pr ? N2 / 3 + 1 : (int)nh;
    // P1: axis-0 inverse of u_hat and of omega_y, omega_z, u_hat -> A
    // (5 fields, w1)
    ColArgs a = col_args();
    a.in = uhat; a.out = p->w1;
    a.rows_in = N0; a.rows_out = M0;
    a.in_row = (long long)N1 * nh; a.out_row = p->aplane;
    a.in_field = p->nk; a.out_field = A_f;
    a.in_member = 3 * p->nk; a.out_member = 5 * A_f;
    a.n_fields = 3; a.n_members = B;
    a.gN0 = N0; a.gN1 = N1; a.gN2 = N2; a.gnh = (int)nh;
    a.ks0 = p->ks[0]; a.ks1 = p->ks[1]; a.ks2 = p->ks[2];
    a.mask23 = pr;
    int ngroups = N1;
    if (pr) {
      a.jcompact = 1;
      a.K1 = N1 / 3;
      ngroups = 2 * a.K1 + 1;
    }
    {
      PassTimer t(p, PF_PASS_INV0, s);
      // the entries past lval are neither loaded nor stored: A's row padding
      // keeps whatever it held, which only the same columns of P2 read.
      // The (j, l) columns are enumerated back to back, so no tile is
      // padded at a group's end (3/2 rule at 96^3: 49 entries per group,
      // which whole tiles per group rounded up to 64; 15% faster).
      // one launch: blocks of fields 0..2 transform u_hat, blocks of 3, 4
      // omega_y and omega_z (omega_x = i (ky w - kz v) needs no axis-0
      // derivative: P2 forms it at its loads, so A holds 5 fields, not 6).
      // With nontemporal stores (whose partial lines cost a read-modify-write
      // of the line) each group starts a new tile and whole tiles are stored
      // into A's padded rows, so every store is whole 128-byte lines.
      a.f_base = 0;
      a.n_fields = 5;
      const int v0 = p->tsel[PF_PASS_INV0];
      const bool lines = col_cache_policy(v0) != 0;
      PF_TRY(col_groups<T>(p, M0, K_INV_UC3, v0, a, ngroups, nh, P, lval, s, 0, lines ? (int)P : 0, !lines));
    }
    // P2: axis-1 inverse, A -> Bbuf (w2), velocity and vorticity, omega_x
    // formed from v and w at the loads
    a = col_args();
    a.in = p->w1; a.out = p->w2;
    a.rows_in = N1; a.rows_out = M1;
    a.in_row = P; a.out_row = P;
    a.in_field = A_f; a.out_field = B_f;
    a.in_member = 5 * A_f; a.out_member = 6 * B_f;
    a.n_fields = 6; a.n_members = B;
    a.ks1 = p->ks[1]; a.ks2 = p->ks[2];
    if (pr) {
      a.prune = 1;
      a.Krow = N1 / 3;
    }
    {
      PassTimer t(p, PF_PASS_INV1, s);
      // whole tiles of the padded rows, loaded and stored (a partial last
      // tile's stores cost more than the extra bytes: 30% at 128^3, 3/2
      // rule).  Columns are independent, so A's padding values stay in
      // their own columns, and B's entries past lval are not read by the
      // last-axis pass.
      const int cap = (int)P;
      PF_TRY(col_planes<T>(p, M1, K_INV_CURL3, p->tsel[PF_PASS_INV1], a, M0, p->aplane, (long long)M1 * P, lval,
                           s, cap, cap));
    }
    // P3: last axis c2r, u x omega, r2c, truncate: Bbuf -> C (w1)
    RowArgs r{};
    r.in = p->w2; r.out = p->w1;
    r.n_rows = M0 * M1; r.N = N2; r.nh = (int)nh;
    r.in_row = P; r.out_row = P;
    r.in_field = B_f; r.out_field = B_f;
    r.in_member = 6 * B_f; r.out_member = 3 * B_f;
    r.n_members = B; r.n_fields = 1;
    r.mask23 = pr; r.scale = 1.0 / M2;
    { PassTimer t(p, PF_PASS_ROW, s); PF_TRY(row<T>(p, M2, ROW_FUSED3, p->tsel[PF_PASS_ROW], r, s)); }
    // P4: axis-1 forward + truncation: C -> D (w2), all nh entries in rows
    // of pitch P.
    a = col_args();
    a.in = p->w1; a.out = p->w2;
    a.rows_in = M1; a.rows_out = N1;
    a.in_row = P; a.out_row = P;
    a.in_field = B_f; a.out_field = D_f;
    a.in_member = 3 * B_f; a.out_member = 3 * D_f;
    a.n_fields = 3; a.n_members = B; a.scale = 1.0 / M1;
    {
      PassTimer t(p, PF_PASS_FWD1, s);
      // whole tiles are loaded from C's padded rows; D's rows get their nh
      // entries, or whole tiles (whole lines) with nontemporal stores
      const int v4 = p->tsel[PF_PASS_FWD1];
      const int cap = col_cache_policy(v4) != 0 ? (int)P : 0;
      PF_TRY(col_planes<T>(p, M1, K_FWD, v4, a, M0, (long long)M1 * P, p->dplane, (int)nh, s, (int)P, cap));
    }
  } else {
    const int N0 = p->N[0], N1 = p->N[1];
    const int M0 = p->M[0], M1 = p->M[1];
    const long long A_f = (long long)M0 * P;
    const int lval = pr ? N1 / 3 + 1 : (int)nh;
    // P1: axis-0 inverse of u_hat and the scalar vorticity, u_hat -> A (3 fields, w1)
    ColArgs a = col_args();
    a.in = uhat; a.out = p->w1;
    a.rows_in = N0; a.rows_out = M0;
    a.in_row = nh; a.out_row = P;
    a.in_field = p->nk; a.out_field = A_f;
    a.in_member = 2 * p->nk; a.out_member = 3 * A_f;
    a.n_members = B;
    a.gN0 = N0; a.gN1 = N1; a.gN2 = N1; a.gnh = (int)nh;
    a.ks0 = p->ks[0]; a.ks1 = p->ks[1];
    a.mask23 = pr;
    {
      PassTimer t(p, PF_PASS_INV0, s);
      // one launch: blocks of fields 0..1 transform u_hat, field 2 omega_hat
      a.n_fields = 3;
      a.f_base = 0;
      PF_TRY(col_groups<T>(p, M0, K_INV_UC2, p->tsel[PF_PASS_INV0], a, 1, 0, 0, lval, s));
    }
    // P2: fused last axis: A -> C (2 fields, w2)
    RowArgs r{};
    r.in = p->w1; r.out = p->w2;
    r.n_rows = M0; r.N = N1; r.nh = (int)nh;
    r.in_row = P; r.out_row = P;
    r.in_field = A_f; r.out_field = A_f;
    r.in_member = 3 * A_f; r.out_member = 2 * A_f;
    r.n_members = B; r.n_fields = 1;
    r.mask23 = pr; r.scale = 1.0 / M1;
    { PassTimer t(p, PF_PASS_ROW, s); PF_TRY(row<T>(p, M1, ROW_FUSED2, p->tsel[PF_PASS_ROW], r, s)); }
  }
  return 0;
}

// The axis-0 forward pass from the front's output (w2) into the state layout:
// `out` receives the unprojected nonlinear term (K_FWD) or, with the stage
// kinds, the stage update is applied to u_hat (out unused).
template <typename T> int last_forward(pf_plan* p, int kind, void* out, const StageArgs* st, hipStream_t s) {
  const long long nh = p->nh, P = p->pitch;
  const int v = p->tsel[kind == K_FWD ? PF_PASS_FWD0 : PF_PASS_FWD0_STAGE];
  const int d = p->dim;
  const int M0 = p->M[0], N0 = p->N[0];
  ColArgs a = col_args();
  a.in = p->w2; a.out = out;
  a.rows_in = M0; a.rows_out = N0;
  a.n_fields = d; a.n_members = p->nact; a.scale = 1.0 / M0;
  a.out_field = p->nk; a.out_member = d * p->nk;
  if (st) a.st = *st;
  if (d == 3) {
    // the N1 rows of nh entries of each plane enumerated back to back: D
    // holds them at pitch P (its planes one cache line further apart, which
    // keeps the tile's rows off a near-power-of-two stride), the state at nh
    const int N1 = p->N[1];
    const long long D_f = (long long)M0 * p->dplane;
    a.in_row = p->dplane; a.out_row = (long long)N1 * nh;
    a.in_field = D_f; a.in_member = d * D_f;
    a.gN1 = N1; a.gnh = (int)nh;
    return col_groups<T>(p, M0, kind, v, a, N1, P, nh, (int)nh, s, 0, 0, true);
  }
  const long long D_f = (long long)M0 * P;
  a.in_row = P; a.out_row = nh;
  a.in_field = D_f; a.in_member = d * D_f;
  a.gN1 = 1; a.gnh = (int)nh;
  // exactly the nh columns (the stage kinds update every enumerated column)
  return col_groups<T>(p, M0, kind, v, a, 1, 0, 0, (int)nh, s, 0, 0, true);
}

// Nonlinear term (unprojected) of u_hat into the plan's R buffer (w1).
template <typename T> int nonlinear_unprojected(pf_plan* p, const void* uhat, hipStream_t s) {
  int rc = nonlinear_front<T>(p, uhat, s);
  if (rc) return rc;
  { PassTimer t(p, PF_PASS_FWD0, s); PF_TRY(last_forward<T>(p, K_FWD, p->w1, nullptr, s)); }
  return 0;
}

// Unpadded backward of nf fields (nf = members x components): in -> out.
template <typename T> int backward_fields(pf_plan* p, const void* in, void* out, int nf, hipStream_t s) {
  const long long nh = p->nh;
  if ((size_t)nf * p->nk * p->csz > p->w1_bytes) return fail(PF_EINVAL, "backward: too many fields for workspace");
  if (p->dim == 3) {
    const int N0 = p->N[0], N1 = p->N[1], N2 = p->N[2];
    ColArgs a = col_args();
    a.in = in; a.out = p->w1;
    a.rows_in = N0; a.rows_out = N0;
    a.in_row = (long long)N1 * nh; a.out_row = a.in_row;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf;
    PF_TRY(col_plain<T>(p, N0, K_INV, a, N1 * (int)nh, 1, 0, 0, s));
    a = col_args();
    a.in = p->w1; a.out = p->w1;
    a.rows_in = N1; a.rows_out = N1;
    a.in_row = nh; a.out_row = nh;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf;
    PF_TRY(col_plain<T>(p, N1, K_INV, a, (int)nh, N0, (long long)N1 * nh, (long long)N1 * nh, s));
    RowArgs r{};
    r.in = p->w1; r.out = out;
    r.n_rows = N0 * N1; r.N = N2; r.nh = (int)nh;
    r.in_row = nh; r.out_row = N2;
    r.in_field = p->nk; r.out_field = p->nphys;
    r.n_fields = nf; r.n_members = 1; r.out_scale = 1.0;
    PF_TRY(row<T>(p, N2, ROW_C2R, 0, r, s));
  } else {
    const int N0 = p->N[0], N1 = p->N[1];
    ColArgs a = col_args();
    a.in = in; a.out = p->w1;
    a.rows_in = N0; a.rows_out = N0;
    a.in_row = nh; a.out_row = nh;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf;
    PF_TRY(col_plain<T>(p, N0, K_INV, a, (int)nh, 1, 0, 0, s));
    RowArgs r{};
    r.in = p->w1; r.out = out;
    r.n_rows = N0; r.N = N1; r.nh = (int)nh;
    r.in_row = nh; r.out_row = N1;
    r.in_field = p->nk; r.out_field = p->nphys;
    r.n_fields = nf; r.n_members = 1; r.out_scale = 1.0;
    PF_TRY(row<T>(p, N1, ROW_C2R, 0, r, s));
  }
  return 0;
}

template <typename T> int forward_fields(pf_plan* p, const void* in, void* out, int nf, hipStream_t s) {
  const long long nh = p->nh;
  if (p->dim == 3) {
    const int N0 = p->N[0], N1 = p->N[1], N2 = p->N[2];
    RowArgs r{};
    r.in = in; r.out = out;
    r.n_rows = N0 * N1; r.N = N2; r.nh = (int)nh;
    r.in_row = N2; r.out_row = nh;
    r.in_field = p->nphys; r.out_field = p->nk;
    r.n_fields = nf; r.n_members = 1; r.scale = 1.0 / N2;
    PF_TRY(row<T>(p, N2, ROW_R2C, 0, r, s));
    ColArgs a = col_args();
    a.in = out; a.out = out;
    a.rows_in = N1; a.rows_out = N1;
    a.in_row = nh; a.out_row = nh;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf; a.scale = 1.0 / N1;
    PF_TRY(col_plain<T>(p, N1, K_FWD, a, (int)nh, N0, (long long)N1 * nh, (long long)N1 * nh, s));
    a = col_args();
    a.in = out; a.out = out;
    a.rows_in = N0; a.rows_out = N0;
    a.in_row = (long long)N1 * nh; a.out_row = a.in_row;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf; a.scale = 1.0 / N0;
    PF_TRY(col_plain<T>(p, N0, K_FWD, a, N1 * (int)nh, 1, 0, 0, s));
  } else {
    const int N0 = p->N[0], N1 = p->N[1];
    RowArgs r{};
    r.in = in; r.out = out;
    r.n_rows = N0; r.N = N1; r.nh = (int)nh;
    r.in_row = N1; r.out_row = nh;
    r.in_field = p->nphys; r.out_field = p->nk;
    r.n_fields = nf; r.n_members = 1; r.scale = 1.0 / N1;
    PF_TRY(row<T>(p, N1, ROW_R2C, 0, r, s));
    ColArgs a = col_args();
    a.in = out; a.out = out;
    a.rows_in = N0; a.rows_out = N0;
    a.in_row = nh; a.out_row = nh;
    a.in_field = p->nk; a.out_field = p->nk;
    a.n_fields = nf; a.scale = 1.0 / N0;
    PF_TRY(col_plain<T>(p, N0, K_FWD, a, (int)nh, 1, 0, 0, s));
  }
  return 0;
}

struct StageCoef {
  double ca, cb, noise_a;
  bool last;
};

int n_stages(int integ) { return integ == PF_RK3 ? 3 : (integ == PF_RK4 ? 4 : 2); }

// integrator tables, time_integrator.py:45-47, :102-107, :170-174, :222-232, :287-288
StageCoef stage_coef(int integ, int st) {
  const double s2 = std::sqrt(2.0), s3 = std::sqrt(3.0);
  switch (integ) {
    case PF_EXPLICIT_PC: return {st == 0 ? 0.5 : 1.0, 0.0, st == 0 ? 0.0 : 1.0, st == 1};
    case PF_IMPLICIT_PC: return {st == 0 ? 0.5 : 1.0, st == 0 ? 0.0 : 0.5, st == 0 ? 0.0 : 1.0, st == 1};
    case PF_RK2: return {st == 0 ? 0.0 : 0.5, st == 0 ? 1.0 : 0.5, 0.0, st == 1};
    case PF_RK3: {
      const double a[3] = {0.0, 0.75, 1.0 / 3.0}, b[3] = {1.0, 0.25, 2.0 / 3.0};
      const double nf[3] = {(2 * s2 + s3) / 5, (-4 * s2 + 3 * s3) / 5, (s2 - 2 * s3) / 10};
      return {a[st], b[st], nf[st], st == 2};
    }
    default: {
      const double a[4] = {0.25, 1.0 / 3.0, 0.5, 1.0};
      return {a[st], 0.0, 0.0, st == 3};
    }
  }
}

bool uses_noise(int integ) { return integ == PF_EXPLICIT_PC || integ == PF_IMPLICIT_PC || integ == PF_RK3; }

template <typename T> int band_energy(pf_plan* p, const void* uhat, hipStream_t s) {
  BandArgs ba{};
  ba.uhat = uhat; ba.coef = p->fcoef; ba.nk = p->nk;
  ba.N0 = p->N[0]; ba.N1 = p->N[1]; ba.N2 = p->dim == 3 ? p->N[2] : p->N[1]; ba.nh = p->nh;
  ba.ks0 = p->ks[0]; ba.ks1 = p->ks[1]; ba.ks2 = p->ks[2];
  ba.kmin2 = p->d.forcing_kmin * p->d.forcing_kmin;
  ba.kmax2 = p->d.forcing_kmax * p->d.forcing_kmax;
  ba.kmax = p->d.forcing_kmax;
  ba.power = p->d.forcing_power;
  { PassTimer t(p, PF_PASS_BAND, s, false); PF_TRY(launch_band<T>(p->dim, ba, p->nact, s)); }
  return 0;
}

// Arguments of one stage kernel launch (noise: raw draws of this step or null).
StageArgs stage_args(const pf_plan* p, void* uhat, int stage, const void* noise, bool noise_on) {
  const StageCoef c = stage_coef(p->d.integrator, stage);
  StageArgs a = geometry(p);
  a.uhat = uhat;
  a.u0 = p->u0;
  a.rhs_nl = p->d.nonlinear ? p->w1 : nullptr;
  a.lin_op = p->lin_op;
  a.corr = p->corr;
  a.noise = noise_on ? noise : nullptr;
  a.seeds = p->seeds;
  a.fcoef = p->d.forcing == PF_FORCING_CONSTANT_POWER ? p->fcoef : nullptr;
  a.integ = p->d.integrator;
  a.stage = stage;
  a.last_stage = c.last;
  a.save_u0 = stage == 0;
  a.ca = c.ca;
  a.cb = c.cb;
  a.dt = p->d.dt;
  a.dt_sqrt = p->d.integrator == PF_RK3 ? std::sqrt(p->d.dt) : std::sqrt(0.5 * p->d.dt);
  a.noise_on = noise_on;
  a.noise_type = p->d.noise_type;
  a.use_rng = noise == nullptr;
  a.noise_a = c.noise_a;
  a.noise_mag = p->d.noise_mag;
  a.noise_norm = 1.0 / std::sqrt((double)p->nphys) / std::sqrt(2.0 - 2.0 / p->dim);
  a.mask_nyq = p->d.mask_nyquist;
  a.fk2min = p->d.forcing_kmin * p->d.forcing_kmin;
  a.fk2max = p->d.forcing_kmax * p->d.forcing_kmax;
  a.step = p->step;
  a.msteps = p->msteps;
  a.mode = 0;
  return a;
}

// Stokes flow with implicit PC and no forcing: the predictor's result is
// never read (the corrector uses u0, L u0 and the noise only), so a step is
// the corrector alone.
bool corrector_only(const pf_plan* p) {
  return !p->d.nonlinear && p->d.forcing == PF_FORCING_NONE && p->d.integrator == PF_IMPLICIT_PC && p->fuse_stage;
}

// Such steps with the on-device noise stream (or none) run many at a time
// with the state in registers (stokes_steps_kernel), unless the stage pass's
// launch variant 1 asks for one launch per step.  Both give the same state
// bit for bit.
bool resident_steps(const pf_plan* p, const void* noise) {
  return corrector_only(p) && noise == nullptr && p->tsel[PF_PASS_STAGE] == 0;
}

// Cache policy of the separate stage kernel: nontemporal loads and stores
// under PF_PASS_STAGE variant 7 (plans whose steps are not corrector-only).
int stage_cache_policy(const pf_plan* p) { return p->tsel[PF_PASS_STAGE] == 7 ? 3 : 0; }

constexpr int kResidentChunk = 200;  // steps per register-resident launch

// n_steps corrector-only steps with the state in registers; acc (nullable)
// += sum over members and components of |u_hat|^2 after every step.
template <typename T> int resident_impl(pf_plan* p, void* uhat, long long n_steps, double* acc, hipStream_t s) {
  const bool noise_on = p->d.noise_mag != 0.0;
  StageArgs a = stage_args(p, uhat, 1, nullptr, noise_on);
  a.save_u0 = 2;
  a.single_draw = 1;
  for (long long done = 0; done < n_steps;) {
    const int k = (int)std::min<long long>(kResidentChunk, n_steps - done);
    { PassTimer t(p, PF_PASS_STAGE, s, false); PF_TRY(launch_stokes_steps<T>(p->dim, a, k, acc, s)); }
    PF_TRY(launch_member_step(p->msteps, p->nact, s, k));
    done += k;
    p->step += k;
  }
  return 0;
}

template <typename T> int step_impl(pf_plan* p, void* uhat, long long n_steps, const void* noise, hipStream_t s) {
  if (resident_steps(p, noise)) return resident_impl<T>(p, uhat, n_steps, nullptr, s);
  const int ns = n_stages(p->d.integrator);
  const bool noise_on = p->d.noise_mag != 0.0 && uses_noise(p->d.integrator);
  const long long noise_block = (long long)p->nact * 2 * p->nc * p->nk;  // complex per step
  const bool fused = p->fuse_stage &&
                     tile_cols<T>(p->M[0], p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, p->tsel[PF_PASS_FWD0_STAGE]) > 0;
  const int first = corrector_only(p) ? 1 : 0;
  for (long long st = 0; st < n_steps; ++st) {
    const void* nz = noise ? (const char*)noise + (size_t)(st * noise_block) * p->csz : nullptr;
    for (int stage = first; stage < ns; ++stage) {
      if (p->d.forcing == PF_FORCING_CONSTANT_POWER) {
        int rc = band_energy<T>(p, uhat, s);
        if (rc) return rc;
      }
      StageArgs a = stage_args(p, uhat, stage, nz, noise_on);
      if (stage == first && first > 0) a.save_u0 = 2;  // u_hat is still u0; nothing to save
      a.single_draw = first == ns - 1;                   // this step's draws feed this stage only
      if (p->d.nonlinear && fused) {
        int rc = nonlinear_front<T>(p, uhat, s);
        if (rc) return rc;
        a.rhs_nl = nullptr;
        PassTimer t(p, PF_PASS_FWD0_STAGE, s);
        PF_TRY(last_forward<T>(p, p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, nullptr, &a, s));
      } else {
        if (p->d.nonlinear) {
          int rc = nonlinear_unprojected<T>(p, uhat, s);
          if (rc) return rc;
        }
        PassTimer t(p, PF_PASS_STAGE, s);
        PF_TRY(launch_stage<T>(p->dim, a, s, stage_cache_policy(p)));
      }
    }
    PF_TRY(launch_member_step(p->msteps, p->nact, s));
    p->step += 1;
  }
  return 0;
}

constexpr int kMaxMulti = 16;  // plans one pf_step_multi call advances together

// Issues the recorded launches of several plans pass by pass (record r of
// every plan before record r + 1 of any, all on one stream, so each plan's
// passes keep their order): two plans whose records at r are the same pass
// kind at different compiled lengths, both at the same launch variant of the
// default geometry (pair_variant: any cache policy), share one paired
// launch; the rest launch one by one.
template <typename T> int issue_logs(LaunchLog* logs, int np, hipStream_t s, int* paired) {
  const int nr = logs[0].n;
  for (int r = 0; r < nr; ++r) {
    bool used[kMaxMulti] = {};
    for (int i = 0; i < np; ++i) {
      if (used[i]) continue;
      const LaunchRec& a = logs[i].r[r];
      int partner = -1;
      if (sizeof(T) == 4 && pair_variant(a.row, a.v)) {
        for (int j = i + 1; j < np && partner < 0; ++j) {
          const LaunchRec& b = logs[j].r[r];
          if (!used[j] && b.v == a.v && b.row == a.row && b.kind == a.kind && b.L != a.L && pair_compiled(a.L, b.L))
            partner = j;
        }
      }
      int rc = -4;
      if (partner >= 0) {
        const LaunchRec& b = logs[partner].r[r];
        rc = a.row ? launch_row_pair<T>(a.kind, a.L, a.ra, a.tw, b.L, b.ra, b.tw, s, a.v)
                   : launch_col_pair<T>(a.kind, a.L, a.ca, a.tw, b.L, b.ca, b.tw, s, a.v);
        if (rc == 0) {
          used[partner] = true;
          if (paired) ++*paired;
        }
      }
      if (rc == -4) rc = a.row ? launch_row<T>(a.L, a.kind, a.ra, a.tw, s, a.v) : launch_col<T>(a.L, a.kind, a.ca, a.tw, s, a.v);
      if (rc) return launch_fail(rc, "issue_logs");
      used[i] = true;
    }
  }
  return 0;
}

// n_steps steps of several plans, their nonlinear passes issued together
// (issue_logs); a plan whose axis-0 forward pass is not fused with its stage
// update runs its stage kernel after the issued passes.  Plans that cannot go
// pass by pass together (other stage counts or dimension, Stokes flow) step
// one after another instead.  Noise is the on-device stream.
template <typename T> int step_multi_impl(pf_plan** ps, void** us, int np, long long n_steps, hipStream_t s,
                                          int* paired) {
  bool together = np >= 2;
  bool fused[kMaxMulti] = {};
  for (int i = 0; i < np && together; ++i) {
    const pf_plan* p = ps[i];
    fused[i] = p->fuse_stage &&
               tile_cols<T>(p->M[0], p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, p->tsel[PF_PASS_FWD0_STAGE]) > 0;
    together = p->d.nonlinear && p->d.integrator == ps[0]->d.integrator && p->dim == ps[0]->dim && !p->prof;
  }
  if (!together) {
    for (int i = 0; i < np; ++i) {
      int rc = step_impl<T>(ps[i], us[i], n_steps, nullptr, s);
      if (rc) return rc;
    }
    return 0;
  }
  const int ns = n_stages(ps[0]->d.integrator);
  for (long long st = 0; st < n_steps; ++st) {
    for (int stage = 0; stage < ns; ++stage) {
      LaunchLog logs[kMaxMulti];
      StageArgs sa[kMaxMulti];
      for (int i = 0; i < np; ++i) {
        pf_plan* p = ps[i];
        if (p->d.forcing == PF_FORCING_CONSTANT_POWER) {
          int rc = band_energy<T>(p, us[i], s);
          if (rc) return rc;
        }
        const bool noise_on = p->d.noise_mag != 0.0 && uses_noise(p->d.integrator);
        StageArgs& a = sa[i];
        a = stage_args(p, us[i], stage, nullptr, noise_on);
        p->rec = &logs[i];
        int rc = nonlinear_front<T>(p, us[i], s);
        if (!rc && fused[i]) {
          StageArgs f = a;
          f.rhs_nl = nullptr;
          rc = last_forward<T>(p, p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, nullptr, &f, s);
        } else if (!rc) {
          rc = last_forward<T>(p, K_FWD, p->w1, nullptr, s);  // R; the stage kernel follows
        }
        p->rec = nullptr;
        if (rc) return rc;
        if (logs[i].overflow) return fail(PF_EINVAL, "step_multi: more launches per stage than recorded");
      }
      bool same = true;
      for (int i = 1; i < np; ++i) {
        same = same && logs[i].n == logs[0].n;
        for (int r = 0; same && r < logs[0].n; ++r)
          same = logs[i].r[r].row == logs[0].r[r].row && logs[i].r[r].kind == logs[0].r[r].kind;
      }
      if (same) {
        int rc = issue_logs<T>(logs, np, s, paired);
        if (rc) return rc;
      } else {
        for (int i = 0; i < np; ++i) {
          int rc = issue_logs<T>(&logs[i], 1, s, nullptr);
          if (rc) return rc;
        }
      }
      for (int i = 0; i < np; ++i)
        if (!fused[i]) PF_TRY(launch_stage<T>(ps[i]->dim, sa[i], s, stage_cache_policy(ps[i])));
    }
    for (int i = 0; i < np; ++i) {
      PF_TRY(launch_member_step(ps[i]->msteps, ps[i]->nact, s));
      ps[i]->step += 1;
    }
  }
  return 0;
}

template <typename T>
int compute_rhs_impl(pf_plan* p, void* uhat, double na, const void* noise, void* rnl, void* rlin, void* rf, void* rn,
                     hipStream_t s) {
  if (p->d.nonlinear) {
    int rc = nonlinear_unprojected<T>(p, uhat, s);
    if (rc) return rc;
  }
  if (p->d.forcing == PF_FORCING_CONSTANT_POWER) {
    int rc = band_energy<T>(p, uhat, s);
    if (rc) return rc;
  }
  StageArgs a = stage_args(p, uhat, 0, noise, p->d.noise_mag != 0.0);
  a.noise_a = na;
  a.mode = 1;
  a.t_nl = rnl; a.t_lin = rlin; a.t_force = rf; a.t_noise = rn;
  PF_TRY(launch_stage<T>(p->dim, a, s));
  return 0;
}

}  // namespace

namespace {

// Whether launch variant v exists for every launch of a pass at this plan's
// lengths (variant 0 always does).
template <typename T> bool variant_ok(const pf_plan* p, int pass, int v) {
  if (v == 0) return true;
  if (v < 0 || v >= kVariants) return false;
  const bool d3 = p->dim == 3;
  switch (pass) {
    case PF_PASS_INV0:
      return tile_cols<T>(p->M[0], d3 ? K_INV_UC3 : K_INV_UC2, v) > 0;
    case PF_PASS_INV1: return d3 && tile_cols<T>(p->M[1], K_INV_CURL3, v) > 0;
    case PF_PASS_ROW: return row_pair_count<T>(p->M[p->dim - 1], d3 ? ROW_FUSED3 : ROW_FUSED2, v) > 0;
    case PF_PASS_FWD1: return d3 && tile_cols<T>(p->M[1], K_FWD, v) > 0;
    case PF_PASS_FWD0: return tile_cols<T>(p->M[0], K_FWD, v) > 0;
    case PF_PASS_FWD0_STAGE: return tile_cols<T>(p->M[0], d3 ? K_FWD_STAGE3 : K_FWD_STAGE2, v) > 0;
    // corrector-only Stokes steps: 0 register-resident, 1 one launch per
    // step; other plans: 7 the separate stage kernel with nontemporal loads
    // and stores
    case PF_PASS_STAGE: return corrector_only(p) ? v == 1 : v == 7;
    default: return false;
  }
}

// Average duration (ms) of the launches recorded for pass `id`; -1 if none.
double avg_ms(pf_plan* p, int id) {
  double t = 0.0;
  for (auto& e : p->ev[id]) {
    if (hipEventSynchronize(e.second) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.first, e.second) != hipSuccess) return -1.0;
    t += ms;
  }
  return p->ev[id].empty() ? -1.0 : t / (double)p->ev[id].size();
}

struct DevBuffer {
  void* ptr = nullptr;
  ~DevBuffer() { if (ptr) (void)hipFree(ptr); }
};

// One evaluation of every tunable pass on the state `u`: the front passes
// and the fused axis-0 forward + stage update (if it exists at this
// variant), then the front passes again and the unfused axis-0 forward and
// stage update, so that each candidate follows the same passes (whose output
// may still sit in the last-level cache) as it does in a step.
template <typename T> int tune_round(pf_plan* p, void* u, bool fused_ok, int stage_cp, hipStream_t s) {
  const int ns = n_stages(p->d.integrator);
  const bool noise_on = p->d.noise_mag != 0.0 && uses_noise(p->d.integrator);
  StageArgs a = stage_args(p, u, ns > 1 ? 1 : 0, nullptr, noise_on);
  a.save_u0 = 0;
  int rc = 0;
  if (fused_ok) {
    if ((rc = nonlinear_front<T>(p, u, s))) return rc;
    StageArgs f = a;
    f.rhs_nl = nullptr;
    PassTimer t(p, PF_PASS_FWD0_STAGE, s);
    PF_TRY(last_forward<T>(p, p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, nullptr, &f, s));
  }
  if ((rc = nonlinear_front<T>(p, u, s))) return rc;
  { PassTimer t(p, PF_PASS_FWD0, s); PF_TRY(last_forward<T>(p, K_FWD, p->w1, nullptr, s)); }
  { PassTimer t(p, PF_PASS_STAGE, s); PF_TRY(launch_stage<T>(p->dim, a, s, stage_cp)); }
  return 0;
}

// One stage of the plan's own launch sequence on u (front passes, then the
// fused axis-0 forward + stage pass or the unfused pair), for timing the
// sequence as a whole.
template <typename T> int sequence_round(pf_plan* p, void* u, hipStream_t s) {
  const int ns = n_stages(p->d.integrator);
  const bool noise_on = p->d.noise_mag != 0.0 && uses_noise(p->d.integrator);
  StageArgs a = stage_args(p, u, ns > 1 ? 1 : 0, nullptr, noise_on);
  a.save_u0 = 0;
  PF_TRY(nonlinear_front<T>(p, u, s));
  if (p->fuse_stage) {
    StageArgs f = a;
    f.rhs_nl = nullptr;
    PF_TRY(last_forward<T>(p, p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, nullptr, &f, s));
  } else {
    PF_TRY(last_forward<T>(p, K_FWD, p->w1, nullptr, s));
    PF_TRY(launch_stage<T>(p->dim, a, s, stage_cache_policy(p)));
  }
  return 0;
}

// Milliseconds per stage of n back-to-back stages of the current plan (the
// lower of two runs), or a negative value on failure.
template <typename T> double sequence_ms(pf_plan* p, void* u, int n, hipStream_t s) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess) return -1.0;
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return -1.0; }
  double best = -1.0;
  for (int rep = 0; rep < 2; ++rep) {
    if (hipEventRecord(e0, s) != hipSuccess) break;
    int rc = 0;
    for (int i = 0; i < n && !rc; ++i) rc = sequence_round<T>(p, u, s);
    float ms = 0.f;
    if (rc || hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      best = -1.0;
      break;
    }
    const double t = ms / (double)n;
    if (best < 0.0 || t < best) best = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

template <typename T> int autotune_impl(pf_plan* p, const void* uhat, int reps, double* ms_out, hipStream_t s) {
  if (ms_out)
    for (int i = 0; i < PF_NPASS; ++i) ms_out[i] = 0.0;
  if (!p->d.nonlinear) return PF_OK;  // the step is one point-wise kernel: nothing to choose
  const size_t sb = (size_t)p->B * p->dim * p->nk * p->csz;
  DevBuffer tmp;
  PF_HIP(hipMalloc(&tmp.ptr, sb));
  PF_HIP(hipMemcpyAsync(tmp.ptr, uhat, sb, hipMemcpyDeviceToDevice, s));
  PF_HIP(hipMemcpyAsync(p->u0, uhat, sb, hipMemcpyDeviceToDevice, s));
  const bool prof0 = p->prof;
  const uint32_t mask0 = p->prof_mask;
  p->prof_mask = 0xffffffffu;
  const int tunable[] = {PF_PASS_INV0, PF_PASS_INV1, PF_PASS_ROW, PF_PASS_FWD1, PF_PASS_FWD0, PF_PASS_FWD0_STAGE};
  double best[PF_NPASS], def[PF_NPASS], tv[PF_NPASS][kVariants];
  int bestv[PF_NPASS];
  for (int i = 0; i < PF_NPASS; ++i) {
    best[i] = def[i] = -1.0;
    bestv[i] = 0;
    for (int v = 0; v < kVariants; ++v) tv[i][v] = -1.0;
  }
  double stage_ms = -1.0, stage_def = -1.0;
  double stage_by_cp[2] = {-1.0, -1.0};  // the stage kernel's best time, ordinary / nontemporal
  int stage_cp = 0;
  int rc = 0;
  for (int v = -1; v < kVariants && !rc; ++v) {  // v = -1: warm-up round at the defaults
    const int vv = v < 0 ? 0 : v;
    for (int id : tunable) p->tsel[id] = variant_ok<T>(p, id, vv) ? vv : 0;
    const bool fused_ok = tile_cols<T>(p->M[0], p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, p->tsel[PF_PASS_FWD0_STAGE]) > 0;
    // the separate stage kernel's cache policy follows the column variant's
    // (its loads and stores nontemporal in the round of variant 7)
    const int scp = col_cache_policy(vv) == 3 ? 3 : 0;
    pf_profile_enable(p, 1);
    for (int r = 0; r < (v < 0 ? 1 : reps) && !rc; ++r) rc = tune_round<T>(p, tmp.ptr, fused_ok, scp, s);
    if (rc || v < 0) continue;
    for (int id : tunable) {
      if (!variant_ok<T>(p, id, v)) continue;
      const double t = avg_ms(p, id);
      tv[id][v] = t;
      if (v == 0) def[id] = t;
      if (t >= 0.0 && (best[id] < 0.0 || t < best[id])) { best[id] = t; bestv[id] = v; }
    }
    const double ts = avg_ms(p, PF_PASS_STAGE);
    if (ts >= 0.0 && v == 0) stage_def = ts;
    if (ts >= 0.0 && (stage_ms < 0.0 || ts < stage_ms)) { stage_ms = ts; stage_cp = scp; }
    double& sb = stage_by_cp[scp ? 1 : 0];
    if (ts >= 0.0 && (sb < 0.0 || ts < sb)) sb = ts;
  }
  pf_profile_enable(p, prof0);
  p->prof_mask = mask0;
  const hipError_t e = hipStreamSynchronize(s);
  if (rc) {
    for (int id : tunable) p->tsel[id] = 0;
    p->tsel[PF_PASS_STAGE] = 0;
    return rc;
  }
  if (e != hipSuccess) return fail(PF_EHIP, std::string("autotune: ") + hipGetErrorString(e));
  // A variant replaces the default only when it is clearly faster (timing
  // noise of a few repetitions must not flip a plan).
  constexpr double kMargin = 0.97;
  for (int id : tunable) {
    if (bestv[id] != 0 && def[id] >= 0.0 && best[id] > kMargin * def[id]) {
      bestv[id] = 0;
      best[id] = def[id];
    }
    p->tsel[id] = bestv[id];
  }
  if (stage_cp != 0 && stage_def >= 0.0 && stage_ms > kMargin * stage_def) {
    stage_cp = 0;
    stage_ms = stage_def;
  }
  p->tsel[PF_PASS_STAGE] = stage_cp ? 7 : 0;
  // fused against unfused axis-0 forward pass + stage update: fused only
  // where it is clearly faster.  The unfused stage kernel streams the state
  // and keeps its time from box to box, where the fused pass's strided stage
  // traffic varies by several percent, so a tie goes to the unfused pair.
  if (best[PF_PASS_FWD0_STAGE] >= 0.0 && best[PF_PASS_FWD0] >= 0.0 && stage_ms >= 0.0)
    p->fuse_stage = best[PF_PASS_FWD0_STAGE] < kMargin * (best[PF_PASS_FWD0] + stage_ms);

  // Refinement on the whole sequence.  The passes' choices interact through
  // the caches: a pass's stores decide what the next pass finds in the
  // 256 MB last-level cache (e.g. at 256^3 the axis-0 forward pass with
  // ordinary stores leaves R there for the stage kernel), which timing each
  // pass within rounds of one variant does not see.  Each other variant of
  // each pass in use (within 15% of the chosen one on its own), the stage
  // kernel's cache policy and the fusion choice are tried on consecutive
  // stages of the plan's own sequence and kept where the sequence gets at
  // least 0.5% faster (the lower of two timings of six stages repeats to
  // about 0.1% at 256^3).
  {
    constexpr int kSeq = 6;
    constexpr double kSeqMargin = 0.995;
    const bool prof1 = p->prof;
    p->prof = false;  // no per-pass events from these rounds
    double base = sequence_ms<T>(p, tmp.ptr, kSeq, s);
    bool changed = false;
    auto trial = [&](auto apply, auto undo) {
      if (base < 0.0) return;
      apply();
      const double t = sequence_ms<T>(p, tmp.ptr, kSeq, s);
      if (t >= 0.0 && t < kSeqMargin * base) {
        base = t;
        changed = true;
      } else {
        undo();
      }
    };
    // a second sweep over the passes where the first changed the plan (a
    // pass's best choice can depend on a later pass's)
    for (int sweep = 0; sweep < 2 && (sweep == 0 || changed); ++sweep) {
      changed = false;
      for (int id : tunable) {
        const bool used = p->fuse_stage ? id != PF_PASS_FWD0 : id != PF_PASS_FWD0_STAGE;
        if (!used) continue;
        for (int v = 0; v < kVariants; ++v) {
          const int cur = p->tsel[id];
          if (v == cur || tv[id][v] < 0.0 || !variant_ok<T>(p, id, v)) continue;
          // variants far slower on their own are not worth a trial
          if (tv[id][cur] > 0.0 && tv[id][v] > 1.15 * tv[id][cur]) continue;
          trial([&] { p->tsel[id] = v; }, [&] { p->tsel[id] = cur; });
        }
      }
    }
    if (!p->fuse_stage && !corrector_only(p)) {
      const int cur = p->tsel[PF_PASS_STAGE];
      trial([&] { p->tsel[PF_PASS_STAGE] = cur ? 0 : 7; }, [&] { p->tsel[PF_PASS_STAGE] = cur; });
    }
    if (tile_cols<T>(p->M[0], p->dim == 3 ? K_FWD_STAGE3 : K_FWD_STAGE2, p->tsel[PF_PASS_FWD0_STAGE]) > 0) {
      const bool cur = p->fuse_stage;
      trial([&] { p->fuse_stage = !cur; }, [&] { p->fuse_stage = cur; });
    }
    p->prof = prof1;
    const hipError_t e2 = hipStreamSynchronize(s);
    if (e2 != hipSuccess) return fail(PF_EHIP, std::string("autotune: ") + hipGetErrorString(e2));
  }
  if (ms_out) {
    for (int id : tunable) ms_out[id] = tv[id][p->tsel[id]] < 0.0 ? 0.0 : tv[id][p->tsel[id]];
    const double st = stage_by_cp[p->tsel[PF_PASS_STAGE] == 7 ? 1 : 0];
    ms_out[PF_PASS_STAGE] = st < 0.0 ? 0.0 : st;
  }
  return PF_OK;
}

// Validates a descriptor and fills the plan's geometry (grid, padded
// lengths, pitches, strides) and the sizes of its two work buffers in
// complex elements.  Touches no device and allocates nothing.
int plan_geometry(pf_plan* p, const pf_desc& d, long long& w1c, long long& w2c) {
  if (d.dim != 2 && d.dim != 3) return fail(PF_EINVAL, "NS solver only supports 2D and 3D problems.");
  if (d.precision != PF_FP32 && d.precision != PF_FP64) return fail(PF_EINVAL, "precision must be PF_FP32 or PF_FP64");
  if (d.dealias != PF_DEALIAS_PAD32 && d.dealias != PF_DEALIAS_TRUNC23) return fail(PF_EINVAL, "invalid dealiasing rule");
  if (d.integrator < PF_EXPLICIT_PC || d.integrator > PF_RK4) return fail(PF_EINVAL, "invalid time integrator");
  if (d.noise_type != PF_NOISE_THERMAL && d.noise_type != PF_NOISE_CORRELATED)
    return fail(PF_EINVAL, "Invalid noise type. Options are 'thermal' or 'correlated'.");
  if (d.forcing != PF_FORCING_NONE && d.forcing != PF_FORCING_CONSTANT_POWER) return fail(PF_EINVAL, "invalid forcing");
  if (d.batch < 1) return fail(PF_EINVAL, "batch must be >= 1");
  if (!(d.dt > 0)) return fail(PF_EINVAL, "dt must be > 0");
  p->d = d;
  p->dim = d.dim;
  p->B = d.batch;
  p->fp64 = d.precision == PF_FP64;
  p->rsz = p->fp64 ? 8 : 4;
  p->csz = 2 * p->rsz;
  p->direct23 = d.dealias == PF_DEALIAS_TRUNC23;
  for (int i = 0; i < d.dim; ++i) {
    const int n = d.n[i];
    if (n < 4 || n % 2) return fail(PF_EINVAL, "grid sizes must be even and >= 4");
    p->N[i] = n;
    p->M[i] = p->direct23 ? n : (3 * n) / 2;
    if (!supported_length(p->N[i]) || !supported_length(p->M[i]))
      return failf(PF_EUNSUPPORTED, "grid size %d (padded %d) is not a compiled transform length", n, p->M[i]);
    if (!(d.length[i] > 0)) return fail(PF_EINVAL, "box lengths must be > 0");
    p->ks[i] = 6.283185307179586476925286766559 / d.length[i];
  }
  if (p->M[d.dim - 1] % 2) return fail(PF_EINVAL, "padded last axis must be even");
  const int last = d.dim - 1;
  p->nh = p->N[last] / 2 + 1;
  p->nk = (long long)p->nh;
  p->nphys = 1;
  for (int i = 0; i < d.dim; ++i) p->nphys *= p->N[i];
  for (int i = 0; i < last; ++i) p->nk *= p->N[i];
  p->nc = d.noise_type == PF_NOISE_THERMAL ? d.dim * d.dim : d.dim;

  const long long B = p->B, nh = p->nh;
  const long long align = 128 / (long long)p->csz;
  p->pitch = (nh + align - 1) / align * align;
  const long long P = p->pitch;
  p->dplane = d.dim == 3 ? (long long)p->N[1] * P + align : 0;
  p->aplane = d.dim == 3 ? (long long)p->N[1] * P + align : 0;
  w1c = B * d.dim * p->nk;
  w2c = 0;
  if (d.nonlinear) {
    if (d.dim == 3) {
      const long long M0 = p->M[0], M1 = p->M[1], N0 = p->N[0], N1 = p->N[1];
      const long long A = B * 5 * M0 * p->aplane, Bb = B * 6 * M0 * M1 * P, C = B * 3 * M0 * M1 * P;
      const long long D = B * 3 * M0 * p->dplane, R = B * 3 * N0 * N1 * nh;
      w1c = std::max(std::max(w1c, A), std::max(C, R));
      w2c = std::max(Bb, D);
    } else {
      const long long M0 = p->M[0], N0 = p->N[0];
      const long long A = B * 3 * M0 * P, C = B * 2 * M0 * P, R = B * 2 * N0 * nh;
      w1c = std::max(w1c, std::max(A, R));
      w2c = C;
    }
  }
  p->nact = p->B;
  return PF_OK;
}

}  // namespace

extern "C" {

const char* pf_last_error(void) { return g_err; }
const char* pf_version(void) { return "periodicflow-hip 0.1 gfx950"; }

int pf_plan_create(const pf_desc* desc, pf_plan** out) {
  if (!desc || !out) return fail(PF_EINVAL, "null argument");
  *out = nullptr;
  const pf_desc& d = *desc;
  pf_plan* p = new pf_plan();
  long long w1c = 0, w2c = 0;
  if (int rc0 = plan_geometry(p, d, w1c, w2c)) {
    delete p;
    return rc0;
  }
  const long long B = p->B;
  int rc = 0;
  p->w1_bytes = (size_t)w1c * p->csz;
  p->w2_bytes = (size_t)w2c * p->csz;
  if ((rc = dev_alloc(p, &p->w1, p->w1_bytes)) ||
      (w2c && (rc = dev_alloc(p, &p->w2, p->w2_bytes))) ||
      (rc = dev_alloc(p, &p->u0, (size_t)B * d.dim * p->nk * p->csz)) ||
      (rc = dev_alloc(p, (void**)&p->seeds, (size_t)B * sizeof(unsigned long long))) ||
      (rc = dev_alloc(p, (void**)&p->msteps, (size_t)B * sizeof(long long))) ||
      (rc = dev_alloc(p, (void**)&p->fcoef, (size_t)B * sizeof(double)))) {
    release(p);
    delete p;
    return rc;
  }
  // correlated noise: the correlation function starts at zero, as the
  // reference's correlation_func (space_solver.py:148), so correlated noise
  // without pf_set_correlation contributes nothing
  if (d.noise_type == PF_NOISE_CORRELATED) {
    hipError_t ez = hipSuccess;
    if ((rc = dev_alloc(p, &p->corr, (size_t)p->nk * p->csz)) ||
        (ez = hipMemset(p->corr, 0, (size_t)p->nk * p->csz)) != hipSuccess) {
      release(p);
      delete p;
      return rc ? rc : hip_fail(ez, "hipMemset correlation");
    }
  }
  for (int i = 0; i < d.dim; ++i) {
    if ((rc = make_twiddles(p, p->N[i])) || (rc = make_twiddles(p, p->M[i]))) {
      release(p);
      delete p;
      return rc;
    }
  }
  p->nact = p->B;
  std::vector<unsigned long long> seeds(B);
  for (long long i = 0; i < B; ++i) seeds[i] = (unsigned long long)i;
  hipError_t e = hipMemcpy(p->seeds, seeds.data(), B * sizeof(unsigned long long), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    std::vector<long long> z(B, 0);
    e = hipMemcpy(p->msteps, z.data(), B * sizeof(long long), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    release(p);
    delete p;
    return hip_fail(e, "hipMemcpy seeds");
  }
  *out = p;
  return PF_OK;
}

int pf_plan_destroy(pf_plan* p) {
  if (!p) return PF_OK;
  (void)hipDeviceSynchronize();
  pf_profile_enable(p, 0);
  release(p);
  delete p;
  return PF_OK;
}

int pf_plan_get_info(const pf_plan* p, pf_plan_info* info) {
  if (!p || !info) return fail(PF_EINVAL, "null argument");
  info->n_fourier = p->nk;
  info->n_physical = p->nphys;
  for (int i = 0; i < 3; ++i) info->padded[i] = i < p->dim ? p->M[i] : 1;
  info->noise_components = p->nc;
  info->workspace_bytes = (int64_t)p->bytes;
  info->state_bytes = (int64_t)p->B * p->dim * p->nk * p->csz;
  return PF_OK;
}

int pf_set_seeds(pf_plan* p, const uint64_t* seeds) {
  if (!p || !seeds) return fail(PF_EINVAL, "null argument");
  PF_HIP(hipMemcpy(p->seeds, seeds, (size_t)p->B * sizeof(uint64_t), hipMemcpyHostToDevice));
  return PF_OK;
}

int pf_set_step_counter(pf_plan* p, int64_t step) {
  if (!p) return fail(PF_EINVAL, "null plan");
  p->step = step;
  std::vector<long long> v((size_t)p->B, (long long)step);
  PF_HIP(hipMemcpy(p->msteps, v.data(), v.size() * sizeof(long long), hipMemcpyHostToDevice));
  return PF_OK;
}

int pf_set_member(pf_plan* p, int32_t member, uint64_t seed, int64_t step) {
  if (!p) return fail(PF_EINVAL, "null plan");
  if (member < 0 || member >= p->B) return fail(PF_EINVAL, "member out of range");
  const unsigned long long sd = seed;
  const long long st = step;
  PF_HIP(hipMemcpy(p->seeds + member, &sd, sizeof(sd), hipMemcpyHostToDevice));
  PF_HIP(hipMemcpy(p->msteps + member, &st, sizeof(st), hipMemcpyHostToDevice));
  return PF_OK;
}

int pf_set_active(pf_plan* p, int32_t n) {
  if (!p) return fail(PF_EINVAL, "null plan");
  if (n < 0 || n > p->B) return fail(PF_EINVAL, "active member count out of range");
  p->nact = n;
  return PF_OK;
}

int pf_get_member_steps(pf_plan* p, int64_t* out) {
  if (!p || !out) return fail(PF_EINVAL, "null argument");
  PF_HIP(hipMemcpy(out, p->msteps, (size_t)p->B * sizeof(long long), hipMemcpyDeviceToHost));
  return PF_OK;
}

int64_t pf_get_step_counter(const pf_plan* p) { return p ? p->step : -1; }

int pf_add_step_counter(pf_plan* p, int64_t delta) {
  if (!p) return fail(PF_EINVAL, "null plan");
  p->step += delta;
  return PF_OK;
}

int pf_set_linear_operator(pf_plan* p, const void* L, void* stream) {
  if (!p) return fail(PF_EINVAL, "null plan");
  if (!L) { p->lin_op = nullptr; return PF_OK; }
  void* dst = nullptr;
  for (void* q : p->owned)
    if (q == p->lin_op) dst = q;
  if (!dst) {
    int rc = dev_alloc(p, &dst, (size_t)p->nk * p->rsz);
    if (rc) return rc;
  }
  PF_HIP(hipMemcpyAsync(dst, L, (size_t)p->nk * p->rsz, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  p->lin_op = dst;
  return PF_OK;
}

int pf_set_correlation(pf_plan* p, const void* C, void* stream) {
  if (!p || !C) return fail(PF_EINVAL, "null argument");
  if (!p->corr) {
    int rc = dev_alloc(p, &p->corr, (size_t)p->nk * p->csz);
    if (rc) return rc;
  }
  PF_HIP(hipMemcpyAsync(p->corr, C, (size_t)p->nk * p->csz, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return PF_OK;
}

int pf_forward(pf_plan* p, const void* u, void* uhat, int32_t ncomp, void* stream) {
  if (!p || !u || !uhat || ncomp < 1) return fail(PF_EINVAL, "invalid argument");
  const int nf = p->nact * ncomp;
  return p->fp64 ? forward_fields<double>(p, u, uhat, nf, (hipStream_t)stream)
                 : forward_fields<float>(p, u, uhat, nf, (hipStream_t)stream);
}

int pf_backward(pf_plan* p, const void* uhat, void* u, int32_t ncomp, void* stream) {
  if (!p || !u || !uhat || ncomp < 1) return fail(PF_EINVAL, "invalid argument");
  // the workspace holds batch * dim fields: transform in chunks
  const long long cap = (long long)(p->w1_bytes / ((size_t)p->nk * p->csz));
  const long long total = (long long)p->nact * ncomp;
  for (long long f0 = 0; f0 < total; f0 += cap) {
    const int nf = (int)std::min(cap, total - f0);
    const char* src = (const char*)uhat + (size_t)f0 * p->nk * p->csz;
    char* dst = (char*)u + (size_t)f0 * p->nphys * p->rsz;
    int rc = p->fp64 ? backward_fields<double>(p, src, dst, nf, (hipStream_t)stream)
                     : backward_fields<float>(p, src, dst, nf, (hipStream_t)stream);
    if (rc) return rc;
  }
  return PF_OK;
}

int pf_project(pf_plan* p, void* uhat, void* stream) {
  if (!p || !uhat) return fail(PF_EINVAL, "invalid argument");
  const StageArgs g = geometry(p);
  const int rc = p->fp64 ? launch_project<double>(p->dim, g, uhat, uhat, 0, 0, p->nact, (hipStream_t)stream)
                         : launch_project<float>(p->dim, g, uhat, uhat, 0, 0, p->nact, (hipStream_t)stream);
  return rc ? launch_fail(rc, "project") : PF_OK;
}

int pf_mask_zero_mode(pf_plan* p, void* uhat, void* stream) {
  if (!p || !uhat) return fail(PF_EINVAL, "invalid argument");
  const StageArgs g = geometry(p);
  const int rc = p->fp64 ? launch_project<double>(p->dim, g, uhat, uhat, 0, 2, p->nact, (hipStream_t)stream)
                         : launch_project<float>(p->dim, g, uhat, uhat, 0, 2, p->nact, (hipStream_t)stream);
  return rc ? launch_fail(rc, "mask_zero_mode") : PF_OK;
}

int pf_rhs_nonlinear(pf_plan* p, const void* uhat, void* rhs, void* stream) {
  if (!p || !uhat || !rhs) return fail(PF_EINVAL, "invalid argument");
  if (!p->d.nonlinear) return fail(PF_EINVAL, "plan was created with nonlinear = 0");
  hipStream_t s = (hipStream_t)stream;
  int rc = p->fp64 ? nonlinear_unprojected<double>(p, uhat, s) : nonlinear_unprojected<float>(p, uhat, s);
  if (rc) return rc;
  const StageArgs g = geometry(p);
  rc = p->fp64 ? launch_project<double>(p->dim, g, p->w1, rhs, p->d.mask_nyquist, 0, p->nact, s)
               : launch_project<float>(p->dim, g, p->w1, rhs, p->d.mask_nyquist, 0, p->nact, s);
  return rc ? launch_fail(rc, "project") : PF_OK;
}

int pf_step(pf_plan* p, void* uhat, int64_t n_steps, const void* noise, void* stream) {
  if (!p || !uhat || n_steps < 0) return fail(PF_EINVAL, "invalid argument");
  if (p->nact == 0 || n_steps == 0) return PF_OK;
  return p->fp64 ? step_impl<double>(p, uhat, n_steps, noise, (hipStream_t)stream)
                 : step_impl<float>(p, uhat, n_steps, noise, (hipStream_t)stream);
}

int pf_step_multi(pf_plan* const* plans, void* const* u_hats, int32_t n_plans, int64_t n_steps, int32_t* paired,
                  void* stream) {
  if (paired) *paired = 0;
  if (!plans || !u_hats || n_plans < 1 || n_plans > kMaxMulti || n_steps < 0)
    return failf(PF_EINVAL, "pf_step_multi: invalid argument (1 <= n_plans <= %d)", kMaxMulti);
  pf_plan* ps[kMaxMulti];
  void* us[kMaxMulti];
  int np = 0;
  for (int i = 0; i < n_plans; ++i) {
    if (!plans[i] || !u_hats[i]) return fail(PF_EINVAL, "pf_step_multi: null plan or state");
    if (plans[i]->fp64 != plans[0]->fp64) return fail(PF_EINVAL, "pf_step_multi: plans of different precision");
    for (int j = 0; j < i; ++j)
      if (plans[j] == plans[i]) return fail(PF_EINVAL, "pf_step_multi: a plan appears twice");
    if (plans[i]->nact == 0) continue;
    ps[np] = plans[i];
    us[np] = u_hats[i];
    ++np;
  }
  if (np == 0 || n_steps == 0) return PF_OK;
  hipStream_t s = (hipStream_t)stream;
  return plans[0]->fp64 ? step_multi_impl<double>(ps, us, np, n_steps, s, paired)
                        : step_multi_impl<float>(ps, us, np, n_steps, s, paired);
}

int pf_step_accumulate_sk(pf_plan* p, void* uhat, int64_t n_steps, double* acc, void* stream) {
  if (!p || !uhat || !acc || n_steps < 0) return fail(PF_EINVAL, "invalid argument");
  if (p->nact == 0 || n_steps == 0) return PF_OK;
  hipStream_t s = (hipStream_t)stream;
  if (resident_steps(p, nullptr))
    return p->fp64 ? resident_impl<double>(p, uhat, n_steps, acc, s) : resident_impl<float>(p, uhat, n_steps, acc, s);
  for (int64_t i = 0; i < n_steps; ++i) {
    int rc = p->fp64 ? step_impl<double>(p, uhat, 1, nullptr, s) : step_impl<float>(p, uhat, 1, nullptr, s);
    if (rc) return rc;
    rc = p->fp64 ? launch_sk<double>(p->dim, uhat, acc, p->nk, p->nact, s)
                 : launch_sk<float>(p->dim, uhat, acc, p->nk, p->nact, s);
    if (rc) return launch_fail(rc, "accumulate_sk");
  }
  return PF_OK;
}

int pf_compute_rhs(pf_plan* p, void* uhat, double a, const void* noise, void* rhs_nonlinear, void* rhs_linear,
                   void* forcing, void* noise_out, void* stream) {
  if (!p || !uhat || !rhs_nonlinear || !rhs_linear || !forcing || !noise_out) return fail(PF_EINVAL, "invalid argument");
  return p->fp64 ? compute_rhs_impl<double>(p, uhat, a, noise, rhs_nonlinear, rhs_linear, forcing, noise_out,
                                            (hipStream_t)stream)
                 : compute_rhs_impl<float>(p, uhat, a, noise, rhs_nonlinear, rhs_linear, forcing, noise_out,
                                           (hipStream_t)stream);
}

int pf_stage_update(pf_plan* p, int32_t stage, void* unew, const void* u0, const void* rhs_nonlinear,
                    void* rhs_linear, const void* forcing, const void* noise, void* stream) {
  if (!p || !unew || !u0 || !rhs_nonlinear || !rhs_linear || !forcing || !noise) return fail(PF_EINVAL, "invalid argument");
  if (stage < 0 || stage >= n_stages(p->d.integrator)) return fail(PF_EINVAL, "stage out of range");
  StageArgs a = stage_args(p, unew, stage, nullptr, false);
  a.u0 = const_cast<void*>(u0);
  a.save_u0 = 0;
  a.mode = 2;
  a.rhs_nl = nullptr;
  a.t_nl = const_cast<void*>(rhs_nonlinear);
  a.t_lin = rhs_linear;
  a.t_force = const_cast<void*>(forcing);
  a.t_noise = const_cast<void*>(noise);
  const int rc = p->fp64 ? launch_stage<double>(p->dim, a, (hipStream_t)stream)
                         : launch_stage<float>(p->dim, a, (hipStream_t)stream);
  return rc ? launch_fail(rc, "stage_update") : PF_OK;
}

int pf_draw_noise(pf_plan* p, void* out, void* stream) {
  if (!p || !out) return fail(PF_EINVAL, "invalid argument");
  const int rc = p->fp64 ? launch_draw<double>(out, p->seeds, p->msteps, p->nk, p->nc, p->nact, (hipStream_t)stream)
                         : launch_draw<float>(out, p->seeds, p->msteps, p->nk, p->nc, p->nact, (hipStream_t)stream);
  return rc ? launch_fail(rc, "draw_noise") : PF_OK;
}

int pf_profile_enable(pf_plan* p, int32_t enable) {
  if (!p) return fail(PF_EINVAL, "null plan");
  for (auto& v : p->ev) {
    for (auto& e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    v.clear();
  }
  p->prof = enable != 0;
  return PF_OK;
}

int pf_profile_select(pf_plan* p, uint32_t pass_mask) {
  if (!p) return fail(PF_EINVAL, "null plan");
  p->prof_mask = pass_mask;
  return PF_OK;
}

int pf_set_fusion(pf_plan* p, int32_t fuse_stage) {
  if (!p) return fail(PF_EINVAL, "null plan");
  p->fuse_stage = fuse_stage != 0;
  return PF_OK;
}

int pf_set_tiles(pf_plan* p, const int32_t* tiles) {
  if (!p || !tiles) return fail(PF_EINVAL, "null argument");
  for (int i = 0; i < PF_NPASS; ++i) {
    const bool ok = p->fp64 ? variant_ok<double>(p, i, tiles[i]) : variant_ok<float>(p, i, tiles[i]);
    if (!ok)
      return fail(PF_EUNSUPPORTED, "launch variant " + std::to_string(tiles[i]) + " of pass " + std::to_string(i) +
                                       " does not exist at this plan's lengths");
  }
  for (int i = 0; i < PF_NPASS; ++i) p->tsel[i] = tiles[i];
  return PF_OK;
}

int pf_get_tiles(const pf_plan* p, int32_t* tiles, int32_t* fuse_stage) {
  if (!p || !tiles) return fail(PF_EINVAL, "null argument");
  for (int i = 0; i < PF_NPASS; ++i) tiles[i] = p->tsel[i];
  if (fuse_stage) *fuse_stage = p->fuse_stage ? 1 : 0;
  return PF_OK;
}

int pf_autotune(pf_plan* p, const void* uhat, int32_t reps, double* ms_out, void* stream) {
  if (!p || !uhat || reps < 1) return fail(PF_EINVAL, "invalid argument");
  if (p->nact == 0) return PF_OK;
  hipStream_t s = (hipStream_t)stream;
  return p->fp64 ? autotune_impl<double>(p, uhat, reps, ms_out, s) : autotune_impl<float>(p, uhat, reps, ms_out, s);
}

int pf_profile_read(pf_plan* p, double* total_ms, int64_t* count, int32_t n) {
  if (!p || !total_ms || !count) return fail(PF_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) {
    total_ms[i] = 0.0;
    count[i] = 0;
    if (i >= PF_NPASS) continue;
    for (auto& e : p->ev[i]) {
      PF_HIP(hipEventSynchronize(e.second));
      float ms = 0.f;
      PF_HIP(hipEventElapsedTime(&ms, e.first, e.second));
      total_ms[i] += ms;
      count[i] += 1;
    }
  }
  return PF_OK;
}

int pf_diagnostics(pf_plan* p, const void* uhat, double* out, void* stream) {
  if (!p || !uhat || !out) return fail(PF_EINVAL, "invalid argument");
  DiagArgs a{};
  a.uhat = uhat; a.out = out; a.nk = p->nk;
  a.N0 = p->N[0]; a.N1 = p->N[1]; a.N2 = p->dim == 3 ? p->N[2] : p->N[1]; a.nh = p->nh;
  a.ks0 = p->ks[0]; a.ks1 = p->ks[1]; a.ks2 = p->ks[2];
  const int rc = p->fp64 ? launch_diag<double>(p->dim, a, p->nact, (hipStream_t)stream)
                         : launch_diag<float>(p->dim, a, p->nact, (hipStream_t)stream);
  return rc ? launch_fail(rc, "diagnostics") : PF_OK;
}

int pf_spectrum(pf_plan* p, const void* uhat, double* out, int32_t nshell, void* stream) {
  if (!p || !uhat || !out || nshell < 1) return fail(PF_EINVAL, "invalid argument");
  DiagArgs a{};
  a.uhat = uhat; a.out = out; a.nk = p->nk;
  a.N0 = p->N[0]; a.N1 = p->N[1]; a.N2 = p->dim == 3 ? p->N[2] : p->N[1]; a.nh = p->nh;
  a.ks0 = p->ks[0]; a.ks1 = p->ks[1]; a.ks2 = p->ks[2];
  const int rc = p->fp64 ? launch_spectrum<double>(p->dim, a, p->nact, nshell, (hipStream_t)stream)
                         : launch_spectrum<float>(p->dim, a, p->nact, nshell, (hipStream_t)stream);
  return rc ? launch_fail(rc, "spectrum") : PF_OK;
}

int pf_accumulate_sk(pf_plan* p, const void* uhat, double* acc, void* stream) {
  if (!p || !uhat || !acc) return fail(PF_EINVAL, "invalid argument");
  const int rc = p->fp64 ? launch_sk<double>(p->dim, uhat, acc, p->nk, p->nact, (hipStream_t)stream)
                         : launch_sk<float>(p->dim, uhat, acc, p->nk, p->nact, (hipStream_t)stream);
  return rc ? launch_fail(rc, "accumulate_sk") : PF_OK;
}

}  // extern "C"

// ------------------------------------------------------ allocation-free entry
// pf_rhs_nonlinear_noalloc keeps its twiddle tables and work buffers in one
// file-scope device array instead of memory it allocates.  Each GPU holds its
// own copy of the array (from when the library's code is loaded there until
// the process ends).  Host code must never pass `scratch` itself to a kernel
// or a copy: the name stands for a host-side shadow.  The device address,
// looked up once per device with hipGetSymbolAddress, is what the kernels get.
#define PF_SCRATCH_BYTES (128ull << 20)
static __device__ __align__(256) unsigned char scratch[PF_SCRATCH_BYTES];

namespace {

constexpr size_t kTwSlot = 32 * 1024;  // bytes per twiddle table (length <= 1536 in FP64: 24 KB)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on its first use there).
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return failf(PF_EINVAL, "device index %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(PF_EHIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_dev[dev] = ptr;
  }
  *out = g_scratch_dev[dev];
  return PF_OK;
}

// The twiddle table of length L (layout of make_twiddles) computed on the
// device in double precision: entry OFF(q) + (r - 1) Ns + k of pass q >= 1 is
// exp(-2 pi i r k / (Ns R)).
template <typename T> __global__ void __launch_bounds__(256) twiddle_kernel(cplx<T>* out, int n, RadixPlan pl) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  int off = 0, ns = pl.r[0];
  for (int q = 1; q < pl.n; ++q) {
    const int R = pl.r[q];
    const int cnt = (R - 1) * ns;
    if (m < off + cnt) {
      const int e = m - off;
      const int r = e / ns + 1, k = e % ns;
      double sn, cs;
      sincospi(2.0 * (double)(r * k) / (double)(ns * R), &sn, &cs);
      out[m] = mk<T>((T)cs, (T)-sn);
      return;
    }
    off += cnt;
    ns *= R;
  }
}

template <typename T> int noalloc_rhs(pf_plan* p, const void* uhat, void* rhs, hipStream_t s) {
  for (int t = 0; t < p->ntw; ++t) {
    const int L = p->tw[t].L, n = tw_size(L);
    if (n == 0) continue;
    hipLaunchKernelGGL(twiddle_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, s, (cplx<T>*)p->tw[t].ptr, n,
                       make_plan(L));
    PF_HIP(hipGetLastError());
  }
  int rc = nonlinear_unprojected<T>(p, uhat, s);
  if (rc) return rc;
  const StageArgs g = geometry(p);
  rc = launch_project<T>(p->dim, g, p->w1, rhs, p->d.mask_nyquist, 0, p->nact, s);
  return rc ? launch_fail(rc, "project") : PF_OK;
}

}  // namespace

extern "C" {

int pf_rhs_nonlinear_noalloc(const pf_desc* desc, const void* u_hat, void* rhs, void* stream) {
  if (!desc || !u_hat || !rhs) return fail(PF_EINVAL, "invalid argument");
  if (!desc->nonlinear) return fail(PF_EINVAL, "descriptor has nonlinear = 0");
  pf_plan p;  // geometry only: its tables stay empty, nothing is allocated
  long long w1c = 0, w2c = 0;
  int rc = plan_geometry(&p, *desc, w1c, w2c);
  if (rc) return rc;
  int lens[6];
  int nl = 0;
  for (int i = 0; i < p.dim; ++i)
    for (int L : {p.N[i], p.M[i]}) {
      bool seen = false;
      for (int t = 0; t < nl; ++t) seen = seen || lens[t] == L;
      if (!seen) lens[nl++] = L;
    }
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  p.w1_bytes = (size_t)w1c * p.csz;
  p.w2_bytes = (size_t)w2c * p.csz;
  const size_t tw_bytes = (size_t)nl * kTwSlot;
  const size_t need = tw_bytes + up(p.w1_bytes) + up(p.w2_bytes);
  if (need > PF_SCRATCH_BYTES)
    return failf(PF_EINVAL,
                 "pf_rhs_nonlinear_noalloc: this grid and batch need %zu bytes of work memory; the static scratch "
                 "holds %llu (create a plan and call pf_rhs_nonlinear instead)",
                 need, (unsigned long long)PF_SCRATCH_BYTES);
  void* base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  char* b = (char*)base;
  for (int t = 0; t < nl; ++t) {
    if ((size_t)tw_size(lens[t]) * p.csz > kTwSlot) return failf(PF_EINVAL, "twiddle table of length %d too large", lens[t]);
    p.tw[p.ntw++] = {lens[t], 0, b + (size_t)t * kTwSlot};
  }
  p.w1 = b + tw_bytes;
  p.w2 = w2c ? b + tw_bytes + up(p.w1_bytes) : nullptr;
  hipStream_t s = (hipStream_t)stream;
  return p.fp64 ? noalloc_rhs<double>(&p, u_hat, rhs, s) : noalloc_rhs<float>(&p, u_hat, rhs, s);
}

}  // extern "C"
