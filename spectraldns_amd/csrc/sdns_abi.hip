// C-ABI of libsdns_hip.so (declared in include/sdns.h): plans, workspace,
// twiddle / wavenumber tables and the orchestration of the pass kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/sdns.h"
#include "sdns_launch.hpp"

using namespace sdns;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(SDNS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

bool strided_supported(int M) {
  switch (M) {
#define SDNS_X(m) case m:
    SDNS_STRIDED_SIZES(SDNS_X)
#undef SDNS_X
    return true;
    default: return false;
  }
}
bool half_supported(int n) {
  switch (n) {
#define SDNS_X(m) case m:
    SDNS_HALF_SIZES(SDNS_X)
#undef SDNS_X
    return true;
    default: return false;
  }
}

// Kept wavenumbers of one axis, as (neg, pos): modes -neg..pos.  Same rules
// as the oracle's keep_ranges.
void keep_ranges(int N, int dealias, int mode, bool r2c, int in[2], int out[2]) {
  const int half = N / 2;
  const int full_neg = r2c ? 0 : half, full_pos = r2c ? half : half - 1;
  out[0] = full_neg; out[1] = full_pos;
  in[0] = full_neg; in[1] = full_pos;
  if (dealias == SDNS_DEALIAS_23) {
    if (mode == SDNS_MODE_DIRECT) {
      const int K = (N + 2) / 3 - 1;  // largest k with 3k < N
      in[0] = r2c ? 0 : K; in[1] = K;
    } else {
      int K = (2 * (half + 1) - 1) / 3;  // largest k with 3k < 2(N/2+1)
      in[1] = out[1] = r2c ? std::min(K, half) : std::min(K, half - 1);
      in[0] = out[0] = r2c ? 0 : std::min(K, half);
    }
  } else if (dealias == SDNS_DEALIAS_32) {
    in[0] = r2c ? 0 : half - 1; in[1] = half - 1;
  }
}

}  // namespace

struct sdns_plan {
  int N[3], M[3], Nh, n2h, nz_plain;  // n2h = M2/2 (z half length), nz_plain = N2/2
  double L[3];
  int precision, dealias, mask_nyquist, mode;
  size_t esz;  // bytes per complex element
  hipStream_t stream = nullptr;
  int in[3][2], out[3][2];
  int nky, nkz, nkyo, nkzo;
  int nkzp, nkzop;  // kz strides of the work buffers, padded to 16 elements
  int var[6] = {0, 0, 0, 0, 0, 0};  // schedule variant per pass (see sdns_plan_set_variant)
  // device tables
  std::map<int, void*> tw;       // forward twiddles by line length
  std::map<int, void*> rtw;      // real-transform twiddles by real length
  void* ktab[3] = {nullptr, nullptr, nullptr};
  // workspace
  size_t szA = 0, szB = 0, szS = 0, szPart = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  bool ws_owned = false;
  char *bufA = nullptr, *bufB = nullptr, *stB = nullptr, *stC = nullptr;
  double* part = nullptr;
  int max_blocks = 0;
  // per-pass event timing
  bool timing = false;
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<Rec> rec;
  std::vector<hipEvent_t> pool;
  int cur_id = 0;
  hipEvent_t cur_a = nullptr;
};

namespace {

// exp(+2 pi i m / len) in long double: exact integer reduction to a quadrant,
// complement to [0, pi/4], then libm cosl/sinl.
void host_cexp2pi(long m, long len, long double* c, long double* s) {
  m %= len;
  if (m < 0) m += len;
  const long a = 4 * m, q = a / len, rem = a - q * len;
  const long double hp = 1.57079632679489661923132169163975144L;
  long double cc, ss;
  if (2 * rem <= len) { const long double phi = hp * (long double)rem / (long double)len; cc = cosl(phi); ss = sinl(phi); }
  else { const long double phi = hp * (long double)(len - rem) / (long double)len; cc = sinl(phi); ss = cosl(phi); }
  switch (q) {
    case 0: *c = cc; *s = ss; break;
    case 1: *c = -ss; *s = cc; break;
    case 2: *c = -cc; *s = -ss; break;
    default: *c = ss; *s = -cc; break;
  }
}

// Forward twiddles exp(-2 pi i m / len): m in [0, len) for complex line
// lengths, m in [0, len/2) for the real-transform table of real length len.
template <typename Rt> int upload_twiddles(sdns_plan* p, int len, bool real_tw) {
  std::map<int, void*>& tab = real_tw ? p->rtw : p->tw;
  if (tab.count(len)) return 0;
  const int cnt = real_tw ? len / 2 : len;
  std::vector<Rt> h(2 * (size_t)cnt);
  for (int m = 0; m < cnt; ++m) {
    long double c, s;
    host_cexp2pi(m, len, &c, &s);
    h[2 * m] = (Rt)c;
    h[2 * m + 1] = (Rt)(-s);
  }
  void* d = nullptr;
  HIP_TRY(hipMalloc(&d, h.size() * sizeof(Rt)));
  HIP_TRY(hipMemcpy(d, h.data(), h.size() * sizeof(Rt), hipMemcpyHostToDevice));
  tab[len] = d;
  return 0;
}

template <typename Rt> int upload_k(sdns_plan* p) {
  for (int ax = 0; ax < 3; ++ax) {
    const int N = p->N[ax];
    const int cnt = ax == 2 ? N / 2 + 1 : N;
    const double fac = 2.0 * M_PI / p->L[ax];
    std::vector<Rt> h(cnt);
    for (int i = 0; i < cnt; ++i) {
      const int k = (ax == 2 || i < N / 2) ? i : i - N;
      h[i] = (Rt)((double)k * fac);
    }
    HIP_TRY(hipMalloc(&p->ktab[ax], cnt * sizeof(Rt)));
    HIP_TRY(hipMemcpy(p->ktab[ax], h.data(), cnt * sizeof(Rt), hipMemcpyHostToDevice));
  }
  return 0;
}

template <typename Rt> int build_tables(sdns_plan* p) {
  int rc;
  if ((rc = upload_twiddles<Rt>(p, p->M[0], false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->M[1], false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->n2h, false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->M[2], true))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->N[0], false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->N[1], false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->nz_plain, false))) return rc;
  if ((rc = upload_twiddles<Rt>(p, p->N[2], true))) return rc;
  return upload_k<Rt>(p);
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

void compute_sizes(sdns_plan* p) {
  const size_t e = p->esz;
  const size_t M0 = p->M[0], M1 = p->M[1];
  const size_t state = 3ull * p->N[0] * p->N[1] * p->Nh;
  const size_t w1 = 5ull * M0 * p->nky * p->nkzp, w3 = 3ull * M0 * M1 * p->nkzop;
  const size_t w2 = 6ull * M0 * M1 * p->nkzp, w4 = 3ull * M0 * p->nkyo * p->nkzop;
  p->szA = align256(std::max(std::max(w1, w3), state) * e);
  p->szB = align256(std::max(std::max(w2, w4), state) * e);
  p->szS = align256(state * e);
  // P5 blocks: columns N1*Nh over the smallest tile of the supported lengths
  p->max_blocks = p->N[1] * p->Nh;  // upper bound (>= 1 column per block)
  p->szPart = align256((size_t)std::max(p->max_blocks, 4096) * 2 * sizeof(double) + 4096);
}

int assign_workspace(sdns_plan* p, void* base, size_t bytes) {
  const size_t need = p->szA + p->szB + 2 * p->szS + p->szPart;
  if (bytes < need) return fail(SDNS_ERR_INVALID, "workspace of %zu bytes, need %zu", bytes, need);
  if (((uintptr_t)base) & 255) return fail(SDNS_ERR_INVALID, "workspace must be 256-byte aligned");
  char* c = (char*)base;
  p->bufA = c; c += p->szA;
  p->bufB = c; c += p->szB;
  p->stB = c; c += p->szS;
  p->stC = c; c += p->szS;
  p->part = (double*)c;
  p->ws = base;
  p->ws_bytes = bytes;
  return 0;
}

int ensure_workspace(sdns_plan* p) {
  if (p->ws) return 0;
  const size_t need = p->szA + p->szB + 2 * p->szS + p->szPart;
  void* d = nullptr;
  if (hipMalloc(&d, need) != hipSuccess) return fail(SDNS_ERR_NOMEM, "hipMalloc of %zu bytes failed", need);
  p->ws_owned = true;
  return assign_workspace(p, d, need);
}

RowMap rmap(int neg, int pos, int base) { return RowMap{neg, pos, base}; }

hipEvent_t pool_event(sdns_plan* p) {
  if (!p->pool.empty()) { hipEvent_t e = p->pool.back(); p->pool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
void tstart(sdns_plan* p, int id) {
  if (!p->timing) return;
  p->cur_a = pool_event(p);
  p->cur_id = id;
  (void)hipEventRecord(p->cur_a, p->stream);
}
void tstop(sdns_plan* p) {
  if (!p->timing) return;
  hipEvent_t b = pool_event(p);
  (void)hipEventRecord(b, p->stream);
  p->rec.push_back({p->cur_id, p->cur_a, b});
}

// ---------------------------------------------------------------- transforms
template <typename Rt> int plain_backward(sdns_plan* p, const void* uh, void* up, int nc) {
  const int N0 = p->N[0], N1 = p->N[1], Nh = p->Nh, N2 = p->N[2];
  const long plane = (long)N1 * Nh, fs = (long)N0 * plane;
  StridedArgs<Rt> a{};
  // x inverse: columns (iy, kz) contiguous
  a.in = uh; a.out = p->bufA; a.tw = p->tw[N0];
  a.Q = (int)plane; a.cdiv = (int)plane; a.cvalid = (int)plane; a.in_hi = 0; a.out_hi = 0;
  a.in_row = plane; a.out_row = plane; a.in_field = fs; a.out_field = fs;
  a.rin = rmap(N0 / 2, N0 / 2 - 1, N0); a.rout = a.rin; a.scale = Rt(1);
  if (launch_strided<Rt>(N0, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", N0);
  // y inverse: columns (x, kz)
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw[N1];
  a.Q = N0 * Nh; a.cdiv = Nh; a.cvalid = Nh; a.in_hi = plane; a.out_hi = plane; a.in_row = Nh; a.out_row = Nh;
  a.rin = rmap(N1 / 2, N1 / 2 - 1, N1); a.rout = a.rin;
  if (launch_strided<Rt>(N1, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", N1);
  // z C2R
  ZArgs<Rt> z{};
  z.in = p->bufB; z.out = up; z.tw = p->tw[N2 / 2]; z.rtw = p->rtw[N2];
  z.L = N0 * N1; z.nk = Nh; z.in_line = Nh; z.in_field = fs; z.out_line = N2 / 2;
  z.out_field = (long)N0 * N1 * (N2 / 2); z.scale = Rt(1);
  if (launch_z<Rt>(N2 / 2, p->var[5], +1, nc, z, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "z length %d", N2);
  return 0;
}

template <typename Rt> int plain_forward(sdns_plan* p, const void* up, void* uh, int nc) {
  const int N0 = p->N[0], N1 = p->N[1], Nh = p->Nh, N2 = p->N[2];
  const long plane = (long)N1 * Nh, fs = (long)N0 * plane;
  ZArgs<Rt> z{};
  z.in = up; z.out = p->bufA; z.tw = p->tw[N2 / 2]; z.rtw = p->rtw[N2];
  z.L = N0 * N1; z.nk = Nh; z.in_line = N2 / 2; z.in_field = (long)N0 * N1 * (N2 / 2);
  z.out_line = Nh; z.out_field = fs; z.scale = Rt(1);
  if (launch_z<Rt>(N2 / 2, p->var[5], -1, nc, z, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "z length %d", N2);
  StridedArgs<Rt> a{};
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw[N1];
  a.Q = N0 * Nh; a.cdiv = Nh; a.cvalid = Nh; a.in_hi = plane; a.out_hi = plane; a.in_row = Nh; a.out_row = Nh;
  a.in_field = fs; a.out_field = fs;
  a.rin = rmap(N1 / 2, N1 / 2 - 1, N1); a.rout = a.rin; a.scale = Rt(1);
  if (launch_strided<Rt>(N1, p->var[5], -1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", N1);
  a.in = p->bufB; a.out = uh; a.tw = p->tw[N0];
  a.Q = (int)plane; a.cdiv = (int)plane; a.cvalid = (int)plane; a.in_hi = 0; a.out_hi = 0; a.in_row = plane;
  a.out_row = plane;
  a.rin = rmap(N0 / 2, N0 / 2 - 1, N0); a.rout = a.rin;
  a.scale = (Rt)(1.0 / ((double)N0 * N1 * N2));
  if (launch_strided<Rt>(N0, p->var[5], -1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", N0);
  return 0;
}

// P1..P4 of one RHS evaluation from the state u (device) into W4 (= bufB).
template <typename Rt> int conv_passes(sdns_plan* p, const void* u) {
  const int M0 = p->M[0], M1 = p->M[1];
  P1Args<Rt> a1{};
  a1.u = u; a1.w1 = p->bufA; a1.tw = p->tw[M0]; a1.kx = (const Rt*)p->ktab[0];
  a1.N0 = p->N[0]; a1.N1 = p->N[1]; a1.Nh = p->Nh;
  a1.x = rmap(p->in[0][0], p->in[0][1], p->N[0]);
  a1.y_neg = p->in[1][0]; a1.y_pos = p->in[1][1]; a1.nkz = p->nkz; a1.nkzp = p->nkzp; a1.Q = p->nky * p->nkzp;
  tstart(p, 0);
  if (launch_p1<Rt>(M0, p->var[0], a1, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", M0);
  tstop(p);
  P2Args<Rt> a2{};
  a2.w1 = p->bufA; a2.w2 = p->bufB; a2.tw = p->tw[M1]; a2.ky = (const Rt*)p->ktab[1]; a2.kz = (const Rt*)p->ktab[2];
  a2.N1 = p->N[1]; a2.M0 = M0; a2.y = rmap(p->in[1][0], p->in[1][1], p->nky); a2.nkz = p->nkz; a2.nkzp = p->nkzp;
  a2.Q = M0 * p->nkzp;
  tstart(p, 1);
  if (launch_p2<Rt>(M1, p->var[1], a2, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  tstop(p);
  P3Args<Rt> a3{};
  a3.w2 = p->bufB; a3.w3 = p->bufA; a3.tw = p->tw[p->n2h]; a3.rtw = p->rtw[p->M[2]];
  a3.nkz = p->nkz; a3.nkzp = p->nkzp; a3.nkzo = p->nkzo; a3.nkzop = p->nkzop; a3.L = M0 * M1;
  tstart(p, 2);
  if (launch_p3<Rt>(p->n2h, p->var[2], a3, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "z length %d", p->M[2]);
  tstop(p);
  StridedArgs<Rt> a4{};
  a4.in = p->bufA; a4.out = p->bufB; a4.tw = p->tw[M1];
  a4.Q = M0 * p->nkzop; a4.cdiv = p->nkzop; a4.cvalid = p->nkzo;
  a4.in_hi = (long)M1 * p->nkzop; a4.in_row = p->nkzop; a4.in_field = (long)M0 * M1 * p->nkzop;
  a4.out_hi = (long)p->nkyo * p->nkzop; a4.out_row = p->nkzop; a4.out_field = (long)M0 * p->nkyo * p->nkzop;
  a4.rin = rmap(M1 / 2, M1 - M1 / 2 - 1, M1);
  a4.rout = rmap(p->out[1][0], p->out[1][1], p->nkyo);
  a4.scale = Rt(1);
  tstart(p, 3);
  if (launch_strided<Rt>(M1, p->var[3], -1, 3, a4, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  tstop(p);
  return 0;
}

template <typename Rt> P5Args<Rt> p5_args(sdns_plan* p, double nu) {
  P5Args<Rt> a{};
  a.w4 = p->bufB; a.tw = p->tw[p->M[0]];
  a.kx = (const Rt*)p->ktab[0]; a.ky = (const Rt*)p->ktab[1]; a.kz = (const Rt*)p->ktab[2];
  a.N0 = p->N[0]; a.N1 = p->N[1]; a.Nh = p->Nh; a.N2 = p->N[2];
  a.x = rmap(p->out[0][0], p->out[0][1], p->N[0]);
  a.y_neg = p->out[1][0]; a.y_pos = p->out[1][1]; a.nkzo = p->nkzo; a.nkzop = p->nkzop;
  a.mask_nyquist = p->mask_nyquist;
  a.nu = (Rt)nu;
  a.scale = (Rt)(1.0 / ((double)p->M[0] * p->M[1] * p->M[2]));
  a.Q = p->N[1] * p->Nh;
  return a;
}

template <typename Rt> int compute_rhs(sdns_plan* p, const void* u, void* rhs, double nu, const void* src, void* ph) {
  int rc;
  if ((rc = conv_passes<Rt>(p, u))) return rc;
  P5Args<Rt> a = p5_args<Rt>(p, nu);
  a.A = const_cast<void*>(u); a.rhs = rhs; a.phat = ph; a.source = src;
  tstart(p, 4);
  if (launch_p5<Rt>(p->M[0], p->var[4], P5_RHS, a, p->stream, nullptr))
    return fail(SDNS_ERR_UNSUPPORTED, "x length %d", p->M[0]);
  tstop(p);
  return 0;
}

template <typename Rt> int convection(sdns_plan* p, const void* u, void* rhs) {
  int rc;
  if ((rc = conv_passes<Rt>(p, u))) return rc;
  P5Args<Rt> a = p5_args<Rt>(p, 0.0);
  a.A = const_cast<void*>(u); a.rhs = rhs; a.conv_only = 1;
  if (launch_p5<Rt>(p->M[0], p->var[4], P5_RHS, a, p->stream, nullptr))
    return fail(SDNS_ERR_UNSUPPORTED, "x length %d", p->M[0]);
  return 0;
}

template <typename Rt> int rk4_steps(sdns_plan* p, void* u, const double* dts, int nsteps, double nu, double* diag,
                                      void* ph) {
  // a, b and dt in the working precision (integrators.py:185-186, config.py:123-127)
  const Rt a[4] = {Rt(1. / 6.), Rt(1. / 3.), Rt(1. / 3.), Rt(1. / 6.)};
  const Rt b[3] = {Rt(0.5), Rt(0.5), Rt(1.)};
  const Rt nur = (Rt)nu;
  for (int s = 0; s < nsteps; ++s) {
    const Rt dt = (Rt)dts[s];
    for (int rk = 0; rk < 4; ++rk) {
      int rc;
      if ((rc = conv_passes<Rt>(p, rk == 0 ? u : (const void*)p->stB))) return rc;
      P5Args<Rt> pa = p5_args<Rt>(p, nu);
      pa.A = u; pa.B = p->stB; pa.Cacc = p->stC;
      pa.adt = a[rk] * dt;
      pa.bdt = rk < 3 ? b[rk] * dt : Rt(0);
      const int mode = rk == 0 ? P5_RK0 : (rk < 3 ? P5_RK12 : P5_RK3);
      pa.partials = (rk == 3 && diag) ? p->part : nullptr;
      pa.phat = (rk == 3 && s == nsteps - 1) ? ph : nullptr;
      int nblk = 0;
      tstart(p, 4);
      if (launch_p5<Rt>(p->M[0], p->var[4], mode, pa, p->stream, &nblk)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", p->M[0]);
      tstop(p);
      if (rk == 3 && diag) {
        tstart(p, 5);
        if (launch_sum_partials(p->part, nblk, diag + 2 * s, 0.5, (double)nur, p->stream))
          return fail(SDNS_ERR_HIP, "partial sum launch");
        tstop(p);
      }
    }
  }
  return 0;
}

template <typename Rt> EwArgs<Rt> ew_args(sdns_plan* p) {
  EwArgs<Rt> a{};
  a.kx = (const Rt*)p->ktab[0]; a.ky = (const Rt*)p->ktab[1]; a.kz = (const Rt*)p->ktab[2];
  a.N0 = p->N[0]; a.N1 = p->N[1]; a.Nh = p->Nh; a.N2 = p->N[2];
  a.n = (long)p->N[0] * p->N[1] * p->Nh;
  a.mask_nyquist = p->mask_nyquist;
  return a;
}

template <typename Rt> int energy(sdns_plan* p, const void* u, double nu, double* out) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  const int nblk = 1024;
  launch_energy<Rt>(a, p->part, nblk, p->stream);
  if (launch_sum_partials(p->part, nblk, out, 0.5, (double)(Rt)nu, p->stream)) return fail(SDNS_ERR_HIP, "sum");
  return 0;
}

template <typename Rt> int force(sdns_plan* p, void* u, double target, double kf, double* sums) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  a.out = u;
  const Rt kf2 = (Rt)(kf * kf);
  double* s = sums ? sums : (double*)((char*)p->part + p->szPart - 4096);
  const int nblk = 1024;
  launch_force_reduce<Rt>(a, kf2, p->part, nblk, p->stream);
  if (launch_sum_partials(p->part, nblk, s, 1.0, 1.0, p->stream)) return fail(SDNS_ERR_HIP, "sum");
  launch_force_scale<Rt>(a, kf2, s, target, p->stream);
  return 0;
}

template <typename Rt> int spectrum(sdns_plan* p, const void* u, int nb, double* out) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  const size_t avail = p->szPart - 4096;
  int nblk = (int)std::min<size_t>(256, avail / (2 * (size_t)nb * sizeof(double)));
  if (nblk < 1) return fail(SDNS_ERR_INVALID, "nbins %d too large", nb);
  launch_spectrum<Rt>(a, nb, p->part, nblk, p->stream);
  if (launch_sum_rows(p->part, nblk, 2 * nb, out, p->stream)) return fail(SDNS_ERR_HIP, "sum rows");
  return 0;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SDNS_ERR_HIP, "kernel launch: %s", hipGetErrorString(e));
  return 0;
}

}  // namespace

namespace sdns {
// Deterministic final sum of per-block (energy, enstrophy) partials.
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ part, int nblk,
                                                      double* __restrict__ out, double e_scale, double o_scale) {
  __shared__ double s[2][256];
  double a0 = 0, a1 = 0;
  for (int i = threadIdx.x; i < nblk; i += 256) { a0 += part[2 * i]; a1 += part[2 * i + 1]; }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { s[0][threadIdx.x] += s[0][threadIdx.x + w]; s[1][threadIdx.x] += s[1][threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s[0][0] * e_scale; out[1] = s[1][0] * o_scale; }
}
int launch_sum_partials(const double* part, int nblk, double* out, double e_scale, double o_scale, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, nblk, out, e_scale, o_scale);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
// out[j] = sum over b of part[b*width + j], fixed order (deterministic).
__global__ void __launch_bounds__(256) k_sum_rows(const double* __restrict__ part, int nblk, int width,
                                                  double* __restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= width) return;
  double s = 0;
  for (int b = 0; b < nblk; ++b) s += part[(long)b * width + j];
  out[j] = s;
}
int launch_sum_rows(const double* part, int nblk, int width, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_rows, dim3((width + 255) / 256), dim3(256), 0, s, part, nblk, width, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
}  // namespace sdns

extern "C" {

const char* sdns_last_error(void) { return g_err.c_str(); }

int sdns_plan_create(sdns_plan** plan, const int32_t N[3], const double L[3], int precision, int dealias,
                     int mask_nyquist, int dealias_mode, void* hip_stream) {
  if (!plan || !N || !L) return fail(SDNS_ERR_INVALID, "null argument");
  *plan = nullptr;
  if (precision != SDNS_SINGLE && precision != SDNS_DOUBLE) return fail(SDNS_ERR_INVALID, "precision %d", precision);
  if (dealias < 0 || dealias > 2) return fail(SDNS_ERR_INVALID, "dealias %d", dealias);
  if (dealias_mode < 0 || dealias_mode > 1) return fail(SDNS_ERR_INVALID, "dealias_mode %d", dealias_mode);
  for (int i = 0; i < 3; ++i) {
    if (N[i] < 4 || N[i] % 2) return fail(SDNS_ERR_UNSUPPORTED, "N[%d] = %d must be even and >= 4", i, N[i]);
    if (!(L[i] > 0)) return fail(SDNS_ERR_INVALID, "L[%d] must be positive", i);
  }
  sdns_plan* p = new sdns_plan();
  for (int i = 0; i < 3; ++i) {
    p->N[i] = N[i];
    p->L[i] = L[i];
    p->M[i] = dealias == SDNS_DEALIAS_32 ? 3 * N[i] / 2 : N[i];
  }
  p->Nh = N[2] / 2 + 1;
  p->n2h = p->M[2] / 2;
  p->nz_plain = N[2] / 2;
  p->precision = precision;
  p->dealias = dealias;
  p->mask_nyquist = mask_nyquist ? 1 : 0;
  p->mode = dealias_mode;
  p->esz = precision == SDNS_DOUBLE ? 16 : 8;
  p->stream = (hipStream_t)hip_stream;
  for (int i = 0; i < 3; ++i) keep_ranges(N[i], dealias, dealias_mode, i == 2, p->in[i], p->out[i]);
  p->nky = p->in[1][0] + p->in[1][1] + 1;
  p->nkz = p->in[2][1] + 1;
  p->nkyo = p->out[1][0] + p->out[1][1] + 1;
  p->nkzo = p->out[2][1] + 1;
  p->nkzp = (p->nkz + 15) / 16 * 16;
  p->nkzop = (p->nkzo + 15) / 16 * 16;
  const int lens[] = {p->M[0], p->M[1], p->N[0], p->N[1]};
  for (int l : lens)
    if (!strided_supported(l)) { delete p; return fail(SDNS_ERR_UNSUPPORTED, "no kernel for line length %d", l); }
  if (!half_supported(p->n2h) || !half_supported(p->nz_plain)) {
    delete p;
    return fail(SDNS_ERR_UNSUPPORTED, "no kernel for z length %d / %d", p->M[2], p->N[2]);
  }
  const long long cols = (long long)std::max(p->M[0], p->M[1]) * std::max(p->M[1], p->N[1]) * p->Nh;
  if (cols > 0x7fffffffLL) { delete p; return fail(SDNS_ERR_UNSUPPORTED, "grid too large"); }
  const int rc = precision == SDNS_DOUBLE ? build_tables<double>(p) : build_tables<float>(p);
  if (rc) { sdns_plan_destroy(p); return rc; }
  compute_sizes(p);
  *plan = p;
  return 0;
}

int sdns_plan_destroy(sdns_plan* p) {
  if (!p) return 0;
  for (auto& r : p->rec) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (auto e : p->pool) (void)hipEventDestroy(e);
  for (auto& kv : p->tw) (void)hipFree(kv.second);
  for (auto& kv : p->rtw) (void)hipFree(kv.second);
  for (int i = 0; i < 3; ++i)
    if (p->ktab[i]) (void)hipFree(p->ktab[i]);
  if (p->ws_owned && p->ws) (void)hipFree(p->ws);
  delete p;
  return 0;
}

int sdns_plan_set_variant(sdns_plan* p, int pass, int variant) {
  if (!p) return fail(SDNS_ERR_INVALID, "null plan");
  const int maxv = pass == 2 ? 3 : 1;
  if (pass < 0 || pass > 5 || variant < 0 || variant > maxv)
    return fail(SDNS_ERR_INVALID, "pass %d variant %d", pass, variant);
  p->var[pass] = variant;
  return 0;
}

int sdns_plan_set_stream(sdns_plan* p, void* s) {
  if (!p) return fail(SDNS_ERR_INVALID, "null plan");
  p->stream = (hipStream_t)s;
  return 0;
}

size_t sdns_workspace_bytes(const sdns_plan* p) {
  return p ? p->szA + p->szB + 2 * p->szS + p->szPart : 0;
}

int sdns_set_workspace(sdns_plan* p, void* d, size_t bytes) {
  if (!p || !d) return fail(SDNS_ERR_INVALID, "null argument");
  if (p->ws_owned && p->ws) { (void)hipFree(p->ws); p->ws_owned = false; }
  p->ws = nullptr;
  return assign_workspace(p, d, bytes);
}

int sdns_forward(sdns_plan* p, const void* up, void* uh, int nc) {
  if (!p || !up || !uh || nc < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  if (nc > 3) return fail(SDNS_ERR_INVALID, "ncomp %d > 3", nc);
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? plain_forward<double>(p, up, uh, nc) : plain_forward<float>(p, up, uh, nc);
  return rc ? rc : check_launch();
}

int sdns_backward(sdns_plan* p, const void* uh, void* up, int nc) {
  if (!p || !up || !uh || nc < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  if (nc > 3) return fail(SDNS_ERR_INVALID, "ncomp %d > 3", nc);
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? plain_backward<double>(p, uh, up, nc) : plain_backward<float>(p, uh, up, nc);
  return rc ? rc : check_launch();
}

static int ew(sdns_plan* p, int op, void* out, const void* a, const void* b, void* aux, double nu, long n) {
  if (!p || !out) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc;
  if (p->precision == SDNS_DOUBLE) {
    EwArgs<double> e = ew_args<double>(p);
    e.out = out; e.a = a; e.b = b; e.aux = aux; e.nu = nu;
    if (n >= 0) e.n = n;
    rc = launch_elementwise<double>(op, e, p->stream);
  } else {
    EwArgs<float> e = ew_args<float>(p);
    e.out = out; e.a = a; e.b = b; e.aux = aux; e.nu = (float)nu;
    if (n >= 0) e.n = n;
    rc = launch_elementwise<float>(op, e, p->stream);
  }
  if (rc) return fail(SDNS_ERR_INVALID, "op %d", op);
  return check_launch();
}

int sdns_cross1(sdns_plan* p, void* c, const void* a, const void* b, size_t npts) {
  if (!a || !b) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_CROSS1, c, a, b, nullptr, 0, (long)npts);
}
int sdns_cross2(sdns_plan* p, void* c, const void* a) {
  if (!a) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_CROSS2, c, a, nullptr, nullptr, 0, -1);
}
int sdns_add_pressure_diffusion(sdns_plan* p, void* rhs, const void* u, double nu, void* ph) {
  if (!u) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_PRESSURE_DIFFUSION, rhs, u, nullptr, ph, nu, -1);
}
int sdns_divergence(sdns_plan* p, void* d, const void* u) {
  if (!u) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_DIVERGENCE, d, u, nullptr, nullptr, 0, -1);
}
int sdns_mul_minus_i(sdns_plan* p, void* out, const void* a, size_t n) {
  if (!a) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_MUL_MINUS_I, out, a, nullptr, nullptr, 0, (long)n);
}
int sdns_mask_nyquist(sdns_plan* p, void* a) { return ew(p, EW_MASK_NYQUIST, a, nullptr, nullptr, nullptr, 0, -1); }

int sdns_compute_rhs(sdns_plan* p, const void* u, void* rhs, double nu, const void* src, void* ph) {
  if (!p || !u || !rhs) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? compute_rhs<double>(p, u, rhs, nu, src, ph)
                                   : compute_rhs<float>(p, u, rhs, nu, src, ph);
  return rc ? rc : check_launch();
}

int sdns_convection(sdns_plan* p, const void* u, void* rhs) {
  if (!p || !u || !rhs) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? convection<double>(p, u, rhs) : convection<float>(p, u, rhs);
  return rc ? rc : check_launch();
}

int sdns_rk4_steps(sdns_plan* p, void* u, void* u_lo, const double* dts, int nsteps, double nu, double* diag,
                   void* ph) {
  if (!p || !u || (!dts && nsteps > 0) || nsteps < 0) return fail(SDNS_ERR_INVALID, "bad argument");
  if (u_lo && p->precision == SDNS_SINGLE) return fail(SDNS_ERR_UNSUPPORTED, "compensated accumulation not built");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? rk4_steps<double>(p, u, dts, nsteps, nu, diag, ph)
                                   : rk4_steps<float>(p, u, dts, nsteps, nu, diag, ph);
  return rc ? rc : check_launch();
}

int sdns_force(sdns_plan* p, void* u, double target, double kf, double* sums) {
  if (!p || !u) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? force<double>(p, u, target, kf, sums) : force<float>(p, u, target, kf, sums);
  return rc ? rc : check_launch();
}

int sdns_spectrum(sdns_plan* p, const void* u, int nbins, double* out) {
  if (!p || !u || !out || nbins < 1 || nbins > 8192) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? spectrum<double>(p, u, nbins, out) : spectrum<float>(p, u, nbins, out);
  return rc ? rc : check_launch();
}

int sdns_timing_enable(sdns_plan* p, int enable) {
  if (!p) return fail(SDNS_ERR_INVALID, "null plan");
  p->timing = enable != 0;
  return 0;
}

int sdns_timing_read(sdns_plan* p, double ms[6], int count[6]) {
  if (!p || !ms || !count) return fail(SDNS_ERR_INVALID, "bad argument");
  for (int i = 0; i < 6; ++i) { ms[i] = 0; count[i] = 0; }
  for (auto& r : p->rec) {
    HIP_TRY(hipEventSynchronize(r.b));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, r.a, r.b));
    const int id = (r.id >= 0 && r.id < 6) ? r.id : 5;
    ms[id] += t;
    count[id] += 1;
    p->pool.push_back(r.a);
    p->pool.push_back(r.b);
  }
  p->rec.clear();
  return 0;
}

int sdns_energy(sdns_plan* p, const void* u, double nu, double* out) {
  if (!p || !u || !out) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? energy<double>(p, u, nu, out) : energy<float>(p, u, nu, out);
  return rc ? rc : check_launch();
}

}  // extern "C"
