// C-ABI of libsdns_hip.so (declared in include/sdns.h): plans, workspace,
// twiddle / wavenumber tables and the orchestration of the pass kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sdns.h"
#include "sdns_launch.hpp"
#include "sdns_closure.hpp"

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
  char *bufA = nullptr, *bufB = nullptr, *stB = nullptr, *stC = nullptr, *stClo = nullptr;
  double* part = nullptr;
  int max_blocks = 0;
  // per-pass event timing
  bool timing = false;
  int timing_mask = 63;  // pass ids recorded while timing (bit i = pass i)
  bool cur_rec = false;
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<Rec> rec;
  std::vector<hipEvent_t> pool;
  int cur_id = 0;
  hipEvent_t cur_a = nullptr;
};

// Learned-closure workspace of one coarse grid (kernels in sdns_closure.hip).
struct sdns_closure {
  int n[3] = {0, 0, 0};
  long P = 0;
  hipStream_t stream = nullptr;
  char* mem = nullptr;
  __bf16 *x0 = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr;
  __bf16 *gyb = nullptr, *gz2 = nullptr, *gz1 = nullptr;
  float *y = nullptr, *gy = nullptr, *t = nullptr, *wpart = nullptr, *bpart = nullptr;
  float *phys = nullptr, *tauphys = nullptr;
  char *tauhat = nullptr, *src = nullptr;  // 6 and 3 complex64 spectral fields
  double *lpart = nullptr, *loss = nullptr;
  sdns::ClosureWeights w{};
  bool have_params = false;
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
  if ((rc = upload_k<Rt>(p))) return rc;
  // hipMemcpy from pageable memory may return before its DMA lands; the plan
  // may run on any stream, so wait for the tables here
  HIP_TRY(hipDeviceSynchronize());
  return 0;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

void compute_sizes(sdns_plan* p) {
  const size_t e = p->esz;
  const size_t M0 = p->M[0], M1 = p->M[1];
  const size_t state = 3ull * p->N[0] * p->N[1] * p->Nh;
  // W1 holds 6 fields when the vorticity is given (sdns_convection_uw)
  const size_t w1 = 6ull * M0 * p->nky * p->nkzp, w3 = 3ull * M0 * M1 * p->nkzop;
  const size_t w2 = 6ull * M0 * M1 * p->nkzp, w4 = 3ull * M0 * p->nkyo * p->nkzop;
  p->szA = align256(std::max(std::max(w1, w3), state) * e);
  p->szB = align256(std::max(std::max(w2, w4), state) * e);
  p->szS = align256(state * e);
  // P5 blocks: columns N1*Nh over the smallest tile of the supported lengths
  p->max_blocks = p->N[1] * p->Nh;  // upper bound (>= 1 column per block)
  p->szPart = align256((size_t)std::max(p->max_blocks, 4096) * 2 * sizeof(double) + 4096);
}

int assign_workspace(sdns_plan* p, void* base, size_t bytes) {
  const size_t need = p->szA + p->szB + 3 * p->szS + p->szPart;
  if (bytes < need) return fail(SDNS_ERR_INVALID, "workspace of %zu bytes, need %zu", bytes, need);
  if (((uintptr_t)base) & 255) return fail(SDNS_ERR_INVALID, "workspace must be 256-byte aligned");
  char* c = (char*)base;
  p->bufA = c; c += p->szA;
  p->bufB = c; c += p->szB;
  p->stB = c; c += p->szS;
  p->stC = c; c += p->szS;
  p->stClo = c; c += p->szS;
  p->part = (double*)c;
  p->ws = base;
  p->ws_bytes = bytes;
  return 0;
}

int ensure_workspace(sdns_plan* p) {
  if (p->ws) return 0;
  const size_t need = p->szA + p->szB + 3 * p->szS + p->szPart;
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
  p->cur_rec = p->timing && ((p->timing_mask >> id) & 1);
  if (!p->cur_rec) return;
  p->cur_a = pool_event(p);
  p->cur_id = id;
  (void)hipEventRecord(p->cur_a, p->stream);
}
void tstop(sdns_plan* p) {
  if (!p->cur_rec) return;
  p->cur_rec = false;
  hipEvent_t b = pool_event(p);
  (void)hipEventRecord(b, p->stream);
  p->rec.push_back({p->cur_id, p->cur_a, b});
}

// ---------------------------------------------------------------- transforms
// With trunc, each pass reads only the plan's kept input modes of its axis (the
// others count as zero): the tensor-product truncation of a 2/3-rule
// dealiased backward transform, inside the transform.
template <typename Rt> int plain_backward(sdns_plan* p, const void* uh, void* up, int nc, bool trunc = false) {
  const int N0 = p->N[0], N1 = p->N[1], Nh = p->Nh, N2 = p->N[2];
  const long plane = (long)N1 * Nh, fs = (long)N0 * plane;
  StridedArgs<Rt> a{};
  // x inverse: columns (iy, kz) contiguous
  a.in = uh; a.out = p->bufA; a.tw = p->tw[N0];
  a.Q = (int)plane; a.cdiv = (int)plane; a.cvalid = (int)plane; a.in_hi = 0; a.out_hi = 0;
  a.in_row = plane; a.out_row = plane; a.in_field = fs; a.out_field = fs;
  a.rout = rmap(N0 / 2, N0 / 2 - 1, N0);
  a.rin = trunc ? rmap(p->in[0][0], p->in[0][1], N0) : a.rout;
  a.scale = Rt(1);
  if (launch_strided<Rt>(N0, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", N0);
  // y inverse: columns (x, kz)
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw[N1];
  a.Q = N0 * Nh; a.cdiv = Nh; a.cvalid = Nh; a.in_hi = plane; a.out_hi = plane; a.in_row = Nh; a.out_row = Nh;
  a.rout = rmap(N1 / 2, N1 / 2 - 1, N1);
  a.rin = trunc ? rmap(p->in[1][0], p->in[1][1], N1) : a.rout;
  if (launch_strided<Rt>(N1, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", N1);
  // z C2R
  ZArgs<Rt> z{};
  z.in = p->bufB; z.out = up; z.tw = p->tw[N2 / 2]; z.rtw = p->rtw[N2];
  z.L = N0 * N1; z.nk = trunc ? p->in[2][1] + 1 : Nh; z.in_line = Nh; z.in_field = fs; z.out_line = N2 / 2;
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

// 3/2-rule transforms between the N-mode spectral array and the padded
// physical grid M = 3N/2 (shenfun's padded VTp.backward / Tp.forward, as
// NS.py:29-32 and MHD.py:30-32 build them): the backward passes read the
// kept input modes (|k| < N/2 on every axis, the Nyquist modes dropped) into
// bins of length-M lines whose other bins are zero; the forward passes keep
// the N-mode band of each padded line (|k| <= N/2 on the R2C axis, -N/2..N/2-1
// on the others) and scale by 1/(M0 M1 M2).  Intermediates: x done
// (nc, M0, N1, Nh) in bufA, y done (nc, M0, M1, Nh) in bufB.
template <typename Rt> int padded_backward(sdns_plan* p, const void* uh, void* up, int nc) {
  const int N0 = p->N[0], N1 = p->N[1], Nh = p->Nh;
  const int M0 = p->M[0], M1 = p->M[1], M2 = p->M[2];
  const long plane = (long)N1 * Nh;
  StridedArgs<Rt> a{};
  a.in = uh; a.out = p->bufA; a.tw = p->tw.at(M0);
  a.Q = (int)plane; a.cdiv = (int)plane; a.cvalid = (int)plane; a.in_hi = 0; a.out_hi = 0;
  a.in_row = plane; a.out_row = plane; a.in_field = (long)N0 * plane; a.out_field = (long)M0 * plane;
  a.rin = rmap(p->in[0][0], p->in[0][1], N0);
  a.rout = rmap(M0 / 2, M0 / 2 - 1, M0);
  a.scale = Rt(1);
  if (launch_strided<Rt>(M0, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", M0);
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw.at(M1);
  a.Q = M0 * Nh; a.cdiv = Nh; a.cvalid = Nh; a.in_hi = plane; a.out_hi = (long)M1 * Nh; a.in_row = Nh;
  a.out_row = Nh; a.in_field = (long)M0 * plane; a.out_field = (long)M0 * M1 * Nh;
  a.rin = rmap(p->in[1][0], p->in[1][1], N1);
  a.rout = rmap(M1 / 2, M1 / 2 - 1, M1);
  if (launch_strided<Rt>(M1, p->var[5], +1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  ZArgs<Rt> z{};
  z.in = p->bufB; z.out = up; z.tw = p->tw.at(M2 / 2); z.rtw = p->rtw.at(M2);
  z.L = M0 * M1; z.nk = p->in[2][1] + 1; z.in_line = Nh; z.in_field = (long)M0 * M1 * Nh; z.out_line = M2 / 2;
  z.out_field = (long)M0 * M1 * (M2 / 2); z.scale = Rt(1);
  if (launch_z<Rt>(M2 / 2, p->var[5], +1, nc, z, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "z length %d", M2);
  return 0;
}

template <typename Rt> int padded_forward(sdns_plan* p, const void* up, void* uh, int nc) {
  const int N0 = p->N[0], N1 = p->N[1], Nh = p->Nh;
  const int M0 = p->M[0], M1 = p->M[1], M2 = p->M[2];
  const long plane = (long)N1 * Nh;
  ZArgs<Rt> z{};
  z.in = up; z.out = p->bufA; z.tw = p->tw.at(M2 / 2); z.rtw = p->rtw.at(M2);
  z.L = M0 * M1; z.nk = p->out[2][1] + 1; z.in_line = M2 / 2; z.in_field = (long)M0 * M1 * (M2 / 2);
  z.out_line = Nh; z.out_field = (long)M0 * M1 * Nh; z.scale = Rt(1);
  if (launch_z<Rt>(M2 / 2, p->var[5], -1, nc, z, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "z length %d", M2);
  StridedArgs<Rt> a{};
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw.at(M1);
  a.Q = M0 * Nh; a.cdiv = Nh; a.cvalid = Nh; a.in_hi = (long)M1 * Nh; a.out_hi = plane; a.in_row = Nh;
  a.out_row = Nh; a.in_field = (long)M0 * M1 * Nh; a.out_field = (long)M0 * plane;
  a.rin = rmap(M1 / 2, M1 / 2 - 1, M1);
  a.rout = rmap(p->out[1][0], p->out[1][1], N1);
  a.scale = Rt(1);
  if (launch_strided<Rt>(M1, p->var[5], -1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  a.in = p->bufB; a.out = uh; a.tw = p->tw.at(M0);
  a.Q = (int)plane; a.cdiv = (int)plane; a.cvalid = (int)plane; a.in_hi = 0; a.out_hi = 0; a.in_row = plane;
  a.out_row = plane; a.in_field = (long)M0 * plane; a.out_field = (long)N0 * plane;
  a.rin = rmap(M0 / 2, M0 / 2 - 1, M0);
  a.rout = rmap(p->out[0][0], p->out[0][1], N0);
  a.scale = (Rt)(1.0 / ((double)M0 * M1 * M2));
  if (launch_strided<Rt>(M0, p->var[5], -1, nc, a, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", M0);
  return 0;
}

// Argument blocks of P1..P4 of one RHS evaluation from the state u (device)
// into W4 (= bufB); shared by the single-box path and the batched ensembles.
template <typename Rt> P1Args<Rt> p1_args(const sdns_plan* p, const void* u) {
  P1Args<Rt> a{};
  a.u = u; a.w1 = p->bufA; a.tw = p->tw.at(p->M[0]); a.kx = (const Rt*)p->ktab[0];
  a.N0 = p->N[0]; a.N1 = p->N[1]; a.Nh = p->Nh;
  a.x = rmap(p->in[0][0], p->in[0][1], p->N[0]);
  a.y_neg = p->in[1][0]; a.y_pos = p->in[1][1]; a.nkz = p->nkz; a.nkzp = p->nkzp; a.Q = p->nky * p->nkzp;
  return a;
}
template <typename Rt> P2Args<Rt> p2_args(const sdns_plan* p) {
  P2Args<Rt> a{};
  a.w1 = p->bufA; a.w2 = p->bufB; a.tw = p->tw.at(p->M[1]); a.ky = (const Rt*)p->ktab[1];
  a.kz = (const Rt*)p->ktab[2];
  a.N1 = p->N[1]; a.M0 = p->M[0]; a.y = rmap(p->in[1][0], p->in[1][1], p->nky); a.nkz = p->nkz; a.nkzp = p->nkzp;
  a.Q = p->M[0] * p->nkzp;
  return a;
}
template <typename Rt> P3Args<Rt> p3_args(const sdns_plan* p) {
  P3Args<Rt> a{};
  a.w2 = p->bufB; a.w3 = p->bufA; a.tw = p->tw.at(p->n2h); a.rtw = p->rtw.at(p->M[2]);
  a.nkz = p->nkz; a.nkzp = p->nkzp; a.nkzo = p->nkzo; a.nkzop = p->nkzop; a.L = p->M[0] * p->M[1];
  return a;
}
template <typename Rt> StridedArgs<Rt> p4_args(const sdns_plan* p) {
  const int M0 = p->M[0], M1 = p->M[1];
  StridedArgs<Rt> a{};
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw.at(M1);
  a.Q = M0 * p->nkzop; a.cdiv = p->nkzop; a.cvalid = p->nkzo;
  a.in_hi = (long)M1 * p->nkzop; a.in_row = p->nkzop; a.in_field = (long)M0 * M1 * p->nkzop;
  a.out_hi = (long)p->nkyo * p->nkzop; a.out_row = p->nkzop; a.out_field = (long)M0 * p->nkyo * p->nkzop;
  a.rin = rmap(M1 / 2, M1 - M1 / 2 - 1, M1);
  a.rout = rmap(p->out[1][0], p->out[1][1], p->nkyo);
  a.scale = Rt(1);
  return a;
}

// P2 with a given vorticity: the plain y-inverse of the six W1 fields.
template <typename Rt> StridedArgs<Rt> p2_given_args(const sdns_plan* p) {
  const int M0 = p->M[0], M1 = p->M[1];
  StridedArgs<Rt> a{};
  a.in = p->bufA; a.out = p->bufB; a.tw = p->tw.at(M1);
  a.Q = M0 * p->nkzp; a.cdiv = p->nkzp; a.cvalid = p->nkz;
  a.in_hi = (long)p->nky * p->nkzp; a.in_row = p->nkzp; a.in_field = (long)M0 * p->nky * p->nkzp;
  a.out_hi = (long)M1 * p->nkzp; a.out_row = p->nkzp; a.out_field = (long)M0 * M1 * p->nkzp;
  a.rin = rmap(p->in[1][0], p->in[1][1], p->nky);
  a.rout = rmap(M1 / 2, M1 - M1 / 2 - 1, M1);
  a.scale = Rt(1);
  return a;
}

// P1..P4 of one convection evaluation from the state u into W4 (= bufB).  w:
// the vorticity of u when given (the velocity-vorticity solver), otherwise
// null and the passes form curl(u).
template <typename Rt> int conv_passes(sdns_plan* p, const void* u, const void* w = nullptr) {
  const int M0 = p->M[0], M1 = p->M[1];
  P1Args<Rt> a1 = p1_args<Rt>(p, u);
  a1.w = w;
  tstart(p, 0);
  if (launch_p1<Rt>(M0, p->var[0], a1, p->stream)) return fail(SDNS_ERR_UNSUPPORTED, "x length %d", M0);
  tstop(p);
  tstart(p, 1);
  if (w) {
    if (launch_strided<Rt>(M1, p->var[1], +1, 6, p2_given_args<Rt>(p), p->stream))
      return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  } else if (launch_p2<Rt>(M1, p->var[1], p2_args<Rt>(p), p->stream)) {
    return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  }
  tstop(p);
  tstart(p, 2);
  if (launch_p3<Rt>(p->n2h, p->var[2], p3_args<Rt>(p), p->stream))
    return fail(SDNS_ERR_UNSUPPORTED, "z length %d", p->M[2]);
  tstop(p);
  tstart(p, 3);
  if (launch_strided<Rt>(M1, p->var[3], -1, 3, p4_args<Rt>(p), p->stream))
    return fail(SDNS_ERR_UNSUPPORTED, "y length %d", M1);
  tstop(p);
  return 0;
}

template <typename Rt> P5Args<Rt> p5_args(const sdns_plan* p, double nu) {
  P5Args<Rt> a{};
  a.w4 = p->bufB; a.tw = p->tw.at(p->M[0]);
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

template <typename Rt> int convection(sdns_plan* p, const void* u, void* rhs, const void* w = nullptr) {
  int rc;
  if ((rc = conv_passes<Rt>(p, u, w))) return rc;
  P5Args<Rt> a = p5_args<Rt>(p, 0.0);
  a.A = const_cast<void*>(u); a.rhs = rhs; a.conv_only = 1;
  if (launch_p5<Rt>(p->M[0], p->var[4], P5_RHS, a, p->stream, nullptr))
    return fail(SDNS_ERR_UNSUPPORTED, "x length %d", p->M[0]);
  return 0;
}

template <typename Rt> EwArgs<Rt> ew_args(sdns_plan* p);

// Network forward on the closure's packed input x0 -> y (FP32, 6 of 8 channels).
// keep_pre: store the pre-activations z1, z2 (the backward pass reads them;
// inference does not).
int closure_net_forward(sdns_closure* c, hipStream_t st, bool keep_pre) {
  ConvArgs a{};
  a.n0 = c->n[0]; a.n1 = c->n[1]; a.n2 = c->n[2];
  a.x = c->x0; a.w = c->w.w1f; a.bias = c->w.b1; a.z = keep_pre ? c->z1 : nullptr; a.a = c->a1;
  if (launch_conv(8, 32, CONV_EPI_SILU, a, st)) return fail(SDNS_ERR_UNSUPPORTED, "closure grid");
  a.x = c->a1; a.w = c->w.w2f; a.bias = c->w.b2; a.z = keep_pre ? c->z2 : nullptr; a.a = c->a2;
  if (launch_conv(32, 32, CONV_EPI_SILU, a, st)) return fail(SDNS_ERR_HIP, "closure layer 2");
  a.x = c->a2; a.w = c->w.w3f; a.bias = c->w.b3; a.z = nullptr; a.a = nullptr; a.y = c->y; a.ystride = 8; a.ncout = 6;
  if (launch_conv(32, 16, CONV_EPI_LINEAR, a, st)) return fail(SDNS_ERR_HIP, "closure layer 3");
  return 0;
}

// Projected, dealiased SGS force -i k_j tau_ij of the spectral state u (single
// precision plan on the closure's grid) into c->src.
int closure_source(sdns_closure* c, sdns_plan* p, const void* u, float in_scale, float out_scale) {
  int rc;
  if ((rc = plain_backward<float>(p, u, c->phys, 3))) return rc;
  launch_pack_input(c->phys, in_scale, c->x0, c->P, p->stream);
  if ((rc = closure_net_forward(c, p->stream, false))) return rc;
  launch_unpack_output(c->y, out_scale, c->tauphys, c->P, p->stream);
  const size_t fbytes = (size_t)p->N[0] * p->N[1] * p->Nh * p->esz;
  if ((rc = plain_forward<float>(p, c->tauphys, c->tauhat, 3))) return rc;
  if ((rc = plain_forward<float>(p, c->tauphys + 3 * c->P, c->tauhat + 3 * fbytes, 3))) return rc;
  EwArgs<float> e = ew_args<float>(p);
  e.a = c->tauhat; e.out = c->src;
  if (launch_elementwise<float>(EW_SGS_FORCE, e, p->stream)) return fail(SDNS_ERR_HIP, "sgs force");
  return 0;
}

// With a closure: its Source is evaluated from the step's initial state and
// held through the four stages (as the reference's Source, set between steps),
// or, with per_stage, from every stage state.
template <typename Rt> int rk4_steps(sdns_plan* p, void* u, void* ulo, const double* dts, int nsteps, double nu,
                                      double* diag, void* ph, const void* src = nullptr, sdns_closure* cl = nullptr,
                                      float in_scale = 1.f, float out_scale = 1.f, int per_stage = 0) {
  // a, b and dt in the working precision (integrators.py:185-186, config.py:123-127)
  const Rt a[4] = {Rt(1. / 6.), Rt(1. / 3.), Rt(1. / 3.), Rt(1. / 6.)};
  const Rt b[3] = {Rt(0.5), Rt(0.5), Rt(1.)};
  const Rt nur = (Rt)nu;
  for (int s = 0; s < nsteps; ++s) {
    const Rt dt = (Rt)dts[s];
    for (int rk = 0; rk < 4; ++rk) {
      int rc;
      const void* stage_in = rk == 0 ? u : (const void*)p->stB;
      if (cl && (per_stage || rk == 0) && (rc = closure_source(cl, p, stage_in, in_scale, out_scale))) return rc;
      if ((rc = conv_passes<Rt>(p, stage_in))) return rc;
      P5Args<Rt> pa = p5_args<Rt>(p, nu);
      pa.A = u; pa.B = p->stB; pa.Cacc = p->stC;
      if (ulo) { pa.Alo = ulo; pa.Clo = p->stClo; }
      pa.source = cl ? (const void*)cl->src : src;
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
  for (int i = 0; i < 3; ++i) a.keep[i] = std::max(p->in[i][0], p->in[i][1]);
  return a;
}

template <typename Rt> int lincomb(sdns_plan* p, void* out, int nterms, const void* const* x, const double* c) {
  LinArgs<Rt> a{};
  a.out = out;
  a.nterms = nterms;
  a.n = 3L * p->N[0] * p->N[1] * p->Nh;
  for (int j = 0; j < nterms; ++j) {
    a.x[j] = x[j];
    a.c[j] = (Rt)c[j];
  }
  if (launch_lincomb<Rt>(a, p->stream)) return fail(SDNS_ERR_INVALID, "nterms %d", nterms);
  return 0;
}

template <typename Rt> int error_norm(sdns_plan* p, const void* err, const void* u0, const void* u1, double atol,
                                      double rtol, double* out) {
  const int nblk = 1024;  // 3 * nblk partials fit the partials buffer (>= 8192 doubles)
  launch_errnorm<Rt>(err, u0, u1, (Rt)atol, (Rt)rtol, (long)p->N[0] * p->N[1] * p->Nh, p->part, nblk, p->stream);
  if (launch_sum_cols(p->part, nblk, 3, out, p->stream)) return fail(SDNS_ERR_HIP, "sum columns");
  return 0;
}

template <typename Rt> int energy(sdns_plan* p, const void* u, double nu, double* out) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  const int nblk = 1024;
  launch_energy<Rt>(a, p->part, nblk, p->stream);
  if (launch_sum_partials(p->part, nblk, out, 0.5, (double)(Rt)nu, p->stream)) return fail(SDNS_ERR_HIP, "sum");
  return 0;
}

template <typename Rt> int force(sdns_plan* p, void* u, void* ulo, double target, double kf, double* sums) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  a.out = u;
  a.b = ulo;
  a.aux = ulo;
  const Rt kf2 = (Rt)(kf * kf);
  double* s = sums ? sums : (double*)((char*)p->part + p->szPart - 4096);
  const int nblk = 1024;
  launch_force_reduce<Rt>(a, kf2, p->part, nblk, p->stream);
  if (launch_sum_partials(p->part, nblk, s, 1.0, 1.0, p->stream)) return fail(SDNS_ERR_HIP, "sum");
  launch_force_scale<Rt>(a, kf2, s, target, p->stream);
  return 0;
}

// Shell sums through `part` (room for nblk * 2 * nb doubles): every block
// writes its whole row of partials, so part needs no clearing.
template <typename Rt> int spectrum(sdns_plan* p, const void* u, int nb, double* part, int nblk, double* out) {
  EwArgs<Rt> a = ew_args<Rt>(p);
  a.a = u;
  launch_spectrum<Rt>(a, nb, part, nblk, p->stream);
  if (launch_sum_rows(part, nblk, 2 * nb, out, p->stream)) return fail(SDNS_ERR_HIP, "sum rows");
  return 0;
}
template <typename Rt> int spectrum(sdns_plan* p, const void* u, int nb, double* out) {
  const size_t avail = p->szPart - 4096;
  int nblk = (int)std::min<size_t>(256, avail / (2 * (size_t)nb * sizeof(double)));
  if (nblk < 1) return fail(SDNS_ERR_INVALID, "nbins %d too large", nb);
  return spectrum<Rt>(p, u, nb, p->part, nblk, out);
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
// Column sums of part[nblk][width] with one block per column: strided per-thread
// sums, then a fixed-order tree in LDS (deterministic, no serial chain over nblk).
__global__ void __launch_bounds__(256) k_sum_cols(const double* __restrict__ part, int nblk, int width,
                                                  double* __restrict__ out) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += 256) s += part[(long)b * width + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}
int launch_sum_cols(const double* part, int nblk, int width, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_cols, dim3(width), dim3(256), 0, s, part, nblk, width, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
int launch_sum_rows(const double* part, int nblk, int width, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_rows, dim3((width + 255) / 256), dim3(256), 0, s, part, nblk, width, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
// Per-box (energy, dissipation) of a batched ensemble: block b sums box b's
// P5 partials part[off[b] .. off[b] + 2*cnt[b]) in a fixed order.
__global__ void __launch_bounds__(256) k_batch_sum(const double* __restrict__ part, const int* __restrict__ off,
                                                   const int* __restrict__ cnt, const double* __restrict__ nu,
                                                   double* __restrict__ out) {
  __shared__ double s[2][256];
  const int b = blockIdx.x;
  const double* pb = part + off[b];
  double a0 = 0, a1 = 0;
  for (int i = threadIdx.x; i < cnt[b]; i += 256) { a0 += pb[2 * i]; a1 += pb[2 * i + 1]; }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) { s[0][threadIdx.x] += s[0][threadIdx.x + w]; s[1][threadIdx.x] += s[1][threadIdx.x + w]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[2 * b] = 0.5 * s[0][0]; out[2 * b + 1] = nu[b] * s[1][0]; }
}
}  // namespace sdns

// ---------------------------------------------------------------- batched ensembles
// Every box keeps its own plan (tables, work buffers, RK stage buffers); the
// batch adds, in one device allocation, the per-box argument blocks of each
// pass, the line length of each box per pass and the block work lists.  One
// RK4 step is 20 batched pass launches plus one reduction, enqueued on the
// batch's own stream and captured once as a hipGraph.
struct sdns_batch {
  int nbox = 0, precision = 0;
  std::vector<sdns_plan*> box;
  hipStream_t user = nullptr, own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  char* dmem = nullptr;
  // device arrays inside dmem
  void* d_p1[2] = {nullptr, nullptr};  // P1 args: stage 0 reads u_hat, stages 1-3 the stage buffer
  void *d_p2 = nullptr, *d_p3 = nullptr, *d_p4 = nullptr;
  void* d_p5[4] = {nullptr, nullptr, nullptr, nullptr};  // P5 args of RK stages 0..3
  int* d_len[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  BatchWork* d_work[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int nwork[5] = {0, 0, 0, 0, 0}, lmax[5] = {0, 0, 0, 0, 0};
  double *d_part = nullptr, *d_diag = nullptr, *d_nu = nullptr;
  int *d_poff = nullptr, *d_pcnt = nullptr;
  size_t args_bytes = 0;  // bytes of one per-box argument array (largest struct * nbox)
  std::vector<int> poff, pcnt;
  std::vector<void*> u;
  std::vector<double> nu, dt;
  bool have_state = false, have_params = false, dirty = true, use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::map<std::pair<int, int>, int> sched;  // (pass, line length) -> schedule variant
};

namespace {

// Work list of a batched pass: items (box, tile) in box order, split into 8
// contiguous chunks of about equal work (weight per tile of each box), chunk
// x executed by the blocks i with i % 8 == x (the XCD that dispatches them),
// so that the tiles of one box share an L2.  Padding items have box = -1.
std::vector<BatchWork> batch_work(const std::vector<int>& ntile, const std::vector<int>& weight) {
  std::vector<BatchWork> items;
  long total = 0;
  for (size_t b = 0; b < ntile.size(); ++b) {
    for (int t = 0; t < ntile[b]; ++t) items.push_back({(int)b, t});
    total += (long)ntile[b] * weight[b];
  }
  std::vector<std::vector<BatchWork>> chunk(8);
  long acc = 0;
  for (const BatchWork& w : items) {
    int x = (int)((acc * 8) / std::max(total, 1L));
    if (x > 7) x = 7;
    chunk[x].push_back(w);
    acc += weight[w.box];
  }
  size_t per = 0;
  for (auto& c : chunk) per = std::max(per, c.size());
  std::vector<BatchWork> out(8 * std::max(per, (size_t)1), BatchWork{-1, 0});
  for (int x = 0; x < 8; ++x)
    for (size_t j = 0; j < chunk[x].size(); ++j) out[8 * j + x] = chunk[x][j];
  return out;
}

bool batch_strided_ok(int M) { return batch_E(M) > 0; }
bool batch_half_ok(int n) { return batch_zE(n) > 0; }

// Schedule variant of pass k for line length l: the one set by
// sdns_batch_set_variant, or 0; a variant whose plan equals variant 0's is 0.
int batch_variant(const sdns_batch* b, int k, int l) {
  const auto it = b->sched.find({k, l});
  if (it == b->sched.end() || it->second == 0) return 0;
  return (k == 2 ? batch_zE(l, 1) != batch_zE(l, 0) : batch_E(l, 1) != batch_E(l, 0)) ? 1 : 0;
}

template <typename Rt> int batch_layout(sdns_batch* b) {
  const int nb = b->nbox;
  std::vector<int> len[5], nt[5], wt[5];
  for (int k = 0; k < 5; ++k) b->lmax[k] = 0;
  for (int i = 0; i < nb; ++i) {
    const sdns_plan* p = b->box[i];
    const int M0 = p->M[0], M1 = p->M[1], n = p->n2h;
    const int l[5] = {M0, M1, n, M1, M0};
    const int q[5] = {p->nky * p->nkzp, M0 * p->nkzp, M0 * M1, M0 * p->nkzop, p->N[1] * p->Nh};
    for (int k = 0; k < 5; ++k) {
      const int v = batch_variant(b, k, l[k]);
      const int per = k == 2 ? batch_LPB<Rt>(l[k], v) : batch_T<Rt>(l[k], v);
      len[k].push_back(l[k] + BATCH_VSTRIDE * v);
      nt[k].push_back((q[k] + per - 1) / per);
      wt[k].push_back(k == 2 ? batch_zE(l[k], v) : batch_E(l[k], v));
      b->lmax[k] = std::max(b->lmax[k], l[k]);
    }
  }
  std::vector<BatchWork> work[5];
  size_t nitems = 0;
  for (int k = 0; k < 5; ++k) {
    work[k] = batch_work(nt[k], wt[k]);
    b->nwork[k] = (int)work[k].size() / 8 * 8;
    nitems += work[k].size();
  }
  b->poff.assign(nb, 0);
  b->pcnt = nt[4];
  int part_total = 0;
  for (int i = 0; i < nb; ++i) { b->poff[i] = part_total; part_total += 2 * nt[4][i]; }
  // one allocation: 9 argument arrays, lengths, work lists, partials, diag, nu, offsets
  size_t argsz = std::max(std::max(sizeof(P1Args<Rt>), sizeof(P2Args<Rt>)),
                          std::max(std::max(sizeof(P3Args<Rt>), sizeof(StridedArgs<Rt>)), sizeof(P5Args<Rt>)));
  b->args_bytes = align256(argsz * nb);
  const size_t bytes = 9 * b->args_bytes + align256(5 * nb * sizeof(int)) + align256(nitems * sizeof(BatchWork)) +
                       align256(part_total * sizeof(double)) + 2 * align256(2 * nb * sizeof(double)) +
                       2 * align256(nb * sizeof(int));
  if (hipMalloc((void**)&b->dmem, bytes) != hipSuccess) return fail(SDNS_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
  char* c = b->dmem;
  void** slots[9] = {&b->d_p1[0], &b->d_p1[1], &b->d_p2, &b->d_p3, &b->d_p4,
                     &b->d_p5[0], &b->d_p5[1], &b->d_p5[2], &b->d_p5[3]};
  for (auto sl : slots) { *sl = c; c += b->args_bytes; }
  int* lens = (int*)c;
  c += align256(5 * nb * sizeof(int));
  BatchWork* wl = (BatchWork*)c;
  c += align256(nitems * sizeof(BatchWork));
  b->d_part = (double*)c; c += align256(part_total * sizeof(double));
  b->d_diag = (double*)c; c += align256(2 * nb * sizeof(double));
  b->d_nu = (double*)c; c += align256(2 * nb * sizeof(double));
  b->d_poff = (int*)c; c += align256(nb * sizeof(int));
  b->d_pcnt = (int*)c;
  std::vector<int> hl;
  for (int k = 0; k < 5; ++k) hl.insert(hl.end(), len[k].begin(), len[k].end());
  HIP_TRY(hipMemcpyAsync(lens, hl.data(), hl.size() * sizeof(int), hipMemcpyHostToDevice, b->own));
  size_t o = 0;
  for (int k = 0; k < 5; ++k) {
    b->d_len[k] = lens + k * nb;
    b->d_work[k] = wl + o;
    HIP_TRY(hipMemcpyAsync(wl + o, work[k].data(), work[k].size() * sizeof(BatchWork), hipMemcpyHostToDevice,
                           b->own));
    o += work[k].size();
  }
  HIP_TRY(hipMemcpyAsync(b->d_poff, b->poff.data(), nb * sizeof(int), hipMemcpyHostToDevice, b->own));
  HIP_TRY(hipMemcpyAsync(b->d_pcnt, b->pcnt.data(), nb * sizeof(int), hipMemcpyHostToDevice, b->own));
  HIP_TRY(hipStreamSynchronize(b->own));
  return 0;
}

// Per-box argument arrays for the current state pointers, nu and dt.
template <typename Rt> int batch_upload(sdns_batch* b) {
  const int nb = b->nbox;
  const Rt a[4] = {Rt(1. / 6.), Rt(1. / 3.), Rt(1. / 3.), Rt(1. / 6.)};
  const Rt bb[3] = {Rt(0.5), Rt(0.5), Rt(1.)};
  std::vector<P1Args<Rt>> h1[2];
  std::vector<P2Args<Rt>> h2;
  std::vector<P3Args<Rt>> h3;
  std::vector<StridedArgs<Rt>> h4;
  std::vector<P5Args<Rt>> h5[4];
  for (int i = 0; i < nb; ++i) {
    const sdns_plan* p = b->box[i];
    h1[0].push_back(p1_args<Rt>(p, b->u[i]));
    h1[1].push_back(p1_args<Rt>(p, p->stB));
    h2.push_back(p2_args<Rt>(p));
    h3.push_back(p3_args<Rt>(p));
    h4.push_back(p4_args<Rt>(p));
    {  // the batched y-forward streams W3 / W4 when its tiles cover whole lines
      StridedArgs<Rt>& q = h4.back();
      const int T = batch_T<Rt>(p->M[1], batch_variant(b, 3, p->M[1]));
      q.stream_in = tiles_whole_lines<Rt>(q.in, T, q.cdiv, q.in_hi, q.in_row, q.in_field);
      q.stream_out = tiles_whole_lines<Rt>(q.out, T, q.cdiv, q.out_hi, q.out_row, q.out_field);
    }
    const Rt dt = (Rt)b->dt[i];
    for (int rk = 0; rk < 4; ++rk) {  // stages 1 and 2 share a but not b
      P5Args<Rt> pa = p5_args<Rt>(p, b->nu[i]);
      pa.A = b->u[i]; pa.B = p->stB; pa.Cacc = p->stC;
      pa.adt = a[rk] * dt;
      pa.bdt = rk < 3 ? bb[rk] * dt : Rt(0);
      pa.partials = rk == 3 ? b->d_part + b->poff[i] : nullptr;
      h5[rk].push_back(pa);
    }
  }
  // copies on the batch stream, which is then drained: the host vectors stay
  // alive until the data are on the device and no replay can see a partial update
  hipStream_t s = b->own;
  HIP_TRY(hipMemcpyAsync(b->d_p1[0], h1[0].data(), nb * sizeof(P1Args<Rt>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_p1[1], h1[1].data(), nb * sizeof(P1Args<Rt>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_p2, h2.data(), nb * sizeof(P2Args<Rt>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_p3, h3.data(), nb * sizeof(P3Args<Rt>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_p4, h4.data(), nb * sizeof(StridedArgs<Rt>), hipMemcpyHostToDevice, s));
  for (int m = 0; m < 4; ++m)
    HIP_TRY(hipMemcpyAsync(b->d_p5[m], h5[m].data(), nb * sizeof(P5Args<Rt>), hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(b->d_nu, b->nu.data(), nb * sizeof(double), hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

// One RK4 step of every box on stream s.
template <typename Rt> int batch_step(sdns_batch* b, hipStream_t s) {
  for (int rk = 0; rk < 4; ++rk) {
    launch_batch_p1<Rt>(b->lmax[0], (const P1Args<Rt>*)b->d_p1[rk == 0 ? 0 : 1], b->d_len[0], b->d_work[0],
                        b->nwork[0], s);
    launch_batch_p2<Rt>(b->lmax[1], (const P2Args<Rt>*)b->d_p2, b->d_len[1], b->d_work[1], b->nwork[1], s);
    launch_batch_p3<Rt>(b->lmax[2], (const P3Args<Rt>*)b->d_p3, b->d_len[2], b->d_work[2], b->nwork[2], s);
    launch_batch_fwd<Rt>(b->lmax[3], (const StridedArgs<Rt>*)b->d_p4, b->d_len[3], b->d_work[3], b->nwork[3], 3, s);
    const int mode = rk == 0 ? P5_RK0 : (rk < 3 ? P5_RK12 : P5_RK3);
    launch_batch_p5<Rt>(b->lmax[4], mode, (const P5Args<Rt>*)b->d_p5[rk], b->d_len[4], b->d_work[4], b->nwork[4], s);
  }
  hipLaunchKernelGGL(k_batch_sum, dim3(b->nbox), dim3(256), 0, s, (const double*)b->d_part, (const int*)b->d_poff,
                     (const int*)b->d_pcnt, (const double*)b->d_nu, b->d_diag);
  return check_launch();
}

int batch_step_any(sdns_batch* b, hipStream_t s) {
  return b->precision == SDNS_DOUBLE ? batch_step<double>(b, s) : batch_step<float>(b, s);
}

void batch_drop_graph(sdns_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int batch_capture(sdns_batch* b) {
  HIP_TRY(hipStreamBeginCapture(b->own, hipStreamCaptureModeThreadLocal));
  const int rc = batch_step_any(b, b->own);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->own, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(SDNS_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return 0;
}

}  // namespace

// ---------------------------------------------------------------- allocation-free spectrum
// Scratch of sdns_spectrum_noalloc: the per-block shell partials, 2^19 doubles
// (4 MiB) on every GPU the library runs on, held until the process ends.  In
// host code `scratch` names a host-side stand-in, never the device memory:
// kernels get the address that hipGetSymbolAddress returns for the current
// device (scratch_on_device).
namespace sdns {
constexpr int kScratchDoubles = 1 << 19;
constexpr int kNoallocMaxBins = 8192;  // 2 * nbins doubles of LDS per block: 128 KiB
__device__ __align__(256) double scratch[kScratchDoubles];
}  // namespace sdns

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_on_device(double** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SDNS_ERR_UNSUPPORTED, "device %d (at most %d)", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(sdns::scratch));
    if (e != hipSuccess || !ptr)
      return fail(SDNS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_dev[dev] = (double*)ptr;
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

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
  const int maxv = pass == 2 ? 5 : (pass == 4 ? 3 : 1);
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
  return p ? p->szA + p->szB + 3 * p->szS + p->szPart : 0;
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

int sdns_backward_dealiased(sdns_plan* p, const void* uh, void* up, int nc) {
  if (!p || !up || !uh || nc < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  if (nc > 3) return fail(SDNS_ERR_INVALID, "ncomp %d > 3", nc);
  int rc = ensure_workspace(p);
  if (rc) return rc;
  if (p->dealias == SDNS_DEALIAS_32)
    rc = p->precision == SDNS_DOUBLE ? padded_backward<double>(p, uh, up, nc) : padded_backward<float>(p, uh, up, nc);
  else
    rc = p->precision == SDNS_DOUBLE ? plain_backward<double>(p, uh, up, nc, true)
                                     : plain_backward<float>(p, uh, up, nc, true);
  return rc ? rc : check_launch();
}

int sdns_forward_dealiased(sdns_plan* p, const void* up, void* uh, int nc) {
  if (!p || !up || !uh || nc < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  if (nc > 3) return fail(SDNS_ERR_INVALID, "ncomp %d > 3", nc);
  int rc = ensure_workspace(p);
  if (rc) return rc;
  if (p->dealias == SDNS_DEALIAS_32)
    rc = p->precision == SDNS_DOUBLE ? padded_forward<double>(p, up, uh, nc) : padded_forward<float>(p, up, uh, nc);
  else
    rc = p->precision == SDNS_DOUBLE ? plain_forward<double>(p, up, uh, nc) : plain_forward<float>(p, up, uh, nc);
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
int sdns_cross2_k_over_k2(sdns_plan* p, void* c, const void* a) {
  if (!a) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_CROSS2_KOK, c, a, nullptr, nullptr, 0, -1);
}
int sdns_add_linear(sdns_plan* p, void* rhs, const void* w_hat, double nu, const void* source) {
  if (!w_hat) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_ADD_LINEAR, rhs, w_hat, source, nullptr, nu, -1);
}
int sdns_vv_rhs_tail(sdns_plan* p, void* rhs, const void* v_hat, const void* w_hat, double nu, const void* source) {
  if (!v_hat || !w_hat) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_VV_TAIL, rhs, v_hat, w_hat, const_cast<void*>(source), nu, -1);
}
int sdns_dealias_truncate(sdns_plan* p, void* out, const void* a) {
  if (!a) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_TRUNCATE, out, a, nullptr, nullptr, 0, -1);
}
int sdns_set_elsasser(sdns_plan* p, void* c_hat, const void* zz_hat) {
  if (!zz_hat) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_SET_ELSASSER, c_hat, zz_hat, nullptr, nullptr, 0, -1);
}
int sdns_gradient_truncated(sdns_plan* p, void* out, const void* a) {
  if (!a) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_GRAD_TRUNC, out, a, nullptr, nullptr, 0, -1);
}
int sdns_dot3(sdns_plan* p, void* out, const void* u, const void* g, size_t npoints) {
  if (!u || !g) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_DOT3, out, u, g, nullptr, 0, (long)npoints);
}
int sdns_uu_products(sdns_plan* p, void* uu, const void* u, size_t npoints) {
  if (!u) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_UU_PRODUCTS, uu, u, nullptr, nullptr, 0, (long)npoints);
}
int sdns_div_uu(sdns_plan* p, void* out, const void* uu_hat, double scale) {
  if (!uu_hat) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_DIV_UU, out, uu_hat, nullptr, nullptr, scale, -1);
}
int sdns_elsasser_products(sdns_plan* p, void* zz, const void* ub, size_t npoints) {
  if (!ub) return fail(SDNS_ERR_INVALID, "null input");
  return ew(p, EW_ELSASSER_PRODUCTS, zz, ub, nullptr, nullptr, 0, (long)npoints);
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

int sdns_convection_uw(sdns_plan* p, const void* u, const void* w, void* out) {
  if (!p || !u || !w || !out) return fail(SDNS_ERR_INVALID, "bad argument");
  if (out == u || out == w) return fail(SDNS_ERR_INVALID, "out may not alias an input");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? convection<double>(p, u, out, w) : convection<float>(p, u, out, w);
  return rc ? rc : check_launch();
}

int sdns_rk4_steps(sdns_plan* p, void* u, void* u_lo, const double* dts, int nsteps, double nu, double* diag,
                   void* ph) {
  return sdns_rk4_steps_source(p, u, u_lo, dts, nsteps, nu, nullptr, diag, ph);
}

int sdns_rk4_steps_source(sdns_plan* p, void* u, void* u_lo, const double* dts, int nsteps, double nu,
                          const void* source, double* diag, void* ph) {
  if (!p || !u || (!dts && nsteps > 0) || nsteps < 0) return fail(SDNS_ERR_INVALID, "bad argument");
  if (source && (source == u || (u_lo && source == u_lo)))
    return fail(SDNS_ERR_INVALID, "the source may not alias the state");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? rk4_steps<double>(p, u, nullptr, dts, nsteps, nu, diag, ph, source)
                                   : rk4_steps<float>(p, u, u_lo, dts, nsteps, nu, diag, ph, source);
  return rc ? rc : check_launch();
}

int sdns_force(sdns_plan* p, void* u, double target, double kf, double* sums) {
  if (!p || !u) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? force<double>(p, u, nullptr, target, kf, sums)
                                    : force<float>(p, u, nullptr, target, kf, sums);
  return rc ? rc : check_launch();
}

int sdns_force_compensated(sdns_plan* p, void* u, void* u_lo, double target, double kf, double* sums) {
  if (!p || !u || !u_lo) return fail(SDNS_ERR_INVALID, "bad argument");
  if (p->precision != SDNS_SINGLE) return fail(SDNS_ERR_INVALID, "the compensated state is single precision");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = force<float>(p, u, u_lo, target, kf, sums);
  return rc ? rc : check_launch();
}

int sdns_spectrum(sdns_plan* p, const void* u, int nbins, double* out) {
  if (!p || !u || !out || nbins < 1 || nbins > 8192) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? spectrum<double>(p, u, nbins, out) : spectrum<float>(p, u, nbins, out);
  return rc ? rc : check_launch();
}

int sdns_spectrum_noalloc(sdns_plan* p, const void* u, int nbins, double* out) {
  if (!p || !u || !out) return fail(SDNS_ERR_INVALID, "bad argument");
  if (nbins < 1 || nbins > kNoallocMaxBins)
    return fail(SDNS_ERR_INVALID, "nbins %d outside 1..%d (the limit of the no-alloc scratch of %d doubles)", nbins,
                kNoallocMaxBins, kScratchDoubles);
  double* part = nullptr;
  int rc = scratch_on_device(&part);
  if (rc) return rc;
  const int nblk = std::min(256, kScratchDoubles / (2 * nbins));  // >= 32
  rc = p->precision == SDNS_DOUBLE ? spectrum<double>(p, u, nbins, part, nblk, out)
                                   : spectrum<float>(p, u, nbins, part, nblk, out);
  return rc ? rc : check_launch();
}

int sdns_timing_enable(sdns_plan* p, int enable) {
  if (!p) return fail(SDNS_ERR_INVALID, "null plan");
  p->timing = enable != 0;
  return 0;
}

int sdns_timing_passes(sdns_plan* p, int mask) {
  if (!p) return fail(SDNS_ERR_INVALID, "null plan");
  if (mask < 0 || mask > 63) return fail(SDNS_ERR_INVALID, "pass mask %d", mask);
  p->timing_mask = mask;
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

int sdns_lincomb(sdns_plan* p, void* out, int nterms, const void* const* x, const double* c) {
  if (!p || !out || !x || !c || nterms < 1 || nterms > SDNS_LIN_MAX)
    return fail(SDNS_ERR_INVALID, "bad argument (1 <= nterms <= %d)", SDNS_LIN_MAX);
  for (int j = 0; j < nterms; ++j)
    if (!x[j]) return fail(SDNS_ERR_INVALID, "null term %d", j);
  const int rc = p->precision == SDNS_DOUBLE ? lincomb<double>(p, out, nterms, x, c) : lincomb<float>(p, out, nterms, x, c);
  return rc ? rc : check_launch();
}

int sdns_error_norm(sdns_plan* p, const void* err, const void* u0, const void* u1, double atol, double rtol,
                    double* out) {
  if (!p || !err || !u0 || !u1 || !out) return fail(SDNS_ERR_INVALID, "bad argument");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = p->precision == SDNS_DOUBLE ? error_norm<double>(p, err, u0, u1, atol, rtol, out)
                                   : error_norm<float>(p, err, u0, u1, atol, rtol, out);
  return rc ? rc : check_launch();
}

// ---------------------------------------------------------------- batched ensembles
int sdns_batch_destroy(sdns_batch* b) {
  if (!b) return 0;
  if (b->own) (void)hipStreamSynchronize(b->own);
  batch_drop_graph(b);
  for (sdns_plan* p : b->box) sdns_plan_destroy(p);
  if (b->dmem) (void)hipFree(b->dmem);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->own) (void)hipStreamDestroy(b->own);
  delete b;
  return 0;
}

int sdns_batch_create(sdns_batch** out, int nbox, const int32_t* N, const double* L, int precision, int dealias,
                      int mask_nyquist, int dealias_mode, void* hip_stream) {
  if (!out || !N || !L || nbox < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  *out = nullptr;
  if (dealias == SDNS_DEALIAS_NONE)
    return fail(SDNS_ERR_UNSUPPORTED, "batched ensembles run the dealiased (2/3 or 3/2 rule) step");
  sdns_batch* b = new sdns_batch();
  b->nbox = nbox;
  b->precision = precision;
  b->user = (hipStream_t)hip_stream;
  auto bail = [&](int rc) { sdns_batch_destroy(b); return rc; };
  if (hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
    return bail(fail(SDNS_ERR_HIP, "stream / event creation failed"));
  for (int i = 0; i < nbox; ++i) {
    sdns_plan* p = nullptr;
    int rc = sdns_plan_create(&p, N + 3 * i, L + 3 * i, precision, dealias, mask_nyquist, dealias_mode, b->own);
    if (rc) return bail(rc);
    b->box.push_back(p);
    if (!batch_strided_ok(p->M[0]) || !batch_strided_ok(p->M[1]) || !batch_half_ok(p->n2h))
      return bail(fail(SDNS_ERR_UNSUPPORTED, "box %d: no batched kernel for lines %d x %d x %d (x, y in 16..128 "
                       "with a power-of-two thread count per line, z/2 in 8..64)", i, p->M[0], p->M[1], p->M[2]));
    if ((rc = ensure_workspace(p))) return bail(rc);
  }
  const int rc = precision == SDNS_DOUBLE ? batch_layout<double>(b) : batch_layout<float>(b);
  if (rc) return bail(rc);
  b->u.assign(nbox, nullptr);
  b->nu.assign(nbox, 0.0);
  b->dt.assign(nbox, 0.0);
  *out = b;
  return 0;
}

int sdns_batch_set_state(sdns_batch* b, void* const* u_hat) {
  if (!b || !u_hat) return fail(SDNS_ERR_INVALID, "bad argument");
  for (int i = 0; i < b->nbox; ++i)
    if (!u_hat[i]) return fail(SDNS_ERR_INVALID, "u_hat[%d] is NULL", i);
  HIP_TRY(hipStreamSynchronize(b->own));
  for (int i = 0; i < b->nbox; ++i) b->u[i] = u_hat[i];
  b->have_state = true;
  b->dirty = true;
  return 0;
}

int sdns_batch_set_params(sdns_batch* b, const double* nu, const double* dt) {
  if (!b || !nu || !dt) return fail(SDNS_ERR_INVALID, "bad argument");
  for (int i = 0; i < b->nbox; ++i)
    if (!(nu[i] >= 0) || !(dt[i] > 0)) return fail(SDNS_ERR_INVALID, "box %d: nu %g, dt %g", i, nu[i], dt[i]);
  HIP_TRY(hipStreamSynchronize(b->own));
  for (int i = 0; i < b->nbox; ++i) { b->nu[i] = nu[i]; b->dt[i] = dt[i]; }
  b->have_params = true;
  b->dirty = true;
  return 0;
}

int sdns_batch_use_graph(sdns_batch* b, int enable) {
  if (!b) return fail(SDNS_ERR_INVALID, "null batch");
  b->use_graph = enable != 0;
  return 0;
}

int sdns_batch_set_stream(sdns_batch* b, void* hip_stream) {
  if (!b) return fail(SDNS_ERR_INVALID, "null batch");
  b->user = (hipStream_t)hip_stream;
  return 0;
}

int sdns_batch_rk4_steps(sdns_batch* b, int nsteps, double* diag) {
  if (!b || nsteps < 0) return fail(SDNS_ERR_INVALID, "bad argument");
  if (!b->have_state || !b->have_params) return fail(SDNS_ERR_INVALID, "set the state and the parameters first");
  if (b->dirty) {
    HIP_TRY(hipStreamSynchronize(b->own));
    const int rc = b->precision == SDNS_DOUBLE ? batch_upload<double>(b) : batch_upload<float>(b);
    if (rc) return rc;
    b->dirty = false;
  }
  HIP_TRY(hipEventRecord(b->ev_in, b->user));
  HIP_TRY(hipStreamWaitEvent(b->own, b->ev_in, 0));
  if (b->use_graph && nsteps > 0 && !b->exec) {
    const int rc = batch_capture(b);
    if (rc) return rc;
  }
  for (int s = 0; s < nsteps; ++s) {
    if (b->use_graph) {
      HIP_TRY(hipGraphLaunch(b->exec, b->own));
    } else {
      const int rc = batch_step_any(b, b->own);
      if (rc) return rc;
    }
  }
  if (diag && nsteps > 0)
    HIP_TRY(hipMemcpyAsync(diag, b->d_diag, 2 * b->nbox * sizeof(double), hipMemcpyDeviceToDevice, b->own));
  HIP_TRY(hipEventRecord(b->ev_out, b->own));
  HIP_TRY(hipStreamWaitEvent(b->user, b->ev_out, 0));
  return 0;
}

int sdns_batch_set_variant(sdns_batch* b, int pass, int length, int variant) {
  if (!b || pass < 0 || pass > 4 || length < 1 || variant < 0 || variant > 1)
    return fail(SDNS_ERR_INVALID, "pass %d, length %d, variant %d", pass, length, variant);
  b->sched[{pass, length}] = variant;
  // new work lists: drain the batch stream, drop the captured step and lay out again
  HIP_TRY(hipStreamSynchronize(b->own));
  batch_drop_graph(b);
  if (b->dmem) (void)hipFree(b->dmem);
  b->dmem = nullptr;
  const int rc = b->precision == SDNS_DOUBLE ? batch_layout<double>(b) : batch_layout<float>(b);
  if (rc) return rc;
  b->dirty = true;
  return 0;
}

int sdns_batch_box_plan(sdns_batch* b, int i, sdns_plan** plan) {
  if (!b || !plan || i < 0 || i >= b->nbox) return fail(SDNS_ERR_INVALID, "bad argument");
  *plan = b->box[i];
  return 0;
}

// ---------------------------------------------------------------- learned closure
size_t sdns_closure_nparams(void) { return (size_t)CL_NPARAM; }

int sdns_closure_destroy(sdns_closure* c) {
  if (!c) return 0;
  if (c->mem) (void)hipFree(c->mem);
  delete c;
  return 0;
}

int sdns_closure_create(sdns_closure** out, const int32_t N[3], void* hip_stream) {
  if (!out || !N) return fail(SDNS_ERR_INVALID, "null argument");
  *out = nullptr;
  const long P = (long)N[0] * N[1] * N[2];
  if (N[0] < 4 || N[1] < 4 || N[2] % 32 || P % 4096)
    return fail(SDNS_ERR_UNSUPPORTED, "closure grid %dx%dx%d: N2 %% 32 == 0 and N0 N1 N2 %% 4096 == 0 required",
                N[0], N[1], N[2]);
  sdns_closure* c = new sdns_closure();
  for (int i = 0; i < 3; ++i) c->n[i] = N[i];
  c->P = P;
  c->stream = (hipStream_t)hip_stream;
  const long nchunk = P / 4096;
  const size_t fbytes = (size_t)N[0] * N[1] * (N[2] / 2 + 1) * 8;  // complex64 spectral field
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
  const size_t o_x0 = take(P * 8 * 2), o_z1 = take(P * 64), o_a1 = take(P * 64), o_z2 = take(P * 64),
               o_a2 = take(P * 64), o_gyb = take(P * 16), o_gz2 = take(P * 64), o_gz1 = take(P * 64),
               o_y = take(P * 32), o_gy = take(P * 32), o_t = take(P * 32),
               o_wp = take(nchunk * 27 * 32 * 32 * 4), o_bp = take(256 * 32 * 4), o_ph = take(3 * P * 4),
               o_tp = take(6 * P * 4), o_th = take(6 * fbytes), o_src = take(3 * fbytes), o_lp = take(256 * 8),
               o_l = take(8), o_w1 = take(27 * 32 * 8 * 2), o_w2 = take(27 * 32 * 32 * 2),
               o_w3 = take(27 * 16 * 32 * 2), o_w3b = take(27 * 32 * 8 * 2), o_w2b = take(27 * 32 * 32 * 2),
               o_b = take(80 * 4);
  if (hipMalloc((void**)&c->mem, off) != hipSuccess) {
    delete c;
    return fail(SDNS_ERR_NOMEM, "hipMalloc of %zu bytes failed", off);
  }
  char* m = c->mem;
  c->x0 = (__bf16*)(m + o_x0); c->z1 = (__bf16*)(m + o_z1); c->a1 = (__bf16*)(m + o_a1);
  c->z2 = (__bf16*)(m + o_z2); c->a2 = (__bf16*)(m + o_a2); c->gyb = (__bf16*)(m + o_gyb);
  c->gz2 = (__bf16*)(m + o_gz2); c->gz1 = (__bf16*)(m + o_gz1);
  c->y = (float*)(m + o_y); c->gy = (float*)(m + o_gy); c->t = (float*)(m + o_t);
  c->wpart = (float*)(m + o_wp); c->bpart = (float*)(m + o_bp); c->phys = (float*)(m + o_ph);
  c->tauphys = (float*)(m + o_tp); c->tauhat = m + o_th; c->src = m + o_src;
  c->lpart = (double*)(m + o_lp); c->loss = (double*)(m + o_l);
  c->w.w1f = (__bf16*)(m + o_w1); c->w.w2f = (__bf16*)(m + o_w2); c->w.w3f = (__bf16*)(m + o_w3);
  c->w.w3b = (__bf16*)(m + o_w3b); c->w.w2b = (__bf16*)(m + o_w2b);
  c->w.b1 = (float*)(m + o_b); c->w.b2 = c->w.b1 + 32; c->w.b3 = c->w.b2 + 32;
  *out = c;
  return 0;
}

int sdns_closure_set_stream(sdns_closure* c, void* hip_stream) {
  if (!c) return fail(SDNS_ERR_INVALID, "null closure");
  c->stream = (hipStream_t)hip_stream;
  return 0;
}

int sdns_closure_set_params(sdns_closure* c, const float* params) {
  if (!c || !params) return fail(SDNS_ERR_INVALID, "bad argument");
  launch_pack_weights(params, c->w, c->stream);
  c->have_params = true;
  return check_launch();
}

int sdns_closure_forward(sdns_closure* c, const float* u_phys, float in_scale, float out_scale, float* tau_phys) {
  if (!c || !u_phys || !tau_phys) return fail(SDNS_ERR_INVALID, "bad argument");
  if (!c->have_params) return fail(SDNS_ERR_INVALID, "closure parameters not set");
  launch_pack_input(u_phys, in_scale, c->x0, c->P, c->stream);
  int rc = closure_net_forward(c, c->stream, false);
  if (rc) return rc;
  launch_unpack_output(c->y, out_scale, tau_phys, c->P, c->stream);
  return check_launch();
}

int sdns_closure_train_step(sdns_closure* c, const float* u_phys, const float* tau_phys, float in_scale,
                            float out_scale, float* grad, double* loss) {
  if (!c || !u_phys || !tau_phys || !grad) return fail(SDNS_ERR_INVALID, "bad argument");
  if (!c->have_params) return fail(SDNS_ERR_INVALID, "closure parameters not set");
  hipStream_t st = c->stream;
  const long P = c->P;
  const int n0 = c->n[0], n1 = c->n[1], n2 = c->n[2];
  launch_pack_input(u_phys, in_scale, c->x0, P, st);
  int rc = closure_net_forward(c, st, true);
  if (rc) return rc;
  launch_pack_target(tau_phys, 1.0f / out_scale, c->t, P, st);
  launch_loss(c->y, c->t, c->gy, c->gyb, P, c->lpart, c->loss, st);
  // layer 3
  int nch = launch_wgrad(8, 32, c->gyb, c->a2, c->wpart, n0, n1, n2, st);
  if (nch < 0) return fail(SDNS_ERR_UNSUPPORTED, "closure grid");
  launch_wgrad_reduce(c->wpart, nch, 8, 32, 6, 32, grad + CL_OW3, st);
  launch_bgrad_f32(c->gy, 8, 6, P, c->bpart, grad + CL_OB3, st);
  ConvArgs a{};
  a.n0 = n0; a.n1 = n1; a.n2 = n2;
  a.x = c->gyb; a.w = c->w.w3b; a.zprev = c->z2; a.a = c->gz2;
  if (launch_conv(8, 32, CONV_EPI_DSILU, a, st)) return fail(SDNS_ERR_HIP, "closure layer 3 input gradient");
  // layer 2
  nch = launch_wgrad(32, 32, c->gz2, c->a1, c->wpart, n0, n1, n2, st);
  launch_wgrad_reduce(c->wpart, nch, 32, 32, 32, 32, grad + CL_OW2, st);
  launch_bgrad_bf16(c->gz2, 32, 32, P, c->bpart, grad + CL_OB2, st);
  a.x = c->gz2; a.w = c->w.w2b; a.zprev = c->z1; a.a = c->gz1;
  if (launch_conv(32, 32, CONV_EPI_DSILU, a, st)) return fail(SDNS_ERR_HIP, "closure layer 2 input gradient");
  // layer 1
  nch = launch_wgrad(32, 8, c->gz1, c->x0, c->wpart, n0, n1, n2, st);
  launch_wgrad_reduce(c->wpart, nch, 32, 8, 32, 3, grad + CL_OW1, st);
  launch_bgrad_bf16(c->gz1, 32, 32, P, c->bpart, grad + CL_OB1, st);
  if (loss) HIP_TRY(hipMemcpyAsync(loss, c->loss, sizeof(double), hipMemcpyDeviceToDevice, st));
  return check_launch();
}

int sdns_adam_step(float* p, const float* g, float* m, float* v, size_t n, float lr, float beta1, float beta2,
                   float eps, int t, void* hip_stream) {
  if (!p || !g || !m || !v || t < 1) return fail(SDNS_ERR_INVALID, "bad argument");
  launch_adam(p, g, m, v, (long)n, lr, beta1, beta2, eps, t, (hipStream_t)hip_stream);
  return check_launch();
}

int sdns_box_filter(sdns_plan* p, const void* u_phys, int r, void* ubar, void* tau) {
  if (!p || !u_phys || !ubar || !tau) return fail(SDNS_ERR_INVALID, "bad argument");
  const int rc = p->precision == SDNS_DOUBLE
                     ? launch_box_filter_f64((const double*)u_phys, p->N[0], p->N[1], p->N[2], r, (double*)ubar,
                                             (double*)tau, p->stream)
                     : launch_box_filter((const float*)u_phys, p->N[0], p->N[1], p->N[2], r, (float*)ubar,
                                         (float*)tau, p->stream);
  if (rc) return fail(SDNS_ERR_INVALID, "filter width %d does not divide the grid", r);
  return check_launch();
}

int sdns_closure_rk4_steps(sdns_closure* c, sdns_plan* p, void* u, const double* dts, int nsteps, double nu,
                           float in_scale, float out_scale, int per_stage, double* diag) {
  if (!c || !p || !u || (!dts && nsteps > 0) || nsteps < 0) return fail(SDNS_ERR_INVALID, "bad argument");
  if (!c->have_params) return fail(SDNS_ERR_INVALID, "closure parameters not set");
  if (p->precision != SDNS_SINGLE) return fail(SDNS_ERR_UNSUPPORTED, "the closure runs in single precision");
  for (int i = 0; i < 3; ++i)
    if (p->N[i] != c->n[i]) return fail(SDNS_ERR_INVALID, "plan and closure grids differ");
  int rc = ensure_workspace(p);
  if (rc) return rc;
  rc = rk4_steps<float>(p, u, nullptr, dts, nsteps, nu, diag, nullptr, nullptr, c, in_scale, out_scale, per_stage);
  return rc ? rc : check_launch();
}

}  // extern "C"
