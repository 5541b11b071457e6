// pgw_lib.hip - host side of the MI355X search core: device state, launch sequence of one
// sky point, and the C ABI declared in include/pgw.h.
//
// One handle = one segment (detector data resident in HBM) + one HIP stream.  job_core keeps
// the reference's per-sky-point semantics (search/network/src-cpu/jobcore.c:189-785); the host
// only does the scalar bookkeeping (grid -> sky coordinates, spindown range, candidate records).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pgw.h"
#include "pgw_kernels2.hpp"

using namespace pgw;

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      if (getenv("PGW_VERBOSE")) fprintf(stderr, "pgw: %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return PGW_ERR_DEVICE;                                                             \
    }                                                                                    \
  } while (0)

namespace {

constexpr int kCandCap = 1 << 20;   // candidates per sky point (device buffer)
constexpr int kCandChunk = 4096;    // candidates copied back with the count, without a second sync

int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

// L = L1 * L2 split (row length L1 in {1024, 2048, 4096}, column length L2 in 16..256)
bool split_len(int L, int& L1, int& L2) {
  switch (L) {
    case 1 << 14: L1 = 1024; L2 = 16; return true;
    case 1 << 15: L1 = 1024; L2 = 32; return true;
    case 1 << 16: L1 = 2048; L2 = 32; return true;
    case 1 << 17: L1 = 2048; L2 = 64; return true;
    case 1 << 18: L1 = 4096; L2 = 64; return true;
    case 1 << 19: L1 = 4096; L2 = 128; return true;
    case 1 << 20: L1 = 4096; L2 = 256; return true;
    default: return false;
  }
}

template <typename T, int P2> constexpr int col_tile() { return ColShape<P2, ColE<T, P2>::E>::C; }

template <typename T> struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    n = count;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) { p = nullptr; return PGW_ERR_NOMEM; }
    return PGW_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

// Twiddle table W_P^m = exp(-2 pi i m / P), m < P, computed in FP64 on the host.
template <typename T> int upload_tw(DevBuf<cplx<T>>& buf, int P, int stride_exp = 0, int count = -1) {
  if (count < 0) count = P;
  std::vector<cplx<T>> h(count);
  for (int m = 0; m < count; m++) {
    double s, c;
    const double mm = (double)m * (double)(1 << stride_exp);
    sincos(2.0 * M_PI * mm / (double)P, &s, &c);
    h[m].x = (T)c;
    h[m].y = (T)(-s);
  }
  int st = buf.alloc(count);
  if (st) return st;
  if (hipMemcpy(buf.p, h.data(), sizeof(cplx<T>) * count, hipMemcpyHostToDevice) != hipSuccess) return PGW_ERR_DEVICE;
  return PGW_OK;
}

template <typename T> struct Plan {
  int L = 0, L1 = 0, L2 = 0, twshift = 0;
  DevBuf<cplx<T>> twc, twr;
  int init(int len) {
    L = len;
    if (!split_len(L, L1, L2)) return PGW_ERR_UNSUPPORTED;
    twshift = 20 - ilog2(L);
    int st = upload_tw<T>(twc, L2);
    if (!st) st = upload_tw<T>(twr, L1);
    return st;
  }
  void release() { twc.release(); twr.release(); }
};

// ---------------------------------------------------------- launch helpers
template <typename T, int DIR, typename Loader>
void launch_col(const Plan<T>& pl, Loader ld, cplx<T>* Z, int nbatch, const cplx<T>* hi, const cplx<T>* lo,
                hipStream_t st) {
#define PGW_COL_CASE(P2)                                                                       \
  case P2:                                                                                     \
    hipLaunchKernelGGL((k_fft_col2<T, P2, DIR, Loader>), dim3(pl.L1 / col_tile<T, P2>(), nbatch), dim3(256), 0, \
                       st, ld, Z, pl.L1, pl.twc.p, hi, lo, pl.twshift);                        \
    break;
  switch (pl.L2) {
    PGW_COL_CASE(16)
    PGW_COL_CASE(32)
    PGW_COL_CASE(64)
    PGW_COL_CASE(128)
    PGW_COL_CASE(256)
  }
#undef PGW_COL_CASE
}

template <typename T, int DIR>
void launch_row(const Plan<T>& pl, cplx<T>* Z, int nbatch, hipStream_t st) {
#define PGW_ROW_CASE(P1)                                                                                    \
  case P1:                                                                                                  \
    hipLaunchKernelGGL((k_fft_row2<T, P1, DIR>), dim3(pl.L2, nbatch), dim3(P1 / 16), 0, st, Z, pl.L2, pl.twr.p); \
    break;
  switch (pl.L1) {
    PGW_ROW_CASE(1024)
    PGW_ROW_CASE(2048)
    PGW_ROW_CASE(4096)
  }
#undef PGW_ROW_CASE
}

template <typename T>
void launch_spin_col(const Plan<T>& pl, const SpinParams& sp, const cplx<T>* xm, const double* shft, cplx<T>* Z,
                     int nt, const cplx<T>* hi, const cplx<T>* lo, bool seq, hipStream_t st) {
#define PGW_SCOL_CASE(P2)                                                                                \
  case P2:                                                                                               \
    if (seq || ColE<T, P2>::SEQ)                                                                         \
      hipLaunchKernelGGL((k_spin_col2<T, P2, true>), dim3(pl.L1 / col_tile<T, P2>(), nt), dim3(256), 0, st, sp, xm, \
                         shft, Z, pl.twc.p, hi, lo);                                                     \
    else                                                                                                 \
      hipLaunchKernelGGL((k_spin_col2<T, P2, false>), dim3(pl.L1 / col_tile<T, P2>(), nt), dim3(256), 0, st, sp, xm, \
                         shft, Z, pl.twc.p, hi, lo);                                                     \
    break;
  switch (pl.L2) {
    PGW_SCOL_CASE(16)
    PGW_SCOL_CASE(32)
    PGW_SCOL_CASE(64)
    PGW_SCOL_CASE(128)
    PGW_SCOL_CASE(256)
  }
#undef PGW_SCOL_CASE
}

template <typename T>
void launch_spin_row(const Plan<T>& pl, const RowParams& rp, const cplx<T>* Z, const double* aabb, T* Ft, T* psum,
                     T* pmax, int nt, int na, hipStream_t st) {
#define PGW_SROW_CASE(P1)                                                                                 \
  case P1:                                                                                                \
    if (na == 2)                                                                                          \
      hipLaunchKernelGGL((k_spin_row2<T, P1, 2>), dim3(pl.L2, nt), dim3(P1 / 16), 0, st, rp, Z, aabb, pl.twr.p, Ft, \
                         psum, pmax);                                                                     \
    else                                                                                                  \
      hipLaunchKernelGGL((k_spin_row2<T, P1, 1>), dim3(pl.L2, nt), dim3(P1 / 16), 0, st, rp, Z, aabb, pl.twr.p, Ft, \
                         psum, pmax);                                                                     \
    break;
  switch (pl.L1) {
    PGW_SROW_CASE(1024)
    PGW_SROW_CASE(2048)
    PGW_SROW_CASE(4096)
  }
#undef PGW_SROW_CASE
}

// lin2ast, search/network/src-cpu/auxi.c:15-24 (host; same expression order)
void lin2ast(double be1, double be2, int pm, double sepsm, double cepsm, double* sinal, double* cosal,
             double* sindel, double* cosdel) {
#pragma clang fp contract(off)
  *sindel = be1 * sepsm - (2 * pm - 3) * sqrt(1. - be1 * be1 - be2 * be2) * cepsm;
  *cosdel = sqrt(1. - (*sindel) * (*sindel));
  *sinal = (be1 - sepsm * (*sindel)) / (cepsm * (*cosdel));
  *cosal = be2 / (*cosdel);
}

}  // namespace

// ---------------------------------------------------------------- handle --
struct pgw_handle {
  virtual ~pgw_handle() {}
  virtual int job_core(int pm, int mm, int nn, pgw_range* r, const pgw_opts* o, int* sgnlc, double* sgnlv, int cap,
                       int* FNum) = 0;
  virtual int job_batch(const int* sky, int nsky, pgw_range* r, const pgw_opts* o, int* sgnlc, double* sgnlv,
                        int cap, int* FNum) = 0;
  virtual int debug_fstat(int pm, int mm, int nn, int ss, const pgw_opts* o, double* F) = 0;
  virtual int debug_resampled(int pm, int mm, int nn, int det, double* a, double* b) = 0;
  virtual int time_stages(int pm, int mm, int nn, int s0, int nt, const pgw_opts* o, int reps, double* ms) = 0;
  double last_ms = 0.0;
};

namespace {

struct SkyState {
  double het0, sgnlt2, sgnlt3;
};

template <typename T> struct Impl final : pgw_handle {
  using T2 = cplx<T>;
  pgw_settings sett;
  int nifo = 0;
  std::vector<double> amod, sig2, sphir, cphir, ssb0;  // per detector
  Plan<T> pf, pi, ps;                                    // nfft fwd, Ninterp inv, L = nfftf
  DevBuf<T2> hi, lo;
  DevBuf<double> xDat, DetSSB, cosmodf, sinmodf, shft, shftf, paa, pbb, aabb, mu;
  DevBuf<T2> R, Zf, Zi, Y, xm, Z;
  DevBuf<double2> u, y2;
  DevBuf<T> Ft, psum, pmax;
  DevBuf<SplineTab> spl;
  DevBuf<Cand> cand;
  DevBuf<int> ncand;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evd[2] = {nullptr, nullptr};  // candidate copy of buffer 0/1 done
  Cand* hc_pin = nullptr;                  // pinned mirrors: 2 x kCandChunk candidates, 2 counts
  int* hn_pin = nullptr;
  int TB = 128;         // templates per launch
  int row_na = 1;        // row pass: 2 = both spectra through one LDS exchange, 1 = one after the other
  bool col_seq = false;  // column pass: spectra share one LDS buffer in turn
  int nblk = 0, nparts = 0, npro = 0;
  int nspl = 0;         // spline knots interpftpad*N
  std::vector<Cand> hcand;

  ~Impl() override {
    pf.release(); pi.release(); ps.release();
    hi.release(); lo.release();
    xDat.release(); DetSSB.release(); cosmodf.release(); sinmodf.release(); shft.release(); shftf.release();
    paa.release(); pbb.release(); aabb.release(); mu.release();
    R.release(); Zf.release(); Zi.release(); Y.release(); xm.release(); Z.release();
    u.release(); y2.release(); Ft.release(); psum.release(); pmax.release(); spl.release();
    cand.release(); ncand.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (auto& e : evd)
      if (e) (void)hipEventDestroy(e);
    if (hc_pin) (void)hipHostFree(hc_pin);
    if (hn_pin) (void)hipHostFree(hn_pin);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init(const pgw_settings* s, const pgw_detector* ifo, int n) {
    sett = *s;
    nifo = n;
    sett.nifo = n;
    const int N = sett.N;
    if (sett.Ninterp != sett.interpftpad * sett.nfft || sett.nfftf != sett.fftpad * sett.nfft) return PGW_ERR_ARG;
    if (sett.interpftpad * N > sett.Ninterp || N > sett.nfft || sett.nmin < 1 || sett.nmax > sett.nfftf ||
        sett.nmax <= sett.nmin)
      return PGW_ERR_ARG;
    if (sett.Ninterp - sett.nfft != sett.nfft) return PGW_ERR_UNSUPPORTED;  // zero-insert map assumes interpftpad = 2
    const int M = sett.nmax - sett.nmin;
    nblk = M / PGW_NAV;
    nparts = nblk + (M % PGW_NAV ? 1 : 0);
    if (nparts > 256) return PGW_ERR_UNSUPPORTED;
    nspl = sett.interpftpad * N;
    // kernel variants; PGW_TUNE="row_na=2,col_seq=0,tb=32" overrides (used when measuring variants)
    if (const char* tune = getenv("PGW_TUNE")) {
      int v;
      if (const char* q = strstr(tune, "row_na=")) if (sscanf(q, "row_na=%d", &v) == 1 && (v == 1 || v == 2)) row_na = v;
      if (const char* q = strstr(tune, "col_seq=")) if (sscanf(q, "col_seq=%d", &v) == 1) col_seq = v != 0;
      if (const char* q = strstr(tune, "tb=")) if (sscanf(q, "tb=%d", &v) == 1 && v >= 1 && v <= 256) TB = v;
    }
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreateWithFlags(&evd[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&evd[1], hipEventDisableTiming));
    if (hipHostMalloc((void**)&hc_pin, sizeof(Cand) * 2 * kCandChunk, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&hn_pin, sizeof(int) * 2, hipHostMallocDefault) != hipSuccess)
      return PGW_ERR_NOMEM;
    int st;
    if ((st = pf.init(sett.nfft)) || (st = pi.init(sett.Ninterp)) || (st = ps.init(sett.nfftf))) return st;
    // W_{2^20}^m two-level tables: hi[h] = W^{1024 h}, lo[l] = W^l
    if ((st = upload_tw<T>(hi, 1 << 20, 10, 1024)) || (st = upload_tw<T>(lo, 1 << 20, 0, 1024))) return st;
    amod.resize(9 * n); sig2.resize(n); sphir.resize(n); cphir.resize(n); ssb0.resize(3 * n);
    std::vector<double> hx((size_t)n * N), hd((size_t)n * 3 * N);
    for (int d = 0; d < n; d++) {
      if (!ifo[d].xDat || !ifo[d].DetSSB || !(ifo[d].sig2 > 0)) return PGW_ERR_ARG;
      memcpy(&amod[9 * d], ifo[d].amod, 9 * sizeof(double));
      sig2[d] = ifo[d].sig2; sphir[d] = ifo[d].sphir; cphir[d] = ifo[d].cphir;
      memcpy(&ssb0[3 * d], ifo[d].DetSSB, 3 * sizeof(double));
      memcpy(&hx[(size_t)d * N], ifo[d].xDat, sizeof(double) * N);
      memcpy(&hd[(size_t)d * 3 * N], ifo[d].DetSSB, sizeof(double) * 3 * N);
    }
    // Earth-rotation auxiliary arrays, init.c:448-455 (host glibc, same values as the CPU path)
    std::vector<double> hc(N), hs(N);
    for (int i = 0; i < N; i++) {
      const double omrt = sett.omr * i;
      hc[i] = cos(omrt);
      hs[i] = sin(omrt);
    }
    npro = (sett.nfft + NT - 1) / NT;
    const size_t L = sett.nfftf;
    if ((st = xDat.alloc((size_t)n * N)) || (st = DetSSB.alloc((size_t)n * 3 * N)) || (st = cosmodf.alloc(N)) ||
        (st = sinmodf.alloc(N)) || (st = shft.alloc((size_t)n * N)) || (st = shftf.alloc((size_t)n * N)) ||
        (st = paa.alloc((size_t)n * npro)) || (st = pbb.alloc((size_t)n * npro)) || (st = aabb.alloc(2)) ||
        (st = R.alloc((size_t)2 * n * sett.nfft)) || (st = Zf.alloc((size_t)2 * n * sett.nfft)) ||
        (st = Zi.alloc((size_t)2 * n * sett.Ninterp)) || (st = Y.alloc((size_t)2 * n * sett.Ninterp)) ||
        (st = u.alloc((size_t)2 * n * (nspl + 1))) || (st = y2.alloc((size_t)2 * n * (nspl + 1))) ||
        (st = xm.alloc((size_t)2 * n * N)) || (st = Z.alloc((size_t)TB * 2 * L)) || (st = Ft.alloc((size_t)TB * L)) ||
        (st = psum.alloc((size_t)TB * nparts * ps.L2)) || (st = pmax.alloc((size_t)TB * nparts * ps.L2)) ||
        (st = mu.alloc((size_t)TB * nparts)) || (st = spl.alloc(1)) || (st = cand.alloc(2 * (size_t)kCandCap)) ||
        (st = ncand.alloc(2)))
      return st;
    HIPCHK(hipMemcpy(xDat.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(DetSSB.p, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(cosmodf.p, hc.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(sinmodf.p, hs.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    // spline recursion coefficients (auxi.c:66-72): invp_i = 2/(c_{i-1}+4), c_i = -invp_i/2, c_0 = 0
    SplineTab tab;
    double c = 0.0;
    tab.c[0] = 0.0;
    tab.invp[0] = 0.0;
    for (int i = 1; i < SPL_NTAB; i++) {
      const double invp = 2. / (c + 4.);
      c = -.5 * invp;
      tab.invp[i] = invp;
      tab.c[i] = c;
    }
    tab.invp_inf = 2. / (c + 4.);
    tab.c_inf = -.5 * tab.invp_inf;
    if (tab.c_inf != c) return PGW_ERR_UNSUPPORTED;  // recursion must have reached its fixed point
    HIPCHK(hipMemcpy(spl.p, &tab, sizeof(tab), hipMemcpyHostToDevice));
    hcand.resize(kCandCap);
    return PGW_OK;
  }

  // Sky-point stage, jobcore.c:243-393.  Returns 1 when the point is inside the disk.
  int sky_stage(int pm, int mm, int nn, SkyState& ss, int& status) {
#pragma clang fp contract(off)
    status = PGW_OK;
    const double* M = sett.M;
    const double al1 = nn * M[10] + mm * M[14];
    const double al2 = nn * M[11] + mm * M[15];
    if ((al1 * al1 + al2 * al2) / (sett.oms * sett.oms) > 1.) return 0;
    double sinalt, cosalt, sindelt, cosdelt;
    lin2ast(al1 / sett.oms, al2 / sett.oms, pm, sett.sepsm, sett.cepsm, &sinalt, &cosalt, &sindelt, &cosdelt);
    ss.sgnlt2 = asin(sindelt);
    ss.sgnlt3 = fmod(atan2(sinalt, cosalt) + 2. * M_PI, 2. * M_PI);
    ss.het0 = fmod(nn * M[8] + mm * M[12], M[0]);
    SkyParams p;
    memset(&p, 0, sizeof(p));
    p.het0 = ss.het0;
    p.oms = sett.oms;
    p.N = sett.N;
    p.nfft = sett.nfft;
    p.nifo = nifo;
    for (int d = 0; d < nifo; d++) {
      SkyDet& q = p.det[d];
      memcpy(q.c, &amod[9 * d], sizeof(q.c));
      q.cosalfr = cosalt * cphir[d] + sinalt * sphir[d];
      q.sinalfr = sinalt * cphir[d] - cosalt * sphir[d];
      q.c2d = cosdelt * cosdelt;
      q.c2sd = sindelt * cosdelt;
      q.sindel = sindelt;
      q.cosdel = cosdelt;
      q.nS0 = cosalt * cosdelt;
      q.nS1 = sinalt * cosdelt;
      q.nS2 = sindelt;
      q.shft1 = q.nS0 * ssb0[3 * d] + q.nS1 * ssb0[3 * d + 1] + q.nS2 * ssb0[3 * d + 2];
      q.sig2 = sig2[d];
    }
    const int nb = 2 * nifo;
    hipLaunchKernelGGL((k_sky_prologue<T>), dim3(npro, nifo), dim3(NT), 0, stream, p, xDat.p, DetSSB.p, cosmodf.p,
                       sinmodf.p, R.p, shft.p, shftf.p, paa.p, pbb.p);
    hipLaunchKernelGGL(k_aabb, dim3(1), dim3(NT), 0, stream, p, paa.p, pbb.p, npro, aabb.p);
    // resampling: forward FFT (nfft), zero-insert, backward FFT (Ninterp), scale (jobcore.c:319-362)
    LoadPlain<T2> lp{R.p, sett.nfft, sett.nfft};
    launch_col<T, -1>(pf, lp, Zf.p, nb, hi.p, lo.p, stream);
    launch_row<T, -1>(pf, Zf.p, nb, stream);
    LoadZeroInsert<T2> lz{Zf.p, sett.nfft, sett.Ninterp, sett.nfft / 2 + 1, pf.L1, pf.L2};
    launch_col<T, +1>(pi, lz, Zi.p, nb, hi.p, lo.p, stream);
    launch_row<T, +1>(pi, Zi.p, nb, stream);
    const int count = std::min(sett.Ninterp, nspl);
    const int rows_k1 = std::min(pi.L1, (count + pi.L2 - 1) / pi.L2);
    dim3 gt((rows_k1 + 31) / 32, (pi.L2 + 31) / 32, nb);
    const T ft = (T)((double)sett.interpftpad / sett.Ninterp);
    hipLaunchKernelGGL((k_untranspose<T>), gt, dim3(NT), 0, stream, Zi.p, Y.p, pi.L1, pi.L2, count, ft);
    // spline on the first interpftpad*N samples, evaluated at interpftpad*(i - shftf[i]) (auxi.c:98-122)
    const int nchunk = (nspl + SPL_K - 1) / SPL_K;
    dim3 gs((nchunk + NT - 1) / NT, nb);
    hipLaunchKernelGGL((k_spline_u<T>), gs, dim3(NT), 0, stream, Y.p, (size_t)sett.Ninterp, nspl, spl.p, u.p,
                       (size_t)(nspl + 1));
    hipLaunchKernelGGL(k_spline_y2, gs, dim3(NT), 0, stream, u.p, (size_t)(nspl + 1), nspl, spl.p, y2.p);
    dim3 ge((sett.N + NT - 1) / NT, nb);
    hipLaunchKernelGGL((k_spline_eval<T, T>), ge, dim3(NT), 0, stream, Y.p, (size_t)sett.Ninterp, y2.p,
                       (size_t)(nspl + 1), shftf.p, sett.N, sett.interpftpad, p, xm.p);
    if (hipGetLastError() != hipSuccess) status = PGW_ERR_DEVICE;
    return 1;
  }

  void spindown_range(int mm, int nn, const pgw_range* r, const pgw_opts* o, int& smin, int& smax) {
#pragma clang fp contract(off)
    const double* M = sett.M;
    smin = r->sst;
    smax = r->spndr[1];
    if (!o->use_range) {  // jobcore.c:424-439
      smin = (int)trunc((sett.Smin - nn * M[9] - mm * M[13]) / M[5]);
      smax = (int)trunc(-(nn * M[9] + mm * M[13] + sett.Smax) / M[5]);
      if (smin > smax) std::swap(smin, smax);
    }
    if (o->s0_flag) smin = smax;
  }

  // launches the template stage for spindowns [s0, s0+nt) into candidate buffer buf;
  // ev (optional) = 4 events bracketing the kernels
  int launch_templates(int mm, int nn, int s0, int nt, const pgw_opts* o, int buf = 0, hipEvent_t* ev = nullptr) {
    const double* M = sett.M;
    SpinParams sp;
    sp.M0 = M[0]; sp.M4 = M[4]; sp.M5 = M[5]; sp.M9 = M[9]; sp.M13 = M[13];
    sp.nn = nn; sp.mm = mm; sp.ss0 = s0; sp.s0_flag = o->s0_flag;
    sp.N = sett.N; sp.L1 = ps.L1; sp.nifo = nifo; sp.twshift = ps.twshift;
    if (ev) HIPCHK(hipEventRecord(ev[0], stream));
    launch_spin_col<T>(ps, sp, xm.p, shft.p, Z.p, nt, hi.p, lo.p, col_seq, stream);
    if (ev) HIPCHK(hipEventRecord(ev[1], stream));
    RowParams rp{ps.L2, sett.nmin, sett.nmax, nblk, nparts};
    launch_spin_row<T>(ps, rp, Z.p, aabb.p, Ft.p, psum.p, pmax.p, nt, row_na, stream);
    if (ev) HIPCHK(hipEventRecord(ev[2], stream));
    ThrParams tp{ps.L1, ps.L2, sett.nmin, sett.nmax, nblk, nparts, o->white_flag, s0, kCandCap, o->trl};
    hipLaunchKernelGGL((k_threshold<T>), dim3(nparts, nt), dim3(64), 0, stream, tp, Ft.p, psum.p, pmax.p, mu.p,
                       cand.p + (size_t)buf * kCandCap, ncand.p + buf);
    if (ev) HIPCHK(hipEventRecord(ev[3], stream));
    return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
  }

  // candidate records in the reference's order and format (jobcore.c:659-695)
  int emit(const SkyState& sk, int mm, int nn, const pgw_opts* o, int nc, int* sgnlc, double* sgnlv, int cap) {
#pragma clang fp contract(off)
    const double* M = sett.M;
    std::sort(hcand.begin(), hcand.begin() + nc,
              [](const Cand& a, const Cand& b) { return a.ss != b.ss ? a.ss < b.ss : a.ii < b.ii; });
    std::vector<double> rec;
    rec.reserve((size_t)nc * PGW_NPAR);
    for (int c = 0; c < nc; c++) {
      const int ss = hcand[c].ss, ii = hcand[c].ii;
      const double Fc = hcand[c].F;
      const double sgnlt1 = (o->s0_flag ? 0. : ss * M[5] + nn * M[9] + mm * M[13]);
      double het1 = fmod(ss * M[4], M[0]);
      if (het1 < 0) het1 += M[0];
      const double sgnl0 = sk.het0 + het1;
      double r[PGW_NPAR];
      r[0] = 2. * M_PI * ii / ((double)sett.fftpad * sett.nfft) + sgnl0;
      r[1] = sgnlt1;
      r[2] = sk.sgnlt2;
      r[3] = sk.sgnlt3;
      r[4] = sqrt(2. * (Fc - sett.nd));
      bool veto = false;
      for (int k = 0; k < sett.numlines_band; k++)
        if (r[0] >= sett.lines[k][0] && r[0] <= sett.lines[k][1]) { veto = true; break; }
      if (!veto) rec.insert(rec.end(), r, r + PGW_NPAR);
    }
    const int nrec = (int)(rec.size() / PGW_NPAR);
    if (*sgnlc + nrec > cap) {
      *sgnlc += nrec;
      return PGW_ERR_CAPACITY;
    }
    memcpy(sgnlv + (size_t)PGW_NPAR * (*sgnlc), rec.data(), sizeof(double) * rec.size());
    *sgnlc += nrec;
    return PGW_OK;
  }

  struct Pending {
    SkyState sk;
    int mm, nn, buf;
  };

  // Queues all device work of one sky point into candidate buffer buf, followed by the async copy
  // of the candidate count and the first kCandChunk candidates; returns 0 outside the disk.
  int launch_sky(int pm, int mm, int nn, pgw_range* r, const pgw_opts* o, int buf, SkyState& sk, int* FNum,
                 int& status) {
    if (!sky_stage(pm, mm, nn, sk, status)) return 0;
    if (status) return 1;
    int smin, smax;
    spindown_range(mm, nn, r, o, smin, smax);
    status = PGW_ERR_DEVICE;
    if (hipMemsetAsync(ncand.p + buf, 0, sizeof(int), stream) != hipSuccess) return 1;
    for (int s0 = smin; s0 <= smax; s0 += TB) {
      const int nt = std::min(TB, smax - s0 + 1);
      if ((status = launch_templates(mm, nn, s0, nt, o, buf))) return 1;
    }
    status = PGW_ERR_DEVICE;
    if (hipMemcpyAsync(hn_pin + buf, ncand.p + buf, sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipMemcpyAsync(hc_pin + (size_t)buf * kCandChunk, cand.p + (size_t)buf * kCandCap, sizeof(Cand) * kCandChunk,
                       hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipEventRecord(evd[buf], stream) != hipSuccess)
      return 1;
    status = PGW_OK;
    if (smax >= smin) *FNum += smax - smin + 1;
    return 1;
  }

  // Waits for a queued sky point's candidates and appends its records.
  int finish(const Pending& pd, const pgw_opts* o, int* sgnlc, double* sgnlv, int cap) {
    HIPCHK(hipEventSynchronize(evd[pd.buf]));
    const int nc = hn_pin[pd.buf];
    if (nc > kCandCap) return PGW_ERR_CAPACITY;
    memcpy(hcand.data(), hc_pin + (size_t)pd.buf * kCandChunk, sizeof(Cand) * std::min(nc, kCandChunk));
    if (nc > kCandChunk) {  // rare: the rest synchronously (buffer pd.buf is not reused before this returns)
      HIPCHK(hipMemcpyAsync(hcand.data() + kCandChunk, cand.p + (size_t)pd.buf * kCandCap + kCandChunk,
                            sizeof(Cand) * (nc - kCandChunk), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    return emit(pd.sk, pd.mm, pd.nn, o, nc, sgnlc, sgnlv, cap);
  }

  // Sky points in order; the device works on point i while the host turns point i-1 into records.
  int run_batch(const int* sky, int nsky, pgw_range* r, const pgw_opts* o, int* sgnlc, double* sgnlv, int cap,
                int* FNum) {
    HIPCHK(hipEventRecord(ev0, stream));
    Pending prev{};
    bool have_prev = false;
    int buf = 0, st = PGW_OK;
    for (int i = 0; i < nsky && !st; i++) {
      Pending cur{};
      cur.mm = sky[3 * i + 1];
      cur.nn = sky[3 * i + 2];
      cur.buf = buf;
      if (!launch_sky(sky[3 * i], cur.mm, cur.nn, r, o, buf, cur.sk, FNum, st)) continue;
      if (st) break;
      if (have_prev) st = finish(prev, o, sgnlc, sgnlv, cap);
      prev = cur;
      have_prev = true;
      buf ^= 1;
    }
    HIPCHK(hipEventRecord(ev1, stream));
    if (have_prev && !st) st = finish(prev, o, sgnlc, sgnlv, cap);
    HIPCHK(hipEventSynchronize(ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
    last_ms = t;
    return st;
  }

  int job_core(int pm, int mm, int nn, pgw_range* r, const pgw_opts* o, int* sgnlc, double* sgnlv, int cap,
               int* FNum) override {
    const int sky[3] = {pm, mm, nn};
    return run_batch(sky, 1, r, o, sgnlc, sgnlv, cap, FNum);
  }

  int job_batch(const int* sky, int nsky, pgw_range* r, const pgw_opts* o, int* sgnlc, double* sgnlv, int cap,
                int* FNum) override {
    return run_batch(sky, nsky, r, o, sgnlc, sgnlv, cap, FNum);
  }

  int debug_fstat(int pm, int mm, int nn, int ss, const pgw_opts* o, double* Fout) override {
    SkyState sk;
    int st;
    if (!sky_stage(pm, mm, nn, sk, st)) return 1;
    if (st) return st;
    HIPCHK(hipMemsetAsync(ncand.p, 0, sizeof(int), stream));
    if ((st = launch_templates(mm, nn, ss, 1, o))) return st;
    const size_t L = sett.nfftf;
    std::vector<T> hF(L);
    std::vector<double> hmu(nparts);
    HIPCHK(hipMemcpyAsync(hF.data(), Ft.p, sizeof(T) * L, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(hmu.data(), mu.p, sizeof(double) * nparts, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int k = sett.nmin; k < sett.nmax; k++) {
      const int part = std::min((k - sett.nmin) / PGW_NAV, nparts - 1);
      Fout[k - sett.nmin] = (double)hF[(size_t)(k % ps.L2) * ps.L1 + k / ps.L2] / hmu[part];
    }
    return PGW_OK;
  }

  int time_stages(int pm, int mm, int nn, int s0, int nt, const pgw_opts* o, int reps, double* ms) override {
    if (nt < 1 || reps < 1) return PGW_ERR_ARG;
    nt = std::min(nt, TB);  // at most one batch
    SkyState sk;
    int st;
    hipEvent_t ev[4];
    for (int k = 0; k < 4; k++) HIPCHK(hipEventCreate(&ev[k]));
    HIPCHK(hipEventRecord(ev0, stream));
    const int inside = sky_stage(pm, mm, nn, sk, st);
    HIPCHK(hipEventRecord(ev1, stream));
    if (!inside) st = 1;
    double acc[4] = {0, 0, 0, 0};
    if (!st) {
      HIPCHK(hipStreamSynchronize(stream));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
      acc[0] = t;
      for (int r = 0; r < reps && !st; r++) {
        HIPCHK(hipMemsetAsync(ncand.p, 0, sizeof(int), stream));
        st = launch_templates(mm, nn, s0, nt, o, 0, ev);
        HIPCHK(hipStreamSynchronize(stream));
        for (int k = 0; k < 3; k++) {
          HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
          acc[k + 1] += t;
        }
      }
    }
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(ev[k]);
    ms[0] = acc[0];
    for (int k = 1; k < 4; k++) ms[k] = acc[k] / reps;
    ms[4] = nt;
    return st;
  }

  int debug_resampled(int pm, int mm, int nn, int det, double* a, double* b) override {
    if (det < 0 || det >= nifo) return PGW_ERR_ARG;
    SkyState sk;
    int st;
    if (!sky_stage(pm, mm, nn, sk, st)) return 1;
    if (st) return st;
    const int N = sett.N;
    std::vector<T2> h((size_t)2 * N);
    HIPCHK(hipMemcpyAsync(h.data(), xm.p + (size_t)2 * det * N, sizeof(T2) * 2 * N, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < N; i++) {
      a[2 * i] = h[i].x; a[2 * i + 1] = h[i].y;
      b[2 * i] = h[N + i].x; b[2 * i + 1] = h[N + i].y;
    }
    return PGW_OK;
  }
};

}  // namespace

// ------------------------------------------------------------------ C ABI --
extern "C" {

int pgw_search_settings(double fpo, double dt, int N, int fftpad, pgw_settings* sett) {
#pragma clang fp contract(off)
  if (!sett || N < 16 || !(dt > 0) || fftpad < 1) return PGW_ERR_ARG;
  memset(sett, 0, sizeof(*sett));
  sett->fpo = fpo;
  sett->dt = dt;
  sett->B = 0.5 / dt;
  sett->oms = 2. * M_PI * fpo * dt;
  sett->omr = 7.2921151467064e-5 * dt;
  sett->N = N;
  sett->nfft = 1 << (int)ceil(log(N) / log(2.));
  sett->Smax = 2. * M_PI * 0.5e-8 * dt * dt;
  sett->Smin = 2. * M_PI * 0.5e-9 * dt * dt;
  sett->nd = 2;
  sett->interpftpad = 2;
  sett->fftpad = fftpad;
  sett->nmin = (int)(sett->fftpad * PGW_NAVB * sett->B);
  sett->nmax = (int)((sett->nfft / 2 - PGW_NAVB * sett->B) * sett->fftpad);
  sett->Ninterp = sett->interpftpad * sett->nfft;
  sett->nfftf = sett->fftpad * sett->nfft;
  sett->sepsm = sin(0.40909280422232897);
  sett->cepsm = cos(0.40909280422232897);
  return PGW_OK;
}

int pgw_rogcvir(double ephi, double egam, double c[9]) {
#pragma clang fp contract(off)
  if (!c) return PGW_ERR_ARG;
  const double sp = sin(ephi), cp = cos(ephi);
  c[0] = .25 * sin(2. * egam) * (1 + sp * sp);
  c[1] = -.5 * cos(2. * egam) * sp;
  c[2] = .5 * sin(2. * egam) * sin(2. * ephi);
  c[3] = -cos(2. * egam) * cp;
  c[4] = .75 * sin(2. * egam) * (cp * cp);
  c[5] = cos(2. * egam) * sp;
  c[6] = .5 * sin(2. * egam) * (1. + sp * sp);
  c[7] = cos(2. * egam) * cp;
  c[8] = .5 * sin(2. * egam) * sin(2. * ephi);
  return PGW_OK;
}

double pgw_sig2(const double* x, int n) {
#pragma clang fp contract(off)
  if (!x || n < 2) return 0.0;
  int nz = 0;
  for (int j = 0; j < n; j++)
    if (!x[j]) nz++;
  if (nz == n) return 0.0;
  const double crf0 = (double)n / (n - nz);
  double mean = 0., var = 0.;
  for (int i = 0; i < n; i++) mean += x[i];
  mean /= n;
  for (int i = 0; i < n; i++) var += (x[i] - mean) * (x[i] - mean);
  var /= (n - 1);
  return crf0 * var;
}

int pgw_init(pgw_handle** h, const pgw_settings* sett, const pgw_detector* ifo, int nifo, int precision) {
  if (!h || !sett || !ifo || nifo < 1 || nifo > PGW_MAX_DETECTORS) return PGW_ERR_ARG;
  *h = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PGW_ERR_NODEVICE;
  pgw_handle* p = nullptr;
  int st;
  if (precision == PGW_PREC_FP64) {
    auto* q = new (std::nothrow) Impl<double>();
    if (!q) return PGW_ERR_NOMEM;
    st = q->init(sett, ifo, nifo);
    p = q;
  } else if (precision == PGW_PREC_FP32) {
    auto* q = new (std::nothrow) Impl<float>();
    if (!q) return PGW_ERR_NOMEM;
    st = q->init(sett, ifo, nifo);
    p = q;
  } else {
    return PGW_ERR_ARG;
  }
  if (st) {
    delete p;
    return st;
  }
  *h = p;
  return PGW_OK;
}

int pgw_job_core(pgw_handle* h, int pm, int mm, int nn, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                 double* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv)) return PGW_ERR_ARG;
  return h->job_core(pm, mm, nn, s_range, opts, sgnlc, sgnlv, sgnlv_cap, FNum);
}

int pgw_job_batch(pgw_handle* h, const int* sky, int nsky, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                  double* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || (!sky && nsky > 0) || nsky < 0 || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv))
    return PGW_ERR_ARG;
  return h->job_batch(sky, nsky, s_range, opts, sgnlc, sgnlv, sgnlv_cap, FNum);
}

int pgw_debug_fstat(pgw_handle* h, int pm, int mm, int nn, int ss, const pgw_opts* opts, double* F) {
  if (!h || !opts || !F) return PGW_ERR_ARG;
  return h->debug_fstat(pm, mm, nn, ss, opts, F);
}

int pgw_debug_resampled(pgw_handle* h, int pm, int mm, int nn, int det, double* xDatma, double* xDatmb) {
  if (!h || !xDatma || !xDatmb) return PGW_ERR_ARG;
  return h->debug_resampled(pm, mm, nn, det, xDatma, xDatmb);
}

int pgw_time_stages(pgw_handle* h, int pm, int mm, int nn, int s0, int nt, const pgw_opts* opts, int reps,
                    double ms[5]) {
  if (!h || !opts || !ms) return PGW_ERR_ARG;
  return h->time_stages(pm, mm, nn, s0, nt, opts, reps, ms);
}

int pgw_copy_bench(size_t bytes, int reps, double* ms) {
  if (!ms || bytes < 16 || reps < 1) return PGW_ERR_ARG;
  const size_t n = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  HIPCHK(hipMalloc(&a, n * sizeof(double4)));
  if (hipMalloc(&b, n * sizeof(double4)) != hipSuccess) { (void)hipFree(a); return PGW_ERR_NOMEM; }
  HIPCHK(hipMemset(a, 0, n * sizeof(double4)));
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int grid = (int)std::min<size_t>((n + NT - 1) / NT, 1 << 20);
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(NT), 0, 0, a, b, n);  // warm-up
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(NT), 0, 0, a, b, n);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return PGW_OK;
}

double pgw_last_gpu_ms(const pgw_handle* h) { return h ? h->last_ms : 0.0; }

void pgw_destroy(pgw_handle* h) { delete h; }

const char* pgw_strerror(int status) {
  switch (status) {
    case PGW_OK: return "ok";
    case PGW_ERR_ARG: return "invalid argument";
    case PGW_ERR_CAPACITY: return "candidate buffer capacity exceeded";
    case PGW_ERR_DEVICE: return "HIP runtime error";
    case PGW_ERR_NOMEM: return "out of memory";
    case PGW_ERR_UNSUPPORTED: return "unsupported size";
    case PGW_ERR_NODEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

}  // extern "C"
