// pgw_lib.hip - host side of the MI355X search core: device state, launch sequences, and the
// C ABI declared in include/pgw.h.
//
// A handle holds one or more segments (detector data resident in HBM) and one HIP stream.  Work
// is a list of (segment, pm, mm, nn) sky points processed in groups: the sky-point stage of each
// point of a group fills one of kSlots slots (resampled series, _tmp1, aa/bb), then template
// launches of up to TB templates run over all spindowns of all points of the group, described by
// per-template descriptors in device memory; each launch's threshold runs in the first
// workgroups of the next launch's column pass.  Points of one group share the F-statistic FFT
// length and bin range, so segments of uneven N with the same nfft batch together.  The host does
// only the scalar bookkeeping of job_core (search/network/src-cpu/jobcore.c:189-785): sky
// coordinates, spindown range, per-template het1/sgnlt1 and the candidate records; it turns the
// records of group g-1 into output while the device works on group g.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/pgw.h"
#include "pgw_fstatnet.hpp"
#include "pgw_fused.hpp"
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

constexpr int kCandCap = 1 << 20;   // candidates per group (device buffer)
constexpr int kCandChunk = 4096;    // candidates copied back with the count, without a second sync
// A group's device candidate buffer: a header Cand {t = count, ii = error word} at [0] (the count
// keeps counting past the capacity; the error word is set by a fused launch whose dependency wait
// ran out), the candidates from [1].  One copy brings the header and the first kCandChunk back.
constexpr int kSlots = 16;          // sky points per group
static_assert(kSlots <= kSpinSlots, "a launch's sky slots fit the kernel-argument table");
constexpr int kDescCap = 1 << 16;   // templates per group
constexpr int kTBMax = 256;         // templates per launch, upper bound
constexpr size_t kScratchDoubles = PGW_FSN_SCRATCH_BYTES / sizeof(double);

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
    release();
    n = count;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) { p = nullptr; return PGW_ERR_NOMEM; }
    return PGW_OK;
  }
  // keeps the buffer if it already holds count elements (the caller has synchronised its users)
  int alloc_min(size_t count) { return p && n >= count ? PGW_OK : alloc(count); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  ~DevBuf() { release(); }
};

template <typename T> struct PinBuf {
  T* p = nullptr;
  int alloc(size_t count) {
    return hipHostMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1), hipHostMallocDefault) == hipSuccess
               ? PGW_OK : PGW_ERR_NOMEM;
  }
  ~PinBuf() {
    if (p) (void)hipHostFree(p);
  }
};

// Twiddle table W_P^m = exp(-2 pi i m / P) (m scaled by 2^stride_exp), computed in FP64 on the host.
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
  // l1max > 0: rows of at most l1max points where the column length stays <= 256 (shorter rows,
  // longer columns)
  int init(int len, int l1max = 0) {
    L = len;
    if (!split_len(L, L1, L2)) return PGW_ERR_UNSUPPORTED;
    while (l1max >= 1024 && L1 > l1max && L2 < 256) {
      L1 /= 2;
      L2 *= 2;
    }
    twshift = 20 - ilog2(L);
    int st = upload_tw<T>(twc, L2);
    if (!st) st = upload_tw<T>(twr, L1);
    return st;
  }
};

// ---------------------------------------------------------- launch helpers
template <typename T, int DIR, bool OTW = true, typename Loader>
void launch_col(const Plan<T>& pl, Loader ld, cplx<T>* Z, int nbatch, const cplx<T>* hi, const cplx<T>* lo,
                hipStream_t st) {
#define PGW_COL_CASE(P2)                                                                       \
  case P2:                                                                                     \
    hipLaunchKernelGGL((k_fft_col2<T, P2, DIR, Loader, OTW>), dim3(pl.L1 / col_tile<T, P2>(), nbatch), dim3(256), 0, \
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

// forward row transform, phase shift and sign, inverse row transform on each row (k_fft_row_shift)
template <typename T>
void launch_row_shift(const Plan<T>& pl, cplx<T>* Z, int nbatch, const cplx<T>* hi, const cplx<T>* lo, int nyqst,
                      hipStream_t st) {
  const int shift1 = 19 - ilog2(pl.L1);  // e^{i pi k1 / L1} = conj(W_{2^20}^{k1 << (20 - log2(2 L1))})
#define PGW_ROW_CASE(P1)                                                                                    \
  case P1:                                                                                                  \
    hipLaunchKernelGGL((k_fft_row_shift<T, P1>), dim3(pl.L2, nbatch), dim3(P1 / 16), 0, st, Z, pl.L2, pl.twr.p, hi, \
                       lo, shift1, nyqst);                                                                  \
    break;
  switch (pl.L1) {
    PGW_ROW_CASE(1024)
    PGW_ROW_CASE(2048)
    PGW_ROW_CASE(4096)
  }
#undef PGW_ROW_CASE
}

// Nl: the largest N of the launch's templates; register rows of the column pass that only see
// zero padding for that N are pruned (three variants: all E rows, E/2, 3E/8).
// tf: a threshold to fold into the launch (tf.nt = 0: none); its rows are set here.
// wide: workgroups of 512 threads (tiles of twice the columns, so each store instruction writes
// row segments of 512 B instead of 256 B) for the FP32 column pass with 8 register rows and the
// spectra in turn (elsewhere 256; FP64 and 16 register rows measured slower with 512)
template <typename T>
void launch_spin_col(const Plan<T>& pl, const SlotBases& sb, const SpinTab& tab, const TmplDesc* desc, cplx<T>* Z,
                     int nt, int Nl, const cplx<T>* hi, const cplx<T>* lo, bool seq, const ThrFuse<T>& tf,
                     hipStream_t st, bool rev = false, bool lf = false, bool nts = false, bool wide = false) {
  const SpinParams sp{pl.L1, pl.twshift, rev ? 1 : 0, nts ? 1 : 0};
  // loads-first demodulation: FP32 with 16 register rows (elsewhere the default kernel runs)
  const bool lf_ = lf && std::is_same<T, float>::value;
#define PGW_SCOL_NZ(P2, SQ, U, NZ)                                                                        \
  do {                                                                                                    \
    if constexpr (std::is_same<T, float>::value && ColE<T, P2>::E == 8 && SQ) {                          \
      if (wide && !lf_) {                                                                                 \
        const long gx_ = pl.L1 / ColShape<P2, ColE<T, P2>::E, 512>::C;                                    \
        ThrFuse<T> tf_ = tf;                                                                              \
        tf_.rows = tf.nt > 0 ? (int)(((long)tf.p.nparts * tf.nt + 8 * gx_ - 1) / (8 * gx_)) : 0;         \
        hipLaunchKernelGGL((k_spin_col2<T, P2, SQ, U, NZ, false, 512>), dim3(gx_, nt + tf_.rows), dim3(512), 0, \
                           st, sp, sb, tab, desc, Z, pl.twc.p, hi, lo, tf_);                              \
        break;                                                                                            \
      }                                                                                                   \
    }                                                                                                     \
    const long gx_ = pl.L1 / col_tile<T, P2>();                                                           \
    ThrFuse<T> tf_ = tf;                                                                                  \
    tf_.rows = tf.nt > 0 ? (int)(((long)tf.p.nparts * tf.nt + 4 * gx_ - 1) / (4 * gx_)) : 0;             \
    if (lf_)                                                                                              \
      hipLaunchKernelGGL((k_spin_col2<T, P2, SQ, U, NZ, std::is_same<T, float>::value && ColE<T, P2>::E == 16>), \
                         dim3(gx_, nt + tf_.rows), dim3(256), 0, st, sp, sb, tab, desc, Z, pl.twc.p, hi, lo, tf_); \
    else                                                                                                  \
      hipLaunchKernelGGL((k_spin_col2<T, P2, SQ, U, NZ>), dim3(gx_, nt + tf_.rows), dim3(256), 0, st, sp, sb, tab, \
                         desc, Z, pl.twc.p, hi, lo, tf_);                                                 \
  } while (0)
#define PGW_SCOL_GO(P2, SQ, U)                                                                           \
  do {                                                                                                   \
    constexpr int E_ = ColE<T, P2>::E;                                                                   \
    const long span = (long)pl.L1 * (P2 / E_); /* n of register row r starts at r * span */              \
    const long rows = (Nl + span - 1) / span;                                                            \
    if (rows <= 3 * E_ / 8) PGW_SCOL_NZ(P2, SQ, U, 3 * E_ / 8);                                          \
    else if (rows <= E_ / 2) PGW_SCOL_NZ(P2, SQ, U, E_ / 2);                                             \
    else PGW_SCOL_NZ(P2, SQ, U, E_);                                                                     \
  } while (0)
#define PGW_SCOL_CASE(P2)                                                                                \
  case P2:                                                                                               \
    if (seq || ColE<T, P2>::SEQ) {                                                                       \
      if (sb.uslot >= 0) PGW_SCOL_GO(P2, true, true); else PGW_SCOL_GO(P2, true, false);                 \
    } else {                                                                                             \
      if (sb.uslot >= 0) PGW_SCOL_GO(P2, false, true); else PGW_SCOL_GO(P2, false, false);               \
    }                                                                                                    \
    break;
  switch (pl.L2) {
    PGW_SCOL_CASE(16)
    PGW_SCOL_CASE(32)
    PGW_SCOL_CASE(64)
    PGW_SCOL_CASE(128)
    PGW_SCOL_CASE(256)
  }
#undef PGW_SCOL_CASE
#undef PGW_SCOL_GO
#undef PGW_SCOL_NZ
}

template <typename T>
void launch_spin_row(const Plan<T>& pl, const RowParams& rp, const double* aabb, const TmplDesc* desc, const cplx<T>* Z, T* Ft, T* psum,
                     T* pmax, int nt, int na, bool allow_half, int row8, hipStream_t st) {
  // nmax <= L/2 (always so with the reference's settings): the upper half of each row's outputs
  // holds no bin of [nmin, nmax)
  const bool half = allow_half && 2L * rp.nmax <= (long)pl.L1 * pl.L2;
#define PGW_SROW_GO(P1, NA, H)                                                                            \
  hipLaunchKernelGGL((k_spin_row2<T, P1, NA, H>), dim3(pl.L2, nt), dim3(P1 / 16), 0, st, rp, aabb, desc, Z, pl.twr.p, \
                     Ft, psum, pmax)
#define PGW_SROW_CASE(P1)                                                                                 \
  case P1:                                                                                                \
    if (na == 2) {                                                                                        \
      if (half) PGW_SROW_GO(P1, 2, true); else PGW_SROW_GO(P1, 2, false);                                 \
    } else {                                                                                              \
      if (half) PGW_SROW_GO(P1, 1, true); else PGW_SROW_GO(P1, 1, false);                                 \
    }                                                                                                     \
    break;
#define PGW_SROW8_GO(P1, H, PF, WL)                                                                      \
  hipLaunchKernelGGL((k_spin_row8<T, P1, H, PF, WL>), dim3(pl.L2, nt), dim3(P1 / 8), 0, st, rp, aabb, desc, Z, pl.twr.p, \
                     Ft, psum, pmax)
#define PGW_SROW8_H(P1, PF, WL)                                                                           \
  do {                                                                                                    \
    if (half) PGW_SROW8_GO(P1, true, PF, WL); else PGW_SROW8_GO(P1, false, PF, WL);                       \
  } while (0)
#define PGW_SROW8_CASE(P1)                                                                                \
  case P1:                                                                                                \
    if (row8 == 1) PGW_SROW8_H(P1, true, false);                                                          \
    else if (row8 == 2) PGW_SROW8_H(P1, false, false);                                                    \
    else if (row8 == 3) PGW_SROW8_H(P1, true, true);                                                      \
    else PGW_SROW8_H(P1, false, true);                                                                    \
    break;
  // 8 elements per thread: 1 second row prefetched, 2 not; 3 and 4 likewise with the
  // transforms' later exchanges inside each wave (row_fft8w)
  if (row8) {
    switch (pl.L1) {
      PGW_SROW8_CASE(1024)
      PGW_SROW8_CASE(2048)
      PGW_SROW8_CASE(4096)
    }
  } else {
    switch (pl.L1) {
      PGW_SROW_CASE(1024)
      PGW_SROW_CASE(2048)
      PGW_SROW_CASE(4096)
    }
  }
#undef PGW_SROW8_CASE
#undef PGW_SROW8_H
#undef PGW_SROW8_GO
#undef PGW_SROW_CASE
#undef PGW_SROW_GO
}

// The template stage of a launch as one persistent kernel (pgw_fused.hpp), FP32 form.  a: every
// field but the job layout and the shape-dependent counts.  Returns false (nothing launched) for
// shapes it is not built for.  Grid: the workgroups that fit the device at once (at most one per
// job).
template <int P2, int P1, int NZ>
void go_fused(FusedArgs& a, const SpinTab& tab, int ncu, hipStream_t st) {
  static int per_cu = 0;
  if (per_cu < 1) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spin_fused<P2, P1, NZ>, 256, 0) != hipSuccess || nb < 1)
      nb = 1;
    per_cu = nb;
  }
  const int grid = std::max(1, std::min(a.total, per_cu * ncu));
  hipLaunchKernelGGL((k_spin_fused<P2, P1, NZ>), dim3(grid), dim3(256), 0, st, a, tab);
}

bool launch_fused(int L1, int L2, int Nl, FusedArgs& a, const SpinTab& tab, int lag, int ncu, hipStream_t st) {
  const int nparts = a.tp.nparts;
#define PGW_FUSED_GO(P2, P1)                                                                  \
  do {                                                                                        \
    constexpr int E_ = ColE<float, P2>::E;                                                    \
    a.ct = P1 / ColShape<P2, E_>::C;                                                          \
    a.rj = P2 / FusedShape<P2, P1>::RJ;                                                       \
    a.tj = (nparts + 3) / 4;                                                                  \
    a.D = std::min(lag, a.nt / 2);                                                            \
    a.o1 = a.D * a.ct;                                                                        \
    a.o2 = a.o1 + a.D * (a.ct + a.rj);                                                        \
    a.o3 = a.o2 + (a.nt - 2 * a.D) * (a.ct + a.rj + a.tj);                                    \
    a.o4 = a.o3 + a.D * (a.rj + a.tj);                                                        \
    a.total = a.o4 + a.D * a.tj;                                                              \
    const long span = (long)P1 * (P2 / E_);                                                   \
    const long rows = (Nl + span - 1) / span;                                                 \
    if (rows <= 3 * E_ / 8) go_fused<P2, P1, 3 * E_ / 8>(a, tab, ncu, st);                    \
    else if (rows <= E_ / 2) go_fused<P2, P1, E_ / 2>(a, tab, ncu, st);                       \
    else go_fused<P2, P1, E_>(a, tab, ncu, st);                                               \
    return true;                                                                              \
  } while (0)
  if (L1 == 1024 && L2 == 16) PGW_FUSED_GO(16, 1024);
  if (L1 == 1024 && L2 == 32) PGW_FUSED_GO(32, 1024);
  if (L1 == 2048 && L2 == 32) PGW_FUSED_GO(32, 2048);
  if (L1 == 2048 && L2 == 64) PGW_FUSED_GO(64, 2048);
  if (L1 == 4096 && L2 == 64) PGW_FUSED_GO(64, 4096);
  if (L1 == 4096 && L2 == 128) PGW_FUSED_GO(128, 4096);
#undef PGW_FUSED_GO
  return false;
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

// Scratch of the allocation-free Fstatnet (pgw_fstatnet_noalloc): the (point, chunk) partial
// sums.  Every GPU the library runs on holds one copy until the process ends; all handles and
// calls on a GPU share it.  Host code reaches it only through the device address that
// scratch_address() obtains per device, never through the array's name.
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of that device's copy of `scratch`

// The address of the current device's `scratch`, looked up once per device.
int scratch_address(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return PGW_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PGW_ERR_DEVICE;
    it = g_scratch_ptr.emplace(dev, static_cast<double*>(p)).first;
  }
  *out = it->second;
  return PGW_OK;
}

}  // namespace

// ---------------------------------------------------------------- handle --
struct pgw_handle {
  virtual ~pgw_handle() {}
  // ssoff/sslist (optional): work item i searches spindowns sslist[ssoff[i] .. ssoff[i+1]) instead
  // of the range's
  // sgnlv: records of PGW_NPAR doubles, or of PGW_NPAR floats when rec_float (the reference's
  // FLOAT_TYPE = float build)
  virtual int run(const int* work, int n, pgw_range* r, const pgw_opts* o, int* sgnlc, void* sgnlv, int cap,
                  int* rec_work, int* FNum, const int* ssoff = nullptr, const int* sslist = nullptr,
                  bool rec_float = false) = 0;
  virtual int debug_fstat(int pm, int mm, int nn, int ss, const pgw_opts* o, double* F) = 0;
  virtual int debug_resampled(int pm, int mm, int nn, int det, double* a, double* b) = 0;
  virtual int time_stages(int seg, int pm, int mm, int nn, int s0, int nt, const pgw_opts* o, int reps,
                          double* ms) = 0;
  virtual int autotune(int seg, int pm, int mm, int nn, const pgw_opts* o, const pgw_range* r, int nsky, int* tb,
                       int* row_na, int* col_seq) = 0;
  virtual int tune_string(char* buf, int len) const = 0;
  virtual int kernel_times(int enable, double* out) = 0;
  virtual int fstatnet(int seg, const double* pts, int n, double* out, double* grad) = 0;
  virtual int fstatnet_noalloc(int seg, const double* d_pts, int n, double* d_out, double* d_grad) = 0;
  virtual int fstatnet_multi(const int* seg, const double* pts, int n, double* out, double* grad) = 0;
  virtual int settings(int seg, pgw_settings* out) const = 0;
  double last_ms = 0.0;
};

namespace {

struct SkyState {
  double het0, sgnlt2, sgnlt3;
};

template <typename T> struct Engine final : pgw_handle {
  using T2 = cplx<T>;

  struct Seg {
    pgw_settings sett;
    std::vector<double> amod, sig2, sphir, cphir, ssb0;
    DevBuf<double> xDat, DetSSB, cosmodf, sinmodf;
    Plan<T>*pf = nullptr, *pi = nullptr, *ps = nullptr;
    int nblk = 0, nparts = 0, npro = 0, nspl = 0;
  };

  // one sky point of a group
  struct Item {
    int work, seg, pm, mm, nn, slot, smin, smax, t0;
    SkyState sk;
    const int* ssl = nullptr;  // explicit spindown list (spotlight), else smin..smax
    int nlist = 0;
    int ns() const { return ssl ? nlist : smax - smin + 1; }
    int ss_at(int j) const { return ssl ? ssl[j] : smin + j; }
  };
  struct Group {
    std::vector<Item> items;
    int ntempl = 0, buf = 0;
    Cand* cbuf = nullptr;  // device candidate buffer, header first (default: cand + buf * (kCandCap + 1))
    int ccap = 0;          // its capacity in candidates (default: cand_cap)
  };

  std::vector<std::unique_ptr<Seg>> segs;
  std::map<int, std::unique_ptr<Plan<T>>> plans;
  int nifo = 0;
  size_t Nmax = 0, nfftmax = 0, Nimax = 0, Lmax = 0, L2max = 0, partsmax = 0, nsplmax = 0, npromax = 0;
  DevBuf<T2> hi, lo;
  // shared sky-stage scratch
  DevBuf<T2> R, Zf, Zi, Y;
  DevBuf<double2> y2;
  DevBuf<double> shftf, paa, pbb;
  DevBuf<SplineTab> spl;
  DevBuf<SkyItem> sky_dev;  // 2 x kSlots
  PinBuf<SkyItem> sky_pin;  // 2 x kSlots
  // sky slots
  DevBuf<T2> xm_slots;
  DevBuf<double> tmp1_slots, aabb_slots;
  size_t slot_xm = 0, slot_tmp1 = 0;
  // template stage
  DevBuf<T2> Z;
  DevBuf<T> Ft, psum, pmax;
  DevBuf<double> mu;
  DevBuf<TmplDesc> desc;  // 2 x kDescCap
  PinBuf<TmplDesc> desc_pin;
  DevBuf<Cand> cand;      // 2 x (1 + kCandCap): header, candidates
  PinBuf<Cand> hc_pin;    // 2 x (1 + kCandChunk)
  std::vector<Cand> hcand;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evd[2] = {nullptr, nullptr};
  int TB = 128, TBalloc = 0;
  int cand_cap = kCandCap;  // candidates per group in the device buffer (PGW_TUNE candcap= lowers it)
  int l1max = 0;            // FFT row length cap (0: the split_len table)
  // FP32 form: demodulation phase in revolutions (PhaseRev; the batching column pass 1 % faster,
  // identical candidates); rev=0 in PGW_TUNE restores the radians form
  bool phase_rev = std::is_same<T, float>::value;
  bool col_lf = false;      // FP32 column pass: loads-first demodulation (autotuned)
  // FP32 column pass with 8 register rows: 512-thread workgroups (column pass 2-9 % faster on the
  // batching and network workloads)
  bool col_wide = true;
  // column pass: spectra written with non-temporal stores (autotuned: faster where a launch's
  // spectra outgrow the last-level cache, slower where the row pass would find them there)
  bool col_nt = false;
  // row pass: spectra read with non-temporal loads (each line is read once; faster or equal in
  // every measured shape, e.g. batching FP64 row pass 0.121 -> 0.100 ms per 128 templates)
  bool row_nt = true;
  // row pass: F written with non-temporal stores (the threshold reads F only in flagged blocks):
  // +0.3-1 % in every measured shape
  bool f_nt = true;
  int group = kSlots;  // sky points per group
  int row_na = 1;
  bool col_seq = true;  // column pass: the two spectra share one LDS buffer in turn
  int spl_k = 4;  // spline outputs per thread (4, 8 or 16)
  bool spl_fuse = true;  // untranspose, spline and its evaluation in one kernel (k_spline_fused)
  bool row_half = true;  // row pass forms only the outputs that can fall in [nmin, nmax)
  // row pass with 8 elements per thread (1 with the second row prefetched, 2 without; 3 and 4
  // likewise with the transforms' later exchanges inside each wave): the FP64 default, 3 (two
  // workgroups per CU instead of one, 6 workgroup barriers per row instead of 14); FP32 keeps 16
  // elements per thread
  int row8 = std::is_same<T, double>::value ? 3 : 0;
  // each launch's threshold folded into the next launch's column pass (the group's last one runs
  // on its own); thr_prev: the threshold still to run (nt = 0: none)
  bool thr_fuse = true;
  ThrFuse<T> thr_prev{};
  // a group's launch sequence (sky stage, copies, template launches) captured into a HIP graph and
  // launched as one; one executable graph per buffer set, updated in place while the sequence keeps
  // its shape
  bool graph = false;
  bool capturing = false;
  // the template stage of a launch as one persistent work-queue kernel (FP32 form; the separate
  // kernels elsewhere); lag: stages between a template's column jobs and its row jobs
  bool fused = false;
  int fused_lag = 8;
  int fused_spin = 1 << 20;  // polls of a dependency counter before the fused kernel gives up (~1 s)
  DevBuf<FusedSync> fsync;
  int ncu = 1;  // compute units of the handle's device
  hipGraphExec_t gexec[2] = {nullptr, nullptr};

  // kernel timing over searches: 4 events per template launch (column, row, threshold; with the
  // threshold folded into the next column pass the last interval is empty)
  int ktime = 0;         // record every ktime-th template launch (0: off)
  long kseq = 0;         // template launches since recording was switched on
  std::vector<hipEvent_t> kpool;
  std::vector<int> knt;  // templates of each timed launch
  size_t kused = 0;
  double kacc[3] = {0, 0, 0};
  double klaunch = 0, ktempl = 0;

  hipEvent_t* kquad(int nt) {
    if (4 * (kused + 1) > kpool.size()) {
      for (int k = 0; k < 4; k++) {
        hipEvent_t e = nullptr;
        // timing only: no system-scope fence (its cache writeback and invalidation would slow
        // the following kernels and enter the measurement)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
        kpool.push_back(e);
      }
      knt.push_back(0);
    }
    knt[kused] = nt;
    return &kpool[4 * kused++];
  }
  // after the stream has drained: add the recorded launches to the totals
  void kcollect() {
    for (size_t q = 0; q < kused; q++) {
      for (int k = 0; k < 3; k++) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, kpool[4 * q + k], kpool[4 * q + k + 1]) == hipSuccess) kacc[k] += t;
      }
      klaunch += 1;
      ktempl += knt[q];
    }
    kused = 0;
  }
  int kernel_times(int enable, double* out) override {
    HIPCHK(hipStreamSynchronize(stream));
    kcollect();
    if (out) {
      for (int k = 0; k < 3; k++) out[k] = kacc[k];
      out[3] = klaunch;
      out[4] = ktempl;
    }
    kacc[0] = kacc[1] = kacc[2] = 0;
    klaunch = ktempl = 0;
    ktime = enable > 0 ? enable : 0;
    kseq = 0;
    return PGW_OK;
  }

  ~Engine() override {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipGraphExec_t& x : gexec)
      if (x) (void)hipGraphExecDestroy(x);
    for (hipEvent_t e : kpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ev0, ev1, evd[0], evd[1]})
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  Plan<T>* plan(int L, int& st) {
    auto it = plans.find(L);
    if (it != plans.end()) return it->second.get();
    auto p = std::make_unique<Plan<T>>();
    st = p->init(L, l1max);
    if (st) return nullptr;
    Plan<T>* q = p.get();
    plans[L] = std::move(p);
    return q;
  }

  int add_segment(const pgw_settings* s, const pgw_detector* ifo) {
    auto sg = std::make_unique<Seg>();
    Seg& g = *sg;
    g.sett = *s;
    g.sett.nifo = nifo;
    const pgw_settings& S = g.sett;
    const int N = S.N;
    if (S.Ninterp != S.interpftpad * S.nfft || S.nfftf != S.fftpad * S.nfft) return PGW_ERR_ARG;
    if (S.interpftpad * N > S.Ninterp || N > S.nfft || S.nmin < 1 || S.nmax > S.nfftf || S.nmax <= S.nmin ||
        !(S.oms > 0) || S.numlines_band < 0 || S.numlines_band > PGW_MAXL)
      return PGW_ERR_ARG;
    if (S.Ninterp - S.nfft != S.nfft) return PGW_ERR_UNSUPPORTED;  // zero-insert map assumes interpftpad = 2
    const int M = S.nmax - S.nmin;
    g.nblk = M / PGW_NAV;
    g.nparts = g.nblk + (M % PGW_NAV ? 1 : 0);
    if (g.nparts > 256) return PGW_ERR_UNSUPPORTED;
    g.nspl = S.interpftpad * N;
    g.npro = (S.nfft + NT - 1) / NT;
    int st = PGW_OK;
    if (!(g.pf = plan(S.nfft, st)) || !(g.pi = plan(S.Ninterp, st)) || !(g.ps = plan(S.nfftf, st)))
      return st ? st : PGW_ERR_UNSUPPORTED;
    g.amod.resize(9 * nifo);
    g.sig2.resize(nifo);
    g.sphir.resize(nifo);
    g.cphir.resize(nifo);
    g.ssb0.resize(3 * nifo);
    std::vector<double> hx((size_t)nifo * N), hd((size_t)nifo * 3 * N);
    for (int d = 0; d < nifo; d++) {
      if (!ifo[d].xDat || !ifo[d].DetSSB || !(ifo[d].sig2 > 0)) return PGW_ERR_ARG;
      memcpy(&g.amod[9 * d], ifo[d].amod, 9 * sizeof(double));
      g.sig2[d] = ifo[d].sig2;
      g.sphir[d] = ifo[d].sphir;
      g.cphir[d] = ifo[d].cphir;
      memcpy(&g.ssb0[3 * d], ifo[d].DetSSB, 3 * sizeof(double));
      memcpy(&hx[(size_t)d * N], ifo[d].xDat, sizeof(double) * N);
      memcpy(&hd[(size_t)d * 3 * N], ifo[d].DetSSB, sizeof(double) * 3 * N);
    }
    // Earth-rotation auxiliary arrays, init.c:448-455 (host glibc, same values as the CPU path)
    std::vector<double> hc(N), hs(N);
    for (int i = 0; i < N; i++) {
      const double omrt = S.omr * i;
      hc[i] = cos(omrt);
      hs[i] = sin(omrt);
    }
    if ((st = g.xDat.alloc(hx.size())) || (st = g.DetSSB.alloc(hd.size())) || (st = g.cosmodf.alloc(N)) ||
        (st = g.sinmodf.alloc(N)))
      return st;
    HIPCHK(hipMemcpy(g.xDat.p, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.DetSSB.p, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.cosmodf.p, hc.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.sinmodf.p, hs.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    Nmax = std::max<size_t>(Nmax, N);
    nfftmax = std::max<size_t>(nfftmax, S.nfft);
    Nimax = std::max<size_t>(Nimax, S.Ninterp);
    Lmax = std::max<size_t>(Lmax, S.nfftf);
    L2max = std::max<size_t>(L2max, g.ps->L2);
    partsmax = std::max<size_t>(partsmax, g.nparts);
    nsplmax = std::max<size_t>(nsplmax, g.nspl);
    npromax = std::max<size_t>(npromax, g.npro);
    segs.push_back(std::move(sg));
    return PGW_OK;
  }

  // template-stage buffers for `tb` templates per launch
  int ensure_tb(int tb) {
    if (tb <= TBalloc) return PGW_OK;
    if (stream) HIPCHK(hipStreamSynchronize(stream));
    int st;
    if ((st = Z.alloc((size_t)tb * 2 * Lmax)) || (st = Ft.alloc((size_t)tb * Lmax)) ||
        (st = psum.alloc((size_t)tb * partsmax * L2max)) || (st = pmax.alloc((size_t)tb * partsmax * L2max)) ||
        (st = mu.alloc((size_t)tb * partsmax)))
      return st;
    TBalloc = tb;
    return PGW_OK;
  }

  int init(const pgw_settings* s, const pgw_detector* ifo, int nseg, int ndet) {
    nifo = ndet;
    int st;
    // kernel variants; PGW_TUNE="row_na=2,col_seq=1,tb=64,group=4,spl_k=4,half=1" overrides (used when
    // measuring variants)
    if (const char* tune = getenv("PGW_TUNE")) {
      int v;
      if (const char* q = strstr(tune, "row_na=")) if (sscanf(q, "row_na=%d", &v) == 1 && (v == 1 || v == 2)) row_na = v;
      if (const char* q = strstr(tune, "col_seq=")) if (sscanf(q, "col_seq=%d", &v) == 1) col_seq = v != 0;
      if (const char* q = strstr(tune, "tb=")) if (sscanf(q, "tb=%d", &v) == 1 && v >= 1 && v <= kTBMax) TB = v;
      if (const char* q = strstr(tune, "group=")) if (sscanf(q, "group=%d", &v) == 1 && v >= 1 && v <= kSlots) group = v;
      if (const char* q = strstr(tune, "spl_k=")) if (sscanf(q, "spl_k=%d", &v) == 1 && (v == 4 || v == 8 || v == 16)) spl_k = v;
      if (const char* q = strstr(tune, "spl_fuse=")) if (sscanf(q, "spl_fuse=%d", &v) == 1) spl_fuse = v != 0;
      if (const char* q = strstr(tune, "half=")) if (sscanf(q, "half=%d", &v) == 1) row_half = v != 0;
      if (const char* q = strstr(tune, "row8=")) if (sscanf(q, "row8=%d", &v) == 1 && v >= 0 && v <= 4) row8 = v;
      if (const char* q = strstr(tune, "thr_fuse=")) if (sscanf(q, "thr_fuse=%d", &v) == 1) thr_fuse = v != 0;
      if (const char* q = strstr(tune, "candcap=")) if (sscanf(q, "candcap=%d", &v) == 1 && v >= 1 && v <= kCandCap) cand_cap = v;
      if (const char* q = strstr(tune, "l1max=")) if (sscanf(q, "l1max=%d", &v) == 1 && (v == 1024 || v == 2048 || v == 4096)) l1max = v;
      if (const char* q = strstr(tune, "rev=")) if (sscanf(q, "rev=%d", &v) == 1) phase_rev = v != 0;
      if (const char* q = strstr(tune, "col_lf=")) if (sscanf(q, "col_lf=%d", &v) == 1) col_lf = v != 0;
      if (const char* q = strstr(tune, "col_nt=")) if (sscanf(q, "col_nt=%d", &v) == 1) col_nt = v != 0;
      if (const char* q = strstr(tune, "row_nt=")) if (sscanf(q, "row_nt=%d", &v) == 1) row_nt = v != 0;
      if (const char* q = strstr(tune, "f_nt=")) if (sscanf(q, "f_nt=%d", &v) == 1) f_nt = v != 0;
      if (const char* q = strstr(tune, "graph=")) if (sscanf(q, "graph=%d", &v) == 1) graph = v != 0;
      if (const char* q = strstr(tune, "col_wide=")) if (sscanf(q, "col_wide=%d", &v) == 1) col_wide = v != 0;
      if (const char* q = strstr(tune, "fused=")) if (sscanf(q, "fused=%d", &v) == 1) fused = v != 0;
      if (const char* q = strstr(tune, "flag=")) if (sscanf(q, "flag=%d", &v) == 1 && v >= 0 && v <= 64) fused_lag = v;
      if (const char* q = strstr(tune, "fspin=")) if (sscanf(q, "fspin=%d", &v) == 1 && v >= 1 && v <= 24) fused_spin = 1 << v;
    }
    // segments after the variants: their FFT plans follow l1max
    for (int k = 0; k < nseg; k++)
      if ((st = add_segment(&s[k], ifo + (size_t)k * ndet))) return st;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // this device's copy of the allocation-free Fstatnet's scratch (the device is current here)
    if ((st = scratch_address(&fsn_scratch))) return st;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreateWithFlags(&evd[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&evd[1], hipEventDisableTiming));
    // W_{2^20}^m two-level tables: hi[h] = W^{1024 h}, lo[l] = W^l
    if ((st = upload_tw<T>(hi, 1 << 20, 10, 1024)) || (st = upload_tw<T>(lo, 1 << 20, 0, 1024))) return st;
    slot_xm = (size_t)2 * nifo * Nmax;
    slot_tmp1 = (size_t)nifo * Nmax;
    // sky-stage scratch for a group of up to kSlots sky points
    const size_t K = kSlots;
    if ((st = R.alloc(K * 2 * nifo * nfftmax)) || (st = Zf.alloc(K * 2 * nifo * nfftmax)) ||
        (st = Zi.alloc(K * 2 * nifo * Nimax)) || (st = Y.alloc(K * 2 * nifo * Nimax)) ||
        (st = y2.alloc(K * 2 * nifo * (nsplmax + 1))) ||
        (st = shftf.alloc(K * nifo * Nmax)) || (st = paa.alloc(K * nifo * npromax)) ||
        (st = pbb.alloc(K * nifo * npromax)) || (st = spl.alloc(1)) || (st = sky_dev.alloc(2 * K)) ||
        (st = sky_pin.alloc(2 * K)) ||
        (st = xm_slots.alloc((size_t)kSlots * slot_xm)) || (st = tmp1_slots.alloc((size_t)kSlots * slot_tmp1)) ||
        (st = aabb_slots.alloc((size_t)kSlots * 2)) || (st = desc.alloc((size_t)2 * kDescCap)) ||
        (st = desc_pin.alloc((size_t)2 * kDescCap)) || (st = cand.alloc((size_t)2 * (kCandCap + 1))) ||
        (st = hc_pin.alloc((size_t)2 * (kCandChunk + 1))) ||
        (st = fsync.alloc(1)) ||
        (st = ensure_tb(TB)))
      return st;
    HIPCHK(hipMemset(fsync.p, 0, sizeof(FusedSync)));  // the fused kernel's counters start at zero
    {
      int dev = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
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

  T2* slot_xm_ptr(int slot) { return xm_slots.p + (size_t)slot * slot_xm; }
  double* slot_tmp1_ptr(int slot) { return tmp1_slots.p + (size_t)slot * slot_tmp1; }
  double* slot_aabb_ptr(int slot) { return aabb_slots.p + (size_t)slot * 2; }

  // Sky coordinates of a grid point (jobcore.c:243-264); 0 when outside the disk.
  int sky_coords(const Seg& g, int pm, int mm, int nn, SkyState& ss, double& sinalt, double& cosalt, double& sindelt,
                 double& cosdelt) {
#pragma clang fp contract(off)
    const double* M = g.sett.M;
    const double al1 = nn * M[10] + mm * M[14];
    const double al2 = nn * M[11] + mm * M[15];
    if ((al1 * al1 + al2 * al2) / (g.sett.oms * g.sett.oms) > 1.) return 0;
    lin2ast(al1 / g.sett.oms, al2 / g.sett.oms, pm, g.sett.sepsm, g.sett.cepsm, &sinalt, &cosalt, &sindelt, &cosdelt);
    ss.sgnlt2 = asin(sindelt);
    ss.sgnlt3 = fmod(atan2(sinalt, cosalt) + 2. * M_PI, 2. * M_PI);
    ss.het0 = fmod(nn * M[8] + mm * M[12], M[0]);
    return 1;
  }

  // Coefficients of one sky point's stage (jobcore.c:270-295); 0 when outside the disk.
  int sky_item(const Seg& g, int pm, int mm, int nn, int slot, SkyState& ss, SkyItem& it) {
#pragma clang fp contract(off)
    double sinalt, cosalt, sindelt, cosdelt;
    if (!sky_coords(g, pm, mm, nn, ss, sinalt, cosalt, sindelt, cosdelt)) return 0;
    const pgw_settings& S = g.sett;
    memset(&it, 0, sizeof(it));
    SkyParams& p = it.p;
    p.het0 = ss.het0;
    p.oms = S.oms;
    p.N = S.N;
    p.nfft = S.nfft;
    p.nifo = nifo;
    for (int d = 0; d < nifo; d++) {
      SkyDet& q = p.det[d];
      memcpy(q.c, &g.amod[9 * d], sizeof(q.c));
      q.cosalfr = cosalt * g.cphir[d] + sinalt * g.sphir[d];
      q.sinalfr = sinalt * g.cphir[d] - cosalt * g.sphir[d];
      q.c2d = cosdelt * cosdelt;
      q.c2sd = sindelt * cosdelt;
      q.sindel = sindelt;
      q.cosdel = cosdelt;
      q.nS0 = cosalt * cosdelt;
      q.nS1 = sinalt * cosdelt;
      q.nS2 = sindelt;
      q.shft1 = q.nS0 * g.ssb0[3 * d] + q.nS1 * g.ssb0[3 * d + 1] + q.nS2 * g.ssb0[3 * d + 2];
      q.sig2 = g.sig2[d];
    }
    it.xDat = g.xDat.p;
    it.DetSSB = g.DetSSB.p;
    it.cosmodf = g.cosmodf.p;
    it.sinmodf = g.sinmodf.p;
    it.xm = slot_xm_ptr(slot);
    it.tmp1 = slot_tmp1_ptr(slot);
    it.aabb = slot_aabb_ptr(slot);
    it.nspl = g.nspl;
    return 1;
  }

  // Sky-point stage (jobcore.c:270-393) of G sky points at once: items[j] (host, prepared by
  // sky_item) are staged through pinned buffer `buf`; every kernel covers all G sky points.
  // Segments of one batch share nfft and Ninterp (compatible()).
  // zero (optional): the group's candidate header, cleared by the first kernel
  int sky_batch(const Seg& g0, const SkyItem* items, int G, int buf, int* zero = nullptr) {
    const pgw_settings& S = g0.sett;
    SkyItem* hp = sky_pin.p + (size_t)buf * kSlots;
    memcpy(hp, items, sizeof(SkyItem) * G);
    const SkyItem* dp = sky_dev.p + (size_t)buf * kSlots;
    HIPCHK(hipMemcpyAsync((void*)dp, hp, sizeof(SkyItem) * G, hipMemcpyHostToDevice, stream));
    int Ng = 0, nsplg = 0;
    for (int j = 0; j < G; j++) {
      Ng = std::max(Ng, items[j].p.N);
      nsplg = std::max(nsplg, items[j].nspl);
    }
    const int nb = 2 * nifo;
    hipLaunchKernelGGL((k_sky_prologue<T>), dim3(g0.npro, nifo, G), dim3(NT), 0, stream, dp, R.p, shftf.p, Nmax,
                       paa.p, pbb.p, zero);
    // resampling: forward FFT (nfft), zero-insert, backward FFT (Ninterp), scale (jobcore.c:319-362)
    LoadPlain<T2> lp{R.p, S.nfft, S.nfft};
    launch_col<T, -1>(*g0.pf, lp, Zf.p, nb * G, hi.p, lo.p, stream);
    const T ft = (T)((double)S.interpftpad / S.Ninterp);
    if (spl_fuse) {
      // odd samples by an nfft-point inverse transform (even ones: R itself): forward row
      // transforms, phase shift and inverse row transforms in one pass, then the column pass with
      // input twiddles (natural-order output); then untranspose, spline and evaluation in one kernel
      // (y and y2 stay in LDS)
      launch_row_shift<T>(*g0.pf, Zf.p, nb * G, hi.p, lo.p, S.nfft / 2 + 1, stream);
      LoadOddTw<T2> lw{Zf.p, hi.p, lo.p, S.nfft, ilog2(g0.pf->L1), g0.pi->twshift};
      launch_col<T, +1, false>(*g0.pf, lw, Zi.p, nb * G, hi.p, lo.p, stream);
      int lg = 0;
      while ((1 << lg) < g0.pf->L2) lg++;
      int* err = zero ? zero + 1 : nullptr;
#define PGW_SPLINE_F(K)                                                                                          \
  hipLaunchKernelGGL((k_spline_fused<T, K, true>), dim3((nsplg + spf_out(K) - 1) / spf_out(K), nb * G), dim3(NT), \
                     0, stream, Zi.p, g0.pf->L1, lg, ft, (size_t)S.nfft, R.p, (size_t)S.nfft, spl.p, shftf.p,    \
                     Nmax, S.interpftpad, dp, nb, paa.p, pbb.p, g0.npro, err)
      if (spl_k == 4) PGW_SPLINE_F(4);
      else if (spl_k == 16) PGW_SPLINE_F(16);
      else PGW_SPLINE_F(8);
#undef PGW_SPLINE_F
      return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
    }
    launch_row<T, -1>(*g0.pf, Zf.p, nb * G, stream);
    LoadZeroInsert<T2> lz{Zf.p, S.nfft, S.Ninterp, S.nfft / 2 + 1, g0.pf->L1, g0.pf->L2};
    launch_col<T, +1>(*g0.pi, lz, Zi.p, nb * G, hi.p, lo.p, stream);
    launch_row<T, +1>(*g0.pi, Zi.p, nb * G, stream);
    const int count = std::min(S.Ninterp, nsplg);
    const int rows_k1 = std::min(g0.pi->L1, (count + g0.pi->L2 - 1) / g0.pi->L2);
    dim3 gt((rows_k1 + 31) / 32, (g0.pi->L2 + 31) / 32, nb * G);
    hipLaunchKernelGGL((k_untranspose<T>), gt, dim3(NT), 0, stream, Zi.p, Y.p, g0.pi->L1, g0.pi->L2, count, ft);
    // spline on the first interpftpad*N samples, evaluated at interpftpad*(i - shftf[i]) (auxi.c:98-122)
#define PGW_SPLINE(K)                                                                                           \
  hipLaunchKernelGGL((k_spline<T, K>), dim3((nsplg + spf_out(K) - 1) / spf_out(K), nb * G), dim3(NT), 0, stream, \
                     Y.p, (size_t)S.Ninterp, spl.p, y2.p, (size_t)(nsplmax + 1), dp, nb, paa.p, pbb.p, g0.npro)
    if (spl_k == 4) PGW_SPLINE(4);
    else if (spl_k == 16) PGW_SPLINE(16);
    else PGW_SPLINE(8);
#undef PGW_SPLINE
    dim3 ge((Ng + NT - 1) / NT, nb * G);
    hipLaunchKernelGGL((k_spline_eval<T, T>), ge, dim3(NT), 0, stream, Y.p, (size_t)S.Ninterp, y2.p,
                       (size_t)(nsplmax + 1), shftf.p, Nmax, S.interpftpad, dp, nb);
    return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
  }

  // One sky point's stage into slot `slot` (diagnostics and stage timing).  Returns 1 when
  // inside the disk; status receives errors.
  int sky_stage(const Seg& g, int pm, int mm, int nn, int slot, SkyState& ss, int& status) {
    SkyItem it;
    status = PGW_OK;
    if (!sky_item(g, pm, mm, nn, slot, ss, it)) return 0;
    HIPCHK(hipStreamSynchronize(stream));  // the pinned staging buffer may still feed a copy
    status = sky_batch(g, &it, 1, 0);
    return 1;
  }

  void spindown_range(const Seg& g, int mm, int nn, const pgw_range* r, const pgw_opts* o, int& smin, int& smax) {
#pragma clang fp contract(off)
    const double* M = g.sett.M;
    smin = r->sst;
    smax = r->spndr[1];
    if (!o->use_range) {  // jobcore.c:424-439
      smin = (int)trunc((g.sett.Smin - nn * M[9] - mm * M[13]) / M[5]);
      smax = (int)trunc(-(nn * M[9] + mm * M[13] + g.sett.Smax) / M[5]);
      if (smin > smax) std::swap(smin, smax);
    }
    if (o->s0_flag) smin = smax;
  }

  static SpinGrid spin_grid(const double* M, int mm, int nn, int ss0, int s0_flag) {
    return SpinGrid{M[0], M[4], M[5], M[9], M[13], mm, nn, ss0, s0_flag};
  }
  static void tmpl_params(const double* M, int mm, int nn, int ss, int s0_flag, double& het1, double& sgnlt1) {
    spin_coeffs(spin_grid(M, mm, nn, 0, s0_flag), ss, het1, sgnlt1);
  }

  // descriptors of item `it`'s templates into host buffer `h` (index t0..)
  void fill_desc(const Item& it, const pgw_opts* o, TmplDesc* h) {
    const Seg& g = *segs[it.seg];
    for (int j = 0; j < it.ns(); j++) {
      const int ss = it.ss_at(j);
      TmplDesc& d = h[it.t0 + j];
      d.slot = it.slot;
      tmpl_params(g.sett.M, it.mm, it.nn, ss, o->s0_flag, d.het1, d.sgnlt1);
      d.N = g.sett.N;
      d.nifo = nifo;
      d.pad = 0;
    }
  }

  // runs a threshold left pending by a folded launch
  int flush_thr() {
    if (thr_prev.nt > 0) {
      const ThrFuse<T>& f = thr_prev;
      hipLaunchKernelGGL((k_threshold<T>), dim3(f.p.nparts, f.nt), dim3(64), 0, stream, f.p, f.Ft, f.psum, f.pmax,
                         f.mu, f.cand, f.ncand);
      thr_prev.nt = 0;
    }
    return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
  }

  // template stage for descriptors [t0, t0 + nt) of buffer buf; ev (optional) = 4 events around the kernels
  // uni: all nt templates belong to one item (its templates j0, j0 + 1, ...), else nullptr
  // fold: leave this launch's threshold pending, to run inside the next column launch (or flush_thr)
  int launch_templates(const Seg& g, int buf, Cand* cb, int ccap, int t0, int nt, const Item* uni, int j0, int Nl,
                       const pgw_opts* o, hipEvent_t* ev = nullptr, bool fold = false) {
    const TmplDesc* d = desc.p + (size_t)buf * kDescCap + t0;
    if (ev) HIPCHK(hipEventRecord(ev[0], stream));
    SlotBases sb{xm_slots.p, tmp1_slots.p, aabb_slots.p, slot_xm, slot_tmp1, nifo, -1, 0, 0};
    SpinTab tab;
    if (uni) {
      sb.uslot = uni->slot;
      sb.uN = segs[uni->seg]->sett.N;
      if (nt <= kSpinTab) {  // else the kernels read het1, sgnlt1 from the descriptors
        sb.tab = 1;
        const SpinGrid q = spin_grid(segs[uni->seg]->sett.M, uni->mm, uni->nn, 0, o->s0_flag);
        for (int i = 0; i < nt; i++) spin_coeffs(q, uni->ss_at(j0 + i), tab.het1[i], tab.sgnlt1[i]);
      }
    } else if (nt <= kSpinTab) {
      // templates of several sky points: their values from the host copy of the descriptors
      const TmplDesc* hd = desc_pin.p + (size_t)buf * kDescCap + t0;
      sb.tab = 1;
      for (int i = 0; i < kSpinSlots; i++) tab.slotN[i] = 0;
      for (int i = 0; i < nt; i++) {
        tab.het1[i] = hd[i].het1;
        tab.sgnlt1[i] = hd[i].sgnlt1;
        tab.slot[i] = (unsigned char)hd[i].slot;
        tab.slotN[hd[i].slot] = hd[i].N;
      }
    }
    RowParams rp{g.ps->L2, g.sett.nmin, g.sett.nmax, g.nblk, g.nparts, ilog2(g.ps->L2), row_nt ? 1 : 0, f_nt ? 1 : 0,
                 sb.uslot, sb.tab, {}};
    if (sb.uslot < 0 && sb.tab)
      for (int i = 0; i < nt; i++) rp.slot[i] = tab.slot[i];
    ThrParams tp{g.ps->L1, g.ps->L2, g.sett.nmin, g.sett.nmax, g.nblk, g.nparts, o->white_flag, t0, ccap, ilog2(g.ps->L2), o->trl};
    if constexpr (std::is_same<T, float>::value) {
      // one persistent kernel for the whole template stage (its threshold included)
      if (fused && phase_rev && !col_lf && row_half && row_na == 1 && nt <= kFusedTB &&
          2L * g.sett.nmax <= (long)g.ps->L1 * g.ps->L2) {
        if (int s = flush_thr()) return s;
        FusedArgs fa{};
        fa.sp = SpinParams{g.ps->L1, g.ps->twshift, 1, col_nt ? 1 : 0};
        fa.sb = sb;
        fa.rp = rp;
        fa.tp = tp;
        fa.desc = d;
        fa.Z = Z.p;
        fa.twc = g.ps->twc.p;
        fa.twr = g.ps->twr.p;
        fa.hi = hi.p;
        fa.lo = lo.p;
        fa.Ft = Ft.p;
        fa.psum = psum.p;
        fa.pmax = pmax.p;
        fa.mu = mu.p;
        fa.cand = cb + 1;
        fa.ncand = &cb->t;
        fa.err = &cb->ii;
        fa.sy = fsync.p;
        fa.nt = nt;
        fa.spin_limit = fused_spin;
        if (launch_fused(g.ps->L1, g.ps->L2, uni ? segs[uni->seg]->sett.N : Nl, fa, tab, fused_lag, ncu, stream)) {
          if (ev) {
            HIPCHK(hipEventRecord(ev[1], stream));
            HIPCHK(hipEventRecord(ev[2], stream));
            HIPCHK(hipEventRecord(ev[3], stream));
          }
          return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
        }
      }
    }
    // the pending threshold reads F and the FStat partials of the previous launch, which the row
    // pass below overwrites only after this column launch has finished
    ThrFuse<T> tf{};
    if (thr_prev.nt > 0) {
      if (fold) tf = thr_prev;
      else if (int s = flush_thr()) return s;
      thr_prev.nt = 0;
    }
    launch_spin_col<T>(*g.ps, sb, tab, d, Z.p, nt, uni ? segs[uni->seg]->sett.N : Nl, hi.p, lo.p, col_seq, tf,
                       stream, phase_rev, col_lf, col_nt, col_wide);
    if (ev) HIPCHK(hipEventRecord(ev[1], stream));
    launch_spin_row<T>(*g.ps, rp, aabb_slots.p, d, Z.p, Ft.p, psum.p, pmax.p, nt, row_na, row_half, row8, stream);
    if (ev) HIPCHK(hipEventRecord(ev[2], stream));
    if (fold) {
      thr_prev = ThrFuse<T>{tp, Ft.p, psum.p, pmax.p, mu.p, cb + 1, &cb->t, nt, 0};
    } else {
      hipLaunchKernelGGL((k_threshold<T>), dim3(g.nparts, nt), dim3(64), 0, stream, tp, Ft.p, psum.p, pmax.p, mu.p,
                         cb + 1, &cb->t);
    }
    if (ev) HIPCHK(hipEventRecord(ev[3], stream));
    return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
  }

  bool compatible(const Seg& a, const Seg& b) const {
    return a.sett.nfftf == b.sett.nfftf && a.sett.nmin == b.sett.nmin && a.sett.nmax == b.sett.nmax &&
           a.sett.fftpad == b.sett.fftpad && a.sett.nfft == b.sett.nfft;
  }

  // Queues a group: sky stages into slots, descriptors, template launches, async candidate copy,
  // and (rec_ev) the event finish_group waits on.
  int launch_group(Group& gr, const pgw_opts* o, bool rec_ev = true) {
    int st = PGW_OK;
    if (!gr.cbuf) {
      gr.cbuf = cand.p + (size_t)gr.buf * (kCandCap + 1);
      gr.ccap = cand_cap;
    }
    SkyItem si[kSlots];
    for (Item& it : gr.items)
      if (!sky_item(*segs[it.seg], it.pm, it.mm, it.nn, it.slot, it.sk, si[it.slot])) return PGW_ERR_DEVICE;
    if ((st = sky_batch(*segs[gr.items[0].seg], si, (int)gr.items.size(), gr.buf, &gr.cbuf->t))) return st;
    TmplDesc* hd = desc_pin.p + (size_t)gr.buf * kDescCap;
    for (const Item& it : gr.items) fill_desc(it, o, hd);
    // launches of at most kSpinTab templates take their templates' values from kernel arguments;
    // only larger launches and the persistent template-stage kernel read descriptors on the device
    if (TB > kSpinTab || fused)
      HIPCHK(hipMemcpyAsync(desc.p + (size_t)gr.buf * kDescCap, hd, sizeof(TmplDesc) * gr.ntempl,
                            hipMemcpyHostToDevice, stream));
    const Seg& g0 = *segs[gr.items[0].seg];
    int Ng = 0;  // largest N of the group
    for (const Item& it : gr.items) Ng = std::max(Ng, segs[it.seg]->sett.N);
    size_t k = 0;
    for (int t0 = 0; t0 < gr.ntempl; t0 += TB) {
      const int nt = std::min(TB, gr.ntempl - t0);
      while (k + 1 < gr.items.size() && gr.items[k + 1].t0 <= t0) k++;
      const Item& it = gr.items[k];
      const bool one = t0 + nt <= it.t0 + it.ns();
      const bool timed = !capturing && ktime && kseq++ % ktime == 0;
      if ((st = launch_templates(g0, gr.buf, gr.cbuf, gr.ccap, t0, nt, one ? &it : nullptr, t0 - it.t0, Ng, o,
                                 timed ? kquad(nt) : nullptr, thr_fuse))) {
        thr_prev.nt = 0;
        return st;
      }
    }
    if ((st = flush_thr())) return st;
    HIPCHK(hipMemcpyAsync(hc_pin.p + (size_t)gr.buf * (kCandChunk + 1), gr.cbuf, sizeof(Cand) * (kCandChunk + 1),
                          hipMemcpyDeviceToHost, stream));
    if (rec_ev) HIPCHK(hipEventRecord(evd[gr.buf], stream));
    return PGW_OK;
  }

  // launch_group's sequence captured from the stream into a graph, then launched as one.  The
  // executable graph of the group's buffer set is updated with the new parameters when the
  // sequence has the same shape, else instantiated again; its previous launch has finished (the
  // group two before this one, same buffer set, was waited for by finish_group).
  int launch_group_graph(Group& gr, const pgw_opts* o) {
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    int st = launch_group(gr, o, false);
    capturing = false;
    hipGraph_t gph = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream, &gph);
    if (st || ec != hipSuccess || !gph) {
      thr_prev.nt = 0;
      if (gph) (void)hipGraphDestroy(gph);
      return st ? st : PGW_ERR_DEVICE;
    }
    hipGraphExec_t& x = gexec[gr.buf];
    if (x) {
      hipGraphNode_t en = nullptr;
      hipGraphExecUpdateResult r = hipGraphExecUpdateError;
      if (hipGraphExecUpdate(x, gph, &en, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(x);
        x = nullptr;
      }
    }
    hipError_t e = hipSuccess;
    if (!x) e = hipGraphInstantiate(&x, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    if (e != hipSuccess) {
      x = nullptr;
      return PGW_ERR_DEVICE;
    }
    HIPCHK(hipGraphLaunch(x, stream));
    HIPCHK(hipEventRecord(evd[gr.buf], stream));
    return PGW_OK;
  }

  // Waits for a group's candidates and appends its records (jobcore.c:659-695) in (sky point, ss, bin) order.
  int finish_group(const Group& gr0, const pgw_opts* o, int* sgnlc, void* sgnlv, int cap, int* rec_work,
                   bool rec_float) {
#pragma clang fp contract(off)
    HIPCHK(hipEventSynchronize(evd[gr0.buf]));
    const Cand* hh = hc_pin.p + (size_t)gr0.buf * (kCandChunk + 1);  // header, then candidates
    if (hh->ii) return PGW_ERR_DEVICE;  // a fused launch's dependency wait ran out
    const int nc = hh->t;  // the device count keeps counting past the buffer's capacity
    const Group* pg = &gr0;
    Group big;
    DevBuf<Cand> bigbuf;
    if (nc > gr0.ccap) {
      // More candidates than the group's device buffer holds: run the group again on its own into
      // a buffer of exactly that many.  The next group (other buffer set) is drained first; the
      // rerun rebuilds this group's sky slots, which that group no longer needs.
      HIPCHK(hipStreamSynchronize(stream));
      int st;
      if ((st = bigbuf.alloc((size_t)std::max(nc, kCandChunk) + 1))) return st;
      big = gr0;
      big.cbuf = bigbuf.p;
      big.ccap = nc;
      if ((st = launch_group(big, o))) return st;
      HIPCHK(hipEventSynchronize(evd[big.buf]));
      if (hh->ii || hh->t != nc) return PGW_ERR_DEVICE;  // the same work found a different count
      if (hcand.size() < (size_t)nc) hcand.resize(nc);
      pg = &big;
    }
    const Group& gr = *pg;
    memcpy(hcand.data(), hh + 1, sizeof(Cand) * std::min(nc, kCandChunk));
    if (nc > kCandChunk) {  // rare: the rest synchronously (the buffer is not reused before this returns)
      HIPCHK(hipMemcpyAsync(hcand.data() + kCandChunk, gr.cbuf + 1 + kCandChunk, sizeof(Cand) * (nc - kCandChunk),
                            hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    std::sort(hcand.begin(), hcand.begin() + nc,
              [](const Cand& a, const Cand& b) { return a.t != b.t ? a.t < b.t : a.ii < b.ii; });
    int st = PGW_OK;
    size_t k = 0;
    for (int c = 0; c < nc; c++) {
      while (k + 1 < gr.items.size() && hcand[c].t >= gr.items[k + 1].t0) k++;
      const Item& it = gr.items[k];
      const Seg& g = *segs[it.seg];
      const pgw_settings& S = g.sett;
      const double* M = S.M;
      const int ss = it.ss_at(hcand[c].t - it.t0), ii = hcand[c].ii;
      double het1, sgnlt1;
      tmpl_params(M, it.mm, it.nn, ss, o->s0_flag, het1, sgnlt1);
      const double sgnl0 = it.sk.het0 + het1;
      double r[PGW_NPAR];
      r[0] = 2. * M_PI * ii / ((double)S.fftpad * S.nfft) + sgnl0;
      r[1] = sgnlt1;
      r[2] = it.sk.sgnlt2;
      r[3] = it.sk.sgnlt3;
      r[4] = sqrt(2. * (hcand[c].F - S.nd));
      bool veto = false;
      for (int l = 0; l < S.numlines_band; l++)
        if (r[0] >= S.lines[l][0] && r[0] <= S.lines[l][1]) { veto = true; break; }
      if (veto) continue;
      if (*sgnlc >= cap) {
        st = PGW_ERR_CAPACITY;
        (*sgnlc)++;
        continue;
      }
      if (rec_float) {  // FLOAT_TYPE = float: each field rounded once, as the reference's assignment does
        float* q = static_cast<float*>(sgnlv) + (size_t)PGW_NPAR * (*sgnlc);
        for (int f = 0; f < PGW_NPAR; f++) q[f] = (float)r[f];
      } else {
        memcpy(static_cast<double*>(sgnlv) + (size_t)PGW_NPAR * (*sgnlc), r, sizeof(r));
      }
      if (rec_work) rec_work[*sgnlc] = it.work;
      (*sgnlc)++;
    }
    return st;
  }

  // Work items (seg, pm, mm, nn) in order, grouped; the host turns group g-1 into records while
  // the device works on group g.
  int run(const int* work, int n, pgw_range* r, const pgw_opts* o, int* sgnlc, void* sgnlv, int cap, int* rec_work,
          int* FNum, const int* ssoff, const int* sslist, bool rec_float) override {
    HIPCHK(hipEventRecord(ev0, stream));
    Group prev;
    bool have_prev = false;
    int buf = 0, st = PGW_OK, i = 0;
    while (i < n && !st) {
      Group gr;
      gr.buf = buf;
      const Seg* g0 = nullptr;
      while (i < n && (int)gr.items.size() < group) {
        const int sgi = work[4 * i];
        if (sgi < 0 || sgi >= (int)segs.size()) { st = PGW_ERR_ARG; break; }
        const Seg& g = *segs[sgi];
        Item it;
        it.work = i;
        it.seg = sgi;
        it.pm = work[4 * i + 1];
        it.mm = work[4 * i + 2];
        it.nn = work[4 * i + 3];
        double a, b, c, d;
        if (!sky_coords(g, it.pm, it.mm, it.nn, it.sk, a, b, c, d)) { i++; continue; }  // outside the disk
        if (sslist) {
          it.ssl = sslist + ssoff[i];
          it.nlist = ssoff[i + 1] - ssoff[i];
        } else {
          spindown_range(g, it.mm, it.nn, r, o, it.smin, it.smax);
        }
        const int ns = std::max(0, it.ns());
        if (ns > kDescCap) { st = PGW_ERR_UNSUPPORTED; break; }
        if (g0 && (!compatible(*g0, g) || gr.ntempl + ns > kDescCap)) break;
        if (!g0) g0 = &g;
        if (ns == 0) { i++; continue; }
        it.slot = (int)gr.items.size();
        it.t0 = gr.ntempl;
        gr.ntempl += ns;
        gr.items.push_back(it);
        *FNum += ns;
        i++;
      }
      if (st || gr.items.empty()) continue;
      if ((st = graph ? launch_group_graph(gr, o) : launch_group(gr, o))) break;
      if (have_prev) st = finish_group(prev, o, sgnlc, sgnlv, cap, rec_work, rec_float);
      prev = std::move(gr);
      have_prev = true;
      buf ^= 1;
    }
    HIPCHK(hipEventRecord(ev1, stream));
    if (have_prev) {
      const int st2 = finish_group(prev, o, sgnlc, sgnlv, cap, rec_work, rec_float);
      if (!st) st = st2;
    }
    HIPCHK(hipEventSynchronize(ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
    last_ms = t;
    if (kused) kcollect();
    return st;
  }

  // one template (segment 0) through the template stage; F normalised as thresholded
  int debug_fstat(int pm, int mm, int nn, int ss, const pgw_opts* o, double* Fout) override {
    const Seg& g = *segs[0];
    Group gr;
    Item it{0, 0, pm, mm, nn, 0, ss, ss, 0, SkyState{}};
    double a, b, c, d;
    if (!sky_coords(g, pm, mm, nn, it.sk, a, b, c, d)) return 1;
    gr.items.push_back(it);
    gr.ntempl = 1;
    gr.buf = 0;
    int st;
    if ((st = launch_group(gr, o))) return st;
    const size_t L = g.sett.nfftf;
    std::vector<T> hF(L);
    std::vector<double> hmu(g.nparts);
    HIPCHK(hipMemcpyAsync(hF.data(), Ft.p, sizeof(T) * L, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(hmu.data(), mu.p, sizeof(double) * g.nparts, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int k = g.sett.nmin; k < g.sett.nmax; k++) {
      const int part = std::min((k - g.sett.nmin) / PGW_NAV, g.nparts - 1);
      Fout[k - g.sett.nmin] = (double)hF[(size_t)(k % g.ps->L2) * g.ps->L1 + k / g.ps->L2] / hmu[part];
    }
    return PGW_OK;
  }

  // Fstatnet (followup.c:413-534) and its gradient at n points of segment sgi, in batches of
  // kFsnBatch points: partial sums per (point, chunk), then a fixed-order final sum per point.
  int settings(int sgi, pgw_settings* out) const override {
    if (sgi < 0 || sgi >= (int)segs.size()) return PGW_ERR_ARG;
    *out = segs[sgi]->sett;
    return PGW_OK;
  }

  DevBuf<double> fsn_pts, fsn_part, fsn_out, fsn_grad;
  double* fsn_scratch = nullptr;  // this device's `scratch` (pgw_fstatnet_noalloc)

  FsnParams fsn_params(const Seg& g) const {
    FsnParams q{};
    q.N = g.sett.N;
    q.nifo = nifo;
    q.nchunk = (g.sett.N + FSN_CHUNK - 1) / FSN_CHUNK;
    q.oms = g.sett.oms;
    for (int d = 0; d < nifo; d++) {
      for (int k = 0; k < 9; k++) q.det[d].amod[k] = g.amod[9 * d + k];
      q.det[d].sphir = g.sphir[d];
      q.det[d].cphir = g.cphir[d];
      q.det[d].w = 1.0 / g.sig2[d];
    }
    return q;
  }

  // the two Fstatnet kernels for m points at device address pts, partial sums in part
  // (m * nchunk * FSN_NSUM doubles), results to device addresses out, grad (optional)
  int fsn_enqueue(const Seg& g, const FsnParams& q, const double* pts, int m, double* part, double* out,
                  double* grad) {
    if (grad)
      hipLaunchKernelGGL(k_fstatnet_part<true>, dim3(m, q.nchunk), dim3(FSN_NT), 0, stream, q, pts, g.xDat.p,
                         g.DetSSB.p, g.cosmodf.p, g.sinmodf.p, part);
    else
      hipLaunchKernelGGL(k_fstatnet_part<false>, dim3(m, q.nchunk), dim3(FSN_NT), 0, stream, q, pts, g.xDat.p,
                         g.DetSSB.p, g.cosmodf.p, g.sinmodf.p, part);
    hipLaunchKernelGGL(k_fstatnet_final, dim3((m + 63) / 64), dim3(64), 0, stream, m, q.nchunk, part, pts, out,
                       grad);
    return hipGetLastError() == hipSuccess ? PGW_OK : PGW_ERR_DEVICE;
  }

  int fstatnet(int sgi, const double* pts, int n, double* out, double* grad) override {
    constexpr int kFsnBatch = 4096;
    if (sgi < 0 || sgi >= (int)segs.size() || n < 0 || (n && (!pts || !out))) return PGW_ERR_ARG;
    const Seg& g = *segs[sgi];
    const FsnParams q = fsn_params(g);
    const int nb = std::min(n, kFsnBatch);
    int st;
    if (nb > 0 && ((st = fsn_pts.alloc_min((size_t)4 * nb)) || (st = fsn_part.alloc_min((size_t)nb * q.nchunk * FSN_NSUM)) ||
                   (st = fsn_out.alloc_min((size_t)11 * nb)) || (st = fsn_grad.alloc_min((size_t)4 * nb))))
      return st;
    HIPCHK(hipEventRecord(ev0, stream));
    for (int p0 = 0; p0 < n; p0 += kFsnBatch) {
      const int m = std::min(kFsnBatch, n - p0);
      HIPCHK(hipMemcpyAsync(fsn_pts.p, pts + 4 * (size_t)p0, sizeof(double) * 4 * m, hipMemcpyHostToDevice, stream));
      if ((st = fsn_enqueue(g, q, fsn_pts.p, m, fsn_part.p, fsn_out.p, grad ? fsn_grad.p : nullptr))) return st;
      HIPCHK(hipMemcpyAsync(out + 11 * (size_t)p0, fsn_out.p, sizeof(double) * 11 * m, hipMemcpyDeviceToHost, stream));
      if (grad)
        HIPCHK(hipMemcpyAsync(grad + 4 * (size_t)p0, fsn_grad.p, sizeof(double) * 4 * m, hipMemcpyDeviceToHost,
                              stream));
      HIPCHK(hipStreamSynchronize(stream));  // the next batch reuses the device buffers
    }
    HIPCHK(hipEventRecord(ev1, stream));
    HIPCHK(hipEventSynchronize(ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
    last_ms = t;
    return PGW_OK;
  }

  // Fstatnet of points spread over segments (point i in segment seg[i]): the points are grouped by
  // segment on the host (stable), uploaded once, each segment's two kernels queued back to back
  // over its own slice of the buffers, the outputs copied back once, in batches of kFsnBatch.
  int fstatnet_multi(const int* seg, const double* pts, int n, double* out, double* grad) override {
    constexpr int kFsnBatch = 4096;
    if (n < 0 || (n && (!seg || !pts || !out))) return PGW_ERR_ARG;
    for (int i = 0; i < n; i++)
      if (seg[i] < 0 || seg[i] >= (int)segs.size()) return PGW_ERR_ARG;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return seg[a] < seg[b]; });
    size_t maxchunk = 0;
    for (const auto& g : segs) maxchunk = std::max(maxchunk, (size_t)((g->sett.N + FSN_CHUNK - 1) / FSN_CHUNK));
    const int nb = std::min(n, kFsnBatch);
    int st;
    if (nb > 0 && ((st = fsn_pts.alloc_min((size_t)4 * nb)) || (st = fsn_part.alloc_min((size_t)nb * maxchunk * FSN_NSUM)) ||
                   (st = fsn_out.alloc_min((size_t)11 * nb)) || (st = fsn_grad.alloc_min((size_t)4 * nb))))
      return st;
    std::vector<double> hp((size_t)4 * nb), ho((size_t)11 * nb), hg((size_t)4 * nb);
    HIPCHK(hipEventRecord(ev0, stream));
    for (int p0 = 0; p0 < n; p0 += kFsnBatch) {
      const int m = std::min(kFsnBatch, n - p0);
      for (int j = 0; j < m; j++) memcpy(&hp[4 * (size_t)j], pts + 4 * (size_t)order[p0 + j], 4 * sizeof(double));
      HIPCHK(hipMemcpyAsync(fsn_pts.p, hp.data(), sizeof(double) * 4 * m, hipMemcpyHostToDevice, stream));
      size_t partoff = 0;
      for (int j0 = 0; j0 < m;) {  // runs of one segment
        const int sgi = seg[order[p0 + j0]];
        int j1 = j0 + 1;
        while (j1 < m && seg[order[p0 + j1]] == sgi) j1++;
        const Seg& g = *segs[sgi];
        const FsnParams q = fsn_params(g);
        if ((st = fsn_enqueue(g, q, fsn_pts.p + 4 * (size_t)j0, j1 - j0, fsn_part.p + partoff, fsn_out.p + 11 * (size_t)j0,
                              grad ? fsn_grad.p + 4 * (size_t)j0 : nullptr)))
          return st;
        partoff += (size_t)(j1 - j0) * q.nchunk * FSN_NSUM;
        j0 = j1;
      }
      HIPCHK(hipMemcpyAsync(ho.data(), fsn_out.p, sizeof(double) * 11 * m, hipMemcpyDeviceToHost, stream));
      if (grad) HIPCHK(hipMemcpyAsync(hg.data(), fsn_grad.p, sizeof(double) * 4 * m, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));  // the next batch reuses the buffers
      for (int j = 0; j < m; j++) {
        memcpy(out + 11 * (size_t)order[p0 + j], &ho[11 * (size_t)j], 11 * sizeof(double));
        if (grad) memcpy(grad + 4 * (size_t)order[p0 + j], &hg[4 * (size_t)j], 4 * sizeof(double));
      }
    }
    HIPCHK(hipEventRecord(ev1, stream));
    HIPCHK(hipEventSynchronize(ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
    last_ms = t;
    return PGW_OK;
  }

  // Fstatnet on caller-owned device buffers with the partial sums in this device's `scratch`: no
  // allocation, one batch, synchronous on the handle's stream.
  int fstatnet_noalloc(int sgi, const double* d_pts, int n, double* d_out, double* d_grad) override {
    if (sgi < 0 || sgi >= (int)segs.size() || n < 0 || (n && (!d_pts || !d_out))) return PGW_ERR_ARG;
    if (n == 0) return PGW_OK;
    const Seg& g = *segs[sgi];
    const FsnParams q = fsn_params(g);
    if ((size_t)n * q.nchunk * FSN_NSUM > kScratchDoubles) return PGW_ERR_SCRATCH;
    if (!fsn_scratch) return PGW_ERR_DEVICE;
    HIPCHK(hipEventRecord(ev0, stream));
    // every (point, chunk) slot of the partial sums is written before it is read: no clearing
    int st;
    if ((st = fsn_enqueue(g, q, d_pts, n, fsn_scratch, d_out, d_grad))) return st;
    HIPCHK(hipEventRecord(ev1, stream));
    HIPCHK(hipEventSynchronize(ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
    last_ms = t;
    return PGW_OK;
  }

  int debug_resampled(int pm, int mm, int nn, int det, double* a, double* b) override {
    if (det < 0 || det >= nifo) return PGW_ERR_ARG;
    const Seg& g = *segs[0];
    SkyState sk;
    int st;
    if (!sky_stage(g, pm, mm, nn, 0, sk, st)) return 1;
    if (st) return st;
    const int N = g.sett.N;
    std::vector<T2> h((size_t)2 * N);
    HIPCHK(hipMemcpyAsync(h.data(), slot_xm_ptr(0) + (size_t)2 * det * N, sizeof(T2) * 2 * N, hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < N; i++) {
      a[2 * i] = h[i].x;
      a[2 * i + 1] = h[i].y;
      b[2 * i] = h[N + i].x;
      b[2 * i + 1] = h[N + i].y;
    }
    return PGW_OK;
  }

  int time_stages(int sgi, int pm, int mm, int nn, int s0, int nt, const pgw_opts* o, int reps, double* ms) override {
    if (nt < 1 || reps < 1 || sgi < 0 || sgi >= (int)segs.size()) return PGW_ERR_ARG;
    nt = std::min(nt, TB);  // at most one launch
    const Seg& g = *segs[sgi];
    Item it{0, sgi, pm, mm, nn, 0, s0, s0 + nt - 1, 0, SkyState{}};
    int st = PGW_OK;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < 4; k++) HIPCHK(hipEventCreateWithFlags(&ev[k], hipEventDisableSystemFence));
    HIPCHK(hipEventRecord(ev0, stream));
    const int inside = sky_stage(g, pm, mm, nn, 0, it.sk, st);
    HIPCHK(hipEventRecord(ev1, stream));
    if (!inside && !st) st = 1;
    double acc[4] = {0, 0, 0, 0};
    if (!st) {
      HIPCHK(hipStreamSynchronize(stream));  // the pinned descriptor buffer may still feed a copy
      fill_desc(it, o, desc_pin.p);
      HIPCHK(hipMemcpyAsync(desc.p, desc_pin.p, sizeof(TmplDesc) * nt, hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
      acc[0] = t;
      for (int rr = 0; rr < reps && !st; rr++) {
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, &cand.p->t, 2);
        st = launch_templates(g, 0, cand.p, cand_cap, 0, nt, &it, 0, g.sett.N, o, ev);
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

  // Times the template stage per template for each (TB, row pass, column pass) variant on the
  // given sky point and keeps the fastest.
  // Times every kernel variant on a small search of sky point (pm, mm, nn) of segment sgi (four
  // copies of it with spindowns 0..511, so launches run back to back as in a search; threshold
  // out of reach so no records) and keeps the fastest.
  // the kernel variants in PGW_TUNE's syntax (setting PGW_TUNE to it reproduces them)
  int tune_string(char* buf, int len) const override {
    if (!buf || len <= 0) return PGW_ERR_ARG;
    const int n = snprintf(buf, (size_t)len,
                           "tb=%d,row_na=%d,row8=%d,col_seq=%d,col_lf=%d,col_nt=%d,row_nt=%d,f_nt=%d,rev=%d,"
                           "half=%d,spl_k=%d,group=%d,thr_fuse=%d,l1max=%d,candcap=%d,graph=%d,fused=%d,flag=%d,col_wide=%d,"
                           "spl_fuse=%d",
                           TB, row_na, row8, col_seq ? 1 : 0, col_lf ? 1 : 0, col_nt ? 1 : 0, row_nt ? 1 : 0,
                           f_nt ? 1 : 0, phase_rev ? 1 : 0,
                           row_half ? 1 : 0, spl_k, group, thr_fuse ? 1 : 0, l1max, cand_cap, graph ? 1 : 0, fused ? 1 : 0,
                           fused_lag, col_wide ? 1 : 0, spl_fuse ? 1 : 0);
    return n >= 0 && n < len ? PGW_OK : PGW_ERR_ARG;
  }

  // r: the caller's spindown range (each work item's templates as a search would form them), or
  // nullptr (512 spindowns per work item); nsky: work items (copies of the sky point) per timed call
  int autotune(int sgi, int pm, int mm, int nn, const pgw_opts* o, const pgw_range* r, int nsky, int* tb_out,
               int* na_out, int* seq_out) override {
    if (sgi < 0 || sgi >= (int)segs.size()) return PGW_ERR_ARG;
    const int tbs[3] = {64, 128, 256};
    // row pass variants: 16 elements per thread with the spectra in turn (na 1), 8 elements per
    // thread with the transforms' later exchanges inside each wave, without / with the second row
    // prefetched (row8 4 / 3), 16 elements per thread with the spectra together (na 2).  FP32
    // keeps the first: the others are slower over a search.  (row8 1 and 2, all exchanges
    // block-wide, are slower than 3 and 4 at every measured shape; PGW_TUNE still selects them.)
    const int rv_na[4] = {1, 1, 1, 2}, rv_r8[4] = {0, 4, 3, 0};
    const int nrv = std::is_same<T, double>::value ? 4 : 1;
    const int NI = nsky;
    constexpr int NS = 512;
    std::vector<int> work(4 * NI), ssoff(NI + 1), sslist((size_t)NI * NS);
    for (int i = 0; i < NI; i++) {
      work[4 * i] = sgi;
      work[4 * i + 1] = pm;
      work[4 * i + 2] = mm;
      work[4 * i + 3] = nn;
      ssoff[i] = i * NS;
      for (int j = 0; j < NS; j++) sslist[(size_t)i * NS + j] = j;
    }
    ssoff[NI] = NI * NS;
    // with the caller's range the searches keep the caller's threshold, so that the candidates'
    // cost (the threshold's run walks, the records) is timed as well; else no candidates
    pgw_opts oq = *o;
    if (!r) oq.trl = 1e300;
    constexpr int kRecCap = 1 << 16;
    std::vector<double> rec(r ? (size_t)kRecCap * PGW_NPAR : 0);
    pgw_range rq;
    memset(&rq, 0, sizeof(rq));
    if (r) rq = *r;
    const int kt_saved = ktime;
    ktime = 0;
    auto per_template = [&](double& ms) {
      int sgnlc = 0, fn = 0;
      pgw_range rr = rq;  // run may update the range it is given
      int st = r ? run(work.data(), NI, &rr, &oq, &sgnlc, rec.data(), kRecCap, nullptr, &fn, nullptr, nullptr, false)
                 : run(work.data(), NI, &rr, &oq, &sgnlc, nullptr, 0, nullptr, &fn, ssoff.data(), sslist.data(), false);
      if (st == PGW_ERR_CAPACITY) st = PGW_OK;  // more records than the buffer holds: timed all the same
      if (st) return st;
      if (fn == 0) return PGW_ERR_ARG;  // outside the sky disk
      ms = last_ms / fn;
      return PGW_OK;
    };
    double best = 1e30;
    int btb = TB, bna = row_na, br8 = row8, st = PGW_OK;
    bool bseq = col_seq, blf = col_lf, bnt = col_nt;
    // loads-first column pass: FP32 at column lengths of 128 and 256 (16 register rows)
    const int nlf = std::is_same<T, float>::value && segs[sgi]->ps->L2 >= 128 ? 2 : 1;
    for (int tb : tbs) {
      if ((st = ensure_tb(tb))) break;
      for (int rv = 0; rv < nrv && !st; rv++)
        for (int seq = 0; seq <= 1 && !st; seq++)
          for (int lf = 0; lf < nlf && !st; lf++)
            for (int nts = 0; nts <= 1 && !st; nts++) {
              TB = tb;
              row_na = rv_na[rv];
              row8 = rv_r8[rv];
              col_seq = seq != 0;
              col_lf = lf != 0;
              col_nt = nts != 0;
              double w, t1, t2;  // one untimed search, then the faster of two
              if ((st = per_template(w)) || (st = per_template(t1)) || (st = per_template(t2))) break;
              const double per = std::min(t1, t2);
              if (per < best) {
                best = per;
                btb = tb;
                bna = rv_na[rv];
                br8 = rv_r8[rv];
                bseq = seq != 0;
                blf = lf != 0;
                bnt = nts != 0;
              }
            }
      if (st) break;
    }
    ktime = kt_saved;
    TB = btb;
    row_na = bna;
    row8 = br8;
    col_seq = bseq;
    col_lf = blf;
    col_nt = bnt;
    if (st) return st;
    if (tb_out) *tb_out = TB;
    if (na_out) *na_out = row8 ? 8 : row_na;
    if (seq_out) *seq_out = col_seq;
    return PGW_OK;
  }
};

template <typename T>
int make_engine(pgw_handle** h, const pgw_settings* s, const pgw_detector* ifo, int nseg, int nifo) {
  auto* q = new (std::nothrow) Engine<T>();
  if (!q) return PGW_ERR_NOMEM;
  const int st = q->init(s, ifo, nseg, nifo);
  if (st) {
    delete q;
    return st;
  }
  *h = q;
  return PGW_OK;
}

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

int pgw_init_multi(pgw_handle** h, const pgw_settings* sett, const pgw_detector* ifo, int nseg, int nifo,
                   int precision) {
  if (!h || !sett || !ifo || nseg < 1 || nifo < 1 || nifo > PGW_MAX_DETECTORS) return PGW_ERR_ARG;
  *h = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PGW_ERR_NODEVICE;
  if (precision == PGW_PREC_FP64) return make_engine<double>(h, sett, ifo, nseg, nifo);
  if (precision == PGW_PREC_FP32) return make_engine<float>(h, sett, ifo, nseg, nifo);
  return PGW_ERR_ARG;
}

int pgw_init(pgw_handle** h, const pgw_settings* sett, const pgw_detector* ifo, int nifo, int precision) {
  return pgw_init_multi(h, sett, ifo, 1, nifo, precision);
}

int pgw_job_core(pgw_handle* h, int pm, int mm, int nn, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                 double* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv)) return PGW_ERR_ARG;
  const int work[4] = {0, pm, mm, nn};
  return h->run(work, 1, s_range, opts, sgnlc, sgnlv, sgnlv_cap, nullptr, FNum);
}

int pgw_job_batch(pgw_handle* h, const int* sky, int nsky, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                  double* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || (!sky && nsky > 0) || nsky < 0 || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv))
    return PGW_ERR_ARG;
  std::vector<int> work((size_t)4 * nsky);
  for (int i = 0; i < nsky; i++) {
    work[4 * i] = 0;
    work[4 * i + 1] = sky[3 * i];
    work[4 * i + 2] = sky[3 * i + 1];
    work[4 * i + 3] = sky[3 * i + 2];
  }
  return h->run(work.data(), nsky, s_range, opts, sgnlc, sgnlv, sgnlv_cap, nullptr, FNum);
}

int pgw_job_core_f(pgw_handle* h, int pm, int mm, int nn, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                   float* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv)) return PGW_ERR_ARG;
  const int work[4] = {0, pm, mm, nn};
  return h->run(work, 1, s_range, opts, sgnlc, sgnlv, sgnlv_cap, nullptr, FNum, nullptr, nullptr, true);
}

int pgw_job_batch_f(pgw_handle* h, const int* sky, int nsky, pgw_range* s_range, const pgw_opts* opts, int* sgnlc,
                    float* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || (!sky && nsky > 0) || nsky < 0 || !s_range || !opts || !sgnlc || !FNum || (sgnlv_cap > 0 && !sgnlv))
    return PGW_ERR_ARG;
  std::vector<int> work((size_t)4 * nsky);
  for (int i = 0; i < nsky; i++) {
    work[4 * i] = 0;
    work[4 * i + 1] = sky[3 * i];
    work[4 * i + 2] = sky[3 * i + 1];
    work[4 * i + 3] = sky[3 * i + 2];
  }
  return h->run(work.data(), nsky, s_range, opts, sgnlc, sgnlv, sgnlv_cap, nullptr, FNum, nullptr, nullptr, true);
}

int pgw_job_batch_multi(pgw_handle* h, const int* work, int nwork, pgw_range* s_range, const pgw_opts* opts,
                        int* sgnlc, double* sgnlv, int sgnlv_cap, int* rec_work, int* FNum) {
  if (!h || (!work && nwork > 0) || nwork < 0 || !s_range || !opts || !sgnlc || !FNum ||
      (sgnlv_cap > 0 && !sgnlv))
    return PGW_ERR_ARG;
  return h->run(work, nwork, s_range, opts, sgnlc, sgnlv, sgnlv_cap, rec_work, FNum);
}

int pgw_autotune(pgw_handle* h, int seg, int pm, int mm, int nn, const pgw_opts* opts, int* tb, int* row_na,
                 int* col_seq) {
  if (!h || !opts) return PGW_ERR_ARG;
  return h->autotune(seg, pm, mm, nn, opts, nullptr, 4, tb, row_na, col_seq);
}

int pgw_autotune_range(pgw_handle* h, int seg, int pm, int mm, int nn, const pgw_opts* opts,
                       const pgw_range* s_range, int nsky, int* tb, int* row_na, int* col_seq) {
  if (!h || !opts || !s_range || nsky < 1 || nsky > 64) return PGW_ERR_ARG;
  return h->autotune(seg, pm, mm, nn, opts, s_range, nsky, tb, row_na, col_seq);
}

int pgw_tune_string(const pgw_handle* h, char* buf, int len) {
  if (!h) return PGW_ERR_ARG;
  return h->tune_string(buf, len);
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
  return h->time_stages(0, pm, mm, nn, s0, nt, opts, reps, ms);
}

int pgw_job_spotlight(pgw_handle* h, int pm, const int* mm, const int* nn, const int* noss, const int* ss, int nsky,
                      const pgw_opts* opts, int* sgnlc, double* sgnlv, int sgnlv_cap, int* FNum) {
  if (!h || !opts || !sgnlc || !FNum || nsky < 0 || (nsky && (!mm || !nn || !noss)) || (sgnlv_cap > 0 && !sgnlv))
    return PGW_ERR_ARG;
  std::vector<int> work(4 * (size_t)nsky), off((size_t)nsky + 1, 0);
  for (int i = 0; i < nsky; i++) {
    if (noss[i] < 0 || (noss[i] > 0 && !ss)) return PGW_ERR_ARG;
    work[4 * i] = 0;
    work[4 * i + 1] = pm;
    work[4 * i + 2] = mm[i];
    work[4 * i + 3] = nn[i];
    off[i + 1] = off[i] + noss[i];
  }
  pgw_range r{};  // unused: every sky point brings its spindown list
  return h->run(work.data(), nsky, &r, opts, sgnlc, sgnlv, sgnlv_cap, nullptr, FNum, off.data(), ss);
}

int pgw_fstatnet(pgw_handle* h, int seg, const double* pts, int npts, double* out, double* grad) {
  if (!h) return PGW_ERR_ARG;
  return h->fstatnet(seg, pts, npts, out, grad);
}

int pgw_fstatnet_multi(pgw_handle* h, const int* seg, const double* pts, int npts, double* out, double* grad) {
  if (!h) return PGW_ERR_ARG;
  return h->fstatnet_multi(seg, pts, npts, out, grad);
}

int pgw_fstatnet_noalloc(pgw_handle* h, int seg, const double* d_pts, int npts, double* d_out, double* d_grad) {
  if (!h) return PGW_ERR_ARG;
  return h->fstatnet_noalloc(seg, d_pts, npts, d_out, d_grad);
}

int pgw_copy_bench(size_t bytes, int reps, double* ms) {
  if (!ms || bytes < 32 || reps < 1) return PGW_ERR_ARG;
  const size_t n = bytes / sizeof(double4);
  double4 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  HIPCHK(hipMalloc(&a, n * sizeof(double4)));
  if (hipMalloc(&b, n * sizeof(double4)) != hipSuccess) {
    (void)hipFree(a);
    return PGW_ERR_NOMEM;
  }
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

int pgw_hbm_probe(int kind, size_t bytes, int reps, double* ms) {
  if (!ms || kind < 0 || kind > 2 || bytes < 4096 || reps < 1) return PGW_ERR_ARG;
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  float* o = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t sq = nullptr;  // a stream of its own (the null stream stays unused)
  int st = PGW_OK;
  auto fail = [&](int s) {
    if (!st) st = s;
  };
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess || hipMalloc(&o, 64) != hipSuccess)
    fail(PGW_ERR_NOMEM);
  if (!st && (hipStreamCreateWithFlags(&sq, hipStreamNonBlocking) != hipSuccess ||
              hipMemsetAsync(a, 0, n * 16, sq) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess))
    fail(PGW_ERR_DEVICE);
  double best = 1e30;
  // the fastest of three grid sizes (a few workgroups per CU up to many): copies and reads peak
  // at small grids, non-temporal writes at large ones
  for (int grid : {2048, 8192, 32768}) {
    if (st) break;
    auto go = [&]() {
      if (kind == 0) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(NT), 0, sq, (const double4*)a, (double4*)b, n / 2);
      else if (kind == 1) hipLaunchKernelGGL(k_read_probe, dim3(grid), dim3(NT), 0, sq, (const float4*)a, n, o);
      else hipLaunchKernelGGL(k_write_probe, dim3(grid), dim3(NT), 0, sq, (float4*)b, n);
    };
    go();  // warm-up
    if (hipEventRecord(e0, sq) != hipSuccess) { fail(PGW_ERR_DEVICE); break; }
    for (int r = 0; r < reps; r++) go();
    float t = 0.f;
    if (hipEventRecord(e1, sq) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess)
      fail(PGW_ERR_DEVICE);
    else
      best = std::min(best, (double)t / reps);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sq) {
    (void)hipStreamSynchronize(sq);
    (void)hipStreamDestroy(sq);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(o);
  if (!st) *ms = best;
  return st;
}

double pgw_last_gpu_ms(const pgw_handle* h) { return h ? h->last_ms : 0.0; }

int pgw_kernel_times(pgw_handle* h, int enable, double out[5]) {
  if (!h) return PGW_ERR_ARG;
  return h->kernel_times(enable, out);
}

int pgw_get_settings(const pgw_handle* h, int seg, pgw_settings* sett) {
  if (!h || !sett) return PGW_ERR_ARG;
  return h->settings(seg, sett);
}

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
    case PGW_ERR_IO: return "trigger or state file I/O error";
    case PGW_ERR_SCRATCH: return "points x chunks exceed the allocation-free entry point's scratch";
    default: return "unknown status";
  }
}

}  // extern "C"
