// pgw_fused.hpp - the template stage of one launch (demodulation + column pass, row pass + F +
// FStat partials, threshold; jobcore.c:456-695) as one persistent work-queue kernel.
//
// A grid of resident workgroups takes jobs from a device counter in ticket order: the column
// tiles of template s, the rows of template s - D and the threshold blocks of template s - 2D form
// stage s.  A row job waits until every column tile of its template has finished, a threshold job
// until every row of its template has; both counters are raised by jobs with lower tickets, which
// workgroups that wait on nothing have already taken, so every wait ends.  Each wait is also bounded
// (spin_limit polls): a wait that runs out marks the launch's results invalid (err) and the job
// goes on, so the kernel always exits.  The last workgroup to leave clears the counters for the
// next launch.
//
// Data handed from one job to another (spectra, F, FStat partials) is written and read at device
// scope (cache policy SC1: the store is performed past this XCD's L2, the load misses a line that
// L2 may hold stale), and each job waits for its stores to complete before its count is raised.
// So no job needs a device-scope fence, which would write back and invalidate the whole L2 of its
// XCD (with such fences every job, the network workload ran 17 times slower).
//
// Built for the FP32 form (phase in revolutions, the spectra in turn through one LDS buffer in
// both passes, half-spectrum row pass); the host uses the separate kernels elsewhere.
#pragma once
#include "pgw_kernels2.hpp"

namespace pgw {

constexpr int kFusedTB = 256;  // templates per fused launch at most
constexpr int kSC1 = 16;       // buffer cache policy: device scope

// waits until this wave's stores have completed (vmcnt 0)
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// counters of a fused launch, all zero between launches
struct FusedSync {
  int ticket;             // next job
  int exitcnt;            // workgroups that have left
  int colcnt[kFusedTB];   // column jobs finished, per template
  int rowcnt[kFusedTB];   // row jobs finished, per template
};

struct FusedArgs {
  SpinParams sp;
  SlotBases sb;
  RowParams rp;
  ThrParams tp;
  const TmplDesc* desc;
  cplx<float>* Z;
  const cplx<float>* twc;
  const cplx<float>* twr;
  const cplx<float>* hi;
  const cplx<float>* lo;
  float* Ft;
  float* psum;
  float* pmax;
  double* mu;
  Cand* cand;
  int* ncand;
  int* err;          // set to 1 when a dependency wait runs out
  FusedSync* sy;
  int nt;            // templates of the launch
  int ct;            // column jobs (tiles) per template
  int rj;            // row jobs per template
  int tj;            // threshold jobs (4 FStat blocks each) per template
  int D;             // lag of the row jobs behind the column jobs, and of the threshold jobs behind the rows
  int o1, o2, o3, o4, total;  // ticket boundaries of the five phases of the stage sequence
  int spin_limit;
};

// job j -> kind (0 column tile, 1 row job, 2 threshold job), template t, index i among the
// template's jobs of that kind.  Stages [0, D): columns; [D, 2D): columns, rows; [2D, nt): all
// three; [nt, nt + D): rows, thresholds; [nt + D, nt + 2D): thresholds (nt >= 2D).
__host__ __device__ __forceinline__ void fused_job(const FusedArgs& a, int j, int& kind, int& t, int& i) {
  if (j < a.o1) {
    const int q = j / a.ct;
    kind = 0, t = q, i = j - q * a.ct;
  } else if (j < a.o2) {
    const int B = a.ct + a.rj, k = j - a.o1, q = k / B, r = k - q * B;
    if (r < a.ct) kind = 0, t = a.D + q, i = r;
    else kind = 1, t = q, i = r - a.ct;
  } else if (j < a.o3) {
    const int B = a.ct + a.rj + a.tj, k = j - a.o2, q = k / B, r = k - q * B;
    if (r < a.ct) kind = 0, t = 2 * a.D + q, i = r;
    else if (r < a.ct + a.rj) kind = 1, t = a.D + q, i = r - a.ct;
    else kind = 2, t = q, i = r - a.ct - a.rj;
  } else if (j < a.o4) {
    const int B = a.rj + a.tj, k = j - a.o3, q = k / B, r = k - q * B;
    if (r < a.rj) kind = 1, t = a.nt - a.D + q, i = r;
    else kind = 2, t = a.nt - 2 * a.D + q, i = r - a.rj;
  } else {
    const int k = j - a.o4, q = k / a.tj;
    kind = 2, t = a.nt - a.D + q, i = k - q * a.tj;
  }
}

// polls *cnt until it reaches need (one lane); a wait that runs out sets *err
__device__ __forceinline__ void fused_wait(const FusedArgs& a, int* cnt, int need) {
  for (int it = 0; __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need; it++) {
    if (it >= a.spin_limit) {
      __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Column job: tile bx of template t (the column pass of k_spin_col2 for the FP32 form, phase in
// revolutions, the spectra in turn through lds).
// tw: the twiddle tables (twc, hi, lo), passed per job
struct FusedTw {
  const cplx<float>* twc;
  const cplx<float>* twr;
  const cplx<float>* hi;
  const cplx<float>* lo;
};

template <int P2, int NZ>
__device__ __forceinline__ void fused_col(const FusedArgs& a, const FusedTw& tw, const SpinTab& tab, int bx, int t,
                                          cplx<float>* lds) {
  using T2 = cplx<float>;
  constexpr int E = ColE<float, P2>::E;
  using S = ColShape<P2, E>;
  const int nifo = a.sb.nifo;
  int slot, N;
  double het1, sgnlt1;
  if (a.sb.uslot >= 0) {  // the launch's templates are spindowns of one sky point
    slot = a.sb.uslot;
    N = a.sb.uN;
    if (a.sb.tab) {
      het1 = tab.het1[t];
      sgnlt1 = tab.sgnlt1[t];
    } else {
      het1 = ld_global(&a.desc[t].het1, 0);
      sgnlt1 = ld_global(&a.desc[t].sgnlt1, 0);
    }
  } else {
    const TmplDesc& d = a.desc[t];
    slot = d.slot;
    N = d.N;
    het1 = d.het1;
    sgnlt1 = d.sgnlt1;
  }
  const BufRes xmb(static_cast<const T2*>(a.sb.xm) + (size_t)slot * a.sb.xm_stride,
                   (uint32_t)(2 * nifo * N * sizeof(T2)));
  const BufRes t1b(a.sb.tmp1 + (size_t)slot * a.sb.tmp1_stride, (uint32_t)(nifo * N * sizeof(double)));
  const int c = threadIdx.x % S::C, g = threadIdx.x / S::C;
  const int n1 = bx * S::C + c;
  const int L1 = a.sp.L1;
  const size_t L = (size_t)L1 * P2;
  T2 x[2][E];
  const PhaseRev pr(het1, sgnlt1);
#pragma unroll
  for (int r = 0; r < E; r++) {
    const int n = n1 + L1 * (g + S::G * r);
    T2 xa = mk<T2>(0, 0), xb = mk<T2>(0, 0);
    if (r < NZ && n < N) {
      const double dn = (double)n;
      const uint32_t vx = (uint32_t)n * (uint32_t)sizeof(T2), vt = (uint32_t)n * (uint32_t)sizeof(double);
#pragma unroll 2
      for (int d = 0; d < nifo; d++) {
        const uint32_t dN = (uint32_t)d * (uint32_t)N;
        const double tmp1 = t1b.load<double>(vt, dN * (uint32_t)sizeof(double));
        const float2 e = pr.eval(dn, tmp1);
        const T2 ex = mk<T2>(e.x, e.y);
        xa = cmac(xa, xmb.load<T2>(vx, 2 * dN * (uint32_t)sizeof(T2)), ex);
        xb = cmac(xb, xmb.load<T2>(vx, (2 * dN + (uint32_t)N) * (uint32_t)sizeof(T2)), ex);
      }
    }
    x[0][r] = xa;
    x[1][r] = xb;
  }
  T2* const Zt = a.Z + (size_t)t * 2 * L;
  const BufRes Zb(Zt, (uint32_t)(2 * L * sizeof(T2)));
  const uint32_t voff = (uint32_t)(n1 + g * S::Q * L1) * (uint32_t)sizeof(T2);
  const uint32_t rowb = (uint32_t)L1 * (uint32_t)sizeof(T2), specb = (uint32_t)(L * sizeof(T2));
  T2 w[E];
  col_out_twiddles<P2, E>(tw_big(tw.hi, tw.lo, (uint32_t)n1 << a.sp.twshift), w);
  constexpr bool PW = E == 16;
  T2 pw[PW ? E : 1];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    T2 (&xs)[1][E] = *reinterpret_cast<T2(*)[1][E]>(&x[s]);
    if (s) __syncthreads();
    col_fft<T2, P2, E, 1, -1, NZ>(xs, lds, tw.twc, PW ? pw : nullptr, s > 0);
#pragma unroll
    for (int q = 0; q < E; q++) {
      const T2 v = cmul(xs[0][q], w[q]);
      const uint32_t so = (uint32_t)s * specb + (uint32_t)col_out_k2c<P2, E>(q) * rowb;
      if (a.sp.nt) Zb.store_c<kSC1 | 2>(v, voff, so);
      else Zb.store_c<kSC1>(v, voff, so);
    }
  }
}

// Row job part: row k2 of template t by the P1/16 threads tid of one group (the row pass of
// k_spin_row2 for FP32, the spectra in turn, the second row prefetched, half spectrum); lds holds
// RowShape<P1>::PADN elements of this group.  Every thread of the workgroup takes the barriers.
template <int P1>
__device__ __forceinline__ void fused_row(const FusedArgs& a, const cplx<float>* __restrict__ twr, int k2, int t, int tid,
                                          cplx<float>* lds) {
  using T = float;
  using T2 = cplx<T>;
  using S = RowShape<P1>;
  constexpr int NT = S::NT;
  const RowParams& p = a.rp;
  const size_t L = (size_t)P1 * p.L2;
  const BufRes Zr(a.Z + (size_t)t * 2 * L, (uint32_t)(2 * L * sizeof(T2)));
  const uint32_t vr = (uint32_t)tid * (uint32_t)sizeof(T2), mb = NT * (uint32_t)sizeof(T2);
  const uint32_t ra = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T2), rb = ra + (uint32_t)(L * sizeof(T2));
  auto ldz = [&](uint32_t off) { return p.nt ? Zr.load<T2, kSC1 | 2>(vr, off) : Zr.load<T2, kSC1>(vr, off); };
  const double* aabb = a.sb.aabb + 2 * (a.sb.uslot >= 0 ? a.sb.uslot : a.desc[t].slot);
  const T raa = (T)(1.0 / ld_global(aabb, 0)), rbb = (T)(1.0 / ld_global(aabb, 1));
  constexpr int MO = 8;  // half spectrum: output slots that can hold a bin of [nmin, nmax)
  T F[MO];
  T2 x[1][16], y[1][16];
#pragma unroll
  for (int m = 0; m < 16; m++) x[0][m] = ldz(ra + mb * m);
#pragma unroll
  for (int m = 0; m < 16; m++) y[0][m] = ldz(rb + mb * m);
  T2 pc[16];
  constexpr bool PB = P1 != 2048;
  T2 pb[PB ? 16 : 1];
  T2* pbp = PB ? pb : nullptr;
  row_fft<T2, P1, 1, -1, 1>(x, lds, twr, pc, false, pbp, tid);
#pragma unroll
  for (int m = 0; m < MO; m++) F[m] = (x[0][m].x * x[0][m].x + x[0][m].y * x[0][m].y) * raa;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 16; m++) x[0][m] = y[0][m];
  row_fft<T2, P1, 1, -1, 1>(x, lds, twr, pc, true, pbp, tid);
#pragma unroll
  for (int m = 0; m < MO; m++) F[m] = F[m] + (x[0][m].x * x[0][m].x + x[0][m].y * x[0][m].y) * rbb;
  __syncthreads();
  T* Fs = reinterpret_cast<T*>(lds);
  const BufRes Fb(a.Ft + (size_t)t * L, (uint32_t)(L * sizeof(T)));
  const uint32_t vf = (uint32_t)tid * (uint32_t)sizeof(T), fo = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T);
#pragma unroll
  for (int m = 0; m < MO; m++) {
    const int k1 = tid + NT * m;
    const int k = p.L2 * k1 + k2;
    const bool ok = (k >= p.nmin && k < p.nmax);
    if (ok) {
      if (p.fnt) Fb.template store<kSC1 | 2>(F[m], vf, fo + (uint32_t)(NT * m * sizeof(T)));
      else Fb.template store<kSC1>(F[m], vf, fo + (uint32_t)(NT * m * sizeof(T)));
    }
    Fs[k1] = ok ? F[m] : (T)0;
  }
  __syncthreads();
  // FStat block partials of this row, as in k_spin_row2: per-thread sums over 16 consecutive bins
  // split at a block boundary, then one thread per block adds them in thread order
  T* rs = Fs + P1;
  if (tid < NT / 2) {
    const int j0 = 16 * tid;
    const int nb = p.nparts - 1;
    const int b0 = min(max((p.L2 * j0 + k2 - p.nmin) >> 12, 0), nb);
    const int jsplit = b0 == nb ? P1 : (PGW_NAV_DEV * (b0 + 1) + p.nmin - k2 + p.L2 - 1) >> p.lgL2;
    T s0 = 0, s1 = 0, m0 = 0, m1 = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int j = j0 + ((i + tid + (tid >> 4)) & 15);
      const T f = Fs[j];
      if (j < jsplit) { s0 += f; m0 = f > m0 ? f : m0; }
      else { s1 += f; m1 = f > m1 ? f : m1; }
    }
    rs[tid] = s0;
    rs[NT + tid] = s1;
    rs[2 * NT + tid] = m0;
    rs[3 * NT + tid] = m1;
  }
  __syncthreads();
  for (int part = tid; part < p.nparts; part += NT) {
    const int lok = p.nmin + PGW_NAV_DEV * part;
    const int hik = (part < p.nblk) ? lok + PGW_NAV_DEV : p.nmax;
    const int jlo = max(0, (lok - k2 + p.L2 - 1) >> p.lgL2);
    const int jhi = min((hik - k2 + p.L2 - 1) >> p.lgL2, P1);
    T sum = 0, mx = 0;
    if (jlo < jhi) {
      const int nb = p.nparts - 1;
      for (int u = jlo >> 4; u <= (jhi - 1) >> 4; u++) {
        const int b0 = min(max((p.L2 * (16 * u) + k2 - p.nmin) >> 12, 0), nb);
        const int h = b0 == part ? 0 : 1;
        sum += rs[h * NT + u];
        const T v = rs[(2 + h) * NT + u];
        mx = v > mx ? v : mx;
      }
    }
    __hip_atomic_store(&a.psum[((size_t)t * p.nparts + part) * p.L2 + k2], sum, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&a.pmax[((size_t)t * p.nparts + part) * p.L2 + k2], mx, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int P2, int P1> struct FusedShape {
  static constexpr int RJ = 256 / RowShape<P1>::NT;  // rows per row job
  static constexpr int NCOL = P2 * ColShape<P2, ColE<float, P2>::E>::C;
  static constexpr int NROW = RJ * RowShape<P1>::PADN;
  static constexpr int NTHR = 4 * THR_LDS;  // doubles: one mean table and bitmap per wave
  static constexpr int N = NCOL > NROW ? (NCOL > NTHR ? NCOL : NTHR) : (NROW > NTHR ? NROW : NTHR);
  // waves per SIMD: the job loop with the row pass inside needs 200-236 VGPRs (at 3 or 4 waves
  // per SIMD it spilled 30-130 registers)
  static constexpr int W = 2;
};

template <int P2, int P1, int NZ>
__global__ __launch_bounds__(256, (FusedShape<P2, P1>::W)) void k_spin_fused(const FusedArgs a, const SpinTab tab) {
  using SH = FusedShape<P2, P1>;
  constexpr int NT = RowShape<P1>::NT;
  __shared__ cplx<float> lds[SH::N];
  __shared__ int job_s;
  for (;;) {
    if (threadIdx.x == 0) job_s = atomicAdd(&a.sy->ticket, 1);
    __syncthreads();
    const int j = __builtin_amdgcn_readfirstlane(job_s);  // uniform: the job's offsets stay scalar
    if (j >= a.total) break;  // uniform
    int kind, t, i;
    fused_job(a, j, kind, t, i);
    // the twiddle tables through registers the compiler cannot see into, so no twiddle load or
    // product is hoisted out of the job loop (it would hold their registers across every job)
    FusedTw tw{a.twc, a.twr, a.hi, a.lo};
    asm volatile("" : "+s"(tw.twc), "+s"(tw.twr), "+s"(tw.hi), "+s"(tw.lo));
    if (kind == 0) {
      fused_col<P2, NZ>(a, tw, tab, i, t, lds);
      wait_stores();  // this wave's spectrum stores performed at device scope before the count
      __syncthreads();
      if (threadIdx.x == 0) atomicAdd(&a.sy->colcnt[t], 1);
    } else if (kind == 1) {
      if (threadIdx.x == 0) fused_wait(a, &a.sy->colcnt[t], a.ct);
      __syncthreads();  // the spectra are read (at device scope) after the count
      // the row group of this wave (a row has P1/16 >= 64 threads: whole waves), uniform in the
      // wave, so the row's offsets stay in scalar registers
      const int sub = __builtin_amdgcn_readfirstlane((int)threadIdx.x / NT);
      fused_row<P1>(a, tw.twr, SH::RJ * i + sub, t, threadIdx.x - sub * NT, lds + sub * RowShape<P1>::PADN);
      wait_stores();  // F and the partials performed at device scope before the count
      __syncthreads();
      if (threadIdx.x == 0) atomicAdd(&a.sy->rowcnt[t], 1);
    } else {
      if (threadIdx.x == 0) fused_wait(a, &a.sy->rowcnt[t], a.rj);
      __syncthreads();
      const int w = threadIdx.x >> 6, part = 4 * i + w;
      if (part < a.tp.nparts)
        thr_wave<float, 32, 16, true>(a.tp, part, t, threadIdx.x & 63, reinterpret_cast<double*>(lds) + w * THR_LDS, a.Ft,
                        a.psum, a.pmax, a.mu, a.cand, a.ncand);
    }
    __syncthreads();  // lds and job_s are reused by the next job
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&a.sy->exitcnt, 1) == (int)gridDim.x - 1) {  // every other workgroup has left
      for (int q = 0; q < a.nt; q++) {
        a.sy->colcnt[q] = 0;
        a.sy->rowcnt[q] = 0;
      }
      a.sy->ticket = 0;
      a.sy->exitcnt = 0;
    }
  }
}

}  // namespace pgw
