// pgw_kernels2.hpp - template-stage and resampling FFT kernels on the register-resident FFT
// blocks of pgw_fft2.hpp.
#pragma once
#include <type_traits>

#include "pgw_fft2.hpp"
#include "pgw_kernels.hpp"

namespace pgw {

#ifndef PGW_ROW64_W
#define PGW_ROW64_W 1
#endif
#ifndef PGW_ROW64_PREF
#define PGW_ROW64_PREF true
#endif

// column-pass register depth per precision and column length (LDS per array = 256*E elements)
template <typename T, int P2> struct ColE;
template <int P2> struct ColE<float, P2> { static constexpr int E = (P2 <= 64 ? 8 : 16); static constexpr bool SEQ = false; };
template <> struct ColE<double, 16> { static constexpr int E = 8; static constexpr bool SEQ = false; };
template <> struct ColE<double, 32> { static constexpr int E = 8; static constexpr bool SEQ = false; };
template <> struct ColE<double, 64> { static constexpr int E = 8; static constexpr bool SEQ = false; };
template <> struct ColE<double, 128> { static constexpr int E = 16; static constexpr bool SEQ = true; };
template <> struct ColE<double, 256> { static constexpr int E = 16; static constexpr bool SEQ = true; };

// Fused spindown demodulation + column pass of the F-statistic FFT (jobcore.c:456-630).
// grid (L1/C, ntemplates), 256 threads.  SEQ: the two spectra share one LDS buffer in turn.  For template t (ss = ss0 + t) and detector d:
//   x_{a,b}[n] += xm[d][{a,b}][n] exp(-i phi_d(n)),  phi_d(n) = het1 n + sgnlt1 (n^2 + 2 n shft_d[n]),
// zero for n >= N; the column FFT and the W_L^{n1 k2} twiddle follow; Z[t][{a,b}][k2][n1].
// NZ: register rows r >= NZ hold n >= L1 P2 NZ / E >= N for every template of the launch
// (zero-padded input), so they are neither demodulated nor transformed.
// tf: the previous launch's threshold, run by the first tf.rows grid rows (none when tf.rows = 0):
// dispatched first, their short dependent loads overlap the column pass's workgroups.
// LF (loads first): detector by detector, the loads of all NZ register rows are issued before any
// of them is used (indexes clamped into the series, rows past N add nothing), so a wave waits on
// its loads once per detector instead of once per row.  Built for FP32 with E = 16 only: at
// E = 8 the extra registers cost waves per SIMD and the pass is 12-15 % slower.
// minimum waves per SIMD of the column pass (FP32 with 8 register rows: PGW_COL32_W; FP64 with 8
// register rows: PGW_COL64_W = 4, i.e. at most 128 VGPRs: the 64-point pass needed 129 and ran at 3
// waves per SIMD; at 4 (12 bytes spilled) it is 3.5 % faster)
#ifndef PGW_COL32_W
#define PGW_COL32_W 1
#endif
#ifndef PGW_COL64_W
#define PGW_COL64_W 4
#endif
template <typename T, int P2> struct ColOcc {
  static constexpr int W = ColE<T, P2>::E != 8 ? 1 : std::is_same<T, float>::value ? PGW_COL32_W : PGW_COL64_W;
};

// NTH: threads per workgroup (256, or 512 for tiles of twice the columns: each store instruction
// then writes row segments twice as long)
template <typename T, int P2, bool SEQ, bool UNI, int NZ = ColE<T, P2>::E, bool LF = false, int NTH = 256>
__global__ __launch_bounds__(NTH, (ColOcc<T, P2>::W)) void k_spin_col2(SpinParams p, SlotBases sb, const SpinTab tab,
                                                   const TmplDesc* __restrict__ desc,
                                                   cplx<T>* __restrict__ Z, const cplx<T>* __restrict__ twc,
                                                   const cplx<T>* __restrict__ hi, const cplx<T>* __restrict__ lo,
                                                   const ThrFuse<T> tf) {
  using T2 = cplx<T>;
  constexpr int E = ColE<T, P2>::E;
  using S = ColShape<P2, E, NTH>;
  __shared__ T2 lds[(SEQ ? 1 : 2) * P2 * S::C];
  static_assert(sizeof(lds) >= NTH / 64 * THR_LDS * sizeof(double), "threshold waves' mean tables fit the LDS");
  if ((int)blockIdx.y < tf.rows) {  // whole workgroups: no barrier is split
    const int w = threadIdx.x >> 6;
    const int item = ((int)blockIdx.y * (int)gridDim.x + (int)blockIdx.x) * (NTH / 64) + w;
    // few loads in flight in the (rare) flagged path, so the column pass keeps its registers
    if (item < tf.p.nparts * tf.nt)
      thr_wave<T, 4, 2>(tf.p, item % tf.p.nparts, item / tf.p.nparts, threadIdx.x & 63,
                  reinterpret_cast<double*>(lds) + w * THR_LDS, tf.Ft, tf.psum, tf.pmax, tf.mu, tf.cand, tf.ncand);
    return;
  }
  const int t = blockIdx.y - tf.rows;
  int slot, N;
  double het1, sgnlt1;
  if constexpr (UNI) {
    slot = sb.uslot;
    N = sb.uN;
    if (sb.tab) {
      het1 = tab.het1[t];
      sgnlt1 = tab.sgnlt1[t];
    } else {
      het1 = ld_global(&desc[t].het1, 0);
      sgnlt1 = ld_global(&desc[t].sgnlt1, 0);
    }
  } else if (sb.tab) {
    slot = tab.slot[t];
    N = tab.slotN[slot];
    het1 = tab.het1[t];
    sgnlt1 = tab.sgnlt1[t];
  } else {
    const TmplDesc& dsc = desc[t];
    slot = dsc.slot;
    N = dsc.N;
    het1 = dsc.het1;
    sgnlt1 = dsc.sgnlt1;
  }
  const int nifo = sb.nifo;
  // the slot's series xm[d][{a,b}][N] and _tmp1[d][N] as raw buffers: lane offset n, uniform
  // offset d N (at most 2^28 bytes: nifo <= 8, N <= 2^20)
  const BufRes xmb(static_cast<const T2*>(sb.xm) + (size_t)slot * sb.xm_stride, (uint32_t)(2 * nifo * N * sizeof(T2)));
  const BufRes t1b(sb.tmp1 + (size_t)slot * sb.tmp1_stride, (uint32_t)(nifo * N * sizeof(double)));
  const int c = threadIdx.x % S::C, g = threadIdx.x / S::C;
  const int n1 = blockIdx.x * S::C + c;
  const int L1 = p.L1;
  const size_t L = (size_t)L1 * P2;
  T2 x[2][E];
  // the demodulation of register row r: REV (FP32 only) accumulates the phase in revolutions
  auto demod = [&](auto rev_tag) {
    constexpr bool REV = decltype(rev_tag)::value;
    [[maybe_unused]] const PhaseRev pr(het1, sgnlt1);
#pragma unroll
    for (int r = 0; r < E; r++) {
      const int n = n1 + L1 * (g + S::G * r);
      T2 xa = mk<T2>(0, 0), xb = mk<T2>(0, 0);
      if (r < NZ && n < N) {
        const double dn = (double)n;
        const uint32_t vx = (uint32_t)n * (uint32_t)sizeof(T2), vt = (uint32_t)n * (uint32_t)sizeof(double);
#pragma unroll 2
        for (int d = 0; d < nifo; d++) {  // two detectors' loads in flight together
          const uint32_t dN = (uint32_t)d * (uint32_t)N;
          // phase = het1 i + sgnlt1 _tmp1 (jobcore.c:530), uncontracted; _tmp1 from the sky slot
          const double tmp1 = t1b.load<double>(vt, dN * (uint32_t)sizeof(double));
          T2 ex;
          if constexpr (REV) {
            const float2 e = pr.eval(dn, tmp1);
            ex = mk<T2>(e.x, e.y);
          }
          else ex = PhaseRot<T>::eval(lin2_nc(het1, dn, sgnlt1, tmp1));
          xa = cmac(xa, xmb.load<T2>(vx, 2 * dN * (uint32_t)sizeof(T2)), ex);
          xb = cmac(xb, xmb.load<T2>(vx, (2 * dN + (uint32_t)N) * (uint32_t)sizeof(T2)), ex);
        }
      }
      x[0][r] = xa;
      x[1][r] = xb;
    }
  };
  auto demod_lf = [&](auto rev_tag) {
    constexpr bool REV = decltype(rev_tag)::value;
    [[maybe_unused]] const PhaseRev pr(het1, sgnlt1);
#pragma unroll
    for (int r = 0; r < E; r++) x[0][r] = x[1][r] = mk<T2>(0, 0);
#pragma unroll 1
    for (int d = 0; d < nifo; d++) {
      const uint32_t dN = (uint32_t)d * (uint32_t)N;
      double tv[NZ];
      T2 av[NZ], bv[NZ];
#pragma unroll
      for (int r = 0; r < NZ; r++) {
        const int nc = min(n1 + L1 * (g + S::G * r), N - 1);
        const uint32_t vx = (uint32_t)nc * (uint32_t)sizeof(T2), vt = (uint32_t)nc * (uint32_t)sizeof(double);
        tv[r] = t1b.load<double>(vt, dN * (uint32_t)sizeof(double));
        av[r] = xmb.load<T2>(vx, 2 * dN * (uint32_t)sizeof(T2));
        bv[r] = xmb.load<T2>(vx, (2 * dN + (uint32_t)N) * (uint32_t)sizeof(T2));
      }
#pragma unroll
      for (int r = 0; r < NZ; r++) {
        const int n = n1 + L1 * (g + S::G * r);
        if (n < N) {
          const double dn = (double)n;
          T2 ex;
          if constexpr (REV) {
            const float2 e = pr.eval(dn, tv[r]);
            ex = mk<T2>(e.x, e.y);
          } else {
            ex = PhaseRot<T>::eval(lin2_nc(het1, dn, sgnlt1, tv[r]));
          }
          x[0][r] = cmac(x[0][r], av[r], ex);
          x[1][r] = cmac(x[1][r], bv[r], ex);
        }
      }
    }
  };
  if constexpr (LF) {
    if (p.rev) demod_lf(std::true_type{});
    else demod_lf(std::false_type{});
  } else if constexpr (sizeof(T) == 4) {
    if (p.rev) demod(std::true_type{});
    else demod(std::false_type{});
  } else {
    demod(std::false_type{});
  }
  // the template's two spectra, 2 L elements (byte offsets fit 32 bits: L <= 2^20).  Output slot s
  // is row k2 = g Q + col_out_k2c(s): lane offset (n1 + g Q L1), uniform offset col_out_k2c(s) L1
  // (FP64 elements, 16 bytes, are stored through a plain pointer at the same offsets)
  T2* const Zt = Z + (size_t)t * 2 * L;
  const BufRes Zb(Zt, (uint32_t)(2 * L * sizeof(T2)));
  const uint32_t voff = (uint32_t)(n1 + g * S::Q * L1) * (uint32_t)sizeof(T2);
  const uint32_t rowb = (uint32_t)L1 * (uint32_t)sizeof(T2), specb = (uint32_t)(L * sizeof(T2));
  auto put = [&](T2 v, uint32_t soff) {
    if constexpr (sizeof(T2) == 8) {
      if (p.nt) Zb.store_nt(v, voff, soff);
      else Zb.store(v, voff, soff);
    } else {
      T2* q = reinterpret_cast<T2*>(reinterpret_cast<char*>(Zt) + voff + soff);
      if (p.nt) {
        using V = double __attribute__((ext_vector_type(2)));
        V w;
        w.x = v.x;
        w.y = v.y;
        __builtin_nontemporal_store(w, reinterpret_cast<V*>(q));
      } else {
        *q = v;
      }
    }
  };
  T2 w[E];
  col_out_twiddles<P2, E, NTH>(tw_big(hi, lo, (uint32_t)n1 << p.twshift), w);
  if constexpr (!SEQ) {
    col_fft<T2, P2, E, 2, -1, NZ, NTH>(x, lds, twc);
#pragma unroll
    for (int s = 0; s < E; s++) {
      const uint32_t k2c = col_out_k2c<P2, E>(s);
      put(cmul(x[0][s], w[s]), k2c * rowb);
      put(cmul(x[1][s], w[s]), specb + k2c * rowb);
    }
  } else {
    // FP32 with 16 rows per thread: the step-1 twiddle powers of the first spectrum serve the
    // second (4 % fewer instructions at the same occupancy; with 8 rows the extra registers cost
    // a wave per SIMD and more time than they save)
    constexpr bool PW = sizeof(T) == 4 && E == 16;
    T2 pw[PW ? E : 1];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      T2 (&xa)[1][E] = *reinterpret_cast<T2(*)[1][E]>(&x[a]);
      if (a) __syncthreads();
      col_fft<T2, P2, E, 1, -1, NZ, NTH>(xa, lds, twc, PW ? pw : nullptr, a > 0);
#pragma unroll
      for (int s = 0; s < E; s++) {
        put(cmul(xa[0][s], w[s]), (uint32_t)a * specb + (uint32_t)col_out_k2c<P2, E>(s) * rowb);
      }
    }
  }
}

// Row pass + F-statistic + FStat partials (jobcore.c:629-651, auxi.c:214-240): grid (L2, ntemplates),
// P1/16 threads.  F(k), k = L2 k1 + k2 in [nmin, nmax), at Ft[t][k2][k1]; psum/pmax[t][part][k2].
// NA = 2 transforms both spectra together; NA = 1 one after the other through the same LDS.
// minimum waves per SIMD for the row pass (register budget 512 / W VGPRs per lane), and whether
// the second spectrum's row is loaded into registers before the first transform (NA = 1)
template <typename T, int P1, int NA> struct RowOcc {
  static constexpr int W = 1;
  static constexpr bool PREF = true;
};
#ifndef PGW_ROW32_W
#define PGW_ROW32_W 1
#endif
template <int P1> struct RowOcc<float, P1, 1> {
  static constexpr int W = PGW_ROW32_W;
  static constexpr bool PREF = true;
};
template <int P1> struct RowOcc<double, P1, 1> {
  static constexpr int W = PGW_ROW64_W;
  static constexpr bool PREF = PGW_ROW64_PREF;
};
// HALF: nmax <= L/2, so only k1 < P1/2 (the thread's outputs m < 8) can hold a bin of
// [nmin, nmax); the other half of the last FFT stage is never formed.
template <typename T, int P1, int NA, bool HALF>
__global__ __launch_bounds__(P1 / 16, (RowOcc<T, P1, NA>::W)) void k_spin_row2(RowParams p, const double* __restrict__ aabb_slots,
                                                      const TmplDesc* __restrict__ desc,
                                                      const cplx<T>* __restrict__ Z,
                                                      const cplx<T>* __restrict__ twr, T* __restrict__ Ft,
                                                      T* __restrict__ psum, T* __restrict__ pmax) {
  using T2 = cplx<T>;
  using S = RowShape<P1>;
  constexpr int NT = S::NT;
  __shared__ T2 lds[NA * S::PADN];
  // templates in reverse launch order: the column pass wrote the last templates' spectra last, so
  // the first row workgroups find them still in the last-level cache (MALL)
  const int k2 = blockIdx.x, t = gridDim.y - 1 - blockIdx.y, tid = threadIdx.x;
  const size_t L = (size_t)P1 * p.L2;
  // rows k2 of the template's two spectra and of its F, as raw buffers: lane offset tid, uniform
  // offset k2 P1 + NT m (at most 2^25 bytes: L <= 2^20)
  const BufRes Zr(Z + (size_t)t * 2 * L, (uint32_t)(2 * L * sizeof(T2)));
  const uint32_t vr = (uint32_t)tid * (uint32_t)sizeof(T2), mb = NT * (uint32_t)sizeof(T2);
  const uint32_t ra = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T2), rb = ra + (uint32_t)(L * sizeof(T2));
  auto ldz = [&](uint32_t off) { return p.nt ? Zr.load<T2, 2>(vr, off) : Zr.load<T2>(vr, off); };
  const double* aabb = aabb_slots + 2 * row_slot(p, desc, t);
  // F = |xa|^2/aa + |xb|^2/bb (jobcore.c:635-638) with the two reciprocals taken once: each
  // product differs from the quotient by at most one rounding
  const T raa = (T)(1.0 / ld_global(aabb, 0)), rbb = (T)(1.0 / ld_global(aabb, 1));
  constexpr int MO = HALF ? 8 : 16;  // output slots that can hold a bin of [nmin, nmax)
  T F[MO];
  if constexpr (NA == 2) {
    T2 x[2][16];
#pragma unroll
    for (int m = 0; m < 16; m++) { x[0][m] = ldz(ra + mb * m); x[1][m] = ldz(rb + mb * m); }
    row_fft<T2, P1, 2, -1>(x, lds, twr);
#pragma unroll
    for (int m = 0; m < MO; m++)
      F[m] = (x[0][m].x * x[0][m].x + x[0][m].y * x[0][m].y) * raa + (x[1][m].x * x[1][m].x + x[1][m].y * x[1][m].y) * rbb;
  } else {
    constexpr bool PREF = RowOcc<T, P1, NA>::PREF;
    T2 x[1][16], y[1][PREF ? 16 : 1];
#pragma unroll
    for (int m = 0; m < 16; m++) x[0][m] = ldz(ra + mb * m);
    if constexpr (PREF) {
#pragma unroll
      for (int m = 0; m < 16; m++) y[0][m] = ldz(rb + mb * m);  // in flight during the first FFT
    }
    // FP32: the stage-C twiddle powers of the first transform serve the second (no more
    // registers at the peak: they live while the prefetched row no longer does)
    constexpr bool PC = sizeof(T) == 4;
    T2 pc[PC ? 16 : 1];
    // and the stage-B powers too (-3 % VALU instructions), except for rows of 2048, whose kernel
    // fits 128 VGPRs (four waves per SIMD) only without them
#ifndef PGW_ROW_PB
#define PGW_ROW_PB 1
#endif
    constexpr bool PB = PC && P1 != 2048 && PGW_ROW_PB;
    T2 pb[PB ? 16 : 1];
    T2* pbp = PB ? pb : nullptr;
    row_fft<T2, P1, 1, -1>(x, lds, twr, PC ? pc : nullptr, false, pbp);
#pragma unroll
    for (int m = 0; m < MO; m++) F[m] = (x[0][m].x * x[0][m].x + x[0][m].y * x[0][m].y) * raa;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; m++) x[0][m] = PREF ? y[0][PREF ? m : 0] : ldz(rb + mb * m);
    row_fft<T2, P1, 1, -1>(x, lds, twr, PC ? pc : nullptr, true, pbp);
#pragma unroll
    for (int m = 0; m < MO; m++) F[m] = F[m] + (x[0][m].x * x[0][m].x + x[0][m].y * x[0][m].y) * rbb;
  }
  __syncthreads();
  T* Fs = reinterpret_cast<T*>(lds);
  const BufRes Fb(Ft + (size_t)t * L, (uint32_t)(L * sizeof(T)));
  const uint32_t vf = (uint32_t)tid * (uint32_t)sizeof(T), fo = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T);
  auto stf = [&](T v, uint32_t soff) {
    if (p.fnt) Fb.template store<2>(v, vf, soff);
    else Fb.store(v, vf, soff);
  };
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k1 = tid + NT * m;
    if (m < MO) {
      const int k = p.L2 * k1 + k2;
      const bool ok = (k >= p.nmin && k < p.nmax);
      if (ok) stf(F[m < MO ? m : 0], fo + (uint32_t)(NT * m * sizeof(T)));
      Fs[k1] = ok ? F[m < MO ? m : 0] : (T)0;
    }
    // (HALF, m >= MO: the upper half of Fs, k1 >= P1/2, holds no bin of [nmin, nmax) and is not read)
  }
  __syncthreads();
  // FStat block partials of this row: block b holds bins k in [nmin + 4096 b, nmin + 4096 (b+1))
  // (the last one up to nmax), i.e. 4096 / L2 >= 16 consecutive k1.  Thread tid sums its 16
  // consecutive bins j = 16 tid + i, which fall into at most two blocks b0 <= b1 (bins outside
  // [nmin, nmax) are 0 and clamped into the first or last block), ...
  T* rs = Fs + P1;   // [2][NT] sums, then [2][NT] maxima
  // HALF: threads whose 16 bins lie in the upper half (tid >= NT/2) have nothing to sum, and no
  // block's bin range reaches their entries of rs (its bins end below nmax <= L/2)
  if (!HALF || tid < NT / 2) {
    const int j0 = 16 * tid;
    const int nb = p.nparts - 1;
    const int b0 = min(max((p.L2 * j0 + k2 - p.nmin) >> 12, 0), nb);
    // first bin of block b0 + 1 (none when b0 is the last block)
    const int jsplit = b0 == nb ? P1 : (PGW_NAV_DEV * (b0 + 1) + p.nmin - k2 + p.L2 - 1) >> p.lgL2;
    T s0 = 0, s1 = 0, m0 = 0, m1 = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      // rotated start (by tid and tid / 16, so lanes 16 apart differ too): lanes spread over the
      // LDS banks
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
  // ... and one thread per block adds the threads whose bins cover it, in thread order.
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
    psum[((size_t)t * p.nparts + part) * p.L2 + k2] = sum;
    pmax[((size_t)t * p.nparts + part) * p.L2 + k2] = mx;
  }
}

// Row pass with 8 elements per thread (row_fft8; P1/8 threads, two workgroups per CU within 128
// VGPRs): the FP64 form of k_spin_row2, same outputs.  The spectra are transformed in turn through
// one LDS array, the second row's loads in flight during the first transform (PREF).
// WL: row_fft8w (one block-wide exchange per transform, the others inside each wave; 6 workgroup
// barriers per row instead of 14); its outputs are a permutation of the threads', so F reaches
// global memory through LDS (padded by one element per P1/64: conflict-free scattered writes).
template <typename T, int P1, bool HALF, bool PREF, bool WL = false>
__global__ __launch_bounds__(P1 / 8, 4) void k_spin_row8(RowParams p, const double* __restrict__ aabb_slots,
                                                         const TmplDesc* __restrict__ desc,
                                                         const cplx<T>* __restrict__ Z,
                                                         const cplx<T>* __restrict__ twr, T* __restrict__ Ft,
                                                         T* __restrict__ psum, T* __restrict__ pmax) {
  using T2 = cplx<T>;
  using S = Row8Shape<P1>;
  constexpr int NT = S::NT;
  __shared__ T2 lds[WL ? Row8wShape<P1>::LDSN : S::PADN];
  // templates in reverse launch order: the column pass wrote the last templates' spectra last, so
  // the first row workgroups find them still in the last-level cache (MALL)
  const int k2 = blockIdx.x, t = gridDim.y - 1 - blockIdx.y, tid = threadIdx.x;
  const size_t L = (size_t)P1 * p.L2;
  // rows k2 of the template's two spectra and of its F, as raw buffers: lane offset tid, uniform
  // offset k2 P1 + NT m (at most 2^25 bytes: L <= 2^20)
  const BufRes Zr(Z + (size_t)t * 2 * L, (uint32_t)(2 * L * sizeof(T2)));
  const uint32_t vr = (uint32_t)tid * (uint32_t)sizeof(T2), mb = NT * (uint32_t)sizeof(T2);
  const uint32_t ra = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T2), rb = ra + (uint32_t)(L * sizeof(T2));
  auto ldz = [&](uint32_t off) { return p.nt ? Zr.load<T2, 2>(vr, off) : Zr.load<T2>(vr, off); };
  const double* aabb = aabb_slots + 2 * row_slot(p, desc, t);
  const T raa = (T)(1.0 / ld_global(aabb, 0)), rbb = (T)(1.0 / ld_global(aabb, 1));
  constexpr int MO = HALF ? 4 : 8;  // output slots that can hold a bin of [nmin, nmax)
  T F[MO];
  T2 x[8], y[PREF ? 8 : 1];
#pragma unroll
  for (int m = 0; m < 8; m++) x[m] = ldz(ra + mb * m);
  if constexpr (PREF) {
#pragma unroll
    for (int m = 0; m < 8; m++) y[m] = ldz(rb + mb * m);
  }
  auto fft = [&]() {
    if constexpr (WL) row_fft8w<T2, P1, -1>(x, lds, twr);
    else row_fft8<T2, P1, -1>(x, lds, twr);
  };
  fft();
#pragma unroll
  for (int m = 0; m < MO; m++) F[m] = (x[m].x * x[m].x + x[m].y * x[m].y) * raa;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 8; m++) x[m] = PREF ? y[PREF ? m : 0] : ldz(rb + mb * m);
  fft();
#pragma unroll
  for (int m = 0; m < MO; m++) F[m] = F[m] + (x[m].x * x[m].x + x[m].y * x[m].y) * rbb;
  __syncthreads();
  T* Fs = reinterpret_cast<T*>(lds);
  // Fs position of row bin k1 (WL: one padding element per P1/64)
  auto fpos = [](int k1) { return WL ? k1 + k1 / (P1 / 64) : k1; };
  const BufRes Fb(Ft + (size_t)t * L, (uint32_t)(L * sizeof(T)));
  const uint32_t vf = (uint32_t)tid * (uint32_t)sizeof(T), fo = (uint32_t)(k2 * P1) * (uint32_t)sizeof(T);
  auto stf = [&](T v, uint32_t soff) {
    if (p.fnt) Fb.template store<2>(v, vf, soff);
    else Fb.store(v, vf, soff);
  };
  if constexpr (WL) {
    // register m holds bin k0 + NT m: through LDS to the thread that stores bin tid + NT m
    const int k0 = row8w_k0<P1>(tid);
#pragma unroll
    for (int m = 0; m < MO; m++) {
      const int k1 = k0 + NT * m;
      const int k = p.L2 * k1 + k2;
      Fs[fpos(k1)] = (k >= p.nmin && k < p.nmax) ? F[m] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MO; m++) {
      const int k1 = tid + NT * m;
      const int k = p.L2 * k1 + k2;
      if (k >= p.nmin && k < p.nmax) stf(Fs[fpos(k1)], fo + (uint32_t)(NT * m * sizeof(T)));
    }
    // (the partials below read Fs and write rs, which the stores above do not touch: no barrier)
  } else {
#pragma unroll
    for (int m = 0; m < 8; m++) {
      const int k1 = tid + NT * m;
      if (m < MO) {
        const int k = p.L2 * k1 + k2;
        const bool ok = (k >= p.nmin && k < p.nmax);
        if (ok) stf(F[m < MO ? m : 0], fo + (uint32_t)(NT * m * sizeof(T)));
        Fs[k1] = ok ? F[m < MO ? m : 0] : (T)0;
      }
      // (HALF, m >= MO: the upper half of Fs, k1 >= P1/2, holds no bin of [nmin, nmax) and is not read)
    }
    __syncthreads();
  }
  // FStat block partials of this row as in k_spin_row2, with runs of 8 consecutive bins per thread
  // (a block spans 4096 / L2 >= 16 consecutive k1, so a run falls into at most two blocks).
  T* rs = Fs + fpos(P1);   // [2][NT] sums, then [2][NT] maxima
  // HALF: threads whose 8 bins lie in the upper half (tid >= NT/2) have nothing to sum, and no
  // block's bin range reaches their entries of rs (its bins end below nmax <= L/2)
  if (!HALF || tid < NT / 2) {
    const int j0 = 8 * tid;
    const int nb = p.nparts - 1;
    const int b0 = min(max((p.L2 * j0 + k2 - p.nmin) >> 12, 0), nb);
    const int jsplit = b0 == nb ? P1 : (PGW_NAV_DEV * (b0 + 1) + p.nmin - k2 + p.L2 - 1) >> p.lgL2;
    T s0 = 0, s1 = 0, m0 = 0, m1 = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int j = j0 + ((i + tid) & 7);  // rotated start: lanes spread over the LDS banks
      const T f = Fs[fpos(j)];
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
      for (int u = jlo >> 3; u <= (jhi - 1) >> 3; u++) {
        const int b0 = min(max((p.L2 * (8 * u) + k2 - p.nmin) >> 12, 0), nb);
        const int h = b0 == part ? 0 : 1;
        sum += rs[h * NT + u];
        const T v = rs[(2 + h) * NT + u];
        mx = v > mx ? v : mx;
      }
    }
    psum[((size_t)t * p.nparts + part) * p.L2 + k2] = sum;
    pmax[((size_t)t * p.nparts + part) * p.L2 + k2] = mx;
  }
}

// Generic column pass (one array per block): grid (L1/C, nbatch).  OTW: the four-step's output
// twiddle W_L^{n1 k2} (false: the column transform alone, for a transform whose twiddles the
// loader applies to its input)
template <typename T, int P2, int DIR, typename Loader, bool OTW = true>
__global__ __launch_bounds__(256) void k_fft_col2(Loader ld, cplx<T>* __restrict__ Z, int L1,
                                                  const cplx<T>* __restrict__ twc, const cplx<T>* __restrict__ hi,
                                                  const cplx<T>* __restrict__ lo, int twshift) {
  using T2 = cplx<T>;
  constexpr int E = ColE<T, P2>::E;
  using S = ColShape<P2, E>;
  __shared__ T2 lds[P2 * S::C];
  const int b = blockIdx.y;
  const int c = threadIdx.x % S::C, g = threadIdx.x / S::C;
  const int n1 = blockIdx.x * S::C + c;
  const size_t L = (size_t)L1 * P2;
  T2 x[1][E];
#pragma unroll
  for (int r = 0; r < E; r++) x[0][r] = ld(b, n1 + L1 * (g + S::G * r));
  T2 w[E];
  if constexpr (OTW) col_out_twiddles<P2, E>(tw_dir<DIR>(tw_big(hi, lo, (uint32_t)n1 << twshift)), w);
  col_fft<T2, P2, E, 1, DIR>(x, lds, twc);
#pragma unroll
  for (int s = 0; s < E; s++) {
    const int k2 = col_out_k2<P2, E>(s);
    Z[(size_t)b * L + (size_t)k2 * L1 + n1] = OTW ? cmul(x[0][s], w[s]) : x[0][s];
  }
}

// Generic row pass, in place: grid (L2, nbatch), P1/16 threads.
template <typename T, int P1, int DIR>
__global__ __launch_bounds__(P1 / 16) void k_fft_row2(cplx<T>* __restrict__ Z, int L2, const cplx<T>* __restrict__ twr) {
  using T2 = cplx<T>;
  using S = RowShape<P1>;
  __shared__ T2 lds[S::PADN];
  T2* row = Z + (size_t)blockIdx.y * P1 * L2 + (size_t)blockIdx.x * P1;
  T2 x[1][16];
#pragma unroll
  for (int m = 0; m < 16; m++) x[0][m] = row[threadIdx.x + S::NT * m];
  row_fft<T2, P1, 1, DIR>(x, lds, twr);
#pragma unroll
  for (int m = 0; m < 16; m++) row[threadIdx.x + S::NT * m] = x[0][m];
}

// Resampling's odd samples, first half (jobcore.c:334-352): on each row k2 of the forward
// spectrum in T-layout (X[L2 k1 + k2] at k1 of row k2), the forward row transform, then X[k] s_k
// e^{i pi k / L} with k = L2 k1 + k2, of which the factor e^{i pi k1 / L1} (and the sign s_k = -1 for
// k >= nyqst) is applied here and e^{i pi k2 / L}, constant along the row, by the column pass's
// loader (LoadOddTw), then the inverse row transform over k1: the inverse transform's first step,
// in place.  grid (L2, nbatch), P1/16 threads; shift1 = 20 - log2(2 P1).
template <typename T, int P1>
__global__ __launch_bounds__(P1 / 16) void k_fft_row_shift(cplx<T>* __restrict__ Z, int L2, const cplx<T>* __restrict__ twr,
                                                           const cplx<T>* __restrict__ hi, const cplx<T>* __restrict__ lo,
                                                           int shift1, int nyqst) {
  using T2 = cplx<T>;
  using S = RowShape<P1>;
  __shared__ T2 lds[S::PADN];
  const int k2 = blockIdx.x, t = threadIdx.x;
  T2* row = Z + (size_t)blockIdx.y * P1 * L2 + (size_t)k2 * P1;
  T2 x[1][16];
#pragma unroll
  for (int m = 0; m < 16; m++) x[0][m] = row[t + S::NT * m];
  row_fft<T2, P1, 1, -1, 2>(x, lds, twr);
#pragma unroll
  for (int m = 0; m < 16; m++) {
    const int k1 = t + S::NT * m;
    const T2 v = cmul(x[0][m], conjg(tw_big(hi, lo, (uint32_t)k1 << shift1)));
    x[0][m] = L2 * k1 + k2 < nyqst ? v : mk<T2>(-v.x, -v.y);
  }
  __syncthreads();  // the forward transform's last LDS reads are done
  row_fft<T2, P1, 1, 1, 2>(x, lds, twr);
#pragma unroll
  for (int m = 0; m < 16; m++) row[t + S::NT * m] = x[0][m];
}

}  // namespace pgw
