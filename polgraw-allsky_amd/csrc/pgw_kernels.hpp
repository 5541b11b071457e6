// pgw_kernels.hpp - HIP kernels of the coherent F-statistic search core (gfx950).
//
// Sky-point stage (once per sky point, amortised over its spindowns; each launch covers all the
// sky points of a group):
//   k_sky_prologue    modvir + shft, _tmp1 + first demodulation (jobcore.c:270-313, settings.c:281-325)
//   k_fft_col2/k_fft_row2 + loaders  resampling FFTs         (jobcore.c:319-362)
//   k_untranspose     T-layout -> natural order, times interpftpad/Ninterp
//   k_spline          natural cubic spline (auxi.c:55-81) and the noise-weighted sums aa, bb
//                     (jobcore.c:375-393); k_spline_eval: splint (auxi.c:83-122), /sig2
// Template stage (per spindown, the hot loop; pgw_kernels2.hpp):
//   k_spin_col2       phase modulation fused into the column pass of the F-statistic FFT
//                     (jobcore.c:456-627)
//   k_spin_row2/8     row pass + F = |xa|^2/aa + |xb|^2/bb + FStat block partials (jobcore.c:629-651)
//   k_threshold       FStat means, normalisation, run-max threshold, candidates (jobcore.c:659-695);
//                     in a search its waves run in the first grid rows of the next launch's
//                     column pass (thr_wave), only a group's last launch has a launch of its own
#pragma once
#include "pgw_common.hpp"

namespace pgw {

constexpr int NT = 256;          // threads per block everywhere
constexpr int PGW_NAV_DEV = 4096; // FStat block length (NAVFSTAT, auxi.h:13)
constexpr int SPL_W = 32;        // spline warm-up (|coefficient| ~ 0.268: 0.268^32 ~ 5e-19)
constexpr int SPL_NTAB = 64;     // spline recursion coefficients tabulated below this index
constexpr int SPL_SB = 8;        // spline steps per scheduling block (LDS reads issued ahead within one)

struct SkyDet {
  double c[9];
  double cosalfr, sinalfr, c2d, c2sd, sindel, cosdel;
  double nS0, nS1, nS2, shft1, sig2;
};

struct SkyParams {
  SkyDet det[8];
  double het0, oms;
  int N, nfft, nifo;
};

// One sky point of a batched sky stage: its coefficients, its segment's inputs and the sky slot
// that receives its resampled series xm, _tmp1 and aa/bb.  Sky points of one batch share
// nfft, Ninterp and the detector count; N may differ.
struct SkyItem {
  SkyParams p;
  const double* xDat;     // [nifo][N]
  const double* DetSSB;   // [nifo][3N]
  const double* cosmodf;  // [N]
  const double* sinmodf;  // [N]
  void* xm;               // slot series cplx<T>[nifo][{a,b}][N]
  double* tmp1;           // slot _tmp1 = i^2 + 2 i shft [nifo][N]
  double* aabb;           // slot aa, bb
  int nspl, pad;          // spline knots interpftpad N
};

// Expressions evaluated in the reference's operation order without fused multiply-add
// (HIP compiles with -ffp-contract=fast-honor-pragmas, so the pragma holds after inlining).
__device__ __forceinline__ double dot3_nc(double a0, double b0, double a1, double b1, double a2, double b2) {
#pragma clang fp contract(off)
  return a0 * b0 + a1 * b1 + a2 * b2;
}
__device__ __forceinline__ double lin2_nc(double a0, double b0, double a1, double b1) {
#pragma clang fp contract(off)
  return a0 * b0 + a1 * b1;
}

// ------------------------------------------------------------- sky stage --

// grid (ceil(nfft/NT), nifo, sky points).  For sky point j: R[j][2d][i] = xDat aa e^{-i phase},
// R[j][2d+1][i] = xDat bb e^{-i phase} for i < N, zero for N <= i < nfft; _tmp1 (slot), shftf
// [j][d][Nstride]; block partial sums of aa^2, bb^2 at paa/pbb [j][d][block].
template <typename T>
// zero (optional): two ints cleared by the first workgroup (the group's candidate count and
// error word, which the template stage of the same group counts up later in the stream)
__global__ __launch_bounds__(NT) void k_sky_prologue(const SkyItem* __restrict__ items, cplx<T>* __restrict__ R,
                                                     double* __restrict__ shftf, size_t Nstride,
                                                     double* __restrict__ paa, double* __restrict__ pbb,
                                                     int* __restrict__ zero) {
  using T2 = cplx<T>;
  const int d = blockIdx.y, j = blockIdx.z;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (zero && (blockIdx.x | blockIdx.y | blockIdx.z) == 0 && threadIdx.x < 2) zero[threadIdx.x] = 0;
  const SkyItem& it = items[j];
  const SkyParams& p = it.p;
  const SkyDet& q = p.det[d];
  const double* __restrict__ xDat = it.xDat;
  const double* __restrict__ DetSSB = it.DetSSB;
  const double* __restrict__ cosmodf = it.cosmodf;
  const double* __restrict__ sinmodf = it.sinmodf;
  double* __restrict__ tmp1 = it.tmp1;
  shftf += (size_t)j * p.nifo * Nstride;
  R += (size_t)j * 2 * p.nifo * p.nfft;
  paa += (size_t)j * p.nifo * gridDim.x;
  pbb += (size_t)j * p.nifo * gridDim.x;
  double a2 = 0.0, b2 = 0.0;
  T2 va = mk<T2>(0, 0), vb = mk<T2>(0, 0);
  if (i < p.N) {
    const double cm = cosmodf[i], sm = sinmodf[i];
    // modvir, settings.c:309-322
    const double c = q.cosalfr * cm + q.sinalfr * sm;
    const double s = q.sinalfr * cm - q.cosalfr * sm;
    const double c2s = 2. * (c * c), cs = c * s;
    const double aa = q.c[0] * (2. - q.c2d) * c2s + q.c[1] * (2. - q.c2d) * 2. * cs + q.c[2] * q.c2sd * c +
                      q.c[3] * q.c2sd * s - q.c[0] * (2. - q.c2d) + q.c[4] * q.c2d;
    const double bb = q.c[5] * q.sindel * c2s + q.c[6] * q.sindel * 2. * cs + q.c[7] * q.cosdel * c +
                      q.c[8] * q.cosdel * s - q.c[5] * q.sindel;
    const double* D = DetSSB + (size_t)d * 3 * p.N + 3 * (size_t)i;
    // shft = n.DetSSB and the phase in the reference's operation order, no contraction:
    // shftf[0] must come out exactly 0 (the spline returns 0 for x < 0)
    const double sh = dot3_nc(q.nS0, D[0], q.nS1, D[1], q.nS2, D[2]);
    // _tmp1 = i^2 + 2 i shft, once per sky point as the reference does (jobcore.c:293), uncontracted
    tmp1[(size_t)d * p.N + i] = lin2_nc((double)i, (double)i, (double)(2 * i), sh);
    shftf[(size_t)d * Nstride + i] = sh - q.shft1;
    double sp, cp;
    sincos(lin2_nc(p.het0, (double)i, p.oms, sh), &sp, &cp);
    const double x = xDat[(size_t)d * p.N + i];
    const double xa = x * aa, xb = x * bb;
    va = mk<T2>((T)(xa * cp), (T)(xa * -sp));
    vb = mk<T2>((T)(xb * cp), (T)(xb * -sp));
    a2 = aa * aa;
    b2 = bb * bb;
  }
  if (i < p.nfft) {
    R[(size_t)(2 * d) * p.nfft + i] = va;
    R[(size_t)(2 * d + 1) * p.nfft + i] = vb;
  }
  __shared__ double red[2][NT / 64];
  a2 = wave_sum(a2);
  b2 = wave_sum(b2);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { red[0][w] = a2; red[1][w] = b2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int k = 0; k < NT / 64; k++) { sa += red[0][k]; sb += red[1][k]; }
    paa[(size_t)d * gridDim.x + blockIdx.x] = sa;
    pbb[(size_t)d * gridDim.x + blockIdx.x] = sb;
  }
}

// one block: aa = sum_d (sum_i aa_d^2)/sig2_d, bb likewise (jobcore.c:375-393)
// aa, bb (jobcore.c:375-393): the per-block partial sums of a^2, b^2 of every detector, summed in
// a fixed order and divided by sig2; one workgroup, red = 2 x NT doubles of LDS.
__device__ __forceinline__ void aabb_reduce(const SkyParams& p, const double* __restrict__ paa,
                                            const double* __restrict__ pbb, int nb, double* __restrict__ aabb,
                                            double (*red)[NT]) {
  double aa = 0.0, bb = 0.0;
  for (int d = 0; d < p.nifo; d++) {
    double sa = 0.0, sb = 0.0;
    for (int k = threadIdx.x; k < nb; k += NT) { sa += paa[(size_t)d * nb + k]; sb += pbb[(size_t)d * nb + k]; }
    red[0][threadIdx.x] = sa;
    red[1][threadIdx.x] = sb;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) { red[0][threadIdx.x] += red[0][threadIdx.x + s]; red[1][threadIdx.x] += red[1][threadIdx.x + s]; }
      __syncthreads();
    }
    aa += red[0][0] / p.det[d].sig2;
    bb += red[1][0] / p.det[d].sig2;
    __syncthreads();
  }
  if (threadIdx.x == 0) { aabb[0] = aa; aabb[1] = bb; }
}

// ----------------------------------------------------- generic 2-pass FFT --

// Plain loader: x[b][n] for n < len, zero beyond.
template <typename T2> struct LoadPlain {
  const T2* src;
  int len, stride;
  __device__ T2 operator()(int b, int n) const {
    return n < len ? src[(size_t)b * stride + n] : mk<T2>(0, 0);
  }
};

// Zero-insert loader for the inverse resampling FFT (jobcore.c:334-352): reads the forward
// spectrum X (length nf, T-layout with row length L1f and L2f rows) as
//   Y[n] = X[n] (n < nyqst), 0 (nyqst <= n < nyqst + ni - nf), X[n - (ni - nf)] (beyond).
template <typename T2> struct LoadZeroInsert {
  const T2* src;
  int nf, ni, nyqst, L1f, L2f;
  __device__ T2 operator()(int b, int n) const {
    int k;
    if (n < nyqst) k = n;
    else if (n < nyqst + ni - nf) return mk<T2>(0, 0);
    else k = n - (ni - nf);
    return src[(size_t)b * nf + (size_t)(k % L2f) * L1f + k / L2f];
  }
};

// Odd samples of the two-fold interpolation (jobcore.c:334-352) without the 2 nf-point transform:
// at 2m+1 the inverse transform of the zero-inserted spectrum is the nf-point inverse transform of
// s_k e^{i pi k / nf} X[k] (s_k = -1 for k >= nyqst: those bins sit above the inserted zeros, at
// frequency k - nf), and at 2m it is nf times the input series.  That inverse transform runs as
// row transforms over k1 of the spectrum's T-layout rows k2 (k = L2 k1 + k2; k_fft_row_shift),
// then twiddles and column transforms over k2 whose outputs are in natural order.  This loader
// feeds the column pass: element n = n1 + L1 k2 of the row-transformed array times
// W_L^{-n1 k2} e^{i pi k2 / L} = conj(W_{2L}^{(2 n1 + 1) k2}) from the shared twiddle tables
// (shift = 20 - log2(2 L)).
template <typename T2> struct LoadOddTw {
  const T2* src;
  const T2* hi;
  const T2* lo;
  int L, lgL1, shift;
  __device__ T2 operator()(int b, int n) const {
    const uint32_t n1 = (uint32_t)n & ((1u << lgL1) - 1u), k2 = (uint32_t)n >> lgL1;
    const uint32_t m = ((2u * n1 + 1u) * k2) & (2u * (uint32_t)L - 1u);
    return cmul(src[(size_t)b * L + n], conjg(tw_big(hi, lo, m << shift)));
  }
};

// T-layout (rows of L1, L2 rows) -> natural order y[b][j] = scale * X[j], j < count.
// grid (ceil(count/ (32*L2)) ... ) handled as tiles of 32 x 32 over the [L2][L1] matrix.
template <typename T>
__global__ __launch_bounds__(NT) void k_untranspose(const cplx<T>* __restrict__ Z, cplx<T>* __restrict__ y, int L1, int L2,
                                                    int count, T scale) {
  using T2 = cplx<T>;
  __shared__ T2 tile[32][33];
  const int b = blockIdx.z;
  const int k1_0 = blockIdx.x * 32, k2_0 = blockIdx.y * 32;
  const size_t L = (size_t)L1 * L2;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int k2 = k2_0 + r, k1 = k1_0 + tx;
    if (k2 < L2 && k1 < L1) tile[r][tx] = Z[(size_t)b * L + (size_t)k2 * L1 + k1];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int k1 = k1_0 + r, k2 = k2_0 + tx;
    const long j = (long)L2 * k1 + k2;
    if (k2 < L2 && k1 < L1 && j < count) {
      const T2 v = tile[tx][r];
      y[(size_t)b * L + j] = mk<T2>(v.x * scale, v.y * scale);
    }
  }
}

// ----------------------------------------------------------------- spline --
// Natural cubic spline on n unit-spaced knots (auxi.c:55-81), chunked: each thread owns K
// outputs (4, 8 or 16) and starts its recursion SPL_W samples early from a zero state (exact when the
// chunk is within SPL_W of the true start).  invp_i, c_i = -0.5 invp_i tabulated for i < SPL_NTAB.
struct SplineTab {
  double invp[SPL_NTAB];
  double c[SPL_NTAB];
  double invp_inf, c_inf;
};

// Both recursions of the spline for spf_out(K) = NT K consecutive outputs of one series per
// workgroup: the input window is read into LDS with coalesced loads, u (forward, chunks of K + 1
// per thread) and then y2 (backward, chunks of K) run from LDS, each chunk starting SPL_W samples
// early from a zero state, and y2 leaves through LDS with coalesced stores.  grid (ceil(n /
// spf_out(K)), series).  Both recursions run a fixed number of fully unrolled steps whose LDS
// addresses are the lane's base plus a constant, so the reads of a block of SPL_SB steps issue
// ahead of the dependent arithmetic.  Lane strides are odd in elements: K + 1 for the forward
// reads of the unpadded y window, and K + 1 in the u array, which is padded by one element per K
// (spk), for the backward reads of chunks of K.
__host__ __device__ constexpr int spf_out(int K) { return NT * K; }            // y2 outputs per workgroup
__host__ __device__ constexpr int spf_uw(int K) { return NT * K + SPL_W; }     // u values (outputs + warm-up above)
__host__ __device__ constexpr int spf_yw(int K) { return NT * (K + 1) + SPL_W + 2; }  // y samples the forward chunks read
template <int K> __device__ __forceinline__ constexpr int spk(int m) { return m + m / K; }
// coefficient tables of the first wave of a series' first workgroup, indexed along its chains:
// invp at i + SPL_W for i in [-SPL_W, 64 (K + 1)), c at k for k in [0, 64 K + SPL_W)
__host__ __device__ constexpr int spl_ti(int K) { return 64 * (K + 1) + SPL_W; }
__host__ __device__ constexpr int spl_tc(int K) { return 64 * K + SPL_W; }

// Forward recursion of one thread: u on [uf, uf + K + 1) within [A0, U1] into Us (at spk<K>(i - A0)),
// from a zero state at uf - SPL_W.  Yw: y at uf - SPL_W - 1 onwards.  HEAD (first wave of a
// series' first workgroup): invp from Ti (Ti[s] for step s), which is 0 for i < 1, so u stays 0
// below the series' start (u_0 = 0, auxi.c:62), and the tabulated invp_i below SPL_NTAB.
template <bool HEAD, typename T, int K>
__device__ __forceinline__ void spline_fwd(const cplx<T>* Yw, double2* Us, const double* Ti, double invp_inf,
                                           int uf, int A0, int U1) {
  constexpr int KF = K + 1;
  auto yat = [&](int m) {
    const cplx<T> v = Yw[m];
    return make_double2((double)v.x, (double)v.y);
  };
  double2 ym = yat(0), y0 = yat(1);
  double ur = 0.0, ui = 0.0;
#pragma unroll
  for (int s = 0; s < SPL_W + KF; s++) {
    const int i = uf - SPL_W + s;
    const double2 yp = yat(s + 2);
    const double invp = HEAD ? Ti[s] : invp_inf;
    const double dr = ym.x - 2. * y0.x + yp.x, di = ym.y - 2. * y0.y + yp.y;
    ur = (-.5 * ur + 3. * dr) * invp;
    ui = (-.5 * ui + 3. * di) * invp;
    if (s >= SPL_W && i <= U1) Us[spk<K>(i - A0)] = make_double2(ur, ui);
    ym = y0;
    y0 = yp;
    if (s % SPL_SB == SPL_SB - 1) __builtin_amdgcn_sched_barrier(0);  // bounds the reads issued ahead
  }
}

// Backward recursion of one thread: y2 on [a, a + K) into out, from a zero state at
// a + K + SPL_W - 1 (and, if above is given, y2 at a + K).  Uw = Us + spk<K>(a - A0).  HEAD: c_k from
// Tc (Tc[m] for k = a + m).  TAIL (last workgroups of a series): steps above n - 2 keep the state at
// 0 (y2_{n-1} = 0).
template <bool HEAD, bool TAIL, int K>
__device__ __forceinline__ void spline_bwd(const double2* Uw, const double* Tc, double c_inf, int a, int n,
                                           double2 (&out)[K], double2* above = nullptr) {
  double zr = 0.0, zi = 0.0;
#pragma unroll
  for (int s = 0; s < SPL_W + K; s++) {
    const int m = K + SPL_W - 1 - s;  // offset in the chunk: spk<K>(a - A0 + m) = spk<K>(a - A0) + spk<K>(m)
    const double c = HEAD ? Tc[m] : c_inf;
    const double2 uk = Uw[spk<K>(m)];
    double nr = c * zr + uk.x, ni = c * zi + uk.y;
    if (TAIL && a + m > n - 2) nr = ni = 0.0;
    zr = nr;
    zi = ni;
    if (s >= SPL_W) out[K - 1 - (s - SPL_W)] = make_double2(zr, zi);
    if (above && s == SPL_W - 1) *above = make_double2(zr, zi);  // y2 at a + K (SPL_W - 1 warm-up steps)
    if (s % SPL_SB == SPL_SB - 1) __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename T, int K>
__global__ __launch_bounds__(NT) void k_spline(const cplx<T>* __restrict__ y, size_t ystride,
                                               const SplineTab* __restrict__ tab, double2* __restrict__ y2,
                                               size_t stride, const SkyItem* __restrict__ items, int nb,
                                               const double* __restrict__ paa, const double* __restrict__ pbb,
                                               int npro) {
  constexpr int KF = K + 1;
  static_assert(NT * KF >= NT * K + SPL_W, "forward chunks cover the warm-up values");
  static_assert(64 * KF >= SPL_NTAB + SPL_W && 64 * K >= SPL_NTAB, "tabulated coefficients stay in the first wave");
  constexpr int SPF_OUT = spf_out(K), SPF_UW = spf_uw(K), SPF_YW = spf_yw(K);
  constexpr int TI = spl_ti(K), TC = spl_tc(K);
  __shared__ cplx<T> Ys[SPF_YW];
  __shared__ double2 Us[spk<K>(SPF_UW) + 1];
  __shared__ double tsh[2 * NT > TI + TC ? 2 * NT : TI + TC];  // aa/bb reduction, then coefficient tables
  const int b = blockIdx.y, tid = threadIdx.x;
  const SkyItem& it = items[b / nb];
  // the sky point's aa, bb, by the first workgroup of its first series (saves a launch)
  if (blockIdx.x == 0 && b % nb == 0) {
    const size_t o = (size_t)(b / nb) * it.p.nifo * npro;
    aabb_reduce(it.p, paa + o, pbb + o, npro, it.aabb, reinterpret_cast<double (*)[NT]>(tsh));
  }
  const int n = it.nspl;
  const int A0 = blockIdx.x * SPF_OUT;
  if (A0 >= n) return;  // beyond a shorter sky point's knots (uniform across the workgroup)
  const int A1 = min(A0 + SPF_OUT, n);         // y2 outputs [A0, A1)
  const int U1 = min(A1 + SPL_W - 1, n - 2);   // u needed on [A0, U1]
  const int Ylo = A0 - SPL_W - 1, Yhi = U1 + 1;  // y window [Ylo, Yhi] (Yhi <= n - 1); below 0 only
                                               // in the first workgroup, where invp = 0 voids it
  const cplx<T>* Y = y + (size_t)b * ystride;
  double2* Y2 = y2 + (size_t)b * stride;
  // the window's loads all in flight before the LDS stores (indices clamped to [0, Yhi]; entries
  // beyond Yhi are read only by steps whose results are discarded)
  constexpr int NYL = (SPF_YW + NT - 1) / NT;
  cplx<T> yv[NYL];
#pragma unroll
  for (int r = 0; r < NYL; r++) yv[r] = Y[min(max(Ylo + tid + r * NT, 0), Yhi)];
#pragma unroll
  for (int r = 0; r < NYL; r++)
    if (tid + r * NT < SPF_YW) Ys[tid + r * NT] = yv[r];
  const bool head = A0 == 0;  // uniform across the workgroup
  if (head) {
    // invp_i, c_i (auxi.c:66-72): 0 at and below i = 0, tabulated below SPL_NTAB, the limit above
    for (int j = tid; j < TI; j += NT) {
      const int i = j - SPL_W;
      tsh[j] = i < 1 ? 0.0 : i < SPL_NTAB ? tab->invp[i] : tab->invp_inf;
    }
    for (int k = tid; k < TC; k += NT) tsh[TI + k] = k < SPL_NTAB ? tab->c[k] : tab->c_inf;
  }
  const double invp_inf = tab->invp_inf, c_inf = tab->c_inf;
  __syncthreads();
  // u_i = (-u_{i-1}/2 + 3 (y_{i-1} - 2 y_i + y_{i+1})) invp_i  (auxi.c:64-72); thread t: u on
  // [A0 + t (K+1), A0 + (t+1) (K+1)) within [A0, U1] (NT (K+1) covers the SPL_W - 1 values above
  // the outputs that the backward warm-ups read)
  const bool w0 = head && tid < 64;  // uniform across each wave
  const bool tail = A0 + SPF_OUT + SPL_W - 1 > n - 2;
  {
    const int uf = A0 + tid * KF;
    if (uf <= U1) {
      const cplx<T>* Yw = Ys + tid * KF;  // y at uf - SPL_W - 1 = Ylo + tid KF
      if (w0) spline_fwd<true, T, K>(Yw, Us, tsh + tid * KF, invp_inf, uf, A0, U1);
      else spline_fwd<false, T, K>(Yw, Us, tsh, invp_inf, uf, A0, U1);
    }
  }
  __syncthreads();
  // y2_k = c_k y2_{k+1} + u_k, y2_{n-1} = 0  (auxi.c:73-80), outputs [a, e) of this thread
  const int a = A0 + tid * K, e = min(a + K, A1);
  double2 out[K];
  if (a < A1) {
    const double2* Uw = Us + spk<K>(a - A0);
    if (w0) spline_bwd<true, true, K>(Uw, tsh + TI + tid * K, c_inf, a, n, out);
    else if (tail) spline_bwd<false, true, K>(Uw, tsh, c_inf, a, n, out);
    else spline_bwd<false, false, K>(Uw, tsh, c_inf, a, n, out);
  }
  __syncthreads();  // every u read is done: Us now stages y2
  if (a < A1) {
#pragma unroll
    for (int j = 0; j < K; j++)
      if (a + j < e) Us[spk<K>(a + j - A0)] = out[j];
  }
  __syncthreads();
  for (int j = tid; j < A1 - A0; j += NT) Y2[A0 + j] = Us[spk<K>(j)];
  if (A1 == n && tid == 0) {  // y2_{n-1} = 0 and the guard entry y2_n read when x == n-1
    Y2[n - 1] = make_double2(0.0, 0.0);
    Y2[n] = make_double2(0.0, 0.0);
  }
}

// out[b][i] = splint(y, y2, n, interpftpad (i - shftf[i])) / sig2   (auxi.c:83-120, jobcore.c:386-389)
// batch b = nb j + 2d + {0: a, 1: b} for sky point j; output stored as the slot's xm[d][{a,b}][i].
template <typename T, typename TO>
__global__ __launch_bounds__(NT) void k_spline_eval(const cplx<T>* __restrict__ y, size_t ystride,
                                                    const double2* __restrict__ y2, size_t y2stride,
                                                    const double* __restrict__ shftf, size_t Nstride,
                                                    int interpftpad, const SkyItem* __restrict__ items, int nb) {
  const int b = blockIdx.y, j = b / nb, bs = b % nb, d = bs >> 1;
  const SkyItem& it = items[j];
  const int N = it.p.N;
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= N) return;
  const int n = interpftpad * N;
  const double x = interpftpad * (i - shftf[((size_t)j * it.p.nifo + d) * Nstride + i]);
  double vr = 0.0, vi = 0.0;
  if (!(x < 0 || x > n - 1)) {
    const int klo = (int)floor(x), khi = klo + 1;
    const double A = khi - x, B = x - klo;
    const cplx<T>* Y = y + (size_t)b * ystride;
    const double2* Y2 = y2 + (size_t)b * y2stride;
    const double ca = A * A * A - A, cb = B * B * B - B;
    const double2 ylo = make_double2((double)Y[klo].x, (double)Y[klo].y);
    // khi == n only at x == n-1 exactly, where its weight B is 0
    const double2 yhi = khi < n ? make_double2((double)Y[khi].x, (double)Y[khi].y) : make_double2(0.0, 0.0);
    const double2 zlo = Y2[klo], zhi = khi < n ? Y2[khi] : make_double2(0.0, 0.0);
    vr = A * ylo.x + B * yhi.x + (ca * zlo.x + cb * zhi.x) / 6.0;
    vi = A * ylo.y + B * yhi.y + (ca * zlo.y + cb * zhi.y) / 6.0;
  }
  const double sig2 = it.p.det[d].sig2;
  static_cast<cplx<TO>*>(it.xm)[(size_t)bs * N + i] = mk<cplx<TO>>((TO)(vr / sig2), (TO)(vi / sig2));
}

// The resampled series' last three steps in one kernel: the inverse FFT's T-layout output Zi (L2
// rows of L1, series b at b * zstride) read as the spline's y window with the scale ft applied
// (k_untranspose; HALF: Zi holds the odd samples only, in natural order (LoadOddTw), and the even
// ones are the input series R), both spline recursions (k_spline), and the evaluation at interpftpad (i -
// shftf[i]) divided by sig2 for the outputs i whose klo falls in this workgroup's knots
// (k_spline_eval): y and y2 stay in LDS.  x_i = interpftpad (i - shftf_i) increases with i, so a
// workgroup's outputs are the i in [min{i : x_i >= A0}, min{i : x_i >= A1}) (from 0 in the first
// workgroup, to N in the last); each bound is found by one wave from a guess at A / interpftpad +
// shftf and a 64-wide window around it.  A bound outside its window (the shift would have to change
// by more than 30 samples over the distance of the shift), or an output whose knots are not this
// workgroup's (abscissae that do not increase with i), sets *err and leaves the outputs
// unwritten: the group's search returns PGW_ERR_DEVICE (the three-kernel path, spl_fuse=0, takes
// any abscissae).  y2 at A1 (the knot above the last output knot) comes from the last chunk's
// warm-up.
template <typename T, int K, bool HALF>
__global__ __launch_bounds__(NT) void k_spline_fused(const cplx<T>* __restrict__ Zi, int L1, int lgL2, T ft,
                                                     size_t zstride, const cplx<T>* __restrict__ R, size_t rstride,
                                                     const SplineTab* __restrict__ tab,
                                                     const double* __restrict__ shftf, size_t Nstride, int interpftpad,
                                                     const SkyItem* __restrict__ items, int nb,
                                                     const double* __restrict__ paa, const double* __restrict__ pbb,
                                                     int npro, int* __restrict__ err) {
  constexpr int KF = K + 1;
  static_assert(NT * KF >= NT * K + SPL_W, "forward chunks cover the warm-up values");
  static_assert(64 * KF >= SPL_NTAB + SPL_W && 64 * K >= SPL_NTAB, "tabulated coefficients stay in the first wave");
  constexpr int SPF_OUT = spf_out(K), SPF_UW = spf_uw(K), SPF_YW = spf_yw(K);
  constexpr int TI = spl_ti(K), TC = spl_tc(K);
  __shared__ cplx<T> Ys[SPF_YW];
  __shared__ double2 Us[spk<K>(SPF_UW) + 1];
  __shared__ double tsh[2 * NT > TI + TC ? 2 * NT : TI + TC];
  __shared__ int ibound[2];
  const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int j = b / nb, bs = b % nb, d = bs >> 1;
  const SkyItem& it = items[j];
  if (blockIdx.x == 0 && bs == 0) {
    const size_t o = (size_t)j * it.p.nifo * npro;
    aabb_reduce(it.p, paa + o, pbb + o, npro, it.aabb, reinterpret_cast<double (*)[NT]>(tsh));
  }
  const int n = it.nspl, N = it.p.N;
  const int A0 = blockIdx.x * SPF_OUT;
  if (A0 >= n) return;  // beyond a shorter sky point's knots (uniform across the workgroup)
  const int A1 = min(A0 + SPF_OUT, n);
  const int U1 = min(A1 + SPL_W - 1, n - 2);
  const int Ylo = A0 - SPL_W - 1, Yhi = U1 + 1;
  const double* sh = shftf + ((size_t)j * it.p.nifo + d) * Nstride;
  // output bounds: wave 0 the lower, wave 1 the upper (guess and window loads in flight early)
  if (w < 2) {
    const int A = w == 0 ? A0 : A1;
    const bool fixed = w == 0 ? A0 == 0 : A1 == n;
    if (fixed) {
      if (lane == 0) ibound[w] = w == 0 ? 0 : N;
    } else {
      const int i0 = min(max(A / interpftpad, 0), N - 1);
      const int g = min(max((int)floor((double)A / interpftpad + sh[i0]) - 31, 0), N - 1);
      const int i = g + lane;
      const double x = i < 0 ? -1.0 : i >= N ? 1e300 : interpftpad * (i - sh[i]);
      const unsigned long long bal = __ballot(x >= (double)A);
      if (lane == 0) {
        if ((bal & 1ull) || bal == 0) {  // the bound is not inside the window
          if (err) *err = 2;
          ibound[w] = -1;
        } else {
          ibound[w] = g + __builtin_ctzll(bal);
        }
      }
    }
  }
  // y window on [jlo, Yhi].  HALF: Zi holds the odd samples y at 2m + 1 in natural order (m) and
  // the even samples y at 2m are the input series R itself.  Else from the T-layout: element (r, c)
  // of the block of rows 0..L2-1 and columns [cA, cA + ncol) holding y on [jlo, Yhi], read row by
  // row (contiguous row segments).
  const int jlo = max(Ylo, 0);
  if (HALF) {
    const cplx<T>* Z = Zi + (size_t)b * zstride;
    const cplx<T>* Rb = R + (size_t)b * rstride;
    for (int m = jlo / 2 + tid; m <= (Yhi - 1) / 2; m += NT) {
      const cplx<T> v = Z[m];
      Ys[2 * m + 1 - Ylo] = mk<cplx<T>>(v.x * ft, v.y * ft);
    }
    for (int m = (jlo + 1) / 2 + tid; m <= Yhi / 2; m += NT) Ys[2 * m - Ylo] = Rb[m];
  } else {
    const int L2 = 1 << lgL2;
    const int cA = jlo >> lgL2, ncol = (Yhi >> lgL2) - cA + 1, nel = ncol * L2;
    const cplx<T>* Z = Zi + (size_t)b * zstride;
    const int dr = NT / ncol, dc = NT % ncol;
    int r = tid / ncol, c = tid % ncol;
    for (int e0 = 0; e0 < nel; e0 += NT) {
      const int jj = (cA + c) * L2 + r;
      if (e0 + tid < nel && jj >= jlo && jj <= Yhi) {
        const cplx<T> v = Z[(size_t)r * L1 + cA + c];
        Ys[jj - Ylo] = mk<cplx<T>>(v.x * ft, v.y * ft);
      }
      c += dc;
      r += dr;
      if (c >= ncol) {
        c -= ncol;
        r++;
      }
    }
  }
  // window entries below 0 (first workgroup) and above Yhi: finite values for discarded steps
  for (int m = tid; m < SPF_YW; m += NT)
    if (m < jlo - Ylo || m > Yhi - Ylo) Ys[m] = mk<cplx<T>>(0, 0);
  const bool head = A0 == 0;
  if (head) {
    for (int q = tid; q < TI; q += NT) {
      const int i = q - SPL_W;
      tsh[q] = i < 1 ? 0.0 : i < SPL_NTAB ? tab->invp[i] : tab->invp_inf;
    }
    for (int k = tid; k < TC; k += NT) tsh[TI + k] = k < SPL_NTAB ? tab->c[k] : tab->c_inf;
  }
  const double invp_inf = tab->invp_inf, c_inf = tab->c_inf;
  __syncthreads();
  const bool w0 = head && tid < 64;
  const bool tail = A0 + SPF_OUT + SPL_W - 1 > n - 2;
  {
    const int uf = A0 + tid * KF;
    if (uf <= U1) {
      const cplx<T>* Yw = Ys + tid * KF;
      if (w0) spline_fwd<true, T, K>(Yw, Us, tsh + tid * KF, invp_inf, uf, A0, U1);
      else spline_fwd<false, T, K>(Yw, Us, tsh, invp_inf, uf, A0, U1);
    }
  }
  __syncthreads();
  const int a = A0 + tid * K, e = min(a + K, A1);
  double2 out[K + 1];  // out[K]: y2 at a + K, from the warm-up
  if (a < A1) {
    const double2* Uw = Us + spk<K>(a - A0);
    double2 o2[K];
    if (w0) spline_bwd<true, true, K>(Uw, tsh + TI + tid * K, c_inf, a, n, o2, &out[K]);
    else if (tail) spline_bwd<false, true, K>(Uw, tsh, c_inf, a, n, o2, &out[K]);
    else spline_bwd<false, false, K>(Uw, tsh, c_inf, a, n, o2, &out[K]);
#pragma unroll
    for (int q = 0; q < K; q++) out[q] = o2[q];
  }
  __syncthreads();  // every u read is done: Us now holds y2 on [A0, A1]
  if (a < A1) {
#pragma unroll
    for (int q = 0; q < K; q++)
      if (a + q < e) Us[spk<K>(a + q - A0)] = out[q];
    if (a + K == A1 && A1 < n) Us[spk<K>(A1 - A0)] = out[K];
  }
  __syncthreads();
  // splint at interpftpad (i - shftf[i]) (auxi.c:83-120), / sig2 (jobcore.c:386-389)
  const int ilo = ibound[0], ihi = ibound[1];
  if (ilo < 0 || ihi < 0) return;
  const double sig2 = it.p.det[d].sig2;
  cplx<T>* xo = static_cast<cplx<T>*>(it.xm) + (size_t)bs * N;
  for (int i = ilo + tid; i < ihi; i += NT) {
    const double x = interpftpad * (i - sh[i]);
    double vr = 0.0, vi = 0.0;
    if (!(x < 0 || x > n - 1)) {
      const int klo = (int)floor(x), khi = klo + 1;
      if (klo < A0 || klo >= A1) {  // the abscissae do not increase with i: not this workgroup's knots
        if (err) *err = 2;
        continue;
      }
      const double Aw = khi - x, Bw = x - klo;
      const double ca = Aw * Aw * Aw - Aw, cb = Bw * Bw * Bw - Bw;
      const cplx<T> yl = Ys[klo - Ylo];
      const double2 ylo = make_double2((double)yl.x, (double)yl.y);
      double2 yhi = make_double2(0.0, 0.0), zhi = make_double2(0.0, 0.0);
      if (khi < n) {  // khi == n only at x == n-1 exactly, where its weight Bw is 0
        const cplx<T> yh = Ys[khi - Ylo];
        yhi = make_double2((double)yh.x, (double)yh.y);
        zhi = Us[spk<K>(khi - A0)];
      }
      const double2 zlo = Us[spk<K>(klo - A0)];
      vr = Aw * ylo.x + Bw * yhi.x + (ca * zlo.x + cb * zhi.x) / 6.0;
      vi = Aw * ylo.y + Bw * yhi.y + (ca * zlo.y + cb * zhi.y) / 6.0;
    }
    xo[i] = mk<cplx<T>>((T)(vr / sig2), (T)(vi / sig2));
  }
}

// ---------------------------------------------------------- template stage --

// One template of a launch: the resampled series of its sky point (a sky slot), its shifts and
// noise-weighted aa/bb, and its frequency offset het1 and spindown sgnlt1 computed on the host
// with the reference's expressions (jobcore.c:456,466-467).  Templates of one launch may come
// from different sky points and segments that share the F-statistic FFT length.
// One template: its sky slot (whose resampled data xm, _tmp1 and aa/bb sums the
// kernels find at slot * stride from the slot bases), spindown phase coefficients and data length.
struct TmplDesc {
  double het1, sgnlt1;
  int slot, N, nifo, pad;
};
static_assert(sizeof(TmplDesc) == 32, "descriptor layout");

// Sky-slot bases: xm cplx<T>[slot][nifo][{a,b}][N], tmp1 double[slot][nifo][N], aabb double[slot][2].
// het1 and sgnlt[1] of spindown ss (jobcore.c:456,466-467), uncontracted.
struct SpinGrid {
  double M0, M4, M5, M9, M13;
  int mm, nn, ss0, s0_flag;
};
inline void spin_coeffs(const SpinGrid& q, int ss, double& het1, double& sgnlt1) {
#pragma clang fp contract(off)
  sgnlt1 = (q.s0_flag ? 0. : ss * q.M5 + q.nn * q.M9 + q.mm * q.M13);
  het1 = fmod(ss * q.M4, q.M0);
  if (het1 < 0) het1 += q.M0;
}

// uslot >= 0: the launch's templates all belong to the sky point in slot uslot with data length
// uN, so the kernels take addresses from these arguments instead of waiting for the descriptor;
// tab = 1: their het1, sgnlt1 are in the kernel-argument table (launches of at most kSpinTab
// templates), else in the descriptors (a load that no data load waits for).
struct SlotBases {
  const void* xm;
  const double* tmp1;
  const double* aabb;
  unsigned long long xm_stride, tmp1_stride;  // elements per slot
  int nifo, uslot, uN, tab;
};

// het1 and sgnlt[1] of the templates of a launch of at most kSpinTab templates, passed by value as
// a kernel argument: a block reads its entry with one scalar load from the argument segment.
constexpr int kSpinTab = 128;
// Launches spanning several sky points (slot of template t in slot[t], data length of a slot in
// slotN) read their templates' sky slots from the table as well, so a group whose launches all
// fit the table needs no descriptors on the device.
constexpr int kSpinSlots = 16;
struct SpinTab {
  double het1[kSpinTab], sgnlt1[kSpinTab];
  unsigned char slot[kSpinTab];
  int slotN[kSpinSlots];
};

struct SpinParams {
  int L1, twshift;
  int rev;  // FP32 form: phase accumulated in revolutions (PhaseRev) instead of radians
  int nt;   // spectra written with non-temporal stores
};

template <typename T> struct PhaseRot;
// sin and cos of x in FP64 without a large-argument path: k = rint(2x / pi), x - k pi/2 by three
// FMAs against pi/2 split in three doubles (the first two products exact inside the FMA, so the
// reduced angle is off by ~1e-16 for |x| < 2^50), then minimax polynomials of degree 13 (sin)
// and 12 (cos) on |r| <= pi/4, and the quadrant k mod 4.  Absolute error ~1.1e-16 (against
// 200-bit arithmetic, |x| up to 1e12); demodulation phases here are below 1e7.
__device__ __forceinline__ void sincos_cw(double x, double& s, double& c) {
  constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  const double k = rint(x * 0.6366197723675814);
  double r = fma(-k, P1, x);
  r = fma(-k, P2, r);
  r = fma(-k, P3, r);
  const double r2 = r * r;
  // sin r = r + r^3 S(r^2), cos r = 1 - r^2 C(r^2): S and C of degree 5, the minimax polynomials
  // of sin and cos on |r| <= pi/4 (Remez exchange in 200-bit arithmetic, sincos_minimax.py;
  // approximation error 1.6e-18 and 5.2e-17, below the FP64 rounding of the result)
  double ps = 1.588858596376266e-10;
  ps = fma(ps, r2, -2.5050600239083346e-08);
  ps = fma(ps, r2, 2.7557312544777556e-06);
  ps = fma(ps, r2, -0.0001984126982593413);
  ps = fma(ps, r2, 0.008333333333316482);
  ps = fma(ps, r2, -0.166666666666666);
  const double sn = fma(r * r2, ps, r);
  double pc = -2.0633590346945927e-09;
  pc = fma(pc, r2, 2.755529638055987e-07);
  pc = fma(pc, r2, -2.480157901658819e-05);
  pc = fma(pc, r2, 0.001388888887172768);
  pc = fma(pc, r2, -0.04166666666650322);
  pc = fma(pc, r2, 0.49999999999999484);
  const double cs = fma(-r2, pc, 1.0);
  // quadrant k mod 4 (k is an integer in FP64; exact for |k| < 2^53)
  const int q = (int)fma(-4.0, floor(k * 0.25), k);
  const double s0 = (q & 1) ? cs : sn, c0 = (q & 1) ? sn : cs;
  s = (q & 2) ? -s0 : s0;
  c = ((q + 1) & 2) ? -c0 : c0;
}

template <> struct PhaseRot<double> {
  // exp(-i phase) in FP64
  __device__ static __forceinline__ double2 eval(double ph) {
    double s, c;
    sincos_cw(ph, s, c);
    return make_double2(c, -s);
  }
};
template <> struct PhaseRot<float> {
  // FP64 phase reduced modulo 2 pi in FP64 (two-term Cody-Waite), then the hardware sin/cos of
  // the reduced angle in revolutions (v_sin_f32 / v_cos_f32 take x and return sin/cos(2 pi x))
  __device__ static __forceinline__ float2 eval(double ph) {
    const double k = rint(ph * 0.15915494309189533577);
    const double r = fma(-k, 6.283185307179586232, ph);
    const double rr = fma(-k, 2.4492935982947064e-16, r);
    const float rev = (float)(rr * 0.15915494309189533577);
    return make_float2(__builtin_amdgcn_cosf(rev), -__builtin_amdgcn_sinf(rev));
  }
};

// FP32 form, phase in revolutions: het1 / 2 pi and sgnlt1 / 2 pi once per thread, then per
// sample u = het1 n / 2 pi + sgnlt1 _tmp1 / 2 pi in FP64 and its fraction u - rint(u) (exact in
// FP64 for |u| < 2^52) to the hardware sin/cos, which take revolutions: four FP64 operations per
// sample fewer than the radians and two-term reduction of PhaseRot<float>.
struct PhaseRev {
  double h, s;  // het1 / 2 pi, sgnlt1 / 2 pi
  __device__ __forceinline__ PhaseRev(double het1, double sgnlt1)
      : h(het1 * 0.15915494309189533577), s(sgnlt1 * 0.15915494309189533577) {}
  __device__ __forceinline__ float2 eval(double dn, double tmp1) const {
    const double u = fma(h, dn, s * tmp1);
    const float f = (float)(u - rint(u));
    return make_float2(__builtin_amdgcn_cosf(f), -__builtin_amdgcn_sinf(f));
  }
};

struct RowParams {
  int L2, nmin, nmax, nblk, nparts, lgL2;  // L2 = 1 << lgL2
  int nt;   // spectra read with non-temporal loads
  int fnt;  // F written with non-temporal stores
  // sky slot of template t: uslot (>= 0: the launch's one slot), else slot[t] (tab = 1) or the
  // descriptor's
  int uslot, tab;
  unsigned char slot[kSpinTab];
};
__device__ __forceinline__ int row_slot(const RowParams& p, const TmplDesc* __restrict__ desc, int t) {
  return p.uslot >= 0 ? p.uslot : p.tab ? (int)p.slot[t] : desc[t].slot;
}

struct Cand {
  int t, ii;   // template index within the group, frequency bin
  double F;    // normalised F at the run maximum
};

struct ThrParams {
  int L1, L2, nmin, nmax, nblk, nparts, white, t0, cap, lgL2;  // L2 = 1 << lgL2
  double trl;
};

// FStat normalisation + run-max threshold (auxi.c:214-240, jobcore.c:659-695).
// grid (nparts, ntemplates), one 64-lane wave per (part, template).  The wave reduces its part's
// L2 row partials in a fixed order (mean mu = sum / (2*4096) for full blocks, 1 for the raw tail
// or --whitenoise) and exits unless max(F)/mu > trl.  A flagged part computes every part's mu and
// scans its own bins; each maximal run of F/mu > trl is emitted by the lane owning its first bin
// (it may continue into later parts), at the last index of the run maximum.
__device__ __forceinline__ double part_mean(const double sum, int part, const ThrParams& p) {
  return (part < p.nblk && !p.white) ? sum / (2. * PGW_NAV_DEV) : 1.0;
}

constexpr int THR_MAXP = 256;  // FStat blocks per template at most
constexpr int THR_LDS = THR_MAXP + 64;  // doubles of LDS per threshold wave: block means, pre-selection bitmap

// One wave (lane = 0..63) thresholds FStat block `part` of template t; mu: THR_LDS doubles of LDS
// owned by this wave.  TJ, QB: loads in flight per batch in the flagged path (bins, block means).
// SC: F and the partials read at device scope (written by other workgroups of the same launch).
// HITF: Fh (PGW_NAV_DEV elements of LDS owned by this wave) keeps the F of the pre-selected bins,
// so that the exact tests and the run walks inside the block read LDS (a bin that is not
// pre-selected is not above trl: the run ends there) and only a run continuing into the next
// block loads from F.
template <typename T, int TJ = (sizeof(T) == 4 ? 32 : 16), int QB = 16, bool SC = false, bool HITF = false>
__device__ __forceinline__ void thr_wave(const ThrParams& p, int part, int t, int lane, double* mu,
                                         const T* __restrict__ Ft, const T* __restrict__ psum,
                                         const T* __restrict__ pmax, double* __restrict__ mu_out,
                                         Cand* __restrict__ cand, int* __restrict__ ncand, T* Fh = nullptr) {
  const size_t L = (size_t)p.L1 * p.L2;
  const T* F = Ft + (size_t)t * L;
  const T* ps0 = psum + (size_t)t * p.nparts * p.L2;
  const T* pm0 = pmax + (size_t)t * p.nparts * p.L2;
  auto ld = [](const T* q) -> T {
    if constexpr (SC) return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *q;
  };
  double sum = 0.0, mx = 0.0;
  for (int k2 = lane; k2 < p.L2; k2 += 64) {
    sum += (double)ld(&ps0[(size_t)part * p.L2 + k2]);
    mx = fmax(mx, (double)ld(&pm0[(size_t)part * p.L2 + k2]));
  }
  sum = wave_sum(sum);
  mx = wave_max(mx);
  const double m = part_mean(sum, part, p);
  if (lane == 0) mu_out[(size_t)t * p.nparts + part] = m;
  if (!(mx / m > p.trl)) return;  // uniform across the wave
  // flagged: pre-select this part's bins with F > trl m (1 - 1e-14), which every bin with
  // F / m > trl passes; the exact test follows per selected bin.  Bin k lies at row k mod L2,
  // column k / L2 of F, so the part's bins are read row by row over the columns [cA, cA + ncol)
  // that hold them, each load instruction covering contiguous row segments (element e = r ncol + c
  // of that block for lane + 64 j, TJ loads in flight per batch).  A pre-selected bin at offset
  // o = k - lok sets bit o / 64 of bitmap word o mod 64: lane l's word then holds the bins
  // lok + l + 64 j that the exact tests below take lane by lane.
  const int lok = p.nmin + PGW_NAV_DEV * part;
  const int hik = (part < p.nblk) ? lok + PGW_NAV_DEV : p.nmax;
  const int mL2 = p.L2 - 1;
  auto at = [&](int k) { return (double)ld(&F[(size_t)(k & mL2) * p.L1 + (k >> p.lgL2)]); };
  unsigned long long* bm = reinterpret_cast<unsigned long long*>(mu + THR_MAXP);
  bm[lane] = 0;
  wave_lds_sync();
  const double pre = p.trl * m * (1. - 1e-14);
  {
    const int cA = lok >> p.lgL2, ncol = ((hik - 1) >> p.lgL2) - cA + 1;
    const int nel = ncol * p.L2;
    const int dr = 64 / ncol, dc = 64 % ncol;  // the step of 64 elements in rows and columns
    int r = lane / ncol, c = lane % ncol;
    // (TJ loads per batch keep the kernel's register count, and so the occupancy of the far
    // more common unflagged waves, moderate)
#pragma unroll 1
    for (int e0 = 0; e0 < nel; e0 += 64 * TJ) {
      T f[TJ];
      int o[TJ];
#pragma unroll
      for (int j = 0; j < TJ; j++) {
        const int k = (cA + c) * p.L2 + r;
        o[j] = (r < p.L2 && k >= lok && k < hik) ? k - lok : -1;
        f[j] = ld(&F[(size_t)min(r, mL2) * p.L1 + cA + c]);
        c += dc;
        r += dr;
        if (c >= ncol) {
          c -= ncol;
          r++;
        }
      }
#pragma unroll
      for (int j = 0; j < TJ; j++)
        if (o[j] >= 0 && (double)f[j] > pre) {
          atomicOr(&bm[o[j] & 63], 1ull << (o[j] >> 6));
          if constexpr (HITF) Fh[o[j]] = f[j];
        }
    }
  }
  wave_lds_sync();
  uint64_t hits = bm[lane];
  // means of all parts with the unflagged waves' arithmetic, QB parts' loads in flight at a time
#pragma unroll 1
  for (int q0 = 0; q0 < p.nparts; q0 += QB) {
    double v[QB];
#pragma unroll
    for (int j = 0; j < QB; j++) v[j] = 0.0;
    for (int k2 = lane; k2 < p.L2; k2 += 64) {
      T x[QB];
#pragma unroll
      for (int j = 0; j < QB; j++) x[j] = ld(&ps0[(size_t)min(q0 + j, p.nparts - 1) * p.L2 + k2]);
#pragma unroll
      for (int j = 0; j < QB; j++) v[j] += (double)x[j];
    }
#pragma unroll
    for (int j = 0; j < QB; j++) {
      const double sq = wave_sum(v[j]);
      if (lane == 0 && q0 + j < p.nparts) mu[q0 + j] = part_mean(sq, q0 + j, p);
    }
  }
  wave_lds_sync();
  // neighbours' pre-selection bits: bin k - 1 and k + 1 of this part are lane - 1 / lane + 1 (same
  // j), or lane 63 (j - 1) / lane 0 (j + 1) across the wave's edge.  A neighbour without its bit is
  // certainly not above trl, so its exact test (a dependent load) is skipped.
  const uint64_t h_lo = __shfl(hits, lane == 0 ? 63 : lane - 1, 64);
  const uint64_t h_hi = __shfl(hits, lane == 63 ? 0 : lane + 1, 64);
  const uint64_t prevm = lane == 0 ? h_lo << 1 : h_lo;
  const uint64_t nextm = lane == 63 ? h_hi >> 1 : h_hi;
  while (hits) {
    const int j = __builtin_ctzll(hits);
    hits &= hits - 1;
    const int k = lok + lane + 64 * j;
    const double fn = (HITF ? (double)Fh[k - lok] : at(k)) / mu[part];
    if (!(fn > p.trl)) continue;
    bool start = (k == p.nmin);
    if (!start && (k == lok || (prevm >> j) & 1)) {  // previous bin in another part, or pre-selected
      const int kp = k - 1;
      const int pp = min((kp - p.nmin) / PGW_NAV_DEV, p.nparts - 1);
      start = !(((HITF && k != lok) ? (double)Fh[kp - lok] : at(kp)) / mu[pp] > p.trl);
    } else if (!start) {
      start = true;
    }
    if (!start) continue;
    int ii = k;
    double Fc = fn;
    bool ends = k + 1 < hik && !((nextm >> j) & 1);  // next bin of this part not pre-selected
    int kw = k + 1;  // first bin of the run walk below
    if constexpr (HITF) {  // the run inside this part, from LDS
      for (; !ends && kw < hik; kw++) {
        const int o = kw - lok;
        const double v = (double)Fh[o] / mu[part];
        if (!((bm[o & 63] >> (o >> 6)) & 1ull) || !(v > p.trl)) ends = true;
        else if (v >= Fc) { ii = kw; Fc = v; }
      }
    }
    // the run's later bins RB at a time, their loads in flight together (addresses clamped below
    // nmax; bins past the run's end are loaded and ignored)
    constexpr int RB = TJ < 8 ? TJ : 8;
    for (int kk0 = kw; !ends && kk0 < p.nmax; kk0 += RB) {
      double v[RB];
#pragma unroll
      for (int q = 0; q < RB; q++) {
        const int kk = min(kk0 + q, p.nmax - 1);
        const int pk = min((kk - p.nmin) / PGW_NAV_DEV, p.nparts - 1);
        v[q] = at(kk) / mu[pk];
      }
      bool done = false;
#pragma unroll
      for (int q = 0; q < RB; q++) {
        if (done) continue;
        if (kk0 + q >= p.nmax || !(v[q] > p.trl)) done = true;
        else if (v[q] >= Fc) { ii = kk0 + q; Fc = v[q]; }
      }
      if (done) break;
    }
    const int idx = atomicAdd(ncand, 1);
    if (idx < p.cap) {
      Cand c;
      c.t = p.t0 + t;
      c.ii = ii;
      c.F = Fc;
      cand[idx] = c;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(64) void k_threshold(ThrParams p, const T* __restrict__ Ft, const T* __restrict__ psum,
                                                  const T* __restrict__ pmax, double* __restrict__ mu_out,
                                                  Cand* __restrict__ cand, int* __restrict__ ncand) {
  __shared__ double mu[THR_LDS];
  if constexpr (sizeof(T) == 4) {  // (FP64: 32 KB of LDS per wave cost more occupancy than it saved)
    __shared__ T Fh[PGW_NAV_DEV];
    thr_wave<T, 32, 16, false, true>(p, blockIdx.x, blockIdx.y, threadIdx.x, mu, Ft, psum, pmax, mu_out, cand, ncand,
                                     Fh);
  } else {
    thr_wave<T>(p, blockIdx.x, blockIdx.y, threadIdx.x, mu, Ft, psum, pmax, mu_out, cand, ncand);
  }
}

// A threshold launch folded into the next column-pass launch: the first `rows` grid rows of that
// launch run it, one (part, template) per wave, beside the column pass's own workgroups.
template <typename T> struct ThrFuse {
  ThrParams p;
  const T* Ft;
  const T* psum;
  const T* pmax;
  double* mu;
  Cand* cand;
  int* ncand;
  int nt;    // templates of the folded threshold (0: none)
  int rows;  // grid rows appended for it
};

// zero n ints (used instead of a memset so the launch sequence can be captured in a graph)
__global__ void k_zero_int(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// plain copy (grid-stride), used to calibrate hardware byte counters
__global__ __launch_bounds__(NT) void k_copy(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) b[i] = a[i];
}

// HBM probes (grid-stride, 16 bytes per lane and step): read only (a sum that is stored only if it
// takes a value the data cannot give, so the loads stay) and write only (non-temporal stores)
__global__ __launch_bounds__(NT) void k_read_probe(const float4* __restrict__ a, size_t n, float* __restrict__ out) {
  float s = 0.f;
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
    const float4 x = a[i];
    s += (x.x + x.y) + (x.z + x.w);
  }
  if (s == -1.f) out[0] = s;  // the buffer holds zeros
}
__global__ __launch_bounds__(NT) void k_write_probe(float4* __restrict__ b, size_t n) {
  using V = float __attribute__((ext_vector_type(4)));
  V z = {0.f, 0.f, 0.f, 0.f};
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
    __builtin_nontemporal_store(z, reinterpret_cast<V*>(b + i));
}

}  // namespace pgw
