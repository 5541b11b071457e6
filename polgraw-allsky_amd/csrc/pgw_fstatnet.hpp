// Exact network F-statistic at continuous parameters and its gradient (the followup's Fstatnet,
// followup/src/followup.c:413-534), for batches of points.  FP64 throughout: the phase reaches
// ~1e6 rad.
#pragma once

#include "pgw_kernels.hpp"

namespace pgw {

struct FsnDet {
  double amod[9];     // c1..c9 of rogcvir
  double sphir, cphir;
  double w;           // 1 / sig2
};

struct FsnParams {
  int N, nifo, nchunk, pad;
  double oms;
  FsnDet det[PGW_MAX_DETECTORS];
};

constexpr int FSN_NT = 256;
constexpr int FSN_PER = 16;                      // samples per thread and detector in a block
constexpr int FSN_CHUNK = FSN_NT * FSN_PER;      // samples per block
constexpr int FSN_NSUM = 26;

// Sums per point (index):
//   0,1 Xa = sum w x a e^{-i phi}; 2,3 Xb (b for a);
//   4..11 dXa/dtheta_k (k = omega, omegadot, delta, alpha; re, im); 12..19 dXb/dtheta_k;
//   20 A = sum w a^2; 21 B; 22,23 dA/d(delta, alpha); 24,25 dB/d(delta, alpha).

// phase = omega (i + shft) + omegadot i i + (oms + 2 omegadot i) shft in the reference's order
// (followup.c:455-457), uncontracted
__device__ __forceinline__ double fsn_phase(double om, double omd, double oms, int i, double shft) {
#pragma clang fp contract(off)
  return om * (double)(i + shft) + (omd * i * i) + ((oms + 2. * omd * i) * shft);
}

// grid (npoints, nchunk), FSN_NT threads: block (p, c) sums samples [c FSN_CHUNK, (c+1) FSN_CHUNK)
// of every detector for point p into part[p][c][FSN_NSUM].  Consecutive blocks are different points
// of the same samples, so a chunk's inputs are read from HBM once and then from L2.  GRAD = false
// accumulates only the sums of F (0-3, 20, 21).
template <bool GRAD>
__global__ __launch_bounds__(FSN_NT) void k_fstatnet_part(FsnParams q, const double* __restrict__ pts,
                                                          const double* __restrict__ xDat,
                                                          const double* __restrict__ DetSSB,
                                                          const double* __restrict__ cosmodf,
                                                          const double* __restrict__ sinmodf,
                                                          double* __restrict__ part) {
  const int p = blockIdx.x, c = blockIdx.y, tid = threadIdx.x;
  const double om = pts[4 * p], omd = pts[4 * p + 1], de = pts[4 * p + 2], al = pts[4 * p + 3];
  double sinal, cosal, sindel, cosdel;
  sincos(al, &sinal, &cosal);
  sincos(de, &sindel, &cosdel);
  const double nS0 = cosal * cosdel, nS1 = sinal * cosdel, nS2 = sindel;
  const double nD0 = -cosal * sindel, nD1 = -sinal * sindel, nD2 = cosdel;  // d nS / d delta
  const double nA0 = -sinal * cosdel, nA1 = cosal * cosdel;                 // d nS / d alpha
  const double c2d = cosdel * cosdel, c2sd = sindel * cosdel, cos2d = cosdel * cosdel - sindel * sindel;
  const int N = q.N;
  double s[FSN_NSUM];
#pragma unroll
  for (int k = 0; k < FSN_NSUM; k++) s[k] = 0.;
  for (int d = 0; d < q.nifo; d++) {
    const FsnDet& D = q.det[d];
    const double cosalfr = cosal * D.cphir + sinal * D.sphir;
    const double sinalfr = sinal * D.cphir - cosal * D.sphir;
    const double* __restrict__ R = DetSSB + (size_t)d * 3 * N;
    const double* __restrict__ x = xDat + (size_t)d * N;
    for (int r = 0; r < FSN_PER; r++) {
      const int i = c * FSN_CHUNK + r * FSN_NT + tid;
      if (i >= N) break;
      const double R0 = R[3 * (size_t)i], R1 = R[3 * (size_t)i + 1], R2 = R[3 * (size_t)i + 2];
      const double shft = dot3_nc(nS0, R0, nS1, R1, nS2, R2);
      double sp, cp;
      // values only: the bounded reduction and minimax series of the column pass (79 instead of 96
      // VGPRs, 6 waves per SIMD instead of 5: 732-740k -> 833-836k points/s); with the gradient
      // the device library's sincos measured faster (471k vs 420k points/s)
      if constexpr (GRAD) sincos(fsn_phase(om, omd, q.oms, i, shft), &sp, &cp);
      else sincos_cw(fsn_phase(om, omd, q.oms, i, shft), sp, cp);
      // modvir at sample i (followup settings.c:298-316) and its delta / alpha derivatives
      const double cm = cosmodf[i], sm = sinmodf[i];
      const double cc = cosalfr * cm + sinalfr * sm, ss = sinalfr * cm - cosalfr * sm;
      const double c2s = 2. * cc * cc, cs = cc * ss, cdiff = cc * cc - ss * ss;
      const double* cf = D.amod;
      const double a = cf[0] * (2. - c2d) * c2s + cf[1] * (2. - c2d) * 2. * cs + cf[2] * c2sd * cc +
                       cf[3] * c2sd * ss - cf[0] * (2. - c2d) + cf[4] * c2d;
      const double b = cf[5] * sindel * c2s + cf[6] * sindel * 2. * cs + cf[7] * cosdel * cc +
                       cf[8] * cosdel * ss - cf[5] * sindel;
      const double xw = x[i] * D.w;
      const double er = xw * cp, ei = -xw * sp;
      const double ur = a * er, ui = a * ei, vr = b * er, vi = b * ei;
      s[0] += ur; s[1] += ui; s[2] += vr; s[3] += vi;
      s[20] += D.w * a * a;
      s[21] += D.w * b * b;
      if constexpr (!GRAD) continue;
      const double da_d = cf[0] * 2. * c2sd * c2s + cf[1] * 2. * c2sd * 2. * cs + cf[2] * cos2d * cc +
                          cf[3] * cos2d * ss - cf[0] * 2. * c2sd - 2. * cf[4] * c2sd;
      const double db_d = cf[5] * cosdel * c2s + cf[6] * cosdel * 2. * cs - cf[7] * sindel * cc -
                          cf[8] * sindel * ss - cf[5] * cosdel;
      const double da_a = cf[0] * (2. - c2d) * (-4. * cs) + cf[1] * (2. - c2d) * 2. * cdiff -
                          cf[2] * c2sd * ss + cf[3] * c2sd * cc;
      const double db_a = cf[5] * sindel * (-4. * cs) + cf[6] * sindel * 2. * cdiff - cf[7] * cosdel * ss +
                          cf[8] * cosdel * cc;
      // d phi / d theta: i + shft, i^2 + 2 i shft, K d shft / d delta, K d shft / d alpha
      const double di = (double)i;
      const double K = om + q.oms + 2. * omd * di;
      const double ph0 = di + shft, ph1 = di * di + 2. * di * shft;
      const double ph2 = K * (nD0 * R0 + nD1 * R1 + nD2 * R2), ph3 = K * (nA0 * R0 + nA1 * R1);
      // d/dtheta (a e) = da e - i a e dphi: (-i)(ur + i ui) = ui - i ur
      s[4] += ph0 * ui;  s[5] -= ph0 * ur;
      s[6] += ph1 * ui;  s[7] -= ph1 * ur;
      s[8] += ph2 * ui + da_d * er;  s[9] += -ph2 * ur + da_d * ei;
      s[10] += ph3 * ui + da_a * er; s[11] += -ph3 * ur + da_a * ei;
      s[12] += ph0 * vi; s[13] -= ph0 * vr;
      s[14] += ph1 * vi; s[15] -= ph1 * vr;
      s[16] += ph2 * vi + db_d * er; s[17] += -ph2 * vr + db_d * ei;
      s[18] += ph3 * vi + db_a * er; s[19] += -ph3 * vr + db_a * ei;
      s[22] += 2. * D.w * a * da_d;
      s[23] += 2. * D.w * a * da_a;
      s[24] += 2. * D.w * b * db_d;
      s[25] += 2. * D.w * b * db_a;
    }
  }
  __shared__ double red[FSN_NT / 64][FSN_NSUM];
  const int wv = tid / 64, ln = tid % 64;
#pragma unroll
  for (int k = 0; k < FSN_NSUM; k++) {
    if (!GRAD && !(k < 4 || k == 20 || k == 21)) continue;  // only the sums of F
    const double v = wave_sum(s[k]);
    if (ln == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (tid < FSN_NSUM) {
    double v = 0.;
    if (GRAD || tid < 4 || tid == 20 || tid == 21) {
      v = red[0][tid];
      for (int w = 1; w < FSN_NT / 64; w++) v += red[w][tid];
    }
    part[((size_t)p * q.nchunk + c) * FSN_NSUM + tid] = v;
  }
}

// One thread per point: chunk partials summed in chunk order, then the reference's outputs
// (followup.c:498-522) and dF/dtheta for F = |Xa|^2/A + |Xb|^2/B.
__global__ __launch_bounds__(64) void k_fstatnet_final(int n, int nchunk, const double* __restrict__ part,
                                                       const double* __restrict__ pts, double* __restrict__ out,
                                                       double* __restrict__ grad) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= n) return;
  double s[FSN_NSUM];
  for (int k = 0; k < FSN_NSUM; k++) s[k] = 0.;
  for (int c = 0; c < nchunk; c++) {
    const double* v = part + ((size_t)p * nchunk + c) * FSN_NSUM;
    for (int k = 0; k < FSN_NSUM; k++) s[k] += v[k];
  }
  const double aa = s[20], bb = s[21];
  const double Fa = s[0] * s[0] + s[1] * s[1], Fb = s[2] * s[2] + s[3] * s[3];
  double* o = out + 11 * (size_t)p;
  o[5] = -((Fa / aa) + (Fb / bb));
  o[0] = 2 * s[0] / aa;
  o[1] = 2 * s[2] / bb;
  o[2] = -2 * s[1] / aa;
  o[3] = -2 * s[3] / bb;
  o[4] = sqrt(2 * (-o[5] - 2));
  for (int k = 0; k < 4; k++) o[6 + k] = pts[4 * (size_t)p + k];
  o[10] = sqrt(4 * s[0] * s[0] + 4 * s[2] * s[2] + 4 * s[1] * s[1] + 4 * s[3] * s[3]);
  if (grad) {
    for (int k = 0; k < 4; k++) {
      const double dA = k >= 2 ? s[22 + (k - 2)] : 0., dB = k >= 2 ? s[24 + (k - 2)] : 0.;
      // 2 Re(conj(X) dX) = 2 (Xr dXr + Xi dXi)
      grad[4 * (size_t)p + k] = 2. * (s[0] * s[4 + 2 * k] + s[1] * s[5 + 2 * k]) / aa - Fa * dA / (aa * aa) +
                                2. * (s[2] * s[12 + 2 * k] + s[3] * s[13 + 2 * k]) / bb - Fb * dB / (bb * bb);
    }
  }
}

}  // namespace pgw
