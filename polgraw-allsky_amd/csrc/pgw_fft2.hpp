// pgw_fft2.hpp - register-resident two-pass FFT building blocks (gfx950, 256-thread blocks).
//
// Same index convention as pgw_common.hpp: L = L1*L2, n = n1 + L1 n2, k = L2 k1 + k2.
//
// Column pass (P2-point transforms over n2, C consecutive columns n1 per block):
//   P2 = E * G.  Thread t owns column c = t % C and group g = t / C (C = 256/G) and holds
//   x[g + G r], r < E, in registers.  Step 1: E-point DFT in registers, times W_P2^{g kr}.
//   One LDS exchange ([g][kr][c], c fastest: conflict-free).  Step 2: G-point DFTs over g,
//   giving X[kr + E kg] in registers.
//
// Row pass (P1 = 16*16*R3 points, NT = P1/16 threads): radix-16, radix-16, radix-R3 Stockham
//   stages with the data in registers and three LDS exchanges padded by one element per 16;
//   thread t ends with X[t + NT m], m < 16 (natural order).
//
// Stage twiddles come from one table entry per thread and stage, raised to the needed powers by
// a product tree of depth <= 4 (error a few ulp); arrays are processed together so they share
// every barrier.
#pragma once
#include "pgw_common.hpp"

namespace pgw {

// w^0..w^(R-1) by a product tree (depth <= 4 for R <= 16)
template <int R, typename T2> __device__ __forceinline__ void tw_powers(T2 w, T2* p) {
  p[0] = mk<T2>(1, 0);
  if constexpr (R > 1) p[1] = w;
#pragma unroll
  for (int r = 2; r < R; r++) {
    const int h = 1 << (31 - __builtin_clz(r));  // highest power of two <= r
    p[r] = (h == r) ? cmul(p[r / 2], p[r / 2]) : cmul(p[h], p[r - h]);
  }
}

// v[r] *= w^r for r < R.  FP32: product-tree powers (few ulp).  FP64: running product (one
// complex register live instead of R; error grows ~R ulp, 2e-15 at R = 16).
template <int R, int DIR, typename T2> __device__ __forceinline__ void twiddle_run(T2* v, T2 w) {
  w = tw_dir<DIR>(w);
  if constexpr (sizeof(decltype(T2::x)) == 8) {
    T2 pr = w;
#pragma unroll
    for (int r = 1; r < R; r++) {
      v[r] = cmul(v[r], pr);
      if (r + 1 < R) pr = cmul(pr, w);
    }
  } else {
    T2 p[R];
    tw_powers<R>(w, p);
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], p[r]);
  }
}

template <int R, int DIR, typename T2> __device__ __forceinline__ void apply_tw(T2* v, T2 w) {
  T2 p[R];
  tw_powers<R>(tw_dir<DIR>(w), p);
#pragma unroll
  for (int r = 1; r < R; r++) v[r] = cmul(v[r], p[r]);
}

template <int R, int DIR, typename T2> __device__ __forceinline__ void apply_tw_pow(T2* v, const T2* p) {
#pragma unroll
  for (int r = 1; r < R; r++) v[r] = cmul(v[r], p[r]);
}

// ----------------------------------------------------------- column pass --
template <int P2, int E, int NTH = 256> struct ColShape {
  static constexpr int G = P2 / E;     // groups per column
  static constexpr int C = NTH / G;    // columns per block of NTH threads
  static constexpr int Q = E / G;      // kr values per thread in step 2
  static_assert(G >= 1 && G <= E && E % G == 0, "column split");
};

// NA arrays x[a][r] (registers) -> X in registers: out[a][q*G + kg] = X[kr + E kg], kr = grp*Q + q.
// lds must hold NA * P2 * C elements.  twc[m] = W_P2^m (forward sign).  One barrier inside; the
// caller must barrier before reusing lds.
// NZ < E: the inputs x[a][r], r >= NZ, are zero (pruned first DFT).  pw (optional): the step-1
// twiddle powers of this thread, computed on the first call (pw_have = false), reused after.
template <typename T2, int P2, int E, int NA, int DIR, int NZ = E, int NTH = 256>
__device__ __forceinline__ void col_fft(T2 (&x)[NA][E], T2* lds, const T2* __restrict__ twc, T2* pw = nullptr,
                                        bool pw_have = false) {
  using S = ColShape<P2, E, NTH>;
  const int c = threadIdx.x % S::C, g = threadIdx.x / S::C;
  const T2 wg = twc[g];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    dftP<E, NZ, DIR>(x[a]);
    if (pw) {  // W_P2^{g kr} from the cached powers
      if (!pw_have && a == 0) tw_powers<E>(tw_dir<DIR>(wg), pw);
      apply_tw_pow<E, DIR>(x[a], pw);
    } else {
      twiddle_run<E, DIR>(x[a], wg);  // W_P2^{g kr}
    }
#pragma unroll
    for (int kr = 0; kr < E; kr++) lds[(size_t)a * P2 * S::C + (g * E + kr) * S::C + c] = x[a][kr];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < NA; a++) {
#pragma unroll
    for (int q = 0; q < S::Q; q++) {
      const int kr = g * S::Q + q;
      T2 v[S::G];
#pragma unroll
      for (int gg = 0; gg < S::G; gg++) v[gg] = lds[(size_t)a * P2 * S::C + (gg * E + kr) * S::C + c];
      if constexpr (S::G > 1) dftR<S::G, DIR>(v);
#pragma unroll
      for (int kg = 0; kg < S::G; kg++) x[a][q * S::G + kg] = v[kg];
    }
  }
}

// Inter-pass twiddles W_L^{n1 k2} for the thread's E output slots from one table entry u = W_L^{n1}:
// k2 = kr + E kg with kr = grp*Q + q, so W = u^{grp Q} * u^q * (u^E)^{kg}  (forward sign).
template <int P2, int E, int NTH = 256, typename T2>
__device__ __forceinline__ void col_out_twiddles(T2 u, T2* w) {
  using S = ColShape<P2, E, NTH>;
  const int g = threadIdx.x / S::C;
  T2 pu[E + 1];  // u^0..u^E
  tw_powers<E>(u, pu);
  pu[E] = cmul(pu[E / 2], pu[E / 2]);
  // u^{g Q}: binary powering over the small exponent (g < G, Q <= E)
  T2 base = mk<T2>(1, 0), sq = u;
  for (int e = g * S::Q; e; e >>= 1) {
    if (e & 1) base = cmul(base, sq);
    sq = cmul(sq, sq);
  }
  T2 pe[S::G];  // (u^E)^kg
  tw_powers<S::G>(pu[E], pe);
#pragma unroll
  for (int q = 0; q < S::Q; q++) {
    const T2 bq = cmul(base, pu[q]);
#pragma unroll
    for (int kg = 0; kg < S::G; kg++) w[q * S::G + kg] = cmul(bq, pe[kg]);
  }
}

// output index k2 of register slot s (= q*G + kg) for the thread
// output slot s of a thread holds row k2 = g Q + col_out_k2c(s): the thread's group g gives the
// first term, the slot alone the second
template <int P2, int E> __device__ __forceinline__ constexpr int col_out_k2c(int s) {
  using S = ColShape<P2, E>;
  return s / S::G + E * (s % S::G);
}
template <int P2, int E> __device__ __forceinline__ int col_out_k2(int s) {
  using S = ColShape<P2, E>;
  const int g = threadIdx.x / S::C;
  return g * S::Q + col_out_k2c<P2, E>(s);
}

// -------------------------------------------------------------- row pass --
__device__ __forceinline__ int rpad(int i) { return i + (i >> 4); }

template <int P1> struct RowShape {
  static constexpr int NT = P1 / 16;
  static constexpr int R3 = P1 / 256;
  static constexpr int PADN = P1 + P1 / 16;  // padded elements per array
  static_assert(R3 == 4 || R3 == 8 || R3 == 16, "row length 1024, 2048 or 4096");
};

// x[a][m] = row element (t + NT m) on entry, X[t + NT m] on exit.  lds holds NA * PADN elements.
// twr[m] = W_P1^m (forward sign).  Barriers inside; ends after a barrier-free read (caller must
// barrier before reusing lds).
// pc (optional): the stage-C twiddle powers of this thread, computed on the first call (pc_have =
// false) and reused on later calls of the same thread (the second spectrum of a row); pb likewise
// the stage-B powers.  t: the thread's index among the row's P1/16 threads (a workgroup may
// transform several rows, one per group of P1/16 threads).
// TAG: distinct instantiations for distinct kernels (a body shared between kernels is optimised once
// before it is inlined, which changed the register allocation of the row pass)
template <typename T2, int P1, int NA, int DIR, int TAG = 0>
__device__ __forceinline__ void row_fft(T2 (&x)[NA][16], T2* lds, const T2* __restrict__ twr, T2* pc = nullptr,
                                        bool pc_have = false, T2* pb = nullptr, int t = (int)threadIdx.x) {
  using S = RowShape<P1>;
  constexpr int NT = S::NT, R3 = S::R3, PADN = S::PADN;
  // stage twiddle bases loaded up front so their latency overlaps stage A
  const T2 wB = twr[(t & 15) * (P1 / 256)];
  T2 wC[16 / R3];
#pragma unroll
  for (int i = 0; i < 16 / R3; i++) wC[i] = twr[t + NT * i];
  // stage A: radix 16, Ns = 1; outputs at 16 t + r
#pragma unroll
  for (int a = 0; a < NA; a++) {
    dftR<16, DIR>(x[a]);
#pragma unroll
    for (int r = 0; r < 16; r++) lds[a * PADN + rpad(16 * t + r)] = x[a][r];
  }
  __syncthreads();
  // stage B: radix 16, Ns = 16; butterfly j = t reads j + NT r
  {
    const int k = t & 15;
#pragma unroll
    for (int a = 0; a < NA; a++) {
#pragma unroll
      for (int r = 0; r < 16; r++) x[a][r] = lds[a * PADN + rpad(t + NT * r)];
      if (pb) {  // W_256^{k r} from the cached powers
        if (!pc_have && a == 0) tw_powers<16>(tw_dir<DIR>(wB), pb);
        apply_tw_pow<16, DIR>(x[a], pb);
      } else {
        twiddle_run<16, DIR>(x[a], wB);  // W_256^{k r}
      }
      dftR<16, DIR>(x[a]);
    }
    __syncthreads();
    const int base = (t - k) * 16 + k;
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
      for (int r = 0; r < 16; r++) lds[a * PADN + rpad(base + 16 * r)] = x[a][r];
    __syncthreads();
  }
  // stage C: radix R3, Ns = 256; butterflies b = t + NT i (i < 16/R3) read b + 256 r, k = b
#pragma unroll
  for (int i = 0; i < 16 / R3; i++) {
    const int b = t + NT * i;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      T2 v[R3];
#pragma unroll
      for (int r = 0; r < R3; r++) v[r] = lds[a * PADN + rpad(b + 256 * r)];
      if (pc) {  // W_P1^{b r} from the cached powers
        if (!pc_have && a == 0) tw_powers<R3>(tw_dir<DIR>(wC[i]), pc + R3 * i);
        apply_tw_pow<R3, DIR>(v, pc + R3 * i);
      } else {
        twiddle_run<R3, DIR>(v, wC[i]);  // W_P1^{b r}
      }
      dftR<R3, DIR>(v);
      // outputs b + 256 r = t + NT (i + (16/R3) r)
#pragma unroll
      for (int r = 0; r < R3; r++) x[a][i + (16 / R3) * r] = v[r];
    }
  }
}

// Row FFT with 8 elements per thread (NT = P1/8 threads): radix-8 Stockham stages with Ns = 1, 8,
// 64, then radix R = P1/512 with Ns = 512, through one padded LDS array of P1 elements (three
// exchanges).  Half the registers of row_fft per thread, for FP64.  x[m] = row element
// (t + NT m) on entry, X[t + NT m] on exit.  Ends after a barrier-free read (caller must barrier
// before reusing lds).
template <int P1> struct Row8Shape {
  static constexpr int NT = P1 / 8;
  static constexpr int RL = P1 / 512;           // last radix
  static constexpr int PADN = P1 + P1 / 16;     // padded elements
  static_assert(RL == 2 || RL == 4 || RL == 8, "row length 1024, 2048 or 4096");
};

template <typename T2, int P1, int DIR>
__device__ __forceinline__ void row_fft8(T2 (&x)[8], T2* lds, const T2* __restrict__ twr) {
  using S = Row8Shape<P1>;
  constexpr int NT = S::NT, RL = S::RL;
  const int t = threadIdx.x;
  const T2 wB = twr[(t & 7) * (P1 / 64)];     // W_64^k, k = t mod 8
  const T2 wC = twr[(t & 63) * (P1 / 512)];   // W_512^k, k = t mod 64
  T2 wD[8 / RL];
#pragma unroll
  for (int i = 0; i < 8 / RL; i++) wD[i] = twr[t + NT * i];  // W_P1^b, b = t + NT i
  // stage A: radix 8, Ns = 1: outputs at 8 t + r
  dftR<8, DIR>(x);
#pragma unroll
  for (int r = 0; r < 8; r++) lds[rpad(8 * t + r)] = x[r];
  __syncthreads();
  // stages B (Ns = 8) and C (Ns = 64): butterfly j = t reads j + NT r, writes (j - k) 8 + k + Ns r
#pragma unroll
  for (int st = 0; st < 2; st++) {
    const int Ns = st == 0 ? 8 : 64;
    const int k = t & (Ns - 1);
#pragma unroll
    for (int r = 0; r < 8; r++) x[r] = lds[rpad(t + NT * r)];
    twiddle_run<8, DIR>(x, st == 0 ? wB : wC);
    dftR<8, DIR>(x);
    __syncthreads();
    const int base = (t - k) * 8 + k;
#pragma unroll
    for (int r = 0; r < 8; r++) lds[rpad(base + Ns * r)] = x[r];
    __syncthreads();
  }
  // stage D: radix RL, Ns = 512: butterflies b = t + NT i (i < 8/RL) read b + 512 r (k = b);
  // outputs b + 512 r = t + NT (i + (8/RL) r)
#pragma unroll
  for (int i = 0; i < 8 / RL; i++) {
    const int b = t + NT * i;
    T2 v[RL];
#pragma unroll
    for (int r = 0; r < RL; r++) v[r] = lds[rpad(b + 512 * r)];
    twiddle_run<RL, DIR>(v, wD[i]);
    dftR<RL, DIR>(v);
#pragma unroll
    for (int r = 0; r < RL; r++) x[i + (8 / RL) * r] = v[r];
  }
}

// Row FFT with 8 elements per thread in which only one exchange crosses waves (NT = P1/8 = 64 R
// threads, R = P1/512 waves; thread t = 64 w + l).  Four-step with P1 = 8 x NT: x[m] = row element
// n = t + NT m on entry, so the 8-point DFT over m runs in registers (k1), times W_P1^{t k1}.
// One block-wide exchange (barrier) hands the NT values of slice k1 to wave k1 / (8/R), slot
// s = k1 mod (8/R).  Each wave then transforms its 8/R slices of length 64 R alone, in its own
// region of lds: radix R over j (n2 = l + 64 j) in registers, times W_{64R}^{l k1''}, leaving
// 8 sequences of 64 values over the lanes (register d = s R + k1''); these are 8 x 8 four-step
// DFTs with two exchanges inside the region (padded q + q/8: conflict-free strided reads), so
// no workgroup barrier follows the first.
// On exit x[m] = X[row8w_k0<P1>(t) + NT m].  Region of wave w: lds[576 w, 576 w + 576); lds must
// hold 576 R elements.  The caller must barrier before lds is written by other waves again.
template <int P1> struct Row8wShape {
  static constexpr int NT = P1 / 8;
  static constexpr int R = P1 / 512;            // waves; radix of the in-wave first step
  static constexpr int SPW = 8 / R;             // slices per wave
  static constexpr int REG = 576;               // elements per wave region (512 + 64 padding)
  static constexpr int LDSN = REG * R;
  static_assert(R == 2 || R == 4 || R == 8, "row length 1024, 2048 or 4096");
};

// first output index of thread t: X[k0 + NT m] is its register m
template <int P1> __device__ __forceinline__ int row8w_k0(int t) {
  using S = Row8wShape<P1>;
  const int w = t >> 6, l = t & 63, d = l >> 3, ka = l & 7;
  const int s = d / S::R, k1pp = d % S::R;
  return w * S::SPW + s + 8 * (k1pp + S::R * ka);
}

template <typename T2, int P1, int DIR>
__device__ __forceinline__ void row_fft8w(T2 (&x)[8], T2* lds, const T2* __restrict__ twr) {
  using S = Row8wShape<P1>;
  constexpr int NT = S::NT, R = S::R, SPW = S::SPW, REG = S::REG;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  // twiddle bases up front: W_P1^t, W_{64R}^l = W_P1^{8 l}, W_64^a = W_P1^{a P1/64}
  const T2 w1 = twr[t];
  const T2 w2 = twr[8 * l];
  const T2 w3 = twr[(l & 7) * (P1 / 64)];
  // step 1: 8-point DFT over n1 = m for n2 = t, times W_P1^{t k1}
  dftR<8, DIR>(x);
  twiddle_run<8, DIR>(x, w1);
  // slice k1 -> wave k1 / SPW, slot k1 % SPW, position n2 = t
#pragma unroll
  for (int k1 = 0; k1 < 8; k1++) lds[(k1 / SPW) * REG + (k1 % SPW) * NT + t] = x[k1];
  __syncthreads();
  T2* reg = lds + w * REG;
#pragma unroll
  for (int s = 0; s < SPW; s++)
#pragma unroll
    for (int j = 0; j < R; j++) x[s * R + j] = reg[s * NT + l + 64 * j];
  wave_lds_sync();
  // radix R over j for each slot, times W_{64R}^{l k1''}: register d = s R + k1'' holds element l
  // of the d-th 64-point sequence
#pragma unroll
  for (int s = 0; s < SPW; s++) {
    dftR<R, DIR>(x + s * R);
    twiddle_run<R, DIR>(x + s * R, w2);
  }
  auto q = [](int i) { return i + (i >> 3); };
  // 64-point DFTs over l = a + 8 b: lane (d, a) gathers b = 0..7
#pragma unroll
  for (int d = 0; d < 8; d++) reg[q(64 * d + l)] = x[d];
  wave_lds_sync();
  const int dd = l >> 3, a = l & 7;
#pragma unroll
  for (int b = 0; b < 8; b++) x[b] = reg[q(64 * dd + a + 8 * b)];
  wave_lds_sync();
  dftR<8, DIR>(x);
  twiddle_run<8, DIR>(x, w3);  // W_64^{a ka}
  // lane (d, ka) gathers a = 0..7
#pragma unroll
  for (int ka = 0; ka < 8; ka++) reg[q(64 * dd + 8 * ka + a)] = x[ka];
  wave_lds_sync();
#pragma unroll
  for (int aa = 0; aa < 8; aa++) x[aa] = reg[q(64 * dd + 8 * a + aa)];
  dftR<8, DIR>(x);
}

}  // namespace pgw
