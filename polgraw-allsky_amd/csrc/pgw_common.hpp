// pgw_common.hpp - device building blocks shared by the search-core kernels:
// complex arithmetic, in-register small DFTs (radix 2/4/8/16) and one Stockham
// autosort stage over data staged in LDS.
//
// FFT layout convention (used by every two-pass transform in this library):
//   L = L1 * L2,  input index n = n1 + L1*n2,  output index k = L2*k1 + k2.
//   column pass: for each n1, an L2-point DFT over n2, times W_L^{n1 k2},
//                stored as Z[k2][n1] (row-major, rows of length L1);
//   row pass:    for each k2, an L1-point DFT over n1 of row Z[k2][.],
//                giving X[L2*k1 + k2] at Z[k2][k1] ("T-layout").
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgw {

template <typename T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <typename T> using cplx = typename vec2<T>::type;

template <typename T2> __device__ __forceinline__ T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
  T2 r; r.x = a; r.y = b; return r;
}
// Loads through a pointer that is known to address device (global) memory.  Pointers read from
// memory (template descriptors) are generic to the compiler, which then emits flat loads; these
// helpers cast to the global address space so the loads are global_load and wait on vmcnt only.
template <typename T> using gptr = const __attribute__((address_space(1))) T*;
__device__ __forceinline__ double ld_global(const double* p, size_t i) { return ((gptr<double>)p)[i]; }
template <typename T2> __device__ __forceinline__ T2 ld_global(const T2* p, size_t i) {
  using V = decltype(T2::x) __attribute__((ext_vector_type(2)));
  const V v = ((gptr<V>)p)[i];
  return mk<T2>(v.x, v.y);
}
// Raw buffer over [base, base + bytes) with a workgroup-uniform base: an access's address is the
// base + a 32-bit per-lane byte offset + a uniform byte offset held in a scalar register, so no
// 64-bit vector address arithmetic is spent per access (a plain pointer access costs one 64-bit
// vector add each).  Resource word 3 = 0x00020000: the gfx9-family raw buffer (32-bit data
// format, no swizzle, no stride).  The hardware range check compares only the lane offset (voff
// plus the instruction offset) with `bytes`, not the scalar offset, so callers keep
// voff + soff + access size <= bytes by their own index bounds (every use below does) and never
// rely on the check.  Stores hold at most 64 bits of data: with 128-bit buffer stores and a
// scalar offset, FP64 spectra came out wrong at the tail of some dispatches (most likely a
// store-data hazard the compiler left uncovered; plain 16-byte pointer stores are exact, and
// the FP64 variant test with 8 templates per launch guards this).
struct BufRes {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ BufRes(const void* base, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000)) {}
  // V = double, float2 or double2 (8 or 16 bytes); AUX = cache policy (0; 2: non-temporal; 16:
  // device scope (SC1), the load misses any line the L2 of this XCD may hold stale)
  template <typename V, int AUX = 0> __device__ __forceinline__ V load(uint32_t voff, uint32_t soff) const {
    if constexpr (sizeof(V) == 8) {
      return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
    } else {
      static_assert(sizeof(V) == 16, "8- or 16-byte loads");
      return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, AUX));
    }
  }
  template <int AUX = 0> __device__ __forceinline__ void store(float v, uint32_t voff, uint32_t soff) const {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, AUX);
  }
  template <int AUX = 0> __device__ __forceinline__ void store(double v, uint32_t voff, uint32_t soff) const {
    using U = uint32_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, (int)voff, (int)soff, AUX);
  }
  __device__ __forceinline__ void store(float2 v, uint32_t voff, uint32_t soff) const {
    using U = uint32_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, (int)voff, (int)soff, 0);
  }
  // cache policy AUX (bits: 1 SC0, 2 NT, 16 SC1)
  template <int AUX> __device__ __forceinline__ void store_c(float2 v, uint32_t voff, uint32_t soff) const {
    using U = uint32_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, (int)voff, (int)soff, AUX);
  }
  // non-temporal (cache policy NT): streamed out without keeping the line in L2
  __device__ __forceinline__ void store_nt(float2 v, uint32_t voff, uint32_t soff) const {
    using U = uint32_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, (int)voff, (int)soff, 2);
  }
};

template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { return mk<T2>(a.x + b.x, a.y + b.y); }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { return mk<T2>(a.x - b.x, a.y - b.y); }
template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return mk<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T2> __device__ __forceinline__ T2 conjg(T2 a) { return mk<T2>(a.x, -a.y); }
// acc + a e with two fused multiply-adds per component (cadd(acc, cmul(a, e)) takes a multiply, a
// fused multiply-add and an add per component, and rounds once more)
template <typename T2> __device__ __forceinline__ T2 cmac(T2 acc, T2 a, T2 e) {
  return mk<T2>(fma(a.x, e.x, fma(-a.y, e.y, acc.x)), fma(a.x, e.y, fma(a.y, e.x, acc.y)));
}
// twiddle from a forward-sign table entry: DIR = -1 forward (as stored), +1 backward (conjugate)
template <int DIR, typename T2> __device__ __forceinline__ T2 tw_dir(T2 w) { return DIR < 0 ? w : conjg(w); }
// multiply by W_4 = exp(DIR * i pi/2): -i (forward) or +i (backward)
template <int DIR, typename T2> __device__ __forceinline__ T2 mul_w4(T2 d) {
  return DIR < 0 ? mk<T2>(d.y, -d.x) : mk<T2>(-d.y, d.x);
}

// ---------------------------------------------------------------- small DFTs
// In-place R-point DFT of v[0..R-1], X[q] = sum_r v[r] exp(DIR 2 pi i r q / R).
template <int DIR, typename T2> __device__ __forceinline__ void dft2(T2* v) {
  T2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR, typename T2> __device__ __forceinline__ void dft4(T2& v0, T2& v1, T2& v2, T2& v3) {
  T2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = mul_w4<DIR>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <int DIR, typename T2> __device__ __forceinline__ void dft4(T2* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }

// multiply by exp(DIR 2 pi i K / 16), K = 0..7
template <int DIR, int K, typename T2> __device__ __forceinline__ T2 rot16(T2 a) {
  using T = decltype(T2::x);
  constexpr double cs[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                            0.38268343236508977173, 0.0, -0.38268343236508977173,
                            -0.70710678118654752440, -0.92387953251128675613};
  constexpr double sn[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440,
                            0.92387953251128675613, 1.0, 0.92387953251128675613,
                            0.70710678118654752440, 0.38268343236508977173};
  constexpr T c = (T)cs[K], s = (T)(DIR * sn[K]);
  return mk<T2>(a.x * c - a.y * s, a.x * s + a.y * c);
}

template <int DIR, typename T2> __device__ __forceinline__ void dft8(T2* v) {
  T2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  T2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  o1 = rot16<DIR, 2>(o1);
  o2 = mul_w4<DIR>(o2);
  o3 = rot16<DIR, 6>(o3);
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// multiply by exp(DIR 2 pi i K / 16), K = 0..15 (K >= 8: the negated rotation by K - 8; the
// negation folds into the following additions as a source modifier)
template <int DIR, int K, typename T2> __device__ __forceinline__ T2 rot16n(T2 a) {
  if constexpr (K == 0) return a;
  else if constexpr (K == 4) return mul_w4<DIR>(a);
  else if constexpr (K < 8) return rot16<DIR, K>(a);
  else {
    const T2 r = rot16n<DIR, K - 8>(a);
    return mk<T2>(-r.x, -r.y);
  }
}

// 16 = 4 x 4: r = q + 4 p, k = k1 + 4 k2; four 4-point DFTs over p, twiddles W16^(q k1), four
// 4-point DFTs over q (four general and four pi/4 rotations: 8 instructions fewer than the
// radix-2 split into two 8-point DFTs)
template <int DIR, typename T2> __device__ __forceinline__ void dft16(T2* v) {
  T2 a0[4], a1[4], a2[4], a3[4];
#pragma unroll
  for (int p = 0; p < 4; p++) { a0[p] = v[4 * p]; a1[p] = v[1 + 4 * p]; a2[p] = v[2 + 4 * p]; a3[p] = v[3 + 4 * p]; }
  dft4<DIR>(a0); dft4<DIR>(a1); dft4<DIR>(a2); dft4<DIR>(a3);
  a1[1] = rot16n<DIR, 1>(a1[1]); a1[2] = rot16n<DIR, 2>(a1[2]); a1[3] = rot16n<DIR, 3>(a1[3]);
  a2[1] = rot16n<DIR, 2>(a2[1]); a2[2] = rot16n<DIR, 4>(a2[2]); a2[3] = rot16n<DIR, 6>(a2[3]);
  a3[1] = rot16n<DIR, 3>(a3[1]); a3[2] = rot16n<DIR, 6>(a3[2]); a3[3] = rot16n<DIR, 9>(a3[3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    T2 w0 = a0[k1], w1 = a1[k1], w2 = a2[k1], w3 = a3[k1];
    dft4<DIR>(w0, w1, w2, w3);
    v[k1] = w0; v[k1 + 4] = w1; v[k1 + 8] = w2; v[k1 + 12] = w3;
  }
}

// a * exp(DIR 2 pi i k / R), R in {4, 8, 16}, 0 < k < R/2 (k a compile-time constant after
// unrolling), as the twiddles of dft4/dft8/dft16
template <int R, int DIR, typename T2> __device__ __forceinline__ T2 rot_r(T2 a, int k) {
  const int K = k * (16 / R);
  switch (K) {
    case 1: return rot16<DIR, 1>(a);
    case 2: return rot16<DIR, 2>(a);
    case 3: return rot16<DIR, 3>(a);
    case 4: return mul_w4<DIR>(a);
    case 5: return rot16<DIR, 5>(a);
    case 6: return rot16<DIR, 6>(a);
    case 7: return rot16<DIR, 7>(a);
    default: return a;
  }
}

// LDS written by some lanes of a wave made visible to the wave's other lanes (no workgroup
// barrier: only the wave's own accesses are ordered)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int R, int DIR, typename T2> __device__ __forceinline__ void dftR(T2* v) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 16) dft16<DIR>(v);
  else static_assert(R == 2, "unsupported radix");
}

// R-point DFT of v[0..R-1] whose inputs v[r], r >= NZ, are zero (input pruning): an even/odd
// split (the one of dft4 and dft8), with sub-transforms of one nonzero input reduced to copies and
// of none to zeros; the result equals dftR's on such inputs to rounding (bit for bit for R <= 8).
template <int R, int NZ, int DIR, typename T2> __device__ __forceinline__ void dftP(T2* v) {
  if constexpr (NZ >= R) {
    dftR<R, DIR>(v);
  } else if constexpr (NZ <= 0) {
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = mk<T2>(0, 0);
  } else if constexpr (NZ == 1) {
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = v[0];
  } else {
    constexpr int H = R / 2;
    T2 e[H], o[H];
#pragma unroll
    for (int i = 0; i < H; i++) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    dftP<H, (NZ + 1) / 2, DIR>(e);
    dftP<H, NZ / 2, DIR>(o);
    if constexpr (NZ / 2 > 0) {
#pragma unroll
      for (int k = 1; k < H; k++) o[k] = rot_r<R, DIR>(o[k], k);
    }
#pragma unroll
    for (int k = 0; k < H; k++) {
      if constexpr (NZ / 2 > 0) {
        v[k] = cadd(e[k], o[k]);
        v[k + H] = csub(e[k], o[k]);
      } else {
        v[k] = e[k];
        v[k + H] = e[k];
      }
    }
  }
}

// ------------------------------------------------------------ Stockham stage
// One radix-R Stockham autosort stage of C interleaved P-point transforms held in
// LDS as s[e*C + c] (e = element index 0..P-1, c = transform 0..C-1), in place.
// Ns = product of the radices of the stages already done.  tw[m] = W_P^m (forward sign).
// All NT threads of the block must call it (two barriers inside).
template <typename T2, int P, int C, int R, int NT, int DIR>
__device__ __forceinline__ void stockham_stage(T2* s, int Ns, const T2* __restrict__ tw) {
  constexpr int NBF = (P / R) * C;               // butterflies in the stage
  constexpr int BPT = (NBF + NT - 1) / NT;       // per thread
  T2 v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = tid + q * NT;
    if (NBF % NT == 0 || b < NBF) {
      const int c = b % C, j = b / C;
#pragma unroll
      for (int r = 0; r < R; r++) v[q][r] = s[(j + r * (P / R)) * C + c];
      const int k = j & (Ns - 1);
      if (Ns > 1 && k) {
        const int step = (P / (Ns * R)) * k;
#pragma unroll
        for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], tw_dir<DIR>(tw[r * step]));
      }
      dftR<R, DIR>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; q++) {
    const int b = tid + q * NT;
    if (NBF % NT == 0 || b < NBF) {
      const int c = b % C, j = b / C;
      const int k = j & (Ns - 1);
      const int idx = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) s[(idx + r * Ns) * C + c] = v[q][r];
    }
  }
  __syncthreads();
}

// Radix plans for the supported transform sizes (product of radices = P).
template <int P> struct RadixPlan;
template <> struct RadixPlan<16> { static constexpr int n = 1; static constexpr int r[4] = {16, 1, 1, 1}; };
template <> struct RadixPlan<32> { static constexpr int n = 2; static constexpr int r[4] = {8, 4, 1, 1}; };
template <> struct RadixPlan<64> { static constexpr int n = 2; static constexpr int r[4] = {8, 8, 1, 1}; };
template <> struct RadixPlan<128> { static constexpr int n = 2; static constexpr int r[4] = {16, 8, 1, 1}; };
template <> struct RadixPlan<256> { static constexpr int n = 2; static constexpr int r[4] = {16, 16, 1, 1}; };
template <> struct RadixPlan<512> { static constexpr int n = 3; static constexpr int r[4] = {8, 8, 8, 1}; };
template <> struct RadixPlan<1024> { static constexpr int n = 3; static constexpr int r[4] = {16, 16, 4, 1}; };
template <> struct RadixPlan<2048> { static constexpr int n = 3; static constexpr int r[4] = {16, 16, 8, 1}; };
template <> struct RadixPlan<4096> { static constexpr int n = 3; static constexpr int r[4] = {16, 16, 16, 1}; };

// Full P-point transform of C interleaved columns in LDS (natural order in and out).
template <typename T2, int P, int C, int NT, int DIR>
__device__ __forceinline__ void lds_fft(T2* s, const T2* __restrict__ tw) {
  using RP = RadixPlan<P>;
  int Ns = 1;
  stockham_stage<T2, P, C, RP::r[0], NT, DIR>(s, Ns, tw);
  Ns *= RP::r[0];
  if constexpr (RP::n > 1) {
    stockham_stage<T2, P, C, RP::r[1], NT, DIR>(s, Ns, tw);
    Ns *= RP::r[1];
  }
  if constexpr (RP::n > 2) {
    stockham_stage<T2, P, C, RP::r[2], NT, DIR>(s, Ns, tw);
    Ns *= RP::r[2];
  }
  if constexpr (RP::n > 3) {
    stockham_stage<T2, P, C, RP::r[3], NT, DIR>(s, Ns, tw);
  }
}

// W_L^m from the two-level table: hi[m >> 10] * lo[m & 1023] (forward sign).
template <typename T2> __device__ __forceinline__ T2 tw_big(const T2* __restrict__ hi, const T2* __restrict__ lo, uint32_t m) {
  return cmul(hi[m >> 10], lo[m & 1023u]);
}

// Wave-level (64 lanes) reductions in a fixed butterfly order: deterministic.
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  return x;
}
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}
__device__ __forceinline__ float wave_max(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmaxf(x, __shfl_xor(x, off, 64));
  return x;
}

}  // namespace pgw
