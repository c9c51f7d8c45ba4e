// Size-specialised FFT passes for the hot grid extents (64, 128, 256, 512).
//
// Each transform of length N is spread over LANES threads holding E = N/LANES
// complex values in registers.  A Stockham pass of radix R gives each thread
// E/R butterflies b = lane + LANES*q (q < E/R): inputs at b + r*N/R, outputs
// at j + r*P with P the product of the earlier radices and j = (b - b%P)*R + b%P.
// Consecutive passes exchange through LDS; the first pass loads straight
// from global memory and the last pass stores straight to it.  The last
// pass of a transform leaves thread `lane` holding positions b + r*N/R_last,
// which are exactly the inputs of a first pass of radix R_last: an inverse
// transform followed by a forward transform with the radix order reversed
// keeps the data in registers in between (used by the x-row kernel).
#pragma once

#include <type_traits>

#include "ddx_kernels.h"

namespace ddx {

// --- plans -------------------------------------------------------------------

template <int N> struct RowPlan;          // x rows: one transform per row team
template <> struct RowPlan<64>  { static constexpr int LANES = 16, E = 4, NP = 3; static constexpr int R[3] = {4, 4, 4}; };
template <> struct RowPlan<128> { static constexpr int LANES = 32, E = 4, NP = 4; static constexpr int R[4] = {4, 4, 4, 2}; };
template <> struct RowPlan<256> { static constexpr int LANES = 64, E = 4, NP = 4; static constexpr int R[4] = {4, 4, 4, 4}; };
template <> struct RowPlan<512> { static constexpr int LANES = 64, E = 8, NP = 3; static constexpr int R[3] = {8, 8, 8}; };

template <int N> struct ColPlan;          // columns: W columns per workgroup
template <> struct ColPlan<64>  { static constexpr int LANES = 8,  E = 8,  W = 16, NP = 2; static constexpr int R[2] = {8, 8}; };
template <> struct ColPlan<128> { static constexpr int LANES = 8,  E = 16, W = 16, NP = 2; static constexpr int R[2] = {16, 8}; };
template <> struct ColPlan<256> { static constexpr int LANES = 16, E = 16, W = 8,  NP = 2; static constexpr int R[2] = {16, 16}; };
template <> struct ColPlan<512> { static constexpr int LANES = 64, E = 8,  W = 4,  NP = 3; static constexpr int R[3] = {8, 8, 8}; };

// Radix variants (V = 1, ddx_plan_set_tuning DDX_TUNE_RADIX), for the
// autotuner to choose: rows of 64..256 points with more points per thread
// and fewer passes, 512-point rows with radix 4 passes (the FP64 rows of 10
// and 12 fields keep 8 points per thread, within the registers), and columns
// with fewer points per thread and more passes.
template <int N> struct RowPlanAlt;
template <> struct RowPlanAlt<64>  { static constexpr int LANES = 8,  E = 8,  NP = 2; static constexpr int R[2] = {8, 8}; };
template <> struct RowPlanAlt<128> { static constexpr int LANES = 16, E = 8,  NP = 3; static constexpr int R[3] = {8, 4, 4}; };
template <> struct RowPlanAlt<256> { static constexpr int LANES = 32, E = 8,  NP = 3; static constexpr int R[3] = {8, 8, 4}; };
template <> struct RowPlanAlt<512> { static constexpr int LANES = 64, E = 8,  NP = 4; static constexpr int R[4] = {8, 4, 4, 4}; };

template <int N> struct ColPlanAlt;
template <> struct ColPlanAlt<64>  { static constexpr int LANES = 16, E = 4,  W = 8, NP = 3; static constexpr int R[3] = {4, 4, 4}; };
template <> struct ColPlanAlt<128> { static constexpr int LANES = 16, E = 8,  W = 8, NP = 3; static constexpr int R[3] = {8, 4, 4}; };
template <> struct ColPlanAlt<256> { static constexpr int LANES = 32, E = 8,  W = 4, NP = 3; static constexpr int R[3] = {8, 8, 4}; };
template <> struct ColPlanAlt<512> { static constexpr int LANES = 32, E = 16, W = 4, NP = 3; static constexpr int R[3] = {16, 8, 4}; };

template <int N, int V> using RowPlanV = typename std::conditional<V == 0, RowPlan<N>, RowPlanAlt<N>>::type;
template <int N, int V> using ColPlanV = typename std::conditional<V == 0, ColPlan<N>, ColPlanAlt<N>>::type;

// Column tiles of double width (WIDE, n = 256 only): FP32 tiles are then 16
// columns wide, so a tile row is a whole 128-byte line, and FP64 tiles of
// the plain passes too (256-byte rows; their exchanges then go through the
// half-size real buffer): 6 % faster y passes at 256^3.  The FP64 spectral
// z pass keeps 8 columns, where 16 were slower.
template <int N, bool WIDE, int V = 0>
constexpr int col_width() {
  return ColPlanV<N, V>::W * (WIDE ? 2 : 1);
}

// LDS index of position m in the exchange written by a pass of radix R after
// P points: runs of P contiguous outputs start every P*R positions; shifting
// run t by t*P (R + 1 odd) spreads the runs of consecutive threads over
// distinct banks, and consecutive positions stay nearly contiguous.
template <int P, int R>
__device__ __forceinline__ int xpad(int m) { return m + (m / (P * R)) * P; }

// The exchange indices are affine in the leg r, so each thread computes one
// base and the legs use constant offsets (immediate LDS offsets): with
// k = b mod P, xpad(out_pos(b, r)) = xpad(out_pos(b, 0)) + r P because
// k + r P < P R; and with S = N / R2 a multiple of P R,
// xpad(in_pos(b, r)) = xpad(in_pos(b, 0)) + r (S + S / R).
template <int N, int P, int R, int R2>
constexpr int RSTEP() { return N / R2 + N / R2 / R; }

// The first exchange of a row transform (P = 1, radix <= 4, one transform
// per buffer) writes positions R b + r and reads lane + S r.  The padding
// above leaves the reads with two-way bank conflicts; this XOR swizzle of
// bits 0-1 by bits 3-4 is a permutation of [0, n) under which both the
// writes and the reads of 8 (or 16) consecutive lanes fall in distinct
// 16-byte bank groups.
__device__ __forceinline__ int xswz(int m) { return m ^ ((m >> 3) & 3) ^ ((m >> 4) & 3); }

template <int N, int WS, int P, int R>
constexpr bool swizzled() { return WS == 1 && P == 1 && R <= 4 && N >= 64 && N <= 256; }

// LDS entries per transform needed by the padded exchanges of a plan
template <int N, class PL>
constexpr int exch_size() {
  int mx = 0;
  for (int i = 0; i < PL::NP; ++i) mx = (N / PL::R[i]) > mx ? (N / PL::R[i]) : mx;
  return N + mx + 1;
}

// Butterflies of one pass with twiddles from a per-thread register cache
// (conjugated for D > 0).  POW = false: c[q*R + r] = forward twiddle of
// butterfly q, leg r.  POW = true: c[q] = the leg-1 twiddle w of butterfly q
// and leg r uses w^r = w^(r/2) w^(r - r/2), a product tree of depth
// ceil(log2 r) (a few ulp from the table value); the cache then takes E/R
// entries instead of E, loaded before the data.
template <typename T, int D, int E, int R, int P, bool POW>
__device__ __forceinline__ void reg_pass_cached(cpx<T> (&v)[E], const cpx<T> (&c)[E]) {
#pragma unroll
  for (int q = 0; q < E / R; ++q) {
    if (P > 1) {
      cpx<T> pw[R];
      if (POW) {
        pw[1] = c[q];
#pragma unroll
        for (int r = 2; r < R; ++r) pw[r] = pw[r / 2] * pw[r - r / 2];
      } else {
#pragma unroll
        for (int r = 1; r < R; ++r) pw[r] = c[q * R + r];
      }
#pragma unroll
      for (int r = 1; r < R; ++r) v[q * R + r] = v[q * R + r] * mk(pw[r].x, D < 0 ? pw[r].y : -pw[r].y);
    }
    dft<D, R>(&v[q * R]);
  }
}

// Butterflies of one pass on register data (twiddle + DFT_R).
template <typename T, int D, int N, int LANES, int E, int R, int P>
__device__ __forceinline__ void reg_pass(cpx<T> (&v)[E], int lane, const cpx<T>* __restrict__ tw) {
#pragma unroll
  for (int q = 0; q < E / R; ++q) {
    const int b = lane + LANES * q;
    if (P > 1) {
      const int k = b % P;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cpx<T> w = tw[r * k * (N / (P * R))];
        v[q * R + r] = v[q * R + r] * mk(w.x, D < 0 ? w.y : -w.y);
      }
    }
    dft<D, R>(&v[q * R]);
  }
}

template <int N, int LANES, int R, int P>
__device__ __forceinline__ int out_pos(int lane, int q, int r) {
  const int b = lane + LANES * q;
  const int k = b % P;
  return (b - k) * R + k + r * P;
}

template <int N, int LANES, int R>
__device__ __forceinline__ int in_pos(int lane, int q, int r) {
  return lane + LANES * q + r * (N / R);
}

// Passes of a transform whose first-pass inputs are in v.  Exchanges go
// through buf at index xpad(pos) * WS + c (WS transforms interleaved).  PL
// is the plan; REV runs its radices in reverse order (forward transforms
// that follow an inverse one in the x-row kernel).
template <typename T, int D, int N, class PL, bool REV, int WS, int I = 0, int P = 1>
__device__ __forceinline__ void run_passes(cpx<T> (&v)[PL::E], int lane, int c, cpx<T>* buf,
                                           const cpx<T>* __restrict__ tw) {
  constexpr int NP = PL::NP;
  constexpr int R = PL::R[REV ? NP - 1 - I : I];
  reg_pass<T, D, N, PL::LANES, PL::E, R, P>(v, lane, tw);
  if constexpr (I + 1 < NP) {
    constexpr int R2 = PL::R[REV ? NP - 2 - I : I + 1];
#pragma unroll
    for (int q = 0; q < PL::E / R; ++q) {
      const int wb = xpad<P, R>(out_pos<N, PL::LANES, R, P>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wb + r * (P * WS)] = v[q * R + r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PL::E / R2; ++q) {
      const int rb = xpad<P, R>(in_pos<N, PL::LANES, R2>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R2; ++r) v[q * R2 + r] = buf[rb + r * (RSTEP<N, P, R, R2>() * WS)];
    }
    __syncthreads();
    run_passes<T, D, N, PL, REV, WS, I + 1, P * R>(v, lane, c, buf, tw);
  }
}

// Register twiddle cache of a plan: pass I of the (REV-ordered) radix
// sequence gets E entries, c[I][q*R + r] = tw[r * (b % P) * N/(P R)]
// (POW = false), or E/R entries c[I][q] = tw[(b % P) * N/(P R)] (POW = true).
template <typename T, int N, class PL, bool REV, bool POW, int I = 0, int P = 1>
__device__ __forceinline__ void fill_twiddles(cpx<T> (&c)[PL::NP][PL::E], int lane,
                                              const cpx<T>* __restrict__ tw) {
  constexpr int NP = PL::NP;
  constexpr int R = PL::R[REV ? NP - 1 - I : I];
#pragma unroll
  for (int q = 0; q < PL::E / R; ++q) {
    const int k = (lane + PL::LANES * q) % P;
    if (POW) {
      c[I][q] = P > 1 ? tw[k * (N / (P * R))] : mk(T(1), T(0));
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r)
        c[I][q * R + r] = (P > 1 && r > 0) ? tw[r * k * (N / (P * R))] : mk(T(1), T(0));
    }
  }
  if constexpr (I + 1 < NP) fill_twiddles<T, N, PL, REV, POW, I + 1, P * R>(c, lane, tw);
}

// run_passes with twiddles from a register cache filled by fill_twiddles
template <typename T, int D, int N, class PL, bool REV, int WS, bool POW, int I = 0, int P = 1>
__device__ __forceinline__ void run_passes_cached(cpx<T> (&v)[PL::E], int lane, int c, cpx<T>* buf,
                                                  const cpx<T> (&twc)[PL::NP][PL::E]) {
  constexpr int NP = PL::NP;
  constexpr int R = PL::R[REV ? NP - 1 - I : I];
  reg_pass_cached<T, D, PL::E, R, P, POW>(v, twc[I]);
  if constexpr (I + 1 < NP) {
    constexpr int R2 = PL::R[REV ? NP - 2 - I : I + 1];
    if constexpr (swizzled<N, WS, P, R>()) {
#pragma unroll
      for (int q = 0; q < PL::E / R; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) buf[xswz(out_pos<N, PL::LANES, R, P>(lane, q, r))] = v[q * R + r];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PL::E / R2; ++q)
#pragma unroll
        for (int r = 0; r < R2; ++r) v[q * R2 + r] = buf[xswz(in_pos<N, PL::LANES, R2>(lane, q, r))];
    } else {
#pragma unroll
      for (int q = 0; q < PL::E / R; ++q) {
        const int wb = xpad<P, R>(out_pos<N, PL::LANES, R, P>(lane, q, 0)) * WS + c;
#pragma unroll
        for (int r = 0; r < R; ++r) buf[wb + r * (P * WS)] = v[q * R + r];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PL::E / R2; ++q) {
        const int rb = xpad<P, R>(in_pos<N, PL::LANES, R2>(lane, q, 0)) * WS + c;
#pragma unroll
        for (int r = 0; r < R2; ++r) v[q * R2 + r] = buf[rb + r * (RSTEP<N, P, R, R2>() * WS)];
      }
    }
    __syncthreads();
    run_passes_cached<T, D, N, PL, REV, WS, POW, I + 1, P * R>(v, lane, c, buf, twc);
  }
}

// run_passes_cached with each exchange done in two halves through a buffer of
// reals (real parts, then imaginary parts): half the LDS of a complex
// exchange, for kernels whose occupancy LDS limits
template <typename T, int D, int N, class PL, bool REV, int WS, bool POW, int I = 0, int P = 1>
__device__ __forceinline__ void run_passes_split(cpx<T> (&v)[PL::E], int lane, int c, T* buf,
                                                 const cpx<T> (&twc)[PL::NP][PL::E]) {
  constexpr int NP = PL::NP;
  constexpr int R = PL::R[REV ? NP - 1 - I : I];
  reg_pass_cached<T, D, PL::E, R, P, POW>(v, twc[I]);
  if constexpr (I + 1 < NP) {
    constexpr int R2 = PL::R[REV ? NP - 2 - I : I + 1];
    T nx[PL::E];
#pragma unroll
    for (int q = 0; q < PL::E / R; ++q) {
      const int wb = xpad<P, R>(out_pos<N, PL::LANES, R, P>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wb + r * (P * WS)] = v[q * R + r].x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PL::E / R2; ++q) {
      const int rb = xpad<P, R>(in_pos<N, PL::LANES, R2>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R2; ++r) nx[q * R2 + r] = buf[rb + r * (RSTEP<N, P, R, R2>() * WS)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PL::E / R; ++q) {
      const int wb = xpad<P, R>(out_pos<N, PL::LANES, R, P>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[wb + r * (P * WS)] = v[q * R + r].y;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PL::E / R2; ++q) {
      const int rb = xpad<P, R>(in_pos<N, PL::LANES, R2>(lane, q, 0)) * WS + c;
#pragma unroll
      for (int r = 0; r < R2; ++r)
        v[q * R2 + r] = mk(nx[q * R2 + r], buf[rb + r * (RSTEP<N, P, R, R2>() * WS)]);
    }
    __syncthreads();
    run_passes_split<T, D, N, PL, REV, WS, POW, I + 1, P * R>(v, lane, c, buf, twc);
  }
}

template <class PL>
constexpr bool uniform_radix() {
  for (int i = 1; i < PL::NP; ++i)
    if (PL::R[i] != PL::R[0]) return false;
  return true;
}

// positions held after the last pass (R_last, P_last = N / R_last): b + r*N/R_last
template <int N, class PL, bool REV>
__device__ __forceinline__ int final_pos(int lane, int q, int r) {
  constexpr int RL = PL::R[REV ? 0 : PL::NP - 1];
  return lane + PL::LANES * q + r * (N / RL);
}

// ---------------------------------------------------------------------------
// x-row kernel: NIN half-spectrum rows -> x-space products -> NOUT rows
// ---------------------------------------------------------------------------

template <int PH, int ND> struct PhysIO;
template <> struct PhysIO<PH_HYDRO, 3> { static constexpr int NIN = 6, NOUT = 3; };
template <> struct PhysIO<PH_HYDRO, 2> { static constexpr int NIN = 3, NOUT = 2; };
template <> struct PhysIO<PH_BOUSS, 3> { static constexpr int NIN = 10, NOUT = 4; };
template <> struct PhysIO<PH_BOUSS, 2> { static constexpr int NIN = 6, NOUT = 3; };
template <> struct PhysIO<PH_MHD, 3>   { static constexpr int NIN = 12, NOUT = 6; };
template <> struct PhysIO<PH_MHD, 2>   { static constexpr int NIN = 6, NOUT = 3; };

// Loads of input pair p of a row (fields 2p and 2p + 1) at the first-pass
// positions of this lane, from clamped in-range entries
template <typename T, int N, class PL, int NIN>
__device__ __forceinline__ void row_loads(cpx<T> (&la)[PL::E], cpx<T> (&lb)[PL::E], const RowArgs<T>& a,
                                          int p, long rl, int lane, int cx) {
  constexpr int LANES = PL::LANES, E = PL::E, R0 = PL::R[0];
  const bool hasb = (2 * p + 1 < NIN);
  const cpx<T>* __restrict__ ia = a.in[2 * p] + rl * a.kxp;
  const cpx<T>* __restrict__ ib = a.in[hasb ? 2 * p + 1 : 2 * p] + rl * a.kxp;
#pragma unroll
  for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
    for (int rr = 0; rr < R0; ++rr) {
      const int m = in_pos<N, LANES, R0>(lane, qq, rr);
      const int idx = m <= cx ? m : (m >= N - cx ? N - m : 0);
      la[qq * R0 + rr] = ia[idx];
      lb[qq * R0 + rr] = hasb ? ib[idx] : mk(T(0), T(0));
    }
}

// XIN: the inputs arrive in x space (a.xin[f], rows of N reals, the values
// a backward transform of the same fields gives), so the c2r passes are
// skipped; a variant of its own, with XF.
template <typename T, int N, int PH, int ND, int V = 0, bool XF = false, bool XIN = false>
__global__ void __launch_bounds__(64) row_phys_fast(const RowArgs<T> a, long nrows) {
  using PL = RowPlanV<N, V>;
  constexpr int LANES = PL::LANES, E = PL::E;
  constexpr int RPB = 64 / LANES;
  constexpr int NIN = PhysIO<PH, ND>::NIN, NOUT = PhysIO<PH, ND>::NOUT;
  constexpr int NPI = (NIN + 1) / 2, NPO = (NOUT + 1) / 2;
  constexpr int R0 = PL::R[0];
  constexpr int BUF = exch_size<N, PL>();
  // FP64 rows of 12 fields park the x-space values of the first two input
  // pairs in LDS until the products (8 KB per row at n = 256), which keeps
  // the kernel within the registers of 3 wavefronts per SIMD
  constexpr int NPARK = (NIN >= 12 && sizeof(T) == 8 && !XIN) ? 4 : 0;
  __shared__ cpx<T> lds[RPB][BUF];
  __shared__ T park[RPB][NPARK > 0 ? NPARK : 1][NPARK > 0 ? N : 1];
  const int team = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const long r = (long)blockIdx.x * RPB + team;
  const bool live = r < nrows;
  cpx<T>* buf = lds[team];
  const int cx = a.kx - 1;
  const cpx<T>* __restrict__ tw = a.tw;
  constexpr bool UNI = uniform_radix<PL>();
  // FP64: leg-1 twiddles only (powers formed in the butterflies), which frees
  // the registers the 12 x-space fields of the MHD row need
  constexpr bool POW = sizeof(T) == 8;
  cpx<T> twi[PL::NP][E];
  fill_twiddles<T, N, PL, false, POW>(twi, lane, tw);

  // Loads are unconditional, from clamped in-range addresses (the last row
  // for a team past the end, entry 0 for positions the 2/3 rule leaves
  // empty), and the unused values are zeroed afterwards: loads under
  // lane-dependent branches each end in a wait for memory, which serialises
  // a row's dozens of loads.  The loads of pair p + 1 are issued before the
  // transform of pair p (two pairs in flight per wavefront).
  const long rl = live ? r : nrows - 1;
  T q[NIN][E];
  constexpr int RLB = PL::R[PL::NP - 1];
  if constexpr (XIN) {
    // register qq * RLB + rr holds the c2r passes' final position of (lane, qq, rr)
#pragma unroll
    for (int f = 0; f < NIN; ++f) {
      const T* __restrict__ xr = a.xin[f] + rl * (long)N;
#pragma unroll
      for (int qq = 0; qq < E / RLB; ++qq)
#pragma unroll
        for (int rr = 0; rr < RLB; ++rr) q[f][qq * RLB + rr] = xr[final_pos<N, PL, false>(lane, qq, rr)];
    }
  } else {
  cpx<T> la[2][E], lb[2][E];
  row_loads<T, N, PL, NIN>(la[0], lb[0], a, 0, rl, lane, cx);
#pragma unroll
  for (int p = 0; p < NPI; ++p) {
    if (p + 1 < NPI) row_loads<T, N, PL, NIN>(la[(p + 1) & 1], lb[(p + 1) & 1], a, p + 1, rl, lane, cx);
    cpx<T> v[E];
    const bool hasb = (2 * p + 1 < NIN);
#pragma unroll
    for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
      for (int rr = 0; rr < R0; ++rr) {
        const int m = in_pos<N, LANES, R0>(lane, qq, rr);
        const bool keep = live && (m <= cx || m >= N - cx);
        const bool mirror = m >= N - cx;
        const cpx<T> va = la[p & 1][qq * R0 + rr], vb = lb[p & 1][qq * R0 + rr];
        // conj for mirrored positions; the k = 0 entry is taken as real.  The
        // k = 0 imaginary parts are scaled by 0 rather than selected away, so
        // that their loads stay unconditional.
        const T z0 = m == 0 ? T(0) : T(1);
        const T ay = (mirror ? -va.y : va.y) * z0;
        const T by = (mirror ? -vb.y : vb.y) * z0;
        v[qq * R0 + rr] = keep ? mk(va.x - by, ay + vb.x) : mk(T(0), T(0));
      }
    run_passes_cached<T, 1, N, PL, false, 1, POW>(v, lane, 0, buf, twi);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (2 * p + 1 < NPARK) {
        // each lane reads back only what it wrote: no synchronisation needed
        park[team][2 * p][e * LANES + lane] = v[e].x;
        park[team][2 * p + 1][e * LANES + lane] = v[e].y;
      } else {
        q[2 * p][e] = v[e].x;
        if (hasb) q[2 * p + 1][e] = v[e].y;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int f = 0; f < NPARK; ++f)
#pragma unroll
    for (int e = 0; e < E; ++e) q[f][e] = park[team][f][e * LANES + lane];
  }

  T o[NOUT][E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    T qi[12], oi[6];
#pragma unroll
    for (int f = 0; f < 12; ++f) qi[f] = f < NIN ? q[f][e] : T(0);
#pragma unroll
    for (int f = 0; f < 6; ++f) oi[f] = T(0);
    row_physics<T>(PH, ND, a.fpr_inv, qi, oi);
#pragma unroll
    for (int f = 0; f < NOUT; ++f) o[f][e] = oi[f];
  }
  // XF: an x-space forcing joins its product before the r2c (register
  // qq * RL + rr holds the c2r passes' final position of (lane, qq, rr)); a
  // variant of its own, so that the kernels without one keep their registers
#pragma unroll
  for (int f = 0; f < NOUT; ++f) {
    if (XF && a.xadd[f]) {
      const T* __restrict__ xa = a.xadd[f] + (live ? r : nrows - 1) * (long)N;
#pragma unroll
      for (int qq = 0; qq < E / RLB; ++qq)
#pragma unroll
        for (int rr = 0; rr < RLB; ++rr)
          o[f][qq * RLB + rr] += xa[final_pos<N, PL, false>(lane, qq, rr)];
    }
  }

  cpx<T> twf[PL::NP][E];
  if constexpr (UNI) {
#pragma unroll
    for (int i = 0; i < PL::NP; ++i)
#pragma unroll
      for (int e = 0; e < E; ++e) twf[i][e] = twi[i][e];
  } else {
    fill_twiddles<T, N, PL, true, POW>(twf, lane, tw);
  }
#pragma unroll
  for (int p = 0; p < NPO; ++p) {
    cpx<T> v[E];
    const bool hasb = (2 * p + 1 < NOUT);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk(o[2 * p][e], hasb ? o[2 * p + 1][e] : T(0));
    run_passes_cached<T, -1, N, PL, true, 1, POW>(v, lane, 0, buf, twf);
    constexpr int RL = PL::R[0];
#pragma unroll
    for (int qq = 0; qq < E / RL; ++qq)
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) buf[final_pos<N, PL, true>(lane, qq, rr)] = v[qq * RL + rr];
    __syncthreads();
    if (live) {
      cpx<T>* __restrict__ oa = a.out[2 * p] + r * a.kxp;
      cpx<T>* __restrict__ ob = hasb ? a.out[2 * p + 1] + r * a.kxp : nullptr;
      for (int k = lane; k < a.kxp; k += LANES) {
        cpx<T> ra = mk(T(0), T(0)), rb = mk(T(0), T(0));
        if (k < a.kx) {
          const cpx<T> zk = buf[k];
          const cpx<T> zm = conj(buf[(N - k) % N]);
          const cpx<T> s = zk + zm, d = zk - zm;
          ra = mk(T(0.5) * s.x, T(0.5) * s.y);
          rb = mk(T(0.5) * d.y, T(-0.5) * d.x);
        }
        st_nt(&oa[k], ra, a.nt);
        if (hasb) st_nt(&ob[k], rb, a.nt);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// plain x-row transforms (ddx_bwd / ddx_fwd): two rows of one field per team,
// the field (of several in one call) in blockIdx.y
// ---------------------------------------------------------------------------

// c2r of rows 2t and 2t + 1: z = A + i B from the two half spectra, one
// inverse transform, x rows = (Re z, Im z).  Loads as in row_phys_fast
// (clamped addresses, zeroed afterwards, the k = 0 imaginary part dropped).
template <typename T, int N, int V = 0>
__global__ void __launch_bounds__(64) row_c2r_fast(const RowArgs<T> a, long nrows) {
  using PL = RowPlanV<N, V>;
  constexpr int LANES = PL::LANES, E = PL::E, RPB = 64 / LANES;
  constexpr int R0 = PL::R[0], RL = PL::R[PL::NP - 1];
  constexpr int BUF = exch_size<N, PL>();
  constexpr bool POW = sizeof(T) == 8;
  __shared__ cpx<T> lds[RPB][BUF];
  const int team = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const long ra = 2 * ((long)blockIdx.x * RPB + team), rb = ra + 1;
  const bool live_a = ra < nrows, live_b = rb < nrows;
  const long la_row = live_a ? ra : nrows - 1, lb_row = live_b ? rb : nrows - 1;
  const int cx = a.kx - 1;
  cpx<T> twc[PL::NP][E];
  fill_twiddles<T, N, PL, false, POW>(twc, lane, a.tw);
  const int f = blockIdx.y;   // the field
  const cpx<T>* __restrict__ ia = a.in[f] + la_row * a.kxp;
  const cpx<T>* __restrict__ ib = a.in[f] + lb_row * a.kxp;
  cpx<T> va[E], vb[E];
#pragma unroll
  for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
    for (int rr = 0; rr < R0; ++rr) {
      const int m = in_pos<N, LANES, R0>(lane, qq, rr);
      const int idx = m <= cx ? m : (m >= N - cx ? N - m : 0);
      va[qq * R0 + rr] = ia[idx];
      vb[qq * R0 + rr] = ib[idx];
    }
  cpx<T> v[E];
#pragma unroll
  for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
    for (int rr = 0; rr < R0; ++rr) {
      const int e = qq * R0 + rr;
      const int m = in_pos<N, LANES, R0>(lane, qq, rr);
      const bool keep = m <= cx || m >= N - cx;
      const bool mirror = m >= N - cx;
      const T z0 = m == 0 ? T(0) : T(1);
      const T ay = (mirror ? -va[e].y : va[e].y) * z0;
      const T by = live_b ? (mirror ? -vb[e].y : vb[e].y) * z0 : T(0);
      const T bx = live_b ? vb[e].x : T(0);
      v[e] = keep ? mk(va[e].x - by, ay + bx) : mk(T(0), T(0));
    }
  run_passes_cached<T, 1, N, PL, false, 1, POW>(v, lane, 0, lds[team], twc);
  T* __restrict__ xa = a.xout[f] + ra * (long)N;
#pragma unroll
  for (int qq = 0; qq < E / RL; ++qq)
#pragma unroll
    for (int rr = 0; rr < RL; ++rr) {
      const int pos = final_pos<N, PL, false>(lane, qq, rr);
      if (live_a) xa[pos] = v[qq * RL + rr].x;
      if (live_b) xa[N + pos] = v[qq * RL + rr].y;
    }
}

// r2c of rows 2t and 2t + 1: z = a + i b, one forward transform, half
// spectra A = (Z_k + conj Z_{N-k}) / 2, B = (Z_k - conj Z_{N-k}) / 2i.  The
// x loads sit at the first-pass positions of the reversed radix order.
template <typename T, int N, int V = 0>
__global__ void __launch_bounds__(64) row_r2c_fast(const RowArgs<T> a, long nrows) {
  using PL = RowPlanV<N, V>;
  constexpr int LANES = PL::LANES, E = PL::E, RPB = 64 / LANES;
  constexpr int RL = PL::R[PL::NP - 1], RF = PL::R[0];
  constexpr int BUF = exch_size<N, PL>();
  constexpr bool POW = sizeof(T) == 8;
  constexpr bool UNI = uniform_radix<PL>();
  __shared__ cpx<T> lds[RPB][BUF];
  const int team = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  const long ra = 2 * ((long)blockIdx.x * RPB + team), rb = ra + 1;
  const bool live_a = ra < nrows, live_b = rb < nrows;
  const long la_row = live_a ? ra : nrows - 1, lb_row = live_b ? rb : nrows - 1;
  cpx<T>* buf = lds[team];
  cpx<T> twf[PL::NP][E];
  fill_twiddles<T, N, PL, true, POW>(twf, lane, a.tw);
  const int f = blockIdx.y;   // the field
  const T* __restrict__ xa = a.xin[f] + la_row * (long)N;
  const T* __restrict__ xb = a.xin[f] + lb_row * (long)N;
  cpx<T> v[E];
#pragma unroll
  for (int qq = 0; qq < E / RL; ++qq)
#pragma unroll
    for (int rr = 0; rr < RL; ++rr) {
      const int pos = final_pos<N, PL, false>(lane, qq, rr);
      const T x0 = xa[pos], x1 = xb[pos];
      v[qq * RL + rr] = mk(x0, live_b ? x1 : T(0));
    }
  run_passes_cached<T, -1, N, PL, true, 1, POW>(v, lane, 0, buf, twf);
#pragma unroll
  for (int qq = 0; qq < E / RF; ++qq)
#pragma unroll
    for (int rr = 0; rr < RF; ++rr) buf[final_pos<N, PL, true>(lane, qq, rr)] = v[qq * RF + rr];
  __syncthreads();
  cpx<T>* __restrict__ oa = a.out[f] + ra * (long)a.kxp;
  for (int k = lane; k < a.kxp; k += LANES) {
    cpx<T> A = mk(T(0), T(0)), B = mk(T(0), T(0));
    if (k < a.kx) {
      const cpx<T> zk = buf[k];
      const cpx<T> zm = conj(buf[(N - k) % N]);
      const cpx<T> s = zk + zm, d = zk - zm;
      A = mk(T(0.5) * s.x, T(0.5) * s.y);
      B = mk(T(0.5) * d.y, T(-0.5) * d.x);
    }
    if (live_a) oa[k] = A;
    if (live_b) oa[a.kxp + k] = B;
  }
  (void)UNI;
}

// ---------------------------------------------------------------------------
// column kernel: W columns per workgroup, LANES threads per column
// ---------------------------------------------------------------------------

// First-pass inputs of one column: the stored rows (compact layouts leave out
// the rows the 2/3 rule removes), or with SPECTRAL the recipe's NT terms
// coef * (i k)^[dim >= 0] * src evaluated on the fly.  Every load is issued
// unconditionally, from a clamped in-range address (a column of the previous
// tile for the lanes past the last column; row 0 for rows left out), and
// unused values are zeroed by a select afterwards: loads under lane-dependent
// or per-term branches each end in a wait for memory, which serialises the
// column's loads.  Wavenumbers: k_x is one table load per lane (the column),
// the slab axis's one per workgroup, and the row axis's one per element,
// needed only by the ROWK variant (a term differentiates along the row axis)
// and shared by both terms.  FP64 passes at n = 128 form the row values from
// the index as the table does, ((ms * (1/n)) * 2) * kny, bit for bit: that
// keeps them at 168 VGPRs, three wavefronts per SIMD (172 with the loads);
// elsewhere the loads, issued with the data, measured faster.
template <typename T, int N, class PL, bool SPECTRAL, int NT, bool ROWK, int W>
__device__ __forceinline__ void col_gather(cpx<T> (&v)[PL::E], const ColArgs<T>& a, int lane, int f,
                                           long slab, int col, bool live) {
  constexpr int LANES = PL::LANES, E = PL::E, R0 = PL::R[0];
  constexpr bool KR_IDX = sizeof(T) == 8 && N == 128;
  // lanes past the last column read the same lane's column of the previous
  // tile: distinct addresses in lines that tile's workgroup has just read
  const int colc = live ? col : (col >= 2 * W ? col - W : a.ncols - 1);
  constexpr int NTA = NT > 0 ? NT : 1;
  const cpx<T>* srcp[NTA];
  bool deriv[NTA], byrow[NTA], byslab[NTA];
  T coef[NTA];
  double kc = 0.0, ks = 0.0;
  if (SPECTRAL) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const Term tm = a.rec[f].t[t];
      srcp[t] = a.src[tm.src];
      deriv[t] = tm.dim >= 0;
      byrow[t] = tm.dim > 0 && tm.dim == a.row_dim;
      byslab[t] = tm.dim > 0 && tm.dim != a.row_dim;
      coef[t] = T(tm.coef);
    }
    kc = a.k_col[colc];
    // passes without a slab table read k_x[0] in its place (never used)
    ks = (a.k_slab ? a.k_slab : a.k_col)[a.k_slab ? slab : 0];
  }
  cpx<T> ld[NTA][E];
  double kr[ROWK ? E : 1];
  bool ok[E];
#pragma unroll
  for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
    for (int rr = 0; rr < R0; ++rr) {
      const int e = qq * R0 + rr;
      const int m = in_pos<N, LANES, R0>(lane, qq, rr);
      int row;
      if (a.in_c < 0) row = m;
      else if (m <= a.in_c) row = m;
      else if (m >= N - a.in_c) row = m - (N - a.in_rows);
      else row = -1;
      ok[e] = live && row >= 0;
      const int rowc = row >= 0 ? row : 0;
      const long off = slab * a.in_ss + (long)rowc * a.in_rs + colc;
      if (SPECTRAL) {
#pragma unroll
        for (int t = 0; t < NT; ++t) ld[t][e] = srcp[t][off];
        if (ROWK) {
          if (KR_IDX) kr[e] = ((double)(m <= a.in_c ? m : m - N) * a.kr_inv) * 2.0 * a.kr_kny;
          else kr[e] = a.k_row[rowc];
        }
      } else {
        ld[0][e] = a.in[f][off];
      }
    }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    cpx<T> x = mk(T(0), T(0));
    if (SPECTRAL) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double kk = (ROWK && byrow[t]) ? kr[ROWK ? e : 0] : (byslab[t] ? ks : kc);
        const cpx<T> sv = deriv[t] ? mul_is(ld[t][e], T(kk)) : ld[t][e];
        x = x + scale(sv, coef[t]);
      }
    } else {
      x = ld[0][e];
    }
    v[e] = ok[e] ? x : mk(T(0), T(0));
  }
}

// Output store of a column pass; nt: as a nontemporal (streaming) store
// (uniform per launch).
template <typename T>
__device__ __forceinline__ void col_store(cpx<T>* p, cpx<T> v, bool nt) {
  st_nt(p, v, nt);
}

template <typename T, int D, int N, bool SPECTRAL, bool WIDE, int V = 0>
__global__ void __launch_bounds__((ColPlanV<N, V>::LANES * col_width<N, WIDE, V>()))
col_fft_fast(const ColArgs<T> a) {
  using PL = ColPlanV<N, V>;
  constexpr int LANES = PL::LANES, E = PL::E, W = col_width<N, WIDE, V>();
  constexpr int R0 = PL::R[0];
  constexpr int RL = PL::R[PL::NP - 1];
  constexpr int BUF = exch_size<N, PL>();
  // FP64 exchanges for n <= 128 go through a buffer of reals in two halves,
  // which lifts the occupancy LDS allowed (8 % faster passes at 128^3); at
  // n = 256 only the 16-column tiles use it (their complex buffer would
  // exceed 64 KB); for 8-column tiles the extra round cost more than the
  // occupancy gained
  constexpr bool SPLIT = sizeof(T) == 8 && (N <= 128 || W == 16);
  __shared__ typename std::conditional<SPLIT, T, cpx<T>>::type lds[BUF * W];
  const int c = threadIdx.x % W, lane = threadIdx.x / W;
  int f, tile;
  long slab;
  if (a.xcd_group == 2) {
    // Workgroups are dealt to the 8 XCDs round robin by linear index L.  XCD
    // x = L mod 8 takes slabs x, x + 8, ..., and walks each slab tile by
    // tile, the F fields of a tile one after another: fields that read the
    // same sources (a spectral prologue) find them in that XCD's L2, and so
    // do the lanes past the last column, which read the previous tile.
    const long L = blockIdx.x, t = L >> 3;
    const long per = (long)a.ntiles * a.nfields;
    slab = (t / per) * 8 + (L & 7);
    if (slab >= a.nslabs) return;   // padding of the last group of 8 slabs
    tile = (int)((t % per) / a.nfields);
    f = (int)(t % a.nfields);
  } else if (a.xcd_group) {
    // L = 8 (floor(U / 8) F + f) + U mod 8 puts the F fields of unit
    // U = (slab, tile) on one XCD, one after another
    const long L = blockIdx.x, t = L >> 3;
    const long U = (t / a.nfields) * 8 + (L & 7);
    if (U >= (long)a.ntiles * a.nslabs) return;   // padding of the last group
    f = (int)(t % a.nfields);
    tile = (int)(U % a.ntiles);
    slab = U / a.ntiles;
  } else {
    f = blockIdx.z;
    tile = blockIdx.x;
    slab = blockIdx.y;
  }
  const int col = tile * W + c;
  const bool live = col < a.ncols;
  // twiddles are loaded before the data (FP64: leg-1 twiddles, powers formed
  // in the butterflies; FP32: every leg), so the second pass does not wait on
  // table loads after the exchange
  constexpr bool POW = sizeof(T) == 8;
  cpx<T> twc[PL::NP][E];
  fill_twiddles<T, N, PL, false, POW>(twc, lane, a.tw);
  cpx<T> v[E];
  if (SPECTRAL) {
    // the term count is uniform: one straight-line gather per count
    const Recipe& rc = a.rec[f];
    const int nterm = rc.nterm;
    const bool rowk = (rc.t[0].dim > 0 && rc.t[0].dim == a.row_dim) ||
                      (nterm >= 2 && rc.t[1].dim > 0 && rc.t[1].dim == a.row_dim);
    if (nterm >= 2) {
      if (rowk) col_gather<T, N, PL, true, 2, true, W>(v, a, lane, f, slab, col, live);
      else col_gather<T, N, PL, true, 2, false, W>(v, a, lane, f, slab, col, live);
    } else {
      if (rowk) col_gather<T, N, PL, true, 1, true, W>(v, a, lane, f, slab, col, live);
      else col_gather<T, N, PL, true, 1, false, W>(v, a, lane, f, slab, col, live);
    }
  } else {
    col_gather<T, N, PL, false, 0, false, W>(v, a, lane, f, slab, col, live);
  }
  if constexpr (SPLIT) run_passes_split<T, D, N, PL, false, W, POW>(v, lane, c, lds, twc);
  else run_passes_cached<T, D, N, PL, false, W, POW>(v, lane, c, lds, twc);
  if (!live) return;
#pragma unroll
  for (int qq = 0; qq < E / RL; ++qq)
#pragma unroll
    for (int rr = 0; rr < RL; ++rr) {
      const int m = final_pos<N, PL, false>(lane, qq, rr);
      int j;
      if (a.out_c < 0) j = m;
      else if (m <= a.out_c) j = m;
      else if (m >= N - a.out_c) j = m - (N - a.out_rows);
      else j = -1;
      if (j >= 0) col_store<T>(&a.out[f][slab * a.out_ss + (long)j * a.out_rs + col], scale(v[qq * RL + rr], a.scale), a.nt);
    }
}

// ---------------------------------------------------------------------------
// forward y pass fused with the k-space epilogue (2D, n_y = 64)
// ---------------------------------------------------------------------------

constexpr int YEPI_W = 8;      // kx columns per workgroup
constexpr int YEPI_WAVES = 4;  // wavefront f transforms product field f
constexpr int YEPI_SPLIT = 2;  // workgroups per column group (blockIdx.z: share of the ky rows)

// YEPI_SPLIT workgroups per YEPI_W kx columns of one member.  Each transforms
// every product field (2 or 3) of the columns: wavefront f runs the y
// transform of field f (the column plan of 64: 8 lanes x 8 columns), so the
// fields are transformed side by side; a wavefront without a field repeats
// field 0 and discards it, which keeps the workgroup barriers uniform.  The
// kept ky modes go to LDS, and then the workgroup runs the epilogue
// (epi_mode) of its share of the ky rows, one mode per thread.  The mode's
// other reads (the sweep's y and F1, the Boussinesq T and buoyancy velocity,
// the wavenumbers) are issued before the transform, so that their latency
// overlaps it; the shares make one mode per thread enough (the transform is
// repeated per share: at 64^2 launch latency and dependent memory round
// trips set the time, not the arithmetic).  This saves the compact
// products' round trip through HBM and, on the 64^2 grids, one launch per
// right-hand side.
// a.U[f]: the x-row products, layout (members, 64, Kxp).
template <typename T>
__global__ void __launch_bounds__(64 * YEPI_WAVES) yepi64_kernel(const EpiArgs<T> a) {
  constexpr int N = 64;
  using PL = ColPlan<N>;
  constexpr int LANES = PL::LANES, E = PL::E, W = YEPI_W;
  static_assert(LANES * W == 64, "one wavefront per field");
  constexpr int R0 = PL::R[0], RL = PL::R[PL::NP - 1];
  constexpr int BUF = exch_size<N, PL>();
  constexpr int KYMAX = 2 * ((N + 2) / 3) + 1;   // bound on the kept rows
  static_assert((KYMAX + YEPI_SPLIT - 1) / YEPI_SPLIT * W <= 64 * YEPI_WAVES, "one mode per thread");
  constexpr bool POW = sizeof(T) == 8;
  __shared__ cpx<T> lds[YEPI_WAVES][BUF * W];
  __shared__ cpx<T> stage[3][KYMAX * W];
  const KGeom& g = a.g;
  const int wave = threadIdx.x / 64, t = threadIdx.x % 64;
  const int c = t % W, lane = t / W;
  const int col = blockIdx.x * W + c;
  const long slab = blockIdx.y;
  const long base = slab * (long)N * g.Kxp + (col < g.Kxp ? col : g.Kxp - 1);
  const int cy = (g.Ky - 1) / 2;
  const bool bouss = a.physics == PH_BOUSS, mhd = a.physics == PH_MHD;
  const int nf = 2 + ((bouss || mhd) ? 1 : 0);
  const int nvar = 2 + (bouss ? 1 : (mhd ? 2 : 0));
  const int f = wave < nf ? wave : 0;
  const cpx<T> z = mk(T(0), T(0));

  // this thread's mode: row j (of this workgroup's share) and column cl
  const int KH = (g.Ky + YEPI_SPLIT - 1) / YEPI_SPLIT;
  const int jj = threadIdx.x / W, cc = threadIdx.x - jj * W;
  const int j = blockIdx.z * KH + jj;
  const int cl = blockIdx.x * W + cc;
  const bool mine = jj < KH && j < g.Ky && cl < g.Kxp;
  const bool live = mine && cl < g.Kx;
  const long idx = ((long)j * g.Kz + slab) * g.Kxp + cl;
  cpx<T> yv[6], fv[6], tval = z, bval = z;
#pragma unroll
  for (int v = 0; v < 6; ++v) yv[v] = fv[v] = z;
  double kx = 0.0, ky = 0.0, kz = 0.0;
  if (live) {
    kx = g.kx[cl];
    ky = g.ky[j];
    kz = g.kz[slab];
    if (bouss) {
      tval = a.state[2][idx];
      bval = (a.bdir == 0 ? a.state[0] : a.state[1])[idx];
    }
    if (a.upd_kind >= 0) {
      const bool hasf = a.upd_kind >= UPD_ETD2RK1;
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (v < nvar) {
          yv[v] = a.upd_y[v][idx];
          if (hasf) fv[v] = a.upd_f1[v][idx];
        }
    }
  }

  cpx<T> twc[PL::NP][E];
  fill_twiddles<T, N, PL, false, POW>(twc, lane, a.tw_y);
  cpx<T> v[E];
  const cpx<T>* __restrict__ src = (f == 0 ? a.U[0] : (f == 1 ? a.U[1] : a.U[2]));
#pragma unroll
  for (int qq = 0; qq < E / R0; ++qq)
#pragma unroll
    for (int rr = 0; rr < R0; ++rr)
      v[qq * R0 + rr] = src[base + (long)in_pos<N, LANES, R0>(lane, qq, rr) * g.Kxp];
  run_passes_cached<T, -1, N, PL, false, W, POW>(v, lane, c, lds[wave], twc);
  if (wave < nf) {
#pragma unroll
    for (int qq = 0; qq < E / RL; ++qq)
#pragma unroll
      for (int rr = 0; rr < RL; ++rr) {
        const int m = final_pos<N, PL, false>(lane, qq, rr);
        const int jm = m <= cy ? m : (m >= N - cy ? m - (N - g.Ky) : -1);
        if (jm >= 0) stage[wave][jm * W + c] = v[qq * RL + rr];
      }
  }
  __syncthreads();
  if (!mine) return;
  if (!live) {
    epi_pad(a, idx, nvar);
    return;
  }
  const int i = j * W + cc;
  const cpx<T> e0 = nf > 2 ? stage[2][i] : z;
  epi_mode(a, idx, kx, ky, kz, stage[0][i], stage[1][i], z, e0, z, z, tval, bval, yv, fv);
}

}  // namespace ddx
