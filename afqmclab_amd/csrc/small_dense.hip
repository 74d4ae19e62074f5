// Per-walker N x N dense kernels: one workgroup per matrix, the matrix register resident (a
// 112 x 112 complex FP64 matrix is 196 KB, more than a CU's LDS), 16 waves, or 8 waves with part
// of the matrix in LDS so that two matrices share a CU.
//
// gj_inverse_logdet: Gauss-Jordan exchange sweep with partial pivoting.  The pivot of column k
//   is the largest |re|+|im| among rows not yet used, first index on ties, as izamax in zgetrf,
//   so the pivots equal LU's u_kk and
//   logdet = sum log|u_kk| + log(prod u_kk/|u_kk| * sign(perm)) reproduces logDeterminant
//   (tensor_2d_common_fun.cpp:75-105).  Writes O^{-1}.
//   Column layout: wave w owns columns w + 16c, lane l rows l + 64r, so a column and its pivot
//   search live in one wave (a DPP reduction, no LDS round trip).  Per step the owner of column
//   k+1 updates that column first, finds its pivot and publishes the multipliers
//   m_i = x_i,k+1 / pivot (0 at the pivot row) through LDS; every wave reads the pivot row of its
//   own columns with readlane.  A row chosen as pivot is not rescaled: it keeps the scale
//   -pivot (the exchange step is linear in each row), which is divided out when the inverse is
//   written.  One barrier per step.  For 129 to 160 rows the third row group (rows 128..159,
//   lanes 0..31) lives in LDS, each wave holding its own columns of it: thirty complex registers
//   per lane would not fit the 128 a 16-wave workgroup allows.  For 65 to 112 rows and more
//   matrices than CUs, 8 waves own 12 or 14 columns each and the last half of them keep their
//   second row group in LDS: up to 128 registers and 68 KB of LDS per matrix let two matrices
//   share a CU, and each hides the other's per-step synchronisation.  Up to 64 rows, 8 or 4 waves per matrix
//   (one row group in registers) let several matrices share a CU.
//   Two more kernels give the same pivots: a 2-D lane layout (both operands of each rank-1 update
//   from LDS) and a kernel in panels of four columns (rank-4 updates, with one wave pivoting the
//   next panel while the others update); the dispatcher picks by size.
// chol_inverse: S = R^H R (upper R, positive real diagonal) and R^{-1} in one sweep, row
//   formulation in one array: after the Jacobi scaling S' = D^-1/2 S D^-1/2, step k leaves
//   row k of the working array unscaled (scale r_kk: R's row in the upper part, the row of
//   A = R'^{-H} in the lower part) and updates the rows below it; the lower part holds the
//   Hermitian Schur complement until its column's step, where adding (S_kk + 1) times the
//   multiplier turns it into A's column (|S_kk| <= 1 after the scaling, so the cancellation
//   costs no accuracy).  Row layout: wave w owns rows w + 16c, lane l columns l + 64r, so a row
//   is published by one wave with coalesced LDS writes and R^{-1} = D^-1/2 A^H is written
//   coalesced.  Returns sum log R_kk (CholeskyQR step of the walker restabilisation).  For 129
//   to 160 columns the third column group (columns 128..159, lanes 0..31) lives in LDS, as in
//   the Gauss-Jordan kernel.
// From 161 to 256 columns both work on a copy of the matrix in device memory (GjWideK, CholWideK).
#include "small_dense.hpp"

#include "group.hpp"

#include <type_traits>
#include <utility>

namespace {

constexpr int NMAX = 160;
constexpr int WAVES = 16;

// f(integral_constant<int, 0>), ..., f(integral_constant<int, E-1>)
template <int E, typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl<E>(f, std::make_integer_sequence<int, E>{});
}

__device__ __forceinline__ double abs1(zd a) { return fabs(a.re) + fabs(a.im); }

// Packed pivot key of a candidate row: the bits of the non-negative score |re|+|im| (monotone as
// an unsigned integer) with the low 8 bits replaced by 255 - row, and bit 63 set.  The maximum
// key is the largest score, ties (to 2^-44 relative) going to the smallest row, as izamax does.
__device__ __forceinline__ unsigned long long pivot_key(zd v, int row) {
  double s = abs1(v);
  if (!(s >= 0.0)) s = 0.0;  // NaN input: still pick a valid unused row (reported singular)
  return ((unsigned long long)__double_as_longlong(s) & ~0xFFull) | (unsigned long long)(255 - row) | (1ull << 63);
}

// x - a*b with two fused multiply-adds per component
__device__ __forceinline__ zd cfms(zd x, zd a, zd b) {
  return cmk<double>(fma(-a.re, b.re, fma(a.im, b.im, x.re)), fma(-a.re, b.im, fma(-a.im, b.re, x.im)));
}

// 1/d for a positive normal d: hardware reciprocal refined by two Newton steps
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ zd readlane_z(zd v, int lane) { return cmk<double>(readlane_d(v.re, lane), readlane_d(v.im, lane)); }
// a wave-uniform value (an LDS broadcast read) moved to scalar registers
__device__ __forceinline__ double readfirstlane_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ zd readfirstlane_z(zd v) { return cmk<double>(readfirstlane_d(v.re), readfirstlane_d(v.im)); }

// one DPP step of a wave maximum: v = max(v, v of the source lane); lanes without a source (or
// in rows outside ROW_MASK) keep their value
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_max(unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
  const unsigned nlo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
  const unsigned nhi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
  const unsigned long long w = ((unsigned long long)nhi << 32) | nlo;
  return w > v ? w : v;
}

// maximum over the 64 lanes, returned uniform: row_shr 1, 2, 4, 8 leave each row's maximum in
// its lane 15; row_bcast 15 and 31 carry it into lane 63
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  v = dpp_max<0x111, 0xf>(v);
  v = dpp_max<0x112, 0xf>(v);
  v = dpp_max<0x114, 0xf>(v);
  v = dpp_max<0x118, 0xf>(v);
  v = dpp_max<0x142, 0xa>(v);
  v = dpp_max<0x143, 0xc>(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Gauss-Jordan, column layout: x[r][c] = element (lane + 64r, wave + WV c); NR * 64 >= N and
// NC * WV >= N.  The last row group keeps its first NL register columns in registers and the
// others in LDS (gl[wave][c - NL][lane], lanes below LL): all of them for 129..160 rows, the
// second half for the 8-wave kernel of 97..112 rows, none otherwise.
template <typename T, int NR, int NC, int WV, int NL>
struct GjK {
  static constexpr int NT = 64 * WV, WPE = 4;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
template <typename T, int NR, int NC, int WV, int NL>
__device__ __forceinline__ void GjK<T, NR, NC, WV, NL>::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  const int bidx = bk.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  typedef cplx<T> ct;
  const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;

  __shared__ zd colbuf[2][64 * NR];
  __shared__ int pivrow[2];
  __shared__ zd pivs[NMAX];
  __shared__ zd rowscale[NMAX];
  __shared__ int perm[NMAX];
  __shared__ int qinv[NMAX];
  __shared__ int singular, parity;

  constexpr int RL = NR - 1;                   // the last row group
  constexpr bool LG = NL < NC;                 // part of it in LDS
  constexpr int LL = 64 * NR <= NMAX ? 64 : NMAX - 64 * RL;  // its lanes that can hold a row
  static_assert(LL > 0 && LL <= 64, "row groups");
  __shared__ zd gl[LG ? WV : 1][LG ? NC - NL : 1][LG ? LL : 1];
  const bool llane = lane < LL;
  // whether element (r, c) is a register (folded at compile time once the loops are unrolled)
  auto reg = [](int r, int c) { return r < RL || c < NL; };
  auto lds = [&](int c, int l) -> zd& { return gl[LG ? wv : 0][LG && c >= NL ? c - NL : 0][LG ? l : 0]; };
  zd x[NR][NC];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = lane + 64 * r, j = wv + WV * c;
      const zd v = (i < N && j < N) ? to_zd(in[i + (long long)j * p.ld_in]) : cmk<double>(0, 0);
      if (reg(r, c)) x[r][c] = v;
      else if (llane) lds(c, lane) = v;
    }
  // element (r, c) -= m * rv
  auto fms = [&](int r, int c, zd m, zd rv) {
    if (reg(r, c)) x[r][c] = cfms(x[r][c], m, rv);
    else if (llane) lds(c, lane) = cfms(lds(c, lane), m, rv);
  };
  auto get = [&](int r, int c) -> zd {
    if (reg(r, c)) return x[r][c];
    return llane ? lds(c, lane) : cmk<double>(0, 0);
  };
  // element (row of lane pl in group r, c), wave-uniform, in scalar registers
  auto row_value = [&](int r, int c, int pl) -> zd {
    if (reg(r, c)) return readlane_z(x[r][c], pl);
    return readfirstlane_z(lds(c, pl));
  };
  unsigned used = 0u;  // bit r: row lane + 64r has been a pivot row
  if (threadIdx.x == 0) { singular = 0; parity = 0; }

  // The owner of column `col` (register column C1 of this wave, already updated through step
  // col-1) picks its pivot among the unused rows and publishes the multipliers and the pivot.
  auto publish = [&](auto c1c, int col, int buf) {
    constexpr int C1 = decltype(c1c)::value;
    unsigned long long key = 0ull;
    zd v[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      v[r] = get(r, C1);
      const int i = lane + 64 * r;
      if (i < N && !((used >> r) & 1u)) {
        const unsigned long long kk = pivot_key(v[r], i);
        key = kk > key ? kk : key;
      }
    }
    key = wave_max_u64(key);
    // bit 63 marks a candidate: at step col there are N - col >= 1 unused rows
    const int bi = 255 - (int)(key & 0xFFull);
    const int pr = (bi >= 0 && bi < N) ? bi : 0;
    const int rr = pr >> 6, pl = pr & 63;
    zd piv = cmk<double>(0, 0);
    static_for<NR>([&](auto rc) {
      constexpr int R = decltype(rc)::value;
      if (rr == R) piv = row_value(R, C1, pl);
    });
    const bool nz = abs1(piv) > 0.0;
    const double id = nz ? recip(piv.re * piv.re + piv.im * piv.im) : 0.0;
    const zd inv = cmk<double>(piv.re * id, -piv.im * id);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lane + 64 * r;
      colbuf[buf][i] = (i == pr) ? cmk<double>(0, 0) : cmul(v[r], inv);
    }
    if (lane == 0) {
      pivrow[buf] = pr;
      pivs[col] = piv;
      perm[col] = pr;
      rowscale[pr] = cmk<double>(-inv.re, -inv.im);
      if (!nz) singular = 1;
    }
  };

  if (wv == 0) publish(std::integral_constant<int, 0>{}, 0, 0);
  __syncthreads();

  // Step k = CB*WV + t: column k is register column CB of wave t, column k+1 register column CB
  // of wave t+1 (or CB+1 of wave 0), so no register array is indexed by a run-time value.
  auto step = [&](auto cbc, int t) {
    constexpr int CB = decltype(cbc)::value;
    const int k = CB * WV + t;
    const int buf = k & 1, nbuf = buf ^ 1;
    const int pr = __builtin_amdgcn_readfirstlane(pivrow[buf]);
    const int rr = pr >> 6, pl = pr & 63;
    zd m[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) m[r] = colbuf[buf][lane + 64 * r];
    if (lane == pl) used |= 1u << rr;
    // the pivot row of this wave's columns (uniform)
    zd rv[NC];
    static_for<NR>([&](auto rc) {
      constexpr int R = decltype(rc)::value;
      if (rr == R) {
#pragma unroll
        for (int c = 0; c < NC; ++c) rv[c] = row_value(R, c, pl);
      }
    });
    const bool next_here = (k + 1 < N) && (wv == (t + 1) % WV);
    const bool next_in_cb = (t + 1 < WV);
    // look-ahead: column k+1 first, then its pivot and multipliers
    if (next_here) {
      if (next_in_cb) {
#pragma unroll
        for (int r = 0; r < NR; ++r) fms(r, CB, m[r], rv[CB]);
        publish(std::integral_constant<int, CB>{}, k + 1, nbuf);
      } else {
        if constexpr (CB + 1 < NC) {
#pragma unroll
          for (int r = 0; r < NR; ++r) fms(r, CB + 1, m[r], rv[CB + 1]);
          publish(std::integral_constant<int, CB + 1>{}, k + 1, nbuf);
        }
      }
    }
    const bool owner = (wv == t);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const bool done = next_here && ((next_in_cb && c == CB) || (!next_in_cb && c == CB + 1));
      if (done || (owner && c == CB)) continue;
#pragma unroll
      for (int r = 0; r < NR; ++r) fms(r, c, m[r], rv[c]);
    }
    // column k becomes the multipliers, -1 at the (unscaled) pivot row
    if (owner) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const zd v = (lane + 64 * r == pr) ? cmk<double>(-1.0, 0.0) : m[r];
        if (reg(r, CB)) x[r][CB] = v;
        else if (llane) lds(CB, lane) = v;
      }
    }
    __syncthreads();
  };
  auto block = [&](auto cbc) {
    constexpr int CB = decltype(cbc)::value;
    for (int t = 0; t < WV && CB * WV + t < N; ++t) step(cbc, t);
  };
  static_for<NC>(block);

  if ((int)threadIdx.x < N) qinv[perm[threadIdx.x]] = threadIdx.x;
  // sign of the row permutation: parity of its inversions, counted in parallel
  {
    int odd = 0;
    for (int e = threadIdx.x; e < N * N; e += 64 * WV) {
      const int i = e / N, j = e - i * N;
      odd ^= (i < j && perm[i] > perm[j]) ? 1 : 0;
    }
    if (odd) atomicXor(&parity, 1);
  }
  // log|det| and the phase of det from the pivots: wave 0, fixed-order trees (deterministic)
  double lognorm = 0.0;
  zd phase = cmk<double>(1.0, 0.0);
  if (wv == 0) {
#pragma unroll
    for (int q = 0; q < (NMAX + 63) / 64; ++q) {
      const int k = lane + 64 * q;
      if (k < N) {
        const zd pv = pivs[k];
        const double ap = cabs(pv);
        const double ia = ap > 0.0 ? 1.0 / ap : 0.0;
        lognorm += log(ap);
        phase = cmul(phase, cmk<double>(pv.re * ia, pv.im * ia));
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lognorm += __shfl_xor(lognorm, o, 64);
      const zd q = cmk<double>(__shfl_xor(phase.re, o, 64), __shfl_xor(phase.im, o, 64));
      phase = cmul(phase, q);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const zd ph = parity ? cmk<double>(-phase.re, -phase.im) : phase;
    zd ld;
    if (singular) {
      ld = cmk<double>(-1e300, 0.0);
    } else {
      ld = cmk<double>(lognorm, atan2(ph.im, ph.re));
      ld.re += log(hypot(ph.re, ph.im));
    }
    p.logdet[bidx] = ld;
    if (p.info) p.info[bidx] = singular;
  }
  if (!p.write_inverse) return;
  // inverse: O^{-1}[qinv[i]][perm[j]] = x[i][j] / (-pivot of row i)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = lane + 64 * r;
    if (i >= N) continue;
    const int row = qinv[i];
    const zd sc = rowscale[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int j = wv + WV * c;
      if (j >= N) continue;
      out[row + (long long)perm[j] * p.ld_out] = from_zd<T>(cmul(get(r, c), sc));
    }
  }
}

// Gauss-Jordan, 2-D lane layout: lane = lc LR + lr of wave w holds rows lr + LR r (r < NR) of the
// columns w + WV (lc + LC c) (c < NC, LC = 64 / LR).  A column lives in the LR lanes of one column
// group of one wave, so its pivot search is still one wave reduction (the other lanes offer key
// 0).  Both operands of the rank-1 update come from LDS: the multipliers of a lane's rows, which the
// column's owner publishes, and the pivot row's values in the lane's columns, which the lanes
// holding the pivot row write and their wave reads back.  Per step and lane that is NR NC complex
// multiply-adds against NR + NC LDS reads and no readlanes (the 1-D layout, lanes = rows, reads NC
// pivot-row values by four readlanes each).  Pivots, the update order of every element and the
// deferred row scaling are those of the 1-D kernel, so the results agree with it bit for bit.
template <typename T, int WV, int LR, int NR, int NC, int WPEV>
struct Gj2K {
  static constexpr int NT = 64 * WV, WPE = WPEV, WPE_MAX = WPEV;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
template <typename T, int WV, int LR, int NR, int NC, int WPEV>
__device__ __forceinline__ void Gj2K<T, WV, LR, NR, NC, WPEV>::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  constexpr int LC = 64 / LR;
  constexpr int ROWS = LR * NR;
  constexpr int COLS = WV * LC * NC;
  constexpr int PER = WV * LC;  // columns per register column index
  static_assert(LR * LC == 64 && WV >= 2 && ROWS <= 256, "2-D layout");
  const int bidx = bk.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int lane = threadIdx.x & 63;
  const int lr = lane % LR, lc = lane / LR;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  typedef cplx<T> ct;
  const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;

  __shared__ zd colbuf[2][ROWS];
  __shared__ zd prow[WV][LC][NR][NC];
  __shared__ int pivrow[2];
  __shared__ zd pivs[COLS];
  __shared__ zd rowscale[ROWS];
  __shared__ int perm[COLS];
  __shared__ int qinv[ROWS];
  __shared__ int singular, parity;

  zd x[NR][NC];
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = lr + LR * r, j = wv + WV * (lc + LC * c);
      x[r][c] = (i < N && j < N) ? to_zd(in[i + (long long)j * p.ld_in]) : cmk<double>(0, 0);
    }
  unsigned used = 0u;  // bit r: row lr + LR r has been a pivot row
  if (threadIdx.x == 0) { singular = 0; parity = 0; }

  // The owner of column `col` (register column C1, column group lcg of this wave, updated through
  // step col-1) picks its pivot among the unused rows and publishes the multipliers and the pivot.
  auto publish = [&](auto c1c, int lcg, int col, int buf) __attribute__((always_inline)) {
    constexpr int C1 = decltype(c1c)::value;
    unsigned long long key = 0ull;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = lr + LR * r;
      if (lc == lcg && i < N && !((used >> r) & 1u)) {
        const unsigned long long kk = pivot_key(x[r][C1], i);
        key = kk > key ? kk : key;
      }
    }
    key = wave_max_u64(key);
    const int bi = 255 - (int)(key & 0xFFull);
    const int pr = (bi >= 0 && bi < N) ? bi : 0;
    const int rr = pr / LR, pl = lcg * LR + pr % LR;
    // every row register read by its own readlane and the result chosen among those values: a
    // choice among the registers themselves (x[rr]) is compiled into a read through a run-time
    // pointer, which moves x out of registers
    zd piv = readlane_z(x[0][C1], pl);
    static_for<NR>([&](auto rc) __attribute__((always_inline)) {
      constexpr int R = decltype(rc)::value;
      if constexpr (R > 0) {
        const zd v = readlane_z(x[R][C1], pl);
        piv.re = (rr == R) ? v.re : piv.re;  // componentwise: a select of the struct copies it through memory
        piv.im = (rr == R) ? v.im : piv.im;
      }
    });
    const bool nz = abs1(piv) > 0.0;
    const double id = nz ? recip(piv.re * piv.re + piv.im * piv.im) : 0.0;
    const zd inv = cmk<double>(piv.re * id, -piv.im * id);
    if (lc == lcg) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int i = lr + LR * r;
        colbuf[buf][i] = (i == pr) ? cmk<double>(0, 0) : cmul(x[r][C1], inv);
      }
    }
    if (lane == 0) {
      pivrow[buf] = pr;
      pivs[col] = piv;
      perm[col] = pr;
      rowscale[pr] = cmk<double>(-inv.re, -inv.im);
      if (!nz) singular = 1;
    }
  };

  if (wv == 0) publish(std::integral_constant<int, 0>{}, 0, 0, 0);
  __syncthreads();

  // Step k = CB PER + t: column k is register column CB of column group t / WV of wave t % WV;
  // column k+1 is register column CB of group (t+1) / WV of wave (t+1) % WV, or register column
  // CB+1 of group 0 of wave 0 when t is the block's last.
  auto step = [&](auto cbc, int t) __attribute__((always_inline)) {
    constexpr int CB = decltype(cbc)::value;
    const int k = CB * PER + t;
    const int buf = k & 1, nbuf = buf ^ 1;
    const int pr = __builtin_amdgcn_readfirstlane(pivrow[buf]);
    const int rr = pr / LR, plr = pr % LR;
    zd m[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) m[r] = colbuf[buf][lr + LR * r];
    // the pivot row's values in this wave's columns, one lane per column group
    // (all NR rows of the pivot row's lanes go to LDS and the readers pick row rr there, for the
    // reason given in publish)
    if (lr == plr) {
#pragma unroll
      for (int R = 0; R < NR; ++R)
#pragma unroll
        for (int c = 0; c < NC; ++c) prow[wv][lc][R][c] = x[R][c];
      used |= 1u << rr;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    zd rv[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) rv[c] = prow[wv][lc][rr][c];
    const int kw = t % WV, kl = t / WV;  // owner of column k
    const bool next_in_cb = t + 1 < PER;
    const int tn = next_in_cb ? t + 1 : 0;
    const bool next_here = (k + 1 < N) && (wv == tn % WV);
    // look-ahead: the register column holding column k+1 first, then its pivot and multipliers
    if (next_here) {
      if (next_in_cb) {
#pragma unroll
        for (int r = 0; r < NR; ++r) x[r][CB] = cfms(x[r][CB], m[r], rv[CB]);
        publish(std::integral_constant<int, CB>{}, tn / WV, k + 1, nbuf);
      } else {
        if constexpr (CB + 1 < NC) {
#pragma unroll
          for (int r = 0; r < NR; ++r) x[r][CB + 1] = cfms(x[r][CB + 1], m[r], rv[CB + 1]);
          publish(std::integral_constant<int, CB + 1>{}, 0, k + 1, nbuf);
        }
      }
    }
    const bool owner = (wv == kw);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const bool done = next_here && ((next_in_cb && c == CB) || (!next_in_cb && c == CB + 1));
      if (done) continue;
      // column k itself (group kl of the owner's register column CB) becomes the multipliers
      const bool keep = owner && c == CB && lc == kl;
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if (!keep) x[r][c] = cfms(x[r][c], m[r], rv[c]);
    }
    if (owner && lc == kl) {
#pragma unroll
      for (int r = 0; r < NR; ++r) x[r][CB] = (lr + LR * r == pr) ? cmk<double>(-1.0, 0.0) : m[r];
    }
    __syncthreads();
  };
  auto block = [&](auto cbc) __attribute__((always_inline)) {
    constexpr int CB = decltype(cbc)::value;
    for (int t = 0; t < PER && CB * PER + t < N; ++t) step(cbc, t);
  };
  static_for<NC>(block);

  for (int i = threadIdx.x; i < N; i += 64 * WV) qinv[perm[i]] = i;
  // sign of the row permutation: parity of its inversions, counted in parallel
  {
    int odd = 0;
    for (int e = threadIdx.x; e < N * N; e += 64 * WV) {
      const int i = e / N, j = e - i * N;
      odd ^= (i < j && perm[i] > perm[j]) ? 1 : 0;
    }
    if (odd) atomicXor(&parity, 1);
  }
  // log|det| and the phase of det from the pivots: wave 0, fixed-order trees (the 1-D kernel's)
  double lognorm = 0.0;
  zd phase = cmk<double>(1.0, 0.0);
  if (wv == 0) {
#pragma unroll
    for (int q = 0; q < (NMAX + 63) / 64; ++q) {
      const int k = lane + 64 * q;
      if (k < N) {
        const zd pv = pivs[k];
        const double ap = cabs(pv);
        const double ia = ap > 0.0 ? 1.0 / ap : 0.0;
        lognorm += log(ap);
        phase = cmul(phase, cmk<double>(pv.re * ia, pv.im * ia));
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lognorm += __shfl_xor(lognorm, o, 64);
      const zd q = cmk<double>(__shfl_xor(phase.re, o, 64), __shfl_xor(phase.im, o, 64));
      phase = cmul(phase, q);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const zd ph = parity ? cmk<double>(-phase.re, -phase.im) : phase;
    zd ld;
    if (singular) {
      ld = cmk<double>(-1e300, 0.0);
    } else {
      ld = cmk<double>(lognorm, atan2(ph.im, ph.re));
      ld.re += log(hypot(ph.re, ph.im));
    }
    p.logdet[bidx] = ld;
    if (p.info) p.info[bidx] = singular;
  }
  if (!p.write_inverse) return;
  // inverse: O^{-1}[qinv[i]][perm[j]] = x[i][j] / (-pivot of row i)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int i = lr + LR * r;
    if (i >= N) continue;
    const int row = qinv[i];
    const zd sc = rowscale[i];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int j = wv + WV * (lc + LC * c);
      if (j >= N) continue;
      out[row + (long long)perm[j] * p.ld_out] = from_zd<T>(cmul(x[r][c], sc));
    }
  }
}

// Gauss-Jordan in panels of four columns.  The exchange steps of one panel compose to a left
// factor L that differs from the identity only in the columns of the panel's pivot rows P, so every
// column outside the panel becomes A - M A[P, :] with M = (I - L)[:, P] = S + I[:, P], S being the
// panel's own columns after its four in-place steps: a panel costs one rank-4 update and one
// barrier instead of four rank-1 steps with a barrier each.
// One wave (w = WV) does all the pivoting: it factorises panel j+1 while the others apply panel j.
// The update waves hand it each panel's columns one phase early (current through panel j-1), it
// applies panel j to them itself, and its serial pivot chain (lanes = rows, NRF row groups, so the
// search is a wave reduction) overlaps the bulk of the update.  Update wave w holds the CPW = 4 NC
// columns CPW w .. CPW w + CPW-1 of all rows, in a 2-D layout: lane lr + 16 lc holds rows lr + 16r
// (r < NR) of columns CPW w + lc + 4c (c < NC).  The pivot rows of its own columns are then in its
// own lanes, which write them to the wave's LDS row buffer for the others to read back, and panel
// b is register column b % NC of wave b / NC.  Per panel and lane: 4 NR NC complex multiply-adds
// against 4 (NR + NC) LDS reads of M and of the pivot rows.
// Pivots follow the rule of the unblocked kernels (largest |re|+|im| among unused rows, first index
// on ties) and the row scaling is deferred the same way; results agree with them to rounding (the
// rank-4 updates sum in another order).
// A workgroup factorises matrices bk.x, bk.x + bk.gx, ...  With two row groups (65 to 112 columns,
// one workgroup per CU) the launch has at most one workgroup per CU, and where the matrix fits
// LDS beside the kernel's own buffers the pivoting wave copies the workgroup's next matrix into a
// staging buffer (direct global-to-LDS loads, 64 x 16 bytes per instruction, a few per panel)
// while the current one is factorised: the next matrix's load is then an LDS read, and no
// workgroup waits for another's stores to drain before it starts.
// Workgroup barrier for LDS data only: waits for the wave's own LDS accesses, not for its
// global loads and stores (nor for loads into LDS in flight, which a workgroup-scope fence would
// wait for); the memory clobber keeps the compiler from moving accesses across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// matrices per workgroup at most (the launch has at least batch / GJB_MAXM workgroups), and the
// flags of a matrix's exponent slot: skipped (not alive), singular
constexpr int GJB_MAXM = 32, GJB_SKIPPED = INT_MIN, GJB_SINGULAR = INT_MIN + 1;
template <typename T, int NR, int NC, int WV, int NRF, int WPEV>
struct GjBK {
  static constexpr int NT = 64 * (WV + 1), WPE = WPEV, WPE_MAX = WPEV;
  static constexpr int NMAX_B = 16 * NR;
  // (FP32 storage only: with FP64 storage at 65 to 80 columns, the one size that fits, 1024
  // inversions of 72 x 72 took 0.309 ms against 0.295 without staging)
  static constexpr bool STAGE = NRF == 2 && sizeof(T) == 4 && sizeof(cplx<T>) * NMAX_B * NMAX_B <= 104 * 1024;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
template <typename T, int NR, int NC, int WV, int NRF, int WPEV>
__device__ __forceinline__ void GjBK<T, NR, NC, WV, NRF, WPEV>::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  constexpr int NB = 4, CPW = 4 * NC;  // panel size, columns per update wave
  constexpr int COLS = CPW * WV, FROWS = 64 * NRF;
  static_assert(16 * NR <= FROWS && NR * NC >= NRF * NB, "row capacity");
  const int N = p.N;
  const int NP = (N + NB - 1) / NB;
  // thread, lane and the lane's row and column in the 2-D layout: set again, opaque to the compiler,
  // for every matrix, so that no address derived from them is hoisted out of the matrix loop into
  // registers held across it
  int tid = threadIdx.x, lane = tid & 63, lr = lane & 15, lc = lane >> 4;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  typedef cplx<T> ct;

  // staging buffer: the next matrix, columns of N elements back to back, as 16-byte chunks
  // (CPC per column); usable when every column starts on a 16-byte boundary
  constexpr int CE = 16 / (int)sizeof(ct);  // elements per chunk
  const bool st = STAGE && (N * (int)sizeof(ct)) % 16 == 0 && (p.ld_in * (long long)sizeof(ct)) % 16 == 0 &&
                  (p.stride_in * (long long)sizeof(ct)) % 16 == 0 && (reinterpret_cast<uintptr_t>(p.in) & 15) == 0;
  const int CPC = N / CE;                       // chunks per column
  const int NI = (CPC * N + 63) / 64;           // load instructions per matrix
  const int IPP = NP > 1 ? (NI + NP - 2) / (NP - 1) : NI;  // per panel phase of the pivoting wave
  // the chunk of this lane in the first instruction: column q0j, chunk q0t of it
  const int q0j = st ? lane / CPC : 0, q0t = st ? lane - q0j * CPC : 0;
  bool staged = false;  // the staging buffer holds matrix bidx

  // the workgroup's LDS as one block, the staging buffer last: the small buffers stay within
  // reach of the 16-bit offsets of LDS instructions (16-byte aligned, for 16-byte LDS accesses)
  struct alignas(16) Lds {
    zd pc[2][NB][FROWS];  // panel columns handed to the pivoting wave
    zd mm[2][NB][FROWS];  // multipliers M of a factorised panel
    zd rb[WV][NB][CPW];   // each update wave's pivot rows of its own columns
    zd fr[NRF][NB];       // the pivoting wave's pivot row
    zd pivs[COLS];
    zd rowscale[FROWS];
    zd dz[GJB_MAXM];      // per matrix of the workgroup: det = dz 2^de (de: or a flag)
    int de[GJB_MAXM];
    int pp[2][NB];        // pivot rows of a factorised panel
    int perm[COLS];
    int qinv[FROWS];
    int singular, parity;
    alignas(16) ct stage[STAGE ? NMAX_B * NMAX_B : 1];
  };
  __shared__ Lds sm;
  static_assert(!STAGE || sizeof(ct) * 4 * NMAX_B * WV <= sizeof(sm.pc) + sizeof(sm.mm), "output buffer");
  auto& pc = sm.pc;
  auto& mm = sm.mm;
  auto& rb = sm.rb;
  auto& fr = sm.fr;
  auto& pivs = sm.pivs;
  auto& rowscale = sm.rowscale;
  auto& pp = sm.pp;
  auto& perm = sm.perm;
  auto& qinv = sm.qinv;
  auto& stage = sm.stage;
  int& singular = sm.singular;
  int& parity = sm.parity;

  const bool upd = wv < WV;
  // update waves: x[r][c] = element (lr + 16r, CPW wv + lc + 4c); panel b is register column
  // b % NC of wave b / NC (lane column lc holding its column lc)
  zd x[NR][NC];
  // A register column chosen at run time is reached through independent branches on the bits of
  // a mask the compiler cannot see through: branches on b % 4 == C would be recognised as
  // exclusive and merged into one with the register read through a run-time pointer.
  auto col_mask = [](int b) __attribute__((always_inline)) {
    unsigned m = (unsigned)__builtin_amdgcn_readfirstlane(1 << (b % NC));
    asm volatile("" : "+s"(m));
    return m;
  };
  auto hand_over = [&](int b, int buf) __attribute__((always_inline)) {
    if (b / NC != wv) return;
    const unsigned sel = col_mask(b);
    static_for<NC>([&](auto cc) __attribute__((always_inline)) {
      constexpr int C = decltype(cc)::value;
      if (sel & (1u << C)) {
#pragma unroll
        for (int r = 0; r < NR; ++r) pc[buf][lc][lr + 16 * r] = x[r][C];
      }
    });
  };

  // pivoting wave: the multipliers of the last panel it factorised are kept in x (slot r NB + q),
  // its pivot rows in pk
  auto fm = [&](auto sc) -> zd& {
    constexpr int S = decltype(sc)::value;
    return x[S / NC][S % NC];
  };
  int pk[NB];
  unsigned used = 0u;  // bit r: row lane + 64r has been a pivot row

  // Factorise panel jn from its handed-over copy, after applying panel jn-1 to it when pre.
  auto factor = [&](int jn, bool pre) __attribute__((always_inline)) {
    const int buf = jn & 1;
    zd y[NRF][NB];
#pragma unroll
    for (int r = 0; r < NRF; ++r)
#pragma unroll
      for (int q = 0; q < NB; ++q) y[r][q] = pc[buf][q][lane + 64 * r];
    if (pre) {
      // y -= M_{jn-1} y[P_{jn-1}, :]; the pivot rows come from the unmodified copy in LDS
      static_for<NB>([&](auto kc) __attribute__((always_inline)) {
        constexpr int K = decltype(kc)::value;
        zd rv[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) rv[q] = pc[buf][q][pk[K]];
        static_for<NRF>([&](auto rc) __attribute__((always_inline)) {
          constexpr int R = decltype(rc)::value;
          const zd mk = fm(std::integral_constant<int, R * NB + K>{});
#pragma unroll
          for (int q = 0; q < NB; ++q) y[R][q] = cfms(y[R][q], mk, rv[q]);
        });
      });
    }
    int pkn[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int col = NB * jn + q;
      pkn[q] = 0;
      if (col >= N) continue;
      unsigned long long key = 0ull;
#pragma unroll
      for (int r = 0; r < NRF; ++r) {
        const int i = lane + 64 * r;
        if (i < N && !((used >> r) & 1u)) {
          const unsigned long long kk = pivot_key(y[r][q], i);
          key = kk > key ? kk : key;
        }
      }
      key = wave_max_u64(key);
      const int bi = 255 - (int)(key & 0xFFull);
      const int pr = (bi >= 0 && bi < N) ? bi : 0;
      const int rr = pr >> 6, pl = pr & 63;
      // the pivot row through LDS: its lane writes all its row groups and every lane reads group
      // rr back (a choice among the registers themselves, y[rr], would be compiled into a read
      // through a run-time pointer and move y out of registers)
      if (lane == pl) {
#pragma unroll
        for (int r = 0; r < NRF; ++r)
#pragma unroll
          for (int q2 = 0; q2 < NB; ++q2) fr[r][q2] = y[r][q2];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      zd prv[NB];
#pragma unroll
      for (int q2 = 0; q2 < NB; ++q2) prv[q2] = fr[rr][q2];
      const zd piv = prv[q];
      const bool nz = abs1(piv) > 0.0;
      const double id = nz ? recip(piv.re * piv.re + piv.im * piv.im) : 0.0;
      const zd inv = cmk<double>(piv.re * id, -piv.im * id);
      zd m[NRF];
#pragma unroll
      for (int r = 0; r < NRF; ++r) m[r] = (lane + 64 * r == pr) ? cmk<double>(0, 0) : cmul(y[r][q], inv);
#pragma unroll
      for (int q2 = 0; q2 < NB; ++q2) {
        if (q2 == q) continue;
#pragma unroll
        for (int r = 0; r < NRF; ++r) y[r][q2] = cfms(y[r][q2], m[r], prv[q2]);
      }
#pragma unroll
      for (int r = 0; r < NRF; ++r) y[r][q] = (lane + 64 * r == pr) ? cmk<double>(-1.0, 0.0) : m[r];
      if (lane == pl) used |= 1u << rr;
      if (lane == 0) {
        pivs[col] = piv;
        perm[col] = pr;
        rowscale[pr] = cmk<double>(-inv.re, -inv.im);
        if (!nz) singular = 1;
      }
      pkn[q] = pr;
      // the next step's pivot-row write must not overtake this step's reads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // M = S + I[:, P], published and kept for the next panel (zero columns past N)
    static_for<NB>([&](auto qc) __attribute__((always_inline)) {
      constexpr int Q = decltype(qc)::value;
      const bool live = NB * jn + Q < N;
      static_for<NRF>([&](auto rc) __attribute__((always_inline)) {
        constexpr int R = decltype(rc)::value;
        const int i = lane + 64 * R;
        zd v = live ? y[R][Q] : cmk<double>(0, 0);
        if (live && i == pkn[Q]) v.re += 1.0;
        fm(std::integral_constant<int, R * NB + Q>{}) = v;
        mm[buf][Q][i] = v;
      });
      pk[Q] = pkn[Q];
      if (lane == 0) pp[buf][Q] = pkn[Q];
    });
  };

  // one staging load of matrix nb: chunks 64 k + lane, the lane's being chunk qt of column qj
  auto stage_load = [&](int nb, int k, int qj, int qt) __attribute__((always_inline)) {
    if (64 * k + lane < CPC * N) {
      const ct* src = reinterpret_cast<const ct*>(p.in) + (long long)nb * p.stride_in + (long long)qj * p.ld_in + CE * qt;
      __builtin_amdgcn_global_load_lds(src, stage + 64 * CE * k, 16, 0, 0);
    }
  };

  int m = 0;  // the matrix's place in the workgroup's sequence; after the loop, their number
  for (int bidx = bk.x; bidx < p.batch; bidx += bk.gx) {
    if (p.alive && !p.alive[bidx / p.alive_div]) {
      if (threadIdx.x == 0) sm.de[m] = GJB_SKIPPED;
      staged = false;
      ++m;
      if (!STAGE) break;  // (without staging the launch has one workgroup per matrix)
      continue;
    }
    const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
    ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;
    const int nb = bidx + bk.gx;
    const bool stage_next = st && NP > 1 && nb < p.batch;
    tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    lane = tid & 63;
    lr = lane & 15;
    lc = lane >> 4;
    // the staging loads of this matrix (issued by the pivoting wave) have landed: vmcnt(0)
    if (!upd && staged) __builtin_amdgcn_s_waitcnt(0x0F70);
    // (also: every wave is done with the previous matrix's buffers)
    lds_barrier();
    if (tid == 0) { singular = 0; parity = 0; }
    used = 0u;
#pragma unroll
    for (int q = 0; q < NB; ++q) pk[q] = 0;
    if (upd) {
      // every load issued unconditionally (element 0 in place of the padding), so that they are
      // all in flight together
      if (staged) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const int i = lr + 16 * r, j = CPW * wv + lc + 4 * c;
            const bool ok = i < N && j < N;
            const zd v = to_zd(stage[ok ? i + j * N : 0]);
            x[r][c] = ok ? v : cmk<double>(0, 0);
          }
      } else {
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const int i = lr + 16 * r, j = CPW * wv + lc + 4 * c;
            const bool ok = i < N && j < N;
            const zd v = to_zd(in[ok ? i + (long long)j * p.ld_in : 0]);
            x[r][c] = ok ? v : cmk<double>(0, 0);
          }
      }
      hand_over(0, 0);
      if (NP > 1) hand_over(1, 1);
    }

    // the pivoting wave's chain is the critical path: its instructions go first on its SIMD
    if (!upd) __builtin_amdgcn_s_setprio(2);
    lds_barrier();
    if (!upd) factor(0, false);
    lds_barrier();
    int sk = 0, sqj = q0j, sqt = q0t;  // the pivoting wave's next staging load
    for (int j = 0; j < NP; ++j) {
      const int buf = j & 1;
      if (upd) {
        int P[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) P[k] = __builtin_amdgcn_readfirstlane(pp[buf][k]);
        // the pivot rows of this wave's columns, written by the lanes holding them: one
        // independent store per row register, the lane condition naming the register (branches
        // that differ only in the register would be merged into a read through a run-time pointer)
        // (also for the padding columns of the last panel, whose pivot row 0 keeps the buffer
        // finite under their zero multipliers)
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          static_for<NR>([&](auto rc) __attribute__((always_inline)) {
            constexpr int R = decltype(rc)::value;
            if (lr + 16 * R == P[k]) {
#pragma unroll
              for (int c = 0; c < NC; ++c) rb[wv][k][lc + 4 * c] = x[R][c];
            }
          });
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // x -= M_j rb, one k at a time (straight-line code, so that the loads of k+1 are issued
        // under the multiply-adds of k; past N the multipliers are zero)
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          zd mk[NR], rk[NC];
#pragma unroll
          for (int r = 0; r < NR; ++r) mk[r] = mm[buf][k][lr + 16 * r];
#pragma unroll
          for (int c = 0; c < NC; ++c) rk[c] = rb[wv][k][lc + 4 * c];
#pragma unroll
          for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int c = 0; c < NC; ++c) x[r][c] = cfms(x[r][c], mk[r], rk[c]);
        }
        // panel j's own columns become S = M - I[:, P] (zero past N)
        if (j / NC == wv) {
          const unsigned sel = col_mask(j);
          static_for<NC>([&](auto cc) __attribute__((always_inline)) {
            constexpr int C = decltype(cc)::value;
            if (sel & (1u << C)) {
              const int pq = pp[buf][lc];
#pragma unroll
              for (int r = 0; r < NR; ++r) {
                zd v = mm[buf][lc][lr + 16 * r];
                if (lr + 16 * r == pq) v.re -= 1.0;
                x[r][C] = (NB * j + lc < N) ? v : cmk<double>(0, 0);
              }
            }
          });
        }
        if (j + 2 < NP) hand_over(j + 2, buf);
        // this phase's reads of the row buffer before the next phase's writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      } else if (j + 1 < NP) {
        if (stage_next) {
          for (int u = 0; u < IPP && sk < NI; ++u, ++sk) {
            stage_load(nb, sk, sqj, sqt);
            // 64 chunks on: at most two column steps, a column holding at least 33 chunks
            sqt += 64;
            if (sqt >= CPC) { sqt -= CPC; ++sqj; }
            if (sqt >= CPC) { sqt -= CPC; ++sqj; }
          }
        }
        factor(j + 1, true);
      }
      lds_barrier();
    }

    if (!upd) __builtin_amdgcn_s_setprio(0);
    for (int i = tid; i < N; i += NT) qinv[perm[i]] = i;
    // sign of the row permutation: parity of its inversions, counted in parallel (pair e = i N + j,
    // stepped by NT = di N + dj)
    {
      int odd = 0;
      const int di = NT / N, dj = NT - di * N;
      int i = tid / N, j = tid - (tid / N) * N;
      for (int e = tid; e < N * N; e += NT) {
        odd ^= (i < j && perm[i] > perm[j]) ? 1 : 0;
        i += di;
        j += dj;
        if (j >= N) { j -= N; ++i; }
      }
      if (odd) __hip_atomic_fetch_xor(&parity, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // det = +-(product of the pivots), kept as a complex mantissa z (larger part in [1/2, 1)) and a
    // power of two: wave 0, lane products then a fixed-order tree; the logarithms are taken after
    // the matrix loop
    zd z = cmk<double>(1.0, 0.0);
    int ez = 0;
    auto renorm = [&]() __attribute__((always_inline)) {
      int e;
      (void)frexp(fmax(fabs(z.re), fabs(z.im)), &e);
      z = cmk<double>(ldexp(z.re, -e), ldexp(z.im, -e));
      ez += e;
    };
    if (wv == 0) {
#pragma unroll
      for (int q = 0; q < (COLS + 63) / 64; ++q) {
        const int k = lane + 64 * q;
        if (k < N) {
          z = cmul(z, pivs[k]);
          renorm();
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const zd zo = cmk<double>(__shfl_xor(z.re, o, 64), __shfl_xor(z.im, o, 64));
        ez += __shfl_xor(ez, o, 64);
        z = cmul(z, zo);
        renorm();
      }
    }
    lds_barrier();
    if (tid == 0) {
      sm.dz[m] = parity ? cmk<double>(-z.re, -z.im) : z;
      sm.de[m] = singular ? GJB_SINGULAR : ez;
      if (p.info) p.info[bidx] = singular;
    }
    // inverse: O^{-1}[qinv[i]][perm[j]] = x[i][j] / (-pivot of row i)
    if (p.write_inverse && upd && STAGE) {
      // four columns at a time through the wave's part of the (now idle) panel buffers, rows put
      // in output order there, so that 16 lanes store 16 consecutive rows of a column
      ct* wb = reinterpret_cast<ct*>(&pc[0][0][0]) + wv * 4 * NMAX_B;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const int i = lr + 16 * r;
          if (i < N) wb[lc * NMAX_B + qinv[i]] = from_zd<T>(cmul(x[r][c], rowscale[i]));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int j = CPW * wv + lc + 4 * c;
        if (j < N) {
          ct* oc = out + (long long)perm[j] * p.ld_out;
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const int k = lr + 16 * r;
            if (k < N) oc[k] = wb[lc * NMAX_B + k];
          }
        }
        // the reads above before the next columns' writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    } else if (p.write_inverse && upd) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const int i = lr + 16 * r;
        if (i >= N) continue;
        const int row = qinv[i];
        const zd sc = rowscale[i];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int j = CPW * wv + lc + 4 * c;
          if (j >= N) continue;
          out[row + (long long)perm[j] * p.ld_out] = from_zd<T>(cmul(x[r][c], sc));
        }
      }
    }
    staged = stage_next;
    ++m;
    if (!STAGE) break;
  }
  // no staging load may still be writing LDS when the workgroup ends
  if (!upd) __builtin_amdgcn_s_waitcnt(0x0F70);
  // the log-determinants of the workgroup's matrices, one per thread
  __syncthreads();
  for (int q = threadIdx.x; q < m; q += NT) {
    const int e = sm.de[q];
    if (e == GJB_SKIPPED) continue;
    const zd zq = sm.dz[q];
    p.logdet[bk.x + q * bk.gx] = e == GJB_SINGULAR ? cmk<double>(-1e300, 0.0)
                                                   : cmk<double>(fma((double)e, M_LN2, log(hypot(zq.re, zq.im))),
                                                                 atan2(zq.im, zq.re));
  }
}

// Cholesky and R^{-1}, row layout: y[c][r] = element (wave + 16c, lane + 64r).
template <typename T, int NR, int NC>
struct CholK {
  static constexpr int NT = 64 * WAVES, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
template <typename T, int NR, int NC>
__device__ __forceinline__ void CholK<T, NR, NC>::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  const int bidx = bk.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  typedef cplx<T> ct;
  const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;

  __shared__ zd rowbuf[2][64 * NR];
  __shared__ double iskkbuf[2];
  __shared__ double dsq[64 * NR];  // D^-1/2
  __shared__ double skk_all[NMAX];  // r'_kk^2
  __shared__ double logr[NMAX];     // log R_kk
  __shared__ int bad;

  // S is Hermitian: element (i, j) is read as conj(S(j, i)), so consecutive lanes read
  // consecutive addresses
  // register column groups; with NR == 3 the third group (columns 128 + lane, lane < 32) is g2[wv]
  constexpr bool G2 = NR == 3;
  constexpr int RR = G2 ? 2 : NR;
  static_assert(!G2 || NMAX <= 160, "the LDS column group holds columns 128..159");
  __shared__ zd g2[G2 ? WAVES : 1][G2 ? NC : 1][32];
  const bool g2lane = G2 && lane < 32;
  zd y[NC][RR];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int i = wv + WAVES * c, j = lane + 64 * r;
      y[c][r] = (i < N && j < N) ? cconj(to_zd(in[j + (long long)i * p.ld_in])) : cmk<double>(0, 0);
    }
  if constexpr (G2) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = wv + WAVES * c, j = lane + 128;
      if (g2lane) g2[wv][c][lane] = (i < N && j < N) ? cconj(to_zd(in[j + (long long)i * p.ld_in])) : cmk<double>(0, 0);
    }
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int i = wv + WAVES * c, j = lane + 64 * r;
      if (i == j && i < N) {
        double d;
        if constexpr (G2) d = r < RR ? y[c][r < RR ? r : 0].re : g2[wv][c][lane].re;
        else d = y[c][r].re;
        if (!(d > 0.0)) { bad = 1; d = 1.0; }
        dsq[i] = 1.0 / sqrt(d);
        logr[i] = 0.5 * log(d);
      }
    }
  __syncthreads();
  double dcol[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int j = lane + 64 * r;
    dcol[r] = j < N ? dsq[j] : 0.0;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = wv + WAVES * c;
    if (i >= N) continue;
    const double di = dsq[i];
#pragma unroll
    for (int r = 0; r < RR; ++r) y[c][r] = cscale(y[c][r], di * dcol[r]);
    if constexpr (G2) {
      if (g2lane) g2[wv][c][lane] = cscale(g2[wv][c][lane], di * dcol[2]);
    }
  }

  // The owner of row k (register row CK of this wave, updated through step k-1) publishes it.
  auto publish = [&](auto ckc, int k, int buf) {
    constexpr int CK = decltype(ckc)::value;
    const int kr = k >> 6, kl = k & 63;
    double skk = 1.0;
    static_for<RR>([&](auto rc) {
      constexpr int R = decltype(rc)::value;
      if (kr == R) skk = readlane_d(y[CK][R].re, kl);
    });
    if constexpr (G2) {
      if (kr == 2) skk = readfirstlane_d(g2[wv][CK][kl].re);
    }
    const bool ok = skk > p.pivot_tol;
    if (!ok) skk = 1.0;
#pragma unroll
    for (int r = 0; r < RR; ++r) {
      const int j = lane + 64 * r;
      rowbuf[buf][j] = (j == k) ? cmk<double>(skk + 1.0, 0.0) : y[CK][r];
    }
    if constexpr (G2) {
      const int j = lane + 128;
      rowbuf[buf][j] = (j == k) ? cmk<double>(skk + 1.0, 0.0) : (g2lane ? g2[wv][CK][lane] : cmk<double>(0, 0));
    }
    if (lane == 0) {
      iskkbuf[buf] = recip(skk);
      skk_all[k] = skk;
      if (!ok) bad = 1;
    }
  };

  if (wv == 0) publish(std::integral_constant<int, 0>{}, 0, 0);
  __syncthreads();

  // Step k = CB*16 + t: row k is register row CB of wave t, row k+1 register row CB of wave t+1
  // (or CB+1 of wave 0).
  auto step = [&](auto cbc, int t) {
    constexpr int CB = decltype(cbc)::value;
    const int k = CB * WAVES + t;
    const int buf = k & 1, nbuf = buf ^ 1;
    const double iskk = iskkbuf[buf];
    zd rk[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) rk[r] = rowbuf[buf][lane + 64 * r];
    const bool next_here = (k + 1 < N) && (wv == (t + 1) % WAVES);
    const bool next_in_cb = (t + 1 < WAVES);
    auto update_row = [&](auto cc) {
      constexpr int C = decltype(cc)::value;
      const int i = wv + WAVES * C;
      const zd mi = cscale(cconj(rowbuf[buf][i]), iskk);
#pragma unroll
      for (int r = 0; r < RR; ++r) y[C][r] = cfms(y[C][r], mi, rk[r]);
      if constexpr (G2) {
        if (g2lane) g2[wv][C][lane] = cfms(g2[wv][C][lane], mi, rk[2]);
      }
    };
    if (next_here) {
      if (next_in_cb) {
        update_row(std::integral_constant<int, CB>{});
        publish(std::integral_constant<int, CB>{}, k + 1, nbuf);
      } else {
        if constexpr (CB + 1 < NC) {
          update_row(std::integral_constant<int, CB + 1>{});
          publish(std::integral_constant<int, CB + 1>{}, k + 1, nbuf);
        }
      }
    }
    static_for<NC>([&](auto cc) {
      constexpr int C = decltype(cc)::value;
      const int i = wv + WAVES * C;
      const bool done = next_here && ((next_in_cb && C == CB) || (!next_in_cb && C == CB + 1));
      if (i > k && i < N && !done) update_row(cc);
    });
    __syncthreads();
  };
  auto block = [&](auto cbc) {
    constexpr int CB = decltype(cbc)::value;
    for (int t = 0; t < WAVES && CB * WAVES + t < N; ++t) step(cbc, t);
  };
  static_for<NC>(block);

  // sum log R_kk (and over k < nsub): wave 0, fixed-order tree (deterministic)
  if (wv == 0) {
    double ls = 0.0, lsub = 0.0;
#pragma unroll
    for (int q = 0; q < (NMAX + 63) / 64; ++q) {
      const int k = lane + 64 * q;
      if (k < N) {
        const double l = logr[k] + 0.5 * log(skk_all[k]);
        ls += l;
        if (k < p.nsub) lsub += l;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ls += __shfl_xor(ls, o, 64);
      lsub += __shfl_xor(lsub, o, 64);
    }
    if (lane == 0) {
      p.logdet[bidx] = cmk<double>(ls, lsub);
      if (p.info) p.info[bidx] = bad;
    }
  }
  // R^{-1}(j, i) = dsq[j] conj(A(i, j)) for j < i with A(i, j) = y(i, j) / r'_ii; dsq[i] / r'_ii
  // on the diagonal; zero below
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = wv + WAVES * c;
    if (i >= N) continue;
    const double ri = 1.0 / sqrt(skk_all[i]);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int j = lane + 64 * r;
      if (j >= N) continue;
      zd yv;
      if constexpr (G2) yv = r < RR ? y[c][r < RR ? r : 0] : g2[wv][c][lane];
      else yv = y[c][r];
      zd v;
      if (j < i) v = cscale(cconj(yv), ri * dcol[r]);
      else if (j == i) v = cmk<double>(ri * dcol[r], 0.0);
      else v = cmk<double>(0, 0);
      out[j + (long long)i * p.ld_out] = from_zd<T>(v);
    }
  }
}

// 161 to 256 columns.  An N x N complex FP64 matrix is then up to 1 MB, more than a CU's registers
// and LDS hold together (512 KB + 160 KB), so one workgroup per matrix works on a copy in device
// memory (the L2 holds the rows and columns of a step) with one barrier per phase of a step:
// correct at these sizes, not tuned.  Thread t takes row t % 256 of columns t / 256 + 4c, so every
// pass over the matrix is coalesced down the columns.
constexpr int NWIDE = 256;
constexpr int WIDE_CSTEP = 64 * WAVES / NWIDE;  // column phases of a 1024-thread workgroup
constexpr int WIDE_UNROLL = 8;  // columns a thread loads before it stores (16 spills at the 64-VGPR limit, 2x slower)

// Gauss-Jordan exchange steps on the FP64 copy W (N x N per matrix in SmallDenseArgs::work), with
// the pivots of the register kernels (largest |re|+|im| among rows not yet used, first row on
// ties).  Exchanging column k with pivot row p: W[p][k] = 1/piv, W[p][j] = -W[p][j]/piv,
// W[i][k] = W[i][k]/piv, W[i][j] -= W[i][k] W[p][j]/piv.  After N steps
// O^{-1}[i][perm[j]] = W[perm[i]][j], perm[k] being column k's pivot row; the pivots equal LU's
// u_kk, and the log-determinant is formed from them and the parity of perm as in the other kernels.
template <typename T>
struct GjWideK {
  static constexpr int NT = 64 * WAVES, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
template <typename T>
__device__ __forceinline__ void GjWideK<T>::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  const int bidx = bk.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  typedef cplx<T> ct;
  const ct* in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;
  zd* W = reinterpret_cast<zd*>(p.work) + (long long)bidx * N * N;
  __shared__ zd prow[NWIDE], mcol[NWIDE], pivs[NWIDE];
  __shared__ int perm[NWIDE], used[NWIDE];
  __shared__ unsigned long long wkey[NWIDE / 64];
  __shared__ int singular, parity, spr;
  __shared__ zd sinv;
  const int i = tid & (NWIDE - 1), c0 = tid / NWIDE;
  if (i < N)
    for (int j = c0; j < N; j += WIDE_CSTEP) W[i + (long long)j * N] = to_zd(in[i + (long long)j * p.ld_in]);
  if (tid < NWIDE) used[tid] = 0;
  if (tid == 0) { singular = 0; parity = 0; }
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    if (tid < NWIDE) {
      unsigned long long key = 0;
      if (tid < N && !used[tid]) key = pivot_key(W[tid + (long long)k * N], tid);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const unsigned hi = __shfl_xor((unsigned)(key >> 32), o, 64), lo = __shfl_xor((unsigned)key, o, 64);
        const unsigned long long q = ((unsigned long long)hi << 32) | lo;
        key = q > key ? q : key;
      }
      if (lane == 0) wkey[wv] = key;
    }
    __syncthreads();
    if (tid == 0) {
      unsigned long long key = wkey[0];
      for (int q = 1; q < (N + 63) / 64; ++q) key = wkey[q] > key ? wkey[q] : key;
      const int pr = 255 - (int)(key & 0xFF);
      const zd piv = W[pr + (long long)k * N];
      const bool nz = abs1(piv) > 0.0;
      const double id = nz ? recip(piv.re * piv.re + piv.im * piv.im) : 0.0;
      sinv = cmk<double>(piv.re * id, -piv.im * id);
      spr = pr; perm[k] = pr; used[pr] = 1; pivs[k] = piv;
      if (!nz) singular = 1;
    }
    __syncthreads();
    const int pr = spr;
    const zd inv = sinv;
    if (tid < N) {
      prow[tid] = W[pr + (long long)tid * N];
      mcol[tid] = tid == pr ? cmk<double>(0, 0) : cmul(W[tid + (long long)k * N], inv);
    }
    __syncthreads();
    if (i < N) {
      // eight independent loads in flight per thread before the stores
      const zd mi = mcol[i];
      for (int j0 = c0; j0 < N; j0 += WIDE_UNROLL * WIDE_CSTEP) {
        zd v[WIDE_UNROLL];
#pragma unroll
        for (int u = 0; u < WIDE_UNROLL; ++u) {
          const int j = j0 + u * WIDE_CSTEP;
          if (j < N) v[u] = W[i + (long long)j * N];
        }
#pragma unroll
        for (int u = 0; u < WIDE_UNROLL; ++u) {
          const int j = j0 + u * WIDE_CSTEP;
          if (j >= N) continue;
          zd r;
          if (j == k) r = i == pr ? inv : mi;
          else if (i == pr) r = cmul(cmk<double>(-prow[j].re, -prow[j].im), inv);
          else r = cfms(v[u], mi, prow[j]);
          W[i + (long long)j * N] = r;
        }
      }
    }
    __syncthreads();
  }
  // parity of the row permutation: inversions counted per row
  if (tid < N) {
    int odd = 0;
    for (int j = tid + 1; j < N; ++j) odd ^= perm[tid] > perm[j] ? 1 : 0;
    if (odd) atomicXor(&parity, 1);
  }
  __syncthreads();
  if (tid == 0) {
    double lognorm = 0.0;
    zd phase = cmk<double>(1.0, 0.0);
    for (int k = 0; k < N; ++k) {
      const zd pv = pivs[k];
      const double ap = cabs(pv);
      const double ia = ap > 0.0 ? 1.0 / ap : 0.0;
      lognorm += log(ap);
      phase = cmul(phase, cmk<double>(pv.re * ia, pv.im * ia));
    }
    if (parity) phase = cmk<double>(-phase.re, -phase.im);
    p.logdet[bidx] = singular ? cmk<double>(-1e300, 0.0)
                              : cmk<double>(lognorm + log(hypot(phase.re, phase.im)), atan2(phase.im, phase.re));
    if (p.info) p.info[bidx] = singular;
  }
  if (!p.write_inverse) return;
  if (i < N) {
    const int row = perm[i];
    for (int j = c0; j < N; j += WIDE_CSTEP) out[i + (long long)perm[j] * p.ld_out] = from_zd<T>(W[row + (long long)j * N]);
  }
}

// Cholesky S = R^H R and R^{-1} in place in the FP64 output (161 to 256 columns).  After the Jacobi
// scaling S' = D^-1/2 S D^-1/2 (a non-positive diagonal counts as a breakdown and scales by 1), the
// right-looking factorisation of S' = R'^H R' keeps R' in the upper triangle and the Hermitian
// trailing matrix above its diagonal; a pivot of S' at or below pivot_tol is a breakdown and is
// replaced by 1.  Thread i then forms column i of X = R'^{-1} by back substitution into row i of the
// lower triangle, and the output is R^{-1} = D^-1/2 X, upper triangular, zero below.  logdet gets
// (sum log R_kk, sum_{k<nsub} log R_kk), info the breakdown flag, as in CholK.
struct CholWideK {
  static constexpr int NT = 64 * WAVES, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const SmallDenseArgs& p);
};
__device__ __forceinline__ void CholWideK::run(const BlockIdx3& bk, const SmallDenseArgs& p) {
  const int bidx = bk.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const long long ld = p.ld_out;
  const int tid = threadIdx.x;
  const zd* in = reinterpret_cast<const zd*>(p.in) + (long long)bidx * p.stride_in;
  zd* W = reinterpret_cast<zd*>(p.out) + (long long)bidx * p.stride_out;
  __shared__ double dsq[NWIDE], logr[NWIDE], skk_all[NWIDE], rdiag[NWIDE];
  __shared__ zd rrow[NWIDE];
  __shared__ int bad;
  const int i = tid & (NWIDE - 1), c0 = tid / NWIDE;
  if (tid == 0) bad = 0;
  if (in != W && i < N)
    for (int j = c0; j < N; j += WIDE_CSTEP) W[i + (long long)j * ld] = in[i + (long long)j * p.ld_in];
  __syncthreads();
  if (tid < N) {
    double d = W[tid + (long long)tid * ld].re;
    if (!(d > 0.0)) { bad = 1; d = 1.0; }
    dsq[tid] = 1.0 / sqrt(d);
    logr[tid] = 0.5 * log(d);
  }
  __syncthreads();
  if (i < N)
    for (int j = c0; j < N; j += WIDE_CSTEP) W[i + (long long)j * ld] = cscale(W[i + (long long)j * ld], dsq[i] * dsq[j]);
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    // row k of R': r'_kk = sqrt(s'_kk), r'_kj = s'_kj / r'_kk (every thread reads the same s'_kk)
    double skk = W[k + (long long)k * ld].re;
    const bool ok = skk > p.pivot_tol;
    if (!ok) skk = 1.0;
    const double rk = sqrt(skk);
    if (tid == 0) {
      skk_all[k] = skk;
      rdiag[k] = rk;
      if (!ok) bad = 1;
    }
    if (tid < N && tid > k) {
      const zd v = cscale(W[k + (long long)tid * ld], 1.0 / rk);
      W[k + (long long)tid * ld] = v;
      rrow[tid] = v;
    }
    __syncthreads();
    if (tid == k) W[k + (long long)k * ld] = cmk<double>(rk, 0.0);
    // trailing upper triangle: s'_ij -= conj(r'_ki) r'_kj for k < i <= j
    if (i < N && i > k) {
      const zd ci = cconj(rrow[i]);
      for (int j0 = c0; j0 < N; j0 += WIDE_UNROLL * WIDE_CSTEP) {
        zd v[WIDE_UNROLL];
#pragma unroll
        for (int u = 0; u < WIDE_UNROLL; ++u) {
          const int j = j0 + u * WIDE_CSTEP;
          if (j >= i && j < N) v[u] = W[i + (long long)j * ld];
        }
#pragma unroll
        for (int u = 0; u < WIDE_UNROLL; ++u) {
          const int j = j0 + u * WIDE_CSTEP;
          if (j >= i && j < N) W[i + (long long)j * ld] = cfms(v[u], ci, rrow[j]);
        }
      }
    }
    __syncthreads();
  }
  // X = R'^{-1}: x_ii = 1 / r'_ii, x_ji = -(sum_{j<l<=i} r'_jl x_li) / r'_jj, into W[i + j ld]
  if (tid < N) {
    const int c = tid;
    const double xd = 1.0 / rdiag[c];
    for (int j = c - 1; j >= 0; --j) {
      zd s = cmul(W[j + (long long)c * ld], cmk<double>(xd, 0.0));
#pragma unroll 8
      for (int l = j + 1; l < c; ++l) s = cadd(s, cmul(W[j + (long long)l * ld], W[c + (long long)l * ld]));
      W[c + (long long)j * ld] = cscale(s, -1.0 / rdiag[j]);
    }
  }
  __syncthreads();
  // R^{-1}(j, i) = dsq[j] x_ji above the diagonal, dsq[i] / r'_ii on it, zero below
  if (i < N)
    for (int j = c0; j < i; j += WIDE_CSTEP) {
      const zd x = W[i + (long long)j * ld];
      W[j + (long long)i * ld] = cscale(x, dsq[j]);
      W[i + (long long)j * ld] = cmk<double>(0, 0);
    }
  if (tid < N) W[tid + (long long)tid * ld] = cmk<double>(dsq[tid] / rdiag[tid], 0.0);
  if (tid == 0) {
    double ls = 0.0, lsub = 0.0;
    for (int k = 0; k < N; ++k) {
      const double l = logr[k] + 0.5 * log(skk_all[k]);
      ls += l;
      if (k < p.nsub) lsub += l;
    }
    p.logdet[bidx] = cmk<double>(ls, lsub);
    if (p.info) p.info[bidx] = bad;
  }
}

// compute units of the current device (cached per device; threads racing on the first call all
// store the same value)
int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cached[dev] = n;
  }
  return cached[dev];
}

template <typename T, int NR, int NC, int WV = WAVES, int NL = NC>
int run_gj(const SmallDenseArgs& a, hipStream_t s) {
  return glaunch<GjK<T, NR, NC, WV, NL>>(a, dim3(a.batch), dim3(64 * WV), 0, s);
}
template <typename T, int WV, int LR, int NR, int NC, int WPE>
int run_gj2(const SmallDenseArgs& a, hipStream_t s) {
  return glaunch<Gj2K<T, WV, LR, NR, NC, WPE>>(a, dim3(a.batch), dim3(64 * WV), 0, s);
}
template <typename T, int NR, int NC, int WV, int NRF, int WPE>
int run_gjb(const SmallDenseArgs& a, hipStream_t s) {
  typedef GjBK<T, NR, NC, WV, NRF, WPE> K;
  // with staging, one workgroup per CU (its registers allow no more) walks over the matrices
  const int grid = K::STAGE ? std::max(std::min(a.batch, cu_count()), (a.batch + GJB_MAXM - 1) / GJB_MAXM) : a.batch;
  return glaunch<K>(a, dim3(grid), dim3(64 * (WV + 1)), 0, s);
}
template <typename T, int NR, int NC>
int run_chol(const SmallDenseArgs& a, hipStream_t s) {
  return glaunch<CholK<T, NR, NC>>(a, dim3(a.batch), dim3(64 * WAVES), 0, s);
}

template <typename T>
int dispatch(int kind, const SmallDenseArgs& a, hipStream_t s) {
  REQUIRE(a.N >= 1 && a.N <= NWIDE, "small dense kernels support 1 <= N <= %d (got %d)", NWIDE, a.N);
  REQUIRE(a.alive_div >= 1, "alive_div must be >= 1");
  if (a.batch <= 0) return 0;
  if (a.N > NMAX) {
    // beyond the register-resident kernels: working copies in device memory
    if (kind == 0) {
      REQUIRE(a.work != nullptr, "Gauss-Jordan of %d columns needs a work buffer of %lld bytes", a.N,
              small_dense_gj_work_bytes(a.N, a.batch));
      return glaunch<GjWideK<T>>(a, dim3(a.batch), dim3(64 * WAVES), 0, s);
    }
    REQUIRE(sizeof(T) == 8, "Cholesky of %d columns runs in FP64 only", a.N);
    REQUIRE(a.ld_out >= a.N && a.stride_out >= a.ld_out * a.N, "Cholesky of %d columns: output layout too small", a.N);
    return glaunch<CholWideK>(a, dim3(a.batch), dim3(64 * WAVES), 0, s);
  }
  if (kind == 0) {
    // Defaults, from 4096 (112, 96 columns), 2048 (64), 1024 (72) and 3072 (40, 24) inversions on
    // MI355X: the 2-D layout up to 32 columns (24: 0.070 ms against 0.101 for the 1-D layout; from
    // 33 columns on it is slower, its registers allowing two waves per SIMD with the LDS round
    // trips of every step on the chain); the panel kernel for 33 to 48 columns (40: 0.148 against
    // 0.262 ms, four matrices per CU) and 65 to 112 (112: 2.16 against 2.34 ms; 72: 0.309 against
    // 0.343; 96: 1.72 against 1.74), where its registers allow one or two matrices per CU; the 1-D
    // layout otherwise (64: 0.29 against 0.42 for the panel kernel, which fits two matrices per CU
    // there against the 1-D kernel's four).  The choice depends on N only, so that a grouped call
    // and the members' own calls do the same arithmetic (the panel kernel's results differ from the
    // others' in rounding).
    // a grouped call launches the matrices of all its members together
    const bool many = a.batch > cu_count() || grouping();
    const bool panels = a.N <= 48 || (a.N > 64 && a.N <= 112);
    const int layout = a.layout ? a.layout : a.N <= 32 ? 2 : panels ? 3 : 1;
    REQUIRE(layout >= 1 && layout <= 3, "Gauss-Jordan layout %d unknown", layout);
    if (layout == 3 && a.N <= 112) {
      // ceil(N / 16) update waves of 16 columns and one pivoting wave
      if (a.N <= 16) return run_gjb<T, 1, 4, 1, 1, 4>(a, s);
      if (a.N <= 32) return run_gjb<T, 2, 4, 2, 1, 4>(a, s);
      if (a.N <= 48) return run_gjb<T, 3, 4, 3, 1, 4>(a, s);
      if (a.N <= 64) return run_gjb<T, 4, 4, 4, 1, 4>(a, s);
      if (a.N <= 80) return run_gjb<T, 5, 4, 5, 2, 2>(a, s);
      if (a.N <= 96) return run_gjb<T, 6, 4, 6, 2, 2>(a, s);
      return run_gjb<T, 7, 4, 7, 2, 2>(a, s);
    }
    if (layout == 2 && a.N <= 128) {
      if (a.N <= 32) return run_gj2<T, 2, 32, 1, 8, 4>(a, s);
      if (a.N <= 64) return run_gj2<T, 4, 32, 2, 8, 2>(a, s);
      if (a.N <= 96) return run_gj2<T, 8, 32, 3, 6, 2>(a, s);
      if (a.N <= 112) return run_gj2<T, 8, 32, 4, 7, 2>(a, s);
      return run_gj2<T, 8, 32, 4, 8, 2>(a, s);
    }
    // up to 64 columns, fewer waves per matrix and more matrices per CU: 8 waves, or 4 once
    // there are more matrices than CUs (2048 x 64^2: 0.28 against 0.36 ms with 16 waves;
    // 1024 x 40^2: 0.10 against 0.12 ms)
    if (a.N <= 32) return many ? run_gj<T, 1, 8, 4, 8>(a, s) : run_gj<T, 1, 4, 8, 4>(a, s);
    if (a.N <= 64) return many ? run_gj<T, 1, 16, 4, 16>(a, s) : run_gj<T, 1, 8, 8, 8>(a, s);
    if (a.N <= 96) return many ? run_gj<T, 2, 12, 8, 6>(a, s) : run_gj<T, 2, 6>(a, s);
    // 8 waves per matrix with half of the second row group in LDS: two matrices share a CU and
    // each hides the other's per-step synchronisation (2048 x 112^2: 1.14 against 1.36 ms), which
    // pays once there are more matrices than CUs
    if (a.N <= 112) return many ? run_gj<T, 2, 14, 8, 7>(a, s) : run_gj<T, 2, 7>(a, s);
    if (a.N <= 128) return run_gj<T, 2, 8>(a, s);
    return run_gj<T, 3, 10, WAVES, 0>(a, s);
  }
  if (a.N <= 16) return run_chol<T, 1, 1>(a, s);
  if (a.N <= 32) return run_chol<T, 1, 2>(a, s);
  if (a.N <= 64) return run_chol<T, 1, 4>(a, s);
  if (a.N <= 96) return run_chol<T, 2, 6>(a, s);
  if (a.N <= 112) return run_chol<T, 2, 7>(a, s);
  if (a.N <= 128) return run_chol<T, 2, 8>(a, s);
  return run_chol<T, 3, 10>(a, s);
}

}  // namespace

long long small_dense_gj_work_bytes(int N, int batch) {
  return N > NMAX ? (long long)sizeof(zd) * N * N * std::max(batch, 0) : 0;
}

template <typename T>
int launch_gj_inverse(const SmallDenseArgs& a, hipStream_t s) { return dispatch<T>(0, a, s); }
template <typename T>
int launch_chol_inverse(const SmallDenseArgs& a, hipStream_t s) { return dispatch<T>(1, a, s); }

template int launch_gj_inverse<double>(const SmallDenseArgs&, hipStream_t);
template int launch_gj_inverse<float>(const SmallDenseArgs&, hipStream_t);
template int launch_chol_inverse<double>(const SmallDenseArgs&, hipStream_t);
template int launch_chol_inverse<float>(const SmallDenseArgs&, hipStream_t);
