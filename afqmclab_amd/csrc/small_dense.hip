// Per-walker N x N dense kernels, register resident, one workgroup per matrix.
//
// The N x N complex FP64 matrix (N <= 160; 112 x 112 is 196 KB, more than a
// CU's LDS) is distributed 2D-cyclically over a TG x TG thread grid: thread
// (tr, tc) owns rows tr + TG*a and columns tc + TG*b.  Per elimination step one
// column travels through LDS (double buffered) and the pivot row by in-wave
// shuffle; the pivot of the next column is found one step ahead by the lanes
// that own it, so each step has one barrier.
//
// gj_inverse_logdet: Gauss-Jordan exchange sweep with partial pivoting (the
//   pivot of column k is the largest |re|+|im| among rows not yet used, first
//   index on ties, as izamax in zgetrf).  The pivots equal LU's u_kk, so
//   logdet = sum log|u_kk| + log(prod u_kk/|u_kk| * sign(perm)) reproduces
//   logDeterminant (tensor_2d_common_fun.cpp:75-105).  Writes O^{-1}.
// chol_inverse: S = R^H R (upper R, positive real diagonal), then R^{-1} by
//   the same sweep without pivoting; returns sum log R_kk (CholeskyQR step of
//   the walker restabilisation).
#include "small_dense.hpp"

#include <type_traits>
#include <utility>

namespace {

constexpr int NMAX = 160;

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

// Gauss-Jordan with look-ahead: the threads that own column k+1 (TG consecutive lanes of one
// wave) update that column first during step k, find its pivot with a reduction inside their
// lane group and publish the column and the pivot row through LDS, then do the rest of their
// update.  The pivot search is thereby off the critical path of the other waves, and each
// step has one barrier.  The pivot row index is uniform over the workgroup (read as a scalar),
// so the pivot row is taken from registers by a scalar branch and an in-wave shuffle.
template <typename T, int TG, int E>
__global__ __launch_bounds__(TG * TG) void gj_la_kernel(SmallDenseArgs p) {
  const int bidx = blockIdx.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int tid = threadIdx.x;
  const int tr = tid % TG, tc = tid / TG;
  typedef cplx<T> ct;
  const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;

  __shared__ zd colbuf[2][NMAX];
  __shared__ int pivrow[2];
  __shared__ zd pivinv[2];
  __shared__ zd pivs[NMAX];
  __shared__ int perm[NMAX];
  __shared__ int qinv[NMAX];
  __shared__ unsigned char seen[NMAX];

  zd x[E][E];
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int c = 0; c < E; ++c) {
      int i = tr + TG * a, j = tc + TG * c;
      x[a][c] = (i < N && j < N) ? to_zd(in[i + (long long)j * p.ld_in]) : cmk<double>(0, 0);
    }
  unsigned used = 0u;  // bit a: row tr + TG*a has been a pivot row
  int singular = 0;

  // The owners of column `col` (TG consecutive lanes) publish it: value(a) gives their entry in
  // register row a; the pivot row is found by a reduction inside the lane group, and the lane
  // holding the pivot publishes its index and inverse (0 for a zero pivot).
  auto publish_col = [&](auto value, int col, int buf) {
    unsigned long long key = 0ull;
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = tr + TG * a;
      if (i < N) {
        const zd v = value(a);
        colbuf[buf][i] = v;
        if (!((used >> a) & 1u)) {
          const unsigned long long kk = pivot_key(v, i);
          key = kk > key ? kk : key;
        }
      }
    }
#pragma unroll
    for (int o = TG / 2; o > 0; o >>= 1) {
      const unsigned long long ok = (unsigned long long)__shfl_xor((long long)key, o, 64);
      key = ok > key ? ok : key;
    }
    // bit 63 marks a candidate: at step col there are N - col >= 1 unused rows
    const int bi = 255 - (int)(key & 0xFFull);
    const int prn = (bi >= 0 && bi < N) ? bi : 0;
    if (tr == prn % TG) {
      const zd pv = colbuf[buf][prn];  // written by this lane above
      pivrow[buf] = prn;
      pivinv[buf] = abs1(pv) > 0.0 ? cinv(pv) : cmk<double>(0, 0);
    }
  };

  if (tc == 0) publish_col([&](int a) { return x[a][0]; }, 0, 0);
  __syncthreads();

  // Step k = CB*TG + t: column k sits in register column CB (a compile-time constant), column
  // k+1 in CB or CB+1, so no register array is ever indexed by a run-time value.
  auto step = [&](auto cbc, int t) {
    constexpr int CB = decltype(cbc)::value;
    const int k = CB * TG + t;
    const int buf = k & 1;
    const int pr = __builtin_amdgcn_readfirstlane(pivrow[buf]);
    const zd piv = colbuf[buf][pr];
    const zd inv = pivinv[buf];
    if (!(abs1(piv) > 0.0)) singular = 1;
    if (tid == 0) {
      perm[k] = pr;
      pivs[k] = piv;
    }
    const int ar = pr / TG, prm = pr % TG;
    const bool prow = (tr == prm);
    if (prow) used |= 1u << ar;
    // rowv[c] = x[pr][j] / piv (ar is uniform, so these are scalar branches around the shuffles);
    // the update below leaves column k to its own fix-up
    zd rowv[E];
    {
      const int src = prm + TG * (tc % (64 / TG));
#pragma unroll
      for (int a = 0; a < E; ++a) {
        if (a == ar) {
#pragma unroll
          for (int c = 0; c < E; ++c) {
            zd v;
            v.re = __shfl(x[a][c].re, src, 64);
            v.im = __shfl(x[a][c].im, src, 64);
            rowv[c] = cmul(v, inv);
          }
        }
      }
    }
    zd cv[E];
#pragma unroll
    for (int a = 0; a < E; ++a) {
      const int i = tr + TG * a;
      cv[a] = i < N ? colbuf[buf][i] : cmk<double>(0, 0);
    }
    // look-ahead: the owners of column k+1 compute its updated values first and publish them
    // (the update below recomputes the same values in place)
    auto lookahead = [&](auto c1c) {
      constexpr int C1 = decltype(c1c)::value;
      publish_col([&](int a) {
        return (prow && a == ar) ? cmk<double>(-rowv[C1].re, -rowv[C1].im) : cfms(x[a][C1], cv[a], rowv[C1]);
      }, k + 1, buf ^ 1);
    };
    if (k + 1 < N && tc == (t + 1) % TG) {
      if (t + 1 < TG) {
        lookahead(std::integral_constant<int, CB>{});
      } else {
        if constexpr (CB + 1 < E) lookahead(std::integral_constant<int, CB + 1>{});
      }
    }
    // rank-1 update, then the pivot row and column k
#pragma unroll
    for (int c = 0; c < E; ++c)
#pragma unroll
      for (int a = 0; a < E; ++a) x[a][c] = cfms(x[a][c], cv[a], rowv[c]);
    if (prow) {
#pragma unroll
      for (int a = 0; a < E; ++a)
        if (a == ar)
#pragma unroll
          for (int c = 0; c < E; ++c) x[a][c] = cmk<double>(-rowv[c].re, -rowv[c].im);
    }
    if (tc == t) {
#pragma unroll
      for (int a = 0; a < E; ++a) x[a][CB] = (tr + TG * a == pr) ? inv : cmul(cv[a], inv);
    }
    __syncthreads();
  };
  auto block = [&](auto cbc) {
    constexpr int CB = decltype(cbc)::value;
    for (int t = 0; t < TG && CB * TG + t < N; ++t) step(cbc, t);
  };
  static_for<E>(block);
  if (tid < N) qinv[perm[tid]] = tid;
  if (tid == 0) {
    double lognorm = 0.0;
    zd phase = cmk<double>(1.0, 0.0);
    if (!singular) {
      for (int k = 0; k < N; ++k) {
        const zd pv = pivs[k];
        const double ap = cabs(pv);
        lognorm += log(ap);
        phase = cmul(phase, cmk<double>(pv.re / ap, pv.im / ap));
      }
    }
    int visited_parity = 0;
    for (int i = 0; i < N; ++i) seen[i] = 0;
    for (int i = 0; i < N; ++i) {
      if (seen[i]) continue;
      int len = 0, j = i;
      while (len < N && !seen[j]) { seen[j] = 1; j = perm[j]; ++len; }
      visited_parity ^= (len - 1) & 1;
    }
    zd ph = visited_parity ? cmk<double>(-phase.re, -phase.im) : phase;
    zd ld;
    if (singular) {
      ld = cmk<double>(-1e300, 0.0);
    } else {
      ld = cmk<double>(lognorm + 0.0, atan2(ph.im, ph.re));
      ld.re += log(hypot(ph.re, ph.im));
    }
    p.logdet[bidx] = ld;
    if (p.info) p.info[bidx] = singular;
  }
  if (!p.write_inverse) return;
  __syncthreads();
#pragma unroll
  for (int a = 0; a < E; ++a) {
    int i = tr + TG * a;
    if (i >= N) continue;
    int row = qinv[i];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      int j = tc + TG * c;
      if (j >= N) continue;
      out[row + (long long)perm[j] * p.ld_out] = from_zd<T>(x[a][c]);
    }
  }
}

template <typename T, int TG, int E>
__global__ __launch_bounds__(TG * TG) void chol_kernel(SmallDenseArgs p) {
  const int bidx = blockIdx.x;
  if (p.alive && !p.alive[bidx / p.alive_div]) return;
  const int N = p.N;
  const int tid = threadIdx.x;
  const int tr = tid % TG, tc = tid / TG;
  typedef cplx<T> ct;
  const ct* __restrict__ in = reinterpret_cast<const ct*>(p.in) + (long long)bidx * p.stride_in;
  ct* __restrict__ out = reinterpret_cast<ct*>(p.out) + (long long)bidx * p.stride_out;

  __shared__ zd colbuf[2][NMAX];
  __shared__ zd rowbuf[2][NMAX];

  zd x[E][E];
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int c = 0; c < E; ++c) {
      int i = tr + TG * a, j = tc + TG * c;
      x[a][c] = (i < N && j < N) ? to_zd(in[i + (long long)j * p.ld_in]) : cmk<double>(0, 0);
    }
  double logsum = 0.0, logsub = 0.0;
  int bad = 0;
  // Cholesky, upper factor: one barrier per step (double-buffered row)
  for (int k = 0; k < N; ++k) {
    const int buf = k & 1;
    if (tr == k % TG) {
      const int ar = k / TG;
#pragma unroll
      for (int a = 0; a < E; ++a) {
        if (a != ar) continue;
#pragma unroll
        for (int c = 0; c < E; ++c) {
          int j = tc + TG * c;
          if (j < N) rowbuf[buf][j] = x[a][c];
        }
      }
    }
    __syncthreads();
    double skk = rowbuf[buf][k].re;
    if (!(skk > 0.0)) { bad = 1; skk = 1.0; }
    const double d = sqrt(skk), id = 1.0 / d;
    if (tid == 0) {
      double l = log(d);
      logsum += l;
      if (k < p.nsub) logsub += l;
    }
#pragma unroll
    for (int a = 0; a < E; ++a) {
      int i = tr + TG * a;
      if (i >= N || i < k) continue;
      zd rki = cscale(rowbuf[buf][i], id);  // R[k,i]
#pragma unroll
      for (int c = 0; c < E; ++c) {
        int j = tc + TG * c;
        if (j >= N || j < k) continue;
        zd rkj = cscale(rowbuf[buf][j], id);
        if (i == k) x[a][c] = rkj;
        else x[a][c] = csub(x[a][c], cmul(cconj(rki), rkj));
      }
    }
  }
  // zero the strictly lower part
#pragma unroll
  for (int a = 0; a < E; ++a)
#pragma unroll
    for (int c = 0; c < E; ++c) {
      int i = tr + TG * a, j = tc + TG * c;
      if (j < i) x[a][c] = cmk<double>(0, 0);
    }
  // R^{-1} by the exchange sweep with pivot (k, k); the pivot row reaches each thread by an
  // in-wave shuffle (as in gj_la_kernel), so one barrier per step
  __syncthreads();
  for (int k = 0; k < N; ++k) {
    const int buf = k & 1;
    if (tc == k % TG) {
      const int kc = k / TG;
#pragma unroll
      for (int a = 0; a < E; ++a) {
        int i = tr + TG * a;
#pragma unroll
        for (int c = 0; c < E; ++c)
          if (c == kc && i < N) colbuf[buf][i] = x[a][c];
      }
    }
    __syncthreads();
    const zd inv = cinv(colbuf[buf][k]);
    zd rowv[E];
    {
      const int ar = k / TG;
      const int src = (k % TG) + TG * (tc % (64 / TG));
#pragma unroll
      for (int c = 0; c < E; ++c) {
        zd v = cmk<double>(0, 0);
#pragma unroll
        for (int a = 0; a < E; ++a)
          if (a == ar) v = x[a][c];
        v.re = __shfl(v.re, src, 64);
        v.im = __shfl(v.im, src, 64);
        rowv[c] = cmul(v, inv);
      }
    }
#pragma unroll
    for (int a = 0; a < E; ++a) {
      int i = tr + TG * a;
      if (i >= N) continue;
      zd ci = colbuf[buf][i];
#pragma unroll
      for (int c = 0; c < E; ++c) {
        int j = tc + TG * c;
        if (j >= N) continue;
        zd r = rowv[c];
        if (i == k) {
          x[a][c] = (j == k) ? inv : cmk<double>(-r.re, -r.im);
        } else if (j == k) {
          x[a][c] = cmul(ci, inv);
        } else {
          x[a][c] = csub(x[a][c], cmul(ci, r));
        }
      }
    }
  }
  if (tid == 0) {
    p.logdet[bidx] = cmk<double>(logsum, logsub);
    if (p.info) p.info[bidx] = bad;
  }
#pragma unroll
  for (int a = 0; a < E; ++a) {
    int i = tr + TG * a;
    if (i >= N) continue;
#pragma unroll
    for (int c = 0; c < E; ++c) {
      int j = tc + TG * c;
      if (j >= N) continue;
      out[i + (long long)j * p.ld_out] = from_zd<T>(x[a][c]);
    }
  }
}

template <typename T, int TG, int E>
int run(int kind, const SmallDenseArgs& a, hipStream_t s) {
  dim3 grid(a.batch), block(TG * TG);
  if (kind == 0) hipLaunchKernelGGL((gj_la_kernel<T, TG, E>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((chol_kernel<T, TG, E>), grid, block, 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int dispatch(int kind, const SmallDenseArgs& a, hipStream_t s) {
  REQUIRE(a.N >= 1 && a.N <= NMAX, "small dense kernels support 1 <= N <= %d (got %d)", NMAX, a.N);
  REQUIRE(a.alive_div >= 1, "alive_div must be >= 1");
  if (a.batch <= 0) return 0;
  if (a.N <= 16) return run<T, 8, 2>(kind, a, s);
  if (a.N <= 32) return run<T, 16, 2>(kind, a, s);
  if (a.N <= 64) return run<T, 16, 4>(kind, a, s);
  if (a.N <= 128) return run<T, 32, 4>(kind, a, s);
  return run<T, 32, 5>(kind, a, s);
}

}  // namespace

template <typename T>
int launch_gj_inverse(const SmallDenseArgs& a, hipStream_t s) { return dispatch<T>(0, a, s); }
template <typename T>
int launch_chol_inverse(const SmallDenseArgs& a, hipStream_t s) { return dispatch<T>(1, a, s); }

template int launch_gj_inverse<double>(const SmallDenseArgs&, hipStream_t);
template int launch_gj_inverse<float>(const SmallDenseArgs&, hipStream_t);
template int launch_chol_inverse<double>(const SmallDenseArgs&, hipStream_t);
template int launch_chol_inverse<float>(const SmallDenseArgs&, hipStream_t);
