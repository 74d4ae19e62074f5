// Phaseless walker kernels (see ph_kernels.hpp for the reference functions).
#include "ph_kernels.hpp"

namespace {

constexpr double PI_REF = 3.14159265358979324;  // random_hao.cpp:79

__device__ double bsum_d(double v, double* sh) {
  const int tid = threadIdx.x;
  v = wave_sum(v);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}
__device__ zd bsum_z(zd v, double* sh) { return cmk<double>(bsum_d(v.re, sh), bsum_d(v.im, sh)); }

// gaussianHao: sqrt(-2 ln u1) cos(2 pi u2), u1 == 0 -> 1e-300 (random_hao.cpp:78-85)
__device__ __forceinline__ double gaussian(double u1, double u2) {
  if (u1 == 0.0) u1 += 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * PI_REF * u2);
}

template <typename T>
__global__ __launch_bounds__(256) void ph_sample_kernel(PhSampleArgs a) {
  const int w = blockIdx.x;
  if (!a.alive[w]) return;
  __shared__ double sh[8];
  const int G = a.G;
  typedef cplx<T> ct;
  const ct* bgw = reinterpret_cast<const ct*>(a.bg) + (long long)w * G;
  ct* xsw = reinterpret_cast<ct*>(a.xs) + (long long)w * G;
  const unsigned long long base = a.sc->counter + (unsigned long long)a.prefix[w] * 2ull * (unsigned long long)G;
  const unsigned long long key = a.sc->key;
  zd f2 = cmk<double>(0, 0), xf = f2, xb = f2;
  for (int g = threadIdx.x; g < G; g += blockDim.x) {
    zd f;
    if (a.force_type == 1) {
      f = a.const_force[g];
    } else {
      // f_g = sqrt(-dt) (bg_g - currentBg_g), rescaled to |f| <= cap (realMaterialMoleculeMeasureFixedSD2sSD2is.cpp:107-113)
      zd d = to_zd(bgw[g]);
      d.re -= a.cur_bg[g];
      f = cmul(d, a.sqrt_minus_dt);
      const double af = cabs(f);
      if (af > a.cap) f = cscale(f, a.cap / af);
    }
    const unsigned long long k = base + 2ull * (unsigned long long)g;
    const double u1 = afqmc_take_uniform(a.ubuf, a.sc, key, k);
    const double u2 = afqmc_take_uniform(a.ubuf, a.sc, key, k + 1);
    const zd x = cmk<double>(gaussian(u1, u2) + f.re, f.im);
    f2 = cadd(f2, cmul(f, f));
    xf = cadd(xf, cmul(x, f));
    xb = cadd(xb, cscale(x, a.cur_bg[g]));
    xsw[g] = from_zd<T>(cmul(a.sqrt_minus_dt, x));
    if (a.force_out) a.force_out[(long long)w * G + g] = f;
    if (a.aux_out) a.aux_out[(long long)w * G + g] = x;
  }
  f2 = bsum_z(f2, sh);
  xf = bsum_z(xf, sh);
  xb = bsum_z(xb, sh);
  if (threadIdx.x == 0) {
    // sample logw = 1/2 sum f^2 - sum x f - sqrt(-dt) sum x bg (choleskyReal.cpp:104-124)
    const zd sq = cmul(a.sqrt_minus_dt, xb);
    const zd slw = cmk<double>(0.5 * f2.re - xf.re - sq.re, 0.5 * f2.im - xf.im - sq.im);
    if (a.sample_logw_out) a.sample_logw_out[w] = slw;
    // second cos projection on Im(sum x f) (afqmcPhaselessProjection.cpp:108-111)
    const double c = cos(xf.im);
    if (c <= 0.0) {
      a.alive[w] = 0;
    } else {
      zd lw = a.logw[w];
      lw.re += log(c) + slw.re + a.dt * a.sc->ET + a.onebody_logw.re;
      lw.im += slw.im + a.onebody_logw.im;
      a.logw[w] = lw;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void taylor_acc_kernel(long long n, const cplx<T>* __restrict__ Tk,
                                                          cplx<T>* __restrict__ W, int* active, int* orders,
                                                          int order, int base, double accuracy, int dynamic,
                                                          int* n_active) {
  const int w = blockIdx.x;
  if (!active[w]) return;
  __shared__ double sh[8];
  double mx = 0.0;
  const cplx<T>* t = Tk + (long long)w * n;
  cplx<T>* o = W + (long long)w * n;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const cplx<T> v = t[i];
    o[i] = cadd(o[i], v);
    const double a = hypot((double)v.re, (double)v.im);
    mx = a > mx ? a : mx;
  }
  // block max
  for (int off = 32; off > 0; off >>= 1) {
    double other = __shfl_xor(mx, off, 64);
    mx = other > mx ? other : mx;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) m = sh[i] > m ? sh[i] : m;
    orders[w] = order;
    const bool done = dynamic ? (order > base && !(m > accuracy)) : (order >= base);
    if (done) {
      active[w] = 0;
    } else {
      atomicAdd(n_active, 1);
    }
  }
}

// tr(D_g D_g) of each Ns x Ns block D_g (rows g Ns .. g Ns + Ns - 1 of the walker's D).  With
// STAGE the block is first read column by column (contiguous rows, coalesced) into LDS with a
// padded stride, so the transposed operand is an LDS read rather than a global load G Ns
// elements apart; both forms add the same products in the same order.
template <typename T, bool STAGE>
__global__ __launch_bounds__(256) void exchange_kernel(int G, int Ns, const cplx<T>* __restrict__ D, zd* ex,
                                                        int accumulate) {
  const int w = blockIdx.y;
  const int g = blockIdx.x;
  const long long ld = (long long)G * Ns;
  const cplx<T>* d = D + (long long)w * ld * Ns + (long long)g * Ns;
  __shared__ double sh[8];
  extern __shared__ unsigned char stage_mem[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(stage_mem);  // tile[c (Ns + 1) + r] = D_g(r, c)
  if (STAGE) {
    for (int e = threadIdx.x; e < Ns * Ns; e += blockDim.x) {
      const int r = e % Ns, c = e / Ns;
      tile[c * (Ns + 1) + r] = d[r + (long long)c * ld];
    }
    __syncthreads();
  }
  zd s = cmk<double>(0, 0);
  for (int e = threadIdx.x; e < Ns * Ns; e += blockDim.x) {
    const int i = e / Ns, j = e % Ns;
    const zd a = to_zd(STAGE ? tile[i * (Ns + 1) + j] : d[j + (long long)i * ld]);
    const zd b = to_zd(STAGE ? tile[j * (Ns + 1) + i] : d[i + (long long)j * ld]);
    s = cadd(s, cmul(a, b));
  }
  s = bsum_z(s, sh);
  if (threadIdx.x == 0) {
    zd* o = ex + (long long)w * G + g;
    *o = accumulate ? cadd(*o, s) : s;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void rowdot_kernel(long long n, const cplx<T>* X, long long strideX, const cplx<T>* P,
                                                     zd* out, int accumulate) {
  const int w = blockIdx.x;
  __shared__ double sh[8];
  zd s = cmk<double>(0, 0);
  for (long long i = threadIdx.x; i < n; i += blockDim.x)
    s = cadd(s, cmul(to_zd(X[(long long)w * strideX + i]), to_zd(P[i])));
  s = bsum_z(s, sh);
  if (threadIdx.x == 0) out[w] = accumulate ? cadd(out[w], s) : s;
}

template <typename T>
__global__ __launch_bounds__(256) void ph_energy_kernel(PhEnergyArgs a) {
  const int w = blockIdx.x;
  zd* out = a.per_walker + (long long)w * 4;
  if (!a.alive[w]) {
    if (threadIdx.x < 4) out[threadIdx.x] = cmk<double>(0, 0);
    return;
  }
  __shared__ double sh[8];
  const cplx<T>* bg = reinterpret_cast<const cplx<T>*>(a.bg) + (long long)w * a.G;
  zd s = cmk<double>(0, 0);
  for (int g = threadIdx.x; g < a.G; g += blockDim.x) {
    const zd b = to_zd(bg[g]);
    s = cadd(s, csub(cmul(b, b), a.ex[(long long)w * a.G + g]));
  }
  s = bsum_z(s, sh);
  if (threadIdx.x == 0) {
    const zd H = cadd(cscale(s, 0.5), a.tkin[w]);  // :223-229
    const zd d = a.den[w];
    out[0] = d;
    out[1] = cmul(H, d);
    out[2] = cmul(a.tkin[w], d);
    out[3] = H;
  }
}

__global__ __launch_bounds__(256) void ph_energy_reduce_kernel(int nw, const zd* per_walker, DevScalars* sc) {
  __shared__ zd part[256][3];
  const int tid = threadIdx.x;
  const int chunk = (nw + 255) / 256;
  zd s0 = cmk<double>(0, 0), s1 = s0, s2 = s0;
  for (int w = tid * chunk; w < min(nw, (tid + 1) * chunk); ++w) {
    s0 = cadd(s0, per_walker[(long long)w * 4]);
    s1 = cadd(s1, per_walker[(long long)w * 4 + 1]);
    s2 = cadd(s2, per_walker[(long long)w * 4 + 2]);
  }
  part[tid][0] = s0; part[tid][1] = s1; part[tid][2] = s2;
  __syncthreads();
  if (tid == 0) {
    zd t0 = cmk<double>(0, 0), t1 = t0, t2 = t0;
    for (int i = 0; i < 256; ++i) { t0 = cadd(t0, part[i][0]); t1 = cadd(t1, part[i][1]); t2 = cadd(t2, part[i][2]); }
    sc->den = cadd(sc->den, t0);
    sc->HNum = cadd(sc->HNum, t1);
    sc->KNum = cadd(sc->KNum, t2);  // TNum
  }
}

template <typename T>
__global__ void ph_g_reduce_kernel(PhEnergyArgs a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G) return;
  const cplx<T>* bg = reinterpret_cast<const cplx<T>*>(a.bg);
  zd sb = cmk<double>(0, 0), se = sb;
  for (int w = 0; w < a.nw; ++w) {
    if (!a.alive[w]) continue;
    const zd d = a.den[w];
    sb = cadd(sb, cmul(to_zd(bg[(long long)w * a.G + g]), d));
    se = cadd(se, cmul(a.ex[(long long)w * a.G + g], d));
  }
  a.bg_acc[g] = cadd(a.bg_acc[g], sb);
  a.ex_acc[g] = cadd(a.ex_acc[g], se);
}

}  // namespace

int launch_ph_sample(const PhSampleArgs& a, hipStream_t s) {
  if (a.fp32) hipLaunchKernelGGL(ph_sample_kernel<float>, dim3(a.nw), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(ph_sample_kernel<double>, dim3(a.nw), dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_taylor_accumulate(int nw, long long n, const void* Tk, void* W, int* active, int* orders, int order,
                             int base, double accuracy, int dynamic, int* n_active, hipStream_t s) {
  hipLaunchKernelGGL(taylor_acc_kernel<T>, dim3(nw), dim3(256), 0, s, n, (const cplx<T>*)Tk, (cplx<T>*)W, active,
                     orders, order, base, accuracy, dynamic, n_active);
  HIP_LAUNCH_CHECK();
  return 0;
}
template int launch_taylor_accumulate<double>(int, long long, const void*, void*, int*, int*, int, int, double, int,
                                              int*, hipStream_t);
template int launch_taylor_accumulate<float>(int, long long, const void*, void*, int*, int*, int, int, double, int,
                                             int*, hipStream_t);

template <typename T>
int launch_exchange_trace(int nw, int G, int Ns, const void* D, zd* ex, int accumulate, hipStream_t s) {
  // blocks whose padded copy fits 64 KB of LDS are staged (Ns up to 63 in FP64, 90 in FP32);
  // larger ones read global memory
  const size_t stage = (size_t)Ns * (Ns + 1) * sizeof(cplx<T>);
  if (stage <= 65536)
    hipLaunchKernelGGL((exchange_kernel<T, true>), dim3(G, nw), dim3(256), stage, s, G, Ns, (const cplx<T>*)D, ex,
                       accumulate);
  else
    hipLaunchKernelGGL((exchange_kernel<T, false>), dim3(G, nw), dim3(256), 0, s, G, Ns, (const cplx<T>*)D, ex,
                       accumulate);
  HIP_LAUNCH_CHECK();
  return 0;
}
template int launch_exchange_trace<double>(int, int, int, const void*, zd*, int, hipStream_t);
template int launch_exchange_trace<float>(int, int, int, const void*, zd*, int, hipStream_t);

template <typename T>
int launch_rowdot_sum(int nw, long long n, const void* X, long long strideX, const void* P, zd* out, int accumulate,
                      hipStream_t s) {
  hipLaunchKernelGGL(rowdot_kernel<T>, dim3(nw), dim3(256), 0, s, n, (const cplx<T>*)X, strideX, (const cplx<T>*)P,
                     out, accumulate);
  HIP_LAUNCH_CHECK();
  return 0;
}
template int launch_rowdot_sum<double>(int, long long, const void*, long long, const void*, zd*, int, hipStream_t);
template int launch_rowdot_sum<float>(int, long long, const void*, long long, const void*, zd*, int, hipStream_t);

int launch_ph_energy(const PhEnergyArgs& a, hipStream_t s) {
  if (a.fp32) hipLaunchKernelGGL(ph_energy_kernel<float>, dim3(a.nw), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(ph_energy_kernel<double>, dim3(a.nw), dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(ph_energy_reduce_kernel, dim3(1), dim3(256), 0, s, a.nw, a.per_walker, a.sc);
  HIP_LAUNCH_CHECK();
  if (a.fp32) hipLaunchKernelGGL(ph_g_reduce_kernel<float>, dim3((a.G + 63) / 64), dim3(64), 0, s, a);
  else hipLaunchKernelGGL(ph_g_reduce_kernel<double>, dim3((a.G + 63) / 64), dim3(64), 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}
