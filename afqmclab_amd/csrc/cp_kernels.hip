// Walker-level kernels of the CP step (afqmcConstraintPathProjection.cpp:37-156),
// the HS sampling of NiupNidn (NiupNidn.cpp:66-322), the mixed energy of
// HubbardSOCMeasureFixedSDSD (:178-254) and the comb (populationControlConfiguration.cpp:24-101).
#include "cp_kernels.hpp"

#include "group.hpp"

#include <algorithm>
#include <type_traits>

unsigned long long afqmc_seed_mix_host(unsigned long long seed) {
  unsigned long long z = seed * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

namespace {

__device__ double block_sum_d(double v, double* sh) {
  const int tid = threadIdx.x;
  v = wave_sum(v);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  const int nwv = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nwv; ++i) t += sh[i];
  __syncthreads();
  return t;
}
__device__ zd block_sum_z(zd v, double* sh) {
  double r = block_sum_d(v.re, sh);
  double i = block_sum_d(v.im, sh);
  return cmk<double>(r, i);
}

// ------------------------------------------------------------------ logO
struct LogoK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const LogOArgs& a);
};
__device__ __forceinline__ void LogoK::run(const BlockIdx3& bk, const LogOArgs& a) {
  const int w = bk.x * blockDim.x + threadIdx.x;
  if (w >= a.nw) return;
  if (a.mode != LOGO_MEASURE && a.mode != LOGO_ONLY && !a.alive[w]) {
    if (a.mode == LOGO_WEIGHT) a.weight[w] = 0.0;
    return;
  }
  if (a.mode == LOGO_MEASURE && !a.alive[w]) {
    a.den[w] = cmk<double>(0, 0);
    return;
  }
  double re = 0.0, ang = 0.0;
  int singular = 0;
  if (a.combine == COMBINE_WRAP && a.nblk > 1) {
    zd ph = cmk<double>(1.0, 0.0);
    for (int b = 0; b < a.nblk; ++b) {
      zd l = a.ldet[b][w];
      re += l.re;
      double s, c;
      sincos(l.im, &s, &c);
      ph = cmul(ph, cmk<double>(c, s));
      if (a.info[b][w]) singular = 1;
    }
    ang = atan2(ph.im, ph.re);
  } else {
    for (int b = 0; b < a.nblk; ++b) {
      zd l = a.ldet[b][w];
      re += l.re;
      ang += l.im;
      if (a.info[b][w]) singular = 1;
    }
  }
  if (singular) { re = -1e300; ang = 0.0; }
  zd lw = a.logw[w];
  zd lo = cmk<double>(a.trial_logw.re + lw.re + re, -a.trial_logw.im + lw.im + ang);
  a.logO[w] = lo;
  if (a.mode == LOGO_STEP) {
    double c = cos(lo.im);
    if (c <= 0.0) {
      a.alive[w] = 0;
    } else {
      a.logw_out[w] = cmk<double>(lw.re + log(c), lw.im - lo.im);
    }
  } else if (a.mode == LOGO_WEIGHT) {
    double o = (lo.re < -700.0) ? 0.0 : exp(lo.re) * cos(lo.im);
    if (o <= 0.0) {
      a.weight[w] = 0.0;
      a.alive[w] = 0;
    } else {
      a.weight[w] = o;
      a.logw_out[w] = cmk<double>(lw.re - lo.re, lw.im - lo.im);
    }
  } else if (a.mode == LOGO_MEASURE) {
    a.den[w] = cexp_d(lo);
  }
}

// ------------------------------------------------------------------ alive prefix
struct AliveScanArgs {
  const int* alive;
  int* prefix;
  DevScalars* sc;
  int nw, pad;
};
struct AliveScanK {
  static constexpr int NT = 1024, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const AliveScanArgs& g);
};
__device__ __forceinline__ void AliveScanK::run(const BlockIdx3& bk, const AliveScanArgs& g) {
  const int nw = g.nw;
  const int* alive = g.alive;
  int* prefix = g.prefix;
  DevScalars* sc = g.sc;

  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int chunk = (nw + 1023) / 1024;
  const int lo = tid * chunk, hi = min(nw, lo + chunk);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += alive[i] ? 1 : 0;
  part[tid] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = (tid >= o) ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int base = part[tid] - s;
  for (int i = lo; i < hi; ++i) {
    prefix[i] = base;
    base += alive[i] ? 1 : 0;
  }
  if (tid == 1023) sc->alive_count = part[1023];
}

// ------------------------------------------------------------------ HS sampling
__device__ __forceinline__ zd gsum(const zd* g, int w, int ntile, int Mb, int row) {
  zd s = cmk<double>(0, 0);
  const zd* p = g + (long long)w * ntile * Mb + row;
  for (int t = 0; t < ntile; ++t) s = cadd(s, p[(long long)t * Mb]);
  return s;
}

struct SiteG { zd gup, gdn, off_i, off_iL; };

__device__ __forceinline__ SiteG site_green(const zd* const* gd, const zd* const* go, const int* ntile, const int* Mb,
                                            int dense, int w, int i, int L) {
  SiteG g;
  if (dense) {
    g.gup = gsum(gd[0], w, ntile[0], Mb[0], i);
    g.gdn = gsum(gd[0], w, ntile[0], Mb[0], i + L);
    if (go[0]) {
      // d1 of X row r is G(r^, r): greenOffDiagonal(i) = G(i, i+L) sits at row i+L
      g.off_i = gsum(go[0], w, ntile[0], Mb[0], i + L);
      g.off_iL = gsum(go[0], w, ntile[0], Mb[0], i);
    } else {
      g.off_i = g.off_iL = cmk<double>(0, 0);
    }
  } else {
    g.gup = gsum(gd[0], w, ntile[0], Mb[0], i);
    g.gdn = gsum(gd[1], w, ntile[1], Mb[1], i);
    g.off_i = g.off_iL = cmk<double>(0, 0);
  }
  return g;
}

__device__ __forceinline__ zd cexp_c(zd a) { return cexp_d(a); }
__device__ __forceinline__ zd ccosh_c(zd a) {
  zd e = cexp_d(a), em = cexp_d(cmk<double>(-a.re, -a.im));
  return cmk<double>(0.5 * (e.re + em.re), 0.5 * (e.im + em.im));
}
__device__ __forceinline__ zd csinh_c(zd a) {
  zd e = cexp_d(a), em = cexp_d(cmk<double>(-a.re, -a.im));
  return cmk<double>(0.5 * (e.re - em.re), 0.5 * (e.im - em.im));
}
__device__ __forceinline__ zd ccos_c(zd a) { return ccosh_c(cmk<double>(-a.im, a.re)); }      // cos z = cosh(iz)
__device__ __forceinline__ zd csin_c(zd a) {                                                     // sin z = -i sinh(iz)
  zd s = csinh_c(cmk<double>(-a.im, a.re));
  return cmk<double>(s.im, -s.re);
}

// The 2x2 site matrix [d00 d01; d10 d11] acting on (up_i, dn_i) of the sample with gamma_i x_i = gx
// (NiupNidn.cpp:269-322): the field exponential times the constant one-body factor constDiag.
__device__ __forceinline__ void hs_site_matrix(int decomp, zd gx, const zd* cdiag, int L, int i, zd& d00, zd& d01,
                                               zd& d10, zd& d11) {
  const zd c00 = cdiag[i], c01 = cdiag[L + i], c10 = cdiag[2 * L + i], c11 = cdiag[3 * L + i];
  switch (decomp) {
    case DECOMP_DENSITY_CHARGE: {
      zd e = cexp_c(gx);
      d00 = cmul(e, c00); d10 = cmul(e, c10); d01 = cmul(e, c01); d11 = cmul(e, c11);
      break;
    }
    case DECOMP_DENSITY_SPIN: {
      zd e = cexp_c(gx), einv = cinv(e);
      d00 = cmul(e, c00); d10 = cmul(einv, c10); d01 = cmul(e, c01); d11 = cmul(einv, c11);
      break;
    }
    case DECOMP_HOP_CHARGE: {
      zd ch = ccosh_c(gx), sn = csinh_c(gx);
      d00 = cadd(cmul(ch, c00), cmul(sn, c10)); d10 = cadd(cmul(sn, c00), cmul(ch, c10));
      d01 = cadd(cmul(ch, c01), cmul(sn, c11)); d11 = cadd(cmul(sn, c01), cmul(ch, c11));
      break;
    }
    default: {
      zd co = ccos_c(gx), si = csin_c(gx);
      d00 = cadd(cmul(co, c00), cmul(si, c10)); d10 = csub(cmul(co, c10), cmul(si, c00));
      d01 = cadd(cmul(co, c01), cmul(si, c11)); d11 = csub(cmul(co, c11), cmul(si, c01));
      break;
    }
  }
}

template <typename T>
struct HsSampleK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const HsSampleArgs& a);
};
template <typename T>
__device__ __forceinline__ void HsSampleK<T>::run(const BlockIdx3& bk, const HsSampleArgs& a) {
  const int w = bk.x;
  if (!a.alive[w]) return;
  __shared__ double sh[16];
  const int L = a.L;
  const unsigned long long base = a.sc->counter + (unsigned long long)a.prefix[w] * (unsigned long long)L;
  const unsigned long long key = a.sc->key;
  double lp = 0.0;
  zd xg = cmk<double>(0, 0);
  typedef cplx<T> ct;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    SiteG g = site_green(a.gd, a.go, a.ntile, a.Mb, a.dense, w, i, L);
    double f;
    if (a.force_type == 1) {
      f = a.const_force[i];
    } else {
      zd bg;
      switch (a.decomp) {
        case DECOMP_DENSITY_CHARGE: bg = cmk<double>(g.gup.re + g.gdn.re - 1.0, g.gup.im + g.gdn.im); break;
        case DECOMP_DENSITY_SPIN: bg = csub(g.gup, g.gdn); break;
        case DECOMP_HOP_CHARGE: bg = cadd(g.off_i, g.off_iL); break;
        default: bg = csub(g.off_i, g.off_iL); break;
      }
      f = cmul(a.gamma[i], bg).re;
      if (f > a.cap) f = a.cap;
      if (f < -a.cap) f = -a.cap;
    }
    const double em = exp(-f), ep = exp(f);
    int x;
    if (a.aux_in) {
      x = a.aux_in[(long long)w * L + i];
    } else {
      const unsigned long long k = base + (unsigned long long)i;
      const double u = afqmc_take_uniform(a.ubuf, a.sc, key, k);
      const double prob = em / (em + ep);
      x = (u < prob) ? -1 : 1;
    }
    lp += x * f - log(ep + em);
    const zd gx = cscale(a.gamma[i], (double)x);
    xg = cadd(xg, gx);
    zd d00, d01, d10, d11;
    hs_site_matrix(a.decomp, gx, a.cdiag, L, i, d00, d01, d10, d11);
    if (a.diag_only) {
      if (a.dense) {
        ct* s0 = reinterpret_cast<ct*>(a.scale[0]) + (long long)w * a.Mb[0];
        s0[i] = from_zd<T>(d00);
        s0[i + L] = from_zd<T>(d11);
      } else {
        reinterpret_cast<ct*>(a.scale[0])[(long long)w * a.Mb[0] + i] = from_zd<T>(d00);
        reinterpret_cast<ct*>(a.scale[1])[(long long)w * a.Mb[1] + i] = from_zd<T>(d11);
      }
    } else {
      ct* d = reinterpret_cast<ct*>(a.d2x2) + (long long)w * 4 * L;
      d[i] = from_zd<T>(d00);
      d[L + i] = from_zd<T>(d01);
      d[2 * L + i] = from_zd<T>(d10);
      d[3 * L + i] = from_zd<T>(d11);
    }
    if (a.force_out) a.force_out[(long long)w * L + i] = f;
    if (a.aux_out) a.aux_out[(long long)w * L + i] = x;
  }
  lp = block_sum_d(lp, sh);
  xg = block_sum_z(xg, sh);
  if (threadIdx.x == 0) {
    zd sl = cmk<double>(a.log_half_L, 0.0);
    if (a.decomp == DECOMP_DENSITY_CHARGE) sl = cmk<double>(sl.re + a.dtUSum * 0.5 - xg.re, -xg.im);
    sl.re -= lp;
    zd lw = a.logw[w];
    lw.re += sl.re + a.dt * a.sc->ET + a.onebody_logw.re;
    lw.im += sl.im + a.onebody_logw.im;
    a.logw[w] = lw;
  }
}

struct AdvanceArgs {
  DevScalars* sc;
  int per_walker, pad;
};
struct AdvanceK {
  static constexpr int NT = 64, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const AdvanceArgs& a);
};
__device__ __forceinline__ void AdvanceK::run(const BlockIdx3& bk, const AdvanceArgs& a) {
  DevScalars* sc = a.sc;
  const int per_walker = a.per_walker;

  sc->counter += (unsigned long long)per_walker * (unsigned long long)sc->alive_count;
  sc->walker_steps += (unsigned long long)sc->alive_count;
}

// ------------------------------------------------------------------ HS 2x2 apply
// NiupNidnSampleSDOperation::applyToRight (NiupNidnSDOperation.cpp:15-41) on dense 2L x N walkers:
// for every column j and site i, (w'(i), w'(i+L)) = [d00 d01; d10 d11]_i (w(i), w(i+L)).  Grid
// (column chunk, walker): the walker's 4 x L site matrices are staged in LDS once per workgroup,
// and each work item moves S consecutive sites of one column as 16-byte vectors (S = 2 complex
// FP32 or 1 complex FP64; S = 1 with 8-byte FP32 accesses when L is odd).  No 64-bit division.
template <typename T>
struct HsApplyArgs {
  const cplx<T>* in;
  cplx<T>* out;
  const cplx<T>* d;
  int L, N, cpb, pad;
};
template <typename T, int S>
struct HsApplyK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const HsApplyArgs<T>& g);
};
template <typename T, int S>
__device__ __forceinline__ void HsApplyK<T, S>::run(const BlockIdx3& bk, const HsApplyArgs<T>& g) {
  const int L = g.L, N = g.N, cpb = g.cpb;
  const cplx<T>* __restrict__ in = g.in;
  cplx<T>* __restrict__ out = g.out;
  const cplx<T>* __restrict__ d = g.d;

  typedef cplx<T> ct;
  struct alignas(S * sizeof(ct)) vt { ct z[S]; };
  extern __shared__ __align__(16) unsigned char hs_lds[];
  vt* sd = reinterpret_cast<vt*>(hs_lds);
  const int w = bk.y, tid = threadIdx.x;
  const int nv = L / S;  // site vectors per half column
  const vt* dw = reinterpret_cast<const vt*>(d + (long long)w * 4 * L);
  for (int k = tid; k < 4 * nv; k += blockDim.x) sd[k] = dw[k];
  __syncthreads();
  const int j0 = bk.x * cpb, j1 = min(N, j0 + cpb);
  const long long base = (long long)w * 2 * L * N;
  auto apply = [&](const vt& a, const vt& b, const vt& d00, const vt& d01, const vt& d10, const vt& d11, vt& up,
                   vt& dn) {
#pragma unroll
    for (int t = 0; t < S; ++t) {
      up.z[t] = cadd(cmul(d00.z[t], a.z[t]), cmul(d01.z[t], b.z[t]));
      dn.z[t] = cadd(cmul(d10.z[t], a.z[t]), cmul(d11.z[t], b.z[t]));
    }
  };
  if (nv <= (int)blockDim.x) {
    // thread -> (column offset, site vector), fixed over the loop, so its site matrices are read
    // from LDS once.  Four columns per iteration with every load issued before the first store:
    // eight 16-byte loads in flight per thread (one column at a time left the CU short of
    // outstanding loads: 0.59-0.62 of the HBM peak).
    constexpr int U = 4;
    const int cpp = blockDim.x / nv, jj = tid / nv, v = tid - jj * nv;
    if (jj < cpp) {
      const vt d00 = sd[v], d01 = sd[nv + v], d10 = sd[2 * nv + v], d11 = sd[3 * nv + v];
      for (int j = j0 + jj; j < j1; j += U * cpp) {
        vt a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (j + u * cpp < j1) {
            const long long o = base + (long long)(j + u * cpp) * 2 * L + (long long)v * S;
            a[u] = *reinterpret_cast<const vt*>(in + o);
            b[u] = *reinterpret_cast<const vt*>(in + o + L);
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (j + u * cpp < j1) {
            const long long o = base + (long long)(j + u * cpp) * 2 * L + (long long)v * S;
            vt up, dn;
            apply(a[u], b[u], d00, d01, d10, d11, up, dn);
            *reinterpret_cast<vt*>(out + o) = up;
            *reinterpret_cast<vt*>(out + o + L) = dn;
          }
        }
      }
    }
  } else {
    for (int j = j0; j < j1; ++j)
      for (int v = tid; v < nv; v += blockDim.x) {
        const long long o = base + (long long)j * 2 * L + (long long)v * S;
        vt up, dn;
        apply(*reinterpret_cast<const vt*>(in + o), *reinterpret_cast<const vt*>(in + o + L), sd[v], sd[nv + v],
              sd[2 * nv + v], sd[3 * nv + v], up, dn);
        *reinterpret_cast<vt*>(out + o) = up;
        *reinterpret_cast<vt*>(out + o + L) = dn;
      }
  }
}

// ------------------------------------------------------------------ comb
// popConfiguration (populationControlConfiguration.cpp:24-101): sequential
// left-to-right sums in one thread, so bin edges round exactly as on the host.
struct CombArgs {
  const double* weight;
  const double* ubuf;
  DevScalars* sc;
  int* table;
  double* prob;
  int* number;
  int L, pad;
};
struct CombK {
  static constexpr int NT = 64, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const CombArgs& g);
};
__device__ __forceinline__ void CombK::run(const BlockIdx3& bk, const CombArgs& g) {
  const int L = g.L;
  const double* weight = g.weight;
  const double* ubuf = g.ubuf;
  DevScalars* sc = g.sc;
  int* table = g.table;
  double* prob = g.prob;
  int* number = g.number;

  if (threadIdx.x != 0) return;
  const unsigned long long key = sc->key;
  const unsigned long long c0 = sc->counter;
  double sum = 0.0, maxW = 0.0, minW = 1e20;
  for (int i = 0; i < L; ++i) {
    double w = weight[i];
    sum += w;
    if (maxW < w) maxW = w;
    if (minW > w) minW = w;
  }
  sc->wmin = minW;
  sc->wmax = maxW;
  sc->wavg = sum / (L * 1.0);
  for (int i = 0; i < L; ++i) {
    prob[i] = weight[i] / sum;
    number[i] = 0;
  }
  int index_old = 0;
  double prob_sum_old = prob[0];
  for (int i = 0; i < L; ++i) {
    const unsigned long long k = c0 + (unsigned long long)i;
    const double u = afqmc_take_uniform(ubuf, sc, key, k);
    const double prob_sum_new = (i + u) / (L * 1.0);
    while (prob_sum_new > prob_sum_old) {
      if (index_old + 1 >= L) break;  // the reference has no guard here
      index_old++;
      prob_sum_old += prob[index_old];
    }
    number[index_old]++;
  }
  sc->counter = c0 + (unsigned long long)L;
  for (int i = 0; i < L; ++i) table[i] = (number[i] != 0) ? i : -1;
  int zero_i = 0, one_i = 0;
  while (zero_i < L && number[zero_i] != 0) zero_i++;
  while (one_i < L && number[one_i] <= 1) one_i++;
  while (zero_i < L && one_i < L) {
    table[zero_i] = one_i;
    number[zero_i]++;
    number[one_i]--;
    while (zero_i < L && number[zero_i] != 0) zero_i++;
    while (one_i < L && number[one_i] <= 1) one_i++;
  }
}

// The same comb with the walkers in LDS: the two running sums (the weight total and the cumulative
// probabilities) stay sequential in one lane, so every partial sum is rounded as the reference
// rounds it, while the teeth search the cumulative sums in parallel.  The reference's scan stops at
// the first bin j (from the previous tooth's bin on) with !(tooth > cumulative[j]); the teeth
// (i + u_i) / L increase with i and the cumulative sums do not decrease, so that bin is the first j
// over the whole array where the predicate holds, found by binary search (clamped at L-1 as the scan).
constexpr int COMB_LDS_MAX = 4096;
struct CombLdsK {
  static constexpr int NT = 1024, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const CombArgs& g);
};
__device__ __forceinline__ void CombLdsK::run(const BlockIdx3& bk, const CombArgs& g) {
  const int L = g.L;
  const double* weight = g.weight;
  const double* ubuf = g.ubuf;
  DevScalars* sc = g.sc;
  int* table = g.table;
  double* prob = g.prob;
  int* number = g.number;

  __shared__ double p[COMB_LDS_MAX];
  __shared__ int num[COMB_LDS_MAX];
  __shared__ double s_sum;
  const int tid = threadIdx.x, T = blockDim.x;
  for (int i = tid; i < L; i += T) {
    p[i] = weight[i];
    num[i] = 0;
  }
  __syncthreads();
  if (tid == 0) {
    double sum = 0.0, maxW = 0.0, minW = 1e20;
    for (int i = 0; i < L; ++i) {
      const double w = p[i];
      sum += w;
      if (maxW < w) maxW = w;
      if (minW > w) minW = w;
    }
    sc->wmin = minW;
    sc->wmax = maxW;
    sc->wavg = sum / (L * 1.0);
    s_sum = sum;
  }
  __syncthreads();
  const double sum = s_sum;
  for (int i = tid; i < L; i += T) {
    const double q = p[i] / sum;
    prob[i] = q;
    p[i] = q;
  }
  __syncthreads();
  if (tid == 0) {
    double acc = p[0];
    for (int j = 1; j < L; ++j) {
      acc += p[j];
      p[j] = acc;
    }
  }
  __syncthreads();
  const unsigned long long key = sc->key;
  const unsigned long long c0 = sc->counter;
  for (int i = tid; i < L; i += T) {
    const double u = afqmc_take_uniform(ubuf, sc, key, c0 + (unsigned long long)i);
    const double tooth = (i + u) / (L * 1.0);
    int lo = 0, hi = L - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (!(tooth > p[mid])) hi = mid;
      else lo = mid + 1;
    }
    atomicAdd(&num[lo], 1);
  }
  __syncthreads();
  if (tid == 0) {
    sc->counter = c0 + (unsigned long long)L;
    for (int i = 0; i < L; ++i) table[i] = (num[i] != 0) ? i : -1;
    int zero_i = 0, one_i = 0;
    while (zero_i < L && num[zero_i] != 0) zero_i++;
    while (one_i < L && num[one_i] <= 1) one_i++;
    while (zero_i < L && one_i < L) {
      table[zero_i] = one_i;
      num[zero_i]++;
      num[one_i]--;
      while (zero_i < L && num[zero_i] != 0) zero_i++;
      while (one_i < L && num[one_i] <= 1) one_i++;
    }
  }
  __syncthreads();
  for (int i = tid; i < L; i += T) number[i] = num[i];
}

template <typename T>
struct ApplyTableArgs {
  cplx<T>* wf0;
  cplx<T>* wf1;
  long long ws0, ws1;
  zd* logw;
  int* alive;
  const int* table;
  int nw, pad;
};
template <typename T>
struct ApplyTableK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const ApplyTableArgs<T>& g);
};
template <typename T>
__device__ __forceinline__ void ApplyTableK<T>::run(const BlockIdx3& bk, const ApplyTableArgs<T>& g) {
  cplx<T>* wf0 = g.wf0;
  cplx<T>* wf1 = g.wf1;
  const long long ws0 = g.ws0, ws1 = g.ws1;
  zd* logw = g.logw;
  int* alive = g.alive;
  const int* table = g.table;

  const int i = bk.y;
  const int src = table[i];
  if (src != i && src >= 0) {
    for (long long e = (long long)bk.x * blockDim.x + threadIdx.x; e < ws0; e += (long long)bk.gx * blockDim.x)
      wf0[(long long)i * ws0 + e] = wf0[(long long)src * ws0 + e];
    if (wf1)
      for (long long e = (long long)bk.x * blockDim.x + threadIdx.x; e < ws1; e += (long long)bk.gx * blockDim.x)
        wf1[(long long)i * ws1 + e] = wf1[(long long)src * ws1 + e];
    if (bk.x == 0 && threadIdx.x == 0) logw[i] = logw[src];
  }
  if (bk.x == 0 && threadIdx.x == 0) alive[i] = 1;
}

// ------------------------------------------------------------------ energies
__global__ __launch_bounds__(256) void energy_kernel(EnergyArgs a) {
  const int w = blockIdx.x;
  __shared__ double sh[16];
  zd* out = a.per_walker + (long long)w * 8;
  if (!a.alive[w]) {
    if (threadIdx.x < 8) out[threadIdx.x] = cmk<double>(0, 0);
    return;
  }
  const int L = a.L;
  zd V = cmk<double>(0, 0), R = V, K = V, nup = V, ndn = V;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    SiteG g = site_green(a.gd, a.go, a.ntile, a.Mb, a.dense, w, i, L);
    zd t = csub(cmul(g.gup, g.gdn), cmul(g.off_i, g.off_iL));
    V = cadd(V, cscale(t, a.U[i]));
    R = cadd(R, cscale(g.gup, -a.mu[i] + a.hz[i] * 0.5));
    R = cadd(R, cscale(g.gdn, -a.mu[i] - a.hz[i] * 0.5));
    R = cadd(R, cmul(cmk<double>(a.hx[i] * 0.5, -a.hy[i] * 0.5), g.off_i));
    R = cadd(R, cmul(cmk<double>(a.hx[i] * 0.5, a.hy[i] * 0.5), g.off_iL));
    nup = cadd(nup, g.gup);
    ndn = cadd(ndn, g.gdn);
  }
  for (int b = 0; b < a.nblk; ++b)
    for (int r = threadIdx.x; r < a.Mb[b]; r += blockDim.x) K = cadd(K, gsum(a.gk[b], w, a.ntile[b], a.Mb[b], r));
  V = block_sum_z(V, sh);
  R = block_sum_z(R, sh);
  K = block_sum_z(K, sh);
  nup = block_sum_z(nup, sh);
  ndn = block_sum_z(ndn, sh);
  if (threadIdx.x == 0) {
    zd d = a.den[w];
    out[0] = d;
    out[1] = cmul(cadd(cadd(K, V), R), d);
    out[2] = cmul(K, d);
    out[3] = cmul(V, d);
    out[4] = cmul(R, d);
    out[5] = cmul(nup, d);
    out[6] = cmul(ndn, d);
    out[7] = cadd(cadd(K, V), R);
  }
}

__global__ __launch_bounds__(256) void energy_reduce_kernel(int nw, const zd* per_walker, DevScalars* sc) {
  // fixed-order reduction over walkers: thread t sums a contiguous chunk, then the block sums in order
  __shared__ zd part[256][7];
  const int tid = threadIdx.x;
  const int chunk = (nw + 255) / 256;
  zd s[7];
  for (int k = 0; k < 7; ++k) s[k] = cmk<double>(0, 0);
  for (int w = tid * chunk; w < min(nw, (tid + 1) * chunk); ++w)
    for (int k = 0; k < 7; ++k) s[k] = cadd(s[k], per_walker[(long long)w * 8 + k]);
  for (int k = 0; k < 7; ++k) part[tid][k] = s[k];
  __syncthreads();
  if (tid == 0) {
    zd t[7];
    for (int k = 0; k < 7; ++k) t[k] = cmk<double>(0, 0);
    for (int i = 0; i < 256; ++i)
      for (int k = 0; k < 7; ++k) t[k] = cadd(t[k], part[i][k]);
    sc->den = cadd(sc->den, t[0]);
    sc->HNum = cadd(sc->HNum, t[1]);
    sc->KNum = cadd(sc->KNum, t[2]);
    sc->VNum = cadd(sc->VNum, t[3]);
    sc->RNum = cadd(sc->RNum, t[4]);
    sc->NupTot = cadd(sc->NupTot, t[5]);
    sc->NdnTot = cadd(sc->NdnTot, t[6]);
  }
}

// per-site densities weighted by den: walker chunks in parallel (blockIdx.y), then the chunk
// partials summed in chunk order, so the result does not depend on scheduling
__global__ void site_partial_kernel(EnergyArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.L) return;
  const int L = a.L;
  const int per = (a.nw + SITE_CHUNKS - 1) / SITE_CHUNKS;
  const int w0 = blockIdx.y * per, w1 = min(a.nw, w0 + per);
  zd su = cmk<double>(0, 0), sd = su, sp = su, sm = su;
  for (int w = w0; w < w1; ++w) {
    if (!a.alive[w]) continue;
    SiteG g = site_green(a.gd, a.go, a.ntile, a.Mb, a.dense, w, i, L);
    zd d = a.den[w];
    su = cadd(su, cmul(g.gup, d));
    sd = cadd(sd, cmul(g.gdn, d));
    sp = cadd(sp, cmul(g.off_i, d));
    sm = cadd(sm, cmul(g.off_iL, d));
  }
  zd* o = a.site_part + (long long)blockIdx.y * 4 * L;
  o[i] = su;
  o[L + i] = sd;
  o[2 * L + i] = sp;
  o[3 * L + i] = sm;
}

__global__ void site_final_kernel(EnergyArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 4 * a.L) return;
  zd s = cmk<double>(0, 0);
  for (int c = 0; c < SITE_CHUNKS; ++c) s = cadd(s, a.site_part[(long long)c * 4 * a.L + e]);
  a.site_acc[e] = cadd(a.site_acc[e], s);
}

// need[w] = alive && max_ij |S_ij - delta_ij| > tol (S = Q^H Q of walker w), skip[w] = alive && !need
template <typename T>
struct OrthoArgs {
  const cplx<T>* S;
  const int* alive;
  int* need;
  double tol;
  int N, pad;
};
template <typename T>
struct OrthoK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const OrthoArgs<T>& g);
};
template <typename T>
__device__ __forceinline__ void OrthoK<T>::run(const BlockIdx3& bk, const OrthoArgs<T>& g) {
  const int N = g.N;
  const cplx<T>* S = g.S;
  const int* alive = g.alive;
  const double tol = g.tol;
  int* need = g.need;

  const int w = bk.x;
  __shared__ double sh[8];
  const cplx<T>* s = S + (long long)w * N * N;
  double m = 0.0;
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int i = e % N, j = e / N;
    const zd v = to_zd(s[e]);
    const double d = fabs(v.re - (i == j ? 1.0 : 0.0)) + fabs(v.im);
    m = (d > m || d != d) ? d : m;  // NaN counts as not orthonormal
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double x = __shfl_xor(m, o, 64);
    m = (x > m || x != x) ? x : m;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) t = (sh[k] > t || sh[k] != sh[k]) ? sh[k] : t;
    const int a = alive ? alive[w] : 1;
    need[w] = a && !(t <= tol);
  }
}

struct QrFlagArgs {
  const int* alive;
  const int* info;
  int* flag;
  int nw, pad;
};
struct QrFlagK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const QrFlagArgs& g);
};
__device__ __forceinline__ void QrFlagK::run(const BlockIdx3& bk, const QrFlagArgs& g) {
  const int nw = g.nw;
  const int* alive = g.alive;
  const int* info = g.info;
  int* flag = g.flag;

  const int w = bk.x * blockDim.x + threadIdx.x;
  if (w < nw) flag[w] = (alive[w] && info[w]) ? 1 : 0;
}

// S_w += c tr(S_w) I for flagged walkers: one workgroup per walker
struct QrShiftArgs {
  zd* S;
  const int* flag;
  double c;
  int N, pad;
};
struct QrShiftK {
  static constexpr int NT = 64, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const QrShiftArgs& g);
};
__device__ __forceinline__ void QrShiftK::run(const BlockIdx3& bk, const QrShiftArgs& g) {
  const int N = g.N;
  zd* S = g.S;
  const int* flag = g.flag;
  const double c = g.c;

  const int w = bk.x;
  if (!flag[w]) return;
  zd* s = S + (long long)w * N * N;
  __shared__ double sh[4];
  double t = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) t += fabs(s[i + (long long)i * N].re);
  t = block_sum_d(t, sh);
  const double shift = c * t;
  for (int i = threadIdx.x; i < N; i += blockDim.x) s[i + (long long)i * N].re += shift;
}

// one workgroup: flagged walkers count as shifted, those whose shifted Cholesky broke down as well
// die (their Gram matrix is singular: a rank-deficient walker); the counts are added by one thread
struct QrKillArgs {
  int* flag;
  int* flag2;
  const int* info;
  int* alive;
  DevScalars* sc;
  int nw, count_shifted;
};
struct QrKillK {
  static constexpr int NT = 1024, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const QrKillArgs& g);
};
__device__ __forceinline__ void QrKillK::run(const BlockIdx3& bk, const QrKillArgs& g) {
  const int nw = g.nw, count_shifted = g.count_shifted;
  int* flag = g.flag;
  int* flag2 = g.flag2;
  const int* info = g.info;
  int* alive = g.alive;
  DevScalars* sc = g.sc;

  __shared__ double sh[16];
  double nshift = 0.0, nkill = 0.0;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    if (!flag[w]) continue;
    nshift += 1.0;
    if (info[w]) {
      alive[w] = 0;
      flag[w] = 0;
      if (flag2) flag2[w] = 0;
      nkill += 1.0;
    }
  }
  nshift = block_sum_d(nshift, sh);
  nkill = block_sum_d(nkill, sh);
  if (threadIdx.x == 0) {
    if (count_shifted) sc->qr_shifted += (unsigned long long)nshift;
    sc->qr_killed += (unsigned long long)nkill;
  }
}

__global__ void set_et_kernel(DevScalars* sc) {
  zd e = cdiv(sc->HNum, sc->den);
  sc->ET = e.re;
}

struct LogwAddArgs {
  zd* logw;
  const int* alive;
  const zd* add0;
  const zd* add1;
  double m0, m1;
  int nw, pad;
};
struct LogwAddK {
  static constexpr int NT = 256, WPE = 1;
  __device__ static void run(const BlockIdx3& bk, const LogwAddArgs& g);
};
__device__ __forceinline__ void LogwAddK::run(const BlockIdx3& bk, const LogwAddArgs& g) {
  const int nw = g.nw;
  zd* logw = g.logw;
  const int* alive = g.alive;
  const zd* add0 = g.add0;
  const zd* add1 = g.add1;
  const double m0 = g.m0, m1 = g.m1;

  const int w = bk.x * blockDim.x + threadIdx.x;
  if (w >= nw || (alive && !alive[w])) return;
  double d = m0 * add0[w].re + m1 * add0[w].im;
  if (add1) d += m0 * add1[w].re + m1 * add1[w].im;
  logw[w].re += d;
}

// ------------------------------------------------------------------ back-propagation
template <typename T>
__global__ __launch_bounds__(256) void bp_field_kernel(BpFieldArgs a) {
  const int w = blockIdx.x;
  __shared__ int anc_sh;
  if (threadIdx.x == 0) {
    int p = a.anc[w];
    if (a.table) p = a.table[p];
    if (p < 0 || p >= a.nw) p = w;  // a comb over zero total weight leaves no valid parent
    a.anc[w] = p;
    anc_sh = p;
  }
  __syncthreads();
  const int L = a.L;
  const int* x = a.aux + (long long)anc_sh * L;
  typedef cplx<T> ct;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    zd d00, d01, d10, d11;
    hs_site_matrix(a.decomp, cscale(a.gamma[i], (double)x[i]), a.cdiag, L, i, d00, d01, d10, d11);
    // applyToLeft multiplies by D^H: [conj d00, conj d10; conj d01, conj d11]
    const zd e00 = cconj(d00), e01 = cconj(d10), e10 = cconj(d01), e11 = cconj(d11);
    if (a.diag_only) {
      if (a.dense) {
        ct* s0 = reinterpret_cast<ct*>(a.scale[0]) + (long long)w * a.Mb[0];
        s0[i] = from_zd<T>(e00);
        s0[i + L] = from_zd<T>(e11);
      } else {
        reinterpret_cast<ct*>(a.scale[0])[(long long)w * a.Mb[0] + i] = from_zd<T>(e00);
        reinterpret_cast<ct*>(a.scale[1])[(long long)w * a.Mb[1] + i] = from_zd<T>(e11);
      }
    } else {
      ct* d = reinterpret_cast<ct*>(a.d2x2) + (long long)w * 4 * L;
      d[i] = from_zd<T>(e00);
      d[L + i] = from_zd<T>(e01);
      d[2 * L + i] = from_zd<T>(e10);
      d[3 * L + i] = from_zd<T>(e11);
    }
  }
}

__global__ void iota_kernel(int n, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

template <typename E>
__global__ void gather_walkers_kernel(long long n, const int* idx, const E* __restrict__ src, E* __restrict__ dst) {
  const int w = blockIdx.y;
  const E* s = src + (idx ? (long long)idx[w] * n : 0);
  E* d = dst + (long long)w * n;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    d[k] = s[k];
}

template <typename T>
__global__ __launch_bounds__(256) void conj_transpose_kernel(int M, int N, long long sstride, long long dstride,
                                                             const int* alive, const cplx<T>* __restrict__ src,
                                                             cplx<T>* __restrict__ dst) {
  const int w = blockIdx.z;
  if (alive && !alive[w]) return;
  __shared__ cplx<T> t[32][33];
  const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const cplx<T>* sp = src + (long long)w * sstride;
  for (int r = ty; r < 32; r += 8) {
    const int m = m0 + tx, n = n0 + r;
    if (m < M && n < N) t[r][tx] = sp[m + (long long)n * M];
  }
  __syncthreads();
  cplx<T>* dp = dst + (long long)w * dstride;
  for (int r = ty; r < 32; r += 8) {
    const int n = n0 + tx, m = m0 + r;
    if (m < M && n < N) dp[n + (long long)m * N] = cconj(t[tx][r]);
  }
}

// G(r, c) of one walker: dense G^T with leading dimension ld0, or the up-up / down-down blocks
template <typename T>
__device__ __forceinline__ zd wick_g(const cplx<T>* g0, const cplx<T>* g1, int dense, int ld0, int L, int r, int c) {
  if (dense) return to_zd(g0[c + (long long)r * ld0]);
  if (r < L && c < L) return to_zd(g0[c + (long long)r * L]);
  if (r >= L && c >= L) return to_zd(g1[(c - L) + (long long)(r - L) * L]);
  return cmk<double>(0, 0);
}

__device__ __forceinline__ void cfma_acc(zd& acc, zd w, zd v) { acc = cadd(acc, cmul(w, v)); }

template <typename T>
__global__ __launch_bounds__(256) void wick_kernel(WickArgs a) {
  const int L = a.L;
  const int ntI = (L + 15) / 16;
  const int I0 = (blockIdx.x % ntI) * 16, J0 = (blockIdx.x / ntI) * 16;
  const int tj = threadIdx.x & 15, ti = threadIdx.x >> 4;
  const int i = I0 + ti, j = J0 + tj;
  const bool act = i < L && j < L;
  // tr[q][r][c] = G(J0 + r + sj L, I0 + c + si L) with q = 2 sj + si
  __shared__ zd tr[4][16][17];
  // dg[side][q][k]: side 0 the i rows, 1 the j rows; q: G(k,k), G(k+L,k+L), G(k,k+L), G(k+L,k)
  __shared__ zd dg[2][4][16];
  zd acc[WICK_Q];
#pragma unroll
  for (int q = 0; q < WICK_Q; ++q) acc[q] = cmk<double>(0, 0);
  const zd zero = cmk<double>(0, 0);
  typedef cplx<T> ct;
  for (int w = a.c0 + (int)blockIdx.y; w < a.c0 + a.cn; w += a.slices) {
    if (!a.alive[w]) continue;
    const zd wt = a.den[w];
    const ct* g0 = reinterpret_cast<const ct*>(a.gt[0]) + (long long)(w - a.c0) * a.gstride[0];
    const ct* g1 = a.dense ? nullptr : reinterpret_cast<const ct*>(a.gt[1]) + (long long)(w - a.c0) * a.gstride[1];
    __syncthreads();
    {
      const int jj = J0 + ti, ii = I0 + tj;
      const bool ok = jj < L && ii < L;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int sj = q >> 1, si = q & 1;
        tr[q][ti][tj] = ok ? wick_g<T>(g0, g1, a.dense, a.ld[0], L, jj + sj * L, ii + si * L) : zero;
      }
    }
    if (threadIdx.x < 128) {
      const int side = threadIdx.x >> 6, q = (threadIdx.x >> 4) & 3, k = threadIdx.x & 15;
      const int kk = (side ? J0 : I0) + k;
      const int ro = (q == 1 || q == 3) ? L : 0, co = (q == 1 || q == 2) ? L : 0;
      dg[side][q][k] = kk < L ? wick_g<T>(g0, g1, a.dense, a.ld[0], L, kk + ro, kk + co) : zero;
    }
    zd gUU = zero, gUD = zero, gDU = zero, gDD = zero;
    if (act) {
      gUU = wick_g<T>(g0, g1, a.dense, a.ld[0], L, i, j);
      gDD = wick_g<T>(g0, g1, a.dense, a.ld[0], L, i + L, j + L);
      if (a.dense) {
        gUD = wick_g<T>(g0, g1, 1, a.ld[0], L, i, j + L);
        gDU = wick_g<T>(g0, g1, 1, a.ld[0], L, i + L, j);
      }
    }
    __syncthreads();
    if (!act) continue;
    const zd tUU = tr[0][tj][ti], tUD = tr[1][tj][ti], tDU = tr[2][tj][ti], tDD = tr[3][tj][ti];
    const zd diU = dg[0][0][ti], diD = dg[0][1][ti], diUD = dg[0][2][ti], diDU = dg[0][3][ti];
    const zd djU = dg[1][0][tj], djD = dg[1][1][tj], djUD = dg[1][2][tj], djDU = dg[1][3][tj];
    cfma_acc(acc[0], wt, gUU);
    cfma_acc(acc[1], wt, gUD);
    cfma_acc(acc[2], wt, gDU);
    cfma_acc(acc[3], wt, gDD);
    cfma_acc(acc[5], wt, csub(cmul(diU, djD), cmul(gUD, tDU)));
    cfma_acc(acc[6], wt, csub(cmul(diD, djU), cmul(gDU, tUD)));
    cfma_acc(acc[10], wt, csub(cmul(gUU, gDD), cmul(gUD, gDU)));
    if (i == j) {
      const zd det = csub(cmul(diU, diD), cmul(diUD, diDU));
      cfma_acc(acc[4], wt, diU);
      cfma_acc(acc[7], wt, diD);
      cfma_acc(acc[8], wt, csub(diU, det));
      cfma_acc(acc[9], wt, csub(diD, det));
    } else {
      cfma_acc(acc[4], wt, csub(cmul(diU, djU), cmul(gUU, tUU)));
      cfma_acc(acc[7], wt, csub(cmul(diD, djD), cmul(gDD, tDD)));
      cfma_acc(acc[8], wt, csub(cmul(diUD, djDU), cmul(gUU, tDD)));
      cfma_acc(acc[9], wt, csub(cmul(diDU, djUD), cmul(gDD, tUU)));
    }
  }
  if (!act) return;
  const long long LL = (long long)L * L, pr = i + (long long)j * L;
  zd* out = a.part + (long long)blockIdx.y * WICK_Q * LL + pr;
#pragma unroll
  for (int q = 0; q < WICK_Q; ++q) out[q * LL] = a.first ? acc[q] : cadd(out[q * LL], acc[q]);
}

__global__ void wick_reduce_kernel(int L, int slices, const zd* part, zd* green, zd* dd, zd* pm, zd* mp, zd* sp) {
  const long long LL = (long long)L * L;
  const long long pr = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= LL) return;
  const int i = (int)(pr % L), j = (int)(pr / L);
  zd v[WICK_Q];
  for (int q = 0; q < WICK_Q; ++q) {
    zd t = cmk<double>(0, 0);
    for (int y = 0; y < slices; ++y) t = cadd(t, part[((long long)y * WICK_Q + q) * LL + pr]);
    v[q] = t;
  }
  const long long L2 = 2 * (long long)L;
  const long long at[4] = {i + j * L2, i + (j + L) * L2, (i + L) + j * L2, (i + L) + (j + L) * L2};
  for (int s = 0; s < 4; ++s) {
    green[at[s]] = cadd(green[at[s]], v[s]);
    dd[at[s]] = cadd(dd[at[s]], v[4 + s]);
  }
  pm[pr] = cadd(pm[pr], v[8]);
  mp[pr] = cadd(mp[pr], v[9]);
  sp[pr] = cadd(sp[pr], v[10]);
}

__global__ void den_sum_kernel(int nw, const zd* den, const int* alive, zd* acc) {
  __shared__ double sh[16];
  zd t = cmk<double>(0, 0);
  for (int w = threadIdx.x; w < nw; w += blockDim.x)
    if (alive[w]) t = cadd(t, den[w]);
  t = block_sum_z(t, sh);
  if (threadIdx.x == 0) acc[0] = cadd(acc[0], t);
}

}  // namespace

int launch_logo(const LogOArgs& a, hipStream_t s) {
  if (a.nw <= 0) return 0;
  return glaunch<LogoK>(a, dim3((a.nw + 255) / 256), dim3(256), 0, s);
}

int launch_alive_scan(int nw, const int* alive, int* prefix, DevScalars* sc, hipStream_t s) {
  AliveScanArgs g;
  memset(&g, 0, sizeof(g));
  g.alive = alive; g.prefix = prefix; g.sc = sc; g.nw = nw;
  return glaunch<AliveScanK>(g, dim3(1), dim3(1024), 0, s);
}

int launch_hs_sample(const HsSampleArgs& a, hipStream_t s) {
  if (a.fp32) return glaunch<HsSampleK<float>>(a, dim3(a.nw), dim3(256), 0, s);
  return glaunch<HsSampleK<double>>(a, dim3(a.nw), dim3(256), 0, s);
}

int launch_advance_counter(DevScalars* sc, int per_walker, hipStream_t s) {
  AdvanceArgs g;
  memset(&g, 0, sizeof(g));
  g.sc = sc; g.per_walker = per_walker;
  return glaunch<AdvanceK>(g, dim3(1), dim3(1), 0, s);
}

template <typename T>
int launch_hs_apply_2x2(int nw, int L, int N, const void* in, void* out, const void* d2x2, hipStream_t s) {
  if ((long long)nw * L * N == 0) return 0;
  // 16-byte vectors of two FP32 complex numbers need an even L (every row offset then even)
  const bool pair = std::is_same<T, float>::value && L % 2 == 0;
  const int S = pair ? 2 : 1;
  const size_t lds = (size_t)4 * L * sizeof(cplx<T>);
  // about 64 site vectors per thread per workgroup, in column chunks of equal width
  const int nv = L / S;
  const int target = std::max(1, std::min(N, (64 * 256) / std::max(nv, 1)));
  const int chunks = (N + target - 1) / target;
  const int cpb = (N + chunks - 1) / chunks;
  dim3 grid((unsigned)((N + cpb - 1) / cpb), (unsigned)nw);
  HsApplyArgs<T> g;
  memset(&g, 0, sizeof(g));
  g.in = reinterpret_cast<const cplx<T>*>(in);
  g.d = reinterpret_cast<const cplx<T>*>(d2x2);
  g.out = reinterpret_cast<cplx<T>*>(out);
  g.L = L; g.N = N; g.cpb = cpb;
  if constexpr (std::is_same<T, float>::value) {
    if (pair) return glaunch<HsApplyK<T, 2>>(g, grid, dim3(256), lds, s);
  }
  return glaunch<HsApplyK<T, 1>>(g, grid, dim3(256), lds, s);
}
template int launch_hs_apply_2x2<double>(int, int, int, const void*, void*, const void*, hipStream_t);
template int launch_hs_apply_2x2<float>(int, int, int, const void*, void*, const void*, hipStream_t);

int launch_comb(int nw, const double* weight, const double* ubuf, DevScalars* sc, int* table, double* prob,
                int* number, hipStream_t s) {
  CombArgs g;
  memset(&g, 0, sizeof(g));
  g.weight = weight; g.ubuf = ubuf; g.sc = sc; g.table = table; g.prob = prob; g.number = number; g.L = nw;
  if (nw <= COMB_LDS_MAX) return glaunch<CombLdsK>(g, dim3(1), dim3(1024), 0, s);
  return glaunch<CombK>(g, dim3(1), dim3(64), 0, s);
}

template <typename T>
int launch_apply_table(int nw, int nblk, void* const* wf, const long long* wsize, zd* logw, int* alive,
                       const int* table, hipStream_t s) {
  ApplyTableArgs<T> g;
  memset(&g, 0, sizeof(g));
  g.wf0 = reinterpret_cast<cplx<T>*>(wf[0]);
  g.wf1 = nblk > 1 ? reinterpret_cast<cplx<T>*>(wf[1]) : nullptr;
  g.ws0 = wsize[0]; g.ws1 = nblk > 1 ? wsize[1] : 0;
  g.logw = logw; g.alive = alive; g.table = table; g.nw = nw;
  return glaunch<ApplyTableK<T>>(g, dim3(8, nw), dim3(256), 0, s);
}
template int launch_apply_table<double>(int, int, void* const*, const long long*, zd*, int*, const int*, hipStream_t);
template int launch_apply_table<float>(int, int, void* const*, const long long*, zd*, int*, const int*, hipStream_t);

int launch_energy(const EnergyArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(energy_kernel, dim3(a.nw), dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(energy_reduce_kernel, dim3(1), dim3(256), 0, s, a.nw, a.per_walker, a.sc);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(site_partial_kernel, dim3((a.L + 63) / 64, SITE_CHUNKS), dim3(64), 0, s, a);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(site_final_kernel, dim3((4 * a.L + 255) / 256), dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_ortho_check(int nw, int N, const zd* S, const int* alive, double tol, int* need, hipStream_t s) {
  if (nw <= 0) return 0;
  OrthoArgs<double> g;
  memset(&g, 0, sizeof(g));
  g.S = S; g.alive = alive; g.need = need; g.tol = tol; g.N = N;
  return glaunch<OrthoK<double>>(g, dim3(nw), dim3(256), 0, s);
}
int launch_ortho_check_f32(int nw, int N, const zf* S, const int* alive, double tol, int* need, hipStream_t s) {
  if (nw <= 0) return 0;
  OrthoArgs<float> g;
  memset(&g, 0, sizeof(g));
  g.S = S; g.alive = alive; g.need = need; g.tol = tol; g.N = N;
  return glaunch<OrthoK<float>>(g, dim3(nw), dim3(256), 0, s);
}

int launch_qr_flag(int nw, const int* alive, const int* info, int* flag, hipStream_t s) {
  if (nw <= 0) return 0;
  QrFlagArgs g;
  memset(&g, 0, sizeof(g));
  g.alive = alive; g.info = info; g.flag = flag; g.nw = nw;
  return glaunch<QrFlagK>(g, dim3((nw + 255) / 256), dim3(256), 0, s);
}

int launch_qr_shift(int nw, int N, zd* S, const int* flag, double c, hipStream_t s) {
  if (nw <= 0) return 0;
  QrShiftArgs g;
  memset(&g, 0, sizeof(g));
  g.S = S; g.flag = flag; g.c = c; g.N = N;
  return glaunch<QrShiftK>(g, dim3(nw), dim3(64), 0, s);
}

int launch_qr_kill(int nw, int* flag, int* flag2, const int* info, int* alive, DevScalars* sc,
                   int count_shifted, hipStream_t s) {
  if (nw <= 0) return 0;
  QrKillArgs g;
  memset(&g, 0, sizeof(g));
  g.flag = flag; g.flag2 = flag2; g.info = info; g.alive = alive; g.sc = sc; g.nw = nw;
  g.count_shifted = count_shifted;
  return glaunch<QrKillK>(g, dim3(1), dim3(1024), 0, s);
}

int launch_set_et_from_energy(DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(set_et_kernel, dim3(1), dim3(1), 0, s, sc);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_logw_add_real(int nw, zd* logw, const int* alive, const zd* add0, const zd* add1, int /*unused*/,
                         double mult0, double mult1, hipStream_t s) {
  LogwAddArgs g;
  memset(&g, 0, sizeof(g));
  g.logw = logw; g.alive = alive; g.add0 = add0; g.add1 = add1; g.m0 = mult0; g.m1 = mult1; g.nw = nw;
  return glaunch<LogwAddK>(g, dim3((nw + 255) / 256), dim3(256), 0, s);
}

int launch_bp_field(const BpFieldArgs& a, hipStream_t s) {
  if (a.nw <= 0) return 0;
  if (a.fp32) hipLaunchKernelGGL(bp_field_kernel<float>, dim3(a.nw), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(bp_field_kernel<double>, dim3(a.nw), dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_iota(int n, int* out, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, out);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_gather_walkers(int nw, long long n, size_t esz, const int* idx, const void* src, void* dst, hipStream_t s) {
  if (nw <= 0 || n <= 0) return 0;
  dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 256), nw);
  if (esz == 16)
    hipLaunchKernelGGL(gather_walkers_kernel<uint4>, grid, dim3(256), 0, s, n, idx, (const uint4*)src, (uint4*)dst);
  else
    hipLaunchKernelGGL(gather_walkers_kernel<uint2>, grid, dim3(256), 0, s, n, idx, (const uint2*)src, (uint2*)dst);
  HIP_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_conj_transpose(int nw, int M, int N, long long sstride, long long dstride, const int* alive,
                          const void* src, void* dst, hipStream_t s) {
  if (nw <= 0 || M <= 0 || N <= 0) return 0;
  dim3 grid((M + 31) / 32, (N + 31) / 32, nw);
  hipLaunchKernelGGL(conj_transpose_kernel<T>, grid, dim3(256), 0, s, M, N, sstride, dstride, alive,
                     reinterpret_cast<const cplx<T>*>(src), reinterpret_cast<cplx<T>*>(dst));
  HIP_LAUNCH_CHECK();
  return 0;
}
template int launch_conj_transpose<double>(int, int, int, long long, long long, const int*, const void*, void*,
                                           hipStream_t);
template int launch_conj_transpose<float>(int, int, int, long long, long long, const int*, const void*, void*,
                                          hipStream_t);

int launch_wick(const WickArgs& a, hipStream_t s) {
  const int nt = (a.L + 15) / 16;
  dim3 grid(nt * nt, a.slices);
  if (a.fp32) hipLaunchKernelGGL(wick_kernel<float>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(wick_kernel<double>, grid, dim3(256), 0, s, a);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_wick_reduce(int L, int slices, const zd* part, zd* acc_green, zd* acc_dd, zd* acc_pm, zd* acc_mp,
                       zd* acc_sp, hipStream_t s) {
  const long long LL = (long long)L * L;
  hipLaunchKernelGGL(wick_reduce_kernel, dim3((unsigned)((LL + 255) / 256)), dim3(256), 0, s, L, slices, part,
                     acc_green, acc_dd, acc_pm, acc_mp, acc_sp);
  HIP_LAUNCH_CHECK();
  return 0;
}

int launch_den_sum(int nw, const zd* den, const int* alive, zd* acc, hipStream_t s) {
  hipLaunchKernelGGL(den_sum_kernel, dim3(1), dim3(256), 0, s, nw, den, alive, acc);
  HIP_LAUNCH_CHECK();
  return 0;
}
