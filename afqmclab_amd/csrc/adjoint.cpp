// Fixed-field propagation of constraint-path walkers and its hand-written adjoint (reverse mode).
//
// Forward, for s = 0 .. nsteps-1 and each spin block b (sigma = +1 up, -1 down):
//   Y_s = D_s W_s,  D_s = diag(exp(sigma gamma_i x_{s,w,i}) constDiag_b,i),  W_{s+1} = B Y_s,
//   B = exp(-dt K_b); every mgs steps W = Q R (CholeskyQR2) with logw += log det R.
// Energy: E = sum_w O_w E_L,w / sum_w O_w, O_w = exp(logw_w) prod_b det(Phi_b^H W_b),
//   E_L = sum_b tr(K_b M_b) + sum_i U_i M_up,ii M_dn,ii, M_b = theta_b Phi_b^H, theta_b = W_b O_b^{-1}
//   (the local energy of HubbardSOCMeasureFixedSDSD.cpp:178-218 in spin-block form).
// Reverse, in the frame of the normalised walkers (W_true,s = Q_s T_s with T_s the product of the
// R factors so far, adjoint Wt_s = dE/dW_true,s T_s^T):
//   Wt_n = c_w [ (E_L,w - E) conj(Phi) + P - conj(Phi) theta^T P ] O^{-T},
//          P = K^T conj(Phi) + diag(U_i M_other,ii) conj(Phi),  c_w = O_w / sum O;
//   at a QR after step s: Wt <- Wt R^{-T};  Ybar = B^T Wt_{s+1};  Wt_s = D_s Ybar;
//   Bbar += Wt_{s+1} (D_s Q_s)^T;  Ubar_i += sum_c Ybar(i,c) Q_s(i,c) D_ii d log D_ii / d U_i.
// The forward states are checkpointed every ckpt steps and each segment is recomputed (bitwise
// identically) during the reverse sweep.
#include "cp_internal.hpp"

using namespace cpi;

namespace {

__device__ __forceinline__ double field(unsigned long long key, int s, int nw, int L, int w, int i) {
  const unsigned long long k = ((unsigned long long)s * nw + w) * (unsigned long long)L + i;
  return afqmc_uniform_dev(key, k) < 0.5 ? -1.0 : 1.0;
}

// scale[b][w][i] = exp(sigma_b gamma_i x) constDiag_b,i
__global__ void ff_scale_kernel(int s, unsigned long long key, int nw, int L, const zd* gamma, const zd* cdiag,
                                zd* up, zd* dn) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nw * L) return;
  const int w = (int)(e / L), i = (int)(e % L);
  const double x = field(key, s, nw, L, w, i);
  const zd gx = cscale(gamma[i], x);
  up[e] = cmul(cexp_d(gx), cdiag[i]);
  dn[e] = cmul(cexp_d(cmk<double>(-gx.re, -gx.im)), cdiag[3 * L + i]);
}

// out[w][c][i] = scale[w][i] * in[w][c][i]
__global__ void rowscale_kernel(long long total, int M, int N, const zd* scale, const zd* in, zd* out) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long per = (long long)M * N;
    const long long w = e / per;
    const int i = (int)(e % M);
    out[e] = cmul(scale[w * M + i], in[e]);
  }
}

// per walker: out (N x M) = in (M x N)^T
__global__ void transpose_kernel(long long total, int M, int N, const zd* in, zd* out) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long per = (long long)M * N;
    const long long w = e / per;
    const long long r = e % per;
    const int i = (int)(r % M), c = (int)(r / M);
    out[w * per + c + (long long)i * N] = in[e];
  }
}

// part[w][i] = sum_c Ybar(i,c) Q(i,c) * D(w,i) * dlogD(w,i)
__global__ void ubar_partial_kernel(int s, unsigned long long key, int nw, int M, int N, double sigma,
                                    const zd* gprime, double dt, const zd* scale, const zd* Yb, const zd* Q, zd* part) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (i >= M) return;
  const long long base = (long long)w * M * N;
  zd acc = cmk<double>(0, 0);
  for (int c = 0; c < N; ++c) acc = cadd(acc, cmul(Yb[base + i + (long long)c * M], Q[base + i + (long long)c * M]));
  const double x = field(key, s, nw, M, w, i);
  const zd dl = cmk<double>(sigma * x * gprime[i].re - 0.5 * dt, sigma * x * gprime[i].im);
  part[(long long)w * M + i] = cmul(cmul(acc, scale[(long long)w * M + i]), dl);
}

// out[i] += sum_w part[w][i] (walker order)
__global__ void sum_walkers_kernel(int nw, int M, const zd* part, zd* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  zd s = cmk<double>(0, 0);
  for (int w = 0; w < nw; ++w) s = cadd(s, part[(long long)w * M + i]);
  out[i] = cadd(out[i], s);
}

// out[e] = sum_w c_w * in[w][e] (e < per), walker order
__global__ void weighted_sum_kernel(int nw, long long per, const zd* c, const zd* in, zd* out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per) return;
  zd s = cmk<double>(0, 0);
  for (int w = 0; w < nw; ++w) s = cadd(s, cmul(c[w], in[(long long)w * per + e]));
  out[e] = s;
}

// out[e] = sum_w in[w][e]
__global__ void plain_sum_kernel(int nw, long long per, const zd* in, zd* out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per) return;
  zd s = cmk<double>(0, 0);
  for (int w = 0; w < nw; ++w) s = cadd(s, in[(long long)w * per + e]);
  out[e] = s;
}

// m[w][i] = sum_j theta(i,j) conj(Phi(i,j));  kin[w] (+)= sum_ij theta(i,j) P2(i,j)
__global__ void mdiag_kernel(int nw, int M, int N, const zd* theta, const zd* phi, zd* m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = blockIdx.y;
  if (i >= M) return;
  const long long base = (long long)w * M * N;
  zd s = cmk<double>(0, 0);
  for (int j = 0; j < N; ++j) s = cadd(s, cmul(theta[base + i + (long long)j * M], cconj(phi[i + (long long)j * M])));
  m[(long long)w * M + i] = s;
}

__global__ void kin_kernel(int M, int N, const zd* theta, const zd* p2, zd* kin, int accumulate) {
  const int w = blockIdx.x;
  __shared__ double sr[256], si[256];
  const long long base = (long long)w * M * N;
  zd s = cmk<double>(0, 0);
  for (long long e = threadIdx.x; e < (long long)M * N; e += blockDim.x) s = cadd(s, cmul(theta[base + e], p2[e]));
  sr[threadIdx.x] = s.re;
  si[threadIdx.x] = s.im;
  __syncthreads();
  if (threadIdx.x == 0) {
    zd t = cmk<double>(0, 0);
    for (int k = 0; k < (int)blockDim.x; ++k) t = cadd(t, cmk<double>(sr[k], si[k]));
    kin[w] = accumulate ? cadd(kin[w], t) : t;
  }
}

// per walker: E_L, logO; then (one thread, walker order) the weights c_w and E
__global__ void weights_kernel(int nw, int L, const double* U, const zd* mup, const zd* mdn, const zd* kin,
                               const zd* logw, const zd* ld0, const zd* ld1, zd* el, zd* logO, zd* c, zd* E) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mx = -1e300;
  for (int w = 0; w < nw; ++w) {
    zd v = kin[w];
    for (int i = 0; i < L; ++i) v = cadd(v, cscale(cmul(mup[(long long)w * L + i], mdn[(long long)w * L + i]), U[i]));
    el[w] = v;
    logO[w] = cadd(logw[w], cadd(ld0[w], ld1[w]));
    if (logO[w].re > mx) mx = logO[w].re;
  }
  zd den = cmk<double>(0, 0), num = den;
  for (int w = 0; w < nw; ++w) {
    const zd o = cexp_d(cmk<double>(logO[w].re - mx, logO[w].im));
    c[w] = o;
    den = cadd(den, o);
    num = cadd(num, cmul(o, el[w]));
  }
  const zd inv = cinv(den);
  for (int w = 0; w < nw; ++w) c[w] = cmul(c[w], inv);
  *E = cmul(num, inv);
}

// P[w](i,j) = wdkT(i,j) + U_i m_other[w][i] conj(Phi(i,j))
__global__ void pmat_kernel(long long total, int M, int N, const zd* wdkT, const double* U, const zd* mother,
                            const zd* phi, zd* P) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long per = (long long)M * N;
    const long long w = e / per, r = e % per;
    const int i = (int)(r % M);
    P[e] = cadd(wdkT[r], cmul(cscale(mother[w * M + i], U[i]), cconj(phi[r])));
  }
}

// Z[w] = c_w ((E_L,w - E) conj(Phi) + P[w])
__global__ void zmat_kernel(long long total, int M, int N, const zd* c, const zd* el, const zd* E, const zd* phi,
                            zd* P) {
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
    const long long per = (long long)M * N;
    const long long w = e / per, r = e % per;
    const zd d = csub(el[w], *E);
    P[e] = cmul(c[w], cadd(cmul(d, cconj(phi[r])), P[e]));
  }
}

__global__ void conj_kernel(long long n, const zd* in, zd* out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out[e] = cconj(in[e]);
}

// u[i] = sum_w c_w mup[w][i] mdn[w][i]
__global__ void udirect_kernel(int nw, int L, const zd* c, const zd* mup, const zd* mdn, zd* u) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  zd s = cmk<double>(0, 0);
  for (int w = 0; w < nw; ++w) s = cadd(s, cmul(c[w], cmul(mup[(long long)w * L + i], mdn[(long long)w * L + i])));
  u[i] = cadd(u[i], s);
}

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 65536); }

// Device scratch of one fixed-field run (spin blocks, FP64).
struct FF {
  afqmc_cp cp;
  afqmc_walkers w;
  int nw, L, nsteps, mgs, ckpt, nseg;
  unsigned long long key;
  long long ws[2];
  DevBuf<zd> ckw[2], ckl;       // checkpoints: walkers, logw
  DevBuf<zd> seg[2];            // Q_s of the current segment
  DevBuf<zd> rinv[2];           // R^{-1} of both CholeskyQR passes of each QR in the segment
  DevBuf<zd> init_logw;
  DevBuf<zd> wt[2], t1[2], t2[2], theta[2], P[2], conjphi[2], oinvT[2], S[2], bb[2];
  DevBuf<zd> m[2], kin, el, logO, c, E, ubar, part, gprime;
  int nrq() const { return ckpt / mgs + 1; }
};

int ff_alloc(FF& f) {
  afqmc_cp cp = f.cp;
  int rc = 0;
  const int nseg_states = f.ckpt;
  for (int b = 0; b < 2 && !rc; ++b) {
    const long long n = f.ws[b] * f.nw;
    const long long Nb = cp->Nb[b];
    rc = f.ckw[b].alloc((size_t)(n * f.nseg));
    if (!rc) rc = f.seg[b].alloc((size_t)(n * nseg_states));
    if (!rc) rc = f.rinv[b].alloc((size_t)(Nb * Nb * f.nw * 2 * f.nrq()));
    if (!rc) rc = f.wt[b].alloc((size_t)n);
    if (!rc) rc = f.t1[b].alloc((size_t)n);
    if (!rc) rc = f.t2[b].alloc((size_t)n);
    if (!rc) rc = f.theta[b].alloc((size_t)n);
    if (!rc) rc = f.P[b].alloc((size_t)n);
    if (!rc) rc = f.conjphi[b].alloc((size_t)f.ws[b]);
    if (!rc) rc = f.oinvT[b].alloc((size_t)(Nb * Nb * f.nw));
    if (!rc) rc = f.S[b].alloc((size_t)(Nb * Nb * f.nw));
    if (!rc) rc = f.bb[b].alloc((size_t)((long long)f.L * f.L * f.nw));
    if (!rc) rc = f.m[b].alloc((size_t)((long long)f.L * f.nw));
  }
  if (!rc) rc = f.ckl.alloc((size_t)f.nw * f.nseg);
  if (!rc) rc = f.init_logw.alloc((size_t)f.nw);
  if (!rc) rc = f.kin.alloc(f.nw);
  if (!rc) rc = f.el.alloc(f.nw);
  if (!rc) rc = f.logO.alloc(f.nw);
  if (!rc) rc = f.c.alloc(f.nw);
  if (!rc) rc = f.E.alloc(1);
  if (!rc) rc = f.ubar.alloc(f.L);
  if (!rc) rc = f.part.alloc((size_t)f.nw * f.L);
  if (!rc) rc = f.gprime.alloc(f.L);
  return rc;
}

GemmArgs gb() {
  GemmArgs a;
  memset(&a, 0, sizeof(a));
  a.alpha_re = 1.0;
  a.scale_group = 1;
  return a;
}

int ff_scales(FF& f, int s) {
  afqmc_cp cp = f.cp;
  hipLaunchKernelGGL(ff_scale_kernel, dim3(grid_for((long long)f.nw * f.L)), dim3(256), 0, cp->ctx->stream, s, f.key,
                     f.nw, f.L, cp->gamma.p, cp->cdiag.p, (zd*)cp->scale[0].p, (zd*)cp->scale[1].p);
  HIP_LAUNCH_CHECK();
  return 0;
}

// CholeskyQR2 of both blocks in place (wf <-> tmp), keeping each pass's R^{-1} in slot q if q >= 0
int ff_qr(FF& f, int q) {
  afqmc_cp cp = f.cp;
  afqmc_walkers w = f.w;
  hipStream_t s = cp->ctx->stream;
  int rc;
  for (int pass = 0; pass < 2; ++pass) {
    for (int b = 0; b < 2; ++b) {
      const int Mb = cp->Mb[b], Nb = cp->Nb[b];
      GemmArgs a = gb();
      a.M = Nb; a.N = Nb; a.K = Mb; a.batch = f.nw;
      a.A = w->wf[b].p; a.lda = Mb; a.strideA = f.ws[b];
      a.B = w->wf[b].p; a.ldb = Mb; a.strideB = f.ws[b];
      a.C = cp->qrS[b].p; a.ldc = Nb; a.strideC = (long long)Nb * Nb;
      if ((rc = launch_cgemm<double>(OP_C, EPI_STORE, a, s))) return rc;
      SmallDenseArgs c;
      memset(&c, 0, sizeof(c));
      c.N = Nb; c.batch = f.nw; c.in = cp->qrS[b].p; c.ld_in = Nb; c.stride_in = (long long)Nb * Nb;
      c.out = cp->qrS[b].p; c.ld_out = Nb; c.stride_out = c.stride_in;
      c.logdet = cp->qrld[b].p; c.info = cp->qrinfo[b].p; c.nsub = Nb; c.alive_div = 1;
      if ((rc = launch_chol_inverse<double>(c, s))) return rc;
      if (q >= 0) {
        const size_t bytes = sizeof(zd) * (size_t)Nb * Nb * f.nw;
        zd* dst = f.rinv[b].p + ((size_t)q * 2 + pass) * (size_t)Nb * Nb * f.nw;
        HIP_TRY(hipMemcpyAsync(dst, cp->qrS[b].p, bytes, hipMemcpyDeviceToDevice, s));
      }
      GemmArgs g = gb();
      g.M = Mb; g.N = Nb; g.K = Nb; g.batch = f.nw;
      g.A = w->wf[b].p; g.lda = Mb; g.strideA = f.ws[b];
      g.B = cp->qrS[b].p; g.ldb = Nb; g.strideB = (long long)Nb * Nb;
      g.C = w->tmp[b].p; g.ldc = Mb; g.strideC = f.ws[b];
      if ((rc = launch_cgemm<double>(OP_N, EPI_STORE, g, s))) return rc;
      swap(w->wf[b], w->tmp[b]);
      if ((rc = launch_logw_add_real(f.nw, w->logw.p, nullptr, cp->qrld[b].p, nullptr, 0, 1.0, 0.0, s))) return rc;
    }
  }
  return 0;
}

// steps s0 .. s1-1 from the current walkers; with store, Q_s goes to seg[s - s0] and the QR
// factors after step s to rinv slot (s - s0) / mgs
int ff_forward(FF& f, int s0, int s1, bool store) {
  afqmc_cp cp = f.cp;
  afqmc_walkers w = f.w;
  hipStream_t s = cp->ctx->stream;
  int rc;
  for (int st = s0; st < s1; ++st) {
    if (store)
      for (int b = 0; b < 2; ++b)
        HIP_TRY(hipMemcpyAsync(f.seg[b].p + (size_t)(st - s0) * f.ws[b] * f.nw, w->wf[b].p,
                               sizeof(zd) * f.ws[b] * f.nw, hipMemcpyDeviceToDevice, s));
    if ((rc = ff_scales(f, st))) return rc;
    if ((rc = propagate_gemm<double>(cp, w, 0, w->wf, w->tmp, true))) return rc;
    for (int b = 0; b < 2; ++b) swap(w->wf[b], w->tmp[b]);
    if ((st + 1) % f.mgs == 0) {
      if ((rc = ff_qr(f, store ? (st - s0) / f.mgs : -1))) return rc;
    }
  }
  return 0;
}

// E and the weights of the current walkers; with adjoint, also Wt_n, mbar sums and Ubar (direct)
int ff_terminal(FF& f, bool adjoint) {
  afqmc_cp cp = f.cp;
  afqmc_walkers w = f.w;
  hipStream_t s = cp->ctx->stream;
  int rc;
  if ((rc = overlap_gemm<double>(cp, w, w->wf))) return rc;
  if ((rc = gj_blocks<double>(cp, w, true, nullptr))) return rc;
  for (int b = 0; b < 2; ++b) {
    const int Mb = cp->Mb[b], Nb = cp->Nb[b];
    GemmArgs a = gb();
    a.M = Mb; a.N = Nb; a.K = Nb; a.batch = f.nw;
    a.A = w->wf[b].p; a.lda = Mb; a.strideA = f.ws[b];
    a.B = cp->ovl[b].p; a.ldb = Nb; a.strideB = (long long)Nb * Nb;
    a.C = f.theta[b].p; a.ldc = Mb; a.strideC = f.ws[b];
    if ((rc = launch_cgemm<double>(OP_N, EPI_STORE, a, s))) return rc;
    hipLaunchKernelGGL(mdiag_kernel, dim3((Mb + 63) / 64, f.nw), dim3(64), 0, s, f.nw, Mb, Nb, f.theta[b].p,
                       (const zd*)cp->phiT[b].p, f.m[b].p);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(kin_kernel, dim3(f.nw), dim3(256), 0, s, Mb, Nb, f.theta[b].p, (const zd*)cp->wdkT[b].p,
                       f.kin.p, b);
    HIP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(weights_kernel, dim3(1), dim3(64), 0, s, f.nw, f.L, cp->U.p, f.m[0].p, f.m[1].p, f.kin.p,
                     w->logw.p, cp->ldet[0].p, cp->ldet[1].p, f.el.p, f.logO.p, f.c.p, f.E.p);
  HIP_LAUNCH_CHECK();
  if (!adjoint) return 0;
  for (int b = 0; b < 2; ++b) {
    const int Mb = cp->Mb[b], Nb = cp->Nb[b];
    const long long n = f.ws[b] * f.nw;
    hipLaunchKernelGGL(conj_kernel, dim3(grid_for(f.ws[b])), dim3(256), 0, s, f.ws[b], (const zd*)cp->phiT[b].p,
                       f.conjphi[b].p);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(pmat_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, (const zd*)cp->wdkT[b].p, cp->U.p,
                       f.m[1 - b].p, (const zd*)cp->phiT[b].p, f.P[b].p);
    HIP_LAUNCH_CHECK();
    // S = theta^T P
    GemmArgs a = gb();
    a.M = Nb; a.N = Nb; a.K = Mb; a.batch = f.nw;
    a.A = f.theta[b].p; a.lda = Mb; a.strideA = f.ws[b];
    a.B = f.P[b].p; a.ldb = Mb; a.strideB = f.ws[b];
    a.C = f.S[b].p; a.ldc = Nb; a.strideC = (long long)Nb * Nb;
    if ((rc = launch_cgemm<double>(OP_T, EPI_STORE, a, s))) return rc;
    // P -= conj(Phi) S
    GemmArgs g = gb();
    g.M = Mb; g.N = Nb; g.K = Nb; g.batch = f.nw;
    g.A = f.conjphi[b].p; g.lda = Mb; g.strideA = 0;
    g.B = f.S[b].p; g.ldb = Nb; g.strideB = (long long)Nb * Nb;
    g.C = f.P[b].p; g.ldc = Mb; g.strideC = f.ws[b];
    g.alpha_re = -1.0; g.beta_one = 1;
    if ((rc = launch_cgemm<double>(OP_N, EPI_STORE, g, s))) return rc;
    hipLaunchKernelGGL(zmat_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, f.c.p, f.el.p, f.E.p,
                       (const zd*)cp->phiT[b].p, f.P[b].p);
    HIP_LAUNCH_CHECK();
    // Wt = Z O^{-T}
    const long long nn = (long long)Nb * Nb * f.nw;
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for(nn)), dim3(256), 0, s, nn, Nb, Nb, (const zd*)cp->ovl[b].p,
                       f.oinvT[b].p);
    HIP_LAUNCH_CHECK();
    GemmArgs z = gb();
    z.M = Mb; z.N = Nb; z.K = Nb; z.batch = f.nw;
    z.A = f.P[b].p; z.lda = Mb; z.strideA = f.ws[b];
    z.B = f.oinvT[b].p; z.ldb = Nb; z.strideB = (long long)Nb * Nb;
    z.C = f.wt[b].p; z.ldc = Mb; z.strideC = f.ws[b];
    if ((rc = launch_cgemm<double>(OP_N, EPI_STORE, z, s))) return rc;
  }
  hipLaunchKernelGGL(udirect_kernel, dim3((f.L + 63) / 64), dim3(64), 0, s, f.nw, f.L, f.c.p, f.m[0].p, f.m[1].p,
                     f.ubar.p);
  HIP_LAUNCH_CHECK();
  return 0;
}

// one reverse step s (Q_s in seg slot s - s0); Wt holds Wt_{s+1} on entry and Wt_s on exit
int ff_reverse_step(FF& f, int st, int s0) {
  afqmc_cp cp = f.cp;
  hipStream_t s = cp->ctx->stream;
  int rc;
  if ((st + 1) % f.mgs == 0) {
    // Wt <- Wt R^{-T} with R^{-1} = R1^{-1} R2^{-1}: first R2^{-T}, then R1^{-T}
    const int q = (st - s0) / f.mgs;
    for (int b = 0; b < 2; ++b) {
      const int Mb = cp->Mb[b], Nb = cp->Nb[b];
      const long long nn = (long long)Nb * Nb * f.nw;
      for (int pass = 1; pass >= 0; --pass) {
        const zd* r = f.rinv[b].p + ((size_t)q * 2 + pass) * (size_t)Nb * Nb * f.nw;
        hipLaunchKernelGGL(transpose_kernel, dim3(grid_for(nn)), dim3(256), 0, s, nn, Nb, Nb, r, f.oinvT[b].p);
        HIP_LAUNCH_CHECK();
        GemmArgs a = gb();
        a.M = Mb; a.N = Nb; a.K = Nb; a.batch = f.nw;
        a.A = f.wt[b].p; a.lda = Mb; a.strideA = f.ws[b];
        a.B = f.oinvT[b].p; a.ldb = Nb; a.strideB = (long long)Nb * Nb;
        a.C = f.t1[b].p; a.ldc = Mb; a.strideC = f.ws[b];
        if ((rc = launch_cgemm<double>(OP_N, EPI_STORE, a, s))) return rc;
        swap(f.wt[b], f.t1[b]);
      }
    }
  }
  if ((rc = ff_scales(f, st))) return rc;
  for (int b = 0; b < 2; ++b) {
    const int Mb = cp->Mb[b], Nb = cp->Nb[b];
    const long long n = f.ws[b] * f.nw;
    const zd* Q = f.seg[b].p + (size_t)(st - s0) * n;
    const zd* scale = (const zd*)cp->scale[b].p;
    // Bbar_w += Wt_{s+1} (D_s Q_s)^T: per walker (Wt^T)^T (Y^T) with op T
    hipLaunchKernelGGL(rowscale_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, scale, Q, f.t1[b].p);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, (const zd*)f.t1[b].p,
                       f.t2[b].p);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(transpose_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, (const zd*)f.wt[b].p,
                       f.t1[b].p);
    HIP_LAUNCH_CHECK();
    GemmArgs a = gb();
    a.M = Mb; a.N = Mb; a.K = Nb; a.batch = f.nw;
    a.A = f.t1[b].p; a.lda = Nb; a.strideA = n / f.nw;
    a.B = f.t2[b].p; a.ldb = Nb; a.strideB = n / f.nw;
    a.C = f.bb[b].p; a.ldc = Mb; a.strideC = (long long)Mb * Mb;
    a.beta_one = 1;
    if ((rc = launch_cgemm<double>(OP_T, EPI_STORE, a, s))) return rc;
    // Ybar = B^T Wt_{s+1}
    GemmArgs y = gb();
    y.M = Mb; y.N = Nb * f.nw; y.K = Mb; y.batch = 1;
    y.A = cp->expK[0][b].p; y.lda = Mb;
    y.B = f.wt[b].p; y.ldb = Mb;
    y.C = f.t1[b].p; y.ldc = Mb;
    if ((rc = launch_cgemm<double>(OP_T, EPI_STORE, y, s))) return rc;
    // Ubar += sum_c Ybar Q D dlogD
    hipLaunchKernelGGL(ubar_partial_kernel, dim3((Mb + 63) / 64, f.nw), dim3(64), 0, s, st, f.key, f.nw, Mb, Nb,
                       b == 0 ? 1.0 : -1.0, (const zd*)f.gprime.p, cp->dt, scale, (const zd*)f.t1[b].p, Q, f.part.p);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(sum_walkers_kernel, dim3((Mb + 63) / 64), dim3(64), 0, s, f.nw, Mb, (const zd*)f.part.p,
                       f.ubar.p);
    HIP_LAUNCH_CHECK();
    // Wt_s = D_s Ybar
    hipLaunchKernelGGL(rowscale_kernel, dim3(grid_for(n)), dim3(256), 0, s, n, Mb, Nb, scale, (const zd*)f.t1[b].p,
                       f.wt[b].p);
    HIP_LAUNCH_CHECK();
  }
  return 0;
}

int ff_check(afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, int ckpt) {
  REQUIRE(cp && w && w->cp == cp, "invalid argument");
  REQUIRE(!cp->fp32, "the fixed-field adjoint runs in FP64 (precision 0)");
  REQUIRE(!cp->dense && cp->nblk == 2 && cp->diag_only && cp->decomp == DECOMP_DENSITY_SPIN,
          "the fixed-field adjoint needs spin-block walkers with the densitySpin decomposition");
  REQUIRE(nsteps >= 0 && mgs > 0 && ckpt > 0 && ckpt % mgs == 0, "need nsteps >= 0 and ckpt a multiple of mgs > 0");
  return 0;
}

int ff_setup(FF& f, afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, int ckpt, uint64_t seed) {
  f.cp = cp; f.w = w; f.nw = w->nw; f.L = cp->L; f.nsteps = nsteps; f.mgs = mgs; f.ckpt = ckpt;
  f.nseg = std::max(1, (nsteps + ckpt - 1) / ckpt);
  f.key = afqmc_seed_mix_host(seed);
  f.ws[0] = w->wsize(0); f.ws[1] = w->wsize(1);
  {
    // the HS factors' U-derivative below assumes constDiag = exp(-dt U_i / 2): mu = h = 0
    std::vector<double> v(f.L);
    const DevBuf<double>* fields[4] = {&cp->mu, &cp->hx, &cp->hy, &cp->hz};
    for (const DevBuf<double>* d : fields) {
      HIP_TRY(hipMemcpy(v.data(), d->p, sizeof(double) * f.L, hipMemcpyDeviceToHost));
      for (double x : v) REQUIRE(x == 0.0, "the fixed-field adjoint supports mu = hx = hy = hz = 0 only");
    }
  }
  int rc = ff_alloc(f);
  if (rc) return rc;
  // d gamma_i / d U_i for gamma = solveCoshxEqExpy(dt U / 2) = log(e - sqrt(e^2 - 1)) of densitySpin
  std::vector<zd> gp(f.L);
  std::vector<double> hU(f.L);
  HIP_TRY(hipMemcpy(hU.data(), cp->U.p, sizeof(double) * f.L, hipMemcpyDeviceToHost));
  for (int i = 0; i < f.L; ++i) {
    // |gamma| = acosh(e) with e = exp(dt U / 2); the sign is that of the gamma in use
    const double e = std::exp(cp->dt * hU[i] * 0.5);
    const double d = e * e > 1.0 ? 0.5 * cp->dt * e / std::sqrt(e * e - 1.0) : 0.0;
    gp[i] = cmk<double>(cp->h_gamma[i].real() < 0.0 ? -d : d, 0.0);
  }
  HIP_TRY(hipMemcpy(f.gprime.p, gp.data(), sizeof(zd) * f.L, hipMemcpyHostToDevice));
  hipStream_t s = cp->ctx->stream;
  HIP_TRY(hipMemsetAsync(f.ubar.p, 0, sizeof(zd) * f.L, s));
  for (int b = 0; b < 2; ++b) HIP_TRY(hipMemsetAsync(f.bb[b].p, 0, sizeof(zd) * (size_t)f.L * f.L * f.nw, s));
  HIP_TRY(hipMemcpyAsync(f.init_logw.p, w->logw.p, sizeof(zd) * f.nw, hipMemcpyDeviceToDevice, s));
  return 0;
}

// checkpoint k <- current walkers (or the reverse)
int ff_ck(FF& f, int k, bool save) {
  hipStream_t s = f.cp->ctx->stream;
  for (int b = 0; b < 2; ++b) {
    zd* ck = f.ckw[b].p + (size_t)k * f.ws[b] * f.nw;
    void* wf = f.w->wf[b].p;
    HIP_TRY(hipMemcpyAsync(save ? (void*)ck : wf, save ? wf : (void*)ck, sizeof(zd) * f.ws[b] * f.nw,
                           hipMemcpyDeviceToDevice, s));
  }
  zd* cl = f.ckl.p + (size_t)k * f.nw;
  HIP_TRY(hipMemcpyAsync(save ? (void*)cl : (void*)f.w->logw.p, save ? (void*)f.w->logw.p : (void*)cl,
                         sizeof(zd) * f.nw, hipMemcpyDeviceToDevice, s));
  return 0;
}

}  // namespace

extern "C" {

int afqmc_cp_fixed_field_energy(afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, uint64_t field_seed,
                                double* energy) {
  int rc = ff_check(cp, w, nsteps, mgs, mgs);
  if (rc) return rc;
  REQUIRE(energy, "null argument");
  NOT_CAPTURING();
  ScopedDevice sd(cp->ctx);
  FF f;
  if ((rc = ff_setup(f, cp, w, nsteps, mgs, mgs, field_seed))) return rc;
  if ((rc = ff_ck(f, 0, true))) return rc;
  if ((rc = ff_forward(f, 0, nsteps, false))) return rc;
  if ((rc = ff_terminal(f, false))) return rc;
  zd E;
  HIP_TRY(hipMemcpyAsync(&E, f.E.p, sizeof(zd), hipMemcpyDeviceToHost, cp->ctx->stream));
  if ((rc = ff_ck(f, 0, false))) return rc;
  HIP_TRY(hipStreamSynchronize(cp->ctx->stream));
  energy[0] = E.re;
  energy[1] = E.im;
  return 0;
}

int afqmc_cp_fixed_field_adjoint(afqmc_cp cp, afqmc_walkers w, int nsteps, int mgs, int ckpt, uint64_t field_seed,
                                 double* energy, double* bbar, double* mbar, double* ubar) {
  int rc = ff_check(cp, w, nsteps, mgs, ckpt);
  if (rc) return rc;
  REQUIRE(energy && bbar && mbar && ubar, "null argument");
  NOT_CAPTURING();
  ScopedDevice sd(cp->ctx);
  hipStream_t s = cp->ctx->stream;
  FF f;
  if ((rc = ff_setup(f, cp, w, nsteps, mgs, ckpt, field_seed))) return rc;
  // forward with checkpoints at 0, ckpt, 2 ckpt, ...
  for (int k = 0; k < f.nseg; ++k) {
    if ((rc = ff_ck(f, k, true))) return rc;
    const int s0 = k * ckpt, s1 = std::min(nsteps, s0 + ckpt);
    if ((rc = ff_forward(f, s0, s1, false))) return rc;
  }
  if ((rc = ff_terminal(f, true))) return rc;
  // reverse sweep, one recomputed segment at a time
  for (int k = f.nseg - 1; k >= 0; --k) {
    const int s0 = k * ckpt, s1 = std::min(nsteps, s0 + ckpt);
    if (s1 <= s0) continue;
    if ((rc = ff_ck(f, k, false))) return rc;
    if ((rc = ff_forward(f, s0, s1, true))) return rc;
    for (int st = s1 - 1; st >= s0; --st)
      if ((rc = ff_reverse_step(f, st, s0))) return rc;
  }
  // outputs
  const int L = f.L;
  std::vector<zd> hb((size_t)L * L), hth;
  DevBuf<zd> bsum, thsum;
  if ((rc = bsum.alloc((size_t)L * L))) return rc;
  for (int b = 0; b < 2; ++b) {
    hipLaunchKernelGGL(plain_sum_kernel, dim3(grid_for((long long)L * L)), dim3(256), 0, s, f.nw, (long long)L * L,
                       (const zd*)f.bb[b].p, bsum.p);
    HIP_LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(hb.data(), bsum.p, sizeof(zd) * L * L, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (size_t e = 0; e < hb.size(); ++e) {
      bbar[2 * ((size_t)b * L * L + e)] = hb[e].re;
      bbar[2 * ((size_t)b * L * L + e) + 1] = hb[e].im;
    }
  }
  // mbar_b = conj(Phi_b) (sum_w c_w theta_w)^T, L x L
  for (int b = 0; b < 2; ++b) {
    const int Nb = cp->Nb[b];
    const long long per = (long long)L * Nb;
    if ((rc = thsum.alloc((size_t)per))) return rc;
    hipLaunchKernelGGL(weighted_sum_kernel, dim3(grid_for(per)), dim3(256), 0, s, f.nw, per, (const zd*)f.c.p,
                       (const zd*)f.theta[b].p, thsum.p);
    HIP_LAUNCH_CHECK();
    hth.resize((size_t)per);
    std::vector<zd> phi((size_t)per);
    HIP_TRY(hipMemcpyAsync(hth.data(), thsum.p, sizeof(zd) * per, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(phi.data(), cp->phiT[b].p, sizeof(zd) * per, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j) {
        // (conj(Phi) Theta^T)(i, j) = sum_k conj(Phi(i,k)) Theta(j,k)
        cd acc = 0.0;
        for (int kk = 0; kk < Nb; ++kk) {
          const zd p = phi[i + (size_t)kk * L], t = hth[j + (size_t)kk * L];
          acc += std::conj(cd(p.re, p.im)) * cd(t.re, t.im);
        }
        mbar[2 * ((size_t)b * L * L + i + (size_t)j * L)] = acc.real();
        mbar[2 * ((size_t)b * L * L + i + (size_t)j * L) + 1] = acc.imag();
      }
  }
  std::vector<zd> hu(L);
  zd E;
  HIP_TRY(hipMemcpyAsync(hu.data(), f.ubar.p, sizeof(zd) * L, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(&E, f.E.p, sizeof(zd), hipMemcpyDeviceToHost, s));
  // restore the input walkers
  if ((rc = ff_ck(f, 0, false))) return rc;
  HIP_TRY(hipMemcpyAsync(w->logw.p, f.init_logw.p, sizeof(zd) * f.nw, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
  for (int i = 0; i < L; ++i) { ubar[2 * i] = hu[i].re; ubar[2 * i + 1] = hu[i].im; }
  energy[0] = E.re;
  energy[1] = E.im;
  return 0;
}

}  // extern "C"
