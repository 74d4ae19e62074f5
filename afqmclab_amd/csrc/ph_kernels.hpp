// Walker-level kernels of the phaseless step (afqmcPhaselessProjection.cpp:75-121), the
// continuous HS of CholeskyReal (choleskyReal.cpp:57-169), the Taylor exponential of
// LogHop2isSD2isOperation (logHop2isSD2isOperation.cpp:99-195) and the mixed estimator of
// RealMaterialMoleculeMeasureFixedSD2sSD2is (:197-284).
#pragma once
#include "cp_kernels.hpp"

struct PhSampleArgs {
  int nw, G, force_type, fp32;
  const void* bg;            // complex T, [nw][G]: sum_s vec(A'_s[g]) . vec(X_s) from the GEMM
  const double* cur_bg;      // G: background of the model (choleskyBg)
  const zd* const_force;     // G
  zd sqrt_minus_dt;
  double cap, dt;
  zd onebody_logw;
  int* alive;
  const int* prefix;
  DevScalars* sc;
  const double* ubuf;
  zd* logw;
  void* xs;                  // complex T, [nw][G]: sqrt(-dt) x_g, the sample GEMM's B operand
  zd* force_out; zd* aux_out; zd* sample_logw_out;  // optional records [nw][G], [nw]
};
int launch_ph_sample(const PhSampleArgs& a, hipStream_t s);

// W' += T (both [nw][n] complex T); for walkers that are active: maxabs of T over the walker,
// then deactivate when order > base and maxabs <= accuracy (LogHop2isSD2isOperation::isConverged)
template <typename T>
int launch_taylor_accumulate(int nw, long long n, const void* Tk, void* W, int* active, int* orders, int order,
                             int base, double accuracy, int dynamic, int* n_active, hipStream_t s);

// ex[w][g] = sum_{i,j < Ns} D(g*Ns + j, i) D(g*Ns + i, j), D per walker (G*Ns x Ns, column-major)
template <typename T>
int launch_exchange_trace(int nw, int G, int Ns, const void* D, zd* ex, int accumulate, hipStream_t s);

// t[w] += sum_{i,j} X(i,j) P(i,j)  (X per walker L x Ns, P shared)
template <typename T>
int launch_rowdot_sum(int nw, long long n, const void* X, long long strideX, const void* P, zd* out, int accumulate,
                      hipStream_t s);

struct PhEnergyArgs {
  int nw, G, fp32;
  const int* alive;
  const zd* den;          // [nw] exp(logO) of the measured walkers
  const zd* tkin;         // [nw]
  const void* bg;         // complex T [nw][G]
  const zd* ex;           // [nw][G]
  zd* per_walker;         // [nw][4]: den, H*den, T*den, H
  zd* bg_acc; zd* ex_acc; // [G] += bg * den, ex * den
  DevScalars* sc;
};
int launch_ph_energy(const PhEnergyArgs& a, hipStream_t s);
