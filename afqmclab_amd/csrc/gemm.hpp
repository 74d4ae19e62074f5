// Batched complex GEMM on MFMA for gfx950.
//
//   C_b = alpha * op(A_b) * diag(s_{b,n}) * B_b          (EPI_STORE)
//   green row-dots of X_b = op(A_b) * B_b                (EPI_GREEN)
//
// op(A) in {N, T, C}; any of A, B may be shared across the batch (stride 0).
// "Giant" mode: one batch whose N dimension concatenates all walkers' columns
// (walkers are stored back to back with the same leading dimension), with the
// per-walker row scaling selected by column group n / scale_group.
#pragma once
#include "common.hpp"

enum { OP_N = 0, OP_T = 1, OP_C = 2, OP_NR = 3 };  // OP_NR: A real (element type T), op N
enum { EPI_STORE = 0, EPI_GREEN = 1 };

struct GemmArgs {
  int M, N, K, batch;
  const void* A; long long lda, strideA;
  const void* B; long long ldb, strideB;
  void* C; long long ldc, strideC;
  double alpha_re, alpha_im;
  int beta_one;            // C += result instead of C = result
  // optional row scaling of B (complex T): s = scale + b*scale_strideB + (n/scale_group)*scale_strideG
  const void* scale; long long scale_strideB, scale_strideG; int scale_group;
  // green epilogue partners (complex T, M x N column-major, leading dim ldp, shared across batch)
  const void* P0;          // trial wf: d0 = sum_j X(i,j) conj(P0(i,j)),  d1 = sum_j X(i,j) conj(P0(i^,j))
  const void* P2;          // (phiT^H K)^T: d2 = sum_j X(i,j) P2(i,j)
  long long ldp;
  int partner_half;        // i^ = i +- partner_half for the off-diagonal pairs (0 = none)
  zd* out0; zd* out1; zd* out2;  // [batch][ceil(N / 64)][M] partial sums, one per 64-column group
  // alive mask: batches with alive[b]==0 are skipped (batched mode only); nullptr = all
  const int* alive;
  // tile order within an XCD: 0 = M tiles fastest (B panels shared), 1 = N tiles fastest (A panels
  // shared; for a large shared A against many narrow B panels)
  int n_fastest;
  // output tile shape (GEMM_TILE_*); the Green epilogue takes the shapes of gemm_tile_green
  int tile;
  // herm: the result is Hermitian (a Gram matrix op(A) B with op C and A = B): only the tiles on
  // or above the diagonal are computed, the others written as their conjugate transposes
  // (square tiles only).  b_upper: B is upper triangular, so the K loop of an output column tile
  // stops at its last column.
  int herm, b_upper;
  // split_bf16 (FP32, op N, plain store only): each FP32 operand split into two BF16 parts and
  // each real product formed as hi*hi + hi*lo + lo*hi with BF16 MFMAs accumulating in FP32
  // (fixed 128x64 tile); about 1e-6 of sum |a b| instead of FP32's 1e-7
  int split_bf16;
  // set by the launcher, not the caller: every operand's byte offsets from a workgroup's base fit
  // in 31 bits, so the kernel loads its whole K tiles through buffer resources
  int fast_loads;
};

// Output tile shapes of the MFMA GEMM: the autotuner picks one per call site.
enum {
  GEMM_TILE_64x64 = 0, GEMM_TILE_32x64 = 1, GEMM_TILE_64x32 = 2, GEMM_TILE_32x32 = 3, GEMM_TILE_128x64 = 4,
  GEMM_TILE_64x64_K32 = 5, GEMM_TILE_64x32_K32 = 6, GEMM_TILE_32x64_K32 = 7,
  // 8 waves (128x128: 2 x 4, 256x64: 4 x 2) and a 64x128 tile of 4 waves, all with a 16-deep K tile
  GEMM_TILE_128x128 = 8, GEMM_TILE_256x64 = 9, GEMM_TILE_64x128 = 10, GEMM_TILE_COUNT = 11
};
inline int gemm_tile_bn(int tile) {
  if (tile == GEMM_TILE_64x32 || tile == GEMM_TILE_32x32 || tile == GEMM_TILE_64x32_K32) return 32;
  if (tile == GEMM_TILE_128x128 || tile == GEMM_TILE_64x128) return 128;
  return 64;
}
// shapes with the Green epilogue: 32 columns per wave and whole 64-column groups
inline bool gemm_tile_green(int tile) { return gemm_tile_bn(tile) == 64 || tile == GEMM_TILE_128x128; }
inline bool gemm_tile_square(int tile) {
  return tile == GEMM_TILE_64x64 || tile == GEMM_TILE_32x32 || tile == GEMM_TILE_64x64_K32 ||
         tile == GEMM_TILE_128x128;
}

template <typename T>
int launch_cgemm(int opa, int epi, const GemmArgs& a, hipStream_t s);
