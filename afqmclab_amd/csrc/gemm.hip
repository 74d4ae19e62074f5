// Batched complex GEMM kernels (MFMA 16x16x4, f64 or f32) for gfx950.
//
// FP32: the complex product takes three real MFMAs per 16x16x4 block (the 3M
// form of Gauss's product): P1 += Are*Bre, P2 += Aim*Bim,
// P3 += (Are+Aim)*(Bre+Bim), accumulated over K, then Cre = P1 - P2,
// Cim = P3 - P1 - P2: 6 real flops per complex multiply-add where the roofline
// counts the algorithmic 8.  FP64: four (Cre += Are*Bre - Aim*Bim,
// Cim += Are*Bim + Aim*Bre); a third set of FP64 accumulators would cost more
// occupancy than the saved MFMA time.  A real A (op NR) needs two products,
// Cre += A*Bre and Cim += A*Bim.  Tiles are staged
// through LDS as separate re/im planes so that each MFMA operand is one LDS
// read per lane (the operand sums are formed in registers); the next K tile is
// prefetched into registers while the current one is multiplied.
#include "gemm.hpp"

#include "group.hpp"

#include <type_traits>

namespace {

// Green tiles with wave tiles of up to 32 x 32: left alone, the compiler gives the 64 x 64 tile
// 216 registers (two waves per SIMD); asked for at least two waves it schedules it in 158 (three
// waves), and more waves let one wave's prologue and epilogue overlap the others' MFMAs, which
// matters because the K loop of a Green tile is only N/16 long.  The larger wave tiles already
// sit at 256 registers.
// One 4-, 8- or 16-byte element through a buffer resource: voff is the lane's byte offset, soff a
// wave-uniform byte offset (the K advance)
template <typename V>
__device__ __forceinline__ V buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  static_assert(sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16, "element size");
  if constexpr (sizeof(V) == 4) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  } else if constexpr (sizeof(V) == 8) {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  }
}
// A raw buffer resource over 2^31 - 1 bytes from base (dword 3 of the gfx9 descriptor: 0x00020000).
// base must be wave-uniform; it is moved to scalar registers explicitly, because a uniform value
// that the compiler computed in vector registers (the tile index's 64-bit division) would
// otherwise be wrapped in a loop over its distinct lane values.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(base);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void* u = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(u, (short)0, 0x7fffffff, 0x00020000);
}

template <typename T, int OPA, int BM, int BN, int BK, int WM, int WN, int EPI, bool SCALE>
struct CgemmK {
  static constexpr int NT = WM * WN * 64;
  static constexpr int WPE = (EPI == EPI_GREEN && (BM / WM / 16) * (BN / WN / 16) <= 4) ? 2 : 1;
  __device__ static void run(const BlockIdx3& bk, const GemmArgs& p);
};
template <typename T, int OPA, int BM, int BN, int BK, int WM, int WN, int EPI, bool SCALE>
__device__ __forceinline__ void CgemmK<T, OPA, BM, BN, BK, WM, WN, EPI, SCALE>::run(const BlockIdx3& bk, const GemmArgs& p) {
  constexpr int NT = WM * WN * 64;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int PAD = 1;
  constexpr int A_ELEMS = BM * BK / NT;
  constexpr int B_ELEMS = BK * BN / NT;
  static_assert(A_ELEMS * NT == BM * BK && B_ELEMS * NT == BK * BN, "tile/thread mismatch");
  typedef typename mfma_traits<T>::acc_t acc_t;
  typedef cplx<T> ct;

  __shared__ T As[2][BK][BM + PAD];
  __shared__ T Bs[2][BK][BN + PAD];

  // Workgroups are handed to the 8 XCDs in turn (workgroup id mod 8).  Renumber the tiles so
  // that each XCD gets a contiguous range (M tile fastest, or N tile fastest when asked): the
  // workgroups that share an operand panel then run on the same XCD and read it through one L2.
  int bx, by, b;
  {
    const unsigned long long gx = bk.gx, gy = bk.gy;
    const unsigned long long total = gx * gy * bk.gz;
    const unsigned long long L = bk.x + gx * (bk.y + gy * bk.z);
    const unsigned long long per = total / 8;
    const unsigned long long lg = (L < per * 8) ? (L % 8) * per + L / 8 : L;
    if (p.n_fastest) {
      by = (int)(lg % gy);
      const unsigned long long r = lg / gy;
      bx = (int)(r % gx);
      b = (int)(r / gx);
    } else {
      bx = (int)(lg % gx);
      const unsigned long long r = lg / gx;
      by = (int)(r % gy);
      b = (int)(r / gy);
    }
  }
  if (p.alive && !p.alive[b]) return;
  if (p.herm && bx > by) return;  // the conjugate transpose of tile (by, bx) fills it
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int m0 = bx * BM, n0 = by * BN;

  const ct* __restrict__ A = reinterpret_cast<const ct*>(p.A) + (long long)b * p.strideA;
  const T* __restrict__ Ar = reinterpret_cast<const T*>(p.A) + (long long)b * p.strideA;  // OP_NR
  const ct* __restrict__ Bm = reinterpret_cast<const ct*>(p.B) + (long long)b * p.strideB;
  const ct* __restrict__ S = SCALE ? reinterpret_cast<const ct*>(p.scale) + (long long)b * p.scale_strideB : nullptr;

  // 3M (FP32, complex A): acc_re = P1, acc_x = P2, acc_im = P3 until the end of the K loop
  constexpr bool GAUSS = std::is_same<T, float>::value && OPA != OP_NR;
  acc_t acc_re[FM][FN], acc_im[FM][FN], acc_x[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc_re[i][j] = acc_t{0, 0, 0, 0};
      acc_im[i][j] = acc_t{0, 0, 0, 0};
      acc_x[i][j] = acc_t{0, 0, 0, 0};
    }

  // The next K tile is loaded into registers before the current one is multiplied and stored to
  // LDS after it: loads are issued unconditionally (out-of-range lanes read the operand's first
  // element) and masked, and the HS scale is applied at the store, so no wait on the loads sits
  // between their issue and the MFMAs.
  ct ra[A_ELEMS], rb[B_ELEMS];
  ct rs[SCALE ? B_ELEMS : 1];
  bool ina[A_ELEMS], inb[B_ELEMS];
  const bool kpos = p.K > 0;

  // Load addressing is fixed over the K loop: a thread's element e sits at a per-thread base plus
  // e times a uniform stride, plus the K offset (k0 * lda for A in op N, k0 otherwise), so a tile
  // costs one 64-bit add per element instead of the index arithmetic.
  constexpr bool A_NMAJ = OPA == OP_N || OPA == OP_NR;
  static_assert(NT % BM == 0 && NT % BK == 0, "load layout");
  constexpr int AE = A_NMAJ ? NT / BM : NT / BK;  // rows of k (op N) or m (op T/C) per element step
  const int a_m = A_NMAJ ? tid % BM : tid / BK;  // first element's m (+ e AE for op T/C)
  const int a_k = A_NMAJ ? tid / BM : tid % BK;  // first element's k (+ e AE for op N)
  const long long a_base = A_NMAJ ? (long long)(m0 + a_m) + (long long)a_k * p.lda
                                  : (long long)a_k + (long long)(m0 + a_m) * p.lda;
  const long long a_estride = (long long)AE * p.lda;
  const long long a_kstride = A_NMAJ ? p.lda : 1;
  constexpr int BE = NT / BK;
  const int b_k = tid % BK, b_n = tid / BK;
  const long long b_base = (long long)b_k + (long long)(n0 + b_n) * p.ldb;
  const long long b_estride = (long long)BE * p.ldb;
  long long s_base[SCALE ? B_ELEMS : 1];
  if (SCALE) {
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e)
      s_base[e] = (long long)((n0 + b_n + e * BE) / p.scale_group) * p.scale_strideG + b_k;
  }

  // Buffer-resource loads (p.fast_loads) for the whole K tiles: each element's byte offset from the
  // workgroup's operand base (row m0 of A, column n0 of B) is fixed over the K loop and the K
  // advance is a scalar offset, so a K tile's loads cost no vector address arithmetic.  Elements
  // outside M or N read the base and are zeroed at the store; a tile wholly inside M and N stores
  // unmasked.  A partial last K tile takes the masked loads.
  constexpr int AESZ = OPA == OP_NR ? (int)sizeof(T) : (int)sizeof(ct);
  // Only where it measured faster on MI355X: FP32 op-N tiles of up to 64 x 64 with a plain store
  // (the propagation GEMM's 64 x 64 tile: 0.93 against 0.99 ms per launch at 16 x 16 sites, 112
  // electrons per spin and 2048 walkers).  The offset registers cost the larger and the FP64 tiles
  // occupancy or spills, and the Green and overlap kernels a wave per SIMD.
  constexpr bool FAST_OK = std::is_same<T, float>::value && OPA == OP_N && EPI == EPI_STORE && BM * BN <= 64 * 64;
  const bool fast = FAST_OK && p.fast_loads != 0;
  const bool tile_in = m0 + BM <= p.M && n0 + BN <= p.N;
  const long long a_wg = A_NMAJ ? (long long)m0 : (long long)m0 * p.lda;
  const __amdgpu_buffer_rsrc_t rsa =
      buf_rsrc(OPA == OP_NR ? (const void*)(Ar + a_wg) : (const void*)(A + a_wg));
  const __amdgpu_buffer_rsrc_t rsb = buf_rsrc(Bm + (long long)n0 * p.ldb);
  const __amdgpu_buffer_rsrc_t rss = buf_rsrc(SCALE ? (const void*)S : (const void*)Bm);
  int a_vo[A_ELEMS], b_vo[B_ELEMS], s_vo[SCALE ? B_ELEMS : 1];
  if (fast) {
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int am = a_m + (A_NMAJ ? 0 : e * AE), ak = a_k + (A_NMAJ ? e * AE : 0);
      const bool in = m0 + am < p.M;
      ina[e] = in;
      const long long o = A_NMAJ ? (long long)am + (long long)ak * p.lda : (long long)ak + (long long)am * p.lda;
      a_vo[e] = in ? (int)(o * AESZ) : 0;
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const int bn = b_n + e * BE;
      const bool in = n0 + bn < p.N;
      inb[e] = in;
      b_vo[e] = in ? (int)(((long long)b_k + (long long)bn * p.ldb) * (long long)sizeof(ct)) : 0;
      if (SCALE) s_vo[e] = in ? (int)(s_base[e] * (long long)sizeof(ct)) : 0;
    }
  }
  auto load_fast = [&](int k0) {
    const int sa = __builtin_amdgcn_readfirstlane((int)((long long)k0 * a_kstride * AESZ));
    const int sb = __builtin_amdgcn_readfirstlane(k0 * (int)sizeof(ct));
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      if (OPA == OP_NR) ra[e] = cmk<T>(buf_load<T>(rsa, a_vo[e], sa), (T)0);
      else ra[e] = buf_load<ct>(rsa, a_vo[e], sa);
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      rb[e] = buf_load<ct>(rsb, b_vo[e], sb);
      if (SCALE) rs[e] = buf_load<ct>(rss, s_vo[e], sb);
    }
  };

  auto load_tiles = [&](int k0) {
    if (!kpos) {
#pragma unroll
      for (int e = 0; e < A_ELEMS; ++e) ina[e] = false;
#pragma unroll
      for (int e = 0; e < B_ELEMS; ++e) inb[e] = false;
      return;
    }
    const long long a_k0 = (long long)k0 * a_kstride;
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      const int gm = m0 + a_m + (A_NMAJ ? 0 : e * AE), gk = k0 + a_k + (A_NMAJ ? e * AE : 0);
      const bool in = gm < p.M && gk < p.K;
      ina[e] = in;
      const long long off = in ? a_base + e * a_estride + a_k0 : 0;
      if (OPA == OP_NR) ra[e] = cmk<T>(Ar[off], (T)0);
      else ra[e] = A[off];
    }
    const bool bk_in = k0 + b_k < p.K;
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      const bool in = bk_in && n0 + b_n + e * BE < p.N;
      inb[e] = in;
      rb[e] = Bm[in ? b_base + e * b_estride + k0 : 0];
      if (SCALE) rs[e] = S[in ? s_base[e] + k0 : 0];
    }
  };
  auto store_tiles_as = [&](auto masked) {
    constexpr bool MASKED = decltype(masked)::value;
#pragma unroll
    for (int e = 0; e < A_ELEMS; ++e) {
      int idx = tid + e * NT;
      int m, k;
      if (OPA == OP_N || OPA == OP_NR) { m = idx % BM; k = idx / BM; }
      else { k = idx % BK; m = idx / BK; }
      const ct v = (!MASKED || ina[e]) ? ra[e] : cmk<T>(0, 0);
      As[0][k][m] = v.re;
      As[1][k][m] = (OPA == OP_C) ? -v.im : v.im;
    }
#pragma unroll
    for (int e = 0; e < B_ELEMS; ++e) {
      int idx = tid + e * NT;
      int k = idx % BK, n = idx / BK;
      ct v = cmk<T>(0, 0);
      if (!MASKED || inb[e]) v = SCALE ? cmul(rs[e], rb[e]) : rb[e];
      Bs[0][k][n] = v.re;
      Bs[1][k][n] = v.im;
    }
  };
  // the masks of the buffer path are fixed over the K loop: a tile inside M and N needs none
  auto store_tiles = [&](bool loaded_fast) {
    if (loaded_fast && tile_in) store_tiles_as(std::false_type{});
    else store_tiles_as(std::true_type{});
  };
  const int nk_whole = p.K / BK;  // K tiles with no rows past K
  auto load = [&](int kt) {
    const bool f = fast && kt < nk_whole;
    if (f) load_fast(kt * BK);
    else load_tiles(kt * BK);
    return f;
  };

  int nk = (p.K + BK - 1) / BK;
  // upper-triangular B: rows k > n of column n are zero, so K tiles past this tile's last
  // column add nothing
  if (p.b_upper) nk = min(nk, (min(n0 + BN, p.K) + BK - 1) / BK);
  bool next_fast = load(0);
  store_tiles(next_fast);
  __syncthreads();
  const int lr = lane & 15, lk = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) next_fast = load(kt + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T are[FM], aim[FM], asum[FM], bre[FN], bim[FN], bsum[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        int m = wm * TM + i * 16 + lr;
        are[i] = As[0][kk + lk][m];
        aim[i] = As[1][kk + lk][m];
        if (GAUSS) asum[i] = are[i] + aim[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int n = wn * TN + j * 16 + lr;
        bre[j] = Bs[0][kk + lk][n];
        bim[j] = Bs[1][kk + lk][n];
        if (GAUSS) bsum[j] = bre[j] + bim[j];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if (OPA == OP_NR) {  // real A: two products
            acc_re[i][j] = mfma_traits<T>::mma(are[i], bre[j], acc_re[i][j]);
            acc_im[i][j] = mfma_traits<T>::mma(are[i], bim[j], acc_im[i][j]);
          } else if (GAUSS) {
            acc_re[i][j] = mfma_traits<T>::mma(are[i], bre[j], acc_re[i][j]);
            acc_x[i][j] = mfma_traits<T>::mma(aim[i], bim[j], acc_x[i][j]);
            acc_im[i][j] = mfma_traits<T>::mma(asum[i], bsum[j], acc_im[i][j]);
          } else {
            acc_re[i][j] = mfma_traits<T>::mma(are[i], bre[j], acc_re[i][j]);
            acc_im[i][j] = mfma_traits<T>::mma(are[i], bim[j], acc_im[i][j]);
            acc_re[i][j] = mfma_traits<T>::mma(-aim[i], bim[j], acc_re[i][j]);
            acc_im[i][j] = mfma_traits<T>::mma(aim[i], bre[j], acc_im[i][j]);
          }
        }
    }
    __syncthreads();
    if (kt + 1 < nk) {
      store_tiles(next_fast);
      __syncthreads();
    }
  }

  if (GAUSS) {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T p1 = acc_re[i][j][r], p2 = acc_x[i][j][r], p3 = acc_im[i][j][r];
          acc_re[i][j][r] = p1 - p2;
          acc_im[i][j][r] = (p3 - p1) - p2;
        }
  }

  if (EPI == EPI_STORE) {
    ct* C = reinterpret_cast<ct*>(p.C) + (long long)b * p.strideC;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int gn = n0 + wn * TN + j * 16 + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int gm = m0 + wm * TM + i * 16 + mfma_traits<T>::row(lk, r);
          if (gm < p.M && gn < p.N) {
            double xr = (double)acc_re[i][j][r], xi = (double)acc_im[i][j][r];
            double yr = p.alpha_re * xr - p.alpha_im * xi;
            double yi = p.alpha_re * xi + p.alpha_im * xr;
            long long off = gm + (long long)gn * p.ldc;
            if (p.beta_one) {
              ct o = C[off];
              yr += (double)o.re;
              yi += (double)o.im;
            }
            C[off] = cmk<T>((T)yr, (T)yi);
            if (p.herm && bx < by) C[gn + (long long)gm * p.ldc] = cmk<T>((T)yr, (T)-yi);
          }
        }
      }
  } else {
    // green epilogue: row dots of X with the partner matrices, reduced over
    // this tile's columns, written as partial sums per column tile.
    const ct* P0 = reinterpret_cast<const ct*>(p.P0);
    const ct* P2 = reinterpret_cast<const ct*>(p.P2);
    const bool use1 = p.partner_half != 0, use2 = P2 != nullptr;
    __shared__ zd red[3][WN][BM];
    // each wave's dots over its 32 columns in the accumulators' precision (FP32 tiles: 32 terms
    // per partial, far below the mixed path's error), the sums across waves and tiles in FP64
    typedef cplx<T> dt;
    dt d0[FM][4], d1[FM][4], d2[FM][4];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        d0[i][r] = cmk<T>(0, 0);
        d1[i][r] = cmk<T>(0, 0);
        d2[i][r] = cmk<T>(0, 0);
      }
    // FP32 accumulators: lane (lk, lr) holds rows 4 lk .. 4 lk + 3 of column lr, which are
    // contiguous in a column-major partner: two 16-byte loads instead of four scattered ones
    constexpr bool ROWS4 = std::is_same<T, float>::value;
    const bool vec_ok = ROWS4 && (p.ldp % 2 == 0);
    const bool vec_partner = vec_ok && (p.partner_half % 4 == 0);
    auto load4 = [&](const ct* P, int gr0, int gn, bool fast, ct* q) {
      const ct* src = P + gr0 + (long long)gn * p.ldp;
      if (fast) {
        const float4 a = *reinterpret_cast<const float4*>(src);
        const float4 c = *reinterpret_cast<const float4*>(src + 2);
        q[0] = cmk<T>((T)a.x, (T)a.y); q[1] = cmk<T>((T)a.z, (T)a.w);
        q[2] = cmk<T>((T)c.x, (T)c.y); q[3] = cmk<T>((T)c.z, (T)c.w);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = src[r];
      }
    };
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int gn = n0 + wn * TN + j * 16 + lr;
        if (ROWS4) {
          const int gm0 = m0 + wm * TM + i * 16 + 4 * lk;
          if (gn < p.N && gm0 + 3 < p.M) {
            ct q0[4], q1[4], q2[4];
            load4(P0, gm0, gn, vec_ok, q0);
            if (use1) {
              int gp0 = gm0 < p.partner_half ? gm0 + p.partner_half : gm0 - p.partner_half;
              if (vec_partner) {
                load4(P0, gp0, gn, true, q1);
              } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                  int gm = gm0 + r;
                  int gp = gm < p.partner_half ? gm + p.partner_half : gm - p.partner_half;
                  q1[r] = P0[gp + (long long)gn * p.ldp];
                }
              }
            }
            if (use2) load4(P2, gm0, gn, vec_ok, q2);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const dt x = cmk<T>(acc_re[i][j][r], acc_im[i][j][r]);
              d0[i][r] = cadd(d0[i][r], cmul(x, cconj(q0[r])));
              if (use1) d1[i][r] = cadd(d1[i][r], cmul(x, cconj(q1[r])));
              if (use2) d2[i][r] = cadd(d2[i][r], cmul(x, q2[r]));
            }
            continue;
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int gm = m0 + wm * TM + i * 16 + mfma_traits<T>::row(lk, r);
          if (gm < p.M && gn < p.N) {
            const dt x = cmk<T>(acc_re[i][j][r], acc_im[i][j][r]);
            const dt q0 = P0[gm + (long long)gn * p.ldp];
            d0[i][r] = cadd(d0[i][r], cmul(x, cconj(q0)));
            if (use1) {
              int gp = gm < p.partner_half ? gm + p.partner_half : gm - p.partner_half;
              const dt q1 = P0[gp + (long long)gn * p.ldp];
              d1[i][r] = cadd(d1[i][r], cmul(x, cconj(q1)));
            }
            if (use2) {
              const dt q2 = P2[gm + (long long)gn * p.ldp];
              d2[i][r] = cadd(d2[i][r], cmul(x, q2));
            }
          }
        }
      }
    // reduce over the 16 lanes that share a row (lane % 16 = column); only the dots in use
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
          d0[i][r].re += __shfl_xor(d0[i][r].re, o, 64);
          d0[i][r].im += __shfl_xor(d0[i][r].im, o, 64);
        }
        if (use1) {
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) {
            d1[i][r].re += __shfl_xor(d1[i][r].re, o, 64);
            d1[i][r].im += __shfl_xor(d1[i][r].im, o, 64);
          }
        }
        if (use2) {
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) {
            d2[i][r].re += __shfl_xor(d2[i][r].re, o, 64);
            d2[i][r].im += __shfl_xor(d2[i][r].im, o, 64);
          }
        }
        if (lr == 0) {
          int m = wm * TM + i * 16 + mfma_traits<T>::row(lk, r);
          red[0][wn][m] = to_zd(d0[i][r]);
          red[1][wn][m] = to_zd(d1[i][r]);
          red[2][wn][m] = to_zd(d2[i][r]);
        }
      }
    __syncthreads();
    // one partial sum per 64-column group (two waves of 32 columns each, summed in wave order), so
    // that every tile shape adds the same terms in the same order
    static_assert(EPI != EPI_GREEN || (TN == 32 && BN % 64 == 0), "Green tiles: 32 columns per wave");
    constexpr int GROUPS = BN / 64, WPG = 64 / TN;
    const int ngroups = (p.N + 63) / 64;
    for (int m = tid; m < BM; m += NT) {
      int gm = m0 + m;
      if (gm >= p.M) continue;
#pragma unroll
      for (int q = 0; q < GROUPS; ++q) {
        const int slot = by * GROUPS + q;
        if (slot >= ngroups) break;
        zd s0 = cmk<double>(0, 0), s1 = s0, s2 = s0;
        for (int w = q * WPG; w < (q + 1) * WPG; ++w) {
          s0 = cadd(s0, red[0][w][m]);
          s1 = cadd(s1, red[1][w][m]);
          s2 = cadd(s2, red[2][w][m]);
        }
        long long off = ((long long)b * ngroups + slot) * p.M + gm;
        p.out0[off] = s0;
        if (p.out1) p.out1[off] = s1;
        if (p.out2) p.out2[off] = s2;
      }
    }
  }
}


// Split-BF16 complex GEMM, op N, plain store (see GemmArgs::split_bf16): 128 x 64 output tile, 4
// waves of 64 x 32, K tiles of 16 (one v_mfma_f32_16x16x16bf16_1k step).  The 3M form as in
// CgemmK: P1 = Re A Re B, P2 = Im A Im B, P3 = (Re A + Im A)(Re B + Im B), each real
// product from three BF16 products (lo*hi, hi*lo, hi*hi).  Each thread loads runs of four
// consecutive k of one row (A) or column (B), splits them once (re, im and re + im, each into hi
// and lo BF16 parts) and stores six packed 8-byte planes, k contiguous; a lane then reads each MFMA
// operand (four k of one row) with one 8-byte LDS load.  MFMA operand layout: lane l holds
// A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]; the accumulator as in the FP32 16x16x4
// form.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
// four values into packed hi and lo BF16 parts
__device__ __forceinline__ void bf16_split4(const float* x, bf16x4_t& hi, bf16x4_t& lo) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const unsigned short h = bf16_rne(x[t]);
    hi[t] = (short)h;
    lo[t] = (short)bf16_rne(x[t] - __uint_as_float((unsigned)h << 16));
  }
}

template <bool SCALE>
__global__ __launch_bounds__(256) void bf16x3_gemm_kernel(GemmArgs p) {
  constexpr int BM = 128, BN = 64, BK = 16, WM = 2, NT = 256;
  constexpr int TM = BM / WM, TN = BN / 2, FM = TM / 16, FN = TN / 16;
  constexpr int AP = BM * (BK / 4) / NT;  // runs of four k per thread: A 2, B 1
  static_assert(BN * (BK / 4) == NT, "one B run per thread");
  typedef cplx<float> ct;
  // [plane: re hi, re lo, im hi, im lo, sum hi, sum lo][row][k]
  __shared__ bf16x4_t As[6][BM][BK / 4];
  __shared__ bf16x4_t Bs[6][BN][BK / 4];

  int bx, by, b;
  {
    const unsigned long long gx = gridDim.x, gy = gridDim.y;
    const unsigned long long total = gx * gy * gridDim.z;
    const unsigned long long L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned long long per = total / 8;
    const unsigned long long lg = (L < per * 8) ? (L % 8) * per + L / 8 : L;
    bx = (int)(lg % gx);
    const unsigned long long r = lg / gx;
    by = (int)(r % gy);
    b = (int)(r / gy);
  }
  if (p.alive && !p.alive[b]) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int m0 = bx * BM, n0 = by * BN;
  const ct* __restrict__ A = reinterpret_cast<const ct*>(p.A) + (long long)b * p.strideA;
  const ct* __restrict__ Bm = reinterpret_cast<const ct*>(p.B) + (long long)b * p.strideB;
  const ct* __restrict__ S = SCALE ? reinterpret_cast<const ct*>(p.scale) + (long long)b * p.scale_strideB : nullptr;

  typedef mfma_traits<float>::acc_t acc_t;
  acc_t p1[FM][FN], p2[FM][FN], p3[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      p1[i][j] = acc_t{0, 0, 0, 0};
      p2[i][j] = acc_t{0, 0, 0, 0};
      p3[i][j] = acc_t{0, 0, 0, 0};
    }

  // A runs: run e of a thread is row a_m, k group a_g + e * (NT / BM); B run: column b_n, k group b_g
  const int a_m = tid % BM, a_g = tid / BM;
  const int b_n = tid / (BK / 4), b_g = tid % (BK / 4);
  const bool a_row_in = m0 + a_m < p.M, b_col_in = n0 + b_n < p.N;
  const ct* a_ptr = A + (m0 + a_m);
  const ct* b_ptr = Bm + (long long)(n0 + b_n) * p.ldb;
  const ct* s_ptr = SCALE ? S + (long long)((n0 + b_n) / p.scale_group) * p.scale_strideG : nullptr;
  ct ra[AP][4], rb[4], rs[SCALE ? 4 : 1];
  bool ina[AP][4], inb[4];
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int e = 0; e < AP; ++e)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gk = k0 + 4 * (a_g + e * (NT / BM)) + t;
        const bool in = a_row_in && gk < p.K;
        ina[e][t] = in;
        ra[e][t] = in ? a_ptr[(long long)gk * p.lda] : A[0];
      }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gk = k0 + 4 * b_g + t;
      const bool in = b_col_in && gk < p.K;
      inb[t] = in;
      rb[t] = in ? b_ptr[gk] : Bm[0];
      if (SCALE) rs[t] = in ? s_ptr[gk] : S[0];
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < AP; ++e) {
      float re[4], im[4], sm[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const ct v = ina[e][t] ? ra[e][t] : cmk<float>(0.f, 0.f);
        re[t] = v.re; im[t] = v.im; sm[t] = v.re + v.im;
      }
      const int g = a_g + e * (NT / BM);
      bf16_split4(re, As[0][a_m][g], As[1][a_m][g]);
      bf16_split4(im, As[2][a_m][g], As[3][a_m][g]);
      bf16_split4(sm, As[4][a_m][g], As[5][a_m][g]);
    }
    float re[4], im[4], sm[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      ct v = cmk<float>(0.f, 0.f);
      if (inb[t]) v = SCALE ? cmul(rs[t], rb[t]) : rb[t];
      re[t] = v.re; im[t] = v.im; sm[t] = v.re + v.im;
    }
    bf16_split4(re, Bs[0][b_n][b_g], Bs[1][b_n][b_g]);
    bf16_split4(im, Bs[2][b_n][b_g], Bs[3][b_n][b_g]);
    bf16_split4(sm, Bs[4][b_n][b_g], Bs[5][b_n][b_g]);
  };

  const int nk = (p.K + BK - 1) / BK;
  const int lr = lane & 15, lq = lane >> 4;
  load_tiles(0);
  store_tiles();
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + 1 < nk) load_tiles((kt + 1) * BK);
    bf16x4_t af[FM][6], bf[FN][6];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int q = 0; q < 6; ++q) af[i][q] = As[q][wm * TM + i * 16 + lr][lq];
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 6; ++q) bf[j][q] = Bs[q][wn * TN + j * 16 + lr][lq];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        p1[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][1], bf[j][0], p1[i][j], 0, 0, 0);
        p1[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][0], bf[j][1], p1[i][j], 0, 0, 0);
        p1[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][0], bf[j][0], p1[i][j], 0, 0, 0);
        p2[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][3], bf[j][2], p2[i][j], 0, 0, 0);
        p2[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][2], bf[j][3], p2[i][j], 0, 0, 0);
        p2[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][2], bf[j][2], p2[i][j], 0, 0, 0);
        p3[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][5], bf[j][4], p3[i][j], 0, 0, 0);
        p3[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][4], bf[j][5], p3[i][j], 0, 0, 0);
        p3[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i][4], bf[j][4], p3[i][j], 0, 0, 0);
      }
    __syncthreads();
    if (kt + 1 < nk) {
      store_tiles();
      __syncthreads();
    }
  }

  ct* C = reinterpret_cast<ct*>(p.C) + (long long)b * p.strideC;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn * TN + j * 16 + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * TM + i * 16 + mfma_traits<float>::row(lq, r);
        if (gm < p.M && gn < p.N) {
          const float x1 = p1[i][j][r], x2 = p2[i][j][r], x3 = p3[i][j][r];
          const double xr = (double)(x1 - x2), xi = (double)((x3 - x1) - x2);
          const double yr = p.alpha_re * xr - p.alpha_im * xi;
          const double yi = p.alpha_re * xi + p.alpha_im * xr;
          C[gm + (long long)gn * p.ldc] = cmk<float>((float)yr, (float)yi);
        }
      }
    }
}

// whether the kernel may load its whole K tiles through buffer resources: every byte offset from a
// workgroup's operand base (plus the K advance) below 2^31
template <typename T, int OPA, int BM, int BN, int BK>
bool fast_loads_ok(const GemmArgs& a, bool scale) {
  if (a.K < BK) return false;
  const long long lim = (1ll << 31) - 1;
  const long long ce = (long long)sizeof(cplx<T>);
  const long long ae = OPA == OP_NR ? (long long)sizeof(T) : ce;
  const bool nmaj = OPA == OP_N || OPA == OP_NR;
  const long long a_span = nmaj ? (long long)BM + (long long)a.K * a.lda : (long long)a.K + (long long)BM * a.lda;
  const long long b_span = (long long)a.K + (long long)BN * a.ldb;
  if (a_span * ae > lim || b_span * ce > lim) return false;
  if (scale) {
    const long long s_span = (long long)((a.N + a.scale_group - 1) / a.scale_group) * a.scale_strideG + a.K;
    if (s_span * ce > lim) return false;
  }
  return true;
}

template <typename T, int OPA, int BM, int BN, int BK, int WM, int WN, int EPI, bool SCALE>
int run(const GemmArgs& a0, hipStream_t s) {
  dim3 grid((a0.M + BM - 1) / BM, (a0.N + BN - 1) / BN, a0.batch);
  if (grid.y > 65535 || grid.z > 65535) {
    afqmc_set_error("cgemm grid too large (%u, %u, %u)", grid.x, grid.y, grid.z);
    return -1;
  }
  GemmArgs a = a0;
  a.fast_loads = fast_loads_ok<T, OPA, BM, BN, BK>(a0, SCALE) ? 1 : 0;
  return glaunch<CgemmK<T, OPA, BM, BN, BK, WM, WN, EPI, SCALE>>(a, grid, dim3(WM * WN * 64), 0, s);
}

template <typename T, int OPA, int EPI, bool SCALE>
int dispatch_tile(const GemmArgs& a, hipStream_t s) {
  // 4 waves (2 x 2) unless the shape says 8; K tile 16 unless K32
  switch (a.tile) {
    case GEMM_TILE_32x64: return run<T, OPA, 32, 64, 16, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_64x32: return run<T, OPA, 64, 32, 16, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_32x32: return run<T, OPA, 32, 32, 16, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_128x64: return run<T, OPA, 128, 64, 16, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_64x64_K32: return run<T, OPA, 64, 64, 32, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_64x32_K32: return run<T, OPA, 64, 32, 32, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_32x64_K32: return run<T, OPA, 32, 64, 32, 2, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_128x128: return run<T, OPA, 128, 128, 16, 2, 4, EPI, SCALE>(a, s);
    case GEMM_TILE_256x64: return run<T, OPA, 256, 64, 16, 4, 2, EPI, SCALE>(a, s);
    case GEMM_TILE_64x128: return run<T, OPA, 64, 128, 16, 2, 2, EPI, SCALE>(a, s);
    default: return run<T, OPA, 64, 64, 16, 2, 2, EPI, SCALE>(a, s);
  }
}

// call sites without tuned shapes use the default tile only (fewer instantiations)
template <typename T, int OPA, int EPI, bool SCALE>
int dispatch_default(const GemmArgs& a, hipStream_t s) {
  return run<T, OPA, 64, 64, 16, 2, 2, EPI, SCALE>(a, s);
}

template <typename T>
int dispatch_green(const GemmArgs& a, hipStream_t s) {
  if (a.tile == GEMM_TILE_32x64) return run<T, OP_N, 32, 64, 16, 2, 2, EPI_GREEN, false>(a, s);
  if (a.tile == GEMM_TILE_128x64) return run<T, OP_N, 128, 64, 16, 2, 2, EPI_GREEN, false>(a, s);
  if (a.tile == GEMM_TILE_64x64_K32) return run<T, OP_N, 64, 64, 32, 2, 2, EPI_GREEN, false>(a, s);
  if (a.tile == GEMM_TILE_32x64_K32) return run<T, OP_N, 32, 64, 32, 2, 2, EPI_GREEN, false>(a, s);
  if (a.tile == GEMM_TILE_256x64) return run<T, OP_N, 256, 64, 16, 4, 2, EPI_GREEN, false>(a, s);
  if (a.tile == GEMM_TILE_128x128) return run<T, OP_N, 128, 128, 16, 2, 4, EPI_GREEN, false>(a, s);
  return run<T, OP_N, 64, 64, 16, 2, 2, EPI_GREEN, false>(a, s);
}

}  // namespace

template <typename T>
int launch_cgemm(int opa, int epi, const GemmArgs& a, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return 0;
  REQUIRE(a.K >= 0, "cgemm: negative K");
  const bool sc = a.scale != nullptr;
  REQUIRE(!sc || a.scale_group > 0, "cgemm: scale_group must be positive");
  REQUIRE(epi == EPI_STORE || (a.out0 && a.P0), "cgemm: green epilogue needs P0/out0");
  REQUIRE(a.tile >= 0 && a.tile < GEMM_TILE_COUNT, "cgemm: unknown tile shape %d", a.tile);
  REQUIRE(epi != EPI_GREEN || gemm_tile_green(a.tile), "cgemm: tile shape %d has no Green epilogue", a.tile);
  REQUIRE(!a.herm || (opa == OP_C && epi == EPI_STORE && !a.beta_one && a.alpha_im == 0.0 && a.M == a.N &&
                      gemm_tile_square(a.tile)),
          "cgemm: a Hermitian result needs op C, a square tile, M == N, real alpha and a plain store");
  REQUIRE(!a.split_bf16 || (std::is_same<T, float>::value && opa == OP_N && epi == EPI_STORE && !a.beta_one &&
                            !a.herm && !a.b_upper),
          "cgemm: split_bf16 takes FP32, op N and a plain store");
  if (a.split_bf16) {
    dim3 grid((a.M + 127) / 128, (a.N + 63) / 64, a.batch);
    if (grid.y > 65535 || grid.z > 65535) {
      afqmc_set_error("cgemm grid too large (%u, %u, %u)", grid.x, grid.y, grid.z);
      return -1;
    }
    if (sc) hipLaunchKernelGGL(bf16x3_gemm_kernel<true>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(bf16x3_gemm_kernel<false>, grid, dim3(256), 0, s, a);
    HIP_LAUNCH_CHECK();
    return 0;
  }
  if (epi == EPI_STORE) {
    if (opa == OP_N) return sc ? dispatch_tile<T, OP_N, EPI_STORE, true>(a, s) : dispatch_tile<T, OP_N, EPI_STORE, false>(a, s);
    if (opa == OP_T) return dispatch_tile<T, OP_T, EPI_STORE, false>(a, s);
    if (opa == OP_C) return dispatch_tile<T, OP_C, EPI_STORE, false>(a, s);
    if (opa == OP_NR) return dispatch_default<T, OP_NR, EPI_STORE, false>(a, s);
  } else {
    if (opa == OP_N) return dispatch_green<T>(a, s);
  }
  afqmc_set_error("cgemm: unsupported op/epilogue combination (%d, %d)", opa, epi);
  return -1;
}

template int launch_cgemm<double>(int, int, const GemmArgs&, hipStream_t);
template int launch_cgemm<float>(int, int, const GemmArgs&, hipStream_t);
