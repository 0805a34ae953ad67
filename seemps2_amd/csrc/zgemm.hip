// Complex128 GEMM on the FP64 matrix cores of gfx950 (v_mfma_f64_16x16x4f64).
//
// C = alpha * op(A) @ op(B) + beta * C for strided batches, row-major storage.
// A complex product is three real MFMA products ("3M", Gauss): t1 += ar*br,
// t2 += ai*bi, t3 += (ar+ai)*(br+bi), then re = t1 - t2, im = t3 - t1 - t2;
// or four ("4M"): re += ar*br - ai*bi, im += ar*bi + ai*br.  3M issues a
// quarter fewer MFMAs for the same output; its rounding error is bounded, like
// 4M's, by a small multiple of u * sum |a||b| (normwise), though the imaginary
// part alone is not resolved to relative precision when it is much smaller
// than the real part.  The FP64 32 x 32 tile uses 3M unless
// SEEMPS_GEMM_SCHEME=4m; the 64 x 64 tile (no registers to spare for a third
// accumulator) and FP32 use 4M.
//
// Operand tiles are staged in LDS split into real and imaginary planes,
// K-major, so each MFMA fragment read is a contiguous run of 16 doubles; the
// op(.) transposition / conjugation is folded into the global->LDS load (no
// transposed copies are materialised).
//
// Tiles: 64 x 64 or 32 x 32 outputs per 256-thread workgroup (4 waves of a
// quarter each), chosen per call to spread the tiles evenly over the CUs;
// K step 16, next K-tile prefetched into registers while the current one is used.
//
// MFMA f64 16x16x4 lane layout (checked on the device): A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, result item r of lane l is D[l/16 + 4r][l%16].
#include "common.h"

namespace seemps {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Arithmetic of the product: FP64 MFMA, or FP32 MFMA with FP32 accumulation
// (operands rounded to FP32 on their way into LDS; C stays complex128).
// Output row of accumulator item r on a lane with k-group lk: f64 16x16x4 gives
// D[lk + 4r][l%16], f32 16x16x4 gives D[4 lk + r][l%16] (both checked on the
// device); the operand layouts are the same.
struct PrecF64 {
  typedef double T;
  typedef d4 V;
  static __device__ __forceinline__ V mfma(T a, T b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
struct PrecF32 {
  typedef float T;
  typedef f4 V;
  static __device__ __forceinline__ V mfma(T a, T b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int BK = 16, PAD = 4;
constexpr int BM = 64, BN = 64;  // largest tile: split-K decisions are taken on it

// Tile shapes of the FP64 kernels (the FP32 kernels use 64 x 64 only).
enum TileShape { TILE_64x64 = 0, TILE_32x32 = 1 };

// Load the (TM x BK) tile of op(A) starting at (m0, k0) into registers:
// TM / 16 complex per thread.  For op in {N, C} A is stored M x K; for {T, H}
// it is stored K x M.
// CHECK = false: the tile lies inside op(A) (no bounds tests).
template <int OP, int TM, bool CHECK>
__device__ __forceinline__ void load_a(const double2* __restrict__ A, int64_t lda, int64_t M,
                                       int64_t K, int64_t m0, int64_t k0,
                                       double2 (&r)[TM / 16]) {
  constexpr int E = TM / 16, TPR = BK / E;  // elements per thread, threads per row
  const int t = threadIdx.x;
  if (OP == SEEMPS_OP_N || OP == SEEMPS_OP_C) {
    const int64_t m = m0 + t / TPR;
    const int64_t kb = k0 + (t % TPR) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t k = kb + u;
      r[u] = (!CHECK || (m < M && k < K)) ? A[m * lda + k] : make_double2(0.0, 0.0);
    }
  } else {
    const int64_t k = k0 + (t >> 4);
    const int64_t mb = m0 + (t & 15) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t m = mb + u;
      r[u] = (!CHECK || (m < M && k < K)) ? A[k * lda + m] : make_double2(0.0, 0.0);
    }
  }
}

template <int OP, int TM, typename T>
__device__ __forceinline__ void store_a(T* __restrict__ s_re, T* __restrict__ s_im,
                                        const double2 (&r)[TM / 16]) {
  constexpr int E = TM / 16, TPR = BK / E, LD = TM + PAD;
  const int t = threadIdx.x;
  const double sg = (OP == SEEMPS_OP_C || OP == SEEMPS_OP_H) ? -1.0 : 1.0;
  if (OP == SEEMPS_OP_N || OP == SEEMPS_OP_C) {
    const int m = t / TPR, kb = (t % TPR) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      s_re[(kb + u) * LD + m] = (T)r[u].x;
      s_im[(kb + u) * LD + m] = (T)(sg * r[u].y);
    }
  } else {
    const int k = t >> 4, mb = (t & 15) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      s_re[k * LD + mb + u] = (T)r[u].x;
      s_im[k * LD + mb + u] = (T)(sg * r[u].y);
    }
  }
}

// (BK x TN) tile of op(B) at (k0, n0).  For op in {N, C} B is stored K x N;
// for {T, H} it is stored N x K.
template <int OP, int TN, bool CHECK>
__device__ __forceinline__ void load_b(const double2* __restrict__ B, int64_t ldb, int64_t K,
                                       int64_t N, int64_t k0, int64_t n0,
                                       double2 (&r)[TN / 16]) {
  constexpr int E = TN / 16, TPR = BK / E;
  const int t = threadIdx.x;
  if (OP == SEEMPS_OP_N || OP == SEEMPS_OP_C) {
    const int64_t k = k0 + (t >> 4);
    const int64_t nb = n0 + (t & 15) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t n = nb + u;
      r[u] = (!CHECK || (k < K && n < N)) ? B[k * ldb + n] : make_double2(0.0, 0.0);
    }
  } else {
    const int64_t n = n0 + t / TPR;
    const int64_t kb = k0 + (t % TPR) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int64_t k = kb + u;
      r[u] = (!CHECK || (k < K && n < N)) ? B[n * ldb + k] : make_double2(0.0, 0.0);
    }
  }
}

template <int OP, int TN, typename T>
__device__ __forceinline__ void store_b(T* __restrict__ s_re, T* __restrict__ s_im,
                                        const double2 (&r)[TN / 16]) {
  constexpr int E = TN / 16, TPR = BK / E, LD = TN + PAD;
  const int t = threadIdx.x;
  const double sg = (OP == SEEMPS_OP_C || OP == SEEMPS_OP_H) ? -1.0 : 1.0;
  if (OP == SEEMPS_OP_N || OP == SEEMPS_OP_C) {
    const int k = t >> 4, nb = (t & 15) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      s_re[k * LD + nb + u] = (T)r[u].x;
      s_im[k * LD + nb + u] = (T)(sg * r[u].y);
    }
  } else {
    const int n = t / TPR, kb = (t % TPR) * E;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      s_re[(kb + u) * LD + n] = (T)r[u].x;
      s_im[(kb + u) * LD + n] = (T)(sg * r[u].y);
    }
  }
}

// Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b mod 8).
// Renumber them so that each XCD takes one contiguous run of output tiles
// (row-major over the tile grid): the runs share A row panels inside one L2.
// A bijection on [0, n) for any n.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t n) {
  constexpr int X = 8;
  const int64_t q = n / X, r = n % X;
  const int64_t x = b % X, i = b / X;
  return x * q + (x < r ? x : r) + i;
}

// Workgroups per CU each tile shape is compiled for (bounds the registers).
// The 3M 32 x 32 kernel holds a third accumulator and the operand sums: five
// workgroups per CU keep it free of spills.
constexpr int tile_occupancy(int tm, int tn, bool g3) {
  return tm == 64 && tn == 64 ? 3 : (g3 ? 5 : 6);
}

// One workgroup of 4 waves per TM x TN output tile; the waves split it 2 x 2,
// each owning (TM/2) x (TN/2) = FA x FB fragments of 16 x 16.  Batch index z
// uses K (the last batch uses Klast: split-K slices share one launch, the
// final slice being shorter).  Tiles are numbered along blockIdx.x,
// row-major over the ceil(M/TM) x ceil(N/TN) grid.
// 64 x 64: at most 168 registers per lane (three waves per SIMD, no spills).
template <int OPA, int OPB, typename P, int TM, int TN, bool G3>
__global__ __launch_bounds__(256, tile_occupancy(TM, TN, G3)) void zgemm_kernel(
    int64_t M, int64_t N, int64_t Kfull, int64_t Klast, double2 alpha,
    const double2* __restrict__ A, int64_t lda, int64_t sA, const double2* __restrict__ B,
    int64_t ldb, int64_t sB, double2 beta, double2* __restrict__ C, int64_t ldc, int64_t sC) {
  typedef typename P::T T;
  typedef typename P::V V;
  constexpr int LA = TM + PAD, LB = TN + PAD;
  constexpr int FA = TM / 32, FB = TN / 32;
  __shared__ T as_re[BK * LA], as_im[BK * LA];
  __shared__ T bs_re[BK * LB], bs_im[BK * LB];

  const int64_t bz = blockIdx.z;
  const int64_t K = (bz == (int64_t)gridDim.z - 1) ? Klast : Kfull;
  A += bz * sA;
  B += bz * sB;
  C += bz * sC;
  const int64_t tiles_n = ceil_div(N, TN);
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t m0 = (tile / tiles_n) * TM;
  const int64_t n0 = (tile % tiles_n) * TN;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * (TM / 2), wn = (wave & 1) * (TN / 2);
  const int li = lane & 15, lk = lane >> 4;

  // 4M: cr, ci hold re, im.  3M: cr, ci, cs hold t1, t2, t3.
  constexpr int SA = G3 ? FA : 1, SB = G3 ? FB : 1;
  V cr[FA][FB], ci[FA][FB], cs[SA][SB];
#pragma unroll
  for (int a = 0; a < FA; ++a)
#pragma unroll
    for (int b = 0; b < FB; ++b) {
      cr[a][b] = V{0, 0, 0, 0};
      ci[a][b] = V{0, 0, 0, 0};
    }
#pragma unroll
  for (int a = 0; a < SA; ++a)
#pragma unroll
    for (int b = 0; b < SB; ++b) cs[a][b] = V{0, 0, 0, 0};

  // K-tiles of interior output tiles are loaded without bounds tests
  const bool interior = (m0 + TM <= M) && (n0 + TN <= N);
  double2 ra[TM / 16], rb[TN / 16];
  auto load = [&](int64_t k0) {
    if (interior && k0 + BK <= K) {
      load_a<OPA, TM, false>(A, lda, M, K, m0, k0, ra);
      load_b<OPB, TN, false>(B, ldb, K, N, k0, n0, rb);
    } else {
      load_a<OPA, TM, true>(A, lda, M, K, m0, k0, ra);
      load_b<OPB, TN, true>(B, ldb, K, N, k0, n0, rb);
    }
  };
  load(0);
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    store_a<OPA, TM, T>(as_re, as_im, ra);
    store_b<OPB, TN, T>(bs_re, bs_im, rb);
    __syncthreads();
    if (k0 + BK < K) load(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int k = ks * 4 + lk;
      T ar[FA], ai[FA], br[FB], bi[FB];
#pragma unroll
      for (int a = 0; a < FA; ++a) {
        ar[a] = as_re[k * LA + wm + a * 16 + li];
        ai[a] = as_im[k * LA + wm + a * 16 + li];
      }
#pragma unroll
      for (int b = 0; b < FB; ++b) {
        br[b] = bs_re[k * LB + wn + b * 16 + li];
        bi[b] = bs_im[k * LB + wn + b * 16 + li];
      }
      if constexpr (G3) {
        // three independent chains: no product waits on the one before it
#pragma unroll
        for (int a = 0; a < SA; ++a)
#pragma unroll
          for (int b = 0; b < SB; ++b) {
            cr[a][b] = P::mfma(ar[a], br[b], cr[a][b]);
            ci[a][b] = P::mfma(ai[a], bi[b], ci[a][b]);
            cs[a][b] = P::mfma(ar[a] + ai[a], br[b] + bi[b], cs[a][b]);
          }
      } else {
        // the real and imaginary chains alternate, so no product waits on the
        // one issued just before it
#pragma unroll
        for (int a = 0; a < FA; ++a)
#pragma unroll
          for (int b = 0; b < FB; ++b) {
            cr[a][b] = P::mfma(ar[a], br[b], cr[a][b]);
            ci[a][b] = P::mfma(ar[a], bi[b], ci[a][b]);
          }
#pragma unroll
        for (int a = 0; a < FA; ++a)
#pragma unroll
          for (int b = 0; b < FB; ++b) {
            cr[a][b] = P::mfma(-ai[a], bi[b], cr[a][b]);
            ci[a][b] = P::mfma(ai[a], br[b], ci[a][b]);
          }
      }
    }
    __syncthreads();
  }

  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
#pragma unroll
  for (int a = 0; a < FA; ++a)
#pragma unroll
    for (int b = 0; b < FB; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + a * 16 + P::row(lk, r);
        const int64_t n = n0 + wn + b * 16 + li;
        if (m < M && n < N) {
          double2 acc;
          if constexpr (G3) {
            const double t1 = cr[a][b][r], t2 = ci[a][b][r];
            acc = make_double2(t1 - t2, ((double)cs[a][b][r] - t1) - t2);
          } else {
            acc = make_double2((double)cr[a][b][r], (double)ci[a][b][r]);
          }
          double2 out = cmul(alpha, acc);
          if (use_beta) out = cfma(beta, C[m * ldc + n], out);
          C[m * ldc + n] = out;
        }
      }
}

// Tile shape for an M x N output in `batch` slices on `num_cu` CUs.  The
// 64 x 64 tile (three workgroups per CU) is the faster one per output element
// once every CU holds four or more of them; below that the 32 x 32 tile (five
// per CU with the three-product scheme, six with four; four times as many
// tiles) spreads the work more evenly and keeps more waves per SIMD to hide
// the LDS and memory latencies.  Measured with four products on the matvec
// and environment GEMMs of chi = 128 and 256 (25-35 % faster with the small
// tile) and on a 2048 x 2048 x 512 product (6 % slower with it).
// SEEMPS_GEMM_TILE = 64x64 or 32x32 fixes the shape (a tuning knob).
int choose_tile(int64_t M, int64_t N, int64_t batch, int num_cu) {
  static const int forced = [] {
    const char* e = getenv("SEEMPS_GEMM_TILE");
    if (!e) return -1;
    if (!strcmp(e, "64x64")) return (int)TILE_64x64;
    if (!strcmp(e, "32x32")) return (int)TILE_32x32;
    return -1;
  }();
  if (forced >= 0) return forced;
  const int64_t cu = num_cu > 0 ? num_cu : 256;
  const int64_t tiles64 = ceil_div(M, 64) * ceil_div(N, 64) * batch;
  return tiles64 >= 4 * cu ? TILE_64x64 : TILE_32x32;
}

// FP64 products use the 3M scheme unless SEEMPS_GEMM_SCHEME=4m (a tuning and
// ablation knob, read once).
bool gemm_scheme_3m() {
  static const bool g3 = [] {
    const char* e = getenv("SEEMPS_GEMM_SCHEME");
    return !(e && !strcmp(e, "4m"));
  }();
  return g3;
}

template <int OPA, int OPB>
void launch(int prec, int tile, int64_t batch, hipStream_t s, int64_t M, int64_t N, int64_t K,
            int64_t Klast, double2 alpha, const double2* A, int64_t lda, int64_t sA,
            const double2* B, int64_t ldb, int64_t sB, double2 beta, double2* C, int64_t ldc,
            int64_t sC) {
  auto grid = [&](int tm, int tn) {
    return dim3((unsigned)(ceil_div(M, tm) * ceil_div(N, tn)), 1u, (unsigned)batch);
  };
  if (prec == SEEMPS_PREC_FP32)
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, PrecF32, 64, 64, false>), grid(64, 64), dim3(256),
                       0, s, M, N, K, Klast, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
  else if (gemm_scheme_3m() && tile == TILE_32x32)
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, PrecF64, 32, 32, true>), grid(32, 32), dim3(256),
                       0, s, M, N, K, Klast, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
  else if (tile == TILE_32x32)
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, PrecF64, 32, 32, false>), grid(32, 32), dim3(256),
                       0, s, M, N, K, Klast, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
  else
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, PrecF64, 64, 64, false>), grid(64, 64), dim3(256),
                       0, s, M, N, K, Klast, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
}

// C[m, n] = alpha * sum_z P[z][m, n] + beta * C[m, n]  (P packed M x N slices).
__global__ void splitk_reduce(int64_t M, int64_t N, int nz, const double2* __restrict__ P,
                              double2 alpha, double2 beta, double2* __restrict__ C, int64_t ldc) {
  const int64_t total = M * N;
  const bool use_beta = (beta.x != 0.0 || beta.y != 0.0);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double2 s = P[e];
    for (int z = 1; z < nz; ++z) s = cadd(s, P[z * total + e]);
    const int64_t m = e / N, n = e % N;
    double2 out = cmul(alpha, s);
    if (use_beta) out = cfma(beta, C[m * ldc + n], out);
    C[m * ldc + n] = out;
  }
}

// flops: algorithmic count of the whole product (for kernel timing).
int dispatch(int prec, int opA, int opB, int64_t batch, hipStream_t stream, int64_t M, int64_t N,
             int64_t K, int64_t Klast, double2 alpha, const double2* a, int64_t lda, int64_t sA,
             const double2* b, int64_t ldb, int64_t sB, double2 beta, double2* c, int64_t ldc,
             int64_t sC, double flops) {
  DeviceInfo info;
  const int ist = device_info(&info);
  if (ist) return ist;
  const int tile = choose_tile(M, N, batch, info.num_cu);
  // the tile index runs along grid.x (at most 2^31 - 1 workgroups)
  SEEMPS_CHECK_ARG(ceil_div(M, 32) * ceil_div(N, 32) < (int64_t)INT32_MAX,
                   "zgemm: output too large");
  TimedLaunch timed(prec == SEEMPS_PREC_FP32 ? "zgemm_kernel_fp32" : "zgemm_kernel", stream,
                    flops);
#define SEEMPS_GEMM_CASE(x, y)                                                              \
  if (opA == x && opB == y) {                                                               \
    launch<x, y>(prec, tile, batch, stream, M, N, K, Klast, alpha, a, lda, sA, b, ldb, sB,   \
                 beta, c, ldc, sC);                                                         \
    return check_launch("zgemm");                                                           \
  }
  SEEMPS_GEMM_CASE(0, 0) SEEMPS_GEMM_CASE(0, 1) SEEMPS_GEMM_CASE(0, 2) SEEMPS_GEMM_CASE(0, 3)
  SEEMPS_GEMM_CASE(1, 0) SEEMPS_GEMM_CASE(1, 1) SEEMPS_GEMM_CASE(1, 2) SEEMPS_GEMM_CASE(1, 3)
  SEEMPS_GEMM_CASE(2, 0) SEEMPS_GEMM_CASE(2, 1) SEEMPS_GEMM_CASE(2, 2) SEEMPS_GEMM_CASE(2, 3)
  SEEMPS_GEMM_CASE(3, 0) SEEMPS_GEMM_CASE(3, 1) SEEMPS_GEMM_CASE(3, 2) SEEMPS_GEMM_CASE(3, 3)
#undef SEEMPS_GEMM_CASE
  set_error("zgemm: invalid op combination");
  return SEEMPS_ERR_INVALID;
}

}  // namespace

// Number of K slices zgemm_ws uses for (M, N, K) (1 = no split).
int64_t splitk_slices(int64_t M, int64_t N, int64_t K) {
  // tuning knob: tile count from which no split is made (SEEMPS_SPLITK_TILES);
  // the slices aim at twice that many workgroups
  static const int64_t limit = [] {
    const char* e = getenv("SEEMPS_SPLITK_TILES");
    const long v = e ? atol(e) : 0;
    return (int64_t)(v > 0 ? v : 192);
  }();
  const int64_t tiles = ceil_div(M, BM) * ceil_div(N, BN);
  if (tiles >= limit || K < 256) return 1;
  int64_t z = ceil_div(2 * limit, tiles);
  if (z > 8) z = 8;
  while (z > 1 && K / z < 64) --z;
  return z;
}

int64_t splitk_workspace(int64_t M, int64_t N, int64_t K) {
  const int64_t z = splitk_slices(M, N, K);
  return z > 1 ? z * M * N : 0;
}

// zgemm with split-K over the caller's workspace when the output has too few
// tiles to fill the device (deterministic: slices reduced in a fixed order).
int zgemm_ws(int opA, int opB, int64_t M, int64_t N, int64_t K, double2 alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, double2 beta, void* C, int64_t ldc,
             double2* ws, int64_t ws_elems, hipStream_t stream, int prec) {
  const int64_t z = splitk_slices(M, N, K);
  if (z <= 1 || ws == nullptr || ws_elems < z * M * N)
    return zgemm(opA, opB, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1, stream,
                 prec);
  int64_t Kc = ceil_div(ceil_div(K, z), BK) * BK;
  const int64_t nz = ceil_div(K, Kc);
  const int64_t Klast = K - (nz - 1) * Kc;
  const int64_t sA = (opA == SEEMPS_OP_N || opA == SEEMPS_OP_C) ? Kc : Kc * lda;
  const int64_t sB = (opB == SEEMPS_OP_N || opB == SEEMPS_OP_C) ? Kc * ldb : Kc;
  int st = dispatch(prec, opA, opB, nz, stream, M, N, Kc, Klast, make_double2(1.0, 0.0),
                    (const double2*)A, lda, sA, (const double2*)B, ldb, sB,
                    make_double2(0.0, 0.0), ws, N, M * N, 8.0 * M * N * K);
  if (st) return st;
  int64_t blocks = ceil_div(M * N, 256);
  if (blocks > 4096) blocks = 4096;
  TimedLaunch timed("splitk_reduce", stream, 0.0);
  hipLaunchKernelGGL(splitk_reduce, dim3((unsigned)blocks), dim3(256), 0, stream, M, N, (int)nz,
                     ws, alpha, beta, (double2*)C, ldc);
  return check_launch("splitk_reduce");
}

// Enqueue a complex GEMM (shared by the other host entry points of the library).
int zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K, double2 alpha, const void* A,
          int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB, double2 beta, void* C,
          int64_t ldc, int64_t sC, int64_t batch, hipStream_t stream, int prec) {
  SEEMPS_CHECK_ARG(opA >= 0 && opA <= 3 && opB >= 0 && opB <= 3, "zgemm: invalid op flags");
  SEEMPS_CHECK_ARG(prec == SEEMPS_PREC_FP64 || prec == SEEMPS_PREC_FP32,
                   "zgemm: invalid precision");
  SEEMPS_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "zgemm: negative dimension");
  if (M == 0 || N == 0 || batch == 0) return SEEMPS_OK;
  SEEMPS_CHECK_ARG(batch <= 65535, "zgemm: batch too large");
  if (K == 0) {
    // Pure scaling of C by beta (or zero fill).
    SEEMPS_CHECK_ARG(ldc == N && (batch == 1 || sC == M * N), "zgemm: K == 0 needs packed C");
    return zaxpby_impl(M * N * batch, make_double2(0.0, 0.0), nullptr, 0, beta, (double2*)C,
                       stream);
  }
  return dispatch(prec, opA, opB, batch, stream, M, N, K, K, alpha, (const double2*)A, lda, sA,
                  (const double2*)B, ldb, sB, beta, (double2*)C, ldc, sC,
                  8.0 * M * N * K * batch);
}

}  // namespace seemps

extern "C" int seemps_zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K, double alpha_re,
                            double alpha_im, const void* A, int64_t lda, int64_t strideA,
                            const void* B, int64_t ldb, int64_t strideB, double beta_re,
                            double beta_im, void* C, int64_t ldc, int64_t strideC,
                            int64_t batch, void* stream) {
  return seemps::zgemm(opA, opB, M, N, K, make_double2(alpha_re, alpha_im), A, lda, strideA, B,
                       ldb, strideB, make_double2(beta_re, beta_im), C, ldc, strideC, batch,
                       (hipStream_t)stream);
}

extern "C" int seemps_zgemm_prec(int precision, int opA, int opB, int64_t M, int64_t N,
                                 int64_t K, double alpha_re, double alpha_im, const void* A,
                                 int64_t lda, int64_t strideA, const void* B, int64_t ldb,
                                 int64_t strideB, double beta_re, double beta_im, void* C,
                                 int64_t ldc, int64_t strideC, int64_t batch, void* stream) {
  return seemps::zgemm(opA, opB, M, N, K, make_double2(alpha_re, alpha_im), A, lda, strideA, B,
                       ldb, strideB, make_double2(beta_re, beta_im), C, ldc, strideC, batch,
                       (hipStream_t)stream, precision);
}

extern "C" int64_t seemps_zgemm_workspace(int64_t M, int64_t N, int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  return seemps::splitk_workspace(M, N, K) * 16;
}

extern "C" int seemps_zgemm_ws(int opA, int opB, int64_t M, int64_t N, int64_t K, double alpha_re,
                               double alpha_im, const void* A, int64_t lda, const void* B,
                               int64_t ldb, double beta_re, double beta_im, void* C, int64_t ldc,
                               void* work, int64_t work_bytes, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(opA >= 0 && opA <= 3 && opB >= 0 && opB <= 3, "zgemm_ws: invalid op flags");
  SEEMPS_CHECK_ARG(M >= 0 && N >= 0 && K >= 0 && work_bytes >= 0, "zgemm_ws: negative size");
  if (M == 0 || N == 0) return SEEMPS_OK;
  if (K == 0)
    return zgemm(opA, opB, M, N, K, make_double2(alpha_re, alpha_im), A, lda, 0, B, ldb, 0,
                 make_double2(beta_re, beta_im), C, ldc, 0, 1, (hipStream_t)stream);
  return zgemm_ws(opA, opB, M, N, K, make_double2(alpha_re, alpha_im), A, lda, B, ldb,
                  make_double2(beta_re, beta_im), C, ldc, (double2*)work, work_bytes / 16,
                  (hipStream_t)stream);
}
