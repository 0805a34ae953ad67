// Grouped complex GEMM: many independent products C_p = op(A_p) op(B_p) of
// uneven shapes in one launch.  The host lists every output tile of every
// problem once (tile_prob[t] = problem of tile t, tiles of a problem in
// row-major order); workgroup t computes tile t.  Tile shapes TM x TN are
// 64 x 64 (four waves) or 32 x 32 (one wave); each wave computes a 32 x 32
// block with the complex "4M" scheme on v_mfma_f64_16x16x4f64, operands
// staged through LDS as real/imaginary planes with the op folded into the
// load.  Used by the batched overlap engine (many MPS of uneven length and
// bond dimension), where the tile shape is chosen by measurement.
#include "common.h"

namespace seemps {
namespace {

// device-side view of SeempsGemmDesc (same layout, typed pointers)
struct GemmDesc {
  int64_t M, N, K;
  const double2* A;
  int64_t lda;
  const double2* B;
  int64_t ldb;
  double2* C;
  int64_t ldc;
  int64_t tiles_n;
  int64_t tile0;
};
static_assert(sizeof(GemmDesc) == sizeof(SeempsGemmDesc), "descriptor layout");

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int GBK = 16;  // K step

template <int OP>
__device__ __forceinline__ double2 load_op(const double2* __restrict__ X, int64_t ld, int64_t r,
                                           int64_t c, int64_t R, int64_t Cn) {
  // element (r, c) of op(X), op(X) being R x Cn; X stored row-major with stride ld
  if (r >= R || c >= Cn) return make_double2(0.0, 0.0);
  double2 v = (OP == SEEMPS_OP_N || OP == SEEMPS_OP_C) ? X[r * ld + c] : X[c * ld + r];
  if (OP == SEEMPS_OP_C || OP == SEEMPS_OP_H) v.y = -v.y;
  return v;
}

template <int OPA, int OPB, int TM, int TN>
__global__ __launch_bounds__((TM / 32) * (TN / 32) * 64) void grouped_zgemm_kernel(
    const GemmDesc* __restrict__ desc, const int32_t* __restrict__ tile_prob) {
  constexpr int NT = (TM / 32) * (TN / 32) * 64;
  constexpr int LA = TM + 1, LB = TN + 1;  // padded LDS rows (k-major)
  __shared__ double as_re[GBK * LA], as_im[GBK * LA];
  __shared__ double bs_re[GBK * LB], bs_im[GBK * LB];
  const int64_t tile = blockIdx.x;
  const GemmDesc d = desc[tile_prob[tile]];
  const int64_t local = tile - d.tile0;
  const int64_t m0 = (local / d.tiles_n) * TM, n0 = (local % d.tiles_n) * TN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave / (TN / 32)) * 32, wn = (wave % (TN / 32)) * 32;
  const int li = lane & 15, lk = lane >> 4;
  d4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  for (int64_t k0 = 0; k0 < d.K; k0 += GBK) {
    // A tile: TM x GBK of op(A); consecutive threads along k for N/C (row-major
    // rows), along m for T/H (the stored matrix's rows run along m)
    for (int e = t; e < TM * GBK; e += NT) {
      int mm, kk;
      if (OPA == SEEMPS_OP_N || OPA == SEEMPS_OP_C) {
        mm = e / GBK;
        kk = e % GBK;
      } else {
        kk = e / TM;
        mm = e % TM;
      }
      const double2 v = load_op<OPA>(d.A, d.lda, m0 + mm, k0 + kk, d.M, d.K);
      as_re[kk * LA + mm] = v.x;
      as_im[kk * LA + mm] = v.y;
    }
    for (int e = t; e < GBK * TN; e += NT) {
      int kk, nn;
      if (OPB == SEEMPS_OP_N || OPB == SEEMPS_OP_C) {
        kk = e / TN;
        nn = e % TN;
      } else {
        nn = e / GBK;
        kk = e % GBK;
      }
      const double2 v = load_op<OPB>(d.B, d.ldb, k0 + kk, n0 + nn, d.K, d.N);
      bs_re[kk * LB + nn] = v.x;
      bs_im[kk * LB + nn] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GBK / 4; ++ks) {
      const int k = ks * 4 + lk;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        ar[a] = as_re[k * LA + wm + a * 16 + li];
        ai[a] = as_im[k * LA + wm + a * 16 + li];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        br[b] = bs_re[k * LB + wn + b * 16 + li];
        bi[b] = bs_im[k * LB + wn + b * 16 + li];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
          cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], cr[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
          ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // f64 16x16x4 output: item r of a lane with k-group lk is row lk + 4 r
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + wm + a * 16 + lk + 4 * r;
        const int64_t n = n0 + wn + b * 16 + li;
        if (m < d.M && n < d.N) d.C[m * d.ldc + n] = make_double2(cr[a][b][r], ci[a][b][r]);
      }
}

template <int OPA, int OPB>
int launch_grouped(int tile_cfg, const GemmDesc* desc, const int32_t* tile_prob, int64_t ntiles,
                   hipStream_t s) {
  if (tile_cfg == SEEMPS_TILE_64x64)
    hipLaunchKernelGGL((grouped_zgemm_kernel<OPA, OPB, 64, 64>), dim3((unsigned)ntiles),
                       dim3(256), 0, s, desc, tile_prob);
  else
    hipLaunchKernelGGL((grouped_zgemm_kernel<OPA, OPB, 32, 32>), dim3((unsigned)ntiles),
                       dim3(64), 0, s, desc, tile_prob);
  return check_launch("grouped_zgemm");
}

}  // namespace
}  // namespace seemps

extern "C" int64_t seemps_grouped_tiles(int tile_cfg, int64_t M, int64_t N) {
  const int64_t t = tile_cfg == SEEMPS_TILE_64x64 ? 64 : 32;
  return seemps::ceil_div(M, t) * seemps::ceil_div(N, t);
}

extern "C" int seemps_zgemm_grouped(int opA, int opB, int tile_cfg, int64_t nprob,
                                    const void* desc, const int32_t* tile_prob, int64_t ntiles,
                                    void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(tile_cfg == SEEMPS_TILE_64x64 || tile_cfg == SEEMPS_TILE_32x32,
                   "zgemm_grouped: invalid tile configuration");
  SEEMPS_CHECK_ARG(nprob >= 0 && ntiles >= 0 && ntiles < (1ll << 31),
                   "zgemm_grouped: invalid problem or tile count");
  if (ntiles == 0) return SEEMPS_OK;
  const GemmDesc* d = (const GemmDesc*)desc;
  hipStream_t s = (hipStream_t)stream;
  TimedLaunch timed("grouped_zgemm", s, 0.0);
#define SEEMPS_GROUPED_CASE(x, y) \
  if (opA == x && opB == y) return launch_grouped<x, y>(tile_cfg, d, tile_prob, ntiles, s);
  SEEMPS_GROUPED_CASE(SEEMPS_OP_N, SEEMPS_OP_N)
  SEEMPS_GROUPED_CASE(SEEMPS_OP_H, SEEMPS_OP_N)
  SEEMPS_GROUPED_CASE(SEEMPS_OP_N, SEEMPS_OP_T)
  SEEMPS_GROUPED_CASE(SEEMPS_OP_T, SEEMPS_OP_N)
  SEEMPS_GROUPED_CASE(SEEMPS_OP_C, SEEMPS_OP_N)
  SEEMPS_GROUPED_CASE(SEEMPS_OP_N, SEEMPS_OP_H)
#undef SEEMPS_GROUPED_CASE
  set_error("zgemm_grouped: op pair (%d, %d) not instantiated", opA, opB);
  return SEEMPS_ERR_INVALID;
}
