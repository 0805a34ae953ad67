// Memory-bound kernels of the hot path: the small-index "local operator"
// contraction shared by MPO application, MPO environments, the DMRG matvec and
// two-site gates; elementwise scalings; deterministic two-pass reductions.
#include "common.h"

namespace seemps {
namespace {

constexpr int LQ = 64;  // q-values per workgroup of zlocal

// out[p, j1, q, j2] (+)= sum_{i0 <= i < i0 + IC} in[p, i, q] Mop[i, j1, j2] for the
// output columns j = j1 * J2 + j2 in [j0, j0 + JC), j0 = blockIdx.y * JC.
// One workgroup handles (p, 64 consecutive q, one j tile); the IC x 64 input
// slab and the IC x JC operator tile are staged in LDS, 4 threads per q split
// the tile's outputs.  The host runs the i tiles one after another, the later
// ones accumulating, so a contracted index of any size fits in LDS.
__global__ __launch_bounds__(256) void zlocal_kernel(int64_t P, int I, int64_t Q, int J1, int J2,
                                                     int64_t nqb, int i0, int IC, int JC,
                                                     const double2* __restrict__ in,
                                                     const double2* __restrict__ Mop,
                                                     double2* __restrict__ out, int accumulate) {
  extern __shared__ double2 smem[];
  const int J = J1 * J2;
  const int j0 = blockIdx.y * JC;
  const int jn = min(JC, J - j0);
  double2* xs = smem;            // [IC][LQ]
  double2* ms = smem + IC * LQ;  // [IC][JC]
  const int64_t p = blockIdx.x / nqb;
  const int64_t q0 = (blockIdx.x % nqb) * LQ;
  const int tid = threadIdx.x;
  for (int e = tid; e < IC * LQ; e += blockDim.x) {
    const int i = e / LQ, ql = e % LQ;
    const int64_t q = q0 + ql;
    xs[e] = (q < Q) ? in[(p * I + i0 + i) * Q + q] : make_double2(0.0, 0.0);
  }
  for (int e = tid; e < IC * jn; e += blockDim.x) {
    const int i = e / jn, jl = e % jn;
    ms[i * JC + jl] = Mop[(int64_t)(i0 + i) * J + j0 + jl];
  }
  __syncthreads();
  const int ql = tid % LQ;
  const int64_t q = q0 + ql;
  if (q >= Q) return;
  for (int jl = tid / LQ; jl < jn; jl += blockDim.x / LQ) {
    double2 acc = make_double2(0.0, 0.0);
    for (int i = 0; i < IC; ++i) acc = cfma(xs[i * LQ + ql], ms[i * JC + jl], acc);
    const int j = j0 + jl;
    const int j1 = j / J2, j2 = j % J2;
    const int64_t o = ((p * J1 + j1) * Q + q) * J2 + j2;
    if (accumulate) acc = cadd(acc, out[o]);
    out[o] = acc;
  }
}

__global__ void zaxpby_kernel(int64_t n, double2 a, const double2* __restrict__ x, int conj_x,
                              double2 b, double2* __restrict__ y, int use_b) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (x) {
      double2 xv = x[k];
      if (conj_x) xv.y = -xv.y;
      v = cmul(a, xv);
    }
    if (use_b) v = cfma(b, y[k], v);
    y[k] = v;
  }
}

// Site-wise product of two MPS tensors (the MPS Hadamard product):
// out[(x, y), k, (u, v)] = A[x, k, u] * B[y, k, v], A (a, d, b), B (c, d, e).
__global__ void zsite_product_kernel(int64_t a, int64_t d, int64_t b, int64_t c, int64_t e,
                                     const double2* __restrict__ A, const double2* __restrict__ B,
                                     double2* __restrict__ out) {
  const int64_t be = b * e, total = a * c * d * be;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = i % be, row = i / be;  // row = (x * c + y) * d + k
    const int64_t k = row % d, xy = row / d;
    const int64_t x = xy / c, y = xy % c, u = col / e, v = col % e;
    out[i] = cmul(A[(x * d + k) * b + u], B[(y * d + k) * e + v]);
  }
}

__global__ void zscale2d_kernel(int64_t rows, int64_t cols, const double* __restrict__ sr,
                                const double* __restrict__ sc, const double2* __restrict__ src,
                                int64_t ld_src, double2* __restrict__ dst, int64_t ld_dst) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / cols, c = e % cols;
    double f = 1.0;
    if (sr) f *= sr[r];
    if (sc) f *= sc[c];
    const double2 v = src[r * ld_src + c];
    dst[r * ld_dst + c] = make_double2(f * v.x, f * v.y);
  }
}

// Block-wide sum of one complex value per thread (256 threads).
__device__ double2 block_sum(double2 v, double2* red) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double2 s = make_double2(0.0, 0.0);
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = cadd(s, red[w]);
  __syncthreads();
  return s;  // valid in thread 0
}

// Pass 1: partial[blk * nv + j] = sum over this block's chunk of conj(V_j) w.
__global__ __launch_bounds__(256) void zmdotc_partial(int64_t n, int nv,
                                                      const double2* __restrict__ V, int64_t ldv,
                                                      const double2* __restrict__ w,
                                                      double2* __restrict__ partial) {
  __shared__ double2 red[4];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = blockIdx.x * chunk;
  const int64_t hi = min(n, lo + chunk);
  for (int j = 0; j < nv; ++j) {
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t k = lo + threadIdx.x; k < hi; k += blockDim.x)
      acc = cfmac(V[j * ldv + k], w[k], acc);
    const double2 s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x * nv + j] = s;
  }
}

// Pass 2: result[j] = sum_blk partial[blk * nv + j], fixed order.
// result[j] = sum_b partial[b * nv + j]: block j, thread t adds the partials
// b = t, t + 256, ..., then a shuffle and LDS tree (a fixed order).  256 threads.
__global__ __launch_bounds__(256) void zmdotc_final(int nblk, int nv,
                                                    const double2* __restrict__ partial,
                                                    double2* __restrict__ result) {
  __shared__ double2 red[4];
  const int j = blockIdx.x;
  double2 s = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s = cadd(s, partial[b * nv + j]);
  for (int off = 32; off > 0; off >>= 1) {
    s.x += __shfl_down(s.x, off, 64);
    s.y += __shfl_down(s.y, off, 64);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) result[j] = cadd(cadd(red[0], red[1]), cadd(red[2], red[3]));
}

}  // namespace

int64_t reduce_blocks(int64_t n) {
  int64_t b = ceil_div(n, 4096);
  if (b < 1) b = 1;
  if (b > 512) b = 512;
  return b;
}

int zlocal_impl(int64_t P, int64_t I, int64_t Q, int64_t J1, int64_t J2, const double2* in,
                const double2* Mop, double2* out, int accumulate, hipStream_t stream) {
  SEEMPS_CHECK_ARG(P >= 0 && I >= 1 && Q >= 0 && J1 >= 1 && J2 >= 1, "zlocal: bad dimensions");
  SEEMPS_CHECK_ARG(I <= 0x7fffffff && J1 * J2 <= 0x7fffffff, "zlocal: dimensions too large");
  if (P == 0 || Q == 0) return SEEMPS_OK;
  const int64_t J = J1 * J2;
  // LDS tiles: IC x 64 inputs and IC x JC operator entries, at most 3072
  // complex values (48 KiB).  When the whole operator fits (the MPO and
  // physical indices of d = 2 chains) one launch does everything; otherwise
  // the contracted index runs in tiles of 32 and the outputs in tiles of 32.
  int64_t IC = I, JC = J;
  if (I * LQ + I * J > 3072) {
    IC = I < 32 ? I : 32;
    JC = (3072 - IC * LQ) / IC;
    if (JC > J) JC = J;
  }
  const int64_t nqb = ceil_div(Q, LQ);
  const int64_t nblocks = P * nqb;
  const int64_t njb = ceil_div(J, JC);
  SEEMPS_CHECK_ARG(nblocks <= 0x7fffffff && njb <= 65535, "zlocal: grid too large");
  const size_t shmem = (size_t)(IC * LQ + IC * JC) * sizeof(double2);
  TimedLaunch timed("zlocal_kernel", stream, 8.0 * P * I * Q * J);
  for (int64_t i0 = 0; i0 < I; i0 += IC) {
    const int64_t ic = (I - i0) < IC ? (I - i0) : IC;
    hipLaunchKernelGGL(zlocal_kernel, dim3((unsigned)nblocks, (unsigned)njb), dim3(256), shmem,
                       stream, P, (int)I, Q, (int)J1, (int)J2, nqb, (int)i0, (int)ic, (int)JC,
                       in, Mop, out, (i0 > 0) ? 1 : accumulate);
    const int st = check_launch("zlocal");
    if (st) return st;
  }
  return SEEMPS_OK;
}

int zaxpby_impl(int64_t n, double2 a, const double2* x, int conj_x, double2 b, double2* y,
                hipStream_t stream) {
  if (n <= 0) return SEEMPS_OK;
  const int use_b = (b.x != 0.0 || b.y != 0.0);
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zaxpby_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, n, a, x, conj_x,
                     b, y, use_b);
  return check_launch("zaxpby");
}

int zscale2d_impl(int64_t rows, int64_t cols, const double* sr, const double* sc,
                  const double2* src, int64_t ld_src, double2* dst, int64_t ld_dst,
                  hipStream_t stream) {
  if (rows <= 0 || cols <= 0) return SEEMPS_OK;
  int64_t blocks = ceil_div(rows * cols, 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zscale2d_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, rows, cols,
                     sr, sc, src, ld_src, dst, ld_dst);
  return check_launch("zscale2d");
}

int zmdotc_impl(int64_t n, int64_t nv, const double2* V, int64_t ldv, const double2* w,
                double2* result, double2* work, hipStream_t stream) {
  SEEMPS_CHECK_ARG(nv >= 1 && nv <= 64, "zmdotc: number of vectors out of range");
  const int64_t nblk = reduce_blocks(n);
  hipLaunchKernelGGL(zmdotc_partial, dim3((unsigned)nblk), dim3(256), 0, stream, n, (int)nv, V,
                     ldv, w, work);
  int st = check_launch("zmdotc_partial");
  if (st) return st;
  hipLaunchKernelGGL(zmdotc_final, dim3((unsigned)nv), dim3(256), 0, stream, (int)nblk, (int)nv,
                     work, result);
  return check_launch("zmdotc_final");
}

}  // namespace seemps

extern "C" int seemps_zlocal(int64_t P, int64_t I, int64_t Q, int64_t J1, int64_t J2,
                             const void* in, const void* Mop, void* out, int accumulate,
                             void* stream) {
  return seemps::zlocal_impl(P, I, Q, J1, J2, (const double2*)in, (const double2*)Mop,
                             (double2*)out, accumulate, (hipStream_t)stream);
}

extern "C" int seemps_zaxpby(int64_t n, double a_re, double a_im, const void* x, int conj_x,
                             double b_re, double b_im, void* y, void* stream) {
  return seemps::zaxpby_impl(n, make_double2(a_re, a_im), (const double2*)x, conj_x,
                             make_double2(b_re, b_im), (double2*)y, (hipStream_t)stream);
}

extern "C" int seemps_zscale2d(int64_t rows, int64_t cols, const double* sr, const double* sc,
                               const void* src, int64_t ld_src, void* dst, int64_t ld_dst,
                               void* stream) {
  return seemps::zscale2d_impl(rows, cols, sr, sc, (const double2*)src, ld_src, (double2*)dst,
                               ld_dst, (hipStream_t)stream);
}

extern "C" int64_t seemps_reduce_workspace(int64_t n) {
  return seemps::reduce_blocks(n) * (int64_t)sizeof(double2);
}

extern "C" int seemps_zdotc(int64_t n, const void* x, const void* y, void* result, void* work,
                            int64_t work_bytes, void* stream) {
  SEEMPS_CHECK_ARG(n >= 0, "zdotc: negative length");
  SEEMPS_CHECK_ARG(work_bytes >= seemps_reduce_workspace(n), "zdotc: workspace too small");
  if (n == 0) {
    return seemps::zaxpby_impl(1, make_double2(0.0, 0.0), nullptr, 0, make_double2(0.0, 0.0),
                               (double2*)result, (hipStream_t)stream);
  }
  return seemps::zmdotc_impl(n, 1, (const double2*)x, n, (const double2*)y, (double2*)result,
                             (double2*)work, (hipStream_t)stream);
}

extern "C" int seemps_zsite_product(int64_t a, int64_t d, int64_t b, int64_t c, int64_t e,
                                    const void* A, const void* B, void* out, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(a >= 1 && d >= 1 && b >= 1 && c >= 1 && e >= 1,
                   "site_product: bad dimensions");
  const int64_t total = a * c * d * b * e;
  int64_t blocks = ceil_div(total, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(zsite_product_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)stream, a, d, b, c, e, (const double2*)A, (const double2*)B,
                     (double2*)out);
  return check_launch("site_product");
}
