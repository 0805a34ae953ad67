// DMRG two-site effective Hamiltonian: operator build, matvec and a Lanczos
// eigensolver whose Krylov recurrence, tridiagonal eigensolve and Ritz vector
// stay on the device (one host synchronisation per restart cycle).
//
// Replaces QuadraticForm.two_site_operator (src/seemps/operators/quadratic.py:
// 187-237) and the scipy eigsh / ARPACK call of _diagonalize_two_site
// (src/seemps/optimization/dmrg.py:15-23).
#include <cmath>
#include <cstddef>
#include <vector>

#include "common.h"

namespace seemps {
namespace {

// Hm[(c,k,l), (ko,lo,d)] = sum_e O1[c, ko, k, e] O2[e, lo, l, d]
__global__ void two_site_op_kernel(int c_, int d1, int e_, int d2, int w2,
                                   const double2* __restrict__ O1, const double2* __restrict__ O2,
                                   double2* __restrict__ Hm) {
  const int rows = c_ * d1 * d2, cols = d1 * d2 * w2;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int row = idx / cols, col = idx % cols;
  const int c = row / (d1 * d2), k = (row / d2) % d1, l = row % d2;
  const int ko = col / (d2 * w2), lo = (col / w2) % d2, d = col % w2;
  double2 acc = make_double2(0.0, 0.0);
  for (int e = 0; e < e_; ++e) {
    const double2 a = O1[((c * d1 + ko) * d1 + k) * e_ + e];
    const double2 b = O2[((e * d2 + lo) * d2 + l) * w2 + d];
    acc = cfma(a, b, acc);
  }
  Hm[idx] = acc;
}

struct PermDesc {
  int ndim;
  int64_t out_dims[8];
  int64_t in_strides_perm[8];  // stride in the input of output axis k
};

__global__ void permute_kernel(int64_t total, PermDesc p, const double2* __restrict__ in,
                               double2* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = e, src = 0;
    for (int k = p.ndim - 1; k >= 0; --k) {
      const int64_t i = rem % p.out_dims[k];
      rem /= p.out_dims[k];
      src += i * p.in_strides_perm[k];
    }
    out[e] = in[src];
  }
}


// x[i] = sum_{j<k} V[j*ldv + i] * y[j]  (y real)
__global__ void combine_kernel(int64_t n, int k, const double2* __restrict__ V, int64_t ldv,
                               const double* __restrict__ y, double2* __restrict__ x) {
  __shared__ double ys[64];
  if (threadIdx.x < k) ys[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int j = 0; j < k; ++j) {
      const double2 v = V[j * ldv + i];
      acc.x += v.x * ys[j];
      acc.y += v.y * ys[j];
    }
    x[i] = acc;
  }
}

// Scalars of one Lanczos cycle kept on the device.
struct LzScalars {
  double alpha[64];
  double beta[64];
  double y[64];
  double2 h[64];
  double2 h2[64];
  double2 nrm;      // <w, w>
  double theta;
  double resid;
  double scale;     // running max |alpha| + |beta| (breakdown reference)
  int kd;           // effective Krylov dimension of the cycle
  int broken;       // breakdown seen in this cycle
  double scl[64];   // scale after step j
  int brk[64];      // breakdown seen up to step j
};

constexpr int LZ_NV = 64;  // stride of the per-block partial dot products

// Block-wide sum of one complex value per thread (256 threads), result valid
// in every thread.
__device__ double2 lz_block_sum(double2 v, double2* red) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t = cadd(t, red[q]);
  return t;
}

// Classical Gram-Schmidt of w against V[0..k) with the DGKS second pass
// (applied only when the first pass removed more than half of |w|^2), in three
// launches over lz_blocks(n) chunks.  Each block keeps the k entries V[j][i] of
// its elements in registers (k <= LZ_KREG), so every launch reads V once.
// Partial sums are reduced redundantly by every block in a fixed order, which
// keeps the result deterministic without an extra launch.
constexpr int LZ_KREG = 24;

// Number of chunks of the orthogonalisation launches.
inline int lz_blocks(int64_t n) {
  int64_t b = ceil_div(n, 1024);
  if (b > 256) b = 256;
  if (b < 1) b = 1;
  return (int)b;
}

// Sum over blocks of per-thread partials (k complex values and one real per
// thread) into part[blockIdx.x * LZ_NV + j] and pn[blockIdx.x].
__device__ void lz_store_partials(const double2* acc, int k, double nrm, double2* red,
                                  double* redn, double2* __restrict__ part,
                                  double* __restrict__ pn) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < LZ_KREG; ++j) {
    if (j < k) {
      double2 v = acc[j];
      for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
      }
      if (lane == 0) red[wave * LZ_KREG + j] = v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, 64);
  if (lane == 0) redn[wave] = nrm;
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6);
  if ((int)threadIdx.x < k) {
    double2 t = make_double2(0.0, 0.0);
    for (int q = 0; q < nw; ++q) t = cadd(t, red[q * LZ_KREG + threadIdx.x]);
    part[blockIdx.x * LZ_NV + threadIdx.x] = t;
  }
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < nw; ++q) t += redn[q];
    pn[blockIdx.x] = t;
  }
}

// h[j] = sum_b part[b * LZ_NV + j] into LDS, j < k (8 k <= blockDim.x): eight
// lanes per j each add every eighth partial, then a three-step shuffle tree.
// The order is fixed, so every block derives the same h.  Called by every
// thread of the block.
__device__ void lz_load_coeffs(int k, int nb, const double2* __restrict__ part, double2* hs) {
  const int j = (int)threadIdx.x >> 3, sub = (int)threadIdx.x & 7;
  double2 acc = make_double2(0.0, 0.0);
  if (j < k)
    for (int b = sub; b < nb; b += 8) acc = cadd(acc, part[b * LZ_NV + j]);
  for (int off = 1; off < 8; off <<= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
  }
  if (j < k && sub == 0) hs[j] = acc;
}

// sum_b pn[b], b < nb, by one whole wavefront: lane l adds pn[l], pn[l + 64],
// ..., then a butterfly, which leaves the same total (same order) in every lane.
__device__ double lz_wave_sum(int nb, const double* __restrict__ pn) {
  double t = 0.0;
  for (int b = (int)threadIdx.x & 63; b < nb; b += 64) t += pn[b];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  return t;
}

// Pass 1: part = chunk sums of conj(V_j) w; pn0 = chunk sums of |w|^2.
__global__ __launch_bounds__(256) void lz_orth1(int64_t n, int k, const double2* __restrict__ V,
                                                const double2* __restrict__ w,
                                                double2* __restrict__ part,
                                                double* __restrict__ pn0) {
  __shared__ double2 red[4 * LZ_KREG];
  __shared__ double redn[4];
  const int64_t chunk = ceil_div(n, (int64_t)gridDim.x);
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double2 acc[LZ_KREG];
#pragma unroll
  for (int j = 0; j < LZ_KREG; ++j) acc[j] = make_double2(0.0, 0.0);
  double nrm = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const double2 x = w[i];
    nrm += cabs2(x);
#pragma unroll
    for (int j = 0; j < LZ_KREG; ++j)
      if (j < k) acc[j] = cfmac(V[j * n + i], x, acc[j]);
  }
  lz_store_partials(acc, k, nrm, red, redn, part, pn0);
}

// Pass 2: h = sum of pass-1 partials (block 0 stores it); w -= V h; then the
// chunk sums of conj(V_j) w (part_out) and |w|^2 (pn1) of the updated w.
__global__ __launch_bounds__(256) void lz_orth2(int64_t n, int k, const double2* __restrict__ V,
                                                double2* __restrict__ w,
                                                const double2* __restrict__ part_in,
                                                double2* __restrict__ part_out,
                                                double* __restrict__ pn1,
                                                double2* __restrict__ h_out) {
  __shared__ double2 hs[LZ_KREG];
  __shared__ double2 red[4 * LZ_KREG];
  __shared__ double redn[4];
  lz_load_coeffs(k, gridDim.x, part_in, hs);
  __syncthreads();
  if (blockIdx.x == 0 && (int)threadIdx.x < k) h_out[threadIdx.x] = hs[threadIdx.x];
  const int64_t chunk = ceil_div(n, (int64_t)gridDim.x);
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double2 acc[LZ_KREG];
#pragma unroll
  for (int j = 0; j < LZ_KREG; ++j) acc[j] = make_double2(0.0, 0.0);
  double nrm = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    double2 vr[LZ_KREG];
    double2 x = w[i];
#pragma unroll
    for (int j = 0; j < LZ_KREG; ++j) {
      if (j < k) {
        vr[j] = V[j * n + i];
        const double2 t = cmul(vr[j], hs[j]);
        x.x -= t.x;
        x.y -= t.y;
      }
    }
    w[i] = x;
    nrm += cabs2(x);
#pragma unroll
    for (int j = 0; j < LZ_KREG; ++j)
      if (j < k) acc[j] = cfmac(vr[j], x, acc[j]);
  }
  lz_store_partials(acc, k, nrm, red, redn, part_out, pn1);
}

// Pass 3 (DGKS): if |w|^2 after pass 1 is at least half of |w|^2 before it,
// the first pass was accurate and w is kept (h2 = 0); otherwise w -= V h2 with
// h2 = the pass-2 coefficients.  pn2 receives the final chunk norms.
__global__ __launch_bounds__(256) void lz_orth3(int64_t n, int k, const double2* __restrict__ V,
                                                double2* __restrict__ w,
                                                const double2* __restrict__ part_in,
                                                const double* __restrict__ pn0,
                                                const double* __restrict__ pn1,
                                                double* __restrict__ pn2,
                                                double2* __restrict__ h2_out) {
  __shared__ double2 hs[LZ_KREG];
  __shared__ double s_nrm[2];
  __shared__ double2 red[4];
  const int nb = gridDim.x;
  if (threadIdx.x < 64) {
    const double t = lz_wave_sum(nb, pn0);
    if (threadIdx.x == 0) s_nrm[0] = t;
  } else if (threadIdx.x < 128) {
    const double t = lz_wave_sum(nb, pn1);
    if (threadIdx.x == 64) s_nrm[1] = t;
  }
  lz_load_coeffs(k, nb, part_in, hs);
  __syncthreads();
  const bool second = !(s_nrm[1] >= 0.5 * s_nrm[0]);
  if (blockIdx.x == 0 && (int)threadIdx.x < k)
    h2_out[threadIdx.x] = second ? hs[threadIdx.x] : make_double2(0.0, 0.0);
  if (!second) {
    if (threadIdx.x == 0) pn2[blockIdx.x] = pn1[blockIdx.x];
    return;
  }
  const int64_t chunk = ceil_div(n, (int64_t)gridDim.x);
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double nrm = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    double2 x = w[i];
#pragma unroll
    for (int j = 0; j < LZ_KREG; ++j) {
      if (j < k) {
        const double2 t = cmul(V[j * n + i], hs[j]);
        x.x -= t.x;
        x.y -= t.y;
      }
    }
    w[i] = x;
    nrm += cabs2(x);
  }
  const double2 t = lz_block_sum(make_double2(nrm, 0.0), red);
  if (threadIdx.x == 0) pn2[blockIdx.x] = t.x;
}

// Step j done: beta_j = ||w|| (from the partial norms, in every block),
// alpha_j = Re(h_j + h2_j), breakdown test, v_{j+1} = w / beta_j (0 after a
// breakdown).  Block 0 records the scalars.
__global__ __launch_bounds__(256) void lz_finish(int64_t n, int nparts, int j,
                                                 const double* __restrict__ pnorm,
                                                 LzScalars* __restrict__ s, double2* __restrict__ w) {
  __shared__ double s_b;
  __shared__ int s_brk;
  const double nrm2 = (threadIdx.x < 64) ? lz_wave_sum(nparts, pnorm) : 0.0;
  if (threadIdx.x == 0) {
    const double a = s->h[j].x + s->h2[j].x;
    const double b = sqrt(fmax(nrm2, 0.0));
    const double prev = (j > 0) ? s->beta[j - 1] : 0.0;
    const double scale_prev = (j > 0) ? s->scl[j - 1] : 0.0;
    const int brk_prev = (j > 0) ? s->brk[j - 1] : 0;
    const double ref = fmax(scale_prev, fabs(a) + prev);
    const int brk = brk_prev || (b <= 1e-12 * ref);
    s_b = b;
    s_brk = brk;
    if (blockIdx.x == 0) {
      s->alpha[j] = a;
      s->beta[j] = brk ? 0.0 : b;
      s->scl[j] = ref;
      s->brk[j] = brk;
      s->broken = brk;
      if (brk && !brk_prev) s->kd = j + 1;
    }
  }
  __syncthreads();
  const double inv = s_brk ? 0.0 : 1.0 / s_b;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = w[i];
    w[i] = make_double2(x.x * inv, x.y * inv);
  }
}

// Scale v in place by 1/sqrt(<v,v>) read from s->nrm.
__global__ void lz_unit(int64_t n, const LzScalars* __restrict__ s, double2* __restrict__ v) {
  const double nn = sqrt(fmax(s->nrm.x, 0.0));
  const double inv = (nn > 0.0) ? 1.0 / nn : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = v[i];
    v[i] = make_double2(x.x * inv, x.y * inv);
  }
}

__global__ void lz_reset(LzScalars* s) {
  if (threadIdx.x == 0) {
    s->broken = 0;
    s->kd = 0;
    s->scale = 0.0;
  }
}

// Sturm count: number of eigenvalues of the leading kd x kd block below x.
__device__ int sturm(const double* a, const double* b, int kd, double x) {
  int cnt = 0;
  double d = a[0] - x;
  if (d < 0.0) ++cnt;
  for (int i = 1; i < kd; ++i) {
    const double dd = (d == 0.0) ? 1e-300 : d;
    d = a[i] - x - b[i - 1] * b[i - 1] / dd;
    if (d < 0.0) ++cnt;
  }
  return cnt;
}

// Lowest eigenpair of the symmetric tridiagonal T = tridiag(beta, alpha, beta)
// of dimension kd (= m, or the breakdown size): 64-way multisection on Sturm
// counts for theta, then inverse iteration (tridiagonal LU with partial
// pivoting, LAPACK dgttrf/dgttrs steps) for y.  One wavefront; alpha and beta
// are staged in LDS first, so the serial Sturm recurrences and the inverse
// iteration wait on no global loads.  Ten rounds narrow the Gershgorin
// interval by 65^10 (about 1e-18 of its width), below the spacing of doubles.
__global__ __launch_bounds__(64) void lz_tridiag(LzScalars* s, int m) {
  __shared__ double lo_s, hi_s;
  __shared__ double a[64], b[64];
  const int lane = threadIdx.x;
  const int kd = s->broken ? s->kd : m;
  a[lane] = (lane < m) ? s->alpha[lane] : 0.0;
  b[lane] = (lane < m) ? s->beta[lane] : 0.0;
  __syncthreads();
  if (lane == 0) {
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < kd; ++i) {
      const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < kd - 1 ? fabs(b[i]) : 0.0);
      lo = fmin(lo, a[i] - r);
      hi = fmax(hi, a[i] + r);
    }
    const double pad = 1e-12 * fmax(fabs(lo), fabs(hi)) + 1e-300;
    lo_s = lo - pad;
    hi_s = hi + pad;
  }
  __syncthreads();
  for (int round = 0; round < 10; ++round) {
    const double lo = lo_s, hi = hi_s;
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const int c = sturm(a, b, kd, x);
    // smallest lane whose point has at least one eigenvalue below it
    const unsigned long long mask = __ballot(c >= 1);
    __syncthreads();
    if (lane == 0) {
      if (mask) {
        const int l = __ffsll((long long)mask) - 1;
        const double nhi = lo + (hi - lo) * (double)(l + 1) / 65.0;
        const double nlo = (l == 0) ? lo : lo + (hi - lo) * (double)l / 65.0;
        lo_s = nlo;
        hi_s = nhi;
      } else {
        lo_s = lo + (hi - lo) * 64.0 / 65.0;
      }
    }
    __syncthreads();
  }
  if (lane != 0) return;
  const double theta = 0.5 * (lo_s + hi_s);
  // Inverse iteration on (T - theta I), one lane; the factors live in LDS
  // (dynamically indexed private arrays would spill to scratch memory).
  __shared__ double d[64], dl[64], du[64], du2[64], y[64];
  __shared__ int piv[64];
  double tnorm = 0.0;
  for (int i = 0; i < kd; ++i) tnorm = fmax(tnorm, fabs(a[i]) + (i < kd - 1 ? fabs(b[i]) : 0.0));
  const double tiny = 2.220446049250313e-16 * fmax(tnorm, 1e-300);
  for (int i = 0; i < kd; ++i) {
    d[i] = a[i] - theta;
    du[i] = (i < kd - 1) ? b[i] : 0.0;
    dl[i] = (i < kd - 1) ? b[i] : 0.0;
    du2[i] = 0.0;
    piv[i] = i;
  }
  for (int i = 0; i < kd - 1; ++i) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] == 0.0) d[i] = tiny;
      const double fact = dl[i] / d[i];
      dl[i] = fact;
      d[i + 1] -= fact * du[i];
      piv[i] = i;
    } else {
      const double fact = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = fact;
      const double temp = du[i];
      du[i] = d[i + 1];
      d[i + 1] = temp - fact * d[i + 1];
      if (i < kd - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -fact * du[i + 1];
      }
      piv[i] = i + 1;
    }
  }
  for (int i = 0; i < kd; ++i)
    if (d[i] == 0.0) d[i] = tiny;
  for (int i = 0; i < kd; ++i) y[i] = 1.0;
  for (int it = 0; it < 4; ++it) {
    for (int i = 0; i < kd - 1; ++i) {
      if (piv[i] == i) {
        y[i + 1] -= dl[i] * y[i];
      } else {
        const double temp = y[i];
        y[i] = y[i + 1];
        y[i + 1] = temp - dl[i] * y[i];
      }
    }
    y[kd - 1] /= d[kd - 1];
    if (kd > 1) y[kd - 2] = (y[kd - 2] - du[kd - 2] * y[kd - 1]) / d[kd - 2];
    for (int i = kd - 3; i >= 0; --i) y[i] = (y[i] - du[i] * y[i + 1] - du2[i] * y[i + 2]) / d[i];
    double nn = 0.0;
    for (int i = 0; i < kd; ++i) nn += y[i] * y[i];
    nn = sqrt(nn);
    for (int i = 0; i < kd; ++i) y[i] /= nn;
  }
  // Fix the sign: largest component positive.
  int imax = 0;
  for (int i = 1; i < kd; ++i)
    if (fabs(y[i]) > fabs(y[imax])) imax = i;
  const double sg = (y[imax] < 0.0) ? -1.0 : 1.0;
  for (int i = 0; i < 64; ++i) s->y[i] = (i < kd) ? sg * y[i] : 0.0;
  s->theta = theta;
  s->resid = (kd == m && !s->broken) ? fabs(b[m - 1] * y[kd - 1]) : 0.0;
}

struct EigWork {
  double2* V;      // (m + 1) x n
  double2* x1;     // a*c*d1*d2*f
  double2* x2;     // a*d1*d2*w2*f
  double2* part;   // reduction partials (64 x nblk)
  double2* part2;  // second-pass partials (64 x nblk)
  double* pnorm;   // partial squared norms (3 x nblk: before pass 1, after it, final)
  double2* ks;     // split-K partials of the matvec GEMMs
  LzScalars* sc;
};

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

// Partial-sum slots needed by both the Lanczos passes and zmdotc_impl.
inline int64_t part_blocks(int64_t n) {
  const int64_t a = lz_blocks(n), b = reduce_blocks(n);
  return a > b ? a : b;
}

// Split-K scratch (elements) of the two matvec GEMMs.
int64_t matvec_ks_elems(int64_t a, int64_t c, int64_t b, int64_t d1, int64_t d2, int64_t w2,
                        int64_t f, int64_t g) {
  const int64_t k1 = splitk_workspace(a * c, d1 * d2 * f, b);
  const int64_t k2 = splitk_workspace(a * d1 * d2, g, w2 * f);
  return k1 > k2 ? k1 : k2;
}

int64_t eig_bytes_full(int64_t chiL, int64_t c, int64_t d1, int64_t d2, int64_t w2, int64_t chiR,
                       int64_t m) {
  const int64_t n = chiL * d1 * d2 * chiR;
  int64_t b = 0;
  b += al256((m + 1) * n * 16);
  b += al256(chiL * c * d1 * d2 * chiR * 16);
  b += al256(chiL * d1 * d2 * w2 * chiR * 16);
  b += al256(64 * part_blocks(n) * 16);
  b += al256(64 * part_blocks(n) * 16);
  b += al256(3 * part_blocks(n) * 8);
  b += al256(matvec_ks_elems(chiL, c, chiL, d1, d2, w2, chiR, chiR) * 16);
  b += al256(sizeof(LzScalars));
  return b;
}

EigWork eig_carve(void* work, int64_t chiL, int64_t c, int64_t d1, int64_t d2, int64_t w2,
                  int64_t chiR, int64_t m) {
  const int64_t n = chiL * d1 * d2 * chiR;
  char* p = (char*)work;
  EigWork w;
  w.V = (double2*)p;
  p += al256((m + 1) * n * 16);
  w.x1 = (double2*)p;
  p += al256(chiL * c * d1 * d2 * chiR * 16);
  w.x2 = (double2*)p;
  p += al256(chiL * d1 * d2 * w2 * chiR * 16);
  w.part = (double2*)p;
  p += al256(64 * part_blocks(n) * 16);
  w.part2 = (double2*)p;
  p += al256(64 * part_blocks(n) * 16);
  w.pnorm = (double*)p;
  p += al256(3 * part_blocks(n) * 8);
  w.ks = (double2*)p;
  p += al256(matvec_ks_elems(chiL, c, chiL, d1, d2, w2, chiR, chiR) * 16);
  w.sc = (LzScalars*)p;
  return w;
}

unsigned grid_for(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// out[a, ko, lo, g] = sum L[a,c,b] Hm[(c,k,l),(ko,lo,d)] v[b,k,l,f] R[g,d,f]
int two_site_matvec_impl(int64_t a, int64_t c, int64_t b, int64_t d1, int64_t d2, int64_t w2,
                         int64_t f, int64_t g, const double2* L, const double2* Hm,
                         const double2* R, const double2* v, double2* out, double2* x1,
                         double2* x2, double2* ks, int64_t ks_elems, hipStream_t stream,
                         int prec) {
  int st;
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  // x1[(a,c), (k,l,f)] = L[(a,c), b] v[b, (k,l,f)]
  if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_N, a * c, d1 * d2 * f, b, one, L, b, v, d1 * d2 * f,
                     zero, x1, d1 * d2 * f, ks, ks_elems, stream, prec)))
    return st;
  // x2[a, (ko,lo,d), f] = sum_{(c,k,l)} x1[a, (c,k,l), f] Hm[(c,k,l), (ko,lo,d)]
  if ((st = zlocal_impl(a, c * d1 * d2, f, d1 * d2 * w2, 1, x1, Hm, x2, 0, stream))) return st;
  // out[(a,ko,lo), g] = x2[(a,ko,lo), (d,f)] R[g, (d,f)]^T
  return zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, a * d1 * d2, g, w2 * f, one, x2, w2 * f, R, w2 * f,
                  zero, out, g, ks, ks_elems, stream, prec);
}

}  // namespace seemps

using namespace seemps;

extern "C" int seemps_two_site_operator(int64_t c, int64_t d1, int64_t e, int64_t d2, int64_t w2,
                                        const void* O1, const void* O2, void* Hm, void* stream) {
  SEEMPS_CHECK_ARG(c >= 1 && d1 >= 1 && e >= 1 && d2 >= 1 && w2 >= 1,
                   "two_site_operator: bad dimensions");
  const int64_t total = c * d1 * d2 * d1 * d2 * w2;
  hipLaunchKernelGGL(two_site_op_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, (int)c, (int)d1, (int)e, (int)d2, (int)w2,
                     (const double2*)O1, (const double2*)O2, (double2*)Hm);
  return check_launch("two_site_operator");
}

extern "C" int seemps_zpermute(int ndim, const int64_t* dims, const int* perm, const void* in,
                               void* out, void* stream) {
  SEEMPS_CHECK_ARG(ndim >= 1 && ndim <= 8, "permute: 1 <= ndim <= 8");
  PermDesc p;
  p.ndim = ndim;
  int64_t strides[8];
  int64_t acc = 1;
  for (int k = ndim - 1; k >= 0; --k) {
    strides[k] = acc;
    acc *= dims[k];
  }
  int seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < ndim; ++k) {
    SEEMPS_CHECK_ARG(perm[k] >= 0 && perm[k] < ndim && !seen[perm[k]], "permute: invalid perm");
    seen[perm[k]] = 1;
    p.out_dims[k] = dims[perm[k]];
    p.in_strides_perm[k] = strides[perm[k]];
  }
  const int64_t total = acc;
  if (total == 0) return SEEMPS_OK;
  hipLaunchKernelGGL(permute_kernel, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream,
                     total, p, (const double2*)in, (double2*)out);
  return check_launch("permute");
}

extern "C" int64_t seemps_two_site_matvec_workspace(int64_t a, int64_t c, int64_t b, int64_t d1,
                                                    int64_t d2, int64_t w2, int64_t f) {
  return al256(a * c * d1 * d2 * f * 16) + al256(a * d1 * d2 * w2 * f * 16) +
         al256(matvec_ks_elems(a, c, b, d1, d2, w2, f, a) * 16);
}

extern "C" int seemps_two_site_matvec(int64_t a, int64_t c, int64_t b, int64_t d1, int64_t d2,
                                      int64_t w2, int64_t f, int64_t g, const void* L,
                                      const void* Hm, const void* R, const void* v, void* out,
                                      void* work, int64_t work_bytes, void* stream) {
  SEEMPS_CHECK_ARG(work_bytes >= seemps_two_site_matvec_workspace(a, c, b, d1, d2, w2, f),
                   "two_site_matvec: workspace too small");
  double2* x1 = (double2*)work;
  double2* x2 = (double2*)((char*)work + al256(a * c * d1 * d2 * f * 16));
  double2* ks = (double2*)((char*)x2 + al256(a * d1 * d2 * w2 * f * 16));
  // the split-K size depends on g too; the workspace query assumed g = a
  const int64_t ks_avail = (work_bytes - ((char*)ks - (char*)work)) / 16;
  return two_site_matvec_impl(a, c, b, d1, d2, w2, f, g, (const double2*)L, (const double2*)Hm,
                              (const double2*)R, (const double2*)v, (double2*)out, x1, x2, ks,
                              ks_avail, (hipStream_t)stream, SEEMPS_PREC_FP64);
}

extern "C" int64_t seemps_eigensolve_workspace(int64_t chiL, int64_t c, int64_t d1, int64_t d2,
                                               int64_t w2, int64_t chiR, int64_t ncv) {
  const int64_t n = chiL * d1 * d2 * chiR;
  const int64_t m = ncv < n ? ncv : n;
  return eig_bytes_full(chiL, c, d1, d2, w2, chiR, m);
}

namespace {

int eigensolve_impl(int prec, int64_t chiL, int64_t c, int64_t d1, int64_t d2, int64_t w2,
                    int64_t chiR, const void* L, const void* Hm, const void* R, const void* v0,
                    void* evec, int64_t ncv, double tol, int64_t maxiter, double* stats,
                    int nstats, void* work, int64_t work_bytes, hipStream_t st_) {
  SEEMPS_CHECK_ARG(prec == SEEMPS_PREC_FP64 || prec == SEEMPS_PREC_FP32 ||
                       prec == SEEMPS_PREC_MIXED,
                   "eigensolve: invalid precision");
  SEEMPS_CHECK_ARG(chiL >= 1 && chiR >= 1 && c >= 1 && w2 >= 1 && d1 >= 1 && d2 >= 1,
                   "eigensolve: bad dimensions");
  SEEMPS_CHECK_ARG(ncv >= 2 && ncv < LZ_KREG, "eigensolve: ncv must be in [2, %d]",
                   LZ_KREG - 1);
  SEEMPS_CHECK_ARG(maxiter >= 1, "eigensolve: maxiter must be positive");
  const int64_t n = chiL * d1 * d2 * chiR;
  const int64_t m = ncv < n ? ncv : n;
  SEEMPS_CHECK_ARG(work_bytes >= eig_bytes_full(chiL, c, d1, d2, w2, chiR, m),
                   "eigensolve: workspace too small");
  if (tol <= 0.0) tol = 2.220446049250313e-16;
  const double eps23 = pow(2.220446049250313e-16, 2.0 / 3.0);
  EigWork w = eig_carve(work, chiL, c, d1, d2, w2, chiR, m);
  const double2* L_ = (const double2*)L;
  const double2* H_ = (const double2*)Hm;
  const double2* R_ = (const double2*)R;
  const unsigned gb = grid_for(n);
  const int nbk = lz_blocks(n);
  int st;
  // V[0] = v0 / ||v0||
  SEEMPS_HIP(hipMemcpyAsync(w.V, v0, n * 16, hipMemcpyDeviceToDevice, st_));
  if ((st = zmdotc_impl(n, 1, w.V, n, w.V, &w.sc->nrm, w.part, st_))) return st;
  hipLaunchKernelGGL(lz_unit, dim3(gb), dim3(256), 0, st_, n, w.sc, w.V);
  if ((st = check_launch("lz_unit"))) return st;
  double theta = 0.0, resid = 0.0;
  double* hp = (double*)pinned_stage(PINNED_BYTES);  // host reads of the scalars
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  static_assert(offsetof(LzScalars, resid) == offsetof(LzScalars, theta) + 8,
                "theta and resid are read together");
  int64_t matvecs = 0, matvecs32 = 0, cycle = 0;
  bool converged = false;
  // FP32 phase: the operator is applied with FP32 products, whose rounding
  // (~1e-7 relative) floors the residual near 1e-7 |theta|.  MIXED enters it
  // only when the start vector's residual (the first step's beta, from an FP64
  // product) exceeds 1e-4 |alpha|, and leaves it at 1e-5 |theta| or when a
  // restart gains less than a factor 4.  A good start vector (a converged
  // DMRG sweep) thus stays in FP64, where one short cycle suffices.
  int mv_prec = prec == SEEMPS_PREC_FP32 ? SEEMPS_PREC_FP32 : SEEMPS_PREC_FP64;
  const double fp32_floor = 1e-5, fp32_start = 1e-4;
  double last_resid = HUGE_VAL;
  // In FP64 cycles the Ritz residual is also tested after steps 4, 6, 8, ...
  // (one single-block tridiagonal solve and a synchronisation each), so that a
  // warm start stops well before ncv products; only for operators large enough
  // that a product outweighs the check.
  const bool early = n >= 8192;
  const int64_t first_check = 4, check_every = 2;
  for (cycle = 0; cycle < maxiter; ++cycle) {
    hipLaunchKernelGGL(lz_reset, dim3(1), dim3(64), 0, st_, w.sc);
    int64_t kd_eff = m;
    for (int64_t j = 0; j < m; ++j) {
      double2* vj = w.V + j * n;
      double2* wv = w.V + (j + 1) * n;
      if ((st = two_site_matvec_impl(chiL, c, chiL, d1, d2, w2, chiR, chiR, L_, H_, R_, vj, wv,
                                     w.x1, w.x2, w.ks,
                                     matvec_ks_elems(chiL, c, chiL, d1, d2, w2, chiR, chiR),
                                     st_, mv_prec)))
        return st;
      ++matvecs;
      if (mv_prec == SEEMPS_PREC_FP32) ++matvecs32;
      // Gram-Schmidt against V[0..j] with the DGKS correction, then the norm:
      // four launches (dots | subtract + dots | conditional subtract | normalise).
      const int kk = (int)(j + 1);
      const double orth = 8.0 * n * kk;
      double* pn0 = w.pnorm;
      double* pn1 = w.pnorm + nbk;
      double* pn2 = w.pnorm + 2 * nbk;
      {
        TimedLaunch timed("lz_orth1", st_, orth);
        hipLaunchKernelGGL(lz_orth1, dim3(nbk), dim3(256), 0, st_, n, kk, w.V, wv, w.part, pn0);
      }
      {
        TimedLaunch timed("lz_orth2", st_, 2.0 * orth);
        hipLaunchKernelGGL(lz_orth2, dim3(nbk), dim3(256), 0, st_, n, kk, w.V, wv, w.part,
                           w.part2, pn1, w.sc->h);
      }
      {
        TimedLaunch timed("lz_orth3", st_, orth);
        hipLaunchKernelGGL(lz_orth3, dim3(nbk), dim3(256), 0, st_, n, kk, w.V, wv, w.part2, pn0,
                           pn1, pn2, w.sc->h2);
      }
      {
        TimedLaunch timed("lz_finish", st_, 0.0);
        hipLaunchKernelGGL(lz_finish, dim3(gb), dim3(256), 0, st_, n, nbk, (int)j, pn2, w.sc,
                           wv);
      }
      if ((st = check_launch("lanczos step"))) return st;
      if (prec == SEEMPS_PREC_MIXED && cycle == 0 && j == 0) {
        double* ab = hp;
        SEEMPS_HIP(hipMemcpyAsync(&ab[0], &w.sc->alpha[0], 8, hipMemcpyDeviceToHost, st_));
        SEEMPS_HIP(hipMemcpyAsync(&ab[1], &w.sc->beta[0], 8, hipMemcpyDeviceToHost, st_));
        SEEMPS_HIP(hipStreamSynchronize(st_));
        if (ab[1] > fp32_start * fmax(eps23, fabs(ab[0]))) mv_prec = SEEMPS_PREC_FP32;
      }
      const int64_t kd = j + 1;
      if (early && mv_prec == SEEMPS_PREC_FP64 && kd < m && kd >= first_check &&
          (kd - first_check) % check_every == 0) {
        {
          TimedLaunch timed("lz_tridiag", st_, 0.0);
          hipLaunchKernelGGL(lz_tridiag, dim3(1), dim3(64), 0, st_, w.sc, (int)kd);
        }
        if ((st = check_launch("lz_tridiag"))) return st;
        double* h = hp;  // theta and resid are adjacent in LzScalars
        SEEMPS_HIP(hipMemcpyAsync(h, &w.sc->theta, 16, hipMemcpyDeviceToHost, st_));
        SEEMPS_HIP(hipStreamSynchronize(st_));
        if (h[1] <= tol * fmax(eps23, fabs(h[0]))) {
          kd_eff = kd;
          break;
        }
      }
    }
    if (kd_eff == m) {
      {
        TimedLaunch timed("lz_tridiag", st_, 0.0);
        hipLaunchKernelGGL(lz_tridiag, dim3(1), dim3(64), 0, st_, w.sc, (int)m);
      }
      if ((st = check_launch("lz_tridiag"))) return st;
    }
    // Ritz vector into V[m] (free: at most V[kd_eff] holds a Lanczos vector),
    // normalise, restart from it.
    double2* x = w.V + m * n;
    hipLaunchKernelGGL(combine_kernel, dim3(gb), dim3(256), 0, st_, n, (int)kd_eff, w.V, n,
                       w.sc->y, x);
    if ((st = check_launch("combine"))) return st;
    if ((st = zmdotc_impl(n, 1, x, n, x, &w.sc->nrm, w.part, st_))) return st;
    hipLaunchKernelGGL(lz_unit, dim3(gb), dim3(256), 0, st_, n, w.sc, x);
    SEEMPS_HIP(hipMemcpyAsync(w.V, x, n * 16, hipMemcpyDeviceToDevice, st_));
    double* h = hp;
    SEEMPS_HIP(hipMemcpyAsync(h, &w.sc->theta, 16, hipMemcpyDeviceToHost, st_));
    SEEMPS_HIP(hipStreamSynchronize(st_));
    theta = h[0];
    resid = h[1];
    const double scale = fmax(eps23, fabs(theta));
    if (mv_prec == SEEMPS_PREC_FP32 && prec == SEEMPS_PREC_MIXED) {
      // An FP32 residual meets the test only for tolerances FP32 can resolve.
      if (tol >= fp32_floor && resid <= tol * scale) {
        converged = true;
        ++cycle;
        break;
      }
      if (resid <= fp32_floor * scale || resid > 0.25 * last_resid) mv_prec = SEEMPS_PREC_FP64;
      last_resid = resid;
      continue;
    }
    if (resid <= tol * scale) {
      converged = true;
      ++cycle;
      break;
    }
  }
  SEEMPS_HIP(hipMemcpyAsync(evec, w.V, n * 16, hipMemcpyDeviceToDevice, st_));
  if (stats) {
    stats[0] = theta;
    stats[1] = resid;
    stats[2] = (double)matvecs;
    stats[3] = (double)cycle;
    if (nstats > 4) stats[4] = (double)matvecs32;
  }
  if (!converged) {
    set_error("eigensolve: no convergence after %lld restarts (residual %g)", (long long)cycle,
              resid);
    return SEEMPS_ERR_NOCONVERGE;
  }
  return SEEMPS_OK;
}

}  // namespace

extern "C" int seemps_two_site_eigensolve(int64_t chiL, int64_t c, int64_t d1, int64_t d2,
                                          int64_t w2, int64_t chiR, const void* L, const void* Hm,
                                          const void* R, const void* v0, void* evec, int64_t ncv,
                                          double tol, int64_t maxiter, double* stats, void* work,
                                          int64_t work_bytes, void* stream) {
  return eigensolve_impl(SEEMPS_PREC_FP64, chiL, c, d1, d2, w2, chiR, L, Hm, R, v0, evec, ncv,
                         tol, maxiter, stats, 4, work, work_bytes, (hipStream_t)stream);
}

extern "C" int seemps_two_site_eigensolve_prec(int precision, int64_t chiL, int64_t c,
                                               int64_t d1, int64_t d2, int64_t w2, int64_t chiR,
                                               const void* L, const void* Hm, const void* R,
                                               const void* v0, void* evec, int64_t ncv,
                                               double tol, int64_t maxiter, double* stats5,
                                               void* work, int64_t work_bytes, void* stream) {
  return eigensolve_impl(precision, chiL, c, d1, d2, w2, chiR, L, Hm, R, v0, evec, ncv, tol,
                         maxiter, stats5, 5, work, work_bytes, (hipStream_t)stream);
}
