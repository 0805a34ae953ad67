// SVD of a matrix whose columns are already nearly orthogonal, by
// simultaneous Jacobi rotations: every iteration forms the Gram matrix
// G = W^H W with one MFMA GEMM, derives for every column pair (i, j) the
// plane rotation that would annihilate G_ij (the one-sided Jacobi rotation
// of that pair), assembles all of them into one matrix Z = I + X (X
// anti-Hermitian, X_ij = t_ij ph_ij from the rotation's tangent and phase),
// replaces Z by its unitary polar factor (Newton-Schulz iterations, GEMMs),
// and rotates [W; V] <- [W; V] Z with one GEMM.  Near convergence the step
// is Jacobi's: the normalised off-diagonal of G falls quadratically from one
// iteration to the next, so a matrix that starts at 1e-3 needs four or five
// iterations, each a handful of GEMMs on the matrix cores instead of a chain
// of small dependent block steps.
//
// Accuracy: each G_ij is a dot product of two columns, with rounding
// relative to the two column norms, so the rotations are as accurate as the
// one-sided Jacobi's; the polar factor is unitary to a few ulp (the
// Newton-Schulz count is chosen on the host from a bound on ||X||_2).
//
// Used by the basis-sketched split (sketch.hip), whose W = op(A) Omega has
// nearly orthogonal columns; a matrix that does not converge within
// POLAR_MAX_ITERS iterations is reported as not converged and the caller
// falls back to the full SVD.
#include <cmath>

#include "common.h"
#include "jacobi_inner.h"

namespace seemps {

namespace {

constexpr int POLAR_MAX_ITERS = 12;
// An iteration that starts below this normalised off-diagonal leaves ~1e-18
// (quadratic convergence), so it is the last one.
constexpr double POLAR_QUAD_STOP = 1e-9;

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct PolarWork {
  double2* XA;     // (Lg + l) x l: rows < Lg the working matrix, then V
  double2* XB;     // ping-pong buffer of XA
  double2* G;      // l x l
  double2* Z;      // l x l
  double2* Z2;     // l x l
  double2* T;      // l x l
  double* stats;   // [0] max normalised off-diagonal, [1] bound on ||X||_2 (bit patterns),
                   // [2] ||W||_F^2, [3] negligible squared column norm
  double* nrm2;    // l squared column norms
  int* perm;       // l: columns by decreasing norm
  int64_t* tiny;   // first output slot with a negligible column (l if none)
  double2* sk;     // split-K scratch
  int64_t sk_elems;
  int64_t bytes;
};

PolarWork polar_carve(int64_t Lg, int64_t l, void* base) {
  PolarWork w;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    void* q = base ? (void*)((char*)base + off) : nullptr;
    off += al256(bytes);
    return q;
  };
  const int64_t H = Lg + l;
  w.XA = (double2*)take(H * l * 16);
  w.XB = (double2*)take(H * l * 16);
  w.G = (double2*)take(l * l * 16);
  w.Z = (double2*)take(l * l * 16);
  w.Z2 = (double2*)take(l * l * 16);
  w.T = (double2*)take(l * l * 16);
  w.stats = (double*)take(8 * 8);
  w.nrm2 = (double*)take(l * 8);
  w.perm = (int*)take(l * 4);
  w.tiny = (int64_t*)take(8);
  int64_t sk = splitk_workspace(l, l, Lg);
  const int64_t s2 = splitk_workspace(l, l, l), s3 = splitk_workspace(H, l, l);
  if (s2 > sk) sk = s2;
  if (s3 > sk) sk = s3;
  w.sk_elems = sk;
  w.sk = (double2*)take(sk * 16);
  w.bytes = off;
  return w;
}

// XA = [W; I]: W (Lg x l, row stride ldw) on top, the l x l identity below.
__global__ void polar_init(int64_t Lg, int64_t l, const double2* __restrict__ W, int64_t ldw,
                           double2* __restrict__ XA) {
  const int64_t total = (Lg + l) * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / l, c = e % l;
    XA[e] = (r < Lg) ? W[r * ldw + c] : make_double2((r - Lg == c) ? 1.0 : 0.0, 0.0);
  }
}

// stats[2] = ||W||_F^2 (fixed summation order), stats[3] = (Lg eps)^2 stats[2];
// stats[0] = stats[1] = 0.  One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void polar_frob(int64_t Lg, int64_t l,
                                                   const double2* __restrict__ XA,
                                                   double* __restrict__ stats) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < Lg * l; e += blockDim.x) acc += cabs2(XA[e]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
    const double le = (double)Lg * 2.220446049250313e-16;
    stats[0] = 0.0;
    stats[1] = 0.0;
    stats[2] = t;
    stats[3] = le * le * t;
  }
}

__global__ void polar_reset(double* __restrict__ stats) {
  if (threadIdx.x == 0) {
    stats[0] = 0.0;
    stats[1] = 0.0;
  }
}

// Row i of Z = I + X from the Hermitian Gram matrix G (upper triangle used;
// the lower one is its mirror): for i < j the pair (i, j) rotation of
// one-sided Jacobi, J = [[cs, sn ph], [-sn conj(ph), cs]] on columns (i, j),
// enters as X_ij = t ph and X_ji = -t conj(ph), t = sn / cs.  Pairs already
// orthogonal to `tol` (relative to the two column norms), or with a
// negligible column, get no rotation.  Also accumulates, by atomic maximum on
// the bit patterns of non-negative doubles, the largest normalised
// off-diagonal and the largest row sum of |X| (= ||X||_1 = ||X||_inf for an
// anti-Hermitian X, a bound on ||X||_2).  One workgroup per row.
__global__ __launch_bounds__(256) void polar_generator(int64_t l, const double2* __restrict__ G,
                                                       double tol, double* __restrict__ stats,
                                                       double2* __restrict__ Z) {
  __shared__ double red_off[4], red_sum[4];
  const int64_t i = blockIdx.x;
  const double tiny2 = stats[3];
  double mx = 0.0, rowsum = 0.0;
  for (int64_t j = threadIdx.x; j < l; j += blockDim.x) {
    double2 z = make_double2(i == j ? 1.0 : 0.0, 0.0);
    if (j != i) {
      const int64_t a = i < j ? i : j, b = i < j ? j : i;
      const double gaa = G[a * l + a].x, gbb = G[b * l + b].x;
      const double2 c = G[a * l + b];
      const double s2 = cabs2(c);
      if (gaa > tiny2 && gbb > tiny2) mx = fmax(mx, sqrt(s2 / gaa / gbb));
      if (s2 > 0.0 && gaa > tiny2 && gbb > tiny2 && s2 > (tol * tol * gaa) * gbb) {
        const double r = jac::rsqrt_nr(s2);  // 1 / |G_ab|
        const double2 ph = make_double2(c.x * r, c.y * r);
        const double tau = (gbb - gaa) * (0.5 * r);
        double t;
        if (fabs(tau) > 1e150) {
          t = 0.5 * jac::rcp_nr(tau);
        } else {
          const double q = fma(tau, tau, 1.0);
          t = (tau >= 0.0 ? 1.0 : -1.0) * jac::rcp_nr(fabs(tau) + q * jac::rsqrt_nr(q));
        }
        // row i: X_ab = t ph (i = a), X_ba = -t conj(ph) (i = b)
        z = (i == a) ? make_double2(t * ph.x, t * ph.y) : make_double2(-t * ph.x, t * ph.y);
        rowsum += fabs(t);
      }
    }
    Z[i * l + j] = z;
  }
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmax(mx, __shfl_down(mx, off, 64));
    rowsum += __shfl_down(rowsum, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red_off[wave] = mx;
    red_sum[wave] = rowsum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0, s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      m = fmax(m, red_off[q]);
      s += red_sum[q];
    }
    // non-negative doubles order like their bit patterns
    atomicMax((unsigned long long*)&stats[0], (unsigned long long)__double_as_longlong(m));
    atomicMax((unsigned long long*)&stats[1], (unsigned long long)__double_as_longlong(s));
  }
}

// nrm2[c] = sum over the first Lg rows of |XA[r][c]|^2.  One workgroup per column.
__global__ __launch_bounds__(256) void polar_norms(int64_t Lg, int64_t l,
                                                   const double2* __restrict__ XA,
                                                   double* __restrict__ nrm2) {
  __shared__ double red[4];
  const int64_t c = blockIdx.x;
  double acc = 0.0;
  for (int64_t r = threadIdx.x; r < Lg; r += blockDim.x) acc += cabs2(XA[r * l + c]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) nrm2[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

// perm[rank] = c by decreasing nrm2, ties by index, NaN last; *tiny = first
// slot whose column is negligible (<= stats[3]), l if none.
__global__ void polar_rank(int64_t l, const double* __restrict__ nrm2,
                           const double* __restrict__ stats, int* __restrict__ perm,
                           int64_t* __restrict__ tiny) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < l;
       c += (int64_t)gridDim.x * blockDim.x) {
    const double nc = isnan(nrm2[c]) ? -1.0 : nrm2[c];
    int rank = 0;
    for (int64_t d = 0; d < l; ++d) {
      const double nd = isnan(nrm2[d]) ? -1.0 : nrm2[d];
      rank += (nd > nc || (nd == nc && d < c)) ? 1 : 0;
    }
    perm[rank] = (int)c;
    if (!(nc > stats[3])) atomicMin((unsigned long long*)tiny, (unsigned long long)rank);
  }
}

__global__ void polar_tiny_reset(int64_t l, int64_t* tiny) {
  if (threadIdx.x == 0) *tiny = l;
}

// U[r][k] = XA[r][perm[k]] / S[k] (row stride l), S[k] = ||column perm[k]||,
// VH[k][q] = conj(XA[Lg + q][perm[k]]).
__global__ void polar_output(int64_t Lg, int64_t l, const double2* __restrict__ XA,
                             const double* __restrict__ nrm2, const int* __restrict__ perm,
                             double2* __restrict__ U, double* __restrict__ S,
                             double2* __restrict__ VH) {
  const int64_t total = (Lg + l) * l;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / l, k = e % l;
    const int c = perm[k];
    const double2 v = XA[r * l + c];
    if (r < Lg) {
      const double s = sqrt(nrm2[c]);
      const double f = s > 0.0 ? 1.0 / s : 0.0;
      U[r * l + k] = make_double2(v.x * f, v.y * f);
      if (r == 0) S[k] = s;
    } else {
      VH[k * l + (r - Lg)] = make_double2(v.x, -v.y);
    }
  }
}

unsigned polar_grid(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Newton-Schulz steps that take the smallest singular value s of the scaled
// Z (all singular values lie in [s, 1]) to 1 within rounding, plus one.
int newton_schulz_steps(double s) {
  int k = 0;
  while (1.0 - s > 1e-16 && k < 60) {
    s = s * (3.0 - s * s) * 0.5;
    ++k;
  }
  return k + 1;
}

}  // namespace

int64_t polar_svd_bytes(int64_t Lg, int64_t l) { return polar_carve(Lg, l, nullptr).bytes; }

// SVD W = U diag(S) VH of the Lg x l matrix W (row stride ldw, Lg >= l, columns
// nearly orthogonal): U Lg x l (row stride l), S descending, VH l x l.  The
// stream is synchronised once per iteration (the host reads the convergence
// measure).  *tiny_dev (device, may be read after the caller's next
// synchronisation) receives the first slot whose singular value is
// negligible (<= Lg eps ||W||_F), l if none: such slots have U columns of
// zero, which the caller must not keep.  Returns SEEMPS_ERR_NOCONVERGE when
// POLAR_MAX_ITERS iterations do not reach the tolerance.  info = {iterations,
// converged}.
int polar_svd(int64_t Lg, int64_t l, const double2* W, int64_t ldw, double2* U, double* S,
              double2* VH, const int64_t** tiny_dev, void* work, int64_t work_bytes,
              int32_t* info, hipStream_t st_) {
  SEEMPS_CHECK_ARG(Lg >= l && l >= 1, "polar_svd: needs Lg >= l >= 1");
  PolarWork w = polar_carve(Lg, l, work);
  SEEMPS_CHECK_ARG(work_bytes >= w.bytes, "polar_svd: workspace too small");
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  const int64_t H = Lg + l;
  const double tol = 32.0 * sqrt((double)Lg) * 2.220446049250313e-16;
  double* hp = (double*)pinned_stage(PINNED_BYTES);
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  int st;
  hipLaunchKernelGGL(polar_init, dim3(polar_grid(H * l)), dim3(256), 0, st_, Lg, l, W, ldw,
                     w.XA);
  if ((st = check_launch("polar_init"))) return st;
  hipLaunchKernelGGL(polar_frob, dim3(1), dim3(1024), 0, st_, Lg, l, w.XA, w.stats);
  if ((st = check_launch("polar_frob"))) return st;
  int it = 0;
  bool converged = false;
  for (it = 0; it < POLAR_MAX_ITERS; ++it) {
    if (it > 0) {
      hipLaunchKernelGGL(polar_reset, dim3(1), dim3(64), 0, st_, w.stats);
      if ((st = check_launch("polar_reset"))) return st;
    }
    // G = W^H W
    if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_N, l, l, Lg, one, w.XA, l, w.XA, l, zero, w.G, l,
                       w.sk, w.sk_elems, st_)))
      return st;
    {
      TimedLaunch timed("polar_generator", st_, 0.0);
      hipLaunchKernelGGL(polar_generator, dim3((unsigned)l), dim3(256), 0, st_, l, w.G, tol,
                         w.stats, w.Z);
      if ((st = check_launch("polar_generator"))) return st;
    }
    SEEMPS_HIP(hipMemcpyAsync(hp, w.stats, 16, hipMemcpyDeviceToHost, st_));
    SEEMPS_HIP(hipStreamSynchronize(st_));
    const double off = hp[0], xnorm = hp[1];
    if (!(off == off)) break;  // NaN: not converged
    if (off < tol || xnorm == 0.0) {
      converged = true;
      break;
    }
    // polar factor of Z: scaled so its singular values lie in (0, 1]
    const double f = 1.0 / sqrt(1.0 + xnorm * xnorm);
    if ((st = zaxpby_impl(l * l, make_double2(f, 0.0), w.Z, 0, zero, w.Z, st_))) return st;
    const int ns = newton_schulz_steps(f);
    for (int k = 0; k < ns; ++k) {
      // T = Z^H Z; Z2 = 1.5 Z - 0.5 Z T
      if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_N, l, l, l, one, w.Z, l, w.Z, l, zero, w.T, l,
                         w.sk, w.sk_elems, st_)))
        return st;
      if ((st = zaxpby_impl(l * l, make_double2(1.5, 0.0), w.Z, 0, zero, w.Z2, st_))) return st;
      if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_N, l, l, l, make_double2(-0.5, 0.0), w.Z, l,
                         w.T, l, one, w.Z2, l, w.sk, w.sk_elems, st_)))
        return st;
      double2* t = w.Z;
      w.Z = w.Z2;
      w.Z2 = t;
    }
    // [W; V] <- [W; V] Z
    if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_N, H, l, l, one, w.XA, l, w.Z, l, zero, w.XB, l,
                       w.sk, w.sk_elems, st_)))
      return st;
    double2* t = w.XA;
    w.XA = w.XB;
    w.XB = t;
    if (off < POLAR_QUAD_STOP) {
      converged = true;
      ++it;
      break;
    }
  }
  if (info) {
    info[0] = it;
    info[1] = converged ? 1 : 0;
  }
  if (!converged) {
    set_error("polar_svd: not converged after %d iterations", it);
    return SEEMPS_ERR_NOCONVERGE;
  }
  hipLaunchKernelGGL(polar_norms, dim3((unsigned)l), dim3(256), 0, st_, Lg, l, w.XA, w.nrm2);
  if ((st = check_launch("polar_norms"))) return st;
  hipLaunchKernelGGL(polar_tiny_reset, dim3(1), dim3(64), 0, st_, l, w.tiny);
  if ((st = check_launch("polar_tiny_reset"))) return st;
  hipLaunchKernelGGL(polar_rank, dim3(polar_grid(l)), dim3(256), 0, st_, l, w.nrm2, w.stats,
                     w.perm, w.tiny);
  if ((st = check_launch("polar_rank"))) return st;
  hipLaunchKernelGGL(polar_output, dim3(polar_grid(H * l)), dim3(256), 0, st_, Lg, l, w.XA,
                     w.nrm2, w.perm, U, S, VH);
  if ((st = check_launch("polar_output"))) return st;
  if (tiny_dev) *tiny_dev = w.tiny;
  return SEEMPS_OK;
}

}  // namespace seemps
