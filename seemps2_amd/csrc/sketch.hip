// Fixed-rank randomized SVD split (the learned-rank truncation path).
//
// For an m x n matrix A and a sketch size l <= min(m, n):
//   Y = A Omega (Omega n x l complex Gaussian), optionally refined by power
//   iterations Y <- A (A^H Y) with re-orthonormalisation, Q = orth(Y)
//   (Householder), B = Q^H A (l x n), full Jacobi SVD of B, U = Q U_B.
// The truncation rule runs on the l singular values of B, with the unsketched
// weight ||A||_F^2 - ||B||_F^2 counted as discarded.  When that weight alone is
// more than the rule may discard and the bond cap does not bind below l (or a
// threshold rule would keep more than l values), the result is the full SVD
// split instead and the fallback flag is set.  Everything stays on the device; the host reads D, the error and the
// flag once, as seemps_svd_split does.
#include <cmath>
#include <initializer_list>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "common.h"
#include "jacobi_inner.h"

namespace seemps {
namespace {

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// out[e] = complex standard Gaussian (Box-Muller on two counter-hash uniforms).
__global__ void gauss_kernel(int64_t count, uint64_t seed, double2* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < count;
       e += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h1 = mix64(seed + 0x9e3779b97f4a7c15ull * (2 * (uint64_t)e + 1));
    const uint64_t h2 = mix64(h1 ^ 0xd1b54a32d192ed03ull);
    const double u1 = ((h1 >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0, 1]
    const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);        // [0, 1)
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    out[e] = make_double2(r * cs * 0.7071067811865476, r * sn * 0.7071067811865476);
  }
}

// Column-major identity block: column c (length len) = e_c, c < ncol.
__global__ void identity_cm(int64_t len, int64_t ncol, double2* __restrict__ out) {
  const int64_t total = len * ncol;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / len, r = e % len;
    out[e] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

// Truncation on the sketch spectrum s[0..l) with the unsketched weight
// frob - sum s^2 counted as discarded (see the file comment).  full = min(m, n).
// resid2 = ||A - Q Q^H A||_F^2 computed from the residual itself: the
// difference ||A||^2 - sum s^2 would carry rounding of order l eps ||A||^2,
// as large as the whole budget of a tolerance near machine epsilon.
__global__ void truncate_sketch_kernel(int64_t l, int64_t full, const double* __restrict__ s,
                                       const double2* __restrict__ frob2,
                                       const double2* __restrict__ resid2, int method, double tol,
                                       int64_t max_bond, int64_t* D_out, double* err_out,
                                       int* fallback_out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double total = frob2->x;
  const double tail = resid2->x > 0.0 ? resid2->x : 0.0;
  int64_t D = l;
  int fallback = 0;
  // A sketch decides a capped truncation only when what it did not see cannot
  // matter: the unsketched weight is negligible against the matrix, or far
  // below the smallest kept value (so no unseen direction competes with the
  // kept ones).  Otherwise (flat spectra, e.g. random circuits) the full SVD
  // decides.
  auto capped_ok = [&](int64_t Dk) {
    const double sD = s[Dk - 1];
    return tail <= 1e-14 * total || tail <= 1e-3 * sD * sD;
  };
  if (method == SEEMPS_TRUNC_NONE) {
    fallback = 1;
  } else if (method == SEEMPS_TRUNC_REL_NORM2) {
    const double max_error = total * tol;
    if (tail > max_error) {
      // the tolerance alone would keep more than l values; the sketch still
      // decides when the bond cap binds at or below l (certified as above)
      if (max_bond <= l) {
        D = max_bond;
        if (!capped_ok(D)) fallback = 1;
      } else {
        fallback = 1;
      }
    } else {
      double acc = tail;
      int64_t k = 0;
      for (int64_t i = l - 1; i >= 1; --i) {  // keep at least one value
        const double x = s[i] * s[i];
        if (acc + x > max_error) break;
        acc += x;
        ++k;
      }
      D = l - k;
    }
  } else {
    const double thr = (method == SEEMPS_TRUNC_REL_SV) ? tol * s[0] : tol;
    const int64_t cap_d = (full < max_bond) ? full : max_bond;
    D = (l < cap_d) ? l : cap_d;
    bool found = false;
    for (int64_t i = 1; i < D; ++i) {
      if (s[i] <= thr) {
        D = i;
        found = true;
        break;
      }
    }
    // the rule would look past the sketch: only the full spectrum can decide
    if (!found && D == l && l < cap_d) fallback = 1;
    // the cap decided below the sketch size: certify as above
    if (!found && D == cap_d && D < l && !capped_ok(D)) fallback = 1;
  }
  if (D > max_bond) D = max_bond;
  if (D < 1) D = 1;
  double e = tail;
  for (int64_t i = l - 1; i >= D; --i) e += s[i] * s[i];
  *D_out = D;
  *err_out = e;
  *fallback_out = fallback;
}

unsigned grid_for(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

constexpr int QRNB = 16;

// Workspace layout of the sketch path (the fallback reuses the whole buffer).
struct SketchWork {
  double2 *Om, *Ycm, *Vall, *Tall, *tau, *Z, *ks, *Qcm, *Zcm, *B, *Ub, *VHb, *U, *red, *frob2,
      *Rres, *resid2;
  double* S;
  int64_t* Ddev;
  double* edev;
  int* fdev;
  void* svdw;
  int64_t svd_bytes;
  double2* sk;  // split-K partial products (zgemm_ws) of the few-tile products
  int64_t sk_elems;
  int64_t bytes;
};

inline int64_t max_of(std::initializer_list<int64_t> xs) {
  int64_t r = 0;
  for (int64_t x : xs) r = x > r ? x : r;
  return r;
}

// base == nullptr only sizes the layout (w.bytes).
SketchWork sketch_carve(int64_t m, int64_t n, int64_t l, void* base) {
  const int64_t mx = m > n ? m : n;
  SketchWork w;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    void* q = base ? (void*)((char*)base + off) : nullptr;
    off += al256(bytes);
    return q;
  };
  w.Om = (double2*)take(l * n * 16);
  w.Ycm = (double2*)take(l * m * 16);
  w.Vall = (double2*)take(l * mx * 16);
  w.Tall = (double2*)take(ceil_div(l, QRNB) * QRNB * QRNB * 16);
  w.tau = (double2*)take(l * 16);
  w.Z = (double2*)take(2 * QRNB * (mx > l ? mx : l) * 16);
  w.ks = (double2*)take(8 * QRNB * l * 16);
  w.Qcm = (double2*)take(l * m * 16);
  w.Zcm = (double2*)take(l * n * 16);
  w.B = (double2*)take(l * n * 16);
  w.Ub = (double2*)take(l * l * 16);
  w.VHb = (double2*)take(l * n * 16);
  w.U = (double2*)take(m * l * 16);
  w.red = (double2*)take(reduce_blocks(m * n) * 16);
  w.frob2 = (double2*)take(16);
  w.Rres = (double2*)take(m * n * 16);
  w.resid2 = (double2*)take(16);
  w.S = (double*)take(l * 8);
  {
    // D, discarded weight and fallback flag side by side: one read to the host
    char* rk = (char*)take(32);
    w.Ddev = (int64_t*)rk;
    w.edev = rk ? (double*)(rk + 8) : nullptr;
    w.fdev = rk ? (int*)(rk + 16) : nullptr;
  }
  w.svd_bytes = svd_workspace_bytes(l, n);
  w.svdw = take(w.svd_bytes);
  w.sk_elems = max_of({splitk_workspace(l, m, n), splitk_workspace(l, n, m),
                       splitk_workspace(m, n, l), splitk_workspace(m, l, l)});
  w.sk = (double2*)take(w.sk_elems * 16);
  w.bytes = off;
  return w;
}

// Orthonormal basis of the l columns (length len) of Xcm, written to Qout
// (column-major, l columns of length len).  Xcm is overwritten.  WS supplies the
// QR scratch (Vall, Tall, tau, Z, ks) sized for l columns of length len.
template <class WS>
int orthonormalize(double2* Xcm, int64_t len, int64_t l, double2* Qout, const WS& w,
                   hipStream_t stream) {
  int st;
  if ((st = qr_factor(Xcm, len, l, QRNB, w.Vall, w.Tall, w.tau, w.Z, w.ks, 8 * QRNB * l, stream)))
    return st;
  hipLaunchKernelGGL(identity_cm, dim3(grid_for(len * l)), dim3(256), 0, stream, len, l, Qout);
  if ((st = check_launch("identity_cm"))) return st;
  return qr_apply_q(Qout, len, l, l, QRNB, w.Vall, w.Tall, w.Z, w.ks, 8 * QRNB * l, stream);
}

// ---------------------------------------------------------------------------
// Split sketched by the bond's existing basis.  In a canonical MPS the
// neighbour of the centre on the side the sweep moves to is an isometry whose
// span is the previous row (or column) space of the two-site matrix; after a
// local update that space moves only slightly, so Omega = [basis | p expansion
// directions] captures the matrix, and W = op(A) Omega has nearly orthogonal
// columns (an isometry times the previous singular values, whatever the gauge
// of the other side): the Jacobi SVD of W needs no QR preconditioner and
// converges in a few sweeps.  The expansion directions (one power step on a
// Gaussian block, projected off the basis twice and orthonormalised) give the
// truncation candidates outside the old space, as the two-site update needs.

// Oh (l x S, row-major) = Omega^H: rows j < k from the basis (side 0: the
// k x S row-orthonormal basis itself; side 1: the adjoint of the S x k
// column-orthonormal basis), rows k + q = conj(column q of Qz), Qz column-major.
__global__ void build_omega_h(int64_t k, int64_t p, int64_t S, int side,
                              const double2* __restrict__ basis, const double2* __restrict__ Qz,
                              double2* __restrict__ Oh) {
  const int64_t total = (k + p) * S;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / S, c = e % S;
    double2 v;
    if (j < k) {
      if (side == 0) {
        v = basis[j * S + c];
      } else {
        const double2 b = basis[c * k + j];
        v = make_double2(b.x, -b.y);
      }
    } else {
      const double2 q = Qz[(j - k) * S + c];
      v = make_double2(q.x, -q.y);
    }
    Oh[e] = v;
  }
}

// dst[i][j] = conj(src[j * ld + i]) * (scale ? scale[i] : 1), i < rows, j < cols.
__global__ void adjoint_scale(int64_t rows, int64_t cols, const double2* __restrict__ src,
                              int64_t ld, const double* __restrict__ scale,
                              double2* __restrict__ dst) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, j = e % cols;
    const double f = scale ? scale[i] : 1.0;
    const double2 v = src[j * ld + i];
    dst[e] = make_double2(f * v.x, -f * v.y);
  }
}

// Eigenvectors of the p x p Hermitian Gram matrix G2 (p <= PW, row-major) of
// the expansion block, by cyclic Jacobi sweeps in LDS (one workgroup of 256
// threads; padding rows and columns are zero and get no rotation), into Q2
// (p x p, row-major, unitary): the expansion directions rotated by Q2 give the
// sketched matrix columns that are orthogonal among themselves.
__global__ __launch_bounds__(256) void expansion_rotation(int p, const double2* __restrict__ G2,
                                                          double2* __restrict__ Q2) {
  using jac::GP;
  using jac::PW;
  __shared__ double2 g[PW * GP], g2[PW * GP], q[PW * GP], rot[PW];
  __shared__ int active[2];
  __shared__ double dmax;
  for (int e = threadIdx.x; e < PW * PW; e += blockDim.x) {
    const int i = e / PW, j = e % PW;
    g[i * GP + j] = (i < p && j < p) ? G2[i * p + j] : make_double2(0.0, 0.0);
    q[i * GP + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < p; ++i) m = fmax(m, g[i * GP + i].x);
    dmax = m;
  }
  __syncthreads();
  // columns below (PW eps)^2 of the largest are left alone; pairs orthogonal
  // to 1e-6 relative get no rotation.  (Stopping at 1e-3 after three sweeps
  // measured twice the rotation iteration's time on these matrices: the
  // expansion images need to be orthogonal well below the basis columns' own
  // coupling.)
  const double tiny2 = (double)PW * PW * 4.9e-32 * dmax;
  double2* cur = g;
  double2* other = g2;
  for (int sweep = 0; sweep < 16; ++sweep) {
    int nact = 0;
    double2* res = jac::inner_jacobi(cur, other, q, rot, active, 1e-6, tiny2, 1, &nact);
    if (res != cur) {
      other = cur;
      cur = res;
    }
    if (nact == 0) break;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < p * p; e += blockDim.x) {
    const int i = e / p, j = e % p;
    Q2[e] = q[i * GP + j];
  }
}

// Workspace of the basis-sketched split: Lg x Sh working orientation (Lg = m,
// Sh = n for side 0; Lg = n, Sh = m for side 1), l = k + p sketch columns.
struct BasisWork {
  double2 *G0t, *Yt, *Zt, *C, *Qz, *Vall, *Tall, *tau, *Z, *ks, *Oh, *W, *Uw, *VHw, *T, *red,
      *frob2, *Rres, *resid2, *G2, *Q2;
  double* S;
  int64_t* Ddev;
  double* edev;
  int* fdev;
  void* svdw;
  int64_t svd_bytes;
  double2* sk;  // split-K partial products (zgemm_ws) of the few-tile products
  int64_t sk_elems;
  int64_t bytes;
};

BasisWork basis_carve(int64_t Lg, int64_t Sh, int64_t k, int64_t p, void* base) {
  const int64_t l = k + p, mx = Lg > Sh ? Lg : Sh, pp = p > 0 ? p : 1;
  BasisWork w;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    void* q = base ? (void*)((char*)base + off) : nullptr;
    off += al256(bytes);
    return q;
  };
  w.G0t = (double2*)take(pp * Sh * 16);
  w.Yt = (double2*)take(pp * Lg * 16);
  w.Zt = (double2*)take(pp * Sh * 16);
  w.C = (double2*)take(pp * k * 16);
  w.Qz = (double2*)take(pp * Sh * 16);
  w.Vall = (double2*)take(pp * Sh * 16);
  w.Tall = (double2*)take(ceil_div(pp, QRNB) * QRNB * QRNB * 16);
  w.tau = (double2*)take(pp * 16);
  w.Z = (double2*)take(2 * QRNB * (Sh > pp ? Sh : pp) * 16);
  w.ks = (double2*)take(8 * QRNB * pp * 16);
  w.Oh = (double2*)take(l * Sh * 16);
  w.W = (double2*)take(Lg * l * 16);
  w.Uw = (double2*)take(Lg * l * 16);
  w.VHw = (double2*)take(l * l * 16);
  w.T = (double2*)take(l * mx * 16);
  w.red = (double2*)take(reduce_blocks(Lg * Sh) * 16);
  w.frob2 = (double2*)take(16);
  w.Rres = (double2*)take(Lg * Sh * 16);
  w.resid2 = (double2*)take(16);
  w.S = (double*)take(l * 8);
  {
    // D, discarded weight and fallback flag side by side: one read to the host
    char* rk = (char*)take(32);
    w.Ddev = (int64_t*)rk;
    w.edev = rk ? (double*)(rk + 8) : nullptr;
    w.fdev = rk ? (int*)(rk + 16) : nullptr;
  }
  w.G2 = (double2*)take(pp * pp * 16);
  w.Q2 = (double2*)take(pp * pp * 16);
  w.svd_bytes = svd_workspace_bytes(Lg, l, false);
  if (polar_svd_bytes(Lg, l) > w.svd_bytes) w.svd_bytes = polar_svd_bytes(Lg, l);
  w.svdw = take(w.svd_bytes);
  // the factor products (D <= l columns) split when D = l fits
  w.sk_elems = max_of({splitk_workspace(pp, Lg, Sh), splitk_workspace(pp, Sh, Lg),
                       splitk_workspace(Lg, pp, Sh), splitk_workspace(pp, pp, Lg),
                       splitk_workspace(pp, Sh, pp),
                       splitk_workspace(pp, k, Sh), splitk_workspace(pp, Sh, k),
                       splitk_workspace(Lg, l, Sh), splitk_workspace(Lg, Sh, l),
                       splitk_workspace(l, Sh, l), splitk_workspace(Sh, l, l)});
  w.sk = (double2*)take(w.sk_elems * 16);
  w.bytes = off;
  return w;
}

constexpr int NF = 32;  // sketch columns = features of the rank predictor

// Features of the rank predictor from the 32 x 32 Gram matrix G = Y^H Y of the
// sketch Y = A Omega, one workgroup of 256 threads: G (scaled by its largest
// diagonal entry) is reduced to a real symmetric tridiagonal by Householder
// reflections, and each eigenvalue is found by 18 rounds of 9-section on
// Sturm counts (8 lanes per eigenvalue, one shift each; the interval shrinks
// 9^18 ~ 1.5e17 times from the Gershgorin bounds).  Sorted descending:
// f[0] = log10(min(m, n)) / 4, f[i] = clamp(log10(sigma_i / sigma_0), -8, 0) / 8
// for i = 1..31 (the Gram of the sketch resolves ratios down to ~sqrt(eps),
// hence the clamp).
__global__ __launch_bounds__(256) void sketch_features_kernel(const double2* __restrict__ G,
                                                              int64_t r, float* __restrict__ f) {
  using jac::GP;
  constexpr int n = NF;
  constexpr double PIVMIN = 1e-290;  // smallest pivot magnitude of a Sturm count
  __shared__ double2 A[n * GP];
  __shared__ double2 v[n], pw[n];   // reflector, then p = beta A v and w = p - K v
  __shared__ double d[n], e2[n];    // tridiagonal: diagonal, squared off-diagonal
  __shared__ double lam[n];         // eigenvalues, ascending
  __shared__ double s_scale, s_beta;
  __shared__ int s_skip[2];
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) mx = fmax(mx, G[i * n + i].x);
    s_scale = mx;
  }
  __syncthreads();
  const double scale = s_scale;
  const double inv = scale > 0.0 ? 1.0 / scale : 0.0;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const double2 g = G[e];
    A[(e / n) * GP + e % n] = make_double2(g.x * inv, g.y * inv);
  }
  __syncthreads();
  for (int j = 0; j < n - 2; ++j) {
    const int L = n - 1 - j;  // trailing rows / columns j + 1 .. n - 1
    if (tid < 64) {
      // reflector of column j below the diagonal: v = x - alpha e_0,
      // alpha = -phase(x_0) |x|, so that H x = alpha e_0 with H = I - beta v v^H
      const double2 x = (lane < L) ? A[(j + 1 + lane) * GP + j] : make_double2(0.0, 0.0);
      double tail2 = (lane >= 1) ? cabs2(x) : 0.0;
      for (int o = 32; o > 0; o >>= 1) tail2 += __shfl_xor(tail2, o, 64);
      const double2 x0 = make_double2(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64));
      const double a02 = cabs2(x0), a0 = sqrt(a02);
      const int skip = !(tail2 > 1e-200);  // column already reduced (or below any resolution)
      if (lane == 0) {
        d[j] = A[j * GP + j].x;
        e2[j] = skip ? a02 : a02 + tail2;
        s_skip[j & 1] = skip;
      }
      if (!skip) {
        const double xn = sqrt(a02 + tail2);
        const double2 ph = a0 > 0.0 ? make_double2(x0.x / a0, x0.y / a0) : make_double2(1.0, 0.0);
        if (lane < L) v[lane] = (lane == 0) ? make_double2(ph.x * (a0 + xn), ph.y * (a0 + xn)) : x;
        if (lane == 0) s_beta = 2.0 / ((a0 + xn) * (a0 + xn) + tail2);
      }
    }
    __syncthreads();
    if (s_skip[j & 1]) continue;  // the flag is double-buffered: wave 0 may run ahead
    const double beta = s_beta;
    {
      // p = beta A v: 8 lanes per row
      const int t = tid >> 3, sub = tid & 7;
      double2 acc = make_double2(0.0, 0.0);
      if (t < L)
        for (int k = sub; k < L; k += 8) acc = cfma(A[(j + 1 + t) * GP + j + 1 + k], v[k], acc);
      for (int o = 1; o < 8; o <<= 1) {
        acc.x += __shfl_xor(acc.x, o, 64);
        acc.y += __shfl_xor(acc.y, o, 64);
      }
      if (t < L && sub == 0) pw[t] = make_double2(beta * acc.x, beta * acc.y);
    }
    __syncthreads();
    if (tid < 64) {
      // w = p - K v, K = (beta / 2) v^H p (real: A is Hermitian)
      const double2 vv = lane < L ? v[lane] : make_double2(0.0, 0.0);
      const double2 pp = lane < L ? pw[lane] : make_double2(0.0, 0.0);
      double dot = vv.x * pp.x + vv.y * pp.y;
      for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
      const double K = 0.5 * beta * dot;
      if (lane < L) pw[lane] = make_double2(pp.x - K * vv.x, pp.y - K * vv.y);
    }
    __syncthreads();
    // trailing block: A <- A - v w^H - w v^H
    for (int e = tid; e < L * L; e += blockDim.x) {
      const int rr = e / L, cc = e % L;
      const double2 vr = v[rr], wr = pw[rr], vc = v[cc], wc = pw[cc];
      double2 a = A[(j + 1 + rr) * GP + j + 1 + cc];
      a.x -= vr.x * wc.x + vr.y * wc.y + wr.x * vc.x + wr.y * vc.y;
      a.y -= vr.y * wc.x - vr.x * wc.y + wr.y * vc.x - wr.x * vc.y;
      A[(j + 1 + rr) * GP + j + 1 + cc] = a;
    }
    __syncthreads();
  }
  if (tid == 0) {
    d[n - 2] = A[(n - 2) * GP + n - 2].x;
    e2[n - 2] = cabs2(A[(n - 1) * GP + n - 2]);
    d[n - 1] = A[(n - 1) * GP + n - 1].x;
  }
  __syncthreads();
  double dd[n], ee[n - 1];
  double glo = 1e300, ghi = -1e300;
#pragma unroll
  for (int i = 0; i < n; ++i) {
    dd[i] = d[i];
    if (i < n - 1) ee[i] = e2[i];
  }
#pragma unroll
  for (int i = 0; i < n; ++i) {
    const double rad = (i > 0 ? sqrt(ee[i - 1]) : 0.0) + (i < n - 1 ? sqrt(ee[i]) : 0.0);
    glo = fmin(glo, dd[i] - rad);
    ghi = fmax(ghi, dd[i] + rad);
  }
  const double span = ghi - glo;
  double lo = glo - 1e-14 * span - 1e-300, hi = ghi + 1e-14 * span + 1e-300;
  // eigenvalue i (ascending) by the 8-lane group of this thread; lane g of the
  // group counts the eigenvalues below lo + (g + 1) h, h = (hi - lo) / 9
  const int i = tid >> 3, g = tid & 7;
  for (int it = 0; it < 18; ++it) {
    const double h = (hi - lo) * (1.0 / 9.0);
    const double x = lo + (double)(g + 1) * h;
    double q = dd[0] - x;
    if (fabs(q) < PIVMIN) q = -PIVMIN;
    int cnt = q < 0.0;
#pragma unroll
    for (int k = 1; k < n; ++k) {
      q = (dd[k] - x) - ee[k - 1] * jac::rcp_nr(q);
      if (fabs(q) < PIVMIN) q = -PIVMIN;
      cnt += q < 0.0;
    }
    // first shift of the group with more than i eigenvalues below it
    const unsigned long long m = __ballot(cnt > i);
    const unsigned bits = (unsigned)(m >> ((lane >> 3) * 8)) & 0xFFu;
    const int jj = bits ? __builtin_ctz(bits) : 8;
    const double nlo = lo + (double)jj * h;
    const double nhi = (jj == 8) ? hi : lo + (double)(jj + 1) * h;
    lo = nlo;
    hi = nhi;
  }
  if (g == 0) lam[i] = 0.5 * (lo + hi) * scale;
  __syncthreads();
  if (tid == 0) f[0] = (float)(log10((double)r) / 4.0);
  if (tid >= 1 && tid < n) {
    // descending order: sigma_k^2 = lam[n - 1 - k]
    const double l0 = lam[n - 1], li = lam[n - 1 - tid];
    double val = (l0 > 0.0 && li > 0.0) ? 0.5 * log10(li / l0) : -8.0;
    val = val < -8.0 ? -8.0 : (val > 0.0 ? 0.0 : val);
    f[tid] = (float)(val / 8.0);
  }
}

}  // namespace
}  // namespace seemps

namespace seemps {
namespace {
// split-K scratch of the spectral sketch's two products (after Omega, Y^T and G)
int64_t spectral_sk_elems(int64_t m, int64_t n) {
  return max_of({splitk_workspace(NF, m, n), splitk_workspace(NF, NF, m)});
}
}  // namespace
}  // namespace seemps

extern "C" int64_t seemps_spectral_sketch_workspace(int64_t m, int64_t n) {
  using namespace seemps;
  auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  return al(NF * n * 16) + al(NF * m * 16) + al(NF * NF * 16) + spectral_sk_elems(m, n) * 16;
}

extern "C" int seemps_spectral_sketch(int64_t m, int64_t n, const void* A, uint64_t seed,
                                      float* features, void* work, int64_t work_bytes,
                                      void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(m >= 1 && n >= 1, "spectral_sketch: empty matrix");
  SEEMPS_CHECK_ARG(work_bytes >= seemps_spectral_sketch_workspace(m, n),
                   "spectral_sketch: workspace too small");
  hipStream_t st_ = (hipStream_t)stream;
  auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  char* p = (char*)work;
  double2* Om = (double2*)p;
  p += al(NF * n * 16);
  double2* Yt = (double2*)p;
  p += al(NF * m * 16);
  double2* G = (double2*)p;
  p += al(NF * NF * 16);
  double2* sk = (double2*)p;
  const int64_t sk_elems = spectral_sk_elems(m, n);
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  int st;
  hipLaunchKernelGGL(gauss_kernel, dim3(grid_for(NF * n)), dim3(256), 0, st_, NF * n, seed, Om);
  if ((st = check_launch("gauss"))) return st;
  // Y^T = Omega^T A^T (32 x m); G = conj(Y^T) (Y^T)^T = Y^H Y
  if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, NF, m, n, one, Om, n, A, n, zero, Yt, m, sk,
                     sk_elems, st_)))
    return st;
  if ((st = zgemm_ws(SEEMPS_OP_C, SEEMPS_OP_T, NF, NF, m, one, Yt, m, Yt, m, zero, G, NF, sk,
                     sk_elems, st_)))
    return st;
  hipLaunchKernelGGL(sketch_features_kernel, dim3(1), dim3(256), 0, st_, G, m < n ? m : n,
                     features);
  return check_launch("sketch_features");
}

extern "C" int64_t seemps_svd_split_sketch_workspace(int64_t m, int64_t n, int64_t l) {
  const int64_t full = seemps::svd_split_bytes(m, n);
  const int64_t r = m < n ? m : n;
  if (l < 1 || l > r) return full;
  const int64_t sk = seemps::sketch_carve(m, n, l, nullptr).bytes;
  return sk > full ? sk : full;
}

extern "C" int seemps_svd_split_sketch(int64_t m, int64_t n, const void* A_, int mode, int method,
                                       double tolerance, int64_t max_bond, int64_t l,
                                       int64_t power_iters, uint64_t seed, void* outA,
                                       void* outB, double* s_out, int64_t* D_host,
                                       double* err_host, int32_t* fallback_host, void* work,
                                       int64_t work_bytes, int32_t* info, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(m >= 1 && n >= 1, "svd_split_sketch: empty matrix");
  SEEMPS_CHECK_ARG(mode == 0 || mode == 1, "svd_split_sketch: invalid mode");
  SEEMPS_CHECK_ARG(power_iters >= 0 && power_iters <= 8, "svd_split_sketch: invalid power_iters");
  SEEMPS_CHECK_ARG(work_bytes >= seemps_svd_split_sketch_workspace(m, n, l),
                   "svd_split_sketch: workspace too small");
  hipStream_t st_ = (hipStream_t)stream;
  const double2* A = (const double2*)A_;
  const int64_t r = m < n ? m : n;
  if (fallback_host) *fallback_host = 0;
  auto full_split = [&]() {
    if (fallback_host) *fallback_host = 1;
    return svd_split_impl(m, n, A, mode, method, tolerance, max_bond, (double2*)outA,
                          (double2*)outB, s_out, D_host, err_host, work, work_bytes, info, st_);
  };
  // nothing to gain from a sketch as wide as the matrix, or no rule to decide with
  if (l < 1 || l >= r || method == SEEMPS_TRUNC_NONE) return full_split();
  const SketchWork w = sketch_carve(m, n, l, work);
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
  int st;
  {
    TimedLaunch phase_range("phase_sketch_range", st_, 0.0);
    // Y^T = Omega^T A^T, stored column-major (l columns of length m)
    hipLaunchKernelGGL(gauss_kernel, dim3(grid_for(l * n)), dim3(256), 0, st_, l * n, seed, w.Om);
    if ((st = check_launch("gauss"))) return st;
    if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, l, m, n, one, w.Om, n, A, n, zero, w.Ycm, m, w.sk,
                       w.sk_elems, st_)))
      return st;
    for (int64_t it = 0; it < power_iters; ++it) {
      if ((st = orthonormalize(w.Ycm, m, l, w.Qcm, w, st_))) return st;
      // Z^T = Q^T conj(A): the columns of A^H Q
      if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_C, l, n, m, one, w.Qcm, m, A, n, zero, w.Zcm, n,
                         w.sk, w.sk_elems, st_)))
        return st;
      if ((st = orthonormalize(w.Zcm, n, l, w.B, w, st_))) return st;  // B holds Qz here
      // Y^T = Qz^T A^T
      if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, l, m, n, one, w.B, n, A, n, zero, w.Ycm, m, w.sk,
                         w.sk_elems, st_)))
        return st;
    }
    if ((st = orthonormalize(w.Ycm, m, l, w.Qcm, w, st_))) return st;
    // B = Q^H A (l x n)
    if ((st = zgemm_ws(SEEMPS_OP_C, SEEMPS_OP_N, l, n, m, one, w.Qcm, m, A, n, zero, w.B, n, w.sk,
                       w.sk_elems, st_)))
      return st;
  }
  int off = 0;
  {
    TimedLaunch phase_svd("phase_sketch_svd", st_, 0.0);
    if ((st = svd_enqueue(l, n, w.B, n, w.Ub, w.S, w.VHb, w.svdw, w.svd_bytes, st_, &off,
                          false)))
      return st;
  }
  if ((st = zmdotc_impl(m * n, 1, A, m * n, A, w.frob2, w.red, st_))) return st;
  // R = A - Q B (m x n), Q^T stored l x m
  SEEMPS_HIP(hipMemcpyAsync(w.Rres, A, m * n * 16, hipMemcpyDeviceToDevice, st_));
  if ((st = zgemm_ws(SEEMPS_OP_T, SEEMPS_OP_N, m, n, l, make_double2(-1.0, 0.0), w.Qcm, m, w.B, n,
                     one, w.Rres, n, w.sk, w.sk_elems, st_)))
    return st;
  if ((st = zmdotc_impl(m * n, 1, w.Rres, m * n, w.Rres, w.resid2, w.red, st_))) return st;
  hipLaunchKernelGGL(truncate_sketch_kernel, dim3(1), dim3(64), 0, st_, l, r, w.S, w.frob2,
                     w.resid2, method, tolerance, max_bond, w.Ddev, w.edev, w.fdev);
  if ((st = check_launch("truncate_sketch"))) return st;
  if ((st = svd_complete_enqueue(l, n, w.Ub, w.VHb, w.svdw, w.Ddev, st_))) return st;
  char* hp = (char*)pinned_stage(PINNED_BYTES);
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  SEEMPS_HIP(hipMemcpyAsync(hp, w.Ddev, 20, hipMemcpyDeviceToHost, st_));
  if ((st = svd_read_ctrl(w.svdw, off, info, st_))) return st;  // synchronises
  int64_t hD = 0;
  double herr = 0.0;
  int hfb = 0;
  memcpy(&hD, hp, 8);
  memcpy(&herr, hp + 8, 8);
  memcpy(&hfb, hp + 16, sizeof(int));
  if (hfb) return full_split();
  const int64_t D = hD;
  // U = Q U_B (m x l): (Q^T)^T U_B with Q^T stored l x m
  if ((st = zgemm_ws(SEEMPS_OP_T, SEEMPS_OP_N, m, D, l, one, w.Qcm, m, w.Ub, l, zero, w.U, l, w.sk,
                     w.sk_elems, st_)))
    return st;
  if (mode == 0) {
    if ((st = zscale2d_impl(m, D, nullptr, nullptr, w.U, l, (double2*)outA, D, st_))) return st;
    if ((st = zscale2d_impl(D, n, w.S, nullptr, w.VHb, n, (double2*)outB, n, st_))) return st;
  } else {
    if ((st = zscale2d_impl(m, D, nullptr, w.S, w.U, l, (double2*)outA, D, st_))) return st;
    if ((st = zscale2d_impl(D, n, nullptr, nullptr, w.VHb, n, (double2*)outB, n, st_))) return st;
  }
  if (s_out) SEEMPS_HIP(hipMemcpyAsync(s_out, w.S, l * 8, hipMemcpyDeviceToDevice, st_));
  *D_host = D;
  *err_host = herr;
  return SEEMPS_OK;
}

extern "C" int64_t seemps_svd_split_basis_workspace(int64_t m, int64_t n, int side, int64_t k,
                                                    int64_t p) {
  const int64_t full = seemps::svd_split_bytes(m, n);
  const int64_t Lg = side == 0 ? m : n, Sh = side == 0 ? n : m;
  if (k < 1 || p < 0 || k + p > (Lg < Sh ? Lg : Sh)) return full;
  const int64_t bk = seemps::basis_carve(Lg, Sh, k, p, nullptr).bytes;
  return bk > full ? bk : full;
}

extern "C" int seemps_svd_split_basis(int64_t m, int64_t n, const void* A_, int side,
                                      const void* basis, int64_t k, int64_t p, uint64_t seed,
                                      int mode, int method, double tolerance, int64_t max_bond,
                                      void* outA, void* outB, double* s_out, int64_t* D_host,
                                      double* err_host, int32_t* fallback_host, void* work,
                                      int64_t work_bytes, int32_t* info, void* stream) {
  using namespace seemps;
  SEEMPS_CHECK_ARG(m >= 1 && n >= 1, "svd_split_basis: empty matrix");
  SEEMPS_CHECK_ARG(mode == 0 || mode == 1, "svd_split_basis: invalid mode");
  SEEMPS_CHECK_ARG(side == 0 || side == 1, "svd_split_basis: invalid side");
  SEEMPS_CHECK_ARG(k >= 1 && p >= 0, "svd_split_basis: invalid sketch sizes");
  SEEMPS_CHECK_ARG(work_bytes >= seemps_svd_split_basis_workspace(m, n, side, k, p),
                   "svd_split_basis: workspace too small");
  hipStream_t st_ = (hipStream_t)stream;
  const double2* A = (const double2*)A_;
  const double2* Bs = (const double2*)basis;
  const int64_t r = m < n ? m : n, l = k + p;
  const int64_t Lg = side == 0 ? m : n, Sh = side == 0 ? n : m;
  if (fallback_host) *fallback_host = 0;
  auto full_split = [&]() {
    if (fallback_host) *fallback_host = 1;
    return svd_split_impl(m, n, A, mode, method, tolerance, max_bond, (double2*)outA,
                          (double2*)outB, s_out, D_host, err_host, work, work_bytes, info, st_);
  };
  if (l > (Lg < Sh ? Lg : Sh) || method == SEEMPS_TRUNC_NONE) return full_split();
  const BasisWork w = basis_carve(Lg, Sh, k, p, work);
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0),
                mone = make_double2(-1.0, 0.0);
  int st;
  {
    TimedLaunch phase("phase_basis_sketch", st_, 0.0);
    if (p > 0) {
      // Expansion block Z = X^H X G0 (Sh x p), X = A (side 0) or A^H (side 1),
      // kept transposed: Zt = Z^T (p x Sh), which is Z column-major.
      hipLaunchKernelGGL(gauss_kernel, dim3(grid_for(p * Sh)), dim3(256), 0, st_, p * Sh, seed,
                         w.G0t);
      if ((st = check_launch("gauss"))) return st;
      if (side == 0) {
        // Yt = G0^T A^T (p x m); Zt = Yt conj(A) (p x n)
        if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, p, m, n, one, w.G0t, n, A, n, zero, w.Yt, m,
                           w.sk, w.sk_elems, st_)))
          return st;
        if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_C, p, n, m, one, w.Yt, m, A, n, zero, w.Zt, n,
                           w.sk, w.sk_elems, st_)))
          return st;
      } else {
        // Yt = G0^T conj(A) (p x n); Zt = Yt A^T (p x m)
        if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_C, p, n, m, one, w.G0t, m, A, n, zero, w.Yt, n,
                           w.sk, w.sk_elems, st_)))
          return st;
        if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, p, m, n, one, w.Yt, n, A, n, zero, w.Zt, m,
                           w.sk, w.sk_elems, st_)))
          return st;
      }
      // X -= Q (Q^H X) for the p columns of Xt (column-major, length Sh);
      // transposed: C = Xt conj(Q) (p x k), Xt -= C Q^T.
      auto project = [&](double2* Xt) -> int {
        int e;
        if (side == 0) {  // Q^H = basis (k x n): conj(Q) = basis^T, Q^T = conj(basis)
          if ((e = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, p, k, n, one, Xt, n, Bs, n, zero, w.C, k,
                            w.sk, w.sk_elems, st_)))
            return e;
          return zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_C, p, n, k, mone, w.C, k, Bs, n, one, Xt, n, w.sk,
                          w.sk_elems, st_);
        }
        // Q = basis (m x k): conj(Q) = conj(basis), Q^T = basis^T
        if ((e = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_C, p, k, m, one, Xt, m, Bs, k, zero, w.C, k, w.sk,
                          w.sk_elems, st_)))
          return e;
        return zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, p, m, k, mone, w.C, k, Bs, k, one, Xt, m, w.sk,
                        w.sk_elems, st_);
      };
      // Project twice, orthonormalise; then once more of each: when A lies (almost)
      // in the basis span, Z is rank-deficient and the Householder factor's extra
      // columns are arbitrary, so they must be projected off the basis again.
      if ((st = project(w.Zt)) || (st = project(w.Zt))) return st;
      if ((st = orthonormalize(w.Zt, Sh, p, w.Qz, w, st_))) return st;
      if ((st = project(w.Qz)) || (st = project(w.Qz))) return st;
      if ((st = orthonormalize(w.Qz, Sh, p, w.Zt, w, st_))) return st;
    }
    if (p > 0 && p <= jac::PW) {
      // Rotate the expansion directions so that their images X Z are
      // orthogonal among themselves (they are already nearly orthogonal to
      // the images of the basis): W = X [basis | Z] then has nearly orthogonal
      // columns throughout and its SVD is the rotation iteration's.
      // W2 = X Z (Lg x p) in Yt, G2 = W2^H W2, Q2 its eigenvectors, Zt <- Q2^T Zt.
      if (side == 0) {
        if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_T, m, p, n, one, A, n, w.Zt, n, zero, w.Yt, p,
                           w.sk, w.sk_elems, st_)))
          return st;
      } else {
        if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_T, n, p, m, one, A, n, w.Zt, m, zero, w.Yt, p,
                           w.sk, w.sk_elems, st_)))
          return st;
      }
      if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_N, p, p, Lg, one, w.Yt, p, w.Yt, p, zero, w.G2, p,
                         w.sk, w.sk_elems, st_)))
        return st;
      hipLaunchKernelGGL(expansion_rotation, dim3(1), dim3(256), 0, st_, (int)p, w.G2, w.Q2);
      if ((st = check_launch("expansion_rotation"))) return st;
      if ((st = zgemm_ws(SEEMPS_OP_T, SEEMPS_OP_N, p, Sh, p, one, w.Q2, p, w.Zt, Sh, zero, w.G0t,
                         Sh, w.sk, w.sk_elems, st_)))
        return st;
      SEEMPS_HIP(hipMemcpyAsync(w.Zt, w.G0t, p * Sh * 16, hipMemcpyDeviceToDevice, st_));
    }
    const double2* Qext = w.Zt;  // p orthonormal columns orthogonal to the basis
    hipLaunchKernelGGL(build_omega_h, dim3(grid_for(l * Sh)), dim3(256), 0, st_, k, p, Sh, side,
                       Bs, Qext, w.Oh);
    if ((st = check_launch("build_omega_h"))) return st;
    // W = X Omega = X Oh^H (Lg x l)
    if (side == 0) {
      if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_H, m, l, n, one, A, n, w.Oh, n, zero, w.W, l, w.sk,
                         w.sk_elems, st_)))
        return st;
    } else {
      if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_H, n, l, m, one, A, n, w.Oh, m, zero, w.W, l, w.sk,
                         w.sk_elems, st_)))
        return st;
    }
  }
  int off = 0;
  // W = X [basis | expansion].  The basis columns are the new two-site matrix
  // seen through the previous singular vectors, orthogonal up to the change of
  // the two-site tensor, and the expansion block was rotated above so that its
  // images are orthogonal among themselves (and nearly so to the rest): the
  // SVD of W is the simultaneous rotation iteration of polar_svd.hip, a few
  // GEMM iterations, for up to jac::PW expansion columns.  The block Jacobi
  // (which converges from any start) takes wider sketches and any matrix the
  // rotation iteration does not finish.  SEEMPS_BASIS_SVD=jacobi uses the
  // block Jacobi throughout.
  static const bool polar_enabled = [] {
    const char* e = getenv("SEEMPS_BASIS_SVD");
    return !(e && strcmp(e, "jacobi") == 0);
  }();
  char* hp = (char*)pinned_stage(PINNED_BYTES);
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  int64_t hD = 0, htiny = l;
  double herr = 0.0;
  int hfb = 0, hconv = 1;
  // SVD of W, residual of the projection, truncation on the sketch spectrum,
  // then one synchronisation that reads D, the discarded weight, the
  // fallback flag and (rotation iteration) its convergence and first
  // negligible slot.
  auto svd_and_truncate = [&](bool polar) -> int {
    int e;
    const int64_t* tiny_dev = nullptr;
    const int* conv_dev = nullptr;
    {
      TimedLaunch phase("phase_basis_svd", st_, 0.0);
      if (polar) {
        if ((e = polar_svd(Lg, l, w.W, l, w.Uw, w.S, w.VHw, &tiny_dev, &conv_dev, w.svdw,
                           w.svd_bytes, st_)))
          return e;
      } else if ((e = svd_enqueue(Lg, l, w.W, l, w.Uw, w.S, w.VHw, w.svdw, w.svd_bytes, st_,
                                  &off, false, false))) {
        return e;
      }
    }
    if ((e = zmdotc_impl(m * n, 1, A, m * n, A, w.frob2, w.red, st_))) return e;
    // R = A - (projection of A on Omega): side 0: A - W Oh; side 1: A - Oh^H W^H
    SEEMPS_HIP(hipMemcpyAsync(w.Rres, A, m * n * 16, hipMemcpyDeviceToDevice, st_));
    if (side == 0) {
      if ((e = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_N, m, n, l, mone, w.W, l, w.Oh, n, one, w.Rres, n,
                        w.sk, w.sk_elems, st_)))
        return e;
    } else {
      if ((e = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_H, m, n, l, mone, w.Oh, m, w.W, l, one, w.Rres, n,
                        w.sk, w.sk_elems, st_)))
        return e;
    }
    if ((e = zmdotc_impl(m * n, 1, w.Rres, m * n, w.Rres, w.resid2, w.red, st_))) return e;
    hipLaunchKernelGGL(truncate_sketch_kernel, dim3(1), dim3(64), 0, st_, l, r, w.S, w.frob2,
                       w.resid2, method, tolerance, max_bond, w.Ddev, w.edev, w.fdev);
    if ((e = check_launch("truncate_sketch"))) return e;
    if (polar) {
      SEEMPS_HIP(hipMemcpyAsync(hp, w.Ddev, 20, hipMemcpyDeviceToHost, st_));
      SEEMPS_HIP(hipMemcpyAsync(hp + 24, tiny_dev, 8, hipMemcpyDeviceToHost, st_));
      SEEMPS_HIP(hipMemcpyAsync(hp + 32, conv_dev, sizeof(int), hipMemcpyDeviceToHost, st_));
      SEEMPS_HIP(hipStreamSynchronize(st_));
      memcpy(&htiny, hp + 24, 8);
      memcpy(&hconv, hp + 32, sizeof(int));
      if (info) {
        info[0] = 0;
        info[1] = hconv;
      }
    } else {
      if ((e = svd_complete_enqueue(Lg, l, w.Uw, w.VHw, w.svdw, w.Ddev, st_, false))) return e;
      SEEMPS_HIP(hipMemcpyAsync(hp, w.Ddev, 20, hipMemcpyDeviceToHost, st_));
      if ((e = svd_read_ctrl(w.svdw, off, info, st_))) return e;  // synchronises
      htiny = l;
      hconv = 1;
    }
    memcpy(&hD, hp, 8);
    memcpy(&herr, hp + 8, 8);
    memcpy(&hfb, hp + 16, sizeof(int));
    return SEEMPS_OK;
  };
  bool use_polar = polar_enabled && p <= jac::PW;
  if ((st = svd_and_truncate(use_polar))) return st;
  if (use_polar && !hconv) {
    if (getenv("SEEMPS_POLAR_TRACE"))
      fprintf(stderr, "[basis split %lld x %lld, k %lld] rotation iteration not converged: "
              "block Jacobi\n", (long long)m, (long long)n, (long long)k);
    use_polar = false;
    if ((st = svd_and_truncate(false))) return st;
  }
  // a kept slot with a negligible singular value has no vector from the
  // rotation iteration: the full SVD (with its orthonormal completion) decides
  if (hfb || htiny < hD) {
    if (use_polar && getenv("SEEMPS_POLAR_TRACE"))
      fprintf(stderr, "[basis split %lld x %lld, k %lld, p %lld] fallback: %s (D %lld, tiny %lld)\n",
              (long long)m, (long long)n, (long long)k, (long long)p,
              hfb ? "sketch weight" : "negligible kept value", (long long)hD, (long long)htiny);
    return full_split();
  }
  const int64_t D = hD;
  const double* sA = (mode == 1) ? w.S : nullptr;  // S into the left factor
  const double* sB = (mode == 0) ? w.S : nullptr;  // S into the right factor
  if (side == 0) {
    // A = Uw S (VHw Oh): left = Uw[:, :D], right = VHw[:D] Oh (D x n)
    if ((st = zgemm_ws(SEEMPS_OP_N, SEEMPS_OP_N, D, n, l, one, w.VHw, l, w.Oh, n, zero, w.T, n,
                       w.sk, w.sk_elems, st_)))
      return st;
    if ((st = zscale2d_impl(m, D, nullptr, sA, w.Uw, l, (double2*)outA, D, st_))) return st;
    if ((st = zscale2d_impl(D, n, sB, nullptr, w.T, n, (double2*)outB, n, st_))) return st;
  } else {
    // A^H = Uw S (VHw Oh): A = (Oh^H VHw^H) S Uw^H; left = Oh^H VHw[:D]^H (m x D),
    // right = Uw[:, :D]^H (D x n)
    if ((st = zgemm_ws(SEEMPS_OP_H, SEEMPS_OP_H, m, D, l, one, w.Oh, m, w.VHw, l, zero, w.T, D,
                       w.sk, w.sk_elems, st_)))
      return st;
    if ((st = zscale2d_impl(m, D, nullptr, sA, w.T, D, (double2*)outA, D, st_))) return st;
    hipLaunchKernelGGL(adjoint_scale, dim3(grid_for(D * n)), dim3(256), 0, st_, D, n, w.Uw, l, sB,
                       (double2*)outB);
    if ((st = check_launch("adjoint_scale"))) return st;
  }
  if (s_out) SEEMPS_HIP(hipMemcpyAsync(s_out, w.S, l * 8, hipMemcpyDeviceToDevice, st_));
  *D_host = D;
  *err_host = herr;
  return SEEMPS_OK;
}
