// Inner Hermitian Jacobi eigensolver on a 32 x 32 matrix held in LDS, shared
// by the block Jacobi SVD (svd.hip, one call per pair and step) and the
// spectral sketch of the learned truncation (sketch.hip).
#pragma once

#include "common.h"

namespace seemps {
namespace jac {

constexpr int PW = 32;       // matrix order (pair width of the block Jacobi)
constexpr int GP = PW + 1;   // padded LDS row stride (complex)

// 1/sqrt(x) and 1/x for positive normal x: the hardware estimates refined by
// two Newton steps each (quadratic convergence from the estimate's error to a
// few ulp), so a rotation costs four short FMA chains instead of the longer
// correctly rounded sqrt / division sequences.
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-(x * y), y, 1.0);  // 1 - x y^2
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-x, y, 1.0);  // 1 - x y
    y = fma(y, e, y);
  }
  return y;
}

// Rotation of the index pair (a, b) of the Hermitian matrix G (row stride GP)
// that annihilates G[a][b]: J = [[cs, sn ph], [-sn conj(ph), cs]], applied as
// G <- J^H G J.  Identity (sn = 0) when the pair is already orthogonal to
// tolerance (|G_ab| <= tol sqrt(G_aa G_bb)) or either column is negligible.
// t = tan(angle) solves t^2 + 2 tau t - 1 = 0 (smaller root); cs = 1/sqrt(1+t^2)
// and sn = t cs keep cs^2 + sn^2 = 1 to a few ulp for any t.
__device__ __forceinline__ void jacobi_rotation(const double2* G, int a, int b, double tol,
                                                double tiny2, double* cs, double* sn,
                                                double2* ph) {
  const double gaa = G[a * GP + a].x, gbb = G[b * GP + b].x;
  const double2 c = G[a * GP + b];
  const double s2 = cabs2(c);
  *cs = 1.0;
  *sn = 0.0;
  *ph = make_double2(1.0, 0.0);
  if (s2 > 0.0 && gaa > tiny2 && gbb > tiny2 && s2 > (tol * tol * gaa) * gbb) {
    const double r = rsqrt_nr(s2);  // 1 / |G_ab|
    *ph = make_double2(c.x * r, c.y * r);
    const double tau = (gbb - gaa) * (0.5 * r);
    double t_;
    if (fabs(tau) > 1e150) {
      t_ = 0.5 * rcp_nr(tau);
    } else {
      const double q = fma(tau, tau, 1.0);
      const double den = fabs(tau) + q * rsqrt_nr(q);
      t_ = (tau >= 0.0 ? 1.0 : -1.0) * rcp_nr(den);
    }
    *cs = rsqrt_nr(fma(t_, t_, 1.0));
    *sn = t_ * (*cs);
  }
}

// Index pair t (0 <= t < PW/2) of round k of the circle schedule over PW indices.
__device__ __forceinline__ void inner_pair(int k, int t, int* a, int* b) {
  auto idx = [&](int x) { return 1 + (x + k) % (PW - 1); };
  if (t == 0) {
    *a = 0;
    *b = idx(0);
  } else {
    *a = idx(t);
    *b = idx(PW - 1 - t);
  }
}

// Index pair t of round k (0 <= k < PW/2) of the cross-half schedule: index t
// of the first half against index PW/2 + (t + k) mod PW/2 of the second, so the
// PW/2 rounds visit every pair across the halves once and none within a half.
__device__ __forceinline__ void cross_pair(int k, int t, int* a, int* b) {
  *a = t;
  *b = PW / 2 + ((t + k) & (PW / 2 - 1));
}

// Inner cyclic Jacobi sweep(s) on the Hermitian PW x PW matrix G (LDS, row
// stride GP; G2 is a second buffer of the same shape), accumulating Q (LDS,
// on entry identity).  256 threads.  Each round: the first 16 lanes derive the
// round's 16 rotations into rot (one FP64 chain on the critical path instead
// of one per wavefront); then thread (p, q) applies J_p^H G[pair p, pair q]
// J_q to one 2x2 block, reading G and writing the other buffer, and updates
// two rows of Q for pair p; a round whose 16 rotations are all the identity
// skips that phase.  Returns the buffer holding the final G.  When nactive is
// given, *nactive (thread-private, the same value in every thread) receives
// the number of rounds that rotated anything.  cross: each sweep is the PW/2
// rounds of the cross-half schedule instead of the PW - 1 rounds of the circle.
__device__ inline double2* inner_jacobi(double2* G, double2* G2, double2* Q, double2* rot, int* active,
                                 double tol, double tiny2, int sweeps,
                                 int* nactive = nullptr, bool cross = false) {
  const int tid = threadIdx.x;
  const int p = tid >> 4, q = tid & 15;
  double2* cur = G;
  double2* nxt = G2;
  int rnd = 0;
  const int rounds = cross ? PW / 2 : PW - 1;
  auto pair_of = [cross](int k, int t, int* a, int* b) {
    if (cross)
      cross_pair(k, t, a, b);
    else
      inner_pair(k, t, a, b);
  };
  for (int sw = 0; sw < sweeps; ++sw) {
    for (int k = 0; k < rounds; ++k) {
      if (tid < 64) {
        int any = 0;
        if (tid < PW / 2) {
          int a, b;
          pair_of(k, tid, &a, &b);
          double cs, sn;
          double2 ph;
          jacobi_rotation(cur, a, b, tol, tiny2, &cs, &sn, &ph);
          rot[2 * tid] = make_double2(cs, sn);
          rot[2 * tid + 1] = ph;
          any = (sn != 0.0);
        }
        any = __any(any);
        if (tid == 0) active[rnd & 1] = any;
      }
      __syncthreads();
      // identity round: G and Q unchanged, buffers not swapped.  The flag is
      // double-buffered so a wavefront still reading it is never overtaken.
      if (!active[(rnd++) & 1]) continue;
      if (nactive) ++*nactive;
      int a, b, c, d;
      pair_of(k, p, &a, &b);
      pair_of(k, q, &c, &d);
      const double2 rp = rot[2 * p], php = rot[2 * p + 1];
      const double2 rq = rot[2 * q], phq = rot[2 * q + 1];
      const double csp = rp.x, snp = rp.y, csq = rq.x, snq = rq.y;
      // Y = X J_q on the block rows (a, b), columns (c, d)
      const double2 xac = cur[a * GP + c], xad = cur[a * GP + d];
      const double2 xbc = cur[b * GP + c], xbd = cur[b * GP + d];
      const double2 pqc = make_double2(phq.x, -phq.y);
      auto right = [&](double2 xc, double2 xd, double2* yc, double2* yd) {
        const double2 t1 = cmul(pqc, xd), t2 = cmul(phq, xc);
        *yc = make_double2(csq * xc.x - snq * t1.x, csq * xc.y - snq * t1.y);
        *yd = make_double2(snq * t2.x + csq * xd.x, snq * t2.y + csq * xd.y);
      };
      double2 yac, yad, ybc, ybd;
      right(xac, xad, &yac, &yad);
      right(xbc, xbd, &ybc, &ybd);
      // Z = J_p^H Y
      const double2 ppc = make_double2(php.x, -php.y);
      auto left = [&](double2 ya, double2 yb, double2* za, double2* zb) {
        const double2 t1 = cmul(php, yb), t2 = cmul(ppc, ya);
        *za = make_double2(csp * ya.x - snp * t1.x, csp * ya.y - snp * t1.y);
        *zb = make_double2(snp * t2.x + csp * yb.x, snp * t2.y + csp * yb.y);
      };
      double2 zac, zbc, zad, zbd;
      left(yac, ybc, &zac, &zbc);
      left(yad, ybd, &zad, &zbd);
      if (p == q && snp != 0.0) {
        // (c, d) == (a, b): the rotated pair is exactly orthogonal
        zad = make_double2(0.0, 0.0);
        zbc = make_double2(0.0, 0.0);
        zac.y = 0.0;
        zbd.y = 0.0;
      }
      nxt[a * GP + c] = zac;
      nxt[a * GP + d] = zad;
      nxt[b * GP + c] = zbc;
      nxt[b * GP + d] = zbd;
      // Q <- Q J_p on rows 2q, 2q + 1
      if (snp != 0.0) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int x = 2 * q + u;
          const double2 xa = Q[x * GP + a], xb = Q[x * GP + b];
          const double2 tb = cmul(ppc, xb), ta = cmul(php, xa);
          Q[x * GP + a] = make_double2(csp * xa.x - snp * tb.x, csp * xa.y - snp * tb.y);
          Q[x * GP + b] = make_double2(snp * ta.x + csp * xb.x, snp * ta.y + csp * xb.y);
        }
      }
      __syncthreads();
      double2* tmp = cur;
      cur = nxt;
      nxt = tmp;
    }
  }
  return cur;
}

}  // namespace jac
}  // namespace seemps
