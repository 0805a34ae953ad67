// FP64 complex SVD (QR-preconditioned one-sided block Jacobi) and the Strategy
// truncation rule.
//
// Pipeline for an m x n matrix A (B = A if m >= n, else B = A^H; B is
// len x ncol with len >= ncol):
//   1. B is copied column-major (column c contiguous).
//   2. If ncol > 16: blocked Householder QR, B = Q R (qr.hip).  The Jacobi then
//      works on X = R^H (ncol x ncol); Jacobi on R^H converges in a handful of
//      sweeps even for graded spectra (Drmac-Veselic preconditioning).
//      Otherwise the Jacobi works on B directly.
//   3. One-sided block Jacobi on the working matrix W (columns of X or B):
//      column blocks of width BW, every pair of blocks visited once per sweep
//      in a round-robin (circle) schedule; per pair the 2BW x 2BW Gram matrix is
//      formed (partial sums over row chunks in different workgroups), an inner
//      Hermitian Jacobi sweep accumulates Q, and W_pq <- W_pq Q, V_pq <- V_pq Q.
//      Stops when every normalised off-diagonal |G_ij| / sqrt(G_ii G_jj) seen in
//      a sweep is below 32 sqrt(len) eps (relative orthogonality), when it was
//      already below 1e-9 (quadratic convergence: that sweep's rotations leave
//      ~1e-18), or when it stalls within 16x of the target.
//   4. sigma_k = ||w_k||, sorted descending; the W side gives the normalised
//      singular vectors of the working matrix, V (times Q in the QR path) the
//      other side.  Exactly-zero singular values get an orthonormal completion.
//
// The Jacobi runs as one cooperative persistent kernel.  With one work item
// per workgroup, steps are ordered by per-pair and per-row-chunk counters and
// one bounded grid barrier per sweep reduces the convergence measure;
// otherwise two bounded grid barriers per block step.  A workgroup that waits
// too long raises an abort flag that every waiter observes, so every wave
// always exits.
//
// Replaces _destructive_svd (src/seemps/cython/svd.pxi:17-83) and the
// truncation rules of src/seemps/cython/truncation.pxi:222-334.
#include <cmath>
#include <cstdlib>

#include "common.h"
#include "jacobi_inner.h"

namespace seemps {

namespace {

using jac::GP;
using jac::inner_jacobi;
using jac::PW;

constexpr int BW = PW / 2;      // block width (columns); PW, GP from jacobi_inner.h
constexpr int RT = 32;          // rows per LDS tile (the 2 x 2 blocked loops assume 32)
constexpr int RTP = RT + 1;     // padded tile row length (complex)
constexpr int RMAX = 16;        // max row chunks per pair
// Relative off-diagonal measure from which one more sweep is known to suffice
// (quadratic convergence: the sweep leaves ~1e-18 against a target ~1e-13).
constexpr double QUAD_STOP = 1e-9;
constexpr int QRNB = 16;        // QR panel width
constexpr int QR_MIN_COLS = 17; // use the QR preconditioner from this many columns
constexpr int MAX_SWEEPS = 60;
constexpr unsigned long long SPIN_LIMIT = 4000000ull;

struct Ctrl {
  unsigned int barrier;       // grid barrier arrivals
  unsigned int abort;         // set when a barrier wait timed out
  unsigned long long off[3];  // per-sweep max normalised off-diagonal (as bits), slot sweep % 3
  int info[4];                // sweeps, converged, aborted, unused
  long long prof[8];          // wall-clock ticks of workgroup 0 per phase
  double frob2;               // ||W||_F^2 (invariant under the rotations)
  double tiny2;               // squared norm below which a column is negligible
  double offhist[64];         // off-diagonal measure of each sweep (diagnostics)
};

struct SvdShape {
  int64_t m, n;         // input shape
  int64_t lenA, ncol;   // working matrix B: ncol columns of length lenA
  int64_t len;          // Jacobi column length (ncol with QR, lenA without)
  int64_t nb, ncolp;    // blocks (even), padded column count
  int64_t P, R;         // pairs per step, row chunks
  int transposed;       // 1 if B = A^H
  int use_qr;
  int cross_inner;      // inner solves after a sweep's first step: cross-block pairs only
};

struct SvdWork {
  double2* Acm;   // ncol x lenA (QR in place)
  double2* Vall;  // ncol x lenA (QR vectors)
  double2* Tall;  // npanels x QRNB x QRNB
  double2* tau;   // ncol
  double2* Z;     // 2 x QRNB x max(ncol, ncolp)
  double2* Y;     // ncol x lenA (Q V)
  double2* ks;    // split-K partials of the QR panel products
  double2* W;     // ncolp x len
  double2* V;     // ncolp x ncolp
  double2* part;  // 2 x P x RMAX x PW x PW (partial Grams, by step parity)
  unsigned int* sync;  // P pair arrival counters, then P x R chunk step counters
  double* colnrm;  // ncol squared column norms of B (QR path)
  int* colperm;    // ncol: B's columns by decreasing norm (QR path)
  double* sigma;  // ncolp
  int* perm;      // ncolp
  Ctrl* ctrl;
};

inline int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }

SvdShape svd_shape(int64_t m, int64_t n, int num_cu, bool precondition = true) {
  SvdShape s;
  s.m = m;
  s.n = n;
  s.transposed = (m < n);
  s.lenA = s.transposed ? n : m;
  s.ncol = s.transposed ? m : n;
  s.use_qr = precondition && (s.ncol >= QR_MIN_COLS);
  s.len = s.use_qr ? s.ncol : s.lenA;
  int64_t nb = ceil_div(s.ncol, BW);
  if (nb < 2) nb = 2;
  nb += (nb & 1);
  s.nb = nb;
  s.ncolp = nb * BW;
  s.P = nb / 2;
  // Row chunks per pair: one workgroup per CU (cooperative residency caps the
  // grid there).  The QR-preconditioned matrix takes about half that while
  // at least 8 chunks remain: fewer, longer chunks shorten each step's
  // reduction and waits, and on 512 columns (16 pairs) R = 8 beat R = 16 over
  // whole DMRG sweeps; on the basis-sketched matrices R = 16 stayed ahead.
  int64_t R = num_cu / s.P;
  if (s.use_qr) {
    int64_t half = num_cu / (2 * s.P);
    if (half < 8) half = 8;
    if (half < R) R = half;
  }
  // tuning knob: cap on the row chunks per pair (SEEMPS_SVD_ROW_CHUNKS)
  static const int64_t r_cap = [] {
    const char* e = getenv("SEEMPS_SVD_ROW_CHUNKS");
    const long v = e ? atol(e) : 0;
    return (int64_t)(v > 0 ? v : RMAX);
  }();
  if (R > r_cap) R = r_cap;
  const int64_t rmax_len = ceil_div(s.len, RT);
  if (R > rmax_len) R = rmax_len;
  if (R > RMAX) R = RMAX;
  if (R < 1) R = 1;
  s.R = R;
  // Inner solves: every block's own pairs are visited at the first step of a
  // sweep (each block meets one partner there); with cross_inner the later
  // steps rotate only the pairs across the two blocks (16 rounds instead of
  // 31).  Measured over whole chi=256 DMRG sweeps this is faster in both the
  // QR-preconditioned and the basis-sketched splits (more sweeps at times,
  // each about half as long); the convergence test is the same either way.
  // SEEMPS_SVD_INNER=full restores the full inner sweep at every step, and
  // SEEMPS_SVD_INNER=hybrid keeps it for the first sweep only (tuning: hybrid
  // converges the sketch matrices in fewer sweeps, but lost over DMRG sweeps).
  static const int inner_mode = [] {
    const char* e = getenv("SEEMPS_SVD_INNER");
    if (e && strcmp(e, "full") == 0) return 0;
    if (e && strcmp(e, "hybrid") == 0) return 2;  // full inner sweeps in sweep 0
    return 1;
  }();
  s.cross_inner = s.nb > 2 ? inner_mode : 0;
  return s;
}

// Split-K scratch of the NB-row products of the QR (factor and apply).
// (Up to 8 slices of NB x ncol, enough for every panel's shape.)
inline int64_t qr_ks_elems(const SvdShape& s) { return 8 * QRNB * s.ncol; }

#define SEEMPS_SVD_REGIONS(X)                                         \
  X(Acm, s.use_qr ? s.ncol * s.lenA : 0, double2)                     \
  X(Vall, s.use_qr ? s.ncol * s.lenA : 0, double2)                    \
  X(Tall, s.use_qr ? ceil_div(s.ncol, QRNB) * QRNB * QRNB : 0, double2) \
  X(tau, s.use_qr ? s.ncol : 0, double2)                              \
  X(Z, s.use_qr ? 2 * QRNB * (s.ncolp > s.lenA ? s.ncolp : s.lenA) : 0, double2) \
  X(Y, s.use_qr ? s.ncol * s.lenA : 0, double2)                       \
  X(ks, s.use_qr ? qr_ks_elems(s) : 0, double2)                       \
  X(W, s.ncolp * s.len, double2)                                      \
  X(V, s.ncolp * s.ncolp, double2)                                    \
  X(part, 2 * s.P * RMAX * PW * PW, double2)                          \
  X(sync, s.P + s.P * s.R, unsigned int)                              \
  X(colnrm, s.use_qr ? s.ncol : 0, double)                            \
  X(colperm, s.use_qr ? s.ncol : 0, int)                              \
  X(sigma, s.ncolp, double)                                           \
  X(perm, s.ncolp, int)                                               \
  X(ctrl, 1, Ctrl)

int64_t svd_bytes(const SvdShape& s) {
  int64_t b = 0;
#define X(name, count, type) b += align256((int64_t)(count) * (int64_t)sizeof(type));
  SEEMPS_SVD_REGIONS(X)
#undef X
  return b;
}

SvdWork svd_carve(const SvdShape& s, void* work) {
  char* p = (char*)work;
  SvdWork w;
#define X(name, count, type)                                   \
  w.name = (type*)p;                                           \
  p += align256((int64_t)(count) * (int64_t)sizeof(type));
  SEEMPS_SVD_REGIONS(X)
#undef X
  return w;
}

// B column-major: dst[c * lenA + r] = A[r][c] (m >= n) or conj(A[c][r]) (m < n),
// zero for padded columns c >= ncol (dst has ncols_dst columns).
__global__ void load_colmajor(SvdShape s, const double2* __restrict__ A, int64_t lda,
                              double2* __restrict__ dst, int64_t ncols_dst,
                              const int* __restrict__ perm) {
  const int64_t total = ncols_dst * s.lenA;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / s.lenA, r = e % s.lenA;
    double2 v = make_double2(0.0, 0.0);
    if (c < s.ncol) {
      const int64_t cs = perm ? perm[c] : c;  // column of B stored at position c
      if (!s.transposed) {
        v = A[r * lda + cs];
      } else {
        v = A[cs * lda + r];
        v.y = -v.y;
      }
    }
    dst[e] = v;
  }
}

// Squared norms of the ncol columns of B (B = A, or A^H when transposed).
__global__ void col_norms(SvdShape s, const double2* __restrict__ A, int64_t lda,
                          double* __restrict__ nrm) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < s.ncol;
       c += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    if (!s.transposed) {
      for (int64_t r = 0; r < s.lenA; ++r) acc += cabs2(A[r * lda + c]);
    } else {
      for (int64_t r = 0; r < s.lenA; ++r) acc += cabs2(A[c * lda + r]);
    }
    nrm[c] = acc;
  }
}

// perm[rank] = c with decreasing norm, ties by index (the QR then sees the
// heaviest columns first, which leaves R^H graded the way one-sided Jacobi
// converges fast on).
__global__ void col_rank(SvdShape s, const double* __restrict__ nrm, int* __restrict__ perm) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < s.ncol;
       c += (int64_t)gridDim.x * blockDim.x) {
    // NaN keys sort last, so the ranks always form a permutation
    const double nc = isnan(nrm[c]) ? -1.0 : nrm[c];
    int rank = 0;
    for (int64_t d = 0; d < s.ncol; ++d) {
      const double nd = isnan(nrm[d]) ? -1.0 : nrm[d];
      rank += (nd > nc || (nd == nc && d < c)) ? 1 : 0;
    }
    perm[rank] = (int)c;
  }
}

// W[c][k] = conj(R[c][k]) = conj(Acm[k][c]) for k >= c (columns of X = R^H).
__global__ void load_w_from_r(SvdShape s, const double2* __restrict__ Acm,
                              double2* __restrict__ W) {
  const int64_t total = s.ncolp * s.len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / s.len, k = e % s.len;
    double2 v = make_double2(0.0, 0.0);
    if (c < s.ncol && k >= c) {
      v = Acm[k * s.lenA + c];
      v.y = -v.y;
    }
    W[e] = v;
  }
}

// V = identity on the real columns; control block reset.
__global__ void init_v_ctrl(SvdShape s, SvdWork w) {
  const int64_t totalV = s.ncolp * s.ncolp;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < totalV;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / s.ncolp, r = e % s.ncolp;
    w.V[e] = make_double2((c == r && c < s.ncol) ? 1.0 : 0.0, 0.0);
  }
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < s.P + s.P * s.R;
       e += (int64_t)gridDim.x * blockDim.x)
    w.sync[e] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.ctrl->barrier = 0;
    w.ctrl->abort = 0;
    w.ctrl->off[0] = 0;
    w.ctrl->off[1] = 0;
    w.ctrl->off[2] = 0;
    for (int i = 0; i < 4; ++i) w.ctrl->info[i] = 0;
    for (int i = 0; i < 8; ++i) w.ctrl->prof[i] = 0;
  }
}

// ctrl->frob2 = ||W||_F^2 (fixed summation order), ctrl->tiny2 = (len eps)^2 frob2.
// One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void svd_frob(SvdShape s, SvdWork w) {
  __shared__ double red[16];
  const int64_t total = s.ncolp * s.len;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < total; e += blockDim.x) acc += cabs2(w.W[e]);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += red[q];
    const double le = (double)s.len * 2.220446049250313e-16;
    w.ctrl->frob2 = t;
    w.ctrl->tiny2 = le * le * t;
  }
}

// Y[c][r] = V[c][r] for r < ncol, 0 below (input of Q V).
__global__ void load_y(SvdShape s, SvdWork w) {
  const int64_t total = s.ncol * s.lenA;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e / s.lenA, r = e % s.lenA;
    w.Y[e] = (r < s.ncol) ? w.V[c * s.ncolp + r] : make_double2(0.0, 0.0);
  }
}

// Bounded grid barrier.  Returns false if the barrier was aborted.
__device__ bool grid_barrier(Ctrl* ctrl, unsigned int target, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // publish, then arrive
    atomicAdd(&ctrl->barrier, 1u);
    unsigned long long spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&ctrl->barrier, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           target) {
      if (__hip_atomic_load(&ctrl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (++spins > SPIN_LIMIT) {
        atomicExch(&ctrl->abort, 1u);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see every arrival's stores
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

// Block indices of pair p in round k of the circle schedule over nb blocks.
__device__ __forceinline__ void pair_blocks(int64_t nb, int64_t k, int64_t p, int64_t* bi,
                                            int64_t* bj) {
  auto idx = [&](int64_t x) { return 1 + (x + k) % (nb - 1); };
  if (p == 0) {
    *bi = 0;
    *bj = idx(0);
  } else {
    *bi = idx(p);
    *bj = idx(nb - 1 - p);
  }
}

// Column index (in W / V) of local column t (0 <= t < PW) of the pair.
__device__ __forceinline__ int64_t pair_col(int64_t bi, int64_t bj, int t) {
  return (t < BW) ? bi * BW + t : bj * BW + (t - BW);
}

// Rotate rows [w0, w1) of the pair's PW columns of W (column stride ldw) and
// rows [v0, v1) of those of V (stride ldv) by Q: M_pq <- M_pq Q.  Both RT-row
// tiles are staged in LDS together (tw, tv), so one load latency covers both.
__device__ void rotate_rows(double2* __restrict__ W, int64_t ldw, int64_t w0, int64_t w1,
                            double2* __restrict__ V, int64_t ldv, int64_t v0, int64_t v1,
                            int64_t bi, int64_t bj, const double2* Q, double2* tw,
                            double2* tv) {
  const int tid = threadIdx.x;
  const int64_t nw = ceil_div(w1 - w0, (int64_t)RT), nv = ceil_div(v1 - v0, (int64_t)RT);
  const int64_t nblk = nw > nv ? nw : nv;
  for (int64_t blk = 0; blk < nblk; ++blk) {
    const int64_t rbw = w0 + blk * RT, rbv = v0 + blk * RT;
    const int nrw = (blk < nw) ? (int)min((int64_t)RT, w1 - rbw) : 0;
    const int nrv = (blk < nv) ? (int)min((int64_t)RT, v1 - rbv) : 0;
    for (int e = tid; e < PW * RT; e += blockDim.x) {
      const int t = e / RT, rr = e % RT;
      tw[t * RTP + rr] = (rr < nrw) ? W[pair_col(bi, bj, t) * ldw + rbw + rr] : make_double2(0, 0);
      tv[t * RTP + rr] = (rr < nrv) ? V[pair_col(bi, bj, t) * ldv + rbv + rr] : make_double2(0, 0);
    }
    __syncthreads();
    {
      // thread (jb, r0): columns jb, jb + 16 by rows r0, r0 + 16 of both tiles
      const int jb = tid >> 4, r0 = tid & 15;
      double2 w00 = make_double2(0.0, 0.0), w01 = w00, w10 = w00, w11 = w00;
      double2 v00 = w00, v01 = w00, v10 = w00, v11 = w00;
#pragma unroll 4
      for (int k = 0; k < PW; ++k) {
        const double2 q0 = Q[k * GP + jb], q1 = Q[k * GP + jb + 16];
        const double2 x0 = tw[k * RTP + r0], x1 = tw[k * RTP + r0 + 16];
        const double2 y0 = tv[k * RTP + r0], y1 = tv[k * RTP + r0 + 16];
        w00 = cfma(x0, q0, w00);
        w01 = cfma(x1, q0, w01);
        w10 = cfma(x0, q1, w10);
        w11 = cfma(x1, q1, w11);
        v00 = cfma(y0, q0, v00);
        v01 = cfma(y1, q0, v01);
        v10 = cfma(y0, q1, v10);
        v11 = cfma(y1, q1, v11);
      }
      double2* W0 = W + pair_col(bi, bj, jb) * ldw + rbw;
      double2* W1 = W + pair_col(bi, bj, jb + 16) * ldw + rbw;
      double2* V0 = V + pair_col(bi, bj, jb) * ldv + rbv;
      double2* V1 = V + pair_col(bi, bj, jb + 16) * ldv + rbv;
      if (r0 < nrw) {
        W0[r0] = w00;
        W1[r0] = w10;
      }
      if (r0 + 16 < nrw) {
        W0[r0 + 16] = w01;
        W1[r0 + 16] = w11;
      }
      if (r0 < nrv) {
        V0[r0] = v00;
        V1[r0] = v10;
      }
      if (r0 + 16 < nrv) {
        V0[r0 + 16] = v01;
        V1[r0 + 16] = v11;
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ long long now_ticks() { return (long long)wall_clock64(); }

// Pair owning block x in round k of the circle schedule (inverse of pair_blocks).
__device__ __forceinline__ int64_t block_owner(int64_t nb, int64_t k, int64_t x) {
  if (x == 0) return 0;
  const int64_t y = ((x - 1 - k) % (nb - 1) + (nb - 1)) % (nb - 1);  // idx(y) == x
  if (y == 0) return 0;
  return (y < nb / 2) ? y : nb - 1 - y;
}

// Phase A: partial Gram matrix of rows [r0, r1) of the pair's columns into dst
// (PW x PW, row-major).  Thread (ib, jb) accumulates the 2 x 2 block of
// entries (ib + 16x, jb + 16y): four LDS reads per row feed four complex FMAs.
__device__ void partial_gram(const SvdShape& s, const SvdWork& w, int64_t bi, int64_t bj,
                             int64_t r0, int64_t r1, double2* __restrict__ dst, double2* tile) {
  static_assert(PW == 32, "the 2 x 2 register blocking assumes 256 threads and PW = 32");
  const int tid = threadIdx.x;
  const int ib = tid >> 4, jb = tid & 15;
  double2 a00 = make_double2(0, 0), a01 = a00, a10 = a00, a11 = a00;
  for (int64_t rb = r0; rb < r1; rb += RT) {
    const int nr = (int)min((int64_t)RT, r1 - rb);
    for (int e = tid; e < PW * RT; e += blockDim.x) {
      const int t = e / RT, rr = e % RT;
      tile[t * RTP + rr] =
          (rr < nr) ? w.W[pair_col(bi, bj, t) * s.len + rb + rr] : make_double2(0, 0);
    }
    __syncthreads();
    const double2* x0 = tile + ib * RTP;
    const double2* x1 = tile + (ib + 16) * RTP;
    const double2* y0 = tile + jb * RTP;
    const double2* y1 = tile + (jb + 16) * RTP;
    for (int rr = 0; rr < nr; ++rr) {
      const double2 u0 = x0[rr], u1 = x1[rr], v0 = y0[rr], v1 = y1[rr];
      a00 = cfmac(u0, v0, a00);
      a01 = cfmac(u0, v1, a01);
      a10 = cfmac(u1, v0, a10);
      a11 = cfmac(u1, v1, a11);
    }
    __syncthreads();
  }
  dst[ib * PW + jb] = a00;
  dst[ib * PW + jb + 16] = a01;
  dst[(ib + 16) * PW + jb] = a10;
  dst[(ib + 16) * PW + jb + 16] = a11;
}

// Phases B + C for item (p, r): G = sum of the R partials at src, the pair's
// off-diagonal measure (r == 0) into *off_slot, inner Jacobi, rotation of this
// chunk's rows of W and V.  prof/pt/t0: phase timing of workgroup 0.
__device__ void pair_update(const SvdShape& s, const SvdWork& w, int64_t bi, int64_t bj,
                            int64_t r, const double2* __restrict__ src,
                            unsigned long long* off_slot, double tol, double tiny2,
                            int inner_sweeps, bool cross, double2* G, double2* G2,
                            double2* Q, double2* rot, int* s_active, double2* tile, double* s_off,
                            bool prof, long long* pt, long long* t0) {
  const int tid = threadIdx.x;
  const int64_t rc = ceil_div(s.len, s.R);
  const int64_t rcv = ceil_div(s.ncolp, s.R);
  // Upper triangle from the partials; the lower one mirrors it, so G is
  // exactly Hermitian.
  for (int e = tid; e < PW * PW; e += blockDim.x) {
    const int i = e / PW, j = e % PW;
    Q[i * GP + j] = make_double2((i == j) ? 1.0 : 0.0, 0.0);
    if (i > j) continue;
    double2 gq[RMAX];
#pragma unroll
    for (int q = 0; q < RMAX; ++q) gq[q] = (q < s.R) ? src[q * PW * PW + e] : make_double2(0.0, 0.0);
    double2 g = gq[0];
#pragma unroll
    for (int q = 1; q < RMAX; ++q) g = cadd(g, gq[q]);
    if (i == j) g.y = 0.0;
    G[i * GP + j] = g;
    G[j * GP + i] = make_double2(g.x, -g.y);
  }
  __syncthreads();
  // max normalised off-diagonal over non-negligible column pairs (every
  // workgroup of the pair derives the same value from the same G)
  double mx = 0.0;
  for (int e = tid; e < PW * PW; e += blockDim.x) {
    const int i = e / PW, j = e % PW;
    if (i < j) {
      const double gi = G[i * GP + i].x, gj = G[j * GP + j].x;
      if (gi > tiny2 && gj > tiny2) {
        const double2 gij = G[i * GP + j];
        mx = fmax(mx, sqrt(cabs2(gij) / gi / gj));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((tid & 63) == 0) s_off[tid >> 6] = mx;
  __syncthreads();
  double pair_off = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) pair_off = fmax(pair_off, s_off[q]);
  if (r == 0 && tid == 0) atomicMax(off_slot, (unsigned long long)__double_as_longlong(pair_off));
  if (pair_off <= tol) {
    // already orthogonal to tolerance: no rotation is applied to this pair
    if (prof) {
      const long long t1 = now_ticks();
      pt[2] += t1 - *t0;
      *t0 = t1;
    }
    __syncthreads();
    return;
  }
  if (prof) {
    const long long t1 = now_ticks();
    pt[2] += t1 - *t0;
    *t0 = t1;
  }
  inner_jacobi(G, G2, Q, rot, s_active, tol, tiny2, inner_sweeps, nullptr, cross);
  if (prof) {
    const long long t1 = now_ticks();
    pt[3] += t1 - *t0;
    *t0 = t1;
  }
  // G2 is free after the inner solve: it holds the V tile
  rotate_rows(w.W, s.len, r * rc, min(s.len, r * rc + rc), w.V, s.ncolp, r * rcv,
              min(s.ncolp, r * rcv + rcv), bi, bj, Q, tile, G2);
  if (prof) {
    const long long t1 = now_ticks();
    pt[4] += t1 - *t0;
    *t0 = t1;
  }
}

// Bounded wait (thread 0 polls, the block follows) until every counter in
// ctr[0..n) has reached target.  Returns false if the wait was aborted.
__device__ bool wait_counters(Ctrl* ctrl, const unsigned int* c0, const unsigned int* c1,
                              unsigned int target, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long spins = 0;
    int ok = 1;
    while (__hip_atomic_load(c0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target ||
           (c1 && __hip_atomic_load(c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)) {
      if (__hip_atomic_load(&ctrl->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = 0;
        break;
      }
      if (++spins > SPIN_LIMIT) {
        atomicExch(&ctrl->abort, 1u);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see the signallers' stores
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

// Publish this block's writes, then add 1 to / set a counter.
// Signals publish this workgroup's stores: the barrier gathers them in thread
// 0, whose release fence precedes the counter update; a wait that observes the
// counter acquires them (and its barrier passes them on to its workgroup).
__device__ void signal_add(unsigned int* c) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(c, 1u);
  }
}
__device__ void signal_set(unsigned int* c, unsigned int v) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(c, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Work item (pair p, row chunk r) of workgroup b when every item has its own
// workgroup.  Workgroups are dispatched round-robin over the XCDs (b mod 8), so
// the items are laid out XCD by XCD: each XCD takes whole pairs first (their
// partial Grams then meet in that XCD's L2 instead of crossing XCDs through
// memory), and the pairs left over fill the remaining slots.  A bijection of
// [0, G) onto the items for every G = P R.
constexpr int64_t NXCD = 8;
__device__ inline void local_item(int64_t b, int64_t G, int64_t R, int64_t* p, int64_t* r) {
  if (G % NXCD) {
    *p = b / R;
    *r = b % R;
    return;
  }
  const int64_t per = G / NXCD, x = b % NXCD, j = b / NXCD;
  const int64_t f = per / R;  // whole pairs per XCD
  if (j < f * R) {
    *p = x * f + j / R;
    *r = j % R;
    return;
  }
  const int64_t q = x * (per - f * R) + (j - f * R);
  *p = NXCD * f + q / R;
  *r = q % R;
}

// One-sided block Jacobi sweeps.  Work item (p, r) = row chunk r of pair p.
// LOCAL: one item per workgroup (gridDim.x == P R); the steps are ordered by
// local counters only -- the R workgroups of a pair meet after their partial
// Grams (pair counter), and a workgroup starts a step once the two pairs that
// held its blocks in the previous step have rotated its row chunk (chunk step
// counters).  One grid barrier per sweep reduces the convergence measure.
// !LOCAL: items are strided over the grid, with two grid barriers per step.
template <bool LOCAL>
__global__ __launch_bounds__(256) void svd_jacobi(SvdShape s, SvdWork w, double tol,
                                                  int max_sweeps) {
  __shared__ double2 tile[PW * RTP];
  __shared__ double2 G[PW * GP];
  __shared__ double2 G2[PW * GP];
  __shared__ double2 Q[PW * GP];
  __shared__ double2 rot[PW];
  __shared__ int s_flag;
  __shared__ int s_active[2];
  __shared__ double s_off[8];

  const int tid = threadIdx.x;
  const bool prof = (blockIdx.x == 0 && tid == 0);
  long long pt[6] = {0, 0, 0, 0, 0, 0};
  long long t0 = 0;
  const int64_t items = s.P * s.R;
  const int64_t rc = ceil_div(s.len, s.R);  // rows of W per chunk
  const int64_t steps = s.nb - 1;
  const int inner_sweeps = (s.nb == 2) ? 3 : 1;
  const double tiny2 = w.ctrl->tiny2;
  const size_t part_stride = (size_t)s.P * RMAX * PW * PW;
  unsigned int* pair_ctr = w.sync;
  unsigned int* chunk_ctr = w.sync + s.P;
  unsigned int barrier_no = 0;
  const unsigned int G_ = gridDim.x;
  int sweep = 0;
  int converged = 0;
  double off_prev = 1.0e300;
  for (; sweep < max_sweeps; ++sweep) {
    // Slot (sweep + 1) % 3 was last read after the grid barrier ending sweep
    // sweep - 2; every workgroup has passed the barrier ending sweep - 1 since.
    if (blockIdx.x == 0 && tid == 0) w.ctrl->off[(sweep + 1) % 3] = 0ull;
    unsigned long long* off_slot = &w.ctrl->off[sweep % 3];
    for (int64_t k = 0; k < steps; ++k) {
      const int64_t g = (int64_t)sweep * steps + k;
      const bool cross = s.cross_inner && k != 0 && (s.cross_inner == 1 || sweep > 0);
      double2* part = w.part + (g & 1) * part_stride;
      if (prof) t0 = now_ticks();
      if (LOCAL) {
        int64_t p, r;
        local_item(blockIdx.x, gridDim.x, s.R, &p, &r);
        int64_t bi, bj;
        pair_blocks(s.nb, k, p, &bi, &bj);
        if (g > 0) {
          const int64_t kp = (k == 0) ? steps - 1 : k - 1;
          const int64_t oi = block_owner(s.nb, kp, bi), oj = block_owner(s.nb, kp, bj);
          if (!wait_counters(w.ctrl, &chunk_ctr[oi * s.R + r], &chunk_ctr[oj * s.R + r],
                             (unsigned int)g, &s_flag))
            goto aborted;
        }
        if (prof) {
          const long long t1 = now_ticks();
          pt[5] += t1 - t0;
          t0 = t1;
        }
        partial_gram(s, w, bi, bj, r * rc, min(s.len, r * rc + rc),
                     part + (p * RMAX + r) * PW * PW, tile);
        if (prof) {
          const long long t1 = now_ticks();
          pt[0] += t1 - t0;
          t0 = t1;
        }
        signal_add(&pair_ctr[p]);
        if (!wait_counters(w.ctrl, &pair_ctr[p], nullptr, (unsigned int)((g + 1) * s.R),
                           &s_flag))
          goto aborted;
        if (prof) {
          const long long t1 = now_ticks();
          pt[1] += t1 - t0;
          t0 = t1;
        }
        pair_update(s, w, bi, bj, r, part + p * RMAX * PW * PW, off_slot, tol, tiny2,
                    inner_sweeps, cross, G, G2, Q, rot, s_active, tile, s_off, prof, pt, &t0);
        signal_set(&chunk_ctr[p * s.R + r], (unsigned int)(g + 1));
      } else {
        for (int64_t it = blockIdx.x; it < items; it += G_) {
          const int64_t p = it / s.R, r = it % s.R;
          int64_t bi, bj;
          pair_blocks(s.nb, k, p, &bi, &bj);
          partial_gram(s, w, bi, bj, r * rc, min(s.len, r * rc + rc),
                       part + (p * RMAX + r) * PW * PW, tile);
        }
        if (prof) {
          const long long t1 = now_ticks();
          pt[0] += t1 - t0;
          t0 = t1;
        }
        if (!grid_barrier(w.ctrl, (++barrier_no) * G_, &s_flag)) goto aborted;
        if (prof) {
          const long long t1 = now_ticks();
          pt[1] += t1 - t0;
          t0 = t1;
        }
        for (int64_t it = blockIdx.x; it < items; it += G_) {
          const int64_t p = it / s.R, r = it % s.R;
          int64_t bi, bj;
          pair_blocks(s.nb, k, p, &bi, &bj);
          pair_update(s, w, bi, bj, r, part + p * RMAX * PW * PW, off_slot, tol, tiny2,
                      inner_sweeps, cross, G, G2, Q, rot, s_active, tile, s_off, prof, pt, &t0);
        }
        if (!grid_barrier(w.ctrl, (++barrier_no) * G_, &s_flag)) goto aborted;
        if (prof) pt[5] += now_ticks() - t0;
      }
    }
    if (LOCAL) {
      // all pairs' measures of this sweep are in *off_slot after the barrier
      if (!grid_barrier(w.ctrl, (++barrier_no) * G_, &s_flag)) goto aborted;
    }
    {
      const unsigned long long bits =
          __hip_atomic_load(off_slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      const double off_sweep = __longlong_as_double((long long)bits);
      if (blockIdx.x == 0 && tid == 0 && sweep < 64) w.ctrl->offhist[sweep] = off_sweep;
      // Converged at the target, or stalled within 16x of it: near convergence
      // Jacobi reduces the measure quadratically, so a sweep that gains less than
      // 4x there is sitting on the rounding floor.  A sweep whose measure (taken
      // before each pair's rotation) was already below QUAD_STOP leaves an
      // off-diagonal of order QUAD_STOP^2, far below tol: the sweep that would
      // only confirm it is skipped.
      if (off_sweep <= tol || off_sweep <= QUAD_STOP ||
          (off_sweep <= 16.0 * tol && off_sweep > 0.25 * off_prev)) {
        converged = 1;
        ++sweep;
        break;
      }
      off_prev = off_sweep;
    }
  }
  if (prof) {
    w.ctrl->info[0] = sweep;
    w.ctrl->info[1] = converged;
    for (int i = 0; i < 6; ++i) w.ctrl->prof[i] = pt[i];
  }
  return;
aborted:
  if (blockIdx.x == 0 && tid == 0) w.ctrl->info[2] = 1;
}

// sigma_c = ||w_c||.
__global__ void svd_sigma(SvdShape s, SvdWork w) {
  __shared__ double red[256];
  for (int64_t c = blockIdx.x; c < s.ncolp; c += gridDim.x) {
    double acc = 0.0;
    for (int64_t r = threadIdx.x; r < s.len; r += blockDim.x) acc += cabs2(w.W[c * s.len + r]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) w.sigma[c] = sqrt(red[0]);
    __syncthreads();
  }
}

// perm[rank] = c with descending sigma, ties by index.  The keys are staged in
// LDS (dynamic, ncolp doubles) when the launch provides it, else read from
// memory.
__global__ void svd_rank(SvdShape s, SvdWork w, int staged) {
  extern __shared__ double keys[];
  // NaN keys sort last, so the ranks always form a permutation
  auto key = [&](int64_t d) -> double {
    const double x = staged ? keys[d] : w.sigma[d];
    return isnan(x) ? -1.0 : x;
  };
  if (staged) {
    for (int64_t d = threadIdx.x; d < s.ncolp; d += blockDim.x) keys[d] = w.sigma[d];
    __syncthreads();
  }
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < s.ncolp;
       c += (int64_t)gridDim.x * blockDim.x) {
    const double sc = key(c);
    int rank = 0;
    for (int64_t d = 0; d < s.ncolp; ++d) {
      const double sd = key(d);
      rank += (sd > sc || (sd == sc && d < c)) ? 1 : 0;
    }
    w.perm[rank] = (int)c;
  }
}

// Output.  W side: normalised columns of the working matrix (length len);
// V side: V (no QR; stride ncolp, length ncol) or Y = Q V (QR; stride lenA,
// length lenA).  wside_u: the W side is U (columns) rather than VH (conj rows).
__global__ void svd_output(SvdShape s, SvdWork w, double2* __restrict__ U, double* __restrict__ S,
                           double2* __restrict__ VH) {
  const int64_t r = s.ncol;
  const int wside_u = (s.use_qr == s.transposed);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t k = t0; k < r; k += stride) S[k] = w.sigma[w.perm[k]];
  const double tiny2 = w.ctrl->tiny2;
  for (int64_t e = t0; e < r * s.len; e += stride) {
    const int64_t k = e / s.len, x = e % s.len;
    const int64_t c = w.perm[k];
    const double sg = w.sigma[c];
    double2 v = w.W[c * s.len + x];
    // negligible columns get an orthonormal completion afterwards
    v = (sg > 0.0 && sg * sg > tiny2) ? make_double2(v.x / sg, v.y / sg) : make_double2(0.0, 0.0);
    // with QR, coordinate x of the W side is B's column colperm[x]
    const int64_t xo = s.use_qr ? (int64_t)w.colperm[x] : x;
    if (wside_u)
      U[xo * r + k] = v;
    else
      VH[k * s.n + xo] = make_double2(v.x, -v.y);
  }
  const double2* src = s.use_qr ? w.Y : w.V;
  const int64_t ld = s.use_qr ? s.lenA : s.ncolp;
  const int64_t vlen = s.use_qr ? s.lenA : s.ncol;
  for (int64_t e = t0; e < r * vlen; e += stride) {
    const int64_t k = e / vlen, x = e % vlen;
    const double2 v = src[(int64_t)w.perm[k] * ld + x];
    if (wside_u)
      VH[k * s.n + x] = make_double2(v.x, -v.y);
    else
      U[x * r + k] = v;
  }
}

// Complete the W-side singular vectors of negligible singular values
// (sigma^2 <= tiny2: at the rounding level of the factorisation, where the
// directions are not determined) to an orthonormal set.  The singular values
// themselves are kept.  For slot k the candidate is the unit vector e_x whose
// coordinate x has the smallest leverage sum_{j<k} |u_j[x]|^2 among the
// vectors already fixed; its residual after projection then has squared norm
// 1 - leverage >= 1 - k/len > 0, so the completion cannot fail.  Projection is
// classical Gram-Schmidt applied twice, one wavefront per dot product.  Only
// slots k < *kmax are completed when kmax is given (the caller keeps no more).
// One workgroup of 1024 threads; coef holds at least ncol complex values.
__global__ __launch_bounds__(1024) void svd_complete(SvdShape s, SvdWork w, double2* __restrict__ U,
                                                     double2* __restrict__ VH,
                                                     const int64_t* __restrict__ kmax) {
  const int64_t r = s.ncol, len = s.len;
  const int wside_u = (s.use_qr == s.transposed);
  // element x of W-side vector k, as a plain (unconjugated) vector component
  auto load = [&](int64_t k, int64_t x) -> double2 {
    if (wside_u) return U[x * r + k];
    const double2 v = VH[k * s.n + x];
    return make_double2(v.x, -v.y);
  };
  auto store = [&](int64_t k, int64_t x, double2 v) {
    if (wside_u)
      U[x * r + k] = v;
    else
      VH[k * s.n + x] = make_double2(v.x, -v.y);
  };
  __shared__ double s_val[16];
  __shared__ int64_t s_idx[16];
  __shared__ int64_t s_best;
  __shared__ double s_nrm;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwaves = (int)(blockDim.x >> 6);
  double2* coef = w.part;
  const double tiny2 = w.ctrl->tiny2;
  int64_t kend = r;
  if (kmax && *kmax < kend) kend = *kmax;
  // Usually no kept column is negligible: find out in parallel first, and
  // leave without the sequential scan when none is.
  __shared__ int s_any;
  if (tid == 0) s_any = 0;
  __syncthreads();
  for (int64_t k = tid; k < kend; k += blockDim.x) {
    const double sg = w.sigma[w.perm[k]];
    if (!(sg > 0.0 && sg * sg > tiny2)) atomicOr(&s_any, 1);
  }
  __syncthreads();
  if (!s_any) return;
  for (int64_t k = 0; k < kend; ++k) {
    const double sg = w.sigma[w.perm[k]];
    if (sg > 0.0 && sg * sg > tiny2) continue;
    // Candidate coordinate: smallest leverage, ties to the smallest index.
    double best = 1.0e300;
    int64_t bx = len;
    for (int64_t x = tid; x < len; x += blockDim.x) {
      double lev = 0.0;
      for (int64_t j = 0; j < k; ++j) lev += cabs2(load(j, x));
      if (lev < best) {
        best = lev;
        bx = x;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(best, off, 64);
      const int64_t ox = __shfl_down(bx, off, 64);
      if (ov < best || (ov == best && ox < bx)) {
        best = ov;
        bx = ox;
      }
    }
    if (lane == 0) {
      s_val[wave] = best;
      s_idx[wave] = bx;
    }
    __syncthreads();
    if (tid == 0) {
      double bv = s_val[0];
      int64_t bi = s_idx[0];
      for (int q = 1; q < nwaves; ++q)
        if (s_val[q] < bv || (s_val[q] == bv && s_idx[q] < bi)) {
          bv = s_val[q];
          bi = s_idx[q];
        }
      s_best = bi;
    }
    __syncthreads();
    const int64_t xb = s_best;
    for (int64_t x = tid; x < len; x += blockDim.x)
      store(k, x, make_double2(x == xb ? 1.0 : 0.0, 0.0));
    __syncthreads();
    for (int pass = 0; pass < 2; ++pass) {
      for (int64_t j = wave; j < k; j += nwaves) {
        double2 acc = make_double2(0.0, 0.0);
        for (int64_t x = lane; x < len; x += 64) acc = cfmac(load(j, x), load(k, x), acc);
        for (int off = 32; off > 0; off >>= 1) {
          acc.x += __shfl_down(acc.x, off, 64);
          acc.y += __shfl_down(acc.y, off, 64);
        }
        if (lane == 0) coef[j] = acc;
      }
      __syncthreads();
      for (int64_t x = tid; x < len; x += blockDim.x) {
        double2 b = load(k, x);
        for (int64_t j = 0; j < k; ++j) {
          const double2 pa = cmul(coef[j], load(j, x));
          b = make_double2(b.x - pa.x, b.y - pa.y);
        }
        store(k, x, b);
      }
      __syncthreads();
    }
    double acc = 0.0;
    for (int64_t x = tid; x < len; x += blockDim.x) acc += cabs2(load(k, x));
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) s_val[wave] = acc;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < nwaves; ++q) t += s_val[q];
      s_nrm = sqrt(t);
    }
    __syncthreads();
    const double nrm = s_nrm;
    for (int64_t x = tid; x < len; x += blockDim.x) {
      const double2 b = load(k, x);
      store(k, x, make_double2(b.x / nrm, b.y / nrm));
    }
    __syncthreads();
  }
}

// Truncation rule: one lane runs the reference's sequential loops (summation
// order identical to truncation.pxi:251-327, products and sums rounded
// separately as in the reference's compiled loops) over an LDS copy of s
// (STAGED, spectra of up to TRUNC_STAGE values) or over s itself.
constexpr int64_t TRUNC_STAGE = 4096;
template <bool STAGED>
__global__ void truncate_kernel(int64_t N, const double* __restrict__ s, int method, double tol,
                                int64_t max_bond, int64_t* D_out, double* err_out) {
#pragma clang fp contract(off)
  extern __shared__ double stage[];
  if (STAGED) {
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) stage[i] = s[i];
    __syncthreads();
  }
  const double* ss = STAGED ? stage : s;
  if (threadIdx.x != 0) return;
  int64_t D = N;
  double err = 0.0;
  if (method == SEEMPS_TRUNC_NONE) {
    D = N;
    err = 0.0;
  } else if (method == SEEMPS_TRUNC_REL_NORM2) {
    // errors[i] = sum of the i smallest s^2, accumulated from the smallest.
    double total = 0.0;
    for (int64_t i = 0; i < N; ++i) {
      const double x = ss[N - 1 - i];
      total += x * x;
    }
    const double max_error = total * tol;
    double acc = 0.0;
    int64_t i_stop = N - 1;
    for (int64_t i = 0; i < N; ++i) {
      if (acc > max_error) {
        i_stop = i - 1;
        break;
      }
      const double x = ss[N - 1 - i];
      acc += x * x;
    }
    D = N - i_stop;
    if (D > max_bond) D = max_bond;
    double e = 0.0;  // errors[N - D]
    for (int64_t i = 0; i < N - D; ++i) {
      const double x = ss[N - 1 - i];
      e += x * x;
    }
    err = e;
  } else {
    const double thr = (method == SEEMPS_TRUNC_REL_SV) ? tol * ss[0] : tol;
    D = (N < max_bond) ? N : max_bond;
    for (int64_t i = 1; i < D; ++i) {
      if (ss[i] <= thr) {
        D = i;
        break;
      }
    }
    double e = 0.0;
    for (int64_t i = D; i < N; ++i) e += ss[i] * ss[i];
    err = e;
  }
  *D_out = D;
  *err_out = err;
}

int cooperative_grid(int64_t items, int* grid) {
  DeviceInfo di;
  int st = device_info(&di);
  if (st) return st;
  static thread_local int occ_dev = -1, occ = 0;
  if (occ_dev != di.device) {
    int o = 0;
    SEEMPS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &o, (const void*)svd_jacobi<true>, 256, 0));
    occ = o;
    occ_dev = di.device;
  }
  int64_t lim = (int64_t)occ * di.num_cu;
  if (lim < 1) {
    set_error("svd: kernel cannot be resident on this device");
    return SEEMPS_ERR_RUNTIME;
  }
  int64_t g = items < lim ? items : lim;
  if (g > 1024) g = 1024;
  *grid = (int)g;
  return SEEMPS_OK;
}

unsigned blocks_for(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

// Enqueue the SVD.  With kmax_dev == nullptr every slot is completed; otherwise
// svd_complete_enqueue must follow, limited to the slots the caller keeps.
int svd_enqueue(int64_t m, int64_t n, const double2* A, int64_t lda, double2* U, double* S,
                double2* VH, void* work, int64_t work_bytes, hipStream_t stream,
                int* info_dev_off, bool complete, bool precondition) {
  DeviceInfo di;
  int st = device_info(&di);
  if (st) return st;
  const SvdShape s = svd_shape(m, n, di.num_cu, precondition);
  SEEMPS_CHECK_ARG(work_bytes >= svd_bytes(s), "svd: workspace too small (%lld < %lld)",
                   (long long)work_bytes, (long long)svd_bytes(s));
  SvdWork w = svd_carve(s, work);
  if (s.use_qr) {
    hipLaunchKernelGGL(col_norms, dim3(blocks_for(s.ncol)), dim3(256), 0, stream, s, A, lda,
                       w.colnrm);
    if ((st = check_launch("col_norms"))) return st;
    hipLaunchKernelGGL(col_rank, dim3(blocks_for(s.ncol)), dim3(256), 0, stream, s, w.colnrm,
                       w.colperm);
    if ((st = check_launch("col_rank"))) return st;
    hipLaunchKernelGGL(load_colmajor, dim3(blocks_for(s.ncol * s.lenA)), dim3(256), 0, stream, s,
                       A, lda, w.Acm, s.ncol, (const int*)w.colperm);
    if ((st = check_launch("load_colmajor"))) return st;
    if ((st = qr_factor(w.Acm, s.lenA, s.ncol, QRNB, w.Vall, w.Tall, w.tau, w.Z, w.ks,
                        qr_ks_elems(s), stream)))
      return st;
    hipLaunchKernelGGL(load_w_from_r, dim3(blocks_for(s.ncolp * s.len)), dim3(256), 0, stream, s,
                       w.Acm, w.W);
    if ((st = check_launch("load_w_from_r"))) return st;
  } else {
    hipLaunchKernelGGL(load_colmajor, dim3(blocks_for(s.ncolp * s.lenA)), dim3(256), 0, stream,
                       s, A, lda, w.W, s.ncolp, (const int*)nullptr);
    if ((st = check_launch("load_colmajor"))) return st;
  }
  hipLaunchKernelGGL(init_v_ctrl, dim3(blocks_for(s.ncolp * s.ncolp)), dim3(256), 0, stream, s, w);
  if ((st = check_launch("init_v_ctrl"))) return st;
  hipLaunchKernelGGL(svd_frob, dim3(1), dim3(1024), 0, stream, s, w);
  if ((st = check_launch("svd_frob"))) return st;
  int grid = 1;
  if ((st = cooperative_grid(s.P * s.R, &grid))) return st;
  // Relative orthogonality target: a few times above the rounding floor of the
  // measured Gram entries (~sqrt(len) eps), so that rounding alone cannot keep
  // the measure above it.
  const double tol = 32.0 * sqrt((double)s.len) * 2.220446049250313e-16;
  int max_sweeps = MAX_SWEEPS;
  void* args[] = {(void*)&s, (void*)&w, (void*)&tol, (void*)&max_sweeps};
  {
    // nominal SVD flops (complex): 4 * (6 M N^2 + 16 N^3), M >= N
    const double M_ = (double)s.lenA, N_ = (double)s.ncol;
    TimedLaunch timed("svd_jacobi", stream, 4.0 * (6.0 * M_ * N_ * N_ + 16.0 * N_ * N_ * N_));
    // one work item per workgroup when they all fit: local step ordering
    const bool local = (grid == s.P * s.R);
    const void* fn = local ? (const void*)svd_jacobi<true> : (const void*)svd_jacobi<false>;
    SEEMPS_HIP(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), args, 0, stream));
  }
  hipLaunchKernelGGL(svd_sigma, dim3((unsigned)(s.ncolp < 1024 ? s.ncolp : 1024)), dim3(256), 0,
                     stream, s, w);
  if ((st = check_launch("svd_sigma"))) return st;
  {
    const int staged = s.ncolp <= 8192;  // up to 64 KB of keys per workgroup
    const size_t shmem = staged ? (size_t)s.ncolp * sizeof(double) : 0;
    hipLaunchKernelGGL(svd_rank, dim3(blocks_for(s.ncolp)), dim3(256), shmem, stream, s, w,
                       staged);
  }
  if ((st = check_launch("svd_rank"))) return st;
  if (s.use_qr) {
    hipLaunchKernelGGL(load_y, dim3(blocks_for(s.ncol * s.lenA)), dim3(256), 0, stream, s, w);
    if ((st = check_launch("load_y"))) return st;
    if ((st = qr_apply_q(w.Y, s.lenA, s.ncol, s.ncol, QRNB, w.Vall, w.Tall, w.Z, w.ks,
                         qr_ks_elems(s), stream)))
      return st;
  }
  const int64_t big = s.ncol * (s.len > s.lenA ? s.len : s.lenA);
  hipLaunchKernelGGL(svd_output, dim3(blocks_for(big)), dim3(256), 0, stream, s, w, U, S, VH);
  if ((st = check_launch("svd_output"))) return st;
  if (complete) {
    TimedLaunch timed("svd_complete", stream, 0.0);
    hipLaunchKernelGGL(svd_complete, dim3(1), dim3(1024), 0, stream, s, w, U, VH,
                       (const int64_t*)nullptr);
    if ((st = check_launch("svd_complete"))) return st;
  }
  if (info_dev_off) *info_dev_off = (int)((char*)w.ctrl - (char*)work);
  return SEEMPS_OK;
}

// Orthonormal completion of the slots k < *kmax_dev (see svd_complete).
int svd_complete_enqueue(int64_t m, int64_t n, double2* U, double2* VH, void* work,
                         const int64_t* kmax_dev, hipStream_t stream, bool precondition) {
  DeviceInfo di;
  int st = device_info(&di);
  if (st) return st;
  const SvdShape s = svd_shape(m, n, di.num_cu, precondition);
  SvdWork w = svd_carve(s, work);
  TimedLaunch timed("svd_complete", stream, 0.0);
  hipLaunchKernelGGL(svd_complete, dim3(1), dim3(1024), 0, stream, s, w, U, VH, kmax_dev);
  return check_launch("svd_complete");
}

int64_t svd_workspace_bytes(int64_t m, int64_t n, bool precondition) {
  // Sized for the largest row split so the size does not depend on the device.
  SvdShape s = svd_shape(m, n, 1 << 20, precondition);
  return svd_bytes(s);
}

int truncate_enqueue(int64_t N, const double* s, int method, double tol, int64_t max_bond,
                     int64_t* D_out, double* err_out, hipStream_t stream) {
  SEEMPS_CHECK_ARG(N >= 1, "truncate: empty spectrum");
  SEEMPS_CHECK_ARG(method >= 0 && method <= 3, "truncate: invalid method %d", method);
  SEEMPS_CHECK_ARG(max_bond >= 1, "truncate: invalid max bond dimension");
  if (N <= TRUNC_STAGE)
    hipLaunchKernelGGL(truncate_kernel<true>, dim3(1), dim3(256), (size_t)N * sizeof(double),
                       stream, N, s, method, tol, max_bond, D_out, err_out);
  else
    hipLaunchKernelGGL(truncate_kernel<false>, dim3(1), dim3(64), 0, stream, N, s, method, tol,
                       max_bond, D_out, err_out);
  return check_launch("truncate");
}

// Read the control block after the stream has been synchronised by the caller's copy.
int svd_read_ctrl(void* work, int off, int32_t* info, hipStream_t stream) {
  static_assert(sizeof(Ctrl) <= PINNED_BYTES - PINNED_CTRL_OFF, "control block too large");
  char* hp = (char*)pinned_stage(PINNED_BYTES);
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  hp += PINNED_CTRL_OFF;
  SEEMPS_HIP(hipMemcpyAsync(hp, (char*)work + off, sizeof(Ctrl), hipMemcpyDeviceToHost, stream));
  SEEMPS_HIP(hipStreamSynchronize(stream));
  Ctrl h;
  memcpy(&h, hp, sizeof(h));
  if (info) {
    info[0] = h.info[0];
    info[1] = h.info[1];
  }
  if (getenv("SEEMPS_SVD_PROFILE")) {
    int dev = 0, rate = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev);  // kHz
    const double us = rate > 0 ? 1e3 / rate : 0.0;
    fprintf(stderr,
            "[svd] sweeps=%d converged=%d  us: gramA=%.1f bar1=%.1f reduce=%.1f inner=%.1f "
            "rotate=%.1f bar2=%.1f\n",
            h.info[0], h.info[1], h.prof[0] * us, h.prof[1] * us, h.prof[2] * us, h.prof[3] * us,
            h.prof[4] * us, h.prof[5] * us);
    fprintf(stderr, "[svd] off by sweep:");
    for (int i = 0; i < h.info[0] && i < 16; ++i) fprintf(stderr, " %.1e", h.offhist[i]);
    fprintf(stderr, "\n");
  }
  if (h.info[2]) {
    set_error("svd: grid barrier timed out (cooperative residency lost)");
    return SEEMPS_ERR_RUNTIME;
  }
  if (!h.info[1]) {
    const int ns = h.info[0] < 64 ? h.info[0] : 64;
    set_error("SVD did not converge in %d sweeps (off-diagonal measure by sweep: %.2e %.2e %.2e "
              "... %.2e %.2e; negligible column norm^2 %.2e)",
              h.info[0], ns > 0 ? h.offhist[0] : 0.0, ns > 1 ? h.offhist[1] : 0.0,
              ns > 2 ? h.offhist[2] : 0.0, ns > 1 ? h.offhist[ns - 2] : 0.0,
              ns > 0 ? h.offhist[ns - 1] : 0.0, h.tiny2);
    return SEEMPS_ERR_NOCONVERGE;
  }
  return SEEMPS_OK;
}

}  // namespace seemps

extern "C" int64_t seemps_svd_workspace(int64_t m, int64_t n) {
  return seemps::svd_workspace_bytes(m, n);
}

extern "C" int seemps_zgesvd(int64_t m, int64_t n, const void* A, int64_t lda, void* U, double* S,
                             void* VH, void* work, int64_t work_bytes, int32_t* info,
                             void* stream) {
  SEEMPS_CHECK_ARG(m >= 1 && n >= 1, "svd: empty matrix");
  SEEMPS_CHECK_ARG(lda >= n, "svd: lda < n");
  int off = 0;
  int st = seemps::svd_enqueue(m, n, (const double2*)A, lda, (double2*)U, S, (double2*)VH, work,
                               work_bytes, (hipStream_t)stream, &off);
  if (st) return st;
  return seemps::svd_read_ctrl(work, off, info, (hipStream_t)stream);
}

extern "C" int seemps_truncate(int64_t N, const double* s, int method, double tolerance,
                               int64_t max_bond, int64_t* D_out, double* err_out, void* stream) {
  return seemps::truncate_enqueue(N, s, method, tolerance, max_bond, D_out, err_out,
                                  (hipStream_t)stream);
}

namespace seemps {

int64_t svd_split_bytes(int64_t m, int64_t n) {
  const int64_t r = m < n ? m : n;
  auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  return al(m * r * 16) + al(r * 8) + al(r * n * 16) + al(64) + svd_workspace_bytes(m, n);
}

int svd_split_impl(int64_t m, int64_t n, const double2* A, int mode, int method, double tolerance,
                   int64_t max_bond, double2* outA, double2* outB, double* s_out,
                   int64_t* D_host, double* err_host, void* work, int64_t work_bytes,
                   int32_t* info, hipStream_t st_) {
  SEEMPS_CHECK_ARG(m >= 1 && n >= 1, "svd_split: empty matrix");
  SEEMPS_CHECK_ARG(mode == 0 || mode == 1, "svd_split: invalid mode");
  SEEMPS_CHECK_ARG(work_bytes >= svd_split_bytes(m, n), "svd_split: workspace too small");
  const int64_t r = m < n ? m : n;
  auto al = [](int64_t x) { return (x + 255) & ~int64_t(255); };
  char* p = (char*)work;
  double2* U = (double2*)p;
  p += al(m * r * 16);
  double* S = (double*)p;
  p += al(r * 8);
  double2* VH = (double2*)p;
  p += al(r * n * 16);
  int64_t* Ddev = (int64_t*)p;
  double* edev = (double*)(p + 8);
  p += al(64);
  int off = 0;
  int st = svd_enqueue(m, n, A, n, U, S, VH, p, work_bytes - (p - (char*)work), st_, &off, false);
  if (st) return st;
  if ((st = truncate_enqueue(r, S, method, tolerance, max_bond, Ddev, edev, st_))) return st;
  if ((st = svd_complete_enqueue(m, n, U, VH, p, Ddev, st_))) return st;
  char* hp = (char*)pinned_stage(PINNED_BYTES);
  SEEMPS_CHECK_ARG(hp != nullptr, "pinned host staging could not be allocated");
  SEEMPS_HIP(hipMemcpyAsync(hp, Ddev, 8, hipMemcpyDeviceToHost, st_));
  SEEMPS_HIP(hipMemcpyAsync(hp + 8, edev, 8, hipMemcpyDeviceToHost, st_));
  if ((st = svd_read_ctrl(p, off, info, st_))) return st;  // synchronises
  int64_t hD = 0;
  double herr = 0.0;
  memcpy(&hD, hp, 8);
  memcpy(&herr, hp + 8, 8);
  const int64_t D = hD;
  if (mode == 0) {
    if ((st = zscale2d_impl(m, D, nullptr, nullptr, U, r, outA, D, st_))) return st;
    if ((st = zscale2d_impl(D, n, S, nullptr, VH, n, outB, n, st_))) return st;
  } else {
    if ((st = zscale2d_impl(m, D, nullptr, S, U, r, outA, D, st_))) return st;
    if ((st = zscale2d_impl(D, n, nullptr, nullptr, VH, n, outB, n, st_))) return st;
  }
  if (s_out) SEEMPS_HIP(hipMemcpyAsync(s_out, S, r * 8, hipMemcpyDeviceToDevice, st_));
  *D_host = D;
  *err_host = herr;
  return SEEMPS_OK;
}

}  // namespace seemps

extern "C" int64_t seemps_svd_split_workspace(int64_t m, int64_t n) {
  return seemps::svd_split_bytes(m, n);
}

extern "C" int seemps_svd_split(int64_t m, int64_t n, const void* A, int mode, int method,
                                double tolerance, int64_t max_bond, void* outA, void* outB,
                                double* s_out, int64_t* D_host, double* err_host, void* work,
                                int64_t work_bytes, int32_t* info, void* stream) {
  return seemps::svd_split_impl(m, n, (const double2*)A, mode, method, tolerance, max_bond,
                                (double2*)outA, (double2*)outB, s_out, D_host, err_host, work,
                                work_bytes, info, (hipStream_t)stream);
}
