// SVD of a matrix whose columns are already nearly orthogonal, by
// simultaneous Jacobi rotations.  Every iteration forms the Gram matrix
// G = W^H W, derives for every column pair (i, j) the plane rotation that
// would annihilate G_ij (the one-sided Jacobi rotation of that pair: tangent
// t and phase ph), assembles all of them into one matrix Z = I + X (X
// anti-Hermitian, X_ij = t ph, X_ji = -t conj(ph)), replaces Z by its unitary
// polar factor (Newton-Schulz iterations Z <- Z (3 I - Z^H Z) / 2 after a
// scaling that puts its singular values in (0, 1]), and rotates
// [W; V] <- [W; V] Z.  Near convergence the step is Jacobi's: the normalised
// off-diagonal of G falls quadratically from one iteration to the next (a
// captured chi = 256 DMRG split goes 2e-2, 4e-3, 5e-5, 1e-6, 5e-10), so a
// handful of iterations of a few GEMMs each replace the long chain of small
// dependent block steps of the block Jacobi.
//
// The whole iteration runs in one cooperative kernel: the products are 16 x 16
// output tiles on FP64 MFMA spread over every workgroup, separated by grid
// barriers, and the convergence measure is read on the device.  The Gram
// product and the generator share one phase: each Gram tile is turned into
// its tile of Z by the workgroup that computed it.
//
// Accuracy: each G_ij is a dot product of two columns, rounded relative to the
// two column norms, so the rotations are as accurate as one-sided Jacobi's;
// the polar factor is unitary to a few ulp (the Newton-Schulz count follows
// from a bound on ||X||_2).
//
// Used by the basis-sketched split at a capped bond (sketch.hip), whose
// W = op(A) basis has nearly orthogonal columns; a matrix that does not
// converge within POLAR_MAX_ITERS iterations is reported as such and the
// caller uses the block Jacobi instead.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <vector>

#include "common.h"
#include "jacobi_inner.h"

namespace seemps {

namespace {

// SEEMPS_POLAR_TRACE: 1 prints a summary per call, 2 also the barrier
// timings (diagnostics; both synchronise)
int trace_level() {
  static const int lvl = [] {
    const char* e = getenv("SEEMPS_POLAR_TRACE");
    return e ? atoi(e) : 0;
  }();
  return lvl;
}
long long* g_trace_buf = nullptr;

constexpr int POLAR_MAX_ITERS = 12;
// Newton-Schulz steps beyond this (a rotation far from the identity: a matrix
// whose columns are not nearly orthogonal) end the iteration as not
// converged, and the caller takes the block Jacobi.
constexpr int POLAR_MAX_NS = 10;
// An iteration that starts below this normalised off-diagonal leaves ~1e-18
// (quadratic convergence), so it is the last one.
constexpr double POLAR_QUAD_STOP = 1e-9;

inline int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct PolarWork {
  double2* XA;     // (Lg + l) x l: rows < Lg the working matrix, then V
  double2* XB;     // ping-pong buffer of XA
  unsigned long long* rowsum;  // 2 x l: per iteration parity, sums of |X| per row (fixed point)
  double2* Z;      // l x l
  double2* Z2;     // l x l
  double2* T;      // l x l
  double* stats;   // [0] max normalised off-diagonal, [1] bound on ||X||_2 (bit patterns),
                   // [2] ||W||_F^2, [3] negligible squared column norm
  double* nrm2;    // l squared column norms
  int* perm;       // l: columns by decreasing norm
  int64_t* tiny;   // first output slot with a negligible column (l if none)
  void* ctrl;      // PolarCtrl of the cooperative iteration
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
  w.rowsum = (unsigned long long*)take(2 * l * 8);
  w.Z = (double2*)take(l * l * 16);
  w.Z2 = (double2*)take(l * l * 16);
  w.T = (double2*)take(l * l * 16);
  w.stats = (double*)take(8 * 8);
  w.nrm2 = (double*)take(l * 8);
  w.perm = (int*)take(l * 4);
  w.tiny = (int64_t*)take(8);
  w.ctrl = take(256);
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

// ---------------------------------------------------------------------------
// The whole iteration in one cooperative kernel: the GEMMs are spread over
// all workgroups as 16 x 16 output tiles (FP64 MFMA), separated by grid
// barriers, and the convergence test is read on the device, so no launch or
// host round trip sits between the small products.

typedef double pd4 __attribute__((ext_vector_type(4)));

struct PolarCtrl {
  unsigned int barrier;   // grid barrier arrivals
  unsigned int abort;     // a barrier wait timed out
  unsigned int released;  // last barrier phase whose arrivals are complete
  int iterations;
  int converged;
  unsigned long long stats[2][2];  // per iteration parity: [0] max off (bits), [1] unused
  // workgroup 0: [0] ticks waiting in barriers, [1] ticks in all, [2] barriers,
  // [3] Newton-Schulz steps, ticks in [4] Gram and generator, [6] Newton-Schulz,
  // [7] rotation phases (each with its closing barrier)
  long long prof[8];
};

constexpr unsigned long long POLAR_SPIN_LIMIT = 4000000ull;

// Bounded grid barrier (all workgroups resident: cooperative launch).  The
// last workgroup to arrive publishes the phase in a separate word, which the
// others poll, so the waiters do not contend with the arrivals.  Returns false
// once any workgroup has given up waiting.
// `trace` (diagnostics, normally null): arrival time of every workgroup at
// each of the first POLAR_TRACE_BARRIERS barriers, then workgroup 0's
// departure times.
constexpr int POLAR_TRACE_BARRIERS = 64;
__device__ bool polar_barrier(PolarCtrl* c, unsigned int* phase, int* s_ok,
                              long long* trace) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int ph = ++*phase;
    const unsigned int target = ph * gridDim.x;
    const long long t0 = (blockIdx.x == 0) ? (long long)wall_clock64() : 0;
    if (trace && ph <= POLAR_TRACE_BARRIERS)
      trace[(ph - 1) * gridDim.x + blockIdx.x] = (long long)wall_clock64();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int before = atomicAdd(&c->barrier, 1u);
    int ok = 1;
    if (before + 1 == target) {
      __hip_atomic_store(&c->released, ph, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long spins = 0;
      while (__hip_atomic_load(&c->released, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ph) {
        if (__hip_atomic_load(&c->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          ok = 0;
          break;
        }
        if (++spins > POLAR_SPIN_LIMIT) {
          atomicExch(&c->abort, 1u);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (blockIdx.x == 0) {
      const long long t1 = (long long)wall_clock64();
      c->prof[0] += t1 - t0;
      c->prof[2] += 1;
      if (trace && ph <= POLAR_TRACE_BARRIERS)
        trace[POLAR_TRACE_BARRIERS * gridDim.x + ph - 1] = t1;
    }
    *s_ok = ok;
  } else {
    ++*phase;
  }
  __syncthreads();
  return *s_ok != 0;
}

// Cout = alpha op(A) B + beta Cin (row-major; op(A) = A (M x K, stride lda) or
// A^H with A stored K x M; B K x N; Cin may be null when beta = 0), by 16 x 16
// output tiles on FP64 MFMA.  The 8 wavefronts of a workgroup form
// 8 / WPT groups of WPT wavefronts; each group takes one tile at a time
// (tiles handed out over the grid), its wavefronts a contiguous share of K
// each, and the partial tiles are summed in LDS in a fixed order.  The
// caller picks WPT per product (polar_wpt): a product with few tiles splits K
// finely, one with many tiles runs several tiles per workgroup at once, so
// that every product takes about one round of tiles with all of a wavefront's
// loads in flight together.  Each wavefront issues the loads of up to 8
// K-steps before their MFMAs.  MFMA f64 16x16x4: lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator item r is
// D[l / 16 + 4 r][l % 16].
//
// HERM_C: the product is Hermitian (A^H A with B = A, beta = 0): only the
// tiles on and above the diagonal (tile row <= tile column) are computed, and
// each is also stored, as its adjoint, below the diagonal.  HERM_B: B is a
// matrix with only those upper tiles stored, so its element (k, j) below the
// diagonal tiles is read as conj(B[j][k]).
constexpr int PG_UNROLL = 8;
constexpr int POLAR_THREADS = 512;
constexpr int POLAR_WAVES = POLAR_THREADS / 64;

// Tile (row, column) of the t-th tile on or above the diagonal of a T x T
// tile grid, numbered row by row.
__device__ __forceinline__ void upper_tile(int64_t t, int64_t T, int64_t* tr, int64_t* tc) {
  int64_t r = 0;
  while (t >= T - r) {
    t -= T - r;
    ++r;
  }
  *tr = r;
  *tc = r + t;
}

// Wavefronts per tile (8, 4 or 2) for a product of `ntiles` tiles and
// `ksteps` K-steps of 4 on `grid` workgroups.  The cost of a choice is its
// rounds of tiles times the time of one round: one memory latency (operands
// written by other workgroups before the barrier come from HBM, a few us) per
// batch of PG_UNROLL K-steps, counted as 12 K-steps of MFMA work, plus the
// K-steps themselves.  Ties go to the finer K split.
__device__ __forceinline__ int polar_wpt(int ntiles, int ksteps, int grid) {
  int best = POLAR_WAVES;
  int best_cost = -1;
  for (int wpt = POLAR_WAVES; wpt >= 2; wpt >>= 1) {
    const int groups = POLAR_WAVES / wpt;
    const int rounds = (ntiles + grid * groups - 1) / (grid * groups);
    const int per = (ksteps + wpt - 1) / wpt;
    const int cost = rounds * (12 * ((per + PG_UNROLL - 1) / PG_UNROLL) + per);
    if (best_cost < 0 || cost < best_cost) {
      best = wpt;
      best_cost = cost;
    }
  }
  return best;
}

template <bool ADJ, bool HERM_C = false, bool HERM_B = false>
__device__ void polar_gemm(int64_t M64, int64_t N64, int64_t K64, const double2* __restrict__ A,
                           int64_t lda64, const double2* __restrict__ B, int64_t ldb64,
                           double alpha, double beta, const double2* __restrict__ Cin,
                           double2* __restrict__ C, int64_t ldc64, double* red) {
  // every operand here has fewer than 2^31 elements (checked by the host):
  // 32-bit offsets from uniform bases keep the address registers few
  const int M = (int)M64, N = (int)N64, K = (int)K64;
  const int lda = (int)lda64, ldb = (int)ldb64, ldc = (int)ldc64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tm = (M + 15) >> 4, tn = (N + 15) >> 4;
  const int li = lane & 15, lk = lane >> 4;
  const int ntiles = HERM_C ? tn * (tn + 1) / 2 : tm * tn;
  const int ksteps = (K + 3) >> 2;
  const int wpt = polar_wpt(ntiles, ksteps, gridDim.x);
  const int groups = POLAR_WAVES / wpt;
  const int grp = wave / wpt, sub = wave % wpt;
  // K range of this wavefront, in whole K-steps of 4
  const int per = (ksteps + wpt - 1) / wpt;
  const int s0 = sub * per, s1 = (s0 + per < ksteps) ? s0 + per : ksteps;
  // partial tiles: slot (grp, sub - 1) of 64 lanes x 8 doubles
  double* part = red + (grp * (wpt - 1)) * 64 * 8;
  for (int base = (int)blockIdx.x * groups; base < ntiles; base += (int)gridDim.x * groups) {
    const int t = base + grp;
    const bool active = t < ntiles;
    int m0 = 0, n0 = 0;
    if (active) {
      if (HERM_C) {
        int64_t tr, tc;
        upper_tile(t, tn, &tr, &tc);
        m0 = (int)tr * 16;
        n0 = (int)tc * 16;
      } else {
        m0 = (t / tn) * 16;
        n0 = (t % tn) * 16;
      }
    }
    const int ia = m0 + li, jb = n0 + li;
    pd4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    for (int sb = s0; active && sb < s1; sb += PG_UNROLL) {
      double2 av[PG_UNROLL], bv[PG_UNROLL];
#pragma unroll
      for (int u = 0; u < PG_UNROLL; ++u) {
        const int k = (sb + u) * 4 + lk;
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
        if (sb + u < s1 && k < K) {
          if (ia < M) a = ADJ ? A[k * lda + ia] : A[ia * lda + k];
          if (jb < N) {
            if (HERM_B && (k >> 4) > (jb >> 4)) {
              const double2 bt = B[jb * ldb + k];
              b = make_double2(bt.x, -bt.y);
            } else {
              b = B[k * ldb + jb];
            }
          }
        }
        av[u] = a;
        bv[u] = b;
      }
#pragma unroll
      for (int u = 0; u < PG_UNROLL; ++u) {
        const double ar = av[u].x, ai = ADJ ? -av[u].y : av[u].y;
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bv[u].x, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bv[u].y, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bv[u].y, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bv[u].x, ci, 0, 0, 0);
      }
    }
    // partial tiles of wavefronts 1.. of each group to LDS; wavefront 0 of the
    // group sums them in order and stores
    if (active && sub > 0) {
      double* dst = part + ((sub - 1) * 64 + lane) * 8;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dst[r] = cr[r];
        dst[4 + r] = ci[r];
      }
    }
    __syncthreads();
    if (active && sub == 0) {
      for (int q = 0; q < wpt - 1; ++q) {
        const double* src = part + (q * 64 + lane) * 8;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cr[r] += src[r];
          ci[r] += src[4 + r];
        }
      }
      const int col = n0 + li;
      if (col < N) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + lk + 4 * r;
          if (row < M) {
            double2 v = make_double2(alpha * cr[r], alpha * ci[r]);
            if (beta != 0.0) {
              const double2 c0 = Cin[row * ldc + col];
              v.x += beta * c0.x;
              v.y += beta * c0.y;
            }
            C[row * ldc + col] = v;
            // HERM_C: the tile below the diagonal too, as the adjoint (the
            // diagonal tiles hold both halves already)
            if (HERM_C && m0 != n0) C[col * ldc + row] = make_double2(v.x, -v.y);
          }
        }
      }
    }
    __syncthreads();
  }
}

// Fixed-point scale of the per-row sums of |X| (|t| <= 1, so a row sums to at
// most l < 2^12): integer atomics add them in any order to the same value.
constexpr double POLAR_ROWSUM_SCALE = 1125899906842624.0;  // 2^50

// The Gram tiles of W (Lg x l) on and above the diagonal, as polar_gemm
// computes them, fused with the rotation generator: every workgroup also
// forms the squared norms of the 32 columns of its tile from the operands it
// loads (the same split of K and summation order in every tile, so a column
// norm is the same number wherever it is used), and turns its tile of G into
// the tile of Z = I + X and its mirror below the diagonal.  For a pair (a, b),
// a < b, the one-sided Jacobi rotation of the pair has tangent t and phase ph
// (from G_aa, G_bb, G_ab); X_ab = t ph, X_ba = -t conj(ph).  Pairs below the
// threshold (or with a negligible column) get 0.  Also: st[0] = max over pairs
// of |G_ab| / sqrt(G_aa G_bb) (atomic max of the bit pattern), rowsum[a] +=
// sum over the row of |t| in fixed point (integer atomics).
__device__ void polar_gram_generator(int Lg, int l, const double2* __restrict__ W,
                                     double2* __restrict__ Z, double tiny2, double tol,
                                     unsigned long long* st, unsigned long long* rowsum,
                                     double* red) {
  constexpr int SLOT = 10;  // partial tile (8) and the two column norms
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tn = (l + 15) >> 4;
  const int li = lane & 15, lk = lane >> 4;
  const int ntiles = tn * (tn + 1) / 2;
  const int ksteps = (Lg + 3) >> 2;
  const int wpt = polar_wpt(ntiles, ksteps, gridDim.x);
  const int groups = POLAR_WAVES / wpt;
  const int grp = wave / wpt, sub = wave % wpt;
  const int per = (ksteps + wpt - 1) / wpt;
  const int s0 = sub * per, s1 = (s0 + per < ksteps) ? s0 + per : ksteps;
  double* part = red + (grp * (wpt - 1)) * 64 * SLOT;
  for (int base = (int)blockIdx.x * groups; base < ntiles; base += (int)gridDim.x * groups) {
    const int t = base + grp;
    const bool active = t < ntiles;
    int m0 = 0, n0 = 0;
    if (active) {
      int64_t tr, tc;
      upper_tile(t, tn, &tr, &tc);
      m0 = (int)tr * 16;
      n0 = (int)tc * 16;
    }
    const int ia = m0 + li, jb = n0 + li;
    pd4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    double na = 0.0, nb = 0.0;
    for (int sb = s0; active && sb < s1; sb += PG_UNROLL) {
      double2 av[PG_UNROLL], bv[PG_UNROLL];
#pragma unroll
      for (int u = 0; u < PG_UNROLL; ++u) {
        const int k = (sb + u) * 4 + lk;
        double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
        if (sb + u < s1 && k < Lg) {
          if (ia < l) a = W[k * l + ia];
          if (jb < l) b = W[k * l + jb];
        }
        av[u] = a;
        bv[u] = b;
      }
#pragma unroll
      for (int u = 0; u < PG_UNROLL; ++u) {
        const double ar = av[u].x, ai = -av[u].y;
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bv[u].x, cr, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bv[u].y, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bv[u].y, ci, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bv[u].x, ci, 0, 0, 0);
        na = fma(av[u].x, av[u].x, fma(av[u].y, av[u].y, na));
        nb = fma(bv[u].x, bv[u].x, fma(bv[u].y, bv[u].y, nb));
      }
    }
    // column norms: the four lanes of a column (lk = 0..3), summed pairwise
    // (the same value lands in all four)
    na += __shfl_xor(na, 16, 64);
    na += __shfl_xor(na, 32, 64);
    nb += __shfl_xor(nb, 16, 64);
    nb += __shfl_xor(nb, 32, 64);
    if (active && sub > 0) {
      double* dst = part + ((sub - 1) * 64 + lane) * SLOT;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        dst[r] = cr[r];
        dst[4 + r] = ci[r];
      }
      dst[8] = na;
      dst[9] = nb;
    }
    __syncthreads();
    if (active && sub == 0) {
      for (int q = 0; q < wpt - 1; ++q) {
        const double* src = part + (q * 64 + lane) * SLOT;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cr[r] += src[r];
          ci[r] += src[4 + r];
        }
        na += src[8];
        nb += src[9];
      }
      // tile entry r of this lane: G_ab, a = m0 + lk + 4 r, b = n0 + li;
      // G_aa from the lane that loaded column a (li = lk + 4 r, lk = 0)
      const int b = jb;
      const double gbb = nb;
      double mx = 0.0, sum_b = 0.0;
      double sum_a[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int a = m0 + lk + 4 * r;
        const double gaa = __shfl(na, lk + 4 * r, 64);
        sum_a[r] = 0.0;
        if (a >= l || b >= l || a > b) continue;
        if (a == b) {
          Z[a * l + a] = make_double2(1.0, 0.0);
          continue;
        }
        double2 z = make_double2(0.0, 0.0);
        const double2 c = make_double2(cr[r], ci[r]);
        const double s2 = cabs2(c);
        if (gaa > tiny2 && gbb > tiny2) mx = fmax(mx, sqrt(s2 / gaa / gbb));
        if (s2 > 0.0 && gaa > tiny2 && gbb > tiny2 && s2 > (tol * tol * gaa) * gbb) {
          const double rr = jac::rsqrt_nr(s2);
          const double2 ph = make_double2(c.x * rr, c.y * rr);
          const double tau = (gbb - gaa) * (0.5 * rr);
          double tt;
          if (fabs(tau) > 1e150) {
            tt = 0.5 * jac::rcp_nr(tau);
          } else {
            const double q = fma(tau, tau, 1.0);
            tt = (tau >= 0.0 ? 1.0 : -1.0) * jac::rcp_nr(fabs(tau) + q * jac::rsqrt_nr(q));
          }
          z = make_double2(tt * ph.x, tt * ph.y);
          sum_a[r] = fabs(tt);
          sum_b += fabs(tt);
        }
        Z[a * l + b] = z;
        Z[b * l + a] = make_double2(-z.x, z.y);
      }
      // row b: the four lanes of column b (lk = 0..3); rows a: the sixteen
      // lanes of each (lk, r)
      sum_b += __shfl_xor(sum_b, 16, 64);
      sum_b += __shfl_xor(sum_b, 32, 64);
      if (lk == 0 && b < l && sum_b > 0.0)
        atomicAdd(&rowsum[b], (unsigned long long)ceil(sum_b * POLAR_ROWSUM_SCALE));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v = sum_a[r];
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        const int a = m0 + lk + 4 * r;
        if (li == 0 && a < l && v > 0.0)
          atomicAdd(&rowsum[a], (unsigned long long)ceil(v * POLAR_ROWSUM_SCALE));
      }
      for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
      if (lane == 0) atomicMax(&st[0], (unsigned long long)__double_as_longlong(mx));
    }
    __syncthreads();
  }
}

// Newton-Schulz steps that take a smallest singular value s (the others lie
// in [s, 1]) to within 2.3e-16 of 1: the step maps [0, 1] into itself with
// fixed point 1, so the whole spectrum is then unitary to a few ulp.  Zero
// when s already is (then Z = I + X is used as it is: Z^H Z = I - X^2 differs
// from I by at most |X|^2 <= 4.6e-16).
__device__ int ns_steps_device(double s) {
  int k = 0;
  while (1.0 - s > 2.3e-16 && k < 60) {
    const double next = s * (3.0 - s * s) * 0.5;
    ++k;
    if (next == s) break;
    s = next;
  }
  return k;
}

// The rotation iteration of the file comment, all of it on the device.
// XA holds [W; I] on entry and [W V; V] on exit (copied back from the
// ping-pong buffer if needed); stats[3] is the negligible squared column norm.
__global__ __launch_bounds__(POLAR_THREADS) void polar_persistent(int64_t Lg, int64_t l, double2* XA,
                                                        double2* XB,
                                                        unsigned long long* rowsum, double2* Z,
                                                        double2* Z2, double2* T,
                                                        const double* __restrict__ stats,
                                                        PolarCtrl* ctrl, double tol,
                                                        int max_iters, long long* trace) {
  __shared__ unsigned long long red_max[POLAR_WAVES];
  __shared__ double red[(POLAR_WAVES - 1) * 64 * 10];
  __shared__ int s_ok;
  unsigned int phase = 0;
  const long long t_start = (long long)wall_clock64();
  const int64_t H = Lg + l;
  const double tiny2 = stats[3];
  double2* cur = XA;
  double2* nxt = XB;
  int it = 0, converged = 0;
  long long t_prev = t_start;
  // workgroup 0 books the time since the previous stamp to prof[slot]
  auto stamp = [&](int slot) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const long long now = (long long)wall_clock64();
      ctrl->prof[slot] += now - t_prev;
      t_prev = now;
    }
  };
  for (it = 0; it < max_iters; ++it) {
    // Gram tiles fused with the generator: Z = I + X, the largest normalised
    // off-diagonal in st[0] and the row sums of |X| in rowsum[it & 1] (both
    // cleared one iteration ago, or before the launch)
    unsigned long long* st = ctrl->stats[it & 1];
    unsigned long long* rs = rowsum + (it & 1) * l;
    polar_gram_generator((int)Lg, (int)l, cur, Z, tiny2, tol, st, rs, red);
    if (blockIdx.x == 0) {
      unsigned long long* rs_next = rowsum + ((it + 1) & 1) * l;
      for (int64_t a = threadIdx.x; a < l; a += blockDim.x) rs_next[a] = 0ull;
      if (threadIdx.x == 0) atomicExch(&ctrl->stats[(it + 1) & 1][0], 0ull);
    }
    if (!polar_barrier(ctrl, &phase, &s_ok, trace)) break;
    stamp(4);
    const double off = __longlong_as_double(
        (long long)__hip_atomic_load(&st[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    // ||X||_2 <= max row sum of |X| (X anti-Hermitian: the 1- and inf-norms agree)
    unsigned long long mrow = 0ull;
    for (int64_t a = threadIdx.x; a < l; a += blockDim.x) mrow = rs[a] > mrow ? rs[a] : mrow;
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long v = __shfl_xor(mrow, o, 64);
      mrow = v > mrow ? v : mrow;
    }
    if ((threadIdx.x & 63) == 0) red_max[threadIdx.x >> 6] = mrow;
    __syncthreads();
    unsigned long long mall = 0ull;
    for (int q = 0; q < POLAR_WAVES; ++q) mall = red_max[q] > mall ? red_max[q] : mall;
    __syncthreads();
    const double xn = (double)mall / POLAR_ROWSUM_SCALE;
    if (!(off == off)) break;  // NaN: not converged
    if (off < tol || xn == 0.0) {
      converged = 1;
      break;
    }
    // polar factor of f Z, f scaling its singular values into (0, 1]; the
    // first Newton-Schulz step applies f through its coefficients:
    // T = f^2 Z^H Z, Z2 = 1.5 f Z - 0.5 f Z T
    const double f = 1.0 / sqrt(1.0 + xn * xn);
    const int ns = ns_steps_device(f);
    if (ns > POLAR_MAX_NS) break;  // not converged
    bool ok = true;
    for (int k = 0; k < ns && ok; ++k) {
      const double g = (k == 0) ? f : 1.0;
      polar_gemm<true, true>(l, l, l, Z, l, Z, l, g * g, 0.0, nullptr, T, l, red);
      if (!(ok = polar_barrier(ctrl, &phase, &s_ok, trace))) break;
      polar_gemm<false>(l, l, l, Z, l, T, l, -0.5 * g, 1.5 * g, Z, Z2, l, red);
      if (!(ok = polar_barrier(ctrl, &phase, &s_ok, trace))) break;
      double2* tmp = Z;
      Z = Z2;
      Z2 = tmp;
    }
    if (!ok) break;
    stamp(6);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctrl->prof[3] += ns;
    polar_gemm<false>(H, l, l, cur, l, Z, l, 1.0, 0.0, nullptr, nxt, l, red);  // [W; V] Z
    if (!polar_barrier(ctrl, &phase, &s_ok, trace)) break;
    stamp(7);
    double2* tmp = cur;
    cur = nxt;
    nxt = tmp;
    if (off < POLAR_QUAD_STOP) {
      converged = 1;
      ++it;
      break;
    }
  }
  if (cur != XA) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < H * l;
         e += (int64_t)gridDim.x * blockDim.x)
      XA[e] = cur[e];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctrl->prof[1] = (long long)wall_clock64() - t_start;
    ctrl->iterations = it;
    ctrl->converged = converged && !__hip_atomic_load(&ctrl->abort, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Control block and both parities of the row sums to zero.
__global__ void polar_ctrl_reset(PolarCtrl* c, int64_t l, unsigned long long* rowsum) {
  for (int64_t a = threadIdx.x; a < 2 * l; a += blockDim.x) rowsum[a] = 0ull;
  if (threadIdx.x == 0) {
    c->barrier = 0;
    c->abort = 0;
    c->released = 0;
    c->iterations = 0;
    c->converged = 0;
    c->stats[0][0] = c->stats[0][1] = c->stats[1][0] = c->stats[1][1] = 0ull;
    for (int q = 0; q < 8; ++q) c->prof[q] = 0;
  }
}

unsigned polar_grid(int64_t n) {
  int64_t b = ceil_div(n, 256);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

int64_t polar_svd_bytes(int64_t Lg, int64_t l) { return polar_carve(Lg, l, nullptr).bytes; }

// SVD W = U diag(S) VH of the Lg x l matrix W (row stride ldw, Lg >= l, columns
// nearly orthogonal): U Lg x l (row stride l), S descending, VH l x l.  All
// work is enqueued on the stream, nothing is synchronised.  Device results the
// caller reads at its next synchronisation: *conv_dev (int) = 1 when the
// iteration converged (otherwise U, S, VH must not be used), *tiny_dev
// (int64) = the first slot whose singular value is negligible
// (<= Lg eps ||W||_F), l if none: such slots have U columns of zero, which the
// caller must not keep.  Returns SEEMPS_ERR_NOCONVERGE without enqueuing
// anything when the cooperative grid cannot be resident.
int polar_svd(int64_t Lg, int64_t l, const double2* W, int64_t ldw, double2* U, double* S,
              double2* VH, const int64_t** tiny_dev, const int** conv_dev, void* work,
              int64_t work_bytes, hipStream_t st_) {
  SEEMPS_CHECK_ARG(Lg >= l && l >= 1, "polar_svd: needs Lg >= l >= 1");
  SEEMPS_CHECK_ARG((Lg + l) * l < (int64_t)INT32_MAX, "polar_svd: matrix too large");
  PolarWork w = polar_carve(Lg, l, work);
  SEEMPS_CHECK_ARG(work_bytes >= w.bytes, "polar_svd: workspace too small");
  static_assert(sizeof(PolarCtrl) <= 256, "control block too large");
  const int64_t H = Lg + l;
  const double tol = 32.0 * sqrt((double)Lg) * 2.220446049250313e-16;
  DeviceInfo di;
  int st = device_info(&di);
  if (st) return st;
  static thread_local int occ_dev = -1, occ = 0;
  if (occ_dev != di.device) {
    int o = 0;
    SEEMPS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, (const void*)polar_persistent,
                                                            POLAR_THREADS, 0));
    occ = o;
    occ_dev = di.device;
  }
  // one workgroup of 8 wavefronts per CU, no more than the widest product's
  // ceil(H/16) ceil(l/16) tiles
  int64_t grid = ceil_div(H, 16) * ceil_div(l, 16);
  if (grid > di.num_cu) grid = di.num_cu;
  if (grid > (int64_t)occ * di.num_cu) grid = (int64_t)occ * di.num_cu;
  if (grid < 1) {
    set_error("polar_svd: cooperative kernel cannot be resident");
    return SEEMPS_ERR_NOCONVERGE;
  }
  hipLaunchKernelGGL(polar_init, dim3(polar_grid(H * l)), dim3(256), 0, st_, Lg, l, W, ldw,
                     w.XA);
  if ((st = check_launch("polar_init"))) return st;
  hipLaunchKernelGGL(polar_frob, dim3(1), dim3(1024), 0, st_, Lg, l, w.XA, w.stats);
  if ((st = check_launch("polar_frob"))) return st;
  PolarCtrl* ctrl = (PolarCtrl*)w.ctrl;
  hipLaunchKernelGGL(polar_ctrl_reset, dim3(1), dim3(256), 0, st_, ctrl, l, w.rowsum);
  if ((st = check_launch("polar_ctrl_reset"))) return st;
  {
    int64_t Lg_ = Lg, l_ = l;
    double tol_ = tol;
    int iters = POLAR_MAX_ITERS;
    const double* stats = w.stats;
    long long* trace_buf = nullptr;
    if (trace_level() >= 2) {  // diagnostics: barrier arrival times (leaked per call)
      const size_t tb = (POLAR_TRACE_BARRIERS * grid + POLAR_TRACE_BARRIERS) * 8;
      SEEMPS_HIP(hipMalloc(&trace_buf, tb));
      SEEMPS_HIP(hipMemsetAsync(trace_buf, 0, tb, st_));
      g_trace_buf = trace_buf;
    }
    void* args[] = {(void*)&Lg_, (void*)&l_, (void*)&w.XA, (void*)&w.XB, (void*)&w.rowsum,
                    (void*)&w.Z, (void*)&w.Z2, (void*)&w.T, (void*)&stats, (void*)&ctrl,
                    (void*)&tol_, (void*)&iters, (void*)&trace_buf};
    // nominal SVD flops of the matrix (as for svd_jacobi): 4 (6 Lg l^2 + 16 l^3)
    const double M_ = (double)Lg, N_ = (double)l;
    TimedLaunch timed("polar_svd", st_, 4.0 * (6.0 * M_ * N_ * N_ + 16.0 * N_ * N_ * N_));
    SEEMPS_HIP(hipLaunchCooperativeKernel((const void*)polar_persistent, dim3((unsigned)grid),
                                          dim3(POLAR_THREADS), args, 0, st_));
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
  if (trace_level() >= 1) {  // diagnostics: synchronises
    PolarCtrl h;
    SEEMPS_HIP(hipMemcpyAsync(&h, ctrl, sizeof(h), hipMemcpyDeviceToHost, st_));
    SEEMPS_HIP(hipStreamSynchronize(st_));
    fprintf(stderr, "[polar_svd %lld x %lld] %d iterations, converged %d, %lld barriers, "
            "%.1f us of %.1f us waiting in barriers (workgroup 0), grid %lld (%d per CU "
            "resident); us in Gram and generator %.1f, %lld Newton-Schulz steps %.1f, "
            "rotation %.1f\n", (long long)Lg, (long long)l, h.iterations, h.converged, h.prof[2],
            h.prof[0] * 0.01, h.prof[1] * 0.01, (long long)grid, occ, h.prof[4] * 0.01,
            h.prof[3], h.prof[6] * 0.01, h.prof[7] * 0.01);
    if (g_trace_buf) {
      std::vector<long long> tb(POLAR_TRACE_BARRIERS * grid + POLAR_TRACE_BARRIERS);
      SEEMPS_HIP(hipMemcpy(tb.data(), g_trace_buf, tb.size() * 8, hipMemcpyDeviceToHost));
      (void)hipFree(g_trace_buf);
      g_trace_buf = nullptr;
      const int nb = (int)std::min<long long>(h.prof[2], POLAR_TRACE_BARRIERS);
      for (int b = 0; b < nb; ++b) {
        long long lo = tb[b * grid], hi = lo;
        int last = 0;
        for (int64_t g = 0; g < grid; ++g) {
          const long long v = tb[b * grid + g];
          lo = std::min(lo, v);
          if (v > hi) { hi = v; last = (int)g; }
        }
        const long long dep = tb[POLAR_TRACE_BARRIERS * grid + b];
        const long long prev = b > 0 ? tb[POLAR_TRACE_BARRIERS * grid + b - 1] : lo;
        fprintf(stderr, "  barrier %2d: phase %6.1f us (wg0 %6.1f), arrival spread %6.1f us "
                "(last wg %3d), release %5.1f us\n", b, (hi - prev) * 0.01,
                (tb[b * grid] - prev) * 0.01, (hi - lo) * 0.01, last, (dep - hi) * 0.01);
      }
    }
  }
  if (tiny_dev) *tiny_dev = w.tiny;
  if (conv_dev) *conv_dev = &ctrl->converged;
  return SEEMPS_OK;
}

}  // namespace seemps
