// Blocked Householder QR of a column-major working matrix (FP64 complex), used
// to precondition the one-sided Jacobi SVD (Jacobi on R^H converges in a few
// sweeps even for strongly graded spectra, where Jacobi on A needs tens).
//
// Storage: A_cm holds ncol columns of length len (len >= ncol), column c at
// A_cm + c * len.  After qr_factor, column c holds R[0..c, c] in rows 0..c and
// the Householder vector v_c (implicit leading 1) in rows c+1..len-1.  Each
// panel of NB columns keeps an explicit copy of its vectors (Vp, NB x L with
// L = len - j0, unit diagonal, zeros above) and the compact-WY factor T
// (NB x NB upper triangular) so that H_j0 ... H_j0+NB-1 = I - V T V^H.
// The panel is factorised by one workgroup; trailing updates and the
// application of Q are complex GEMMs on the FP64 matrix cores.
//
// Reflector convention (LAPACK zlarfg): H = I - tau v v^H with
// H^H [alpha; x] = [beta; 0], beta real.
#include <cmath>

#include "common.h"

namespace seemps {
namespace {

constexpr int QR_THREADS = 256;

__device__ double2 block_reduce_sum(double2 v, double2* red) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 s = make_double2(0.0, 0.0);
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = cadd(s, red[w]);
    red[8] = s;
  }
  __syncthreads();
  const double2 s = red[8];
  __syncthreads();
  return s;
}

// Factorise the panel of nb columns starting at j0; write Vp (nb x L) and T.
__global__ __launch_bounds__(QR_THREADS) void qr_panel_kernel(double2* __restrict__ A, int64_t len,
                                                              int64_t j0, int nb,
                                                              double2* __restrict__ tau,
                                                              double2* __restrict__ Vp,
                                                              double2* __restrict__ T) {
  __shared__ double2 red[16];
  __shared__ double2 s_tau[32];
  const int tid = threadIdx.x;
  const int64_t L = len - j0;
  for (int j = 0; j < nb; ++j) {
    const int64_t c = j0 + j;
    double2* col = A + c * len;
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t r = c + 1 + tid; r < len; r += blockDim.x) acc.x += cabs2(col[r]);
    const double xnorm2 = block_reduce_sum(acc, red).x;
    const double2 alpha = col[c];
    double2 t = make_double2(0.0, 0.0);
    double beta = alpha.x;
    double2 scal = make_double2(0.0, 0.0);
    if (!(xnorm2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(cabs2(alpha) + xnorm2);
      beta = (alpha.x >= 0.0) ? -nrm : nrm;
      t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      // scal = 1 / (alpha - beta)
      const double2 d = make_double2(alpha.x - beta, alpha.y);
      const double dd = cabs2(d);
      scal = make_double2(d.x / dd, -d.y / dd);
    }
    __syncthreads();
    if (!(t.x == 0.0 && t.y == 0.0)) {
      for (int64_t r = c + 1 + tid; r < len; r += blockDim.x) col[r] = cmul(scal, col[r]);
    }
    __syncthreads();
    if (tid == 0) {
      col[c] = make_double2(beta, 0.0);
      tau[c] = t;
      s_tau[j] = t;
    }
    __syncthreads();
    // Apply H^H = I - conj(t) v v^H to the remaining panel columns.
    if (!(t.x == 0.0 && t.y == 0.0)) {
      for (int jj = j + 1; jj < nb; ++jj) {
        double2* cc = A + (j0 + jj) * len;
        const double2 ccc = cc[c];  // read before thread 0 rewrites it below
        double2 a = make_double2(0.0, 0.0);
        for (int64_t r = c + 1 + tid; r < len; r += blockDim.x) a = cfmac(col[r], cc[r], a);
        double2 w = block_reduce_sum(a, red);
        w = cadd(w, ccc);  // v[c] = 1
        const double2 f = cmul(make_double2(t.x, -t.y), w);  // conj(t) * (v^H a)
        if (tid == 0) cc[c] = make_double2(ccc.x - f.x, ccc.y - f.y);
        for (int64_t r = c + 1 + tid; r < len; r += blockDim.x) {
          const double2 p = cmul(f, col[r]);
          cc[r] = make_double2(cc[r].x - p.x, cc[r].y - p.y);
        }
        __syncthreads();
      }
    }
  }
  // Explicit V (nb x L): unit diagonal, zeros above.
  for (int64_t e = tid; e < (int64_t)nb * L; e += blockDim.x) {
    const int j = (int)(e / L);
    const int64_t rr = e % L;
    double2 v;
    if (rr < j)
      v = make_double2(0.0, 0.0);
    else if (rr == j)
      v = make_double2(1.0, 0.0);
    else
      v = A[(j0 + j) * len + j0 + rr];
    Vp[e] = v;
  }
  __syncthreads();
  // T (zlarft, forward, columnwise): T[j][j] = tau_j,
  // T[0:j, j] = -tau_j * T[0:j, 0:j] * (V[:, 0:j]^H v_j).
  __shared__ double2 s_y[32];
  __shared__ double2 s_T[32 * 32];
  for (int e = tid; e < nb * nb; e += blockDim.x) s_T[e] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    for (int i = 0; i < j; ++i) {
      double2 a = make_double2(0.0, 0.0);
      for (int64_t rr = j + tid; rr < L; rr += blockDim.x)
        a = cfmac(Vp[(int64_t)i * L + rr], Vp[(int64_t)j * L + rr], a);
      const double2 y = block_reduce_sum(a, red);
      if (tid == 0) s_y[i] = y;
    }
    __syncthreads();
    if (tid == 0) {
      const double2 tj = s_tau[j];
      for (int i = 0; i < j; ++i) {
        double2 acc2 = make_double2(0.0, 0.0);
        for (int k = i; k < j; ++k) acc2 = cfma(s_T[i * nb + k], s_y[k], acc2);
        const double2 v = cmul(tj, acc2);
        s_T[i * nb + j] = make_double2(-v.x, -v.y);
      }
      s_T[j * nb + j] = tj;
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += blockDim.x) T[e] = s_T[e];
}

// Panel factorisation with the whole panel resident in LDS (nb <= 16 columns
// of L <= QR_LDS_ROWS rows, one thread per row): one load, two barriers per
// column (dot products and norm | reflector and updates, row by row), the T
// factor from wavefront-parallel dot products, one write-back.  Same outputs
// as qr_panel_kernel.
constexpr int QR_LDS_THREADS = 1024;
constexpr int QR_LDS_ROWS = 560;
constexpr int QR_LDS_NB = 16;

__global__ __launch_bounds__(QR_LDS_THREADS) void qr_panel_lds_kernel(
    double2* __restrict__ A, int64_t len, int64_t j0, int nb, double2* __restrict__ tau,
    double2* __restrict__ Vp, double2* __restrict__ T) {
  __shared__ double2 P[QR_LDS_NB * QR_LDS_ROWS];  // column j at P + j * L
  __shared__ double2 s_T[QR_LDS_NB * QR_LDS_NB];
  __shared__ double2 s_y[QR_LDS_NB * QR_LDS_NB];  // y[i][j] = V_i^H V_j, i < j
  __shared__ double2 s_w[QR_LDS_NB];    // unscaled column-j dot products
  __shared__ double2 s_ccj[QR_LDS_NB];  // trailing columns' row-j entries
  __shared__ double2 s_tau[QR_LDS_NB];
  __shared__ double2 s_alpha;
  __shared__ double s_xn;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nwaves = QR_LDS_THREADS / 64;
  const int L = (int)(len - j0);
  for (int e = tid; e < nb * L; e += QR_LDS_THREADS) {
    const int j = e / L, rr = e % L;
    P[e] = A[(j0 + j) * len + j0 + rr];
  }
  __syncthreads();
  // Two barriers per column.  Phase 1 only reads: wavefront q < nt forms the
  // dot product of the unscaled column j with trailing column j + 1 + q (rows
  // below j) and keeps that column's row-j entry; wavefront nt forms the norm
  // of column j below the diagonal and keeps alpha = column j's row j.  Phase 2:
  // thread rr owns row rr of the panel: it scales row rr of the reflector and
  // applies every trailing column's update to row rr, so no two threads touch
  // the same entry and nothing is read after being written.
  for (int j = 0; j < nb; ++j) {
    double2* col = P + j * L;
    const int nt = nb - j - 1;  // trailing columns (<= 15 < nwaves)
    if (wave < nt) {
      const double2* cc = P + (j + 1 + wave) * L;
      double2 a = make_double2(0.0, 0.0);
      for (int rr = j + 1 + lane; rr < L; rr += 64) a = cfmac(col[rr], cc[rr], a);
      for (int off = 32; off > 0; off >>= 1) {
        a.x += __shfl_down(a.x, off, 64);
        a.y += __shfl_down(a.y, off, 64);
      }
      if (lane == 0) {
        s_w[wave] = a;
        s_ccj[wave] = cc[j];
      }
    } else if (wave == nt) {
      double acc = 0.0;
      for (int rr = j + 1 + lane; rr < L; rr += 64) acc += cabs2(col[rr]);
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
      if (lane == 0) {
        s_xn = acc;
        s_alpha = col[j];
      }
    }
    __syncthreads();  // #1
    const double2 alpha = s_alpha;
    const double xnorm2 = s_xn;
    double2 t = make_double2(0.0, 0.0);
    double beta = alpha.x;
    double2 scal = make_double2(0.0, 0.0);
    const bool reflect = !(xnorm2 == 0.0 && alpha.y == 0.0);
    if (reflect) {
      const double nrm = sqrt(cabs2(alpha) + xnorm2);
      beta = (alpha.x >= 0.0) ? -nrm : nrm;
      t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 d = make_double2(alpha.x - beta, alpha.y);
      const double dd = cabs2(d);
      scal = make_double2(d.x / dd, -d.y / dd);
    }
    const int rr = tid;
    if (rr >= j && rr < L) {
      if (reflect) {
        // v = [1; scal col[j+1:]];  f_q = conj(t) v^H c_q = conj(t) (c_q[j] + conj(scal) w_q)
        const double2 ct = make_double2(t.x, -t.y), cs = make_double2(scal.x, -scal.y);
        const double2 v = (rr == j) ? make_double2(1.0, 0.0) : cmul(scal, col[rr]);
        for (int q = 0; q < nt; ++q) {
          const double2 f = cmul(ct, cadd(s_ccj[q], cmul(cs, s_w[q])));
          double2* cc = P + (j + 1 + q) * L;
          const double2 p = cmul(f, v);
          cc[rr] = make_double2(cc[rr].x - p.x, cc[rr].y - p.y);
        }
        if (rr > j) col[rr] = v;
      }
      if (rr == j) {
        col[j] = make_double2(beta, 0.0);
        s_tau[j] = t;
        tau[j0 + j] = t;
      }
    }
    __syncthreads();  // #2
  }
  // y[i][j] = conj(V_i[j]) + sum_{rr > j} conj(V_i[rr]) V_j[rr], i < j
  for (int q = wave; q < nb * nb; q += nwaves) {
    const int i = q / nb, j = q % nb;
    if (i >= j) continue;
    const double2* ci = P + i * L;
    const double2* cj = P + j * L;
    double2 a = make_double2(0.0, 0.0);
    for (int rr = j + 1 + lane; rr < L; rr += 64) a = cfmac(ci[rr], cj[rr], a);
    for (int off = 32; off > 0; off >>= 1) {
      a.x += __shfl_down(a.x, off, 64);
      a.y += __shfl_down(a.y, off, 64);
    }
    if (lane == 0) s_y[i * QR_LDS_NB + j] = cadd(a, make_double2(ci[j].x, -ci[j].y));
  }
  for (int e = tid; e < QR_LDS_NB * QR_LDS_NB; e += QR_LDS_THREADS) s_T[e] = make_double2(0, 0);
  __syncthreads();
  // T (zlarft, forward, columnwise): T[j][j] = tau_j,
  // T[0:j, j] = -tau_j * T[0:j, 0:j] * y[0:j, j].
  for (int j = 0; j < nb; ++j) {
    if (tid < j) {
      const int i = tid;
      double2 acc2 = make_double2(0.0, 0.0);
      for (int k = i; k < j; ++k) acc2 = cfma(s_T[i * QR_LDS_NB + k], s_y[k * QR_LDS_NB + j], acc2);
      const double2 v = cmul(s_tau[j], acc2);
      s_T[i * QR_LDS_NB + j] = make_double2(-v.x, -v.y);
    } else if (tid == j) {
      s_T[j * QR_LDS_NB + j] = s_tau[j];
    }
    __syncthreads();
  }
  for (int e = tid; e < nb * nb; e += QR_LDS_THREADS)
    T[e] = s_T[(e / nb) * QR_LDS_NB + e % nb];
  for (int e = tid; e < nb * L; e += QR_LDS_THREADS) {
    const int j = e / L, rr = e % L;
    const double2 p = P[e];
    A[(j0 + j) * len + j0 + rr] = p;
    Vp[e] = (rr < j) ? make_double2(0.0, 0.0) : (rr == j) ? make_double2(1.0, 0.0) : p;
  }
}

}  // namespace

// In-place blocked QR of A_cm (ncol columns of length len, len >= ncol).
// Work: Vall (ncol * len complex), Tall (npanels * NB * NB), tau (ncol), Z (NB * ncol * 2),
// ks (ks_elems complex, split-K partials of the NB-row products; may be 0).
int qr_factor(double2* A, int64_t len, int64_t ncol, int nb_max, double2* Vall, double2* Tall,
              double2* tau, double2* Z, double2* ks, int64_t ks_elems, hipStream_t stream) {
  int st;
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0),
                mone = make_double2(-1.0, 0.0);
  for (int64_t j0 = 0, p = 0; j0 < ncol; j0 += nb_max, ++p) {
    const int nb = (int)((ncol - j0) < nb_max ? (ncol - j0) : nb_max);
    const int64_t L = len - j0;
    double2* Vp = Vall + j0 * len;  // nb x L block (fits: nb * L <= nb * len)
    double2* T = Tall + p * nb_max * nb_max;
    if (L <= QR_LDS_ROWS && nb <= QR_LDS_NB) {
      TimedLaunch timed("qr_panel_lds_kernel", stream, 0.0);
      hipLaunchKernelGGL(qr_panel_lds_kernel, dim3(1), dim3(QR_LDS_THREADS), 0, stream, A, len,
                         j0, nb, tau, Vp, T);
      if ((st = check_launch("qr_panel_lds"))) return st;
    } else {
      TimedLaunch timed("qr_panel_kernel", stream, 0.0);
      hipLaunchKernelGGL(qr_panel_kernel, dim3(1), dim3(QR_THREADS), 0, stream, A, len, j0, nb,
                         tau, Vp, T);
      if ((st = check_launch("qr_panel"))) return st;
    }
    const int64_t ntr = ncol - j0 - nb;
    if (ntr <= 0) continue;
    double2* At = A + (j0 + nb) * len + j0;  // ntr x L, row stride len
    double2* Z1 = Z;                          // nb x ntr
    double2* Z2 = Z + nb_max * ncol;          // nb x ntr
    // Z1 = conj(Vp) At^T ; Z2 = T^H Z1 ; At -= Z2^T Vp
    if ((st = zgemm_ws(SEEMPS_OP_C, SEEMPS_OP_T, nb, ntr, L, one, Vp, L, At, len, zero, Z1, ntr,
                       ks, ks_elems, stream)))
      return st;
    if ((st = zgemm(SEEMPS_OP_H, SEEMPS_OP_N, nb, ntr, nb, one, T, nb, 0, Z1, ntr, 0, zero, Z2, ntr,
                    0, 1, stream)))
      return st;
    if ((st = zgemm(SEEMPS_OP_T, SEEMPS_OP_N, ntr, L, nb, mone, Z2, ntr, 0, Vp, L, 0, one, At, len,
                    0, 1, stream)))
      return st;
  }
  return SEEMPS_OK;
}

// Y <- Q Y for Y column-major (ny columns of length len), Q from qr_factor.
int qr_apply_q(double2* Y, int64_t len, int64_t ny, int64_t ncol, int nb_max, const double2* Vall,
               const double2* Tall, double2* Z, double2* ks, int64_t ks_elems,
               hipStream_t stream) {
  int st;
  const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0),
                mone = make_double2(-1.0, 0.0);
  const int64_t npan = ceil_div(ncol, nb_max);
  for (int64_t p = npan - 1; p >= 0; --p) {
    const int64_t j0 = p * nb_max;
    const int nb = (int)((ncol - j0) < nb_max ? (ncol - j0) : nb_max);
    const int64_t L = len - j0;
    const double2* Vp = Vall + j0 * len;
    const double2* T = Tall + p * nb_max * nb_max;
    double2* Yt = Y + j0;  // ny x L, row stride len
    double2* Z1 = Z;
    double2* Z2 = Z + nb_max * (ny > ncol ? ny : ncol);
    if ((st = zgemm_ws(SEEMPS_OP_C, SEEMPS_OP_T, nb, ny, L, one, Vp, L, Yt, len, zero, Z1, ny, ks,
                       ks_elems, stream)))
      return st;
    if ((st = zgemm(SEEMPS_OP_N, SEEMPS_OP_N, nb, ny, nb, one, T, nb, 0, Z1, ny, 0, zero, Z2, ny, 0,
                    1, stream)))
      return st;
    if ((st = zgemm(SEEMPS_OP_T, SEEMPS_OP_N, ny, L, nb, mone, Z2, ny, 0, Vp, L, 0, one, Yt, len, 0,
                    1, stream)))
      return st;
  }
  return SEEMPS_OK;
}

}  // namespace seemps
