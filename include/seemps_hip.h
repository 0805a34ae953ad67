/*
 * seemps_hip.h -- C ABI of libseemps_hip.so, the gfx950 (MI355X) kernels of the
 * SeeMPS tensor-train hot path.
 *
 * Conventions
 *   - Every pointer to tensor data is a DEVICE pointer (HIP global memory).
 *   - complex128 = interleaved (re, im) doubles; tensors are row-major (C order),
 *     with the layouts of the reference: MPS A[a, i, b], MPO W[a, i_out, j_in, b],
 *     MPS environments rho[bra, ket], MPO environments L/R[bra, mpo, ket].
 *   - `stream` is a hipStream_t (NULL = default stream).  All work is enqueued on
 *     it; entry points that must return a size to the host (the kept bond
 *     dimension D, eigensolver statistics) synchronise that stream once.
 *   - Scratch memory is caller-provided (`work`, `work_bytes`); the library never
 *     allocates persistent device memory.  `*_workspace()` gives the bytes needed.
 *   - Return value: SEEMPS_OK or an error code; seemps_last_error() describes it.
 *     The Python layer maps SEEMPS_ERR_INVALID -> ValueError,
 *     SEEMPS_ERR_NOCONVERGE -> scipy.linalg.LinAlgError, SEEMPS_ERR_RUNTIME ->
 *     RuntimeError, as the reference does (src/seemps/cython/svd.pxi:78-83,
 *     contractions.pxi:37, environments.pxi:17,48).
 *
 * Citations are to the reference repository (juanjosegarciaripoll/seemps2).
 */
#ifndef SEEMPS_HIP_H
#define SEEMPS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEEMPS_OK 0
#define SEEMPS_ERR_INVALID 1
#define SEEMPS_ERR_NOCONVERGE 2
#define SEEMPS_ERR_RUNTIME 3
#define SEEMPS_ERR_WORKSPACE 4

/* Operand flags.  0/1/2 are GemmOrder.NORMAL/TRANSPOSE/ADJOINT of the reference
 * (src/seemps/cython/gemm.pxi:8-11, _python_core.py:48-54); 3 = conjugate. */
#define SEEMPS_OP_N 0
#define SEEMPS_OP_T 1
#define SEEMPS_OP_H 2
#define SEEMPS_OP_C 3

/* Arithmetic of the contractions (no counterpart in the reference, which is
 * FP64 throughout).  FP32: operands rounded to FP32 on load, FP32 MFMA products
 * and accumulation, complex128 in memory. */
#define SEEMPS_PREC_FP64 0
#define SEEMPS_PREC_FP32 1
/* Eigensolver only: FP32 contractions until the Ritz residual reaches
 * 1e-5 |theta|, FP64 from there to the requested tolerance. */
#define SEEMPS_PREC_MIXED 2

/* Truncation method codes: Truncation.* of src/seemps/cython/truncation.pxi:1-8. */
#define SEEMPS_TRUNC_NONE 0
#define SEEMPS_TRUNC_REL_SV 1
#define SEEMPS_TRUNC_REL_NORM2 2
#define SEEMPS_TRUNC_ABS_SV 3

/* Message for the last error raised on the calling host thread. */
const char* seemps_last_error(void);
/* ABI version of this header (incremented on incompatible changes). */
int seemps_abi_version(void);
/* Free the library's process-lifetime host resources: the pinned staging
 * buffers of the device-to-host reads and the pooled timing events (pending
 * timings are drained first).  Call it before the HIP runtime shuts down (the
 * Python layer registers it with atexit).  The library stays usable: later
 * calls allocate again.  No counterpart in the reference. */
int seemps_release(void);

/* ---------------------------------------------------------------------------
 * Contractions
 * ------------------------------------------------------------------------- */

/* C[b] = alpha * op(A[b]) @ op(B[b]) + beta * C[b], complex128, b < batch.
 * op(A) is M x K, op(B) is K x N; lda/ldb/ldc are row strides (elements) of the
 * stored matrices, stride* the batch strides (elements).  beta == 0 never reads C.
 * Replaces: _gemm (src/seemps/cython/gemm.pxi:16-122), _contract_last_and_first
 * (contractions.pxi:6-38), the two GEMMs of _update_left/right_environment
 * (environments.pxi:10-62) and np.matmul in AntilinearForm.tensor2site
 * (state/antilinear.py:96). */
int seemps_zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K,
                 double alpha_re, double alpha_im,
                 const void* A, int64_t lda, int64_t strideA,
                 const void* B, int64_t ldb, int64_t strideB,
                 double beta_re, double beta_im,
                 void* C, int64_t ldc, int64_t strideC,
                 int64_t batch, void* stream);

/* Single (batch 1) seemps_zgemm that splits K over a caller workspace when the
 * output has too few tiles to fill the device (slices reduced in a fixed
 * order: deterministic).  seemps_zgemm_workspace gives the bytes (0 = no
 * split pays; any work pointer is then accepted). */
int64_t seemps_zgemm_workspace(int64_t M, int64_t N, int64_t K);
int seemps_zgemm_ws(int opA, int opB, int64_t M, int64_t N, int64_t K,
                    double alpha_re, double alpha_im, const void* A, int64_t lda,
                    const void* B, int64_t ldb, double beta_re, double beta_im,
                    void* C, int64_t ldc, void* work, int64_t work_bytes, void* stream);

/* seemps_zgemm with the product in `precision` (SEEMPS_PREC_FP64 or
 * SEEMPS_PREC_FP32).  No counterpart in the reference: the FP32 form is the
 * contraction inside the mixed-precision eigensolver. */
int seemps_zgemm_prec(int precision, int opA, int opB, int64_t M, int64_t N, int64_t K,
                      double alpha_re, double alpha_im,
                      const void* A, int64_t lda, int64_t strideA,
                      const void* B, int64_t ldb, int64_t strideB,
                      double beta_re, double beta_im,
                      void* C, int64_t ldc, int64_t strideC,
                      int64_t batch, void* stream);

/* Site-wise product of two MPS tensors (the MPS Hadamard product, one site):
 * out[(x, y), k, (u, v)] = A[x, k, u] B[y, k, v], A (a, d, b), B (c, d, e),
 * out (a*c, d, b*e).  Replaces the broadcast product of MPS.__mul__(MPS)
 * (state/mps.py:201-210). */
int seemps_zsite_product(int64_t a, int64_t d, int64_t b, int64_t c, int64_t e,
                         const void* A, const void* B, void* out, void* stream);

/* Grouped complex GEMM (no counterpart in the reference: the batching of many
 * MPS of uneven size).  C_p = op(A) op(B) for every problem p of desc (a
 * device array of nprob SeempsGemmDesc, row-major operands as in
 * seemps_zgemm, alpha = 1, beta = 0); tile_prob (device, ntiles entries)
 * lists the problem of every output tile, each problem's tiles contiguous
 * from desc[p].tile0 in row-major tile order with desc[p].tiles_n tiles per
 * tile row (seemps_grouped_tiles gives a problem's tile count).  tile_cfg:
 * SEEMPS_TILE_64x64 (256 threads) or SEEMPS_TILE_32x32 (64 threads).  Op
 * pairs: (N,N), (H,N), (N,T), (T,N), (C,N), (N,H).  Allocates nothing and
 * does not synchronise, so it can be captured in a HIP graph. */
#define SEEMPS_TILE_64x64 0
#define SEEMPS_TILE_32x32 1
typedef struct {
  int64_t M, N, K;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  int64_t tiles_n;
  int64_t tile0;
} SeempsGemmDesc;
int64_t seemps_grouped_tiles(int tile_cfg, int64_t M, int64_t N);
int seemps_zgemm_grouped(int opA, int opB, int tile_cfg, int64_t nprob, const void* desc,
                         const int32_t* tile_prob, int64_t ntiles, void* stream);

/* Local-operator contraction, complex128:
 *   out[p, j1, q, j2] = sum_i in[p, i, q] * Mop[i, j1, j2]   (+ out if accumulate)
 * with in[P, I, Q], Mop[I, J1, J2], out[P, J1, Q, J2] contiguous (any sizes: the
 * contracted index runs in LDS tiles, accumulated in order).  Replaces the small-index steps of _mpo_multiply_tensor
 * (src/seemps/operators/mpo.py:26-45, "cjd,aijb->caidb"), of the MPO environment
 * updates (state/environments.py:19-42), of the two-site matvec
 * (operators/quadratic.py:204-214) and of _contract_nrjl_ijk_klm
 * (cython/contractions.pxi:102-132). */
int seemps_zlocal(int64_t P, int64_t I, int64_t Q, int64_t J1, int64_t J2,
                  const void* in, const void* Mop, void* out, int accumulate,
                  void* stream);

/* y[k] = a * op(x[k]) + b * y[k] for k < n (op = N or C); b == 0 never reads y. */
int seemps_zaxpby(int64_t n, double a_re, double a_im, const void* x, int conj_x,
                  double b_re, double b_im, void* y, void* stream);

/* dst[r, c] = sr[r] * src[r, c] * sc[c] for r < rows, c < cols, with optional
 * real scalings sr / sc (NULL = 1) and row strides ld_src / ld_dst (elements).
 * Used for the s*V / U*s products of schmidt.pxi:40,81,145,171. */
int seemps_zscale2d(int64_t rows, int64_t cols, const double* sr, const double* sc,
                    const void* src, int64_t ld_src, void* dst, int64_t ld_dst,
                    void* stream);

/* ---------------------------------------------------------------------------
 * Reductions (deterministic two-pass wavefront reductions)
 * ------------------------------------------------------------------------- */

int64_t seemps_reduce_workspace(int64_t n);
/* result[0] = sum_k conj(x[k]) * y[k] (device complex128).  np.vdot as used by
 * simplify (state/simplification.py:125-126) and _join_environments
 * (cython/environments.pxi:79-85). */
int seemps_zdotc(int64_t n, const void* x, const void* y, void* result,
                 void* work, int64_t work_bytes, void* stream);

/* seemps_zdotc without the workspace argument.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * partial sums go to a library-owned static device array of 512 complex
 * values (8 KiB per GPU, held until the process ends), enough for every n.
 * Allocates nothing.  Calls that overlap on one GPU share that array and
 * race: serialise them (one stream, or events). */
int seemps_zdotc_noalloc(int64_t n, const void* x, const void* y, void* result, void* stream);

/* ---------------------------------------------------------------------------
 * Truncation and SVD
 * ------------------------------------------------------------------------- */

/* Truncation rule on a descending FP64 spectrum s[0..N): writes the kept size
 * D_out[0] (device int64) and the squared error err_out[0] (device double).
 * Replaces destructively_truncate_vector / Strategy._truncate
 * (src/seemps/cython/truncation.pxi:222-334), same loops and summation order. */
int seemps_truncate(int64_t N, const double* s, int method, double tolerance,
                    int64_t max_bond, int64_t* D_out, double* err_out, void* stream);

/* Economy SVD A = U diag(S) VH of an m x n complex128 matrix (row stride lda),
 * S descending, U m x r (row stride r), VH r x n (row stride n), r = min(m, n).
 * One-sided block Jacobi in FP64 on one device (cooperative persistent kernel).
 * info (host, may be NULL) receives {sweeps, converged}.  Returns
 * SEEMPS_ERR_NOCONVERGE if the sweep limit is reached.
 * Replaces _destructive_svd (src/seemps/cython/svd.pxi:17-83). */
int64_t seemps_svd_workspace(int64_t m, int64_t n);
int seemps_zgesvd(int64_t m, int64_t n, const void* A, int64_t lda, void* U,
                  double* S, void* VH, void* work, int64_t work_bytes, int32_t* info,
                  void* stream);

/* SVD + truncation + split of an m x n matrix (the two-site tensor reshaped as
 * (a*d1, d2*b)).  mode 0 = left-orthogonal split: outA = U[:, :D] (m x D),
 * outB = diag(s[:D]) VH[:D] (D x n) -- _left_orth_2site (schmidt.pxi:123-147),
 * __update_in_canonical_form_right (schmidt.pxi:18-42).
 * mode 1 = right-orthogonal split: outA = U[:, :D] diag(s[:D]), outB = VH[:D]
 * -- _right_orth_2site (schmidt.pxi:150-174), __update_in_canonical_form_left
 * (schmidt.pxi:59-82).  outA must hold m*r elements, outB r*n, s_out r doubles
 * (may be NULL).  D_host / err_host receive D and the squared truncation error
 * (the stream is synchronised once). */
int64_t seemps_svd_split_workspace(int64_t m, int64_t n);
int seemps_svd_split(int64_t m, int64_t n, const void* A, int mode, int method,
                     double tolerance, int64_t max_bond, void* outA, void* outB,
                     double* s_out, int64_t* D_host, double* err_host, void* work,
                     int64_t work_bytes, int32_t* info, void* stream);

/* ---------------------------------------------------------------------------
 * DMRG two-site effective operator and eigensolver
 * ------------------------------------------------------------------------- */

/* out[a, ko, lo, g] = sum L[a, c, b] H[c, ko, k, lo, l, d] v[b, k, l, f] R[g, d, f]
 * where the MPO pair enters as Hm[(c, k, l), (ko, lo, d)] (w*d1*d2 x d1*d2*w2).
 * Shapes: L (a, c, b), v (b, d1, d2, f), R (g, w2, f).  Work: two intermediates.
 * Replaces QuadraticForm.two_site_operator._matvec (operators/quadratic.py:187-214). */
int64_t seemps_two_site_matvec_workspace(int64_t a, int64_t c, int64_t b, int64_t d1,
                                         int64_t d2, int64_t w2, int64_t f);
int seemps_two_site_matvec(int64_t a, int64_t c, int64_t b, int64_t d1, int64_t d2,
                           int64_t w2, int64_t f, int64_t g, const void* L,
                           const void* Hm, const void* R, const void* v, void* out,
                           void* work, int64_t work_bytes, void* stream);

/* Hm for the MPO pair (O1, O2): Hm[(c,k,l),(ko,lo,d)] = sum_e O1[c,ko,k,e] O2[e,lo,l,d]
 * (the precomputed H12 reshape of quadratic.py:103,198-201).  O1 is c x d1 x d1 x e,
 * O2 is e x d2 x d2 x w2. */
int seemps_two_site_operator(int64_t c, int64_t d1, int64_t e, int64_t d2, int64_t w2,
                             const void* O1, const void* O2, void* Hm, void* stream);

/* out = in.permute(perm) for a contiguous complex128 tensor of ndim <= 8 dims
 * (out axis k is input axis perm[k]); dims/perm are host arrays. */
int seemps_zpermute(int ndim, const int64_t* dims, const int* perm, const void* in,
                    void* out, void* stream);

/* Lowest eigenpair of the Hermitian two-site operator above for bra = ket
 * (a = b = chiL, g = f = chiR), Lanczos with full re-orthogonalisation (classical
 * Gram-Schmidt with the DGKS correction), Krylov dimension ncv (2..23), explicit
 * restarts from the Ritz vector, started from v0; stops when the Ritz residual
 * |beta_k y_k| <= tol * max(eps^(2/3), |theta|) (ARPACK's test; tol = 0 means
 * machine epsilon), tested at the end of each cycle (k = ncv) and, for
 * operators of dimension >= 8192, also after steps k = 4, 6, 8, ... so a warm
 * start stops before ncv products.  The Krylov recurrence, the tridiagonal eigensolve (FP64
 * bisection + inverse iteration) and the Ritz vector stay on the device; the
 * stream is synchronised once per restart.  evec receives the unit-norm
 * eigenvector; stats (host, may be NULL) = {eigenvalue, residual, matvecs,
 * restarts}.  Replaces _diagonalize_two_site (optimization/dmrg.py:15-23), i.e.
 * scipy eigsh -> ARPACK (scipy/sparse/linalg/_eigen/arpack/arpack.py:1565-1583). */
int64_t seemps_eigensolve_workspace(int64_t chiL, int64_t c, int64_t d1, int64_t d2,
                                    int64_t w2, int64_t chiR, int64_t ncv);
int seemps_two_site_eigensolve(int64_t chiL, int64_t c, int64_t d1, int64_t d2, int64_t w2,
                               int64_t chiR, const void* L, const void* Hm, const void* R,
                               const void* v0, void* evec, int64_t ncv, double tol,
                               int64_t maxiter, double* stats, void* work,
                               int64_t work_bytes, void* stream);
/* seemps_two_site_eigensolve with the contractions of the operator in
 * `precision`.  SEEMPS_PREC_MIXED: when the start vector's residual (from the
 * first, FP64 product) exceeds 1e-4 |alpha_0|, restarts run with FP32
 * contractions until the Ritz residual reaches 1e-5 max(eps^(2/3), |theta|) (or
 * stops falling); FP64 restarts then continue to the requested tol, so the
 * result meets the same test as the FP64 solver.  SEEMPS_PREC_FP32 stays in FP32 (tol below ~1e-6 cannot be met).
 * Orthogonalisation, tridiagonal solve and Ritz vectors are FP64 in all modes.
 * stats[4] (when stats has 5 slots, i.e. stats5 != 0) receives the number of
 * FP32 matvecs.  No counterpart in the reference. */
int seemps_two_site_eigensolve_prec(int precision, int64_t chiL, int64_t c, int64_t d1,
                                    int64_t d2, int64_t w2, int64_t chiR, const void* L,
                                    const void* Hm, const void* R, const void* v0, void* evec,
                                    int64_t ncv, double tol, int64_t maxiter, double* stats5,
                                    void* work, int64_t work_bytes, void* stream);

/* Fixed-rank randomized variant of seemps_svd_split (no counterpart in the
 * reference: the learned-rank truncation).  Sketch Y = A Omega with l columns
 * (Omega complex Gaussian from `seed`), power_iters refinements
 * Y <- A (A^H Y), Q = orth(Y) (Householder), B = Q^H A, Jacobi SVD of B,
 * U = Q U_B.  The truncation rule runs on B's l singular values with the
 * unsketched weight ||A||_F^2 - ||B||_F^2 counted as discarded.  When that
 * weight exceeds what the rule may discard while max_bond > l, when the bond
 * cap decides at or below l but the unsketched weight is neither negligible
 * (<= 1e-14 ||A||_F^2) nor far below the smallest kept value (<= 1e-3 s_D^2),
 * when a threshold rule would look past l values, for SEEMPS_TRUNC_NONE, or
 * when l >= min(m, n), the full
 * seemps_svd_split result is returned and *fallback_host = 1.  Otherwise same
 * outputs as seemps_svd_split (s_out receives l values). */
int64_t seemps_svd_split_sketch_workspace(int64_t m, int64_t n, int64_t l);
int seemps_svd_split_sketch(int64_t m, int64_t n, const void* A, int mode, int method,
                            double tolerance, int64_t max_bond, int64_t l, int64_t power_iters,
                            uint64_t seed, void* outA, void* outB, double* s_out,
                            int64_t* D_host, double* err_host, int32_t* fallback_host,
                            void* work, int64_t work_bytes, int32_t* info, void* stream);

/* Split sketched by an isometry of the bond (no counterpart in the reference:
 * the learned-rank truncation).  side 0: basis is k x n with orthonormal rows
 * (the right-canonical neighbour a sweep to the right replaces); side 1: m x k
 * with orthonormal columns (the left-canonical neighbour of a sweep to the
 * left).  Omega = [basis^H | p expansion directions] (side 0; side 1 alike on
 * A^H), the expansion directions being Z = X^H X G (G complex Gaussian from
 * `seed`, X = A or A^H) projected off the basis and orthonormalised (twice
 * each); W = X Omega; one-sided Jacobi SVD of W without QR preconditioning;
 * truncation on W's k + p singular values with ||A||_F^2 - ||W||_F^2 counted
 * as discarded, and the same fallback to seemps_svd_split as
 * seemps_svd_split_sketch (also when k + p > min(m, n)).  Outputs as
 * seemps_svd_split (s_out receives k + p values). */
int64_t seemps_svd_split_basis_workspace(int64_t m, int64_t n, int side, int64_t k, int64_t p);
int seemps_svd_split_basis(int64_t m, int64_t n, const void* A, int side, const void* basis,
                           int64_t k, int64_t p, uint64_t seed, int mode, int method,
                           double tolerance, int64_t max_bond, void* outA, void* outB,
                           double* s_out, int64_t* D_host, double* err_host,
                           int32_t* fallback_host, void* work, int64_t work_bytes,
                           int32_t* info, void* stream);

/* Features of the learned rank predictor for an m x n matrix A (row-major,
 * lda = n): 32 sketch columns Y = A Omega (Omega from `seed`), eigenvalues of
 * Y^H Y, features[0] = log10(min(m, n)) / 4 and features[i] =
 * clamp(log10(s_i / s_0), -8, 0) / 8 (s = sketch singular values), FP32 on
 * the device.  No counterpart in the reference. */
int64_t seemps_spectral_sketch_workspace(int64_t m, int64_t n);
int seemps_spectral_sketch(int64_t m, int64_t n, const void* A, uint64_t seed, float* features,
                           void* work, int64_t work_bytes, void* stream);

/* Rank-predictor MLP (FP32, no counterpart in the reference):
 * y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3 for B samples of F features
 * with H hidden units; params = W1 (H x F), b1, W2 (H x H), b2, w3 (H), b3.
 * Forward keeps the activations h1, h2 (B x H) for the backward pass, which
 * returns the gradients of the mean squared error against `target` (B) in the
 * same layout as params and the loss itself (device float). */
int64_t seemps_mlp_param_count(int F, int H);
int64_t seemps_mlp_scratch_count(int B, int H);
int seemps_mlp_forward(int B, int F, int H, const float* params, const float* x, float* h1,
                       float* h2, float* y, void* stream);
int seemps_mlp_backward(int B, int F, int H, const float* params, const float* x,
                        const float* h1, const float* h2, const float* y, const float* target,
                        float* grads, float* loss, float* scratch, void* stream);
/* Adam update of n parameters (step >= 1, bias-corrected). */
int seemps_adam_step(int64_t n, float* params, const float* grads, float* m, float* v, float lr,
                     float beta1, float beta2, float eps, int64_t step, void* stream);

/* ---------------------------------------------------------------------------
 * Kernel timing (no counterpart in the reference: measurement support for
 * benchmarks).  The library brackets each launch of the selected kernels with
 * HIP events on the launch stream and accumulates the elapsed times and the
 * algorithmic flops of each launch (8 per complex multiply-add; the SVD at
 * the nominal count 4 * (6 m n^2 + 16 n^3), m >= n).
 * ------------------------------------------------------------------------- */

/* filter: comma-separated kernel names (e.g. "svd_jacobi,zgemm_kernel"), "*"
 * for every instrumented kernel, NULL or "" to switch timing off.  Resets the
 * records. */
int seemps_timing_enable(const char* filter);
/* Launch `launches` empty kernels on `stream`, each bracketed like the
 * instrumented launches (record name "timing_calibration" when enabled): the
 * per-launch cost of the event bracketing itself, for subtraction. */
int seemps_timing_calibrate(int launches, void* stream);
/* Number of kernels with records.  Waits for the pending events. */
int seemps_timing_count(void);
/* Record `index` (< seemps_timing_count()): kernel name (NUL-terminated,
 * truncated to name_len), launches, summed milliseconds and summed flops. */
int seemps_timing_get(int index, char* name, int64_t name_len, int64_t* launches,
                      double* total_ms, double* flops);

#ifdef __cplusplus
}
#endif

#endif /* SEEMPS_HIP_H */
