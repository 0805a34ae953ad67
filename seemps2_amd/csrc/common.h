// Shared definitions for the gfx950 kernels of the SeeMPS hot path.
//
// Storage conventions (all row-major, as NumPy / torch C-order):
//   complex128 values are interleaved (re, im) pairs, accessed as double2.
//   MPS tensor A[a, i, b], MPO tensor W[a, i_out, j_in, b],
//   MPO environments L[bra, mpo, ket], R[bra, mpo, ket].
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/seemps_hip.h"

namespace seemps {

// Last error message of the calling host thread (returned by seemps_last_error).
void set_error(const char* fmt, ...);

#define SEEMPS_CHECK_ARG(cond, ...)                          \
  do {                                                       \
    if (!(cond)) {                                           \
      ::seemps::set_error(__VA_ARGS__);                      \
      return SEEMPS_ERR_INVALID;                             \
    }                                                        \
  } while (0)

#define SEEMPS_HIP(call)                                                   \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      ::seemps::set_error("%s failed: %s", #call, hipGetErrorString(e_));  \
      return SEEMPS_ERR_RUNTIME;                                           \
    }                                                                      \
  } while (0)

// Check the most recent launch.
int check_launch(const char* what);

// Brackets one launch with HIP events when timing is enabled for `name`
// (seemps_timing_enable): construct right before the launch, on its stream.
class TimedLaunch {
 public:
  TimedLaunch(const char* name, hipStream_t stream, double flops);
  ~TimedLaunch();
  TimedLaunch(const TimedLaunch&) = delete;
  TimedLaunch& operator=(const TimedLaunch&) = delete;

 private:
  int slot_;
  hipEvent_t start_;
  hipStream_t stream_;
  double flops_;
};

struct cplx {
  double re, im;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  // a * b + c
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 c) {
  // conj(a) * b + c
  return make_double2(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Pinned host buffer of at least `bytes` for small device-to-host reads (one
// per host thread; nullptr if it cannot be allocated).  The caller copies into
// it on its stream and synchronises before reading.
void* pinned_stage(size_t bytes);
// Layout of the staging buffer: scalars at [0, PINNED_CTRL_OFF), the SVD
// control block at [PINNED_CTRL_OFF, PINNED_BYTES).
constexpr size_t PINNED_CTRL_OFF = 1024;
constexpr size_t PINNED_BYTES = 4096;

// Device properties cached per device (CU count, cooperative limits).
struct DeviceInfo {
  int device;
  int num_cu;
};
int device_info(DeviceInfo* info);

// Internal enqueue functions shared between the translation units.
// prec: SEEMPS_PREC_FP64 (FP64 MFMA) or SEEMPS_PREC_FP32 (operands rounded to FP32,
// FP32 MFMA and accumulation; A, B and C stay complex128 in memory).
int zgemm(int opA, int opB, int64_t M, int64_t N, int64_t K, double2 alpha, const void* A,
          int64_t lda, int64_t sA, const void* B, int64_t ldb, int64_t sB, double2 beta, void* C,
          int64_t ldc, int64_t sC, int64_t batch, hipStream_t stream,
          int prec = SEEMPS_PREC_FP64);
// Split-K GEMM over caller scratch (ws holds splitk_workspace(M, N, K) elements);
// falls back to zgemm when splitting does not pay or ws is too small.
int zgemm_ws(int opA, int opB, int64_t M, int64_t N, int64_t K, double2 alpha, const void* A,
             int64_t lda, const void* B, int64_t ldb, double2 beta, void* C, int64_t ldc,
             double2* ws, int64_t ws_elems, hipStream_t stream, int prec = SEEMPS_PREC_FP64);
int64_t splitk_workspace(int64_t M, int64_t N, int64_t K);
int zaxpby_impl(int64_t n, double2 a, const double2* x, int conj_x, double2 b, double2* y,
                hipStream_t stream);
int zlocal_impl(int64_t P, int64_t I, int64_t Q, int64_t J1, int64_t J2, const double2* in,
                const double2* Mop, double2* out, int accumulate, hipStream_t stream);
int zscale2d_impl(int64_t rows, int64_t cols, const double* sr, const double* sc,
                  const double2* src, int64_t ld_src, double2* dst, int64_t ld_dst,
                  hipStream_t stream);
// Number of partial sums used by the two-pass reductions for length n.
int64_t reduce_blocks(int64_t n);
// result[j] = sum_k conj(V[j*ldv + k]) * w[k], j < nv (nv <= 64); work holds
// reduce_blocks(n) * nv partials.
int zmdotc_impl(int64_t n, int64_t nv, const double2* V, int64_t ldv, const double2* w,
                double2* result, double2* work, hipStream_t stream);

// SVD and QR building blocks (svd.hip, qr.hip).
// precondition = false skips the QR preconditioner (for matrices whose columns
// are already nearly orthogonal); the workspace query, the SVD and the
// completion must then all be called with the same flag.
int svd_enqueue(int64_t m, int64_t n, const double2* A, int64_t lda, double2* U, double* S,
                double2* VH, void* work, int64_t work_bytes, hipStream_t stream,
                int* info_dev_off, bool complete = true, bool precondition = true);
int svd_complete_enqueue(int64_t m, int64_t n, double2* U, double2* VH, void* work,
                         const int64_t* kmax_dev, hipStream_t stream, bool precondition = true);
int64_t svd_workspace_bytes(int64_t m, int64_t n, bool precondition = true);
int truncate_enqueue(int64_t N, const double* s, int method, double tol, int64_t max_bond,
                     int64_t* D_out, double* err_out, hipStream_t stream);
int svd_read_ctrl(void* work, int off, int32_t* info, hipStream_t stream);
int64_t svd_split_bytes(int64_t m, int64_t n);
int svd_split_impl(int64_t m, int64_t n, const double2* A, int mode, int method, double tolerance,
                   int64_t max_bond, double2* outA, double2* outB, double* s_out,
                   int64_t* D_host, double* err_host, void* work, int64_t work_bytes,
                   int32_t* info, hipStream_t stream);
// SVD of an Lg x l matrix with nearly orthogonal columns by simultaneous
// Jacobi rotations applied as unitary polar factors (polar_svd.hip).
int64_t polar_svd_bytes(int64_t Lg, int64_t l);
int polar_svd(int64_t Lg, int64_t l, const double2* W, int64_t ldw, double2* U, double* S,
              double2* VH, const int64_t** tiny_dev, const int** conv_dev, void* work,
              int64_t work_bytes, hipStream_t stream);
int qr_factor(double2* A, int64_t len, int64_t ncol, int nb_max, double2* Vall, double2* Tall,
              double2* tau, double2* Z, double2* ks, int64_t ks_elems, hipStream_t stream);
int qr_apply_q(double2* Y, int64_t len, int64_t ny, int64_t ncol, int nb_max, const double2* Vall,
               const double2* Tall, double2* Z, double2* ks, int64_t ks_elems,
               hipStream_t stream);

}  // namespace seemps
