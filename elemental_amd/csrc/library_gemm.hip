// The rank-NB trailing update of the mixed-precision LU as a plain library GEMM: rocBLAS
// gemm_ex with fp16 / bf16 operands, fp32 C and fp32 accumulation, on the caller's stream.
// The operands are the same K-major packs the hand-written kernel reads (row r of an operand
// is K contiguous halves); alpha and beta live in device memory (the power-of-two scale of
// the fp16 operand is only known on the device). On MI355X at 31744 x 31744 x 1024 with
// random data this runs at 663-678 TFLOP/s against 493-509 TFLOP/s for lpgemm_kernel; the
// rank-128 updates stay on lpgemm_kernel, which is as fast there.
#include <rocblas/rocblas.h>

#include <climits>
#include <cstdio>
#include <map>
#include <mutex>

#include "elx_common.hpp"

namespace {

std::mutex g_rb_mu;
std::map<int, rocblas_handle> g_rb;  // one handle per device

elx_status_t rb_handle(rocblas_handle* out) {
    int dev = 0;
    ELX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_rb_mu);
    auto it = g_rb.find(dev);
    if (it == g_rb.end()) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            std::fprintf(stderr, "elemental_amd: rocblas_create_handle failed\n");
            return ELX_ERROR_HIP;
        }
        it = g_rb.emplace(dev, h).first;
    }
    *out = it->second;
    return ELX_SUCCESS;
}

// out[0] = -2^-(mul * e), out[1] = out[0] * 2^-11, out[2] = 1 (e = 0 when e_dev is null)
__global__ void gemm_alphas_kernel(const int* e_dev, int mul, float* out) {
    const int e = e_dev ? *e_dev * mul : 0;
    const float a = -__builtin_ldexpf(1.0f, -e);
    out[0] = a;
    out[1] = a * (1.0f / 2048.0f);
    out[2] = 1.0f;
}

}  // namespace

namespace elx {

elx_status_t launch_gemm_alphas(hipStream_t s, const int* e_dev, int mul, float* out) {
    hipLaunchKernelGGL(gemm_alphas_kernel, dim3(1), dim3(1), 0, s, e_dev, mul, out);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t library_gemm_lowp(hipStream_t s, int prec, i64 m, i64 n, i64 k, const float* alpha_dev,
                               const uint16_t* A, i64 lda, const uint16_t* B, i64 ldb,
                               const float* beta_dev, float* C, i64 ldc) {
    if (m <= 0 || n <= 0 || k <= 0) return ELX_SUCCESS;
    ELX_ARG(prec == 0 || prec == 1);
    ELX_ARG(m < INT_MAX && n < INT_MAX && k < INT_MAX && lda < INT_MAX && ldb < INT_MAX &&
            ldc < INT_MAX && lda >= k && ldb >= k && ldc >= m);
    rocblas_handle h = nullptr;
    ELX_CHECK(rb_handle(&h));
    ProfScope prof_(s, kKLpgemm, 2.0 * (double)m * (double)n * (double)k);
    std::lock_guard<std::mutex> lk(g_rb_mu);  // the handle's stream and pointer mode
    if (rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_device) != rocblas_status_success)
        return ELX_ERROR_HIP;
    const rocblas_datatype dt = (prec == 0) ? rocblas_datatype_f16_r : rocblas_datatype_bf16_r;
    // C(m x n) += alpha * A_pack^T B_pack: as column-major matrices the packs are k x m and
    // k x n
    const rocblas_status st = rocblas_gemm_ex(
        h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
        (rocblas_int)k, alpha_dev, A, dt, (rocblas_int)lda, B, dt, (rocblas_int)ldb, beta_dev, C,
        rocblas_datatype_f32_r, (rocblas_int)ldc, C, rocblas_datatype_f32_r, (rocblas_int)ldc,
        rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
    if (st != rocblas_status_success) {
        std::fprintf(stderr, "elemental_amd: rocblas_gemm_ex failed (%d)\n", (int)st);
        return ELX_ERROR_HIP;
    }
    return ELX_SUCCESS;
}

elx_status_t library_dgemm(hipStream_t s, bool ta, bool tb, i64 m, i64 n, i64 k, double alpha,
                           const double* A, i64 lda, const double* B, i64 ldb, double beta,
                           double* C, i64 ldc, const double* coef_dev) {
    ELX_ARG(m < INT_MAX && n < INT_MAX && k < INT_MAX && lda < INT_MAX && ldb < INT_MAX &&
            ldc < INT_MAX);
    rocblas_handle h = nullptr;
    ELX_CHECK(rb_handle(&h));
    std::lock_guard<std::mutex> lk(g_rb_mu);
    if (rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, coef_dev ? rocblas_pointer_mode_device
                                             : rocblas_pointer_mode_host) != rocblas_status_success)
        return ELX_ERROR_HIP;
    const rocblas_status st = rocblas_dgemm(
        h, ta ? rocblas_operation_transpose : rocblas_operation_none,
        tb ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
        (rocblas_int)k, coef_dev ? coef_dev : &alpha, A, (rocblas_int)lda, B, (rocblas_int)ldb,
        coef_dev ? coef_dev + 1 : &beta, C, (rocblas_int)ldc);
    if (st != rocblas_status_success) {
        std::fprintf(stderr, "elemental_amd: rocblas_dgemm failed (%d)\n", (int)st);
        return ELX_ERROR_HIP;
    }
    return ELX_SUCCESS;
}

elx_status_t library_dsyrk(hipStream_t s, bool upper, bool trans, i64 n, i64 k, double alpha,
                           const double* A, i64 lda, double beta, double* C, i64 ldc,
                           const double* coef_dev) {
    ELX_ARG(n < INT_MAX && k < INT_MAX && lda < INT_MAX && ldc < INT_MAX);
    rocblas_handle h = nullptr;
    ELX_CHECK(rb_handle(&h));
    std::lock_guard<std::mutex> lk(g_rb_mu);
    if (rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, coef_dev ? rocblas_pointer_mode_device
                                             : rocblas_pointer_mode_host) != rocblas_status_success)
        return ELX_ERROR_HIP;
    const rocblas_status st = rocblas_dsyrk(
        h, upper ? rocblas_fill_upper : rocblas_fill_lower,
        trans ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)n,
        (rocblas_int)k, coef_dev ? coef_dev : &alpha, A, (rocblas_int)lda,
        coef_dev ? coef_dev + 1 : &beta, C, (rocblas_int)ldc);
    if (st != rocblas_status_success) {
        std::fprintf(stderr, "elemental_amd: rocblas_dsyrk failed (%d)\n", (int)st);
        return ELX_ERROR_HIP;
    }
    return ELX_SUCCESS;
}

elx_status_t library_strsm(hipStream_t s, bool upper, bool trans, bool unit, i64 m, i64 n,
                           const float* A, i64 lda, float* B, i64 ldb) {
    if (m <= 0 || n <= 0) return ELX_SUCCESS;
    ELX_ARG(m < INT_MAX && n < INT_MAX && lda < INT_MAX && ldb < INT_MAX);
    rocblas_handle h = nullptr;
    ELX_CHECK(rb_handle(&h));
    std::lock_guard<std::mutex> lk(g_rb_mu);
    if (rocblas_set_stream(h, s) != rocblas_status_success ||
        rocblas_set_pointer_mode(h, rocblas_pointer_mode_host) != rocblas_status_success)
        return ELX_ERROR_HIP;
    const float one = 1.0f;
    const rocblas_status st = rocblas_strsm(
        h, rocblas_side_left, upper ? rocblas_fill_upper : rocblas_fill_lower,
        trans ? rocblas_operation_transpose : rocblas_operation_none,
        unit ? rocblas_diagonal_unit : rocblas_diagonal_non_unit, (rocblas_int)m, (rocblas_int)n,
        &one, A, (rocblas_int)lda, B, (rocblas_int)ldb);
    if (st != rocblas_status_success) {
        std::fprintf(stderr, "elemental_amd: rocblas_strsm failed (%d)\n", (int)st);
        return ELX_ERROR_HIP;
    }
    return ELX_SUCCESS;
}

void library_gemm_release() {
    std::lock_guard<std::mutex> lk(g_rb_mu);
    for (auto& kv : g_rb) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(kv.first);
        (void)rocblas_destroy_handle(kv.second);
        (void)hipSetDevice(cur);
    }
    g_rb.clear();
}

}  // namespace elx
