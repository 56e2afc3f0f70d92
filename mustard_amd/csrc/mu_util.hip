// mu_util.hip -- generators, residuals, norms and small elementwise kernels
// (test matrices, refinement steps and verification on the device).
#include <algorithm>
#include <climits>
#include <cstring>

#include "mu_ops.hpp"
#include "mu_runtime.hpp"

namespace mu {

// ---------------------------------------------------------------------------
// generators, residuals and small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t sm64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void k_gen_uniform(double *M, size_t count, uint64_t seed, double lo, double hi)
{
    const uint64_t s0 = sm64(seed);
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x) {
        const uint64_t z = sm64(s0 + k);
        M[k] = lo + (hi - lo) * ((double)(z >> 11) * (1.0 / 9007199254740992.0));
    }
}

int gen_uniform_device(double *M, size_t count, uint64_t seed, double lo, double hi,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_uniform, dim3(2048), dim3(256), 0, s, M, count, seed, lo, hi);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// A <- A / div + shift I
__global__ void k_add_diag_scale(double *A, int n, size_t lda, double div, double shift)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        double *p = A + i + j * lda;
        *p = *p / div + (i == j ? shift : 0.0);
    }
}

int add_diag_scale(double *A, int n, int lda, double div, double shift, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_diag_scale, dim3(2048), dim3(256), 0, s, A, n, (size_t)lda, div,
                       shift);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_symmetrize_lower(double *A, int n, size_t lda)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        if (i < j)
            A[i + j * lda] = A[j + i * lda];
    }
}

int symmetrize_lower(double *A, int n, int lda, hipStream_t s)
{
    hipLaunchKernelGGL(k_symmetrize_lower, dim3(2048), dim3(256), 0, s, A, n, (size_t)lda);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_fill_int(int32_t *p, size_t count, int32_t v)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = v;
}

int fill_int(int32_t *p, size_t count, int32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_int, dim3(64), dim3(256), 0, s, p, count, v);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// F packed (LU or Cholesky) -> explicit L (unit lower for LU) and U (upper
// for LU; unused for Cholesky).
__global__ void k_split_lu(const double *F, int n, size_t ldf, double *L, double *U, size_t ld,
                           int chol)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double f = F[i + j * ldf];
        if (chol) {
            L[i + j * ld] = i >= j ? f : 0.0;
        } else {
            L[i + j * ld] = i > j ? f : (i == j ? 1.0 : 0.0);
            U[i + j * ld] = i <= j ? f : 0.0;
        }
    }
}

int split_lu(const double *F, int n, int ldf, double *L, double *U, int ld, int chol,
             hipStream_t s)
{
    hipLaunchKernelGGL(k_split_lu, dim3(4096), dim3(256), 0, s, F, n, (size_t)ldf, L, U,
                       (size_t)ld, chol);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_sub_frob2(const double *A, size_t lda, const double *B, size_t ldb, int m,
                            int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        const double d = B ? A[i + j * lda] - B[i + j * ldb] : A[i + j * lda];
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

int sub_frob2(const double *A, int lda, const double *B, int ldb, int m, int n, double *acc,
              hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_frob2, dim3(2048), dim3(256), 0, s, A, (size_t)lda, B, (size_t)ldb,
                       m, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_sum_abs(const double *A, size_t lda, int m, int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x)
        s += fabs(A[k % m + (k / m) * lda]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

int sum_abs(const double *A, int lda, int m, int n, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_abs, dim3(2048), dim3(256), 0, s, A, (size_t)lda, m, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Column-major copy: workgroup (row block of 1024, column chunk) walks its
// columns; each thread moves 4 rows of a column per step (no index division
// per element, 8 KB contiguous per wave and column).
__global__ void k_copy_matrix(const double *src, size_t lds, double *dst, size_t ldd, int m,
                              int n)
{
    const int i0 = blockIdx.x * 1024 + threadIdx.x;
    for (int j = blockIdx.y; j < n; j += gridDim.y) {
        const double *s = src + (size_t)j * lds;
        double *d = dst + (size_t)j * ldd;
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + 256 * q;
            v[q] = i < m ? s[i] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = i0 + 256 * q;
            if (i < m)
                d[i] = v[q];
        }
    }
}

int copy_matrix(const double *src, int lds, double *dst, int ldd, int m, int n, hipStream_t s)
{
    if (m <= 0 || n <= 0)
        return 0;
    const int gx = (m + 1023) / 1024;
    const int gy = std::min(n, std::max(1, 16384 / gx));
    hipLaunchKernelGGL(k_copy_matrix, dim3(gx, gy), dim3(256), 0, s, src, (size_t)lds, dst,
                       (size_t)ldd, m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Column maxima of |M| (m x n): out[j] = max_i |M(i, j)|.  Non-negative
// doubles order like their bit patterns, so an integer atomic max combines
// the per-thread maxima; out must be zeroed first.
// The maximum is taken on the bit patterns of |x|: they order like the
// values, and a NaN's pattern sorts above inf, so a NaN anywhere in the
// column makes its maximum NaN (fmax would drop it).
__global__ void k_colmax_abs(const double *M, size_t ld, int m, int n, unsigned long long *out)
{
    const int j = blockIdx.y;
    unsigned long long v = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(M[i + j * ld]));
        v = b > v ? b : v;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x])
            red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out + j, red[0]);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Residuals of a batch of packed problems, one workgroup per problem p of
// order n = ns[p]: A_p (n x n, leading dimension n) at A + a_off[p], x_p, b_p
// and r_p at v_off[p] of X, B and R.  r = b - A x in FP64, and the four
// maxima of the LAPACK stopping test, ||r||_inf, ||x||_inf, ||A||_inf and
// ||b||_inf, into norms[4p .. 4p+3].  A thread owns rows tid, tid + 256, ...
// and walks them along the columns (a wave reads 64 consecutive rows of a
// column); maxima on the bit patterns of non-negative values (a NaN's
// pattern sorts above inf, so NaN propagates).
__global__ __launch_bounds__(256) void k_batch_residual(const int *ns, const long long *a_off,
                                                        const long long *v_off, const double *A,
                                                        const double *X, const double *B,
                                                        double *R, double *norms)
{
    const int p = blockIdx.x, n = ns[p];
    const double *a = A + a_off[p];
    const double *x = X + v_off[p], *b = B + v_off[p];
    double *r = R + v_off[p];
    unsigned long long mr = 0, mx = 0, ma = 0, mb = 0;
    auto bits = [](double v) { return (unsigned long long)__double_as_longlong(fabs(v)); };
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double acc = 0.0, asum = 0.0;
        for (int j = 0; j < n; j++) {
            const double v = a[i + (size_t)j * n];
            acc = fma(v, x[j], acc);
            asum += fabs(v);
        }
        const double ri = b[i] - acc;
        r[i] = ri;
        mr = max(mr, bits(ri));
        mx = max(mx, bits(x[i]));
        ma = max(ma, bits(asum));
        mb = max(mb, bits(b[i]));
    }
    __shared__ unsigned long long red[4][256];
    red[0][threadIdx.x] = mr;
    red[1][threadIdx.x] = mx;
    red[2][threadIdx.x] = ma;
    red[3][threadIdx.x] = mb;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; q++)
                red[q][threadIdx.x] = max(red[q][threadIdx.x], red[q][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 4)
        norms[4 * p + threadIdx.x] = __longlong_as_double((long long)red[threadIdx.x][0]);
}

int batch_residual(int nprob, const int *ns, const long long *a_off, const long long *v_off,
                   const double *A, const double *X, const double *B, double *R, double *norms,
                   hipStream_t s)
{
    if (nprob <= 0)
        return 0;
    hipLaunchKernelGGL(k_batch_residual, dim3(nprob), dim3(256), 0, s, ns, a_off, v_off, A, X, B,
                       R, norms);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int colmax_abs(const double *M, int ld, int m, int n, double *out, hipStream_t s)
{
    if (n <= 0)
        return 0;
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
    hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, o, n, 0ull);
    const int gx = std::max(1, std::min(256, (m + 255) / 256));
    hipLaunchKernelGGL(k_colmax_abs, dim3(gx, n), dim3(256), 0, s, M, (size_t)ld, m, n, o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ||A||_inf = max_i sum_j |A(i, j)|.  Pass 1: thread (row i, column chunk c)
// sums |A(i, j)| over the chunk's INF_KC columns into W[c m + i] (a wave
// reads 64 consecutive rows of a column: coalesced; eight loads in flight).
// Pass 2: per row, the chunks in order, then the maximum (an atomic max of
// non-negative doubles, order-independent, on their bit patterns).
constexpr int INF_KC = 256;

__global__ void k_rowsum_partial(const double *A, size_t lda, int m, int n, double *W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j0 = blockIdx.y * INF_KC, j1 = min(n, j0 + INF_KC);
    if (i >= m)
        return;
    const double *a = A + i;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int j = j0;
    for (; j + 8 <= j1; j += 8)
#pragma unroll
        for (int u = 0; u < 8; u++)
            s[u] += fabs(a[(size_t)(j + u) * lda]);
    for (; j < j1; j++)
        s[0] += fabs(a[(size_t)j * lda]);
    W[(size_t)blockIdx.y * m + i] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

__global__ void k_rowsum_max(const double *W, int m, int nchunk, unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sum = 0.0;
    if (i < m)
        for (int c = 0; c < nchunk; c++)
            sum += W[(size_t)c * m + i];
    __shared__ double red[256];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

size_t inf_norm_work_doubles(int m, int n) { return (size_t)((n + INF_KC - 1) / INF_KC) * m; }

// k_rowsum_partial that also copies the chunk it reads to D (the copy of the
// input a factorization overwrites): one pass over A for both.  Groups of 8
// columns are loaded one group ahead of the stores (unconditional loads
// from clamped columns, the stores and sums masked), so a thread keeps a
// group in flight while it writes the previous one.
__global__ void k_copy_rowsum_partial(const double *A, size_t lda, double *D, size_t ldd, int m,
                                      int n, double *W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j0 = blockIdx.y * INF_KC, j1 = min(n, j0 + INF_KC);
    if (i >= m)
        return;
    const double *a = A + i;
    double *d = D + i;
    double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double v[2][8];
    auto ld = [&](double (&x)[8], int jj) {
#pragma unroll
        for (int u = 0; u < 8; u++)
            x[u] = a[(size_t)min(jj + u, j1 - 1) * lda];
    };
    auto put = [&](const double (&x)[8], int jj) {
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (jj + u < j1) {
                d[(size_t)(jj + u) * ldd] = x[u];
                s[u] += fabs(x[u]);
            }
    };
    ld(v[0], j0);
    for (int j = j0; j < j1; j += 16) {
        ld(v[1], j + 8);
        put(v[0], j);
        if (j + 16 < j1)
            ld(v[0], j + 16);
        put(v[1], j + 8);
    }
    W[(size_t)blockIdx.y * m + i] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

int copy_inf_norm(const double *A, int lda, double *D, int ldd, int m, int n, double *out,
                  double *work, hipStream_t s)
{
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, s, o, 1, 0ull);
    if (m <= 0 || n <= 0)
        return hipGetLastError() == hipSuccess ? 0 : -1;
    const int nchunk = (n + INF_KC - 1) / INF_KC;
    hipLaunchKernelGGL(k_copy_rowsum_partial, dim3((m + 255) / 256, nchunk), dim3(256), 0, s, A,
                       (size_t)lda, D, (size_t)ldd, m, n, work);
    hipLaunchKernelGGL(k_rowsum_max, dim3((m + 255) / 256), dim3(256), 0, s, work, m, nchunk, o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int inf_norm(const double *A, int lda, int m, int n, double *out, double *work, hipStream_t s)
{
    unsigned long long *o = reinterpret_cast<unsigned long long *>(out);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, s, o, 1, 0ull);
    if (m <= 0 || n <= 0)
        return hipGetLastError() == hipSuccess ? 0 : -1;
    const int nchunk = (n + INF_KC - 1) / INF_KC;
    hipLaunchKernelGGL(k_rowsum_partial, dim3((m + 255) / 256, nchunk), dim3(256), 0, s, A,
                       (size_t)lda, m, n, work);
    hipLaunchKernelGGL(k_rowsum_max, dim3((m + 255) / 256), dim3(256), 0, s, work, m, nchunk, o);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Column-major (ld lds) <-> stored by rows (ld n): the m x n right-hand
// sides of the reduced-precision solves, whose steps read 16 columns of a
// row as one 128-byte line.
__global__ void k_to_rows(const double *S, size_t lds, double *D, int m, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x)
        D[k] = S[(k / n) + (k % n) * lds];
}
__global__ void k_from_rows(const double *S, double *D, size_t ldd, int m, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x)
        D[(k / n) + (k % n) * ldd] = S[k];
}
int to_rows(const double *S, int lds, double *D, int m, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_rows, dim3(512), dim3(256), 0, s, S, (size_t)lds, D, m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
int from_rows(const double *S, double *D, int ldd, int m, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_from_rows, dim3(512), dim3(256), 0, s, S, D, (size_t)ldd, m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// X += D (m x n)
__global__ void k_add_matrix(double *X, size_t ldx, const double *D, size_t ldd, int m, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        X[i + j * ldx] += D[i + j * ldd];
    }
}

int add_matrix(double *X, int ldx, const double *D, int ldd, int m, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_add_matrix, dim3(1024), dim3(256), 0, s, X, (size_t)ldx, D, (size_t)ldd,
                       m, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Residual R = B - A X for up to 16 right-hand sides, split over K so the
// n x n read of A (the whole cost) runs on thousands of workgroups.  Pass 1:
// workgroup (256 rows, K chunk c) forms P_c = A[rows, chunk] X[chunk, :]:
// thread r reads row r of A down the chunk (a wave reads 64 consecutive rows
// of a column: 512 contiguous bytes) and keeps 16 sums, one per right-hand
// side, against X rows staged in LDS 128 at a time (read as broadcasts).
// The stores go to the partial buffer P[c] (column-major, leading dimension
// m).  Pass 2 sums the chunks in a fixed order, so the result does not
// depend on scheduling.
constexpr int RES_KC = 1024;  // K chunk of one workgroup
constexpr int RES_KS = 128;   // X rows staged in LDS at a time

__global__ __launch_bounds__(256) void k_residual_partial(const double *A, size_t lda,
                                                          const double *X, size_t ldx, int m,
                                                          int n, int nrhs, double *P)
{
    __shared__ double xs[RES_KS * 16];
    const int tid = threadIdx.x, c = blockIdx.y;
    const int r = blockIdx.x * 256 + tid;
    const bool rok = r < m;
    const double *a = A + min(r, m - 1);
    const int k0 = c * RES_KC, k1 = min(n, k0 + RES_KC);
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; j++)
        acc[j] = 0.0;
    for (int ks = k0; ks < k1; ks += RES_KS) {
        // stage X[ks .. ks + RES_KS)[0 .. 16) (zero past the chunk / nrhs)
#pragma unroll
        for (int q = 0; q < RES_KS * 16 / 256; q++) {
            const int idx = tid + 256 * q, kq = idx >> 4, j = idx & 15;
            const int kk = ks + kq;
            xs[idx] = (kk < k1 && j < nrhs) ? X[(size_t)kk + (size_t)j * ldx] : 0.0;
        }
        lds_barrier();
        const int kn = min(RES_KS, k1 - ks);
        for (int kq = 0; kq < kn; kq += 8) {
            double av[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int kk = ks + kq + u;
                av[u] = (rok && kq + u < kn) ? a[(size_t)min(kk, n - 1) * lda] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double *xr = xs + (kq + u) * 16;
#pragma unroll
                for (int j = 0; j < 16; j++)
                    acc[j] = fma(av[u], xr[j], acc[j]);
            }
        }
        lds_barrier();
    }
    if (rok) {
        double *p = P + (size_t)c * m * 16 + r;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < nrhs)
                p[(size_t)j * m] = acc[j];
    }
}

__global__ void k_residual_reduce(const double *B, size_t ldb, const double *P, int m, int nrhs,
                                  int nchunk, double *R, size_t ldr)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)m * nrhs;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % m, j = k / m;
        double sum = 0.0;
        for (int c = 0; c < nchunk; c++)
            sum += P[(size_t)c * m * 16 + i + j * m];
        R[i + j * ldr] = B[i + j * ldb] - sum;
    }
}

size_t residual_partial_doubles(int m, int n)
{
    return (size_t)((n + RES_KC - 1) / RES_KC) * m * 16;
}

int residual(const double *A, int lda, const double *X, int ldx, const double *B, int ldb,
             double *R, int ldr, int m, int n, int nrhs, double *P, hipStream_t s)
{
    if (m <= 0 || nrhs <= 0)
        return 0;
    const int nchunk = (n + RES_KC - 1) / RES_KC;
    for (int j0 = 0; j0 < nrhs; j0 += 16) {
        const int w = std::min(16, nrhs - j0);
        if (n > 0)
            hipLaunchKernelGGL(k_residual_partial, dim3((m + 255) / 256, nchunk), dim3(256), 0,
                               s, A, (size_t)lda, X + (size_t)j0 * ldx, (size_t)ldx, m, n, w, P);
        hipLaunchKernelGGL(k_residual_reduce, dim3(1024), dim3(256), 0, s,
                           B + (size_t)j0 * ldb, (size_t)ldb, P, m, w, n > 0 ? nchunk : 0,
                           R + (size_t)j0 * ldr, (size_t)ldr);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Conditioned test matrices: A = diag(s), then A <- H A H for each reflector.
__global__ void k_diag_geometric(double *A, size_t lda, int n, double kappa)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        A[i + j * lda] = i == j ? (n > 1 ? pow(kappa, -(double)i / (double)(n - 1)) : 1.0) : 0.0;
    }
}

// acc[0] += x . y
__global__ void k_dot(const double *x, const double *y, int n, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += x[i] * y[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(acc, red[0]);
}

// A <- H A H with H = I - tau v v^T, tau = 2 / v.v, given w = A v and
// c = v.w (sc[0] = v.v, sc[1] = v.w):
//   A - tau (v w^T + w v^T) + tau^2 c v v^T.
// The products are rounded separately (contraction into fused multiply-adds
// is off here), so A(i, j) and A(j, i) stay bit-for-bit equal.
__global__ void k_reflect2(double *A, size_t lda, int n, const double *v, const double *w,
                           const double *sc)
{
#pragma clang fp contract(off)
    const double tau = 2.0 / sc[0], c = sc[1];
    const double t2c = (tau * tau) * c;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double p1 = v[i] * w[j], p2 = w[i] * v[j];
        const double s1 = p1 + p2;
        const double s2 = v[i] * v[j];
        const double d1 = tau * s1, d2 = t2c * s2;
        A[i + j * lda] = (A[i + j * lda] - d1) + d2;
    }
}

// dst = tril(src) with the diagonal scaled by dscale and zeros above it
__global__ void k_tril_scaled(const double *src, size_t lds, double *dst, size_t ldd, int n,
                              double dscale)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        const double v = src[i + j * lds];
        dst[i + j * ldd] = i > j ? v : i == j ? dscale * v : 0.0;
    }
}

int tril_scaled(const double *src, int lds, double *dst, int ldd, int n, double dscale,
                hipStream_t s)
{
    hipLaunchKernelGGL(k_tril_scaled, dim3(4096), dim3(256), 0, s, src, (size_t)lds, dst,
                       (size_t)ldd, n, dscale);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// dst = (src + src^T) / 2 (out of place)
__global__ void k_sym_avg(const double *src, size_t lds, double *dst, size_t ldd, int n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)n * n;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t i = k % n, j = k / n;
        dst[i + j * ldd] = 0.5 * (src[i + j * lds] + src[j + i * lds]);
    }
}

int sym_avg(const double *src, int lds, double *dst, int ldd, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_sym_avg, dim3(4096), dim3(256), 0, s, src, (size_t)lds, dst, (size_t)ldd,
                       n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// t consecutive b x b identity tiles (tile k at D + k b^2, leading dimension b)
__global__ void k_identity_tiles(double *D, int b, int t)
{
    const size_t total = (size_t)b * b * t;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (size_t)gridDim.x * blockDim.x) {
        const size_t e = k % ((size_t)b * b);
        D[k] = (e % b) == (e / b) ? 1.0 : 0.0;
    }
}

int identity_tiles(double *D, int b, int t, hipStream_t s)
{
    hipLaunchKernelGGL(k_identity_tiles, dim3(2048), dim3(256), 0, s, D, b, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int gen_cond_start(double *A, int n, int lda, double kappa, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_geometric, dim3(4096), dim3(256), 0, s, A, (size_t)lda, n, kappa);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dot_acc(const double *x, const double *y, int n, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot, dim3(std::max(1, std::min(1024, (n + 255) / 256))), dim3(256), 0, s,
                       x, y, n, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int reflect2(double *A, int lda, int n, const double *v, const double *w, const double *sc,
             hipStream_t s)
{
    hipLaunchKernelGGL(k_reflect2, dim3(4096), dim3(256), 0, s, A, (size_t)lda, n, v, w, sc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mu

