// mu_ops.hpp -- gfx950 work-item bodies of the tile operations.
//
// One work item runs on one workgroup of 256 threads (4 wave64s).  The same
// bodies serve the per-unit grid kernels of the hipGraph mode and the
// persistent scheduler kernel.
//
// Arithmetic replaced (all column-major, leading dimension per operand):
//   GEMM  cublasGemmEx 64F  (mustard/lu_mustard.cu:511-522, cholesky_mustard.cu:376-387)
//   SYRK  cublasDsyrk LOWER (mustard/cholesky_mustard.cu:327-333)
//   TRSM  cublasDtrsm       (mustard/lu_mustard.cu:413-422, 459-468; cholesky_mustard.cu:281-290)
//   GETRF cusolverDnDgetrf, devIpiv = NULL (mustard/lu_mustard.cu:360-369)
//   POTRF cusolverDnDpotrf LOWER (mustard/cholesky_mustard.cu:238-246)
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mu_types.h"

namespace mu {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Operand pointers come from descriptors in memory, so the compiler cannot
// infer their address space; they always point to device global memory, and
// saying so turns flat accesses into global (vector memory) instructions.
#define MU_G __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ MU_G T *gp(T *p)
{
    return (MU_G T *)p;
}

// Workgroup barrier for LDS hand-offs only.  __syncthreads() is a
// workgroup-scope fence plus s_barrier, and the fence waits for every
// outstanding global load as well (vmcnt(0)), which would drain the
// prefetches the loops below keep in flight across their barriers.  No item
// body passes data between waves through global memory, so waiting for the
// LDS operations (lgkmcnt(0)) before the barrier is enough; the "memory"
// clobber keeps the compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// v_mfma_f64_16x16x4_f64.  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// it receives D[4*e + l/16][l%16] in element e (register e holds rows 4e..4e+3).
__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// GEMM / SYRK output block: C[i0:i0+64, j0:j0+64] = beta C + alpha op(A) op(B)
// ---------------------------------------------------------------------------
// Four waves in a 2 x 2 grid, each owning a 32 x 32 sub-block (2 x 2 MFMA
// tiles).  K runs in blocks of 8, double-buffered in LDS.  The kernel that
// runs GEMM-only batches needs 20 KB of LDS and at most 128 VGPRs per lane,
// so four workgroups share a CU and their waves hide each other's latency.
//
// LDS layouts.  An operand whose stored columns run along the output
// dimension ("i-contiguous": A, or B^T) is kept as S[k][i] with a pitch of
// PI doubles; one whose stored columns run along k ("k-contiguous": A^T, or B)
// is kept as S[i][k] with a pitch of PK doubles.
static_assert(GEMM_BM == 64 && GEMM_BN == 64 && GEMM_BK == 8, "GEMM loaders assume 64 x 64 x 8");
constexpr int PI = GEMM_BM + 16;
constexpr int PK = GEMM_BK + 1;
constexpr int LDS_I = GEMM_BK * PI;   // doubles per i-contiguous buffer
constexpr int LDS_K = GEMM_BM * PK;   // doubles per k-contiguous buffer

// The loaders below issue every load unconditionally, from an address clamped
// into the operand; the values that fall outside it are zeroed only when the
// registers are written to LDS.  With no branch around a load and no use of
// its result before the LDS store, the compiler's wait counting stays exact
// and a prefetch stays in flight until the half-step that stores it.  VEC
// (16-byte aligned operands, even leading dimensions and even m, n, k, set by
// the lowering) loads element pairs, which then never straddle an edge.

// i-contiguous operand: element (r, k) at p[r + k*ld], r < rows, k < kcols.
// Thread t covers rows 2(t%32), 2(t%32)+1 of column t/32.
template <bool VEC>
__device__ __forceinline__ void ld_icontig(double (&v)[2], const MU_G double *p, int ld, int rows,
                                           int kcols)
{
    const int t = threadIdx.x, r = (t & 31) * 2, c = t >> 5;
    const int cc = max(0, min(c, kcols - 1));
    if (VEC) {
        const int rr = max(0, min(r, rows - 2));
        const v2d x = *reinterpret_cast<const MU_G v2d *>(p + rr + (size_t)cc * ld);
        v[0] = x[0];
        v[1] = x[1];
    } else {
        v[0] = p[min(r, rows - 1) + (size_t)cc * ld];
        v[1] = p[min(r + 1, rows - 1) + (size_t)cc * ld];
    }
}

__device__ __forceinline__ void st_icontig(double *S, const double (&v)[2], int rows, int kcols)
{
    const int t = threadIdx.x, r = (t & 31) * 2, c = t >> 5;
    const bool cin = c < kcols;
    v2d x;
    x[0] = (cin && r < rows) ? v[0] : 0.0;
    x[1] = (cin && r + 1 < rows) ? v[1] : 0.0;
    *reinterpret_cast<v2d *>(S + c * PI + r) = x;
}

// k-contiguous operand: element (r, k) at p[k + r*ld].
// Thread t covers k = 2(t%4), 2(t%4)+1 of row t/4.
template <bool VEC>
__device__ __forceinline__ void ld_kcontig(double (&v)[2], const MU_G double *p, int ld, int rows,
                                           int kcols)
{
    const int t = threadIdx.x, k = (t & 3) * 2, r = t >> 2;
    const int rr = min(r, rows - 1);
    if (VEC) {
        const int kk = max(0, min(k, kcols - 2));
        const v2d x = *reinterpret_cast<const MU_G v2d *>(p + kk + (size_t)rr * ld);
        v[0] = x[0];
        v[1] = x[1];
    } else {
        v[0] = p[max(0, min(k, kcols - 1)) + (size_t)rr * ld];
        v[1] = p[max(0, min(k + 1, kcols - 1)) + (size_t)rr * ld];
    }
}

__device__ __forceinline__ void st_kcontig(double *S, const double (&v)[2], int rows, int kcols)
{
    const int t = threadIdx.x, k = (t & 3) * 2, r = t >> 2;
    const bool rin = r < rows;
    double *d = S + r * PK + k;
    d[0] = (rin && k < kcols) ? v[0] : 0.0;
    d[1] = (rin && k + 1 < kcols) ? v[1] : 0.0;
}

// op(A)(i, k) read from the LDS copy.
template <bool TA>
__device__ __forceinline__ double frag_a(const double *S, int i, int k)
{
    return TA ? S[i * PK + k] : S[k * PI + i];
}
// op(B)(k, j) read from the LDS copy.
template <bool TB>
__device__ __forceinline__ double frag_b(const double *S, int k, int j)
{
    return TB ? S[k * PI + j] : S[j * PK + k];
}

template <bool TA, bool TB, bool VEC>
__device__ __attribute__((always_inline)) inline void gemm_block(const UnitDesc &u, int bi, int bj, bool lower_only, double *smem)
{
    constexpr int BM = GEMM_BM, BN = GEMM_BN, BK = GEMM_BK;
    constexpr int SA = TA ? LDS_K : LDS_I;
    constexpr int SB = TB ? LDS_I : LDS_K;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;  // 2 x 2 waves of 32 x 32
    const int i0 = bi * BM, j0 = bj * BN;
    const int rows = min(BM, u.m - i0), cols = min(BN, u.n - j0);
    const int K = u.k;
    // op(A)(i, k): TA ? a[k + i*lda] : a[i + k*lda];  op(B)(k, j): TB ? b[j + k*ldb] : b[k + j*ldb]
    const MU_G double *A = TA ? gp(u.a) + (size_t)i0 * u.lda : gp(u.a) + i0;
    const MU_G double *B = TB ? gp(u.b) + j0 : gp(u.b) + (size_t)j0 * u.ldb;
    double *As0 = smem, *As1 = smem + SA, *Bs0 = smem + 2 * SA, *Bs1 = smem + 2 * SA + SB;

    v4d acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            acc[x][y] = v4d{0.0, 0.0, 0.0, 0.0};

    // Global loads run two k-blocks ahead of the MFMAs: block kb+2 is fetched
    // while block kb is computed and block kb+1 waits in registers for its
    // LDS buffer.
    double ra0[2], rb0[2], ra1[2], rb1[2];
    auto load = [&](double (&ra)[2], double (&rb)[2], int k0) {
        // a block past K is read (and zeroed) at k-offset 0, inside the operand
        const int kc = K - k0, ks = kc > 0 ? k0 : 0;
        if (TA)
            ld_kcontig<VEC>(ra, A + ks, u.lda, rows, kc);
        else
            ld_icontig<VEC>(ra, A + (size_t)ks * u.lda, u.lda, rows, kc);
        if (TB)
            ld_icontig<VEC>(rb, B + (size_t)ks * u.ldb, u.ldb, cols, kc);
        else
            ld_kcontig<VEC>(rb, B + ks, u.ldb, cols, kc);
    };
    // store the block loaded for k-offset k0 (zeroing what lies outside)
    auto store = [&](const double (&ra)[2], const double (&rb)[2], double *As, double *Bs,
                     int k0) {
        const int kc = K - k0;
        if (TA)
            st_kcontig(As, ra, rows, kc);
        else
            st_icontig(As, ra, rows, kc);
        if (TB)
            st_icontig(Bs, rb, cols, kc);
        else
            st_kcontig(Bs, rb, cols, kc);
    };
    const int li = lane & 15, lk = lane >> 4;
    auto compute = [&](const double *As, const double *Bs) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; x++)
                fa[x] = frag_a<TA>(As, wm * 32 + x * 16 + li, kk + lk);
#pragma unroll
            for (int y = 0; y < 2; y++)
                fb[y] = frag_b<TB>(Bs, kk + lk, wn * 32 + y * 16 + li);
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] = mfma_f64(fa[x], fb[y], acc[x][y]);
        }
    };

    // k-blocks in pairs, one per LDS buffer; with an odd count the last
    // block is left in buffer 0 by the last pair (or the prologue) and runs
    // alone.  Loads past the last block read clamped addresses, load zeros
    // and are never used.
    const int nkb = (K + BK - 1) / BK;
    if (K <= 0)
        goto epilogue;  // C = beta C: the operands are not read at all
    load(ra0, rb0, 0);
    store(ra0, rb0, As0, Bs0, 0);
    load(ra1, rb1, BK);
    lds_barrier();
    for (int kb = 0; kb + 1 < nkb; kb += 2) {
        load(ra0, rb0, (kb + 2) * BK);
        compute(As0, Bs0);
        store(ra1, rb1, As1, Bs1, (kb + 1) * BK);
        lds_barrier();
        load(ra1, rb1, (kb + 3) * BK);
        compute(As1, Bs1);
        store(ra0, rb0, As0, Bs0, (kb + 2) * BK);
        lds_barrier();
    }
    if (nkb & 1)
        compute(As0, Bs0);

epilogue:
    // Epilogue: all 16 C values of a lane are loaded before any is combined
    // and stored, so their latencies overlap.
    const double alpha = u.alpha, beta = u.beta;
    double cv[2][2][4];
    bool ok[2][2][4];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = wm * 32 + x * 16 + 4 * e + lk;
                const int c = wn * 32 + y * 16 + li;
                ok[x][y][e] = r < rows && c < cols && (!lower_only || i0 + r >= j0 + c);
                cv[x][y][e] = (ok[x][y][e] && beta != 0.0)
                                  ? gp(u.c)[(i0 + r) + (size_t)(j0 + c) * u.ldc]
                                  : 0.0;
            }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = wm * 32 + x * 16 + 4 * e + lk;
                const int c = wn * 32 + y * 16 + li;
                if (ok[x][y][e])
                    gp(u.c)[(i0 + r) + (size_t)(j0 + c) * u.ldc] =
                        beta * cv[x][y][e] + alpha * acc[x][y][e];
            }
}

// ---------------------------------------------------------------------------
// Reduced-precision GEMM / SYRK block (units flagged F_LOW_F32 or F_LOW_BF16):
// the FP64 operands are rounded to FP32 (or BF16, round-to-nearest-even) as
// they are staged into LDS, the products run on FP32-accumulating MFMA
// (v_mfma_f32_16x16x4f32 / v_mfma_f32_16x16x16bf16_1k), and the result is
// applied to the FP64 C: C = beta C + alpha (double) acc.  Same 64 x 64 block
// and 2 x 2 waves as the FP64 kernel; K runs in blocks of 16; both operands
// are kept k-contiguous in LDS (S[row][k]) so a bf16 lane reads its four k
// values with one 8-byte load.
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

// Row of element e of a lane's FP32 16 x 16 MFMA result (lane group lk).
__device__ __forceinline__ int f32_row(int e, int lk) { return 4 * lk + e; }

__device__ __forceinline__ unsigned short to_bf16(double x)
{
    // round to FP32, then to BF16 by round-to-nearest-even on the dropped bits
    unsigned u = __float_as_uint((float)x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <int P>
struct LowPrec;
typedef float v2f __attribute__((ext_vector_type(2)));
typedef unsigned short v2us __attribute__((ext_vector_type(2)));
template <>
struct LowPrec<1> {
    using T = float;
    using T2 = v2f;
    static constexpr int PK = 17;
    static __device__ __forceinline__ float cv(double x) { return (float)x; }
    static __device__ __forceinline__ float cv(float x) { return x; }
};
template <>
struct LowPrec<2> {
    using T = unsigned short;
    using T2 = v2us;
    static constexpr int PK = 20;
    static __device__ __forceinline__ unsigned short cv(double x) { return to_bf16(x); }
    static __device__ __forceinline__ unsigned short cv(unsigned short x) { return x; }
};

template <bool TA, bool TB, bool VEC, int P, bool SRC>
__device__ __attribute__((always_inline)) inline void gemm_block_low(const UnitDesc &u, int bi, int bj,
                                                                     bool lower_only, double *smem)
{
    using LP = LowPrec<P>;
    using T = typename LP::T;
    // operand storage: FP64, or (SRC) already rounded to T by an OP_CVT unit
    using S = typename std::conditional<SRC, T, double>::type;
    using S2 = typename std::conditional<SRC, typename LP::T2, v2d>::type;
    constexpr int BM = GEMM_BM, BK = 16, PKL = LP::PK, SBUF = BM * PKL;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    const int i0 = bi * BM, j0 = bj * GEMM_BN;
    const int rows = min(BM, u.m - i0), cols = min(GEMM_BN, u.n - j0);
    const int K = u.k;
    const MU_G S *A = TA ? (const MU_G S *)u.a + (size_t)i0 * u.lda : (const MU_G S *)u.a + i0;
    const MU_G S *B = TB ? (const MU_G S *)u.b + j0 : (const MU_G S *)u.b + (size_t)j0 * u.ldb;
    T *SL = reinterpret_cast<T *>(smem);
    T *As[2] = {SL, SL + 2 * SBUF};
    T *Bs[2] = {SL + SBUF, SL + 3 * SBUF};

    // One operand block (64 rows x 16 k) per thread: two element pairs.
    // i-contiguous storage (element (r,k) at p[r + k*ld]): pair = rows r, r+1.
    // k-contiguous storage (element (r,k) at p[k + r*ld]): pair = k, k+1.
    auto ld_op = [&](S (&v)[4], const MU_G S *p, int ld, int nrow, int kc, bool icontig) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int idx = tid + WG_THREADS * q;
            if (icontig) {
                const int r = (idx & 31) * 2, c = idx >> 5;
                const int cc = max(0, min(c, kc - 1));
                if (VEC) {
                    const S2 x = *reinterpret_cast<const MU_G S2 *>(p + max(0, min(r, nrow - 2)) +
                                                                   (size_t)cc * ld);
                    v[2 * q] = x[0];
                    v[2 * q + 1] = x[1];
                } else {
                    v[2 * q] = p[min(r, nrow - 1) + (size_t)cc * ld];
                    v[2 * q + 1] = p[min(r + 1, nrow - 1) + (size_t)cc * ld];
                }
            } else {
                const int k = (idx & 7) * 2, r = idx >> 3;
                const int rr = min(r, nrow - 1);
                if (VEC) {
                    const S2 x = *reinterpret_cast<const MU_G S2 *>(p + max(0, min(k, kc - 2)) +
                                                                   (size_t)rr * ld);
                    v[2 * q] = x[0];
                    v[2 * q + 1] = x[1];
                } else {
                    v[2 * q] = p[max(0, min(k, kc - 1)) + (size_t)rr * ld];
                    v[2 * q + 1] = p[max(0, min(k + 1, kc - 1)) + (size_t)rr * ld];
                }
            }
        }
    };
    const T zero = LP::cv(0.0);
    auto st_op = [&](T *D, const S (&v)[4], int nrow, int kc, bool icontig) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int idx = tid + WG_THREADS * q;
            if (icontig) {
                const int r = (idx & 31) * 2, c = idx >> 5;
                const bool cin = c < kc;
                D[r * PKL + c] = (cin && r < nrow) ? LP::cv(v[2 * q]) : zero;
                D[(r + 1) * PKL + c] = (cin && r + 1 < nrow) ? LP::cv(v[2 * q + 1]) : zero;
            } else {
                const int k = (idx & 7) * 2, r = idx >> 3;
                const bool rin = r < nrow;
                D[r * PKL + k] = (rin && k < kc) ? LP::cv(v[2 * q]) : zero;
                D[r * PKL + k + 1] = (rin && k + 1 < kc) ? LP::cv(v[2 * q + 1]) : zero;
            }
        }
    };
    S ra0[4], rb0[4], ra1[4], rb1[4];
    auto load = [&](S (&ra)[4], S (&rb)[4], int k0) {
        const int kc = K - k0, ks = kc > 0 ? k0 : 0;
        if (TA)
            ld_op(ra, A + ks, u.lda, rows, kc, false);
        else
            ld_op(ra, A + (size_t)ks * u.lda, u.lda, rows, kc, true);
        if (TB)
            ld_op(rb, B + (size_t)ks * u.ldb, u.ldb, cols, kc, true);
        else
            ld_op(rb, B + ks, u.ldb, cols, kc, false);
    };
    auto store = [&](const S (&ra)[4], const S (&rb)[4], int buf, int k0) {
        const int kc = K - k0;
        st_op(As[buf], ra, rows, kc, !TA);
        st_op(Bs[buf], rb, cols, kc, TB);
    };
    v4f acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
            acc[x][y] = v4f{0.f, 0.f, 0.f, 0.f};
    auto compute = [&](int buf) {
        const T *Sa = As[buf], *Sb = Bs[buf];
        if (P == 1) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                float fa[2], fb[2];
#pragma unroll
                for (int x = 0; x < 2; x++)
                    fa[x] = Sa[(wm * 32 + x * 16 + li) * PKL + kk + lk];
#pragma unroll
                for (int y = 0; y < 2; y++)
                    fb[y] = Sb[(wn * 32 + y * 16 + li) * PKL + kk + lk];
#pragma unroll
                for (int x = 0; x < 2; x++)
#pragma unroll
                    for (int y = 0; y < 2; y++)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[x], fb[y], acc[x][y], 0,
                                                                         0, 0);
            }
        } else {
            v4s fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; x++)
                fa[x] = *reinterpret_cast<const v4s *>(Sa + (wm * 32 + x * 16 + li) * PKL + 4 * lk);
#pragma unroll
            for (int y = 0; y < 2; y++)
                fb[y] = *reinterpret_cast<const v4s *>(Sb + (wn * 32 + y * 16 + li) * PKL + 4 * lk);
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
                for (int y = 0; y < 2; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[x], fb[y], acc[x][y], 0,
                                                                          0, 0);
        }
    };

    const int nk = ((K + BK - 1) / BK + 1) & ~1;
    if (K > 0) {
        load(ra0, rb0, 0);
        store(ra0, rb0, 0, 0);
        load(ra1, rb1, BK);
        lds_barrier();
        for (int kb = 0; kb < nk; kb += 2) {
            load(ra0, rb0, (kb + 2) * BK);
            compute(0);
            store(ra1, rb1, 1, (kb + 1) * BK);
            lds_barrier();
            load(ra1, rb1, (kb + 3) * BK);
            compute(1);
            store(ra0, rb0, 0, (kb + 2) * BK);
            lds_barrier();
        }
    }
    const double alpha = u.alpha, beta = u.beta;
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = wm * 32 + x * 16 + f32_row(e, lk);
                const int c = wn * 32 + y * 16 + li;
                if (r < rows && c < cols && (!lower_only || i0 + r >= j0 + c)) {
                    MU_G double *cp = gp(u.c) + (i0 + r) + (size_t)(j0 + c) * u.ldc;
                    const double old = beta != 0.0 ? *cp : 0.0;
                    *cp = beta * old + alpha * (double)acc[x][y][e];
                }
            }
}

// ---------------------------------------------------------------------------
// Reduced-precision GEMM / SYRK block with both operands k-contiguous in
// their storage precision (F_KVEC: op(A) = A^T of a panel the conversion
// stored transposed, op(B) = B), C = beta C + alpha op(A) op(B) on a
// 128 x 128 output block (units flagged F_BLK128, made for the level-batched
// graph).  Each of the 4 waves owns a 64 x 64 quarter, 4 x 4 MFMA tiles, so
// every LDS fragment feeds four MFMAs.  BF16 runs on gfx950's
// v_mfma_f32_16x16x32_bf16 (8 k-values per lane, lane group lk holding
// k = 8 lk .. 8 lk + 7 of A's row and B's column), FP32 on
// v_mfma_f32_16x16x4f32; FP32 accumulation, the result applied to the FP64
// C.  The k-slices are long (BF16: 64, FP32: 32) and every operand row
// reaches LDS as whole 16-byte vectors (LDS rows padded by 16 bytes),
// global -> registers -> LDS one slice ahead of the MFMAs.
// ---------------------------------------------------------------------------
template <int P>
__device__ __attribute__((always_inline)) inline void gemm_block_lowk128(const UnitDesc &u,
                                                                         int bi, int bj,
                                                                         bool lower_only,
                                                                         double *smem)
{
    using T = typename LowPrec<P>::T;
    typedef unsigned int v4u __attribute__((ext_vector_type(4)));
    typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
    constexpr int BM = 128;
    constexpr int ES = sizeof(T), VE = 16 / ES;        // elements per 16 bytes
    constexpr int BK = P == 2 ? 64 : 32, PK = BK + VE;  // k-slice, LDS row pitch (elements)
    constexpr int CPR = BK / VE, NL = BM * BK / VE / WG_THREADS;  // chunks per row, loads/thread
    static_assert(NL * WG_THREADS * VE == BM * BK, "loader split");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, li = lane & 15, lk = lane >> 4;
    const int i0 = bi * BM, j0 = bj * BM;
    const int rows = min(BM, u.m - i0), cols = min(BM, u.n - j0);
    const int K = u.k;
    const MU_G T *A = (const MU_G T *)u.a + (size_t)i0 * u.lda;  // op(A) row r at A + r lda
    const MU_G T *B = (const MU_G T *)u.b + (size_t)j0 * u.ldb;  // op(B) column c at B + c ldb
    T *SL = reinterpret_cast<T *>(smem);
    T *As[2] = {SL, SL + 2 * BM * PK};
    T *Bs[2] = {SL + BM * PK, SL + 3 * BM * PK};

    v4u ra0[NL], rb0[NL], ra1[NL], rb1[NL];
    // chunk ch of row r of the k-slice at k0; a chunk past K reads the
    // slice's last full chunk (or the row's first) and is zeroed when stored
    auto load = [&](v4u (&xa)[NL], v4u (&xb)[NL], int k0) {
        const int kc = K - k0, ks = kc > 0 ? k0 : 0;
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int idx = tid + WG_THREADS * q, r = idx / CPR, ch = idx % CPR;
            const int kk = min(ch * VE, max(kc - VE, 0));
            xa[q] = *(const MU_G v4u *)(A + (size_t)min(r, rows - 1) * u.lda + ks + kk);
            xb[q] = *(const MU_G v4u *)(B + (size_t)min(r, cols - 1) * u.ldb + ks + kk);
        }
    };
    auto store = [&](const v4u (&xa)[NL], const v4u (&xb)[NL], int buf, int k0) {
        const int kc = K - k0;
        const v4u zero = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int q = 0; q < NL; q++) {
            const int idx = tid + WG_THREADS * q, r = idx / CPR, ch = idx % CPR;
            const bool kin = ch * VE < kc;
            *(v4u *)(As[buf] + r * PK + ch * VE) = (kin && r < rows) ? xa[q] : zero;
            *(v4u *)(Bs[buf] + r * PK + ch * VE) = (kin && r < cols) ? xb[q] : zero;
        }
    };
    v4f acc[4][4];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
            acc[x][y] = v4f{0.f, 0.f, 0.f, 0.f};
    auto compute = [&](int buf) {
        const T *Sa = As[buf] + (wm * 64 + li) * PK, *Sb = Bs[buf] + (wn * 64 + li) * PK;
        if constexpr (P == 2) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 32) {
                v8bf fa[4], fb[4];
#pragma unroll
                for (int x = 0; x < 4; x++)
                    fa[x] = __builtin_bit_cast(
                        v8bf, *reinterpret_cast<const v4u *>(Sa + x * 16 * PK + kk + 8 * lk));
#pragma unroll
                for (int y = 0; y < 4; y++)
                    fb[y] = __builtin_bit_cast(
                        v8bf, *reinterpret_cast<const v4u *>(Sb + y * 16 * PK + kk + 8 * lk));
#pragma unroll
                for (int x = 0; x < 4; x++)
#pragma unroll
                    for (int y = 0; y < 4; y++)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[x], fb[y], acc[x][y],
                                                                            0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                float fa[4], fb[4];
#pragma unroll
                for (int x = 0; x < 4; x++)
                    fa[x] = Sa[x * 16 * PK + kk + lk];
#pragma unroll
                for (int y = 0; y < 4; y++)
                    fb[y] = Sb[y * 16 * PK + kk + lk];
#pragma unroll
                for (int x = 0; x < 4; x++)
#pragma unroll
                    for (int y = 0; y < 4; y++)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[x], fb[y], acc[x][y], 0, 0, 0);
            }
        }
    };
    const int nkb = (K + BK - 1) / BK;
    if (K > 0) {
        load(ra0, rb0, 0);
        store(ra0, rb0, 0, 0);
        load(ra1, rb1, BK);
        lds_barrier();
        for (int kb = 0; kb + 1 < nkb; kb += 2) {
            load(ra0, rb0, (kb + 2) * BK);
            compute(0);
            store(ra1, rb1, 1, (kb + 1) * BK);
            lds_barrier();
            load(ra1, rb1, (kb + 3) * BK);
            compute(1);
            store(ra0, rb0, 0, (kb + 2) * BK);
            lds_barrier();
        }
        if (nkb & 1)
            compute(0);
    }
    const double alpha = u.alpha, beta = u.beta;
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = wm * 64 + x * 16 + f32_row(e, lk);
                const int c = wn * 64 + y * 16 + li;
                if (r < rows && c < cols && (!lower_only || i0 + r >= j0 + c)) {
                    MU_G double *cp = gp(u.c) + (i0 + r) + (size_t)(j0 + c) * u.ldc;
                    const double old = beta != 0.0 ? *cp : 0.0;
                    *cp = beta * old + alpha * (double)acc[x][y][e];
                }
            }
}
// LDS of gemm_block_lowk128: 2 buffers x (128 + 128) rows x (BK + VE) elements = 72 KB
constexpr int LDS_GEMM_LOWK128 = 2 * 256 * (64 + 8) * 2 / 8;

// ---------------------------------------------------------------------------
// Panel triangular solve.  Solves T Y = Z in place for one panel of up to 16
// right-hand sides, T lower triangular of order n <= TRSM_NMAX:
//   left-lower-unit   X = L^-1 B  : T = L,   Z = B (columns)
//   right-upper       X = B U^-1  : T = U^T, Z = B^T (rows of B)
//   right-lower-trans X = B L^-T  : T = L,   Z = B^T (rows of B)
// With REV the row order of T and Z is reversed, which turns an upper
// triangular solve (back substitution) into the same lower-triangular form.
// Z is scaled by alpha on load.  The panel lives in LDS; each 16-row step solves its diagonal block with one
// thread per right-hand side and updates the rows below with FP64 MFMA.
// ---------------------------------------------------------------------------
template <bool TT, bool ZR, bool UNIT, bool REV, int NMAX = TRSM_NMAX>
__device__ __attribute__((always_inline)) inline void trsm_panel(const UnitDesc &u, int panel, double *smem)
{
    // NMAX: largest triangle the caller's LDS holds (TRSM_NMAX in the generic
    // kernel, SMALL_MAX in k_batch_trsm_small)
    const int n = u.m, p0 = panel * TRSM_W, w = min(TRSM_W, u.n - p0);
    const int np = (n + 15) & ~15;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *Y = smem;              // [np][16]
    double *Td = smem + NMAX * 16; // [16][17]
    const MU_G double *t = gp(u.a);
    const size_t ldt = u.lda, ldz = u.ldc;
    MU_G double *z = gp(u.c);
    const double alpha = u.alpha;
    // 32-bit offsets with 24-bit multiplies (full rate, unlike 64-bit
    // products): the lowering guarantees (n + 1) * ldt < 2^31 for triangles
    const unsigned ldt32 = (unsigned)ldt;
    auto tv = [&](int r, int c) -> double {
        if (REV) {
            r = n - 1 - r;
            c = n - 1 - c;
        }
        const unsigned off = TT ? (unsigned)c + __umul24((unsigned)r, ldt32)
                                : (unsigned)r + __umul24((unsigned)c, ldt32);
        return t[off];
    };
    auto zp = [&](int r, int c) -> MU_G double * {
        if (REV)
            r = n - 1 - r;
        return ZR ? z + (p0 + c) + r * ldz : z + r + (p0 + c) * ldz;
    };

    for (int idx = tid; idx < np * 16; idx += WG_THREADS) {
        int r, c;
        if (ZR) {
            c = idx & 15;
            r = idx >> 4;
        } else {
            r = idx % np;
            c = idx / np;
        }
        Y[r * 16 + c] = (r < n && c < w) ? alpha * *zp(r, c) : 0.0;
    }
    lds_barrier();

    // Each 16-row step jb: stage the diagonal block (reciprocals of its
    // diagonal in place of the entries, so the solve multiplies), solve it
    // with one thread per right-hand side, then update the rows below,
    // Y[rb] -= T[rb, jb] Y[jb], with FP64 MFMA.  The triangle fragments of
    // step jb+16 are loaded during step jb, row block by row block right
    // after that block's MFMAs, so they are consumed in the order they were
    // issued and at most 64 loads per lane are in flight (the hardware's
    // in-order vector-memory counter tracks 63).
    const int li = lane & 15, lk = lane >> 4;
    const int r16 = tid >> 4, c16 = tid & 15;
    constexpr int MAXB = NMAX / 64 > 0 ? NMAX / 64 : 1;
    // Loads are unconditional, from addresses clamped into the triangle;
    // values outside it are zeroed where they are used (see the header of
    // the GEMM loaders for why no load sits behind a branch).
    auto load_diag = [&](int jb) -> double {
        return tv(min(jb + r16, n - 1), min(jb + c16, n - 1));
    };
    auto load_block = [&](double (&f)[4], int jb, int i) {
        const int rb = jb + 16 + 16 * wave + 64 * i;
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            f[kk] = tv(min(rb + li, n - 1), min(jb + 4 * kk + lk, n - 1));
    };
    auto step = [&](int jb, double dv, const double (&fa)[MAXB][4], double &dn,
                    double (&fn)[MAXB][4]) {
        dv = (jb + r16 < n && c16 <= r16) ? dv : 0.0;
        if (!UNIT && c16 == r16 && jb + r16 < n)
            dv = 1.0 / dv;
        Td[r16 * 17 + c16] = dv;
        lds_barrier();
        if (tid < 16) {
            const int nb = min(16, n - jb);
            double y[16];
#pragma unroll
            for (int r = 0; r < 16; r++)
                y[r] = Y[(jb + r) * 16 + tid];
#pragma unroll
            for (int r = 0; r < 16; r++) {
                if (r < nb) {
                    double s = y[r];
#pragma unroll
                    for (int q = 0; q < r; q++)
                        s -= Td[r * 17 + q] * y[q];
                    if (!UNIT)
                        s *= Td[r * 17 + r];
                    y[r] = s;
                }
            }
#pragma unroll
            for (int r = 0; r < 16; r++)
                if (r < nb)
                    Y[(jb + r) * 16 + tid] = y[r];
        }
        lds_barrier();
        dn = load_diag(jb + 16);
        double fb[4];
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            fb[kk] = Y[(jb + 4 * kk + lk) * 16 + li];
#pragma unroll
        for (int i = 0; i < MAXB; i++) {
            const int rb = jb + 16 + 16 * wave + 64 * i;
            if (rb < np) {
                v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const bool in = rb + li < n && jb + 4 * kk + lk < n;
                    acc = mfma_f64(in ? fa[i][kk] : 0.0, fb[kk], acc);
                }
#pragma unroll
                for (int e = 0; e < 4; e++)
                    Y[(rb + 4 * e + lk) * 16 + li] -= acc[e];
            }
            load_block(fn[i], jb + 16, i);
        }
        lds_barrier();
    };
    double fa0[MAXB][4], fa1[MAXB][4];
    double d0 = load_diag(0), d1 = 0.0;
#pragma unroll
    for (int i = 0; i < MAXB; i++)
        load_block(fa0[i], 0, i);
    for (int jb = 0; jb < n; jb += 32) {
        step(jb, d0, fa0, d1, fa1);
        if (jb + 16 >= n)
            break;
        step(jb + 16, d1, fa1, d0, fa0);
    }

    for (int idx = tid; idx < np * 16; idx += WG_THREADS) {
        int r, c;
        if (ZR) {
            c = idx & 15;
            r = idx >> 4;
        } else {
            r = idx % np;
            c = idx / np;
        }
        if (r < n && c < w)
            *zp(r, c) = Y[r * 16 + c];
    }
}

// ---------------------------------------------------------------------------
// Inverted diagonal blocks.  For a triangle T of order n > SMALL_MAX (read in
// the canonical lower form the panel solve uses: flags TT, UNIT, REV as in
// trsm_panel), item `it` inverts diagonal blocks 16*it .. 16*it+15 of 16 x 16
// and stores inv(T_jj) row-major at c + 256 j.  Rows and columns of the last
// block past n are the identity.  Thread t: block t/16, column t%16, by
// forward substitution against the unit vector (divisions, as LAPACK's
// dtrti2 does; this runs once per triangle).
// ---------------------------------------------------------------------------
template <bool TT, bool UNIT, bool REV>
__device__ __attribute__((always_inline)) inline void trinv_blocks(const UnitDesc &u, int it,
                                                                   double *smem)
{
    const int n = u.m, tid = threadIdx.x, lb = tid >> 4, c = tid & 15;
    const int j = it * TRINV_BLOCKS + lb, j0 = 16 * j;
    const MU_G double *t = gp(u.a);
    const size_t ldt = u.lda;
    const unsigned ldt32 = (unsigned)ldt;
    auto tv = [&](int r, int q) -> double {
        if (REV) {
            r = n - 1 - r;
            q = n - 1 - q;
        }
        const unsigned off = TT ? (unsigned)q + __umul24((unsigned)r, ldt32)
                                : (unsigned)r + __umul24((unsigned)q, ldt32);
        return t[off];
    };
    double *D = smem + lb * 16 * 17;
    const bool live = j0 < n;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const bool in = live && j0 + r < n && j0 + c < n;
        double v;
        if (r == c)
            v = (UNIT || !in) ? 1.0 : tv(j0 + r, j0 + c);
        else
            v = (r > c && in) ? tv(j0 + r, j0 + c) : 0.0;
        D[r * 17 + c] = v;
    }
    lds_barrier();
    double x[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        double s = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < r; q++)
            s -= D[r * 17 + q] * x[q];
        x[r] = UNIT ? s : s / D[r * 17 + r];
    }
    if (live) {
        MU_G double *w = gp(u.c) + (size_t)j * 256;
#pragma unroll
        for (int r = 0; r < 16; r++)
            w[r * 16 + c] = x[r];
    }
}

// ---------------------------------------------------------------------------
// Panel triangular solve with inverted diagonal blocks (triangles of order
// n > SMALL_MAX; the TRINV unit before it left inv(T_jj) at u.b).  Same
// problem and flags as trsm_panel.  Step jb: every wave forms the solved
// block X_jb = inv(T_jj) Y_jb with four MFMAs -- the result's register layout
// is exactly the B operand of the update MFMAs, so it never goes back
// through LDS -- wave 0 stores it to Z, and each wave updates its row blocks
// below, Y[rb] -= T[rb, jb] X_jb.  One barrier per step.  The next step's
// inverse block and triangle fragments are loaded during the current step
// (the latter row block by row block, right after that block's MFMAs).
// ---------------------------------------------------------------------------
template <bool TT, bool ZR, bool UNIT, bool REV, int NW = 4, int NMAX = TRSM_NMAX>
__device__ __attribute__((always_inline)) inline void trsm_panel_inv(const UnitDesc &u, int panel,
                                                                     double *smem)
{
    // NW waves per workgroup (4 in the generic kernel, 8 in k_batch_trsm:
    // two waves per SIMD hide each other's MFMA and load latency); NMAX the
    // largest triangle the caller's LDS holds
    static_assert(NMAX % (64 * NW) == 0, "row blocks come in groups of four per wave");
    const int n = u.m, p0 = panel * TRSM_W, w = min(TRSM_W, u.n - p0);
    const int np = (n + 15) & ~15, nblk = np >> 4;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    double *Y = smem;  // [np][16]
    const MU_G double *t = gp(u.a);
    const MU_G double *winv = gp(u.b);
    const size_t ldt = u.lda, ldz = u.ldc;
    MU_G double *z = gp(u.c);
    const double alpha = u.alpha;
    // 32-bit offsets with 24-bit multiplies (full rate, unlike 64-bit
    // products): the lowering guarantees (n + 1) * ldt < 2^31 for triangles
    const unsigned ldt32 = (unsigned)ldt;
    auto tv = [&](int r, int c) -> double {
        if (REV) {
            r = n - 1 - r;
            c = n - 1 - c;
        }
        const unsigned off = TT ? (unsigned)c + __umul24((unsigned)r, ldt32)
                                : (unsigned)r + __umul24((unsigned)c, ldt32);
        return t[off];
    };
    auto zp = [&](int r, int c) -> MU_G double * {
        if (REV)
            r = n - 1 - r;
        return ZR ? z + (p0 + c) + r * ldz : z + r + (p0 + c) * ldz;
    };

    for (int idx = tid; idx < np * 16; idx += 64 * NW) {
        int r, c;
        if (ZR) {
            c = idx & 15;
            r = idx >> 4;
        } else {
            r = idx % np;
            c = idx / np;
        }
        Y[r * 16 + c] = (r < n && c < w) ? alpha * *zp(r, c) : 0.0;
    }
    lds_barrier();

    constexpr int MAXB = NMAX / (16 * NW);
    auto load_inv = [&](double (&wv)[4], int jb) {
        const MU_G double *b = winv + (size_t)min(jb >> 4, nblk - 1) * 256 + li * 16 + lk;
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            wv[kk] = b[4 * kk];
    };
    auto load_block = [&](double (&f)[4], int jb, int i) {
        const int rb = jb + 16 + 16 * wave + 16 * NW * i;
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            f[kk] = tv(min(rb + li, n - 1), min(jb + 4 * kk + lk, n - 1));
    };
    auto step = [&](int jb, const double (&wv)[4], const double (&fa)[MAXB][4], double (&wn)[4],
                    double (&fn)[MAXB][4]) {
        double yb[4];
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            yb[kk] = Y[(jb + 4 * kk + lk) * 16 + li];
        v4d xs = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
            xs = mfma_f64(wv[kk], yb[kk], xs);
        // xs[e] = X[jb + 4e + lk][li]
        if (wave == 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = jb + 4 * e + lk;
                if (r < n && li < w)
                    *zp(r, li) = xs[e];
            }
        }
        load_inv(wn, jb + 16);
        // Row blocks in groups of four under one (wave-uniform) test, so the
        // four MFMA chains of a group interleave: the accumulators start as
        // the Y rows and take -T[rb, jb] X_jb.  A block of a live group that
        // lies past the panel reads clamped rows and is not written back.
#pragma unroll
        for (int g = 0; g < MAXB / 4; g++) {
            const int rb0 = jb + 16 + 16 * wave + 64 * NW * g;
            if (rb0 < np) {
                v4d acc[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int rr = min(rb0 + 16 * NW * q, np - 16);
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        acc[q][e] = Y[(rr + 4 * e + lk) * 16 + li];
                }
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const bool in = rb0 + 16 * NW * q + li < n && jb + 4 * kk + lk < n;
                        acc[q] = mfma_f64(in ? -fa[4 * g + q][kk] : 0.0, xs[kk], acc[q]);
                    }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int rb = rb0 + 16 * NW * q;
                    if (rb < np)
#pragma unroll
                        for (int e = 0; e < 4; e++)
                            Y[(rb + 4 * e + lk) * 16 + li] = acc[q][e];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++)
                load_block(fn[4 * g + q], jb + 16, 4 * g + q);
        }
        lds_barrier();
    };
    double fa0[MAXB][4], fa1[MAXB][4], w0[4], w1[4];
    load_inv(w0, 0);
#pragma unroll
    for (int i = 0; i < MAXB; i++)
        load_block(fa0[i], 0, i);
    for (int jb = 0; jb < n; jb += 32) {
        step(jb, w0, fa0, w1, fa1);
        if (jb + 16 >= n)
            break;
        step(jb + 16, w1, fa1, w0, fa0);
    }
}

// ---------------------------------------------------------------------------
// Small factorizations of an n x n tile (n <= SMALL_MAX) held in LDS.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void report_info(const UnitDesc &u, int col)
{
    if (u.info)
        __hip_atomic_fetch_min(u.info, u.info_base + col + 1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Blocked diagonal factorizations of an n x n tile (n <= SMALL_MAX), right-
// looking over panels of 16 columns: no-pivot LU (GETRF) and lower Cholesky
// (POTRF).  The tile sits in registers in the FP64 MFMA result layout --
// 16 x 16 blocks, block-column j held by wave j % NW, so with 4 waves a
// lane holds 2 block-columns x 8 blocks x 4 doubles and with 8 waves one
// (element e of lane l of a block: row 4e + l/16, column l%16).  Panel p:
//   * its owner wave (p % NW) copies block-column p to LDS and factors the
//     panel alone, one lane per row (rows l and l + 64), the pivot row (LU)
//     or the column's entries of the diagonal block (Cholesky) broadcast by
//     readlane -- no barrier inside the panel; for LU it also forms the
//     inverse of the unit lower 16 x 16 diagonal block;
//   * after one barrier every wave updates its block-columns right of p on
//     MFMA: LU forms U12 = inv(L11) A12 (four MFMAs; the result layout is
//     the B operand of the next MFMAs) and A22 -= L21 U12; Cholesky
//     A22 -= L21 L21^T on the lower blocks; L21 is read from the panel in
//     LDS as the A operand;
//   * a second barrier frees the panel buffer.
// Rows and columns past n are the identity, so the padded factorization is
// the n x n one; only the n x n part (Cholesky: its lower triangle) is
// stored.  Zero pivots / non-positive minors report LAPACK-style info and
// the factorization continues (the factors then hold inf / nan).
// ---------------------------------------------------------------------------
constexpr int BF_LD = 17;                          // LDS row stride of the panel
constexpr int LDS_SMALL_PANEL = SMALL_MAX * BF_LD;  // 128 x 16 panel (17 KB)
// + inv(L11) for LU, + 1 / diag(L11) for Cholesky
constexpr int LDS_SMALL_BLK = LDS_SMALL_PANEL + 16 * BF_LD + 16;

// 1/x and 1/sqrt(x) in FP64 from the hardware estimates (v_rcp_f64,
// v_rsq_f64) and two Newton steps each (r <- r + r (1 - x r), r <- r + r
// (1/2 - x r^2 / 2)), within an ulp or two of the rounded results: the
// panel factorizations' per-column reciprocal and square root, which as IEEE
// division / sqrt sequences were most of a panel's instructions.  A zero
// pivot gives NaN (not the division's inf) and a negative one NaN: either way
// the factors carry non-finite values, and the column is reported as info.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x)
{
    const double h = 0.5 * x;
    double r = __builtin_amdgcn_rsq(x);
    r = fma(r, fma(-h * r, r, 0.5), r);
    return fma(r, fma(-h * r, r, 0.5), r);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <bool CHOL, int NW = 4>
__device__ __attribute__((always_inline)) inline void factor_small_blk(const UnitDesc &u,
                                                                       double *smem)
{
    // NW waves (4 or 8); block-column j is held by wave j % NW
    static_assert(NW == 4 || NW == 8, "4 or 8 waves");
    constexpr int JJ = 8 / NW;
    const int n = u.m, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int q = lane >> 4, cl = lane & 15;
    const int nb = (n + 15) >> 4;
    MU_G double *a = gp(u.c);
    const size_t lda = u.ldc;
    double *P = smem;                     // panel: row r (0..127), column c at P[r * 17 + c]
    double *Li = smem + LDS_SMALL_PANEL;  // LU: inv(L11), row-major, stride 17
    double *Rc = Li + 16 * BF_LD;         // Cholesky: 1 / diag(L11) of the panel
    // the diagonal factorization is the serial chain of the whole DAG: its
    // waves issue ahead of the GEMM waves sharing their SIMDs
    __builtin_amdgcn_s_setprio(3);

    v4d R[JJ][8];  // block (i, w + NW jj)
#pragma unroll
    for (int jj = 0; jj < JJ; jj++)
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = 16 * i + 4 * e + q, c = 16 * (w + NW * jj) + cl;
                double v = r == c ? 1.0 : 0.0;
                if (r < n && c < n && (!CHOL || r >= c))
                    v = a[r + (size_t)c * lda];
                R[jj][i][e] = v;
            }

    for (int p = 0; p < nb; p++) {
        const int jp = p / NW;  // the owner's register column of block-column p
        if (w == p % NW) {
            // ---- panel: block-column p, rows 16p .. 127, to LDS ----
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (i >= p) {
                    const v4d blk = (JJ > 1 && jp) ? R[JJ - 1][i] : R[0][i];
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        P[(16 * i + 4 * e + q) * BF_LD + cl] = blk[e];
                }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            const int p0 = 16 * p;
            // sweep 1: rows p0 .. p0 + 63, lane l on row p0 + l; the pivot
            // row of column c is lane c's
            {
                const int r = p0 + lane;
                const bool in = r < SMALL_MAX;
                // lanes past the tile read the last row (their results are
                // dropped): no branch around the loads
                const double *src = P + (in ? r : SMALL_MAX - 1) * BF_LD;
                double x[16];
#pragma unroll
                for (int c = 0; c < 16; c++)
                    x[c] = src[c];
                // the first failing column of the panel, uniform; reported
                // once after the sweep instead of a branch in every column
                int bad = 16;
#pragma unroll
                for (int c = 0; c < 16; c++) {
                    const double piv = readlane_d(x[c], c);
                    if (!CHOL) {
                        bad = min(bad, piv == 0.0 ? c : 16);
                        const double l = lane > c ? x[c] * rcp_nr(piv) : 0.0;
                        if (lane > c)
                            x[c] = l;
                        // lane c's own row keeps its values (multiplier 0)
#pragma unroll
                        for (int k = c + 1; k < 16; k++)
                            x[k] = fma(-l, readlane_d(x[k], c), x[k]);
                    } else {
                        bad = min(bad, piv > 0.0 ? 16 : c);
                        const double rd = rsq_nr(piv);
                        Rc[c] = rd;
                        // every lane, no masks: lane c gets piv * rd, the
                        // lanes above it only change the diagonal block's
                        // upper part, which nothing reads
                        x[c] *= rd;
#pragma unroll
                        for (int k = c + 1; k < 16; k++)  // L[p0 + k][c] is lane k's
                            x[k] = fma(-x[c], readlane_d(x[c], k), x[k]);
                    }
                }
                if (bad < 16 && lane == 0)
                    report_info(u, p0 + bad);
#pragma unroll
                for (int c = 0; c < 16; c++)
                    if (in)
                        P[r * BF_LD + c] = x[c];
            }
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            // sweep 2: rows p0 + 64 .. 127 against the finished diagonal
            // block in LDS (LU: L21 = A21 inv(U11); Cholesky: A21 inv(L11^T))
            if (p0 + 64 < SMALL_MAX) {
                const int r = p0 + 64 + lane;
                const bool in = r < SMALL_MAX;
                const double *src = P + (in ? r : SMALL_MAX - 1) * BF_LD;
                double y[16];
#pragma unroll
                for (int c = 0; c < 16; c++)
                    y[c] = src[c];
#pragma unroll
                for (int c = 0; c < 16; c++) {
                    const double *dr = P + (p0 + c) * BF_LD;  // LU: row c of U11
                    if (!CHOL) {
                        y[c] *= rcp_nr(dr[c]);  // the multiplier of sweep 1's rows
#pragma unroll
                        for (int k = c + 1; k < 16; k++)
                            y[k] = fma(-y[c], dr[k], y[k]);
                    } else {
                        y[c] *= Rc[c];  // 1 / L11[c][c], as sweep 1 scaled its rows
#pragma unroll
                        for (int k = c + 1; k < 16; k++)  // L11[k][c]
                            y[k] = fma(-y[c], P[(p0 + k) * BF_LD + c], y[k]);
                    }
                }
#pragma unroll
                for (int c = 0; c < 16; c++)
                    if (in)
                        P[r * BF_LD + c] = y[c];
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            }
            if (!CHOL && lane < 16) {
                // column `lane` of inv(L11) (unit lower): forward substitution
                double z[16];
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    double s = r == lane ? 1.0 : 0.0;
#pragma unroll
                    for (int k = 0; k < r; k++)
                        s = fma(-P[(p0 + r) * BF_LD + k], z[k], s);
                    z[r] = s;
                }
#pragma unroll
                for (int r = 0; r < 16; r++)
                    Li[r * BF_LD + lane] = z[r];
            }
            // the factored panel back into the owner's registers
#pragma unroll
            for (int i = 0; i < 8; i++)
                if (i >= p) {
                    v4d blk;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        blk[e] = P[(16 * i + 4 * e + q) * BF_LD + cl];
                    if (CHOL && i == p)  // keep the diagonal block's upper part as loaded
#pragma unroll
                        for (int e = 0; e < 4; e++)
                            if (4 * e + q < cl)
                                blk[e] = 0.0;
                    if (JJ > 1 && jp)
                        R[JJ - 1][i] = blk;
                    else
                        R[0][i] = blk;
                }
        }
        lds_barrier();
        // ---- trailing update of this wave's block-columns j > p ----
#pragma unroll
        for (int jj = 0; jj < JJ; jj++) {
            const int j = w + NW * jj;
            if (j <= p || j >= nb)
                continue;
            if (!CHOL) {
                // U12 = inv(L11) A12, block (p, j)
                v4d up = R[jj][0];
#pragma unroll
                for (int i = 1; i < 8; i++)
                    if (i == p)
                        up = R[jj][i];
                v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
                    acc = mfma_f64(Li[cl * BF_LD + 4 * kk + q], up[kk], acc);
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (i == p)
                        R[jj][i] = acc;
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (i > p && i < nb) {
                        v4d c4 = R[jj][i];
#pragma unroll
                        for (int kk = 0; kk < 4; kk++)
                            c4 = mfma_f64(-P[(16 * i + cl) * BF_LD + 4 * kk + q], acc[kk], c4);
                        R[jj][i] = c4;
                    }
            } else {
                double bj[4];
#pragma unroll
                for (int kk = 0; kk < 4; kk++)
                    bj[kk] = P[(16 * j + cl) * BF_LD + 4 * kk + q];
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (i >= j && i < nb) {
                        v4d c4 = R[jj][i];
#pragma unroll
                        for (int kk = 0; kk < 4; kk++)
                            c4 = mfma_f64(-P[(16 * i + cl) * BF_LD + 4 * kk + q], bj[kk], c4);
                        R[jj][i] = c4;
                    }
            }
        }
        lds_barrier();
    }

#pragma unroll
    for (int jj = 0; jj < JJ; jj++)
#pragma unroll
        for (int i = 0; i < 8; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = 16 * i + 4 * e + q, c = 16 * (w + NW * jj) + cl;
                if (r < n && c < n && (!CHOL || r >= c))
                    a[r + (size_t)c * lda] = R[jj][i][e];
            }
    __builtin_amdgcn_s_setprio(0);
}

// LDS (doubles) one workgroup needs for any item.
constexpr int LDS_GEMM = 2 * (LDS_I > LDS_K ? LDS_I : LDS_K) * 2;  // 2 operands x 2 buffers
constexpr int LDS_TRSM = TRSM_NMAX * 16 + 16 * 17;
constexpr int LDS_SMALL = LDS_SMALL_BLK;
static_assert(TRINV_BLOCKS * 16 * 17 <= LDS_TRSM, "TRINV staging fits the generic LDS");
static_assert(TRINV_BLOCKS * 16 == WG_THREADS, "TRINV: one thread per block column");
static_assert(4 * GEMM_BM * 20 * 2 <= LDS_GEMM * 8 && 4 * GEMM_BM * 17 * 4 <= LDS_GEMM * 8,
              "reduced-precision GEMM buffers fit the GEMM LDS");
constexpr int LDS_DOUBLES = LDS_GEMM > LDS_TRSM ? (LDS_GEMM > LDS_SMALL ? LDS_GEMM : LDS_SMALL)
                                                : (LDS_TRSM > LDS_SMALL ? LDS_TRSM : LDS_SMALL);

// Order of the output blocks of one GEMM/SYRK unit.  Consecutive items run
// at the same time (and round-robin over the XCDs), so they should share
// operand strips: items walk groups of BLOCK_GROUP block rows column by
// column, rows fastest inside a group.  A group's A strips (32 x 64 rows x K)
// stay in L2 while its columns go by, and each B strip is read by the group's
// 32 blocks in a row.  A unit of at most BLOCK_GROUP block rows (one tile
// GEMM) is plain column-major.  Measured on the merged LU trailing update
// (22000 x 22000 x 1000): 55.9 TFLOP/s column-major, 58.6 with groups of 32.
constexpr int BLOCK_GROUP = 32;

// GEMM item i of a gm x gn block grid -> (bi, bj).
__device__ __forceinline__ void gemm_block_order(int gm, int gn, int i, int &bi, int &bj)
{
    const int per = BLOCK_GROUP * gn, first = (i / per) * BLOCK_GROUP;
    const int gsz = min(gm - first, BLOCK_GROUP), local = i - (first / BLOCK_GROUP) * per;
    bi = first + local % gsz;
    bj = local / gsz;
}

// SYRK item i -> (bi, bj), bi >= bj, over the lower blocks of a g x g grid:
// groups of BLOCK_GROUP block rows; inside a group the columns left of it
// (full height) and then its own lower triangle, column by column.
__device__ __forceinline__ void syrk_block(int g, int i, int &bi, int &bj)
{
    int r0 = 0;
    for (;;) {
        const int gsz = min(g - r0, BLOCK_GROUP);
        const int count = r0 * gsz + gsz * (gsz + 1) / 2;
        if (i < count || r0 + gsz >= g) {
            if (i < r0 * gsz) {
                bj = i / gsz;
                bi = r0 + i % gsz;
                return;
            }
            i -= r0 * gsz;
            int c = 0;
            while (i >= gsz - c) {
                i -= gsz - c;
                c++;
            }
            bj = r0 + c;
            bi = r0 + c + i;
            return;
        }
        i -= count;
        r0 += gsz;
    }
}

template <bool TA, bool TB>
__device__ __forceinline__ void gemm_dispatch_vec(const UnitDesc &u, int bi, int bj, bool lo,
                                                  double *smem)
{
    if (u.flags & (F_LOW_F32 | F_LOW_BF16)) {
#define MU_LOW_CASE(P, SRC)                                               \
    if (u.flags & F_VEC)                                                  \
        gemm_block_low<TA, TB, true, P, SRC>(u, bi, bj, lo, smem);        \
    else                                                                  \
        gemm_block_low<TA, TB, false, P, SRC>(u, bi, bj, lo, smem);
        const bool src = (u.flags & F_SRC_LOW) != 0;
        if (u.flags & F_LOW_F32) {
            if (src) {
                MU_LOW_CASE(1, true)
            } else {
                MU_LOW_CASE(1, false)
            }
        } else {
            if (src) {
                MU_LOW_CASE(2, true)
            } else {
                MU_LOW_CASE(2, false)
            }
        }
#undef MU_LOW_CASE
    } else if (u.flags & F_VEC) {
        gemm_block<TA, TB, true>(u, bi, bj, lo, smem);
    } else {
        gemm_block<TA, TB, false>(u, bi, bj, lo, smem);
    }
}

// GEMM or SYRK item `it` of unit `u` (the only kinds k_batch_gemm runs).
// ---------------------------------------------------------------------------
// GEMM block for narrow outputs (units flagged F_NARROW: n <= 16 columns, the
// right-hand sides of a substitution step; FP64, op(B) = B).  An item is one
// 16-row tile x all n columns, so a level of these (one 1000-row step is 63
// items) spreads over many CUs.  The 4 waves split K (k-chunks of 32 taken
// round-robin) and read the MFMA fragments straight from global memory --
// lane (li, lk) supplies op(A)[row li][k + lk] and B[k + lk][col li] -- with
// the next chunk's loads issued before the current chunk's MFMAs; the four
// partial tiles are added through LDS.  A 64 x 64 block would spend 3/4 of
// its MFMAs on absent columns and leave most CUs idle.
// ---------------------------------------------------------------------------
constexpr int NARROW_ROWS = 16;  // output rows per narrow item

template <bool TA>
__device__ __attribute__((always_inline)) inline void gemm_narrow(const UnitDesc &u, int bi,
                                                                  double *smem)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int i0 = bi * NARROW_ROWS, K = u.k, ncol = u.n;
    const MU_G double *a = gp(u.a);
    const MU_G double *b = gp(u.b);
    const size_t lda = u.lda, ldb = u.ldb;
    const bool cok = li < ncol;
    // Every load is issued unconditionally from an address clamped into the
    // operand (a lane's row and column clamped, k clamped to K - 1), so no
    // branch separates a load from the others in flight and the wait before
    // a chunk's MFMAs covers only that chunk's loads.  Rows and columns past
    // the block are computed but never stored; the k-steps past K are zeroed
    // in both operands as they enter the MFMA.
    const MU_G double *bc = b + (size_t)min(li, ncol - 1) * ldb;
    const int rc = min(i0 + li, u.m - 1);
    const MU_G double *ar = TA ? a + (size_t)rc * lda : a + rc;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    constexpr int KU = 8;  // k-steps of 4 per chunk
    double fa[2][KU], fb[2][KU];
    auto load = [&](double (&ga)[KU], double (&gb)[KU], int k) {
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const int kc = min(k + 4 * q + lk, K - 1);
            gb[q] = bc[kc];
            ga[q] = TA ? ar[kc] : ar[(size_t)kc * lda];
        }
    };
    auto mma = [&](const double (&ga)[KU], const double (&gb)[KU], int k) {
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const bool kok = k + 4 * q + lk < K;
            acc = mfma_f64(kok ? ga[q] : 0.0, kok ? gb[q] : 0.0, acc);
        }
    };
    constexpr int CH = 4 * KU;  // k per chunk
    int k = wave * CH;
    if (K > 0 && k < K) {
        load(fa[0], fb[0], k);
        for (;;) {
            const int kn = k + 4 * CH;
            if (kn < K) {
                load(fa[1], fb[1], kn);
                mma(fa[0], fb[0], k);
            } else {
                mma(fa[0], fb[0], k);
                break;
            }
            const int kn2 = kn + 4 * CH;
            if (kn2 < K) {
                load(fa[0], fb[0], kn2);
                mma(fa[1], fb[1], kn);
            } else {
                mma(fa[1], fb[1], kn);
                break;
            }
            k = kn2;
        }
    }
    // waves 1..3 hand their partial tiles to wave 0: LDS slot [wave - 1][e][lane]
    if (wave > 0)
#pragma unroll
        for (int e = 0; e < 4; e++)
            smem[((wave - 1) * 4 + e) * 64 + lane] = acc[e];
    lds_barrier();
    if (wave != 0)
        return;
    const double alpha = u.alpha, beta = u.beta;
    MU_G double *c = gp(u.c);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        // acc[e] = D[4e + lk][li]: row i0 + 4e + lk, column li
        const int row = i0 + 4 * e + lk;
        const double v = acc[e] + (smem[e * 64 + lane] + smem[(4 + e) * 64 + lane] +
                                   smem[(8 + e) * 64 + lane]);
        if (row < u.m && cok) {
            MU_G double *cp = c + row + (size_t)li * u.ldc;
            *cp = (beta != 0.0 ? beta * *cp : 0.0) + alpha * v;
        }
    }
}
static_assert(3 * 4 * 64 <= LDS_GEMM, "narrow GEMM reduction fits the GEMM LDS");

// ---------------------------------------------------------------------------
// Narrow GEMM with an FP32 A (F_NARROW | F_A32): C = beta C + alpha A B for
// A m x k stored FP32 (column-major, lda in elements), B k x n FP64 (n <= 16)
// and C FP64 -- a substitution step of a refinement solve on FP32 copies of
// the factors, where A is the whole traffic.  An item is 64 output rows x
// every column x one k-slice of NARROW32_KSLICE (u.gn slices): short slices
// put many items in flight at once, so a step costs a few memory round
// trips instead of one per 128 k.  Lane (li, lk) owns rows 4 li .. 4 li + 3
// of the item: per k-step of 4 it loads one float4 of A (16 lanes cover 256
// contiguous bytes of a column, 4 columns per instruction) and one element
// B[k + lk][li], rounded to FP32.  Row e of the lane's group feeds MFMA e
// (v_mfma_f32_16x16x4f32), whose output row index li stands for item row
// 4 li + e.  The 4 waves split the slice in chunks of 32 taken round-robin
// (loads of the next chunk in flight over the current chunk's MFMAs); the
// partial tiles meet in LDS and each wave finishes 16 of the 64 rows.  One
// slice: the FP32 sums are added into C in FP64.  Several: each item stores
// its 64 x 16 partial tile, and the item that finds itself last of its row
// block on the block's arrival counter adds the slices' tiles in slice order
// (FP64; the same order whatever the arrival order, so results do not
// depend on scheduling), updates C and resets the counter.  VEC: a 16-byte
// aligned, lda and m multiples of 4.
// ---------------------------------------------------------------------------
constexpr int NARROW32_ROWS = 64;  // output rows per FP32-A narrow item

template <bool VEC>
__device__ __attribute__((always_inline)) inline void gemm_narrow_a32(const UnitDesc &u, int it,
                                                                      double *smem)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int bi = it % u.gm, sl = it / u.gm, ns = u.gn;
    const int i0 = bi * NARROW32_ROWS, m = u.m, ncol = u.n;
    // this item's k range [k0, K): all of k when the unit is not split
    const int k0 = sl * NARROW32_KSLICE, K = ns == 1 ? u.k : min(u.k, k0 + NARROW32_KSLICE);
    const MU_G float *a = (const MU_G float *)u.a;
    const MU_G double *b = gp(u.b);
    const size_t lda = u.lda, ldb = u.ldb;
    const bool cok = li < ncol;
    // B(k, j) at b[k sk + j sj], C(i, j) at c[i ci + j cj]: columns, or rows
    // with F_BROW (a k-step of the 16 x 4 B fragment is then 512 contiguous
    // bytes instead of 16 columns' 32)
    const bool brow = (u.flags & F_BROW) != 0;
    const size_t sk = brow ? ldb : 1, sj = brow ? 1 : ldb;
    const size_t ci = brow ? (size_t)u.ldc : 1, cj = brow ? 1 : (size_t)u.ldc;
    // loads unconditional from clamped addresses, the k-steps past the
    // slice zeroed as they enter the MFMA (as in gemm_narrow)
    const MU_G double *bc = b + (size_t)min(li, ncol - 1) * sj;
    const int r0 = i0 + 4 * li;  // this lane's rows r0 .. r0 + 3
    // VEC: m % 4 == 0, so the lane's 4 rows are all inside or all outside
    const int r0c = min(r0, m - 4);
    v4f acc[4];
#pragma unroll
    for (int e = 0; e < 4; e++)
        acc[e] = v4f{0.f, 0.f, 0.f, 0.f};
    constexpr int KU = 8;  // k-steps of 4 per chunk
    v4f fa[2][KU];
    double fb[2][KU];
    auto load = [&](v4f (&ga)[KU], double (&gb)[KU], int k) {
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const int kc = min(k + 4 * q + lk, K - 1);
            gb[q] = bc[(size_t)kc * sk];
            if (VEC) {
                ga[q] = *(const MU_G v4f *)(a + r0c + (size_t)kc * lda);
            } else {
                const MU_G float *ac = a + (size_t)kc * lda;
                ga[q] = v4f{ac[min(r0, m - 1)], ac[min(r0 + 1, m - 1)], ac[min(r0 + 2, m - 1)],
                            ac[min(r0 + 3, m - 1)]};
            }
        }
    };
    auto mma = [&](const v4f (&ga)[KU], const double (&gb)[KU], int k) {
#pragma unroll
        for (int q = 0; q < KU; q++) {
            const bool kok = k + 4 * q + lk < K;
            const float bv = kok ? (float)gb[q] : 0.f;
            const v4f av = kok ? ga[q] : v4f{0.f, 0.f, 0.f, 0.f};
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bv, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bv, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bv, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bv, acc[3], 0, 0, 0);
        }
    };
    constexpr int CH = 4 * KU;  // k per chunk
    int k = k0 + wave * CH;
    if (k < K) {
        load(fa[0], fb[0], k);
        for (;;) {
            const int kn = k + 4 * CH;
            if (kn < K) {
                load(fa[1], fb[1], kn);
                mma(fa[0], fb[0], k);
            } else {
                mma(fa[0], fb[0], k);
                break;
            }
            const int kn2 = kn + 4 * CH;
            if (kn2 < K) {
                load(fa[0], fb[0], kn2);
                mma(fa[1], fb[1], kn);
            } else {
                mma(fa[1], fb[1], kn);
                break;
            }
            k = kn2;
        }
    }
    // partial tiles to LDS: slot [wave][e][e'][lane], acc[e][e'] = D_e[4 lk + e'][li],
    // item row 4 (4 lk + e') + e, column li
    float *sm = reinterpret_cast<float *>(smem);
#pragma unroll
    for (int e = 0; e < 4; e++)
#pragma unroll
        for (int f = 0; f < 4; f++)
            sm[((wave * 4 + e) * 4 + f) * 64 + lane] = acc[e][f];
    lds_barrier();
    // wave w finishes e' = w: rows 16 lk + 4 w + e (e = 0..3) of column li
    float s[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        s[e] = 0.f;
#pragma unroll
        for (int w = 0; w < 4; w++)
            s[e] += sm[((w * 4 + e) * 4 + wave) * 64 + lane];
    }
    const double alpha = u.alpha, beta = u.beta;
    MU_G double *c = gp(u.c);
    const int rb = i0 + 16 * lk + 4 * wave;
    auto store = [&](const double (&v)[4]) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = rb + e;
            if (row < m && cok) {
                MU_G double *cp = c + (size_t)row * ci + (size_t)li * cj;
                *cp = (beta != 0.0 ? beta * *cp : 0.0) + alpha * v[e];
            }
        }
    };
    if (ns == 1) {
        const double v[4] = {s[0], s[1], s[2], s[3]};
        store(v);
        return;
    }
    // several slices: publish this slice's tile, count the arrival
    MU_G float *part = gp(u.part);
    int32_t *cnt = reinterpret_cast<int32_t *>(u.part + (size_t)u.gm * ns * 1024) + bi;
#pragma unroll
    for (int e = 0; e < 4; e++)
        part[((size_t)bi * ns + sl) * 1024 + (wave * 4 + e) * 64 + lane] = s[e];
    // each wave waits for its stores to reach L2 and the barrier puts all of
    // them before thread 0's agent-scope release (one L2 write-back per
    // item) and count; the last arrival's acquire (one L2 invalidation per
    // row block) and the barrier after it come before every wave's loads
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    int *flag = reinterpret_cast<int *>(sm + 4 * 4 * 4 * 64);
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const int old = __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == ns - 1)
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *flag = old;
    }
    __syncthreads();
    if (*flag != ns - 1)
        return;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < ns; q++)
#pragma unroll
        for (int e = 0; e < 4; e++)
            v[e] += (double)part[((size_t)bi * ns + q) * 1024 + (wave * 4 + e) * 64 + lane];
    store(v);
    if (threadIdx.x == 0)
        __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static_assert(4 * 4 * 4 * 64 + 1 <= 2 * LDS_GEMM, "FP32-A narrow reduction fits the GEMM LDS");

// Narrow GEMM item (F_NARROW units only; kept out of run_gemm_item so the
// 64 x 64 block kernels keep their register budget).
__device__ __attribute__((always_inline)) inline void run_gemm_narrow_item(const UnitDesc &u,
                                                                           int it, double *smem)
{
    if (u.flags & F_A32) {
        if (u.flags & F_VEC)
            gemm_narrow_a32<true>(u, it, smem);
        else
            gemm_narrow_a32<false>(u, it, smem);
        return;
    }
    if (u.flags & F_TA)
        gemm_narrow<true>(u, it, smem);
    else
        gemm_narrow<false>(u, it, smem);
}

// One GEMM output block of a unit without F_NARROW, any transposition.
__device__ __attribute__((always_inline)) inline void gemm_any_block(const UnitDesc &u, int bi,
                                                                     int bj, double *smem)
{
    switch (u.flags & (F_TA | F_TB)) {
    case 0: gemm_dispatch_vec<false, false>(u, bi, bj, false, smem); break;
    case F_TB: gemm_dispatch_vec<false, true>(u, bi, bj, false, smem); break;
    case F_TA: gemm_dispatch_vec<true, false>(u, bi, bj, false, smem); break;
    default: gemm_dispatch_vec<true, true>(u, bi, bj, false, smem); break;
    }
}

// GEMM/SYRK item of a unit without F_NARROW.  GROUPED: the unit's items
// cover grp x grp output blocks (the persistent kernel's lowering); the
// level kernels run one block per item and keep the single-block form.
template <bool GROUPED = false>
__device__ __attribute__((always_inline)) inline void run_gemm_item(const UnitDesc &u, int it,
                                                                    double *smem)
{
    if (u.op == OP_SYRK) {
        int bi, bj;
        syrk_block(u.gm, it, bi, bj);
        // A A^T: A read directly (op(B) = A^T), or, for a reduced-precision
        // panel stored transposed, A^T read as op(A) = (A^T)^T
        if (u.flags & F_TA)
            gemm_dispatch_vec<true, false>(u, bi, bj, bi == bj, smem);
        else
            gemm_dispatch_vec<false, true>(u, bi, bj, bi == bj, smem);
        return;
    }
    if (!GROUPED || u.grp <= 1) {
        int bi, bj;
        gemm_block_order(u.gm, u.gn, it, bi, bj);
        gemm_any_block(u, bi, bj, smem);
        return;
    }
    // g x g output blocks one after the other, a barrier between them (the
    // next block's staging overwrites the LDS the previous one read last)
    const int g = u.grp;
    const int gm = (u.gm + g - 1) / g;
    const int bi0 = (it % gm) * g, bj0 = (it / gm) * g;
#pragma unroll 1
    for (int q = 0; q < g * g; q++) {
        const int bi = bi0 + q % g, bj = bj0 + q / g;
        if (bi >= u.gm || bj >= u.gn)
            continue;
        if (q)
            lds_barrier();
        gemm_any_block(u, bi, bj, smem);
    }
}

// GEMM or SYRK item of a unit flagged F_KVEC | F_BLK128 (k_batch_gemm_low;
// the graph builder gives every unit of that kernel 128 x 128 blocks).
__device__ __attribute__((always_inline)) inline void run_gemm_lowk_item(const UnitDesc &u, int it,
                                                                         double *smem)
{
    int bi, bj;
    bool lo = false;
    if (u.op == OP_SYRK) {
        syrk_block(u.gm, it, bi, bj);
        lo = bi == bj;
    } else {
        gemm_block_order(u.gm, u.gn, it, bi, bj);
    }
    if (u.flags & F_LOW_BF16)
        gemm_block_lowk128<2>(u, bi, bj, lo, smem);
    else
        gemm_block_lowk128<1>(u, bi, bj, lo, smem);
}

// FP64 block -> FP32 / BF16 storage: item = one 64 x 64 block of the m x n
// source (c, ldc), written to (b, ldb).
// F_CVT_T: the block is written transposed (element (r, c) at b[c + r*ldb]).
// Each wave turns its own 64 x 16 column slice around through LDS (its own
// 16 x 65 patch, no workgroup barrier): it reads the FP64 columns (a wave
// reads 64 consecutive rows), and writes rows of 16 consecutive output
// elements.
template <class T>
__device__ __forceinline__ void cvt_block_t(const UnitDesc &u, int i0, int j0, double *smem)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T *tl = reinterpret_cast<T *>(smem) + w * 16 * 65;  // [column][row], pitch 65
    const MU_G double *src = gp(u.c);
    const int r = i0 + lane;
    const bool both = (u.flags & F_CVT_BOTH) != 0;
    MU_G T *plain = (MU_G T *)u.a;
#pragma unroll
    for (int c = 0; c < 16; c++) {
        const int cc = j0 + 16 * w + c;
        const bool in = r < u.m && cc < u.n;
        const double v = in ? src[r + (size_t)cc * u.ldc] : 0.0;
        const T cvv = LowPrec<sizeof(T) == 2 ? 2 : 1>::cv(v);
        tl[c * 65 + lane] = cvv;
        if (both && in)
            plain[r + (size_t)cc * u.lda] = cvv;
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    MU_G T *dst = (MU_G T *)u.b;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int rr = 4 * q + (lane >> 4), c = lane & 15;
        const int gr = i0 + rr, gc = j0 + 16 * w + c;
        if (gr < u.m && gc < u.n)
            dst[gc + (size_t)gr * u.ldb] = tl[c * 65 + rr];
    }
}

__device__ __attribute__((always_inline)) inline void cvt_block(const UnitDesc &u, int it,
                                                                double *smem)
{
    const int bi = it % u.gm, bj = it / u.gm;
    const int i0 = bi * 64, j0 = bj * 64;
    if (u.flags & F_CVT_T) {
        if (u.flags & F_LOW_BF16)
            cvt_block_t<unsigned short>(u, i0, j0, smem);
        else
            cvt_block_t<float>(u, i0, j0, smem);
        return;
    }
    const MU_G double *src = gp(u.c);
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int idx = threadIdx.x + WG_THREADS * q, r = i0 + (idx & 63), c = j0 + (idx >> 6);
        if (r < u.m && c < u.n) {
            const double v = src[r + (size_t)c * u.ldc];
            if (u.flags & F_LOW_BF16)
                ((MU_G unsigned short *)u.b)[r + (size_t)c * u.ldb] = to_bf16(v);
            else
                ((MU_G float *)u.b)[r + (size_t)c * u.ldb] = (float)v;
        }
    }
}

// TRSM item with inverted diagonal blocks on NW waves (k_batch_trsm,
// k_batch_trsm_mid).
template <int NW, int NMAX = TRSM_NMAX>
__device__ __attribute__((always_inline)) inline void run_trsm_inv_item(const UnitDesc &u, int it,
                                                                        double *smem)
{
    switch ((u.flags >> 3) & 0x27) {  // T_TRANS, Z_ROWS, UNIT, REV
#define MU_TRSM_INV_CASE(bits)                                                                 \
    case bits:                                                                                 \
        trsm_panel_inv<(bits & 1) != 0, (bits & 2) != 0, (bits & 4) != 0, (bits & 0x20) != 0, \
                       NW, NMAX>(u, it, smem);                                                 \
        break;
        MU_TRSM_INV_CASE(0x00) MU_TRSM_INV_CASE(0x01) MU_TRSM_INV_CASE(0x02) MU_TRSM_INV_CASE(0x03)
        MU_TRSM_INV_CASE(0x04) MU_TRSM_INV_CASE(0x05) MU_TRSM_INV_CASE(0x06) MU_TRSM_INV_CASE(0x07)
        MU_TRSM_INV_CASE(0x20) MU_TRSM_INV_CASE(0x21) MU_TRSM_INV_CASE(0x22) MU_TRSM_INV_CASE(0x23)
        MU_TRSM_INV_CASE(0x24) MU_TRSM_INV_CASE(0x25) MU_TRSM_INV_CASE(0x26) MU_TRSM_INV_CASE(0x27)
#undef MU_TRSM_INV_CASE
    default: break;
    }
}

// TRINV item (inverses of TRINV_BLOCKS diagonal blocks).
__device__ __attribute__((always_inline)) inline void run_trinv_item(const UnitDesc &u, int it,
                                                                     double *smem)
{
    switch ((u.flags >> 3) & 0x25) {  // T_TRANS, UNIT, REV
    case 0x00: trinv_blocks<false, false, false>(u, it, smem); break;
    case 0x01: trinv_blocks<true, false, false>(u, it, smem); break;
    case 0x04: trinv_blocks<false, true, false>(u, it, smem); break;
    case 0x05: trinv_blocks<true, true, false>(u, it, smem); break;
    case 0x20: trinv_blocks<false, false, true>(u, it, smem); break;
    case 0x21: trinv_blocks<true, false, true>(u, it, smem); break;
    case 0x24: trinv_blocks<false, true, true>(u, it, smem); break;
    default: trinv_blocks<true, true, true>(u, it, smem); break;
    }
}

// Panel solve of a triangle of order <= SMALL_MAX without inverted blocks
// (k_batch_trsm_small: the panel in 18.5 KB of LDS).
__device__ __attribute__((always_inline)) inline void run_trsm_small_item(const UnitDesc &u, int it,
                                                                          double *smem)
{
    switch ((u.flags >> 3) & 0x27) {  // T_TRANS, Z_ROWS, UNIT, REV
#define MU_TRSM_SMALL_CASE(bits)                                                           \
    case bits:                                                                             \
        trsm_panel<(bits & 1) != 0, (bits & 2) != 0, (bits & 4) != 0, (bits & 0x20) != 0, \
                   SMALL_MAX>(u, it, smem);                                                \
        break;
        MU_TRSM_SMALL_CASE(0x00) MU_TRSM_SMALL_CASE(0x01) MU_TRSM_SMALL_CASE(0x02)
        MU_TRSM_SMALL_CASE(0x03) MU_TRSM_SMALL_CASE(0x04) MU_TRSM_SMALL_CASE(0x05)
        MU_TRSM_SMALL_CASE(0x06) MU_TRSM_SMALL_CASE(0x07) MU_TRSM_SMALL_CASE(0x20)
        MU_TRSM_SMALL_CASE(0x21) MU_TRSM_SMALL_CASE(0x22) MU_TRSM_SMALL_CASE(0x23)
        MU_TRSM_SMALL_CASE(0x24) MU_TRSM_SMALL_CASE(0x25) MU_TRSM_SMALL_CASE(0x26)
        MU_TRSM_SMALL_CASE(0x27)
#undef MU_TRSM_SMALL_CASE
    default: break;
    }
}
constexpr int LDS_TRSM_SMALL = SMALL_MAX * 16 + 16 * 17;

// Small GETRF/POTRF or CVT item (the only kinds k_batch_small runs), on NW
// waves; a CVT item uses the first four.
template <int NW = 4>
__device__ __attribute__((always_inline)) inline void run_small_item(const UnitDesc &u, int it,
                                                                     double *smem)
{
    if (u.op == OP_GETRF)
        factor_small_blk<false, NW>(u, smem);
    else if (u.op == OP_POTRF)
        factor_small_blk<true, NW>(u, smem);
    else if (threadIdx.x < WG_THREADS)
        cvt_block(u, it, smem);
}

// Execute local item `it` of unit `u`.  Always inlined, so that LDS stays
// addressed as LDS and no call stack is needed in the persistent kernel.
__device__ __attribute__((always_inline)) inline void run_item(const UnitDesc &u, int it,
                                                               double *smem)
{
    switch (u.op) {
    case OP_GEMM:
        if (u.flags & F_NARROW)
            run_gemm_narrow_item(u, it, smem);
        else
            run_gemm_item(u, it, smem);
        break;
    case OP_SYRK: run_gemm_item(u, it, smem); break;
    case OP_TRSM:
        switch ((u.flags >> 3) & 0x67) {  // T_TRANS, Z_ROWS, UNIT, REV, INV
#define MU_TRSM_CASE(bits)                                                                      \
    case bits:                                                                                  \
        trsm_panel<(bits & 1) != 0, (bits & 2) != 0, (bits & 4) != 0, (bits & 0x20) != 0>(     \
            u, it, smem);                                                                       \
        break;                                                                                  \
    case bits | 0x40:                                                                           \
        trsm_panel_inv<(bits & 1) != 0, (bits & 2) != 0, (bits & 4) != 0, (bits & 0x20) != 0>( \
            u, it, smem);                                                                       \
        break;
            MU_TRSM_CASE(0x00) MU_TRSM_CASE(0x01) MU_TRSM_CASE(0x02) MU_TRSM_CASE(0x03)
            MU_TRSM_CASE(0x04) MU_TRSM_CASE(0x05) MU_TRSM_CASE(0x06) MU_TRSM_CASE(0x07)
            MU_TRSM_CASE(0x20) MU_TRSM_CASE(0x21) MU_TRSM_CASE(0x22) MU_TRSM_CASE(0x23)
            MU_TRSM_CASE(0x24) MU_TRSM_CASE(0x25) MU_TRSM_CASE(0x26) MU_TRSM_CASE(0x27)
#undef MU_TRSM_CASE
        default: break;
        }
        break;
    case OP_GETRF: factor_small_blk<false>(u, smem); break;
    case OP_POTRF: factor_small_blk<true>(u, smem); break;
    case OP_CVT: cvt_block(u, it, smem); break;
    case OP_TRINV: run_trinv_item(u, it, smem); break;
    default: break;
    }
}

}  // namespace mu
