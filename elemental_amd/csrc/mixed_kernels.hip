// Kernels of the mixed-precision solvers: precision conversion, diagonal-block inverses,
// flag-chained persistent triangular solves, and the FP64 residual with its norms.
#include <algorithm>
#include <climits>

#include "elx_common.hpp"
#include "exchange.hpp"
#include "mixed.hpp"

namespace {

constexpr long long SPIN_LIMIT = 1 << 20;

template <typename T>
__device__ __forceinline__ void st_x(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_x(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The working copy W = A^T in fp32: W[j + i*ldw] = float(A(i,j)), i.e. A stored row by row, so
// that the factorization's row interchanges move contiguous memory; the padding W[j + i*ldw],
// n <= j < ldw, is zeroed. rowabs[chunk][i] = sum_{j in chunk} |A(i,j)|, summed in column
// order (chunk = blockIdx.y). A workgroup owns 64 rows and walks its chunk's columns 64 at a
// time through an LDS tile, so the column-major reads and the transposed writes are both
// coalesced.
__global__ __launch_bounds__(256) void convert_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                                      float* __restrict__ W, i64 ldw,
                                                      double* __restrict__ rowabs, int cols_per_chunk) {
    __shared__ double tl[64][65];  // tl[column][row]
    const int t = threadIdx.x, tx = t & 63, ty = t >> 6;
    const i64 i0 = (i64)blockIdx.x * 64;
    const i64 j0 = (i64)blockIdx.y * cols_per_chunk;
    const i64 j1 = std::min<i64>(n, j0 + cols_per_chunk);
    double sum = 0.0;  // row i0 + t, for t < 64
    for (i64 jb = j0; jb < j1; jb += 64) {
        const int nc = (int)std::min<i64>(64, j1 - jb);
        for (int q = ty; q < 64; q += 4) {
            const i64 i = i0 + tx;
            tl[q][tx] = (i < n && q < nc) ? A[i + (jb + q) * lda] : 0.0;
        }
        __syncthreads();
        if (t < 64)
            for (int q = 0; q < nc; ++q) sum += fabs(tl[q][t]);
        for (int r = ty; r < 64; r += 4) {
            const i64 i = i0 + r;
            if (i < n && tx < nc) W[(jb + tx) + i * ldw] = (float)tl[tx][r];
        }
        __syncthreads();
    }
    if (t < 64 && i0 + t < n) rowabs[(i64)blockIdx.y * n + i0 + t] = sum;
    const int pad = (int)(ldw - n);
    if (blockIdx.y == gridDim.y - 1 && pad > 0)
        for (int idx = t; idx < 64 * pad; idx += 256) {
            const i64 i = i0 + idx / pad;
            if (i < n) W[(n + idx % pad) + i * ldw] = 0.0f;
        }
}

// Inverse of the 128 x 128 diagonal blocks of a unit-lower (upper == 0) or non-unit upper
// (upper == 1) triangular factor stored row by row in W (element (i, k) at W[k + i*ldw]).
// Block b's inverse goes to Inv + b*128*128 (column-major, ld 128; rows and columns >= nb
// zero). One workgroup per block, blocked by 32: the block is loaded as a lower triangle L
// (U^T for the upper factor, whose inverse is then the transpose) stored by columns, each
// wave inverts one 32 x 32 diagonal block by right-looking substitution (lane j holds column
// j in registers; step k reads column k of L as 16-byte vectors and updates every later
// entry), and the off-diagonal blocks follow level by level (distance d from the diagonal, all
// blocks of a level at once): X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj).
__global__ __launch_bounds__(256) void diag_inverse_kernel(i64 n, const float* __restrict__ W,
                                                           i64 ldw, int upper,
                                                           float* __restrict__ Inv, i64 b0) {
    extern __shared__ __align__(16) float sm[];
    constexpr int LS = 132;       // column stride of L (16-byte aligned columns)
    constexpr int XS = 129;       // row stride of X
    float* L = sm;                // L(i, k) at L[k*LS + i]
    float* X = sm + 128 * LS;     // X(i, k) at X[i*XS + k]
    float* S = X + 128 * XS;      // [3][32][33] block sums of one level
    const i64 b = b0 + blockIdx.x;
    const i64 r0 = b * 128;
    const int nb = (int)std::min<i64>(128, n - r0);
    const int t = threadIdx.x;
    for (int idx = t; idx < 128 * 128; idx += 256) {
        int i, k;
        if (!upper) {
            k = idx & 127;  // W[(r0+k) + (r0+i)*ldw]: contiguous in k
            i = idx >> 7;
        } else {
            i = idx & 127;  // W[(r0+i) + (r0+k)*ldw]: contiguous in i
            k = idx >> 7;
        }
        float v = 0.0f;
        if (i < nb && k < nb) {
            if (i == k) v = upper ? W[(r0 + i) + (r0 + i) * ldw] : 1.0f;
            else if (k < i) v = upper ? W[(r0 + i) + (r0 + k) * ldw] : W[(r0 + k) + (r0 + i) * ldw];
        } else if (i == k) {
            v = 1.0f;  // identity beyond nb
        }
        L[k * LS + i] = v;
        X[i * XS + k] = 0.0f;
    }
    __syncthreads();
    {
        typedef float f4v __attribute__((ext_vector_type(4)));
        const int w = t >> 6, j = t & 63, base = 32 * w;
        if (j < 32) {
            float x[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) x[i] = (i == j) ? 1.0f : 0.0f;
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const float* col = L + (base + k) * LS + base;  // L(base.., base+k)
                float cv[32];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const f4v v4 = *reinterpret_cast<const f4v*>(col + 4 * q);
                    cv[4 * q] = v4[0];
                    cv[4 * q + 1] = v4[1];
                    cv[4 * q + 2] = v4[2];
                    cv[4 * q + 3] = v4[3];
                }
                const float xk = x[k] / cv[k];
                x[k] = xk;
#pragma unroll
                for (int i = k + 1; i < 32; ++i) x[i] = __builtin_fmaf(-cv[i], xk, x[i]);
            }
#pragma unroll
            for (int i = 0; i < 32; ++i) X[(base + i) * XS + base + j] = x[i];
        }
    }
    __syncthreads();
    const int r = t >> 3, c0 = 4 * (t & 7);  // this thread's 4 entries of each 32 x 32 block
    for (int d = 1; d < 4; ++d) {
        const int nbl = 4 - d;  // blocks (bj + d, bj), bj = 0 .. nbl-1
        for (int bj = 0; bj < nbl; ++bj) {
            const int bi = bj + d;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (int kk = 32 * bj; kk < 32 * bi; ++kk) {
                const float l = L[kk * LS + 32 * bi + r];
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(l, X[kk * XS + 32 * bj + c0 + e], acc[e]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) S[(bj * 32 + r) * 33 + c0 + e] = acc[e];
        }
        __syncthreads();
        for (int bj = 0; bj < nbl; ++bj) {
            const int bi = bj + d;
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
            for (int m = 0; m < 32; ++m) {
                const float xv = X[(32 * bi + r) * XS + 32 * bi + m];
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = __builtin_fmaf(xv, S[(bj * 32 + m) * 33 + c0 + e], acc[e]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) X[(32 * bi + r) * XS + 32 * bj + c0 + e] = -acc[e];
        }
        __syncthreads();
    }
    float* out = Inv + b * 128 * 128;
    for (int idx = t; idx < 128 * 128; idx += 256) {
        const int i = idx & 127, k = idx >> 7;
        float v = 0.0f;
        if (i < nb && k < nb) v = upper ? X[k * XS + i] : X[i * XS + k];
        out[i + k * 128] = v;
    }
}

// X := Inv * X for one 128-row block X (nb x ncols) of a matrix stored row by row: X(k, j) at
// B[j + k*ldb]. Inv is the 128 x 128 inverse of a diagonal block (column-major, rows and
// columns >= nb zero). This is the triangular solve L11^{-1} A12 of the blocked LU as a product
// with the precomputed inverse: every column's result depends only on that column, so a
// workgroup reads its 64 columns (contiguous in every row) into LDS, multiplies, and writes
// them back in place, the product on the FP32 matrix cores (v_mfma_f32_16x16x4f32); the 16
// lanes of a result row group write 64 contiguous bytes.
constexpr int kTrmmLdi = 132, kTrmmLdb = 68;  // padded LDS strides (no bank conflicts)
__global__ __launch_bounds__(256) void trmm_inv_kernel(int nb, i64 ncols, const float* __restrict__ Inv,
                                                       float* B, i64 ldb, unsigned* maxbits) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    extern __shared__ __align__(16) float sm[];
    float* Is = sm;                        // 128 x 128, Is[i + k*kTrmmLdi] = Inv(i, k)
    float* Bs = sm + 128 * kTrmmLdi;       // 128 x 64, Bs[k*kTrmmLdb + j]
    const int t = threadIdx.x;
    const i64 c0 = (i64)blockIdx.x * 64;
    const int nc = (int)std::min<i64>(64, ncols - c0);
    for (int idx = t; idx < 128 * 128 / 4; idx += 256) {
        const int i4 = idx & 31, k = idx >> 5;
        *reinterpret_cast<f4v*>(Is + 4 * i4 + k * kTrmmLdi) = reinterpret_cast<const f4v*>(Inv)[idx];
    }
    for (int idx = t; idx < 128 * 64; idx += 256) {
        const int j = idx & 63, k = idx >> 6;
        Bs[k * kTrmmLdb + j] = (k < nb && j < nc) ? B[(c0 + j) + k * ldb] : 0.0f;
    }
    __syncthreads();
    // X = Inv B on the FP32 matrix cores: wave w owns rows 32w .. 32w+31 (two 16-row tiles) and
    // the 64 columns (four tiles); MFMA operands Inv(i, k) with lane i = l%16, k = l/16, and
    // B(k, j) with j = l%16
    const int lane = t & 63, wave = t >> 6, l16 = lane & 15, lq = lane >> 4;
    f4v acc[2][4];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[u][ct] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
    for (int k0 = 0; k0 < 128; k0 += 4) {
        const int k = k0 + lq;
        float a[2], b[4];
#pragma unroll
        for (int u = 0; u < 2; ++u) a[u] = Is[(32 * wave + 16 * u + l16) + k * kTrmmLdi];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) b[ct] = Bs[k * kTrmmLdb + 16 * ct + l16];
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                acc[u][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[ct], acc[u][ct], 0, 0, 0);
    }
    // lane l, register r: row 16*tile + 4*(l/16) + r, column 16*ct + l%16
    float mx = 0.0f;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 32 * wave + 16 * u + 4 * lq + r;
            if (i >= nb) continue;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int j = 16 * ct + l16;
                if (j < nc) {
                    B[(c0 + j) + i * ldb] = acc[u][ct][r];
                    mx = fmaxf(mx, fabsf(acc[u][ct][r]));
                }
            }
        }
    if (maxbits) {
        // the largest |entry| written, folded into *maxbits for the next update's exponent:
        // one atomic per workgroup (the grid's atomics on the one word serialise)
        __shared__ float wmx[4];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
        if ((t & 63) == 0) wmx[t >> 6] = mx;
        __syncthreads();
        if (t == 0)
            atomicMax(maxbits, __float_as_uint(fmaxf(fmaxf(wmx[0], wmx[1]), fmaxf(wmx[2], wmx[3]))));
    }
}

// S = the full symmetric matrix whose lower (upper == 0) or upper triangle is stored in A
// (column-major, lda); S has leading dimension n. 32 x 32 tiles through LDS, so the
// mirrored (transposed) half is read and written coalesced.
__global__ __launch_bounds__(256) void symmetrize_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                                         int upper, double* __restrict__ S) {
    __shared__ double tl[32][33];
    const i64 bi = blockIdx.x, bj = blockIdx.y;  // output tile (rows bi, columns bj)
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // the output tile lies in the stored triangle, or mirrors tile (bj, bi)
    const bool direct = upper ? (bi <= bj) : (bi >= bj);
    const i64 ri = (direct ? bi : bj) * 32, ci = (direct ? bj : bi) * 32;
    for (int q = ty; q < 32; q += 8) {
        const i64 r = ri + tx, c = ci + q;
        tl[q][tx] = (r < n && c < n) ? A[r + c * lda] : 0.0;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const i64 r = bi * 32 + tx, c = bj * 32 + q;
        if (r >= n || c >= n) continue;
        // element (r, c): stored directly when it is in the triangle, else as (c, r)
        const bool in_tri = upper ? (r <= c) : (r >= c);
        double v;
        if (direct) v = in_tri ? tl[q][tx] : tl[tx][q];  // diagonal tile: mirror inside it
        else v = tl[tx][q];
        S[r + c * n] = v;
    }
}

// The Cholesky factor L (lower triangle of W, column-major, ldw) into the unit-lower / upper
// pair of the LU-based triangular solves, stored row by row like the LU factors (element
// (i, j) of a factor at W[j + i*ldw]): L_u = L D^{-1} (strict lower), U = D L^T (upper,
// diagonal d_i^2), D = diag(L) given in d. L(r, c), r > c, sits at W[r + c*ldw], which is where
// U(c, r) = d_c L(r, c) belongs; L_u(r, c) = L(r, c) / d_c goes to the mirrored W[c + r*ldw].
// 32 x 32 tiles of the lower triangle, tile (bi, bj) also writing its mirror (bj, bi).
__global__ __launch_bounds__(256) void chol_to_lu_kernel(i64 n, float* W, i64 ldw, const float* __restrict__ d,
                                                         int ntiles) {
    __shared__ float tl[32][33];
    // lower tile index -> (bi, bj), bi >= bj
    int bi = (int)((__builtin_sqrtf(8.0f * (float)blockIdx.x + 1.0f) - 1.0f) * 0.5f);
    while ((bi + 1) * (bi + 2) / 2 <= (int)blockIdx.x) ++bi;
    while (bi * (bi + 1) / 2 > (int)blockIdx.x) --bi;
    const int bj = (int)blockIdx.x - bi * (bi + 1) / 2;
    (void)ntiles;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int q = ty; q < 32; q += 8) {
        const i64 r = (i64)bi * 32 + tx, c = (i64)bj * 32 + q;
        tl[q][tx] = (r < n && c < n) ? W[r + c * ldw] : 0.0f;  // tl[col][row]
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        // lower part: (r, c) = (bi*32+tx, bj*32+q)
        const i64 r = (i64)bi * 32 + tx, c = (i64)bj * 32 + q;
        if (r < n && c < n) {
            if (r > c) W[r + c * ldw] = d[c] * tl[q][tx];
            else if (r == c) W[r + c * ldw] = d[r] * d[r];
        }
        // mirrored part: (r2, c2) = (bj*32+tx, bi*32+q), r2 < c2, = L(c2, r2) / d_r2
        const i64 r2 = (i64)bj * 32 + tx, c2 = (i64)bi * 32 + q;
        if (r2 < n && c2 < n && r2 < c2) W[r2 + c2 * ldw] = tl[tx][q] / d[r2];
    }
}

__global__ void diag_extract_kernel(i64 n, const float* __restrict__ W, i64 ldw, float* d) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = W[i + i * ldw];
}

__global__ void iota_kernel(i64 n, int32_t* p) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

// Chunk q of solution block b: {x[2q], x[2q+1], epoch, tag}. The epoch word rejects chunks of
// earlier launches (and the zeroed area) exactly; the tag over the values, (b, q) and the epoch
// rejects a partly written chunk.
__device__ __forceinline__ uint32_t xblock_tag(uint32_t a, uint32_t b, uint32_t id, uint32_t epoch) {
    return epoch ^ elx::xchg::mix32(a, b, id ^ 0x5BD1E995u);
}

// Persistent chained triangular solve with 128-row blocks, fp32, one right-hand side, for
// factors stored row by row (element (i, j) at W[j + i*ldw]).
// lower == 1: L x = y (unit lower, forward); lower == 0: U x = y (upper, backward).
// Workgroup g handles blocks in solve order g, g+G, ...; block b accumulates
// sum_c T(b,c) x_c over the blocks c earlier in solve order as their x_c are published,
// subtracts it from y_b, multiplies by the precomputed inverse of its diagonal block and
// publishes x_b: to x (read by later kernels) and as 64 tagged 16-byte chunks in uncached
// memory (xc), which the other workgroups poll directly -- a block is complete when all its
// tags match the launch's epoch, so no store-completion wait or flag sits on the chain. Each
// wave checks one block per poll (up to four at once); each thread streams 16-byte slices of
// T (16 rows x 4 columns per block; the 32 lanes of a row read 512 contiguous bytes), and the
// tile of the next block in solve order is loaded while it is awaited. Each block only waits
// on blocks earlier in solve order, handled by co-resident workgroups, so every wait ends; a
// bounded spin reports failure instead. Requires ldw % 4 == 0 and the padding columns
// n..ldw-1 of every row zero.
__global__ __launch_bounds__(256) void trsv_kernel(i64 n, const float* __restrict__ W, i64 ldw,
                                                   const float* __restrict__ Inv, int lower,
                                                   const float* __restrict__ y, float* x,
                                                   elx::xchg::u4* xc, uint32_t epoch, int* err) {
    using elx::xchg::u4;
    typedef float f4v __attribute__((ext_vector_type(4)));
    __shared__ float xs[4][128];
    __shared__ float red[8][128];
    __shared__ float part[128][33];
    __shared__ int s_rdy[2][4];
    const int t = threadIdx.x, rq = 4 * (t & 31), jg = t >> 5;  // (rq, jg): the Inv product
    const int cq = 4 * (t & 31), rg = t >> 5;                   // (cq, rg): the tile product
    const int wave = t >> 6, lane = t & 63;
    const i64 nblk = (n + 127) / 128;
    const uint32_t nchunks = (uint32_t)(nblk * 64);
    int par = 0;
    for (i64 q = blockIdx.x; q < nblk; q += gridDim.x) {
        const i64 b = lower ? q : (nblk - 1 - q);
        const i64 r0 = b * 128;
        const int nb = (int)std::min<i64>(128, n - r0);
        const float* inv = Inv + b * 128 * 128;
        f4v iv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            iv[u] = *reinterpret_cast<const f4v*>(inv + rq + (jg + 8 * u) * 128);
        // acc[u]: this thread's share (columns cq..cq+3 of every tile) of row rg + 8u
        float acc[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc[u] = 0.f;
        i64 qc = 0;
        long long spins = 0;
        bool bail = false;
        f4v pw[16];
        i64 pf = -1;
        auto load_tile = [&](i64 qq, f4v (&w)[16]) {
            const i64 col = (lower ? qq : (nblk - 1 - qq)) * 128 + cq;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const i64 row = r0 + rg + 8 * u;
                w[u] = (row < n && col < n) ? *reinterpret_cast<const f4v*>(W + col + row * ldw)
                                            : f4v{0.f, 0.f, 0.f, 0.f};
            }
        };
        if (q > 0) {
            load_tile(0, pw);
            pf = 0;
        }
        while (qc < q && !bail) {
            // wave w checks block qc + w: lane l reads chunk l (x[2l], x[2l+1])
            const i64 qq = qc + wave;
            bool ready = false;
            if (qq < q) {
                const i64 cblk = lower ? qq : (nblk - 1 - qq);
                const uint32_t id = (uint32_t)(cblk * 64 + lane);
                const u4 ch = elx::xchg::ld16(xc, nchunks, id);
                ready = (ch[2] == epoch && ch[3] == xblock_tag(ch[0], ch[1], id, epoch));
                xs[wave][2 * lane] = __uint_as_float(ch[0]);
                xs[wave][2 * lane + 1] = __uint_as_float(ch[1]);
            }
            const bool wready = __all(ready ? 1 : 0);
            if (lane == 0) s_rdy[par][wave] = wready ? 1 : 0;
            __syncthreads();
            int k = 0;
            while (k < 4 && qc + k < q && s_rdy[par][k]) ++k;
            par ^= 1;  // a wave is at most one poll ahead of another: the parities never collide
            if (k == 0) {
                ++spins;
                if ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                    bail = true;
                if (spins > SPIN_LIMIT) {
                    if (t == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    bail = true;
                }
                continue;
            }
            for (int kk = 0; kk < k; ++kk) {
                const i64 qk = qc + kk;
                f4v w[16];
                if (qk == pf) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) w[u] = pw[u];
                } else {
                    load_tile(qk, w);
                }
                const f4v xv = *reinterpret_cast<const f4v*>(&xs[kk][cq]);
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    acc[u] = __builtin_fmaf(w[u][0], xv[0], acc[u]);
                    acc[u] = __builtin_fmaf(w[u][1], xv[1], acc[u]);
                    acc[u] = __builtin_fmaf(w[u][2], xv[2], acc[u]);
                    acc[u] = __builtin_fmaf(w[u][3], xv[3], acc[u]);
                }
            }
            qc += k;
            if (qc < q) {
                load_tile(qc, pw);
                pf = qc;
            }
            __syncthreads();  // every use of xs before the next poll overwrites it
        }
        // v = y_b - sum_c T(b,c) x_c: the 32 column shares of each row summed in a fixed order
#pragma unroll
        for (int u = 0; u < 16; ++u) part[rg + 8 * u][t & 31] = acc[u];
        __syncthreads();
        if (t < 128) {
            float sacc = 0.f;
#pragma unroll
            for (int g2 = 0; g2 < 32; ++g2) sacc += part[t][g2];
            xs[0][t] = (t < nb) ? (y[r0 + t] - sacc) : 0.0f;
        }
        __syncthreads();
        // x_b = Inv_b v
        f4v s4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const float vv = xs[0][jg + 8 * u];
            s4[0] = __builtin_fmaf(iv[u][0], vv, s4[0]);
            s4[1] = __builtin_fmaf(iv[u][1], vv, s4[1]);
            s4[2] = __builtin_fmaf(iv[u][2], vv, s4[2]);
            s4[3] = __builtin_fmaf(iv[u][3], vv, s4[3]);
        }
        red[jg][rq + 0] = s4[0];
        red[jg][rq + 1] = s4[1];
        red[jg][rq + 2] = s4[2];
        red[jg][rq + 3] = s4[3];
        __syncthreads();
        if (t < 64) {
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int g2 = 0; g2 < 8; ++g2) {
                s0 += red[g2][2 * t];
                s1 += red[g2][2 * t + 1];
            }
            if (2 * t >= nb) s0 = 0.f;
            if (2 * t + 1 >= nb) s1 = 0.f;
            if (2 * t < nb) x[r0 + 2 * t] = s0;
            if (2 * t + 1 < nb) x[r0 + 2 * t + 1] = s1;
            const uint32_t id = (uint32_t)(b * 64 + t);
            const uint32_t a0 = __float_as_uint(s0), a1 = __float_as_uint(s1);
            elx::xchg::st16(xc, nchunks, id, u4{a0, a1, epoch, xblock_tag(a0, a1, id, epoch)});
        }
        __syncthreads();
    }
}

// y(i) = float(scale_in * r(p[i])) (p may be null = identity)
__global__ void gather_to_f32_kernel(i64 n, const int32_t* __restrict__ p, const double* __restrict__ r,
                                     const double* __restrict__ sr, float* __restrict__ y) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const i64 k = p ? p[i] : i;
    y[i] = (float)(sr ? r[k] * sr[k] : r[k]);
}

// x(i) (+)= double(d(i))
// Y(i, k) = float(R(p[i], q[k]) * sr[p[i]]) for the block refinement (p, q, sr null =
// identity, identity, no scaling)
__global__ void gather_rows_f32_kernel(i64 n, const int32_t* __restrict__ p,
                                       const int32_t* __restrict__ q, const double* __restrict__ R,
                                       i64 ldr, const double* __restrict__ sr, float* __restrict__ Y,
                                       i64 ldy) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const i64 k = blockIdx.y;
    if (i >= n) return;
    const i64 kc = q ? q[k] : k;
    const i64 r = p ? p[i] : i;
    const double v = R[r + kc * ldr];
    Y[i + k * ldy] = (float)(sr ? v * sr[r] : v);
}
// X(i, q[k]) = [X(i, q[k]) +] double(Y(i, k)) * sc[i]   (sc null: no scaling)
__global__ void update_f64_2d_kernel(i64 n, const int32_t* __restrict__ q,
                                     const float* __restrict__ Y, i64 ldy,
                                     const double* __restrict__ sc, double* __restrict__ X,
                                     i64 ldx, int accumulate) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    const i64 k = blockIdx.y;
    if (i >= n) return;
    const i64 kc = q ? q[k] : k;
    double v = (double)Y[i + k * ldy];
    if (sc) v *= sc[i];
    X[i + kc * ldx] = accumulate ? X[i + kc * ldx] + v : v;
}
// out[k] = max_i |M(i, k)|, one workgroup per column
__global__ __launch_bounds__(256) void col_absmax_kernel(i64 n, const double* __restrict__ M, i64 ldm,
                                                         double* __restrict__ out) {
    const i64 k = blockIdx.x;
    double mx = 0.0;
    for (i64 i = threadIdx.x; i < n; i += 256) mx = fmax(mx, fabs(M[i + k * ldm]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    __shared__ double w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
}
__global__ void update_f64_kernel(i64 n, const float* d, const double* __restrict__ sc,
                                  double* __restrict__ x, int accumulate) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = (double)__hip_atomic_load(const_cast<float*>(d + i), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    if (sc) v *= sc[i];
    x[i] = accumulate ? x[i] + v : v;
}

// Residual partials: part[chunk][i] = sum_{j in chunk} A(i,j) x(j)  (FP64, one RHS)
__global__ __launch_bounds__(256) void gemv_partial_kernel(i64 n, const double* __restrict__ A, i64 lda,
                                                           const double* __restrict__ x,
                                                           double* __restrict__ part,
                                                           int cols_per_chunk) {
    __shared__ double xs[1024];
    const i64 j0 = (i64)blockIdx.y * cols_per_chunk;
    const int nc = (int)std::min<i64>(cols_per_chunk, n - j0);
    for (int q = threadIdx.x; q < nc; q += 256) xs[q] = x[j0 + q];
    __syncthreads();
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* a = A + i + j0 * lda;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = 0;
    for (; j + 4 <= nc; j += 4) {
        s0 = __builtin_fma(a[(i64)(j + 0) * lda], xs[j + 0], s0);
        s1 = __builtin_fma(a[(i64)(j + 1) * lda], xs[j + 1], s1);
        s2 = __builtin_fma(a[(i64)(j + 2) * lda], xs[j + 2], s2);
        s3 = __builtin_fma(a[(i64)(j + 3) * lda], xs[j + 3], s3);
    }
    for (; j < nc; ++j) s0 = __builtin_fma(a[(i64)j * lda], xs[j], s0);
    part[(i64)blockIdx.y * n + i] = (s0 + s1) + (s2 + s3);
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// r(i) = b(i) - sum_chunk part; maxima of |r|, |x|, |b| into norms[0..2] (zeroed first).
__global__ __launch_bounds__(256) void residual_finish_kernel(i64 n, int nch, const double* __restrict__ part,
                                                              const double* __restrict__ b,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ r, double* norms) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    double rv = 0.0, xv = 0.0, bv = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int c = 0; c < nch; ++c) s += part[(i64)c * n + i];
        rv = b ? b[i] - s : s;
        r[i] = rv;
        rv = fabs(rv);
        xv = fabs(x[i]);
        bv = b ? fabs(b[i]) : 0.0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        rv = fmax(rv, __shfl_xor(rv, off));
        xv = fmax(xv, __shfl_xor(xv, off));
        bv = fmax(bv, __shfl_xor(bv, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(norms + 0, rv);
        atomic_max_nonneg(norms + 1, xv);
        atomic_max_nonneg(norms + 2, bv);
    }
}

// max_i sum_c rowabs[c][i] -> *out (zeroed first)
__global__ __launch_bounds__(256) void rowabs_max_kernel(i64 n, int nch, const double* __restrict__ rowabs,
                                                         double* out) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n)
        for (int c = 0; c < nch; ++c) v += rowabs[(i64)c * n + i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, v);
}

// max_i |v_i| over a column (generic helper), into *out (zeroed first)
__global__ __launch_bounds__(256) void vec_absmax_kernel(i64 n, const double* __restrict__ v, double* out) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    double m = (i < n) ? fabs(v[i]) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, m);
}

// part[chunk][c] = sum_{i in chunk} V(i,c) w(i), c < k (chunk = 1024 rows = blockIdx.x)
__global__ __launch_bounds__(256) void multidot_partial_kernel(i64 n, int k, const double* __restrict__ V,
                                                               i64 ldv, const double* __restrict__ w,
                                                               double* __restrict__ part) {
    __shared__ double red[4];
    const i64 i0 = (i64)blockIdx.x * 1024;
    const int t = threadIdx.x;
    for (int c = 0; c < k; ++c) {
        double s = 0.0;
        for (int q = t; q < 1024; q += 256) {
            const i64 i = i0 + q;
            if (i < n) s = __builtin_fma(V[i + c * ldv], w[i], s);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if ((t & 63) == 0) red[t >> 6] = s;
        __syncthreads();
        if (t == 0) part[(i64)blockIdx.x * k + c] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}

// h[c] = sum_chunks part[chunk][c]
__global__ void multidot_finish_kernel(int nchunk, int k, const double* __restrict__ part, double* h) {
    const int c = threadIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int q = 0; q < nchunk; ++q) s += part[(i64)q * k + c];
    h[c] = s;
}

// w(i) += alpha * sum_c V(i,c) h[c]
__global__ __launch_bounds__(256) void multiaxpy_kernel(i64 n, int k, double alpha, const double* __restrict__ V,
                                                        i64 ldv, const double* __restrict__ h,
                                                        double* __restrict__ w) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s = __builtin_fma(V[i + c * ldv], h[c], s);
    w[i] = __builtin_fma(alpha, s, w[i]);
}

// y(i) += alpha * x(i)
__global__ void daxpy_kernel(i64 n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = __builtin_fma(alpha, x[i], y[i]);
}

// y(i) = alpha * x(i)
__global__ void dscal_copy_kernel(i64 n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = alpha * x[i];
}

}  // namespace

namespace elx {

elx_status_t launch_convert(hipStream_t s, i64 n, const double* A, i64 lda, float* W, i64 ldw,
                            double* rowabs, int nch) {
    const int cpc = (int)((n + nch - 1) / nch);
    ELX_ARG(ldw >= n && ldw - n < 1024);
    dim3 grid(cdiv(n, 64), (unsigned)nch);
    hipLaunchKernelGGL(convert_kernel, grid, dim3(256), 0, s, n, A, lda, W, ldw, rowabs, cpc);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_rowabs_max(hipStream_t s, i64 n, int nch, const double* rowabs, double* out) {
    hipLaunchKernelGGL(rowabs_max_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nch, rowabs, out);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_vec_absmax(hipStream_t s, i64 n, const double* v, double* out) {
    hipLaunchKernelGGL(vec_absmax_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, v, out);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_diag_inverse(hipStream_t s, i64 n, const float* W, i64 ldw, int upper,
                                 float* Inv, i64 b0, i64 count) {
    const i64 nblk = (count > 0) ? count : (n + 127) / 128 - b0;
    ELX_ARG(b0 >= 0 && b0 + nblk <= (n + 127) / 128);
    if (nblk <= 0) return ELX_SUCCESS;
    const size_t lds = (128 * 132 + 128 * 129 + 3 * 32 * 33) * sizeof(float);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&diag_inverse_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(diag_inverse_kernel, dim3((unsigned)nblk), dim3(256), lds, s, n, W, ldw,
                       upper, Inv, b0);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_symmetrize(hipStream_t s, i64 n, const double* A, i64 lda, int upper, double* S) {
    if (n <= 0) return ELX_SUCCESS;
    const unsigned T = (unsigned)cdiv(n, 32);
    hipLaunchKernelGGL(symmetrize_kernel, dim3(T, T), dim3(256), 0, s, n, A, lda, upper, S);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_chol_to_lu(hipStream_t s, i64 n, float* W, i64 ldw, float* d) {
    if (n <= 0) return ELX_SUCCESS;
    hipLaunchKernelGGL(diag_extract_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, W, ldw, d);
    ELX_HIP(hipGetLastError());
    const i64 T = (n + 31) / 32;
    const i64 ntiles = T * (T + 1) / 2;
    ELX_ARG(ntiles < (1ll << 31));
    hipLaunchKernelGGL(chol_to_lu_kernel, dim3((unsigned)ntiles), dim3(256), 0, s, n, W, ldw, d, (int)ntiles);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_iota(hipStream_t s, i64 n, int32_t* p) {
    if (n <= 0) return ELX_SUCCESS;
    hipLaunchKernelGGL(iota_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, p);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_trmm_inv(hipStream_t s, i64 nb, i64 ncols, const float* Inv, float* B, i64 ldb,
                             unsigned* maxbits) {
    ELX_ARG(nb >= 0 && nb <= 128 && ncols >= 0 && ldb >= nb);
    if (nb == 0 || ncols == 0) return ELX_SUCCESS;
    ProfScope prof_(s, kKTrsm, (double)nb * (double)nb * (double)ncols);
    const size_t lds = (128 * kTrmmLdi + 128 * kTrmmLdb) * sizeof(float);
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&trmm_inv_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
        attr = true;
    }
    hipLaunchKernelGGL(trmm_inv_kernel, dim3((unsigned)cdiv(ncols, 64)), dim3(256), lds, s, (int)nb,
                       ncols, Inv, B, ldb, maxbits);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_trsv(hipStream_t s, i64 n, const float* W, i64 ldw, const float* Inv,
                         int lower, const float* y, float* x, void* xchunks, uint32_t epoch,
                         int* err) {
    const DeviceInfo& di = device_info();
    const i64 nblk = (n + 127) / 128;
    ELX_ARG(ldw % 4 == 0 && nblk * 64 < ((i64)1 << 26));
    const unsigned G = (unsigned)std::min<i64>(nblk, di.num_cus);
    ProfScope prof_(s, kKTrsv, 0.5 * (double)n * (double)n * sizeof(float));
    elx::xchg::u4* xc = static_cast<elx::xchg::u4*>(xchunks);
    void* args[] = {&n, &W, &ldw, &Inv, &lower, &y, &x, &xc, &epoch, &err};
    ELX_CHECK(launch_resident(reinterpret_cast<const void*>(&trsv_kernel), G, 256, args, 0, s));
    return ELX_SUCCESS;
}

elx_status_t launch_gather_to_f32(hipStream_t s, i64 n, const int32_t* p, const double* r,
                                  const double* sr, float* y) {
    hipLaunchKernelGGL(gather_to_f32_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, p, r, sr, y);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_gather_rows_f32(hipStream_t s, i64 n, i64 nr, const int32_t* p,
                                   const int32_t* q, const double* R, i64 ldr, const double* sr,
                                   float* Y, i64 ldy) {
    if (n <= 0 || nr <= 0) return ELX_SUCCESS;
    ELX_ARG(nr < 65536);
    hipLaunchKernelGGL(gather_rows_f32_kernel, dim3(cdiv(n, 256), (unsigned)nr), dim3(256), 0, s, n,
                       p, q, R, ldr, sr, Y, ldy);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_update_f64_2d(hipStream_t s, i64 n, i64 nr, const int32_t* q, const float* Y,
                                  i64 ldy, const double* sc, double* X, i64 ldx, int accumulate) {
    if (n <= 0 || nr <= 0) return ELX_SUCCESS;
    ELX_ARG(nr < 65536);
    hipLaunchKernelGGL(update_f64_2d_kernel, dim3(cdiv(n, 256), (unsigned)nr), dim3(256), 0, s, n, q,
                       Y, ldy, sc, X, ldx, accumulate);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_col_absmax(hipStream_t s, i64 n, i64 nr, const double* M, i64 ldm, double* out) {
    if (nr <= 0) return ELX_SUCCESS;
    hipLaunchKernelGGL(col_absmax_kernel, dim3((unsigned)nr), dim3(256), 0, s, n, M, ldm, out);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_update_f64(hipStream_t s, i64 n, const float* d, const double* sc, double* x,
                              int accumulate) {
    hipLaunchKernelGGL(update_f64_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d, sc, x,
                       accumulate);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_multidot(hipStream_t s, i64 n, int k, const double* V, i64 ldv,
                             const double* w, double* part, double* h) {
    const unsigned nchunk = cdiv(n, 1024);
    hipLaunchKernelGGL(multidot_partial_kernel, dim3(nchunk), dim3(256), 0, s, n, k, V, ldv, w, part);
    ELX_HIP(hipGetLastError());
    hipLaunchKernelGGL(multidot_finish_kernel, dim3(1), dim3(64), 0, s, (int)nchunk, k, part, h);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_multiaxpy(hipStream_t s, i64 n, int k, double alpha, const double* V, i64 ldv,
                              const double* h, double* w) {
    hipLaunchKernelGGL(multiaxpy_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, k, alpha, V, ldv, h, w);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_axpy(hipStream_t s, i64 n, double alpha, const double* x, double* y) {
    hipLaunchKernelGGL(daxpy_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, alpha, x, y);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_scal_copy(hipStream_t s, i64 n, double alpha, const double* x, double* y) {
    hipLaunchKernelGGL(dscal_copy_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, alpha, x, y);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_residual(hipStream_t s, i64 n, const double* A, i64 lda, const double* x,
                             const double* b, double* r, double* part, int nch, double* norms) {
    const int cpc = (int)((n + nch - 1) / nch);
    if (cpc > 1024) return ELX_ERROR_INVALID_ARG;
    ProfScope prof_(s, kKResidual, (double)n * (double)n * sizeof(double));
    dim3 grid(cdiv(n, 256), (unsigned)nch);
    hipLaunchKernelGGL(gemv_partial_kernel, grid, dim3(256), 0, s, n, A, lda, x, part, cpc);
    ELX_HIP(hipGetLastError());
    hipLaunchKernelGGL(residual_finish_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, n, nch, part, b, x,
                       r, norms);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

}  // namespace elx
