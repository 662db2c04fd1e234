// Low-precision GEMM for the trailing updates of the mixed-precision factorizations:
//   C(fp32, column-major) += alpha * A * B^T
// where A (M x K) and B (N x K) are fp16 or bf16 matrices stored K-major (row r of A is
// K contiguous halves at A + r*lda), and the products accumulate in fp32 on the CDNA4
// matrix cores (v_mfma_f32_16x16x32_f16 / _bf16).
//
// Tiling: one 512-thread workgroup owns a 256 x 256 tile of C; its eight waves (2 along M,
// 4 along N) each own 128 x 64 as 8 x 4 MFMA accumulators (128 fp32 per lane). K advances 32
// at a time through a double-buffered LDS stage (rows padded to 40 halves so the 16-byte
// fragment reads of 16 consecutive rows hit disjoint banks), with the next slab's 16-byte
// global loads in flight while the current one feeds the MFMAs. The MFMA operands are fed as
// (B^T, A^T) so a lane's four results lie in one column of C: stores go out as 64-byte runs.
//
// Pack kernels produce the K-major operands from fp32 column-major data, with an exact
// power-of-two scale (and sign) applied before rounding; absmax / exponent kernels choose the
// scale that keeps an fp16 operand below 2^15.
#include <algorithm>

#include "elx_common.hpp"

namespace {

constexpr int BM = 256, BN = 256, BK = 32, NT = 512;
// K-slab depth of the kernel: 32, or 64 for the rank-NB updates (K a multiple of 64: half the
// barriers and slab hand-offs per flop); LDS rows padded by 8 halves so the 16-byte fragment
// reads of 16 consecutive rows hit disjoint banks
template <int BKV>
constexpr int ldk() {
    return BKV + 8;
}
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

template <int PREC>
struct HalfT;
template <>
struct HalfT<0> {
    typedef _Float16 T;
};
template <>
struct HalfT<1> {
    typedef __bf16 T;
};

template <int PREC>
__device__ __forceinline__ f4 mfma(u4 a, u4 b, f4 c) {
    if constexpr (PREC == 0)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                      __builtin_bit_cast(h8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                       __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

// One K-slab of a K-major operand: rows r0..r0+255, k0..k0+BKV-1; thread t loads BKV bytes
// of row t/2 (halves (BKV/2)*(t%2) ..). Rows >= R load zeros. The operand's K extent is padded
// to a multiple of BKV with zeros by the pack kernels, so k never runs past the row.
template <int BKV>
__device__ __forceinline__ void load_slab(const uint16_t* __restrict__ P, i64 ld, i64 r0, i64 R,
                                          i64 k0, u4 (&r)[BKV / 16]) {
    const int t = threadIdx.x;
    const i64 row = r0 + (t >> 1);
    if (row < R) {
        const u4* src = reinterpret_cast<const u4*>(P + row * ld + k0 + (BKV / 2) * (t & 1));
#pragma unroll
        for (int q = 0; q < BKV / 16; ++q) r[q] = src[q];
    } else {
#pragma unroll
        for (int q = 0; q < BKV / 16; ++q) r[q] = u4{0, 0, 0, 0};
    }
}

// The same slab from an FP32 K-major operand (element (r, k) at P[r*ld + k]), converted to the
// MFMA type as it is loaded (the conversion the pack kernels use); k >= K reads as zero. The
// row must be readable up to the next multiple of 32 past K.
template <int PREC>
__device__ __forceinline__ void load_slab_f32(const float* __restrict__ P, i64 ld, i64 r0, i64 R,
                                              i64 k0, i64 K, u4 (&r)[2]) {
    typedef typename HalfT<PREC>::T HT;
    const int t = threadIdx.x;
    const i64 row = r0 + (t >> 1);
    const i64 kb = k0 + 16 * (t & 1);
    if (row < R) {
        const float4* src = reinterpret_cast<const float4*>(P + row * ld + kb);
        float v[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 f = src[q];
            v[4 * q] = f.x;
            v[4 * q + 1] = f.y;
            v[4 * q + 2] = f.z;
            v[4 * q + 3] = f.w;
        }
        uint32_t w[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float a = (kb + 2 * i < K) ? v[2 * i] : 0.0f;
            const float b = (kb + 2 * i + 1 < K) ? v[2 * i + 1] : 0.0f;
            w[i] = (uint32_t)__builtin_bit_cast(uint16_t, (HT)a) |
                   ((uint32_t)__builtin_bit_cast(uint16_t, (HT)b) << 16);
        }
        r[0] = u4{w[0], w[1], w[2], w[3]};
        r[1] = u4{w[4], w[5], w[6], w[7]};
    } else {
        r[0] = u4{0, 0, 0, 0};
        r[1] = u4{0, 0, 0, 0};
    }
}

template <int BKV>
__device__ __forceinline__ void store_slab(uint16_t* S, const u4 (&r)[BKV / 16]) {
    const int t = threadIdx.x;
    u4* dst = reinterpret_cast<u4*>(S + (t >> 1) * ldk<BKV>() + (BKV / 2) * (t & 1));
#pragma unroll
    for (int q = 0; q < BKV / 16; ++q) dst[q] = r[q];
}

// BF32: B is the FP32 operand itself (element (n, k) at B[n*ldb + k]), converted while staged
// (32-deep slabs only).
template <int PREC, bool BF32, int BKV>
__global__ __launch_bounds__(NT) void lpgemm_kernel(i64 M, i64 N, i64 K, float alpha,
                                                    const uint16_t* __restrict__ A, i64 lda,
                                                    const void* __restrict__ Bv, i64 ldb,
                                                    float* __restrict__ C, i64 ldc,
                                                    const int* ea_dev, const int* eb_dev,
                                                    int tiles_m, int tiles_n) {
    static_assert(!BF32 || BKV == 32, "the FP32 operand is staged in 32-deep slabs");
    constexpr int LDK = ldk<BKV>(), RV = BKV / 16;
    __shared__ __align__(16) uint16_t As[2][BM * LDK];
    __shared__ __align__(16) uint16_t Bs[2][BN * LDK];

    const int total = tiles_m * tiles_n;
    const int per_xcd = (total + 7) / 8;
    const int lin = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (lin >= total) return;
    constexpr int GROUP = 4;
    const int group = lin / (GROUP * tiles_n);
    const int first_m = group * GROUP;
    const int gsize = min(tiles_m - first_m, GROUP);
    const int tm = first_m + (lin % (GROUP * tiles_n)) % gsize;
    const int tn = (lin % (GROUP * tiles_n)) / gsize;
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 128, wn = (wave >> 1) * 64;
    const int r16 = lane & 15, kq = lane >> 4;

    f4 acc[4][8];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int mi = 0; mi < 8; ++mi) acc[ni][mi] = f4{0.f, 0.f, 0.f, 0.f};

    const uint16_t* B = static_cast<const uint16_t*>(Bv);
    const float* Bf = static_cast<const float*>(Bv);
    u4 ra[RV], rb[RV];
    load_slab<BKV>(A, lda, m0, M, 0, ra);
    if constexpr (BF32) load_slab_f32<PREC>(Bf, ldb, n0, N, 0, K, rb);
    else load_slab<BKV>(B, ldb, n0, N, 0, rb);
    store_slab<BKV>(As[0], ra);
    store_slab<BKV>(Bs[0], rb);
    __syncthreads();
    int cur = 0;
    for (i64 k0 = 0; k0 < K; k0 += BKV) {
        const bool more = (k0 + BKV < K);
        if (more) {
            load_slab<BKV>(A, lda, m0, M, k0 + BKV, ra);
            if constexpr (BF32) load_slab_f32<PREC>(Bf, ldb, n0, N, k0 + BKV, K, rb);
            else load_slab<BKV>(B, ldb, n0, N, k0 + BKV, rb);
        }
        const uint16_t* as = As[cur];
        const uint16_t* bs = Bs[cur];
#pragma unroll
        for (int ks = 0; ks < BKV; ks += 32) {
            u4 af[8];
#pragma unroll
            for (int mi = 0; mi < 8; ++mi)
                af[mi] = *reinterpret_cast<const u4*>(as + (wm + mi * 16 + r16) * LDK + ks + 8 * kq);
#pragma unroll
            for (int ni = 0; ni < 4; ++ni) {
                const u4 bf = *reinterpret_cast<const u4*>(bs + (wn + ni * 16 + r16) * LDK + ks + 8 * kq);
#pragma unroll
                for (int mi = 0; mi < 8; ++mi) acc[ni][mi] = mfma<PREC>(bf, af[mi], acc[ni][mi]);
            }
        }
        if (more) {
            store_slab<BKV>(As[cur ^ 1], ra);
            store_slab<BKV>(Bs[cur ^ 1], rb);
        }
        __syncthreads();
        cur ^= 1;
    }

    // Lane holds D(i = n, j = m) with n = n0+wn+ni*16+4*kq+r, m = m0+wm+mi*16+r16.
    const int ea = ea_dev ? *ea_dev : 0, eb = eb_dev ? *eb_dev : 0;
    const float a = alpha * __builtin_ldexpf(1.0f, -(ea + eb));
    // C += a * acc, in CVN column groups of 32 values per lane: all of a round's loads are
    // issued before any of its stores, so they are in flight together (a load may not move
    // above a store to C that could alias it, which would make each read-modify-write wait out
    // a full memory round trip). The 64-deep form has the registers of its larger operand
    // slabs free by now and takes two groups per round.
    constexpr int CVN = BKV == 64 ? 2 : 1;
#pragma unroll
    for (int n0i = 0; n0i < 4; n0i += CVN) {
        float cv[CVN][4][8];
#pragma unroll
        for (int g = 0; g < CVN; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + (n0i + g) * 16 + 4 * kq + r;
#pragma unroll
                for (int mi = 0; mi < 8; ++mi) {
                    const i64 m = m0 + wm + mi * 16 + r16;
                    cv[g][r][mi] = (n < N && m < M) ? C[m + n * ldc] : 0.0f;
                }
            }
#pragma unroll
        for (int g = 0; g < CVN; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn + (n0i + g) * 16 + 4 * kq + r;
#pragma unroll
                for (int mi = 0; mi < 8; ++mi) {
                    const i64 m = m0 + wm + mi * 16 + r16;
                    if (n < N && m < M)
                        C[m + n * ldc] = __builtin_fmaf(a, acc[n0i + g][mi][r], cv[g][r][mi]);
                }
            }
    }
}

// ---------------------------------------------------------------------------------------
// Pack: H(r, k) = cvt(sign * 2^e * S(r, k)) for r < R, k < K; zeros for K <= k < Kpad.
// LO (fp16 only): the low part instead, cvt((v - float(cvt(v))) * 2^11) with v = sign * 2^e *
// S(r, k); the difference is exact in fp32, so hi + 2^-11 lo carries v to about 22 bits and a
// product of two such pairs (hi*hi + hi*lo + lo*hi) is FP32-accurate.
// S(r, k) = S[r + k*lds] (column-major, r fastest: TRANS = false) or S[k + r*lds]
// (TRANS = true). The column-major case goes through a 64x64 LDS transpose so both the reads
// and the K-major writes are coalesced.
// ---------------------------------------------------------------------------------------
// e such that max * 2^e < 2^15 (fp16 headroom for K-term fp32 accumulation), clamped.
__device__ __forceinline__ int scale_exp_of(float mx) {
    int e = 0;
    if (mx > 0.0f && isfinite(mx)) {
        int p;
        (void)frexpf(mx, &p);  // mx < 2^p
        e = 15 - p;
        e = max(-100, min(100, e));
    }
    return e;
}

template <int PREC, bool TRANS, bool LO>
__device__ __forceinline__ uint16_t pack_one(float v) {
    typedef typename HalfT<PREC>::T HT;
    HT h = (HT)v;
    if constexpr (LO) h = (HT)((v - (float)h) * 2048.0f);
    return __builtin_bit_cast(uint16_t, h);
}

// mx_in (may be null): the exponent comes from the largest magnitude folded into *mx_in by the
// kernel that produced S (instead of from e_dev); workgroup (0, 0) then stores it at e_out for
// the GEMM and clears *mx_clear, the word the next producer folds into.
template <int PREC, bool TRANS, bool LO>
__global__ __launch_bounds__(256) void pack_kernel(i64 R, i64 K, i64 Kpad, const float* __restrict__ S,
                                                   i64 lds, float sign, const int* e_dev,
                                                   uint16_t* __restrict__ H, i64 ldh,
                                                   const unsigned* mx_in, unsigned* mx_clear,
                                                   int* e_out) {
    __shared__ float tile[64][65];
    int e = e_dev ? *e_dev : 0;
    if (mx_in) {
        e = scale_exp_of(__uint_as_float(*mx_in));
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
            *e_out = e;
            *mx_clear = 0u;
        }
    }
    const float sc = sign * __builtin_ldexpf(1.0f, e);
    const i64 r0 = (i64)blockIdx.x * 64, k0 = (i64)blockIdx.y * 64;
    const int t = threadIdx.x;
    if (!TRANS) {
        // read: 64 k-columns x 64 consecutive rows
        for (int idx = t; idx < 64 * 64; idx += 256) {
            const int rr = idx & 63, kk = idx >> 6;
            const i64 r = r0 + rr, k = k0 + kk;
            tile[kk][rr] = (r < R && k < K) ? S[r + k * lds] : 0.0f;
        }
        __syncthreads();
        for (int idx = t; idx < 64 * 64; idx += 256) {
            const int kk = idx & 63, rr = idx >> 6;
            const i64 r = r0 + rr, k = k0 + kk;
            if (r < R && k < Kpad) H[r * ldh + k] = pack_one<PREC, TRANS, LO>(sc * tile[kk][rr]);
        }
    } else {
        for (int idx = t; idx < 64 * 64; idx += 256) {
            const int kk = idx & 63, rr = idx >> 6;
            const i64 r = r0 + rr, k = k0 + kk;
            if (r < R && k < Kpad) {
                const float v = (k < K) ? S[k + r * lds] : 0.0f;
                H[r * ldh + k] = pack_one<PREC, TRANS, LO>(sc * v);
            }
        }
    }
}

// e_out = scale_exp_of(max |S(i,j)|) over an m x n column-major block. Workgroups take whole
// columns (rows walked by the 256 threads, no index division) and fold their maxima into
// ctl[0] (as the bits of a non-negative float); the last workgroup to finish (ctl[1] counts)
// turns the maximum into the exponent. ctl[0..1] must start at zero.
__global__ __launch_bounds__(256) void absmax_exp_kernel(i64 m, i64 n, const float* __restrict__ S,
                                                         i64 lds, unsigned* ctl, int* e_out) {
    float mx = 0.0f;
    for (i64 j = blockIdx.x; j < n; j += gridDim.x) {
        const float* col = S + j * lds;
        for (i64 i = threadIdx.x; i < m; i += 256) mx = fmaxf(mx, fabsf(col[i]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    __shared__ float w[4];
    __shared__ bool last;
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        mx = fmaxf(fmaxf(w[0], w[1]), fmaxf(w[2], w[3]));
        atomicMax(ctl, __float_as_uint(mx));
        __threadfence();
        last = (atomicAdd(ctl + 1, 1u) == gridDim.x - 1);
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        const unsigned bits = __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *e_out = scale_exp_of(__uint_as_float(bits));
        // every other workgroup has added its maximum and its count: leave both words zero
        // for the next launch
        __hip_atomic_store(ctl, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(ctl + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

}  // namespace

namespace elx {

elx_status_t launch_lpgemm_packed(hipStream_t s, int prec, i64 m, i64 n, i64 k, float alpha,
                                  const uint16_t* A, i64 lda, const uint16_t* B, i64 ldb,
                                  float* C, i64 ldc, const int* ea_dev, const int* eb_dev) {
    if (m <= 0 || n <= 0 || k <= 0) return ELX_SUCCESS;
    ELX_ARG(lda % 8 == 0 && ldb % 8 == 0 && lda >= ((k + BK - 1) / BK) * BK &&
            ldb >= ((k + BK - 1) / BK) * BK);
    const i64 tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    ELX_ARG(tm * tn < ((i64)1 << 30));
    const unsigned blocks = (unsigned)(((tm * tn + 7) / 8) * 8);
    const void* Bv = B;
    // 64-deep slabs for the rank-NB updates (their K extent is a multiple of 64)
    const bool deep = k >= 512 && k % 64 == 0;
    ProfScope prof_(s, kKLpgemm, 2.0 * (double)m * (double)n * (double)k);
#define ELX_LPG(P, D)                                                                              \
    hipLaunchKernelGGL((lpgemm_kernel<P, false, D>), dim3(blocks), dim3(NT), 0, s, m, n, k, alpha, \
                       A, lda, Bv, ldb, C, ldc, ea_dev, eb_dev, (int)tm, (int)tn)
    if (prec == 0 && deep) ELX_LPG(0, 64);
    else if (prec == 0) ELX_LPG(0, 32);
    else if (deep) ELX_LPG(1, 64);
    else ELX_LPG(1, 32);
#undef ELX_LPG
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_lpgemm_bf32(hipStream_t s, int prec, i64 m, i64 n, i64 k, float alpha,
                                const uint16_t* A, i64 lda, const float* B, i64 ldb, float* C,
                                i64 ldc, const int* ea_dev) {
    if (m <= 0 || n <= 0 || k <= 0) return ELX_SUCCESS;
    ELX_ARG(lda % 8 == 0 && ldb % 4 == 0 && lda >= ((k + BK - 1) / BK) * BK &&
            ldb >= ((k + BK - 1) / BK) * BK && reinterpret_cast<uintptr_t>(B) % 16 == 0);
    const i64 tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    ELX_ARG(tm * tn < ((i64)1 << 30));
    ProfScope prof_(s, kKLpgemm, 2.0 * (double)m * (double)n * (double)k);
    const unsigned blocks = (unsigned)(((tm * tn + 7) / 8) * 8);
    const void* Bv = B;
    const int* eb = nullptr;
    if (prec == 0)
        hipLaunchKernelGGL((lpgemm_kernel<0, true, 32>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, Bv, ldb, C, ldc, ea_dev, eb, (int)tm, (int)tn);
    else
        hipLaunchKernelGGL((lpgemm_kernel<1, true, 32>), dim3(blocks), dim3(NT), 0, s, m, n, k,
                           alpha, A, lda, Bv, ldb, C, ldc, ea_dev, eb, (int)tm, (int)tn);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

elx_status_t launch_pack(hipStream_t s, int prec, bool trans, i64 R, i64 K, i64 Kpad,
                         const float* S, i64 lds, float sign, const int* e_dev, uint16_t* H,
                         i64 ldh, bool lo, const unsigned* mx_in, unsigned* mx_clear,
                         int* e_out) {
    if (R <= 0 || Kpad <= 0) return ELX_SUCCESS;
    ELX_ARG(ldh >= Kpad && (prec == 0 || !lo));
    dim3 grid(cdiv(R, 64), cdiv(Kpad, 64));
    ELX_ARG(grid.y <= 65535);
    ProfScope prof_(s, kKPack, (double)R * (double)Kpad * 6.0);
#define ELX_PACK(P, TR, LO_)                                                                   \
    hipLaunchKernelGGL((pack_kernel<P, TR, LO_>), grid, dim3(256), 0, s, R, K, Kpad, S, lds, sign, \
                       e_dev, H, ldh, mx_in, mx_clear, e_out)
    if (prec == 0) {
        if (lo) {
            if (trans) ELX_PACK(0, true, true);
            else ELX_PACK(0, false, true);
        } else {
            if (trans) ELX_PACK(0, true, false);
            else ELX_PACK(0, false, false);
        }
    } else {
        if (trans) ELX_PACK(1, true, false);
        else ELX_PACK(1, false, false);
    }
#undef ELX_PACK
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

// Exponent e (device int) with max|S| * 2^e < 2^15; ctl must point at two zeroed unsigneds.
elx_status_t launch_scale_exponent(hipStream_t s, i64 m, i64 n, const float* S, i64 lds,
                                   unsigned* ctl, int* e_dev) {
    const unsigned blocks = (unsigned)std::max<i64>(1, std::min<i64>(n, 2048));
    hipLaunchKernelGGL(absmax_exp_kernel, dim3(blocks), dim3(256), 0, s, std::max<i64>(m, 0),
                       std::max<i64>(n, 0), S, lds, ctl, e_dev);
    ELX_HIP(hipGetLastError());
    return ELX_SUCCESS;
}

}  // namespace elx
