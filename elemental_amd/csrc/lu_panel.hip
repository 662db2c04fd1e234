// Partial-pivoting LU panel as one cooperative, LDS-resident kernel.
//
// Restates lu::Panel (src/lapack_like/factor/LU/Panel.hpp:15-60): for every column c of the
// mp x w panel, take the first row of largest |a(i,c)| among rows c..mp-1 (blas::MaxInd =
// idamax, src/core/imports/blas/MaxInd.hpp:26-46), swap it with row c, scale the column
// below by the reciprocal of the pivot (x * (1/alpha), Panel.hpp:54-55) and apply the rank-1
// update a(i,j) += l(i) * (-u(j)) as one FMA per entry (dger, Panel.hpp:56-58).
//
// Design for MI355X: the panel is spread over G workgroups (at most one per CU), each holding
// R consecutive rows in LDS for the whole panel, so the 32 MiB panel of an n=32768 FP64 LU
// never goes back to HBM between columns. Rows are never moved during the panel: a row's
// "label" is its position in the swapped order, and a swap relabels two rows. Each column
// step is one all-to-all exchange: every workgroup publishes its local pivot candidate
// (|value|, label, and the candidate row) with an agent-scope release, waits until all
// workgroups have published for this column, and then every workgroup reduces the same
// candidates to the same winner (largest magnitude, ties to the smallest label -- the
// first index idamax would return). At the end each workgroup writes its rows to their
// final positions, and the list of moved rows (at most 2w) is emitted so the columns
// outside the panel can be permuted in one gather pass.
#include <algorithm>
#include <climits>

#include "elx_common.hpp"

namespace {

constexpr int NT = 256;
constexpr int WMAX = 128;                 // maximum panel width
constexpr long long SPIN_LIMIT = 1 << 20;  // bounded wait: every wave exits
constexpr int RPL = 5;                     // rows per lane held in registers (R <= 320)

template <typename T>
struct PanelArgs {
    i64 mp;
    int w;
    T* A;
    i64 lda;
    int32_t* dests;
    i64 goff;
    int2* pairs;
    int32_t* info;
    int G;
    int R;
    int64_t* flags;
    int64_t epoch_base;
    T* slot_val;     // [2][G] signed pivot candidate value
    int* slot_lbl;   // [2][G] candidate label (INT_MAX when none)
    T* slot_row;     // [2][G][WMAX] candidate row (columns c..w-1 valid)
};

// The exchange area lives in uncached device memory (allocated with
// hipDeviceMallocUncached), so loads and stores go straight to memory and are coherent across
// the XCDs without L2 write-back / invalidate. Accesses are relaxed agent-scope atomics (the
// compiler may neither cache nor merge them); ordering of a workgroup's data before its flag
// comes from waiting for the stores to complete (s_waitcnt) before the flag store.
template <typename T>
__device__ __forceinline__ void st_x(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_x(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_complete() { __builtin_amdgcn_s_waitcnt(0); }

template <typename T>
__device__ __forceinline__ bool better(T v, int l, T bv, int bl) {
    return (v > bv) || (v == bv && l < bl);
}

// Block-wide argmax of (v, l) with payload q; result broadcast to all threads.
template <typename T>
__device__ __forceinline__ void block_best(T& v, int& l, int& q, T* sv, int* sl, int* sq) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(v, off);
        const int ol = __shfl_xor(l, off);
        const int oq = __shfl_xor(q, off);
        if (better(ov, ol, v, l)) {
            v = ov;
            l = ol;
            q = oq;
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sv[wave] = v;
        sl[wave] = l;
        sq[wave] = q;
    }
    __syncthreads();
    v = sv[0];
    l = sl[0];
    q = sq[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w)
        if (better(sv[w], sl[w], v, l)) {
            v = sv[w];
            l = sl[w];
            q = sq[w];
        }
    __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(NT) void lu_panel_kernel(PanelArgs<T> p) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int R = p.R, W = p.w, LD = W + 1;
    T* tile = reinterpret_cast<T*>(smem_raw);   // R x LD, row-major
    T* urow = tile + (size_t)R * LD;            // WMAX
    T* sv = urow + WMAX;                        // NT/64
    int* lbl = reinterpret_cast<int*>(sv + NT / 64);  // R
    int* sl = lbl + R;                          // NT/64
    int* sq = sl + NT / 64;                     // NT/64
    int* sflag = sq + NT / 64;                  // 1

    const int g = blockIdx.x, t = threadIdx.x, G = p.G;
    const i64 row0 = (i64)g * R;
    const int rows = (int)((p.mp - row0) < R ? (p.mp - row0) : R);

    for (int idx = t; idx < rows * W; idx += NT) {
        const int r = idx % rows, j = idx / rows;
        tile[r * LD + j] = p.A[row0 + r + (i64)j * p.lda];
    }
    for (int r = t; r < rows; r += NT) lbl[r] = (int)(row0 + r);
    __syncthreads();

    bool aborted = false;
    for (int c = 0; c < W; ++c) {
        const int s = c & 1;
        // Local candidate: largest |a(r,c)| over rows not yet used as pivots.
        T bv = T(-1);
        int bl = INT_MAX, br = -1;
        for (int r = t; r < rows; r += NT) {
            const int L = lbl[r];
            if (L >= c) {
                const T v = fabs(tile[r * LD + c]);
                if (better(v, L, bv, bl)) {
                    bv = v;
                    bl = L;
                    br = r;
                }
            }
        }
        block_best(bv, bl, br, sv, sl, sq);

        // Publish the candidate (value with sign, label, row) for this column, then the flag.
        T* myrow = p.slot_row + ((size_t)(s * G + g)) * WMAX;
        if (br >= 0)
            for (int j = c + t; j < W; j += NT) st_x(&myrow[j], tile[br * LD + j]);
        if (t == 0) {
            st_x(&p.slot_val[s * G + g], (br >= 0) ? tile[br * LD + c] : T(0));
            st_x(&p.slot_lbl[s * G + g], bl);
        }
        stores_complete();
        __syncthreads();
        const int64_t target = p.epoch_base + c + 1;
        if (t == 0) st_x(&p.flags[g], target);

        // Wait until every workgroup has published column c.
        long long spins = 0;
        while (true) {
            int ok = 1;
            for (int q = t; q < G; q += NT) ok &= (ld_x(&p.flags[q]) >= target);
            if (__syncthreads_and(ok)) break;
            if (++spins > SPIN_LIMIT) {
                aborted = true;
                break;
            }
        }
        if (aborted) break;

        // Every workgroup reduces the same G candidates to the same winner.
        T gv = T(-1);
        int gl = INT_MAX, gw = -1;
        for (int q = t; q < G; q += NT) {
            const int L = ld_x(&p.slot_lbl[s * G + q]);
            if (L != INT_MAX) {
                const T v = fabs(ld_x(&p.slot_val[s * G + q]));
                if (better(v, L, gv, gl)) {
                    gv = v;
                    gl = L;
                    gw = q;
                }
            }
        }
        block_best(gv, gl, gw, sv, sl, sq);
        if (gw < 0) {
            // Every remaining candidate is NaN: no row qualifies (idamax never selects a
            // NaN). Record the step without a swap and flag the panel.
            if (g == 0 && t == 0) {
                p.dests[c] = (int32_t)(p.goff + c);
                atomicCAS(p.info, 0, (int32_t)(p.goff + c + 1));
            }
            __syncthreads();
            continue;
        }
        const int piv = gl;
        const T* wrow = p.slot_row + ((size_t)(s * G + gw)) * WMAX;
        for (int j = c + t; j < W; j += NT) urow[j] = ld_x(&wrow[j]);
        // Swap = relabel: the row at position c moves to position piv ...
        for (int r = t; r < rows; r += NT)
            if (lbl[r] == c) lbl[r] = piv;
        __syncthreads();
        // ... and the winner becomes row c.
        if (gw == g && t == 0) lbl[br] = c;
        if (g == 0 && t == 0) p.dests[c] = (int32_t)(p.goff + piv);
        __syncthreads();

        const T alpha = urow[c];
        if (alpha == T(0)) {
            // Exactly zero pivot: the reference throws SingularMatrixException here.
            if (g == 0 && t == 0) atomicCAS(p.info, 0, (int32_t)(p.goff + c + 1));
            continue;
        }
        const T inv = T(1) / alpha;
        // Multipliers l = a(r,c) * (1/alpha) into registers (every wave needs them) ...
        const int lane = t & 63, wv = t >> 6;
        T lq[RPL];
        bool act[RPL];
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            const int r = lane + 64 * q;
            act[q] = (r < rows) && (lbl[r] > c);
            lq[q] = act[q] ? tile[r * LD + c] * inv : T(0);
        }
        __syncthreads();
        // ... then wave 0 stores them and all waves apply the rank-1 update, wave wv on
        // columns j = c+1+wv (mod 4).
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
            if (!act[q]) continue;
            const int r = lane + 64 * q;
            if (wv == 0) tile[r * LD + c] = lq[q];
            for (int j = c + 1 + wv; j < W; j += NT / 64)
                tile[r * LD + j] = __builtin_fma(-lq[q], urow[j], tile[r * LD + j]);
        }
        __syncthreads();
    }

    if (aborted) {
        if (t == 0) atomicCAS(p.info, 0, -1);
        return;
    }
    // Rows to their final positions; the moved-row list for the other columns.
    for (int idx = t; idx < rows * W; idx += NT) {
        const int r = idx % rows, j = idx / rows;
        p.A[(i64)lbl[r] + (i64)j * p.lda] = tile[r * LD + j];
    }
    for (int r = t; r < rows; r += NT) {
        const int phys = (int)(row0 + r), L = lbl[r];
        if (L < W) p.pairs[L] = (L != phys) ? int2{L, phys} : int2{-1, -1};
        if (phys < W) p.pairs[W + phys] = (L >= W) ? int2{L, phys} : int2{-1, -1};
    }
    (void)sflag;
}

template <typename T>
size_t panel_lds_bytes(int R, int W) {
    return sizeof(T) * ((size_t)R * (W + 1) + WMAX + NT / 64) + sizeof(int) * ((size_t)R + 2 * (NT / 64) + 1);
}

template <typename T>
elx_status_t launch_panel(hipStream_t s, i64 mp, i64 w, T* A, i64 lda, int32_t* dests, i64 goff,
                          int2* pairs, int32_t* info) {
    ELX_ARG(w >= 1 && w <= WMAX && mp >= w);
    const elx::DeviceInfo& di = elx::device_info();
    const size_t lds_cap = (size_t)di.max_lds_per_block;
    const i64 RMIN = 16;
    i64 G = std::min<i64>(di.num_cus, (mp + RMIN - 1) / RMIN);
    if (G < 1) G = 1;
    i64 R = (mp + G - 1) / G;
    G = (mp + R - 1) / R;  // every workgroup owns at least one row
    const size_t lds = panel_lds_bytes<T>((int)R, (int)w);
    if (lds > lds_cap || R > 64 * RPL) {
        std::fprintf(stderr,
                     "elemental_amd: LU panel of height %lld needs %zu bytes of LDS per CU "
                     "(limit %zu)\n",
                     (long long)mp, lds, lds_cap);
        return ELX_ERROR_NOT_SUPPORTED;
    }
    elx::PanelSync* ps = nullptr;
    ELX_CHECK(elx::panel_sync_get(di.num_cus, sizeof(T) * (size_t)2 * di.num_cus * (WMAX + 1) +
                                                  sizeof(int) * 2 * di.num_cus,
                                  &ps));
    PanelArgs<T> a;
    a.mp = mp;
    a.w = (int)w;
    a.A = A;
    a.lda = lda;
    a.dests = dests;
    a.goff = goff;
    a.pairs = pairs;
    a.info = info;
    a.G = (int)G;
    a.R = (int)R;
    a.flags = ps->flags;
    a.epoch_base = (++ps->epoch) * 1024;
    a.slot_val = reinterpret_cast<T*>(ps->slots);
    a.slot_row = a.slot_val + 2 * di.num_cus;
    a.slot_lbl = reinterpret_cast<int*>(a.slot_row + (size_t)2 * di.num_cus * WMAX);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&lu_panel_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cap);
        (void)hipGetLastError();
        attr_set = true;
    }
    void* args[] = {&a};
    ELX_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&lu_panel_kernel<T>),
                                       dim3((unsigned)G), dim3(NT), args, (unsigned)lds, s));
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

elx_status_t launch_lu_panel_f64(hipStream_t s, i64 mp, i64 w, double* A, i64 lda,
                                 int32_t* dests, i64 goff, int2* pairs, int32_t* info) {
    return launch_panel<double>(s, mp, w, A, lda, dests, goff, pairs, info);
}

elx_status_t launch_lu_panel_f32(hipStream_t s, i64 mp, i64 w, float* A, i64 lda,
                                 int32_t* dests, i64 goff, int2* pairs, int32_t* info) {
    return launch_panel<float>(s, mp, w, A, lda, dests, goff, pairs, info);
}

}  // namespace elx
