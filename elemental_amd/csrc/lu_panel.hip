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
// (value, label and look-ahead data in 16-byte records, and the candidate row in 16-byte
// chunks) to uncached memory, every workgroup reads all G records and reduces them to the
// same winner (largest magnitude, ties to the smallest label -- the first index idamax would
// return). The winner's row is read while the next column's records are exchanged, and the
// rank-1 update of the columns right of the next one is applied one step late, so one
// memory round trip per column sits on the critical path. Records and chunks carry a tag
// derived from the column's sequence number and their own contents, so a reader recognises
// stale or partly written data and simply reads again: no store-completion waits or
// separate flags. At the end each workgroup writes its rows to their final positions, and
// the list of moved rows (at most 2w) is emitted so the columns outside the panel can be
// permuted in one gather pass.
#include <algorithm>
#include <climits>

#include "elx_common.hpp"

namespace {

constexpr int NT = 256;
constexpr int WMAX = 128;                 // maximum panel width
constexpr long long SPIN_LIMIT = 1 << 20;  // bounded wait: every wave exits
constexpr int RPL = 5;                     // rows per lane held in registers (R <= 320)
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

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
    uint32_t seq_base;  // sequence number of column 0 of this launch
    u4* rec;            // [2][G][3] candidate records
    u4* rows;           // [3][G][WMAX] candidate row chunks
};

// The exchange area lives in uncached device memory (allocated with
// hipDeviceMallocUncached): loads and stores go to memory and are coherent across the XCDs
// without cache maintenance. Each 16-byte record / chunk is written and read with one
// volatile 16-byte access and validated by its tag.
__device__ __forceinline__ uint32_t mix32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ (c + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}
__device__ __forceinline__ void st16(u4* p, u4 v) { *reinterpret_cast<volatile u4*>(p) = v; }
__device__ __forceinline__ u4 ld16(const u4* p) { return *reinterpret_cast<const volatile u4*>(p); }

// Candidate record: value (as double bits), label, tag.
__device__ __forceinline__ u4 make_rec(double v, int lbl, uint32_t seq) {
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)bits, hi = (uint32_t)(bits >> 32);
    return u4{lo, hi, (uint32_t)lbl, seq ^ mix32(lo, hi, (uint32_t)lbl)};
}
__device__ __forceinline__ bool rec_ok(u4 r, uint32_t seq) { return r[3] == (seq ^ mix32(r[0], r[1], r[2])); }
__device__ __forceinline__ double rec_val(u4 r) {
    return __longlong_as_double((long long)(((uint64_t)r[1] << 32) | r[0]));
}

// Row chunk q: VPC values (one double, or three floats) and a tag.
template <typename T>
struct Chunk;
template <>
struct Chunk<double> {
    static constexpr int VPC = 1;
    __device__ static u4 make(const double* v, int q, uint32_t seq) {
        const uint64_t bits = (uint64_t)__double_as_longlong(v[0]);
        const uint32_t lo = (uint32_t)bits, hi = (uint32_t)(bits >> 32);
        return u4{lo, hi, (uint32_t)q, seq ^ mix32(lo, hi, (uint32_t)q)};
    }
    __device__ static bool ok(u4 c, int q, uint32_t seq) {
        return c[2] == (uint32_t)q && c[3] == (seq ^ mix32(c[0], c[1], (uint32_t)q));
    }
    __device__ static void get(u4 c, double* v) {
        v[0] = __longlong_as_double((long long)(((uint64_t)c[1] << 32) | c[0]));
    }
};
template <>
struct Chunk<float> {
    static constexpr int VPC = 3;
    __device__ static u4 make(const float* v, int q, uint32_t seq) {
        const uint32_t a = __float_as_uint(v[0]), b = __float_as_uint(v[1]), c = __float_as_uint(v[2]);
        return u4{a, b, c, seq ^ mix32(a, b, c ^ ((uint32_t)q * 0x27D4EB2Fu))};
    }
    __device__ static bool ok(u4 c, int q, uint32_t seq) {
        return c[3] == (seq ^ mix32(c[0], c[1], c[2] ^ ((uint32_t)q * 0x27D4EB2Fu)));
    }
    __device__ static void get(u4 c, float* v) {
        v[0] = __uint_as_float(c[0]);
        v[1] = __uint_as_float(c[1]);
        v[2] = __uint_as_float(c[2]);
    }
};

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
    T* lv = urow + WMAX;                        // R + WMAX + R + 3*NT: pipeline buffers
    T* sv = lv + 2 * (size_t)R + WMAX + 3 * NT; // NT/64
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

    typedef Chunk<T> CK;
    const int q1 = (W + CK::VPC - 1) / CK::VPC;
    // Deferred-update pipeline. At the start of step c every live row (label >= c) holds
    // column c updated through step c-1 and columns >= c+1 updated through step c-2: the
    // step c-1 update of those columns waits for the step c-1 pivot row, which is fetched
    // while step c's records are exchanged. The record for column c carries, for its
    // candidate r: a_r[c] (exact, for the pivot choice), its label, and in a look-ahead chunk
    // a_r[c+1] through step c-2 and r's step c-1 multiplier; every reader forms the pivot
    // row's entry u_c[c+1] = a_r[c+1] - l_r u_{c-1}[c+1] with the FMA the owner would use,
    // and likewise the rest of u_c from the fetched row one step later. Each element still
    // receives its updates one FMA at a time in step order, exactly as in lu::Panel.
    T* up = urow;             // u_{c-1}, columns >= c          [WMAX]
    T* un = lv + R;           // u_c being formed               [WMAX]
    T* lp = lv;               // step c-1 multipliers per row   [R]
    T* lc = un + WMAX;        // step c multipliers per row     [R]
    T* la_v = lc + R;         // per record: signed candidate value, look-ahead value and
    T* la_a = la_v + NT;      // multiplier                     [3 x NT]
    T* la_l = la_a + NT;
    for (int j = t; j < WMAX; j += NT) up[j] = T(0);
    for (int r = t; r < rows; r += NT) lp[r] = T(0);

    auto publish = [&](int cc, int cand, int cand_lbl, T aprev_v, T lprev_v) {
        const uint32_t seq = p.seq_base + (uint32_t)cc;
        const int sp = cc & 1;
        u4* myrow = p.rows + ((size_t)((cc % 3) * G + g)) * WMAX;
        if (cand >= 0)
            for (int q = (cc + 1) / CK::VPC + t; q < q1; q += NT) {
                T v[CK::VPC];
#pragma unroll
                for (int e = 0; e < CK::VPC; ++e) {
                    const int j = q * CK::VPC + e;
                    v[e] = (j < W) ? tile[cand * LD + j] : T(0);
                }
                st16(&myrow[q], CK::make(v, q, seq));
            }
        if (t == 0) {
            u4* rec = p.rec + 3 * (size_t)(sp * G + g);
            const double a = (double)aprev_v, l = (double)lprev_v;
            st16(&rec[1], Chunk<double>::make(&a, 0, seq));
            st16(&rec[2], Chunk<double>::make(&l, 1, seq));
            st16(&rec[0], make_rec((cand >= 0) ? (double)tile[cand * LD + cc] : 0.0, cand_lbl, seq));
        }
    };

    // Prologue: the candidate of column 0 (nothing pending).
    int br = -1, bl = INT_MAX;
    {
        T bv = T(-1);
        for (int r = t; r < rows; r += NT) {
            const T v = fabs(tile[r * LD + 0]);
            if (better(v, lbl[r], bv, bl)) {
                bv = v;
                bl = lbl[r];
                br = r;
            }
        }
        block_best(bv, bl, br, sv, sl, sq);
        publish(0, br, bl, (br >= 0 && W > 1) ? tile[br * LD + 1] : T(0), T(0));
    }

    bool aborted = false;
    long long spins = 0;
    // the pivot row of the previous step, in flight: chunk registers, owner, its multiplier
    u4 fch = u4{0u, 0u, 0u, 0u};
    int f_w = -1, f_q0 = 0;
    uint32_t f_seq = 0;
    T f_l = T(0);
    for (int c = 0; c < W; ++c) {
        const int s = c & 1;
        const uint32_t seq = p.seq_base + (uint32_t)c;
        // (1) every workgroup's record for column c, reduced to the same winner
        T gv = T(-1);
        int gl = INT_MAX, gw = -1;
        while (true) {
            gv = T(-1);
            gl = INT_MAX;
            gw = -1;
            int ok = 1;
            for (int q = t; q < G; q += NT) {
                const u4* rq = p.rec + 3 * (size_t)(s * G + q);
                const u4 r = ld16(&rq[0]);
                const u4 ra = ld16(&rq[1]);
                const u4 rl = ld16(&rq[2]);
                if (!rec_ok(r, seq) || !Chunk<double>::ok(ra, 0, seq) || !Chunk<double>::ok(rl, 1, seq)) {
                    ok = 0;
                    continue;
                }
                double a, l;
                Chunk<double>::get(ra, &a);
                Chunk<double>::get(rl, &l);
                const T val = (T)rec_val(r);
                la_v[q] = val;
                la_a[q] = (T)a;
                la_l[q] = (T)l;
                const int L = (int)r[2];
                if (L != INT_MAX) {
                    const T v = fabs(val);
                    if (better(v, L, gv, gl)) {
                        gv = v;
                        gl = L;
                        gw = q;
                    }
                }
            }
            if (__syncthreads_and(ok)) break;
            if (++spins > SPIN_LIMIT) {
                aborted = true;
                break;
            }
        }
        if (aborted) break;
        block_best(gv, gl, gw, sv, sl, sq);
        const bool none = (gw < 0);  // every remaining candidate is NaN (idamax selects none)
        const int piv = none ? c : gl;
        const T alpha = none ? T(0) : la_v[gw];
        const T aprev_w = none ? T(0) : la_a[gw];
        const T lprev_w = none ? T(0) : la_l[gw];
        // (2) the previous pivot row arrives: form u_{c-1} from its published values (through
        // step c-3) and its step c-2 multiplier, then make it current
        if (f_w >= 0) {
            const int q = f_q0 + t;
            const bool mine_q = (q < q1);
            while (true) {
                int ok = 1;
                if (mine_q && !CK::ok(fch, q, f_seq)) {
                    fch = ld16(p.rows + ((size_t)(((c - 1) % 3) * G + f_w)) * WMAX + q);
                    ok = CK::ok(fch, q, f_seq) ? 1 : 0;
                }
                if (__syncthreads_and(ok)) break;
                if (++spins > SPIN_LIMIT) {
                    aborted = true;
                    break;
                }
            }
            if (aborted) break;
            if (mine_q) {
                T v[CK::VPC];
                CK::get(fch, v);
#pragma unroll
                for (int e = 0; e < CK::VPC; ++e) {
                    const int j = q * CK::VPC + e;
                    if (j >= c && j < W) un[j] = __builtin_fma(-f_l, up[j], v[e]);
                }
            }
            __syncthreads();
            for (int j = c + t; j < W; j += NT) up[j] = un[j];
        } else if (c > 0) {
            for (int j = c + t; j < W; j += NT) up[j] = T(0);  // no pivot row (all NaN)
        }
        // (3) issue the fetch of this step's pivot row (consumed at the next step)
        f_w = none ? -1 : gw;
        f_seq = seq;
        f_l = lprev_w;
        f_q0 = (c + 1) / CK::VPC;
        if (f_w >= 0 && f_q0 + t < q1)
            fch = ld16(p.rows + ((size_t)((c % 3) * G + f_w)) * WMAX + f_q0 + t);
        // (4) swap = relabel: the row at position c moves to position piv, the winner becomes
        // row c
        const bool mine = (!none && gw == g);
        for (int r = t; r < rows; r += NT) {
            const int L = lbl[r];
            if (mine && r == br) lbl[r] = c;
            else if (!none && L == c) lbl[r] = piv;
        }
        if (g == 0 && t == 0) {
            p.dests[c] = (int32_t)(p.goff + piv);
            if (none || alpha == T(0)) atomicCAS(p.info, 0, (int32_t)(p.goff + c + 1));
        }
        __syncthreads();
        // An exactly zero pivot (the reference throws SingularMatrixException here) leaves
        // the column unscaled and the rest of the panel unchanged.
        const bool upd = (alpha != T(0));
        const T inv = upd ? T(1) / alpha : T(0);
        const int c1 = c + 1;
        const T upc1 = (c1 < W) ? up[c1] : T(0);
        const T uc1 = __builtin_fma(-lprev_w, upc1, aprev_w);  // u_c[c+1]
        // (5) multipliers; column c+1 through step c; candidate of column c+1
        br = -1;
        bl = INT_MAX;
        {
            T bv = T(-1);
            for (int r = t; r < rows; r += NT) {
                const int L = lbl[r];
                if (L < c) continue;
                T a = (c1 < W) ? __builtin_fma(-lp[r], upc1, tile[r * LD + c1]) : T(0);
                T l = T(0);
                if (L > c && upd) {
                    l = tile[r * LD + c] * inv;
                    tile[r * LD + c] = l;
                    a = __builtin_fma(-l, uc1, a);
                }
                lc[r] = l;
                if (c1 < W) {
                    tile[r * LD + c1] = a;
                    if (L > c) {
                        const T v = fabs(a);
                        if (better(v, L, bv, bl)) {
                            bv = v;
                            bl = L;
                            br = r;
                        }
                    }
                }
            }
            block_best(bv, bl, br, sv, sl, sq);
        }
        if (c1 < W) {
            // (6) bring the candidate's columns >= c+2 through step c-1, publish column c+1
            if (br >= 0)
                for (int j = c1 + 1 + t; j < W; j += NT)
                    tile[br * LD + j] = __builtin_fma(-lp[br], up[j], tile[br * LD + j]);
            __syncthreads();
            publish(c1, br, bl, (br >= 0 && c1 + 1 < W) ? tile[br * LD + c1 + 1] : T(0),
                    (br >= 0) ? lc[br] : T(0));
            // (7) the step c-1 update of columns >= c+2 for the other live rows, while the
            // exchange is in flight
            const int lane = t & 63, wv = t >> 6;
            for (int r = lane; r < rows; r += 64) {
                if (lbl[r] < c || r == br) continue;
                const T l = lp[r];
                for (int j = c1 + 1 + wv; j < W; j += NT / 64)
                    tile[r * LD + j] = __builtin_fma(-l, up[j], tile[r * LD + j]);
            }
        }
        // the step c multipliers become the pending ones
        __syncthreads();
        for (int r = t; r < rows; r += NT) lp[r] = lc[r];
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
    return sizeof(T) * ((size_t)R * (W + 3) + 2 * WMAX + 3 * NT + NT / 64) +
           sizeof(int) * ((size_t)R + 2 * (NT / 64) + 1);
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
    ELX_ARG(G <= NT);      // one record per thread in the exchange
    const size_t lds = panel_lds_bytes<T>((int)R, (int)w);
    if (lds > lds_cap || R > 64 * RPL) {
        std::fprintf(stderr,
                     "elemental_amd: LU panel of height %lld needs %zu bytes of LDS per CU "
                     "(limit %zu)\n",
                     (long long)mp, lds, lds_cap);
        return ELX_ERROR_NOT_SUPPORTED;
    }
    elx::PanelSync* ps = nullptr;
    elx::ProfScope prof_(s, elx::kKPanel, 2.0 * (double)mp * (double)w * sizeof(T));
    const size_t rec_bytes = sizeof(u4) * 3 * 2 * (size_t)di.num_cus;
    // Row chunks rotate over three slots: a winner's row of column c is read until the
    // readers reach column c+1, while its owner may already publish column c+2.
    const size_t row_bytes = sizeof(u4) * 3 * (size_t)di.num_cus * WMAX;
    ELX_CHECK(elx::panel_sync_get(di.num_cus, rec_bytes + row_bytes, &ps));
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
    a.seq_base = (uint32_t)((++ps->epoch) * 1024);
    a.rec = reinterpret_cast<u4*>(ps->slots);
    a.rows = a.rec + 3 * 2 * (size_t)di.num_cus;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&lu_panel_kernel<T>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cap);
        (void)hipGetLastError();
        attr_set = true;
    }
    ELX_CHECK(elx::coop_fits(reinterpret_cast<const void*>(&lu_panel_kernel<T>), NT, lds, G));
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
