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
#include <cstdlib>
#include <cstring>

#include "elx_common.hpp"
#include "exchange.hpp"

namespace {

constexpr int NT = 256;
constexpr int WMAX = 128;                 // maximum panel width
constexpr long long SPIN_LIMIT = 1 << 20;  // bounded wait: every wave exits
constexpr int RPL = 5;                     // rows per lane held in registers (R <= 320)
constexpr int XCD_WMAX = 32;               // one-XCD panel: at most 32 columns,
constexpr int XCD_RMIN = 32;               // 32 to 1024 rows per workgroup
constexpr int XCD_RMAX = 1024;
constexpr int REP = 4;                     // copies of each record (readers spread over them)
using elx::xchg::u4;

template <typename T>
struct PanelArgs {
    i64 mp;
    int w;
    T* A;
    i64 lda;
    int rowmajor;  // element (r, j) at A[r*lda + j] (1) or A[r + j*lda] (0)
    int32_t* dests;
    i64 goff;
    int2* pairs;
    int32_t* info;
    int G;
    int R;
    uint32_t seq_base;  // sequence number of column 0 of this launch
    u4* rec;            // [REP][2][G][RU] candidate records
    u4* rows;           // [3][G][WMAX] candidate row chunks
};

// The exchange area is uncached device memory accessed through the tagged 16-byte units of
// exchange.hpp.
using elx::xchg::ld16;
using elx::xchg::mix32;
using elx::xchg::st16;

// Candidate record, three 16-byte units, each checked on its own for the column's sequence
// number: {value (double bits), label, seq ^ hash} and {look-ahead entry (double bits), seq,
// hash} and {previous multiplier (double bits), seq, hash}. The units are stored by separate
// instructions whose order of arrival is not fixed, so each carries the sequence number (as a
// plain word where there is room): a unit left from an earlier column is rejected exactly,
// never through a hash comparison, and the hash only guards against a unit read while half
// written.
constexpr int RU = 3;  // 16-byte units per record
__device__ __forceinline__ uint32_t rec0_tag(u4 r0, uint32_t seq) {
    return seq ^ mix32(r0[0], r0[1], r0[2]);
}
__device__ __forceinline__ u4 rec_unit(uint32_t lo, uint32_t hi, uint32_t seq, uint32_t salt) {
    return u4{lo, hi, seq, mix32(lo, hi ^ salt, seq)};
}
__device__ __forceinline__ bool rec_unit_ok(u4 r, uint32_t seq, uint32_t salt) {
    return r[2] == seq && r[3] == mix32(r[0], r[1] ^ salt, seq);
}
__device__ __forceinline__ uint32_t dlo(double v) { return (uint32_t)(uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint32_t dhi(double v) { return (uint32_t)((uint64_t)__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double dbits(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
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

// Wave-wide argmax of (v, l) with payload q (q >= 0 marks a candidate; v >= 0, l < 2^22 - 1,
// q < 1024): every lane ends with the wave's best. The candidate is packed into an order key
// -- value bits (non-negative floats order like their bit patterns), then 2^22-1-l, then q --
// reduced by DPP butterflies inside each 16-lane row and readlanes across the four rows.
struct OKey {
    uint32_t v1, v0, l;  // value bits (high word zero for float), label key | q
};
__device__ __forceinline__ bool okey_gt(OKey a, OKey b) {
    const uint64_t av = ((uint64_t)a.v1 << 32) | a.v0, bv = ((uint64_t)b.v1 << 32) | b.v0;
    return av > bv || (av == bv && a.l > b.l);
}
template <int CTRL, bool WIDE>
__device__ __forceinline__ OKey okey_step(OKey a) {
    OKey o;
    o.v1 = WIDE ? (uint32_t)__builtin_amdgcn_update_dpp(0, (int)a.v1, CTRL, 0xF, 0xF, false) : 0u;
    o.v0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)a.v0, CTRL, 0xF, 0xF, false);
    o.l = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)a.l, CTRL, 0xF, 0xF, false);
    return okey_gt(o, a) ? o : a;
}
template <bool WIDE>
__device__ __forceinline__ OKey okey_lane(OKey a, int lane) {
    OKey o;
    o.v1 = WIDE ? (uint32_t)__builtin_amdgcn_readlane((int)a.v1, lane) : 0u;
    o.v0 = (uint32_t)__builtin_amdgcn_readlane((int)a.v0, lane);
    o.l = (uint32_t)__builtin_amdgcn_readlane((int)a.l, lane);
    return o;
}
template <typename T>
__device__ __forceinline__ void wave_best(T& v, int& l, int& q) {
    constexpr bool WIDE = sizeof(T) == 8;
    OKey k{0u, 0u, 0u};
    if (q >= 0) {
        if constexpr (WIDE) {
            const uint64_t bits = (uint64_t)__double_as_longlong((double)v);
            k.v1 = (uint32_t)(bits >> 32);
            k.v0 = (uint32_t)bits;
        } else {
            k.v0 = __float_as_uint((float)v);
        }
        k.l = ((0x3FFFFFu - (uint32_t)l) << 10) | (uint32_t)q;
    }
    k = okey_step<0xB1, WIDE>(k);   // quad_perm [1,0,3,2]
    k = okey_step<0x4E, WIDE>(k);   // quad_perm [2,3,0,1]
    k = okey_step<0x141, WIDE>(k);  // row_half_mirror
    k = okey_step<0x140, WIDE>(k);  // row_mirror
    const OKey r0 = okey_lane<WIDE>(k, 0), r1 = okey_lane<WIDE>(k, 16);
    const OKey r2 = okey_lane<WIDE>(k, 32), r3 = okey_lane<WIDE>(k, 48);
    const OKey b01 = okey_gt(r1, r0) ? r1 : r0, b23 = okey_gt(r3, r2) ? r3 : r2;
    k = okey_gt(b23, b01) ? b23 : b01;
    if (k.l == 0u) {
        v = T(-1);
        l = INT_MAX;
        q = -1;
    } else {
        if constexpr (WIDE)
            v = (T)__longlong_as_double((long long)(((uint64_t)k.v1 << 32) | k.v0));
        else
            v = (T)__uint_as_float(k.v0);
        l = (int)(0x3FFFFFu - (k.l >> 10));
        q = (int)(k.l & 1023u);
    }
}

// Best of the NW per-wave results in LDS.
template <typename T>
__device__ __forceinline__ void pick(const T* wv, const int* wl, const int* wq, T& v, int& l, int& q) {
    v = wv[0];
    l = wl[0];
    q = wq[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w)
        if (better(wv[w], wl[w], v, l)) {
            v = wv[w];
            l = wl[w];
            q = wq[w];
        }
}

// XM = 0: the workgroups are spread over all XCDs and exchange through uncached memory.
// XM = 1: the workgroups sit on one XCD (those with blockIdx % 8 == 0 of a grid dealt
// round-robin over the 8 XCDs; the rest return at once) and exchange through that XCD's L2,
// a few hundred nanoseconds per hop instead of the microseconds of a trip to memory.
template <typename T, int XM>
__global__ __launch_bounds__(NT) void lu_panel_kernel(PanelArgs<T> p) {
    constexpr int LDB = XM ? elx::xchg::kXcdLoad : elx::xchg::kSysScope;
    constexpr int STB = XM ? elx::xchg::kXcdStore : elx::xchg::kSysScope;
    constexpr int NW = NT / 64;
    constexpr int PW = NW - 1;  // waves that read the exchange; the last wave publishes
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int R = p.R, W = p.w, LD = W + 1;
    T* tile = reinterpret_cast<T*>(smem_raw);  // R x LD, row-major
    T* up = tile + (size_t)R * LD;             // WMAX: pivot row of the previous step
    T* lp = up + WMAX;                         // R: multipliers of the previous step
    T* lc = lp + R;                            // R: multipliers of this step
    T* la_v = lc + R;                          // NT per record: candidate value,
    T* la_a = la_v + NT;                       //   look-ahead entry,
    T* la_l = la_a + NT;                       //   previous multiplier
    T* wb_v = la_l + NT;                       // [2][NW] per-wave bests of the record scan
    T* rb_v = wb_v + 2 * NW;                   // [NW] per-wave bests of the row scan
    int* lbl = reinterpret_cast<int*>(rb_v + NW);  // R
    int* wb_l = lbl + R;                       // [2][NW]
    int* wb_q = wb_l + 2 * NW;                 // [2][NW]
    int* rb_l = wb_q + 2 * NW;                 // [NW]
    int* rb_q = rb_l + NW;                     // [NW]
    int* wb_ok = rb_q + NW;                    // [2][NW] per-wave votes

    if (XM && (blockIdx.x & 7) != 0) return;  // not on the panel's XCD
    const int g = XM ? (int)(blockIdx.x >> 3) : (int)blockIdx.x, t = threadIdx.x, G = p.G;
    const int wave = t >> 6, lane = t & 63;
    const i64 row0 = (i64)g * R;
    const int rows = (int)((p.mp - row0) < R ? (p.mp - row0) : R);

    // An earlier panel of this factorization met an exactly zero pivot (or gave up): the
    // reference has thrown, so this panel changes nothing and leaves identity swaps and an
    // empty move list.
    if (*p.info != 0) {
        for (int i = t; i < 2 * W; i += NT) p.pairs[i] = int2{-1, -1};
        for (int c = t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
        return;
    }

    if (rows == 128 && W == 128) {
        // the full tile: loads issued 32 at a time before their LDS stores (a load-store loop
        // would wait out each load's latency in turn); consecutive threads walk a row (row
        // major) or a column (column major)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            T v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + NT * (32 * h + u), a = idx & 127, b = idx >> 7;
                v[u] = p.rowmajor ? p.A[(row0 + b) * p.lda + a] : p.A[row0 + a + (i64)b * p.lda];
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + NT * (32 * h + u), a = idx & 127, b = idx >> 7;
                if (p.rowmajor) tile[b * LD + a] = v[u];
                else tile[a * LD + b] = v[u];
            }
        }
    } else if (p.rowmajor) {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int j = idx % W, r = idx / W;
            tile[r * LD + j] = p.A[(row0 + r) * p.lda + j];
        }
    } else {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int r = idx % rows, j = idx / rows;
            tile[r * LD + j] = p.A[row0 + r + (i64)j * p.lda];
        }
    }
    for (int r = t; r < rows; r += NT) {
        lbl[r] = (int)(row0 + r);
        lp[r] = T(0);
    }
    for (int j = t; j < WMAX; j += NT) up[j] = T(0);
    __syncthreads();

    typedef Chunk<T> CK;
    const int q1 = (W + CK::VPC - 1) / CK::VPC;
    const uint32_t nrec = (uint32_t)REP * 2u * (uint32_t)RU * (uint32_t)G;  // [REP][2][G][RU]
    const uint32_t nrows = 3u * (uint32_t)G * WMAX;  // row chunks in use: [3][G][WMAX]
    // Deferred-update pipeline. At the start of step c every live row (label >= c) holds
    // column c updated through step c-1 and columns >= c+1 updated through step c-2: the
    // step c-1 update of those columns waits for the step c-1 pivot row, which is fetched
    // while step c's records are exchanged. The record for column c carries, for its
    // candidate r: a_r[c] (exact, for the pivot choice), its label, and in two look-ahead
    // chunks a_r[c+1] through step c-2 and r's step c-1 multiplier; every reader forms the
    // pivot row's entry u_c[c+1] = a_r[c+1] - l_r u_{c-1}[c+1] with the FMA the owner would
    // use, and likewise the rest of u_c from the fetched row one step later. Each element
    // still receives its updates one FMA at a time in step order, exactly as in lu::Panel.
    //
    // Roles: waves 0..PW-1 read the exchange (records, pivot row); the last wave issues every
    // store to it, so no reader ever waits for store acknowledgements (on CDNA the vector
    // memory counter covers loads and stores alike); all waves share the LDS work.

    // Block-wide AND with a single barrier: each wave's vote goes to its slot of the current
    // parity. Every vote ends in a barrier, so no wave is ever more than one vote ahead of
    // another and the two parities never collide (the per-wave bests of the record scan use
    // the same parity).
    auto vote_all = [&](bool pred, int par) -> bool {
        const bool w = __all(pred ? 1 : 0);
        if (lane == 0) wb_ok[par * NW + wave] = w ? 1 : 0;
        __syncthreads();
        bool all = true;
#pragma unroll
        for (int i = 0; i < NW; ++i) all = all && (wb_ok[par * NW + i] != 0);
        return all;
    };

    // Publish column cc (called by the last wave): bring the candidate's columns >= cc+1
    // through step cc-2, store its row chunks, then its records.
    auto publish = [&](int cc, int cand, int cand_lbl) {
        const uint32_t seq = p.seq_base + (uint32_t)cc;
        // The record first (every workgroup waits for it); the row chunks are only read by
        // the winner's readers one step later, and carry their own tags.
        const T aprev = (cand >= 0 && cc + 1 < W)
                            ? __builtin_fma(-lp[cand], up[cc + 1], tile[cand * LD + cc + 1])
                            : T(0);
        if (lane < REP) {
            const uint32_t rec = (uint32_t)RU * (uint32_t)(((lane * 2) + (cc & 1)) * G + g);
            const double a = (double)aprev, l = (cand >= 0) ? (double)lc[cand] : 0.0;
            const double val = (cand >= 0) ? (double)tile[cand * LD + cc] : 0.0;
            u4 r0 = u4{dlo(val), dhi(val), (uint32_t)cand_lbl, 0u};
            r0[3] = rec0_tag(r0, seq);
            st16<STB>(p.rec, nrec, rec + 1, rec_unit(dlo(a), dhi(a), seq, 0x5bd1e995u));
            st16<STB>(p.rec, nrec, rec + 2, rec_unit(dlo(l), dhi(l), seq, 0x1b873593u));
            st16<STB>(p.rec, nrec, rec, r0);
        }
        if (cand >= 0) {
            const T lb = lp[cand];
            const uint32_t myrow = (uint32_t)((cc % 3) * G + g) * WMAX;
            for (int q = (cc + 1) / CK::VPC + lane; q < q1; q += 64) {
                T v[CK::VPC];
#pragma unroll
                for (int e = 0; e < CK::VPC; ++e) {
                    const int j = q * CK::VPC + e;
                    if (j > cc && j < W) {
                        v[e] = __builtin_fma(-lb, up[j], tile[cand * LD + j]);
                        tile[cand * LD + j] = v[e];
                    } else {
                        v[e] = (j < W) ? tile[cand * LD + j] : T(0);
                    }
                }
                st16<STB>(p.rows, nrows, myrow + q, CK::make(v, q, seq));
            }
        }
    };

    // Prologue: the candidate of column 0 (nothing pending: lp = 0, up = 0).
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
        wave_best(bv, bl, br);
        if (lane == 0) {
            rb_v[wave] = bv;
            rb_l[wave] = bl;
            rb_q[wave] = br;
        }
        __syncthreads();
        pick(rb_v, rb_l, rb_q, bv, bl, br);
        if (wave == NW - 1) {
            // lc is read for the candidate's multiplier: none yet
            for (int r = lane; r < rows; r += 64) lc[r] = T(0);
            publish(0, br, bl);
        }
    }

    // The bounded waits: each stage counts its own polls from zero. Stage (1) is decided per
    // wave and combined by a vote; stage (2) runs the same iterations in every wave (each ends
    // in a vote), so every wave reaches the same decision at the same barrier.
    bool aborted = false;
    int stop_col = -1;  // column of an exactly zero pivot (the panel stops there)
    int iter = 0;
    // the pivot row of the previous step, in flight: chunk registers, owner, its multiplier
    u4 fch = u4{0u, 0u, 0u, 0u};
    int f_w = -1, f_q0 = 0;
    uint32_t f_seq = 0;
    T f_l = T(0);
    for (int c = 0; c < W; ++c) {
        const int s = c & 1;
        const uint32_t seq = p.seq_base + (uint32_t)c;
        // (1) every workgroup's record for column c, reduced to the same winner. Each polling
        // wave re-reads only its missing records until all of them carry the column's tag,
        // without waiting for the other waves; one vote then combines the waves.
        T gv = T(-1);
        int gl = INT_MAX, gw = -1;
        {
            T bv = T(-1);
            int bL = INT_MAX, bq = -1;
            bool wave_aborted = false;
            long long spins = 0;
            if (wave < PW) {
                constexpr int STRIDE = PW * 64;
                bool got0 = !(t < G), got1 = !(t + STRIDE < G);
                while (true) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const bool got = k ? got1 : got0;
                        if (got) continue;
                        const int q = t + k * STRIDE;
                        const uint32_t rq = (uint32_t)RU * (uint32_t)((((g % REP) * 2) + s) * G + q);
                        const u4 r0 = ld16<LDB>(p.rec, nrec, rq);
                        const u4 r1 = ld16<LDB>(p.rec, nrec, rq + 1);
                        const u4 r2 = ld16<LDB>(p.rec, nrec, rq + 2);
                        if (r0[3] != rec0_tag(r0, seq) || !rec_unit_ok(r1, seq, 0x5bd1e995u) ||
                            !rec_unit_ok(r2, seq, 0x1b873593u))
                            continue;
                        if (k) got1 = true;
                        else got0 = true;
                        const T val = (T)dbits(r0[0], r0[1]);
                        la_v[q] = val;
                        la_a[q] = (T)dbits(r1[0], r1[1]);
                        la_l[q] = (T)dbits(r2[0], r2[1]);
                        const int L = (int)r0[2];
                        if (L != INT_MAX) {
                            const T v = fabs(val);
                            if (better(v, L, bv, bL)) {
                                bv = v;
                                bL = L;
                                bq = q;
                            }
                        }
                    }
                    if (__all((got0 && got1) ? 1 : 0)) break;
                    if (++spins > SPIN_LIMIT) {
                        wave_aborted = true;
                        break;
                    }
                }
            }
            wave_best(bv, bL, bq);
            const int par = iter & 1;
            ++iter;
            if (lane == 0) {
                wb_v[par * NW + wave] = bv;
                wb_l[par * NW + wave] = bL;
                wb_q[par * NW + wave] = bq;
            }
            if (vote_all(!wave_aborted, par))
                pick(wb_v + par * NW, wb_l + par * NW, wb_q + par * NW, gv, gl, gw);
            else
                aborted = true;
        }
        if (aborted) break;
        const bool none = (gw < 0);  // every remaining candidate is NaN (idamax selects none)
        const int piv = none ? c : gl;
        const T alpha = none ? T(0) : la_v[gw];
        const T aprev_w = none ? T(0) : la_a[gw];
        const T lprev_w = none ? T(0) : la_l[gw];
        // (2) the previous pivot row (fetched during the last step) becomes u_{c-1}: its
        // published values (through step c-3) receive the step c-2 update, in place in up
        if (f_w >= 0) {
            const int q = f_q0 + t;
            bool done = !(t < 128 && q < q1);
            long long spins = 0;
            while (true) {
                if (!done) {
                    if (!CK::ok(fch, q, f_seq))
                        fch = ld16<LDB>(p.rows, nrows, (uint32_t)(((c - 1) % 3) * G + f_w) * WMAX + q);
                    if (CK::ok(fch, q, f_seq)) {
                        T v[CK::VPC];
                        CK::get(fch, v);
#pragma unroll
                        for (int e = 0; e < CK::VPC; ++e) {
                            const int j = q * CK::VPC + e;
                            if (j >= c && j < W) up[j] = __builtin_fma(-f_l, up[j], v[e]);
                        }
                        done = true;
                    }
                }
                const int par = iter & 1;
                ++iter;
                if (vote_all(done, par)) break;
                if (++spins > SPIN_LIMIT) {
                    aborted = true;
                    break;
                }
            }
            if (aborted) break;
        } else {
            if (c > 0)
                for (int j = c + t; j < W; j += NT) up[j] = T(0);  // no pivot row (all NaN)
            __syncthreads();
        }
        // An exactly zero pivot (or a column whose candidates are all NaN): lu::Panel throws
        // SingularMatrixException here (Panel.hpp:52-53), after the row interchange of this
        // column (none: the pivot is the current row) and with every earlier step applied
        // to the whole panel. Finish the pending step c-1 update of columns c+1.. (u_{c-1} is
        // complete in up now) and stop; the columns outside the panel are not permuted.
        if (none || alpha == T(0)) {
            const int wr = W - c - 1;
            if (c > 0 && wr > 0)
                for (int idx = t; idx < rows * wr; idx += NT) {
                    const int r = idx / wr, j = c + 1 + idx % wr;
                    if (lbl[r] >= c) tile[r * LD + j] = __builtin_fma(-lp[r], up[j], tile[r * LD + j]);
                }
            if (g == 0 && t == 0) {
                p.dests[c] = (int32_t)(p.goff + piv);
                atomicCAS(p.info, 0, (int32_t)(p.goff + c + 1));
            }
            stop_col = c;
            break;
        }
        // (3) issue the fetch of this step's pivot row (consumed at the next step)
        f_w = none ? -1 : gw;
        f_seq = seq;
        f_l = lprev_w;
        f_q0 = (c + 1) / CK::VPC;
        if (f_w >= 0 && t < 128 && f_q0 + t < q1)
            fch = ld16<LDB>(p.rows, nrows, (uint32_t)((c % 3) * G + f_w) * WMAX + f_q0 + t);
        // (4) per row: the swap as a relabel (the row at position c moves to position piv,
        // the winner becomes row c), the multiplier, column c+1 through step c, and the
        // candidate of column c+1. An exactly zero pivot (the reference throws
        // SingularMatrixException here) leaves the column unscaled.
        const bool mine = (!none && gw == g);
        const int brp = br;
        const bool upd = (alpha != T(0));
        const T inv = upd ? T(1) / alpha : T(0);
        const int c1 = c + 1;
        const T upc1 = (c1 < W) ? up[c1] : T(0);
        const T uc1 = __builtin_fma(-lprev_w, upc1, aprev_w);  // u_c[c+1]
        {
            T bv = T(-1);
            int bL = INT_MAX, bR = -1;
            for (int r = t; r < rows; r += NT) {
                int L = lbl[r];
                if (L < c) continue;
                if (mine && r == brp) {
                    L = c;
                    lbl[r] = c;
                } else if (!none && L == c) {
                    L = piv;
                    lbl[r] = piv;
                }
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
                        if (better(v, L, bv, bL)) {
                            bv = v;
                            bL = L;
                            bR = r;
                        }
                    }
                }
            }
            if (g == 0 && t == 0) p.dests[c] = (int32_t)(p.goff + piv);
            wave_best(bv, bL, bR);
            if (lane == 0) {
                rb_v[wave] = bv;
                rb_l[wave] = bL;
                rb_q[wave] = bR;
            }
            __syncthreads();
            pick(rb_v, rb_l, rb_q, bv, bl, br);
        }
        if (c1 < W) {
            // (5) publish column c+1 (last wave) while the others start (6)
            if (wave == NW - 1) publish(c1, br, bl);
            // (6) the step c-1 update of columns >= c+2 for the other live rows
            const int j0 = c1 + 1;
            if (j0 < W)
                for (int r = lane; r < rows; r += 64) {
                    if (r == br || lbl[r] < c) continue;
                    const T l = lp[r];
                    T* row = tile + r * LD;
                    int j = j0 + wave;
                    for (; j + 3 * NW < W; j += 4 * NW) {
                        const T x0 = row[j], x1 = row[j + NW], x2 = row[j + 2 * NW], x3 = row[j + 3 * NW];
                        const T u0 = up[j], u1 = up[j + NW], u2 = up[j + 2 * NW], u3 = up[j + 3 * NW];
                        row[j] = __builtin_fma(-l, u0, x0);
                        row[j + NW] = __builtin_fma(-l, u1, x1);
                        row[j + 2 * NW] = __builtin_fma(-l, u2, x2);
                        row[j + 3 * NW] = __builtin_fma(-l, u3, x3);
                    }
                    for (; j < W; j += NW) row[j] = __builtin_fma(-l, up[j], row[j]);
                }
        }
        // the step c multipliers become the pending ones (the next step's exchange barrier
        // orders every use of the old ones before they are overwritten)
        T* tmp = lp;
        lp = lc;
        lc = tmp;
    }
    __syncthreads();

    if (aborted) {
        // The panel gave up (a peer never arrived): nothing was written back, so A is
        // unchanged. Every workgroup leaves the same identity swaps and an empty move list (the
        // launches queued after the panel then move nothing) and info = -1.
        for (int i = t; i < 2 * W; i += NT) p.pairs[i] = int2{-1, -1};
        for (int c = t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
        if (t == 0) atomicCAS(p.info, 0, -1);
        return;
    }
    // Rows to their final positions; the moved-row list for the other columns.
    if (rows == 128 && W == 128) {
#pragma unroll 16
        for (int u = 0; u < 64; ++u) {
            const int idx = t + NT * u, a = idx & 127, b = idx >> 7;
            if (p.rowmajor) p.A[(i64)lbl[b] * p.lda + a] = tile[b * LD + a];
            else p.A[(i64)lbl[a] + (i64)b * p.lda] = tile[a * LD + b];
        }
    } else if (p.rowmajor) {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int j = idx % W, r = idx / W;
            p.A[(i64)lbl[r] * p.lda + j] = tile[r * LD + j];
        }
    } else {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int r = idx % rows, j = idx / rows;
            p.A[(i64)lbl[r] + (i64)j * p.lda] = tile[r * LD + j];
        }
    }
    if (stop_col >= 0) {
        for (int i = t; i < 2 * W; i += NT) p.pairs[i] = int2{-1, -1};
        for (int c = stop_col + 1 + t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
        return;
    }
    for (int r = t; r < rows; r += NT) {
        const int phys = (int)(row0 + r), L = lbl[r];
        if (L < W) p.pairs[L] = (L != phys) ? int2{L, phys} : int2{-1, -1};
        if (phys < W) p.pairs[W + phys] = (L >= W) ? int2{L, phys} : int2{-1, -1};
    }
}

template <typename T>
size_t panel_lds_bytes(int R, int W) {
    return sizeof(T) * ((size_t)R * (W + 3) + WMAX + 3 * NT + 3 * (NT / 64)) +
           sizeof(int) * ((size_t)R + 8 * (NT / 64));
}

template <typename T>
elx_status_t launch_panel(hipStream_t s, i64 mp, i64 w, T* A, i64 lda, int rowmajor,
                          int32_t* dests, i64 goff, int2* pairs, int32_t* info, bool xcd) {
    ELX_ARG(w >= 1 && w <= WMAX && mp >= w);
    const elx::DeviceInfo& di = elx::device_info();
    const size_t lds_cap = (size_t)di.max_lds_per_block;
    i64 G, R;
    if (xcd) {
        // one XCD's CUs, one workgroup each, at least 32 rows per workgroup
        ELX_ARG(w <= XCD_WMAX);
        const i64 cus = std::max(di.num_cus / 8, 1);
        G = std::min<i64>(cus, (mp + XCD_RMIN - 1) / XCD_RMIN);
        if (G < 1) G = 1;
        R = (mp + G - 1) / G;
        G = (mp + R - 1) / R;
        ELX_ARG(R <= XCD_RMAX);
    } else {
        // A step costs about the same for 16 to 128 rows per workgroup, while the exchange
        // gets slower with more workgroups: short panels use fewer, fuller workgroups.
        const i64 RMIN = 128;
        G = std::min<i64>(di.num_cus, (mp + RMIN - 1) / RMIN);
        if (G < 1) G = 1;
        R = (mp + G - 1) / G;
        G = (mp + R - 1) / R;  // every workgroup owns at least one row
        ELX_ARG(R <= 64 * RPL);
    }
    ELX_ARG(G <= NT);  // one record per thread in the exchange
    ELX_ARG(mp < (1 << 22) - 1);
    const size_t lds = panel_lds_bytes<T>((int)R, (int)w);
    if (lds > lds_cap) {
        std::fprintf(stderr,
                     "elemental_amd: LU panel of height %lld needs %zu bytes of LDS per CU "
                     "(limit %zu)\n",
                     (long long)mp, lds, lds_cap);
        return ELX_ERROR_NOT_SUPPORTED;
    }
    elx::PanelSync* ps = nullptr;
    elx::ProfScope prof_(s, elx::kKPanel, 2.0 * (double)mp * (double)w * sizeof(T));
    const size_t rec_bytes = sizeof(u4) * REP * 2 * RU * (size_t)di.num_cus;
    // Row chunks rotate over three slots: a winner's row of column c is read until the
    // readers reach column c+1, while its owner may already publish column c+2.
    const size_t row_bytes = sizeof(u4) * 3 * (size_t)di.num_cus * WMAX;
    if (xcd) ELX_CHECK(elx::xcd_sync_get(di.num_cus, rec_bytes + row_bytes, &ps));
    else ELX_CHECK(elx::panel_sync_get(di.num_cus, rec_bytes + row_bytes, &ps));
    PanelArgs<T> a;
    a.mp = mp;
    a.w = (int)w;
    a.A = A;
    a.lda = lda;
    a.rowmajor = rowmajor;
    a.dests = dests;
    a.goff = goff;
    a.pairs = pairs;
    a.info = info;
    a.G = (int)G;
    a.R = (int)R;
    a.seq_base = (uint32_t)((++ps->epoch) * 1024);
    a.rec = reinterpret_cast<u4*>(ps->slots);
    a.rows = a.rec + REP * 2 * RU * (size_t)di.num_cus;
    const void* fn = xcd ? reinterpret_cast<const void*>(&lu_panel_kernel<T, 1>)
                         : reinterpret_cast<const void*>(&lu_panel_kernel<T, 0>);
    static bool attr_set[2] = {false, false};
    if (!attr_set[xcd ? 1 : 0]) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_cap);
        (void)hipGetLastError();
        attr_set[xcd ? 1 : 0] = true;
    }
    void* args[] = {&a};
    ELX_CHECK(elx::launch_resident(fn, xcd ? 8 * G : G, NT, args, lds, s));
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

// Panels exchange their candidates through uncached memory across all XCDs unless
// ELX_PANEL_EXCHANGE=xcd selects the one-XCD exchange through L2 for panels of at most 32
// columns (read at every call, so a process can switch between two complete panels).
bool xcd_panels_enabled() {
    const char* v = std::getenv("ELX_PANEL_EXCHANGE");
    return v && std::strcmp(v, "xcd") == 0;
}

template <typename T>
bool xcd_fits(i64 mp, i64 w) {
    const DeviceInfo& di = device_info();
    const i64 cus = std::max(di.num_cus / 8, 1);
    if (w < 1 || w > XCD_WMAX || mp < w || mp > XCD_RMAX * cus) return false;
    const i64 G = std::min<i64>(cus, (mp + XCD_RMIN - 1) / XCD_RMIN);
    const i64 R = (mp + G - 1) / G;
    return panel_lds_bytes<T>((int)R, (int)w) <= (size_t)di.max_lds_per_block;
}

bool lu_panel_xcd_fits(i64 mp, i64 w) { return xcd_fits<float>(mp, w); }

elx_status_t launch_lu_panel_f64(hipStream_t s, i64 mp, i64 w, double* A, i64 lda,
                                 int32_t* dests, i64 goff, int2* pairs, int32_t* info) {
    return launch_panel<double>(s, mp, w, A, lda, 0, dests, goff, pairs, info, false);
}

// Panels of at most 32 columns (Blocksize() <= 32) run on one XCD where their rows fit.
elx_status_t launch_lu_panel(hipStream_t s, i64 mp, i64 w, double* A, i64 lda, int32_t* dests,
                             i64 goff, int2* pairs, int32_t* info) {
    const bool xcd = xcd_panels_enabled() && xcd_fits<double>(mp, w);
    return launch_panel<double>(s, mp, w, A, lda, 0, dests, goff, pairs, info, xcd);
}

elx_status_t launch_lu_panel(hipStream_t s, i64 mp, i64 w, float* A, i64 lda, int32_t* dests,
                             i64 goff, int2* pairs, int32_t* info) {
    const bool xcd = xcd_panels_enabled() && xcd_fits<float>(mp, w);
    return launch_panel<float>(s, mp, w, A, lda, 0, dests, goff, pairs, info, xcd);
}

elx_status_t launch_lu_panel_f32_rowmajor(hipStream_t s, i64 mp, i64 w, float* A, i64 lda,
                                          int32_t* dests, i64 goff, int2* pairs, int32_t* info,
                                          bool xcd) {
    return launch_panel<float>(s, mp, w, A, lda, 1, dests, goff, pairs, info, xcd);
}



}  // namespace elx
