// Tournament-pivoting LU panel for the mixed-precision factorization (communication-avoiding
// LU, "CALU": L. Grigori, J. Demmel, H. Xiang, SIAM J. Matrix Anal. Appl. 32(4), 2011).
//
// The fp32 working matrix of the mixed-precision solve only has to be a good preconditioner:
// FP64 refinement recovers the accuracy, so its pivot order need not be lu::Panel's. The plain
// FP64 path keeps partial pivoting (lu_panel.hip); this kernel serves the low-precision path,
// where partial pivoting's one all-to-all exchange per column (a memory round trip per column,
// 32768 of them for n = 32768) is most of the factorization time.
//
// The mp x w panel (w <= 128, stored row by row: element (r, j) at A[r*lda + j]) is spread over
// G workgroups, R <= 128 consecutive rows each, held in LDS for the whole panel. It is factored
// WS (16 or 32) columns at a time. For each WS-column block:
//  1. leaf: each workgroup chooses WS of its live rows by Gaussian elimination with partial
//     pivoting on their values in the block's columns (one wave, two rows per lane), and
//     publishes them (label and row) to uncached memory;
//  2. merge: every F-th workgroup (F = 512/WS) gathers the candidates of F leaves and chooses
//     WS of them the same way, on their original values (two rounds: four waves over 128
//     candidates each, then one wave over the survivors); with more than F leaves a second
//     level (workgroup 0) merges the level-1 winners;
//  3. root: workgroup 0's last elimination leaves the winners' L11 and U11; it reads the
//     winners' remaining columns A12 from their leaves' publications, forms
//     U12 = L11^{-1} A12 and the LAPACK-style interchanges of the winners into positions
//     k0..k0+WS-1, and publishes the block rows [L11\U11 | U12];
//  4. every workgroup relabels its rows by those interchanges (rows never move during the
//     panel: a row's label is its position), overwrites its winners with the published rows,
//     and for every other live row solves l U11 = a(k0:k0+WS) and updates
//     a(k0+WS:w) -= l U12.
// At the end each workgroup writes its rows to their final positions and the panel emits the
// same outputs as the partial-pivoting panel: the interchanges (dests, offset goff), the list
// of moved rows (pairs) for the columns outside the panel, and info.
//
// Exchange units are 16 bytes, each written and read with one buffer access and carrying a tag
// (the block's sequence number xor a hash of its contents and position), so a reader
// recognises stale or partly written units and reads them again; every wait is bounded and a
// panel that gives up leaves A unchanged and info = -1. A block whose U11 has a zero or
// non-finite diagonal entry stops the panel with info = goff + column + 1 (the mixed solve
// then falls back to FP64).
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "elx_common.hpp"
#include "exchange.hpp"

namespace {

using elx::xchg::ld16;
using elx::xchg::mix32;
using elx::xchg::st16;
using elx::xchg::u4;

constexpr int NT = 256;
constexpr int NW = NT / 64;
constexpr int WMAX = 128;                  // maximum panel width
constexpr int RMAX = 128;                  // rows per workgroup (two per lane of one wave)
constexpr int CAND = 512;                  // candidates of one merge (four waves of 128)
constexpr int VU = (WMAX + 2) / 3;         // value units of a full published row
constexpr int FULLU = 1 + VU;              // units per published row: header + values
constexpr int UL = WMAX + 4;               // row stride of the block rows in LDS
constexpr long long SPIN_LIMIT = 1 << 20;  // bounded waits: every wave exits
constexpr long long kCaluMinOrder = 4096;  // default: tournament panels from this order on
constexpr long long kCaluMaxRhs = 4;      // ... and for at most this many right-hand sides

template <int WS>
struct Blk {
    static constexpr int F = CAND / WS;       // fan-in of a merge
    static constexpr int SVU = (WS + 2) / 3;  // value units covering the block's columns
    static constexpr int SELU = 1 + SVU;      // units per candidate of a level-1 result
    static constexpr int MUPT = (CAND * SELU + NT - 1) / NT;  // merge units per thread
    static constexpr int RUPT = (WS * FULLU + NT - 1) / NT;   // root-result units per thread
    static constexpr int AUPT = (WS * VU + NT - 1) / NT;      // winner-row units per thread
    static_assert(MUPT <= 32 && RUPT <= 32 && AUPT <= 32, "one mask bit per unit");
};

struct CaluArgs {
    i64 mp;
    int w;
    float* A;
    i64 lda;
    int32_t* dests;
    i64 goff;
    int2* pairs;
    int32_t* info;
    int G;
    int R;
    int levels;         // 1: the root merges the leaves; 2: level-1 merges in between
    uint32_t seq_base;  // sequence number of block 0 of this launch
    u4* leaf;           // [G][WS][FULLU]
    u4* node;           // [ceil(G/F)][WS][SELU]
    u4* root;           // [WS][FULLU]
    uint32_t n_leaf, n_node, n_root;  // units in each area
};

__device__ __forceinline__ uint32_t salt(int kind, int unit) {
    return 0x27D4EB2Fu * (uint32_t)(kind * 256 + unit + 1);
}
__device__ __forceinline__ u4 make_unit(uint32_t a, uint32_t b, uint32_t c, uint32_t seq,
                                        uint32_t sl) {
    return u4{a, b, c, seq ^ mix32(a, b, c ^ sl)};
}
__device__ __forceinline__ bool unit_ok(u4 u, uint32_t seq, uint32_t sl) {
    return u[3] == (seq ^ mix32(u[0], u[1], u[2] ^ sl));
}
__device__ __forceinline__ uint32_t fb(float v) { return __float_as_uint(v); }
__device__ __forceinline__ float bf(uint32_t v) { return __uint_as_float(v); }

// Largest 32-bit key of the wave, in every lane: DPP butterflies inside each 16-lane row, then
// the four rows through readlane.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t k) {
    k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0xB1, 0xF, 0xF, false));
    k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x4E, 0xF, 0xF, false));
    k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x141, 0xF, 0xF, false));
    k = max(k, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k, 0x140, 0xF, 0xF, false));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)k, 0);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane((int)k, 16);
    const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane((int)k, 32);
    const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane((int)k, 48);
    return max(max(r0, r1), max(r2, r3));
}

// Compile-time loop: f(std::integral_constant<int, i>) for i = B .. E-1, so that register
// arrays are only ever indexed by constants.
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// One wave chooses ks (<= WS) of its candidates by Gaussian elimination with partial pivoting
// on their values v: candidate q sits in lane q % 64, register set q / 64 (q < 128), act marks
// the candidates. The pivot of each step is the candidate of largest magnitude (compared on
// the top 24 bits of |v|; NaN counts as zero), the first one on a tie. order[k] (LDS, written
// by lane k) receives the candidate chosen at step k. With keep the eliminated values stay in
// v: the candidate chosen at step k then holds its row of L (multipliers, columns < k) and of
// U (columns >= k). step[i] is the step at which register set i's candidate was chosen (-1:
// not chosen).
template <int WS>
__device__ __forceinline__ void wave_select(float (&v)[2][WS], bool (&act)[2], int ks, bool keep,
                                            int* order, int (&step)[2]) {
    // Straight-line steps inside (no divergent branches): rows that are not eliminated get a
    // zero multiplier, and the candidate chosen at step k is kept in lane k and stored at the
    // end. Only the wave-uniform exit past ks steps branches.
    const int lane = threadIdx.x & 63;
    step[0] = step[1] = -1;
    int qk = -1;
    bool live = true;  // wave-uniform: candidates remain
    static_for<0, WS>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        if (!live || k >= ks) return;
        uint32_t key = 0u;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float a = v[i][k];
            const float m = (a == a) ? fabsf(a) : 0.0f;
            const uint32_t ki = (fb(m) & 0xFFFFFF00u) | (uint32_t)(255 - (i * 64 + lane));
            key = act[i] ? max(key, ki) : key;
        }
        key = wave_max_u32(key);
        if (key == 0u) {  // no candidate left
            live = false;
            return;
        }
        const int q = 255 - (int)(key & 0xFFu);
        const int pl = q & 63, ps = q >> 6;
        float u[WS];
        static_for<k, WS>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const float x = (ps == 0) ? v[0][j] : v[1][j];
            u[j] = bf((uint32_t)__builtin_amdgcn_readlane((int)fb(x), pl));
        });
        // the hardware reciprocal (1 ulp): the multipliers only have to be those of a good
        // preconditioner, and the step is a dependent chain in one wave
        const float piv = u[k];
        const float inv = (piv != 0.0f) ? __builtin_amdgcn_rcpf(piv) : 0.0f;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bool chosen = (i == ps) && (lane == pl);
            const bool elim = act[i] && !chosen;
            step[i] = chosen ? k : step[i];
            act[i] = elim;
            const float l = elim ? v[i][k] * inv : 0.0f;
            if (keep) v[i][k] = elim ? l : v[i][k];
            static_for<k + 1, WS>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                v[i][j] = __builtin_fmaf(-l, u[j], v[i][j]);
            });
        }
        qk = (lane == k) ? q : qk;
    });
    if (lane < ks && qk >= 0) order[lane] = qk;
}

// LDS layout (byte offsets, 16-byte aligned sections).
struct Lds {
    size_t tile, lbl, live, stg_v, stg_lab, stg_src, surv, comp, fin, ublk, lab_w, piv, rdiag,
        misc, total;
};
__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }
template <int WS>
__host__ __device__ inline Lds lds_layout(int R, int W) {
    Lds L;
    size_t o = 0;
    L.tile = o;    o = al16(o + sizeof(float) * (size_t)R * (W + 1));
    L.lbl = o;     o = al16(o + sizeof(int) * RMAX);
    L.live = o;    o = al16(o + sizeof(int) * RMAX);
    L.stg_v = o;   o = al16(o + sizeof(float) * (size_t)CAND * (WS + 1));
    L.stg_lab = o; o = al16(o + sizeof(int) * CAND);
    L.stg_src = o; o = al16(o + sizeof(int) * CAND);
    L.surv = o;    o = al16(o + sizeof(int) * (NW * WS + NW));
    L.comp = o;    o = al16(o + sizeof(int) * NW * WS);
    L.fin = o;     o = al16(o + sizeof(int) * WS);
    L.ublk = o;    o = al16(o + sizeof(float) * (size_t)WS * UL);
    L.lab_w = o;   o = al16(o + sizeof(int) * WS);
    L.piv = o;     o = al16(o + sizeof(int) * WS);
    L.rdiag = o;   o = al16(o + sizeof(float) * WS);
    L.misc = o;    o = al16(o + sizeof(int) * 16);
    L.total = o;
    return L;
}

template <int WS>
__global__ __launch_bounds__(NT) void calu_panel_kernel(CaluArgs p) {
    using B = Blk<WS>;
    constexpr int F = B::F;
    extern __shared__ __align__(16) unsigned char smem[];
    const int R = p.R, W = p.w, LD = W + 1, G = p.G;
    const Lds L = lds_layout<WS>(R, W);
    float* tile = reinterpret_cast<float*>(smem + L.tile);        // [R][LD]
    int* lbl = reinterpret_cast<int*>(smem + L.lbl);              // [RMAX] row labels
    int* live = reinterpret_cast<int*>(smem + L.live);            // [RMAX] live tile rows
    float* stg_v = reinterpret_cast<float*>(smem + L.stg_v);      // [CAND][WS+1] values
    int* stg_lab = reinterpret_cast<int*>(smem + L.stg_lab);      // [CAND] labels (-1: none)
    int* stg_src = reinterpret_cast<int*>(smem + L.stg_src);      // [CAND] source
    int* surv = reinterpret_cast<int*>(smem + L.surv);            // [NW][WS], then NW counts
    int* comp = reinterpret_cast<int*>(smem + L.comp);            // [NW*WS] compacted
    int* fin = reinterpret_cast<int*>(smem + L.fin);              // [WS] final choice
    float* ublk = reinterpret_cast<float*>(smem + L.ublk);        // [WS][UL] block rows
    int* lab_w = reinterpret_cast<int*>(smem + L.lab_w);          // [WS] winner labels
    int* piv = reinterpret_cast<int*>(smem + L.piv);              // [WS] interchanges
    float* rdiag = reinterpret_cast<float*>(smem + L.rdiag);      // [WS] 1 / u_kk
    // misc: [0] gave up, [1] live rows, [2] bad pivot (column + 1), [3] survivors,
    // [4] chosen count
    int* misc = reinterpret_cast<int*>(smem + L.misc);

    const int g = (int)blockIdx.x, t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const i64 row0 = (i64)g * R;
    const int rows = (int)((p.mp - row0) < R ? (p.mp - row0) : R);

    // An earlier panel of this factorization stopped: identity interchanges, no moves.
    if (*p.info != 0) {
        for (int i = t; i < 2 * W; i += NT) p.pairs[i] = int2{-1, -1};
        for (int c = t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
        return;
    }
    if (W == 128 && rows == 128) {
        // the full tile: loads issued 32 at a time before their LDS stores (a load-store loop
        // would wait out each load's latency in turn)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float v[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + NT * (32 * h + u), j = idx & 127, r = idx >> 7;
                v[u] = p.A[(row0 + r) * p.lda + j];
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                const int idx = t + NT * (32 * h + u), j = idx & 127, r = idx >> 7;
                tile[r * LD + j] = v[u];
            }
        }
    } else {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int j = idx % W, r = idx / W;
            tile[r * LD + j] = p.A[(row0 + r) * p.lda + j];
        }
    }
    for (int r = t; r < RMAX; r += NT) lbl[r] = (r < rows) ? (int)(row0 + r) : INT_MIN;
    if (t < 16) misc[t] = 0;
    __syncthreads();

    bool aborted = false;
    int stop_k0 = -1;
    const int nblocks = (W + WS - 1) / WS;
    const bool is_node = (p.levels == 2) && (g % F == 0);
    const bool is_root = (g == 0);

    // Wait until every unit of a set is present: unit idx (< total) of this thread's share is
    // read from `ar` at addr(idx) until its tag matches, then handed to sink(idx, unit). All
    // threads take part; returns false if any wave gave up.
    auto gather = [&](auto upt_tag, const u4* ar, uint32_t nar, int total, uint32_t seq,
                      auto addr, auto sl, auto sink) -> bool {
        constexpr int UPT = decltype(upt_tag)::value;
        constexpr uint32_t ALL = (UPT >= 32) ? 0xFFFFFFFFu : ((1u << UPT) - 1u);
        uint32_t got = 0u;
#pragma unroll
        for (int i = 0; i < UPT; ++i)
            if (t + NT * i >= total) got |= 1u << i;
        long long spins = 0;
        bool gave_up = false;
        while (!__all(got == ALL ? 1 : 0)) {
            if (++spins > SPIN_LIMIT) {
                gave_up = true;
                break;
            }
            u4 un[UPT];
#pragma unroll
            for (int i = 0; i < UPT; ++i)
                if (!(got & (1u << i))) un[i] = ld16(ar, nar, addr(t + NT * i));
#pragma unroll
            for (int i = 0; i < UPT; ++i) {
                if (got & (1u << i)) continue;
                const int idx = t + NT * i;
                if (unit_ok(un[i], seq, sl(idx))) {
                    sink(idx, un[i]);
                    got |= 1u << i;
                }
            }
        }
        if (gave_up && lane == 0) misc[0] = 1;
        __syncthreads();
        return misc[0] == 0;
    };

    // Candidates in stg_* positions [0, ncand) chosen down to min(ws, live candidates) by one
    // or two rounds of wave eliminations (called by every thread). The choice lands in fin[] as
    // candidate positions, in step order; with keep (the root) the winners' eliminated rows go
    // to ublk[k][0, WS) and their labels to lab_w.
    auto merge = [&](int ncand, int ws, bool keep) -> int {
        const int* survp = nullptr;
        int nsurv = 0;
        if (ncand > 128) {
            // round 1: wave w over candidates [128w, 128w + 128)
            float v[2][WS];
            bool act[2];
            int stp[2];
            int cnt = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int q = wave * 128 + i * 64 + lane;
                act[i] = (q < ncand) && (stg_lab[q] >= 0);
#pragma unroll
                for (int j = 0; j < WS; ++j) v[i][j] = act[i] ? stg_v[q * (WS + 1) + j] : 0.0f;
                cnt += __popcll(__ballot(act[i] ? 1 : 0));
            }
            const int ks = min(ws, cnt);
            wave_select<WS>(v, act, ks, false, surv + wave * WS, stp);
            if (lane == 0) surv[NW * WS + wave] = ks;
            __syncthreads();
            if (t == 0) {
                int n = 0;
                for (int w2 = 0; w2 < NW; ++w2)
                    for (int k = 0; k < surv[NW * WS + w2]; ++k)
                        comp[n++] = w2 * 128 + surv[w2 * WS + k];
                misc[3] = n;
            }
            __syncthreads();
            nsurv = misc[3];
            survp = comp;
        }
        // second merge round (or the only one): wave 0 over at most 128 candidates
        if (wave == 0) {
            float v[2][WS];
            bool act[2];
            int stp[2];
            int cpos[2];
            int cnt = 0;
            const int nc = survp ? nsurv : ncand;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int q = i * 64 + lane;
                cpos[i] = (q < nc) ? (survp ? survp[q] : q) : -1;
                act[i] = (cpos[i] >= 0) && (stg_lab[cpos[i]] >= 0);
#pragma unroll
                for (int j = 0; j < WS; ++j)
                    v[i][j] = act[i] ? stg_v[cpos[i] * (WS + 1) + j] : 0.0f;
                cnt += __popcll(__ballot(act[i] ? 1 : 0));
            }
            const int ks = min(ws, cnt);
            int* ord = surv;  // wave 0's local order (round 1 is over)
            wave_select<WS>(v, act, ks, keep, ord, stp);
            if (keep) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    if (stp[i] < 0 || stp[i] >= ks) continue;
#pragma unroll
                    for (int j = 0; j < WS; ++j) ublk[stp[i] * UL + j] = v[i][j];
                    lab_w[stp[i]] = stg_lab[cpos[i]];
                }
            }
            if (lane < ks) {
                const int q = ord[lane];
                fin[lane] = survp ? survp[q] : q;
            }
            if (lane == 0) misc[4] = ks;
        }
        __syncthreads();
        return misc[4];
    };

    // sink for candidate headers and block values (leaf or level-1 units of SELU per candidate)
    auto cand_sink = [&](int idx, u4 un) {
        const int cand = idx / B::SELU, u = idx % B::SELU;
        if (u == 0) {
            stg_lab[cand] = (un[0] == 0xFFFFFFFFu) ? -1 : (int)un[0];
            stg_src[cand] = (int)un[1];
        } else {
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int j = 3 * (u - 1) + e;
                if (j < WS) stg_v[cand * (WS + 1) + j] = bf(un[e]);
            }
        }
    };

    for (int b = 0; b < nblocks; ++b) {
        const int k0 = b * WS;
        const int ws = min(WS, W - k0);
        const int nu = (W - k0 + 2) / 3;    // value units of a row from column k0 on
        const int nuw = max(nu, B::SVU);    // units written per chosen leaf row
        const uint32_t seq = p.seq_base + (uint32_t)b;

        // ---- 1. leaf: choose among this workgroup's live rows, publish
        if (wave == 0) {
            int m = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int r = h * 64 + lane;
                const bool lv = (r < rows) && (lbl[r] >= k0);
                const unsigned long long bal = __ballot(lv ? 1 : 0);
                if (lv) live[m + __popcll(bal & ((1ull << lane) - 1ull))] = r;
                m += __popcll(bal);
            }
            if (lane == 0) misc[1] = m;
        }
        __syncthreads();
        const int m = misc[1];
        for (int idx = t; idx < m * WS; idx += NT) {
            const int q = idx / WS, j = idx % WS;
            const int r = live[q];
            stg_v[q * (WS + 1) + j] = (j < ws) ? tile[r * LD + k0 + j] : 0.0f;
            if (j == 0) {
                stg_lab[q] = lbl[r];
                stg_src[q] = r;
            }
        }
        __syncthreads();
        const int ks_leaf = merge(m, ws, false);
        // published row: header {label, source = g*WS + slot, 0}, then the row from column k0
        // (every slot's header and block values are written, zeros for an empty slot; the rest
        // of a row only for a chosen one)
        {
            const uint32_t base = (uint32_t)g * WS * FULLU;
            for (int idx = t; idx < WS * (1 + nuw); idx += NT) {
                const int slot = idx / (1 + nuw), u = idx % (1 + nuw);
                const bool full = slot < ks_leaf;
                if (!full && u > B::SVU) continue;
                const int r = full ? stg_src[fin[slot]] : 0;
                u4 un;
                if (u == 0) {
                    un = make_unit(full ? (uint32_t)lbl[r] : 0xFFFFFFFFu,
                                   (uint32_t)(g * WS + slot), 0u, seq, salt(0, 0));
                } else {
                    float x[3];
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int c = k0 + 3 * (u - 1) + e;
                        x[e] = (full && c < W) ? tile[r * LD + c] : 0.0f;
                    }
                    un = make_unit(fb(x[0]), fb(x[1]), fb(x[2]), seq, salt(0, u));
                }
                st16(p.leaf, p.n_leaf, base + (uint32_t)(slot * FULLU + u), un);
            }
        }
        __syncthreads();

        // ---- 2. level-1 merge of F leaves
        if (is_node) {
            const int nsrc = min(F, G - g);
            auto addr = [&](int idx) -> uint32_t {
                const int cand = idx / B::SELU, u = idx % B::SELU;
                return (uint32_t)((g * WS + cand) * FULLU + u);
            };
            auto sl = [&](int idx) -> uint32_t { return salt(0, idx % B::SELU); };
            if (!gather(std::integral_constant<int, B::MUPT>{}, p.leaf, p.n_leaf,
                        nsrc * WS * B::SELU, seq, addr, sl, cand_sink)) {
                aborted = true;
                break;
            }
            const int ks = merge(nsrc * WS, ws, false);
            const uint32_t base = (uint32_t)(g / F) * WS * B::SELU;
            for (int idx = t; idx < WS * B::SELU; idx += NT) {
                const int slot = idx / B::SELU, u = idx % B::SELU;
                const bool full = slot < ks;
                const int c = full ? fin[slot] : 0;
                u4 un;
                if (u == 0) {
                    un = make_unit(full ? (uint32_t)stg_lab[c] : 0xFFFFFFFFu,
                                   full ? (uint32_t)stg_src[c] : 0u, 0u, seq, salt(1, 0));
                } else {
                    float x[3];
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int j = 3 * (u - 1) + e;
                        x[e] = (full && j < WS) ? stg_v[c * (WS + 1) + j] : 0.0f;
                    }
                    un = make_unit(fb(x[0]), fb(x[1]), fb(x[2]), seq, salt(1, u));
                }
                st16(p.node, p.n_node, base + (uint32_t)idx, un);
            }
            __syncthreads();
        }

        // ---- 3. root: final choice, U12, interchanges, publish
        if (is_root) {
            int ncand;
            bool ok;
            if (p.levels == 1) {
                ncand = G * WS;
                auto addr = [&](int idx) -> uint32_t {
                    const int cand = idx / B::SELU, u = idx % B::SELU;
                    return (uint32_t)(cand * FULLU + u);
                };
                auto sl = [&](int idx) -> uint32_t { return salt(0, idx % B::SELU); };
                ok = gather(std::integral_constant<int, B::MUPT>{}, p.leaf, p.n_leaf,
                            ncand * B::SELU, seq, addr, sl, cand_sink);
            } else {
                ncand = ((G + F - 1) / F) * WS;
                auto addr = [&](int idx) -> uint32_t { return (uint32_t)idx; };
                auto sl = [&](int idx) -> uint32_t { return salt(1, idx % B::SELU); };
                ok = gather(std::integral_constant<int, B::MUPT>{}, p.node, p.n_node,
                            ncand * B::SELU, seq, addr, sl, cand_sink);
            }
            if (!ok) {
                aborted = true;
                break;
            }
            const int ks = merge(ncand, ws, true);
            // the winners' columns right of the block, from their leaves' publications
            const int wr = W - k0 - ws;
            if (wr > 0) {
                const int u_lo = 1 + WS / 3;  // first unit holding a column >= WS
                const int nun = nu - u_lo + 1;
                auto addr = [&](int idx) -> uint32_t {
                    const int k = idx / nun, u = u_lo + idx % nun;
                    return (uint32_t)(stg_src[fin[k]] * FULLU + u);
                };
                auto sl = [&](int idx) -> uint32_t { return salt(0, u_lo + idx % nun); };
                auto sink = [&](int idx, u4 un) {
                    const int k = idx / nun, u = u_lo + idx % nun;
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int c = 3 * (u - 1) + e;
                        if (c >= ws && c < W - k0) ublk[k * UL + c] = bf(un[e]);
                    }
                };
                if (!gather(std::integral_constant<int, B::AUPT>{}, p.leaf, p.n_leaf, ks * nun,
                            seq, addr, sl, sink)) {
                    aborted = true;
                    break;
                }
                // U12 = L11^{-1} A12, one column per thread (a block with columns to its right
                // is full: ks == ws == WS)
                for (int j = t; j < wr && ks == WS; j += NT) {
                    float x[WS];
                    static_for<0, WS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        float sum = ublk[k * UL + WS + j];
                        static_for<0, k>([&](auto ic) {
                            constexpr int i = decltype(ic)::value;
                            sum = __builtin_fmaf(-ublk[k * UL + i], x[i], sum);
                        });
                        x[k] = sum;
                    });
                    static_for<0, WS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        ublk[k * UL + WS + j] = x[k];
                    });
                }
            }
            if (wave == 0) {
                // LAPACK-style interchanges: step k swaps position k0+k with the winner's
                // current position. Lane k holds winner k's position; at step k the lanes
                // above k that sit at position k0+k move to winner k's old position.
                int cur = (lane < ks) ? lab_w[lane] : -1;
                for (int k = 0; k < ks; ++k) {
                    const int pk = __builtin_amdgcn_readlane(cur, k);
                    if (lane > k && cur == k0 + k) cur = pk;
                }
                if (lane < ks) piv[lane] = cur;
                // the first zero or non-finite diagonal entry of U11
                const float d = (lane < ks) ? ublk[lane * UL + lane] : 1.0f;
                const unsigned long long badm = __ballot((!__builtin_isfinite(d) || d == 0.0f) ? 1 : 0);
                if (lane == 0)
                    misc[2] = (ks < ws) ? ks + 1 : (badm ? (int)__builtin_ctzll(badm) + 1 : 0);
            }
            __syncthreads();
            // publish: row k = header {winner label, interchange target, bad}, then
            // [L11\U11 | U12] row k from column k0
            const uint32_t bad = (uint32_t)misc[2];
            for (int idx = t; idx < ws * (1 + nu); idx += NT) {
                const int k = idx / (1 + nu), u = idx % (1 + nu);
                u4 un;
                if (u == 0) {
                    un = make_unit((uint32_t)(k < ks ? lab_w[k] : 0),
                                   (uint32_t)(k < ks ? piv[k] : 0), bad, seq, salt(2, 0));
                } else {
                    float x[3];
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int c = 3 * (u - 1) + e;
                        x[e] = (k < ks && c < W - k0) ? ublk[k * UL + c] : 0.0f;
                    }
                    un = make_unit(fb(x[0]), fb(x[1]), fb(x[2]), seq, salt(2, u));
                }
                st16(p.root, p.n_root, (uint32_t)(k * FULLU + u), un);
            }
        }

        // ---- 4. every workgroup: the block rows, interchanges, update
        if (!is_root) {
            auto addr = [&](int idx) -> uint32_t {
                const int k = idx / (1 + nu), u = idx % (1 + nu);
                return (uint32_t)(k * FULLU + u);
            };
            auto sl = [&](int idx) -> uint32_t { return salt(2, idx % (1 + nu)); };
            auto sink = [&](int idx, u4 un) {
                const int k = idx / (1 + nu), u = idx % (1 + nu);
                if (u == 0) {
                    lab_w[k] = (int)un[0];
                    piv[k] = (int)un[1];
                    if (k == 0) misc[2] = (int)un[2];
                } else {
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const int c = 3 * (u - 1) + e;
                        if (c < W - k0) ublk[k * UL + c] = bf(un[e]);
                    }
                }
            };
            if (!gather(std::integral_constant<int, B::RUPT>{}, p.root, p.n_root, ws * (1 + nu),
                        seq, addr, sl, sink)) {
                aborted = true;
                break;
            }
        } else {
            __syncthreads();
        }
        if (misc[2] != 0) {
            // zero or non-finite pivot: stop before this block's interchanges (rows keep their
            // labels; the interchanges from column k0 on are the identity)
            stop_k0 = k0;
            if (g == 0 && t == 0) atomicCAS(p.info, 0, (int32_t)(p.goff + k0 + misc[2]));
            break;
        }
        if (t < ws) rdiag[t] = 1.0f / ublk[t * UL + t];
        if (g == 0 && t < ws) p.dests[k0 + t] = (int32_t)(p.goff + piv[t]);
        __syncthreads();
        // relabel; winners take the published rows, the other live rows are updated (two
        // threads per row: both solve for l, each updates every other column)
        {
            const int r = t & (RMAX - 1), h = t >> 7;
            int Lr = (r < rows) ? lbl[r] : INT_MIN;
            const bool lv = Lr >= k0;
            if (lv) {
                for (int k = 0; k < ws; ++k) {
                    const int pk = piv[k];
                    if (Lr == k0 + k) Lr = pk;
                    else if (Lr == pk) Lr = k0 + k;
                }
            }
            const bool winner = lv && Lr < k0 + ws;
            const bool other = lv && !winner;
            float* row = tile + (size_t)r * LD;
            float l[WS];
            if (other) {
                // l U11 = a(k0:k0+ws) by forward substitution (a narrow last block reads its
                // columns clamped and keeps only its ws multipliers)
                static_for<0, WS>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    float sum = row[k0 + min(k, ws - 1)];
                    static_for<0, k>([&](auto ic) {
                        constexpr int i = decltype(ic)::value;
                        sum = __builtin_fmaf(-l[i], ublk[i * UL + k], sum);
                    });
                    l[k] = sum * rdiag[k];
                });
            }
            __syncthreads();  // every read of lbl[r] and of a(k0:k0+ws) precedes the writes
            if (lv && h == 0) lbl[r] = Lr;
            if (winner) {
                const int k = Lr - k0;
                for (int c = h; c < W - k0; c += 2) row[k0 + c] = ublk[k * UL + c];
            } else if (other) {
                // a(k0+ws:) -= l U12, four columns at a time (a block with columns to its right
                // is full, ws == WS), U12 read as broadcast float4
                const int wr = W - k0 - ws;
                for (int c = 4 * h; c < wr; c += 8) {
                    float a[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) a[e] = (c + e < wr) ? row[k0 + WS + c + e] : 0.0f;
                    static_for<0, WS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        const float4 u = *reinterpret_cast<const float4*>(&ublk[k * UL + WS + c]);
                        a[0] = __builtin_fmaf(-l[k], u.x, a[0]);
                        a[1] = __builtin_fmaf(-l[k], u.y, a[1]);
                        a[2] = __builtin_fmaf(-l[k], u.z, a[2]);
                        a[3] = __builtin_fmaf(-l[k], u.w, a[3]);
                    });
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (c + e < wr) row[k0 + WS + c + e] = a[e];
                }
                if (h == 0) {
                    static_for<0, WS>([&](auto kc) {
                        constexpr int k = decltype(kc)::value;
                        if (k < ws) row[k0 + k] = l[k];
                    });
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();

    if (aborted) {
        // a peer never arrived: A is unchanged, identity interchanges, no moves, info = -1
        for (int i = t; i < 2 * W; i += NT) p.pairs[i] = int2{-1, -1};
        for (int c = t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
        if (t == 0) atomicCAS(p.info, 0, -1);
        return;
    }
    // rows to their final positions; the moved-row list for the other columns
    if (W == 128 && rows == 128) {
#pragma unroll 16
        for (int u = 0; u < 64; ++u) {
            const int idx = t + NT * u, j = idx & 127, r = idx >> 7;
            p.A[(i64)lbl[r] * p.lda + j] = tile[r * LD + j];
        }
    } else {
        for (int idx = t; idx < rows * W; idx += NT) {
            const int j = idx % W, r = idx / W;
            p.A[(i64)lbl[r] * p.lda + j] = tile[r * LD + j];
        }
    }
    if (stop_k0 >= 0 && g == 0)
        for (int c = stop_k0 + t; c < W; c += NT) p.dests[c] = (int32_t)(p.goff + c);
    for (int r = t; r < rows; r += NT) {
        const int phys = (int)(row0 + r), Lr = lbl[r];
        if (Lr < W) p.pairs[Lr] = (Lr != phys) ? int2{Lr, phys} : int2{-1, -1};
        if (phys < W) p.pairs[W + phys] = (Lr >= W) ? int2{Lr, phys} : int2{-1, -1};
    }
}

struct Plan {
    i64 G, R;
};
Plan plan(i64 mp) {
    i64 G = (mp + RMAX - 1) / RMAX;
    const i64 R = (mp + G - 1) / G;
    G = (mp + R - 1) / R;  // every workgroup owns at least one row
    return Plan{G, R};
}

template <int WS>
bool fits(i64 mp, i64 w) {
    const elx::DeviceInfo& di = elx::device_info();
    if (w < 1 || w > WMAX || mp < w || mp >= (i64(1) << 30)) return false;
    const Plan pl = plan(mp);
    if (pl.G > di.num_cus || pl.G > (i64)Blk<WS>::F * Blk<WS>::F) return false;
    return lds_layout<WS>((int)pl.R, (int)w).total <= (size_t)di.max_lds_per_block;
}

template <int WS>
elx_status_t launch_calu(hipStream_t s, i64 mp, i64 w, float* A, i64 lda, int32_t* dests,
                         i64 goff, int2* pairs, int32_t* info) {
    using B = Blk<WS>;
    const elx::DeviceInfo& di = elx::device_info();
    ELX_ARG(w >= 1 && w <= WMAX && mp >= w);
    if (!fits<WS>(mp, w)) return ELX_ERROR_NOT_SUPPORTED;
    const Plan pl = plan(mp);
    const Lds L = lds_layout<WS>((int)pl.R, (int)w);
    // areas sized for the largest block width and grid, shared with the partial-pivoting panel
    const size_t n_leaf = (size_t)di.num_cus * 32 * FULLU;
    const size_t n_node = (size_t)di.num_cus * 32 * Blk<16>::SELU;
    const size_t n_root = (size_t)32 * FULLU;
    elx::PanelSync* ps = nullptr;
    ELX_CHECK(elx::panel_sync_get(di.num_cus, sizeof(u4) * (n_leaf + n_node + n_root), &ps));
    elx::ProfScope prof_(s, elx::kKPanel, 2.0 * (double)mp * (double)w * sizeof(float));
    CaluArgs a;
    a.mp = mp;
    a.w = (int)w;
    a.A = A;
    a.lda = lda;
    a.dests = dests;
    a.goff = goff;
    a.pairs = pairs;
    a.info = info;
    a.G = (int)pl.G;
    a.R = (int)pl.R;
    a.levels = (pl.G <= B::F) ? 1 : 2;
    a.seq_base = (uint32_t)((++ps->epoch) * 1024);
    a.leaf = reinterpret_cast<u4*>(ps->slots);
    a.node = a.leaf + n_leaf;
    a.root = a.node + n_node;
    a.n_leaf = (uint32_t)n_leaf;
    a.n_node = (uint32_t)n_node;
    a.n_root = (uint32_t)n_root;
    const void* fn = reinterpret_cast<const void*>(&calu_panel_kernel<WS>);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  di.max_lds_per_block);
        (void)hipGetLastError();
        attr_set = true;
    }
    void* args[] = {&a};
    ELX_CHECK(elx::launch_resident(fn, pl.G, NT, args, L.total, s));
    return ELX_SUCCESS;
}

}  // namespace

namespace elx {

int mixed_panel_mode() {
    const char* v = std::getenv("ELX_MIXED_PANEL");
    if (v && std::strcmp(v, "partial") == 0) return 0;
    if (v && std::strcmp(v, "calu32") == 0) return 32;
    return 16;
}

// Tournament pivoting pays where the factorization dominates the solve (large systems, whose
// panels spread over many workgroups), and its factors precondition less well (multipliers
// above 1), so by default only systems of at least this order use it; an explicit
// ELX_MIXED_PANEL=calu16 / calu32 uses it for every system. ELX_CALU_MIN_ORDER overrides the
// default order.
// Right-hand sides up to which the default uses tournament panels: each IR/GMRES step costs
// about 3.3 ms per column at n = 32768 and the tournament's factors need about five more of
// them per column, against about 80 ms saved in the panels (ELX_CALU_MAX_RHS overrides; an
// explicit ELX_MIXED_PANEL=calu16 / calu32 has no limit).
i64 mixed_panel_max_rhs() {
    const char* v = std::getenv("ELX_MIXED_PANEL");
    if (v && (std::strcmp(v, "calu16") == 0 || std::strcmp(v, "calu32") == 0)) return INT64_MAX;
    const char* m = std::getenv("ELX_CALU_MAX_RHS");
    if (m && *m) return std::max<i64>(0, std::atoll(m));
    return kCaluMaxRhs;
}

i64 mixed_panel_min_order() {
    const char* v = std::getenv("ELX_MIXED_PANEL");
    if (v && (std::strcmp(v, "calu16") == 0 || std::strcmp(v, "calu32") == 0)) return 0;
    const char* m = std::getenv("ELX_CALU_MIN_ORDER");
    if (m && *m) return std::max<i64>(0, std::atoll(m));
    return kCaluMinOrder;
}

bool calu_panel_fits(i64 mp, i64 w, int ws) {
    return ws == 16 ? fits<16>(mp, w) : (ws == 32 ? fits<32>(mp, w) : false);
}

elx_status_t launch_calu_panel_f32_rowmajor(hipStream_t s, i64 mp, i64 w, float* A, i64 lda,
                                            int32_t* dests, i64 goff, int2* pairs, int32_t* info,
                                            int ws) {
    if (ws == 16) return launch_calu<16>(s, mp, w, A, lda, dests, goff, pairs, info);
    if (ws == 32) return launch_calu<32>(s, mp, w, A, lda, dests, goff, pairs, info);
    return ELX_ERROR_INVALID_ARG;
}

}  // namespace elx
