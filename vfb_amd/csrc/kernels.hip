// HIP kernels (gfx950, wave64) of the CS enumeration hot path.
//
// One epoch = one size level of one or more instances:
//   k_expand  : every candidate of the level.  One lane per candidate; the
//               four operators that read the same operand rows are fused
//               (~,X,F,G per row; &,|,lUr,rUl per pair).  Computes the CS words
//               trace by trace, folds them into the 128-bit key and the target
//               mismatch bits, then probes the hash set.  Candidates whose key
//               is new to the set are appended to the fresh list; on a direct
//               level they mark their own byte of the winner map instead, and
//               one that lowers a slot's ordinal of this epoch marks the
//               displaced ordinal's byte of the loser map.
//               Reference: processOperator + applyOperator + makeRlxUnqChkCSs +
//               processUniqueCS + insertInCache (LTL_BS.cu:30-110, 183-238,
//               251-269, 299-329, 383-446).
//   k_resolve : a fresh candidate survives iff it holds the lowest ordinal
//               of its key (slot.ord == own ordinal); survivors set their bit in
//               the level's winner bitmap; matching survivors atomicMin `found`.
//               (A direct level has no fresh list and no resolve.)
//   k_count / k_scan : pack the bytes to the bitmap (direct: marked and not
//               displaced, and the lowest matching one is `found`), popcount
//               blocks of it, exclusive scan.
//   k_store   : stable compaction: survivors in ordinal order, CS recomputed
//               and written column-major with their ordinal as parent record
//               (replaces the Thrust remove/copy passes, LTL_BS.cu:489-530).
#include <hip/hip_runtime.h>

#include "kernels_dev.h"

namespace vfb {

// G = lanes per candidate: 1 for large levels (throughput), kWideLanes for
// small ones (eight times more wavefronts, an eighth of the serial trace loop)
template <int G>
__global__ __launch_bounds__(256) void k_expand(const InstDesc* __restrict__ D, const Task* __restrict__ tasks,
                                                int ntasks, uint64_t total_waves) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * wpb;
    begin_claims();
    for (uint64_t w = (uint64_t)blockIdx.x * wpb + wib; w < total_waves; w += stride) {
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tasks[mid].wave_begin <= w) lo = mid; else hi = mid - 1;
        }
        const Task& tk = tasks[lo];
        const InstDesc& d = D[tk.inst];
        const uint64_t wl = w - tk.wave_begin;
        if (d.check_only) {
            // onTheFly level or the probe of a level: target check only
            if (tk.kind == 0) expand_unary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
            else expand_binary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
        } else if (tk.kind == 0) {
            if (d.key_kind == KEY_MIX) expand_unary_part<KEY_MIX, G>(d, tk, wl, lane);
            else if (d.key_kind == KEY_FIELD) expand_unary_part<KEY_FIELD, G>(d, tk, wl, lane);
            else expand_unary_part<KEY_STRIDE, G>(d, tk, wl, lane);
        } else {
            if (d.key_kind == KEY_MIX) expand_binary_part<KEY_MIX, G>(d, tk, wl, lane);
            else if (d.key_kind == KEY_FIELD) expand_binary_part<KEY_FIELD, G>(d, tk, wl, lane);
            else expand_binary_part<KEY_STRIDE, G>(d, tk, wl, lane);
        }
    }
    flush_claims(D, tasks, ntasks);
}

// ---------------------------------------------------------------------------
// Traces longer than one word (InstDesc::mw).  A trace's words are evaluated
// from its last word back so that X, F, G and U see its later steps: X takes
// bit 0 of the next word into bit 63, F of a word is all ones once a later
// word holds a step, G of a word is zero once a later step fails, and (l U r)
// at step t is the in-word value or, if l holds from t to the word's end, the
// value at the next word's step 0.  Key parts and masks are per column.
// ---------------------------------------------------------------------------

constexpr uint64_t kPhi = 0x9e3779b97f4a7c15ull;   // SplitMix64 increment

// Folds column ci's word into the key.  KEY_MIX mixes word w of a trace as
// dmix64(word + w * kPhi): for w = 0 the reference's per-trace mix
// (LTL_BS.cu:78-88); the field and sample keys take the trace's bits at the
// reference's global positions (LTL_BS.cu:45-76, 94-106).
template <int KEY>
__device__ __forceinline__ void key_acc_col(Key& k, uint64_t c, const ColInfo& ci, const InstDesc& d) {
    if constexpr (KEY == KEY_MIX) {
        const uint64_t x = dmix64(c + (uint64_t)ci.word * kPhi);
        if (ci.trace < d.half) k.hi ^= x; else k.lo ^= x;
    } else if constexpr (KEY == KEY_FIELD) {
        const int m = ci.kbits;
        if (m > 0) {
            const uint64_t f = c & ci.kmask;
            const int j = ci.koff;
            if (j < 63) {
                k.lo |= (f << j) & 0x7fffffffffffffffull;
                if (j + m > 63) k.hi |= f >> (63 - j);
            } else {
                k.hi |= f << (j - 63);
            }
        }
    } else {
        const int n = ci.kbits;
        const int s = d.stride;
        int j = ci.koff;
        for (int q = 0; q < n; ++q, ++j) {
            const uint64_t b = (c >> (ci.kfirst + q * s)) & 1ull;
            if (j < 63) k.lo |= b << j; else k.hi |= b << (j - 63);
        }
    }
}

// ~, X, F, G of one row over multi-word traces (SEL as in expand_unary)
template <int KEY, bool CHECK, int G, int SEL>
__device__ __forceinline__ void expand_unary_mw(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    constexpr int W = 64 / G;
    const int sub = lane % G;
    const uint64_t x = wl * W + (uint64_t)(lane / G);
    const bool active = x < tk.nl;
    const uint64_t xr = active ? x : 0;
    const gu64* __restrict__ L = glob(d.lvl[tk.lvl_l]);
    Key k[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    uint32_t mm = 0;
    for (int t = sub; t < d.T; t += G) {
        const int c0 = d.tcol[t];
        const uint64_t want = t < d.P ? 1ull : 0ull;
        uint64_t xin = 0;
        bool later_any = false, later_all = true;
        for (int c = d.tcol[t + 1] - 1; c >= c0; --c) {
            const ColInfo& ci = d.cinfo[c];
            const uint64_t m = ci.lenmask;
            const uint64_t a = L[(uint64_t)c * tk.nl + xr];
            const uint64_t v0 = ~a & m;                                  // ~
            const uint64_t v1 = (a >> 1) | (xin << 63);                  // X
            const uint64_t v2 = later_any ? m : suffix_or(a);            // F
            const uint64_t v3 = later_all ? (a & ~suffix_or(v0)) : 0ull; // G
            xin = a & 1ull;
            later_any = later_any || a != 0ull;
            later_all = later_all && a == m;
            if (c == c0)
                mm |= miss(v0, want, 1ull) | (miss(v1, want, 1ull) << 1) | (miss(v2, want, 1ull) << 2) |
                      (miss(v3, want, 1ull) << 3);
            if constexpr (!CHECK) {
                if constexpr ((SEL & 1) != 0) key_acc_col<KEY>(k[0], v0, ci, d);
                if constexpr ((SEL & 0xe) != 0) {
                    key_acc_col<KEY>(k[1], v1, ci, d);
                    key_acc_col<KEY>(k[2], v2, ci, d);
                    key_acc_col<KEY>(k[3], v3, ci, d);
                }
            }
        }
    }
    const uint64_t ord[4] = {tk.base[0] + x, tk.base[1] + x, tk.base[2] + x, tk.base[3] + x};
    const uint32_t win = CHECK ? 0xfu : in_window(d, ord);
    finish_group<KEY, CHECK, G>(d, k, ord, mm, active ? ((uint32_t)SEL & tk.part & win) : 0u, lane);
}

// &, |, l U r, r U l of one pair over multi-word traces
template <int KEY, bool CHECK, int G, int SEL>
__device__ __forceinline__ void expand_binary_mw(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    constexpr int W = 64 / G;
    const int sub = lane % G;
    const uint64_t l = wl / tk.nchunk;
    const uint64_t rc = wl - l * tk.nchunk;
    const uint64_t r = rc * W + (uint64_t)(lane / G);
    // both operands from one level: pair (r, l), r > l, repeats the CS of
    // pair (l, r) at a higher ordinal (& | commute, and its l U r / r U l are
    // (l, r)'s r U l / l U r), and (l, l) gives l itself, an older key; none
    // of them can be a winner or the lowest match, so they are skipped
    const bool diag = tk.lvl_l == tk.lvl_r;
    if (diag && rc * W + W - 1 <= l) return;
    const bool active = r < tk.nr && (!diag || r > l);
    const uint64_t rr = active ? r : 0;
    const gu64* __restrict__ L = glob(d.lvl[tk.lvl_l]);
    const gu64* __restrict__ R = glob(d.lvl[tk.lvl_r]);
    Key k[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    uint32_t mm = 0;
    constexpr bool kU = CHECK || (SEL & 0xc) != 0;
    for (int t = sub; t < d.T; t += G) {
        const int c0 = d.tcol[t];
        const uint64_t want = t < d.P ? 1ull : 0ull;
        uint64_t u2 = 0, u3 = 0;
        for (int c = d.tcol[t + 1] - 1; c >= c0; --c) {
            const ColInfo& ci = d.cinfo[c];
            const uint64_t a = L[(uint64_t)c * tk.nl + l];
            const uint64_t b = R[(uint64_t)c * tk.nr + rr];
            const uint64_t v0 = a & b, v1 = a | b;
            uint64_t v2 = 0, v3 = 0;
            if constexpr (kU) {
                v2 = until_op(a, b) | (u2 ? ~suffix_or(~a) : 0ull);
                v3 = until_op(b, a) | (u3 ? ~suffix_or(~b) : 0ull);
                u2 = v2 & 1ull;
                u3 = v3 & 1ull;
            }
            if (c == c0)
                mm |= miss(v0, want, 1ull) | (miss(v1, want, 1ull) << 1) | (miss(v2, want, 1ull) << 2) |
                      (miss(v3, want, 1ull) << 3);
            if constexpr (!CHECK) {
                if constexpr ((SEL & 1) != 0) key_acc_col<KEY>(k[0], v0, ci, d);
                if constexpr ((SEL & 2) != 0) key_acc_col<KEY>(k[1], v1, ci, d);
                if constexpr ((SEL & 0xc) != 0) {
                    key_acc_col<KEY>(k[2], v2, ci, d);
                    key_acc_col<KEY>(k[3], v3, ci, d);
                }
            }
        }
    }
    const uint64_t pair = l * tk.nr + r;
    const uint64_t ord[4] = {tk.base[0] + pair, tk.base[1] + pair, tk.base[2] + 2 * pair,
                             tk.base[2] + 2 * pair + 1};
    const uint32_t win = CHECK ? 0xfu : in_window(d, ord);
    finish_group<KEY, CHECK, G>(d, k, ord, mm, active ? ((uint32_t)SEL & tk.part & win) : 0u, lane);
}

template <int KEY, int G>
__device__ __forceinline__ void expand_unary_mw_part(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    if (tk.part == SEL_NOT) expand_unary_mw<KEY, false, G, SEL_NOT>(d, tk, wl, lane);
    else if (tk.part == SEL_XFG) expand_unary_mw<KEY, false, G, SEL_XFG>(d, tk, wl, lane);
    else expand_unary_mw<KEY, false, G, SEL_ALL>(d, tk, wl, lane);
}

template <int KEY, int G>
__device__ __forceinline__ void expand_binary_mw_part(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    switch (tk.part) {
        case SEL_AND: expand_binary_mw<KEY, false, G, SEL_AND>(d, tk, wl, lane); break;
        case SEL_ANDOR: expand_binary_mw<KEY, false, G, SEL_ANDOR>(d, tk, wl, lane); break;
        case SEL_ORU: expand_binary_mw<KEY, false, G, SEL_ORU>(d, tk, wl, lane); break;
        case SEL_U: expand_binary_mw<KEY, false, G, SEL_U>(d, tk, wl, lane); break;
        default: expand_binary_mw<KEY, false, G, SEL_ALL>(d, tk, wl, lane); break;
    }
}

// k_expand for epochs holding an instance with multi-word traces: the same
// work mapping, each instance by its own path.  A separate kernel, so that the
// one-word kernel's register allocation is that of its own path alone.
template <int G>
__global__ __launch_bounds__(256) void k_expand_mw(const InstDesc* __restrict__ D, const Task* __restrict__ tasks,
                                                   int ntasks, uint64_t total_waves) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * wpb;
    begin_claims();
    for (uint64_t w = (uint64_t)blockIdx.x * wpb + wib; w < total_waves; w += stride) {
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tasks[mid].wave_begin <= w) lo = mid; else hi = mid - 1;
        }
        const Task& tk = tasks[lo];
        const InstDesc& d = D[tk.inst];
        const uint64_t wl = w - tk.wave_begin;
        if (!d.mw) {
            if (d.check_only) {
                if (tk.kind == 0) expand_unary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
                else expand_binary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
            } else if (tk.kind == 0) {
                if (d.key_kind == KEY_MIX) expand_unary_part<KEY_MIX, G>(d, tk, wl, lane);
                else if (d.key_kind == KEY_FIELD) expand_unary_part<KEY_FIELD, G>(d, tk, wl, lane);
                else expand_unary_part<KEY_STRIDE, G>(d, tk, wl, lane);
            } else {
                if (d.key_kind == KEY_MIX) expand_binary_part<KEY_MIX, G>(d, tk, wl, lane);
                else if (d.key_kind == KEY_FIELD) expand_binary_part<KEY_FIELD, G>(d, tk, wl, lane);
                else expand_binary_part<KEY_STRIDE, G>(d, tk, wl, lane);
            }
        } else if (d.check_only) {
            if (tk.kind == 0) expand_unary_mw<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
            else expand_binary_mw<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
        } else if (tk.kind == 0) {
            if (d.key_kind == KEY_MIX) expand_unary_mw_part<KEY_MIX, G>(d, tk, wl, lane);
            else if (d.key_kind == KEY_FIELD) expand_unary_mw_part<KEY_FIELD, G>(d, tk, wl, lane);
            else expand_unary_mw_part<KEY_STRIDE, G>(d, tk, wl, lane);
        } else {
            if (d.key_kind == KEY_MIX) expand_binary_mw_part<KEY_MIX, G>(d, tk, wl, lane);
            else if (d.key_kind == KEY_FIELD) expand_binary_mw_part<KEY_FIELD, G>(d, tk, wl, lane);
            else expand_binary_mw_part<KEY_STRIDE, G>(d, tk, wl, lane);
        }
    }
    flush_claims(D, tasks, ntasks);
}

// Target check of tk.lrows (= K) left rows against one right row per lane,
// one-word rows: each right word is loaded once for the K left rows, so the
// larger operand level is streamed ceil(nl / K) times instead of nl.  Per
// group of 8 traces one load instruction fetches the K left words of all 8
// (lane 8j + i: row l0 + i of the group's trace j; K = 16 takes two), and
// each word reaches the lanes with v_readlane; the group's trace numbers come
// from the wavefront's copy of the check order (4 per lane, read once), so
// the group's right and left words are in flight together.  The lowest
// matching ordinal of the wave goes to ctr->found.
__device__ __forceinline__ uint64_t readlane64(uint64_t x, int src) {
    return ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(x >> 32), src) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((uint32_t)x, src);
}

template <int K>
__device__ __forceinline__ void check_binary_rows(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    static_assert(K == 1 || K == 4 || K == 8 || K == 16, "left rows per wavefront");
    constexpr int NQ = (K + 7) / 8;            // left-word loads per trace group
    const uint64_t lg = wl / tk.nchunk;
    const uint64_t rc = wl - lg * tk.nchunk;
    const uint64_t l0 = lg * K;
    const uint64_t r = rc * 64 + (uint64_t)lane;
    // one operand level: pairs (l, r) with l >= r repeat lower ordinals (as
    // in expand_binary) and are skipped
    const bool diag = tk.lvl_l == tk.lvl_r;
    if (diag && rc * 64 + 63 <= l0) return;
    const bool active = r < tk.nr;
    const uint64_t rr = active ? r : 0;
    const uint64_t nl = tk.nl;
    const gu64* __restrict__ L = glob(d.lvl[tk.lvl_l]);
    const gu64* __restrict__ R = glob(d.lvl[tk.lvl_r]);
    // the check order, 4 traces per lane (lanes 0..kMaxT/4-1)
    static_assert(kMaxT <= 256 && offsetof(InstDesc, check_order) % 4 == 0, "check order in 32-bit lanes");
    const uint32_t ordw = lane < kMaxT / 4 ? ((const uint32_t*)d.check_order)[lane] : 0u;
    // mismatch bits per left row, starting with the candidates that do not
    // count (no pair, or not in the task's part): once every candidate of
    // the wavefront has missed a trace, the rest of the traces are skipped
    uint32_t mm[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
        mm[k] = (!active || l0 + k >= nl || (diag && r <= l0 + k)) ? 0xfu : (~(uint32_t)tk.part & 0xfu);
    const int T = d.T;
    const int jl = lane >> 3, il = lane & 7;
    for (int t0 = 0; t0 < T; t0 += kTraceGroup) {
        const uint64_t o8 = ((uint64_t)__builtin_amdgcn_readlane(ordw, t0 / 4 + 1) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane(ordw, t0 / 4);
        uint64_t bv[kTraceGroup], av[NQ];
#pragma unroll
        for (int j = 0; j < kTraceGroup; ++j)
            if (t0 + j < T) bv[j] = R[((o8 >> (8 * j)) & 0xffu) * tk.nr + rr];
        {
            // rows past the level repeat its last row (masked by mm)
            const uint64_t tl = (o8 >> (8 * jl)) & 0xffu;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const uint64_t row = l0 + (uint64_t)(q * 8 + il) < nl ? l0 + (uint64_t)(q * 8 + il) : nl - 1;
                av[q] = t0 + jl < T && q * 8 + il < K ? L[tl * nl + row] : 0ull;
            }
        }
#pragma unroll
        for (int j = 0; j < kTraceGroup; ++j) {
            if (t0 + j >= T) break;
            const int t = (int)((o8 >> (8 * j)) & 0xffu);
            const uint64_t b = bv[j];
            const uint64_t want = t < d.P ? 1ull : 0ull;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint64_t a = readlane64(av[k / 8], j * 8 + (k & 7));
                mm[k] |= bit0_mismatch(a & b & 1ull, (a | b) & 1ull, until_bit0(a, b), until_bit0(b, a), want);
            }
        }
        uint32_t all = 0xfu;
#pragma unroll
        for (int k = 0; k < K; ++k) all &= mm[k];
        if (!__ballot(all != 0xfu)) break;
    }
    uint64_t best = kEmpty;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (!active || l0 + k >= nl || (diag && r <= l0 + k)) continue;
        const uint64_t pair = (l0 + k) * tk.nr + r;
        const uint64_t ord[4] = {tk.base[0] + pair, tk.base[1] + pair, tk.base[2] + 2 * pair,
                                 tk.base[2] + 2 * pair + 1};
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (((tk.part >> q) & 1u) && !((mm[k] >> q) & 1u) && ord[q] < best) best = ord[q];
    }
    if (__ballot(best != kEmpty) && best != kEmpty) min64(glob(&d.ctr->found), best);
}

// Lowest level-local ordinal among the candidates of wave wl of task tk
// (W = candidates per wave): ~ X F G of rows x0..; & | of pairs p0.., whose
// Until candidates are 2p and 2p + 1 of their block.
__device__ __forceinline__ uint64_t wave_min_ordinal(const Task& tk, uint64_t wl, uint64_t W) {
    uint64_t m = kEmpty;
    if (tk.kind == 0) {
        const uint64_t x0 = wl * W;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (((tk.part >> q) & 1u) && tk.base[q] + x0 < m) m = tk.base[q] + x0;
    } else {
        const uint64_t lg = wl / tk.nchunk;
        const uint64_t p0 = lg * (tk.lrows > 1 ? tk.lrows : 1) * tk.nr + (wl - lg * tk.nchunk) * W;
        if ((tk.part & 1u) && tk.base[0] + p0 < m) m = tk.base[0] + p0;
        if ((tk.part & 2u) && tk.base[1] + p0 < m) m = tk.base[1] + p0;
        if ((tk.part & 0xcu) && tk.base[2] + 2 * p0 < m) m = tk.base[2] + 2 * p0;
    }
    return m;
}

template <int G>
__global__ __launch_bounds__(256) void k_check(const InstDesc* __restrict__ D, const Task* __restrict__ tasks,
                                               int ntasks, uint64_t total_waves) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * wpb;
    for (uint64_t w = (uint64_t)blockIdx.x * wpb + wib; w < total_waves; w += stride) {
        int lo = 0, hi = ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tasks[mid].wave_begin <= w) lo = mid; else hi = mid - 1;
        }
        const Task& tk = tasks[lo];
        const InstDesc& d = D[tk.inst];
        const uint64_t wl = w - tk.wave_begin;
        const uint64_t found = __hip_atomic_load(glob(&d.ctr->found), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (found <= wave_min_ordinal(tk, wl, 64 / G)) continue;
        if (G == 1 && tk.kind == 1 && tk.lrows == 8) {
            check_binary_rows<8>(d, tk, wl, lane);
        } else if (G == 1 && tk.kind == 1 && tk.lrows == 4) {
            check_binary_rows<4>(d, tk, wl, lane);
        } else if (G == 1 && tk.kind == 1 && tk.lrows == 16) {
            check_binary_rows<16>(d, tk, wl, lane);
        } else if (d.mw) {
            if (tk.kind == 0) expand_unary_mw<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
            else expand_binary_mw<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
        } else {
            if (tk.kind == 0) expand_unary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
            else expand_binary<KEY_MIX, true, G, SEL_ALL>(d, tk, wl, lane);
        }
    }
}

// index of the segment holding flat index g (segments sorted by begin)
template <class S>
__device__ __forceinline__ int seg_of(const S* segs, int n, uint64_t g) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].begin <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Inserts the rows of level 1 (the atoms) with ordinals 0..A-1 so that later
// levels deduplicate against them (hashSetsInitialisation, LTL_BS.cu:136-160).
template <int KEY>
__device__ void insert_row(const InstDesc& d, uint64_t x, uint32_t n, unsigned int* fail) {
    const gu64* L = glob(d.lvl[1]);
    Key k{0, 0};
    if (d.mw) {
        for (int c = 0; c < d.C; ++c) key_acc_col<KEY>(k, L[(uint64_t)c * n + x], d.cinfo[c], d);
    } else {
        // 16 independent row loads in flight before their keys fold in
        // (key_acc folds by XOR/OR, so the grouping leaves the key unchanged)
        constexpr int B = 16;
        for (int t0 = 0; t0 < d.T; t0 += B) {
            uint64_t w[B];
#pragma unroll
            for (int j = 0; j < B; ++j) w[j] = t0 + j < d.T ? L[(uint64_t)(t0 + j) * n + x] : 0ull;
#pragma unroll
            for (int j = 0; j < B; ++j)
                if (t0 + j < d.T) key_acc<KEY>(k, w[j], d, t0 + j);
        }
    }
    key_fin(k);
    uint32_t slot;
    int claimed = 0;
    uint64_t displaced;
    if (probe_insert(glob(d.table), d.table_mask, k.hi, k.lo, x, 0, &slot, &claimed, &displaced) == PR_FAIL) or32(glob(fail), 1u);
}

// BS stage: bit 0 of every trace of each stored row, packed into one word
// (transfer, LTL_BS.cu:808-822); rows are columns of n words, trace t's
// step 0 in column tcol[t] (column t when tcol is null)
__global__ void k_bs_words(const unsigned long long* __restrict__ rows, uint32_t n, int T,
                           const int32_t* __restrict__ tcol, unsigned long long* __restrict__ out) {
    const uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t w = 0;
    for (int t = 0; t < T; ++t) w |= (rows[(uint64_t)(tcol ? tcol[t] : t) * n + r] & 1ull) << t;
    out[r] = w;
}

// distinct keys (hi = 0, lo = keys[i]) with ordinals 0..n-1 into a set
__global__ void k_insert_keys(Slot* tab, uint64_t mask, const unsigned long long* __restrict__ keys, uint64_t n,
                              unsigned int* fail) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t slot;
    int claimed = 0;
    uint64_t displaced;
    if (probe_insert(glob(tab), mask, 0ull, keys[i], i, 0, &slot, &claimed, &displaced) == PR_FAIL) or32(glob(fail), 1u);
}

// the atoms of many instances in one launch (segment k = instance k's atoms)
__global__ void k_insert_atoms_multi(const InstDesc* __restrict__ D, const Seg* __restrict__ segs, int nseg,
                                     uint64_t total, unsigned int* fail) {
    const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (g >= total) return;
    const Seg& sg = segs[seg_of(segs, nseg, g)];
    const InstDesc& d = D[sg.inst];
    const uint64_t x = g - sg.begin;
    const uint32_t n = d.nrows[1];
    if (d.key_kind == KEY_MIX) insert_row<KEY_MIX>(d, x, n, fail);
    else if (d.key_kind == KEY_FIELD) insert_row<KEY_FIELD>(d, x, n, fail);
    else insert_row<KEY_STRIDE>(d, x, n, fail);
}

__global__ void k_insert_atoms(const InstDesc* __restrict__ D, uint32_t inst, uint32_t n, unsigned int* fail) {
    const uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (x >= n) return;
    const InstDesc& d = D[inst];
    if (d.key_kind == KEY_MIX) insert_row<KEY_MIX>(d, x, n, fail);
    else if (d.key_kind == KEY_FIELD) insert_row<KEY_FIELD>(d, x, n, fail);
    else insert_row<KEY_STRIDE>(d, x, n, fail);
}

// Moves one key of a completed epoch (ord < keep_below) into a fresh table.
// The keys are distinct, so a key takes the first slot whose hi it claims
// from EMPTY, alone (one CAS); lo and ord are plain stores, read only by
// later kernels.  Lookups stay valid: they skip slots whose (hi, lo) differ.
__device__ __forceinline__ void rehash_slot(const Slot& s, GLB Slot* tab, uint64_t mask, uint64_t keep_below,
                                            unsigned int* fail) {
    if (s.hi == kEmpty || s.lo == kEmpty || s.ord >= keep_below) return;
    uint64_t j = slot_hash(s.hi, s.lo) & mask;
    for (int n = 0; n < kProbeLimit; ++n, j = (j + 1) & mask) {
        if (ld_any(&tab[j].hi) != kEmpty) continue;
        if (cas64(&tab[j].hi, kEmpty, s.hi) == kEmpty) {
            tab[j].lo = s.lo;
            tab[j].ord = s.ord;
            return;
        }
    }
    or32(glob(fail), 1u);
}

__global__ void k_rehash(const Slot* __restrict__ old, uint64_t old_cap, Slot* tab, uint64_t mask,
                         uint64_t keep_below, unsigned int* fail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < old_cap; i += stride)
        rehash_slot(old[i], glob(tab), mask, keep_below, fail);
}

// the rehashes of many instances' tables in one launch
__global__ void k_rehash_multi(const RehashSeg* __restrict__ segs, int n, uint64_t total, unsigned int* fail) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total; g += stride) {
        const RehashSeg& sg = segs[seg_of(segs, n, g)];
        const GLB Slot& os = glob(sg.old)[g - sg.begin];
        rehash_slot(Slot{os.hi, os.lo, os.ord, 0}, glob(sg.tab), sg.mask, sg.keep_below, fail);
    }
}

// fills and copies of many buffers in one launch over a flat space of chunks
// of 8 words (64 bytes; segment k holds chunks begin_k ..).  A wavefront
// takes 64 chunks (4 KB): inside one fill segment as four 16-byte stores per
// lane, each store of the wavefront 1 KB contiguous (every destination is
// 256-byte aligned; 64 contiguous bytes per lane, the other layout, measured
// 3.7 against 5.5-5.8 TB/s of writes); chunks of copies, segment tails and
// boundaries one per lane, word by word
__global__ void k_fill_multi(const OpSeg* __restrict__ segs, int n, uint64_t total_chunks) {
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t wv = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t w = wv; w * 64 < total_chunks; w += nwv) {
        const uint64_t g0 = w * 64;
        const uint64_t g1 = g0 + 64 < total_chunks ? g0 + 64 : total_chunks;
        const OpSeg& s0 = segs[seg_of(segs, n, g0)];
        if (!s0.src && (g1 - s0.begin) * 8 <= s0.words) {
            const u64x2 v = {s0.value, s0.value};
            GLB u64x2* d2 = (GLB u64x2*)(glob(s0.dst) + (g0 - s0.begin) * 8);
            const uint32_t units = (uint32_t)(g1 - g0) * 4;
            for (uint32_t u = (uint32_t)lane; u < units; u += 64) d2[u] = v;
            continue;
        }
        const uint64_t g = g0 + (uint64_t)lane;
        if (g >= g1) continue;
        const OpSeg& sg = segs[seg_of(segs, n, g)];
        const uint64_t i0 = (g - sg.begin) * 8;
        const uint64_t m = sg.words - i0 < 8 ? sg.words - i0 : 8;
        gu64* dst = glob(sg.dst) + i0;
        for (uint64_t j = 0; j < m; ++j) dst[j] = sg.src ? glob(sg.src)[i0 + j] : sg.value;
    }
}

// n words of value v: 16-byte stores, coalesced across the wavefront (p is
// 256-byte aligned; an odd last word on its own)
__global__ void k_fill(unsigned long long* p, uint64_t n, unsigned long long v) {
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const u64x2 v2 = {v, v};
    GLB u64x2* p2 = (GLB u64x2*)glob(p);
    const uint64_t n2 = n / 2;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n2; i += stride) p2[i] = v2;
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) glob(p)[n - 1] = v;
}

// per-instance offsets of the fresh lists in one flat index space, and its
// size, computed on the device so that the resolve launch needs no host sync;
// one workgroup scans 256 instances at a time
__global__ __launch_bounds__(256) void k_fresh_prefix(const InstDesc* __restrict__ D, Seg* segs, int n,
                                                      unsigned long long* total) {
    __shared__ unsigned long long wsum[4];
    __shared__ unsigned long long carry;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 256) {
        const int k = base + (int)threadIdx.x;
        unsigned long long c = 0;
        if (k < n && !D[k].check_only) {
            const unsigned long long f = D[k].ctr->fresh_count;
            c = f < D[k].fresh_cap ? f : D[k].fresh_cap;
        }
        const unsigned long long inc = wave_incl_scan(c);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        unsigned long long pre = carry;
        for (int q = 0; q < w; ++q) pre += wsum[q];
        if (k < n) {
            segs[k].inst = (uint32_t)k;
            segs[k].begin = pre + inc - c;
        }
        __syncthreads();
        if (threadIdx.x == 255) carry = pre + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Winner test of fresh-list entry e: the candidate still holds the lowest
__global__ __launch_bounds__(256) void k_resolve(const InstDesc* __restrict__ D, const Seg* __restrict__ segs,
                                                 int nseg, const unsigned long long* __restrict__ total_p,
                                                 int single) {
    uint64_t total;
    if (single) {
        const uint64_t c = D[0].ctr->fresh_count;
        total = c < D[0].fresh_cap ? c : D[0].fresh_cap;
    } else {
        total = *total_p;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < total; g += stride) {
        const Seg& sg = segs[seg_find(segs, nseg, g)];
        resolve_entry(D[sg.inst], g - sg.begin);
    }
}

// instance's winner map; one wavefront per chunk
__global__ __launch_bounds__(256) void k_count(const InstDesc* __restrict__ D, const Seg* __restrict__ segs,
                                               int nseg, uint64_t total_chunks) {
    const int lane = threadIdx.x & 63;
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * wpb;
    for (uint64_t g = blockIdx.x * wpb + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); g < total_chunks;
         g += stride) {
        const Seg& sg = segs[seg_find(segs, nseg, g)];
        const InstDesc& d = D[sg.inst];
        const uint64_t c = g - sg.begin;
        uint32_t v = pack_word(d, c * kChunkWords + lane, keep_marks(d));
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) glob(d.blkcount)[c] = v;
    }
}

__global__ __launch_bounds__(256) void k_scan(const InstDesc* __restrict__ D, const uint32_t* __restrict__ insts) {
    scan_one(D[insts[blockIdx.x]]);
}

// Resolve and scan of one small level of one instance in a single workgroup
// (one launch instead of two): the fresh list is resolved by the block, then
// scanned after a barrier.
__global__ __launch_bounds__(256) void k_resolve_scan1(const InstDesc* __restrict__ D) {
    const InstDesc& d = D[0];
    const uint64_t c = d.ctr->fresh_count;
    const uint64_t total = c < d.fresh_cap ? c : d.fresh_cap;
    for (uint64_t e = threadIdx.x; e < total; e += blockDim.x) resolve_entry(d, e);
    __syncthreads();
    scan_one(d);
}

// stable compaction of the survivors.  Work item = (chunk of 4096 ordinals,
// one of SP interleaved survivor subsets, slice of kStoreTraces
// traces), one wavefront each: the chunk's set bits
// become an LDS list of 12-bit offsets, then the 64 lanes write consecutive
// rows of the slice's columns (coalesced stores).  Slice 0 also writes the
// parent ordinals.
template <int SP>
__global__ __launch_bounds__(256) void k_store(const InstDesc* __restrict__ D, const Seg* __restrict__ segs,
                                               int nseg, uint64_t total_items, StorePatch patch) {
    __shared__ uint16_t lists[4][64 * kChunkWords];
    const int lane = threadIdx.x & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint16_t* list = lists[wv];
    const uint64_t wpb = blockDim.x >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * wpb;
    for (uint64_t g = blockIdx.x * wpb + wv; g < total_items; g += stride) {
        // nseg == 0: one instance, descriptor 0, outputs from `patch`
        const int si = nseg ? seg_find(segs, nseg, g) : 0;
        const InstDesc& d = D[nseg ? segs[si].inst : 0];
        const uint64_t item = g - (nseg ? segs[si].begin : 0);
        gu64* const out_rows = glob(nseg ? d.out_rows : patch.out_rows);
        gu64* const out_par = glob(nseg ? d.out_par : patch.out_par);
        const uint64_t out_n = nseg ? d.out_n : patch.out_n;
        store_item<SP>(d, item, list, out_rows, out_par, out_n, lane);
    }
}

// out[i] = *addr[i]: the parent ordinals the host needs to spell out the
// formulas found in one step, read in one launch
__global__ void k_gather(const unsigned long long* const* __restrict__ addr, uint64_t n,
                         unsigned long long* __restrict__ out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = *glob(addr[i]);
}

// ---------------------------------------------------------------------------
// launch wrappers (host)
// ---------------------------------------------------------------------------

hipError_t launch_gather(const unsigned long long* const* addr, uint64_t n, unsigned long long* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, addr, n, out);
    return hipGetLastError();
}

hipError_t launch_expand(const InstDesc* D, const Task* tasks, int ntasks, uint64_t total_waves, int wide, int mw,
                         int grid, hipStream_t st) {
    if (mw < 0) {   // every instance of the epoch checks only
        if (wide)
            hipLaunchKernelGGL(k_check<kWideLanes>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
        else
            hipLaunchKernelGGL(k_check<1>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
        return hipGetLastError();
    }
    if (mw && wide)
        hipLaunchKernelGGL(k_expand_mw<kWideLanes>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
    else if (mw)
        hipLaunchKernelGGL(k_expand_mw<1>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
    else if (wide)
        hipLaunchKernelGGL(k_expand<kWideLanes>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
    else
        hipLaunchKernelGGL(k_expand<1>, dim3(grid), dim3(256), 0, st, D, tasks, ntasks, total_waves);
    return hipGetLastError();
}

hipError_t launch_insert_atoms_multi(const InstDesc* D, const Seg* segs, int nseg, uint64_t total, unsigned int* fail,
                                     hipStream_t st) {
    hipLaunchKernelGGL(k_insert_atoms_multi, dim3((unsigned)((total + 63) / 64)), dim3(64), 0, st, D, segs, nseg, total,
                       fail);
    return hipGetLastError();
}

hipError_t launch_bs_words(const unsigned long long* rows, uint32_t n, int T, const int32_t* tcol,
                           unsigned long long* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_bs_words, dim3((n + 255) / 256), dim3(256), 0, st, rows, n, T, tcol, out);
    return hipGetLastError();
}

hipError_t launch_insert_keys(Slot* tab, uint64_t mask, const unsigned long long* keys, uint64_t n, unsigned int* fail,
                              hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_insert_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tab, mask, keys, n, fail);
    return hipGetLastError();
}

hipError_t launch_fill_multi(const OpSeg* segs, int n, uint64_t total, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_multi, dim3(grid), dim3(256), 0, st, segs, n, total);
    return hipGetLastError();
}

hipError_t launch_rehash_multi(const RehashSeg* segs, int n, uint64_t total, unsigned int* fail, int grid,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_rehash_multi, dim3(grid), dim3(256), 0, st, segs, n, total, fail);
    return hipGetLastError();
}

hipError_t launch_insert_atoms(const InstDesc* D, uint32_t inst, uint32_t n, unsigned int* fail, hipStream_t st) {
    hipLaunchKernelGGL(k_insert_atoms, dim3((n + 63) / 64), dim3(64), 0, st, D, inst, n, fail);
    return hipGetLastError();
}

hipError_t launch_rehash(const Slot* old, uint64_t old_cap, Slot* tab, uint64_t mask, uint64_t keep_below,
                         unsigned int* fail, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_rehash, dim3(grid), dim3(256), 0, st, old, old_cap, tab, mask, keep_below, fail);
    return hipGetLastError();
}

hipError_t launch_fill(unsigned long long* p, uint64_t n, unsigned long long v, int grid, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fresh_prefix(const InstDesc* D, Seg* segs, int n, unsigned long long* total, hipStream_t st) {
    hipLaunchKernelGGL(k_fresh_prefix, dim3(1), dim3(256), 0, st, D, segs, n, total);
    return hipGetLastError();
}

hipError_t launch_resolve(const InstDesc* D, const Seg* segs, int nseg, const unsigned long long* total, int single,
                          int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_resolve, dim3(grid), dim3(256), 0, st, D, segs, nseg, total, single);
    return hipGetLastError();
}

hipError_t launch_count(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_chunks, int grid,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_count, dim3(grid), dim3(256), 0, st, D, segs, nseg, total_chunks);
    return hipGetLastError();
}

hipError_t launch_resolve_scan1(const InstDesc* D, hipStream_t st) {
    hipLaunchKernelGGL(k_resolve_scan1, dim3(1), dim3(256), 0, st, D);
    return hipGetLastError();
}

hipError_t launch_scan(const InstDesc* D, const uint32_t* insts, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_scan, dim3(n), dim3(256), 0, st, D, insts);
    return hipGetLastError();
}

hipError_t launch_store(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_items, StorePatch patch,
                        int small, int grid, hipStream_t st) {
    if (small)
        hipLaunchKernelGGL(k_store<kStorePartsSmall>, dim3(grid), dim3(256), 0, st, D, segs, nseg, total_items, patch);
    else
        hipLaunchKernelGGL(k_store<kStoreParts>, dim3(grid), dim3(256), 0, st, D, segs, nseg, total_items, patch);
    return hipGetLastError();
}

}  // namespace vfb
