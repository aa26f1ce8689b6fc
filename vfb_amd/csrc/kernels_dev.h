// Device functions of the CS enumeration shared by the level kernels
// (kernels.hip) and the persistent level-chain kernel (chain.hip): memory
// helpers, the operators, the key fold, the hash-set probe, the fused
// expansion of one task's wavefront, winner resolution, the bitmap scan
// and candidate decoding.  References are given per function.
#pragma once

#include <hip/hip_runtime.h>

#include "engine_dev.h"

namespace vfb {

// Device memory reached through descriptors is global memory.  These casts
// let the compiler emit global loads, stores and atomics (a scalar base and a
// 32-bit offset) instead of flat ones, which need 64-bit addresses in vector
// registers and also count against the scalar-load counter, so that waiting
// on a scalar load waited on them too.
#define GLB __attribute__((address_space(1)))
typedef GLB unsigned long long gu64;
typedef GLB unsigned int gu32;
template <class T>
__device__ __forceinline__ GLB T* glob(T* p) {
    return (GLB T*)p;
}
template <class T>
__device__ __forceinline__ const GLB T* glob(const T* p) {
    return (const GLB T*)p;
}

// relaxed atomics at agent scope (those of atomicCAS, atomicMin, atomicAdd
// and atomicOr), on global memory
__device__ __forceinline__ uint64_t cas64(gu64* p, uint64_t expect, uint64_t desired) {
    unsigned long long e = expect;
    __hip_atomic_compare_exchange_strong(p, &e, (unsigned long long)desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return e;   // the value found (= expect when the exchange happened)
}
__device__ __forceinline__ uint64_t min64(gu64* p, uint64_t v) {
    return __hip_atomic_fetch_min(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t add64(gu64* p, uint64_t v) {
    return __hip_atomic_fetch_add(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void or32(gu32* p, uint32_t v) {
    (void)__hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(gu64* p, uint64_t v) {
    // a relaxed store at agent scope: seen by ld_dev() on every XCD
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t dmix64(uint64_t x) {
    // SplitMix64 finaliser (LTL_BS.cu:80-83)
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t suffix_or(uint64_t x) {
    x |= x >> 1; x |= x >> 2; x |= x >> 4;
    x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}

// l U r (the reference doubles the carry over shifts, LTL_BS.cu:222-233).
// Here one 64-bit add: with the words bit-reversed the future of a step is
// the lower bits, and "r holds, or l holds and l U r holds one step later"
// is the carry chain of an adder with generate g = rev(r) and propagate
// rev(l): the carries of g + (g | rev(l)), taken out of each position, are
// rev(l U r).  Steps past a trace are 0 in both words, so they stay 0.
__device__ __forceinline__ uint64_t until_op(uint64_t l, uint64_t r) {
    const uint64_t g = __builtin_bitreverse64(r);
    const uint64_t b = g | __builtin_bitreverse64(l);
    const uint64_t s = g + b;
    const uint64_t carry_out = s < g ? 1ull : 0ull;   // out of the top bit
    const uint64_t c = s ^ g ^ b;                      // carries into each bit
    return __builtin_bitreverse64((c >> 1) | (carry_out << 63));
}

__device__ __forceinline__ uint64_t apply_op(int op, uint64_t a, uint64_t b, uint64_t m) {
    switch (op) {
    case VFB_OP_NOT: return ~a & m;
    case VFB_OP_AND: return a & b;
    case VFB_OP_OR: return a | b;
    case VFB_OP_NEXT: return a >> 1;
    case VFB_OP_FINALLY: return suffix_or(a);
    case VFB_OP_GLOBALLY: return a & ~suffix_or(~a & m);
    default: return until_op(a, b);
    }
}

// bit 0 of l U r: some k with r_k and l_j for all j < k.  l ^ (l + 1) is the
// mask of bits 0..k0, k0 the first step where l fails (all bits when l holds
// everywhere), so r must hold somewhere in it (until_op evaluated at time 0)
__device__ __forceinline__ bool until_bit0(uint64_t l, uint64_t r) {
    return (r & (l ^ (l + 1ull))) != 0ull;
}

__device__ __forceinline__ uint32_t bit0_mismatch(uint64_t b0, uint64_t b1, bool b2, bool b3, uint64_t want) {
    return (uint32_t)(b0 ^ want) | ((uint32_t)(b1 ^ want) << 1) | ((uint32_t)((uint64_t)b2 ^ want) << 2) |
           ((uint32_t)((uint64_t)b3 ^ want) << 3);
}

struct Key {
    uint64_t hi, lo;
};

// 1 if candidate word c misses the target of its trace: bit 0 against the
// trace's label, or (BS stage) the whole word against the positives mask
__device__ __forceinline__ uint32_t miss(uint64_t c, uint64_t want, uint64_t cmask) {
    return ((c ^ want) & cmask) != 0ull ? 1u : 0u;
}

// Folds trace t's CS word into the key (LTL_BS.cu:30-110).  Trace-uniform
// parameters (offsets, widths, masks) were precomputed on the host.
template <int KEY>
__device__ __forceinline__ void key_acc(Key& k, uint64_t c, const InstDesc& d, int t) {
    if constexpr (KEY == KEY_MIX) {
        const uint64_t x = dmix64(c);
        if (t < d.half) k.hi ^= x; else k.lo ^= x;
    } else if constexpr (KEY == KEY_FIELD) {
        const int m = d.kbits[t];
        if (m > 0) {
            const uint64_t f = c & d.kmask[t];
            const int j = d.koff[t];
            if (j < 63) {
                k.lo |= (f << j) & 0x7fffffffffffffffull;
                if (j + m > 63) k.hi |= f >> (63 - j);
            } else {
                k.hi |= f << (j - 63);
            }
        }
    } else {
        const int n = d.kbits[t];
        const int s = d.stride;
        int j = d.koff[t];
        for (int q = 0; q < n; ++q, ++j) {
            const uint64_t b = (c >> (q * s)) & 1ull;
            if (j < 63) k.lo |= b << j; else k.hi |= b << (j - 63);
        }
    }
}

__device__ __forceinline__ void key_fin(Key& k) {
    // -1 and -2 are one half-key in the reference (hash_set.cuh:198-204)
    if (k.hi == kEmpty) k.hi = kEmpty - 1;
    if (k.lo == kEmpty) k.lo = kEmpty - 1;
}

__device__ __forceinline__ uint64_t slot_hash(uint64_t hi, uint64_t lo) {
    return dmix64(hi ^ dmix64(lo + 0x9e3779b97f4a7c15ull));
}

__device__ __forceinline__ uint64_t ld_any(const gu64* p) {
    // may return a stale value; only EMPTY can be stale because every field
    // of a slot is written once (hi, lo) or only decreases (ord)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ uint64_t ld_dev(const gu64* p) {
    // device-coherent read of a slot's ordinal: a candidate that sees a smaller
    // ordinal of this epoch skips its atomicMin (hot keys stay uncontended)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t bcast64(uint64_t v, int src) {
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

enum { PR_DUP = 0, PR_FRESH = 1, PR_FAIL = 2 };

// Claims (new keys) of the workgroup not yet added to its instance's counter:
// with InstDesc::defer_claims the expansion kernels add them up here and
// flush once per workgroup (flush_claims).  One counter per instance takes
// every wavefront's update otherwise, and atomics on one address run at
// about 83 M/s on MI355X (measured), below the rate the expansion of a large
// level issues them.  The same holds for the fresh candidates a direct
// level counts.
__shared__ unsigned long long g_pending_claims, g_pending_fresh;

__device__ __forceinline__ void count_claims(const InstDesc& d, uint32_t ctot) {
    if (d.defer_claims) {
        atomicAdd(&g_pending_claims, (unsigned long long)ctot);
    } else {
        const unsigned long long before = add64(glob(&d.ctr->claims), ctot);
        if (before + ctot > d.claim_limit) or32(glob(&d.ctr->overflow), 4u);
    }
}

__device__ __forceinline__ void count_fresh(const InstDesc& d, uint32_t ftot) {
    if (d.defer_claims) atomicAdd(&g_pending_fresh, (unsigned long long)ftot);
    else add64(glob(&d.ctr->fresh_real), ftot);
}

// start and end of an expansion kernel: the workgroup's deferred counts into
// the counters of the launch's one instance, with the load-limit check the
// undeferred update makes (the last workgroup to flush sees the total)
__device__ __forceinline__ void begin_claims() {
    if (threadIdx.x == 0) {
        g_pending_claims = 0;
        g_pending_fresh = 0;
    }
    __syncthreads();
}
__device__ __forceinline__ void flush_claims(const InstDesc* D, const Task* tasks, int ntasks) {
    __syncthreads();
    if (threadIdx.x == 0 && ntasks > 0) {
        const InstDesc& d = D[tasks[0].inst];
        const unsigned long long c = g_pending_claims;
        if (c) {
            const unsigned long long before = add64(glob(&d.ctr->claims), c);
            if (before + c > d.claim_limit) or32(glob(&d.ctr->overflow), 4u);
        }
        if (g_pending_fresh) add64(glob(&d.ctr->fresh_real), g_pending_fresh);
    }
}

// Insert-or-find of a 128-bit key with lowest-ordinal tracking, lock-free:
// a slot is claimed by a 64-bit CAS of its ord from EMPTY to the candidate's
// ordinal; the claimer alone then writes hi and lo (agent-scope stores, no
// wait) and is done.  A prober that finds ord taken and hi or lo still EMPTY
// reads them again until the claimer's stores show (the re-reads count
// against kProbeLimit); with both shown it compares the key, and on a match
// lowers ord.  Every prober of a key thus stops at the same slot.  Keys from
// earlier epochs carry ord < epoch_base and are never modified.  The first
// slot's fields (th, tl, to) are passed in, loaded ahead by the caller.
// PR_FRESH only for a candidate that may still hold the key's lowest
// ordinal: one that sees a smaller ordinal of this epoch has certainly lost.
// A fresh candidate that lowered the ordinal returns the one it displaced in
// *displaced (kEmpty otherwise).
__device__ __forceinline__ int probe_from(GLB Slot* tab, uint64_t mask, uint64_t hi, uint64_t lo, uint64_t gord,
                                          uint64_t epoch_base, uint64_t i, uint64_t th, uint64_t tl,
                                          uint64_t to, uint32_t* slot_out, int* claimed, uint64_t* displaced) {
    for (int n = 0; n < kProbeLimit; ++n) {
        GLB Slot* s = tab + i;
        if (to == kEmpty) {
            to = cas64(&s->ord, kEmpty, gord);
            if (to == kEmpty) {
                st_dev(&s->hi, hi);
                st_dev(&s->lo, lo);
                *claimed += 1;
                *slot_out = (uint32_t)i;
                return PR_FRESH;
            }
        }
        if (th == kEmpty || tl == kEmpty) {
            th = ld_dev(&s->hi);          // the claimer's key is on its way
            tl = ld_dev(&s->lo);
            continue;
        }
        if (th == hi && tl == lo) {
            *slot_out = (uint32_t)i;
            if (to < gord) return PR_DUP;      // older key, or a smaller ordinal of this epoch
            const uint64_t old = min64(&s->ord, gord);
            if (old < gord) return PR_DUP;
            *displaced = old;
            return PR_FRESH;
        }
        i = (i + 1) & mask;
        s = tab + i;
        th = ld_any(&s->hi);
        tl = ld_any(&s->lo);
        to = ld_dev(&s->ord);
    }
    return PR_FAIL;
}

__device__ __forceinline__ int probe_insert(GLB Slot* tab, uint64_t mask, uint64_t hi, uint64_t lo,
                                            uint64_t gord, uint64_t epoch_base,
                                            uint32_t* slot_out, int* claimed, uint64_t* displaced) {
    const uint64_t i = slot_hash(hi, lo) & mask;
    GLB Slot* s = tab + i;
    return probe_from(tab, mask, hi, lo, gord, epoch_base, i, ld_any(&s->hi), ld_any(&s->lo), ld_any(&s->ord),
                      slot_out, claimed, displaced);
}

__device__ __forceinline__ uint32_t wave_excl_scan_u32(uint32_t v, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(x, off);
        if (lane >= off) x += u;
    }
    *total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
}

// Probes the candidates of this lane (up to 4) as four interleaved state
// machines -- per slot visited, one round of ord-CAS (a claimer stores its key
// and is done) and one of atomicMin, each round issuing the atomics of all
// four candidates together so that their latencies overlap -- then appends the
// fresh ones with one atomic per wavefront (a direct level marks them in the
// winner and loser bytes instead).  Per-candidate semantics are those
// of probe_from().
__device__ __forceinline__ void finish_candidates(const InstDesc& d, const Key* key, const uint64_t* ord,
                                                  uint32_t mismatch, uint32_t enable, int lane) {
    GLB Slot* tab = glob(d.table);
    const uint64_t mask = d.table_mask;
    uint64_t idx[4], th[4], tl[4], to[4], disp[4];
    uint32_t act = enable, fresh = 0, claimed = 0, lowered = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        idx[q] = slot_hash(key[q].hi, key[q].lo) & mask;
        th[q] = tl[q] = to[q] = kEmpty;
        if ((act >> q) & 1u) {
            th[q] = ld_any(&tab[idx[q]].hi);
            tl[q] = ld_any(&tab[idx[q]].lo);
            to[q] = ld_dev(&tab[idx[q]].ord);
        }
    }
    for (int n = 0; act && n < kProbeLimit; ++n) {
        uint64_t r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (((act >> q) & 1u) && to[q] == kEmpty) r[q] = cas64(&tab[idx[q]].ord, kEmpty, d.epoch_base + ord[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (((act >> q) & 1u) && to[q] == kEmpty) {
                if (r[q] == kEmpty) {                   // claimed: the key goes in, fresh
                    st_dev(&tab[idx[q]].hi, key[q].hi);
                    st_dev(&tab[idx[q]].lo, key[q].lo);
                    ++claimed;
                    fresh |= 1u << q;
                    act &= ~(1u << q);
                } else {
                    to[q] = r[q];
                }
            }
        uint32_t need_min = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (((act >> q) & 1u) && th[q] == key[q].hi && tl[q] == key[q].lo) {
                const uint64_t gord = d.epoch_base + ord[q];
                if (to[q] < gord) act &= ~(1u << q);      // older key, or a smaller ordinal of this epoch
                else need_min |= 1u << q;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if ((need_min >> q) & 1u) r[q] = min64(&tab[idx[q]].ord, d.epoch_base + ord[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if ((need_min >> q) & 1u) {
                act &= ~(1u << q);
                if (r[q] >= d.epoch_base + ord[q]) {
                    fresh |= 1u << q;
                    lowered |= 1u << q;
                    disp[q] = r[q];
                }
            }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if ((act >> q) & 1u) {
                if (th[q] == kEmpty || tl[q] == kEmpty) {   // the claimer's key is on its way
                    th[q] = ld_dev(&tab[idx[q]].hi);
                    tl[q] = ld_dev(&tab[idx[q]].lo);
                } else {                                     // another key: next slot
                    idx[q] = (idx[q] + 1) & mask;
                    th[q] = ld_any(&tab[idx[q]].hi);
                    tl[q] = ld_any(&tab[idx[q]].lo);
                    to[q] = ld_dev(&tab[idx[q]].ord);
                }
            }
    }
    if (d.direct) {
        // the candidate's own byte, and the byte of the ordinal it displaced
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if ((fresh >> q) & 1u) glob(d.winbyte)[ord[q]] = ((mismatch >> q) & 1u) ? 1 : 3;
            if ((lowered >> q) & 1u) glob(d.lostbyte)[disp[q] - d.epoch_base] = 1;
        }
    }
    // one fresh-list reservation and one claims update per wavefront
    uint32_t ftot = 0, ctot = 0;
    const uint32_t fbase = wave_excl_scan_u32((uint32_t)__popc(fresh), &ftot);
    (void)wave_excl_scan_u32(claimed, &ctot);
    const uint64_t fail = __ballot(act != 0);
    if (ftot | ctot | (fail ? 1u : 0u)) {
        if (lane == 0) {
            if (ctot) count_claims(d, ctot);
            if (fail) or32(glob(&d.ctr->overflow), 1u);
            if (d.direct && ftot) count_fresh(d, ftot);
        }
        if (d.direct) return;
        unsigned long long fb = 0;
        if (ftot && lane == 0) fb = add64(glob(&d.ctr->fresh_count), ftot);
        const uint64_t f0 = bcast64(fb, 0);
        uint32_t k = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if ((fresh >> q) & 1u) {
                const uint64_t e = f0 + fbase + k++;
                if (e < d.fresh_cap) {
                    const uint64_t flag = ((mismatch >> q) & 1u) ? 0ull : kMatchBit;
                    glob(d.fresh_ord)[e] = (d.epoch_base + ord[q]) | flag;
                    glob(d.fresh_slot)[e] = (uint32_t)idx[q];
                } else {
                    or32(glob(&d.ctr->overflow), 2u);
                }
            }
        }
    }
}

// onTheFly last round (processOnTheFly, LTL_BS.cu:331-349): no key, no
// storage; the lowest matching ordinal wins
__device__ __forceinline__ void check_candidates(const InstDesc& d, const uint64_t* ord, uint32_t mismatch,
                                                 uint32_t enable) {
    uint64_t best = kEmpty;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (((enable >> q) & 1u) && !((mismatch >> q) & 1u) && ord[q] < best) best = ord[q];
    if (__ballot(best != kEmpty) && best != kEmpty) min64(glob(&d.ctr->found), best);
}

// One candidate per lane (small-level mode): probe, then the fresh-list
// append and claims update with one atomic each per wavefront.
__device__ __forceinline__ void finish_one(const InstDesc& d, const Key& key, uint64_t ord, uint32_t mismatch,
                                           bool en, int lane) {
    uint32_t slot = 0xffffffffu;
    int claimed = 0;
    int r = PR_DUP;
    uint64_t displaced = kEmpty;
    if (en)
        r = probe_insert(glob(d.table), d.table_mask, key.hi, key.lo, d.epoch_base + ord, d.epoch_base, &slot,
                         &claimed, &displaced);
    const bool fresh = r == PR_FRESH;
    if (d.direct && fresh) {
        glob(d.winbyte)[ord] = mismatch ? 1 : 3;
        if (displaced != kEmpty) glob(d.lostbyte)[displaced - d.epoch_base] = 1;
    }
    uint32_t ftot = 0, ctot = 0;
    const uint32_t fpos = wave_excl_scan_u32(fresh ? 1u : 0u, &ftot);
    (void)wave_excl_scan_u32((uint32_t)claimed, &ctot);
    const uint64_t fail = __ballot(r == PR_FAIL);
    if (ftot | ctot | (fail ? 1u : 0u)) {
        if (lane == 0) {
            if (ctot) count_claims(d, ctot);
            if (fail) or32(glob(&d.ctr->overflow), 1u);
            if (d.direct && ftot) count_fresh(d, ftot);
        }
        if (d.direct) return;
        unsigned long long fb = 0;
        if (ftot && lane == 0) fb = add64(glob(&d.ctr->fresh_count), ftot);
        const uint64_t base = bcast64(fb, 0) + fpos;
        if (fresh) {
            if (base < d.fresh_cap) {
                glob(d.fresh_ord)[base] = (d.epoch_base + ord) | (mismatch ? 0ull : kMatchBit);
                glob(d.fresh_slot)[base] = slot;
            } else {
                or32(glob(&d.ctr->overflow), 2u);
            }
        }
    }
}

// candidates q (of the four) whose level-local ordinal lies in the pass's
// window [ord_lo, ord_hi]
__device__ __forceinline__ uint32_t in_window(const InstDesc& d, const uint64_t* ord) {
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) m |= (ord[q] >= d.ord_lo && ord[q] <= d.ord_hi ? 1u : 0u) << q;
    return m;
}

// Traces t0..t0+7 of the target-check order, t0 a multiple of 8, as one
// aligned 8-byte load of InstDesc::check_order (a wavefront-uniform value, so
// the operand loads of the 8 traces issue together instead of each waiting on
// a load of its trace number); trace j is byte j.
static_assert(offsetof(InstDesc, check_order) % 8 == 0 && kMaxT % 8 == 0 && kTraceGroup == 8,
              "check_order in aligned groups of one trace group");
__device__ __forceinline__ uint64_t check_order8(const InstDesc& d, int t0) {
    const uint64_t v = *(const uint64_t*)(d.check_order + t0);
    // every lane loaded the same word: scalar registers, so that loads
    // indexed by these traces (the left rows of k_check) are scalar too
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}

// Per-trace accumulation of one candidate's CS word c.  KEY_MIX keys are
// the XOR of dmix64(word) over the traces below InstDesc::half (hi) and over
// the rest (lo): the loops run those two trace ranges one after the other, so
// a word folds into the running half `acc` with no per-trace select.  The
// other keys place the word's bits by trace (key_acc).  The target check
// keeps the OR of (c ^ want): bit 0 decides a miss (label at time 0), every
// bit in the BS stage (the word against the positives mask).
// (The BS stage's one-word rows have at most 63 bits and so exact field keys:
// a KEY_MIX instance never has bs_mode, and its check needs the low word only.)
template <int KEY>
__device__ __forceinline__ void fold_word(Key& k, uint64_t& acc, uint64_t& mq, uint64_t c, uint64_t want,
                                          const InstDesc& d, int t) {
    if constexpr (KEY == KEY_MIX) {
        mq |= (uint64_t)((uint32_t)c ^ (uint32_t)want);
        acc ^= dmix64(c);
    } else {
        mq |= c ^ want;
        key_acc<KEY>(k, c, d, t);
    }
}

// First trace of the second range (lo half) of a KEY_MIX fold; T (one range)
// otherwise
template <int KEY, bool CHECK>
__device__ __forceinline__ int trace_split(const InstDesc& d) {
    if constexpr (KEY == KEY_MIX && !CHECK) return d.half < 0 ? 0 : (d.half < d.T ? d.half : d.T);
    return d.T;
}

// End of the first range: its running half is the key's hi, and the lane's
// next trace (t = sub mod G) is the first at or after tmid
template <int KEY, int G>
__device__ __forceinline__ void next_split(Key* k, uint64_t* acc, int& tb, int tmid, int sub) {
    if constexpr (KEY == KEY_MIX) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            k[q].hi = acc[q];
            acc[q] = 0;
        }
    }
    tb = tmid + ((sub - tmid) & (G - 1));
}

// The lo half of a KEY_MIX key, and the miss bits of the four candidates
template <int KEY>
__device__ __forceinline__ uint32_t fold_end(Key* k, const uint64_t* acc, const uint64_t* mq, const InstDesc& d) {
    const uint64_t fm = KEY != KEY_MIX && d.bs_mode ? ~0ull : 1ull;
    uint32_t mm = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if constexpr (KEY == KEY_MIX) k[q].lo = acc[q];
        mm |= ((mq[q] & fm) != 0ull ? 1u : 0u) << q;
    }
    return mm;
}

// Where a candidate probed by one lane goes: the device-wide hash set
// (finish_one).  The LDS level chain passes its own sink with the same
// per-candidate semantics (chain.hip).
struct GlobalSink {
    static constexpr bool kGlobal = true;   // finish_candidates (G = 1) probes the device set too
    static __device__ __forceinline__ void finish(const InstDesc& d, const Key& key, uint64_t ord, uint32_t mismatch,
                                                  bool en, int lane) {
        finish_one(d, key, ord, mismatch, en, lane);
    }
};

// Combines the per-lane partial keys / mismatch bits of the G lanes that
// share a candidate (G > 1: each lane folded every G-th trace), then probes
// (or, in a check-only level, checks) the four candidates.
template <int KEY, bool CHECK, int G, class Sink = GlobalSink>
__device__ __forceinline__ void finish_group(const InstDesc& d, Key* k, const uint64_t* ord, uint32_t mm,
                                             uint32_t enable, int lane) {
    if constexpr (G > 1) {
#pragma unroll
        for (int off = 1; off < G; off <<= 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint64_t h = __shfl_xor(k[q].hi, off), l = __shfl_xor(k[q].lo, off);
                if constexpr (KEY == KEY_MIX) { k[q].hi ^= h; k[q].lo ^= l; }
                else { k[q].hi |= h; k[q].lo |= l; }
            }
            mm |= __shfl_xor(mm, off);
        }
    }
    const int sub = lane % G;
    if constexpr (CHECK) {
        check_candidates(d, ord, mm, sub == 0 ? enable : 0u);
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) key_fin(k[q]);
    if constexpr (G == 1) {
        static_assert(Sink::kGlobal, "one lane per candidate probes the device-wide set only");
        finish_candidates(d, k, ord, mm, enable, lane);
    } else {
        // lane `sub` (< 4) of the group probes candidate `sub` alone
        const int q = sub & 3;
        // mask blend rather than k[q], which would force the array to scratch
        Key k1{0, 0};
        uint64_t o1 = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t m = 0ull - (uint64_t)(q == i);
            k1.hi |= k[i].hi & m;
            k1.lo |= k[i].lo & m;
            o1 |= ord[i] & m;
        }
        const bool en1 = sub < 4 && ((enable >> q) & 1u);
        Sink::finish(d, k1, o1, (mm >> q) & 1u, en1, lane);
    }
}

// ~, X, F, G of one row (SEL: mask of the four computed).  G lanes per row;
// lane `sub` folds traces sub, sub+G, ... in groups whose loads are issued
// before their first use.
template <int KEY, bool CHECK, int G, int SEL, class Sink = GlobalSink>
__device__ __forceinline__ void expand_unary(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    constexpr int W = 64 / G;
    const int sub = lane % G;
    const uint64_t x = wl * W + (uint64_t)(lane / G);
    const bool active = x < tk.nl;
    const uint64_t xr = active ? x : 0;
    const gu64* __restrict__ L = glob(d.lvl[tk.lvl_l]);
    Key k[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    uint32_t mm = 0;
    // target checks only: candidates that do not count are settled from the
    // start, and the traces left are skipped once every candidate this lane
    // sees (and every other lane's) has missed one
    const uint32_t dead = CHECK ? (active ? (~(uint32_t)tk.part & 0xfu) : 0xfu) : 0u;
    const int T = d.T;
    // the running key half of KEY_MIX and the OR of (word ^ target) per candidate
    uint64_t acc[4] = {0, 0, 0, 0}, mq[4] = {0, 0, 0, 0};
    const int tmid = trace_split<KEY, CHECK>(d);
    int tb = sub;
    for (int sg = 0; sg < (tmid < T ? 2 : 1); ++sg) {
        const int te = sg == 0 ? tmid : T;
        for (int t0 = tb; t0 < te; t0 += G * kTraceGroup) {
            if constexpr (CHECK) {
                if (!__ballot((mm | dead) != 0xfu)) break;
            }
            uint64_t av[kTraceGroup];
            // (a check visits the traces in check_order; keys need them in order)
            const uint64_t o8 = CHECK && G == 1 ? check_order8(d, t0) : 0;
            auto trace = [&](int j) -> int {
                if constexpr (!CHECK) return t0 + j * G;
                else if constexpr (G == 1) return (int)((o8 >> (8 * j)) & 0xffu);
                else return d.check_order[t0 + j * G];
            };
#pragma unroll
            for (int j = 0; j < kTraceGroup; ++j)
                if (t0 + j * G < te) av[j] = L[(uint64_t)trace(j) * tk.nl + xr];
#pragma unroll
            for (int j = 0; j < kTraceGroup; ++j) {
                if (t0 + j * G >= te) break;
                const int t = trace(j);
                const uint64_t a = av[j];
                const uint64_t m = d.lenmask[t];
                const uint64_t want = (KEY != KEY_MIX || CHECK) && d.bs_mode ? d.bs_target : (t < d.P ? 1ull : 0ull);
                if constexpr (CHECK) {
                    // only bit 0 (time 0) matters: closed forms of the operators there
                    mm |= bit0_mismatch(~a & 1ull, (a >> 1) & 1ull, a != 0, a == m, want);
                    continue;
                }
                const uint64_t c0 = ~a & m;                      // ~
                if constexpr ((SEL & 1) != 0) fold_word<KEY>(k[0], acc[0], mq[0], c0, want, d, t);
                if constexpr ((SEL & 0xe) != 0) {
                    fold_word<KEY>(k[1], acc[1], mq[1], a >> 1, want, d, t);                // X
                    fold_word<KEY>(k[2], acc[2], mq[2], suffix_or(a), want, d, t);          // F
                    fold_word<KEY>(k[3], acc[3], mq[3], a & ~suffix_or(c0), want, d, t);    // G
                }
            }
        }
        if (sg == 0) next_split<KEY, G>(k, acc, tb, tmid, sub);
    }
    if constexpr (!CHECK) mm = fold_end<KEY>(k, acc, mq, d);
    const uint64_t ord[4] = {tk.base[0] + x, tk.base[1] + x, tk.base[2] + x, tk.base[3] + x};
    // SEL_ALL also serves any other mask (the cost function's): tk.part enables
    const uint32_t win = CHECK ? 0xfu : in_window(d, ord);
    finish_group<KEY, CHECK, G, Sink>(d, k, ord, mm, active ? ((uint32_t)SEL & tk.part & win) : 0u, lane);
}

template <int KEY, int G, class Sink = GlobalSink>
__device__ __forceinline__ void expand_unary_part(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    if (tk.part == SEL_NOT) expand_unary<KEY, false, G, SEL_NOT, Sink>(d, tk, wl, lane);
    else if (tk.part == SEL_XFG) expand_unary<KEY, false, G, SEL_XFG, Sink>(d, tk, wl, lane);
    else expand_unary<KEY, false, G, SEL_ALL, Sink>(d, tk, wl, lane);
}

// &, |, l U r, r U l of one pair (left row uniform per wavefront); SEL is
// the mask of the four computed
template <int KEY, bool CHECK, int G, int SEL, class Sink = GlobalSink>
__device__ __forceinline__ void expand_binary(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
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
    // target checks only: early exit as in expand_unary
    const uint32_t dead = CHECK ? (active ? (~(uint32_t)tk.part & 0xfu) : 0xfu) : 0u;
    const int T = d.T;
    uint64_t acc[4] = {0, 0, 0, 0}, mq[4] = {0, 0, 0, 0};
    const int tmid = trace_split<KEY, CHECK>(d);
    int tb = sub;
    for (int sg = 0; sg < (tmid < T ? 2 : 1); ++sg) {
        const int te = sg == 0 ? tmid : T;
        for (int t0 = tb; t0 < te; t0 += G * kTraceGroup) {
            if constexpr (CHECK) {
                if (!__ballot((mm | dead) != 0xfu)) break;
            }
            uint64_t av[kTraceGroup], bv[kTraceGroup];
            const uint64_t o8 = CHECK && G == 1 ? check_order8(d, t0) : 0;
            auto trace = [&](int j) -> int {
                if constexpr (!CHECK) return t0 + j * G;
                else if constexpr (G == 1) return (int)((o8 >> (8 * j)) & 0xffu);
                else return d.check_order[t0 + j * G];
            };
#pragma unroll
            for (int j = 0; j < kTraceGroup; ++j)
                if (t0 + j * G < te) {
                    const uint64_t tt = (uint64_t)trace(j);
                    av[j] = L[tt * tk.nl + l];
                    bv[j] = R[tt * tk.nr + rr];
                }
#pragma unroll
            for (int j = 0; j < kTraceGroup; ++j) {
                if (t0 + j * G >= te) break;
                const int t = trace(j);
                const uint64_t a = av[j], b = bv[j];
                const uint64_t want = (KEY != KEY_MIX || CHECK) && d.bs_mode ? d.bs_target : (t < d.P ? 1ull : 0ull);
                if constexpr (CHECK) {
                    // the probe and check-only levels evaluate every operator
                    mm |= bit0_mismatch(a & b & 1ull, (a | b) & 1ull, until_bit0(a, b), until_bit0(b, a), want);
                    continue;
                }
                if constexpr ((SEL & 1) != 0) fold_word<KEY>(k[0], acc[0], mq[0], a & b, want, d, t);
                if constexpr ((SEL & 2) != 0) fold_word<KEY>(k[1], acc[1], mq[1], a | b, want, d, t);
                if constexpr ((SEL & 0xc) != 0) {
                    fold_word<KEY>(k[2], acc[2], mq[2], until_op(a, b), want, d, t);
                    fold_word<KEY>(k[3], acc[3], mq[3], until_op(b, a), want, d, t);
                }
            }
        }
        if (sg == 0) next_split<KEY, G>(k, acc, tb, tmid, sub);
    }
    if constexpr (!CHECK) mm = fold_end<KEY>(k, acc, mq, d);
    const uint64_t pair = l * tk.nr + r;
    const uint64_t ord[4] = {tk.base[0] + pair, tk.base[1] + pair, tk.base[2] + 2 * pair,
                             tk.base[2] + 2 * pair + 1};
    const uint32_t win = CHECK ? 0xfu : in_window(d, ord);
    finish_group<KEY, CHECK, G, Sink>(d, k, ord, mm, active ? ((uint32_t)SEL & tk.part & win) : 0u, lane);
}

template <int KEY, int G, class Sink = GlobalSink>
__device__ __forceinline__ void expand_binary_part(const InstDesc& d, const Task& tk, uint64_t wl, int lane) {
    switch (tk.part) {
        case SEL_AND: expand_binary<KEY, false, G, SEL_AND, Sink>(d, tk, wl, lane); break;
        case SEL_ANDOR: expand_binary<KEY, false, G, SEL_ANDOR, Sink>(d, tk, wl, lane); break;
        case SEL_ORU: expand_binary<KEY, false, G, SEL_ORU, Sink>(d, tk, wl, lane); break;
        case SEL_U: expand_binary<KEY, false, G, SEL_U, Sink>(d, tk, wl, lane); break;
        default: expand_binary<KEY, false, G, SEL_ALL, Sink>(d, tk, wl, lane); break;
    }
}

__device__ __forceinline__ int seg_find(const Seg* segs, int nseg, uint64_t g) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].begin <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Winner test of fresh-list entry e: the candidate still holds its key's
// lowest ordinal.  The last pass of a level marks it in the byte map; every
// pass reports a matching winner below found_below.
__device__ __forceinline__ void resolve_entry(const InstDesc& d, uint64_t e) {
    const uint64_t of = glob(d.fresh_ord)[e];
    // a pass before the last only looks for a settled match
    if (!d.mark && !(of & kMatchBit)) return;
    const uint32_t slot = glob(d.fresh_slot)[e];
    const uint64_t o = of & ~kMatchBit;
    if (glob(d.table)[slot].ord == o) {
        const uint64_t local = o - d.epoch_base;
        // each ordinal appears at most once in the fresh list: a plain byte
        // store, packed to bits (and cleared) by the count / scan kernels
        if (d.mark) glob(d.winbyte)[local] = 1;
        if ((of & kMatchBit) && local < d.found_below) min64(glob(&d.ctr->found), local);
    }
}

// bit k of the result = byte k of x, for bytes that are 0 or 1
__device__ __forceinline__ uint64_t gather8(uint64_t x) {
    x &= 0x0101010101010101ull;
    x = (x | (x >> 7)) & 0x0003000300030003ull;
    x = (x | (x >> 14)) & 0x0000000F0000000Full;
    return (x | (x >> 28)) & 0xFFull;
}

// A direct level's first pass (found_below = m0 + 1) builds the candidates up
// to the probe's match m0, the only matching candidate among them.  When m0
// lost, another pass builds the rest and packs the level's bytes again, so
// this pass leaves them in place (keep).  Otherwise it packs and clears them
// but m0's, which every packing thread reads for this test, and the scan
// clears that byte last (scan_one).  Every other pass packs and clears.
__device__ __forceinline__ bool keep_marks(const InstDesc& d) {
    if (!d.direct || d.found_below == ~0ull) return false;
    const uint64_t m0 = d.found_below - 1;
    return !(glob(d.winbyte)[m0] == 3 && glob(d.lostbyte)[m0] == 0);
}

// packs the 64 winner bytes of bitmap word w into the word and clears them
// (unless keep); returns the word's popcount.  A direct level's winners are
// the candidates marked fresh and not displaced, and its lowest matching
// winner below found_below is reported here, reduced over the wavefront
// (every lane calls this for one word).
__device__ __forceinline__ uint32_t pack_word(const InstDesc& d, uint64_t w, bool keep) {
    typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
    uint64_t bits = 0, match = 0;
    if (w < d.nwords) {
        GLB u64x2* src = (GLB u64x2*)(d.winbyte + w * 64);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const u64x2 v = src[q];
            bits |= gather8(v.x) << (16 * q);
            bits |= gather8(v.y) << (16 * q + 8);
            if (d.direct) {
                match |= gather8(v.x >> 1) << (16 * q);
                match |= gather8(v.y >> 1) << (16 * q + 8);
            }
        }
        if (d.direct) {
            GLB u64x2* lost = (GLB u64x2*)(d.lostbyte + w * 64);
            uint64_t lb = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const u64x2 v = lost[q];
                lb |= gather8(v.x) << (16 * q);
                lb |= gather8(v.y) << (16 * q + 8);
            }
            bits &= ~lb;
            match &= bits;
            if (!keep) {
#pragma unroll
                for (int q = 0; q < 4; ++q) lost[q] = u64x2{0ull, 0ull};
            }
        }
        // the map starts every level cleared
        if (!keep) {
            u64x2 z[4] = {{0ull, 0ull}, {0ull, 0ull}, {0ull, 0ull}, {0ull, 0ull}};
            if (d.direct && d.found_below != ~0ull && (d.found_below - 1) >> 6 == w) {
                const uint32_t b = (uint32_t)((d.found_below - 1) & 63);   // m0 stays marked (3)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if ((int)(b >> 4) == q) {
                        if ((b >> 3) & 1u) z[q].y = 3ull << (8 * (b & 7));
                        else z[q].x = 3ull << (8 * (b & 7));
                    }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) src[q] = z[q];
        }
        glob(d.winbits)[w] = bits;
    }
    if (d.direct) {
        uint64_t f = match ? w * 64 + (uint64_t)__builtin_ctzll(match) : kEmpty;
        if (f >= d.found_below) f = kEmpty;
        if (__ballot(f != kEmpty)) {
            for (int off = 32; off > 0; off >>= 1) {
                const uint64_t o = __shfl_xor(f, off);
                f = o < f ? o : f;
            }
            if ((threadIdx.x & 63) == 0) min64(glob(&d.ctr->found), f);
        }
    }
    return (uint32_t)__popcll(bits);
}

// packs and counts each chunk (64 bitmap words = 4096 ordinals) of every

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    return v;
}

// One workgroup (256 threads) per instance: exclusive scan of the chunk
// counts, the level's survivor total, and the rank of the returned formula
// among the survivors.
__device__ void scan_one(const InstDesc& d) {
    __shared__ uint64_t wsum[4];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool keep = keep_marks(d);
    if (d.nblk <= (uint64_t)kSmallChunks) {
        // small bitmap: count the chunks here (one wavefront per chunk)
        for (uint64_t c = w; c < d.nblk; c += 4) {
            uint32_t v = pack_word(d, c * kChunkWords + lane, keep);
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) d.blkcount[c] = v;
        }
    }
    __syncthreads();
    // kScanPer consecutive chunks per thread: a level of 2^30 candidates has
    // 2^18 chunks, which one chunk per thread scanned in ~1000 rounds of
    // three barriers (2.2 ms)
    constexpr int kScanPer = 8;
    for (uint64_t base = 0; base < d.nblk; base += 256 * kScanPer) {
        const uint64_t i0 = base + (uint64_t)threadIdx.x * kScanPer;
        uint64_t v[kScanPer];
        uint64_t tsum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            v[j] = i0 + j < d.nblk ? d.blkcount[i0 + j] : 0;
            tsum += v[j];
        }
        const uint64_t inc = wave_incl_scan(tsum);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint64_t pre = carry;
        for (int k = 0; k < w; ++k) pre += wsum[k];
        uint64_t run = pre + inc - tsum;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            if (i0 + j < d.nblk) d.blkprefix[i0 + j] = run;
            run += v[j];
        }
        __syncthreads();
        if (threadIdx.x == 255) carry = pre + inc;
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        // set by atomics earlier in this kernel (fused form) or before it
        const uint64_t f = __hip_atomic_load(&d.ctr->found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t rank = 0;
        if (f != kEmpty) {
            const uint64_t fw = f >> 6, fc = fw / kChunkWords;
            const uint64_t w = fc * kChunkWords + lane;
            uint64_t bits = w < fw ? d.winbits[w] : (w == fw ? d.winbits[w] & ((1ull << (f & 63)) - 1ull) : 0ull);
            uint32_t v = (uint32_t)__popcll(bits);
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            rank = d.blkprefix[fc] + v;
        }
        if (threadIdx.x == 0) {
            // the first pass's m0, kept marked while the bytes were packed
            if (d.direct && d.found_below != ~0ull && !keep) glob(d.winbyte)[d.found_below - 1] = 0;
            d.ctr->winners = carry;
            d.ctr->found_rank = rank;
            if (d.ctr_host) {
                // the epoch's outcome straight to host memory: no read-back copy
                Ctr* h = d.ctr_host;
                h->fresh_count = d.ctr->fresh_count;
                h->fresh_real = d.ctr->fresh_real;
                h->claims = d.ctr->claims;
                h->found = f;
                h->winners = carry;
                h->found_rank = rank;
                h->overflow = d.ctr->overflow;
            }
        }
    }
}

// operands of the candidate with level-local ordinal o
struct Cand {
    int op;
    const gu64* L;
    const gu64* R;
    uint64_t nl, nr, l, r;
    bool swap;
};

__device__ __forceinline__ int unit_of(const InstDesc& d, uint64_t o) {
    int lo = 0, hi = d.nunits - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d.units[mid].base <= o) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ Cand decode_unit(const InstDesc& d, const Unit& u, uint64_t o) {
    Cand c;
    c.op = u.op;
    c.L = glob(d.lvl[u.lvl_l]);
    c.R = glob(d.lvl[u.lvl_r]);
    c.nl = u.nl;
    c.nr = u.nr;
    c.swap = false;
    const uint64_t k = o - u.base;
    if (u.op == VFB_OP_NOT || u.op == VFB_OP_NEXT || u.op == VFB_OP_FINALLY || u.op == VFB_OP_GLOBALLY) {
        c.l = k;
        c.r = 0;
        c.R = c.L;
    } else {
        const uint64_t p = u.op == VFB_OP_UNTIL ? k >> 1 : k;
        c.l = p / u.nr;
        c.r = p - c.l * u.nr;
        c.swap = u.op == VFB_OP_UNTIL && (k & 1);
    }
    return c;
}

__device__ __forceinline__ Cand decode_from(const InstDesc& d, uint64_t o, int u) {
    while (u + 1 < d.nunits && d.units[u + 1].base <= o) ++u;
    return decode_unit(d, d.units[u], o);
}

__device__ __forceinline__ uint64_t cand_word(const Cand& c, int t, uint64_t m) {
    const uint64_t a = c.L[(uint64_t)t * c.nl + c.l];
    if (c.op == VFB_OP_NOT || c.op == VFB_OP_NEXT || c.op == VFB_OP_FINALLY || c.op == VFB_OP_GLOBALLY)
        return apply_op(c.op, a, 0, m);
    const uint64_t b = c.R[(uint64_t)t * c.nr + c.r];
    if (c.op == VFB_OP_UNTIL) return c.swap ? until_op(b, a) : until_op(a, b);
    return apply_op(c.op, a, b, m);
}

// Traces t0..t1-1 of candidate c into rows `row` of out (column-major, out_n
// rows): the operand words of kStoreGroup traces are loaded before any of
// them is stored, so their loads are in flight together (a store between
// them would otherwise order each trace's loads after the last one's store).
constexpr int kStoreGroup = 8;
__device__ __forceinline__ void cand_words_out(const Cand& c, int t0, int t1, const uint64_t* lenmask, gu64* out,
                                               uint64_t out_n, uint64_t row) {
    const bool unary = c.op == VFB_OP_NOT || c.op == VFB_OP_NEXT || c.op == VFB_OP_FINALLY || c.op == VFB_OP_GLOBALLY;
    for (int t = t0; t < t1; t += kStoreGroup) {
        uint64_t a[kStoreGroup], b[kStoreGroup];
#pragma unroll
        for (int j = 0; j < kStoreGroup; ++j) {
            a[j] = t + j < t1 ? c.L[(uint64_t)(t + j) * c.nl + c.l] : 0ull;
            b[j] = t + j < t1 && !unary ? c.R[(uint64_t)(t + j) * c.nr + c.r] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < kStoreGroup; ++j) {
            if (t + j >= t1) break;
            uint64_t v;
            if (c.op == VFB_OP_UNTIL) v = c.swap ? until_op(b[j], a[j]) : until_op(a[j], b[j]);
            else v = apply_op(c.op, a[j], b[j], lenmask[t + j]);
            out[(uint64_t)(t + j) * out_n + row] = v;
        }
    }
}

// trace t of candidate c over its words (multi-word rows), last word first
__device__ __forceinline__ void store_trace_mw(const InstDesc& d, const Cand& cd, int t, gu64* out,
                                               uint64_t out_n, uint64_t row) {
    const int c0 = d.tcol[t];
    const bool unary = cd.op == VFB_OP_NOT || cd.op == VFB_OP_NEXT || cd.op == VFB_OP_FINALLY ||
                       cd.op == VFB_OP_GLOBALLY;
    uint64_t carry = 0;
    bool later_any = false, later_all = true;
    for (int c = d.tcol[t + 1] - 1; c >= c0; --c) {
        const uint64_t m = d.cinfo[c].lenmask;
        uint64_t a = cd.L[(uint64_t)c * cd.nl + cd.l];
        uint64_t b = unary ? 0ull : cd.R[(uint64_t)c * cd.nr + cd.r];
        if (cd.swap) { const uint64_t x = a; a = b; b = x; }   // r U l
        uint64_t v;
        switch (cd.op) {
        case VFB_OP_NOT: v = ~a & m; break;
        case VFB_OP_AND: v = a & b; break;
        case VFB_OP_OR: v = a | b; break;
        case VFB_OP_NEXT: v = (a >> 1) | (carry << 63); carry = a & 1ull; break;
        case VFB_OP_FINALLY: v = later_any ? m : suffix_or(a); later_any = later_any || a != 0ull; break;
        case VFB_OP_GLOBALLY:
            v = later_all ? (a & ~suffix_or(~a & m)) : 0ull;
            later_all = later_all && a == m;
            break;
        default:
            v = until_op(a, b) | (carry ? ~suffix_or(~a) : 0ull);
            carry = v & 1ull;
            break;
        }
        out[(uint64_t)c * out_n + row] = v;
    }
}

// One work item of the stable compaction (k_store; the chain's store phase):
// chunk c of 4096 ordinals, survivor subset `part` of SP, a slice of
// d.store_traces traces.  The chunk's set bits become an LDS list of 12-bit
// offsets, then the 64 lanes write consecutive rows of the slice's columns
// (coalesced stores); slice 0 also writes the parent ordinals.
template <int SP>
__device__ __forceinline__ void store_item(const InstDesc& d, uint64_t item, uint16_t* list, gu64* out_rows,
                                           gu64* out_par, uint64_t out_n, int lane) {
    const int st = d.store_traces;
    const uint32_t nsl = (uint32_t)((d.T + st - 1) / st);
    const uint64_t c = item / (nsl * SP);
    const uint32_t rem = (uint32_t)(item - c * nsl * SP);
    const uint32_t part = rem / nsl;
    const int t0 = (int)(rem - part * nsl) * st;
    const int t1 = t0 + st < d.T ? t0 + st : d.T;
    const uint32_t n = glob(d.blkcount)[c];
    const uint64_t rank0 = glob(d.blkprefix)[c];
    if (n == 0 || rank0 >= out_n) return;
    const uint64_t w = c * kChunkWords + lane;
    uint64_t bits = w < d.nwords ? glob(d.winbits)[w] : 0ull;
    const uint64_t pc = __popcll(bits);
    uint32_t pos = (uint32_t)(wave_incl_scan(pc) - pc);
    while (bits) {
        const int q = __ffsll((unsigned long long)bits) - 1;
        bits &= bits - 1;
        list[pos++] = (uint16_t)(lane * 64 + q);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const uint64_t obase = c * (uint64_t)(kChunkWords * 64);
    const uint32_t lim = rank0 + n <= out_n ? n : (uint32_t)(out_n - rank0);
    // the chunk's 4096 ordinals span few units: find the first one once
    const int u0 = unit_of(d, obase);
    for (uint32_t i = part * 64 + lane; i < lim; i += 64 * SP) {
        const uint64_t o = obase + list[i];
        const Cand cd = decode_from(d, o, u0);
        const uint64_t row = rank0 + i;
        if (d.mw) {
            for (int t = t0; t < t1; ++t) store_trace_mw(d, cd, t, out_rows, out_n, row);
        } else {
            cand_words_out(cd, t0, t1, d.lenmask, out_rows, out_n, row);
        }
        if (t0 == 0) out_par[row] = o;
    }
    __builtin_amdgcn_wave_barrier();
}

}  // namespace vfb
