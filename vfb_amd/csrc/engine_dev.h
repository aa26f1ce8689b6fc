// Device-visible data structures of the CS enumeration engine.
//
// HBM layout (per instance):
//   level block s : column-major uint64 words, word(c, row) = lvl[s][c * nrows[s] + row]
//                   so a wavefront whose 64 lanes take 64 consecutive rows reads
//                   512 contiguous bytes per column (coalesced), and a uniform row
//                   (the left operand of a binary wave) is a broadcast.  Column c
//                   is trace c when every trace has at most 64 steps; otherwise
//                   trace t of n steps owns ceil(n/64) consecutive columns
//                   tcol[t] .. tcol[t+1]-1, word w holding steps 64w..64w+63.
//   parents       : one uint64 per stored row = its level-local canonical ordinal
//                   (decodes to operator, operands), replacing the reference's
//                   left/right index arrays and segment table (LTL_BS.cu:1189-1205).
//   hash set      : open addressing, 32-byte slots {hi, lo, ord, pad}, power-of-two
//                   capacity, linear probing.  Replaces WarpCore's two-level
//                   64-bit scheme (LTL_BS.cu:251-269, hash_set.cuh:192-254).
//   epoch scratch : fresh list (ordinal|match flag, slot), winner bitmap over the
//                   level's ordinal space, per-4096-word block counts / prefixes.
#pragma once

#include <cstdint>

#include "../../include/vfb.h"

namespace vfb {

constexpr int kMaxT = VFB_MAX_TRACES;
constexpr int kMaxS = VFB_MAX_SIZE;
constexpr int kMaxUnits = 4 + 3 * (kMaxS / 2);   // ~,X,F,G + (&,|,U) per split
constexpr uint64_t kEmpty = ~0ull;
constexpr uint64_t kMatchBit = 1ull << 63;
constexpr int kChunkWords = 64;        // bitmap words per scan chunk (4096 ordinals)
constexpr int kProbeLimit = 4096;
constexpr int kStoreTraces = 8;       // traces per k_store work item (levels below kStoreLargeLevel)
constexpr uint64_t kStoreLargeLevel = 1ull << 22;  // levels of this many candidates take vfb_tuning.store_traces
constexpr int kStoreParts = 4;        // survivor subsets per chunk in k_store
constexpr int kStorePartsSmall = 16;  // the same when every level of the launch has at most
constexpr uint64_t kStoreSmallBlk = 16;  // this many chunks (more, shorter wavefronts)
constexpr int kWideLanes = 8;         // lanes per candidate in k_expand's small-level mode
constexpr uint64_t kWideMaxWaves = 512;  // levels below this many wavefronts use it (default)
constexpr int kTraceGroup = 8;        // traces whose operand loads k_expand issues together
constexpr int kCheckRows = 8;         // left rows per wavefront of k_check's binary tasks
constexpr int kSmallChunks = 64;      // bitmaps up to this many chunks are counted in k_scan
constexpr uint64_t kSmallFresh = 4096;  // fresh lists up to this long are resolved in the scan's workgroup
constexpr uint64_t kMinWinWords = 1024; // winner maps are allocated for at least 64K candidates
constexpr uint64_t kMinSlots = 1ull << 17; // hash sets start at 4 MB
constexpr uint64_t kDirectMin = 1ull << 18;  // default vfb_tuning.direct_min (InstDesc::direct)

enum KeyKind : int { KEY_FIELD = 0, KEY_STRIDE = 1, KEY_MIX = 3 };

struct alignas(32) Slot {
    unsigned long long hi, lo, ord, pad;
};

// One column of a row whose traces span several words (InstDesc::mw): its
// trace, its word within the trace, the mask of its valid steps and its part
// of the key (same meaning as InstDesc's per-trace kmask / koff / kbits;
// kfirst = bit of the column's first KEY_STRIDE sample).
struct ColInfo {
    uint64_t lenmask;
    uint64_t kmask;
    int16_t koff, kbits, kfirst, word;
    int32_t trace, pad;
};

struct Ctr {
    unsigned long long fresh_count;  // entries appended to the fresh list this epoch (none with InstDesc::direct)
    unsigned long long claims;       // slots claimed this epoch (= distinct new keys)
    unsigned long long found;        // lowest matching winner, level-local ordinal
    unsigned long long winners;      // survivors of the epoch (set by the scan)
    unsigned long long found_rank;   // survivors ordered before `found`
    unsigned int overflow;           // 1 probe limit, 2 fresh list full, 4 load limit
    unsigned int pad;
    unsigned long long fresh_real;   // with InstDesc::direct: the fresh candidates (marked, not listed)
};

// One segment of a level's canonical candidate order: operator `op` applied to
// rows of level lvl_l (and lvl_r for binary operators).
struct Unit {
    int op;
    int lvl_l, lvl_r;
    uint32_t nl, nr;
    uint64_t base;    // first level-local ordinal
    uint64_t count;   // candidates (Until: 2 per pair)
};

struct InstDesc {
    int T, P, key_kind, half;
    int C;            // words per row (= T unless mw)
    int mw;           // some trace spans several words: cinfo / tcol describe the columns
    int stride;       // KEY_STRIDE sampling stride
    int level;        // level being built
    int nunits;
    int check_only;   // onTheFly level: target check only, no keys, no storage
    int mark;         // resolve: 1 sets the winner bitmap (last pass), 0 only reports a match
    int bs_mode;      // BS stage: one word per row, a match is the word == bs_target
    int store_traces; // traces per k_store work item
    int defer_claims; // the launch holds this instance alone: k_expand adds up its claims per
                      // workgroup in LDS and checks the set's load limit at the workgroup's end
    int direct;       // winners marked during the expansion: a fresh candidate sets its byte in
                      // winbyte (bit 0, bit 1 = match), one that lowers a slot's ordinal of this
                      // epoch sets the displaced ordinal's byte in lostbyte; no fresh list, no resolve
    int pad_desc;
    uint64_t lenmask[kMaxT];
    uint64_t kmask[kMaxT];      // KEY_FIELD: low bits of the trace word that enter the key
    int16_t koff[kMaxT];        // first key bit of the trace's field / samples
    int16_t kbits[kMaxT];       // field width (KEY_FIELD) or number of samples (KEY_STRIDE)
    uint8_t check_order[kMaxT]; // target checks visit the traces positives and negatives
                                // alternating, so that most candidates miss one early
    const ColInfo* cinfo;       // mw: [C] column descriptors
    const int32_t* tcol;        // mw: [T + 1] first column of each trace
    const unsigned long long* lvl[kMaxS + 2];
    uint32_t nrows[kMaxS + 2];
    Slot* table;
    uint64_t table_mask;
    uint64_t claim_limit;       // claims allowed this epoch before the set must grow
    uint64_t epoch_base;        // global ordinal of the level's first candidate
    uint64_t found_below;       // resolve: only matches below this level-local ordinal are reported
    uint64_t ord_lo, ord_hi;    // dedup passes build only candidates with level-local ordinals in [ord_lo, ord_hi]
    uint64_t bs_target;         // BS stage: the positive-traces mask
    unsigned long long* fresh_ord;
    uint32_t* fresh_slot;
    uint64_t fresh_cap;
    unsigned char* winbyte;      // one byte per candidate, set by the resolve (plain stores), or
                                 // by the expansion (direct)
    unsigned char* lostbyte;     // direct: one byte per candidate, set when it was displaced
    unsigned long long* winbits; // the same, packed to bits by the count / scan kernels
    uint64_t nwords;
    uint32_t* blkcount;
    unsigned long long* blkprefix;
    uint64_t nblk;
    unsigned long long* out_rows;
    unsigned long long* out_par;
    uint64_t out_n;             // rows stored of this level (survivors past it are dropped)
    Ctr* ctr;
    Ctr* ctr_host;              // host-coherent copy the scan writes at the end of the epoch (or null)
    Unit units[kMaxUnits];
};

// Fused work item: all operators that read the same operand rows.
//   kind 0: rows of level lvl_l -> ~, X, F, G     (base[0..3])
//   kind 1: pairs (level lvl_l x level lvl_r) -> &, |, lUr, rUl  (base[0..2])
// `part` is the mask of the task's four candidates that are computed: bit q
// selects base[q]'s operator (~ X F G, or & | lUr rUl).  The masks in use:
enum {
    SEL_ALL = 0xf,
    SEL_NOT = 0x1, SEL_XFG = 0xe,                               // unary
    SEL_AND = 0x1, SEL_ANDOR = 0x3, SEL_ORU = 0xe, SEL_U = 0xc  // binary
};
struct Task {
    uint32_t inst;
    uint8_t kind, lvl_l, lvl_r, part;
    uint32_t nl, nr;
    uint32_t nchunk;            // ceil(nr / candidates per wavefront)
    uint16_t ops;               // candidates of base[] that exist at this level (cost function)
    uint8_t lrows;              // k_check, binary: left rows per wavefront (one right word load
                                // serves them all); 0 or 1 otherwise
    uint8_t pad8;
    uint64_t base[4];
    uint64_t wave_begin;        // prefix over tasks, in wavefronts
};

// One segment of a batched fill / copy: dst[i] = src ? src[i] : value for
// i < words, as chunks [begin, begin + ceil(words / 8)) of the flat chunk
// space (8 words per chunk).
struct OpSeg {
    unsigned long long* dst;
    const unsigned long long* src;
    uint64_t words;
    uint64_t value;
    uint64_t begin;
};

// One segment of a batched rehash: slots [begin, begin + old_cap) of the flat
// index space are the old table's, moved into tab if ord < keep_below.
struct RehashSeg {
    const Slot* old;
    Slot* tab;
    uint64_t mask;
    uint64_t keep_below;
    uint64_t begin;
};

// Output of a level's store, passed by value when one instance is stored
// (its descriptor is then the one already on the device from the epoch).
struct StorePatch {
    unsigned long long* out_rows;
    unsigned long long* out_par;
    uint64_t out_n;
};

// Segment of a flat index space belonging to one instance.
struct Seg {
    uint32_t inst, pad;
    uint64_t begin;             // prefix in the flat space
};

// ---------------------------------------------------------------------------
// Persistent level chain (k_chain): one launch builds, for each instance, the
// small levels that follow its last completed one -- expansion, resolve, scan
// and store, separated by barriers of the instance's team of workgroups --
// and returns to the host at the first level that is too large, that holds a
// match, or that the host must build itself.
// ---------------------------------------------------------------------------
enum ChainExit : int32_t {
    CH_BEFORE = 0,   // stopped before level+nlev+1 (too large, max size, or its set must grow)
    CH_FOUND = 1,    // level+nlev+1 holds the returned formula (counters in ChainOut)
    CH_REDO = 2,     // level+nlev+1 must be built again by the host (overflow bits: 1 probe
                     // limit, 2 the host's fresh list would overflow, 4 set load)
    CH_ERROR = 3     // a team barrier timed out
};
constexpr int kChainThreads = 256;
constexpr int kChainMaxTeam = 64;               // workgroups per instance, at most
constexpr uint64_t kChainMaxCands = 1ull << 18; // a chain level's winner bitmap fits LDS (64 chunks)
constexpr uint64_t kChainSoloMax = 64;          // levels this small are built by a team's first workgroup alone
constexpr int kMaxTasks = 1 + kMaxS / 2;        // ~XFG task + one per split
// the level chain in LDS (k_chain_lds): its set's slots, the keys it takes
// (load 50 %), and the candidates of a level it builds (one 4096-ordinal chunk)
constexpr int kLdsChainSlots = 4096;
constexpr uint64_t kLdsChainKeys = kLdsChainSlots / 2;
constexpr uint64_t kLdsChainCands = 4096;

// per instance, uploaded for one launch
struct ChainIn {
    InstDesc d;                 // descriptor with lvl / nrows of levels 1..level
    unsigned long long* slab;   // rows of the chain's levels, back to back (256-byte aligned blocks)
    uint64_t slab_bytes;
    unsigned long long* fresh_ord;   // workspace of n_max candidates
    uint32_t* fresh_slot;
    uint32_t* bitmap[2];        // winners of team levels of even / odd size, n_max bits each (zero)
    uint64_t n_max;             // a level enters the chain only with at most this many candidates
    uint64_t solo_max;          // a level of at most this many is built by the team's first workgroup
    uint64_t rows_left;         // rows the solve may still store (max_rows), ~0 for no cap
    uint64_t committed;         // keys in the set
    uint64_t cap;               // slots of the set
    uint64_t gbase;             // global ordinal of the next level's first candidate
    uint64_t fresh_cap;         // the instance's own fresh list (as the host sizes it level by level)
    uint64_t pass_min;          // levels the host probes first (vfb_tuning.pass_min) are left to it
    double win_ratio;           // survivors / candidates of the last completed level
    double fresh_ratio;         // fresh entries / candidates of the last completed level
    double margin;              // the set holds the next level's survivors estimated as
                                // candidates * win_ratio * margin (vfb_tuning.set_margin_pct)
    uint32_t pct;               // set load (percent) at which the host grows the set
    int32_t level;              // last completed level
    int32_t last_level;         // the chain builds levels up to this one
    int32_t neg_type;
    Ctr ctr[2];                 // counters of even / odd levels
    unsigned long long bar;     // team barrier arrivals
};

// per instance, read back after the launch
struct ChainOut {
    int32_t exit, nlev;         // ChainExit; levels built and stored
    uint32_t overflow, err;     // CH_REDO's overflow bits; 1 if a barrier timed out
    uint64_t found, found_rank; // CH_FOUND: level-local ordinal, survivors ordered before it
    uint64_t N_exit, W_exit, fresh_exit;   // candidates / survivors / fresh entries of the exit level
    uint64_t N[kMaxS], W[kMaxS], fresh[kMaxS], off[kMaxS];   // chain level k is level + 1 + k
    // device clock (100 MHz) of the team's first workgroup: launch, then per
    // level the end of its plan, expansion, resolve, scan and store phases
    uint64_t t0, t[kMaxS][5];
};

}  // namespace vfb
