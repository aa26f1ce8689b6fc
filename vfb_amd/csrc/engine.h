// Internal: the CS enumeration engine shared by the solve entry points and
// the continuous-batching server.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "engine_dev.h"
#include "vfb_internal.h"

namespace vfb {

hipError_t launch_expand(const InstDesc* D, const Task* tasks, int ntasks, uint64_t total_waves, int wide, int mw, int grid,
                         hipStream_t st);
hipError_t launch_insert_atoms(const InstDesc* D, uint32_t inst, uint32_t n, unsigned int* fail, hipStream_t st);
hipError_t launch_rehash(const Slot* old, uint64_t old_cap, Slot* tab, uint64_t mask, uint64_t keep_below,
                         unsigned int* fail, int grid, hipStream_t st);
hipError_t launch_fill(unsigned long long* p, uint64_t n, unsigned long long v, int grid, hipStream_t st);
hipError_t launch_fresh_prefix(const InstDesc* D, Seg* segs, int n, unsigned long long* total, hipStream_t st);
hipError_t launch_resolve(const InstDesc* D, const Seg* segs, int nseg, const unsigned long long* total, int single,
                          int grid, hipStream_t st);
hipError_t launch_count(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_chunks, int grid,
                        hipStream_t st);
hipError_t launch_scan(const InstDesc* D, const uint32_t* insts, int n, hipStream_t st);
hipError_t launch_resolve_scan1(const InstDesc* D, hipStream_t st);
hipError_t launch_insert_atoms_multi(const InstDesc* D, const Seg* segs, int nseg, uint64_t total, unsigned int* fail,
                                     hipStream_t st);
hipError_t launch_fill_multi(const OpSeg* segs, int n, uint64_t total, int grid, hipStream_t st);
hipError_t launch_bs_words(const unsigned long long* rows, uint32_t n, int T, const int32_t* tcol,
                           unsigned long long* out, hipStream_t st);
hipError_t launch_insert_keys(Slot* tab, uint64_t mask, const unsigned long long* keys, uint64_t n, unsigned int* fail,
                              hipStream_t st);
hipError_t launch_rehash_multi(const RehashSeg* segs, int n, uint64_t total, unsigned int* fail, int grid,
                               hipStream_t st);
hipError_t launch_gather(const unsigned long long* const* addr, uint64_t n, unsigned long long* out, hipStream_t st);
// small: every level of the launch has at most kStoreSmallBlk chunks (kStorePartsSmall subsets)
hipError_t launch_store(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_items, StorePatch patch, int small,
                        int grid, hipStream_t st);

// persistent level chain (chain.hip): teams of G workgroups per instance
// (G > 1: a cooperative launch, else teams of one); workgroups of k_chain
// resident per CU
hipError_t launch_chain(ChainIn* ins, ChainOut* outs, int nteams, int G, hipStream_t st, int* G_used);
int chain_blocks_per_cu();
// the level chain in LDS (chain.hip): one workgroup per instance, its hash
// set in LDS (kLdsChain* in engine_dev.h)
hipError_t launch_chain_lds(ChainIn* ins, ChainOut* outs, int n, hipStream_t st);

struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string& s) : std::runtime_error(s) {}
};

#define CK(x)                                                                                   \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw DeviceError(std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

static uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// device arena: one hipMalloc, first-fit free list with coalescing
// ---------------------------------------------------------------------------
class Arena {
  public:
    explicit Arena(size_t bytes) : size_(bytes), owned_(true) {
        CK(hipMalloc(&base_, bytes));
        free_[0] = bytes;
    }
    // over device memory the caller owns (the no-allocation entry point's scratch)
    Arena(void* base, size_t bytes) : base_(base), size_(bytes), owned_(false) { free_[0] = bytes; }
    ~Arena() {
        if (owned_) (void)hipFree(base_);
    }
    void* alloc(size_t bytes) {
        bytes = std::max<size_t>(align256(bytes), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                size_t off = it->first, sz = it->second;
                free_.erase(it);
                if (sz > bytes) free_[off + bytes] = sz - bytes;
                used_[off] = bytes;
                return (char*)base_ + off;
            }
        }
        return nullptr;
    }
    // gives the tail of allocation p beyond `bytes` back to the free list
    void shrink(void* p, size_t bytes) {
        const size_t off = (char*)p - (char*)base_;
        auto it = used_.find(off);
        if (it == used_.end() || bytes >= it->second) return;
        const size_t tail = it->second - bytes;
        it->second = bytes;
        used_[off + bytes] = tail;
        release((char*)base_ + off + bytes);
    }
    void release(void* p) {
        if (!p) return;
        size_t off = (char*)p - (char*)base_;
        auto it = used_.find(off);
        if (it == used_.end()) return;
        size_t sz = it->second;
        used_.erase(it);
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + sz == nx->first) { sz += nx->second; nx = free_.erase(nx); }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) { pv->second += sz; return; }
        }
        free_[off] = sz;
    }
    size_t largest_free() const {
        size_t m = 0;
        for (auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    size_t size() const { return size_; }

  private:
    void* base_ = nullptr;
    size_t size_;
    bool owned_;
    std::map<size_t, size_t> free_;
    std::map<size_t, size_t> used_;
};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;   // not from the arena (an instance's resident atom rows): never released
};

// budget (the solve's own limit) or arena (memory shared with a batch) exhausted
struct CapacityError : std::runtime_error {
    bool arena;
    CapacityError(const std::string& s, bool a) : std::runtime_error(s), arena(a) {}
};

// an engine over fixed memory (vfb_solve_noalloc) needs more than it holds:
// the solve fails instead of switching to the onTheFly round, so that every
// result it returns equals vfb_solve's
struct ScratchExhausted : std::runtime_error {
    explicit ScratchExhausted(const std::string& s) : std::runtime_error(s) {}
};

// ---------------------------------------------------------------------------
// per-instance state
// ---------------------------------------------------------------------------
struct Inst {
    const vfb_instance* in = nullptr;
    vfb_result* res = nullptr;
    int max_size = 0;
    uint64_t budget = 0, used = 0, max_rows = 0;
    InstDesc h{};
    bool done = false;
    bool deferred = false;      // batch/server: evicted for memory, restarted later
    bool parked = false;        // batch: waiting, state kept, to go on alone after the others
    bool need_alone = false;    // deferred because the arena could not hold it beside the others
    uint64_t prio = 0;          // admission order: older instances evict younger ones
    bool check_only = false;    // this level is the onTheFly last round
    bool next_check_only = false;
    int level = 0;              // last completed level
    uint64_t gbase = 0;         // global ordinal of the next level's first candidate
    uint64_t stored = 0;        // rows stored over all levels
    std::vector<Buf> rows;      // per level: T columns of nrows words, then nrows parent ordinals
    std::vector<uint32_t> nrows;
    std::vector<std::vector<Unit>> units;
    Buf table;
    uint64_t cap = 0, committed = 0;
    Buf fresh_ord, fresh_slot, winbyte, winbits, blkcount, blkprefix;
    Buf cols;                   // multi-word rows: ColInfo[C] then tcol[T + 1]
    uint64_t fresh_cap = 0, nwords_cap = 0, nblk_cap = 0;
    double win_ratio = 1.0, fresh_ratio = 1.0;
    uint64_t lost_words = 0;    // a direct level's loser bytes (in fresh_ord) cleared for this many bitmap words
    uint64_t N = 0;             // candidates of the level being built
    std::vector<Task> tasks;
    Ctr ctr{};
    // A large level is first probed (target check only, no keys).  Without a
    // match it is built in one fused pass; with one it is built operator
    // block by operator block in canonical order and stops at the block that
    // settles the match.
    // BS stage (LTL_BS.cu:808-1067): Boolean combinations of this instance's
    // stored time-0 vectors, run as a second instance over one-word rows
    int cost[8] = {1, 1, 1, 1, 1, 1, 1, 1};  // p ~ & | X F G U (LTL_BS's c1..c8); level = cost
    int neg_type = 1;           // LTL_BS's NegType: 2 = ~ of atoms only
    bool unit_cost = true;
    int bs_max = 0;             // vfb_options.bs_max_size
    bool bs_mode = false;       // this instance is a BS stage
    bool bs_pending = false;    // the CS search ended without a match: start the BS stage
    Inst* ltl = nullptr;        // BS stage: the CS instance whose rows are its leaves
    int ltl_levels = 0;         // BS stage: leaf levels 1..ltl_levels (parents = CS rows)
    std::unique_ptr<Inst> bs;   // CS instance: its BS stage
    std::unique_ptr<vfb_instance> bs_in;   // BS stage: its one-trace instance
    int pass = 0;               // next pass of the level (enum Pass in engine.cpp)
    uint8_t usel = 0xf, bsel = 0xf;     // operators of the pass: unary / binary task masks
    uint8_t first_usel = 0xf, first_bsel = 0xf;
    uint64_t m0 = 0;            // the probe's lowest match: the first pass builds ordinals 0..m0
    uint64_t evaluated = 0;     // candidates of this level evaluated so far
    std::vector<Buf> slabs;     // level blocks built by the persistent chain (rows[] borrow from them)
    int chain_skip = 0;         // a level the chain refused or must not redo: built by the host
    // scratch sizes the host path would have reached over the chain's levels
    // (prepare_scratch grows to at least these at the next host level)
    uint64_t floor_fcap = 0, floor_nwords = 0, floor_nblk = 0;
};

class Engine {
  public:
    explicit Engine(size_t arena_bytes);
    // Over fixed device memory: no HIP allocation after construction (the
    // arena, descriptors and flag are carved from `dev`; host staging is
    // static), plain solves only, no timing events
    struct Fixed {
        void* dev;
        size_t bytes;
    };
    explicit Engine(const Fixed& m);
    ~Engine();
    void run(std::vector<Inst*>& insts, const vfb_options& opt);
    // incremental use (server): begin() sets the options; start() sets up new
    // instances and appends the unfinished ones to act; step() builds one
    // level for every instance in act and removes those that finished
    void begin(const vfb_options& opt, bool batch);
    void start(std::vector<Inst*>& fresh, std::vector<Inst*>& act);
    void step(std::vector<Inst*>& act);
    void drain() { sync(); }
    void abandon(Inst& I);
    void forget(std::vector<Inst*>& v);
    // members that cannot get arena memory while others hold it are deferred
    // (need_alone) instead of going on under the plain limit
    void set_defer_alone(bool b) { defer_alone_ = b; }
    size_t arena_size() const { return arena_->size(); }

  private:
    vfb_options opt_;
    bool fixed_ = false;          // built over fixed memory (Engine(const Fixed&))
    bool batch_ = false;
    hipStream_t st_ = nullptr;
    int cus_ = 256;
    std::unique_ptr<Arena> arena_;
    char* d_step_ = nullptr;      // per-phase packed descriptors / counters / tasks / segments
    size_t step_cap_ = 0;
    unsigned int* d_flag_ = nullptr;   // probe-limit failure of rehash / atom insertion
    unsigned int* h_flag_ = nullptr;   // its pinned host copy
    Ctr* h_ctr_ = nullptr;        // pinned read-back of an epoch's counters
    size_t h_ctr_cap_ = 0;
    bool flag_dirty_ = false;
    std::vector<void*> pending_free_;
    // batch mode: fills, uploads and rehashes of many instances are queued and
    // flushed as one launch each before the next expansion
    struct FillOp { Inst* I; unsigned long long* dst; size_t staged; uint64_t words; uint64_t value; };
    struct RehashOp { Inst* I; const Slot* old; uint64_t old_cap; Slot* tab; uint64_t mask; uint64_t keep; };
    std::vector<FillOp> fills_;
    std::vector<RehashOp> rehashes_;
    // old tables of queued rehashes: released to the arena after the flush
    // that launches their rehash (then at the following sync)
    struct RehashSrc { Inst* I; void* p; };
    std::vector<RehashSrc> rehash_src_;
    std::vector<unsigned long long> staged_;  // upload data of copy ops (staged != SIZE_MAX)
    char* d_ops_ = nullptr;
    size_t ops_cap_ = 0;
    char* pin_ = nullptr;         // pinned staging for uploads, recycled at sync points
    size_t pin_cap_ = 0, pin_used_ = 0;
    std::vector<hipEvent_t> ev_pool_;
    struct Timed {
        hipEvent_t a, b;
        int kind;                 // 0 expand, 1 resolve, 2 store, 3 probe, 4 level chain
        std::vector<Inst*> who;
    };
    std::vector<Timed> timed_;
    std::vector<Inst*> live_;     // started, not finished: eviction candidates
    Inst* epoch_single_ = nullptr; // the one instance whose descriptor is at d_step_[0]
    bool last_epoch_probe_only_ = false;   // the last epoch's expansion was timed as a probe
    vfb_tuning tun_{};            // launch / pass parameters, copied at begin()
    bool defer_alone_ = false;    // batch: an instance that cannot get arena memory is deferred
    bool park_ = false;           // batch: members reaching a level of tun_.park_min candidates wait
    std::vector<Inst*> parked_;   // those members, resumed one at a time after the rest
    // persistent level chain: workspace of the launch's instances (fresh
    // lists, winner maps, chunk counts; not part of any solve's budget),
    // laid out winner bitmaps first; the first chain_wb_clean_ bytes are zero
    char* d_chain_ws_ = nullptr;
    size_t chain_ws_cap_ = 0, chain_wb_clean_ = 0;
    ChainOut* h_chain_out_ = nullptr;   // pinned read-back
    size_t h_chain_out_cap_ = 0;
    int chain_resident_ = 0;      // k_chain workgroups resident on the device at once
    // formulas found in one step: their parent ordinals, read in batched
    // gathers (one per depth of the formula trees) instead of one copy per node
    struct FormulaRef { Inst* I; int level; uint64_t o; };
    std::unordered_map<const unsigned long long*, unsigned long long> par_cache_;
    unsigned long long* d_gather_ = nullptr;   // addresses, then values
    unsigned long long* h_gather_ = nullptr;   // pinned read-back
    size_t gather_cap_ = 0;

    void sync();
    void h2d(void* dst, const void* src, size_t bytes);
    char* step_buf(size_t bytes);
    void dalloc(Inst& I, Buf& b, size_t bytes);
    void dfree(Inst& I, Buf& b);
    void release_after_sync(Inst& I, Buf& b);
    void ensure(Inst& I, Buf& b, size_t bytes);
    void release_all(Inst& I);
    void setup(Inst& I);
    void build_units(Inst& I);
    void prepare_scratch(Inst& I);
    void ensure_table(Inst& I, uint64_t want_keys, bool force_rebuild);
    void finish(Inst& I, vfb_status st);
    void start_bs(Inst& I);
    void defer(Inst& I);
    bool on_capacity(Inst& I, const CapacityError& e);
    std::string node_string(Inst& I, int level, uint64_t row);
    void walk_cand(Inst& I, int level, uint64_t o, std::vector<const unsigned long long*>& need);
    void walk_node(Inst& I, int level, uint64_t row, std::vector<const unsigned long long*>& need);
    void prefetch_parents(const std::vector<FormulaRef>& refs);
    std::string cand_string(Inst& I, int level, uint64_t o, int* top_op);
    void fill(Inst& I, void* p, uint64_t words, uint64_t v);
    void upload(Inst& I, void* dst, const uint64_t* src, uint64_t words);
    void flush_ops();
    void purge_ops(const Inst& I);
    void time_begin(hipEvent_t* a, int kind);
    void time_end(hipEvent_t a, int kind, const std::vector<Inst*>& who);
    void collect_times();
    void epoch(std::vector<Inst*>& v);
    void store(std::vector<Inst*>& v);
    void prepare_or_check(Inst& I, std::vector<Inst*>& step);
    bool chain_admits(const Inst& I, uint64_t n_max) const;
    bool lds_chain_admits(const Inst& I) const;
    bool direct_admits(const Inst& I) const;
    bool probed(const Inst& I) const;
    uint64_t set_keys_for(const Inst& I, uint64_t n) const;
    uint64_t set_slots_for(uint64_t want_keys) const;
    bool grow_for_pass(Inst& I, uint64_t n);
    double set_margin() const { return tun_.set_margin_pct / 100.0; }
    uint64_t fresh_exact(const Inst& I) const;
    void chain(std::vector<Inst*>& v, int G, uint64_t n_max, bool lds = false);
    bool chain_adopt(Inst& I, const ChainOut& o);
};

// The process-wide engine (its arena holds all device memory of every solve)
// and the mutex that serialises its use.  engine_acquire() creates the engine,
// or re-creates it with a larger arena when want_bytes exceeds the current one.
std::mutex& engine_mutex();
vfb_tuning current_tuning();   // caller holds engine_mutex()
bool tuning_valid(const vfb_tuning& t);
Engine& engine_acquire(size_t want_bytes);
bool engine_pinned();   // true while a server holds requests in the engine

}  // namespace vfb
