// Host driver of the CS enumeration: the level scheduler of LTLI
// (LTL_BS.cu:1069-1532) rebuilt around whole-level epochs.
//
// For every active instance a step builds size level s:
//   1. candidate units in the reference's order (~, & splits, | splits, X, F,
//      G, U splits; LTL_BS.cu:1236-1488) and fused tasks over them;
//   2. k_expand over all tasks of all instances (one launch);
//   3. sync: on table / fresh-list overflow the set is rebuilt larger and the
//      epoch re-run (results never depend on capacity);
//   4. k_resolve, k_count, k_scan; sync: survivors, lowest matching ordinal;
//   5. found -> formula rebuilt from parent ordinals; else k_store writes the
//      level's survivors in ordinal order into a new level block.
// All device memory of a solve comes from one arena (a single hipMalloc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_dev.h"
#include "vfb_internal.h"

namespace vfb {

hipError_t launch_expand(const InstDesc* D, const Task* tasks, int ntasks, uint64_t total_waves, int grid,
                         hipStream_t st);
hipError_t launch_insert_atoms(const InstDesc* D, uint32_t inst, uint32_t n, hipStream_t st);
hipError_t launch_rehash(const Slot* old, uint64_t old_cap, Slot* tab, uint64_t mask, uint64_t keep_below,
                         unsigned int* fail, int grid, hipStream_t st);
hipError_t launch_fill(unsigned long long* p, uint64_t n, unsigned long long v, int grid, hipStream_t st);
hipError_t launch_resolve(const InstDesc* D, const Seg* segs, int nseg, uint64_t total, int grid, hipStream_t st);
hipError_t launch_count(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_blocks, int grid,
                        hipStream_t st);
hipError_t launch_scan(const InstDesc* D, const uint32_t* insts, int n, hipStream_t st);
hipError_t launch_store(const InstDesc* D, const Seg* segs, int nseg, uint64_t total_blocks, int grid,
                        hipStream_t st);

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

// ---------------------------------------------------------------------------
// device arena: one hipMalloc, first-fit free list with coalescing
// ---------------------------------------------------------------------------
class Arena {
  public:
    explicit Arena(size_t bytes) : size_(bytes) {
        CK(hipMalloc(&base_, bytes));
        free_[0] = bytes;
    }
    ~Arena() { (void)hipFree(base_); }
    void* alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                size_t off = it->first, sz = it->second;
                free_.erase(it);
                if (sz > bytes) free_[off + bytes] = sz - bytes;
                used_[off] = bytes;
                in_use_ += bytes;
                return (char*)base_ + off;
            }
        }
        return nullptr;
    }
    void release(void* p) {
        if (!p) return;
        size_t off = (char*)p - (char*)base_;
        auto it = used_.find(off);
        if (it == used_.end()) return;
        size_t sz = it->second;
        used_.erase(it);
        in_use_ -= sz;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + sz == nx->first) { sz += nx->second; nx = free_.erase(nx); }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) { pv->second += sz; return; }
        }
        free_[off] = sz;
    }
    size_t in_use() const { return in_use_; }
    size_t size() const { return size_; }

  private:
    void* base_ = nullptr;
    size_t size_;
    size_t in_use_ = 0;
    std::map<size_t, size_t> free_;
    std::map<size_t, size_t> used_;
};

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct CapacityError : std::runtime_error {
    explicit CapacityError(const std::string& s) : std::runtime_error(s) {}
};

// ---------------------------------------------------------------------------
// per-instance state
// ---------------------------------------------------------------------------
struct Inst {
    const vfb_instance* in = nullptr;
    vfb_result* res = nullptr;
    int max_size = 0;
    uint64_t budget = 0, used = 0;
    InstDesc h{};
    bool done = false;
    int level = 0;                   // last completed level
    uint64_t gbase = 0;              // global ordinal of the next level's first candidate
    std::vector<Buf> rows, pars;     // per level
    std::vector<uint32_t> nrows;
    std::vector<std::vector<Unit>> units;
    Buf table;
    uint64_t cap = 0, committed = 0;
    Buf fresh_ord, fresh_slot, winbits, blkcount, blkprefix;
    uint64_t fresh_cap = 0, nwords_cap = 0, nblk_cap = 0;
    double win_ratio = 1.0, fresh_ratio = 1.0;
    // current epoch
    uint64_t N = 0;
    std::vector<Task> tasks;
    Ctr ctr{};
    bool redo = false;
};

class Engine {
  public:
    explicit Engine(size_t arena_bytes);
    ~Engine();
    void run(std::vector<Inst*>& insts, const vfb_options& opt);
    size_t arena_size() const { return arena_->size(); }

  private:
    vfb_options opt_;
    hipStream_t st_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    int cus_ = 256;
    std::unique_ptr<Arena> arena_;
    InstDesc* d_desc_ = nullptr;
    size_t desc_cap_ = 0;
    Ctr* d_ctr_ = nullptr;
    size_t ctr_cap_ = 0;
    void* d_aux_ = nullptr;
    size_t aux_cap_ = 0;
    unsigned int* d_flag_ = nullptr;
    char* pin_ = nullptr;        // pinned staging for host->device uploads,
    size_t pin_cap_ = 0;         // recycled only at stream sync points
    size_t pin_used_ = 0;

    void sync();
    void h2d(void* dst, const void* src, size_t bytes);

    void dalloc(Inst& I, Buf& b, size_t bytes);
    void dfree(Inst& I, Buf& b);
    void ensure(Inst& I, Buf& b, size_t bytes);
    void* aux(size_t bytes);
    void setup(Inst& I, uint32_t idx);
    void prepare_level(Inst& I);
    void ensure_table(Inst& I, uint64_t want_keys, bool force_rebuild);
    void upload_descs(std::vector<Inst*>& v);
    void finish(Inst& I, vfb_status st);
    std::string node_string(Inst& I, int level, uint64_t row);
    std::string cand_string(Inst& I, int level, uint64_t o, int* top_op);
    void fill(void* p, uint64_t words, uint64_t v);
    float timed_begin();
    void timed_end(double* acc);
};

Engine::Engine(size_t arena_bytes) {
    vfb_options_default(&opt_);
    CK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    CK(hipEventCreate(&ev_[0]));
    CK(hipEventCreate(&ev_[1]));
    int dev = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    arena_ = std::make_unique<Arena>(arena_bytes);
    CK(hipMalloc(&d_flag_, sizeof(unsigned int)));
}

Engine::~Engine() {
    if (d_desc_) (void)hipFree(d_desc_);
    if (d_ctr_) (void)hipFree(d_ctr_);
    if (d_aux_) (void)hipFree(d_aux_);
    if (d_flag_) (void)hipFree(d_flag_);
    if (pin_) (void)hipHostFree(pin_);
    arena_.reset();
    if (ev_[0]) (void)hipEventDestroy(ev_[0]);
    if (ev_[1]) (void)hipEventDestroy(ev_[1]);
    if (st_) (void)hipStreamDestroy(st_);
}

void Engine::dalloc(Inst& I, Buf& b, size_t bytes) {
    size_t rb = (bytes + 255) & ~size_t(255);
    if (I.used + rb > I.budget) throw CapacityError("per-solve HBM budget exhausted");
    void* p = arena_->alloc(rb);
    if (!p) throw CapacityError("device arena exhausted");
    b.p = p;
    b.bytes = rb;
    I.used += rb;
}

void Engine::dfree(Inst& I, Buf& b) {
    if (b.p) {
        arena_->release(b.p);
        I.used -= b.bytes;
    }
    b = Buf{};
}

void Engine::ensure(Inst& I, Buf& b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return;
    dfree(I, b);
    dalloc(I, b, bytes);
}

void Engine::sync() {
    CK(hipStreamSynchronize(st_));
    pin_used_ = 0;
}

void Engine::h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    const size_t need = (bytes + 255) & ~size_t(255);
    if (pin_used_ + need > pin_cap_) {
        sync();
        if (need > pin_cap_) {
            if (pin_) CK(hipHostFree(pin_));
            pin_ = nullptr;
            pin_cap_ = std::max(need, std::max<size_t>(pin_cap_ * 2, 1 << 22));
            CK(hipHostMalloc((void**)&pin_, pin_cap_, hipHostMallocDefault));
        }
    }
    std::memcpy(pin_ + pin_used_, src, bytes);
    CK(hipMemcpyAsync(dst, pin_ + pin_used_, bytes, hipMemcpyHostToDevice, st_));
    pin_used_ += need;
}

void* Engine::aux(size_t bytes) {
    if (bytes > aux_cap_) {
        sync();
        if (d_aux_) CK(hipFree(d_aux_));
        aux_cap_ = std::max(bytes, aux_cap_ * 2);
        CK(hipMalloc(&d_aux_, aux_cap_));
    }
    return d_aux_;
}

void Engine::fill(void* p, uint64_t words, uint64_t v) {
    int grid = (int)std::min<uint64_t>((words + 255) / 256, (uint64_t)cus_ * 8);
    CK(launch_fill((unsigned long long*)p, words, v, std::max(grid, 1), st_));
}

float Engine::timed_begin() {
    if (opt_.timing) CK(hipEventRecord(ev_[0], st_));
    return 0.f;
}

void Engine::timed_end(double* acc) {
    if (!opt_.timing) return;
    CK(hipEventRecord(ev_[1], st_));
    CK(hipEventSynchronize(ev_[1]));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
    *acc += ms;
}

// key parameters per trace (LTL_BS.cu:30-110), computed once per instance
static void key_params(const vfb_instance& in, int relaxed_type, InstDesc& h) {
    const int T = in.T();
    int lenSum = 0;
    for (int t = 0; t < T; ++t) lenSum += in.lens[t];
    for (int t = 0; t < T; ++t) {
        h.lenmask[t] = len_mask(in.lens[t]);
        h.kmask[t] = 0;
        h.koff[t] = 0;
        h.kbits[t] = 0;
    }
    h.half = T / 2;
    h.stride = 1;
    if (lenSum <= 126) {
        h.key_kind = KEY_FIELD;
        int j = 0;
        for (int t = 0; t < T; ++t) {
            h.koff[t] = (int16_t)j;
            h.kbits[t] = (int16_t)in.lens[t];
            h.kmask[t] = len_mask(in.lens[t]);
            j += in.lens[t];
        }
    } else if (relaxed_type == 1) {
        h.key_kind = KEY_STRIDE;
        h.stride = lenSum / 126;
        int j = 0;
        for (int t = 0; t < T; ++t) {
            int n = (in.lens[t] - 1) / h.stride + 1;
            int use = std::max(0, std::min(n, 126 - j));
            h.koff[t] = (int16_t)std::min(j, 126);
            h.kbits[t] = (int16_t)use;
            j += n;
        }
    } else if (relaxed_type == 2) {
        h.key_kind = KEY_FIELD;
        int j = 0;
        for (int t = 0; t < T; ++t) {
            int mb = (126 * in.lens[t] + lenSum) / lenSum;
            int use = std::max(0, std::min(mb, 126 - j));
            h.koff[t] = (int16_t)std::min(j, 126);
            h.kbits[t] = (int16_t)use;
            h.kmask[t] = len_mask(use);
            j += mb;
        }
    } else {
        h.key_kind = KEY_MIX;
    }
}

void Engine::setup(Inst& I, uint32_t idx) {
    const vfb_instance& in = *I.in;
    const int A = in.A(), T = in.T();
    std::memset(&I.h, 0, sizeof(I.h));
    I.h.T = T;
    I.h.P = in.P;
    key_params(in, opt_.relaxed_type, I.h);
    I.rows.assign(kMaxS + 2, Buf{});
    I.pars.assign(kMaxS + 2, Buf{});
    I.nrows.assign(kMaxS + 2, 0);
    I.units.assign(kMaxS + 2, {});
    vfb_result& R = *I.res;
    // level 1: the atoms, all stored, checked in order (LTL_BS.cu:1127-1163)
    for (int a = 0; a < A; ++a) {
        R.candidates[1]++;
        bool ok = true;
        for (int t = 0; t < T && ok; ++t) {
            bool b = in.atoms[(size_t)a * T + t] & 1;
            if ((t < in.P) != b) ok = false;
        }
        if (ok) {
            R.min_size = 1;
            R.top_op = VFB_OP_ATOM;
            R.found_ordinal = (uint64_t)a;
            R.formula = strdup(in.names[a].c_str());
            R.level_winners[1] = (uint64_t)A;
            R.level_candidates[1] = (uint64_t)A;
            I.done = true;
            R.status = VFB_OK;
            return;
        }
        R.distinct[1]++;
    }
    R.level_winners[1] = (uint64_t)A;
    R.level_candidates[1] = (uint64_t)A;
    std::vector<uint64_t> col((size_t)A * T);
    for (int a = 0; a < A; ++a)
        for (int t = 0; t < T; ++t) col[(size_t)t * A + a] = in.atoms[(size_t)a * T + t] & len_mask(in.lens[t]);
    dalloc(I, I.rows[1], col.size() * 8);
    h2d(I.rows[1].p, col.data(), col.size() * 8);
    I.nrows[1] = (uint32_t)A;
    I.level = 1;
    I.gbase = (uint64_t)A;
    (void)idx;
}

void Engine::prepare_level(Inst& I) {
    const int s = I.level + 1;
    auto& U = I.units[s];
    U.clear();
    uint64_t base = 0;
    auto add = [&](int op, int ll, int lr) {
        Unit u{};
        u.op = op;
        u.lvl_l = ll;
        u.lvl_r = lr;
        u.nl = I.nrows[ll];
        u.nr = lr ? I.nrows[lr] : 1;
        u.base = base;
        uint64_t pairs = (uint64_t)u.nl * (lr ? u.nr : 1);
        u.count = op == VFB_OP_UNTIL ? 2 * pairs : pairs;
        base += u.count;
        U.push_back(u);
    };
    add(VFB_OP_NOT, s - 1, 0);
    for (int i = 1; 2 * i <= s - 1; ++i) add(VFB_OP_AND, i, s - 1 - i);
    for (int i = 1; 2 * i <= s - 1; ++i) add(VFB_OP_OR, i, s - 1 - i);
    add(VFB_OP_NEXT, s - 1, 0);
    add(VFB_OP_FINALLY, s - 1, 0);
    add(VFB_OP_GLOBALLY, s - 1, 0);
    for (int i = 1; 2 * i <= s - 1; ++i) add(VFB_OP_UNTIL, i, s - 1 - i);
    I.N = base;

    // fused tasks
    I.tasks.clear();
    const int nsplit = (s - 1) / 2;
    if (I.nrows[s - 1]) {
        Task t{};
        t.kind = 0;
        t.lvl_l = (uint8_t)(s - 1);
        t.nl = I.nrows[s - 1];
        t.nr = 1;
        t.nchunk = 1;
        t.base[0] = U[0].base;
        t.base[1] = U[1 + 2 * nsplit].base;
        t.base[2] = U[2 + 2 * nsplit].base;
        t.base[3] = U[3 + 2 * nsplit].base;
        I.tasks.push_back(t);
    }
    for (int i = 1; i <= nsplit; ++i) {
        const int j = s - 1 - i;
        if (!I.nrows[i] || !I.nrows[j]) continue;
        Task t{};
        t.kind = 1;
        t.lvl_l = (uint8_t)i;
        t.lvl_r = (uint8_t)j;
        t.nl = I.nrows[i];
        t.nr = I.nrows[j];
        t.nchunk = (t.nr + 63) / 64;
        t.base[0] = U[i].base;                   // & split i
        t.base[1] = U[nsplit + i].base;          // | split i
        t.base[2] = U[4 + 2 * nsplit + i - 1].base;  // U split i
        I.tasks.push_back(t);
    }

    // descriptor fields for this epoch
    InstDesc& h = I.h;
    h.level = s;
    h.nunits = (int)U.size();
    for (size_t k = 0; k < U.size(); ++k) h.units[k] = U[k];
    for (int l = 1; l < s; ++l) {
        h.lvl[l] = (const unsigned long long*)I.rows[l].p;
        h.nrows[l] = I.nrows[l];
    }
    h.epoch_base = I.gbase;
    h.ord_lo = 0;
    h.ord_hi = I.N;

    // scratch sized from the previous level's ratios, grown on overflow
    const uint64_t N = I.N;
    const uint64_t nwords = (N + 63) / 64;
    const uint64_t nblk = (nwords + kChunkWords - 1) / kChunkWords;
    uint64_t fcap = std::min<uint64_t>(N, std::max<uint64_t>(1 << 16, (uint64_t)(N * std::min(1.0, I.fresh_ratio * 1.5)) + 4096));
    if (fcap > I.fresh_cap) {
        ensure(I, I.fresh_ord, fcap * 8);
        ensure(I, I.fresh_slot, fcap * 4);
        I.fresh_cap = fcap;
    }
    if (nwords > I.nwords_cap) {
        ensure(I, I.winbits, std::max<uint64_t>(nwords, 1) * 8);
        I.nwords_cap = nwords;
    }
    if (nblk > I.nblk_cap) {
        ensure(I, I.blkcount, std::max<uint64_t>(nblk, 1) * 4);
        ensure(I, I.blkprefix, std::max<uint64_t>(nblk, 1) * 8);
        I.nblk_cap = nblk;
    }
    h.fresh_ord = (unsigned long long*)I.fresh_ord.p;
    h.fresh_slot = (uint32_t*)I.fresh_slot.p;
    h.fresh_cap = I.fresh_cap;
    h.winbits = (unsigned long long*)I.winbits.p;
    h.nwords = nwords;
    h.blkcount = (uint32_t*)I.blkcount.p;
    h.blkprefix = (unsigned long long*)I.blkprefix.p;
    h.nblk = nblk;
    fill(I.winbits.p, nwords, 0);

    const uint64_t est_new = std::min<uint64_t>(N, std::max<uint64_t>(1 << 16, (uint64_t)(N * std::min(1.0, I.win_ratio * 2.0))));
    ensure_table(I, I.committed + est_new, false);
}

// (re)builds the hash set so that it holds want_keys at load <= 1/2; keeps
// only keys of completed epochs (ord < epoch_base)
void Engine::ensure_table(Inst& I, uint64_t want_keys, bool force_rebuild) {
    uint64_t need = next_pow2(std::max<uint64_t>(4096, 2 * want_keys));
    if (need > (1ull << 32)) need = 1ull << 32;
    if (I.table.p && need <= I.cap && !force_rebuild) {
        I.h.claim_limit = I.cap * 7 / 10 > I.committed ? I.cap * 7 / 10 - I.committed : 0;
        return;
    }
    if (force_rebuild && need <= I.cap) need = I.cap;
    Buf nt;
    dalloc(I, nt, need * sizeof(Slot));
    fill(nt.p, need * 4, kEmpty);
    if (I.table.p) {
        CK(hipMemsetAsync(d_flag_, 0, sizeof(unsigned int), st_));
        int grid = (int)std::min<uint64_t>((I.cap + 255) / 256, (uint64_t)cus_ * 8);
        CK(launch_rehash((const Slot*)I.table.p, I.cap, (Slot*)nt.p, need - 1, I.h.epoch_base ? I.h.epoch_base : I.gbase,
                         d_flag_, grid, st_));
        unsigned int fail = 0;
        CK(hipMemcpyAsync(&fail, d_flag_, sizeof fail, hipMemcpyDeviceToHost, st_));
        sync();
        if (fail) throw CapacityError("hash set rebuild failed");
        dfree(I, I.table);
    }
    I.table = nt;
    I.cap = need;
    I.h.table = (Slot*)nt.p;
    I.h.table_mask = need - 1;
    I.h.claim_limit = need * 7 / 10 > I.committed ? need * 7 / 10 - I.committed : 0;
    I.res->table_slots = need;
}

void Engine::upload_descs(std::vector<Inst*>& v) {
    const size_t n = v.size();
    if (n > desc_cap_) {
        sync();
        if (d_desc_) CK(hipFree(d_desc_));
        CK(hipMalloc(&d_desc_, n * sizeof(InstDesc)));
        desc_cap_ = n;
    }
    if (n > ctr_cap_) {
        sync();
        if (d_ctr_) CK(hipFree(d_ctr_));
        CK(hipMalloc(&d_ctr_, n * sizeof(Ctr)));
        ctr_cap_ = n;
    }
    std::vector<InstDesc> hd(n);
    for (size_t k = 0; k < n; ++k) {
        v[k]->h.ctr = d_ctr_ + k;
        hd[k] = v[k]->h;
    }
    h2d(d_desc_, hd.data(), n * sizeof(InstDesc));
}

std::string Engine::node_string(Inst& I, int level, uint64_t row) {
    if (level == 1) return I.in->names[row];
    uint64_t o = 0;
    CK(hipMemcpy(&o, (const uint64_t*)I.pars[level].p + row, 8, hipMemcpyDeviceToHost));
    return cand_string(I, level, o, nullptr);
}

// formula text of candidate o of level `level` (ltl_nocache.cu:289-317)
std::string Engine::cand_string(Inst& I, int level, uint64_t o, int* top_op) {
    static const char opch[7] = {'~', '&', '|', 'X', 'F', 'G', 'U'};
    const auto& U = I.units[level];
    size_t k = 0;
    for (size_t q = 0; q < U.size(); ++q)
        if (U[q].count && U[q].base <= o) k = q;
    const Unit& u = U[k];
    if (top_op) *top_op = u.op;
    const uint64_t x = o - u.base;
    if (u.op == VFB_OP_NOT || u.op == VFB_OP_NEXT || u.op == VFB_OP_FINALLY || u.op == VFB_OP_GLOBALLY)
        return std::string(1, opch[u.op]) + "(" + node_string(I, u.lvl_l, x) + ")";
    uint64_t p = u.op == VFB_OP_UNTIL ? x >> 1 : x;
    uint64_t l = p / u.nr, r = p % u.nr;
    std::string a = node_string(I, u.lvl_l, l), b = node_string(I, u.lvl_r, r);
    if (u.op == VFB_OP_UNTIL && (x & 1)) std::swap(a, b);
    return "(" + a + ") " + opch[u.op] + " (" + b + ")";
}

void Engine::finish(Inst& I, vfb_status st) {
    I.done = true;
    I.res->status = st;
    if (!I.res->formula) I.res->formula = strdup("Not found !");
    for (auto* v : {&I.rows, &I.pars})
        for (auto& b : *v) dfree(I, b);
    dfree(I, I.table);
    dfree(I, I.fresh_ord);
    dfree(I, I.fresh_slot);
    dfree(I, I.winbits);
    dfree(I, I.blkcount);
    dfree(I, I.blkprefix);
}

void Engine::run(std::vector<Inst*>& all, const vfb_options& opt) {
    opt_ = opt;
    // setup: atoms, level-1 check, initial tables
    std::vector<Inst*> act;
    for (auto* I : all) {
        try {
            setup(*I, 0);
            if (I->done) continue;
            ensure_table(*I, (uint64_t)I->in->A() * 4, false);
            act.push_back(I);
        } catch (CapacityError&) {
            finish(*I, VFB_CAPACITY);
        }
    }
    if (!act.empty()) {
        for (auto* I : act) { I->h.epoch_base = 0; I->h.level = 1; }
        upload_descs(act);
        CK(hipMemsetAsync(d_ctr_, 0, act.size() * sizeof(Ctr), st_));
        for (size_t k = 0; k < act.size(); ++k) {
            act[k]->h.lvl[1] = (const unsigned long long*)act[k]->rows[1].p;
            act[k]->h.nrows[1] = act[k]->nrows[1];
        }
        upload_descs(act);
        for (size_t k = 0; k < act.size(); ++k)
            CK(launch_insert_atoms(d_desc_, (uint32_t)k, act[k]->nrows[1], st_));
        std::vector<Ctr> hc(act.size());
        CK(hipMemcpyAsync(hc.data(), d_ctr_, hc.size() * sizeof(Ctr), hipMemcpyDeviceToHost, st_));
        sync();
        for (size_t k = 0; k < act.size(); ++k) {
            if (hc[k].overflow) throw DeviceError("atom insertion overflow");
            act[k]->committed = hc[k].claims;
        }
    }

    while (!act.empty()) {
        // ---- 1. units, tasks, scratch ----
        std::vector<Inst*> step;
        for (auto* I : act) {
            if (I->done) continue;
            try {
                prepare_level(*I);
                step.push_back(I);
            } catch (CapacityError&) {
                finish(*I, VFB_CAPACITY);
            }
        }
        if (step.empty()) break;

        // ---- 2-3. expand with retry on overflow ----
        std::vector<Inst*> pend = step;
        for (int attempt = 0; !pend.empty(); ++attempt) {
            if (attempt > 12) throw DeviceError("expand did not converge");
            upload_descs(pend);
            std::vector<Ctr> init(pend.size());
            for (auto& c : init) { c = Ctr{}; c.found = kEmpty; }
            h2d(d_ctr_, init.data(), init.size() * sizeof(Ctr));
            std::vector<Task> tasks;
            uint64_t waves = 0;
            for (size_t k = 0; k < pend.size(); ++k)
                for (Task t : pend[k]->tasks) {
                    t.inst = (uint32_t)k;
                    t.wave_begin = waves;
                    waves += t.kind == 0 ? (t.nl + 63) / 64 : (uint64_t)t.nl * t.nchunk;
                    tasks.push_back(t);
                }
            if (waves) {
                Task* d_tasks = (Task*)aux(tasks.size() * sizeof(Task));
                h2d(d_tasks, tasks.data(), tasks.size() * sizeof(Task));
                int grid = (int)std::min<uint64_t>((waves + 3) / 4, (uint64_t)cus_ * 8);
                timed_begin();
                CK(launch_expand(d_desc_, d_tasks, (int)tasks.size(), waves, grid, st_));
                for (auto* I : pend) I->res->expand_launches++;
                if (opt_.timing) {
                    double ms = 0;
                    timed_end(&ms);
                    for (auto* I : pend) I->res->ms_expand += ms;
                }
            }
            std::vector<Ctr> hc(pend.size());
            CK(hipMemcpyAsync(hc.data(), d_ctr_, hc.size() * sizeof(Ctr), hipMemcpyDeviceToHost, st_));
            sync();
            std::vector<Inst*> again;
            for (size_t k = 0; k < pend.size(); ++k) {
                Inst& I = *pend[k];
                I.ctr = hc[k];
                if (!hc[k].overflow) continue;
                try {
                    I.res->retries++;
                    if (hc[k].overflow & 2) {
                        uint64_t fcap = std::min<uint64_t>(I.N, std::max<uint64_t>(I.fresh_cap * 4, hc[k].fresh_count + 4096));
                        ensure(I, I.fresh_ord, fcap * 8);
                        ensure(I, I.fresh_slot, fcap * 4);
                        I.fresh_cap = fcap;
                        I.h.fresh_ord = (unsigned long long*)I.fresh_ord.p;
                        I.h.fresh_slot = (uint32_t*)I.fresh_slot.p;
                        I.h.fresh_cap = fcap;
                    }
                    uint64_t want = (hc[k].overflow & 5) ? (I.cap / 2) * 4 : I.cap / 2;
                    ensure_table(I, want, true);
                    again.push_back(&I);
                } catch (CapacityError&) {
                    finish(I, VFB_CAPACITY);
                }
            }
            pend.swap(again);
        }
        std::vector<Inst*> live;
        for (auto* I : step)
            if (!I->done) live.push_back(I);
        if (live.empty()) break;
        upload_descs(live);
        {
            // restore the counters of the successful expand (upload reset nothing,
            // but instances re-run individually had their own slot index)
            std::vector<Ctr> hc(live.size());
            for (size_t k = 0; k < live.size(); ++k) hc[k] = live[k]->ctr;
            h2d(d_ctr_, hc.data(), hc.size() * sizeof(Ctr));
        }

        // ---- 4. resolve, count, scan ----
        std::vector<Seg> fsegs, bsegs;
        uint64_t ftot = 0, btot = 0;
        std::vector<uint32_t> ids;
        for (size_t k = 0; k < live.size(); ++k) {
            Inst& I = *live[k];
            if (I.ctr.fresh_count) { fsegs.push_back(Seg{(uint32_t)k, 0, ftot}); ftot += I.ctr.fresh_count; }
            if (I.h.nblk) { bsegs.push_back(Seg{(uint32_t)k, 0, btot}); btot += I.h.nblk; }
            ids.push_back((uint32_t)k);
        }
        {
            size_t off1 = fsegs.size() * sizeof(Seg);
            size_t off2 = off1 + bsegs.size() * sizeof(Seg);
            char* a = (char*)aux(off2 + ids.size() * 4 + 64);
            if (!fsegs.empty()) h2d(a, fsegs.data(), off1);
            if (!bsegs.empty()) h2d(a + off1, bsegs.data(), off2 - off1);
            h2d(a + off2, ids.data(), ids.size() * 4);
            timed_begin();
            if (ftot) {
                int grid = (int)std::min<uint64_t>((ftot + 255) / 256, (uint64_t)cus_ * 16);
                CK(launch_resolve(d_desc_, (const Seg*)a, (int)fsegs.size(), ftot, grid, st_));
            }
            if (btot) {
                int grid = (int)std::min<uint64_t>((btot + 3) / 4, (uint64_t)cus_ * 8);
                CK(launch_count(d_desc_, (const Seg*)(a + off1), (int)bsegs.size(), btot, grid, st_));
            }
            CK(launch_scan(d_desc_, (const uint32_t*)(a + off2), (int)ids.size(), st_));
            if (opt_.timing) {
                double ms = 0;
                timed_end(&ms);
                for (auto* I : live) I->res->ms_resolve += ms;
            }
            std::vector<Ctr> hc(live.size());
            CK(hipMemcpyAsync(hc.data(), d_ctr_, hc.size() * sizeof(Ctr), hipMemcpyDeviceToHost, st_));
            sync();
            for (size_t k = 0; k < live.size(); ++k) live[k]->ctr = hc[k];
        }

        // ---- 5. found / store ----
        std::vector<Inst*> store;
        for (auto* Ip : live) {
            Inst& I = *Ip;
            vfb_result& R = *I.res;
            const int s = I.level + 1;
            R.level_candidates[s] = I.N;
            R.level_winners[s] = I.ctr.winners;
            R.fresh_total += I.ctr.fresh_count;
            for (const Unit& u : I.units[s]) {
                const uint64_t a = (u.op == VFB_OP_AND || u.op == VFB_OP_OR) ? 2 : 1;
                R.bytes_expand += u.count * (a * 8 * (uint64_t)I.h.T + 16);
            }
            if (opt_.verbose)
                std::fprintf(stderr, "Cost %-2d | candidates: %-14llu | fresh: %-12llu | distinct: %-12llu | table: %llu\n", s,
                             (unsigned long long)I.N, (unsigned long long)I.ctr.fresh_count,
                             (unsigned long long)I.ctr.winners, (unsigned long long)I.cap);
            if (I.ctr.found != kEmpty) {
                R.candidates[s] = I.ctr.found + 1;
                R.distinct[s] = I.ctr.found_rank;
                R.min_size = s;
                R.found_ordinal = I.ctr.found;
                std::string f = cand_string(I, s, I.ctr.found, &R.top_op);
                R.formula = strdup(f.c_str());
                finish(I, VFB_OK);
                continue;
            }
            R.candidates[s] = I.N;
            R.distinct[s] = I.ctr.winners;
            R.levels_done = s;
            R.bytes_store += I.ctr.winners * (8 * (uint64_t)I.h.T + 24);
            if (s >= I.max_size) {
                finish(I, VFB_NOT_FOUND);
                continue;
            }
            try {
                const uint64_t n = I.ctr.winners;
                if (n > 0xffffffffull) throw CapacityError("level too large");
                dalloc(I, I.rows[s], std::max<uint64_t>(n, 1) * I.h.T * 8);
                dalloc(I, I.pars[s], std::max<uint64_t>(n, 1) * 8);
                I.nrows[s] = (uint32_t)n;
                I.h.out_rows = (unsigned long long*)I.rows[s].p;
                I.h.out_par = (unsigned long long*)I.pars[s].p;
                I.h.out_n = n;
                store.push_back(&I);
            } catch (CapacityError&) {
                finish(I, VFB_CAPACITY);
            }
        }
        if (!store.empty()) {
            upload_descs(store);
            std::vector<Seg> segs;
            uint64_t tot = 0;
            for (size_t k = 0; k < store.size(); ++k)
                if (store[k]->nrows[store[k]->level + 1] && store[k]->h.nblk) {
                    segs.push_back(Seg{(uint32_t)k, 0, tot});
                    tot += store[k]->h.nblk * (uint64_t)((store[k]->h.T + kStoreTraces - 1) / kStoreTraces);
                }
            if (tot) {
                Seg* d = (Seg*)aux(segs.size() * sizeof(Seg));
                h2d(d, segs.data(), segs.size() * sizeof(Seg));
                int grid = (int)std::min<uint64_t>((tot + 3) / 4, (uint64_t)cus_ * 8);
                timed_begin();
                CK(launch_store(d_desc_, d, (int)segs.size(), tot, grid, st_));
                if (opt_.timing) {
                    double ms = 0;
                    timed_end(&ms);
                    for (auto* I : store) I->res->ms_store += ms;
                }
            }
            for (auto* I : store) {
                const uint64_t N = std::max<uint64_t>(I->N, 1);
                I->win_ratio = (double)I->ctr.winners / N;
                I->fresh_ratio = (double)I->ctr.fresh_count / N;
                I->committed += I->ctr.winners;
                I->gbase += I->N;
                I->level += 1;
            }
        }
        std::vector<Inst*> nxt;
        for (auto* I : act)
            if (!I->done) nxt.push_back(I);
        act.swap(nxt);
    }
    sync();
}

static std::mutex g_mu;
static std::unique_ptr<Engine> g_engine;

static vfb_status solve_many(const vfb_instance* const* insts, size_t n, const vfb_options* opt_in, vfb_result* out) {
    vfb_options opt;
    vfb_options_default(&opt);
    if (opt_in) opt = *opt_in;
    if (opt.relaxed_type < 1 || opt.relaxed_type > 3) opt.relaxed_type = 3;
    for (size_t k = 0; k < n; ++k) {
        std::memset(&out[k], 0, sizeof(vfb_result));
        out[k].status = VFB_NOT_FOUND;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Inst> st(n);
    std::vector<Inst*> run;
    for (size_t k = 0; k < n; ++k) {
        std::string why;
        if (!insts[k] || validate(*insts[k], &why) != VFB_OK) {
            out[k].status = VFB_INVALID_INPUT;
            out[k].formula = strdup("see_the_error");
            continue;
        }
        int ms = opt.max_size;
        if (ms < 1 || ms > VFB_MAX_SIZE) {
            out[k].status = VFB_INVALID_INPUT;
            out[k].formula = strdup("see_the_error");
            continue;
        }
        st[k].in = insts[k];
        st[k].res = &out[k];
        st[k].max_size = ms;
        run.push_back(&st[k]);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    try {
        size_t fr = 0, tot = 0;
        CK(hipMemGetInfo(&fr, &tot));
        // one arena per process, sized once (85% of free HBM, or the budget
        // when that is larger than what the current arena holds)
        size_t want = opt.hbm_budget ? (size_t)opt.hbm_budget : 0;
        if (!g_engine || (want && want > g_engine->arena_size())) {
            g_engine.reset();
            CK(hipMemGetInfo(&fr, &tot));
            size_t bytes = want ? std::min(want, (size_t)(fr * 0.95)) : (size_t)(fr * 0.85);
            g_engine = std::make_unique<Engine>(bytes);
        }
        uint64_t per = opt.hbm_budget ? opt.hbm_budget : g_engine->arena_size();
        for (auto* I : run) I->budget = per;
        g_engine->run(run, opt);
    } catch (CapacityError& e) {
        for (auto* I : run)
            if (!I->done) { I->res->status = VFB_CAPACITY; if (!I->res->formula) I->res->formula = strdup("Not found !"); }
    } catch (std::exception& e) {
        std::fprintf(stderr, "vfb: device error: %s\n", e.what());
        for (auto* I : run) {
            if (I->res->status != VFB_OK || !I->done) {
                I->res->status = VFB_DEVICE_ERROR;
                if (!I->res->formula) I->res->formula = strdup("see_the_error");
            }
        }
        g_engine.reset();
        return VFB_DEVICE_ERROR;
    }
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t k = 0; k < n; ++k) {
        out[k].ms_total = ms;
        if (!out[k].formula) out[k].formula = strdup("Not found !");
    }
    return n == 1 ? out[0].status : VFB_OK;
}

}  // namespace vfb

using namespace vfb;

extern "C" vfb_status vfb_solve(const vfb_instance* inst, const vfb_options* opt, vfb_result* out) {
    if (!out) return VFB_INVALID_INPUT;
    const vfb_instance* one[1] = {inst};
    return solve_many(one, 1, opt, out);
}

extern "C" vfb_status vfb_solve_batch(const vfb_instance* const* insts, size_t n, const vfb_options* opt,
                                      vfb_result* out) {
    if (!insts || !out) return VFB_INVALID_INPUT;
    if (n == 0) return VFB_OK;
    return solve_many(insts, n, opt, out);
}

extern "C" vfb_status vfb_device_info(char* name, size_t cap, int* cus) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return VFB_DEVICE_ERROR;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return VFB_DEVICE_ERROR;
    if (name && cap) std::snprintf(name, cap, "%s (%s)", p.name, p.gcnArchName);
    if (cus) *cus = p.multiProcessorCount;
    return VFB_OK;
}
