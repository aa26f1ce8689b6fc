// Host driver of the CS enumeration: the level scheduler of LTLI
// (LTL_BS.cu:1069-1532) rebuilt around whole-level epochs.
//
// A step builds one size level s for every active instance:
//   1. candidate units in the reference's order (~, & splits, | splits, X, F,
//      G, U splits; LTL_BS.cu:1236-1488) and fused tasks over them;
//   2. one packed upload (descriptors, counters, tasks), then k_expand over all
//      tasks of all instances, k_fresh_prefix, k_resolve, k_count and k_scan,
//      all enqueued back to back; ONE sync;
//   3. instances whose hash set or fresh list overflowed get a larger set and
//      re-run the epoch (results never depend on capacity);
//   4. found -> formula rebuilt from parent ordinals; else k_store writes the
//      level's survivors, in ordinal order, into a new level block (async).
// Overflow (the reference's onTheFly / lastRound, LTL_BS.cu:331-349, 516-522,
// 1223-1233, 1490-1491): when the stored rows would pass max_rows or the HBM
// budget, the level is stored up to that point in ordinal order and the next
// level is only checked against the target, then the search stops.
// All device memory of a solve comes from one process-wide arena.
#include "engine.h"

#include <algorithm>
#include <unordered_set>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace vfb {

// process-wide kernel counters (vfb_kernel_stats_get); updated under the
// library's lock, like everything the engines do
static vfb_kernel_stats g_kstats{};

static void drop_live(std::vector<Inst*>& live, Inst* I) {
    live.erase(std::remove(live.begin(), live.end(), I), live.end());
}

Engine::Engine(size_t arena_bytes) {
    vfb_options_default(&opt_);
    CK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    int dev = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    arena_ = std::make_unique<Arena>(arena_bytes);
    // probe-limit failures of rehash / atom insertion: a device word the
    // kernels set, copied back at the next stream sync after such a launch
    CK(hipMalloc((void**)&d_flag_, 256));
    CK(hipMemset(d_flag_, 0, sizeof(unsigned int)));
    CK(hipHostMalloc((void**)&h_flag_, sizeof(unsigned int), hipHostMallocDefault));
    *h_flag_ = 0;
    tun_ = current_tuning();
    chain_resident_ = chain_blocks_per_cu() * cus_;
}

// Host side of a fixed-memory engine: static (pageable) staging for uploads,
// counters and the failure flag.  All fixed-memory engines share them; calls
// are serialised by the library's lock and every call ends with a stream
// sync, after which nothing in them is in flight.
static constexpr size_t kFixedStage = 4u << 20;
static constexpr size_t kFixedStep = 1u << 20;      // descriptor area carved from the scratch
static constexpr size_t kFixedCtrs = 64;
alignas(256) static char g_fixed_stage[kFixedStage];
static Ctr g_fixed_ctr[kFixedCtrs];
static unsigned int g_fixed_flag;

Engine::Engine(const Fixed& m) {
    fixed_ = true;
    vfb_options_default(&opt_);
    CK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    int dev = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev));
    // [0, 256): failure flag; [256, 256 + kFixedStep): descriptors; the rest: arena
    const size_t head = 256 + kFixedStep;
    if (m.bytes < head + (16u << 20)) throw DeviceError("fixed engine memory too small");
    char* base = (char*)m.dev;
    d_flag_ = (unsigned int*)base;
    d_step_ = base + 256;
    step_cap_ = kFixedStep;
    arena_ = std::make_unique<Arena>(base + head, m.bytes - head);
    h_flag_ = &g_fixed_flag;
    *h_flag_ = 0;
    h_ctr_ = g_fixed_ctr;
    h_ctr_cap_ = kFixedCtrs;
    pin_ = g_fixed_stage;
    pin_cap_ = kFixedStage;
    // the scratch keeps the previous process state's contents: clear the flag
    CK(launch_fill((unsigned long long*)d_flag_, 1, 0, 1, st_));
    CK(hipStreamSynchronize(st_));
    tun_ = current_tuning();
}

Engine::~Engine() {
    (void)hipStreamSynchronize(st_);
    if (fixed_) {   // nothing of it was allocated but the stream and events
        for (auto e : ev_pool_) (void)hipEventDestroy(e);
        for (auto& t : timed_) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
        arena_.reset();
        if (st_) (void)hipStreamDestroy(st_);
        return;
    }
    if (d_step_) (void)hipFree(d_step_);
    if (d_ops_) (void)hipFree(d_ops_);
    if (h_flag_) (void)hipHostFree(h_flag_);
    if (h_ctr_) (void)hipHostFree(h_ctr_);
    if (d_flag_) (void)hipFree(d_flag_);
    if (pin_) (void)hipHostFree(pin_);
    if (d_chain_ws_) (void)hipFree(d_chain_ws_);
    if (d_gather_) (void)hipFree(d_gather_);
    if (h_gather_) (void)hipHostFree(h_gather_);
    if (h_chain_out_) (void)hipHostFree(h_chain_out_);
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    for (auto& t : timed_) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    arena_.reset();
    if (st_) (void)hipStreamDestroy(st_);
}

void Engine::sync() {
    if (flag_dirty_) CK(hipMemcpyAsync(h_flag_, d_flag_, sizeof(unsigned int), hipMemcpyDeviceToHost, st_));
    flag_dirty_ = false;
    CK(hipStreamSynchronize(st_));
    pin_used_ = 0;
    collect_times();
    for (auto& pb : pending_free_) arena_->release(pb);
    pending_free_.clear();
    if (*h_flag_) throw DeviceError("hash set insertion exceeded the probe limit");
}

// a buffer still read by enqueued kernels goes back to the arena at the next
// sync; its bytes stop counting against the instance now
void Engine::release_after_sync(Inst& I, Buf& b) {
    if (b.p && !b.borrowed) {
        pending_free_.push_back(b.p);
        I.used -= b.bytes;
    }
    b = Buf{};
}

void Engine::h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (fixed_) {
        // static staging: copies larger than it go in pieces, each after
        // the stream has drained the staging
        const char* s = (const char*)src;
        char* d = (char*)dst;
        while (bytes) {
            if (pin_used_ + 256 > pin_cap_) sync();
            const size_t n = std::min(bytes, pin_cap_ - pin_used_);
            std::memcpy(pin_ + pin_used_, s, n);
            CK(hipMemcpyAsync(d, pin_ + pin_used_, n, hipMemcpyHostToDevice, st_));
            pin_used_ += align256(n);
            s += n;
            d += n;
            bytes -= n;
        }
        return;
    }
    const size_t need = align256(bytes);
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

// Device scratch for one phase's packed upload.  Phases run in stream order,
// so a later phase may overwrite it; growing it waits for the stream first.
char* Engine::step_buf(size_t bytes) {
    if (bytes > step_cap_) {
        if (fixed_) throw ScratchExhausted("descriptor area of the fixed scratch");
        sync();
        if (d_step_) CK(hipFree(d_step_));
        step_cap_ = std::max(bytes, step_cap_ * 2);
        CK(hipMalloc((void**)&d_step_, step_cap_));
    }
    return d_step_;
}

void Engine::time_begin(hipEvent_t* a, int kind) {
    *a = nullptr;
    // kinds: 0 expand, 1 resolve + scan, 2 store, 3 probe, 4 level chain
    if (!opt_.timing || (opt_.timing == 2 && kind != 0 && kind != 4)) return;
    if (ev_pool_.empty()) {
        hipEvent_t e;
        CK(hipEventCreate(&e));
        ev_pool_.push_back(e);
    }
    *a = ev_pool_.back();
    ev_pool_.pop_back();
    CK(hipEventRecord(*a, st_));
}

void Engine::time_end(hipEvent_t a, int kind, const std::vector<Inst*>& who) {
    if (!a) return;
    hipEvent_t b;
    if (ev_pool_.empty()) CK(hipEventCreate(&b));
    else { b = ev_pool_.back(); ev_pool_.pop_back(); }
    CK(hipEventRecord(b, st_));
    timed_.push_back(Timed{a, b, kind, who});
}

// reads finished event pairs (called after a stream sync, so no extra waits)
void Engine::collect_times() {
    for (auto& t : timed_) {
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, t.a, t.b));
        (t.kind == 0   ? g_kstats.ms_expand
         : t.kind == 1 ? g_kstats.ms_resolve
         : t.kind == 2 ? g_kstats.ms_store
         : t.kind == 3 ? g_kstats.ms_probe
                       : g_kstats.ms_chain) += ms;
        for (Inst* I : t.who) {
            if (t.kind == 4) break;   // the chain has no per-result time
            double* acc = t.kind == 0   ? &I->res->ms_expand
                          : t.kind == 1 ? &I->res->ms_resolve
                          : t.kind == 2 ? &I->res->ms_store
                                        : &I->res->ms_probe;
            *acc += ms;
        }
        ev_pool_.push_back(t.a);
        ev_pool_.push_back(t.b);
    }
    timed_.clear();
}

void Engine::dalloc(Inst& I, Buf& b, size_t bytes) {
    size_t rb = std::max<size_t>(align256(bytes), 256);
    if (I.used + rb > I.budget) throw CapacityError("per-solve HBM budget exhausted", false);
    void* p = arena_->alloc(rb);
    if (!p && fixed_) throw ScratchExhausted("device scratch");
    while (!p && batch_) {
        // Shared arena full: evict the youngest other instance that holds
        // memory, so the oldest always progresses as it would alone.  The
        // evicted one restarts later with the same result.
        Inst* victim = nullptr;
        bool older_holds = false;
        for (Inst* v : live_) {
            if (v == &I || v->done || v->used == 0) continue;
            if (v->prio < I.prio) older_holds = true;
            else if (!victim || v->prio > victim->prio) victim = v;
        }
        if (!victim) {
            if (older_holds) throw CapacityError("device arena exhausted", true);
            // buffers released since the last sync may free enough
            if (!pending_free_.empty()) {
                sync();
                p = arena_->alloc(rb);
                continue;
            }
            // solved again alone, as the plain solve would be (its arena then
            // holds none of the others' holes): by the batch after the others,
            // by the server with nothing else admitted.  Already alone: it
            // goes on under the plain limit
            if (defer_alone_) {
                I.need_alone = true;
                throw CapacityError("device arena exhausted", true);
            }
            break;
        }
        sync();
        defer(*victim);
        p = arena_->alloc(rb);
    }
    if (!p) throw CapacityError("device arena exhausted", false);
    b.p = p;
    b.bytes = rb;
    I.used += rb;
}

void Engine::dfree(Inst& I, Buf& b) {
    if (b.p && !b.borrowed) {
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

void Engine::release_all(Inst& I) {
    purge_ops(I);
    for (auto& b : I.rows) dfree(I, b);
    dfree(I, I.table);
    dfree(I, I.fresh_ord);
    dfree(I, I.fresh_slot);
    dfree(I, I.winbits);
    dfree(I, I.winbyte);
    dfree(I, I.blkcount);
    dfree(I, I.blkprefix);
    dfree(I, I.cols);
    for (auto& b : I.slabs) dfree(I, b);
    I.slabs.clear();
    I.fresh_cap = I.nwords_cap = I.nblk_cap = I.lost_words = 0;
    I.cap = 0;
}

void Engine::fill(Inst& I, void* p, uint64_t words, uint64_t v) {
    if (batch_) {
        fills_.push_back(FillOp{&I, (unsigned long long*)p, SIZE_MAX, words, v});
        return;
    }
    int grid = (int)std::min<uint64_t>((words + 255) / 256, (uint64_t)cus_ * 8);
    CK(launch_fill((unsigned long long*)p, words, v, std::max(grid, 1), st_));
}

void Engine::upload(Inst& I, void* dst, const uint64_t* src, uint64_t words) {
    if (!batch_) {
        h2d(dst, src, words * 8);
        return;
    }
    fills_.push_back(FillOp{&I, (unsigned long long*)dst, staged_.size(), words, 0});
    staged_.insert(staged_.end(), src, src + words);
}

// drops the queued operations of an instance whose buffers are released
void Engine::purge_ops(const Inst& I) {
    fills_.erase(std::remove_if(fills_.begin(), fills_.end(), [&](const FillOp& f) { return f.I == &I; }),
                 fills_.end());
    rehashes_.erase(std::remove_if(rehashes_.begin(), rehashes_.end(), [&](const RehashOp& r) { return r.I == &I; }),
                    rehashes_.end());
    // sources of dropped rehashes: only kernels already enqueued may read them
    for (auto& r : rehash_src_)
        if (r.I == &I) pending_free_.push_back(r.p);
    rehash_src_.erase(std::remove_if(rehash_src_.begin(), rehash_src_.end(), [&](const RehashSrc& r) { return r.I == &I; }),
                      rehash_src_.end());
}

// one upload of the segment tables (and staged data), one fill / copy launch,
// one rehash launch; fills first, since rehash targets are fresh tables
void Engine::flush_ops() {
    if (fills_.empty() && rehashes_.empty()) return;
    const size_t nf = fills_.size(), nr = rehashes_.size();
    const size_t o_rs = align256(nf * sizeof(OpSeg));
    const size_t o_st = o_rs + align256(nr * sizeof(RehashSeg));
    const size_t bytes = o_st + staged_.size() * 8;
    if (bytes > ops_cap_) {
        sync();
        if (d_ops_) CK(hipFree(d_ops_));
        ops_cap_ = std::max(bytes, ops_cap_ * 2);
        CK(hipMalloc((void**)&d_ops_, ops_cap_));
    }
    std::vector<char> host(bytes, 0);
    OpSeg* fs = (OpSeg*)host.data();
    uint64_t ftot = 0;
    for (size_t k = 0; k < nf; ++k) {
        const FillOp& f = fills_[k];
        const unsigned long long* src =
            f.staged == SIZE_MAX ? nullptr : (const unsigned long long*)(d_ops_ + o_st) + f.staged;
        fs[k] = OpSeg{f.dst, src, f.words, f.value, ftot};
        ftot += (f.words + 7) / 8;   // k_fill_multi takes 8 words per thread
    }
    RehashSeg* rs = (RehashSeg*)(host.data() + o_rs);
    uint64_t rtot = 0;
    for (size_t k = 0; k < nr; ++k) {
        const RehashOp& r = rehashes_[k];
        rs[k] = RehashSeg{r.old, r.tab, r.mask, r.keep, rtot};
        rtot += r.old_cap;
    }
    if (!staged_.empty()) std::memcpy(host.data() + o_st, staged_.data(), staged_.size() * 8);
    h2d(d_ops_, host.data(), bytes);
    if (ftot) {
        const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((ftot + 255) / 256, (uint64_t)cus_ * 8));
        CK(launch_fill_multi((const OpSeg*)d_ops_, (int)nf, ftot, grid, st_));
    }
    if (rtot) {
        const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((rtot + 255) / 256, (uint64_t)cus_ * 8));
        CK(launch_rehash_multi((const RehashSeg*)(d_ops_ + o_rs), (int)nr, rtot, d_flag_, grid, st_));
        flag_dirty_ = true;
    }
    // the rehash sources go back to the arena once the launch above has run
    for (auto& r : rehash_src_) pending_free_.push_back(r.p);
    rehash_src_.clear();
    fills_.clear();
    rehashes_.clear();
    staged_.clear();
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
    // positive (0..P-1) and negative traces alternating, the longer list's
    // rest at the end
    for (int i = 0, p = 0, n = in.P; i < T; ++i) {
        const bool pos = p < in.P && (n >= T || (i & 1) == 0);
        h.check_order[i] = (uint8_t)(pos ? p++ : n++);
    }
    h.half = T / 2;
    h.stride = 1;
    if (lenSum <= 126) {                         // exact concatenation, :94-106
        h.key_kind = KEY_FIELD;
        int j = 0;
        for (int t = 0; t < T; ++t) {
            h.koff[t] = (int16_t)j;
            h.kbits[t] = (int16_t)in.lens[t];
            h.kmask[t] = len_mask(in.lens[t]);
            j += in.lens[t];
        }
    } else if (relaxed_type == 1) {              // strided samples, :45-59
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
    } else if (relaxed_type == 2) {              // leading bits per trace, :61-76
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
    } else {                                     // hashed, :78-88
        h.key_kind = KEY_MIX;
    }
}

// Key and mask parameters per column for rows whose traces span several
// words: the same key as key_params (LTL_BS.cu:30-110), each trace's field or
// samples split at its word boundaries; type 3 mixes each word on its own.
static void col_params(const vfb_instance& in, int relaxed_type, const InstDesc& h, std::vector<ColInfo>* out) {
    const int T = in.T();
    int lenSum = 0;
    for (int t = 0; t < T; ++t) lenSum += in.lens[t];
    out->clear();
    int j = 0;   // first key bit of the trace
    for (int t = 0; t < T; ++t) {
        const int len = in.lens[t], nw = trace_words(len);
        int field = 0, nsamp = 0, off = std::min(j, 126);
        if (h.key_kind == KEY_FIELD) {
            const int mb = lenSum <= 126 ? len : (126 * len + lenSum) / lenSum;
            field = lenSum <= 126 ? len : std::max(0, std::min(mb, 126 - j));
            j += mb;
        } else if (h.key_kind == KEY_STRIDE) {
            const int n = (len - 1) / h.stride + 1;
            nsamp = std::max(0, std::min(n, 126 - j));
            j += n;
        }
        for (int w = 0; w < nw; ++w) {
            ColInfo ci{};
            ci.trace = t;
            ci.word = (int16_t)w;
            ci.lenmask = w == nw - 1 ? len_mask(len - 64 * w) : ~0ull;
            if (h.key_kind == KEY_FIELD) {
                const int m = std::max(0, std::min(64, field - 64 * w));
                ci.kbits = (int16_t)m;
                ci.kmask = len_mask(m);
                ci.koff = (int16_t)(off + 64 * w);
            } else if (h.key_kind == KEY_STRIDE) {
                // samples q of the trace at steps q * stride inside this word
                const int s = h.stride;
                const int q0 = (64 * w + s - 1) / s;
                const int end = std::min(len, 64 * w + 64);
                const int q1 = std::min((end + s - 1) / s, nsamp);
                ci.kbits = (int16_t)std::max(0, q1 - q0);
                ci.kfirst = (int16_t)(q0 * s - 64 * w);
                ci.koff = (int16_t)std::min(off + q0, 126);
            }
            out->push_back(ci);
        }
    }
}

// Level-1 rows of an instance, column-major: word c of atom a at [c * A + a];
// steps past a trace are 0
static void atom_rows(const vfb_instance& in, std::vector<uint64_t>* rows) {
    const int A = in.A(), T = in.T(), C = in.C();
    const std::vector<int> col = in.cols();
    rows->assign((size_t)A * C, 0);
    for (int a = 0; a < A; ++a)
        for (int t = 0; t < T; ++t)
            for (int c = col[t]; c < col[t + 1]; ++c) {
                const uint64_t m = c + 1 == col[t + 1] ? len_mask(in.lens[t] - 64 * (c - col[t])) : ~0ull;
                (*rows)[(size_t)c * A + a] = in.atoms[(size_t)a * C + c] & m;
            }
}

void Engine::setup(Inst& I) {
    const vfb_instance& in = *I.in;
    const int A = in.A(), T = in.T();
    const int c1 = I.cost[0];           // the atoms' level (their cost)
    const int C = in.C();
    const std::vector<int> col = in.cols();
    std::memset(&I.h, 0, sizeof(I.h));
    I.h.T = T;
    I.h.P = in.P;
    I.h.C = C;
    I.h.mw = C > T ? 1 : 0;
    key_params(in, opt_.relaxed_type, I.h);
    I.rows.assign(kMaxS + 2, Buf{});
    I.nrows.assign(kMaxS + 2, 0);
    I.units.assign(kMaxS + 2, {});
    vfb_result& R = *I.res;
    // level 1: the atoms, all stored, checked in order (LTL_BS.cu:1127-1163)
    for (int a = 0; a < A; ++a) {
        R.candidates[c1]++;
        bool ok = true;
        for (int t = 0; t < T && ok; ++t) {
            bool b = in.atoms[(size_t)a * C + col[t]] & 1;
            if ((t < in.P) != b) ok = false;
        }
        if (ok) {
            R.min_size = c1;
            R.top_op = VFB_OP_ATOM;
            R.found_ordinal = (uint64_t)a;
            R.formula = strdup(in.names[a].c_str());
            R.level_winners[c1] = (uint64_t)A;
            R.level_candidates[c1] = (uint64_t)A;
            I.done = true;
            R.status = VFB_OK;
            return;
        }
        R.distinct[c1]++;
    }
    R.level_winners[c1] = (uint64_t)A;
    R.level_candidates[c1] = (uint64_t)A;
    R.levels_done = c1;
    if (I.h.mw) {
        // column descriptors, then the first column of each trace
        std::vector<ColInfo> ci;
        col_params(in, opt_.relaxed_type, I.h, &ci);
        const size_t n_ci = ci.size() * sizeof(ColInfo) / 8;
        std::vector<uint64_t> blob(n_ci + (T + 2) / 2, 0);
        std::memcpy(blob.data(), ci.data(), ci.size() * sizeof(ColInfo));
        int32_t* tc = (int32_t*)(blob.data() + n_ci);
        for (int t = 0; t <= T; ++t) tc[t] = col[t];
        dalloc(I, I.cols, blob.size() * 8);
        upload(I, I.cols.p, blob.data(), blob.size());
        I.h.cinfo = (const ColInfo*)I.cols.p;
        I.h.tcol = (const int32_t*)((const uint64_t*)I.cols.p + n_ci);
    }
    if (in.dev.p) {
        // resident rows (vfb_instance_upload): read in place, never written
        I.rows[c1] = Buf{in.dev.p, 0, true};
    } else {
        std::vector<uint64_t> rows;
        atom_rows(in, &rows);
        dalloc(I, I.rows[c1], rows.size() * 8);
        upload(I, I.rows[c1].p, rows.data(), rows.size());
    }
    I.nrows[c1] = (uint32_t)A;
    I.stored = (uint64_t)A;
    I.level = c1;
    I.gbase = (uint64_t)A;
}

// candidate units of the next level in canonical order, and fused tasks
void Engine::build_units(Inst& I) {
    const int s = I.level + 1;
    // LTL_BS's cost function (LTL_BS.cu:1236-1488): level = cost; p c1, ~ c2,
    // & c3, | c4, X c5, F c6, G c7, U c8 (unit costs: level = size)
    const int* c = I.cost;
    const int c1 = c[0];
    auto& U = I.units[s];
    U.clear();
    uint64_t base = 0;
    auto add = [&](int op, int ll, int lr, bool on) {
        Unit u{};
        u.op = op;
        u.lvl_l = on ? ll : 0;
        u.lvl_r = on ? lr : 0;
        u.nl = on ? I.nrows[ll] : 0;
        u.nr = on && lr ? I.nrows[lr] : 1;
        u.base = base;
        uint64_t pairs = (uint64_t)u.nl * (lr ? u.nr : 1);
        u.count = op == VFB_OP_UNTIL ? 2 * pairs : pairs;
        base += u.count;
        U.push_back(u);
        return U.size() - 1;
    };
    // canonical order: ~, & splits, | splits, X, F, G, U splits; an operator
    // with no operand level at this cost keeps an empty unit
    const int lnot = s - c[1];
    const size_t q_not = add(VFB_OP_NOT, lnot, 0, lnot >= c1 && (I.neg_type != 2 || lnot == c1));
    std::vector<size_t> q_and, q_or, q_until;
    for (int i = c1; 2 * i <= s - c[2]; ++i) q_and.push_back(add(VFB_OP_AND, i, s - i - c[2], true));
    for (int i = c1; 2 * i <= s - c[3]; ++i) q_or.push_back(add(VFB_OP_OR, i, s - i - c[3], true));
    size_t q_un[4] = {q_not, SIZE_MAX, SIZE_MAX, SIZE_MAX};
    if (!I.bs_mode) {                  // the BS stage combines with ~ & | only
        const int op3[3] = {VFB_OP_NEXT, VFB_OP_FINALLY, VFB_OP_GLOBALLY};
        for (int k = 0; k < 3; ++k) {
            const int lv = s - c[4 + k];
            q_un[1 + k] = add(op3[k], lv, 0, lv >= c1);
        }
        for (int i = c1; 2 * i <= s - c[7]; ++i) q_until.push_back(add(VFB_OP_UNTIL, i, s - i - c[7], true));
    }
    I.N = base;

    // fused tasks: the operators that read the same operand rows
    I.tasks.clear();
    std::vector<Task> un;
    for (int k = 0; k < 4; ++k) {      // ~ X F G by operand level
        if (q_un[k] == SIZE_MAX || !U[q_un[k]].count) continue;
        const Unit& u = U[q_un[k]];
        Task* t = nullptr;
        for (auto& x : un)
            if (x.lvl_l == u.lvl_l) t = &x;
        if (!t) {
            un.push_back(Task{});
            t = &un.back();
            t->kind = 0;
            t->lvl_l = (uint8_t)u.lvl_l;
            t->nl = u.nl;
            t->nr = 1;
            t->nchunk = 1;
        }
        t->ops = (uint16_t)(t->ops | (1u << k));
        t->base[k] = u.base;
    }
    std::vector<Task> bin;
    auto bin_add = [&](size_t q, uint32_t bits, int slot) {
        const Unit& u = U[q];
        if (!u.count) return;
        Task* t = nullptr;
        for (auto& x : bin)
            if (x.lvl_l == u.lvl_l && x.lvl_r == u.lvl_r) t = &x;
        if (!t) {
            bin.push_back(Task{});
            t = &bin.back();
            t->kind = 1;
            t->lvl_l = (uint8_t)u.lvl_l;
            t->lvl_r = (uint8_t)u.lvl_r;
            t->nl = u.nl;
            t->nr = u.nr;
            t->nchunk = (t->nr + 63) / 64;
        }
        t->ops = (uint16_t)(t->ops | bits);
        t->base[slot] = u.base;
    };
    for (size_t q : q_and) bin_add(q, 0x1, 0);
    for (size_t q : q_or) bin_add(q, 0x2, 1);
    for (size_t q : q_until) bin_add(q, 0xc, 2);
    I.tasks.insert(I.tasks.end(), un.begin(), un.end());
    I.tasks.insert(I.tasks.end(), bin.begin(), bin.end());

    InstDesc& h = I.h;
    h.level = s;
    h.check_only = I.check_only ? 1 : 0;
    h.nunits = (int)U.size();
    for (size_t k = 0; k < U.size(); ++k) h.units[k] = U[k];
    for (int l = 1; l < s; ++l) {
        h.lvl[l] = (const unsigned long long*)I.rows[l].p;
        h.nrows[l] = I.nrows[l];
    }
    h.epoch_base = I.gbase;
    h.found_below = ~0ull;
    h.mark = 1;
    // large levels: many items anyway, so each survivor is decoded once for
    // more traces; smaller levels keep more, shorter items
    h.store_traces = I.N >= kStoreLargeLevel ? (int)tun_.store_traces : kStoreTraces;
}

// Passes of a level.  P_FUSED: every operator in one pass.  P_PROBE: target
// check only (bit 0 of every candidate, no keys, no set), which finds m0, the
// lowest matching candidate.  A level without one cannot hold the answer and
// is built by P_FUSED.  Otherwise P_FIRST builds the candidates with ordinals
// 0..m0 (the probe's match is then settled if it is a winner, as it is unless
// a relaxed key collides with a lower ordinal), and P_REST those above m0.
// Every candidate below the returned formula is built, so counts and ranks
// are those of a fused pass.
enum Pass { P_FUSED = 0, P_PROBE = 1, P_FIRST = 2, P_REST = 3 };

static size_t nsplit_of(const Inst& I) { return (I.units[I.level + 1].size() - 4) / 3; }

// operators of the first pass for a match at m0: the canonical blocks up to
// m0's (the kernels skip what an operator mask leaves out)
static void first_pass_for(const Inst& I, uint64_t m0, uint8_t* us, uint8_t* bs) {
    const auto& U = I.units[I.level + 1];
    const size_t ns = nsplit_of(I);
    const uint64_t b_and = U[1].base, b_or = U[ns + 1].base, b_x = U[2 * ns + 1].base, b_u = U[2 * ns + 4].base;
    if (m0 < b_and) { *us = SEL_NOT; *bs = 0; }
    else if (m0 < b_or) { *us = SEL_NOT; *bs = SEL_AND; }
    else if (m0 < b_x) { *us = SEL_NOT; *bs = SEL_ANDOR; }
    else if (m0 < b_u) { *us = SEL_ALL; *bs = SEL_ANDOR; }
    else { *us = SEL_ALL; *bs = SEL_ALL; }
}

// level-local ordinal window [lo, hi] of the current pass
static void pass_window(const Inst& I, uint64_t* lo, uint64_t* hi) {
    *lo = 0;
    *hi = kEmpty;
    if (I.pass == P_FIRST) *hi = I.m0;
    else if (I.pass == P_REST) *lo = I.m0 + 1;
}

// waves (of W candidates, nchunk set for W) of task t that hold one of its
// selected candidates (t.part) with an ordinal at most hi: a prefix, since a
// task's ordinals grow with its rows (unary) or left-major pairs (binary)
static uint64_t waves_upto(const Task& t, uint64_t hi, uint64_t W) {
    bool any = false;
    uint64_t pmax = 0;
    auto upd = [&](uint64_t base, uint64_t per) {
        if (hi < base) return;
        const uint64_t p = (hi - base) / per;
        if (!any || p > pmax) pmax = p;
        any = true;
    };
    if (t.kind == 0) {
        for (int q = 0; q < 4; ++q)
            if ((t.part >> q) & 1u) upd(t.base[q], 1);
        if (!any) return 0;
        return std::min<uint64_t>(pmax, (uint64_t)t.nl - 1) / W + 1;
    }
    if (t.part & 1u) upd(t.base[0], 1);
    if (t.part & 2u) upd(t.base[1], 1);
    if (t.part & 0xcu) upd(t.base[2], 2);   // l U r and r U l of pair p are 2p, 2p + 1
    if (!any) return 0;
    const uint64_t p = std::min<uint64_t>(pmax, (uint64_t)t.nl * t.nr - 1);
    return (p / t.nr) * t.nchunk + (p % t.nr) / W + 1;
}

// whether unit q of the level is computed with unary / binary task masks us / bs
static bool unit_selected(size_t q, size_t ns, uint8_t us, uint8_t bs) {
    if (q == 0) return us & 1;                        // ~
    if (q <= ns) return bs & 1;                       // &
    if (q <= 2 * ns) return bs & 2;                   // |
    if (q <= 2 * ns + 3) return us & (1u << (q - 2 * ns));   // X F G
    return bs & 0xc;                                  // U, both orientations
}

// no keys and no set: an onTheFly level, or the probe of a level
static bool checks_only(const Inst& I) { return I.check_only || I.pass == P_PROBE; }

// candidates and algorithmic bytes (a*8T + 16 per candidate, no key slot in a
// check-only level or a probe) of pass p
static void pass_work(const Inst& I, uint64_t* cand, uint64_t* bytes) {
    const auto& U = I.units[I.level + 1];
    const size_t ns = nsplit_of(I);
    const bool all = I.pass == P_FUSED || I.pass == P_PROBE;
    uint64_t lo, hi;
    pass_window(I, &lo, &hi);
    *cand = *bytes = 0;
    for (size_t q = 0; q < U.size(); ++q) {
        if (!all && !unit_selected(q, ns, I.usel, I.bsel)) continue;
        const Unit& u = U[q];
        // the unit's ordinals [base, base + count) inside [lo, hi]
        const uint64_t s = std::max(u.base, lo);
        const uint64_t e = hi == kEmpty ? u.base + u.count : std::min(u.base + u.count, hi + 1);
        const uint64_t n = e > s ? e - s : 0;
        const uint64_t a = (u.op == VFB_OP_AND || u.op == VFB_OP_OR) ? 2 : 1;
        *cand += n;
        *bytes += n * (a * 8 * (uint64_t)I.h.C + (checks_only(I) ? 0 : 16));
    }
}

// Levels that mark their winners directly (InstDesc::direct): a fresh
// candidate marks its byte and a candidate that displaces another of the
// epoch marks that one lost, so no fresh list is written and no resolve reads
// the set again.  The engine with fixed scratch keeps the fresh list.
bool Engine::direct_admits(const Inst& I) const { return !fixed_ && I.N >= tun_.direct_min; }

// Whether a level is first probed (target checks only, k_check) before it
// is built: a deduplicating level of a unit-cost search with at least
// vfb_tuning.pass_min candidates and an operator block to split at.
bool Engine::probed(const Inst& I) const {
    return !I.check_only && !I.bs_mode && I.unit_cost && tun_.pass_min && I.N >= tun_.pass_min && nsplit_of(I) >= 1;
}

// keys the set makes room for ahead of a level that builds n candidates
uint64_t Engine::set_keys_for(const Inst& I, uint64_t n) const {
    return std::min<uint64_t>(n, std::max<uint64_t>(4096, (uint64_t)(n * std::min(1.0, I.win_ratio * set_margin()))));
}

// slots of a set for want_keys keys at load <= max_load_pct
uint64_t Engine::set_slots_for(uint64_t want_keys) const {
    // at least kMinSlots (or the tuned start size): the small levels of a
    // solve then share one set, with no fill and rehash as it grows
    const uint64_t start = std::max<uint64_t>(kMinSlots, tun_.table_min_slots);
    const uint64_t need = next_pow2(std::max<uint64_t>(start, want_keys * 100 / tun_.max_load_pct + 1));
    // no headroom: plain, batched and served solves size their sets alike, so
    // their HBM-budget decisions (and hence their results) agree
    return std::min<uint64_t>(need, 1ull << 32);
}

// fresh candidates of the level's epoch
uint64_t Engine::fresh_exact(const Inst& I) const {
    return I.h.direct ? I.ctr.fresh_real : I.ctr.fresh_count;
}

// The set of a probed level, grown for the n candidates its next pass
// builds.  False when the instance left the batch for want of memory; a
// level whose set cannot grow is checked only (as prepare_or_check does),
// and then I.check_only is set and its pass is P_FUSED.
bool Engine::grow_for_pass(Inst& I, uint64_t n) {
    try {
        ensure_table(I, I.committed + set_keys_for(I, n), false);
    } catch (CapacityError& e) {
        if (on_capacity(I, e)) return false;
        I.check_only = true;
        build_units(I);
        I.pass = P_FUSED;
        I.usel = I.bsel = SEL_ALL;
    }
    return true;
}

// scratch and hash set for a deduplicating level, sized from the previous
// level's ratios and grown on overflow
void Engine::prepare_scratch(Inst& I) {
    InstDesc& h = I.h;
    const uint64_t N = I.N;
    const uint64_t nwords = (N + 63) / 64;
    const uint64_t nblk = (nwords + kChunkWords - 1) / kChunkWords;
    // at least what the host would hold had it built the levels the level
    // chain built (the floors), so that memory use does not depend on the chain
    uint64_t fcap = std::min<uint64_t>(
        N, std::max<uint64_t>(4096, (uint64_t)(N * std::min(1.0, I.fresh_ratio * 1.5)) + 4096));
    fcap = std::max(fcap, I.floor_fcap);
    const uint64_t want_nw = std::max(nwords, I.floor_nwords);
    const uint64_t want_nb = std::max(nblk, I.floor_nblk);
    I.floor_fcap = I.floor_nwords = I.floor_nblk = 0;
    const bool direct = direct_admits(I);
    if (fcap > I.fresh_cap) {
        ensure(I, I.fresh_ord, fcap * 8);
        ensure(I, I.fresh_slot, fcap * 4);
        I.fresh_cap = fcap;
        I.lost_words = 0;
    }
    if (want_nw > I.nwords_cap) {
        // at least 64K candidates' worth, so that a solve's small levels
        // share one clearing fill
        const uint64_t nw = std::max<uint64_t>(want_nw, kMinWinWords);
        ensure(I, I.winbits, nw * 8);
        ensure(I, I.winbyte, nw * 64);
        // cleared once here; the count / scan kernels clear what they pack
        fill(I, I.winbyte.p, nw * 8, 0);
        I.nwords_cap = nw;
    }
    if (direct) {
        // the loser bytes live in the fresh list's memory, which a direct
        // level does not use: the level's memory (and so every budget
        // decision) is that of a level with a fresh list, unless its bytes
        // need more
        if (I.fresh_ord.bytes < nwords * 64) {
            ensure(I, I.fresh_ord, nwords * 64);
            I.lost_words = 0;
        }
        if (I.lost_words < nwords) {
            // cleared here as far as the level reaches (below lost_words they
            // are clear); the count / scan kernels clear what they pack
            fill(I, (char*)I.fresh_ord.p + I.lost_words * 64, (nwords - I.lost_words) * 8, 0);
            I.lost_words = nwords;
        }
    } else {
        I.lost_words = 0;   // the fresh list overwrites them
    }
    if (want_nb > I.nblk_cap) {
        ensure(I, I.blkcount, std::max<uint64_t>(want_nb, 1) * 4);
        ensure(I, I.blkprefix, std::max<uint64_t>(want_nb, 1) * 8);
        I.nblk_cap = want_nb;
    }
    h.fresh_ord = (unsigned long long*)I.fresh_ord.p;
    h.fresh_slot = (uint32_t*)I.fresh_slot.p;
    h.fresh_cap = I.fresh_cap;
    h.winbits = (unsigned long long*)I.winbits.p;
    h.winbyte = (unsigned char*)I.winbyte.p;
    h.lostbyte = direct ? (unsigned char*)I.fresh_ord.p : nullptr;
    h.direct = direct ? 1 : 0;
    h.nwords = nwords;
    h.blkcount = (uint32_t*)I.blkcount.p;
    h.blkprefix = (unsigned long long*)I.blkprefix.p;
    h.nblk = nblk;
    const uint64_t want = I.committed + set_keys_for(I, N);
    if (probed(I)) {
        // The set grows once the probe has bounded the candidates the level
        // builds (up to its match, or all of them), so a found level's set
        // is sized for its first pass.  The HBM-budget decision is still
        // that of the whole level's set, as if it were allocated here.
        const uint64_t need = set_slots_for(want);
        if (need > I.cap && I.used + align256(need * sizeof(Slot)) > I.budget)
            throw CapacityError("per-solve HBM budget exhausted", false);
        ensure_table(I, I.committed, false);
    } else {
        ensure_table(I, want, false);
    }
}

// (re)builds the hash set for want_keys keys at load <= max_load_pct; keeps only keys
// of completed epochs (ord < epoch_base)
void Engine::ensure_table(Inst& I, uint64_t want_keys, bool force_rebuild) {
    const uint64_t pct = tun_.max_load_pct;
    uint64_t need = set_slots_for(want_keys);
    if (I.table.p && need <= I.cap && !force_rebuild) {
        I.h.claim_limit = I.cap * pct / 100 > I.committed ? I.cap * pct / 100 - I.committed : 0;
        return;
    }
    if (force_rebuild && need <= I.cap) need = I.cap;
    Buf nt;
    dalloc(I, nt, need * sizeof(Slot));
    fill(I, nt.p, need * 4, kEmpty);
    if (I.table.p) {
        // the probe-limit flag is read at the next stream sync (check_flag)
        if (batch_) {
            // the rehash is only queued: the old table must outlive the flush
            // that launches it (a sync before then must not hand it back)
            rehashes_.push_back(RehashOp{&I, (const Slot*)I.table.p, I.cap, (Slot*)nt.p, need - 1, I.gbase});
            rehash_src_.push_back(RehashSrc{&I, I.table.p});
            I.used -= I.table.bytes;
            I.table = Buf{};
        } else {
            int grid = (int)std::min<uint64_t>((I.cap + 255) / 256, (uint64_t)cus_ * 8);
            CK(launch_rehash((const Slot*)I.table.p, I.cap, (Slot*)nt.p, need - 1, I.gbase, d_flag_, grid, st_));
            flag_dirty_ = true;
            release_after_sync(I, I.table);
        }
    }
    I.table = nt;
    I.cap = need;
    I.h.table = (Slot*)nt.p;
    I.h.table_mask = need - 1;
    I.h.claim_limit = need * pct / 100 > I.committed ? need * pct / 100 - I.committed : 0;
    I.res->table_slots = need;
}

// device address of the parent ordinal of row `row` of level `level`
static const unsigned long long* parent_addr(const Inst& I, int level, uint64_t row) {
    return (const unsigned long long*)I.rows[level].p + (uint64_t)I.h.C * I.nrows[level] + row;
}

// the operand nodes of candidate o of level `level` (as cand_string decodes
// it), collecting the parent ordinals not yet in par_cache_
void Engine::walk_cand(Inst& I, int level, uint64_t o, std::vector<const unsigned long long*>& need) {
    const auto& U = I.units[level];
    size_t k = 0;
    for (size_t q = 0; q < U.size(); ++q)
        if (U[q].count && U[q].base <= o) k = q;
    const Unit& u = U[k];
    const uint64_t x = o - u.base;
    if (u.op == VFB_OP_NOT || u.op == VFB_OP_NEXT || u.op == VFB_OP_FINALLY || u.op == VFB_OP_GLOBALLY) {
        walk_node(I, u.lvl_l, x, need);
        return;
    }
    const uint64_t p = u.op == VFB_OP_UNTIL ? x >> 1 : x;
    walk_node(I, u.lvl_l, p / u.nr, need);
    walk_node(I, u.lvl_r, p % u.nr, need);
}

void Engine::walk_node(Inst& I, int level, uint64_t row, std::vector<const unsigned long long*>& need) {
    if (level == I.cost[0] && !I.bs_mode) return;
    const unsigned long long* a = parent_addr(I, level, row);
    const auto it = par_cache_.find(a);
    if (it == par_cache_.end()) {
        need.push_back(a);
        return;
    }
    if (I.bs_mode && level <= I.ltl_levels) walk_node(*I.ltl, level, it->second, need);
    else walk_cand(I, level, it->second, need);
}

// Reads the parent ordinals of every node of the formulas refs will spell
// out, one tree depth at a time: a gather launch and one copy per depth for
// all of them, so that node_string then reads none itself.
void Engine::prefetch_parents(const std::vector<FormulaRef>& refs) {
    par_cache_.clear();
    if (fixed_ || refs.empty()) return;   // a fixed engine allocates nothing: node_string copies
    for (int depth = 0; depth <= 2 * kMaxS; ++depth) {
        std::vector<const unsigned long long*> need;
        for (const auto& f : refs) walk_cand(*f.I, f.level, f.o, need);
        if (need.empty()) return;
        std::sort(need.begin(), need.end());
        need.erase(std::unique(need.begin(), need.end()), need.end());
        const size_t n = need.size();
        if (n > gather_cap_) {
            sync();
            if (d_gather_) CK(hipFree(d_gather_));
            if (h_gather_) CK(hipHostFree(h_gather_));
            d_gather_ = nullptr;
            h_gather_ = nullptr;
            gather_cap_ = std::max<size_t>(n, 2 * gather_cap_);
            CK(hipMalloc((void**)&d_gather_, 2 * gather_cap_ * 8));
            CK(hipHostMalloc((void**)&h_gather_, gather_cap_ * 8, hipHostMallocDefault));
        }
        h2d(d_gather_, need.data(), n * 8);
        CK(launch_gather((const unsigned long long* const*)d_gather_, n, d_gather_ + gather_cap_, st_));
        CK(hipMemcpyAsync(h_gather_, d_gather_ + gather_cap_, n * 8, hipMemcpyDeviceToHost, st_));
        sync();
        for (size_t i = 0; i < n; ++i) par_cache_[need[i]] = h_gather_[i];
    }
}

std::string Engine::node_string(Inst& I, int level, uint64_t row) {
    if (level == I.cost[0] && !I.bs_mode) return I.in->names[row];
    uint64_t o = 0;
    const unsigned long long* a = parent_addr(I, level, row);
    const auto it = par_cache_.find(a);
    if (it != par_cache_.end()) o = it->second;
    else CK(hipMemcpy(&o, a, 8, hipMemcpyDeviceToHost));
    // a BS leaf is the CS row it was taken from
    if (I.bs_mode && level <= I.ltl_levels) return node_string(*I.ltl, level, o);
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
    drop_live(live_, &I);
    I.done = true;
    I.res->status = st;
    if (!I.res->formula) I.res->formula = strdup("Not found !");
    release_all(I);
    if (I.ltl) release_all(*I.ltl);   // a BS stage's leaves are its CS instance's rows
}

void Engine::defer(Inst& I) {
    drop_live(live_, &I);
    release_all(I);
    I.done = true;
    I.deferred = true;
    if (I.ltl) {                       // the whole solve is redone, CS stage first
        release_all(*I.ltl);
        I.ltl->deferred = true;
    }
}

// The BS stage of LTL_BS (LTL_BS.cu:808-1067, 1508-1525) for a CS instance
// whose levels 1..max_size are stored and hold no match: bit 0 of every trace
// of every stored row becomes one word (k_bs_words); the distinct words, in
// level-then-row order (the first occurrence kept, LTL_BS.cu:112-133), are
// the leaf levels of a one-trace instance whose levels max_size+1.. are the
// ~ & | combinations (unit costs; the | cost of Q6 not copied), deduplicated
// exactly (key = the word) and matched when the word is the positives mask.
void Engine::start_bs(Inst& I) {
    const int L = I.level, T = I.h.T;
    uint64_t total = 0;
    for (int l = 1; l <= L; ++l) total += I.nrows[l];
    Buf words;
    dalloc(I, words, std::max<uint64_t>(total, 1) * 8);
    uint64_t off = 0;
    for (int l = 1; l <= L; ++l) {
        CK(launch_bs_words((const unsigned long long*)I.rows[l].p, I.nrows[l], T, I.h.mw ? I.h.tcol : nullptr,
                           (unsigned long long*)words.p + off, st_));
        off += I.nrows[l];
    }
    std::vector<uint64_t> w(total);
    if (total) CK(hipMemcpyAsync(w.data(), words.p, total * 8, hipMemcpyDeviceToHost, st_));
    sync();
    dfree(I, words);
    // first occurrence of each word, level by level (row order within a level)
    auto B = std::make_unique<Inst>();
    B->bs_in = std::make_unique<vfb_instance>();
    B->bs_in->names = {"bs"};
    B->bs_in->lens = {T};
    B->in = B->bs_in.get();
    B->res = I.res;
    B->max_size = I.bs_max;
    B->budget = I.budget;
    B->prio = I.prio;
    B->bs_mode = true;
    B->ltl = &I;
    B->ltl_levels = L;
    std::memset(&B->h, 0, sizeof(B->h));
    key_params(*B->bs_in, 3, B->h);      // one trace of T bits: the exact key is the word
    if (B->h.key_kind != KEY_FIELD)      // the expansion checks BS rows only with field keys
        throw std::logic_error("BS stage without an exact field key");
    B->h.T = 1;
    B->h.C = 1;
    B->h.P = 0;
    B->h.bs_mode = 1;
    B->h.bs_target = (1ull << I.in->P) - 1;
    B->rows.assign(kMaxS + 2, Buf{});
    B->nrows.assign(kMaxS + 2, 0);
    B->units.assign(kMaxS + 2, {});
    std::unordered_set<uint64_t> seen;
    seen.reserve(total * 2 + 16);
    std::vector<uint64_t> keys;
    off = 0;
    live_.push_back(B.get());
    drop_live(live_, &I);
    I.done = true;                      // the CS stage hands over; its rows stay
    I.bs_pending = false;
    Inst& b = *B;
    I.bs = std::move(B);
    try {
        for (int l = 1; l <= L; ++l) {
            std::vector<uint64_t> vals, src;
            for (uint32_t r = 0; r < I.nrows[l]; ++r)
                if (seen.insert(w[off + r]).second) {
                    vals.push_back(w[off + r]);
                    src.push_back(r);
                }
            off += I.nrows[l];
            b.nrows[l] = (uint32_t)vals.size();
            if (vals.empty()) continue;
            // one column of words, then the parents: the CS row of each leaf
            std::vector<uint64_t> blk(vals);
            blk.insert(blk.end(), src.begin(), src.end());
            dalloc(b, b.rows[l], blk.size() * 8);
            upload(b, b.rows[l].p, blk.data(), blk.size());
            keys.insert(keys.end(), vals.begin(), vals.end());
        }
        b.level = L;
        b.stored = b.gbase = b.committed = keys.size();
        ensure_table(b, keys.size() + 4096, false);
        Buf kb;
        dalloc(b, kb, std::max<size_t>(keys.size(), 1) * 8);
        upload(b, kb.p, keys.data(), keys.size());
        flush_ops();
        CK(launch_insert_keys(b.h.table, b.h.table_mask, (const unsigned long long*)kb.p, keys.size(), d_flag_, st_));
        flag_dirty_ = true;
        release_after_sync(b, kb);
    } catch (CapacityError& e) {
        if (!on_capacity(b, e)) finish(b, VFB_CAPACITY);
    }
}

// true when the instance leaves the batch (arena memory is shared there)
bool Engine::on_capacity(Inst& I, const CapacityError& e) {
    if (e.arena) { defer(I); return true; }
    return false;
}

void Engine::epoch(std::vector<Inst*>& v) {
    flush_ops();
    const size_t n = v.size();
    // packed upload: descriptors | counters | tasks | fresh segments | total | chunk segments | ids
    // small epochs use several lanes per candidate (shorter serial trace loop)
    // tasks of each instance's current pass
    auto in_pass = [](const Inst& I, Task& t) {
        uint32_t sel;
        if (I.bs_mode) sel = t.kind == 0 ? SEL_NOT : SEL_ANDOR;   // ~ & | only
        else if (I.pass == P_FUSED || I.pass == P_PROBE) sel = SEL_ALL;
        else sel = t.kind == 0 ? I.usel : I.bsel;
        t.part = (uint8_t)(sel & t.ops);
        return t.part != 0;
    };
    // an epoch of target checks only (probes, onTheFly levels) runs k_check,
    // timed apart from the dedup expansion
    bool all_check = true;
    for (auto* I : v) all_check &= checks_only(*I);
    // waves of a task at W candidates per wave; a first pass stops at m0; a
    // binary task of k_check takes t.lrows left rows per wave
    auto task_waves = [](const Inst& I, Task& t, uint64_t W) -> uint64_t {
        t.nchunk = (uint32_t)((t.nr + W - 1) / W);
        const uint64_t lr = t.lrows > 1 ? t.lrows : 1;
        const uint64_t full = t.kind == 0 ? (t.nl + W - 1) / W : (uint64_t)((t.nl + lr - 1) / lr) * t.nchunk;
        return I.pass == P_FIRST ? std::min(full, waves_upto(t, I.m0, W)) : full;
    };
    uint64_t waves1 = 0;
    for (size_t k = 0; k < n; ++k)
        for (Task t : v[k]->tasks)
            if (in_pass(*v[k], t)) waves1 += task_waves(*v[k], t, 64);
    const int wide = waves1 < tun_.wide_max_waves ? 1 : 0;
    const uint64_t per_wave = wide ? 64 / kWideLanes : 64;
    std::vector<Task> tasks;
    uint64_t waves = 0;
    // one right-operand word for kCheckRows left rows: target checks of
    // one-word rows at one lane per candidate
    for (size_t k = 0; k < n; ++k)
        for (Task t : v[k]->tasks) {
            if (!in_pass(*v[k], t)) continue;
            t.lrows = (uint8_t)(all_check && !wide && t.kind == 1 && !v[k]->h.mw && !v[k]->h.bs_mode ? tun_.check_rows : 1);
            const uint64_t tw = task_waves(*v[k], t, per_wave);
            if (!tw) continue;
            t.inst = (uint32_t)k;
            t.wave_begin = waves;
            waves += tw;
            tasks.push_back(t);
        }
    std::vector<uint32_t> ids;
    std::vector<Seg> bsegs;
    uint64_t btot = 0;
    for (size_t k = 0; k < n; ++k) {
        if (checks_only(*v[k])) continue;
        ids.push_back((uint32_t)k);
        // small bitmaps are counted inside k_scan
        if (v[k]->h.nblk > (uint64_t)kSmallChunks) { bsegs.push_back(Seg{(uint32_t)k, 0, btot}); btot += v[k]->h.nblk; }
    }
    const size_t o_ctr = align256(n * sizeof(InstDesc));
    const size_t o_task = o_ctr + align256(n * sizeof(Ctr));
    const size_t o_fseg = o_task + align256(std::max<size_t>(tasks.size(), 1) * sizeof(Task));
    const size_t o_tot = o_fseg + align256(n * sizeof(Seg));
    const size_t o_bseg = o_tot + 256;
    const size_t o_ids = o_bseg + align256(std::max<size_t>(bsegs.size(), 1) * sizeof(Seg));
    const size_t bytes = o_ids + align256(std::max<size_t>(ids.size(), 1) * 4);
    char* d = step_buf(bytes);
    // host-coherent counters: when every instance's epoch ends with a scan,
    // the scan writes them there and no read-back copy is needed
    if (n > h_ctr_cap_) {
        if (fixed_) throw ScratchExhausted("counter staging of the fixed engine");
        sync();
        if (h_ctr_) CK(hipHostFree(h_ctr_));
        h_ctr_ = nullptr;
        h_ctr_cap_ = std::max<size_t>(n, std::max<size_t>(2 * h_ctr_cap_, 64));
        CK(hipHostMalloc((void**)&h_ctr_, h_ctr_cap_ * sizeof(Ctr), hipHostMallocCoherent));
    }
    // (a fixed engine's counters are pageable host memory: always copied back)
    bool all_scan = !fixed_;
    for (auto* I : v) all_scan &= !checks_only(*I);
    std::vector<char> host(bytes, 0);
    InstDesc* hd = (InstDesc*)host.data();
    Ctr* hc = (Ctr*)(host.data() + o_ctr);
    for (size_t k = 0; k < n; ++k) {
        Inst& I = *v[k];
        I.h.ctr = (Ctr*)(d + o_ctr) + k;
        I.h.mark = 1;
        I.h.found_below = I.pass == P_FIRST ? I.m0 + 1 : ~0ull;
        pass_window(I, &I.h.ord_lo, &I.h.ord_hi);
        hd[k] = I.h;
        hd[k].ctr_host = all_scan ? h_ctr_ + k : nullptr;
        hd[k].check_only = checks_only(I) ? 1 : 0;
        // the launch holds this instance alone: its claims (and a direct
        // level's fresh candidates) are added up per workgroup
        hd[k].defer_claims = n == 1 ? 1 : 0;
        if (I.pass != P_REST) {
            hc[k] = Ctr{};          // first pass of the level's epoch: fresh counters
            hc[k].found = kEmpty;
        } else {                    // later passes continue the epoch
            hc[k] = I.ctr;
        }
    }
    if (!tasks.empty()) std::memcpy(host.data() + o_task, tasks.data(), tasks.size() * sizeof(Task));
    if (!bsegs.empty()) std::memcpy(host.data() + o_bseg, bsegs.data(), bsegs.size() * sizeof(Seg));
    if (!ids.empty()) std::memcpy(host.data() + o_ids, ids.data(), ids.size() * 4);
    h2d(d, host.data(), bytes);
    epoch_single_ = n == 1 ? v[0] : nullptr;
    const InstDesc* D = (const InstDesc*)d;

    hipEvent_t ev = nullptr;
    const int ekind = all_check ? 3 : 0;
    if (waves) {
        time_begin(&ev, ekind);
        int grid = (int)std::min<uint64_t>((waves + 3) / 4, (uint64_t)cus_ * tun_.expand_blocks_per_cu);
        bool mw = false;
        for (auto* I : v) mw |= I->h.mw != 0;
        CK(launch_expand(D, (const Task*)(d + o_task), (int)tasks.size(), waves, wide, all_check ? -1 : mw ? 1 : 0,
                         grid, st_));
        time_end(ev, ekind, v);
        for (auto* I : v) I->res->expand_launches++;
        (ekind == 0 ? g_kstats.expand_launches : g_kstats.probe_launches)++;
    }
    last_epoch_probe_only_ = all_check;
    bool any_dedup = false;
    for (auto* I : v) any_dedup |= !checks_only(*I);
    if (any_dedup) {
        std::vector<Inst*> dd;
        for (auto* I : v)
            if (!checks_only(*I)) dd.push_back(I);
        time_begin(&ev, 1);
        unsigned long long* tot = (unsigned long long*)(d + o_tot);
        // one instance: k_resolve reads its fresh count itself (segment 0 at 0)
        const bool single = n == 1;
        // one small instance at its last pass: resolve and scan in one workgroup
        const Inst& I0 = *v[0];
        bool all_direct = true;
        for (auto* I : dd) all_direct &= I->h.direct != 0;
        if (single && !checks_only(I0) && !I0.h.direct && I0.h.nblk <= (uint64_t)kSmallChunks &&
            I0.h.fresh_cap <= kSmallFresh) {
            CK(launch_resolve_scan1(D, st_));
        } else if (all_direct) {
            // winners marked by the expansion: pack, count and scan only
            if (btot) {
                int grid = (int)std::min<uint64_t>((btot + 3) / 4, (uint64_t)cus_ * 8);
                CK(launch_count(D, (const Seg*)(d + o_bseg), (int)bsegs.size(), btot, grid, st_));
            }
            if (!ids.empty()) CK(launch_scan(D, (const uint32_t*)(d + o_ids), (int)ids.size(), st_));
        } else {
            if (!single) CK(launch_fresh_prefix(D, (Seg*)(d + o_fseg), (int)n, tot, st_));
            uint64_t fcap = 0;
            for (auto* I : dd) fcap += I->h.direct ? 0 : I->h.fresh_cap;
            const int rgrid = (int)std::max<uint64_t>(
                1, std::min<uint64_t>((fcap + 255) / 256, (uint64_t)cus_ * tun_.resolve_blocks_per_cu));
            CK(launch_resolve(D, (const Seg*)(d + o_fseg), (int)n, tot, single ? 1 : 0, rgrid, st_));
            if (btot) {
                int grid = (int)std::min<uint64_t>((btot + 3) / 4, (uint64_t)cus_ * 8);
                CK(launch_count(D, (const Seg*)(d + o_bseg), (int)bsegs.size(), btot, grid, st_));
            }
            if (!ids.empty()) CK(launch_scan(D, (const uint32_t*)(d + o_ids), (int)ids.size(), st_));
        }
        time_end(ev, 1, dd);
    }
    // otherwise (a check-only pass) an asynchronous read-back; the stream sync
    // is the level's one host round trip
    if (!all_scan) CK(hipMemcpyAsync(h_ctr_, d + o_ctr, n * sizeof(Ctr), hipMemcpyDeviceToHost, st_));
    sync();
    for (size_t k = 0; k < n; ++k) v[k]->ctr = h_ctr_[k];
}

void Engine::store(std::vector<Inst*>& v) {
    const size_t n = v.size();
    std::vector<Seg> segs;
    uint64_t tot = 0;
    // levels of few chunks split each chunk's survivors into more subsets
    bool small = true;
    for (size_t k = 0; k < n; ++k) small &= !v[k]->h.out_n || v[k]->h.nblk <= kStoreSmallBlk;
    const uint64_t parts = small ? kStorePartsSmall : kStoreParts;
    for (size_t k = 0; k < n; ++k) {
        Inst& I = *v[k];
        if (!I.h.out_n || !I.h.nblk) continue;
        segs.push_back(Seg{(uint32_t)k, 0, tot});
        tot += I.h.nblk * parts * (uint64_t)((I.h.T + I.h.store_traces - 1) / I.h.store_traces);
    }
    if (!tot) return;
    hipEvent_t ev = nullptr;
    const int grid = (int)std::min<uint64_t>((tot + 3) / 4, (uint64_t)cus_ * tun_.store_blocks_per_cu);
    if (n == 1 && epoch_single_ == v[0]) {
        // the epoch left this instance's descriptor at d_step_[0]: no upload
        const Inst& I = *v[0];
        const StorePatch p{I.h.out_rows, I.h.out_par, I.h.out_n};
        time_begin(&ev, 2);
        CK(launch_store((const InstDesc*)d_step_, nullptr, 0, tot, p, small ? 1 : 0, grid, st_));
        time_end(ev, 2, v);
        g_kstats.store_launches++;
        return;
    }
    const size_t o_seg = align256(n * sizeof(InstDesc));
    const size_t bytes = o_seg + align256(segs.size() * sizeof(Seg));
    char* d = step_buf(bytes);
    std::vector<char> host(bytes, 0);
    for (size_t k = 0; k < n; ++k) ((InstDesc*)host.data())[k] = v[k]->h;
    std::memcpy(host.data() + o_seg, segs.data(), segs.size() * sizeof(Seg));
    h2d(d, host.data(), bytes);
    epoch_single_ = nullptr;
    time_begin(&ev, 2);
    CK(launch_store((const InstDesc*)d, (const Seg*)(d + o_seg), (int)segs.size(), tot, StorePatch{}, small ? 1 : 0, grid,
                    st_));
    time_end(ev, 2, v);
    g_kstats.store_launches++;
}

// Scratch and set for a level the host builds (prepare_scratch); a level
// without the memory to deduplicate it is checked only, and the search stops.
void Engine::prepare_or_check(Inst& I, std::vector<Inst*>& step) {
    if (!I.check_only) {
        try {
            prepare_scratch(I);
        } catch (CapacityError& e) {
            if (on_capacity(I, e)) return;
            if (I.bs_mode) { finish(I, VFB_CAPACITY); return; }
            // no memory to deduplicate this level: check it only, then stop
            I.check_only = true;
            build_units(I);
        }
    }
    step.push_back(&I);
}

// Whether the level chain may build the next level (I.N candidates, units
// already built): a deduplicating level of a unit-cost CS search over
// one-word traces, below max_size, of at most n_max candidates, within the
// row cap, not one the host would probe first (vfb_tuning.pass_min), and
// one for which the host would not grow the hash set (prepare_scratch's
// ensure_table, the same arithmetic as in k_chain).
bool Engine::chain_admits(const Inst& I, uint64_t n_max) const {
    const int s = I.level + 1;
    if (I.bs_mode || !I.unit_cost || I.h.mw || I.check_only || s >= I.max_size || s == I.chain_skip || !I.table.p)
        return false;
    if (I.N == 0 || I.N > n_max) return false;
    if (I.max_rows && I.N > (I.max_rows > I.stored ? I.max_rows - I.stored : 0)) return false;
    if (probed(I)) return false;   // probed first, on the host
    const uint64_t est = set_keys_for(I, I.N);
    const uint64_t start = std::max<uint64_t>(kMinSlots, tun_.table_min_slots);
    return start <= I.cap && (I.committed + est) * 100 / tun_.max_load_pct + 1 <= I.cap;
}

// Whether the LDS level chain may build the next level: chain_admits' rules
// with its limits (the level's candidates, and the set's keys with all of
// them new), over rows the workgroup can key (every stored level's rows
// together at most the set's keys).
bool Engine::lds_chain_admits(const Inst& I) const {
    if (!chain_admits(I, kLdsChainCands) || I.committed + I.N > kLdsChainKeys) return false;
    uint64_t rows = 0;
    for (int s = 1; s <= I.level; ++s) rows += I.nrows[s];
    return rows <= kLdsChainKeys;
}

// One launch of the persistent level chain over instances v (k_chain in
// chain.hip), then one stream sync, after which each instance takes over the
// levels the chain built as if the host had built them.  Each instance gets a
// slab for the rows of those levels (given back beyond what they use); one
// whose budget or the arena has no room keeps building on the host, and is
// left in v.
void Engine::chain(std::vector<Inst*>& v, int G, uint64_t n_max, bool lds) {
    std::vector<Inst*> ch, rest;
    for (auto* I : v) {
        const uint64_t row_bytes = 8 * (uint64_t)I->h.C + 8;
        const size_t slab = align256(n_max * 3 / 2 * row_bytes) + 256 * (size_t)kMaxS;
        const uint64_t room = I->budget > I->used ? I->budget - I->used : 0;
        // only where the budget leaves far more than the chain's levels
        // take, so that no budget decision could have cut one of them
        void* p = room >= 4 * (uint64_t)slab + (64ull << 20) ? arena_->alloc(slab) : nullptr;
        if (!p) {
            rest.push_back(I);
            continue;
        }
        I->slabs.push_back(Buf{p, slab, false});
        I->used += slab;
        ch.push_back(I);
    }
    v.swap(rest);
    if (ch.empty()) return;
    const size_t n = ch.size();
    // workspace: the two winner bitmaps of every instance first, then per
    // instance its fresh list
    const size_t bmb = align256((n_max + 31) / 32 * 4), fo = align256(n_max * 8), fs = align256(n_max * 4);
    const size_t wb = 2 * bmb, per = fo + fs;
    const size_t wb_all = n * wb, need = wb_all + n * per;
    if (need > chain_ws_cap_) {
        sync();
        if (d_chain_ws_) CK(hipFree(d_chain_ws_));
        d_chain_ws_ = nullptr;
        chain_ws_cap_ = std::max(need, 2 * chain_ws_cap_);
        CK(hipMalloc((void**)&d_chain_ws_, chain_ws_cap_));
        chain_wb_clean_ = 0;
    }
    if (wb_all > chain_wb_clean_)
        CK(launch_fill((unsigned long long*)d_chain_ws_, wb_all / 8, 0, cus_ * 4, st_));
    if (n > h_chain_out_cap_) {
        sync();
        if (h_chain_out_) CK(hipHostFree(h_chain_out_));
        h_chain_out_ = nullptr;
        h_chain_out_cap_ = std::max<size_t>(n, 2 * h_chain_out_cap_);
        CK(hipHostMalloc((void**)&h_chain_out_, h_chain_out_cap_ * sizeof(ChainOut), hipHostMallocDefault));
    }
    const size_t o_out = align256(n * sizeof(ChainIn));
    const size_t bytes = o_out + align256(n * sizeof(ChainOut));
    char* d = step_buf(bytes);
    std::vector<char> host(bytes, 0);
    ChainIn* hin = (ChainIn*)host.data();
    ChainOut* hout = (ChainOut*)(host.data() + o_out);
    for (size_t k = 0; k < n; ++k) {
        const Inst& I = *ch[k];
        ChainIn& c = hin[k];
        c.d = I.h;
        c.slab = (unsigned long long*)I.slabs.back().p;
        c.slab_bytes = I.slabs.back().bytes;
        char* w = d_chain_ws_ + wb_all + k * per;
        c.bitmap[0] = (uint32_t*)(d_chain_ws_ + k * wb);
        c.bitmap[1] = (uint32_t*)(d_chain_ws_ + k * wb + bmb);
        c.fresh_ord = (unsigned long long*)w;
        c.fresh_slot = (uint32_t*)(w + fo);
        c.n_max = n_max;
        c.solo_max = std::min<uint64_t>(tun_.chain_cands, kChainSoloMax);
        c.rows_left = I.max_rows ? (I.max_rows > I.stored ? I.max_rows - I.stored : 0) : ~0ull;
        c.committed = I.committed;
        c.cap = I.cap;
        c.gbase = I.gbase;
        c.fresh_cap = std::max(I.fresh_cap, I.floor_fcap);
        c.pass_min = tun_.pass_min;
        c.win_ratio = I.win_ratio;
        c.fresh_ratio = I.fresh_ratio;
        c.pct = tun_.max_load_pct;
        c.margin = set_margin();
        c.level = I.level;
        c.last_level = I.max_size - 1;
        c.neg_type = I.neg_type;
        for (auto& x : c.ctr) x.found = kEmpty;
        hout[k].exit = CH_ERROR;
    }
    h2d(d, host.data(), bytes);
    epoch_single_ = nullptr;
    hipEvent_t ev = nullptr;
    time_begin(&ev, 4);
    int G_used = G;
    if (lds) CK(launch_chain_lds((ChainIn*)d, (ChainOut*)(d + o_out), (int)n, st_));
    else CK(launch_chain((ChainIn*)d, (ChainOut*)(d + o_out), (int)n, G, st_, &G_used));
    time_end(ev, 4, ch);
    g_kstats.chain_launches++;
    if (G_used != G) g_kstats.chain_solo_launches++;
    CK(hipMemcpyAsync(h_chain_out_, d + o_out, n * sizeof(ChainOut), hipMemcpyDeviceToHost, st_));
    chain_wb_clean_ = 0;
    sync();
    for (size_t k = 0; k < n; ++k)
        if (h_chain_out_[k].err || h_chain_out_[k].exit == CH_ERROR)
            throw DeviceError(lds ? "level chain in LDS: a key found no slot in the device set"
                                  : "level chain: a team barrier timed out");
    chain_wb_clean_ = wb_all;   // a team leaves both its bitmaps cleared
    std::vector<FormulaRef> found;
    for (size_t k = 0; k < n; ++k)
        if (chain_adopt(*ch[k], h_chain_out_[k])) found.push_back(FormulaRef{ch[k], ch[k]->level + 1, h_chain_out_[k].found});
    // the formulas found inside the chain, spelled out after batched reads
    prefetch_parents(found);
    for (auto& f : found) {
        vfb_result& R = *f.I->res;
        std::string s = cand_string(*f.I, f.level, f.o, &R.top_op);
        R.formula = strdup(s.c_str());
        finish(*f.I, VFB_OK);
    }
    par_cache_.clear();
}

// The levels one instance's team built (o.nlev, stored in its slab) become
// its levels, with the counters, results and host state a host-built level
// leaves; then its exit: a match (true: the caller spells out the formula and
// finishes the instance), or a level for the host to build again (its keys in
// the set dropped, as after an overflow there), or the next level, which the
// host builds.
bool Engine::chain_adopt(Inst& I, const ChainOut& o) {
    vfb_result& R = *I.res;
    if (opt_.verbose >= 2) {
        // per level: microseconds from the launch to the end of each phase
        std::fprintf(stderr, "chain: exit %d after %d levels (overflow %u)\n", o.exit, o.nlev, o.overflow);
        for (int j = 0; j <= o.nlev && j < kMaxS; ++j)
            std::fprintf(stderr, "chain level %d: plan %.1f expand %.1f resolve %.1f scan %.1f store %.1f us\n",
                         I.level + 1 + j, (o.t[j][0] - o.t0) * 0.01, (o.t[j][1] - o.t0) * 0.01,
                         (o.t[j][2] - o.t0) * 0.01, (o.t[j][3] - o.t0) * 0.01, (o.t[j][4] - o.t0) * 0.01);
    }
    Buf& slab = I.slabs.back();
    auto level_work = [&]() {   // the fused pass's algorithmic bytes (pass_work)
        I.pass = P_FUSED;
        uint64_t c = 0, b = 0;
        pass_work(I, &c, &b);
        R.bytes_expand += b;
        g_kstats.bytes_chain += b;
        R.chain_levels++;
        g_kstats.chain_levels++;
    };
    uint64_t end = 0;
    for (int j = 0; j < o.nlev; ++j) {
        const int s = I.level + 1;
        if (j) build_units(I);
        if (I.N != o.N[j]) throw DeviceError("level chain: a level's candidates differ from the host's units");
        // the scratch prepare_scratch would have grown to for this level
        const uint64_t N = I.N;
        const uint64_t fcap = std::min<uint64_t>(
            N, std::max<uint64_t>(4096, (uint64_t)(N * std::min(1.0, I.fresh_ratio * 1.5)) + 4096));
        const uint64_t nwords = (N + 63) / 64;
        I.floor_fcap = std::max(I.floor_fcap, fcap);
        I.floor_nwords = std::max(I.floor_nwords, nwords);
        I.floor_nblk = std::max<uint64_t>(I.floor_nblk, (nwords + kChunkWords - 1) / kChunkWords);
        level_work();
        const uint64_t W = o.W[j];
        R.level_candidates[s] = N;
        R.level_winners[s] = W;
        R.fresh_total += o.fresh[j];
        R.candidates[s] = N;
        R.distinct[s] = W;
        R.levels_done = s;
        R.bytes_store += W * (8 * (uint64_t)I.h.C + 24);
        const uint64_t row_bytes = 8 * (uint64_t)I.h.C + 8;
        const uint64_t bytes = W ? std::max<uint64_t>(256, align256(W * row_bytes)) : 0;
        I.nrows[s] = (uint32_t)W;
        I.rows[s] = W ? Buf{(char*)slab.p + o.off[j], bytes, true} : Buf{};
        if (W) end = o.off[j] + bytes;
        const uint64_t Nd = std::max<uint64_t>(N, 1);
        I.win_ratio = (double)W / Nd;
        I.fresh_ratio = (double)o.fresh[j] / Nd;
        I.committed += W;
        I.stored += W;
        I.gbase += N;
        I.level = s;
    }
    // the slab keeps the blocks of the levels built
    if (end) {
        arena_->shrink(slab.p, end);
        I.used -= slab.bytes - end;
        slab.bytes = end;
    } else {
        dfree(I, slab);
        I.slabs.pop_back();
    }
    const int s = I.level + 1;
    if (o.exit == CH_BEFORE) {
        // the host builds level s; a level the chain refused outright is
        // not offered to it again
        if (o.nlev == 0) I.chain_skip = s;
        return false;
    }
    if (o.nlev) build_units(I);
    if (I.N != o.N_exit) throw DeviceError("level chain: a level's candidates differ from the host's units");
    if (o.exit == CH_FOUND) {
        level_work();
        R.level_candidates[s] = I.N;
        R.level_winners[s] = o.W_exit;
        R.fresh_total += o.fresh_exit;
        R.candidates[s] = o.found + 1;
        R.distinct[s] = o.found_rank;
        R.min_size = s;
        R.found_ordinal = o.found;
        R.bs_stage = 0;
        return true;
    }
    // CH_REDO: the host builds level s again; its keys in the set are dropped
    // by rebuilding the set from the completed levels' keys, grown as the
    // host grows it after the same overflow
    R.retries++;
    I.chain_skip = s;
    try {
        const uint64_t keys = I.cap * tun_.max_load_pct / 100;
        ensure_table(I, (o.overflow & 5) ? keys * 4 : 0, true);
    } catch (CapacityError& e) {
        if (!on_capacity(I, e)) I.next_check_only = true;   // cannot rebuild: checked only, as on the host
    }
    return false;
}

void Engine::begin(const vfb_options& opt, bool batch) {
    opt_ = opt;
    batch_ = batch;
    defer_alone_ = false;
    park_ = false;
    parked_.clear();
    tun_ = current_tuning();
}

void Engine::run(std::vector<Inst*>& all, const vfb_options& opt) {
    begin(opt, all.size() > 1);
    defer_alone_ = true;           // vfb_solve_batch: memory-starved members are solved alone
    park_ = batch_ && tun_.park_min > 0;
    std::vector<Inst*> act;
    start(all, act);
    for (;;) {
        while (!act.empty()) step(act);
        // parked members go on one at a time, oldest first: each fills the
        // device alone, as its plain solve would
        parked_.erase(std::remove_if(parked_.begin(), parked_.end(), [](Inst* I) { return I->done; }),
                      parked_.end());
        if (parked_.empty()) break;
        auto it = std::min_element(parked_.begin(), parked_.end(),
                                   [](const Inst* a, const Inst* b) { return a->prio < b->prio; });
        Inst* I = *it;
        parked_.erase(it);
        I->parked = false;
        act.push_back(I);
    }
    park_ = false;
    sync();
    for (Inst* I : all) drop_live(live_, I);
}

void Engine::abandon(Inst& I) {
    sync();
    drop_live(live_, &I);
    release_all(I);
    I.done = true;
}

// After an error inside a call: waits for the stream, clears the failure
// flag, and removes every trace of these instances (queued operations, timing
// records, device memory), leaving the engine -- and the other instances it
// holds, such as a server's requests -- usable.
void Engine::forget(std::vector<Inst*>& v) {
    try {
        sync();
    } catch (std::exception&) {
    }
    *h_flag_ = 0;
    flag_dirty_ = false;
    (void)launch_fill((unsigned long long*)d_flag_, 1, 0, 1, st_);
    (void)hipStreamSynchronize(st_);
    std::vector<Inst*> all;
    for (Inst* I : v) {
        all.push_back(I);
        if (I->bs) all.push_back(I->bs.get());
    }
    for (auto& t : timed_)
        for (Inst* I : all) t.who.erase(std::remove(t.who.begin(), t.who.end(), I), t.who.end());
    for (Inst* I : all) {
        drop_live(live_, I);
        parked_.erase(std::remove(parked_.begin(), parked_.end(), I), parked_.end());
        if (epoch_single_ == I) epoch_single_ = nullptr;
        release_all(*I);
        I->done = true;
    }
}

void Engine::start(std::vector<Inst*>& all, std::vector<Inst*>& out) {
    std::vector<Inst*> act;
    for (auto* I : all) {
        live_.push_back(I);
        try {
            setup(*I);
            if (I->done) {   // an atom separates the traces (LTL_BS.cu:1160-1162)
                drop_live(live_, I);
                continue;
            }
            ensure_table(*I, (uint64_t)I->in->A() * 4, false);
            act.push_back(I);
        } catch (CapacityError& e) {
            if (!on_capacity(*I, e)) finish(*I, VFB_CAPACITY);
        }
    }
    // atoms into the hash sets (hashSetsInitialisation, LTL_BS.cu:1217)
    act.erase(std::remove_if(act.begin(), act.end(), [](Inst* I) { return I->done; }), act.end());
    if (!act.empty()) {
        flush_ops();    // atom rows uploaded, tables cleared
        const size_t n = act.size();
        const size_t o_ctr = align256(n * sizeof(InstDesc));
        const size_t o_seg = o_ctr + align256(n * sizeof(Ctr));
        const size_t bytes = o_seg + align256(n * sizeof(Seg));
        char* d = step_buf(bytes);
        std::vector<char> host(bytes, 0);
        uint64_t total = 0;
        for (size_t k = 0; k < n; ++k) {
            Inst& I = *act[k];
            const int c1 = I.cost[0];     // the atoms' level
            I.h.ctr = (Ctr*)(d + o_ctr) + k;
            InstDesc& hd = ((InstDesc*)host.data())[k];
            hd = I.h;
            hd.lvl[1] = (const unsigned long long*)I.rows[c1].p;   // the atom kernel reads slot 1
            hd.nrows[1] = I.nrows[c1];
            ((Seg*)(host.data() + o_seg))[k] = Seg{(uint32_t)k, 0, total};
            total += I.nrows[c1];
        }
        h2d(d, host.data(), bytes);
        // every instance's atoms in one launch
        CK(launch_insert_atoms_multi((const InstDesc*)d, (const Seg*)(d + o_seg), (int)n, total, d_flag_, st_));
        flag_dirty_ = true;
        for (size_t k = 0; k < n; ++k) {
            act[k]->committed = act[k]->nrows[act[k]->cost[0]];   // at most A distinct atom keys
            if (!act[k]->done) out.push_back(act[k]);
        }
    }
}

void Engine::step(std::vector<Inst*>& act) {
    {
        // ---- 1. units, tasks, scratch; small levels to the level chain ----
        std::vector<Inst*> step, pre;
        size_t running = 0;
        for (auto* I : act) running += I->done ? 0 : 1;
        const bool chain_on = tun_.chain_blocks > 0 && !fixed_ && chain_resident_ > 0;
        const bool lds_on = tun_.lds_chain && !fixed_;
        std::vector<Inst*> pre_lds;
        for (auto* I : act) {
            if (I->done) continue;
            I->check_only = I->next_check_only;
            build_units(*I);
            if (park_ && running > 1 && !I->bs_mode && I->N >= tun_.park_min) {
                // a level this large fills the device alone: the member
                // waits, its rows and set kept, until the others are done
                I->parked = true;
                I->res->parked = 1;
                parked_.push_back(I);
                continue;
            }
            if (lds_on && lds_chain_admits(*I)) {
                pre_lds.push_back(I);
                continue;
            }
            if (chain_on && chain_admits(*I, kChainMaxCands)) {
                pre.push_back(I);
                continue;
            }
            prepare_or_check(*I, step);
        }
        if (!pre_lds.empty()) {
            // the smallest levels, one workgroup per instance with its set
            // in LDS; members it could not take (no room for their slab)
            // build this level on the host
            chain(pre_lds, 1, kLdsChainCands, true);
            for (auto* I : pre_lds) prepare_or_check(*I, step);
        }
        if (!pre.empty()) {
            // a team of G workgroups per instance, as many as the device
            // holds at once among them; levels of up to G * chain_cands
            const int G = (int)std::min<uint64_t>(
                tun_.chain_blocks, std::max<uint64_t>(1, (uint64_t)chain_resident_ / pre.size()));
            const uint64_t n_max = std::min<uint64_t>(kChainMaxCands, (uint64_t)G * tun_.chain_cands);
            std::vector<Inst*> ch;
            for (auto* I : pre) {
                if (chain_admits(*I, n_max)) ch.push_back(I);
                else prepare_or_check(*I, step);
            }
            if (!ch.empty()) chain(ch, G, n_max);
            // members the chain could not take (no room for their slab)
            for (auto* I : ch) prepare_or_check(*I, step);
        }
        step.erase(std::remove_if(step.begin(), step.end(), [](Inst* I) { return I->done; }), step.end());
        if (step.empty()) {
            // every member is done, parked, or went on inside the level chain
            std::vector<Inst*> nxt;
            for (auto* I : act)
                if (!I->done && !I->parked) nxt.push_back(I);
            act.swap(nxt);
            return;
        }

        // ---- 2-3. epochs: the passes of each level; a level whose set or
        // list overflowed is grown and built again from its first pass ----
        for (auto* I : step) {
            // large deduplicating levels are probed first
            I->pass = probed(*I) ? P_PROBE : P_FUSED;
            I->usel = I->bsel = SEL_ALL;
            I->evaluated = 0;
        }
        std::vector<Inst*> pend = step;
        for (int round = 0; !pend.empty(); ++round) {
            if (round > 64) throw DeviceError("epoch did not converge");
            pend.erase(std::remove_if(pend.begin(), pend.end(), [](Inst* I) { return I->done; }), pend.end());
            if (pend.empty()) break;
            epoch(pend);
            std::vector<Inst*> again;
            for (auto* Ip : pend) {
                Inst& I = *Ip;
                if (I.done) continue;
                if (I.ctr.overflow) {
                    I.res->retries++;
                    if (I.h.direct) {
                        // the level's marks so far: it is built again from its first pass
                        fill(I, I.winbyte.p, I.h.nwords * 8, 0);
                        fill(I, I.h.lostbyte, I.h.nwords * 8, 0);
                    }
                    try {
                        if (I.ctr.overflow & 2) {
                            uint64_t fcap = std::min<uint64_t>(
                                I.N, std::max<uint64_t>(I.fresh_cap * 4, I.ctr.fresh_count + 4096));
                            ensure(I, I.fresh_ord, fcap * 8);
                            ensure(I, I.fresh_slot, fcap * 4);
                            I.fresh_cap = fcap;
                            I.lost_words = 0;
                            I.h.fresh_ord = (unsigned long long*)I.fresh_ord.p;
                            I.h.fresh_slot = (uint32_t*)I.fresh_slot.p;
                            I.h.fresh_cap = fcap;
                        }
                        uint64_t keys = I.cap * tun_.max_load_pct / 100;
                        ensure_table(I, (I.ctr.overflow & 5) ? keys * 4 : keys, true);
                        // the level is built again from its first dedup pass
                        if (I.pass == P_REST) {
                            I.pass = P_FIRST;
                            I.usel = I.first_usel;
                            I.bsel = I.first_bsel;
                        }
                    } catch (CapacityError& e) {
                        if (on_capacity(I, e)) continue;
                        if (I.bs_mode) { finish(I, VFB_CAPACITY); continue; }
                        I.check_only = true;   // cannot grow: check the level only
                        build_units(I);
                        I.pass = P_FUSED;
                    }
                    I.evaluated = 0;
                    again.push_back(&I);
                    continue;
                }
                uint64_t c = 0, b = 0;
                pass_work(I, &c, &b);
                if (I.pass == P_PROBE) {
                    // no match anywhere in the level: it cannot hold the
                    // answer, one fused pass; else the blocks up to the match
                    I.res->probe_candidates += c;
                    // bytes counted with the kernel whose time they share
                    (last_epoch_probe_only_ ? g_kstats.bytes_probe : g_kstats.bytes_expand) += b;
                    // the set, for the candidates the level builds next
                    if (!grow_for_pass(I, I.ctr.found != kEmpty ? I.ctr.found + 1 : I.N)) continue;
                    if (I.check_only) {
                        again.push_back(&I);
                        continue;
                    }
                    if (I.ctr.found != kEmpty) {
                        uint8_t us = SEL_ALL, bs = SEL_ALL;
                        first_pass_for(I, I.ctr.found, &us, &bs);
                        I.pass = P_FIRST;
                        I.m0 = I.ctr.found;
                        I.usel = I.first_usel = us;
                        I.bsel = I.first_bsel = bs;
                    } else {
                        I.pass = P_FUSED;
                    }
                    again.push_back(&I);
                    continue;
                }
                I.evaluated += c;
                if (I.check_only && last_epoch_probe_only_) {
                    // an onTheFly level ran in k_check, like a probe
                    I.res->probe_candidates += c;
                    g_kstats.bytes_probe += b;
                } else {
                    I.res->bytes_expand += b;
                    g_kstats.bytes_expand += b;
                }
                if (I.pass == P_FIRST && I.ctr.found == kEmpty) {
                    // the probe's match was no winner (its relaxed key was
                    // taken by a lower ordinal): the candidates above it
                    I.pass = P_REST;
                    I.usel = SEL_ALL;
                    I.bsel = SEL_ALL;
                    const uint64_t want = I.committed + set_keys_for(I, I.N);
                    if (set_slots_for(want) > I.cap) {
                        // the set was sized for the first pass: the whole
                        // level's set, and the level built again from its
                        // first pass (a rehash keeps completed levels' keys)
                        if (I.h.direct) {
                            fill(I, I.winbyte.p, I.h.nwords * 8, 0);
                            fill(I, I.h.lostbyte, I.h.nwords * 8, 0);
                        }
                        I.res->retries++;
                        if (!grow_for_pass(I, I.N)) continue;
                        if (!I.check_only) {
                            I.pass = P_FIRST;
                            I.usel = I.first_usel;
                            I.bsel = I.first_bsel;
                        }
                        I.evaluated = 0;
                    }
                    again.push_back(&I);
                }
            }
            pend.swap(again);
        }

        // ---- 4. results, overflow decisions, stores ----
        {
            // the formulas found this step: their nodes' parent ordinals in
            // batched reads before they are spelled out below
            std::vector<FormulaRef> fr;
            for (auto* Ip : step)
                if (!Ip->done && Ip->ctr.found != kEmpty) fr.push_back(FormulaRef{Ip, Ip->level + 1, Ip->ctr.found});
            prefetch_parents(fr);
        }
        std::vector<Inst*> st;
        for (auto* Ip : step) {
            Inst& I = *Ip;
            if (I.done) continue;
            vfb_result& R = *I.res;
            const int s = I.level + 1;
            R.level_candidates[s] = I.evaluated;
            if (opt_.verbose)
                std::fprintf(stderr, "Cost %-2d | %s | candidates: %-14llu | fresh: %-12llu | distinct: %-12llu | table: %llu\n",
                             s, I.check_only ? "check" : "dedup", (unsigned long long)I.N,
                             (unsigned long long)fresh_exact(I), (unsigned long long)I.ctr.winners,
                             (unsigned long long)I.cap);
            if (I.check_only) {
                // onTheFly last round (processOnTheFly, LTL_BS.cu:331-349)
                if (I.ctr.found != kEmpty) {
                    R.candidates[s] = I.ctr.found + 1;
                    R.min_size = s;
                    R.found_ordinal = I.ctr.found;
                    std::string f = cand_string(I, s, I.ctr.found, &R.top_op);
                    R.formula = strdup(f.c_str());
                    finish(I, VFB_OK);
                } else {
                    R.candidates[s] = I.N;
                    finish(I, VFB_CAPACITY);
                }
                continue;
            }
            R.level_winners[s] = I.ctr.winners;
            R.fresh_total += fresh_exact(I);
            if (I.ctr.found != kEmpty) {
                R.candidates[s] = I.ctr.found + 1;
                R.distinct[s] = I.ctr.found_rank;
                R.min_size = s;
                R.found_ordinal = I.ctr.found;
                R.bs_stage = I.bs_mode ? 1 : 0;
                std::string f = cand_string(I, s, I.ctr.found, &R.top_op);
                R.formula = strdup(f.c_str());
                finish(I, VFB_OK);
                continue;
            }
            R.candidates[s] = I.N;
            R.distinct[s] = I.ctr.winners;
            R.levels_done = s;
            if (s >= I.max_size) {
                // nothing up to max_size: the BS stage if asked for (its rows
                // are this level's too), else the search ends
                const bool bs = !I.bs_mode && I.unit_cost && I.bs_max > I.max_size && I.bs_max <= VFB_MAX_SIZE &&
                                I.h.T <= 63 && R.otf_level == 0;
                if (!bs) {
                    finish(I, VFB_NOT_FOUND);
                    continue;
                }
                I.bs_pending = true;
            }
            // rows to keep: all survivors, or as many as max_rows / the budget allow
            const uint64_t W = I.ctr.winners;
            const uint64_t row_bytes = 8 * (uint64_t)I.h.C + 8;
            uint64_t keep = W;
            if (I.max_rows) keep = std::min(keep, I.max_rows > I.stored ? I.max_rows - I.stored : 0);
            uint64_t avail = I.budget > I.used ? I.budget - I.used : 0;
            if (fixed_ && align256(std::min<uint64_t>(keep, avail > 256 ? (avail - 256) / row_bytes : 0) *
                                   row_bytes) > arena_->largest_free())
                throw ScratchExhausted("device scratch");   // not the budget: the scratch
            if (!batch_) avail = std::min<uint64_t>(avail, arena_->largest_free());
            keep = std::min<uint64_t>(keep, avail > 256 ? (avail - 256) / row_bytes : 0);
            keep = std::min<uint64_t>(keep, 0xffffffffull);
            bool evicted = false;
            while (keep > 0) {
                try {
                    dalloc(I, I.rows[s], keep * row_bytes);
                    break;
                } catch (CapacityError& e) {
                    if (on_capacity(I, e)) { evicted = true; break; }
                    // alone in a fragmented arena: keep what the largest block holds
                    const uint64_t lf = arena_->largest_free();
                    const uint64_t fit = lf > 256 ? (lf - 256) / row_bytes : 0;
                    keep = std::min<uint64_t>(fit, keep - 1);
                }
            }
            if (evicted) continue;
            if (keep < W && (I.bs_pending || I.bs_mode || !I.unit_cost)) {
                // the BS stage needs every row; a BS level, or a level under a
                // cost function, has no check-only round
                finish(I, I.bs_pending ? VFB_NOT_FOUND : VFB_CAPACITY);
                continue;
            }
            if (keep < W) {
                I.next_check_only = true;
                R.otf_level = s;
                R.otf_stored = keep;
            }
            R.bytes_store += keep * (8 * (uint64_t)I.h.C + 24);
            g_kstats.bytes_store += keep * (8 * (uint64_t)I.h.C + 24);
            I.nrows[s] = (uint32_t)keep;
            I.h.out_n = keep;
            if (keep) {
                I.h.out_rows = (unsigned long long*)I.rows[s].p;
                I.h.out_par = (unsigned long long*)I.rows[s].p + (uint64_t)I.h.C * keep;
                st.push_back(&I);
            }
            const uint64_t N = std::max<uint64_t>(I.N, 1);
            I.win_ratio = (double)W / N;
            I.fresh_ratio = (double)fresh_exact(I) / N;
            I.committed += W;
            I.stored += keep;
            I.gbase += I.N;
            I.level += 1;
        }
        par_cache_.clear();
        st.erase(std::remove_if(st.begin(), st.end(), [](Inst* I) { return I->done; }), st.end());
        if (!st.empty()) store(st);
    }
    std::vector<Inst*> nxt;
    for (auto* I : act) {
        if (!I->done && I->bs_pending) {
            try {
                start_bs(*I);
            } catch (CapacityError& e) {
                if (!on_capacity(*I, e)) finish(*I, VFB_CAPACITY);
            }
            if (I->bs && !I->bs->done) nxt.push_back(I->bs.get());
            continue;
        }
        if (!I->done && !I->parked) nxt.push_back(I);
    }
    act.swap(nxt);
}

static std::mutex g_mu;
static std::unique_ptr<Engine> g_engine;

// process-wide tuning; read by engines at begin(), under g_mu
static vfb_tuning g_tuning = [] {
    vfb_tuning t;
    vfb_tuning_default(&t);
    return t;
}();

vfb_tuning current_tuning() { return g_tuning; }

std::mutex& engine_mutex() { return g_mu; }

static uint64_t g_arena_bytes = 0;   // vfb_set_arena_bytes; 0 = 85% of the free HBM

Engine& engine_acquire(size_t want) {
    if (!g_engine || (want && want > g_engine->arena_size() && !engine_pinned())) {
        g_engine.reset();
        size_t fr = 0, tot = 0;
        CK(hipMemGetInfo(&fr, &tot));
        size_t bytes = g_arena_bytes ? std::min<size_t>(g_arena_bytes, (size_t)(fr * 0.95)) : (size_t)(fr * 0.85);
        if (want > bytes) bytes = std::min(want, (size_t)(fr * 0.95));
        g_engine = std::make_unique<Engine>(bytes);
    }
    return *g_engine;
}

void dev_rows_free(void* p) { (void)hipFree(p); }

static void reset_result(vfb_result* r) {
    std::free(r->formula);
    std::memset(r, 0, sizeof(vfb_result));
    r->status = VFB_NOT_FOUND;
}

// LTL_BS's cost function: all zero (unit costs) or all >= 1 with the atoms'
// cost within max_size; NegType 0, 1 or 2
static bool costs_valid(const vfb_options& o) {
    bool zero = true, pos = true;
    for (int k = 0; k < 8; ++k) {
        zero &= o.cost[k] == 0;
        pos &= o.cost[k] >= 1;
    }
    if (!zero && (!pos || o.cost[0] > o.max_size)) return false;
    return o.neg_type >= 0 && o.neg_type <= 2;
}

static void set_costs(Inst& I, const vfb_options& o) {
    bool zero = true;
    for (int k = 0; k < 8; ++k) zero &= o.cost[k] == 0;
    I.unit_cost = true;
    for (int k = 0; k < 8; ++k) {
        I.cost[k] = zero ? 1 : o.cost[k];
        I.unit_cost &= I.cost[k] == 1;
    }
    I.neg_type = o.neg_type == 2 ? 2 : 1;
}

// ---------------------------------------------------------------------------
// vfb_solve_noalloc: every device buffer of the solve lives in this array.
// Host code never passes its name to a kernel or copy: the array's device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and the fixed engine of that device carves everything from it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[VFB_NOALLOC_SCRATCH_BYTES];

constexpr int kNoAllocDevices = 64;
struct NoAllocDevice {
    void* ptr = nullptr;             // device address of `scratch` on this device
    std::unique_ptr<Engine> eng;     // fixed engine over it
};
static NoAllocDevice g_noalloc[kNoAllocDevices];   // under g_mu

// set-up on first use per device: the symbol's address, then the engine
static Engine& noalloc_engine() {
    int dev = 0;
    CK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) throw DeviceError("device ordinal out of range");
    NoAllocDevice& d = g_noalloc[dev];
    if (!d.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        d.ptr = p;
    }
    if (!d.eng) d.eng = std::make_unique<Engine>(Engine::Fixed{d.ptr, (size_t)VFB_NOALLOC_SCRATCH_BYTES});
    return *d.eng;
}

static vfb_status solve_many(const vfb_instance* const* insts, size_t n, const vfb_options* opt_in, vfb_result* out,
                             bool noalloc = false) {
    vfb_options opt;
    vfb_options_default(&opt);
    if (opt_in) opt = *opt_in;
    if (opt.relaxed_type < 1 || opt.relaxed_type > 3) opt.relaxed_type = 3;
    if (noalloc) opt.timing = 0;   // no events to create
    for (size_t k = 0; k < n; ++k) {
        std::memset(&out[k], 0, sizeof(vfb_result));
        out[k].status = VFB_NOT_FOUND;
    }
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Inst> st(n);
    std::vector<Inst*> run;
    for (size_t k = 0; k < n; ++k) {
        std::string why;
        if (!insts[k] || validate(*insts[k], &why) != VFB_OK || opt.max_size < 1 || opt.max_size > VFB_MAX_SIZE ||
            !costs_valid(opt)) {
            out[k].status = VFB_INVALID_INPUT;
            out[k].formula = strdup("see_the_error");
            continue;
        }
        st[k].in = insts[k];
        st[k].res = &out[k];
        st[k].max_size = opt.max_size;
        st[k].max_rows = opt.max_rows;
        st[k].bs_max = opt.bs_max_size;
        set_costs(st[k], opt);
        st[k].prio = k;
        run.push_back(&st[k]);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (opt.search == VFB_SEARCH_BRUTE) {
        // ltl_nocache's syntactic search, one instance after another
        vfb_status worst = VFB_OK;
        for (auto* I : run) {
            const vfb_status s = brute_solve(*I->in, I->max_size, opt.verbose, I->res);
            if (s == VFB_INVALID_INPUT && worst == VFB_OK) worst = s;
            if (s == VFB_DEVICE_ERROR) {
                std::free(I->res->formula);
                I->res->status = VFB_DEVICE_ERROR;
                I->res->formula = strdup("see_the_error");
                worst = VFB_DEVICE_ERROR;
            }
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (size_t k = 0; k < n; ++k) {
            out[k].ms_total = ms;
            if (!out[k].formula) out[k].formula = strdup("Not found !");
        }
        if (worst == VFB_DEVICE_ERROR) return VFB_DEVICE_ERROR;
        if (worst == VFB_INVALID_INPUT && n == 1) return VFB_INVALID_INPUT;
        return n == 1 ? out[0].status : VFB_OK;
    }
    Engine* used = nullptr;
    try {
        Engine& eng = noalloc ? noalloc_engine() : engine_acquire(opt.hbm_budget);
        used = &eng;
        uint64_t per = opt.hbm_budget ? std::min<uint64_t>(opt.hbm_budget, eng.arena_size()) : eng.arena_size();
        for (auto* I : run) I->budget = per;
        eng.run(run, opt);
        // batch members that could not get arena memory are solved alone
        for (auto* I : run) {
            if (!I->deferred) continue;
            Inst fresh;
            fresh.in = I->in;
            fresh.res = I->res;
            fresh.max_size = I->max_size;
            fresh.max_rows = I->max_rows;
            fresh.bs_max = I->bs_max;
            set_costs(fresh, opt);
            fresh.budget = per;
            const uint32_t restarts = fresh.res->restarts + 1;
            reset_result(fresh.res);
            fresh.res->restarts = restarts;
            fresh.prio = I->prio;
            *I = std::move(fresh);   // in place: an error below is cleaned up with the rest
            std::vector<Inst*> one{I};
            eng.run(one, opt);
        }
    } catch (ScratchExhausted& e) {
        // only a fixed engine throws this: the solve needs more than its scratch
        std::fprintf(stderr, "vfb_solve_noalloc: the solve needs more than the %llu-byte device scratch (%s)\n",
                     (unsigned long long)VFB_NOALLOC_SCRATCH_BYTES, e.what());
        used->forget(run);
        for (auto* I : run) {
            reset_result(I->res);
            I->res->status = VFB_CAPACITY;
            I->res->formula = strdup("see_the_error");
        }
        return VFB_CAPACITY;
    } catch (std::exception& e) {
        std::fprintf(stderr, "vfb: device error: %s\n", e.what());
        for (auto* I : run) {
            if (!I->done || I->res->status != VFB_OK) {
                std::free(I->res->formula);
                I->res->status = VFB_DEVICE_ERROR;
                I->res->formula = strdup("see_the_error");
            }
        }
        // a server's requests live in the engine: release only this call's
        // instances; otherwise start over with a new engine
        if (noalloc || (engine_pinned() && g_engine)) {
            if (used) used->forget(run);
        } else {
            g_engine.reset();
        }
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

extern "C" vfb_status vfb_solve_noalloc(const vfb_instance* inst, const vfb_options* opt, vfb_result* out) {
    if (!out) return VFB_INVALID_INPUT;
    if (opt && opt->search != VFB_SEARCH_CS) {   // the brute-force search allocates its own buffers
        std::memset(out, 0, sizeof *out);
        out->status = VFB_INVALID_INPUT;
        out->formula = strdup("see_the_error");
        return VFB_INVALID_INPUT;
    }
    const vfb_instance* one[1] = {inst};
    return solve_many(one, 1, opt, out, true);
}

extern "C" vfb_status vfb_solve_batch(const vfb_instance* const* insts, size_t n, const vfb_options* opt,
                                      vfb_result* out) {
    if (!insts || !out) return VFB_INVALID_INPUT;
    if (n == 0) return VFB_OK;
    return solve_many(insts, n, opt, out);
}

extern "C" vfb_status vfb_instance_upload(vfb_instance* inst) {
    if (!inst) return VFB_INVALID_INPUT;
    std::string why;
    if (validate(*inst, &why) != VFB_OK) return VFB_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(g_mu);
    if (inst->dev.p) return VFB_OK;
    std::vector<uint64_t> rows;
    atom_rows(*inst, &rows);
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(rows.size(), 1) * 8) != hipSuccess) {
        (void)hipGetLastError();
        return VFB_DEVICE_ERROR;
    }
    if (hipMemcpy(p, rows.data(), rows.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(p);
        return VFB_DEVICE_ERROR;
    }
    inst->dev.p = p;
    inst->dev.words = rows.size();
    return VFB_OK;
}

extern "C" vfb_status vfb_set_arena_bytes(uint64_t bytes) {
    if (bytes != 0 && bytes < (64ull << 20)) return VFB_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(g_mu);
    if (engine_pinned()) return VFB_INVALID_INPUT;
    g_arena_bytes = bytes;
    g_engine.reset();   // re-created with this arena at the next solve
    return VFB_OK;
}

extern "C" void vfb_kernel_stats_get(vfb_kernel_stats* out, int reset) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (out) *out = g_kstats;
    if (reset) g_kstats = vfb_kernel_stats{};
}

extern "C" void vfb_tuning_default(vfb_tuning* t) {
    if (!t) return;
    t->pass_min = 1ull << 22;
    t->wide_max_waves = kWideMaxWaves;
    t->expand_blocks_per_cu = 64;
    t->resolve_blocks_per_cu = 16;
    t->store_blocks_per_cu = 64;
    t->park_min = 1ull << 24;
    t->table_min_slots = kMinSlots;
    t->max_load_pct = 70;
    t->store_traces = 64;
    t->check_rows = kCheckRows;
    t->chain_blocks = 0;
    t->chain_cands = 1024;
    t->lds_chain = 1;
    t->set_margin_pct = 130;
    t->direct_min = kDirectMin;
}

// the ranges vfb_set_tuning accepts (vfb_tuning_load applies the same)
bool vfb::tuning_valid(const vfb_tuning& t) {
    auto ok = [](uint32_t b) { return b >= 1 && b <= 64; };
    return ok(t.expand_blocks_per_cu) && ok(t.resolve_blocks_per_cu) && ok(t.store_blocks_per_cu) &&
           t.table_min_slots <= (1ull << 32) && t.max_load_pct >= 50 && t.max_load_pct <= 90 && t.store_traces >= 1 &&
           t.store_traces <= 128 && (t.check_rows == 1 || t.check_rows == 4 || t.check_rows == 8 || t.check_rows == 16) &&
           t.chain_blocks <= (uint32_t)kChainMaxTeam && t.chain_cands >= 64 && t.chain_cands <= 65536 &&
           t.lds_chain <= 1 && t.set_margin_pct >= 1 && t.set_margin_pct <= 1000 &&
           t.direct_min >= 1;
}

extern "C" void vfb_get_tuning(vfb_tuning* t) {
    if (!t) return;
    std::lock_guard<std::mutex> lk(g_mu);
    *t = g_tuning;
}

extern "C" vfb_status vfb_set_tuning(const vfb_tuning* t) {
    if (!t || !tuning_valid(*t)) return VFB_INVALID_INPUT;
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuning = *t;
    return VFB_OK;
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
