// Persistent level chain (gfx950): the small levels of many instances in one
// launch, with no host round trip between levels.
//
// Each instance gets a team of G workgroups (G = 1 in large batches, up to
// kChainMaxTeam for a plain solve).  A team builds the levels that follow its
// instance's last completed one, each as the host path builds it with one
// fused pass (LTL_BS.cu:1226-1493 restated by Engine::step):
//   plan     units in canonical order and fused tasks (every workgroup, in LDS)
//   expand   ~ X F G per row, & | lUr rUl per pair, keys, probe, fresh list
//   resolve  fresh entries that hold their key's lowest ordinal
//   scan     pack + count + scan of the winner map (the team's first workgroup)
//   store    survivors in ordinal order into the instance's slab
// with a team barrier between phases.  A level enters the chain only if it
// has at most n_max candidates and the host would not grow the hash set for
// it; the team stops before a level that does not, or at a match, or when a
// level overflows (the host builds that level again).  Everything the
// kernels of the host path compute per candidate is shared code
// (kernels_dev.h), so a chain level is bit-identical to a host level.
//
// Team barriers: G = 1 is __syncthreads.  Otherwise thread 0 of each
// workgroup makes the workgroup's writes visible device-wide (release fence),
// counts its arrival on the instance's counter and waits, with a bounded
// spin, until all G have arrived (acquire fence).  The launch is cooperative,
// so every workgroup of a team is resident.  A spin that runs out sets the
// error flag, and every workgroup leaves at its next barrier.
#include <hip/hip_runtime.h>

#include "kernels_dev.h"

namespace vfb {

constexpr uint32_t kSpinLimit = 1u << 22;

constexpr uint32_t kBmWords = (uint32_t)(kChainMaxCands / 32);   // winner bitmap of a level, 32-bit words
constexpr uint32_t kChunkW32 = 128;                               // 32-bit words per 4096-ordinal chunk
constexpr int kMaxChunks = (int)(kChainMaxCands / 4096);

struct ChainShared {
    InstDesc d;                 // the instance's descriptor, updated level by level
    Task tasks[kMaxTasks];
    uint32_t bm[kBmWords];      // the level's winners (bit = level-local ordinal)
    uint32_t chunk_pre[kMaxChunks + 1];   // survivors ordered before each chunk
    uint32_t item_pre[kMaxChunks + 1];    // store items (64 survivors each) before each chunk
    uint32_t tsum[kChainThreads / 64];    // block scan: per-wavefront totals
    uint16_t wpre[kChainThreads / 64][64];   // store: survivors before each lane's 64 bits
    uint64_t N;                 // candidates of the level
    uint64_t waves;             // wavefronts of its expansion
    uint64_t W, found_rank;     // survivors of the level; those ordered before the match
    int ntasks;
    int ok;                     // outcome of the last team barrier
};

// Units of level s in the canonical order of the host's build_units (unit
// costs: ~, & splits, | splits, X, F, G, U splits; LTL_BS.cu:1236-1488) and
// the fused tasks over them, with kWideLanes lanes per candidate.
template <class SH>
__device__ void chain_plan(SH& sh, int s, int neg_type) {
    InstDesc& d = sh.d;
    uint64_t base = 0;
    int nu = 0;
    auto add = [&](int op, int ll, int lr, bool on) {
        Unit& u = d.units[nu++];
        u.op = op;
        u.lvl_l = on ? ll : 0;
        u.lvl_r = on ? lr : 0;
        u.nl = on ? d.nrows[ll] : 0;
        u.nr = on && lr ? d.nrows[lr] : 1;
        u.base = base;
        const uint64_t pairs = (uint64_t)u.nl * (lr ? u.nr : 1);
        u.count = op == VFB_OP_UNTIL ? 2 * pairs : pairs;
        base += u.count;
    };
    const int p = s - 1;        // operand level of ~ X F G; splits (i, p - i), 2i <= p
    const int ns = p / 2;
    add(VFB_OP_NOT, p, 0, p >= 1 && (neg_type != 2 || p == 1));
    for (int i = 1; i <= ns; ++i) add(VFB_OP_AND, i, p - i, true);
    for (int i = 1; i <= ns; ++i) add(VFB_OP_OR, i, p - i, true);
    add(VFB_OP_NEXT, p, 0, p >= 1);
    add(VFB_OP_FINALLY, p, 0, p >= 1);
    add(VFB_OP_GLOBALLY, p, 0, p >= 1);
    for (int i = 1; i <= ns; ++i) add(VFB_OP_UNTIL, i, p - i, true);
    d.nunits = nu;
    sh.N = base;

    constexpr uint64_t W = 64 / kWideLanes;
    int nt = 0;
    uint64_t waves = 0;
    {
        Task& t = sh.tasks[nt];
        t = Task{};
        const int q[4] = {0, 2 * ns + 1, 2 * ns + 2, 2 * ns + 3};   // ~ X F G
        for (int k = 0; k < 4; ++k)
            if (d.units[q[k]].count) {
                t.ops = (uint16_t)(t.ops | (1u << k));
                t.base[k] = d.units[q[k]].base;
            }
        if (t.ops) {
            t.kind = 0;
            t.lvl_l = (uint8_t)p;
            t.nl = d.nrows[p];
            t.nr = 1;
            t.nchunk = 1;
            t.part = (uint8_t)t.ops;
            t.lrows = 1;
            t.wave_begin = waves;
            waves += (t.nl + W - 1) / W;
            ++nt;
        }
    }
    for (int i = 1; i <= ns; ++i) {
        const Unit& a = d.units[i];
        const Unit& o = d.units[ns + i];
        const Unit& u = d.units[2 * ns + 3 + i];
        const uint32_t ops = (a.count ? 0x1u : 0u) | (o.count ? 0x2u : 0u) | (u.count ? 0xcu : 0u);
        if (!ops) continue;
        Task& t = sh.tasks[nt++];
        t = Task{};
        t.kind = 1;
        t.lvl_l = (uint8_t)i;
        t.lvl_r = (uint8_t)(p - i);
        t.nl = d.nrows[i];
        t.nr = d.nrows[p - i];
        t.ops = (uint16_t)ops;
        t.part = (uint8_t)ops;
        t.base[0] = a.base;
        t.base[1] = o.base;
        t.base[2] = u.base;
        t.nchunk = (uint32_t)((t.nr + W - 1) / W);
        t.lrows = 1;
        t.wave_begin = waves;
        waves += (uint64_t)t.nl * t.nchunk;
    }
    sh.ntasks = nt;
    sh.waves = waves;
}

// Winner test of fresh entry e (resolve_entry's): the candidate still holds
// its key's lowest ordinal -- the slot's ordinal read at agent scope, since
// the atomicMin that set it may have come from another workgroup.  Winners
// set their bit in `bm` (LDS for a level of one workgroup, else the team's
// bitmap in global memory); a matching winner lowers ctr->found.
template <bool LDS>
__device__ __forceinline__ void chain_resolve(const InstDesc& d, uint64_t e, uint32_t* bm) {
    const uint64_t of = glob(d.fresh_ord)[e];
    const uint32_t slot = glob(d.fresh_slot)[e];
    const uint64_t o = of & ~kMatchBit;
    if (ld_dev(&glob(d.table)[slot].ord) == o) {
        const uint32_t local = (uint32_t)(o - d.epoch_base);
        if constexpr (LDS) atomicOr(&bm[local >> 5], 1u << (local & 31));
        else (void)__hip_atomic_fetch_or(glob(&bm[local >> 5]), 1u << (local & 31), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        if (of & kMatchBit) min64(glob(&d.ctr->found), local);
    }
}

// Counts the winner bitmap in LDS (words 0..nw-1; every workgroup of a team
// the same): survivors W, survivors before each 4096-ordinal chunk, store
// items (64 survivors) before each chunk, and the rank of `found`.
__device__ void chain_scan(ChainShared& sh, uint32_t nw, uint64_t found) {
    const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
    constexpr uint32_t per = kBmWords / kChainThreads;   // words per thread (32: a quarter chunk)
    if (t == 0) sh.found_rank = 0;
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < per; ++i) {
        const uint32_t x = (uint32_t)t * per + i;
        if (x < nw) cnt += (uint32_t)__popc(sh.bm[x]);
    }
    uint32_t inc = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
    }
    if (lane == 63) sh.tsum[w] = inc;
    __syncthreads();
    uint32_t pre = inc - cnt;
    for (int k = 0; k < w; ++k) pre += sh.tsum[k];
    if ((t & 3) == 0) sh.chunk_pre[t >> 2] = pre;        // 4 threads per chunk
    if (t == kChainThreads - 1) {
        sh.chunk_pre[kMaxChunks] = pre + cnt;
        sh.W = pre + cnt;
    }
    if (found != kEmpty && (uint32_t)(found >> 5) / per == (uint32_t)t) {
        const uint32_t fw = (uint32_t)(found >> 5);
        uint32_t r = pre;
        for (uint32_t x = (uint32_t)t * per; x < fw; ++x) r += (uint32_t)__popc(sh.bm[x]);
        r += (uint32_t)__popc(sh.bm[fw] & ((1u << (found & 31)) - 1u));
        sh.found_rank = r;
    }
    __syncthreads();
    if (t == 0) {
        uint32_t acc = 0;
        for (int c = 0; c < kMaxChunks; ++c) {
            sh.item_pre[c] = acc;
            acc += (sh.chunk_pre[c + 1] - sh.chunk_pre[c] + 63) / 64;
        }
        sh.item_pre[kMaxChunks] = acc;
    }
    __syncthreads();
}

// Store of one item: survivors part*64 .. part*64+63 of chunk c (lane i takes
// the (part*64+i)-th set bit), traces t0..t1-1, as rows chunk_pre[c] + j of
// the level block (column-major, out_n = W rows); slice 0 also writes the
// parent ordinals.  The store_item of kernels_dev.h with the chunk's bits in
// LDS and a per-lane select instead of a list.
template <class SH>
__device__ __forceinline__ void chain_store_item(const SH& sh, uint32_t c, uint32_t part, int t0, int t1,
                                                 gu64* out_rows, gu64* out_par, uint64_t out_n, uint16_t* wpre,
                                                 int lane) {
    const InstDesc& d = sh.d;
    const uint32_t* cw = sh.bm + (uint64_t)c * kChunkW32;
    const uint64_t bits = ((uint64_t)cw[2 * lane + 1] << 32) | cw[2 * lane];
    const uint32_t pc = (uint32_t)__popcll(bits);
    uint32_t inc = pc;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
    }
    wpre[lane] = (uint16_t)(inc - pc);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const uint32_t n = sh.chunk_pre[c + 1] - sh.chunk_pre[c];
    const uint32_t j = part * 64 + (uint32_t)lane;
    if (j < n) {
        int lo = 0, hi = 63;   // the last lane whose bits start at or before survivor j
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (wpre[mid] <= j) lo = mid; else hi = mid - 1;
        }
        uint64_t b = ((uint64_t)cw[2 * lo + 1] << 32) | cw[2 * lo];
        for (uint32_t k = j - wpre[lo]; k; --k) b &= b - 1;
        const uint64_t o = (uint64_t)c * 4096 + (uint64_t)lo * 64 + (uint64_t)__ffsll((unsigned long long)b) - 1;
        const Cand cd = decode_from(d, o, unit_of(d, (uint64_t)c * 4096));
        const uint64_t row = sh.chunk_pre[c] + j;
        cand_words_out(cd, t0, t1, d.lenmask, out_rows, out_n, row);   // loads of 8 traces issued together
        if (t0 == 0) out_par[row] = o;
    }
    __builtin_amdgcn_wave_barrier();
}

// Barrier of the G workgroups of one team; false once any workgroup of the
// launch has given up waiting (then every workgroup leaves).
__device__ bool team_sync(ChainIn& io, ChainOut& out, int G, uint32_t& nbar, ChainShared& sh) {
    if (G == 1) {
        __syncthreads();
        return true;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();   // this workgroup's writes of the phase, visible device-wide
        ++nbar;
        add64(glob(&io.bar), 1);
        const uint64_t target = (uint64_t)nbar * (uint64_t)G;
        int ok = 1;
        for (uint32_t spins = 0; ld_dev(glob(&io.bar)) < target; ++spins) {
            if (spins > kSpinLimit || __hip_atomic_load(glob(&out.err), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                or32(glob(&out.err), 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();   // the other workgroups' writes, visible here
        sh.ok = ok;
    }
    __syncthreads();
    return sh.ok != 0;
}

__device__ __forceinline__ void store_u64(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(glob(p), (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kChainThreads) void k_chain(ChainIn* ins, ChainOut* outs, int G) {
    __shared__ ChainShared sh;
    const int team = (int)blockIdx.x / G;
    const int rank = (int)blockIdx.x - team * G;
    ChainIn& io = ins[team];
    ChainOut& out = outs[team];
    InstDesc& d = sh.d;
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr uint64_t kWaves = kChainThreads / 64;
    {
        const uint64_t* src = (const uint64_t*)&io.d;
        uint64_t* dst = (uint64_t*)&d;
        for (uint32_t i = threadIdx.x; i < sizeof(InstDesc) / 8; i += kChainThreads) dst[i] = src[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d.fresh_ord = io.fresh_ord;
        d.fresh_slot = io.fresh_slot;
        d.fresh_cap = io.n_max;
        d.ctr_host = nullptr;
        d.check_only = 0;
        d.mark = 1;
        d.bs_mode = 0;
        d.found_below = kEmpty;
        d.ord_lo = 0;
        d.ord_hi = kEmpty;
        d.defer_claims = 0;   // claims go to the counter at once (read back at the level's end)
        d.direct = 0;
    }
    const bool lead = rank == 0 && threadIdx.x == 0;
    if (lead) out.t0 = wall_clock64();
    uint32_t nbar = 0;
    int level = io.level;
    uint64_t committed = io.committed, gbase = io.gbase, rows_left = io.rows_left, used = 0;
    uint64_t host_fcap = io.fresh_cap;
    double win_ratio = io.win_ratio, fresh_ratio = io.fresh_ratio;
    const uint64_t load_limit = io.cap * io.pct / 100;
    const uint32_t bm_words = (uint32_t)((io.n_max + 31) / 32);
    for (int k = 0;; ++k) {
        const int s = level + 1;
        __syncthreads();
        if (threadIdx.x == 0) chain_plan(sh, s, io.neg_type);
        __syncthreads();
        const uint64_t N = sh.N;
        {
            // the other parity's global bitmap, for level s + 1, cleared by
            // the team: it was last read at level s - 1
            uint32_t* nb = io.bitmap[(s + 1) & 1];
            for (uint32_t i = (uint32_t)rank * kChainThreads + threadIdx.x; i < bm_words;
                 i += (uint32_t)G * kChainThreads)
                nb[i] = 0u;
        }
        const uint64_t row_bytes = 8 * (uint64_t)d.T + 8;
        const uint64_t nbytes = (N * row_bytes + 255) & ~255ull;   // the most the level's rows take
        // admission, decided alike in every workgroup: a level within the
        // size limit whose rows fit the row cap and the slab.  The host grows
        // the set where the committed keys plus the level's expected new keys
        // pass its load limit (Engine::prepare_scratch / ensure_table): such
        // a level is left to the host, with the same arithmetic
        bool stop = s > io.last_level || k >= kMaxS || N == 0 || N > io.n_max || N > rows_left ||
                    used + (nbytes < 256 ? 256 : nbytes) > io.slab_bytes ||
                    (io.pass_min && N >= io.pass_min && s >= 3);   // probed on the host (a split exists)
        if (!stop) {
            const double f = win_ratio * io.margin;
            const double m = f < 1.0 ? f : 1.0;
            uint64_t est = (uint64_t)((double)N * m);
            if (est < 4096) est = 4096;
            if (est > N) est = N;
            stop = (committed + est) * 100 / io.pct + 1 > io.cap;
        }
        if (stop) {
            if (lead) {
                out.exit = CH_BEFORE;
                out.nlev = k;
            }
            return;
        }
        {
            // the fresh list the host path would hold for this level
            // (prepare_scratch): a level whose new keys pass it overflows
            // there and is built again, so it is left to the host
            const double f = fresh_ratio * 1.5;
            uint64_t fcap = (uint64_t)((double)N * (f < 1.0 ? f : 1.0)) + 4096;
            if (fcap > N) fcap = N;
            if (fcap > host_fcap) host_fcap = fcap;
        }
        // a level of at most solo_max candidates is built by the team's
        // first workgroup alone (no team barrier inside it)
        const bool solo = G == 1 || N <= io.solo_max;
        const bool active = !solo || rank == 0;
        const uint32_t nw = (uint32_t)((N + 31) / 32);
        const uint32_t nw_chunks = (nw + kChunkW32 - 1) / kChunkW32 * kChunkW32;
        uint32_t* const gbm = io.bitmap[s & 1];
        if (lead) out.t[k][0] = wall_clock64();
        if (threadIdx.x == 0) {
            d.level = s;
            d.epoch_base = gbase;
            d.claim_limit = load_limit > committed ? load_limit - committed : 0;
            d.ctr = &io.ctr[s & 1];
        }
        if (solo && rank == 0)   // the LDS bitmap a solo level resolves into
            for (uint32_t i = threadIdx.x; i < nw_chunks; i += kChainThreads) sh.bm[i] = 0u;
        __syncthreads();

        // ---- expansion: the wavefronts of the active workgroups ----
        if (active) {
            const uint64_t nb = solo ? 1 : (uint64_t)G, rb = solo ? 0 : (uint64_t)rank;
            for (uint64_t w = rb * kWaves + wv; w < sh.waves; w += nb * kWaves) {
                int lo = 0, hi = sh.ntasks - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (sh.tasks[mid].wave_begin <= w) lo = mid; else hi = mid - 1;
                }
                const Task& tk = sh.tasks[lo];
                const uint64_t wl = w - tk.wave_begin;
                if (tk.kind == 0) {
                    if (d.key_kind == KEY_MIX) expand_unary_part<KEY_MIX, kWideLanes>(d, tk, wl, lane);
                    else if (d.key_kind == KEY_FIELD) expand_unary_part<KEY_FIELD, kWideLanes>(d, tk, wl, lane);
                    else expand_unary_part<KEY_STRIDE, kWideLanes>(d, tk, wl, lane);
                } else {
                    if (d.key_kind == KEY_MIX) expand_binary_part<KEY_MIX, kWideLanes>(d, tk, wl, lane);
                    else if (d.key_kind == KEY_FIELD) expand_binary_part<KEY_FIELD, kWideLanes>(d, tk, wl, lane);
                    else expand_binary_part<KEY_STRIDE, kWideLanes>(d, tk, wl, lane);
                }
            }
        }
        if (solo) __syncthreads();
        else if (!team_sync(io, out, G, nbar, sh)) return;
        if (lead) out.t[k][1] = wall_clock64();

        // ---- resolve: winners into the bitmap ----
        const uint64_t fc = ld_dev(glob(&d.ctr->fresh_count));
        const uint64_t nf = fc < io.n_max ? fc : io.n_max;
        if (solo) {
            if (rank == 0)
                for (uint64_t e = threadIdx.x; e < nf; e += kChainThreads) chain_resolve<true>(d, e, sh.bm);
            __syncthreads();
        } else {
            for (uint64_t e = (uint64_t)rank * kChainThreads + threadIdx.x; e < nf; e += (uint64_t)G * kChainThreads)
                chain_resolve<false>(d, e, gbm);
            if (!team_sync(io, out, G, nbar, sh)) return;
            // every workgroup counts the level's winners from its own copy
            for (uint32_t i = threadIdx.x; i < nw_chunks; i += kChainThreads) sh.bm[i] = i < nw ? gbm[i] : 0u;
            __syncthreads();
        }
        if (lead) out.t[k][2] = wall_clock64();

        // ---- count, and the level's outcome (alike in every active workgroup) ----
        uint32_t ovf = 0;
        uint64_t found = kEmpty;
        if (active) {
            found = ld_dev(glob(&d.ctr->found));
            chain_scan(sh, nw, found);
            ovf = __hip_atomic_load(glob(&d.ctr->overflow), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
                  (fc > host_fcap ? 2u : 0u);
        }
        if (lead) out.t[k][3] = wall_clock64();
        const uint64_t W = active ? sh.W : 0;
        uint64_t bytes = W ? (W * row_bytes + 255) & ~255ull : 0;   // as the host allocates a level's rows
        if (W && bytes < 256) bytes = 256;
        unsigned long long* rows = io.slab + used / 8;

        // ---- store: survivors in ordinal order (unless the level ends the chain) ----
        if (active && !ovf && found == kEmpty && W) {
            const uint32_t nsl = (uint32_t)((d.T + kStoreTraces - 1) / kStoreTraces);
            const uint32_t nch = (uint32_t)((N + 4095) / 4096);
            const uint64_t items = (uint64_t)sh.item_pre[nch] * nsl;
            gu64* const orows = glob(rows);
            gu64* const opar = glob(rows + (uint64_t)d.T * W);
            const uint64_t nb = solo ? 1 : (uint64_t)G, rb = solo ? 0 : (uint64_t)rank;
            for (uint64_t g = rb * kWaves + wv; g < items; g += nb * kWaves) {
                const uint32_t ci = (uint32_t)(g / nsl);
                const int t0 = (int)(g - (uint64_t)ci * nsl) * kStoreTraces;
                const int t1 = t0 + kStoreTraces < d.T ? t0 + kStoreTraces : d.T;
                int lo = 0, hi = (int)nch - 1;   // the chunk holding item ci
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (sh.item_pre[mid] <= ci) lo = mid; else hi = mid - 1;
                }
                chain_store_item(sh, (uint32_t)lo, ci - sh.item_pre[lo], t0, t1, orows, opar, W, sh.wpre[wv], lane);
            }
        }
        if (lead) {
            // the level's outcome for the other workgroups and the host
            store_u64(&d.ctr->winners, W);
            store_u64(&d.ctr->found_rank, sh.found_rank);
            if (ovf) __hip_atomic_fetch_or(glob(&d.ctr->overflow), ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (!team_sync(io, out, G, nbar, sh)) return;
        if (lead) out.t[k][4] = wall_clock64();
        const uint64_t Wt = ld_dev(glob(&d.ctr->winners));
        const uint64_t ft = ld_dev(glob(&d.ctr->found));
        const uint32_t ot = __hip_atomic_load(glob(&d.ctr->overflow), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (ot || ft != kEmpty) {
            if (!solo)   // this level's bitmap starts the next launch cleared
                for (uint32_t i = (uint32_t)rank * kChainThreads + threadIdx.x; i < nw; i += (uint32_t)G * kChainThreads)
                    gbm[i] = 0u;
            if (lead) {
                out.exit = ot ? CH_REDO : CH_FOUND;
                out.nlev = k;
                out.overflow = ot;
                out.found = ft;
                out.found_rank = ld_dev(glob(&d.ctr->found_rank));
                out.N_exit = N;
                out.W_exit = Wt;
                out.fresh_exit = fc;
            }
            return;
        }
        uint64_t bt = Wt ? (Wt * row_bytes + 255) & ~255ull : 0;
        if (Wt && bt < 256) bt = 256;
        if (lead) {
            out.N[k] = N;
            out.W[k] = Wt;
            out.fresh[k] = fc;
            out.off[k] = used;
            // the counters of level s + 2 start cleared (no workgroup reads
            // those of level s past this point)
            Ctr& nx = io.ctr[(s + 1) & 1];
            store_u64(&nx.fresh_count, 0);
            store_u64(&nx.claims, 0);
            store_u64(&nx.found, kEmpty);
            store_u64(&nx.winners, 0);
            store_u64(&nx.found_rank, 0);
            __hip_atomic_store(glob(&nx.overflow), 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x == 0) {
            d.lvl[s] = rows;
            d.nrows[s] = (uint32_t)Wt;
        }
        committed += Wt;
        gbase += N;
        rows_left -= Wt;
        used += bt;
        win_ratio = (double)Wt / (double)(N > 1 ? N : 1);
        fresh_ratio = (double)fc / (double)(N > 1 ? N : 1);
        level = s;
    }
}

// ---------------------------------------------------------------------------
// The level chain in LDS (k_chain_lds): one workgroup per instance builds the
// smallest levels with the instance's whole hash set in its LDS.  A level of
// a few hundred candidates is bound by the latency of its dependent memory
// operations; in the device-wide set each candidate waits on global loads and
// atomics (slot, claim, fresh-list append), here on LDS ones, and the resolve
// needs no fresh list: a slot holding a key of the level holds its lowest
// ordinal, the survivor.  At the start the set is filled with the keys of the
// instance's stored rows (levels 1..level), recomputed from their words --
// every key of the device set, whose count `committed` the workgroup checks;
// at the end the keys of the levels it completed go into the device set,
// each claiming an empty slot, with the ordinal the host path would have left
// there.  A level enters only if its candidates, and the set's keys with all
// of them new, fit (kLCands, kLKeys), under the same admission rules as
// k_chain; a level whose new keys pass the device set's load limit or the
// host's fresh list is left to the host (CH_REDO), its keys not written back.
// ---------------------------------------------------------------------------
constexpr int kLThreads = 512;                  // threads of an LDS chain workgroup (8 wavefronts)
constexpr int kLSlots = kLdsChainSlots;         // LDS set slots (24 B each)
constexpr uint64_t kLKeys = kLdsChainKeys;
constexpr uint64_t kLCands = kLdsChainCands;
static_assert(kLCands <= 4096, "an LDS level is one 4096-ordinal chunk");

struct LSlot {
    unsigned long long hi, lo, ord;
};

struct LdsShared {
    InstDesc d;
    Task tasks[kMaxTasks];
    LSlot set[kLSlots];
    uint32_t bm[kChunkW32];        // the level's winners (bit = level-local ordinal)
    uint32_t mbm[kChunkW32];       // its candidates with no target mismatch
    uint32_t chunk_pre[2];         // survivors before the level's one chunk, and in it
    uint16_t wpre[kLThreads / 64][64];
    uint64_t N, waves;
    uint32_t claims, fresh, ovf, found, nkeys, W, found_rank;
    int ntasks;
};

__shared__ LdsShared g_lsh;

__device__ __forceinline__ unsigned long long lds_ld(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned long long lds_cas(unsigned long long* p, unsigned long long e,
                                                      unsigned long long v) {
    __hip_atomic_compare_exchange_strong(p, &e, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return e;
}

// Insert-or-find in the LDS set with the device set's protocol (probe_from):
// a slot is claimed by a CAS of its ord, after which the claimer alone writes
// hi and lo; a prober that finds ord taken and the key not yet shown reads it
// again on its next step; a match lowers ord.  Returns PR_FRESH for a claim
// or a lowered ord, PR_DUP otherwise, PR_FAIL past the probe limit.  Every
// lane leaves the loop at its end only, so a claimer's stores are issued in
// the step of its claim even while a lane of its own wavefront waits on them
// (with an exit from inside the loop, the waiting lanes' path may run first
// and spin while the claimer's is held back to the loop's exit).
__device__ __forceinline__ int lds_probe(LSlot* set, unsigned long long hi, unsigned long long lo,
                                         unsigned long long gord, bool* claimed) {
    uint32_t i = (uint32_t)slot_hash(hi, lo) & (kLSlots - 1);
    unsigned long long to = lds_ld(&set[i].ord), th = lds_ld(&set[i].hi), tl = lds_ld(&set[i].lo);
    int r = PR_FAIL;
    bool done = false;
    for (int n = 0; n < kProbeLimit && !done; ++n) {
        if (to == kEmpty) {
            to = lds_cas(&set[i].ord, kEmpty, gord);
            if (to == kEmpty) {
                lds_st(&set[i].hi, hi);
                lds_st(&set[i].lo, lo);
                *claimed = true;
                r = PR_FRESH;
                done = true;
            }
        }
        if (done) {
        } else if (th == kEmpty || tl == kEmpty) {   // the claimer's key is on its way
            th = lds_ld(&set[i].hi);
            tl = lds_ld(&set[i].lo);
        } else if (th == hi && tl == lo) {
            r = to < gord || __hip_atomic_fetch_min(&set[i].ord, gord, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP) < gord
                    ? PR_DUP : PR_FRESH;
            done = true;
        } else {
            i = (i + 1) & (kLSlots - 1);
            to = lds_ld(&set[i].ord);
            th = lds_ld(&set[i].hi);
            tl = lds_ld(&set[i].lo);
        }
    }
    return r;
}

// The expansion's per-candidate end in the LDS chain (finish_one's counters
// and flags, kept in LDS): a candidate with no mismatch marks its ordinal.
struct LdsSink {
    static constexpr bool kGlobal = false;
    static __device__ __forceinline__ void finish(const InstDesc& d, const Key& key, uint64_t ord, uint32_t mismatch,
                                                  bool en, int lane) {
        (void)lane;
        if (!en) return;
        LdsShared& sh = g_lsh;
        if (!mismatch) atomicOr(&sh.mbm[ord >> 5], 1u << (ord & 31));
        bool claimed = false;
        const int r = lds_probe(sh.set, key.hi, key.lo, d.epoch_base + ord, &claimed);
        if (claimed) atomicAdd(&sh.claims, 1u);
        if (r == PR_FRESH) atomicAdd(&sh.fresh, 1u);
        if (r == PR_FAIL) atomicOr(&sh.ovf, 1u);
    }
};

// the key of a stored row (levels 1..level), as the expansion folded it
template <int KEY>
__device__ __forceinline__ Key row_key(const InstDesc& d, const gu64* rows, uint64_t nrows, uint64_t row) {
    Key k{0, 0};
    for (int t = 0; t < d.T; ++t) key_acc<KEY>(k, rows[(uint64_t)t * nrows + row], d, t);
    key_fin(k);
    return k;
}

__global__ __launch_bounds__(kLThreads) void k_chain_lds(ChainIn* ins, ChainOut* outs) {
    LdsShared& sh = g_lsh;
    ChainIn& io = ins[blockIdx.x];
    ChainOut& out = outs[blockIdx.x];
    InstDesc& d = sh.d;
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr uint32_t kWaves = kLThreads / 64;
    {
        const uint64_t* src = (const uint64_t*)&io.d;
        uint64_t* dst = (uint64_t*)&d;
        for (uint32_t i = threadIdx.x; i < sizeof(InstDesc) / 8; i += kLThreads) dst[i] = src[i];
    }
    for (int i = tid; i < kLSlots; i += kLThreads) sh.set[i] = LSlot{kEmpty, kEmpty, kEmpty};
    __syncthreads();
    const bool lead = tid == 0;
    if (lead) {
        out.t0 = wall_clock64();
        d.check_only = 0;
        d.mark = 1;
        d.bs_mode = 0;
        d.found_below = kEmpty;
        d.ord_lo = 0;
        d.ord_hi = kEmpty;
        d.ctr_host = nullptr;
        d.defer_claims = 0;
        d.direct = 0;
        sh.nkeys = 0;
        sh.ovf = 0;
    }
    // ---- the set: keys of the stored rows, with ordinal 0 (older than any level's) ----
    uint64_t nrows_all = 0;
    for (int s = 1; s <= io.level; ++s) nrows_all += d.nrows[s];
    __syncthreads();
    if (nrows_all <= kLKeys) {
        for (int s = 1; s <= io.level; ++s) {
            const gu64* rows = glob((const unsigned long long*)d.lvl[s]);
            const uint64_t n = d.nrows[s];
            for (uint64_t r = (uint64_t)tid; r < n; r += kLThreads) {
                Key k;
                if (d.key_kind == KEY_MIX) k = row_key<KEY_MIX>(d, rows, n, r);
                else if (d.key_kind == KEY_FIELD) k = row_key<KEY_FIELD>(d, rows, n, r);
                else k = row_key<KEY_STRIDE>(d, rows, n, r);
                bool claimed = false;
                if (lds_probe(sh.set, k.hi, k.lo, 0, &claimed) == PR_FAIL) atomicOr(&sh.ovf, 1u);
                if (claimed) atomicAdd(&sh.nkeys, 1u);
            }
        }
    }
    __syncthreads();
    // the rows' keys are the device set's only if there are as many
    if (nrows_all > kLKeys || sh.ovf || sh.nkeys != io.committed) {
        if (lead) {
            out.exit = CH_BEFORE;
            out.nlev = 0;
        }
        return;
    }

    int level = io.level;
    uint64_t committed = io.committed, gbase = io.gbase, rows_left = io.rows_left, used = 0;
    const uint64_t gbase0 = io.gbase;
    uint64_t host_fcap = io.fresh_cap;
    double win_ratio = io.win_ratio, fresh_ratio = io.fresh_ratio;
    const uint64_t load_limit = io.cap * io.pct / 100;
    for (int k = 0;; ++k) {
        const int s = level + 1;
        __syncthreads();
        if (tid == 0) chain_plan(sh, s, io.neg_type);
        __syncthreads();
        const uint64_t N = sh.N;
        const uint64_t row_bytes = 8 * (uint64_t)d.T + 8;
        const uint64_t nbytes = (N * row_bytes + 255) & ~255ull;
        // admission (alike in every thread): k_chain's, and the LDS set's room
        bool stop = s > io.last_level || k >= kMaxS || N == 0 || N > kLCands || N > io.n_max || N > rows_left ||
                    committed + N > kLKeys || used + (nbytes < 256 ? 256 : nbytes) > io.slab_bytes ||
                    (io.pass_min && N >= io.pass_min && s >= 3);
        if (!stop) {
            const double f = win_ratio * io.margin;
            const double m = f < 1.0 ? f : 1.0;
            uint64_t est = (uint64_t)((double)N * m);
            if (est < 4096) est = 4096;
            if (est > N) est = N;
            stop = (committed + est) * 100 / io.pct + 1 > io.cap;
        }
        if (stop) {
            if (lead) {
                out.exit = CH_BEFORE;
                out.nlev = k;
            }
            break;
        }
        {
            const double f = fresh_ratio * 1.5;
            uint64_t fcap = (uint64_t)((double)N * (f < 1.0 ? f : 1.0)) + 4096;
            if (fcap > N) fcap = N;
            if (fcap > host_fcap) host_fcap = fcap;
        }
        if (lead) out.t[k][0] = wall_clock64();
        if (tid == 0) {
            d.level = s;
            d.epoch_base = gbase;
            sh.claims = 0;
            sh.fresh = 0;
            sh.found = ~0u;
        }
        for (int i = tid; i < (int)kChunkW32; i += kLThreads) {
            sh.bm[i] = 0u;
            sh.mbm[i] = 0u;
        }
        __syncthreads();

        // ---- expansion: keys into the LDS set ----
        for (uint64_t w = wv; w < sh.waves; w += kWaves) {
            int lo = 0, hi = sh.ntasks - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sh.tasks[mid].wave_begin <= w) lo = mid; else hi = mid - 1;
            }
            const Task& tk = sh.tasks[lo];
            const uint64_t wl = w - tk.wave_begin;
            if (tk.kind == 0) {
                if (d.key_kind == KEY_MIX) expand_unary_part<KEY_MIX, kWideLanes, LdsSink>(d, tk, wl, lane);
                else if (d.key_kind == KEY_FIELD) expand_unary_part<KEY_FIELD, kWideLanes, LdsSink>(d, tk, wl, lane);
                else expand_unary_part<KEY_STRIDE, kWideLanes, LdsSink>(d, tk, wl, lane);
            } else {
                if (d.key_kind == KEY_MIX) expand_binary_part<KEY_MIX, kWideLanes, LdsSink>(d, tk, wl, lane);
                else if (d.key_kind == KEY_FIELD) expand_binary_part<KEY_FIELD, kWideLanes, LdsSink>(d, tk, wl, lane);
                else expand_binary_part<KEY_STRIDE, kWideLanes, LdsSink>(d, tk, wl, lane);
            }
        }
        __syncthreads();
        if (lead) out.t[k][1] = wall_clock64();

        // ---- resolve: a slot holding a key of this level names its survivor ----
        for (int i = tid; i < kLSlots; i += kLThreads) {
            const unsigned long long o = sh.set[i].ord;
            if (o != kEmpty && o >= gbase) {
                const uint32_t local = (uint32_t)(o - gbase);
                atomicOr(&sh.bm[local >> 5], 1u << (local & 31));
            }
        }
        __syncthreads();
        for (int i = tid; i < (int)kChunkW32; i += kLThreads) {
            const uint32_t m = sh.bm[i] & sh.mbm[i];
            if (m) atomicMin(&sh.found, (uint32_t)i * 32u + (uint32_t)(__ffs((int)m) - 1));
        }
        __syncthreads();
        if (lead) out.t[k][2] = wall_clock64();
        // ---- survivors and the rank of a match (wavefront 0, two words per lane) ----
        if (wv == 0) {
            const uint32_t f = sh.found;
            uint32_t cnt = 0, below = 0;
            for (int j = 0; j < 2; ++j) {
                const uint32_t x = (uint32_t)lane * 2u + (uint32_t)j;
                const uint32_t b = sh.bm[x];
                cnt += (uint32_t)__popc(b);
                if (f != ~0u) {
                    const uint32_t fw = f >> 5;
                    if (x < fw) below += (uint32_t)__popc(b);
                    else if (x == fw) below += (uint32_t)__popc(b & ((1u << (f & 31)) - 1u));
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                cnt += __shfl_xor(cnt, off);
                below += __shfl_xor(below, off);
            }
            if (lane == 0) {
                sh.W = cnt;
                sh.found_rank = below;
                sh.chunk_pre[0] = 0;
                sh.chunk_pre[1] = cnt;
            }
        }
        __syncthreads();
        if (lead) out.t[k][3] = wall_clock64();
        const uint64_t W = sh.W;
        const uint64_t fc = sh.fresh;
        const uint32_t ovf = sh.ovf | (committed + sh.claims > load_limit ? 4u : 0u) | (fc > host_fcap ? 2u : 0u);
        if (ovf || sh.found != ~0u) {
            if (lead) {
                out.exit = ovf ? CH_REDO : CH_FOUND;
                out.nlev = k;
                out.overflow = ovf;
                out.found = sh.found;
                out.found_rank = sh.found_rank;
                out.N_exit = N;
                out.W_exit = W;
                out.fresh_exit = fc;
            }
            break;
        }
        // ---- store: survivors in ordinal order into the slab ----
        uint64_t bt = W ? (W * row_bytes + 255) & ~255ull : 0;
        if (W && bt < 256) bt = 256;
        unsigned long long* rows = io.slab + used / 8;
        if (W) {
            // 64 survivors per item, the traces split into as many slices
            // as leave every wavefront an item (a slice at least 8 traces)
            const uint32_t parts = (uint32_t)((W + 63) / 64);
            const uint32_t most = (uint32_t)((d.T + kStoreGroup - 1) / kStoreGroup);
            uint32_t nsl = kWaves / parts;
            nsl = nsl < 1 ? 1 : (nsl > most ? most : nsl);
            const int sw = (d.T + (int)nsl - 1) / (int)nsl;
            const uint64_t items = (uint64_t)parts * nsl;
            gu64* const orows = glob(rows);
            gu64* const opar = glob(rows + (uint64_t)d.T * W);
            for (uint64_t g = wv; g < items; g += kWaves) {
                const uint32_t part = (uint32_t)(g / nsl);
                const int t0 = (int)(g - (uint64_t)part * nsl) * sw;
                const int t1 = t0 + sw < d.T ? t0 + sw : d.T;
                chain_store_item(sh, 0u, part, t0, t1, orows, opar, W, sh.wpre[wv], lane);
            }
        }
        __syncthreads();   // the level's rows, visible to this workgroup's next loads
        if (lead) {
            out.t[k][4] = wall_clock64();
            out.N[k] = N;
            out.W[k] = W;
            out.fresh[k] = fc;
            out.off[k] = used;
        }
        if (tid == 0) {
            d.lvl[s] = rows;
            d.nrows[s] = (uint32_t)W;
        }
        committed += W;
        gbase += N;
        rows_left -= W;
        used += bt;
        win_ratio = (double)W / (double)(N > 1 ? N : 1);
        fresh_ratio = (double)fc / (double)(N > 1 ? N : 1);
        level = s;
    }
    // ---- the completed levels' keys into the device set ----
    __syncthreads();
    GLB Slot* tab = glob(d.table);
    const uint64_t mask = d.table_mask;
    for (int i = tid; i < kLSlots; i += kLThreads) {
        const unsigned long long o = sh.set[i].ord;
        if (o == kEmpty || o < gbase0 || o >= gbase) continue;
        const unsigned long long hi = sh.set[i].hi, lo = sh.set[i].lo;
        uint64_t j = slot_hash(hi, lo) & mask;
        int n = 0;
        // a new key of distinct value: the first empty slot is its
        for (; n < kProbeLimit && cas64(&tab[j].ord, kEmpty, o) != kEmpty; ++n) j = (j + 1) & mask;
        if (n == kProbeLimit) {
            or32(glob(&out.err), 1u);
            continue;
        }
        st_dev(&tab[j].hi, hi);
        st_dev(&tab[j].lo, lo);
    }
}

hipError_t launch_chain_lds(ChainIn* ins, ChainOut* outs, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_chain_lds, dim3((unsigned)n), dim3(kLThreads), 0, st, ins, outs);
    return hipGetLastError();
}

// Teams of G > 1 need every workgroup resident at once: a cooperative
// launch.  The runtime refuses one while another process uses the device;
// then every team is one workgroup (no inter-workgroup barrier), which builds
// the same levels, more slowly.  *G_used receives the team size launched.
hipError_t launch_chain(ChainIn* ins, ChainOut* outs, int nteams, int G, hipStream_t st, int* G_used) {
    if (G > 1) {
        void* args[] = {&ins, &outs, &G};
        const hipError_t e = hipLaunchCooperativeKernel((const void*)k_chain, dim3((unsigned)(nteams * G)),
                                                        dim3(kChainThreads), args, 0, st);
        if (e == hipSuccess) {
            *G_used = G;
            return e;
        }
        (void)hipGetLastError();
    }
    *G_used = 1;
    hipLaunchKernelGGL(k_chain, dim3(nteams), dim3(kChainThreads), 0, st, ins, outs, 1);
    return hipGetLastError();
}

int chain_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)k_chain, kChainThreads, 0) != hipSuccess) return 0;
    return n;
}

}  // namespace vfb
