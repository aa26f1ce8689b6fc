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

struct ChainShared {
    InstDesc d;                 // the instance's descriptor, updated level by level
    Task tasks[kMaxTasks];
    uint64_t N;                 // candidates of the level
    uint64_t waves;             // wavefronts of its expansion
    int ntasks;
    int ok;                     // outcome of the last team barrier
};

// Units of level s in the canonical order of the host's build_units (unit
// costs: ~, & splits, | splits, X, F, G, U splits; LTL_BS.cu:1236-1488) and
// the fused tasks over them, with kWideLanes lanes per candidate.
__device__ void chain_plan(ChainShared& sh, int s, int neg_type) {
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

// Winner test of fresh entry e (resolve_entry's), with the slot's ordinal
// read at agent scope: the atomicMin that set it may have come from another
// workgroup of this launch.
__device__ __forceinline__ void chain_resolve_entry(const InstDesc& d, uint64_t e) {
    const uint64_t of = glob(d.fresh_ord)[e];
    const uint32_t slot = glob(d.fresh_slot)[e];
    const uint64_t o = of & ~kMatchBit;
    if (ld_dev(&glob(d.table)[slot].ord) == o) {
        const uint64_t local = o - d.epoch_base;
        glob(d.winbyte)[local] = 1;
        if (of & kMatchBit) min64(glob(&d.ctr->found), local);
    }
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

__global__ __launch_bounds__(kChainThreads) void k_chain(ChainIn* ins, ChainOut* outs,
                                                         int G) {
    __shared__ ChainShared sh;
    __shared__ uint16_t lists[kChainThreads / 64][64 * kChunkWords];
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
        d.winbyte = io.winbyte;
        d.winbits = io.winbits;
        d.blkcount = io.blkcount;
        d.blkprefix = io.blkprefix;
        d.ctr_host = nullptr;
        d.check_only = 0;
        d.mark = 1;
        d.bs_mode = 0;
        d.found_below = kEmpty;
        d.ord_lo = 0;
        d.ord_hi = kEmpty;
        d.store_traces = kStoreTraces;
    }
    const bool lead = rank == 0 && threadIdx.x == 0;
    uint32_t nbar = 0;
    int level = io.level;
    uint64_t committed = io.committed, gbase = io.gbase, rows_left = io.rows_left, used = 0;
    uint64_t host_fcap = io.fresh_cap;
    double win_ratio = io.win_ratio, fresh_ratio = io.fresh_ratio;
    const uint64_t load_limit = io.cap * io.pct / 100;
    for (int k = 0;; ++k) {
        const int s = level + 1;
        __syncthreads();
        if (threadIdx.x == 0) chain_plan(sh, s, io.neg_type);
        __syncthreads();
        const uint64_t N = sh.N;
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
            const double f = win_ratio * 1.3;
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
        if (threadIdx.x == 0) {
            d.level = s;
            d.epoch_base = gbase;
            d.claim_limit = load_limit > committed ? load_limit - committed : 0;
            d.ctr = &io.ctr[s & 1];
            d.nwords = (N + 63) / 64;
            d.nblk = (d.nwords + kChunkWords - 1) / kChunkWords;
        }
        __syncthreads();

        // ---- expansion: the team's wavefronts over the level's tasks ----
        for (uint64_t w = (uint64_t)rank * kWaves + wv; w < sh.waves; w += (uint64_t)G * kWaves) {
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
        if (!team_sync(io, out, G, nbar, sh)) return;

        // ---- resolve ----
        const uint64_t fc = ld_dev(glob(&d.ctr->fresh_count));
        const uint64_t nf = fc < io.n_max ? fc : io.n_max;
        for (uint64_t e = (uint64_t)rank * kChainThreads + threadIdx.x; e < nf; e += (uint64_t)G * kChainThreads)
            chain_resolve_entry(d, e);
        if (!team_sync(io, out, G, nbar, sh)) return;

        // ---- pack, count and scan: the team's first workgroup ----
        if (rank == 0) scan_one(d);
        if (!team_sync(io, out, G, nbar, sh)) return;

        // ---- outcome, read alike by every workgroup ----
        const uint32_t ovf = __hip_atomic_load(glob(&d.ctr->overflow), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
                             (fc > host_fcap ? 2u : 0u);
        const uint64_t found = ld_dev(glob(&d.ctr->found));
        const uint64_t W = ld_dev(glob(&d.ctr->winners));
        uint64_t bytes = W ? (W * row_bytes + 255) & ~255ull : 0;   // as the host allocates a level's rows
        if (W && bytes < 256) bytes = 256;
        if (ovf || found != kEmpty) {
            if (lead) {
                out.exit = ovf ? CH_REDO : CH_FOUND;
                out.nlev = k;
                out.overflow = ovf;
                out.found = found;
                out.found_rank = ld_dev(glob(&d.ctr->found_rank));
                out.N_exit = N;
                out.W_exit = W;
                out.fresh_exit = fc;
            }
            return;
        }

        // ---- store: survivors in ordinal order, then the level is done ----
        unsigned long long* rows = io.slab + used / 8;
        {
            const uint64_t nsl = (uint64_t)((d.T + kStoreTraces - 1) / kStoreTraces);
            gu64* const orows = glob(rows);
            gu64* const opar = glob(rows + (uint64_t)d.T * W);
            if (d.nblk <= kStoreSmallBlk) {
                const uint64_t items = d.nblk * kStorePartsSmall * nsl;
                for (uint64_t g = (uint64_t)rank * kWaves + wv; g < items; g += (uint64_t)G * kWaves)
                    store_item<kStorePartsSmall>(d, g, lists[wv], orows, opar, W, lane);
            } else {
                const uint64_t items = d.nblk * kStoreParts * nsl;
                for (uint64_t g = (uint64_t)rank * kWaves + wv; g < items; g += (uint64_t)G * kWaves)
                    store_item<kStoreParts>(d, g, lists[wv], orows, opar, W, lane);
            }
        }
        if (lead) {
            out.N[k] = N;
            out.W[k] = W;
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
            d.nrows[s] = (uint32_t)W;
        }
        committed += W;
        gbase += N;
        rows_left -= W;
        used += bytes;
        win_ratio = (double)W / (double)(N > 1 ? N : 1);
        fresh_ratio = (double)fc / (double)(N > 1 ? N : 1);
        level = s;
        if (!team_sync(io, out, G, nbar, sh)) return;
    }
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
