// tuning.cpp -- tuning cache files and the project autotuner.
//
// The tuned parameters (vfb_tuning) are launch shapes and the pass threshold
// of the engine; none of them changes a result, which the autotuner checks on
// its sample for every setting it times.  No reference counterpart: the
// reference fixes 1024-thread blocks and one launch per temp-buffer chunk
// (LTL_BS.cu:383-446, 1226-1493).

#include "../../include/vfb.h"

namespace vfb {
bool tuning_valid(const vfb_tuning& t);   // engine.cpp
}

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Field {
    const char* key;
    bool wide;   // uint64_t field
    size_t off;
};

const Field kFields[] = {
    {"pass_min", true, offsetof(vfb_tuning, pass_min)},
    {"wide_max_waves", false, offsetof(vfb_tuning, wide_max_waves)},
    {"expand_blocks_per_cu", false, offsetof(vfb_tuning, expand_blocks_per_cu)},
    {"resolve_blocks_per_cu", false, offsetof(vfb_tuning, resolve_blocks_per_cu)},
    {"store_blocks_per_cu", false, offsetof(vfb_tuning, store_blocks_per_cu)},
    {"park_min", true, offsetof(vfb_tuning, park_min)},
    {"table_min_slots", true, offsetof(vfb_tuning, table_min_slots)},
    {"max_load_pct", false, offsetof(vfb_tuning, max_load_pct)},
    {"store_traces", false, offsetof(vfb_tuning, store_traces)},
    {"check_rows", false, offsetof(vfb_tuning, check_rows)},
    {"chain_blocks", false, offsetof(vfb_tuning, chain_blocks)},
    {"chain_cands", false, offsetof(vfb_tuning, chain_cands)},
    {"lds_chain", false, offsetof(vfb_tuning, lds_chain)},
    {"set_margin_pct", false, offsetof(vfb_tuning, set_margin_pct)},
    {"direct_min", true, offsetof(vfb_tuning, direct_min)},
};

uint64_t get(const vfb_tuning& t, const Field& f) {
    const char* p = (const char*)&t + f.off;
    if (f.wide) return *(const uint64_t*)p;
    return *(const uint32_t*)p;
}

void put(vfb_tuning& t, const Field& f, uint64_t v) {
    char* p = (char*)&t + f.off;
    if (f.wide) *(uint64_t*)p = v;
    else *(uint32_t*)p = (uint32_t)v;
}

// the fields that must agree between two solves of one instance under any
// tuning (the per-level work counters of the found level may not)
bool same_result(const vfb_result& a, const vfb_result& b) {
    if (a.status != b.status || a.min_size != b.min_size || a.top_op != b.top_op ||
        a.found_ordinal != b.found_ordinal || a.otf_level != b.otf_level || a.otf_stored != b.otf_stored)
        return false;
    if (std::memcmp(a.distinct, b.distinct, sizeof(a.distinct)) ||
        std::memcmp(a.candidates, b.candidates, sizeof(a.candidates)))
        return false;
    const char* fa = a.formula ? a.formula : "";
    const char* fb = b.formula ? b.formula : "";
    return std::strcmp(fa, fb) == 0;
}

// one timed run of the sample; results into out (freed by the caller)
double run_sample(const vfb_instance* const* sample, size_t n, const vfb_options* opt, int batched,
                  std::vector<vfb_result>& out, vfb_status* st) {
    for (auto& r : out) vfb_result_free(&r);
    out.assign(n, vfb_result{});
    *st = VFB_OK;
    const auto t0 = std::chrono::steady_clock::now();
    if (batched) {
        *st = vfb_solve_batch(sample, n, opt, out.data());
    } else {
        for (size_t k = 0; k < n; ++k) {
            const vfb_status s = vfb_solve(sample[k], opt, &out[k]);
            if (s == VFB_DEVICE_ERROR || s == VFB_INVALID_INPUT) *st = s;
        }
    }
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" vfb_status vfb_tuning_save(const vfb_tuning* t, const char* path) {
    if (!t || !path) return VFB_INVALID_INPUT;
    FILE* f = std::fopen(path, "w");
    if (!f) return VFB_IO_ERROR;
    char dev[256] = "unknown";
    int cus = 0;
    (void)vfb_device_info(dev, sizeof dev, &cus);
    std::fprintf(f, "# vfb launch tuning (speed only; results do not depend on it)\n");
    std::fprintf(f, "device %s\ncus %d\n", dev, cus);
    for (const Field& fl : kFields) std::fprintf(f, "%s %llu\n", fl.key, (unsigned long long)get(*t, fl));
    const bool ok = std::fclose(f) == 0;
    return ok ? VFB_OK : VFB_IO_ERROR;
}

extern "C" vfb_status vfb_tuning_load(const char* path, vfb_tuning* t) {
    if (!t || !path) return VFB_INVALID_INPUT;
    FILE* f = std::fopen(path, "r");
    if (!f) return VFB_IO_ERROR;
    vfb_tuning v;
    vfb_tuning_default(&v);
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char key[128];
        unsigned long long val = 0;
        if (line[0] == '#' || std::sscanf(line, "%127s %llu", key, &val) != 2) continue;
        for (const Field& fl : kFields)
            if (std::strcmp(key, fl.key) == 0) put(v, fl, val);
    }
    std::fclose(f);
    if (!vfb::tuning_valid(v)) return VFB_INVALID_INPUT;
    *t = v;
    return VFB_OK;
}

extern "C" vfb_status vfb_autotune(const vfb_instance* const* sample, size_t n, const vfb_options* opt,
                                   int batched, int reps, const char* cache_path, vfb_tuning* best,
                                   double* ms_default, double* ms_best) {
    if (!sample || n == 0) return VFB_INVALID_INPUT;
    if (reps < 1) reps = 1;
    vfb_tuning cur;
    vfb_get_tuning(&cur);
    std::vector<vfb_result> ref, got;
    vfb_status st;
    (void)run_sample(sample, n, opt, batched, ref, &st);   // warm-up, and the reference results
    if (st == VFB_DEVICE_ERROR || st == VFB_INVALID_INPUT) {
        for (auto& r : ref) vfb_result_free(&r);
        return st;
    }
    vfb_status err = VFB_OK;
    const char* axis_key = "start";
    uint64_t axis_val = 0;
    auto time_setting = [&](const vfb_tuning& t) -> double {
        if (vfb_set_tuning(&t) != VFB_OK) return 1e300;
        double b = 1e300;
        for (int r = 0; r < reps && err == VFB_OK; ++r) {
            const double ms = run_sample(sample, n, opt, batched, got, &st);
            if (st == VFB_DEVICE_ERROR) { err = st; break; }
            for (size_t k = 0; k < n; ++k)
                if (!same_result(ref[k], got[k])) {
                    std::fprintf(stderr, "vfb_autotune: instance %zu differs under %s = %llu\n", k, axis_key,
                                 (unsigned long long)axis_val);
                    err = VFB_DEVICE_ERROR;
                    break;
                }
            if (ms < b) b = ms;
        }
        return b;
    };
    const double t_default = time_setting(cur);
    double t_best = t_default;
    static const uint64_t cand_pass[] = {0, 1ull << 20, 1ull << 22, 1ull << 24, 1ull << 26};
    static const uint64_t cand_wide[] = {0, 128, 256, 512, 1024, 2048, 8192};
    static const uint64_t cand_expand[] = {8, 16, 32, 64};
    static const uint64_t cand_resolve[] = {8, 16, 32, 64};
    static const uint64_t cand_store[] = {8, 16, 32, 64};
    static const uint64_t cand_park[] = {0, 1ull << 22, 1ull << 24, 1ull << 26};
    // (table_min_slots and max_load_pct are not searched: they size the sets,
    // and a level whose set and scratch do not fit the HBM budget is checked
    // only, so they can change where a budget-limited solve is cut)
    static const uint64_t cand_st[] = {4, 8, 16, 32, 64};
    static const uint64_t cand_rows[] = {1, 4, 8, 16};
    static const uint64_t cand_chain[] = {0, 4, 16, 64};
    static const uint64_t cand_chain_cands[] = {256, 1024, 4096};
    static const uint64_t cand_lds[] = {0, 1};
    struct Axis {
        const Field* f;
        const uint64_t* v;
        size_t nv;
    } axes[] = {{&kFields[0], cand_pass, 5}, {&kFields[1], cand_wide, 7}, {&kFields[2], cand_expand, 4},
                {&kFields[3], cand_resolve, 4}, {&kFields[4], cand_store, 4}, {&kFields[5], cand_park, 4},
                {&kFields[8], cand_st, 5},      {&kFields[9], cand_rows, 4},   {&kFields[10], cand_chain, 4},
                {&kFields[11], cand_chain_cands, 3}, {&kFields[12], cand_lds, 2}};
    // every setting timed, for the cache file: "# tried <key> <value> <ms>"
    std::string log = "# tried start - " + std::to_string(t_default) + "\n";
    for (const Axis& ax : axes) {
        for (size_t i = 0; i < ax.nv && err == VFB_OK; ++i) {
            if (get(cur, *ax.f) == ax.v[i]) continue;
            vfb_tuning t = cur;
            put(t, *ax.f, ax.v[i]);
            axis_key = ax.f->key;
            axis_val = ax.v[i];
            const double ms = time_setting(t);
            log += std::string("# tried ") + ax.f->key + " " + std::to_string(ax.v[i]) + " " + std::to_string(ms) + "\n";
            if (ms < t_best * 0.98) {   // a clear win only: timings of equal settings jitter
                t_best = ms;
                cur = t;
            }
        }
    }
    for (auto& r : ref) vfb_result_free(&r);
    for (auto& r : got) vfb_result_free(&r);
    (void)vfb_set_tuning(&cur);
    if (err != VFB_OK) return err;
    if (best) *best = cur;
    if (ms_default) *ms_default = t_default;
    if (ms_best) *ms_best = t_best;
    if (cache_path) {
        const vfb_status s = vfb_tuning_save(&cur, cache_path);
        if (s != VFB_OK) return s;
        FILE* f = std::fopen(cache_path, "a");
        if (!f) return VFB_IO_ERROR;
        std::fputs(log.c_str(), f);
        return std::fclose(f) == 0 ? VFB_OK : VFB_IO_ERROR;
    }
    return VFB_OK;
}
