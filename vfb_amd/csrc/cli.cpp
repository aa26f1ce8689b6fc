// vfb <json_file> [max_formula_size]
// vfb <input_file> <c1> <c2> <c3> <c4> <c5> <c6> <c7> <c8> <maxCost> <BSMaxCost> <RlxUnqChkTyp> <NegType>
//
// First form: drop-in for the reference binary's process contract
// (ltl_nocache.cu:448-487): reads the JSON trace file, prints
// "<file>, <formula>" on stdout (or "<file>, Not found !"), progress on
// stderr, and exits 0 like the reference (errors are reported on stderr).
// The size bound is optional here and defaults to 32, the reference's
// maximum (ltl_nocache.cu:11).
//
// Second form: LTL_BS's own command line (LTL_BS.cu:1609-1712): its text
// trace format, cost function, BS stage bound, relaxed key type and NegType;
// it prints the traces, the cost function and `LTL: "<formula>"` in its
// syntax ("(l)&(r)", "~(x)").  Sizes here are bounded by 32 (maxCost) and
// the BS stage runs with unit costs only.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vfb.h"

namespace {

// LTL_BS's output syntax: binary operators without spaces (LTLToString,
// LTL_BS.cu:591-650); ltl_nocache's text has them as ") op ("
std::string ltlbs_syntax(const std::string& f) {
    std::string out;
    for (size_t i = 0; i < f.size(); ++i) {
        if (f[i] == ')' && i + 4 < f.size() && f[i + 1] == ' ' && f[i + 3] == ' ' && f[i + 4] == '(' &&
            (f[i + 2] == '&' || f[i + 2] == '|' || f[i + 2] == 'U')) {
            out += ')';
            out += f[i + 2];
            out += '(';
            i += 4;
            continue;
        }
        out += f[i];
    }
    return out;
}

void print_traces(const char* title, const vfb_instance* inst, int t0, int t1) {
    int A = 0, T = 0, P = 0;
    vfb_instance_shape(inst, &A, &T, &P);
    const int W = vfb_instance_words(inst);
    std::vector<int> lens(T), col(T + 1, 0);
    std::vector<uint64_t> atoms((size_t)A * W);
    vfb_instance_copy(inst, lens.data(), atoms.data());
    for (int t = 0; t < T; ++t) col[t + 1] = col[t] + (lens[t] + 63) / 64;
    std::printf("\n%s: \n", title);
    for (int t = t0; t < t1; ++t) {
        std::printf("\t");
        for (int k = 0; k < lens[t]; ++k) {
            std::string s;
            for (int a = 0; a < A; ++a)
                if ((atoms[(size_t)a * W + col[t] + k / 64] >> (k % 64)) & 1) {
                    if (!s.empty()) s += ", ";
                    s += vfb_instance_atom_name(inst, a);
                }
            std::printf("{%s}\t", s.c_str());
        }
        std::printf("\n");
    }
}

int ltlbs_main(int argc, char** argv) {
    bool err = false;
    for (int i = 2; i < argc - 2; ++i) {
        const int v = std::atoi(argv[i]);
        if (v <= 0 || v > SHRT_MAX) {
            std::printf("Argument number %d, \"%s\", should be a positive short integer.\n", i, argv[i]);
            err = true;
        }
    }
    const int rlx = std::atoi(argv[12]), neg = std::atoi(argv[13]);
    if (rlx < 1 || rlx > 3) {
        std::printf("Argument number 12, RlxUnqChkTyp = \"%s\", should be 1, 2, or 3.\n", argv[12]);
        err = true;
    }
    if (neg < 1 || neg > 2) {
        std::printf("Argument number 13, NegType = \"%s\", should be 1, or 2.\n", argv[13]);
        err = true;
    }
    if (err) return 0;
    vfb_instance* inst = nullptr;
    if (vfb_instance_load_text(argv[1], &inst) != VFB_OK) return 0;
    vfb_options opt;
    vfb_options_default(&opt);
    for (int k = 0; k < 8; ++k) opt.cost[k] = (uint16_t)std::atoi(argv[k + 2]);
    const int max_cost = std::atoi(argv[10]), bs_max = std::atoi(argv[11]);
    if (max_cost > VFB_MAX_SIZE) {
        std::printf("In this version, maxCost can be at most %d. It is currently %d.\n", VFB_MAX_SIZE, max_cost);
        vfb_instance_free(inst);
        return 0;
    }
    opt.max_size = max_cost;
    opt.bs_max_size = bs_max > max_cost && bs_max <= VFB_MAX_SIZE ? bs_max : 0;
    opt.relaxed_type = rlx;
    opt.neg_type = neg;
    opt.verbose = std::getenv("VFB_QUIET") ? 0 : 1;
    vfb_result res;
    const vfb_status st = vfb_solve(inst, &opt, &res);
    if (st == VFB_INVALID_INPUT || st == VFB_DEVICE_ERROR) {
        std::fprintf(stderr, "vfb: %s\n", vfb_status_string(st));
    } else {
        int A = 0, T = 0, P = 0;
        vfb_instance_shape(inst, &A, &T, &P);
        print_traces("Positive", inst, 0, P);
        print_traces("Negative", inst, P, T);
        std::printf("\nCost Function: p:%u, ~:%u, &:%u, |:%u, X:%u, F:%u, G:%u, U:%u", opt.cost[0], opt.cost[1],
                    opt.cost[2], opt.cost[3], opt.cost[4], opt.cost[5], opt.cost[6], opt.cost[7]);
        std::printf("\nNumber of Traces: %d", T);
        const std::string f = st == VFB_OK && res.formula ? ltlbs_syntax(res.formula) : std::string("not_found");
        std::printf("\n\nLTL: \"%s\"\n", f.c_str());
    }
    vfb_result_free(&res);
    vfb_instance_free(inst);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 14) return ltlbs_main(argc, argv);
    if (argc < 2 || argc > 3) {
        std::fprintf(stderr, "Arguments should be in the form of\n");
        std::fprintf(stderr, "-----------------------------------------------------------------\n");
        std::fprintf(stderr, "%s <input_file_address> [max_formula_size]\n", argv[0]);
        std::fprintf(stderr, "%s <input_file_address> <c1> <c2> <c3> <c4> <c5> <c6> <c7> <c8> <maxCost> "
                             "<BSMaxCost> <RlxUnqChkTyp> <NegType>\n", argv[0]);
        std::fprintf(stderr, "-----------------------------------------------------------------\n");
        return 0;
    }
    int max_size = VFB_MAX_SIZE;
    if (argc == 3) {
        max_size = std::atoi(argv[2]);
        if (max_size < 1 || max_size > VFB_MAX_SIZE) {
            std::fprintf(stderr, "Argument maxLen = %s should be between 1 and %d\n", argv[2], VFB_MAX_SIZE);
            return 0;
        }
    }
    vfb_instance* inst = nullptr;
    if (vfb_instance_load_json(argv[1], &inst) != VFB_OK) return 0;
    vfb_options opt;
    vfb_options_default(&opt);
    opt.max_size = max_size;
    opt.verbose = std::getenv("VFB_QUIET") ? 0 : 1;
    if (const char* rt = std::getenv("VFB_RELAXED_TYPE")) opt.relaxed_type = std::atoi(rt);
    // VFB_SEARCH=brute: ltl_nocache's own syntactic search (its formula, its counters)
    if (const char* se = std::getenv("VFB_SEARCH")) opt.search = std::strcmp(se, "brute") == 0 ? VFB_SEARCH_BRUTE : VFB_SEARCH_CS;
    vfb_result res;
    vfb_status st = vfb_solve(inst, &opt, &res);
    if (st == VFB_INVALID_INPUT || st == VFB_DEVICE_ERROR) {
        std::fprintf(stderr, "vfb: %s\n", vfb_status_string(st));
    } else {
        if (st == VFB_CAPACITY) std::fprintf(stderr, "vfb: %s before a formula was found\n", vfb_status_string(st));
        std::printf("%s, %s\n", argv[1], res.formula ? res.formula : "Not found !");
    }
    vfb_result_free(&res);
    vfb_instance_free(inst);
    return 0;
}
