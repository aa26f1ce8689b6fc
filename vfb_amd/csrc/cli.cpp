// vfb <json_file> [max_formula_size]
//
// Drop-in for the reference binary's process contract (ltl_nocache.cu:448-487):
// reads the JSON trace file, prints "<file>, <formula>" on stdout (or
// "<file>, Not found !"), progress on stderr, and exits 0 like the reference
// (errors are reported on stderr).  The size bound is optional here and
// defaults to 32, the reference's maximum (ltl_nocache.cu:11).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/vfb.h"

int main(int argc, char** argv) {
    if (argc < 2 || argc > 3) {
        std::fprintf(stderr, "Arguments should be in the form of\n");
        std::fprintf(stderr, "-----------------------------------------------------------------\n");
        std::fprintf(stderr, "%s <input_file_address> [max_formula_size]\n", argv[0]);
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
