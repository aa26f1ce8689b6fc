// Small C-ABI helpers: options defaults, result ownership, status strings.
#include <cstdlib>
#include <cstring>

#include "vfb_internal.h"

extern "C" void vfb_options_default(vfb_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->max_size = VFB_MAX_SIZE;  // ltl_nocache.cu:11 maxFormulaSize
    o->relaxed_type = 3;         // the hashed key, LTL_BS.cu:78-88
    o->verbose = 0;
    o->timing = 0;
    o->hbm_budget = 0;
}

extern "C" void vfb_result_free(vfb_result* r) {
    if (!r) return;
    std::free(r->formula);
    r->formula = nullptr;
}

extern "C" const char* vfb_status_string(vfb_status s) {
    switch (s) {
    case VFB_OK: return "ok";
    case VFB_NOT_FOUND: return "not found";
    case VFB_INVALID_INPUT: return "invalid input";
    case VFB_CAPACITY: return "capacity exhausted";
    case VFB_DEVICE_ERROR: return "device error";
    case VFB_IO_ERROR: return "i/o error";
    }
    return "unknown";
}
