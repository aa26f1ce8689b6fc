// Internal host-side types shared by the library's translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vfb.h"

// 64-bit words of a trace of len steps (word w holds steps 64w..64w+63)
inline int trace_words(int len) { return (len + 63) / 64; }

namespace vfb {
void dev_rows_free(void* p);   // hipFree of a resident atom-row buffer (engine.cpp)
}

// Device copy of an instance's atom rows (vfb_instance_upload).  A copy or an
// assignment of the instance never carries it: the rows it holds belong to
// the handle they were uploaded for.
struct DevRows {
    void* p = nullptr;
    size_t words = 0;
    DevRows() = default;
    DevRows(const DevRows&) {}
    DevRows& operator=(const DevRows&) {
        release();
        return *this;
    }
    ~DevRows() { release(); }
    void release() {
        if (p) vfb::dev_rows_free(p);
        p = nullptr;
        words = 0;
    }
};

struct vfb_instance {
    std::vector<std::string> names;   // AP names, in input order
    std::vector<int> lens;            // per trace, positives first
    std::vector<uint64_t> atoms;      // [a * C() + cols()[t] + w]: word w of trace t
                                      // (= [a * T + t] when every trace has <= 64 steps)
    int P = 0;                        // positive traces
    std::string planted;              // generating formula (generated instances)
    uint64_t seed = 0;                // generating seed
    int default_max_size = 0;         // config default (generated instances)
    DevRows dev;                      // resident atom rows, if uploaded
    int A() const { return (int)names.size(); }
    int T() const { return (int)lens.size(); }
    // words per row: the traces' words in trace order
    int C() const {
        int c = 0;
        for (int n : lens) c += trace_words(n);
        return c;
    }
    // first word of each trace, and C() last ([T + 1])
    std::vector<int> cols() const {
        std::vector<int> c(lens.size() + 1, 0);
        for (size_t t = 0; t < lens.size(); ++t) c[t + 1] = c[t] + trace_words(lens[t]);
        return c;
    }
};

namespace vfb {

// Validates limits; returns VFB_OK or VFB_INVALID_INPUT with a message.
vfb_status validate(const vfb_instance& in, std::string* why);

// ltl_nocache's syntactic brute-force search on the GPU (brute.hip)
vfb_status brute_solve(const vfb_instance& in, int max_size, int verbose, vfb_result* R);

inline uint64_t len_mask(int len) {
    return len >= 64 ? ~uint64_t(0) : ((uint64_t(1) << len) - 1);
}

// Bit-parallel operator semantics (LTL_BS.cu:183-238), host side: used for
// the atom check, planted-formula labels and formula re-verification.
uint64_t host_apply(int op, uint64_t l, uint64_t r, int len);

// Evaluates RPN tokens (atoms 0..A-1, operators 100+op) on trace t.
uint64_t eval_rpn(const vfb_instance& in, const std::vector<int>& rpn, int t);

// Renders RPN in ltl_nocache.cu:289-317 syntax.
std::string rpn_to_string(const std::vector<int>& rpn, const std::vector<std::string>& names);

// Parses the library's own formula syntax back to RPN; false on error.
bool parse_formula(const std::string& s, const std::vector<std::string>& names,
                   std::vector<int>* rpn);

}  // namespace vfb
