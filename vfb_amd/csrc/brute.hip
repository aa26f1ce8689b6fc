// brute.hip -- the syntactic brute-force search of ltl_nocache on the GPU.
//
// ltl_nocache numbers the formulas of each size (generateMatrix,
// ltl_nocache.cu:248-287; numberToFormula, 169-224), evaluates each one from
// scratch on every trace (evaluateRPN, 78-161) and returns the first, in
// numbering order, that holds at time 0 on every positive trace and on no
// negative one (LTL, 319-446).  Here one lane decodes and evaluates one
// formula number; a launch covers a chunk of numbers, keeps the lowest
// matching one with atomicMin, and the host stops after the first chunk
// holding one, so the result is the reference's first match.

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/vfb.h"
#include "vfb_internal.h"

namespace vfb {

constexpr int kBruteMaxT = 128;
constexpr uint64_t kNone = ~0ull;

// per-instance data of the search, by value in constant-like global memory
struct BruteDesc {
    int A, T, P, size;
    uint64_t lenmask[kBruteMaxT];
    uint64_t nf[(VFB_MAX_SIZE + 1) * 8];   // formulas of each size per operator (ltl_nocache.cu:248-287)
};

__device__ __forceinline__ uint64_t b_suffix_or(uint64_t x) {
    x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}

__device__ __forceinline__ uint64_t b_until(uint64_t l, uint64_t r) {
    r |= l & (r >> 1);  l &= l >> 1;
    r |= l & (r >> 2);  l &= l >> 2;
    r |= l & (r >> 4);  l &= l >> 4;
    r |= l & (r >> 8);  l &= l >> 8;
    r |= l & (r >> 16); l &= l >> 16;
    r |= l & (r >> 32);
    return r;
}

// numberToFormula: formula number n of the given size into postfix tokens
// (0..A-1 atoms; 100 + vfb_op operators); an explicit work stack replaces
// the reference's recursion
__device__ void decode(const BruteDesc& d, uint64_t n, int size, int* f) {
    const int opmap[7] = {VFB_OP_NOT, VFB_OP_AND, VFB_OP_OR, VFB_OP_NEXT, VFB_OP_FINALLY, VFB_OP_GLOBALLY,
                          VFB_OP_UNTIL};
    uint64_t sn[VFB_MAX_SIZE + 2];
    int ssz[VFB_MAX_SIZE + 2], ssh[VFB_MAX_SIZE + 2];
    int sp = 0;
    sn[0] = n; ssz[0] = size; ssh[0] = 0;
    while (sp >= 0) {
        uint64_t m = sn[sp];
        int sz = ssz[sp];
        const int sh = ssh[sp];
        --sp;
        if (sz == 1) { f[sh] = (int)m; continue; }
        uint64_t part = 0;
        int op = 0;
        while (m >= part + d.nf[sz * 8 + op]) part += d.nf[sz * 8 + op++];
        m -= part;
        --sz;
        f[sh + sz] = 100 + opmap[op];
        if (op == 0 || op == 3 || op == 4 || op == 5) {
            ++sp; sn[sp] = m; ssz[sp] = sz; ssh[sp] = sh;
        } else {
            part = 0;
            int fi = 1;
            while (m >= part + d.nf[fi * 8 + 7] * d.nf[(sz - fi) * 8 + 7]) {
                part += d.nf[fi * 8 + 7] * d.nf[(sz - fi) * 8 + 7];
                ++fi;
            }
            m -= part;
            const uint64_t nr = d.nf[(sz - fi) * 8 + 7];
            ++sp; sn[sp] = m / nr; ssz[sp] = fi; ssh[sp] = sh;
            ++sp; sn[sp] = m % nr; ssz[sp] = sz - fi; ssh[sp] = sh + fi;
        }
    }
}

// evaluateRPN on trace t (atoms[a * T + t], masked to the trace length)
__device__ uint64_t eval_rpn(const BruteDesc& d, const uint64_t* __restrict__ atoms, const int* f, int n, int t) {
    uint64_t stk[VFB_MAX_SIZE + 2];
    int sp = -1;
    const uint64_t m = d.lenmask[t];
    for (int k = 0; k < n; ++k) {
        const int x = f[k];
        if (x < 100) { stk[++sp] = atoms[(uint64_t)x * d.T + t] & m; continue; }
        const int op = x - 100;
        if (op == VFB_OP_AND || op == VFB_OP_OR || op == VFB_OP_UNTIL) {
            const uint64_t r = stk[sp--];
            const uint64_t l = stk[sp];
            stk[sp] = op == VFB_OP_AND ? (l & r) : op == VFB_OP_OR ? (l | r) : b_until(l, r);
        } else {
            const uint64_t a = stk[sp];
            stk[sp] = op == VFB_OP_NOT ? (~a & m)
                    : op == VFB_OP_NEXT ? (a >> 1)
                    : op == VFB_OP_FINALLY ? b_suffix_or(a)
                    : (a & ~b_suffix_or(~a & m));
        }
    }
    return stk[0];
}

__global__ __launch_bounds__(256) void k_brute(const BruteDesc* __restrict__ D, const uint64_t* __restrict__ atoms,
                                               uint64_t begin, uint64_t count, unsigned long long* found) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const BruteDesc& d = *D;
    const uint64_t n = begin + i;
    int f[VFB_MAX_SIZE + 2];
    decode(d, n, d.size, f);
    bool ok = true;
    for (int t = 0; t < d.T && ok; ++t) {
        const uint64_t w = eval_rpn(d, atoms, f, d.size, t);
        ok = ((w & 1ull) != 0) == (t < d.P);
    }
    if (ok) atomicMin(found, (unsigned long long)n);
}

// host copies of the numbering, for the returned formula's text
static void host_decode(const std::vector<uint64_t>& nf, uint64_t n, int size, std::vector<int>& f) {
    static const int opmap[7] = {VFB_OP_NOT, VFB_OP_AND, VFB_OP_OR, VFB_OP_NEXT, VFB_OP_FINALLY, VFB_OP_GLOBALLY,
                                 VFB_OP_UNTIL};
    f.assign(size, 0);
    struct Item { uint64_t n; int size, shift; };
    std::vector<Item> stk{{n, size, 0}};
    while (!stk.empty()) {
        Item it = stk.back();
        stk.pop_back();
        if (it.size == 1) { f[it.shift] = (int)it.n; continue; }
        uint64_t m = it.n, part = 0;
        int op = 0, sz = it.size;
        while (m >= part + nf[sz * 8 + op]) part += nf[sz * 8 + op++];
        m -= part;
        --sz;
        f[it.shift + sz] = 100 + opmap[op];
        if (op == 0 || op == 3 || op == 4 || op == 5) {
            stk.push_back({m, sz, it.shift});
        } else {
            part = 0;
            int fi = 1;
            while (m >= part + nf[fi * 8 + 7] * nf[(sz - fi) * 8 + 7]) {
                part += nf[fi * 8 + 7] * nf[(sz - fi) * 8 + 7];
                ++fi;
            }
            m -= part;
            const uint64_t nr = nf[(sz - fi) * 8 + 7];
            stk.push_back({m / nr, fi, it.shift});
            stk.push_back({m % nr, sz - fi, it.shift + fi});
        }
    }
}

// LTLToString (ltl_nocache.cu:289-317)
static std::string rpn_text(const vfb_instance& in, const std::vector<int>& f) {
    static const char opch[7] = {'~', '&', '|', 'X', 'F', 'G', 'U'};
    std::vector<std::string> stk;
    for (int x : f) {
        if (x < 100) { stk.push_back(in.names[x]); continue; }
        const int op = x - 100;
        if (op == VFB_OP_AND || op == VFB_OP_OR || op == VFB_OP_UNTIL) {
            std::string r = stk.back();
            stk.pop_back();
            stk.back() = "(" + stk.back() + ") " + opch[op] + " (" + r + ")";
        } else {
            stk.back() = std::string(1, opch[op]) + "(" + stk.back() + ")";
        }
    }
    return stk.size() == 1 ? stk[0] : std::string("Error : Incorrect Formula.");
}

// generateMatrix (ltl_nocache.cu:248-287), saturating instead of wrapping
static std::vector<uint64_t> num_formulas(int A, int max_size) {
    std::vector<uint64_t> nf((size_t)(max_size + 1) * 8, 0);
    auto mul = [](uint64_t a, uint64_t b) { return (b && a > ~0ull / b) ? ~0ull : a * b; };
    auto add = [](uint64_t a, uint64_t b) { return a > ~0ull - b ? ~0ull : a + b; };
    if (max_size >= 1) nf[1 * 8 + 7] = (uint64_t)A;
    if (max_size >= 2) {
        nf[2 * 8 + 0] = nf[2 * 8 + 3] = nf[2 * 8 + 4] = nf[2 * 8 + 5] = (uint64_t)A;
        nf[2 * 8 + 7] = 4 * (uint64_t)A;
    }
    for (int i = 3; i <= max_size; ++i) {
        const uint64_t un = nf[(i - 1) * 8 + 7];
        uint64_t comm = 0, ncomm = 0;
        for (int j = 1; j <= (i - 1) / 2; ++j) comm = add(comm, mul(nf[j * 8 + 7], nf[(i - j - 1) * 8 + 7]));
        for (int j = 1; j <= i - 2; ++j) ncomm = add(ncomm, mul(nf[j * 8 + 7], nf[(i - j - 1) * 8 + 7]));
        nf[i * 8 + 0] = un; nf[i * 8 + 1] = comm; nf[i * 8 + 2] = comm;
        nf[i * 8 + 3] = un; nf[i * 8 + 4] = un; nf[i * 8 + 5] = un;
        nf[i * 8 + 6] = ncomm;
        nf[i * 8 + 7] = add(add(mul(4, un), mul(2, comm)), ncomm);
    }
    return nf;
}

#define BCK(x)                                                                 \
    do {                                                                       \
        if ((x) != hipSuccess) { ok = false; goto done; }                      \
    } while (0)

// The search of LTL(...) (ltl_nocache.cu:319-446) for sizes 1..max_size.
// VFB_CAPACITY if the count of some size overflows 64 bits before a match.
vfb_status brute_solve(const vfb_instance& in, int max_size, int verbose, vfb_result* R) {
    const int A = in.A(), T = in.T();
    for (int t = 0; t < T; ++t)
        if (in.lens[t] > 64) {   // ltl_nocache's traces are one word (ltl_nocache.cu:12)
            std::fprintf(stderr, "vfb: the brute-force search takes traces of at most 64 steps\n");
            R->status = VFB_INVALID_INPUT;
            return VFB_INVALID_INPUT;
        }
    // atoms first (ltl_nocache.cu:367-378)
    for (int a = 0; a < A; ++a) {
        R->candidates[1]++;
        bool match = true;
        for (int t = 0; t < T && match; ++t) match = ((in.atoms[(size_t)a * T + t] & 1) != 0) == (t < in.P);
        if (match) {
            R->min_size = 1;
            R->top_op = VFB_OP_ATOM;
            R->found_ordinal = (uint64_t)a;
            R->formula = strdup(in.names[a].c_str());
            R->status = VFB_OK;
            return VFB_OK;
        }
    }
    R->levels_done = 1;
    const std::vector<uint64_t> nf = num_formulas(A, max_size);
    BruteDesc h{};
    h.A = A; h.T = T; h.P = in.P;
    for (int t = 0; t < T; ++t) h.lenmask[t] = len_mask(in.lens[t]);
    std::memcpy(h.nf, nf.data(), nf.size() * 8);
    BruteDesc* dD = nullptr;
    uint64_t* dA = nullptr;
    unsigned long long* dF = nullptr;
    bool ok = true;
    vfb_status st = VFB_NOT_FOUND;
    constexpr uint64_t kChunk = 1ull << 22;
    BCK(hipMalloc((void**)&dD, sizeof(BruteDesc)));
    BCK(hipMalloc((void**)&dA, in.atoms.size() * 8));
    BCK(hipMalloc((void**)&dF, 8));
    BCK(hipMemcpy(dA, in.atoms.data(), in.atoms.size() * 8, hipMemcpyHostToDevice));
    for (int s = 2; s <= max_size && st == VFB_NOT_FOUND; ++s) {
        const uint64_t total = nf[s * 8 + 7];
        if (total == ~0ull) { st = VFB_CAPACITY; break; }
        h.size = s;
        BCK(hipMemcpy(dD, &h, sizeof(BruteDesc), hipMemcpyHostToDevice));
        if (verbose) std::fprintf(stderr, "Size %-2d | formulas: %llu\n", s, (unsigned long long)total);
        for (uint64_t b = 0; b < total; b += kChunk) {
            const uint64_t cnt = std::min<uint64_t>(kChunk, total - b);
            unsigned long long none = kNone, f = kNone;
            BCK(hipMemcpy(dF, &none, 8, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_brute, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, 0, dD, dA, b, cnt, dF);
            BCK(hipGetLastError());
            BCK(hipMemcpy(&f, dF, 8, hipMemcpyDeviceToHost));
            if (f != kNone) {
                std::vector<int> tok;
                host_decode(nf, f, s, tok);
                R->candidates[s] += f - b + 1;
                R->min_size = s;
                R->top_op = tok[s - 1] - 100;
                R->found_ordinal = f;
                R->formula = strdup(rpn_text(in, tok).c_str());
                st = VFB_OK;
                break;
            }
            R->candidates[s] += cnt;
        }
        if (st == VFB_NOT_FOUND) R->levels_done = s;
    }
done:
    if (dD) (void)hipFree(dD);
    if (dA) (void)hipFree(dA);
    if (dF) (void)hipFree(dF);
    if (!ok) return VFB_DEVICE_ERROR;
    R->status = st;
    return st;
}

}  // namespace vfb
