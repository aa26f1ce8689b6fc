// Instances: the reference JSON trace format, bit-packed construction,
// validation, and the seeded synthetic generator.
//
// JSON schema and AP encoding follow ltl_nocache.cu:35-76: the AP names list
// fixes the atom order, a trace's length is the length of its value arrays,
// positive traces come first.  The schema is fixed, so a small reader of our
// own replaces nlohmann/json.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "vfb_internal.h"

namespace vfb {

// ---------------------------------------------------------------------------
// operator semantics (host side), LTL_BS.cu:183-238
// ---------------------------------------------------------------------------

static inline uint64_t suffix_or(uint64_t x) {
    x |= x >> 1; x |= x >> 2; x |= x >> 4;
    x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}

uint64_t host_apply(int op, uint64_t l, uint64_t r, int len) {
    switch (op) {
    case VFB_OP_NOT: return ~l & len_mask(len);
    case VFB_OP_AND: return l & r;
    case VFB_OP_OR: return l | r;
    case VFB_OP_NEXT: return l >> 1;
    case VFB_OP_FINALLY: return suffix_or(l);
    case VFB_OP_GLOBALLY: return l & ~suffix_or(~l & len_mask(len));
    case VFB_OP_UNTIL:
        for (int k = 1; k <= 32; k <<= 1) {
            r |= l & (r >> k);
            if (k < 32) l &= l >> k;
        }
        return r;
    }
    return 0;
}

uint64_t eval_rpn(const vfb_instance& in, const std::vector<int>& rpn, int t) {
    std::vector<uint64_t> st;
    const int len = in.lens[t];
    for (int x : rpn) {
        if (x < 100) { st.push_back(in.atoms[(size_t)x * in.T() + t] & len_mask(len)); continue; }
        const int op = x - 100;
        if (op == VFB_OP_AND || op == VFB_OP_OR || op == VFB_OP_UNTIL) {
            uint64_t r = st.back(); st.pop_back();
            st.back() = host_apply(op, st.back(), r, len);
        } else {
            st.back() = host_apply(op, st.back(), 0, len);
        }
    }
    return st.empty() ? 0 : st.back();
}

static const char kOpCh[7] = {'~', '&', '|', 'X', 'F', 'G', 'U'};

std::string rpn_to_string(const std::vector<int>& rpn, const std::vector<std::string>& names) {
    std::vector<std::string> st;
    for (int x : rpn) {
        if (x < 100) { st.push_back(names[x]); continue; }
        const int op = x - 100;
        if (op == VFB_OP_AND || op == VFB_OP_OR || op == VFB_OP_UNTIL) {
            std::string r = st.back(); st.pop_back();
            st.back() = "(" + st.back() + ") " + kOpCh[op] + " (" + r + ")";
        } else {
            st.back() = std::string(1, kOpCh[op]) + "(" + st.back() + ")";
        }
    }
    return st.size() == 1 ? st[0] : std::string("Error : Incorrect Formula.\n");
}

namespace {
struct FormulaParser {
    const std::string& s;
    const std::vector<std::string>& names;
    size_t i = 0;
    std::vector<int>* out;
    bool expect(const char* lit) {
        size_t n = std::strlen(lit);
        if (s.compare(i, n, lit) != 0) return false;
        i += n;
        return true;
    }
    bool parse() {
        if (i >= s.size()) return false;
        if (s[i] == '(') {
            ++i;
            if (!parse()) return false;
            if (!expect(") ")) return false;
            if (i >= s.size()) return false;
            char c = s[i++];
            int op = c == '&' ? VFB_OP_AND : c == '|' ? VFB_OP_OR : c == 'U' ? VFB_OP_UNTIL : -1;
            if (op < 0 || !expect(" (") || !parse() || !expect(")")) return false;
            out->push_back(100 + op);
            return true;
        }
        if (i + 1 < s.size() && s[i + 1] == '(' && std::strchr("~XFG", s[i])) {
            char c = s[i];
            i += 2;
            int op = c == '~' ? VFB_OP_NOT : c == 'X' ? VFB_OP_NEXT : c == 'F' ? VFB_OP_FINALLY : VFB_OP_GLOBALLY;
            if (!parse() || !expect(")")) return false;
            out->push_back(100 + op);
            return true;
        }
        int best = -1;
        size_t blen = 0;
        for (size_t a = 0; a < names.size(); ++a)
            if (names[a].size() > blen && s.compare(i, names[a].size(), names[a]) == 0) {
                best = (int)a;
                blen = names[a].size();
            }
        if (best < 0) return false;
        i += blen;
        out->push_back(best);
        return true;
    }
};
}  // namespace

bool parse_formula(const std::string& s, const std::vector<std::string>& names, std::vector<int>* rpn) {
    rpn->clear();
    FormulaParser p{s, names, 0, rpn};
    return p.parse() && p.i == s.size();
}

vfb_status validate(const vfb_instance& in, std::string* why) {
    char buf[256];
    if (in.A() < 1 || in.A() > VFB_MAX_ATOMS) {
        std::snprintf(buf, sizeof buf, "supports 1..%d atomic propositions, got %d", VFB_MAX_ATOMS, in.A());
        *why = buf;
        return VFB_INVALID_INPUT;
    }
    if (in.T() < 1 || in.T() > VFB_MAX_TRACES) {
        std::snprintf(buf, sizeof buf, "supports 1..%d traces, got %d", VFB_MAX_TRACES, in.T());
        *why = buf;
        return VFB_INVALID_INPUT;
    }
    if (in.P < 0 || in.P > in.T()) { *why = "bad positive trace count"; return VFB_INVALID_INPUT; }
    for (int t = 0; t < in.T(); ++t)
        if (in.lens[t] < 1 || in.lens[t] > VFB_MAX_LEN) {
            std::snprintf(buf, sizeof buf, "trace lengths must be 1..%d, trace %d has %d", VFB_MAX_LEN, t, in.lens[t]);
            *why = buf;
            return VFB_INVALID_INPUT;
        }
    if (in.atoms.size() != (size_t)in.A() * in.C()) { *why = "atom table size mismatch"; return VFB_INVALID_INPUT; }
    return VFB_OK;
}

// ---------------------------------------------------------------------------
// minimal JSON reader for the fixed schema
// ---------------------------------------------------------------------------

namespace {
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } k = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal* get(const std::string& key) const {
        for (auto& kv : obj) if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    const char* e;
    std::string err;
    void ws() { while (p < e && std::isspace((unsigned char)*p)) ++p; }
    bool str(std::string* out) {
        if (p >= e || *p != '"') return false;
        ++p;
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) return false;
                char c = *p++;
                switch (c) {
                case 'n': out->push_back('\n'); break;
                case 't': out->push_back('\t'); break;
                case 'r': out->push_back('\r'); break;
                case 'b': out->push_back('\b'); break;
                case 'f': out->push_back('\f'); break;
                case 'u': {
                    if (e - p < 4) return false;
                    unsigned v = (unsigned)std::strtoul(std::string(p, p + 4).c_str(), nullptr, 16);
                    p += 4;
                    if (v < 0x80) out->push_back((char)v);
                    else if (v < 0x800) { out->push_back((char)(0xC0 | (v >> 6))); out->push_back((char)(0x80 | (v & 0x3F))); }
                    else { out->push_back((char)(0xE0 | (v >> 12))); out->push_back((char)(0x80 | ((v >> 6) & 0x3F))); out->push_back((char)(0x80 | (v & 0x3F))); }
                    break;
                }
                default: out->push_back(c);
                }
            } else {
                out->push_back(*p++);
            }
        }
        if (p >= e) return false;
        ++p;
        return true;
    }
    bool val(JVal* v, int depth) {
        if (depth > 64) return false;
        ws();
        if (p >= e) return false;
        char c = *p;
        if (c == '{') {
            v->k = JVal::Obj;
            ++p; ws();
            if (p < e && *p == '}') { ++p; return true; }
            for (;;) {
                ws();
                std::string key;
                if (!str(&key)) return false;
                ws();
                if (p >= e || *p != ':') return false;
                ++p;
                v->obj.emplace_back(key, JVal());
                if (!val(&v->obj.back().second, depth + 1)) return false;
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; return true; }
                return false;
            }
        }
        if (c == '[') {
            v->k = JVal::Arr;
            ++p; ws();
            if (p < e && *p == ']') { ++p; return true; }
            for (;;) {
                v->arr.emplace_back();
                if (!val(&v->arr.back(), depth + 1)) return false;
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; return true; }
                return false;
            }
        }
        if (c == '"') { v->k = JVal::Str; return str(&v->str); }
        if (e - p >= 4 && std::strncmp(p, "true", 4) == 0) { v->k = JVal::Bool; v->num = 1; p += 4; return true; }
        if (e - p >= 5 && std::strncmp(p, "false", 5) == 0) { v->k = JVal::Bool; v->num = 0; p += 5; return true; }
        if (e - p >= 4 && std::strncmp(p, "null", 4) == 0) { v->k = JVal::Null; p += 4; return true; }
        char* end = nullptr;
        std::string tmp(p, std::min<size_t>(64, e - p));
        double d = std::strtod(tmp.c_str(), &end);
        if (end == tmp.c_str()) return false;
        v->k = JVal::Num;
        v->num = d;
        p += end - tmp.c_str();
        return true;
    }
};

bool json_int(const JVal* v, long long* out) {
    if (!v || (v->k != JVal::Num && v->k != JVal::Bool)) return false;
    *out = (long long)v->num;
    return true;
}
}  // namespace

}  // namespace vfb

using namespace vfb;

extern "C" vfb_status vfb_instance_load_json(const char* path, vfb_instance** out) {
    if (!path || !out) return VFB_INVALID_INPUT;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "Error: Could not open file %s\n", path);
        return VFB_IO_ERROR;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    JParser jp{text.data(), text.data() + text.size(), ""};
    JVal root;
    if (!jp.val(&root, 0) || root.k != JVal::Obj) {
        std::fprintf(stderr, "Error: %s is not valid JSON\n", path);
        return VFB_INVALID_INPUT;
    }
    const JVal* aps = root.get("atomic_propositions");
    const JVal* pos = root.get("positive_traces");
    const JVal* neg = root.get("negative_traces");
    if (!aps || aps->k != JVal::Arr || !pos || pos->k != JVal::Arr || !neg || neg->k != JVal::Arr) {
        std::fprintf(stderr, "Error: %s lacks atomic_propositions / positive_traces / negative_traces\n", path);
        return VFB_INVALID_INPUT;
    }
    auto inst = std::make_unique<vfb_instance>();
    for (auto& a : aps->arr) {
        if (a.k != JVal::Str) return VFB_INVALID_INPUT;
        inst->names.push_back(a.str);
    }
    const int A = inst->A();
    if (A < 1 || A > VFB_MAX_ATOMS) {
        std::fprintf(stderr, "This version supports at most %d variables.\n", VFB_MAX_ATOMS);
        return VFB_INVALID_INPUT;
    }
    std::vector<const JVal*> traces;
    for (auto& t : pos->arr) traces.push_back(&t);
    for (auto& t : neg->arr) traces.push_back(&t);
    inst->P = (int)pos->arr.size();
    const int T = (int)traces.size();
    if (T < 1 || T > VFB_MAX_TRACES) {
        std::fprintf(stderr, "This version supports at most %d samples.\n", VFB_MAX_TRACES);
        return VFB_INVALID_INPUT;
    }
    inst->lens.assign(T, 0);
    for (int t = 0; t < T; ++t) {
        const JVal* tr = traces[t];
        if (tr->k != JVal::Obj) return VFB_INVALID_INPUT;
        int len = -1;
        for (int a = 0; a < A; ++a) {
            const JVal* vals = tr->get(inst->names[a]);
            if (!vals || vals->k != JVal::Arr) {
                std::fprintf(stderr, "Error: trace %d lacks values for %s\n", t, inst->names[a].c_str());
                return VFB_INVALID_INPUT;
            }
            const int n = (int)vals->arr.size();
            if (len < 0) len = n;
            if (n != len || n < 1 || n > VFB_MAX_LEN) {
                std::fprintf(stderr, "This version supports traces of size 1..%d with equal lengths per AP.\n", VFB_MAX_LEN);
                return VFB_INVALID_INPUT;
            }
        }
        inst->lens[t] = len;
    }
    // a trace of n steps is ceil(n / 64) words, step k in bit k % 64 of word k / 64
    const int C = inst->C();
    const std::vector<int> col = inst->cols();
    inst->atoms.assign((size_t)A * C, 0);
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < A; ++a) {
            const JVal* vals = traces[t]->get(inst->names[a]);
            for (int k = 0; k < inst->lens[t]; ++k) {
                long long b;
                if (!json_int(&vals->arr[k], &b)) return VFB_INVALID_INPUT;
                if (b) inst->atoms[(size_t)a * C + col[t] + k / 64] |= uint64_t(1) << (k % 64);
            }
        }
    // the count fields are redundant; when present they must agree
    long long v;
    if (json_int(root.get("number_atomic_propositions"), &v) && v != A) return VFB_INVALID_INPUT;
    if (json_int(root.get("number_traces"), &v) && v != T) return VFB_INVALID_INPUT;
    if (json_int(root.get("number_positive_traces"), &v) && v != inst->P) return VFB_INVALID_INPUT;
    const JVal* gf = root.get("generating_formula");
    if (gf && gf->k == JVal::Str) inst->planted = gf->str;
    std::string why;
    vfb_status st = validate(*inst, &why);
    if (st != VFB_OK) {
        std::fprintf(stderr, "%s\n", why.c_str());
        return st;
    }
    *out = inst.release();
    return VFB_OK;
}

// LTL_BS's text trace format (readFile, LTL_BS.cu:1535-1607): one trace per
// line, positives, a line "---", negatives, a line "---", and last the
// alphabet line (the letters a-z, in the order of each step's values).  A
// trace is steps separated by ';', a step the 0/1 values of the letters
// separated by ','.  APs are ordered as the sorted letters (the reference
// keeps them in a std::set<char>).
extern "C" vfb_status vfb_instance_load_text(const char* path, vfb_instance** out) {
    if (!path || !out) return VFB_INVALID_INPUT;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "Failed to open the input file.\n");
        return VFB_IO_ERROR;
    }
    std::vector<std::string> lines;
    for (std::string line; std::getline(f, line);) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        lines.push_back(line);
    }
    size_t last = lines.size();
    while (last > 0 && lines[last - 1].empty()) --last;
    if (last == 0) return VFB_INVALID_INPUT;
    std::string alpha;
    for (char c : lines[last - 1])
        if (c >= 'a' && c <= 'z') alpha += c;
    std::string sorted = alpha;
    std::sort(sorted.begin(), sorted.end());
    sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    if (sorted.empty()) return VFB_INVALID_INPUT;
    // a trace: per step, the letters whose value is 1 (as readFile builds it)
    auto parse = [&](const std::string& line) {
        std::vector<std::string> trace;
        std::string token;
        size_t j = 0;
        for (char c : line) {
            if (c == ';') {
                trace.push_back(token);
                token.clear();
                j = 0;
            } else if (c == ',') {
                continue;
            } else {
                if (c == '1' && j < alpha.size()) token += alpha[j];
                ++j;
            }
        }
        trace.push_back(token);
        return trace;
    };
    std::vector<std::vector<std::string>> pos, neg;
    size_t i = 0;
    for (; i < last - 1 && lines[i] != "---"; ++i) pos.push_back(parse(lines[i]));
    for (++i; i < last - 1 && lines[i] != "---"; ++i) neg.push_back(parse(lines[i]));
    auto inst = std::make_unique<vfb_instance>();
    for (char c : sorted) inst->names.push_back(std::string(1, c));
    const int A = (int)sorted.size();
    inst->P = (int)pos.size();
    std::vector<const std::vector<std::string>*> all;
    for (auto& t : pos) all.push_back(&t);
    for (auto& t : neg) all.push_back(&t);
    const int T = (int)all.size();
    for (int t = 0; t < T; ++t) inst->lens.push_back((int)all[t]->size());
    std::string why;
    for (int n : inst->lens)
        if (n < 1 || n > VFB_MAX_LEN) {
            std::fprintf(stderr, "This version supports traces of 1..%d steps.\n", VFB_MAX_LEN);
            return VFB_INVALID_INPUT;
        }
    const int C = inst->C();
    const std::vector<int> col = inst->cols();
    inst->atoms.assign((size_t)A * C, 0);
    for (int t = 0; t < T; ++t) {
        const auto& tr = *all[t];
        for (size_t k = 0; k < tr.size(); ++k)
            for (char c : tr[k]) {
                const int a = (int)(std::lower_bound(sorted.begin(), sorted.end(), c) - sorted.begin());
                inst->atoms[(size_t)a * C + col[t] + k / 64] |= 1ull << (k % 64);
            }
    }
    if (validate(*inst, &why) != VFB_OK) {
        std::fprintf(stderr, "%s\n", why.c_str());
        return VFB_INVALID_INPUT;
    }
    *out = inst.release();
    return VFB_OK;
}

extern "C" vfb_status vfb_instance_from_bits(int num_atoms, const char* const* atom_names, int num_pos,
                                             int num_neg, const int* lens, const uint64_t* atoms,
                                             vfb_instance** out) {
    if (!out || !lens || !atoms || num_atoms < 1 || num_pos < 0 || num_neg < 0) return VFB_INVALID_INPUT;
    *out = nullptr;
    auto inst = std::make_unique<vfb_instance>();
    for (int a = 0; a < num_atoms; ++a)
        inst->names.push_back(atom_names && atom_names[a] ? std::string(atom_names[a]) : "a" + std::to_string(a));
    const int T = num_pos + num_neg;
    if (T < 1 || T > VFB_MAX_TRACES) return VFB_INVALID_INPUT;
    for (int t = 0; t < T; ++t)
        if (lens[t] < 1 || lens[t] > 64) return VFB_INVALID_INPUT;   // one word per trace here
    inst->P = num_pos;
    inst->lens.assign(lens, lens + T);
    inst->atoms.assign(atoms, atoms + (size_t)num_atoms * T);
    std::string why;
    vfb_status st = validate(*inst, &why);
    if (st != VFB_OK) return st;
    for (int a = 0; a < num_atoms; ++a)
        for (int t = 0; t < T; ++t) inst->atoms[(size_t)a * T + t] &= len_mask(inst->lens[t]);
    *out = inst.release();
    return VFB_OK;
}

extern "C" vfb_status vfb_instance_from_words(int num_atoms, const char* const* atom_names, int num_pos,
                                              int num_neg, const int* lens, const uint64_t* words,
                                              vfb_instance** out) {
    if (!out || !lens || !words || num_atoms < 1 || num_pos < 0 || num_neg < 0) return VFB_INVALID_INPUT;
    *out = nullptr;
    const int T = num_pos + num_neg;
    if (T < 1 || T > VFB_MAX_TRACES) return VFB_INVALID_INPUT;
    for (int t = 0; t < T; ++t)
        if (lens[t] < 1 || lens[t] > VFB_MAX_LEN) return VFB_INVALID_INPUT;
    auto inst = std::make_unique<vfb_instance>();
    for (int a = 0; a < num_atoms; ++a)
        inst->names.push_back(atom_names && atom_names[a] ? std::string(atom_names[a]) : "a" + std::to_string(a));
    inst->P = num_pos;
    inst->lens.assign(lens, lens + T);
    const int C = inst->C();
    const std::vector<int> col = inst->cols();
    inst->atoms.assign(words, words + (size_t)num_atoms * C);
    std::string why;
    vfb_status st = validate(*inst, &why);
    if (st != VFB_OK) return st;
    // steps past a trace's end are cleared
    for (int a = 0; a < num_atoms; ++a)
        for (int t = 0; t < T; ++t) {
            const int last = col[t + 1] - 1;
            inst->atoms[(size_t)a * C + last] &= len_mask(inst->lens[t] - 64 * (last - col[t]));
        }
    *out = inst.release();
    return VFB_OK;
}

static void json_escape(std::string* o, const std::string& s) {
    o->push_back('"');
    for (char c : s) {
        if (c == '"' || c == '\\') { o->push_back('\\'); o->push_back(c); }
        else if (c == '\n') *o += "\\n";
        else o->push_back(c);
    }
    o->push_back('"');
}

extern "C" vfb_status vfb_instance_write_json(const vfb_instance* in, const char* path) {
    if (!in || !path) return VFB_INVALID_INPUT;
    std::string o = "{\n";
    const int T = in->T(), C = in->C();
    const std::vector<int> col = in->cols();
    auto traces = [&](int t0, int t1) {
        o += "[";
        for (int t = t0; t < t1; ++t) {
            o += t > t0 ? ",\n    {" : "\n    {";
            for (int a = 0; a < in->A(); ++a) {
                o += a ? ", " : "";
                json_escape(&o, in->names[a]);
                o += ": [";
                for (int k = 0; k < in->lens[t]; ++k) {
                    o += k ? "," : "";
                    o += (in->atoms[(size_t)a * C + col[t] + k / 64] >> (k % 64)) & 1 ? "1" : "0";
                }
                o += "]";
            }
            o += "}";
        }
        o += "\n  ]";
    };
    o += "  \"positive_traces\": ";
    traces(0, in->P);
    o += ",\n  \"negative_traces\": ";
    traces(in->P, T);
    o += ",\n  \"smallest_known_formula\": \"\",\n  \"generating_formula\": ";
    json_escape(&o, in->planted);
    o += ",\n  \"generating_seed\": ";
    json_escape(&o, in->seed ? std::to_string(in->seed) : std::string());
    o += ",\n  \"name\": \"\",\n  \"atomic_propositions\": [";
    for (int a = 0; a < in->A(); ++a) {
        o += a ? ", " : "";
        json_escape(&o, in->names[a]);
    }
    int maxlen = 0;
    for (int l : in->lens) maxlen = std::max(maxlen, l);
    o += "],\n  \"number_atomic_propositions\": " + std::to_string(in->A());
    o += ",\n  \"number_traces\": " + std::to_string(T);
    o += ",\n  \"number_positive_traces\": " + std::to_string(in->P);
    o += ",\n  \"number_negative_traces\": " + std::to_string(T - in->P);
    o += ",\n  \"max_length_traces\": " + std::to_string(maxlen);
    o += ",\n  \"trace_type\": \"finite\"\n}\n";
    std::ofstream f(path, std::ios::binary);
    if (!f) return VFB_IO_ERROR;
    f << o;
    return f.good() ? VFB_OK : VFB_IO_ERROR;
}

extern "C" void vfb_instance_shape(const vfb_instance* in, int* A, int* T, int* P) {
    if (A) *A = in ? in->A() : 0;
    if (T) *T = in ? in->T() : 0;
    if (P) *P = in ? in->P : 0;
}

extern "C" int vfb_instance_words(const vfb_instance* in) { return in ? in->C() : 0; }

extern "C" void vfb_instance_copy(const vfb_instance* in, int* lens, uint64_t* atoms) {
    if (!in) return;
    if (lens) std::copy(in->lens.begin(), in->lens.end(), lens);
    if (atoms) std::copy(in->atoms.begin(), in->atoms.end(), atoms);
}

extern "C" const char* vfb_instance_atom_name(const vfb_instance* in, int a) {
    if (!in || a < 0 || a >= in->A()) return nullptr;
    return in->names[a].c_str();
}

extern "C" const char* vfb_instance_planted(const vfb_instance* in) { return in ? in->planted.c_str() : ""; }
extern "C" int vfb_instance_max_size(const vfb_instance* in) { return in ? in->default_max_size : 0; }
extern "C" void vfb_instance_free(vfb_instance* in) { delete in; }

// ---------------------------------------------------------------------------
// seeded synthetic generator (no reference counterpart)
// ---------------------------------------------------------------------------

namespace {
typedef unsigned __int128 u128;

struct Rng {
    uint64_t s;
    uint64_t next() {  // SplitMix64
        s += 0x9e3779b97f4a7c15ULL;
        uint64_t x = s;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        return x ^ (x >> 31);
    }
    uint64_t below(uint64_t n) { return (uint64_t)(((u128)next() * n) >> 64); }
    int uniform(int a, int b) { return a + (int)below((uint64_t)(b - a + 1)); }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    bool bern(double p) { return unit() < p; }
};

// syntactic formula counts per size and operator (ltl_nocache.cu:248-287),
// in 128-bit arithmetic
std::vector<std::array<u128, 8>> num_forms(int A, int max_size) {
    std::vector<std::array<u128, 8>> nf(max_size + 1);
    for (auto& r : nf) r.fill(0);
    nf[1][7] = A;
    if (max_size >= 2) { nf[2][0] = nf[2][3] = nf[2][4] = nf[2][5] = A; nf[2][7] = 4 * (u128)A; }
    for (int i = 3; i <= max_size; ++i) {
        u128 un = nf[i - 1][7], comm = 0, ncomm = 0;
        for (int j = 1; j <= (i - 1) / 2; ++j) comm += nf[j][7] * nf[i - j - 1][7];
        for (int j = 1; j <= i - 2; ++j) ncomm += nf[j][7] * nf[i - j - 1][7];
        nf[i][0] = nf[i][3] = nf[i][4] = nf[i][5] = un;
        nf[i][1] = nf[i][2] = comm;
        nf[i][6] = ncomm;
        nf[i][7] = 4 * un + 2 * comm + ncomm;
    }
    return nf;
}

// decodes formula number n of the given size (numberToFormula, ltl_nocache.cu:169-224)
std::vector<int> decode_formula(const std::vector<std::array<u128, 8>>& nf, u128 n, int size) {
    static const int opmap[7] = {VFB_OP_NOT, VFB_OP_AND, VFB_OP_OR, VFB_OP_NEXT, VFB_OP_FINALLY, VFB_OP_GLOBALLY, VFB_OP_UNTIL};
    std::vector<int> f(size, 0);
    struct E { u128 n; int size, shift; };
    std::vector<E> st{{n, size, 0}};
    while (!st.empty()) {
        E e = st.back(); st.pop_back();
        if (e.size == 1) { f[e.shift] = (int)e.n; continue; }
        u128 part = 0;
        int op = 0;
        while (e.n >= part + nf[e.size][op]) part += nf[e.size][op++];
        e.n -= part;
        int sz = e.size - 1;
        f[e.shift + sz] = 100 + opmap[op];
        if (op == 0 || op == 3 || op == 4 || op == 5) {
            st.push_back({e.n, sz, e.shift});
        } else {
            u128 p2 = 0;
            int fi = 1;
            while (e.n >= p2 + nf[fi][7] * nf[sz - fi][7]) { p2 += nf[fi][7] * nf[sz - fi][7]; ++fi; }
            e.n -= p2;
            u128 nr = nf[sz - fi][7];
            st.push_back({e.n / nr, fi, e.shift});
            st.push_back({e.n % nr, sz - fi, e.shift + fi});
        }
    }
    return f;
}

struct GenSpec {
    int A, T, P, lmin, lmax, size, max_size;
    double p;
    bool need_until;
    bool fixed_cfg1;
};

GenSpec spec_for(int config, Rng& r) {
    GenSpec g{};
    switch (config) {
    case 1: g = {3, 10, 5, 4, 8, 5, 10, 0.5, false, true}; break;
    case 2: g = {5, 40, 20, 16, 32, 9, 14, 0.3, true, false}; break;
    case 3: g = {10, 63, 32, 48, 64, 11, 18, 0.5, false, false}; break;
    default: {
        g.A = r.uniform(2, 8);
        g.T = r.uniform(8, 63);
        g.P = r.uniform(1, g.T - 1);
        g.lmin = 8; g.lmax = 64;
        g.p = 0.2 + 0.4 * r.unit();
        g.size = r.uniform(5, 13);
        g.max_size = 16;
        g.need_until = false;
        g.fixed_cfg1 = false;
    }
    }
    return g;
}
}  // namespace

extern "C" vfb_status vfb_instance_generate(int config, uint64_t index, vfb_instance** out) {
    if (!out || config < 1 || config > 5) return VFB_INVALID_INPUT;
    *out = nullptr;
    Rng rng{((uint64_t)config << 32) | index};
    const uint64_t seed = rng.s;
    // the served stream (5) draws from the batched mix (4) with its own seeds
    GenSpec g = spec_for(config == 5 ? 4 : config, rng);
    auto inst = std::make_unique<vfb_instance>();
    for (int a = 0; a < g.A; ++a) inst->names.push_back("a" + std::to_string(a));
    inst->default_max_size = g.max_size;
    inst->seed = seed;
    const auto nf = num_forms(g.A, g.size);

    for (int attempt = 0; attempt < 1000; ++attempt) {
        std::vector<int> f;
        if (g.fixed_cfg1) {
            f = {0, 1, 100 + VFB_OP_NEXT, 100 + VFB_OP_AND, 100 + VFB_OP_FINALLY};  // F(a0 & X a1)
        } else {
            for (;;) {
                u128 r = ((u128)rng.next() << 64) | rng.next();
                f = decode_formula(nf, r % nf[g.size][7], g.size);
                bool has_u = false;
                for (int x : f) has_u |= x == 100 + VFB_OP_UNTIL;
                if (!g.need_until || has_u) break;
            }
        }
        // rejection-sample traces into the positive / negative quotas
        std::vector<std::pair<int, std::vector<uint64_t>>> posv, negv;
        vfb_instance one;
        one.names = inst->names;
        one.lens.assign(1, 0);
        one.atoms.assign(g.A, 0);
        for (int draw = 0; draw < 100000 && ((int)posv.size() < g.P || (int)negv.size() < g.T - g.P); ++draw) {
            int len = rng.uniform(g.lmin, g.lmax);
            std::vector<uint64_t> w(g.A, 0);
            for (int a = 0; a < g.A; ++a)
                for (int k = 0; k < len; ++k)
                    if (rng.bern(g.p)) w[a] |= uint64_t(1) << k;
            one.lens[0] = len;
            one.atoms = w;
            bool label = eval_rpn(one, f, 0) & 1;
            if (label && (int)posv.size() < g.P) posv.emplace_back(len, w);
            else if (!label && (int)negv.size() < g.T - g.P) negv.emplace_back(len, w);
        }
        if ((int)posv.size() < g.P || (int)negv.size() < g.T - g.P) continue;
        inst->P = g.P;
        inst->lens.clear();
        for (auto& v : posv) inst->lens.push_back(v.first);
        for (auto& v : negv) inst->lens.push_back(v.first);
        inst->atoms.assign((size_t)g.A * g.T, 0);
        for (int t = 0; t < g.T; ++t) {
            const auto& w = t < g.P ? posv[t].second : negv[t - g.P].second;
            for (int a = 0; a < g.A; ++a) inst->atoms[(size_t)a * g.T + t] = w[a];
        }
        inst->planted = rpn_to_string(f, inst->names);
        *out = inst.release();
        return VFB_OK;
    }
    return VFB_INVALID_INPUT;
}
