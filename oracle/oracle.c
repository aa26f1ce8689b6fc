/*
 * oracle.c -- CPU restatement of the reference's LTL characteristic-set (CS)
 * enumeration.  TEST INFRASTRUCTURE ONLY: loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.  Plain C11, single-threaded, written for clarity.
 *
 * Semantics restated (reference paths relative to its repository root):
 *   operators            LTL_BS.cu:183-238 (identical: ltl_nocache.cu:95-150)
 *   key function         LTL_BS.cu:30-110  (exact concatenation, relaxed 1/2/3)
 *   invalid-key merging  warpcore/hash_set.cuh:198-204 + LTL_BS.cu:262-265
 *   level/operator order LTL_BS.cu:1226-1493 (NegType 1, unit costs)
 *   target check         LTL_BS.cu:316-319 (only unique CS are checked)
 *   atoms at level 1     LTL_BS.cu:1127-1163 (every AP stored, Q5)
 *   formula text         ltl_nocache.cu:289-317
 *   brute force          ltl_nocache.cu:78-287, 319-446
 *
 * Deliberate, documented differences from the reference binary:
 *   - Until evaluates every (l, r) pair in both orientations (the reference's
 *     processOperator only evaluates even linear pair indices, LTL_BS.cu:397).
 *   - No all-ones sentinel rows, so traces of length 64 are allowed.
 *   - Traces may be longer than one word (up to ORA_MAX_LEN steps): a trace
 *     of n steps is ceil(n/64) words, word w holding steps 64w..64w+63, and a
 *     CS row is every trace's words in trace order (the reference stops at 64).
 *     The operators carry across words (ora_apply_trace); the exact key and
 *     the relaxed types 1 and 2 take the trace's bits at the same global
 *     positions as the reference's formulas; type 3 mixes word w as
 *     mix64(word + w * 0x9e3779b97f4a7c15), which is the reference's per-trace
 *     mix64(word) for one-word traces.
 *   - Among candidates with equal keys the lowest canonical ordinal survives,
 *     and the returned formula is the lowest-ordinal matching survivor
 *     (the reference takes whichever thread's atomicCAS lands first).
 *   - The row cap truncates the overflowing level in ordinal order after the
 *     whole level was deduplicated and checked (the reference stops deduping
 *     mid-level at a device-dependent point); the next level is check-only.
 *
 * The folding of a type-3 half-key equal to -1 onto -2 (at the end of
 * ora_key, hash_set.cuh:198-204) is pinned by constructed instances whose
 * trace words invert SplitMix64 so that one CS's half-key is -1 and
 * another's is -2 (tests/fold_instance.py, tests/test_oracle.py).
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ------------------------------------------------------------------ */
/* primitive arithmetic                                                */
/* ------------------------------------------------------------------ */

uint64_t ora_mask(int len)
{
    /* (1<<len)-1 of LTL_BS.cu:193, with len==64 giving all ones
       (the reference relies on a 64-bit shift there, ltl_nocache.cu:100) */
    return len >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << len) - 1);
}

static uint64_t suffix_or(uint64_t x)
{
    /* LTL_BS.cu:209-211 */
    x |= x >> 1; x |= x >> 2; x |= x >> 4;
    x |= x >> 8; x |= x >> 16; x |= x >> 32;
    return x;
}

uint64_t ora_apply(int op, uint64_t l, uint64_t r, int len)
{
    switch (op) {
    case ORA_NOT:      return ~l & ora_mask(len);          /* LTL_BS.cu:191-195 */
    case ORA_AND:      return l & r;                        /* :196-199 */
    case ORA_OR:       return l | r;                        /* :200-203 */
    case ORA_NEXT:     return l >> 1;                       /* :204-207 */
    case ORA_FINALLY:  return suffix_or(l);                 /* :208-213 */
    case ORA_GLOBALLY: return l & ~suffix_or(~l & ora_mask(len)); /* :214-221 */
    case ORA_UNTIL: {                                       /* :222-233 */
        r |= l & (r >> 1);  l &= l >> 1;
        r |= l & (r >> 2);  l &= l >> 2;
        r |= l & (r >> 4);  l &= l >> 4;
        r |= l & (r >> 8);  l &= l >> 8;
        r |= l & (r >> 16); l &= l >> 16;
        r |= l & (r >> 32);
        return r;
    }
    default: return 0;
    }
}

int ora_words(int len)
{
    return (len + 63) / 64;
}

/* column layout: trace t's words are columns col[t] .. col[t+1]-1 */
static int col_prefix(const int* len, int T, int* col)
{
    col[0] = 0;
    for (int t = 0; t < T; ++t) col[t + 1] = col[t] + ora_words(len[t]);
    return col[T];
}

/* One operator over the words of one trace of len steps (word 0 = steps
   0..63).  One word is LTL_BS.cu:183-238 as restated in ora_apply.  Longer
   traces run from the last word back, carrying what the later steps decide:
   X takes bit 0 of the next word into bit 63; F of a word is all ones once a
   later word holds a step; G of a word is zero once a later step fails;
   (l U r) at step t of a word is its in-word value or, when l holds from t to
   the word's end, the value at the next word's step 0. */
void ora_apply_trace(int op, const uint64_t* l, const uint64_t* r, int len, uint64_t* out)
{
    const int n = ora_words(len);
    if (n == 1) { out[0] = ora_apply(op, l[0], r ? r[0] : 0, len); return; }
    uint64_t carry = 0;
    int later_any = 0, later_all = 1;
    for (int w = n - 1; w >= 0; --w) {
        const uint64_t m = w == n - 1 ? ora_mask(len - 64 * w) : ~(uint64_t)0;
        const uint64_t a = l[w], b = r ? r[w] : 0;
        uint64_t v = 0;
        switch (op) {
        case ORA_NOT: v = ~a & m; break;
        case ORA_AND: v = a & b; break;
        case ORA_OR: v = a | b; break;
        case ORA_NEXT: v = (a >> 1) | (carry << 63); carry = a & 1; break;
        case ORA_FINALLY: v = later_any ? m : suffix_or(a); later_any |= a != 0; break;
        case ORA_GLOBALLY:
            v = later_all ? (a & ~suffix_or(~a & m)) : 0;
            later_all &= a == m;
            break;
        case ORA_UNTIL:
            v = ora_apply(ORA_UNTIL, a, b, 64) | (carry ? ~suffix_or(~a) : 0);
            carry = v & 1;
            break;
        default: break;
        }
        out[w] = v;
    }
}

/* op over every trace of a row (columns as laid out by col_prefix) */
static void row_apply(int op, const uint64_t* l, const uint64_t* r, const int* len, const int* col, int T,
                      uint64_t* out)
{
    for (int t = 0; t < T; ++t)
        ora_apply_trace(op, l + col[t], r ? r + col[t] : NULL, len[t], out + col[t]);
}

uint64_t ora_mix64(uint64_t x)
{
    /* SplitMix64 output finaliser (Steele, Lea, Flood 2014), LTL_BS.cu:80-83 */
    x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
    x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
    return x ^ (x >> 31);
}

static void put_bit(int j, uint64_t* hi, uint64_t* lo)
{
    if (j < 63) *lo |= (uint64_t)1 << j;
    else if (j < 126) *hi |= (uint64_t)1 << (j - 63);
}

/* bit k (step k) of trace i: word k/64 of its columns */
#define TBIT(i, k) ((cs[col[i] + (k) / 64] >> ((k) % 64)) & 1)

void ora_key(const uint64_t* cs, const int* len, int T, int relaxed_type,
             uint64_t* hi_out, uint64_t* lo_out)
{
    /* LTL_BS.cu:30-110, bit-serial exactly as written there */
    uint64_t hi = 0, lo = 0;
    int lenSum = 0;
    int col[ORA_MAX_TRACES + 1];
    col_prefix(len, T, col);
    for (int i = 0; i < T; ++i) lenSum += len[i];

    if (lenSum > 126) {
        if (relaxed_type == 1) {                 /* :45-59 */
            const int stride = lenSum / 126;
            int j = 0;
            for (int i = 0; i < T && j < 126; ++i)
                for (int k = 0; k < len[i] && j < 126; k += stride, ++j)
                    if (TBIT(i, k)) put_bit(j, &hi, &lo);
        } else if (relaxed_type == 2) {          /* :61-76 */
            int j = 0;
            for (int i = 0; i < T && j < 126; ++i) {
                int maxbits = (126 * len[i] + lenSum) / lenSum;
                /* steps past the trace read as 0 (the reference's shifted-out bitPtr) */
                for (int k = 0; k < maxbits && j < 126; ++k, ++j)
                    if (k < 64 * (col[i + 1] - col[i]) && TBIT(i, k)) put_bit(j, &hi, &lo);
            }
        } else {                                 /* :78-88 (type 3) */
            for (int i = 0; i < T; ++i)
                for (int w = 0; w < col[i + 1] - col[i]; ++w) {
                    uint64_t x = ora_mix64(cs[col[i] + w] + (uint64_t)w * UINT64_C(0x9e3779b97f4a7c15));
                    if (i < T / 2) hi ^= x; else lo ^= x;
                }
        }
    } else {                                     /* :94-106 exact */
        int j = 0;
        for (int i = 0; i < T; ++i)
            for (int k = 0; k < len[i]; ++k, ++j)
                if (TBIT(i, k)) put_bit(j, &hi, &lo);
    }
    /* WarpCore rejects the empty (-1) and tombstone (-2) keys and the caller
       then uses slot id 0 for both (hash_set.cuh:198-204, LTL_BS.cu:262-265),
       so -1 and -2 denote the same half-key.  Fold -1 onto -2. */
    if (hi == ~(uint64_t)0) hi = ~(uint64_t)1;
    if (lo == ~(uint64_t)0) lo = ~(uint64_t)1;
    *hi_out = hi;
    *lo_out = lo;
}
#undef TBIT

/* ------------------------------------------------------------------ */
/* simple open-addressing sets                                         */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t* hi;
    uint64_t* lo;     /* lo == 0 && hi == 0 && !used is tracked by used[] */
    uint8_t* used;
    size_t cap, n;
} keyset;

static uint64_t hash2(uint64_t a, uint64_t b)
{
    return ora_mix64(a ^ ora_mix64(b + UINT64_C(0x9e3779b97f4a7c15)));
}

static int keyset_init(keyset* s, size_t cap)
{
    s->cap = cap; s->n = 0;
    s->hi = (uint64_t*)calloc(cap, 8);
    s->lo = (uint64_t*)calloc(cap, 8);
    s->used = (uint8_t*)calloc(cap, 1);
    return s->hi && s->lo && s->used;
}

static void keyset_free(keyset* s)
{
    free(s->hi); free(s->lo); free(s->used);
    memset(s, 0, sizeof(*s));
}

static int keyset_insert(keyset* s, uint64_t hi, uint64_t lo); /* fwd */

static int keyset_grow(keyset* s)
{
    keyset t;
    if (!keyset_init(&t, s->cap * 2)) return 0;
    for (size_t i = 0; i < s->cap; ++i)
        if (s->used[i]) keyset_insert(&t, s->hi[i], s->lo[i]);
    keyset_free(s);
    *s = t;
    return 1;
}

/* returns 1 if inserted (new), 0 if present, -1 on allocation failure */
static int keyset_insert(keyset* s, uint64_t hi, uint64_t lo)
{
    if (2 * (s->n + 1) > s->cap && !keyset_grow(s)) return -1;
    size_t m = s->cap - 1, i = (size_t)hash2(hi, lo) & m;
    while (s->used[i]) {
        if (s->hi[i] == hi && s->lo[i] == lo) return 0;
        i = (i + 1) & m;
    }
    s->used[i] = 1; s->hi[i] = hi; s->lo[i] = lo; s->n++;
    return 1;
}

/* ------------------------------------------------------------------ */
/* CS enumeration (LTL_BS.cu:1069-1493, Until fixed, lowest ordinal wins) */
/* ------------------------------------------------------------------ */

typedef struct {
    int T;              /* words per row (the columns of all traces) */
    uint64_t* cs;       /* rows*T */
    int32_t* left;
    int32_t* right;
    uint8_t* op;
    size_t rows, cap;
} rowstore;

static int rows_push(rowstore* st, const uint64_t* cs, int op, int l, int r)
{
    if (st->rows == st->cap) {
        size_t nc = st->cap ? st->cap * 2 : 1024;
        uint64_t* a = (uint64_t*)realloc(st->cs, nc * st->T * 8);
        if (!a) return 0;
        st->cs = a;
        int32_t* b = (int32_t*)realloc(st->left, nc * 4);
        if (!b) return 0;
        st->left = b;
        b = (int32_t*)realloc(st->right, nc * 4);
        if (!b) return 0;
        st->right = b;
        uint8_t* c = (uint8_t*)realloc(st->op, nc);
        if (!c) return 0;
        st->op = c;
        st->cap = nc;
    }
    memcpy(st->cs + st->rows * st->T, cs, st->T * 8);
    st->left[st->rows] = l; st->right[st->rows] = r; st->op[st->rows] = (uint8_t)op;
    st->rows++;
    return 1;
}

typedef struct {
    char* buf;
    size_t n, cap;
    int overflow;
} sbuf;

static void sb_put(sbuf* b, const char* s)
{
    size_t k = strlen(s);
    if (b->n + k + 1 > b->cap) { b->overflow = 1; return; }
    memcpy(b->buf + b->n, s, k);
    b->n += k;
    b->buf[b->n] = 0;
}

static const char OPCH[7] = { '~', '&', '|', 'X', 'F', 'G', 'U' };

static void atom_name(int a, const char* const* names, char* tmp)
{
    if (names) snprintf(tmp, 128, "%s", names[a]);
    else snprintf(tmp, 128, "p%d", a);
}

/* ltl_nocache.cu:289-317: atoms by name, unary op(x), binary (l) op (r) */
static void row_to_string(const rowstore* st, int row, const char* const* names, sbuf* b)
{
    char tmp[160];
    int op = st->op[row];
    if (op == ORA_ATOM) { atom_name(st->left[row], names, tmp); sb_put(b, tmp); return; }
    if (op == ORA_AND || op == ORA_OR || op == ORA_UNTIL) {
        sb_put(b, "(");
        row_to_string(st, st->left[row], names, b);
        snprintf(tmp, sizeof tmp, ") %c (", OPCH[op]);
        sb_put(b, tmp);
        row_to_string(st, st->right[row], names, b);
        sb_put(b, ")");
    } else {
        snprintf(tmp, sizeof tmp, "%c(", OPCH[op]);
        sb_put(b, tmp);
        row_to_string(st, st->left[row], names, b);
        sb_put(b, ")");
    }
}

static int matches_target(const uint64_t* cs, const int* col, int T, int P)
{
    /* LTL_BS.cu:316-318: step 0 of every trace (bit 0 of its first word) */
    for (int i = 0; i < P; ++i) if (!(cs[col[i]] & 1)) return 0;
    for (int i = P; i < T; ++i) if (cs[col[i]] & 1) return 0;
    return 1;
}

static uint64_t row_hash(const uint64_t* cs, int T)
{
    uint64_t h = UINT64_C(0x243f6a8885a308d3);
    for (int i = 0; i < T; ++i) h = ora_mix64(h ^ cs[i]) + (uint64_t)i;
    return h;
}

/* full-row set (semantic dedup) keyed by row index into the store */
typedef struct {
    int64_t* idx;
    size_t cap, n;
} rowset;

static int rowset_find_or_add(rowset* s, const rowstore* st, const uint64_t* cs,
                              int64_t new_idx, int add)
{
    int T = st->T;
    if (add && 2 * (s->n + 1) > s->cap) {
        size_t nc = s->cap ? s->cap * 2 : 1024;
        int64_t* t = (int64_t*)malloc(nc * 8);
        if (!t) return -1;
        for (size_t i = 0; i < nc; ++i) t[i] = -1;
        for (size_t i = 0; i < s->cap; ++i) if (s->idx[i] >= 0) {
            size_t j = (size_t)row_hash(st->cs + s->idx[i] * T, T) & (nc - 1);
            while (t[j] >= 0) j = (j + 1) & (nc - 1);
            t[j] = s->idx[i];
        }
        free(s->idx);
        s->idx = t; s->cap = nc;
    }
    if (!s->cap) return 1;
    size_t m = s->cap - 1, j = (size_t)row_hash(cs, T) & m;
    while (s->idx[j] >= 0) {
        if (memcmp(st->cs + s->idx[j] * T, cs, T * 8) == 0) return 0;
        j = (j + 1) & m;
    }
    if (add) { s->idx[j] = new_idx; s->n++; }
    return 1;
}

typedef struct {
    const ora_instance* in;
    const int* col;      /* column prefix of the traces */
    int key_mode;
    uint64_t max_rows;
    const char* const* names;
    rowstore st;
    keyset ks;
    rowset rs;
    ora_result* out;
    int level;
    uint64_t local_ord;
    int done;            /* 1 found, 3 capacity, -1 alloc failure */
    int check_only;      /* onTheFly level: target check only */
    int max_size;
    int otf;             /* rows were dropped at the cap */
} enum_ctx;

static void set_found(enum_ctx* c, const uint64_t* cs, int op, int l, int r, uint64_t ord)
{
    /* build the formula text from the candidate's parents */
    rowstore* st = &c->st;
    if (!rows_push(st, cs, op, l, r)) { c->done = -1; return; }
    sbuf b = { c->out->formula, 0, sizeof c->out->formula, 0 };
    c->out->formula[0] = 0;
    row_to_string(st, (int)st->rows - 1, c->names, &b);
    st->rows--;
    c->out->min_size = c->level;
    c->out->top_op = op;
    c->out->found_ordinal = ord;
    c->done = 1;
}

/* processes one candidate in canonical order; returns nonzero to stop */
static int candidate(enum_ctx* c, const uint64_t* cs, int op, int l, int r)
{
    const ora_instance* in = c->in;
    int T = in->T;
    uint64_t ord = c->local_ord++;
    c->out->candidates[c->level]++;
    if (c->check_only) {
        /* processOnTheFly, LTL_BS.cu:331-349 */
        if (matches_target(cs, c->col, T, in->P)) { set_found(c, cs, op, l, r, ord); return 1; }
        return 0;
    }
    int fresh;
    if (c->key_mode == ORA_KEY_FULLROW) {
        fresh = rowset_find_or_add(&c->rs, &c->st, cs, 0, 0);
    } else {
        uint64_t hi, lo;
        ora_key(cs, in->len, T, c->key_mode, &hi, &lo);
        fresh = keyset_insert(&c->ks, hi, lo);
    }
    if (fresh < 0) { c->done = -1; return 1; }
    if (!fresh) return 0;
    if (matches_target(cs, c->col, T, in->P)) { set_found(c, cs, op, l, r, ord); return 1; }
    c->out->distinct[c->level]++;
    if (c->st.rows >= c->max_rows && c->level < c->max_size) {
        /* cache full: the survivor is checked but not stored (LTL_BS.cu:516-522) */
        if (!c->otf) { c->otf = 1; c->out->otf_level = c->level; }
        return 0;
    }
    if (c->key_mode == ORA_KEY_FULLROW) {
        if (rowset_find_or_add(&c->rs, &c->st, cs, (int64_t)c->st.rows, 1) < 0) { c->done = -1; return 1; }
    }
    if (!rows_push(&c->st, cs, op, l, r)) { c->done = -1; return 1; }
    return 0;
}

/* ------------------------------------------------------------------ */
/* BS (Boolean-set) stage: LTL_BS.cu:1508-1525 (transfer of bit 0 of every  */
/* stored CS), 112-133 (dedup of those vectors), 162-181 (~ & | on one    */
/* word), 351-381 and 870-1067 (levels by cost, target = positives mask). */
/* Unit costs.  Q6 (| priced as &) and Q7 (one-block init) are not        */
/* copied; level boundaries are taken after the dedup (the reference      */
/* keeps the pre-dedup LTL start points); the lowest ordinal wins.        */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t* val;
    uint8_t* kind;      /* 0 leaf (a = stored CS row), 1 ~, 2 &, 3 | */
    int64_t* a;
    int64_t* b;
    size_t n, cap;
} bsnodes;

static int bs_push(bsnodes* v, uint64_t val, int kind, int64_t a, int64_t b)
{
    if (v->n == v->cap) {
        size_t nc = v->cap ? v->cap * 2 : 1024;
        uint64_t* x = (uint64_t*)realloc(v->val, nc * 8);
        if (!x) return 0;
        v->val = x;
        uint8_t* k = (uint8_t*)realloc(v->kind, nc);
        if (!k) return 0;
        v->kind = k;
        int64_t* y = (int64_t*)realloc(v->a, nc * 8);
        if (!y) return 0;
        v->a = y;
        y = (int64_t*)realloc(v->b, nc * 8);
        if (!y) return 0;
        v->b = y;
        v->cap = nc;
    }
    v->val[v->n] = val; v->kind[v->n] = (uint8_t)kind; v->a[v->n] = a; v->b[v->n] = b;
    v->n++;
    return 1;
}

static void bs_text(const bsnodes* v, const rowstore* st, int64_t k, const char* const* names, sbuf* b)
{
    switch (v->kind[k]) {
    case 0: row_to_string(st, (int)v->a[k], names, b); return;
    case 1: sb_put(b, "~("); bs_text(v, st, v->a[k], names, b); sb_put(b, ")"); return;
    default:
        sb_put(b, "(");
        bs_text(v, st, v->a[k], names, b);
        sb_put(b, v->kind[k] == 2 ? ") & (" : ") | (");
        bs_text(v, st, v->b[k], names, b);
        sb_put(b, ")");
    }
}

/* BS levels max_size+1 .. bs_max over the rows of levels 1..max_size
   (start[l] = first row of level l); returns 1 if a formula was found */
static int bs_stage(const ora_instance* in, const int* col, const rowstore* st, const size_t* start, int max_size,
                    int bs_max, const char* const* names, ora_result* out)
{
    const int T = in->T, P = in->P, W = st->T;
    const uint64_t tmask = T >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << T) - 1);
    const uint64_t pmask = ((uint64_t)1 << P) - 1;
    bsnodes v;
    memset(&v, 0, sizeof v);
    keyset seen;
    if (!keyset_init(&seen, 1024)) return 0;
    size_t bstart[ORA_MAX_SIZE + 2];
    int found = 0;
    /* the transferred vectors, first occurrence kept, level by level */
    for (int l = 1; l <= max_size; ++l) {
        bstart[l] = v.n;
        for (size_t r = start[l]; r < start[l + 1]; ++r) {
            uint64_t w = 0;
            for (int t = 0; t < T; ++t) w |= (st->cs[r * W + col[t]] & 1) << t;
            if (keyset_insert(&seen, 0, w) == 1) bs_push(&v, w, 0, (int64_t)r, 0);
        }
    }
    bstart[max_size + 1] = v.n;
    for (int s = max_size + 1; s <= bs_max && !found; ++s) {
        uint64_t ord = 0;
        for (int kind = 1; kind <= 3 && !found; ++kind) {
            const int nsplit = kind == 1 ? 1 : (s - 1) / 2;
            for (int i = 1; i <= nsplit && !found; ++i) {
                const int ll = kind == 1 ? s - 1 : i, lr = s - 1 - i;
                const size_t l0 = bstart[ll], l1 = bstart[ll + 1];
                const size_t r0 = kind == 1 ? 0 : bstart[lr], r1 = kind == 1 ? 1 : bstart[lr + 1];
                for (size_t x = l0; x < l1 && !found; ++x)
                    for (size_t y = r0; y < r1 && !found; ++y) {
                        const uint64_t w = kind == 1 ? (~v.val[x] & tmask)
                                         : kind == 2 ? (v.val[x] & v.val[y]) : (v.val[x] | v.val[y]);
                        out->candidates[s]++;
                        const uint64_t o = ord++;
                        if (keyset_insert(&seen, 0, w) != 1) continue;
                        if (w == pmask) {   /* BSInsertInCache, LTL_BS.cu:290-291 */
                            bs_push(&v, w, kind, (int64_t)x, kind == 1 ? 0 : (int64_t)y);
                            sbuf b = { out->formula, 0, sizeof out->formula, 0 };
                            out->formula[0] = 0;
                            bs_text(&v, st, (int64_t)v.n - 1, names, &b);
                            out->min_size = s;
                            out->top_op = kind == 1 ? ORA_NOT : kind == 2 ? ORA_AND : ORA_OR;
                            out->found_ordinal = o;
                            out->bs_stage = 1;
                            found = 1;
                            break;
                        }
                        out->distinct[s]++;
                        bs_push(&v, w, kind, (int64_t)x, kind == 1 ? 0 : (int64_t)y);
                    }
            }
        }
        bstart[s + 1] = v.n;
        if (!found) out->levels_done = s;
    }
    keyset_free(&seen);
    free(v.val); free(v.kind); free(v.a); free(v.b);
    return found;
}

static const int UNIT_COST[8] = { 1, 1, 1, 1, 1, 1, 1, 1 };

static int enumerate_core(const ora_instance* in, int max_size, int key_mode, uint64_t max_rows, int bs_max,
                          const int* cost, int neg_type, const char* const* names, ora_result* out);

int ora_enumerate(const ora_instance* in, int max_size, int key_mode,
                  uint64_t max_rows, const char* const* names, ora_result* out)
{
    return enumerate_core(in, max_size, key_mode, max_rows, 0, UNIT_COST, 1, names, out);
}

int ora_enumerate_bs(const ora_instance* in, int max_size, int key_mode, int bs_max_size,
                     const char* const* names, ora_result* out)
{
    return enumerate_core(in, max_size, key_mode, 0, bs_max_size, UNIT_COST, 1, names, out);
}

int ora_enumerate_costs(const ora_instance* in, int max_cost, int key_mode, const int* cost, int neg_type,
                        const char* const* names, ora_result* out)
{
    for (int k = 0; k < 8; ++k)
        if (cost[k] < 1) { memset(out, 0, sizeof *out); out->status = 2; return 2; }
    if (neg_type != 1 && neg_type != 2) { memset(out, 0, sizeof *out); out->status = 2; return 2; }
    return enumerate_core(in, max_cost, key_mode, 0, 0, cost, neg_type, names, out);
}

static int enumerate_core(const ora_instance* in, int max_size, int key_mode, uint64_t max_rows, int bs_max,
                          const int* cost, int neg_type, const char* const* names, ora_result* out)
{
    memset(out, 0, sizeof *out);
    if (in->T < 1 || in->T > ORA_MAX_TRACES || in->A < 1 || in->A > ORA_MAX_ATOMS ||
        max_size < 1 || max_size > ORA_MAX_SIZE) { out->status = 2; return 2; }
    const int T = in->T;
    for (int t = 0; t < T; ++t)
        if (in->len[t] < 1 || in->len[t] > ORA_MAX_LEN) { out->status = 2; return 2; }
    int col[ORA_MAX_TRACES + 1];
    const int W = col_prefix(in->len, T, col);   /* words per row */
    enum_ctx c;
    memset(&c, 0, sizeof c);
    c.in = in; c.col = col; c.key_mode = key_mode; c.names = names; c.out = out;
    c.max_rows = max_rows ? max_rows : UINT64_MAX;
    c.max_size = max_size;
    c.st.T = W;
    if (key_mode != ORA_KEY_FULLROW && !keyset_init(&c.ks, 1024)) return -1;

    uint64_t* cs = (uint64_t*)malloc(W * 8);
    size_t start[ORA_MAX_SIZE + 2];
    /* level = cost (LTL_BS.cu:1085-1493): p c1, ~ c2, & c3, | c4, X c5, F c6,
       G c7, U c8; unit costs make it the formula size */
    const int c1 = cost[0], c2 = cost[1], c3 = cost[2], c4 = cost[3];
    const int c5 = cost[4], c6 = cost[5], c7 = cost[6], c8 = cost[7];
    const int unit = c1 == 1 && c2 == 1 && c3 == 1 && c4 == 1 && c5 == 1 && c6 == 1 && c7 == 1 && c8 == 1;
    if (c1 > max_size) { out->status = 2; free(cs); keyset_free(&c.ks); return 2; }
    for (int l = 0; l <= c1; ++l) start[l] = 0;

    /* level c1: every AP is stored (LTL_BS.cu:1127-1163) and its key inserted
       (hashSetsInitialisation, LTL_BS.cu:136-160, 1217) */
    c.level = c1;
    for (int a = 0; a < in->A && !c.done; ++a) {
        for (int t = 0; t < T; ++t)
            for (int w = col[t]; w < col[t + 1]; ++w)
                cs[w] = in->atoms[(size_t)a * W + w] & (w + 1 == col[t + 1] ? ora_mask(in->len[t] - 64 * (w - col[t]))
                                                                              : ~(uint64_t)0);
        out->candidates[c1]++;
        c.local_ord++;
        if (matches_target(cs, col, T, in->P)) {
            char tmp[160];
            atom_name(a, names, tmp);
            snprintf(out->formula, sizeof out->formula, "%s", tmp);
            out->min_size = c1; out->top_op = ORA_ATOM; out->found_ordinal = (uint64_t)a;
            c.done = 1;
            break;
        }
        if (key_mode == ORA_KEY_FULLROW) {
            rowset_find_or_add(&c.rs, &c.st, cs, (int64_t)c.st.rows, 1);
        } else {
            uint64_t hi, lo;
            ora_key(cs, in->len, T, key_mode, &hi, &lo);
            keyset_insert(&c.ks, hi, lo);
        }
        rows_push(&c.st, cs, ORA_ATOM, a, 0);
        out->distinct[c1]++;
    }
    if (!c.done) out->levels_done = c1;
    start[c1 + 1] = c.st.rows;

    int last_round = 0;
    for (int s = c1 + 1; s <= max_size && !c.done; ++s) {
        if (c.otf) { last_round = 1; c.check_only = 1; }
        c.level = s;
        c.local_ord = 0;
        const uint64_t* CS = NULL;
#define ROW(i) (c.st.cs + (size_t)(i) * W)
        /* ~ over level s-c2 (LTL_BS.cu:1236-1272; NegType 2: atoms only) */
        if (s - c2 >= c1 && (neg_type == 1 || s - c2 == c1))
            for (size_t x = start[s - c2]; x < start[s - c2 + 1] && !c.done; ++x) {
                CS = ROW(x);
                row_apply(ORA_NOT, CS, NULL, in->len, col, T, cs);
                candidate(&c, cs, ORA_NOT, (int)x, 0);
            }
        /* & then | over splits c1 <= i, 2i <= s - cost, left-major (LTL_BS.cu:1274-1346) */
        for (int op = ORA_AND; op <= ORA_OR && !c.done; ++op) {
            const int co = op == ORA_AND ? c3 : c4;
            for (int i = c1; 2 * i <= s - co && !c.done; ++i) {
                int j = s - i - co;
                for (size_t l = start[i]; l < start[i + 1] && !c.done; ++l)
                    for (size_t r = start[j]; r < start[j + 1] && !c.done; ++r) {
                        const uint64_t* L = ROW(l);
                        const uint64_t* R = ROW(r);
                        row_apply(op, L, R, in->len, col, T, cs);
                        candidate(&c, cs, op, (int)l, (int)r);
                    }
            }
        }
        /* X, F, G over levels s-c5, s-c6, s-c7 (LTL_BS.cu:1348-1451) */
        for (int op = ORA_NEXT; op <= ORA_GLOBALLY && !c.done; ++op) {
            const int lv = s - (op == ORA_NEXT ? c5 : op == ORA_FINALLY ? c6 : c7);
            if (lv < c1) continue;
            for (size_t x = start[lv]; x < start[lv + 1] && !c.done; ++x) {
                CS = ROW(x);
                row_apply(op, CS, NULL, in->len, col, T, cs);
                candidate(&c, cs, op, (int)x, 0);
            }
        }
        /* U over splits, each pair in both orientations (LTL_BS.cu:1453-1488,
           424-442, with every pair evaluated) */
        for (int i = c1; 2 * i <= s - c8 && !c.done; ++i) {
            int j = s - i - c8;
            for (size_t l = start[i]; l < start[i + 1] && !c.done; ++l)
                for (size_t r = start[j]; r < start[j + 1] && !c.done; ++r) {
                    const uint64_t* L = ROW(l);
                    const uint64_t* R = ROW(r);
                    row_apply(ORA_UNTIL, L, R, in->len, col, T, cs);
                    if (candidate(&c, cs, ORA_UNTIL, (int)l, (int)r)) break;
                    L = ROW(l); R = ROW(r);   /* store may have moved */
                    row_apply(ORA_UNTIL, R, L, in->len, col, T, cs);
                    candidate(&c, cs, ORA_UNTIL, (int)r, (int)l);
                }
        }
#undef ROW
        if (!c.done && !last_round) out->levels_done = s;
        start[s + 1] = c.st.rows;
        if (out->otf_level == s) out->otf_stored = start[s + 1] - start[s];
        if (last_round) {
            if (!c.done) c.done = 3;
            break;
        }
    }

    /* nothing up to max_size: Boolean combinations of the stored CS (BS) */
    if (!c.done && !c.otf && unit && bs_max > max_size && bs_max <= ORA_MAX_SIZE && T <= 63 &&
        bs_stage(in, col, &c.st, start, max_size, bs_max, names, out))
        c.done = 1;
    free(cs);
    keyset_free(&c.ks);
    free(c.rs.idx);
    free(c.st.cs); free(c.st.left); free(c.st.right); free(c.st.op);
    if (c.done == -1) { out->status = -1; return -1; }
    if (c.done == 1) { out->status = 0; return 0; }
    if (c.done == 3) { out->status = 3; snprintf(out->formula, sizeof out->formula, "Not found !"); return 3; }
    out->status = 1;
    snprintf(out->formula, sizeof out->formula, "Not found !");
    return 1;
}

/* ------------------------------------------------------------------ */
/* brute-force syntactic enumeration (ltl_nocache.cu:78-287, 319-446)  */
/* ------------------------------------------------------------------ */

void ora_num_formulas(int A, int max_size, uint64_t* nf)
{
    /* generateMatrix, ltl_nocache.cu:248-287 */
    memset(nf, 0, sizeof(uint64_t) * (max_size + 1) * 8);
    if (max_size >= 1) nf[1 * 8 + 7] = (uint64_t)A;
    if (max_size >= 2) {
        nf[2 * 8 + 0] = nf[2 * 8 + 3] = nf[2 * 8 + 4] = nf[2 * 8 + 5] = (uint64_t)A;
        nf[2 * 8 + 7] = 4 * (uint64_t)A;
    }
    for (int i = 3; i <= max_size; ++i) {
        uint64_t un = nf[(i - 1) * 8 + 7], comm = 0, ncomm = 0;
        for (int j = 1; j <= (i - 1) / 2; ++j) comm += nf[j * 8 + 7] * nf[(i - j - 1) * 8 + 7];
        for (int j = 1; j <= i - 2; ++j) ncomm += nf[j * 8 + 7] * nf[(i - j - 1) * 8 + 7];
        nf[i * 8 + 0] = un; nf[i * 8 + 1] = comm; nf[i * 8 + 2] = comm;
        nf[i * 8 + 3] = un; nf[i * 8 + 4] = un; nf[i * 8 + 5] = un;
        nf[i * 8 + 6] = ncomm;
        nf[i * 8 + 7] = 4 * un + 2 * comm + ncomm;
    }
}

/* numberToFormula, ltl_nocache.cu:169-224: tokens 0..A-1 atoms, 100+op ops */
static void number_to_rpn(const uint64_t* nf, uint64_t n, int size, int* f)
{
    static const int opmap[7] = { ORA_NOT, ORA_AND, ORA_OR, ORA_NEXT, ORA_FINALLY, ORA_GLOBALLY, ORA_UNTIL };
    struct { uint64_t n; int size, shift; } stk[ORA_MAX_SIZE + 2];
    int sp = 0;
    stk[0].n = n; stk[0].size = size; stk[0].shift = 0;
    while (sp >= 0) {
        uint64_t m = stk[sp].n; int sz = stk[sp].size, sh = stk[sp].shift;
        sp--;
        if (sz == 1) { f[sh] = (int)m; continue; }
        uint64_t part = 0; int op = 0;
        while (m >= part + nf[sz * 8 + op]) part += nf[sz * 8 + op++];
        m -= part;
        sz--;
        f[sh + sz] = 100 + opmap[op];
        if (op == 0 || op == 3 || op == 4 || op == 5) {
            ++sp; stk[sp].n = m; stk[sp].size = sz; stk[sp].shift = sh;
        } else {
            part = 0; int fi = 1;
            while (m >= part + nf[fi * 8 + 7] * nf[(sz - fi) * 8 + 7]) {
                part += nf[fi * 8 + 7] * nf[(sz - fi) * 8 + 7];
                fi++;
            }
            m -= part;
            uint64_t nr = nf[(sz - fi) * 8 + 7];
            ++sp; stk[sp].n = m / nr; stk[sp].size = fi; stk[sp].shift = sh;
            ++sp; stk[sp].n = m % nr; stk[sp].size = sz - fi; stk[sp].shift = sh + fi;
        }
    }
}

uint64_t ora_eval_rpn(const ora_instance* in, const int* tok, int n, int t)
{
    /* evaluateRPN, ltl_nocache.cu:78-161 */
    uint64_t stk[ORA_MAX_SIZE + 2];
    int sp = -1;
    int len = in->len[t];
    for (int k = 0; k < n; ++k) {
        int x = tok[k];
        if (x < 100) { stk[++sp] = in->atoms[x * in->T + t] & ora_mask(len); continue; }
        int op = x - 100;
        if (op == ORA_AND || op == ORA_OR || op == ORA_UNTIL) {
            uint64_t r = stk[sp--];
            stk[sp] = ora_apply(op, stk[sp], r, len);
        } else {
            stk[sp] = ora_apply(op, stk[sp], 0, len);
        }
    }
    return stk[0];
}

static void rpn_to_string(const int* tok, int n, const char* const* names, char* out, size_t cap)
{
    /* LTLToString, ltl_nocache.cu:289-317 */
    char** stk = (char**)calloc(n + 1, sizeof(char*));
    int sp = -1;
    for (int k = 0; k < n; ++k) {
        int x = tok[k];
        char tmp[160];
        if (x < 100) {
            atom_name(x, names, tmp);
            stk[++sp] = strdup(tmp);
            continue;
        }
        int op = x - 100;
        if (op == ORA_AND || op == ORA_OR || op == ORA_UNTIL) {
            char* r = stk[sp--];
            char* l = stk[sp];
            size_t m = strlen(l) + strlen(r) + 16;
            char* s = (char*)malloc(m);
            snprintf(s, m, "(%s) %c (%s)", l, OPCH[op], r);
            free(l); free(r);
            stk[sp] = s;
        } else {
            char* l = stk[sp];
            size_t m = strlen(l) + 8;
            char* s = (char*)malloc(m);
            snprintf(s, m, "%c(%s)", OPCH[op], l);
            free(l);
            stk[sp] = s;
        }
    }
    snprintf(out, cap, "%s", sp == 0 ? stk[0] : "Error : Incorrect Formula.\n");
    for (int k = 0; k <= sp; ++k) free(stk[k]);
    free(stk);
}

int ora_bruteforce(const ora_instance* in, int max_size, const char* const* names, ora_result* out)
{
    memset(out, 0, sizeof *out);
    if (max_size < 1 || max_size > ORA_MAX_SIZE) { out->status = 2; return 2; }
    int T = in->T;
    for (int t = 0; t < T; ++t)   /* ltl_nocache's traces are one word */
        if (in->len[t] < 1 || in->len[t] > 64) { out->status = 2; return 2; }
    /* atoms first (ltl_nocache.cu:367-378) */
    for (int a = 0; a < in->A; ++a) {
        out->candidates[1]++;
        int ok = 1;
        for (int t = 0; t < T && ok; ++t) {
            int b = (int)(in->atoms[a * T + t] & 1);
            if ((t < in->P) != (b == 1)) ok = 0;
        }
        if (ok) {
            char tmp[160];
            atom_name(a, names, tmp);
            snprintf(out->formula, sizeof out->formula, "%s", tmp);
            out->min_size = 1; out->top_op = ORA_ATOM; out->status = 0;
            return 0;
        }
    }
    uint64_t* nf = (uint64_t*)malloc(sizeof(uint64_t) * (max_size + 1) * 8);
    ora_num_formulas(in->A, max_size, nf);
    int f[ORA_MAX_SIZE + 2];
    for (int s = 2; s <= max_size; ++s) {
        uint64_t total = nf[s * 8 + 7];
        for (uint64_t n = 0; n < total; ++n) {
            number_to_rpn(nf, n, s, f);
            out->candidates[s]++;
            int ok = 1;
            for (int t = 0; t < T && ok; ++t) {
                uint64_t w = ora_eval_rpn(in, f, s, t);
                if ((t < in->P) != ((w & 1) == 1)) ok = 0;
            }
            if (ok) {
                rpn_to_string(f, s, names, out->formula, sizeof out->formula);
                out->min_size = s;
                out->top_op = f[s - 1] - 100;
                out->found_ordinal = n;
                out->status = 0;
                free(nf);
                return 0;
            }
        }
        out->levels_done = s;
    }
    free(nf);
    out->status = 1;
    snprintf(out->formula, sizeof out->formula, "Not found !");
    return 1;
}
