/*
 * cpu_omp.c -- multi-core CPU build of the CS enumeration, used ONLY as the
 * CPU baseline of bench.py and checked against ora_enumerate() by the CPU
 * tests.  It is test/measurement infrastructure like the rest of oracle/:
 * the product path (libvfb.so) never links or calls it.
 *
 * Same arithmetic as ora_enumerate (oracle.c): operators ora_apply
 * (LTL_BS.cu:183-238), keys ora_key (LTL_BS.cu:30-110), canonical candidate
 * order (LTL_BS.cu:1236-1488, every U pair in both orientations), the lowest
 * ordinal of each key wins and only winners are checked against the target
 * (LTL_BS.cu:251-269, 299-330).  The parallel form:
 *   - a level's candidates are split into blocks of ordinals that OpenMP
 *     threads take dynamically; each candidate's key is inserted into one
 *     shared open-addressing set whose slots hold (hi, lo, ordinal): hi and lo
 *     are claimed once by compare-and-swap, the ordinal only decreases;
 *   - candidates that held their key's lowest ordinal when they inserted it
 *     are listed per thread; after the level, those whose ordinal is still
 *     the slot's are the winners (a bitmap), and the lowest matching winner
 *     is the returned formula;
 *   - large levels run in three passes in canonical order (~ X F G, & |, U)
 *     and stop after the pass that settles a match, as the GPU engine does;
 *   - winners are stored in ordinal order by a prefix count over the bitmap.
 * Row caps (onTheFly) are not supported here: max_rows must be 0.
 */
#define _POSIX_C_SOURCE 200809L
#include "oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EMPTY UINT64_MAX
#define MATCH_BIT (UINT64_C(1) << 63)

typedef struct {
    uint64_t hi, lo, ord;
} oslot;

static uint64_t slot_hash(uint64_t hi, uint64_t lo)
{
    return ora_mix64(hi ^ ora_mix64(lo + UINT64_C(0x9e3779b97f4a7c15)));
}

/* 1: this ordinal is (for now) the key's lowest of the level; 0: an older
   key or a lower ordinal holds it; -1: no free slot.  *claimed counts the
   slots this call took from EMPTY. */
static int set_insert(oslot* tab, uint64_t mask, uint64_t hi, uint64_t lo, uint64_t gord, uint64_t* slot_out,
                      uint64_t* claimed)
{
    uint64_t i = slot_hash(hi, lo) & mask;
    for (uint64_t n = 0; n <= mask; ++n, i = (i + 1) & mask) {
        oslot* s = &tab[i];
        uint64_t h = __atomic_load_n(&s->hi, __ATOMIC_ACQUIRE);
        if (h == EMPTY) {
            uint64_t e = EMPTY;
            if (__atomic_compare_exchange_n(&s->hi, &e, hi, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
                h = hi;
                ++*claimed;
            } else {
                h = e;
            }
        }
        if (h != hi) continue;
        uint64_t l = __atomic_load_n(&s->lo, __ATOMIC_ACQUIRE);
        if (l == EMPTY) {
            uint64_t e = EMPTY;
            if (__atomic_compare_exchange_n(&s->lo, &e, lo, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) l = lo;
            else l = e;
        }
        if (l != lo) continue;
        *slot_out = i;
        uint64_t o = __atomic_load_n(&s->ord, __ATOMIC_ACQUIRE);
        while (o > gord) {
            if (__atomic_compare_exchange_n(&s->ord, &o, gord, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) return 1;
        }
        return 0;
    }
    return -1;
}

typedef struct {
    int op;
    int lvl_l, lvl_r;
    uint64_t nl, nr;     /* rows of the operand levels (nr = 1 for unary) */
    uint64_t base, count;
} unit;

typedef struct {
    uint64_t* ord;       /* global ordinal | MATCH_BIT */
    uint64_t* slot;
    size_t n, cap;
} flist;

static int flist_push(flist* f, uint64_t ord, uint64_t slot)
{
    if (f->n == f->cap) {
        size_t nc = f->cap ? f->cap * 2 : 4096;
        uint64_t* a = (uint64_t*)realloc(f->ord, nc * 8);
        if (!a) return 0;
        f->ord = a;
        a = (uint64_t*)realloc(f->slot, nc * 8);
        if (!a) return 0;
        f->slot = a;
        f->cap = nc;
    }
    f->ord[f->n] = ord;
    f->slot[f->n] = slot;
    f->n++;
    return 1;
}

typedef struct {
    const ora_instance* in;
    int key_mode;
    uint64_t* rows[ORA_MAX_SIZE + 2];    /* per level: nrows x T words, row-major */
    uint8_t* pop[ORA_MAX_SIZE + 2];      /* per row: operator */
    uint32_t* pl[ORA_MAX_SIZE + 2];      /* left parent row (in its level) */
    uint32_t* pr[ORA_MAX_SIZE + 2];      /* right parent row */
    uint8_t* plv[ORA_MAX_SIZE + 2];      /* levels of the parents */
    uint8_t* prv[ORA_MAX_SIZE + 2];
    uint64_t nrows[ORA_MAX_SIZE + 2];
    unit units[64];
    int nunits;
} octx;

static int matches(const uint64_t* cs, int T, int P)
{
    for (int t = 0; t < T; ++t)
        if ((int)(cs[t] & 1) != (t < P)) return 0;
    return 1;
}

/* CS of candidate x of unit u, and its parents */
static void cand_cs(const octx* c, const unit* u, uint64_t x, uint64_t* cs, uint32_t* l, uint32_t* r, int* swap)
{
    const int T = c->in->T;
    const int* len = c->in->len;
    *swap = 0;
    if (u->op == ORA_NOT || u->op == ORA_NEXT || u->op == ORA_FINALLY || u->op == ORA_GLOBALLY) {
        const uint64_t* a = c->rows[u->lvl_l] + x * T;
        for (int t = 0; t < T; ++t) cs[t] = ora_apply(u->op, a[t], 0, len[t]);
        *l = (uint32_t)x;
        *r = 0;
        return;
    }
    uint64_t p = u->op == ORA_UNTIL ? x >> 1 : x;
    uint64_t li = p / u->nr, ri = p % u->nr;
    const uint64_t* a = c->rows[u->lvl_l] + li * T;
    const uint64_t* b = c->rows[u->lvl_r] + ri * T;
    if (u->op == ORA_UNTIL && (x & 1)) {
        for (int t = 0; t < T; ++t) cs[t] = ora_apply(ORA_UNTIL, b[t], a[t], len[t]);
        *swap = 1;
    } else {
        for (int t = 0; t < T; ++t) cs[t] = ora_apply(u->op, a[t], b[t], len[t]);
    }
    *l = (uint32_t)li;
    *r = (uint32_t)ri;
}

static int unit_of(const octx* c, uint64_t o)
{
    int k = 0;
    for (int q = 0; q < c->nunits; ++q)
        if (c->units[q].count && c->units[q].base <= o) k = q;
    return k;
}

typedef struct {
    char* p;
    size_t n, cap;
} sb;

static void sb_put(sb* b, const char* s)
{
    size_t k = strlen(s);
    if (b->n + k + 1 > b->cap) k = b->cap > b->n + 1 ? b->cap - b->n - 1 : 0;
    memcpy(b->p + b->n, s, k);
    b->n += k;
    b->p[b->n] = 0;
}

static const char OPC[7] = {'~', '&', '|', 'X', 'F', 'G', 'U'};

/* ltl_nocache.cu:289-317 syntax */
static void node_text(const octx* c, int lvl, uint64_t row, const char* const* names, sb* b)
{
    char tmp[160];
    if (lvl == 1) {
        if (names) sb_put(b, names[row]);
        else { snprintf(tmp, sizeof tmp, "p%llu", (unsigned long long)row); sb_put(b, tmp); }
        return;
    }
    const int op = c->pop[lvl][row];
    if (op == ORA_NOT || op == ORA_NEXT || op == ORA_FINALLY || op == ORA_GLOBALLY) {
        tmp[0] = OPC[op]; tmp[1] = '('; tmp[2] = 0;
        sb_put(b, tmp);
        node_text(c, c->plv[lvl][row], c->pl[lvl][row], names, b);
        sb_put(b, ")");
        return;
    }
    sb_put(b, "(");
    node_text(c, c->plv[lvl][row], c->pl[lvl][row], names, b);
    snprintf(tmp, sizeof tmp, ") %c (", OPC[op]);
    sb_put(b, tmp);
    node_text(c, c->prv[lvl][row], c->pr[lvl][row], names, b);
    sb_put(b, ")");
}

static void cand_text(const octx* c, uint64_t o, const char* const* names, sb* b, int* top)
{
    const unit* u = &c->units[unit_of(c, o)];
    uint64_t cs[ORA_MAX_TRACES];
    uint32_t l, r;
    int swap;
    cand_cs(c, u, o - u->base, cs, &l, &r, &swap);
    *top = u->op;
    char tmp[8];
    if (u->op == ORA_NOT || u->op == ORA_NEXT || u->op == ORA_FINALLY || u->op == ORA_GLOBALLY) {
        tmp[0] = OPC[u->op]; tmp[1] = '('; tmp[2] = 0;
        sb_put(b, tmp);
        node_text(c, u->lvl_l, l, names, b);
        sb_put(b, ")");
        return;
    }
    int la = u->lvl_l, ra = u->lvl_r;
    uint32_t lr = l, rr = r;
    if (swap) { la = u->lvl_r; ra = u->lvl_l; lr = r; rr = l; }
    sb_put(b, "(");
    node_text(c, la, lr, names, b);
    snprintf(tmp, sizeof tmp, ") %c (", OPC[u->op]);
    sb_put(b, tmp);
    node_text(c, ra, rr, names, b);
    sb_put(b, ")");
}

/* a set of new_cap slots holding the keys of `old` with ordinal < keep_below */
static oslot* set_rebuild(const oslot* old, uint64_t old_cap, uint64_t new_cap, uint64_t keep_below)
{
    oslot* nt = (oslot*)malloc(new_cap * sizeof(oslot));
    if (!nt) return NULL;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < new_cap; ++i) nt[i].hi = nt[i].lo = nt[i].ord = EMPTY;
    int fail = 0;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < old_cap; ++i)
        if (old[i].hi != EMPTY && old[i].lo != EMPTY && old[i].ord < keep_below) {
            uint64_t slot, cl = 0;
            if (set_insert(nt, new_cap - 1, old[i].hi, old[i].lo, old[i].ord, &slot, &cl) < 0)
                __atomic_store_n(&fail, 1, __ATOMIC_RELAXED);
        }
    if (fail) { free(nt); return NULL; }
    return nt;
}

static uint64_t next_pow2(uint64_t x)
{
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

/* pass membership of unit q (nsplit binary splits): 1 ~XFG, 2 &|, 3 U */
static int unit_pass(int q, int nsplit)
{
    if (q == 0 || (q >= 2 * nsplit + 1 && q <= 2 * nsplit + 3)) return 1;
    if (q >= 1 && q <= 2 * nsplit) return 2;
    return 3;
}

int ora_enumerate_omp(const ora_instance* in, int max_size, int key_mode, uint64_t pass_min,
                      const char* const* names, ora_result* out)
{
    memset(out, 0, sizeof *out);
    if (in->T < 1 || in->T > ORA_MAX_TRACES || in->A < 1 || in->A > ORA_MAX_ATOMS || max_size < 1 ||
        max_size > ORA_MAX_SIZE || key_mode == ORA_KEY_FULLROW) {
        out->status = 2;
        return 2;
    }
    for (int t = 0; t < in->T; ++t)   /* one word per trace */
        if (in->len[t] < 1 || in->len[t] > 64) { out->status = 2; return 2; }
    const int T = in->T, A = in->A, P = in->P;
    octx c;
    memset(&c, 0, sizeof c);
    c.in = in;
    c.key_mode = key_mode;
    int status = 1;
    uint64_t cap = 1 << 12;
    oslot* tab = NULL;
    uint64_t committed = 0;     /* keys in the set */
    double win_ratio = 1.0;     /* new keys / candidates of the last level */
    uint64_t gbase = 0;         /* global ordinal of the next level's first candidate */
    const int nth = omp_get_max_threads();
    flist* fl = (flist*)calloc((size_t)nth, sizeof(flist));
    uint64_t* winbits = NULL;
    uint64_t* wordrank = NULL;

    /* level 1: the atoms, all stored, keys inserted with ordinals 0..A-1 */
    c.rows[1] = (uint64_t*)malloc((size_t)A * T * 8);
    c.pop[1] = (uint8_t*)calloc((size_t)A, 1);
    c.nrows[1] = (uint64_t)A;
    tab = (oslot*)malloc(cap * sizeof(oslot));
    memset(tab, 0xff, cap * sizeof(oslot));
    for (int a = 0; a < A; ++a) {
        uint64_t* cs = c.rows[1] + (size_t)a * T;
        for (int t = 0; t < T; ++t) cs[t] = in->atoms[(size_t)a * T + t] & ora_mask(in->len[t]);
        out->candidates[1]++;
        if (matches(cs, T, P)) {
            snprintf(out->formula, sizeof out->formula, "%s", names ? names[a] : "p");
            out->min_size = 1;
            out->top_op = ORA_ATOM;
            out->found_ordinal = (uint64_t)a;
            status = 0;
            goto done;
        }
        uint64_t hi, lo, slot, cl = 0;
        ora_key(cs, in->len, T, key_mode, &hi, &lo);
        (void)set_insert(tab, cap - 1, hi, lo, (uint64_t)a, &slot, &cl);
        out->distinct[1]++;
    }
    committed = (uint64_t)A;
    gbase = (uint64_t)A;
    out->levels_done = 1;

    for (int s = 2; s <= max_size; ++s) {
        /* units in canonical order */
        const int nsplit = (s - 1) / 2;
        c.nunits = 0;
        uint64_t N = 0;
#define ADD(OP, LL, LR)                                                                     \
    do {                                                                                    \
        unit* u = &c.units[c.nunits++];                                                     \
        u->op = (OP); u->lvl_l = (LL); u->lvl_r = (LR);                                     \
        u->nl = c.nrows[LL]; u->nr = (LR) ? c.nrows[LR] : 1;                                \
        u->base = N;                                                                        \
        u->count = u->nl * u->nr * ((OP) == ORA_UNTIL ? 2 : 1);                             \
        N += u->count;                                                                      \
    } while (0)
        ADD(ORA_NOT, s - 1, 0);
        for (int i = 1; i <= nsplit; ++i) ADD(ORA_AND, i, s - 1 - i);
        for (int i = 1; i <= nsplit; ++i) ADD(ORA_OR, i, s - 1 - i);
        ADD(ORA_NEXT, s - 1, 0);
        ADD(ORA_FINALLY, s - 1, 0);
        ADD(ORA_GLOBALLY, s - 1, 0);
        for (int i = 1; i <= nsplit; ++i) ADD(ORA_UNTIL, i, s - 1 - i);
#undef ADD
        /* the set is sized from the previous level's share of new keys and
           grown (the level restarted) if this level claims more slots */
        uint64_t est = (uint64_t)((double)N * (win_ratio * 1.5 < 1.0 ? win_ratio * 1.5 : 1.0)) + 4096;
        if (est > N + 16) est = N + 16;
        uint64_t need = next_pow2((uint64_t)((double)(committed + est) / 0.6) + 16);
        const uint64_t nwords = (N + 63) / 64;
        const int passes = (pass_min && N >= pass_min && nsplit >= 1) ? 3 : 1;
        uint64_t found = EMPTY;
        int fail = 0;
        for (int attempt = 0; attempt < 8; ++attempt) {
            if (need > (UINT64_C(1) << 31)) { fail = 1; break; }   /* 48 GiB of slots */
            if (need > cap || attempt > 0) {
                oslot* nt = set_rebuild(tab, cap, need, gbase);
                if (!nt) { fail = 1; break; }
                free(tab);
                tab = nt;
                cap = need;
            }
            free(winbits);
            winbits = (uint64_t*)calloc(nwords + 1, 8);
            for (int k = 0; k < nth; ++k) fl[k].n = 0;
            const uint64_t limit = cap * 8 / 10 > committed ? cap * 8 / 10 - committed : 0;
            uint64_t claims = 0;
            fail = 0;
            found = EMPTY;
            for (int pass = passes == 1 ? 0 : 1; pass <= (passes == 1 ? 0 : 3); ++pass) {
                /* blocks of ordinals of this pass's units */
                enum { BLK = 4096 };
                uint64_t nb = 0;
                for (int q = 0; q < c.nunits; ++q)
                    if (pass == 0 || unit_pass(q, nsplit) == pass) nb += (c.units[q].count + BLK - 1) / BLK;
                uint64_t* bq = (uint64_t*)malloc((nb + 1) * 8 * 2);
                uint64_t m = 0;
                for (int q = 0; q < c.nunits; ++q)
                    if (pass == 0 || unit_pass(q, nsplit) == pass)
                        for (uint64_t x = 0; x < c.units[q].count; x += BLK) {
                            bq[2 * m] = (uint64_t)q;
                            bq[2 * m + 1] = x;
                            ++m;
                        }
#pragma omp parallel
                {
                    const int me = omp_get_thread_num();
                    uint64_t cs[ORA_MAX_TRACES];
#pragma omp for schedule(dynamic, 1)
                    for (uint64_t b = 0; b < m; ++b) {
                        if (__atomic_load_n(&fail, __ATOMIC_RELAXED)) continue;
                        const unit* u = &c.units[bq[2 * b]];
                        const uint64_t x0 = bq[2 * b + 1];
                        const uint64_t x1 = x0 + BLK < u->count ? x0 + BLK : u->count;
                        uint64_t cl = 0;
                        for (uint64_t x = x0; x < x1; ++x) {
                            uint32_t l, r;
                            int sw;
                            cand_cs(&c, u, x, cs, &l, &r, &sw);
                            uint64_t hi, lo, slot;
                            ora_key(cs, in->len, T, key_mode, &hi, &lo);
                            const uint64_t o = u->base + x;
                            const int f = set_insert(tab, cap - 1, hi, lo, gbase + o, &slot, &cl);
                            if (f < 0) { __atomic_store_n(&fail, 2, __ATOMIC_RELAXED); break; }
                            if (f && !flist_push(&fl[me], (gbase + o) | (matches(cs, T, P) ? MATCH_BIT : 0), slot))
                                __atomic_store_n(&fail, 1, __ATOMIC_RELAXED);
                        }
                        if (cl && __atomic_add_fetch(&claims, cl, __ATOMIC_RELAXED) > limit)
                            __atomic_store_n(&fail, 2, __ATOMIC_RELAXED);
                    }
                }
                free(bq);
                if (fail) break;
                /* a match below the evaluated prefix is settled */
                const uint64_t below = pass == 1 ? c.units[1].base
                                     : pass == 2 ? c.units[2 * nsplit + 4].base : EMPTY;
                for (int k = 0; k < nth; ++k) {
                    const flist* f = &fl[k];
                    uint64_t fk = EMPTY;
#pragma omp parallel for reduction(min : fk) schedule(static)
                    for (size_t e = 0; e < f->n; ++e) {
                        const uint64_t of = f->ord[e], o = of & ~MATCH_BIT;
                        if ((of & MATCH_BIT) && tab[f->slot[e]].ord == o && o - gbase < below && o - gbase < fk)
                            fk = o - gbase;
                    }
                    if (fk < found) found = fk;
                }
                if (found != EMPTY) break;
            }
            if (fail != 2) break;
            need = cap * 2;                /* set too small: grow, build the level again */
        }
        if (fail) { status = 3; break; }
        /* winners: fresh entries still holding their slot's ordinal */
        for (int k = 0; k < nth; ++k) {
            const flist* f = &fl[k];
#pragma omp parallel for schedule(static)
            for (size_t e = 0; e < f->n; ++e) {
                const uint64_t o = f->ord[e] & ~MATCH_BIT;
                if (tab[f->slot[e]].ord == o) {
                    const uint64_t x = o - gbase;
                    __atomic_fetch_or(&winbits[x >> 6], UINT64_C(1) << (x & 63), __ATOMIC_RELAXED);
                }
            }
        }
        free(wordrank);
        wordrank = (uint64_t*)malloc((nwords + 1) * 8);
        uint64_t W = 0;
        for (uint64_t w = 0; w < nwords; ++w) { wordrank[w] = W; W += (uint64_t)__builtin_popcountll(winbits[w]); }
        if (found != EMPTY) {
            const uint64_t fw = found >> 6;
            out->distinct[s] = wordrank[fw] + (uint64_t)__builtin_popcountll(winbits[fw] & ((UINT64_C(1) << (found & 63)) - 1));
            out->candidates[s] = found + 1;
            out->min_size = s;
            out->found_ordinal = found;
            sb b = {out->formula, 0, sizeof out->formula};
            out->formula[0] = 0;
            cand_text(&c, found, names, &b, &out->top_op);
            status = 0;
            break;
        }
        out->candidates[s] = N;
        out->distinct[s] = W;
        out->levels_done = s;
        if (s == max_size) break;
        /* store the winners in ordinal order */
        c.rows[s] = (uint64_t*)malloc((W ? W : 1) * (size_t)T * 8);
        c.pop[s] = (uint8_t*)malloc(W ? W : 1);
        c.pl[s] = (uint32_t*)malloc((W ? W : 1) * 4);
        c.pr[s] = (uint32_t*)malloc((W ? W : 1) * 4);
        c.plv[s] = (uint8_t*)malloc(W ? W : 1);
        c.prv[s] = (uint8_t*)malloc(W ? W : 1);
        if (!c.rows[s] || !c.pl[s] || !c.pr[s]) { status = 3; break; }
#pragma omp parallel for schedule(dynamic, 64)
        for (uint64_t w = 0; w < nwords; ++w) {
            uint64_t bits = winbits[w], k = wordrank[w];
            while (bits) {
                const int j = __builtin_ctzll(bits);
                bits &= bits - 1;
                const uint64_t o = w * 64 + (uint64_t)j;
                const unit* u = &c.units[unit_of(&c, o)];
                uint32_t l, r;
                int sw;
                cand_cs(&c, u, o - u->base, c.rows[s] + k * T, &l, &r, &sw);
                c.pop[s][k] = (uint8_t)u->op;
                c.pl[s][k] = sw ? r : l;
                c.pr[s][k] = sw ? l : r;
                c.plv[s][k] = (uint8_t)(sw ? u->lvl_r : u->lvl_l);
                c.prv[s][k] = (uint8_t)(sw ? u->lvl_l : u->lvl_r);
                ++k;
            }
        }
        c.nrows[s] = W;
        committed += W;
        win_ratio = N ? (double)W / (double)N : 1.0;
        gbase += N;
    }
done:
    if (status != 0) snprintf(out->formula, sizeof out->formula, "Not found !");
    out->status = status;
    for (int l = 0; l < ORA_MAX_SIZE + 2; ++l) {
        free(c.rows[l]); free(c.pop[l]); free(c.pl[l]); free(c.pr[l]); free(c.plv[l]); free(c.prv[l]);
    }
    for (int k = 0; k < nth; ++k) { free(fl[k].ord); free(fl[k].slot); }
    free(fl);
    free(tab);
    free(winbits);
    free(wordrank);
    return status;
}

int ora_omp_threads(void) { return omp_get_max_threads(); }

/* CPU baseline of a batch: n independent instances, OpenMP dynamic
   scheduling over instances, each solved by the sequential ora_enumerate
   (no row cap).  Results as ora_enumerate's. */
int ora_enumerate_many_omp(const ora_instance* const* in, const char* const* const* names, int n, int max_size,
                           int key_mode, ora_result* out)
{
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int i = 0; i < n; ++i)
        bad |= ora_enumerate(in[i], max_size, key_mode, 0, names[i], &out[i]) < 0;
    return bad ? -1 : 0;
}
