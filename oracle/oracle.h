/*
 * oracle.h -- CPU restatement of the reference's LTL characteristic-set (CS)
 * enumeration, used ONLY as test infrastructure (the checker).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library.  The product path (libvfb.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Pinning (what checks each function
 * against something other than itself) is listed per function; see the
 * tests in tests/test_oracle.py.
 */
#ifndef VFB_ORACLE_H
#define VFB_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_MAX_TRACES 128
#define ORA_MAX_ATOMS 64
#define ORA_MAX_SIZE 32
#define ORA_MAX_LEN 4096    /* steps per trace; one 64-bit word per 64 steps */

/* operator codes, in the reference's segment order (LTL_BS.cu:17) */
enum { ORA_NOT = 0, ORA_AND = 1, ORA_OR = 2, ORA_NEXT = 3, ORA_FINALLY = 4,
       ORA_GLOBALLY = 5, ORA_UNTIL = 6, ORA_ATOM = 7 };

/* dedup key modes */
enum { ORA_KEY_AUTO1 = 1, ORA_KEY_AUTO2 = 2, ORA_KEY_AUTO3 = 3, /* exact if lenSum<=126, else relaxed type n */
       ORA_KEY_FULLROW = 100 /* semantic dedup on the full CS row (no key function) */ };

typedef struct {
    int A;                  /* number of atomic propositions */
    int T;                  /* number of traces (positives first) */
    int P;                  /* number of positive traces */
    const int* len;         /* [T] trace lengths, 1..ORA_MAX_LEN */
    const uint64_t* atoms;  /* [A*W], W = sum over traces of ora_words(len): trace t's
                               words follow those of traces < t; bit k of word w of
                               trace t (atoms[a*W + col_t + w]) = AP a holds at step
                               64w+k.  With every len <= 64: atoms[a*T+t]. */
} ora_instance;

typedef struct {
    int status;             /* 0 found, 1 not found, 3 capacity */
    int min_size;           /* size of the returned formula (0 if none) */
    int top_op;             /* operator code of the returned formula's root */
    int levels_done;        /* highest level whose candidates were all processed */
    uint64_t distinct[ORA_MAX_SIZE + 1];   /* survivors stored per level (found level: before the match) */
    uint64_t candidates[ORA_MAX_SIZE + 1]; /* candidates generated per level (found level: through the match) */
    uint64_t found_ordinal; /* level-local ordinal of the returned formula */
    int otf_level;          /* level whose rows were truncated at max_rows (0: none) */
    int bs_stage;           /* 1: the formula was found by the BS stage */
    uint64_t otf_stored;    /* rows of that level that were stored */
    char formula[8192];     /* ltl_nocache.cu:289-317 syntax */
} ora_result;

/* ---- primitive arithmetic ---- */
uint64_t ora_mask(int len);
uint64_t ora_apply(int op, uint64_t l, uint64_t r, int len);
uint64_t ora_mix64(uint64_t x);
int ora_words(int len);     /* words of a trace of len steps: ceil(len / 64) */
/* op over the ora_words(len) words of one trace (r NULL for unary ops) */
void ora_apply_trace(int op, const uint64_t* l, const uint64_t* r, int len, uint64_t* out);
/* key of a CS row of T traces (their words in trace order, see ora_instance) */
void ora_key(const uint64_t* cs, const int* len, int T, int relaxed_type,
             uint64_t* hi, uint64_t* lo);

/* ---- enumerators ---- */
/* max_rows (0 = unlimited) caps the stored rows, the reference's
   LTLCacheCapacity (LTL_BS.cu:1182): survivors of the level that overflows
   are still checked, rows past the cap are dropped in ordinal order, the next
   level is checked without dedup or storage over the stored rows (onTheFly,
   LTL_BS.cu:331-349, 516-522) and the search stops there (lastRound,
   LTL_BS.cu:1229-1233, 1490-1491) with status 3 if nothing matched. */
int ora_enumerate(const ora_instance* in, int max_size, int key_mode,
                  uint64_t max_rows, const char* const* names, ora_result* out);
/* ora_enumerate without a row cap, then, if nothing matched up to max_size
   and T <= 63, the BS stage over every stored CS up to bs_max_size
   (LTL_BS.cu:808-1067, 1508-1525; unit costs; Q6/Q7 not copied; level
   boundaries after the dedup; the lowest ordinal wins).  BS levels are
   reported in distinct[] / candidates[] at sizes max_size+1 .. */
int ora_enumerate_bs(const ora_instance* in, int max_size, int key_mode, int bs_max_size,
                     const char* const* names, ora_result* out);
/* ora_enumerate with LTL_BS's cost function (LTL_BS.cu:1085-1493): level =
   cost, cost[8] = p ~ & | X F G U (all >= 1), neg_type 1 (~ of any formula)
   or 2 (~ of atoms only); no row cap, no BS stage */
int ora_enumerate_costs(const ora_instance* in, int max_cost, int key_mode, const int* cost, int neg_type,
                        const char* const* names, ora_result* out);
int ora_bruteforce(const ora_instance* in, int max_size,
                   const char* const* names, ora_result* out);
/* multi-core CPU build of ora_enumerate (cpu_omp.c, OpenMP): same results
   (status, min_size, formula, distinct, candidates, found_ordinal) for
   key modes 1-3, no row cap and traces of at most 64 steps; levels with at least pass_min candidates
   (0: none) are built in three passes that stop at a settled match.  It is
   the CPU baseline of the benchmark, pinned against ora_enumerate. */
int ora_enumerate_omp(const ora_instance* in, int max_size, int key_mode, uint64_t pass_min,
                      const char* const* names, ora_result* out);
int ora_omp_threads(void);
/* CPU baseline of a batch (cpu_omp.c): OpenMP dynamic scheduling over the n
   instances, each by ora_enumerate (no row cap); 0, or -1 on a failure */
int ora_enumerate_many_omp(const ora_instance* const* in, const char* const* const* names, int n, int max_size,
                           int key_mode, ora_result* out);

/* closed-form count of syntactic formulas of each size, ltl_nocache.cu:248-287 */
void ora_num_formulas(int A, int max_size, uint64_t* out /* [(max_size+1)*8] */);

/* evaluate an RPN formula (tokens: 0..A-1 atoms, 100+op operators) on trace t
   with the reference's bit-parallel operators; returns the CS word (traces of
   at most 64 steps, as ltl_nocache's; ora_bruteforce returns status 2 for
   longer ones) */
uint64_t ora_eval_rpn(const ora_instance* in, const int* tokens, int n, int t);

#ifdef __cplusplus
}
#endif
#endif
