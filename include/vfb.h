/*
 * vfb.h -- C-ABI of the MI355X-native LTL synthesizer (libvfb.so).
 *
 * The reference has no library or FFI layer: its boundary is the process
 * (`ltl_nocache <json_file> <maxLen>`, ltl_nocache.cu:448-487) and, inside,
 * one monolithic host function `std::string LTLI(...)` (LTL_BS.cu:1069-1079)
 * that owns every device buffer.  The entry points below are what a binding
 * of that path needs: plain C types, caller-owned handles, status codes in
 * place of the reference's sentinel strings ("see_the_error",
 * LTL_BS.cu:1105-1110; ltl_nocache.cu:326-348, 480-482).
 *
 * Threading: one solve call at a time per process-wide device context is not
 * required; independent calls on different threads are serialised inside
 * the library.  No exception crosses this boundary.
 */
#ifndef VFB_H
#define VFB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFB_MAX_SIZE 32     /* maximum formula size (ltl_nocache.cu:11) */
#define VFB_MAX_TRACES 128  /* ltl_nocache allows 100 (ltl_nocache.cu:10), LTL_BS 63 (LTL_BS.cu:13) */
#define VFB_MAX_ATOMS 64
#define VFB_MAX_LEN 4096    /* steps per trace: ceil(len / 64) 64-bit CS words per trace
                               (the reference stops at one word, ltl_nocache.cu:12) */

typedef enum {
    VFB_OK = 0,             /* formula found */
    VFB_NOT_FOUND = 1,      /* no formula up to max_size ("Not found !", ltl_nocache.cu:439) */
    VFB_INVALID_INPUT = 2,  /* limits exceeded or malformed input (LTL_BS.cu:1105-1110) */
    VFB_CAPACITY = 3,       /* HBM budget, cache or hash set exhausted (never silent) */
    VFB_DEVICE_ERROR = 4,   /* HIP runtime error, or the HIP library is unusable */
    VFB_IO_ERROR = 5        /* file could not be read or written */
} vfb_status;

/* operator codes in the reference's segment order (LTL_BS.cu:17) */
typedef enum {
    VFB_OP_NOT = 0, VFB_OP_AND = 1, VFB_OP_OR = 2, VFB_OP_NEXT = 3,
    VFB_OP_FINALLY = 4, VFB_OP_GLOBALLY = 5, VFB_OP_UNTIL = 6, VFB_OP_ATOM = 7
} vfb_op;

typedef struct vfb_instance vfb_instance;

/* ---- instances --------------------------------------------------------- */

/* Reads the reference's JSON trace format (replaces readJsonFile,
   ltl_nocache.cu:35-76).  Positive traces come first. */
vfb_status vfb_instance_load_json(const char* path, vfb_instance** out);

/* Reads LTL_BS's text trace format (replaces readFile, LTL_BS.cu:1535-1607):
   positive traces, "---", negative traces, "---", then the alphabet line;
   a trace is steps separated by ';', a step the 0/1 values of the alphabet's
   letters separated by ','.  APs are the sorted letters. */
vfb_status vfb_instance_load_text(const char* path, vfb_instance** out);

/* Builds an instance from bit-packed traces of at most 64 steps:
   atoms[a * T + t] bit k is the value of AP a at step k of trace t; traces
   0..num_pos-1 are positive (the layout the reference builds at
   LTL_BS.cu:1127-1163). */
vfb_status vfb_instance_from_bits(int num_atoms, const char* const* atom_names,
                                  int num_pos, int num_neg, const int* lens,
                                  const uint64_t* atoms, vfb_instance** out);

/* The same for traces of any length up to VFB_MAX_LEN: trace t of n steps is
   ceil(n / 64) words, word w holding steps 64w..64w+63 (step k in bit k % 64),
   and a trace's words follow those of the traces before it, so that
   words[a * W + first(t) + w] is word w of AP a on trace t, W = the sum of all
   traces' words (vfb_instance_words).  With every trace at most 64 steps this
   is vfb_instance_from_bits.  No reference counterpart (one word per trace
   there). */
vfb_status vfb_instance_from_words(int num_atoms, const char* const* atom_names,
                                   int num_pos, int num_neg, const int* lens,
                                   const uint64_t* words, vfb_instance** out);

/* Seeded synthetic instance: config 1 plain, 2 dedup stress, 3 limits,
   4 batched mix (5, the served stream, draws from config 4).  Instance k of
   config c uses SplitMix64 seed (c << 32) | k.  No reference counterpart. */
vfb_status vfb_instance_generate(int config, uint64_t index, vfb_instance** out);

/* Writes the reference JSON schema (ExampleLTLF.json) including
   generating_formula / generating_seed for generated instances. */
vfb_status vfb_instance_write_json(const vfb_instance* inst, const char* path);

/* Read-only accessors.  lens: T ints; atoms: A*W words in the layout of
   vfb_instance_from_words (A*T, that of vfb_instance_from_bits, when every
   trace has at most 64 steps). */
void vfb_instance_shape(const vfb_instance* inst, int* num_atoms, int* num_traces,
                        int* num_pos);
int vfb_instance_words(const vfb_instance* inst);   /* W: words of one CS row */
void vfb_instance_copy(const vfb_instance* inst, int* lens, uint64_t* atoms);
const char* vfb_instance_atom_name(const vfb_instance* inst, int a);
const char* vfb_instance_planted(const vfb_instance* inst); /* "" if none */
int vfb_instance_max_size(const vfb_instance* inst);        /* config default, 0 if none */

/* Copies the instance's atom rows (its level-1 CS rows, A*W words) to device
   memory once.  Later vfb_solve calls on this handle read them there instead
   of uploading them per call (the reference uploads them in every LTLI call,
   LTL_BS.cu:1216); so do batches (vfb_solve_batch).  The server works on
   its own copy of a submitted instance and uploads the rows itself.
   The device copy lives until vfb_instance_free and never travels with a copy
   of the instance.  VFB_DEVICE_ERROR without a usable GPU. */
vfb_status vfb_instance_upload(vfb_instance* inst);
void vfb_instance_free(vfb_instance* inst);

/* ---- solving ----------------------------------------------------------- */

typedef struct {
    int max_size;          /* 1..32 (LTLI maxCost, LTL_BS.cu:1071) */
    int relaxed_type;      /* key when total trace length > 126: 1, 2 or 3 (LTL_BS.cu:39-92) */
    int verbose;           /* 1: per-level progress lines on stderr */
    int timing;            /* 1: time each kernel class with HIP events; 2: only the
                              combine+key+dedup kernel (ms_expand; two events per level) */
    uint64_t hbm_budget;   /* bytes of device memory per solve; 0 = the process arena
                              (85% of the HBM free at first use) */
    uint64_t max_rows;     /* cap on stored CS rows, the reference's LTLCacheCapacity
                              (LTL_BS.cu:1182); 0 = limited by hbm_budget only */
    int search;            /* VFB_SEARCH_CS (default): the CS enumeration of LTL_BS;
                              VFB_SEARCH_BRUTE: ltl_nocache's syntactic brute force
                              (formulas in its numbering order, each evaluated from
                              scratch, ltl_nocache.cu:78-287, 319-446): the same
                              formula text as ltl_nocache; candidates[s] counts the
                              formulas checked, distinct[] stays 0 */
    int bs_max_size;       /* > max_size: when nothing matches up to max_size (status
                              VFB_NOT_FOUND, no row cap reached) and T <= 63, the BS stage
                              of LTL_BS (LTL_BS.cu:808-1067, 1508-1525) searches Boolean
                              combinations (~ & |) of the time-0 vectors of every stored
                              CS up to this size (unit costs); 0 = off.  Only with
                              VFB_SEARCH_CS, and not in the server */
    uint16_t cost[8];      /* LTL_BS's cost function (LTL_BS.cu:1085-1493, its CLI c1..c8):
                              costs of p ~ & | X F G U; the levels are then costs and
                              max_size is maxCost.  All 0 = unit costs (formula size).
                              Non-unit costs: no probe, no onTheFly round (a capped
                              level ends the search with VFB_CAPACITY), no BS stage */
    int neg_type;          /* LTL_BS's NegType: 1 (or 0) ~ of any formula, 2 ~ of atoms only */
    int pad_opt;
} vfb_options;

#define VFB_SEARCH_CS 0
#define VFB_SEARCH_BRUTE 1

typedef struct {
    vfb_status status;
    int min_size;          /* size of the returned formula (0 if none) */
    int top_op;            /* vfb_op of its root */
    int levels_done;       /* highest size level processed completely */
    uint64_t distinct[VFB_MAX_SIZE + 1];   /* CS stored per level; at the found level,
                                              survivors ordered before the returned formula */
    uint64_t candidates[VFB_MAX_SIZE + 1]; /* candidates per level; at the found level,
                                              through the returned formula */
    uint64_t level_winners[VFB_MAX_SIZE + 1]; /* distinct CS per processed level, all of it;
                                                 at the found level, of its evaluated part */
    uint64_t level_candidates[VFB_MAX_SIZE + 1]; /* candidates evaluated per level: all of a
                                                    completed level; at the found level, those
                                                    of the passes run (vfb_tuning.pass_min) */
    uint64_t found_ordinal; /* canonical ordinal of the returned formula in its level */
    char* formula;         /* ltl_nocache.cu:289-317 syntax; "Not found !" if none */
    double ms_total;       /* wall time of the solve (host clock) */
    double ms_expand;      /* HIP-event time of the combine+key+dedup kernels */
    double ms_resolve;     /* winner resolution + scan kernels */
    double ms_store;       /* survivor compaction kernels */
    uint64_t expand_launches;
    uint64_t table_slots;  /* final hash-set capacity */
    uint64_t retries;      /* epochs re-run after a table or list grew */
    uint64_t bytes_expand; /* algorithmic bytes of the combine+key+dedup kernel:
                              (a*8T + 16) per candidate, a = 2 for & and |, else 1 */
    uint64_t bytes_store;  /* algorithmic bytes of survivor storage: 8T + 24 per survivor */
    uint64_t fresh_total;  /* candidates whose key was new to the set at expand time */
    int otf_level;         /* level whose rows were truncated (onTheFly, LTL_BS.cu:516-522);
                              the next level was then checked only and the search
                              stopped (status VFB_CAPACITY if nothing matched); 0: none */
    int bs_stage;          /* 1: the formula was found by the BS stage (vfb_options.bs_max_size);
                              distinct[] / candidates[] then continue past max_size with
                              its levels */
    uint64_t otf_stored;   /* rows of otf_level that were stored */
    double ms_probe;       /* HIP-event time of the target-check probes of large levels
                              (vfb_tuning.pass_min); not part of ms_expand */
    uint64_t probe_candidates; /* candidates those probes checked */
    uint32_t restarts;     /* batch / server: times the solve was evicted for arena memory
                              and started again from its atoms (0 for vfb_solve) */
    uint32_t parked;       /* vfb_solve_batch: 1 if it waited for the others at a level of
                              vfb_tuning.park_min candidates and then went on alone */
    uint64_t chain_levels; /* levels built inside the persistent level chain
                              (vfb_tuning.chain_blocks), with no host round trip */
} vfb_result;

void vfb_options_default(vfb_options* opt);

/* Plain single-instance solve: the drop-in for LTLI (LTL_BS.cu:1069-1532). */
vfb_status vfb_solve(const vfb_instance* inst, const vfb_options* opt, vfb_result* out);

/* Device scratch of vfb_solve_noalloc, per GPU (256 MiB). */
#define VFB_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* vfb_solve without any device or pinned-memory allocation, for callers
   that cannot allocate (serving loops that forbid hipMalloc, harnesses that
   own all device memory).  No counterpart in the reference.  Every device
   buffer of the solve -- level rows, hash set, fresh lists, winner maps,
   descriptors -- comes from one static device array of
   VFB_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from its first
   use on that GPU until the process ends; host staging is static too.  The
   caller still passes the instance and the result.  Results equal
   vfb_solve's with the same options; a solve that needs more device memory
   than the scratch returns VFB_CAPACITY (formula "see_the_error", a message
   on stderr) instead of being cut short, and search = VFB_SEARCH_BRUTE is
   refused (VFB_INVALID_INPUT).  The timing option is ignored.  The array
   keeps the previous call's contents, and the solve clears what must start
   empty; calls are serialised by the library's lock (two calls overlapping
   on one GPU would otherwise race on the one array). */
vfb_status vfb_solve_noalloc(const vfb_instance* inst, const vfb_options* opt, vfb_result* out);

/* Many independent instances stepped together, one level per step, in shared
   launches.  out[i] is identical to vfb_solve(insts[i]).  No reference
   counterpart (the reference solves one instance per process). */
vfb_status vfb_solve_batch(const vfb_instance* const* insts, size_t n,
                           const vfb_options* opt, vfb_result* out);

void vfb_result_free(vfb_result* res);
const char* vfb_status_string(vfb_status s);

/* ---- serving: continuous batching -------------------------------------- */

/* A request loop on a background thread.  Every iteration builds one size
   level for every active request in shared launches (requests at different
   levels interleave), then admits queued requests while fewer than
   max_active run and the device memory held by the running ones is below
   hbm_budget (0 = the process arena).  A request's result is identical to
   vfb_solve() of the same instance with the same options.  No reference
   counterpart (the reference solves one instance per process). */
typedef struct vfb_server vfb_server;

vfb_status vfb_server_create(const vfb_options* opt, uint64_t hbm_budget, int max_active, vfb_server** out);

/* Queues a copy of inst; the ticket identifies the request. */
vfb_status vfb_server_submit(vfb_server* srv, const vfb_instance* inst, uint64_t* ticket);

/* Non-blocking: when the request has finished, *done = 1, *out receives its
   result (ms_total = completion minus submission) and the ticket is released;
   otherwise *done = 0.  The return value is then the result's status.
   A ticket is consumed once: VFB_INVALID_INPUT (and *done = 0, *out
   untouched) for an unknown or already consumed ticket. */
vfb_status vfb_server_poll(vfb_server* srv, uint64_t ticket, vfb_result* out, int* done);

/* Blocks until the request has finished, then as vfb_server_poll.  A caller
   still blocked when another caller consumes the ticket, or when
   vfb_server_destroy stops the server, returns VFB_INVALID_INPUT with *out
   untouched; vfb_server_destroy returns only after such callers have left. */
vfb_status vfb_server_wait(vfb_server* srv, uint64_t ticket, vfb_result* out);

/* Requests waiting, running, and device bytes held by the running ones;
   peak_active is the largest number of requests that ran together. */
void vfb_server_stats(vfb_server* srv, uint64_t* queued, uint64_t* active, uint64_t* bytes,
                      uint64_t* peak_active);

/* Stops the loop (requests still queued or running are dropped). */
void vfb_server_destroy(vfb_server* srv);

/* ---- tuning and the autotuner ------------------------------------------ */

/* Launch and scheduling parameters of the process-wide engine.  They change
   speed only -- every result is identical under every valid setting -- except
   table_min_slots and max_load_pct, which size the hash sets: like hbm_budget,
   they take part in deciding whether a level's set and scratch fit the HBM
   budget (a level that does not fit is checked only, the onTheFly round), so
   a solve that reaches the budget can be cut at another level.  No
   reference counterpart (the reference hard-codes 1024-thread blocks and one
   launch per temp-buffer chunk, LTL_BS.cu:383-446, 1226-1493). */
typedef struct {
    uint64_t pass_min;             /* levels with at least this many candidates are first
                                      probed (target check only, no keys); a level holding
                                      a match is then built one canonical operator block at
                                      a time (~, &, |, X F G, U) and stops at the block that
                                      settles it, any other level in one fused pass;
                                      0 = never probe */
    uint32_t wide_max_waves;       /* an expansion of fewer wavefronts gives each
                                      candidate 8 lanes (0 = never) */
    uint32_t expand_blocks_per_cu; /* 256-thread workgroups per CU, expand launch, 1..64 */
    uint32_t resolve_blocks_per_cu;/* same, winner resolution, 1..64 */
    uint32_t store_blocks_per_cu;  /* same, survivor storage, 1..64 */
    uint64_t park_min;             /* vfb_solve_batch: a member whose next level has at least
                                      this many candidates waits, its state kept, until the
                                      others are done, then goes on alone (oldest first), so
                                      such members neither share the device nor evict one
                                      another mid-solve; 0 = never park */
    uint64_t table_min_slots;      /* hash sets start at least this large (rounded up to a power
                                      of two, at most 2^32; default 2^17 = 4 MB); a set is
                                      otherwise grown by rehash as its keys need */
    uint32_t max_load_pct;         /* hash-set load (percent of slots) at which a set is grown,
                                      50..90 (default 70) */
    uint32_t store_traces;         /* traces per survivor-storage work item on levels of at least
                                      2^22 candidates, 1..128 (default 64; smaller levels take 8) */
    uint32_t check_rows;           /* left rows per wavefront in target-check launches
                                      (1, 4, 8 or 16; default 8) */
    uint32_t chain_blocks;         /* persistent level chain: at most this many 256-thread
                                      workgroups per instance (a batch gives each member
                                      what the device holds at once divided among them,
                                      at least 1); 0 = no chain, levels return to the host
                                      one by one; 0..64 (default 0: on MI355X a chain level
                                      takes about as long as a host-built level, round trip
                                      included, so plain solves run without it) */
    uint32_t chain_cands;          /* a level is built inside the chain when it has at most
                                      chain_cands candidates per workgroup of its instance
                                      (and at most 2^18), its solve has unit costs and
                                      one-word traces, and its hash set need not grow;
                                      64..65536 (default 1024) */
    uint32_t lds_chain;            /* 1: the smallest levels of a solve (at most 4096
                                      candidates, its set at most 2048 keys with all of them
                                      new) are built one after another by one workgroup with
                                      the solve's whole hash set in LDS, before any other
                                      path; 0: off (default 1) */
    uint32_t set_margin_pct;       /* a level's hash set is sized for the keys it holds plus
                                      the level's candidates times the last level's survivor
                                      ratio times set_margin_pct / 100; a level that claims
                                      more is built again in a larger set (a retry).  Memory
                                      and retries only, never a result; 1..1000 (default 130) */
    uint32_t pad_tuning;
    uint64_t direct_min;           /* a level of at least direct_min candidates keeps no
                                      fresh list: the expansion marks each new key's
                                      candidate, and the one a lower ordinal displaces, in
                                      byte maps that the count kernel packs (no resolve
                                      pass over the set); smaller levels list their new
                                      keys and resolve them; >= 1 (default 2^18) */
} vfb_tuning;

void vfb_tuning_default(vfb_tuning* t);
/* Current settings of the process (the defaults until vfb_set_tuning). */
void vfb_get_tuning(vfb_tuning* t);
/* VFB_INVALID_INPUT (and no change) for a value out of range. */
vfb_status vfb_set_tuning(const vfb_tuning* t);
/* Text file of "key value" lines, with the device it was tuned on. */
vfb_status vfb_tuning_save(const vfb_tuning* t, const char* path);
vfb_status vfb_tuning_load(const char* path, vfb_tuning* t);

/* Project autotuner: coordinate search over the vfb_tuning fields, timing
   `reps` runs of the sample per setting -- vfb_solve_batch over all of it
   (batched = 1) or vfb_solve of each instance in turn (batched = 0) -- and
   keeping the fastest.  Every setting's results are checked against those of
   the first; a difference returns VFB_DEVICE_ERROR.  On return the best
   setting is active, stored in *best, and saved to cache_path if not NULL;
   *ms_default / *ms_best receive the sample's time under the starting and
   the chosen settings. */
vfb_status vfb_autotune(const vfb_instance* const* sample, size_t n, const vfb_options* opt,
                        int batched, int reps, const char* cache_path, vfb_tuning* best,
                        double* ms_default, double* ms_best);

/* Size of the device arena that every solve of the process allocates its
   device memory from (levels, hash sets, scratch): 0 = 85% of the HBM free
   at first use (the default), else at least 64 MiB.  The engine is made
   again with it at the next solve, so no call may be running; batch members
   and served requests that do not fit together evict one another and are
   restarted, with unchanged results.  VFB_INVALID_INPUT while a server runs
   or below 64 MiB.  Replaces the reference's sizing of its cache from the
   device's memory in every LTLI call (LTL_BS.cu:1179-1183). */
vfb_status vfb_set_arena_bytes(uint64_t bytes);

/* Kernel time and algorithmic bytes of every solve of the process, counted
   once per launch (the per-result ms_* fields give each instance of a
   batch the whole time of the launches it shared).  Times come from the HIP
   events of solves run with vfb_options.timing (2: the combine+dedup kernel
   only); bytes and launches are counted always.  No reference counterpart. */
typedef struct {
    double ms_expand;            /* combine+key+dedup kernel (k_expand), dedup passes */
    uint64_t expand_launches;
    uint64_t bytes_expand;       /* (a*8W + 16) per candidate of those launches (probe
                                    candidates sharing a launch counted at a*8W) */
    double ms_probe;             /* launches that only probed (target check) */
    uint64_t probe_launches;
    uint64_t bytes_probe;        /* a*8W per probed candidate */
    double ms_resolve;
    double ms_store;
    uint64_t store_launches;
    uint64_t bytes_store;        /* 8W + 24 per stored row */
    double ms_chain;             /* the persistent level chain (k_chain) */
    uint64_t chain_launches;
    uint64_t chain_levels;       /* instance levels it built */
    uint64_t bytes_chain;        /* (a*8W + 16) per candidate of those levels */
    uint64_t chain_solo_launches; /* of chain_launches: teams of one workgroup because the
                                    runtime refused the cooperative launch (another process
                                    on the device) */
} vfb_kernel_stats;

/* Copies the counters into *out (if not NULL), then clears them if reset. */
void vfb_kernel_stats_get(vfb_kernel_stats* out, int reset);

/* Device name and compute-unit count, for reports.  Returns VFB_DEVICE_ERROR
   when no HIP device is usable. */
vfb_status vfb_device_info(char* name, size_t name_cap, int* num_cus);

#ifdef __cplusplus
}
#endif
#endif
