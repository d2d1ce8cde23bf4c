/*
 * pps_capi — C ABI of the MI355X postflop solver (exported by libppsg.so).
 *
 * The reference exposes its solver through pybind11 (python/pps_native.cpp)
 * over the C++ class pps::gpu::GpuPostflopSolver (cuda/gpu_cfr.h:141-264).
 * These are the entry points a C FFI (ctypes, cgo, JNI, N-API) binds to
 * reach the same solver: plain pointers and sizes, status codes, and
 * pps_last_error() for the message. Each entry cites the reference member
 * it stands for. Entries marked "no reference counterpart" are additions
 * (compile-only introspection and row access for parity checks).
 *
 * Status: 0 = ok, -1 = error (pps_last_error, thread-local), >= 0 counts
 * where documented. A solver handle is not thread-safe.
 */
#ifndef PPS_CAPI_H
#define PPS_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pps_spot {
  int32_t n_board;          /* 3, 4 or 5 */
  uint8_t board[5];         /* suit*13 + rank */
  double oop[1326];         /* combo weights, combo index (a<b) ascending */
  double ip[1326];
  int64_t pot, stack;
} pps_spot;                 /* PostflopSpot, cuda/gpu_cfr.h:35-42 */

typedef struct pps_bet_config {
  int32_t n_bets;
  double bets[16];          /* pot fractions */
  int32_t n_raises;
  double raises[16];        /* previous-bet multiples */
  int32_t bet_allin, bet_geo, max_raises;
  double add_allin, force_allin, merging;   /* 1.5, 0.15, 0.1 by default */
} pps_bet_config;           /* pf::BetConfig, src/postflop_cfr.h:24-42 */

typedef struct pps_solver pps_solver;

const char* pps_last_error(void);

/* ---- integer layer ---- */
/* handKey(evaluateN(cards, n)): src/cards.cpp:30-165 + cuda/gpu_cfr.cu:616-621 */
uint64_t pps_hand_key(const uint8_t* cards, int n);
/* parseRange: src/range.cpp:176-254; writes 1326 weights */
int pps_parse_range(const char* text, double* w1326);
/* pf::betActions: src/postflop_cfr.cpp:18-152; out = (kind, amount) pairs,
   returns the action count */
int pps_bet_actions(int64_t pot_base, int64_t stack, const pps_bet_config* cfg, int64_t sc0,
                    int64_t sc1, int actor, int after_allin, int num_bets, int num_streets,
                    int64_t* out, int cap);
/* discount coefficients of schedule index t: cuda/gpu_cfr.cu:585-614 */
int pps_schedule(const char* algo, int t, int iters, double out3[3]);

/* ---- compile only (no device; no reference counterpart) ----
   out: [0] nodes [1] max depth [2] decide nodes [3] row total [4] showdown
   [5] fold [6] chance [7] board classes; *bytes = algorithmic bytes/iter */
int pps_compile_info(const pps_spot* spot, const pps_bet_config* cfg, int64_t out[8],
                     double* bytes);
/* Addition: pps_compile_info of the tree compiled with option flags
   (PPS_SUIT_ISOMORPHISM). */
int pps_compile_info_ex(const pps_spot* spot, const pps_bet_config* cfg, int flags,
                        int64_t out[8], double* bytes);

/* ---- solver: GpuPostflopSolver ---- */
int pps_solver_create(const pps_spot* spot, const pps_bet_config* cfg,
                      pps_solver** out);                      /* cuda/gpu_cfr.h:145 */
int pps_solver_load(const char* path, pps_solver** out);     /* cuda/gpu_cfr.h:150 */
/* Addition: pps_solver_create with option flags (PPS_SUIT_ISOMORPHISM:
   include/pps/gpu_cfr.h SolverOptions::suitIsomorphism). */
#define PPS_SUIT_ISOMORPHISM 1
int pps_solver_create_ex(const pps_spot* spot, const pps_bet_config* cfg, int flags,
                         pps_solver** out);
void pps_solver_destroy(pps_solver* s);                       /* cuda/gpu_cfr.h:151 */
int pps_solver_solve(pps_solver* s, int iters, const char* algo,
                     double target);                          /* cuda/gpu_cfr.h:161 */
int pps_solver_continue(pps_solver* s, int iters, double target); /* cuda/gpu_cfr.h:169 */
int pps_solver_reset(pps_solver* s);                          /* cuda/gpu_cfr.h:173 */
int pps_solver_iterations_run(pps_solver* s);                 /* cuda/gpu_cfr.h:177 */
int64_t pps_solver_total_iterations(pps_solver* s);           /* cuda/gpu_cfr.h:181 */
/* ev0, ev1, expl, pair mass (x2^20 as a double) */
int pps_solver_stats(pps_solver* s, double out4[4]);          /* cuda/gpu_cfr.h:185 */
int pps_solver_root_strategy(pps_solver* s, double* out, int cap);   /* cuda/gpu_cfr.h:188 */
int pps_solver_node_strategy(pps_solver* s, int idx, double* out,
                             int cap);                        /* cuda/gpu_cfr.h:197 */
int pps_solver_num_decide_nodes(pps_solver* s);               /* cuda/gpu_cfr.h:201 */
/* info6: node id, player, depth, n_board, n_combos, n_actions;
   actions: (kind, amount, child) triples */
int pps_solver_decide_node(pps_solver* s, int idx, int32_t info6[6], int64_t* actions,
                           int cap);                          /* cuda/gpu_cfr.h:202 */
/* combos[n_combos], freqs[n_actions * n_combos]; returns n_combos */
int pps_solver_combo_strategy(pps_solver* s, int idx, int32_t* combos, double* freqs,
                              int cap);                       /* cuda/gpu_cfr.h:209 */
int pps_solver_num_nodes(pps_solver* s);                      /* cuda/gpu_cfr.h:239 */
int pps_solver_max_depth(pps_solver* s);                      /* cuda/gpu_cfr.h:240 */
int pps_solver_num_base_combos(pps_solver* s, int player);    /* cuda/gpu_cfr.h:229 */
int pps_solver_save(pps_solver* s, const char* path);         /* cuda/gpu_cfr.h:245 */

/* ---- parity / measurement hooks (no reference counterpart) ---- */
/* which 0 = regret, 1 = strategy sum; rows concatenated over decide nodes
   in creation order, action-major; returns the row total (out == NULL:
   size query) */
int64_t pps_solver_read_rows(pps_solver* s, int which, float* out, int64_t cap);
int pps_solver_write_rows(pps_solver* s, int which, const float* in, int64_t n);
/* one CFR half-step for traverser tr with schedule index t of `algo` */
int pps_solver_half_step(pps_solver* s, int tr, int t, const char* algo, int planned);
/* root value walk: mode 1 EV, 2 BR; fills the root cfv row; returns value */
int pps_solver_root_value(pps_solver* s, int tr, int mode, float* root_cfv, double* value);
double pps_solver_iteration_bytes(pps_solver* s);
/* per-launch event timing of `iters` more iterations (profileIterations),
   as ppsg_profile: 13 values (forward incl. fold/showdown launches,
   backward, per-terminal-kind launches); detail != 0: 17 values with the
   per-kind re-runs of the forward and backward launches, after which
   the regret and strategy rows and the iteration count are restored
   (temporary device memory: 8 bytes per regret entry) */
int pps_solver_profile(pps_solver* s, int iters, int detail, double* out);
/* ppsg_info of the solver's engine: 6 values */
int pps_solver_engine_info(pps_solver* s, int64_t out6[6]);
/* The device engine (include/ppsg.h) the solver drives and its slot there:
   the solver's own one-spot engine (slot 0), or the arena of the server that
   serves it. For callers of the ppsg_* calls on it, e.g. the root value sums
   (ppsg_batch_root_sums). The handle stays owned by the solver or server; a
   served solver's engine is driven by the server's thread while a job runs.
   Error for an empty tree. No reference counterpart. */
int pps_solver_engine(pps_solver* s, struct ppsg_engine** engine, int* slot);
double pps_solver_last_solve_ms(pps_solver* s);

/* ---- many spots on one GPU (no reference counterpart) ----
   Continuous batching: every admitted spot stays resident in one device
   arena and all live spots advance together, `check_every` iterations per
   step, with one batched exploitability check after each step. A spot
   retires at expl <= target_frac * pot or at max_iters; requests are
   admitted in arrival order as slots and arena memory free up. A spot's
   result does not depend on the rest of the batch and equals
   pps_solver_solve(iterations, algo) of the same spot bit for bit.
   (The reference serves one solve at a time: python/pps_api.py.) */
typedef struct pps_serve_options {
  double target_frac;       /* 0.005 */
  int32_t check_every;      /* 32 */
  int32_t max_iters;        /* 4096 (rounded up to check_every) */
  int64_t arena_bytes;      /* 0: 80% of free device memory */
  int32_t max_spots;        /* resident slots, 1024 */
  int32_t max_live;         /* 0: memory and slots only; 1: one at a time */
  int32_t compile_threads;  /* 8 */
  int32_t real_time;        /* 1: honour arrival_ms; 0: all arrive at once */
  int32_t prefill;          /* 1: admit all that fit before the first step */
  int32_t flags;            /* option flags of every request (PPS_SUIT_ISOMORPHISM) */
  const int32_t* shapes;    /* per-depth launch shapes (a ppsg_autotune_shapes
                               table, e.g. from pps_batch_profile) or NULL */
  int32_t n_shapes;         /* entries in shapes */
  int32_t term_split;       /* terminal split depth (PPSG_TUNE_TERM_SPLIT, e.g.
                               from pps_batch_profile); 0: one stream */
  const int32_t* request_flags; /* n option flags, one per request (they
                               replace `flags` for it), or NULL */
} pps_serve_options;

typedef struct pps_serve_result {
  int32_t status;           /* 0 solved, -1 rejected (see error) */
  int32_t converged;        /* reached the target before max_iters */
  int32_t iterations;
  int32_t n_root_actions;
  double expl, ev0, ev1;    /* as pps_solver_stats */
  double root_strategy[32]; /* as pps_solver_root_strategy */
  int64_t nodes, footprint; /* compiled nodes, arena bytes */
  double arrival_ms, admit_ms, retire_ms;  /* stream clock */
  char error[96];
} pps_serve_result;

typedef struct pps_serve_report {
  double wall_ms;           /* stream start to last retirement */
  double busy_ms;           /* steps + checks + retirements: spots resident */
  double device_ms, walk_ms, admit_ms, compact_ms;
  int64_t iterations;       /* sum of per-spot iterations */
  int32_t steps, peak_live, compactions, group_size;
  /* the first PPS_SERVE_TRACE steps: live spots, device ms, check ms */
  double step_live[256], step_device_ms[256], step_check_ms[256];
} pps_serve_report;
#define PPS_SERVE_TRACE 256

/* Fills options with the defaults above. */
void pps_serve_defaults(pps_serve_options* opt);
/* Solves n spots (all with bet config cfgs[i], schedule `algo`: "dcfr" or
   "dcfr-g2"); arrival_ms may be NULL (all at 0). out[n]; report may be
   NULL. Returns 0, or -1 on an engine error (pps_last_error). */
int pps_serve(int n, const pps_spot* spots, const pps_bet_config* cfgs, const double* arrival_ms,
              const char* algo, const pps_serve_options* opt, pps_serve_result* out,
              pps_serve_report* report);
/* Launch-level timing of n spots resident in one arena (all admitted at
   once). After the autotuner's walks, `warmup` batched iterations are
   replayed untimed, then `iters` timed: out = [spots, iters, graph ms per batched iteration, forward ms,
   forward launches, backward ms, backward launches, forward bytes and
   backward bytes per batched iteration (algorithmic f32 rows), kernels
   per iteration, work items per iteration, arena bytes, (detail != 0)
   backward decide and backward chance ms per iteration run alone, fold
   and showdown launch ms per iteration, the terminal group size the
   autotuner chose and its EV-walk ms for group sizes 1, 2, 4, 8, 16, 32,
   (detail) forward decide and forward chance ms per iteration run alone,
   the algorithmic bytes per batched iteration of forward decide, forward
   chance, backward decide, backward chance, fold, showdown, the fold
   and showdown launch counts over the timed iterations, the terminal split
   depth the autotuner chose and its ms per dry half-step pair for splits
   0, 3, 4, 5, 6, 7, 8, 9]. The per-depth launch shapes are autotuned first
   (ppsg_autotune_shapes); their table goes to `shapes` (if not NULL and it
   holds shapes_cap >= 1 + 6 * depths entries) for
   pps_serve_options.shapes; the split (after the shapes) is
   pps_serve_options.term_split. */
int pps_batch_profile(int n, const pps_spot* spots, const pps_bet_config* cfgs, const char* algo,
                      int iters, int warmup, int64_t arena_bytes, int detail, double out[42],
                      int32_t* shapes, int32_t shapes_cap);
/* Device bytes the compiled spot takes in an arena (host only). */
int pps_compile_footprint(const pps_spot* spot, const pps_bet_config* cfg, int64_t* bytes);

/* What a solution file (GpuPostflopSolver::save, PPS1 v1) holds, read and
   validated on the host without loading its rows or touching a device; an
   invalid file is an error naming the field (pps_last_error). No reference
   counterpart (the reference reads files only to load them). */
typedef struct pps_solution_info {
  int32_t n_board;
  uint8_t board[5];
  int64_t pot, stack;
  int32_t n_bets, n_raises;   /* bet and raise size list lengths */
  int64_t nodes, iterations;  /* tree size and schedule index of the rows */
  int64_t rows;               /* entries in each of the two row tables */
  char algo[33];
} pps_solution_info;
int pps_solution_inspect(const char* path, pps_solution_info* out);

/* ---- spot server (no reference counterpart as a call; it replaces the
   global solve lock of the reference's HTTP layer, python/pps_api.py:165,
   by batching) ----
   Spots registered with a server stay resident in its device arena; a
   driver thread advances every spot with a running job together. A served
   spot is a pps_solver: every pps_solver_* query works on it (between the
   server's steps), pps_solver_solve / pps_solver_continue run as batched
   jobs of the server and return when the job ends, with results equal bit
   for bit to a solver of its own. pps_solver_destroy removes it from the
   server. Calls from many threads at once are safe; jobs started from
   different threads share the GPU's launches. */
typedef struct pps_server pps_server;
typedef struct pps_server_options {
  int64_t arena_bytes;  /* 0: 80 % of free device memory */
  int32_t max_spots;    /* resident spots at most (256) */
  int32_t max_step;     /* iterations per batched step at most (32) */
} pps_server_options;
typedef struct pps_job {
  int32_t running;
  int32_t status;   /* PPS_JOB_* */
  int32_t kind;     /* 0 none, 1 solve, 2 continue */
  int32_t pad;
  int64_t start;    /* total iterations when the job began */
  int64_t done;     /* iterations run by the job */
  int64_t total;    /* iterations asked for */
  double target;    /* exploitability target in chips, <= 0: none */
} pps_job;
enum { PPS_JOB_NONE = 0, PPS_JOB_RUNNING = 1, PPS_JOB_DONE = 2, PPS_JOB_CANCELLED = 3,
       PPS_JOB_ERROR = 4 };
typedef struct pps_server_counters {
  int64_t steps, iterations;  /* batched steps; spot-iterations run */
  double device_ms, check_ms;
  int32_t resident, running, peak_running, pad;
  int64_t arena_bytes, free_bytes;
} pps_server_counters;
/* pps_server_add / pps_server_load return this when the arena is full */
#define PPS_ARENA_FULL 2

int pps_server_create(const pps_server_options* opt, pps_server** out); /* opt may be NULL */
/* Releases the handle; the server stops when its last served solver is
   destroyed too (running jobs are cancelled then). */
void pps_server_destroy(pps_server* srv);
int pps_server_add(pps_server* srv, const pps_spot* spot, const pps_bet_config* cfg,
                   pps_solver** out);
int pps_server_add_ex(pps_server* srv, const pps_spot* spot, const pps_bet_config* cfg, int flags,
                      pps_solver** out);
int pps_server_load(pps_server* srv, const char* path, pps_solver** out);
int pps_server_stats(pps_server* srv, pps_server_counters* out);
/* Jobs of a served solver: start and return at once (-1 with a running
   job); pps_solver_job reads the job (wait != 0: after it ends); cancel
   returns 1 if a running job was asked to stop, 0 if none ran. A job's
   error message is left in pps_last_error by pps_solver_job, which then
   returns -1. */
int pps_solver_start_solve(pps_solver* s, int iters, const char* algo, double target);
int pps_solver_start_continue(pps_solver* s, int iters, double target);
int pps_solver_job(pps_solver* s, int wait, pps_job* out);
int pps_solver_cancel(pps_solver* s);

#ifdef __cplusplus
}
#endif
#endif
