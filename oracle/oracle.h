/*
 * CPU oracle for the DCFR postflop solve path — TEST INFRASTRUCTURE ONLY.
 *
 * This is an independent CPU restatement of the reference solver's
 * semantics, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker. Nothing in the product library links,
 * loads or calls it.
 *
 * Restated reference behaviour (file:line into the reference tree):
 *   - card ids, combo enumeration            src/cards.h:13-17, src/range.cpp:10-45
 *   - hand ranking + 64-bit hand key         src/cards.cpp:30-165, cuda/gpu_cfr.cu:616-621
 *       (restated as best-of-C(n,5) over a 5-card classifier, a different
 *        algorithm with the same ordering and tiebreak convention)
 *   - range parser incl. its quirks          src/range.cpp:176-254
 *   - bet abstraction                        src/postflop_cfr.cpp:18-152
 *   - tree rules                             cuda/gpu_cfr.cu:1116-1326
 *   - DCFR half-step arithmetic (f32 rows)   cuda/gpu_cfr.cu:227-579
 *       (restated as a recursive per-node walk; showdown via the
 *        strength sweep of postflop-solver game/evaluation.rs:152-250;
 *        chance sums accumulated in f64 as postflop-solver's
 *        solver.rs:191-199, one rounding to f32)
 *   - discount schedules                     cuda/gpu_cfr.cu:585-614
 *   - f64 walks, pair mass, exploitability   cuda/gpu_cfr.cu:871-923,2036-2107
 *   - seeded-strategy probe profiles         cuda/gpu_cfr.cu:2062-2097
 *
 * Pins (see tests/): golden vectors from the reference's own C++ (hand keys,
 * parsed ranges, bet-action lists), the reference's tiny turn ground truth
 * (tools/tiny_check.cpp: uniform, mode 7, mode 9), textbook hand-category
 * counts, the postflop-solver degenerate known answers (game/tests.rs),
 * and a per-pair brute-force DCFR on tiny spots.
 */
#ifndef PPS_ORACLE_H
#define PPS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_spot {
  int n_board;            /* 3, 4 or 5 */
  int board[5];           /* card ids suit*13+rank */
  double w0[1326];        /* OOP combo weights */
  double w1[1326];        /* IP combo weights */
  int64_t pot, stack;
  int n_bets;             /* pot fractions */
  double bets[16];
  int n_raises;           /* previous-bet multiples */
  double raises[16];
  int bet_allin, bet_geo, max_raises;
  double add_allin, force_allin, merging; /* thresholds (1.5, 0.15, 0.1) */
} orc_spot;

/* integer layer */
uint64_t orc_hand_key(const int* cards, int n);         /* n in 5..7 */
int orc_parse_range(const char* s, double* w1326, char* err, int errcap);
/* bet actions: out pairs (kind, amount); returns count or -1 */
int orc_bet_actions(int64_t pot_base, int64_t stack, const orc_spot* cfg,
                    int64_t sc0, int64_t sc1, int actor, int after_allin,
                    int num_bets, int num_streets, int64_t* out, int cap);
void orc_schedule(const char* algo, int t, int iters, double out[3]);

/* all 5-card hands: category counts [0..8] and distinct keys [9] */
void orc_count5(int64_t out[10]);

/* solver */
typedef struct orc_solver orc_solver;
orc_solver* orc_create(const orc_spot* spot);  /* NULL on empty ranges */
void orc_destroy(orc_solver* s);
void orc_set_threads(int n);
/* counts: nodes, max depth, decide nodes, regret row total, showdown,
   fold, chance node counts, distinct boards */
void orc_counts(const orc_solver* s, int64_t out[8]);
int orc_solve(orc_solver* s, int iters, const char* algo);       /* fresh */
int orc_continue(orc_solver* s, int iters);                      /* dcfr */
/* one CFR half-step for traverser tr with schedule index t (rows only;
   the iteration counter is not advanced) */
int orc_half_step(orc_solver* s, int tr, int t, const char* algo, int planned);
/* ev0, ev1, expl, pair mass */
int orc_stats(orc_solver* s, double out[4]);
/* root value per walk: tr, mode 1=EV 2=BR, fills the root cfv (f32) */
double orc_root_value(orc_solver* s, int tr, int mode, float* root_cfv);
int orc_root_strategy(orc_solver* s, double* out, int cap);
/* one walk (tr, mode 1=EV 2=BR) recording node's traverser values (its
   list order) and the opponent reach arriving there; counts out */
int orc_node_values(orc_solver* s, int node, int tr, int mode, float* cfv, float* reach, int cap,
                    int* n_tr, int* n_opp);
/* the cards of each entry of a player's list at node (2 per entry) */
int orc_node_cards(const orc_solver* s, int node, int player, int* cards, int cap);
int orc_node_strategy(orc_solver* s, int decide_idx, double* out, int cap);
/* rows: which 0 = regret, 1 = strategy sum; concatenated over decide
   nodes in creation order, action-major per node */
int64_t orc_rows(orc_solver* s, int which, float* out, int64_t cap);
int64_t orc_set_rows(orc_solver* s, int which, const float* in, int64_t n);
/* decide node meta: node id, player, depth, n_board, na, n_combos */
int orc_decide_meta(const orc_solver* s, int idx, int out[6]);
/* seeded probe profile into the strategy-sum rows (modes 1..9) */
int orc_seed_strategy(orc_solver* s, int mode);
/* per-iteration algorithmic f32 row bytes of the compiled tree */
double orc_alg_bytes(const orc_solver* s);

#ifdef __cplusplus
}
#endif
#endif
