/*
 * ppsg — C ABI of the MI355X DCFR engine (libppsg.so).
 *
 * Plain pointers and sizes; no C++ or GPU-runtime types. Every call returns
 * 0 on success or a negative code; ppsg_last_error() gives the message for
 * the calling thread. A handle is not thread-safe: one thread drives it.
 *
 * This is the seam under the reference's engine API: the host compiler
 * (pps::gpu::compileSpot) produces a ppsg_tree_desc, and the C++
 * GpuPostflopSolver (include/pps/gpu_cfr.h), the gpu_pfflop CLI and the pps
 * Python module drive the engine through these calls. Each entry point
 * names the reference internals it replaces (cuda/gpu_cfr.cu).
 *
 * Tree description. Combo lists are per "board class" (the board after a
 * given sequence of dealt cards): every node of a class shares its lists,
 * which are the player's base list minus combos holding a dealt card, in
 * base order. Per-player arrays below are indexed [player][...]; a class's
 * n entries start at cls_list_off[p][cls].
 */
#ifndef PPSG_H
#define PPSG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PPSG_DECIDE = 0, PPSG_SHOWDOWN = 1, PPSG_FOLD = 2, PPSG_CHANCE = 3 };
enum { PPSG_MODE_CFR = 0, PPSG_MODE_EV = 1, PPSG_MODE_BR = 2 };
enum { PPSG_REGRET = 0, PPSG_STRAT = 1, PPSG_REACH0 = 2, PPSG_REACH1 = 3, PPSG_CFV = 4 };
enum { PPSG_TUNE_GROUP = 0, PPSG_TUNE_TERM_SPLIT = 1, PPSG_TUNE_ADD_GROUP = 2 };

/* Row offsets (reg_off, cfv_off, reach_off, in floats) are multiples of
   64: every f32 row starts on a 256-byte line and is padded to
   PPSG_ROW_STRIDE(n) floats, which the engine checks. */
#define PPSG_ROW_STRIDE(n) (((int64_t)(n) + 63) / 64 * 64)

typedef struct ppsg_tree_desc {
  int32_t n_base[2];               /* base list sizes (root class lists) */
  /* board classes */
  int32_t n_class;
  const int32_t* cls_n[2];         /* list size per class */
  const int64_t* cls_list_off[2];  /* first entry of the class's list */
  const int32_t* cls_board_n;      /* cards on the board per class (3..5) */
  int64_t list_total[2];
  const uint16_t* list_cards[2];   /* c1 | c2 << 8 per list entry */
  const int16_t* list_same[2];     /* slot of the same two cards in the other
                                      player's list of the class, or -1 */
  const uint16_t* to_parent[2];    /* slot in the parent class (class > 0) */
  const int64_t* from_parent_off[2]; /* per class: start in from_parent */
  int64_t from_parent_total[2];
  const int16_t* from_parent[2];   /* per parent slot: slot in this class or -1 */
  const int32_t* seg_off[2];       /* 53 per class: per-card segment bounds */
  const uint16_t* seg_slot[2];     /* 2n per class (at 2*cls_list_off): slots
                                      holding each card; strength order on
                                      5-card classes, slot order elsewhere */
  const uint16_t* sorted[2];       /* 5-card classes: slots by strength */
  const uint32_t* sd_meta[2];      /* 5-card classes, 4 words per traverser
                                      slot (at 4*cls_list_off): le | te<<16,
                                      A1 | B1<<16, A2 | B2<<16, c1 | c2<<8 |
                                      (same+1)<<16 */
  /* nodes, in the reference's creation order */
  int64_t n_nodes;
  int32_t max_depth;
  const uint8_t* kind;
  const int8_t* owner;             /* decide: player; fold: folder */
  const uint8_t* na;
  const int32_t* depth;
  const int32_t* cls;
  const int32_t* child_base;       /* into child_flat */
  int64_t n_child;
  const int32_t* child_flat;
  const int64_t* reg_off;          /* decide: regret/strategy block start;
                                      action a's row at reg_off + a *
                                      PPSG_ROW_STRIDE(n_actor) */
  const int64_t* cfv_off;
  const int64_t* reach_off[2];     /* per traverser layout (aliased rows);
                                      the children of an opponent-acting
                                      decide node PPSG_ROW_STRIDE(n) apart */
  const float* term_val;           /* 8 per node: [tr][win, tie, lose, -] */
  int64_t reg_total, cfv_total, reach_total[2];
  const float* root_reach[2];      /* layout tr root row: opponent weights */
  /* chance deals, per parent class (0 / -1 for classes without a chance
     node): the deal_n[c] cards the class's chance nodes deal (ascending);
     deal_sib[deal_off[c] + b] = the child (index among the chance node's
     children) whose values serve dealt card b (b itself without suit
     isomorphism); deal_map[p][deal_map_off[p][c] + b * n + s] = slot in that
     child's list whose value counts for parent slot s (suit-swapped for a
     card served by another card's child), or -1 where the card blocks s */
  const int32_t* deal_n;
  const int64_t* deal_off;
  int64_t deal_total;
  const uint8_t* deal_sib;
  const int64_t* deal_map_off[2];
  int64_t deal_map_total[2];
  const int16_t* deal_map[2];
  /* base weights (n_base[p] each, f64): the root value sums */
  const double* weight[2];
} ppsg_tree_desc;

typedef struct ppsg_engine ppsg_t;

/* Uploads the compiled tree and allocates the f32 rows.
   Replaces GpuPostflopSolver::init's uploads (cuda/gpu_cfr.cu:1515-1826). */
int ppsg_create(const ppsg_tree_desc* desc, ppsg_t** out);
void ppsg_destroy(ppsg_t* h);

/* Discount table: 3 floats (pos, neg, avg) per schedule index [0, n).
   Replaces the coefTab upload (cuda/gpu_cfr.cu:1884-1893,2018-2031). */
int ppsg_set_schedule(ppsg_t* h, const float* coef, int64_t n);
/* Device-resident schedule index (cuda/gpu_cfr.cu:218-220,2837-2839). */
int ppsg_set_counter(ppsg_t* h, int64_t t);
int ppsg_get_counter(ppsg_t* h, int64_t* t);
/* Zero regret + strategy rows and the counter (cuda/gpu_cfr.cu:1842-1855). */
int ppsg_reset(ppsg_t* h);
/* Enqueues n DCFR iterations (one captured graph per iteration) and waits.
   Replaces captureGraph + replayChunks (cuda/gpu_cfr.cu:844-868,1858-1872). */
int ppsg_iterate(ppsg_t* h, int n);
/* One CFR half-step for traverser tr at the current counter, outside the
   graph (rows only; the counter is not advanced). */
int ppsg_half_step(ppsg_t* h, int tr);
/* One EV (mode 1) or best-response (mode 2) half-step for traverser tr
   against the average strategy; copies the root cfv row (n_base[tr]
   floats) to host. Replaces rootValue's device half (cuda/gpu_cfr.cu:894-916). */
int ppsg_walk(ppsg_t* h, int tr, int mode, float* root_cfv);
/* Row transfer: buf is PPSG_REGRET .. PPSG_CFV; offsets in floats. */
int ppsg_read(ppsg_t* h, int buf, int64_t off, int64_t n, float* host);
int ppsg_write(ppsg_t* h, int buf, int64_t off, int64_t n, const float* host);
/* Device time (ms) of the last ppsg_iterate, from events on its stream. */
int ppsg_last_iterate_ms(ppsg_t* h, double* ms);
/* Engine facts: [0] kernels per iteration, [1] arena bytes,
   [2] work items per iteration, [3] graph captured (0/1), [4] forward and
   [5] backward algorithmic f32 row bytes per iteration (fold and showdown
   rows count as forward: one terminal launch per half-step evaluates
   them after the forward reach launches). */
int ppsg_info(ppsg_t* h, int64_t* out, int cap);
/* Runs `iters` CFR iterations launch by launch (no graph), timing every
   launch with events on the engine stream: out = [forward ms, forward
   launches, backward ms, backward launches, iterations]; the forward
   figures include the terminal launch. Advances the counter like
   ppsg_iterate. With cap >= 13: out[9..10] fold and [11..12] showdown
   launches (ms, launches; part of the forward figures). With cap >= 17 the
   forward and backward launches are also re-run restricted to one node
   kind: out[5..6] backward decide, [7..8] backward chance, [13..14]
   forward decide, [15..16] forward chance; the backward re-runs apply the
   updates again, so in that mode every live spot's regret and strategy
   tables and its counter are saved first (temporary device memory, 8
   bytes per regret entry) and restored after: the rows and counters are
   left as they were.
   (No reference counterpart; the reference's GPU_CFR_PROFILE mode plays
   this role, cuda/gpu_cfr.cu:1895-1987.) */
int ppsg_profile(ppsg_t* h, int iters, double* out, int cap);
/* Per-depth timing of `iters` CFR iterations run launch by launch (events on
   the engine stream, one sync per launch, so launch-bound trees read slower
   than from the graph; ratios are what to compare). out[0] = D (depths),
   then per depth d: forward us for traverser 0, backward us for 0, forward
   for 1, backward for 1 (out[1 + 4d ..]), then fold us for 0 and 1, showdown
   us (both widths) for 0 and 1; sums over the iterations. cap >= 1 + 4D + 4.
   Advances the counter like ppsg_iterate (these are real iterations).
   Replaces the GPU_CFR_PROFILE timing of cuda/gpu_cfr.cu:1895-1987. */
int ppsg_profile_depths(ppsg_t* h, int iters, double* out, int cap);
/* Calibration for hardware byte counters (no reference counterpart): a
   plain device copy of `bytes` (multiple of 16), `reps` launches on the
   null stream; mean launch time out. */
int ppsg_copy_probe(int64_t bytes, int reps, double* ms_per_launch);
const char* ppsg_last_error(void);

/* ---- batches: many resident spots in one device arena ----
 * No reference counterpart: the reference solves one spot per solver
 * object (cuda/gpu_cfr.h:141-264), one CUDA graph each. A batch handle
 * is an engine whose arena holds any number of compiled spots; every
 * step runs one DCFR iteration of every ACTIVE spot, depth d of all of
 * them in the same launch, so the launches per iteration do not grow
 * with the number of spots. Each spot keeps its own schedule table and
 * counter, so a spot's rows evolve exactly as they would alone (sums
 * have a fixed order: results are bitwise independent of the batch).
 * ppsg_create is a batch of one spot in slot 0; the single-spot calls
 * above act on slot 0 of such a handle (ppsg_iterate, ppsg_profile and
 * ppsg_info cover every active spot of any handle). */
/* Device bytes one spot of this description takes in an arena. */
int ppsg_footprint(const ppsg_tree_desc* desc, int64_t* bytes);
/* Allocates an arena of arena_bytes and room for max_spots spots. */
int ppsg_batch_create(int64_t arena_bytes, int max_spots, ppsg_t** out);
/* Uploads a spot (active, zero rows, counter 0, no schedule yet); returns
   -2 when no slot or no large enough arena block is free, even after
   compacting (nothing changed). */
int ppsg_batch_add(ppsg_t* h, const ppsg_tree_desc* desc, int* slot);
/* Retires a spot at once; its slot and arena block are reused after the
   next compaction (immediately when its items are not listed). */
int ppsg_batch_remove(ppsg_t* h, int slot);
/* Inactive (parked) spots stay resident but are skipped by steps and
   walks; removal and parking take effect at once (the spot's work items
   are masked until the engine next compacts its work lists). */
int ppsg_batch_set_active(ppsg_t* h, int slot, int active);
int ppsg_batch_set_schedule(ppsg_t* h, int slot, const float* coef, int64_t n);
int ppsg_batch_set_counter(ppsg_t* h, int slot, int64_t t);
int ppsg_batch_get_counter(ppsg_t* h, int slot, int64_t* t);
int ppsg_batch_reset(ppsg_t* h, int slot);
/* n DCFR iterations of every active spot (one graph replay each). */
int ppsg_batch_step(ppsg_t* h, int n);
/* EV / BR half-step of every active spot; copies the root cfv row of each
   listed (active) spot to root_cfv[i] (n_base[tr] floats; NULL skips). */
int ppsg_batch_walk(ppsg_t* h, int tr, int mode, int n_slots, const int* slots,
                    float* const* root_cfv);
/* EV / BR half-step of the listed spots only, live or parked (the others'
   rows are untouched); copies each listed spot's root cfv row to
   root_cfv[i] (NULL entries skip). For queries on one spot of a batch. */
int ppsg_batch_walk_only(ppsg_t* h, int tr, int mode, int n_slots, const int* slots,
                         float* const* root_cfv);
int ppsg_batch_read(ppsg_t* h, int slot, int buf, int64_t off, int64_t n, float* host);
/* EV or BR walk of one spot (live or parked) restricted to a subtree:
   the forward launches of the `path` nodes (the ancestors whose reach
   feeds it) and of the nodes in [lo, hi), its terminal nodes, and the
   backward launches of [lo, hi). Node ids are pre-order, so the subtree of
   node u is [u, one past its last descendant). Afterwards the cfv rows of
   the subtree and the reach rows of its nodes hold exactly what a full walk
   (ppsg_batch_walk) leaves there; other spots and other nodes' cfv rows
   are not touched. Replaces the whole-tree walks behind the node EV and
   runout queries (cuda/gpu_cfr.cu:2318-2684). */
int ppsg_batch_walk_range(ppsg_t* h, int slot, int tr, int mode, int64_t lo, int64_t hi,
                          const int64_t* path, int n_path);
/* Root value sums of the listed spots (live or parked; a slot may repeat):
   sums[k] = sum over the spot's base list of w_tr[i] * root_cfv[i], in index
   order, in f64 with each multiply and add rounded on its own (the host
   loop `s += w[i] * (double)cfv[i]` bit for bit), where root_cfv is the
   spot's root cfv row as the last launch for traverser tr left it (a walk
   or half-step of tr, e.g. ppsg_batch_walk with no output rows). The
   per-spot descriptors and the results are staged in device memory the
   handle allocates on first use and keeps (grown as needed). Replaces the
   host-side reductions of root rows copied back one spot at a time
   (cuda/gpu_cfr.cu:2036-2107, rootValue / explNow). */
int ppsg_batch_root_sums(ppsg_t* h, int tr, int n_slots, const int* slots, double* sums);
/* ppsg_batch_root_sums without allocating anything. No reference
   counterpart: it is for callers that cannot allocate device memory (code
   under graph capture, tight serving loops, harnesses that forbid it). The
   staging lives in a fixed device array of this library (128 KiB per GPU,
   held until the process ends); at most 4096 spots per call (more is an
   error and changes nothing). The array is shared by every call on a GPU
   and keeps the last call's contents; each call writes every descriptor
   and result it reads, so nothing has to start at zero. Calls that overlap
   on one GPU (from several threads or handles) race on the array. */
int ppsg_batch_root_sums_noalloc(ppsg_t* h, int tr, int n_slots, const int* slots,
                                 double* sums);
int ppsg_batch_write(ppsg_t* h, int slot, int buf, int64_t off, int64_t n, const float* host);
/* Free arena bytes in total and in the largest block (either may be NULL). */
int ppsg_batch_free_bytes(ppsg_t* h, int64_t* total, int64_t* largest);
/* Drops retired spots' masked items from the work lists (and frees their
   slots and arena blocks) when they exceed max_dead_fraction of all listed
   items; *compacted = 1 if it did (may be NULL). */
int ppsg_batch_compact(ppsg_t* h, double max_dead_fraction, int* compacted);
/* Algorithmic f32 row bytes of one iteration of the live spots by node
   kind: forward decide, forward chance, backward decide, backward chance,
   fold, showdown. */
int ppsg_kind_bytes(ppsg_t* h, double out[6]);
/* Launch shape (no reference counterpart; results do not depend on it).
   PPSG_TUNE_GROUP: fold/showdown nodes of one board class evaluated by one
   workgroup (1..64; default 4 for a batch; ppsg_create picks 2, 4 or 32
   from the spot's mean terminal nodes per board class: below 24, up to
   400, above). Larger groups load a class's gather list and per-combo
   metadata once for more nodes and launch fewer workgroups; smaller groups
   spread a small batch over more CUs. Sets the size of every spot (their
   terminal lists are re-cut) and of the spots added later.
   PPSG_TUNE_ADD_GROUP: the group size (1..64) of the spots added from now
   on only; the listed spots keep theirs (nothing is re-cut), so a stream
   can follow its live count without rebuilding the lists.
   PPSG_TUNE_TERM_SPLIT: depth k (0..32767; default 0). The fold and
   showdown nodes at depth <= k get work lists of their own, launched on a
   second stream as soon as the forward launch of depth k - 1 has produced
   their reach rows, beside the deeper forward, terminal and backward
   launches (which do not read them), and joined before the backward launch
   of depth k - 1. The terminal kernels are bound by LDS and VALU work, the
   update kernels by HBM, so a large batch gains from running them side by
   side; 0 keeps every launch on one stream (one spot: launch latency
   dominates). */
int ppsg_tune(ppsg_t* h, int param, int value);
/* Times n candidate group sizes on the live spots (reps EV walks of both
   traversers each, which change no rows), keeps the fastest; *chosen and
   ms[i] (per rep) may be NULL. The same as ppsg_autotune_param with
   PPSG_TUNE_GROUP. */
int ppsg_autotune(ppsg_t* h, const int* candidates, int n, int reps, int* chosen, double* ms);
/* Times n candidate values of a ppsg_tune parameter on the live spots and
   keeps the fastest: group sizes over EV walks, terminal splits over CFR
   half-steps whose backward launches store the old regret and strategy
   values (reach and cfv rows are rewritten by every half-step before they
   are read, so no solve state changes). reps half-steps of both traversers
   per candidate; *chosen and ms[i] (per rep) may be NULL. No reference
   counterpart. */
int ppsg_autotune_param(ppsg_t* h, int param, const int* candidates, int n, int reps, int* chosen,
                        double* ms);
/* Per-depth launch shapes (combos per lane of forward decide items from
   {1, 2, 4}, forward chance items from {2, 4, 8}, backward decide items from
   {1, 2}), picked for the live spots by timing each depth's launch of each
   kind at every candidate: the forward launches rewrite reach rows only and
   the backward ones run a dry update (the CFR update's loads and stores,
   writing the old values back), so no spot's rows change. table (if not
   NULL, cap >= 1 + 6 * depths) receives [depths, then per traverser and
   depth: decide, chance, backward]. Keeps the chosen shapes. No reference
   counterpart (the reference launches one block per node). */
int ppsg_autotune_shapes(ppsg_t* h, int reps, int* table, int cap);
/* Sets the shapes from such a table (n entries; depths beyond it keep the
   defaults 2 / 4 / 1) and rebuilds the work lists. */
int ppsg_set_shapes(ppsg_t* h, const int* table, int n);
/* Free and total memory of the current device (for sizing an arena). */
int ppsg_device_memory(int64_t* free_bytes, int64_t* total_bytes);

#ifdef __cplusplus
}
#endif
#endif
