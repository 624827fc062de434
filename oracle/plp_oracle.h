/* plp_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference 3D pallet-loading decoder, used as
 * the parity checker for the HIP path.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every function cites the reference file:line it follows.  Coordinates are
 * int cm; a placed box is 7 ints (x, y, z, w, d, h, item_type); an AABB is
 * 6 ints (x, y, z, x2, y2, z2).
 */
#ifndef PLP_ORACLE_H
#define PLP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int l, w, h;          /* item dimensions */
    int lf, wf, hf;       /* may stand on that axis (0/1) */
    int qty;
} orc_item;

typedef struct {
    int L, W, H;
    int n_types;
    orc_item *items;
} orc_instance;

/* thpack "original" format, parser.py:67-139 / plp_common.hpp:68-141.
 * Returns the number of instances (>= 1) or -1 on a format error (err
 * receives the message).  Free with orc_free_instances. */
int orc_parse(const char *text, orc_instance **out, char *err, int errlen);
void orc_free_instances(orc_instance *v, int n);

/* Orientation list, deduplicated in parser order: plp_common.hpp:269-280,
 * solver.py:75-83 + parser.py:142-156.  out[6][3] = (w, d, h); returns count. */
int orc_orientations(const orc_item *it, int out[6][3]);

/* (type, copy) step sequence, plp_common.hpp:335-351 / solver.py:169-177.
 * order 0 = volume, 1 = input.  Writes the item type of every step; returns
 * the step count. */
int orc_item_sequence(const orc_instance *inst, int order, int *step_type, int cap);

/* CPython random.Random(int) MT19937, plp_common.hpp:164-250. */
typedef struct {
    uint32_t mt[624];
    int mti;
} orc_rng;

void orc_rng_seed(orc_rng *r, uint64_t seed);
uint32_t orc_rng_next(orc_rng *r);
double orc_rng_random(orc_rng *r);
uint32_t orc_rng_getrandbits(orc_rng *r, int k);
int orc_rng_randrange(orc_rng *r, int n);

/* C2 -> C3 -> C1 predicate, plp_common.hpp:297-321 / solver.py:198-218. */
int orc_feasible_one(int x, int y, int z, int w, int d, int h,
                     const int *aabb6, int n_aabb, int L, int W, int H);

/* bits[ci * n_ors + oi] for every pair, plp_gpu.cpp:18-32. */
void orc_check_all(const int *cand_xyz, int n_cand, const int *ors_wdh, int n_ors,
                   const int *aabb6, int n_aabb, int L, int W, int H, uint8_t *bits);

typedef struct {
    uint64_t pair_box;   /* sum over steps of n_cand * n_ors * n_aabb */
    uint64_t pairs;      /* sum over steps of n_cand * n_ors */
    int steps;           /* steps that evaluated feasibility */
    int max_cand;        /* largest extreme-point list seen */
} orc_work;

/* One decode pass, plp_common.hpp:404-456 / solver.py:180-240.  placed7 gets
 * (x, y, z, w, d, h, type) per box; returns the number placed, -1 if cap is
 * too small.  work may be NULL. */
int orc_decode_pass(const orc_instance *inst, int order, int first_fit, double eps,
                    orc_rng *rng, int *placed7, int cap, orc_work *work);

/* Per-box FSR, plp_common.hpp:356-390 / solver.py:109-141. */
void orc_compute_fsr(const int *placed7, int n, double mu, double *fsr);
/* Fraction of boxes with FSR >= 1, plp_common.hpp:393-399. */
double orc_stability_score(const int *placed7, int n, double mu);

typedef struct {
    int n_boxes;
    long long volume;
    double score;
    int fallback;
    int n_passes;
    uint64_t seed;
    int *placements;     /* n_boxes * 7 in JSON order (type, x, y, z, w, d, h) */
    orc_work work;       /* summed over all passes */
} orc_result;

/* Fixed-pass run + D4' selection, plp_common.hpp:471-523 / solver.py:243-271
 * (the pass loop stops after n_passes instead of a wall budget).  use_theta 0
 * accepts every packing.  placements must hold 7 * (step count) ints. */
int orc_run(const orc_instance *inst, uint64_t seed, int n_passes, double eps,
            int use_theta, double theta, int order, int first_fit, double mu,
            orc_result *out);

/* ---- block-building engine (blockproof/plp_block.cpp), plp_block_oracle.c */

/* One block shape: cnt = a*b*k copies of item ti in orientation oi as a
 * bw x bd x bh stack (BlockShape, plp_block.cpp:27-33). */
typedef struct {
    int ti, oi, a, b, k;
    long long cnt;
    int bw, bd, bh;
} orc_shape;

/* Catalogue of run_block (plp_block.cpp:193-218): shapes in generation order;
 * by_type[type_off[t] .. type_off[t+1]) lists type t's shapes in
 * shape_precedence order (plp_block.cpp:38-45).  Returns the shape count or
 * -1 if cap is too small. */
int orc_block_catalogue(const orc_instance *inst, int amax, int bmax, int kmax, orc_shape *cat,
                        int cap, int *by_type, int *type_off);

/* block_pass (plp_block.cpp:95-204); pick_rule 0 first, 1 support, 2 contact.
 * items7 gets the decomposed items (x, y, z, w, d, h, type); returns their
 * count or -1. */
int orc_block_pass(const orc_instance *inst, const orc_shape *cat, const int *by_type,
                   const int *type_off, double eps, int pick_rule, int topk, orc_rng *rng,
                   int *items7, int cap, long long *volume);

/* Fixed-pass run_block with online D4' (plp_block.cpp:220-278). */
int orc_block_run(const orc_instance *inst, uint64_t seed, int n_passes, double eps,
                  int use_theta, double theta, int amax, int bmax, int kmax, int topk,
                  int pick_rule, double mu, orc_result *out);

/* CPython float repr, plp_common.hpp:535-590.  Returns the length. */
int orc_py_repr(double x, char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif
