/* plp_block_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's block-building engine
 * (blockproof/plp_block.cpp), the parity checker for the HIP block decoder.
 * It reuses the decoder oracle's parser, orientations, MT19937, feasibility
 * predicate and FSR (plp_oracle.c).  Each function cites the reference
 * file:line it follows.
 */
#include <stdlib.h>
#include <string.h>

#include "plp_oracle.h"

/* shape_precedence, plp_block.cpp:38-45: most copies first, then tallest
 * stack, then widest base, ties by orientation index. */
/* per thread: the comparators below read them, and callers may run
 * decoders on several threads at once */
static _Thread_local const orc_shape *g_cat;

static int cmp_precedence(const void *pa, const void *pb) {
    const orc_shape *x = &g_cat[*(const int *)pa], *y = &g_cat[*(const int *)pb];
    if (x->cnt != y->cnt) return x->cnt > y->cnt ? -1 : 1;
    if (x->bh != y->bh) return x->bh > y->bh ? -1 : 1;
    if (x->a != y->a) return x->a > y->a ? -1 : 1;
    if (x->b != y->b) return x->b > y->b ? -1 : 1;
    return x->oi < y->oi ? -1 : (x->oi > y->oi);
}

/* Placement order, plp_block.cpp:156-165: block volume desc, then type,
 * orientation asc, stack height, base a, base b desc. */
static int cmp_place(const void *pa, const void *pb) {
    const orc_shape *s = &g_cat[*(const int *)pa], *t = &g_cat[*(const int *)pb];
    const long long vs = (long long)s->bw * s->bd * s->bh, vt = (long long)t->bw * t->bd * t->bh;
    if (vs != vt) return vs > vt ? -1 : 1;
    if (s->ti != t->ti) return s->ti < t->ti ? -1 : 1;
    if (s->oi != t->oi) return s->oi < t->oi ? -1 : 1;
    if (s->k != t->k) return s->k > t->k ? -1 : 1;
    if (s->a != t->a) return s->a > t->a ? -1 : 1;
    if (s->b != t->b) return s->b > t->b ? -1 : 1;
    return 0;
}

static _Thread_local const orc_instance *g_inst;

/* Type order of the plan, plp_block.cpp:134-138: item volume desc, index asc. */
static int cmp_type_volume(const void *pa, const void *pb) {
    const int x = *(const int *)pa, y = *(const int *)pb;
    const orc_item *a = &g_inst->items[x], *b = &g_inst->items[y];
    const long long vx = (long long)a->l * a->w * a->h, vy = (long long)b->l * b->w * b->h;
    if (vx != vy) return vx > vy ? -1 : 1;
    return x < y ? -1 : (x > y);
}

/* Catalogue of run_block, plp_block.cpp:193-218. */
int orc_block_catalogue(const orc_instance *inst, int amax, int bmax, int kmax, orc_shape *cat,
                        int cap, int *by_type, int *type_off) {
    int n = 0;
    for (int ti = 0; ti < inst->n_types; ti++) {
        int ors[6][3];
        const int no = orc_orientations(&inst->items[ti], ors);
        type_off[ti] = n;
        for (int oi = 0; oi < no; oi++) {
            const int w = ors[oi][0], d = ors[oi][1], h = ors[oi][2];
            for (int a = 1; a <= amax && a * w <= inst->L; a++)
                for (int b = 1; b <= bmax && b * d <= inst->W; b++)
                    for (int k = 1; k <= kmax && k * h <= inst->H; k++) {
                        const long long cnt = (long long)a * b * k;
                        if (cnt > inst->items[ti].qty) break;
                        if (n >= cap) return -1;
                        orc_shape *s = &cat[n];
                        s->ti = ti; s->oi = oi; s->a = a; s->b = b; s->k = k; s->cnt = cnt;
                        s->bw = a * w; s->bd = b * d; s->bh = k * h;
                        by_type[n] = n;
                        n++;
                    }
        }
    }
    type_off[inst->n_types] = n;
    g_cat = cat;
    for (int ti = 0; ti < inst->n_types; ti++)
        qsort(by_type + type_off[ti], (size_t)(type_off[ti + 1] - type_off[ti]), sizeof(int),
              cmp_precedence);
    return n;
}

static int cmp_pt3(const void *a, const void *b) {
    const int *p = (const int *)a, *q = (const int *)b;
    if (p[2] != q[2]) return p[2] < q[2] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    return 0;
}

/* contact_stats, plp_block.cpp:64-81. */
static void contact_stats(int x, int y, int z, const orc_shape *s, const int *aabb, int n_box,
                          long long *bottom, long long *side) {
    const long long bw = s->bw, bd = s->bd, bh = s->bh;
    long long bt = z == 0 ? bw * bd : 0, sd = 0;
    for (int i = 0; i < n_box; i++) {
        const int *b = aabb + 6 * i;
        const long long ox = (x + bw < b[3] ? x + bw : b[3]) - (x > b[0] ? x : b[0]);
        const long long oy = (y + bd < b[4] ? y + bd : b[4]) - (y > b[1] ? y : b[1]);
        const long long oz = (z + bh < b[5] ? z + bh : b[5]) - (z > b[2] ? z : b[2]);
        if (b[5] == z && ox > 0 && oy > 0) bt += ox * oy;
        if (oy > 0 && oz > 0 && (b[3] == x || b[0] == x + bw)) sd += oy * oz;
        if (ox > 0 && oz > 0 && (b[4] == y || b[1] == y + bd)) sd += ox * oz;
    }
    *bottom = bt;
    *side = sd;
}

/* block_pass, plp_block.cpp:95-186.  items7 receives the decomposed item
 * boxes (x, y, z, w, d, h, type) in block order; returns their count, or -1
 * when a buffer is too small. */
int orc_block_pass(const orc_instance *inst, const orc_shape *cat, const int *by_type,
                   const int *type_off, double eps, int pick_rule, int topk, orc_rng *rng,
                   int *items7, int cap, long long *volume) {
    const int L = inst->L, W = inst->W, H = inst->H, nt = inst->n_types;
    long long total = 0;
    for (int t = 0; t < nt; t++) total += inst->items[t].qty;
    const int max_blocks = (int)(total > 0 ? total : 1);
    int *rem = malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
    int *has_ors = malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
    int *types = malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
    int *plan = malloc(sizeof(int) * (size_t)max_blocks);
    int *blk = malloc(sizeof(int) * 10 * (size_t)max_blocks);   /* x y z w d h ti a b k */
    int *aabb = malloc(sizeof(int) * 6 * (size_t)max_blocks);
    const int cand_cap = 1 + 8 * (max_blocks + 1);
    int *cand = malloc(sizeof(int) * 3 * (size_t)cand_cap);
    int *nc = malloc(sizeof(int) * 3 * (size_t)cand_cap);
    int *feas = malloc(sizeof(int) * (size_t)cand_cap);
    int n_blk = 0, n_cand = 1, rc = 0;
    cand[0] = cand[1] = cand[2] = 0;
    for (int t = 0; t < nt; t++) {
        int ors[6][3];
        rem[t] = inst->items[t].qty;
        has_ors[t] = orc_orientations(&inst->items[t], ors) > 0;
    }
    g_inst = inst;
    g_cat = cat;
    for (;;) {
        /* plan (plp_block.cpp:129-153) */
        int n_types = 0;
        for (int t = 0; t < nt; t++)
            if (rem[t] > 0 && has_ors[t]) types[n_types++] = t;
        if (n_types == 0) break;
        qsort(types, (size_t)n_types, sizeof(int), cmp_type_volume);
        if (eps > 0.0 && orc_rng_random(rng) >= 1.0 - eps)
            for (int i = n_types - 1; i > 0; i--) {
                const int j = orc_rng_randrange(rng, i + 1);
                const int tmp = types[i];
                types[i] = types[j];
                types[j] = tmp;
            }
        int n_plan = 0;
        for (int q = 0; q < n_types; q++) {
            const int t = types[q];
            long long r = rem[t];
            const int *shp = by_type + type_off[t];
            const int size = type_off[t + 1] - type_off[t];
            while (r > 0) {
                int start = size;
                while (start > 0 && cat[shp[start - 1]].cnt <= r) start--;
                if (start == size) break;
                const int tk = topk < size - start ? topk : size - start;
                int pk = 0;
                if (eps > 0.0 && orc_rng_random(rng) >= 1.0 - eps) pk = orc_rng_randrange(rng, tk);
                if (n_plan >= max_blocks) { rc = -1; goto out; }
                plan[n_plan++] = shp[start + pk];
                r -= cat[shp[start + pk]].cnt;
            }
        }
        if (n_plan == 0) break;
        qsort(plan, (size_t)n_plan, sizeof(int), cmp_place);
        /* placement (plp_block.cpp:168-183) */
        int any = 0;
        for (int q = 0; q < n_plan; q++) {
            const orc_shape *s = &cat[plan[q]];
            if (rem[s->ti] < s->cnt) continue;
            int nf = 0;
            for (int i = 0; i < n_cand; i++)
                if (orc_feasible_one(cand[3 * i], cand[3 * i + 1], cand[3 * i + 2], s->bw, s->bd,
                                     s->bh, aabb, n_blk, L, W, H))
                    feas[nf++] = i;
            if (nf == 0) continue;
            int pick;
            if (eps > 0.0 && orc_rng_random(rng) >= 1.0 - eps) {
                pick = feas[orc_rng_randrange(rng, nf)];
            } else if (pick_rule == 0) {
                pick = feas[0];
            } else {
                long long k0 = -1, k1 = -1;
                int bi = 0;
                for (int i = 0; i < nf; i++) {
                    const int *p = cand + 3 * feas[i];
                    long long sup, side;
                    contact_stats(p[0], p[1], p[2], s, aabb, n_blk, &sup, &side);
                    const long long p0 = pick_rule == 1 ? sup : sup + side;
                    const long long p1 = pick_rule == 1 ? side : sup;
                    if (p0 > k0 || (p0 == k0 && p1 > k1)) { k0 = p0; k1 = p1; bi = i; }
                }
                pick = feas[bi];
            }
            /* place (plp_block.cpp:109-127) */
            const int x = cand[3 * pick], y = cand[3 * pick + 1], z = cand[3 * pick + 2];
            int *b = blk + 10 * n_blk;
            b[0] = x; b[1] = y; b[2] = z; b[3] = s->bw; b[4] = s->bd; b[5] = s->bh;
            b[6] = s->ti; b[7] = s->a; b[8] = s->b; b[9] = s->k;
            int *ab = aabb + 6 * n_blk;
            ab[0] = x; ab[1] = y; ab[2] = z; ab[3] = x + s->bw; ab[4] = y + s->bd; ab[5] = z + s->bh;
            n_blk++;
            int m = 0;
            for (int i = 0; i < n_cand; i++) {
                const int px = cand[3 * i], py = cand[3 * i + 1], pz = cand[3 * i + 2];
                if (x < px && px < x + s->bw && y < py && py < y + s->bd && z < pz && pz < z + s->bh)
                    continue;
                nc[3 * m] = px; nc[3 * m + 1] = py; nc[3 * m + 2] = pz;
                m++;
            }
            for (int cx = 0; cx < 2; cx++)
                for (int cy = 0; cy < 2; cy++)
                    for (int cz = 0; cz < 2; cz++) {
                        nc[3 * m] = cx ? x + s->bw : x;
                        nc[3 * m + 1] = cy ? y + s->bd : y;
                        nc[3 * m + 2] = cz ? z + s->bh : z;
                        m++;
                    }
            qsort(nc, (size_t)m, 3 * sizeof(int), cmp_pt3);
            n_cand = 0;
            for (int i = 0; i < m; i++) {
                if (n_cand > 0 && cmp_pt3(&nc[3 * i], &cand[3 * (n_cand - 1)]) == 0) continue;
                memcpy(&cand[3 * n_cand], &nc[3 * i], 3 * sizeof(int));
                n_cand++;
            }
            rem[s->ti] -= (int)s->cnt;
            any = 1;
        }
        if (!any) break;
    }
    /* decompose (plp_block.cpp:188-201) */
    {
        int n = 0;
        long long vol = 0;
        for (int q = 0; q < n_blk; q++) {
            const int *b = blk + 10 * q;
            const int iw = b[3] / b[7], id = b[4] / b[8], ih = b[5] / b[9];
            for (int i = 0; i < b[7]; i++)
                for (int j = 0; j < b[8]; j++)
                    for (int t = 0; t < b[9]; t++) {
                        if (n >= cap) { rc = -1; goto out; }
                        int *o = items7 + 7 * n;
                        o[0] = b[0] + i * iw; o[1] = b[1] + j * id; o[2] = b[2] + t * ih;
                        o[3] = iw; o[4] = id; o[5] = ih; o[6] = b[6];
                        vol += (long long)iw * id * ih;
                        n++;
                    }
        }
        *volume = vol;
        rc = n;
    }
out:
    free(rem); free(has_ors); free(types); free(plan); free(blk); free(aabb);
    free(cand); free(nc); free(feas);
    return rc;
}

/* Fixed-pass run_block, plp_block.cpp:220-278 (D4' kept online; the pass loop
 * stops after n_passes instead of a wall budget).  out->placements must hold
 * 7 * (total item count) ints; out->work is not filled. */
int orc_block_run(const orc_instance *inst, uint64_t seed, int n_passes, double eps,
                  int use_theta, double theta, int amax, int bmax, int kmax, int topk,
                  int pick_rule, double mu, orc_result *out) {
    long long total = 0;
    for (int t = 0; t < inst->n_types; t++) total += inst->items[t].qty;
    const int cap = (int)(total > 0 ? total : 1);
    const int cat_cap = inst->n_types * 6 * amax * bmax * kmax + 1;
    orc_shape *cat = malloc(sizeof(orc_shape) * (size_t)cat_cap);
    int *by_type = malloc(sizeof(int) * (size_t)cat_cap);
    int *type_off = malloc(sizeof(int) * (size_t)(inst->n_types + 1));
    int *cur = malloc(sizeof(int) * 7 * (size_t)cap);
    int *best_acc = malloc(sizeof(int) * 7 * (size_t)cap);
    int *best_all = malloc(sizeof(int) * 7 * (size_t)cap);
    int rc = 0;
    if (orc_block_catalogue(inst, amax, bmax, kmax, cat, cat_cap, by_type, type_off) < 0) {
        rc = -1;
        goto done;
    }
    orc_rng rng;
    orc_rng_seed(&rng, seed);
    int n_acc = -1, n_all = -1;
    long long v_acc = -1, v_all = -1;
    double s_acc = 0.0, s_all = 0.0;
    memset(&out->work, 0, sizeof(out->work));
    for (int p = 0; p < n_passes; p++) {
        long long vol = 0;
        const int n = orc_block_pass(inst, cat, by_type, type_off, eps, pick_rule, topk, &rng, cur,
                                     cap, &vol);
        if (n < 0) { rc = -1; goto done; }
        const double sc = orc_stability_score(cur, n, mu);
        const int upd_all = vol > v_all || (vol == v_all && n > n_all);
        const int upd_acc = (!use_theta || sc >= theta) &&
                            (n_acc < 0 || vol > v_acc || (vol == v_acc && n > n_acc));
        if (upd_acc) {
            n_acc = n; v_acc = vol; s_acc = sc;
            memcpy(best_acc, cur, sizeof(int) * 7 * (size_t)n);
        }
        if (upd_all) {
            n_all = n; v_all = vol; s_all = sc;
            memcpy(best_all, cur, sizeof(int) * 7 * (size_t)n);
        }
    }
    {
        const int *src = n_acc >= 0 ? best_acc : best_all;
        out->n_boxes = n_acc >= 0 ? n_acc : n_all;
        out->volume = n_acc >= 0 ? v_acc : v_all;
        out->score = n_acc >= 0 ? s_acc : s_all;
        out->fallback = n_acc < 0;
        out->n_passes = n_passes;
        out->seed = seed;
        for (int i = 0; i < out->n_boxes; i++) {
            const int *b = src + 7 * i;
            int *o = out->placements + 7 * i;
            o[0] = b[6]; o[1] = b[0]; o[2] = b[1]; o[3] = b[2]; o[4] = b[3]; o[5] = b[4]; o[6] = b[5];
        }
    }
done:
    free(cat); free(by_type); free(type_off); free(cur); free(best_acc); free(best_all);
    return rc;
}
