/* plp_oracle.c -- TEST INFRASTRUCTURE ONLY: the parity checker.
 *
 * A plain-C restatement of the reference decoder (plp/solver.py, restated
 * bit-compatibly in gpu/plp_common.hpp).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  Build with -ffp-contract=off
 * (gpu/Makefile:6-10): every double expression keeps the reference's operand
 * order and no multiply-add is fused.
 *
 * Pinned by (tests/test_oracle_*.py):
 *   parse           parser KATs (tests/test_parser.py:31-157) + manifest stats
 *   orientations    orientation KATs (tests/test_solver.py:58-68)
 *   rng             CPython random.Random vectors written by tests/golden/make_golden.py
 *   feasible_one    closed forms (tests/test_geometry.py:15-85, 2s<wd) + brute force
 *   decode_pass     the parity/ref_{cpu,host,gpu} JSON files + the reference g++ build
 *   compute_fsr     hand KATs (tests/test_solver.py:107-145) + reference per-box bits
 *   run             reference g++ build driven for a fixed pass count
 *   py_repr         CPython repr() vectors
 */
#include "plp_oracle.h"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ parser */

typedef struct {
    char **tok;   /* tokens of one non-blank line */
    int n;
} line_t;

static void set_err(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = '\0';
    }
}

/* strtoll with the whole token consumed, as plp_common.hpp:86-92. */
static int tok_int(const char *s, long long *v) {
    char *end = NULL;
    errno = 0;
    long long r = strtoll(s, &end, 10);
    if (end == s || *end != '\0') return -1;
    *v = r;
    return 0;
}

static void free_lines(line_t *lines, int n) {
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < lines[i].n; j++) free(lines[i].tok[j]);
        free(lines[i].tok);
    }
    free(lines);
}

/* parser.py:67-139 (format and error cases) in the token form of
 * plp_common.hpp:68-141. */
int orc_parse(const char *text, orc_instance **out, char *err, int errlen) {
    int cap = 64, nl = 0;
    line_t *lines = (line_t *)calloc((size_t)cap, sizeof(line_t));
    const char *p = text;
    while (*p) {
        const char *e = p;
        while (*e && *e != '\n') e++;
        line_t ln = {NULL, 0};
        int tcap = 0;
        const char *q = p;
        while (q < e) {
            while (q < e && isspace((unsigned char)*q)) q++;
            if (q >= e) break;
            const char *s = q;
            while (q < e && !isspace((unsigned char)*q)) q++;
            if (ln.n == tcap) {
                tcap = tcap ? tcap * 2 : 8;
                ln.tok = (char **)realloc(ln.tok, (size_t)tcap * sizeof(char *));
            }
            char *t = (char *)malloc((size_t)(q - s) + 1);
            memcpy(t, s, (size_t)(q - s));
            t[q - s] = '\0';
            ln.tok[ln.n++] = t;
        }
        if (ln.n > 0) {
            if (nl == cap) {
                cap *= 2;
                lines = (line_t *)realloc(lines, (size_t)cap * sizeof(line_t));
            }
            lines[nl++] = ln;
        }
        p = *e ? e + 1 : e;
    }
    int rc = -1;
    orc_instance *insts = NULL;
    long long P = 0;
    int pos = 1, made = 0;
    if (nl == 0) { set_err(err, errlen, "empty file"); goto done; }
    if (tok_int(lines[0].tok[0], &P) != 0) { set_err(err, errlen, "line 0 must be an integer count"); goto done; }
    if (P <= 0) { set_err(err, errlen, "instance count must be >= 1"); goto done; }
    if (P > 1000000) { set_err(err, errlen, "instance count too large"); goto done; }
    insts = (orc_instance *)calloc((size_t)P, sizeof(orc_instance));
    for (long long k = 0; k < P; k++) {
        if (pos >= nl || lines[pos].n > 2) { set_err(err, errlen, "malformed problem line"); goto done; }
        pos++;
        if (pos >= nl || lines[pos].n != 3) { set_err(err, errlen, "container line must have 3 dimensions"); goto done; }
        long long L, W, H;
        if (tok_int(lines[pos].tok[0], &L) || tok_int(lines[pos].tok[1], &W) || tok_int(lines[pos].tok[2], &H)) {
            set_err(err, errlen, "not an integer (container dimension)"); goto done;
        }
        if (L <= 0 || W <= 0 || H <= 0) { set_err(err, errlen, "container dimensions must be positive"); goto done; }
        pos++;
        if (pos >= nl || lines[pos].n != 1) { set_err(err, errlen, "type-count line must be a single integer"); goto done; }
        long long n;
        if (tok_int(lines[pos].tok[0], &n)) { set_err(err, errlen, "not an integer (type count)"); goto done; }
        pos++;
        orc_instance *I = &insts[made];
        I->L = (int)L; I->W = (int)W; I->H = (int)H;
        I->n_types = 0;
        I->items = (orc_item *)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_item));
        made++;
        for (long long t = 1; t <= n; t++) {
            if (pos >= nl) { set_err(err, errlen, "instance truncated"); goto done; }
            line_t *ln = &lines[pos++];
            if (ln->n != 8) { set_err(err, errlen, "item line must have 8 tokens"); goto done; }
            long long idx, dl, dw, dh, qty;
            if (tok_int(ln->tok[0], &idx)) { set_err(err, errlen, "not an integer (item index)"); goto done; }
            if (idx != t) { set_err(err, errlen, "item index must be sequential"); goto done; }
            if (tok_int(ln->tok[1], &dl) || tok_int(ln->tok[3], &dw) || tok_int(ln->tok[5], &dh)) {
                set_err(err, errlen, "not an integer (item dimension)"); goto done;
            }
            if (dl <= 0 || dw <= 0 || dh <= 0) { set_err(err, errlen, "item dimensions must be positive"); goto done; }
            for (int f = 2; f <= 6; f += 2)
                if (strcmp(ln->tok[f], "0") != 0 && strcmp(ln->tok[f], "1") != 0) {
                    set_err(err, errlen, "orientation flags must be 0/1"); goto done;
                }
            if (tok_int(ln->tok[7], &qty)) { set_err(err, errlen, "not an integer (item qty)"); goto done; }
            if (qty < 0) { set_err(err, errlen, "item qty must be >= 0"); goto done; }
            orc_item *it = &I->items[I->n_types++];
            it->l = (int)dl; it->w = (int)dw; it->h = (int)dh;
            it->lf = ln->tok[2][0] == '1';
            it->wf = ln->tok[4][0] == '1';
            it->hf = ln->tok[6][0] == '1';
            it->qty = (int)qty;
        }
    }
    if (pos != nl) { set_err(err, errlen, "trailing data after last instance"); goto done; }
    rc = made;
    *out = insts;
    insts = NULL;
done:
    if (insts) orc_free_instances(insts, made);
    free_lines(lines, nl);
    return rc;
}

void orc_free_instances(orc_instance *v, int n) {
    if (!v) return;
    for (int i = 0; i < n; i++) free(v[i].items);
    free(v);
}

/* ------------------------------------------------------------ orientations */

/* plp_common.hpp:269-280: lf -> (w,h,l),(h,w,l); wf -> (l,h,w),(h,l,w);
 * hf -> (l,w,h),(w,l,h); duplicates dropped, first occurrence kept. */
int orc_orientations(const orc_item *it, int out[6][3]) {
    int cand[6][3], nc = 0, n = 0;
    if (it->lf) {
        cand[nc][0] = it->w; cand[nc][1] = it->h; cand[nc][2] = it->l; nc++;
        cand[nc][0] = it->h; cand[nc][1] = it->w; cand[nc][2] = it->l; nc++;
    }
    if (it->wf) {
        cand[nc][0] = it->l; cand[nc][1] = it->h; cand[nc][2] = it->w; nc++;
        cand[nc][0] = it->h; cand[nc][1] = it->l; cand[nc][2] = it->w; nc++;
    }
    if (it->hf) {
        cand[nc][0] = it->l; cand[nc][1] = it->w; cand[nc][2] = it->h; nc++;
        cand[nc][0] = it->w; cand[nc][1] = it->l; cand[nc][2] = it->h; nc++;
    }
    for (int i = 0; i < nc; i++) {
        int dup = 0;
        for (int j = 0; j < n; j++)
            if (out[j][0] == cand[i][0] && out[j][1] == cand[i][1] && out[j][2] == cand[i][2]) dup = 1;
        if (!dup) {
            out[n][0] = cand[i][0]; out[n][1] = cand[i][1]; out[n][2] = cand[i][2];
            n++;
        }
    }
    return n;
}

/* ----------------------------------------------------------- item sequence */

typedef struct {
    long long vol;
    int type, copy;
} step_t;

static int cmp_step_volume(const void *a, const void *b) {
    const step_t *x = (const step_t *)a, *y = (const step_t *)b;
    if (x->vol != y->vol) return x->vol > y->vol ? -1 : 1;   /* -volume */
    if (x->type != y->type) return x->type < y->type ? -1 : 1;
    if (x->copy != y->copy) return x->copy < y->copy ? -1 : 1;
    return 0;
}

/* plp_common.hpp:335-351: key (-volume, type, copy) for 'volume'. */
int orc_item_sequence(const orc_instance *inst, int order, int *step_type, int cap) {
    long long total = 0;
    for (int t = 0; t < inst->n_types; t++) total += inst->items[t].qty;
    if (total > cap) return -1;
    step_t *s = (step_t *)malloc(sizeof(step_t) * (size_t)(total > 0 ? total : 1));
    int n = 0;
    for (int t = 0; t < inst->n_types; t++) {
        const orc_item *it = &inst->items[t];
        for (int c = 0; c < it->qty; c++) {
            s[n].vol = (long long)it->l * it->w * it->h;
            s[n].type = t;
            s[n].copy = c;
            n++;
        }
    }
    if (order == 0) qsort(s, (size_t)n, sizeof(step_t), cmp_step_volume);
    for (int i = 0; i < n; i++) step_type[i] = s[i].type;
    free(s);
    return n;
}

/* --------------------------------------------------------------------- rng */

/* CPython init_genrand / init_by_array (plp_common.hpp:204-232). */
static void rng_init_genrand(orc_rng *r, uint32_t s) {
    r->mt[0] = s;
    for (int i = 1; i < 624; i++)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->mti = 624;
}

/* random.Random(int seed): key = little-endian 32-bit chunks of |seed|
 * (plp_common.hpp:170-178). */
void orc_rng_seed(orc_rng *r, uint64_t seed) {
    uint32_t key[2];
    int keyused = 0;
    if (seed == 0) {
        key[0] = 0;
        keyused = 1;
    } else {
        while (seed) {
            key[keyused++] = (uint32_t)seed;
            seed >>= 32;
        }
    }
    rng_init_genrand(r, 19650218u);
    int i = 1, j = 0;
    int k = 624 > keyused ? 624 : keyused;
    for (; k; k--) {
        r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++;
        j++;
        if (i >= 624) { r->mt[0] = r->mt[623]; i = 1; }
        if (j >= keyused) j = 0;
    }
    for (k = 623; k; k--) {
        r->mt[i] = (r->mt[i] ^ ((r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= 624) { r->mt[0] = r->mt[623]; i = 1; }
    }
    r->mt[0] = 0x80000000u;
    r->mti = 624;
}

/* genrand_uint32: twist every 624 words, then temper (plp_common.hpp:234-249). */
uint32_t orc_rng_next(orc_rng *r) {
    if (r->mti >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
            r->mt[i] = r->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        r->mti = 0;
    }
    uint32_t y = r->mt[r->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* random(): (a*2^26 + b) / 2^53 with a = g>>5, b = g>>6 (plp_common.hpp:182-186). */
double orc_rng_random(orc_rng *r) {
    uint32_t a = orc_rng_next(r) >> 5;
    uint32_t b = orc_rng_next(r) >> 6;
    return (a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

uint32_t orc_rng_getrandbits(orc_rng *r, int k) { return orc_rng_next(r) >> (32 - k); }

/* randrange(n): k = bit_length(n), redraw while r >= n (plp_common.hpp:189-194). */
int orc_rng_randrange(orc_rng *r, int n) {
    int k = 0;
    for (unsigned t = (unsigned)n; t; t >>= 1) k++;
    int v = (int)orc_rng_getrandbits(r, k);
    while (v >= n) v = (int)orc_rng_getrandbits(r, k);
    return v;
}

/* ------------------------------------------------------------- feasibility */

static long long mn(long long a, long long b) { return a < b ? a : b; }
static long long mx(long long a, long long b) { return a < b ? b : a; }

/* plp_common.hpp:297-321: C2 bounds, C3 support >= TAU (integer areas, one
 * double division), C1 strict-interior overlap. */
int orc_feasible_one(int x, int y, int z, int w, int d, int h,
                     const int *aabb6, int n_aabb, int L, int W, int H) {
    if (x + w > L || y + d > W || z + h > H) return 0;                      /* C2 */
    if (z > 0) {
        long long s = 0;
        for (int i = 0; i < n_aabb; i++) {
            const int *b = aabb6 + 6 * i;
            if (b[5] == z) {
                long long ox = mn((long long)x + w, b[3]) - mx(x, b[0]);
                if (ox > 0) {
                    long long oy = mn((long long)y + d, b[4]) - mx(y, b[1]);
                    if (oy > 0) s += ox * oy;
                }
            }
        }
        if ((double)s / (double)((long long)w * d) < 0.5) return 0;          /* C3 */
    }
    for (int i = 0; i < n_aabb; i++) {
        const int *b = aabb6 + 6 * i;
        if (x < b[3] && b[0] < x + w && y < b[4] && b[1] < y + d && z < b[5] && b[2] < z + h)
            return 0;                                                        /* C1 */
    }
    return 1;
}

void orc_check_all(const int *cand_xyz, int n_cand, const int *ors_wdh, int n_ors,
                   const int *aabb6, int n_aabb, int L, int W, int H, uint8_t *bits) {
    for (int ci = 0; ci < n_cand; ci++)
        for (int oi = 0; oi < n_ors; oi++)
            bits[ci * n_ors + oi] = (uint8_t)orc_feasible_one(
                cand_xyz[3 * ci], cand_xyz[3 * ci + 1], cand_xyz[3 * ci + 2],
                ors_wdh[3 * oi], ors_wdh[3 * oi + 1], ors_wdh[3 * oi + 2],
                aabb6, n_aabb, L, W, H);
}

/* ------------------------------------------------------------- decode pass */

static int cmp_pt(const void *a, const void *b) {
    const int *p = (const int *)a, *q = (const int *)b;
    if (p[2] != q[2]) return p[2] < q[2] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    return 0;
}

/* plp_common.hpp:404-456.  The candidate list is kept sorted by (z, y, x) and
 * unique (sort_dedup_pt, plp_common.hpp:284-291). */
int orc_decode_pass(const orc_instance *inst, int order, int first_fit, double eps,
                    orc_rng *rng, int *placed7, int cap, orc_work *work) {
    const int L = inst->L, W = inst->W, H = inst->H;
    long long total = 0;
    for (int t = 0; t < inst->n_types; t++) total += inst->items[t].qty;
    int *seq = (int *)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    int n_steps = orc_item_sequence(inst, order, seq, (int)total);
    int (*ors)[6][3] = malloc(sizeof(int[6][3]) * (size_t)(inst->n_types > 0 ? inst->n_types : 1));
    int *n_ors = (int *)malloc(sizeof(int) * (size_t)(inst->n_types > 0 ? inst->n_types : 1));
    for (int t = 0; t < inst->n_types; t++) n_ors[t] = orc_orientations(&inst->items[t], ors[t]);

    int cand_cap = 1 + 8 * (n_steps + 1);
    int *cand = (int *)malloc(sizeof(int) * 3 * (size_t)cand_cap);
    int *nc = (int *)malloc(sizeof(int) * 3 * (size_t)cand_cap);
    int *aabb = (int *)malloc(sizeof(int) * 6 * (size_t)(n_steps > 0 ? n_steps : 1));
    uint8_t *bits = (uint8_t *)malloc((size_t)cand_cap * 6);
    int *feas = (int *)malloc(sizeof(int) * (size_t)cand_cap * 6);
    int n_cand = 1, n_box = 0, rc = 0;
    cand[0] = cand[1] = cand[2] = 0;

    for (int s = 0; s < n_steps; s++) {
        int ti = seq[s];
        int no = n_ors[ti];
        if (no == 0) continue;
        orc_check_all(cand, n_cand, &ors[ti][0][0], no, aabb, n_box, L, W, H, bits);
        if (work) {
            work->pair_box += (uint64_t)n_cand * (uint64_t)no * (uint64_t)n_box;
            work->pairs += (uint64_t)n_cand * (uint64_t)no;
            work->steps++;
            if (n_cand > work->max_cand) work->max_cand = n_cand;
        }
        int nf = 0;
        for (int p = 0; p < n_cand * no; p++)
            if (bits[p]) feas[nf++] = p;                    /* candidate-major */
        if (nf == 0) continue;
        int pick = feas[0];
        if (!(first_fit || eps == 0.0)) {
            if (orc_rng_random(rng) >= 1.0 - eps) pick = feas[orc_rng_randrange(rng, nf)];
        }
        int ci = pick / no, oi = pick % no;
        int x = cand[3 * ci], y = cand[3 * ci + 1], z = cand[3 * ci + 2];
        int w = ors[ti][oi][0], d = ors[ti][oi][1], h = ors[ti][oi][2];
        if (n_box >= cap) { rc = -1; goto out; }
        int *pb = placed7 + 7 * n_box;
        pb[0] = x; pb[1] = y; pb[2] = z; pb[3] = w; pb[4] = d; pb[5] = h; pb[6] = ti;
        int *ab = aabb + 6 * n_box;
        ab[0] = x; ab[1] = y; ab[2] = z; ab[3] = x + w; ab[4] = y + d; ab[5] = z + h;
        n_box++;
        /* extreme points: drop those strictly inside the new box, add its 8
         * corners, sort + unique (plp_common.hpp:439-453). */
        int m = 0;
        for (int i = 0; i < n_cand; i++) {
            int px = cand[3 * i], py = cand[3 * i + 1], pz = cand[3 * i + 2];
            if (x < px && px < x + w && y < py && py < y + d && z < pz && pz < z + h) continue;
            nc[3 * m] = px; nc[3 * m + 1] = py; nc[3 * m + 2] = pz;
            m++;
        }
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) {
                    nc[3 * m] = a ? x + w : x;
                    nc[3 * m + 1] = b ? y + d : y;
                    nc[3 * m + 2] = c ? z + h : z;
                    m++;
                }
        qsort(nc, (size_t)m, 3 * sizeof(int), cmp_pt);
        n_cand = 0;
        for (int i = 0; i < m; i++) {
            if (n_cand > 0 && cmp_pt(&nc[3 * i], &cand[3 * (n_cand - 1)]) == 0) continue;
            cand[3 * n_cand] = nc[3 * i];
            cand[3 * n_cand + 1] = nc[3 * i + 1];
            cand[3 * n_cand + 2] = nc[3 * i + 2];
            n_cand++;
        }
    }
    rc = n_box;
out:
    free(seq); free(ors); free(n_ors); free(cand); free(nc); free(aabb); free(bits); free(feas);
    return rc;
}

/* --------------------------------------------------------------------- FSR */

/* (z desc, index asc) == Python sorted(range(n), key=-z), stable: keys
 * (INT_MAX - z) << 32 | index sort ascending into that order (no shared
 * state, so concurrent calls are safe). */
static int cmp_u64(const void *a, const void *b) {
    unsigned long long x = *(const unsigned long long *)a, y = *(const unsigned long long *)b;
    return x < y ? -1 : (x > y);
}

/* plp_common.hpp:356-390 / solver.py:109-141. */
void orc_compute_fsr(const int *placed7, int n, double mu, double *fsr) {
    if (n <= 0) return;
    double *m = (double *)malloc(sizeof(double) * (size_t)n);
    double *load = (double *)malloc(sizeof(double) * (size_t)n);
    int *idx = (int *)malloc(sizeof(int) * (size_t)n);
    int *aj = (int *)malloc(sizeof(int) * (size_t)n);
    long long *aa = (long long *)malloc(sizeof(long long) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const int *b = placed7 + 7 * i;
        m[i] = 50.0 * (double)((long long)b[3] * b[4] * b[5]) / 1.0e6;   /* RHO * V / 1e6 */
        load[i] = m[i];
        idx[i] = i;
    }
    unsigned long long *key = (unsigned long long *)malloc(sizeof(unsigned long long) * (size_t)n);
    for (int i = 0; i < n; i++)
        key[i] = ((unsigned long long)(0x7fffffffu - (unsigned)placed7[7 * i + 2]) << 32) | (unsigned)i;
    qsort(key, (size_t)n, sizeof *key, cmp_u64);
    for (int i = 0; i < n; i++) idx[i] = (int)(key[i] & 0xffffffffu);
    free(key);
    for (int t = 0; t < n; t++) {
        int k = idx[t];
        const int *bk = placed7 + 7 * k;
        int na = 0;
        long long total = 0;
        for (int j = 0; j < n; j++) {
            const int *bj = placed7 + 7 * j;
            if (j == k || bj[2] + bj[5] != bk[2]) continue;
            long long ox = mn((long long)bk[0] + bk[3], (long long)bj[0] + bj[3]) - mx(bk[0], bj[0]);
            if (ox <= 0) continue;
            long long oy = mn((long long)bk[1] + bk[4], (long long)bj[1] + bj[4]) - mx(bk[1], bj[1]);
            if (oy > 0) {
                aj[na] = j;
                aa[na] = ox * oy;
                na++;
            }
        }
        for (int q = 0; q < na; q++) total += aa[q];
        if (total > 0)
            for (int q = 0; q < na; q++)
                load[aj[q]] += load[k] * (double)aa[q] / (double)total;
    }
    for (int i = 0; i < n; i++) fsr[i] = (mu / 0.3) * load[i] / m[i];
    free(m); free(load); free(idx); free(aj); free(aa);
}

double orc_stability_score(const int *placed7, int n, double mu) {
    if (n <= 0) return 0.0;
    double *f = (double *)malloc(sizeof(double) * (size_t)n);
    orc_compute_fsr(placed7, n, mu, f);
    int c = 0;
    for (int i = 0; i < n; i++)
        if (f[i] >= 1.0) c++;
    free(f);
    return (double)c / (double)n;
}

/* --------------------------------------------------------------------- run */

/* plp_common.hpp:471-523 with a fixed pass count: one PyRng for all passes;
 * D4': accepted pool (score >= theta, or all when theta is off) if non-empty,
 * else every packing with fallback set; max (volume, n_boxes), strict >, so
 * the earliest winner is kept. */
int orc_run(const orc_instance *inst, uint64_t seed, int n_passes, double eps,
            int use_theta, double theta, int order, int first_fit, double mu,
            orc_result *out) {
    long long total = 0;
    for (int t = 0; t < inst->n_types; t++) total += inst->items[t].qty;
    int cap = (int)(total > 0 ? total : 1);
    orc_rng rng;
    orc_rng_seed(&rng, seed);
    int *cur = (int *)malloc(sizeof(int) * 7 * (size_t)cap);
    int *best_acc = (int *)malloc(sizeof(int) * 7 * (size_t)cap);
    int *best_all = (int *)malloc(sizeof(int) * 7 * (size_t)cap);
    int n_acc = -1, n_all = -1;
    long long v_acc = 0, v_all = 0;
    double s_acc = 0.0, s_all = 0.0;
    memset(&out->work, 0, sizeof(out->work));
    for (int p = 0; p < n_passes; p++) {
        int n = orc_decode_pass(inst, order, first_fit, eps, &rng, cur, cap, &out->work);
        if (n < 0) { free(cur); free(best_acc); free(best_all); return -1; }
        long long vol = 0;
        for (int i = 0; i < n; i++) vol += (long long)cur[7 * i + 3] * cur[7 * i + 4] * cur[7 * i + 5];
        double sc = orc_stability_score(cur, n, mu);
        if (n_all < 0 || vol > v_all || (vol == v_all && n > n_all)) {
            n_all = n; v_all = vol; s_all = sc;
            memcpy(best_all, cur, sizeof(int) * 7 * (size_t)n);
        }
        if (!use_theta || sc >= theta) {
            if (n_acc < 0 || vol > v_acc || (vol == v_acc && n > n_acc)) {
                n_acc = n; v_acc = vol; s_acc = sc;
                memcpy(best_acc, cur, sizeof(int) * 7 * (size_t)n);
            }
        }
    }
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
    free(cur); free(best_acc); free(best_all);
    return 0;
}

/* -------------------------------------------------------------------- repr */

/* CPython float repr (plp_common.hpp:535-590): the shortest digit string that
 * reads back to the same double, written fixed iff -4 < decpt <= 16 (value =
 * 0.d1d2.. * 10^decpt), else d[.ddd]e[+-]XX. */
int orc_py_repr(double x, char *buf, int len) {
    char tmp[64];
    if (x == 0.0) return snprintf(buf, (size_t)len, "%s", signbit(x) ? "-0.0" : "0.0");
    if (isnan(x)) return snprintf(buf, (size_t)len, "nan");
    if (isinf(x)) return snprintf(buf, (size_t)len, "%s", x < 0 ? "-inf" : "inf");
    int neg = x < 0;
    double ax = neg ? -x : x;
    int p;
    for (p = 1; p <= 17; p++) {
        snprintf(tmp, sizeof tmp, "%.*e", p - 1, ax);
        if (strtod(tmp, NULL) == ax) break;
    }
    char *e = strchr(tmp, 'e');
    int e10 = atoi(e + 1);
    char digits[32];
    int nd = 0;
    for (char *c = tmp; c < e; c++)
        if (*c != '.') digits[nd++] = *c;
    while (nd > 1 && digits[nd - 1] == '0') nd--;
    digits[nd] = '\0';
    int decpt = e10 + 1;
    char out[96];
    int o = 0;
    if (neg) out[o++] = '-';
    if (decpt > -4 && decpt <= 16) {
        if (decpt <= 0) {
            out[o++] = '0'; out[o++] = '.';
            for (int i = 0; i < -decpt; i++) out[o++] = '0';
            for (int i = 0; i < nd; i++) out[o++] = digits[i];
        } else if (decpt >= nd) {
            for (int i = 0; i < nd; i++) out[o++] = digits[i];
            for (int i = nd; i < decpt; i++) out[o++] = '0';
            out[o++] = '.'; out[o++] = '0';
        } else {
            for (int i = 0; i < decpt; i++) out[o++] = digits[i];
            out[o++] = '.';
            for (int i = decpt; i < nd; i++) out[o++] = digits[i];
        }
        out[o] = '\0';
    } else {
        int ex = decpt - 1;
        out[o++] = digits[0];
        if (nd > 1) {
            out[o++] = '.';
            for (int i = 1; i < nd; i++) out[o++] = digits[i];
        }
        o += snprintf(out + o, sizeof out - (size_t)o, "e%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    }
    return snprintf(buf, (size_t)len, "%s", out);
}
