/* plp_oracle_cli.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Command line over the C oracle, shaped like the reference CLI
 * (gpu/plp_gpu.cpp:80-143) with a fixed pass count instead of a wall budget:
 *
 *   plp_oracle <instance> [--seed N] [--passes P] [--eps F] [--theta F|--no-theta]
 *              [--order volume|input] [--first-fit] [--mu F] [--K K] [--threads T]
 *              [--bench]
 *
 * Prints one JSON line per decoder row (seeds seed..seed+K-1, time_s 0.0), or
 * with --bench a timing line (bench.py's cpu_baseline leg, kind "port").
 */
#include "plp_oracle.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    const orc_instance *inst;
    long long seed;
    int K, passes, use_theta, order, first_fit;
    double eps, theta, mu;
    int next;
    long long checksum;
    pthread_mutex_t mu_lock;
} job_t;

static void print_result(const orc_result *r) {
    char sc[64];
    orc_py_repr(r->score, sc, sizeof sc);
    printf("{\"n_boxes\": %d, \"volume_cm3\": %lld, \"stability_score\": %s, \"fallback\": %s, "
           "\"n_passes\": %d, \"time_s\": 0.0, \"seed\": %llu, \"placements\": [",
           r->n_boxes, r->volume, sc, r->fallback ? "true" : "false", r->n_passes,
           (unsigned long long)r->seed);
    for (int i = 0; i < r->n_boxes; i++) {
        const int *p = r->placements + 7 * i;
        printf("%s[%d, %d, %d, %d, %d, %d, %d]", i ? ", " : "", p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    }
    printf("]}\n");
}

static int total_qty(const orc_instance *inst) {
    int t = 0;
    for (int i = 0; i < inst->n_types; i++) t += inst->items[i].qty;
    return t > 0 ? t : 1;
}

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    orc_result r;
    r.placements = (int *)malloc(sizeof(int) * 7 * (size_t)total_qty(j->inst));
    for (;;) {
        pthread_mutex_lock(&j->mu_lock);
        int k = j->next++;
        pthread_mutex_unlock(&j->mu_lock);
        if (k >= j->K) break;
        orc_run(j->inst, (uint64_t)(j->seed + k), j->passes, j->eps, j->use_theta, j->theta,
                j->order, j->first_fit, j->mu, &r);
        pthread_mutex_lock(&j->mu_lock);
        j->checksum += r.volume;
        pthread_mutex_unlock(&j->mu_lock);
    }
    free(r.placements);
    return NULL;
}

static char *read_file(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)n + 1);
    size_t got = fread(buf, 1, (size_t)n, f);
    buf[got] = '\0';
    fclose(f);
    return buf;
}

int main(int argc, char **argv) {
    const char *path = NULL;
    job_t j;
    memset(&j, 0, sizeof j);
    j.K = 1; j.passes = 1; j.use_theta = 1; j.eps = 0.3; j.theta = 0.9; j.mu = 0.2;
    int threads = 1, bench = 0;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
#define NEXT() (i + 1 < argc ? argv[++i] : (fprintf(stderr, "missing value for %s\n", a), exit(2), ""))
        if (!strcmp(a, "--seed")) j.seed = atoll(NEXT());
        else if (!strcmp(a, "--passes")) j.passes = atoi(NEXT());
        else if (!strcmp(a, "--eps")) j.eps = strtod(NEXT(), NULL);
        else if (!strcmp(a, "--theta")) j.theta = strtod(NEXT(), NULL);
        else if (!strcmp(a, "--no-theta")) j.use_theta = 0;
        else if (!strcmp(a, "--order")) {
            const char *o = NEXT();
            if (!strcmp(o, "volume")) j.order = 0;
            else if (!strcmp(o, "input")) j.order = 1;
            else { fprintf(stderr, "unknown order %s\n", o); return 2; }
        } else if (!strcmp(a, "--first-fit")) j.first_fit = 1;
        else if (!strcmp(a, "--mu")) j.mu = strtod(NEXT(), NULL);
        else if (!strcmp(a, "--K")) j.K = atoi(NEXT());
        else if (!strcmp(a, "--threads")) threads = atoi(NEXT());
        else if (!strcmp(a, "--bench")) bench = 1;
        else if (a[0] != '-' && !path) path = a;
        else { fprintf(stderr, "unknown flag: %s\n", a); return 2; }
    }
    if (!path || j.K < 1 || j.passes < 1 || threads < 1) { fprintf(stderr, "usage: see header\n"); return 2; }
    char *text = read_file(path);
    if (!text) { fprintf(stderr, "instance error: cannot open %s\n", path); return 1; }
    orc_instance *insts = NULL;
    char err[256];
    int n = orc_parse(text, &insts, err, sizeof err);
    free(text);
    if (n != 1) {
        fprintf(stderr, "instance error: %s\n", n < 0 ? err : "expected exactly 1 instance");
        return 1;
    }
    j.inst = &insts[0];
    if (bench) {
        pthread_mutex_init(&j.mu_lock, NULL);
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
        for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &j);
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        double el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
        double packings = (double)j.K * j.passes;
        printf("{\"packings\": %.0f, \"seconds\": %.6f, \"packings_per_s\": %.3f, \"threads\": %d, "
               "\"checksum_volume\": %lld}\n", packings, el, packings / el, threads, j.checksum);
        free(th);
    } else {
        orc_result r;
        r.placements = (int *)malloc(sizeof(int) * 7 * (size_t)total_qty(j.inst));
        for (int k = 0; k < j.K; k++) {
            if (orc_run(j.inst, (uint64_t)(j.seed + k), j.passes, j.eps, j.use_theta, j.theta,
                        j.order, j.first_fit, j.mu, &r) != 0) {
                fprintf(stderr, "oracle run failed\n");
                return 1;
            }
            print_result(&r);
        }
        free(r.placements);
    }
    orc_free_instances(insts, n);
    return 0;
}
