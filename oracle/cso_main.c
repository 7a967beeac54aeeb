/*
 * cso_main.c -- command-line front end of the CPU oracle (test infrastructure).
 *
 * Accepts the covsearch argv contract of cov/engine/covsearch2e.c:1042-1095
 * (-q -n -R -M -s --iters --tabu --stall --kick --cand --random-init --in
 * --out --threads --quiet -t, --preset accepted and ignored because every
 * preset runs the same search) and prints the same RESULT line
 * (:1585-1589).  Time fields are wall/CPU seconds of the search loop.
 * Extra: --kick-order 0|1 (order of the two draws in a kick perturbation),
 * --trace FILE (binary int64 records, CSO_TR_W per step).
 */
#include "covsearch_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static const char *DIG = "0123456789abcdefghijklmnopqrstuvwxyz";

static double wall_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double cpu_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_PROCESS_CPUTIME_ID, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* code file: '#' comments, digits in base 36 or separated integers */
static int read_code(const char *path, int n, uint8_t **out)
{
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); exit(1); }
    char line[4096];
    int m = 0, cap = 1024;
    uint8_t *buf = malloc((size_t)cap * n);
    while (fgets(line, sizeof line, f)) {
        char *h = strchr(line, '#');
        if (h) *h = 0;
        uint8_t tmp[64];
        int len = 0;
        int spaced = strchr(line, ' ') || strchr(line, ',') || strchr(line, '\t');
        if (spaced) {
            for (char *tok = strtok(line, " ,\t\r\n"); tok && len < 64; tok = strtok(NULL, " ,\t\r\n"))
                tmp[len++] = (uint8_t)atoi(tok);
        } else {
            for (char *c = line; *c && *c != '\n' && *c != '\r'; c++) {
                const char *d = strchr(DIG, *c);
                if (!d) continue;
                if (len < 64) tmp[len++] = (uint8_t)(d - DIG);
            }
        }
        if (len == 0) continue;
        if (len != n) { fprintf(stderr, "%s: bad codeword length %d\n", path, len); exit(1); }
        if (m == cap) { cap *= 2; buf = realloc(buf, (size_t)cap * n); }
        memcpy(buf + (size_t)m * n, tmp, n);
        m++;
    }
    fclose(f);
    *out = buf;
    return m;
}

static void write_code(const char *path, const uint8_t *cw, int M, int n, int q)
{
    char tmp[4096];
    snprintf(tmp, sizeof tmp, "%s.part", path);
    FILE *f = fopen(tmp, "w");
    if (!f) { perror(tmp); return; }
    for (int i = 0; i < M; i++) {
        for (int j = 0; j < n; j++) {
            if (q <= 36) fputc(DIG[cw[(size_t)i * n + j]], f);
            else fprintf(f, "%d%s", cw[(size_t)i * n + j], j + 1 < n ? " " : "");
        }
        fputc('\n', f);
    }
    fclose(f);
    if (rename(tmp, path) != 0) perror(path);
}

int main(int argc, char **argv)
{
    int q = 0, n = 0, R = -1, M = 0, tabu = 0, kick = 2, cand = 100000, greedy = 1;
    int threads = 1, kick_order = 0;
    long long iters = -1, stall = 5000;
    unsigned long long seed = 12345;
    const char *in = NULL, *out = NULL, *trace_path = NULL;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
#define ARG(x) (!strcmp(a, x) && i + 1 < argc)
        if (ARG("-q")) q = atoi(argv[++i]);
        else if (ARG("-n")) n = atoi(argv[++i]);
        else if (ARG("-R")) R = atoi(argv[++i]);
        else if (ARG("-M")) M = atoi(argv[++i]);
        else if (ARG("-s")) seed = strtoull(argv[++i], NULL, 10);
        else if (ARG("-t")) ++i;
        else if (ARG("--iters")) iters = atoll(argv[++i]);
        else if (ARG("--tabu")) tabu = atoi(argv[++i]);
        else if (ARG("--stall")) stall = atoll(argv[++i]);
        else if (ARG("--kick")) kick = atoi(argv[++i]);
        else if (ARG("--cand")) cand = atoi(argv[++i]);
        else if (ARG("--threads")) threads = atoi(argv[++i]);
        else if (ARG("--in")) in = argv[++i];
        else if (ARG("--out")) out = argv[++i];
        else if (ARG("--preset")) ++i;
        else if (ARG("--kick-order")) kick_order = atoi(argv[++i]);
        else if (ARG("--trace")) trace_path = argv[++i];
        else if (!strcmp(a, "--random-init")) greedy = 0;
        else if (!strcmp(a, "--quiet")) ;
        else { fprintf(stderr, "unknown/incomplete option %s\n", a); return 2; }
#undef ARG
    }
    if (q < 2 || n < 1 || R < 0 || M < 1) { fprintf(stderr, "usage: cso -q Q -n N -R R -M M ...\n"); return 2; }
    if (M > 65535) { fprintf(stderr, "M > 65535 not supported (uint16 counters)\n"); return 2; }
    cso_t *s = cso_create(q, n, R, M, seed);
    if (!s) { fprintf(stderr, "bad parameters\n"); return 2; }
    cso_set_kick_order(s, kick_order);
    cso_set_threads(s, threads);
    uint8_t *sc = NULL;
    int ms = 0;
    if (in) ms = read_code(in, n, &sc);
    if (cso_init(s, sc, ms, greedy) != 0) { fprintf(stderr, "bad seed code\n"); return 2; }
    free(sc);

    int64_t *trace = NULL;
    if (trace_path && iters >= 0) trace = calloc((size_t)(iters ? iters : 1) * CSO_TR_W, sizeof(int64_t));
    double w0 = wall_now(), c0 = cpu_now();
    int64_t ran = cso_run(s, iters, tabu, stall, kick, cand, trace);
    double el = wall_now() - w0, cpu = cpu_now() - c0;
    long long it = (long long)cso_iters(s);
    printf("RESULT q=%d n=%d R=%d M=%d uncovered=%lld iters=%lld kicks=%lld "
           "time=%.2f rate=%.0f cpu=%.3f cpurate=%.0f target=-1 "
           "ttt_cpu=-1.0000 ttt_wall=-1.0000 ttt_it=-1\n",
           q, n, R, M, (long long)cso_best_uncovered(s), it, (long long)cso_kicks(s),
           el, it / (el > 0 ? el : 1), cpu, it / (cpu > 0 ? cpu : 1));
    if (trace) {
        FILE *f = fopen(trace_path, "wb");
        if (f) { fwrite(trace, sizeof(int64_t) * CSO_TR_W, (size_t)ran, f); fclose(f); }
        free(trace);
    }
    if (out) {
        uint8_t *buf = malloc((size_t)M * n);
        cso_distinct_fill(s, cso_best(s), M, buf);
        write_code(out, buf, M, n, q);
        free(buf);
    }
    int rc = cso_best_uncovered(s) == 0 ? 0 : 1;
    cso_destroy(s);
    return rc;
}
