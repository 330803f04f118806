/* A host of the C ABI with no Python: the runaway-electron swarm in dry air built from raw
 * processes (pt_sb_build, pt_table_build, pt_process_desc), uploaded, stepped and reported
 * as one line of JSON.  Usage: swarm <brem_SB dir> <n> <steps> <seed> [graph]
 * (graph: the steps run as replays of a captured graph, pt_run_graph, instead of pt_run) */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "particulator_hip.h"

#define EV 1.6021766208e-19
#define CHECK(ctx, x)                                                                 \
    do {                                                                              \
        int rc_ = (x);                                                                \
        if (rc_ != PT_OK) {                                                           \
            fprintf(stderr, "%s: %d %s\n", #x, rc_, pt_last_error(ctx));             \
            exit(1);                                                                  \
        }                                                                             \
    } while (0)

typedef struct {
    int Z, nx, ny;
    double *x, *y, *v, *log_energy, *tcs, *data;
} sb_t;

/* a Geant4 2D table: "k nx ny", nx values, ny values, ny rows of nx values */
static void read_sb(const char* dir, int Z, int ncum, sb_t* s) {
    char path[1024];
    snprintf(path, sizeof path, "%s/br%d", dir, Z);
    FILE* f = fopen(path, "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(1);
    }
    int k;
    if (fscanf(f, "%d %d %d", &k, &s->nx, &s->ny) != 3) exit(1);
    s->Z = Z;
    s->x = malloc(sizeof(double) * s->nx);
    s->y = malloc(sizeof(double) * s->ny);
    s->v = malloc(sizeof(double) * s->nx * s->ny);
    for (int i = 0; i < s->nx; i++)
        if (fscanf(f, "%lf", &s->x[i]) != 1) exit(1);
    for (int i = 0; i < s->ny; i++)
        if (fscanf(f, "%lf", &s->y[i]) != 1) exit(1);
    for (int i = 0; i < s->nx * s->ny; i++)
        if (fscanf(f, "%lf", &s->v[i]) != 1) exit(1);
    fclose(f);
    s->log_energy = malloc(sizeof(double) * s->ny);
    s->tcs = malloc(sizeof(double) * s->ny);
    s->data = malloc(sizeof(double) * s->ny * ncum);
    CHECK(NULL, pt_sb_build(Z, s->x, s->nx, s->y, s->ny, s->v, ncum, 1e2 * EV, 5e7 * EV, 1e6 * EV,
                            s->log_energy, s->tcs, s->data));
}

int main(int argc, char** argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: swarm <brem_SB dir> <n> <steps> <seed> [graph]\n");
        return 2;
    }
    const int graph = argc > 5 && strcmp(argv[5], "graph") == 0;
    const int ncum = 1000, n = atoi(argv[2]), steps = atoi(argv[3]);
    const uint64_t seed = strtoull(argv[4], NULL, 10);
    const double cut = 1e4 * EV, dt = 1e-12, efield = 1e6;
    /* dry air at 1 atm and 273 K: N2, O2, Ar number densities */
    const double nair = 101325.0 / (1.38064852e-23 * 273.0);
    const double nN2 = 0.781 * nair, nO2 = 0.210 * nair, nAr = 0.009 * nair;
    const int Zs[3] = {7, 8, 18};
    const double atoms[3] = {2 * nN2, 2 * nO2, nAr};
    sb_t sb[3];
    for (int i = 0; i < 3; i++) read_sb(argv[1], Zs[i], ncum, &sb[i]);
    /* electron processes: elastic and bremsstrahlung per atom, RBEB per molecular orbital
     * (src/rbeb.jl:25-48), Moller for argon */
    static const double orbN2[5][3] = {{409.5, 603.3, 4}, {41.72, 71.13, 2}, {21.00, 63.18, 2},
                                       {17.07, 44.30, 4}, {15.58, 54.91, 2}};
    static const double orbO2[6][3] = {{543.8, 796.2, 4}, {46.19, 79.73, 2}, {29.82, 90.92, 2},
                                       {19.64, 59.89, 4}, {19.79, 71.84, 2}, {12.07, 84.88, 2}};
    pt_process procs[24];
    int sbslot[24];
    int L = 0;
    memset(procs, 0, sizeof procs);
    for (int i = 0; i < 3; i++, L++) {
        procs[L].kind = PT_PROC_ELASTIC;
        procs[L].Z = Zs[i];
        procs[L].coef = atoms[i];
        sbslot[L] = 0;
    }
    for (int i = 0; i < 3; i++, L++) {
        procs[L].kind = PT_PROC_SB;
        procs[L].Z = Zs[i];
        procs[L].coef = atoms[i];
        procs[L].ntab = sb[i].ny;
        procs[L].tab_x = sb[i].log_energy;
        procs[L].tab_y = sb[i].tcs;
        sbslot[L] = i;
    }
    for (int i = 0; i < 5; i++, L++) {
        procs[L].kind = PT_PROC_RBEB;
        procs[L].coef = nN2;
        procs[L].B = orbN2[i][0] * EV;
        procs[L].U = orbN2[i][1] * EV;
        procs[L].N = (int)orbN2[i][2];
        sbslot[L] = 0;
    }
    for (int i = 0; i < 6; i++, L++) {
        procs[L].kind = PT_PROC_RBEB;
        procs[L].coef = nO2;
        procs[L].B = orbO2[i][0] * EV;
        procs[L].U = orbO2[i][1] * EV;
        procs[L].N = (int)orbO2[i][2];
        sbslot[L] = 0;
    }
    procs[L].kind = PT_PROC_MOLLER;
    procs[L].Z = 18;
    procs[L].coef = nAr;
    procs[L].tcut = cut;
    sbslot[L] = 0;
    L++;
    const int K = 32, order = 3;
    const double xmax = 1e3 * EV * 262144.0;
    const double fdt = EV * efield * dt;
    double* rate = malloc(sizeof(double) * (K + 1) * L * order);
    double* rb = malloc(sizeof(double) * (K + 1) * order);
    int32_t perm[24];
    CHECK(NULL, pt_table_build(procs, L, PT_ELECTRON, &fdt, 1, 1.1, K, order, xmax, rate, rb, perm, NULL));
    pt_proc_desc desc[24];
    for (int j = 0; j < L; j++) CHECK(NULL, pt_process_desc(&procs[perm[j]], sbslot[perm[j]], &desc[j]));

    pt_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.capacity[PT_ELECTRON] = (uint64_t)(4 * n + 4096);
    cfg.capacity[PT_PHOTON] = 0;
    cfg.capacity[PT_POSITRON] = 0;
    cfg.seed = seed;
    cfg.tracked[PT_ELECTRON] = 1;
    cfg.nswarm = 1;
    cfg.tally_bins = 64;
    cfg.tally_lo = 1e4 * EV;
    cfg.tally_hi = 1e7 * EV;
    pt_ctx* ctx = NULL;
    CHECK(NULL, pt_create(&cfg, &ctx));
    for (int i = 0; i < 3; i++)
        CHECK(ctx, pt_upload_sb(ctx, i, Zs[i], sb[i].ny, ncum, sb[i].log_energy, sb[i].data));
    pt_table_desc td;
    memset(&td, 0, sizeof td);
    td.nprocs = L;
    td.order = order;
    td.nintervals = K;
    td.nswarm = 1;
    td.xmax = xmax;
    td.energy_cut = cut;
    td.rate = rate;
    td.ratebound = rb;
    td.procs = desc;
    CHECK(ctx, pt_upload_table(ctx, PT_ELECTRON, &td));
    pt_field_desc fd;
    memset(&fd, 0, sizeof fd);
    fd.e.kind = PT_FIELD_HOMOGENEOUS;
    fd.e.v[2] = -efield;
    fd.b.kind = PT_FIELD_HOMOGENEOUS;
    CHECK(ctx, pt_set_field(ctx, 0, &fd));
    CHECK(ctx, pt_init_particles(ctx, PT_ELECTRON, (size_t)n, 1e6 * EV, 1e6 * EV, 0, 0, 0, 0.0));
    CHECK(ctx, pt_init(ctx));
    if (graph)
        CHECK(ctx, pt_run_graph(ctx, 0.0, dt, steps, steps));
    else
        CHECK(ctx, pt_run(ctx, 0.0, dt, steps, 8, steps));
    pt_pop_stats st;
    CHECK(ctx, pt_stats(ctx, PT_ELECTRON, &st));
    printf("{\"nactive\": %llu, \"weight\": %.17g, \"mean_energy\": %.17g, \"max_energy\": %.17g}\n",
           (unsigned long long)st.nactive, st.weight, st.mean_energy, st.max_energy);
    pt_destroy(ctx);
    return 0;
}
