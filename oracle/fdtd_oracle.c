/*
 * CPU oracle for the js_fdtd Yee time-step loop -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker that the HIP backend is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; no product code
 * path may call it.  It restates, in plain C with IEEE double arithmetic and no FMA
 * contraction (build with -ffp-contract=off), the reference's per-step algorithm:
 *
 *   updateH                js_fdtd/assembly/fdtd_kernels.ts:20-114
 *   updateMagneticCPML     js_fdtd/src/cpml.js:456-724
 *   captureSampledCurrents js_fdtd/src/sampling.js:204-245 (as an ordered term list)
 *   updateE                js_fdtd/assembly/fdtd_kernels.ts:116-210
 *   updateElectricCPML     js_fdtd/src/cpml.js:727-1003
 *   injectVoltageSources   js_fdtd/src/sources.js:271-288
 *   captureSampledVoltages js_fdtd/src/sampling.js:295-307
 *   captureMagneticFields  js_fdtd/src/sampling.js:144-198 (point H probes)
 *   captureElectricFields  js_fdtd/src/sampling.js:251-292 (point E probes)
 *   accumulateFarfieldDFT  js_fdtd/src/farfield.js:90-248
 *   step order             js_fdtd/src/fdtdSolver.js:84-140, fdtdWorker.js:57-125
 *
 * Memory model.  Like the reference's WASM engine (js_fdtd/src/grid.js:125-162) the
 * 6 fields and 18 coefficient arrays live in ONE contiguous buffer in the order
 * Hx,Hy,Hz,Ex,Ey,Ez,Cexe,...,Chzex.  updateE's i = nx plane therefore reads past the
 * end of Hz into Ex and past the end of Hy into Hz exactly as the WASM code does.
 * Threaded mode splits the bulk loops into x-slabs like fdtdWorker.js:22-30 but runs
 * all of Ex before Ey/Ez, which is the single-thread outcome of the Ex(0) race.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAX_FACES 6
#define MAX_OBS 16

enum { HX, HY, HZ, EX, EY, EZ };
/* coefficient ids in the buffer order of grid.js:144-162 */
enum { CEXE = 6, CEXHZ, CEXHY, CEYE, CEYHX, CEYHZ, CEZE, CEZHY, CEZHX,
       CHXH, CHXEZ, CHXEY, CHYH, CHYEX, CHYEZ, CHZH, CHZEY, CHZEX, NARR };

typedef struct {
    int present, nc, m_start, e_start;
    double *b_e, *a_e, *b_m, *a_m;
    double *psi[4], *cpsi[4];
    long n[4];
} Face;

typedef struct { int dir, count; int *idx; double *coef, *vpe; int nsteps; } VSource;
typedef struct { int dir, count; int *idx; double csvf; } SVoltage;
typedef struct { int count, negate; int *comp, *index; double *weight; } SCurrent;
typedef struct { int magnetic, comp, i, j, k; } Probe;   /* comp 0..2 = x,y,z; 3 = 'm' */

typedef struct {
    int nx, ny, nz;
    long len[NARR], off[NARR];
    double *buf;
    Face face[MAX_FACES];
    VSource vs[MAX_OBS]; int nvs;
    SVoltage sv[MAX_OBS]; int nsv;
    SCurrent si[MAX_OBS]; int nsi;
    Probe pr[MAX_OBS]; int npr;
    /* NTFF */
    int nfreq, li, lj, lk, ui, uj, uk, ni, nj, nk, steps;
    double *phase;                 /* [steps][nfreq][4] */
    double *ff[24][2];             /* [re/im], nfreq * face size */
    /* threads */
    int nthreads;
} Oracle;

static double *A(Oracle *o, int w) { return o->buf + o->off[w]; }

static void comp_dims(const Oracle *o, int c, int d[3]) {
    static const int ext[6][3] = {{1,0,0},{0,1,0},{0,0,1},{0,1,1},{1,0,1},{1,1,0}};
    d[0] = o->nx + ext[c][0]; d[1] = o->ny + ext[c][1]; d[2] = o->nz + ext[c][2];
}

Oracle *oracle_create(int nx, int ny, int nz) {
    Oracle *o = calloc(1, sizeof(Oracle));
    if (!o) return NULL;
    o->nx = nx; o->ny = ny; o->nz = nz;
    static const int owner[NARR] = {HX,HY,HZ,EX,EY,EZ, EX,EX,EX, EY,EY,EY, EZ,EZ,EZ,
                                    HX,HX,HX, HY,HY,HY, HZ,HZ,HZ};
    long total = 0;
    for (int w = 0; w < NARR; w++) {
        int d[3]; comp_dims(o, owner[w], d);
        o->len[w] = (long)d[0] * d[1] * d[2];
        o->off[w] = total;
        total += o->len[w];
    }
    o->buf = calloc((size_t)total + 8192, sizeof(double));   /* slack like the WASM pages */
    if (!o->buf) { free(o); return NULL; }
    o->nthreads = 1;
    return o;
}

double *oracle_array(Oracle *o, int which) { return (which >= 0 && which < NARR) ? A(o, which) : NULL; }
long oracle_array_len(Oracle *o, int which) { return (which >= 0 && which < NARR) ? o->len[which] : -1; }

static double *dup(const double *src, long n) {
    double *p = malloc(sizeof(double) * (n > 0 ? n : 1));
    if (p && src) memcpy(p, src, sizeof(double) * n);
    return p;
}

/* Face slots follow cpml.js: x {Hy,Hz,Ey,Ez}, y {Hx,Hz,Ez,Ex}, z {Hx,Hy,Ex,Ey}. */
static const int SLOT_COMP[3][4] = {{HY, HZ, EY, EZ}, {HX, HZ, EZ, EX}, {HX, HY, EX, EY}};
static const int SLOT_SRC[3][4]  = {{EZ, EY, HZ, HY}, {EZ, EX, HX, HZ}, {EY, EX, HY, HX}};

int oracle_add_cpml(Oracle *o, int face, int nc, const double *b_e, const double *a_e,
                    const double *b_m, const double *a_m, const double *cpsi0, const double *cpsi1,
                    const double *cpsi2, const double *cpsi3, int m_start, int e_start) {
    if (face < 0 || face >= MAX_FACES) return 1;
    Face *f = &o->face[face];
    int axis = face / 2;
    f->present = 1; f->nc = nc; f->m_start = m_start; f->e_start = e_start;
    f->b_e = dup(b_e, nc); f->a_e = dup(a_e, nc); f->b_m = dup(b_m, nc); f->a_m = dup(a_m, nc);
    const double *cp[4] = {cpsi0, cpsi1, cpsi2, cpsi3};
    for (int s = 0; s < 4; s++) {
        int d[3]; comp_dims(o, SLOT_COMP[axis][s], d);
        d[axis] = nc;
        f->n[s] = (long)d[0] * d[1] * d[2];
        f->cpsi[s] = dup(cp[s], f->n[s]);
        f->psi[s] = calloc(f->n[s] > 0 ? f->n[s] : 1, sizeof(double));
    }
    return 0;
}

double *oracle_cpml_psi(Oracle *o, int face, int slot) {
    return (face >= 0 && face < MAX_FACES && o->face[face].present && slot >= 0 && slot < 4)
        ? o->face[face].psi[slot] : NULL;
}

int oracle_add_vsource(Oracle *o, int dir, const int *idx, const double *coef, int count,
                       const double *vpe, int nsteps) {
    if (o->nvs >= MAX_OBS) return 1;
    VSource *v = &o->vs[o->nvs++];
    v->dir = dir; v->count = count; v->nsteps = nsteps;
    v->idx = malloc(sizeof(int) * (count > 0 ? count : 1)); memcpy(v->idx, idx, sizeof(int) * count);
    v->coef = dup(coef, count); v->vpe = dup(vpe, nsteps);
    return 0;
}

int oracle_add_svoltage(Oracle *o, int dir, const int *idx, int count, double csvf) {
    if (o->nsv >= MAX_OBS) return 1;
    SVoltage *s = &o->sv[o->nsv++];
    s->dir = dir; s->count = count; s->csvf = csvf;
    s->idx = malloc(sizeof(int) * (count > 0 ? count : 1)); memcpy(s->idx, idx, sizeof(int) * count);
    return 0;
}

int oracle_add_scurrent(Oracle *o, const int *comp, const int *index, const double *weight,
                        int count, int negate) {
    if (o->nsi >= MAX_OBS) return 1;
    SCurrent *s = &o->si[o->nsi++];
    s->count = count; s->negate = negate;
    s->comp = malloc(sizeof(int) * (count > 0 ? count : 1)); memcpy(s->comp, comp, sizeof(int) * count);
    s->index = malloc(sizeof(int) * (count > 0 ? count : 1)); memcpy(s->index, index, sizeof(int) * count);
    s->weight = dup(weight, count);
    return 0;
}

int oracle_set_farfield(Oracle *o, int nfreq, const int *box, const double *phase, int steps) {
    o->nfreq = nfreq;
    o->li = box[0]; o->lj = box[1]; o->lk = box[2]; o->ui = box[3]; o->uj = box[4]; o->uk = box[5];
    o->ni = o->ui - o->li; o->nj = o->uj - o->lj; o->nk = o->uk - o->lk;
    o->steps = steps;
    o->phase = dup(phase, (long)steps * nfreq * 4);
    long sz[3] = {(long)o->nj * o->nk, (long)o->ni * o->nk, (long)o->ni * o->nj};
    for (int a = 0; a < 24; a++) {
        long n = sz[a / 8] * nfreq;
        o->ff[a][0] = calloc(n > 0 ? n : 1, sizeof(double));
        o->ff[a][1] = calloc(n > 0 ? n : 1, sizeof(double));
    }
    return 0;
}

double *oracle_ff(Oracle *o, int which, int imag) {
    return (which >= 0 && which < 24 && o->nfreq > 0) ? o->ff[which][imag ? 1 : 0] : NULL;
}

void oracle_destroy(Oracle *o) {
    if (!o) return;
    for (int f = 0; f < MAX_FACES; f++) {
        Face *F = &o->face[f];
        if (!F->present) continue;
        free(F->b_e); free(F->a_e); free(F->b_m); free(F->a_m);
        for (int s = 0; s < 4; s++) { free(F->psi[s]); free(F->cpsi[s]); }
    }
    for (int i = 0; i < o->nvs; i++) { free(o->vs[i].idx); free(o->vs[i].coef); free(o->vs[i].vpe); }
    for (int i = 0; i < o->nsv; i++) free(o->sv[i].idx);
    for (int i = 0; i < o->nsi; i++) { free(o->si[i].comp); free(o->si[i].index); free(o->si[i].weight); }
    if (o->nfreq > 0) {
        free(o->phase);
        for (int a = 0; a < 24; a++) { free(o->ff[a][0]); free(o->ff[a][1]); }
    }
    free(o->buf);
    free(o);
}

/* ---------------------------------------------------------------- bulk updates -- */
/* Each function covers the x-slab [i0,i1) of its component, like the WASM kernels'
 * partition arguments; i-ranges are clipped to the component's own bounds. */

static void bulk_h_slab(Oracle *o, int i0x, int i1x, int i0, int i1) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nyp1 = ny + 1, nzp1 = nz + 1;
    double *Hx = A(o, HX), *Hy = A(o, HY), *Hz = A(o, HZ);
    const double *Ex = A(o, EX), *Ey = A(o, EY), *Ez = A(o, EZ);
    const double *Chxh = A(o, CHXH), *Chxey = A(o, CHXEY), *Chxez = A(o, CHXEZ);
    const double *Chyh = A(o, CHYH), *Chyez = A(o, CHYEZ), *Chyex = A(o, CHYEX);
    const double *Chzh = A(o, CHZH), *Chzex = A(o, CHZEX), *Chzey = A(o, CHZEY);
    (void)nx;
    for (int i = i0x; i < i1x; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long p = ((long)i * ny + j) * nz + k;
                long ey = ((long)i * ny + j) * nzp1 + k;
                long ez0 = ((long)i * nyp1 + j) * nz + k, ez1 = ez0 + nz;
                Hx[p] = Chxh[p] * Hx[p] + Chxey[p] * (Ey[ey + 1] - Ey[ey]) + Chxez[p] * (Ez[ez1] - Ez[ez0]);
            }
    for (int i = i0; i < i1; i++)
        for (int j = 0; j < nyp1; j++)
            for (int k = 0; k < nz; k++) {
                long p = ((long)i * nyp1 + j) * nz + k;
                long ez0 = p, ez1 = ((long)(i + 1) * nyp1 + j) * nz + k;
                long ex = ((long)i * nyp1 + j) * nzp1 + k;
                Hy[p] = Chyh[p] * Hy[p] + Chyez[p] * (Ez[ez1] - Ez[ez0]) + Chyex[p] * (Ex[ex + 1] - Ex[ex]);
            }
    for (int i = i0; i < i1; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nzp1; k++) {
                long p = ((long)i * ny + j) * nzp1 + k;
                long ex0 = ((long)i * nyp1 + j) * nzp1 + k, ex1 = ex0 + nzp1;
                long ey0 = p, ey1 = ((long)(i + 1) * ny + j) * nzp1 + k;
                Hz[p] = Chzh[p] * Hz[p] + Chzex[p] * (Ex[ex1] - Ex[ex0]) + Chzey[p] * (Ey[ey1] - Ey[ey0]);
            }
}

static void bulk_ex_slab(Oracle *o, int i0, int i1) {
    const int ny = o->ny, nz = o->nz, nyp1 = ny + 1, nzp1 = nz + 1;
    double *Ex = A(o, EX);
    const double *Hy = A(o, HY), *Hz = A(o, HZ);
    const double *Cexe = A(o, CEXE), *Cexhz = A(o, CEXHZ), *Cexhy = A(o, CEXHY);
    for (int i = i0; i < i1; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                long p = ((long)i * nyp1 + j) * nzp1 + k;
                long hz1 = ((long)i * ny + j) * nzp1 + k, hz0 = hz1 - nzp1;
                long hy = ((long)i * nyp1 + j) * nz + k;
                Ex[p] = Cexe[p] * Ex[p] + Cexhz[p] * (Hz[hz1] - Hz[hz0]) + Cexhy[p] * (Hy[hy] - Hy[hy - 1]);
            }
}

/* Ey/Ez over i in [i0,i1) of [1, iend); iend = nx+1 reproduces the reference's ghost
 * plane (reads Hz[nx,..] / Hy[nx,..] out of their arrays), iend = nx stops there. */
static void bulk_eyz_slab(Oracle *o, int i0, int i1) {
    const int ny = o->ny, nz = o->nz, nyp1 = ny + 1, nzp1 = nz + 1;
    double *Ey = A(o, EY), *Ez = A(o, EZ);
    const double *Hx = A(o, HX), *Hy = A(o, HY), *Hz = A(o, HZ);
    const double *Ceye = A(o, CEYE), *Ceyhx = A(o, CEYHX), *Ceyhz = A(o, CEYHZ);
    const double *Ceze = A(o, CEZE), *Cezhy = A(o, CEZHY), *Cezhx = A(o, CEZHX);
    if (i0 < 1) i0 = 1;
    for (int i = i0; i < i1; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 1; k < nz; k++) {
                long p = ((long)i * ny + j) * nzp1 + k;
                long hx = ((long)i * ny + j) * nz + k;
                long hz1 = p, hz0 = p - (long)ny * nzp1;
                Ey[p] = Ceye[p] * Ey[p] + Ceyhx[p] * (Hx[hx] - Hx[hx - 1]) + Ceyhz[p] * (Hz[hz1] - Hz[hz0]);
            }
    for (int i = i0; i < i1; i++)
        for (int j = 1; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                long p = ((long)i * nyp1 + j) * nz + k;
                long hy1 = p, hy0 = p - (long)nyp1 * nz;
                long hx1 = ((long)i * ny + j) * nz + k, hx0 = hx1 - nz;
                Ez[p] = Ceze[p] * Ez[p] + Cezhy[p] * (Hy[hy1] - Hy[hy0]) + Cezhx[p] * (Hx[hx1] - Hx[hx0]);
            }
}

/* ------------------------------------------------------------- thread x-slabs -- */
typedef struct { Oracle *o; int id, phase, iend; pthread_barrier_t *bar; } Job;

static void partition(int size, int id, int total, int *s, int *e) {
    int chunk = (size + total - 1) / total;           /* fdtdWorker.js:22-27 */
    *s = id * chunk < size ? id * chunk : size;
    *e = *s + chunk < size ? *s + chunk : size;
}

static void *slab_worker(void *arg) {
    Job *jb = arg;
    Oracle *o = jb->o;
    int s0, e0, s1, e1;
    partition(o->nx, jb->id, o->nthreads, &s0, &e0);
    partition(jb->iend, jb->id, o->nthreads, &s1, &e1);
    if (jb->phase == 0) {
        int sx, ex; partition(o->nx + 1, jb->id, o->nthreads, &sx, &ex);
        bulk_h_slab(o, sx, ex, s0, e0);
    } else {
        bulk_ex_slab(o, s0, e0);
        pthread_barrier_wait(jb->bar);
        bulk_eyz_slab(o, s1, e1);
    }
    return NULL;
}

static void run_slabs(Oracle *o, int phase, int iend) {
    int n = o->nthreads;
    if (n <= 1) {
        if (phase == 0) bulk_h_slab(o, 0, o->nx + 1, 0, o->nx);
        else { bulk_ex_slab(o, 0, o->nx); bulk_eyz_slab(o, 1, iend); }
        return;
    }
    pthread_t th[256]; Job jobs[256]; pthread_barrier_t bar;
    if (n > 256) n = 256;
    pthread_barrier_init(&bar, NULL, (unsigned)n);
    for (int t = 0; t < n; t++) {
        jobs[t] = (Job){o, t, phase, iend, &bar};
        pthread_create(&th[t], NULL, slab_worker, &jobs[t]);
    }
    for (int t = 0; t < n; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&bar);
}

void oracle_set_threads(Oracle *o, int n) { o->nthreads = n < 1 ? 1 : (n > 256 ? 256 : n); }
void oracle_update_h(Oracle *o) { run_slabs(o, 0, 0); }
/* boundary_mode 0: reference (updates the i = nx plane), 1: stop at i = nx-1 */
void oracle_update_e(Oracle *o, int boundary_mode) {
    run_slabs(o, 1, boundary_mode == 0 ? o->nx + 1 : o->nx);
}

/* --------------------------------------------------------------------- CPML -- */
/* One face, H (kind 0) or E (kind 1): all Psi of the face, then the corrections, in
 * the face order the caller imposes (xn,xp,yn,yp,zn,zp).  Generic over the axis: the
 * Psi array has the component's shape with extent nc along the face axis; the source
 * field has one more (H kind) or one fewer (E kind) sample along that axis and the
 * same extents across it. */
static void cpml_face(Oracle *o, int face, int kind) {
    Face *f = &o->face[face];
    if (!f->present) return;
    const int axis = face / 2, nc = f->nc;
    for (int s = kind * 2; s < kind * 2 + 2; s++) {
        int c = SLOT_COMP[axis][s], src = SLOT_SRC[axis][s];
        int dc[3], ds[3];
        comp_dims(o, c, dc); comp_dims(o, src, ds);
        int dp[3] = {dc[0], dc[1], dc[2]}; dp[axis] = nc;
        int lim[3] = {dc[0], dc[1], dc[2]}; lim[axis] = 1;   /* axis coordinate set from ci */
        double *F = A(o, c); const double *S = A(o, src);
        double *psi = f->psi[s]; const double *cpsi = f->cpsi[s];
        const double *b = kind == 0 ? f->b_m : f->b_e, *a = kind == 0 ? f->a_m : f->a_e;
        const int start = kind == 0 ? f->m_start : f->e_start;
        const int o_hi = kind == 0 ? 1 : 0, o_lo = kind == 0 ? 0 : -1;
        /* Psi recurrence */
        for (int ci = 0; ci < nc; ci++) {
            int plane = start + ci;
            for (int u = 0; u < lim[0]; u++) for (int v = 0; v < lim[1]; v++) for (int w = 0; w < lim[2]; w++) {
                int q[3] = {u, v, w}; q[axis] = plane + o_hi;
                int r[3] = {u, v, w}; r[axis] = plane + o_lo;
                int t[3] = {u, v, w}; t[axis] = ci;
                long pi = ((long)t[0] * dp[1] + t[1]) * dp[2] + t[2];
                long s1 = ((long)q[0] * ds[1] + q[1]) * ds[2] + q[2];
                long s0 = ((long)r[0] * ds[1] + r[1]) * ds[2] + r[2];
                psi[pi] = b[ci] * psi[pi] + a[ci] * (S[s1] - S[s0]);
            }
        }
        /* field correction */
        for (int ci = 0; ci < nc; ci++) {
            int plane = start + ci;
            for (int u = 0; u < lim[0]; u++) for (int v = 0; v < lim[1]; v++) for (int w = 0; w < lim[2]; w++) {
                int t[3] = {u, v, w}; t[axis] = ci;
                int g[3] = {u, v, w}; g[axis] = plane;
                long pi = ((long)t[0] * dp[1] + t[1]) * dp[2] + t[2];
                long fp = ((long)g[0] * dc[1] + g[1]) * dc[2] + g[2];
                F[fp] += cpsi[pi] * psi[pi];
            }
        }
    }
}

void oracle_update_h_cpml(Oracle *o) { for (int f = 0; f < MAX_FACES; f++) cpml_face(o, f, 0); }
void oracle_update_e_cpml(Oracle *o) { for (int f = 0; f < MAX_FACES; f++) cpml_face(o, f, 1); }

/* ------------------------------------------------------ sources and samplers -- */
void oracle_inject(Oracle *o, int ts) {
    for (int n = 0; n < o->nvs; n++) {
        VSource *v = &o->vs[n];
        double *E = A(o, EX + v->dir);
        double val = v->vpe[ts];
        for (int m = 0; m < v->count; m++) E[v->idx[m]] += v->coef[m] * val;
    }
}

void oracle_capture_voltages(Oracle *o, double *out) {
    for (int n = 0; n < o->nsv; n++) {
        SVoltage *s = &o->sv[n];
        const double *E = A(o, EX + s->dir);
        double sum = 0;
        for (int m = 0; m < s->count; m++) sum += E[s->idx[m]];
        out[n] = s->csvf * sum;
    }
}

void oracle_capture_currents(Oracle *o, double *out) {
    for (int n = 0; n < o->nsi; n++) {
        SCurrent *s = &o->si[n];
        double sv = 0;
        for (int m = 0; m < s->count; m++) sv += s->weight[m] * A(o, HX + s->comp[m])[s->index[m]];
        out[n] = s->negate ? -sv : sv;
    }
}

/* Point probe at the 0-based cell (i,j,k) = (is-1, js-1, ks-1).  Returns -1 when a
 * probe slot is missing. */
int oracle_add_probe(Oracle *o, int magnetic, int comp, int i, int j, int k) {
    if (o->npr >= MAX_OBS) return -1;
    Probe *p = &o->pr[o->npr++];
    p->magnetic = magnetic; p->comp = comp; p->i = i; p->j = j; p->k = k;
    return 0;
}

/* Element (i,j,k) of component c in the reference's flat layout. */
static double at(Oracle *o, int c, int i, int j, int k) {
    int d[3]; comp_dims(o, c, d);
    return A(o, c)[((long)i * d[1] + j) * d[2] + k];
}

/* captureMagneticFields (sampling.js:144-198): 4-point averages. */
static double h_avg(Oracle *o, int c, int is, int js, int ks) {
    if (c == 0) return 0.25 * (at(o, HX, is, js-1, ks-1) + at(o, HX, is, js, ks-1) +
                               at(o, HX, is, js-1, ks) + at(o, HX, is, js, ks));
    if (c == 1) return 0.25 * (at(o, HY, is-1, js, ks-1) + at(o, HY, is, js, ks-1) +
                               at(o, HY, is-1, js, ks) + at(o, HY, is, js, ks));
    return 0.25 * (at(o, HZ, is-1, js-1, ks) + at(o, HZ, is, js-1, ks) +
                   at(o, HZ, is-1, js, ks) + at(o, HZ, is, js, ks));
}

/* captureElectricFields (sampling.js:251-292): 2-point averages. */
static double e_avg(Oracle *o, int c, int is, int js, int ks) {
    if (c == 0) return 0.5 * (at(o, EX, is-1, js, ks) + at(o, EX, is, js, ks));
    if (c == 1) return 0.5 * (at(o, EY, is, js-1, ks) + at(o, EY, is, js, ks));
    return 0.5 * (at(o, EZ, is, js, ks-1) + at(o, EZ, is, js, ks));
}

/* The probes of one kind (magnetic or electric), in upload order, into out[p] for
 * every probe p of that kind. */
void oracle_capture_probes(Oracle *o, int magnetic, double *out) {
    for (int n = 0; n < o->npr; n++) {
        Probe *p = &o->pr[n];
        if (p->magnetic != magnetic) continue;
        double (*avg)(Oracle *, int, int, int, int) = magnetic ? h_avg : e_avg;
        if (p->comp < 3) {
            out[n] = avg(o, p->comp, p->i, p->j, p->k);
        } else {
            double x = avg(o, 0, p->i, p->j, p->k), y = avg(o, 1, p->i, p->j, p->k), z = avg(o, 2, p->i, p->j, p->k);
            out[n] = sqrt(x*x + y*y + z*z);
        }
    }
}

/* ---------------------------------------------------------------------- NTFF -- */
void oracle_ff_accumulate(Oracle *o, int ts) {
    if (o->nfreq <= 0) return;
    const int ny = o->ny, nz = o->nz, nyp1 = ny + 1, nzp1 = nz + 1;
    const double *Hx = A(o, HX), *Hy = A(o, HY), *Hz = A(o, HZ);
    const double *Ex = A(o, EX), *Ey = A(o, EY), *Ez = A(o, EZ);
#define IHX(i, j, k) ((((long)(i)) * ny + (j)) * nz + (k))
#define IHY(i, j, k) ((((long)(i)) * nyp1 + (j)) * nz + (k))
#define IHZ(i, j, k) ((((long)(i)) * ny + (j)) * nzp1 + (k))
#define IEX(i, j, k) ((((long)(i)) * nyp1 + (j)) * nzp1 + (k))
#define IEY(i, j, k) ((((long)(i)) * ny + (j)) * nzp1 + (k))
#define IEZ(i, j, k) ((((long)(i)) * nyp1 + (j)) * nz + (k))
#define ACC(a, p, c, s, t) do { o->ff[a][0][p] += (c) * (t); o->ff[a][1][p] += (s) * (t); } while (0)
    const int LI = o->li - 1, LJ = o->lj - 1, LK = o->lk - 1, UI = o->ui - 1, UJ = o->uj - 1, UK = o->uk - 1;
    const long sx = (long)o->nj * o->nk, sy = (long)o->ni * o->nk, sz = (long)o->ni * o->nj;
    for (int mi = 0; mi < o->nfreq; mi++) {
        const double *ph = o->phase + ((long)ts * o->nfreq + mi) * 4;
        const double ch = ph[0], sh = ph[1], ce = ph[2], se = ph[3];
        for (int a = 0; a < o->nj; a++) for (int b = 0; b < o->nk; b++) {   /* xp, xn */
            long p = mi * sx + (long)a * o->nk + b; int j0 = LJ + a, k0 = LK + b;
            double t;
            t = 0.5 * (Ez[IEZ(UI, j0, k0)] + Ez[IEZ(UI, j0 + 1, k0)]);                         ACC(2, p, ce, se, t);
            t = -0.5 * (Ey[IEY(UI, j0, k0)] + Ey[IEY(UI, j0, k0 + 1)]);                        ACC(3, p, ce, se, t);
            t = -0.25 * (Hz[IHZ(UI, j0, k0)] + Hz[IHZ(UI, j0, k0 + 1)] + Hz[IHZ(UI - 1, j0, k0)] + Hz[IHZ(UI - 1, j0, k0 + 1)]);
            ACC(0, p, ch, sh, t);
            t = 0.25 * (Hy[IHY(UI, j0, k0)] + Hy[IHY(UI, j0 + 1, k0)] + Hy[IHY(UI - 1, j0, k0)] + Hy[IHY(UI - 1, j0 + 1, k0)]);
            ACC(1, p, ch, sh, t);
        }
        for (int a = 0; a < o->nj; a++) for (int b = 0; b < o->nk; b++) {
            long p = mi * sx + (long)a * o->nk + b; int j0 = LJ + a, k0 = LK + b;
            double t;
            t = -0.5 * (Ez[IEZ(LI, j0, k0)] + Ez[IEZ(LI, j0 + 1, k0)]);                        ACC(6, p, ce, se, t);
            t = 0.5 * (Ey[IEY(LI, j0, k0)] + Ey[IEY(LI, j0, k0 + 1)]);                         ACC(7, p, ce, se, t);
            t = 0.25 * (Hz[IHZ(LI, j0, k0)] + Hz[IHZ(LI, j0, k0 + 1)] + Hz[IHZ(LI - 1, j0, k0)] + Hz[IHZ(LI - 1, j0, k0 + 1)]);
            ACC(4, p, ch, sh, t);
            t = -0.25 * (Hy[IHY(LI, j0, k0)] + Hy[IHY(LI, j0 + 1, k0)] + Hy[IHY(LI - 1, j0, k0)] + Hy[IHY(LI - 1, j0 + 1, k0)]);
            ACC(5, p, ch, sh, t);
        }
        for (int a = 0; a < o->ni; a++) for (int b = 0; b < o->nk; b++) {   /* yp, yn */
            long p = mi * sy + (long)a * o->nk + b; int i0 = LI + a, k0 = LK + b;
            double t;
            t = -0.5 * (Ez[IEZ(i0, UJ, k0)] + Ez[IEZ(i0 + 1, UJ, k0)]);                        ACC(10, p, ce, se, t);
            t = 0.5 * (Ex[IEX(i0, UJ, k0)] + Ex[IEX(i0, UJ, k0 + 1)]);                         ACC(11, p, ce, se, t);
            t = -0.25 * (Hx[IHX(i0, UJ, k0)] + Hx[IHX(i0 + 1, UJ, k0)] + Hx[IHX(i0, UJ - 1, k0)] + Hx[IHX(i0 + 1, UJ - 1, k0)]);
            ACC(9, p, ch, sh, t);
            t = 0.25 * (Hz[IHZ(i0, UJ, k0)] + Hz[IHZ(i0, UJ, k0 + 1)] + Hz[IHZ(i0, UJ - 1, k0)] + Hz[IHZ(i0, UJ - 1, k0 + 1)]);
            ACC(8, p, ch, sh, t);
        }
        for (int a = 0; a < o->ni; a++) for (int b = 0; b < o->nk; b++) {
            long p = mi * sy + (long)a * o->nk + b; int i0 = LI + a, k0 = LK + b;
            double t;
            t = 0.5 * (Ez[IEZ(i0, LJ, k0)] + Ez[IEZ(i0 + 1, LJ, k0)]);                         ACC(14, p, ce, se, t);
            t = -0.5 * (Ex[IEX(i0, LJ, k0)] + Ex[IEX(i0, LJ, k0 + 1)]);                        ACC(15, p, ce, se, t);
            t = 0.25 * (Hx[IHX(i0, LJ, k0)] + Hx[IHX(i0 + 1, LJ, k0)] + Hx[IHX(i0, LJ - 1, k0)] + Hx[IHX(i0 + 1, LJ - 1, k0)]);
            ACC(13, p, ch, sh, t);
            t = -0.25 * (Hz[IHZ(i0, LJ, k0)] + Hz[IHZ(i0, LJ, k0 + 1)] + Hz[IHZ(i0, LJ - 1, k0)] + Hz[IHZ(i0, LJ - 1, k0 + 1)]);
            ACC(12, p, ch, sh, t);
        }
        for (int a = 0; a < o->ni; a++) for (int b = 0; b < o->nj; b++) {   /* zp, zn */
            long p = mi * sz + (long)a * o->nj + b; int i0 = LI + a, j0 = LJ + b;
            double t, t1, t2;
            t = 0.5 * (Ey[IEY(i0, j0, UK)] + Ey[IEY(i0 + 1, j0, UK)]);                         ACC(18, p, ce, se, t);
            t = -0.5 * (Ex[IEX(i0, j0, UK)] + Ex[IEX(i0, j0 + 1, UK)]);                        ACC(19, p, ce, se, t);
            /* Jx first accumulates the Hx average, then swaps it for the Hy average
             * with a subtract-then-add (farfield.js:210-226); Jy gets the Hx average. */
            t1 = 0.25 * (Hx[IHX(i0, j0, UK)] + Hx[IHX(i0 + 1, j0, UK)] + Hx[IHX(i0, j0, UK - 1)] + Hx[IHX(i0 + 1, j0, UK - 1)]);
            ACC(16, p, ch, sh, t1);
            t2 = -0.25 * (Hy[IHY(i0, j0, UK)] + Hy[IHY(i0, j0 + 1, UK)] + Hy[IHY(i0, j0, UK - 1)] + Hy[IHY(i0, j0 + 1, UK - 1)]);
            o->ff[16][0][p] = o->ff[16][0][p] - ch * t1 + ch * t2;
            o->ff[16][1][p] = o->ff[16][1][p] - sh * t1 + sh * t2;
            ACC(17, p, ch, sh, t1);
        }
        for (int a = 0; a < o->ni; a++) for (int b = 0; b < o->nj; b++) {
            long p = mi * sz + (long)a * o->nj + b; int i0 = LI + a, j0 = LJ + b;
            double t;
            t = -0.5 * (Ey[IEY(i0, j0, LK)] + Ey[IEY(i0 + 1, j0, LK)]);                        ACC(22, p, ce, se, t);
            t = 0.5 * (Ex[IEX(i0, j0, LK)] + Ex[IEX(i0, j0 + 1, LK)]);                         ACC(23, p, ce, se, t);
            t = 0.25 * (Hy[IHY(i0, j0, LK)] + Hy[IHY(i0, j0 + 1, LK)] + Hy[IHY(i0, j0, LK - 1)] + Hy[IHY(i0, j0 + 1, LK - 1)]);
            ACC(20, p, ch, sh, t);
            t = -0.25 * (Hx[IHX(i0, j0, LK)] + Hx[IHX(i0 + 1, j0, LK)] + Hx[IHX(i0, j0, LK - 1)] + Hx[IHX(i0 + 1, j0, LK - 1)]);
            ACC(21, p, ch, sh, t);
        }
    }
#undef ACC
}

/* ------------------------------------------------------------------ the loop -- */
/* Steps [ts0, ts0+count).  v_out / i_out: [count][n_sv] and [count][n_si] (may be
 * NULL).  Returns 0. */
/* As oracle_run, also filling p_out [count][n_probes] (may be NULL) with the point
 * probes, H ones after the H update and E ones after the injection
 * (fdtdSolver.js:96,116). */
int oracle_run_probes(Oracle *o, int ts0, int count, int boundary_mode, double *v_out, double *i_out,
                      double *p_out) {
    double vbuf[MAX_OBS], ibuf[MAX_OBS], pbuf[MAX_OBS];
    for (int s = 0; s < count; s++) {
        int ts = ts0 + s;
        oracle_update_h(o);
        oracle_update_h_cpml(o);
        oracle_capture_probes(o, 1, pbuf);
        oracle_capture_currents(o, ibuf);
        oracle_update_e(o, boundary_mode);
        oracle_update_e_cpml(o);
        oracle_inject(o, ts);
        oracle_capture_probes(o, 0, pbuf);
        oracle_capture_voltages(o, vbuf);
        oracle_ff_accumulate(o, ts);
        if (v_out) for (int n = 0; n < o->nsv; n++) v_out[(long)s * o->nsv + n] = vbuf[n];
        if (i_out) for (int n = 0; n < o->nsi; n++) i_out[(long)s * o->nsi + n] = ibuf[n];
        if (p_out) for (int n = 0; n < o->npr; n++) p_out[(long)s * o->npr + n] = pbuf[n];
    }
    return 0;
}

int oracle_run(Oracle *o, int ts0, int count, int boundary_mode, double *v_out, double *i_out) {
    return oracle_run_probes(o, ts0, count, boundary_mode, v_out, i_out, NULL);
}
