/*
 * A host-only stand-in for libfdtd_hip.so that implements the handful of C ABI calls
 * the addon's handle tests need (init, destroy, multi_create/run/synchronize/destroy,
 * last_error, device_count).  Contexts come from a static pool and keep their slot
 * after fdtd_hip_destroy, so a batch that runs a destroyed context gets an error
 * from multi_run_async instead of touching freed memory.  Every call appends a token
 * to an event log, which is written as JSON to $FDTD_HIP_STUB_REPORT when the process
 * exits: i<k> init of slot k, d<k> destroy of slot k, C batch create, R batch run,
 * X batch run refused, D batch destroy, M device_count (a marker the driver places),
 * L<name> a call that takes caller arrays reached the library.  Every context reports
 * fixed object counts (fdtd_hip_get_counts: 1 sampled voltage, 2 sampled currents,
 * 3 probes, 2 NTFF frequencies), so a driver can check the addon's length checks.
 * Used only by tests/test_addon_handles.py, which builds the addon against it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fdtd_hip.h"

#define POOL 64

struct FdtdHipContext { int alive; int destroyed; };
struct FdtdHipMulti { int alive; int n; FdtdHipContext *ctx[POOL]; };

static FdtdHipContext pool[POOL];
static FdtdHipMulti mpool[POOL];
static char err[256];
static char events[4096];

static void event(const char *tok, int k) {
  char t[16];
  snprintf(t, sizeof t, k >= 0 ? "%s%d " : "%s ", tok, k);
  if (strlen(events) + strlen(t) < sizeof events) strcat(events, t);
}

static int stub_fail(const char *msg) {
  snprintf(err, sizeof err, "%s", msg);
  return FDTD_HIP_ERR;
}

const char *fdtd_hip_last_error(void) { return err; }
int fdtd_hip_device_count(int *count) { event("M", -1); *count = 1; return FDTD_HIP_OK; }

int fdtd_hip_init(const FdtdHipDims *dims, FdtdHipContext **out) {
  (void)dims;
  for (int q = 0; q < POOL; ++q)
    if (!pool[q].alive && !pool[q].destroyed) {
      pool[q].alive = 1;
      event("i", q);
      *out = &pool[q];
      return FDTD_HIP_OK;
    }
  return stub_fail("init: stub pool exhausted");
}

void fdtd_hip_destroy(FdtdHipContext *c) {
  if (!c) return;
  c->alive = 0;
  c->destroyed += 1;     /* > 1 would be a double destroy */
  event("d", (int)(c - pool));
}

int fdtd_hip_multi_create(FdtdHipContext *const *ctxs, int n, const FdtdHipTuning *t, FdtdHipMulti **out) {
  (void)t;
  if (n < 1 || n > POOL) return stub_fail("multi_create: bad count");
  for (int q = 0; q < POOL; ++q)
    if (!mpool[q].alive) {
      mpool[q].alive = 1;
      mpool[q].n = n;
      event("C", -1);
      for (int v = 0; v < n; ++v) mpool[q].ctx[v] = ctxs[v];
      *out = &mpool[q];
      return FDTD_HIP_OK;
    }
  return stub_fail("multi_create: stub pool exhausted");
}

int fdtd_hip_multi_run_async(FdtdHipMulti *m, int ts, int count) {
  (void)ts; (void)count;
  if (!m->alive) return stub_fail("multi_run: batch destroyed");
  for (int v = 0; v < m->n; ++v)
    if (!m->ctx[v]->alive) {
      event("X", -1);
      return stub_fail("multi_run: a context of the batch was destroyed");
    }
  event("R", -1);
  return FDTD_HIP_OK;
}

int fdtd_hip_multi_synchronize(FdtdHipMulti *m) { return m->alive ? FDTD_HIP_OK : stub_fail("sync: batch destroyed"); }
void fdtd_hip_multi_destroy(FdtdHipMulti *m) {
  if (!m) return;
  m->alive = 0;
  event("D", -1);
}

int fdtd_hip_get_counts(FdtdHipContext *c, FdtdHipCounts *out) {
  (void)c;
  out->nSampledVoltages = 1; out->nSampledCurrents = 2; out->nProbes = 3;
  out->nSources = 1; out->nFarfieldFreq = 2; out->nPortDftFreq = 0;
  return FDTD_HIP_OK;
}

int fdtd_hip_run_batch(FdtdHipContext *c, int ts, int n, double *sv, double *si) {
  (void)c; (void)ts; (void)n; (void)sv; (void)si;
  event("Lrun", -1);
  return FDTD_HIP_OK;
}

int fdtd_hip_read_traces(FdtdHipContext *c, int ts, int n, double *sv, double *si) {
  (void)c; (void)ts; (void)n; (void)sv; (void)si;
  event("Ltraces", -1);
  return FDTD_HIP_OK;
}

int fdtd_hip_read_probes(FdtdHipContext *c, int ts, int n, double *out) {
  (void)c; (void)ts; (void)n; (void)out;
  event("Lprobes", -1);
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_cpml(FdtdHipContext *c, int faceId, int nc, const double *b_e, const double *a_e,
                         const double *b_m, const double *a_m, const double *CPsi0, long long n0,
                         const double *CPsi1, long long n1, const double *CPsi2, long long n2,
                         const double *CPsi3, long long n3, int m_start, int e_start) {
  (void)c; (void)faceId; (void)nc; (void)b_e; (void)a_e; (void)b_m; (void)a_m; (void)CPsi0; (void)n0;
  (void)CPsi1; (void)n1; (void)CPsi2; (void)n2; (void)CPsi3; (void)n3; (void)m_start; (void)e_start;
  event("Lcpml", -1);
  return FDTD_HIP_OK;
}

int fdtd_hip_directivity(FdtdHipContext *c, const double *freqHz, const double *radPower, int nAngles,
                         const double *theta, const double *phi, double *dataTheta, double *dataPhi) {
  (void)c; (void)freqHz; (void)radPower; (void)nAngles; (void)theta; (void)phi; (void)dataTheta; (void)dataPhi;
  event("Ldir", -1);
  return FDTD_HIP_OK;
}

__attribute__((destructor)) static void report(void) {
  const char *path = getenv("FDTD_HIP_STUB_REPORT");
  if (!path) return;
  int alive = 0, twice = 0;
  for (int q = 0; q < POOL; ++q) {
    alive += pool[q].alive;
    twice += pool[q].destroyed > 1;
  }
  FILE *f = fopen(path, "w");
  if (!f) return;
  fprintf(f, "{\"alive\": %d, \"destroyed_twice\": %d, \"events\": \"%s\"}\n", alive, twice, events);
  fclose(f);
}
