/*
 * napi_addon.c -- thin Node-API (C API, node_api.h) glue between the Node backend
 * (js/backends/hipBackend.mjs) and the C ABI of include/fdtd_hip.h.
 *
 * Plays the role of js_fdtd/src/backends/cuda/native/binding.cpp:1-186: typed arrays
 * cross zero-copy, every failing library call becomes a thrown JS Error whose message
 * is "<call>: <fdtd_hip_last_error()>".  A problem is an external handle instead of
 * the reference addon's process-global state.
 *
 *   deviceCount() -> number
 *   init({nx,ny,nz,dx,dy,dz,dt,numberOfTimeSteps,precision,boundary}) -> handle
 *   uploadCoeff(h, which, Float64Array)
 *   uploadCpml(h, faceId, {nc,b_e,a_e,b_m,a_m,CPsi0..CPsi3,m_start,e_start})
 *   uploadVSource(h, dir, Int32Array idx, Float64Array coef, Float64Array waveform)
 *   uploadSVoltage(h, dir, Int32Array idx, csvf)
 *   uploadSCurrent(h, Int32Array comp, Int32Array index, Float64Array weight, negate)
 *   uploadFarfield(h, nFreq, Int32Array box, Float64Array phase)
 *   uploadPortDft(h, Float64Array freqHz)
 *   runBatch(h, tsStart, count, Float64Array|null sv, Float64Array|null si)
 *   readField(h, which, Float64Array); writeField(h, which, Float64Array)
 *   setTuning(h, {planesPerWorkgroup, cellsPerThread, coefficientStorage, fuseCpml})
 *   getTuning(h) -> tuning; autotune(h, stepsPerCandidate) -> tuning
 *   readFarfield(h, which, Float64Array re, Float64Array im)
 *   readPortDft(h, Float64Array re, Float64Array im)
 *   radiatedPower(h, Float64Array out[nFreq])
 *   directivity(h, freqHz, radPower|null, theta, phi, dataTheta, dataPhi)
 *   setTiming(h, bool); kernelTimes(h) -> {ms:[6], launches:[6]}
 *   destroy(h)                     (throws while a live batch holds h)
 *   readTraces(h, tsStart, count, Float64Array|null sv, Float64Array|null si)
 *   uploadFieldProbe(h, magnetic, component, i, j, k); readProbes(h, tsStart, count, Float64Array)
 *   multiCreate([h, ...], tuning|null) -> batch handle   (fdtd_hip_multi_*: many
 *   multiRun(m, tsStart, count); multiSynchronize(m)       problems, one launch per
 *   multiAutotune(m, stepsPerCandidate) -> tuning          phase)
 *   multiDestroy(m)
 * Handles stay valid until destroyed (a second destroy throws); destroy a batch before
 * the problems it steps.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fdtd_hip.h"

#define CALL(env, expr)                                                    \
  do {                                                                     \
    if ((expr) != napi_ok) { throw_napi(env, #expr); return NULL; }        \
  } while (0)

static void throw_napi(napi_env env, const char *what) {
  const napi_extended_error_info *info = NULL;
  napi_get_last_error_info(env, &info);
  char buf[512];
  snprintf(buf, sizeof buf, "N-API failure in %s: %s", what,
           info && info->error_message ? info->error_message : "unknown");
  napi_throw_error(env, NULL, buf);
}

static napi_value throw_lib(napi_env env, const char *ctx) {
  char buf[1200];
  snprintf(buf, sizeof buf, "%s: %s", ctx, fdtd_hip_last_error());
  napi_throw_error(env, NULL, buf);
  return NULL;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

/* Typed array data pointer of the expected element type, or NULL (with a thrown
 * TypeError) when the value is not one.  allow_null accepts JS null/undefined. */
static void *typed(napi_env env, napi_value v, napi_typedarray_type want, size_t *len, int allow_null, int *ok) {
  napi_valuetype vt;
  *ok = 0;
  if (len) *len = 0;
  if (napi_typeof(env, v, &vt) != napi_ok) return NULL;
  if (allow_null && (vt == napi_null || vt == napi_undefined)) { *ok = 1; return NULL; }
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) { napi_throw_type_error(env, NULL, "expected a typed array"); return NULL; }
  napi_typedarray_type t;
  size_t n;
  void *data;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, &n, &data, &ab, &off) != napi_ok) return NULL;
  if (t != want) {
    napi_throw_type_error(env, NULL, want == napi_float64_array ? "expected a Float64Array" : "expected an Int32Array");
    return NULL;
  }
  if (len) *len = n;
  *ok = 1;
  return data;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) { napi_throw_type_error(env, NULL, "missing arguments"); return 0; }
  return 1;
}

/* The context's object counts (fdtd_hip_get_counts), with a thrown error on failure. */
static int counts_of(napi_env env, FdtdHipContext *c, FdtdHipCounts *out, const char *who) {
  if (fdtd_hip_get_counts(c, out) != FDTD_HIP_OK) { throw_lib(env, who); return 0; }
  return 1;
}

/* RangeError unless a typed array of n elements holds per * count values. */
static int need(napi_env env, size_t n, long long per, int count, const char *what) {
  if (count < 0 || (long long)n >= per * (long long)count) return 1;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %zu elements, need %lld (%lld x %d steps)", what, n, per * (long long)count, per, count);
  napi_throw_range_error(env, NULL, buf);
  return 0;
}

static int get_int(napi_env env, napi_value v, int *out) {
  if (napi_get_value_int32(env, v, out) != napi_ok) { napi_throw_type_error(env, NULL, "expected a number"); return 0; }
  return 1;
}

static int get_double(napi_env env, napi_value v, double *out) {
  if (napi_get_value_double(env, v, out) != napi_ok) { napi_throw_type_error(env, NULL, "expected a number"); return 0; }
  return 1;
}

static int prop(napi_env env, napi_value obj, const char *key, napi_value *out) {
  if (napi_get_named_property(env, obj, key, out) != napi_ok) { throw_napi(env, key); return 0; }
  return 1;
}

static int prop_int(napi_env env, napi_value obj, const char *key, int *out) {
  napi_value v;
  return prop(env, obj, key, &v) && get_int(env, v, out);
}

static int prop_double(napi_env env, napi_value obj, const char *key, double *out) {
  napi_value v;
  return prop(env, obj, key, &v) && get_double(env, v, out);
}

/* A handle is an external around a HandleBox.  destroy()/multiDestroy() free the
 * library object and clear the box, so a second call on the same handle throws
 * instead of touching freed memory; the box itself goes with the JS value, and a
 * handle collected without destroy() frees its library object then.
 *
 * A batch borrows its problems' contexts, so a problem box counts the live batches
 * that hold it.  destroy() of a problem in a live batch throws.  A problem handle
 * collected while a batch still holds it keeps its context (and box) until the last
 * such batch is released, whatever order the collector finalizes them in. */
typedef struct HandleBox {
  void *p;
  int kind;                  /* 1 = problem (FdtdHipContext), 2 = batch (FdtdHipMulti) */
  int batches;               /* problem: live batches that hold it */
  int collected;             /* problem: its JS value is gone, release on the last batch */
  struct HandleBox **members; /* batch: the problem boxes it holds */
  uint32_t nmembers;
} HandleBox;

static void problem_release(HandleBox *b) {
  if (b->p) fdtd_hip_destroy((FdtdHipContext *)b->p);
  b->p = NULL;
  free(b);
}

/* Destroys a batch's library object and lets go of its problems. */
static void batch_release(HandleBox *b) {
  if (b->p) fdtd_hip_multi_destroy((FdtdHipMulti *)b->p);
  b->p = NULL;
  for (uint32_t q = 0; q < b->nmembers; ++q) {
    HandleBox *m = b->members[q];
    if (--m->batches == 0 && m->collected) problem_release(m);
  }
  free(b->members);
  b->members = NULL;
  b->nmembers = 0;
}

static void box_finalize(napi_env env, void *data, void *hint) {
  (void)env; (void)hint;
  HandleBox *b = (HandleBox *)data;
  if (b->kind == 1) {
    if (b->batches > 0) b->collected = 1;
    else problem_release(b);
  } else {
    batch_release(b);
    free(b);
  }
}

static napi_value make_handle(napi_env env, void *p, int kind) {
  HandleBox *b = (HandleBox *)calloc(1, sizeof *b);
  napi_value ext;
  if (!b) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  b->p = p;
  b->kind = kind;
  if (napi_create_external(env, b, box_finalize, NULL, &ext) != napi_ok) {
    b->p = NULL;
    free(b);
    throw_napi(env, "napi_create_external");
    return NULL;
  }
  return ext;
}

static HandleBox *box_of(napi_env env, napi_value v, int kind) {
  void *p = NULL;
  napi_valuetype vt;
  if (napi_typeof(env, v, &vt) != napi_ok || vt != napi_external || napi_get_value_external(env, v, &p) != napi_ok ||
      !p || ((HandleBox *)p)->kind != kind || !((HandleBox *)p)->p) {
    napi_throw_type_error(env, NULL, kind == 1 ? "expected a live problem handle" : "expected a live batch handle");
    return NULL;
  }
  return (HandleBox *)p;
}

static FdtdHipContext *handle(napi_env env, napi_value v) {
  HandleBox *b = box_of(env, v, 1);
  return b ? (FdtdHipContext *)b->p : NULL;
}

static napi_value DeviceCount(napi_env env, napi_callback_info info) {
  (void)info;
  int n = 0;
  if (fdtd_hip_device_count(&n) != FDTD_HIP_OK) n = 0;
  napi_value r;
  CALL(env, napi_create_int32(env, n, &r));
  return r;
}

static napi_value Init(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  FdtdHipDims d;
  memset(&d, 0, sizeof d);
  if (!prop_int(env, argv[0], "nx", &d.nx) || !prop_int(env, argv[0], "ny", &d.ny) ||
      !prop_int(env, argv[0], "nz", &d.nz) || !prop_double(env, argv[0], "dx", &d.dx) ||
      !prop_double(env, argv[0], "dy", &d.dy) || !prop_double(env, argv[0], "dz", &d.dz) ||
      !prop_double(env, argv[0], "dt", &d.dt) ||
      !prop_int(env, argv[0], "numberOfTimeSteps", &d.numberOfTimeSteps) ||
      !prop_int(env, argv[0], "precision", &d.precision) || !prop_int(env, argv[0], "boundary", &d.boundary))
    return NULL;
  FdtdHipContext *ctx = NULL;
  if (fdtd_hip_init(&d, &ctx) != FDTD_HIP_OK) return throw_lib(env, "init");
  napi_value ext = make_handle(env, ctx, 1);
  if (!ext) fdtd_hip_destroy(ctx);
  return ext;
}

static napi_value UploadCoeff(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int which, ok;
  size_t n;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &which)) return NULL;
  double *p = typed(env, argv[2], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  if (fdtd_hip_upload_coeff(c, which, p, (long long)n) != FDTD_HIP_OK) return throw_lib(env, "uploadCoeff");
  return undefined(env);
}

static napi_value UploadCpml(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int face, nc, ms, es, ok;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &face)) return NULL;
  napi_value o = argv[2], v;
  if (!prop_int(env, o, "nc", &nc) || !prop_int(env, o, "m_start", &ms) || !prop_int(env, o, "e_start", &es)) return NULL;
  const char *pk[4] = {"b_e", "a_e", "b_m", "a_m"};
  double *prof[4];
  for (int q = 0; q < 4; ++q) {
    size_t n;
    if (!prop(env, o, pk[q], &v)) return NULL;
    prof[q] = typed(env, v, napi_float64_array, &n, 0, &ok);
    if (!ok) return NULL;
    if (nc < 0 || n < (size_t)nc) {
      char buf[160];
      snprintf(buf, sizeof buf, "uploadCpml: profile %s has %zu entries, need nc = %d", pk[q], n, nc);
      napi_throw_range_error(env, NULL, buf);
      return NULL;
    }
  }
  const char *ck[4] = {"CPsi0", "CPsi1", "CPsi2", "CPsi3"};
  double *cp[4];
  size_t cn[4];
  for (int q = 0; q < 4; ++q) {
    if (!prop(env, o, ck[q], &v)) return NULL;
    cp[q] = typed(env, v, napi_float64_array, &cn[q], 0, &ok);
    if (!ok) return NULL;
  }
  if (fdtd_hip_upload_cpml(c, face, nc, prof[0], prof[1], prof[2], prof[3], cp[0], (long long)cn[0], cp[1],
                           (long long)cn[1], cp[2], (long long)cn[2], cp[3], (long long)cn[3], ms, es) != FDTD_HIP_OK)
    return throw_lib(env, "uploadCpml");
  return undefined(env);
}

static napi_value UploadVSource(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  int dir, ok;
  size_t ni, nc, nw;
  if (!get_args(env, info, 5, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &dir)) return NULL;
  int *idx = typed(env, argv[2], napi_int32_array, &ni, 0, &ok);
  if (!ok) return NULL;
  double *coef = typed(env, argv[3], napi_float64_array, &nc, 0, &ok);
  if (!ok) return NULL;
  double *wave = typed(env, argv[4], napi_float64_array, &nw, 0, &ok);
  if (!ok) return NULL;
  if (ni != nc) { napi_throw_range_error(env, NULL, "uploadVSource: idx and coef lengths differ"); return NULL; }
  if (fdtd_hip_upload_vsource(c, dir, idx, coef, (int)ni, wave, (int)nw) != FDTD_HIP_OK) return throw_lib(env, "uploadVSource");
  return undefined(env);
}

static napi_value UploadSVoltage(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  int dir, ok;
  size_t n;
  double csvf;
  if (!get_args(env, info, 4, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &dir)) return NULL;
  int *idx = typed(env, argv[2], napi_int32_array, &n, 0, &ok);
  if (!ok || !get_double(env, argv[3], &csvf)) return NULL;
  if (fdtd_hip_upload_svoltage(c, dir, idx, (int)n, csvf) != FDTD_HIP_OK) return throw_lib(env, "uploadSVoltage");
  return undefined(env);
}

static napi_value UploadSCurrent(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  int neg, ok;
  size_t n1, n2, n3;
  if (!get_args(env, info, 5, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  int *comp = typed(env, argv[1], napi_int32_array, &n1, 0, &ok);
  if (!ok) return NULL;
  int *index = typed(env, argv[2], napi_int32_array, &n2, 0, &ok);
  if (!ok) return NULL;
  double *w = typed(env, argv[3], napi_float64_array, &n3, 0, &ok);
  if (!ok || !get_int(env, argv[4], &neg)) return NULL;
  if (n1 != n2 || n1 != n3) { napi_throw_range_error(env, NULL, "uploadSCurrent: term lists differ in length"); return NULL; }
  if (fdtd_hip_upload_scurrent(c, comp, index, w, (int)n1, neg) != FDTD_HIP_OK) return throw_lib(env, "uploadSCurrent");
  return undefined(env);
}

static napi_value UploadFarfield(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  int nf, ok;
  size_t nb, np;
  if (!get_args(env, info, 4, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &nf)) return NULL;
  int *box = typed(env, argv[2], napi_int32_array, &nb, 0, &ok);
  if (!ok) return NULL;
  double *ph = typed(env, argv[3], napi_float64_array, &np, 0, &ok);
  if (!ok) return NULL;
  if (nb != 6 || nf <= 0) { napi_throw_range_error(env, NULL, "uploadFarfield: box needs 6 entries"); return NULL; }
  if (fdtd_hip_upload_farfield(c, nf, box, ph, (int)(np / (4 * (size_t)nf))) != FDTD_HIP_OK) return throw_lib(env, "uploadFarfield");
  return undefined(env);
}

static napi_value UploadPortDft(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  int ok;
  size_t n;
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  double *f = typed(env, argv[1], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  if (fdtd_hip_upload_port_dft(c, (int)n, f) != FDTD_HIP_OK) return throw_lib(env, "uploadPortDft");
  return undefined(env);
}

static napi_value RunBatch(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  int ts, count, ok;
  if (!get_args(env, info, 5, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &ts) || !get_int(env, argv[2], &count)) return NULL;
  size_t nv, ni;
  double *sv = typed(env, argv[3], napi_float64_array, &nv, 1, &ok);
  if (!ok) return NULL;
  double *si = typed(env, argv[4], napi_float64_array, &ni, 1, &ok);
  if (!ok) return NULL;
  FdtdHipCounts k;
  if (!counts_of(env, c, &k, "runBatch")) return NULL;
  if ((sv && !need(env, nv, k.nSampledVoltages, count, "runBatch: voltage traces")) ||
      (si && !need(env, ni, k.nSampledCurrents, count, "runBatch: current traces")))
    return NULL;
  if (fdtd_hip_run_batch(c, ts, count, sv, si) != FDTD_HIP_OK) return throw_lib(env, "runBatch");
  return undefined(env);
}

static napi_value ReadField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int which, ok;
  size_t n;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &which)) return NULL;
  double *p = typed(env, argv[2], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  if (fdtd_hip_read_field(c, which, p, (long long)n) != FDTD_HIP_OK) return throw_lib(env, "readField");
  return undefined(env);
}

static napi_value ReadFarfield(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  int which, ok;
  size_t n1, n2;
  if (!get_args(env, info, 4, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &which)) return NULL;
  double *re = typed(env, argv[2], napi_float64_array, &n1, 0, &ok);
  if (!ok) return NULL;
  double *im = typed(env, argv[3], napi_float64_array, &n2, 0, &ok);
  if (!ok) return NULL;
  if (n1 != n2) { napi_throw_range_error(env, NULL, "readFarfield: re/im lengths differ"); return NULL; }
  if (fdtd_hip_read_farfield(c, which, re, im, (long long)n1) != FDTD_HIP_OK) return throw_lib(env, "readFarfield");
  return undefined(env);
}

static napi_value ReadPortDft(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int ok;
  size_t n1, n2;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  double *re = typed(env, argv[1], napi_float64_array, &n1, 0, &ok);
  if (!ok) return NULL;
  double *im = typed(env, argv[2], napi_float64_array, &n2, 0, &ok);
  if (!ok) return NULL;
  if (n1 != n2) { napi_throw_range_error(env, NULL, "readPortDft: re/im lengths differ"); return NULL; }
  if (fdtd_hip_read_port_dft(c, re, im, (long long)n1) != FDTD_HIP_OK) return throw_lib(env, "readPortDft");
  return undefined(env);
}

static napi_value RadiatedPower(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  int ok;
  size_t n;
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  double *out = typed(env, argv[1], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  if (fdtd_hip_radiated_power(c, out, (int)n) != FDTD_HIP_OK) return throw_lib(env, "radiatedPower");
  return undefined(env);
}

static napi_value Directivity(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  int ok;
  size_t nf, np, na, nb, nt, nq;
  if (!get_args(env, info, 7, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  double *freq = typed(env, argv[1], napi_float64_array, &nf, 0, &ok);
  if (!ok) return NULL;
  double *prad = typed(env, argv[2], napi_float64_array, &np, 1, &ok);
  if (!ok) return NULL;
  double *th = typed(env, argv[3], napi_float64_array, &na, 0, &ok);
  if (!ok) return NULL;
  double *ph = typed(env, argv[4], napi_float64_array, &nb, 0, &ok);
  if (!ok) return NULL;
  double *dt = typed(env, argv[5], napi_float64_array, &nt, 0, &ok);
  if (!ok) return NULL;
  double *dp = typed(env, argv[6], napi_float64_array, &nq, 0, &ok);
  if (!ok) return NULL;
  FdtdHipCounts k;
  if (!counts_of(env, c, &k, "directivity")) return NULL;
  if (nf != (size_t)k.nFarfieldFreq) {
    char buf[160];
    snprintf(buf, sizeof buf, "directivity: %zu frequencies, the far-field box has %d", nf, k.nFarfieldFreq);
    napi_throw_range_error(env, NULL, buf);
    return NULL;
  }
  if (na != nb || nt != nq || nt != nf * na || (prad && np != nf)) {
    napi_throw_range_error(env, NULL, "directivity: array lengths do not match nFreq x nAngles");
    return NULL;
  }
  if (fdtd_hip_directivity(c, freq, prad, (int)na, th, ph, dt, dp) != FDTD_HIP_OK) return throw_lib(env, "directivity");
  return undefined(env);
}

static napi_value SetTiming(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bool on = false;
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  napi_get_value_bool(env, argv[1], &on);
  if (fdtd_hip_set_timing(c, on ? 1 : 0) != FDTD_HIP_OK) return throw_lib(env, "setTiming");
  return undefined(env);
}

static napi_value KernelTimes(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  FdtdHipKernelTimes t;
  if (fdtd_hip_get_kernel_times(c, &t) != FDTD_HIP_OK) return throw_lib(env, "kernelTimes");
  napi_value obj, ms, la, x;
  CALL(env, napi_create_object(env, &obj));
  CALL(env, napi_create_array_with_length(env, 6, &ms));
  CALL(env, napi_create_array_with_length(env, 6, &la));
  for (uint32_t q = 0; q < 6; ++q) {
    CALL(env, napi_create_double(env, t.ms[q], &x));
    CALL(env, napi_set_element(env, ms, q, x));
    CALL(env, napi_create_double(env, (double)t.launches[q], &x));
    CALL(env, napi_set_element(env, la, q, x));
  }
  CALL(env, napi_set_named_property(env, obj, "ms", ms));
  CALL(env, napi_set_named_property(env, obj, "launches", la));
  return obj;
}

static napi_value WriteField(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int which, ok;
  size_t n;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &which)) return NULL;
  double *p = typed(env, argv[2], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  if (fdtd_hip_write_field(c, which, p, (long long)n) != FDTD_HIP_OK) return throw_lib(env, "writeField");
  return undefined(env);
}

static napi_value tuning_object(napi_env env, const FdtdHipTuning *t) {
  napi_value obj, x;
  CALL(env, napi_create_object(env, &obj));
  CALL(env, napi_create_int32(env, t->planes_per_workgroup, &x));
  CALL(env, napi_set_named_property(env, obj, "planesPerWorkgroup", x));
  CALL(env, napi_create_int32(env, t->cells_per_thread, &x));
  CALL(env, napi_set_named_property(env, obj, "cellsPerThread", x));
  CALL(env, napi_create_int32(env, t->coefficient_storage, &x));
  CALL(env, napi_set_named_property(env, obj, "coefficientStorage", x));
  CALL(env, napi_create_int32(env, t->fuse_cpml, &x));
  CALL(env, napi_set_named_property(env, obj, "fuseCpml", x));
  return obj;
}

/* setTuning(h, {planesPerWorkgroup, cellsPerThread, coefficientStorage, fuseCpml});
 * absent or zero members select the library defaults. */
static napi_value SetTuning(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  FdtdHipTuning t;
  memset(&t, 0, sizeof t);
  const char *keys[4] = {"planesPerWorkgroup", "cellsPerThread", "coefficientStorage", "fuseCpml"};
  int *dst[4] = {&t.planes_per_workgroup, &t.cells_per_thread, &t.coefficient_storage, &t.fuse_cpml};
  for (int q = 0; q < 4; ++q) {
    bool has = false;
    if (napi_has_named_property(env, argv[1], keys[q], &has) != napi_ok) { throw_napi(env, keys[q]); return NULL; }
    if (has && !prop_int(env, argv[1], keys[q], dst[q])) return NULL;
  }
  if (fdtd_hip_set_tuning(c, &t) != FDTD_HIP_OK) return throw_lib(env, "setTuning");
  return undefined(env);
}

static napi_value GetTuning(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  FdtdHipTuning t;
  if (fdtd_hip_get_tuning(c, &t) != FDTD_HIP_OK) return throw_lib(env, "getTuning");
  return tuning_object(env, &t);
}

/* autotune(h, stepsPerCandidate) -> chosen tuning; only before the first runBatch. */
static napi_value Autotune(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  int steps;
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &steps)) return NULL;
  FdtdHipTuning t;
  if (fdtd_hip_autotune(c, steps, &t) != FDTD_HIP_OK) return throw_lib(env, "autotune");
  return tuning_object(env, &t);
}

static napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  HandleBox *b = box_of(env, argv[0], 1);
  if (!b) return NULL;
  if (b->batches > 0) {
    napi_throw_error(env, NULL, "destroy: the problem is part of a live batch; call multiDestroy on the batch first");
    return NULL;
  }
  fdtd_hip_destroy((FdtdHipContext *)b->p);
  b->p = NULL;
  return undefined(env);
}

static napi_value ReadTraces(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  int ts, count, ok;
  if (!get_args(env, info, 5, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &ts) || !get_int(env, argv[2], &count)) return NULL;
  size_t nv, ni;
  double *sv = typed(env, argv[3], napi_float64_array, &nv, 1, &ok);
  if (!ok) return NULL;
  double *si = typed(env, argv[4], napi_float64_array, &ni, 1, &ok);
  if (!ok) return NULL;
  FdtdHipCounts k;
  if (!counts_of(env, c, &k, "readTraces")) return NULL;
  if ((sv && !need(env, nv, k.nSampledVoltages, count, "readTraces: voltage traces")) ||
      (si && !need(env, ni, k.nSampledCurrents, count, "readTraces: current traces")))
    return NULL;
  if (fdtd_hip_read_traces(c, ts, count, sv, si) != FDTD_HIP_OK) return throw_lib(env, "readTraces");
  return undefined(env);
}

static napi_value UploadFieldProbe(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  int a[5];
  if (!get_args(env, info, 6, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c) return NULL;
  for (int q = 0; q < 5; ++q)
    if (!get_int(env, argv[q + 1], &a[q])) return NULL;
  if (fdtd_hip_upload_field_probe(c, a[0], a[1], a[2], a[3], a[4]) != FDTD_HIP_OK)
    return throw_lib(env, "uploadFieldProbe");
  return undefined(env);
}

static napi_value ReadProbes(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  int ts, count, ok;
  if (!get_args(env, info, 4, argv)) return NULL;
  FdtdHipContext *c = handle(env, argv[0]);
  if (!c || !get_int(env, argv[1], &ts) || !get_int(env, argv[2], &count)) return NULL;
  size_t n;
  double *out = typed(env, argv[3], napi_float64_array, &n, 0, &ok);
  if (!ok) return NULL;
  FdtdHipCounts k;
  if (!counts_of(env, c, &k, "readProbes") || !need(env, n, k.nProbes, count, "readProbes")) return NULL;
  if (fdtd_hip_read_probes(c, ts, count, out) != FDTD_HIP_OK) return throw_lib(env, "readProbes");
  return undefined(env);
}

static FdtdHipMulti *multi_handle(napi_env env, napi_value v) {
  HandleBox *b = box_of(env, v, 2);
  return b ? (FdtdHipMulti *)b->p : NULL;
}

static int tuning_from(napi_env env, napi_value obj, FdtdHipTuning *t) {
  const char *keys[4] = {"planesPerWorkgroup", "cellsPerThread", "coefficientStorage", "fuseCpml"};
  int *dst[4] = {&t->planes_per_workgroup, &t->cells_per_thread, &t->coefficient_storage, &t->fuse_cpml};
  memset(t, 0, sizeof *t);
  for (int q = 0; q < 4; ++q) {
    bool has = false;
    if (napi_has_named_property(env, obj, keys[q], &has) != napi_ok) { throw_napi(env, keys[q]); return 0; }
    if (has && !prop_int(env, obj, keys[q], dst[q])) return 0;
  }
  return 1;
}

static napi_value MultiCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  bool is_array = false;
  uint32_t n = 0;
  if (napi_is_array(env, argv[0], &is_array) != napi_ok || !is_array ||
      napi_get_array_length(env, argv[0], &n) != napi_ok || n == 0) {
    napi_throw_type_error(env, NULL, "expected a non-empty array of problem handles");
    return NULL;
  }
  FdtdHipContext *ctx[4096];
  HandleBox *box[4096];
  if (n > 4096) { napi_throw_range_error(env, NULL, "at most 4096 problems per batch"); return NULL; }
  for (uint32_t q = 0; q < n; ++q) {
    napi_value e;
    CALL(env, napi_get_element(env, argv[0], q, &e));
    if (!(box[q] = box_of(env, e, 1))) return NULL;
    ctx[q] = (FdtdHipContext *)box[q]->p;
  }
  napi_valuetype vt;
  CALL(env, napi_typeof(env, argv[1], &vt));
  FdtdHipTuning t;
  const FdtdHipTuning *tp = NULL;
  if (vt != napi_null && vt != napi_undefined) {
    if (!tuning_from(env, argv[1], &t)) return NULL;
    tp = &t;
  }
  HandleBox **members = (HandleBox **)malloc(n * sizeof *members);
  if (!members) { napi_throw_error(env, NULL, "out of host memory"); return NULL; }
  FdtdHipMulti *m = NULL;
  if (fdtd_hip_multi_create(ctx, (int)n, tp, &m) != FDTD_HIP_OK) {
    free(members);
    return throw_lib(env, "multiCreate");
  }
  napi_value ext = make_handle(env, m, 2);
  if (!ext) {
    fdtd_hip_multi_destroy(m);
    free(members);
    return NULL;
  }
  HandleBox *mb = NULL;
  if (napi_get_value_external(env, ext, (void **)&mb) != napi_ok || !mb) {
    free(members);     /* the box owns m and releases it when collected */
    throw_napi(env, "napi_get_value_external");
    return NULL;
  }
  for (uint32_t q = 0; q < n; ++q) {
    members[q] = box[q];
    ++box[q]->batches;
  }
  mb->members = members;
  mb->nmembers = n;
  return ext;
}

static napi_value MultiRun(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  int ts, count;
  if (!get_args(env, info, 3, argv)) return NULL;
  FdtdHipMulti *m = multi_handle(env, argv[0]);
  if (!m || !get_int(env, argv[1], &ts) || !get_int(env, argv[2], &count)) return NULL;
  if (fdtd_hip_multi_run_async(m, ts, count) != FDTD_HIP_OK) return throw_lib(env, "multiRun");
  return undefined(env);
}

static napi_value MultiSynchronize(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  FdtdHipMulti *m = multi_handle(env, argv[0]);
  if (!m) return NULL;
  if (fdtd_hip_multi_synchronize(m) != FDTD_HIP_OK) return throw_lib(env, "multiSynchronize");
  return undefined(env);
}

static napi_value MultiAutotune(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  int steps;
  if (!get_args(env, info, 2, argv)) return NULL;
  FdtdHipMulti *m = multi_handle(env, argv[0]);
  if (!m || !get_int(env, argv[1], &steps)) return NULL;
  FdtdHipTuning t;
  if (fdtd_hip_multi_autotune(m, steps, &t) != FDTD_HIP_OK) return throw_lib(env, "multiAutotune");
  return tuning_object(env, &t);
}

static napi_value MultiDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  HandleBox *b = box_of(env, argv[0], 2);
  if (!b) return NULL;
  batch_release(b);
  return undefined(env);
}

static napi_value Register(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
    {"deviceCount", NULL, DeviceCount, NULL, NULL, NULL, napi_enumerable, NULL},
    {"init", NULL, Init, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadCoeff", NULL, UploadCoeff, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadCpml", NULL, UploadCpml, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadVSource", NULL, UploadVSource, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadSVoltage", NULL, UploadSVoltage, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadSCurrent", NULL, UploadSCurrent, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadFarfield", NULL, UploadFarfield, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadPortDft", NULL, UploadPortDft, NULL, NULL, NULL, napi_enumerable, NULL},
    {"runBatch", NULL, RunBatch, NULL, NULL, NULL, napi_enumerable, NULL},
    {"readField", NULL, ReadField, NULL, NULL, NULL, napi_enumerable, NULL},
    {"readFarfield", NULL, ReadFarfield, NULL, NULL, NULL, napi_enumerable, NULL},
    {"readPortDft", NULL, ReadPortDft, NULL, NULL, NULL, napi_enumerable, NULL},
    {"writeField", NULL, WriteField, NULL, NULL, NULL, napi_enumerable, NULL},
    {"setTuning", NULL, SetTuning, NULL, NULL, NULL, napi_enumerable, NULL},
    {"getTuning", NULL, GetTuning, NULL, NULL, NULL, napi_enumerable, NULL},
    {"autotune", NULL, Autotune, NULL, NULL, NULL, napi_enumerable, NULL},
    {"radiatedPower", NULL, RadiatedPower, NULL, NULL, NULL, napi_enumerable, NULL},
    {"directivity", NULL, Directivity, NULL, NULL, NULL, napi_enumerable, NULL},
    {"setTiming", NULL, SetTiming, NULL, NULL, NULL, napi_enumerable, NULL},
    {"kernelTimes", NULL, KernelTimes, NULL, NULL, NULL, napi_enumerable, NULL},
    {"destroy", NULL, Destroy, NULL, NULL, NULL, napi_enumerable, NULL},
    {"readTraces", NULL, ReadTraces, NULL, NULL, NULL, napi_enumerable, NULL},
    {"uploadFieldProbe", NULL, UploadFieldProbe, NULL, NULL, NULL, napi_enumerable, NULL},
    {"readProbes", NULL, ReadProbes, NULL, NULL, NULL, napi_enumerable, NULL},
    {"multiCreate", NULL, MultiCreate, NULL, NULL, NULL, napi_enumerable, NULL},
    {"multiRun", NULL, MultiRun, NULL, NULL, NULL, napi_enumerable, NULL},
    {"multiSynchronize", NULL, MultiSynchronize, NULL, NULL, NULL, napi_enumerable, NULL},
    {"multiAutotune", NULL, MultiAutotune, NULL, NULL, NULL, napi_enumerable, NULL},
    {"multiDestroy", NULL, MultiDestroy, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  if (napi_define_properties(env, exports, sizeof d / sizeof d[0], d) != napi_ok) {
    throw_napi(env, "define_properties");
    return NULL;
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Register)
