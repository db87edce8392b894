// fdtd_hip.hip -- host side of the gfx950 Yee time step behind the C ABI of
// include/fdtd_hip.h.  Owns device memory, converts between the reference array
// layouts and the device layout (see fdtd_kernels.hpp), and sequences the per-step
// kernels on one HIP stream.

#include "../../include/fdtd_hip.h"
#include "fdtd_kernels.hpp"
#include "ntff_post.hpp"
#include "ntff_block.hpp"
#include "multi_kernels.hpp"
#include "e_tail.hpp"
#include "earlystop_net.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace fdtd;

namespace {

thread_local std::string g_last_error = "no error";

int fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return FDTD_HIP_ERR;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) return fail("%s: %s (%s:%d)", #expr, hipGetErrorString(_e),  \
                                      __FILE__, __LINE__);                             \
  } while (0)

// Component shapes as offsets from (nx, ny, nz), indexed Hx,Hy,Hz,Ex,Ey,Ez.
const int kExt[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
// Owning component of each coefficient id (reference buffer order).
const int kCoeffOwner[NCOEFF] = {3, 3, 3, 4, 4, 4, 5, 5, 5, 0, 0, 0, 1, 1, 1, 2, 2, 2};
// CPML slot -> component, per face axis (x: Hy,Hz,Ey,Ez; y: Hx,Hz,Ez,Ex; z: Hx,Hy,Ex,Ey).
const int kSlotComp[3][4] = {{1, 2, 4, 5}, {0, 2, 5, 3}, {0, 1, 3, 4}};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace

// Fixed device workspace of fdtd_hip_directivity_noalloc (staged inputs, wave numbers,
// angle trigonometry, radiated power, outputs), laid out as DirLayout for the maximum
// sizes: 2*64 + 64 + 2*8192 + 4*8192 + 2*65536 doubles, about 1.4 MB per device, held
// until the process ends.  Host code never uses the symbol's name as an address: each
// device's copy is found once with hipGetSymbolAddress (noalloc_register).
constexpr long NOALLOC_DOUBLES =
    3L * FDTD_HIP_NOALLOC_MAX_FREQ + 6L * FDTD_HIP_NOALLOC_MAX_ANGLES + 2L * FDTD_HIP_NOALLOC_MAX_OUTPUTS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;

// Six field arrays of at most this size stay in the 256 MB last-level cache between the
// two phases of a step (the IFA's take 19 MB in FP32), so the stencils use plain loads
// and stores there: on the IFA that took the FP32 step from 43.2 to 39.2 us.  Larger
// grids stream their own fields past the cache.
constexpr size_t CACHE_RESIDENT_FIELD_BYTES = 64ull << 20;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on `device` (current), looked up once per device.
int noalloc_register(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail("init: device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return FDTD_HIP_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return fail("init: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  g_noalloc_ptr[device] = (double *)ptr;
  return FDTD_HIP_OK;
}

double *noalloc_workspace(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

}  // namespace

struct FdtdHipContext {
  FdtdHipDims d{};
  int SJ = 0, SI = 0;
  long long ntot = 0;           // elements per uniform array
  size_t esz = 8;               // bytes per field element
  int device = 0;
  hipStream_t stream = nullptr;
  void *fields = nullptr;       // 6 uniform arrays
  void *coeffs = nullptr;       // 18 uniform arrays
  size_t stride_bytes = 0;      // bytes between consecutive uniform arrays

  struct HostFace {
    int nc = 0, m_start = 0, e_start = 0;
    void *prof = nullptr;       // b_e, a_e, b_m, a_m (T, nc each)
    void *psi[4] = {};
    void *cpsi[4] = {};
    int cpsi_mask[4] = {-1, -1, -1, -1};   // 0: every CPsi element of the slot is equal
    long long n[4] = {};
  } face[6];

  std::vector<void *> small_allocs;   // port tables, traces
  PortState ports{};
  std::vector<int> probe_host;        // [npr][PROBE_INTS], uploaded with the traces
  // E elements (component 0..2 = Ex..Ez, i, j, k) the port epilogue writes (sources)
  // and reads (voltage observers, E probes): whether the E-phase tail may run as one
  // launch depends on them.
  std::vector<ECell> tail_e_write, tail_e_read;
  int e_tail_ok = 0;                  // e_tail_mergeable, as of the current run
  int e_tail_merged = 0;              // the last step ran the E-phase tail as one launch
  int traces_ready = 0;

  int nff = 0;
  FarfieldState ff{};
  double *ff_buf = nullptr;
  long long ff_sizes[24] = {};
  double *phase = nullptr;
  int phase_steps = 0;

  int ndft = 0;
  double *dft_w = nullptr, *dft_acc = nullptr;

  // K-blocked NTFF (matrix-core) state: sample buffer and the current block.
  int ntff_mode = FDTD_HIP_NTFF_AUTO;
  void *ntff_buf = nullptr;
  long ntff_spad = 0;
  long ntff_base[25] = {}, ntff_size[24] = {};
  int blk_t0 = 0;          // step of row 0 of the current block
  int blk_rows = 0;        // rows gathered into the current block
  int blk_flushed = 0;     // rows already folded into the accumulators

  // Captured graph of GRAPH_STEPS steps, replayed by run_batch; its kernels take their
  // time step from *d_step (set before a run of replays, advanced by the graph itself).
  int use_graphs = 1;
  int *d_step = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_key[8] = {-1, -1, -1, -1, -1, -1, -1, -1};

  FdtdHipTuning tune{0, 0, 0, 0};  // requested launch shape / coefficient storage (0 = default)
  int coef_mode = -1;           // CoefMode in use; -1 = not decided yet
  int coef_built = 0;           // deduplicated tables built (idx width in coef_idx_bytes)
  int coef_idx_bytes = 0;       // 0 = triples do not fit 16 bits
  long long steps_run = 0;
  std::vector<double> host_coef[NCOEFF];   // uniform-layout copies kept for deduplication
  void *coef_idx = nullptr;     // 6 index arrays (uint8 or uint16), uniform layout with guards
  void *coef_tab[6] = {};
  int coef_n[6] = {};
  int timing = 0;
  int cache_resident = 0;       // field state small enough to stay in the on-die caches
  hipEvent_t ev[8] = {};
  FdtdHipKernelTimes times{};
  // CPML pass tiles of both phases for each CPML placement, in one device buffer,
  // rebuilt when a face is uploaded; cpml_tiles / ncpml name the current placement's.
  CpmlTile *tile_buf = nullptr;
  size_t tile_cap = 0;
  int faces_version = 0, tiles_version = -1;
  size_t tile_off[CPML_FUSE_XY + 1][2] = {}, tile_n[CPML_FUSE_XY + 1][2] = {};
  const CpmlTile *cpml_tiles[2] = {};
  long ncpml[2] = {};

  // Plane 0 of array c / w (one guard plane into its allocation slot).
  void *field_ptr(int c) const { return (char *)fields + c * stride_bytes + (size_t)SI * esz; }
  void *coeff_ptr(int w) const { return (char *)coeffs + w * stride_bytes + (size_t)SI * esz; }
  size_t idx_stride(int esz_idx) const { return ((size_t)(d.nx + 3) * SI * esz_idx + 255) / 256 * 256; }
  void comp_dims(int c, int out[3]) const {
    out[0] = d.nx + kExt[c][0];
    out[1] = d.ny + kExt[c][1];
    out[2] = d.nz + kExt[c][2];
  }
  long long comp_len(int c) const {
    int t[3];
    comp_dims(c, t);
    return (long long)t[0] * t[1] * t[2];
  }
  // Device offset of a reference-layout flat index of component c.
  long long to_dev(int c, long long f) const {
    int t[3];
    comp_dims(c, t);
    long long plane = (long long)t[1] * t[2];
    long long i = f / plane, r = f - i * plane, j = r / t[2], k = r - j * t[2];
    return i * SI + j * SJ + k;
  }
  // (component, i, j, k) of device offset `off` of E component `comp` (0..2).
  ECell e_cell(int comp, long long off) const {
    const int i = (int)(off / SI), r = (int)(off - (long long)i * SI), j = r / SJ;
    return {comp, i, j, r - j * SJ};
  }
  bool fp32() const { return d.precision == FDTD_HIP_FP32; }
};

namespace {

// A captured graph bakes in every pointer and launch parameter of a step: any upload,
// retuning or coefficient rebuild drops it, and the next batch captures a new one.
void drop_graph(FdtdHipContext *c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
}

template <typename T>
Grid<T> make_grid(const FdtdHipContext *c) {
  Grid<T> g{};
  g.nx = c->d.nx; g.ny = c->d.ny; g.nz = c->d.nz;
  g.SJ = c->SJ; g.SI = c->SI;
  g.Hx = (T *)c->field_ptr(0); g.Hy = (T *)c->field_ptr(1); g.Hz = (T *)c->field_ptr(2);
  g.Ex = (T *)c->field_ptr(3); g.Ey = (T *)c->field_ptr(4); g.Ez = (T *)c->field_ptr(5);
  for (int w = 0; w < NCOEFF; ++w) g.C[w] = (const T *)c->coeff_ptr(w);
  const int isz = c->coef_idx_bytes ? c->coef_idx_bytes : 1;
  for (int q = 0; q < 6; ++q) {
    g.ct.idx[q] = c->coef_idx ? (const char *)c->coef_idx + q * c->idx_stride(isz) + (size_t)c->SI * isz : nullptr;
    g.ct.tab[q] = c->coef_tab[q];
    g.ct.n[q] = c->coef_n[q];
  }
  return g;
}

template <typename T>
FaceSet<T> make_faces(const FdtdHipContext *c) {
  FaceSet<T> fs{};
  for (int f = 0; f < 6; ++f) {
    const auto &h = c->face[f];
    Face<T> &F = fs.f[f];
    F.nc = h.nc; F.m_start = h.m_start; F.e_start = h.e_start;
    const T *prof = (const T *)h.prof;
    F.b_e = prof; F.a_e = prof ? prof + h.nc : nullptr;
    F.b_m = prof ? prof + 2 * h.nc : nullptr; F.a_m = prof ? prof + 3 * h.nc : nullptr;
    for (int s = 0; s < 4; ++s) {
      F.psi[s] = (T *)h.psi[s];
      F.cpsi[s] = (const T *)h.cpsi[s];
      F.cmask[s] = h.cpsi_mask[s];
    }
  }
  return fs;
}

// Reference layout (component c, FP64 or FP32 source) -> uniform device array of T.
template <typename T, typename S>
void pack_uniform(const FdtdHipContext *c, int comp, const S *src, std::vector<T> &dst) {
  dst.assign(c->ntot, T(0));
  int t[3];
  c->comp_dims(comp, t);
  for (long long i = 0; i < t[0]; ++i)
    for (long long j = 0; j < t[1]; ++j) {
      const S *row = src + (i * t[1] + j) * t[2];
      T *out = dst.data() + i * c->SI + j * c->SJ;
      for (long long k = 0; k < t[2]; ++k) out[k] = (T)row[k];
    }
}

template <typename T>
void unpack_uniform(const FdtdHipContext *c, int comp, const std::vector<T> &src, double *dst) {
  int t[3];
  c->comp_dims(comp, t);
  for (long long i = 0; i < t[0]; ++i)
    for (long long j = 0; j < t[1]; ++j) {
      const T *row = src.data() + i * c->SI + j * c->SJ;
      double *out = dst + (i * t[1] + j) * t[2];
      for (long long k = 0; k < t[2]; ++k) out[k] = (double)row[k];
    }
}

int dev_alloc(FdtdHipContext *c, void **p, size_t bytes) {
  HIP_TRY(hipMalloc(p, bytes ? bytes : 16));
  HIP_TRY(hipMemsetAsync(*p, 0, bytes ? bytes : 16, c->stream));
  return FDTD_HIP_OK;
}

// Whether every element of v has the same bits as the first (false when v is empty).
template <typename T>
bool all_equal_bits(const std::vector<T> &v) {
  if (v.empty()) return false;
  for (const T &x : v)
    if (std::memcmp(&x, &v[0], sizeof(T)) != 0) return false;
  return true;
}

template <typename T>
int upload_vec(FdtdHipContext *c, void *dst, const std::vector<T> &v) {
  HIP_TRY(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

template <typename S>
int upload_coeff_impl(FdtdHipContext *c, int which, const S *host, long long n) {
  if (!c) return fail("upload_coeff: null context");
  drop_graph(c);
  if (which < 0 || which >= NCOEFF) return fail("upload_coeff: coefficient id %d out of range", which);
  const int comp = kCoeffOwner[which];
  if (n != c->comp_len(comp))
    return fail("upload_coeff: array %d has %lld elements, expected %lld", which, n, c->comp_len(comp));
  pack_uniform(c, comp, host, c->host_coef[which]);
  c->coef_mode = -1;
  if (c->coef_built) {                       // tables no longer match the arrays
    for (auto &t : c->coef_tab) { (void)hipFree(t); t = nullptr; }
    (void)hipFree(c->coef_idx);
    c->coef_idx = nullptr;
    c->coef_built = 0;
    c->coef_idx_bytes = 0;
    for (int &n : c->coef_n) n = 0;
  }
  if (c->fp32()) {
    std::vector<float> v(c->host_coef[which].begin(), c->host_coef[which].end());
    return upload_vec(c, c->coeff_ptr(which), v);
  }
  return upload_vec(c, c->coeff_ptr(which), c->host_coef[which]);
}

template <typename T>
int upload_small(FdtdHipContext *c, const std::vector<T> &v, T **out) {
  void *p = nullptr;
  if (dev_alloc(c, &p, v.size() * sizeof(T))) return FDTD_HIP_ERR;
  c->small_allocs.push_back(p);
  if (!v.empty() && upload_vec(c, p, v)) return FDTD_HIP_ERR;
  *out = (T *)p;
  return FDTD_HIP_OK;
}

int ensure_traces(FdtdHipContext *c) {
  if (c->traces_ready) return FDTD_HIP_OK;
  const int steps = c->d.numberOfTimeSteps;
  std::vector<double> zv((size_t)std::max(1, c->ports.nsv) * steps, 0.0);
  std::vector<double> zi((size_t)std::max(1, c->ports.nsi) * steps, 0.0);
  std::vector<double> zp((size_t)std::max(1, c->ports.npr) * steps, 0.0);
  std::vector<int> pr(c->probe_host);
  if (pr.empty()) pr.assign(PROBE_INTS, 0);
  int *prd = nullptr;
  if (upload_small(c, zv, &c->ports.vtr) || upload_small(c, zi, &c->ports.itr) ||
      upload_small(c, zp, &c->ports.prtr) || upload_small(c, pr, &prd))
    return FDTD_HIP_ERR;
  c->ports.pr = prd;
  c->ports.steps = steps;
  c->ports.dt = c->d.dt;
  // The running port DFT covers the observers uploaded by now; no more can follow.
  if (c->ndft && !c->dft_acc) {
    const int ntr = std::max(1, c->ports.nsv + c->ports.nsi);
    if (upload_small(c, std::vector<double>((size_t)2 * ntr * c->ndft, 0.0), &c->dft_acc)) return FDTD_HIP_ERR;
  }
  c->traces_ready = 1;
  return FDTD_HIP_OK;
}

// Coefficient ids of each component's (self, curl1, curl2) triple, Hx..Ez.
const int kTriple[6][3] = {{CHXH, CHXEY, CHXEZ}, {CHYH, CHYEZ, CHYEX}, {CHZH, CHZEX, CHZEY},
                           {CEXE, CEXHZ, CEXHY}, {CEYE, CEYHX, CEYHZ}, {CEZE, CEZHY, CEZHX}};

struct TripleKey {
  uint64_t a, b, c;
  bool operator==(const TripleKey &o) const { return a == o.a && b == o.b && c == o.c; }
};
struct TripleHash {
  size_t operator()(const TripleKey &k) const {
    return (size_t)(k.a * 0x9E3779B97F4A7C15ULL ^ (k.b + 0x632BE59BD9B4E019ULL) * 0xC2B2AE3D27D4EB4FULL ^ k.c);
  }
};

// Deduplicate each component's coefficient triples over its Yee shape (table order =
// first occurrence) once, into 8-bit indices when every component has at most 256
// distinct triples and 16-bit ones up to 65536; then pick the storage the tuning asks
// for (default: the indices when they exist, else the per-cell arrays).
int build_coefficient_tables(FdtdHipContext *c) {
  if (c->coef_built) return FDTD_HIP_OK;
  for (int w = 0; w < NCOEFF; ++w)
    if (c->host_coef[w].empty()) return fail("run_batch: coefficient array %d was never uploaded", w);
  std::vector<std::vector<uint32_t>> index(6);
  std::vector<std::vector<double>> table(6);
  size_t maxn = 0;
  for (int q = 0; q < 6 && maxn <= 65536; ++q) {
    int t[3];
    c->comp_dims(q, t);
    std::unordered_map<TripleKey, uint32_t, TripleHash> seen;
    index[q].assign(c->ntot, 0);
    const double *a0 = c->host_coef[kTriple[q][0]].data(), *a1 = c->host_coef[kTriple[q][1]].data(),
                 *a2 = c->host_coef[kTriple[q][2]].data();
    for (long long i = 0; i < t[0] && seen.size() <= 65536; ++i)
      for (long long j = 0; j < t[1]; ++j)
        for (long long k = 0; k < t[2]; ++k) {
          const long long o = i * c->SI + j * c->SJ + k;
          TripleKey key;
          std::memcpy(&key.a, a0 + o, 8); std::memcpy(&key.b, a1 + o, 8); std::memcpy(&key.c, a2 + o, 8);
          auto it = seen.find(key);
          uint32_t id;
          if (it == seen.end()) {
            id = (uint32_t)seen.size();
            seen.emplace(key, id);
            table[q].push_back(a0[o]); table[q].push_back(a1[o]); table[q].push_back(a2[o]);
          } else {
            id = it->second;
          }
          index[q][o] = id;
        }
    maxn = std::max(maxn, seen.size());
  }
  c->coef_built = 1;
  if (maxn > 65536) { c->coef_idx_bytes = 0; return FDTD_HIP_OK; }
  const int isz = maxn <= (size_t)U8_TABLE ? 1 : 2;
  const size_t stride = c->idx_stride(isz);
  if (dev_alloc(c, &c->coef_idx, 6 * stride)) return FDTD_HIP_ERR;
  for (int q = 0; q < 6; ++q) {
    void *dst = (char *)c->coef_idx + q * stride + (size_t)c->SI * isz;
    if (isz == 1) {
      std::vector<uint8_t> v(index[q].begin(), index[q].end());
      if (upload_vec(c, dst, v)) return FDTD_HIP_ERR;
    } else {
      std::vector<uint16_t> v(index[q].begin(), index[q].end());
      if (upload_vec(c, dst, v)) return FDTD_HIP_ERR;
    }
    c->coef_n[q] = (int)(table[q].size() / 3);
    if (c->fp32()) {
      std::vector<float> v(table[q].begin(), table[q].end());
      if (dev_alloc(c, &c->coef_tab[q], v.size() * 4) || upload_vec(c, c->coef_tab[q], v)) return FDTD_HIP_ERR;
    } else {
      if (dev_alloc(c, &c->coef_tab[q], table[q].size() * 8) || upload_vec(c, c->coef_tab[q], table[q])) return FDTD_HIP_ERR;
    }
  }
  c->coef_idx_bytes = isz;
  return FDTD_HIP_OK;
}

int finalize_coefficients(FdtdHipContext *c) {
  if (c->coef_mode >= 0) return FDTD_HIP_OK;
  if (c->tune.coefficient_storage == 1) { c->coef_mode = COEF_ARRAYS; return FDTD_HIP_OK; }
  if (build_coefficient_tables(c)) return FDTD_HIP_ERR;
  c->coef_mode = c->coef_idx_bytes == 1 ? COEF_U8 : (c->coef_idx_bytes == 2 ? COEF_U16 : COEF_ARRAYS);
  return FDTD_HIP_OK;
}

// Tiles of 256 plane offsets x `chunk` x-planes for `planes` planes.
int cells_per_thread(const FdtdHipContext *c) {
  const int w = c->tune.cells_per_thread;
  return w > 0 ? w : 1;
}

MarchGeom march_geom(const FdtdHipContext *c, int planes) {
  MarchGeom m;
  m.ncol = c->SI / (256 * cells_per_thread(c));
  m.chunk = c->tune.planes_per_workgroup > 0 ? c->tune.planes_per_workgroup : 2;
  m.nchunk = (planes + m.chunk - 1) / m.chunk;
  m.per_xcd = (m.ncol * m.nchunk + 7) / 8;
  return m;
}

int fuse_cpml(const FdtdHipContext *c) { return c->tune.fuse_cpml; }

// Four cells per thread exist for FP32 only (check_tuning refuses them for FP64).
// NT: streaming loads and stores of the stencils' own fields (see stv_stream).
template <typename T, int CM, int FC, bool NT>
void launch_h_wn(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  const dim3 grid(8 * m.per_xcd), block(256);
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) { hipLaunchKernelGGL((k_update_h<T, CM, 4, FC, NT>), grid, block, 0, c->stream, g, m, fs); return; }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_h<T, CM, 2, FC, NT>), grid, block, 0, c->stream, g, m, fs);
  else hipLaunchKernelGGL((k_update_h<T, CM, 1, FC, NT>), grid, block, 0, c->stream, g, m, fs);
}

template <typename T, int CM, int FC, bool NT>
void launch_e_wn(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  const dim3 grid(8 * m.per_xcd), block(256);
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) {
      hipLaunchKernelGGL((k_update_e<T, CM, 4, FC, NT>), grid, block, 0, c->stream, g, m, ghost, fs);
      return;
    }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_e<T, CM, 2, FC, NT>), grid, block, 0, c->stream, g, m, ghost, fs);
  else hipLaunchKernelGGL((k_update_e<T, CM, 1, FC, NT>), grid, block, 0, c->stream, g, m, ghost, fs);
}

// Plain (cached) accesses only with the separate CPML pass and with the x/y faces
// fused, the placements small grids use; the others always stream.
template <typename T, int CM, int FC>
void launch_h_w(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  if constexpr (FC == CPML_SEPARATE || FC == CPML_FUSE_XY) {
    if (c->cache_resident) { launch_h_wn<T, CM, FC, false>(c, g, m, fs); return; }
  }
  launch_h_wn<T, CM, FC, true>(c, g, m, fs);
}

template <typename T, int CM, int FC>
void launch_e_w(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  if constexpr (FC == CPML_SEPARATE || FC == CPML_FUSE_XY) {
    if (c->cache_resident) { launch_e_wn<T, CM, FC, false>(c, g, m, ghost, fs); return; }
  }
  launch_e_wn<T, CM, FC, true>(c, g, m, ghost, fs);
}

template <typename T, int FC>
void launch_h_f(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  if (c->coef_mode == COEF_U8) launch_h_w<T, COEF_U8, FC>(c, g, m, fs);
  else if (c->coef_mode == COEF_U16) launch_h_w<T, COEF_U16, FC>(c, g, m, fs);
  else launch_h_w<T, COEF_ARRAYS, FC>(c, g, m, fs);
}

template <typename T, int FC>
void launch_e_f(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  if (c->coef_mode == COEF_U8) launch_e_w<T, COEF_U8, FC>(c, g, m, ghost, fs);
  else if (c->coef_mode == COEF_U16) launch_e_w<T, COEF_U16, FC>(c, g, m, ghost, fs);
  else launch_e_w<T, COEF_ARRAYS, FC>(c, g, m, ghost, fs);
}

// The CPML faces of a context as the tile builder takes them.
CpmlLayout cpml_layout(const FdtdHipContext *c) {
  CpmlLayout L{};
  L.n[0] = c->d.nx; L.n[1] = c->d.ny; L.n[2] = c->d.nz;
  for (int f = 0; f < 6; ++f) {
    L.nc[f] = c->face[f].nc;
    L.m_start[f] = c->face[f].m_start;
    L.e_start[f] = c->face[f].e_start;
  }
  return L;
}

// CPML_SPLIT_Z launches: the z-slab march workgroups first, then the stencil tiles.
// The z-slab march of H covers planes 0..nx, that of E planes 0..nx-1 (the E wall
// rows come with plane 0).
template <typename T, int CM>
void launch_h_split(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  const ZslabGeom z = zslab_geom(cpml_layout(c), 0, c->d.nx + 1, m.chunk);
  const dim3 grid(z.ntiles + 8 * m.per_xcd), block(256);
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) { hipLaunchKernelGGL((k_update_h_split<T, CM, 4>), grid, block, 0, c->stream, g, m, z, fs); return; }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_h_split<T, CM, 2>), grid, block, 0, c->stream, g, m, z, fs);
  else hipLaunchKernelGGL((k_update_h_split<T, CM, 1>), grid, block, 0, c->stream, g, m, z, fs);
}

template <typename T, int CM>
void launch_e_split(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  const ZslabGeom z = zslab_geom(cpml_layout(c), 1, c->d.nx, m.chunk);
  const dim3 grid(z.ntiles + 8 * m.per_xcd), block(256);
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) {
      hipLaunchKernelGGL((k_update_e_split<T, CM, 4>), grid, block, 0, c->stream, g, m, z, ghost, fs);
      return;
    }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_e_split<T, CM, 2>), grid, block, 0, c->stream, g, m, z, ghost, fs);
  else hipLaunchKernelGGL((k_update_e_split<T, CM, 1>), grid, block, 0, c->stream, g, m, z, ghost, fs);
}

// CPML_SPLIT_ALL launches: the CPML tiles (padded to a multiple of 8 workgroups), then
// the stencil tiles.
template <typename T, int CM>
void launch_h_all(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  const int nt = (int)c->ncpml[0];
  const dim3 grid(((nt + 7) & ~7) + 8 * m.per_xcd), block(256);
  const CpmlTile *t = c->cpml_tiles[0];
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) { hipLaunchKernelGGL((k_update_h_all<T, CM, 4>), grid, block, 0, c->stream, g, m, fs, t, nt); return; }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_h_all<T, CM, 2>), grid, block, 0, c->stream, g, m, fs, t, nt);
  else hipLaunchKernelGGL((k_update_h_all<T, CM, 1>), grid, block, 0, c->stream, g, m, fs, t, nt);
}

template <typename T, int CM>
void launch_e_all(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  const int nt = (int)c->ncpml[1];
  const dim3 grid(((nt + 7) & ~7) + 8 * m.per_xcd), block(256);
  const CpmlTile *t = c->cpml_tiles[1];
  const int w = cells_per_thread(c);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) {
      hipLaunchKernelGGL((k_update_e_all<T, CM, 4>), grid, block, 0, c->stream, g, m, ghost, fs, t, nt);
      return;
    }
  }
  if (w == 2) hipLaunchKernelGGL((k_update_e_all<T, CM, 2>), grid, block, 0, c->stream, g, m, ghost, fs, t, nt);
  else hipLaunchKernelGGL((k_update_e_all<T, CM, 1>), grid, block, 0, c->stream, g, m, ghost, fs, t, nt);
}

template <typename T>
void launch_h(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, const FaceSet<T> &fs) {
  switch (fuse_cpml(c)) {
    case CPML_FUSE_XY: launch_h_f<T, CPML_FUSE_XY>(c, g, m, fs); break;
    case CPML_SPLIT_ALL:
      if (c->coef_mode == COEF_U8) launch_h_all<T, COEF_U8>(c, g, m, fs);
      else if (c->coef_mode == COEF_U16) launch_h_all<T, COEF_U16>(c, g, m, fs);
      else launch_h_all<T, COEF_ARRAYS>(c, g, m, fs);
      break;
    case CPML_FUSE_ALL: launch_h_f<T, CPML_FUSE_ALL>(c, g, m, fs); break;
    case CPML_FUSE_Z: launch_h_f<T, CPML_FUSE_Z>(c, g, m, fs); break;
    case CPML_SPLIT_Z:
      if (c->coef_mode == COEF_U8) launch_h_split<T, COEF_U8>(c, g, m, fs);
      else if (c->coef_mode == COEF_U16) launch_h_split<T, COEF_U16>(c, g, m, fs);
      else launch_h_split<T, COEF_ARRAYS>(c, g, m, fs);
      break;
    default: launch_h_f<T, CPML_SEPARATE>(c, g, m, fs); break;
  }
}

template <typename T>
void launch_e(FdtdHipContext *c, const Grid<T> &g, const MarchGeom &m, int ghost, const FaceSet<T> &fs) {
  switch (fuse_cpml(c)) {
    case CPML_FUSE_XY: launch_e_f<T, CPML_FUSE_XY>(c, g, m, ghost, fs); break;
    case CPML_SPLIT_ALL:
      if (c->coef_mode == COEF_U8) launch_e_all<T, COEF_U8>(c, g, m, ghost, fs);
      else if (c->coef_mode == COEF_U16) launch_e_all<T, COEF_U16>(c, g, m, ghost, fs);
      else launch_e_all<T, COEF_ARRAYS>(c, g, m, ghost, fs);
      break;
    case CPML_FUSE_ALL: launch_e_f<T, CPML_FUSE_ALL>(c, g, m, ghost, fs); break;
    case CPML_FUSE_Z: launch_e_f<T, CPML_FUSE_Z>(c, g, m, ghost, fs); break;
    case CPML_SPLIT_Z:
      if (c->coef_mode == COEF_U8) launch_e_split<T, COEF_U8>(c, g, m, ghost, fs);
      else if (c->coef_mode == COEF_U16) launch_e_split<T, COEF_U16>(c, g, m, ghost, fs);
      else launch_e_split<T, COEF_ARRAYS>(c, g, m, ghost, fs);
      break;
    default: launch_e_f<T, CPML_SEPARATE>(c, g, m, ghost, fs); break;
  }
}

// Tiles of both phases for the current placement, building every placement's lists
// when the faces changed.  Tile records hold coordinates in 16 bits.
int set_cpml_tiles(FdtdHipContext *c) {
  if (c->tiles_version != c->faces_version) {
    if (std::max({c->d.nx, c->d.ny, c->d.nz}) > CPML_TILE_MAX_AXIS)
      return fail("CPML pass: grid axes of more than %d cells are not supported", CPML_TILE_MAX_AXIS);
    const CpmlLayout L = cpml_layout(c);
    std::vector<CpmlTile> all;
    for (int mode : {CPML_SEPARATE, CPML_FUSE_Z, CPML_FUSE_ALL, CPML_FUSE_XY})
      for (int ph = 0; ph < 2; ++ph) {
        c->tile_off[mode][ph] = all.size();
        build_cpml_tiles(L, ph, mode, all);
        c->tile_n[mode][ph] = all.size() - c->tile_off[mode][ph];
      }
    if (all.size() > c->tile_cap) {
      (void)hipFree(c->tile_buf);
      c->tile_buf = nullptr;
      c->tile_cap = 0;
      HIP_TRY(hipMalloc((void **)&c->tile_buf, all.size() * sizeof(CpmlTile)));
      c->tile_cap = all.size();
    }
    if (!all.empty() && upload_vec(c, c->tile_buf, all)) return FDTD_HIP_ERR;
    c->tiles_version = c->faces_version;
  }
  const int mode = cpml_tile_mode(fuse_cpml(c));
  for (int ph = 0; ph < 2; ++ph) {
    c->cpml_tiles[ph] = c->tile_buf ? c->tile_buf + c->tile_off[mode][ph] : nullptr;
    c->ncpml[ph] = (long)c->tile_n[mode][ph];
  }
  return FDTD_HIP_OK;
}

// Whether the step's E-phase tail may run as one launch (k_e_tail).  The NTFF reads
// E only on the surface of its box [LI,UI] x [LJ,UJ] x [LK,UK] (0-based planes).
bool e_tail_mergeable(const FdtdHipContext *c) {
  const FarfieldState &f = c->ff;
  const int lo[3] = {f.LI, f.LJ, f.LK}, hi[3] = {f.UI, f.UJ, f.UK};
  return e_tail_disjoint(cpml_layout(c), c->tail_e_write, c->tail_e_read, c->nff > 0, lo, hi);
}

bool ntff_blocked(const FdtdHipContext *c) {
  if (c->nff == 0) return false;
  return c->ntff_mode == FDTD_HIP_NTFF_BLOCKED || (c->ntff_mode == FDTD_HIP_NTFF_AUTO && c->fp32());
}

// Sample ranges of the 24 accumulators in the blocked buffer, each padded to a
// multiple of 16 so a matrix-core tile never spans two accumulators.
int ensure_ntff_buffer(FdtdHipContext *c) {
  if (c->ntff_buf) return FDTD_HIP_OK;
  const FarfieldState &f = c->ff;
  const long sz[3] = {(long)f.nj * f.nk, (long)f.ni * f.nk, (long)f.ni * f.nj};
  long pos = 0;
  for (int a = 0; a < 24; ++a) {
    c->ntff_base[a] = pos;
    c->ntff_size[a] = sz[a / 8];
    pos += (sz[a / 8] + 15) / 16 * 16;
  }
  c->ntff_base[24] = pos;
  c->ntff_spad = pos;
  c->blk_t0 = c->blk_rows = c->blk_flushed = 0;
  return dev_alloc(c, &c->ntff_buf, (size_t)NTFF_K * pos * c->esz);
}

template <typename T>
NtffBlock<T> ntff_block(const FdtdHipContext *c) {
  NtffBlock<T> nb{};
  nb.buf = (T *)c->ntff_buf;
  nb.spad = c->ntff_spad;
  for (int a = 0; a < 25; ++a) nb.base[a] = c->ntff_base[a];
  for (int a = 0; a < 24; ++a) nb.size[a] = c->ntff_size[a];
  return nb;
}

// Fold rows [r0, r1) of the block whose first step is step_of(tsp, ts0).
template <typename T>
void ntff_fold(FdtdHipContext *c, const int *tsp, int ts0, int r0, int r1) {
  const NtffBlock<T> nb = ntff_block<T>(c);
  // Persistent grid: a few workgroups per CU, each looping over sample tiles.
  const unsigned grid = (unsigned)std::min<long>((c->ntff_spad / 16 + 3) / 4, 2048);
  for (int f0 = 0; f0 < c->nff; f0 += NTFF_FREQ_GROUP) {
    if (2 * (c->nff - f0) <= 16)
      hipLaunchKernelGGL((k_ntff_mfma<T, 1>), dim3(grid), dim3(256), 0, c->stream, c->ff, nb, tsp, ts0, r0, r1, f0);
    else
      hipLaunchKernelGGL((k_ntff_mfma<T, 2>), dim3(grid), dim3(256), 0, c->stream, c->ff, nb, tsp, ts0, r0, r1, f0);
  }
}

// Fold rows [blk_flushed, blk_rows) of the current block into the accumulators.
template <typename T>
int ntff_flush(FdtdHipContext *c) {
  if (c->blk_rows <= c->blk_flushed) return FDTD_HIP_OK;
  ntff_fold<T>(c, nullptr, c->blk_t0, c->blk_flushed, c->blk_rows);
  c->blk_flushed = c->blk_rows;
  if (c->blk_rows == NTFF_K) c->blk_t0 += NTFF_K, c->blk_rows = c->blk_flushed = 0;
  return FDTD_HIP_OK;
}

// The launches of one time step.  Direct: tsp == nullptr and ts is the step.  Under
// graph capture: ts is the step's offset within the graph and *tsp the graph's first
// step; graphs start on a 64-step NTFF block, so offset = block row.
template <typename T>
int launch_step(FdtdHipContext *c, int ts, const int *tsp = nullptr) {
  const Grid<T> g = make_grid<T>(c);
  const FaceSet<T> fs = make_faces<T>(c);
  const dim3 block(256);
  const int ghost = c->d.boundary == FDTD_HIP_BOUNDARY_REFERENCE;
  const MarchGeom mh = march_geom(c, c->d.nx + 1), me = march_geom(c, ghost ? c->d.nx : c->d.nx + 1);
  const bool t = c->timing;
  // CPML tiles of the separate pass (none when the bulk launches hold them).
  const bool pass = cpml_pass_separate(fuse_cpml(c));
  const long pass_h = pass ? c->ncpml[0] : 0, pass_e = pass ? c->ncpml[1] : 0;

  if (t) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  launch_h<T>(c, g, mh, fs);
  if (t) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  if (pass_h)
    hipLaunchKernelGGL(k_cpml_h<T>, dim3((unsigned)pass_h), block, 0, c->stream, g, fs, c->cpml_tiles[0]);
  if (t) HIP_TRY(hipEventRecord(c->ev[2], c->stream));
  launch_e<T>(c, g, me, ghost, fs);
  if (t) HIP_TRY(hipEventRecord(c->ev[3], c->stream));
  // The E-phase tail: E CPML, port epilogue (source injection, captures, running port
  // DFT), NTFF.  One launch when e_tail_mergeable holds and per-kernel timing is off.
  const bool ports = c->ports.nvs + c->ports.nsv + c->ports.nsi + c->ports.npr > 0;
  PortState ps = c->ports;
  ps.ndft = c->ndft; ps.w = c->dft_w; ps.dft = c->dft_acc;
  const long total =
      c->nff > 0 ? 2 * ((long)c->ff.nj * c->ff.nk + (long)c->ff.ni * c->ff.nk + (long)c->ff.ni * c->ff.nj) : 0;
  const int nkind = c->nff > 0 ? (ntff_blocked(c) ? 2 : 1) : 0;
  int row = ts;
  if (nkind == 2 && !tsp) {
    // Steps must follow on from the block; a jump (a new run) starts a new block.
    if (ts != c->blk_t0 + c->blk_rows) {
      if (ntff_flush<T>(c)) return FDTD_HIP_ERR;
      c->blk_t0 = ts;
      c->blk_rows = c->blk_flushed = 0;
    }
    row = c->blk_rows;
  }
  const NtffBlock<T> nb = nkind == 2 ? ntff_block<T>(c) : NtffBlock<T>{};
  c->e_tail_merged = !t && c->e_tail_ok;
  if (c->e_tail_merged) {
    const int nf = (int)((total + 255) / 256), pp = ports ? 1 : 0;
    const long nblk = pass_e + pp + nf;
    if (nblk > 0) {
      const dim3 grid((unsigned)nblk);
      if (nkind == 2)
        hipLaunchKernelGGL((k_e_tail<T, 2>), grid, block, 0, c->stream, g, fs, c->cpml_tiles[1], pp, nf, ps, c->ff, nb, row, tsp, ts);
      else if (nkind == 1)
        hipLaunchKernelGGL((k_e_tail<T, 1>), grid, block, 0, c->stream, g, fs, c->cpml_tiles[1], pp, nf, ps, c->ff, nb, row, tsp, ts);
      else
        hipLaunchKernelGGL((k_e_tail<T, 0>), grid, block, 0, c->stream, g, fs, c->cpml_tiles[1], pp, nf, ps, c->ff, nb, row, tsp, ts);
    }
  } else {
    if (pass_e)
      hipLaunchKernelGGL(k_cpml_e<T>, dim3((unsigned)pass_e), block, 0, c->stream, g, fs, c->cpml_tiles[1]);
    if (t) HIP_TRY(hipEventRecord(c->ev[4], c->stream));
    if (ports) hipLaunchKernelGGL(k_port_epilogue<T>, dim3(1), block, 0, c->stream, g, ps, tsp, ts);
    if (t) HIP_TRY(hipEventRecord(c->ev[5], c->stream));
    if (nkind == 2)
      hipLaunchKernelGGL(k_ntff_gather<T>, dim3((unsigned)((total + 255) / 256)), block, 0, c->stream, g, c->ff, nb, row);
    else if (nkind == 1)
      hipLaunchKernelGGL(k_ntff_direct<T>, dim3((unsigned)((total + 255) / 256)), block, 0, c->stream, g, c->ff, tsp, ts);
  }
  if (nkind == 2) {
    if (tsp) {
      if (ts == NTFF_K - 1) ntff_fold<T>(c, tsp, 0, 0, NTFF_K);
    } else {
      ++c->blk_rows;
      if (c->blk_rows == NTFF_K && ntff_flush<T>(c)) return FDTD_HIP_ERR;
    }
  }
  if (t) HIP_TRY(hipEventRecord(c->ev[6], c->stream));
  HIP_TRY(hipGetLastError());
  if (t) {
    // events 0..6 bracket: H bulk, H CPML, E bulk, E CPML, port epilogue, NTFF
    static const int slot[6] = {0, 4, 1, 5, 2, 3};
    HIP_TRY(hipEventSynchronize(c->ev[6]));
    for (int q = 0; q < 6; ++q) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, c->ev[q], c->ev[q + 1]));
      c->times.ms[slot[q]] += ms;
      c->times.launches[slot[q]] += 1;
    }
  }
  return FDTD_HIP_OK;
}

constexpr int GRAPH_STEPS = NTFF_K;

// Everything a captured step depends on besides the device data itself.
void graph_key(const FdtdHipContext *c, int key[8]) {
  key[0] = c->tune.planes_per_workgroup; key[1] = c->tune.cells_per_thread;
  key[2] = c->tune.coefficient_storage; key[3] = c->coef_mode;
  key[4] = ntff_blocked(c) ? 1 : 0; key[5] = c->ndft; key[6] = c->tune.fuse_cpml;
  key[7] = c->e_tail_ok;
}

// Capture GRAPH_STEPS steps plus the step advance into an executable graph.
template <typename T>
int build_graph(FdtdHipContext *c) {
  drop_graph(c);
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = FDTD_HIP_OK;
  for (int k = 0; k < GRAPH_STEPS && !rc; ++k) rc = launch_step<T>(c, k, c->d_step);
  if (!rc) hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, c->stream, c->d_step, GRAPH_STEPS);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(c->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (end != hipSuccess) return fail("graph capture failed: %s", hipGetErrorString(end));
  const hipError_t inst = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (inst != hipSuccess) { c->graph = nullptr; return fail("graph instantiation failed: %s", hipGetErrorString(inst)); }
  graph_key(c, c->graph_key);
  return FDTD_HIP_OK;
}

// Steps [ts, ts + count): whole 64-step graphs where the NTFF block state allows,
// direct launches for the rest.
template <typename T>
int run_steps(FdtdHipContext *c, int ts, int count) {
  bool base_set = false;
  for (int s = 0; s < count;) {
    const int t = ts + s;
    const bool aligned = !ntff_blocked(c) || (c->blk_rows == 0 && c->blk_t0 == t);
    if (c->use_graphs && !c->timing && count - s >= GRAPH_STEPS && aligned) {
      int key[8];
      graph_key(c, key);
      if (!c->graph || std::memcmp(key, c->graph_key, sizeof key) != 0) {
        if (build_graph<T>(c)) return FDTD_HIP_ERR;
        base_set = false;
      }
      if (!base_set) hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, c->stream, c->d_step, t);
      base_set = true;
      HIP_TRY(hipGraphLaunch(c->graph, c->stream));
      if (ntff_blocked(c)) c->blk_t0 = t + GRAPH_STEPS;
      s += GRAPH_STEPS;
    } else {
      if (launch_step<T>(c, t)) return FDTD_HIP_ERR;
      base_set = false;
      ++s;
    }
  }
  return FDTD_HIP_OK;
}

}  // namespace

// =================================================================== C ABI ====

extern "C" {

const char *fdtd_hip_last_error(void) { return g_last_error.c_str(); }

int fdtd_hip_device_count(int *count) {
  if (!count) return fail("device_count: null pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  int usable = 0;
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, i) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      ++usable;
  }
  *count = usable;
  return FDTD_HIP_OK;
}

int fdtd_hip_init(const FdtdHipDims *dims, FdtdHipContext **out) {
  if (!dims || !out) return fail("init: null argument");
  *out = nullptr;
  const FdtdHipDims &d = *dims;
  if (d.nx < 2 || d.ny < 2 || d.nz < 2) return fail("init: grid %dx%dx%d too small", d.nx, d.ny, d.nz);
  if (d.precision != FDTD_HIP_FP64 && d.precision != FDTD_HIP_FP32) return fail("init: unknown precision %d", d.precision);
  if (d.boundary != FDTD_HIP_BOUNDARY_REFERENCE && d.boundary != FDTD_HIP_BOUNDARY_MATLAB)
    return fail("init: unknown boundary mode %d", d.boundary);
  if (d.numberOfTimeSteps < 1) return fail("init: numberOfTimeSteps must be positive");
  const long long ntot = (long long)(d.nx + 3) * ((long long)(d.ny + 1) * (d.nz + 4) + 1023);
  if (ntot >= (1LL << 31) - 1) return fail("init: %lld cells exceed the 32-bit device indexing", ntot);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail("init: no HIP device");

  auto *c = new FdtdHipContext();
  c->d = d;
  c->SJ = (d.nz + 1 + 3) / 4 * 4;
  c->SI = ((d.ny + 1) * c->SJ + 1023) / 1024 * 1024;
  c->ntot = (long long)(d.nx + 1) * c->SI;
  c->esz = c->fp32() ? 4 : 8;
  // Each uniform array: guard plane, nx+1 planes, guard plane.
  c->stride_bytes = ((size_t)(d.nx + 3) * c->SI * c->esz + 255) / 256 * 256;
  c->cache_resident = 6 * c->stride_bytes <= CACHE_RESIDENT_FIELD_BYTES;
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  int rc = noalloc_register(c->device);
  if (!rc && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = fail("init: stream creation failed");
  if (!rc) rc = dev_alloc(c, &c->fields, 6 * c->stride_bytes);
  if (!rc) rc = dev_alloc(c, &c->coeffs, NCOEFF * c->stride_bytes);
  if (!rc) rc = dev_alloc(c, (void **)&c->d_step, 64);
  for (int q = 0; !rc && q < 8; ++q)
    if (hipEventCreate(&c->ev[q]) != hipSuccess) rc = fail("init: event creation failed");
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail("init: synchronize failed");
  if (rc) {
    fdtd_hip_destroy(c);
    return rc;
  }
  *out = c;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_coeff(FdtdHipContext *c, int which, const double *host, long long n) {
  if (!host) return fail("upload_coeff: null array");
  return upload_coeff_impl(c, which, host, n);
}

int fdtd_hip_upload_coeff_f32(FdtdHipContext *c, int which, const float *host, long long n) {
  if (!host) return fail("upload_coeff_f32: null array");
  if (c && !c->fp32()) return fail("upload_coeff_f32: context is FP64; FP32 input would lose the reference bits");
  return upload_coeff_impl(c, which, host, n);
}

int fdtd_hip_upload_cpml(FdtdHipContext *c, int faceId, int nc, const double *b_e, const double *a_e,
                         const double *b_m, const double *a_m, const double *CPsi0, long long n0,
                         const double *CPsi1, long long n1, const double *CPsi2, long long n2,
                         const double *CPsi3, long long n3, int m_start, int e_start) {
  if (!c) return fail("upload_cpml: null context");
  drop_graph(c);
  if (faceId < 0 || faceId > 5) return fail("upload_cpml: face id %d out of range", faceId);
  const int axis = faceId / 2;
  const int n = axis == 0 ? c->d.nx : (axis == 1 ? c->d.ny : c->d.nz);
  if (nc <= 0 || 2 * nc >= n) return fail("upload_cpml: nc=%d invalid for an axis of %d cells", nc, n);
  if (m_start < 0 || m_start + nc > n) return fail("upload_cpml: H planes [%d,%d) outside [0,%d)", m_start, m_start + nc, n);
  if (e_start < 1 || e_start + nc > n) return fail("upload_cpml: E planes [%d,%d) outside [1,%d)", e_start, e_start + nc, n);
  if (!b_e || !a_e || !b_m || !a_m) return fail("upload_cpml: null profile");
  auto &F = c->face[faceId];
  if (F.nc) return fail("upload_cpml: face %d already uploaded", faceId);
  const double *cp[4] = {CPsi0, CPsi1, CPsi2, CPsi3};
  const long long nn[4] = {n0, n1, n2, n3};
  for (int s = 0; s < 4; ++s) {
    int t[3];
    c->comp_dims(kSlotComp[axis][s], t);
    t[axis] = nc;
    const long long want = (long long)t[0] * t[1] * t[2];
    if (!cp[s] || nn[s] != want) return fail("upload_cpml: face %d slot %d has %lld elements, expected %lld", faceId, s, nn[s], want);
  }
  F.m_start = m_start;
  F.e_start = e_start;
  const double *pr[4] = {b_e, a_e, b_m, a_m};
  if (c->fp32()) {
    std::vector<float> prof(4 * nc);
    for (int q = 0; q < 4; ++q) for (int ci = 0; ci < nc; ++ci) prof[q * nc + ci] = (float)pr[q][ci];
    float *p = nullptr;
    if (upload_small(c, prof, &p)) return FDTD_HIP_ERR;
    F.prof = p;
  } else {
    std::vector<double> prof(4 * nc);
    for (int q = 0; q < 4; ++q) for (int ci = 0; ci < nc; ++ci) prof[q * nc + ci] = pr[q][ci];
    double *p = nullptr;
    if (upload_small(c, prof, &p)) return FDTD_HIP_ERR;
    F.prof = p;
  }
  for (int s = 0; s < 4; ++s) {
    F.n[s] = nn[s];
    if (dev_alloc(c, &F.psi[s], nn[s] * c->esz) || dev_alloc(c, &F.cpsi[s], nn[s] * c->esz)) return FDTD_HIP_ERR;
    // The whole array is uploaded either way; a slot whose elements are all equal (bit
    // for bit, in the context's precision) is read at element 0 only.
    if (c->fp32()) {
      std::vector<float> v(cp[s], cp[s] + nn[s]);
      F.cpsi_mask[s] = all_equal_bits(v) ? 0 : -1;
      if (upload_vec(c, F.cpsi[s], v)) return FDTD_HIP_ERR;
    } else {
      std::vector<double> v(cp[s], cp[s] + nn[s]);
      F.cpsi_mask[s] = all_equal_bits(v) ? 0 : -1;
      if (upload_vec(c, F.cpsi[s], v)) return FDTD_HIP_ERR;
    }
  }
  F.nc = nc;
  ++c->faces_version;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_vsource(FdtdHipContext *c, int dir, const int *idx, const double *coef, int count,
                            const double *waveform, int nSteps) {
  if (!c) return fail("upload_vsource: null context");
  drop_graph(c);
  if (dir < 0 || dir > 2) return fail("upload_vsource: direction %d out of range", dir);
  if (c->ports.nvs >= MAX_PORT_OBJS) return fail("upload_vsource: at most %d sources", MAX_PORT_OBJS);
  if (count < 0 || (count > 0 && (!idx || !coef))) return fail("upload_vsource: bad edge list");
  if (!waveform || nSteps < c->d.numberOfTimeSteps)
    return fail("upload_vsource: waveform has %d samples, need %d", nSteps, c->d.numberOfTimeSteps);
  const int comp = 3 + dir;
  std::vector<int> di(count);
  for (int n = 0; n < count; ++n) {
    if (idx[n] < 0 || idx[n] >= c->comp_len(comp)) return fail("upload_vsource: edge index %d outside E%c", idx[n], "xyz"[dir]);
    for (int m = 0; m < n; ++m)
      if (idx[m] == idx[n]) return fail("upload_vsource: duplicate edge index %d", idx[n]);
    di[n] = (int)c->to_dev(comp, idx[n]);
  }
  for (int n = 0; n < count; ++n) c->tail_e_write.push_back(c->e_cell(dir, di[n]));
  VSourceDev &v = c->ports.vs[c->ports.nvs];
  v.dir = dir; v.count = count;
  int *pi = nullptr; double *pc = nullptr, *pw = nullptr;
  if (upload_small(c, di, &pi)) return FDTD_HIP_ERR;
  if (upload_small(c, std::vector<double>(coef, coef + count), &pc)) return FDTD_HIP_ERR;
  if (upload_small(c, std::vector<double>(waveform, waveform + c->d.numberOfTimeSteps), &pw)) return FDTD_HIP_ERR;
  v.idx = pi; v.coef = pc; v.wave = pw;
  c->ports.nvs++;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_svoltage(FdtdHipContext *c, int dir, const int *idx, int count, double Csvf) {
  if (!c) return fail("upload_svoltage: null context");
  drop_graph(c);
  if (dir < 0 || dir > 2) return fail("upload_svoltage: direction %d out of range", dir);
  if (c->ports.nsv >= MAX_PORT_OBJS) return fail("upload_svoltage: at most %d observers", MAX_PORT_OBJS);
  if (c->traces_ready) return fail("upload_svoltage: observers must be uploaded before the first step");
  if (count < 0 || (count > 0 && !idx)) return fail("upload_svoltage: bad edge list");
  const int comp = 3 + dir;
  std::vector<int> di(count);
  for (int n = 0; n < count; ++n) {
    if (idx[n] < 0 || idx[n] >= c->comp_len(comp)) return fail("upload_svoltage: edge index %d outside E%c", idx[n], "xyz"[dir]);
    di[n] = (int)c->to_dev(comp, idx[n]);
  }
  for (int n = 0; n < count; ++n) c->tail_e_read.push_back(c->e_cell(dir, di[n]));
  SVoltageDev &s = c->ports.sv[c->ports.nsv];
  int *pi = nullptr;
  if (upload_small(c, di, &pi)) return FDTD_HIP_ERR;
  s.dir = dir; s.count = count; s.idx = pi; s.csvf = Csvf;
  c->ports.nsv++;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_scurrent(FdtdHipContext *c, const int *comp, const int *index, const double *weight,
                             int count, int negate) {
  if (!c) return fail("upload_scurrent: null context");
  drop_graph(c);
  if (c->ports.nsi >= MAX_PORT_OBJS) return fail("upload_scurrent: at most %d observers", MAX_PORT_OBJS);
  if (c->traces_ready) return fail("upload_scurrent: observers must be uploaded before the first step");
  if (count < 0 || (count > 0 && (!comp || !index || !weight))) return fail("upload_scurrent: bad term list");
  std::vector<int> dc(count), di(count);
  for (int n = 0; n < count; ++n) {
    if (comp[n] < 0 || comp[n] > 2) return fail("upload_scurrent: term %d names H component %d", n, comp[n]);
    if (index[n] < 0 || index[n] >= c->comp_len(comp[n])) return fail("upload_scurrent: term %d index %d out of range", n, index[n]);
    dc[n] = comp[n];
    di[n] = (int)c->to_dev(comp[n], index[n]);
  }
  SCurrentDev &s = c->ports.si[c->ports.nsi];
  int *pc = nullptr, *pi = nullptr; double *pw = nullptr;
  if (upload_small(c, dc, &pc) || upload_small(c, di, &pi) ||
      upload_small(c, std::vector<double>(weight, weight + count), &pw)) return FDTD_HIP_ERR;
  s.count = count; s.negate = negate ? 1 : 0; s.comp = pc; s.idx = pi; s.w = pw;
  c->ports.nsi++;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_farfield(FdtdHipContext *c, int nFreq, const int *box, const double *phase, int nSteps) {
  if (!c) return fail("upload_farfield: null context");
  drop_graph(c);
  if (c->nff) return fail("upload_farfield: already uploaded");
  if (nFreq <= 0 || !box || !phase) return fail("upload_farfield: bad arguments");
  if (nSteps < c->d.numberOfTimeSteps) return fail("upload_farfield: phase table has %d steps, need %d", nSteps, c->d.numberOfTimeSteps);
  const int li = box[0], lj = box[1], lk = box[2], ui = box[3], uj = box[4], uk = box[5];
  // Reads reach planes LI-1..UI (+1 along the face) on every axis.
  if (li < 2 || lj < 2 || lk < 2 || ui > c->d.nx || uj > c->d.ny || uk > c->d.nz || ui <= li || uj <= lj || uk <= lk)
    return fail("upload_farfield: box [%d,%d]x[%d,%d]x[%d,%d] does not fit the %dx%dx%d grid",
                li, ui, lj, uj, lk, uk, c->d.nx, c->d.ny, c->d.nz);
  FarfieldState &f = c->ff;
  f.nfreq = nFreq;
  f.LI = li - 1; f.LJ = lj - 1; f.LK = lk - 1; f.UI = ui - 1; f.UJ = uj - 1; f.UK = uk - 1;
  f.ni = ui - li; f.nj = uj - lj; f.nk = uk - lk;
  long long sz[3] = {(long long)f.nj * f.nk, (long long)f.ni * f.nk, (long long)f.ni * f.nj};
  long long total = 0;
  for (int a = 0; a < 24; ++a) { c->ff_sizes[a] = sz[a / 8] * nFreq; total += 2 * c->ff_sizes[a]; }
  void *p = nullptr;
  if (dev_alloc(c, &p, total * sizeof(double))) return FDTD_HIP_ERR;
  c->ff_buf = (double *)p;
  f.acc = c->ff_buf;
  double *ph = nullptr;
  if (upload_small(c, std::vector<double>(phase, phase + (size_t)c->d.numberOfTimeSteps * nFreq * 4), &ph))
    return FDTD_HIP_ERR;
  f.phase = ph;
  c->nff = nFreq;
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_port_dft(FdtdHipContext *c, int nFreq, const double *freqHz) {
  if (!c) return fail("upload_port_dft: null context");
  drop_graph(c);
  if (c->ndft) return fail("upload_port_dft: already uploaded");
  if (nFreq <= 0 || !freqHz) return fail("upload_port_dft: bad arguments");
  std::vector<double> w(nFreq);
  for (int f = 0; f < nFreq; ++f) w[f] = 2 * M_PI * freqHz[f];
  if (upload_small(c, w, &c->dft_w)) return FDTD_HIP_ERR;
  c->ndft = nFreq;
  // Accumulators for the observers present at the first step (ensure_traces).
  if (c->traces_ready) {
    const int ntr = std::max(1, c->ports.nsv + c->ports.nsi);
    if (upload_small(c, std::vector<double>((size_t)2 * ntr * nFreq, 0.0), &c->dft_acc)) return FDTD_HIP_ERR;
  }
  return FDTD_HIP_OK;
}

int fdtd_hip_get_counts(FdtdHipContext *c, FdtdHipCounts *out) {
  if (!c || !out) return fail("get_counts: null argument");
  out->nSampledVoltages = c->ports.nsv;
  out->nSampledCurrents = c->ports.nsi;
  out->nProbes = c->ports.npr;
  out->nSources = c->ports.nvs;
  out->nFarfieldFreq = c->nff;
  out->nPortDftFreq = c->ndft;
  return FDTD_HIP_OK;
}

int fdtd_hip_run_batch_async(FdtdHipContext *c, int tsStart, int count) {
  if (!c) return fail("run_batch: null context");
  if (tsStart < 0 || count < 0 || tsStart + count > c->d.numberOfTimeSteps)
    return fail("run_batch: steps [%d,%d) outside [0,%d)", tsStart, tsStart + count, c->d.numberOfTimeSteps);
  if (c->ndft && c->ports.nsv + c->ports.nsi == 0) return fail("run_batch: port DFT without observers");
  if (ensure_traces(c) || finalize_coefficients(c)) return FDTD_HIP_ERR;
  if (set_cpml_tiles(c)) return FDTD_HIP_ERR;
  if (ntff_blocked(c) && ensure_ntff_buffer(c)) return FDTD_HIP_ERR;
  c->e_tail_ok = e_tail_mergeable(c) ? 1 : 0;
  if (c->fp32() ? run_steps<float>(c, tsStart, count) : run_steps<double>(c, tsStart, count)) return FDTD_HIP_ERR;
  // Leave the accumulators complete at the end of every batch.
  if (ntff_blocked(c) && (c->fp32() ? ntff_flush<float>(c) : ntff_flush<double>(c))) return FDTD_HIP_ERR;
  HIP_TRY(hipGetLastError());
  c->steps_run += count;
  return FDTD_HIP_OK;
}

int fdtd_hip_read_traces(FdtdHipContext *c, int tsStart, int count, double *svOut, double *siOut) {
  if (!c) return fail("read_traces: null context");
  if (tsStart < 0 || count < 0 || tsStart + count > c->d.numberOfTimeSteps)
    return fail("read_traces: steps [%d,%d) outside [0,%d)", tsStart, tsStart + count, c->d.numberOfTimeSteps);
  if (!c->traces_ready && (svOut || siOut)) return fail("read_traces: nothing has run yet");
  for (int o = 0; svOut && o < c->ports.nsv; ++o)
    HIP_TRY(hipMemcpyAsync(svOut + (size_t)o * count, c->ports.vtr + (size_t)o * c->ports.steps + tsStart,
                           count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  for (int o = 0; siOut && o < c->ports.nsi; ++o)
    HIP_TRY(hipMemcpyAsync(siOut + (size_t)o * count, c->ports.itr + (size_t)o * c->ports.steps + tsStart,
                           count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_upload_field_probe(FdtdHipContext *c, int magnetic, int component, int i, int j, int k) {
  if (!c) return fail("upload_field_probe: null context");
  drop_graph(c);
  if (magnetic != 0 && magnetic != 1) return fail("upload_field_probe: magnetic must be 0 or 1");
  if (component < 0 || component > 3) return fail("upload_field_probe: component %d out of range (0..3)", component);
  if (c->traces_ready) return fail("upload_field_probe: probes must be uploaded before the first step");
  if (c->ports.npr >= MAX_FIELD_PROBES) return fail("upload_field_probe: at most %d probes", MAX_FIELD_PROBES);
  // Averaged elements (sampling.js:144-292): E along its own axis back one cell; H over
  // the two transverse axes back one cell, in the reference's summation order.
  static const int kE[3][2][3] = {{{-1, 0, 0}, {0, 0, 0}}, {{0, -1, 0}, {0, 0, 0}}, {{0, 0, -1}, {0, 0, 0}}};
  static const int kH[3][4][3] = {{{0, -1, -1}, {0, 0, -1}, {0, -1, 0}, {0, 0, 0}},
                                  {{-1, 0, -1}, {0, 0, -1}, {-1, 0, 0}, {0, 0, 0}},
                                  {{-1, -1, 0}, {0, -1, 0}, {-1, 0, 0}, {0, 0, 0}}};
  int d[PROBE_INTS] = {magnetic, component, magnetic ? 4 : 2, 0};
  for (int cc = 0; cc < 3; ++cc) {
    if (component < 3 && cc != component) continue;
    const int comp = (magnetic ? 0 : 3) + cc;
    int dims[3];
    c->comp_dims(comp, dims);
    for (int t = 0; t < (magnetic ? 4 : 2); ++t) {
      const int *o = magnetic ? kH[cc][t] : kE[cc][t];
      const int ii = i + o[0], jj = j + o[1], kk = k + o[2];
      if (ii < 0 || ii >= dims[0] || jj < 0 || jj >= dims[1] || kk < 0 || kk >= dims[2])
        return fail("upload_field_probe: %c%c(%d,%d,%d) outside its %dx%dx%d array", magnetic ? 'H' : 'E',
                    "xyz"[cc], ii, jj, kk, dims[0], dims[1], dims[2]);
      d[4 + 4 * cc + t] = (int)c->to_dev(comp, ((long long)ii * dims[1] + jj) * dims[2] + kk);
      if (!magnetic) c->tail_e_read.push_back(c->e_cell(cc, d[4 + 4 * cc + t]));
    }
  }
  c->probe_host.insert(c->probe_host.end(), d, d + PROBE_INTS);
  c->ports.npr++;
  return FDTD_HIP_OK;
}

int fdtd_hip_read_probes(FdtdHipContext *c, int tsStart, int count, double *out) {
  if (!c || !out) return fail("read_probes: null argument");
  if (tsStart < 0 || count < 0 || tsStart + count > c->d.numberOfTimeSteps)
    return fail("read_probes: steps [%d,%d) outside [0,%d)", tsStart, tsStart + count, c->d.numberOfTimeSteps);
  if (!c->traces_ready) return fail("read_probes: nothing has run yet");
  for (int p = 0; p < c->ports.npr; ++p)
    HIP_TRY(hipMemcpyAsync(out + (size_t)p * count, c->ports.prtr + (size_t)p * c->ports.steps + tsStart,
                           count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_run_batch(FdtdHipContext *c, int tsStart, int count, double *svOut, double *siOut) {
  if (fdtd_hip_run_batch_async(c, tsStart, count)) return FDTD_HIP_ERR;
  return fdtd_hip_read_traces(c, tsStart, count, svOut, siOut);
}

int fdtd_hip_get_step_info(FdtdHipContext *c, int *e_tail_merged) {
  if (!c || !e_tail_merged) return fail("get_step_info: null argument");
  *e_tail_merged = c->e_tail_merged;
  return FDTD_HIP_OK;
}

int fdtd_hip_query(FdtdHipContext *c, int *idle) {
  if (!c || !idle) return fail("query: null argument");
  const hipError_t e = hipStreamQuery(c->stream);
  if (e != hipSuccess && e != hipErrorNotReady) return fail("query: %s", hipGetErrorString(e));
  *idle = e == hipSuccess ? 1 : 0;
  return FDTD_HIP_OK;
}

int fdtd_hip_read_field(FdtdHipContext *c, int which, double *host, long long n) {
  if (!c || !host) return fail("read_field: null argument");
  if (which < 0 || which > 5) return fail("read_field: field id %d out of range", which);
  if (n != c->comp_len(which)) return fail("read_field: %lld elements, expected %lld", n, c->comp_len(which));
  if (c->fp32()) {
    std::vector<float> v(c->ntot);
    HIP_TRY(hipMemcpyAsync(v.data(), c->field_ptr(which), c->ntot * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unpack_uniform(c, which, v, host);
  } else {
    std::vector<double> v(c->ntot);
    HIP_TRY(hipMemcpyAsync(v.data(), c->field_ptr(which), c->ntot * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unpack_uniform(c, which, v, host);
  }
  return FDTD_HIP_OK;
}

int fdtd_hip_write_field(FdtdHipContext *c, int which, const double *host, long long n) {
  if (!c || !host) return fail("write_field: null argument");
  if (which < 0 || which > 5) return fail("write_field: field id %d out of range", which);
  if (n != c->comp_len(which)) return fail("write_field: %lld elements, expected %lld", n, c->comp_len(which));
  if (c->fp32()) {
    std::vector<float> v;
    pack_uniform(c, which, host, v);
    return upload_vec(c, c->field_ptr(which), v);
  }
  std::vector<double> v;
  pack_uniform(c, which, host, v);
  return upload_vec(c, c->field_ptr(which), v);
}

static int psi_check(FdtdHipContext *c, int faceId, int slot, long long n, const char *who) {
  if (!c) return fail("%s: null context", who);
  if (faceId < 0 || faceId > 5 || slot < 0 || slot > 3) return fail("%s: face %d slot %d out of range", who, faceId, slot);
  if (!c->face[faceId].nc) return fail("%s: face %d not uploaded", who, faceId);
  if (n != c->face[faceId].n[slot]) return fail("%s: %lld elements, expected %lld", who, n, c->face[faceId].n[slot]);
  return FDTD_HIP_OK;
}

int fdtd_hip_read_psi(FdtdHipContext *c, int faceId, int slot, double *host, long long n) {
  if (!host) return fail("read_psi: null array");
  if (psi_check(c, faceId, slot, n, "read_psi")) return FDTD_HIP_ERR;
  if (c->fp32()) {
    std::vector<float> v(n);
    HIP_TRY(hipMemcpyAsync(v.data(), c->face[faceId].psi[slot], n * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (long long q = 0; q < n; ++q) host[q] = v[q];
  } else {
    HIP_TRY(hipMemcpyAsync(host, c->face[faceId].psi[slot], n * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return FDTD_HIP_OK;
}

int fdtd_hip_write_psi(FdtdHipContext *c, int faceId, int slot, const double *host, long long n) {
  if (!host) return fail("write_psi: null array");
  if (psi_check(c, faceId, slot, n, "write_psi")) return FDTD_HIP_ERR;
  if (c->fp32()) return upload_vec(c, c->face[faceId].psi[slot], std::vector<float>(host, host + n));
  return upload_vec(c, c->face[faceId].psi[slot], std::vector<double>(host, host + n));
}

int fdtd_hip_read_farfield(FdtdHipContext *c, int which, double *re, double *im, long long n) {
  if (!c || !re || !im) return fail("read_farfield: null argument");
  if (!c->nff) return fail("read_farfield: no far-field box uploaded");
  if (which < 0 || which >= 24) return fail("read_farfield: accumulator %d out of range", which);
  if (n != c->ff_sizes[which]) return fail("read_farfield: %lld elements, expected %lld", n, c->ff_sizes[which]);
  const FarfieldState &f = c->ff;
  const double *re_d = f.acc + ff_offset(which, f.nfreq, (long)f.nj * f.nk, (long)f.ni * f.nk, (long)f.ni * f.nj);
  HIP_TRY(hipMemcpyAsync(re, re_d, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(im, re_d + n, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_read_port_dft(FdtdHipContext *c, double *re, double *im, long long n) {
  if (!c || !re || !im) return fail("read_port_dft: null argument");
  if (!c->ndft) return fail("read_port_dft: no port DFT uploaded");
  const long long want = (long long)(c->ports.nsv + c->ports.nsi) * c->ndft;
  if (n != want) return fail("read_port_dft: %lld elements, expected %lld", n, want);
  if (!c->dft_acc) {                     // no step has run: every transform is zero
    std::fill(re, re + n, 0.0);
    std::fill(im, im + n, 0.0);
    return FDTD_HIP_OK;
  }
  std::vector<double> v(2 * n);
  HIP_TRY(hipMemcpyAsync(v.data(), c->dft_acc, 2 * n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (long long q = 0; q < n; ++q) { re[q] = v[2 * q] * c->d.dt; im[q] = v[2 * q + 1] * c->d.dt; }
  return FDTD_HIP_OK;
}

extern "C++" {
namespace {

NtffPost ntff_post_params(const FdtdHipContext *c) {
  const FarfieldState &f = c->ff;
  NtffPost p;
  p.acc = f.acc;
  p.nfreq = f.nfreq; p.ni = f.ni; p.nj = f.nj; p.nk = f.nk;
  p.li = f.LI + 1; p.lj = f.LJ + 1; p.lk = f.LK + 1;
  p.ui = f.UI + 1; p.uj = f.UJ + 1; p.uk = f.UK + 1;
  p.dx = c->d.dx; p.dy = c->d.dy; p.dz = c->d.dz;
  return p;
}

// Device scratch of one post-processing call, freed when the call returns.
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() { for (void *p : ptrs) (void)hipFree(p); }
  template <typename T>
  T *put(FdtdHipContext *c, const T *host, size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T) + 16) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    if (host && hipMemcpyAsync(p, host, n * sizeof(T), hipMemcpyHostToDevice, c->stream) != hipSuccess) return nullptr;
    return (T *)p;
  }
};

}  // namespace
}  // extern "C++"

int fdtd_hip_radiated_power(FdtdHipContext *c, double *out, int nFreq) {
  if (!c || !out) return fail("radiated_power: null argument");
  if (!c->nff) return fail("radiated_power: no far-field box uploaded");
  if (nFreq != c->nff) return fail("radiated_power: %d frequencies, the box has %d", nFreq, c->nff);
  Scratch s;
  double *d = s.put<double>(c, nullptr, nFreq);
  if (!d) return fail("radiated_power: device allocation failed");
  hipLaunchKernelGGL(k_ntff_power, dim3(nFreq), dim3(256), 0, c->stream, ntff_post_params(c), d, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, d, nFreq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

// Directivity of a context into device workspace `ws` laid out by DirLayout for
// capacities (nf_cap, na_cap): inputs staged from the host, the trigonometry of the
// angles and the wave numbers on the device, the radiated power (unless given) and the
// directivity kernel, outputs copied back.  Shared by the allocating and the
// no-allocation entry points, which therefore give identical results.
struct DirLayout {
  long kw, prad, freq, theta, phi, ang, dth, dph, total;
  DirLayout(long nf_cap, long na_cap, long out_cap) {
    kw = 0; prad = nf_cap; freq = 2 * nf_cap; theta = 3 * nf_cap; phi = theta + na_cap;
    ang = phi + na_cap; dth = ang + 4 * na_cap; dph = dth + out_cap; total = dph + out_cap;
  }
};

// Directivity at the nf NTFF frequencies f0 .. f0+nf-1 (freqHz and radPower hold nf
// values each; radPower null: computed here).
static int directivity_impl(FdtdHipContext *c, int f0, int nf, const double *freqHz, const double *radPower,
                            int nAngles, const double *theta, const double *phi, double *dataTheta,
                            double *dataPhi, double *ws, const DirLayout &L) {
  const size_t nout = (size_t)nf * nAngles;
  HIP_TRY(hipMemcpyAsync(ws + L.freq, freqHz, nf * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(ws + L.theta, theta, nAngles * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(ws + L.phi, phi, nAngles * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (radPower)
    HIP_TRY(hipMemcpyAsync(ws + L.prad, radPower, nf * sizeof(double), hipMemcpyHostToDevice, c->stream));
  else
    hipLaunchKernelGGL(k_ntff_power, dim3(nf), dim3(256), 0, c->stream, ntff_post_params(c), ws + L.prad, f0);
  const int nprep = std::max(nf, nAngles);
  hipLaunchKernelGGL(k_directivity_prep, dim3((nprep + 255) / 256), dim3(256), 0, c->stream, ws + L.freq,
                     ws + L.theta, ws + L.phi, nf, nAngles, ws + L.kw, ws + L.ang);
  constexpr int NA = 4;
  hipLaunchKernelGGL(k_ntff_directivity<NA>, dim3((nAngles + NA - 1) / NA, nf), dim3(256), 0, c->stream,
                     ntff_post_params(c), ws + L.kw, ws + L.prad, ws + L.ang, nAngles, ws + L.dth, ws + L.dph, f0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(dataTheta, ws + L.dth, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(dataPhi, ws + L.dph, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

static int directivity_check(FdtdHipContext *c, const double *freqHz, const double *theta, const double *phi,
                             const double *dataTheta, const double *dataPhi, int nAngles, const char *who) {
  if (!c || !freqHz || !theta || !phi || !dataTheta || !dataPhi) return fail("%s: null argument", who);
  if (!c->nff) return fail("%s: no far-field box uploaded", who);
  if (nAngles <= 0) return fail("%s: %d angles", who, nAngles);
  return FDTD_HIP_OK;
}

int fdtd_hip_directivity(FdtdHipContext *c, const double *freqHz, const double *radPower, int nAngles,
                         const double *theta, const double *phi, double *dataTheta, double *dataPhi) {
  if (directivity_check(c, freqHz, theta, phi, dataTheta, dataPhi, nAngles, "directivity")) return FDTD_HIP_ERR;
  const DirLayout L(c->nff, nAngles, (long)c->nff * nAngles);
  Scratch s;
  double *ws = s.put<double>(c, nullptr, (size_t)L.total);
  if (!ws) return fail("directivity: device allocation failed");
  return directivity_impl(c, 0, c->nff, freqHz, radPower, nAngles, theta, phi, dataTheta, dataPhi, ws, L);
}

int fdtd_hip_directivity_at(FdtdHipContext *c, int f, double freqHz, const double *radPower, int nAngles,
                            const double *theta, const double *phi, double *dataTheta, double *dataPhi) {
  if (directivity_check(c, &freqHz, theta, phi, dataTheta, dataPhi, nAngles, "directivity_at")) return FDTD_HIP_ERR;
  if (f < 0 || f >= c->nff) return fail("directivity_at: frequency %d outside [0,%d)", f, c->nff);
  const DirLayout L(1, nAngles, nAngles);
  Scratch s;
  double *ws = s.put<double>(c, nullptr, (size_t)L.total);
  if (!ws) return fail("directivity_at: device allocation failed");
  return directivity_impl(c, f, 1, &freqHz, radPower, nAngles, theta, phi, dataTheta, dataPhi, ws, L);
}

int fdtd_hip_directivity_noalloc(FdtdHipContext *c, const double *freqHz, const double *radPower, int nAngles,
                                 const double *theta, const double *phi, double *dataTheta, double *dataPhi) {
  if (directivity_check(c, freqHz, theta, phi, dataTheta, dataPhi, nAngles, "directivity_noalloc")) return FDTD_HIP_ERR;
  if (c->nff > FDTD_HIP_NOALLOC_MAX_FREQ || nAngles > FDTD_HIP_NOALLOC_MAX_ANGLES ||
      (long)c->nff * nAngles > FDTD_HIP_NOALLOC_MAX_OUTPUTS)
    return fail("directivity_noalloc: %d frequencies x %d angles exceed the fixed workspace (at most %d frequencies, "
                "%d angles, %d outputs); use fdtd_hip_directivity", c->nff, nAngles, FDTD_HIP_NOALLOC_MAX_FREQ,
                FDTD_HIP_NOALLOC_MAX_ANGLES, FDTD_HIP_NOALLOC_MAX_OUTPUTS);
  double *ws = noalloc_workspace(c->device);
  if (!ws) return fail("directivity_noalloc: no workspace address for device %d", c->device);
  const DirLayout L(FDTD_HIP_NOALLOC_MAX_FREQ, FDTD_HIP_NOALLOC_MAX_ANGLES, FDTD_HIP_NOALLOC_MAX_OUTPUTS);
  return directivity_impl(c, 0, c->nff, freqHz, radPower, nAngles, theta, phi, dataTheta, dataPhi, ws, L);
}

int fdtd_hip_set_ntff_mode(FdtdHipContext *c, int mode) {
  if (!c) return fail("set_ntff_mode: null context");
  drop_graph(c);
  if (mode < FDTD_HIP_NTFF_AUTO || mode > FDTD_HIP_NTFF_BLOCKED) return fail("set_ntff_mode: unknown mode %d", mode);
  const bool was = ntff_blocked(c);
  // Rows gathered in blocked mode are folded in before the switch.
  if (was && c->ntff_buf && (c->fp32() ? ntff_flush<float>(c) : ntff_flush<double>(c))) return FDTD_HIP_ERR;
  c->ntff_mode = mode;
  if (ntff_blocked(c) && !was) c->blk_rows = c->blk_flushed = 0, c->blk_t0 = -1;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_set_graphs(FdtdHipContext *c, int enable) {
  if (!c) return fail("set_graphs: null context");
  drop_graph(c);
  c->use_graphs = enable ? 1 : 0;
  return FDTD_HIP_OK;
}

int fdtd_hip_get_ntff_mode(FdtdHipContext *c, int *blocked) {
  if (!c || !blocked) return fail("get_ntff_mode: null argument");
  *blocked = ntff_blocked(c) ? 1 : 0;
  return FDTD_HIP_OK;
}

static int check_tuning(const FdtdHipContext *c, const FdtdHipTuning *t, const char *who) {
  if (t->planes_per_workgroup < 0 || t->planes_per_workgroup > 4096)
    return fail("%s: %d planes per workgroup out of range", who, t->planes_per_workgroup);
  const int maxw = c->fp32() ? 4 : 2;
  if (t->cells_per_thread != 0 && t->cells_per_thread != 1 && t->cells_per_thread != 2 &&
      !(t->cells_per_thread == 4 && maxw == 4))
    return fail("%s: %d cells per thread unsupported (1, 2%s)", who, t->cells_per_thread, maxw == 4 ? " or 4" : "");
  if (t->coefficient_storage < 0 || t->coefficient_storage > 1)
    return fail("%s: coefficient storage %d unknown", who, t->coefficient_storage);
  if (t->fuse_cpml < CPML_SEPARATE || t->fuse_cpml > CPML_FUSE_XY)
    return fail("%s: fuse_cpml %d unknown (0 to 5)", who, t->fuse_cpml);
  return FDTD_HIP_OK;
}

int fdtd_hip_set_tuning(FdtdHipContext *c, const FdtdHipTuning *t) {
  if (!c || !t) return fail("set_tuning: null argument");
  drop_graph(c);
  if (check_tuning(c, t, "set_tuning")) return FDTD_HIP_ERR;
  if (t->coefficient_storage != c->tune.coefficient_storage) c->coef_mode = -1;
  c->tune = *t;
  return FDTD_HIP_OK;
}

int fdtd_hip_get_tuning(FdtdHipContext *c, FdtdHipTuning *t) {
  if (!c || !t) return fail("get_tuning: null argument");
  *t = c->tune;
  return FDTD_HIP_OK;
}

// Zero the dynamic state (fields, CPML Psi, NTFF and port-DFT accumulators, traces).
static int reset_state(FdtdHipContext *c) {
  HIP_TRY(hipMemsetAsync(c->fields, 0, 6 * c->stride_bytes, c->stream));
  for (auto &f : c->face)
    for (int s = 0; s < 4; ++s)
      if (f.psi[s]) HIP_TRY(hipMemsetAsync(f.psi[s], 0, f.n[s] * c->esz, c->stream));
  if (c->ff_buf) {
    long long total = 0;
    for (int a = 0; a < 24; ++a) total += 2 * c->ff_sizes[a];
    HIP_TRY(hipMemsetAsync(c->ff_buf, 0, total * sizeof(double), c->stream));
  }
  if (c->dft_acc) HIP_TRY(hipMemsetAsync(c->dft_acc, 0, 2 * sizeof(double) * std::max(1, c->ports.nsv + c->ports.nsi) * c->ndft, c->stream));
  if (c->ntff_buf) HIP_TRY(hipMemsetAsync(c->ntff_buf, 0, (size_t)NTFF_K * c->ntff_spad * c->esz, c->stream));
  c->blk_t0 = c->blk_rows = c->blk_flushed = 0;
  if (c->traces_ready) {
    HIP_TRY(hipMemsetAsync(c->ports.vtr, 0, sizeof(double) * std::max(1, c->ports.nsv) * c->ports.steps, c->stream));
    HIP_TRY(hipMemsetAsync(c->ports.itr, 0, sizeof(double) * std::max(1, c->ports.nsi) * c->ports.steps, c->stream));
    HIP_TRY(hipMemsetAsync(c->ports.prtr, 0, sizeof(double) * std::max(1, c->ports.npr) * c->ports.steps, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_autotune(FdtdHipContext *c, int steps_per_candidate, FdtdHipTuning *chosen) {
  if (!c) return fail("autotune: null context");
  drop_graph(c);
  if (c->steps_run) return fail("autotune: must run before the first time step (it resets the state)");
  if (steps_per_candidate < 1) steps_per_candidate = 2;
  const int steps = std::min(steps_per_candidate + 1, c->d.numberOfTimeSteps);
  if (ensure_traces(c) || build_coefficient_tables(c)) return FDTD_HIP_ERR;
  if (set_cpml_tiles(c)) return FDTD_HIP_ERR;
  if (ntff_blocked(c) && ensure_ntff_buffer(c)) return FDTD_HIP_ERR;
  const FdtdHipTuning saved = c->tune;
  const int saved_timing = c->timing;
  std::vector<FdtdHipTuning> cands;
  const int ws[3] = {1, 2, 4};
  const int planes[4] = {1, 2, 4, 8};
  for (int cs = 0; cs < 2; ++cs) {
    if (cs == 0 && !c->coef_idx_bytes) continue;
    for (int w : ws) {
      if (w == 4 && !c->fp32()) continue;
      for (int pl : planes)
        for (int fc = CPML_SEPARATE; fc <= CPML_FUSE_XY; ++fc) cands.push_back(FdtdHipTuning{pl, w, cs, fc});
    }
  }
  double best = 1e300;
  FdtdHipTuning best_t = saved;
  for (const auto &t : cands) {
    c->tune = t;
    if (set_cpml_tiles(c)) return FDTD_HIP_ERR;
    c->coef_mode = -1;
    if (finalize_coefficients(c)) return FDTD_HIP_ERR;
    c->timing = 1;
    c->times = FdtdHipKernelTimes{};
    for (int s = 0; s < steps; ++s) {
      if (s == 1) c->times = FdtdHipKernelTimes{};     // first step warms up
      const int rc = c->fp32() ? launch_step<float>(c, s) : launch_step<double>(c, s);
      if (rc) return rc;
    }
    const double ms = c->times.ms[0] + c->times.ms[1] + c->times.ms[4] + c->times.ms[5];
    if (ms < best) { best = ms; best_t = t; }
  }
  c->tune = best_t;
  c->coef_mode = -1;
  c->timing = saved_timing;
  c->times = FdtdHipKernelTimes{};
  if (finalize_coefficients(c) || reset_state(c)) return FDTD_HIP_ERR;
  if (chosen) *chosen = best_t;
  return FDTD_HIP_OK;
}

int fdtd_hip_get_coefficient_storage(FdtdHipContext *c, int *mode, int *entries) {
  if (!c || !mode) return fail("get_coefficient_storage: null argument");
  if (finalize_coefficients(c)) return FDTD_HIP_ERR;
  *mode = c->coef_mode;
  if (entries) for (int q = 0; q < 6; ++q) entries[q] = c->coef_n[q];
  return FDTD_HIP_OK;
}

int fdtd_hip_set_timing(FdtdHipContext *c, int enable) {
  if (!c) return fail("set_timing: null context");
  c->timing = enable ? 1 : 0;
  c->times = FdtdHipKernelTimes{};
  return FDTD_HIP_OK;
}

int fdtd_hip_get_kernel_times(FdtdHipContext *c, FdtdHipKernelTimes *out) {
  if (!c || !out) return fail("get_kernel_times: null argument");
  *out = c->times;
  return FDTD_HIP_OK;
}

int fdtd_hip_synchronize(FdtdHipContext *c) {
  if (!c) return fail("synchronize: null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return FDTD_HIP_OK;
}

void fdtd_hip_destroy(FdtdHipContext *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &f : c->face) {
    for (int s = 0; s < 4; ++s) { (void)hipFree(f.psi[s]); (void)hipFree(f.cpsi[s]); }
  }
  for (void *p : c->small_allocs) (void)hipFree(p);
  for (void *p : c->coef_tab) (void)hipFree(p);
  (void)hipFree(c->coef_idx);
  (void)hipFree(c->tile_buf);
  (void)hipFree(c->ff_buf);
  (void)hipFree(c->ntff_buf);
  drop_graph(c);
  (void)hipFree(c->d_step);
  (void)hipFree(c->fields);
  (void)hipFree(c->coeffs);
  for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}


// ------------------------------------------------------------- many problems --

struct FdtdHipMulti {
  std::vector<FdtdHipContext *> ctx;
  int device = 0;
  bool fp32 = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  FdtdHipTuning tune{0, 0, 0, CPML_SEPARATE};
  int cm = -1;                  // coefficient storage shared by every variant
  int ntff = 0;                 // 0: no variant has NTFF, 1: direct, 2: blocked
  bool ports = false;
  bool e_tail = false;          // every variant's E-phase tail may run as one launch
  // Device descriptors and workgroup maps, with the host images last uploaded.
  void *d_desc = nullptr;
  unsigned *d_map = nullptr;    // [H tiles | E tiles | H CPML | E CPML | NTFF blocks]
  size_t desc_cap = 0, map_cap = 0;
  std::vector<char> h_desc;
  std::vector<unsigned> h_map;
  int nh = 0, ne = 0, nch = 0, nce = 0, nfb = 0;   // H / E tiles, H / E CPML blocks, NTFF blocks
  int *d_step = nullptr;
  hipGraphExec_t graph = nullptr;
  int use_graphs = 1;
  std::vector<hipStream_t> saved;   // the contexts' own streams while they are lent
};

extern "C++" {
namespace {

void multi_drop_graph(FdtdHipMulti *m) {
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  m->graph = nullptr;
}

// Lend every context's stream to the batch: the batch stream first waits for the
// work already queued on each, and the contexts then launch on the batch stream.
int multi_borrow(FdtdHipMulti *m) {
  m->saved.resize(m->ctx.size());
  for (size_t v = 0; v < m->ctx.size(); ++v) {
    FdtdHipContext *c = m->ctx[v];
    HIP_TRY(hipEventRecord(m->ev, c->stream));
    HIP_TRY(hipStreamWaitEvent(m->stream, m->ev, 0));
    m->saved[v] = c->stream;
    c->stream = m->stream;
  }
  return FDTD_HIP_OK;
}

// Give the streams back; each waits for everything the batch queued so far.
int multi_return(FdtdHipMulti *m) {
  if (m->saved.size() != m->ctx.size()) return FDTD_HIP_OK;
  HIP_TRY(hipEventRecord(m->ev, m->stream));
  for (size_t v = 0; v < m->ctx.size(); ++v) {
    m->ctx[v]->stream = m->saved[v];
    HIP_TRY(hipStreamWaitEvent(m->saved[v], m->ev, 0));
  }
  m->saved.clear();
  return FDTD_HIP_OK;
}

template <typename T>
void put_map(std::vector<unsigned> &map, int v, long n) {
  for (long q = 0; q < n; ++q) map.push_back(((unsigned)v << MULTI_LOCAL_BITS) | (unsigned)q);
}

// Build the deduplicated coefficient tables of every variant that still lacks them,
// on up to 8 host threads: the work is host hashing plus uploads on each context's own
// stream (about 10 ms per IFA variant, which one thread would run back to back).
// A failure's message comes back to the calling thread.
int build_tables_threaded(const std::vector<FdtdHipContext *> &ctx, bool indexed) {
  if (!indexed) return FDTD_HIP_OK;
  std::vector<FdtdHipContext *> todo;
  for (FdtdHipContext *c : ctx)
    if (!c->coef_built) todo.push_back(c);
  if (todo.size() < 2) return FDTD_HIP_OK;
  std::atomic<size_t> next{0};
  std::vector<int> rc(todo.size(), FDTD_HIP_OK);
  std::vector<std::string> msg(todo.size());
  auto worker = [&] {
    for (size_t k; (k = next.fetch_add(1)) < todo.size();) {
      FdtdHipContext *c = todo[k];
      if (hipSetDevice(c->device) != hipSuccess) { rc[k] = FDTD_HIP_ERR; msg[k] = "hipSetDevice failed"; continue; }
      try {
        rc[k] = build_coefficient_tables(c);
        if (rc[k]) msg[k] = g_last_error;
      } catch (const std::exception &e) {     // e.g. out of host memory: report, do not terminate
        rc[k] = FDTD_HIP_ERR;
        msg[k] = std::string("coefficient tables: ") + e.what();
      }
    }
  };
  const size_t nth = std::min<size_t>(todo.size(), 8);
  std::vector<std::thread> pool;
  for (size_t t = 0; t < nth; ++t) pool.emplace_back(worker);
  for (auto &t : pool) t.join();
  for (size_t k = 0; k < todo.size(); ++k)
    if (rc[k]) return fail("%s", msg[k].c_str());
  return FDTD_HIP_OK;
}

// Apply the batch tuning to every context, make the coefficient storage uniform,
// and (re)upload the descriptors and maps when they changed.
template <typename T>
int multi_prepare(FdtdHipMulti *m) {
  if (build_tables_threaded(m->ctx, m->tune.coefficient_storage != 1)) return FDTD_HIP_ERR;
  for (int pass = 0; pass < 2; ++pass) {
    for (FdtdHipContext *c : m->ctx) {
      FdtdHipTuning t = m->tune;
      if (pass == 1) t.coefficient_storage = 1;
      if (std::memcmp(&t, &c->tune, sizeof t) != 0) {
        c->tune = t;
        c->coef_mode = -1;
        drop_graph(c);
      }
      if (ensure_traces(c) || finalize_coefficients(c)) return FDTD_HIP_ERR;
      if (set_cpml_tiles(c)) return FDTD_HIP_ERR;
      if (ntff_blocked(c) && ensure_ntff_buffer(c)) return FDTD_HIP_ERR;
    }
    bool same = true;
    for (FdtdHipContext *c : m->ctx) same = same && c->coef_mode == m->ctx[0]->coef_mode;
    if (same) break;
    if (pass == 1) return fail("multi: coefficient storage differs between variants");
  }
  m->cm = m->ctx[0]->coef_mode;
  m->ntff = 0;
  m->ports = false;
  m->e_tail = true;
  for (FdtdHipContext *c : m->ctx) {
    m->e_tail = m->e_tail && e_tail_mergeable(c);
    if (c->nff) {
      const int mode = ntff_blocked(c) ? 2 : 1;
      if (m->ntff && m->ntff != mode) return fail("multi: variants mix blocked and direct NTFF accumulation");
      m->ntff = mode;
    }
    if (c->ports.nvs + c->ports.nsv + c->ports.nsi + c->ports.npr > 0) m->ports = true;
  }

  const int n = (int)m->ctx.size();
  std::vector<VariantDesc<T>> desc(n);
  std::vector<unsigned> hmap, emap, chmap, cemap, fmap;
  for (int v = 0; v < n; ++v) {
    FdtdHipContext *c = m->ctx[v];
    VariantDesc<T> &d = desc[v];
    std::memset(&d, 0, sizeof d);
    d.g = make_grid<T>(c);
    d.fs = make_faces<T>(c);
    d.ghost = c->d.boundary == FDTD_HIP_BOUNDARY_REFERENCE;
    d.mh = march_geom(c, c->d.nx + 1);
    d.me = march_geom(c, d.ghost ? c->d.nx : c->d.nx + 1);
    d.th = c->cpml_tiles[0];
    d.te = c->cpml_tiles[1];
    d.ps = c->ports;
    d.ps.ndft = c->ndft; d.ps.w = c->dft_w; d.ps.dft = c->dft_acc;
    d.has_ports = c->ports.nvs + c->ports.nsv + c->ports.nsi + c->ports.npr > 0;
    d.ff = c->ff;
    if (ntff_blocked(c)) d.nb = ntff_block<T>(c);
    const long th = (long)d.mh.ncol * d.mh.nchunk, te = (long)d.me.ncol * d.me.nchunk;
    long tf = 0;
    if (c->nff) {
      const FarfieldState &f = c->ff;
      tf = (2 * ((long)f.nj * f.nk + (long)f.ni * f.nk + (long)f.ni * f.nj) + 255) / 256;
    }
    const long tch = c->ncpml[0], tce = c->ncpml[1];
    if (th > (long)MULTI_LOCAL_MASK || te > (long)MULTI_LOCAL_MASK || tf > (long)MULTI_LOCAL_MASK ||
        tch > (long)MULTI_LOCAL_MASK || tce > (long)MULTI_LOCAL_MASK)
      return fail("multi: variant %d is too large for a batched launch", v);
    put_map<T>(hmap, v, th);
    put_map<T>(emap, v, te);
    put_map<T>(chmap, v, tch);
    put_map<T>(cemap, v, tce);
    put_map<T>(fmap, v, tf);
  }
  m->nh = (int)hmap.size();
  m->ne = (int)emap.size();
  m->nch = (int)chmap.size();
  m->nce = (int)cemap.size();
  m->nfb = (int)fmap.size();
  std::vector<unsigned> map(hmap);
  for (const auto *part : {&emap, &chmap, &cemap, &fmap}) map.insert(map.end(), part->begin(), part->end());
  // The launch choices a captured graph bakes in besides the device images.
  map.push_back((unsigned)m->cm);
  map.push_back((unsigned)cells_per_thread(m->ctx[0]));
  map.push_back((unsigned)m->ntff);
  map.push_back(m->ports ? 1u : 0u);
  map.push_back((unsigned)m->tune.fuse_cpml);
  map.push_back(m->e_tail ? 1u : 0u);
  const size_t dbytes = desc.size() * sizeof(VariantDesc<T>);
  const bool changed = dbytes != m->h_desc.size() || std::memcmp(desc.data(), m->h_desc.data(), dbytes) != 0 ||
                       map != m->h_map;
  if (!changed) return FDTD_HIP_OK;
  multi_drop_graph(m);
  HIP_TRY(hipStreamSynchronize(m->stream));     // no queued launch still reads the old images
  if (dbytes > m->desc_cap) {
    (void)hipFree(m->d_desc);
    m->d_desc = nullptr;
    HIP_TRY(hipMalloc(&m->d_desc, dbytes));
    m->desc_cap = dbytes;
  }
  if (map.size() * sizeof(unsigned) > m->map_cap || !m->d_map) {
    (void)hipFree(m->d_map);
    m->d_map = nullptr;
    const size_t bytes = std::max<size_t>(map.size(), 1) * sizeof(unsigned);
    HIP_TRY(hipMalloc((void **)&m->d_map, bytes));
    m->map_cap = bytes;
  }
  m->h_desc.assign((const char *)desc.data(), (const char *)desc.data() + dbytes);
  m->h_map = map;
  HIP_TRY(hipMemcpyAsync(m->d_desc, m->h_desc.data(), dbytes, hipMemcpyHostToDevice, m->stream));
  if (!map.empty())
    HIP_TRY(hipMemcpyAsync(m->d_map, m->h_map.data(), map.size() * sizeof(unsigned), hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));     // the host images may change before the next run
  return FDTD_HIP_OK;
}

template <typename T, int CM, int W, int FC>
void multi_launch_hw(FdtdHipMulti *m, bool e_phase) {
  const VariantDesc<T> *vd = (const VariantDesc<T> *)m->d_desc;
  const int nt = e_phase ? m->ne : m->nh;
  const unsigned *map = m->d_map + (e_phase ? m->nh : 0);
  const int per_xcd = (nt + 7) / 8;
  if (e_phase)
    hipLaunchKernelGGL((k_multi_update_e<T, CM, W, FC>), dim3(8 * per_xcd), dim3(256), 0, m->stream, vd, map, nt,
                       per_xcd);
  else
    hipLaunchKernelGGL((k_multi_update_h<T, CM, W, FC>), dim3(8 * per_xcd), dim3(256), 0, m->stream, vd, map, nt,
                       per_xcd);
}

template <typename T, int CM, int FC>
void multi_launch_hc(FdtdHipMulti *m, bool e_phase) {
  const int w = cells_per_thread(m->ctx[0]);
  if constexpr (sizeof(T) == 4) {
    if (w == 4) { multi_launch_hw<T, CM, 4, FC>(m, e_phase); return; }
  }
  if (w == 2) multi_launch_hw<T, CM, 2, FC>(m, e_phase);
  else multi_launch_hw<T, CM, 1, FC>(m, e_phase);
}

template <typename T, int FC>
void multi_launch_hf(FdtdHipMulti *m, bool e_phase) {
  if (m->cm == COEF_U8) multi_launch_hc<T, COEF_U8, FC>(m, e_phase);
  else if (m->cm == COEF_U16) multi_launch_hc<T, COEF_U16, FC>(m, e_phase);
  else multi_launch_hc<T, COEF_ARRAYS, FC>(m, e_phase);
}

// The stencil of one phase for every variant, then its CPML pass when one is left.
template <typename T>
void multi_launch_phase(FdtdHipMulti *m, bool e_phase) {
  switch (m->tune.fuse_cpml) {
    case CPML_FUSE_XY: multi_launch_hf<T, CPML_FUSE_XY>(m, e_phase); break;
    case CPML_FUSE_ALL: multi_launch_hf<T, CPML_FUSE_ALL>(m, e_phase); break;
    case CPML_FUSE_Z: multi_launch_hf<T, CPML_FUSE_Z>(m, e_phase); break;
    default: multi_launch_hf<T, CPML_SEPARATE>(m, e_phase); break;
  }
  const VariantDesc<T> *vd = (const VariantDesc<T> *)m->d_desc;
  const int nb = e_phase ? m->nce : m->nch;
  const unsigned *map = m->d_map + m->nh + m->ne + (e_phase ? m->nch : 0);
  if (nb == 0 || (e_phase && m->e_tail)) return;     // the E CPML tiles then run in k_multi_e_tail
  if (e_phase) hipLaunchKernelGGL(k_multi_cpml_e<T>, dim3((unsigned)nb), dim3(256), 0, m->stream, vd, map);
  else hipLaunchKernelGGL(k_multi_cpml_h<T>, dim3((unsigned)nb), dim3(256), 0, m->stream, vd, map);
}

// One time step of every variant (see launch_step for ts / tsp).
template <typename T>
int multi_step(FdtdHipMulti *m, int ts, const int *tsp = nullptr) {
  const VariantDesc<T> *vd = (const VariantDesc<T> *)m->d_desc;
  if (m->nh) multi_launch_phase<T>(m, false);
  if (m->ne) multi_launch_phase<T>(m, true);
  const unsigned *fmap = m->d_map + m->nh + m->ne + m->nch + m->nce;
  int row = ts;
  if (m->ntff == 2 && !tsp) {
    // Every variant's block state moves in lockstep (multi_run restarts them together).
    for (FdtdHipContext *c : m->ctx)
      if (c->nff) { row = c->blk_rows; break; }
  }
  for (FdtdHipContext *c : m->ctx) c->e_tail_merged = m->e_tail ? 1 : 0;
  if (m->e_tail) {
    // E CPML tiles, port epilogues and NTFF blocks in one launch (k_multi_e_tail).
    const int np = m->ports ? (int)m->ctx.size() : 0, nf = m->ntff ? m->nfb : 0;
    const long nblk = (long)np + nf + m->nce;
    const unsigned *cmap = m->d_map + m->nh + m->ne + m->nch;
    if (nblk > 0) {
      const dim3 grid((unsigned)nblk);
      if (m->ntff == 2)
        hipLaunchKernelGGL((k_multi_e_tail<T, 2>), grid, dim3(256), 0, m->stream, vd, np, fmap, nf, cmap, row, tsp, ts);
      else if (m->ntff == 1)
        hipLaunchKernelGGL((k_multi_e_tail<T, 1>), grid, dim3(256), 0, m->stream, vd, np, fmap, nf, cmap, row, tsp, ts);
      else
        hipLaunchKernelGGL((k_multi_e_tail<T, 0>), grid, dim3(256), 0, m->stream, vd, np, fmap, nf, cmap, row, tsp, ts);
    }
  } else {
    if (m->ports)
      hipLaunchKernelGGL(k_multi_port_epilogue<T>, dim3((unsigned)m->ctx.size()), dim3(256), 0, m->stream, vd, tsp, ts);
    if (m->ntff == 1)
      hipLaunchKernelGGL(k_multi_ntff_direct<T>, dim3((unsigned)m->nfb), dim3(256), 0, m->stream, vd, fmap, tsp, ts);
    else if (m->ntff == 2)
      hipLaunchKernelGGL(k_multi_ntff_gather<T>, dim3((unsigned)m->nfb), dim3(256), 0, m->stream, vd, fmap, row);
  }
  if (m->ntff == 2) {
    for (FdtdHipContext *c : m->ctx) {
      if (!c->nff) continue;
      if (tsp) {
        if (ts == NTFF_K - 1) ntff_fold<T>(c, tsp, 0, 0, NTFF_K);
      } else {
        ++c->blk_rows;
        if (c->blk_rows == NTFF_K && ntff_flush<T>(c)) return FDTD_HIP_ERR;
      }
    }
  }
  HIP_TRY(hipGetLastError());
  return FDTD_HIP_OK;
}

template <typename T>
int multi_build_graph(FdtdHipMulti *m) {
  multi_drop_graph(m);
  HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  int rc = FDTD_HIP_OK;
  for (int k = 0; k < GRAPH_STEPS && !rc; ++k) rc = multi_step<T>(m, k, m->d_step);
  if (!rc) hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, m->stream, m->d_step, GRAPH_STEPS);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(m->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (end != hipSuccess) return fail("multi: graph capture failed: %s", hipGetErrorString(end));
  const hipError_t inst = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (inst != hipSuccess) { m->graph = nullptr; return fail("multi: graph instantiation failed: %s", hipGetErrorString(inst)); }
  return FDTD_HIP_OK;
}

// Steps [ts, ts + count) of every variant, on the batch stream (contexts' streams lent).
template <typename T>
int multi_steps(FdtdHipMulti *m, int ts, int count, bool graphs) {
  if (m->ntff == 2) {
    // Continue the open blocks when every variant's block ends right at ts with the
    // same row count (as a single context continues its own), else fold what is left
    // and start new blocks at ts.
    int rows = -1;
    bool cont = true;
    for (FdtdHipContext *c : m->ctx) {
      if (!c->nff) continue;
      if (rows < 0) rows = c->blk_rows;
      cont = cont && c->blk_rows == rows && c->blk_t0 + c->blk_rows == ts;
    }
    for (FdtdHipContext *c : m->ctx) {
      if (!c->nff || cont) continue;
      if (ntff_flush<T>(c)) return FDTD_HIP_ERR;
      c->blk_t0 = ts;
      c->blk_rows = c->blk_flushed = 0;
    }
  }
  bool base_set = false;
  for (int s = 0; s < count;) {
    const int t = ts + s;
    bool aligned = true;
    for (FdtdHipContext *c : m->ctx)
      if (c->nff && ntff_blocked(c)) aligned = aligned && c->blk_rows == 0 && c->blk_t0 == t;
    if (graphs && count - s >= GRAPH_STEPS && aligned) {
      if (!m->graph) {
        if (multi_build_graph<T>(m)) return FDTD_HIP_ERR;
        base_set = false;
      }
      if (!base_set) hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, m->stream, m->d_step, t);
      base_set = true;
      HIP_TRY(hipGraphLaunch(m->graph, m->stream));
      for (FdtdHipContext *c : m->ctx) if (c->nff && ntff_blocked(c)) c->blk_t0 = t + GRAPH_STEPS;
      s += GRAPH_STEPS;
    } else {
      if (multi_step<T>(m, t)) return FDTD_HIP_ERR;
      base_set = false;
      ++s;
    }
  }
  if (m->ntff == 2)
    for (FdtdHipContext *c : m->ctx)
      if (c->nff && ntff_flush<T>(c)) return FDTD_HIP_ERR;
  HIP_TRY(hipGetLastError());
  return FDTD_HIP_OK;
}

// Steps [ts, ts+count) of every variant with the variants' streams already lent
// (multi_borrow): descriptors for the current tuning, then the launches.
int multi_run_lent(FdtdHipMulti *m, int ts, int count, bool graphs) {
  const int rc = m->fp32 ? multi_prepare<float>(m) : multi_prepare<double>(m);
  if (rc) return rc;
  return m->fp32 ? multi_steps<float>(m, ts, count, graphs) : multi_steps<double>(m, ts, count, graphs);
}

int multi_run(FdtdHipMulti *m, int ts, int count, bool graphs) {
  for (FdtdHipContext *c : m->ctx)
    if (ts < 0 || count < 0 || ts + count > c->d.numberOfTimeSteps)
      return fail("multi_run: steps [%d,%d) outside [0,%d) of a variant", ts, ts + count, c->d.numberOfTimeSteps);
  HIP_TRY(hipSetDevice(m->device));
  if (multi_borrow(m)) return FDTD_HIP_ERR;
  const int rc = multi_run_lent(m, ts, count, graphs);
  const int rc2 = multi_return(m);
  if (rc || rc2) return FDTD_HIP_ERR;
  for (FdtdHipContext *c : m->ctx) c->steps_run += count;
  return FDTD_HIP_OK;
}

}  // namespace
}  // extern "C++"

void fdtd_hip_multi_destroy(FdtdHipMulti *m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  multi_drop_graph(m);
  (void)hipFree(m->d_desc);
  (void)hipFree(m->d_map);
  (void)hipFree(m->d_step);
  if (m->ev) (void)hipEventDestroy(m->ev);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int fdtd_hip_multi_create(FdtdHipContext *const *ctxs, int n, const FdtdHipTuning *tuning, FdtdHipMulti **out) {
  if (!ctxs || !out) return fail("multi_create: null argument");
  *out = nullptr;
  if (n < 1 || n > MULTI_MAX_VARIANTS) return fail("multi_create: %d variants (1..%d)", n, MULTI_MAX_VARIANTS);
  for (int v = 0; v < n; ++v) {
    if (!ctxs[v]) return fail("multi_create: variant %d is null", v);
    for (int u = 0; u < v; ++u)
      if (ctxs[u] == ctxs[v]) return fail("multi_create: variant %d repeats variant %d", v, u);
    if (ctxs[v]->device != ctxs[0]->device) return fail("multi_create: variants on different devices");
    if (ctxs[v]->fp32() != ctxs[0]->fp32()) return fail("multi_create: variants mix FP32 and FP64");
  }
  FdtdHipTuning t{0, 0, 0, CPML_SEPARATE};
  if (tuning) {
    t = *tuning;
    if (check_tuning(ctxs[0], &t, "multi_create")) return FDTD_HIP_ERR;
    if (t.fuse_cpml == CPML_SPLIT_Z || t.fuse_cpml == CPML_SPLIT_ALL)
      return fail("multi_create: CPML placements 3 and 4 (split slabs) are single-problem only");
  }
  HIP_TRY(hipSetDevice(ctxs[0]->device));
  FdtdHipMulti *m = new FdtdHipMulti();
  m->ctx.assign(ctxs, ctxs + n);
  m->device = ctxs[0]->device;
  m->fp32 = ctxs[0]->fp32();
  m->tune = t;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&m->ev, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&m->d_step, sizeof(int)) != hipSuccess) {
    fdtd_hip_multi_destroy(m);
    return fail("multi_create: stream, event or step counter allocation failed");
  }
  *out = m;
  return FDTD_HIP_OK;
}

int fdtd_hip_multi_run_async(FdtdHipMulti *m, int tsStart, int count) {
  if (!m) return fail("multi_run: null batch");
  return multi_run(m, tsStart, count, m->use_graphs != 0);
}

int fdtd_hip_multi_set_graphs(FdtdHipMulti *m, int enable) {
  if (!m) return fail("multi_set_graphs: null batch");
  m->use_graphs = enable ? 1 : 0;
  return FDTD_HIP_OK;
}

int fdtd_hip_multi_synchronize(FdtdHipMulti *m) {
  if (!m) return fail("multi_synchronize: null batch");
  HIP_TRY(hipStreamSynchronize(m->stream));
  return FDTD_HIP_OK;
}

int fdtd_hip_multi_get_tuning(FdtdHipMulti *m, FdtdHipTuning *out) {
  if (!m || !out) return fail("multi_get_tuning: null argument");
  *out = m->tune;
  return FDTD_HIP_OK;
}

int fdtd_hip_multi_autotune(FdtdHipMulti *m, int steps_per_candidate, FdtdHipTuning *chosen) {
  if (!m) return fail("multi_autotune: null batch");
  for (FdtdHipContext *c : m->ctx)
    if (c->steps_run) return fail("multi_autotune: must run before the first time step (it resets the state)");
  if (steps_per_candidate < 1) steps_per_candidate = 2;
  int steps = steps_per_candidate + 1;
  for (FdtdHipContext *c : m->ctx) steps = std::min(steps, c->d.numberOfTimeSteps);
  HIP_TRY(hipSetDevice(m->device));
  // Per-cell coefficient arrays (storage 1) read 9 values per cell and phase where the
  // 8-bit indices read 3 bytes, for the same arithmetic; they are timed only when some
  // variant's tables need wider indices.
  if (build_tables_threaded(m->ctx, true)) return FDTD_HIP_ERR;
  bool all_u8 = true;
  for (FdtdHipContext *c : m->ctx) all_u8 = all_u8 && c->coef_idx_bytes == 1;
  std::vector<FdtdHipTuning> cands;
  for (int cs = 0; cs < (all_u8 ? 1 : 2); ++cs)
    for (int w : {1, 2, 4}) {
      if (w == 4 && !m->fp32) continue;
      for (int pl : {1, 2, 4, 8})
        for (int fc : {CPML_SEPARATE, CPML_FUSE_Z, CPML_FUSE_ALL, CPML_FUSE_XY})
          cands.push_back(FdtdHipTuning{pl, w, cs, fc});
    }
  hipEvent_t ev[3] = {};
  for (hipEvent_t &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      for (hipEvent_t x : ev) if (x) (void)hipEventDestroy(x);
      return fail("multi_autotune: event creation failed");
    }
  auto run_steps = [m](int ts, int n) {
    return m->fp32 ? multi_steps<float>(m, ts, n, false) : multi_steps<double>(m, ts, n, false);
  };
  double best = 1e300;       // ms per timed step
  FdtdHipTuning best_t = m->tune;
  // The variants' streams are lent once for the whole search.
  int rc = multi_borrow(m);
  for (size_t q = 0; !rc && q < cands.size(); ++q) {
    const FdtdHipTuning &t = cands[q];
    m->tune = t;
    if ((rc = m->fp32 ? multi_prepare<float>(m) : multi_prepare<double>(m))) break;
    // Step 0 warms up (and is timed too: a candidate whose warm-up step takes more than
    // 1.5 times the best timed step so far cannot win and is not timed further); steps
    // 1.. are timed.
    float warm = 0.f, ms = 0.f;
    if (hipEventRecord(ev[0], m->stream) != hipSuccess || (rc = run_steps(0, 1)) ||
        hipEventRecord(ev[1], m->stream) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
        hipEventElapsedTime(&warm, ev[0], ev[1]) != hipSuccess) {
      if (!rc) rc = fail("multi_autotune: event timing failed");
      break;
    }
    double per = warm;
    if (steps >= 2) {
      if (warm > 1.5 * best) continue;
      if (hipEventRecord(ev[0], m->stream) != hipSuccess) { rc = fail("multi_autotune: event record failed"); break; }
      if ((rc = run_steps(1, steps - 1))) break;
      if (hipEventRecord(ev[2], m->stream) != hipSuccess || hipEventSynchronize(ev[2]) != hipSuccess ||
          hipEventElapsedTime(&ms, ev[0], ev[2]) != hipSuccess) { rc = fail("multi_autotune: event timing failed"); break; }
      per = ms / (steps - 1);
    }
    if (per < best) { best = per; best_t = t; }
  }
  if (multi_return(m) && !rc) rc = FDTD_HIP_ERR;
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  m->tune = best_t;
  for (FdtdHipContext *c : m->ctx) {
    c->steps_run = 0;
    if (!rc && reset_state(c)) rc = FDTD_HIP_ERR;
  }
  if (rc) return rc;
  if (chosen) *chosen = best_t;
  return FDTD_HIP_OK;
}

// ------------------------------------------------------------ early-stop network --

struct FdtdHipEarlyStop {
  hipStream_t stream = nullptr;
  uint16_t *w = nullptr;
  float *b = nullptr;
  float *act[3] = {};      // input, ping, pong
  int cap = 0;             // batch capacity of the activation buffers
};

static long long es_weight_count() {
  long long n = 0;
  int cin = ES_IN_CH;
  for (int l = 0; l < 4; ++l) { n += (long long)ES_CH * cin * ES_KW; cin = ES_CH; }
  return n + (long long)ES_HIDDEN * ES_CH * ES_LEN[4] + (long long)ES_OUT * ES_HIDDEN;
}

static long long es_bias_count() { return 4LL * ES_CH + ES_HIDDEN + ES_OUT; }

void fdtd_hip_es_destroy(FdtdHipEarlyStop *n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  (void)hipFree(n->w);
  (void)hipFree(n->b);
  for (float *a : n->act) (void)hipFree(a);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

int fdtd_hip_es_create(const unsigned short *weights, long long nweights, const float *biases, long long nbiases,
                       FdtdHipEarlyStop **out) {
  if (!weights || !biases || !out) return fail("es_create: null argument");
  if (nweights != es_weight_count()) return fail("es_create: %lld weights, expected %lld", nweights, es_weight_count());
  if (nbiases != es_bias_count()) return fail("es_create: %lld biases, expected %lld", nbiases, es_bias_count());
  FdtdHipEarlyStop *n = new FdtdHipEarlyStop;
  *out = nullptr;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&n->w, nweights * 2) != hipSuccess || hipMalloc(&n->b, nbiases * 4) != hipSuccess ||
      hipMemcpy(n->w, weights, nweights * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(n->b, biases, nbiases * 4, hipMemcpyHostToDevice) != hipSuccess) {
    fdtd_hip_es_destroy(n);
    return fail("es_create: device allocation or copy failed");
  }
  *out = n;
  return FDTD_HIP_OK;
}

int fdtd_hip_es_predict(FdtdHipEarlyStop *n, int batch, const float *x, float *y) {
  if (!n || !x || !y) return fail("es_predict: null argument");
  if (batch <= 0) return fail("es_predict: batch %d", batch);
  const long in_sz = (long)ES_IN_CH * ES_LEN[0];
  const long mid_sz = std::max<long>((long)ES_CH * ES_LEN[1], ES_OUT);
  if (batch > n->cap) {
    for (float *&a : n->act) { (void)hipFree(a); a = nullptr; }
    if (hipMalloc(&n->act[0], (size_t)batch * in_sz * 4) != hipSuccess ||
        hipMalloc(&n->act[1], (size_t)batch * mid_sz * 4) != hipSuccess ||
        hipMalloc(&n->act[2], (size_t)batch * mid_sz * 4) != hipSuccess) {
      n->cap = 0;
      return fail("es_predict: device allocation failed");
    }
    n->cap = batch;
  }
  HIP_TRY(hipMemcpyAsync(n->act[0], x, (size_t)batch * in_sz * 4, hipMemcpyHostToDevice, n->stream));
  const uint16_t *w = n->w;
  const float *b = n->b;
  float *src = n->act[0], *dst = n->act[1];
  int cin = ES_IN_CH;
  for (int l = 0; l < 4; ++l) {
    const long outs = (long)batch * ES_CH * ES_LEN[l + 1];
    hipLaunchKernelGGL(k_es_conv, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0, n->stream, src, w, b, dst,
                       batch, cin, ES_LEN[l], ES_LEN[l + 1]);
    w += (long)ES_CH * cin * ES_KW;
    b += ES_CH;
    cin = ES_CH;
    src = dst;
    dst = dst == n->act[1] ? n->act[2] : n->act[1];
  }
  const int K1 = ES_CH * ES_LEN[4];
  hipLaunchKernelGGL(k_es_dense, dim3((unsigned)(((long)batch * ES_HIDDEN * 64 + 255) / 256)), dim3(256), 0, n->stream,
                     src, w, b, dst, batch, K1, ES_HIDDEN, 1);
  w += (long)ES_HIDDEN * K1;
  b += ES_HIDDEN;
  src = dst;
  dst = dst == n->act[1] ? n->act[2] : n->act[1];
  hipLaunchKernelGGL(k_es_dense, dim3((unsigned)(((long)batch * ES_OUT * 64 + 255) / 256)), dim3(256), 0, n->stream,
                     src, w, b, dst, batch, ES_HIDDEN, ES_OUT, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(y, dst, (size_t)batch * ES_OUT * 4, hipMemcpyDeviceToHost, n->stream));
  HIP_TRY(hipStreamSynchronize(n->stream));
  return FDTD_HIP_OK;
}

}  // extern "C"
