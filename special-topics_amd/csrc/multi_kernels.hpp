// One launch per phase over many problems ("variants") of uneven grid size.
//
// Each variant keeps its own device arrays (fields, coefficients, CPML psi, ports,
// NTFF).  Its descriptor -- the same Grid / FaceSet / MarchGeom / PortState /
// FarfieldState values the single-problem kernels take as arguments -- sits in a
// device array, and a per-workgroup map names the variant and the variant-local
// tile a workgroup owns.  The bodies are the single-problem ones, so every variant's
// result is bit-identical to running it alone with the same tuning.
//
// Stencil tiles of all variants are numbered variant-major, column-fastest within a
// variant, and dealt to the 8 XCDs in contiguous ranges exactly as march_tile deals
// one problem's tiles, so neighbouring columns of a variant still share an L2.  The
// CPML placement (separate pass, z slabs fused, all faces fused) is one choice for
// the whole batch; the separate pass is then one more batched launch per phase.
#pragma once

namespace fdtd {

template <typename T>
struct VariantDesc {
  Grid<T> g;
  FaceSet<T> fs;
  MarchGeom mh, me;
  const CpmlTile *th, *te;   // CPML pass tiles of each phase (none when every face is fused)
  int ghost;
  int has_ports;
  PortState ps;
  FarfieldState ff;
  NtffBlock<T> nb;
};

// Pointers read from a descriptor in memory are generic to the compiler, which then
// emits flat loads and stores.  Rebuilding each as a global-address-space pointer
// (from its integer value, a conversion the optimiser keeps) and converting that to
// a generic one lets address-space inference turn every access derived from it into a
// global one.
template <typename P>
__device__ __forceinline__ P *as_global(P *p) {
  typedef __attribute__((address_space(1))) P GP;
  return (P *)(GP *)(unsigned long long)p;
}

template <typename T>
__device__ __forceinline__ Grid<T> global_grid(const Grid<T> &in) {
  Grid<T> g = in;
  g.Hx = as_global(g.Hx); g.Hy = as_global(g.Hy); g.Hz = as_global(g.Hz);
  g.Ex = as_global(g.Ex); g.Ey = as_global(g.Ey); g.Ez = as_global(g.Ez);
#pragma unroll
  for (int w = 0; w < NCOEFF; ++w) g.C[w] = as_global(g.C[w]);
#pragma unroll
  for (int q = 0; q < 6; ++q) { g.ct.idx[q] = as_global(g.ct.idx[q]); g.ct.tab[q] = as_global(g.ct.tab[q]); }
  return g;
}

template <typename T>
__device__ __forceinline__ FaceSet<T> global_faces(const FaceSet<T> &in) {
  FaceSet<T> fs = in;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    Face<T> &F = fs.f[f];
    F.b_e = as_global(F.b_e); F.a_e = as_global(F.a_e); F.b_m = as_global(F.b_m); F.a_m = as_global(F.a_m);
#pragma unroll
    for (int s = 0; s < 4; ++s) { F.psi[s] = as_global(F.psi[s]); F.cpsi[s] = as_global(F.cpsi[s]); }
  }
  return fs;
}

// Map entry of a workgroup: variant in the high 12 bits, variant-local tile or block
// in the low 20.
constexpr int MULTI_LOCAL_BITS = 20;
constexpr unsigned MULTI_LOCAL_MASK = (1u << MULTI_LOCAL_BITS) - 1;
constexpr int MULTI_MAX_VARIANTS = 1 << (32 - MULTI_LOCAL_BITS);

// Stencil tile of this workgroup (XCD-dealt as in march_tile) -> map entry; false
// past the last tile.
__device__ __forceinline__ bool multi_tile(const unsigned *__restrict__ map, int ntiles, int per_xcd,
                                           unsigned &entry) {
  const int L = blockIdx.x;
  const int tile = (L & 7) * per_xcd + (L >> 3);
  if (tile >= ntiles) return false;
  entry = map[tile];
  return true;
}

template <typename T, int CM, int W, int FC>
__global__ void __launch_bounds__(256) k_multi_update_h(const VariantDesc<T> *__restrict__ vd,
                                                        const unsigned *__restrict__ map, int ntiles, int per_xcd) {
  __shared__ TableCache<T, CM> tc;
  unsigned entry;
  if (!multi_tile(map, ntiles, per_xcd, entry)) return;
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const int local = (int)(entry & MULTI_LOCAL_MASK);
  const int col = local % d.mh.ncol, ch = local / d.mh.ncol;
  const Grid<T> g = global_grid(d.g);
  const FaceSet<T> fs = global_faces(d.fs);
  if constexpr (early_tables<CM, FC>()) {
    TablePending<T, CM> tp;
    table_issue<T, CM>(g, 0, tp);
    const int p = (col * 256 + (int)threadIdx.x) * W;
    const int j = p / g.SJ;
    const bool active = j <= g.ny;                         // else plane padding
    update_h_cells<T, CM, W, FC, true, true>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, tc, active ? j : 0,
                                             active ? p - j * g.SJ : 0, ch * d.mh.chunk, d.mh.chunk, fs, &tp,
                                             active);
  } else {
    load_tables<T, CM>(g, 0, tc);
    update_h_body<T, CM, W, FC>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, d.mh, tc, col, ch, fs);
  }
}

template <typename T, int CM, int W, int FC>
__global__ void __launch_bounds__(256) k_multi_update_e(const VariantDesc<T> *__restrict__ vd,
                                                        const unsigned *__restrict__ map, int ntiles, int per_xcd) {
  __shared__ TableCache<T, CM> tc;
  unsigned entry;
  if (!multi_tile(map, ntiles, per_xcd, entry)) return;
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const int local = (int)(entry & MULTI_LOCAL_MASK);
  const int col = local % d.me.ncol, ch = local / d.me.ncol;
  const Grid<T> g = global_grid(d.g);
  const FaceSet<T> fs = global_faces(d.fs);
  if constexpr (early_tables<CM, FC>()) {
    TablePending<T, CM> tp;
    table_issue<T, CM>(g, 3, tp);
    const int p = (col * 256 + (int)threadIdx.x) * W;
    const int j = p / g.SJ;
    const bool active = j <= g.ny;                         // else plane padding
    update_e_cells<T, CM, W, FC, true, true>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, tc, active ? j : 0,
                                             active ? p - j * g.SJ : 0, ch * d.me.chunk, d.me.chunk, d.ghost, fs,
                                             &tp, active);
  } else {
    load_tables<T, CM>(g, 3, tc);
    update_e_body<T, CM, W, FC>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, d.me, tc, col, ch, d.ghost, fs);
  }
}

// The CPML pass: one workgroup per tile of the variant's tile list.
template <typename T>
__global__ void __launch_bounds__(256) k_multi_cpml_h(const VariantDesc<T> *__restrict__ vd,
                                                      const unsigned *__restrict__ map) {
  const unsigned entry = map[blockIdx.x];
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const CpmlTile t = as_global(d.th)[entry & MULTI_LOCAL_MASK];
  cpml_h_tile<T>(global_grid(d.g), global_faces(d.fs), t, threadIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_multi_cpml_e(const VariantDesc<T> *__restrict__ vd,
                                                      const unsigned *__restrict__ map) {
  const unsigned entry = map[blockIdx.x];
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const CpmlTile t = as_global(d.te)[entry & MULTI_LOCAL_MASK];
  cpml_e_tile<T>(global_grid(d.g), global_faces(d.fs), t, threadIdx.x);
}

// One block per variant; variants without sources or observers return at once.
template <typename T>
__global__ void __launch_bounds__(256) k_multi_port_epilogue(const VariantDesc<T> *__restrict__ vd, const int *tsp,
                                                             int off) {
  const VariantDesc<T> &d = vd[blockIdx.x];
  if (!d.has_ports) return;
  port_epilogue_body<T>(global_grid(d.g), d.ps, step_of(tsp, off));
}

// NTFF surface work: each variant's threads padded to whole blocks; map entries give
// the variant and its local block.
template <typename T>
__global__ void __launch_bounds__(256) k_multi_ntff_gather(const VariantDesc<T> *__restrict__ vd,
                                                           const unsigned *__restrict__ map, int row) {
  const unsigned entry = map[blockIdx.x];
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const long q = (long)(entry & MULTI_LOCAL_MASK) * blockDim.x + threadIdx.x;
  ntff_gather_body<T>(global_grid(d.g), d.ff, d.nb, row, q);
}

template <typename T>
__global__ void __launch_bounds__(256) k_multi_ntff_direct(const VariantDesc<T> *__restrict__ vd,
                                                           const unsigned *__restrict__ map, const int *tsp, int off) {
  const unsigned entry = map[blockIdx.x];
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const long q = (long)(entry & MULTI_LOCAL_MASK) * blockDim.x + threadIdx.x;
  ntff_direct_body<T>(global_grid(d.g), d.ff, step_of(tsp, off), q);
}

// The batched E-phase tail in one launch, as k_e_tail for one problem: the port
// epilogue blocks (one per variant), then the NTFF blocks (NTFF = 1 direct, 2 gather)
// through their map, then the E CPML tiles through theirs.  Used only when every
// variant's E CPML pass, epilogue and NTFF touch disjoint field elements; variants
// never share memory.
template <typename T, int NTFF>
__global__ void __launch_bounds__(256) k_multi_e_tail(const VariantDesc<T> *__restrict__ vd, int nports,
                                                      const unsigned *__restrict__ fmap, int nntff,
                                                      const unsigned *__restrict__ cmap, int row,
                                                      const int *tsp, int off) {
  int b = (int)blockIdx.x;
  if (b < nports) {
    const VariantDesc<T> &d = vd[b];
    if (d.has_ports) port_epilogue_body<T>(global_grid(d.g), d.ps, step_of(tsp, off));
    return;
  }
  b -= nports;
  if (b < nntff) {
    const unsigned entry = fmap[b];
    const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
    const long q = (long)(entry & MULTI_LOCAL_MASK) * blockDim.x + threadIdx.x;
    if (NTFF == 1) ntff_direct_body<T>(global_grid(d.g), d.ff, step_of(tsp, off), q);
    if (NTFF == 2) ntff_gather_body<T>(global_grid(d.g), d.ff, d.nb, row, q);
    return;
  }
  const unsigned entry = cmap[b - nntff];
  const VariantDesc<T> &d = vd[entry >> MULTI_LOCAL_BITS];
  const CpmlTile t = as_global(d.te)[entry & MULTI_LOCAL_MASK];
  cpml_e_tile<T>(global_grid(d.g), global_faces(d.fs), t, threadIdx.x);
}

}  // namespace fdtd
