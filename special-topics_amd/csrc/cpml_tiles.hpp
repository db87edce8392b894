// Work decomposition of the separate CPML pass: which cells it owns for each CPML
// placement and how they are cut into workgroup tiles, and which E elements it may
// change (so the step's other E-phase work can share its launch).  Plain C++17 with
// no HIP dependency, so a host-only program can check it; the per-lane decode is
// constexpr, which device code may call as well.
//
// The pass covers every cell a CPML face of the phase acts on that the bulk kernels
// leave to it.  Each axis is cut at n, n + 1 and every face's slab bounds, so every
// product box of the pieces lies wholly inside or outside each face slab: its face
// set is one mask.  Boxes are cut into tiles of one 256-lane workgroup: 1 << ks lanes
// along the contiguous axis k, 1 << js rows along j and the remaining
// 256 >> (ks + js) planes along i.  A tile record holds
//   x = i0 | mask << 16 | ks << 22 | js << 26   (mask bit 2 * axis + side, n = 0, p = 1)
//   y = j0 | jend << 16,  z = k0 | kend << 16,  w = iend
// so a workgroup decodes its cells from one scalar load, and which faces and terms
// apply is uniform across it.
#pragma once
#include <algorithm>
#include <array>
#include <vector>

namespace fdtd {

// CPML placement (a tuning choice; results are identical): the separate k_cpml pass
// for every face, the z-face slabs fused into the bulk kernels, every face fused, or
// the z-face slab cells (bulk update and every CPML term) in workgroups of their own
// within the bulk kernels' launch -- FUSE_Z's split of the work, without z-slab cells
// in the bulk kernels' waves.  SPLIT_Z leaves the CPML pass the same cells as FUSE_Z.
// SPLIT_ALL makes each phase one launch: the stencil tiles leave every slab cell alone,
// and the CPML tiles of the separate pass's list, in workgroups of the same launch, give
// each of their cells its bulk update and then its CPML terms (the E wall plane goes
// with the cell of plane 0 at the same (j, k)).  FUSE_XY applies the x- and y-face
// terms inside the bulk kernels, right after each cell's bulk update (whole x planes
// and whole rows share them, so a wave seldom diverges), and leaves the pass only the
// z-face terms of the z-slab cells; z is the last face of every component's order.
constexpr int CPML_SEPARATE = 0, CPML_FUSE_Z = 1, CPML_FUSE_ALL = 2, CPML_SPLIT_Z = 3, CPML_SPLIT_ALL = 4,
              CPML_FUSE_XY = 5;

// The tile list (build_cpml_tiles mode) the CPML pass uses under a placement.
constexpr int cpml_tile_mode(int placement) {
  return placement == CPML_SPLIT_Z ? CPML_FUSE_Z : (placement == CPML_SPLIT_ALL ? CPML_SEPARATE : placement);
}

// Whether a separate CPML pass runs after the bulk kernels under a placement.
constexpr bool cpml_pass_separate(int placement) { return placement != CPML_SPLIT_ALL; }

struct CpmlTile {
  unsigned x, y, z, w;
};

// The cell of `lane` in tile t; false for lanes past the tile's box.
constexpr bool cpml_tile_cell(const CpmlTile &t, int lane, unsigned &mask, int &i, int &j, int &k) {
  const unsigned ks = (t.x >> 22) & 15u, js = (t.x >> 26) & 15u;
  mask = (t.x >> 16) & 63u;
  k = (int)(t.z & 0xffffu) + (lane & ((1 << ks) - 1));
  j = (int)(t.y & 0xffffu) + ((lane >> ks) & ((1 << js) - 1));
  i = (int)(t.x & 0xffffu) + (lane >> (ks + js));
  return k < (int)(t.z >> 16) && j < (int)(t.y >> 16) && i < (int)t.w;
}

// Grid size and the CPML faces (xn, xp, yn, yp, zn, zp; nc == 0 when absent) with the
// first H and E plane of each slab.
struct CpmlLayout {
  int n[3];
  int nc[6], m_start[6], e_start[6];
};

// Largest grid axis the 16-bit tile coordinates hold (planes 0..n).
constexpr int CPML_TILE_MAX_AXIS = 65534;

struct CpmlPiece {
  int lo, hi;
  unsigned mask;
};

// The pieces of axis a for phase (0 = H, 1 = E): [0, n + 1) cut at n and at every
// face slab bound, each piece with the mask bits of the faces whose slab holds it.
inline std::vector<CpmlPiece> cpml_pieces(const CpmlLayout &L, int phase, int a) {
  const int n = L.n[a];
  std::vector<int> cut = {0, n, n + 1};
  for (int s = 0; s < 2; ++s) {
    const int f = 2 * a + s;
    if (!L.nc[f]) continue;
    const int lo = phase == 0 ? L.m_start[f] : L.e_start[f];
    cut.push_back(lo);
    cut.push_back(lo + L.nc[f]);
  }
  std::sort(cut.begin(), cut.end());
  cut.erase(std::unique(cut.begin(), cut.end()), cut.end());
  std::vector<CpmlPiece> out;
  for (size_t q = 0; q + 1 < cut.size(); ++q) {
    CpmlPiece p{cut[q], cut[q + 1], 0u};
    for (int s = 0; s < 2; ++s) {
      const int f = 2 * a + s;
      const int lo = phase == 0 ? L.m_start[f] : L.e_start[f];
      if (L.nc[f] && lo <= p.lo && p.hi <= lo + L.nc[f]) p.mask |= 1u << f;
    }
    out.push_back(p);
  }
  return out;
}

// The tiles of the cells the CPML pass owns for one phase and placement: every cell
// of a face slab with CPML_SEPARATE; with CPML_FUSE_Z all but the z-slab cells of the
// planes the bulk kernels cover (H: all; E: i < nx); none with CPML_FUSE_ALL; with
// CPML_FUSE_XY the z-slab cells, with only their z-face bits in the mask.  Lanes
// go along k (the box's k extent rounded up to a power of two, at most 64), then rows
// along j (its j extent rounded up, up to what is left of 256), then planes along i.
// Workgroups are dealt round-robin to the 8 XCDs, and tiles that read the same lines
// should run on one XCD together.  A piece's tiles are listed in chunks of up to eight
// row groups (j tiles): for each group of eight, for each plane, for each k tile, one
// chunk.  In full chunks a row group's k tiles (which share the sectors at their
// boundaries) and its consecutive planes (which read each other's plane as the x
// neighbour) are eight positions apart, i.e. on one XCD.  A piece with fewer than
// eight row groups fills its chunks with the row groups of several planes, which keeps
// a row's k tiles on one XCD; its planes' x neighbours then sit on other XCDs, which
// matters only for x-face tiles, and those come from pieces with many row groups.  Within one (x, y) piece pair
// the chunks of the z pieces are interleaved: the low-z chunk of a plane is followed by
// its high-z chunk, and a row's high-z cells and the next row's low-z cells are a few
// dozen bytes apart in memory, so both tiles of a pair share the lines in one XCD's L2.
// At 256^3 the interleave took the H pass from 46.4 to 42.8 us and keeping a row's k
// tiles on one XCD to 39.7 us.  The grid axes must not exceed CPML_TILE_MAX_AXIS.
inline void build_cpml_tiles(const CpmlLayout &L, int phase, int mode, std::vector<CpmlTile> &out) {
  if (mode == CPML_FUSE_ALL) return;
  std::vector<CpmlPiece> P[3];
  for (int a = 0; a < 3; ++a) P[a] = cpml_pieces(L, phase, a);
  const int i_end = phase == 0 ? L.n[0] + 1 : L.n[0];
  auto log2_up = [](int v, int cap) {
    int s = 0;
    while ((1 << s) < v && s < cap) ++s;
    return s;
  };
  for (const auto &px : P[0])
    for (const auto &py : P[1]) {
      std::vector<std::vector<std::vector<CpmlTile>>> chunks;   // per z piece, per (group, plane, k tile)
      for (const auto &pz : P[2]) {
        unsigned mask = px.mask | py.mask | pz.mask;
        if (!mask) continue;
        if (mode == CPML_FUSE_Z && pz.mask && px.hi <= i_end) continue;
        if (mode == CPML_FUSE_XY) {
          if (!pz.mask) continue;
          mask = pz.mask;
        }
        const int ks = log2_up(pz.hi - pz.lo, 6);
        const int js = log2_up(py.hi - py.lo, 8 - ks);
        const int kt = 1 << ks, jt = 1 << js, it = 256 >> (ks + js);
        const int nbk = (pz.hi - pz.lo + kt - 1) / kt, nbj = (py.hi - py.lo + jt - 1) / jt;
        // Planes per chunk: one with eight or more row groups, else enough to fill it.
        const int per = nbj >= 8 ? 1 : 8 / nbj;
        chunks.emplace_back();
        for (int g0 = 0; g0 < nbj; g0 += 8)
          for (int i0 = px.lo; i0 < px.hi; i0 += it * per)
            for (int kb = 0; kb < nbk; ++kb) {
              std::vector<CpmlTile> chunk;
              for (int ib = i0; ib < std::min(i0 + it * per, px.hi); ib += it)
                for (int jb = g0; jb < std::min(g0 + 8, nbj); ++jb) {
                  const int j0 = py.lo + jb * jt, k0 = pz.lo + kb * kt;
                  const int j1 = std::min(j0 + jt, py.hi), k1 = std::min(k0 + kt, pz.hi);
                  const int i1 = std::min(ib + it, px.hi);
                  chunk.push_back(CpmlTile{(unsigned)ib | mask << 16 | (unsigned)ks << 22 | (unsigned)js << 26,
                                           (unsigned)j0 | (unsigned)j1 << 16, (unsigned)k0 | (unsigned)k1 << 16,
                                           (unsigned)i1});
                }
              chunks.back().push_back(std::move(chunk));
            }
      }
      size_t most = 0;
      for (const auto &c : chunks) most = std::max(most, c.size());
      for (size_t q = 0; q < most; ++q)
        for (const auto &c : chunks)
          if (q < c.size()) out.insert(out.end(), c[q].begin(), c[q].end());
    }
}

// Lanes of the z-slab march (CPML_SPLIT_Z) of one phase: the cells (j, k) of the two
// z-face slabs over rows 0..rows-1, as segments of hi_n + lo_n lanes: segment s holds
// row s-1's high-z slab cells, then row s's low-z slab cells, which lie a few dozen
// bytes apart in the field arrays.  Workgroups of 256 lanes march `chunk` x planes;
// tiles are numbered lane-group fastest and dealt to the 8 XCDs in contiguous ranges,
// like the bulk kernels' tiles.  ntiles is a multiple of 8.
struct ZslabGeom {
  int lo_start, lo_n, hi_start, hi_n;   // slab k ranges (n = 0 when the face is absent)
  int rows, lanes;                      // rows 0..rows-1; (rows + 1) * (lo_n + hi_n) lanes
  int ncol, chunk, nchunk, per_xcd, ntiles;
};

// The (j, k) of lane q; false for lanes past the rows.
constexpr bool zslab_lane(const ZslabGeom &z, int q, int &j, int &k) {
  const int per = z.lo_n + z.hi_n;
  if (per == 0 || q >= z.lanes) return false;
  const int seg = q / per, w = q - seg * per;
  if (w < z.hi_n) { j = seg - 1; k = z.hi_start + w; }
  else { j = seg; k = z.lo_start + (w - z.hi_n); }
  return j >= 0 && j < z.rows;
}

// The march tile (lane group, x chunk) workgroup L handles; false past the tiles.
constexpr bool zslab_tile(const ZslabGeom &z, int L, int &col, int &ch) {
  const int t = (L & 7) * z.per_xcd + (L >> 3);
  if (t >= z.ncol * z.nchunk) return false;
  col = t % z.ncol;
  ch = t / z.ncol;
  return true;
}

// Z-slab march geometry of phase (0 = H, 1 = E) over `planes` x planes.
inline ZslabGeom zslab_geom(const CpmlLayout &L, int phase, int planes, int chunk) {
  ZslabGeom z{};
  const int lo = 4, hi = 5;
  z.lo_n = L.nc[lo];
  z.lo_start = phase == 0 ? L.m_start[lo] : L.e_start[lo];
  z.hi_n = L.nc[hi];
  z.hi_start = phase == 0 ? L.m_start[hi] : L.e_start[hi];
  if (z.lo_n && z.hi_n && z.lo_start <= z.hi_start + z.hi_n && z.hi_start <= z.lo_start + z.lo_n) {
    // Overlapping or touching slabs: one range, so no cell is marched twice.
    const int a = std::min(z.lo_start, z.hi_start), b = std::max(z.lo_start + z.lo_n, z.hi_start + z.hi_n);
    z.lo_start = a; z.lo_n = b - a; z.hi_n = 0;
  }
  z.rows = L.n[1] + 1;
  z.lanes = (z.rows + 1) * (z.lo_n + z.hi_n);
  z.ncol = (z.lanes + 255) / 256;
  z.chunk = chunk > 0 ? chunk : 1;
  z.nchunk = (planes + z.chunk - 1) / z.chunk;
  z.per_xcd = (z.ncol * z.nchunk + 7) / 8;
  z.ntiles = z.lanes ? 8 * z.per_xcd : 0;
  return z;
}

// Whether the E CPML pass may change E component `comp` (0..2 = Ex..Ez) at (i, j, k):
// the coordinate along one of the component's two transverse axes lies in an E slab of
// that axis.  Conservative at the grid edges, where the component may not exist.
inline bool cpml_e_touches(const CpmlLayout &L, int comp, int i, int j, int k) {
  const int co[3] = {i, j, k};
  for (int a = 0; a < 3; ++a) {
    if (a == comp) continue;
    for (int s = 0; s < 2; ++s) {
      const int f = 2 * a + s;
      if (L.nc[f] && co[a] >= L.e_start[f] && co[a] < L.e_start[f] + L.nc[f]) return true;
    }
  }
  return false;
}

// An E element (component 0..2 = Ex..Ez, i, j, k).
typedef std::array<int, 4> ECell;

// Whether the E CPML pass, the port epilogue and the NTFF step of a time step touch
// disjoint E elements, so that they may run concurrently in one launch: no element the
// epilogue writes (sources) or reads (observers, probes) may change in the E CPML
// pass; with an NTFF box [lo, hi] (0-based planes, both ends included) the box may not
// reach into an E slab (its reads lie on the box surface), and no element the
// epilogue writes may lie on that surface.
inline bool e_tail_disjoint(const CpmlLayout &L, const std::vector<ECell> &writes, const std::vector<ECell> &reads,
                            bool ntff, const int lo[3], const int hi[3]) {
  for (const auto *v : {&writes, &reads})
    for (const ECell &e : *v)
      if (cpml_e_touches(L, e[0], e[1], e[2], e[3])) return false;
  if (!ntff) return true;
  for (int f = 0; f < 6; ++f) {
    const int a = f / 2;
    if (L.nc[f] && L.e_start[f] <= hi[a] && lo[a] < L.e_start[f] + L.nc[f]) return false;
  }
  for (const ECell &e : writes) {
    bool inside = true, on_face = false;
    for (int a = 0; a < 3; ++a) {
      const int v = e[1 + a];
      inside = inside && v >= lo[a] && v <= hi[a];
      on_face = on_face || v == lo[a] || v == hi[a];
    }
    if (inside && on_face) return false;
  }
  return true;
}

}  // namespace fdtd
