// Host check of the CPML pass's work decomposition (csrc/cpml_tiles.hpp), built with
// g++ by test_cpml_tiles.py.  For each layout, phase and CPML placement it decodes
// every lane of every tile exactly as the kernels do and checks, against a
// brute-force enumeration of the grid, that each cell the pass owns is visited once,
// with the mask of exactly the faces whose slab holds it, and nothing else; likewise
// the z-slab march lanes of the split placement (check_zslab).
//
//   cpml_tiles_check            prints one line per case, then OK (exit 0) or FAIL
#include "../special-topics_amd/csrc/cpml_tiles.hpp"

#include <cstdio>
#include <random>
#include <vector>

using namespace fdtd;

static unsigned faces_of(const CpmlLayout &L, int phase, const int co[3]) {
  unsigned m = 0;
  for (int f = 0; f < 6; ++f) {
    const int lo = phase == 0 ? L.m_start[f] : L.e_start[f];
    if (L.nc[f] && co[f / 2] >= lo && co[f / 2] < lo + L.nc[f]) m |= 1u << f;
  }
  return m;
}

static int check(const CpmlLayout &L, const char *name) {
  int bad = 0;
  const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
  const long plane = (long)(ny + 1) * (nz + 1);
  for (int mode : {CPML_SEPARATE, CPML_FUSE_Z, CPML_FUSE_ALL, CPML_FUSE_XY})
    for (int ph = 0; ph < 2; ++ph) {
      std::vector<CpmlTile> tiles;
      build_cpml_tiles(L, ph, mode, tiles);
      std::vector<unsigned char> seen((size_t)(nx + 1) * plane, 0);
      long lanes = 0, cells = 0;
      for (const CpmlTile &t : tiles) {
        for (int lane = 0; lane < 256; ++lane) {
          unsigned mask = 0;
          int i = 0, j = 0, k = 0;
          if (!cpml_tile_cell(t, lane, mask, i, j, k)) continue;
          ++lanes;
          if (i < 0 || i > nx || j < 0 || j > ny || k < 0 || k > nz) {
            if (bad++ < 5) std::printf("  %s mode %d phase %d: lane outside the grid (%d,%d,%d)\n", name, mode, ph, i, j, k);
            continue;
          }
          const int co[3] = {i, j, k};
          // CPML_FUSE_XY: the pass applies only the z-face terms.
          const unsigned want = mode == CPML_FUSE_XY ? faces_of(L, ph, co) & (3u << 4) : faces_of(L, ph, co);
          if (mask != want && bad++ < 5)
            std::printf("  %s mode %d phase %d: mask %x at (%d,%d,%d), expected %x\n", name, mode, ph, mask, i, j, k,
                        want);
          if (seen[i * plane + (long)j * (nz + 1) + k]++ && bad++ < 5)
            std::printf("  %s mode %d phase %d: (%d,%d,%d) visited twice\n", name, mode, ph, i, j, k);
          // The E pass changes component c where its tile mask holds a face of one of c's
          // transverse axes; cpml_e_touches must say so exactly.
          for (int comp = 0; ph == 1 && mode == CPML_SEPARATE && comp < 3; ++comp) {
            const bool term = (mask & ~(3u << (2 * comp))) != 0;
            if (term != cpml_e_touches(L, comp, i, j, k) && bad++ < 5)
              std::printf("  %s: cpml_e_touches(E%c, %d,%d,%d) = %d\n", name, "xyz"[comp], i, j, k, !term);
          }
        }
      }
      const int i_end = ph == 0 ? nx + 1 : nx;
      for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
          for (int k = 0; k <= nz; ++k) {
            const int co[3] = {i, j, k};
            const unsigned m = faces_of(L, ph, co);
            bool owned = m != 0;
            if (mode == CPML_FUSE_ALL) owned = false;
            if (mode == CPML_FUSE_Z && (m & (3u << 4)) && i < i_end) owned = false;
            if (mode == CPML_FUSE_XY) owned = (m & (3u << 4)) != 0;
            cells += owned;
            if (ph == 1 && !m)
              for (int comp = 0; comp < 3; ++comp)
                if (cpml_e_touches(L, comp, i, j, k) && bad++ < 5)
                  std::printf("  %s: cpml_e_touches(E%c, %d,%d,%d) outside every slab\n", name, "xyz"[comp], i, j, k);
            if (owned != (seen[i * plane + (long)j * (nz + 1) + k] != 0) && bad++ < 5)
              std::printf("  %s mode %d phase %d: (%d,%d,%d) owned=%d but visited=%d\n", name, mode, ph, i, j, k, owned,
                          seen[i * plane + (long)j * (nz + 1) + k]);
          }
      std::printf("%s mode %d phase %d: %zu tiles, %ld cells, lane use %.3f\n", name, mode, ph, tiles.size(), cells,
                  tiles.empty() ? 1.0 : (double)lanes / (256.0 * tiles.size()));
    }
  return bad;
}

// CPML_SPLIT_Z: the z-slab march workgroups of each phase visit every (i, j, k) with k
// in a z slab of the phase, j in 0..ny and i in their planes (H: 0..nx, E: 0..nx-1)
// exactly once, for several x chunk depths; together with the bulk kernels' cells
// outside the z slabs and the FUSE_Z tile list this is FUSE_Z's decomposition.
static int check_zslab(const CpmlLayout &L, const char *name) {
  int bad = 0;
  const int nx = L.n[0], ny = L.n[1], nz = L.n[2];
  for (int ph = 0; ph < 2; ++ph)
    for (int chunk : {1, 2, 3, 8}) {
      const int planes = ph == 0 ? nx + 1 : nx;
      const ZslabGeom z = zslab_geom(L, ph, planes, chunk);
      std::vector<unsigned char> seen((size_t)(nx + 1) * (ny + 1) * (nz + 1), 0);
      if (z.ntiles % 8) { std::printf("  %s: %d z-slab tiles, not a multiple of 8\n", name, z.ntiles); ++bad; }
      for (int Lw = 0; Lw < z.ntiles; ++Lw) {
        int col = 0, ch = 0;
        if (!zslab_tile(z, Lw, col, ch)) continue;
        for (int lane = 0; lane < 256; ++lane) {
          int j = 0, k = 0;
          if (!zslab_lane(z, col * 256 + lane, j, k)) continue;
          for (int i = ch * z.chunk; i < std::min(ch * z.chunk + z.chunk, planes); ++i) {
            if (j < 0 || j > ny || k < 0 || k > nz) {
              if (bad++ < 5) std::printf("  %s zslab phase %d: lane outside (%d,%d,%d)\n", name, ph, i, j, k);
              continue;
            }
            if (seen[((size_t)i * (ny + 1) + j) * (nz + 1) + k]++ && bad++ < 5)
              std::printf("  %s zslab phase %d: (%d,%d,%d) twice\n", name, ph, i, j, k);
          }
        }
      }
      long cells = 0;
      for (int i = 0; i <= nx; ++i)
        for (int j = 0; j <= ny; ++j)
          for (int k = 0; k <= nz; ++k) {
            const int co[3] = {i, j, k};
            const bool want = (faces_of(L, ph, co) & (3u << 4)) && i < planes;
            cells += want;
            if (want != (seen[((size_t)i * (ny + 1) + j) * (nz + 1) + k] != 0) && bad++ < 5)
              std::printf("  %s zslab phase %d chunk %d: (%d,%d,%d) want=%d\n", name, ph, chunk, i, j, k, want);
          }
      if (chunk == 2)
        std::printf("%s zslab phase %d: %d tiles, %ld cells, lane use %.3f\n", name, ph, z.ntiles, cells,
                    (double)z.lanes * planes / (256.0 * z.ncol * planes));
    }
  return bad;
}

// Slabs as the reference's builder places them: H planes from 0 and up to n, E planes
// from 1 and up to n (npml cells each), optionally with some faces absent.
static CpmlLayout standard(int nx, int ny, int nz, int npml, unsigned absent) {
  CpmlLayout L{};
  L.n[0] = nx; L.n[1] = ny; L.n[2] = nz;
  for (int f = 0; f < 6; ++f) {
    if (absent & (1u << f)) continue;
    const int n = L.n[f / 2];
    L.nc[f] = npml;
    L.m_start[f] = f % 2 == 0 ? 0 : n - npml;
    L.e_start[f] = f % 2 == 0 ? 1 : n - npml;
  }
  return L;
}

// e_tail_disjoint on constructed cases, one per condition.
static int check_e_tail() {
  const CpmlLayout L = standard(40, 40, 40, 6, 0);     // E slabs [1,7) and [34,40) on every axis
  const int lo[3] = {10, 10, 10}, hi[3] = {30, 30, 30};
  struct Case { const char *what; std::vector<ECell> w, r; bool ntff; int lo0; bool want; };
  const Case cases[] = {
      {"interior source and observer", {{2, 20, 20, 20}}, {{2, 20, 20, 21}}, true, 10, true},
      {"no NTFF, interior ports", {{0, 20, 20, 20}}, {}, false, 10, true},
      {"source in a z slab (Ex has z terms)", {{0, 20, 20, 3}}, {}, false, 10, false},
      {"Ez in an x slab", {{2, 36, 20, 20}}, {}, false, 10, false},
      {"Ez in a z slab only (no z terms)", {{2, 20, 20, 3}}, {}, false, 10, true},
      {"observer in a y slab", {}, {{0, 20, 38, 20}}, false, 10, false},
      {"NTFF box reaching an x slab", {{2, 20, 20, 20}}, {}, true, 6, false},
      {"source on the NTFF surface", {{1, 10, 20, 20}}, {}, true, 10, false},
      {"source on a surface edge", {{1, 30, 30, 20}}, {}, true, 10, false},
      {"source outside the NTFF box", {{1, 32, 20, 20}}, {}, true, 10, true},
      {"observer on the NTFF surface (reads only)", {}, {{1, 10, 20, 20}}, true, 10, true},
  };
  int bad = 0;
  for (const Case &c : cases) {
    const int l[3] = {c.lo0, lo[1], lo[2]};
    const bool got = e_tail_disjoint(L, c.w, c.r, c.ntff, l, hi);
    std::printf("e_tail_disjoint: %s: %d\n", c.what, got);
    if (got != c.want) { std::printf("  expected %d\n", c.want); ++bad; }
  }
  return bad;
}

// Tile order at 256^3, separate pass: each plane's chunk of low-z tiles (face mask zn
// only) is followed directly by the same plane's chunk of high-z tiles (zp only), one
// tile per XCD each, so the two tiles that read a row's last and the next row's first
// cells run on one XCD together.  Checked over the y-interior rows, whose chunks hold
// eight tiles (the single-row piece at j = n has one-tile chunks).
static int check_z_interleave() {
  const CpmlLayout L = standard(256, 256, 256, 8, 0);
  int bad = 0;
  for (int ph = 0; ph < 2; ++ph) {
    std::vector<CpmlTile> t;
    build_cpml_tiles(L, ph, CPML_SEPARATE, t);
    int pairs = 0;
    for (size_t q = 0; q + 8 < t.size(); ++q) {
      const unsigned m = (t[q].x >> 16) & 63u;
      if (m != (1u << 4)) continue;                          // a zn-only tile
      if ((t[q].y & 0xffffu) < 8 || (t[q].y >> 16) > 248) continue;   // full chunks: y-interior rows
      const CpmlTile &u = t[q + 8];
      const bool same_rows = (u.x & 0xffffu) == (t[q].x & 0xffffu) && u.y == t[q].y;
      if (((u.x >> 16) & 63u) != (1u << 5) || !same_rows) {
        if (bad++ < 5) std::printf("  phase %d: tile %zu (zn, plane %u) is not followed 8 later by its zp tile\n", ph,
                                   q, t[q].x & 0xffffu);
      } else {
        ++pairs;
      }
    }
    std::printf("256^3 phase %d: %d zn/zp tile pairs 8 apart\n", ph, pairs);
    if (pairs == 0) ++bad;
  }
  return bad;
}

int main() {
  int bad = check_e_tail();
  bad += check_z_interleave();
  bad += check(standard(64, 72, 80, 8, 0), "64x72x80 npml 8");
  bad += check(standard(39, 136, 136, 8, 0), "39x136x136 npml 8");
  bad += check(standard(21, 17, 70, 4, 1u << 3), "21x17x70 npml 4 no yp");
  bad += check(standard(9, 11, 13, 2, (1u << 0) | (1u << 5)), "9x11x13 npml 2 no xn, zp");
  bad += check(standard(12, 12, 12, 0, 63u), "12^3 no CPML");
  // The largest axis the 16-bit tile coordinates hold (planes 0..n with n = 65534).
  bad += check(standard(CPML_TILE_MAX_AXIS, 4, 5, 1, 0), "65534x4x5 npml 1");
  bad += check(standard(4, 5, CPML_TILE_MAX_AXIS, 1, 0), "4x5x65534 npml 1");
  // Seeded random layouts within the bounds upload_cpml accepts (2 nc < n, H planes in
  // [0, n), E planes in [1, n)); slabs may sit anywhere, even overlap.
  std::mt19937 rng(7);
  for (int r = 0; r < 40; ++r) {
    CpmlLayout L{};
    for (int a = 0; a < 3; ++a) L.n[a] = 5 + (int)(rng() % 40);
    for (int f = 0; f < 6; ++f) {
      if (rng() % 4 == 0) continue;
      const int n = L.n[f / 2];
      const int nc = 1 + (int)(rng() % ((n - 1) / 2));
      L.nc[f] = nc;
      L.m_start[f] = (int)(rng() % (n - nc + 1));
      L.e_start[f] = 1 + (int)(rng() % (n - nc));
    }
    char name[64];
    std::snprintf(name, sizeof name, "random %d (%dx%dx%d)", r, L.n[0], L.n[1], L.n[2]);
    bad += check(L, name);
    bad += check_zslab(L, name);
  }
  bad += check_zslab(standard(64, 72, 80, 8, 0), "64x72x80 npml 8");
  bad += check_zslab(standard(39, 136, 136, 8, 0), "39x136x136 npml 8");
  bad += check_zslab(standard(256, 256, 256, 8, 0), "256^3 npml 8");
  bad += check_zslab(standard(21, 17, 70, 4, 1u << 5), "21x17x70 npml 4 no zp");
  bad += check_zslab(standard(12, 12, 12, 0, 63u), "12^3 no CPML");
  std::printf(bad ? "FAIL (%d)\n" : "OK\n", bad);
  return bad != 0;
}
