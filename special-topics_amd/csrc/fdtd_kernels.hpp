// Device code of the gfx950 Yee time step.
//
// Device layout.  All six field components and all 18 coefficient arrays share one
// "uniform" shape (nx+1) x (ny+1) x (nz+1), k fastest: element (i,j,k) of every
// component sits at i*SI + j*SJ + k with SJ = nz+1 and SI = (ny+1)*SJ rounded up to a
// multiple of 64 elements (so every 256-wide tile starts on a 128-byte line).  Each
// component occupies the sub-box of its own Yee shape; the other slots stay zero and
// are never written.  One zero guard plane precedes plane 0 and one follows plane nx.  One flat offset therefore addresses Hx..Ez and their coefficients at the
// same cell, the curl neighbours are at +-1, +-SJ, +-SI for every component, and the
// plane i = 0 of Hz / Ex coincides with the reference's flat layout of those arrays
// (both have rows of nz+1), which is what the ghost-plane reads need.
//
// Arithmetic.  The FP64 instantiation evaluates every update with the reference's
// expression tree and no contraction, (C0*F + C1*dA) + C2*dB, Psi = b*Psi + a*d,
// F + CPsi*Psi (js_fdtd/assembly/fdtd_kernels.ts:54-56, js_fdtd/src/cpml.js:470,484),
// so results are bit-identical to the reference's FP64 engine.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cpml_tiles.hpp"

namespace fdtd {

enum CoeffId {
  CEXE, CEXHZ, CEXHY, CEYE, CEYHX, CEYHZ, CEZE, CEZHY, CEZHX,
  CHXH, CHXEZ, CHXEY, CHYH, CHYEX, CHYEZ, CHZH, CHZEY, CHZEX, NCOEFF
};

// Deduplicated coefficients: per component (Hx,Hy,Hz,Ex,Ey,Ez) an index array in the
// uniform layout and a table of n (self, curl1, curl2) triples holding the exact
// values of the per-cell arrays.  Real grids have a few hundred distinct triples at
// most (materials x interface averages x CPML grading), so the per-cell coefficient
// traffic drops from three values to one 1- or 2-byte index.
struct CoefTables {
  const void *idx[6];
  const void *tab[6];
  int n[6];
};

// Coefficient storage modes (kernel template argument).
enum CoefMode { COEF_ARRAYS = 0, COEF_U8 = 1, COEF_U16 = 2 };
constexpr int U8_TABLE = 256;

template <typename T>
struct Grid {
  int nx, ny, nz;
  int SJ, SI;
  T *Hx, *Hy, *Hz, *Ex, *Ey, *Ez;
  const T *C[NCOEFF];
  CoefTables ct;
};

// LDS copy of the three tables a kernel uses (COEF_U8 only).
template <typename T, int CM>
struct TableCache {
  T v[CM == COEF_U8 ? 3 * 3 * U8_TABLE : 1];
};

template <typename T, int CM>
__device__ __forceinline__ void load_tables(const Grid<T> &g, int first, TableCache<T, CM> &tc) {
  if (CM != COEF_U8) return;
  // Components unrolled so every table index is a constant (the Grid stays in registers).
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const T *tab = (const T *)g.ct.tab[first + comp];
    const int n3 = 3 * g.ct.n[first + comp];
    for (int r = threadIdx.x; r < 3 * U8_TABLE; r += blockDim.x) tc.v[comp * 3 * U8_TABLE + r] = r < n3 ? tab[r] : T(0);
  }
  __syncthreads();
}

// The same fill in two halves: table_issue loads a thread's share of the three tables
// into registers, the stencil then issues its first plane's loads, and table_commit
// stores the share to LDS and waits at the workgroup barrier.  The barrier's fences
// cover LDS only, so the plane's loads stay in flight across it and the table's round
// trip overlaps theirs.
template <typename T, int CM>
struct TablePending {
  T v[CM == COEF_U8 ? 3 * 3 : 1];
};

template <typename T, int CM>
__device__ __forceinline__ void table_issue(const Grid<T> &g, int first, TablePending<T, CM> &tp) {
  if (CM != COEF_U8) return;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const T *tab = (const T *)g.ct.tab[first + comp];
    const int n3 = 3 * g.ct.n[first + comp];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int r = (int)threadIdx.x + 256 * q;
      tp.v[comp * 3 + q] = r < n3 ? tab[r] : T(0);
    }
  }
}

template <typename T, int CM>
__device__ __forceinline__ void table_commit(TableCache<T, CM> &tc, const TablePending<T, CM> &tp) {
  if (CM != COEF_U8) return;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
#pragma unroll
    for (int q = 0; q < 3; ++q) tc.v[comp * 3 * U8_TABLE + (int)threadIdx.x + 256 * q] = tp.v[comp * 3 + q];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Whether a stencil kernel overlaps its table fill with its first plane's loads: 8-bit
// tables in LDS and a placement whose stencil skips no cell (the separate pass or the
// x/y faces fused), so that every thread of the workgroup reaches the barrier at the
// same point of its first plane.  A workgroup marches only a few planes, so the fill's
// round trip was a large part of its life: at 256^3 the stencils went from 113 / 117 to
// 106 / 109 us, on the IFA from 11.5 / 10.8 to 9.9 / 9.5 us.
template <int CM, int FC>
constexpr bool early_tables() {
  return CM == COEF_U8 && (FC == CPML_SEPARATE || FC == CPML_FUSE_XY);
}

// Cross-lane neighbours by a DPP wavefront shift (an operand modifier of a VALU move, no
// memory access): with DPP_WAVE_SHL1 each lane receives x from the next lane, with
// DPP_WAVE_SHR1 from the previous one; the lane at the wavefront's end (63 / 0), whose
// neighbour is in another wavefront, keeps `edge`.
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;

template <int CTRL>
__device__ __forceinline__ float dpp_shift(float x, float edge) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(x), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_shift(double x, double edge) {
  const long long xi = __double_as_longlong(x), ei = __double_as_longlong(edge);
  const int lo = __builtin_amdgcn_update_dpp((int)ei, (int)xi, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ei >> 32), (int)(xi >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Whether a stencil whose lanes hold consecutive W-cell vectors of a row (ROW_LANES: the
// march, not the z-slab march of CPML_SPLIT_Z, whose lanes pair the ends of two rows)
// takes the k neighbour just past (H) or before (E) its vector from the adjacent lane
// instead of loading it.  A scalar load of it spans as many cache lines per
// wavefront as a vector load (the lanes are W elements apart) for a quarter of the bytes,
// and the stencils are bound by their memory instructions.  The adjacent lane covers that
// element whenever it is used: for H the element lies in the same row as a cell with
// k < nz, whose next lane is then active; for E the previous lane is active whenever this
// one is (threads leave the march from the end of the column).  The placements that hand
// slab cells to other workgroups let lanes skip planes, so they keep the loads.
template <int FC>
constexpr bool k_neighbour_by_shift() { return FC != CPML_SPLIT_Z && FC != CPML_SPLIT_ALL; }

// W consecutive cells of one row, loaded and stored as one W*sizeof(T)-byte access.
template <typename T, int W>
struct alignas(sizeof(T) * W) Vec {
  T v[W];
  __device__ __forceinline__ T &operator[](int e) { return v[e]; }
  __device__ __forceinline__ const T &operator[](int e) const { return v[e]; }
};

template <typename T, int W>
__device__ __forceinline__ Vec<T, W> ldv(const T *__restrict__ p) { return *(const Vec<T, W> *)p; }
template <typename T, int W>
__device__ __forceinline__ void stv(T *__restrict__ p, const Vec<T, W> &v) { *(Vec<T, W> *)p = v; }

// Store of W cells.  NT (streaming) marks the own field of a stencil as written once
// and read next by the other phase from HBM, so it does not evict the neighbour
// planes; problems whose whole state fits in the on-die caches use plain stores
// instead, so the next kernel finds the field there.
template <typename T, int W, bool NT = true>
__device__ __forceinline__ void stv_stream(T *__restrict__ p, const Vec<T, W> &v) {
  if (!NT) {
    stv<T, W>(p, v);
  } else if (W * sizeof(T) == 16) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 x;
    __builtin_memcpy(&x, &v, 16);
    __builtin_nontemporal_store(x, (f4 *)p);
  } else if (W * sizeof(T) == 8) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    f2 x;
    __builtin_memcpy(&x, &v, 8);
    __builtin_nontemporal_store(x, (f2 *)p);
  } else {
#pragma unroll
    for (int e = 0; e < W; ++e) __builtin_nontemporal_store(v[e], p + e);
  }
}

template <typename T, int W, bool NT = true>
__device__ __forceinline__ Vec<T, W> ldv_stream(const T *__restrict__ p) {
  Vec<T, W> r;
  if (!NT) {
    r = ldv<T, W>(p);
  } else if (W * sizeof(T) == 16) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 x = __builtin_nontemporal_load((const f4 *)p);
    __builtin_memcpy(&r, &x, 16);
  } else if (W * sizeof(T) == 8) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 x = __builtin_nontemporal_load((const f2 *)p);
    __builtin_memcpy(&r, &x, 8);
  } else {
#pragma unroll
    for (int e = 0; e < W; ++e) r[e] = __builtin_nontemporal_load(p + e);
  }
  return r;
}

// The coefficient indices of W cells as loaded (8- or 16-bit storage), decoded only in
// coeff_lookup, so that an index load issued ahead of its use is not waited for at once:
// the 8-bit indices of the W cells packed in raw[0] (byte e = cell e), or one 16-bit
// index per element.
template <int CM, int W>
struct CoefIdx {
  uint32_t raw[CM == COEF_U16 ? W : 1];
};

// Index loads of component `first + local` for the W cells starting at offset c:
// streaming (NT) or plain, as the fields.
template <int CM, int W, bool NT = true>
__device__ __forceinline__ CoefIdx<CM, W> coeff_index(const CoefTables &ct, int first, int local, int c) {
  auto ld = [](auto *q) { return NT ? __builtin_nontemporal_load(q) : *q; };
  CoefIdx<CM, W> r;
  if (CM == COEF_U8) {
    const uint8_t *ip = (const uint8_t *)ct.idx[first + local] + c;
    if (W == 4) r.raw[0] = ld((const uint32_t *)ip);
    else if (W == 2) r.raw[0] = ld((const uint16_t *)ip);
    else r.raw[0] = ld(ip);
  } else {
    const uint16_t *ip = (const uint16_t *)ct.idx[first + local] + c;
#pragma unroll
    for (int e = 0; e < (CM == COEF_U16 ? W : 1); ++e) r.raw[e] = ld(ip + e);
  }
  return r;
}

// The (self, curl1, curl2) coefficients of component `first + local` at the W indices:
// from the LDS copy (8-bit) or the table in global memory (16-bit).
template <typename T, int CM, int W>
__device__ __forceinline__ void coeff_lookup(const CoefTables &ct, const TableCache<T, CM> &tc, int first, int local,
                                             const CoefIdx<CM, W> &ix, Vec<T, W> &c0, Vec<T, W> &c1,
                                             Vec<T, W> &c2) {
  const T *t = CM == COEF_U8 ? tc.v + local * 3 * U8_TABLE : (const T *)ct.tab[first + local];
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const int id = 3 * (int)(CM == COEF_U8 ? (ix.raw[0] >> (8 * e)) & 0xff : ix.raw[CM == COEF_U16 ? e : 0]);
    c0[e] = t[id];
    c1[e] = t[id + 1];
    c2[e] = t[id + 2];
  }
}

// The (self, curl1, curl2) coefficients of component `first + local` for the W cells
// starting at offset c.
template <typename T, int CM, int W, bool NT = true>
__device__ __forceinline__ void coeffs(const CoefTables &ct, const TableCache<T, CM> &tc, int first, int local,
                                       int c, const T *__restrict__ a0, const T *__restrict__ a1,
                                       const T *__restrict__ a2, Vec<T, W> &c0, Vec<T, W> &c1, Vec<T, W> &c2) {
  if (CM == COEF_ARRAYS) {
    c0 = ldv_stream<T, W, NT>(a0 + c);
    c1 = ldv_stream<T, W, NT>(a1 + c);
    c2 = ldv_stream<T, W, NT>(a2 + c);
    return;
  }
  coeff_lookup<T, CM, W>(ct, tc, first, local, coeff_index<CM, W, NT>(ct, first, local, c), c0, c1, c2);
}

// One CPML face.  CPsi of slot s is read at element q & cmask[s]: cmask is 0 when every
// element of the slot holds the same value (a face whose slab is one material, as CPML
// slabs usually are), so all reads hit element 0 in cache instead of streaming the
// array; it is -1 otherwise.
template <typename T>
struct Face {
  int nc, m_start, e_start;
  int cmask[4];
  const T *b_e, *a_e, *b_m, *a_m;
  T *psi[4];
  const T *cpsi[4];
};

template <typename T>
struct FaceSet {
  Face<T> f[6];   // xn, xp, yn, yp, zn, zp; nc == 0 when absent
};

// Psi offset of (i,j,k) in slot ci of a face along AXIS; d1, d2 are the component's
// own j and k extents.
template <int AXIS>
__device__ __forceinline__ int psi_offset(int i, int j, int k, int ci, int nc, int d1, int d2) {
  if (AXIS == 0) return (ci * d1 + j) * d2 + k;
  if (AXIS == 1) return (i * nc + ci) * d2 + k;
  return (i * d1 + j) * nc + ci;
}

// Slot (0..nc-1) of face axis AXIS that (i,j,k) falls in, for the n face (s = 0) or
// the p face (s = 1); -1 when outside.
template <typename T, int AXIS, bool MAG>
__device__ __forceinline__ int cpml_slot(const FaceSet<T> &fs, int s, int i, int j, int k) {
  const Face<T> &F = fs.f[2 * AXIS + s];
  const int coord = AXIS == 0 ? i : (AXIS == 1 ? j : k);
  const int ci = coord - (MAG ? F.m_start : F.e_start);
  return (unsigned)ci < (unsigned)F.nc ? ci : -1;
}

// Psi recurrence and field correction of one face slot (cpml.js:470,484 order).
template <typename T, int AXIS, int SLOT, bool MAG>
__device__ __forceinline__ T cpml_step(const FaceSet<T> &fs, int s, int ci, T val, T diff,
                                       int i, int j, int k, int d1, int d2) {
  const Face<T> &F = fs.f[2 * AXIS + s];
  const int q = psi_offset<AXIS>(i, j, k, ci, F.nc, d1, d2);
  const T b = MAG ? F.b_m[ci] : F.b_e[ci];
  const T a = MAG ? F.a_m[ci] : F.a_e[ci];
  const T ps = b * F.psi[SLOT][q] + a * diff;
  F.psi[SLOT][q] = ps;
  return val + F.cpsi[SLOT][q & F.cmask[SLOT]] * ps;
}

// One component's CPML terms on the value v: faces of axis A1 (n then p) then of axis
// A2, each with its curl difference computed on demand.
// NOZ leaves out the z faces (CPML_FUSE_XY: the separate pass applies them after).
template <typename T, int A1, int S1, int A2, int S2, bool MAG, bool NOZ = false, typename D1, typename D2>
__device__ __forceinline__ T cpml_terms(const FaceSet<T> &fs, T v, int i, int j, int k, int d1, int d2,
                                        D1 diff1, D2 diff2) {
  constexpr bool SKIP2 = NOZ && A2 == 2;
  const int a0 = cpml_slot<T, A1, MAG>(fs, 0, i, j, k), a1 = cpml_slot<T, A1, MAG>(fs, 1, i, j, k);
  const int b0 = SKIP2 ? -1 : cpml_slot<T, A2, MAG>(fs, 0, i, j, k);
  const int b1 = SKIP2 ? -1 : cpml_slot<T, A2, MAG>(fs, 1, i, j, k);
  if (a0 >= 0 || a1 >= 0) {
    const T d = diff1();
    if (a0 >= 0) v = cpml_step<T, A1, S1, MAG>(fs, 0, a0, v, d, i, j, k, d1, d2);
    if (a1 >= 0) v = cpml_step<T, A1, S1, MAG>(fs, 1, a1, v, d, i, j, k, d1, d2);
  }
  if (b0 >= 0 || b1 >= 0) {
    const T d = diff2();
    if (b0 >= 0) v = cpml_step<T, A2, S2, MAG>(fs, 0, b0, v, d, i, j, k, d1, d2);
    if (b1 >= 0) v = cpml_step<T, A2, S2, MAG>(fs, 1, b1, v, d, i, j, k, d1, d2);
  }
  return v;
}

// The terms of the faces along axis A for the W cells of a stencil vector, for the two
// components P and Q of the phase whose terms include that axis (slots SP and SQ of the
// face; own extents p1 x p2 and q1 x q2).  Bit e of okp / okq says that cell e of P / Q
// takes its terms here.  Staged: every Psi, CPsi and profile load of the W cells and
// both components is issued first, then the recurrences, corrections and Psi stores
// follow.  (Per term, load-store-load would cost two dependent round trips, chained
// across cells: the compiler may not move a CPsi load above a Psi store.)
// A lane stages one side of the axis: the n side when one of its cells lies in the n
// slab, else the p side.  Its cells' Psi elements and (along z) profile entries are then
// consecutive, base + e.  p-side terms of a lane that staged the n side (a vector that
// reaches both slabs, or a cell in both: an axis shorter than its two slabs) follow,
// unstaged, after the n side's.  The components are independent, and each still takes
// its faces in the reference's order (cpml.js:456-1003) when the axes are applied x, then
// y, then z.
template <typename T, int W, int A, bool MAG, int SP, int SQ>
__device__ __forceinline__ void fused_axis_terms(const FaceSet<T> &fs, int i, int j, int k0, Vec<T, W> &vp,
                                                 Vec<T, W> &vq, const Vec<T, W> &dp, const Vec<T, W> &dq,
                                                 unsigned okp, unsigned okq, int p1, int p2, int q1, int q2) {
#pragma clang fp contract(off)
  unsigned inn = 0, inp = 0;
#pragma unroll
  for (int e = 0; e < W; ++e) {
    if (cpml_slot<T, A, MAG>(fs, 0, i, j, k0 + e) >= 0) inn |= 1u << e;
    if (cpml_slot<T, A, MAG>(fs, 1, i, j, k0 + e) >= 0) inp |= 1u << e;
  }
  const unsigned ok = okp | okq;
  if (!((inn | inp) & ok)) return;
  const bool s = !(inn & ok);                              // the side this lane stages
  const unsigned fast = (s ? inp : inn) & ok, slow = s ? 0u : inp & ok;
  const Face<T> &F = fs.f[2 * A + (s ? 1 : 0)];
  const int ci0 = (A == 0 ? i : (A == 1 ? j : k0)) - (MAG ? F.m_start : F.e_start);
  const int qp0 = psi_offset<A>(i, j, k0, ci0, F.nc, p1, p2), qq0 = psi_offset<A>(i, j, k0, ci0, F.nc, q1, q2);
  T *__restrict__ psP = F.psi[SP] + qp0;
  T *__restrict__ psQ = F.psi[SQ] + qq0;
  const int mp = F.cmask[SP], mq = F.cmask[SQ];
  const T *__restrict__ cpP = F.cpsi[SP] + (qp0 & mp);
  const T *__restrict__ cpQ = F.cpsi[SQ] + (qq0 & mq);
  const T *__restrict__ bb = (MAG ? F.b_m : F.b_e) + ci0;
  const T *__restrict__ aa = (MAG ? F.a_m : F.a_e) + ci0;
  constexpr int NB = A == 2 ? W : 1;                       // profile entries per lane
  T psp[W], cpp[W], psq[W], cpq[W], b[NB], a[NB];
#pragma unroll
  for (int e = 0; e < W; ++e) {
    if (!(fast >> e & 1u)) continue;
    if (A == 2 || !(fast & ((1u << e) - 1u))) {            // along z every cell, else the first
      b[A == 2 ? e : 0] = bb[A == 2 ? e : 0];
      a[A == 2 ? e : 0] = aa[A == 2 ? e : 0];
    }
    if (okp >> e & 1u) { psp[e] = psP[e]; cpp[e] = cpP[e & mp]; }
    if (okq >> e & 1u) { psq[e] = psQ[e]; cpq[e] = cpQ[e & mq]; }
  }
#pragma unroll
  for (int e = 0; e < W; ++e) {
    if (!(fast >> e & 1u)) continue;
    const T be = b[A == 2 ? e : 0], ae = a[A == 2 ? e : 0];
    if (okp >> e & 1u) {
      const T ps = be * psp[e] + ae * dp[e];
      psP[e] = ps;
      vp[e] = vp[e] + cpp[e] * ps;
    }
    if (okq >> e & 1u) {
      const T ps = be * psq[e] + ae * dq[e];
      psQ[e] = ps;
      vq[e] = vq[e] + cpq[e] * ps;
    }
  }
  if (slow) {
#pragma unroll
    for (int e = 0; e < W; ++e) {
      if (!(slow >> e & 1u)) continue;
      const int cq = cpml_slot<T, A, MAG>(fs, 1, i, j, k0 + e);
      if (okp >> e & 1u) vp[e] = cpml_step<T, A, SP, MAG>(fs, 1, cq, vp[e], dp[e], i, j, k0 + e, p1, p2);
      if (okq >> e & 1u) vq[e] = cpml_step<T, A, SQ, MAG>(fs, 1, cq, vq[e], dq[e], i, j, k0 + e, q1, q2);
    }
  }
}

// Whether coordinate v along AXIS lies in a slab of that axis's faces (MAG: H planes,
// else E planes).
template <typename T, int AXIS, bool MAG>
__device__ __forceinline__ bool in_slab(const FaceSet<T> &fs, int v) {
  return cpml_slot<T, AXIS, MAG>(fs, 0, v, v, v) >= 0 || cpml_slot<T, AXIS, MAG>(fs, 1, v, v, v) >= 0;
}

// Whether the bulk kernel applies the CPML terms of (i,j,k) itself.
template <typename T, int FC, bool MAG>
__device__ __forceinline__ bool fused_cpml_cell(const FaceSet<T> &fs, int i, int j, int k) {
  if (FC == CPML_FUSE_Z) return in_slab<T, 2, MAG>(fs, k);
  if (FC == CPML_FUSE_ALL)
    return in_slab<T, 0, MAG>(fs, i) || in_slab<T, 1, MAG>(fs, j) || in_slab<T, 2, MAG>(fs, k);
  if (FC == CPML_FUSE_XY) return in_slab<T, 0, MAG>(fs, i) || in_slab<T, 1, MAG>(fs, j);
  return false;
}

template <typename T>
struct CpmlDiff {
  const T *P;
  int hi, lo;
};

template <typename T>
struct CpmlStage {
  T v, u1, w1, u2, w2;
  T ps[4], cp[4];
  T b[4], a[4];     // staged 1-D profile values (FP32 only; see cpml_stage_profile)
};

// Whether the 1-D profile values b and a are staged with the other loads or read at
// store time.  They come from arrays of nc entries that stay in cache.  In FP32 the
// kernel's SGPRs set its occupancy, so staging them costs nothing and saves a round
// trip (measured: reading them at store time made the FP32 pass 5% slower).  In FP64
// staging them holds 16 more VGPRs per component, and reading them at store time takes
// the pass from 133 VGPRs (3 waves per SIMD) to 89 (5 waves), 6% faster.
template <typename T>
constexpr bool cpml_stage_profile() { return sizeof(T) == 4; }

template <typename T, int A, int SLOT, bool MAG>
__device__ __forceinline__ void cpml_load_face(const FaceSet<T> &fs, int s, int i, int j, int k, int d1, int d2,
                                               T &ps, T &cp, T &b, T &a) {
  const Face<T> &F = fs.f[2 * A + s];
  const int ci = (A == 0 ? i : (A == 1 ? j : k)) - (MAG ? F.m_start : F.e_start);
  const int q = psi_offset<A>(i, j, k, ci, F.nc, d1, d2);
  ps = F.psi[SLOT][q];
  cp = F.cpsi[SLOT][q & F.cmask[SLOT]];
  if (cpml_stage_profile<T>()) {
    b = MAG ? F.b_m[ci] : F.b_e[ci];
    a = MAG ? F.a_m[ci] : F.a_e[ci];
  }
}

template <typename T, int A, int SLOT, bool MAG>
__device__ __forceinline__ T cpml_store_face(const FaceSet<T> &fs, int s, int i, int j, int k, int d1, int d2,
                                             T v, T diff, T ps, T cp, T b, T a) {
  const Face<T> &F = fs.f[2 * A + s];
  const int ci = (A == 0 ? i : (A == 1 ? j : k)) - (MAG ? F.m_start : F.e_start);
  if (!cpml_stage_profile<T>()) {
    b = MAG ? F.b_m[ci] : F.b_e[ci];
    a = MAG ? F.a_m[ci] : F.a_e[ci];
  }
  const T p = b * ps + a * diff;
  F.psi[SLOT][psi_offset<A>(i, j, k, ci, F.nc, d1, d2)] = p;
  return v + cp * p;
}

// Psi, CPsi (and in FP32 the profile values) of the faces in `mask` of one component
// whose terms come from the faces of axes A1 and A2.
template <typename T, int A1, int S1, int A2, int S2, bool MAG>
__device__ __forceinline__ void cpml_stage_faces(const FaceSet<T> &fs, unsigned mask, int i, int j, int k, int d1,
                                                 int d2, CpmlStage<T> &L) {
  if (mask & (1u << (2 * A1))) cpml_load_face<T, A1, S1, MAG>(fs, 0, i, j, k, d1, d2, L.ps[0], L.cp[0], L.b[0], L.a[0]);
  if (mask & (2u << (2 * A1))) cpml_load_face<T, A1, S1, MAG>(fs, 1, i, j, k, d1, d2, L.ps[1], L.cp[1], L.b[1], L.a[1]);
  if (mask & (1u << (2 * A2))) cpml_load_face<T, A2, S2, MAG>(fs, 0, i, j, k, d1, d2, L.ps[2], L.cp[2], L.b[2], L.a[2]);
  if (mask & (2u << (2 * A2))) cpml_load_face<T, A2, S2, MAG>(fs, 1, i, j, k, d1, d2, L.ps[3], L.cp[3], L.b[3], L.a[3]);
}

// The staged faces' Psi recurrences (curl differences e1 along A1, e2 along A2) and
// field corrections applied to v, in the reference's order; returns the field value.
template <typename T, int A1, int S1, int A2, int S2, bool MAG>
__device__ __forceinline__ T cpml_apply_faces(const FaceSet<T> &fs, unsigned mask, int i, int j, int k, int d1, int d2,
                                              T v, T e1, T e2, const CpmlStage<T> &L) {
#pragma clang fp contract(off)
  if (mask & (1u << (2 * A1)))
    v = cpml_store_face<T, A1, S1, MAG>(fs, 0, i, j, k, d1, d2, v, e1, L.ps[0], L.cp[0], L.b[0], L.a[0]);
  if (mask & (2u << (2 * A1)))
    v = cpml_store_face<T, A1, S1, MAG>(fs, 1, i, j, k, d1, d2, v, e1, L.ps[1], L.cp[1], L.b[1], L.a[1]);
  if (mask & (1u << (2 * A2)))
    v = cpml_store_face<T, A2, S2, MAG>(fs, 0, i, j, k, d1, d2, v, e2, L.ps[2], L.cp[2], L.b[2], L.a[2]);
  if (mask & (2u << (2 * A2)))
    v = cpml_store_face<T, A2, S2, MAG>(fs, 1, i, j, k, d1, d2, v, e2, L.ps[3], L.cp[3], L.b[3], L.a[3]);
  return v;
}

// Face bits of each axis in a tile mask.
constexpr unsigned CPML_X = 3u, CPML_Y = 3u << 2, CPML_Z = 3u << 4;

// Launch geometry shared by the two stencil kernels.  A thread updates W consecutive
// cells of one row (rows are padded to a multiple of 4 elements, so the W cells never
// straddle a row); a workgroup's 256 threads cover 256*W consecutive plane offsets
// (a "column"; planes are padded to a multiple of 1024 elements, so every column
// starts on a 128-byte line and all vector accesses, including the +-SJ and +-SI
// neighbours, are aligned).  A tile is one column x `chunk` x-planes; the workgroup
// marches through its planes keeping the x-neighbour values it needs next in
// registers.  Tiles are numbered column-fastest and dealt to the 8 XCDs in contiguous
// ranges (workgroup L runs on XCD L % 8), so the rows a tile reads from its
// neighbour columns were fetched into the same XCD's L2 by the tile next to it.
struct MarchGeom {
  int ncol;       // columns per plane
  int nchunk;     // x chunks
  int chunk;      // planes per chunk
  int per_xcd;    // tiles per XCD
};

__device__ __forceinline__ bool march_tile_at(const MarchGeom &m, int L, int &col, int &ch) {
  const int tile = (L & 7) * m.per_xcd + (L >> 3);
  if (tile >= m.ncol * m.nchunk) return false;
  col = tile % m.ncol;
  ch = tile / m.ncol;
  return true;
}

__device__ __forceinline__ bool march_tile(const MarchGeom &m, int &col, int &ch) {
  return march_tile_at(m, blockIdx.x, col, ch);
}

// Bit e set when cell k0 + e of a W-cell vector lies in a z-face slab of the phase
// (MAG: H slabs, else E slabs): the cells a CPML_SPLIT_Z bulk kernel leaves to the
// z-slab workgroups.
template <typename T, int W, bool MAG>
__device__ __forceinline__ unsigned zslab_bits(const FaceSet<T> &fs, int k0) {
  unsigned b = 0;
#pragma unroll
  for (int e = 0; e < W; ++e)
    if (in_slab<T, 2, MAG>(fs, k0 + e)) b |= 1u << e;
  return b;
}

// The cells of row j (W cells from k0) that a stencil tile of placement FC leaves to
// other workgroups, before the x-plane is known: CPML_SPLIT_Z leaves z-slab cells,
// CPML_SPLIT_ALL every cell of a y or z slab (x slabs: split_plane_bits).
template <typename T, int W, int FC, bool MAG>
__device__ __forceinline__ unsigned split_row_bits(const FaceSet<T> &fs, int j, int k0) {
  if (FC == CPML_SPLIT_ALL && in_slab<T, 1, MAG>(fs, j)) return (1u << W) - 1;
  if (FC == CPML_SPLIT_Z || FC == CPML_SPLIT_ALL) return zslab_bits<T, W, MAG>(fs, k0);
  return 0u;
}

// The same for plane i: CPML_SPLIT_ALL also leaves every cell of an x-slab plane.
template <typename T, int W, int FC, bool MAG>
__device__ __forceinline__ unsigned split_plane_bits(const FaceSet<T> &fs, unsigned row_bits, int i) {
  if (FC == CPML_SPLIT_ALL && in_slab<T, 0, MAG>(fs, i)) return (1u << W) - 1;
  return row_bits;
}

// The field and coefficient arrays as no-alias pointers: the kernel bodies below take
// them as __restrict__ parameters so the compiler may issue the next plane's loads
// ahead of this plane's stores.
template <typename T>
struct Fields {
  T *__restrict__ Hx; T *__restrict__ Hy; T *__restrict__ Hz;
  T *__restrict__ Ex; T *__restrict__ Ey; T *__restrict__ Ez;
};

// H bulk update; W consecutive cells of a row per thread.  Hx over i in [0,nx], Hy/Hz
// over i in [0,nx).  The reference applies the CPML terms after the whole bulk update
// (js_fdtd/src/fdtdWorker.js:62-79); a cell's terms depend only on its own bulk value,
// its psi and E, so applying them to the cell right after its bulk update (FC != 0)
// gives the same result as the separate k_cpml pass.  Every array has a zero guard
// plane before plane 0 and after plane nx, so all neighbour loads are unconditional;
// elements outside a component's Yee shape are written back unchanged.
// ROW_LANES: the lanes hold consecutive W-cell vectors of a row (k_neighbour_by_shift).
// EARLY (early_tables): the kernel has issued its table loads into *tp and the LDS table
// is filled here, by table_commit, once the first plane's loads are out.  Every thread of
// the workgroup calls in, those past the plane's last row (active false) with j = k0 = 0
// so that their loads stay in bounds; they leave after the barrier.
template <typename T, int CM, int W, int FC, bool NT = true, bool EARLY = false, bool ROW_LANES = true>
__device__ __forceinline__ void update_h_cells(
    T *__restrict__ Hx, T *__restrict__ Hy, T *__restrict__ Hz,
    const T *__restrict__ Ex, const T *__restrict__ Ey, const T *__restrict__ Ez,
    const Grid<T> &g, TableCache<T, CM> &tc, int j, int k0, int i0, int chunk, const FaceSet<T> &fs,
    const TablePending<T, CM> *tp = nullptr, bool active = true) {
#pragma clang fp contract(off)
  typedef Vec<T, W> V;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int p = j * SJ + k0;
  const int i1 = min(i0 + chunk, nx + 1);
  constexpr unsigned ALL = (1u << W) - 1;
  constexpr bool FUSED = FC == CPML_FUSE_Z || FC == CPML_FUSE_ALL || FC == CPML_FUSE_XY;
  constexpr bool NOZ = FC == CPML_FUSE_XY;
  // CPML_SPLIT_Z / CPML_SPLIT_ALL: cells of the slabs these placements hand to other
  // workgroups of the launch are neither updated nor stored here.
  const unsigned row_skip = split_row_bits<T, W, FC, true>(fs, j, k0);
  if (row_skip == ALL) return;
  int c = i0 * SI + p;
  V ey_i = ldv<T, W>(Ey + c), ez_i = ldv<T, W>(Ez + c);   // E at plane i, carried forward
  for (int i = i0; i < i1; ++i, c += SI) {
    const bool inner = i < nx;
    const unsigned skip = split_plane_bits<T, W, FC, true>(fs, row_skip, i);
    const V ey_n = ldv<T, W>(Ey + c + SI), ez_n = ldv<T, W>(Ez + c + SI);
    if (skip == ALL) {
      ey_i = ey_n;
      ez_i = ez_n;
      continue;
    }
    const V ex_i = ldv<T, W>(Ex + c);
    const V ez_j = ldv<T, W>(Ez + c + SJ), ex_j = ldv<T, W>(Ex + c + SJ);
    // First cell past the vector: from the next lane (lane 63 loads it).
    constexpr bool SHIFT = ROW_LANES && k_neighbour_by_shift<FC>();
    const bool edge = !SHIFT || (threadIdx.x & 63) == 63;
    T ey_w = T(0), ex_w = T(0);
    if (edge) {
      ey_w = Ey[c + W];
      ex_w = Ex[c + W];
    }
    V hx = ldv_stream<T, W, NT>(Hx + c), hy = ldv_stream<T, W, NT>(Hy + c), hz = ldv_stream<T, W, NT>(Hz + c);
    V c0, c1, c2, c3, c4, c5, c6, c7, c8;
    if (EARLY) {
      const CoefIdx<CM, W> x0 = coeff_index<CM, W, NT>(g.ct, 0, 0, c), x1 = coeff_index<CM, W, NT>(g.ct, 0, 1, c),
                           x2 = coeff_index<CM, W, NT>(g.ct, 0, 2, c);
      if (i == i0) {
        table_commit<T, CM>(tc, *tp);
        if (!active) return;
      }
      coeff_lookup<T, CM, W>(g.ct, tc, 0, 0, x0, c0, c1, c2);
      coeff_lookup<T, CM, W>(g.ct, tc, 0, 1, x1, c3, c4, c5);
      coeff_lookup<T, CM, W>(g.ct, tc, 0, 2, x2, c6, c7, c8);
    } else {
      coeffs<T, CM, W, NT>(g.ct, tc, 0, 0, c, g.C[CHXH], g.C[CHXEY], g.C[CHXEZ], c0, c1, c2);
      coeffs<T, CM, W, NT>(g.ct, tc, 0, 1, c, g.C[CHYH], g.C[CHYEZ], g.C[CHYEX], c3, c4, c5);
      coeffs<T, CM, W, NT>(g.ct, tc, 0, 2, c, g.C[CHZH], g.C[CHZEX], g.C[CHZEY], c6, c7, c8);
    }
    if constexpr (SHIFT) {
      ey_w = dpp_shift<DPP_WAVE_SHL1>(ey_i[0], ey_w);
      ex_w = dpp_shift<DPP_WAVE_SHL1>(ex_i[0], ex_w);
    }
    // The six E differences of the curls; the CPML terms use the same ones (E does not
    // change during the H phase).
    V dy_hx, dz_hx, dx_hy, dz_hy, dx_hz, dy_hz;
#pragma unroll
    for (int e = 0; e < W; ++e) {
      const int k = k0 + e;
      const T ey_k = e + 1 < W ? ey_i[e + 1 < W ? e + 1 : 0] : ey_w;
      const T ex_k = e + 1 < W ? ex_i[e + 1 < W ? e + 1 : 0] : ex_w;
      dz_hx[e] = ey_k - ey_i[e];       // Hx: dEy along z
      dy_hx[e] = ez_j[e] - ez_i[e];    //     dEz along y
      dx_hy[e] = ez_n[e] - ez_i[e];    // Hy: dEz along x
      dz_hy[e] = ex_k - ex_i[e];       //     dEx along z
      dy_hz[e] = ex_j[e] - ex_i[e];    // Hz: dEx along y
      dx_hz[e] = ey_n[e] - ey_i[e];    //     dEy along x
      if (j < ny && k < nz) {                              // Hx (nx+1, ny, nz)
        T h = c0[e] * hx[e] + c1[e] * dz_hx[e] + c2[e] * dy_hx[e];
        hx[e] = h;
      }
      if (inner && k < nz) {                               // Hy (nx, ny+1, nz)
        T h = c3[e] * hy[e] + c4[e] * dx_hy[e] + c5[e] * dz_hy[e];
        hy[e] = h;
      }
      if (inner && j < ny && k <= nz) {                    // Hz (nx, ny, nz+1)
        T h = c6[e] * hz[e] + c7[e] * dy_hz[e] + c8[e] * dx_hz[e];
        hz[e] = h;
      }
    }
    if (FUSED) {
      // Cells the bulk kernel owns get their CPML terms here (CPML_FUSE_XY: those of the
      // x and y faces), right after their bulk update and in the reference's face order;
      // k_cpml skips them (CPML_FUSE_XY: applies the z-face terms after).  Hx takes its y
      // then z faces, Hy x then z, Hz x then y: one staged round per axis.
      unsigned okx = 0, oky = 0, okz = 0;
#pragma unroll
      for (int e = 0; e < W; ++e) {
        const int k = k0 + e;
        if (k > nz || !fused_cpml_cell<T, FC, true>(fs, i, j, k)) continue;
        if (j < ny && k < nz) okx |= 1u << e;
        if (inner && k < nz) oky |= 1u << e;
        if (inner && j < ny) okz |= 1u << e;
      }
      if (okx | oky | okz) {
        fused_axis_terms<T, W, 0, true, 0, 1>(fs, i, j, k0, hy, hz, dx_hy, dx_hz, oky, okz, ny + 1, nz, ny, nz + 1);
        fused_axis_terms<T, W, 1, true, 0, 1>(fs, i, j, k0, hx, hz, dy_hx, dy_hz, okx, okz, ny, nz, ny, nz + 1);
        if (!NOZ)
          fused_axis_terms<T, W, 2, true, 0, 1>(fs, i, j, k0, hx, hy, dz_hx, dz_hy, okx, oky, ny, nz, ny + 1, nz);
      }
    }
    if (skip) {
#pragma unroll
      for (int e = 0; e < W; ++e) {
        if (skip >> e & 1u) continue;
        Hx[c + e] = hx[e];
        if (inner) { Hy[c + e] = hy[e]; Hz[c + e] = hz[e]; }
      }
    } else {
      stv_stream<T, W, NT>(Hx + c, hx);
      if (inner) { stv_stream<T, W, NT>(Hy + c, hy); stv_stream<T, W, NT>(Hz + c, hz); }
    }
    ey_i = ey_n;
    ez_i = ez_n;
  }
}

// The stencil tile (col, ch) of the H bulk update: W consecutive cells of a row per
// thread, over the tile's x planes.
template <typename T, int CM, int W, int FC, bool NT = true>
__device__ __forceinline__ void update_h_body(
    T *__restrict__ Hx, T *__restrict__ Hy, T *__restrict__ Hz,
    const T *__restrict__ Ex, const T *__restrict__ Ey, const T *__restrict__ Ez,
    const Grid<T> &g, const MarchGeom &m, TableCache<T, CM> &tc, int col, int ch, const FaceSet<T> &fs) {
  const int p = (col * 256 + (int)threadIdx.x) * W;
  const int j = p / g.SJ;
  if (j > g.ny) return;                                    // plane padding
  update_h_cells<T, CM, W, FC, NT>(Hx, Hy, Hz, Ex, Ey, Ez, g, tc, j, p - j * g.SJ, ch * m.chunk, m.chunk, fs);
}

template <typename T, int CM, int W, int FC, bool NT = true>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, MarchGeom m, FaceSet<T> fs) {
  __shared__ TableCache<T, CM> tc;
  int col, ch;
  if (!march_tile(m, col, ch)) return;
  if constexpr (early_tables<CM, FC>()) {
    TablePending<T, CM> tp;
    table_issue<T, CM>(g, 0, tp);
    const int p = (col * 256 + (int)threadIdx.x) * W;
    const int j = p / g.SJ;
    const bool active = j <= g.ny;                         // else plane padding
    update_h_cells<T, CM, W, FC, NT, true>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, tc, active ? j : 0,
                                           active ? p - j * g.SJ : 0, ch * m.chunk, m.chunk, fs, &tp, active);
  } else {
    load_tables<T, CM>(g, 0, tc);
    update_h_body<T, CM, W, FC, NT>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, m, tc, col, ch, fs);
  }
}

// Bulk update of the i = nx wall's Ey/Ez at (j, k) under the reference boundary (see
// update_e_cells); ex0 is Ex(0, j, k) after this step's bulk update, before its CPML
// terms.
template <typename T>
__device__ __forceinline__ void e_wall_bulk(const Grid<T> &g, T *__restrict__ Ey, T *__restrict__ Ez,
                                            const T *__restrict__ Hx, const T *__restrict__ Hy,
                                            const T *__restrict__ Hz, int j, int k, T ex0) {
#pragma clang fp contract(off)
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int cg = nx * SI + j * SJ + k;
  const bool jin = j >= 1 && j < ny, kin = k >= 1 && k < nz;
  if (j < ny && kin) {                                   // Ey(nx, j, k)
    const T dHx = Hx[cg] - Hx[cg - 1];
    const T dHz = ex0 - Hz[cg - SI];
    T v = g.C[CEYE][cg] * Ey[cg] + g.C[CEYHX][cg] * dHx + g.C[CEYHZ][cg] * dHz;
    Ey[cg] = v;
  }
  if (jin && k < nz) {                                   // Ez(nx, j, k)
    const int f = j * nz + k;                            // reference flat index into Hz
    const int fr = f / (nz + 1);
    const T dHy = Hz[fr * SJ + (f - fr * (nz + 1))] - Hy[cg - SI];
    const T dHx = Hx[cg] - Hx[cg - SJ];
    T v = g.C[CEZE][cg] * Ez[cg] + g.C[CEZHY][cg] * dHy + g.C[CEZHX][cg] * dHx;
    Ez[cg] = v;
  }
}

// CPML terms of the i = nx wall's Ey/Ez at (j, k), after its bulk update (or none,
// under the matlab boundary), in the reference's face order (NOZ: without the z faces).
template <typename T, bool NOZ = false>
__device__ __forceinline__ void e_wall_terms(const Grid<T> &g, const FaceSet<T> &fs, T *__restrict__ Ey,
                                             T *__restrict__ Ez, const T *__restrict__ Hx, const T *__restrict__ Hy,
                                             const T *__restrict__ Hz, int j, int k) {
#pragma clang fp contract(off)
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int ce = nx * SI + j * SJ + k;
  if (k > nz || !fused_cpml_cell<T, NOZ ? CPML_FUSE_XY : CPML_FUSE_ALL, false>(fs, nx, j, k)) return;
  if (j < ny)                                            // Ey(nx): x then z faces
    Ey[ce] = cpml_terms<T, 0, 2, 2, 3, false, NOZ>(fs, Ey[ce], nx, j, k, ny, nz + 1,
        [&] { return Hz[ce] - Hz[ce - SI]; }, [&] { return Hx[ce] - Hx[ce - 1]; });
  if (k < nz)                                            // Ez(nx): x then y faces
    Ez[ce] = cpml_terms<T, 0, 3, 1, 2, false, NOZ>(fs, Ez[ce], nx, j, k, ny + 1, nz,
        [&] { return Hy[ce] - Hy[ce - SI]; }, [&] { return Hx[ce] - Hx[ce - SJ]; });
}

// E bulk update; W consecutive cells of a row per thread.  With ghost != 0 (reference
// boundary) the E planes are i in [0,nx) and the thread at i = 0 also updates Ey/Ez on
// the i = nx wall the way the reference's updateE does: Ey there differences
// "Hz[nx]", which is Ex(0,j,k) after this step's bulk update (the CPML terms come
// later, in k_cpml), and Ez differences "Hy[nx]", which is the reference-layout flat
// element j*nz+k of Hz.  With ghost == 0 the wall is left to the CPML terms.  With
// every face fused (FC == CPML_FUSE_ALL) the same thread then applies the wall's CPML
// terms, which read only H.  Elements the bulk does not update keep their value.
// EARLY, tp and active: as for update_h_cells.
template <typename T, int CM, int W, int FC, bool NT = true, bool EARLY = false, bool ROW_LANES = true>
__device__ __forceinline__ void update_e_cells(
    T *__restrict__ Ex, T *__restrict__ Ey, T *__restrict__ Ez,
    const T *__restrict__ Hx, const T *__restrict__ Hy, const T *__restrict__ Hz,
    const Grid<T> &g, TableCache<T, CM> &tc, int j, int k0, int i0, int chunk, int ghost,
    const FaceSet<T> &fs, const TablePending<T, CM> *tp = nullptr, bool active = true) {
#pragma clang fp contract(off)
  typedef Vec<T, W> V;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int p = j * SJ + k0;
  const int i1 = min(i0 + chunk, nx);
  constexpr unsigned ALL = (1u << W) - 1;
  constexpr bool FUSED = FC == CPML_FUSE_Z || FC == CPML_FUSE_ALL || FC == CPML_FUSE_XY;
  constexpr bool NOZ = FC == CPML_FUSE_XY;
  // CPML_SPLIT_Z / CPML_SPLIT_ALL: cells (wall cells included) of the slabs these
  // placements hand to other workgroups of the launch are left alone here.
  const unsigned row_skip = split_row_bits<T, W, FC, false>(fs, j, k0);
  if (row_skip == ALL) return;
  // The wall cells of this row go with the cells of plane 0.
  const unsigned skip0 = split_plane_bits<T, W, FC, false>(fs, row_skip, 0);
  const bool jin = j >= 1 && j < ny;
  int c = i0 * SI + p;
  V hz_m = ldv<T, W>(Hz + c - SI), hy_m = ldv<T, W>(Hy + c - SI);   // H at plane i-1
  V ex0 = V{};     // Ex at plane 0 after its bulk update, before any CPML term (the wall reads it)
  for (int i = i0; i < i1; ++i, c += SI) {
    const bool iin = i >= 1 && i < nx;
    const unsigned skip = split_plane_bits<T, W, FC, false>(fs, row_skip, i);
    const V hz_i = ldv<T, W>(Hz + c), hy_i = ldv<T, W>(Hy + c), hx_i = ldv<T, W>(Hx + c);
    if (skip == ALL) {
      hz_m = hz_i;
      hy_m = hy_i;
      continue;
    }
    const V hz_j = ldv<T, W>(Hz + c - SJ), hx_j = ldv<T, W>(Hx + c - SJ);
    // Last cell before the vector: from the previous lane (lane 0 loads it).
    constexpr bool SHIFT = ROW_LANES && k_neighbour_by_shift<FC>();
    const bool edge = !SHIFT || (threadIdx.x & 63) == 0;
    T hy_w = T(0), hx_w = T(0);
    if (edge) {
      hy_w = Hy[c - 1];
      hx_w = Hx[c - 1];
    }
    V ex = ldv_stream<T, W, NT>(Ex + c), ey = ldv_stream<T, W, NT>(Ey + c), ez = ldv_stream<T, W, NT>(Ez + c);
    V c0, c1, c2, c3, c4, c5, c6, c7, c8;
    if (EARLY) {
      const CoefIdx<CM, W> x0 = coeff_index<CM, W, NT>(g.ct, 3, 0, c), x1 = coeff_index<CM, W, NT>(g.ct, 3, 1, c),
                           x2 = coeff_index<CM, W, NT>(g.ct, 3, 2, c);
      if (i == i0) {
        table_commit<T, CM>(tc, *tp);
        if (!active) return;
      }
      coeff_lookup<T, CM, W>(g.ct, tc, 3, 0, x0, c0, c1, c2);
      coeff_lookup<T, CM, W>(g.ct, tc, 3, 1, x1, c3, c4, c5);
      coeff_lookup<T, CM, W>(g.ct, tc, 3, 2, x2, c6, c7, c8);
    } else {
      coeffs<T, CM, W, NT>(g.ct, tc, 3, 0, c, g.C[CEXE], g.C[CEXHZ], g.C[CEXHY], c0, c1, c2);
      coeffs<T, CM, W, NT>(g.ct, tc, 3, 1, c, g.C[CEYE], g.C[CEYHX], g.C[CEYHZ], c3, c4, c5);
      coeffs<T, CM, W, NT>(g.ct, tc, 3, 2, c, g.C[CEZE], g.C[CEZHY], g.C[CEZHX], c6, c7, c8);
    }
    if constexpr (SHIFT) {
      hy_w = dpp_shift<DPP_WAVE_SHR1>(hy_i[W - 1], hy_w);
      hx_w = dpp_shift<DPP_WAVE_SHR1>(hx_i[W - 1], hx_w);
    }
    // The six H differences of the curls, shared with the CPML terms (H does not change
    // during the E phase).
    V dy_ex, dz_ex, dz_ey, dx_ey, dx_ez, dy_ez;
#pragma unroll
    for (int e = 0; e < W; ++e) {
      const int k = k0 + e;
      const bool kin = k >= 1 && k < nz;
      const T hy_k = e > 0 ? hy_i[e > 0 ? e - 1 : 0] : hy_w;
      const T hx_k = e > 0 ? hx_i[e > 0 ? e - 1 : 0] : hx_w;
      dy_ex[e] = hz_i[e] - hz_j[e];    // Ex: dHz along y
      dz_ex[e] = hy_i[e] - hy_k;       //     dHy along z
      dz_ey[e] = hx_i[e] - hx_k;       // Ey: dHx along z
      dx_ey[e] = hz_i[e] - hz_m[e];    //     dHz along x
      dx_ez[e] = hy_i[e] - hy_m[e];    // Ez: dHy along x
      dy_ez[e] = hx_i[e] - hx_j[e];    //     dHx along y
      if (i < nx && k <= nz) {                             // Ex (nx, ny+1, nz+1)
        bool t = jin && kin;
        T v = ex[e];
        if (t) v = c0[e] * v + c1[e] * dy_ex[e] + c2[e] * dz_ex[e];
        ex[e] = v;
      }
      if (j < ny && k <= nz) {                             // Ey (nx+1, ny, nz+1)
        bool t = iin && kin;
        T v = ey[e];
        if (t) v = c3[e] * v + c4[e] * dz_ey[e] + c5[e] * dx_ey[e];
        ey[e] = v;
      }
      if (k < nz) {                                        // Ez (nx+1, ny+1, nz)
        bool t = iin && jin;
        T v = ez[e];
        if (t) v = c6[e] * v + c7[e] * dx_ez[e] + c8[e] * dy_ez[e];
        ez[e] = v;
      }
    }
    if (i == 0) ex0 = ex;
    if (FUSED) {
      // As in update_h_cells; ex0 keeps the pre-CPML Ex that the ghost plane reads.  Ex
      // takes its y then z faces, Ey x then z, Ez x then y.
      unsigned okx = 0, oky = 0, okz = 0;
#pragma unroll
      for (int e = 0; e < W; ++e) {
        const int k = k0 + e;
        if (k > nz || !fused_cpml_cell<T, FC, false>(fs, i, j, k)) continue;
        if (i < nx) okx |= 1u << e;
        if (j < ny) oky |= 1u << e;
        if (k < nz) okz |= 1u << e;
      }
      if (okx | oky | okz) {
        fused_axis_terms<T, W, 0, false, 2, 3>(fs, i, j, k0, ey, ez, dx_ey, dx_ez, oky, okz, ny, nz + 1, ny + 1, nz);
        fused_axis_terms<T, W, 1, false, 2, 3>(fs, i, j, k0, ez, ex, dy_ez, dy_ex, okz, okx, ny + 1, nz, ny + 1,
                                               nz + 1);
        if (!NOZ)
          fused_axis_terms<T, W, 2, false, 2, 3>(fs, i, j, k0, ex, ey, dz_ex, dz_ey, okx, oky, ny + 1, nz + 1, ny,
                                                 nz + 1);
      }
    }
    if (skip) {
#pragma unroll
      for (int e = 0; e < W; ++e) {
        if (skip >> e & 1u) continue;
        if (i < nx) Ex[c + e] = ex[e];
        Ey[c + e] = ey[e];
        Ez[c + e] = ez[e];
      }
    } else {
      if (i < nx) stv_stream<T, W, NT>(Ex + c, ex);
      stv_stream<T, W, NT>(Ey + c, ey);
      stv_stream<T, W, NT>(Ez + c, ez);
    }

    hz_m = hz_i;
    hy_m = hy_i;
  }
  // The i = nx wall, once per column by the chunk that holds plane 0.  Only this thread
  // writes the wall's Ey/Ez in this phase, and H does not change, so the wall may follow
  // the march; ex0 is Ex(0) as the bulk left it.  Kept out of the march loop, its
  // registers do not add to the loop's.  With CPML_FUSE_ALL / CPML_SPLIT_ALL the same
  // thread then applies the wall's CPML terms.
  constexpr bool WALL_TERMS = FC == CPML_FUSE_ALL || FC == CPML_SPLIT_ALL || FC == CPML_FUSE_XY;
  if ((ghost || WALL_TERMS) && i0 == 0) {
#pragma unroll
    for (int e = 0; e < W; ++e) {
      if (skip0 >> e & 1u) continue;
      if (ghost) e_wall_bulk<T>(g, Ey, Ez, Hx, Hy, Hz, j, k0 + e, ex0[e]);
      if (WALL_TERMS) e_wall_terms<T, FC == CPML_FUSE_XY>(g, fs, Ey, Ez, Hx, Hy, Hz, j, k0 + e);
    }
  }
}

// The stencil tile (col, ch) of the E bulk update (see update_h_body).
template <typename T, int CM, int W, int FC, bool NT = true>
__device__ __forceinline__ void update_e_body(
    T *__restrict__ Ex, T *__restrict__ Ey, T *__restrict__ Ez,
    const T *__restrict__ Hx, const T *__restrict__ Hy, const T *__restrict__ Hz,
    const Grid<T> &g, const MarchGeom &m, TableCache<T, CM> &tc, int col, int ch, int ghost,
    const FaceSet<T> &fs) {
  const int p = (col * 256 + (int)threadIdx.x) * W;
  const int j = p / g.SJ;
  if (j > g.ny) return;                                    // plane padding
  update_e_cells<T, CM, W, FC, NT>(Ex, Ey, Ez, Hx, Hy, Hz, g, tc, j, p - j * g.SJ, ch * m.chunk, m.chunk, ghost, fs);
}

template <typename T, int CM, int W, int FC, bool NT = true>
__global__ void __launch_bounds__(256) k_update_e(Grid<T> g, MarchGeom m, int ghost, FaceSet<T> fs) {
  __shared__ TableCache<T, CM> tc;
  int col, ch;
  if (!march_tile(m, col, ch)) return;
  if constexpr (early_tables<CM, FC>()) {
    TablePending<T, CM> tp;
    table_issue<T, CM>(g, 3, tp);
    const int p = (col * 256 + (int)threadIdx.x) * W;
    const int j = p / g.SJ;
    const bool active = j <= g.ny;                         // else plane padding
    update_e_cells<T, CM, W, FC, NT, true>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, tc, active ? j : 0,
                                           active ? p - j * g.SJ : 0, ch * m.chunk, m.chunk, ghost, fs, &tp,
                                           active);
  } else {
    load_tables<T, CM>(g, 3, tc);
    update_e_body<T, CM, W, FC, NT>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, m, tc, col, ch, ghost, fs);
  }
}

// CPML_SPLIT_Z: one launch of the bulk update in which workgroups [0, z.ntiles) march
// the z-slab cells (bulk update, then every CPML term, one cell per lane; the E wall's
// bulk update for those rows with the plane-0 value of Ex they computed) and the rest
// are the stencil tiles, which leave the z-slab cells alone.  The two sets write
// disjoint cells and read only the other field, so their order does not matter.
template <typename T, int CM, int W>
__global__ void __launch_bounds__(256) k_update_h_split(Grid<T> g, MarchGeom m, ZslabGeom z, FaceSet<T> fs) {
  __shared__ TableCache<T, CM> tc;
  const int L = blockIdx.x;
  int col, ch;
  if (L < z.ntiles) {
    if (!zslab_tile(z, L, col, ch)) return;
    load_tables<T, CM>(g, 0, tc);
    int j, k;
    if (!zslab_lane(z, col * 256 + (int)threadIdx.x, j, k)) return;
    update_h_cells<T, CM, 1, CPML_FUSE_Z, true, false, false>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, tc, j, k, ch * z.chunk, z.chunk, fs);
  } else {
    if (!march_tile_at(m, L - z.ntiles, col, ch)) return;
    load_tables<T, CM>(g, 0, tc);
    update_h_body<T, CM, W, CPML_SPLIT_Z>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, m, tc, col, ch, fs);
  }
}

template <typename T, int CM, int W>
__global__ void __launch_bounds__(256) k_update_e_split(Grid<T> g, MarchGeom m, ZslabGeom z, int ghost, FaceSet<T> fs) {
  __shared__ TableCache<T, CM> tc;
  const int L = blockIdx.x;
  int col, ch;
  if (L < z.ntiles) {
    if (!zslab_tile(z, L, col, ch)) return;
    load_tables<T, CM>(g, 3, tc);
    int j, k;
    if (!zslab_lane(z, col * 256 + (int)threadIdx.x, j, k)) return;
    update_e_cells<T, CM, 1, CPML_FUSE_Z, true, false, false>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, tc, j, k, ch * z.chunk, z.chunk,
                                          ghost, fs);
  } else {
    if (!march_tile_at(m, L - z.ntiles, col, ch)) return;
    load_tables<T, CM>(g, 3, tc);
    update_e_body<T, CM, W, CPML_SPLIT_Z>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, m, tc, col, ch, ghost, fs);
  }
}

// ------------------------------------------------------------------ CPML pass --

// The pass's tiles (CpmlTile, cpml_tile_cell) are described in cpml_tiles.hpp.

// One component's CPML work at one cell, staged: every load first (field, the
// operands of the curl differences, Psi, CPsi, profile), then the differences, the
// Psi recurrences and field corrections in the reference's order (faces of axis A1 n
// then p, then of A2; cpml.js:470,484) and the stores.  Issuing the loads of all
// three components before any arithmetic or store keeps them in flight together.
// The curl difference along an axis is P[c + hi] - P[c + lo].
template <typename T, int A1, int S1, int A2, int S2, bool MAG>
__device__ __forceinline__ void cpml_stage_load(const FaceSet<T> &fs, unsigned mask, const T *F, int c, int i, int j,
                                                int k, int d1, int d2, CpmlDiff<T> x1, CpmlDiff<T> x2,
                                                CpmlStage<T> &L) {
  L.v = F[c];
  if (mask & (3u << (2 * A1))) { L.u1 = x1.P[c + x1.hi]; L.w1 = x1.P[c + x1.lo]; }
  if (mask & (3u << (2 * A2))) { L.u2 = x2.P[c + x2.hi]; L.w2 = x2.P[c + x2.lo]; }
  cpml_stage_faces<T, A1, S1, A2, S2, MAG>(fs, mask, i, j, k, d1, d2, L);
}

template <typename T, int A1, int S1, int A2, int S2, bool MAG>
__device__ __forceinline__ void cpml_stage_store(const FaceSet<T> &fs, unsigned mask, T *F, int c, int i, int j,
                                                 int k, int d1, int d2, const CpmlStage<T> &L) {
#pragma clang fp contract(off)
  F[c] = cpml_apply_faces<T, A1, S1, A2, S2, MAG>(fs, mask, i, j, k, d1, d2, L.v, L.u1 - L.w1, L.u2 - L.w2, L);
}


// Magnetic CPML terms of one tile's H elements, faces in the reference's order per
// component (x, then y, then z; cpml.js:456-724).  Runs after the H bulk update, as
// the reference's updateMagneticCPML does; the E differences are this step's E,
// which no kernel changes in between.
template <typename T>
__device__ __forceinline__ void cpml_h_tile(const Grid<T> &g, const FaceSet<T> &fs, const CpmlTile &t, int lane) {
#pragma clang fp contract(off)
  unsigned mask;
  int i, j, k;
  if (!cpml_tile_cell(t, lane, mask, i, j, k)) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int c = i * SI + j * SJ + k;
  const T *Ex = g.Ex, *Ey = g.Ey, *Ez = g.Ez;
  const bool hx = (mask & (CPML_Y | CPML_Z)) && j < ny && k < nz;   // Hx: y then z faces
  const bool hy = (mask & (CPML_X | CPML_Z)) && i < nx && k < nz;   // Hy: x then z faces
  const bool hz = (mask & (CPML_X | CPML_Y)) && i < nx && j < ny;   // Hz: x then y faces
  CpmlStage<T> Lx, Ly, Lz;
  if (hx)
    cpml_stage_load<T, 1, 0, 2, 0, true>(fs, mask, g.Hx, c, i, j, k, ny, nz, {Ez, SJ, 0}, {Ey, 1, 0}, Lx);
  if (hy)
    cpml_stage_load<T, 0, 0, 2, 1, true>(fs, mask, g.Hy, c, i, j, k, ny + 1, nz, {Ez, SI, 0}, {Ex, 1, 0}, Ly);
  if (hz)
    cpml_stage_load<T, 0, 1, 1, 1, true>(fs, mask, g.Hz, c, i, j, k, ny, nz + 1, {Ey, SI, 0}, {Ex, SJ, 0}, Lz);
  if (hx) cpml_stage_store<T, 1, 0, 2, 0, true>(fs, mask, g.Hx, c, i, j, k, ny, nz, Lx);
  if (hy) cpml_stage_store<T, 0, 0, 2, 1, true>(fs, mask, g.Hy, c, i, j, k, ny + 1, nz, Ly);
  if (hz) cpml_stage_store<T, 0, 1, 1, 1, true>(fs, mask, g.Hz, c, i, j, k, ny, nz + 1, Lz);
}

// Electric CPML terms (cpml.js:727-1003), after the E bulk update.  Covers the whole
// E extents, so the i = nx wall of Ey/Ez gets its z / y terms here whichever
// boundary mode updated (or skipped) its bulk value.
template <typename T>
__device__ __forceinline__ void cpml_e_tile(const Grid<T> &g, const FaceSet<T> &fs, const CpmlTile &t, int lane) {
#pragma clang fp contract(off)
  unsigned mask;
  int i, j, k;
  if (!cpml_tile_cell(t, lane, mask, i, j, k)) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int c = i * SI + j * SJ + k;
  const T *Hx = g.Hx, *Hy = g.Hy, *Hz = g.Hz;
  const bool ex = (mask & (CPML_Y | CPML_Z)) && i < nx;              // Ex: y then z faces
  const bool ey = (mask & (CPML_X | CPML_Z)) && j < ny;              // Ey: x then z faces
  const bool ez = (mask & (CPML_X | CPML_Y)) && k < nz;              // Ez: x then y faces
  CpmlStage<T> Lx, Ly, Lz;
  if (ex)
    cpml_stage_load<T, 1, 3, 2, 2, false>(fs, mask, g.Ex, c, i, j, k, ny + 1, nz + 1, {Hz, 0, -SJ}, {Hy, 0, -1}, Lx);
  if (ey)
    cpml_stage_load<T, 0, 2, 2, 3, false>(fs, mask, g.Ey, c, i, j, k, ny, nz + 1, {Hz, 0, -SI}, {Hx, 0, -1}, Ly);
  if (ez)
    cpml_stage_load<T, 0, 3, 1, 2, false>(fs, mask, g.Ez, c, i, j, k, ny + 1, nz, {Hy, 0, -SI}, {Hx, 0, -SJ}, Lz);
  if (ex) cpml_stage_store<T, 1, 3, 2, 2, false>(fs, mask, g.Ex, c, i, j, k, ny + 1, nz + 1, Lx);
  if (ey) cpml_stage_store<T, 0, 2, 2, 3, false>(fs, mask, g.Ey, c, i, j, k, ny, nz + 1, Ly);
  if (ez) cpml_stage_store<T, 0, 3, 1, 2, false>(fs, mask, g.Ez, c, i, j, k, ny + 1, nz, Lz);
}

template <typename T>
__global__ void __launch_bounds__(256) k_cpml_h(Grid<T> g, FaceSet<T> fs, const CpmlTile *__restrict__ tiles) {
  cpml_h_tile<T>(g, fs, tiles[blockIdx.x], threadIdx.x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_cpml_e(Grid<T> g, FaceSet<T> fs, const CpmlTile *__restrict__ tiles) {
  cpml_e_tile<T>(g, fs, tiles[blockIdx.x], threadIdx.x);
}

// ------------------------------------------------- one launch per phase (SPLIT_ALL) --

// H cell (i, j, k) of a CPML tile under CPML_SPLIT_ALL: its bulk update (the
// expressions of update_h_cells), then the CPML terms of every face in the tile's
// mask (those of cpml_h_tile), one cell per lane.  All loads are issued first.
template <typename T, int CM>
__device__ __forceinline__ void h_cell_split_all(const Grid<T> &g, const FaceSet<T> &fs, const TableCache<T, CM> &tc,
                                                 unsigned mask, int i, int j, int k) {
#pragma clang fp contract(off)
  typedef Vec<T, 1> V;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  const int c = i * SI + j * SJ + k;
  const T *__restrict__ Ex = g.Ex;
  const T *__restrict__ Ey = g.Ey;
  const T *__restrict__ Ez = g.Ez;
  const bool inner = i < nx;
  const bool ux = j < ny && k < nz, uy = inner && k < nz, uz = inner && j < ny;   // k <= nz in a tile
  const bool px = ux && (mask & (CPML_Y | CPML_Z)), py = uy && (mask & (CPML_X | CPML_Z)),
             pz = uz && (mask & (CPML_X | CPML_Y));
  T hx = g.Hx[c], hy = g.Hy[c], hz = g.Hz[c];
  const T ex = Ex[c], ey = Ey[c], ez = Ez[c];
  const T ey_k = Ey[c + 1], ex_k = Ex[c + 1], ez_j = Ez[c + SJ], ex_j = Ex[c + SJ];
  const T ez_n = Ez[c + SI], ey_n = Ey[c + SI];
  V c0, c1, c2, c3, c4, c5, c6, c7, c8;
  coeffs<T, CM, 1>(g.ct, tc, 0, 0, c, g.C[CHXH], g.C[CHXEY], g.C[CHXEZ], c0, c1, c2);
  coeffs<T, CM, 1>(g.ct, tc, 0, 1, c, g.C[CHYH], g.C[CHYEZ], g.C[CHYEX], c3, c4, c5);
  coeffs<T, CM, 1>(g.ct, tc, 0, 2, c, g.C[CHZH], g.C[CHZEX], g.C[CHZEY], c6, c7, c8);
  CpmlStage<T> Lx, Ly, Lz;
  if (px) cpml_stage_faces<T, 1, 0, 2, 0, true>(fs, mask, i, j, k, ny, nz, Lx);
  if (py) cpml_stage_faces<T, 0, 0, 2, 1, true>(fs, mask, i, j, k, ny + 1, nz, Ly);
  if (pz) cpml_stage_faces<T, 0, 1, 1, 1, true>(fs, mask, i, j, k, ny, nz + 1, Lz);
  const T dz_hx = ey_k - ey, dy_hx = ez_j - ez;
  const T dx_hy = ez_n - ez, dz_hy = ex_k - ex;
  const T dy_hz = ex_j - ex, dx_hz = ey_n - ey;
  if (ux) hx = c0[0] * hx + c1[0] * dz_hx + c2[0] * dy_hx;
  if (uy) hy = c3[0] * hy + c4[0] * dx_hy + c5[0] * dz_hy;
  if (uz) hz = c6[0] * hz + c7[0] * dy_hz + c8[0] * dx_hz;
  if (px) hx = cpml_apply_faces<T, 1, 0, 2, 0, true>(fs, mask, i, j, k, ny, nz, hx, dy_hx, dz_hx, Lx);
  if (py) hy = cpml_apply_faces<T, 0, 0, 2, 1, true>(fs, mask, i, j, k, ny + 1, nz, hy, dx_hy, dz_hy, Ly);
  if (pz) hz = cpml_apply_faces<T, 0, 1, 1, 1, true>(fs, mask, i, j, k, ny, nz + 1, hz, dx_hz, dy_hz, Lz);
  if (ux) g.Hx[c] = hx;
  if (uy) g.Hy[c] = hy;
  if (uz) g.Hz[c] = hz;
}

// E cell (i, j, k) of a CPML tile under CPML_SPLIT_ALL (update_e_cells' bulk
// expressions, then cpml_e_tile's terms).  Cells of the i = nx wall are left to the
// cell of plane 0 at the same (j, k), which needs Ex(0) before its CPML terms: that
// lane does the wall's bulk update (reference boundary) and then its CPML terms.
template <typename T, int CM>
__device__ __forceinline__ void e_cell_split_all(const Grid<T> &g, const FaceSet<T> &fs, const TableCache<T, CM> &tc,
                                                 unsigned mask, int i, int j, int k, int ghost) {
#pragma clang fp contract(off)
  typedef Vec<T, 1> V;
  const int nx = g.nx, ny = g.ny, nz = g.nz, SJ = g.SJ, SI = g.SI;
  if (i >= nx) return;
  const int c = i * SI + j * SJ + k;
  const T *__restrict__ Hx = g.Hx;
  const T *__restrict__ Hy = g.Hy;
  const T *__restrict__ Hz = g.Hz;
  const bool iin = i >= 1, jin = j >= 1 && j < ny, kin = k >= 1 && k < nz;
  const bool ux = jin && kin, uy = j < ny && iin && kin, uz = k < nz && iin && jin;
  const bool px = mask & (CPML_Y | CPML_Z), py = j < ny && (mask & (CPML_X | CPML_Z)),
             pz = k < nz && (mask & (CPML_X | CPML_Y));
  T ex = g.Ex[c], ey = g.Ey[c], ez = g.Ez[c];
  const T hx = Hx[c], hy = Hy[c], hz = Hz[c];
  const T hz_j = Hz[c - SJ], hx_j = Hx[c - SJ], hy_k = Hy[c - 1], hx_k = Hx[c - 1];
  const T hz_m = Hz[c - SI], hy_m = Hy[c - SI];
  V c0, c1, c2, c3, c4, c5, c6, c7, c8;
  coeffs<T, CM, 1>(g.ct, tc, 3, 0, c, g.C[CEXE], g.C[CEXHZ], g.C[CEXHY], c0, c1, c2);
  coeffs<T, CM, 1>(g.ct, tc, 3, 1, c, g.C[CEYE], g.C[CEYHX], g.C[CEYHZ], c3, c4, c5);
  coeffs<T, CM, 1>(g.ct, tc, 3, 2, c, g.C[CEZE], g.C[CEZHY], g.C[CEZHX], c6, c7, c8);
  CpmlStage<T> Lx, Ly, Lz;
  if (px) cpml_stage_faces<T, 1, 3, 2, 2, false>(fs, mask, i, j, k, ny + 1, nz + 1, Lx);
  if (py) cpml_stage_faces<T, 0, 2, 2, 3, false>(fs, mask, i, j, k, ny, nz + 1, Ly);
  if (pz) cpml_stage_faces<T, 0, 3, 1, 2, false>(fs, mask, i, j, k, ny + 1, nz, Lz);
  const T dy_ex = hz - hz_j, dz_ex = hy - hy_k;
  const T dz_ey = hx - hx_k, dx_ey = hz - hz_m;
  const T dx_ez = hy - hy_m, dy_ez = hx - hx_j;
  if (ux) ex = c0[0] * ex + c1[0] * dy_ex + c2[0] * dz_ex;
  if (uy) ey = c3[0] * ey + c4[0] * dz_ey + c5[0] * dx_ey;
  if (uz) ez = c6[0] * ez + c7[0] * dx_ez + c8[0] * dy_ez;
  const T ex0 = ex;
  if (px) ex = cpml_apply_faces<T, 1, 3, 2, 2, false>(fs, mask, i, j, k, ny + 1, nz + 1, ex, dy_ex, dz_ex, Lx);
  if (py) ey = cpml_apply_faces<T, 0, 2, 2, 3, false>(fs, mask, i, j, k, ny, nz + 1, ey, dx_ey, dz_ey, Ly);
  if (pz) ez = cpml_apply_faces<T, 0, 3, 1, 2, false>(fs, mask, i, j, k, ny + 1, nz, ez, dx_ez, dy_ez, Lz);
  g.Ex[c] = ex;
  if (j < ny) g.Ey[c] = ey;
  if (k < nz) g.Ez[c] = ez;
  if (i == 0) {
    if (ghost) e_wall_bulk<T>(g, g.Ey, g.Ez, Hx, Hy, Hz, j, k, ex0);
    e_wall_terms<T>(g, fs, g.Ey, g.Ez, Hx, Hy, Hz, j, k);
  }
}

// CPML_SPLIT_ALL: one launch per phase.  Workgroups [0, ntiles) are the CPML tiles
// (bulk update and CPML terms of every slab cell), padded to a multiple of 8 so the
// stencil tiles after them keep their XCD placement; the stencil tiles leave every
// slab cell alone.  The two sets write disjoint cells and read only the other field.
template <typename T, int CM, int W>
__global__ void __launch_bounds__(256) k_update_h_all(Grid<T> g, MarchGeom m, FaceSet<T> fs,
                                                      const CpmlTile *__restrict__ tiles, int ntiles) {
  __shared__ TableCache<T, CM> tc;
  const int L = blockIdx.x, nt8 = (ntiles + 7) & ~7;
  if (L < nt8) {
    if (L >= ntiles) return;
    load_tables<T, CM>(g, 0, tc);
    unsigned mask;
    int i, j, k;
    if (!cpml_tile_cell(tiles[L], (int)threadIdx.x, mask, i, j, k)) return;
    h_cell_split_all<T, CM>(g, fs, tc, mask, i, j, k);
  } else {
    int col, ch;
    if (!march_tile_at(m, L - nt8, col, ch)) return;
    load_tables<T, CM>(g, 0, tc);
    update_h_body<T, CM, W, CPML_SPLIT_ALL>(g.Hx, g.Hy, g.Hz, g.Ex, g.Ey, g.Ez, g, m, tc, col, ch, fs);
  }
}

template <typename T, int CM, int W>
__global__ void __launch_bounds__(256) k_update_e_all(Grid<T> g, MarchGeom m, int ghost, FaceSet<T> fs,
                                                      const CpmlTile *__restrict__ tiles, int ntiles) {
  __shared__ TableCache<T, CM> tc;
  const int L = blockIdx.x, nt8 = (ntiles + 7) & ~7;
  if (L < nt8) {
    if (L >= ntiles) return;
    load_tables<T, CM>(g, 3, tc);
    unsigned mask;
    int i, j, k;
    if (!cpml_tile_cell(tiles[L], (int)threadIdx.x, mask, i, j, k)) return;
    e_cell_split_all<T, CM>(g, fs, tc, mask, i, j, k, ghost);
  } else {
    int col, ch;
    if (!march_tile_at(m, L - nt8, col, ch)) return;
    load_tables<T, CM>(g, 3, tc);
    update_e_body<T, CM, W, CPML_SPLIT_ALL>(g.Ex, g.Ey, g.Ez, g.Hx, g.Hy, g.Hz, g, m, tc, col, ch, ghost, fs);
  }
}

// ---------------------------------------------------------------- port epilogue --

constexpr int MAX_PORT_OBJS = 8;

struct VSourceDev { int dir, count; const int *idx; const double *coef, *wave; };
struct SVoltageDev { int dir, count; const int *idx; double csvf; };
struct SCurrentDev { int count, negate; const int *comp, *idx; const double *w; };

// A point field probe: magnetic flag, component (0..2, or 3 = magnitude), then the
// device offsets of the 2 (E) or 4 (H) averaged elements of each of Fx, Fy, Fz.
constexpr int PROBE_INTS = 16;
constexpr int MAX_FIELD_PROBES = 64;

struct PortState {
  int nvs, nsv, nsi;
  VSourceDev vs[MAX_PORT_OBJS];
  SVoltageDev sv[MAX_PORT_OBJS];
  SCurrentDev si[MAX_PORT_OBJS];
  double *vtr, *itr;      // [nsv][steps], [nsi][steps]
  int npr;                // point field probes
  const int *pr;          // [npr][PROBE_INTS]
  double *prtr;           // [npr][steps]
  int steps;
  int ndft;               // running DFT frequencies (0 = off)
  const double *w;        // angular frequencies
  double *dft;            // [(nsv+nsi)][ndft][2]
  double dt;
};

// Ordered sum of up to `count` terms whose values lane n of the wavefront holds for
// term base+n: each lane loads one term (a coalesced gather), then the terms are
// folded left to right through wavefront shuffles, so the FP64 result is the
// reference's sequential sum.
__device__ __forceinline__ double ordered_wave_sum(double acc, double mine, int nterms) {
  for (int m = 0; m < nterms; ++m) acc += __shfl(mine, m, 64);
  return acc;
}

// Time step of a launch: `off` alone when launched directly, or the device-held step
// of the first step of a replayed graph plus the kernel's offset within the graph.
__device__ __forceinline__ int step_of(const int *tsp, int off) { return off + (tsp ? *tsp : 0); }

// Value of a point field probe (captureElectricFields / captureMagneticFields,
// js_fdtd/src/sampling.js:144-292): the FP64 average of 2 E or 4 H elements, summed
// left to right as the reference does, or the magnitude of the three averages.
template <typename T>
__device__ __forceinline__ double probe_value(const Grid<T> &g, const int *d) {
#pragma clang fp contract(off)
  const bool mag = d[0] != 0;
  const T *F[3] = {mag ? g.Hx : g.Ex, mag ? g.Hy : g.Ey, mag ? g.Hz : g.Ez};
  auto avg = [&](int c) -> double {
    const int *ix = d + 4 + 4 * c;
    const T *f = F[c];
    if (!mag) return 0.5 * ((double)f[ix[0]] + (double)f[ix[1]]);
    return 0.25 * ((((double)f[ix[0]] + (double)f[ix[1]]) + (double)f[ix[2]]) + (double)f[ix[3]]);
  };
  if (d[1] < 3) return avg(d[1]);
  const double x = avg(0), y = avg(1), z = avg(2);
  return sqrt(x * x + y * y + z * z);
}

// Source injection, then port voltages and currents and the point field probes, then
// the running DFTs of the port quantities, for step ts; run by one whole block.  The injection precedes the captures as in
// fdtdSolver.js:111-117.
template <typename T>
__device__ __forceinline__ void port_epilogue_body(const Grid<T> &g, const PortState &ps, int ts) {
#pragma clang fp contract(off)
  __shared__ double s_val[2 * MAX_PORT_OBJS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
  T *E[3] = {g.Ex, g.Ey, g.Ez};
  const T *H[3] = {g.Hx, g.Hy, g.Hz};

  for (int s = 0; s < ps.nvs; ++s) {
    const VSourceDev &v = ps.vs[s];
    const double val = v.wave[ts];
    for (int n = tid; n < v.count; n += blockDim.x) {
      T *e = E[v.dir] + v.idx[n];
      *e = T((double)*e + v.coef[n] * val);
    }
    __syncthreads();
  }

  // One wavefront per observer.
  for (int o = wave; o < ps.nsv + ps.nsi; o += nwaves) {
    double acc = 0.0;
    if (o < ps.nsv) {
      const SVoltageDev &sv = ps.sv[o];
      for (int base = 0; base < sv.count; base += 64) {
        const int n = base + lane;
        const double x = n < sv.count ? (double)E[sv.dir][sv.idx[n]] : 0.0;
        acc = ordered_wave_sum(acc, x, min(64, sv.count - base));
      }
      acc = sv.csvf * acc;
      if (lane == 0) { ps.vtr[(long)o * ps.steps + ts] = acc; s_val[o] = acc; }
    } else {
      const SCurrentDev &si = ps.si[o - ps.nsv];
      for (int base = 0; base < si.count; base += 64) {
        const int n = base + lane;
        const double x = n < si.count ? si.w[n] * (double)H[si.comp[n]][si.idx[n]] : 0.0;
        acc = ordered_wave_sum(acc, x, min(64, si.count - base));
      }
      if (si.negate) acc = -acc;
      if (lane == 0) { ps.itr[(long)(o - ps.nsv) * ps.steps + ts] = acc; s_val[o] = acc; }
    }
  }
  // Probes read the fields after the E update and the injection; H does not change
  // after the H phase, so H probes see the values captureMagneticFields sees.
  for (int p = tid; p < ps.npr; p += blockDim.x)
    ps.prtr[(long)p * ps.steps + ts] = probe_value<T>(g, ps.pr + (long)p * PROBE_INTS);
  if (ps.ndft == 0) return;
  __syncthreads();
  const int ntr = ps.nsv + ps.nsi;
  const double tv = (ts + 1) * ps.dt;
  const double ti = (ts + 1) * ps.dt + -ps.dt / 2;
  for (int q = tid; q < ntr * ps.ndft; q += blockDim.x) {
    const int o = q / ps.ndft, f = q - o * ps.ndft;
    const double x = s_val[o];
    double sn, cs;
    sincos(ps.w[f] * (o < ps.nsv ? tv : ti), &sn, &cs);
    double *d = ps.dft + 2 * (long)q;
    d[0] += x * cs;
    d[1] -= x * sn;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_port_epilogue(Grid<T> g, PortState ps, const int *tsp, int off) {
  port_epilogue_body<T>(g, ps, step_of(tsp, off));
}

// ------------------------------------------------------------------------ NTFF --

struct FarfieldState {
  int nfreq;
  int LI, LJ, LK, UI, UJ, UK;    // 0-based box planes
  int ni, nj, nk;                // face extents
  const double *phase;           // [steps][nfreq][4]
  double *acc;                   // 24 accumulators, each re[nfreq*size] then im[...]
};

// Accumulator a (0..23: faces xp,xn,yp,yn,zp,zn x {J0,J1,M0,M1}) starts here; its
// imaginary part follows after nfreq*faceSize entries.
__host__ __device__ __forceinline__ long ff_offset(int a, int nfreq, long sx, long sy, long sz) {
  const int grp = a >> 3;
  const long size = (grp == 0 ? sx : (grp == 1 ? sy : sz)) * nfreq;
  const long start = grp == 0 ? 0 : (grp == 1 ? 16 * sx * nfreq : 16 * (sx + sy) * nfreq);
  return start + (long)(a & 7) * 2 * size;
}

// Direct running DFT of the Huygens-box tangential fields, one thread per face cell,
// all frequencies in a loop.  Face and component mapping follow farfield.js:115-246
// (including the zp-face Jx subtract-then-add sequence, :210-226).
template <typename T>
__device__ __forceinline__ void ntff_direct_body(const Grid<T> &g, const FarfieldState &fs, int ts, long q) {
#pragma clang fp contract(off)
  const long sx = (long)fs.nj * fs.nk, sy = (long)fs.ni * fs.nk, sz = (long)fs.ni * fs.nj;
  const long total = 2 * (sx + sy + sz);
  if (q >= total) return;
  const int SJ = g.SJ, SI = g.SI;
  auto at = [&](int i, int j, int k) { return i * SI + j * SJ + k; };
  int face;          // 0 xp, 1 xn, 2 yp, 3 yn, 4 zp, 5 zn
  long cell;
  if (q < 2 * sx) { face = q < sx ? 0 : 1; cell = q - (face == 0 ? 0 : sx); }
  else if (q < 2 * sx + 2 * sy) { long r = q - 2 * sx; face = r < sy ? 2 : 3; cell = r - (face == 2 ? 0 : sy); }
  else { long r = q - 2 * sx - 2 * sy; face = r < sz ? 4 : 5; cell = r - (face == 4 ? 0 : sz); }

  // tJ0, tJ1 (H-based, phase h) and tM0, tM1 (E-based, phase e) of this cell, and the
  // accumulator ids they feed.
  T tJ0, tJ1, tM0, tM1;
  int aJ0, aJ1, aM0, aM1;
  T tJx2 = T(0);    // zp only
  if (face <= 1) {
    const int a = (int)(cell / fs.nk), b = (int)(cell - (long)a * fs.nk);
    const int j0 = fs.LJ + a, k0 = fs.LK + b;
    const int I = face == 0 ? fs.UI : fs.LI;
    const T sgn = face == 0 ? T(1) : T(-1);
    tM0 = (sgn * T(0.5)) * (g.Ez[at(I, j0, k0)] + g.Ez[at(I, j0 + 1, k0)]);
    tM1 = (-sgn * T(0.5)) * (g.Ey[at(I, j0, k0)] + g.Ey[at(I, j0, k0 + 1)]);
    tJ0 = (-sgn * T(0.25)) * (g.Hz[at(I, j0, k0)] + g.Hz[at(I, j0, k0 + 1)] + g.Hz[at(I - 1, j0, k0)] + g.Hz[at(I - 1, j0, k0 + 1)]);
    tJ1 = (sgn * T(0.25)) * (g.Hy[at(I, j0, k0)] + g.Hy[at(I, j0 + 1, k0)] + g.Hy[at(I - 1, j0, k0)] + g.Hy[at(I - 1, j0 + 1, k0)]);
    const int base = face == 0 ? 0 : 4;
    aJ0 = base; aJ1 = base + 1; aM0 = base + 2; aM1 = base + 3;
  } else if (face <= 3) {
    const int a = (int)(cell / fs.nk), b = (int)(cell - (long)a * fs.nk);
    const int i0 = fs.LI + a, k0 = fs.LK + b;
    const int J = face == 2 ? fs.UJ : fs.LJ;
    const T sgn = face == 2 ? T(1) : T(-1);
    tM0 = (-sgn * T(0.5)) * (g.Ez[at(i0, J, k0)] + g.Ez[at(i0 + 1, J, k0)]);
    tM1 = (sgn * T(0.5)) * (g.Ex[at(i0, J, k0)] + g.Ex[at(i0, J, k0 + 1)]);
    tJ1 = (-sgn * T(0.25)) * (g.Hx[at(i0, J, k0)] + g.Hx[at(i0 + 1, J, k0)] + g.Hx[at(i0, J - 1, k0)] + g.Hx[at(i0 + 1, J - 1, k0)]);
    tJ0 = (sgn * T(0.25)) * (g.Hz[at(i0, J, k0)] + g.Hz[at(i0, J, k0 + 1)] + g.Hz[at(i0, J - 1, k0)] + g.Hz[at(i0, J - 1, k0 + 1)]);
    const int base = face == 2 ? 8 : 12;
    aJ0 = base; aJ1 = base + 1; aM0 = base + 2; aM1 = base + 3;
  } else {
    const int a = (int)(cell / fs.nj), b = (int)(cell - (long)a * fs.nj);
    const int i0 = fs.LI + a, j0 = fs.LJ + b;
    const int K = face == 4 ? fs.UK : fs.LK;
    const T sgn = face == 4 ? T(1) : T(-1);
    tM0 = (sgn * T(0.5)) * (g.Ey[at(i0, j0, K)] + g.Ey[at(i0 + 1, j0, K)]);
    tM1 = (-sgn * T(0.5)) * (g.Ex[at(i0, j0, K)] + g.Ex[at(i0, j0 + 1, K)]);
    const T avgHx = g.Hx[at(i0, j0, K)] + g.Hx[at(i0 + 1, j0, K)] + g.Hx[at(i0, j0, K - 1)] + g.Hx[at(i0 + 1, j0, K - 1)];
    const T avgHy = g.Hy[at(i0, j0, K)] + g.Hy[at(i0, j0 + 1, K)] + g.Hy[at(i0, j0, K - 1)] + g.Hy[at(i0, j0 + 1, K - 1)];
    if (face == 4) {
      tJ0 = T(0.25) * avgHx;          // Jx, replaced below by -0.25*avgHy
      tJ1 = T(0.25) * avgHx;          // Jy
      tJx2 = T(-0.25) * avgHy;
    } else {
      tJ0 = T(0.25) * avgHy;          // Jx
      tJ1 = T(-0.25) * avgHx;         // Jy
    }
    const int base = face == 4 ? 16 : 20;
    aJ0 = base; aJ1 = base + 1; aM0 = base + 2; aM1 = base + 3;
  }

  const long fsz = face <= 1 ? sx : (face <= 3 ? sy : sz);
  const long fsize = fsz * fs.nfreq;
  double *rJ0 = fs.acc + ff_offset(aJ0, fs.nfreq, sx, sy, sz), *iJ0 = rJ0 + fsize;
  double *rJ1 = fs.acc + ff_offset(aJ1, fs.nfreq, sx, sy, sz), *iJ1 = rJ1 + fsize;
  double *rM0 = fs.acc + ff_offset(aM0, fs.nfreq, sx, sy, sz), *iM0 = rM0 + fsize;
  double *rM1 = fs.acc + ff_offset(aM1, fs.nfreq, sx, sy, sz), *iM1 = rM1 + fsize;
  const double j0 = (double)tJ0, j1 = (double)tJ1, m0 = (double)tM0, m1 = (double)tM1;
  const double j2 = (double)tJx2;
  for (int f = 0; f < fs.nfreq; ++f) {
    const double *ph = fs.phase + ((long)ts * fs.nfreq + f) * 4;
    const double ch = ph[0], sh = ph[1], ce = ph[2], se = ph[3];
    const long pos = f * fsz + cell;
    rM0[pos] = rM0[pos] + ce * m0; iM0[pos] = iM0[pos] + se * m0;
    rM1[pos] = rM1[pos] + ce * m1; iM1[pos] = iM1[pos] + se * m1;
    if (face == 4) {
      const double r = rJ0[pos] + ch * j0, m = iJ0[pos] + sh * j0;
      rJ0[pos] = r - ch * j0 + ch * j2;
      iJ0[pos] = m - sh * j0 + sh * j2;
    } else {
      rJ0[pos] = rJ0[pos] + ch * j0; iJ0[pos] = iJ0[pos] + sh * j0;
    }
    rJ1[pos] = rJ1[pos] + ch * j1; iJ1[pos] = iJ1[pos] + sh * j1;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_ntff_direct(Grid<T> g, FarfieldState fs, const int *tsp, int off) {
  ntff_direct_body<T>(g, fs, step_of(tsp, off), (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Advances the device-held step by `by` at the end of a replayed graph.
__global__ void k_advance_step(int *tsp, int by) {
  if (threadIdx.x == 0) *tsp += by;
}

// Sets the device-held step before a run of graph replays.
__global__ void k_set_step(int *tsp, int v) {
  if (threadIdx.x == 0) *tsp = v;
}

}  // namespace fdtd
