// Device code of the gfx950 Yee time step.
//
// Device layout.  All six field components and all 18 coefficient arrays share one
// "uniform" shape (nx+1) x (ny+1) x (nz+1), k fastest: element (i,j,k) of every
// component sits at i*SI + j*SJ + k with SJ = nz+1, SI = (ny+1)*SJ.  Each component
// occupies the sub-box of its own Yee shape; the other slots stay zero and are never
// written.  One flat offset therefore addresses Hx..Ez and their coefficients at the
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

namespace fdtd {

enum CoeffId {
  CEXE, CEXHZ, CEXHY, CEYE, CEYHX, CEYHZ, CEZE, CEZHY, CEZHX,
  CHXH, CHXEZ, CHXEY, CHYH, CHYEX, CHYEZ, CHZH, CHZEY, CHZEX, NCOEFF
};

template <typename T>
struct Grid {
  int nx, ny, nz;
  int SJ, SI;
  T *Hx, *Hy, *Hz, *Ex, *Ey, *Ez;
  const T *C[NCOEFF];
};

// One CPML face.  Psi/CPsi keep the reference layout: the component's shape with the
// extent along the face axis replaced by nc.
template <typename T>
struct Face {
  int nc, m_start, e_start;
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

// Convolution term of the two faces (n then p) along AXIS for slot SLOT.  MAG selects
// the magnetic profiles and H planes.  `diff` is the curl difference the face's Psi
// integrates, identical to the one in the bulk update.
template <typename T, int AXIS, int SLOT, bool MAG>
__device__ __forceinline__ T cpml_term(const FaceSet<T> &fs, T val, T diff,
                                       int i, int j, int k, int d1, int d2) {
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const Face<T> &F = fs.f[2 * AXIS + s];
    const int coord = AXIS == 0 ? i : (AXIS == 1 ? j : k);
    const int ci = coord - (MAG ? F.m_start : F.e_start);
    if ((unsigned)ci < (unsigned)F.nc) {
      const int q = psi_offset<AXIS>(i, j, k, ci, F.nc, d1, d2);
      const T b = MAG ? F.b_m[ci] : F.b_e[ci];
      const T a = MAG ? F.a_m[ci] : F.a_e[ci];
      const T ps = b * F.psi[SLOT][q] + a * diff;
      F.psi[SLOT][q] = ps;
      val = val + F.cpsi[SLOT][q] * ps;
    }
  }
  return val;
}

// H update, bulk and magnetic CPML fused, one thread per (i, flat plane offset p).
// Hx over i in [0,nx], Hy/Hz over i in [0,nx).  Per component the CPML faces are
// applied in the reference's face order (x, then y, then z).
template <typename T>
__global__ void __launch_bounds__(256) k_update_h(Grid<T> g, FaceSet<T> fs) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (p >= g.SI) return;
  const int j = p / g.SJ;
  const int k = p - j * g.SJ;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int c = i * g.SI + p;

  if (j < ny && k < nz) {                                  // Hx (nx+1, ny, nz)
    const T dEy = g.Ey[c + 1] - g.Ey[c];
    const T dEz = g.Ez[c + g.SJ] - g.Ez[c];
    T h = g.C[CHXH][c] * g.Hx[c] + g.C[CHXEY][c] * dEy + g.C[CHXEZ][c] * dEz;
    h = cpml_term<T, 1, 0, true>(fs, h, dEz, i, j, k, ny, nz);
    h = cpml_term<T, 2, 0, true>(fs, h, dEy, i, j, k, ny, nz);
    g.Hx[c] = h;
  }
  if (i < nx) {
    if (k < nz) {                                          // Hy (nx, ny+1, nz)
      const T dEz = g.Ez[c + g.SI] - g.Ez[c];
      const T dEx = g.Ex[c + 1] - g.Ex[c];
      T h = g.C[CHYH][c] * g.Hy[c] + g.C[CHYEZ][c] * dEz + g.C[CHYEX][c] * dEx;
      h = cpml_term<T, 0, 0, true>(fs, h, dEz, i, j, k, ny + 1, nz);
      h = cpml_term<T, 2, 1, true>(fs, h, dEx, i, j, k, ny + 1, nz);
      g.Hy[c] = h;
    }
    if (j < ny) {                                          // Hz (nx, ny, nz+1)
      const T dEx = g.Ex[c + g.SJ] - g.Ex[c];
      const T dEy = g.Ey[c + g.SI] - g.Ey[c];
      T h = g.C[CHZH][c] * g.Hz[c] + g.C[CHZEX][c] * dEx + g.C[CHZEY][c] * dEy;
      h = cpml_term<T, 0, 1, true>(fs, h, dEy, i, j, k, ny, nz + 1);
      h = cpml_term<T, 1, 1, true>(fs, h, dEx, i, j, k, ny, nz + 1);
      g.Hz[c] = h;
    }
  }
}

// E update, bulk and electric CPML fused.  With ghost != 0 (reference boundary) the
// grid covers i in [0,nx) and the thread at i = 0 also updates Ey/Ez on the i = nx
// wall the way the reference's updateE does: Ey there differences "Hz[nx]", which is
// Ex(0,j,k) after this step's bulk update and before its CPML terms, and Ez
// differences "Hy[nx]", which is the reference-layout flat element j*nz+k of Hz.
// With ghost == 0 the grid covers i in [0,nx] and the wall gets CPML terms only.
template <typename T>
__global__ void __launch_bounds__(256) k_update_e(Grid<T> g, FaceSet<T> fs, int ghost) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (p >= g.SI) return;
  const int j = p / g.SJ;
  const int k = p - j * g.SJ;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int SJ = g.SJ, SI = g.SI;
  const int c = i * SI + p;
  const bool kin = k >= 1 && k < nz, jin = j >= 1 && j < ny, iin = i >= 1 && i < nx;

  T exb = T(0);
  if (i < nx) {                                            // Ex (nx, ny+1, nz+1)
    const T dHz = jin ? g.Hz[c] - g.Hz[c - SJ] : T(0);
    const T dHy = kin ? g.Hy[c] - g.Hy[c - 1] : T(0);
    T e = g.Ex[c];
    if (jin && kin) e = g.C[CEXE][c] * e + g.C[CEXHZ][c] * dHz + g.C[CEXHY][c] * dHy;
    exb = e;
    e = cpml_term<T, 1, 3, false>(fs, e, dHz, i, j, k, ny + 1, nz + 1);
    e = cpml_term<T, 2, 2, false>(fs, e, dHy, i, j, k, ny + 1, nz + 1);
    if (j <= ny && k <= nz) g.Ex[c] = e;
  }
  if (j < ny) {                                            // Ey (nx+1, ny, nz+1)
    const T dHx = kin ? g.Hx[c] - g.Hx[c - 1] : T(0);
    const T dHz = iin ? g.Hz[c] - g.Hz[c - SI] : T(0);
    T e = g.Ey[c];
    if (iin && kin) e = g.C[CEYE][c] * e + g.C[CEYHX][c] * dHx + g.C[CEYHZ][c] * dHz;
    e = cpml_term<T, 0, 2, false>(fs, e, dHz, i, j, k, ny, nz + 1);
    e = cpml_term<T, 2, 3, false>(fs, e, dHx, i, j, k, ny, nz + 1);
    g.Ey[c] = e;
  }
  if (k < nz) {                                            // Ez (nx+1, ny+1, nz)
    const T dHy = iin ? g.Hy[c] - g.Hy[c - SI] : T(0);
    const T dHx = jin ? g.Hx[c] - g.Hx[c - SJ] : T(0);
    T e = g.Ez[c];
    if (iin && jin) e = g.C[CEZE][c] * e + g.C[CEZHY][c] * dHy + g.C[CEZHX][c] * dHx;
    e = cpml_term<T, 0, 3, false>(fs, e, dHy, i, j, k, ny + 1, nz);
    e = cpml_term<T, 1, 2, false>(fs, e, dHx, i, j, k, ny + 1, nz);
    g.Ez[c] = e;
  }

  if (ghost && i == 0) {
    const int cg = nx * SI + p;
    if (j < ny && kin) {                                   // Ey(nx, j, k)
      const T dHx = g.Hx[cg] - g.Hx[cg - 1];
      const T dHz = exb - g.Hz[cg - SI];
      T e = g.C[CEYE][cg] * g.Ey[cg] + g.C[CEYHX][cg] * dHx + g.C[CEYHZ][cg] * dHz;
      e = cpml_term<T, 2, 3, false>(fs, e, dHx, nx, j, k, ny, nz + 1);
      g.Ey[cg] = e;
    }
    if (jin && k < nz) {                                   // Ez(nx, j, k)
      const T dHy = g.Hz[j * nz + k] - g.Hy[cg - SI];
      const T dHx = g.Hx[cg] - g.Hx[cg - SJ];
      T e = g.C[CEZE][cg] * g.Ez[cg] + g.C[CEZHY][cg] * dHy + g.C[CEZHX][cg] * dHx;
      e = cpml_term<T, 1, 2, false>(fs, e, dHx, nx, j, k, ny + 1, nz);
      g.Ez[cg] = e;
    }
  }
}

// ---------------------------------------------------------------- port epilogue --

constexpr int MAX_PORT_OBJS = 8;

struct VSourceDev { int dir, count; const int *idx; const double *coef, *wave; };
struct SVoltageDev { int dir, count; const int *idx; double csvf; };
struct SCurrentDev { int count, negate; const int *comp, *idx; const double *w; };

struct PortState {
  int nvs, nsv, nsi;
  VSourceDev vs[MAX_PORT_OBJS];
  SVoltageDev sv[MAX_PORT_OBJS];
  SCurrentDev si[MAX_PORT_OBJS];
  double *vtr, *itr;      // [nsv][steps], [nsi][steps]
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

// Source injection, then port voltages and currents, then their running DFTs.
// One block; the injection precedes the captures as in fdtdSolver.js:111-117.
template <typename T>
__global__ void __launch_bounds__(256) k_port_epilogue(Grid<T> g, PortState ps, int ts) {
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
__global__ void __launch_bounds__(256) k_ntff_direct(Grid<T> g, FarfieldState fs, int ts) {
#pragma clang fp contract(off)
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
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

}  // namespace fdtd
