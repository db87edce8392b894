// Near-to-far-field post-processing on the GPU, FP64, from the device-resident
// Huygens-box accumulators (layout of FarfieldState / ff_offset).
//
//   k_ntff_power:       P(f) = 0.5 * sum over faces of +-dA * Re(M x conj(J)).n
//                       (js_fdtd/src/farfield.js:257-314)
//   k_ntff_directivity: N (from J) and L (from M) radiation vectors with the phase
//                       k r^.r' of every surface cell, then
//                       D_theta = k^2/(8 pi eta0 P) |L_phi + eta0 N_theta|^2,
//                       D_phi   = k^2/(8 pi eta0 P) |L_theta - eta0 N_phi|^2
//                       (farfield.js:328-607)
//
// The projections and phases are the reference's per cell; the sums run in parallel,
// so results agree with the sequential JavaScript to rounding (about 1e-13 relative),
// not bit for bit.
#pragma once

namespace fdtd {

struct NtffPost {
  const double *acc;        // 24 accumulators (ff_offset layout)
  int nfreq, ni, nj, nk;    // face extents
  double li, lj, lk, ui, uj, uk;   // 1-based box planes, as farfield.js holds them
  double dx, dy, dz;
};

// Surface cells: x faces (nj*nk), then y faces (ni*nk), then z faces (ni*nj); each
// cell stands for the pair of opposite faces.
__device__ __forceinline__ void ntff_cell(const NtffPost &p, long q, int &axis, long &cell, int &r, int &c) {
  const long sx = (long)p.nj * p.nk, sy = (long)p.ni * p.nk;
  if (q < sx) { axis = 0; cell = q; r = (int)(q / p.nk); c = (int)(q - (long)r * p.nk); }
  else if (q < sx + sy) { axis = 1; cell = q - sx; r = (int)(cell / p.nk); c = (int)(cell - (long)r * p.nk); }
  else { axis = 2; cell = q - sx - sy; r = (int)(cell / p.nj); c = (int)(cell - (long)r * p.nj); }
}

// Block-wide sum of V doubles per thread (256 threads = 4 wavefronts); thread 0 gets
// the totals in out[].
template <int V>
__device__ __forceinline__ void block_sum(double (&v)[V], double *lds /* 4*V */) {
#pragma unroll
  for (int q = 0; q < V; ++q)
    for (int off = 32; off > 0; off >>= 1) v[q] += __shfl_xor(v[q], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < V; ++q) lds[wave * V + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < V; ++q) v[q] = ((lds[q] + lds[V + q]) + lds[2 * V + q]) + lds[3 * V + q];
  }
}

// One block per frequency f0 + blockIdx.x of the accumulators; out[blockIdx.x].
__global__ void __launch_bounds__(256) k_ntff_power(NtffPost p, double *out, int f0) {
  __shared__ double lds[4];
  const int f = f0 + (int)blockIdx.x;
  const long sx = (long)p.nj * p.nk, sy = (long)p.ni * p.nk, sz = (long)p.ni * p.nj;
  double acc[1] = {0.0};
  for (long q = threadIdx.x; q < sx + sy + sz; q += blockDim.x) {
    int axis, r, c;
    long cell;
    ntff_cell(p, q, axis, cell, r, c);
    const long fsz = axis == 0 ? sx : (axis == 1 ? sy : sz);
    const double dA = axis == 0 ? p.dy * p.dz : (axis == 1 ? p.dx * p.dz : p.dx * p.dy);
    // Re(M1 conj J0) - Re(M0 conj J1) on x and z faces; its negative on y faces.
    const double kind = axis == 1 ? -1.0 : 1.0;
    const long pos = (long)f * fsz + cell;
    for (int side = 0; side < 2; ++side) {
      const int base = axis * 8 + side * 4;
      const double *J0 = p.acc + ff_offset(base + 0, p.nfreq, sx, sy, sz);
      const double *J1 = p.acc + ff_offset(base + 1, p.nfreq, sx, sy, sz);
      const double *M0 = p.acc + ff_offset(base + 2, p.nfreq, sx, sy, sz);
      const double *M1 = p.acc + ff_offset(base + 3, p.nfreq, sx, sy, sz);
      const long im = fsz * p.nfreq;
      const double t = M1[pos] * J0[pos] + M1[pos + im] * J0[pos + im] - M0[pos] * J1[pos] - M0[pos + im] * J1[pos + im];
      acc[0] += (side == 0 ? 1.0 : -1.0) * kind * dA * t;
    }
  }
  block_sum<1>(acc, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = 0.5 * acc[0];
}

// Wave numbers k = 2 pi f / c and the angle table [nA][4] = sin th, cos th, sin ph,
// cos ph that k_ntff_directivity takes.
__global__ void __launch_bounds__(256) k_directivity_prep(const double *freq, const double *theta, const double *phi,
                                                          int nf, int nA, double *kwave, double *ang) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nf) kwave[t] = 2 * M_PI * freq[t] / 2.99792458e8;
  if (t < nA) {
    double s, c;
    sincos(theta[t], &s, &c);
    ang[4 * t] = s; ang[4 * t + 1] = c;
    sincos(phi[t], &s, &c);
    ang[4 * t + 2] = s; ang[4 * t + 3] = c;
  }
}

// One block per (frequency, group of NA angles): frequency f0 + blockIdx.y of the
// accumulators, entry blockIdx.y of kwave, prad and the outputs.
template <int NA>
__global__ void __launch_bounds__(256) k_ntff_directivity(NtffPost p, const double *kwave, const double *prad,
                                                          const double *ang /* [nA][4]: sin th, cos th, sin ph, cos ph */,
                                                          int nA, double *dth, double *dph, int f0) {
  __shared__ double lds[4 * 8 * NA];
  const int f = blockIdx.y, fa = f0 + (int)blockIdx.y, a0 = blockIdx.x * NA;
  const long sx = (long)p.nj * p.nk, sy = (long)p.ni * p.nk, sz = (long)p.ni * p.nj;
  const double k = kwave[f];
  const double ci = 0.5 * (p.ui + p.li), cj = 0.5 * (p.uj + p.lj), ck = 0.5 * (p.uk + p.lk);
  double st[NA], ct[NA], sp[NA], cp[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int aa = min(a0 + a, nA - 1);
    st[a] = ang[4 * aa]; ct[a] = ang[4 * aa + 1]; sp[a] = ang[4 * aa + 2]; cp[a] = ang[4 * aa + 3];
  }
  // Per angle: N_theta, N_phi, L_theta, L_phi, re and im.
  double s[8 * NA];
#pragma unroll
  for (int q = 0; q < 8 * NA; ++q) s[q] = 0.0;

  for (long q = threadIdx.x; q < sx + sy + sz; q += blockDim.x) {
    int axis, r, c;
    long cell;
    ntff_cell(p, q, axis, cell, r, c);
    const long fsz = axis == 0 ? sx : (axis == 1 ? sy : sz);
    const long pos = (long)fa * fsz + cell, im = fsz * p.nfreq;
    const double dA = axis == 0 ? p.dy * p.dz : (axis == 1 ? p.dx * p.dz : p.dx * p.dy);
    // Cell-centre offsets from the box centre along the two surface axes.
    const double ri = (p.li - 1) + r - ci + 0.5, rj = (p.lj - 1) + r - cj + 0.5;
    const double cjo = (p.lj - 1) + c - cj + 0.5, cko = (p.lk - 1) + c - ck + 0.5;
    for (int side = 0; side < 2; ++side) {
      const int base = axis * 8 + side * 4;
      const double *J0 = p.acc + ff_offset(base + 0, p.nfreq, sx, sy, sz) + pos;
      const double *J1 = p.acc + ff_offset(base + 1, p.nfreq, sx, sy, sz) + pos;
      const double *M0 = p.acc + ff_offset(base + 2, p.nfreq, sx, sy, sz) + pos;
      const double *M1 = p.acc + ff_offset(base + 3, p.nfreq, sx, sy, sz) + pos;
      const double j0r = J0[0], j0i = J0[im], j1r = J1[0], j1i = J1[im];
      const double m0r = M0[0], m0i = M0[im], m1r = M1[0], m1i = M1[im];
      double ox, oy, oz;
      if (axis == 0) { ox = (side == 0 ? p.ui : p.li) - ci; oy = rj; oz = cko; }
      else if (axis == 1) { ox = ri; oy = (side == 0 ? p.uj : p.lj) - cj; oz = cko; }
      else { ox = ri; oy = cjo; oz = (side == 0 ? p.uk : p.lk) - ck; }
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const double ph = ox * p.dx * st[a] * cp[a] + oy * p.dy * st[a] * sp[a] + oz * p.dz * ct[a];
        double sn, cs;
        sincos(k * ph, &sn, &cs);
        // N_theta = J0*A1 + J1*A2, N_phi = J0*B1 + J1*B2 (L likewise from M).
        double A1, A2, B1, B2;
        if (axis == 0) { A1 = dA * (ct[a] * sp[a]); A2 = -(dA * st[a]); B1 = dA * cp[a]; B2 = 0.0; }
        else if (axis == 1) { A1 = dA * (ct[a] * cp[a]); A2 = -(dA * st[a]); B1 = -(dA * sp[a]); B2 = 0.0; }
        else { A1 = dA * (ct[a] * cp[a]); A2 = dA * (ct[a] * sp[a]); B1 = -(dA * sp[a]); B2 = dA * cp[a]; }
        double xr = j0r * A1 + j1r * A2, xi = j0i * A1 + j1i * A2;
        s[8 * a + 0] += xr * cs - xi * sn; s[8 * a + 1] += xr * sn + xi * cs;
        xr = j0r * B1 + j1r * B2; xi = j0i * B1 + j1i * B2;
        s[8 * a + 2] += xr * cs - xi * sn; s[8 * a + 3] += xr * sn + xi * cs;
        xr = m0r * A1 + m1r * A2; xi = m0i * A1 + m1i * A2;
        s[8 * a + 4] += xr * cs - xi * sn; s[8 * a + 5] += xr * sn + xi * cs;
        xr = m0r * B1 + m1r * B2; xi = m0i * B1 + m1i * B2;
        s[8 * a + 6] += xr * cs - xi * sn; s[8 * a + 7] += xr * sn + xi * cs;
      }
    }
  }
  block_sum<8 * NA>(s, lds);
  if (threadIdx.x == 0) {
    const double eta0 = 376.73031346177066;
    const double scale = k * k / (8 * M_PI * eta0 * prad[f]);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (a0 + a >= nA) break;
      const double tr = s[8 * a + 6] + eta0 * s[8 * a + 0], ti = s[8 * a + 7] + eta0 * s[8 * a + 1];
      const double pr = s[8 * a + 4] - eta0 * s[8 * a + 2], pi = s[8 * a + 5] - eta0 * s[8 * a + 3];
      dth[(long)f * nA + a0 + a] = scale * (tr * tr + ti * ti);
      dph[(long)f * nA + a0 + a] = scale * (pr * pr + pi * pi);
    }
  }
}

}  // namespace fdtd
