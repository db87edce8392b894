// K-blocked running DFT of the Huygens-box surface currents on the matrix cores.
//
// The direct form (k_ntff_direct) reads and writes every FP64 accumulator of every
// frequency on every step.  Here each step only stores its tangential samples, one
// row of a [K][samples] buffer, and every K steps one launch folds the block into the
// accumulators as a matrix product:
//
//   acc[n][s] += sum_k P[n][k] * t[k][s],   n = (frequency, re/im), s = surface sample
//
// with P the FP64 phase table of the block's steps (cos and -sin of w*dt*(ts+1/2) for
// the H-based J samples, of w*dt*(ts+1) for the E-based M samples; farfield.js:101-110).
// The product runs on v_mfma_f64_16x16x4f64 in FP64; the samples are the field
// precision (FP32 in the fast path).  The result equals the step-by-step sums to FP64
// rounding; the zp-face Jx subtract-then-add sequence (farfield.js:211-226) is taken at
// its mathematical value, so this form is not the bit-exact FP64 drop-in.
#pragma once

namespace fdtd {

constexpr int NTFF_K = 64;          // steps per block
constexpr int NTFF_FREQ_GROUP = 16; // frequencies per launch (two 16-column tiles of P)

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

template <typename T>
struct NtffBlock {
  T *buf;                 // [NTFF_K][spad] samples
  long spad;              // padded sample count (multiple of 16)
  long base[25];          // first sample of accumulator a (multiples of 16); base[24] = spad
  long size[24];          // surface cells of accumulator a
};

// Gather: the four tangential values of every face cell of this step into row `row`.
// Face and component mapping as k_ntff_direct (farfield.js:115-246).
template <typename T>
__device__ __forceinline__ void ntff_gather_body(const Grid<T> &g, const FarfieldState &fs, const NtffBlock<T> &nb,
                                                 int row, long q) {
#pragma clang fp contract(off)
  const long sx = (long)fs.nj * fs.nk, sy = (long)fs.ni * fs.nk, sz = (long)fs.ni * fs.nj;
  if (q >= 2 * (sx + sy + sz)) return;
  const int SJ = g.SJ, SI = g.SI;
  auto at = [&](int i, int j, int k) { return i * SI + j * SJ + k; };
  int face;
  long cell;
  if (q < 2 * sx) { face = q < sx ? 0 : 1; cell = q - (face == 0 ? 0 : sx); }
  else if (q < 2 * sx + 2 * sy) { long r = q - 2 * sx; face = r < sy ? 2 : 3; cell = r - (face == 2 ? 0 : sy); }
  else { long r = q - 2 * sx - 2 * sy; face = r < sz ? 4 : 5; cell = r - (face == 4 ? 0 : sz); }
  T tJ0, tJ1, tM0, tM1;
  if (face <= 1) {
    const int a = (int)(cell / fs.nk), b = (int)(cell - (long)a * fs.nk);
    const int j0 = fs.LJ + a, k0 = fs.LK + b;
    const int I = face == 0 ? fs.UI : fs.LI;
    const T sgn = face == 0 ? T(1) : T(-1);
    tM0 = (sgn * T(0.5)) * (g.Ez[at(I, j0, k0)] + g.Ez[at(I, j0 + 1, k0)]);
    tM1 = (-sgn * T(0.5)) * (g.Ey[at(I, j0, k0)] + g.Ey[at(I, j0, k0 + 1)]);
    tJ0 = (-sgn * T(0.25)) * (g.Hz[at(I, j0, k0)] + g.Hz[at(I, j0, k0 + 1)] + g.Hz[at(I - 1, j0, k0)] + g.Hz[at(I - 1, j0, k0 + 1)]);
    tJ1 = (sgn * T(0.25)) * (g.Hy[at(I, j0, k0)] + g.Hy[at(I, j0 + 1, k0)] + g.Hy[at(I - 1, j0, k0)] + g.Hy[at(I - 1, j0 + 1, k0)]);
  } else if (face <= 3) {
    const int a = (int)(cell / fs.nk), b = (int)(cell - (long)a * fs.nk);
    const int i0 = fs.LI + a, k0 = fs.LK + b;
    const int J = face == 2 ? fs.UJ : fs.LJ;
    const T sgn = face == 2 ? T(1) : T(-1);
    tM0 = (-sgn * T(0.5)) * (g.Ez[at(i0, J, k0)] + g.Ez[at(i0 + 1, J, k0)]);
    tM1 = (sgn * T(0.5)) * (g.Ex[at(i0, J, k0)] + g.Ex[at(i0, J, k0 + 1)]);
    tJ1 = (-sgn * T(0.25)) * (g.Hx[at(i0, J, k0)] + g.Hx[at(i0 + 1, J, k0)] + g.Hx[at(i0, J - 1, k0)] + g.Hx[at(i0 + 1, J - 1, k0)]);
    tJ0 = (sgn * T(0.25)) * (g.Hz[at(i0, J, k0)] + g.Hz[at(i0, J, k0 + 1)] + g.Hz[at(i0, J - 1, k0)] + g.Hz[at(i0, J - 1, k0 + 1)]);
  } else {
    const int a = (int)(cell / fs.nj), b = (int)(cell - (long)a * fs.nj);
    const int i0 = fs.LI + a, j0 = fs.LJ + b;
    const int K = face == 4 ? fs.UK : fs.LK;
    const T sgn = face == 4 ? T(1) : T(-1);
    tM0 = (sgn * T(0.5)) * (g.Ey[at(i0, j0, K)] + g.Ey[at(i0 + 1, j0, K)]);
    tM1 = (-sgn * T(0.5)) * (g.Ex[at(i0, j0, K)] + g.Ex[at(i0, j0 + 1, K)]);
    const T avgHx = g.Hx[at(i0, j0, K)] + g.Hx[at(i0 + 1, j0, K)] + g.Hx[at(i0, j0, K - 1)] + g.Hx[at(i0 + 1, j0, K - 1)];
    const T avgHy = g.Hy[at(i0, j0, K)] + g.Hy[at(i0, j0 + 1, K)] + g.Hy[at(i0, j0, K - 1)] + g.Hy[at(i0, j0 + 1, K - 1)];
    if (face == 4) { tJ0 = T(-0.25) * avgHy; tJ1 = T(0.25) * avgHx; }
    else { tJ0 = T(0.25) * avgHy; tJ1 = T(-0.25) * avgHx; }
  }
  const int a0 = face * 4;
  T *r = nb.buf + (long)row * nb.spad;
  r[nb.base[a0 + 0] + cell] = tJ0;
  r[nb.base[a0 + 1] + cell] = tJ1;
  r[nb.base[a0 + 2] + cell] = tM0;
  r[nb.base[a0 + 3] + cell] = tM1;
}

template <typename T>
__global__ void __launch_bounds__(256) k_ntff_gather(Grid<T> g, FarfieldState fs, NtffBlock<T> nb, int row) {
  ntff_gather_body<T>(g, fs, nb, row, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Fold rows [r0, r1) of the block starting at step ts0 (see step_of) into the accumulators of
// frequencies [f0, f0 + NT*8).  A wavefront handles 16-sample tiles (tiles never
// straddle accumulators); P in LDS (NT = 2: 32 KiB), zero outside [r0, r1) and beyond
// the last frequency.
template <typename T, int NT>
__global__ void __launch_bounds__(256) k_ntff_mfma(FarfieldState fs, NtffBlock<T> nb, const int *tsp, int ts_off,
                                                   int r0, int r1, int f0) {
  constexpr int NP = NT * 16;
  const int ts0 = step_of(tsp, ts_off);
  __shared__ double P[2 * NTFF_K * NP];   // J (H phase) then M (E phase), [k][column]
  const int ncol = 2 * (fs.nfreq - f0);    // columns left from f0 (only the first NP used)
  for (int q = threadIdx.x; q < 2 * NTFF_K * NP; q += blockDim.x) {
    const int type = q / (NTFF_K * NP), k = (q / NP) % NTFF_K, n = q % NP;
    double v = 0.0;
    if (k >= r0 && k < r1 && n < ncol)
      v = fs.phase[((long)(ts0 + k) * fs.nfreq + f0 + n / 2) * 4 + type * 2 + (n & 1)];
    P[q] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, kl = lane >> 4;
  const long sxs = (long)fs.nj * fs.nk, sys = (long)fs.ni * fs.nk, szs = (long)fs.ni * fs.nj;
  const long ntiles = nb.spad / 16;
  // Persistent: the phase tile above is filled once per workgroup, then each wavefront
  // walks 16-sample tiles with a stride of the whole grid.
  for (long tile = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < ntiles;
       tile += (long)gridDim.x * (blockDim.x >> 6)) {
    const long s0 = tile * 16;
    int a = 0;
    while (s0 >= nb.base[a + 1]) ++a;
    const double *Pt = P + ((a & 3) >= 2 ? NTFF_K * NP : 0);   // accumulators 2, 3 of a face are M
    mfma_d4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    const T *col = nb.buf + s0 + (lane & 15);
    // Fully unrolled, so all 16 sample loads of the tile are issued before the first
    // matrix-core step; the fold is HBM-bound and needs the loads in flight together.
#pragma unroll
    for (int kk = 0; kk < NTFF_K; kk += 4) {
      const double b = (double)col[(long)(kk + kl) * nb.spad];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double av = Pt[(kk + kl) * NP + t * 16 + (lane & 15)];
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b, acc[t], 0, 0, 0);
      }
    }
    // D[row = lane/16 + 4v][col = lane%16]: row = column n of P, col = sample.
    const long fsz = a < 8 ? sxs : (a < 16 ? sys : szs);
    const long cell = s0 - nb.base[a] + (lane & 15);
    if (cell < nb.size[a]) {
      double *re = fs.acc + ff_offset(a, fs.nfreq, sxs, sys, szs);
      const long im = fsz * fs.nfreq;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int n = t * 16 + kl + 4 * v;
          if (n < ncol) {
            double *d = re + (n & 1) * im + (long)(f0 + (n >> 1)) * fsz + cell;
            *d += acc[t][v];
          }
        }
      }
    }
  }
}

}  // namespace fdtd
