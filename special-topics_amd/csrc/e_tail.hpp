// The E-phase tail of one time step as a single launch: one block for the port
// epilogue (source injection, port and probe captures, running port DFT), then the
// NTFF blocks (NTFF = 1: direct running DFT; 2: blocked gather into row `row`), then
// the E CPML tiles.  Separately these are three launches; a small grid (the IFA's 721k
// cells) spends several microseconds of every step on the two short ones.  The
// epilogue is one block with a serial chain of a few microseconds, so it is dealt
// first and runs beside the CPML tiles instead of after them.  The host uses this form
// only when it has shown the three touch disjoint field elements: every E element the
// epilogue writes or reads and every E element the NTFF reads lies outside the E CPML
// slabs, and no injected E element lies on the NTFF surface.  H does not change in
// this phase.
#pragma once

namespace fdtd {

template <typename T, int NTFF>
__global__ void __launch_bounds__(256) k_e_tail(Grid<T> g, FaceSet<T> fs, const CpmlTile *__restrict__ tiles,
                                                int ports, int nntff, PortState ps, FarfieldState ff,
                                                NtffBlock<T> nb, int row, const int *tsp, int off) {
  int b = (int)blockIdx.x;
  if (ports) {
    if (b == 0) {
      port_epilogue_body<T>(g, ps, step_of(tsp, off));
      return;
    }
    --b;
  }
  if (b < nntff) {
    const long q = (long)b * blockDim.x + threadIdx.x;
    if (NTFF == 1) ntff_direct_body<T>(g, ff, step_of(tsp, off), q);
    if (NTFF == 2) ntff_gather_body<T>(g, ff, nb, row, q);
    return;
  }
  cpml_e_tile<T>(g, fs, tiles[b - nntff], (int)threadIdx.x);
}

}  // namespace fdtd
