// Inference of the early-stop network (architecture in fdtdhip/earlystop_net.py):
// four Conv1d(k=5, stride=2, pad=2) + ReLU layers of 64 channels over a [2][375]
// input, then Linear(1536 -> 512) + ReLU and Linear(512 -> 1376).  Weights are BF16
// bit patterns, activations and accumulation FP32.
#pragma once

#include <cstdint>

namespace fdtd {

constexpr int ES_IN_CH = 2, ES_CH = 64, ES_KW = 5, ES_HIDDEN = 512, ES_OUT = 1376;
constexpr int ES_LEN[5] = {375, 188, 94, 47, 24};

__device__ __forceinline__ float bf16f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// out[b][co][t] = relu(bias[co] + sum_{ci,k} w[co][ci][k] * in[b][ci][2t + k - 2]).
__global__ void __launch_bounds__(256) k_es_conv(const float *in, const uint16_t *w, const float *bias, float *out,
                                                 int batch, int cin, int lin, int lout) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)batch * ES_CH * lout) return;
  const int t = (int)(q % lout), co = (int)((q / lout) % ES_CH), b = (int)(q / ((long)lout * ES_CH));
  float acc = bias[co];
  const float *x = in + (long)b * cin * lin;
  const uint16_t *wc = w + (long)co * cin * ES_KW;
  for (int ci = 0; ci < cin; ++ci) {
#pragma unroll
    for (int k = 0; k < ES_KW; ++k) {
      const int s = 2 * t + k - ES_KW / 2;
      if (s >= 0 && s < lin) acc += bf16f(wc[ci * ES_KW + k]) * x[ci * lin + s];
    }
  }
  out[q] = acc > 0.f ? acc : 0.f;
}

// out[b][n] = act(bias[n] + sum_k w[n][k] * in[b][k]); one wavefront per output.
__global__ void __launch_bounds__(256) k_es_dense(const float *in, const uint16_t *w, const float *bias, float *out,
                                                  int batch, int K, int N, int relu) {
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= (long)batch * N) return;
  const int n = (int)(wave % N), b = (int)(wave / N);
  const float *x = in + (long)b * K;
  const uint16_t *wr = w + (long)n * K;
  float acc = 0.f;
  for (int k = lane; k < K; k += 64) acc += bf16f(wr[k]) * x[k];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) {
    const float v = acc + bias[n];
    out[(long)b * N + n] = relu && v < 0.f ? 0.f : v;
  }
}

}  // namespace fdtd
