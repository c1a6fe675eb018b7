// learned.hip -- gfx950 kernels of the learned conjunct ordering: per-batch
// column sketches and a small FP32 MLP (forward, backward, Adam) that
// predicts a range conjunct's selectivity and cost from them.
//
// The network is tiny (24 -> 64 -> 64 -> 2), so the kernels favour plain,
// deterministic code over throughput: one thread per output element, fixed
// summation order (no atomics), so the gradients are reproducible and can be
// compared element by element with a host implementation.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bsql/interp.h"
#include "learned.h"

namespace bsql_learned {

// ------------------------------------------------------------------ sketch

__device__ __forceinline__ double read_value(const void* data, int type, int64_t r) {
  switch (type) {
    case BSQL_INT8: return (double)((const int8_t*)data)[r];
    case BSQL_INT16: return (double)((const int16_t*)data)[r];
    case BSQL_INT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return (double)((const int32_t*)data)[r];
    case BSQL_UINT8: case BSQL_BOOL8: return (double)((const uint8_t*)data)[r];
    case BSQL_UINT16: return (double)((const uint16_t*)data)[r];
    case BSQL_UINT32: return (double)((const uint32_t*)data)[r];
    case BSQL_FLOAT32: return (double)((const float*)data)[r];
    case BSQL_FLOAT64: return ((const double*)data)[r];
    default: return (double)((const int64_t*)data)[r];
  }
}

// One workgroup of SKETCH_SAMPLE threads per column: thread i reads row
// i * rows / SKETCH_SAMPLE. Output per column: null fraction, min, max
// (valid sample values) and the SKETCH_BINS-bin histogram of the valid
// sample over [min, max], as fractions of the valid sample.
__global__ __launch_bounds__(SKETCH_SAMPLE) void sketch_kernel(const SketchCol* __restrict__ cols, int64_t rows,
                                                               float* __restrict__ out) {
  __shared__ double s_v[SKETCH_SAMPLE];
  __shared__ int s_ok[SKETCH_SAMPLE];
  __shared__ double s_min, s_max;
  __shared__ int s_bins[SKETCH_BINS];
  const SketchCol c = cols[blockIdx.x];
  const int i = threadIdx.x;
  const int64_t r = rows > 0 ? (int64_t)i * rows / SKETCH_SAMPLE : 0;
  const int64_t b = (int64_t)c.vbit + r;
  const bool ok = rows > 0 && (!c.valid || ((c.valid[b >> 5] >> (b & 31)) & 1u));
  s_v[i] = ok ? read_value(c.data, c.type, r) : 0.0;
  s_ok[i] = ok ? 1 : 0;
  if (i < SKETCH_BINS) s_bins[i] = 0;
  __syncthreads();
  if (i == 0) {
    // fixed-order scan: the sketch is a pure function of the sample
    double mn = 0.0, mx = 0.0;
    int n = 0;
    for (int k = 0; k < SKETCH_SAMPLE; ++k) {
      if (!s_ok[k]) continue;
      const double v = s_v[k];
      if (n == 0 || v < mn) mn = v;
      if (n == 0 || v > mx) mx = v;
      ++n;
    }
    s_min = mn;
    s_max = mx;
    float* o = out + (size_t)blockIdx.x * SKETCH_DIM;
    o[0] = rows > 0 ? 1.0f - (float)n / (float)SKETCH_SAMPLE : 0.0f;
    o[1] = (float)mn;
    o[2] = (float)mx;
  }
  __syncthreads();
  if (ok) {
    const double span = s_max - s_min;
    int bin = span > 0.0 ? (int)((s_v[i] - s_min) / span * SKETCH_BINS) : 0;
    bin = bin < 0 ? 0 : (bin >= SKETCH_BINS ? SKETCH_BINS - 1 : bin);
    atomicAdd(&s_bins[bin], 1);  // integer counts: order does not matter
  }
  __syncthreads();
  if (i < SKETCH_BINS) {
    int n = 0;
    for (int k = 0; k < SKETCH_SAMPLE; ++k) n += s_ok[k];
    out[(size_t)blockIdx.x * SKETCH_DIM + 3 + i] = n ? (float)s_bins[i] / (float)n : 0.0f;
  }
}

// --------------------------------------------------------------------- MLP

// y[b, o] = act(sum_i x[b, i] * w[i, o] + bias[o])
__global__ void dense_forward(const float* __restrict__ x, const float* __restrict__ w,
                              const float* __restrict__ bias, float* __restrict__ y, int batch, int in,
                              int out, int relu) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= batch * out) return;
  const int b = t / out, o = t % out;
  float acc = bias[o];
  for (int i = 0; i < in; ++i) acc += x[(size_t)b * in + i] * w[(size_t)i * out + o];
  y[t] = relu && acc < 0.0f ? 0.0f : acc;
}

// Loss gradient at the output (batch x 2): column 0 is the selectivity
// logit (binary cross-entropy against the observed fraction), column 1 the
// log cost (squared error). Mean over the batch; per-sample losses in loss.
__global__ void output_grad(const float* __restrict__ y, const float* __restrict__ target,
                            float* __restrict__ dy, float* __restrict__ loss, int batch, float cost_weight) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const float z = y[2 * b], t = target[2 * b];
  const float p = 1.0f / (1.0f + expf(-z));
  const float d = y[2 * b + 1] - target[2 * b + 1];
  dy[2 * b] = (p - t) / (float)batch;
  dy[2 * b + 1] = cost_weight * 2.0f * d / (float)batch;
  // -t log p - (1 - t) log(1 - p), written stably
  const float bce = fmaxf(z, 0.0f) - z * t + log1pf(expf(-fabsf(z)));
  loss[b] = bce + cost_weight * d * d;
}

// dw[i, o] = sum_b x[b, i] * dy[b, o]; db[o] = sum_b dy[b, o]
__global__ void dense_backward_w(const float* __restrict__ x, const float* __restrict__ dy,
                                 float* __restrict__ dw, float* __restrict__ db, int batch, int in, int out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < in * out) {
    const int i = t / out, o = t % out;
    float acc = 0.0f;
    for (int b = 0; b < batch; ++b) acc += x[(size_t)b * in + i] * dy[(size_t)b * out + o];
    dw[t] = acc;
  } else if (t < in * out + out) {
    const int o = t - in * out;
    float acc = 0.0f;
    for (int b = 0; b < batch; ++b) acc += dy[(size_t)b * out + o];
    db[o] = acc;
  }
}

// dx[b, i] = (sum_o dy[b, o] * w[i, o]) * relu'(x_out[b, i]) where x_out is
// the ReLU output of the layer below (0 where it was clamped)
__global__ void dense_backward_x(const float* __restrict__ dy, const float* __restrict__ w,
                                 const float* __restrict__ act, float* __restrict__ dx, int batch, int in,
                                 int out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= batch * in) return;
  const int b = t / in, i = t % in;
  float acc = 0.0f;
  for (int o = 0; o < out; ++o) acc += dy[(size_t)b * out + o] * w[(size_t)i * out + o];
  dx[t] = act[t] > 0.0f ? acc : 0.0f;
}

__global__ void adam_step(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

}  // namespace bsql_learned

// ------------------------------------------------------------------ launchers

namespace bsql_learned {

static unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_sketch(const SketchCol* d_cols, int n_cols, int64_t rows, float* d_out, hipStream_t stream) {
  if (n_cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(sketch_kernel, dim3((unsigned)n_cols), dim3(SKETCH_SAMPLE), 0, stream, d_cols, rows, d_out);
  return hipGetLastError();
}

hipError_t mlp_forward(const MlpShape& s, const float* params, const float* x, int batch, float* h1, float* h2,
                       float* y, hipStream_t stream) {
  const float* w1 = params + s.off_w1;
  const float* b1 = params + s.off_b1;
  const float* w2 = params + s.off_w2;
  const float* b2 = params + s.off_b2;
  const float* w3 = params + s.off_w3;
  const float* b3 = params + s.off_b3;
  hipLaunchKernelGGL(dense_forward, dim3(blocks_for((int64_t)batch * s.hidden)), dim3(256), 0, stream, x, w1, b1,
                     h1, batch, s.in, s.hidden, 1);
  hipLaunchKernelGGL(dense_forward, dim3(blocks_for((int64_t)batch * s.hidden)), dim3(256), 0, stream, h1, w2, b2,
                     h2, batch, s.hidden, s.hidden, 1);
  hipLaunchKernelGGL(dense_forward, dim3(blocks_for((int64_t)batch * s.out)), dim3(256), 0, stream, h2, w3, b3, y,
                     batch, s.hidden, s.out, 0);
  return hipGetLastError();
}

hipError_t mlp_backward(const MlpShape& s, const float* params, const float* x, const float* h1, const float* h2,
                        const float* y, const float* target, int batch, float cost_weight, float* dy, float* dh2,
                        float* dh1, float* grads, float* loss, hipStream_t stream) {
  const float* w2 = params + s.off_w2;
  const float* w3 = params + s.off_w3;
  hipLaunchKernelGGL(output_grad, dim3(blocks_for(batch)), dim3(256), 0, stream, y, target, dy, loss, batch,
                     cost_weight);
  hipLaunchKernelGGL(dense_backward_w, dim3(blocks_for((int64_t)s.hidden * s.out + s.out)), dim3(256), 0, stream,
                     h2, dy, grads + s.off_w3, grads + s.off_b3, batch, s.hidden, s.out);
  hipLaunchKernelGGL(dense_backward_x, dim3(blocks_for((int64_t)batch * s.hidden)), dim3(256), 0, stream, dy, w3,
                     h2, dh2, batch, s.hidden, s.out);
  hipLaunchKernelGGL(dense_backward_w, dim3(blocks_for((int64_t)s.hidden * s.hidden + s.hidden)), dim3(256), 0,
                     stream, h1, dh2, grads + s.off_w2, grads + s.off_b2, batch, s.hidden, s.hidden);
  hipLaunchKernelGGL(dense_backward_x, dim3(blocks_for((int64_t)batch * s.hidden)), dim3(256), 0, stream, dh2, w2,
                     h1, dh1, batch, s.hidden, s.hidden);
  hipLaunchKernelGGL(dense_backward_w, dim3(blocks_for((int64_t)s.in * s.hidden + s.hidden)), dim3(256), 0,
                     stream, x, dh1, grads + s.off_w1, grads + s.off_b1, batch, s.in, s.hidden);
  return hipGetLastError();
}

hipError_t mlp_adam(float* params, const float* grads, float* m, float* v, int n, float lr, int step,
                    hipStream_t stream) {
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.0f, c2 = 1.0f;
  for (int i = 0; i < step; ++i) { c1 *= b1; c2 *= b2; }
  hipLaunchKernelGGL(adam_step, dim3(blocks_for(n)), dim3(256), 0, stream, params, grads, m, v, n, lr, b1, b2,
                     1e-8f, 1.0f - c1, 1.0f - c2);
  return hipGetLastError();
}

}  // namespace bsql_learned
