// learned.h -- shapes shared by the learned-ordering kernels and their host
// side: column sketches and the MLP parameter layout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// sketch of one column: null fraction, min, max, then the histogram
#define SKETCH_SAMPLE 1024
#define SKETCH_BINS 16
#define SKETCH_DIM (3 + SKETCH_BINS)

namespace bsql_learned {

struct SketchCol {
  const void* data;        // already advanced to the view's first row
  const uint32_t* valid;   // whole words; vbit = remaining bit offset
  int32_t vbit;
  int32_t type;
};

// Parameters in one buffer: w1 (in x hidden), b1, w2 (hidden x hidden), b2,
// w3 (hidden x out), b3; weights row-major by input.
struct MlpShape {
  int in, hidden, out;
  int off_w1, off_b1, off_w2, off_b2, off_w3, off_b3, n_params;
  static MlpShape make(int in, int hidden, int out) {
    MlpShape s{in, hidden, out, 0, 0, 0, 0, 0, 0, 0};
    s.off_w1 = 0;
    s.off_b1 = s.off_w1 + in * hidden;
    s.off_w2 = s.off_b1 + hidden;
    s.off_b2 = s.off_w2 + hidden * hidden;
    s.off_w3 = s.off_b2 + hidden;
    s.off_b3 = s.off_w3 + hidden * out;
    s.n_params = s.off_b3 + out;
    return s;
  }
};

hipError_t launch_sketch(const SketchCol* d_cols, int n_cols, int64_t rows, float* d_out, hipStream_t stream);
hipError_t mlp_forward(const MlpShape& s, const float* params, const float* x, int batch, float* h1, float* h2,
                       float* y, hipStream_t stream);
hipError_t mlp_backward(const MlpShape& s, const float* params, const float* x, const float* h1, const float* h2,
                        const float* y, const float* target, int batch, float cost_weight, float* dy, float* dh2,
                        float* dh1, float* grads, float* loss, hipStream_t stream);
hipError_t mlp_adam(float* params, const float* grads, float* m, float* v, int n, float lr, int step,
                    hipStream_t stream);

}  // namespace bsql_learned
