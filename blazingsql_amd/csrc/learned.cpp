// learned.cpp -- C ABI of the learned conjunct ordering: column sketches,
// the FP32 MLP (create, forward, gradients, Adam training step), conjunct
// features, and the ordering it drives. The plain static-order path never
// depends on any of this: an order only permutes the conjuncts, whose
// conjunction (the selected rows) is the same in every order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/bsql/interp.h"

namespace bsql {
bool process_exiting();
}  // namespace bsql
#include "capi_internal.h"
#include "learned.h"
#include "program.hpp"

using namespace bsql_learned;
using namespace bsql;

struct bsql_mlp {
  MlpShape shape;
  float* params = nullptr;  // device
  float* grads = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  int step = 0;
  // activations of the last forward (resized per batch)
  float *h1 = nullptr, *h2 = nullptr, *y = nullptr, *dy = nullptr, *dh1 = nullptr, *dh2 = nullptr;
  int cap = 0;
  ~bsql_mlp() {
    if (bsql::process_exiting()) return;
    for (float* p : {params, grads, m, v, h1, h2, y, dy, dh1, dh2})
      if (p) (void)hipFree(p);
  }
  hipError_t reserve(int batch) {
    if (batch <= cap) return hipSuccess;
    for (float** p : {&h1, &h2, &y, &dy, &dh1, &dh2}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    const size_t H = (size_t)shape.hidden * batch, O = (size_t)shape.out * batch;
    hipError_t e = hipMalloc((void**)&h1, H * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h2, H * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&dh1, H * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&dh2, H * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&y, O * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&dy, O * 4);
    if (e == hipSuccess) cap = batch;
    return e;
  }
};

#define BSQL_TRY(...)                                              \
  try {                                                            \
    __VA_ARGS__                                                    \
  } catch (const bsql::Error& e) {                                 \
    return set_error(e.code, e.what());                            \
  } catch (const std::bad_alloc&) {                                \
    return set_error(BSQL_ERR_OOM, "host allocation failed");      \
  } catch (const std::exception& e) {                              \
    return set_error(BSQL_ERR_INVALID, e.what());                  \
  }

extern "C" {

int bsql_sketch(const bsql_column* cols, int n_cols, int64_t num_rows, float* d_out, void* stream) {
  BSQL_TRY({
    if (n_cols < 0 || (n_cols > 0 && (!cols || !d_out)) || num_rows < 0) fail("invalid arguments");
    std::vector<SketchCol> sc((size_t)n_cols);
    for (int k = 0; k < n_cols; ++k) {
      const bsql_column& c = cols[k];
      if (c.type_id == BSQL_STRING || type_size(c.type_id) == 0) fail("sketches are of fixed-width columns");
      if (c.size < num_rows) fail("column shorter than num_rows");
      sc[(size_t)k].data = (const char*)c.data + c.offset * type_size(c.type_id);
      sc[(size_t)k].valid = c.validity ? c.validity + (c.offset >> 5) : nullptr;
      sc[(size_t)k].vbit = (int32_t)(c.offset & 31);
      sc[(size_t)k].type = c.type_id;
    }
    SketchCol* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sc.size() * sizeof(SketchCol) + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(d, sc.data(), sc.size() * sizeof(SketchCol), hipMemcpyHostToDevice,
                                            (hipStream_t)stream);
    if (e == hipSuccess) e = launch_sketch(d, n_cols, num_rows, d_out, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (d) (void)hipFree(d);
    if (e != hipSuccess) return device_error(e, "bsql_sketch");
    return BSQL_OK;
  })
}

int bsql_mlp_create(uint64_t seed, bsql_mlp** out) {
  BSQL_TRY({
    if (!out) fail("invalid arguments");
    auto m = std::make_unique<bsql_mlp>();
    m->shape = MlpShape::make(BSQL_FEATURE_DIM, BSQL_MLP_HIDDEN, 2);
    const int n = m->shape.n_params;
    // He-style uniform init for the ReLU layers, small last layer
    std::vector<float> p((size_t)n, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int count, float scale) {
      std::uniform_real_distribution<float> u(-1.0f, 1.0f);
      const float a = scale * std::sqrt(6.0f / (float)fan_in);
      for (int i = 0; i < count; ++i) p[(size_t)(off + i)] = a * u(rng);
    };
    const MlpShape& s = m->shape;
    fill(s.off_w1, s.in, s.in * s.hidden, 1.0f);
    fill(s.off_w2, s.hidden, s.hidden * s.hidden, 1.0f);
    fill(s.off_w3, s.hidden, s.hidden * s.out, 0.3f);
    hipError_t e = hipSuccess;
    for (float** q : {&m->params, &m->grads, &m->m, &m->v})
      if (e == hipSuccess) e = hipMalloc((void**)q, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpy(m->params, p.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m, 0, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemset(m->v, 0, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemset(m->grads, 0, (size_t)n * 4);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_create");
    *out = m.release();
    return BSQL_OK;
  })
}

int bsql_mlp_num_params(const bsql_mlp* m) { return m ? m->shape.n_params : -1; }

int bsql_mlp_get_params(const bsql_mlp* m, float* host) {
  BSQL_TRY({
    if (!m || !host) fail("invalid arguments");
    hipError_t e = hipMemcpy(host, m->params, (size_t)m->shape.n_params * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_get_params");
    return BSQL_OK;
  })
}

int bsql_mlp_set_params(bsql_mlp* m, const float* host) {
  BSQL_TRY({
    if (!m || !host) fail("invalid arguments");
    hipError_t e = hipMemcpy(m->params, host, (size_t)m->shape.n_params * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_set_params");
    m->step = 0;
    const size_t bytes = (size_t)m->shape.n_params * 4;
    if (hipMemset(m->m, 0, bytes) != hipSuccess || hipMemset(m->v, 0, bytes) != hipSuccess)
      fail("resetting optimizer state failed");
    return BSQL_OK;
  })
}

int bsql_mlp_forward(bsql_mlp* m, const float* d_x, int batch, float* d_y, void* stream) {
  BSQL_TRY({
    if (!m || !d_x || !d_y || batch <= 0) fail("invalid arguments");
    hipError_t e = m->reserve(batch);
    if (e == hipSuccess) e = mlp_forward(m->shape, m->params, d_x, batch, m->h1, m->h2, d_y, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_forward");
    return BSQL_OK;
  })
}

static hipError_t forward_backward(bsql_mlp* m, const float* d_x, const float* d_target, int batch,
                                   float cost_weight, float* d_loss, hipStream_t stream) {
  hipError_t e = m->reserve(batch);
  if (e == hipSuccess) e = mlp_forward(m->shape, m->params, d_x, batch, m->h1, m->h2, m->y, stream);
  if (e == hipSuccess)
    e = mlp_backward(m->shape, m->params, d_x, m->h1, m->h2, m->y, d_target, batch, cost_weight, m->dy, m->dh2,
                     m->dh1, m->grads, d_loss, stream);
  return e;
}

int bsql_mlp_gradients(bsql_mlp* m, const float* d_x, const float* d_target, int batch, float cost_weight,
                       float* d_grads, float* d_loss, void* stream) {
  BSQL_TRY({
    if (!m || !d_x || !d_target || !d_grads || !d_loss || batch <= 0) fail("invalid arguments");
    hipError_t e = forward_backward(m, d_x, d_target, batch, cost_weight, d_loss, (hipStream_t)stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_grads, m->grads, (size_t)m->shape.n_params * 4, hipMemcpyDeviceToDevice,
                         (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_gradients");
    return BSQL_OK;
  })
}

int bsql_mlp_train_step(bsql_mlp* m, const float* d_x, const float* d_target, int batch, float lr,
                        float cost_weight, float* d_loss, void* stream) {
  BSQL_TRY({
    if (!m || !d_x || !d_target || !d_loss || batch <= 0) fail("invalid arguments");
    hipError_t e = forward_backward(m, d_x, d_target, batch, cost_weight, d_loss, (hipStream_t)stream);
    if (e == hipSuccess)
      e = mlp_adam(m->params, m->grads, m->m, m->v, m->shape.n_params, lr, ++m->step, (hipStream_t)stream);
    if (e != hipSuccess) return device_error(e, "bsql_mlp_train_step");
    return BSQL_OK;
  })
}

void bsql_mlp_free(bsql_mlp* m) { delete m; }

int bsql_conjunct_features(const bsql_program* cpred, const int32_t* in_types, int n_in, const float* sketch,
                           float* features, int cap) {
  try {
    if (!cpred || !in_types || !sketch || n_in <= 0) fail("invalid arguments");
    const Program& pred = *cpred;
    const std::vector<int32_t> st = slot_types_for(pred, in_types, n_in);
    std::vector<int32_t> table(in_types, in_types + n_in);
    const std::vector<DevRange> rs = conjunct_ranges(pred, table.data(), n_in);
    const int n = (int)rs.size();
    if (!features || cap < n) return n;
    for (int j = 0; j < n; ++j) {
      const DevRange& r = rs[(size_t)j];
      const float* sk = sketch + (size_t)r.col * SKETCH_DIM;
      float* f = features + (size_t)j * BSQL_FEATURE_DIM;
      const double mn = sk[1], mx = sk[2], span = mx - mn;
      auto norm = [&](double v, double none) {
        if (!(span > 0.0)) return v < mn ? -0.5 : (v > mx ? 1.5 : 0.5);
        double t = (v - mn) / span;
        if (std::isnan(t)) return none;
        return std::min(1.5, std::max(-0.5, t));
      };
      double lo, hi;
      bool has_lo, has_hi;
      if (r.kind & RK_F) {
        std::memcpy(&lo, &r.lo, 8);
        std::memcpy(&hi, &r.hi, 8);
        has_lo = lo != -std::numeric_limits<double>::infinity();
        has_hi = hi != std::numeric_limits<double>::infinity();
      } else {
        const double mul = r.mul ? (double)r.mul : 1.0;
        lo = (double)r.lo / mul;
        hi = (double)r.hi / mul;
        has_lo = r.lo != INT64_MIN;
        has_hi = r.hi != INT64_MAX;
      }
      f[0] = sk[0];
      for (int b = 0; b < SKETCH_BINS; ++b) f[1 + b] = sk[3 + b];
      f[17] = (float)((r.kind & RK_NE) ? norm(lo, 0.5) : (has_lo ? norm(lo, -0.5) : -0.5));
      f[18] = (float)((r.kind & RK_NE) ? norm(lo, 0.5) : (has_hi ? norm(hi, 1.5) : 1.5));
      f[19] = (r.kind & RK_NE) ? 1.0f : 0.0f;
      f[20] = (r.kind & RK_F) ? 1.0f : 0.0f;
      f[21] = (float)type_size(table[(size_t)r.col]) / 8.0f;
      f[22] = (r.kind & RK_EMPTY) ? 1.0f : 0.0f;
      // the sketch's own estimate: the histogram's mass over [lo, hi] (a
      // uniform density inside each bin), of the non-null rows
      double est;
      if (r.kind & RK_EMPTY) {
        est = 0.0;
      } else if (r.kind & RK_NE) {
        est = 1.0;
      } else {
        const double a = f[17], b = f[18];
        est = 0.0;
        for (int k = 0; k < SKETCH_BINS; ++k) {
          const double l = (double)k / SKETCH_BINS, h = (double)(k + 1) / SKETCH_BINS;
          const double ov = std::min(b, h) - std::max(a, l);
          if (ov > 0.0) est += sk[3 + k] * ov * SKETCH_BINS;
        }
        if (!(span > 0.0)) est = (a <= 0.5 && b >= 0.5) ? 1.0 : 0.0;
      }
      f[23] = (float)(est * (1.0 - (double)sk[0]));
    }
    (void)st;
    return n;
  } catch (const bsql::Error& e) {
    return -set_error(e.code, e.what());
  } catch (const std::exception& e) {
    return -set_error(BSQL_ERR_INVALID, e.what());
  }
}

int bsql_order_by_prediction(bsql_program* pred, const float* sel, const float* cost, int n_conj) {
  BSQL_TRY({
    if (!pred || !sel || !cost || n_conj <= 0) fail("invalid arguments");
    // cheap, selective conjuncts first: ascending cost / (1 - selectivity)
    std::vector<int32_t> order((size_t)n_conj);
    std::iota(order.begin(), order.end(), 0);
    auto rank = [&](int j) { return (double)cost[j] / std::max(1e-3, 1.0 - (double)sel[j]); };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rank(a) < rank(b); });
    return bsql_program_set_conjunct_order(pred, order.data(), n_conj);
  })
}

int bsql_order_conjuncts(bsql_program* pred, const int32_t* in_types, int n_in, const float* sel,
                         const float* cost, int n_conj, int32_t* order_out) {
  BSQL_TRY({
    if (!pred || !in_types || n_in <= 0 || !sel || !cost || n_conj <= 0) fail("invalid arguments");
    std::vector<int32_t> order((size_t)n_conj);
    std::iota(order.begin(), order.end(), 0);
    std::vector<DevRange> rs;
    try {
      rs = conjunct_ranges(*pred, in_types, n_in);
    } catch (const Error&) {
      rs.clear();  // not all range tests: the rank rule below
    }
    if ((int)rs.size() == n_conj && n_conj <= 16) {
      // Expected cost of an order, the conjuncts' selectivities taken as
      // independent: conjunct j after a prefix passing fraction p of the rows
      // costs cost_j * (lines(p) + kLatency * slices(p)), where lines(p) is
      // the fraction of its 128-byte lines still holding a live row (its
      // loads are predicated per lane) and slices(p) the fraction of
      // 1024-row slices still alive (one memory round trip per range per
      // live slice). Exact minimum over all orders by dynamic programming on
      // the set of conjuncts placed first.
      constexpr double kLatency = 0.2;
      const int k = n_conj;
      std::vector<int> esc((size_t)k);  // log2 of the element size: 128 >> esc rows per line
      for (int j = 0; j < k; ++j) {
        const int es = type_size(in_types[rs[(size_t)j].col]);
        esc[(size_t)j] = es >= 8 ? 3 : (es >= 4 ? 2 : (es >= 2 ? 1 : 0));
      }
      const size_t full = (size_t)1 << k;
      std::vector<double> pass(full, 1.0), best(full, 0.0);
      std::vector<int8_t> last(full, -1);
      // per subset placed first: lines(p) for 128, 64, 32, 16 rows per line
      // and slices(p), p = its pass fraction (powers by repeated squaring)
      std::vector<double> fac((size_t)5 * full);
      auto factors = [&](size_t m) {
        const double q = 1.0 - pass[m];
        double q16 = q * q;
        q16 *= q16;
        q16 *= q16;
        q16 *= q16;
        const double q32 = q16 * q16, q64 = q32 * q32, q128 = q64 * q64;
        double q1024 = q128 * q128;
        q1024 *= q1024;
        q1024 *= q1024;
        double* f = &fac[5 * m];
        f[0] = 1.0 - q128;
        f[1] = 1.0 - q64;
        f[2] = 1.0 - q32;
        f[3] = 1.0 - q16;
        f[4] = 1.0 - q1024;
      };
      factors(0);
      for (size_t m = 1; m < full; ++m) {
        const int j0 = __builtin_ctzll(m);
        pass[m] = pass[m & (m - 1)] * std::min(1.0, std::max(0.0, (double)sel[j0]));
        best[m] = std::numeric_limits<double>::infinity();
        for (int j = 0; j < k; ++j) {
          if (!((m >> j) & 1u)) continue;
          const size_t before = m & ~((size_t)1 << j);
          const double* f = &fac[5 * before];
          const double c = best[before] + std::max(0.0, (double)cost[j]) * (f[esc[(size_t)j]] + kLatency * f[4]);
          if (c < best[m]) {
            best[m] = c;
            last[m] = (int8_t)j;
          }
        }
        factors(m);
      }
      for (size_t m = full - 1, pos = (size_t)k; m; m &= ~((size_t)1 << last[m])) order[--pos] = last[m];
    } else {
      auto rank = [&](int j) { return (double)cost[j] / std::max(1e-3, 1.0 - (double)sel[j]); };
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rank(a) < rank(b); });
    }
    if (order_out) std::copy(order.begin(), order.end(), order_out);
    return bsql_program_set_conjunct_order(pred, order.data(), n_conj);
  })
}

}  // extern "C"
