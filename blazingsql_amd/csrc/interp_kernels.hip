// interp_kernels.hip -- gfx950 kernels of the row-parallel expression
// interpreter, its boolean-mask stream compaction, and the fused
// filter+project pass, with their host launchers. The device code shared
// with kernels generated for one program is in vm_device.h.
//
// Execution model (replaces transformKernel / InterpreterFunctor,
// engine/src/Interpreter/interpreter_ops.cuh:165-964):
//   * a workgroup of 256 lanes owns a tile of 256*R consecutive rows; lane t
//     owns rows [t*R, t*R+R) of the tile, so every column load is one
//     contiguous 4R- or 8R-byte vector per lane (128-bit loads for R >= 4
//     int32 or R >= 2 int64/double) and a wave reads one contiguous span;
//   * the program is wave-uniform: it is read from the kernel argument block
//     through the scalar cache and dispatched with scalar branches;
//   * the slot buffer of the reference (8 bytes per slot per row, in LDS) is a
//     register file reg[NREG][R] in VGPRs, indexed with the wave-uniform
//     register number (s_set_gpr_idx relative addressing), or, for batched
//     launches with BSQL_LAUNCH_LDS_REGS, in LDS at one address add per
//     operand; validity is one bit per register per row in a VGPR;
//   * compaction (filters): a selection kernel writes 64-bit ballots and a
//     count per wave slice and a total per workgroup block, with no order
//     between workgroups; slice_ids turns them into the selected row ids in
//     row order; the output pass gathers those rows, 64 per instruction.
//     Optionally (BSQL_LAUNCH_FUSED) a range-test filter runs as one kernel:
//     wave tiles taken from ticket counters, decoupled look-back for the
//     output position, output validity assembled in LDS and written as whole
//     words, with an AND/OR atomic pair for the two words a tile shares.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <tuple>

#include "vm_device.h"

namespace bsql_dev {

// ------------------------------------------------------- string outputs
//
// A string output is written by the VM as one slice per row (start byte |
// length << 32, into the chars of its source column). Two kernels turn the
// slices into an Arrow string column: offsets (exclusive scan of the lengths
// of valid rows, decoupled look-back across 4096-row tiles, offsets[count] =
// total), then the chars. Null rows get length 0.
constexpr int STR_R = 16;
constexpr int STR_TILE = BLOCK * STR_R;

__device__ __forceinline__ bool row_valid(const uint32_t* valid, int64_t row) {
  return !valid || ((gp<uint32_t>(valid)[row >> 5] >> (row & 31)) & 1u);
}

__global__ __launch_bounds__(BLOCK) void str_offsets_kernel(const uint64_t* __restrict__ slices,
                                                            const uint32_t* __restrict__ valid,
                                                            int32_t* __restrict__ offsets,
                                                            const int64_t* __restrict__ d_count,
                                                            int64_t count_h, CompactArgs ca) {
  __shared__ uint32_t s_ticket;
  __shared__ int64_t s_wave_total[WAVES];
  __shared__ int64_t s_prefix;
  if (threadIdx.x == 0) s_ticket = atomicAdd(ca.ticket, 1u);
  __syncthreads();
  const int64_t tile = s_ticket;
  int64_t count = count_h;
  if (d_count) {
    count = *d_count;
    if (count > ca.capacity) count = ca.capacity;
  }
  const int64_t tiles = count > 0 ? (count + STR_TILE - 1) / STR_TILE : 1;
  if (tile >= tiles) return;  // rows beyond the count: nothing to do
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t r0 = tile * STR_TILE + (int64_t)tid * STR_R;
  const GLOBAL uint64_t* sl = gp<uint64_t>(slices);
  uint32_t len[STR_R];
  int64_t mine = 0;
#pragma unroll
  for (int k = 0; k < STR_R; ++k) {
    const int64_t row = r0 + k;
    len[k] = (row < count && row_valid(valid, row)) ? (uint32_t)(sl[row] >> 32) : 0u;
    mine += len[k];
  }
  // inclusive scan of the lane sums over the wave
  int64_t incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t t = __shfl_up(incl, d, 64);
    if (lane >= d) incl += t;
  }
  if (lane == 63) s_wave_total[wave] = incl;
  __syncthreads();
  int64_t wave_off = 0, tile_total = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    const int64_t t = s_wave_total[w];
    if (w < wave) wave_off += t;
    tile_total += t;
  }
  if (wave == 0) {
    const int64_t excl = look_back(ca, tile, tile_total, lane);
    if (lane == 0) s_prefix = excl;
  }
  __syncthreads();
  int64_t pos = s_prefix + wave_off + incl - mine;
  GLOBAL int32_t* off = gpw<int32_t>(offsets);
#pragma unroll
  for (int k = 0; k < STR_R; ++k) {
    const int64_t row = r0 + k;
    if (row < count) off[row] = (int32_t)pos;
    pos += len[k];
  }
  if (tile == tiles - 1 && tid == BLOCK - 1) off[count] = (int32_t)(s_prefix + tile_total);
}

__global__ __launch_bounds__(BLOCK) void str_copy_kernel(const uint64_t* __restrict__ slices,
                                                         const uint32_t* __restrict__ valid,
                                                         const int32_t* __restrict__ offsets,
                                                         const uint8_t* __restrict__ src,
                                                         uint8_t* __restrict__ dst,
                                                         const int64_t* __restrict__ d_count,
                                                         int64_t count_h, int64_t capacity) {
  int64_t count = count_h;
  if (d_count) {
    count = *d_count;
    if (count > capacity) count = capacity;
  }
  const GLOBAL uint64_t* sl = gp<uint64_t>(slices);
  const GLOBAL int32_t* off = gp<int32_t>(offsets);
  const GLOBAL uint8_t* s = gp<uint8_t>(src);
  GLOBAL uint8_t* d = gpw<uint8_t>(dst);
  for (int64_t row = (int64_t)blockIdx.x * BLOCK + threadIdx.x; row < count;
       row += (int64_t)gridDim.x * BLOCK) {
    if (!row_valid(valid, row)) continue;
    const uint64_t v = sl[row];
    const uint32_t start = (uint32_t)v, n = (uint32_t)(v >> 32);
    const int64_t o = off[row];
    for (uint32_t i = 0; i < n; ++i) d[o + i] = s[start + i];
  }
}

__global__ void fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace bsql_dev

// ------------------------------------------------------------------ launchers

namespace bsql {

using namespace bsql_dev;

static int g_num_cus = 0;

// Workgroups of a persistent launch: enough to fill every CU several times.
static int64_t persistent_grid(int64_t tiles) {
  if (g_num_cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      g_num_cus = n;
    else
      g_num_cus = 256;
  }
  const int64_t cap = (int64_t)g_num_cus * 8;
  return tiles < cap ? tiles : cap;
}

// Workgroups of a persistent execute launch over num_rows rows, B lanes x R
// rows per tile (the interpreter's and a generated kernel's).
int64_t execute_grid(int64_t num_rows, int rows_per_lane, int block) {
  const int64_t tile = (int64_t)block * rows_per_lane;
  return persistent_grid((num_rows + tile - 1) / tile);
}

// Sets n words at p to v (stream-ordered; a kernel, so it may be captured).
hipError_t fill_words(uint64_t* p, int64_t n, uint64_t v, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n, v);
  return hipGetLastError();
}

// (NREG, R, B) variants compiled: register file size x rows per lane x
// workgroup size. 64-lane workgroups serve programs whose spill slots need
// more LDS than a 256-lane workgroup can have.
#define BSQL_VARIANTS(X) \
  X(8, 2, 256) X(8, 4, 256) X(16, 1, 256) X(16, 2, 256) X(16, 1, 64) X(16, 2, 64) X(8, 4, 64)
// the batched interpreter's variants with the register file in LDS (block *
// NREG * R * 8 bytes of it per workgroup)
#define BSQL_LF_VARIANTS(X) \
  X(8, 2, 256) X(16, 1, 256) X(8, 2, 64) X(8, 4, 64) X(16, 1, 64) X(16, 2, 64)

static int64_t regfile_lds_bytes(int nreg, int block, int rows) { return (int64_t)block * nreg * rows * 8; }

// LDS a workgroup may have for spill slots (the static arrays of the
// selection kernels take < 1 KB of the 64 KB).
static int64_t max_spill_lds() { return 60 * 1024; }

// Picks the variant: the smallest register file that holds the program, and
// the rows per lane requested (0 = the default for that register file), for
// workgroup size `block`; false if no variant holds it (or its spill slots).
static bool pick(int nreg, int rpt_req, int block, int n_spill, int* NR, int* RR, bool lf = false) {
  struct V { int n, r; };
  const V vs[] = {{8, 4}, {16, 2}, {8, 2}, {16, 1}};
  for (const V& v : vs) {
    if (v.n < nreg) continue;
    if (rpt_req > 0 && v.r != rpt_req) continue;
    bool compiled = false;
#define X(N, RPT, BL) if (N == v.n && RPT == v.r && BL == block) compiled = true;
    if (lf) {
      BSQL_LF_VARIANTS(X)
    } else {
      BSQL_VARIANTS(X)
    }
#undef X
    if (!compiled) continue;
    const int64_t lds = spill_lds_bytes(n_spill, block, v.r) + (lf ? regfile_lds_bytes(v.n, block, v.r) : 0);
    if (lds > max_spill_lds()) continue;
    *NR = v.n;
    *RR = v.r;
    return true;
  }
  if (rpt_req > 0) return pick(nreg, 0, block, n_spill, NR, RR, lf);
  return false;
}

bool variant_fits(int nreg, int rpt, int block, int n_spill, bool lf) {
  int NR, RR;
  if (!pick(nreg, rpt, block, n_spill, &NR, &RR, lf)) return false;
  return rpt == 0 || RR == rpt;
}

// Rows per lane of the variant a launch with this shape uses (0: none).
int variant_rows(int nreg, int rpt, int block, int n_spill, bool lf) {
  int NR, RR;
  return pick(nreg, rpt, block, n_spill, &NR, &RR, lf) ? RR : 0;
}

// LF: the register file in LDS after the spill slots
template <int NREG, int R, int B, bool LF = false>
static hipError_t launch_execute_t(const DevPlan* d_plan, int64_t num_rows, int n_spill, hipStream_t stream) {
  const int64_t tile = (int64_t)B * R;
  const int64_t grid = persistent_grid((num_rows + tile - 1) / tile);
  if (grid == 0) return hipSuccess;
  const unsigned lds = (unsigned)(spill_lds_bytes(n_spill, B, R) + (LF ? regfile_lds_bytes(NREG, B, R) : 0));
  hipLaunchKernelGGL((vm_rows<NREG, R, false, B, PlanSrc, LF>), dim3((unsigned)grid), dim3(B), lds, stream, d_plan,
                     (const uint32_t*)nullptr, (const int64_t*)nullptr, num_rows);
  return hipGetLastError();
}

// chunks per wave of the selection tiles: 4096-row tiles for the VM,
// 8192-row tiles for range tests
template <int R> struct VmChunks { static constexpr int S = 16 / R; };
// range selection: 4 rows per chunk; 2, 4 or 8 chunks per wave (8, 16 or 32
// rows per lane: sel_rpt 8, 16, 32), 4 by default
constexpr int RANGE_R = 4;
static int range_chunks_of(int sel_rpt) { return sel_rpt == 8 ? 2 : (sel_rpt == 32 ? 8 : 4); }

template <int NREG, int R, int B, bool LF = false>
static hipError_t launch_select_vm_t(const DevPlan* d_plan, const SelArgs& sa, int64_t grid, int n_spill,
                                     hipStream_t stream) {
  const unsigned lds = (unsigned)(spill_lds_bytes(n_spill, B, R) + (LF ? regfile_lds_bytes(NREG, B, R) : 0));
  hipLaunchKernelGGL((select_vm<NREG, R, VmChunks<R>::S, B, PlanSrc, LF>), dim3((unsigned)grid), dim3(B), lds,
                     stream, d_plan, sa);
  return hipGetLastError();
}

template <int R, int S>
static hipError_t launch_slice_ids(const SelArgs& sa, uint32_t* idx, int64_t* d_count, int64_t grid,
                                   hipStream_t stream) {
  hipLaunchKernelGGL((slice_ids<R, S>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, sa, idx, d_count);
  return hipGetLastError();
}

template <int NREG, int R, int B, bool LF = false>
static hipError_t launch_gather_t(const DevPlan* d_plan, int64_t num_rows, const uint32_t* idx,
                                  const int64_t* d_count, int64_t capacity, int n_spill, hipStream_t stream) {
  const int64_t tile = (int64_t)B * R;
  const int64_t grid = persistent_grid((num_rows + tile - 1) / tile);
  if (grid == 0) return hipSuccess;
  const unsigned lds = (unsigned)(spill_lds_bytes(n_spill, B, R) + (LF ? regfile_lds_bytes(NREG, B, R) : 0));
  hipLaunchKernelGGL((vm_rows<NREG, R, true, B, PlanSrc, LF>), dim3((unsigned)grid), dim3(B), lds, stream, d_plan,
                     idx, d_count, capacity);
  return hipGetLastError();
}

// `d_plan` is a device copy of `plan` (made by the caller on `stream`).
// lf: the variant with the register file in LDS (BSQL_LAUNCH_LDS_REGS); the
// VGPR one where no LDS variant holds the program.
hipError_t launch_execute(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt, int block,
                          hipStream_t stream, bool lf) {
  int NR, RR;
  if (lf && pick(nreg, rpt, block, plan.n_spill, &NR, &RR, true)) {
#define X(N, RPT, BL)                      \
  if (NR == N && RR == RPT && block == BL) \
    return launch_execute_t<N, RPT, BL, true>(d_plan, plan.num_rows, plan.n_spill, stream);
    BSQL_LF_VARIANTS(X)
#undef X
  }
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) return launch_execute_t<N, RPT, BL>(d_plan, plan.num_rows, plan.n_spill, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Rows per selection slice and ballots per slice (R * S) of the kernel a
// plan selects with.
static bool slice_shape(const DevPlan& plan, int nreg, int rpt, int block, int64_t* rows, int* ballots) {
  if (plan.n_ranges > 0) {
    const int chunks = range_chunks_of(rpt);
    *rows = (int64_t)64 * RANGE_R * chunks;
    *ballots = RANGE_R * chunks;
    return true;
  }
  int NR, RR;
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return false;
  *rows = 64 * 16;  // R * VmChunks<R>::S == 16 rows per lane
  *ballots = 16;
  return true;
}

// Workgroups of `kernel` that are resident on the device at once (its
// occupancy per CU x CUs), or 8 per CU when the runtime cannot say.
static int64_t resident_workgroups(const void* kernel, int block, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void*, int, size_t>, int64_t> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(std::make_tuple(dev, kernel, block, lds));
  if (it != cache.end()) return it->second;
  (void)persistent_grid(1);  // learns the CU count
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds) != hipSuccess || per_cu <= 0)
    per_cu = 8;
  const int64_t r = (int64_t)per_cu * g_num_cus;
  cache.emplace(std::make_tuple(dev, kernel, block, lds), r);
  return r;
}

// Workgroups of the selection launch (and of its slice_ids): no more than are
// resident at once, so that no second round of workgroups trails the first
// (each takes a contiguous block of slices, ceil(slices / grid) of them).
static int64_t selection_grid(const DevPlan& plan, int nreg, int sel_rpt, int block, int64_t n_slices, bool lf) {
  const void* k = nullptr;
  int wg = BLOCK;
  size_t lds = 0;
  int NR, RR;
  if (plan.n_ranges > 0) {
    switch (range_chunks_of(sel_rpt)) {
      case 2: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 2>); break;
      case 8: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 8>); break;
      default: k = reinterpret_cast<const void*>(&select_ranges<RANGE_R, 4>);
    }
  } else if (lf && pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR, true)) {
    wg = block;
    lds = (size_t)(spill_lds_bytes(plan.n_spill, block, RR) + regfile_lds_bytes(NR, block, RR));
#define X(N, RPT, BL)                      \
  if (NR == N && RR == RPT && block == BL) \
    k = reinterpret_cast<const void*>(&select_vm<N, RPT, VmChunks<RPT>::S, BL, PlanSrc, true>);
    BSQL_LF_VARIANTS(X)
#undef X
  } else {
    if (pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR)) {
      wg = block;
      lds = (size_t)spill_lds_bytes(plan.n_spill, block, RR);
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) k = reinterpret_cast<const void*>(&select_vm<N, RPT, VmChunks<RPT>::S, BL>);
      BSQL_VARIANTS(X)
#undef X
    }
  }
  const int waves = wg / 64;
  int64_t grid = persistent_grid((n_slices + waves - 1) / waves);
  if (k) {
    const int64_t r = resident_workgroups(k, wg, lds);
    if (grid > r) grid = r;
  }
  return grid;
}

// (lf: the selection kernel the launch will use, whose residency sizes the
// grid and so the block totals; the same value as the launch's)
int64_t select_scratch_words(const DevPlan& plan, int nreg, int rpt, int block, bool lf) {
  int64_t rows;
  int ballots;
  if (!slice_shape(plan, nreg, rpt, block, &rows, &ballots)) return -1;
  const int64_t n = (plan.num_rows + rows - 1) / rows;
  const int64_t grid = selection_grid(plan, nreg, rpt, block, n, lf);
  return n * ballots + (n + 2 * grid + 1) / 2;
}

// The output phase of a filter: ops [split, n_ops) over the selected row ids
// (written by the selection phase, count in *d_count).
static hipError_t launch_compact_gather(const DevPlan& plan, const DevPlan* d_plan, int nreg, int rpt, int block,
                                        const uint32_t* idx, const int64_t* d_count, int64_t capacity,
                                        hipStream_t stream, bool lf) {
  if (plan.split >= plan.n_ops) return hipSuccess;
  int NR, RR;
  if (lf && pick(nreg, rpt, block, plan.n_spill, &NR, &RR, true)) {
#define X(N, RPT, BL)                                                                                    \
  if (NR == N && RR == RPT && block == BL)                                                               \
    return launch_gather_t<N, RPT, BL, true>(d_plan, plan.num_rows < capacity ? plan.num_rows : capacity, \
                                             idx, d_count, capacity, plan.n_spill, stream);
    BSQL_LF_VARIANTS(X)
#undef X
  }
  if (!pick(nreg, rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL)                                                                                         \
  if (NR == N && RR == RPT && block == BL)                                                                    \
    return launch_gather_t<N, RPT, BL>(d_plan, plan.num_rows < capacity ? plan.num_rows : capacity, idx,      \
                                       d_count, capacity, plan.n_spill, stream);
  BSQL_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Filter (+ project): selection kernel -> ids of the selected rows in `idx`
// (num_rows entries) and their count in *d_count, then the output ops over
// those rows. `sel` holds select_scratch_words(...) words. `phases`: bit 0 =
// selection (+ scan + ids), bit 1 = output phase (a caller that sizes its
// outputs from the count runs them as two calls on one stream, with idx and
// *d_count kept in between). Nothing is cleared first: every slice writes
// its ballots and count, and the scan writes *d_count.
// Whether a filter (+ project) plan runs as the fused one-pass kernel: a
// range-test predicate, output ops that fit the 8-register file without
// spill slots, at most FUSED_MAX_OUT outputs.
bool fused_eligible(const DevPlan& plan, int nreg) {
  return plan.n_ranges > 0 && plan.split < plan.n_ops && plan.n_out <= FUSED_MAX_OUT && nreg <= 8 &&
         plan.n_spill == 0;
}

// Words of scratch the fused kernel uses: one status word per tile, the
// ticket counters.
int64_t fused_scratch_words(int64_t num_rows) { return (num_rows + FUSED_TILE - 1) / FUSED_TILE + 4; }

#define BSQL_FUSED_NREG 8
#define BSQL_FUSED_R 2
#define BSQL_FUSED_SH 4

static hipError_t launch_fused(const DevPlan& plan, const DevPlan* d_plan, uint64_t* words, int64_t* d_count,
                               int64_t capacity, hipStream_t stream) {
  const int64_t tiles = (plan.num_rows + FUSED_TILE - 1) / FUSED_TILE;
  const int64_t n_words = tiles + (FUSED_COUNTERS + 1) / 2;
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((n_words + 255) / 256)), dim3(256), 0, stream, words, n_words, 0ULL);
  if (tiles == 0) {
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, stream, (uint64_t*)d_count, (int64_t)1, 0ULL);
    return hipGetLastError();
  }
  FusedArgs fa{words, (uint32_t*)(words + tiles), d_count, capacity, tiles};
  const size_t lds = (size_t)fused_lds_bytes(plan.n_out);
  const void* k = reinterpret_cast<const void*>(
      &select_project_fused<BSQL_FUSED_NREG, BSQL_FUSED_R, BSQL_FUSED_SH>);
  int64_t grid = resident_workgroups(k, 64, lds);
  if (grid > tiles) grid = tiles;
  // every ticket counter has workgroups
  grid = (grid + FUSED_COUNTERS - 1) / FUSED_COUNTERS * FUSED_COUNTERS;
  hipLaunchKernelGGL((select_project_fused<BSQL_FUSED_NREG, BSQL_FUSED_R, BSQL_FUSED_SH>), dim3((unsigned)grid),
                     dim3(64), (unsigned)lds, stream, d_plan, fa);
  return hipGetLastError();
}

hipError_t launch_compact(const DevPlan& plan, const DevPlan* d_plan, int nreg, int sel_rpt, int rpt, int block,
                          uint64_t* sel, uint32_t* idx, int64_t* d_count, int64_t capacity,
                          hipStream_t stream, int phases, bool fused, bool lf) {
  if (fused && phases == 3) return launch_fused(plan, d_plan, sel, d_count, capacity, stream);
  if (!(phases & 1))
    return launch_compact_gather(plan, d_plan, nreg, rpt, block, idx, d_count, capacity, stream, lf);
  int64_t rows;
  int ballots;
  if (!slice_shape(plan, nreg, sel_rpt, block, &rows, &ballots)) return hipErrorInvalidValue;
  const int64_t n = (plan.num_rows + rows - 1) / rows;
  const int64_t grid = selection_grid(plan, nreg, sel_rpt, block, n, lf);
  uint32_t* counts = (uint32_t*)(sel + n * ballots);
  SelArgs sa{sel, counts, counts + n, counts + n + grid, n, grid ? (n + grid - 1) / grid : 0};
  hipError_t e = hipSuccess;
  const bool ids = plan.split < plan.n_ops;
  if (n > 0) {
    e = hipErrorInvalidValue;
    if (plan.n_ranges > 0) {
      switch (range_chunks_of(sel_rpt)) {
        case 2:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 2>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
          break;
        case 8:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 8>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
          break;
        default:
          hipLaunchKernelGGL((select_ranges<RANGE_R, 4>), dim3((unsigned)grid), dim3(BLOCK), 0, stream, d_plan, sa);
      }
      e = hipGetLastError();
    } else {
      int NR, RR;
      if (lf && pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR, true)) {
#define X(N, RPT, BL)                      \
  if (NR == N && RR == RPT && block == BL) \
    e = launch_select_vm_t<N, RPT, BL, true>(d_plan, sa, grid, plan.n_spill, stream);
        BSQL_LF_VARIANTS(X)
#undef X
      } else {
        if (!pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR)) return hipErrorInvalidValue;
#define X(N, RPT, BL) \
  if (NR == N && RR == RPT && block == BL) e = launch_select_vm_t<N, RPT, BL>(d_plan, sa, grid, plan.n_spill, stream);
        BSQL_VARIANTS(X)
#undef X
      }
    }
    if (e != hipSuccess) return e;
  }
  if (ids && n > 0) {
    // the ids kernel also sums the block totals and writes the row total
    if (plan.n_ranges > 0) {
      const int chunks = range_chunks_of(sel_rpt);
      e = chunks == 2 ? launch_slice_ids<RANGE_R, 2>(sa, idx, d_count, grid, stream)
          : chunks == 8 ? launch_slice_ids<RANGE_R, 8>(sa, idx, d_count, grid, stream)
                        : launch_slice_ids<RANGE_R, 4>(sa, idx, d_count, grid, stream);
    } else {
      // the ballot layout of the selection kernel that ran: its rows per lane
      int NR, RR;
      if (!(lf && pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR, true)) &&
          !pick(nreg, sel_rpt, block, plan.n_spill, &NR, &RR))
        return hipErrorInvalidValue;
      e = RR == 1 ? launch_slice_ids<1, VmChunks<1>::S>(sa, idx, d_count, grid, stream)
          : RR == 2 ? launch_slice_ids<2, VmChunks<2>::S>(sa, idx, d_count, grid, stream)
                    : launch_slice_ids<4, VmChunks<4>::S>(sa, idx, d_count, grid, stream);
    }
  } else {
    hipLaunchKernelGGL(block_scan, dim3(1), dim3(SCAN_THREADS), 0, stream, (const uint32_t*)sa.block_sums,
                       sa.block_base, n > 0 ? grid : (int64_t)0, d_count);
    e = hipGetLastError();
  }
  if (e != hipSuccess || !(phases & 2)) return e;
  return launch_compact_gather(plan, d_plan, nreg, rpt, block, idx, d_count, capacity, stream, lf);
}

// Batched launch: `plans` (device) holds n_batches plans, `tile_start`
// (device) n_batches + 1 prefix counts of tiles of block*rpt rows. No
// allocation and no copy: the launch can be captured in a hipGraph.
int batch_tile_rows(int nreg, int rpt, int block, bool lf) {
  int NR, RR;
  if (!pick(nreg, rpt, block, 0, &NR, &RR, lf)) return -1;
  return block * RR;
}

// lf: the variant with the register file in LDS
hipError_t launch_batches(const DevPlan* plans, const int64_t* tile_start, int n_batches, int64_t n_items,
                          int nreg, int rpt, int block, int n_spill, bool lf, hipStream_t stream) {
  int NR, RR;
  if (!pick(nreg, rpt, block, n_spill, &NR, &RR, lf)) return hipErrorInvalidValue;
  const int64_t grid = persistent_grid(n_items);
  if (grid == 0) return hipSuccess;
#define X(N, RPT, BL)                                                                                  \
  if (!lf && NR == N && RR == RPT && block == BL) {                                                    \
    hipLaunchKernelGGL((vm_batches<N, RPT, BL>), dim3((unsigned)grid), dim3(BL),                       \
                       (unsigned)spill_lds_bytes(n_spill, BL, RPT), stream, plans, tile_start, n_batches); \
    return hipGetLastError();                                                                          \
  }
  BSQL_VARIANTS(X)
#undef X
#define X(N, RPT, BL)                                                                                  \
  if (lf && NR == N && RR == RPT && block == BL) {                                                     \
    hipLaunchKernelGGL((vm_batches<N, RPT, BL, true>), dim3((unsigned)grid), dim3(BL),                 \
                       (unsigned)(spill_lds_bytes(n_spill, BL, RPT) + regfile_lds_bytes(N, BL, RPT)), stream, \
                       plans, tile_start, n_batches);                                                  \
    return hipGetLastError();                                                                          \
  }
  BSQL_LF_VARIANTS(X)
#undef X
  return hipErrorInvalidValue;
}

// Assembles string output columns from the slices the VM stored: offsets
// then chars. `status` holds string_tiles(capacity) + 1 words; `d_count`
// (device) or `count` (host) gives the rows.
int64_t string_tiles(int64_t capacity) { return capacity > 0 ? (capacity + STR_TILE - 1) / STR_TILE : 1; }

static void launch_offsets(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                           const int64_t* d_count, int64_t count, int64_t capacity, uint64_t* status,
                           hipStream_t stream) {
  const int64_t tiles = string_tiles(capacity);
  CompactArgs ca{status, (uint32_t*)(status + tiles), nullptr, capacity, tiles};
  hipLaunchKernelGGL(fill_u64, dim3((unsigned)((tiles + 1 + 255) / 256)), dim3(256), 0, stream, status,
                     tiles + 1, 0ULL);
  hipLaunchKernelGGL(str_offsets_kernel, dim3((unsigned)tiles), dim3(BLOCK), 0, stream, slices, valid, offsets,
                     d_count, count, ca);
}

// Offsets (count + 1 of them) of a string column from row lengths stored as
// length << 32 in `slices` (null rows: 0). `status`: string_tiles(count) + 1 words.
hipError_t launch_string_offsets(const uint64_t* slices, const uint32_t* valid, int32_t* offsets, int64_t count,
                                 uint64_t* status, hipStream_t stream) {
  launch_offsets(slices, valid, offsets, nullptr, count, count, status, stream);
  return hipGetLastError();
}

hipError_t launch_string_output(const uint64_t* slices, const uint32_t* valid, int32_t* offsets,
                                const uint8_t* src_chars, uint8_t* dst_chars, const int64_t* d_count,
                                int64_t count, int64_t capacity, uint64_t* status, hipStream_t stream) {
  launch_offsets(slices, valid, offsets, d_count, count, capacity, status, stream);
  const int64_t blocks = persistent_grid((capacity + BLOCK - 1) / BLOCK);
  if (blocks > 0)
    hipLaunchKernelGGL(str_copy_kernel, dim3((unsigned)blocks), dim3(BLOCK), 0, stream, slices, valid, offsets,
                       src_chars, dst_chars, d_count, count, capacity);
  return hipGetLastError();
}

}  // namespace bsql
