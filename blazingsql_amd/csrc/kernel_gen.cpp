// kernel_gen.cpp -- kernels generated for one program.
//
// The interpreter reads every op, column type and range from the launch plan
// and dispatches on them for every tile. For a program launched with
// BSQL_LAUNCH_GENERATED this file writes a small HIP source that
// instantiates the same device templates (vm_device.h) with a source of
// constant op fields, column types, validity presence and range kinds
// (literals, column addresses and range bounds still come from the plan at
// launch), compiles it with hipRTC for the device's architecture, and keeps
// the code object for the process. The compiler folds the dispatch away, so
// the same op semantics run straight-line with only the registers the
// program uses. The source text is the cache key: programs that differ only
// in literals, row counts or column addresses share one kernel.
#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <functional>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "program.hpp"

namespace bsql {

// interp_kernels.hip
int64_t execute_grid(int64_t num_rows, int rows_per_lane, int block);
hipError_t fill_words(uint64_t* p, int64_t n, uint64_t v, hipStream_t stream);

namespace {

// The device headers, as built into the library (gen_headers.inc is written
// by the build from vm_device.h, devplan.h, pow10.h and interp.h, system
// includes removed): kGenHeaderCount, kGenHeaderNames, kGenHeaderTexts.
#include "gen_headers.inc"

// What hipRTC does not predefine of the C headers the device code uses.
const char* kPrelude =
    "typedef __hip_internal::int8_t int8_t;\n"
    "typedef __hip_internal::uint8_t uint8_t;\n"
    "typedef __hip_internal::int16_t int16_t;\n"
    "typedef __hip_internal::uint16_t uint16_t;\n"
    "typedef __hip_internal::int32_t int32_t;\n"
    "typedef __hip_internal::uint32_t uint32_t;\n"
    "typedef __hip_internal::int64_t int64_t;\n"
    "typedef __hip_internal::uint64_t uint64_t;\n"
    "typedef __SIZE_TYPE__ size_t;\n"
    "typedef __UINTPTR_TYPE__ uintptr_t;\n"
    "#define INT64_MAX __INT64_MAX__\n"
    "#define INT64_MIN (-__INT64_MAX__ - 1)\n"
    "#define INT32_MAX __INT32_MAX__\n"
    "#define INT32_MIN (-__INT32_MAX__ - 1)\n";

// The one-pass filter + project kernel's tile: kTileWaves wave slices of
// kChunks chunks of 64 lanes x kRows rows (512-row slices, 4096-row tiles).
// Measured on the 100M-row scan (outputs staged in LDS) over (rows per
// lane, chunks, waves): few rows per lane keep the registers (and the kept
// column's values) small enough for many waves per CU, and 8-wave tiles
// keep the look-backs few; (2, 4, 8) ran fastest (0.566 ms against 0.60 to
// 0.62 for (2, 4, 16), 0.68 for (2, 8, 4), 0.70 for (4, 4, 4)).
constexpr int kRows = 2, kChunks = 4, kTileWaves = 8;
constexpr int kTile = 64 * kRows * kChunks * kTileWaves;

// A one-pass kernel's tile: `waves` wave slices of `chunks` chunks of 64
// lanes x `rows` rows.
struct TileShape {
  int rows, chunks, waves;
  int64_t slice() const { return 64LL * rows * chunks; }
  int64_t tile() const { return slice() * waves; }
};
constexpr TileShape kOnePass{kRows, kChunks, kTileWaves};
// The batched kernel's tile (8192 rows): every query's outputs load, compute
// and stage once per chunk, and a query's output phase waits on its loads
// before the next query's, so more rows per lane and fewer chunks keep more
// loads in flight per wait. Measured over 8 queries on 64M copula rows
// (rho 0 / 0.5 / 0.9, ms): (2, 4, 8) 3.70 / 3.84 / 3.71, (4, 2, 8) 2.94 /
// 3.13 / 2.95, (8, 1, 8) 3.00 / 3.16 / 2.91, (8, 4, 4) 2.57 / 2.87 / 2.71,
// (8, 2, 8) 2.52 / 2.91 / 2.66; staging 16 or 64 KB of rows instead of 32
// KB ran no faster.
constexpr TileShape kBatch{8, 2, 8};

// LDS for a tile's staged output rows (8-byte slots): 32 KB holds 2048 rows
// of two outputs, half of a tile; with the rest of the kernel's LDS a
// workgroup stays within 64 KB. Measured on the 100M-row scan (12 % of
// each 4096-row tile selected): staging 1024 or 2048 rows ran alike.
constexpr int64_t kStageBytes = 32768, kMaxLds = 65536;

struct Loaded {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

struct Kernel {
  std::string code;                 // code object
  std::map<int, Loaded> on_device;  // device -> module
};

std::mutex g_mu;
std::map<std::string, std::unique_ptr<Kernel>> g_kernels;  // name expression + source -> kernel
// check mode (bsql_generated_kernel_check): compile for this architecture,
// count the code bytes, load and launch nothing
thread_local const char* t_check_arch = nullptr;
thread_local int64_t t_check_bytes = 0;
int64_t g_compiled = 0, g_launches = 0;
double g_compile_seconds = 0;

[[noreturn]] void device_fail(const std::string& m) { throw Error(BSQL_ERR_DEVICE, m); }

std::string arch_of(int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) device_fail("generated kernel: no device properties");
  std::string a = prop.gcnArchName;
  const size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

std::string compile(const std::string& src, const std::string& name_expr, const std::string& arch,
                    std::string* lowered) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "bsql_generated.hip", kGenHeaderCount, kGenHeaderTexts,
                          kGenHeaderNames) != HIPRTC_SUCCESS)
    device_fail("generated kernel: hiprtcCreateProgram failed");
  struct Destroy {
    hiprtcProgram* p;
    ~Destroy() { hiprtcDestroyProgram(p); }
  } destroy{&prog};
  hiprtcAddNameExpression(prog, name_expr.c_str());
  const std::string arch_opt = "--offload-arch=" + arch;
  // the library's own build flags: one rounding per FP64 + - * /, no
  // packed FP32 instructions
  const char* opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-DBSQL_RTC=1"};
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    device_fail("generated kernel did not compile: " + log.substr(0, 2000));
  }
  const char* low = nullptr;
  if (hiprtcGetLoweredName(prog, name_expr.c_str(), &low) != HIPRTC_SUCCESS || !low)
    device_fail("generated kernel: no lowered name");
  *lowered = low;
  size_t size = 0;
  if (hiprtcGetCodeSize(prog, &size) != HIPRTC_SUCCESS || size == 0) device_fail("generated kernel: no code");
  std::string code(size, '\0');
  hiprtcGetCode(prog, &code[0]);
  return code;
}

// The kernel for (name expression, source) on the current device, compiled
// and loaded on first use.
// BSQL_GENERATED_DUMP=<dir>: every generated source is also written there
// (the kernel's name expression in its first line), for reading and for
// compiling it by hand.
void dump_source(const std::string& name_expr, const std::string& src) {
  const char* dir = std::getenv("BSQL_GENERATED_DUMP");
  if (!dir || !*dir) return;
  const std::string path = std::string(dir) + "/bsql_gen_" + std::to_string(std::hash<std::string>()(src) ^
                                                                            std::hash<std::string>()(name_expr)) +
                           ".hip";
  if (FILE* f = std::fopen(path.c_str(), "w")) {
    std::fprintf(f, "// %s\n%s", name_expr.c_str(), src.c_str());
    std::fclose(f);
  }
}

hipFunction_t kernel_for(const std::string& name_expr, const std::string& src) {
  dump_source(name_expr, src);
  if (t_check_arch) {
    std::string lowered;
    t_check_bytes += (int64_t)compile(src, name_expr, t_check_arch, &lowered).size();
    return nullptr;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) device_fail("generated kernel: no current device");
  std::lock_guard<std::mutex> lk(g_mu);
  const std::string key = name_expr + "\n" + src;
  auto it = g_kernels.find(key);
  if (it == g_kernels.end()) {
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<Kernel> k(new Kernel);
    std::string lowered;
    k->code = compile(src, name_expr, arch_of(dev), &lowered);
    k->code.append(1, '\0');
    k->code.append(lowered);  // the lowered name rides after the code object
    g_compile_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++g_compiled;
    it = g_kernels.emplace(key, std::move(k)).first;
  }
  Kernel& k = *it->second;
  auto d = k.on_device.find(dev);
  if (d == k.on_device.end()) {
    const size_t cut = k.code.rfind('\0');
    const std::string lowered = k.code.substr(cut + 1);
    Loaded l;
    if (hipModuleLoadData(&l.mod, k.code.data()) != hipSuccess) device_fail("generated kernel: module load failed");
    if (hipModuleGetFunction(&l.fn, l.mod, lowered.c_str()) != hipSuccess)
      device_fail("generated kernel: function not found");
    d = k.on_device.emplace(dev, l).first;
  }
  ++g_launches;
  return d->second.fn;
}

bool is_loadn_like(uint8_t code) { return code == D_LOADN || code == D_SPILL || code == D_FILL; }

// Input columns the ops [q0, q1) load.  (Range bounds stay launch arguments;
// a range's column, kind and unit factor `mul` are part of its shape.)
std::vector<bool> loaded_columns(const DevPlan& plan, const std::vector<DevOp>& ops, int q0, int q1) {
  std::vector<bool> used((size_t)plan.n_in, false);
  for (int q = q0; q < q1; ++q) {
    const DevOp& o = ops[(size_t)q];
    if (o.code == D_LOADN) {
      for (int j = 0; j < o.aux; ++j) {
        const int c = (int)((o.imm >> (8 * j)) & 0xff);
        if (c < plan.n_in) used[(size_t)c] = true;
      }
    } else if (o.code == D_LOADS || o.code == D_LOAD) {
      if (o.a < plan.n_in) used[o.a] = true;
    }
  }
  return used;
}

// The constant source of one program: columns, outputs, ranges and ops.
// nt[j]: loads of input column j are non-temporal.
// The element kind of the range tests over a column of type t.
const char* range_kind(int t) {
  switch (t) {
    case BSQL_INT8: return "K_I8";
    case BSQL_INT16: return "K_I16";
    case BSQL_INT32: case BSQL_TIMESTAMP_DAYS: case BSQL_DURATION_DAYS: return "K_I32";
    case BSQL_UINT8: return "K_U8";
    case BSQL_UINT16: return "K_U16";
    case BSQL_UINT32: return "K_U32";
    case BSQL_BOOL8: return "K_B8";
    case BSQL_FLOAT32: return "K_F32";
    case BSQL_FLOAT64: return "K_F64";
    default: return "K_I64";
  }
}

// eager: the range tests load every column before the first test (keep_col's
// values kept by its first range).
// stage_rows: rows of a one-pass tile's outputs staged in LDS; name: the
// struct's name (several programs in one kernel use one struct each).
std::string source_struct(const DevPlan& plan, const std::vector<DevOp>& ops, const std::vector<bool>& nt,
                          int keep_col = -1, bool eager = false, int stage_rows = 0, const char* name = "Src") {
  std::ostringstream s;
  const std::string ranges = std::string(name) + "Range";
  s << "namespace bsql_gen {\nusing namespace bsql_dev;\ntemplate <int r> struct " << ranges << " {};\n";
  bool kept = false;
  for (int r = 0; eager && r < plan.n_ranges; ++r) {
    const int c = plan.ranges[r].col;
    const bool keep_here = c == keep_col && !kept;
    kept = kept || keep_here;
    s << "template <> struct " << ranges << "<" << r << "> {\n  typedef " << range_kind(plan.in[c].type)
      << " K;\n  static constexpr int col = " << c << ";\n  static constexpr bool keep = "
      << (keep_here ? "true" : "false") << ";\n};\n";
  }
  s << "struct " << name << " {\n  static constexpr bool kPlanOps = false;\n"
    << "  static constexpr int kKeepCol = " << keep_col << ";\n"
    << "  static constexpr bool kEager = " << (eager ? "true" : "false") << ";\n"
    << "  static constexpr int kRanges = " << (eager ? plan.n_ranges : 0) << ";\n"
    << "  static constexpr int kStageRows = " << stage_rows << ";\n"
    << "  template <int r> using RangeAt = " << ranges << "<r>;\n";
  s << "  __device__ static __forceinline__ DevInCol in(const DevPlan& p, int j) {\n"
       "    DevInCol c = p.in[j];\n    switch (j) {\n";
  for (int j = 0; j < plan.n_in; ++j) {
    const DevInCol& c = plan.in[j];
    s << "      case " << j << ": c.type = " << c.type << "; c.aligned16 = " << (c.aligned16 ? 1 : 0) << ";";
    if (!c.valid) s << " c.valid = nullptr;";
    s << " break;\n";
  }
  s << "      default: break;\n    }\n    return c;\n  }\n";
  s << "  __device__ static __forceinline__ DevOutCol out(const DevPlan& p, int o) {\n"
       "    DevOutCol c = p.out[o];\n    switch (o) {\n";
  for (int o = 0; o < plan.n_out; ++o) {
    s << "      case " << o << ": c.type = " << plan.out[o].type << ";";
    if (!plan.out[o].valid) s << " c.valid = nullptr;";
    s << " break;\n";
  }
  s << "      default: break;\n    }\n    return c;\n  }\n";
  s << "  __device__ static __forceinline__ DevRange range(const DevPlan& p, int r) {\n"
       "    DevRange g = p.ranges[r];\n    switch (r) {\n";
  for (int r = 0; r < plan.n_ranges; ++r)
    s << "      case " << r << ": g.col = " << plan.ranges[r].col << "; g.kind = " << plan.ranges[r].kind
      << "; g.mul = " << (long long)plan.ranges[r].mul << "LL; break;\n";
  s << "      default: break;\n    }\n    return g;\n  }\n";
  s << "  __device__ static __forceinline__ int n_ranges(const DevPlan&) { return " << plan.n_ranges << "; }\n";
  s << "  __device__ static __forceinline__ int split(const DevPlan&) { return " << plan.split << "; }\n";
  s << "  __device__ static __forceinline__ int n_ops(const DevPlan&) { return " << plan.n_ops << "; }\n";
  s << "  __device__ static __forceinline__ bool nt(int j, bool dflt) {\n    switch (j) {\n";
  for (int j = 0; j < plan.n_in; ++j) s << "      case " << j << ": return " << (nt[(size_t)j] ? "true" : "false") << ";\n";
  s << "      default: return dflt;\n    }\n  }\n";
  s << "  template <class F>\n"
       "  __device__ static __forceinline__ void ops(const DevPlan& plan, int q0, int q1, F&& f) {\n";
  for (int q = 0; q < plan.n_ops; ++q) {
    const DevOp& o = ops[(size_t)q];
    s << "    if (q0 <= " << q << " && " << q << " < q1) {\n      DevOp op;\n"
      << "      op.code = " << (int)o.code << "; op.dst = " << (int)o.dst << "; op.a = " << (int)o.a
      << "; op.b = " << (int)o.b << "; op.flags = " << (int)o.flags << "; op.aux = " << (int)o.aux
      << "; op.aux2 = " << (int)o.aux2 << "; op.pad = 0;\n";
    // structural immediates (column lists, spill slots) are compiled in;
    // literals and device pointers are read from the plan at launch
    if (is_loadn_like(o.code)) s << "      op.imm = " << (long long)o.imm << "LL;\n";
    else s << "      op.imm = load_op(plan, " << q << ").imm;\n";
    s << "      if (f(op)) return;\n    }\n";
  }
  s << "  }\n};\n}  // namespace bsql_gen\n";
  return s.str();
}

std::string full_source(const std::string& body) {
  return std::string(kPrelude) + "#include \"vm_device.h\"\n" + body;
}

}  // namespace

bool generated_filter_project_eligible(const DevPlan& plan, const std::vector<DevOp>& ops) {
  if (plan.n_ranges <= 0 || plan.split >= plan.n_ops || plan.n_spill != 0) return false;
  if (plan.n_out > BSQL_DEV_MAX_OUT) return false;
  for (int q = plan.split; q < plan.n_ops; ++q)
    if (ops[(size_t)q].code == D_RAND) return false;  // RAND numbers rows by output position
  return true;
}

namespace {

// How the one-pass kernel runs one program: the source of its constants
// (struct `name`) and the dynamic LDS it needs.
struct OnePassShape {
  std::string source;
  int64_t lds = 0;
};

// shared[j]: input column j is read by other programs of the same kernel
// too (its loads stay in the caches for them).
// batch: for the batched kernel, whose selection is generated separately
// (no range columns kept, no eager range loads).
OnePassShape one_pass_shape(const DevPlan& plan, const std::vector<DevOp>& ops, const char* name,
                            const std::vector<bool>* shared = nullptr, bool batch = false,
                            TileShape ts = kOnePass) {
  // the output ops' columns that a range test reads: the widest one stays
  // in registers, the others are read again through the cache
  std::vector<bool> out_cols = loaded_columns(plan, ops, plan.split, plan.n_ops);
  std::vector<bool> nt((size_t)plan.n_in, true);
  if (shared)
    for (int j = 0; j < plan.n_in; ++j) nt[(size_t)j] = !(*shared)[(size_t)j];
  int keep = -1;
  for (int r = 0; r < plan.n_ranges; ++r) {
    const int c = plan.ranges[r].col;
    if (c < 0 || c >= plan.n_in || !out_cols[(size_t)c] || plan.in[c].type == BSQL_STRING) continue;
    nt[(size_t)c] = false;
    if (keep < 0 || type_size(plan.in[c].type) > type_size(plan.in[keep].type)) keep = c;
  }
  if (keep >= 0) nt[(size_t)keep] = !(shared && (*shared)[(size_t)keep]);
  if (batch) keep = -1;
  const int64_t slice = ts.slice();
  // every range column in flight at once when their elements fit the
  // registers (at most 16 bytes per row over the ranges)
  int range_bytes = 0;
  for (int r = 0; r < plan.n_ranges; ++r) range_bytes += type_size(plan.in[plan.ranges[r].col].type);
  const bool eager = !batch && plan.n_ranges <= 8 && range_bytes <= 16;
  // rows staged per tile: whole waves, within the LDS budget
  const int vw = (int)(slice / 32 + 2);
  int cap = 0;
  if (plan.n_out > 0) {
    cap = (int)std::min<int64_t>(ts.tile(), kStageBytes / (8 * plan.n_out) / 64 * 64);
    while (cap > 0 && fp_lds_bytes(plan.n_out, cap, ts.waves, vw) > kMaxLds) cap -= 64;
  }
  OnePassShape sh;
  sh.source = source_struct(plan, ops, nt, keep, eager, cap, name);
  sh.lds = fp_lds_bytes(plan.n_out, cap, ts.waves, vw);
  return sh;
}

int64_t one_pass_tiles(int64_t num_rows, TileShape ts = kOnePass) {
  const int64_t tiles = (num_rows + ts.tile() - 1) / ts.tile();
  if (tiles > (int64_t)INT32_MAX) fail("a generated filter handles at most 2^31-1 tiles per call");
  return tiles;
}

}  // namespace

hipError_t launch_generated_filter_project(const DevPlan& plan, const std::vector<DevOp>& ops,
                                           const DevPlan* d_plan, uint64_t* status, int64_t* d_count,
                                           int64_t capacity, hipStream_t stream) {
  const int64_t tiles = one_pass_tiles(plan.num_rows);
  if (tiles == 0) return fill_words((uint64_t*)d_count, 1, 0, stream);
  std::ostringstream name;
  name << "&bsql_dev::filter_project_rows<bsql_gen::Src, 16, " << kRows << ", " << kChunks << ", " << kTileWaves
       << ">";
  const OnePassShape sh = one_pass_shape(plan, ops, "Src");
  const hipFunction_t fn = kernel_for(name.str(), full_source(sh.source));
  if (!fn) return hipSuccess;  // check mode
  hipError_t e = fill_words(status, tiles, 0, stream);
  if (e != hipSuccess) return e;
  FusedArgs fa{status, nullptr, d_count, capacity, tiles};
  void* args[] = {(void*)&d_plan, (void*)&fa};
  return hipModuleLaunchKernel(fn, (unsigned)tiles, 1, 1, 64 * kTileWaves, 1, 1, (unsigned)sh.lds, stream, args,
                               nullptr);
}

namespace {

// The selection of a batch of queries (struct Batch): for each table column
// some range reads, in `order`, one load of its chunks for the lanes where a
// query testing it still has a row (cached when `cached[c]`: the queries'
// outputs read it again), then every range on it applied to its query's
// rows. A column that no lane needs is not read.
std::string batch_select_source(const std::vector<DevPlan>& plans, const std::vector<std::vector<int>>& in_maps,
                                const std::vector<int>& order, const std::map<int, bool>& cached) {
  const int nq = (int)plans.size();
  std::ostringstream s;
  s << "namespace bsql_gen {\nstruct Batch {\n  template <int R, int S, bool FULL>\n"
    << "  __device__ static __forceinline__ void select(const MultiArgs& ma, int64_t row0, int64_t num_rows,\n"
    << "                                                uint32_t (&alive)[" << nq << "]) {\n"
    << "    constexpr uint32_t RMASK = (1u << R) - 1u;\n";
  for (int c : order) {
    struct Use {
      int q, r, j;
    };
    std::vector<Use> uses;
    for (int q = 0; q < nq; ++q)
      for (int r = 0; r < plans[(size_t)q].n_ranges; ++r) {
        const int j = plans[(size_t)q].ranges[r].col;
        if (in_maps[(size_t)q][(size_t)j] == c) uses.push_back({q, r, j});
      }
    if (uses.empty()) continue;
    const DevPlan& p0 = plans[(size_t)uses[0].q];
    std::ostringstream need;
    std::vector<bool> seen((size_t)nq, false);
    for (const Use& u : uses) {
      if (seen[(size_t)u.q]) continue;
      seen[(size_t)u.q] = true;
      need << (need.tellp() > 0 ? " | " : "") << "alive[" << u.q << "]";
    }
    auto ci = cached.find(c);
    const bool cache = ci != cached.end() && ci->second;
    s << "    {  // table column " << c << "\n"
      << "      const uint32_t need = " << need.str() << ";\n"
      << "      if (__ballot(need != 0) != 0) {\n"
      << "        typedef " << range_kind(p0.in[uses[0].j].type) << " K;\n"
      << "        typedef K::T CT;\n"
      << "        const DevInCol c = Src" << uses[0].q << "::in(*ma.plan[" << uses[0].q << "], " << uses[0].j << ");\n"
      << "        CT v[S][R];\n        uint32_t vm = 0;\n"
      << "#pragma unroll\n        for (int k = 0; k < S; ++k) {\n"
      << "          const int64_t r0 = row0 + (int64_t)k * 64 * R;\n"
      << "          int n = R;\n"
      << "          if (!FULL) {\n"
      << "            const int64_t left = num_rows - r0;\n"
      << "            n = left <= 0 ? 0 : (left >= R ? R : (int)left);\n"
      << "          }\n"
      << "          if (((need >> (k * R)) & RMASK) == 0) n = 0;\n"
      << "          load_raw<CT, R, " << (cache ? "false" : "true") << ">(c, r0, n, v[k]);\n"
      << "          vm |= load_valid<R>(c, r0, n) << (k * R);\n"
      << "        }\n";
    for (const Use& u : uses) {
      s << "        {\n          const DevRange rg = Src" << u.q << "::range(*ma.plan[" << u.q << "], " << u.r << ");\n"
        << "          uint32_t pass = 0;\n"
        << "          if (!(rg.kind & RK_EMPTY)) {\n"
        << "#pragma unroll\n          for (int k = 0; k < S; ++k) pass |= range_bits<K, R>(rg, v[k]) << (k * R);\n"
        << "          }\n"
        << "          alive[" << u.q << "] &= vm & pass;\n        }\n";
    }
    s << "      }\n    }\n";
  }
  s << "  }\n};\n}  // namespace bsql_gen\n";
  return s.str();
}

}  // namespace

// in_maps[q][j]: the table column that input j of plans[q] is.
hipError_t launch_generated_filter_project_batch(const std::vector<DevPlan>& plans,
                                                 const std::vector<std::vector<DevOp>>& ops,
                                                 const std::vector<std::vector<int>>& in_maps,
                                                 const std::vector<const DevPlan*>& d_plans, uint64_t* status,
                                                 int64_t* d_counts, const std::vector<int64_t>& capacity,
                                                 hipStream_t stream) {
  const int nq = (int)plans.size();
  if (nq < 1 || nq > BSQL_MAX_BATCH_QUERIES) fail("a batch holds 1 to 16 queries");
  const TileShape ts = kBatch;
  const int64_t tiles = one_pass_tiles(plans[0].num_rows, ts);
  if (tiles == 0) return fill_words((uint64_t*)d_counts, nq, 0, stream);
  // who reads each table column: the selection (once, for all queries) and
  // each query's outputs; a column read again after the selection is
  // loaded through the caches
  std::map<int, int> readers;
  std::map<int, int> first;  // table column -> earliest position of a range on it in a query's order
  std::map<int, int> uses;
  for (int q = 0; q < nq; ++q) {
    const DevPlan& pl = plans[(size_t)q];
    const std::vector<bool> out_cols = loaded_columns(pl, ops[(size_t)q], pl.split, pl.n_ops);
    for (int j = 0; j < pl.n_in; ++j)
      if (out_cols[(size_t)j]) ++readers[in_maps[(size_t)q][(size_t)j]];
    for (int r = 0; r < pl.n_ranges; ++r) {
      const int c = in_maps[(size_t)q][(size_t)pl.ranges[r].col];
      auto f = first.find(c);
      if (f == first.end() || r < f->second) first[c] = r;
      ++uses[c];
    }
  }
  std::map<int, bool> cached;
  std::vector<int> order;
  for (auto& f : first) {
    order.push_back(f.first);
    cached[f.first] = readers[f.first] > 0;
    ++readers[f.first];
  }
  // the columns in the order the predicates' (learned) conjunct orders
  // put them: earliest first, then the ones more queries test
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return first[a] != first[b] ? first[a] < first[b] : uses[a] > uses[b];
  });
  std::ostringstream name, body;
  name << "&bsql_dev::filter_project_batch<16, " << ts.rows << ", " << ts.chunks << ", " << ts.waves
       << ", bsql_gen::Batch";
  int64_t lds = 0;
  for (int q = 0; q < nq; ++q) {
    const std::string sname = "Src" + std::to_string(q);
    const DevPlan& pl = plans[(size_t)q];
    std::vector<bool> shared((size_t)pl.n_in, false);
    for (int j = 0; j < pl.n_in; ++j) shared[(size_t)j] = readers[in_maps[(size_t)q][(size_t)j]] > 1;
    const OnePassShape sh = one_pass_shape(pl, ops[(size_t)q], sname.c_str(), &shared, true, ts);
    body << sh.source;
    lds = std::max(lds, sh.lds);
    name << ", bsql_gen::" << sname;
  }
  name << ">";
  body << batch_select_source(plans, in_maps, order, cached);
  const hipFunction_t fn = kernel_for(name.str(), full_source(body.str()));
  if (!fn) return hipSuccess;  // check mode
  // each query's status words: tiles of them from status + q * tiles
  hipError_t e = fill_words(status, tiles * nq, 0, stream);
  if (e != hipSuccess) return e;
  MultiArgs ma;
  std::memset(&ma, 0, sizeof(ma));
  for (int q = 0; q < nq; ++q) {
    ma.plan[q] = d_plans[(size_t)q];
    ma.fa[q] = FusedArgs{status + (size_t)q * (size_t)tiles, nullptr, d_counts + q, capacity[(size_t)q], tiles};
  }
  void* args[] = {(void*)&ma};
  return hipModuleLaunchKernel(fn, (unsigned)tiles, 1, 1, 64 * ts.waves, 1, 1, (unsigned)lds, stream, args,
                               nullptr);
}

int64_t generated_filter_project_scratch_words(int64_t num_rows) {
  const int64_t tile = std::min(kOnePass.tile(), kBatch.tile());
  return (num_rows + tile - 1) / tile + 1;
}

bool generated_execute_eligible(const DevPlan& plan, const std::vector<DevOp>& ops) {
  (void)ops;
  return plan.split == 0 && plan.n_ops > 0;
}

hipError_t launch_generated_execute(const DevPlan& plan, const std::vector<DevOp>& ops, const DevPlan* d_plan,
                                    int nreg, int rpt, int block, hipStream_t stream) {
  if (rpt != 1 && rpt != 2 && rpt != 4) fail("generated kernel: rows per lane must be 1, 2 or 4");
  const int64_t grid = execute_grid(plan.num_rows, rpt, block);
  if (grid == 0) return hipSuccess;
  std::vector<bool> nt((size_t)plan.n_in, true);
  std::ostringstream name;
  name << "&bsql_dev::vm_rows<" << (nreg <= 8 ? 8 : 16) << ", " << rpt << ", false, " << block << ", bsql_gen::Src>";
  const hipFunction_t fn = kernel_for(name.str(), full_source(source_struct(plan, ops, nt)));
  if (!fn) return hipSuccess;  // check mode
  const uint32_t* idx = nullptr;
  const int64_t* cnt = nullptr;
  int64_t cap = plan.num_rows;
  void* args[] = {(void*)&d_plan, (void*)&idx, (void*)&cnt, (void*)&cap};
  const unsigned lds = (unsigned)spill_lds_bytes(plan.n_spill, block, rpt);
  return hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, (unsigned)block, 1, 1, lds, stream, args, nullptr);
}

void set_generated_check(const char* arch) {
  t_check_arch = arch;
  t_check_bytes = 0;
}
bool generated_check_mode() { return t_check_arch != nullptr; }
int64_t generated_check_bytes() { return t_check_bytes; }

void generated_stats(int64_t* kernels, int64_t* launches, double* compile_seconds) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (kernels) *kernels = g_compiled;
  if (launches) *launches = g_launches;
  if (compile_seconds) *compile_seconds = g_compile_seconds;
}

}  // namespace bsql
