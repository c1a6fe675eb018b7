// noalloc.hip -- the fixed device scratch of the allocation-free filter +
// project entry point (bsql_filter_project_noalloc), and the kernel that
// writes a launch plan into it.
//
// The scratch is one file-scope device array per GPU, held until the process
// ends. Host code never uses the array's name as an address: each device's
// copy is looked up once with hipGetSymbolAddress (with that device current)
// and that pointer is what kernels and copies receive. Every call on a GPU
// shares the array, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>

namespace bsql {

// 128 MiB: the selection ballots, the selected row ids (4 bytes a row) and
// the launch plan of a filter over up to about 32M rows.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;

}  // namespace bsql

static __device__ __align__(256) uint64_t scratch[bsql::kNoAllocScratchBytes / 8];

namespace bsql {
namespace {

std::mutex g_mu;
std::map<int, uint64_t*> g_ptr;  // device -> address of its `scratch`

// Plan words carried as a kernel argument (under the 4 KB argument limit).
constexpr int kChunkWords = 480;
struct WordChunk {
  uint64_t w[kChunkWords];
};

__global__ __launch_bounds__(256) void write_words(uint64_t* __restrict__ dst, WordChunk c, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = c.w[i];
}

}  // namespace

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

// The calling thread's current device's copy of `scratch`.
hipError_t noalloc_scratch(uint64_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ptr.find(dev);
  if (it == g_ptr.end()) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    it = g_ptr.emplace(dev, (uint64_t*)p).first;
  }
  *out = it->second;
  return hipSuccess;
}

// Copies n host words to device memory at dst with kernel launches whose
// arguments carry the words (stream-ordered; no host buffer is read later,
// so the copy can be captured in a graph and replayed).
hipError_t write_words_async(uint64_t* dst, const uint64_t* src, size_t n, hipStream_t stream) {
  for (size_t at = 0; at < n; at += kChunkWords) {
    WordChunk c;
    const int m = (int)(n - at < (size_t)kChunkWords ? n - at : (size_t)kChunkWords);
    for (int i = 0; i < m; ++i) c.w[i] = src[at + (size_t)i];
    hipLaunchKernelGGL(write_words, dim3(1), dim3(256), 0, stream, dst + at, c, m);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace bsql
