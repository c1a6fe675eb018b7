// Operator-surface test: Filter -> Projection -> BindableTableScan kernels of
// bsql/kernels.hpp over device tables, as the reference's execution graph
// drives them (one batch per do_process, the result in the output cache),
// plus RETRY on an allocation failure (inputs handed back untouched) and
// FAIL on an invalid expression. Row values are small closed forms so that
// the expected results are evident. Exit status 0 = every check passed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bsql/kernels.hpp"

using bsql::Table;
using bsql::execution::BatchedProjection;
using bsql::execution::BindableTableScan;
using bsql::execution::Filter;
using bsql::execution::Projection;
using bsql::execution::QueueSink;
using bsql::execution::task_status;

static int g_failures = 0;
#define CHECK(cond, what)                                             \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__); \
      ++g_failures;                                                   \
    }                                                                 \
  } while (0)

static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::printf("HIP error in %s: %s\n", what, hipGetErrorString(e));
    std::exit(2);
  }
}

template <class T>
static std::vector<T> read_column(const bsql_table* t, int k, hipStream_t s) {
  bsql_column c{};
  bsql_table_column(t, k, &c);
  std::vector<T> v((size_t)c.size);
  if (c.size) hip_ok(hipMemcpyAsync(v.data(), (const T*)c.data + c.offset, v.size() * sizeof(T),
                                    hipMemcpyDeviceToHost, s), "copy out");
  hip_ok(hipStreamSynchronize(s), "sync");
  return v;
}

static std::vector<bool> read_valid(const bsql_table* t, int k, hipStream_t s) {
  bsql_column c{};
  bsql_table_column(t, k, &c);
  std::vector<bool> v((size_t)c.size, true);
  if (!c.validity || !c.size) return v;
  std::vector<uint32_t> w((size_t)((c.size + 31) / 32));
  hip_ok(hipMemcpyAsync(w.data(), c.validity, w.size() * 4, hipMemcpyDeviceToHost, s), "copy validity");
  hip_ok(hipStreamSynchronize(s), "sync");
  for (int64_t i = 0; i < c.size; ++i) v[(size_t)i] = (w[(size_t)(i / 32)] >> (i % 32)) & 1u;
  return v;
}

static std::vector<std::string> read_strings(const bsql_table* t, int k, hipStream_t s) {
  bsql_column c{};
  bsql_table_column(t, k, &c);
  std::vector<int32_t> off((size_t)c.size + 1);
  hip_ok(hipMemcpyAsync(off.data(), c.offsets + c.offset, off.size() * 4, hipMemcpyDeviceToHost, s), "offsets");
  hip_ok(hipStreamSynchronize(s), "sync");
  std::vector<char> chars((size_t)std::max(off.back(), 1));
  if (off.back() > 0)
    hip_ok(hipMemcpyAsync(chars.data(), c.data, (size_t)off.back(), hipMemcpyDeviceToHost, s), "chars");
  hip_ok(hipStreamSynchronize(s), "sync");
  std::vector<std::string> v;
  for (int64_t i = 0; i < c.size; ++i)
    v.emplace_back(chars.data() + off[(size_t)i], (size_t)(off[(size_t)i + 1] - off[(size_t)i]));
  return v;
}

// a: int32 i % 100; b: float64 i / 1000 (row i null when i % 7 == 3)
static Table make_batch(int64_t n, hipStream_t s) {
  const int32_t types[2] = {BSQL_INT32, BSQL_FLOAT64};
  const char* names[2] = {"a", "b"};
  const uint8_t nullable[2] = {0, 1};
  bsql_table* t = nullptr;
  if (bsql_table_create(types, names, nullable, nullptr, 2, n, s, &t) != BSQL_OK) {
    std::printf("table_create: %s\n", bsql_last_error());
    std::exit(2);
  }
  std::vector<int32_t> a((size_t)n);
  std::vector<double> b((size_t)n);
  std::vector<uint32_t> w((size_t)((n + 31) / 32), 0);
  for (int64_t i = 0; i < n; ++i) {
    a[(size_t)i] = (int32_t)(i % 100);
    b[(size_t)i] = (double)i / 1000.0;
    if (i % 7 != 3) w[(size_t)(i / 32)] |= 1u << (i % 32);
  }
  if (n == 0) return Table(t);
  bsql_column ca{}, cb{};
  bsql_table_column(t, 0, &ca);
  bsql_table_column(t, 1, &cb);
  hip_ok(hipMemcpyAsync(ca.data, a.data(), a.size() * 4, hipMemcpyHostToDevice, s), "fill a");
  hip_ok(hipMemcpyAsync(cb.data, b.data(), b.size() * 8, hipMemcpyHostToDevice, s), "fill b");
  hip_ok(hipMemcpyAsync(cb.validity, w.data(), w.size() * 4, hipMemcpyHostToDevice, s), "fill validity");
  hip_ok(hipStreamSynchronize(s), "sync");
  return Table(t);
}

static int g_alloc_calls = 0;
static int failing_alloc(void*, void**, size_t, void*) {
  ++g_alloc_calls;
  return 1;  // out of memory
}
static void failing_free(void*, void*, size_t, void*) {}

int main() {
  hipStream_t s;
  hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
  const int64_t n = 100000;
  auto sink = std::make_shared<QueueSink>();

  // Filter: WHERE a > 90 AND b < 50.0 (a null b drops the row)
  Filter filter("LogicalFilter(condition=[AND(>($0, 90), <($1, 50.0:DECIMAL(3, 1)))])");
  std::vector<Table> in;
  in.push_back(make_batch(n, s));
  auto r = filter.do_process(std::move(in), sink, s, {});
  CHECK(r.status == task_status::SUCCESS, "filter succeeds");
  Table filtered = sink->pull();
  CHECK(filtered != nullptr, "filter output in the cache");
  std::vector<int64_t> expect_rows;
  for (int64_t i = 0; i < n; ++i)
    if (i % 100 > 90 && i % 7 != 3 && (double)i / 1000.0 < 50.0) expect_rows.push_back(i);
  CHECK(bsql_table_num_rows(filtered.get()) == (int64_t)expect_rows.size(), "filter row count");
  {
    auto a = read_column<int32_t>(filtered.get(), 0, s);
    auto b = read_column<double>(filtered.get(), 1, s);
    bool same = a.size() == expect_rows.size();
    for (size_t j = 0; same && j < a.size(); ++j)
      same = a[j] == (int32_t)(expect_rows[j] % 100) && b[j] == (double)expect_rows[j] / 1000.0;
    CHECK(same, "filtered values in row order");
    CHECK(std::string(bsql_table_name(filtered.get(), 1)) == "b", "filter keeps names");
    bsql_column ca{};
    bsql_table_column(filtered.get(), 0, &ca);
    CHECK(ca.validity == nullptr, "non-nullable input stays non-nullable");
  }

  // Projection: x = a * 2 + 1, y = b (the filtered batch is consumed)
  Projection proj("LogicalProject(x=[+(*($0, 2), 1)], y=[$1])");
  CHECK(proj.bypass(filtered.get()) == 0, "computing projection is not a bypass");
  std::vector<Table> in2;
  in2.push_back(std::move(filtered));
  r = proj.do_process(std::move(in2), sink, s, {});
  CHECK(r.status == task_status::SUCCESS, "projection succeeds");
  Table projected = sink->pull();
  CHECK(projected != nullptr, "projection output in the cache");
  {
    auto x = read_column<int32_t>(projected.get(), 0, s);
    auto vx = read_valid(projected.get(), 0, s);
    bool same = x.size() == expect_rows.size();
    for (size_t j = 0; same && j < x.size(); ++j) same = x[j] == (int32_t)(expect_rows[j] % 100) * 2 + 1 && vx[j];
    CHECK(same, "projected values");
    CHECK(std::string(bsql_table_name(projected.get(), 0)) == "x", "projection names");
  }

  // Identity projection: moved through, renamed
  Projection rename("LogicalProject(p=[$0], q=[$1])");
  CHECK(rename.bypass(projected.get()) == 2, "identity projection with new names is a renaming bypass");
  bsql_table* before = projected.get();
  std::vector<Table> in3;
  in3.push_back(std::move(projected));
  r = rename.do_process(std::move(in3), sink, s, {});
  Table renamed = sink->pull();
  CHECK(r.status == task_status::SUCCESS && renamed.get() == before, "bypass moves the same table");
  CHECK(std::string(bsql_table_name(renamed.get(), 1)) == "q", "bypass renames");

  // BindableTableScan with a pushed-down filter over the projected columns
  BindableTableScan scan(
      "BindableTableScan(table=[[main, t]], filters=[[<($0, 5)]], projects=[[0, 1]], aliases=[[u, v]])");
  std::vector<Table> in4;
  in4.push_back(make_batch(1000, s));
  r = scan.do_process(std::move(in4), sink, s, {});
  Table scanned = sink->pull();
  CHECK(r.status == task_status::SUCCESS && scanned, "scan succeeds");
  CHECK(bsql_table_num_rows(scanned.get()) == 50, "scan filter row count");
  CHECK(std::string(bsql_table_name(scanned.get(), 0)) == "u", "scan aliases");

  // Projection building strings (columns made before the interpreter runs):
  // label = 'a=' || CAST(a AS VARCHAR), up = UPPER(label), and a filter on it
  Projection strs("LogicalProject(label=[||('a=', CAST($0):VARCHAR)], up=[UPPER(||('x', CAST($0):VARCHAR))])");
  std::vector<Table> in8;
  in8.push_back(make_batch(1000, s));
  r = strs.do_process(std::move(in8), sink, s, {});
  Table labelled = sink->pull();
  CHECK(r.status == task_status::SUCCESS && labelled, "string projection succeeds");
  {
    auto label = read_strings(labelled.get(), 0, s);
    auto up = read_strings(labelled.get(), 1, s);
    bool same = label.size() == 1000 && up.size() == 1000;
    for (size_t i = 0; same && i < label.size(); ++i)
      same = label[i] == "a=" + std::to_string(i % 100) && up[i] == "X" + std::to_string(i % 100);
    CHECK(same, "built strings");
  }
  Filter on_label("LogicalFilter(condition=[=(SUBSTRING($0, 3), '42')])");
  std::vector<Table> in9;
  in9.push_back(std::move(labelled));
  r = on_label.do_process(std::move(in9), sink, s, {});
  Table picked = sink->pull();
  CHECK(r.status == task_status::SUCCESS && picked && bsql_table_num_rows(picked.get()) == 10,
        "filter on a built string column");

  // RETRY: the allocator reports out of memory; the input comes back
  std::vector<Table> in5;
  in5.push_back(make_batch(n, s));
  bsql_table* batch5 = in5[0].get();
  bsql_set_allocator(failing_alloc, failing_free, nullptr);
  r = filter.do_process(std::move(in5), sink, s, {});
  bsql_set_allocator(nullptr, nullptr, nullptr);
  CHECK(r.status == task_status::RETRY, "allocation failure is RETRY");
  CHECK(g_alloc_calls > 0, "the installed allocator was used");
  CHECK(r.inputs.size() == 1 && r.inputs[0].get() == batch5, "RETRY hands the input back");
  CHECK(sink->size() == 0, "nothing cached on RETRY");
  // ... and the retried task succeeds once memory is there
  r = filter.do_process(std::move(r.inputs), sink, s, {});
  Table retried = sink->pull();
  CHECK(r.status == task_status::SUCCESS && retried &&
            bsql_table_num_rows(retried.get()) == (int64_t)expect_rows.size(),
        "retry succeeds");

  // Batched projection: 24 waiting batches of uneven sizes (one empty) taken
  // from the queue within a byte budget and projected in one launch; every
  // batch's result in input order, values from the closed forms.
  {
    auto queue = std::make_shared<QueueSink>();
    std::vector<int64_t> sizes;
    for (int i = 0; i < 24; ++i) {
      const int64_t rows = i == 5 ? 0 : (int64_t)(1000 + 37 * i * i);
      sizes.push_back(rows);
      queue->add_to_cache(make_batch(rows, s));
    }
    std::vector<Table> batches = queue->pull_batches((size_t)1 << 40);
    CHECK(batches.size() == 24 && queue->size() == 0, "pull_batches takes every batch within the budget");
    BatchedProjection bproj(
        "LogicalProject(x=[+(*($0, 2), 1)], y=[CASE(IS NOT NULL($1), $1, -1.0:DOUBLE)], z=[*($1, $0)])");
    auto bsink = std::make_shared<QueueSink>();
    r = bproj.do_process(std::move(batches), bsink, s, {});
    CHECK(r.status == task_status::SUCCESS && bsink->size() == 24, "batched projection succeeds");
    bool same = true;
    for (size_t i = 0; i < sizes.size(); ++i) {
      Table t = bsink->pull();
      if (!t || bsql_table_num_rows(t.get()) != sizes[i]) {
        same = false;
        continue;
      }
      auto x = read_column<int32_t>(t.get(), 0, s);
      auto y = read_column<double>(t.get(), 1, s);
      auto z = read_column<double>(t.get(), 2, s);
      auto vz = read_valid(t.get(), 2, s);
      for (int64_t j = 0; same && j < sizes[i]; ++j) {
        const bool null_b = j % 7 == 3;
        const double b = (double)j / 1000.0;
        same = x[(size_t)j] == (int32_t)(j % 100) * 2 + 1 && y[(size_t)j] == (null_b ? -1.0 : b) &&
               vz[(size_t)j] == !null_b && (null_b || z[(size_t)j] == b * (double)(j % 100));
      }
    }
    CHECK(same, "batched projection values per batch");
    // a budget takes a prefix (at least one batch)
    for (int i = 0; i < 3; ++i) queue->add_to_cache(make_batch(1000, s));
    const int64_t one = bsql_table_size_in_bytes(queue->pull_batches(0)[0].get());
    CHECK(one == 1000 * 4 + 1000 * 8 + 128, "table size in bytes: data plus the padded validity mask");
    CHECK(queue->pull_batches((size_t)(2 * one + 10)).size() == 2 && queue->size() == 0,
          "pull_batches stops at the budget");
    // RETRY: every input handed back, nothing cached; then the retry succeeds
    std::vector<Table> again;
    for (int i = 0; i < 3; ++i) again.push_back(make_batch(500 + i, s));
    std::vector<bsql_table*> ptrs;
    for (Table& t : again) ptrs.push_back(t.get());
    bsql_set_allocator(failing_alloc, failing_free, nullptr);
    r = bproj.do_process(std::move(again), bsink, s, {});
    bsql_set_allocator(nullptr, nullptr, nullptr);
    bool back = r.status == task_status::RETRY && r.inputs.size() == 3;
    for (size_t i = 0; back && i < 3; ++i) back = r.inputs[i].get() == ptrs[i];
    CHECK(back && bsink->size() == 0, "batched RETRY hands every input back");
    r = bproj.do_process(std::move(r.inputs), bsink, s, {});
    CHECK(r.status == task_status::SUCCESS && bsink->size() == 3, "batched retry succeeds");
    while (bsink->size()) bsink->pull();
  }

  // FAIL: a column the batch does not have
  Filter bad("LogicalFilter(condition=[>($5, 1)])");
  std::vector<Table> in6;
  in6.push_back(make_batch(10, s));
  r = bad.do_process(std::move(in6), sink, s, {});
  CHECK(r.status == task_status::FAIL && !r.what.empty() && r.inputs.empty(), "invalid expression is FAIL");

  // Empty batch: an empty table of the same schema, even for a bad predicate
  std::vector<Table> in7;
  in7.push_back(make_batch(0, s));
  r = bad.do_process(std::move(in7), sink, s, {});
  Table empty = sink->pull();
  CHECK(r.status == task_status::SUCCESS && empty && bsql_table_num_rows(empty.get()) == 0 &&
            bsql_table_num_columns(empty.get()) == 2,
        "empty input gives an empty table");

  hip_ok(hipStreamSynchronize(s), "final sync");
  retried.reset();
  picked.reset();
  scanned.reset();
  renamed.reset();
  empty.reset();
  hip_ok(hipStreamSynchronize(s), "frees");
  hip_ok(hipStreamDestroy(s), "stream destroy");
  if (g_failures) {
    std::printf("%d check(s) failed\n", g_failures);
    return 1;
  }
  std::printf("operator kernels: all checks passed\n");
  return 0;
}
