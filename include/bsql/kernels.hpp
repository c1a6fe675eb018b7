// bsql/kernels.hpp -- the operator surface of the filter/project path in C++,
// header-only over the C ABI of bsql/table.h.
//
// Mirrors the reference's execution kernels for this path:
//   task_status / task_result      engine/src/execution_kernels/kernel.h:10-20
//   kernel::do_process             engine/src/execution_kernels/kernel.h:199-203
//   Filter::do_process             engine/src/execution_kernels/BatchProcessing.cpp:454-473
//   Projection::do_process + run   engine/src/execution_kernels/BatchProcessing.cpp:365-448
//   BindableTableScan::do_process  engine/src/execution_kernels/BatchProcessing.cpp:261-284
//   ConcatenatingCacheMachine      engine/src/cache_machine/CacheMachine.h:149-175, .cpp:866-900
//     (QueueSink::pull_batches + BatchedProjection: the waiting batches in one launch)
// A task hands each kernel one batch (a table) with the stream it runs on; the
// kernel calls the processor function and maps its outcome: BSQL_OK ->
// SUCCESS (the result goes to the output sink), BSQL_ERR_OOM -> RETRY with
// the inputs handed back (the processor functions leave their inputs
// untouched when they fail), anything else -> FAIL with the message.
// Unlike the reference, which ignores the stream it is given
// (BatchProcessing.cpp:373, 462), every launch is ordered on that stream.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "table.h"

namespace bsql {

// Owning handle of a bsql_table (ral::frame::BlazingTable).
struct TableDeleter {
  void operator()(bsql_table* t) const { bsql_table_free(t); }
};
using Table = std::unique_ptr<bsql_table, TableDeleter>;

namespace execution {

enum class task_status { SUCCESS, RETRY, FAIL };

struct task_result {
  task_status status;
  std::string what;
  std::vector<Table> inputs;  // handed back on RETRY
};

// Where a kernel puts its output batches (the role of the CacheMachine the
// reference's kernels write to).
class CacheSink {
 public:
  virtual ~CacheSink() = default;
  virtual void add_to_cache(Table table) = 0;
};

// A FIFO of tables; thread-safe.
class QueueSink : public CacheSink {
 public:
  void add_to_cache(Table table) override {
    std::lock_guard<std::mutex> lk(mu_);
    q_.push_back(std::move(table));
  }
  Table pull() {
    std::lock_guard<std::mutex> lk(mu_);
    if (q_.empty()) return Table();
    Table t = std::move(q_.front());
    q_.pop_front();
    return t;
  }
  size_t size() const {
    std::lock_guard<std::mutex> lk(mu_);
    return q_.size();
  }
  // The batches a consumer takes at once, as ConcatenatingCacheMachine::
  // pullCacheData collects them (CacheMachine.cpp:866-900): popped while the
  // total with the next batch stays within max_bytes (at least one), in
  // arrival order, and not concatenated: the batched projection runs them in
  // one launch as they are.
  std::vector<Table> pull_batches(size_t max_bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    std::vector<Table> out;
    size_t total = 0;
    while (!q_.empty()) {
      const size_t next = (size_t)std::max<int64_t>(bsql_table_size_in_bytes(q_.front().get()), 0);
      if (!out.empty() && total + next > max_bytes) break;
      total += next;
      out.push_back(std::move(q_.front()));
      q_.pop_front();
    }
    return out;
  }

 private:
  mutable std::mutex mu_;
  std::deque<Table> q_;
};

// A computation unit of the execution graph with its logical-plan text.
class Kernel {
 public:
  explicit Kernel(std::string expression) : expression_(std::move(expression)) {}
  virtual ~Kernel() = default;
  virtual task_result do_process(std::vector<Table> inputs, std::shared_ptr<CacheSink> output,
                                 hipStream_t stream, const std::map<std::string, std::string>& args) = 0;
  const std::string& expression() const { return expression_; }

 protected:
  // SUCCESS / RETRY / FAIL from a processor function's return code.
  static task_result outcome(int rc, std::vector<Table>& inputs) {
    if (rc == BSQL_OK) return {task_status::SUCCESS, std::string(), {}};
    if (rc == BSQL_ERR_OOM) return {task_status::RETRY, bsql_last_error(), std::move(inputs)};
    return {task_status::FAIL, bsql_last_error(), {}};
  }
  std::string expression_;
};

// LogicalFilter(condition=[...]): one output batch per input batch.
class Filter : public Kernel {
 public:
  using Kernel::Kernel;
  task_result do_process(std::vector<Table> inputs, std::shared_ptr<CacheSink> output, hipStream_t stream,
                         const std::map<std::string, std::string>&) override {
    if (inputs.empty() || !inputs[0]) return {task_status::FAIL, "Filter: no input batch", {}};
    bsql_table* out = nullptr;
    const int rc = bsql_process_filter(inputs[0].get(), expression_.c_str(), &out, stream);
    if (rc == BSQL_OK) output->add_to_cache(Table(out));
    return outcome(rc, inputs);
  }
};

// LogicalProject(name=[expr], ...). An identity projection ($0..$n-1) moves
// the batch through, renamed when the names differ, without a task (the
// bypass of Projection::run).
class Projection : public Kernel {
 public:
  using Kernel::Kernel;
  // The bypass decision for a batch with these names: 0 = evaluate,
  // 1 = pass through, 2 = pass through renamed.
  int bypass(const bsql_table* batch) const {
    std::vector<const char*> names;
    for (int k = 0; k < bsql_table_num_columns(batch); ++k) names.push_back(bsql_table_name(batch, k));
    int b = 0, w = 0;
    if (bsql_project_bypass(expression_.c_str(), names.data(), (int)names.size(), &b, &w) != BSQL_OK) return 0;
    return b ? (w ? 2 : 1) : 0;
  }
  task_result do_process(std::vector<Table> inputs, std::shared_ptr<CacheSink> output, hipStream_t stream,
                         const std::map<std::string, std::string>&) override {
    if (inputs.empty() || !inputs[0]) return {task_status::FAIL, "Projection: no input batch", {}};
    bsql_table* out = nullptr;
    const int rc = bsql_process_project(inputs[0].get(), expression_.c_str(), &out, stream);
    if (rc == BSQL_OK) {
      (void)inputs[0].release();  // consumed by the projection
      output->add_to_cache(Table(out));
    }
    return outcome(rc, inputs);
  }
};

// The Projection kernel over a set of waiting batches of one schema: one
// task for all of them, the program compiled once and every batch computed
// in one persistent launch (bsql_process_project_batches). SUCCESS puts each
// batch's result in the output cache in input order; RETRY hands every input
// back (none is consumed); FAIL reports the message. An identity projection
// moves the batches through as Projection does.
class BatchedProjection : public Projection {
 public:
  using Projection::Projection;
  task_result do_process(std::vector<Table> inputs, std::shared_ptr<CacheSink> output, hipStream_t stream,
                         const std::map<std::string, std::string>&) override {
    if (inputs.empty()) return {task_status::FAIL, "Projection: no input batch", {}};
    std::vector<bsql_table*> in;
    for (Table& t : inputs) {
      if (!t) return {task_status::FAIL, "Projection: a null input batch", {}};
      in.push_back(t.get());
    }
    std::vector<bsql_table*> out(in.size(), nullptr);
    const int rc = bsql_process_project_batches(in.data(), (int)in.size(), expression_.c_str(), out.data(), stream);
    if (rc == BSQL_OK) {
      for (Table& t : inputs) (void)t.release();  // consumed (or moved through) by the projection
      for (bsql_table* t : out) output->add_to_cache(Table(t));
    }
    return outcome(rc, inputs);
  }
};

// BindableTableScan(table=[[...]], filters=[[...]], projects=[[...]],
// aliases=[[...]]): batches arrive with the projected columns already
// loaded; a pushed-down filter runs on them ($k = k-th projected column)
// and the scan's aliases name the result.
class BindableTableScan : public Kernel {
 public:
  using Kernel::Kernel;
  bool filterable() const { return expression_.find("filters") != std::string::npos; }
  task_result do_process(std::vector<Table> inputs, std::shared_ptr<CacheSink> output, hipStream_t stream,
                         const std::map<std::string, std::string>&) override {
    if (inputs.empty() || !inputs[0]) return {task_status::FAIL, "BindableTableScan: no input batch", {}};
    if (!filterable()) {
      const int rc = bsql_table_fix_column_aliases(inputs[0].get(), expression_.c_str());
      if (rc != BSQL_OK) return outcome(rc, inputs);
      output->add_to_cache(std::move(inputs[0]));
      return {task_status::SUCCESS, std::string(), {}};
    }
    bsql_table* out = nullptr;
    int rc = bsql_process_filter(inputs[0].get(), expression_.c_str(), &out, stream);
    if (rc != BSQL_OK) return outcome(rc, inputs);
    Table result(out);
    rc = bsql_table_fix_column_aliases(result.get(), expression_.c_str());
    if (rc != BSQL_OK) return outcome(rc, inputs);
    output->add_to_cache(std::move(result));
    return {task_status::SUCCESS, std::string(), {}};
  }
};

}  // namespace execution
}  // namespace bsql
