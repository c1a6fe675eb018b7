/*
 * bsql/table.h -- owning tables and the processor functions of the filter and
 * project operators, over the interpreter of bsql/interp.h.
 *
 * These replace the second level of the reference's operator stack: the
 * functions its Filter, Projection and BindableTableScan kernels call from
 * do_process (engine/src/execution_kernels/LogicalFilter.h:14-20,
 * LogicalProject.h:24-30), which take a table and a Calcite query part and
 * return a newly owned table. The first level (the kernels and their
 * SUCCESS / RETRY / FAIL task results) is the header-only C++ layer in
 * bsql/kernels.hpp.
 *
 * Conventions are those of bsql/interp.h: device buffers, explicit streams,
 * return codes (BSQL_ERR_OOM = retryable, the inputs are untouched).
 * Functions that must know a row count before allocating (process_filter)
 * synchronize their stream once; nothing synchronizes the device.
 */
#ifndef BSQL_TABLE_H
#define BSQL_TABLE_H

#include <stddef.h>
#include <stdint.h>

#include "interp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device memory ------------------------------------------------------
 * Every buffer a table owns comes from this allocator. The default is the
 * HIP stream-ordered allocator (hipMallocAsync / hipFreeAsync on the given
 * stream). A host with its own pool (the reference allocates through RMM)
 * installs it here. The allocation returns 0, or 1 when out of memory
 * (reported below as BSQL_ERR_OOM, retryable), or any other value for a
 * device error (BSQL_ERR_DEVICE). NULL functions restore the default. A
 * table keeps the allocator it was created with for its frees. */
typedef int (*bsql_alloc_fn)(void* ctx, void** ptr, size_t bytes, void* stream);
typedef void (*bsql_free_fn)(void* ctx, void* ptr, size_t bytes, void* stream);
int bsql_set_allocator(bsql_alloc_fn alloc, bsql_free_fn release, void* ctx);

/* ---- tables ----------------------------------------------------------------
 * bsql_table <=> ral::frame::BlazingTable (LogicPrimitives.h:23-89): columns
 * plus names. A table owns the buffers it allocated and borrows the ones it
 * was built from (bsql_table_from_views), like BlazingColumnOwner and
 * BlazingColumnView. Buffers are freed on the table's stream: the last
 * stream passed to a function that read or wrote it (or set with
 * bsql_table_record_stream), which must still exist when it is freed. */
typedef struct bsql_table bsql_table;

/* A table of `num_rows` rows with uninitialized data, allocated on `stream`:
 * fixed-width columns of types[k]; validity words iff nullable[k] (NULL:
 * none); STRING columns get num_rows+1 offsets (none for an empty column)
 * and chars_bytes[k] bytes of chars (chars_bytes may be NULL when there is no
 * STRING column). */
int bsql_table_create(const int32_t* types, const char* const* names, const uint8_t* nullable,
                      const int64_t* chars_bytes, int n_cols, int64_t num_rows, void* stream,
                      bsql_table** out);
/* A table that borrows caller-owned column views (nothing is copied or
 * freed). `names` may be NULL (empty names). */
int bsql_table_from_views(const bsql_column* cols, const char* const* names, int n_cols, int64_t num_rows,
                          bsql_table** out);
void bsql_table_free(bsql_table* table);
int64_t bsql_table_num_rows(const bsql_table* table);
int bsql_table_num_columns(const bsql_table* table);
/* View of column k (device pointers stay owned by the table). */
int bsql_table_column(const bsql_table* table, int k, bsql_column* out);
/* Name of column k (owned by the table), or NULL. */
const char* bsql_table_name(const bsql_table* table, int k);
int bsql_table_set_names(bsql_table* table, const char* const* names, int n);
int bsql_table_record_stream(bsql_table* table, void* stream);
/* Device bytes of the table's columns: data (chars and offsets for strings)
 * plus validity masks (BlazingTable::sizeInBytes,
 * engine/src/execution_kernels/LogicPrimitives.cpp:75-110); -1 for NULL. */
int64_t bsql_table_size_in_bytes(const bsql_table* table);

/* ---- processor functions ------------------------------------------------- */

/* ral::processor::process_filter (LogicalFilter.cpp:31-52): the predicate is
 * the query part's `condition=[...]`, else its `filters=[[...]]`
 * (get_named_expression, expression_utils.cpp:537-547), over `in`'s columns;
 * it must be BOOL8. Empty input gives an empty table of the same schema.
 * The output has the selected rows of every column in order, the input's
 * names, and validity iff the input column has validity. Synchronizes
 * `stream` once, to size the output. */
int bsql_process_filter(const bsql_table* in, const char* query_part, bsql_table** out, void* stream);

/* ral::processor::process_project (LogicalProject.cpp:943-969): the query
 * part's `name=[expr]` list (get_query_part + get_expressions_from_expression_list,
 * expression_utils.cpp:764-771, 1012-1066). On success `in` is consumed (its
 * buffers move into *out or are freed; do not use it again); on failure it is
 * untouched and still the caller's. A projection that lists every column in
 * order ($0, $1, ...) moves the input through, renamed
 * (bypassingProject, expression_utils.cpp:1115-1150, as Projection::run does,
 * BatchProcessing.cpp:394-406). */
int bsql_process_project(bsql_table* in, const char* query_part, bsql_table** out, void* stream);

/* Projection over several batches of one schema (a cache's waiting batches,
 * as ConcatenatingCacheMachine::pullCacheData collects them,
 * engine/src/cache_machine/CacheMachine.cpp:866-900, for the Projection kernel,
 * engine/src/execution_kernels/BatchProcessing.cpp:365-448): the program is
 * compiled once and every batch's outputs are computed in ONE persistent
 * launch (bsql_batch_*), without concatenating the batches. outs[i] is batch
 * i's result; on success every input is consumed (as bsql_process_project
 * consumes its input). On BSQL_ERR_OOM nothing is consumed and nothing is
 * returned (the caller retries the whole set). Programs the batched launch
 * does not run (string steps or outputs) and batches of different schemas
 * are evaluated batch by batch. */
int bsql_process_project_batches(bsql_table* const* ins, int n, const char* query_part, bsql_table** outs,
                                 void* stream);

/* ral::processor::evaluate_expressions (LogicalProject.cpp:796-941): one
 * output column per expression, each with validity (the reference's outputs
 * are always nullable); names are empty. */
int bsql_evaluate_expressions(const bsql_table* in, const char* const* exprs, int n_exprs, bsql_table** out,
                              void* stream);

/* ral::processor::applyBooleanFilter (LogicalFilter.cpp:22-29): stable
 * compaction of `in` by a BOOL8 mask of in's row count; a null or false mask
 * row drops the row. Synchronizes `stream` once. */
int bsql_apply_boolean_filter(const bsql_table* in, const bsql_column* mask, bsql_table** out, void* stream);

/* ---- the interpreter's other callers ------------------------------------- */

/* The input column of one aggregation, as the aggregation kernels compute it
 * (compute_aggregations_without_groupby / with_groupby, GroupBy.cpp:226-249
 * and :329-352). expr is the text between the aggregation's parentheses in
 * the LogicalAggregate plan. kind: BSQL_AGG_COUNT_VALID for COUNT(..),
 * BSQL_AGG_COUNT_ALL for COUNT(*), 0 for the others. COUNT over several
 * columns becomes the sum of their INTEGER casts
 * (modify_multi_column_count_expression), so that its nulls are the rows
 * where any of them is null; one column or a column number is that column,
 * borrowed; COUNT(*) the first column; anything else the expression
 * evaluated (evaluate_expressions). *out: a one-column table (its column
 * borrows the input's when it is one of them). */
enum { BSQL_AGG_OTHER = 0, BSQL_AGG_COUNT_VALID = 1, BSQL_AGG_COUNT_ALL = 2 };
int bsql_aggregation_input(const bsql_table* in, const char* expr, int kind, bsql_table** out, void* stream);

/* ral::skip_data::process_skipdata_for_table (SkipDataProcessor.cpp:242-346):
 * which row groups a table scan must read. metadata: the row-group
 * statistics table, column "min_<i>_<name>" / "max_<i>_<name>" for table
 * column i (names[i]) where it has statistics, the last two columns the file
 * and row-group ids. The scan's filter (condition=[..] or filters=[[..]]
 * of table_scan, over its projects=[[..]] columns) is rewritten over the
 * statistics (bsql_skip_data_filter) and evaluated by the interpreter; *out
 * is the table of the last two metadata columns filtered by it. When no row
 * group can be skipped (no filter, or nothing of it survives the rewrite),
 * *out is NULL and *read_all is 1. */
int bsql_process_skipdata_for_table(const bsql_table* metadata, const char* const* names, int n_names,
                                    const char* table_scan, bsql_table** out, int* read_all, void* stream);

/* The rewrite of a filter over the statistics table (projected column i:
 * minimum at column 2i, maximum at 2i + 1; has_stats[i] != 0 when column i
 * has statistics), the expression text process_skipdata_for_table
 * evaluates: drop_value and apply_skip_data_rules (SkipDataProcessor.cpp:
 * 232-246). An empty result: no row group can be skipped. Writes at most
 * cap-1 bytes and a NUL, and the full length to *len. */
int bsql_skip_data_filter(const char* filter, const uint8_t* has_stats, int n_cols, char* out, size_t cap,
                          size_t* len);

/* ---- query-part helpers of the operators ---------------------------------- */

/* bypassingProject (expression_utils.cpp:1115-1150) for a table with these
 * column names: *bypass = 1 if the projection lists $0..$n-1 in order,
 * *with_aliases = 1 if it also renames. */
int bsql_project_bypass(const char* query_part, const char* const* names, int n_names, int* bypass,
                        int* with_aliases);
/* The projection's aliases as the table's names (bypass with aliases). */
int bsql_table_set_project_aliases(bsql_table* table, const char* query_part);
/* fix_column_aliases (expression_utils.cpp:512-535): the scan's
 * `aliases=[[...]]` replace the first names. */
int bsql_table_fix_column_aliases(bsql_table* table, const char* query_part);
/* get_projections_wrapper (BatchProcessing.cpp:30-46) over get_projections
 * (expression_utils.cpp:549-580): the column indices a BindableTableScan of
 * a `num_columns`-column table loads (all of them when the query part lists
 * none). Returns the count; entries are written up to cap. A filter of that
 * scan indexes this projected list ($k = k-th loaded column). Minus an error
 * code on failure. */
int bsql_scan_projections(const char* query_part, int num_columns, int32_t* cols, int cap);

#ifdef __cplusplus
}
#endif

#endif /* BSQL_TABLE_H */
