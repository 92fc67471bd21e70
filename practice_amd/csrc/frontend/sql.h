// SQL subset front end: parser, translator and predicate-pushdown optimizer.
//
// The reference parses with hyrise/sql-parser (not available here) and
// translates with QueryTranslator (src/core/translator.cpp:958-1100); this is
// a small recursive-descent parser of the SSB dialect that yields the same
// operator-tree shapes:
//   FROM t1, t2, ..., tn  ->  right-deep HashJoin(build = tk, probe = rest)
//                             with t1 innermost (translator.cpp:533-550)
//   WHERE                 ->  FilterNode above the join tree
//   aggregates / GROUP BY ->  AggregateNode (group columns first, then the
//                             aggregates in SELECT order)
//   ORDER BY / LIMIT      ->  SortLimitNode with output-slot keys
// BETWEEN is rejected as in translator.cpp:459-463.
#pragma once

#include <memory>
#include <string>

#include "../plan/operators.h"

namespace db {

// Parse + translate one SELECT statement into the naive operator tree.
std::unique_ptr<OperatorNode> translateSql(const std::string& sql);

// PredicatePushdownRule (src/core/optimizer_rules.cpp:185-307): split the
// WHERE conjunction, push single-side conjuncts below each join, and attach
// cross-table equi-key conjuncts to HashJoinNode::join_condition.
class Optimizer {
public:
    explicit Optimizer(const Catalog* catalog = nullptr) : catalog_(catalog) {}
    std::unique_ptr<OperatorNode> optimize(std::unique_ptr<OperatorNode> plan) const;

private:
    const Catalog* catalog_;
};

// translateSql + Optimizer::optimize.
std::unique_ptr<OperatorNode> buildOptimizedTree(const std::string& sql, const Catalog* catalog);

}  // namespace db
