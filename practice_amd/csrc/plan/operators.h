// Operator tree of the optimized plan.
//
// Mirrors the reference plan API (include/core/operators.h:81-384): node kinds,
// class names, public fields (children_[0] of a HashJoinNode is the build side,
// children_[1] the probe side), AggregateDef, the result-slot layout
// (visibleTupleSize / storageTupleSize) and calculateResultSize.  The HIP code
// generator consumes exactly these fields.
#pragma once

#include <algorithm>
#include <cctype>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "expressions.h"
#include "memory.h"

namespace db {

enum class OperatorType { TABLE_SCAN, FILTER, HASH_JOIN, AGGREGATE, PROJECTION, SORT_LIMIT };

class OperatorVisitor;

class OperatorNode {
protected:
    std::vector<std::unique_ptr<OperatorNode>> children_;

public:
    OperatorNode* parent_ = nullptr;  // non-owning back pointer, set by addChild
    virtual ~OperatorNode() = default;
    virtual OperatorType getType() const = 0;
    virtual void accept(OperatorVisitor& visitor) const = 0;
    void addChild(std::unique_ptr<OperatorNode> child) {
        child->parent_ = this;
        children_.push_back(std::move(child));
    }
    const std::vector<std::unique_ptr<OperatorNode>>& getChildren() const { return children_; }
    std::vector<std::unique_ptr<OperatorNode>>& getChildrenMutable() { return children_; }
};

class TableScanNode final : public OperatorNode {
public:
    std::string table_name;
    std::string table_alias;
    explicit TableScanNode(std::string name, std::string alias = "")
        : table_name(std::move(name)), table_alias(std::move(alias)) {}
    OperatorType getType() const override { return OperatorType::TABLE_SCAN; }
    void accept(OperatorVisitor& visitor) const override;
};

class FilterNode final : public OperatorNode {
public:
    std::unique_ptr<ExprNode> predicate;
    explicit FilterNode(std::unique_ptr<ExprNode> pred) : predicate(std::move(pred)) {}
    OperatorType getType() const override { return OperatorType::FILTER; }
    void accept(OperatorVisitor& visitor) const override;
};

// children_[0] = build side (dimension), children_[1] = probe side.
class HashJoinNode final : public OperatorNode {
public:
    std::unique_ptr<ExprNode> join_condition;  // filled by the optimizer
    explicit HashJoinNode(std::unique_ptr<ExprNode> cond = nullptr) : join_condition(std::move(cond)) {}
    OperatorType getType() const override { return OperatorType::HASH_JOIN; }
    void accept(OperatorVisitor& visitor) const override;
};

class ProjectionNode final : public OperatorNode {
public:
    std::vector<std::unique_ptr<ExprNode>> select_exprs;
    std::vector<std::string> output_aliases;
    OperatorType getType() const override { return OperatorType::PROJECTION; }
    void accept(OperatorVisitor& visitor) const override;
};

struct SortKeyDef {
    std::size_t column_index = 0;
    bool descending = false;
};

class SortLimitNode final : public OperatorNode {
public:
    std::vector<SortKeyDef> sort_keys;
    bool has_limit = false;
    std::size_t limit = 0;
    OperatorType getType() const override { return OperatorType::SORT_LIMIT; }
    void accept(OperatorVisitor& visitor) const override;
};

struct AggregateDef {
    std::string func_name;               // COUNT, SUM, MIN, MAX, AVG
    std::unique_ptr<ExprNode> agg_expr;  // StarExpr only for COUNT(*)

    AggregateDef(std::string name, std::unique_ptr<ExprNode> expr) : func_name(std::move(name)), agg_expr(std::move(expr)) {
        std::transform(func_name.begin(), func_name.end(), func_name.begin(), ::toupper);
        if (!(isCount() || isSum() || isMin() || isMax() || isAvg()))
            throw std::runtime_error("Unsupported aggregate function: " + func_name);
        if (!agg_expr)
            throw std::runtime_error("Aggregate function " + func_name + " requires an argument; use COUNT(*) for row count");
        if (hasStarArgument() && !isCount())
            throw std::runtime_error("Aggregate function " + func_name + " does not accept '*' according to SQL semantics");
    }
    bool isCount() const { return func_name == "COUNT"; }
    bool isSum() const { return func_name == "SUM"; }
    bool isMin() const { return func_name == "MIN"; }
    bool isMax() const { return func_name == "MAX"; }
    bool isAvg() const { return func_name == "AVG"; }
    bool hasStarArgument() const { return agg_expr && agg_expr->getType() == ExprType::STAR; }
    // AVG keeps a hidden non-null count slot (operators.h:271-276).
    bool needsNonNullCount() const { return isAvg(); }

    AggregateDef(const AggregateDef&) = delete;
    AggregateDef& operator=(const AggregateDef&) = delete;
    AggregateDef(AggregateDef&&) = default;
    AggregateDef& operator=(AggregateDef&&) = default;
};

class AggregateNode final : public OperatorNode {
public:
    std::vector<std::unique_ptr<ExprNode>> group_by_exprs;
    std::vector<AggregateDef> aggregates;
    std::unique_ptr<ExprNode> having_predicate;  // over result slots (__RESULT__.__result_col_<k>)
    std::vector<std::string> output_aliases;

    OperatorType getType() const override { return OperatorType::AGGREGATE; }
    void accept(OperatorVisitor& visitor) const override;

    uint64_t visibleTupleSize() const { return group_by_exprs.size() + aggregates.size(); }
    uint64_t hiddenCountSlotCount() const {
        uint64_t n = 0;
        for (const auto& a : aggregates) n += a.needsNonNullCount() ? 1 : 0;
        return n;
    }
    bool needsHiddenCountSlot() const { return hiddenCountSlotCount() != 0; }
    uint64_t storageTupleSize() const {
        uint64_t slots = group_by_exprs.size() + aggregates.size() + hiddenCountSlotCount();
        return slots == 0 ? 1 : slots;
    }
    // Result buffer size in u64 slots: product of group cardinalities (+1 for a
    // nullable group column) times the storage tuple (operators.h:331-369).
    uint64_t calculateResultSize(const Catalog& catalog) const;
};

class OperatorVisitor {
public:
    virtual ~OperatorVisitor() = default;
    virtual void visit(const TableScanNode& node) = 0;
    virtual void visit(const FilterNode& node) = 0;
    virtual void visit(const HashJoinNode& node) = 0;
    virtual void visit(const AggregateNode& node) = 0;
    virtual void visit(const ProjectionNode& node) = 0;
    virtual void visit(const SortLimitNode& node) = 0;
};

inline void TableScanNode::accept(OperatorVisitor& v) const { v.visit(*this); }
inline void FilterNode::accept(OperatorVisitor& v) const { v.visit(*this); }
inline void HashJoinNode::accept(OperatorVisitor& v) const { v.visit(*this); }
inline void ProjectionNode::accept(OperatorVisitor& v) const { v.visit(*this); }
inline void SortLimitNode::accept(OperatorVisitor& v) const { v.visit(*this); }
inline void AggregateNode::accept(OperatorVisitor& v) const { v.visit(*this); }

// Integer outputs of a CASE used as GROUP BY key (operators.h:20-52).
void collectCaseIntegerOutputs(const ExprNode* expr, std::vector<int64_t>& values, bool& may_be_null);

// S-expression of an optimized plan, as read by the oracle.
void printPlanSexpr(std::ostream& out, const OperatorNode* node);
// Indented human-readable tree.
void printPlanTree(std::ostream& out, const OperatorNode* node, int indent = 0);

// Re-links every parent_ pointer below `root` (the optimizer moves subtrees).
void repairParents(OperatorNode* root);

}  // namespace db
