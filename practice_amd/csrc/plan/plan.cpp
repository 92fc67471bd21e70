// Plan helpers: names, S-expression / tree printers, result-size formula.
#include <iomanip>
#include <sstream>

#include "operators.h"

namespace db {

const char* exprTypeName(ExprType t) {
    switch (t) {
        case ExprType::COLUMN_REF: return "COLUMN_REF";
        case ExprType::LITERAL_INT: return "LITERAL_INT";
        case ExprType::LITERAL_FLOAT: return "LITERAL_FLOAT";
        case ExprType::LITERAL_NULL: return "LITERAL_NULL";
        case ExprType::OP_AND: return "OP_AND";
        case ExprType::OP_OR: return "OP_OR";
        case ExprType::OP_NOT: return "OP_NOT";
        case ExprType::OP_EQ: return "OP_EQ";
        case ExprType::OP_NEQ: return "OP_NEQ";
        case ExprType::OP_LT: return "OP_LT";
        case ExprType::OP_LTE: return "OP_LTE";
        case ExprType::OP_GT: return "OP_GT";
        case ExprType::OP_GTE: return "OP_GTE";
        case ExprType::OP_ADD: return "OP_ADD";
        case ExprType::OP_SUB: return "OP_SUB";
        case ExprType::OP_MUL: return "OP_MUL";
        case ExprType::OP_DIV: return "OP_DIV";
        case ExprType::OP_IS_NULL: return "OP_IS_NULL";
        case ExprType::OP_IS_NOT_NULL: return "OP_IS_NOT_NULL";
        case ExprType::CASE_WHEN: return "CASE_WHEN";
        case ExprType::STAR: return "STAR";
    }
    return "UNKNOWN_EXPR_TYPE";
}

bool isResultSlotRef(const ExprNode* e, int& slot) {
    slot = -1;
    if (!e || e->getType() != ExprType::COLUMN_REF) return false;
    const auto* c = static_cast<const ColumnRefExpr*>(e);
    static const std::string prefix = "__result_col_";
    if (c->table_name != "__RESULT__" || c->column_name.rfind(prefix, 0) != 0) return false;
    try {
        slot = std::stoi(c->column_name.substr(prefix.size()));
    } catch (...) {
        return false;
    }
    return slot >= 0;
}

static const char* sexprOp(ExprType t) {
    switch (t) {
        case ExprType::OP_AND: return "and";
        case ExprType::OP_OR: return "or";
        case ExprType::OP_NOT: return "not";
        case ExprType::OP_EQ: return "eq";
        case ExprType::OP_NEQ: return "ne";
        case ExprType::OP_LT: return "lt";
        case ExprType::OP_LTE: return "le";
        case ExprType::OP_GT: return "gt";
        case ExprType::OP_GTE: return "ge";
        case ExprType::OP_ADD: return "add";
        case ExprType::OP_SUB: return "sub";
        case ExprType::OP_MUL: return "mul";
        case ExprType::OP_DIV: return "div";
        case ExprType::OP_IS_NULL: return "isnull";
        case ExprType::OP_IS_NOT_NULL: return "isnotnull";
        default: return "?";
    }
}

void printExprSexpr(std::ostream& out, const ExprNode* e) {
    if (!e) {
        out << "()";
        return;
    }
    int slot = -1;
    switch (e->getType()) {
        case ExprType::COLUMN_REF:
            if (isResultSlotRef(e, slot)) out << "(slot " << slot << ")";
            else out << "(col " << static_cast<const ColumnRefExpr*>(e)->column_name << ")";
            return;
        case ExprType::LITERAL_INT: out << "(int " << static_cast<const LiteralIntExpr*>(e)->value << ")"; return;
        case ExprType::LITERAL_FLOAT: {
            std::ostringstream s;
            s << std::setprecision(17) << static_cast<const LiteralFloatExpr*>(e)->value;
            out << "(float " << s.str() << ")";
            return;
        }
        case ExprType::LITERAL_NULL: out << "(null)"; return;
        case ExprType::STAR: out << "(star)"; return;
        case ExprType::CASE_WHEN: {
            const auto* c = static_cast<const CaseWhenExpr*>(e);
            out << "(case ";
            printExprSexpr(out, c->condition.get());
            out << " ";
            printExprSexpr(out, c->then_expr.get());
            out << " ";
            printExprSexpr(out, c->else_expr.get());
            out << ")";
            return;
        }
        default: {
            const auto* b = static_cast<const BinaryExpr*>(e);
            out << "(" << sexprOp(b->op_type) << " ";
            printExprSexpr(out, b->left.get());
            if (b->right) {
                out << " ";
                printExprSexpr(out, b->right.get());
            }
            out << ")";
            return;
        }
    }
}

void printPlanSexpr(std::ostream& out, const OperatorNode* n) {
    switch (n->getType()) {
        case OperatorType::TABLE_SCAN:
            out << "(scan " << static_cast<const TableScanNode*>(n)->table_name << ")";
            return;
        case OperatorType::FILTER:
            out << "(filter ";
            printExprSexpr(out, static_cast<const FilterNode*>(n)->predicate.get());
            out << " ";
            printPlanSexpr(out, n->getChildren()[0].get());
            out << ")";
            return;
        case OperatorType::HASH_JOIN:
            out << "(join ";
            printExprSexpr(out, static_cast<const HashJoinNode*>(n)->join_condition.get());
            out << " ";
            printPlanSexpr(out, n->getChildren()[0].get());
            out << " ";
            printPlanSexpr(out, n->getChildren()[1].get());
            out << ")";
            return;
        case OperatorType::AGGREGATE: {
            const auto* a = static_cast<const AggregateNode*>(n);
            out << "(agg (groups";
            for (const auto& g : a->group_by_exprs) {
                out << " ";
                printExprSexpr(out, g.get());
            }
            out << ") (aggs";
            for (const auto& d : a->aggregates) {
                out << " (" << d.func_name << " ";
                printExprSexpr(out, d.agg_expr.get());
                out << ")";
            }
            out << ") (having ";
            printExprSexpr(out, a->having_predicate.get());
            out << ") ";
            printPlanSexpr(out, n->getChildren()[0].get());
            out << ")";
            return;
        }
        case OperatorType::SORT_LIMIT: {
            const auto* s = static_cast<const SortLimitNode*>(n);
            out << "(sort (keys";
            for (const auto& k : s->sort_keys) out << " (" << k.column_index << " " << (k.descending ? "desc" : "asc") << ")";
            out << ") (limit " << (s->has_limit ? static_cast<long long>(s->limit) : -1LL) << ") ";
            printPlanSexpr(out, n->getChildren()[0].get());
            out << ")";
            return;
        }
        case OperatorType::PROJECTION: {
            const auto* p = static_cast<const ProjectionNode*>(n);
            out << "(project (exprs";
            for (const auto& e : p->select_exprs) {
                out << " ";
                printExprSexpr(out, e.get());
            }
            out << ") ";
            printPlanSexpr(out, n->getChildren()[0].get());
            out << ")";
            return;
        }
    }
}

void printPlanTree(std::ostream& out, const OperatorNode* n, int indent) {
    out << std::string(static_cast<size_t>(indent) * 2, ' ');
    switch (n->getType()) {
        case OperatorType::TABLE_SCAN: out << "TableScan " << static_cast<const TableScanNode*>(n)->table_name; break;
        case OperatorType::FILTER:
            out << "Filter ";
            printExprSexpr(out, static_cast<const FilterNode*>(n)->predicate.get());
            break;
        case OperatorType::HASH_JOIN:
            out << "HashJoin ";
            printExprSexpr(out, static_cast<const HashJoinNode*>(n)->join_condition.get());
            break;
        case OperatorType::AGGREGATE: out << "Aggregate"; break;
        case OperatorType::PROJECTION: out << "Projection"; break;
        case OperatorType::SORT_LIMIT: out << "SortLimit"; break;
    }
    out << "\n";
    for (const auto& c : n->getChildren()) printPlanTree(out, c.get(), indent + 1);
}

void repairParents(OperatorNode* root) {
    for (auto& c : root->getChildrenMutable()) {
        c->parent_ = root;
        repairParents(c.get());
    }
}

void collectCaseIntegerOutputs(const ExprNode* expr, std::vector<int64_t>& values, bool& may_be_null) {
    if (!expr) {
        may_be_null = true;
        return;
    }
    switch (expr->getType()) {
        case ExprType::LITERAL_INT: values.push_back(static_cast<const LiteralIntExpr*>(expr)->value); return;
        case ExprType::LITERAL_NULL: may_be_null = true; return;
        case ExprType::CASE_WHEN: {
            const auto* c = static_cast<const CaseWhenExpr*>(expr);
            collectCaseIntegerOutputs(c->then_expr.get(), values, may_be_null);
            collectCaseIntegerOutputs(c->else_expr.get(), values, may_be_null);
            return;
        }
        default: throw std::runtime_error("GROUP BY CASE currently requires integer THEN/ELSE result expressions");
    }
}

uint64_t AggregateNode::calculateResultSize(const Catalog& catalog) const {
    if (group_by_exprs.empty()) return storageTupleSize();
    uint64_t groups = 1;
    for (const auto& g : group_by_exprs) {
        uint64_t card = 1;
        if (g->getType() == ExprType::COLUMN_REF) {
            const auto& col = static_cast<const ColumnRefExpr*>(g.get())->column_name;
            const std::string t = catalog.tableOfColumn(col);
            if (catalog.hasTable(t)) {
                const auto& meta = catalog.getTableMetadata(t);
                card = meta.hasColumnStats(col) ? meta.getColumnStats(col).cardinality_ : meta.getSize();
                if (meta.isColumnNullable(col)) ++card;
            }
        } else if (g->getType() == ExprType::CASE_WHEN) {
            std::vector<int64_t> vals;
            bool may_null = false;
            collectCaseIntegerOutputs(g.get(), vals, may_null);
            if (!vals.empty()) {
                const auto mm = std::minmax_element(vals.begin(), vals.end());
                card = static_cast<uint64_t>(*mm.second - *mm.first + 1) + (may_null ? 1 : 0);
            }
            if (card == 0) card = 1;
        }
        groups *= card;
    }
    return groups * storageTupleSize();
}

}  // namespace db
