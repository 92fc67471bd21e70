// HIP code generator: plan analysis + kernel source emission.
//
// Analysis restates the reference generator's decisions (visitor.cpp):
//   - build side = first child that is a (filtered) dimension scan (4466-4528)
//   - hash-table domain: PART/SUPPLIER/CUSTOMER len = table rows, min 0;
//     DDATE len 61131, min 19920101; otherwise the first *key* column's range
//     (computeBuildInfo 3864-3917)
//   - one payload column per dimension: the first dimension column needed by
//     an ancestor, aggregate columns before filter columns (3967-4013)
//   - fact filters: AND chains of `col op int` take the fast path on raw values
//     (no validity); anything else the 3VL path (597-667, 1823-1893)
//   - build filters: AND/OR of `col op lit` on raw values (4047-4201)
//   - payload registers carry no validity (4573-4579)
//   - perfect GROUP BY hash, Horner over catalog (min, distinct count), +1 and
//     NULL code 0 for catalog-nullable columns (2056-2127)
//   - aggregate input: leftmost leaf cast to unsigned long long (1947-2018)
//   - sparse output eligibility (2215-2231), result types / nullability (77-185)
// Emission is MI355X-native: the fact kernel streams 128-bit loads, keeps one
// flag bit per item, probes compact tables staged in LDS or resident in L2,
// and privatizes GROUP BY aggregates in LDS (u64 ds atomics) before one global
// merge per block.
#include "hip_codegen.h"
#include "vm_expr.h"
#include "../kernels/aot_kernels.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <regex>
#include <sstream>

namespace db {

std::string signatureOf(const std::string& text) {
    uint64_t h = 0xCBF29CE484222325ULL;
    for (unsigned char c : text) {
        h ^= c;
        h *= 0x100000001B3ULL;
    }
    char buf[17];
    snprintf(buf, sizeof buf, "%016llx", static_cast<unsigned long long>(h));
    return buf;
}

namespace {

[[noreturn]] void unsupported(const std::string& what) { throw std::runtime_error(what); }

// result column name of an expression without an alias
// (defaultExpressionColumnName, execution.cpp:540-557)
std::string defaultColumnName(const ExprNode* e) {
    if (!e) return "expr";
    switch (e->getType()) {
        case ExprType::COLUMN_REF: return static_cast<const ColumnRefExpr*>(e)->column_name;
        case ExprType::STAR: return "*";
        case ExprType::CASE_WHEN: return "case_when";
        case ExprType::LITERAL_INT:
        case ExprType::LITERAL_FLOAT: return "literal";
        case ExprType::LITERAL_NULL: return "NULL";
        default: return "expr";
    }
}

bool litTrue(const std::string& s) { return s == "1" || s == "true" || s == "(1)" || s == "(true)"; }
bool litFalse(const std::string& s) { return s == "0" || s == "false" || s == "(0)" || s == "(false)"; }

std::string andTerms(std::initializer_list<std::string> terms) {
    std::string out;
    for (const auto& t : terms) {
        if (t.empty() || litTrue(t)) continue;
        if (litFalse(t)) return "0";
        if (!out.empty()) out += " && ";
        out += t;
    }
    return out.empty() ? "1" : "(" + out + ")";
}
std::string boolAnd(const std::string& a, const std::string& b) {
    if (litFalse(a) || litFalse(b)) return "0";
    if (litTrue(a)) return b;
    if (litTrue(b)) return a;
    return "(" + a + " && " + b + ")";
}
std::string boolOr(const std::string& a, const std::string& b) {
    if (litTrue(a) || litTrue(b)) return "1";
    if (litFalse(a)) return b;
    if (litFalse(b)) return a;
    return "(" + a + " || " + b + ")";
}
std::string boolNot(const std::string& a) {
    if (litTrue(a)) return "0";
    if (litFalse(a)) return "1";
    return "(!(" + a + "))";
}

const char* cmpOp(ExprType t) {
    switch (t) {
        case ExprType::OP_EQ: return "==";
        case ExprType::OP_NEQ: return "!=";
        case ExprType::OP_LT: return "<";
        case ExprType::OP_LTE: return "<=";
        case ExprType::OP_GT: return ">";
        default: return ">=";
    }
}
const char* safeFn(ExprType t) {
    switch (t) {
        case ExprType::OP_ADD: return "chip::safe_add";
        case ExprType::OP_SUB: return "chip::safe_sub";
        case ExprType::OP_MUL: return "chip::safe_mul";
        case ExprType::OP_DIV: return "chip::safe_div";
        case ExprType::OP_EQ: return "chip::safe_eq";
        case ExprType::OP_NEQ: return "chip::safe_neq";
        case ExprType::OP_LT: return "chip::safe_lt";
        case ExprType::OP_LTE: return "chip::safe_lte";
        case ExprType::OP_GT: return "chip::safe_gt";
        default: return "chip::safe_gte";
    }
}
ExprType swapCmp(ExprType t) {
    switch (t) {
        case ExprType::OP_LT: return ExprType::OP_GT;
        case ExprType::OP_LTE: return ExprType::OP_GTE;
        case ExprType::OP_GT: return ExprType::OP_LT;
        case ExprType::OP_GTE: return ExprType::OP_LTE;
        default: return t;
    }
}

void collectColumns(const ExprNode* e, std::vector<std::string>& out) {
    if (!e) return;
    switch (e->getType()) {
        case ExprType::COLUMN_REF: {
            const auto& n = static_cast<const ColumnRefExpr*>(e)->column_name;
            if (std::find(out.begin(), out.end(), n) == out.end()) out.push_back(n);
            return;
        }
        case ExprType::CASE_WHEN: {
            const auto* c = static_cast<const CaseWhenExpr*>(e);
            collectColumns(c->condition.get(), out);
            collectColumns(c->then_expr.get(), out);
            collectColumns(c->else_expr.get(), out);
            return;
        }
        default:
            if (isBinaryOp(e->getType())) {
                const auto* b = static_cast<const BinaryExpr*>(e);
                collectColumns(b->left.get(), out);
                collectColumns(b->right.get(), out);
            }
    }
}

bool containsFloat(const ExprNode* e) {
    if (!e) return false;
    if (e->getType() == ExprType::LITERAL_FLOAT) return true;
    if (e->getType() == ExprType::CASE_WHEN) {
        const auto* c = static_cast<const CaseWhenExpr*>(e);
        return containsFloat(c->condition.get()) || containsFloat(c->then_expr.get()) || containsFloat(c->else_expr.get());
    }
    if (isBinaryOp(e->getType())) {
        const auto* b = static_cast<const BinaryExpr*>(e);
        return containsFloat(b->left.get()) || containsFloat(b->right.get());
    }
    return false;
}

enum class FilterPath { Fast, Universal, Unsupported };

// validateFilterFastPathPredicate (visitor.cpp:597-667)
FilterPath classifyFilter(const ExprNode* e, std::string& why) {
    if (!e) return FilterPath::Fast;
    const ExprType t = e->getType();
    if (t == ExprType::OP_AND) {
        const auto* b = static_cast<const BinaryExpr*>(e);
        const FilterPath l = classifyFilter(b->left.get(), why);
        if (l == FilterPath::Unsupported) return l;
        const FilterPath r = classifyFilter(b->right.get(), why);
        if (r == FilterPath::Unsupported) return r;
        return (l == FilterPath::Universal || r == FilterPath::Universal) ? FilterPath::Universal : FilterPath::Fast;
    }
    if (t == ExprType::OP_OR || t == ExprType::OP_NOT || t == ExprType::OP_IS_NULL || t == ExprType::OP_IS_NOT_NULL)
        return FilterPath::Universal;
    if (!isComparison(t)) {
        why = std::string("Unsupported filter predicate node: ") + exprTypeName(t);
        return FilterPath::Unsupported;
    }
    const auto* c = static_cast<const BinaryExpr*>(e);
    const ExprType lt = c->left->getType(), rt = c->right->getType();
    if (lt == ExprType::COLUMN_REF && rt == ExprType::LITERAL_INT) return FilterPath::Fast;
    if (lt == ExprType::COLUMN_REF && rt == ExprType::COLUMN_REF) return FilterPath::Universal;
    if (isArithmetic(lt) || isArithmetic(rt)) return FilterPath::Universal;
    if (lt == ExprType::LITERAL_FLOAT || rt == ExprType::LITERAL_FLOAT) {
        why = "Unsupported float literal in fast-path predicate";
        return FilterPath::Unsupported;
    }
    why = std::string("Unsupported comparison shape in fast-path predicate, op=") + exprTypeName(t) +
          ", left=" + exprTypeName(lt) + ", right=" + exprTypeName(rt);
    return FilterPath::Unsupported;
}

bool buildFilterSupported(const ExprNode* e) {
    if (!e) return true;
    const ExprType t = e->getType();
    if (t == ExprType::OP_AND || t == ExprType::OP_OR) {
        const auto* b = static_cast<const BinaryExpr*>(e);
        return buildFilterSupported(b->left.get()) && buildFilterSupported(b->right.get());
    }
    if (!isComparison(t)) return false;
    const auto* c = static_cast<const BinaryExpr*>(e);
    auto isLit = [](const ExprNode* x) {
        return x->getType() == ExprType::LITERAL_INT || x->getType() == ExprType::LITERAL_FLOAT;
    };
    return (c->left->getType() == ExprType::COLUMN_REF && isLit(c->right.get())) ||
           (c->right->getType() == ExprType::COLUMN_REF && isLit(c->left.get()));
}

// std::to_string(double) as C source sees it ("%f")
double printedDouble(double v) {
    char buf[512];
    snprintf(buf, sizeof buf, "%f", v);
    return strtod(buf, nullptr);
}

int bitsFor(uint64_t range) {  // bits to encode [0, range]
    int b = 0;
    while (b < 64 && (range >> b) != 0) ++b;
    return b;
}

struct Op {
    int kind;  // 0 filter, 1 join
    const ExprNode* pred;
    bool fast;
    int join;
};

class Gen {
public:
    Gen(const Catalog& c, const CodegenOptions& o) : cat_(c), opt_(o) {}

    QueryProgram run(const OperatorNode* root) {
        std::ostringstream sx;
        printPlanSexpr(sx, root);
        prog_.plan_sexpr = sx.str();
        analyze(root);
        emit();
        // a batched template is identified by its per-tile body and by the
        // columns it binds as well: different pipelines can share identical
        // build / flag kernels, and plans over different columns can share
        // identical text (a launch binds one column set per template)
        std::string binding;
        if (!prog_.batch_body.empty()) {
            binding = "|" + prog_.fact_table;
            for (const auto& c : prog_.fact_cols) binding += "|f:" + c;
            for (const auto& c : prog_.dim_cols) binding += "|d:" + c;
        }
        prog_.signature = signatureOf(prog_.batch_body.empty()
                                          ? prog_.source
                                          : prog_.source + prog_.batch_body + prog_.batch_mask_build + prog_.batch_mask_body +
                                                prog_.batch_mask_body_deferred + binding);
        return std::move(prog_);
    }

private:
    const Catalog& cat_;
    CodegenOptions opt_;
    QueryProgram prog_;
    const AggregateNode* agg_ = nullptr;
    const SortLimitNode* sort_ = nullptr;
    const ProjectionNode* proj_ = nullptr;
    std::vector<Op> ops_;
    std::vector<const ExprNode*> build_pred_;
    bool build_filtered_ = false;  // join being analysed has a build-side filter
    int flag_w_ = 0;                // LDS flag bits per group (0: LDS aggregation impossible)
    std::vector<int> flag_dyn_, flag_static_;
    std::map<std::string, int> fact_slot_;
    std::map<std::string, int> dim_slot_;
    unsigned visible_mask_ = 0;  // joins whose payload registers exist at this point of the pipeline
    // GROUP BY hash parameters
    std::vector<int64_t> gmin_;
    std::vector<uint64_t> gcard_;
    std::vector<bool> gnullable_;
    std::vector<bool> gdense_;

    // ---------------------------------------------------------------- catalog
    std::string tableOf(const std::string& col) const { return cat_.tableOfColumn(col); }
    const TableMetadata& meta(const std::string& t) const { return cat_.getTableMetadata(t); }
    bool nullable(const std::string& col) const { return cat_.columnNullable(col); }
    const ColumnStatistics* stats(const std::string& col) const {
        const std::string t = tableOf(col);
        if (!cat_.hasTable(t)) return nullptr;
        const auto& m = meta(t);
        return m.hasColumnStats(col) ? &m.getColumnStats(col) : nullptr;
    }
    bool uniqueKey(const std::string& table, const std::string& col) const {
        const auto& m = meta(table);
        if (m.isColumnPrimaryKey(col) || m.isColumnUnique(col)) return true;
        const auto* s = stats(col);
        return s && s->cardinality_ > 0 && s->cardinality_ == m.getSize();
    }
    bool isDimOp(const OperatorNode* n) const {
        const TableScanNode* scan = nullptr;
        if (n->getType() == OperatorType::TABLE_SCAN) scan = static_cast<const TableScanNode*>(n);
        else if (n->getType() == OperatorType::FILTER && !n->getChildren().empty() &&
                 n->getChildren()[0]->getType() == OperatorType::TABLE_SCAN)
            scan = static_cast<const TableScanNode*>(n->getChildren()[0].get());
        return scan && cat_.hasTable(scan->table_name) && !meta(scan->table_name).isFactTable();
    }

    // ---------------------------------------------------------------- analysis
    void analyze(const OperatorNode* root) {
        const OperatorNode* n = root;
        if (n->getType() == OperatorType::SORT_LIMIT) {
            sort_ = static_cast<const SortLimitNode*>(n);
            n = n->getChildren()[0].get();
        }
        if (n->getType() == OperatorType::PROJECTION) {
            if (opt_.vm) unsupported("interpreted tier: projections run compiled");
            if (opt_.batch_body) unsupported("projections do not run in the shared scan");
            proj_ = static_cast<const ProjectionNode*>(n);
        } else if (n->getType() != OperatorType::AGGREGATE) {
            unsupported("plan root must be an aggregate or a projection");
        } else {
            agg_ = static_cast<const AggregateNode*>(n);
        }

        std::vector<const OperatorNode*> chain, build;
        const OperatorNode* cur = n->getChildren()[0].get();
        while (cur->getType() != OperatorType::TABLE_SCAN) {
            if (cur->getType() == OperatorType::FILTER) {
                chain.push_back(cur);
                build.push_back(nullptr);
                cur = cur->getChildren()[0].get();
            } else if (cur->getType() == OperatorType::HASH_JOIN) {
                const auto& ch = cur->getChildren();
                const OperatorNode *b = nullptr, *p = nullptr;
                if (isDimOp(ch[0].get())) b = ch[0].get(), p = ch[1].get();
                else if (isDimOp(ch[1].get())) b = ch[1].get(), p = ch[0].get();
                else unsupported("hash join without a dimension build side");
                chain.push_back(cur);
                build.push_back(b);
                cur = p;
            } else {
                unsupported("unsupported operator below the aggregate");
            }
        }
        const auto* scan = static_cast<const TableScanNode*>(cur);
        if (!cat_.hasTable(scan->table_name)) unsupported("Table not found: " + scan->table_name);
        if (!meta(scan->table_name).isFactTable()) unsupported("the probe pipeline must scan the fact table");
        prog_.fact_table = scan->table_name;
        prog_.fact_rows = static_cast<int64_t>(meta(scan->table_name).getSize());

        for (int i = static_cast<int>(chain.size()) - 1; i >= 0; --i) {
            const OperatorNode* node = chain[static_cast<size_t>(i)];
            Op op{};
            if (node->getType() == OperatorType::FILTER) {
                const auto* f = static_cast<const FilterNode*>(node);
                std::string why;
                const FilterPath path = classifyFilter(f->predicate.get(), why);
                if (path == FilterPath::Unsupported)
                    unsupported("Filter predicate is not translatable before JIT C++ compilation: " + why);
                op.kind = 0;
                op.pred = f->predicate.get();
                op.fast = path == FilterPath::Fast;
            } else {
                JoinProgram jp;
                const ExprNode* bpred = nullptr;
                buildInfo(static_cast<const HashJoinNode*>(node), build[static_cast<size_t>(i)], jp, bpred);
                for (const auto& other : prog_.joins)
                    if (other.dim_table == jp.dim_table) unsupported("dimension " + jp.dim_table + " joined twice");
                op.kind = 1;
                op.join = static_cast<int>(prog_.joins.size());
                prog_.joins.push_back(jp);
                build_pred_.push_back(bpred);
            }
            ops_.push_back(op);
        }
        if (prog_.joins.size() > CQ_MAX_JOINS_HOST) unsupported("too many joins");
        if (proj_) analyzeProjection();
        else analyzeAggregate();
    }

    // consumeProjectionVector (visitor.cpp:3466-3548): the output columns,
    // each expression's result type and nullability as for GROUP BY columns
    std::vector<const ExprNode*> proj_exprs_;
    std::vector<std::unique_ptr<ExprNode>> proj_star_;
    void analyzeProjection() {
        for (const auto& e : proj_->select_exprs) {
            if (!e) continue;
            if (e->getType() != ExprType::STAR) {
                proj_exprs_.push_back(e.get());
                continue;
            }
            // every column of every scanned table, tables in plan order
            // (collectTableScans: depth first, children in order), dimension
            // columns through row ids (consumeProjectionVector, visitor.cpp:3475-3490)
            std::vector<std::string> tables;
            std::function<void(const OperatorNode*)> scans = [&](const OperatorNode* n) {
                if (n->getType() == OperatorType::TABLE_SCAN) tables.push_back(static_cast<const TableScanNode*>(n)->table_name);
                for (const auto& c : n->getChildren()) scans(c.get());
            };
            scans(proj_->getChildren()[0].get());
            for (const auto& t : tables)
                for (const auto& c : meta(t).getColumnNames()) {
                    proj_star_.push_back(std::make_unique<ColumnRefExpr>(c));
                    proj_exprs_.push_back(proj_star_.back().get());
                }
        }
        if (proj_exprs_.empty()) unsupported("projection without columns");
        if (proj_exprs_.size() > 64) unsupported("more than 64 projection columns");
        for (const auto& jp : prog_.joins)
            if (jp.fk_alts.size() > 1) unsupported("OR-key joins in a projection are not supported");
        prog_.projection = true;
        prog_.ngroups = 0;
        prog_.naggs = 0;
        prog_.grouped = false;
        prog_.visible_ts = prog_.storage_ts = static_cast<int>(proj_exprs_.size());
        prog_.groups = 0;
        prog_.fast_sparse = false;
        for (const ExprNode* e : proj_exprs_) {
            OutputColumn oc;
            oc.type = resultType(e);
            oc.nullable = resultNullable(e);
            prog_.out.push_back(oc);
        }
        for (size_t c = 0; c < prog_.out.size(); ++c) {
            if (c < proj_->output_aliases.size()) prog_.out[c].name = proj_->output_aliases[c];
            if (prog_.out[c].name.empty()) prog_.out[c].name = defaultColumnName(proj_exprs_[c]);
        }
        prog_.has_sort_node = sort_ != nullptr;
        if (sort_) {
            prog_.sort_keys = sort_->sort_keys;
            prog_.has_limit = sort_->has_limit;
            prog_.limit = sort_->limit;
            for (const auto& k : prog_.sort_keys)
                if (k.column_index >= static_cast<size_t>(prog_.visible_ts))
                    unsupported("ORDER BY key index exceeds generated result width");
        }
    }
    // resultLogicalTypeForExpr (visitor.cpp:77-107)
    static LogicalType resultType(const ExprNode* e) {
        switch (e->getType()) {
            case ExprType::OP_EQ: case ExprType::OP_NEQ: case ExprType::OP_LT: case ExprType::OP_LTE:
            case ExprType::OP_GT: case ExprType::OP_GTE: case ExprType::OP_AND: case ExprType::OP_OR:
            case ExprType::OP_NOT: case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
            case ExprType::LITERAL_NULL:
                return LogicalType::UInt64;
            case ExprType::LITERAL_FLOAT: return LogicalType::Float64;
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                const auto x = resultType(c->then_expr.get()), y = resultType(c->else_expr.get());
                if (x == LogicalType::Float64 || y == LogicalType::Float64) return LogicalType::Float64;
                if (x == LogicalType::Int64 || y == LogicalType::Int64) return LogicalType::Int64;
                return LogicalType::UInt64;
            }
            default: return LogicalType::Int64;
        }
    }
    bool resultNullable(const ExprNode* e) const {
        if (!e) return true;
        switch (e->getType()) {
            case ExprType::LITERAL_NULL: return true;
            case ExprType::STAR: case ExprType::LITERAL_INT: case ExprType::LITERAL_FLOAT:
            case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
                return false;
            case ExprType::COLUMN_REF: return nullable(static_cast<const ColumnRefExpr*>(e)->column_name);
            case ExprType::OP_NOT: return resultNullable(static_cast<const BinaryExpr*>(e)->left.get());
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                return resultNullable(c->then_expr.get()) || resultNullable(c->else_expr.get());
            }
            default: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                return resultNullable(b->left.get()) || resultNullable(b->right.get());
            }
        }
    }
    static constexpr size_t CQ_MAX_JOINS_HOST = 8;

    void buildInfo(const HashJoinNode* join, const OperatorNode* build_side, JoinProgram& jp, const ExprNode*& bpred) {
        const TableScanNode* scan = build_side->getType() == OperatorType::TABLE_SCAN
                                        ? static_cast<const TableScanNode*>(build_side)
                                        : static_cast<const TableScanNode*>(build_side->getChildren()[0].get());
        jp.dim_table = scan->table_name;
        const auto& m = meta(jp.dim_table);
        jp.dim_rows = static_cast<int64_t>(m.getSize());
        bpred = build_side->getType() == OperatorType::FILTER ? static_cast<const FilterNode*>(build_side)->predicate.get() : nullptr;
        if (bpred && !buildFilterSupported(bpred)) unsupported("Unsupported build-side filter predicate for block codegen");
        build_filtered_ = bpred != nullptr;
        const std::string& t = jp.dim_table;
        if (t == "PART" || t == "SUPPLIER" || t == "CUSTOMER") {
            jp.len = jp.dim_rows;
            jp.kmin = 0;
        } else if (t == "DDATE") {
            jp.len = 61131;
            jp.kmin = 19920101;
        } else {
            jp.len = 0;
            for (const auto& c : m.getColumnNames()) {
                if (c.find("key") != std::string::npos || c.find("Key") != std::string::npos) {
                    if (m.hasColumnStats(c)) {
                        const auto& s = m.getColumnStats(c);
                        const int64_t range = s.max_value_ - s.min_value_ + 1;
                        if (range > 0) jp.kmin = s.min_value_, jp.len = range;
                        break;
                    }
                }
            }
            if (jp.len == 0) jp.len = jp.dim_rows;
        }
        if (jp.len <= 0) unsupported("empty dimension table " + t);
        // keys from the join condition
        std::function<void(const ExprNode*)> keys = [&](const ExprNode* e) {
            if (!e) return;
            if (e->getType() == ExprType::OP_EQ) {
                const auto* b = static_cast<const BinaryExpr*>(e);
                if (b->left->getType() == ExprType::COLUMN_REF && b->right->getType() == ExprType::COLUMN_REF) {
                    const auto& l = static_cast<const ColumnRefExpr*>(b->left.get())->column_name;
                    const auto& r = static_cast<const ColumnRefExpr*>(b->right.get())->column_name;
                    // every key pair adds a fact key alternative (visitor.cpp:3921-3945)
                    const std::string pk = tableOf(l) == t ? l : r, fk = tableOf(l) == t ? r : l;
                    if (!jp.pk_col.empty() && jp.pk_col != pk)
                        unsupported("OR-key join over two dimension keys (" + jp.pk_col + ", " + pk + ")");
                    jp.pk_col = pk;
                    jp.fk_alts.push_back(fk);
                    if (jp.fk_col.empty()) jp.fk_col = fk;
                }
            } else if (e->getType() == ExprType::OP_AND || e->getType() == ExprType::OP_OR) {
                const auto* b = static_cast<const BinaryExpr*>(e);
                keys(b->left.get());
                keys(b->right.get());
            }
        };
        keys(join->join_condition.get());
        if (jp.pk_col.empty()) unsupported("hash join without an equi-key condition");
        for (const auto& fk : jp.fk_alts)
            if (tableOf(fk) != prog_.fact_table && !prog_.fact_table.empty())
                unsupported("join key " + fk + " is not a fact-table column");
        if (jp.fk_alts.size() > 1) {
            if (opt_.vm) unsupported("interpreted tier: OR-key joins run on compiled modules");
            if (opt_.batch_body) unsupported("OR-key joins do not run in the shared scan");
            for (const auto& other : prog_.joins)
                if (&other != &jp && other.fk_alts.size() > 1) unsupported("more than one OR-key join in a plan");
            if (jp.fk_alts.size() > 8) unsupported("more than 8 key alternatives in one join");
        }
        if (!uniqueKey(t, jp.pk_col)) {
            // MHT when the catalog proves the key non-unique (visitor.cpp:
            // 3946-3961); it always carries the build row id (3963-3970)
            const auto* ks = stats(jp.pk_col);
            if (!ks || ks->cardinality_ == 0 || ks->cardinality_ >= m.getSize())
                unsupported("build key " + jp.pk_col + " is neither unique nor known to be non-unique in the catalog");
            if (opt_.vm) unsupported("interpreted tier: multi-value joins run on compiled modules");
            if (opt_.batch_body) unsupported("multi-value joins do not run in the shared scan");
            if (jp.fk_alts.size() > 1) unsupported("OR-key join on a non-unique build key");
            for (const auto& other : prog_.joins)
                if (&other != &jp && (other.mht || other.fk_alts.size() > 1))
                    unsupported("a multi-value join with another multi-value or OR-key join in one plan");
            if (jp.len > (int64_t(1) << 30) || jp.dim_rows > (int64_t(1) << 31) - 1)
                unsupported("multi-value table too large");
            jp.mht = true;
            jp.rowid = true;
            jp.pair = true;
            const uint64_t len = static_cast<uint64_t>(jp.len);
            const uint64_t sums = 2 * static_cast<uint64_t>(cq_scan_block_count(jp.len) + 2);
            jp.mht_cnt = 4 * len;
            jp.mht_off = 6 * len;
            jp.mht_sums = 8 * len;
            jp.mht_pos = 8 * len + sums;
            jp.mht_rows = 9 * len + sums;
            jp.ht_words = jp.mht_rows + static_cast<uint64_t>(jp.dim_rows);
            return;
        }
        // a projection above that reads this dimension gathers its columns
        // by row id (projectionRequiresRowIdPayload, visitor.cpp:3300-3317)
        for (const OperatorNode* p = join->parent_; p; p = p->parent_) {
            if (p->getType() != OperatorType::PROJECTION) continue;
            for (const auto& e : static_cast<const ProjectionNode*>(p)->select_exprs) {
                if (!e) continue;
                if (e->getType() == ExprType::STAR) jp.rowid = true;
                std::vector<std::string> cols;
                collectColumns(e.get(), cols);
                for (const auto& c : cols) jp.rowid = jp.rowid || tableOf(c) == t;
            }
        }
        if (jp.rowid) {
            jp.pair = true;
            chooseLayout(jp);
            return;
        }
        // payload: aggregate columns first, then filter columns above the join
        std::vector<std::string> aggc, fltc;
        for (const OperatorNode* p = join->parent_; p; p = p->parent_) {
            if (p->getType() == OperatorType::AGGREGATE) {
                const auto* a = static_cast<const AggregateNode*>(p);
                for (const auto& d : a->aggregates) collectColumns(d.agg_expr.get(), aggc);
                for (const auto& g : a->group_by_exprs) collectColumns(g.get(), aggc);
            } else if (p->getType() == OperatorType::FILTER) {
                collectColumns(static_cast<const FilterNode*>(p)->predicate.get(), fltc);
            }
        }
        for (const auto* lst : {&aggc, &fltc}) {
            for (const auto& c : *lst)
                if (tableOf(c) == t && c != jp.pk_col) {
                    jp.payload_col = c;
                    break;
                }
            if (!jp.payload_col.empty()) break;
        }
        jp.pair = !jp.payload_col.empty();
        chooseLayout(jp);
    }

    void chooseLayout(JoinProgram& jp) const {
        jp.compact = false;
        const auto* ks = stats(jp.pk_col);
        if (opt_.allow_compact_tables && ks && !nullable(jp.pk_col) && uniqueKey(jp.dim_table, jp.pk_col)) {
            const int64_t lo = ks->min_value_ - jp.kmin, hi = ks->max_value_ - jp.kmin;
            if (lo >= 0 && hi - lo + 1 <= jp.len && hi < (int64_t(1) << 31)) {
                jp.jlo = static_cast<int>(lo / jp.len);
                const int jhi = static_cast<int>(hi / jp.len);
                jp.tb = jhi > jp.jlo ? 1 : 0;
                int code_bits = 0;
                jp.pmin = 0;
                bool ok = true;
                if (jp.rowid) {
                    code_bits = bitsFor(static_cast<uint64_t>(std::max<int64_t>(jp.dim_rows - 1, 0)));
                } else if (jp.pair) {
                    const auto* ps = stats(jp.payload_col);
                    if (!ps) ok = false;
                    else {
                        jp.pmin = ps->min_value_;
                        code_bits = bitsFor(static_cast<uint64_t>(ps->max_value_ - ps->min_value_));
                    }
                }
                int need = 1 + jp.tb + code_bits;
                int eb = 1;
                while (eb < need) eb <<= 1;
                if (ok && eb <= 32) {
                    jp.compact = true;
                    jp.eb = eb;
                }
            }
        }
        if (jp.compact) jp.ht_words = (static_cast<uint64_t>(jp.len) * static_cast<uint64_t>(jp.eb) + 31) / 32;
        else jp.ht_words = 2 * static_cast<uint64_t>(jp.len);
        // key-indexed candidate: same key conditions as the compact layout
        jp.dense_ok = false;
        if (opt_.allow_dense_tables && jp.compact) {
            const auto* ks2 = stats(jp.pk_col);
            jp.dmin = ks2->min_value_;
            jp.dspan = ks2->max_value_ - ks2->min_value_ + 1;
            const auto& tm = meta(jp.dim_table);
            const bool all_present = !build_filtered_ && static_cast<uint64_t>(jp.dspan) == tm.getSize() &&
                                     ks2->cardinality_ == tm.getSize();
            jp.doff = all_present ? 0 : 1;
            uint64_t range = 0;
            if (jp.rowid) {
                range = static_cast<uint64_t>(std::max<int64_t>(jp.dim_rows - 1, 0));
            } else if (jp.pair) {
                const auto* ps = stats(jp.payload_col);
                range = static_cast<uint64_t>(ps->max_value_ - ps->min_value_);
            }
            jp.dbits = jp.pair ? bitsFor(range + static_cast<uint64_t>(jp.doff)) : (all_present ? 0 : 1);
            jp.dense_ok = jp.dbits <= 31 && jp.dspan > 0 && jp.dspan <= (int64_t(1) << 31);
            // key 0 against the zero-filled reference: which key owns its slot
            jp.zmode = 0;
            jp.zx = 0;
            if (jp.kmin <= 0) {
                const int64_t r = (-jp.kmin) % jp.len;
                const int64_t first = jp.kmin + r;  // smallest key >= kmin in key 0's slot
                int64_t t = jp.dmin > first ? (jp.dmin - first + jp.len - 1) / jp.len : 0;
                const int64_t k0 = first + t * jp.len;
                if (k0 > jp.dmin + jp.dspan - 1) jp.zmode = 1;
                else {
                    jp.zmode = k0 == 0 ? 3 : 2;
                    jp.zx = static_cast<int>(k0 - jp.dmin);
                }
                // every key of [dmin, dmin + dspan) present: the slot owner
                // is always in the table, so key 0 matches exactly when it
                // is a key itself (mode 3, its own entry) and never otherwise
                // (mode 2, where 0 lies outside the key range): the plain
                // entry lookup already says so
                if (jp.doff == 0 && jp.zmode >= 2) {
                    jp.zmode = 0;
                    jp.zx = 0;
                }
            }
            // calendar keys (yyyymmdd dates): index by (year, month, day),
            // 372 entries a year instead of 10,000, when that is smaller.
            // Some of those entries are no date, so code 0 marks absent keys.
            jp.cal = 0;
            if (jp.dense_ok && opt_.calendar_tables && ks2->is_calendar_ && jp.zmode == 0) {
                const int64_t y0 = ks2->min_value_ / 10000, y1 = ks2->max_value_ / 10000;
                const int64_t span = (y1 - y0 + 1) * 372;
                if (y0 > 0 && span < jp.dspan) {
                    jp.cal = static_cast<int>(y0);
                    jp.dmin = y0 * 10000 + 101;
                    jp.dspan = span;
                    jp.doff = 1;
                    jp.dbits = jp.pair ? bitsFor(range + 1) : 1;
                    jp.dense_ok = jp.dbits <= 31;
                }
            }
        }
    }

    static int pow2Width(int bits) {
        int w = bits <= 0 ? 0 : 1;
        while (w && w < bits) w <<= 1;
        return w;
    }
    static uint64_t denseWords(const JoinProgram& jp, int width) {
        if (width == 0) return 0;
        return (static_cast<uint64_t>(jp.dspan) * static_cast<uint64_t>(width) + 31) / 32 + 1;  // + padding word
    }
    // Final table layout once LDS placement is known: key-indexed when it is
    // no larger than the compact table (exact widths in LDS, power-of-two
    // widths in HBM so a probe is one load).
    void finalizeLayout(JoinProgram& jp, bool lds) {
        jp.lds = lds;
        jp.dense = false;
        if (jp.dense_ok) {
            const int w = lds ? jp.dbits : pow2Width(jp.dbits);
            const uint64_t dw = denseWords(jp, w);
            const uint64_t cw = (static_cast<uint64_t>(jp.len) * static_cast<uint64_t>(jp.eb) + 31) / 32;
            if (w <= 31 && dw <= cw + 1) {  // + the padding word: at equal widths the key-indexed probe is cheaper
                jp.dense = true;
                jp.db = w;
                jp.ht_words = dw;
                return;
            }
        }
        if (jp.compact) jp.ht_words = (static_cast<uint64_t>(jp.len) * static_cast<uint64_t>(jp.eb) + 31) / 32;
    }
    int64_t densePbase(const JoinProgram& jp) const { return (jp.pair && !jp.rowid ? jp.pmin : 0) - jp.doff; }
    // LDS bytes the table needs when placed there (0: not placeable)
    uint64_t ldsBytes(const JoinProgram& jp) const {
        if (jp.dense_ok) return denseWords(jp, jp.dbits) * 4;
        if (jp.compact) return ((static_cast<uint64_t>(jp.len) * static_cast<uint64_t>(jp.eb) + 31) / 32) * 4;
        return 0;
    }

    void analyzeAggregate() {
        const auto* a = agg_;
        prog_.ngroups = static_cast<int>(a->group_by_exprs.size());
        prog_.naggs = static_cast<int>(a->aggregates.size());
        prog_.visible_ts = static_cast<int>(a->visibleTupleSize());
        prog_.storage_ts = static_cast<int>(a->storageTupleSize());
        prog_.grouped = prog_.ngroups > 0;
        prog_.has_having = a->having_predicate != nullptr;
        int hidden = 0;
        for (int i = 0; i < prog_.naggs; ++i) {
            const auto& d = a->aggregates[static_cast<size_t>(i)];
            prog_.agg_funcs.push_back(d.func_name);
            prog_.hidden_slot.push_back(d.isAvg() ? prog_.visible_ts + hidden++ : -1);
            if (d.isAvg()) prog_.has_avg = true;
            if (d.isMin()) prog_.min_slots.push_back(prog_.ngroups + i);
            if (!d.isCount() && containsFloat(d.agg_expr.get()))
                unsupported("floating-point aggregate inputs are not supported");
        }
        // output layout
        std::function<LogicalType(const ExprNode*)> outType = [&](const ExprNode* e) -> LogicalType {
            switch (e->getType()) {
                case ExprType::OP_EQ: case ExprType::OP_NEQ: case ExprType::OP_LT: case ExprType::OP_LTE:
                case ExprType::OP_GT: case ExprType::OP_GTE: case ExprType::OP_AND: case ExprType::OP_OR:
                case ExprType::OP_NOT: case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
                case ExprType::LITERAL_NULL:
                    return LogicalType::UInt64;
                case ExprType::LITERAL_FLOAT: return LogicalType::Float64;
                case ExprType::CASE_WHEN: {
                    const auto* c = static_cast<const CaseWhenExpr*>(e);
                    const auto x = outType(c->then_expr.get()), y = outType(c->else_expr.get());
                    if (x == LogicalType::Float64 || y == LogicalType::Float64) return LogicalType::Float64;
                    if (x == LogicalType::Int64 || y == LogicalType::Int64) return LogicalType::Int64;
                    return LogicalType::UInt64;
                }
                default: return LogicalType::Int64;
            }
        };
        std::function<bool(const ExprNode*)> outNullable = [&](const ExprNode* e) -> bool {
            if (!e) return true;
            switch (e->getType()) {
                case ExprType::LITERAL_NULL: return true;
                case ExprType::STAR: case ExprType::LITERAL_INT: case ExprType::LITERAL_FLOAT:
                case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
                    return false;
                case ExprType::COLUMN_REF: return nullable(static_cast<const ColumnRefExpr*>(e)->column_name);
                case ExprType::OP_NOT: return outNullable(static_cast<const BinaryExpr*>(e)->left.get());
                case ExprType::CASE_WHEN: {
                    const auto* c = static_cast<const CaseWhenExpr*>(e);
                    return outNullable(c->then_expr.get()) || outNullable(c->else_expr.get());
                }
                default: {
                    const auto* b = static_cast<const BinaryExpr*>(e);
                    return outNullable(b->left.get()) || outNullable(b->right.get());
                }
            }
        };
        for (int g = 0; g < prog_.ngroups; ++g) {
            OutputColumn oc;
            oc.type = outType(a->group_by_exprs[static_cast<size_t>(g)].get());
            oc.nullable = outNullable(a->group_by_exprs[static_cast<size_t>(g)].get());
            prog_.out.push_back(oc);
        }
        for (int i = 0; i < prog_.naggs; ++i) {
            const auto& d = a->aggregates[static_cast<size_t>(i)];
            OutputColumn oc;
            oc.type = d.isAvg() ? LogicalType::Float64 : LogicalType::UInt64;
            oc.nullable = d.isCount() ? false : outNullable(d.agg_expr.get());
            prog_.out.push_back(oc);
        }
        // names: the alias, else the reference's default (a group column's
        // name, an aggregate's function; execution.cpp:540-579)
        for (size_t c = 0; c < prog_.out.size(); ++c) {
            if (c < a->output_aliases.size()) prog_.out[c].name = a->output_aliases[c];
            if (!prog_.out[c].name.empty()) continue;
            const size_t ng = a->group_by_exprs.size();
            prog_.out[c].name = c < ng ? defaultColumnName(a->group_by_exprs[c].get()) : a->aggregates[c - ng].func_name;
        }
        // sparse output eligibility
        bool sparse = !prog_.has_avg && !prog_.has_having;
        for (const auto& oc : prog_.out) sparse = sparse && !oc.nullable;
        for (const auto& f : prog_.agg_funcs) sparse = sparse && (f == "COUNT" || f == "SUM");
        if (sparse && prog_.grouped) {
            bool nonzero = false;
            for (const auto& g : a->group_by_exprs) {
                if (g->getType() == ExprType::COLUMN_REF) {
                    const auto* s = stats(static_cast<const ColumnRefExpr*>(g.get())->column_name);
                    if (s && s->min_value_ != 0) nonzero = true;
                }
            }
            for (const auto& f : prog_.agg_funcs) nonzero = nonzero || f == "COUNT";
            sparse = nonzero;
        }
        prog_.has_sort_node = sort_ != nullptr;
        prog_.fast_sparse = sparse && !prog_.has_sort_node && !prog_.has_having;
        if (sort_) {
            prog_.sort_keys = sort_->sort_keys;
            prog_.has_limit = sort_->has_limit;
            prog_.limit = sort_->limit;
            for (const auto& k : prog_.sort_keys)
                if (k.column_index >= static_cast<size_t>(prog_.visible_ts))
                    unsupported("ORDER BY key index exceeds generated result width");
        }
        // perfect hash domain
        prog_.groups = 1;
        for (const auto& gexpr : a->group_by_exprs) {
            int64_t mn = 0;
            uint64_t card = 1;
            bool nul = false, dense = false;
            if (gexpr->getType() == ExprType::COLUMN_REF) {
                const auto& col = static_cast<const ColumnRefExpr*>(gexpr.get())->column_name;
                const auto* s = stats(col);
                // nullable from the catalog alone (visitor.cpp:2077): a
                // dimension column arriving in a payload register (valid
                // expression 1) still takes the +1 term and card + 1
                // (visitor.cpp:2105-2108), so its code 0 is never used
                nul = nullable(col);
                if (s) {
                    mn = s->min_value_;
                    card = s->cardinality_;
                    dense = s->max_value_ - s->min_value_ + 1 == static_cast<int64_t>(s->cardinality_);
                }
            } else if (gexpr->getType() == ExprType::CASE_WHEN) {
                std::vector<int64_t> vals;
                bool may_null = false;
                collectCaseIntegerOutputs(gexpr.get(), vals, may_null);
                if (vals.empty() && !may_null) unsupported("GROUP BY CASE has no integer output values");
                if (!vals.empty()) {
                    const auto mm = std::minmax_element(vals.begin(), vals.end());
                    mn = *mm.first;
                    card = static_cast<uint64_t>(*mm.second - *mm.first + 1);
                }
                nul = true;
                dense = true;
            } else {
                unsupported(std::string("GROUP BY expression is not supported by perfect hash codegen: ") +
                            exprTypeName(gexpr->getType()));
            }
            if (nul) card += 1;
            if (card == 0) card = 1;
            gmin_.push_back(mn);
            gcard_.push_back(card);
            gnullable_.push_back(nul);
            gdense_.push_back(dense);
            prog_.groups *= static_cast<int64_t>(card);
            if (prog_.groups > (int64_t(1) << 31)) unsupported("GROUP BY hash domain exceeds int32");
        }
    }

    // ---------------------------------------------------------------- literals
    std::string lit(int64_t v) {  // C type of std::to_string(v) as a literal
        const int s = static_cast<int>(prog_.lits.size());
        if (s >= CQ_MAX_LITS_HOST) unsupported("too many literals in one query");
        prog_.lits.push_back(v);
        if (v >= -2147483647LL && v <= 2147483647LL) return "((int)P.lit[" + std::to_string(s) + "])";
        return "P.lit[" + std::to_string(s) + "]";
    }
    std::string litInt32(int64_t v) {  // literal converted to an int compare value
        const int s = static_cast<int>(prog_.lits.size());
        if (s >= CQ_MAX_LITS_HOST) unsupported("too many literals in one query");
        prog_.lits.push_back(v);
        return "((int)P.lit[" + std::to_string(s) + "])";
    }
    std::string flit(double v) {
        const int s = static_cast<int>(prog_.flits.size());
        if (s >= CQ_MAX_FLITS_HOST) unsupported("too many floating-point literals in one query");
        prog_.flits.push_back(printedDouble(v));
        return "P.flit[" + std::to_string(s) + "]";
    }
    static constexpr int CQ_MAX_LITS_HOST = 48;
    static constexpr int CQ_MAX_FLITS_HOST = 8;

    // ---------------------------------------------------------------- columns in the scan
    int factSlot(const std::string& col) {
        auto it = fact_slot_.find(col);
        if (it != fact_slot_.end()) return it->second;
        const int s = static_cast<int>(prog_.fact_cols.size());
        if (s >= 16) unsupported("too many fact columns in one query");
        prog_.fact_cols.push_back(col);
        prog_.fact_nullable.push_back(nullable(col));
        fact_slot_[col] = s;
        return s;
    }
    int dimSlot(const std::string& col) {
        auto it = dim_slot_.find(col);
        if (it != dim_slot_.end()) return it->second;
        const int s = static_cast<int>(prog_.dim_cols.size());
        if (s >= 24) unsupported("too many dimension columns in one query");
        prog_.dim_cols.push_back(col);
        prog_.dim_nullable.push_back(nullable(col));
        dim_slot_[col] = s;
        return s;
    }
    int payloadJoin(const std::string& col) const {
        for (int j = 0; j < static_cast<int>(prog_.joins.size()); ++j)
            if (((visible_mask_ >> j) & 1u) && prog_.joins[static_cast<size_t>(j)].payload_col == col) return j;
        return -1;
    }
    int rowidJoin(const std::string& col) const {
        for (size_t j = 0; j < prog_.joins.size(); ++j)
            if (prog_.joins[j].rowid && prog_.joins[j].dim_table == tableOf(col)) return static_cast<int>(j);
        return -1;
    }
    std::string colValue(const std::string& col) {
        if (tableOf(col) == prog_.fact_table) return "c" + std::to_string(factSlot(col)) + "[i]";
        if (const int rj = rowidJoin(col); rj >= 0)
            return "P.dcol[" + std::to_string(dimSlot(col)) + "][p" + std::to_string(rj) + "[i]]";
        const int j = payloadJoin(col);
        if (j >= 0) return "p" + std::to_string(j) + "[i]";
        unsupported("column '" + col + "' of dimension " + tableOf(col) +
                    " is not available in the probe pipeline (only one payload column per joined dimension)");
    }
    std::string colValid(const std::string& col) {
        if (tableOf(col) == prog_.fact_table) {
            const int s = factSlot(col);
            return prog_.fact_nullable[static_cast<size_t>(s)] ? "((v" + std::to_string(s) + " >> i) & 1u)" : "1";
        }
        if (const int rj = rowidJoin(col); rj >= 0)  // gathered by row id with its validity
            return nullable(col) ? "chip::bit32(P.dvalid[" + std::to_string(dimSlot(col)) + "], p" + std::to_string(rj) + "[i])" : "1";
        if (payloadJoin(col) >= 0) return "1";  // payload registers carry no validity
        colValue(col);                           // raises the diagnostic
        return "1";
    }

    // expressionValueNoEmit / translateInlineExpr (visitor.cpp:837-920, 1110-1216)
    std::string vexpr(const ExprNode* e) {
        switch (e->getType()) {
            case ExprType::COLUMN_REF: return colValue(static_cast<const ColumnRefExpr*>(e)->column_name);
            case ExprType::LITERAL_INT: return lit(static_cast<const LiteralIntExpr*>(e)->value);
            case ExprType::LITERAL_FLOAT: return flit(static_cast<const LiteralFloatExpr*>(e)->value);
            case ExprType::LITERAL_NULL: return "0";
            case ExprType::STAR: return "1";
            case ExprType::OP_IS_NULL:
            case ExprType::OP_IS_NOT_NULL: {
                const std::string v = vvalid(static_cast<const BinaryExpr*>(e)->left.get());
                const bool isnull = e->getType() == ExprType::OP_IS_NULL;
                if (litTrue(v)) return isnull ? "0" : "1";
                if (litFalse(v)) return isnull ? "1" : "0";
                return isnull ? "(!(" + v + "))" : "(" + v + ")";
            }
            case ExprType::OP_NOT: return boolNot(vexpr(static_cast<const BinaryExpr*>(e)->left.get()));
            case ExprType::OP_AND: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = vexpr(b->left.get());
                return boolAnd(l, vexpr(b->right.get()));
            }
            case ExprType::OP_OR: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = vexpr(b->left.get());
                return boolOr(l, vexpr(b->right.get()));
            }
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                const std::string choose = andTerms({vvalid(c->condition.get()), vexpr(c->condition.get())});
                const std::string t = vexpr(c->then_expr.get());
                const std::string f = vexpr(c->else_expr.get());
                return "((" + choose + ") ? (" + t + ") : (" + f + "))";
            }
            default: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = vexpr(b->left.get());
                const std::string r = vexpr(b->right.get());
                return std::string(safeFn(b->op_type)) + "(" + l + ", " + r + ")";
            }
        }
    }

    // expressionValidExpr (visitor.cpp:921-991)
    std::string vvalid(const ExprNode* e) {
        switch (e->getType()) {
            case ExprType::LITERAL_NULL: return "0";
            case ExprType::STAR: case ExprType::LITERAL_INT: case ExprType::LITERAL_FLOAT:
            case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
                return "1";
            case ExprType::COLUMN_REF: return colValid(static_cast<const ColumnRefExpr*>(e)->column_name);
            case ExprType::OP_NOT: return vvalid(static_cast<const BinaryExpr*>(e)->left.get());
            case ExprType::OP_AND:
            case ExprType::OP_OR: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string lv = vvalid(b->left.get()), rv = vvalid(b->right.get());
                const std::string l = vexpr(b->left.get()), r = vexpr(b->right.get());
                if (e->getType() == ExprType::OP_AND)
                    return boolOr(boolOr(andTerms({lv, boolNot(l)}), andTerms({rv, boolNot(r)})), andTerms({lv, rv}));
                return boolOr(boolOr(andTerms({lv, l}), andTerms({rv, r})), andTerms({lv, rv}));
            }
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                const std::string sel = andTerms({vvalid(c->condition.get()), vexpr(c->condition.get())});
                return boolOr(andTerms({sel, vvalid(c->then_expr.get())}), andTerms({boolNot(sel), vvalid(c->else_expr.get())}));
            }
            default: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                return andTerms({vvalid(b->left.get()), vvalid(b->right.get())});
            }
        }
    }

    // translateMathExprPush (visitor.cpp:1947-2018): leftmost leaf cast to u64
    std::string mexpr(const ExprNode* e, bool cast) {
        switch (e->getType()) {
            case ExprType::COLUMN_REF: {
                const std::string r = colValue(static_cast<const ColumnRefExpr*>(e)->column_name);
                return cast ? "(unsigned long long)" + r : r;
            }
            case ExprType::LITERAL_INT: return lit(static_cast<const LiteralIntExpr*>(e)->value);
            case ExprType::LITERAL_FLOAT: return flit(static_cast<const LiteralFloatExpr*>(e)->value);
            case ExprType::STAR: return cast ? "(unsigned long long)1" : "1";
            case ExprType::OP_ADD: case ExprType::OP_SUB: case ExprType::OP_MUL: case ExprType::OP_DIV: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = mexpr(b->left.get(), cast);
                const std::string r = mexpr(b->right.get(), false);
                switch (e->getType()) {
                    case ExprType::OP_ADD: return "(" + l + " + " + r + ")";
                    case ExprType::OP_SUB: return "(" + l + " - " + r + ")";
                    case ExprType::OP_MUL: return "(" + l + " * " + r + ")";
                    default: return "chip::cdiv(" + l + ", " + r + ")";
                }
            }
            case ExprType::OP_IS_NULL:
            case ExprType::OP_IS_NOT_NULL: {
                const std::string v = vexpr(e);
                return cast ? "(unsigned long long)" + v : v;
            }
            case ExprType::LITERAL_NULL: return "0";
            default: {
                const std::string v = vexpr(e);
                return cast ? "(unsigned long long)(" + v + ")" : v;
            }
        }
    }

    // Row flags of the items of a lane: a u32 bit mask `fl` (plain kernels)
    // or, in shared batched templates, `bool fl[IPT]` (wave-wide lane masks
    // in scalar registers, so flag logic runs on the scalar unit)
    bool bool_flags_ = false;
    std::string flGet() const { return bool_flags_ ? "(u32)fl[i]" : "((fl >> i) & 1u)"; }
    std::string flClear() const { return bool_flags_ ? "fl[i] = false;" : "fl &= ~(1u << i);"; }
    std::string flAny() const { return bool_flags_ ? "chip::any_of(fl)" : "fl != 0u"; }
    // keep the row only if the 0/1 int expression `c` is 1
    std::string flKeep(const std::string& c) const {
        return bool_flags_ ? "fl[i] = fl[i] & ((" + c + ") != 0);" : "fl &= ~((u32)((" + c + ") ^ 1) << i);";
    }

    // fast path: AND chain of (int)col op (int)literal, raw values
    // (bitwise: 0/1 terms combined with &, no short-circuit branches)
    std::string fastCond(const ExprNode* e, bool bitwise = false) {
        if (e->getType() == ExprType::OP_AND) {
            const auto* b = static_cast<const BinaryExpr*>(e);
            const std::string l = fastCond(b->left.get(), bitwise);
            return l + (bitwise ? " & " : " && ") + fastCond(b->right.get(), bitwise);
        }
        const auto* c = static_cast<const BinaryExpr*>(e);
        const auto& col = static_cast<const ColumnRefExpr*>(c->left.get())->column_name;
        if (tableOf(col) != prog_.fact_table)
            unsupported("fast-path filter on dimension column " + col + " above its join");
        const std::string v = "c" + std::to_string(factSlot(col)) + "[i]";
        const std::string t = v + " " + cmpOp(c->op_type) + " " + litInt32(static_cast<const LiteralIntExpr*>(c->right.get())->value);
        return bitwise ? "((int)(" + t + "))" : "(" + t + ")";
    }

    // build-side predicate on raw dimension values
    std::string buildCond(const ExprNode* e) {
        return buildCondWith(e, [&](const std::string& name) { return "P.dcol[" + std::to_string(dimSlot(name)) + "][r]"; });
    }
    // the same predicate over another source of the raw dimension values
    // bitwise: 0/1 terms combined with & and | (no short-circuit branches)
    std::string buildCondWith(const ExprNode* e, const std::function<std::string(const std::string&)>& value,
                              bool bitwise = false) {
        const ExprType t = e->getType();
        if (t == ExprType::OP_AND || t == ExprType::OP_OR) {
            const auto* b = static_cast<const BinaryExpr*>(e);
            const std::string l = buildCondWith(b->left.get(), value, bitwise);
            const char* op = bitwise ? (t == ExprType::OP_AND ? " & " : " | ") : (t == ExprType::OP_AND ? " && " : " || ");
            return "(" + l + op + buildCondWith(b->right.get(), value, bitwise) + ")";
        }
        const auto* c = static_cast<const BinaryExpr*>(e);
        const bool col_left = c->left->getType() == ExprType::COLUMN_REF;
        const ExprNode* col = col_left ? c->left.get() : c->right.get();
        const ExprNode* l = col_left ? c->right.get() : c->left.get();
        const ExprType op = col_left ? c->op_type : swapCmp(c->op_type);
        const auto& name = static_cast<const ColumnRefExpr*>(col)->column_name;
        std::string litcode;
        if (l->getType() == ExprType::LITERAL_INT) litcode = litInt32(static_cast<const LiteralIntExpr*>(l)->value);
        else litcode = "((int)" + flit(static_cast<const LiteralFloatExpr*>(l)->value) + ")";
        if (bitwise) return "((int)(" + value(name) + " " + cmpOp(op) + " " + litcode + "))";
        return "(" + value(name) + " " + cmpOp(op) + " " + litcode + ")";
    }

    // ---------------------------------------------------------------- emission
    // Plain modules take the parameter block by value (kernarg segment);
    // modules of the batched path read it from the device array of blocks so
    // a captured graph stays valid when only constants change.
    std::string paramDecl() const {
        return opt_.batch_body ? "const cq_query_params* __restrict__ PP" : "const cq_query_params P";
    }
    std::string paramAlias() const { return opt_.batch_body ? "    const cq_query_params& P = *PP;\n" : ""; }

    std::string groupHash() {
        const auto* a = agg_;
        std::string h;
        for (size_t g = 0; g < a->group_by_exprs.size(); ++g) {
            const ExprNode* e = a->group_by_exprs[g].get();
            const std::string v = vexpr(e);
            const std::string mn = gmin_[g] >= 0 ? std::to_string(gmin_[g])
                                                  : std::to_string(static_cast<uint64_t>(gmin_[g])) + "ULL";
            std::string term;
            if (gnullable_[g]) term = "((" + vvalid(e) + ") ? (" + v + " - " + mn + " + 1) : 0)";
            else term = "(" + v + " - " + mn + ")";
            h = h.empty() ? term : "(" + h + " * " + std::to_string(gcard_[g]) + " + " + term + ")";
        }
        // Every term of a dense domain lies in [0, card) (catalog min / max /
        // distinct count bound every value, the NULL code is 0), so the
        // Horner value lies in [0, groups) and the reference's `% groups`
        // is the identity: left out.
        bool dense = true;
        for (bool d : gdense_) dense = dense && d;
        if (dense) return "(int)(" + h + ")";
        return "(int)((" + h + ") % " + std::to_string(prog_.groups) + ")";
    }

    // Fraction of dimension rows a build filter keeps, from its structure
    // and catalog cardinalities only (never from the constants, so queries
    // that differ in constants keep one plan signature).
    double buildSelectivity(const ExprNode* e) const {
        if (!e) return 1.0;
        const ExprType t = e->getType();
        if (t == ExprType::OP_AND || t == ExprType::OP_OR) {
            const auto* b = static_cast<const BinaryExpr*>(e);
            const double l = buildSelectivity(b->left.get()), r = buildSelectivity(b->right.get());
            return t == ExprType::OP_AND ? l * r : std::min(1.0, l + r);
        }
        const auto* c = static_cast<const BinaryExpr*>(e);
        const ExprNode* col = c->left->getType() == ExprType::COLUMN_REF ? c->left.get() : c->right.get();
        const auto* st = stats(static_cast<const ColumnRefExpr*>(col)->column_name);
        const double card = st && st->cardinality_ > 0 ? static_cast<double>(st->cardinality_) : 10.0;
        switch (c->op_type) {
            case ExprType::OP_EQ: return 1.0 / card;
            case ExprType::OP_NEQ: return 1.0 - 1.0 / card;
            default: return 0.5;
        }
    }
    // Probe cost relative to an LDS probe: tables beyond the 32 KiB L1 go to L2.
    double probeCost(const JoinProgram& jp) const { return jp.lds ? 1.0 : (jp.ht_words * 4 <= 32 * 1024 ? 2.0 : 8.0); }

    // Execution order of the pipeline: every filter and probe only removes
    // rows, so the order does not change the result.  Filters on fact
    // columns first, then probes by rows removed per unit cost (the classic
    // predicate-ordering rank), then filters that read payload registers.
    std::vector<Op> pipelineOrder() const {
        std::vector<Op> first, joins, last;
        for (const Op& op : ops_) {
            if (op.kind == 1) {
                joins.push_back(op);
                continue;
            }
            std::vector<std::string> cols;
            collectColumns(op.pred, cols);
            bool fact_only = true;
            for (const auto& c : cols) fact_only = fact_only && tableOf(c) == prog_.fact_table;
            (fact_only ? first : last).push_back(op);
        }
        std::stable_sort(joins.begin(), joins.end(), [&](const Op& a, const Op& b) {
            const JoinProgram& ja = prog_.joins[static_cast<size_t>(a.join)];
            const JoinProgram& jb = prog_.joins[static_cast<size_t>(b.join)];
            const double ra = (1.0 - buildSelectivity(build_pred_[static_cast<size_t>(a.join)])) / probeCost(ja);
            const double rb = (1.0 - buildSelectivity(build_pred_[static_cast<size_t>(b.join)])) / probeCost(jb);
            if (ra != rb) return ra > rb;
            return probeCost(ja) < probeCost(jb);
        });
        std::vector<Op> out = first;
        out.insert(out.end(), joins.begin(), joins.end());
        out.insert(out.end(), last.begin(), last.end());
        return out;
    }

    // Multi-value table: cq_build_<j> counts the build rows per slot
    // (BlockBuildMHT_Count), the runtime scans the counts into first
    // positions, cq_mhtw_<j> places each row id and writes the slot's
    // directory entry (BlockBuildMHT_Write, join.h:301-364).  Build filter
    // and key validity as for the other tables (visitor.cpp:4295-4464).
    void emitMhtBuild(std::ostringstream& src, int j, const std::string& cond) {
        const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
        auto head = [&](const std::string& name) {
            src << "extern \"C\" __global__ __launch_bounds__(256) void " << name << "(" << paramDecl() << ") {\n"
                << paramAlias() << "    using namespace chip;\n"
                << "    const long long n = P.dim_rows[" << j << "];\n"
                << "    u32* const T = P.ht[" << j << "];\n"
                << "    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long long)gridDim.x * 256) {\n";
            if (!cond.empty()) src << "        if (!" << cond << ") continue;\n";
            if (nullable(jp.pk_col)) src << "        if (!chip::bit32(P.dvalid[" << jp.pk_dslot << "], r)) continue;\n";
            src << "        const int k = P.dcol[" << jp.pk_dslot << "][r];\n"
                << "        long long s;\n"
                << "        if (!chip::mht_slot(k, " << jp.kmin << ", " << jp.len << "LL, s)) continue;\n";
        };
        head("cq_build_" + std::to_string(j));
        src << "        atomicAdd((u64*)(T + " << jp.mht_cnt << ") + s, 1ull);\n    }\n}\n\n";
        head("cq_mhtw_" + std::to_string(j));
        src << "        const u32 first = (u32)((const u64*)(T + " << jp.mht_off << "))[s];\n"
            << "        const u32 count = (u32)((const u64*)(T + " << jp.mht_cnt << "))[s];\n"
            << "        const u32 at = atomicAdd(T + " << jp.mht_pos << " + s, 1u);\n"
            << "        T[" << jp.mht_rows << " + first + at] = (u32)r;\n"
            << "        v4u d; d.x = (u32)k; d.y = first; d.z = count; d.w = 0u;\n"
            << "        *(v4u*)(T + 4 * s) = d;\n    }\n}\n\n";
    }

    void emitBuild(std::ostringstream& src, int j) {
        JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
        const ExprNode* bp = build_pred_[static_cast<size_t>(j)];
        std::string cond = bp ? buildCond(bp) : "";
        jp.pk_dslot = dimSlot(jp.pk_col);
        if (jp.mht) {
            emitMhtBuild(src, j, cond);
            return;
        }
        if (jp.pair && !jp.rowid) jp.payload_dslot = dimSlot(jp.payload_col);
        src << "extern \"C\" __global__ __launch_bounds__(256) void cq_build_" << j << "(" << paramDecl() << ") {\n"
            << paramAlias()
            << "    const long long n = P.dim_rows[" << j << "];\n"
            << "    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long long)gridDim.x * 256) {\n";
        if (!cond.empty()) src << "        if (!" << cond << ") continue;\n";
        if (nullable(jp.pk_col))
            src << "        if (!chip::bit32(P.dvalid[" << jp.pk_dslot << "], r)) continue;\n";
        src << "        const int k = P.dcol[" << jp.pk_dslot << "][r];\n";
        src << "        const int pay = "
            << (jp.rowid ? std::string("(int)r") : jp.pair ? "P.dcol[" + std::to_string(jp.payload_dslot) + "][r]" : "0") << ";\n";
        if (jp.dense)
            src << "        chip::ht_insert_dense(P.ht[" << j << "], k, " << jp.dmin << ", " << jp.dspan << ", " << jp.db << ", "
                << densePbase(jp) << ", pay, " << jp.cal << ");\n";
        else if (jp.compact)
            src << "        chip::ht_insert_compact(P.ht[" << j << "], k, " << jp.kmin << ", " << jp.len << ", " << jp.eb << ", "
                << jp.tb << ", " << jp.jlo << ", " << jp.pmin << ", pay);\n";
        else
            src << "        chip::ht_insert_ref((int*)P.ht[" << j << "], k, " << jp.kmin << ", " << jp.len << ", "
                << (jp.pair ? "true" : "false") << ", pay);\n";
        src << "    }\n}\n\n";
    }

    void emit() {
        std::ostringstream src;
        src << "// generated by crystal-hip for plan " << "\n";
        src << "#include \"crystal_hip.h\"\n\n";

        const bool grouped = prog_.grouped;
        // LDS aggregation when the domain decodes from the group index
        const bool batch = opt_.batch_body;
        computeFlagLayout();
        if (opt_.vm) {
            emitVm();
            return;
        }
        if (batch && sharedEligible()) {
            emitBatchShared(src);
            return;
        }
        bool lds_agg = grouped && opt_.allow_lds_aggregation && !batch && flag_w_ > 0 && prog_.storage_ts <= 64;
        for (size_t g = 0; g < gdense_.size(); ++g) lds_agg = lds_agg && gdense_[g];
        // scalar aggregates: per-wave partials of the block reduction (16
        // waves at most, 8 + 4 bytes per aggregate and hidden count)
        const uint64_t agg_lds = grouped ? aggLdsBytes() : static_cast<uint64_t>(16 * 12 * 2 * std::max(1, prog_.naggs));
        if (agg_lds > opt_.lds_budget) lds_agg = false;
        prog_.lds_aggregation = lds_agg;
        uint64_t lds_used = lds_agg ? agg_lds : 0;
        const uint64_t red_lds = grouped ? 0 : agg_lds;  // outside the launch-shape choice below
        lds_used += red_lds;
        // LDS placement of the dimension tables: the subset that avoids the
        // most HBM-side gathers within the budget (tables beyond the 32 KiB
        // L1 weigh more: their probes go to L2), then the final layouts
        const int nj = static_cast<int>(prog_.joins.size());
        int best_mask = 0;
        if (!batch && opt_.allow_lds_tables) {
            int64_t best_w = 0;
            uint64_t best_b = 0;
            for (int m = 1; m < (1 << nj); ++m) {
                uint64_t b = 0;
                int64_t w = 0;
                bool ok = true;
                for (int j = 0; j < nj && ok; ++j) {
                    if (!((m >> j) & 1)) continue;
                    const uint64_t lb = ldsBytes(prog_.joins[static_cast<size_t>(j)]);
                    if (lb == 0) ok = false;
                    b += lb;
                    w += lb > 32 * 1024 ? 8 : 1;
                }
                if (!ok || lds_used + b > opt_.lds_budget) continue;
                if (w > best_w || (w == best_w && b < best_b)) {
                    best_w = w;
                    best_b = b;
                    best_mask = m;
                }
            }
        }
        for (int j = 0; j < nj; ++j) {
            auto& jp = prog_.joins[static_cast<size_t>(j)];
            finalizeLayout(jp, (best_mask >> j) & 1);
            if (jp.lds) lds_used += jp.ht_words * 4;
        }
        // builds follow the final layouts (the probes below use the same)
        for (int j = 0; j < nj; ++j) emitBuild(src, j);
        prog_.scan_lds_bytes = lds_used;
        int bt = opt_.block_threads;
        if (bt <= 0) bt = lds_used - red_lds > 40 * 1024 ? 1024 : (lds_used - red_lds > 0 ? 512 : 256);
        prog_.block_threads = bt;
        const int vpt = std::max(1, std::min(8, opt_.vectors_per_thread));
        prog_.vectors_per_thread = vpt;
        const int ipt = 4 * vpt;

        // ---- pipeline body first: it decides which fact columns are loaded
        std::ostringstream body, orbody;
        visible_mask_ = 0;
        // an OR-key or multi-value join runs last, with the filters that
        // read its payload
        int multi = -1, mhtj = -1;
        for (size_t j = 0; j < prog_.joins.size(); ++j) {
            if (prog_.joins[j].fk_alts.size() > 1) multi = static_cast<int>(j);
            if (prog_.joins[j].mht) mhtj = static_cast<int>(j);
        }
        const int last = multi >= 0 ? multi : mhtj;
        std::vector<Op> dependent;
        const std::vector<Op> order = pipelineOrder();
        // Key-indexed HBM probes whose earlier probes remove no rows (no
        // build filter): their loads are issued with the first probe, so the
        // L2 round trip overlaps the LDS probes in between.  Rows are only
        // ever removed, so loading for rows an LDS probe drops later reads
        // entries nobody uses and changes no result.
        uint32_t hoisted = 0;
        if (opt_.hoist_gathers && !opt_.batch_body && !opt_.vm) {
            bool ahead_keeps_all = true;
            size_t probes_before = 0;
            for (const Op& op : order) {
                if (op.kind != 1 || op.join == last) continue;
                const JoinProgram& jp = prog_.joins[static_cast<size_t>(op.join)];
                if (jp.dense && !jp.lds && useRsrc(jp) && ahead_keeps_all && probes_before > 0)
                    hoisted |= 1u << op.join;
                ahead_keeps_all = ahead_keeps_all && build_pred_[static_cast<size_t>(op.join)] == nullptr;
                ++probes_before;
            }
        }
        auto probeArgs = [&](const JoinProgram& jp, std::string& key, std::string& kv) {
            const int fs = factSlot(jp.fk_col);
            key = "c" + std::to_string(fs) + "[i]";
            kv = prog_.fact_nullable[static_cast<size_t>(fs)] ? "((v" + std::to_string(fs) + " >> i) & 1u)" : "";
        };
        bool issued = false;
        for (const Op& op : order) {
            if (op.kind == 1 && hoisted && !issued) {
                for (size_t j = 0; j < prog_.joins.size(); ++j) {
                    if (!((hoisted >> j) & 1u)) continue;
                    std::string key, kv;
                    probeArgs(prog_.joins[j], key, kv);
                    emitProbe(body, static_cast<int>(j), key, kv, 1);
                }
                issued = true;
            }
            if (last >= 0 && op.kind == 1 && op.join == last) continue;
            if (last >= 0 && op.kind == 0) {
                std::vector<std::string> cols;
                collectColumns(op.pred, cols);
                bool dep = false;
                for (const auto& c : cols) dep = dep || tableOf(c) == prog_.joins[static_cast<size_t>(last)].dim_table;
                if (dep) {
                    dependent.push_back(op);
                    continue;
                }
            }
            if (op.kind == 0) {
                emitFilterOp(body, op);
            } else {
                std::string key, kv;
                probeArgs(prog_.joins[static_cast<size_t>(op.join)], key, kv);
                emitProbe(body, op.join, key, kv, ((hoisted >> op.join) & 1u) ? 2 : 0);
            }
        }
        if (multi >= 0) {
            // OR-key join: the probe, the filters that read its payload and
            // the aggregation once per key alternative (a row's later
            // alternative is skipped where it equals a valid earlier key, so
            // each matching dimension row counts once; visitor.cpp:4581-4638)
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(multi)];
            const size_t na = jp.fk_alts.size();
            std::vector<int> slots;
            for (const auto& fk : jp.fk_alts) slots.push_back(factSlot(fk));
            auto key = [&](size_t k) { return "c" + std::to_string(slots[k]) + "[i]"; };
            auto keyValid = [&](size_t k) {
                return prog_.fact_nullable[static_cast<size_t>(slots[k])] ? "((v" + std::to_string(slots[k]) + " >> i) & 1u)"
                                                                          : std::string("1u");
            };
            orbody << "        // " << jp.dim_table << " joined on";
            for (const auto& fk : jp.fk_alts) orbody << " " << fk;
            orbody << ": the rest of the pipeline once per key alternative\n"
                   << "        int kor[IPT];\n"
                   << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                   << "            u32 ok = 1u;\n";
            for (size_t k = 0; k < na; ++k)
                orbody << "            " << (k ? "else if" : "if") << " (alt == " << k << ") { kor[i] = " << key(k) << "; ok = " << keyValid(k)
                       << "; }\n";
            for (size_t k = 0; k + 1 < na; ++k)
                orbody << "            if (alt > " << k << " && " << keyValid(k) << " && " << key(k) << " == kor[i]) ok = 0u;\n";
            orbody << "            if (!ok) " << flClear() << "\n"
                   << "        }\n";
            emitProbe(orbody, multi, "kor[i]", "");
            for (const Op& op : dependent) emitFilterOp(orbody, op);
        }
        // multi-value join: directory lookup per item (mlocate), then per
        // match index m the build row ids and the filters that read them
        // (mbody, inside `for (m ...)`, where `continue` moves to the next m)
        std::ostringstream mlocate, mbody;
        if (mhtj >= 0) {
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(mhtj)];
            const std::string J = std::to_string(mhtj);
            const int fs = factSlot(jp.fk_col);
            std::string live = flGet();
            if (prog_.fact_nullable[static_cast<size_t>(fs)]) live = "(" + live + " && ((v" + std::to_string(fs) + " >> i) & 1u))";
            mlocate << "        // " << jp.dim_table << " joined on non-unique " << jp.pk_col
                    << ": the rest of the pipeline once per matching build row\n"
                    << "        u32 mn[IPT], mf[IPT], mmax = 0u;\n"
                    << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                    << "            mn[i] = 0u; mf[i] = 0u;\n"
                    << "            if (" << live << ") mn[i] = chip::mht_find(P.ht[" << J << "], c" << fs << "[i], " << jp.kmin << ", "
                    << jp.len << "LL, mf[i]);\n"
                    << "            mmax = mn[i] > mmax ? mn[i] : mmax;\n"
                    << "        }\n";
            mbody << "            fl = 0u;\n"
                  << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) if (m < mn[i]) fl |= 1u << i;\n"
                  << "            int p" << J << "[IPT];\n"
                  << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) p" << J << "[i] = ((fl >> i) & 1u) ? (int)P.ht[" << J
                  << "][" << jp.mht_rows << " + mf[i] + m] : 0;\n";
            visible_mask_ |= 1u << mhtj;
            for (const Op& op : dependent) emitFilterOp(mbody, op);
        }
        if (prog_.projection) {
            if (mhtj >= 0) {
                prog_.rows_unordered = true;
                emitProjectionExpanded(src, body.str(), mlocate.str(), mbody.str(), bt, vpt);
                prog_.source = src.str();
                setAlgorithmicBytes();
                return;
            }
            emitProjection(src, body.str(), bt, vpt, ipt);
            prog_.source = src.str();
            setAlgorithmicBytes();
            return;
        }

        // ---- aggregation
        std::ostringstream aggcode, epilogue, prologue;
        const auto* a = agg_;
        if (!grouped) emitScalarAggregate(aggcode, prologue, epilogue);
        else if (lds_agg) emitLdsAggregate(aggcode, prologue, epilogue);
        else emitGlobalAggregate(aggcode);
        if (batch) {
            std::ostringstream tile;
            if (!grouped) emitScalarAggregateTile(tile);
            else tile << aggcode.str();
            prog_.batch_body = body.str() + tile.str();
        }
        for (int i = 0; i < prog_.naggs && !grouped; ++i)
            if (prog_.agg_funcs[static_cast<size_t>(i)] == "COUNT") prog_.count_slots_scalar.push_back(i);

        if (prog_.keys_kernel) emitKeysKernel(src);
        // ---- kernel (the batched path runs the body inside cq_batch_scan)
        if (batch) {
            emitDenseCount(src);
            prog_.source = src.str();
            setAlgorithmicBytes();
            return;
        }
        src << "#define BT " << bt << "\n#define VPT " << vpt << "\n#define IPT " << ipt << "\n#define TILE (BT * IPT)\n\n";
        src << "extern \"C\" __global__ __launch_bounds__(BT) void cq_scan(const cq_query_params P) {\n";
        src << "    using namespace chip;\n";
        src << "    const int tid = threadIdx.x;\n";
        for (const auto& jp : prog_.joins) {
            if (!jp.lds) continue;
            const int j = static_cast<int>(&jp - prog_.joins.data());
            src << "    __shared__ u32 s_ht" << j << "[" << jp.ht_words << "];\n"
                << "    for (int k = tid; k < " << jp.ht_words << "; k += BT) s_ht" << j << "[k] = P.ht[" << j << "][k];\n";
        }
        emitTableRsrc(src);
        src << prologue.str();
        src << "    __syncthreads();\n";
        const size_t nf = prog_.fact_cols.size();
        if (!opt_.prefetch || nf * static_cast<size_t>(ipt) > 48) {
            src << "    for (long long tile = blockIdx.x; tile < P.num_tiles; tile += gridDim.x) {\n";
            src << "        const long long base = P.row_begin + tile * (long long)TILE;\n";
            for (size_t s = 0; s < nf; ++s) {
                src << "        int c" << s << "[IPT];\n";
                if (prog_.fact_nullable[s]) src << "        u32 v" << s << " = 0u;\n";
            }
            src << "        u32 fl = 0u;\n";
            emitTileLoad(src, "        ", "", "base", "fl");
        } else {
            // software pipeline: the next tile's columns are in flight while
            // this tile's probes and aggregation run
            for (size_t s = 0; s < nf; ++s) {
                src << "    int nc" << s << "[IPT];\n";
                if (prog_.fact_nullable[s]) src << "    u32 nv" << s << " = 0u;\n";
            }
            src << "    u32 nfl = 0u;\n"
                << "    if ((long long)blockIdx.x < P.num_tiles) {\n"
                << "        const long long nbase = P.row_begin + (long long)blockIdx.x * (long long)TILE;\n";
            emitTileLoad(src, "        ", "n", "nbase", "nfl");
            src << "    }\n";
            src << "    for (long long tile = blockIdx.x; tile < P.num_tiles; tile += gridDim.x) {\n";
            for (size_t s = 0; s < nf; ++s) {
                src << "        int c" << s << "[IPT];\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) c" << s << "[i] = nc" << s
                    << "[i];\n";
                if (prog_.fact_nullable[s]) src << "        const u32 v" << s << " = nv" << s << "; nv" << s << " = 0u;\n";
            }
            src << "        u32 fl = nfl; nfl = 0u;\n"
                << "        if (tile + gridDim.x < P.num_tiles) {\n"
                << "            const long long nbase = P.row_begin + (tile + gridDim.x) * (long long)TILE;\n";
            emitTileLoad(src, "            ", "n", "nbase", "nfl");
            src << "        }\n";
        }
        src << body.str();
        if (multi >= 0) {
            src << "        const u32 fl_or = fl;\n"
                << "#pragma unroll 1\n"
                << "        for (int alt = 0; alt < " << prog_.joins[static_cast<size_t>(multi)].fk_alts.size() << "; ++alt) {\n"
                << "        fl = fl_or;\n"
                << orbody.str() << aggcode.str() << "        }\n";
        } else if (mhtj >= 0) {
            src << mlocate.str()
                << "#pragma unroll 1\n"
                << "        for (u32 m = 0u; __ballot(m < mmax) != 0ull; ++m) {\n"
                << mbody.str() << aggcode.str() << "        }\n";
        } else {
            src << aggcode.str();
        }
        src << "    }\n";
        src << epilogue.str();
        src << "}\n\n";

        emitDenseCount(src);
        prog_.source = src.str();
        (void)a;
        setAlgorithmicBytes();
    }

    // Projection: cq_pcount stores the live rows of each (tile, vector, wave);
    // after an exclusive scan of those counts cq_pwrite writes every live row
    // at its offset, so the output is in scan (row) order.  No block barrier:
    // a wave that finds no live row in a tile leaves the tile early.
    void emitProjection(std::ostringstream& src, const std::string& body, int bt, int vpt, int ipt) {
        const int waves = bt / 64;
        prog_.proj_waves = waves;
        prog_.block_threads = bt;
        prog_.vectors_per_thread = vpt;
        (void)ipt;
        const int ts = prog_.visible_ts;
        // output expressions first: they decide the fact columns each tile loads
        std::vector<std::string> vals, oks;
        for (int c = 0; c < ts; ++c) {
            vals.push_back(vexpr(proj_exprs_[static_cast<size_t>(c)]));
            oks.push_back(vvalid(proj_exprs_[static_cast<size_t>(c)]));
        }
        src << "#define BT " << bt << "\n#define VPT " << vpt << "\n#define IPT " << 4 * vpt << "\n#define TILE (BT * IPT)\n#define WAVES "
            << waves << "\n\n";
        auto head = [&](const char* name) {
            src << "extern \"C\" __global__ __launch_bounds__(BT) void " << name << "(const cq_query_params P) {\n"
                << "    using namespace chip;\n"
                << "    const int tid = threadIdx.x;\n"
                << "    const int wave = tid >> 6, lane = tid & 63;\n";
            for (const auto& jp : prog_.joins) {
                if (!jp.lds) continue;
                const int j = static_cast<int>(&jp - prog_.joins.data());
                src << "    __shared__ u32 s_ht" << j << "[" << jp.ht_words << "];\n"
                    << "    for (int k = tid; k < " << jp.ht_words << "; k += BT) s_ht" << j << "[k] = P.ht[" << j << "][k];\n";
            }
            emitTableRsrc(src);
            src << "    __syncthreads();\n"
                << "    for (long long tile = blockIdx.x; tile < P.num_tiles; tile += gridDim.x) {\n"
                << "        const long long base = P.row_begin + tile * (long long)TILE;\n";
            for (size_t c = 0; c < prog_.fact_cols.size(); ++c) {
                src << "        int c" << c << "[IPT];\n";
                if (prog_.fact_nullable[c]) src << "        u32 v" << c << " = 0u;\n";
            }
            src << "        u32 fl = 0u;\n";
            emitTileLoad(src, "        ", "", "base", "fl");
            src << body;
        };
        head("cq_pcount");
        src << "#pragma unroll\n        for (int v = 0; v < VPT; ++v) {\n"
            << "            const u64 n = wave_sum((u64)__popc((fl >> (4 * v)) & 0xfu));\n"
            << "            if (lane == 0) P.tile_count[(tile * VPT + v) * WAVES + wave] = n;\n"
            << "        }\n    }\n    (void)lane;\n}\n\n";
        head("cq_pwrite");
        src << "#pragma unroll\n        for (int v = 0; v < VPT; ++v) {\n"
            << "            const u32 m4 = (fl >> (4 * v)) & 0xfu;\n"
            << "            const u32 cnt = __popc(m4);\n"
            << "            const u64 below = (1ull << lane) - 1ull;\n"
            << "            const u32 pre = __popcll(__ballot(cnt & 1u) & below) + 2u * __popcll(__ballot((cnt >> 1) & 1u) & below) +\n"
            << "                            4u * __popcll(__ballot((cnt >> 2) & 1u) & below);\n"
            << "            const u64 first = P.tile_off[(tile * VPT + v) * WAVES + wave] + pre;\n"
            << "#pragma unroll\n            for (int k = 0; k < 4; ++k) {\n"
            << "                const int i = 4 * v + k;\n"
            << "                if (!((m4 >> k) & 1u)) continue;\n"
            << "                const u64 row = first + (u64)__popc(m4 & ((1u << k) - 1u));\n";
        for (int c = 0; c < ts; ++c) {
            const std::string& v = vals[static_cast<size_t>(c)];
            std::string cast;
            switch (prog_.out[static_cast<size_t>(c)].type) {
                case LogicalType::Float64: cast = "(u64)__double_as_longlong((double)(" + v + "))"; break;
                case LogicalType::Int64: cast = "(u64)(long long)(" + v + ")"; break;
                default: cast = "(u64)(" + v + ")"; break;
            }
            src << "                P.result[row * " << ts << " + " << c << "] = " << cast << ";\n";
            const std::string& ok = oks[static_cast<size_t>(c)];
            if (litFalse(ok)) continue;
            const std::string set = "{ const u64 b = row * " + std::to_string(ts) + " + " + std::to_string(c) +
                                    "; atomicOr(&P.result_valid[b >> 6], 1ull << (b & 63)); }";
            if (litTrue(ok)) src << "                " << set << "\n";
            else src << "                if (" << ok << ") " << set << "\n";
        }
        src << "            }\n        }\n    }\n}\n\n";
    }

    // Projection through a multi-value join: as emitProjection, with each
    // live row expanded into one output row per matching build row.  The
    // count pass stores each wave's surviving (row, match) pairs per (tile,
    // vector); the write pass counts them again per item, ranks them in the
    // wave (exclusive lane scan) and writes them in (lane, item, match) order.
    void emitProjectionExpanded(std::ostringstream& src, const std::string& body, const std::string& mlocate,
                                const std::string& mbody, int bt, int vpt) {
        const int waves = bt / 64;
        prog_.proj_waves = waves;
        prog_.block_threads = bt;
        prog_.vectors_per_thread = vpt;
        const int ts = prog_.visible_ts;
        std::vector<std::string> vals, oks;
        for (int c = 0; c < ts; ++c) {
            vals.push_back(vexpr(proj_exprs_[static_cast<size_t>(c)]));
            oks.push_back(vvalid(proj_exprs_[static_cast<size_t>(c)]));
        }
        src << "#define BT " << bt << "\n#define VPT " << vpt << "\n#define IPT " << 4 * vpt << "\n#define TILE (BT * IPT)\n#define WAVES "
            << waves << "\n\n";
        const std::string loop = "#pragma unroll 1\n        for (u32 m = 0u; __ballot(m < mmax) != 0ull; ++m) {\n";
        auto head = [&](const char* name) {
            src << "extern \"C\" __global__ __launch_bounds__(BT) void " << name << "(const cq_query_params P) {\n"
                << "    using namespace chip;\n"
                << "    const int tid = threadIdx.x;\n"
                << "    const int wave = tid >> 6, lane = tid & 63;\n";
            for (const auto& jp : prog_.joins) {
                if (!jp.lds) continue;
                const int j = static_cast<int>(&jp - prog_.joins.data());
                src << "    __shared__ u32 s_ht" << j << "[" << jp.ht_words << "];\n"
                    << "    for (int k = tid; k < " << jp.ht_words << "; k += BT) s_ht" << j << "[k] = P.ht[" << j << "][k];\n";
            }
            emitTableRsrc(src);
            src << "    __syncthreads();\n"
                << "    for (long long tile = blockIdx.x; tile < P.num_tiles; tile += gridDim.x) {\n"
                << "        const long long base = P.row_begin + tile * (long long)TILE;\n";
            for (size_t c = 0; c < prog_.fact_cols.size(); ++c) {
                src << "        int c" << c << "[IPT];\n";
                if (prog_.fact_nullable[c]) src << "        u32 v" << c << " = 0u;\n";
            }
            src << "        u32 fl = 0u;\n";
            emitTileLoad(src, "        ", "", "base", "fl");
            src << body << mlocate
                << "        u32 sc[IPT];\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) sc[i] = 0u;\n"
                << loop << mbody
                << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) sc[i] += (fl >> i) & 1u;\n"
                << "        }\n";
        };
        head("cq_pcount");
        src << "#pragma unroll\n        for (int v = 0; v < VPT; ++v) {\n"
            << "            const u64 n = wave_sum((u64)(sc[4 * v] + sc[4 * v + 1] + sc[4 * v + 2] + sc[4 * v + 3]));\n"
            << "            if (lane == 0) P.tile_count[(tile * VPT + v) * WAVES + wave] = n;\n"
            << "        }\n    }\n    (void)lane;\n}\n\n";
        head("cq_pwrite");
        src << "        u64 at[IPT];\n"
            << "#pragma unroll\n        for (int v = 0; v < VPT; ++v) {\n"
            << "            u64 run = P.tile_off[(tile * VPT + v) * WAVES + wave] +\n"
            << "                      wave_exclusive(sc[4 * v] + sc[4 * v + 1] + sc[4 * v + 2] + sc[4 * v + 3]);\n"
            << "#pragma unroll\n            for (int k = 0; k < 4; ++k) { at[4 * v + k] = run; run += sc[4 * v + k]; }\n"
            << "        }\n"
            << loop << mbody
            << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) {\n"
            << "                if (!((fl >> i) & 1u)) continue;\n"
            << "                const u64 row = at[i]++;\n";
        for (int c = 0; c < ts; ++c) {
            const std::string& v = vals[static_cast<size_t>(c)];
            std::string cast;
            switch (prog_.out[static_cast<size_t>(c)].type) {
                case LogicalType::Float64: cast = "(u64)__double_as_longlong((double)(" + v + "))"; break;
                case LogicalType::Int64: cast = "(u64)(long long)(" + v + ")"; break;
                default: cast = "(u64)(" + v + ")"; break;
            }
            src << "                P.result[row * " << ts << " + " << c << "] = " << cast << ";\n";
            const std::string& ok = oks[static_cast<size_t>(c)];
            if (litFalse(ok)) continue;
            const std::string set = "{ const u64 b = row * " + std::to_string(ts) + " + " + std::to_string(c) +
                                    "; atomicOr(&P.result_valid[b >> 6], 1ull << (b & 63)); }";
            if (litTrue(ok)) src << "                " << set << "\n";
            else src << "                if (" << ok << ") " << set << "\n";
        }
        src << "            }\n        }\n    }\n    (void)lane;\n}\n\n";
    }

    void emitFilterOp(std::ostringstream& body, const Op& op) {
        if (op.fast) {
            const std::string c = fastCond(op.pred);
            body << "        // filter (fast path, raw values)\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                 << "            if (!(" << c << ")) " << flClear() << "\n";
        } else {
            const std::string c = andTerms({vvalid(op.pred), vexpr(op.pred)});
            body << "        // filter (three-valued logic)\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                 << "            if (!(" << c << ")) " << flClear() << "\n";
        }
        body << "        if (__ballot(" << flAny() << ") == 0ull) continue;\n";
    }

    // key-indexed table in HBM read through a buffer resource (rt<j>, set up
    // at the top of the kernel by emitTableRsrc)
    bool useRsrc(const JoinProgram& jp) const {
        const bool stream = opt_.stream_large_tables && !jp.lds && jp.ht_words * 4 > 32 * 1024;
        return jp.dense && !jp.lds && opt_.buffer_tables && !opt_.batch_body && !stream && opt_.split_probes &&
               jp.ht_words * 4 < (uint64_t(1) << 31);
    }
    void emitTableRsrc(std::ostringstream& src) const {
        for (size_t j = 0; j < prog_.joins.size(); ++j)
            if (useRsrc(prog_.joins[j]))
                src << "    const chip::table_rsrc rt" << j << " = chip::make_table_rsrc(P.ht[" << j << "], "
                    << prog_.joins[j].ht_words * 4 << "u);\n";
    }

    // Probe of join `join` with the per-row key expression `key` (of item i);
    // keyValid: the key's validity bit expression, "" when it cannot be NULL.
    // part 1 emits only the locate and load loops of a key-indexed HBM
    // probe (hoisted ahead of LDS work), part 2 the decode that completes it.
    void emitProbe(std::ostringstream& body, int join, const std::string& key, const std::string& keyValid, int part = 0) {
        const JoinProgram& jp = prog_.joins[static_cast<size_t>(join)];
        const std::string j = std::to_string(join);
        if (part != 0) {
            const std::string al = (jp.db & (jp.db - 1)) == 0 ? "true" : "false";
            if (part == 1) {
                std::string live = flGet();
                if (!keyValid.empty()) live = "(" + live + " && " + keyValid + ")";
                body << "        // " << jp.dim_table << " on " << jp.fk_col << ": table loads issued ahead of the LDS probes\n"
                     << "        u32 q" << j << "[IPT], x" << j << "[IPT], l" << j << "[IPT], h" << j << "[IPT];\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                     << "            q" << j << "[i] = chip::dense_locate(" << key << ", " << live << ", " << jp.dmin << ", "
                     << jp.dspan << ", " << jp.zmode << ", " << jp.zx << ", " << jp.cal << ", x" << j << "[i]);\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                     << "            chip::dense_words_rsrc<" << al << ">(rt" << j << ", x" << j << "[i], " << jp.db << ", l" << j
                     << "[i], h" << j << "[i]);\n";
                return;
            }
            body << "        // probe " << jp.dim_table << " on " << jp.fk_col << " (loads issued above)\n"
                 << "        int p" << j << "[IPT];\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            p" << j << "[i] = 0;\n"
                 << "            if (!chip::dense_decode<" << al << ", " << jp.zmode << ">(q" << j << "[i], x" << j << "[i], l" << j << "[i], h" << j
                 << "[i], " << jp.db << ", " << densePbase(jp) << ", " << (jp.doff ? "true" : "false") << ", p" << j
                 << "[i])) " << flClear() << "\n"
                 << "        }\n"
                 << "        if (__ballot(" << flAny() << ") == 0ull) continue;\n";
            visible_mask_ |= 1u << join;
            return;
        }
        const std::string table = jp.lds ? "s_ht" + j : "P.ht[" + j + "]";
        // tables beyond L1 (32 KiB) that stay in HBM may be read with
        // streaming loads (L1 keeps the small ones)
        const bool stream = opt_.stream_large_tables && !jp.lds && jp.ht_words * 4 > 32 * 1024;
        const std::string st = stream ? "true" : "false";
        body << "        // probe " << jp.dim_table << " on " << (key.rfind("kor", 0) == 0 ? "the alternative's key" : jp.fk_col) << "\n";
        body << "        int p" << j << "[IPT];\n";
        if (opt_.split_probes && (jp.dense || jp.compact)) {
            // locate every row's entry, issue all loads, then decode:
            // one table round trip per tile instead of one per row
            std::string live = flGet();
            if (!keyValid.empty()) live = "(" + live + " && " + keyValid + ")";
            body << "        {\n";
            if (jp.dense) {
                const std::string al = (jp.db & (jp.db - 1)) == 0 ? "true" : "false";
                body << "        u32 q" << j << "[IPT], x" << j << "[IPT], l" << j << "[IPT], h" << j << "[IPT];\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                     << "            q" << j << "[i] = chip::dense_locate(" << key << ", " << live << ", " << jp.dmin << ", "
                     << jp.dspan << ", " << jp.zmode << ", " << jp.zx << ", " << jp.cal << ", x" << j << "[i]);\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n";
                if (useRsrc(jp))
                    body << "            chip::dense_words_rsrc<" << al << ">(rt" << j << ", x" << j << "[i], " << jp.db << ", l" << j
                         << "[i], h" << j << "[i]);\n";
                else
                    body << "            chip::dense_words<" << al << ", " << st << ">(" << table << ", x" << j << "[i], " << jp.db
                         << ", l" << j << "[i], h" << j << "[i]);\n";
                body
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                     << "            p" << j << "[i] = 0;\n"
                     << "            if (!chip::dense_decode<" << al << ", " << jp.zmode << ">(q" << j << "[i], x" << j << "[i], l" << j << "[i], h" << j
                     << "[i], " << jp.db << ", " << densePbase(jp) << ", " << (jp.doff ? "true" : "false") << ", p" << j
                     << "[i])) " << flClear() << "\n"
                     << "        }\n";
            } else {
                body << "        int s" << j << "[IPT]; u32 w" << j << "[IPT];\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n"
                     << "            w" << j << "[i] = chip::compact_locate(" << key << ", " << live << ", " << jp.kmin << ", "
                     << jp.len << ", " << jp.eb << ", s" << j << "[i]);\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) w" << j << "[i] = chip::ht_word<" << st << ">("
                     << table << ", w" << j << "[i]);\n"
                     << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                     << "            p" << j << "[i] = 0;\n"
                     << "            if (!chip::compact_decode(w" << j << "[i], s" << j << "[i], " << key << ", " << jp.kmin << ", "
                     << jp.len << ", " << jp.eb << ", " << jp.tb << ", " << jp.jlo << ", " << jp.pmin << ", p" << j << "[i])) "
                     << flClear() << "\n"
                     << "        }\n";
            }
            body << "        }\n";
        } else {
            body << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            p" << j << "[i] = 0;\n"
                 << "            if (!" << flGet() << ") continue;\n";
            if (!keyValid.empty()) body << "            if (!" << keyValid << ") { " << flClear() << " continue; }\n";
            if (jp.dense)
                body << "            if (!chip::ht_probe_dense<" << ((jp.db & (jp.db - 1)) == 0 ? "true" : "false") << ", " << st << ">("
                     << table << ", " << key << ", " << jp.dmin << ", " << jp.dspan << ", " << jp.db << ", " << densePbase(jp) << ", "
                     << (jp.doff ? "true" : "false") << ", " << jp.zmode << ", " << jp.zx << ", p" << j << "[i], " << jp.cal << ")) "
                     << flClear() << "\n";
            else if (jp.compact)
                body << "            if (!chip::ht_probe_compact<" << st << ">(" << table << ", " << key << ", " << jp.kmin << ", " << jp.len
                     << ", " << jp.eb << ", " << jp.tb << ", " << jp.jlo << ", " << jp.pmin << ", p" << j << "[i])) " << flClear() << "\n";
            else
                body << "            if (!chip::ht_probe_ref((const int*)" << table << ", " << key << ", " << jp.kmin << ", " << jp.len
                     << ", " << (jp.pair ? "true" : "false") << ", p" << j << "[i])) " << flClear() << "\n";
            body << "        }\n";
        }
        body << "        if (__ballot(" << flAny() << ") == 0ull) continue;\n";
        visible_mask_ |= 1u << join;
    }

    // Loads one tile of every fact column into <pre>c<s>[IPT] / <pre>v<s>
    // (validity bits) and sets the row-present mask `flv` (full tiles: 128-bit
    // nontemporal loads; the ragged last tile: guarded scalar loads).
    void emitTileLoad(std::ostringstream& src, const std::string& ind, const std::string& pre, const std::string& base,
                      const std::string& flv) {
        const size_t nf = prog_.fact_cols.size();
        const int ipt = 4 * prog_.vectors_per_thread;
        src << ind << "if (" << base << " + TILE <= P.row_end) {\n"
            << "#pragma unroll\n" << ind << "    for (int v = 0; v < VPT; ++v) {\n"
            << ind << "        const long long r = " << base << " + (long long)v * (BT * 4) + tid * 4;\n";
        for (size_t s = 0; s < nf; ++s) {
            const std::string c = pre + "c" + std::to_string(s);
            src << ind << "        { const v4i x = __builtin_nontemporal_load((const v4i*)(P.col[" << s << "] + r)); " << c
                << "[4 * v] = x.x; " << c << "[4 * v + 1] = x.y; " << c << "[4 * v + 2] = x.z; " << c << "[4 * v + 3] = x.w; }\n";
            if (prog_.fact_nullable[s])
                src << ind << "        " << pre << "v" << s << " |= ((P.valid[" << s << "][r >> 5] >> (r & 31)) & 0xfu) << (4 * v);\n";
        }
        src << ind << "    }\n"
            << ind << "    " << flv << " = " << (ipt == 32 ? "0xffffffffu" : "((1u << IPT) - 1u)") << ";\n"
            << ind << "} else {\n"
            << "#pragma unroll\n" << ind << "    for (int i = 0; i < IPT; ++i) {\n"
            << ind << "        const long long r = " << base << " + (long long)(i >> 2) * (BT * 4) + tid * 4 + (i & 3);\n"
            << ind << "        const bool in = r < P.row_end;\n"
            << ind << "        if (in) " << flv << " |= 1u << i;\n";
        for (size_t s = 0; s < nf; ++s) {
            src << ind << "        " << pre << "c" << s << "[i] = in ? P.col[" << s << "][r] : 0;\n";
            if (prog_.fact_nullable[s])
                src << ind << "        if (in) " << pre << "v" << s << " |= chip::bit32(P.valid[" << s << "], r) << i;\n";
        }
        src << ind << "    }\n" << ind << "}\n";
    }

    // ------------------------------------------------ interpreted tier
    // The same expressions the HIP source would contain, compiled to
    // register code for the ahead-of-time interpreter (no hipRTC).
    // The linear form of an interpreted plan (cq_lin_program), when it has
    // one; otherwise lin.enabled stays 0 and the bytecode runs.  fcmp_lits:
    // the P.lit index of every fast-path compare in emission order;
    // agg_lit_base: per aggregate, the first P.lit index its value
    // expression took (its literals follow in left-to-right order).
    void analyzeLinear(cq_vm_program* vp, const std::vector<int>& fcmp_lits, const std::vector<int>& agg_lit_base) {
        cq_lin_program L;
        std::memset(&L, 0, sizeof L);
        L.join = -1;
        auto pos = [&](const std::string& col) -> int {  // column position, -1: no room
            const int s = factSlot(col);
            for (int k = 0; k < L.ncols; ++k)
                if (L.slot[k] == s) return k;
            if (L.ncols >= CQ_LIN_COLS) return -1;
            L.slot[L.ncols] = s;
            L.nullable[L.ncols] = prog_.fact_nullable[static_cast<size_t>(s)] ? 1 : 0;
            return L.ncols++;
        };
        if (prog_.projection || prog_.joins.size() > 1) return;
        if (prog_.naggs > CQ_LIN_AGGS || (!prog_.grouped && prog_.storage_ts > CQ_LIN_RACC)) return;
        if (!prog_.joins.empty()) {
            const JoinProgram& jp = prog_.joins[0];
            if (!jp.dense || jp.mht || jp.rowid || jp.fk_alts.size() != 1) return;
            L.join = 0;
        }
        size_t next_fcmp = 0;
        for (const Op& op : pipelineOrder()) {
            if (op.kind == 1) continue;
            if (!op.fast) return;
            std::vector<const BinaryExpr*> leaves;
            andLeaves(op.pred, leaves);
            for (const BinaryExpr* lf : leaves) {
                const auto& col = static_cast<const ColumnRefExpr*>(lf->left.get())->column_name;
                if (L.nterms >= CQ_LIN_TERMS || next_fcmp >= fcmp_lits.size()) return;
                cq_lin_term& t = L.term[L.nterms++];
                if ((t.col = pos(col)) < 0) return;
                switch (lf->op_type) {
                    case ExprType::OP_EQ: t.op = CQ_OP_EQ; break;
                    case ExprType::OP_NEQ: t.op = CQ_OP_NE; break;
                    case ExprType::OP_LT: t.op = CQ_OP_LT; break;
                    case ExprType::OP_LTE: t.op = CQ_OP_LE; break;
                    case ExprType::OP_GT: t.op = CQ_OP_GT; break;
                    case ExprType::OP_GTE: t.op = CQ_OP_GE; break;
                    default: return;
                }
                t.lit = fcmp_lits[next_fcmp++];
            }
        }
        // filtered columns take the first positions; terms grouped by
        // position (every term must hold, so their order is free)
        L.nfilter_cols = L.ncols;
        std::stable_sort(L.term, L.term + L.nterms, [](const cq_lin_term& x, const cq_lin_term& y) { return x.col < y.col; });
        for (int c = 0, t = 0; c < CQ_LIN_COLS; ++c) {
            while (t < L.nterms && L.term[t].col == c) ++t;
            L.term_end[c] = t;
        }
        if (L.join == 0 && (L.join_col = pos(prog_.joins[0].fk_col)) < 0) return;
        // a small table is probed from LDS (the budget is what is left
        // beside the staged columns and LDS group aggregates)
        if (L.join == 0 && opt_.allow_lds_tables && prog_.joins[0].ht_words <= CQ_LIN_LDS_TABLE_WORDS)
            L.join_lds_words = static_cast<int>(prog_.joins[0].ht_words);
        // one operand: a fact column, the join's payload or (with `litc`)
        // the next literal of the expression, in the type lit() gave it
        auto operand = [&](const ExprNode* e, cq_lin_step& st, int* litc) -> bool {
            if (e->getType() == ExprType::LITERAL_INT) {
                if (!litc) return false;
                const int64_t v = static_cast<const LiteralIntExpr*>(e)->value;
                st.kind = CQ_LIN_LIT;
                st.idx = (*litc)++;
                st.i32 = (v >= -2147483647LL && v <= 2147483647LL) ? 1 : 0;
                return true;
            }
            if (e->getType() != ExprType::COLUMN_REF) return false;
            const auto& col = static_cast<const ColumnRefExpr*>(e)->column_name;
            if (tableOf(col) == prog_.fact_table) {
                st.kind = CQ_LIN_COL;
                return (st.idx = pos(col)) >= 0;
            }
            if (L.join == 0 && prog_.joins[0].payload_col == col) {
                st.kind = CQ_LIN_PAY;
                st.idx = 0;
                return true;
            }
            return false;
        };
        // left-deep chain: the leftmost leaf a column, right operands leaves
        std::function<bool(const ExprNode*, cq_lin_agg&, int*)> chain = [&](const ExprNode* e, cq_lin_agg& g, int* litc) -> bool {
            const ExprType t = e->getType();
            if (t == ExprType::COLUMN_REF) {
                if (g.nsteps >= CQ_LIN_SRCS) return false;
                return operand(e, g.step[g.nsteps++], nullptr);
            }
            if (t != ExprType::OP_ADD && t != ExprType::OP_SUB && t != ExprType::OP_MUL) return false;
            const auto* b = static_cast<const BinaryExpr*>(e);
            if (!chain(b->left.get(), g, litc) || g.nsteps >= CQ_LIN_SRCS) return false;
            cq_lin_step& st = g.step[g.nsteps++];
            st.op = t == ExprType::OP_ADD ? CQ_OP_ADD : (t == ExprType::OP_SUB ? CQ_OP_SUB : CQ_OP_MUL);
            return operand(b->right.get(), st, litc);
        };
        // validity of arithmetic over leaves: AND of the nullable fact
        // columns' bits (literals and payload registers are always valid)
        std::function<bool(const ExprNode*, cq_lin_agg&)> validity = [&](const ExprNode* e, cq_lin_agg& g) -> bool {
            const ExprType t = e->getType();
            if (t == ExprType::LITERAL_INT) return true;
            if (t == ExprType::COLUMN_REF) {
                cq_lin_step st;
                std::memset(&st, 0, sizeof st);
                if (!operand(e, st, nullptr)) return false;
                if (st.kind != CQ_LIN_COL || !L.nullable[st.idx]) return true;
                for (int k = 0; k < g.nvalid; ++k)
                    if (g.valid_col[k] == st.idx) return true;
                if (g.nvalid >= CQ_LIN_SRCS) return false;
                g.valid_col[g.nvalid++] = st.idx;
                return true;
            }
            if (t != ExprType::OP_ADD && t != ExprType::OP_SUB && t != ExprType::OP_MUL) return false;
            const auto* b = static_cast<const BinaryExpr*>(e);
            return validity(b->left.get(), g) && validity(b->right.get(), g);
        };
        const auto* a = agg_;
        L.naggs = prog_.naggs;
        for (int i = 0; i < prog_.naggs; ++i) {
            const auto& d = a->aggregates[static_cast<size_t>(i)];
            cq_lin_agg& g = L.agg[i];
            if (d.hasStarArgument()) continue;  // constant 1, always valid
            if (!validity(d.agg_expr.get(), g)) return;
            if (d.isCount()) continue;
            int litc = agg_lit_base[static_cast<size_t>(i)];
            if (!chain(d.agg_expr.get(), g, &litc)) return;
        }
        if (prog_.grouped) {
            if (prog_.ngroups > CQ_LIN_KEYS || !vp->decodable) return;
            L.ngroups = prog_.ngroups;
            for (int g = 0; g < prog_.ngroups; ++g) {
                const ExprNode* e = a->group_by_exprs[static_cast<size_t>(g)].get();
                cq_lin_step st;
                std::memset(&st, 0, sizeof st);
                if (!operand(e, st, nullptr)) return;
                L.key[g].kind = st.kind;
                L.key[g].col = st.idx;
                L.key[g].valid_col = (st.kind == CQ_LIN_COL && L.nullable[st.idx]) ? st.idx : -1;
            }
        }
        L.enabled = 1;
        vp->lin = L;
    }

    void emitVm() {
        auto vp = std::make_shared<cq_vm_program>();
        std::memset(vp.get(), 0, sizeof(cq_vm_program));
        std::vector<cq_vm_instr> code;
        vm::Builder b(code);
        const int nj = static_cast<int>(prog_.joins.size());
        for (int j = 0; j < nj; ++j) finalizeLayout(prog_.joins[static_cast<size_t>(j)], false);
        prog_.block_threads = 128;
        prog_.vectors_per_thread = 1;
        vp->njoins = nj;
        // dimension tables: layout and build-filter programs
        for (int j = 0; j < nj; ++j) {
            JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            cq_vm_join& J = vp->joins[j];
            J.layout = jp.dense ? CQ_L_DENSE : (jp.compact ? CQ_L_COMPACT : CQ_L_REF);
            J.fk_slot = factSlot(jp.fk_col);
            J.pk_dslot = jp.pk_dslot = dimSlot(jp.pk_col);
            J.payload_dslot = jp.pair ? (jp.payload_dslot = dimSlot(jp.payload_col)) : -1;
            J.pk_nullable = nullable(jp.pk_col) ? 1 : 0;
            J.fk_nullable = prog_.fact_nullable[static_cast<size_t>(J.fk_slot)] ? 1 : 0;
            J.kmin = static_cast<int>(jp.kmin);
            J.len = static_cast<int>(jp.len);
            J.eb = jp.eb, J.tb = jp.tb, J.jlo = jp.jlo, J.pmin = static_cast<int>(jp.pmin);
            J.dmin = static_cast<int>(jp.dmin), J.dspan = static_cast<int>(jp.dspan), J.db = jp.db;
            J.pbase = static_cast<int>(densePbase(jp)), J.doff = jp.doff, J.zmode = jp.zmode, J.zx = jp.zx, J.cal = jp.cal;
            J.build_begin = static_cast<int>(code.size());
            if (const ExprNode* bp = build_pred_[static_cast<size_t>(j)]) {
                vm::Val c = b.compile(buildCond(bp));
                b.emit(CQ_OP_FILTER, c.type, 0, 0, c.reg, 0, 0, 0);
                b.release(c);
            }
            J.build_end = static_cast<int>(code.size());
        }
        // fact pipeline in the plan's execution order
        vp->scan_begin = static_cast<int>(code.size());
        std::vector<int> pay_reg(static_cast<size_t>(nj));
        for (int j = 0; j < nj; ++j) {
            pay_reg[static_cast<size_t>(j)] = b.allocPermanent();
            b.setPayloadReg(j, pay_reg[static_cast<size_t>(j)]);
        }
        visible_mask_ = 0;
        std::vector<int> fcmp_lits, agg_lit_base;
        for (const Op& op : pipelineOrder()) {
            if (op.kind == 0) {
                if (op.fast) {
                    // one fused compare-and-filter per `col op literal` of the AND chain
                    std::vector<const BinaryExpr*> leaves;
                    andLeaves(op.pred, leaves);
                    for (const BinaryExpr* lf : leaves) {
                        const auto& col = static_cast<const ColumnRefExpr*>(lf->left.get())->column_name;
                        if (tableOf(col) != prog_.fact_table)
                            unsupported("fast-path filter on dimension column " + col + " above its join");
                        const int k = static_cast<int>(prog_.lits.size());
                        (void)litInt32(static_cast<const LiteralIntExpr*>(lf->right.get())->value);
                        fcmp_lits.push_back(k);
                        int cmp = CQ_OP_EQ;
                        switch (lf->op_type) {
                            case ExprType::OP_EQ: cmp = CQ_OP_EQ; break;
                            case ExprType::OP_NEQ: cmp = CQ_OP_NE; break;
                            case ExprType::OP_LT: cmp = CQ_OP_LT; break;
                            case ExprType::OP_LTE: cmp = CQ_OP_LE; break;
                            case ExprType::OP_GT: cmp = CQ_OP_GT; break;
                            case ExprType::OP_GTE: cmp = CQ_OP_GE; break;
                            default: unsupported("fast-path comparison");
                        }
                        b.emit(CQ_OP_FCMP, cmp, 0, 0, 0, 0, k, factSlot(col));
                    }
                } else {
                    vm::Val c = b.compile(andTerms({vvalid(op.pred), vexpr(op.pred)}));
                    b.emit(CQ_OP_FILTER, c.type, 0, 0, c.reg, 0, 0, 0);
                    b.release(c);
                }
            } else {
                const JoinProgram& jp = prog_.joins[static_cast<size_t>(op.join)];
                const int fs = factSlot(jp.fk_col);
                if (prog_.fact_nullable[static_cast<size_t>(fs)]) {
                    vm::Val v = b.compile("((v" + std::to_string(fs) + " >> i) & 1u)");
                    b.emit(CQ_OP_FILTER, v.type, 0, 0, v.reg, 0, 0, 0);
                    b.release(v);
                }
                vm::Val k = b.compile("c" + std::to_string(fs) + "[i]");
                b.emit(CQ_OP_PROBE, CQ_T_I32, 0, pay_reg[static_cast<size_t>(op.join)], k.reg, 0, 0, op.join);
                b.release(k);
                visible_mask_ |= 1u << op.join;
            }
            b.emit(CQ_OP_EXIT, 0, 0, 0, 0, 0, 0, 0);
        }
        // aggregation inputs (left in registers for the interpreter's epilogue)
        const auto* a = agg_;
        vp->naggs = prog_.naggs;
        vp->ngroups = prog_.ngroups;
        vp->storage_ts = prog_.storage_ts;
        vp->visible_ts = prog_.visible_ts;
        vp->fast_sparse = prog_.fast_sparse ? 1 : 0;
        vp->groups = prog_.groups;
        if (prog_.naggs > CQ_VM_MAX_AGGS || prog_.ngroups > CQ_VM_MAX_AGGS || prog_.storage_ts > 31)
            unsupported("interpreted tier: too many aggregate slots");
        if (prog_.grouped) {
            vm::Val h = b.compile(groupHash());
            vp->group_reg = h.reg;
            for (int g = 0; g < prog_.ngroups; ++g) {
                const ExprNode* e = a->group_by_exprs[static_cast<size_t>(g)].get();
                vp->key_val[g] = b.toType(b.compile(vexpr(e)), CQ_T_U64).reg;
                const std::string v = vvalid(e);
                vp->key_valid[g] = litTrue(v) ? -1 : b.compile(v).reg;
                vp->key_min[g] = gmin_[static_cast<size_t>(g)];
                vp->key_card[g] = static_cast<long long>(gcard_[static_cast<size_t>(g)]);
                vp->key_nullable[g] = gnullable_[static_cast<size_t>(g)] ? 1 : 0;
            }
            bool dec = true;
            for (bool d : gdense_) dec = dec && d;
            vp->decodable = dec ? 1 : 0;
        }
        for (int i = 0; i < prog_.naggs; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            vp->agg_fn[i] = f == "MIN" ? CQ_VM_MIN : (f == "MAX" ? CQ_VM_MAX : (f == "COUNT" ? CQ_VM_COUNT : CQ_VM_SUM));
            vp->agg_hidden[i] = prog_.hidden_slot[static_cast<size_t>(i)];
            agg_lit_base.push_back(static_cast<int>(prog_.lits.size()));
            const AggRow r = aggRow(i);
            vp->agg_val[i] = f == "COUNT" ? -1 : b.compile(r.value).reg;
            vp->agg_valid[i] = litTrue(r.valid) ? -1 : b.compile(r.valid).reg;
        }
        vp->scan_end = static_cast<int>(code.size());
        // HAVING over the visible result slots of a group
        vp->having_reg = -1;
        if (a->having_predicate) {
            b.resetColumnCache();
            vp->having_begin = static_cast<int>(code.size());
            vm::Val hv = b.compile(havingExpr());
            vp->having_reg = hv.reg;
            vp->having_type = hv.type;
            vp->having_end = static_cast<int>(code.size());
        }
        for (int i = 0; i < prog_.naggs && !prog_.grouped; ++i)
            if (prog_.agg_funcs[static_cast<size_t>(i)] == "COUNT") prog_.count_slots_scalar.push_back(i);
        vp->ncode = static_cast<int>(code.size());
        std::copy(code.begin(), code.end(), vp->code);
        vp->nregs = b.maxRegs();
        // group aggregates in LDS when the domain decodes and fits
        const uint64_t nas = static_cast<uint64_t>(prog_.storage_ts - prog_.ngroups);
        const uint64_t lds = static_cast<uint64_t>(prog_.groups) * (nas * 8 + 4);
        prog_.vm_lds_bytes = (prog_.grouped && vp->decodable && lds <= 64 * 1024) ? lds : 0;
        analyzeLinear(vp.get(), fcmp_lits, agg_lit_base);
        prog_.vm = vp;
        std::ostringstream src;
        src << "// interpreted: " << code.size() << " instructions, " << b.maxRegs() << " registers"
            << (vp->lin.enabled ? ", linear form" : "") << "\n";
        for (size_t k = 0; k < code.size(); ++k)
            src << "// " << k << ": op " << int(code[k].op) << " t" << int(code[k].t) << "/" << int(code[k].t2) << " d" << int(code[k].d)
                << " a" << int(code[k].a) << " b" << int(code[k].b) << " c" << int(code[k].c) << " imm " << code[k].imm << "\n";
        prog_.source = src.str();
        setAlgorithmicBytes();
    }

    // ------------------------------------------------ batched templates
    // Shared-probe form: possible when every join has the key-indexed
    // layout and its attributes (payload + build-filter columns) pack into
    // one 32-bit entry.
    bool sharedEligible() {
        if (!opt_.batch_shared_tables) return false;
        for (size_t j = 0; j < prog_.joins.size(); ++j) {
            JoinProgram& jp = prog_.joins[j];
            if (!jp.dense_ok) return false;
            jp.s_cols.clear();
            jp.s_min.clear();
            jp.s_bits.clear();
            jp.s_off.clear();
            jp.s_payload = -1;
            std::vector<std::string> cols;
            if (jp.pair) cols.push_back(jp.payload_col);
            if (build_pred_[j]) collectColumns(build_pred_[j], cols);
            const auto& tm = meta(jp.dim_table);
            const auto* ks = stats(jp.pk_col);
            jp.s_present = (static_cast<uint64_t>(jp.dspan) == tm.getSize() && ks->cardinality_ == tm.getSize()) ? 0 : 1;
            auto bitsOf = [&](const std::string& c) -> int {
                const auto* cs = stats(c);
                return cs ? bitsFor(static_cast<uint64_t>(cs->max_value_ - cs->min_value_)) : 64;
            };
            int need = jp.s_present;
            std::vector<std::string> packed;
            for (const auto& c : cols) {
                if (std::find(packed.begin(), packed.end(), c) != packed.end()) continue;
                if (!stats(c) || tableOf(c) != jp.dim_table) return false;
                packed.push_back(c);
                need += bitsOf(c);
            }
            if (need > 32) return false;
            // The dimension's other small attributes (at most 32 values)
            // join the entry when they fit its 32 bits and the table stays
            // small (its power-of-two width unchanged, or at most 1 MiB):
            // templates that read different attributes of one dimension then
            // build the same table, and a batch probes it once for all of
            // them.  Attributes are packed in catalog column order.
            for (const auto& c : tm.getColumnNames()) {
                if (c == jp.pk_col || std::find(packed.begin(), packed.end(), c) != packed.end() || !stats(c)) continue;
                const int b = bitsOf(c);
                if (b > 5 || need + b > 32) continue;
                const int w0 = pow2Width(need), w1 = pow2Width(need + b);
                if (w1 != w0 && static_cast<uint64_t>(jp.dspan) * static_cast<uint64_t>(w1) > 8ull * (1u << 20)) continue;
                packed.push_back(c);
                need += b;
            }
            std::vector<std::string> order = tm.getColumnNames();
            std::stable_sort(packed.begin(), packed.end(), [&](const std::string& x, const std::string& y) {
                return std::find(order.begin(), order.end(), x) < std::find(order.begin(), order.end(), y);
            });
            int off = jp.s_present;
            for (const auto& c : packed) {
                const auto* cs = stats(c);
                const int b = bitsOf(c);
                if (c == jp.payload_col) jp.s_payload = static_cast<int>(jp.s_cols.size());
                jp.s_cols.push_back(c);
                jp.s_min.push_back(cs->min_value_);
                jp.s_bits.push_back(b);
                jp.s_off.push_back(off);
                off += b;
            }
            if (off > 32) return false;
        }
        return true;
    }
    // attribute `a` of join j from entry expression `e`, as the raw int value
    std::string sharedAttr(const JoinProgram& jp, size_t a, const std::string& e) const {
        const int b = jp.s_bits[a];
        const std::string code = b == 0 ? "0" : "(int)((" + e + " >> " + std::to_string(jp.s_off[a]) + ") & " +
                                                    std::to_string(b == 32 ? 0xffffffffu : ((1u << b) - 1u)) + "u)";
        return "((int)" + std::to_string(jp.s_min[a]) + " + " + code + ")";
    }
    std::string sharedAttrOf(const JoinProgram& jp, const std::string& col, const std::string& e) const {
        for (size_t a = 0; a < jp.s_cols.size(); ++a)
            if (jp.s_cols[a] == col) return sharedAttr(jp, a, e);
        throw std::runtime_error("internal: attribute " + col + " not packed");
    }

    void emitBatchShared(std::ostringstream& src) {
        prog_.shared_tables = true;
        prog_.batch_bool_flags = true;
        bool_flags_ = true;
        const int nj = static_cast<int>(prog_.joins.size());
        for (int j = 0; j < nj; ++j) {
            JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            int bits = jp.s_present;
            for (int b : jp.s_bits) bits += b;
            jp.shared = true;
            jp.dense = true;
            jp.lds = false;
            jp.db = pow2Width(bits);
            jp.doff = jp.s_present;
            jp.ht_words = denseWords(jp, jp.db);
            // tables of equal key: equal contents (every row, packed attributes)
            std::ostringstream k;
            k << jp.dim_table << "|" << jp.pk_col << "|" << jp.dmin << "|" << jp.dspan << "|" << jp.cal << "|" << jp.db << "|"
              << jp.s_present;
            for (size_t a = 0; a < jp.s_cols.size(); ++a) k << "|" << jp.s_cols[a] << ":" << jp.s_min[a] << ":" << jp.s_bits[a] << ":" << jp.s_off[a];
            jp.s_key = k.str();
        }
        prog_.block_threads = opt_.block_threads > 0 ? opt_.block_threads : 256;
        prog_.vectors_per_thread = std::max(1, std::min(8, opt_.vectors_per_thread));

        // template builds: every dimension row with its packed attributes
        for (int j = 0; j < nj; ++j) {
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            const int pk = dimSlot(jp.pk_col);
            src << "extern \"C\" __global__ __launch_bounds__(256) void cq_build_" << j << "(" << paramDecl() << ") {\n"
                << paramAlias() << "    const long long n = P.dim_rows[" << j << "];\n";
            if (jp.db == 0) {
                src << "    (void)n;\n}\n\n";
                continue;
            }
            src << "    for (long long r = (long long)blockIdx.x * 256 + threadIdx.x; r < n; r += (long long)gridDim.x * 256) {\n";
            if (nullable(jp.pk_col)) src << "        if (!chip::bit32(P.dvalid[" << pk << "], r)) continue;\n";
            src << "        const unsigned int x = chip::dense_index(P.dcol[" << pk << "][r], " << jp.dmin << ", " << jp.dspan << ", "
                << jp.cal << ");\n"
                << "        if (x >= " << jp.dspan << "u) continue;\n"
                << "        unsigned int e = " << jp.s_present << "u;\n";
            for (size_t a = 0; a < jp.s_cols.size(); ++a)
                if (jp.s_bits[a] > 0)
                    src << "        e |= ((unsigned int)(P.dcol[" << dimSlot(jp.s_cols[a]) << "][r] - (int)" << jp.s_min[a] << ")) << "
                        << jp.s_off[a] << ";\n";
            src << "        const unsigned int b = x * " << jp.db << "u;\n"
                << "        atomicOr(&P.ht[" << j << "][b >> 5], e << (b & 31));\n"
                << "    }\n}\n\n";
        }

        // template-level probes, once per tile for all of its queries
        std::ostringstream sh;
        prog_.batch_shared_parts.assign(static_cast<size_t>(nj), std::string());
        for (int j = 0; j < nj; ++j) {
            JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            const int fs = factSlot(jp.fk_col);
            const std::string J = std::to_string(j);
            // e: entry, m: the key is in the table, z: probe key 0 (the
            // reference's empty-slot corner); m and z are lane masks
            if (opt_.split_probes) {
                // entry index of every row first, then all loads, then the
                // present test (a dead row reads entry 0)
                std::ostringstream t;
                t << "        {\n        u32 @X@[IPT];\n"
                  << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                  << "            @X@[i] = 0u; @M@[i] = false; @Z@[i] = false;\n"
                  << "            if (!((tile_fl >> i) & 1u)) continue;\n"
                  << "@FVTEST@"
                  << "            const int fk = @FK@;\n";
                if (jp.zmode != 0) {
                    const std::string zx = jp.zmode == 1 ? "0u" : jp.zmode == 2 ? std::to_string(jp.zx) + "u"
                                                                             : "(u32)(0 - " + std::to_string(jp.dmin) + ")";
                    t << "            if (fk == 0) { @Z@[i] = true; @X@[i] = " << zx << "; continue; }\n";
                }
                t << "            const u32 x = chip::dense_index(fk, " << jp.dmin << ", " << jp.dspan << ", " << jp.cal << ");\n"
                  << "            if (x >= " << jp.dspan << "u) continue;\n"
                  << "            @X@[i] = x; @M@[i] = true;\n"
                  << "        }\n";
                if (jp.db == 0) t << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) @E@[i] = 0u;\n";
                else
                    t << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) @E@[i] = chip::dense_entry_aligned(@TAB@, @X@[i], " << jp.db
                      << ");\n";
                t << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n";
                if (jp.s_present) t << "            if (@M@[i] && !(@E@[i] & 1u)) @M@[i] = false;\n";
                if (jp.zmode == 1) t << "            if (@Z@[i]) @E@[i] = 0u;\n";
                t << "            if (!@M@[i] && !@Z@[i]) @E@[i] = 0u;\n"
                  << "        }\n        }\n";
                jp.s_probe = t.str();
                std::string part = "        // shared probe " + jp.dim_table + " on " + jp.fk_col + "\n" + "        u32 e" + J +
                                   "[IPT]; bool m" + J + "[IPT], z" + J + "[IPT];\n" + jp.s_probe;
                const std::string fv = prog_.fact_nullable[static_cast<size_t>(fs)] ? "((v" + std::to_string(fs) + " >> i) & 1u)" : "";
                auto subst = [](std::string text, const std::string& from, const std::string& to) {
                    for (size_t at = text.find(from); at != std::string::npos; at = text.find(from, at + to.size())) text.replace(at, from.size(), to);
                    return text;
                };
                part = subst(part, "@FVTEST@", fv.empty() ? "" : "            if (!" + fv + ") continue;\n");
                part = subst(part, "@X@", "x" + J);
                part = subst(part, "@M@", "m" + J);
                part = subst(part, "@Z@", "z" + J);
                part = subst(part, "@E@", "e" + J);
                part = subst(part, "@FK@", "c" + std::to_string(fs) + "[i]");
                part = subst(part, "@TAB@", "cq_sht" + J);
                prog_.batch_shared_parts[static_cast<size_t>(j)] = part;
                sh << part;
                continue;
            }
            jp.s_probe.clear();
            std::ostringstream one;
            one << "        // shared probe " << jp.dim_table << " on " << jp.fk_col << "\n"
               << "        u32 e" << J << "[IPT]; bool m" << J << "[IPT], z" << J << "[IPT];\n"
               << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
               << "            e" << J << "[i] = 0u; m" << J << "[i] = false; z" << J << "[i] = false;\n"
               << "            if (!((tile_fl >> i) & 1u)) continue;\n";
            if (prog_.fact_nullable[static_cast<size_t>(fs)]) one << "            if (!((v" << fs << " >> i) & 1u)) continue;\n";
            one << "            const int fk = c" << fs << "[i];\n";
            auto entry = [&](const std::string& x) {
                return jp.db == 0 ? std::string("0u") : "chip::dense_entry_aligned(cq_sht" + J + ", " + x + ", " + std::to_string(jp.db) + ")";
            };
            if (jp.zmode != 0) {
                std::string ze = "0u";
                if (jp.zmode == 2) ze = entry(std::to_string(jp.zx) + "u");
                if (jp.zmode == 3) ze = entry("(u32)(0 - " + std::to_string(jp.dmin) + ")");
                one << "            if (fk == 0) { z" << J << "[i] = true; e" << J << "[i] = " << ze << "; continue; }\n";
            }
            one << "            const u32 x = chip::dense_index(fk, " << jp.dmin << ", " << jp.dspan << ", " << jp.cal << ");\n"
               << "            if (x >= " << jp.dspan << "u) continue;\n"
               << "            const u32 e = " << entry("x") << ";\n";
            if (jp.s_present) one << "            if (!(e & 1u)) continue;\n";
            one << "            e" << J << "[i] = e; m" << J << "[i] = true;\n"
               << "        }\n";
            prog_.batch_shared_parts[static_cast<size_t>(j)] = one.str();
            sh << one.str();
        }
        prog_.batch_shared = sh.str();

        // per-query pipeline: fact filters, this query's build filters on
        // the attributes, payload-dependent filters, aggregation
        std::ostringstream body;
        visible_mask_ = 0;
        std::vector<Op> order = pipelineOrder();
        std::vector<Op> fact_f, late_f;
        for (const Op& op : order) {
            if (op.kind == 1) continue;
            std::vector<std::string> cols;
            collectColumns(op.pred, cols);
            bool fact_only = true;
            for (const auto& c : cols) fact_only = fact_only && tableOf(c) == prog_.fact_table;
            (fact_only ? fact_f : late_f).push_back(op);
        }
        auto emitFilter = [&](const Op& op) {
            body << "        // filter\n#pragma unroll\n        for (int i = 0; i < IPT; ++i)\n";
            if (op.fast) body << "            " << flKeep(fastCond(op.pred, true)) << "\n";
            else body << "            if (!(" << andTerms({vvalid(op.pred), vexpr(op.pred)}) << ")) " << flClear() << "\n";
            body << "        if (__ballot(" << flAny() << ") == 0ull) continue;\n";
        };
        for (const Op& op : fact_f) emitFilter(op);
        for (int j = 0; j < nj; ++j) {
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            const std::string J = std::to_string(j);
            const ExprNode* bp = build_pred_[static_cast<size_t>(j)];
            // straight-line 0/1 arithmetic: the build filter of this query on
            // the packed attributes, no short-circuit branches
            const std::string cond =
                bp ? buildCondWith(bp, [&](const std::string& col) { return sharedAttrOf(jp, col, "e"); }, true) : "1";
            const std::string ins = jp.s_present ? "((int)(e & 1u) & " + cond + ")" : cond;
            const std::string pay = jp.s_payload >= 0 ? sharedAttr(jp, static_cast<size_t>(jp.s_payload), "e") : "0";
            // The payload register of a row only matters while the row is
            // alive, so it is the entry's payload (0 for probe key 0 unless
            // the empty-slot match takes the build row's): independent of
            // the query except in zmode 3, hoisted out of the query loop.
            body << "        // join " << jp.dim_table << ": shared probe, this query's build filter\n"
                 << "        int p" << J << "[IPT];\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            const u32 e = e" << J << "[i];\n"
                 << "            bool keep = m" << J << "[i] & ((" << cond << ") != 0);\n";
            if (jp.zmode == 1) body << "            keep = keep | z" << J << "[i];\n";
            if (jp.zmode == 2) body << "            keep = z" << J << "[i] ? ((" << ins << ") == 0) : keep;\n";
            if (jp.zmode == 3) body << "            keep = keep | z" << J << "[i];\n";
            body << "            fl[i] = fl[i] & keep;\n";
            if (jp.zmode == 3)
                body << "            p" << J << "[i] = z" << J << "[i] ? ((" << ins << ") ? " << pay << " : 0) : " << pay << ";\n";
            else if (jp.zmode != 0)
                body << "            p" << J << "[i] = z" << J << "[i] ? 0 : " << pay << ";\n";
            else
                body << "            p" << J << "[i] = " << pay << ";\n";
            body << "        }\n"
                 << "        if (__ballot(" << flAny() << ") == 0ull) continue;\n";
            visible_mask_ |= 1u << j;
        }
        for (const Op& op : late_f) emitFilter(op);

        // aggregation: scalars per tile, small group domains in the query's
        // LDS region, the rest with global atomics
        std::ostringstream aggcode, pro, epi;
        if (!prog_.grouped && opt_.allow_lds_aggregation && prog_.naggs <= 32) {
            emitScalarAggregateLds(aggcode, pro, epi);
            prog_.lds_aggregation = true;
            prog_.batch_lds_stride = 8 * scalarLdsWords();
            prog_.batch_lds_init = pro.str();
            prog_.batch_lds_flush = epi.str();
        } else if (!prog_.grouped) {
            emitScalarAggregateTile(aggcode);
        } else {
            bool lds = opt_.allow_lds_aggregation && flag_w_ > 0 && prog_.storage_ts <= 64;
            for (bool d : gdense_) lds = lds && d;
            const uint64_t bytes = (aggLdsBytes() + 7) & ~uint64_t(7);
            if (lds && bytes <= 4096) {
                emitLdsAggregate(aggcode, pro, epi, true);
                prog_.lds_aggregation = true;
                prog_.batch_lds_stride = bytes;
                prog_.batch_lds_init = pro.str();
                prog_.batch_lds_flush = epi.str();
            } else if (opt_.allow_lds_aggregation && hashAggEligible() && hashAggBytes() <= 4096) {
                emitLdsHashAggregate(aggcode, pro, epi);
                prog_.lds_aggregation = true;
                prog_.batch_lds_stride = hashAggBytes();
                prog_.batch_hash_slot_bytes = hashSlotBytes();
                prog_.batch_lds_init = pro.str();
                prog_.batch_lds_flush = epi.str();
            } else {
                emitGlobalAggregate(aggcode);
            }
        }
        for (int i = 0; i < prog_.naggs && !prog_.grouped; ++i)
            if (prog_.agg_funcs[static_cast<size_t>(i)] == "COUNT") prog_.count_slots_scalar.push_back(i);
        prog_.batch_body = hoistParams(body.str() + aggcode.str());
        bool_flags_ = false;
        {
            // the mask form is an option on top of the per-query form: if it
            // cannot be built (e.g. too many constants), the template keeps
            // the per-query form only
            const size_t nlits = prog_.lits.size();
            const unsigned vis = visible_mask_;
            try {
                emitMaskForm(fact_f, late_f);
            } catch (const std::exception&) {
                prog_.lits.resize(nlits);
                prog_.batch_mask = false;
                prog_.batch_mask_tables.clear();
                prog_.batch_mask_build.clear();
                prog_.batch_mask_body.clear();
                prog_.batch_mask_body_deferred.clear();
                prog_.batch_mask_deferred.clear();
            }
            visible_mask_ = vis;
        }
        if (prog_.keys_kernel) emitKeysKernel(src);
        emitDenseCount(src);
        prog_.source = src.str();
        setAlgorithmicBytes();
    }

    // ------------------------------------------------ query-mask form
    // A shared template whose per-query work is only (a) fast-path filters
    // on fact columns with a small value domain and (b) build filters on a
    // few packed attribute bits evaluates, per launch, each such predicate
    // for every value of its domain and every query of the template into
    // an LDS table of query bit masks.  A row's accepting queries are then
    // the AND of a few table lookups, and aggregation visits only (row,
    // query) pairs that pass: the cost per row no longer grows with the
    // number of queries.  Results are those of the per-query form (the
    // tables hold the same predicates, evaluated exactly on every value the
    // column or attribute can take).
    struct MaskFact {
        int slot = 0;
        int64_t min = 0, domain = 0;
        std::vector<const BinaryExpr*> leaves;
    };
    static constexpr int64_t kMaskDomain = 1024;
    static constexpr int kMaskAttrBits = 10;

    static void andLeaves(const ExprNode* e, std::vector<const BinaryExpr*>& out) {
        if (e->getType() == ExprType::OP_AND) {
            const auto* b = static_cast<const BinaryExpr*>(e);
            andLeaves(b->left.get(), out);
            andLeaves(b->right.get(), out);
            return;
        }
        out.push_back(static_cast<const BinaryExpr*>(e));
    }

    void emitMaskForm(const std::vector<Op>& fact_f, const std::vector<Op>& late_f) {
        prog_.batch_mask = false;
        prog_.batch_mask_deferred.clear();
        prog_.batch_mask_body_deferred.clear();
        if (!opt_.batch_masks || !late_f.empty()) return;
        // aggregation kinds with a per-row insert for one query
        const bool scalar_lds = !prog_.grouped && prog_.batch_lds_stride > 0;
        const bool hash_lds = prog_.grouped && prog_.batch_hash_slot_bytes > 0;
        const bool dense_lds = prog_.grouped && prog_.batch_lds_stride > 0 && !hash_lds;
        const bool global = prog_.grouped && prog_.batch_lds_stride == 0;
        if (!scalar_lds && !hash_lds && !dense_lds && !global) return;
        // fact filters: fast path on non-nullable columns with small domains
        std::map<std::string, MaskFact> facts;
        std::vector<std::string> fact_order;
        for (const Op& op : fact_f) {
            if (!op.fast) return;
            std::vector<const BinaryExpr*> leaves;
            andLeaves(op.pred, leaves);
            for (const BinaryExpr* b : leaves) {
                const auto& col = static_cast<const ColumnRefExpr*>(b->left.get())->column_name;
                const int slot = factSlot(col);
                if (prog_.fact_nullable[static_cast<size_t>(slot)]) return;
                const auto* cs = stats(col);
                if (!cs || cs->max_value_ < cs->min_value_ || cs->max_value_ - cs->min_value_ + 1 > kMaskDomain) return;
                if (!facts.count(col)) {
                    fact_order.push_back(col);
                    facts[col].slot = slot;
                    facts[col].min = cs->min_value_;
                    facts[col].domain = cs->max_value_ - cs->min_value_ + 1;
                }
                facts[col].leaves.push_back(b);
            }
        }
        // joins: attribute bits their build filters read
        const int nj = static_cast<int>(prog_.joins.size());
        std::vector<int> jlo(static_cast<size_t>(nj), 0), jbits(static_cast<size_t>(nj), -1);
        for (int j = 0; j < nj; ++j) {
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            if (jp.zmode == 3) return;
            const ExprNode* bp = build_pred_[static_cast<size_t>(j)];
            if (!bp) continue;
            std::vector<std::string> cols;
            collectColumns(bp, cols);
            int lo = 32, hi = 0;
            for (const auto& c : cols)
                for (size_t a = 0; a < jp.s_cols.size(); ++a)
                    if (jp.s_cols[a] == c && jp.s_bits[a] > 0) {
                        lo = std::min(lo, jp.s_off[a]);
                        hi = std::max(hi, jp.s_off[a] + jp.s_bits[a]);
                    }
            if (hi <= lo) lo = hi = 0;
            if (hi - lo > kMaskAttrBits) return;
            jlo[static_cast<size_t>(j)] = lo;
            jbits[static_cast<size_t>(j)] = hi - lo;
        }
        // aggregate inputs and the group index must not depend on a query's constants
        auto constantFree = [](const std::string& x) { return x.find("P.lit[") == std::string::npos && x.find("P.flit[") == std::string::npos; };
        const auto* a = agg_;
        const bool decode = keysDecodable();
        std::string gh = prog_.grouped ? groupHash() : "0";
        if (!constantFree(gh)) return;
        for (int i = 0; i < prog_.naggs; ++i) {
            const AggRow r = aggRow(i);
            if (!constantFree(r.value) || !constantFree(r.valid)) return;
        }
        for (int g = 0; g < prog_.ngroups; ++g) {
            const ExprNode* e = a->group_by_exprs[static_cast<size_t>(g)].get();
            if (!constantFree(vvalid(e)) || (!decode && !constantFree(vexpr(e)))) return;
        }

        // ---- tables and their build (per launch, every block)
        std::ostringstream build, body;
        prog_.batch_mask_tables.clear();
        build << "    {\n        const int q0 = B.tq_begin[@T@], nq = B.tq_end[@T@] - q0;\n";
        for (size_t k = 0; k < fact_order.size(); ++k) {
            const MaskFact& f = facts[fact_order[k]];
            const std::string name = "cq_m@T@_f" + std::to_string(k);
            prog_.batch_mask_tables.emplace_back(name, static_cast<int>(f.domain));
            std::string cond;
            for (const BinaryExpr* b : f.leaves)
                cond += (cond.empty() ? "" : " && ") + std::string("(cv ") + cmpOp(b->op_type) + " " +
                        litInt32(static_cast<const LiteralIntExpr*>(b->right.get())->value) + ")";
            build << "        for (int v = tid; v < " << f.domain << "; v += BT) {\n"
                  << "            const int cv = (int)(" << f.min << "ll + v);\n"
                  << "            u64 m = 0ull;\n"
                  << "            for (int q = 0; q < nq; ++q) {\n"
                  << "                const cq_query_params& P = B.qp[q0 + q];\n"
                  << "                if (" << cond << ") m |= 1ull << q;\n"
                  << "            }\n"
                  << "            " << name << "[v] = m;\n"
                  << "        }\n";
        }
        for (int j = 0; j < nj; ++j) {
            if (jbits[static_cast<size_t>(j)] < 0) continue;
            const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
            const std::string name = "cq_m@T@_j" + std::to_string(j);
            const int n = 1 << jbits[static_cast<size_t>(j)];
            prog_.batch_mask_tables.emplace_back(name, n);
            const std::string cond =
                buildCondWith(build_pred_[static_cast<size_t>(j)], [&](const std::string& col) { return sharedAttrOf(jp, col, "e"); }, true);
            build << "        for (int v = tid; v < " << n << "; v += BT) {\n"
                  << "            const u32 e = (u32)v << " << jlo[static_cast<size_t>(j)] << ";\n"
                  << "            u64 m = 0ull;\n"
                  << "            for (int q = 0; q < nq; ++q) {\n"
                  << "                const cq_query_params& P = B.qp[q0 + q];\n"
                  << "                if ((" << cond << ") != 0) m |= 1ull << q;\n"
                  << "            }\n"
                  << "            " << name << "[v] = m;\n"
                  << "        }\n";
        }
        build << "    }\n";

        // ---- per tile: accepting queries of every item
        body << "        // query masks: the template's queries that accept each item\n"
             << "        const int cq_nq = B.tq_end[@T@] - B.tq_begin[@T@];\n"
             << "        const u64 cq_all = cq_nq >= 64 ? ~0ull : ((1ull << cq_nq) - 1ull);\n"
             << "        u64 qm[IPT];\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) qm[i] = ((tile_fl >> i) & 1u) ? cq_all : 0ull;\n";
        for (size_t k = 0; k < fact_order.size(); ++k) {
            const MaskFact& f = facts[fact_order[k]];
            body << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            const unsigned v = (unsigned)(c" << f.slot << "[i] - (int)" << f.min << ");\n"
                 << "            qm[i] &= v < " << f.domain << "u ? cq_m@T@_f" << k << "[v] : 0ull;\n"
                 << "        }\n";
        }
        // Joins whose table can never sit in LDS are probed last, and only
        // for the items some query still accepts after the fact filters and
        // the other joins: each item skipped saves an L2 gather.  Worth it
        // only when something before them can reject items.  The deferred
        // probe waits on the masks, so its L2 round trip no longer overlaps
        // them: a launch of several templates (one block per CU, each
        // wave running every template per tile) keeps the early probe
        // (batch_mask_body), a launch of one template takes the deferred
        // body (batch_mask_body_deferred).
        prog_.batch_mask_deferred.clear();
        std::vector<int> jorder;
        {
            std::vector<int> late;
            bool rejects = !fact_order.empty();
            for (int j = 0; j < nj; ++j) {
                const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
                const bool big = opt_.defer_probes && jp.shared && !jp.s_probe.empty() && 4 * jp.ht_words > kBatchTableLdsBytes;
                if (big) late.push_back(j);
                else {
                    jorder.push_back(j);
                    rejects = rejects || jbits[static_cast<size_t>(j)] >= 0;
                }
            }
            for (int j : late) {
                jorder.push_back(j);
                if (rejects) prog_.batch_mask_deferred.push_back(j);
            }
        }
        auto joinSection = [&](bool defer) {
            std::ostringstream js;
            for (int j : jorder) {
                const JoinProgram& jp = prog_.joins[static_cast<size_t>(j)];
                const std::string J = std::to_string(j);
                if (defer && std::find(prog_.batch_mask_deferred.begin(), prog_.batch_mask_deferred.end(), j) !=
                                 prog_.batch_mask_deferred.end()) {
                    // the shared probe text, live items only
                    const int fs = factSlot(jp.fk_col);
                    std::string t = jp.s_probe;
                    const std::string live = "            if (!((tile_fl >> i) & 1u)) continue;\n";
                    const size_t at = t.find(live);
                    if (at == std::string::npos) throw std::logic_error("internal: shared probe without its live-item test");
                    t.replace(at, live.size(), "            if (qm[i] == 0ull) continue;\n");
                    auto subst = [](std::string text, const std::string& from, const std::string& to) {
                        for (size_t k = text.find(from); k != std::string::npos; k = text.find(from, k + to.size()))
                            text.replace(k, from.size(), to);
                        return text;
                    };
                    const std::string fv = prog_.fact_nullable[static_cast<size_t>(fs)] ? "((v" + std::to_string(fs) + " >> i) & 1u)" : "";
                    t = subst(t, "@FVTEST@", fv.empty() ? "" : "            if (!" + fv + ") continue;\n");
                    t = subst(t, "@X@", "x" + J);
                    t = subst(t, "@M@", "m" + J);
                    t = subst(t, "@Z@", "z" + J);
                    t = subst(t, "@E@", "e" + J);
                    t = subst(t, "@FK@", "c" + std::to_string(fs) + "[i]");
                    t = subst(t, "@TAB@", "cq_sht" + J);
                    js << "        // " << jp.dim_table << " on " << jp.fk_col << ": probed for the items some query still accepts\n"
                       << "        u32 e" << J << "[IPT]; bool m" << J << "[IPT], z" << J << "[IPT];\n" << t;
                }
                const std::string pay = jp.s_payload >= 0 ? sharedAttr(jp, static_cast<size_t>(jp.s_payload), "e") : "0";
                const bool filt = jbits[static_cast<size_t>(j)] >= 0;
                const std::string idx = "((e >> " + std::to_string(jlo[static_cast<size_t>(j)]) + ") & " +
                                        std::to_string((1u << std::max(0, jbits[static_cast<size_t>(j)])) - 1u) + "u)";
                const std::string acc = filt ? "cq_m@T@_j" + J + "[" + idx + "]" : "~0ull";
                js << "        int p" << J << "[IPT];\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                   << "            const u32 e = e" << J << "[i];\n"
                   << "            u64 jm = m" << J << "[i] ? " << acc << " : 0ull;\n";
                if (jp.zmode == 1) js << "            if (z" << J << "[i]) jm = ~0ull;\n";
                if (jp.zmode == 2) {
                    // probe key 0 matches the reference's empty slot: the queries
                    // that did not insert the key sharing that slot
                    const std::string ins = jp.s_present ? "((e & 1u) ? " + acc + " : 0ull)" : acc;
                    js << "            if (z" << J << "[i]) jm = ~" << ins << ";\n";
                }
                js << "            qm[i] &= jm;\n";
                if (jp.zmode != 0) js << "            p" << J << "[i] = z" << J << "[i] ? 0 : " << pay << ";\n";
                else js << "            p" << J << "[i] = " << pay << ";\n";
                js << "        }\n";
            }
            return js.str();
        };
        const std::string head = body.str();
        const std::string joins_early = joinSection(false);
        const std::string joins_deferred = prog_.batch_mask_deferred.empty() ? std::string() : joinSection(true);
        body.str("");
        body.clear();
        visible_mask_ = (1u << nj) - 1u;

        // ---- aggregation over (row, query) pairs
        const int ng = prog_.ngroups, sts = prog_.storage_ts, nas = sts - ng;
        const int64_t G = prog_.groups;
        std::vector<int> key_vals, key_valids;
        for (int g = 0; g < ng; ++g) {
            const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
            if (!litTrue(v) && !litFalse(v)) key_valids.push_back(g);
            if (!decode && (global || hash_lds)) key_vals.push_back(g);
        }
        body << "        u32 pend = 0u;\n        int gh[IPT];\n";
        for (int g : key_vals) body << "        u64 gkv" << g << "[IPT];\n";
        for (int g : key_valids) body << "        u32 gkok" << g << " = 0u;\n";
        for (int i = 0; i < prog_.naggs; ++i) body << "        u64 gav" << i << "[IPT]; u32 gaok" << i << " = 0u;\n";
        body << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
             << "            const int h = " << gh << ";\n"
             << "            gh[i] = h;\n"
             << "            if (qm[i] != 0ull" << (prog_.grouped ? " && (unsigned)h < " + std::to_string(G) + "u" : "") << ") pend |= 1u << i;\n";
        for (int g : key_vals) body << "            gkv" << g << "[i] = (u64)(" << vexpr(a->group_by_exprs[static_cast<size_t>(g)].get()) << ");\n";
        for (int g : key_valids)
            body << "            if (" << vvalid(a->group_by_exprs[static_cast<size_t>(g)].get()) << ") gkok" << g << " |= 1u << i;\n";
        for (int i = 0; i < prog_.naggs; ++i) {
            const AggRow r = aggRow(i);
            body << "            gav" << i << "[i] = " << r.value << ";\n";
            if (litTrue(r.valid)) body << "            gaok" << i << " |= 1u << i;\n";
            else if (!litFalse(r.valid)) body << "            if (" << r.valid << ") gaok" << i << " |= 1u << i;\n";
        }
        body << "        }\n"
             << "        const int lane = tid & 63;\n        (void)lane;\n"
             << "        while (__ballot(pend != 0u) != 0ull) {\n"
             << "            if (pend == 0u) continue;\n"
             << "            const int it = __builtin_ctz(pend);\n"
             << "            u64 qmi = 0ull;\n            int h = 0;\n";
        for (int g : key_vals) body << "            u64 kv" << g << " = 0ull;\n";
        for (int i = 0; i < prog_.naggs; ++i) body << "            u64 av" << i << " = 0ull;\n";
        body << "#pragma unroll\n            for (int j = 0; j < IPT; ++j)\n                if (j == it) {\n"
             << "                    qmi = qm[j];\n                    h = gh[j];\n";
        for (int g : key_vals) body << "                    kv" << g << " = gkv" << g << "[j];\n";
        for (int i = 0; i < prog_.naggs; ++i) body << "                    av" << i << " = gav" << i << "[j];\n";
        body << "                }\n"
             << "            const int qb = __builtin_ctzll(qmi);\n"
             << "            qmi &= qmi - 1ull;\n"
             << "#pragma unroll\n            for (int j = 0; j < IPT; ++j)\n                if (j == it) qm[j] = qmi;\n"
             << "            if (qmi == 0ull) pend &= ~(1u << it);\n"
             << "            const cq_query_params& P = B.qp[B.tq_begin[@T@] + qb];\n"
             << "            (void)P;\n";
        if (prog_.batch_lds_stride > 0)
            body << "            u64* s_agg = (u64*)((char*)cq_lds + B.lds_off[@T@] + (unsigned)qb * B.lds_stride[@T@]);\n"
                 << "            u32* s_flag = (u32*)(s_agg + " << prog_.groups * (prog_.storage_ts - prog_.ngroups) << ");\n"
                 << "            const unsigned cq_hbits = B.hbits[@T@];\n"
                 << "            (void)s_flag; (void)cq_hbits;\n";
        auto keyValidAt = [&](int g) -> std::string {
            const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
            if (litTrue(v) || litFalse(v)) return v;
            return "((gkok" + std::to_string(g) + " >> it) & 1u)";
        };
        auto aggValidAt = [&](int i) -> std::string {
            const std::string v = aggRow(i).valid;
            if (litTrue(v) || litFalse(v)) return v;
            return "((gaok" + std::to_string(i) + " >> it) & 1u)";
        };
        auto aggValueAt = [](int i) { return "av" + std::to_string(i); };
        auto keyValueAt = [](int g) { return "kv" + std::to_string(g); };
        const std::string ind = "            ";
        if (scalar_lds) {
            // per-lane partial slots of the query (emitScalarAggregateLds layout)
            const int na = prog_.naggs;
            std::vector<int> cnt_base(static_cast<size_t>(na), -1);
            int nh = 0;
            for (int i = 0; i < na; ++i)
                if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) cnt_base[static_cast<size_t>(i)] = 64 * (na + nh++);
            const uint64_t seen_word = scalarLdsWords() - 1;
            for (int i = 0; i < na; ++i) {
                const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
                const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
                body << ind << "if (" << aggValidAt(i) << ") {\n"
                     << ind << "    " << atom << "(&s_agg[" << 64 * i << " + lane], av" << i << ");\n";
                if (f != "COUNT" && cnt_base[static_cast<size_t>(i)] < 0)
                    body << ind << "    atomicOr((u32*)(s_agg + " << seen_word << "), " << (1u << i) << "u);\n";
                if (cnt_base[static_cast<size_t>(i)] >= 0)
                    body << ind << "    atomicAdd(&s_agg[" << cnt_base[static_cast<size_t>(i)] << " + lane], 1ull);\n";
                body << ind << "}\n";
            }
        } else if (dense_lds) {
            // emitLdsAggregate layout: u64 slots per group, W flag bits per group
            const int W = flag_w_;
            const int per = 32 / W;
            int lp = 0;
            while ((1 << lp) < per) ++lp;
            body << ind << "u32 want = 1u;\n";
            for (int g : flag_dyn_) {
                if (g >= ng) continue;
                body << ind << "if (" << keyValidAt(g) << ") want |= 1u << " << dynBit(g) << ";\n";
            }
            for (int i = 0; i < prog_.naggs; ++i) {
                const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
                const int slot = ng + i;
                const std::string ref = "s_agg[h * " + std::to_string(nas) + " + " + std::to_string(slot - ng) + "]";
                body << ind << "if (" << aggValidAt(i) << ") {\n";
                if (f == "MIN") body << ind << "    atomicMin(&" << ref << ", av" << i << ");\n";
                else if (f == "MAX") body << ind << "    atomicMax(&" << ref << ", av" << i << ");\n";
                else body << ind << "    atomicAdd(&" << ref << ", av" << i << ");\n";
                const int hs = prog_.hidden_slot[static_cast<size_t>(i)];
                if (hs >= 0) body << ind << "    atomicAdd(&s_agg[h * " << nas << " + " << (hs - ng) << "], 1ull);\n";
                const int db = dynBit(slot);
                if (db > 0) body << ind << "    want |= 1u << " << db << ";\n";
                body << ind << "}\n";
            }
            body << ind << "{ u32* fw = &s_flag[h >> " << lp << "]; const u32 sh = (u32)(h & " << (per - 1) << ") * " << W
                 << "u; if (((*fw >> sh) & want) != want) atomicOr(fw, want << sh); }\n";
        } else if (hash_lds) {
            body << ind << "u32* s_hk = (u32*)s_agg;\n"
                 << ind << "u32* s_hf = s_hk + (1u << cq_hbits);\n"
                 << ind << "u64* s_hv = s_agg + (1u << cq_hbits);\n"
                 << ind << "u32 slot = ((u32)h * 2654435761u) >> (32u - cq_hbits);\n"
                 << ind << "int found = -1;\n"
                 << ind << "for (int t = 0; t < 8; ++t) {\n"
                 << ind << "    u32 cur = s_hk[slot];\n"
                 << ind << "    if (cur == 0u) { const u32 old = atomicCAS(&s_hk[slot], 0u, (u32)h + 1u); cur = old == 0u ? (u32)h + 1u : old; }\n"
                 << ind << "    if (cur == (u32)h + 1u) { found = (int)slot; break; }\n"
                 << ind << "    slot = (slot + 1u) & ((1u << cq_hbits) - 1u);\n"
                 << ind << "}\n"
                 << ind << "if (found >= 0) {\n"
                 << ind << "    u32 want = 1u;\n";
            for (int g : flag_dyn_) {
                if (g >= ng || prog_.fast_sparse) continue;
                body << ind << "    if (" << keyValidAt(g) << ") want |= 1u << " << dynBit(g) << ";\n";
            }
            for (int i = 0; i < prog_.naggs; ++i) {
                const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
                const std::string ref = "s_hv[found * " + std::to_string(nas) + " + " + std::to_string(i) + "]";
                body << ind << "    if (" << aggValidAt(i) << ") {\n";
                if (f == "MIN") body << ind << "        atomicMin(&" << ref << ", av" << i << ");\n";
                else if (f == "MAX") body << ind << "        atomicMax(&" << ref << ", av" << i << ");\n";
                else body << ind << "        atomicAdd(&" << ref << ", av" << i << ");\n";
                const int hs = prog_.hidden_slot[static_cast<size_t>(i)];
                if (hs >= 0) body << ind << "        atomicAdd(&s_hv[found * " << nas << " + " << (hs - ng) << "], 1ull);\n";
                const int db = prog_.fast_sparse ? -1 : dynBit(ng + i);
                if (db > 0) body << ind << "        want |= 1u << " << db << ";\n";
                body << ind << "    }\n";
            }
            body << ind << "    if ((s_hf[found] & want) != want) atomicOr(&s_hf[found], want);\n"
                 << ind << "} else {\n"
                 << ind << "    const long long out = (long long)h * " << sts << ";\n";
            globalRowWith(body, ind + "    ", keyValueAt, keyValidAt, aggValidAt, aggValueAt);
            body << ind << "}\n";
        } else {
            body << ind << "const long long out = (long long)h * " << sts << ";\n";
            globalRowWith(body, ind, keyValueAt, keyValidAt, aggValidAt, aggValueAt);
        }
        body << "        }\n";
        prog_.batch_mask_build = build.str();
        prog_.batch_mask_body = head + joins_early + body.str();
        prog_.batch_mask_body_deferred = joins_deferred.empty() ? std::string() : head + joins_deferred + body.str();
        prog_.batch_mask = true;
    }

    // The per-query body reads its constants from the query's parameter block
    // in device memory: load each one once at the top of the body (scalar
    // loads issued together) instead of at every use.
    static std::string hoistParams(const std::string& code) {
        static const std::regex use(R"(P\.(lit|flit|ht)\[(\d+)\]|P\.result_valid\b|P\.result\b)");
        std::map<std::string, std::string> decl;  // replacement -> declaration
        std::string out;
        auto it = std::sregex_iterator(code.begin(), code.end(), use);
        size_t last = 0;
        for (; it != std::sregex_iterator(); ++it) {
            const std::smatch& m = *it;
            std::string name, d;
            if (m[1].matched) {
                const std::string k = m[1].str();
                name = "cq_" + k + m[2].str();
                const char* type = k == "flit" ? "const double " : (k == "ht" ? "unsigned int* const " : "const long long ");
                d = type + name + " = " + m.str() + ";";
            } else if (m.str() == "P.result_valid") {
                name = "cq_result_valid";
                d = "unsigned long long* const cq_result_valid = P.result_valid;";
            } else {
                name = "cq_result";
                d = "unsigned long long* const cq_result = P.result;";
            }
            decl[name] = d;
            out += code.substr(last, static_cast<size_t>(m.position()) - last) + name;
            last = static_cast<size_t>(m.position() + m.length());
        }
        out += code.substr(last);
        std::string head;
        for (const auto& kv : decl) head += "        " + kv.second + "\n";
        if (!head.empty()) head += "        (void)0;\n";
        return head + out;
    }

    // algorithmic bytes of one execution: fact columns + bitmaps + dimension columns
    void setAlgorithmicBytes() {
        const size_t nf = prog_.fact_cols.size();
        const uint64_t N = static_cast<uint64_t>(prog_.fact_rows);
        uint64_t bytes = 0;
        for (size_t s = 0; s < nf; ++s) {
            bytes += 4 * N;
            if (prog_.fact_nullable[s]) bytes += 8 * ((N + 63) / 64);
        }
        for (const auto& c : prog_.dim_cols) bytes += 4 * static_cast<uint64_t>(meta(tableOf(c)).getSize());
        prog_.algorithmic_bytes = bytes;
    }

    // value stored into / accumulated for aggregate `i` of the current row
    struct AggRow {
        std::string valid, value;
    };
    AggRow aggRow(int i) {
        const auto& d = agg_->aggregates[static_cast<size_t>(i)];
        AggRow r;
        r.valid = d.hasStarArgument() ? "1" : vvalid(d.agg_expr.get());
        r.value = d.isCount() ? "(unsigned long long)1" : "(unsigned long long)(" + mexpr(d.agg_expr.get(), true) + ")";
        return r;
    }

    void emitScalarAggregate(std::ostringstream& code, std::ostringstream& pro, std::ostringstream& epi) {
        const int na = prog_.naggs;
        for (int i = 0; i < na; ++i) {
            pro << "    u64 acc" << i << " = " << (prog_.agg_funcs[static_cast<size_t>(i)] == "MIN" ? "~0ull" : "0ull") << ";\n"
                << "    u32 seen" << i << " = 0u;\n";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) pro << "    u64 cnt" << i << " = 0ull;\n";
        }
        code << "        // scalar aggregates\n#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
             << "            if (!" << flGet() << ") continue;\n";
        for (int i = 0; i < na; ++i) {
            const AggRow r = aggRow(i);
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            code << "            if (" << r.valid << ") {\n"
                 << "                const u64 x = " << r.value << ";\n";
            if (f == "MIN") code << "                acc" << i << " = x < acc" << i << " ? x : acc" << i << ";\n";
            else if (f == "MAX") code << "                acc" << i << " = x > acc" << i << " ? x : acc" << i << ";\n";
            else code << "                acc" << i << " += x;\n";
            code << "                seen" << i << " = 1u;\n";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) code << "                cnt" << i << " += 1ull;\n";
            code << "            }\n";
        }
        code << "        }\n";
        // one atomic per block and aggregate: wave reductions, then the
        // first wave reduces the waves' partials (same-address global atomics
        // from every wave serialize in L2); identity partials are skipped
        // (the slot already holds the identity: 0, or ~0 for MIN)
        int nslots = na;
        for (int i = 0; i < na; ++i) nslots += prog_.hidden_slot[static_cast<size_t>(i)] >= 0 ? 1 : 0;
        epi << "    const int lane = tid & 63, wid = tid >> 6;\n"
            << "    __shared__ u64 s_red[" << nslots << "][BT / 64];\n"
            << "    __shared__ u32 s_any[" << na << "][BT / 64];\n";
        int hs = na;
        for (int i = 0; i < na; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const char* red = f == "MIN" ? "wave_min" : (f == "MAX" ? "wave_max" : "wave_sum");
            epi << "    { const u64 w = " << red << "(acc" << i << "); const u32 any = __ballot(seen" << i << " != 0u) != 0ull ? 1u : 0u;\n"
                << "      if (lane == 0) { s_red[" << i << "][wid] = w; s_any[" << i << "][wid] = any; } }\n";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0)
                epi << "    { const u64 c = wave_sum(cnt" << i << "); if (lane == 0) s_red[" << hs++ << "][wid] = c; }\n";
        }
        epi << "    __syncthreads();\n"
            << "    if (wid == 0) {\n";
        hs = na;
        for (int i = 0; i < na; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const char* red = f == "MIN" ? "wave_min" : (f == "MAX" ? "wave_max" : "wave_sum");
            const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
            const char* idn = f == "MIN" ? "~0ull" : "0ull";
            epi << "        { const u64 w = " << red << "(lane < BT / 64 ? s_red[" << i << "][lane] : " << idn << ");\n"
                << "          const bool any = __ballot(lane < BT / 64 && s_any[" << i << "][lane] != 0u) != 0ull;\n"
                << "          if (lane == 0 && w != " << idn << ") " << atom << "(&P.result[" << i << "], w);\n";
            if (f != "COUNT" && prog_.hidden_slot[static_cast<size_t>(i)] < 0)
                epi << "          if (lane == 0 && any) chip::set_bit64(P.result_valid, " << i << ");\n";
            else
                epi << "          (void)any;\n";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) {
                epi << "          { const u64 c = wave_sum(lane < BT / 64 ? s_red[" << hs++ << "][lane] : 0ull); if (lane == 0 && c) atomicAdd(&P.result["
                    << prog_.hidden_slot[static_cast<size_t>(i)] << "], c); }\n";
            }
            epi << "        }\n";
            if (f == "COUNT")
                epi << "        if (blockIdx.x == 0 && lane == 0 && P.row_end > P.row_begin) chip::set_bit64(P.result_valid, " << i << ");\n";
        }
        epi << "    }\n";
    }

    // scalar aggregates reduced per tile (batched shared scan)
    void emitScalarAggregateTile(std::ostringstream& code) {
        const int na = prog_.naggs;
        code << "        {\n";
        for (int i = 0; i < na; ++i) {
            code << "            u64 acc" << i << " = " << (prog_.agg_funcs[static_cast<size_t>(i)] == "MIN" ? "~0ull" : "0ull")
                 << "; u32 seen" << i << " = 0u;";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) code << " u64 cnt" << i << " = 0ull;";
            code << "\n";
        }
        code << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) {\n"
             << "                if (!" << flGet() << ") continue;\n";
        for (int i = 0; i < na; ++i) {
            const AggRow r = aggRow(i);
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            code << "                if (" << r.valid << ") { const u64 x = " << r.value << "; ";
            if (f == "MIN") code << "acc" << i << " = x < acc" << i << " ? x : acc" << i << ";";
            else if (f == "MAX") code << "acc" << i << " = x > acc" << i << " ? x : acc" << i << ";";
            else code << "acc" << i << " += x;";
            code << " seen" << i << " = 1u;";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) code << " cnt" << i << " += 1ull;";
            code << " }\n";
        }
        code << "            }\n            const int lane = tid & 63;\n";
        for (int i = 0; i < na; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const char* red = f == "MIN" ? "wave_min" : (f == "MAX" ? "wave_max" : "wave_sum");
            const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
            const char* nz = f == "MIN" ? "w != ~0ull" : "w != 0ull";
            code << "            { const u64 w = " << red << "(acc" << i << "); const bool any = __ballot(seen" << i
                 << " != 0u) != 0ull;\n              if (lane == 0 && " << nz << ") " << atom << "(&P.result[" << i << "], w);\n";
            if (f != "COUNT" && prog_.hidden_slot[static_cast<size_t>(i)] < 0)
                code << "              if (lane == 0 && any) chip::set_bit64(P.result_valid, " << i << ");\n";
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0)
                code << "              { const u64 c = wave_sum(cnt" << i << "); if (lane == 0 && c) atomicAdd(&P.result["
                     << prog_.hidden_slot[static_cast<size_t>(i)] << "], c); }\n";
            code << "            }\n";
        }
        code << "        }\n";
    }

    // scalar aggregates of a batched query: per-lane partials in the query's
    // LDS region (one u64 per aggregate and wave lane, plus the AVG counts
    // and a word of "row seen" bits), added with one LDS atomic per wave and
    // tile (no cross-lane reduction per tile); the merge reduces the 64 lane
    // slots and issues one global atomic per aggregate and block
    uint64_t scalarLdsWords() const {
        int nh = 0;
        for (int i = 0; i < prog_.naggs; ++i) nh += prog_.hidden_slot[static_cast<size_t>(i)] >= 0;
        return 64 * static_cast<uint64_t>(prog_.naggs + nh) + 1;
    }
    void emitScalarAggregateLds(std::ostringstream& code, std::ostringstream& pro, std::ostringstream& epi) {
        const int na = prog_.naggs;
        std::vector<int> cnt_base(static_cast<size_t>(na), -1);
        int nh = 0;
        for (int i = 0; i < na; ++i)
            if (prog_.hidden_slot[static_cast<size_t>(i)] >= 0) cnt_base[static_cast<size_t>(i)] = 64 * (na + nh++);
        const uint64_t words = scalarLdsWords();
        const uint64_t seen_word = words - 1;
        std::string mins;
        for (int i = 0; i < na; ++i)
            if (prog_.agg_funcs[static_cast<size_t>(i)] == "MIN") mins += (mins.empty() ? "" : " || ") + std::string("(k >> 6) == ") + std::to_string(i);
        pro << "    for (int k = tid; k < " << words << "; k += BT) s_agg[k] = "
            << (mins.empty() ? "0ull" : "(k < " + std::to_string(64 * na) + " && (" + mins + ")) ? ~0ull : 0ull") << ";\n";
        code << "        {\n";
        for (int i = 0; i < na; ++i) {
            code << "            u64 acc" << i << " = " << (prog_.agg_funcs[static_cast<size_t>(i)] == "MIN" ? "~0ull" : "0ull")
                 << "; u32 seen" << i << " = 0u;";
            if (cnt_base[static_cast<size_t>(i)] >= 0) code << " u64 cnt" << i << " = 0ull;";
            code << "\n";
        }
        // predicated (every expression is total: safe division, wrapping
        // arithmetic), so the loop has no branches
        code << "#pragma unroll\n            for (int i = 0; i < IPT; ++i) {\n"
             << "                const u32 row = " << flGet() << ";\n";
        for (int i = 0; i < na; ++i) {
            const AggRow r = aggRow(i);
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            code << "                { const u32 on = row & (u32)(" << r.valid << "); const u64 x = " << r.value << "; ";
            if (f == "MIN") code << "acc" << i << " = (on && x < acc" << i << ") ? x : acc" << i << ";";
            else if (f == "MAX") code << "acc" << i << " = (on && x > acc" << i << ") ? x : acc" << i << ";";
            else code << "acc" << i << " += on ? x : 0ull;";
            code << " seen" << i << " |= on;";
            if (cnt_base[static_cast<size_t>(i)] >= 0) code << " cnt" << i << " += (u64)on;";
            code << " }\n";
        }
        code << "            }\n            const int lane = tid & 63;\n";
        for (int i = 0; i < na; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
            code << "            if (__ballot(seen" << i << " != 0u) != 0ull) {\n"
                 << "              " << atom << "(&s_agg[" << 64 * i << " + lane], acc" << i << ");\n";
            if (f != "COUNT" && cnt_base[static_cast<size_t>(i)] < 0)
                code << "              if (lane == 0) atomicOr((u32*)(s_agg + " << seen_word << "), " << (1u << i) << "u);\n";
            if (cnt_base[static_cast<size_t>(i)] >= 0)
                code << "              atomicAdd(&s_agg[" << cnt_base[static_cast<size_t>(i)] << " + lane], cnt" << i << ");\n";
            code << "            }\n";
        }
        code << "        }\n";
        epi << "    if (tid < 64) {\n"
            << "        const u32 seen = *(const u32*)(s_agg + " << seen_word << ");\n"
            << "        (void)seen;\n";
        for (int i = 0; i < na; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const char* red = f == "MIN" ? "wave_min" : (f == "MAX" ? "wave_max" : "wave_sum");
            const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
            const char* nz = f == "MIN" ? "x != ~0ull" : "x != 0ull";
            epi << "        { const u64 x = " << red << "(s_agg[" << 64 * i << " + tid]); if (tid == 0 && " << nz << ") " << atom
                << "(&P.result[" << i << "], x); }\n";
            if (f != "COUNT" && cnt_base[static_cast<size_t>(i)] < 0)
                epi << "        if (tid == 0 && ((seen >> " << i << ") & 1u)) chip::set_bit64(P.result_valid, " << i << ");\n";
            if (cnt_base[static_cast<size_t>(i)] >= 0)
                epi << "        { const u64 c = wave_sum(s_agg[" << cnt_base[static_cast<size_t>(i)] << " + tid]); if (tid == 0 && c) atomicAdd(&P.result["
                    << prog_.hidden_slot[static_cast<size_t>(i)] << "], c); }\n";
        }
        epi << "    }\n";
    }

    // per-row bit mask of storage slots whose validity bit the row sets
    void rowSlotBits(std::ostringstream& code, const std::string& indent, std::vector<AggRow>& rows) {
        const auto* a = agg_;
        code << indent << "u32 want = 0x80000000u;\n";
        if (!prog_.fast_sparse) {
            for (int g = 0; g < prog_.ngroups; ++g) {
                const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
                if (litTrue(v)) code << indent << "want |= 1u << " << g << ";\n";
                else if (!litFalse(v)) code << indent << "if (" << v << ") want |= 1u << " << g << ";\n";
            }
        }
        rows.clear();
        for (int i = 0; i < prog_.naggs; ++i) rows.push_back(aggRow(i));
    }

    // LDS flag bits per group: bit 0 = group hit, then one bit per storage
    // slot whose validity depends on the row ("dynamic"); slots valid on
    // every hit ("static") are derived from the hit bit at the merge.
    void computeFlagLayout() {
        flag_dyn_.clear();
        flag_static_.clear();
        flag_w_ = 0;
        if (!prog_.grouped) return;
        const auto* a = agg_;
        // aggregation runs after every probe: all payload registers visible
        const unsigned saved_visible = visible_mask_;
        visible_mask_ = (1u << prog_.joins.size()) - 1u;
        if (!prog_.fast_sparse) {
            for (int g = 0; g < prog_.ngroups; ++g) {
                const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
                if (litTrue(v)) flag_static_.push_back(g);
                else if (!litFalse(v)) flag_dyn_.push_back(g);
            }
            for (int i = 0; i < prog_.naggs; ++i) {
                const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
                const int slot = prog_.ngroups + i;
                if (f == "COUNT") flag_static_.push_back(slot);
                else if (prog_.hidden_slot[static_cast<size_t>(i)] < 0) {
                    const std::string v = aggRow(i).valid;
                    if (litTrue(v)) flag_static_.push_back(slot);
                    else if (!litFalse(v)) flag_dyn_.push_back(slot);
                }
            }
        }
        visible_mask_ = saved_visible;
        const int need = 1 + static_cast<int>(flag_dyn_.size());
        if (need > 32) return;  // no LDS aggregation
        flag_w_ = 1;
        while (flag_w_ < need) flag_w_ <<= 1;
    }
    uint64_t aggLdsBytes() const {
        const uint64_t G = static_cast<uint64_t>(prog_.groups);
        const uint64_t nas = static_cast<uint64_t>(prog_.storage_ts - prog_.ngroups);
        const uint64_t w = static_cast<uint64_t>(std::max(flag_w_, 1));
        return G * nas * 8 + ((G * w + 31) / 32) * 4;
    }
    int dynBit(int slot) const {
        for (size_t k = 0; k < flag_dyn_.size(); ++k)
            if (flag_dyn_[k] == slot) return 1 + static_cast<int>(k);
        return -1;
    }

    // batch_form: no __shared__ declarations and no block barrier; the
    // caller provides s_agg / s_flag pointers into its per-query LDS region
    void emitLdsAggregate(std::ostringstream& code, std::ostringstream& pro, std::ostringstream& epi, bool batch_form = false) {
        const int ng = prog_.ngroups, sts = prog_.storage_ts, nas = sts - ng;
        const int64_t G = prog_.groups;
        const int W = flag_w_;
        const int per = 32 / W;
        int lp = 0;
        while ((1 << lp) < per) ++lp;
        const int64_t FW = (G * W + 31) / 32;
        const std::string mask = W == 32 ? "0xffffffffu" : std::to_string((1u << W) - 1u) + "u";
        if (!batch_form)
            pro << "    __shared__ u64 s_agg[" << G * nas << "];\n"
                << "    __shared__ u32 s_flag[" << FW << "];  // " << W << " flag bits per group\n";
        // MIN slots start at ~0
        std::string init = "0ull";
        if (!prog_.min_slots.empty()) {
            std::string cond;
            for (int s : prog_.min_slots) cond += (cond.empty() ? "" : " || ") + std::string("k % ") + std::to_string(nas) + " == " + std::to_string(s - ng);
            init = "((" + cond + ") ? ~0ull : 0ull)";
        }
        pro << "    for (int k = tid; k < " << G * nas << "; k += BT) s_agg[k] = " << init << ";\n"
            << "    for (int k = tid; k < " << FW << "; k += BT) s_flag[k] = 0u;\n";
        const auto* a = agg_;
        std::function<std::string(int)> keyValid = [&](int g) { return vvalid(a->group_by_exprs[static_cast<size_t>(g)].get()); };
        std::function<std::string(int)> aggValid = [&](int i) { return aggRow(i).valid; };
        std::function<std::string(int)> aggValue = [&](int i) { return aggRow(i).value; };
        if (batch_form) {
            // batched: live rows are sparse, one queued row per lane per round
            code << "        // GROUP BY: perfect hash, LDS-privatized aggregates\n";
            const RowQueue rq = emitRowQueue(code);
            keyValid = rq.keyValid;
            aggValid = rq.aggValid;
            aggValue = rq.aggValue;
            code << "            {\n"
                 << "            u32 want = 1u;\n";
        } else if (opt_.split_probes) {
            // group slots and flag words of every row first (all flag reads
            // in flight at once), then the updates
            code << "        // GROUP BY: perfect hash, LDS-privatized aggregates\n"
                 << "        int gh[IPT]; u32 gw[IPT];\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            gh[i] = -1;\n"
                 << "            if (!" << flGet() << ") continue;\n"
                 << "            const int h = " << groupHash() << ";\n"
                 << "            if ((unsigned)h < " << G << "u) gh[i] = h;\n"
                 << "        }\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) gw[i] = s_flag[(gh[i] < 0 ? 0 : gh[i]) >> " << lp << "];\n"
                 << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
                 << "            if (gh[i] < 0) continue;\n"
                 << "            const int h = gh[i];\n"
                 << "            u32 want = 1u;\n";
        } else {
            code << "        // GROUP BY: perfect hash, LDS-privatized aggregates\n#pragma unroll\n"
                 << "        for (int i = 0; i < IPT; ++i) {\n"
                 << "            if (!" << flGet() << ") continue;\n"
                 << "            const int h = " << groupHash() << ";\n"
                 << "            if ((unsigned)h >= " << G << "u) continue;\n"
                 << "            u32 want = 1u;\n";
        }
        for (int g : flag_dyn_) {
            if (g >= ng) continue;
            code << "            if (" << keyValid(g) << ") want |= 1u << " << dynBit(g) << ";\n";
        }
        for (int i = 0; i < prog_.naggs; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const int slot = ng + i;
            const std::string ref = "s_agg[h * " + std::to_string(nas) + " + " + std::to_string(slot - ng) + "]";
            const std::string value = aggValue(i);
            code << "            if (" << aggValid(i) << ") {\n";
            if (f == "MIN") code << "                atomicMin(&" << ref << ", " << value << ");\n";
            else if (f == "MAX") code << "                atomicMax(&" << ref << ", " << value << ");\n";
            else code << "                atomicAdd(&" << ref << ", " << value << ");\n";
            const int hs = prog_.hidden_slot[static_cast<size_t>(i)];
            if (hs >= 0) code << "                atomicAdd(&s_agg[h * " << nas << " + " << (hs - ng) << "], 1ull);\n";
            const int db = dynBit(slot);
            if (db > 0) code << "                want |= 1u << " << db << ";\n";
            code << "            }\n";
        }
        if (!batch_form && opt_.split_probes)
            code << "            { const u32 sh = (u32)(h & " << (per - 1) << ") * " << W
                 << "u; if (((gw[i] >> sh) & want) != want) atomicOr(&s_flag[h >> " << lp << "], want << sh); }\n"
                 << "        }\n";
        else
            code << "            { u32* fw = &s_flag[h >> " << lp << "]; const u32 sh = (u32)(h & " << (per - 1) << ") * " << W
                 << "u; if (((*fw >> sh) & want) != want) atomicOr(fw, want << sh); }\n"
                 << "        }\n";
        if (batch_form) code << "        }\n";  // round loop
        // merge into the global slots once per block
        if (!batch_form) epi << "    __syncthreads();\n";
        epi << "    for (int g = tid; g < " << G << "; g += BT) {\n"
            << "        const u32 f = (s_flag[g >> " << lp << "] >> ((u32)(g & " << (per - 1) << ") * " << W << "u)) & " << mask
            << ";\n"
            << "        if (!(f & 1u)) continue;\n"
            << "        const long long base = (long long)g * " << sts << ";\n"
            << "        int rem = g;\n";
        for (int g = ng - 1; g >= 0; --g) {
            epi << "        const int d" << g << " = rem % " << gcard_[static_cast<size_t>(g)] << "; rem /= "
                << gcard_[static_cast<size_t>(g)] << ";\n";
        }
        for (int g = 0; g < ng; ++g) {
            const std::string mn = std::to_string(gmin_[static_cast<size_t>(g)]);
            std::string val;
            if (gnullable_[static_cast<size_t>(g)]) val = "(d" + std::to_string(g) + " == 0 ? 0ll : (long long)" + mn + " + d" + std::to_string(g) + " - 1)";
            else val = "((long long)" + mn + " + d" + std::to_string(g) + ")";
            epi << "        P.result[base + " << g << "] = (u64)" << val << ";\n";
        }
        for (int s = ng; s < sts; ++s) {
            std::string f = "SUM";
            for (int i = 0; i < prog_.naggs; ++i)
                if (ng + i == s) f = prog_.agg_funcs[static_cast<size_t>(i)];
            const std::string v = "s_agg[g * " + std::to_string(nas) + " + " + std::to_string(s - ng) + "]";
            if (f == "MIN") epi << "        { const u64 x = " << v << "; if (x != ~0ull) atomicMin(&P.result[base + " << s << "], x); }\n";
            else if (f == "MAX") epi << "        { const u64 x = " << v << "; if (x) atomicMax(&P.result[base + " << s << "], x); }\n";
            else epi << "        { const u64 x = " << v << "; if (x) atomicAdd(&P.result[base + " << s << "], x); }\n";
        }
        if (!prog_.fast_sparse && (!flag_static_.empty() || !flag_dyn_.empty())) {
            // the group's validity bits are contiguous: one or two atomicOr
            unsigned long long smask = 0;
            for (int s : flag_static_) smask |= 1ull << s;
            epi << "        u64 vm = " << smask << "ull;\n";
            for (int s : flag_dyn_) epi << "        vm |= (u64)((f >> " << dynBit(s) << ") & 1u) << " << s << ";\n";
            epi << "        { const long long b0 = base; const u32 o = (u32)(b0 & 63);\n"
                << "          atomicOr(&P.result_valid[b0 >> 6], vm << o);\n"
                << "          if (o + " << sts << "u > 64u) atomicOr(&P.result_valid[(b0 >> 6) + 1], vm >> (64u - o)); }\n";
        }
        epi << "    }\n";
    }

    // Global aggregation can leave the group keys to cq_keys when they
    // decode from the group index and some slot is valid on every hit (that
    // slot's validity bit then marks the group as hit).
    bool keysDecodable() const {
        if (!prog_.grouped || prog_.fast_sparse || flag_w_ == 0 || prog_.storage_ts > 64 || flag_static_.empty()) return false;
        for (bool d : gdense_)
            if (!d) return false;
        return true;
    }

    void emitGlobalAggregate(std::ostringstream& code) {
        const int sts = prog_.storage_ts;
        const int64_t G = prog_.groups;
        prog_.keys_kernel = keysDecodable();
        code << "        // GROUP BY: perfect hash, global u64 atomics per row\n#pragma unroll\n"
             << "        for (int i = 0; i < IPT; ++i) {\n"
             << "            if (!" << flGet() << ") continue;\n"
             << "            const int h = " << groupHash() << ";\n"
             << "            if ((unsigned)h >= " << G << "u) continue;\n"
             << "            const long long out = (long long)h * " << sts << ";\n";
        globalRow(code, "            ");
        code << "        }\n";
    }

    // per-row global aggregation statements (h / out in scope)
    void globalRow(std::ostringstream& code, const std::string& ind) {
        const auto* a = agg_;
        globalRowWith(
            code, ind, [&](int g) { return vexpr(a->group_by_exprs[static_cast<size_t>(g)].get()); },
            [&](int g) { return vvalid(a->group_by_exprs[static_cast<size_t>(g)].get()); },
            [&](int i) { return aggRow(i).valid; }, [&](int i) { return aggRow(i).value; });
    }
    // the same with the row's group key values / validity and aggregate
    // validity / values given as expressions (e.g. picked from item arrays)
    void globalRowWith(std::ostringstream& code, const std::string& ind, const std::function<std::string(int)>& keyValue,
                       const std::function<std::string(int)>& keyValid, const std::function<std::string(int)>& aggValid,
                       const std::function<std::string(int)>& aggValue) {
        const int ng = prog_.ngroups;
        const bool decode = keysDecodable();
        auto setbit = [&](const std::string& idx) {
            return "{ const long long b = " + idx + "; unsigned long long* w = &P.result_valid[b >> 6]; const u64 m = 1ull << (b & 63); if (!(*w & m)) atomicOr(w, m); }";
        };
        auto isStatic = [&](int slot) {
            return std::find(flag_static_.begin(), flag_static_.end(), slot) != flag_static_.end();
        };
        for (int g = 0; g < ng; ++g) {
            if (decode) {
                if (isStatic(g)) continue;  // written by cq_keys
                const std::string v = keyValid(g);
                if (!litFalse(v)) code << ind << "if (" << v << ") " << setbit("out + " + std::to_string(g)) << "\n";
                continue;
            }
            code << ind << "P.result[out + " << g << "] = (u64)(" << keyValue(g) << ");\n";
            if (!prog_.fast_sparse) {
                const std::string v = keyValid(g);
                if (litTrue(v)) code << ind << setbit("out + " + std::to_string(g)) << "\n";
                else if (!litFalse(v)) code << ind << "if (" << v << ") " << setbit("out + " + std::to_string(g)) << "\n";
            }
        }
        for (int i = 0; i < prog_.naggs; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const int slot = ng + i;
            code << ind << "if (" << aggValid(i) << ") {\n";
            const char* atom = f == "MIN" ? "atomicMin" : (f == "MAX" ? "atomicMax" : "atomicAdd");
            code << ind << "    " << atom << "(&P.result[out + " << slot << "], " << aggValue(i) << ");\n";
            const int hs = prog_.hidden_slot[static_cast<size_t>(i)];
            if (hs >= 0) code << ind << "    atomicAdd(&P.result[out + " << hs << "], 1ull);\n";
            else if (f != "COUNT" && !prog_.fast_sparse && !(decode && isStatic(slot)))
                code << ind << "    " << setbit("out + " + std::to_string(slot)) << "\n";
            code << ind << "}\n";
            if (f == "COUNT" && !prog_.fast_sparse && !decode) code << ind << setbit("out + " + std::to_string(slot)) << "\n";
        }
        if (decode) code << ind << setbit("out + " + std::to_string(flag_static_.front())) << "  // hit\n";
    }

    // Rows of one query are sparse after the joins: instead of running a
    // (branchy) per-row aggregate update once per item slot with a few lanes
    // active, every lane queues its live items and each round handles one
    // item per lane (rounds = the most live items any lane holds, usually
    // one).  Emits the queue and opens the round loop with the row's group
    // index in `h` and accessors for its key / aggregate values; the caller
    // closes the loop with "        }\n".
    struct RowQueue {
        std::function<std::string(int)> keyValue, keyValid, aggValid, aggValue;
    };
    RowQueue emitRowQueue(std::ostringstream& code) {
        const auto* a = agg_;
        const int ng = prog_.ngroups;
        const int64_t G = prog_.groups;
        const bool decode = keysDecodable();
        std::vector<int> key_vals, key_valids;  // group columns whose value / validity the row provides
        for (int g = 0; g < ng; ++g) {
            const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
            if (!litTrue(v) && !litFalse(v)) key_valids.push_back(g);
            if (!decode) key_vals.push_back(g);
        }
        code << "        u32 pend = 0u;\n        int gh[IPT];\n";
        for (int g : key_vals) code << "        u64 gkv" << g << "[IPT];\n";
        for (int g : key_valids) code << "        u32 gkok" << g << " = 0u;\n";
        for (int i = 0; i < prog_.naggs; ++i) code << "        u64 gav" << i << "[IPT]; u32 gaok" << i << " = 0u;\n";
        code << "#pragma unroll\n        for (int i = 0; i < IPT; ++i) {\n"
             << "            const int h = " << groupHash() << ";\n"
             << "            gh[i] = h;\n"
             << "            if (" << flGet() << " && (unsigned)h < " << G << "u) pend |= 1u << i;\n";
        for (int g : key_vals) code << "            gkv" << g << "[i] = (u64)(" << vexpr(a->group_by_exprs[static_cast<size_t>(g)].get()) << ");\n";
        for (int g : key_valids)
            code << "            if (" << vvalid(a->group_by_exprs[static_cast<size_t>(g)].get()) << ") gkok" << g << " |= 1u << i;\n";
        for (int i = 0; i < prog_.naggs; ++i) {
            const AggRow r = aggRow(i);
            code << "            gav" << i << "[i] = " << r.value << ";\n";
            if (litTrue(r.valid)) code << "            gaok" << i << " |= 1u << i;\n";
            else if (!litFalse(r.valid)) code << "            if (" << r.valid << ") gaok" << i << " |= 1u << i;\n";
        }
        code << "        }\n"
             << "        while (__ballot(pend != 0u) != 0ull) {\n"
             << "            if (pend == 0u) continue;\n"
             << "            const int it = __builtin_ctz(pend);\n"
             << "            pend &= pend - 1u;\n"
             << "            int h = 0;\n";
        for (int g : key_vals) code << "            u64 kv" << g << " = 0ull;\n";
        for (int i = 0; i < prog_.naggs; ++i) code << "            u64 av" << i << " = 0ull;\n";
        code << "#pragma unroll\n            for (int j = 0; j < IPT; ++j)\n                if (j == it) {\n"
             << "                    h = gh[j];\n";
        for (int g : key_vals) code << "                    kv" << g << " = gkv" << g << "[j];\n";
        for (int i = 0; i < prog_.naggs; ++i) code << "                    av" << i << " = gav" << i << "[j];\n";
        code << "                }\n";
        RowQueue q;
        q.keyValue = [](int g) { return "kv" + std::to_string(g); };
        q.aggValue = [](int i) { return "av" + std::to_string(i); };
        q.keyValid = [this, a](int g) -> std::string {
            const std::string v = vvalid(a->group_by_exprs[static_cast<size_t>(g)].get());
            if (litTrue(v) || litFalse(v)) return v;
            return "((gkok" + std::to_string(g) + " >> it) & 1u)";
        };
        q.aggValid = [this](int i) -> std::string {
            const std::string v = aggRow(i).valid;
            if (litTrue(v) || litFalse(v)) return v;
            return "((gaok" + std::to_string(i) + " >> it) & 1u)";
        };
        return q;
    }

    // Batched queries whose dense domain does not fit their LDS region: an
    // open-addressing table per query in LDS (group index + 1 as key, linear
    // probing up to 8 slots); rows of groups that find no slot take the
    // global path.  The table has 1 << cq_hbits slots, chosen per launch from
    // the LDS left over by the other templates (at least kHashMinBits).  Keys
    // come back from the group index: written by the merge (sparse output)
    // or by cq_keys.
    static constexpr int kHashMinBits = 7;
    bool hashAggEligible() {
        if (!prog_.grouped || prog_.storage_ts > 64) return false;
        for (bool d : gdense_)
            if (!d) return false;
        return prog_.fast_sparse || keysDecodable();
    }
    uint64_t hashSlotBytes() const {
        const uint64_t nas = static_cast<uint64_t>(prog_.storage_ts - prog_.ngroups);
        return 8 + 8 * nas;
    }
    uint64_t hashAggBytes() const { return (uint64_t(1) << kHashMinBits) * hashSlotBytes(); }
    void emitLdsHashAggregate(std::ostringstream& code, std::ostringstream& pro, std::ostringstream& epi) {
        const int ng = prog_.ngroups, sts = prog_.storage_ts, nas = sts - ng;
        prog_.keys_kernel = !prog_.fast_sparse;
        // region: u32 keys[S], u32 flags[S], u64 vals[S * nas]
        std::string init = "0ull";
        if (!prog_.min_slots.empty()) {
            std::string cond;
            for (int s : prog_.min_slots) cond += (cond.empty() ? "" : " || ") + std::string("k % ") + std::to_string(nas) + " == " + std::to_string(s - ng);
            init = "((" + cond + ") ? ~0ull : 0ull)";
        }
        pro << "    for (int k = tid; k < (2 << cq_hbits); k += BT) ((u32*)s_agg)[k] = 0u;\n"
            << "    for (int k = tid; k < (" << nas << " << cq_hbits); k += BT) s_agg[(1 << cq_hbits) + k] = " << init << ";\n";
        code << "        // GROUP BY: LDS hash table of this query, global atomics on overflow\n";
        const RowQueue rq = emitRowQueue(code);
        const auto& keyValidAt = rq.keyValid;
        const auto& aggValidAt = rq.aggValid;
        code << "            u32* s_hk = (u32*)s_agg;\n"
             << "            u32* s_hf = s_hk + (1u << cq_hbits);\n"
             << "            u64* s_hv = s_agg + (1u << cq_hbits);\n"
             << "            u32 slot = ((u32)h * 2654435761u) >> (32u - cq_hbits);\n"
             << "            int found = -1;\n"
             << "            for (int t = 0; t < 8; ++t) {\n"
             << "                u32 cur = s_hk[slot];\n"
             << "                if (cur == 0u) { const u32 old = atomicCAS(&s_hk[slot], 0u, (u32)h + 1u); cur = old == 0u ? (u32)h + 1u : old; }\n"
             << "                if (cur == (u32)h + 1u) { found = (int)slot; break; }\n"
             << "                slot = (slot + 1u) & ((1u << cq_hbits) - 1u);\n"
             << "            }\n"
             << "            if (found >= 0) {\n"
             << "                u32 want = 1u;\n";
        for (int g : flag_dyn_) {
            if (g >= ng || prog_.fast_sparse) continue;
            code << "                if (" << keyValidAt(g) << ") want |= 1u << " << dynBit(g) << ";\n";
        }
        for (int i = 0; i < prog_.naggs; ++i) {
            const std::string& f = prog_.agg_funcs[static_cast<size_t>(i)];
            const std::string ref = "s_hv[found * " + std::to_string(nas) + " + " + std::to_string(i) + "]";
            code << "                if (" << aggValidAt(i) << ") {\n";
            if (f == "MIN") code << "                    atomicMin(&" << ref << ", av" << i << ");\n";
            else if (f == "MAX") code << "                    atomicMax(&" << ref << ", av" << i << ");\n";
            else code << "                    atomicAdd(&" << ref << ", av" << i << ");\n";
            const int hs = prog_.hidden_slot[static_cast<size_t>(i)];
            if (hs >= 0) code << "                    atomicAdd(&s_hv[found * " << nas << " + " << (hs - ng) << "], 1ull);\n";
            const int db = prog_.fast_sparse ? -1 : dynBit(ng + i);
            if (db > 0) code << "                    want |= 1u << " << db << ";\n";
            code << "                }\n";
        }
        code << "                if ((s_hf[found] & want) != want) atomicOr(&s_hf[found], want);\n"
             << "            } else {\n"
             << "                const long long out = (long long)h * " << sts << ";\n";
        globalRowWith(code, "                ", rq.keyValue, keyValidAt, aggValidAt, rq.aggValue);
        code << "            }\n"
             << "        }\n";
        // merge: aggregates (and sparse keys) of every used slot
        epi << "    for (int slot = tid; slot < (1 << cq_hbits); slot += BT) {\n"
            << "        const u32* s_hk = (const u32*)s_agg;\n"
            << "        const u32 key = s_hk[slot];\n"
            << "        if (key == 0u) continue;\n"
            << "        const long long g = (long long)key - 1;\n"
            << "        const u32 f = s_hk[(1 << cq_hbits) + slot];\n"
            << "        const long long base = g * " << sts << ";\n";
        if (prog_.fast_sparse) {
            epi << "        long long rem = g;\n";
            for (int g = ng - 1; g >= 0; --g)
                epi << "        const long long d" << g << " = rem % " << gcard_[static_cast<size_t>(g)] << "; rem /= "
                    << gcard_[static_cast<size_t>(g)] << ";\n";
            for (int g = 0; g < ng; ++g) {
                const std::string mn = std::to_string(gmin_[static_cast<size_t>(g)]);
                std::string val;
                if (gnullable_[static_cast<size_t>(g)]) val = "(d" + std::to_string(g) + " == 0 ? 0ll : (long long)" + mn + " + d" + std::to_string(g) + " - 1)";
                else val = "((long long)" + mn + " + d" + std::to_string(g) + ")";
                epi << "        P.result[base + " << g << "] = (u64)" << val << ";\n";
            }
        }
        for (int s2 = ng; s2 < sts; ++s2) {
            std::string f = "SUM";
            for (int i = 0; i < prog_.naggs; ++i)
                if (ng + i == s2) f = prog_.agg_funcs[static_cast<size_t>(i)];
            const std::string v = "s_agg[(1 << cq_hbits) + slot * " + std::to_string(nas) + " + " + std::to_string(s2 - ng) + "]";
            if (f == "MIN") epi << "        { const u64 x = " << v << "; if (x != ~0ull) atomicMin(&P.result[base + " << s2 << "], x); }\n";
            else if (f == "MAX") epi << "        { const u64 x = " << v << "; if (x) atomicMax(&P.result[base + " << s2 << "], x); }\n";
            else epi << "        { const u64 x = " << v << "; if (x) atomicAdd(&P.result[base + " << s2 << "], x); }\n";
        }
        if (!prog_.fast_sparse) {
            unsigned long long smask = 0;
            for (int sl : flag_static_) smask |= 1ull << sl;
            epi << "        u64 vm = " << smask << "ull;\n";
            for (int sl : flag_dyn_) epi << "        vm |= (u64)((f >> " << dynBit(sl) << ") & 1u) << " << sl << ";\n";
            epi << "        { const u32 o = (u32)(base & 63);\n"
                << "          atomicOr(&P.result_valid[base >> 6], vm << o);\n"
                << "          if (o + " << sts << "u > 64u) atomicOr(&P.result_valid[(base >> 6) + 1], vm >> (64u - o)); }\n";
        } else {
            epi << "        (void)f;\n";
        }
        epi << "    }\n";
    }

    void emitKeysKernel(std::ostringstream& src) {
        emitKeysKernelAs(src, false);
        // batched templates: the same for every query of the template in one
        // launch (blockIdx.y = query, parameter blocks in device memory)
        if (opt_.batch_body) emitKeysKernelAs(src, true);
    }
    void emitKeysKernelAs(std::ostringstream& src, bool batched) {
        const int ng = prog_.ngroups, sts = prog_.storage_ts;
        const int64_t G = prog_.groups;
        unsigned long long smask = 0;
        for (int s : flag_static_) smask |= 1ull << s;
        if (batched)
            src << "extern \"C\" __global__ __launch_bounds__(256) void cq_keys_b(const cq_query_params* __restrict__ QP) {\n"
                << "    const cq_query_params& P = QP[blockIdx.y];\n";
        else
            src << "extern \"C\" __global__ __launch_bounds__(256) void cq_keys(" << paramDecl() << ") {\n" << paramAlias();
        src << "    using namespace chip;\n"
            << "    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;\n"
            << "    if (g >= " << G << "ll) return;\n"
            << "    const long long base = g * " << sts << ";\n"
            << "    if (!chip::bit64(P.result_valid, base + " << flag_static_.front() << ")) return;\n"
            << "    long long rem = g;\n";
        for (int g = ng - 1; g >= 0; --g)
            src << "    const long long d" << g << " = rem % " << gcard_[static_cast<size_t>(g)] << "; rem /= " << gcard_[static_cast<size_t>(g)]
                << ";\n";
        for (int g = 0; g < ng; ++g) {
            const std::string mn = std::to_string(gmin_[static_cast<size_t>(g)]);
            std::string val;
            if (gnullable_[static_cast<size_t>(g)]) val = "(d" + std::to_string(g) + " == 0 ? 0ll : (long long)" + mn + " + d" + std::to_string(g) + " - 1)";
            else val = "((long long)" + mn + " + d" + std::to_string(g) + ")";
            src << "    P.result[base + " << g << "] = (u64)" << val << ";\n";
        }
        src << "    const u64 vm = " << smask << "ull; const u32 o = (u32)(base & 63);\n"
            << "    atomicOr(&P.result_valid[base >> 6], vm << o);\n"
            << "    if (o + " << sts << "u > 64u) atomicOr(&P.result_valid[(base >> 6) + 1], vm >> (64u - o));\n"
            << "}\n\n";
    }

    // HAVING over result slots (aggregateTupleExprValue/Valid, visitor.cpp:2277-2450)
    std::string slotValue(int slot) {
        const LogicalType t = prog_.out[static_cast<size_t>(slot)].type;
        const std::string raw = "R[base + " + std::to_string(slot) + "]";
        if (t == LogicalType::Float64) return "chip::bits_to_double(" + raw + ")";
        if (t == LogicalType::Int64) return "(long long)(" + raw + ")";
        return "(u64)(" + raw + ")";
    }
    std::string hvalue(const ExprNode* e) {
        int slot = -1;
        if (isResultSlotRef(e, slot)) {
            if (slot >= prog_.visible_ts) unsupported("HAVING slot out of range");
            return slotValue(slot);
        }
        switch (e->getType()) {
            case ExprType::LITERAL_INT: return lit(static_cast<const LiteralIntExpr*>(e)->value);
            case ExprType::LITERAL_FLOAT: return flit(static_cast<const LiteralFloatExpr*>(e)->value);
            case ExprType::LITERAL_NULL: return "0";
            case ExprType::STAR: return "1";
            case ExprType::COLUMN_REF: unsupported("HAVING codegen received non-result ColumnRef");
            case ExprType::OP_IS_NULL:
            case ExprType::OP_IS_NOT_NULL: {
                const std::string v = hvalid(static_cast<const BinaryExpr*>(e)->left.get());
                const bool isnull = e->getType() == ExprType::OP_IS_NULL;
                if (litTrue(v)) return isnull ? "0" : "1";
                if (litFalse(v)) return isnull ? "1" : "0";
                return isnull ? "(!(" + v + "))" : "(" + v + ")";
            }
            case ExprType::OP_NOT: return boolNot(hvalue(static_cast<const BinaryExpr*>(e)->left.get()));
            case ExprType::OP_AND:
            case ExprType::OP_OR: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = hvalue(b->left.get());
                const std::string r = hvalue(b->right.get());
                return e->getType() == ExprType::OP_AND ? boolAnd(l, r) : boolOr(l, r);
            }
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                const std::string cv = hvalid(c->condition.get());
                const std::string cond = hvalue(c->condition.get());
                const std::string sel = litTrue(cv) ? "(" + cond + ")" : "((" + cv + ") && (" + cond + "))";
                const std::string t = hvalue(c->then_expr.get());
                const std::string f = hvalue(c->else_expr.get());
                return "((" + sel + ") ? (" + t + ") : (" + f + "))";
            }
            default: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string l = hvalue(b->left.get());
                const std::string r = hvalue(b->right.get());
                return std::string(safeFn(b->op_type)) + "(" + l + ", " + r + ")";
            }
        }
    }
    std::string hvalid(const ExprNode* e) {
        int slot = -1;
        if (isResultSlotRef(e, slot)) return "chip::bit64(RV, base + " + std::to_string(slot) + ")";
        switch (e->getType()) {
            case ExprType::LITERAL_NULL: return "0";
            case ExprType::STAR: case ExprType::LITERAL_INT: case ExprType::LITERAL_FLOAT:
            case ExprType::OP_IS_NULL: case ExprType::OP_IS_NOT_NULL:
                return "1";
            case ExprType::OP_NOT: return hvalid(static_cast<const BinaryExpr*>(e)->left.get());
            case ExprType::OP_AND:
            case ExprType::OP_OR: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                const std::string lv = hvalid(b->left.get()), rv = hvalid(b->right.get());
                const std::string l = hvalue(b->left.get()), r = hvalue(b->right.get());
                if (e->getType() == ExprType::OP_AND)
                    return boolOr(boolOr(andTerms({lv, boolNot(l)}), andTerms({rv, boolNot(r)})), andTerms({lv, rv}));
                return boolOr(boolOr(andTerms({lv, l}), andTerms({rv, r})), andTerms({lv, rv}));
            }
            case ExprType::CASE_WHEN: {
                const auto* c = static_cast<const CaseWhenExpr*>(e);
                const std::string cv = hvalid(c->condition.get());
                const std::string cond = hvalue(c->condition.get());
                const std::string sel = litTrue(cv) ? "(" + cond + ")" : "((" + cv + ") && (" + cond + "))";
                return "(((" + sel + ") && (" + hvalid(c->then_expr.get()) + ")) || (!(" + sel + ") && (" +
                       hvalid(c->else_expr.get()) + ")))";
            }
            case ExprType::COLUMN_REF: unsupported("HAVING validity codegen received non-result ColumnRef");
            default: {
                const auto* b = static_cast<const BinaryExpr*>(e);
                return andTerms({hvalid(b->left.get()), hvalid(b->right.get())});
            }
        }
    }

    std::string havingExpr() {
        std::string having = "1";
        if (agg_->having_predicate) {
            const std::string v = hvalid(agg_->having_predicate.get());
            const std::string x = hvalue(agg_->having_predicate.get());
            having = litTrue(v) ? "(" + x + ")" : (litFalse(v) ? "0" : "((" + v + ") && (" + x + "))");
        }
        return having;
    }

    void emitDenseCount(std::ostringstream& src) {
        emitDenseCountAs(src, false);
        if (opt_.batch_body) emitDenseCountAs(src, true);
    }
    void emitDenseCountAs(std::ostringstream& src, bool batched) {
        const int vts = prog_.visible_ts;
        if (batched)
            src << "extern \"C\" __global__ __launch_bounds__(256) void cq_dense_count_b(const cq_query_params* __restrict__ QP, "
                   "const cq_shape_desc* __restrict__ SD) {\n"
                << "    const cq_query_params& P = QP[blockIdx.y];\n"
                << "    const unsigned long long* __restrict__ R = SD[blockIdx.y].res;\n"
                << "    const unsigned long long* __restrict__ RV = SD[blockIdx.y].rv;\n"
                << "    unsigned long long* counts = SD[blockIdx.y].fl;\n"
                << "    const long long G = SD[blockIdx.y].G;\n";
        else
            src << "extern \"C\" __global__ __launch_bounds__(256) void cq_dense_count(" << paramDecl()
                << ", const unsigned long long* __restrict__ R, const unsigned long long* __restrict__ RV, "
                   "unsigned long long* counts, long long G) {\n"
                << paramAlias();
        src << "    using namespace chip;\n"
            << "    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;\n"
            << "    if (g >= G) return;\n"
            << "    const long long base = g * " << vts << ";\n";
        if (prog_.ngroups == 0) {
            src << "    bool present = true;\n";
        } else {
            src << "    bool present = false;\n"
                << "    for (int c = 0; c < " << vts << " && !present; ++c)\n"
                << "        present = chip::bit64(RV, base + c) != 0u || R[base + c] != 0ull;\n";
        }
        const std::string having = havingExpr();
        src << "    counts[g] = (present && (" << having << ")) ? 1ull : 0ull;\n"
            << "}\n";
    }
};

}  // namespace

HipOperatorVisitor::HipOperatorVisitor(const Catalog& catalog, CodegenOptions options)
    : catalog_(catalog), options_(options) {}

void HipOperatorVisitor::visit(const TableScanNode&) {}
void HipOperatorVisitor::visit(const FilterNode&) {}
void HipOperatorVisitor::visit(const HashJoinNode&) {}
void HipOperatorVisitor::visit(const AggregateNode& node) {
    program_ = Gen(catalog_, options_).run(&node);
    done_ = true;
}
void HipOperatorVisitor::visit(const ProjectionNode& node) {
    program_ = Gen(catalog_, options_).run(&node);
    done_ = true;
}
void HipOperatorVisitor::visit(const SortLimitNode& node) {
    program_ = Gen(catalog_, options_).run(&node);
    done_ = true;
}
QueryProgram HipOperatorVisitor::takeProgram() {
    if (!done_) throw std::runtime_error("HipOperatorVisitor: no program generated");
    done_ = false;
    return std::move(program_);
}

QueryProgram generateQueryProgram(const OperatorNode* root, const Catalog& catalog, const CodegenOptions& options) {
    HipOperatorVisitor v(catalog, options);
    root->accept(v);
    return v.takeProgram();
}

}  // namespace db
