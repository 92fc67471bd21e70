// Expression tree of the optimized plan.
//
// Same node kinds, class names and public fields as the reference plan API
// (include/core/expressions.h:13-218) so that a front end written against the
// reference builds the same trees here.  Implementation is this project's.
#pragma once

#include <cstdint>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>

namespace db {

enum class ExprType {
    COLUMN_REF,
    LITERAL_INT,
    LITERAL_FLOAT,
    LITERAL_NULL,
    OP_AND,
    OP_OR,
    OP_NOT,
    OP_EQ,
    OP_NEQ,
    OP_LT,
    OP_LTE,
    OP_GT,
    OP_GTE,
    OP_ADD,
    OP_SUB,
    OP_MUL,
    OP_DIV,
    OP_IS_NULL,
    OP_IS_NOT_NULL,
    CASE_WHEN,
    STAR,
};

class ExprVisitor;

class ExprNode {
public:
    virtual ~ExprNode() = default;
    virtual ExprType getType() const = 0;
    virtual void accept(ExprVisitor& visitor) const = 0;
    virtual std::unique_ptr<ExprNode> clone() const = 0;
};

class ColumnRefExpr final : public ExprNode {
public:
    std::string table_name;   // may be empty
    std::string column_name;  // always set
    explicit ColumnRefExpr(std::string column, std::string table = "")
        : table_name(std::move(table)), column_name(std::move(column)) {}
    ExprType getType() const override { return ExprType::COLUMN_REF; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override {
        return std::make_unique<ColumnRefExpr>(column_name, table_name);
    }
};

class LiteralIntExpr final : public ExprNode {
public:
    int64_t value;
    explicit LiteralIntExpr(int64_t v) : value(v) {}
    ExprType getType() const override { return ExprType::LITERAL_INT; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override { return std::make_unique<LiteralIntExpr>(value); }
};

class LiteralFloatExpr final : public ExprNode {
public:
    double value;
    explicit LiteralFloatExpr(double v) : value(v) {}
    ExprType getType() const override { return ExprType::LITERAL_FLOAT; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override { return std::make_unique<LiteralFloatExpr>(value); }
};

class LiteralNullExpr final : public ExprNode {
public:
    ExprType getType() const override { return ExprType::LITERAL_NULL; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override { return std::make_unique<LiteralNullExpr>(); }
};

// Binary (and, with right == nullptr, unary NOT / IS [NOT] NULL) operator.
class BinaryExpr final : public ExprNode {
public:
    ExprType op_type;
    std::unique_ptr<ExprNode> left;
    std::unique_ptr<ExprNode> right;
    BinaryExpr(ExprType op, std::unique_ptr<ExprNode> l, std::unique_ptr<ExprNode> r)
        : op_type(op), left(std::move(l)), right(std::move(r)) {}
    ExprType getType() const override { return op_type; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override {
        return std::make_unique<BinaryExpr>(op_type, left ? left->clone() : nullptr,
                                            right ? right->clone() : nullptr);
    }
};

// Searched CASE with one WHEN arm; more arms nest in else_expr.
class CaseWhenExpr final : public ExprNode {
public:
    std::unique_ptr<ExprNode> condition;
    std::unique_ptr<ExprNode> then_expr;
    std::unique_ptr<ExprNode> else_expr;
    CaseWhenExpr(std::unique_ptr<ExprNode> c, std::unique_ptr<ExprNode> t, std::unique_ptr<ExprNode> e)
        : condition(std::move(c)), then_expr(std::move(t)), else_expr(std::move(e)) {
        if (!condition || !then_expr || !else_expr)
            throw std::runtime_error("CASE WHEN requires condition, THEN expression, and ELSE expression");
    }
    ExprType getType() const override { return ExprType::CASE_WHEN; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override {
        return std::make_unique<CaseWhenExpr>(condition->clone(), then_expr->clone(), else_expr->clone());
    }
};

class StarExpr final : public ExprNode {
public:
    ExprType getType() const override { return ExprType::STAR; }
    void accept(ExprVisitor& visitor) const override;
    std::unique_ptr<ExprNode> clone() const override { return std::make_unique<StarExpr>(); }
};

class ExprVisitor {
public:
    virtual ~ExprVisitor() = default;
    virtual void visit(const ColumnRefExpr& node) = 0;
    virtual void visit(const LiteralIntExpr& node) = 0;
    virtual void visit(const LiteralFloatExpr& node) = 0;
    virtual void visit(const LiteralNullExpr& node) = 0;
    virtual void visit(const BinaryExpr& node) = 0;
    virtual void visit(const CaseWhenExpr& node) = 0;
    virtual void visit(const StarExpr& node) = 0;
};

inline void ColumnRefExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void LiteralIntExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void LiteralFloatExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void LiteralNullExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void BinaryExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void CaseWhenExpr::accept(ExprVisitor& v) const { v.visit(*this); }
inline void StarExpr::accept(ExprVisitor& v) const { v.visit(*this); }

inline bool isComparison(ExprType t) { return t >= ExprType::OP_EQ && t <= ExprType::OP_GTE; }
inline bool isArithmetic(ExprType t) { return t >= ExprType::OP_ADD && t <= ExprType::OP_DIV; }
inline bool isBinaryOp(ExprType t) { return t >= ExprType::OP_AND && t <= ExprType::OP_IS_NOT_NULL; }

const char* exprTypeName(ExprType t);

// Prints the expression as an S-expression, the wire form read by the oracle.
void printExprSexpr(std::ostream& out, const ExprNode* e);

// Result slot reference inside HAVING: ColumnRefExpr("__result_col_<k>", "__RESULT__").
bool isResultSlotRef(const ExprNode* e, int& slot);

}  // namespace db
