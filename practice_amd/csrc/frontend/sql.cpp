// SQL subset parser, translator and predicate pushdown.
#include "sql.h"

#include <cctype>
#include <cstdlib>
#include <functional>
#include <set>
#include <sstream>
#include <vector>

namespace db {
namespace {

// ---------------------------------------------------------------- lexer

enum class Tok { End, Ident, Int, Float, Str, Sym };

struct Token {
    Tok kind = Tok::End;
    std::string text;  // identifiers upper-cased in `upper`
    std::string upper;
};

std::vector<Token> lex(const std::string& s) {
    std::vector<Token> out;
    size_t i = 0;
    while (i < s.size()) {
        const char c = s[i];
        if (std::isspace(static_cast<unsigned char>(c))) { ++i; continue; }
        if (c == '-' && i + 1 < s.size() && s[i + 1] == '-') {  // comment
            while (i < s.size() && s[i] != '\n') ++i;
            continue;
        }
        Token t;
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            size_t j = i;
            while (j < s.size() && (std::isalnum(static_cast<unsigned char>(s[j])) || s[j] == '_')) ++j;
            t.kind = Tok::Ident;
            t.text = s.substr(i, j - i);
            i = j;
        } else if (std::isdigit(static_cast<unsigned char>(c)) || (c == '.' && i + 1 < s.size() && std::isdigit(static_cast<unsigned char>(s[i + 1])))) {
            size_t j = i;
            bool is_float = false;
            while (j < s.size() && (std::isdigit(static_cast<unsigned char>(s[j])) || s[j] == '.' || s[j] == 'e' || s[j] == 'E' ||
                                    ((s[j] == '+' || s[j] == '-') && (s[j - 1] == 'e' || s[j - 1] == 'E')))) {
                if (s[j] == '.' || s[j] == 'e' || s[j] == 'E') is_float = true;
                ++j;
            }
            t.kind = is_float ? Tok::Float : Tok::Int;
            t.text = s.substr(i, j - i);
            i = j;
        } else if (c == '\'') {
            size_t j = i + 1;
            while (j < s.size() && s[j] != '\'') ++j;
            t.kind = Tok::Str;
            t.text = s.substr(i + 1, j - i - 1);
            i = j + 1;
        } else {
            static const char* two[] = {"<=", ">=", "<>", "!="};
            t.kind = Tok::Sym;
            bool matched = false;
            for (const char* op : two) {
                if (s.compare(i, 2, op) == 0) {
                    t.text = op;
                    i += 2;
                    matched = true;
                    break;
                }
            }
            if (!matched) t.text = std::string(1, c), ++i;
        }
        t.upper = t.text;
        for (auto& ch : t.upper) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
        out.push_back(std::move(t));
    }
    out.push_back(Token{});
    return out;
}

// ---------------------------------------------------------------- parse tree

struct PExpr {
    enum Kind { Col, Int, Float, Null, Star, Bin, Not, IsNull, IsNotNull, Case, Func } kind = Null;
    std::string table, name;  // Col: table/column; Func: upper-case name
    int64_t ival = 0;
    double fval = 0;
    ExprType op = ExprType::OP_EQ;
    std::vector<std::unique_ptr<PExpr>> args;  // Bin: 2; Not/IsNull: 1; Func: args
    // Case: when/then pairs in args (cond0, then0, cond1, then1, ...), else in else_
    std::unique_ptr<PExpr> else_;
};
using P = std::unique_ptr<PExpr>;

std::string printP(const PExpr* e) {
    std::ostringstream o;
    switch (e->kind) {
        case PExpr::Col: o << e->name; break;
        case PExpr::Int: o << e->ival; break;
        case PExpr::Float: o << e->fval; break;
        case PExpr::Null: o << "NULL"; break;
        case PExpr::Star: o << "*"; break;
        case PExpr::Bin: o << "(" << printP(e->args[0].get()) << " " << exprTypeName(e->op) << " " << printP(e->args[1].get()) << ")"; break;
        case PExpr::Not: o << "NOT(" << printP(e->args[0].get()) << ")"; break;
        case PExpr::IsNull: o << "ISNULL(" << printP(e->args[0].get()) << ")"; break;
        case PExpr::IsNotNull: o << "ISNOTNULL(" << printP(e->args[0].get()) << ")"; break;
        case PExpr::Case:
            o << "CASE";
            for (const auto& a : e->args) o << " " << printP(a.get());
            if (e->else_) o << " ELSE " << printP(e->else_.get());
            break;
        case PExpr::Func:
            o << e->name << "(";
            for (size_t i = 0; i < e->args.size(); ++i) o << (i ? "," : "") << printP(e->args[i].get());
            o << ")";
            break;
    }
    return o.str();
}

struct SelectItem {
    P expr;
    std::string alias;
};
struct OrderItem {
    P expr;
    bool desc = false;
};
struct Select {
    std::vector<SelectItem> items;
    std::vector<std::pair<std::string, std::string>> from;  // table, alias
    P where;
    std::vector<P> group_by;
    P having;
    std::vector<OrderItem> order_by;
    bool has_limit = false;
    int64_t limit = 0;
};

class Parser {
public:
    explicit Parser(const std::string& sql) : toks_(lex(sql)) {}

    Select parseSelect() {
        Select s;
        expectKw("SELECT");
        do {
            SelectItem it;
            it.expr = parseExpr();
            if (acceptKw("AS")) it.alias = expectIdent();
            else if (peek().kind == Tok::Ident && !isKeyword(peek().upper)) it.alias = expectIdent();
            s.items.push_back(std::move(it));
        } while (acceptSym(","));
        expectKw("FROM");
        do {
            std::string t = expectIdent();
            std::string alias;
            if (acceptKw("AS")) alias = expectIdent();
            else if (peek().kind == Tok::Ident && !isKeyword(peek().upper)) alias = expectIdent();
            for (auto& ch : t) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
            for (auto& ch : alias) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
            s.from.emplace_back(t, alias);
        } while (acceptSym(","));
        if (acceptKw("WHERE")) s.where = parseExpr();
        if (acceptKw("GROUP")) {
            expectKw("BY");
            do s.group_by.push_back(parseExpr());
            while (acceptSym(","));
        }
        if (acceptKw("HAVING")) s.having = parseExpr();
        if (acceptKw("ORDER")) {
            expectKw("BY");
            do {
                OrderItem o;
                o.expr = parseExpr();
                if (acceptKw("DESC")) o.desc = true;
                else acceptKw("ASC");
                s.order_by.push_back(std::move(o));
            } while (acceptSym(","));
        }
        if (acceptKw("LIMIT")) {
            if (peek().kind != Tok::Int) fail("LIMIT expects an integer");
            s.has_limit = true;
            s.limit = std::strtoll(next().text.c_str(), nullptr, 10);
            if (s.limit < 0) fail("LIMIT must be non-negative");
        }
        acceptSym(";");
        if (peek().kind != Tok::End) fail("unexpected token '" + peek().text + "'");
        return s;
    }

private:
    std::vector<Token> toks_;
    size_t pos_ = 0;

    [[noreturn]] void fail(const std::string& m) const { throw std::runtime_error("SQL parse error: " + m); }
    const Token& peek(size_t k = 0) const { return toks_[std::min(pos_ + k, toks_.size() - 1)]; }
    const Token& next() { return toks_[pos_ < toks_.size() - 1 ? pos_++ : pos_]; }
    static bool isKeyword(const std::string& u) {
        static const std::set<std::string> kw = {"SELECT", "FROM", "WHERE", "GROUP", "BY", "HAVING", "ORDER", "LIMIT",
                                                 "AND", "OR", "NOT", "AS", "ASC", "DESC", "IS", "NULL", "CASE", "WHEN",
                                                 "THEN", "ELSE", "END", "BETWEEN", "IN", "LIKE", "ON", "JOIN"};
        return kw.count(u) > 0;
    }
    bool acceptKw(const char* kw) {
        if (peek().kind == Tok::Ident && peek().upper == kw) { ++pos_; return true; }
        return false;
    }
    void expectKw(const char* kw) {
        if (!acceptKw(kw)) fail(std::string("expected ") + kw + " near '" + peek().text + "'");
    }
    bool acceptSym(const char* s) {
        if (peek().kind == Tok::Sym && peek().text == s) { ++pos_; return true; }
        return false;
    }
    void expectSym(const char* s) {
        if (!acceptSym(s)) fail(std::string("expected '") + s + "' near '" + peek().text + "'");
    }
    std::string expectIdent() {
        if (peek().kind != Tok::Ident) fail("expected identifier near '" + peek().text + "'");
        return next().text;
    }

    static P bin(ExprType op, P a, P b) {
        auto e = std::make_unique<PExpr>();
        e->kind = PExpr::Bin;
        e->op = op;
        e->args.push_back(std::move(a));
        e->args.push_back(std::move(b));
        return e;
    }
    static P unary(PExpr::Kind k, P a) {
        auto e = std::make_unique<PExpr>();
        e->kind = k;
        e->args.push_back(std::move(a));
        return e;
    }

    P parseExpr() { return parseOr(); }
    P parseOr() {
        P l = parseAnd();
        while (acceptKw("OR")) l = bin(ExprType::OP_OR, std::move(l), parseAnd());
        return l;
    }
    P parseAnd() {
        P l = parseNot();
        while (acceptKw("AND")) l = bin(ExprType::OP_AND, std::move(l), parseNot());
        return l;
    }
    P parseNot() {
        if (acceptKw("NOT")) {
            P a = parseNot();
            return unary(PExpr::Not, std::move(a));
        }
        return parseCmp();
    }
    P parseCmp() {
        P l = parseAdd();
        if (acceptKw("IS")) {
            const bool neg = acceptKw("NOT");
            expectKw("NULL");
            return unary(neg ? PExpr::IsNotNull : PExpr::IsNull, std::move(l));
        }
        if (peek().kind == Tok::Ident && (peek().upper == "BETWEEN" || (peek().upper == "NOT" && peek(1).upper == "BETWEEN")))
            fail("BETWEEN is not supported. Rewrite as col >= low AND col <= high.");
        if (peek().kind == Tok::Sym) {
            const std::string s = peek().text;
            ExprType op;
            if (s == "=") op = ExprType::OP_EQ;
            else if (s == "<>" || s == "!=") op = ExprType::OP_NEQ;
            else if (s == "<") op = ExprType::OP_LT;
            else if (s == "<=") op = ExprType::OP_LTE;
            else if (s == ">") op = ExprType::OP_GT;
            else if (s == ">=") op = ExprType::OP_GTE;
            else return l;
            ++pos_;
            return bin(op, std::move(l), parseAdd());
        }
        return l;
    }
    P parseAdd() {
        P l = parseMul();
        for (;;) {
            if (acceptSym("+")) l = bin(ExprType::OP_ADD, std::move(l), parseMul());
            else if (acceptSym("-")) l = bin(ExprType::OP_SUB, std::move(l), parseMul());
            else return l;
        }
    }
    P parseMul() {
        P l = parseUnary();
        for (;;) {
            if (acceptSym("*")) l = bin(ExprType::OP_MUL, std::move(l), parseUnary());
            else if (acceptSym("/")) l = bin(ExprType::OP_DIV, std::move(l), parseUnary());
            else return l;
        }
    }
    P parseUnary() {
        if (acceptSym("-")) {
            P a = parseUnary();
            if (a->kind == PExpr::Int) { a->ival = -a->ival; return a; }
            if (a->kind == PExpr::Float) { a->fval = -a->fval; return a; }
            fail("unary minus is only supported on numeric literals");
        }
        if (acceptSym("+")) return parseUnary();
        return parsePrimary();
    }
    P parsePrimary() {
        const Token& t = peek();
        auto e = std::make_unique<PExpr>();
        if (t.kind == Tok::Int) {
            e->kind = PExpr::Int;
            e->ival = std::strtoll(next().text.c_str(), nullptr, 10);
            return e;
        }
        if (t.kind == Tok::Float) {
            e->kind = PExpr::Float;
            e->fval = std::strtod(next().text.c_str(), nullptr);
            return e;
        }
        if (t.kind == Tok::Str) fail("string literals are not supported (columns are integer encoded)");
        if (acceptSym("*")) {
            e->kind = PExpr::Star;
            return e;
        }
        if (acceptSym("(")) {
            P inner = parseExpr();
            expectSym(")");
            return inner;
        }
        if (t.kind != Tok::Ident) fail("unexpected token '" + t.text + "'");
        if (acceptKw("NULL")) {
            e->kind = PExpr::Null;
            return e;
        }
        if (acceptKw("CASE")) return parseCase();
        std::string name = next().text;
        if (acceptSym("(")) {
            e->kind = PExpr::Func;
            e->name = name;
            for (auto& ch : e->name) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
            if (!acceptSym(")")) {
                do e->args.push_back(parseExpr());
                while (acceptSym(","));
                expectSym(")");
            }
            return e;
        }
        e->kind = PExpr::Col;
        if (acceptSym(".")) {
            e->table = name;
            for (auto& ch : e->table) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
            e->name = expectIdent();
        } else {
            e->name = name;
        }
        return e;
    }
    P parseCase() {
        auto e = std::make_unique<PExpr>();
        e->kind = PExpr::Case;
        P subject;
        if (!(peek().kind == Tok::Ident && peek().upper == "WHEN")) subject = parseExpr();
        while (acceptKw("WHEN")) {
            P cond = parseExpr();
            if (subject) {
                // simple CASE -> searched comparison
                auto copy = clonePtr(subject.get());
                cond = bin(ExprType::OP_EQ, std::move(copy), std::move(cond));
            }
            expectKw("THEN");
            e->args.push_back(std::move(cond));
            e->args.push_back(parseExpr());
        }
        if (e->args.empty()) fail("CASE without WHEN");
        if (acceptKw("ELSE")) e->else_ = parseExpr();
        expectKw("END");
        return e;
    }
    static P clonePtr(const PExpr* s) {
        auto e = std::make_unique<PExpr>();
        e->kind = s->kind;
        e->table = s->table;
        e->name = s->name;
        e->ival = s->ival;
        e->fval = s->fval;
        e->op = s->op;
        for (const auto& a : s->args) e->args.push_back(clonePtr(a.get()));
        if (s->else_) e->else_ = clonePtr(s->else_.get());
        return e;
    }
};

bool isAggName(const std::string& u) { return u == "COUNT" || u == "SUM" || u == "MIN" || u == "MAX" || u == "AVG"; }

// ---------------------------------------------------------------- translate

class Translator {
public:
    explicit Translator(Select& s) : s_(s) {}

    std::unique_ptr<ExprNode> expr(const PExpr* e) {
        switch (e->kind) {
            case PExpr::Col: return std::make_unique<ColumnRefExpr>(e->name, resolveTable(e->table));
            case PExpr::Int: return std::make_unique<LiteralIntExpr>(e->ival);
            case PExpr::Float: return std::make_unique<LiteralFloatExpr>(e->fval);
            case PExpr::Null: return std::make_unique<LiteralNullExpr>();
            case PExpr::Star: return std::make_unique<StarExpr>();
            case PExpr::Bin: return std::make_unique<BinaryExpr>(e->op, expr(e->args[0].get()), expr(e->args[1].get()));
            case PExpr::Not: return std::make_unique<BinaryExpr>(ExprType::OP_NOT, expr(e->args[0].get()), nullptr);
            case PExpr::IsNull: return std::make_unique<BinaryExpr>(ExprType::OP_IS_NULL, expr(e->args[0].get()), nullptr);
            case PExpr::IsNotNull:
                return std::make_unique<BinaryExpr>(ExprType::OP_IS_NOT_NULL, expr(e->args[0].get()), nullptr);
            case PExpr::Case: return caseExpr(e, 0);
            case PExpr::Func:
                if (isAggName(e->name)) throw std::runtime_error("QueryTranslator: aggregate function in an unsupported position");
                throw std::runtime_error("QueryTranslator: unsupported scalar function reference: " + e->name);
        }
        throw std::runtime_error("QueryTranslator: unsupported expression");
    }

    std::unique_ptr<OperatorNode> translate() {
        std::unique_ptr<OperatorNode> root;
        for (size_t i = 0; i < s_.from.size(); ++i) {
            auto scan = std::make_unique<TableScanNode>(s_.from[i].first, s_.from[i].second);
            if (!root) {
                root = std::move(scan);
            } else {
                auto join = std::make_unique<HashJoinNode>(nullptr);
                join->addChild(std::move(scan));  // build side
                join->addChild(std::move(root));  // probe side
                root = std::move(join);
            }
        }
        if (s_.where) {
            auto f = std::make_unique<FilterNode>(expr(s_.where.get()));
            f->addChild(std::move(root));
            root = std::move(f);
        }
        bool has_agg = !s_.group_by.empty();
        for (const auto& it : s_.items)
            if (it.expr->kind == PExpr::Func && isAggName(it.expr->name)) has_agg = true;
        if (has_agg) {
            auto agg = std::make_unique<AggregateNode>();
            for (const auto& g : s_.group_by) agg->group_by_exprs.push_back(groupExpr(g.get()));
            for (const auto& it : s_.items) {
                if (!(it.expr->kind == PExpr::Func && isAggName(it.expr->name))) continue;
                if (it.expr->args.size() != 1)
                    throw std::runtime_error("QueryTranslator: aggregate function " + it.expr->name +
                                             " requires exactly one SQL argument; use COUNT(*) for row count");
                agg->aggregates.emplace_back(it.expr->name, expr(it.expr->args[0].get()));
            }
            agg->output_aliases.resize(agg->visibleTupleSize());
            for (size_t i = 0; i < s_.items.size(); ++i) {
                const int slot = selectSlot(i, *agg);
                if (slot >= 0 && !s_.items[i].alias.empty()) agg->output_aliases[static_cast<size_t>(slot)] = s_.items[i].alias;
            }
            if (s_.having) agg->having_predicate = havingExpr(s_.having.get(), *agg);
            agg_ = agg.get();
            agg->addChild(std::move(root));
            root = std::move(agg);
        } else {
            auto proj = std::make_unique<ProjectionNode>();
            for (const auto& it : s_.items) {
                proj->select_exprs.push_back(expr(it.expr.get()));
                proj->output_aliases.push_back(it.alias);
            }
            proj->addChild(std::move(root));
            root = std::move(proj);
        }
        if (!s_.order_by.empty() || s_.has_limit) {
            auto sl = std::make_unique<SortLimitNode>();
            sl->has_limit = s_.has_limit;
            sl->limit = static_cast<size_t>(s_.limit);
            for (const auto& o : s_.order_by) {
                SortKeyDef k;
                k.column_index = static_cast<size_t>(orderSlot(o.expr.get()));
                k.descending = o.desc;
                sl->sort_keys.push_back(k);
            }
            sl->addChild(std::move(root));
            root = std::move(sl);
        }
        return root;
    }

private:
    Select& s_;
    const AggregateNode* agg_ = nullptr;

    std::string resolveTable(const std::string& qual) const {
        if (qual.empty()) return "";
        for (const auto& f : s_.from)
            if (f.first == qual || f.second == qual) return f.first;
        return qual;
    }

    std::unique_ptr<ExprNode> groupExpr(const PExpr* g) {
        // GROUP BY <alias> refers to a select item
        if (g->kind == PExpr::Col && g->table.empty()) {
            for (const auto& it : s_.items)
                if (!it.alias.empty() && it.alias == g->name && !(it.expr->kind == PExpr::Func)) return expr(it.expr.get());
        }
        return expr(g);
    }

    std::unique_ptr<ExprNode> caseExpr(const PExpr* e, size_t arm) {
        auto cond = expr(e->args[arm].get());
        auto then_v = expr(e->args[arm + 1].get());
        std::unique_ptr<ExprNode> else_v;
        if (arm + 2 < e->args.size()) else_v = caseExpr(e, arm + 2);
        else if (e->else_) else_v = expr(e->else_.get());
        else else_v = std::make_unique<LiteralNullExpr>();
        return std::make_unique<CaseWhenExpr>(std::move(cond), std::move(then_v), std::move(else_v));
    }

    // output slot of select item i in the aggregate layout
    int selectSlot(size_t i, const AggregateNode& agg) const {
        const PExpr* e = s_.items[i].expr.get();
        if (e->kind == PExpr::Func && isAggName(e->name)) {
            int k = 0;
            for (size_t j = 0; j < i; ++j)
                if (s_.items[j].expr->kind == PExpr::Func && isAggName(s_.items[j].expr->name)) ++k;
            return static_cast<int>(agg.group_by_exprs.size()) + k;
        }
        const std::string key = printP(e);
        for (size_t g = 0; g < s_.group_by.size(); ++g) {
            const PExpr* ge = s_.group_by[g].get();
            if (printP(ge) == key) return static_cast<int>(g);
            if (ge->kind == PExpr::Col && ge->name == s_.items[i].alias) return static_cast<int>(g);
        }
        return -1;
    }

    int orderSlot(const PExpr* e) const {
        const AggregateNode* agg = agg_;
        if (e->kind == PExpr::Int) {
            const int64_t k = e->ival;
            if (k < 1 || static_cast<size_t>(k) > s_.items.size()) throw std::runtime_error("ORDER BY position out of range");
            if (!agg) return static_cast<int>(k - 1);
            const int slot = selectSlot(static_cast<size_t>(k - 1), *agg);
            if (slot < 0) throw std::runtime_error("ORDER BY position does not name an output column");
            return slot;
        }
        if (e->kind == PExpr::Col && e->table.empty()) {
            for (size_t i = 0; i < s_.items.size(); ++i)
                if (!s_.items[i].alias.empty() && s_.items[i].alias == e->name)
                    return agg ? selectSlot(i, *agg) : static_cast<int>(i);
        }
        const std::string key = printP(e);
        if (agg) {
            for (size_t g = 0; g < s_.group_by.size(); ++g)
                if (printP(s_.group_by[g].get()) == key) return static_cast<int>(g);
            for (size_t i = 0; i < s_.items.size(); ++i)
                if (printP(s_.items[i].expr.get()) == key) {
                    const int slot = selectSlot(i, *agg);
                    if (slot >= 0) return slot;
                }
        } else {
            for (size_t i = 0; i < s_.items.size(); ++i)
                if (printP(s_.items[i].expr.get()) == key) return static_cast<int>(i);
        }
        throw std::runtime_error("ORDER BY expression is not an output column: " + key);
    }

    std::unique_ptr<ExprNode> slotRef(int slot) const {
        return std::make_unique<ColumnRefExpr>("__result_col_" + std::to_string(slot), "__RESULT__");
    }

    // HAVING over visible result slots (aggregate calls, group columns, aliases)
    std::unique_ptr<ExprNode> havingExpr(const PExpr* e, const AggregateNode& agg) {
        if (e->kind == PExpr::Func && isAggName(e->name)) {
            const std::string key = printP(e);
            for (size_t i = 0; i < s_.items.size(); ++i)
                if (printP(s_.items[i].expr.get()) == key) return slotRef(selectSlot(i, agg));
            throw std::runtime_error("HAVING aggregate must also appear in the SELECT list: " + key);
        }
        if (e->kind == PExpr::Col) {
            for (size_t i = 0; i < s_.items.size(); ++i)
                if (!s_.items[i].alias.empty() && s_.items[i].alias == e->name) return slotRef(selectSlot(i, agg));
            for (size_t g = 0; g < s_.group_by.size(); ++g)
                if (printP(s_.group_by[g].get()) == printP(e)) return slotRef(static_cast<int>(g));
            throw std::runtime_error("HAVING column must be a GROUP BY column or an output alias: " + e->name);
        }
        switch (e->kind) {
            case PExpr::Bin:
                return std::make_unique<BinaryExpr>(e->op, havingExpr(e->args[0].get(), agg), havingExpr(e->args[1].get(), agg));
            case PExpr::Not: return std::make_unique<BinaryExpr>(ExprType::OP_NOT, havingExpr(e->args[0].get(), agg), nullptr);
            case PExpr::IsNull: return std::make_unique<BinaryExpr>(ExprType::OP_IS_NULL, havingExpr(e->args[0].get(), agg), nullptr);
            case PExpr::IsNotNull:
                return std::make_unique<BinaryExpr>(ExprType::OP_IS_NOT_NULL, havingExpr(e->args[0].get(), agg), nullptr);
            case PExpr::Case: {
                std::function<std::unique_ptr<ExprNode>(size_t)> arm = [&](size_t a) -> std::unique_ptr<ExprNode> {
                    auto c = havingExpr(e->args[a].get(), agg);
                    auto t = havingExpr(e->args[a + 1].get(), agg);
                    std::unique_ptr<ExprNode> el;
                    if (a + 2 < e->args.size()) el = arm(a + 2);
                    else if (e->else_) el = havingExpr(e->else_.get(), agg);
                    else el = std::make_unique<LiteralNullExpr>();
                    return std::make_unique<CaseWhenExpr>(std::move(c), std::move(t), std::move(el));
                };
                return arm(0);
            }
            default: return expr(e);
        }
    }
};

// ---------------------------------------------------------------- pushdown

std::string columnTable(const ColumnRefExpr* c, const Catalog* cat) {
    if (!c->table_name.empty()) return c->table_name;
    return cat ? cat->tableOfColumn(c->column_name) : Catalog::prefixTable(c->column_name);
}

void exprTables(const ExprNode* e, const Catalog* cat, std::set<std::string>& out) {
    if (!e) return;
    switch (e->getType()) {
        case ExprType::COLUMN_REF: {
            const std::string t = columnTable(static_cast<const ColumnRefExpr*>(e), cat);
            if (!t.empty()) out.insert(t);
            return;
        }
        case ExprType::CASE_WHEN: {
            const auto* c = static_cast<const CaseWhenExpr*>(e);
            exprTables(c->condition.get(), cat, out);
            exprTables(c->then_expr.get(), cat, out);
            exprTables(c->else_expr.get(), cat, out);
            return;
        }
        default:
            if (isBinaryOp(e->getType())) {
                const auto* b = static_cast<const BinaryExpr*>(e);
                exprTables(b->left.get(), cat, out);
                exprTables(b->right.get(), cat, out);
            }
    }
}

void planTables(const OperatorNode* n, std::set<std::string>& out) {
    if (n->getType() == OperatorType::TABLE_SCAN) out.insert(static_cast<const TableScanNode*>(n)->table_name);
    for (const auto& c : n->getChildren()) planTables(c.get(), out);
}

void flattenAnd(std::unique_ptr<ExprNode>& e, std::vector<std::unique_ptr<ExprNode>>& out) {
    if (!e) return;
    if (e->getType() == ExprType::OP_AND) {
        auto* b = static_cast<BinaryExpr*>(e.get());
        flattenAnd(b->left, out);
        flattenAnd(b->right, out);
        e.reset();
        return;
    }
    out.push_back(std::move(e));
}

std::unique_ptr<ExprNode> andChain(std::vector<std::unique_ptr<ExprNode>>& v) {
    auto r = std::move(v[0]);
    for (size_t i = 1; i < v.size(); ++i) r = std::make_unique<BinaryExpr>(ExprType::OP_AND, std::move(r), std::move(v[i]));
    return r;
}

bool keyLike(const std::string& c) {
    return c.find("key") != std::string::npos || c.find("Key") != std::string::npos || c == "lo_orderdate" ||
           c == "lo_commitdate";
}

bool hashJoinPredicate(const ExprNode* e, const Catalog* cat) {
    if (!e) return false;
    if (e->getType() == ExprType::OP_EQ) {
        const auto* b = static_cast<const BinaryExpr*>(e);
        if (!b->left || !b->right || b->left->getType() != ExprType::COLUMN_REF || b->right->getType() != ExprType::COLUMN_REF)
            return false;
        const auto* l = static_cast<const ColumnRefExpr*>(b->left.get());
        const auto* r = static_cast<const ColumnRefExpr*>(b->right.get());
        const std::string lt = columnTable(l, cat), rt = columnTable(r, cat);
        if (lt.empty() || rt.empty() || lt == rt) return false;
        return keyLike(l->column_name) || keyLike(r->column_name);
    }
    if (e->getType() == ExprType::OP_OR) {
        const auto* b = static_cast<const BinaryExpr*>(e);
        return hashJoinPredicate(b->left.get(), cat) && hashJoinPredicate(b->right.get(), cat);
    }
    return false;
}

bool allIn(const std::set<std::string>& a, const std::set<std::string>& b) {
    for (const auto& x : a)
        if (!b.count(x)) return false;
    return true;
}

void pushdown(std::unique_ptr<OperatorNode>& node, const Catalog* cat);

void tryPushdown(std::unique_ptr<OperatorNode>& node, const Catalog* cat) {
    if (!node || node->getType() != OperatorType::FILTER) return;
    auto* filter = static_cast<FilterNode*>(node.get());
    if (filter->getChildren().empty() || !filter->predicate) return;
    auto& child = filter->getChildrenMutable()[0];
    if (child->getType() != OperatorType::HASH_JOIN) return;
    auto* join = static_cast<HashJoinNode*>(child.get());
    auto& jc = join->getChildrenMutable();
    if (jc.size() < 2) return;
    std::set<std::string> lt, rt;
    planTables(jc[0].get(), lt);
    planTables(jc[1].get(), rt);

    std::vector<std::unique_ptr<ExprNode>> conds, left, right, rest;
    flattenAnd(filter->predicate, conds);
    for (auto& c : conds) {
        std::set<std::string> t;
        exprTables(c.get(), cat, t);
        if (!t.empty() && allIn(t, lt)) left.push_back(std::move(c));
        else if (!t.empty() && allIn(t, rt)) right.push_back(std::move(c));
        else rest.push_back(std::move(c));
    }
    std::vector<std::unique_ptr<ExprNode>> remaining;
    for (auto& c : rest) {
        std::set<std::string> t;
        exprTables(c.get(), cat, t);
        bool tl = false, tr = false;
        for (const auto& x : t) {
            tl = tl || lt.count(x);
            tr = tr || rt.count(x);
        }
        if (tl && tr && hashJoinPredicate(c.get(), cat)) {
            if (join->join_condition)
                join->join_condition = std::make_unique<BinaryExpr>(ExprType::OP_AND, std::move(join->join_condition), std::move(c));
            else
                join->join_condition = std::move(c);
        } else {
            remaining.push_back(std::move(c));
        }
    }
    if (left.empty() && right.empty() && !join->join_condition) {
        filter->predicate = andChain(remaining);
        return;
    }
    if (!left.empty()) {
        auto f = std::make_unique<FilterNode>(andChain(left));
        f->addChild(std::move(jc[0]));
        jc[0] = std::move(f);
        pushdown(jc[0], cat);
    }
    if (!right.empty()) {
        auto f = std::make_unique<FilterNode>(andChain(right));
        f->addChild(std::move(jc[1]));
        jc[1] = std::move(f);
        pushdown(jc[1], cat);
    }
    if (remaining.empty()) node = std::move(child);
    else filter->predicate = andChain(remaining);
}

void pushdown(std::unique_ptr<OperatorNode>& node, const Catalog* cat) {
    if (!node) return;
    for (auto& c : node->getChildrenMutable()) pushdown(c, cat);
    tryPushdown(node, cat);
}

}  // namespace

std::unique_ptr<OperatorNode> translateSql(const std::string& sql) {
    if (sql.find_first_not_of(" \t\r\n;") == std::string::npos) throw std::runtime_error("Empty query");
    Parser p(sql);
    Select s = p.parseSelect();
    Translator t(s);
    auto root = t.translate();
    repairParents(root.get());
    return root;
}

std::unique_ptr<OperatorNode> Optimizer::optimize(std::unique_ptr<OperatorNode> plan) const {
    pushdown(plan, catalog_);
    repairParents(plan.get());
    return plan;
}

std::unique_ptr<OperatorNode> buildOptimizedTree(const std::string& sql, const Catalog* catalog) {
    return Optimizer(catalog).optimize(translateSql(sql));
}

}  // namespace db
