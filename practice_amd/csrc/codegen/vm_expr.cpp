// C expression -> typed register code for the interpreted tier.
#include "vm_expr.h"

#include <cctype>
#include <climits>
#include <cstdlib>
#include <stdexcept>

namespace db {
namespace vm {

namespace {

[[noreturn]] void fail(const std::string& why, const std::string& expr) {
    throw std::runtime_error("interpreted tier: " + why + " in `" + expr + "`");
}

bool isInt(int t) { return t != CQ_T_F64; }
bool isSigned(int t) { return t == CQ_T_I32 || t == CQ_T_I64; }
int rankOf(int t) { return (t == CQ_T_I32 || t == CQ_T_U32) ? 1 : 2; }

// usual arithmetic conversions (after integer promotion; bool is int here)
int arith(int a, int b) {
    if (a == CQ_T_F64 || b == CQ_T_F64) return CQ_T_F64;
    if (a == b) return a;
    if (isSigned(a) == isSigned(b)) return rankOf(a) >= rankOf(b) ? a : b;
    const int u = isSigned(a) ? b : a, s = isSigned(a) ? a : b;
    if (rankOf(u) >= rankOf(s)) return u;
    return s;  // int64 holds every u32
}

struct Tok {
    enum Kind { End, Ident, Num, Punct } kind = End;
    std::string text;
};

class Parser {
public:
    Parser(Builder& b, const std::map<int, int>& payload, std::map<int, int>& cols, std::map<int, int>& vcols,
           const std::string& s)
        : b_(b), payload_(payload), cols_(cols), vcols_(vcols), s_(s) {
        next();
    }
    Val parseAll() {
        Val v = ternary();
        if (tok_.kind != Tok::End) fail("trailing text near `" + tok_.text + "`", s_);
        return v;
    }

private:
    Builder& b_;
    const std::map<int, int>& payload_;
    std::map<int, int>& cols_;
    std::map<int, int>& vcols_;
    const std::string& s_;
    size_t pos_ = 0;
    Tok tok_;

    void next() {
        while (pos_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[pos_]))) ++pos_;
        tok_ = Tok{};
        if (pos_ >= s_.size()) return;
        const char c = s_[pos_];
        if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
            size_t e = pos_;
            while (e < s_.size()) {
                if (std::isalnum(static_cast<unsigned char>(s_[e])) || s_[e] == '_') ++e;
                else if (s_.compare(e, 2, "::") == 0) e += 2;
                else if (s_[e] == '.' && e + 1 < s_.size() && std::isalpha(static_cast<unsigned char>(s_[e + 1]))) ++e;
                else break;
            }
            tok_ = {Tok::Ident, s_.substr(pos_, e - pos_)};
            pos_ = e;
            return;
        }
        if (std::isdigit(static_cast<unsigned char>(c))) {
            size_t e = pos_;
            while (e < s_.size() && std::isalnum(static_cast<unsigned char>(s_[e]))) ++e;
            tok_ = {Tok::Num, s_.substr(pos_, e - pos_)};
            pos_ = e;
            return;
        }
        static const char* two[] = {"&&", "||", "==", "!=", "<=", ">=", ">>", "<<"};
        for (const char* t : two)
            if (s_.compare(pos_, 2, t) == 0) {
                tok_ = {Tok::Punct, t};
                pos_ += 2;
                return;
            }
        tok_ = {Tok::Punct, std::string(1, c)};
        ++pos_;
    }
    bool is(const char* p) const { return tok_.kind == Tok::Punct && tok_.text == p; }
    void expect(const char* p) {
        if (!is(p)) fail(std::string("expected `") + p + "` near `" + tok_.text + "`", s_);
        next();
    }
    long long number() {
        if (tok_.kind != Tok::Num) fail("expected a number near `" + tok_.text + "`", s_);
        const long long v = std::strtoll(tok_.text.c_str(), nullptr, 10);
        next();
        return v;
    }

    Val imm(long long v, int type) {
        Val r;
        r.type = type;
        r.reg = b_.emit(CQ_OP_IMM, type, 0, -1, 0, 0, 0, v);
        return r;
    }
    Val binop(int op, Val a, Val bb, int type) {
        a = b_.toType(a, type);
        bb = b_.toType(bb, type);
        Val r;
        r.type = (op >= CQ_OP_EQ && op <= CQ_OP_GE) ? CQ_T_I32 : type;
        r.reg = b_.emit(op, type, type, -1, a.reg, bb.reg, 0, 0);
        b_.release(a);
        b_.release(bb);
        return r;
    }
    Val logical(int op, Val a, Val bb) {
        Val r;
        r.type = CQ_T_I32;
        r.reg = b_.emit(op, a.type, bb.type, -1, a.reg, bb.reg, 0, 0);
        b_.release(a);
        b_.release(bb);
        return r;
    }

    Val ternary() {
        Val c = lor();
        if (!is("?")) return c;
        next();
        Val t = ternary();
        expect(":");
        Val f = ternary();
        const int type = arith(t.type, f.type);
        t = b_.toType(t, type);
        f = b_.toType(f, type);
        Val r;
        r.type = type;
        r.reg = b_.emit(CQ_OP_SEL, type, c.type, -1, t.reg, f.reg, c.reg, 0);
        b_.release(c);
        b_.release(t);
        b_.release(f);
        return r;
    }
    Val lor() {
        Val a = land();
        while (is("||")) {
            next();
            a = logical(CQ_OP_OR, a, land());
        }
        return a;
    }
    Val land() {
        Val a = band();
        while (is("&&")) {
            next();
            a = logical(CQ_OP_AND, a, band());
        }
        return a;
    }
    Val band() {
        Val a = equality();
        while (is("&")) fail("bitwise & outside a validity test", s_);
        return a;
    }
    Val equality() {
        Val a = relational();
        while (is("==") || is("!=")) {
            const int op = is("==") ? CQ_OP_EQ : CQ_OP_NE;
            next();
            Val b2 = relational();
            a = binop(op, a, b2, arith(a.type, b2.type));
        }
        return a;
    }
    Val relational() {
        Val a = additive();
        while (is("<") || is("<=") || is(">") || is(">=")) {
            const int op = is("<") ? CQ_OP_LT : is("<=") ? CQ_OP_LE : is(">") ? CQ_OP_GT : CQ_OP_GE;
            next();
            Val b2 = additive();
            a = binop(op, a, b2, arith(a.type, b2.type));
        }
        return a;
    }
    Val additive() {
        Val a = multiplicative();
        while (is("+") || is("-")) {
            const int op = is("+") ? CQ_OP_ADD : CQ_OP_SUB;
            next();
            Val b2 = multiplicative();
            a = binop(op, a, b2, arith(a.type, b2.type));
        }
        return a;
    }
    Val multiplicative() {
        Val a = unary();
        while (is("*") || is("/") || is("%")) {
            const int op = is("*") ? CQ_OP_MUL : is("/") ? CQ_OP_DIV : CQ_OP_MOD;
            next();
            Val b2 = unary();
            const int t = arith(a.type, b2.type);
            if (op == CQ_OP_MOD && !isInt(t)) fail("% on double", s_);
            a = binop(op, a, b2, t);
        }
        return a;
    }

    // (type) casts recognised by their keywords
    bool castType(int& type) {
        size_t save = pos_;
        Tok saved = tok_;
        if (!is("(")) return false;
        next();
        std::string name;
        while (tok_.kind == Tok::Ident) {
            name += (name.empty() ? "" : " ") + tok_.text;
            next();
        }
        static const std::map<std::string, int> types = {
            {"int", CQ_T_I32},           {"unsigned", CQ_T_U32},       {"unsigned int", CQ_T_U32}, {"u32", CQ_T_U32},
            {"long long", CQ_T_I64},     {"i64", CQ_T_I64},            {"unsigned long long", CQ_T_U64},
            {"u64", CQ_T_U64},           {"double", CQ_T_F64}};
        auto it = types.find(name);
        if (it != types.end() && is(")")) {
            next();
            type = it->second;
            return true;
        }
        pos_ = save;
        tok_ = saved;
        return false;
    }

    Val unary() {
        if (is("!")) {
            next();
            Val a = unary();
            Val r;
            r.type = CQ_T_I32;
            r.reg = b_.emit(CQ_OP_NOT, a.type, 0, -1, a.reg, 0, 0, 0);
            b_.release(a);
            return r;
        }
        if (is("-")) {
            next();
            Val a = unary();
            const int t = arith(a.type, CQ_T_I32);
            return binop(CQ_OP_SUB, imm(0, CQ_T_I32), a, t);
        }
        int type = 0;
        if (castType(type)) return b_.toType(unary(), type);
        return primary();
    }

    int indexOf(const std::string& ident, const char* prefix) {
        return std::atoi(ident.c_str() + std::string(prefix).size());
    }
    bool identIs(const std::string& id, const char* prefix) {
        const std::string p(prefix);
        if (id.compare(0, p.size(), p) != 0 || id.size() == p.size()) return false;
        for (size_t k = p.size(); k < id.size(); ++k)
            if (!std::isdigit(static_cast<unsigned char>(id[k]))) return false;
        return true;
    }
    void expectIdent(const char* name) {
        if (tok_.kind != Tok::Ident || tok_.text != name) fail(std::string("expected `") + name + "`", s_);
        next();
    }
    // "[base + k]" / "base + k" slot index of HAVING code
    long long slotIndex() {
        expectIdent("base");
        expect("+");
        return number();
    }

    Val primary() {
        if (tok_.kind == Tok::Num) {
            std::string t = tok_.text;
            const long long v = number();
            std::string suffix;
            for (char c : t)
                if (std::isalpha(static_cast<unsigned char>(c))) suffix += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
            int type;
            if (suffix == "ull" || suffix == "llu") type = CQ_T_U64;
            else if (suffix == "ll") type = CQ_T_I64;
            else if (suffix == "u") type = v <= 0xffffffffLL ? CQ_T_U32 : CQ_T_U64;
            else type = v <= INT_MAX ? CQ_T_I32 : CQ_T_I64;
            if (t.size() > 1 && t[0] == '0' && std::isdigit(static_cast<unsigned char>(t[1]))) fail("octal literal", s_);
            if (suffix == "ull" || suffix == "llu") {
                const unsigned long long uv = std::strtoull(t.c_str(), nullptr, 10);
                return imm(static_cast<long long>(uv), type);
            }
            return imm(v, type);
        }
        if (is("(")) {
            next();
            // ((v<N> >> i) & 1u): validity bit of fact column N
            if (tok_.kind == Tok::Ident && identIs(tok_.text, "v")) {
                const int slot = indexOf(tok_.text, "v");
                next();
                expect(">>");
                expectIdent("i");
                expect(")");
                expect("&");
                if (tok_.kind != Tok::Num || (tok_.text != "1u" && tok_.text != "1")) fail("validity mask", s_);
                next();
                Val r;
                r.type = CQ_T_U32;
                auto it = vcols_.find(slot);
                if (it == vcols_.end()) {
                    const int reg = b_.allocPermanent();
                    b_.emit(CQ_OP_VALID, CQ_T_U32, 0, reg, 0, 0, 0, slot);
                    it = vcols_.emplace(slot, reg).first;
                }
                r.reg = it->second;
                return r;
            }
            Val v = ternary();
            expect(")");
            return v;
        }
        if (tok_.kind != Tok::Ident) fail("unexpected `" + tok_.text + "`", s_);
        const std::string id = tok_.text;
        next();
        if (identIs(id, "c")) {  // c<N>[i]
            const int slot = indexOf(id, "c");
            expect("[");
            expectIdent("i");
            expect("]");
            auto it = cols_.find(slot);
            if (it == cols_.end()) {
                const int reg = b_.allocPermanent();
                b_.emit(CQ_OP_COL, CQ_T_I32, 0, reg, 0, 0, 0, slot);
                it = cols_.emplace(slot, reg).first;
            }
            Val r;
            r.reg = it->second;
            r.type = CQ_T_I32;
            return r;
        }
        if (identIs(id, "p")) {  // p<N>[i]: payload register of join N
            const int j = indexOf(id, "p");
            expect("[");
            expectIdent("i");
            expect("]");
            auto it = payload_.find(j);
            if (it == payload_.end()) fail("payload of a join not probed yet", s_);
            Val r;
            r.reg = it->second;
            r.type = CQ_T_I32;
            return r;
        }
        if (id == "P.lit" || id == "P.flit") {
            expect("[");
            const long long k = number();
            expect("]");
            Val r;
            r.type = id == "P.lit" ? CQ_T_I64 : CQ_T_F64;
            r.reg = b_.emit(id == "P.lit" ? CQ_OP_LIT : CQ_OP_FLIT, r.type, 0, -1, 0, 0, 0, k);
            return r;
        }
        if (id == "P.dcol") {  // P.dcol[s][r]
            expect("[");
            const long long sl = number();
            expect("]");
            expect("[");
            expectIdent("r");
            expect("]");
            Val r;
            r.type = CQ_T_I32;
            r.reg = b_.emit(CQ_OP_DCOL, CQ_T_I32, 0, -1, 0, 0, 0, sl);
            return r;
        }
        if (id == "R") {  // R[base + k]: raw result slot
            expect("[");
            const long long k = slotIndex();
            expect("]");
            Val r;
            r.type = CQ_T_U64;
            r.reg = b_.emit(CQ_OP_SLOT, CQ_T_U64, 0, -1, 0, 0, 0, k);
            return r;
        }
        if (id == "chip::bits_to_double") {
            expect("(");
            expectIdent("R");
            expect("[");
            const long long k = slotIndex();
            expect("]");
            expect(")");
            Val r;
            r.type = CQ_T_F64;
            r.reg = b_.emit(CQ_OP_SLOT, CQ_T_F64, 0, -1, 0, 0, 0, k);
            return r;
        }
        if (id == "chip::bit64") {  // chip::bit64(RV, base + k)
            expect("(");
            expectIdent("RV");
            expect(",");
            const long long k = slotIndex();
            expect(")");
            Val r;
            r.type = CQ_T_U32;
            r.reg = b_.emit(CQ_OP_SLOTV, CQ_T_U32, 0, -1, 0, 0, 0, k);
            return r;
        }
        if (id == "chip::bit32") {  // chip::bit32(P.dvalid[s], r)
            expect("(");
            expectIdent("P.dvalid");
            expect("[");
            const long long sl = number();
            expect("]");
            expect(",");
            expectIdent("r");
            expect(")");
            Val r;
            r.type = CQ_T_U32;
            r.reg = b_.emit(CQ_OP_DVALID, CQ_T_U32, 0, -1, 0, 0, 0, sl);
            return r;
        }
        if (id.compare(0, 11, "chip::safe_") == 0 || id == "chip::cdiv") {
            expect("(");
            Val a = ternary();
            expect(",");
            Val b2 = ternary();
            expect(")");
            if (id == "chip::cdiv") return binop(CQ_OP_DIV, a, b2, arith(a.type, b2.type));
            // prom(): int -> int64 (an unsigned int argument has no unique overload)
            auto prom = [&](int t) {
                if (t == CQ_T_U32) fail("ambiguous promotion of an unsigned int", s_);
                return t == CQ_T_I32 ? CQ_T_I64 : t;
            };
            const int t = arith(prom(a.type), prom(b2.type));
            const std::string fn = id.substr(11);
            static const std::map<std::string, int> ops = {{"add", CQ_OP_ADD}, {"sub", CQ_OP_SUB}, {"mul", CQ_OP_MUL},
                                                           {"div", CQ_OP_DIV}, {"eq", CQ_OP_EQ},   {"neq", CQ_OP_NE},
                                                           {"lt", CQ_OP_LT},   {"gt", CQ_OP_GT},   {"lte", CQ_OP_LE},
                                                           {"gte", CQ_OP_GE}};
            auto it = ops.find(fn);
            if (it == ops.end()) fail("unknown helper " + id, s_);
            return binop(it->second, a, b2, t);
        }
        fail("unknown identifier " + id, s_);
    }
};

}  // namespace

int Builder::alloc() {
    for (int r = 0; r < CQ_VM_MAX_REGS; ++r)
        if (!used_[static_cast<size_t>(r)]) {
            used_[static_cast<size_t>(r)] = true;
            max_regs_ = std::max(max_regs_, r + 1);
            return r;
        }
    throw std::runtime_error("interpreted tier: expression needs more than 32 registers");
}

int Builder::allocPermanent() {
    const int r = alloc();
    permanent_[static_cast<size_t>(r)] = true;
    return r;
}

void Builder::release(const Val& v) {
    if (v.reg >= 0 && !permanent_[static_cast<size_t>(v.reg)]) used_[static_cast<size_t>(v.reg)] = false;
}

int Builder::emit(int op, int t, int t2, int d, int a, int b, int c, long long imm) {
    if (code_.size() >= CQ_VM_MAX_CODE) throw std::runtime_error("interpreted tier: program longer than 512 instructions");
    if (d < 0) d = alloc();
    cq_vm_instr in{};
    in.op = static_cast<unsigned char>(op);
    in.t = static_cast<unsigned char>(t);
    in.t2 = static_cast<unsigned char>(t2);
    in.d = static_cast<unsigned char>(d);
    in.a = static_cast<unsigned char>(a);
    in.b = static_cast<unsigned char>(b);
    in.c = static_cast<unsigned char>(c);
    in.imm = static_cast<int>(static_cast<unsigned long long>(imm) & 0xffffffffull);
    in.imm2 = static_cast<unsigned int>(static_cast<unsigned long long>(imm) >> 32);
    code_.push_back(in);
    return d;
}

Val Builder::toType(Val v, int type) {
    if (v.type == type) return v;
    Val r;
    r.type = type;
    r.reg = emit(CQ_OP_CONV, type, v.type, -1, v.reg, 0, 0, 0);
    release(v);
    return r;
}

Val Builder::compile(const std::string& expr) {
    Parser p(*this, payload_, cols_, vcols_, expr);
    return p.parseAll();
}

}  // namespace vm
}  // namespace db
