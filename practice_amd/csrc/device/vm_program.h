// Bytecode of the interpreted tier: a plan the code generator accepts,
// compiled to register code that ahead-of-time kernels execute (no hipRTC).
// Shared by the host compiler (codegen/hip_codegen.cpp, VM mode) and the
// device interpreter (kernels/vm_kernels.hip).  Plain C layout.
//
// Values live in 64-bit registers and carry the C type the generated HIP
// source would give them (the interpreter follows the same conversions):
// CQ_T_I32 (int, also bool), CQ_T_U32, CQ_T_I64, CQ_T_U64 (two's
// complement / raw bits) and CQ_T_F64 (IEEE bits).
#ifndef CQ_VM_PROGRAM_H
#define CQ_VM_PROGRAM_H

#define CQ_VM_MAX_CODE 512
#define CQ_VM_MAX_REGS 32
#define CQ_VM_MAX_AGGS 32

enum cq_vm_type { CQ_T_I32 = 0, CQ_T_U32 = 1, CQ_T_I64 = 2, CQ_T_U64 = 3, CQ_T_F64 = 4 };

enum cq_vm_op {
    CQ_OP_END = 0,
    CQ_OP_COL,      // d = fact column a (int32) of the row
    CQ_OP_VALID,    // d = validity bit of fact column a
    CQ_OP_DCOL,     // d = dimension column a (int32) of the row (build programs)
    CQ_OP_DVALID,   // d = validity bit of dimension column a
    CQ_OP_PAY,      // d = payload register of join a (0 before / without a payload)
    CQ_OP_LIT,      // d = P.lit[imm] as type t (CQ_T_I32 truncates like (int))
    CQ_OP_FLIT,     // d = P.flit[imm]
    CQ_OP_IMM,      // d = 64-bit immediate (imm | imm2 << 32) of type t
    CQ_OP_SLOT,     // d = result slot imm of the group (HAVING programs) as type t
    CQ_OP_SLOTV,    // d = validity bit of result slot imm
    CQ_OP_CONV,     // d = (t)a where a has type t2
    CQ_OP_ADD, CQ_OP_SUB, CQ_OP_MUL, CQ_OP_DIV, CQ_OP_MOD,  // d = a op b in type t
    CQ_OP_EQ, CQ_OP_NE, CQ_OP_LT, CQ_OP_LE, CQ_OP_GT, CQ_OP_GE,  // d = (a op b) in type t, as int 0/1
    CQ_OP_NOT,      // d = !a            (a of type t)
    CQ_OP_AND,      // d = a && b        (a of type t, b of type t2)
    CQ_OP_OR,       // d = a || b
    CQ_OP_SEL,      // d = c ? a : b     (c of type t2; a, b already of the result type)
    CQ_OP_FILTER,   // row survives only if a != 0 (type t)
    CQ_OP_PROBE,    // probe join imm with key register a; payload into the join's payload register
    CQ_OP_EXIT,     // the row leaves the pipeline if no row of the wave is alive (early exit)
    CQ_OP_FCMP      // fast-path filter: row survives only if (int)fact column imm  <t>  (int)P.lit[c]
                    // (t one of CQ_OP_EQ..CQ_OP_GE; raw values, no validity)
};

struct cq_vm_instr {
    unsigned char op, t, t2, d;
    unsigned char a, b, c, pad;
    int imm;
    unsigned int imm2;
};

// dimension table of one join, in the layout the code generator chose
enum cq_vm_layout { CQ_L_REF = 0, CQ_L_COMPACT = 1, CQ_L_DENSE = 2 };
struct cq_vm_join {
    int layout;
    int fk_slot;          // fact column of the probe key
    int pk_dslot;         // dimension key column
    int payload_dslot;    // -1 without payload
    int pk_nullable, fk_nullable;
    int kmin, len;        // reference slot function
    int eb, tb, jlo, pmin;             // compact
    int dmin, dspan, db, pbase, doff, zmode, zx;  // key-indexed
    int cal;                           // calendar index from this first year (0: key - dmin)
    int build_begin, build_end;        // build-filter program (code range; empty: every row)
};

enum cq_vm_agg_fn { CQ_VM_SUM = 0, CQ_VM_MIN = 1, CQ_VM_MAX = 2, CQ_VM_COUNT = 3 };

// Linear form of the same plan, when it has one (cq_lin_program.enabled):
// fast-path fact filters only (AND of `(int)column op (int)literal` on raw
// values), at most one join, on a key-indexed table, group keys that are
// fact or payload columns of dense domains, and aggregate inputs that are
// left-deep chains in unsigned long long (the leftmost leaf a column cast to
// u64, every right operand a column, the payload or a literal, each
// converted to u64 as C converts it).  Literals stay P.lit references, so
// one program serves every query of its shape.  An ahead-of-time kernel
// evaluates the form with no per-instruction dispatch and produces the
// values the bytecode would.
#define CQ_LIN_COLS 8
#define CQ_LIN_TERMS 8
#define CQ_LIN_AGGS 4
#define CQ_LIN_SRCS 4
#define CQ_LIN_KEYS 2
#define CQ_LIN_RACC 8    // scalar result slots (aggregates + AVG counts)
#define CQ_LIN_LDS_TABLE_WORDS 4096  // largest join table probed from LDS (16 KiB)
struct cq_lin_term {
    int col;   // column position (cq_lin_program.slot)
    int op;    // CQ_OP_EQ .. CQ_OP_GE
    int lit;   // P.lit index
    int pad;
};
// one step of a left-deep chain: x = x op operand (step 0: x = operand)
enum cq_lin_kind { CQ_LIN_COL = 0, CQ_LIN_PAY = 1, CQ_LIN_LIT = 2 };
struct cq_lin_step {
    int op;    // CQ_OP_ADD / CQ_OP_SUB / CQ_OP_MUL (ignored for step 0)
    int kind;  // cq_lin_kind
    int idx;   // column position, or P.lit index
    int i32;   // literal read as (int)P.lit (else as long long)
};
struct cq_lin_agg {
    int nsteps;                         // 0: the constant 1 (COUNT)
    int nvalid;                         // validity: AND of these column positions' bits
    int valid_col[CQ_LIN_SRCS];
    struct cq_lin_step step[CQ_LIN_SRCS];
};
struct cq_lin_key {
    int kind, col;                      // CQ_LIN_COL (position) or CQ_LIN_PAY
    int valid_col;                      // nullable fact column: its position, else -1
    int pad;
};
struct cq_lin_program {
    int enabled;
    int ncols;                          // column positions: the filters' first
    int nfilter_cols;                   // positions [0, nfilter_cols) hold every filtered column
    int slot[CQ_LIN_COLS];              // fact column slot (P.col / P.valid) per position
    int nullable[CQ_LIN_COLS];
    int nterms;                         // sorted by column position
    int term_end[CQ_LIN_COLS];          // terms of position c: [term_end[c - 1], term_end[c])
    struct cq_lin_term term[CQ_LIN_TERMS];
    int join;                           // -1: none; else the join (key-indexed layout) probed
    int join_col;                       // position of its fact key
    int join_lds_words;                 // > 0: its table (this many words) is copied to LDS and probed there
    int naggs;                          // = cq_vm_program.naggs
    struct cq_lin_agg agg[CQ_LIN_AGGS];
    int ngroups;
    struct cq_lin_key key[CQ_LIN_KEYS];
};

// Several linear-form programs over the same fact rows in one pass
// (k_lin_multi): the union of their columns is staged once per tile and
// every program reads it.  Program k's linear form indexes the union
// positions (col / valid); the programs and their parameter blocks are
// arrays in device memory beside this descriptor.  LDS offsets are bytes
// into the kernel's dynamic LDS after the staged columns; CQ_LIN_NONE: the
// join table is probed in HBM / the group aggregates go to global atomics.
#define CQ_LIN_BATCH 16
#define CQ_LIN_NONE 0xffffffffu
struct cq_lin_batch {
    int nprog;
    int ncols;
    long long row_begin, row_end;
    const int* col[CQ_LIN_COLS];
    const unsigned int* valid[CQ_LIN_COLS];  // bitmap of a nullable column, else null
    unsigned int tab_off[CQ_LIN_BATCH];
    unsigned int grp_off[CQ_LIN_BATCH];
};

struct cq_vm_program {
    int ncode;
    struct cq_vm_instr code[CQ_VM_MAX_CODE];
    int nregs;
    int njoins;
    struct cq_vm_join joins[8];
    int scan_begin, scan_end;          // per-row pipeline
    // aggregation: per aggregate the value / validity registers after the
    // pipeline (value already converted to u64), function, hidden count slot
    int naggs, ngroups, storage_ts, visible_ts;
    int agg_fn[CQ_VM_MAX_AGGS];
    int agg_val[CQ_VM_MAX_AGGS], agg_valid[CQ_VM_MAX_AGGS];  // registers (-1: constant 1 / always valid)
    int agg_hidden[CQ_VM_MAX_AGGS];                          // AVG count slot or -1
    int agg_static[CQ_VM_MAX_AGGS];                          // validity set on every hit (COUNT)
    // GROUP BY: group index register, key value / validity registers
    int group_reg;
    long long groups;
    int key_val[CQ_VM_MAX_AGGS], key_valid[CQ_VM_MAX_AGGS];
    // decode of a dense domain (keys rebuilt from the group index at merge)
    int decodable;
    long long key_min[CQ_VM_MAX_AGGS];
    long long key_card[CQ_VM_MAX_AGGS];
    int key_nullable[CQ_VM_MAX_AGGS];
    int fast_sparse;
    // HAVING (one "row" per group)
    int having_begin, having_end, having_reg;  // having_reg < 0: no HAVING
    int having_type;
    struct cq_lin_program lin;
};

#endif
