// Interpreted tier: compiles the C expressions the HIP code generator emits
// (the same strings hipRTC would compile) into typed register code for the
// ahead-of-time interpreter (device/vm_program.h).  Types follow C: literal
// typing, integer promotion, the usual arithmetic conversions, int32 / u32
// wrap-around (-fwrapv), and the generator's helpers (chip::safe_* promote
// int to int64 first, chip::cdiv divides with x / 0 == 0).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "../device/vm_program.h"

namespace db {
namespace vm {

struct Val {
    int reg = -1;
    int type = CQ_T_I32;
};

class Builder {
public:
    explicit Builder(std::vector<cq_vm_instr>& code) : code_(code) {}
    // registers holding the payload of each join (read by "p<j>[i]")
    void setPayloadReg(int join, int reg) { payload_[join] = reg; }
    int allocPermanent();
    // compile `expr`; the value stays in a register until release()
    Val compile(const std::string& expr);
    Val toType(Val v, int type);  // C conversion
    void release(const Val& v);
    int emit(int op, int t, int t2, int d, int a, int b, int c, long long imm);
    int maxRegs() const { return max_regs_; }
    void resetColumnCache() { cols_.clear(), vcols_.clear(); }

private:
    std::vector<cq_vm_instr>& code_;
    std::map<int, int> payload_;
    std::map<int, int> cols_, vcols_;  // fact slot -> register of its value / validity
    std::vector<bool> used_ = std::vector<bool>(CQ_VM_MAX_REGS, false);
    std::vector<bool> permanent_ = std::vector<bool>(CQ_VM_MAX_REGS, false);
    int max_regs_ = 0;
    int alloc();
};

}  // namespace vm
}  // namespace db
