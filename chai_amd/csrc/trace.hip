// trace.hip -- fused-trace compiler (host) and VM kernel (device).
//
// Replaces the reference's JIT middle and back end for elementwise /
// reduction / matmul statements: MakeStmt splitting (jitstmt/MakeStmt.cpp:
// 579-726), kernelization (jitkrnl/WorkTrace.cpp:62-370,
// jitkrnl/GroupTogether.cpp:178-201), OpenCL text emission
// (enqueue/EnqueueIdSpace.cpp:19-320, EnqueueReduce.cpp:21-222,
// EnqueueAST.cpp:95-740) and the per-trace JustInTimeMemo.
//
// Grouping rules (kept from the reference): matmul is its own kernel;
// elementwise statements over one index space fuse; a reduction is fused into
// the kernel that computes its operand (the reference instead emits a separate
// serial reduction kernel); a statement reading a reduction result of the
// current group starts a new group, and so does a statement gathering
// (gather1_floor / gather2_floor) from a variable the current group computes.
// Within a group, TEMP variables live in VM registers; a variable crossing
// groups, or gathered from, is stored.
#include "common.hpp"
#include "gemm_internal.hpp"
#include "trace_vm.hpp"
#include "jit_headers.inc"

#include <hip/hiprtc.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace chai_hip {
// program_cl.hip: launch of an inline OpenCL kernel by handle
int cl_launch_kernel(void* kernel, const chai_cl_arg* args, uint32_t nargs, uint32_t dims, const uint64_t* global,
                     const uint64_t* local, hipStream_t st);

// ===================================================================== device

#define VM_SLOTS(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11)
constexpr int VM_NSLOTS = 12;
#define VM_RD_CASE(n) case n: v = R##n; break;
#define VM_WR_CASE(n) case n: R##n = v; break;
#define VM_DECL(n) V4 R##n = {{0, 0, 0, 0}};

__global__ __launch_bounds__(VM_THREADS) void trace_vm_kernel(VmLaunchArgs L)
{
    const uint64_t slot = blockIdx.y;
    VM_SLOTS(VM_DECL)
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    // table entries are picked with constant indices (a runtime index into the
    // by-value argument block would make the compiler copy it to scratch)
#define VM_VAR_CASE(n) case n: return L.vars[n];
#define VM_RNG_CASE(n) case n: return L.rng[n];
    auto var_at = [&](uint32_t k) -> VmVar {
        switch (k) { VM_VAR_CASE(0) VM_VAR_CASE(1) VM_VAR_CASE(2) VM_VAR_CASE(3) VM_VAR_CASE(4)
                     VM_VAR_CASE(5) VM_VAR_CASE(6) VM_VAR_CASE(7) VM_VAR_CASE(8) VM_VAR_CASE(9)
                     VM_VAR_CASE(10) VM_VAR_CASE(11) VM_VAR_CASE(12) VM_VAR_CASE(13) VM_VAR_CASE(14)
                     VM_VAR_CASE(15) VM_VAR_CASE(16) VM_VAR_CASE(17) VM_VAR_CASE(18) VM_VAR_CASE(19)
                     VM_VAR_CASE(20) VM_VAR_CASE(21) VM_VAR_CASE(22)
                     default: return L.vars[23]; }
    };
    auto rng_at = [&](uint32_t k) -> VmRngDims {
        switch (k) { VM_RNG_CASE(0) VM_RNG_CASE(1) VM_RNG_CASE(2) VM_RNG_CASE(3) VM_RNG_CASE(4)
                     VM_RNG_CASE(5) VM_RNG_CASE(6) default: return L.rng[7]; }
    };
#define VM_CHAIN_CASE(n) case n + 1: return L.chains[n];
    auto chain_at = [&](uint32_t k) -> VmChain {   // k = chain index + 1; 0 = none
        switch (k) { case 0: return VmChain{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0, 0, 0, 0};
                     VM_CHAIN_CASE(0) VM_CHAIN_CASE(1) VM_CHAIN_CASE(2) VM_CHAIN_CASE(3)
                     VM_CHAIN_CASE(4) VM_CHAIN_CASE(5) VM_CHAIN_CASE(6) default: return L.chains[7]; }
    };
    const uint64_t stride = (uint64_t)gridDim.x * VM_THREADS;
    for (uint64_t v = (uint64_t)blockIdx.x * VM_THREADS + threadIdx.x; v < L.vps; v += stride) {
        const uint64_t y = v / L.vpr;
        const uint64_t x0 = (v - y * L.vpr) * 4;
        const int nval = (L.W - x0) < 4 ? (int)(L.W - x0) : 4;
        for (uint32_t pc = 0; pc < L.nins; pc++) {
            const VmIns I = L.prog[pc];
            V4 a, b, c, v;
            bool wr = true;
            switch (I.s1) { VM_SLOTS(VM_RD_CASE) }
            a = v;
            switch (I.op) {
                case VM_LOAD:
                    if (I.p2) v = vm_load_c(var_at(I.a), chain_at(I.p2), slot, y, x0, nval);
                    else v = vm_load(var_at(I.a), slot, y, x0, nval, L.W, L.H);
                    break;
                case VM_IMM: v.l[0] = v.l[1] = v.l[2] = v.l[3] = I.u; break;
                case VM_RNG: {
                    const RngArgs ra = {I.a, I.b, I.c, I.p, I.u, I.d0, I.d1};
                    v = vm_rng(ra, rng_at(I.e), slot, y, x0, nval, chain_at(I.p2));
                    break;
                }
                case VM_INDEX: {
                    const VmChain C = chain_at(I.p2);
                    for (int i = 0; i < 4; i++)
                        v.l[i] = from_double(I.p, (double)(I.a == 0 ? chain_mod(x0 + i, C.W, C.nW)
                                                                    : chain_mod(y, C.H, C.nH)));
                    break;
                }
                case VM_CVT:
                    for (int i = 0; i < 4; i++) v.l[i] = from_double(I.p, as_double(I.p2, a.l[i]));
                    break;
                case VM_UN: v = vm_un(I.a, I.p, a); break;
                case VM_BIN:
                    switch (I.s2) { VM_SLOTS(VM_RD_CASE) }
                    v = vm_bin(I.a, I.p, a, v);
                    break;
                case VM_TER:
                    switch (I.s2) { VM_SLOTS(VM_RD_CASE) }
                    b = v;
                    switch (I.s3) { VM_SLOTS(VM_RD_CASE) }
                    c = v;
                    v = vm_ter(I.a, I.p, a, b, c);
                    break;
                case VM_SEL:
                    switch (I.s2) { VM_SLOTS(VM_RD_CASE) }
                    b = v;
                    switch (I.s3) { VM_SLOTS(VM_RD_CASE) }
                    c = v;
                    for (int i = 0; i < 4; i++) v.l[i] = as_double(I.p2, a.l[i]) != 0.0 ? b.l[i] : c.l[i];
                    break;
                case VM_GATHER:
                    switch (I.s2) { VM_SLOTS(VM_RD_CASE) }
                    v = vm_gather(var_at(I.a), chain_at(I.p2), slot, y, I.e == 2, I.b, a, I.c, v, nval);
                    break;
                case VM_STORE:
                    vm_store(var_at(I.a), a, slot, y, x0, nval, L.W, L.H);
                    wr = false;
                    break;
                case VM_ACC: {
                    double s = 0.0;
                    for (int i = 0; i < 4; i++) if (i < nval) s += as_double(I.p, a.l[i]);
                    if (I.a == 0) acc0 += s;
                    else if (I.a == 1) acc1 += s;
                    else if (I.a == 2) acc2 += s;
                    else acc3 += s;
                    wr = false;
                    break;
                }
                default: v = a; break;   // VM_MOV
            }
            if (wr) switch (I.dst) { VM_SLOTS(VM_WR_CASE) }
        }
    }
    if (L.nacc) {
        const uint64_t base = ((uint64_t)slot * gridDim.x + blockIdx.x) * L.nacc;
        double r = block_sum_f64<VM_THREADS>(acc0);
        if (threadIdx.x == 0) L.partial[base + 0] = r;
        if (L.nacc > 1) { r = block_sum_f64<VM_THREADS>(acc1); if (threadIdx.x == 0) L.partial[base + 1] = r; }
        if (L.nacc > 2) { r = block_sum_f64<VM_THREADS>(acc2); if (threadIdx.x == 0) L.partial[base + 2] = r; }
        if (L.nacc > 3) { r = block_sum_f64<VM_THREADS>(acc3); if (threadIdx.x == 0) L.partial[base + 3] = r; }
    }
}

struct RedOut {
    void* out;
    uint32_t prec;
    double divisor;
};

// fixed-order combine of the per-block partials: one block per (slot, accumulator)
__global__ __launch_bounds__(VM_THREADS) void trace_reduce_final(const double* __restrict__ partial,
                                                                 uint32_t blocks, uint32_t nacc,
                                                                 RedOut o0, RedOut o1, RedOut o2, RedOut o3)
{
    const uint32_t slot = blockIdx.x, k = blockIdx.y;
    double s = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += VM_THREADS)
        s += partial[((uint64_t)slot * blocks + b) * nacc + k];
    s = block_sum_f64<VM_THREADS>(s);
    if (threadIdx.x == 0) {
        const RedOut o = k == 0 ? o0 : k == 1 ? o1 : k == 2 ? o2 : o3;
        const double r = o.divisor != 1.0 ? s / o.divisor : s;
        const uint64_t bits = from_double(o.prec, r);
        if (o.prec == CHAI_F64) ((uint64_t*)o.out)[slot] = bits;
        else ((uint32_t*)o.out)[slot] = (uint32_t)bits;
    }
}

// ======================================================================= host

namespace {

bool is_fp(int p) { return p == CHAI_F32 || p == CHAI_F64; }

// opcode classes (bytecode/ByteCodes.hpp)
bool is_scalar(uint32_t c) { return c >= 5 && c <= 8; }
bool is_convert(uint32_t c) { return c >= 1 && c <= 4; }
bool is_rng(uint32_t c) { return c >= 110 && c <= 115; }
bool is_index(uint32_t c) { return c >= 42 && c <= 49; }
bool is_red(uint32_t c) { return c == 130 || c == 131 || c == 132; }
bool is_mm(uint32_t c) { return c == 120 || c == 121; }
bool is_gather(uint32_t c) { return c == 40 || c == 41; }
bool is_litdata(uint32_t c) { return c >= 68 && c <= 82; }
bool is_cl(uint32_t c) { return c == 50; }          // kernel_from_opencl
bool is_tr(uint32_t c) { return c == 122; }         // transpose
// an inline OpenCL kernel statement: its last node is kernel_from_opencl
bool is_cl_stmt(const chai_trace& t, const chai_trace_stmt& st)
{
    return st.count > 0 && (uint64_t)st.first + st.count <= t.num_nodes && is_cl(t.nodes[st.first + st.count - 1].code);
}

int prec_of_suffix(int k) { return k == 0 ? CHAI_U32 : k == 1 ? CHAI_I32 : k == 2 ? CHAI_F32 : CHAI_F64; }

// ones1/ones2/zeros1/zeros2 leaves (bytecode/ByteCodes.hpp:90-105): the value
// and type the reference's translator gives them (runtime/Translator.cpp:
// 116-131, jitast/AstLitdata.cpp).  The dispatch table keeps the last
// registration of a code, and zeros2_f64 is declared as 73, the code of
// ones2_i32, so 73 is an f64 zero.
void litdata(uint32_t c, int& prec, double& v)
{
    if (c == 73) { prec = CHAI_F64; v = 0.0; return; }
    prec = prec_of_suffix((int)((c - 68) % 4));
    v = c < 76 ? 1.0 : 0.0;
}

bool is_unary(uint32_t c, bool& fp_only)
{
    static const uint32_t fp1[] = {223, 261, 256, 219, 220, 221, 222, 237, 238, 239, 240, 257, 214, 262,
                                   203, 200, 206, 259, 215, 263, 212, 253, 255, 266, 211, 218, 217,
                                   201, 202, 204, 205, 208, 209, 216, 233, 235, 241, 246, 260, 264, 265,
                                   301, 305, 308, 507, 508, 509, 510, 513};
    fp_only = false;
    if (c == 10 || c == 15 || c == 16 || c == 400 || c == 404 || c == 225) return true;
    for (uint32_t f : fp1)
        if (c == f) { fp_only = true; return true; }
    return false;
}

bool is_binary(uint32_t c, int& kind)   // kind 0: any, 1: fp only, 2: int only
{
    static const uint32_t any[] = {11, 12, 13, 14, 32, 33, 34, 35, 36, 37, 302, 303};
    static const uint32_t fp[] = {227, 228, 248, 207, 229, 232, 213, 224, 210, 234, 243, 244, 247, 249,
                                  250, 251, 254, 306, 500, 501, 502, 503, 504, 505, 506, 511, 512};
    static const uint32_t in[] = {17, 18, 19, 20, 21, 22, 30, 31, 401, 402, 405, 411, 412, 413, 414, 415};
    for (uint32_t x : any) if (c == x) { kind = 0; return true; }
    for (uint32_t x : fp) if (c == x) { kind = 1; return true; }
    for (uint32_t x : in) if (c == x) { kind = 2; return true; }
    return false;
}

bool is_ternary(uint32_t c)
{
    return c == 23 || c == 300 || c == 226 || c == 242 || c == 304 || c == 307 || c == 406 || c == 407 ||
           c == 408;
}

struct Dims {
    uint64_t W = 1, H = 1, S = 1;
    bool operator==(const Dims& o) const { return W == o.W && H == o.H && S == o.S; }
    bool operator!=(const Dims& o) const { return !(*this == o); }
};

struct ENode {
    uint32_t code;
    const chai_trace_node* src;
    std::vector<int> kids;
    int prec = 0;
    Dims d;
};

struct Stmt {
    uint32_t lhs;
    std::vector<ENode> nodes;   // nodes[root] last
    int root;
    int kind;                   // 0 elementwise, 1 reduction, 2 matmul
    Dims space;                 // index space of the kernel computing it
};

struct Group {
    int kind;                   // 0 VM, 2 matmul
    Dims space;
    std::vector<int> stmts;
    // VM
    std::vector<VmIns> prog;
    std::vector<uint32_t> var_of_slot;   // VmVar table: trace variable per entry
    std::vector<VmRngDims> rng;
    std::vector<VmChain> chains;         // broadcast chains of leaf instructions
    std::vector<uint32_t> red_vars;      // output variable of each accumulator
    std::vector<double> red_div;
    VmIns* dprog = nullptr;
    // JIT: the same program as straight-line HIP, compiled with hipRTC
    std::string jit_src, jit_log;
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

}  // namespace

// a trace held by the library: the caller's copy, or a rewritten one
struct Versioned {
    std::vector<chai_trace_var> vars;
    std::vector<chai_trace_stmt> stmts;
    std::vector<chai_trace_node> nodes;
    chai_trace t;
    void adopt(const chai_trace& in)
    {
        vars.assign(in.vars, in.vars + in.num_vars);
        stmts.assign(in.stmts, in.stmts + in.num_stmts);
        nodes.assign(in.nodes, in.nodes + in.num_nodes);
        t.vars = vars.data(); t.num_vars = (uint32_t)vars.size();
        t.stmts = stmts.data(); t.num_stmts = (uint32_t)stmts.size();
        t.nodes = nodes.data(); t.num_nodes = (uint32_t)nodes.size();
    }
};

struct TraceProgram {
    // the program's own copies of the trace (statements point into them)
    Versioned own, spl, ver;
    std::vector<chai_trace_var> vars;
    uint32_t nuser = 0;               // variables with caller buffers (the rest are versions)
    std::vector<Stmt> stmts;
    std::vector<Group> groups;
    std::vector<bool> needs_mem;      // TEMP variables that must be stored
    int ninstr = 0;
    ~TraceProgram()
    {
        for (auto& g : groups) {
            if (g.dprog) (void)hipFree(g.dprog);
            if (g.mod) (void)hipModuleUnload(g.mod);
        }
    }
};

namespace {

int fail(const char* fmt, ...)
{
    char buf[300];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_error("trace_compile: %s", buf);
    return CHAI_ERR_ARG;
}

Dims dmax(const Dims& a, const Dims& b)
{
    Dims r;
    r.W = a.W > b.W ? a.W : b.W;
    r.H = a.H > b.H ? a.H : b.H;
    r.S = a.S > b.S ? a.S : b.S;
    return r;
}

// result shape of matmul(A, B) with the operands' transpose flags, the
// oracle's dispatch: matrix * vector, vector * matrix, matrix * matrix.  With
// slots, instance s multiplies A's instance s mod A.slots by B's instance
// s mod B.slots (interp/BaseInterp.cpp:79-98, interp/InterpMatmulBase.cpp:
// 113-118): max(A.slots, B.slots) results.
Dims mm_dims(const Dims& a, const Dims& b, bool ta, bool tb)
{
    Dims r;
    r.S = a.S > b.S ? a.S : b.S;
    if (b.H == 1 && !tb && a.H > 1) {
        r.W = ta ? a.W : a.H;
    } else if (a.H == 1 && !ta && b.H > 1) {
        r.W = tb ? b.H : b.W;
    } else {
        r.W = tb ? b.H : b.W;
        r.H = ta ? a.W : a.H;
    }
    return r;
}

// operand count of a node, -1 for a code the GPU back end does not know
int node_arity(uint32_t c)
{
    bool fp_only;
    int bkind;
    if (c == CHAI_TRACE_VAR || is_scalar(c) || is_rng(c) || is_index(c) || is_litdata(c)) return 0;
    if (is_convert(c) || c == 131 || c == 132 || is_tr(c) || is_unary(c, fp_only)) return 1;
    if (c == 130 || c == 120 || c == 40 || is_binary(c, bkind)) return 2;
    if (c == 121 || c == 41 || is_ternary(c)) return 3;
    return -1;
}

// types and shapes of one postfix expression (no checks against a variable)
int parse_nodes(const chai_trace& t, const chai_trace_node* nodes, uint32_t count, Stmt& s)
{
    std::vector<int> stack;
    for (uint32_t i = 0; i < count; i++) {
        const chai_trace_node& n = nodes[i];
        ENode e;
        e.code = n.code;
        e.src = &n;
        int arity = 0;
        bool fp_only = false;
        int bkind = 0;
        if (n.code == CHAI_TRACE_VAR) {
            if (n.arg >= t.num_vars) return fail("read of variable %u of %u", n.arg, t.num_vars);
            const chai_trace_var& v = t.vars[n.arg];
            e.prec = (int)v.prec;
            e.d.W = v.W; e.d.H = v.H; e.d.S = v.slots;
        } else if (is_scalar(n.code)) {
            e.prec = prec_of_suffix((int)n.code - 5);
        } else if (is_rng(n.code)) {
            static const int rp[] = {CHAI_F32, CHAI_F64, CHAI_U32, CHAI_I32, CHAI_F32, CHAI_F64};
            e.prec = rp[n.code - 110];
            e.d.W = n.W; e.d.H = n.H ? n.H : 1; e.d.S = n.slots ? n.slots : 1;
            if (n.arg != CHAI_RNG_PHILOX && n.arg != CHAI_RNG_THREEFRY) return fail("rng variant %u", n.arg);
        } else if (is_index(n.code)) {
            e.prec = prec_of_suffix((int)((n.code - 42) % 4));
            e.d.W = n.W; e.d.H = n.H ? n.H : 1; e.d.S = n.slots ? n.slots : 1;
        } else if (is_litdata(n.code)) {
            double v;
            litdata(n.code, e.prec, v);
            e.d.W = n.W; e.d.H = n.H ? n.H : 1;
            if (n.W == 0) return fail("ones/zeros leaf with zero width");
        } else if (is_convert(n.code)) {
            arity = 1;
            e.prec = prec_of_suffix((int)n.code - 1);
        } else if (is_unary(n.code, fp_only)) {
            arity = 1;
        } else if (is_binary(n.code, bkind)) {
            arity = 2;
        } else if (is_ternary(n.code)) {
            arity = 3;
        } else if (n.code == 131 || n.code == 132 || is_tr(n.code)) {
            arity = 1;
        } else if (n.code == 130) {
            arity = 2;
        } else if (is_gather(n.code)) {
            arity = n.code == 40 ? 2 : 3;
        } else if (n.code == 120) {
            arity = 2;
        } else if (n.code == 121) {
            arity = 3;
        } else {
            return fail("bytecode %u is not supported by the GPU back end", n.code);
        }
        if ((int)stack.size() < arity) return fail("stack underflow at bytecode %u", n.code);
        e.kids.assign(stack.end() - arity, stack.end());
        stack.resize(stack.size() - arity);
        // typing
        if (is_gather(n.code)) {
            // shape and type of the data array, which must be a variable in
            // memory (the reference's JIT gathers from a data variable too,
            // jitstmt/MakeStmt.cpp:656-673); subscripts keep their own types
            const ENode& a = s.nodes[e.kids[0]];
            if (a.code != CHAI_TRACE_VAR) return fail("gather data must be a variable");
            if (a.src->flags & 1) return fail("gather from a transposed variable is not supported");
            e.prec = a.prec;
            e.d = a.d;
        } else if (arity == 1 && !is_convert(n.code)) {
            const ENode& k = s.nodes[e.kids[0]];
            e.prec = k.prec;
            e.d = k.d;
            if (fp_only && !is_fp(k.prec)) return fail("bytecode %u needs a floating point operand", n.code);
            if (n.code == 15 || n.code == 16 || n.code == 404) {
                if (is_fp(k.prec)) return fail("bytecode %u needs an integer operand", n.code);
            }
            if (n.code == 131 || n.code == 132) { e.d.W = 1; e.d.H = 1; }
            if (is_tr(n.code)) { e.d.W = k.d.H; e.d.H = k.d.W; }   // interp/InterpTranspose.cpp
        } else if (is_convert(n.code)) {
            e.d = s.nodes[e.kids[0]].d;
        } else if (arity == 2) {
            const ENode& a = s.nodes[e.kids[0]];
            const ENode& b = s.nodes[e.kids[1]];
            e.prec = a.prec > b.prec ? a.prec : b.prec;
            if (n.code == 120) {
                e.d = mm_dims(a.d, b.d, a.src->flags & 1, b.src->flags & 1);
            } else {
                e.d = dmax(a.d, b.d);
                if (bkind == 1 && !is_fp(e.prec)) return fail("bytecode %u needs floating point", n.code);
                if (bkind == 2 && is_fp(e.prec)) return fail("bytecode %u needs integers", n.code);
                if (n.code == 130) { e.d.W = 1; e.d.H = 1; }
            }
        } else if (arity == 3) {
            const ENode& a = s.nodes[e.kids[0]];
            const ENode& b = s.nodes[e.kids[1]];
            const ENode& c = s.nodes[e.kids[2]];
            if (n.code == 23) {        // cond: type and dims from the two branches (InterpCond.cpp)
                e.prec = b.prec > c.prec ? b.prec : c.prec;
                e.d = dmax(b.d, c.d);
            } else if (n.code == 121) {
                e.prec = c.prec;
                e.d = c.d;
            } else {
                e.prec = a.prec > b.prec ? a.prec : b.prec;
                e.prec = e.prec > c.prec ? e.prec : c.prec;
                e.d = dmax(dmax(a.d, b.d), c.d);
                if ((n.code == 226 || n.code == 242 || n.code == 304 || n.code == 307) && !is_fp(e.prec))
                    return fail("bytecode %u needs floating point", n.code);
                if ((n.code == 406 || n.code == 407 || n.code == 408) && is_fp(e.prec))
                    return fail("bytecode %u needs integers", n.code);
            }
        }
        s.nodes.push_back(e);
        stack.push_back((int)s.nodes.size() - 1);
    }
    if (stack.size() != 1) return fail("statement leaves %zu values", stack.size());
    s.root = stack[0];
    return CHAI_OK;
}

// An inline OpenCL kernel statement (api/ProgramCL.cpp:202-219): argument
// nodes then the kernel_from_opencl root.  Arguments are checked against the
// kernel's metadata here, so a bad call fails at compile time.
int parse_cl_stmt(const chai_trace& t, const chai_trace_stmt& cs, Stmt& s)
{
    const chai_trace_node& r = t.nodes[cs.first + cs.count - 1];
    const uint32_t na = cs.count - 1;
    if (r.arg != na) return fail("inline kernel statement: %u arguments declared, %u nodes", r.arg, na);
    if (!r.seed) return fail("inline kernel statement without a kernel handle");
    if (r.vec_len < 1 || r.vec_len > 2) return fail("inline kernel work space must be 1- or 2-dimensional");
    std::vector<chai_cl_arg> args(na);
    bool first_array = true;
    ENode root;
    root.code = r.code;
    root.src = &r;
    for (uint32_t k = 0; k < na; k++) {
        const chai_trace_node& n = t.nodes[cs.first + k];
        ENode e;
        e.code = n.code;
        e.src = &n;
        chai_cl_arg& a = args[k];
        std::memset(&a, 0, sizeof a);
        if (n.code == CHAI_TRACE_VAR) {
            if (n.arg >= t.num_vars) return fail("read of variable %u of %u", n.arg, t.num_vars);
            const chai_trace_var& v = t.vars[n.arg];
            e.prec = (int)v.prec;
            e.d.W = v.W; e.d.H = v.H; e.d.S = v.slots;
            a.kind = CHAI_CL_ARG_ARRAY;
            a.prec = v.prec;
            if (first_array && n.arg != cs.lhs)
                return fail("an inline kernel statement assigns its first array argument");
            first_array = false;
        } else if (is_scalar(n.code)) {
            e.prec = prec_of_suffix((int)n.code - 5);
            a.kind = CHAI_CL_ARG_SCALAR;
            a.prec = (uint32_t)e.prec;
            if (e.prec == CHAI_U32) a.v.u = (uint32_t)(int64_t)n.s0;
            else if (e.prec == CHAI_I32) a.v.i = (int32_t)n.s0;
            else if (e.prec == CHAI_F32) a.v.f = (float)n.s0;
            else a.v.d = n.s0;
        } else if (n.code == CHAI_TRACE_CL_LOCAL) {
            a.kind = CHAI_CL_ARG_LOCAL;
            a.v.count = n.W;
        } else {
            return fail("bytecode %u cannot be an inline kernel argument", n.code);
        }
        root.kids.push_back((int)s.nodes.size());
        s.nodes.push_back(e);
    }
    if (first_array) return fail("an inline kernel statement needs an array argument");
    s.nodes.push_back(root);
    s.root = (int)s.nodes.size() - 1;
    s.kind = 3;
    const uint64_t g[2] = {r.W, r.H}, l[2] = {(uint64_t)r.s0, (uint64_t)r.s1};
    return chai_hip_cl_check((void*)r.seed, args.data(), na, r.vec_len, g, l);
}

int parse_stmt(const chai_trace& t, const chai_trace_stmt& cs, Stmt& s)
{
    s.lhs = cs.lhs;
    if (cs.lhs >= t.num_vars) return fail("statement assigns variable %u of %u", cs.lhs, t.num_vars);
    if ((uint64_t)cs.first + cs.count > t.num_nodes || cs.count == 0) return fail("bad node range");
    if (is_cl_stmt(t, cs)) return parse_cl_stmt(t, cs, s);
    const int rc = parse_nodes(t, t.nodes + cs.first, cs.count, s);
    if (rc) return rc;
    const ENode& r = s.nodes[s.root];
    s.kind = is_red(r.code) ? 1 : is_mm(r.code) ? 2 : is_tr(r.code) ? 4 : 0;
    // reductions / matmul / transpose only at the root
    for (size_t i = 0; i + 1 < s.nodes.size(); i++)
        if (is_red(s.nodes[i].code) || is_mm(s.nodes[i].code) || is_tr(s.nodes[i].code))
            return fail("reduction, matmul or transpose below the root of a statement (split it first)");
    const chai_trace_var& lv = t.vars[s.lhs];
    if (s.kind == 0) {
        s.space = r.d;
        if (r.d.W != lv.W || r.d.H != lv.H || r.d.S != lv.slots)
            return fail("statement shape %llux%llux%llu differs from its variable",
                        (unsigned long long)r.d.W, (unsigned long long)r.d.H, (unsigned long long)r.d.S);
    } else if (s.kind == 1) {
        Dims in = s.nodes[r.kids[0]].d;
        if (r.code == 130) in = dmax(in, s.nodes[r.kids[1]].d);
        s.space = in;
        if (lv.W != 1 || lv.H != 1 || lv.slots != in.S) return fail("reduction result must be 1x1xslots");
    } else if (s.kind == 4) {
        const ENode& k = s.nodes[r.kids[0]];
        if (k.code != CHAI_TRACE_VAR || (k.src->flags & 1)) return fail("transpose operand must be a variable");
        if (r.d.W != lv.W || r.d.H != lv.H || r.d.S != lv.slots) return fail("transpose result has the wrong shape");
    } else {
        for (int k : r.kids)
            if (s.nodes[k].code != CHAI_TRACE_VAR) return fail("matmul operands must be variables");
        // slots: each operand has one instance (shared by every product, the
        // reference's sameDataA / sameDataB, enqueue/EnqueueMatmulAuto.cpp:
        // 57-68) or as many as the result
        if (lv.slots != r.d.S) return fail("matmul result has %llu slots, its variable %llu",
                                           (unsigned long long)r.d.S, (unsigned long long)lv.slots);
        for (int k : r.kids) {
            const uint64_t ks = s.nodes[k].d.S;
            if (ks != 1 && ks != lv.slots) return fail("matmul operand with %llu of %llu slots",
                                                      (unsigned long long)ks, (unsigned long long)lv.slots);
        }
    }
    return CHAI_OK;
}

struct Emitter {
    const chai_trace& t;
    TraceProgram& P;
    Group& G;
    std::map<uint32_t, int> reg_of_var;   // temp variable -> register slot
    std::map<uint32_t, uint32_t> vtab;    // trace variable -> VmVar table index
    bool temp_used[VM_NSLOTS] = {};      // register slots holding live temporaries
    int max_stack = 0;
    std::vector<Dims> path;               // dims of the operators above the node being emitted

    // Broadcast chain of a leaf with dims `leaf` under the operators in `path`
    // (see VmChain); the leaf's own modulus is part of the chain only when
    // include_leaf (index leaves; loads and rng apply it themselves).
    // Returns the chain index + 1, 0 for none, -1 on overflow.
    int chain_for(const Dims& leaf, bool include_leaf)
    {
        VmChain c;
        std::memset(&c, 0, sizeof c);
        auto one = [&](uint64_t Dims::*f, uint64_t space, uint64_t* out, uint32_t& n) -> bool {
            std::vector<uint64_t> st;
            auto push = [&](uint64_t v) {
                if (v >= space) return;                     // coordinates are already below it
                while (!st.empty() && st.back() % v == 0) st.pop_back();
                st.push_back(v);
            };
            for (const Dims& a : path) push(a.*f);
            push(leaf.*f);
            if (!include_leaf && !st.empty() && st.back() == leaf.*f) st.pop_back();
            if (st.size() > 3) return false;
            n = (uint32_t)st.size();
            for (size_t i = 0; i < st.size(); i++) out[i] = st[i];
            return true;
        };
        if (!one(&Dims::W, G.space.W, c.W, c.nW) || !one(&Dims::H, G.space.H, c.H, c.nH) ||
            !one(&Dims::S, G.space.S, c.S, c.nS))
            return -1;
        if ((c.nW | c.nH | c.nS) == 0) return 0;
        for (size_t i = 0; i < G.chains.size(); i++)
            if (std::memcmp(&G.chains[i], &c, sizeof c) == 0) return (int)i + 1;
        if (G.chains.size() >= 8) return -1;
        G.chains.push_back(c);
        return (int)G.chains.size();
    }

    int min_temp() const
    {
        for (int s = 0; s < VM_NSLOTS; s++)
            if (temp_used[s]) return s;
        return VM_NSLOTS;
    }
    // highest free slot above `above`, or -1
    int alloc_temp(int above)
    {
        for (int s = VM_NSLOTS - 1; s > above; s--)
            if (!temp_used[s]) { temp_used[s] = true; return s; }
        return -1;
    }

    uint32_t var_index(uint32_t v)
    {
        auto it = vtab.find(v);
        if (it != vtab.end()) return it->second;
        const uint32_t k = (uint32_t)G.var_of_slot.size();
        G.var_of_slot.push_back(v);
        vtab[v] = k;
        return k;
    }

    VmIns ins(uint16_t op, int p, int dst)
    {
        VmIns I;
        std::memset(&I, 0, sizeof I);
        I.op = op;
        I.p = (uint8_t)p;
        I.dst = (uint8_t)dst;
        return I;
    }

    // value of node n in slot >= tgt (or in a temp slot); returns the slot
    int emit(const Stmt& s, int n, int tgt, int& err)
    {
        if (tgt > max_stack) max_stack = tgt;
        if (tgt >= min_temp()) { err = fail("expression too deep for the VM registers"); return 0; }
        const ENode& e = s.nodes[n];
        const chai_trace_node& src = *e.src;
        int chain = 0;
        if (e.code == CHAI_TRACE_VAR || is_rng(e.code) || is_index(e.code)) {
            chain = chain_for(e.d, is_index(e.code));
            if (chain < 0) { err = fail("broadcast nesting too deep for the VM"); return 0; }
        }
        if (e.code == CHAI_TRACE_VAR) {
            auto it = reg_of_var.find(src.arg);
            if (it != reg_of_var.end()) {
                if (chain) { err = fail("register-resident variable under a broadcast"); return 0; }
                return it->second;
            }
            VmIns I = ins(VM_LOAD, e.prec, tgt);
            I.p2 = (uint8_t)chain;
            I.a = var_index(src.arg);
            G.prog.push_back(I);
            return tgt;
        }
        if (is_scalar(e.code) || is_litdata(e.code)) {
            // a constant: its dims only shape the statement, every element is equal
            double val = src.s0;
            if (is_litdata(e.code)) { int p; litdata(e.code, p, val); }
            VmIns I = ins(VM_IMM, e.prec, tgt);
            switch (e.prec) {
                case CHAI_U32: I.u = (uint32_t)(int64_t)val; break;
                case CHAI_I32: I.u = (uint32_t)(int32_t)(int64_t)val; break;
                case CHAI_F32: { float f = (float)val; uint32_t b; std::memcpy(&b, &f, 4); I.u = b; break; }
                default: std::memcpy(&I.u, &val, 8); break;
            }
            G.prog.push_back(I);
            return tgt;
        }
        if (is_rng(e.code)) {
            VmIns I = ins(VM_RNG, e.prec, tgt);
            I.p2 = (uint8_t)chain;
            I.a = src.arg;
            const int svl = e.prec == CHAI_F64 ? 2 : 4;
            const int vl = src.vec_len ? (int)src.vec_len : svl;
            if (e.prec == CHAI_F64 ? (vl != 1 && vl != 2) : (vl != 1 && vl != 2 && vl != 4)) {
                err = fail("rng vector length %d", vl);
                return 0;
            }
            I.b = (uint32_t)vl;
            const bool normal = e.code == 110 || e.code == 111;
            const bool lim = !normal && (e.code == 114 || e.code == 115) && (src.s0 != 0.0 || src.s1 != 0.0);
            I.c = (normal ? 1u : 0u) | (lim ? 2u : 0u);
            I.u = src.seed;
            I.d0 = src.s0;
            I.d1 = src.s1;
            VmRngDims R;
            std::memset(&R, 0, sizeof R);
            R.W = e.d.W; R.H = e.d.H; R.S = e.d.S;
            R.vpr = pad_width(e.prec, e.d.W) / (uint64_t)vl;
            R.same = (e.d == G.space) ? 1u : 0u;
            size_t ri = 0;
            while (ri < G.rng.size() && std::memcmp(&G.rng[ri], &R, sizeof R) != 0) ri++;
            if (ri == G.rng.size()) {
                if (G.rng.size() >= 8) { err = fail("more than 8 rng shapes in one kernel"); return 0; }
                G.rng.push_back(R);
            }
            I.e = (uint32_t)ri;
            G.prog.push_back(I);
            return tgt;
        }
        if (is_index(e.code)) {
            VmIns I = ins(VM_INDEX, e.prec, tgt);
            I.p2 = (uint8_t)chain;
            I.a = src.arg ? 1 : 0;
            G.prog.push_back(I);
            return tgt;
        }
        if (is_gather(e.code)) {
            // subscripts are evaluated at the gather's own coordinate (its dims
            // join the broadcast path) in their own types; the data variable
            // is addressed, never loaded as a value
            const int chain_g = chain_for(e.d, false);
            if (chain_g < 0) { err = fail("broadcast nesting too deep for the VM"); return 0; }
            path.push_back(e.d);
            const int sx = emit(s, e.kids[1], tgt, err);
            if (err) return 0;
            int sy = sx;
            if (e.code == 41) {
                sy = emit(s, e.kids[2], sx == tgt ? tgt + 1 : tgt, err);
                if (err) return 0;
            }
            path.pop_back();
            VmIns I = ins(VM_GATHER, e.prec, tgt);
            I.p2 = (uint8_t)chain_g;
            I.a = var_index(s.nodes[e.kids[0]].src->arg);
            I.s1 = (uint8_t)sx;
            I.s2 = (uint8_t)sy;
            I.b = (uint32_t)s.nodes[e.kids[1]].prec;
            I.c = e.code == 41 ? (uint32_t)s.nodes[e.kids[2]].prec : 0u;
            I.e = e.code == 41 ? 2u : 1u;
            G.prog.push_back(I);
            return tgt;
        }
        // operators: evaluate the operands into consecutive slots, converting to
        // the operation's precision where it differs
        int want[3] = {e.prec, e.prec, e.prec};
        if (is_convert(e.code)) want[0] = s.nodes[e.kids[0]].prec;
        if (e.code == 23) want[0] = s.nodes[e.kids[0]].prec;   // condition keeps its type
        int slots[3] = {0, 0, 0};
        int next = tgt;
        path.push_back(e.d);
        for (size_t k = 0; k < e.kids.size(); k++) {
            int sl = emit(s, e.kids[k], next, err);
            if (err) return 0;
            const int kp = s.nodes[e.kids[k]].prec;
            if (kp != want[k]) {
                VmIns C = ins(VM_CVT, want[k], next);
                C.p2 = (uint8_t)kp;
                C.s1 = (uint8_t)sl;
                G.prog.push_back(C);
                sl = next;
            }
            slots[k] = sl;
            if (sl == next) next++;
        }
        path.pop_back();
        VmIns I = ins(VM_MOV, e.prec, tgt);
        I.s1 = (uint8_t)slots[0];
        I.s2 = (uint8_t)slots[1];
        I.s3 = (uint8_t)slots[2];
        I.a = e.code;
        if (is_convert(e.code)) {
            I.op = VM_CVT;
            I.p2 = (uint8_t)s.nodes[e.kids[0]].prec;
        } else if (e.code == 23) {
            I.op = VM_SEL;
            I.p2 = (uint8_t)s.nodes[e.kids[0]].prec;
        } else if (e.kids.size() == 1) {
            I.op = VM_UN;
        } else if (e.kids.size() == 2) {
            I.op = VM_BIN;
        } else {
            I.op = VM_TER;
        }
        G.prog.push_back(I);
        return tgt;
    }
};

}  // namespace

// Reassignment.  Chai arrays are mutable ("modl += gather2_floor(...)" in a
// loop, sample/kirch.cpp), so a trace may assign an OUTPUT or TEMP variable
// several times and every read sees the latest assignment before it.  Each
// assignment but the last gets a fresh TEMP variable (a version, appended
// after the caller's variables); the last writes the variable itself, whose
// buffer is therefore never read before it is written.  Reading such a
// variable before its first assignment is refused.

// Statement splitting (jitstmt/MakeStmt.cpp:579-726).  A reduction or matmul
// below the root of a statement, a matmul operand that is not a variable and
// gather data that is not a variable each become a statement of their own,
// placed before the statement and assigning a fresh TEMP that the statement
// then reads.  Inner expressions split first, so nested reductions work.
static int split_trace(const chai_trace& in, Versioned& V, const chai_trace*& out)
{
    out = &in;
    bool need = false;
    for (uint32_t i = 0; i < in.num_stmts && !need; i++) {
        const chai_trace_stmt& st = in.stmts[i];
        if ((uint64_t)st.first + st.count > in.num_nodes || st.count == 0) return fail("bad node range");
        if (is_cl_stmt(in, st)) continue;
        for (uint32_t k = 0; k < st.count; k++) {
            const uint32_t c = in.nodes[st.first + k].code;
            if (((is_red(c) || is_mm(c)) && k + 1 < st.count) || is_mm(c) || is_gather(c) || is_tr(c)) need = true;
        }
    }
    if (!need) return CHAI_OK;
    V.vars.assign(in.vars, in.vars + in.num_vars);
    // move cur[b, e) into a statement of its own; returns the VAR node reading its result
    auto hoist = [&](std::vector<chai_trace_node>& cur, size_t b, size_t e, chai_trace_node& var) -> int {
        std::vector<chai_trace_node> sub(cur.begin() + b, cur.begin() + e);
        chai_trace view;
        std::memset(&view, 0, sizeof view);
        view.vars = V.vars.data();
        view.num_vars = (uint32_t)V.vars.size();
        Stmt ts;
        const int rc = parse_nodes(view, sub.data(), (uint32_t)sub.size(), ts);
        if (rc) return rc;
        const ENode& r = ts.nodes[ts.root];
        chai_trace_var tv;
        tv.prec = (uint32_t)r.prec;
        tv.role = CHAI_VAR_TEMP;
        tv.W = r.d.W; tv.H = r.d.H; tv.slots = r.d.S;
        const uint32_t v = (uint32_t)V.vars.size();
        V.vars.push_back(tv);
        chai_trace_stmt ns;
        ns.lhs = v;
        ns.first = (uint32_t)V.nodes.size();
        ns.count = (uint32_t)sub.size();
        ns.pad = 0;
        V.nodes.insert(V.nodes.end(), sub.begin(), sub.end());
        V.stmts.push_back(ns);
        std::memset(&var, 0, sizeof var);
        var.code = CHAI_TRACE_VAR;
        var.arg = v;
        var.flags = sub.back().flags & 1u;        // a transposed operand stays transposed
        cur.erase(cur.begin() + b, cur.begin() + e);
        cur.insert(cur.begin() + b, var);
        return CHAI_OK;
    };
    for (uint32_t i = 0; i < in.num_stmts; i++) {
        const chai_trace_stmt& st = in.stmts[i];
        if (is_cl_stmt(in, st)) {            // inline kernel: arguments are variables and scalars already
            chai_trace_stmt ns = st;
            ns.first = (uint32_t)V.nodes.size();
            V.nodes.insert(V.nodes.end(), in.nodes + st.first, in.nodes + st.first + st.count);
            V.stmts.push_back(ns);
            continue;
        }
        std::vector<chai_trace_node> cur;
        std::vector<size_t> starts;                // subtree start in cur, one per stack entry
        for (uint32_t k = 0; k < st.count; k++) {
            const chai_trace_node& n = in.nodes[st.first + k];
            const int ar = node_arity(n.code);
            if (ar < 0) return fail("bytecode %u is not supported by the GPU back end", n.code);
            if ((int)starts.size() < ar) return fail("stack underflow at bytecode %u", n.code);
            const size_t base = starts.size() - (size_t)ar;
            // operands that must be variables: all of a matmul's, a gather's data
            const int nv = is_mm(n.code) ? ar : (is_gather(n.code) || is_tr(n.code)) ? 1 : 0;
            for (int j = nv - 1; j >= 0; j--) {
                const size_t b = starts[base + j];
                const size_t e = (size_t)j + 1 < (size_t)ar ? starts[base + j + 1] : cur.size();
                if (e - b == 1 && cur[b].code == CHAI_TRACE_VAR) continue;
                chai_trace_node var;
                const int rc = hoist(cur, b, e, var);
                if (rc) return rc;
                for (size_t q = base + j + 1; q < starts.size(); q++) starts[q] -= e - b - 1;
            }
            const size_t start = ar ? starts[base] : cur.size();
            starts.resize(base);
            cur.push_back(n);
            starts.push_back(start);
            if ((is_red(n.code) || is_mm(n.code) || is_tr(n.code)) && k + 1 < st.count) {
                chai_trace_node var;
                const int rc = hoist(cur, start, cur.size(), var);
                if (rc) return rc;
            }
        }
        chai_trace_stmt ns = st;
        ns.first = (uint32_t)V.nodes.size();
        ns.count = (uint32_t)cur.size();
        V.nodes.insert(V.nodes.end(), cur.begin(), cur.end());
        V.stmts.push_back(ns);
    }
    V.t.vars = V.vars.data(); V.t.num_vars = (uint32_t)V.vars.size();
    V.t.stmts = V.stmts.data(); V.t.num_stmts = (uint32_t)V.stmts.size();
    V.t.nodes = V.nodes.data(); V.t.num_nodes = (uint32_t)V.nodes.size();
    out = &V.t;
    return CHAI_OK;
}

static int version_trace(const chai_trace& in, Versioned& V, const chai_trace*& out)
{
    out = &in;
    std::vector<uint32_t> nassign(in.num_vars, 0);
    bool multi = false;
    for (uint32_t i = 0; i < in.num_stmts; i++) {
        const uint32_t v = in.stmts[i].lhs;
        if (v >= in.num_vars) return fail("statement assigns variable %u of %u", v, in.num_vars);
        if (is_cl_stmt(in, in.stmts[i])) continue;      // an inline kernel mutates in place
        if (++nassign[v] > 1) multi = true;
    }
    if (!multi) return CHAI_OK;
    V.vars.assign(in.vars, in.vars + in.num_vars);
    V.stmts.assign(in.stmts, in.stmts + in.num_stmts);
    V.nodes.assign(in.nodes, in.nodes + in.num_nodes);
    std::vector<uint32_t> cur(in.num_vars), done(in.num_vars, 0);
    for (uint32_t v = 0; v < in.num_vars; v++) cur[v] = v;
    for (auto& st : V.stmts) {
        if ((uint64_t)st.first + st.count > in.num_nodes) return fail("bad node range");
        for (uint32_t k = 0; k < st.count; k++) {
            chai_trace_node& n = V.nodes[st.first + k];
            if (n.code != CHAI_TRACE_VAR) continue;
            if (n.arg >= in.num_vars) return fail("read of variable %u of %u", n.arg, in.num_vars);
            if (nassign[n.arg] > 1 && !done[n.arg])
                return fail("variable %u is read before its first assignment", n.arg);
            n.arg = cur[n.arg];
        }
        const uint32_t v = st.lhs;
        if (is_cl_stmt(in, st)) {       // mutates the current version of its arrays in place
            st.lhs = cur[v];
            continue;
        }
        if (nassign[v] > 1 && ++done[v] < nassign[v]) {
            chai_trace_var tv = in.vars[v];
            tv.role = CHAI_VAR_TEMP;
            cur[v] = (uint32_t)V.vars.size();
            V.vars.push_back(tv);
            st.lhs = cur[v];
        } else {
            cur[v] = v;
        }
    }
    V.t.vars = V.vars.data(); V.t.num_vars = (uint32_t)V.vars.size();
    V.t.stmts = V.stmts.data(); V.t.num_stmts = (uint32_t)V.stmts.size();
    V.t.nodes = V.nodes.data(); V.t.num_nodes = (uint32_t)V.nodes.size();
    out = &V.t;
    return CHAI_OK;
}

// Statement-to-kernel grouping and VM code generation.
// stmt_cap (empty, or one entry per statement): a fused kernel whose first
// statement has a nonzero cap takes statements only while their nodes total
// at most the cap (see chai_hip_trace_compile)
static int compile_trace(const chai_trace& t, TraceProgram& P, const std::vector<size_t>& stmt_cap)
{
    P.vars.assign(t.vars, t.vars + t.num_vars);
    for (const auto& v : P.vars) {
        if (!valid_prec((int)v.prec)) return fail("variable precision %u", v.prec);
        if (v.W == 0 || v.H == 0 || v.slots == 0) return fail("empty variable");
        if (v.slots > 65535) return fail("more than 65535 slots");
    }
    std::vector<int> writes(t.num_vars, 0);
    P.stmts.resize(t.num_stmts);
    for (uint32_t i = 0; i < t.num_stmts; i++) {
        int rc = parse_stmt(t, t.stmts[i], P.stmts[i]);
        if (rc) return rc;
        if (P.stmts[i].kind == 3) continue;             // in-place kernel, not an assignment
        if (P.vars[P.stmts[i].lhs].role == CHAI_VAR_INPUT) return fail("statement assigns an input");
        if (writes[P.stmts[i].lhs]++) return fail("variable %u assigned twice", P.stmts[i].lhs);
    }
    // group statements
    std::vector<int> group_of_stmt(t.num_stmts, -1);
    std::vector<int> red_group_of_var(t.num_vars, -1);
    std::vector<int> group_of_var(t.num_vars, -1);
    size_t group_nodes = 0, group_cap = 0;
    for (uint32_t i = 0; i < t.num_stmts; i++) {
        const Stmt& s = P.stmts[i];
        bool reads_red = false;
        for (const auto& e : s.nodes)
            if (e.code == CHAI_TRACE_VAR && !P.groups.empty() &&
                red_group_of_var[e.src->arg] == (int)P.groups.size() - 1)
                reads_red = true;
        // a gather reads its data at arbitrary subscripts: data computed in the
        // current kernel must be complete first, so the statement starts a new one
        for (const auto& e : s.nodes)
            if (is_gather(e.code) && !P.groups.empty() &&
                group_of_var[s.nodes[e.kids[0]].src->arg] == (int)P.groups.size() - 1)
                reads_red = true;
        const bool join = s.kind != 2 && s.kind != 3 && s.kind != 4 && !P.groups.empty() && P.groups.back().kind == 0 &&
                          P.groups.back().space == s.space && !reads_red &&
                          (s.kind == 0 || std::count_if(P.groups.back().stmts.begin(), P.groups.back().stmts.end(),
                                                        [&](int q) { return P.stmts[q].kind == 1; }) < VM_MAX_ACC) &&
                          (group_cap == 0 || group_nodes + s.nodes.size() <= group_cap);
        if (!join) {
            Group g;
            g.kind = s.kind == 2 || s.kind == 3 || s.kind == 4 ? s.kind : 0;
            g.space = s.space;
            P.groups.push_back(g);
            group_nodes = 0;
            group_cap = stmt_cap.empty() ? 0 : stmt_cap[i];
        }
        group_nodes += s.nodes.size();
        P.groups.back().stmts.push_back((int)i);
        group_of_stmt[i] = (int)P.groups.size() - 1;
        group_of_var[s.lhs] = (int)P.groups.size() - 1;
        if (s.kind == 1) red_group_of_var[s.lhs] = (int)P.groups.size() - 1;
    }
    // which TEMP variables cross groups (or feed matmul) and need memory
    P.needs_mem.assign(t.num_vars, false);
    std::vector<int> def_group(t.num_vars, -1);
    for (uint32_t i = 0; i < t.num_stmts; i++)
        if (P.stmts[i].kind != 3) def_group[P.stmts[i].lhs] = group_of_stmt[i];
    for (uint32_t i = 0; i < t.num_stmts; i++) {
        const Stmt& s = P.stmts[i];
        for (const auto& e : s.nodes) {
            if (e.code != CHAI_TRACE_VAR) continue;
            const uint32_t v = e.src->arg;
            if (P.vars[v].role != CHAI_VAR_TEMP) continue;
            if (def_group[v] < 0) return fail("temp variable %u read before it is assigned", v);
            if (def_group[v] != group_of_stmt[i] || s.kind == 2 || s.kind == 3 || s.kind == 4) P.needs_mem[v] = true;
        }
        for (const auto& e : s.nodes)
            if (is_gather(e.code)) P.needs_mem[s.nodes[e.kids[0]].src->arg] = true;
        if (s.kind == 1 && P.vars[s.lhs].role == CHAI_VAR_TEMP) P.needs_mem[s.lhs] = true;
        if ((s.kind == 2 || s.kind == 4) && P.vars[s.lhs].role == CHAI_VAR_TEMP) P.needs_mem[s.lhs] = true;
    }
    // VM code per group
    for (auto& G : P.groups) {
        if (G.kind != 0) continue;
        Emitter E{t, P, G};
        // last statement of the group reading each variable: its register is
        // released after that statement
        std::map<uint32_t, int> last_use;
        for (int si : G.stmts)
            for (const auto& e : P.stmts[si].nodes)
                if (e.code == CHAI_TRACE_VAR) last_use[e.src->arg] = si;
        auto free_dead = [&](int si) {
            for (auto it = E.reg_of_var.begin(); it != E.reg_of_var.end();) {
                auto lu = last_use.find(it->first);
                if (lu != last_use.end() && lu->second <= si) {
                    E.temp_used[it->second] = false;
                    it = E.reg_of_var.erase(it);
                } else {
                    ++it;
                }
            }
        };
        for (int si : G.stmts) {
            const Stmt& s = P.stmts[si];
            int err = 0;
            const ENode& r = s.nodes[s.root];
            if (s.kind == 1) {
                int sl;
                int p = r.prec;
                if (r.code == 130) {      // dot = sum of the promoted products
                    ENode mul = r;
                    (void)mul;
                    int a = E.emit(s, r.kids[0], 0, err);
                    if (err) return err;
                    const int ap = s.nodes[r.kids[0]].prec;
                    if (ap != p) {
                        VmIns C = E.ins(VM_CVT, p, 0); C.p2 = (uint8_t)ap; C.s1 = (uint8_t)a;
                        G.prog.push_back(C); a = 0;
                    }
                    const int nb = a == 0 ? 1 : 0;
                    int b = E.emit(s, r.kids[1], nb, err);
                    if (err) return err;
                    const int bp = s.nodes[r.kids[1]].prec;
                    if (bp != p) {
                        VmIns C = E.ins(VM_CVT, p, nb); C.p2 = (uint8_t)bp; C.s1 = (uint8_t)b;
                        G.prog.push_back(C); b = nb;
                    }
                    VmIns M = E.ins(VM_BIN, p, 0); M.a = 13; M.s1 = (uint8_t)a; M.s2 = (uint8_t)b;
                    G.prog.push_back(M);
                    sl = 0;
                } else {
                    sl = E.emit(s, r.kids[0], 0, err);
                    if (err) return err;
                }
                VmIns A = E.ins(VM_ACC, p, 0);
                A.a = (uint32_t)G.red_vars.size();
                A.s1 = (uint8_t)sl;
                G.prog.push_back(A);
                G.red_vars.push_back(s.lhs);
                G.red_div.push_back(r.code == 131 ? (double)(s.space.W * s.space.H) : 1.0);
                free_dead(si);
                continue;
            }
            int sl = E.emit(s, s.root, 0, err);
            if (err) return err;
            const chai_trace_var& lv = P.vars[s.lhs];
            if ((int)lv.prec != r.prec) {
                VmIns C = E.ins(VM_CVT, (int)lv.prec, 0);
                C.p2 = (uint8_t)r.prec;
                C.s1 = (uint8_t)sl;
                G.prog.push_back(C);
                sl = 0;
            }
            if (lv.role == CHAI_VAR_OUTPUT || P.needs_mem[s.lhs]) {
                VmIns S = E.ins(VM_STORE, (int)lv.prec, 0);
                S.a = E.var_index(s.lhs);
                S.s1 = (uint8_t)sl;
                G.prog.push_back(S);
            }
            // later statements of this group read it from a register
            bool used_later = false;
            for (int sj : G.stmts)
                if (sj > si)
                    for (const auto& e : P.stmts[sj].nodes)
                        if (e.code == CHAI_TRACE_VAR && e.src->arg == s.lhs) used_later = true;
            if (used_later) {
                const int ts = E.alloc_temp(0);
                if (ts < 0) return fail("too many live temporaries for the VM registers");
                VmIns Mv = E.ins(VM_MOV, (int)lv.prec, ts);
                Mv.s1 = (uint8_t)sl;
                G.prog.push_back(Mv);
                E.reg_of_var[s.lhs] = ts;
            }
            free_dead(si);
        }
        if (G.var_of_slot.size() > 24) return fail("more than 24 arrays in one fused kernel");
        P.ninstr += (int)G.prog.size();
    }
    return CHAI_OK;
}

static bool jit_enabled();
static size_t jit_max_instructions();
static size_t jit_rolled_size(const std::vector<VmIns>& prog);
static void jit_build(const TraceProgram& P, Group& G);

// device side of compilation: upload VM programs, JIT-compile when enabled
static int materialize(TraceProgram& P)
{
    for (auto& G : P.groups) {
        if (G.kind != 0) continue;
        CHAI_HIP_TRY(hipMalloc((void**)&G.dprog, sizeof(VmIns) * G.prog.size()));
        CHAI_HIP_TRY(hipMemcpy(G.dprog, G.prog.data(), sizeof(VmIns) * G.prog.size(), hipMemcpyHostToDevice));
        if (!jit_enabled()) continue;
        if (jit_rolled_size(G.prog) <= jit_max_instructions()) {
            jit_build(P, G);
        } else {
            // hipRTC compile time grows faster than linearly with the straight-line
            // code (measured: 0.5 s at 16 statements, 7.4 s at 512, 98 s at 2048), so
            // a kernel this long runs on the VM, which needs no compile
            char b[160];
            snprintf(b, sizeof b, "not compiled: %zu VM instructions (after loop rolling) exceed the JIT limit "
                     "of %zu (CHAI_TRACE_JIT_MAX); runs on the trace VM", jit_rolled_size(G.prog),
                     jit_max_instructions());
            G.jit_log = b;
        }
    }
    return CHAI_OK;
}

static int run_matmul(const TraceProgram& P, const Stmt& s, void* const* bufs, hipStream_t st)
{
    const ENode& r = s.nodes[s.root];
    const uint32_t va = s.nodes[r.kids[0]].src->arg, vb = s.nodes[r.kids[1]].src->arg;
    const bool ta = s.nodes[r.kids[0]].src->flags & 1, tb = s.nodes[r.kids[1]].src->flags & 1;
    const chai_trace_var &A = P.vars[va], &B = P.vars[vb], &Cv = P.vars[s.lhs];
    const bool general = r.code == 121;
    const uint64_t S = Cv.slots;                      // instances: A / B have 1 or S
    if (general && s.nodes[r.kids[2]].src->arg != s.lhs) {
        // C operand is a different variable: copy it into the output first
        // (one instance of it serves every slot)
        const uint32_t vc = s.nodes[r.kids[2]].src->arg;
        const size_t bytes = prec_size((int)Cv.prec) * Cv.W * Cv.H;
        if (P.vars[vc].prec != Cv.prec) return fail("matmulG C operand precision differs from its result");
        for (uint64_t k = 0; k < S; k++)
            CHAI_HIP_TRY(hipMemcpyAsync((char*)bufs[s.lhs] + k * bytes,
                                        (const char*)bufs[vc] + (P.vars[vc].slots == 1 ? 0 : k * bytes), bytes,
                                        hipMemcpyDeviceToDevice, st));
    }
    // row-major arrays: W = columns, H = rows.  A vector has H == 1.
    const uint64_t am = ta ? A.W : A.H, ak = ta ? A.H : A.W;
    const uint64_t bk = tb ? B.W : B.H, bn = tb ? B.H : B.W;
    const int pc = (int)Cv.prec;
    const uint64_t sa = A.slots == 1 ? 0 : A.W * A.H, sb = B.slots == 1 ? 0 : B.W * B.H, sc = Cv.W * Cv.H;
    if (B.H == 1 && !tb && A.H > 1) {                 // matrix * vector
        if (ak != B.W) return fail("matrix-vector inner dimension mismatch");
        return chai_hip_gemv((int)A.prec, (int)B.prec, pc, ta, ta ? A.H : A.H, A.W, bufs[va], A.W,
                             bufs[vb], bufs[s.lhs], general, r.src->s0, r.src->s1, S, sa == 0, sa, sb, sc, st);
    }
    if (A.H == 1 && !ta && B.H > 1) {                 // vector * matrix = B^T x
        if (A.W != bk) return fail("vector-matrix inner dimension mismatch");
        return chai_hip_gemv((int)B.prec, (int)A.prec, pc, !tb, B.H, B.W, bufs[vb], B.W, bufs[va],
                             bufs[s.lhs], general, r.src->s0, r.src->s1, S, sb == 0, sb, sa, sc, st);
    }
    if (ak != bk) return fail("matmul inner dimension mismatch");
    if (Cv.H != am || Cv.W != bn) return fail("matmul result has the wrong shape");
    chai_gemm_desc d;
    std::memset(&d, 0, sizeof d);
    d.prec_a = (int)A.prec; d.prec_b = (int)B.prec; d.prec_c = pc;
    d.trans_a = ta; d.trans_b = tb;
    d.M = am; d.N = bn; d.K = ak;
    d.lda = A.W; d.ldb = B.W; d.ldc = Cv.W;
    d.general = general; d.alpha = r.src->s0; d.beta = r.src->s1;
    d.batch = S;
    d.same_a = sa == 0 && S > 1; d.same_b = sb == 0 && S > 1;
    d.stride_a = sa; d.stride_b = sb; d.stride_c = sc;
    d.mode = CHAI_GEMM_AUTO;
    d.variant = -1;
    return chai_hip_gemm(&d, bufs[va], bufs[vb], bufs[s.lhs], st);
}

// T = transpose(V) (interp/InterpTranspose.cpp:13-60): out[s][x][y] =
// in[s][y][x].  HBM-bound: 64 x 64 tiles through LDS (one padded column, so
// the transposed reads are conflict free), reads and writes coalesced.
template <typename T>
__global__ __launch_bounds__(256) void transpose_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                        uint64_t W, uint64_t H)
{
    __shared__ T tile[64][65];
    const uint64_t slot = blockIdx.z;
    in += slot * W * H;
    out += slot * W * H;
    const uint64_t x0 = (uint64_t)blockIdx.x * 64, y0 = (uint64_t)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const uint64_t y = y0 + r, x = x0 + tx;
        if (y < H && x < W) tile[r][tx] = in[y * W + x];
    }
    __syncthreads();
    // output row = input column x0 + r, output column = input row y0 + tx
    for (int r = ty; r < 64; r += 4) {
        const uint64_t oy = x0 + r, ox = y0 + tx;
        if (oy < W && ox < H) out[oy * H + ox] = tile[tx][r];
    }
}

static int run_transpose(const TraceProgram& P, const Stmt& s, void* const* bufs, hipStream_t st)
{
    const ENode& r = s.nodes[s.root];
    const uint32_t vi = s.nodes[r.kids[0]].src->arg;
    const chai_trace_var& V = P.vars[vi];
    const dim3 grid(ceil_div_u(V.W, 64), ceil_div_u(V.H, 64), (unsigned)V.slots);
    if (grid.y > 65535) return fail("transpose of %llu rows", (unsigned long long)V.H);
    if (V.prec == CHAI_F64)
        transpose_kernel<uint64_t><<<grid, 256, 0, st>>>((const uint64_t*)bufs[vi], (uint64_t*)bufs[s.lhs], V.W, V.H);
    else
        transpose_kernel<uint32_t><<<grid, 256, 0, st>>>((const uint32_t*)bufs[vi], (uint32_t*)bufs[s.lhs], V.W, V.H);
    CHAI_LAUNCH_CHECK("transpose");
    return CHAI_OK;
}

// an inline OpenCL kernel statement on the trace's buffers
static int run_cl(const TraceProgram& P, const Stmt& s, void* const* bufs, hipStream_t st)
{
    const ENode& r = s.nodes[s.root];
    std::vector<chai_cl_arg> args(r.kids.size());
    for (size_t k = 0; k < r.kids.size(); k++) {
        const ENode& e = s.nodes[r.kids[k]];
        chai_cl_arg& a = args[k];
        std::memset(&a, 0, sizeof a);
        if (e.code == CHAI_TRACE_VAR) {
            a.kind = CHAI_CL_ARG_ARRAY;
            a.prec = (uint32_t)e.prec;
            a.v.ptr = bufs[e.src->arg];
        } else if (e.code == CHAI_TRACE_CL_LOCAL) {
            a.kind = CHAI_CL_ARG_LOCAL;
            a.v.count = e.src->W;
        } else {
            a.kind = CHAI_CL_ARG_SCALAR;
            a.prec = (uint32_t)e.prec;
            const double v = e.src->s0;
            if (e.prec == CHAI_U32) a.v.u = (uint32_t)(int64_t)v;
            else if (e.prec == CHAI_I32) a.v.i = (int32_t)v;
            else if (e.prec == CHAI_F32) a.v.f = (float)v;
            else a.v.d = v;
        }
    }
    const uint64_t g[2] = {r.src->W, r.src->H}, l[2] = {(uint64_t)r.src->s0, (uint64_t)r.src->s1};
    return cl_launch_kernel((void*)r.src->seed, args.data(), (uint32_t)args.size(), r.src->vec_len, g, l, st);
}

// ------------------------------------------------------------ trace JIT
// The typed VM program of a group, emitted as straight-line HIP: every VM
// instruction becomes one statement on register variables R0..R11, array
// shapes and instruction fields become literals, so the compiler specialises
// every load, op and store (the role the reference's OpenCL text emission
// plays, enqueue/EnqueueIdSpace.cpp:19-320).  Semantics are the VM's own
// device functions (trace_vm.hpp), so JIT and VM agree bit for bit.
static int g_jit_mode = -1;   // -1: read CHAI_TRACE_JIT once; 0 off; 1 on

// longest fused kernel (VM instructions) compiled by the JIT; longer ones run on the VM
static size_t jit_max_instructions()
{
    static const size_t lim = [] {
        const char* e = getenv("CHAI_TRACE_JIT_MAX");
        const long v = e ? strtol(e, nullptr, 10) : 0;
        return v > 0 ? (size_t)v : (size_t)1024;
    }();
    return lim;
}

static bool jit_enabled()
{
    if (g_jit_mode < 0) {
        const char* e = getenv("CHAI_TRACE_JIT");
        g_jit_mode = (e && e[0] == '0') ? 0 : 1;
    }
    return g_jit_mode == 1;
}

static std::string hexd(double d)
{
    char b[64];
    snprintf(b, sizeof b, "%a", d);
    if (d != d) return "__builtin_nan(\"\")";
    if (d == __builtin_inf()) return "__builtin_inf()";
    if (d == -__builtin_inf()) return "(-__builtin_inf())";
    return b;
}

// ------------------------------------------------------------ loop rolling
// A traced C++ loop reaches the compiler unrolled (the reference folds such
// runs into loops, jitbc/ByteTrace.cpp:211-387).  In a fused kernel's VM
// program the repetitions are equal up to a fixed permutation pi of register
// slots: repetition j + 1 is repetition j with every register r renamed pi(r).
// The JIT then emits the first repetition once in a loop and, after each
// iteration, moves R[r] <- R[pi(r)]; the values stay those of the unrolled
// program (R_loop[r] = R_unrolled[pi^j(r)] before iteration j), and a final
// move restores the unrolled register assignment.  The VM runs the unrolled
// program as it is.
struct JitSeg {
    size_t start;
    int L, k;                                   // pattern length, repetitions (k <= 1: one instruction)
    std::array<int, VM_NSLOTS> pi;
};

// register operands an instruction reads or writes, in a fixed order
static int ins_regs(const VmIns& I, int* r)
{
    switch (I.op) {
        case VM_LOAD: case VM_IMM: case VM_RNG: case VM_INDEX: r[0] = I.dst; return 1;
        case VM_CVT: case VM_UN: case VM_MOV: r[0] = I.dst; r[1] = I.s1; return 2;
        case VM_BIN: case VM_GATHER: r[0] = I.dst; r[1] = I.s1; r[2] = I.s2; return 3;
        case VM_TER: case VM_SEL: r[0] = I.dst; r[1] = I.s1; r[2] = I.s2; r[3] = I.s3; return 4;
        case VM_STORE: case VM_ACC: r[0] = I.s1; return 1;
        default: r[0] = I.dst; r[1] = I.s1; return 2;
    }
}

static bool same_but_regs(const VmIns& x, const VmIns& y)
{
    return x.op == y.op && x.p == y.p && x.p2 == y.p2 && x.a == y.a && x.b == y.b && x.c == y.c && x.e == y.e &&
           x.u == y.u && std::memcmp(&x.d0, &y.d0, 8) == 0 && std::memcmp(&x.d1, &y.d1, 8) == 0;
}

// extend the partial map m (-1 = unset) from prog[a, a+L) to prog[b, b+L); false on a mismatch
static bool match_rename(const std::vector<VmIns>& prog, size_t a, size_t b, int L, std::array<int, VM_NSLOTS>& m,
                         std::array<int, VM_NSLOTS>& mi)
{
    for (int q = 0; q < L; q++) {
        const VmIns &x = prog[a + q], &y = prog[b + q];
        if (!same_but_regs(x, y)) return false;
        int rx[4], ry[4];
        const int n = ins_regs(x, rx);
        ins_regs(y, ry);
        for (int t = 0; t < n; t++) {
            if (rx[t] >= VM_NSLOTS || ry[t] >= VM_NSLOTS) return false;
            if (m[rx[t]] < 0 && mi[ry[t]] < 0) { m[rx[t]] = ry[t]; mi[ry[t]] = rx[t]; }
            else if (m[rx[t]] != ry[t] || mi[ry[t]] != rx[t]) return false;
        }
    }
    return true;
}

static std::vector<JitSeg> jit_plan(const std::vector<VmIns>& prog)
{
    std::vector<JitSeg> plan;
    const size_t n = prog.size();
    size_t i = 0;
    while (i < n) {
        JitSeg best{i, 1, 1, {}};
        size_t best_cov = 1;
        for (int L = 1; L <= 64 && i + 2 * (size_t)L <= n; L++) {
            std::array<int, VM_NSLOTS> m, mi;
            m.fill(-1);
            mi.fill(-1);
            if (!match_rename(prog, i, i + L, L, m, mi)) continue;
            // complete the partial renaming to a permutation of all slots
            std::array<int, VM_NSLOTS> pi = m;
            std::vector<int> free_dst;
            for (int r = 0; r < VM_NSLOTS; r++) if (mi[r] < 0) free_dst.push_back(r);
            size_t f = 0;
            for (int r = 0; r < VM_NSLOTS; r++) if (pi[r] < 0) pi[r] = free_dst[f++];
            // further repetitions must follow the same permutation
            int k = 2;
            while (i + (size_t)(k + 1) * L <= n) {
                std::array<int, VM_NSLOTS> m2, mi2;
                m2.fill(-1);
                mi2.fill(-1);
                if (!match_rename(prog, i + (size_t)(k - 1) * L, i + (size_t)k * L, L, m2, mi2)) break;
                bool ok = true;
                for (int r = 0; r < VM_NSLOTS && ok; r++) if (m2[r] >= 0 && m2[r] != pi[r]) ok = false;
                if (!ok) break;
                k++;
            }
            const size_t cov = (size_t)k * L;
            if (cov > best_cov && cov >= 8) { best = JitSeg{i, L, k, pi}; best_cov = cov; }
        }
        plan.push_back(best);
        i += best_cov;
    }
    return plan;
}

// R[r] <- R[src[r]] for every slot at once
static std::string perm_code(const std::array<int, VM_NSLOTS>& src)
{
    std::string s, t;
    char b[96];
    for (int r = 0; r < VM_NSLOTS; r++)
        if (src[r] != r) { snprintf(b, sizeof b, " const V4 t%d = R%d;", r, src[r]); s += b; }
    if (s.empty()) return "";
    for (int r = 0; r < VM_NSLOTS; r++)
        if (src[r] != r) { snprintf(b, sizeof b, " R%d = t%d;", r, r); t += b; }
    return "        {" + s + t + " }\n";
}

// instructions the JIT source holds after rolling
static size_t jit_rolled_size(const std::vector<VmIns>& prog)
{
    size_t n = 0;
    for (const JitSeg& g : jit_plan(prog)) n += g.k <= 1 ? 1 : (size_t)g.L;
    return n;
}

static std::string jit_source(const TraceProgram& P, const Group& G)
{
    const uint64_t W = G.space.W, H = G.space.H, S = G.space.S;
    const uint64_t vpr = (W + 3) / 4, vps = vpr * H;
    std::string s;
    char b[1024];
    s += "#include \"trace_vm.hpp\"\n";
    for (size_t k = 0; k < G.var_of_slot.size(); k++) {
        snprintf(b, sizeof b, "// L.vars[%zu] = trace variable %u\n", k, G.var_of_slot[k]);
        s += b;
    }
    s += "extern \"C\" __global__ __launch_bounds__(256) void chai_trace_kernel(chai_hip::VmLaunchArgs L)\n{\n";
    s += "    using namespace chai_hip;\n    const uint64_t slot = blockIdx.y;\n";
    for (int r = 0; r < VM_NSLOTS; r++) { snprintf(b, sizeof b, "    V4 R%d = {{0, 0, 0, 0}};\n", r); s += b; }
    s += "    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;\n";
    snprintf(b, sizeof b,
             "    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < %lluull; v += (uint64_t)gridDim.x * 256) {\n"
             "        const uint64_t y = v / %lluull;\n        const uint64_t x0 = (v - y * %lluull) * 4;\n"
             "        const int nval = (%lluull - x0) < 4 ? (int)(%lluull - x0) : 4;\n",
             (unsigned long long)vps, (unsigned long long)vpr, (unsigned long long)vpr,
             (unsigned long long)W, (unsigned long long)W);
    s += b;
    // one VM instruction as a line of HIP
    const auto emit_ins = [&](const VmIns& I) -> std::string {
        const auto var_lit = [&](uint32_t k, bool in) {
            const uint32_t v = G.var_of_slot[k];
            const chai_trace_var& tv = P.vars[v];
            const int same = (tv.W == W && tv.H == H && tv.slots == S) ? 1 : 0;
            char t[256];
            snprintf(t, sizeof t, "VmVar{%s, %s, %lluull, %lluull, %lluull, %uu, %uu}",
                     in ? ("L.vars[" + std::to_string(k) + "].in").c_str() : "nullptr",
                     in ? "nullptr" : ("L.vars[" + std::to_string(k) + "].out").c_str(),
                     (unsigned long long)tv.W, (unsigned long long)tv.H, (unsigned long long)tv.slots,
                     tv.prec, same);
            return std::string(t);
        };
        const auto chain_lit = [&](uint32_t k) {
            if (k == 0) return std::string("VmChain{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, 0u, 0u, 0u, 0u}");
            const VmChain& c = G.chains[k - 1];
            char t[400];
            snprintf(t, sizeof t,
                     "VmChain{{%lluull, %lluull, %lluull}, {%lluull, %lluull, %lluull}, {%lluull, %lluull, %lluull}, "
                     "%uu, %uu, %uu, 0u}",
                     (unsigned long long)c.W[0], (unsigned long long)c.W[1], (unsigned long long)c.W[2],
                     (unsigned long long)c.H[0], (unsigned long long)c.H[1], (unsigned long long)c.H[2],
                     (unsigned long long)c.S[0], (unsigned long long)c.S[1], (unsigned long long)c.S[2], c.nW, c.nH,
                     c.nS);
            return std::string(t);
        };
        switch (I.op) {
            case VM_LOAD:
                if (I.p2)
                    snprintf(b, sizeof b, "        R%d = vm_load_c(%s, %s, slot, y, x0, nval);\n", I.dst,
                             var_lit(I.a, true).c_str(), chain_lit(I.p2).c_str());
                else
                    snprintf(b, sizeof b, "        R%d = vm_load(%s, slot, y, x0, nval, %lluull, %lluull);\n", I.dst,
                             var_lit(I.a, true).c_str(), (unsigned long long)W, (unsigned long long)H);
                break;
            case VM_IMM:
                snprintf(b, sizeof b, "        R%d.l[0] = R%d.l[1] = R%d.l[2] = R%d.l[3] = %lluull;\n", I.dst, I.dst,
                         I.dst, I.dst, (unsigned long long)I.u);
                break;
            case VM_RNG: {
                const VmRngDims& R = G.rng[I.e];
                snprintf(b, sizeof b,
                         "        R%d = vm_rng(RngArgs{%uu, %uu, %uu, %uu, %lluull, %s, %s}, "
                         "VmRngDims{%lluull, %lluull, %lluull, %lluull, %uu, 0u}, slot, y, x0, nval, %s);\n",
                         I.dst, I.a, I.b, I.c, (unsigned)I.p, (unsigned long long)I.u, hexd(I.d0).c_str(),
                         hexd(I.d1).c_str(), (unsigned long long)R.W, (unsigned long long)R.H,
                         (unsigned long long)R.S, (unsigned long long)R.vpr, R.same, chain_lit(I.p2).c_str());
                break;
            }
            case VM_INDEX:
                if (I.p2)
                    snprintf(b, sizeof b,
                             "        { const VmChain C = %s; for (int i = 0; i < 4; i++) R%d.l[i] = "
                             "from_double(%u, (double)(%s)); }\n",
                             chain_lit(I.p2).c_str(), I.dst, (unsigned)I.p,
                             I.a == 0 ? "chain_mod(x0 + i, C.W, C.nW)" : "chain_mod(y, C.H, C.nH)");
                else
                    snprintf(b, sizeof b,
                             "        for (int i = 0; i < 4; i++) R%d.l[i] = from_double(%u, (double)(%s));\n",
                             I.dst, (unsigned)I.p, I.a == 0 ? "x0 + i" : "y");
                break;
            case VM_CVT:
                snprintf(b, sizeof b,
                         "        { V4 t = R%d; for (int i = 0; i < 4; i++) R%d.l[i] = from_double(%u, as_double(%u, t.l[i])); }\n",
                         I.s1, I.dst, (unsigned)I.p, (unsigned)I.p2);
                break;
            case VM_UN:
                snprintf(b, sizeof b, "        R%d = vm_un(%uu, %u, R%d);\n", I.dst, I.a, (unsigned)I.p, I.s1);
                break;
            case VM_BIN:
                snprintf(b, sizeof b, "        R%d = vm_bin(%uu, %u, R%d, R%d);\n", I.dst, I.a, (unsigned)I.p, I.s1, I.s2);
                break;
            case VM_TER:
                snprintf(b, sizeof b, "        R%d = vm_ter(%uu, %u, R%d, R%d, R%d);\n", I.dst, I.a, (unsigned)I.p, I.s1,
                         I.s2, I.s3);
                break;
            case VM_SEL:
                snprintf(b, sizeof b,
                         "        { V4 c = R%d, p = R%d, q = R%d; for (int i = 0; i < 4; i++) R%d.l[i] = as_double(%u, c.l[i]) != 0.0 ? p.l[i] : q.l[i]; }\n",
                         I.s1, I.s2, I.s3, I.dst, (unsigned)I.p2);
                break;
            case VM_GATHER:
                snprintf(b, sizeof b, "        R%d = vm_gather(%s, %s, slot, y, %d, %uu, R%d, %uu, R%d, nval);\n",
                         I.dst, var_lit(I.a, true).c_str(), chain_lit(I.p2).c_str(), I.e == 2 ? 1 : 0, I.b, I.s1,
                         I.c, I.s2);
                break;
            case VM_STORE:
                snprintf(b, sizeof b, "        vm_store(%s, R%d, slot, y, x0, nval, %lluull, %lluull);\n",
                         var_lit(I.a, false).c_str(), I.s1, (unsigned long long)W, (unsigned long long)H);
                break;
            case VM_ACC:
                snprintf(b, sizeof b,
                         "        { double t = 0.0; for (int i = 0; i < 4; i++) if (i < nval) t += as_double(%u, R%d.l[i]); acc%u += t; }\n",
                         (unsigned)I.p, I.s1, I.a);
                break;
            default:
                snprintf(b, sizeof b, "        R%d = R%d;\n", I.dst, I.s1);
                break;
        }
        return std::string(b);
    };
    // repeated runs (a traced loop) become one copy of the pattern in a for loop
    const std::vector<JitSeg> plan = jit_plan(G.prog);
    for (const JitSeg& g : plan) {
        if (g.k <= 1) {
            s += emit_ins(G.prog[g.start]);
            continue;
        }
        snprintf(b, sizeof b, "        for (int it = 0; it < %d; ++it) {\n", g.k);
        s += b;
        for (int q = 0; q < g.L; q++) s += "    " + emit_ins(G.prog[g.start + q]);
        s += "    " + perm_code(g.pi);         // R[r] <- R[pi(r)]
        s += "        }\n";
        // after k iterations R_loop[r] = R_unrolled[pi^k(r)]: put values back
        std::array<int, VM_NSLOTS> pk;
        for (int r = 0; r < VM_NSLOTS; r++) pk[r] = r;
        for (int it = 0; it < g.k; it++)
            for (int r = 0; r < VM_NSLOTS; r++) pk[r] = g.pi[pk[r]];
        std::array<int, VM_NSLOTS> inv;
        for (int r = 0; r < VM_NSLOTS; r++) inv[pk[r]] = r;
        s += perm_code(inv);                       // R[q] <- R_loop[inv(q)]
    }
    s += "    }\n";
    if (!G.red_vars.empty()) {
        snprintf(b, sizeof b, "    const uint64_t base = ((uint64_t)slot * gridDim.x + blockIdx.x) * %zu;\n", G.red_vars.size());
        s += b;
        for (size_t k = 0; k < G.red_vars.size(); k++) {
            snprintf(b, sizeof b,
                     "    { const double r = block_sum_f64<256>(acc%zu); if (threadIdx.x == 0) L.partial[base + %zu] = r; }\n", k, k);
            s += b;
        }
    }
    s += "}\n";
    return s;
}

static const char* kJitPrelude =
    "typedef __hip_internal::uint8_t uint8_t;\ntypedef __hip_internal::uint16_t uint16_t;\n"
    "typedef __hip_internal::uint32_t uint32_t;\ntypedef __hip_internal::uint64_t uint64_t;\n"
    "typedef __hip_internal::int32_t int32_t;\ntypedef __hip_internal::int64_t int64_t;\n"
    "typedef unsigned long uintptr_t;\n#define INT32_MAX 2147483647\n#define INT32_MIN (-2147483647 - 1)\n"
    "enum { CHAI_U32 = 0, CHAI_I32 = 1, CHAI_F32 = 4, CHAI_F64 = 8, CHAI_RNG_PHILOX = 1230, CHAI_RNG_THREEFRY = 1231 };\n";

// Compile one group's source; on failure the group keeps running on the VM.
static void jit_build(const TraceProgram& P, Group& G)
{
    G.jit_src = jit_source(P, G);
    const std::string full = std::string(kJitPrelude) + G.jit_src;
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, full.c_str(), "chai_trace.hip", kJitNumHeaders, kJitHeaders, kJitHeaderNames) !=
        HIPRTC_SUCCESS) {
        G.jit_log = "hiprtcCreateProgram failed";
        return;
    }
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, 8, opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (ls > 1) {
        G.jit_log.resize(ls);
        hiprtcGetProgramLog(prog, &G.jit_log[0]);
    }
    if (rc == HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetCodeSize(prog, &n);
        std::vector<char> code(n);
        hiprtcGetCode(prog, code.data());
        if (hipModuleLoadData(&G.mod, code.data()) == hipSuccess &&
            hipModuleGetFunction(&G.fn, G.mod, "chai_trace_kernel") == hipSuccess) {
            // compiled and loaded
        } else {
            G.fn = nullptr;
            G.jit_log += "\nhipModuleLoadData / GetFunction failed";
        }
    }
    hiprtcDestroyProgram(&prog);
}

static int launch_program(const TraceProgram& P, void* const* user, hipStream_t st)
{
    const uint32_t nv = (uint32_t)P.vars.size();
    std::vector<void*> bufs(nv, nullptr);
    std::vector<void*> owned;
    for (uint32_t v = 0; v < nv; v++) {
        bufs[v] = (user && v < P.nuser) ? user[v] : nullptr;
        if (P.vars[v].role != CHAI_VAR_TEMP && !bufs[v]) return fail("variable %u has no buffer", v);
        if (P.vars[v].role == CHAI_VAR_TEMP && P.needs_mem[v] && !bufs[v]) {
            const size_t bytes = prec_size((int)P.vars[v].prec) * P.vars[v].W * P.vars[v].H * P.vars[v].slots;
            CHAI_HIP_TRY(hipMallocAsync(&bufs[v], bytes, st));
            owned.push_back(bufs[v]);
        }
    }
    int rc = CHAI_OK;
    for (const auto& G : P.groups) {
        if (G.kind == 2) {
            rc = run_matmul(P, P.stmts[G.stmts[0]], bufs.data(), st);
            if (rc) break;
            continue;
        }
        if (G.kind == 3) {
            rc = run_cl(P, P.stmts[G.stmts[0]], bufs.data(), st);
            if (rc) break;
            continue;
        }
        if (G.kind == 4) {
            rc = run_transpose(P, P.stmts[G.stmts[0]], bufs.data(), st);
            if (rc) break;
            continue;
        }
        VmLaunchArgs L;
        std::memset(&L, 0, sizeof L);
        L.prog = G.dprog;
        L.nins = (uint32_t)G.prog.size();
        L.nacc = (uint32_t)G.red_vars.size();
        L.W = G.space.W; L.H = G.space.H; L.S = G.space.S;
        L.vpr = (G.space.W + 3) / 4;
        L.vps = L.vpr * G.space.H;
        for (size_t k = 0; k < G.var_of_slot.size(); k++) {
            const uint32_t v = G.var_of_slot[k];
            VmVar& V = L.vars[k];
            V.in = bufs[v];
            V.out = bufs[v];
            V.W = P.vars[v].W; V.H = P.vars[v].H; V.S = P.vars[v].slots;
            V.prec = P.vars[v].prec;
            V.same = (V.W == L.W && V.H == L.H && V.S == L.S) ? 1u : 0u;
        }
        for (size_t k = 0; k < G.rng.size(); k++) L.rng[k] = G.rng[k];
        for (size_t k = 0; k < G.chains.size(); k++) L.chains[k] = G.chains[k];
        // one vector per thread for elementwise kernels (a grid-stride loop only past
        // 2^22 blocks); reductions keep at most 1024 partials per slot
        uint64_t bps = (L.vps + VM_THREADS - 1) / VM_THREADS;
        const uint64_t cap = L.nacc ? 1024 : (1ull << 22);
        if (bps > cap) bps = cap;
        if (bps < 1) bps = 1;
        double* partial = nullptr;
        if (L.nacc) {
            CHAI_HIP_TRY(hipMallocAsync((void**)&partial, sizeof(double) * L.S * bps * L.nacc, st));
            L.partial = partial;
        }
        hipError_t e;
        if (G.fn) {
            void* args[] = {&L};
            e = hipModuleLaunchKernel(G.fn, (unsigned)bps, (unsigned)L.S, 1, VM_THREADS, 1, 1, 0, st, args, nullptr);
        } else {
            trace_vm_kernel<<<dim3((unsigned)bps, (unsigned)L.S), VM_THREADS, 0, st>>>(L);
            e = hipGetLastError();
        }
        if (e == hipSuccess && L.nacc) {
            RedOut o[4];
            std::memset(o, 0, sizeof o);
            for (uint32_t k = 0; k < L.nacc; k++) {
                o[k].out = bufs[G.red_vars[k]];
                o[k].prec = P.vars[G.red_vars[k]].prec;
                o[k].divisor = G.red_div[k];
            }
            trace_reduce_final<<<dim3((unsigned)L.S, L.nacc), VM_THREADS, 0, st>>>(partial, (uint32_t)bps, L.nacc,
                                                                                   o[0], o[1], o[2], o[3]);
            e = hipGetLastError();
        }
        if (partial) (void)hipFreeAsync(partial, st);
        if (e != hipSuccess) { rc = hip_fail(e, "trace_vm launch"); break; }
    }
    for (void* p : owned) (void)hipFreeAsync(p, st);
    return rc;
}

// A fused kernel longer than the JIT limit (after loop rolling) would run on
// the trace VM, an order of magnitude slower than compiled code.  When such a
// kernel holds several statements, its statements are regrouped into shorter
// kernels (cap on their nodes, halved until each fits or holds one
// statement); the variables that then cross kernels get memory as any other
// crossing temporary.  Kernels of one statement stay whole.
static int fit_jit_limit(const chai_trace& t, TraceProgram& P)
{
    std::vector<size_t> cap;
    for (int attempt = 0; attempt < 12; attempt++) {
        bool over = false;
        std::vector<size_t> next = cap.empty() ? std::vector<size_t>(t.num_stmts, 0) : cap;
        for (const Group& G : P.groups) {
            if (G.kind != 0 || G.stmts.size() < 2 || jit_rolled_size(G.prog) <= jit_max_instructions()) continue;
            size_t nodes = 0;
            for (int si : G.stmts) nodes += P.stmts[si].nodes.size();
            const size_t c = std::max<size_t>(1, (cap.empty() || !cap[G.stmts[0]] ? nodes : cap[G.stmts[0]]) / 2);
            for (int si : G.stmts) next[si] = c;
            over = true;
        }
        if (!over) return CHAI_OK;
        cap = next;
        P.stmts.clear();
        P.groups.clear();
        P.needs_mem.clear();
        P.ninstr = 0;
        const int rc = compile_trace(t, P, cap);
        if (rc) return rc;
    }
    return CHAI_OK;
}

// ------------------------------------------------ compiled-trace cache (JIT memo)
static uint64_t fnv_hash(const void* p, size_t n, uint64_t h)
{
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

struct CacheEntry {
    std::shared_ptr<TraceProgram> prog;
    std::string key;
};
static std::mutex g_cache_mu;
static std::map<uint64_t, std::vector<CacheEntry>> g_cache;
static uint64_t g_cache_hits = 0, g_cache_misses = 0, g_cache_entries = 0;

struct Handle {
    std::shared_ptr<TraceProgram> prog;
};

}  // namespace chai_hip

using namespace chai_hip;

extern "C" int chai_hip_trace_compile(const chai_trace* trace, void** handle)
{
    if (!trace || !handle) { set_error("trace_compile: null argument"); return CHAI_ERR_ARG; }
    *handle = nullptr;
    if (!trace->num_stmts || !trace->vars || !trace->stmts || !trace->nodes) {
        set_error("trace_compile: empty trace");
        return CHAI_ERR_ARG;
    }
    std::string key;
    key.append((const char*)trace->vars, sizeof(chai_trace_var) * trace->num_vars);
    key.append((const char*)trace->stmts, sizeof(chai_trace_stmt) * trace->num_stmts);
    key.append((const char*)trace->nodes, sizeof(chai_trace_node) * trace->num_nodes);
    key.push_back(jit_enabled() ? 'J' : 'V');
    const uint64_t h = fnv_hash(key.data(), key.size(), 1469598103934665603ull);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache.find(h);
        if (it != g_cache.end())
            for (auto& e : it->second)
                if (e.key == key) {
                    *handle = new Handle{e.prog};
                    g_cache_hits++;
                    return CHAI_OK;
                }
        g_cache_misses++;
    }
    auto prog = std::make_shared<TraceProgram>();
    prog->nuser = trace->num_vars;
    prog->own.adopt(*trace);
    const chai_trace *ts = nullptr, *tv = nullptr;
    int rc = split_trace(prog->own.t, prog->spl, ts);
    if (rc == CHAI_OK) rc = version_trace(*ts, prog->ver, tv);
    if (rc == CHAI_OK) rc = compile_trace(*tv, *prog, {});
    if (rc == CHAI_OK && jit_enabled()) rc = fit_jit_limit(*tv, *prog);
    if (rc == CHAI_OK) rc = materialize(*prog);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_cache[h].push_back(CacheEntry{prog, key});
        g_cache_entries++;
    }
    *handle = new Handle{prog};
    return CHAI_OK;
}

extern "C" int chai_hip_trace_cache_stats(uint64_t* hits, uint64_t* misses, uint64_t* entries)
{
    std::lock_guard<std::mutex> lk(g_cache_mu);
    if (hits) *hits = g_cache_hits;
    if (misses) *misses = g_cache_misses;
    if (entries) *entries = g_cache_entries;
    return CHAI_OK;
}

extern "C" int chai_hip_trace_launch(void* handle, void* const* buffers, void* stream)
{
    if (!handle) { set_error("trace_launch: null handle"); return CHAI_ERR_STATE; }
    return launch_program(*static_cast<Handle*>(handle)->prog, buffers, as_stream(stream));
}

extern "C" int chai_hip_trace_info(void* handle, int* num_kernels, int* num_instructions)
{
    if (!handle) { set_error("trace_info: null handle"); return CHAI_ERR_STATE; }
    const TraceProgram& P = *static_cast<Handle*>(handle)->prog;
    if (num_kernels) *num_kernels = (int)P.groups.size();
    if (num_instructions) *num_instructions = P.ninstr;
    return CHAI_OK;
}

extern "C" int chai_hip_trace_release(void* handle)
{
    delete static_cast<Handle*>(handle);
    return CHAI_OK;
}

extern "C" int chai_hip_trace_set_jit(int on)
{
    g_jit_mode = on ? 1 : 0;
    return CHAI_OK;
}

static int copy_out(const std::string& s, char* buf, size_t len)
{
    if (buf && len) {
        const size_t n = s.size() < len - 1 ? s.size() : len - 1;
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (int)s.size();
}

extern "C" int chai_hip_trace_kernel_source(void* handle, int kernel, char* buf, size_t len, int* jitted)
{
    if (!handle) { set_error("trace_kernel_source: null handle"); return CHAI_ERR_STATE; }
    const TraceProgram& P = *static_cast<Handle*>(handle)->prog;
    if (kernel < 0 || kernel >= (int)P.groups.size()) { set_error("trace_kernel_source: no kernel %d", kernel); return CHAI_ERR_ARG; }
    const Group& G = P.groups[kernel];
    if (jitted) *jitted = G.fn ? 1 : 0;
    return copy_out(G.fn ? G.jit_src : G.jit_log, buf, len) >= 0 ? CHAI_OK : CHAI_ERR_ARG;
}

extern "C" int chai_hip_trace_emit_source(const chai_trace* trace, int kernel, char* buf, size_t len)
{
    if (!trace) { set_error("trace_emit_source: null trace"); return CHAI_ERR_ARG; }
    TraceProgram P;
    Versioned spl, ver;
    const chai_trace *ts = nullptr, *tv = nullptr;
    int rc = split_trace(*trace, spl, ts);
    if (rc == CHAI_OK) rc = version_trace(*ts, ver, tv);
    if (rc == CHAI_OK) rc = compile_trace(*tv, P, {});
    if (rc == CHAI_OK) rc = fit_jit_limit(*tv, P);    // the kernels the JIT engine compiles
    if (rc) return rc;
    if (kernel < 0 || kernel >= (int)P.groups.size() || P.groups[kernel].kind != 0) {
        set_error("trace_emit_source: kernel %d is not a fused VM kernel", kernel);
        return CHAI_ERR_ARG;
    }
    copy_out(std::string(kJitPrelude) + jit_source(P, P.groups[kernel]), buf, len);
    return CHAI_OK;
}

extern "C" int chai_hip_trace_infer(chai_trace_var* vars, uint32_t num_vars, const chai_trace_stmt* stmts,
                                    uint32_t num_stmts, const chai_trace_node* nodes, uint32_t num_nodes)
{
    if (!vars || (num_stmts && (!stmts || !nodes))) { set_error("trace_infer: null argument"); return CHAI_ERR_ARG; }
    chai_trace t;
    t.vars = vars; t.num_vars = num_vars;
    t.stmts = stmts; t.num_stmts = num_stmts;
    t.nodes = nodes; t.num_nodes = num_nodes;
    for (uint32_t i = 0; i < num_stmts; i++) {
        const chai_trace_stmt& cs = stmts[i];
        if (cs.lhs >= num_vars) return fail("statement %u assigns variable %u of %u", i, cs.lhs, num_vars);
        if ((uint64_t)cs.first + cs.count > num_nodes || cs.count == 0) return fail("statement %u: bad node range", i);
        if (is_cl_stmt(t, cs) || vars[cs.lhs].W != 0) continue;
        for (uint32_t k = cs.first; k < cs.first + cs.count; k++)
            if (nodes[k].code == CHAI_TRACE_VAR && nodes[k].arg < num_vars && vars[nodes[k].arg].W == 0)
                return fail("statement %u reads variable %u before its shape is known", i, nodes[k].arg);
        Stmt s;
        const int rc = parse_nodes(t, nodes + cs.first, cs.count, s);
        if (rc) return rc;
        const ENode& r = s.nodes[s.root];
        chai_trace_var& v = vars[cs.lhs];
        v.prec = (uint32_t)r.prec;
        v.W = r.d.W; v.H = r.d.H; v.slots = r.d.S;
    }
    return CHAI_OK;
}

extern "C" int chai_hip_trace_check(const chai_trace* trace)
{
    if (!trace || !trace->num_stmts || !trace->vars || !trace->stmts || !trace->nodes) {
        set_error("trace_check: empty trace");
        return CHAI_ERR_ARG;
    }
    TraceProgram P;
    Versioned spl, ver;
    const chai_trace *ts = nullptr, *tv = nullptr;
    int rc = split_trace(*trace, spl, ts);
    if (rc == CHAI_OK) rc = version_trace(*ts, ver, tv);
    if (rc == CHAI_OK) rc = compile_trace(*tv, P, {});
    return rc;
}
