// DeviceMI355X.cpp -- lowering a VectorTrace to a chai_trace and running it
// on one MI355X through libchai_mi355x.so.
//
// The statements of a VectorTrace (common/VectorTrace.hpp:28-119) are
// "variable@version = RHS", each RHS a Stak<BC> (bytecode/BC.hpp) that the
// interpreter evaluates in order (runtime/Interpreter.cpp:344-372,
// interp/DispatchInterp.cpp): a variable BC names an array pushed before the
// next operation; scalar and pointer BCs are that operation's own arguments,
// popped last-pushed first (interp/BaseInterp.cpp:152-183); an operation pops
// its array operands from the value stack (the first operand on top,
// api/ArrayFun.cpp:45-52) and pushes its result.  The lowering replays that
// stack discipline and emits each operation as a node of the chai_trace
// (include/chai_mi355x.h), whose postfix order lists operands first-to-last.
//
// Host memory conventions of the reference, kept on both sides:
//   - arrays from make1 / make2 hold the caller's rows at the padded row
//     pitch PrecType::padWidth (memmgr/MemInterpAlloc.cpp:146-186);
//   - arrays computed by a device are dense W x H rows in a back memory of
//     the padded size, zero filled (memmgr/MemBase.cpp:13-30), and a 1 x 1
//     result keeps its vector-length tail zero, which read_scalar sums
//     (api/ArrayFP.cpp:86-104);
//   - readouts copy W x H dense elements (interp/InterpReadout.cpp:13-40).
#include "DeviceMI355X.hpp"

#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <stack>
#include <utility>
#include <vector>

#include "BackMem.hpp"
#include "FrontMem.hpp"
#include "MemInterp.hpp"
#include "PrecType.hpp"
#include "VectorNut.hpp"
#include "VectorTrace.hpp"
#include "chai/BC.hpp"

#include "../../include/chai_mi355x.h"

using namespace std;

namespace chai_internal {

////////////////////////////////////////
// the C ABI, resolved from the shared library at initialisation

struct MI355XBackend
{
    void* so;
    int (*init)(int);
    const char* (*last_error)(void);
    int (*alloc)(size_t, void**);
    int (*free_)(void*);
    int (*h2d)(void*, const void*, size_t, void*);
    int (*d2h)(void*, const void*, size_t, void*);
    int (*sync)(void*);
    int (*infer)(chai_trace_var*, uint32_t, const chai_trace_stmt*, uint32_t, const chai_trace_node*, uint32_t);
    int (*compile)(const chai_trace*, void**);
    int (*launch)(void*, void* const*, void*);
    int (*release)(void*);
};

namespace {

template <typename F>
bool resolve(void* so, const char* name, F& fn)
{
    fn = reinterpret_cast<F>(dlsym(so, name));
    return fn != NULL;
}

MI355XBackend* load_backend(string& why)
{
    const char* path = getenv("CHAI_MI355X_LIB");
    if (!path || !*path) path = "libchai_mi355x.so";
    void* so = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!so) {
        why = string("cannot load ") + path + ": " + dlerror();
        return NULL;
    }
    MI355XBackend* be = new MI355XBackend;
    be->so = so;
    const bool ok = resolve(so, "chai_hip_init", be->init) &&
                    resolve(so, "chai_hip_last_error", be->last_error) &&
                    resolve(so, "chai_hip_alloc", be->alloc) &&
                    resolve(so, "chai_hip_free", be->free_) &&
                    resolve(so, "chai_hip_h2d", be->h2d) &&
                    resolve(so, "chai_hip_d2h", be->d2h) &&
                    resolve(so, "chai_hip_sync", be->sync) &&
                    resolve(so, "chai_hip_trace_infer", be->infer) &&
                    resolve(so, "chai_hip_trace_compile", be->compile) &&
                    resolve(so, "chai_hip_trace_launch", be->launch) &&
                    resolve(so, "chai_hip_trace_release", be->release);
    if (!ok) {
        why = string(path) + " lacks a chai_hip_* entry point";
        dlclose(so);
        delete be;
        return NULL;
    }
    return be;
}

size_t elem_size(const size_t prec) { return prec == PrecType::Double ? 8 : 4; }

// an array read by the trace that lives in host memory already
struct HostArray {
    vector< FrontMem* > fronts;    // one per slot (gathered trace)
    bool padded;                   // make1 / make2 data: rows at the padded pitch
};

struct Readout {
    uint32_t var;                  // trace variable read out
    vector< FrontMem* > dst;       // one per slot
};

struct Assigned {
    uint32_t var;                  // trace variable holding the result
    uint32_t variable, version;    // the VectorTrace's name for it
};

struct Sub {                       // an expression: postfix nodes, operands first
    vector< chai_trace_node > n;
    bool transposed;
    Sub() : transposed(false) {}
};

struct Item {                      // a scalar or pointer argument BC
    bool ptr;
    double value;
    size_t stmt;
};

// per-opcode operand counts (runtime/Interpreter.cpp:44-286 registrations):
// arguments taken from the bytecode stack and arrays taken from the values
bool op_counts(const uint32_t c, size_t& items, size_t& arrays)
{
    items = 0;
    arrays = 0;
    if (c >= 1 && c <= 4) { arrays = 1; return true; }          // convert_*
    if (c >= 5 && c <= 8) { items = 1; return true; }           // scalar_*
    if (c == 10 || c == 15 || c == 16) { arrays = 1; return true; }
    if ((c >= 11 && c <= 14) || (c >= 17 && c <= 22) || (c >= 30 && c <= 37)) { arrays = 2; return true; }
    if (c == 23) { arrays = 3; return true; }                   // operatorCONDEXPR
    if (c == 40) { arrays = 2; return true; }                   // gather1_floor
    if (c == 41) { arrays = 3; return true; }                   // gather2_floor
    if (c >= 42 && c <= 45) { items = 2; return true; }         // index1_*: dim, W
    if (c >= 46 && c <= 49) { items = 3; return true; }         // index2_*: dim, W, H
    if (c >= 60 && c <= 67) { items = 4; return true; }         // make*: W, H, slots, data
    if (c >= 68 && c <= 71) { items = 1; return true; }         // ones1_*: W
    if (c >= 72 && c <= 75) { items = 2; return true; }         // ones2_* (73: zeros2_f64): W, H
    if (c >= 76 && c <= 79) { items = 1; return true; }         // zeros1_*
    if (c >= 80 && c <= 82) { items = 2; return true; }         // zeros2_*
    if (c >= 90 && c <= 101) { items = 1; arrays = 1; return true; }    // read_scalar / read1 / read2
    if (c == 110 || c == 111 || c == 112 || c == 113) { items = 5; return true; }  // rng: variant, seed, W, H, slots
    if (c == 114 || c == 115) { items = 7; return true; }       // + min, max
    if (c == 120) { arrays = 2; return true; }                  // matmul
    if (c == 121) { items = 2; arrays = 3; return true; }       // matmulG: alpha, beta
    if (c == 122) { arrays = 1; return true; }                  // transpose
    if (c == 130) { arrays = 2; return true; }                  // dot_product
    if (c == 131 || c == 132) { arrays = 1; return true; }      // mean, sum
    // math, common, integer and relational functions (InterpFun1/2/3)
    static const uint32_t unary[] = {
        200, 201, 202, 203, 204, 205, 206, 208, 209, 211, 212, 214, 215, 216, 217, 218, 219, 220, 221,
        222, 223, 225, 233, 235, 237, 238, 239, 240, 241, 246, 253, 255, 256, 257, 259, 260, 261, 262,
        263, 264, 265, 266, 301, 305, 308, 400, 404, 507, 508, 509, 510, 513};
    static const uint32_t binary[] = {
        207, 210, 213, 224, 227, 228, 229, 232, 234, 243, 244, 247, 248, 249, 250, 251, 254, 302, 303,
        306, 401, 402, 405, 411, 412, 413, 414, 415, 500, 501, 502, 503, 504, 505, 506, 511, 512};
    static const uint32_t ternary[] = {226, 242, 300, 304, 307, 406, 407, 408};
    for (size_t i = 0; i < sizeof unary / sizeof unary[0]; i++)
        if (c == unary[i]) { arrays = 1; return true; }
    for (size_t i = 0; i < sizeof binary / sizeof binary[0]; i++)
        if (c == binary[i]) { arrays = 2; return true; }
    for (size_t i = 0; i < sizeof ternary / sizeof ternary[0]; i++)
        if (c == ternary[i]) { arrays = 3; return true; }
    return false;
}

class Lowering : public Visit<BC>
{
    VectorTrace& _vt;
    stack< Sub > _vals;
    vector< Item > _items;
    size_t _ops;                   // operations of the current statement
    bool _ok;
    string _why;

public:
    vector< chai_trace_var > vars;
    vector< chai_trace_stmt > stmts;
    vector< chai_trace_node > nodes;
    map< pair< uint32_t, uint32_t >, uint32_t > varOf;   // (variable, version) -> trace variable
    map< uint32_t, HostArray > host;                      // trace variable -> host data
    vector< Readout > readouts;
    vector< Assigned > assigned;                          // every assignment, in order
    vector< FrontMem* > nutReads;

    typedef int (*InferFn)(chai_trace_var*, uint32_t, const chai_trace_stmt*, uint32_t, const chai_trace_node*,
                           uint32_t);
    InferFn infer;
    const char* (*inferError)(void);

    Lowering(VectorTrace& vt) : _vt(vt), _ops(0), _ok(true), infer(NULL), inferError(NULL) {}

    bool ok(void) const { return _ok; }
    const string& why(void) const { return _why; }

    void refuse(const string& s)
    {
        if (_ok) _why = s;
        _ok = false;
    }

    uint32_t addVar(const uint32_t prec, const uint32_t role, const uint64_t W, const uint64_t H, const uint64_t S)
    {
        chai_trace_var v;
        memset(&v, 0, sizeof v);
        v.prec = prec;
        v.role = role;
        v.W = W;
        v.H = H;
        v.slots = S;
        vars.push_back(v);
        return (uint32_t)vars.size() - 1;
    }

    // a host array read by the trace becomes an input variable
    bool hostVar(const vector< FrontMem* >& fronts, const bool padded, uint32_t& var)
    {
        if (fronts.empty()) { refuse("array without memory"); return false; }
        const FrontMem* f0 = fronts.front();
        for (size_t i = 0; i < fronts.size(); i++) {
            const FrontMem* f = fronts[i];
            if (f->slots() != 1 || f->W() != f0->W() || f->H() != f0->H() || f->prec() != f0->prec()) {
                refuse("vector array data (slots inside one trace) is not lowered");
                return false;
            }
            if (!f->ptrMem()) { refuse("array memory not allocated"); return false; }
        }
        var = addVar((uint32_t)f0->prec(), CHAI_VAR_INPUT, f0->W(), f0->H(), fronts.size());
        HostArray h;
        h.fronts = fronts;
        h.padded = padded;
        host[var] = h;
        return true;
    }

    static void materialize(Sub& s)
    {
        s.n.push_back(node(122));
        s.transposed = false;
    }

    static chai_trace_node node(const uint32_t code)
    {
        chai_trace_node n;
        memset(&n, 0, sizeof n);
        n.code = code;
        return n;
    }

    void visit(const uint32_t variable, const uint32_t version)
    {
        if (!_ok) return;
        const pair< uint32_t, uint32_t > key(variable, version);
        uint32_t var;
        if (varOf.count(key)) {
            var = varOf[key];
        } else {
            // assigned by an earlier trace: its memory is in the variable's nut
            if (!_vt.vectorNuts().count(variable)) { refuse("variable without a nut"); return; }
            const vector< FrontMem* > fronts = _vt.vectorNuts()[variable]->getNutMem(version);
            if (!hostVar(fronts, fronts.empty() ? false : fronts.front()->dataPtr() != NULL, var)) return;
            varOf[key] = var;
            nutReads.insert(nutReads.end(), fronts.begin(), fronts.end());
        }
        Sub s;
        chai_trace_node n = node(CHAI_TRACE_VAR);
        n.arg = var;
        s.n.push_back(n);
        _vals.push(s);
    }

    void visit(const double scalar)
    {
        Item it;
        it.ptr = false;
        it.value = scalar;
        it.stmt = 0;
        _items.push_back(it);
    }

    void visit(void* ptr)
    {
        Item it;
        it.ptr = true;
        it.value = 0;
        it.stmt = reinterpret_cast< size_t >(ptr);
        _items.push_back(it);
    }

    void visit(const uint32_t c)
    {
        if (!_ok) return;
        _ops++;
        size_t nitems, narrays;
        if (!op_counts(c, nitems, narrays)) {
            char buf[64];
            snprintf(buf, sizeof buf, "bytecode %u is not lowered", c);
            refuse(buf);
            return;
        }
        if (_items.size() < nitems || _vals.size() < narrays) { refuse("bytecode stack underflow"); return; }
        vector< Item > a;                                   // a[0] = last pushed
        for (size_t i = 0; i < nitems; i++) { a.push_back(_items.back()); _items.pop_back(); }
        vector< Sub > k;                                    // k[0] = first operand
        for (size_t i = 0; i < narrays; i++) { k.push_back(_vals.top()); _vals.pop(); }
        for (size_t i = 0; i < nitems; i++)
            if (a[i].ptr != ((c >= 60 && c <= 67 && i == 3) || (c >= 90 && c <= 101))) {
                refuse("unexpected argument kind");
                return;
            }

        Sub out;
        if (c >= 60 && c <= 67) {                           // make1 / make2: host data
            if (!_vt.frontMem().count(a[3].stmt)) { refuse("make data missing"); return; }
            uint32_t var;
            if (!hostVar(_vt.frontMem()[a[3].stmt], true, var)) return;
            chai_trace_node n = node(CHAI_TRACE_VAR);
            n.arg = var;
            out.n.push_back(n);
            _vals.push(out);
            return;
        }
        if (c >= 90 && c <= 101) {                          // readout into host memory
            if (k[0].n.size() != 1 || k[0].n[0].code != CHAI_TRACE_VAR || k[0].transposed) {
                refuse("readout of an expression");
                return;
            }
            if (!_vt.frontMem().count(a[0].stmt)) { refuse("readout memory missing"); return; }
            Readout r;
            r.var = k[0].n[0].arg;
            r.dst = _vt.frontMem()[a[0].stmt];
            readouts.push_back(r);
            return;                                         // pushes no value
        }
        if (c == 122) {                                     // transpose: a flag on a matmul operand
            out = k[0];
            out.transposed = !out.transposed;
            _vals.push(out);
            return;
        }
        // a transpose stays a flag on a matmul operand (the product reads the
        // operand transposed, jitstmt/MakeStmt.cpp:704-720); anywhere else it
        // becomes a transpose node of its own
        for (size_t i = 0; i < narrays; i++)
            if (k[i].transposed && c != 120 && c != 121) materialize(k[i]);
        if (c >= 1 && c <= 4 && k[0].n.size() == 1 && k[0].n[0].code == CHAI_TRACE_VAR && !k[0].transposed) {
            // converting an array to its own precision returns the same array
            // (interp/InterpConvert.cpp:21-27): assignments "A = B" alias
            static const uint32_t cp[] = {CHAI_U32, CHAI_I32, CHAI_F32, CHAI_F64};
            if (vars[k[0].n[0].arg].prec == cp[c - 1]) {
                _vals.push(k[0]);
                return;
            }
        }
        chai_trace_node n = node(c);
        if (c >= 5 && c <= 8) {
            n.s0 = a[0].value;
        } else if (c >= 42 && c <= 49) {
            n.arg = (uint32_t)a[0].value;
            n.W = (uint64_t)a[1].value;
            n.H = nitems == 3 ? (uint64_t)a[2].value : 1;
            n.slots = 1;
        } else if (c >= 68 && c <= 82) {
            n.W = (uint64_t)a[0].value;
            n.H = nitems == 2 ? (uint64_t)a[1].value : 1;
            n.slots = 1;
        } else if (c >= 110 && c <= 115) {
            n.arg = (uint32_t)a[0].value;
            n.seed = (uint64_t)a[1].value;
            n.W = (uint64_t)a[2].value;
            n.H = (uint64_t)a[3].value;
            n.slots = (uint64_t)a[4].value;
            if (nitems == 7) { n.s0 = a[5].value; n.s1 = a[6].value; }
        } else if (c == 121) {
            n.s0 = a[0].value;
            n.s1 = a[1].value;
        }
        for (size_t i = 0; i < narrays; i++) {
            if (k[i].transposed) k[i].n.back().flags ^= 1u;
            out.n.insert(out.n.end(), k[i].n.begin(), k[i].n.end());
        }
        out.n.push_back(n);
        _vals.push(out);
    }

    // one statement "variable@version = RHS"
    void statement(const size_t i)
    {
        if (!_ok) return;
        while (!_vals.empty()) _vals.pop();
        _items.clear();
        _ops = 0;
        _vt.getRHS(i).ascend(*this);
        if (!_ok) return;
        const uint32_t variable = _vt.getVariable(i), version = _vt.getVersion(i);
        // a readout pushes nothing; a statement without an operation leaves
        // its variables unpushed (interp/DispatchInterp.cpp:68-88): neither
        // assigns the statement's array
        if (_vals.empty() || _ops == 0) return;
        if (_vals.size() != 1 || !_items.empty()) { refuse("statement leaves extra values"); return; }
        Sub s = _vals.top();
        _vals.pop();
        if (s.transposed) materialize(s);
        const pair< uint32_t, uint32_t > key(variable, version);
        Assigned asg;
        asg.variable = variable;
        asg.version = version;
        if (s.n.size() == 1 && s.n[0].code == CHAI_TRACE_VAR) {
            // make1 / make2, or an alias: the statement's array is that array
            asg.var = s.n[0].arg;
        } else {
            // computed: a temporary until liveness says it must be stored
            asg.var = addVar(0, CHAI_VAR_TEMP, 0, 0, 0);
            // forceVecLength (common/ClientTrace.cpp:38-89, enqueue/JustInTimeMemo.cpp:
            // 59-88): the vector length the caller fixed for the variable selects
            // the RNG block of a random array assigned to it
            map< uint32_t, int >::const_iterator fv = _vt.forceVecLength().find(variable);
            if (fv != _vt.forceVecLength().end() && fv->second > 0 && s.n.back().code >= 110 &&
                s.n.back().code <= 115)
                s.n.back().vec_len = (uint32_t)fv->second;
            chai_trace_stmt st;
            memset(&st, 0, sizeof st);
            st.lhs = asg.var;
            st.first = (uint32_t)nodes.size();
            st.count = (uint32_t)s.n.size();
            nodes.insert(nodes.end(), s.n.begin(), s.n.end());
            stmts.push_back(st);
            // its shape and type now, so later statements can see them
            if (infer(&vars[0], (uint32_t)vars.size(), &stmts[0], (uint32_t)stmts.size(), &nodes[0],
                      (uint32_t)nodes.size()) != CHAI_OK) {
                refuse(string("shape inference: ") + inferError());
                return;
            }
        }
        varOf[key] = asg.var;
        assigned.push_back(asg);
    }

    // Which assigned arrays outlive the trace (common/ClientTrace.cpp:114-139,
    // the variables constructed and not destroyed): the last assignment of
    // each live variable is stored and becomes the variable's nut memory;
    // every other statement result is a temporary of the fused kernels.
    // Arrays read out are stored too.
    void liveness(void)
    {
        map< uint32_t, size_t > last;                // variable -> its last assignment
        for (size_t j = 0; j < assigned.size(); j++) last[assigned[j].variable] = j;
        vector< Assigned > keep;
        for (size_t j = 0; j < assigned.size(); j++) {
            const Assigned& a = assigned[j];
            if (last[a.variable] != j || !_vt.liveVariables().count(a.variable)) continue;
            if (!host.count(a.var)) vars[a.var].role = CHAI_VAR_OUTPUT;
            keep.push_back(a);
        }
        assigned.swap(keep);
        for (size_t r = 0; r < readouts.size(); r++)
            if (!host.count(readouts[r].var)) vars[readouts[r].var].role = CHAI_VAR_OUTPUT;
    }

};

// dense [slots][H][W] staging of a host array
void gather_host(const HostArray& h, const chai_trace_var& v, vector< char >& out)
{
    const size_t es = elem_size(v.prec), row = v.W * es;
    const size_t pitch = (h.padded ? PrecType::padWidth(v.prec, v.W) : v.W) * es;
    out.resize(v.slots * v.H * row);
    for (size_t s = 0; s < v.slots; s++) {
        const char* src = static_cast< const char* >(h.fronts[s]->ptrMem());
        for (size_t y = 0; y < v.H; y++) memcpy(&out[(s * v.H + y) * row], src + y * pitch, row);
    }
}

}  // namespace

////////////////////////////////////////
// device

DeviceMI355X::DeviceMI355X(const int deviceNum, const int ordinal)
    : DeviceBase(deviceNum),
      _ordinal(ordinal),
      _be(NULL),
      _traces(0),
      _fallbacks(0) { }

DeviceMI355X::~DeviceMI355X(void)
{
    if (_be) {
        dlclose(_be->so);
        delete _be;
    }
}

void DeviceMI355X::sub_initDevice(MemInterp&)
{
    string why;
    MI355XBackend* be = load_backend(why);
    if (be && be->init(_ordinal) != CHAI_OK) {
        why = string("chai_hip_init: ") + be->last_error();
        dlclose(be->so);
        delete be;
        be = NULL;
    }
    _be = be;
    _why = why;
    if (!_be && getenv("CHAI_MI355X_VERBOSE"))
        fprintf(stderr, "DeviceMI355X %d: not available (%s)\n", getDeviceNum(), _why.c_str());
}

bool DeviceMI355X::sub_evaluate(VectorTrace& vt)
{
    const bool verbose = getenv("CHAI_MI355X_VERBOSE") != NULL;
    if (!_be) {
        _fallbacks++;
        return false;
    }
    Lowering L(vt);
    L.infer = _be->infer;
    L.inferError = _be->last_error;
    for (size_t i = 0; i < vt.traceLength() && L.ok(); i++) L.statement(i);
    if (L.ok()) L.liveness();
    // CHAI_MI355X_DUMP=<dir>: write each lowered trace (three counts, then
    // the variable, statement and node arrays) for offline inspection
    const char* dump = getenv("CHAI_MI355X_DUMP");
    if (dump && L.ok() && !L.stmts.empty()) {
        char path[512];
        snprintf(path, sizeof path, "%s/trace_%016llx_%zu.bin", dump, (unsigned long long)vt.hashCode(),
                 _traces + _fallbacks);
        FILE* f = fopen(path, "wb");
        if (f) {
            const uint32_t n[3] = {(uint32_t)L.vars.size(), (uint32_t)L.stmts.size(), (uint32_t)L.nodes.size()};
            fwrite(n, sizeof n, 1, f);
            fwrite(&L.vars[0], sizeof(chai_trace_var), L.vars.size(), f);
            fwrite(&L.stmts[0], sizeof(chai_trace_stmt), L.stmts.size(), f);
            fwrite(&L.nodes[0], sizeof(chai_trace_node), L.nodes.size(), f);
            fclose(f);
        }
    }
    void* prog = NULL;
    if (L.ok() && !L.stmts.empty()) {
        chai_trace t;
        t.vars = &L.vars[0];
        t.num_vars = (uint32_t)L.vars.size();
        t.stmts = &L.stmts[0];
        t.num_stmts = (uint32_t)L.stmts.size();
        t.nodes = &L.nodes[0];
        t.num_nodes = (uint32_t)L.nodes.size();
        if (_be->compile(&t, &prog) != CHAI_OK) L.refuse(string("compile: ") + _be->last_error());
    }
    if (!L.ok()) {
        if (verbose) fprintf(stderr, "DeviceMI355X %d: trace %llx refused (%s)\n", getDeviceNum(),
                             (unsigned long long)vt.hashCode(), L.why().c_str());
        _fallbacks++;
        return false;
    }

    // device buffers: inputs uploaded, every other variable written by the trace
    bool good = true;
    vector< void* > bufs(L.vars.size(), (void*)NULL);
    vector< char > stage;
    for (size_t v = 0; v < L.vars.size() && good; v++) {
        const chai_trace_var& tv = L.vars[v];
        const size_t bytes = tv.W * tv.H * tv.slots * elem_size(tv.prec);
        good = _be->alloc(bytes ? bytes : 8, &bufs[v]) == CHAI_OK;
        if (good && L.host.count((uint32_t)v)) {
            gather_host(L.host[(uint32_t)v], tv, stage);
            good = _be->h2d(bufs[v], &stage[0], stage.size(), NULL) == CHAI_OK;
        }
    }
    if (good && prog) good = _be->launch(prog, &bufs[0], NULL) == CHAI_OK;
    if (good) good = _be->sync(NULL) == CHAI_OK;

    // read back every stored array into memory the interpreter would have made
    // (memmgr/MemBase.cpp:13-30), once per array however many variables alias it
    map< uint32_t, vector< FrontMem* > > results;
    for (size_t j = 0; j < L.assigned.size() && good; j++) {
        const uint32_t var = L.assigned[j].var;
        if (L.host.count(var) || results.count(var)) continue;
        const chai_trace_var& tv = L.vars[var];
        const size_t es = elem_size(tv.prec), frontBytes = tv.W * tv.H * es;
        stage.resize(frontBytes * tv.slots);
        good = _be->d2h(&stage[0], bufs[var], stage.size(), NULL) == CHAI_OK;
        if (!good) break;
        BackMem* back = getMemInterp().allocBackMem(tv.prec, tv.W, tv.H, tv.slots);
        if (!back) { good = false; break; }
        vector< FrontMem* > fronts;
        for (size_t s = 0; s < tv.slots; s++) {
            FrontMem* m = getMemInterp().allocFrontMem(tv.prec, tv.W, tv.H, back, s);
            memcpy(m->ptrMem(), &stage[s * frontBytes], frontBytes);
            fronts.push_back(m);
        }
        results[var] = fronts;
    }
    for (size_t r = 0; r < L.readouts.size() && good; r++) {
        const Readout& ro = L.readouts[r];
        const chai_trace_var& tv = L.vars[ro.var];
        const size_t es = elem_size(tv.prec), frontBytes = tv.W * tv.H * es;
        if (ro.dst.size() != tv.slots) { good = false; break; }
        stage.resize(frontBytes * tv.slots);
        good = _be->d2h(&stage[0], bufs[ro.var], stage.size(), NULL) == CHAI_OK;
        for (size_t s = 0; s < tv.slots && good; s++) memcpy(ro.dst[s]->ptrMem(), &stage[s * frontBytes], frontBytes);
    }
    for (size_t v = 0; v < bufs.size(); v++)
        if (bufs[v]) _be->free_(bufs[v]);
    if (prog) _be->release(prog);

    if (!good) {
        if (verbose) fprintf(stderr, "DeviceMI355X %d: trace %llx failed on the device (%s)\n", getDeviceNum(),
                             (unsigned long long)vt.hashCode(), _be->last_error());
        for (map< uint32_t, vector< FrontMem* > >::iterator it = results.begin(); it != results.end(); it++)
            for (size_t s = 0; s < it->second.size(); s++) delete it->second[s]->release();
        _fallbacks++;
        return false;
    }

    // publish: the live variables' arrays become their nut memory (the
    // interpreter's same-precision copies share one array the same way)
    for (size_t j = 0; j < L.assigned.size(); j++) {
        const Assigned& a = L.assigned[j];
        const vector< FrontMem* >& fronts = L.host.count(a.var) ? L.host[a.var].fronts : results[a.var];
        vt.vectorNuts()[a.variable]->setNutMem(a.version, fronts);
        if (!L.host.count(a.var)) fronts.front()->getBackMem()->needMTSafe();
    }
    // the interpreter releases every array it read (no-op for nut-held memory)
    for (size_t i = 0; i < L.nutReads.size(); i++) delete L.nutReads[i]->release();
    _traces++;
    return true;
}

}; // namespace chai_internal
