// program_cl.hip -- the ProgramCL inline-kernel hook on ROCm.
//
// A Chai program may carry its own OpenCL C kernels (api/ProgramCL.cpp:133-376)
// and launch them on its arrays between traced statements
// (enqueue/EnqueueAST.cpp:432-489).  The reference hands the text to the
// OpenCL driver (vendor/OCLdevice.cpp:940-968) and binds the arrays' buffers
// with clSetKernelArg (memmgr/MemTransJIT.cpp:161-260).  Here the same text is
// compiled by ROCm's code-object manager (comgr) for gfx950: OpenCL C 1.2 with
// the device libraries to bitcode, code generation, link to an executable
// code object.  The code object's own metadata (amdhsa.kernels) gives each
// kernel's explicit arguments (offset, size, value kind, address space, type
// name), which is how arguments are checked and packed; nothing is parsed
// from the source here.  The code object is loaded with hipModuleLoadData and
// launched with hipModuleLaunchKernel on the caller's stream, on the back
// end's own device buffers, so an inline kernel and the fused trace kernels
// share memory without copies.
//
// Programs are cached by a 64-bit hash of their text and options, as the
// reference keys them by TEA::hash of the text (api/ProgramCL.cpp:362-376),
// and live until chai_hip_shutdown / process exit.
//
// __local pointer arguments are given their byte count at launch time, as
// clSetKernelArg(k, i, bytes, NULL) does.  Their LDS addresses are laid out
// after the kernel's static LDS (each aligned to 16 bytes) by a generated
// wrapper: the program is recompiled once per distinct set of local sizes
// with a kernel that declares the __local arrays and calls the user kernel.
#include <hip/hip_runtime.h>
#include <amd_comgr/amd_comgr.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chai_mi355x.h"
#include "common.hpp"

namespace chai_hip {
namespace {

struct ClArg {
    uint32_t offset = 0, size = 0;
    int kind = CHAI_CL_ARG_OTHER;      // CHAI_CL_ARG_*
    std::string type_name, name, value_kind, address_space;
};

struct ClProgram;

struct ClKernel {
    std::string name;
    ClProgram* prog = nullptr;
    std::vector<ClArg> args;           // explicit arguments only
    uint32_t kernarg_size = 0, group_static = 0, max_wg = 0;
    hipFunction_t fn = nullptr;
};

struct ClProgram {
    std::string text, options;
    std::vector<char> code;            // executable code object
    std::string log;
    std::map<std::string, ClKernel> kernels;
    hipModule_t mod = nullptr;
    int device = -1;
    // wrappers for kernels with __local pointer arguments, per local-size set
    std::map<std::string, std::unique_ptr<ClProgram>> wrapped;
};

// comgr is opened at run time from the ROCm installation (ROCM_PATH, else
// /opt/rocm), not linked: a process may already hold another copy of the
// library under the same soname (PyTorch bundles one) that lacks clang's
// OpenCL headers (opencl-c.h).  dlopen by full path loads the ROCm copy beside
// it, with its own symbols.
#define CG_FUNCS(X) X(action_data_count) X(action_data_get_data) X(action_info_set_isa_name) X(action_info_set_language) X(action_info_set_logging) X(action_info_set_option_list) X(create_action_info) X(create_data) X(create_data_set) X(data_set_add) X(destroy_action_info) X(destroy_data_set) X(destroy_metadata) X(do_action) X(get_data) X(get_data_metadata) X(get_metadata_list_size) X(get_metadata_string) X(index_list_metadata) X(metadata_lookup) X(release_data) X(set_data) X(set_data_name)

struct ComgrApi {
#define CG_FN(n) decltype(&::amd_comgr_##n) n = nullptr;
    CG_FUNCS(CG_FN)
#undef CG_FN
    bool ok = false;
    std::string err;
};

const ComgrApi& cg()
{
    static const ComgrApi api = [] {
        ComgrApi a;
        const char* root = getenv("ROCM_PATH");
        const std::string path = std::string(root && *root ? root : "/opt/rocm") + "/lib/libamd_comgr.so.3";
        void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) h = dlopen("libamd_comgr.so.3", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            const char* e = dlerror();
            a.err = e ? e : "dlopen failed";
            return a;
        }
#define CG_SYM(n)                                                    \
    a.n = reinterpret_cast<decltype(a.n)>(dlsym(h, "amd_comgr_" #n)); \
    if (!a.n) {                                                      \
        a.err = "missing amd_comgr_" #n;                             \
        return a;                                                    \
    }
        CG_FUNCS(CG_SYM)
#undef CG_SYM
        a.ok = true;
        return a;
    }();
    return api;
}

std::mutex g_mu;
std::map<uint64_t, std::unique_ptr<ClProgram>> g_programs;

uint64_t hash_text(const std::string& a, const std::string& b)
{
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : a) h = (h ^ c) * 1099511628211ull;
    h = (h ^ 0xffu) * 1099511628211ull;
    for (unsigned char c : b) h = (h ^ c) * 1099511628211ull;
    return h;
}

int cfail(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_error("%s", buf);
    return CHAI_ERR_ARG;
}

// ---- comgr metadata helpers
std::string md_string(amd_comgr_metadata_node_t n)
{
    size_t len = 0;
    if (cg().get_metadata_string(n, &len, nullptr) != AMD_COMGR_STATUS_SUCCESS) return "";
    std::string s(len, '\0');
    if (cg().get_metadata_string(n, &len, &s[0]) != AMD_COMGR_STATUS_SUCCESS) return "";
    while (!s.empty() && s.back() == '\0') s.pop_back();
    return s;
}

bool md_lookup_string(amd_comgr_metadata_node_t map, const char* key, std::string& out)
{
    amd_comgr_metadata_node_t v;
    if (cg().metadata_lookup(map, key, &v) != AMD_COMGR_STATUS_SUCCESS) return false;
    out = md_string(v);
    cg().destroy_metadata(v);
    return true;
}

uint32_t md_lookup_u32(amd_comgr_metadata_node_t map, const char* key)
{
    std::string s;
    return md_lookup_string(map, key, s) ? (uint32_t)strtoul(s.c_str(), nullptr, 10) : 0u;
}

int parse_metadata(amd_comgr_data_t exe, ClProgram& P)
{
    amd_comgr_metadata_node_t root, kernels;
    if (cg().get_data_metadata(exe, &root) != AMD_COMGR_STATUS_SUCCESS) return cfail("no code-object metadata");
    if (cg().metadata_lookup(root, "amdhsa.kernels", &kernels) != AMD_COMGR_STATUS_SUCCESS) {
        cg().destroy_metadata(root);
        return cfail("code object lists no kernels");
    }
    size_t nk = 0;
    cg().get_metadata_list_size(kernels, &nk);
    for (size_t i = 0; i < nk; i++) {
        amd_comgr_metadata_node_t k, args;
        if (cg().index_list_metadata(kernels, i, &k) != AMD_COMGR_STATUS_SUCCESS) continue;
        ClKernel K;
        md_lookup_string(k, ".name", K.name);
        K.kernarg_size = md_lookup_u32(k, ".kernarg_segment_size");
        K.group_static = md_lookup_u32(k, ".group_segment_fixed_size");
        K.max_wg = md_lookup_u32(k, ".max_flat_workgroup_size");
        if (cg().metadata_lookup(k, ".args", &args) == AMD_COMGR_STATUS_SUCCESS) {
            size_t na = 0;
            cg().get_metadata_list_size(args, &na);
            for (size_t j = 0; j < na; j++) {
                amd_comgr_metadata_node_t a;
                if (cg().index_list_metadata(args, j, &a) != AMD_COMGR_STATUS_SUCCESS) continue;
                ClArg A;
                md_lookup_string(a, ".value_kind", A.value_kind);
                A.offset = md_lookup_u32(a, ".offset");
                A.size = md_lookup_u32(a, ".size");
                md_lookup_string(a, ".type_name", A.type_name);
                md_lookup_string(a, ".name", A.name);
                md_lookup_string(a, ".address_space", A.address_space);
                cg().destroy_metadata(a);
                if (A.value_kind.compare(0, 7, "hidden_") == 0) continue;   // filled by the runtime
                if (A.value_kind == "global_buffer") A.kind = CHAI_CL_ARG_ARRAY;
                else if (A.value_kind == "dynamic_shared_pointer") A.kind = CHAI_CL_ARG_LOCAL;
                else if (A.value_kind == "by_value") A.kind = CHAI_CL_ARG_SCALAR;
                else A.kind = CHAI_CL_ARG_OTHER;                     // images, samplers, queues
                K.args.push_back(A);
            }
            cg().destroy_metadata(args);
        }
        cg().destroy_metadata(k);
        if (!K.name.empty()) {
            K.prog = &P;
            P.kernels[K.name] = K;
        }
    }
    cg().destroy_metadata(kernels);
    cg().destroy_metadata(root);
    return CHAI_OK;
}

// ---- comgr compile: OpenCL C 1.2 + device libs -> bitcode -> relocatable -> executable
#define COMGR_TRY(call, what)                                                             \
    do {                                                                                  \
        if ((call) != AMD_COMGR_STATUS_SUCCESS) { rc = cfail("comgr: %s failed", what); goto done; } \
    } while (0)

void append_logs(amd_comgr_data_set_t set, std::string& log)
{
    size_t n = 0;
    if (cg().action_data_count(set, AMD_COMGR_DATA_KIND_LOG, &n) != AMD_COMGR_STATUS_SUCCESS) return;
    for (size_t i = 0; i < n; i++) {
        amd_comgr_data_t d;
        if (cg().action_data_get_data(set, AMD_COMGR_DATA_KIND_LOG, i, &d) != AMD_COMGR_STATUS_SUCCESS) continue;
        size_t len = 0;
        cg().get_data(d, &len, nullptr);
        std::string s(len, '\0');
        cg().get_data(d, &len, &s[0]);
        log += s;
        cg().release_data(d);
    }
}

int compile_cl(ClProgram& P)
{
    if (!cg().ok) return cfail("cannot load ROCm's code-object manager (libamd_comgr.so.3): %s", cg().err.c_str());
    int rc = CHAI_OK;
    amd_comgr_data_t src = {0}, exe = {0};
    amd_comgr_data_set_t in = {0}, bc = {0}, reloc = {0}, out = {0};
    amd_comgr_action_info_t ai = {0};
    bool have_src = false, have_in = false, have_bc = false, have_reloc = false, have_out = false, have_ai = false,
         have_exe = false;
    // options: kernel-argument metadata (type and parameter names), no packed
    // FP32 instructions (the target feature is recorded in every function's
    // attributes, so code generation honours it too), plus the caller's
    std::vector<std::string> opts = {"-cl-kernel-arg-info", "-O3", "-Xclang", "-target-feature", "-Xclang",
                                     "-packed-fp32-ops"};
    {
        size_t p = 0;
        const std::string& o = P.options;
        while (p < o.size()) {
            while (p < o.size() && o[p] == ' ') p++;
            size_t e = p;
            while (e < o.size() && o[e] != ' ') e++;
            if (e > p) opts.push_back(o.substr(p, e - p));
            p = e;
        }
    }
    std::vector<const char*> optp;
    for (auto& s : opts) optp.push_back(s.c_str());

    COMGR_TRY(cg().create_data(AMD_COMGR_DATA_KIND_SOURCE, &src), "create_data");
    have_src = true;
    COMGR_TRY(cg().set_data(src, P.text.size(), P.text.data()), "set_data");
    COMGR_TRY(cg().set_data_name(src, "chai_program.cl"), "set_data_name");
    COMGR_TRY(cg().create_data_set(&in), "create_data_set");
    have_in = true;
    COMGR_TRY(cg().data_set_add(in, src), "data_set_add");
    COMGR_TRY(cg().create_action_info(&ai), "create_action_info");
    have_ai = true;
    COMGR_TRY(cg().action_info_set_language(ai, AMD_COMGR_LANGUAGE_OPENCL_1_2), "set_language");
    COMGR_TRY(cg().action_info_set_isa_name(ai, "amdgcn-amd-amdhsa--gfx950"), "set_isa_name");
    COMGR_TRY(cg().action_info_set_option_list(ai, optp.data(), optp.size()), "set_option_list");
    COMGR_TRY(cg().action_info_set_logging(ai, true), "set_logging");
    COMGR_TRY(cg().create_data_set(&bc), "create_data_set");
    have_bc = true;
    if (cg().do_action(AMD_COMGR_ACTION_COMPILE_SOURCE_WITH_DEVICE_LIBS_TO_BC, ai, in, bc) !=
        AMD_COMGR_STATUS_SUCCESS) {
        append_logs(bc, P.log);
        rc = cfail("OpenCL C compile failed:\n%s", P.log.c_str());
        goto done;
    }
    append_logs(bc, P.log);
    COMGR_TRY(cg().create_data_set(&reloc), "create_data_set");
    have_reloc = true;
    {
        // code generation runs the optimisation pipeline only when asked to
        // (an empty option list compiles the bitcode at -O0: nothing inlined,
        // every value spilled to lanes); the target feature is repeated so
        // no packed FP32 instruction is selected
        static const char* cgopts[] = {"-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
        COMGR_TRY(cg().action_info_set_option_list(ai, cgopts, sizeof cgopts / sizeof cgopts[0]), "set_option_list");
    }
    if (cg().do_action(AMD_COMGR_ACTION_CODEGEN_BC_TO_RELOCATABLE, ai, bc, reloc) != AMD_COMGR_STATUS_SUCCESS) {
        append_logs(reloc, P.log);
        rc = cfail("code generation failed:\n%s", P.log.c_str());
        goto done;
    }
    COMGR_TRY(cg().create_data_set(&out), "create_data_set");
    have_out = true;
    if (cg().do_action(AMD_COMGR_ACTION_LINK_RELOCATABLE_TO_EXECUTABLE, ai, reloc, out) !=
        AMD_COMGR_STATUS_SUCCESS) {
        append_logs(out, P.log);
        rc = cfail("link failed:\n%s", P.log.c_str());
        goto done;
    }
    COMGR_TRY(cg().action_data_get_data(out, AMD_COMGR_DATA_KIND_EXECUTABLE, 0, &exe), "get executable");
    have_exe = true;
    {
        size_t len = 0;
        COMGR_TRY(cg().get_data(exe, &len, nullptr), "get_data");
        P.code.resize(len);
        COMGR_TRY(cg().get_data(exe, &len, P.code.data()), "get_data");
    }
    rc = parse_metadata(exe, P);
done:
    if (have_exe) cg().release_data(exe);
    if (have_src) cg().release_data(src);
    if (have_ai) cg().destroy_action_info(ai);
    if (have_in) cg().destroy_data_set(in);
    if (have_bc) cg().destroy_data_set(bc);
    if (have_reloc) cg().destroy_data_set(reloc);
    if (have_out) cg().destroy_data_set(out);
    return rc;
}

int load_module(ClProgram& P)
{
    int dev = 0;
    CHAI_HIP_TRY(hipGetDevice(&dev));
    if (P.mod && P.device == dev) return CHAI_OK;
    if (P.mod) return cfail("inline OpenCL program was loaded on device %d, not %d", P.device, dev);
    CHAI_HIP_TRY(hipModuleLoadData(&P.mod, P.code.data()));
    P.device = dev;
    for (auto& kv : P.kernels) CHAI_HIP_TRY(hipModuleGetFunction(&kv.second.fn, P.mod, kv.first.c_str()));
    return CHAI_OK;
}

int get_program(const char* text, size_t len, const char* options, ClProgram*& out)
{
    std::string t(text, len), o(options ? options : "");
    const uint64_t h = hash_text(t, o);
    auto it = g_programs.find(h);
    if (it != g_programs.end() && it->second->text == t && it->second->options == o) {
        out = it->second.get();
        return CHAI_OK;
    }
    std::unique_ptr<ClProgram> P(new ClProgram);
    P->text = t;
    P->options = o;
    const int rc = compile_cl(*P);
    if (rc) return rc;
    out = P.get();
    g_programs[h] = std::move(P);
    return CHAI_OK;
}

const char* ocl_scalar_type(int prec)
{
    return prec == CHAI_U32 ? "uint" : prec == CHAI_I32 ? "int" : prec == CHAI_F32 ? "float" : "double";
}

// The parameter list text of kernel `name` in the program text (between the
// parentheses after the name), split at top-level commas.
bool kernel_params(const std::string& text, const std::string& name, std::vector<std::string>& params)
{
    size_t p = 0;
    while ((p = text.find(name, p)) != std::string::npos) {
        const bool left_ok = p == 0 || !(isalnum((unsigned char)text[p - 1]) || text[p - 1] == '_');
        size_t q = p + name.size();
        while (q < text.size() && isspace((unsigned char)text[q])) q++;
        if (left_ok && q < text.size() && text[q] == '(') {
            // the name must follow "kernel ... void"
            const size_t k1 = text.rfind("kernel", p);
            if (k1 != std::string::npos) {
                int depth = 0;
                std::string cur;
                for (size_t i = q; i < text.size(); i++) {
                    const char c = text[i];
                    if (c == '(') { if (depth++ == 0) continue; }
                    if (c == ')') { if (--depth == 0) { params.push_back(cur); return true; } }
                    if (c == ',' && depth == 1) { params.push_back(cur); cur.clear(); continue; }
                    cur += c;
                }
            }
        }
        p = q;
    }
    return false;
}

// A wrapper kernel for a kernel with __local pointer arguments: the wrapper
// takes the other arguments, declares one 16-byte-aligned __local array per
// local argument (sizes fixed at compile time) and calls the user kernel.
int wrapped_program(ClProgram& P, const ClKernel& K, const chai_cl_arg* args, ClProgram*& W, std::string& wname)
{
    std::string key;
    for (size_t i = 0; i < K.args.size(); i++)
        if (K.args[i].kind == CHAI_CL_ARG_LOCAL) key += std::to_string(i) + ":" + std::to_string(args[i].v.count) + ";";
    key = K.name + "|" + key;
    wname = "chai_local_wrap_" + K.name;
    auto it = P.wrapped.find(key);
    if (it != P.wrapped.end()) { W = it->second.get(); return CHAI_OK; }
    std::vector<std::string> params;
    if (!kernel_params(P.text, K.name, params) || params.size() != K.args.size())
        return cfail("cannot find the parameter list of kernel %s in the program text", K.name.c_str());
    std::string decl, call, locals;
    for (size_t i = 0; i < K.args.size(); i++) {
        const std::string nm = "chai_a" + std::to_string(i);
        if (K.args[i].kind == CHAI_CL_ARG_LOCAL) {
            const uint64_t bytes = args[i].v.count;
            const uint64_t words = (bytes + 15) / 16;
            locals += "  __local uint4 " + nm + "[" + std::to_string(words ? words : 1) + "];\n";
            // cast to the parameter's own pointer type: strip the parameter name
            std::string ty = params[i];
            size_t e = ty.find_last_not_of(" \t\n\r");
            size_t b = e;
            while (b != std::string::npos && b > 0 && (isalnum((unsigned char)ty[b - 1]) || ty[b - 1] == '_')) b--;
            ty = ty.substr(0, b);
            call += (call.empty() ? "" : ", ") + std::string("(") + ty + ")" + nm;
        } else {
            decl += (decl.empty() ? "" : ", ") + params[i];
            // the parameter's own name
            std::string ps = params[i];
            size_t e = ps.find_last_not_of(" \t\n\r");
            size_t b = e;
            while (b > 0 && (isalnum((unsigned char)ps[b - 1]) || ps[b - 1] == '_')) b--;
            call += (call.empty() ? "" : ", ") + ps.substr(b, e - b + 1);
        }
    }
    std::string text = P.text + "\n__kernel void " + wname + "(" + decl + ")\n{\n" + locals + "  " + K.name + "(" +
                       call + ");\n}\n";
    std::unique_ptr<ClProgram> NP(new ClProgram);
    NP->text = text;
    NP->options = P.options;
    const int rc = compile_cl(*NP);
    if (rc) return rc;
    if (!NP->kernels.count(wname)) return cfail("wrapper kernel %s missing from its code object", wname.c_str());
    W = NP.get();
    P.wrapped[key] = std::move(NP);
    return CHAI_OK;
}

// Check a launch against the kernel's metadata and, unless `dry`, enqueue it.
// A dry run needs no device (it may compile a wrapper) and accepts null buffers.
int launch_cl(ClProgram& P, const ClKernel& K, const chai_cl_arg* args, uint32_t nargs, uint32_t dims,
              const uint64_t* global, const uint64_t* local, hipStream_t st, bool dry)
{
    if (nargs != K.args.size())
        return cfail("kernel %s takes %zu arguments, %u given", K.name.c_str(), K.args.size(), nargs);
    if (dims < 1 || dims > 2) return cfail("work spaces are 1- or 2-dimensional");
    uint64_t g[3] = {1, 1, 1}, l[3] = {1, 1, 1};
    for (uint32_t d = 0; d < dims; d++) {
        g[d] = global[d];
        l[d] = local ? local[d] : 0;
        if (l[d] == 0) l[d] = g[d] % 64 == 0 ? 64 : 1;      // no local size given: 64 when it divides
        if (g[d] == 0) return cfail("empty global work size");
        if (g[d] % l[d]) return cfail("global size %llu is not a multiple of the local size %llu (OpenCL 1.2 rule)",
                                      (unsigned long long)g[d], (unsigned long long)l[d]);
        if (g[d] / l[d] > 0x7fffffffull) return cfail("too many work groups");
    }
    if (l[0] * l[1] > 1024 || (K.max_wg && l[0] * l[1] > K.max_wg))
        return cfail("local size %llu exceeds the kernel's limit", (unsigned long long)(l[0] * l[1]));
    bool has_local = false;
    for (uint32_t i = 0; i < nargs; i++) {
        const ClArg& A = K.args[i];
        const chai_cl_arg& a = args[i];
        if (A.kind == CHAI_CL_ARG_OTHER) return cfail("argument %u of %s (%s) is not supported", i, K.name.c_str(),
                                                      A.value_kind.c_str());
        if ((int)a.kind != A.kind)
            return cfail("argument %u of %s is a %s, the call passes kind %u", i, K.name.c_str(),
                         A.value_kind.c_str(), a.kind);
        if (A.kind == CHAI_CL_ARG_ARRAY && (A.size != 8 || (!dry && !a.v.ptr)))
            return cfail("argument %u of %s: null or malformed buffer", i, K.name.c_str());
        if (A.kind == CHAI_CL_ARG_SCALAR) {
            const uint32_t need = a.prec == CHAI_F64 ? 8u : 4u;
            if (!valid_prec((int)a.prec) || A.size != need)
                return cfail("argument %u of %s is %u bytes (%s), the call passes a %s", i, K.name.c_str(), A.size,
                             A.type_name.c_str(), ocl_scalar_type((int)a.prec));
        }
        if (A.kind == CHAI_CL_ARG_LOCAL) {
            if (a.v.count == 0 || a.v.count > 65536) return cfail("local argument %u: %llu bytes", i,
                                                                  (unsigned long long)a.v.count);
            has_local = true;
        }
    }
    if (has_local) {
        ClProgram* W = nullptr;
        std::string wname;
        int rc = wrapped_program(P, K, args, W, wname);
        if (rc) return rc;
        std::vector<chai_cl_arg> rest;
        for (uint32_t i = 0; i < nargs; i++)
            if (K.args[i].kind != CHAI_CL_ARG_LOCAL) rest.push_back(args[i]);
        return launch_cl(*W, W->kernels[wname], rest.data(), (uint32_t)rest.size(), dims, g, l, st, dry);
    }
    if (dry) return CHAI_OK;
    const int rc = load_module(P);
    if (rc) return rc;
    // kernelParams: one pointer per explicit argument; the value sits at the
    // start of the argument's union (pointer, 32-bit or 64-bit scalar)
    std::vector<void*> params(nargs ? nargs : 1);
    for (uint32_t i = 0; i < nargs; i++) params[i] = const_cast<void*>(static_cast<const void*>(&args[i].v));
    CHAI_HIP_TRY(hipModuleLaunchKernel(K.fn, (unsigned)(g[0] / l[0]), (unsigned)(g[1] / l[1]), 1, (unsigned)l[0],
                                       (unsigned)l[1], 1, 0, st, params.data(), nullptr));
    return CHAI_OK;
}

}  // namespace
}  // namespace chai_hip

using namespace chai_hip;

extern "C" int chai_hip_cl_build(const char* text, size_t len, const char* options, void** program)
{
    if (!text || !program) return cfail("null program text or handle");
    std::lock_guard<std::mutex> lk(g_mu);
    ClProgram* P = nullptr;
    const int rc = get_program(text, len, options, P);
    if (rc) return rc;
    *program = P;
    return CHAI_OK;
}

extern "C" int chai_hip_cl_build_log(void* program, char* buf, size_t len)
{
    if (!program || !buf || !len) return cfail("null argument");
    const std::string& s = static_cast<ClProgram*>(program)->log;
    snprintf(buf, len, "%s", s.c_str());
    return CHAI_OK;
}

extern "C" int chai_hip_cl_code_object(void* program, void* buf, size_t len, size_t* size)
{
    if (!program || !size) return cfail("null argument");
    const std::vector<char>& code = static_cast<ClProgram*>(program)->code;
    *size = code.size();
    if (buf && len) std::memcpy(buf, code.data(), std::min(len, code.size()));
    return CHAI_OK;
}

extern "C" int chai_hip_cl_kernel_args(void* program, const char* name, int* num_args, chai_cl_arg_info* info,
                                       int max_info)
{
    if (!program || !name || !num_args) return cfail("null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    ClProgram& P = *static_cast<ClProgram*>(program);
    auto it = P.kernels.find(name);
    if (it == P.kernels.end()) return cfail("program has no kernel named %s", name);
    const ClKernel& K = it->second;
    *num_args = (int)K.args.size();
    for (int i = 0; i < (int)K.args.size() && i < max_info && info; i++) {
        std::memset(&info[i], 0, sizeof info[i]);
        info[i].kind = (uint32_t)K.args[i].kind;
        info[i].size = K.args[i].size;
        snprintf(info[i].type_name, sizeof info[i].type_name, "%s", K.args[i].type_name.c_str());
        snprintf(info[i].name, sizeof info[i].name, "%s", K.args[i].name.c_str());
    }
    return CHAI_OK;
}

extern "C" int chai_hip_cl_launch(void* program, const char* name, const chai_cl_arg* args, uint32_t nargs,
                                  uint32_t dims, const uint64_t* global, const uint64_t* local, void* stream)
{
    if (!program || !name || (nargs && !args) || !global) return cfail("null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    ClProgram& P = *static_cast<ClProgram*>(program);
    auto it = P.kernels.find(name);
    if (it == P.kernels.end()) return cfail("program has no kernel named %s", name);
    return launch_cl(P, it->second, args, nargs, dims, global, local, (hipStream_t)stream, false);
}

extern "C" int chai_hip_cl_kernel(void* program, const char* name, void** kernel)
{
    if (!program || !name || !kernel) return cfail("null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    ClProgram& P = *static_cast<ClProgram*>(program);
    auto it = P.kernels.find(name);
    if (it == P.kernels.end()) return cfail("program has no kernel named %s", name);
    *kernel = &it->second;
    return CHAI_OK;
}

extern "C" int chai_hip_cl_check(void* kernel, const chai_cl_arg* args, uint32_t nargs, uint32_t dims,
                                 const uint64_t* global, const uint64_t* local)
{
    if (!kernel || (nargs && !args) || !global) return cfail("null argument");
    std::lock_guard<std::mutex> lk(g_mu);
    ClKernel& K = *static_cast<ClKernel*>(kernel);
    return launch_cl(*K.prog, K, args, nargs, dims, global, local, nullptr, true);
}

namespace chai_hip {
// launch by kernel handle (the trace compiler's inline-kernel statements)
int cl_launch_kernel(void* kernel, const chai_cl_arg* args, uint32_t nargs, uint32_t dims, const uint64_t* global,
                     const uint64_t* local, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    ClKernel& K = *static_cast<ClKernel*>(kernel);
    return launch_cl(*K.prog, K, args, nargs, dims, global, local, st, false);
}
}  // namespace chai_hip
