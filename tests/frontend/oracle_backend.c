/*
 * oracle_backend.c -- TEST INFRASTRUCTURE: a stand-in for libchai_mi355x.so
 * that evaluates lowered traces with the CPU oracle (oracle/trace.c) in host
 * memory.  Loaded by the MI355X device of the reference's scheduler through
 * CHAI_MI355X_LIB in the CPU tests, it checks the device's lowering of real
 * VectorTraces (chai_amd/device/DeviceMI355X.cpp) against the reference's
 * own interpreter without a GPU.  Shape inference is the real library's
 * (chai_hip_trace_infer is host code), loaded from CHAI_REAL_LIB.
 *
 * Implements the entry points the device resolves: init, last_error, alloc,
 * free, h2d, d2h, sync, trace_infer, trace_compile, trace_launch,
 * trace_release.  Compile keeps a deep copy of the trace; launch runs
 * chai_or_trace_eval on the host buffers.
 */
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/chai_mi355x.h"
#include "../../oracle/chai_oracle.h"

static char g_err[256];

const char* chai_hip_last_error(void) { return g_err; }
int chai_hip_init(int ordinal) { (void)ordinal; return CHAI_OK; }

int chai_hip_alloc(size_t bytes, void** out)
{
    *out = calloc(1, bytes ? bytes : 8);
    return *out ? CHAI_OK : CHAI_ERR_NOMEM;
}

int chai_hip_free(void* p) { free(p); return CHAI_OK; }

int chai_hip_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    (void)stream;
    memcpy(dst, src, bytes);
    return CHAI_OK;
}

int chai_hip_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    (void)stream;
    memcpy(dst, src, bytes);
    return CHAI_OK;
}

int chai_hip_sync(void* stream) { (void)stream; return CHAI_OK; }

typedef int (*infer_fn)(chai_trace_var*, uint32_t, const chai_trace_stmt*, uint32_t, const chai_trace_node*,
                        uint32_t);
typedef const char* (*err_fn)(void);

typedef int (*check_fn)(const chai_trace*);

static void* real_lib(void)
{
    static void* real = NULL;
    const char* path = getenv("CHAI_REAL_LIB");
    if (!real && path) real = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!real) snprintf(g_err, sizeof g_err, "oracle backend: cannot load CHAI_REAL_LIB");
    return real;
}

int chai_hip_trace_infer(chai_trace_var* vars, uint32_t nv, const chai_trace_stmt* stmts, uint32_t ns,
                         const chai_trace_node* nodes, uint32_t nn)
{
    void* real = real_lib();
    if (!real) return CHAI_ERR_STATE;
    infer_fn f = (infer_fn)dlsym(real, "chai_hip_trace_infer");
    err_fn e = (err_fn)dlsym(real, "chai_hip_last_error");
    const int rc = f(vars, nv, stmts, ns, nodes, nn);
    if (rc != CHAI_OK) snprintf(g_err, sizeof g_err, "%s", e());
    return rc;
}

typedef struct {
    chai_trace t;
    chai_trace_var* vars;
    chai_trace_stmt* stmts;
    chai_trace_node* nodes;
} Prog;

int chai_hip_trace_compile(const chai_trace* t, void** handle)
{
    /* the real library's front-end checks first: refuse what the GPU refuses */
    void* real = real_lib();
    if (!real) return CHAI_ERR_STATE;
    check_fn chk = (check_fn)dlsym(real, "chai_hip_trace_check");
    err_fn e = (err_fn)dlsym(real, "chai_hip_last_error");
    if (chk(t) != CHAI_OK) {
        snprintf(g_err, sizeof g_err, "%s", e());
        return CHAI_ERR_ARG;
    }
    Prog* p = (Prog*)calloc(1, sizeof(Prog));
    p->vars = (chai_trace_var*)malloc(sizeof(chai_trace_var) * t->num_vars);
    p->stmts = (chai_trace_stmt*)malloc(sizeof(chai_trace_stmt) * t->num_stmts);
    p->nodes = (chai_trace_node*)malloc(sizeof(chai_trace_node) * t->num_nodes);
    memcpy(p->vars, t->vars, sizeof(chai_trace_var) * t->num_vars);
    memcpy(p->stmts, t->stmts, sizeof(chai_trace_stmt) * t->num_stmts);
    memcpy(p->nodes, t->nodes, sizeof(chai_trace_node) * t->num_nodes);
    p->t.vars = p->vars; p->t.num_vars = t->num_vars;
    p->t.stmts = p->stmts; p->t.num_stmts = t->num_stmts;
    p->t.nodes = p->nodes; p->t.num_nodes = t->num_nodes;
    *handle = p;
    return CHAI_OK;
}

int chai_hip_trace_launch(void* handle, void* const* buffers, void* stream)
{
    (void)stream;
    Prog* p = (Prog*)handle;
    if (chai_or_trace_eval((const struct chai_trace*)&p->t, buffers) != 0) {
        snprintf(g_err, sizeof g_err, "oracle backend: trace evaluation failed");
        return CHAI_ERR_ARG;
    }
    return CHAI_OK;
}

int chai_hip_trace_release(void* handle)
{
    Prog* p = (Prog*)handle;
    if (!p) return CHAI_OK;
    free(p->vars);
    free(p->stmts);
    free(p->nodes);
    free(p);
    return CHAI_OK;
}
