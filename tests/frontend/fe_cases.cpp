// fe_cases.cpp -- Chai programs written against the reference's own API
// (api/chai.h), run by the reference's scheduler.  TEST PROGRAM: built by
// tests/frontend/Makefile against the reference library compiled here
// (oracle/frontend.mk) and the MI355X device (chai_amd/device/).
//
//   fe_cases <case> <dir>
//   fe_cases mm_time <n>       (the interpreter's own speed on an n^3 product)
//
// reads the case's inputs from <dir>/<case>.<name>.bin (raw little-endian
// arrays, shapes fixed per case below), evaluates the case with chai.h and
// writes each result to <dir>/<case>.out.<name>.bin.  Which devices run it is
// the runtime's business: the interpreter always exists; CHAI_MI355X=0 adds
// the MI355X device, which takes every trace it can and hands the rest back.
#include <chai/chai.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

using namespace chai;
using namespace std;

static string g_dir, g_case;

template <typename X>
static vector<X> load(const char* name, size_t n)
{
    vector<X> v(n);
    const string path = g_dir + "/" + g_case + "." + name + ".bin";
    FILE* f = fopen(path.c_str(), "rb");
    if (!f || fread(&v[0], sizeof(X), n, f) != n) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
    return v;
}

template <typename X>
static void save(const char* name, const X* p, size_t n)
{
    const string path = g_dir + "/" + g_case + ".out." + name + ".bin";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || fwrite(p, sizeof(X), n, f) != n) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

template <typename X>
static void save(const char* name, const vector<X>& v) { save(name, &v[0], v.size()); }

// config 1: C = matmul(A, B), f32 256 x 256 x 256
static void case_mm_f32(void)
{
    const size_t n = 256;
    vector<float> a = load<float>("A", n * n), b = load<float>("B", n * n), c(n * n);
    {
        Arrayf32 A = make2(n, n, &a[0]);
        Arrayf32 B = make2(n, n, &b[0]);
        Arrayf32 C = matmul(A, B);
        C.read2(&c[0], c.size() * sizeof(float));
    }
    save("C", c);
}

// operand transposes folded into the product, general form, matrix-vector
static void case_mm_shapes(void)
{
    const size_t M = 96, K = 160, N = 72;
    vector<float> a = load<float>("A", M * K), at = load<float>("At", K * M), b = load<float>("B", K * N),
                  bt = load<float>("Bt", N * K), c0 = load<float>("C0", M * N), x = load<float>("x", K);
    vector<float> c1(M * N), c2(M * N), c3(M * N), c4(M * N), y(M);
    {
        Arrayf32 A = make2(K, M, &a[0]);
        Arrayf32 At = make2(M, K, &at[0]);
        Arrayf32 B = make2(N, K, &b[0]);
        Arrayf32 Bt = make2(K, N, &bt[0]);
        Arrayf32 C0 = make2(N, M, &c0[0]);
        Arrayf32 X = make1(K, &x[0]);
        Arrayf32 C1 = matmul(transpose(At), B);
        Arrayf32 C2 = matmul(A, transpose(Bt));
        Arrayf32 C3 = matmul(transpose(At), transpose(Bt));
        Arrayf32 C4 = matmul(A, B, C0, 0.75f, -1.5f);
        Arrayf32 Y = matmul(A, X);
        C1.read2(&c1[0], c1.size() * sizeof(float));
        C2.read2(&c2[0], c2.size() * sizeof(float));
        C3.read2(&c3[0], c3.size() * sizeof(float));
        C4.read2(&c4[0], c4.size() * sizeof(float));
        Y.read1(&y[0], y.size() * sizeof(float));
    }
    save("C1", c1);
    save("C2", c2);
    save("C3", c3);
    save("C4", c4);
    save("y", y);
}

// integer and mixed products: the product in the promoted C type, the sum
// in double, one conversion to the result type (interp/InterpMatmulBase.hpp)
static void case_mm_int(void)
{
    const size_t M = 64, K = 48, N = 40;
    vector<uint32_t> ua = load<uint32_t>("UA", M * K), ub = load<uint32_t>("UB", K * N);
    vector<int32_t> ia = load<int32_t>("IA", M * K), ib = load<int32_t>("IB", K * N);
    vector<float> fb = load<float>("FB", K * N);
    vector<uint32_t> cuu(M * N);
    vector<int32_t> cii(M * N), cui(M * N);
    vector<float> cif(M * N);
    {
        Arrayu32 UA = make2(K, M, &ua[0]);
        Arrayu32 UB = make2(N, K, &ub[0]);
        Arrayi32 IA = make2(K, M, &ia[0]);
        Arrayi32 IB = make2(N, K, &ib[0]);
        Arrayf32 FB = make2(N, K, &fb[0]);
        Arrayu32 CUU = matmul(UA, UB);
        Arrayi32 CII = matmul(IA, IB);
        Arrayi32 CUI = matmul(UA, IB);
        Arrayf32 CIF = matmul(IA, FB);
        CUU.read2(&cuu[0], cuu.size() * sizeof(uint32_t));
        CII.read2(&cii[0], cii.size() * sizeof(int32_t));
        CUI.read2(&cui[0], cui.size() * sizeof(int32_t));
        CIF.read2(&cif[0], cif.size() * sizeof(float));
    }
    save("CUU", cuu);
    save("CII", cii);
    save("CUI", cui);
    save("CIF", cif);
}

static void case_mm_f64(void)
{
    const size_t M = 64, K = 80, N = 48;
    vector<double> a = load<double>("A", M * K), b = load<double>("B", K * N), c(M * N);
    {
        Arrayf64 A = make2(K, M, &a[0]);
        Arrayf64 B = make2(N, K, &b[0]);
        Arrayf64 C = matmul(A, B);
        C.read2(&c[0], c.size() * sizeof(double));
    }
    save("C", c);
}

// elementwise expressions fused into one kernel, with IEEE-exact ops only
static void case_ew(void)
{
    const size_t W = 64, H = 48;
    vector<float> a = load<float>("A", W * H), b = load<float>("B", W * H);
    vector<float> e(W * H), f(W * H), g(W * H);
    {
        Arrayf32 A = make2(W, H, &a[0]);
        Arrayf32 B = make2(W, H, &b[0]);
        Arrayf32 E = (A * 0.5f + B) * (A - 2.0f) / (fabs(B) + 1.0f);
        Arrayf32 F = sqrt(fabs(A * B)) + chai::max(A, B) - (A > B);
        Arrayf32 G = fma(A, B, E) + floor(A * 8.0f) + (E < F);
        E.read2(&e[0], e.size() * sizeof(float));
        F.read2(&f[0], f.size() * sizeof(float));
        G.read2(&g[0], g.size() * sizeof(float));
    }
    save("E", e);
    save("F", f);
    save("G", g);
}

// reductions read out as scalars
static void case_red(void)
{
    const size_t W = 256, H = 96;
    vector<float> a = load<float>("A", W * H), b = load<float>("B", W * H);
    vector<int32_t> ia = load<int32_t>("IA", W * H);
    float r[4];
    int32_t ri[1];
    {
        Arrayf32 A = make2(W, H, &a[0]);
        Arrayf32 B = make2(W, H, &b[0]);
        Arrayi32 IA = make2(W, H, &ia[0]);
        Arrayf32 S0 = sum(A);
        r[0] = S0.read_scalar();
        Arrayf32 S1 = mean(B * 2.0f);
        r[1] = S1.read_scalar();
        Arrayf32 S2 = dot_product(A, B);
        r[2] = S2.read_scalar();
        Arrayf32 S3 = sum(A * A + B);
        r[3] = S3.read_scalar();
        Arrayi32 S4 = sum(IA);
        ri[0] = S4.read_scalar();
    }
    save("r", r, 4);
    save("ri", ri, 1);
}

// integer elementwise semantics
static void case_ints(void)
{
    const size_t W = 128, H = 16;
    vector<uint32_t> ua = load<uint32_t>("UA", W * H), ub = load<uint32_t>("UB", W * H);
    vector<int32_t> ia = load<int32_t>("IA", W * H), ib = load<int32_t>("IB", W * H);
    vector<uint32_t> x(W * H);
    vector<int32_t> y(W * H);
    {
        Arrayu32 UA = make2(W, H, &ua[0]);
        Arrayu32 UB = make2(W, H, &ub[0]);
        Arrayi32 IA = make2(W, H, &ia[0]);
        Arrayi32 IB = make2(W, H, &ib[0]);
        Arrayu32 X = ((UA ^ UB) + (UA << 3u)) * UB - (UB >> 5u) + rotate(UA, UB) + mul_hi(UA, UB);
        Arrayi32 Y = (IA * IB + IA / (abs(IB) + 1)) - (IA % (abs(IB) + 7)) + add_sat(IA, IB) + (IA < IB);
        X.read2(&x[0], x.size() * sizeof(uint32_t));
        Y.read2(&y[0], y.size() * sizeof(int32_t));
    }
    save("X", x);
    save("Y", y);
}

// a five-point stencil of gathers (sample/filter.cpp's form) and index leaves
static void case_gather(void)
{
    const size_t W = 64, H = 32;
    vector<float> a = load<float>("A", W * H), o(W * H);
    {
        Arrayf32 A = make2(W, H, &a[0]);
        Arrayf32 X = index_f32(0, W, H);
        Arrayf32 Y = index_f32(1, W, H);
        Arrayf32 O = 0.5f * A + 0.125f * (gather2_floor(A, X + 1, Y) + gather2_floor(A, X - 1, Y) +
                                          gather2_floor(A, X, Y + 1) + gather2_floor(A, X, Y - 1));
        O.read2(&o[0], o.size() * sizeof(float));
    }
    save("O", o);
}

// several traces: arrays computed by one trace are read by the next ones
static void case_multi(void)
{
    const size_t W = 128, H = 64;
    vector<float> a = load<float>("A", W * H), b = load<float>("B", W * H);
    vector<float> d(W * H), e(W * H);
    float s[2];
    {
        Arrayf32 A = make2(W, H, &a[0]);
        Arrayf32 B = make2(W, H, &b[0]);
        Arrayf32 D = A + B;
        D.read2(&d[0], d.size() * sizeof(float));
        Arrayf32 S0 = sum(D);
        s[0] = S0.read_scalar();
        Arrayf32 E = D * D - A;
        for (int i = 0; i < 5; i++) E += 0.5f * B;
        E.read2(&e[0], e.size() * sizeof(float));
        Arrayf32 S1 = sum(E - D);
        s[1] = S1.read_scalar();
    }
    save("D", d);
    save("E", e);
    save("s", s, 2);
}

// the interpreter's own speed: C = matmul(A, B) on n x n f32 (uniform in
// [-1, 1) from a fixed LCG), timed from the statements to the readout that
// evaluates them; prints "seconds <t> flops <2 n^3>"
static void case_mm_time(size_t n)
{
    vector<float> a(n * n), b(n * n), c(n * n);
    uint32_t s = 42u;
    for (size_t i = 0; i < n * n; i++) { s = s * 1664525u + 1013904223u; a[i] = (float)(s >> 8) * 0x1p-23f - 1.0f; }
    for (size_t i = 0; i < n * n; i++) { s = s * 1664525u + 1013904223u; b[i] = (float)(s >> 8) * 0x1p-23f - 1.0f; }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    {
        Arrayf32 A = make2(n, n, &a[0]);
        Arrayf32 B = make2(n, n, &b[0]);
        Arrayf32 C = matmul(A, B);
        C.read2(&c[0], c.size() * sizeof(float));
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    const double dt = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    printf("seconds %.6f flops %.6e checksum %.6e\n", dt, 2.0 * n * n * n, (double)c[n * n / 2]);
}

int main(int argc, char* argv[])
{
    if (argc == 3 && string(argv[1]) == "mm_time") {
        init("");
        case_mm_time((size_t)atol(argv[2]));
        shutdown();
        return 0;
    }
    if (argc != 3) {
        fprintf(stderr, "usage: %s <case> <dir> | mm_time <n>\n", argv[0]);
        return 1;
    }
    g_case = argv[1];
    g_dir = argv[2];
    init("");                        // the interpreter is always present
    if (g_case == "mm_f32") case_mm_f32();
    else if (g_case == "mm_shapes") case_mm_shapes();
    else if (g_case == "mm_int") case_mm_int();
    else if (g_case == "mm_f64") case_mm_f64();
    else if (g_case == "ew") case_ew();
    else if (g_case == "red") case_red();
    else if (g_case == "ints") case_ints();
    else if (g_case == "gather") case_gather();
    else if (g_case == "multi") case_multi();
    else {
        fprintf(stderr, "unknown case %s\n", g_case.c_str());
        shutdown();
        return 1;
    }
    shutdown();
    return 0;
}
