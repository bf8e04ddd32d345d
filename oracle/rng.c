/*
 * rng.c -- oracle for Chai's counter-based RNG stream.  TEST INFRASTRUCTURE.
 *
 * Restates:
 *   - Philox2x32/4x32/2x64-10 and Threefry2x32/4x32/2x64-20 as the reference
 *     emits them for its OpenCL JIT (Random123/R123Philox.cpp:34-286,
 *     Random123/R123Threefry.cpp:131-290; Random123 1.06, D. E. Shaw Research,
 *     vendored in the reference as text emitters).
 *   - key/counter layout per generated vector (R123Philox.cpp:373-381,
 *     R123Threefry.cpp:377-387): key = {vidx, 0xdecafbad[, 0xfacebead,
 *     0x12345678]}, ctr = {seed, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d}
 *     (64-bit: key = {vidx[, 0xdecafbad]}, ctr = {seed, 0xf00dcafe}).
 *   - vidx = linear vector index over the width-padded array including slot
 *     tiles (kernel/Subscript.cpp:359-380, misc/PrecType.cpp:120-133).
 *   - integer words returned raw; float = (float)(int32)w * 2^-32 + 0.5,
 *     double = (double)(int64)w * 2^-64 + 0.5 (R123Philox.cpp:405-428,
 *     R123Util.cpp:98).
 *   - limits: u * (hi - lo) [+ lo when lo != 0] (enqueue/EnqueueAST.cpp:534-545).
 *   - Box-Muller normals: vecLen 4 -> [r0 sin(2pi u1), r0 cos(2pi u1),
 *     r2 sin(2pi u3), r2 cos(2pi u3)], r = sqrt(-2 ln u) (R123Philox.cpp:480-507).
 *
 * Deliberate differences from the reference, fixed here as the canonical stream:
 *   - 2*pi is the exact double constant, not the 6-digit "3.14159" the JIT prints.
 *   - log/sin/cos are this file's own routines (series given below, no vendor
 *     math library): double precision for the f64 stream, single precision for
 *     the f32 stream, every operation an IEEE op, so host and device agree bit
 *     for bit.
 *   - The reference interpreter's RNG (interp/InterpRNG.cpp, glibc random_r,
 *     seeded from an address) is not reproducible and is not followed.
 *
 *   - Threefry4x32: Chai's emitter mixes word pairs (0,1),(2,3) in every round
 *     (R123Threefry.cpp:238-255) and drops Random123's (0,3),(2,1) permutation
 *     on odd rounds; the stream follows Chai's emitter.
 *
 * Pinned by: the Random123 1.06 published known-answer vectors for all six
 * block functions (tests/golden/r123_kat.json; Threefry4x32 through the
 * standard form chai_or_threefry4x32_20_r123, which shares key schedule and
 * rotations with Chai's variant), and rocRAND's independent Philox4x32-10 and
 * Threefry2x64-20 (tests/golden/rng_kat.json, from tests/golden/make_rng_kat.cpp).
 * Chai's permutation-free Threefry4x32 itself has no independent implementation:
 * beyond the shared parts it is "parity unpinned".
 * Uniform conversion pinned by closed-form edge cases; log/sincos pinned
 * against glibc double log/sin/cos (<= 2 ulp) and by the closed-form values
 * sin(2pi*k/8).
 */
#include <math.h>
#include <string.h>
#include "chai_oracle.h"


/* ------------------------------------------------------------------ Philox */

static uint32_t mulhilo32(uint32_t a, uint32_t b, uint32_t* hi)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    return (uint32_t)p;
}

static uint64_t mulhilo64(uint64_t a, uint64_t b, uint64_t* hi)
{
    unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
    *hi = (uint64_t)(p >> 64);
    return (uint64_t)p;
}

void chai_or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, hi1;
        uint32_t lo0 = mulhilo32(0xD2511F53u, c0, &hi0);
        uint32_t lo1 = mulhilo32(0xCD9E8D57u, c2, &hi1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void chai_or_philox2x32_10(const uint32_t ctr[2], const uint32_t key[1], uint32_t out[2])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], k0 = key[0];
    for (int r = 0; r < 10; r++) {
        if (r) k0 += 0x9E3779B9u;
        uint32_t hi;
        uint32_t lo = mulhilo32(0xd256d193u, c0, &hi);
        uint32_t n0 = hi ^ k0 ^ c1;
        c0 = n0; c1 = lo;
    }
    out[0] = c0; out[1] = c1;
}

void chai_or_philox2x64_10(const uint64_t ctr[2], const uint64_t key[1], uint64_t out[2])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], k0 = key[0];
    for (int r = 0; r < 10; r++) {
        if (r) k0 += 0x9E3779B97F4A7C15ull;
        uint64_t hi;
        uint64_t lo = mulhilo64(0xD2B74407B1CE6E93ull, c0, &hi);
        uint64_t n0 = hi ^ k0 ^ c1;
        c0 = n0; c1 = lo;
    }
    out[0] = c0; out[1] = c1;
}

/* ---------------------------------------------------------------- Threefry */

static uint32_t rotl32(uint32_t x, unsigned n) { return (x << (n & 31)) | (x >> ((32 - n) & 31)); }
static uint64_t rotl64(uint64_t x, unsigned n) { return (x << (n & 63)) | (x >> ((64 - n) & 63)); }

/* Rotation schedules of Threefish as used by Random123 (Salmon et al.,
 * SC'11, "Parallel random numbers: as easy as 1, 2, 3", table of rotation
 * constants for the 4x32, 2x32 and 2x64 variants). */
static const unsigned R4x32[8][2] = {{10, 26}, {11, 21}, {13, 27}, {23, 5},
                                     {6, 20},  {17, 11}, {25, 10}, {18, 20}};
static const unsigned R2x32[8] = {13, 15, 26, 6, 17, 29, 16, 24};
static const unsigned R2x64[8] = {16, 42, 12, 31, 16, 32, 24, 21};

/* permute = 1: Random123's Threefry4x32-20 (odd rounds mix (0,3),(2,1)).
 * permute = 0: the variant Chai's emitter writes (R123Threefry.cpp:238-255 mixes
 * (0,1),(2,3) in every round); this one defines Chai's stream. */
static void threefry4x32_20_impl(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4],
                                 int permute)
{
    uint32_t ks[5];
    ks[4] = 0x1BD11BDAu;
    uint32_t X[4];
    for (int i = 0; i < 4; i++) { ks[i] = key[i]; ks[4] ^= key[i]; X[i] = ctr[i] + ks[i]; }
    for (int r = 0; r < 20; r++) {
        const unsigned* R = R4x32[r & 7];
        if (!permute || (r & 1) == 0) {
            X[0] += X[1]; X[1] = rotl32(X[1], R[0]); X[1] ^= X[0];
            X[2] += X[3]; X[3] = rotl32(X[3], R[1]); X[3] ^= X[2];
        } else {
            X[0] += X[3]; X[3] = rotl32(X[3], R[0]); X[3] ^= X[0];
            X[2] += X[1]; X[1] = rotl32(X[1], R[1]); X[1] ^= X[2];
        }
        if ((r & 3) == 3) {
            uint32_t s = (uint32_t)((r + 1) >> 2);
            for (int i = 0; i < 4; i++) X[i] += ks[(s + i) % 5];
            X[3] += s;
        }
    }
    memcpy(out, X, sizeof X);
}

void chai_or_threefry4x32_20(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4])
{
    threefry4x32_20_impl(ctr, key, out, 0);
}

void chai_or_threefry4x32_20_r123(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4])
{
    threefry4x32_20_impl(ctr, key, out, 1);
}

void chai_or_threefry2x32_20(const uint32_t ctr[2], const uint32_t key[2], uint32_t out[2])
{
    uint32_t ks[3];
    ks[2] = 0x1BD11BDAu;
    uint32_t X[2];
    for (int i = 0; i < 2; i++) { ks[i] = key[i]; ks[2] ^= key[i]; X[i] = ctr[i] + ks[i]; }
    for (int r = 0; r < 20; r++) {
        X[0] += X[1]; X[1] = rotl32(X[1], R2x32[r & 7]); X[1] ^= X[0];
        if ((r & 3) == 3) {
            uint32_t s = (uint32_t)((r + 1) >> 2);
            X[0] += ks[s % 3]; X[1] += ks[(s + 1) % 3]; X[1] += s;
        }
    }
    out[0] = X[0]; out[1] = X[1];
}

void chai_or_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2])
{
    uint64_t ks[3];
    ks[2] = 0x1BD11BDAA9FC1A22ull;
    uint64_t X[2];
    for (int i = 0; i < 2; i++) { ks[i] = key[i]; ks[2] ^= key[i]; X[i] = ctr[i] + ks[i]; }
    for (int r = 0; r < 20; r++) {
        X[0] += X[1]; X[1] = rotl64(X[1], R2x64[r & 7]); X[1] ^= X[0];
        if ((r & 3) == 3) {
            uint64_t s = (uint64_t)((r + 1) >> 2);
            X[0] += ks[s % 3]; X[1] += ks[(s + 1) % 3]; X[1] += s;
        }
    }
    out[0] = X[0]; out[1] = X[1];
}

/* ------------------------------------------------------- Box-Muller math */

static double from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* log(x) = e*ln2 + 2*atanh(s), s = f/(2+f), f = m-1, m in [sqrt(1/2), sqrt(2)).
 * 2*atanh(s) = 2s + s*z*P(z), z = s^2, P(z) = sum_{k=1..10} 2/(2k+1) z^(k-1)
 * (Taylor series of atanh; |s| <= 0.1716 so the truncation is below 2^-56).
 * ln2 is split hi/lo so e*ln2_hi is exact (fdlibm's split). */
double chai_or_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (isinf(x)) return x;
    uint64_t b = to_bits(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {                       /* subnormal */
        x = x * 0x1p54;
        b = to_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 0x1.8618618618618p-4;
    p = fma(p, z, 0x1.af286bca1af28p-4);
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-4);
    p = fma(p, z, 0x1.1111111111111p-3);
    p = fma(p, z, 0x1.3b13b13b13b14p-3);
    p = fma(p, z, 0x1.745d1745d1746p-3);
    p = fma(p, z, 0x1.c71c71c71c71cp-3);
    p = fma(p, z, 0x1.2492492492492p-2);
    p = fma(p, z, 0x1.999999999999ap-2);
    p = fma(p, z, 0x1.5555555555555p-1);
    double sz = s * z;
    double logm = fma(sz, p, 2.0 * s);
    double de = (double)e;
    return fma(de, 0x1.62e42fee00000p-1, fma(de, 0x1.a39ef35793c76p-33, logm));
}

/* sin(2*pi*u), cos(2*pi*u): quadrant q = rint(4u), r = u - q/4 (exact),
 * x = 2*pi*r in [-pi/4, pi/4], Taylor series through x^17 (sin) and x^18 (cos). */
void chai_or_sincos2pi(double u, double* sp, double* cp)
{
    double q = rint(4.0 * u);
    double r = u - 0.25 * q;
    int qi = ((int)(long long)q) & 3;
    double x = r * 0x1.921fb54442d18p+2;
    double x2 = x * x;
    double ps = 0x1.952c77030ad4ap-49;
    ps = fma(ps, x2, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, x2, 0x1.6124613a86d09p-33);
    ps = fma(ps, x2, -0x1.ae64567f544e4p-26);
    ps = fma(ps, x2, 0x1.71de3a556c734p-19);
    ps = fma(ps, x2, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, x2, 0x1.1111111111111p-7);
    ps = fma(ps, x2, -0x1.5555555555555p-3);
    double sn = fma(x * x2, ps, x);
    double pc = -0x1.6827863b97d97p-53;
    pc = fma(pc, x2, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, x2, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, x2, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, x2, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, x2, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, x2, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, x2, 0x1.5555555555555p-5);
    pc = fma(pc, x2, -0.5);
    double cs = fma(x2, pc, 1.0);
    double s, c;
    switch (qi) {
        case 0: s = sn;  c = cs;  break;
        case 1: s = cs;  c = -sn; break;
        case 2: s = -sn; c = -cs; break;
        default: s = -cs; c = sn; break;
    }
    *sp = s; *cp = c;
}

/* f32 normals are evaluated in f32 (the f32 stream's own routines; every
 * operation an IEEE f32 op with explicit fmaf and correctly rounded divide /
 * sqrt, exactly as the device code does):
 * log(x) = e*ln2 + 2s + s z P(z), s = f/(2+f), z = s^2, P(z) = sum_{k=1..5} 2/(2k+1) z^(k-1);
 * sin/cos(2 pi u) from q = rint(4u), t = u - q/4, series in t with coefficients
 * (2 pi)^k / k! through t^9 / t^10 (truncation below 2^-27 for |t| <= 1/8). */
static float f_from_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t f_to_bits(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

float chai_or_logf(float x)
{
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (isinf(x)) return x;
    uint32_t b = f_to_bits(x);
    int e = (int)((b >> 23) & 0xff);
    if (e == 0) {
        x = x * 0x1p25f;
        b = f_to_bits(x);
        e = (int)((b >> 23) & 0xff) - 25;
    }
    e -= 127;
    float m = f_from_bits((b & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 0x1.745d18p-3f;
    p = fmaf(p, z, 0x1.c71c72p-3f);
    p = fmaf(p, z, 0x1.24924ap-2f);
    p = fmaf(p, z, 0x1.99999ap-2f);
    p = fmaf(p, z, 0x1.555556p-1f);
    float logm = fmaf(s * z, p, 2.0f * s);
    float de = (float)e;
    return fmaf(de, 0x1.62e4p-1f, fmaf(de, 0x1.7f7d1cp-20f, logm));
}

void chai_or_sincos2pif(float u, float* sp, float* cp)
{
    float q = rintf(4.0f * u);
    float t = u - 0.25f * q;
    int qi = ((int)q) & 3;
    float t2 = t * t;
    float ps = 0x1.507834p+5f;
    ps = fmaf(ps, t2, -0x1.32d2ccp+6f);
    ps = fmaf(ps, t2, 0x1.466bc6p+6f);
    ps = fmaf(ps, t2, -0x1.4abbcep+5f);
    float sn = fmaf(t * t2, ps, t * 0x1.921fb6p+2f);
    float pc = -0x1.a6d1f2p+4f;
    pc = fmaf(pc, t2, 0x1.e1f506p+5f);
    pc = fmaf(pc, t2, -0x1.55d3c8p+6f);
    pc = fmaf(pc, t2, 0x1.03c1f0p+6f);
    pc = fmaf(pc, t2, -0x1.3bd3ccp+4f);
    float cs = fmaf(t2, pc, 1.0f);
    *sp = qi == 0 ? sn : qi == 1 ? cs : qi == 2 ? -sn : -cs;
    *cp = qi == 0 ? cs : qi == 1 ? -sn : qi == 2 ? -cs : sn;
}

/* ------------------------------------------------------------ the stream */

static size_t std_vec_len(int prec) { return prec == 8 ? 2 : 4; }

static float u32_to_unit_f32(uint32_t w)
{
    float t = (float)(int32_t)w;
    t = t * 0x1p-32f;
    return t + 0.5f;
}

static double u64_to_unit_f64(uint64_t w)
{
    double t = (double)(int64_t)w;
    t = t * 0x1p-64;
    return t + 0.5;
}

/* one generated vector of up to 4 32-bit words */
static void block32(int variant, size_t vl, uint32_t vidx, uint32_t seed, uint32_t w[4])
{
    if (vl == 4) {
        if (variant == 1231) {
            const uint32_t k[4] = {vidx, 0xdecafbadu, 0xfacebeadu, 0x12345678u};
            const uint32_t c[4] = {seed, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du};
            chai_or_threefry4x32_20(c, k, w);
        } else {
            const uint32_t k[2] = {vidx, 0xdecafbadu};
            const uint32_t c[4] = {seed, 0xf00dcafeu, 0xdeadbeefu, 0xbeeff00du};
            chai_or_philox4x32_10(c, k, w);
        }
    } else {
        const uint32_t c[2] = {seed, 0xf00dcafeu};
        if (variant == 1231) {
            const uint32_t k[2] = {vidx, 0xdecafbadu};
            chai_or_threefry2x32_20(c, k, w);
        } else {
            const uint32_t k[1] = {vidx};
            chai_or_philox2x32_10(c, k, w);
        }
        w[2] = w[3] = 0;
    }
}

static void block64(int variant, uint64_t vidx, uint64_t seed, uint64_t w[2])
{
    const uint64_t c[2] = {seed, 0xf00dcafeull};
    if (variant == 1231) {
        const uint64_t k[2] = {vidx, 0xdecafbadull};
        chai_or_threefry2x64_20(c, k, w);
    } else {
        const uint64_t k[1] = {vidx};
        chai_or_philox2x64_10(c, k, w);
    }
}

/* The raw words of one generated vector: vl words of the 32-bit stream
 * (pr = 32, vl = 1, 2, 4) or 2 words of the 64-bit stream (pr = 64) for
 * vector index vidx and generator seed (the tid / idx arguments of the
 * reference's *_uniform_* wrappers, Random123/R123Philox.cpp:366-381,
 * Random123/R123Threefry.cpp:377-387). */
int chai_or_rng_words(int variant, int pr, int vl, uint64_t vidx, uint64_t seed, uint64_t* out)
{
    if (variant != 1230 && variant != 1231) return -1;
    if (pr == 64) {
        uint64_t w[2];
        block64(variant, vidx, seed, w);
        out[0] = w[0]; out[1] = w[1];
        return 0;
    }
    if (pr != 32 || (vl != 1 && vl != 2 && vl != 4)) return -1;
    uint32_t w[4];
    block32(variant, (size_t)vl, (uint32_t)vidx, (uint32_t)seed, w);
    for (int i = 0; i < vl; i++) out[i] = w[i];
    return 0;
}

int chai_or_rng_fill(int variant, int prec, int vec_len, uint64_t seed,
                     size_t W, size_t H, size_t slots, int normal,
                     int has_limits, double lo, double hi, void* out)
{
    if (variant != 1230 && variant != 1231) return -1;
    if (prec != 0 && prec != 1 && prec != 4 && prec != 8) return -1;
    size_t svl = std_vec_len(prec);
    size_t vl = vec_len ? (size_t)vec_len : svl;
    if (prec == 8 ? (vl != 1 && vl != 2) : (vl != 1 && vl != 2 && vl != 4)) return -1;
    if (normal && prec != 4 && prec != 8) return -1;
    size_t Wp = (W % svl) ? svl * (1 + W / svl) : W;   /* PrecType::padWidth */
    size_t vpr = Wp / vl;                               /* vectors per row */

    for (size_t s = 0; s < slots; s++)
    for (size_t y = 0; y < H; y++)
    for (size_t x0 = 0; x0 < W; x0 += vl) {
        size_t vidx = s * H * vpr + y * vpr + x0 / vl;
        size_t base = (s * H + y) * W + x0;
        size_t n = (W - x0 < vl) ? W - x0 : vl;
        if (prec == 8) {
            uint64_t w[2];
            block64(variant, (uint64_t)vidx, seed, w);
            double u0 = u64_to_unit_f64(w[0]), u1 = u64_to_unit_f64(w[1]);
            double v[2];
            if (normal) {
                double r = sqrt(-2.0 * chai_or_log(u0)), sn, cs;
                chai_or_sincos2pi(u1, &sn, &cs);
                v[0] = r * sn; v[1] = r * cs;
            } else {
                v[0] = u0; v[1] = u1;
                if (has_limits) {
                    double d = hi - lo;
                    for (int i = 0; i < 2; i++) { v[i] = v[i] * d; if (lo != 0.0) v[i] = v[i] + lo; }
                }
            }
            for (size_t i = 0; i < n; i++) ((double*)out)[base + i] = v[i];
        } else {
            uint32_t w[4];
            block32(variant, vl, (uint32_t)vidx, (uint32_t)seed, w);
            if (prec == 0) {
                for (size_t i = 0; i < n; i++) ((uint32_t*)out)[base + i] = w[i];
            } else if (prec == 1) {
                for (size_t i = 0; i < n; i++) ((int32_t*)out)[base + i] = (int32_t)w[i];
            } else if (!normal) {
                float v[4];
                for (int i = 0; i < 4; i++) v[i] = u32_to_unit_f32(w[i]);
                if (has_limits) {
                    float d = (float)(hi - lo), l = (float)lo;
                    for (int i = 0; i < 4; i++) { v[i] = v[i] * d; if (lo != 0.0) v[i] = v[i] + l; }
                }
                for (size_t i = 0; i < n; i++) ((float*)out)[base + i] = v[i];
            } else {
                float v[4];
                for (int p = 0; p < 2; p++) {
                    float r = sqrtf(-2.0f * chai_or_logf(u32_to_unit_f32(w[2 * p]))), sn, cs;
                    chai_or_sincos2pif(u32_to_unit_f32(w[2 * p + 1]), &sn, &cs);
                    v[2 * p] = r * sn;
                    v[2 * p + 1] = r * cs;
                }
                for (size_t i = 0; i < n; i++) ((float*)out)[base + i] = v[i];
            }
        }
    }
    return 0;
}
