// make_rng_kat.cpp -- writes tests/golden/rng_kat.json from rocRAND's
// independent host implementations of Philox4x32-10 and Threefry2x64-20
// (ROCm 7.2 /opt/rocm/include/rocrand).  The fixture pins the oracle's block
// functions (oracle/rng.c).  rocRAND's threefry4x32_20 is not used: its round
// permutation (round_idx & 2) matches neither Random123 nor Chai's emitter.
//
// Build and run (host only, no GPU needed):
//   hipcc -O1 -o /tmp/make_rng_kat tests/golden/make_rng_kat.cpp && /tmp/make_rng_kat > tests/golden/rng_kat.json
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
#include <rocrand/rocrand_threefry2x64_20.h>

// Expose rocRAND's protected block functions.
struct Philox : rocrand_device::philox4x32_10_engine {
    static uint4 block(uint4 c, uint2 k) { Philox p; return p.ten_rounds(c, k); }
};
struct Threefry2 : rocrand_device::threefry2x64_20_engine {
    static ulonglong2 block(ulonglong2 c, ulonglong2 k) { return threefry_rounds(c, k); }
};

static unsigned lcg(unsigned& s) { s = s * 1664525u + 1013904223u; return s; }

int main()
{
    const unsigned special[3][6] = {
        {0u, 0u, 0u, 0u, 0u, 0u},
        {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
        {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u}};
    unsigned s = 12345u;
    std::printf("{\n \"source\": \"rocRAND (ROCm 7.2) host implementations\",\n");
    std::printf(" \"philox4x32_10\": [\n");
    for (int i = 0; i < 19; i++) {
        unsigned v[6];
        for (int j = 0; j < 6; j++) v[j] = i < 3 ? special[i][j] : lcg(s);
        uint4 r = Philox::block(make_uint4(v[0], v[1], v[2], v[3]), make_uint2(v[4], v[5]));
        std::printf("  {\"ctr\": [%u, %u, %u, %u], \"key\": [%u, %u], \"out\": [%u, %u, %u, %u]}%s\n",
                    v[0], v[1], v[2], v[3], v[4], v[5], r.x, r.y, r.z, r.w, i == 18 ? "" : ",");
    }
    std::printf(" ],\n \"threefry2x64_20\": [\n");
    for (int i = 0; i < 19; i++) {
        unsigned long long v[4];
        for (int j = 0; j < 4; j++)
            v[j] = i == 0 ? 0ull : i == 1 ? ~0ull : ((unsigned long long)lcg(s) << 32) | lcg(s);
        ulonglong2 c; c.x = v[0]; c.y = v[1];
        ulonglong2 k; k.x = v[2]; k.y = v[3];
        ulonglong2 r = Threefry2::block(c, k);
        std::printf("  {\"ctr\": [%llu, %llu], \"key\": [%llu, %llu], \"out\": [%llu, %llu]}%s\n",
                    v[0], v[1], v[2], v[3], r.x, r.y, i == 18 ? "" : ",");
    }
    std::printf(" ]\n}\n");
    return 0;
}
