// r123_host.cl -- host-callable entry points onto the Random123 OpenCL text
// the reference emits (oracle/_ref/r123_emitted.cl, written by r123_emit),
// compiled as OpenCL C 1.2 for x86-64 by `make -C oracle ref` into
// oracle/_ref/libchai_r123.so.
// TEST INFRASTRUCTURE: used only by tests/ to pin the oracle's block
// functions (Philox4x32/2x32-10, Threefry4x32/2x32/2x64-20 as Chai emits
// them) and the stream's key / counter layout.  Only integer code is called:
// the text's floating-point wrappers use OpenCL built-ins (convert_float4,
// log, sin) that have no host implementation, and Philox2x64 uses mul_hi.
#include "r123_emitted.cl"

void chai_ref_philox4x32(const uint* c, const uint* k, uint* o)
{
    philox4x32_ctr_t C = {{c[0], c[1], c[2], c[3]}};
    philox4x32_key_t K = {{k[0], k[1]}};
    philox4x32_ctr_t R = philox4x32(C, K);
    for (int i = 0; i < 4; i++) o[i] = R.v[i];
}

void chai_ref_philox2x32(const uint* c, const uint* k, uint* o)
{
    philox2x32_ctr_t C = {{c[0], c[1]}};
    philox2x32_key_t K = {{k[0]}};
    philox2x32_ctr_t R = philox2x32(C, K);
    o[0] = R.v[0];
    o[1] = R.v[1];
}

void chai_ref_threefry4x32(const uint* c, const uint* k, uint* o)
{
    threefry4x32_ctr_t C = {{c[0], c[1], c[2], c[3]}};
    threefry4x32_key_t K = {{k[0], k[1], k[2], k[3]}};
    threefry4x32_ctr_t R = threefry4x32(C, K);
    for (int i = 0; i < 4; i++) o[i] = R.v[i];
}

void chai_ref_threefry2x32(const uint* c, const uint* k, uint* o)
{
    threefry2x32_ctr_t C = {{c[0], c[1]}};
    threefry2x32_key_t K = {{k[0], k[1]}};
    threefry2x32_ctr_t R = threefry2x32(C, K);
    o[0] = R.v[0];
    o[1] = R.v[1];
}

void chai_ref_threefry2x64(const ulong* c, const ulong* k, ulong* o)
{
    threefry2x64_ctr_t C = {{c[0], c[1]}};
    threefry2x64_key_t K = {{k[0], k[1]}};
    threefry2x64_ctr_t R = threefry2x64(C, K);
    o[0] = R.v[0];
    o[1] = R.v[1];
}

// The u32 words of one stream vector through the text's own uniform
// wrappers (variant 1230 Philox / 1231 Threefry, vector length 1, 2 or 4):
// these carry the key {tid, ...} and counter {idx, ...} constants.
int chai_ref_stream_words(int variant, int vl, uint tid, uint idx, uint* o)
{
    if (variant == 1230 && vl == 4) {
        uint4 r = philox_uniform_0_4(tid, idx);
        o[0] = r.s0; o[1] = r.s1; o[2] = r.s2; o[3] = r.s3;
    } else if (variant == 1230 && vl == 2) {
        uint2 r = philox_uniform_0_2(tid, idx);
        o[0] = r.s0; o[1] = r.s1;
    } else if (variant == 1230 && vl == 1) {
        o[0] = philox_uniform_0_1(tid, idx);
    } else if (variant == 1231 && vl == 4) {
        uint4 r = threefry_uniform_0_4(tid, idx);
        o[0] = r.s0; o[1] = r.s1; o[2] = r.s2; o[3] = r.s3;
    } else if (variant == 1231 && vl == 2) {
        uint2 r = threefry_uniform_0_2(tid, idx);
        o[0] = r.s0; o[1] = r.s1;
    } else if (variant == 1231 && vl == 1) {
        o[0] = threefry_uniform_0_1(tid, idx);
    } else {
        return -1;
    }
    return 0;
}
