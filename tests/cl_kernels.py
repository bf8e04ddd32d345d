"""OpenCL C kernels for the inline-kernel (ProgramCL) tests, each doing what a
reference sample or ProgramCL feature exercises:

- ADD_SCALAR: a[i] += s, the kernel of sample/inline_cl.cpp;
- BITONIC: one pass of a bitonic sorting network over uint keys, the
  computation of sample/sorting/bitonic.cpp;
- AMERICAN: a recombining binomial tree for American calls in double, with
  static __local arrays, the computation of
  sample/finance/american_option_batch.cpp (sizes from -D options);
- BLOCK_SUM: per-work-group sums through a __local pointer argument
  (ProgramCL's LOCAL_MEM argument kind, api/ProgramCL.cpp:221-240);
- SCALE_T: a 2-D work space (ProgramCL::Kernel::operator() with two global
  and two local sizes, api/ProgramCL.cpp:336-342);
- PROBE: work-item functions written to clamped indices (never out of bounds
  whatever the ids are), to check the work-space mapping first.
"""

ADD_SCALAR = """
__kernel void add_scalar(__global float* a, float s)
{
    a[get_global_id(0)] += s;
}
"""

BITONIC = """
__kernel void bitonic_pass(__global uint* keys, const uint stage, const uint pass,
                           const uint width, const uint ascending)
{
    const uint t = get_global_id(0);
    const uint dist = 1u << (stage - pass);
    const uint lo = (t % dist) + (t / dist) * 2u * dist;
    const uint hi = lo + dist;
    uint up = ascending;
    if ((t >> stage) & 1u) up = 1u - up;
    const uint a = keys[lo], b = keys[hi];
    const uint mn = min(a, b), mx = max(a, b);
    keys[lo] = up ? mn : mx;
    keys[hi] = up ? mx : mn;
}
"""

AMERICAN = """
#pragma OPENCL EXTENSION cl_khr_fp64 : enable
__kernel void american_call(__global double* value, __global const double* strike,
                            __global const double* vol, const double s0, const double r,
                            const double T)
{
    __local double px[(STEPS + 1) * GROUP];
    __local double opt[(STEPS + 1) * GROUP];
    const uint base = (STEPS + 1) * get_local_id(0);
    const uint g = get_global_id(0);
    const double dt = T / STEPS;
    const double growth = exp(r * dt);
    const double u = exp(vol[g] * sqrt(dt));
    const double d = 1.0 / u;
    const double p = (growth - d) / (u - d);
    const double K = strike[g];
    double s = s0 * pown(d, STEPS);
    for (int i = 0; i <= STEPS; i++) {
        px[base + i] = s;
        opt[base + i] = max(s - K, 0.0);
        s = s * (u * u);
    }
    for (int n = STEPS; n > 0; n--)
        for (int j = 0; j < n; j++) {
            const double cont = (p * opt[base + j + 1] + (1.0 - p) * opt[base + j]) / growth;
            px[base + j] = d * px[base + j + 1];
            opt[base + j] = max(px[base + j] - K, cont);
        }
    value[g] = opt[base];
}
"""

BLOCK_SUM = """
__kernel void block_sum(__global float* out, __global const float* in, __local float* tmp,
                        const uint n)
{
    const uint l = get_local_id(0);
    const uint g = get_global_id(0);
    tmp[l] = g < n ? in[g] : 0.0f;
    barrier(CLK_LOCAL_MEM_FENCE);
    if (l == 0) {
        float s = 0.0f;
        for (uint i = 0; i < get_local_size(0); i++) s += tmp[i];
        out[get_group_id(0)] = s;
    }
}
"""

SCALE_T = """
__kernel void scale_t(__global float* out, __global const float* in, const float s,
                      const uint W, const uint H)
{
    const uint x = get_global_id(0), y = get_global_id(1);
    out[y * W + x] = in[x * H + y] * s;
}
"""

PROBE = """
__kernel void probe(__global uint* out, const uint n)
{
    const uint g = min((uint)get_global_id(0), n - 1u);
    out[4u * g + 0u] = (uint)get_global_id(0);
    out[4u * g + 1u] = (uint)get_local_id(0);
    out[4u * g + 2u] = (uint)get_group_id(0);
    out[4u * g + 3u] = (uint)get_local_size(0) + 1000u * (uint)get_global_size(0);
}
"""
