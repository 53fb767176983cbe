// solver_mode 1: every distance constraint of an iteration in one launch, the corrections summed
// per vertex in 32-bit fixed point (1e-6 m units) with integer atomics, then averaged and
// over-relaxed by a per-vertex launch that also clears the sums for the next iteration.
//
// Reference semantics: crates/cloth_core/src/shaders/distance_atomic.wgsl:78-182
// (solve_distance_atomic, apply_atomic_accum), dispatched per iteration in place of the distance
// colours (crates/cloth_core/src/simulation/dispatch.rs:1436-1446).
//
// Integer sums do not depend on the order the atomics land in, so the result is deterministic
// and equal to a serial sum.  Each constraint record is the 16-byte {v0, v1 | shear << 31, rest,
// tension compliance} of the coloured path (with the compression compliance in `cc` when the two
// differ); the sums are 16 bytes per vertex {dx, dy, dz, count}.
#include "tc_kernels.h"

namespace tc {

template <bool FIRST, bool SPLIT>
__global__ void __launch_bounds__(256) k_distance_jacobi(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                         const uint4 *__restrict__ rec, const float *__restrict__ ccomp,
                                                         float *__restrict__ lam, uint32_t nd, int4 *__restrict__ acc,
                                                         SubstepArgs a) {
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nd) return;
    const uint4 r = rec[k];
    const uint32_t i0 = r.x, i1 = r.y & 0x7FFFFFFFu;
    const float4 q0 = P[i0], q1 = P[i1];
    const float w0 = q0.w, w1 = q1.w, wsum = w0 + w1;
    if (wsum <= 1e-7f) { if (FIRST) lam[k] = 0.0f; return; }
    const f3 delta = xyz(q0) - xyz(q1);
    const float dist = length(delta);
    if (dist <= 1e-7f) { if (FIRST) lam[k] = 0.0f; return; }
    const float c = dist - __uint_as_float(r.z);
    float comp = __uint_as_float(r.w);
    if (SPLIT && !(c >= 0.0f)) comp = ccomp[k];
    if (comp >= 1e9f) { if (FIRST) lam[k] = 0.0f; return; }
    const f3 dir = delta / dist;
    const float dt = fmaxf(a.dt, 1e-8f);
    const float alpha = comp / (dt * dt);
    const bool shear = (r.y >> 31) != 0u;
    const float dc = shear ? a.shear_damping : (c >= 0.0f ? a.tension_damping : a.compression_damping);
    const float gamma = fminf(dc * 1e-5f / dt, 1e4f);
    const f3 rel = (xyz(q0) - xyz(X[i0])) - (xyz(q1) - xyz(X[i1]));
    const float vproj = dot(dir, rel);
    const float l = FIRST ? 0.0f : lam[k];
    const float dl = (-c - alpha * l - gamma * vproj) / ((1.0f + gamma) * wsum + alpha);
    lam[k] = l + dl * 0.25f;
    const float s0 = w0 * dl, s1 = -w1 * dl;
    auto fx = [](float v) { return int(fminf(fmaxf(v * 1000000.0f, -50000000.0f), 50000000.0f)); };
    atomicAdd(&acc[i0].x, fx(dir.x * s0));
    atomicAdd(&acc[i0].y, fx(dir.y * s0));
    atomicAdd(&acc[i0].z, fx(dir.z * s0));
    atomicAdd(&acc[i0].w, 1);
    atomicAdd(&acc[i1].x, fx(dir.x * s1));
    atomicAdd(&acc[i1].y, fx(dir.y * s1));
    atomicAdd(&acc[i1].z, fx(dir.z * s1));
    atomicAdd(&acc[i1].w, 1);
}

__global__ void __launch_bounds__(256) k_jacobi_apply(float4 *__restrict__ P, int4 *__restrict__ acc, uint32_t n,
                                                      SubstepArgs a) {
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int4 s = acc[v];
    const uint32_t count = uint32_t(s.w);
    if (count == 0u) return;
    const float scale = (1.0f / 1000000.0f) / float(count);
    float4 p = P[v];
    p.x += (float(s.x) * scale) * 1.25f;
    p.y += (float(s.y) * scale) * 1.25f;
    p.z += (float(s.z) * scale) * 1.25f;
    P[v] = p;
    acc[v] = make_int4(0, 0, 0, 0);
}

hipError_t launch_distance_jacobi(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, const float *cc,
                                  float *lam, uint32_t nd, int4 *acc, uint32_t nv, const SubstepArgs &a, bool first,
                                  bool split) {
    if (nd) {
        const dim3 g((nd + 255) / 256), b(256);
        if (first) {
            if (split) hipLaunchKernelGGL((k_distance_jacobi<true, true>), g, b, 0, s, P, X, rec, cc, lam, nd, acc, a);
            else hipLaunchKernelGGL((k_distance_jacobi<true, false>), g, b, 0, s, P, X, rec, cc, lam, nd, acc, a);
        } else {
            if (split) hipLaunchKernelGGL((k_distance_jacobi<false, true>), g, b, 0, s, P, X, rec, cc, lam, nd, acc, a);
            else hipLaunchKernelGGL((k_distance_jacobi<false, false>), g, b, 0, s, P, X, rec, cc, lam, nd, acc, a);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!nv) return hipSuccess;
    hipLaunchKernelGGL(k_jacobi_apply, dim3((nv + 255) / 256), dim3(256), 0, s, P, acc, nv, a);
    return hipGetLastError();
}

}  // namespace tc
