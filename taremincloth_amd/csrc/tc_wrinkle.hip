// Wrinkle field (drape guide): an external radial acceleration around a bone, read from an RGBA8
// texture unrolled over the bone's cylinder (u around the axis, repeating; v along it, clamped).
// R: valley strength, G: crest strength, B / A: valley / crest target radius in [r_min, r_max].
// Valleys pull a vertex inwards once it lies beyond the valley radius, crests push it out while it
// lies inside the crest radius, each ramping up over its window; the net acceleration changes the
// velocity and the predicted position of the substep.
//
// Reference semantics (paths relative to the reference repository root):
//   kernel       crates/cloth_core/src/shaders/wrinkle_field.wgsl:58-154
//   parameters   crates/cloth_core/src/simulation/types.rs:110-139, mod.rs:2997-3130,
//                crates/cloth_py/src/lib.rs:3050-3137
//   dispatch     after predict, before the solver iterations (dispatch.rs:151-165)
//
// The reference samples the texture with the hardware's bilinear filter; here the bilinear weights
// are computed in fp32 from the texel-centre convention (x = u W - 0.5), and the angle comes from
// atan2_det, so the oracle reproduces every bit.
#include "tc_sim_impl.h"

#include <cstring>

namespace tc {

__device__ __forceinline__ float4 wrinkle_sample(const uchar4 *__restrict__ tex, uint32_t w, uint32_t h, float u,
                                                 float v) {
    const float fx = u * float(w) - 0.5f, fy = v * float(h) - 0.5f;
    const float x0f = floorf(fx), y0f = floorf(fy);
    const float tx = fx - x0f, ty = fy - y0f;
    const int x0 = int(x0f), y0 = int(y0f);
    const int W = int(w), H = int(h);
    auto wrap = [W](int x) { const int m = x % W; return m < 0 ? m + W : m; };
    auto clampi = [H](int y) { return y < 0 ? 0 : (y > H - 1 ? H - 1 : y); };
    const int xa = wrap(x0), xb = wrap(x0 + 1), ya = clampi(y0), yb = clampi(y0 + 1);
    auto texel = [&](int x, int y) {
        const uchar4 c = tex[size_t(y) * w + x];
        return make_float4(float(c.x) / 255.0f, float(c.y) / 255.0f, float(c.z) / 255.0f, float(c.w) / 255.0f);
    };
    const float4 t00 = texel(xa, ya), t10 = texel(xb, ya), t01 = texel(xa, yb), t11 = texel(xb, yb);
    auto lerp2 = [&](float a, float b, float c, float d) {
        return (a * (1.0f - tx) + b * tx) * (1.0f - ty) + (c * (1.0f - tx) + d * tx) * ty;
    };
    return make_float4(lerp2(t00.x, t10.x, t01.x, t11.x), lerp2(t00.y, t10.y, t01.y, t11.y),
                       lerp2(t00.z, t10.z, t01.z, t11.z), lerp2(t00.w, t10.w, t01.w, t11.w));
}

__global__ void __launch_bounds__(256) k_wrinkle(float4 *__restrict__ P, float4 *__restrict__ V, uint32_t n,
                                                 const uchar4 *__restrict__ tex, uint32_t tw, uint32_t th,
                                                 WrinkleArgs A, float dt_sub) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float EPS = 1e-7f, TWO_PI = 6.2831855f;
    const float4 q = P[i];
    if (q.w <= 0.0f) return;
    const f3 p = xyz(q);
    const f3 axis = mk3(A.axis[0], A.axis[1], A.axis[2]);
    const f3 diff = p - mk3(A.origin[0], A.origin[1], A.origin[2]);
    const float z = dot(diff, axis);
    if (z < A.z_min || z > A.z_max) return;
    const f3 rv = diff - axis * z;
    const float r = length(rv);
    if (r <= 1e-5f) return;
    if (A.influence_radius > 0.0f && A.r_max > 0.0f && r > A.r_max + A.influence_radius) return;
    const f3 nrm = mk3(A.normal[0], A.normal[1], A.normal[2]);
    float theta = atan2_det(dot(rv, mk3(A.binormal[0], A.binormal[1], A.binormal[2])), dot(rv, nrm));
    if (theta < 0.0f) theta += TWO_PI;
    const float uq = theta / TWO_PI;
    const float u = uq - truncf(uq);  // WGSL f32 %: x - y trunc(x / y)
    const float zl = fmaxf(A.z_max - A.z_min, EPS);
    const float v = clampf((z - A.z_min) / zl, 0.0f, 1.0f);
    const float4 t = wrinkle_sample(tex, tw, th, u, v);
    const float rl = fmaxf(A.r_max - A.r_min, EPS);
    const float tv_r = A.r_min + t.z * rl, tc_r = A.r_min + t.w * rl;
    const float dt = fmaxf(dt_sub, 1e-8f);
    float a_pull = 0.0f, a_push = 0.0f;
    if (t.x > 0.0f) {
        const float vr = fmaxf(tv_r, A.bone_radius + 0.002f);
        const float margin = fmaxf(0.0f, r - vr);
        a_pull = t.x * A.blend * A.stiffness * clampf(margin / fmaxf(A.valley_window, 0.001f), 0.0f, 1.0f);
    }
    if (t.y > 0.0f) {
        const float ex = fmaxf(0.0f, tc_r - r);
        a_push = t.y * A.blend * A.stiffness * clampf(ex / fmaxf(A.crest_window, 0.001f), 0.0f, 1.0f);
    }
    const float a = a_push - a_pull;
    if (!(fabsf(a) > EPS)) return;
    const f3 dir = r > EPS ? rv / r : nrm;
    const f3 dv = dir * (a * dt);
    const float4 vv = V[i];
    V[i] = make_float4(vv.x + dv.x, vv.y + dv.y, vv.z + dv.z, vv.w);
    const f3 dp = dv * dt;
    P[i] = make_float4(q.x + dp.x, q.y + dp.y, q.z + dp.z, q.w);
}

}  // namespace tc

hipError_t enqueue_wrinkle(tc_sim *s, float dt_sub) {
    if (!s->wrinkle_on || !s->wrinkle_tex.p || !s->mesh.nv) return hipSuccess;
    const uint32_t n = s->mesh.nv;
    tc::WrinkleArgs w = s->wrinkle;
    w.blend *= s->priority_scale();  // dispatch.rs:53-60: the sewing priority scales the blend
    hipLaunchKernelGGL(tc::k_wrinkle, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->P.p, s->V.p, n,
                       s->wrinkle_tex.p, s->wrinkle_w, s->wrinkle_h, w, dt_sub);
    return hipGetLastError();
}

extern "C" {

int tc_sim_set_enable_wrinkle_field(tc_sim *s, int enable) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    s->wrinkle_on = enable != 0;
    s->touch();
    return TC_OK;
}

int tc_sim_wrinkle_field_enabled(const tc_sim *s, int *enabled) {
    if (!s || !enabled) return fail(TC_ERR_VALUE, "null argument");
    *enabled = s->wrinkle_on ? 1 : 0;
    return TC_OK;
}

int tc_sim_set_wrinkle_field_params(tc_sim *s, const float *frame, float z_min, float z_max, float r_min, float r_max,
                                    float valley_window, float crest_window) {
    if (!s || !frame) return fail(TC_ERR_VALUE, "null argument");
    tc::WrinkleArgs &w = s->wrinkle;
    std::memcpy(w.origin, frame, 12); w.influence_radius = frame[3];
    std::memcpy(w.axis, frame + 4, 12); w.bone_radius = frame[7];
    std::memcpy(w.normal, frame + 8, 12); w.stiffness = frame[11];
    std::memcpy(w.binormal, frame + 12, 12); w.blend = frame[15];
    w.z_min = z_min; w.z_max = z_max; w.r_min = r_min; w.r_max = r_max;
    w.valley_window = valley_window; w.crest_window = crest_window;
    s->touch();
    return TC_OK;
}

int tc_sim_get_wrinkle_field_params(const tc_sim *s, float *out) {
    if (!s || !out) return fail(TC_ERR_VALUE, "null argument");
    const tc::WrinkleArgs &w = s->wrinkle;
    const float v[22] = {w.origin[0], w.origin[1], w.origin[2], w.influence_radius, w.axis[0], w.axis[1], w.axis[2],
                         w.bone_radius, w.normal[0], w.normal[1], w.normal[2], w.stiffness, w.binormal[0],
                         w.binormal[1], w.binormal[2], w.blend, w.z_min, w.z_max, w.r_min, w.r_max, w.valley_window,
                         w.crest_window};
    std::memcpy(out, v, sizeof(v));
    return TC_OK;
}

int tc_sim_set_wrinkle_field_texture_2d(tc_sim *s, uint32_t width, uint32_t height, const uint8_t *rgba, size_t nbytes,
                                        float z_min, float z_max, float r_min, float r_max) {
    if (!s) return fail(TC_ERR_VALUE, "null simulator");
    if (width == 0 || height == 0 || nbytes == 0 || !rgba) {  // mod.rs:3033-3039: the texture goes
        s->wrinkle_tex.release();
        s->wrinkle_w = s->wrinkle_h = 0;
        s->touch();
        return TC_OK;
    }
    if (uint64_t(width) * height * 4u != nbytes)
        return fail(TC_ERR_VALUE, "texture_bytes must hold width * height * 4 bytes");
    TC_HIP(hipStreamSynchronize(s->stream));
    TC_HIP(s->wrinkle_tex.upload(reinterpret_cast<const uchar4 *>(rgba), size_t(width) * height, s->stream));
    TC_HIP(hipStreamSynchronize(s->stream));
    s->wrinkle_w = width;
    s->wrinkle_h = height;
    s->wrinkle.z_min = z_min; s->wrinkle.z_max = z_max; s->wrinkle.r_min = r_min; s->wrinkle.r_max = r_max;
    s->touch();
    return TC_OK;
}

}  // extern "C"
