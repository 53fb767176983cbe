// Core XPBD substep kernels for gfx950: predict, per-colour distance and
// bending projection, collider resolve, velocity update, readback packing and
// the constraint-residual reduction.
//
// Reference semantics (paths relative to the reference repository root):
//   predict      crates/cloth_core/src/shaders/predict.wgsl:33-55
//   distance     crates/cloth_core/src/shaders/distance.wgsl:60-124
//   bending      crates/cloth_core/src/shaders/bending.wgsl:59-126
//   collision    crates/cloth_core/src/shaders/collision.wgsl:83-108, 273-652
//   update_vel   crates/cloth_core/src/shaders/update_vel.wgsl:33-53
//   extract      crates/cloth_core/src/shaders/extract_positions.wgsl:14-26
//
// The projection kernels are HBM/gather bound: one thread per constraint of a
// colour, a 16-byte constraint record {v0, v1 | type<<31, rest, compliance}
// and two 16-byte gathers of P.  No two constraints of one colour share a
// vertex, so the scatter back to P needs no atomics.
#include "tc_kernels.h"
#include <algorithm>

namespace tc {

__global__ void __launch_bounds__(256) k_predict(float4 *__restrict__ P, const float4 *__restrict__ X,
                                                 float4 *__restrict__ V, uint32_t n, SubstepArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = fmaxf(a.dt, 1e-8f);
    const float4 x = X[i];
    const float w = P[i].w;
    float4 v = V[i];
    f3 p;
    if (w <= 0.0f) {
        v.x = 0.0f; v.y = 0.0f; v.z = 0.0f;
        p = xyz(x);
    } else {
        const float keep = fmaxf(0.0f, 1.0f - a.damping * dt);
        v.x = v.x * keep + a.gx * dt;
        v.y = v.y * keep + a.gy * dt;
        v.z = v.z * keep + a.gz * dt;
        p = mk3(x.x + v.x * dt, x.y + v.y * dt, x.z + v.z * dt);
    }
    V[i] = v;
    P[i] = make_float4(p.x, p.y, p.z, w);
}

template <bool FIRST, bool DAMP, bool SPLIT>
__global__ void __launch_bounds__(256) k_distance(float4 *__restrict__ P, const float4 *__restrict__ X,
                                                  const uint4 *__restrict__ rec, const float *__restrict__ ccomp,
                                                  float *__restrict__ lam, uint32_t off, uint32_t cnt,
                                                  SubstepArgs a) {
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const uint32_t k = off + t;
    const uint4 r = rec[k];
    const uint32_t i0 = r.x, i1 = r.y & 0x7FFFFFFFu;
    const float4 q0 = P[i0], q1 = P[i1];
    const float w0 = q0.w, w1 = q1.w, wsum = w0 + w1;
    // lambda is zero at the substep start; the first sweep stores it even when it skips
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
    const float l = FIRST ? 0.0f : lam[k];
    float dl;
    if (DAMP) {
        const bool shear = (r.y >> 31) != 0u;
        const float dc = shear ? a.shear_damping : (c >= 0.0f ? a.tension_damping : a.compression_damping);
        const float gamma = fminf(dc * 1e-5f / dt, 1e4f);
        const f3 rel = (xyz(q0) - xyz(X[i0])) - (xyz(q1) - xyz(X[i1]));
        const float vproj = dot(dir, rel);
        dl = (-c - alpha * l - gamma * vproj) / ((1.0f + gamma) * wsum + alpha);
    } else {
        dl = (-c - alpha * l) / (wsum + alpha);
    }
    lam[k] = l + dl;
    const f3 c0 = dir * (w0 * dl);
    const f3 c1 = dir * (-w1 * dl);
    P[i0] = make_float4(q0.x + c0.x, q0.y + c0.y, q0.z + c0.z, w0);
    P[i1] = make_float4(q1.x + c1.x, q1.y + c1.y, q1.z + c1.z, w1);
}

// gate (meshes with sewing seams, else null): {v0, v1, gate distance bits, seam flag} per constraint;
// a seam constraint (bending.wgsl:92-104) pulls only while its opposite vertices are closer than the
// unfolded rest length and its first sewn pair is within the gate distance.
template <bool FIRST, bool DAMP>
__global__ void __launch_bounds__(256) k_bending(float4 *__restrict__ P, const float4 *__restrict__ X,
                                                 const uint4 *__restrict__ rec, float *__restrict__ lam,
                                                 uint32_t off, uint32_t cnt, SubstepArgs a,
                                                 const uint4 *__restrict__ gate) {
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const uint32_t k = off + t;
    const uint4 r = rec[k];
    const float comp = __uint_as_float(r.w);
    if (comp >= 1e9f) { if (FIRST) lam[k] = 0.0f; return; }
    const uint32_t i2 = r.x, i3 = r.y;
    const float4 q2 = P[i2], q3 = P[i3];
    const float w2 = q2.w, w3 = q3.w, wsum = w2 + w3;
    if (wsum <= 1e-7f) { if (FIRST) lam[k] = 0.0f; return; }
    const f3 delta = xyz(q2) - xyz(q3);
    const float dist = length(delta);
    if (dist <= 1e-7f) { if (FIRST) lam[k] = 0.0f; return; }
    if (gate) {
        const uint4 g = gate[k];
        if (g.w != 0u) {
            if (dist >= __uint_as_float(r.z)) { if (FIRST) lam[k] = 0.0f; return; }
            const float gap = length(xyz(P[g.x]) - xyz(P[g.y]));
            if (gap > __uint_as_float(g.z)) { if (FIRST) lam[k] = 0.0f; return; }
        }
    }
    const f3 dir = delta / dist;
    const float dt = fmaxf(a.dt, 1e-8f);
    const float alpha = comp / (dt * dt);
    const float c = dist - __uint_as_float(r.z);
    const float l = FIRST ? 0.0f : lam[k];
    float dl;
    if (DAMP) {
        const float gamma = fminf(a.bending_damping * 1e-5f / dt, 1e4f);
        const f3 rel = (xyz(q2) - xyz(X[i2])) - (xyz(q3) - xyz(X[i3]));
        const float vproj = dot(dir, rel);
        dl = (-c - alpha * l - gamma * vproj) / ((1.0f + gamma) * wsum + alpha);
    } else {
        dl = (-c - alpha * l) / (wsum + alpha);
    }
    lam[k] = l + dl;
    const f3 c2 = dir * (w2 * dl);
    const f3 c3 = dir * (-w3 * dl);
    P[i2] = make_float4(q2.x + c2.x, q2.y + c2.y, q2.z + c2.z, w2);
    P[i3] = make_float4(q3.x + c3.x, q3.y + c3.y, q3.z + c3.z, w3);
}

// collision.wgsl:83-108
__device__ __forceinline__ void contact_response(f3 target, f3 n, float fr, float rest, float dt, f3 p_init, f3 &xn,
                                                 f3 &p) {
    const f3 total = target - xn;
    const f3 tang = total - n * dot(total, n);
    p = target - tang * clampf(fr, 0.0f, 1.0f);
    const float edt = fmaxf(dt, 1e-6f);
    const float vin = dot((p_init - xn) / edt, n);
    float vout = 0.0f;
    if (vin < 0.0f) vout = -clampf(rest, 0.0f, 1.0f) * vin;
    const float dxn = dot(target - xn, n) - vout * edt;
    xn = xn + n * dxn;
}

// collision.wgsl:273-319 (closest point and face flag)
__device__ __forceinline__ f3 closest_on_tri(f3 p, f3 a, f3 b, f3 c, bool &is_face) {
    is_face = false;
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
    const f3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
    const float denom = 1.0f / (va + vb + vc);
    is_face = true;
    return a + ab * (vb * denom) + ac * (vc * denom);
}

// collision.wgsl:322-361 (Moller-Trumbore with 5% edge tolerance)
__device__ __forceinline__ bool seg_tri(f3 po, f3 pn, f3 a, f3 b, f3 c, float &t_out) {
    const f3 d = pn - po;
    const float ld = dot(d, d);
    if (ld < 1e-12f) return false;
    const f3 e1 = b - a, e2 = c - a;
    const f3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    const float tn = dot(e1, e1) * dot(e2, e2);
    if (det * det <= tn * ld * 1e-10f || tn <= 1e-20f) return false;
    const float inv = 1.0f / det;
    const f3 tv = po - a;
    const float u = dot(tv, pv) * inv;
    if (u < -0.05f || u > 1.05f) return false;
    const f3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < -0.05f || (u + v) > 1.05f) return false;
    const float t = dot(e2, qv) * inv;
    if (t >= 0.0f && t <= 1.0f) { t_out = t; return true; }
    return false;
}

__device__ __forceinline__ float seg_sphere_dist_sq(f3 a, f3 b, f3 c) {  // collision.wgsl:364-374
    const f3 ab = b - a;
    const float ab2 = dot(ab, ab);
    if (ab2 <= 1e-12f) { const f3 d = a - c; return dot(d, d); }
    const float t = clampf(dot(c - a, ab) / ab2, 0.0f, 1.0f);
    const f3 d = (a + ab * t) - c;
    return dot(d, d);
}

// collision.wgsl:377-652, decoupled (once per substep), direct triangle scan
// in the Morton order the host sorted the triangles into.
// Trilinear sample of the (distance, weight) volume at normalised coordinates, clamp-to-edge
// addressing with texel centres at (i + 0.5) / dim -- the filtering a linear sampler applies.
__device__ __forceinline__ float2 sdf_sample(const SdfArgs &s, float u, float v, float w) {
    const float x = u * float(s.w) - 0.5f, y = v * float(s.h) - 0.5f, z = w * float(s.d) - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const float tx = x - fx0, ty = y - fy0, tz = z - fz0;
    const int32_t xi = f2i_sat(fx0), yi = f2i_sat(fy0), zi = f2i_sat(fz0);
    const int32_t x0 = min(max(xi, 0), int32_t(s.w) - 1), x1 = min(max(xi + 1, 0), int32_t(s.w) - 1);
    const int32_t y0 = min(max(yi, 0), int32_t(s.h) - 1), y1 = min(max(yi + 1, 0), int32_t(s.h) - 1);
    const int32_t z0 = min(max(zi, 0), int32_t(s.d) - 1), z1 = min(max(zi + 1, 0), int32_t(s.d) - 1);
    auto half_bits = [](uint32_t b) {
        const uint16_t u = uint16_t(b);
        _Float16 h;
        __builtin_memcpy(&h, &u, 2);
        return (float)h;
    };
    auto texel = [&](int32_t i, int32_t j, int32_t k) {
        const uint32_t t = s.tex[(size_t(k) * s.h + uint32_t(j)) * s.w + uint32_t(i)];
        return make_float2(half_bits(t & 0xFFFFu), half_bits(t >> 16));
    };
    auto lerp2 = [](float2 a, float2 b, float t) { return make_float2(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t); };
    const float2 c00 = lerp2(texel(x0, y0, z0), texel(x1, y0, z0), tx);
    const float2 c10 = lerp2(texel(x0, y1, z0), texel(x1, y1, z0), tx);
    const float2 c01 = lerp2(texel(x0, y0, z1), texel(x1, y0, z1), tx);
    const float2 c11 = lerp2(texel(x0, y1, z1), texel(x1, y1, z1), tx);
    return lerp2(lerp2(c00, c10, ty), lerp2(c01, c11, ty), tz);
}

// matrix (given as m[row][column]) applied the way the reference's shader applies the uploaded
// array: result[k] = sum_j m[j][k] * v[j]
__device__ __forceinline__ f3 sdf_xform(const float *m, f3 v, float w) {
    return mk3(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
               m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w);
}

// solve_bone_sdf_collision (collision.wgsl:116-265)
__device__ void sdf_collision(const SdfArgs &s, uint32_t ntris_mesh, f3 &p, f3 &xn, float dt, f3 p_initial, float vth) {
    if (s.nb == 0) return;
    float best_dist = 1e6f, best_fr = 0.0f, best_rest = 0.0f, best_target = 0.0f;
    f3 best_n = mk3(0.0f, 0.0f, 0.0f);
    uint32_t hits = 0;
    for (uint32_t b = 0; b < s.nb; ++b) {
        const float *I = s.info + 20 * size_t(b);
        const float radius = I[15];
        if (radius > 0.0f) {
            const f3 dw = p - mk3(I[3], I[7], I[11]);
            const float rt = radius + vth;
            if (dot(dw, dw) > rt * rt) continue;
        }
        const float *wm = s.xf + 32 * size_t(b), *im = wm + 16;
        const f3 pl = sdf_xform(im, p, 1.0f);
        const float target = vth + I[17];
        const float margin = target + 0.05f;
        if (pl.x < I[0] - margin || pl.x > I[4] + margin || pl.y < I[1] - margin || pl.y > I[5] + margin ||
            pl.z < I[2] - margin || pl.z > I[6] + margin)
            continue;
        const f3 sc = mk3(I[8], I[9], I[10]), of = mk3(I[12], I[13], I[14]);
        const f3 uvw = mk3(pl.x * sc.x + of.x, pl.y * sc.y + of.y, pl.z * sc.z + of.z);
        const f3 umin = mk3(I[0] * sc.x + of.x, I[1] * sc.y + of.y, I[2] * sc.z + of.z);
        const f3 umax = mk3(I[4] * sc.x + of.x, I[5] * sc.y + of.y, I[6] * sc.z + of.z);
        if (uvw.x < umin.x || uvw.x > umax.x || uvw.y < umin.y || uvw.y > umax.y || uvw.z < umin.z || uvw.z > umax.z)
            continue;
        const f3 eps_slice = (umax - umin) * (s.nb == 1u ? 0.002f : 0.015f);
        const f3 lo = umin + eps_slice, hi = umax - eps_slice;
        const f3 us = mk3(clampf(uvw.x, lo.x, hi.x), clampf(uvw.y, lo.y, hi.y), clampf(uvw.z, lo.z, hi.z));
        const float2 sv = sdf_sample(s, us.x, us.y, us.z);
        const float raw = sv.x, alpha = sv.y;
        const float threshold = fmaxf(I[19], 0.01f);
        float retract = 0.0f;
        if (ntris_mesh > 0u) {
            const float t = clampf((alpha - threshold) / (0.7f - threshold), 0.0f, 1.0f);
            retract = (1.0f - t * t * (3.0f - 2.0f * t)) * 0.06f;
        }
        const float eff = alpha > threshold ? raw + retract : 10.0f;
        if (eff < target) {
            const f3 eu = sc * 0.005f;
            const float dxp = sdf_sample(s, clampf(uvw.x + eu.x, lo.x, hi.x), us.y, us.z).x;
            const float dxn = sdf_sample(s, clampf(uvw.x - eu.x, lo.x, hi.x), us.y, us.z).x;
            const float dyp = sdf_sample(s, us.x, clampf(uvw.y + eu.y, lo.y, hi.y), us.z).x;
            const float dyn = sdf_sample(s, us.x, clampf(uvw.y - eu.y, lo.y, hi.y), us.z).x;
            const float dzp = sdf_sample(s, us.x, us.y, clampf(uvw.z + eu.z, lo.z, hi.z)).x;
            const float dzn = sdf_sample(s, us.x, us.y, clampf(uvw.z - eu.z, lo.z, hi.z)).x;
            float dx = dxp - dxn, dy = dyp - dyn, dz = dzp - dzn;
            if (fabsf(dx) < 1e-6f) dx = uvw.x < (lo.x + hi.x) * 0.5f ? dxp - raw : raw - dxn;
            if (fabsf(dy) < 1e-6f) dy = uvw.y < (lo.y + hi.y) * 0.5f ? dyp - raw : raw - dyn;
            if (fabsf(dz) < 1e-6f) dz = uvw.z < (lo.z + hi.z) * 0.5f ? dzp - raw : raw - dzn;
            const f3 g = mk3(dx, dy, dz);
            const float gl = length(g);
            f3 nl = mk3(0.0f, 0.0f, 1.0f);
            if (gl > 1e-6f) {
                nl = g / gl;
            } else {
                const f3 dc = pl - mk3((I[0] + I[4]) * 0.5f, (I[1] + I[5]) * 0.5f, (I[2] + I[6]) * 0.5f);
                const float lc = length(dc);
                if (lc > 1e-6f) nl = dc / lc;
            }
            const f3 nw0 = sdf_xform(wm, nl, 0.0f);
            const f3 nw = nw0 / length(nw0);
            if (hits == 0u || eff < best_dist) {
                best_dist = eff;
                best_n = nw;
                best_fr = I[16];
                best_rest = I[18];
                best_target = target;
            }
            ++hits;
        }
    }
    if (hits > 0u && best_dist < best_target) {
        const float pen = best_target - best_dist;
        contact_response(p + best_n * pen, best_n, best_fr, best_rest, dt, p_initial, xn, p);
    }
}

// Whether the segment a->b may pass within `th` of the box [lo, hi]: the gap between the
// segment's box and [lo, hi], compared with a slack of 1e-5 relative + 1e-5 m so that no
// triangle whose own bounding-sphere test (below) would pass is skipped.
__device__ __forceinline__ bool seg_box_near(f3 a, f3 b, float4 lo, float4 hi, float th) {
    const f3 smin = min3(a, b), smax = max3(a, b);
    const float gx = fmaxf(0.0f, fmaxf(lo.x - smax.x, smin.x - hi.x));
    const float gy = fmaxf(0.0f, fmaxf(lo.y - smax.y, smin.y - hi.y));
    const float gz = fmaxf(0.0f, fmaxf(lo.z - smax.z, smin.z - hi.z));
    const float lim = th * 1.00001f + 1e-5f;
    return !(gx * gx + gy * gy + gz * gz > lim * lim);
}

__global__ void __launch_bounds__(256) k_collision(float4 *__restrict__ P, float4 *__restrict__ X, uint32_t n,
                                                   const uint32_t *__restrict__ order,
                                                   const Collider *__restrict__ cols, uint32_t ncols,
                                                   const MeshTri *__restrict__ tris,
                                                   const float4 *__restrict__ bounds, uint32_t ntris,
                                                   const float4 *__restrict__ cbox, int single_sided_recovery,
                                                   float dt, SdfArgs sdf, uint32_t vpw) {
    // A wavefront takes vpw vertices (lanes 0..vpw-1); slot k handles vertex order[k], a spatially
    // compact order, so that the lanes of a wavefront need the same triangle clusters.  Every lane
    // of a wavefront stays to the end: lanes without a vertex (or with a pinned one) still take
    // part in staging and in the cluster votes.
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint32_t slot = (blockIdx.x * (blockDim.x >> 6) + wv) * vpw + lane;
    const bool has = lane < vpw && slot < n;
    const uint32_t i = has ? (order ? order[slot] : slot) : 0u;
    const float4 q = has ? P[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool active = has && q.w > 0.0f;
    const float4 x = active ? X[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    f3 p = xyz(q);
    const f3 p_init = p, x_old = xyz(x);
    f3 xn = x_old;
    const float th = x.w;
    for (uint32_t c = 0; active && c < ncols; ++c) {
        const Collider col = cols[c];
        const f3 A = mk3(col.ax, col.ay, col.az);
        if (col.type == 0u) {
            const float tr = col.radius + th;
            const f3 d = p - A;
            const float dist = length(d);
            if (dist < tr) {
                f3 nrm = mk3(0.0f, 0.0f, 1.0f);
                if (dist > 1e-7f) nrm = d / dist;
                contact_response(A + nrm * tr, nrm, col.friction, col.restitution, dt, p_init, xn, p);
            }
        } else if (col.type == 1u) {
            const f3 B = mk3(col.bx, col.by, col.bz);
            const float tr = col.radius + th;
            const f3 ab = B - A;
            const float l2 = dot(ab, ab);
            float t = 0.0f;
            if (l2 > 1e-7f) t = clampf(dot(p - A, ab) / l2, 0.0f, 1.0f);
            const f3 cl = A + ab * t;
            const f3 d = p - cl;
            const float dist = length(d);
            if (dist < tr) {
                f3 nrm = mk3(0.0f, 0.0f, 1.0f);
                if (dist > 1e-7f) nrm = d / dist;
                contact_response(cl + nrm * tr, nrm, col.friction, col.restitution, dt, p_init, xn, p);
            }
        } else if (col.type == 2u) {
            const f3 B = mk3(col.bx, col.by, col.bz);
            const f3 nrm = B / length(B);
            const float dist = dot(p - A, nrm);
            if (dist < th) contact_response(p + nrm * (th - dist), nrm, col.friction, col.restitution, dt, p_init, xn, p);
        }
    }
    // bit 0: a front contact was applied; bit 1: a recovery candidate exists (one word: as two
    // separate flags the compiler merged their stores through a pointer select into scratch)
    uint32_t contact_bits = 0u;
    float best_depth = 999999.0f;
    f3 best_pos = p, best_n = mk3(0.0f, 0.0f, 0.0f);
    float best_fr = 0.0f, best_rest = 0.0f;
    // collision.wgsl:463-542 for one triangle, in the scan's order
    auto triangle = [&](const MeshTri &T) {
        const f3 t0 = mk3(T.p0x, T.p0y, T.p0z), t1 = mk3(T.p1x, T.p1y, T.p1z), t2 = mk3(T.p2x, T.p2y, T.p2z);
        const float target = th + T.thickness;
        const f3 cr = cross(t1 - t0, t2 - t0);
        const float cl = length(cr);
        if (cl < 1e-7f) return;
        const f3 fn = cr / cl;
        const bool single = (T.flags & 1u) != 0u;
        float th_hit;
        if (seg_tri(x_old, p, t0, t1, t2, th_hit)) {
            const f3 hit = x_old + (p - x_old) * th_hit;
            f3 nrm = fn;
            if (dot(p - x_old, fn) > 0.0f && !single) nrm = -fn;
            contact_response(hit + nrm * target, nrm, T.friction, T.restitution, dt, p_init, xn, p);
            contact_bits |= 1u;
        }
        bool is_face;
        const f3 qp = closest_on_tri(p, t0, t1, t2, is_face);
        const f3 d = p - qp;
        const float dist = length(d);
        const float sd = dot(d, fn);
        if (single) {
            if (sd >= 0.0f) {
                if (dist < target) {
                    f3 nrm = fn;
                    if (dist > 1e-7f) nrm = d / dist;
                    contact_response(p + nrm * (target - dist), nrm, T.friction, T.restitution, dt, p_init, xn, p);
                    contact_bits |= 1u;
                }
            } else if (single_sided_recovery && (T.flags & 2u) == 0u) {
                if (is_face && dist < target * 3.0f && dist <= fabsf(sd) * 1.1f + 1e-7f && dist < best_depth) {
                    best_depth = dist;
                    best_n = fn;
                    best_pos = p + fn * (target - sd);
                    best_fr = T.friction;
                    best_rest = T.restitution;
                    contact_bits |= 2u;
                }
            }
        } else if (dist < target && sd >= -target) {
            f3 nrm = fn;
            if (dist > 1e-7f) nrm = d / dist;
            contact_response(p + nrm * (target - dist), nrm, T.friction, T.restitution, dt, p_init, xn, p);
            contact_bits |= 1u;
        }
    };
    // Triangles in Morton order, in clusters of 32 and super-clusters of 32 clusters with boxes
    // around their bounding spheres scaled by 1.5 (cbox: super-cluster boxes, then cluster
    // boxes, lo/hi pairs).  A (super-)cluster is skipped when no lane's current segment comes
    // near its box.  The boxes of 32 super-clusters, the boxes of one super-cluster's clusters,
    // and a cluster's bounding spheres and triangles are staged in LDS by the whole wavefront.
    // Inside a cluster each lane marks the triangles whose bounding-sphere test it passes and
    // works through its own marks in order, all lanes in step, so that the wavefront pays for
    // the longest list instead of every triangle that any lane needs.  A contact that moves p
    // renews the lane's marks after that triangle.  Each vertex therefore meets its triangles in
    // the scan's order with the p the scan would have, and every contact is applied as in a scan
    // of all triangles.
    __shared__ float4 s_sbox[4][64];
    __shared__ float4 s_cbox[4][64];
    __shared__ float4 s_bound[4][32];
    __shared__ MeshTri s_tri[4][32];
    const uint32_t ncl = (ntris + 31u) / 32u, nsc = (ncl + 31u) / 32u;
    const float4 *sbox = cbox, *clbox = cbox + 2 * size_t(nsc);
    auto marks = [&](uint32_t from, uint32_t cnt) -> uint32_t {
        uint32_t m = 0;
        for (uint32_t kk = from; kk < cnt; ++kk) {
            const float4 b = s_bound[wv][kk];
            const float rr = b.w * 1.5f + th;
            if (!(seg_sphere_dist_sq(x_old, p, xyz(b)) > rr * rr)) m |= 1u << kk;
        }
        return m;
    };
    for (uint32_t s0 = 0; s0 < nsc; s0 += 32u) {
        const uint32_t s1 = min(nsc, s0 + 32u);
        __builtin_amdgcn_wave_barrier();  // the previous chunk's reads come first
        if (lane < 2u * (s1 - s0)) s_sbox[wv][lane] = sbox[2 * size_t(s0) + lane];
        __builtin_amdgcn_wave_barrier();
        for (uint32_t sc = s0; sc < s1; ++sc) {
            if (!__any(active && seg_box_near(x_old, p, s_sbox[wv][2u * (sc - s0)], s_sbox[wv][2u * (sc - s0) + 1u],
                                              th)))
                continue;
            const uint32_t c0 = sc * 32u, c1 = min(ncl, c0 + 32u);
            __builtin_amdgcn_wave_barrier();
            if (lane < 2u * (c1 - c0)) s_cbox[wv][lane] = clbox[2 * size_t(c0) + lane];
            __builtin_amdgcn_wave_barrier();
            for (uint32_t ci = c0; ci < c1; ++ci) {
                if (!__any(active && seg_box_near(x_old, p, s_cbox[wv][2u * (ci - c0)], s_cbox[wv][2u * (ci - c0) + 1u],
                                                  th)))
                    continue;
                const uint32_t t0i = ci * 32u, cnt = min(ntris - t0i, 32u);
                __builtin_amdgcn_wave_barrier();  // the previous cluster's reads come first
                if (lane < cnt) s_bound[wv][lane] = bounds[t0i + lane];
                {
                    const float4 *src = reinterpret_cast<const float4 *>(tris + t0i);
                    float4 *dst = reinterpret_cast<float4 *>(&s_tri[wv][0]);
                    if (lane < cnt * 4u) dst[lane] = src[lane];
                    if (lane + 64u < cnt * 4u) dst[lane + 64u] = src[lane + 64u];
                }
                __builtin_amdgcn_wave_barrier();
                uint32_t m = active ? marks(0u, cnt) : 0u;
                while (__any(m != 0u)) {
                    if (m != 0u) {
                        const uint32_t kk = uint32_t(__ffs(m) - 1);
                        m &= m - 1u;
                        const f3 pb = p;
                        triangle(s_tri[wv][kk]);
                        if (__float_as_uint(p.x) != __float_as_uint(pb.x) ||
                            __float_as_uint(p.y) != __float_as_uint(pb.y) ||
                            __float_as_uint(p.z) != __float_as_uint(pb.z))
                            m = marks(kk + 1u, cnt);
                    }
                }
            }
        }
    }
    if (!active) return;
    if (contact_bits == 2u) contact_response(best_pos, best_n, best_fr, best_rest, dt, p_init, xn, p);
    // SDF colliders last, with the current p as the reference for the velocity (collision.wgsl:645-647)
    sdf_collision(sdf, ntris, p, xn, dt, p, th);
    P[i] = make_float4(p.x, p.y, p.z, q.w);
    X[i] = make_float4(xn.x, xn.y, xn.z, x.w);
}

__global__ void __launch_bounds__(256) k_update_vel(const float4 *__restrict__ P, float4 *__restrict__ X,
                                                    float4 *__restrict__ V, uint32_t n, float dt_sub) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = fmaxf(dt_sub, 1e-8f);
    const float4 p = P[i];
    const float4 x = X[i];
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p.w > 0.0f) {
        v.x = (p.x - x.x) / dt;
        v.y = (p.y - x.y) / dt;
        v.z = (p.z - x.z) / dt;
    }
    V[i] = v;
    X[i] = make_float4(p.x, p.y, p.z, x.w);
}

__global__ void k_extract(const float4 *__restrict__ X, float *__restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = X[i];
    out[3 * i] = x.x;
    out[3 * i + 1] = x.y;
    out[3 * i + 2] = x.z;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t value, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// Residual h_j = C_j + alpha~_j lambda_j + gamma_j grad C_j . (p - x) over distance and
// bending constraints, accumulated in double: the numerator of the damped projection's
// update (distance.wgsl:60-124, bending.wgsl:59-126), which the iterations drive to zero.
// gamma = min(damping * 1e-5 / dt, 1e4) is zero without constraint damping, and then X is not
// read.  Pass 1: each block strides over the constraints and writes (sum h^2, sum L0^2,
// max |h|/L0) to its slot of `part`; pass 2 (one block) sums the slots in block order, so the
// result does not depend on scheduling.
constexpr uint32_t RES_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}

template <bool DAMP>
__global__ void __launch_bounds__(256) k_residual_part(const float4 *__restrict__ P, const float4 *__restrict__ X,
                                                       const uint4 *__restrict__ drec,
                                                       const float *__restrict__ dcc, const float *__restrict__ dlam,
                                                       uint32_t nd, const uint4 *__restrict__ brec,
                                                       const float *__restrict__ blam, uint32_t nb, SubstepArgs sa,
                                                       int split, double *__restrict__ part,
                                                       const uint32_t *__restrict__ run_flag) {
    const float dt_sub = sa.dt;
    if (run_flag && *run_flag == 0u) return;  // the previous check passed: nothing to measure
    __shared__ double sh[3][4];
    double hh = 0.0, ll = 0.0, mx = 0.0;
    const double dt = dt_sub > 1e-8f ? (double)dt_sub : 1e-8;
    const double inv_dt2 = 1.0 / (dt * dt);
    const uint32_t n = nd + nb;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const bool isd = t < nd;
        const uint32_t k = isd ? t : t - nd;
        const uint4 r = isd ? drec[k] : brec[k];
        const uint32_t a = r.x, b = isd ? (r.y & 0x7FFFFFFFu) : r.y;
        const float4 qa = P[a], qb = P[b];
        const double L = __uint_as_float(r.z);
        const double ws = (double)qa.w + (double)qb.w;
        const double dx = (double)qa.x - qb.x, dy = (double)qa.y - qb.y, dz = (double)qa.z - qb.z;
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        double comp = __uint_as_float(r.w);
        if (isd && split && d - L < 0.0) comp = dcc[k];
        if (ws > 1e-7 && d > 1e-7 && comp < 1e9) {
            double h = (d - L) + comp * inv_dt2 * (double)(isd ? dlam[k] : blam[k]);
            if (DAMP) {
                const bool shear = isd && (r.y >> 31) != 0u;
                const float dc = !isd ? sa.bending_damping
                                      : (shear ? sa.shear_damping
                                               : (d - L >= 0.0 ? sa.tension_damping : sa.compression_damping));
                if (dc > 0.0f) {
                    const double gamma = fmin((double)dc * (double)1e-5f / dt, 1e4);
                    const float4 xa = X[a], xb = X[b];
                    const double rx = ((double)qa.x - xa.x) - ((double)qb.x - xb.x);
                    const double ry = ((double)qa.y - xa.y) - ((double)qb.y - xb.y);
                    const double rz = ((double)qa.z - xa.z) - ((double)qb.z - xb.z);
                    h += gamma * (dx * rx + dy * ry + dz * rz) / d;
                }
            }
            hh += h * h;
            ll += L * L;
            if (L > 0.0) mx = fmax(mx, fabs(h) / L);
        }
    }
    hh = wave_sum(hh);
    ll = wave_sum(ll);
    mx = wave_max(mx);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = hh; sh[1][w] = ll; sh[2][w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
        part[3 * blockIdx.x + 1] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
        part[3 * blockIdx.x + 2] = fmax(fmax(sh[2][0], sh[2][1]), fmax(sh[2][2], sh[2][3]));
    }
}

// out[0] = sum h^2, out[1] = sum L0^2, out[2] = max |h|/L0, out[3] = relative residual.
// With gate != nullptr: gate[0] = 1 if the relative residual exceeds `threshold` (else 0); the
// first check of a substep adds gate[1] += 1 (substeps checked), gate[2] += gate[0] (substeps that
// fell back) and gate[3] += gate[0] (fallback chunks); a chained check adds gate[3] += gate[0]; the
// final check after the budget's last chunk adds gate[4] += gate[0] (substeps that end above it).
__global__ void __launch_bounds__(64) k_residual_final(const double *__restrict__ part, uint32_t nblocks,
                                                       double *__restrict__ out, uint32_t *__restrict__ gate,
                                                       float threshold, const uint32_t *run_flag,
                                                       uint32_t *gate_host, int final_check) {
    // a chained check (run_flag set) after a check that passed keeps the gate closed
    if (run_flag && *run_flag == 0u) {
        if (gate_host && threadIdx.x == 0) *gate_host = 0u;
        return;
    }
    double hh = 0.0, ll = 0.0, mx = 0.0;
    for (uint32_t b = threadIdx.x; b < nblocks; b += 64) {
        hh += part[3 * b + 0];
        ll += part[3 * b + 1];
        mx = fmax(mx, part[3 * b + 2]);
    }
    hh = wave_sum(hh);
    ll = wave_sum(ll);
    mx = wave_max(mx);
    if (threadIdx.x == 0) {
        const double rel = ll > 0.0 ? sqrt(hh / ll) : 0.0;
        out[0] = hh;
        out[1] = ll;
        out[2] = mx;
        out[3] = rel;
        if (gate) {
            const uint32_t fail = rel > (double)threshold ? 1u : 0u;
            gate[0] = fail;
            // pinned host copy of the flag (host-coherent at the kernel's end)
            if (gate_host) *gate_host = fail;
            if (final_check) {
                gate[4] += fail;  // the budget is spent and the substep ends above the threshold
            } else if (run_flag) {
                gate[3] += fail;  // another fallback chunk
            } else {
                gate[1] += 1u;    // substeps checked
                gate[2] += fail;  // substeps that fell back
                gate[3] += fail;
            }
        }
    }
}

static inline uint32_t grid_for(size_t n, uint32_t block) { return uint32_t((n + block - 1) / block); }

hipError_t launch_predict(hipStream_t s, float4 *P, const float4 *X, float4 *V, uint32_t n, const SubstepArgs &a,
                          uint32_t block) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_predict, dim3(grid_for(n, block)), dim3(block), 0, s, P, X, V, n, a);
    return hipGetLastError();
}

hipError_t launch_distance(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, const float *cc, float *lam,
                           uint32_t off, uint32_t cnt, const SubstepArgs &a, bool first, bool split, uint32_t block) {
    if (!cnt) return hipSuccess;
    const bool damp = a.tension_damping != 0.0f || a.compression_damping != 0.0f || a.shear_damping != 0.0f;
    const dim3 g(grid_for(cnt, block)), b(block);
#define TC_DIST(F, D, S) hipLaunchKernelGGL((k_distance<F, D, S>), g, b, 0, s, P, X, rec, cc, lam, off, cnt, a)
    if (first) {
        if (damp) { if (split) TC_DIST(true, true, true); else TC_DIST(true, true, false); }
        else { if (split) TC_DIST(true, false, true); else TC_DIST(true, false, false); }
    } else {
        if (damp) { if (split) TC_DIST(false, true, true); else TC_DIST(false, true, false); }
        else { if (split) TC_DIST(false, false, true); else TC_DIST(false, false, false); }
    }
#undef TC_DIST
    return hipGetLastError();
}

hipError_t launch_bending(hipStream_t s, float4 *P, const float4 *X, const uint4 *rec, float *lam, uint32_t off,
                          uint32_t cnt, const SubstepArgs &a, bool first, uint32_t block, const uint4 *gate) {
    if (!cnt) return hipSuccess;
    const bool damp = a.bending_damping != 0.0f;
    const dim3 g(grid_for(cnt, block)), b(block);
    if (first) {
        if (damp) hipLaunchKernelGGL((k_bending<true, true>), g, b, 0, s, P, X, rec, lam, off, cnt, a, gate);
        else hipLaunchKernelGGL((k_bending<true, false>), g, b, 0, s, P, X, rec, lam, off, cnt, a, gate);
    } else {
        if (damp) hipLaunchKernelGGL((k_bending<false, true>), g, b, 0, s, P, X, rec, lam, off, cnt, a, gate);
        else hipLaunchKernelGGL((k_bending<false, false>), g, b, 0, s, P, X, rec, lam, off, cnt, a, gate);
    }
    return hipGetLastError();
}

hipError_t launch_collision(hipStream_t s, float4 *P, float4 *X, uint32_t n, const uint32_t *order,
                            const Collider *cols, uint32_t ncols, const MeshTri *tris, const float4 *bounds,
                            uint32_t ntris, const float4 *cbox, int recovery, float dt, uint32_t block,
                            const SdfArgs &sdf) {
    if (!n) return hipSuccess;
    if (block % 64u != 0u || block > 256u) return hipErrorInvalidValue;  // whole wavefronts, 4 LDS slots
    // Vertices per wavefront.  Against a triangle mesh each wavefront is a latency-bound chain
    // (cluster votes, then every lane's candidate triangles in order), and the chain grows with
    // the cloth area its lanes cover: a 16k-vertex cloth on 81,920 triangles takes 1.11 ms at 64
    // vertices per wavefront, 0.44 ms at 16 and 0.27 ms at 4.  So the wavefronts get fewer
    // vertices until there are about 8,192 of them (at least 4 vertices each).  Primitives and
    // SDFs are per-lane work: there, 64.
    uint32_t vpw = 64;
    if (ntris) {
        vpw = 4;
        while (vpw < 64u && size_t(vpw) * 2u * 8192u <= n) vpw *= 2u;
    }
    hipLaunchKernelGGL(k_collision, dim3(grid_for(n, (block / 64u) * vpw)), dim3(block), 0, s, P, X, n, order, cols,
                       ncols, tris, bounds, ntris, cbox, recovery, dt, sdf, vpw);
    return hipGetLastError();
}

hipError_t launch_update_vel(hipStream_t s, const float4 *P, float4 *X, float4 *V, uint32_t n, float dt,
                             uint32_t block) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_update_vel, dim3(grid_for(n, block)), dim3(block), 0, s, P, X, V, n, dt);
    return hipGetLastError();
}

hipError_t launch_extract_positions(hipStream_t s, const float4 *X, float *out, uint32_t n) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_extract, dim3(grid_for(n, 256)), dim3(256), 0, s, X, out, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t value, size_t n) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

hipError_t launch_residual(hipStream_t s, const float4 *P, const float4 *X, const uint4 *drec, const float *dcc,
                           const float *dlam, uint32_t nd, const uint4 *brec, const float *blam, uint32_t nb,
                           const SubstepArgs &a, bool split, double *part, double *out, uint32_t *gate,
                           float threshold, const uint32_t *run_flag, uint32_t *gate_host, bool final_check) {
    const size_t n = size_t(nd) + nb;
    const uint32_t blocks = uint32_t(std::min<size_t>(RES_BLOCKS, std::max<size_t>(grid_for(n, 256), 1)));
    SubstepArgs ra = a;
    ra.run_flag = nullptr;
    const bool damp = a.tension_damping > 0.0f || a.compression_damping > 0.0f || a.shear_damping > 0.0f ||
                      a.bending_damping > 0.0f;
    if (damp)
        hipLaunchKernelGGL(k_residual_part<true>, dim3(blocks), dim3(256), 0, s, P, X, drec, dcc, dlam, nd, brec,
                           blam, nb, ra, split ? 1 : 0, part, run_flag);
    else
        hipLaunchKernelGGL(k_residual_part<false>, dim3(blocks), dim3(256), 0, s, P, X, drec, dcc, dlam, nd, brec,
                           blam, nb, ra, split ? 1 : 0, part, run_flag);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(64), 0, s, part, blocks, out, gate, threshold, run_flag,
                       gate_host, final_check ? 1 : 0);
    return hipGetLastError();
}

}  // namespace tc
