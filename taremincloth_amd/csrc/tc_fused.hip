// One solver iteration (all distance colours, all sewing colours, then all
// bending colours, the order of dispatch.rs:167-205) per launch, one workgroup
// per cone tile (plan in tc_tiles.cpp).
//
// Per workgroup: load the tile's cone of P into LDS (16 B per vertex), replay
// the selected constraints of each colour phase from LDS with a block barrier
// between phases, write back the tile's own vertices and the lambdas of the
// constraints it owns.  The projection arithmetic is the same operation
// sequence as k_distance / k_bending (distance.wgsl:60-124, bending.wgsl:59-126), or, with the
// fast projection switched on, the same formula with the square root and the divisions replaced
// by refined hardware reciprocals (not bit-identical; within 1e-6 relative per projection).
//
// Positions and lambdas are double-buffered across launches (the cones of
// neighbouring tiles overlap, so a tile must not read what another tile of the
// same launch writes).  Items (12 B): word0 = local v0 | local v1 << 16,
// word1 = constraint index (26 bits) | compliance class << 26 | flag << 30 |
// owner << 31, word2 = rest length bits.  The flag marks a shear constraint
// among the distance items and a seam-gated constraint among the bending items;
// a gated item's word2 holds the local indices of its hinge vertices instead
// (rest length from the bending record).  Sewing items carry only the index:
// their current rest length is in the sewing record, and the lock may rewrite
// it; only the owner writes, and every tile replaying the item takes the lock
// branch alike, where the old value is not read.
//
// Tiles come out of a depth-first coordinate bisection, so consecutive tiles are spatial neighbours; the
// block -> tile map keeps a contiguous run of tiles on each of the 8 XCDs
// (workgroups are dealt to XCDs round-robin) so overlapping cone reads of P,
// lambda and items share that XCD's L2.
#include "tc_kernels.h"

namespace tc {

__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblocks) {
    const uint32_t per = nblocks / 8u;  // nblocks is a multiple of 8
    return (b % 8u) * per + b / 8u;
}

// 1/sqrt(x) and 1/x from the hardware approximations (about 1 ulp) refined by one Newton-Raphson
// step each: y <- y (3 - x y^2) / 2 and r <- r (2 - x r).
__device__ __forceinline__ float rsqrt_nr(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    return y * (1.5f - (0.5f * x) * (y * y));
}
__device__ __forceinline__ float rcp_nr(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return r * (2.0f - x * r);
}

// Projection of one item from LDS (the k_distance / k_bending arithmetic).
// Class table in LDS (per-lane indexed): [0] dtc [16] dcc [32] bcomp [48] dalpha_t [64] dalpha_c [80] balpha
template <bool DIST, bool DAMP, bool FAST, bool GATE>
__device__ __forceinline__ void project_item(float4 *__restrict__ sp, const float4 *__restrict__ sx, uint32_t w0,
                                             uint32_t w1, uint32_t w2, float l, float *__restrict__ lam_out,
                                             const float *__restrict__ cls_tab, const SubstepArgs &a, float dt,
                                             const uint4 *__restrict__ brec, const uint4 *__restrict__ bgate) {
    const bool gated = GATE && !DIST && ((w1 >> 30) & 1u) != 0u;
    float rest = __uint_as_float(w2);
    const uint32_t l0 = w0 & 0xFFFFu, l1 = w0 >> 16;
    const uint32_t k = w1 & 0x03FFFFFFu, cls = (w1 >> 26) & 15u;
    const bool shear = ((w1 >> 30) & 1u) != 0u, owner = (w1 >> 31) != 0u;
    const float4 q0 = sp[l0], q1 = sp[l1];
    const float wa = q0.w, wb = q1.w, wsum = wa + wb;
    // alpha~ = compliance / dt^2 is precomputed per compliance class (same operations, same bits)
    float alpha;
    if (DIST) {
        // an owner that skips the projection still carries lambda into the output buffer
        if (wsum <= 1e-7f) { if (owner) lam_out[k] = l; return; }
    } else {
        if (cls_tab[32 + cls] >= 1e9f || wsum <= 1e-7f) { if (owner) lam_out[k] = l; return; }
        alpha = cls_tab[80 + cls];
    }
    const f3 delta = xyz(q0) - xyz(q1);
    float d, inv_d = 0.0f;
    if (FAST) {
        // one reciprocal square root gives the length and the direction; |delta| <= 1e-7 as d^2 <= 1e-14
        const float dsq = dot(delta, delta);
        if (dsq <= 1e-14f) { if (owner) lam_out[k] = l; return; }
        inv_d = rsqrt_nr(dsq);
        d = dsq * inv_d;
    } else {
        d = length(delta);
        if (d <= 1e-7f) { if (owner) lam_out[k] = l; return; }
    }
    if (gated) {  // the seam gate of k_bending: only while closer than rest and the seam nearly closed
        rest = __uint_as_float(brec[k].z);
        if (d >= rest) { if (owner) lam_out[k] = l; return; }
        const float gap = length(xyz(sp[w2 & 0xFFFFu]) - xyz(sp[w2 >> 16]));
        if (gap > __uint_as_float(bgate[k].z)) { if (owner) lam_out[k] = l; return; }
    }
    const float cval = d - rest;
    if (DIST) {
        const bool tens = cval >= 0.0f;
        if (cls_tab[(tens ? 0 : 16) + cls] >= 1e9f) { if (owner) lam_out[k] = l; return; }
        alpha = cls_tab[(tens ? 48 : 64) + cls];
    }
    const f3 dir = FAST ? delta * inv_d : delta / d;
    float dl;
    if (DAMP) {
        const float dc = DIST ? (shear ? a.shear_damping : (cval >= 0.0f ? a.tension_damping : a.compression_damping))
                              : a.bending_damping;
        const float gamma = fminf(dc * 1e-5f / dt, 1e4f);
        const f3 rel = (xyz(q0) - xyz(sx[l0])) - (xyz(q1) - xyz(sx[l1]));
        const float num = -cval - alpha * l - gamma * dot(dir, rel), den = (1.0f + gamma) * wsum + alpha;
        dl = FAST ? num * rcp_nr(den) : num / den;
    } else {
        const float num = -cval - alpha * l, den = wsum + alpha;
        dl = FAST ? num * rcp_nr(den) : num / den;
    }
    if (owner) lam_out[k] = l + dl;
    const f3 c0 = dir * (wa * dl), c1 = dir * (-wb * dl);
    sp[l0] = make_float4(q0.x + c0.x, q0.y + c0.y, q0.z + c0.z, wa);
    sp[l1] = make_float4(q1.x + c1.x, q1.y + c1.y, q1.z + c1.z, wb);
}

// One sewing item from LDS: the k_sewing arithmetic (sewing.wgsl:56-110), exact in both modes.
__device__ __forceinline__ void project_sew(float4 *__restrict__ sp, uint32_t w0, uint32_t w1, float l,
                                            float *__restrict__ lam_out, uint4 *__restrict__ srec,
                                            const float4 *__restrict__ X, const FusedArgs &fa, float dt) {
    const uint32_t l0 = w0 & 0xFFFFu, l1 = w0 >> 16;
    const uint32_t k = w1 & 0x03FFFFFFu;
    const bool owner = (w1 >> 31) != 0u;
    const float4 q0 = sp[l0], q1 = sp[l1];
    const float wa = q0.w, wb = q1.w, wsum = wa + wb;
    if (wsum <= 1e-7f) { if (owner) lam_out[k] = l; return; }
    const f3 delta = xyz(q0) - xyz(q1);
    const float dist = length(delta);
    if (dist <= 1e-7f) { if (owner) lam_out[k] = l; return; }
    const f3 dir = delta / dist;
    const uint4 r = srec[k];
    float target = __uint_as_float(r.z), comp = fa.sew_comp;
    const float lock_thresh = fmaxf(X[r.x].w + X[r.y].w, fa.sew_lock_dist);  // X.w: vertex thickness
    if (fa.sew_lock > 0.5f && dist <= lock_thresh) {
        comp = 0.0f;
        target = __uint_as_float(r.w);
        if (owner) srec[k].z = r.w;
    }
    const float alpha = comp / (dt * dt);
    const float c = dist - target;
    const float dl = (-c - alpha * l) / (wsum + alpha);
    if (owner) lam_out[k] = l + dl;
    const f3 c0 = dir * (wa * dl), c1 = dir * (wb * dl);
    sp[l0] = make_float4(q0.x + c0.x, q0.y + c0.y, q0.z + c0.z, wa);
    sp[l1] = make_float4(q1.x - c1.x, q1.y - c1.y, q1.z - c1.z, wb);
}

// The first item of every phase a thread owns is software-pipelined: its
// words are loaded one phase ahead and its lambda right after the current
// phase's LDS work, so global latency overlaps the barrier.  Further items of a
// phase (phases larger than the block) are loaded in place.
// SEW: the plan has sewing phases or seam-gated bending items (a separate instantiation, so plain
// cloth keeps the shorter per-phase code of distance and bending alone)
template <bool FIRST, bool DAMP, bool FAST, bool SEW>
__global__ void __launch_bounds__(1024) k_tile_iteration(
    const float4 *__restrict__ Pin, float4 *__restrict__ Pout, const float4 *__restrict__ X,
    const uint32_t *__restrict__ ext_off, const uint32_t *__restrict__ ext_vid, const uint32_t *__restrict__ n_own,
    const uint32_t *__restrict__ ph_off, const uint32_t *__restrict__ items, const float *__restrict__ dlam_in,
    float *__restrict__ dlam_out, const float *__restrict__ blam_in, float *__restrict__ blam_out,
    const float *__restrict__ slam_in, float *__restrict__ slam_out, uint4 *__restrict__ srec,
    const uint4 *__restrict__ brec, const uint4 *__restrict__ bgate, FusedArgs fa, SubstepArgs a) {
    extern __shared__ float4 sp[];
    __shared__ float cls_tab[96];
    if (a.run_flag && *a.run_flag == 0u) return;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= fa.ntiles) return;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    if (tid < 16) {
        cls_tab[tid] = fa.dtc[tid];
        cls_tab[16 + tid] = fa.dcc[tid];
        cls_tab[32 + tid] = fa.bcomp[tid];
        cls_tab[48 + tid] = fa.dalpha_t[tid];
        cls_tab[64 + tid] = fa.dalpha_c[tid];
        cls_tab[80 + tid] = fa.balpha[tid];
    }
    const uint32_t *po = ph_off + size_t(tile) * (fa.nphases + 1);
    const uint32_t L = fa.nphases, ndp = fa.nd_phases, nds = SEW ? fa.nd_phases + fa.ns_phases : fa.nd_phases;
    // lambda of item index k in phase c: a branch per buffer (a select between the buffer pointers
    // would turn the loads into flat loads, whose counter the LDS waits then share)
    auto lam_at = [&](uint32_t c, uint32_t k) -> float {
        if (c < ndp) return dlam_in[k];
        if (SEW && c < nds) return slam_in[k];
        return blam_in[k];
    };
    // A thread's first item of a phase moves through a two-deep pipeline: its words are loaded
    // two phases ahead and its lambda one phase ahead, so neither load waits at a phase start.
    auto load_item = [&](uint32_t c, uint32_t &w0, uint32_t &w1, uint32_t &w2) -> bool {
        if (c >= L) return false;
        const uint32_t i = po[c] + tid;
        if (i >= po[c + 1]) return false;
        w0 = items[3 * size_t(i)]; w1 = items[3 * size_t(i) + 1]; w2 = items[3 * size_t(i) + 2];
        return true;
    };
    uint32_t c0 = 0, c1 = 0, c2 = 0, n0 = 0, n1 = 0, n2 = 0;
    bool h_cur = load_item(0, c0, c1, c2);
    bool h_nxt = load_item(1, n0, n1, n2);
    const uint32_t e0 = ext_off[tile], ne = ext_off[tile + 1] - e0;
    float4 *sx = sp + fa.max_ext;  // DAMP only: committed positions of the cone
    for (uint32_t l = tid; l < ne; l += bs) {
        const uint32_t g = ext_vid[e0 + l];
        sp[l] = Pin[g];
        if (DAMP) sx[l] = X[g];
    }
    float lc = 0.0f;
    if (!FIRST && h_cur) lc = lam_at(0, c1 & 0x03FFFFFFu);
    __syncthreads();
    const float dt = fmaxf(a.dt, 1e-8f);
    for (uint32_t c = 0; c < L; ++c) {
        const uint32_t kind = c < ndp ? 0u : (SEW && c < nds) ? 1u : 2u;  // distance, sewing, bending
        // the next phase's lambda (its item arrived during the previous phase) and the item after
        float ln = 0.0f;
        if (!FIRST && h_nxt) ln = lam_at(c + 1, n1 & 0x03FFFFFFu);
        uint32_t m0 = 0, m1 = 0, m2 = 0;
        const bool h_nn = load_item(c + 2, m0, m1, m2);
        if (h_cur) {
            if (kind == 0u) project_item<true, DAMP, FAST, SEW>(sp, sx, c0, c1, c2, lc, dlam_out, cls_tab, a, dt, brec, bgate);
            else if (kind == 2u) project_item<false, DAMP, FAST, SEW>(sp, sx, c0, c1, c2, lc, blam_out, cls_tab, a, dt, brec, bgate);
            else if (SEW) project_sew(sp, c0, c1, lc, slam_out, srec, X, fa, dt);
        }
        for (uint32_t i = po[c] + tid + bs; i < po[c + 1]; i += bs) {
            const uint32_t w0 = items[3 * size_t(i)], w1 = items[3 * size_t(i) + 1], w2 = items[3 * size_t(i) + 2];
            const float l = FIRST ? 0.0f : lam_at(c, w1 & 0x03FFFFFFu);
            if (kind == 0u) project_item<true, DAMP, FAST, SEW>(sp, sx, w0, w1, w2, l, dlam_out, cls_tab, a, dt, brec, bgate);
            else if (kind == 2u) project_item<false, DAMP, FAST, SEW>(sp, sx, w0, w1, w2, l, blam_out, cls_tab, a, dt, brec, bgate);
            else if (SEW) project_sew(sp, w0, w1, l, slam_out, srec, X, fa, dt);
        }
        c0 = n0; c1 = n1; c2 = n2; h_cur = h_nxt; lc = ln;
        n0 = m0; n1 = m1; n2 = m2; h_nxt = h_nn;
        __syncthreads();
    }
    const uint32_t own = n_own[tile];
    for (uint32_t l = tid; l < own; l += bs) Pout[ext_vid[e0 + l]] = sp[l];
}

hipError_t launch_tile_iteration(hipStream_t s, const float4 *Pin, float4 *Pout, const float4 *X,
                                 const FusedPlanDev &p, const float *dlam_in, float *dlam_out, const float *blam_in,
                                 float *blam_out, const float *slam_in, float *slam_out, const FusedArgs &fa,
                                 const SubstepArgs &a, bool first, uint32_t block) {
    if (!fa.ntiles) return hipSuccess;
    if (fa.ns_phases && (!p.srec || !slam_in || !slam_out)) return hipErrorInvalidValue;
    FusedArgs fx = fa;
    fused_alphas(fx, a.dt);
    const bool damp = a.tension_damping != 0.0f || a.compression_damping != 0.0f || a.shear_damping != 0.0f ||
                      a.bending_damping != 0.0f;
    const uint32_t nblocks = ((fa.ntiles + 7u) / 8u) * 8u;
    const size_t lds = size_t(fa.max_ext) * sizeof(float4) * (damp ? 2 : 1);
    const bool sew = fa.ns_phases != 0u || p.bgate != nullptr;
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(block), lds, s, Pin, Pout, X, p.ext_off, p.ext_vid, p.n_own,
                           p.ph_off, p.items, dlam_in, dlam_out, blam_in, blam_out, slam_in, slam_out, p.srec, p.brec,
                           p.bgate, fx, a);
    };
    const int v = (first ? 8 : 0) | (damp ? 4 : 0) | (fa.fast ? 2 : 0) | (sew ? 1 : 0);
    switch (v) {
#define TC_TILE(V) case V: go(k_tile_iteration<(V & 8) != 0, (V & 4) != 0, (V & 2) != 0, (V & 1) != 0>); break;
        TC_TILE(0) TC_TILE(1) TC_TILE(2) TC_TILE(3) TC_TILE(4) TC_TILE(5) TC_TILE(6) TC_TILE(7)
        TC_TILE(8) TC_TILE(9) TC_TILE(10) TC_TILE(11) TC_TILE(12) TC_TILE(13) TC_TILE(14) TC_TILE(15)
#undef TC_TILE
    }
    return hipGetLastError();
}

void fused_alphas(FusedArgs &fa, float dt_sub) {
    const float dt = fmaxf(dt_sub, 1e-8f);
    const float dt2 = dt * dt;
    for (int c = 0; c < 16; ++c) {
        fa.dalpha_t[c] = fa.dtc[c] / dt2;
        fa.dalpha_c[c] = fa.dcc[c] / dt2;
        fa.balpha[c] = fa.bcomp[c] / dt2;
    }
}

hipError_t set_tile_lds_limit(size_t bytes) {
#define TC_TILE(V) reinterpret_cast<const void *>(&k_tile_iteration<(V & 8) != 0, (V & 4) != 0, (V & 2) != 0, (V & 1) != 0>),
    const void *ks[] = {TC_TILE(0) TC_TILE(1) TC_TILE(2) TC_TILE(3) TC_TILE(4) TC_TILE(5) TC_TILE(6) TC_TILE(7)
                        TC_TILE(8) TC_TILE(9) TC_TILE(10) TC_TILE(11) TC_TILE(12) TC_TILE(13) TC_TILE(14) TC_TILE(15)};
#undef TC_TILE
    for (const void *k : ks) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace tc
