// One solver iteration (all distance colours, then all bending colours) per
// launch, one workgroup per cone tile (plan in tc_tiles.cpp).
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
// word1 = constraint index (26 bits) | compliance class << 26 | shear << 30 |
// owner << 31, word2 = rest length bits.
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
template <bool DIST, bool DAMP, bool FAST>
__device__ __forceinline__ void project_item(float4 *__restrict__ sp, const float4 *__restrict__ sx, uint32_t w0,
                                             uint32_t w1, uint32_t w2, float l, float *__restrict__ lam_out,
                                             const float *__restrict__ cls_tab, const SubstepArgs &a, float dt) {
    const float rest = __uint_as_float(w2);
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

// The first item of every phase a thread owns is software-pipelined: its
// words are loaded one phase ahead and its lambda right after the current
// phase's LDS work, so global latency overlaps the barrier.  Further items of a
// phase (phases larger than the block) are loaded in place.
template <bool FIRST, bool DAMP, bool FAST>
__global__ void __launch_bounds__(1024) k_tile_iteration(
    const float4 *__restrict__ Pin, float4 *__restrict__ Pout, const float4 *__restrict__ X,
    const uint32_t *__restrict__ ext_off, const uint32_t *__restrict__ ext_vid, const uint32_t *__restrict__ n_own,
    const uint32_t *__restrict__ ph_off, const uint32_t *__restrict__ items, const float *__restrict__ dlam_in,
    float *__restrict__ dlam_out, const float *__restrict__ blam_in, float *__restrict__ blam_out, FusedArgs fa,
    SubstepArgs a) {
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
    const uint32_t L = fa.nphases, ndp = fa.nd_phases;
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
    if (!FIRST && h_cur) lc = (ndp > 0 ? dlam_in : blam_in)[c1 & 0x03FFFFFFu];
    __syncthreads();
    const float dt = fmaxf(a.dt, 1e-8f);
    for (uint32_t c = 0; c < L; ++c) {
        const bool dist = c < ndp;
        // the next phase's lambda (its item arrived during the previous phase) and the item after
        float ln = 0.0f;
        if (!FIRST && h_nxt) ln = (c + 1 < ndp ? dlam_in : blam_in)[n1 & 0x03FFFFFFu];
        uint32_t m0 = 0, m1 = 0, m2 = 0;
        const bool h_nn = load_item(c + 2, m0, m1, m2);
        const float *lam_in = dist ? dlam_in : blam_in;
        float *lam_out = dist ? dlam_out : blam_out;
        if (h_cur) {
            if (dist) project_item<true, DAMP, FAST>(sp, sx, c0, c1, c2, lc, lam_out, cls_tab, a, dt);
            else project_item<false, DAMP, FAST>(sp, sx, c0, c1, c2, lc, lam_out, cls_tab, a, dt);
        }
        for (uint32_t i = po[c] + tid + bs; i < po[c + 1]; i += bs) {
            const uint32_t w0 = items[3 * size_t(i)], w1 = items[3 * size_t(i) + 1], w2 = items[3 * size_t(i) + 2];
            const float l = FIRST ? 0.0f : lam_in[w1 & 0x03FFFFFFu];
            if (dist) project_item<true, DAMP, FAST>(sp, sx, w0, w1, w2, l, lam_out, cls_tab, a, dt);
            else project_item<false, DAMP, FAST>(sp, sx, w0, w1, w2, l, lam_out, cls_tab, a, dt);
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
                                 float *blam_out, const FusedArgs &fa, const SubstepArgs &a, bool first,
                                 uint32_t block) {
    if (!fa.ntiles) return hipSuccess;
    FusedArgs fx = fa;
    fused_alphas(fx, a.dt);
    const bool damp = a.tension_damping != 0.0f || a.compression_damping != 0.0f || a.shear_damping != 0.0f ||
                      a.bending_damping != 0.0f;
    const uint32_t nblocks = ((fa.ntiles + 7u) / 8u) * 8u;
    const size_t lds = size_t(fa.max_ext) * sizeof(float4) * (damp ? 2 : 1);
#define TC_TILE(K)                                                                                               \
    hipLaunchKernelGGL(K, dim3(nblocks), dim3(block), lds, s, Pin, Pout, X, p.ext_off, p.ext_vid, p.n_own,        \
                       p.ph_off, p.items, dlam_in, dlam_out, blam_in, blam_out, fx, a)
    if (fa.fast) {
        if (first) { if (damp) TC_TILE((k_tile_iteration<true, true, true>)); else TC_TILE((k_tile_iteration<true, false, true>)); }
        else { if (damp) TC_TILE((k_tile_iteration<false, true, true>)); else TC_TILE((k_tile_iteration<false, false, true>)); }
    } else {
        if (first) { if (damp) TC_TILE((k_tile_iteration<true, true, false>)); else TC_TILE((k_tile_iteration<true, false, false>)); }
        else { if (damp) TC_TILE((k_tile_iteration<false, true, false>)); else TC_TILE((k_tile_iteration<false, false, false>)); }
    }
#undef TC_TILE
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
    const void *ks[] = {
        reinterpret_cast<const void *>(&k_tile_iteration<true, true, false>),
        reinterpret_cast<const void *>(&k_tile_iteration<true, false, false>),
        reinterpret_cast<const void *>(&k_tile_iteration<false, true, false>),
        reinterpret_cast<const void *>(&k_tile_iteration<false, false, false>),
        reinterpret_cast<const void *>(&k_tile_iteration<true, true, true>),
        reinterpret_cast<const void *>(&k_tile_iteration<true, false, true>),
        reinterpret_cast<const void *>(&k_tile_iteration<false, true, true>),
        reinterpret_cast<const void *>(&k_tile_iteration<false, false, true>),
    };
    for (const void *k : ks) {
        const hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace tc
