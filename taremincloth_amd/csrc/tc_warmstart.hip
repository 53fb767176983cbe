// Learned warm start kernels (gfx950): features, BF16 MFMA forward / backward of
// the message-passing network, weight-gradient GEMMs, Adam, and the apply /
// bookkeeping passes around the projection loop.  Layer algebra in tc_ws.h.
//
// MFMA used throughout: v_mfma_f32_16x16x32_bf16 (one wave, 16x32 A, 32x16 B,
// 16x16 FP32 C).  Lane l = 16 g + c holds A row c, B column c, each with the
// eight K values 8g..8g+7, and C rows 4g..4g+3 of column c.  The forward layers
// compute the transposed product (weights as A, activations as B) so that a
// lane's four results are consecutive features of one row.
#include <algorithm>
#include <cstring>

#include "tc_device.h"
#include "tc_ws.h"

namespace tc {
namespace ws {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
    for (int i = 0; i < 8; ++i) z[i] = (bf16)0.0f;
    return z;
}

NetLayout make_layout(uint32_t rounds) {
    NetLayout nl;
    nl.rounds = rounds;
    uint32_t off = 0, n = 0;
    auto add = [&](uint32_t K, uint32_t N) {
        nl.L[n++] = {off, K, N};
        off += K * N + N;
    };
    add(F, H);
    add(H, H);
    for (uint32_t r = 0; r < rounds; ++r) {
        add(2 * H, H);
        add(H, H);
    }
    add(H, H);
    add(H, O);
    nl.nlayers = n;
    nl.nparams = off;
    return nl;
}

// ---------------------------------------------------------------- features
__device__ __forceinline__ float soft(float x) { return x / (1.0f + fabsf(x) * (1.0f / 16.0f)); }

// Sum over the incident pairs k of vertex i of d_k (|d_k| - rest_k) / |d_k|, d_k = p_j - p0, in
// incidence order.  The first FEAT_NB neighbours' ids, rest lengths and positions are all
// requested before any is used (ids from index 0 and positions of vertex i stand in for the
// slots past the degree, so every load is in bounds), which keeps the chain of dependent
// loads two deep; further neighbours follow one by one.
constexpr int FEAT_NB = 8;
__device__ __forceinline__ f3 strain_sum(const FeatArgs &a, uint32_t i, f3 p0, uint32_t k0, uint32_t k1,
                                         const uint32_t *__restrict__ nbr, const float *__restrict__ rest) {
    const float bw = a.beta;
    const uint32_t n = k1 - k0;
    uint32_t js[FEAT_NB];
    float rs[FEAT_NB];
#pragma unroll
    for (int k = 0; k < FEAT_NB; ++k) {
        const uint32_t q = uint32_t(k) < n ? k0 + k : 0u;
        js[k] = nbr[q];
        rs[k] = rest[q];
    }
    float4 pj[FEAT_NB], dj[FEAT_NB];
#pragma unroll
    for (int k = 0; k < FEAT_NB; ++k) {
        const uint32_t j = uint32_t(k) < n ? js[k] : i;
        pj[k] = a.P[j];
        dj[k] = a.D[j];
    }
    f3 acc = mk3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < FEAT_NB; ++k) {
        const f3 d = mk3(pj[k].x + bw * dj[k].x, pj[k].y + bw * dj[k].y, pj[k].z + bw * dj[k].z) - p0;
        const float len = length(d);
        if (uint32_t(k) < n && len > 1e-12f) acc = acc + d * ((len - rs[k]) / len);
    }
    for (uint32_t k = k0 + FEAT_NB; k < k1; ++k) {
        const uint32_t j = nbr[k];
        const float4 p = a.P[j], dd = a.D[j];
        const f3 d = mk3(p.x + bw * dd.x, p.y + bw * dd.y, p.z + bw * dd.z) - p0;
        const float len = length(d);
        if (len > 1e-12f) acc = acc + d * ((len - rest[k]) / len);
    }
    return acc;
}

__global__ void __launch_bounds__(256) k_features(FeatArgs a, bf16 *__restrict__ feat) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nv) return;
    const float4 pi = a.P[i], di = a.D[i];
    const float bw = a.beta;
    const f3 p0 = mk3(pi.x + bw * di.x, pi.y + bw * di.y, pi.z + bw * di.z);
    const float is = 1.0f / a.s;
    float f[F];
    for (int k = 0; k < F; ++k) f[k] = 0.0f;
    f[0] = soft(di.x * is);
    f[1] = soft(di.y * is);
    f[2] = soft(di.z * is);
    const uint32_t d0 = a.d_off[i], d1 = a.d_off[i + 1];
    const f3 sd = strain_sum(a, i, p0, d0, d1, a.d_nbr, a.d_rest);
    f[3] = soft(sd.x * is);
    f[4] = soft(sd.y * is);
    f[5] = soft(sd.z * is);
    const uint32_t b0 = a.b_off[i], b1 = a.b_off[i + 1];
    const f3 sb = strain_sum(a, i, p0, b0, b1, a.b_nbr, a.b_rest);
    f[6] = soft(sb.x * is);
    f[7] = soft(sb.y * is);
    f[8] = soft(sb.z * is);
    f[9] = pi.w > 0.0f ? 1.0f : 0.0f;
    f[10] = float(d1 - d0) * 0.125f;
    f[11] = float(b1 - b0) * 0.125f;
    // nearest primitive collider surface (signed, minus the vertex thickness) and its normal
    const Collider *cols = static_cast<const Collider *>(a.cols);
    const float th = a.X[i].w;
    float best = 3.0e38f;
    f3 bn = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t c = 0; c < a.ncols; ++c) {
        const Collider col = cols[c];
        const f3 A = mk3(col.ax, col.ay, col.az);
        float sd;
        f3 n;
        if (col.type == 2u) {
            const f3 B = mk3(col.bx, col.by, col.bz);
            n = B / length(B);
            sd = dot(p0 - A, n);
        } else {
            f3 q = A;
            if (col.type == 1u) {
                const f3 ab = mk3(col.bx, col.by, col.bz) - A;
                const float l2 = dot(ab, ab);
                float t = 0.0f;
                if (l2 > 1e-7f) t = clampf(dot(p0 - A, ab) / l2, 0.0f, 1.0f);
                q = A + ab * t;
            }
            const f3 d = p0 - q;
            const float dist = length(d);
            n = dist > 1e-7f ? d / dist : mk3(0.0f, 0.0f, 1.0f);
            sd = dist - col.radius;
        }
        sd = sd - th;
        if (sd < best) { best = sd; bn = n; }
    }
    const float r = a.ncols ? fminf(best / a.lscale, 4.0f) : 4.0f;
    f[12] = soft(r);
    if (r < 4.0f) { f[13] = bn.x; f[14] = bn.y; f[15] = bn.z; }
    bf16x8 *out = reinterpret_cast<bf16x8 *>(feat + size_t(i) * F);
    for (int q = 0; q < F / 8; ++q) {
        bf16x8 v;
        for (int k = 0; k < 8; ++k) v[k] = (bf16)f[q * 8 + k];
        out[q] = v;
    }
}

__global__ void k_copy_pred(const float4 *__restrict__ P, float4 *__restrict__ Ppred, uint32_t nv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) Ppred[i] = P[i];
}

__global__ void k_apply(float4 *__restrict__ P, float4 *__restrict__ Ppred, const float4 *__restrict__ D,
                        const float *__restrict__ y, float s, float beta, uint32_t nv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float4 p = P[i];
    Ppred[i] = p;
    if (p.w > 0.0f) {
        const float4 d = D[i];
        float ox = 0.0f, oy = 0.0f, oz = 0.0f;
        if (y) {
            ox = y[size_t(i) * O + 0];
            oy = y[size_t(i) * O + 1];
            oz = y[size_t(i) * O + 2];
        }
        p.x = p.x + (beta * d.x + s * ox);
        p.y = p.y + (beta * d.y + s * oy);
        p.z = p.z + (beta * d.z + s * oz);
        P[i] = p;
    }
}

__global__ void k_update_prev(const float4 *__restrict__ P, const float4 *__restrict__ Ppred, float4 *__restrict__ D,
                              float4 *__restrict__ T, float s, float beta, uint32_t nv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const float4 p = P[i], q = Ppred[i];
    const bool free_v = q.w > 0.0f;
    float4 nd = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (free_v) nd = make_float4(p.x - q.x, p.y - q.y, p.z - q.z, 0.0f);
    if (T) {
        const float4 od = D[i];
        const float is = 1.0f / s;
        T[i] = free_v ? make_float4((nd.x - beta * od.x) * is, (nd.y - beta * od.y) * is, (nd.z - beta * od.z) * is,
                                    1.0f)
                      : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    D[i] = nd;
}

// ---------------------------------------------------------------- forward
enum { M_ENC = 0, M_DEC = 2 };

struct MlpFwd {
    uint32_t rows, ntiles;
    const bf16 *xin;     // ENC: features (rows x F); DEC: h (rows x H)
    const bf16 *w1t;
    const float *b1;
    const bf16 *w2t;
    const float *b2;
    bf16 *hout;
    float *yout;
    bf16 *xsave, *asave;
};

template <int K1, int N2>
constexpr size_t fwd_lds() {
    return size_t(2) * (TR * (K1 + 8) + TR * (H + 8));
}

// 16-byte chunk q of a tile's input row block: row r0 + q / (K1/8), column chunk q % (K1/8).
template <int K1, int MODE>
__device__ __forceinline__ bf16x8 load_x_chunk(const MlpFwd &a, uint32_t r0, int q) {
    const int rr = q / (K1 / 8), ch = q % (K1 / 8);
    const uint32_t row = r0 + rr;
    if (row >= a.rows) return zero8();
    return *reinterpret_cast<const bf16x8 *>(&a.xin[size_t(row) * K1 + ch * 8]);
}

// Fused two-layer block over tiles of 64 rows: X -> relu(X W1 + b1) -> (.) W2 + b2.
// Both products are computed transposed (weights as the MFMA A operand, activations as
// B), so a lane's four results are consecutive features of one row.  Wave w owns the
// 64 output columns 64*(w&1).. of the hidden layer and the 32 rows 32*(w>>1).. of the
// tile; its weight fragments stay in registers for the whole persistent launch, so LDS
// only carries the activations.  The next tile's input is loaded into registers while
// the current tile computes.
template <int K1, int N2, int MODE>
__global__ void __launch_bounds__(256) k_mlp_fwd(MlpFwd a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int LK = K1 + 8, LH = H + 8;
    constexpr int XCH = TR * K1 / 8 / 256;  // 16-byte input chunks per thread and tile
    constexpr int KS1 = K1 / 32, KS2 = H / 32;
    constexpr int CT2 = N2 == H ? 4 : 1;     // layer-2 column tiles per wave
    bf16 *xs = reinterpret_cast<bf16 *>(smem);  // TR x LK
    bf16 *as = xs + TR * LK;                    // TR x LH
    __shared__ __align__(16) float b1s[H], b2s[N2];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    const int cg = w & 1, rg = w >> 1;
    bf16x8 wf1[4][KS1], wf2[CT2][KS2];
    for (int ct = 0; ct < 4; ++ct)
        for (int ks = 0; ks < KS1; ++ks)
            wf1[ct][ks] = *reinterpret_cast<const bf16x8 *>(&a.w1t[(64 * cg + 16 * ct + c) * K1 + ks * 32 + g * 8]);
    for (int ct = 0; ct < CT2; ++ct) {
        const int n = N2 == H ? 64 * cg + 16 * ct + c : c;
        for (int ks = 0; ks < KS2; ++ks)
            wf2[ct][ks] = *reinterpret_cast<const bf16x8 *>(&a.w2t[n * H + ks * 32 + g * 8]);
    }
    if (tid < H) b1s[tid] = a.b1[tid];
    if (tid < N2) b2s[tid] = a.b2[tid];
    bf16x8 nx[XCH];
    // A tile's output is stored during the next tile, right after that tile's input load is
    // issued, so the wait for the input never waits for stores issued just before it.  The
    // first pass stores zeros into its own tile's rows, which it overwrites later.
    constexpr int OCH = TR * (H / 8) / 256;  // 16-byte output chunks per thread (hidden output)
    bf16x8 ov[OCH];
    for (int u = 0; u < OCH; ++u) ov[u] = zero8();
    float4 yv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // decoder output of this lane's row
    uint32_t t = blockIdx.x, tp = t;
    if (t < a.ntiles)
        for (int u = 0; u < XCH; ++u) nx[u] = load_x_chunk<K1, MODE>(a, t * TR, tid + 256 * u);
    auto store_out = [&](uint32_t tile) {
        if (MODE == M_DEC) {
            const uint32_t row = tile * TR + 16 * w + c;
            if (row < a.rows) *reinterpret_cast<float4 *>(&a.yout[size_t(row) * O + 4 * g]) = yv;
        } else {
            for (int u = 0; u < OCH; ++u) {
                const int q = tid + 256 * u, r2 = q / (H / 8), ch = q % (H / 8);
                const uint32_t row2 = tile * TR + r2;
                if (row2 < a.rows) *reinterpret_cast<bf16x8 *>(&a.hout[size_t(row2) * H + ch * 8]) = ov[u];
            }
        }
    };
    for (; t < a.ntiles; t += gridDim.x) {
        const uint32_t r0 = t * TR;
        __syncthreads();
        for (int u = 0; u < XCH; ++u) {
            const int q = tid + 256 * u, rr = q / (K1 / 8), ch = q % (K1 / 8);
            *reinterpret_cast<bf16x8 *>(&xs[rr * LK + ch * 8]) = nx[u];
        }
        const uint32_t tn = t + gridDim.x;
        if (tn < a.ntiles)
            for (int u = 0; u < XCH; ++u) nx[u] = load_x_chunk<K1, MODE>(a, tn * TR, tid + 256 * u);
        store_out(tp);
        tp = t;
        __syncthreads();
        if (a.xsave) {
            for (int q = tid; q < TR * (K1 / 8); q += 256) {
                const int rr = q / (K1 / 8), ch = q % (K1 / 8);
                const uint32_t row = r0 + rr;
                if (row < a.rows)
                    *reinterpret_cast<bf16x8 *>(&a.xsave[size_t(row) * K1 + ch * 8]) =
                        *reinterpret_cast<const bf16x8 *>(&xs[rr * LK + ch * 8]);
            }
        }
        // layer 1: hidden units 64cg + 16ct + 4g + j of rows 32rg + 16rt + c
        for (int rt = 0; rt < 2; ++rt) {
            const int rr = 32 * rg + 16 * rt + c;
            f32x4 acc[4];
            for (int ct = 0; ct < 4; ++ct) acc[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            // all of the row's B fragments first, so that the LDS reads overlap instead of each
            // waiting in front of its four products
            bf16x8 xf[KS1];
            for (int ks = 0; ks < KS1; ++ks) xf[ks] = *reinterpret_cast<const bf16x8 *>(&xs[rr * LK + ks * 32 + g * 8]);
            for (int ks = 0; ks < KS1; ++ks)
                for (int ct = 0; ct < 4; ++ct) acc[ct] = mfma(wf1[ct][ks], xf[ks], acc[ct]);
            for (int ct = 0; ct < 4; ++ct) {
                const int col = 64 * cg + 16 * ct + 4 * g;
                const float4 bias = *reinterpret_cast<const float4 *>(&b1s[col]);
                bf16x4 v;
                v[0] = (bf16)fmaxf(acc[ct][0] + bias.x, 0.0f);
                v[1] = (bf16)fmaxf(acc[ct][1] + bias.y, 0.0f);
                v[2] = (bf16)fmaxf(acc[ct][2] + bias.z, 0.0f);
                v[3] = (bf16)fmaxf(acc[ct][3] + bias.w, 0.0f);
                *reinterpret_cast<bf16x4 *>(&as[rr * LH + col]) = v;
            }
        }
        __syncthreads();
        if (a.asave) {
            for (int q = tid; q < TR * (H / 8); q += 256) {
                const int rr = q / (H / 8), ch = q % (H / 8);
                const uint32_t row = r0 + rr;
                if (row < a.rows)
                    *reinterpret_cast<bf16x8 *>(&a.asave[size_t(row) * H + ch * 8]) =
                        *reinterpret_cast<const bf16x8 *>(&as[rr * LH + ch * 8]);
            }
        }
        if (MODE == M_DEC) {
            // 16 outputs: wave w takes rows 16w..16w+15
            const int rr = 16 * w + c;
            f32x4 acc2 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < KS2; ++ks)
                acc2 = mfma(wf2[0][ks], *reinterpret_cast<const bf16x8 *>(&as[rr * LH + ks * 32 + g * 8]), acc2);
            const float4 bias = *reinterpret_cast<const float4 *>(&b2s[4 * g]);
            yv = make_float4(acc2[0] + bias.x, acc2[1] + bias.y, acc2[2] + bias.z, acc2[3] + bias.w);
        } else {
            f32x4 acc2[2][4];
            bf16x8 af[2][KS2];
            for (int rt = 0; rt < 2; ++rt)
                for (int ks = 0; ks < KS2; ++ks)
                    af[rt][ks] = *reinterpret_cast<const bf16x8 *>(&as[(32 * rg + 16 * rt + c) * LH + ks * 32 + g * 8]);
            for (int rt = 0; rt < 2; ++rt) {
                for (int ct = 0; ct < 4; ++ct) acc2[rt][ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int ks = 0; ks < KS2; ++ks)
                    for (int ct = 0; ct < 4; ++ct) acc2[rt][ct] = mfma(wf2[ct][ks], af[rt][ks], acc2[rt][ct]);
            }
            __syncthreads();  // the partner wave of this row group has read `as`
            for (int rt = 0; rt < 2; ++rt) {
                const int rr = 32 * rg + 16 * rt + c;
                for (int ct = 0; ct < 4; ++ct) {
                    const int col = 64 * cg + 16 * ct + 4 * g;
                    const float4 bias = *reinterpret_cast<const float4 *>(&b2s[col]);
                    float v0 = acc2[rt][ct][0] + bias.x, v1 = acc2[rt][ct][1] + bias.y, v2 = acc2[rt][ct][2] + bias.z,
                          v3 = acc2[rt][ct][3] + bias.w;
                    bf16x4 o;
                    o[0] = (bf16)v0; o[1] = (bf16)v1; o[2] = (bf16)v2; o[3] = (bf16)v3;
                    *reinterpret_cast<bf16x4 *>(&as[rr * LH + col]) = o;
                }
            }
            __syncthreads();
            for (int u = 0; u < OCH; ++u) {
                const int q = tid + 256 * u, r2 = q / (H / 8), ch = q % (H / 8);
                ov[u] = *reinterpret_cast<const bf16x8 *>(&as[r2 * LH + ch * 8]);
            }
        }
    }
    if (blockIdx.x < a.ntiles) store_out(tp);
}

// ---------------------------------------------------------------- message-passing round
// h' = h + W2 relu(W1 [h, m] + b1) + b2 with m[i] = mean over the neighbours j of h[j],
// in one persistent launch over tiles of 64 rows.  The mean is formed while the tile's
// input is staged (FP32 sum in neighbour order, times 1/deg, rounded to BF16), so it is
// never written to or read back from HBM.
//
// 512 threads, 8 waves: wave w owns the 32 output columns 32*(w&3).. of both layers and
// the 32 rows 32*(w>>2).. of the tile, so its weight fragments (96 registers) leave room
// for two waves per SIMD.  Thread q stages the 16-byte column chunk q&15 of rows q>>4 and
// (q>>4)+32.  The neighbour gather runs three tiles ahead in three stages, one per
// tile: the CSR range, then the neighbour ids, then the rows themselves, so every load
// depends only on data that arrived during an earlier tile's compute.  Up to DMAX (4 or 6,
// the smallest that covers the graph's largest degree) neighbours per row are prefetched;
// any beyond are read when the tile is staged.
// The buffers the kernel writes (hout, xsave, asave) hold whole tiles (padded_rows), so
// every store is unconditional.  Two barriers per tile: after staging, and between the
// layers; the outputs leave from the layer-2 accumulators.
struct RoundFwd {
    uint32_t rows, ntiles;
    const bf16 *h;               // rows x H
    const uint32_t *off, *nbr;   // neighbour lists (CSR)
    const bf16 *w1t;
    const float *b1;
    const bf16 *w2t;
    const float *b2;
    bf16 *hout;
    bf16 *xsave, *asave;         // training saves of [h, m] and of the hidden activation
};
constexpr int RND_THREADS = 512;
constexpr int RND_U = TR * (H / 8) / RND_THREADS;  // staged (row, chunk) pairs per thread
constexpr size_t round_lds() { return size_t(2) * (TR * (2 * H + 8) + TR * (H + 8)); }

template <bool SAVE, int DMAX>
__global__ void __launch_bounds__(RND_THREADS) k_round_fwd(RoundFwd a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int K1 = 2 * H, LK = K1 + 8, LH = H + 8, KS1 = K1 / 32, KS2 = H / 32;
    bf16 *xs = reinterpret_cast<bf16 *>(smem);  // TR x LK: [h, m]
    bf16 *as = xs + TR * LK;                    // TR x LH: hidden, then output
    __shared__ __align__(16) float b1s[H], b2s[H];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    const int cg = w & 3, rg = w >> 2;
    bf16x8 wf1[2][KS1], wf2[2][KS2];
    for (int ct = 0; ct < 2; ++ct) {
        const int n = 32 * cg + 16 * ct + c;
        for (int ks = 0; ks < KS1; ++ks)
            wf1[ct][ks] = *reinterpret_cast<const bf16x8 *>(&a.w1t[n * K1 + ks * 32 + g * 8]);
        for (int ks = 0; ks < KS2; ++ks)
            wf2[ct][ks] = *reinterpret_cast<const bf16x8 *>(&a.w2t[n * H + ks * 32 + g * 8]);
    }
    if (tid < H) {
        b1s[tid] = a.b1[tid];
        b2s[tid] = a.b2[tid];
    }
    const int qc = tid & 15, qr = tid >> 4;
    const uint32_t G = gridDim.x;
    // stage A: CSR range; stage B: neighbour ids; stage C: own row chunk and neighbour chunks.
    // Every load is issued unconditionally from an address that is always valid (row 0 for a
    // row past the end, the row itself for a neighbour slot past its degree), and a stage
    // only selects on the previous stage's values once they have arrived, so no load sits
    // in a branch and the compiler waits only for the data it is about to use.
    uint32_t sa_b[RND_U], sa_e[RND_U], sa_ok = 0;
    uint32_t sb_id[RND_U][DMAX], sb_b[RND_U], sb_n[RND_U];
    uint32_t sc_b[RND_U], sc_n[RND_U], sc_ok = 0;
    bf16x8 sc_self[RND_U], sc_nb[RND_U][DMAX];
    auto stage_a = [&](uint32_t t) {
        sa_ok = 0;
        for (int u = 0; u < RND_U; ++u) {
            const uint32_t row = t * TR + qr + 32 * u;
            const bool ok = t < a.ntiles && row < a.rows;
            const uint32_t r = ok ? row : 0u;
            sa_b[u] = a.off[r];
            sa_e[u] = a.off[r + 1];
            sa_ok |= uint32_t(ok) << u;
        }
    };
    auto stage_b = [&]() {
        for (int u = 0; u < RND_U; ++u) {
            const bool ok = (sa_ok >> u) & 1u;
            const uint32_t b0 = ok ? sa_b[u] : 0u, n = ok ? sa_e[u] - sa_b[u] : 0u;
            sb_b[u] = b0;
            sb_n[u] = n;
            for (int k = 0; k < DMAX; ++k) sb_id[u][k] = a.nbr[uint32_t(k) < n ? b0 + k : 0u];
        }
    };
    auto stage_c = [&](uint32_t t) {
        sc_ok = 0;
        for (int u = 0; u < RND_U; ++u) {
            const uint32_t row = t * TR + qr + 32 * u;
            const bool ok = t < a.ntiles && row < a.rows;
            const uint32_t r = ok ? row : 0u;
            sc_ok |= uint32_t(ok) << u;
            sc_b[u] = sb_b[u];
            sc_n[u] = sb_n[u];
            sc_self[u] = *reinterpret_cast<const bf16x8 *>(&a.h[size_t(r) * H + qc * 8]);
            for (int k = 0; k < DMAX; ++k) {
                const uint32_t j = uint32_t(k) < sb_n[u] ? sb_id[u][k] : r;
                sc_nb[u][k] = *reinterpret_cast<const bf16x8 *>(&a.h[size_t(j) * H + qc * 8]);
            }
        }
    };
    uint32_t t = blockIdx.x;
    stage_a(t);
    stage_b();
    stage_c(t);
    stage_a(t + G);
    stage_b();
    stage_a(t + 2 * G);
    for (; t < a.ntiles; t += G) {
        const uint32_t r0 = t * TR;
        // the previous tile read `xs` last before its final barrier, so staging needs none
        for (int u = 0; u < RND_U; ++u) {
            const int rr = qr + 32 * u;
            float acc[8];
            for (int i = 0; i < 8; ++i) acc[i] = 0.0f;
            const uint32_t n = sc_n[u];
            for (int k = 0; k < DMAX; ++k)
                if (uint32_t(k) < n)
                    for (int i = 0; i < 8; ++i) acc[i] += (float)sc_nb[u][k][i];
            for (uint32_t k = DMAX; k < n; ++k) {
                const bf16x8 v = *reinterpret_cast<const bf16x8 *>(&a.h[size_t(a.nbr[sc_b[u] + k]) * H + qc * 8]);
                for (int i = 0; i < 8; ++i) acc[i] += (float)v[i];
            }
            if (n > 0) {
                const float inv = 1.0f / float(n);
                for (int i = 0; i < 8; ++i) acc[i] *= inv;
            }
            bf16x8 m;
            for (int i = 0; i < 8; ++i) m[i] = (bf16)acc[i];
            *reinterpret_cast<bf16x8 *>(&xs[rr * LK + qc * 8]) = (sc_ok >> u) & 1u ? sc_self[u] : zero8();
            *reinterpret_cast<bf16x8 *>(&xs[rr * LK + H + qc * 8]) = m;
        }
        stage_c(t + G);
        stage_b();
        stage_a(t + 3 * G);
        __syncthreads();
        if (SAVE) {
            for (int u = 0; u < TR * (K1 / 8) / RND_THREADS; ++u) {
                const int q = tid + RND_THREADS * u, rr = q / (K1 / 8), ch = q % (K1 / 8);
                *reinterpret_cast<bf16x8 *>(&a.xsave[size_t(r0 + rr) * K1 + ch * 8]) =
                    *reinterpret_cast<const bf16x8 *>(&xs[rr * LK + ch * 8]);
            }
        }
        // layer 1: hidden units 32cg + 16ct + 4g + j of rows 32rg + 16rt + c.  The residual
        // inputs of the lane's outputs are read here too: after the next barrier no wave reads
        // `xs` again, so the next tile may be staged over it without another barrier.
        bf16x4 hs[2][2];
        for (int rt = 0; rt < 2; ++rt) {
            const int rr = 32 * rg + 16 * rt + c;
            f32x4 acc[2];
            for (int ct = 0; ct < 2; ++ct) acc[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            bf16x8 xf[KS1];
            for (int ks = 0; ks < KS1; ++ks) xf[ks] = *reinterpret_cast<const bf16x8 *>(&xs[rr * LK + ks * 32 + g * 8]);
            for (int ks = 0; ks < KS1; ++ks)
                for (int ct = 0; ct < 2; ++ct) acc[ct] = mfma(wf1[ct][ks], xf[ks], acc[ct]);
            for (int ct = 0; ct < 2; ++ct) {
                const int col = 32 * cg + 16 * ct + 4 * g;
                const float4 bias = *reinterpret_cast<const float4 *>(&b1s[col]);
                bf16x4 v;
                v[0] = (bf16)fmaxf(acc[ct][0] + bias.x, 0.0f);
                v[1] = (bf16)fmaxf(acc[ct][1] + bias.y, 0.0f);
                v[2] = (bf16)fmaxf(acc[ct][2] + bias.z, 0.0f);
                v[3] = (bf16)fmaxf(acc[ct][3] + bias.w, 0.0f);
                *reinterpret_cast<bf16x4 *>(&as[rr * LH + col]) = v;
                hs[rt][ct] = *reinterpret_cast<const bf16x4 *>(&xs[rr * LK + col]);
            }
        }
        __syncthreads();
        if (SAVE) {
            for (int u = 0; u < TR * (H / 8) / RND_THREADS; ++u) {
                const int q = tid + RND_THREADS * u, rr = q / (H / 8), ch = q % (H / 8);
                *reinterpret_cast<bf16x8 *>(&a.asave[size_t(r0 + rr) * H + ch * 8]) =
                    *reinterpret_cast<const bf16x8 *>(&as[rr * LH + ch * 8]);
            }
        }
        // layer 2 and the residual
        f32x4 acc2[2][2];
        for (int rt = 0; rt < 2; ++rt) {
            bf16x8 af[KS2];
            for (int ks = 0; ks < KS2; ++ks)
                af[ks] = *reinterpret_cast<const bf16x8 *>(&as[(32 * rg + 16 * rt + c) * LH + ks * 32 + g * 8]);
            for (int ct = 0; ct < 2; ++ct) acc2[rt][ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < KS2; ++ks)
                for (int ct = 0; ct < 2; ++ct) acc2[rt][ct] = mfma(wf2[ct][ks], af[ks], acc2[rt][ct]);
        }
        // layer-2 outputs go straight from the accumulators to HBM (the rows are padded to whole
        // tiles); `as` is next written by the next tile's layer 1, after the staging barrier
        for (int rt = 0; rt < 2; ++rt) {
            const int rr = 32 * rg + 16 * rt + c;
            for (int ct = 0; ct < 2; ++ct) {
                const int col = 32 * cg + 16 * ct + 4 * g;
                const float4 bias = *reinterpret_cast<const float4 *>(&b2s[col]);
                bf16x4 o;
                o[0] = (bf16)((float)hs[rt][ct][0] + (acc2[rt][ct][0] + bias.x));
                o[1] = (bf16)((float)hs[rt][ct][1] + (acc2[rt][ct][1] + bias.y));
                o[2] = (bf16)((float)hs[rt][ct][2] + (acc2[rt][ct][2] + bias.z));
                o[3] = (bf16)((float)hs[rt][ct][3] + (acc2[rt][ct][3] + bias.w));
                *reinterpret_cast<bf16x4 *>(&a.hout[size_t(r0 + rr) * H + col]) = o;
            }
        }
    }
}

// ---------------------------------------------------------------- backward of one block
struct MlpBwd {
    uint32_t rows, ntiles;
    const float *dy;     // rows x N2
    const bf16 *asave;   // rows x H
    const bf16 *w2;      // H x N2 row-major
    const bf16 *w1;      // K1 x H row-major
    bf16 *dyb;           // rows x N2
    bf16 *dzb;           // rows x H
    float *dx;           // rows x K1 (nullptr: not needed)
};

template <int K1, int N2>
constexpr size_t bwd_lds() {
    return size_t(2) * (H * ((N2 < 32 ? 32 : N2) + 8) + K1 * (H + 8) + TR * ((N2 < 32 ? 32 : N2) + 8) + TR * (H + 8));
}

// dA = dY W2^T ; dZ = dA * [A > 0] ; dX = dZ W1^T.  Operands rounded to BF16 where they enter MFMA.
template <int K1, int N2>
__global__ void __launch_bounds__(256) k_mlp_bwd(MlpBwd a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NP = N2 < 32 ? 32 : N2, LY = NP + 8, LH = H + 8;
    bf16 *w2s = reinterpret_cast<bf16 *>(smem);  // H x LY   (row = hidden unit, col = output)
    bf16 *w1s = w2s + H * LY;                      // K1 x LH  (row = input, col = hidden unit)
    bf16 *ys = w1s + K1 * LH;                      // TR x LY
    bf16 *zs = ys + TR * LY;                       // TR x LH
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    for (int q = tid; q < H * NP / 8; q += 256) {
        const int n = q / (NP / 8), k8 = q % (NP / 8);
        bf16x8 v = zero8();
        if (k8 * 8 < N2) v = *reinterpret_cast<const bf16x8 *>(&a.w2[n * N2 + k8 * 8]);
        *reinterpret_cast<bf16x8 *>(&w2s[n * LY + k8 * 8]) = v;
    }
    if (a.dx) {
        for (int q = tid; q < K1 * H / 8; q += 256) {
            const int n = q / (H / 8), k8 = q % (H / 8);
            *reinterpret_cast<bf16x8 *>(&w1s[n * LH + k8 * 8]) = *reinterpret_cast<const bf16x8 *>(&a.w1[n * H + k8 * 8]);
        }
    }
    for (uint32_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const uint32_t r0 = t * TR;
        __syncthreads();
        for (int q = tid; q < TR * (NP / 4); q += 256) {
            const int rr = q / (NP / 4), c4 = (q % (NP / 4)) * 4;
            const uint32_t row = r0 + rr;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (row < a.rows && c4 < N2) v = *reinterpret_cast<const float4 *>(&a.dy[size_t(row) * N2 + c4]);
            bf16 *dst = &ys[rr * LY + c4];
            dst[0] = (bf16)v.x; dst[1] = (bf16)v.y; dst[2] = (bf16)v.z; dst[3] = (bf16)v.w;
            if (row < a.rows && c4 < N2) {
                bf16 *o = &a.dyb[size_t(row) * N2 + c4];
                o[0] = dst[0]; o[1] = dst[1]; o[2] = dst[2]; o[3] = dst[3];
            }
        }
        for (int q = tid; q < TR * (H / 8); q += 256) {
            const int rr = q / (H / 8), ch = q % (H / 8);
            const uint32_t row = r0 + rr;
            bf16x8 v = zero8();
            if (row < a.rows) v = *reinterpret_cast<const bf16x8 *>(&a.asave[size_t(row) * H + ch * 8]);
            *reinterpret_cast<bf16x8 *>(&zs[rr * LH + ch * 8]) = v;
        }
        __syncthreads();
        f32x4 acc[H / 16];
        for (int ct = 0; ct < H / 16; ++ct) acc[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < NP / 32; ++ks) {
            const bf16x8 af = *reinterpret_cast<const bf16x8 *>(&ys[(16 * w + c) * LY + ks * 32 + g * 8]);
            for (int ct = 0; ct < H / 16; ++ct) {
                const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(&w2s[(ct * 16 + c) * LY + ks * 32 + g * 8]);
                acc[ct] = mfma(af, bfr, acc[ct]);
            }
        }
        for (int ct = 0; ct < H / 16; ++ct)
            for (int j = 0; j < 4; ++j) {
                bf16 &z = zs[(16 * w + g * 4 + j) * LH + ct * 16 + c];
                z = (float)z > 0.0f ? (bf16)acc[ct][j] : (bf16)0.0f;
            }
        __syncthreads();
        for (int q = tid; q < TR * (H / 8); q += 256) {
            const int rr = q / (H / 8), ch = q % (H / 8);
            const uint32_t row = r0 + rr;
            if (row < a.rows)
                *reinterpret_cast<bf16x8 *>(&a.dzb[size_t(row) * H + ch * 8]) =
                    *reinterpret_cast<const bf16x8 *>(&zs[rr * LH + ch * 8]);
        }
        if (a.dx) {
            f32x4 acc3[K1 / 16];
            for (int ct = 0; ct < K1 / 16; ++ct) acc3[ct] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < H / 32; ++ks) {
                const bf16x8 af = *reinterpret_cast<const bf16x8 *>(&zs[(16 * w + c) * LH + ks * 32 + g * 8]);
                for (int ct = 0; ct < K1 / 16; ++ct) {
                    const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(&w1s[(ct * 16 + c) * LH + ks * 32 + g * 8]);
                    acc3[ct] = mfma(af, bfr, acc3[ct]);
                }
            }
            for (int ct = 0; ct < K1 / 16; ++ct)
                for (int j = 0; j < 4; ++j) {
                    const uint32_t row = r0 + 16 * w + g * 4 + j;
                    if (row < a.rows) a.dx[size_t(row) * K1 + ct * 16 + c] = acc3[ct][j];
                }
        }
    }
}

// ---------------------------------------------------------------- weight gradients
// partial[block] = X^T dZ over the block's rows (K x N) followed by the column sums of dZ (N).
struct GemmTN {
    uint32_t rows, rows_per_block;
    const bf16 *x;   // rows x K
    const bf16 *dz;  // rows x N
    float *partial;  // blocks x (K*N + N)
};

template <int K, int N>
__global__ void __launch_bounds__(256) k_gemm_tn(GemmTN a) {
    constexpr int CH = 32, LC = CH + 8;
    __shared__ __align__(16) bf16 xt[K * LC];
    __shared__ __align__(16) bf16 zt[N * LC];
    constexpr int NT = N / 16, TILES = (K / 16) * NT, TPW = (TILES + 3) / 4;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
    f32x4 acc[TPW];
    for (int u = 0; u < TPW; ++u) acc[u] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;
    const uint32_t rb = blockIdx.x * a.rows_per_block;
    const uint32_t re = std::min(a.rows, rb + a.rows_per_block);
    for (uint32_t r0 = rb; r0 < re; r0 += CH) {
        __syncthreads();
        for (int q = tid; q < CH * (K / 8); q += 256) {
            const int rr = q / (K / 8), k8 = q % (K / 8);
            const uint32_t row = r0 + rr;
            bf16x8 v = zero8();
            if (row < re) v = *reinterpret_cast<const bf16x8 *>(&a.x[size_t(row) * K + k8 * 8]);
            for (int i = 0; i < 8; ++i) xt[(k8 * 8 + i) * LC + rr] = v[i];
        }
        for (int q = tid; q < CH * (N / 8); q += 256) {
            const int rr = q / (N / 8), k8 = q % (N / 8);
            const uint32_t row = r0 + rr;
            bf16x8 v = zero8();
            if (row < re) v = *reinterpret_cast<const bf16x8 *>(&a.dz[size_t(row) * N + k8 * 8]);
            for (int i = 0; i < 8; ++i) zt[(k8 * 8 + i) * LC + rr] = v[i];
        }
        __syncthreads();
        if (tid < N)
            for (int rr = 0; rr < CH; ++rr) bsum += (float)zt[tid * LC + rr];
        for (int u = 0; u < TPW; ++u) {
            const int tt = w + 4 * u;
            if (tt < TILES) {
                const int mt = tt / NT, nt = tt % NT;
                const bf16x8 af = *reinterpret_cast<const bf16x8 *>(&xt[(mt * 16 + c) * LC + g * 8]);
                const bf16x8 bfr = *reinterpret_cast<const bf16x8 *>(&zt[(nt * 16 + c) * LC + g * 8]);
                acc[u] = mfma(af, bfr, acc[u]);
            }
        }
    }
    float *out = a.partial + size_t(blockIdx.x) * (K * N + N);
    for (int u = 0; u < TPW; ++u) {
        const int tt = w + 4 * u;
        if (tt < TILES) {
            const int mt = tt / NT, nt = tt % NT;
            for (int j = 0; j < 4; ++j) out[(mt * 16 + g * 4 + j) * N + nt * 16 + c] = acc[u][j];
        }
    }
    if (tid < N) out[K * N + tid] = bsum;
}

struct LayerTable {
    uint32_t n;
    Layer L[MAX_LAYERS];
};

// grad[i] = sum over blocks of the partials of parameter i (block order).
__global__ void k_reduce_grad(LayerTable lt, const float *__restrict__ partial, uint32_t blocks, uint32_t nparams,
                              float *__restrict__ grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nparams) return;
    uint32_t li = 0;
    while (li + 1 < lt.n && i >= lt.L[li + 1].off) ++li;
    const Layer L = lt.L[li];
    const uint32_t size = L.K * L.N + L.N, p = i - L.off;
    const float *base = partial + size_t(blocks) * L.off + p;
    float s = 0.0f;
    for (uint32_t b = 0; b < blocks; ++b) s += base[size_t(b) * size];
    grad[i] = s;
}

// Graph transpose of the mean aggregation: dh = dh_next + dx[:, :H] + sum_j dx[j, H:] / deg(j).
__global__ void k_graph_bwd(uint32_t rows, const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                            const float *__restrict__ dhn, const float *__restrict__ dx, float *__restrict__ dh) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows * (H / 4)) return;
    const uint32_t row = q / (H / 4), c4 = (q % (H / 4)) * 4;
    const float4 a = *reinterpret_cast<const float4 *>(&dhn[size_t(row) * H + c4]);
    const float4 b = *reinterpret_cast<const float4 *>(&dx[size_t(row) * 2 * H + c4]);
    float4 v = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
    for (uint32_t k = off[row]; k < off[row + 1]; ++k) {
        const uint32_t j = nbr[k];
        const float inv = 1.0f / float(off[j + 1] - off[j]);
        const float4 u = *reinterpret_cast<const float4 *>(&dx[size_t(j) * 2 * H + H + c4]);
        v.x += u.x * inv; v.y += u.y * inv; v.z += u.z * inv; v.w += u.w * inv;
    }
    *reinterpret_cast<float4 *>(&dh[size_t(row) * H + c4]) = v;
}

// Loss = sum over free vertices of |y[0:3] - t|^2 / n_free; dy = dLoss/dy.
__global__ void __launch_bounds__(256) k_loss(uint32_t rows, const float *__restrict__ y, const float4 *__restrict__ t,
                                              float inv_nfree, float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double sh[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double lsum = 0.0;
    if (i < rows) {
        const float4 tt = t[i];
        const bool m = tt.w > 0.0f;
        const float tv[3] = {tt.x, tt.y, tt.z};
        for (int c = 0; c < O; ++c) {
            float d = 0.0f;
            if (c < 3 && m) d = y[size_t(i) * O + c] - tv[c];
            dy[size_t(i) * O + c] = 2.0f * d * inv_nfree;
            lsum += double(d) * double(d);
        }
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + (sh[2] + sh[3])) * double(inv_nfree);
}

__device__ __forceinline__ void write_bf16_copies(const LayerTable &lt, uint32_t i, float th, bf16 *w, bf16 *wt) {
    uint32_t li = 0;
    while (li + 1 < lt.n && i >= lt.L[li + 1].off) ++li;
    const Layer L = lt.L[li];
    const uint32_t p = i - L.off;
    if (p < L.K * L.N) {
        const uint32_t k = p / L.N, n = p % L.N;
        w[i] = (bf16)th;
        wt[L.off + n * L.K + k] = (bf16)th;
    }
}

__global__ void k_adam(LayerTable lt, AdamArgs a, float *__restrict__ theta, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ g, bf16 *__restrict__ w, bf16 *__restrict__ wt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float gi = g[i];
    const float mi = a.beta1 * m[i] + (1.0f - a.beta1) * gi;
    const float vi = a.beta2 * v[i] + (1.0f - a.beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float th = theta[i] - a.lr * (mi / a.bc1) / (sqrtf(vi / a.bc2) + a.eps);
    theta[i] = th;
    write_bf16_copies(lt, i, th, w, wt);
}

__global__ void k_refresh(LayerTable lt, uint32_t n, const float *__restrict__ theta, bf16 *__restrict__ w,
                          bf16 *__restrict__ wt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) write_bf16_copies(lt, i, theta[i], w, wt);
}

// ---------------------------------------------------------------- host side
namespace {
inline uint32_t cdiv(size_t n, uint32_t b) { return uint32_t((n + b - 1) / b); }
LayerTable table(const NetLayout &nl) {
    LayerTable lt;
    lt.n = nl.nlayers;
    for (uint32_t k = 0; k < nl.nlayers; ++k) lt.L[k] = nl.L[k];
    return lt;
}
constexpr uint32_t PERSIST = 256;  // one workgroup per CU for the LDS-heavy layers

template <int K1, int N2, int MODE>
hipError_t fwd_block(hipStream_t st, const NetDev &nd, const Layer &l1, const Layer &l2, uint32_t rows,
                     const bf16 *xin, bf16 *hout, float *yout, bf16 *xsave, bf16 *asave) {
    MlpFwd a;
    a.rows = rows;
    a.ntiles = cdiv(rows, TR);
    a.xin = xin;
    a.w1t = nd.wt + l1.off;
    a.b1 = nd.theta + l1.off + l1.K * l1.N;
    a.w2t = nd.wt + l2.off;
    a.b2 = nd.theta + l2.off + l2.K * l2.N;
    a.hout = hout;
    a.yout = yout;
    a.xsave = xsave;
    a.asave = asave;
    const size_t lds = fwd_lds<K1, N2>();
    // two workgroups per CU of the persistent grid
    hipLaunchKernelGGL((k_mlp_fwd<K1, N2, MODE>), dim3(std::min(a.ntiles, PERSIST * 2)), dim3(256), lds, st, a);
    return hipGetLastError();
}

hipError_t round_fwd(hipStream_t st, const NetDev &nd, const Layer &l1, const Layer &l2, const Graph &g,
                     uint32_t rows, const bf16 *h, bf16 *hout, bf16 *xsave, bf16 *asave) {
    RoundFwd a;
    a.rows = rows;
    a.ntiles = cdiv(rows, TR);
    a.h = h;
    a.off = g.off;
    a.nbr = g.nbr;
    a.w1t = nd.wt + l1.off;
    a.b1 = nd.theta + l1.off + l1.K * l1.N;
    a.w2t = nd.wt + l2.off;
    a.b2 = nd.theta + l2.off + l2.K * l2.N;
    a.hout = hout;
    a.xsave = xsave;
    a.asave = asave;
    // one workgroup (8 waves) per CU: the weight fragments take about half of each wave's registers
    const dim3 grid(std::min(a.ntiles, PERSIST)), block(RND_THREADS);
    const bool save = xsave && asave, d4 = g.max_degree <= 4;
    if (save && d4) hipLaunchKernelGGL((k_round_fwd<true, 4>), grid, block, round_lds(), st, a);
    else if (save) hipLaunchKernelGGL((k_round_fwd<true, 6>), grid, block, round_lds(), st, a);
    else if (d4) hipLaunchKernelGGL((k_round_fwd<false, 4>), grid, block, round_lds(), st, a);
    else hipLaunchKernelGGL((k_round_fwd<false, 6>), grid, block, round_lds(), st, a);
    return hipGetLastError();
}

template <int K1, int N2>
hipError_t bwd_block(hipStream_t st, const NetDev &nd, const Layer &l1, const Layer &l2, uint32_t rows,
                     const float *dy, const bf16 *asave, bf16 *dyb, bf16 *dzb, float *dx) {
    MlpBwd a;
    a.rows = rows;
    a.ntiles = cdiv(rows, TR);
    a.dy = dy;
    a.asave = asave;
    a.w2 = nd.w + l2.off;
    a.w1 = nd.w + l1.off;
    a.dyb = dyb;
    a.dzb = dzb;
    a.dx = dx;
    const size_t lds = bwd_lds<K1, N2>();
    hipLaunchKernelGGL((k_mlp_bwd<K1, N2>), dim3(std::min(a.ntiles, PERSIST)), dim3(256), lds, st, a);
    return hipGetLastError();
}

template <int K, int N>
hipError_t gemm(hipStream_t st, const Layer &L, uint32_t rows, uint32_t blocks, const bf16 *x, const bf16 *dz,
                float *partial) {
    GemmTN a;
    a.rows = rows;
    a.rows_per_block = ((cdiv(rows, blocks) + 31) / 32) * 32;
    a.x = x;
    a.dz = dz;
    a.partial = partial + size_t(blocks) * L.off;
    hipLaunchKernelGGL((k_gemm_tn<K, N>), dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename Kern>
hipError_t allow_lds(Kern k, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(bytes));
}
}  // namespace

hipError_t set_lds_limits() {
    hipError_t e;
    if ((e = allow_lds(&k_mlp_fwd<F, H, M_ENC>, fwd_lds<F, H>())) != hipSuccess) return e;
    if ((e = allow_lds(&k_round_fwd<false, 4>, round_lds())) != hipSuccess) return e;
    if ((e = allow_lds(&k_round_fwd<false, 6>, round_lds())) != hipSuccess) return e;
    if ((e = allow_lds(&k_round_fwd<true, 4>, round_lds())) != hipSuccess) return e;
    if ((e = allow_lds(&k_round_fwd<true, 6>, round_lds())) != hipSuccess) return e;
    if ((e = allow_lds(&k_mlp_fwd<H, O, M_DEC>, fwd_lds<H, O>())) != hipSuccess) return e;
    if ((e = allow_lds(&k_mlp_bwd<H, O>, bwd_lds<H, O>())) != hipSuccess) return e;
    if ((e = allow_lds(&k_mlp_bwd<2 * H, H>, bwd_lds<2 * H, H>())) != hipSuccess) return e;
    return allow_lds(&k_mlp_bwd<F, H>, bwd_lds<F, H>());
}

uint32_t gemm_blocks(uint32_t rows) { return std::max<uint32_t>(1, std::min<uint32_t>(256, cdiv(rows, 256))); }
size_t partial_floats(const NetLayout &nl, uint32_t blocks) { return size_t(blocks) * nl.nparams; }
uint32_t loss_blocks(uint32_t rows) { return std::max<uint32_t>(1, cdiv(rows, 256)); }

hipError_t launch_features(hipStream_t st, const FeatArgs &a, bf16 *feat) {
    if (!a.nv) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3(cdiv(a.nv, 256)), dim3(256), 0, st, a, feat);
    return hipGetLastError();
}
hipError_t launch_copy_pred(hipStream_t st, const float4 *P, float4 *Ppred, uint32_t nv) {
    if (!nv) return hipSuccess;
    hipLaunchKernelGGL(k_copy_pred, dim3(cdiv(nv, 256)), dim3(256), 0, st, P, Ppred, nv);
    return hipGetLastError();
}
hipError_t launch_apply(hipStream_t st, float4 *P, float4 *Ppred, const float4 *D, const float *y, float s,
                        float beta, uint32_t nv) {
    if (!nv) return hipSuccess;
    hipLaunchKernelGGL(k_apply, dim3(cdiv(nv, 256)), dim3(256), 0, st, P, Ppred, D, y, s, beta, nv);
    return hipGetLastError();
}
hipError_t launch_update_prev(hipStream_t st, const float4 *P, const float4 *Ppred, float4 *D, float4 *T, float s,
                              float beta, uint32_t nv) {
    if (!nv) return hipSuccess;
    hipLaunchKernelGGL(k_update_prev, dim3(cdiv(nv, 256)), dim3(256), 0, st, P, Ppred, D, T, s, beta, nv);
    return hipGetLastError();
}

hipError_t forward(hipStream_t st, const NetLayout &nl, const NetDev &nd, const Graph &g, uint32_t rows,
                   const FwdBufs &b, hipEvent_t *ev) {
    if (!rows) return hipSuccess;
    hipError_t e;
    int k = 0;
    auto mark = [&]() -> hipError_t { return ev ? hipEventRecord(ev[k++], st) : hipSuccess; };
    if ((e = mark()) != hipSuccess) return e;
    if ((e = fwd_block<F, H, M_ENC>(st, nd, nl.L[nl.enc1()], nl.L[nl.enc2()], rows, b.feat, b.h[0], nullptr,
                                    nullptr, b.training ? b.a_enc : nullptr)) != hipSuccess) return e;
    if ((e = mark()) != hipSuccess) return e;
    for (uint32_t r = 0; r < nl.rounds; ++r) {
        if ((e = round_fwd(st, nd, nl.L[nl.r1(r)], nl.L[nl.r2(r)], g, rows, b.h[r], b.h[r + 1],
                           b.training ? b.x_r[r] : nullptr, b.training ? b.a_r[r] : nullptr)) != hipSuccess) return e;
        if ((e = mark()) != hipSuccess) return e;
    }
    if ((e = fwd_block<H, O, M_DEC>(st, nd, nl.L[nl.dec1()], nl.L[nl.dec2()], rows, b.h[nl.rounds], nullptr,
                                    b.y, nullptr, b.training ? b.a_dec : nullptr)) != hipSuccess) return e;
    return mark();
}

hipError_t loss_only(hipStream_t st, uint32_t rows, const float *y, const float4 *t, float inv_nfree, float *dy,
                     double *loss_part) {
    if (!rows) return hipSuccess;
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks(rows)), dim3(256), 0, st, rows, y, t, inv_nfree, dy, loss_part);
    return hipGetLastError();
}

hipError_t loss_and_backward(hipStream_t st, const NetLayout &nl, const NetDev &nd, const Graph &g, uint32_t rows,
                             const FwdBufs &f, const float4 *t, float inv_nfree, const BwdBufs &b) {
    if (!rows) return hipSuccess;
    hipError_t e;
    const uint32_t blocks = b.partial_blocks;
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks(rows)), dim3(256), 0, st, rows, f.y, t, inv_nfree, b.dy,
                       b.loss_part);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // decoder
    float *dh = b.dh0, *dh_other = b.dh1;
    if ((e = bwd_block<H, O>(st, nd, nl.L[nl.dec1()], nl.L[nl.dec2()], rows, b.dy, f.a_dec, b.dyb, b.dzb, dh)) !=
        hipSuccess) return e;
    if ((e = gemm<H, O>(st, nl.L[nl.dec2()], rows, blocks, f.a_dec, b.dyb, b.partial)) != hipSuccess) return e;
    if ((e = gemm<H, H>(st, nl.L[nl.dec1()], rows, blocks, f.h[nl.rounds], b.dzb, b.partial)) != hipSuccess) return e;
    // rounds, last to first
    for (uint32_t rr = nl.rounds; rr-- > 0;) {
        if ((e = bwd_block<2 * H, H>(st, nd, nl.L[nl.r1(rr)], nl.L[nl.r2(rr)], rows, dh, f.a_r[rr], b.dyb, b.dzb,
                                     b.dx)) != hipSuccess) return e;
        if ((e = gemm<H, H>(st, nl.L[nl.r2(rr)], rows, blocks, f.a_r[rr], b.dyb, b.partial)) != hipSuccess) return e;
        if ((e = gemm<2 * H, H>(st, nl.L[nl.r1(rr)], rows, blocks, f.x_r[rr], b.dzb, b.partial)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_graph_bwd, dim3(cdiv(size_t(rows) * (H / 4), 256)), dim3(256), 0, st, rows, g.off, g.nbr,
                           dh, b.dx, dh_other);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(dh, dh_other);
    }
    // encoder
    if ((e = bwd_block<F, H>(st, nd, nl.L[nl.enc1()], nl.L[nl.enc2()], rows, dh, f.a_enc, b.dyb, b.dzb, nullptr)) !=
        hipSuccess) return e;
    if ((e = gemm<H, H>(st, nl.L[nl.enc2()], rows, blocks, f.a_enc, b.dyb, b.partial)) != hipSuccess) return e;
    if ((e = gemm<F, H>(st, nl.L[nl.enc1()], rows, blocks, f.feat, b.dzb, b.partial)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_grad, dim3(cdiv(nl.nparams, 256)), dim3(256), 0, st, table(nl), b.partial, blocks,
                       nl.nparams, b.grad);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t st, const NetLayout &nl, const AdamArgs &a, float *theta, float *m, float *v,
                       const float *g, bf16 *w, bf16 *wt) {
    hipLaunchKernelGGL(k_adam, dim3(cdiv(a.n, 256)), dim3(256), 0, st, table(nl), a, theta, m, v, g, w, wt);
    return hipGetLastError();
}

hipError_t launch_refresh(hipStream_t st, const NetLayout &nl, const float *theta, bf16 *w, bf16 *wt) {
    hipLaunchKernelGGL(k_refresh, dim3(cdiv(nl.nparams, 256)), dim3(256), 0, st, table(nl), nl.nparams, theta, w, wt);
    return hipGetLastError();
}

}  // namespace ws
}  // namespace tc
