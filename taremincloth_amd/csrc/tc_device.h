// Device-side helpers and layouts shared by the XPBD kernels (gfx950, wave64).
//
// Vertex state lives in HBM as three separate float4 streams (structure of
// arrays, one 16-byte vector per vertex and stream):
//   X[i] = (x, y, z, thickness)     committed position + contact radius
//   P[i] = (px, py, pz, inv_mass)   predicted / corrected position + weight
//   V[i] = (vx, vy, vz, 0)          velocity
// A constraint gathers P of its two vertices with one 16-byte load each, so
// position and inverse mass arrive in the same cache line.
//
// All arithmetic follows the operation order of the reference WGSL shaders
// (crates/cloth_core/src/shaders/*.wgsl); the library is compiled with
// -ffp-contract=off so that no FMA contraction changes the rounding.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tc {

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 xyz(const float4 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
// WGSL mix(a, b, t) = a * (1 - t) + b * t
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ f3 mix3(f3 a, f3 b, float t) { return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t)}; }
__device__ __forceinline__ f3 min3(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ f3 max3(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

// WGSL i32(f32) is a saturating truncation.
__device__ __forceinline__ int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483520.0f) return 2147483647;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
struct i3 { int32_t x, y, z; };
__device__ __forceinline__ i3 cell_of(f3 p, float cell) {
    return {f2i_sat(floorf(p.x / cell)), f2i_sat(floorf(p.y / cell)), f2i_sat(floorf(p.z / cell))};
}
// spatial_grid_sort.wgsl:25-31
__device__ __forceinline__ uint32_t hash_cell(i3 c, uint32_t table) {
    const uint32_t n = ((uint32_t)c.x * 73856093u) ^ ((uint32_t)c.y * 19349663u) ^ ((uint32_t)c.z * 83492791u);
    return n % table;
}

// atan2 by octant reduction: the ratio min/max of |x|, |y| in [0, 1], folded to |a| <= tan(pi/8)
// by atan(a) = pi/4 + atan((a - 1) / (a + 1)), then the odd Taylor series of atan to a^15
// (truncation below 2e-8 rad), and the octant restored.  The same operations on the host give the
// same bits, which a library atan2f does not promise.
__host__ __device__ __forceinline__ float atan2_det(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    if (!(mx > 0.0f)) return 0.0f;
    float a = mn / mx, off = 0.0f;
    if (a > 0.41421356f) { a = (a - 1.0f) / (a + 1.0f); off = 0.78539816f; }
    const float s = a * a;
    float p = -1.0f / 15.0f;
    p = p * s + 1.0f / 13.0f;
    p = p * s - 1.0f / 11.0f;
    p = p * s + 1.0f / 9.0f;
    p = p * s - 1.0f / 7.0f;
    p = p * s + 1.0f / 5.0f;
    p = p * s - 1.0f / 3.0f;
    p = p * s + 1.0f;
    float r = off + a * p;
    if (ay > ax) r = 1.5707964f - r;
    if (x < 0.0f) r = 3.1415927f - r;
    if (y < 0.0f) r = -r;
    return r;
}

// Collider primitive (GpuCollider, mesh.rs:65-74): type 0 sphere, 1 capsule, 2 plane.
struct Collider {
    uint32_t type;
    float friction, restitution, pad0;
    float ax, ay, az, radius;
    float bx, by, bz, pad1;
};
static_assert(sizeof(Collider) == 48, "GpuCollider is 48 bytes (mesh.rs:65-74)");

// Mesh collider triangle (GpuMeshTriangle, mesh.rs:78-87) + bounding sphere.
struct MeshTri {
    float p0x, p0y, p0z, friction;
    float p1x, p1y, p1z, thickness;
    float p2x, p2y, p2z, restitution;
    uint32_t flags; float pad0, pad1, pad2;
};
static_assert(sizeof(MeshTri) == 64, "the collider stages a triangle as 4 float4");

// Unique edge of the mesh collider (GpuColliderEdge, mesh.rs): end points, the largest
// thickness and mean friction of its triangles, the normalised sum of their normals, bit 0 of
// flags set when every one of them is single-sided.
struct ColliderEdge {
    float p0x, p0y, p0z, thickness;
    float p1x, p1y, p1z, friction;
    float nx, ny, nz;
    uint32_t flags;
};
static_assert(sizeof(ColliderEdge) == 48, "GpuColliderEdge is 48 bytes");
constexpr uint32_t COLLIDER_EDGE_TABLE = 32768;  // spatial_hash.rs:7 (DEFAULT_HASH_TABLE_SIZE)

// Self-collision accumulator (SelfCollisionAccum, simulation/types.rs:145-178)
struct Accum {
    int32_t dx, dy, dz; uint32_t count;
    int32_t pdx, pdy, pdz; uint32_t pcount;
    int32_t cpx, cpy, cpz, cnx, cny, cnz;
    uint32_t untangle, ee_count;
};
static_assert(sizeof(Accum) == 64, "SelfCollisionAccum is 64 bytes (simulation/types.rs:145-178)");

}  // namespace tc
