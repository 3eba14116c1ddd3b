// Mesh voxelisation by ray parity: kernel voxelise_mesh of the reference
// (Solver/Kernel/OpenCL/voxelise_mesh.cl:24-48, one launch per geometry entry
// at Solver/Kernel/voxelise_mesh.py:111-135).  Every interior cell casts a ray
// from its centre along the fixed direction D = normalize(1, 0.3713907,
// 0.2178311) and counts the triangles it hits (ray_hit, :5-22); an odd count
// marks the cell.
//
// The reference tests every cell against every triangle.  Here a block owns a
// tile of 8 x 4 x 8 cells and tests its cells only against the triangles a
// ray from the tile can hit, with the reference's arithmetic per test, so the
// mask is the brute-force mask bit for bit:
//   * mesh_prepare_kernel computes per triangle what ray_hit derives from the
//     triangle alone (a, e1, e2, p = D x e2, det, 1/det; the same operations,
//     so the same bits) and a culling box: the triangle's footprint projected
//     along D onto the plane x = 0, in coordinates s = y - x Dy/Dx and
//     t = z - x Dz/Dx (a ray keeps its (s, t)), and its largest x, y, z (D
//     is positive on every axis and a hit lies ahead of the centre).  The box
//     is widened by a rounding margin, 2e-5 (1 + E^2/|det|)(R + 2E) for edge
//     length E and coordinate magnitude R: ray_hit's float u and v can
//     differ from the exact barycentrics by about 10 eps (R E + |u| E^2)/|det|,
//     the margin is some 16 times that, converted to a distance.  A triangle
//     with |det| < 1e-7 never hits (ray_hit returns 0 before any other test)
//     and gets an empty box; one whose margin would not be small (E^2/|det|
//     near 1/eps) or that holds a non-finite coordinate gets an unbounded
//     box and is tested by every cell.
//   * mesh_parity_kernel walks the triangles in chunks of 256: each thread
//     tests one triangle's box against the tile's (s, t) range and lower
//     corner, the survivors are compacted into LDS, and every cell runs the
//     exact test against each survivor (LDS broadcast reads).
// The cull only ever removes a triangle whose exact test is 0 for every cell
// of the tile, so hit counts, and with them the parities, are unchanged.
#include <cfloat>

#include "common.hpp"

namespace clq {

struct MeshTri {
    float a[3], e1[3], e2[3], p[3], inv;
    float pad[3];
};
static_assert(sizeof(MeshTri) == 64, "MeshTri is 64 bytes");

struct MeshBox {
    float smin, smax, tmin, tmax, xmax, ymax, zmax, pad;
};

constexpr int kMeshTX = 8, kMeshTY = 4, kMeshTZ = 8;  // tile: 256 cells
constexpr int kMeshThreads = kMeshTX * kMeshTY * kMeshTZ;

__device__ __forceinline__ void mesh_cross(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ float mesh_dot(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__global__ __launch_bounds__(256) void mesh_prepare_kernel(const float *__restrict__ tris, int64_t count, float d0,
                                                           float d1, float d2, float rs, float rt, float reach,
                                                           MeshTri *__restrict__ out, MeshBox *__restrict__ box) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= count) return;
    const float d[3] = {d0, d1, d2};
    float v[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) v[c] = tris[9 * q + c];
    MeshTri m;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        m.a[a] = v[a];
        m.e1[a] = v[3 + a] - v[a];
        m.e2[a] = v[6 + a] - v[a];
        m.pad[a] = 0.0f;
    }
    mesh_cross(d, m.e2, m.p);
    const float det = mesh_dot(m.e1, m.p);
    m.inv = 1.0f / det;
    out[q] = m;
    MeshBox b;
    b.pad = 0.0f;
    bool finite = isfinite(det);
    float big = 0.0f;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        finite = finite && isfinite(v[c]);
        big = fmaxf(big, fabsf(v[c]));
    }
    if (finite && fabsf(det) < 1.0e-7f) {  // ray_hit returns 0 for every ray
        b.smin = b.tmin = FLT_MAX;
        b.smax = b.tmax = -FLT_MAX;
        b.xmax = b.ymax = b.zmax = -FLT_MAX;
        box[q] = b;
        return;
    }
    float e3[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) e3[a] = v[6 + a] - v[3 + a];
    const float E = sqrtf(fmaxf(fmaxf(mesh_dot(m.e1, m.e1), mesh_dot(m.e2, m.e2)), mesh_dot(e3, e3)));
    const float cond = E * E / fabsf(det);
    const float margin = 2.0e-5f * (1.0f + cond) * (fmaxf(reach, big) + 2.0f * E);
    if (!finite || !(cond < 1.0e5f) || !(margin < 1.0e30f)) {  // test it everywhere
        b.smin = b.tmin = -FLT_MAX;
        b.smax = b.tmax = FLT_MAX;
        b.xmax = b.ymax = b.zmax = FLT_MAX;
        box[q] = b;
        return;
    }
    float smin = FLT_MAX, smax = -FLT_MAX, tmin = FLT_MAX, tmax = -FLT_MAX;
    float xm = -FLT_MAX, ym = -FLT_MAX, zm = -FLT_MAX;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float x = v[3 * c], y = v[3 * c + 1], z = v[3 * c + 2];
        const float s = y - x * rs, t = z - x * rt;
        smin = fminf(smin, s);
        smax = fmaxf(smax, s);
        tmin = fminf(tmin, t);
        tmax = fmaxf(tmax, t);
        xm = fmaxf(xm, x);
        ym = fmaxf(ym, y);
        zm = fmaxf(zm, z);
    }
    // (s, t) distances are at most 1 / Dx < 2 times perpendicular ones
    const float ms = 2.0f * margin;
    b.smin = smin - ms;
    b.smax = smax + ms;
    b.tmin = tmin - ms;
    b.tmax = tmax + ms;
    b.xmax = xm + margin;
    b.ymax = ym + margin;
    b.zmax = zm + margin;
    box[q] = b;
}

// ray_hit (voxelise_mesh.cl:5-22) with the triangle's own terms precomputed
__device__ __forceinline__ int mesh_hit(const float o[3], const float d[3], const MeshTri &m) {
    float t[3], q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] = o[a] - m.a[a];
    const float u = mesh_dot(t, m.p) * m.inv;
    if (u < 0.0f || u > 1.0f) return 0;
    mesh_cross(t, m.e1, q);
    const float v = mesh_dot(d, q) * m.inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    return mesh_dot(m.e2, q) * m.inv > 1.0e-6f;
}

__global__ __launch_bounds__(kMeshThreads) void mesh_parity_kernel(const MeshTri *__restrict__ tris,
                                                                  const MeshBox *__restrict__ boxes, int64_t count,
                                                                  int nx, int ny, int nz, float delta, float o0,
                                                                  float o1, float o2, float d0, float d1, float d2,
                                                                  float rs, float rt, uint8_t *__restrict__ mask) {
    __shared__ MeshTri list[kMeshThreads];
    __shared__ int wave_count[kMeshThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ntz = (nz + kMeshTZ - 1) / kMeshTZ, nty = (ny + kMeshTY - 1) / kMeshTY;
    const int tile = blockIdx.x;
    const int ti = tile / (nty * ntz), tj = (tile / ntz) % nty, tk = tile % ntz;
    const int i0 = ti * kMeshTX, j0 = tj * kMeshTY, k0 = tk * kMeshTZ;
    const int i = i0 + tid / (kMeshTY * kMeshTZ), j = j0 + (tid / kMeshTZ) % kMeshTY, k = k0 + tid % kMeshTZ;
    const float d[3] = {d0, d1, d2};
    // the cell centre of voxelise_mesh.cl:39-40 (the fused multiply-add of
    // fp-contract=on)
    const float o[3] = {fmaf((float)i + 0.5f, delta, o0), fmaf((float)j + 0.5f, delta, o1),
                        fmaf((float)k + 0.5f, delta, o2)};
    // the tile's (s, t) range and lower corner from its corner cells in the grid
    const int i1 = min(i0 + kMeshTX - 1, nx - 1), j1 = min(j0 + kMeshTY - 1, ny - 1), k1 = min(k0 + kMeshTZ - 1, nz - 1);
    const float xl = fmaf((float)i0 + 0.5f, delta, o0), xh = fmaf((float)i1 + 0.5f, delta, o0);
    const float yl = fmaf((float)j0 + 0.5f, delta, o1), yh = fmaf((float)j1 + 0.5f, delta, o1);
    const float zl = fmaf((float)k0 + 0.5f, delta, o2), zh = fmaf((float)k1 + 0.5f, delta, o2);
    const float smin = yl - xh * rs, smax = yh - xl * rs, tmin = zl - xh * rt, tmax = zh - xl * rt;
    int parity = 0;
    for (int64_t base = 0; base < count; base += kMeshThreads) {
        const int64_t q = base + tid;
        bool keep = false;
        if (q < count) {
            const MeshBox b = boxes[q];
            keep = b.smax >= smin && b.smin <= smax && b.tmax >= tmin && b.tmin <= tmax && b.xmax >= xl &&
                   b.ymax >= yl && b.zmax >= zl;
        }
        const unsigned long long bal = __ballot(keep);
        if (lane == 0) wave_count[wave] = __popcll(bal);
        __syncthreads();
        int offset = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kMeshThreads / 64; ++w) {
            offset += w < wave ? wave_count[w] : 0;
            total += wave_count[w];
        }
        if (keep) list[offset + __popcll(bal & ((1ull << lane) - 1ull))] = tris[q];
        __syncthreads();
        for (int s = 0; s < total; ++s) parity ^= mesh_hit(o, d, list[s]);
        __syncthreads();  // the list is rewritten by the next chunk
    }
    if (i >= 1 && j >= 1 && k >= 1 && i < nx - 1 && j < ny - 1 && k < nz - 1 && parity)
        mask[gidx(i, j, k, ny, nz)] = 1;
}

// normalize((float3)(1.0f, 0.3713907f, 0.2178311f)) as v / sqrt(dot(v, v))
void mesh_direction(float d[3]) {
    const float v[3] = {1.0f, 0.3713907f, 0.2178311f};
    const float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int a = 0; a < 3; ++a) d[a] = v[a] / len;
}

size_t mesh_scratch_bytes(int64_t count) { return (size_t)count * (sizeof(MeshTri) + sizeof(MeshBox)); }

void launch_voxelise_mesh(const float *tris, int64_t count, void *scratch, int nx, int ny, int nz, float delta,
                          const float origin[3], uint8_t *mask, hipStream_t s) {
    if (count <= 0) return;
    float d[3];
    mesh_direction(d);
    const float rs = d[1] / d[0], rt = d[2] / d[0];
    // magnitude of the cell-centre coordinates (the rays' origins)
    float reach = 0.0f;
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
        reach = fmaxf(reach, fmaxf(fabsf(origin[a]), fabsf(origin[a] + (float)n[a] * delta)));
    MeshTri *prep = (MeshTri *)scratch;
    MeshBox *box = (MeshBox *)(prep + count);
    hipLaunchKernelGGL(mesh_prepare_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, tris, count, d[0],
                       d[1], d[2], rs, rt, reach, prep, box);
    const int tiles = ((nx + kMeshTX - 1) / kMeshTX) * ((ny + kMeshTY - 1) / kMeshTY) * ((nz + kMeshTZ - 1) / kMeshTZ);
    hipLaunchKernelGGL(mesh_parity_kernel, dim3((unsigned)tiles), dim3(kMeshThreads), 0, s, prep, box, count, nx, ny,
                       nz, delta, origin[0], origin[1], origin[2], d[0], d[1], d[2], rs, rt, mask);
}

}  // namespace clq
