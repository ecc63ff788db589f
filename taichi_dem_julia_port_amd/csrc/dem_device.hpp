// dem_device.hpp -- device-side math of the DEM step for gfx950.
//
// Every floating-point expression keeps the operand order of the reference
// (lucifer1004/taichi_dem_julia_port, src/kernels.jl, src/utils.jl) and the library is built with
// -ffp-contract=off, so FP64 results are the reference's unfused IEEE results.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dem {

struct d3 { double x, y, z; };

__device__ __forceinline__ d3 mk3(double x, double y, double z) { return d3{x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ d3 operator/(d3 a, double s) { return mk3(a.x / s, a.y / s, a.z / s); }
// ((a1 b1 + a2 b2) + a3 b3)
__device__ __forceinline__ double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double norm(d3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// StaticArrays normalize: inv(norm(a)) * a
__device__ __forceinline__ d3 normalize(d3 a) { double s = 1.0 / norm(a); return s * a; }
__device__ __forceinline__ d3 xyz(double4 v) { return mk3(v.x, v.y, v.z); }

// 3x3 matrix, row-major.
struct m3 { double a[3][3]; };

__device__ __forceinline__ d3 mul(const m3& A, d3 u) {
    return mk3(A.a[0][0] * u.x + A.a[0][1] * u.y + A.a[0][2] * u.z,
               A.a[1][0] * u.x + A.a[1][1] * u.y + A.a[1][2] * u.z,
               A.a[2][0] * u.x + A.a[2][1] * u.y + A.a[2][2] * u.z);
}

// Inverse from cross products of the columns, scaled by the determinant.
__device__ __forceinline__ m3 inverse(const m3& A) {
    d3 c0 = mk3(A.a[0][0], A.a[1][0], A.a[2][0]);
    d3 c1 = mk3(A.a[0][1], A.a[1][1], A.a[2][1]);
    d3 c2 = mk3(A.a[0][2], A.a[1][2], A.a[2][2]);
    d3 r0 = cross(c1, c2);
    double det = dot(c0, r0);
    c0 = c0 / det;
    r0 = r0 / det;
    d3 r1 = cross(c2, c0);
    d3 r2 = cross(c0, c1);
    m3 B;
    B.a[0][0] = r0.x; B.a[0][1] = r0.y; B.a[0][2] = r0.z;
    B.a[1][0] = r1.x; B.a[1][1] = r1.y; B.a[1][2] = r1.z;
    B.a[2][0] = r2.x; B.a[2][1] = r2.y; B.a[2][2] = r2.z;
    return B;
}

// Rotation taking unit vector a onto the local x axis (Rodrigues form), kernels.jl:145-165.
__device__ __forceinline__ m3 contact_frame(d3 a, double tol) {
    const d3 ex = mk3(1.0, 0.0, 0.0);
    d3 v = cross(a, ex);
    double s = norm(v);
    double c = dot(a, ex);
    m3 R;
    if (s < tol) {
        double sg = c > 0.0 ? 1.0 : -1.0;
        R.a[0][0] = sg;  R.a[0][1] = 0.0; R.a[0][2] = 0.0;
        R.a[1][0] = 0.0; R.a[1][1] = 1.0; R.a[1][2] = 0.0;
        R.a[2][0] = 0.0; R.a[2][1] = 0.0; R.a[2][2] = sg;
        return R;
    }
    m3 K;
    K.a[0][0] = 0.0;  K.a[0][1] = -v.z; K.a[0][2] = v.y;
    K.a[1][0] = v.z;  K.a[1][1] = 0.0;  K.a[1][2] = -v.x;
    K.a[2][0] = -v.y; K.a[2][1] = v.x;  K.a[2][2] = 0.0;
    const double omc = 1.0 - c;
    const double s2 = s * s;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double k2 = K.a[r][0] * K.a[0][q] + K.a[r][1] * K.a[1][q] + K.a[r][2] * K.a[2][q];
            double id = (r == q) ? 1.0 : 0.0;
            R.a[r][q] = (id + K.a[r][q]) + (k2 * omc) / s2;
        }
    }
    return R;
}

// Rotation matrix of a (not necessarily unit) quaternion (w, x, y, z).
__device__ __forceinline__ m3 quat_matrix(double4 q) {
    const double w = q.x, x = q.y, y = q.z, z = q.w;
    const double nn = 1.0 / (w * w + x * x + y * y + z * z);
    m3 M;
    M.a[0][0] = 1.0 - 2.0 * (y * y + z * z) * nn;
    M.a[0][1] = 2.0 * (x * y - z * w) * nn;
    M.a[0][2] = 2.0 * (x * z + y * w) * nn;
    M.a[1][0] = 2.0 * (x * y + z * w) * nn;
    M.a[1][1] = 1.0 - 2.0 * (x * x + z * z) * nn;
    M.a[1][2] = 2.0 * (y * z - x * w) * nn;
    M.a[2][0] = 2.0 * (x * z - y * w) * nn;
    M.a[2][1] = 2.0 * (y * z + x * w) * nn;
    M.a[2][2] = 1.0 - 2.0 * (x * x + y * y) * nn;
    return M;
}

// ---- cell hashing (utils.jl:1-16) ------------------------------------------------------------

// Spread the low 21 bits of v to every third bit.
__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

struct HashGrid {
    double dmin[3];
    double cs;
    uint64_t H;     // buckets
    uint64_t Hmask; // H - 1 when H is a power of two, else 0
};

// 1-based cell (ceil), utils.jl:1.
__device__ __forceinline__ void cell_of(const HashGrid& g, d3 p, int32_t c[3]) {
    c[0] = (int32_t)ceil((p.x - g.dmin[0]) / g.cs);
    c[1] = (int32_t)ceil((p.y - g.dmin[1]) / g.cs);
    c[2] = (int32_t)ceil((p.z - g.dmin[2]) / g.cs);
}

// 0-based bucket of a 1-based cell: (M(c - 1) + 1) mod H  (utils.jl:2 minus one).
__device__ __forceinline__ uint32_t bucket_of(const HashGrid& g, int32_t cx, int32_t cy, int32_t cz) {
    uint64_t m = spread21((uint64_t)(int64_t)(cx - 1)) | (spread21((uint64_t)(int64_t)(cy - 1)) << 1) |
                 (spread21((uint64_t)(int64_t)(cz - 1)) << 2);
    m += 1;
    return (uint32_t)(g.Hmask ? (m & g.Hmask) : (m % g.H));
}

// Eight window buckets in the reference order [self, x, y, z, xy, xz, yz, xyz] (utils.jl:4-16).
__device__ __forceinline__ void window_of(const HashGrid& g, d3 p, uint32_t w[8], int32_t c[3],
                                          double ctr[3]) {
    cell_of(g, p, c);
    ctr[0] = g.dmin[0] + ((double)c[0] - 0.5) * g.cs;
    ctr[1] = g.dmin[1] + ((double)c[1] - 0.5) * g.cs;
    ctr[2] = g.dmin[2] + ((double)c[2] - 0.5) * g.cs;
    const int32_t dx = p.x > ctr[0] ? 1 : -1;
    const int32_t dy = p.y > ctr[1] ? 1 : -1;
    const int32_t dz = p.z > ctr[2] ? 1 : -1;
    w[0] = bucket_of(g, c[0], c[1], c[2]);
    w[1] = bucket_of(g, c[0] + dx, c[1], c[2]);
    w[2] = bucket_of(g, c[0], c[1] + dy, c[2]);
    w[3] = bucket_of(g, c[0], c[1], c[2] + dz);
    w[4] = bucket_of(g, c[0] + dx, c[1] + dy, c[2]);
    w[5] = bucket_of(g, c[0] + dx, c[1], c[2] + dz);
    w[6] = bucket_of(g, c[0], c[1] + dy, c[2] + dz);
    w[7] = bucket_of(g, c[0] + dx, c[1] + dy, c[2] + dz);
}

__device__ __forceinline__ void window_of(const HashGrid& g, d3 p, uint32_t w[8]) {
    int32_t c[3];
    double ctr[3];
    window_of(g, p, w, c, ctr);
}

// Reference bucket of a position (utils.jl:1-2 minus one).
__device__ __forceinline__ uint32_t ref_bucket(const HashGrid& g, d3 p) {
    int32_t c[3];
    cell_of(g, p, c);
    return bucket_of(g, c[0], c[1], c[2]);
}

// Is the (lo, hi) pair in the reference's cp_list: is hi's bucket one of lo's window buckets?
__device__ __forceinline__ bool ref_candidate(const HashGrid& g, d3 lo, d3 hi) {
    uint32_t w[8];
    window_of(g, lo, w);
    const uint32_t b = ref_bucket(g, hi);
    bool in = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) in |= (w[k] == b);
    return in;
}

// Dense cell grid used for sorting and the broad phase when the domain is compact: cells of
// side cell >= the largest contact distance, linear index x-fastest, positions outside the
// domain clamped into the border cells (clamping keeps touching pairs in adjacent cells).
struct DenseGrid {
    double origin[3];
    double inv_cell;
    double cell;  // 1 / inv_cell: the cell edge
    double inv_dx, inv_dxy;  // 1 / dims[0], 1 / (dims[0] dims[1]) for decoding a cell index
    int32_t dims[3];
    uint32_t cells;
};

// a / d for a < 2^31 from a correctly rounded reciprocal: the product is within 2^-20 of a / d,
// so the truncated quotient is off by at most one and the remainder fixes it.
__device__ __forceinline__ uint32_t div_by(uint32_t a, uint32_t d, double inv) {
    uint32_t q = (uint32_t)((double)a * inv);
    const int64_t r = (int64_t)a - (int64_t)q * (int64_t)d;
    return r < 0 ? q - 1u : (r >= (int64_t)d ? q + 1u : q);
}

// Cell coordinates of a linear dense-cell index.
__device__ __forceinline__ void dense_coords(const DenseGrid& g, uint32_t cell, int32_t c[3]) {
    const uint32_t dx = (uint32_t)g.dims[0], dxy = dx * (uint32_t)g.dims[1];
    const uint32_t z = div_by(cell, dxy, g.inv_dxy);
    const uint32_t rem = cell - z * dxy;
    const uint32_t y = div_by(rem, dx, g.inv_dx);
    c[0] = (int32_t)(rem - y * dx);
    c[1] = (int32_t)y;
    c[2] = (int32_t)z;
}

__device__ __forceinline__ uint32_t dense_cell(const DenseGrid& g, d3 p, int32_t c[3]) {
    const double r[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // clamped to the grid; a non-finite coordinate goes to cell 0
        const double f = floor((r[a] - g.origin[a]) * g.inv_cell);
        c[a] = f >= 0.0 ? (f < (double)g.dims[a] ? (int32_t)f : g.dims[a] - 1) : 0;
    }
    return (uint32_t)c[0] + (uint32_t)g.dims[0] * ((uint32_t)c[1] + (uint32_t)g.dims[1] * (uint32_t)c[2]);
}

// ---- wave-level helpers ----------------------------------------------------------------------

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One atomic per wavefront.
__device__ __forceinline__ void wave_add(unsigned long long* dst, uint64_t v) {
    uint64_t s = wave_sum_u64(v);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(dst, (unsigned long long)s);
}

}  // namespace dem
