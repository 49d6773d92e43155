// Device voxeliser: the generalised winding number and the narrow-band distance at every cell
// centre of a grid, one thread per cell, walking the host tree of geom.cpp (flattened) in the
// same order with the same operations: the reference's voxelize and signed_distance_band
// (src/geom/voxel_grid.cpp:81-138, winding_number.cpp:155-307).
//
// The distance uses only additions, multiplications, divisions, square roots and comparisons,
// all correctly rounded on the device as on the host, so it is the host's bit for bit. The
// winding number also calls atan2, whose device and host results may differ in the last bits:
// the tree walk (its far-field decisions use norms and divisions only) is the same, so the two
// sums differ by a few ulps per solid angle. Cells whose device number lies within 1e-9 of the
// threshold are therefore decided again by the host evaluation; every other cell is decided
// the same way by both, and the mask equals the host's (and the reference's) bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <numbers>
#include <string>
#include <vector>

#include "../../include/orrery_lbm.h"
#include "geom_tree.h"

using lbm::geom::FlatNode;

namespace {

constexpr int kStack = 64;
// Device winding numbers within this distance of the threshold are re-decided on the host
// (device and host sums differ by far less: a few ulps of each solid angle).
constexpr double kRecheck = 1e-9;

struct D3 {
    double x, y, z;
};
__device__ inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline D3 divs(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double squared_norm(D3 a) { return dot(a, a); }

// std::hypot(x, y, z) of the host library: the largest magnitude times the square root of the
// sum of the squared ratios.
__device__ inline double norm(D3 v) {
    const double x = fabs(v.x);
    const double y = fabs(v.y);
    const double z = fabs(v.z);
    const double a = x < y ? (y < z ? z : y) : (x < z ? z : x);
    if (a != 0.0) {  // true for NaN, as the host's conversion to bool
        return a * sqrt((x / a) * (x / a) + (y / a) * (y / a) + (z / a) * (z / a));
    }
    return 0.0;
}

__device__ inline D3 point(const double* p) { return {p[0], p[1], p[2]}; }

// GridSpec::centre: origin + (index + 1/2) spacing.
__device__ inline D3 cell_centre(D3 origin, double h, std::uint64_t c, std::uint64_t nx, std::uint64_t ny) {
    const std::uint64_t i = c % nx;
    const std::uint64_t j = (c / nx) % ny;
    const std::uint64_t k = c / (nx * ny);
    return add(origin, D3{(static_cast<double>(i) + 0.5) * h, (static_cast<double>(j) + 0.5) * h,
                          (static_cast<double>(k) + 0.5) * h});
}

// Van Oosterom and Strackee (1983), on unit vectors.
__device__ double solid_angle(D3 p, D3 a, D3 b, D3 c) {
    const D3 ra = sub(a, p);
    const D3 rb = sub(b, p);
    const D3 rc = sub(c, p);
    const double la = norm(ra);
    const double lb = norm(rb);
    const double lc = norm(rc);
    if (!(la > 0.0) || !(lb > 0.0) || !(lc > 0.0)) {
        return 0.0;
    }
    const D3 ua = divs(ra, la);
    const D3 ub = divs(rb, lb);
    const D3 uc = divs(rc, lc);
    const double numerator = dot(ua, cross(ub, uc));
    const double denominator = 1.0 + dot(ua, ub) + dot(ub, uc) + dot(uc, ua);
    return 2.0 * atan2(numerator, denominator);
}

// The dipole tree walk of the host (children pushed left then right, popped last first). A
// tree deeper than the stack (64 levels: far beyond any mesh that fits in memory) is reported.
__global__ __launch_bounds__(256) void winding_cells(const FlatNode* __restrict__ nodes, const double* __restrict__ tri,
                                                     double scale, double beta, D3 origin, double h, std::uint64_t nx,
                                                     std::uint64_t ny, std::uint64_t n, double* __restrict__ out,
                                                     int* __restrict__ overflowed) {
    const std::uint64_t c = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) {
        return;
    }
    const D3 p = cell_centre(origin, h, c, nx, ny);
    std::uint32_t stack[kStack];
    int top = 0;
    stack[top++] = 0;
    double sum = 0.0;
    while (top > 0) {
        const FlatNode& node = nodes[stack[--top]];
        const D3 d = sub(point(node.centre), p);
        const double dist = norm(d);
        if (dist > beta * node.radius) {
            const double flux = dot(point(node.normal), divs(d, dist));
            if (flux != 0.0) {
                const double ratio = scale / dist;
                sum += flux * ratio * ratio;
            }
            continue;
        }
        if (node.leaf != 0U) {
            for (std::uint32_t i = node.first; i < node.first + node.count; ++i) {
                const double* t = tri + 9U * static_cast<std::size_t>(i);
                sum += solid_angle(p, point(t), point(t + 3), point(t + 6));
            }
            continue;
        }
        if (top + 2 > kStack) {
            overflowed[0] = 1;
            break;
        }
        stack[top++] = node.left;
        stack[top++] = node.right;
    }
    out[c] = sum / (4.0 * std::numbers::pi);
}

// std::max({lo - p, 0, p - hi}): the first largest of the three.
__device__ inline double gap(double lo, double p, double hi) {
    double r = lo - p;
    if (r < 0.0) r = 0.0;
    const double c = p - hi;
    if (r < c) r = c;
    return r;
}

__device__ inline double box_distance_squared(const FlatNode& b, D3 p) {
    const double dx = gap(b.lo[0], p.x, b.hi[0]);
    const double dy = gap(b.lo[1], p.y, b.hi[1]);
    const double dz = gap(b.lo[2], p.z, b.hi[2]);
    return dx * dx + dy * dy + dz * dz;
}

__device__ D3 closest_on_segment(D3 p, D3 a, D3 b) {
    const D3 ab = sub(b, a);
    const double length2 = squared_norm(ab);
    if (!(length2 > 0.0)) {
        return a;
    }
    double t = dot(sub(p, a), ab) / length2;  // std::clamp(t, 0, 1)
    t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
    return add(a, mul(ab, t));
}

__device__ D3 closest_on_edges(D3 p, D3 a, D3 b, D3 c) {
    D3 best = closest_on_segment(p, a, b);
    double best2 = squared_norm(sub(best, p));
    const D3 q1 = closest_on_segment(p, b, c);
    const D3 q2 = closest_on_segment(p, c, a);
    const double d1 = squared_norm(sub(q1, p));
    if (d1 < best2) {
        best = q1;
        best2 = d1;
    }
    const double d2 = squared_norm(sub(q2, p));
    if (d2 < best2) {
        best = q2;
    }
    return best;
}

// Ericson, Real-Time Collision Detection, 5.1.5 (the host's branch order).
__device__ D3 closest_on_triangle(D3 p, D3 a, D3 b, D3 c) {
    const D3 ab = sub(b, a);
    const D3 ac = sub(c, a);
    if (!(squared_norm(cross(ab, ac)) > 0.0)) {
        return closest_on_edges(p, a, b, c);
    }
    const D3 ap = sub(p, a);
    const double d1 = dot(ab, ap);
    const double d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        return a;
    }
    const D3 bp = sub(p, b);
    const double d3 = dot(ab, bp);
    const double d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        return add(a, mul(ab, d1 / (d1 - d3)));
    }
    const D3 cp = sub(p, c);
    const double d5 = dot(ab, cp);
    const double d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        return add(a, mul(ac, d2 / (d2 - d6)));
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    }
    const double sum = va + vb + vc;
    if (!(sum > 0.0)) {
        return closest_on_edges(p, a, b, c);
    }
    return add(add(a, mul(ab, vb / sum)), mul(ac, vc / sum));
}

// Nearest-first walk of the host's distance query, bounded by the band; signed by the mask.
__global__ __launch_bounds__(256) void distance_cells(const FlatNode* __restrict__ nodes, const double* __restrict__ tri,
                                                      D3 origin, double h, std::uint64_t nx, std::uint64_t ny,
                                                      std::uint64_t n, const std::uint8_t* __restrict__ solid,
                                                      double band, double* __restrict__ distance,
                                                      std::uint8_t* __restrict__ in_band, int* __restrict__ overflowed) {
    const std::uint64_t c = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) {
        return;
    }
    const D3 p = cell_centre(origin, h, c, nx, ny);
    double best2 = band * band;
    std::uint32_t stack[kStack];
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        const FlatNode& node = nodes[stack[--top]];
        if (box_distance_squared(node, p) >= best2) {
            continue;
        }
        if (node.leaf != 0U) {
            for (std::uint32_t i = node.first; i < node.first + node.count; ++i) {
                const double* t = tri + 9U * static_cast<std::size_t>(i);
                const D3 q = closest_on_triangle(p, point(t), point(t + 3), point(t + 6));
                const double d2 = squared_norm(sub(q, p));
                best2 = d2 < best2 ? d2 : best2;
            }
            continue;
        }
        if (top + 2 > kStack) {
            overflowed[0] = 1;
            break;
        }
        const double dl = box_distance_squared(nodes[node.left], p);
        const double dr = box_distance_squared(nodes[node.right], p);
        const bool left_first = dl <= dr;
        stack[top++] = left_first ? node.right : node.left;
        stack[top++] = left_first ? node.left : node.right;
    }
    const double d = sqrt(best2);
    if (in_band != nullptr) {
        in_band[c] = static_cast<std::uint8_t>(d < band);
    }
    distance[c] = (solid[c] != 0 ? -1.0 : 1.0) * d;
}

int device_fail(const char* what, hipError_t e) {
    (void)lbm::geom::fail(std::string(what) + ": " + hipGetErrorString(e));
    return LBM_EDEVICE;
}

// Device buffers of one call, freed on every path.
struct Buffers {
    std::vector<void*> held;
    ~Buffers() {
        for (void* p : held) {
            (void)hipFree(p);
        }
    }
    template <class P> hipError_t alloc(P** out, std::size_t bytes) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, bytes == 0 ? 8 : bytes);
        if (e == hipSuccess) {
            held.push_back(p);
            *out = static_cast<P*>(p);
        }
        return e;
    }
};

hipError_t upload_tree(Buffers& buf, const lbm::geom::FlatTree& t, FlatNode** nodes, double** tri) {
    hipError_t e = buf.alloc(nodes, t.nodes.size() * sizeof(FlatNode));
    if (e == hipSuccess) e = buf.alloc(tri, t.tri.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(*nodes, t.nodes.data(), t.nodes.size() * sizeof(FlatNode), hipMemcpyHostToDevice);
    if (e == hipSuccess && !t.tri.empty()) {
        e = hipMemcpy(*tri, t.tri.data(), t.tri.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    return e;
}

unsigned blocks_of(std::uint64_t n) { return static_cast<unsigned>((n + 255U) / 256U); }

} // namespace

extern "C" {

int lbm_voxelize_device(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny,
                        uint64_t nz, double threshold, uint8_t* solid, double* winding, size_t* rechecked) {
    if (m == nullptr || origin == nullptr || solid == nullptr) {
        return lbm::geom::fail("voxelize: null argument");
    }
    if (!lbm::geom::check_grid(origin, spacing, nx, ny, nz)) {
        return LBM_EINVAL;
    }
    if (!std::isfinite(threshold)) {
        return lbm::geom::fail("voxelize: threshold must be finite");
    }
    const std::uint64_t n = nx * ny * nz;
    lbm::geom::FlatTree tree;
    lbm::geom::flatten_tree(m, tree);
    std::vector<double> w(n, 0.0);
    if (!tree.nodes.empty()) {
        Buffers buf;
        FlatNode* nodes = nullptr;
        double* tri = nullptr;
        double* out = nullptr;
        int* overflow = nullptr;
        hipError_t e = upload_tree(buf, tree, &nodes, &tri);
        if (e == hipSuccess) e = buf.alloc(&out, n * sizeof(double));
        if (e == hipSuccess) e = buf.alloc(&overflow, sizeof(int));
        if (e == hipSuccess) e = hipMemset(overflow, 0, sizeof(int));
        if (e != hipSuccess) {
            return device_fail("voxelize (device): buffers", e);
        }
        hipLaunchKernelGGL(winding_cells, dim3(blocks_of(n)), dim3(256), 0, nullptr, nodes, tri, tree.scale, tree.beta,
                           D3{origin[0], origin[1], origin[2]}, spacing, nx, ny, n, out, overflow);
        int over = 0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(w.data(), out, n * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(&over, overflow, sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            return device_fail("voxelize (device)", e);
        }
        if (over != 0) {
            return lbm::geom::fail("voxelize (device): the mesh tree is deeper than the device walk's stack");
        }
    }
    // Cells near the threshold: the host evaluation decides.
    std::vector<std::uint64_t> near;
    std::vector<double> points;
    for (std::uint64_t c = 0; c < n; ++c) {
        if (!(std::fabs(w[c] - threshold) >= kRecheck)) {
            near.push_back(c);
            const std::uint64_t i = c % nx;
            const std::uint64_t j = (c / nx) % ny;
            const std::uint64_t k = c / (nx * ny);
            points.push_back(origin[0] + (static_cast<double>(i) + 0.5) * spacing);
            points.push_back(origin[1] + (static_cast<double>(j) + 0.5) * spacing);
            points.push_back(origin[2] + (static_cast<double>(k) + 0.5) * spacing);
        }
    }
    if (!near.empty()) {
        std::vector<double> exact(near.size());
        const int rc = lbm_winding_numbers(m, points.data(), near.size(), exact.data());
        if (rc != LBM_OK) {
            return rc;
        }
        for (std::size_t q = 0; q < near.size(); ++q) {
            w[near[q]] = exact[q];
        }
    }
    for (std::uint64_t c = 0; c < n; ++c) {
        solid[c] = static_cast<std::uint8_t>(w[c] >= threshold);
        if (winding != nullptr) {
            winding[c] = w[c];
        }
    }
    if (rechecked != nullptr) {
        *rechecked = near.size();
    }
    return LBM_OK;
}

int lbm_signed_distance_band_device(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx,
                                    uint64_t ny, uint64_t nz, const uint8_t* solid, double band_cells,
                                    double* distance, uint8_t* in_band) {
    if (m == nullptr || origin == nullptr || solid == nullptr || distance == nullptr) {
        return lbm::geom::fail("signed_distance_band: null argument");
    }
    if (!std::isfinite(band_cells) || !(band_cells > 0.0)) {
        return lbm::geom::fail("signed_distance_band: band must be positive");
    }
    if (!lbm::geom::check_grid(origin, spacing, nx, ny, nz)) {
        return LBM_EINVAL;
    }
    const double band = band_cells * spacing;
    if (!std::isfinite(band)) {
        return lbm::geom::fail("signed_distance_band: band is not finite");
    }
    const std::uint64_t n = nx * ny * nz;
    lbm::geom::FlatTree tree;
    lbm::geom::flatten_tree(m, tree);
    if (tree.nodes.empty()) {
        for (std::uint64_t c = 0; c < n; ++c) {
            distance[c] = (solid[c] != 0 ? -1.0 : 1.0) * band;
            if (in_band != nullptr) {
                in_band[c] = 0;
            }
        }
        return LBM_OK;
    }
    Buffers buf;
    FlatNode* nodes = nullptr;
    double* tri = nullptr;
    std::uint8_t* dsolid = nullptr;
    double* dist = nullptr;
    std::uint8_t* band_mask = nullptr;
    int* overflow = nullptr;
    hipError_t e = upload_tree(buf, tree, &nodes, &tri);
    if (e == hipSuccess) e = buf.alloc(&dsolid, n);
    if (e == hipSuccess) e = buf.alloc(&dist, n * sizeof(double));
    if (e == hipSuccess) e = buf.alloc(&band_mask, n);
    if (e == hipSuccess) e = buf.alloc(&overflow, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(dsolid, solid, n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(overflow, 0, sizeof(int));
    if (e != hipSuccess) {
        return device_fail("signed_distance_band (device): buffers", e);
    }
    hipLaunchKernelGGL(distance_cells, dim3(blocks_of(n)), dim3(256), 0, nullptr, nodes, tri,
                       D3{origin[0], origin[1], origin[2]}, spacing, nx, ny, n, dsolid, band, dist, band_mask,
                       overflow);
    int over = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(distance, dist, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && in_band != nullptr) e = hipMemcpy(in_band, band_mask, n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&over, overflow, sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        return device_fail("signed_distance_band (device)", e);
    }
    if (over != 0) {
        return lbm::geom::fail("signed_distance_band (device): the mesh tree is deeper than the device walk's stack");
    }
    return LBM_OK;
}

} // extern "C"
