// Host geometry for the lattice masks: body meshes (solids of revolution, the blunt capsule, the
// wedge, the box, the NACA 4-digit wing), rigid transforms, and the voxelisation of a mesh by
// its generalised winding number (hierarchical far field) plus the narrow-band signed distance.
//
// Masks decide solid nodes by thresholding an FP64 winding number, so they are integer outputs
// that must equal the reference's bit for bit. The arithmetic therefore follows
// src/geom/primitives.cpp, naca.cpp, winding_number.cpp and voxel_grid.cpp expression for
// expression (same libm calls, same summation orders, the same std::nth_element partition of
// the tree build), and this file is compiled by the host C++ compiler with the reference's
// floating-point policy (-O2 -ffp-contract=off -fno-fast-math). The voxel loops run on all host
// threads; every cell is computed independently, so the result does not depend on the split.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <numbers>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/orrery_lbm.h"
#include "geom_tree.h"

namespace {

constexpr double kPi = std::numbers::pi;

thread_local std::string g_geom_error;

int geom_fail(const std::string& m) {
    g_geom_error = m;
    return LBM_EINVAL;
}

struct V3 {
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(const V3& a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(const V3& a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    return a;
}
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(const V3& a) { return std::hypot(a.x, a.y, a.z); }
inline double squared_norm(const V3& a) { return dot(a, a); }
inline bool finite(const V3& a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

using Tri = std::array<std::uint32_t, 3>;

struct Box {
    V3 min;
    V3 max;
    V3 extent() const { return max - min; }
    V3 centre() const { return (min + max) * 0.5; }
};

// ---- solids of revolution (primitives.cpp:24-290) ------------------------------------------
struct PP {
    double r = 0.0;
    double z = 0.0;
};

struct Segment {
    bool arc = false;
    PP start;
    PP end;
    PP centre;
    double radius = 0.0;
    double phi0 = 0.0;
    double phi1 = 0.0;
};

PP arc_point(PP c, double radius, double phi) { return {c.r + radius * std::cos(phi), c.z + radius * std::sin(phi)}; }

Segment line(PP a, PP b) {
    Segment s;
    s.start = a;
    s.end = b;
    return s;
}

Segment arc(PP c, double radius, double phi0, double phi1) {
    Segment s;
    s.arc = true;
    s.centre = c;
    s.radius = radius;
    s.phi0 = phi0;
    s.phi1 = phi1;
    s.start = arc_point(c, radius, phi0);
    s.end = arc_point(c, radius, phi1);
    return s;
}

int arc_pieces(const Segment& s, double arc_step) {
    const double pieces = std::ceil(std::abs(s.phi1 - s.phi0) / arc_step);
    return std::max(1, static_cast<int>(std::min(pieces, 1e6)));
}

struct Mesh {
    std::vector<V3> v;
    std::vector<Tri> t;
};

// RevolutionProfile::polyline + mesh (primitives.cpp:191-290)
bool revolution_mesh(const std::vector<Segment>& segments, int around, double arc_step, Mesh& m) {
    if (around < 3 || around > 100000 || !(arc_step > 1e-6) || !(arc_step <= kPi)) {
        geom_fail("revolution mesh: around must be in [3, 100000] and arc_step in (1e-6, pi]");
        return false;
    }
    std::vector<PP> p;
    p.push_back(segments.front().start);
    for (const Segment& s : segments) {
        if (!s.arc) {
            p.push_back(s.end);
            continue;
        }
        const int pieces = arc_pieces(s, arc_step);
        for (int i = 1; i < pieces; ++i) {
            const double t = static_cast<double>(i) / static_cast<double>(pieces);
            p.push_back(arc_point(s.centre, s.radius, s.phi0 + (s.phi1 - s.phi0) * t));
        }
        p.push_back(s.end);
    }
    p.front().r = 0.0;
    p.back().r = 0.0;
    const std::size_t rings = p.size() - 2;
    if (rings == 0) {
        geom_fail("revolution mesh: profile has no point off the axis");
        return false;
    }
    const auto n = static_cast<std::size_t>(around);
    m.v.clear();
    m.t.clear();
    m.v.push_back({0.0, 0.0, p.front().z});
    for (std::size_t i = 1; i <= rings; ++i) {
        for (std::size_t j = 0; j < n; ++j) {
            const double theta = 2.0 * kPi * static_cast<double>(j) / static_cast<double>(n);
            m.v.push_back({p[i].r * std::cos(theta), p[i].r * std::sin(theta), p[i].z});
        }
    }
    m.v.push_back({0.0, 0.0, p.back().z});
    const std::uint32_t bottom = 0;
    const auto top = static_cast<std::uint32_t>(m.v.size() - 1);
    const auto ring = [n](std::size_t i, std::size_t j) { return static_cast<std::uint32_t>(1 + (i - 1) * n + (j % n)); };
    for (std::size_t j = 0; j < n; ++j) {
        m.t.push_back({bottom, ring(1, j + 1), ring(1, j)});
    }
    for (std::size_t i = 1; i < rings; ++i) {
        for (std::size_t j = 0; j < n; ++j) {
            const std::uint32_t a = ring(i, j);
            const std::uint32_t b = ring(i, j + 1);
            const std::uint32_t c = ring(i + 1, j + 1);
            const std::uint32_t d = ring(i + 1, j);
            m.t.push_back({a, b, c});
            m.t.push_back({a, c, d});
        }
    }
    for (std::size_t j = 0; j < n; ++j) {
        m.t.push_back({ring(rings, j), ring(rings, j + 1), top});
    }
    return true;
}

// capsule_layout + capsule_profile (primitives.cpp:347-426)
std::vector<Segment> capsule_segments(double max_diameter, double nose_radius, double shoulder_radius,
                                      double afterbody_half_angle, double top_diameter) {
    const double radius = 0.5 * max_diameter;
    const double cr = radius - shoulder_radius;
    const double d = nose_radius - shoulder_radius;
    const double cz = nose_radius - std::sqrt(d * d - cr * cr);
    const PP shoulder_centre{cr, cz};
    const double tangent_phi = std::atan2(cz - nose_radius, cr);
    const PP cone_start = arc_point(shoulder_centre, shoulder_radius, afterbody_half_angle);
    const double top_r = 0.5 * top_diameter;
    const double top_z = cone_start.z + (cone_start.r - top_r) / std::tan(afterbody_half_angle);
    std::vector<Segment> segments;
    segments.push_back(arc({0.0, nose_radius}, nose_radius, -kPi / 2.0, tangent_phi));
    Segment shoulder = arc(shoulder_centre, shoulder_radius, tangent_phi, afterbody_half_angle);
    shoulder.start = segments.back().end;
    segments.push_back(shoulder);
    segments.push_back(line(cone_start, {top_r, top_z}));
    if (top_r > 0.0) {
        segments.push_back(line({top_r, top_z}, {0.0, top_z}));
    }
    return segments;
}

// ---- NACA 4-digit wing (naca.cpp:98-257) -------------------------------------------------------
struct Naca {
    double m;
    double p;
    double t;
    bool cambered() const { return m != 0.0 && p > 0.0 && p < 1.0; }
    double half_thickness(double x) const {
        return 5.0 * t * (0.2969 * std::sqrt(x) + x * (-0.1260 + x * (-0.3516 + x * (0.2843 - 0.1036 * x))));
    }
    double camber(double x) const {
        if (!cambered()) {
            return 0.0;
        }
        if (x < p) {
            return m / (p * p) * (2.0 * p * x - x * x);
        }
        return m / ((1.0 - p) * (1.0 - p)) * (1.0 - 2.0 * p + 2.0 * p * x - x * x);
    }
    double camber_slope(double x) const {
        if (!cambered()) {
            return 0.0;
        }
        const double d = x < p ? p * p : (1.0 - p) * (1.0 - p);
        return 2.0 * m / d * (p - x);
    }
};

// Naca4::section (naca.cpp:155-172): 2 n points (x, z) per unit chord, counter-clockwise from
// the trailing edge over the upper surface, cosine-spaced stations.
std::vector<PP> naca_section(const Naca& a, int n) {
    const auto count = static_cast<std::size_t>(n);
    std::vector<PP> section(2 * count); // (x, z) in r, z
    section[0] = {1.0, 0.0};
    section[count] = {0.0, 0.0};
    for (std::size_t s = 1; s < count; ++s) {
        const double x = 0.5 * (1.0 - std::cos(kPi * static_cast<double>(s) / n));
        const double yt = a.half_thickness(x);
        const double theta = std::atan(a.camber_slope(x));
        const double yc = a.camber(x);
        section[count - s] = {x - yt * std::sin(theta), yc + yt * std::cos(theta)};
        section[count + s] = {x + yt * std::sin(theta), yc - yt * std::cos(theta)};
    }
    return section;
}

void naca_wing(const Naca& a, double chord, double span, int n, int span_divisions, Mesh& m) {
    const auto count = static_cast<std::size_t>(n);
    const std::vector<PP> section = naca_section(a, n);
    const std::size_t ring = section.size();
    const auto strips = static_cast<std::size_t>(span_divisions);
    m.v.clear();
    m.t.clear();
    for (std::size_t j = 0; j <= strips; ++j) {
        const double y = span * (static_cast<double>(j) / static_cast<double>(strips) - 0.5);
        for (const PP& q : section) {
            m.v.push_back({chord * q.r, y, chord * q.z});
        }
    }
    const auto at = [ring](std::size_t j, std::size_t k) {
        return static_cast<std::uint32_t>(j * ring + (k == ring ? 0 : k));
    };
    for (std::size_t j = 0; j < strips; ++j) {
        for (std::size_t k = 0; k < ring; ++k) {
            const std::uint32_t a0 = at(j, k);
            const std::uint32_t b0 = at(j, k + 1);
            const std::uint32_t c0 = at(j + 1, k + 1);
            const std::uint32_t d0 = at(j + 1, k);
            m.t.push_back({a0, c0, b0});
            m.t.push_back({a0, d0, c0});
        }
    }
    const std::size_t stations = count;
    const auto upper = [stations](std::size_t s) { return stations - s; };
    const auto lower = [stations, ring](std::size_t s) { return stations + s == ring ? 0 : stations + s; };
    for (std::size_t s = 0; s < stations; ++s) {
        std::vector<std::array<std::size_t, 3>> cap;
        if (s > 0) {
            cap.push_back({upper(s + 1), upper(s), lower(s)});
        }
        if (s + 1 < stations) {
            cap.push_back({upper(s + 1), lower(s), lower(s + 1)});
        }
        for (const auto& c : cap) {
            m.t.push_back({at(0, c[0]), at(0, c[1]), at(0, c[2])});
            m.t.push_back({at(strips, c[0]), at(strips, c[2]), at(strips, c[1])});
        }
    }
}

// ---- winding number and distance (winding_number.cpp:21-307) ----------------------------------
Box bounds_of(const Mesh& m) {
    if (m.v.empty()) {
        return {};
    }
    Box b{m.v.front(), m.v.front()};
    for (const V3& v : m.v) {
        b.min = {std::min(b.min.x, v.x), std::min(b.min.y, v.y), std::min(b.min.z, v.z)};
        b.max = {std::max(b.max.x, v.x), std::max(b.max.y, v.y), std::max(b.max.z, v.z)};
    }
    return b;
}

double box_distance_squared(const Box& b, const V3& p) {
    const double dx = std::max({b.min.x - p.x, 0.0, p.x - b.max.x});
    const double dy = std::max({b.min.y - p.y, 0.0, p.y - b.max.y});
    const double dz = std::max({b.min.z - p.z, 0.0, p.z - b.max.z});
    return dx * dx + dy * dy + dz * dz;
}

Box merge(const Box& a, const Box& b) {
    return {{std::min(a.min.x, b.min.x), std::min(a.min.y, b.min.y), std::min(a.min.z, b.min.z)},
            {std::max(a.max.x, b.max.x), std::max(a.max.y, b.max.y), std::max(a.max.z, b.max.z)}};
}

V3 closest_on_segment(const V3& p, const V3& a, const V3& b) {
    const V3 ab = b - a;
    const double length2 = squared_norm(ab);
    if (!(length2 > 0.0)) {
        return a;
    }
    const double t = std::clamp(dot(p - a, ab) / length2, 0.0, 1.0);
    return a + ab * t;
}

V3 closest_on_edges(const V3& p, const V3& a, const V3& b, const V3& c) {
    V3 best = closest_on_segment(p, a, b);
    double best2 = squared_norm(best - p);
    for (const V3& q : {closest_on_segment(p, b, c), closest_on_segment(p, c, a)}) {
        const double d2 = squared_norm(q - p);
        if (d2 < best2) {
            best = q;
            best2 = d2;
        }
    }
    return best;
}

// Ericson, Real-Time Collision Detection, 5.1.5
V3 closest_on_triangle(const V3& p, const V3& a, const V3& b, const V3& c) {
    const V3 ab = b - a;
    const V3 ac = c - a;
    if (!(squared_norm(cross(ab, ac)) > 0.0)) {
        return closest_on_edges(p, a, b, c);
    }
    const V3 ap = p - a;
    const double d1 = dot(ab, ap);
    const double d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        return a;
    }
    const V3 bp = p - b;
    const double d3 = dot(ab, bp);
    const double d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        return a + ab * (d1 / (d1 - d3));
    }
    const V3 cp = p - c;
    const double d5 = dot(ab, cp);
    const double d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        return a + ac * (d2 / (d2 - d6));
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
    }
    const double sum = va + vb + vc;
    if (!(sum > 0.0)) {
        return closest_on_edges(p, a, b, c);
    }
    return a + ab * (vb / sum) + ac * (vc / sum);
}

// Van Oosterom and Strackee (1983), on unit vectors.
double solid_angle(const V3& p, const V3& a, const V3& b, const V3& c) {
    const V3 ra = a - p;
    const V3 rb = b - p;
    const V3 rc = c - p;
    const double la = norm(ra);
    const double lb = norm(rb);
    const double lc = norm(rc);
    if (!(la > 0.0) || !(lb > 0.0) || !(lc > 0.0)) {
        return 0.0;
    }
    const V3 ua = ra / la;
    const V3 ub = rb / lb;
    const V3 uc = rc / lc;
    const double numerator = dot(ua, cross(ub, uc));
    const double denominator = 1.0 + dot(ua, ub) + dot(ub, uc) + dot(uc, ua);
    return 2.0 * std::atan2(numerator, denominator);
}

// Bounding-volume tree with the dipole far field of Barill et al. (2018), beta = 2, leaves of 8.
class Tree {
public:
    explicit Tree(const Mesh& m, double beta = 2.0, std::size_t leaf = 8) : m_(m), beta_(beta) {
        const V3 e = bounds_of(m).extent();
        const double extent = std::max({e.x, e.y, e.z});
        scale_ = (!(extent > 0.0) || !std::isfinite(extent)) ? 1.0 : std::ldexp(1.0, std::ilogb(extent));
        order_.assign(m.t.size(), 0U);
        std::iota(order_.begin(), order_.end(), 0U);
        if (!order_.empty()) {
            nodes_.reserve(2 * (m.t.size() / leaf + 1));
            build(0, static_cast<std::uint32_t>(order_.size()), leaf);
        }
    }

    double winding(const V3& p) const {
        if (nodes_.empty()) {
            return 0.0;
        }
        std::array<std::uint32_t, 128> stack{};
        std::size_t top = 0;
        stack[top++] = 0;
        double sum = 0.0;
        while (top > 0) {
            const Node& node = nodes_[stack[--top]];
            const V3 d = node.centre - p;
            const double dist = norm(d);
            if (dist > beta_ * node.radius) {
                const double flux = dot(node.area_normal, d / dist);
                if (flux != 0.0) {
                    const double ratio = scale_ / dist;
                    sum += flux * ratio * ratio;
                }
                continue;
            }
            if (node.leaf) {
                for (std::uint32_t i = node.first; i < node.first + node.count; ++i) {
                    const Tri& t = m_.t[order_[i]];
                    sum += solid_angle(p, m_.v[t[0]], m_.v[t[1]], m_.v[t[2]]);
                }
                continue;
            }
            stack[top++] = node.left;
            stack[top++] = node.right;
        }
        return sum / (4.0 * kPi);
    }

    double distance(const V3& p, double max_distance) const {
        double best2 = max_distance * max_distance;
        if (nodes_.empty()) {
            return max_distance;
        }
        std::array<std::uint32_t, 128> stack{};
        std::size_t top = 0;
        stack[top++] = 0;
        while (top > 0) {
            const Node& node = nodes_[stack[--top]];
            if (box_distance_squared(node.box, p) >= best2) {
                continue;
            }
            if (node.leaf) {
                for (std::uint32_t i = node.first; i < node.first + node.count; ++i) {
                    const Tri& t = m_.t[order_[i]];
                    const V3 q = closest_on_triangle(p, m_.v[t[0]], m_.v[t[1]], m_.v[t[2]]);
                    best2 = std::min(best2, squared_norm(q - p));
                }
                continue;
            }
            const double dl = box_distance_squared(nodes_[node.left].box, p);
            const double dr = box_distance_squared(nodes_[node.right].box, p);
            const bool left_first = dl <= dr;
            stack[top++] = left_first ? node.right : node.left;
            stack[top++] = left_first ? node.left : node.right;
        }
        return std::sqrt(best2);
    }

    void flatten(lbm::geom::FlatTree& out) const {
        out.scale = scale_;
        out.beta = beta_;
        out.nodes.resize(nodes_.size());
        for (std::size_t k = 0; k < nodes_.size(); ++k) {
            const Node& n = nodes_[k];
            lbm::geom::FlatNode& f = out.nodes[k];
            f = lbm::geom::FlatNode{{n.centre.x, n.centre.y, n.centre.z},
                                    {n.area_normal.x, n.area_normal.y, n.area_normal.z},
                                    n.radius,
                                    {n.box.min.x, n.box.min.y, n.box.min.z},
                                    {n.box.max.x, n.box.max.y, n.box.max.z},
                                    n.first,
                                    n.count,
                                    n.left,
                                    n.right,
                                    n.leaf ? 1U : 0U,
                                    0U};
        }
        out.tri.resize(9 * order_.size());
        for (std::size_t i = 0; i < order_.size(); ++i) {
            const Tri& t = m_.t[order_[i]];
            for (int c = 0; c < 3; ++c) {
                const V3& v = m_.v[t[c]];
                out.tri[9 * i + 3 * c + 0] = v.x;
                out.tri[9 * i + 3 * c + 1] = v.y;
                out.tri[9 * i + 3 * c + 2] = v.z;
            }
        }
    }

private:
    struct Node {
        Box box;
        V3 centre;
        V3 area_normal;
        double radius = 0.0;
        std::uint32_t first = 0;
        std::uint32_t count = 0;
        std::uint32_t left = 0;
        std::uint32_t right = 0;
        bool leaf = true;
    };

    V3 centroid(std::uint32_t t) const {
        const Tri& tri = m_.t[t];
        return (m_.v[tri[0]] + m_.v[tri[1]] + m_.v[tri[2]]) / 3.0;
    }

    std::uint32_t build(std::uint32_t first, std::uint32_t count, std::size_t leaf) {
        const auto index = static_cast<std::uint32_t>(nodes_.size());
        nodes_.emplace_back();
        Node node;
        node.first = first;
        node.count = count;
        node.box = Box{m_.v[m_.t[order_[first]][0]], m_.v[m_.t[order_[first]][0]]};
        const V3 c0 = centroid(order_[first]);
        Box centres{c0, c0};
        double area = 0.0;
        V3 weighted;
        for (std::uint32_t i = first; i < first + count; ++i) {
            const Tri& tri = m_.t[order_[i]];
            for (std::uint32_t v : tri) {
                node.box = merge(node.box, Box{m_.v[v], m_.v[v]});
            }
            const V3 c = centroid(order_[i]);
            centres = merge(centres, Box{c, c});
            const V3& v0 = m_.v[tri[0]];
            const V3 n2 = cross((m_.v[tri[1]] - v0) / scale_, (m_.v[tri[2]] - v0) / scale_);
            const double a = 0.5 * norm(n2);
            node.area_normal += n2 * 0.5;
            weighted += c * a;
            area += a;
        }
        node.centre = area > 0.0 ? weighted / area : node.box.centre();
        if (count <= leaf) {
            node.leaf = true;
            for (std::uint32_t i = first; i < first + count; ++i) {
                for (std::uint32_t v : m_.t[order_[i]]) {
                    node.radius = std::max(node.radius, norm(m_.v[v] - node.centre));
                }
            }
            nodes_[index] = node;
            return index;
        }
        const V3 e = centres.extent();
        const int axis = e.x >= e.y && e.x >= e.z ? 0 : (e.y >= e.z ? 1 : 2);
        const auto key = [&](std::uint32_t t) {
            const V3 c = centroid(t);
            return axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
        };
        const std::uint32_t half = count / 2;
        std::nth_element(order_.begin() + first, order_.begin() + first + half, order_.begin() + first + count,
                         [&](std::uint32_t a, std::uint32_t b) {
                             const double ka = key(a);
                             const double kb = key(b);
                             return ka != kb ? ka < kb : a < b;
                         });
        node.leaf = false;
        node.left = build(first, half, leaf);
        node.right = build(first + half, count - half, leaf);
        for (std::uint32_t child : {node.left, node.right}) {
            const Node& c = nodes_[child];
            node.radius = std::max(node.radius, norm(c.centre - node.centre) + c.radius);
        }
        nodes_[index] = node;
        return index;
    }

    const Mesh& m_;
    double beta_;
    double scale_ = 1.0;
    std::vector<std::uint32_t> order_;
    std::vector<Node> nodes_;
};

// Runs body(k) for k in [0, n) on the host threads.
template <class F> void parallel_slabs(std::size_t n, F&& body) {
    const unsigned hw = std::max(1U, std::thread::hardware_concurrency());
    const std::size_t workers = std::min<std::size_t>(hw, n);
    if (workers <= 1) {
        for (std::size_t k = 0; k < n; ++k) {
            body(k);
        }
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(workers);
    for (std::size_t w = 0; w < workers; ++w) {
        pool.emplace_back([&, w] {
            for (std::size_t k = w; k < n; k += workers) {
                body(k);
            }
        });
    }
    for (auto& t : pool) {
        t.join();
    }
}

// GridSpec::centre (voxel_grid.cpp:12-16)
V3 centre_of(const double* origin, double spacing, std::size_t i, std::size_t j, std::size_t k) {
    return V3{origin[0], origin[1], origin[2]} + V3{(static_cast<double>(i) + 0.5) * spacing,
                                                    (static_cast<double>(j) + 0.5) * spacing,
                                                    (static_cast<double>(k) + 0.5) * spacing};
}

bool grid_ok(const double* origin, double spacing, std::uint64_t nx, std::uint64_t ny, std::uint64_t nz) {
    const std::uint64_t max = std::uint64_t{1} << 27U;
    if (nx == 0 || ny == 0 || nz == 0 || nx > max / ny || nx * ny > max / nz) {
        geom_fail("voxelize: grid empty or too large");
        return false;
    }
    if (!std::isfinite(origin[0]) || !std::isfinite(origin[1]) || !std::isfinite(origin[2]) ||
        !std::isfinite(spacing) || !(spacing > 0.0) || !finite(centre_of(origin, spacing, nx - 1, ny - 1, nz - 1))) {
        geom_fail("voxelize: grid origin must be finite and spacing positive and finite");
        return false;
    }
    return true;
}

} // namespace

struct lbm_mesh {
    Mesh m;
};

namespace lbm::geom {

void flatten_tree(const lbm_mesh* m, FlatTree& out) {
    const Tree tree(m->m);
    tree.flatten(out);
}

bool check_grid(const double* origin, double spacing, std::uint64_t nx, std::uint64_t ny, std::uint64_t nz) {
    return grid_ok(origin, spacing, nx, ny, nz);
}

int fail(const std::string& message) { return geom_fail(message); }

} // namespace lbm::geom

extern "C" {

const char* lbm_geom_last_error(void) { return g_geom_error.c_str(); }

int lbm_mesh_create(const double* vertices, size_t nv, const uint32_t* triangles, size_t nt, lbm_mesh** out) {
    if (out == nullptr || (nv > 0 && vertices == nullptr) || (nt > 0 && triangles == nullptr)) {
        return geom_fail("lbm_mesh_create: null argument");
    }
    auto* h = new (std::nothrow) lbm_mesh();
    if (h == nullptr) {
        return LBM_ENOMEM;
    }
    h->m.v.resize(nv);
    for (std::size_t k = 0; k < nv; ++k) {
        h->m.v[k] = {vertices[3 * k], vertices[3 * k + 1], vertices[3 * k + 2]};
    }
    h->m.t.resize(nt);
    for (std::size_t k = 0; k < nt; ++k) {
        for (int a = 0; a < 3; ++a) {
            if (triangles[3 * k + a] >= nv) {
                delete h;
                return geom_fail("lbm_mesh_create: triangle index out of range");
            }
            h->m.t[k][a] = triangles[3 * k + a];
        }
    }
    *out = h;
    return LBM_OK;
}

int lbm_mesh_destroy(lbm_mesh* m) {
    delete m;
    return LBM_OK;
}

int lbm_mesh_size(const lbm_mesh* m, size_t* nv, size_t* nt) {
    if (m == nullptr) {
        return geom_fail("null mesh");
    }
    if (nv != nullptr) *nv = m->m.v.size();
    if (nt != nullptr) *nt = m->m.t.size();
    return LBM_OK;
}

int lbm_mesh_data(const lbm_mesh* m, double* vertices, uint32_t* triangles) {
    if (m == nullptr) {
        return geom_fail("null mesh");
    }
    if (vertices != nullptr) {
        for (std::size_t k = 0; k < m->m.v.size(); ++k) {
            vertices[3 * k] = m->m.v[k].x;
            vertices[3 * k + 1] = m->m.v[k].y;
            vertices[3 * k + 2] = m->m.v[k].z;
        }
    }
    if (triangles != nullptr) {
        for (std::size_t k = 0; k < m->m.t.size(); ++k) {
            for (int a = 0; a < 3; ++a) {
                triangles[3 * k + a] = m->m.t[k][a];
            }
        }
    }
    return LBM_OK;
}

// kind: 0 sphere (a = radius), 1 cylinder (a = radius, b = height), 2 cone (a = radius,
// b = height), 3 capsule (a..e = max diameter, nose radius, shoulder radius, afterbody half
// angle, top diameter).
int lbm_mesh_revolution(int32_t kind, double a, double b, double c, double d, double e, int32_t around,
                        double arc_step, lbm_mesh** out) {
    if (out == nullptr) {
        return geom_fail("null output");
    }
    std::vector<Segment> segments;
    const auto positive = [](double v) { return std::isfinite(v) && v > 0.0; };
    switch (kind) {
        case 0: // sphere_profile (primitives.cpp:292-295)
            if (!positive(a)) return geom_fail("sphere radius must be positive and finite");
            segments.push_back(arc({0.0, 0.0}, a, -kPi / 2.0, kPi / 2.0));
            break;
        case 1: // cylinder_profile (:297-303)
            if (!positive(a) || !positive(b)) return geom_fail("cylinder radius and height must be positive and finite");
            segments = {line({0.0, 0.0}, {a, 0.0}), line({a, 0.0}, {a, b}), line({a, b}, {0.0, b})};
            break;
        case 2: // cone_profile (:305-310)
            if (!positive(a) || !positive(b)) return geom_fail("cone radius and height must be positive and finite");
            segments = {line({0.0, 0.0}, {a, 0.0}), line({a, 0.0}, {0.0, b})};
            break;
        case 3: { // capsule_profile (:411-426), CapsuleShape::validate (:381-409)
            for (double v : {a, b, c, d}) {
                if (!positive(v)) return geom_fail("capsule parameters must be positive and finite");
            }
            if (!std::isfinite(e) || e < 0.0) return geom_fail("capsule top diameter must be non-negative and finite");
            if (!(d < kPi / 2.0)) return geom_fail("afterbody half angle must be below 90 degrees");
            if (!(c < 0.5 * a)) return geom_fail("shoulder radius must be below the maximum radius");
            if (!(b > 0.5 * a)) return geom_fail("nose radius must exceed the maximum radius");
            segments = capsule_segments(a, b, c, d, e);
            break;
        }
        default:
            return geom_fail("unknown solid of revolution");
    }
    auto* h = new (std::nothrow) lbm_mesh();
    if (h == nullptr) {
        return LBM_ENOMEM;
    }
    if (!revolution_mesh(segments, around, arc_step, h->m)) {
        delete h;
        return LBM_EINVAL;
    }
    *out = h;
    return LBM_OK;
}

// Wedge::mesh (primitives.cpp:590-604)
int lbm_mesh_wedge(double length, double half_angle, double span, lbm_mesh** out) {
    if (out == nullptr || !std::isfinite(length) || !(length > 0.0) || !std::isfinite(span) || !(span > 0.0) ||
        !std::isfinite(half_angle) || !(half_angle > 0.0) || !(half_angle < kPi / 2.0)) {
        return geom_fail("wedge: length and span positive, half angle in (0, 90) degrees");
    }
    auto* h = new (std::nothrow) lbm_mesh();
    if (h == nullptr) {
        return LBM_ENOMEM;
    }
    const double hh = length * std::tan(half_angle);
    const double s = 0.5 * span;
    h->m.v = {{0.0, -s, 0.0}, {0.0, s, 0.0}, {length, -s, hh}, {length, s, hh}, {length, -s, -hh}, {length, s, -hh}};
    h->m.t = {{0, 2, 3}, {0, 3, 1}, {0, 1, 5}, {0, 5, 4}, {4, 5, 3}, {4, 3, 2}, {0, 4, 2}, {1, 3, 5}};
    *out = h;
    return LBM_OK;
}

// box_mesh (primitives.cpp:557-571)
int lbm_mesh_box(double lx, double ly, double lz, lbm_mesh** out) {
    const auto positive = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (out == nullptr || !positive(lx) || !positive(ly) || !positive(lz)) {
        return geom_fail("box sizes must be positive and finite");
    }
    auto* h = new (std::nothrow) lbm_mesh();
    if (h == nullptr) {
        return LBM_ENOMEM;
    }
    for (int i = 0; i < 8; ++i) {
        h->m.v.push_back({(i & 1) != 0 ? 0.5 * lx : -0.5 * lx, (i & 2) != 0 ? 0.5 * ly : -0.5 * ly,
                          (i & 4) != 0 ? 0.5 * lz : -0.5 * lz});
    }
    h->m.t = {{0, 4, 6}, {0, 6, 2}, {1, 3, 7}, {1, 7, 5}, {0, 1, 5}, {0, 5, 4},
              {2, 6, 7}, {2, 7, 3}, {0, 2, 3}, {0, 3, 1}, {4, 5, 7}, {4, 7, 6}};
    *out = h;
    return LBM_OK;
}

// naca_wing (naca.cpp:184-257) of Naca4{max_camber, camber_position, thickness}.
int lbm_mesh_naca_wing(double max_camber, double camber_position, double thickness, double chord, double span,
                       int32_t n, int32_t span_divisions, lbm_mesh** out) {
    if (out == nullptr || !std::isfinite(chord) || !(chord > 0.0) || !std::isfinite(span) || !(span > 0.0)) {
        return geom_fail("naca_wing: chord and span must be positive and finite");
    }
    if (span_divisions < 1 || span_divisions > 100000 || n < 3 || n > 1000000) {
        return geom_fail("naca_wing: n in [3, 1000000] and span_divisions in [1, 100000]");
    }
    auto* h = new (std::nothrow) lbm_mesh();
    if (h == nullptr) {
        return LBM_ENOMEM;
    }
    naca_wing(Naca{max_camber, camber_position, thickness}, chord, span, n, span_divisions, h->m);
    *out = h;
    return LBM_OK;
}

// Naca4::section (naca.cpp:155-172) of Naca4{max_camber, camber_position, thickness}.
int lbm_naca_section(double max_camber, double camber_position, double thickness, int32_t n, double* points) {
    if (points == nullptr) {
        return geom_fail("naca_section: null output");
    }
    if (n < 3 || n > 1000000) {
        return geom_fail("Naca4::section: n must be in [3, 1000000]");
    }
    const std::vector<PP> section = naca_section(Naca{max_camber, camber_position, thickness}, n);
    for (std::size_t k = 0; k < section.size(); ++k) {
        points[2 * k] = section[k].r;
        points[2 * k + 1] = section[k].z;
    }
    return LBM_OK;
}

// polygon_mask (src/aero/lbm/solid_mask.cpp:45-73): even-odd crossings of the row through each
// node, node (x, y) solid when inside the polygon (x, z) given in lattice units.
int lbm_polygon_mask(uint64_t nx, uint64_t ny, const double* points, size_t n, uint8_t* solid) {
    if (solid == nullptr || (n > 0 && points == nullptr)) {
        return geom_fail("polygon_mask: null argument");
    }
    if (nx == 0 || ny == 0 || nx > (std::uint64_t{1} << 27U) / ny) {
        return geom_fail("solid mask: dimensions must be at least 1 and the node count at most geom::max_grid_cells");
    }
    std::fill(solid, solid + nx * ny, std::uint8_t{0});
    if (n < 3) {
        return LBM_OK;
    }
    for (std::uint64_t y = 0; y < ny; ++y) {
        const auto py = static_cast<double>(y);
        for (std::uint64_t x = 0; x < nx; ++x) {
            const auto px = static_cast<double>(x);
            bool inside = false;
            for (std::size_t a = 0, b = n - 1; a < n; b = a++) {
                const double ax = points[2 * a], az = points[2 * a + 1];
                const double bx = points[2 * b], bz = points[2 * b + 1];
                if ((az > py) != (bz > py)) {
                    const double crossing = ax + (py - az) * (bx - ax) / (bz - az);
                    if (px < crossing) {
                        inside = !inside;
                    }
                }
            }
            if (inside) {
                solid[x + nx * y] = 1;
            }
        }
    }
    return LBM_OK;
}

// Rigid transform of the vertices: v <- R (v - pivot) + translation, R row-major 3x3.
int lbm_mesh_transform(lbm_mesh* m, const double rotation[9], const double pivot[3], const double translation[3]) {
    if (m == nullptr || rotation == nullptr || pivot == nullptr || translation == nullptr) {
        return geom_fail("null argument");
    }
    for (V3& v : m->m.v) {
        const V3 p = v - V3{pivot[0], pivot[1], pivot[2]};
        v = {rotation[0] * p.x + rotation[1] * p.y + rotation[2] * p.z + translation[0],
             rotation[3] * p.x + rotation[4] * p.y + rotation[5] * p.z + translation[1],
             rotation[6] * p.x + rotation[7] * p.y + rotation[8] * p.z + translation[2]};
    }
    return LBM_OK;
}

int lbm_mesh_bounds(const lbm_mesh* m, double lo[3], double hi[3]) {
    if (m == nullptr) {
        return geom_fail("null mesh");
    }
    const Box b = bounds_of(m->m);
    lo[0] = b.min.x;
    lo[1] = b.min.y;
    lo[2] = b.min.z;
    hi[0] = b.max.x;
    hi[1] = b.max.y;
    hi[2] = b.max.z;
    return LBM_OK;
}

// voxelize (voxel_grid.cpp:81-102): solid where the winding number at the cell centre is at
// least `threshold`; `winding` (optional) receives the numbers.
int lbm_voxelize(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny, uint64_t nz,
                 double threshold, uint8_t* solid, double* winding) {
    if (m == nullptr || origin == nullptr || solid == nullptr) {
        return geom_fail("voxelize: null argument");
    }
    if (!grid_ok(origin, spacing, nx, ny, nz)) {
        return LBM_EINVAL;
    }
    if (!std::isfinite(threshold)) {
        return geom_fail("voxelize: threshold must be finite");
    }
    const Tree tree(m->m);
    parallel_slabs(nz, [&](std::size_t k) {
        for (std::size_t j = 0; j < ny; ++j) {
            for (std::size_t i = 0; i < nx; ++i) {
                const std::size_t c = i + nx * (j + ny * k);
                const double w = tree.winding(centre_of(origin, spacing, i, j, k));
                if (winding != nullptr) {
                    winding[c] = w;
                }
                solid[c] = static_cast<std::uint8_t>(w >= threshold);
            }
        }
    });
    return LBM_OK;
}

// signed_distance_band (voxel_grid.cpp:104-138): exact distance to the surface within
// band_cells * spacing, negative in solid voxels; `in_band` optional.
int lbm_signed_distance_band(const lbm_mesh* m, const double origin[3], double spacing, uint64_t nx, uint64_t ny,
                             uint64_t nz, const uint8_t* solid, double band_cells, double* distance, uint8_t* in_band) {
    if (m == nullptr || origin == nullptr || solid == nullptr || distance == nullptr) {
        return geom_fail("signed_distance_band: null argument");
    }
    if (!std::isfinite(band_cells) || !(band_cells > 0.0)) {
        return geom_fail("signed_distance_band: band must be positive");
    }
    if (!grid_ok(origin, spacing, nx, ny, nz)) {
        return LBM_EINVAL;
    }
    const double band = band_cells * spacing;
    if (!std::isfinite(band)) {
        return geom_fail("signed_distance_band: band is not finite");
    }
    const Tree tree(m->m);
    parallel_slabs(nz, [&](std::size_t k) {
        for (std::size_t j = 0; j < ny; ++j) {
            for (std::size_t i = 0; i < nx; ++i) {
                const std::size_t c = i + nx * (j + ny * k);
                const double sign = solid[c] != 0 ? -1.0 : 1.0;
                const double d = tree.distance(centre_of(origin, spacing, i, j, k), band);
                if (in_band != nullptr) {
                    in_band[c] = static_cast<std::uint8_t>(d < band);
                }
                distance[c] = sign * d;
            }
        }
    });
    return LBM_OK;
}

// Winding number at arbitrary points (tests and the threshold-adjacent count).
int lbm_winding_numbers(const lbm_mesh* m, const double* points, size_t n, double* out) {
    if (m == nullptr || points == nullptr || out == nullptr) {
        return geom_fail("null argument");
    }
    const Tree tree(m->m);
    for (std::size_t k = 0; k < n; ++k) {
        out[k] = tree.winding({points[3 * k], points[3 * k + 1], points[3 * k + 2]});
    }
    return LBM_OK;
}

} // extern "C"
