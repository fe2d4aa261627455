// C ABI of libdmap_amd.so (include/dmap_amd.h): contexts, resident tiles,
// objects, and the intersection entry points. Host-side set-up of each
// triangle runs in double here, once per object; the per-ray work runs in
// the HIP kernels of pyramid.hip and traverse.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace
{
thread_local std::string g_error;
}

int dmap_amd_internal::fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

namespace
{

// ------------------------------------------------------------ host math

struct V2
{
    double x, y;
};
struct V3
{
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline double comp(V3 a, int q) { return q == 0 ? a.x : q == 1 ? a.y : a.z; }

float f_down(double x)
{
    float f = (float)x;
    return (double)f <= x ? f : std::nextafter(f, -INFINITY);
}
float f_up(double x)
{
    float f = (float)x;
    return (double)f >= x ? f : std::nextafter(f, INFINITY);
}

// Affine form over two cell symbols (affine_form.h:40-61).
struct AF
{
    double c = 0, a = 0, b = 0, k = 0;
    double dev() const { return std::abs(a) + std::abs(b) + k; }
};
AF mul(const AF &x, const AF &r)
{
    AF o;
    double u = x.dev(), v = r.dev();
    o.c = x.c * r.c;
    o.a = x.c * r.a + r.c * x.a;
    o.b = x.c * r.b + r.c * x.b;
    o.k = std::abs(r.c) * x.k + std::abs(x.c) * r.k + u * v;
    return o;
}
AF scale(double s, const AF &x) { return {x.c * s, x.a * s, x.b * s, x.k * std::abs(s)}; }
AF add(const AF &x, const AF &y) { return {x.c + y.c, x.a + y.a, x.b + y.b, x.k + y.k}; }

// AffineVec3::normalize (affine_form.cpp:64-95). The squared length's
// remainder bounds the cross term 2 sum_q c_q k_q e_q by 2 sum_q |c_q| k_q,
// its maximum over independent e_q in [-1, 1]; affine_form.cpp:83 writes
// 2 |c . k|, which equals it when every c_q k_q has one sign and is too
// small otherwise. Every caller here passes k = 0, where both are 0.
void normalize(AF N[3])
{
    double dev[3], cc = 0, cu = 0, cv = 0, ck = 0, dd = 0;
    for (int q = 0; q < 3; ++q) {
        dev[q] = N[q].dev();
        dd += dev[q] * dev[q];
    }
    for (int q = 0; q < 3; ++q) {
        cc += N[q].c * N[q].c;
        cu += N[q].c * N[q].a;
        cv += N[q].c * N[q].b;
        ck += std::abs(N[q].c) * N[q].k;
    }
    double offset = 0.5 * dd;
    AF l{cc + offset, 2.0 * cu, 2.0 * cv, 2.0 * ck + offset};
    double lo = l.c - l.dev(), hi = l.c + l.dev();
    if (lo <= 0.0) {
        for (int q = 0; q < 3; ++q)
            N[q] = AF{0, 0, 0, 1};
        return;
    }
    auto g = [](double x) { return 1.0 / std::sqrt(x); };
    double alpha = -0.5 * g(hi) * g(hi) * g(hi);
    double beta = 0.5 * (g(lo) + g(hi) - alpha * (lo + hi));
    double delta = 0.5 * std::abs(g(lo) - g(hi) - alpha * (lo - hi));
    AF rs{alpha * l.c + beta, alpha * l.a, alpha * l.b, std::abs(alpha) * l.k + delta};
    for (int q = 0; q < 3; ++q)
        N[q] = mul(rs, N[q]);
}

// The host-side description of one triangle, in double.
struct HostTri
{
    V3 p0, p1, p2, q0, e1, e2, m0, Mu, Mv;
    V2 t0, t1, t2;
    double J[4];
};

bool make_tri(const double *pos, const double *nrm, const double *uv, HostTri &t)
{
    // BaseTriangle, general chart (displaced_surface.cpp:54-91).
    t.p0 = {pos[0], pos[1], pos[2]};
    t.p1 = {pos[3], pos[4], pos[5]};
    t.p2 = {pos[6], pos[7], pos[8]};
    V3 n0 = {nrm[0], nrm[1], nrm[2]}, n1 = {nrm[3], nrm[4], nrm[5]}, n2 = {nrm[6], nrm[7], nrm[8]};
    t.t0 = {uv[0], uv[1]};
    t.t1 = {uv[2], uv[3]};
    t.t2 = {uv[4], uv[5]};
    double a = t.t1.x - t.t0.x, b = t.t2.x - t.t0.x, c = t.t1.y - t.t0.y, d = t.t2.y - t.t0.y;
    double det = a * d - b * c;
    if (!(std::abs(det) > 1e-30))
        return false;
    t.J[0] = d / det;
    t.J[1] = -b / det;
    t.J[2] = -c / det;
    t.J[3] = a / det;
    V3 eb1 = t.p1 - t.p0, eb2 = t.p2 - t.p0, Mb1 = n1 - n0, Mb2 = n2 - n0;
    t.e1 = t.J[0] * eb1 + t.J[2] * eb2;
    t.e2 = t.J[1] * eb1 + t.J[3] * eb2;
    t.Mu = t.J[0] * Mb1 + t.J[2] * Mb2;
    t.Mv = t.J[1] * Mb1 + t.J[3] * Mb2;
    double c0 = -(t.J[0] * t.t0.x + t.J[1] * t.t0.y), c1 = -(t.J[2] * t.t0.x + t.J[3] * t.t0.y);
    t.q0 = t.p0 + (c0 * eb1 + c1 * eb2);
    t.m0 = n0 + (c0 * Mb1 + c1 * Mb2);
    return true;
}

// Fills a triangle's device record. Returns false for a degenerate frame or
// footprint; *root_level_out receives the root level found (also when it
// is too high for the walk, MAX_ROOT_LEVEL, which then returns false).
bool fill_record(const HostTri &h, int n_leaf, TriRecord &R, int *root_level_out)
{
    *root_level_out = -1;
    std::memset(&R, 0, sizeof R);
    const V3 *vs[6] = {&h.q0, &h.e1, &h.e2, &h.m0, &h.Mu, &h.Mv};
    double *ds[6] = {R.q0, R.e1, R.e2, R.m0, R.Mu, R.Mv};
    for (int v = 0; v < 6; ++v)
        for (int q = 0; q < 3; ++q)
            ds[v][q] = comp(*vs[v], q);
    // TangentFrame (node_bounds.cpp:13-23): to_tangent = [e1 e2 ng]^-1.
    V3 ng = cross(h.e1, h.e2);
    double ngl = norm(ng);
    if (!(ngl > 0.0))
        return false;
    ng = (1.0 / ngl) * ng;
    double m[3][3];
    for (int r = 0; r < 3; ++r) {
        m[r][0] = comp(h.e1, r);
        m[r][1] = comp(h.e2, r);
        m[r][2] = comp(ng, r);
    }
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1], c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2],
           c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    if (!(std::abs(det) > 0.0))
        return false;
    double inv = 1.0 / det;
    double T[9] = {c00 * inv,
                   (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv,
                   (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv,
                   c01 * inv,
                   (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv,
                   (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv,
                   c02 * inv,
                   (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv,
                   (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv};
    for (int q = 0; q < 9; ++q) {
        R.T[q] = T[q];
        R.f_T[q] = (float)T[q];
    }
    for (int q = 0; q < 4; ++q)
        R.J[q] = h.J[q];
    R.t0[0] = h.t0.x;
    R.t0[1] = h.t0.y;
    // UvTriangle (uv_clip.cpp:10-25).
    V2 t[3] = {h.t0, h.t1, h.t2};
    double signed_area = 0.5 * ((h.t1.x - h.t0.x) * (h.t2.y - h.t0.y) - (h.t1.y - h.t0.y) * (h.t2.x - h.t0.x));
    double orient = signed_area >= 0.0 ? 1.0 : -1.0;
    for (int k = 0; k < 3; ++k) {
        V2 e = {t[(k + 1) % 3].x - t[k].x, t[(k + 1) % 3].y - t[k].y};
        double nx = orient * -e.y, ny = orient * e.x;
        R.dn_d[2 * k] = nx;
        R.dn_d[2 * k + 1] = ny;
        R.dd_d[k] = nx * t[k].x + ny * t[k].y;
        R.dn[2 * k] = (float)nx;
        R.dn[2 * k + 1] = (float)ny;
        R.dd[k] = (float)R.dd_d[k];
    }
    R.dlo_d[0] = std::min(std::min(h.t0.x, h.t1.x), h.t2.x);
    R.dlo_d[1] = std::min(std::min(h.t0.y, h.t1.y), h.t2.y);
    R.dhi_d[0] = std::max(std::max(h.t0.x, h.t1.x), h.t2.x);
    R.dhi_d[1] = std::max(std::max(h.t0.y, h.t1.y), h.t2.y);
    for (int q = 0; q < 2; ++q) {
        R.dlo[q] = (float)R.dlo_d[q];
        R.dhi[q] = (float)R.dhi_d[q];
    }
    // footprint_roots (descent_sampler.cpp:66-83) and the root cells of
    // the DisplacedIntersector ctor (displaced_intersector.cpp:139-149).
    int level = 0;
    int64_t i0 = 0, j0 = 0;
    while (true) {
        double w = std::ldexp(1.0, level) / n_leaf;
        int64_t ilo = (int64_t)std::floor(R.dlo_d[0] / w), ihi = (int64_t)std::ceil(R.dhi_d[0] / w) - 1;
        int64_t jlo = (int64_t)std::floor(R.dlo_d[1] / w), jhi = (int64_t)std::ceil(R.dhi_d[1] / w) - 1;
        if (ihi - ilo <= 1 && jhi - jlo <= 1) {
            i0 = ilo;
            j0 = jlo;
            break;
        }
        if (++level > 40)
            return false;
    }
    double w = std::ldexp(1.0, level) / n_leaf;
    int64_t i_lo = (int64_t)std::floor(R.dlo_d[0] / w), i_hi = (int64_t)std::ceil(R.dhi_d[0] / w) - 1;
    int64_t j_lo = (int64_t)std::floor(R.dlo_d[1] / w), j_hi = (int64_t)std::ceil(R.dhi_d[1] / w) - 1;
    if (!(i_lo == i0 && j_lo == j0 && i_hi <= i0 + 1 && j_hi <= j0 + 1))
        return false;
    if (std::abs(i_lo) > (1 << 28) || std::abs(j_lo) > (1 << 28))
        return false;
    // The walk keeps four trail bits for each level from the root block's
    // down to level 0 in one 64-bit word: at most 16 levels, so a root
    // level of at most MAX_ROOT_LEVEL = 15.
    *root_level_out = level;
    if (level > MAX_ROOT_LEVEL)
        return false;
    R.root_level = level;
    R.n_roots = 0;
    for (int64_t j = j_lo; j <= j_hi; ++j)
        for (int64_t i = i_lo; i <= i_hi; ++i) {
            R.root_i[R.n_roots] = (int)i;
            R.root_j[R.n_roots] = (int)j;
            ++R.n_roots;
        }
    // A constant interpolated normal gives one normalized form for every
    // cell (node_bounds.cpp:100-103 with Mu = Mv = 0).
    R.const_normal = (h.Mu.x == 0.0 && h.Mu.y == 0.0 && h.Mu.z == 0.0 && h.Mv.x == 0.0 && h.Mv.y == 0.0 &&
                      h.Mv.z == 0.0)
                         ? 1
                         : 0;
    {
        AF N[3] = {AF{h.m0.x, 0, 0, 0}, AF{h.m0.y, 0, 0, 0}, AF{h.m0.z, 0, 0, 0}};
        normalize(N);
        double kmax = 0;
        for (int q = 0; q < 3; ++q) {
            R.nhat_d[q] = N[q].c;
            R.nhat[q] = (float)N[q].c;
            kmax = std::max(kmax, N[q].k);
        }
        R.nhat_kd = kmax;
        R.nhat_k = f_up(kmax);
        for (int r = 0; r < 3; ++r) {
            double c = 0, k = 0;
            for (int q = 0; q < 3; ++q) {
                c += T[3 * r + q] * N[q].c;
                k += std::abs(T[3 * r + q]) * N[q].k;
            }
            R.mc_d[r] = c;
            R.mk_d[r] = k;
            R.mc[r] = (float)c;
            R.mk[r] = f_up((k + std::abs(c - (double)R.mc[r])) * (1.0 + 0x1p-50));
            // the double sum's rounding is inside the 2^-50 bump
            R.mck[r] = f_up((std::abs((double)R.mc[r]) + (double)R.mk[r]) * (1.0 + 0x1p-50));
        }
    }
    for (int q = 0; q < 3; ++q) {
        R.f_m0[q] = (float)comp(h.m0, q);
        R.f_Mu[q] = (float)comp(h.Mu, q);
        R.f_Mv[q] = (float)comp(h.Mv, q);
    }
    // residual_tolerance = 1e-10 x mean edge (displaced_intersector.cpp:150,
    // displaced_surface.cpp:168-171).
    double me = (norm(h.p1 - h.p0) + norm(h.p2 - h.p0) + norm(h.p2 - h.p1)) / 3.0;
    R.residual_tolerance = 1e-10 * me;
    return true;
}

// triangle_aabb (node_bounds.cpp:251-269), then widened for fp32 rays as
// rtc_bounds does (displaced_intersector.cpp:454-464).
void triangle_box(const HostTri &tri, double h_min, double h_max, float out[6])
{
    const V2 t[3] = {tri.t0, tri.t1, tri.t2};
    AF h{0.5 * (h_min + h_max), 0, 0, 0.5 * (h_max - h_min)};
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; ++k) {
        V2 c = {0.5 * t[k].x + 0.25 * t[(k + 1) % 3].x + 0.25 * t[(k + 2) % 3].x,
                0.5 * t[k].y + 0.25 * t[(k + 1) % 3].y + 0.25 * t[(k + 2) % 3].y};
        V2 du = {0.25 * (t[(k + 1) % 3].x - t[k].x), 0.25 * (t[(k + 1) % 3].y - t[k].y)};
        V2 dv = {0.25 * (t[(k + 2) % 3].x - t[k].x), 0.25 * (t[(k + 2) % 3].y - t[k].y)};
        V3 P = tri.q0 + c.x * tri.e1 + c.y * tri.e2;
        V3 Pa = du.x * tri.e1 + du.y * tri.e2, Pb = dv.x * tri.e1 + dv.y * tri.e2;
        V3 M = tri.m0 + c.x * tri.Mu + c.y * tri.Mv;
        V3 Ma = du.x * tri.Mu + du.y * tri.Mv, Mb = dv.x * tri.Mu + dv.y * tri.Mv;
        AF N[3], Pf[3];
        for (int q = 0; q < 3; ++q) {
            N[q] = AF{comp(M, q), comp(Ma, q), comp(Mb, q), 0};
            Pf[q] = AF{comp(P, q), comp(Pa, q), comp(Pb, q), 0};
        }
        normalize(N);
        for (int q = 0; q < 3; ++q) {
            AF S = add(Pf[q], mul(h, N[q]));
            lo[q] = std::min(lo[q], S.c - S.dev());
            hi[q] = std::max(hi[q], S.c + S.dev());
        }
    }
    double mag = 0;
    for (int q = 0; q < 3; ++q)
        mag = std::max(mag, std::max(std::abs(lo[q]), std::abs(hi[q])));
    double pad = 8.0 * (double)FLT_EPSILON * (1.0 + mag);
    for (int q = 0; q < 3; ++q) {
        out[q] = f_down(lo[q] - pad);
        out[3 + q] = f_up(hi[q] + pad);
    }
}

bool is_pow2(int n) { return n >= 1 && (n & (n - 1)) == 0; }

// Top-level BVH over the triangles' boxes, a complete binary tree in heap
// order: the triangles of a subtree are split at their median box centre
// along the longest axis of the centres' spread, the lower (m + 1) / 2 going
// to child 2i, so a subtree of capacity C (a power of two) never receives
// more than C triangles. It stands in for the reference's embree scene over
// the same boxes (ObjectIntersector::build_bvh, displaced_intersector.cpp:540-554).
// Writes the subtree's box (empty: lo > hi) to blo, bhi. Nodes are built
// with their children's boxes as lo / hi and stored as centre / radius
// (bvh_centre_radius).
struct BvhBuildNode
{
    float lo[2][3], hi[2][3];
    int axis, empty;
};
void build_bvh_node(std::vector<BvhBuildNode> &nodes, std::vector<int> &leaf, const std::vector<float> &boxes,
                    std::vector<int> &idx, uint32_t node, int lo, int hi, uint32_t P, float blo[3], float bhi[3])
{
    for (int a = 0; a < 3; ++a) {
        blo[a] = INFINITY;
        bhi[a] = -INFINITY;
    }
    if (node >= P) {
        leaf[node - P] = hi > lo ? idx[lo] : -1;
        if (hi > lo)
            for (int a = 0; a < 3; ++a) {
                blo[a] = boxes[6 * idx[lo] + a];
                bhi[a] = boxes[6 * idx[lo] + 3 + a];
            }
        return;
    }
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int q = lo; q < hi; ++q) {
        const float *b = &boxes[6 * idx[q]];
        for (int a = 0; a < 3; ++a) {
            float c = 0.5f * (b[a] + b[3 + a]);
            clo[a] = std::min(clo[a], c);
            chi[a] = std::max(chi[a], c);
        }
    }
    int axis = 0;
    if (hi - lo >= 2)
        for (int a = 1; a < 3; ++a)
            if (chi[a] - clo[a] > chi[axis] - clo[axis])
                axis = a;
    int mid = lo + (hi - lo + 1) / 2;
    if (hi - lo >= 2)
        std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi, [&](int x, int y) {
            return boxes[6 * x + axis] + boxes[6 * x + 3 + axis] < boxes[6 * y + axis] + boxes[6 * y + 3 + axis];
        });
    BvhBuildNode nd{};
    nd.axis = axis;
    nd.empty = (mid > lo ? 0 : 1) | (hi > mid ? 0 : 2);
    build_bvh_node(nodes, leaf, boxes, idx, 2 * node, lo, mid, P, nd.lo[0], nd.hi[0]);
    build_bvh_node(nodes, leaf, boxes, idx, 2 * node + 1, mid, hi, P, nd.lo[1], nd.hi[1]);
    for (int c = 0; c < 2; ++c)
        for (int a = 0; a < 3; ++a) {
            if (nd.empty & (1 << c)) {
                nd.lo[c][a] = 0.f;
                nd.hi[c][a] = 0.f;
                continue;
            }
            blo[a] = std::min(blo[a], nd.lo[c][a]);
            bhi[a] = std::max(bhi[a], nd.hi[c][a]);
        }
    nodes[node] = nd;
}

// The device form of a box [lo, hi]: fp32 centre c and radius r per axis
// with [c - r, c + r] holding [lo, hi]. The centre is rounded to fp32, the
// radius (formed in double) rounded up and then raised by one more ulp,
// which covers the rounding of the double difference.
void bvh_centre_radius(const float lo[3], const float hi[3], float c[3], float r[3])
{
    for (int a = 0; a < 3; ++a) {
        c[a] = (float)(0.5 * ((double)lo[a] + (double)hi[a]));
        const double rd = std::max((double)hi[a] - (double)c[a], (double)c[a] - (double)lo[a]);
        float rf = (float)rd;
        if ((double)rf < rd)
            rf = std::nextafter(rf, INFINITY);
        r[a] = std::nextafter(rf, INFINITY);
    }
}

// Leaves of the heap-ordered BVH over n triangles: the power of two >= n.
int bvh_leaves_for(int n_tri)
{
    int P = 1;
    while (P < n_tri)
        P *= 2;
    return P;
}

} // namespace

static_assert(sizeof(dmap_node_record) == sizeof(NodeRecordDev) && sizeof(NodeRecordDev) == 48,
              "dmap_node_record and NodeRecordDev share one layout");

namespace
{

// Runs body(k) for k in [0, n) on up to 16 host threads (chunks of
// consecutive k); the per-triangle set-up of large meshes.
template <class F> void parallel_for(int n, F body)
{
    const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int threads = n < 4096 ? 1 : hw;
    if (threads == 1) {
        for (int k = 0; k < n; ++k)
            body(k);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi)
            pool.emplace_back([=] {
                for (int k = lo; k < hi; ++k)
                    body(k);
            });
    }
    for (auto &th : pool)
        th.join();
}

} // namespace

extern "C" {

const char *dmap_last_error(void) { return g_error.c_str(); }
const char *dmap_version(void) { return "dmap_amd 0.1 gfx950"; }

// Scratch of dmap_intersect_closest_noalloc: the ray-queue head (256 B), the
// ray reorder buffers (sort keys, permutation, radix sort temporary storage;
// about 24 B per ray) and the raw records (32 B per ray) of one call of at
// most NOALLOC_MAX_RAYS rays.
// Every device the library loads on holds one copy for the life of the
// process. Host code must not use `scratch` itself as an address: its
// device address comes from hipGetSymbolAddress, per device, in
// dmap_ctx_create, and is kept in dmap_ctx::scratch.
constexpr int64_t NOALLOC_MAX_RAYS = (int64_t)1 << 20;
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

int dmap_ctx_create(int device, dmap_ctx **out)
{
    if (!out)
        return fail(DMAP_ERR_INVALID, "dmap_ctx_create: out is NULL");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(DMAP_ERR_INVALID, "dmap_ctx_create: device %d of %d", device, count);
    HIP_TRY(hipSetDevice(device));
    void *sp = nullptr;
    hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp)
        return fail(DMAP_ERR_HIP, "dmap_ctx_create: scratch address on device %d: %s", device, hipGetErrorString(e));
    // The workspace pool: device memory of this device, kept once used
    // (the default pool returns it at each synchronization, and a call
    // after a synchronization then maps it again: 2.54 against 2.49 ms per
    // 4M-ray call, and about 0.2 ms per path-traced frame, which
    // synchronizes once per bounce).
    hipMemPoolProps props;
    std::memset(&props, 0, sizeof props);
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device;
    hipMemPool_t pool = nullptr;
    if ((e = hipMemPoolCreate(&pool, &props)) != hipSuccess)
        return fail(DMAP_ERR_HIP, "dmap_ctx_create: workspace pool on device %d: %s", device, hipGetErrorString(e));
    uint64_t keep = UINT64_MAX;
    if ((e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) != hipSuccess) {
        (void)hipMemPoolDestroy(pool);
        return fail(DMAP_ERR_HIP, "dmap_ctx_create: workspace pool threshold: %s", hipGetErrorString(e));
    }
    dmap_ctx *c = new dmap_ctx;
    c->device = device;
    c->scratch = sp;
    c->pool = pool;
    *out = c;
    return DMAP_OK;
}

void dmap_ctx_destroy(dmap_ctx *ctx)
{
    if (!ctx)
        return;
    if (ctx->pool && hipSetDevice(ctx->device) == hipSuccess) {
        // Workspaces freed on a stream go back to the pool when that work
        // ends: wait for it before the pool goes.
        (void)hipDeviceSynchronize();
        (void)hipMemPoolDestroy(ctx->pool);
    }
    delete ctx;
}

int dmap_device_alloc(dmap_ctx *ctx, int64_t bytes, void **ptr)
{
    if (!ctx || !ptr || bytes < 0)
        return fail(DMAP_ERR_INVALID, "dmap_device_alloc: bad argument");
    *ptr = nullptr;
    if (bytes == 0)
        return DMAP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(ptr, (size_t)bytes);
    if (e != hipSuccess)
        return fail(DMAP_ERR_MEMORY, "dmap_device_alloc: %lld bytes: %s", (long long)bytes, hipGetErrorString(e));
    return DMAP_OK;
}

int dmap_device_free(dmap_ctx *ctx, void *ptr)
{
    if (!ctx)
        return fail(DMAP_ERR_INVALID, "dmap_device_free: NULL context");
    if (!ptr)
        return DMAP_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipFree(ptr));
    return DMAP_OK;
}

int dmap_memcpy(void *dst, const void *src, int64_t bytes, int kind, void *stream)
{
    if (bytes < 0 || kind < 0 || kind > 2 || (bytes > 0 && (!dst || !src)))
        return fail(DMAP_ERR_INVALID, "dmap_memcpy: bad argument");
    if (bytes == 0)
        return DMAP_OK;
    const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
    if (stream) {
        HIP_TRY(hipMemcpyAsync(dst, src, (size_t)bytes, k[kind], (hipStream_t)stream));
    } else {
        HIP_TRY(hipMemcpy(dst, src, (size_t)bytes, k[kind]));
    }
    return DMAP_OK;
}

int dmap_stream_synchronize(void *stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return DMAP_OK;
}

static void tile_free(dmap_tile *t)
{
    if (!t)
        return;
    (void)hipFree(t->texels);
    (void)hipFree(t->nodes);
    (void)hipFree(t->mm);
    (void)hipFree(t->tp);
    (void)hipFree(t->rr);
    (void)hipFree(t->scratch);
    delete t;
}

static int tile_build(dmap_tile *t, hipStream_t stream)
{
    HIP_TRY(hipSetDevice(t->ctx->device));
    PyramidOut P{t->mm, t->tp, t->rr, t->n, t->n_levels, t->strength, t->offset};
    hipError_t e = t->build == 0 ? launch_build_fold(P, t->nodes, t->scratch, stream)
                                 : launch_build_direct(P, t->nodes, stream);
    HIP_TRY(e);
    return DMAP_OK;
}

// The root cell's [h_min, h_max], served above the root; synchronizes `stream`.
static int read_root(dmap_tile *t, hipStream_t stream)
{
    float2 root;
    HIP_TRY(hipMemcpyAsync(&root, t->mm + level_base(1), sizeof root, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    t->root_hmin = root.x;
    t->root_hmax = root.y;
    return DMAP_OK;
}

int dmap_tile_create(dmap_ctx *ctx, const float *texels, int n, int texels_on_device, int repeat, double strength,
                     double offset, int build, void *stream_, dmap_tile **out)
{
    if (!ctx || !texels || !out)
        return fail(DMAP_ERR_INVALID, "dmap_tile_create: NULL argument");
    if (!is_pow2(n) || n > 16384)
        return fail(DMAP_ERR_INVALID, "dmap_tile_create: the tile needs 2^L texels per side (<= 16384), got %d", n);
    if (build != 0 && build != 1)
        return fail(DMAP_ERR_INVALID, "dmap_tile_create: unknown pyramid build %d, expected 0 fold or 1 direct", build);
    if (!std::isfinite(strength) || !std::isfinite(offset))
        return fail(DMAP_ERR_INVALID, "dmap_tile_create: non-finite strength or offset");
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(ctx->device));
    dmap_tile *t = new dmap_tile;
    t->ctx = ctx;
    t->n = n;
    t->n_levels = 1;
    while ((1 << t->n_levels) <= n)
        ++t->n_levels;
    t->repeat = repeat ? 1 : 0;
    t->build = build;
    t->strength = strength;
    t->offset = offset;
    t->cells = pyramid_cells((uint32_t)n);
    size_t nn = (size_t)n * n, nodes = (size_t)(n + 1) * (n + 1);
    size_t scratch = fold_scratch_doubles(n);
    hipError_t e;
    if ((e = hipMalloc(&t->texels, nn * 4)) != hipSuccess || (e = hipMalloc(&t->nodes, nodes * 4)) != hipSuccess ||
        (e = hipMalloc(&t->mm, t->cells * 8)) != hipSuccess || (e = hipMalloc(&t->tp, t->cells * 16)) != hipSuccess ||
        (e = hipMalloc(&t->rr, t->cells * 8)) != hipSuccess ||
        (e = hipMalloc(&t->scratch, scratch * 8)) != hipSuccess) {
        tile_free(t);
        return fail(DMAP_ERR_MEMORY, "dmap_tile_create: device allocation failed (%s)", hipGetErrorString(e));
    }
    t->bytes = (int64_t)(nn * 4 + nodes * 4 + t->cells * 32 + scratch * 8);
    e = hipMemcpyAsync(t->texels, texels, nn * 4, texels_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                       stream);
    if (e == hipSuccess)
        e = launch_close_tile(t->texels, t->nodes, n, t->repeat, stream);
    if (e != hipSuccess) {
        tile_free(t);
        return fail(DMAP_ERR_HIP, "dmap_tile_create: %s", hipGetErrorString(e));
    }
    int rc = tile_build(t, stream);
    if (rc == DMAP_OK)
        rc = read_root(t, stream);
    if (rc != DMAP_OK) {
        tile_free(t);
        return rc;
    }
    *out = t;
    return DMAP_OK;
}

int dmap_tile_set_texels(dmap_tile *tile, const float *texels, int texels_on_device, void *stream_)
{
    if (!tile || !texels)
        return fail(DMAP_ERR_INVALID, "dmap_tile_set_texels: NULL argument");
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(tile->ctx->device));
    size_t nn = (size_t)tile->n * tile->n;
    HIP_TRY(hipMemcpyAsync(tile->texels, texels, nn * 4,
                           texels_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_close_tile(tile->texels, tile->nodes, tile->n, tile->repeat, stream));
    int rc = tile_build(tile, stream);
    return rc == DMAP_OK ? read_root(tile, stream) : rc;
}

int dmap_tile_rebuild(dmap_tile *tile, void *stream)
{
    if (!tile)
        return fail(DMAP_ERR_INVALID, "dmap_tile_rebuild: NULL tile");
    return tile_build(tile, (hipStream_t)stream);
}

int dmap_tile_info(const dmap_tile *tile, int *n_leaf, int *n_levels, int64_t *device_bytes)
{
    if (!tile)
        return fail(DMAP_ERR_INVALID, "dmap_tile_info: NULL tile");
    if (n_leaf)
        *n_leaf = tile->n;
    if (n_levels)
        *n_levels = tile->n_levels;
    if (device_bytes)
        *device_bytes = tile->bytes;
    return DMAP_OK;
}

int dmap_tile_read_level(const dmap_tile *tile, int level, float *out)
{
    if (!tile || !out)
        return fail(DMAP_ERR_INVALID, "dmap_tile_read_level: NULL argument");
    if (level < 0 || level >= tile->n_levels)
        return fail(DMAP_ERR_INVALID, "dmap_tile_read_level: level %d of %d", level, tile->n_levels);
    HIP_TRY(hipSetDevice(tile->ctx->device));
    uint32_t m = (uint32_t)tile->n >> level;
    size_t cnt = (size_t)m * m, base = level_base(m);
    std::vector<float2> mm(cnt), rr(cnt);
    std::vector<float4> tp(cnt);
    HIP_TRY(hipMemcpy(mm.data(), tile->mm + base, cnt * 8, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(tp.data(), tile->tp + base, cnt * 16, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rr.data(), tile->rr + base, cnt * 8, hipMemcpyDeviceToHost));
    for (uint32_t j = 0; j < m; ++j)
        for (uint32_t i = 0; i < m; ++i) {
            size_t src = morton2(i, j), dst = (size_t)j * m + i;
            out[0 * cnt + dst] = tp[src].x;
            out[1 * cnt + dst] = tp[src].y;
            out[2 * cnt + dst] = tp[src].z;
            out[3 * cnt + dst] = tp[src].w;
            out[4 * cnt + dst] = rr[src].x;
            out[5 * cnt + dst] = rr[src].y;
            out[6 * cnt + dst] = mm[src].x;
            out[7 * cnt + dst] = mm[src].y;
        }
    return DMAP_OK;
}

int dmap_tile_read_nodes(const dmap_tile *tile, float *out)
{
    if (!tile || !out)
        return fail(DMAP_ERR_INVALID, "dmap_tile_read_nodes: NULL argument");
    HIP_TRY(hipSetDevice(tile->ctx->device));
    size_t nodes = (size_t)(tile->n + 1) * (tile->n + 1);
    HIP_TRY(hipMemcpy(out, tile->nodes, nodes * 4, hipMemcpyDeviceToHost));
    return DMAP_OK;
}

void dmap_tile_destroy(dmap_tile *tile)
{
    if (tile)
        (void)hipSetDevice(tile->ctx->device);
    tile_free(tile);
}

static PyramidView view_of(const dmap_tile *t)
{
    PyramidView v;
    v.mm = t->mm;
    v.tp = t->tp;
    v.rr = t->rr;
    v.nodes = t->nodes;
    v.n_leaf = t->n;
    v.n_levels = t->n_levels;
    v.repeat = t->repeat;
    v.scale = (float)t->strength;
    v.offset = (float)t->offset;
    v.scale_d = t->strength;
    v.offset_d = t->offset;
    v.root_hmin = t->root_hmin;
    v.root_hmax = t->root_hmax;
    return v;
}

static void object_free(dmap_object *o)
{
    if (!o)
        return;
    (void)hipFree(o->tris);
    (void)hipFree(o->boxes);
    (void)hipFree(o->bvh);
    delete o;
}

// Objects of at least this many triangles get the device BVH builder when
// the caller leaves the choice to the library (DMAP_BVH_AUTO).
constexpr int DEVICE_BVH_MIN_TRIANGLES = 1 << 14;

// The triangles' height ranges over the tile's current pyramid, their
// boxes, the top-level BVH and the object's device view (triangle_height_range,
// triangle_aabb, node_bounds.cpp:218-269; ObjectIntersector::build_bvh,
// displaced_intersector.cpp:540-576). Synchronizes the device.
static int object_bounds(dmap_object *o, const std::vector<HostTri> &tris)
{
    const int n_tri = o->n_tri;
    const dmap_tile *tile = o->tile;
    hipError_t e;
    float *h_range = nullptr;
    e = hipMalloc(&h_range, sizeof(float) * 2 * n_tri);
    if (e == hipSuccess)
        e = hipMemcpy(o->tris, o->host.data(), sizeof(TriRecord) * n_tri, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_height_range(view_of(tile), o->tris, n_tri, h_range, nullptr);
    o->h_range_host.resize(2 * (size_t)n_tri);
    if (e == hipSuccess)
        e = hipMemcpy(o->h_range_host.data(), h_range, sizeof(float) * 2 * n_tri, hipMemcpyDeviceToHost);
    (void)hipFree(h_range);
    if (e != hipSuccess)
        return fail(DMAP_ERR_HIP, "object bounds: %s", hipGetErrorString(e));
    o->boxes_host.resize(6 * (size_t)n_tri);
    for (int k = 0; k < n_tri; ++k)
        if (!(o->h_range_host[2 * k] <= o->h_range_host[2 * k + 1]))
            return fail(DMAP_ERR_INVALID, "object bounds: triangle %d's footprint holds no cell", k);
    parallel_for(n_tri, [&](int k) {
        float lo = o->h_range_host[2 * k], hi = o->h_range_host[2 * k + 1];
        triangle_box(tris[k], lo, hi, &o->boxes_host[6 * k]);
        o->host[k].h_lo = lo;
        o->host[k].h_hi = hi;
        for (int q = 0; q < 3; ++q) {
            o->host[k].aabb_lo[q] = o->boxes_host[6 * k + q];
            o->host[k].aabb_hi[q] = o->boxes_host[6 * k + 3 + q];
        }
    });
    e = hipMemcpy(o->boxes, o->boxes_host.data(), sizeof(float) * 6 * n_tri, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(DMAP_ERR_HIP, "object bounds: %s", hipGetErrorString(e));
    const int P = o->bvh_leaves;
    const bool on_device = o->bvh_builder == DMAP_BVH_DEVICE ||
                           (o->bvh_builder == DMAP_BVH_AUTO && n_tri >= DEVICE_BVH_MIN_TRIANGLES);
    if (on_device) {
        const size_t ws_bytes = bvh_build_workspace_bytes(n_tri, P);
        void *ws = nullptr;
        e = hipMalloc(&ws, ws_bytes);
        float tree_box[6];
        if (e == hipSuccess)
            e = bvh_build_device(o->boxes, n_tri, P, o->bvh, (int *)(o->bvh + P), tree_box, ws, ws_bytes, nullptr);
        (void)hipFree(ws);
        if (e != hipSuccess)
            return fail(DMAP_ERR_HIP, "object bounds: device BVH build: %s", hipGetErrorString(e));
        for (int a = 0; a < 3; ++a) {
            o->bound_lo[a] = tree_box[a];
            o->bound_hi[a] = tree_box[3 + a];
        }
    } else {
        std::vector<int> idx(n_tri);
        for (int k = 0; k < n_tri; ++k)
            idx[k] = k;
        std::vector<BvhBuildNode> build((size_t)P, BvhBuildNode{});
        std::vector<int> leaf((size_t)P, -1);
        BvhBuildNode &root = build[0];
        build_bvh_node(build, leaf, o->boxes_host, idx, 1, 0, n_tri, (uint32_t)P, root.lo[0], root.hi[0]);
        root.empty = 2;
        std::vector<BvhNode> nodes((size_t)P, BvhNode{});
        for (int i = 0; i < P; ++i) {
            const BvhBuildNode &b = build[(size_t)i];
            BvhNode &nd = nodes[(size_t)i];
            nd.axis = b.axis;
            nd.empty = b.empty;
            for (int c = 0; c < 2; ++c)
                if (!(b.empty & (1 << c)))
                    bvh_centre_radius(b.lo[c], b.hi[c], nd.ctr[c], nd.rad[c]);
        }
        for (int a = 0; a < 3; ++a) {
            o->bound_lo[a] = root.lo[0][a];
            o->bound_hi[a] = root.hi[0][a];
        }
        e = hipMemcpy(o->bvh, nodes.data(), sizeof(BvhNode) * P, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(o->bvh + P, leaf.data(), sizeof(int) * P, hipMemcpyHostToDevice);
        if (e != hipSuccess)
            return fail(DMAP_ERR_HIP, "object bounds: %s", hipGetErrorString(e));
    }
    o->const_normal = 1;
    for (int k = 0; k < n_tri; ++k)
        o->const_normal &= o->host[k].const_normal;
    o->view.pyr = view_of(tile);
    o->view.tris = o->tris;
    o->view.boxes = o->boxes;
    o->view.n_tri = n_tri;
    o->view.const_normal = o->const_normal;
    o->view.bvh = o->bvh;
    o->view.bvh_leaf = (const int *)(o->bvh + P);
    o->view.bvh_leaves = P;
    e = hipMemcpy(o->tris, o->host.data(), sizeof(TriRecord) * n_tri, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(DMAP_ERR_HIP, "object bounds: %s", hipGetErrorString(e));
    return DMAP_OK;
}

int dmap_object_create(dmap_ctx *ctx, dmap_tile *tile, int n_tri, const double *positions, const double *normals,
                       const double *uvs, dmap_object **out)
{
    return dmap_object_create_ex(ctx, tile, n_tri, positions, normals, uvs, DMAP_BVH_AUTO, out);
}

int dmap_object_create_ex(dmap_ctx *ctx, dmap_tile *tile, int n_tri, const double *positions, const double *normals,
                          const double *uvs, int bvh_builder, dmap_object **out)
{
    if (!ctx || !tile || !positions || !normals || !uvs || !out || n_tri <= 0)
        return fail(DMAP_ERR_INVALID, "dmap_object_create: bad argument");
    if (bvh_builder != DMAP_BVH_AUTO && bvh_builder != DMAP_BVH_HOST && bvh_builder != DMAP_BVH_DEVICE)
        return fail(DMAP_ERR_INVALID, "dmap_object_create: unknown BVH builder %d", bvh_builder);
    if (n_tri > (1 << 30))
        return fail(DMAP_ERR_INVALID, "dmap_object_create: at most 2^30 triangles per object");
    if (tile->ctx->device != ctx->device)
        return fail(DMAP_ERR_INVALID, "dmap_object_create: tile lives on another device");
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<HostTri> tris(n_tri);
    dmap_object *o = new dmap_object;
    o->ctx = ctx;
    o->tile = tile;
    o->n_tri = n_tri;
    o->bvh_builder = bvh_builder;
    o->host.resize(n_tri);
    // Per triangle: its chart and record; the first failure (lowest k) is
    // reported.
    std::vector<int> bad(n_tri, 0), root_levels(n_tri, -1);
    parallel_for(n_tri, [&](int k) {
        if (!make_tri(positions + 9 * (size_t)k, normals + 9 * (size_t)k, uvs + 6 * (size_t)k, tris[k]))
            bad[k] = 1;
        else if (!fill_record(tris[k], tile->n, o->host[k], &root_levels[k]))
            bad[k] = 2;
    });
    for (int k = 0; k < n_tri; ++k) {
        if (bad[k] == 1) {
            delete o;
            return fail(DMAP_ERR_INVALID, "dmap_object_create: triangle %d is degenerate in the texture plane", k);
        }
        if (bad[k] == 2) {
            const int root_level = root_levels[k];
            delete o;
            if (root_level > MAX_ROOT_LEVEL)
                return fail(DMAP_ERR_INVALID,
                            "dmap_object_create: triangle %d needs root level %d (its uv span over a %d-texel tile); "
                            "the walk supports root levels up to %d",
                            k, root_level, tile->n, MAX_ROOT_LEVEL);
            return fail(DMAP_ERR_INVALID, "dmap_object_create: triangle %d has a degenerate frame or footprint", k);
        }
    }
    hipError_t e;
    if ((e = hipMalloc(&o->tris, sizeof(TriRecord) * n_tri)) != hipSuccess ||
        (e = hipMalloc(&o->boxes, sizeof(float) * 6 * n_tri)) != hipSuccess) {
        object_free(o);
        return fail(DMAP_ERR_MEMORY, "dmap_object_create: %s", hipGetErrorString(e));
    }
    o->positions.assign(positions, positions + 9 * (size_t)n_tri);
    o->normals.assign(normals, normals + 9 * (size_t)n_tri);
    o->uvs.assign(uvs, uvs + 6 * (size_t)n_tri);
    o->bvh_leaves = bvh_leaves_for(n_tri);
    e = hipMalloc(&o->bvh, (sizeof(BvhNode) + sizeof(int)) * (size_t)o->bvh_leaves);
    if (e != hipSuccess) {
        object_free(o);
        return fail(DMAP_ERR_MEMORY, "dmap_object_create: %s", hipGetErrorString(e));
    }
    int rc = object_bounds(o, tris);
    if (rc != DMAP_OK) {
        object_free(o);
        return rc;
    }
    *out = o;
    return DMAP_OK;
}

int dmap_object_refresh(dmap_object *obj)
{
    if (!obj)
        return fail(DMAP_ERR_INVALID, "dmap_object_refresh: NULL object");
    HIP_TRY(hipSetDevice(obj->ctx->device));
    HIP_TRY(hipDeviceSynchronize());
    std::vector<HostTri> tris(obj->n_tri);
    parallel_for(obj->n_tri, [&](int k) {
        make_tri(obj->positions.data() + 9 * (size_t)k, obj->normals.data() + 9 * (size_t)k,
                 obj->uvs.data() + 6 * (size_t)k, tris[k]);
    });
    return object_bounds(obj, tris);
}

int dmap_object_triangle_info(const dmap_object *obj, int k, int *root_level, int *n_roots, int32_t *roots,
                              float *aabb, float *h_range)
{
    if (!obj || k < 0 || k >= obj->n_tri)
        return fail(DMAP_ERR_INVALID, "dmap_object_triangle_info: bad triangle %d", k);
    const TriRecord &R = obj->host[k];
    if (root_level)
        *root_level = R.root_level;
    if (n_roots)
        *n_roots = R.n_roots;
    if (roots)
        for (int r = 0; r < 4; ++r) {
            roots[2 * r] = r < R.n_roots ? R.root_i[r] : 0;
            roots[2 * r + 1] = r < R.n_roots ? R.root_j[r] : 0;
        }
    if (aabb)
        std::memcpy(aabb, &obj->boxes_host[6 * k], 6 * sizeof(float));
    if (h_range) {
        h_range[0] = obj->h_range_host[2 * k];
        h_range[1] = obj->h_range_host[2 * k + 1];
    }
    return DMAP_OK;
}

void dmap_object_destroy(dmap_object *obj)
{
    if (obj)
        (void)hipSetDevice(obj->ctx->device);
    object_free(obj);
}

static int check_options(const dmap_options *opt, dmap_options &o)
{
    o = dmap_options{2, 1, 0, 0, DMAP_ORDER_NEAR_FIRST, 1, 0};
    if (opt)
        o = *opt;
    if (o.bound < 0 || o.bound > 2 || o.leaf < 0 || o.leaf > 1 || o.precision < 0 || o.precision > 1 ||
        o.walk_start < 0 || o.walk_start > 1 || o.reorder < 0 || o.reorder > 2)
        return fail(DMAP_ERR_INVALID,
                    "bad traversal options (bound %d, leaf %d, precision %d, reorder %d, walk_start %d)", o.bound,
                    o.leaf, o.precision, o.reorder, o.walk_start);
    return DMAP_OK;
}

// Kernel timing: when on, each trace records a pair of events around the
// traversal kernel; dmap_kernel_timing_read sums their durations.
static std::mutex g_timing_mu;
static bool g_timing = false;
static std::vector<std::pair<hipEvent_t, hipEvent_t>> g_timing_events;

// fixed_ws: NULL to allocate the call's workspace on the stream; else the
// device address of a caller-provided workspace of fixed_bytes (the
// no-allocation entry point's scratch), used instead, with no allocation,
// free or timing event.
// Batches under this many rays are traced in their given order (a reorder
// would cost more than it saves).
constexpr int64_t REORDER_MIN_RAYS = 4096;

static int trace(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits, uint8_t *occluded,
                 int prim_mask, const dmap_options *options, const dmap_ray_stats *stats, void *stream,
                 void *fixed_ws = nullptr, size_t fixed_bytes = 0)
{
    if (!obj || !rays)
        return fail(DMAP_ERR_INVALID, "intersect: NULL object or rays");
    if (n < 0)
        return fail(DMAP_ERR_INVALID, "intersect: negative ray count");
    if (n == 0)
        return DMAP_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->tmin || !rays->tmax)
        return fail(DMAP_ERR_INVALID, "intersect: a ray array is NULL");
    if (prim_mask < -1 || prim_mask >= obj->n_tri)
        return fail(DMAP_ERR_INVALID, "intersect: prim_mask %d of %d triangles", prim_mask, obj->n_tri);
    dmap_options o;
    int rc = check_options(options, o);
    if (rc != DMAP_OK)
        return rc;
    HIP_TRY(hipSetDevice(obj->ctx->device));
    TraceParams p;
    std::memset(&p, 0, sizeof p);
    p.obj = obj->view;
    p.prim_mask = prim_mask;
    p.bound = o.bound;
    p.slab_max_level = o.slab_max_level;
    p.leaf = o.leaf;
    p.polish = o.polish ? 1 : 0;
    p.walk_start = o.walk_start;
    p.any_hit = occluded ? 1 : 0;
    p.n_rays = n;
    p.rays = RaysDev{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    if (hits)
        p.hits = HitsDev{hits->t,  hits->s,  hits->tt,   hits->bu,      hits->bv,     hits->nx,
                         hits->ny, hits->nz, hits->prim, hits->texel_i, hits->texel_j};
    p.occluded = occluded;
    if (stats) {
        if ((stats->records == nullptr) != (stats->record_count == nullptr) || stats->record_capacity < 0)
            return fail(DMAP_ERR_INVALID, "intersect: node-test records need records, record_count and a "
                                          "capacity >= 0 together");
        p.stats = StatsDev{stats->triangle_tests,       stats->node_tests,          stats->leaf_tests,
                           stats->sub_squares,          stats->newton_runs,         stats->newton_iters,
                           stats->node_tests_after_hit, stats->leaf_tests_after_hit, stats->node_tests_per_level,
                           reinterpret_cast<NodeRecordDev *>(stats->records), stats->record_capacity,
                           reinterpret_cast<unsigned long long *>(stats->record_count)};
    }
    hipStream_t s = (hipStream_t)stream;
    if (n > 0x7fffffffLL)
        return fail(DMAP_ERR_INVALID, "intersect: at most 2^31 - 1 rays per call");
    // Per-call workspace, stream-ordered: the ray queue head, and the
    // buffers of the ray order when one is asked for (batches under 4,096
    // rays are traced as given).
    const int order = n >= REORDER_MIN_RAYS ? o.reorder : DMAP_ORDER_GIVEN;
    size_t rbytes = order == DMAP_ORDER_SORT         ? (reorder_workspace_bytes(n) + 255) / 256 * 256
                    : order == DMAP_ORDER_NEAR_FIRST ? (near_first_workspace_bytes(n) + 255) / 256 * 256
                                                     : 0;
    // A closest-hit call's raw records, after the queue head and the reorder
    // buffers.
    size_t raw_bytes = occluded ? 0 : sizeof(RawHit) * (size_t)n;
    void *ws = fixed_ws;
    if (fixed_ws) {
        if (256 + rbytes + raw_bytes > fixed_bytes)
            return fail(DMAP_ERR_INVALID, "intersect: %lld rays need %zu B of scratch, more than its %zu B",
                        (long long)n, 256 + rbytes + raw_bytes, fixed_bytes);
    } else {
        HIP_TRY(hipMallocFromPoolAsync(&ws, 256 + rbytes + raw_bytes, obj->ctx->pool, s));
    }
    p.queue = (unsigned *)ws;
    if (raw_bytes)
        p.raw = (RawHit *)((char *)ws + 256 + rbytes);
    if (order != DMAP_ORDER_GIVEN) {
        uint32_t *perm = nullptr;
        hipError_t e = order == DMAP_ORDER_SORT
                           ? reorder_rays(p.rays, n, obj->bound_lo, obj->bound_hi, (char *)ws + 256, rbytes, &perm, s)
                           : near_first_rays(p.rays, n, obj->bound_lo, obj->bound_hi, (char *)ws + 256, rbytes, &perm,
                                             s);
        if (e != hipSuccess) {
            if (!fixed_ws)
                (void)hipFreeAsync(ws, s);
            return fail(DMAP_ERR_HIP, "intersect: ray reorder: %s", hipGetErrorString(e));
        }
        p.perm = perm;
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (!fixed_ws) {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        if (g_timing && hipEventCreate(&ev[0]) == hipSuccess && hipEventCreate(&ev[1]) == hipSuccess)
            g_timing_events.emplace_back(ev[0], ev[1]);
    }
    hipError_t e = launch_trace(p, o.precision, s, ev[1] ? ev[0] : nullptr, ev[1]);
    if (!fixed_ws)
        (void)hipFreeAsync(ws, s);
    HIP_TRY(e);
    return DMAP_OK;
}

int dmap_kernel_timing(int enable)
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing = enable != 0;
    return DMAP_OK;
}

int dmap_kernel_timing_read(double *trace_ms, int64_t *launches)
{
    if (!trace_ms || !launches)
        return fail(DMAP_ERR_INVALID, "dmap_kernel_timing_read: NULL argument");
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    {
        std::lock_guard<std::mutex> lk(g_timing_mu);
        evs.swap(g_timing_events);
    }
    double sum = 0.0;
    int rc = DMAP_OK;
    for (auto &pr : evs) {
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(pr.second);
        if (e == hipSuccess)
            e = hipEventElapsedTime(&ms, pr.first, pr.second);
        if (e != hipSuccess && rc == DMAP_OK)
            rc = fail(DMAP_ERR_HIP, "dmap_kernel_timing_read: %s", hipGetErrorString(e));
        sum += ms;
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    *trace_ms = sum;
    *launches = (int64_t)evs.size();
    return rc;
}

int dmap_ray_order(const dmap_object *obj, const dmap_rays *rays, int64_t n, int reorder, uint32_t *perm,
                   void *stream)
{
    if (!obj || !rays || !perm)
        return fail(DMAP_ERR_INVALID, "dmap_ray_order: NULL object, rays or perm");
    if (n < 0 || n > 0x7fffffffLL)
        return fail(DMAP_ERR_INVALID, "dmap_ray_order: %lld rays, 0 .. 2^31 - 1 per call", (long long)n);
    if (reorder != DMAP_ORDER_GIVEN && reorder != DMAP_ORDER_SORT && reorder != DMAP_ORDER_NEAR_FIRST)
        return fail(DMAP_ERR_INVALID, "dmap_ray_order: order %d", reorder);
    if (n == 0)
        return DMAP_OK;
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz || !rays->tmin || !rays->tmax)
        return fail(DMAP_ERR_INVALID, "dmap_ray_order: a ray array is NULL");
    HIP_TRY(hipSetDevice(obj->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const int order = n >= REORDER_MIN_RAYS ? reorder : DMAP_ORDER_GIVEN;
    if (order == DMAP_ORDER_GIVEN) {
        HIP_TRY(identity_order(perm, n, s));
        return DMAP_OK;
    }
    const RaysDev R{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    const size_t bytes = order == DMAP_ORDER_SORT ? reorder_workspace_bytes(n) : near_first_workspace_bytes(n);
    void *ws = nullptr;
    HIP_TRY(hipMallocFromPoolAsync(&ws, bytes, obj->ctx->pool, s));
    uint32_t *out = nullptr;
    hipError_t e = order == DMAP_ORDER_SORT ? reorder_rays(R, n, obj->bound_lo, obj->bound_hi, ws, bytes, &out, s)
                                            : near_first_rays(R, n, obj->bound_lo, obj->bound_hi, ws, bytes, &out, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(perm, out, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    const hipError_t ef = hipFreeAsync(ws, s);
    if (e == hipSuccess)
        e = ef;
    if (e != hipSuccess)
        return fail(DMAP_ERR_HIP, "dmap_ray_order: %s", hipGetErrorString(e));
    return DMAP_OK;
}

int dmap_intersect_closest(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                           int prim_mask, const dmap_options *options, const dmap_ray_stats *stats, void *stream)
{
    if (n == 0 && obj && rays)
        return DMAP_OK;
    if (!hits || !hits->t || !hits->s || !hits->tt || !hits->prim || !hits->texel_i || !hits->texel_j)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_closest: hits t, s, tt, prim, texel_i and texel_j are required");
    return trace(obj, rays, n, hits, nullptr, prim_mask, options, stats, stream);
}

int dmap_intersect_closest_noalloc(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                                   int prim_mask, const dmap_options *options, const dmap_ray_stats *stats,
                                   void *stream)
{
    if (n == 0 && obj && rays)
        return DMAP_OK;
    if (!hits || !hits->t || !hits->s || !hits->tt || !hits->prim || !hits->texel_i || !hits->texel_j)
        return fail(DMAP_ERR_INVALID,
                    "dmap_intersect_closest_noalloc: hits t, s, tt, prim, texel_i and texel_j are required");
    if (n > NOALLOC_MAX_RAYS)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_closest_noalloc: %lld rays, at most %lld per call",
                    (long long)n, (long long)NOALLOC_MAX_RAYS);
    if (!obj || !obj->ctx->scratch)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_closest_noalloc: NULL object or context without scratch");
    return trace(obj, rays, n, hits, nullptr, prim_mask, options, stats, stream, obj->ctx->scratch,
                 NOALLOC_SCRATCH_BYTES);
}

int dmap_intersect_adjoint(const dmap_object *obj, const dmap_rays *rays, int64_t n, const float *dl_dt,
                           const float *dl_dnx, const float *dl_dny, const float *dl_dnz, double *grad_texels,
                           const dmap_options *options, void *stream)
{
    if (!obj || !rays || !grad_texels)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_adjoint: NULL object, rays or gradient");
    bool any_n = dl_dnx || dl_dny || dl_dnz, all_n = dl_dnx && dl_dny && dl_dnz;
    if (any_n != all_n)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_adjoint: give all three normal gradients or none");
    dmap_options o;
    int rc = check_options(options, o);
    if (rc != DMAP_OK)
        return rc;
    if (o.leaf != 0)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_adjoint: the adjoint follows the Newton leaf only");
    if (n == 0)
        return DMAP_OK;
    if (n < 0 || n > 0x7fffffffLL)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_adjoint: bad ray count %lld", (long long)n);
    HIP_TRY(hipSetDevice(obj->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    // The closest hits are recomputed, not kept from the forward pass.
    void *ws = nullptr;
    HIP_TRY(hipMallocFromPoolAsync(&ws, 44 * (size_t)n, obj->ctx->pool, s));
    float *f = (float *)ws;
    int32_t *ip = (int32_t *)(f + 8 * n);
    dmap_hits h{f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n, f + 7 * n, ip, ip + n, ip + 2 * n};
    rc = trace(obj, rays, n, &h, nullptr, -1, &o, nullptr, stream);
    if (rc == DMAP_OK) {
        AdjointParams ap;
        std::memset(&ap, 0, sizeof ap);
        ap.obj = obj->view;
        ap.rays = RaysDev{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
        ap.hits = HitsDev{h.t, h.s, h.tt, h.bu, h.bv, h.nx, h.ny, h.nz, h.prim, h.texel_i, h.texel_j};
        ap.dl_dt = dl_dt;
        ap.dl_dn[0] = dl_dnx;
        ap.dl_dn[1] = dl_dny;
        ap.dl_dn[2] = dl_dnz;
        ap.grad = grad_texels;
        ap.n_rays = n;
        hipError_t e = launch_adjoint(ap, s);
        if (e != hipSuccess)
            rc = fail(DMAP_ERR_HIP, "dmap_intersect_adjoint: %s", hipGetErrorString(e));
    }
    (void)hipFreeAsync(ws, s);
    return rc;
}

int dmap_hit_record(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                    const dmap_uv_map *uv_map, int twosided, float *record, void *stream)
{
    return dmap_hit_record_world(obj, rays, n, hits, uv_map, twosided, nullptr, nullptr, record, stream);
}

int dmap_hit_record_world(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                          const dmap_uv_map *uv_map, int twosided, const double *object_to_world,
                          const dmap_ray_diffs *diffs, float *record, void *stream)
{
    if (!obj || !rays || !hits || !record)
        return fail(DMAP_ERR_INVALID, "dmap_hit_record: NULL argument");
    Instance inst{};
    if (object_to_world) {
        // m^-T from the cofactors: (m^-1)^T = cof(m) / det(m).
        const double *A = object_to_world;
        const double m[9] = {A[0], A[1], A[2], A[4], A[5], A[6], A[8], A[9], A[10]};
        const double c[9] = {m[4] * m[8] - m[5] * m[7], m[5] * m[6] - m[3] * m[8], m[3] * m[7] - m[4] * m[6],
                             m[2] * m[7] - m[1] * m[8], m[0] * m[8] - m[2] * m[6], m[1] * m[6] - m[0] * m[7],
                             m[1] * m[5] - m[2] * m[4], m[2] * m[3] - m[0] * m[5], m[0] * m[4] - m[1] * m[3]};
        const double det = m[0] * c[0] + m[1] * c[1] + m[2] * c[2];
        if (!(std::isfinite(det) && det != 0.0))
            return fail(DMAP_ERR_INVALID, "dmap_hit_record_world: the object-to-world matrix is singular");
        for (int q = 0; q < 9; ++q) {
            inst.m[q] = m[q];
            inst.nt[q] = c[q] / det;
        }
        inst.t[0] = A[3];
        inst.t[1] = A[7];
        inst.t[2] = A[11];
        inst.on = 1;
    }
    RayDiffs rd{};
    if (diffs) {
        const float *pl[12] = {diffs->rx_ox, diffs->rx_oy, diffs->rx_oz, diffs->rx_dx, diffs->rx_dy, diffs->rx_dz,
                               diffs->ry_ox, diffs->ry_oy, diffs->ry_oz, diffs->ry_dx, diffs->ry_dy, diffs->ry_dz};
        for (int q = 0; q < 12; ++q) {
            if (!pl[q])
                return fail(DMAP_ERR_INVALID, "dmap_hit_record_world: a ray differential array is NULL");
            rd.p[q] = pl[q];
        }
        rd.on = 1;
    }
    if (!hits->prim || !hits->bu || !hits->bv || (twosided && (!rays->dx || !rays->dy || !rays->dz)))
        return fail(DMAP_ERR_INVALID, "dmap_hit_record: hits prim, bu, bv (and ray directions when twosided) needed");
    if (n < 0)
        return fail(DMAP_ERR_INVALID, "dmap_hit_record: negative count");
    UvMap uv{1.0, 1.0, 0.0, 0.0, 0};
    if (uv_map) {
        uv = UvMap{uv_map->scale_u, uv_map->scale_v, uv_map->offset_u, uv_map->offset_v, uv_map->flip_v ? 1 : 0};
        if (!(uv.scale_u != 0.0 && uv.scale_v != 0.0 && std::isfinite(uv.scale_u) && std::isfinite(uv.scale_v)))
            return fail(DMAP_ERR_INVALID, "dmap_hit_record: uv scales must be finite and non-zero");
    }
    if (n == 0)
        return DMAP_OK;
    HIP_TRY(hipSetDevice(obj->ctx->device));
    RaysDev R{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    HitsDev H{hits->t,  hits->s,  hits->tt,   hits->bu,      hits->bv,     hits->nx,
              hits->ny, hits->nz, hits->prim, hits->texel_i, hits->texel_j};
    HIP_TRY(launch_hit_record(obj->view, R, H, n, uv, twosided, inst, rd, record, (hipStream_t)stream));
    return DMAP_OK;
}

int dmap_intersect_any(const dmap_object *obj, const dmap_rays *rays, int64_t n, uint8_t *occluded, int prim_mask,
                       const dmap_options *options, void *stream)
{
    if (n == 0 && obj && rays)
        return DMAP_OK;
    if (!occluded)
        return fail(DMAP_ERR_INVALID, "dmap_intersect_any: occluded is NULL");
    return trace(obj, rays, n, nullptr, occluded, prim_mask, options, nullptr, stream);
}

} // extern "C"
