// CPU oracle for the ray / displaced-surface intersection path.
//
// TEST INFRASTRUCTURE ONLY: a double-precision restatement of the
// reference's CPU intersector, used as the parity checker by tests/, by
// __graft_entry__.smoke(), and as the timed host-core baseline by bench.py.
// The product library never links or calls this file.
//
// Each function names the reference file:line (relative to the reference
// repository root, under code/cpp/src/) whose behaviour it restates. The
// restatement keeps the reference's operation order where the result is a
// floating-point value another check compares against, so that the golden
// pyramid vectors match to rounding.

#include "dmap_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace orc
{

// ---------------------------------------------------------------- vectors

struct V2
{
    double x = 0.0, y = 0.0;
};
struct V3
{
    double x = 0.0, y = 0.0, z = 0.0;
    double operator[](int a) const { return a == 0 ? x : a == 1 ? y : z; }
    double &operator[](int a) { return a == 0 ? x : a == 1 ? y : z; }
};

static inline V2 v2(double x, double y) { return V2{x, y}; }
static inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
static inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
static inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
static inline V2 operator*(double s, V2 a) { return {s * a.x, s * a.y}; }
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
static inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
// Eigen's normalized(): divide by the norm when it is positive.
static inline V3 normalized(V3 a)
{
    double z = dot(a, a);
    return z > 0.0 ? a / std::sqrt(z) : a;
}

struct M3
{
    double m[3][3];
    V3 row(int r) const { return {m[r][0], m[r][1], m[r][2]}; }
    V3 mul(V3 v) const { return {dot(row(0), v), dot(row(1), v), dot(row(2), v)}; }
};

// 3x3 inverse by the adjugate (the cofactor form Eigen uses for fixed 3x3).
static M3 inverse(const M3 &a)
{
    const double(*m)[3] = a.m;
    double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    double inv = 1.0 / det;
    M3 r;
    r.m[0][0] = c00 * inv;
    r.m[1][0] = c01 * inv;
    r.m[2][0] = c02 * inv;
    r.m[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    r.m[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    r.m[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    r.m[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    r.m[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    r.m[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    return r;
}

// ------------------------------------------------------------ height grid

// HeightGrid (displaced_surface.h:40-56, displaced_surface.cpp:12-47).
struct Grid
{
    int W = 0, H = 0;
    double scale = 1.0;
    const double *values = nullptr;
    double offset = 0.0;
    bool repeat = false;

    double value(int j, int i) const { return values[(size_t)j * W + i]; }

    void cell(double x, double y, int &i, int &j, double &s, double &t) const
    {
        if (repeat) {
            x -= std::floor(x);
            y -= std::floor(y);
        } else {
            x = std::clamp(x, 0.0, 1.0);
            y = std::clamp(y, 0.0, 1.0);
        }
        double fx = x * (W - 1), fy = y * (H - 1);
        i = std::min((int)fx, W - 2);
        j = std::min((int)fy, H - 2);
        s = fx - i;
        t = fy - j;
    }
    double h(double x, double y) const
    {
        int i, j;
        double s, t;
        cell(x, y, i, j, s, t);
        return scale * ((1.0 - s) * (1.0 - t) * value(j, i) + s * (1.0 - t) * value(j, i + 1) +
                        (1.0 - s) * t * value(j + 1, i) + s * t * value(j + 1, i + 1)) +
               offset;
    }
    V2 grad(double x, double y) const
    {
        int i, j;
        double s, t;
        cell(x, y, i, j, s, t);
        double ds = (1.0 - t) * (value(j, i + 1) - value(j, i)) + t * (value(j + 1, i + 1) - value(j + 1, i));
        double dt = (1.0 - s) * (value(j + 1, i) - value(j, i)) + s * (value(j + 1, i + 1) - value(j, i + 1));
        return v2(scale * (ds * (W - 1)), scale * (dt * (H - 1)));
    }
};

// --------------------------------------------------------- base triangle

// BaseTriangle with a general chart (displaced_surface.cpp:54-91).
struct Tri
{
    V3 q0, e1, e2, m0, Mu, Mv, p0, p1, p2;
    V2 t0, t1, t2;
    double J[2][2];

    V3 P(double s, double t) const { return q0 + s * e1 + t * e2; }
    V3 M(double s, double t) const { return m0 + s * Mu + t * Mv; }
    V2 bary(double s, double t) const
    {
        V2 d = v2(s, t) - t0;
        return v2(J[0][0] * d.x + J[0][1] * d.y, J[1][0] * d.x + J[1][1] * d.y);
    }
    V2 param(double u, double v) const { return t0 + u * (t1 - t0) + v * (t2 - t0); }
};

static bool make_tri(const double *in, Tri &tri)
{
    V3 p0 = v3(in[0], in[1], in[2]), p1 = v3(in[3], in[4], in[5]), p2 = v3(in[6], in[7], in[8]);
    V3 n0 = v3(in[9], in[10], in[11]), n1 = v3(in[12], in[13], in[14]), n2 = v3(in[15], in[16], in[17]);
    V2 t0 = v2(in[18], in[19]), t1 = v2(in[20], in[21]), t2 = v2(in[22], in[23]);
    tri.p0 = p0;
    tri.p1 = p1;
    tri.p2 = p2;
    tri.t0 = t0;
    tri.t1 = t1;
    tri.t2 = t2;
    V3 eb1 = p1 - p0, eb2 = p2 - p0, Mb1 = n1 - n0, Mb2 = n2 - n0;
    // T = [t1 - t0 | t2 - t0], J = T^-1 (displaced_surface.cpp:62-68).
    double a = t1.x - t0.x, b = t2.x - t0.x, c = t1.y - t0.y, d = t2.y - t0.y;
    double det = a * d - b * c;
    if (!(std::abs(det) > 1e-30))
        return false;
    tri.J[0][0] = d / det;
    tri.J[0][1] = -b / det;
    tri.J[1][0] = -c / det;
    tri.J[1][1] = a / det;
    // E J and D J, and c = -J t0 (displaced_surface.cpp:73-85).
    auto col = [&](V3 x, V3 y, int k) { return tri.J[0][k] * x + tri.J[1][k] * y; };
    tri.e1 = col(eb1, eb2, 0);
    tri.e2 = col(eb1, eb2, 1);
    tri.Mu = col(Mb1, Mb2, 0);
    tri.Mv = col(Mb1, Mb2, 1);
    V2 cc = v2(-(tri.J[0][0] * t0.x + tri.J[0][1] * t0.y), -(tri.J[1][0] * t0.x + tri.J[1][1] * t0.y));
    tri.q0 = p0 + (cc.x * eb1 + cc.y * eb2);
    tri.m0 = n0 + (cc.x * Mb1 + cc.y * Mb2);
    return true;
}

// normal_frame_at (displaced_surface.cpp:99-108).
struct NFrame
{
    V3 N, Nu, Nv;
};
static NFrame normal_frame_at(const Tri &tri, double s, double t)
{
    V3 M = tri.M(s, t);
    double len = norm(M);
    NFrame f;
    f.N = M / len;
    f.Nu = (tri.Mu - dot(f.N, tri.Mu) * f.N) / len;
    f.Nv = (tri.Mv - dot(f.N, tri.Mv) * f.N) / len;
    return f;
}

// displaced_position (displaced_surface.cpp:162-166).
static V3 displaced_position(const Tri &tri, const Grid &g, double s, double t)
{
    NFrame f = normal_frame_at(tri, s, t);
    return tri.P(s, t) + g.h(s, t) * f.N;
}

// mean_edge (displaced_surface.cpp:168-171).
static double mean_edge(const Tri &tri)
{
    return (norm(tri.p1 - tri.p0) + norm(tri.p2 - tri.p0) + norm(tri.p2 - tri.p1)) / 3.0;
}

// ---------------------------------------------------------------- pyramid

struct Node
{
    double h0, gu, gv, r, ru, rv, h_min, h_max;
};

struct Level
{
    int m = 0;
    std::vector<double> ch[ORC_CHANNELS];
    explicit Level(int m_) : m(m_)
    {
        for (auto &c : ch)
            c.assign((size_t)m_ * m_, 0.0);
    }
};

// leaf_level (taylor_pyramid.cpp:45-67): the exact node of each texel.
static Level leaf_level(const std::vector<double> &V, int nodes, int n)
{
    Level out(n);
    double w = 1.0 / n, s = 0.5 * w;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double v00 = V[(size_t)j * nodes + i], v10 = V[(size_t)j * nodes + i + 1];
            double v01 = V[(size_t)(j + 1) * nodes + i], v11 = V[(size_t)(j + 1) * nodes + i + 1];
            size_t k = (size_t)j * n + i;
            out.ch[ORC_H0][k] = 0.25 * (v00 + v10 + v01 + v11);
            out.ch[ORC_GU][k] = 0.5 * ((v10 - v00) + (v11 - v01)) / w;
            out.ch[ORC_GV][k] = 0.5 * ((v01 - v00) + (v11 - v10)) / w;
            double c3 = (v00 - v10 - v01 + v11) / (w * w);
            out.ch[ORC_R][k] = std::abs(c3) * s * s;
            out.ch[ORC_RU][k] = std::abs(c3) * s;
            out.ch[ORC_RV][k] = std::abs(c3) * s;
            out.ch[ORC_HMIN][k] = std::min(std::min(v00, v10), std::min(v01, v11));
            out.ch[ORC_HMAX][k] = std::max(std::max(v00, v10), std::max(v01, v11));
        }
    return out;
}

// fold (taylor_pyramid.cpp:77-138): the conservative 4-to-1 fold.
static Level fold(const Level &c)
{
    int m = c.m / 2;
    Level out(m);
    double s = 0.5 / m;
    const double off[2] = {-0.5 * s, 0.5 * s};
    for (int J = 0; J < m; ++J)
        for (int I = 0; I < m; ++I) {
            size_t kp = (size_t)J * m + I;
            double gu_hi = -INFINITY, gu_lo = INFINITY, gv_hi = -INFINITY, gv_lo = INFINITY;
            double h_min = INFINITY, h_max = -INFINITY;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    size_t kc = (size_t)(2 * J + a) * c.m + (2 * I + b);
                    gu_hi = std::max(gu_hi, c.ch[ORC_GU][kc] + c.ch[ORC_RU][kc]);
                    gu_lo = std::min(gu_lo, c.ch[ORC_GU][kc] - c.ch[ORC_RU][kc]);
                    gv_hi = std::max(gv_hi, c.ch[ORC_GV][kc] + c.ch[ORC_RV][kc]);
                    gv_lo = std::min(gv_lo, c.ch[ORC_GV][kc] - c.ch[ORC_RV][kc]);
                    h_min = std::min(h_min, c.ch[ORC_HMIN][kc]);
                    h_max = std::max(h_max, c.ch[ORC_HMAX][kc]);
                }
            double gu_p = 0.5 * (gu_hi + gu_lo), gv_p = 0.5 * (gv_hi + gv_lo);
            out.ch[ORC_GU][kp] = gu_p;
            out.ch[ORC_RU][kp] = 0.5 * (gu_hi - gu_lo);
            out.ch[ORC_GV][kp] = gv_p;
            out.ch[ORC_RV][kp] = 0.5 * (gv_hi - gv_lo);
            out.ch[ORC_HMIN][kp] = h_min;
            out.ch[ORC_HMAX][kp] = h_max;
            double upper = -INFINITY, lower = INFINITY;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    size_t kc = (size_t)(2 * J + a) * c.m + (2 * I + b);
                    double dx = off[b], dy = off[a];
                    double du_g = c.ch[ORC_GU][kc] - gu_p, dv_g = c.ch[ORC_GV][kc] - gv_p;
                    double base = c.ch[ORC_H0][kc] - c.ch[ORC_GU][kc] * dx - c.ch[ORC_GV][kc] * dy;
                    for (int cu = 0; cu < 2; ++cu)
                        for (int cv = 0; cv < 2; ++cv) {
                            double corner = base + du_g * (dx + off[cu]) + dv_g * (dy + off[cv]);
                            upper = std::max(upper, corner + c.ch[ORC_R][kc]);
                            lower = std::min(lower, corner - c.ch[ORC_R][kc]);
                        }
                }
            out.ch[ORC_H0][kp] = 0.5 * (upper + lower);
            out.ch[ORC_R][kp] = 0.5 * (upper - lower);
        }
    return out;
}

// direct_level (taylor_pyramid.cpp:151-201) with texel_coeffs (:20-39).
static Level direct_level(const std::vector<double> &V, int nodes, int n, int level, const std::vector<double> &c1,
                          const std::vector<double> &c2, const std::vector<double> &c3)
{
    int k = 1 << level, m = n >> level;
    Level out(m);
    double s_texel = 0.5 / n;
    for (int J = 0; J < m; ++J)
        for (int I = 0; I < m; ++I) {
            size_t kp = (size_t)J * m + I;
            double gu_hi = -INFINITY, gu_lo = INFINITY, gv_hi = -INFINITY, gv_lo = INFINITY;
            for (int b = 0; b < k; ++b)
                for (int a = 0; a < k; ++a) {
                    size_t kt = (size_t)(J * k + b) * n + (I * k + a);
                    double spread = std::abs(c3[kt]) * s_texel;
                    gu_hi = std::max(gu_hi, c1[kt] + spread);
                    gu_lo = std::min(gu_lo, c1[kt] - spread);
                    gv_hi = std::max(gv_hi, c2[kt] + spread);
                    gv_lo = std::min(gv_lo, c2[kt] - spread);
                }
            double gu = 0.5 * (gu_hi + gu_lo), gv = 0.5 * (gv_hi + gv_lo);
            out.ch[ORC_GU][kp] = gu;
            out.ch[ORC_RU][kp] = 0.5 * (gu_hi - gu_lo);
            out.ch[ORC_GV][kp] = gv;
            out.ch[ORC_RV][kp] = 0.5 * (gv_hi - gv_lo);
            double upper = -INFINITY, lower = INFINITY, h_min = INFINITY, h_max = -INFINITY;
            for (int b = 0; b <= k; ++b) {
                double dv = (b - 0.5 * k) / n;
                for (int a = 0; a <= k; ++a) {
                    double du = (a - 0.5 * k) / n;
                    double value = V[(size_t)(J * k + b) * nodes + (I * k + a)];
                    double residual = value - gu * du - gv * dv;
                    upper = std::max(upper, residual);
                    lower = std::min(lower, residual);
                    h_min = std::min(h_min, value);
                    h_max = std::max(h_max, value);
                }
            }
            out.ch[ORC_H0][kp] = 0.5 * (upper + lower);
            out.ch[ORC_R][kp] = 0.5 * (upper - lower);
            out.ch[ORC_HMIN][kp] = h_min;
            out.ch[ORC_HMAX][kp] = h_max;
        }
    return out;
}

struct Pyramid
{
    int n_leaf = 0, n_levels = 0;
    std::vector<Level> levels;

    double cell_width(int level) const { return std::ldexp(1.0, level) / n_leaf; }

    // TaylorPyramid ctor (taylor_pyramid.cpp:215-251).
    bool build(const double *values, int nodes, double scale, int mode, double offset)
    {
        int n = nodes - 1;
        if (!(n >= 1 && (n & (n - 1)) == 0))
            return false;
        n_leaf = n;
        n_levels = 1;
        while ((1 << n_levels) <= n)
            ++n_levels;
        std::vector<double> V((size_t)nodes * nodes);
        for (size_t k = 0; k < V.size(); ++k)
            V[k] = values[k] * scale;
        levels.clear();
        if (mode == 0) {
            levels.push_back(leaf_level(V, nodes, n));
            while (levels.back().m > 1)
                levels.push_back(fold(levels.back()));
        } else {
            size_t cnt = (size_t)n * n;
            std::vector<double> c1(cnt), c2(cnt), c3(cnt);
            double w = 1.0 / n;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double v00 = V[(size_t)j * nodes + i], v10 = V[(size_t)j * nodes + i + 1];
                    double v01 = V[(size_t)(j + 1) * nodes + i], v11 = V[(size_t)(j + 1) * nodes + i + 1];
                    size_t k = (size_t)j * n + i;
                    c1[k] = 0.5 * ((v10 - v00) + (v11 - v01)) / w;
                    c2[k] = 0.5 * ((v01 - v00) + (v11 - v10)) / w;
                    c3[k] = (v00 - v10 - v01 + v11) / (w * w);
                }
            for (int level = 0; level < n_levels; ++level)
                levels.push_back(direct_level(V, nodes, n, level, c1, c2, c3));
        }
        if (offset != 0.0)
            for (Level &l : levels)
                for (size_t k = 0; k < l.ch[0].size(); ++k) {
                    l.ch[ORC_H0][k] += offset;
                    l.ch[ORC_HMIN][k] += offset;
                    l.ch[ORC_HMAX][k] += offset;
                }
        return true;
    }

    // node (taylor_pyramid.cpp:259-274): wrapped indices, min-max above the root.
    Node node(int level, int64_t i, int64_t j) const
    {
        if (level >= n_levels) {
            const Level &root = levels.back();
            double lo = root.ch[ORC_HMIN][0], hi = root.ch[ORC_HMAX][0];
            return {0.5 * (lo + hi), 0.0, 0.0, 0.5 * (hi - lo), root.ch[ORC_RU][0], root.ch[ORC_RV][0], lo, hi};
        }
        const Level &l = levels[level];
        auto wrap = [&](int64_t a) {
            int64_t r = a % l.m;
            return r < 0 ? r + l.m : r;
        };
        size_t k = (size_t)wrap(j) * l.m + wrap(i);
        return {l.ch[0][k], l.ch[1][k], l.ch[2][k], l.ch[3][k], l.ch[4][k], l.ch[5][k], l.ch[6][k], l.ch[7][k]};
    }
};

// ------------------------------------------------------- affine arithmetic

struct Interval
{
    double lo = 0.0, hi = 0.0;
    double magnitude() const { return std::max(std::abs(lo), std::abs(hi)); }
    bool contains(double x) const { return lo <= x && x <= hi; }
};
static Interval ipt(double x) { return {x, x}; }
// affine_form.cpp:10-16.
static Interval operator+(Interval a, Interval b) { return {a.lo + b.lo, a.hi + b.hi}; }
static Interval operator*(double s, Interval a) { return s >= 0.0 ? Interval{s * a.lo, s * a.hi} : Interval{s * a.hi, s * a.lo}; }

// AffineForm (affine_form.h:40-61, affine_form.cpp:18-62).
struct AF
{
    double c = 0.0, a = 0.0, b = 0.0, k = 0.0;
    Interval interval() const
    {
        double e = std::abs(a) + std::abs(b) + k;
        return {c - e, c + e};
    }
};
static AF af(double c, double a = 0.0, double b = 0.0, double k = 0.0) { return AF{c, a, b, k}; }
static AF operator+(AF x, const AF &r) { return {x.c + r.c, x.a + r.a, x.b + r.b, x.k + r.k}; }
static AF operator*(AF x, const AF &r)
{
    double u = std::abs(x.a) + std::abs(x.b) + x.k;
    double v = std::abs(r.a) + std::abs(r.b) + r.k;
    double c0 = x.c;
    AF o;
    o.c = c0 * r.c;
    o.a = c0 * r.a + r.c * x.a;
    o.b = c0 * r.b + r.c * x.b;
    o.k = std::abs(r.c) * x.k + std::abs(c0) * r.k + u * v;
    return o;
}
static AF operator*(double s, AF x) { return {x.c * s, x.a * s, x.b * s, x.k * std::abs(s)}; }

// rsqrt, Min-Range (affine_form.cpp:64-76). The caller guarantees lo > 0.
static AF rsqrt(const AF &f)
{
    Interval r = f.interval();
    double lo = r.lo, hi = r.hi;
    auto g = [](double x) { return 1.0 / std::sqrt(x); };
    double alpha = -0.5 * g(hi) * g(hi) * g(hi);
    double beta = 0.5 * (g(lo) + g(hi) - alpha * (lo + hi));
    double delta = 0.5 * std::abs(g(lo) - g(hi) - alpha * (lo - hi));
    return AF{alpha * f.c + beta, alpha * f.a, alpha * f.b, std::abs(alpha) * f.k + delta};
}

struct AV3
{
    AF x, y, z;
    const AF &operator[](int a) const { return a == 0 ? x : a == 1 ? y : z; }
};
static AV3 av3(V3 c, V3 a = V3{}, V3 b = V3{}, V3 k = V3{})
{
    return {AF{c.x, a.x, b.x, k.x}, AF{c.y, a.y, b.y, k.y}, AF{c.z, a.z, b.z, k.z}};
}
static AV3 operator+(const AV3 &p, const AV3 &q) { return {p.x + q.x, p.y + q.y, p.z + q.z}; }
static AV3 operator*(const AF &s, const AV3 &v) { return {s * v.x, s * v.y, s * v.z}; }
static AF dot(V3 a, const AV3 &b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static AV3 mul(const M3 &A, const AV3 &v) { return {dot(A.row(0), v), dot(A.row(1), v), dot(A.row(2), v)}; }

// sq_length (affine_form.cpp:78-85).
static AF sq_length(const AV3 &v)
{
    V3 xc = v3(v.x.c, v.y.c, v.z.c), xu = v3(v.x.a, v.y.a, v.z.a), xv = v3(v.x.b, v.y.b, v.z.b),
       xk = v3(v.x.k, v.y.k, v.z.k);
    V3 dev = v3(std::abs(xu.x) + std::abs(xv.x) + xk.x, std::abs(xu.y) + std::abs(xv.y) + xk.y,
                std::abs(xu.z) + std::abs(xv.z) + xk.z);
    double offset = 0.5 * dot(dev, dev);
    return AF{dot(xc, xc) + offset, 2.0 * dot(xc, xu), 2.0 * dot(xc, xv), 2.0 * std::abs(dot(xc, xk)) + offset};
}

// normalize (affine_form.cpp:87-95).
static void normalize(AV3 &v)
{
    AF l = sq_length(v);
    if (l.interval().lo <= 0.0) {
        v.x = v.y = v.z = AF{0.0, 0.0, 0.0, 1.0};
        return;
    }
    v = rsqrt(l) * v;
}

// ------------------------------------------------------------- uv domain

// UvTriangle (uv_clip.cpp:10-25).
struct UvTri
{
    V2 t[3], n[3];
    double d[3];
    V2 lo, hi;
    void init(V2 t0, V2 t1, V2 t2)
    {
        t[0] = t0;
        t[1] = t1;
        t[2] = t2;
        double signed_area = 0.5 * ((t1 - t0).x * (t2 - t0).y - (t1 - t0).y * (t2 - t0).x);
        double orient = signed_area >= 0.0 ? 1.0 : -1.0;
        for (int k = 0; k < 3; ++k) {
            V2 e = t[(k + 1) % 3] - t[k];
            n[k] = orient * v2(-e.y, e.x);
            d[k] = dot2(n[k], t[k]);
        }
        lo = v2(std::min(std::min(t0.x, t1.x), t2.x), std::min(std::min(t0.y, t1.y), t2.y));
        hi = v2(std::max(std::max(t0.x, t1.x), t2.x), std::max(std::max(t0.y, t1.y), t2.y));
    }
    bool inside(V2 p) const
    {
        for (int k = 0; k < 3; ++k)
            if (dot2(n[k], p) < d[k])
                return false;
        return true;
    }
};

enum Overlap
{
    OUTSIDE = 0,
    INSIDE = 1,
    STRADDLE = 2
};

// classify_square (uv_clip.cpp:27-44).
static Overlap classify_square(const UvTri &tri, V2 c, double s)
{
    if (c.x + s <= tri.lo.x || c.x - s >= tri.hi.x || c.y + s <= tri.lo.y || c.y - s >= tri.hi.y)
        return OUTSIDE;
    bool all_inside = true;
    for (int k = 0; k < 3; ++k) {
        double proj = dot2(tri.n[k], c);
        double ext = (std::abs(tri.n[k].x) + std::abs(tri.n[k].y)) * s;
        if (proj + ext <= tri.d[k])
            return OUTSIDE;
        if (proj - ext < tri.d[k])
            all_inside = false;
    }
    return all_inside ? INSIDE : STRADDLE;
}

// footprint_roots (descent_sampler.cpp:66-83) with its helpers (:15-27).
static void footprint_roots(const UvTri &dom, int n_leaf, int &level, int64_t &i0, int64_t &j0)
{
    level = 0;
    while (true) {
        double w = std::ldexp(1.0, level) / n_leaf;
        int64_t ilo = (int64_t)std::floor(dom.lo.x / w), ihi = (int64_t)std::ceil(dom.hi.x / w) - 1;
        int64_t jlo = (int64_t)std::floor(dom.lo.y / w), jhi = (int64_t)std::ceil(dom.hi.y / w) - 1;
        if (ihi - ilo <= 1 && jhi - jlo <= 1) {
            i0 = ilo;
            j0 = jlo;
            return;
        }
        ++level;
    }
}

// ------------------------------------------------------------ node bounds

// TangentFrame (node_bounds.cpp:13-23).
struct Frame
{
    M3 to_tangent;
    V3 origin, ng;
    V3 point(V3 p) const { return to_tangent.mul(p - origin); }
    V3 direction(V3 d) const { return to_tangent.mul(d); }
};
static Frame tangent_frame(const Tri &tri)
{
    Frame f;
    f.origin = tri.q0;
    f.ng = normalized(cross(tri.e1, tri.e2));
    M3 from;
    for (int r = 0; r < 3; ++r) {
        from.m[r][0] = tri.e1[r];
        from.m[r][1] = tri.e2[r];
        from.m[r][2] = f.ng[r];
    }
    f.to_tangent = inverse(from);
    return f;
}

// interval_tolerance (node_bounds.cpp:128).
static double interval_tolerance(double t_scale) { return 1e-9 * std::max(1.0, t_scale); }

struct Box
{
    V3 lo, hi;
    // Aabb3::clip (node_bounds.cpp:25-44).
    bool clip(V3 o, V3 d, double &t0, double &t1) const
    {
        for (int a = 0; a < 3; ++a) {
            if (d[a] != 0.0) {
                double ta = (lo[a] - o[a]) / d[a], tb = (hi[a] - o[a]) / d[a];
                if (ta > tb)
                    std::swap(ta, tb);
                t0 = std::max(t0, ta);
                t1 = std::min(t1, tb);
            } else if (o[a] < lo[a] || o[a] > hi[a]) {
                return false;
            }
        }
        if (t0 > t1 + interval_tolerance(std::max(std::abs(t0), std::abs(t1))))
            return false;
        if (t0 > t1)
            std::swap(t0, t1);
        return true;
    }
    void expand(const Box &o)
    {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], o.lo[a]);
            hi[a] = std::max(hi[a], o.hi[a]);
        }
    }
};

static Box hull(const AV3 &S)
{
    Box b;
    for (int a = 0; a < 3; ++a) {
        Interval r = S[a].interval();
        b.lo[a] = r.lo;
        b.hi[a] = r.hi;
    }
    return b;
}

struct NodeBound
{
    int level = 0;
    int64_t i = 0, j = 0;
    Node node;
    V2 centre, rect_lo, rect_hi;
    AV3 m;
    Box box;
};

// node_bound (node_bounds.cpp:83-109) with normal_form (:65-68).
static NodeBound node_bound(const Tri &tri, const Frame &frame, const Pyramid &pyr, int level, int64_t i, int64_t j)
{
    NodeBound nb;
    nb.level = level;
    nb.i = i;
    nb.j = j;
    nb.node = pyr.node(level, i, j);
    double w = pyr.cell_width(level);
    nb.centre = v2((i + 0.5) * w, (j + 0.5) * w);
    V2 tlo = v2(std::min(std::min(tri.t0.x, tri.t1.x), tri.t2.x), std::min(std::min(tri.t0.y, tri.t1.y), tri.t2.y));
    V2 thi = v2(std::max(std::max(tri.t0.x, tri.t1.x), tri.t2.x), std::max(std::max(tri.t0.y, tri.t1.y), tri.t2.y));
    nb.rect_lo = v2(std::max(i * w, tlo.x), std::max(j * w, tlo.y));
    nb.rect_hi = v2(std::min((i + 1) * w, thi.x), std::min((j + 1) * w, thi.y));
    V2 c = 0.5 * (nb.rect_lo + nb.rect_hi);
    V2 half = 0.5 * (nb.rect_hi - nb.rect_lo);
    half.x = std::max(half.x, 0.0);
    half.y = std::max(half.y, 0.0);
    AV3 N = av3(tri.m0 + c.x * tri.Mu + c.y * tri.Mv, half.x * tri.Mu, half.y * tri.Mv);
    normalize(N);
    nb.m = mul(frame.to_tangent, N);
    AF h{0.5 * (nb.node.h_min + nb.node.h_max), 0.0, 0.0, 0.5 * (nb.node.h_max - nb.node.h_min)};
    AV3 p = av3(v3(c.x, c.y, 0.0), v3(half.x, 0.0, 0.0), v3(0.0, half.y, 0.0));
    nb.box = hull(p + h * nb.m);
    return nb;
}

struct Slab
{
    Interval A, B, C, D;
    double R = 0.0;
    Interval eval(V3 p) const { return ((p.z * A + p.x * B) + p.y * C) + D; }
};

// slab_form (node_bounds.cpp:111-122).
static Slab slab_form(const NodeBound &nb)
{
    const Node &n = nb.node;
    Interval mx = nb.m.x.interval(), my = nb.m.y.interval(), mz = nb.m.z.interval();
    Slab f;
    f.A = (ipt(1.0) + n.gu * mx) + n.gv * my;
    f.B = (-n.gu) * mz;
    f.C = (-n.gv) * mz;
    f.D = (-(n.h0 - n.gu * nb.centre.x - n.gv * nb.centre.y)) * mz;
    f.R = n.r * mz.magnitude();
    return f;
}

struct TRay
{
    V3 o, d;
    double t_min = 0.0, t_max = INFINITY;
};

// box_interval (node_bounds.cpp:130-135).
static bool box_interval(const NodeBound &nb, const TRay &ray, double &t0, double &t1)
{
    t0 = ray.t_min;
    t1 = ray.t_max;
    return nb.box.clip(ray.o, ray.d, t0, t1);
}

// slab_interval (node_bounds.cpp:137-179).
static bool slab_interval(const NodeBound &nb, const TRay &ray, double &t0, double &t1)
{
    Box column = nb.box;
    column.lo.z = -INFINITY;
    column.hi.z = INFINITY;
    t0 = ray.t_min;
    t1 = ray.t_max;
    if (!column.clip(ray.o, ray.d, t0, t1))
        return false;
    V3 anchor = ray.o + t0 * ray.d;
    double span = t1 - t0;
    Slab f = slab_form(nb);
    Interval F0 = f.eval(anchor);
    Interval F1 = (ray.d.z * f.A + ray.d.x * f.B) + ray.d.y * f.C;
    double s0 = 0.0, s1 = span;
    if (F1.lo > 0.0)
        s1 = std::min(s1, (f.R - F0.lo) / F1.lo);
    else if (F1.lo < 0.0)
        s0 = std::max(s0, (f.R - F0.lo) / F1.lo);
    else if (F0.lo > f.R)
        return false;
    if (F1.hi > 0.0)
        s0 = std::max(s0, (-f.R - F0.hi) / F1.hi);
    else if (F1.hi < 0.0)
        s1 = std::min(s1, (-f.R - F0.hi) / F1.hi);
    else if (F0.hi < -f.R)
        return false;
    if (s0 > s1 + interval_tolerance(std::abs(t0) + span))
        return false;
    if (s0 > s1)
        std::swap(s0, s1);
    t1 = t0 + s1;
    t0 = t0 + s0;
    return true;
}

// both_interval (node_bounds.cpp:181-197).
static bool both_interval(const NodeBound &nb, const TRay &ray, double &t0, double &t1)
{
    double b0, b1, s0, s1;
    if (!box_interval(nb, ray, b0, b1) || !slab_interval(nb, ray, s0, s1))
        return false;
    t0 = std::max(b0, s0);
    t1 = std::min(b1, s1);
    double tol = interval_tolerance(std::max(std::abs(t0), std::abs(t1)));
    if (t0 > t1 + tol)
        return false;
    if (t0 > t1)
        std::swap(t0, t1);
    return true;
}

struct Options
{
    int bound = 2; // traversal_options.h:32, Both
    int slab_max_level = 1;
    int leaf = 0; // Newton
};

// node_interval (node_bounds.cpp:199-211).
static bool node_interval(const NodeBound &nb, const TRay &ray, const Options &opt, double &t0, double &t1)
{
    bool read_slab = nb.level <= opt.slab_max_level;
    switch (opt.bound) {
    case 0:
        return box_interval(nb, ray, t0, t1);
    case 1:
        return read_slab ? slab_interval(nb, ray, t0, t1) : box_interval(nb, ray, t0, t1);
    default:
        return read_slab ? both_interval(nb, ray, t0, t1) : box_interval(nb, ray, t0, t1);
    }
}

// height_range_walk (node_bounds.cpp:218-234).
static void height_range_walk(const UvTri &dom, const Pyramid &pyr, int level, int64_t i, int64_t j, double &h_min,
                              double &h_max)
{
    double w = pyr.cell_width(level);
    V2 c = v2((i + 0.5) * w, (j + 0.5) * w);
    Overlap ov = classify_square(dom, c, 0.5 * w);
    if (ov == OUTSIDE)
        return;
    if (ov == INSIDE || level == 0) {
        Node n = pyr.node(level, i, j);
        h_min = std::min(h_min, n.h_min);
        h_max = std::max(h_max, n.h_max);
        return;
    }
    for (int ch = 0; ch < 4; ++ch)
        height_range_walk(dom, pyr, level - 1, 2 * i + (ch & 1), 2 * j + (ch >> 1), h_min, h_max);
}

// triangle_aabb (node_bounds.cpp:251-269).
static Box triangle_aabb(const Tri &tri, double h_min, double h_max)
{
    const V2 t[3] = {tri.t0, tri.t1, tri.t2};
    AF h{0.5 * (h_min + h_max), 0.0, 0.0, 0.5 * (h_max - h_min)};
    Box box{v3(INFINITY, INFINITY, INFINITY), v3(-INFINITY, -INFINITY, -INFINITY)};
    for (int k = 0; k < 3; ++k) {
        V2 c = (0.5 * t[k] + 0.25 * t[(k + 1) % 3]) + 0.25 * t[(k + 2) % 3];
        V2 du = 0.25 * (t[(k + 1) % 3] - t[k]), dv = 0.25 * (t[(k + 2) % 3] - t[k]);
        AV3 P = av3(tri.P(c.x, c.y), du.x * tri.e1 + du.y * tri.e2, dv.x * tri.e1 + dv.y * tri.e2);
        AV3 N = av3(tri.M(c.x, c.y), du.x * tri.Mu + du.y * tri.Mv, dv.x * tri.Mu + dv.y * tri.Mv);
        normalize(N);
        box.expand(hull(P + h * N));
    }
    return box;
}

// ---------------------------------------------------------- intersector

struct Cell
{
    int level = 0;
    int64_t i = 0, j = 0;
    bool operator!=(const Cell &r) const { return level != r.level || i != r.i || j != r.j; }
};

// floor_div2 / floor_mod2 and the walk (displaced_intersector.cpp:18-19, 85-121).
static int64_t floor_div2(int64_t x) { return x >= 0 ? x / 2 : -((-x + 1) / 2); }
static int floor_mod2(int64_t x) { return (int)(x - 2 * floor_div2(x)); }
static void cell_up(Cell &c)
{
    ++c.level;
    c.i = floor_div2(c.i);
    c.j = floor_div2(c.j);
}
static void cell_down(Cell &c, bool si, bool sj)
{
    --c.level;
    c.i = 2 * c.i + (si ? 1 : 0);
    c.j = 2 * c.j + (sj ? 1 : 0);
}
static void cell_next(Cell &c, bool si, bool sj, int root_level)
{
    while (true) {
        int code = 2 * floor_mod2(c.i + (si ? 1 : 0)) + floor_mod2(c.j + (sj ? 1 : 0));
        if (code == 1) {
            c.j += sj ? 1 : -1;
            c.i += si ? -1 : 1;
            return;
        }
        if (code == 3) {
            cell_up(c);
            if (c.level > root_level)
                return;
            continue;
        }
        c.j += sj ? -1 : 1;
        return;
    }
}

// plane_basis (displaced_intersector.cpp:22-27).
static void plane_basis(V3 n, V3 &d1, V3 &d2)
{
    V3 a = std::abs(n.x) < 0.9 ? v3(1.0, 0.0, 0.0) : v3(0.0, 1.0, 0.0);
    d1 = normalized(cross(a, n));
    d2 = cross(n, d1);
}

// Möller–Trumbore (displaced_intersector.cpp:30-49).
static bool ray_triangle(V3 o, V3 d, double t_min, double t_max, V3 A, V3 B, V3 C, double &t, double &b1, double &b2)
{
    V3 e1 = B - A, e2 = C - A;
    V3 p = cross(d, e2);
    double det = dot(e1, p);
    if (det == 0.0)
        return false;
    double inv = 1.0 / det;
    V3 s = o - A;
    b1 = dot(s, p) * inv;
    if (b1 < 0.0 || b1 > 1.0)
        return false;
    V3 q = cross(s, e1);
    b2 = dot(d, q) * inv;
    if (b2 < 0.0 || b1 + b2 > 1.0)
        return false;
    t = dot(e2, q) * inv;
    return t > t_min && t < t_max;
}

struct Hit
{
    double t = INFINITY;
    V2 st, bary;
    V3 ng;
    int64_t i = 0, j = 0;
};

struct Stats
{
    int64_t v[ORC_STATS] = {};
    // When set, every node test of the walk is appended
    // (TraversalStats::records, displaced_intersector.h:101-102), for ray
    // `ray` on triangle `prim`.
    std::vector<orc_node_record> *records = nullptr;
    int32_t ray = 0, prim = 0;
};

// record_node_test (displaced_intersector.cpp:159-172): the box's and the
// slab's decisions with the walk's current t_max, then the three interval
// lengths with t_max = inf (-1 when that test fails).
static void record_node_test(const NodeBound &nb, TRay ray, int level, int64_t i, int64_t j, Stats &st)
{
    orc_node_record rec{};
    rec.ray = st.ray;
    rec.prim = st.prim;
    rec.level = level;
    rec.seq = (int32_t)(st.v[ORC_ST_NODE] - 1);
    rec.i = i;
    rec.j = j;
    double t0, t1;
    rec.box_pass = box_interval(nb, ray, t0, t1);
    rec.slab_pass = slab_interval(nb, ray, t0, t1);
    ray.t_max = INFINITY;
    rec.box_length = box_interval(nb, ray, t0, t1) ? t1 - t0 : -1.0;
    rec.slab_length = slab_interval(nb, ray, t0, t1) ? t1 - t0 : -1.0;
    rec.both_length = both_interval(nb, ray, t0, t1) ? t1 - t0 : -1.0;
    st.records->push_back(rec);
}

// restricted (displaced_intersector.cpp:72-81).
static AF restricted(const AF &f, double a, double b, double rho)
{
    double alpha = 2.0 * a - 1.0, beta = 2.0 * b - 1.0, r = 2.0 * rho;
    return AF{f.c + f.a * alpha + f.b * beta, f.a * r, f.b * r, f.k};
}
static AV3 restricted(const AV3 &v, double a, double b, double rho)
{
    return {restricted(v.x, a, b, rho), restricted(v.y, a, b, rho), restricted(v.z, a, b, rho)};
}

struct LeafBound
{
    V2 lo;
    double w = 0.0, h00 = 0.0, h10 = 0.0, h01 = 0.0, h11 = 0.0;
    AV3 N_hat;
    V3 o, d, dn, d1, d2;
    double d_len = 0.0;
};

struct Intersector
{
    const Tri *tri = nullptr;
    const Pyramid *pyr = nullptr;
    Grid field;
    Frame frame;
    UvTri domain;
    int root_level = 0;
    Cell roots[4];
    int n_roots = 0;
    double residual_tolerance = 0.0;
    int max_iterations = 20;   // displaced_intersector.h:129
    int subdivision_depth = 5; // displaced_intersector.h:132

    // DisplacedIntersector ctor (displaced_intersector.cpp:135-151).
    void init(const Tri &t, const Pyramid &p, const Grid &g)
    {
        tri = &t;
        pyr = &p;
        field = g;
        frame = tangent_frame(t);
        domain.init(t.t0, t.t1, t.t2);
        int64_t i0, j0;
        footprint_roots(domain, p.n_leaf, root_level, i0, j0);
        double w = p.cell_width(root_level);
        int64_t i_lo = (int64_t)std::floor(domain.lo.x / w), i_hi = (int64_t)std::ceil(domain.hi.x / w) - 1;
        int64_t j_lo = (int64_t)std::floor(domain.lo.y / w), j_hi = (int64_t)std::ceil(domain.hi.y / w) - 1;
        n_roots = 0;
        for (int64_t j = j_lo; j <= j_hi; ++j)
            for (int64_t i = i_lo; i <= i_hi; ++i)
                roots[n_roots++] = Cell{root_level, i, j};
        residual_tolerance = 1e-10 * mean_edge(t);
    }

    // newton_from (displaced_intersector.cpp:325-388).
    bool newton_from(V3 o, V3 d, double t_min, double t_max, int64_t i, int64_t j, V2 guess, Hit &hit,
                     int64_t *iters) const
    {
        double w = pyr->cell_width(0);
        V2 lo = v2(i * w, j * w), hi = v2((i + 1) * w, (j + 1) * w);
        double d_len = norm(d);
        V3 dn = d / d_len, d1, d2;
        plane_basis(dn, d1, d2);
        double prev_err2 = INFINITY;
        int err_streak = 0, behind_streak = 0, invalid_streak = 0;
        for (int iter = 0; iter < max_iterations; ++iter) {
            if (iters)
                ++*iters;
            double s = guess.x, t = guess.y;
            NFrame f = normal_frame_at(*tri, s, t);
            double h = field.h(s, t);
            V2 gh = field.grad(s, t);
            V3 S = tri->P(s, t) + h * f.N;
            V3 delta = S - o;
            V2 F = v2(dot(delta, d1), dot(delta, d2));
            double err2 = F.x * F.x + F.y * F.y;
            double along = dot(delta, dn);
            err_streak = err2 > prev_err2 ? err_streak + 1 : 0;
            behind_streak = along < 0.0 ? behind_streak + 1 : 0;
            if (err_streak >= 2 || behind_streak >= 2)
                return false;
            prev_err2 = err2;
            V3 Su = (tri->e1 + gh.x * f.N) + h * f.Nu;
            V3 Sv = (tri->e2 + gh.y * f.N) + h * f.Nv;
            if (err2 < residual_tolerance * residual_tolerance) {
                if (along < 0.0 || !domain.inside(guess))
                    return false;
                double t_hit = along / d_len;
                if (t_hit <= t_min || t_hit >= t_max)
                    return false;
                hit.t = t_hit;
                hit.st = guess;
                hit.bary = tri->bary(s, t);
                hit.ng = normalized(cross(Su, Sv));
                hit.i = i;
                hit.j = j;
                return true;
            }
            double J00 = dot(d1, Su), J01 = dot(d1, Sv), J10 = dot(d2, Su), J11 = dot(d2, Sv);
            double det = J00 * J11 - J01 * J10;
            if (det == 0.0)
                return false;
            V2 step = v2((J11 * F.x - J01 * F.y) / det, (-J10 * F.x + J00 * F.y) / det);
            guess = guess - step;
            bool outside = guess.x < lo.x || guess.x > hi.x || guess.y < lo.y || guess.y > hi.y || !domain.inside(guess);
            if (outside) {
                if (++invalid_streak >= 3)
                    return false;
                guess = v2(std::min(std::max(guess.x, lo.x), hi.x), std::min(std::max(guess.y, lo.y), hi.y));
            } else {
                invalid_streak = 0;
            }
        }
        return false;
    }

    // sub_square_bound (displaced_intersector.cpp:266-290).
    bool sub_square_bound(const LeafBound &leaf, double a, double b, double rho, double t_min, double t_max,
                          double &t_lo) const
    {
        V2 c = leaf.lo + leaf.w * v2(a, b);
        if (classify_square(domain, c, leaf.w * rho) == OUTSIDE)
            return false;
        double T = ((leaf.h11 - leaf.h10) - leaf.h01) + leaf.h00;
        double ha = leaf.h10 - leaf.h00, hb = leaf.h01 - leaf.h00;
        AF h{((leaf.h00 + ha * a) + hb * b) + T * a * b, (ha + T * b) * rho, (hb + T * a) * rho,
             std::abs(T) * rho * rho};
        AV3 P = av3(tri->P(c.x, c.y), leaf.w * rho * tri->e1, leaf.w * rho * tri->e2);
        AV3 S = P + h * restricted(leaf.N_hat, a, b, rho);
        AF F1 = dot(leaf.d1, S), F2 = dot(leaf.d2, S);
        F1.c -= dot(leaf.d1, leaf.o);
        F2.c -= dot(leaf.d2, leaf.o);
        if (!F1.interval().contains(0.0) || !F2.interval().contains(0.0))
            return false;
        AF along = dot(leaf.dn, S);
        along.c -= dot(leaf.dn, leaf.o);
        Interval range = along.interval();
        t_lo = range.lo / leaf.d_len;
        return t_lo < t_max && range.hi / leaf.d_len > t_min;
    }

    // search_leaf (displaced_intersector.cpp:292-323).
    void search_leaf(const LeafBound &leaf, int64_t i, int64_t j, int depth, double a, double b, double rho,
                     double t_min, double &t_max, Hit &hit, bool &found, Stats *st) const
    {
        if (depth == subdivision_depth) {
            Hit h;
            if (st)
                ++st->v[ORC_ST_NEWTON_RUNS];
            if (newton_from(leaf.o, leaf.d, t_min, t_max, i, j, leaf.lo + leaf.w * v2(a, b), h,
                            st ? &st->v[ORC_ST_NEWTON_ITERS] : nullptr)) {
                hit = h;
                t_max = h.t;
                found = true;
            }
            return;
        }
        struct Child
        {
            double a, b, t_lo;
        } ch[4];
        int n = 0;
        double r = 0.5 * rho;
        for (int q = 0; q < 4; ++q) {
            double ca = a + ((q & 1) ? r : -r), cb = b + ((q & 2) ? r : -r), t_lo;
            if (st)
                ++st->v[ORC_ST_SUB];
            if (sub_square_bound(leaf, ca, cb, r, t_min, t_max, t_lo))
                ch[n++] = {ca, cb, t_lo};
        }
        std::stable_sort(ch, ch + n, [](const Child &x, const Child &y) { return x.t_lo < y.t_lo; });
        for (int q = 0; q < n; ++q) {
            if (ch[q].t_lo >= t_max)
                break;
            search_leaf(leaf, i, j, depth + 1, ch[q].a, ch[q].b, r, t_min, t_max, hit, found, st);
        }
    }

    // intersect_newton (displaced_intersector.cpp:239-264).
    bool intersect_newton(V3 o, V3 d, double t_min, double t_max, int64_t i, int64_t j, Hit &hit, Stats *st) const
    {
        LeafBound leaf;
        leaf.w = pyr->cell_width(0);
        leaf.lo = v2(i * leaf.w, j * leaf.w);
        leaf.h00 = field.h(leaf.lo.x, leaf.lo.y);
        leaf.h10 = field.h(leaf.lo.x + leaf.w, leaf.lo.y);
        leaf.h01 = field.h(leaf.lo.x, leaf.lo.y + leaf.w);
        leaf.h11 = field.h(leaf.lo.x + leaf.w, leaf.lo.y + leaf.w);
        V2 c = leaf.lo + v2(0.5 * leaf.w, 0.5 * leaf.w);
        leaf.N_hat = av3(tri->m0 + c.x * tri->Mu + c.y * tri->Mv, 0.5 * leaf.w * tri->Mu, 0.5 * leaf.w * tri->Mv);
        normalize(leaf.N_hat);
        leaf.o = o;
        leaf.d = d;
        leaf.d_len = norm(d);
        leaf.dn = d / leaf.d_len;
        plane_basis(leaf.dn, leaf.d1, leaf.d2);
        bool found = false;
        search_leaf(leaf, i, j, 0, 0.5, 0.5, 0.5, t_min, t_max, hit, found, st);
        return found;
    }

    // intersect_two_triangle (displaced_intersector.cpp:390-426).
    bool intersect_two_triangle(V3 o, V3 d, double t_min, double t_max, int64_t i, int64_t j, Hit &hit) const
    {
        double w = pyr->cell_width(0);
        V2 st00 = v2(i * w, j * w), st10 = v2((i + 1) * w, j * w), st01 = v2(i * w, (j + 1) * w),
           st11 = v2((i + 1) * w, (j + 1) * w);
        V3 S00 = displaced_position(*tri, field, st00.x, st00.y);
        V3 S10 = displaced_position(*tri, field, st10.x, st10.y);
        V3 S01 = displaced_position(*tri, field, st01.x, st01.y);
        V3 S11 = displaced_position(*tri, field, st11.x, st11.y);
        const V3 *A[2] = {&S00, &S10}, *B[2] = {&S10, &S11}, *C[2] = {&S01, &S01};
        const V2 *a[2] = {&st00, &st10}, *b[2] = {&st10, &st11}, *c[2] = {&st01, &st01};
        bool found = false;
        for (int k = 0; k < 2; ++k) {
            double t, b1, b2;
            if (!ray_triangle(o, d, t_min, t_max, *A[k], *B[k], *C[k], t, b1, b2))
                continue;
            V2 st = ((1.0 - b1 - b2) * *a[k] + b1 * *b[k]) + b2 * *c[k];
            if (!domain.inside(st))
                continue;
            hit.t = t;
            hit.st = st;
            hit.bary = tri->bary(st.x, st.y);
            hit.ng = normalized(cross(*B[k] - *A[k], *C[k] - *A[k]));
            hit.i = i;
            hit.j = j;
            t_max = t;
            found = true;
        }
        return found;
    }

    // intersect (displaced_intersector.cpp:175-237).
    bool intersect(V3 o, V3 d, double t_min, double t_max, const Options &opt, Hit &hit, Stats *st) const
    {
        TRay ray{frame.point(o), frame.direction(d), t_min, t_max};
        bool si = ray.d.x < 0.0, sj = ray.d.y < 0.0;
        bool found = false;
        int64_t nodes_after = 0, leaves_after = 0; // since the latest hit (displaced_intersector.cpp:181)
        for (int r = 0; r < n_roots; ++r) {
            Cell cur = roots[r], end = roots[r];
            cell_next(end, si, sj, root_level);
            while (cur != end) {
                double w = pyr->cell_width(cur.level);
                V2 c = v2((cur.i + 0.5) * w, (cur.j + 0.5) * w);
                if (classify_square(domain, c, 0.5 * w) == OUTSIDE) {
                    cell_next(cur, si, sj, root_level);
                    continue;
                }
                if (st) {
                    ++st->v[ORC_ST_NODE];
                    ++st->v[ORC_ST_LEVEL0 + std::min(cur.level, (int)ORC_MAX_LEVELS - 1)];
                    ++nodes_after;
                }
                NodeBound nb = node_bound(*tri, frame, *pyr, cur.level, cur.i, cur.j);
                ray.t_max = t_max;
                if (st && st->records)
                    record_node_test(nb, ray, cur.level, cur.i, cur.j, *st);
                double t0, t1;
                if (!node_interval(nb, ray, opt, t0, t1)) {
                    cell_next(cur, si, sj, root_level);
                    continue;
                }
                if (cur.level > 0) {
                    cell_down(cur, si, sj);
                    continue;
                }
                if (st) {
                    ++st->v[ORC_ST_LEAF];
                    ++leaves_after;
                }
                Hit lh;
                bool leaf = opt.leaf == 0 ? intersect_newton(o, d, t_min, t_max, cur.i, cur.j, lh, st)
                                          : intersect_two_triangle(o, d, t_min, t_max, cur.i, cur.j, lh);
                if (leaf && lh.t < t_max) {
                    hit = lh;
                    t_max = hit.t;
                    found = true;
                    nodes_after = 0;
                    leaves_after = 0;
                }
                cell_next(cur, si, sj, root_level);
            }
        }
        // displaced_intersector.cpp:232-235.
        if (st && found) {
            st->v[ORC_ST_NODE_AFTER_HIT] += nodes_after;
            st->v[ORC_ST_LEAF_AFTER_HIT] += leaves_after;
        }
        return found;
    }
};

struct Scene
{
    std::vector<double> nodes;
    Grid grid;
    Pyramid pyr;
    std::vector<Tri> tris;
    std::vector<Intersector> its;
    std::vector<Box> boxes;
    std::vector<double> h_lo, h_hi;

    // ObjectIntersector::intersect_flat (displaced_intersector.cpp:578-599).
    bool intersect(V3 o, V3 d, double t_min, double t_max, const Options &opt, Hit &hit, int &triangle, Stats *st,
                   bool any) const
    {
        bool found = false;
        triangle = -1;
        for (size_t k = 0; k < its.size(); ++k) {
            double t0 = t_min, t1 = t_max;
            if (!boxes[k].clip(o, d, t0, t1))
                continue;
            if (st) {
                ++st->v[ORC_ST_TRIANGLE];
                st->prim = (int32_t)k;
            }
            Hit h;
            if (its[k].intersect(o, d, t_min, t_max, opt, h, st)) {
                hit = h;
                triangle = (int)k;
                t_max = h.t;
                found = true;
                if (any)
                    return true;
            }
        }
        return found;
    }
};

// Dynamic chunks of 1024 rays over a std::thread pool: the host-side
// analogue of ks::parallel_for over rays (traversal_test_util.cpp:13-19).
template <typename F> static void parallel_rays(int64_t n, int nthreads, F &&fn)
{
    if (nthreads <= 0)
        nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    const int64_t chunk = 1024;
    if (nthreads == 1 || n <= chunk) {
        for (int64_t r = 0; r < n; ++r)
            fn(r);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&]() {
            while (true) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n)
                    break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t r = b; r < e; ++r)
                    fn(r);
            }
        });
    for (auto &th : pool)
        th.join();
}

static void write_hit(const Hit &h, double *out)
{
    out[0] = h.t;
    out[1] = h.st.x;
    out[2] = h.st.y;
    out[3] = h.bary.x;
    out[4] = h.bary.y;
    out[5] = h.ng.x;
    out[6] = h.ng.y;
    out[7] = h.ng.z;
    out[8] = (double)h.i;
    out[9] = (double)h.j;
}

} // namespace orc

using namespace orc;

extern "C" {

int orc_pyramid_levels(int nodes)
{
    int n = nodes - 1, L = 1;
    while ((1 << L) <= n)
        ++L;
    return L;
}

int64_t orc_pyramid_cells(int nodes)
{
    int64_t n = nodes - 1, cells = 0;
    for (int64_t m = n; m >= 1; m >>= 1)
        cells += m * m;
    return cells;
}

int orc_pyramid_build(const double *values, int nodes, double scale, double offset, int build, double *out)
{
    Pyramid p;
    if (!p.build(values, nodes, scale, build, offset))
        return -1;
    size_t pos = 0;
    for (const Level &l : p.levels)
        for (int c = 0; c < ORC_CHANNELS; ++c) {
            std::memcpy(out + pos, l.ch[c].data(), l.ch[c].size() * sizeof(double));
            pos += l.ch[c].size();
        }
    return p.n_levels;
}

void orc_close_tile(const double *texels, int n, int repeat, double *out)
{
    int W = n + 1;
    for (int j = 0; j < W; ++j) {
        int tj = repeat ? j % n : std::min(j, n - 1);
        for (int i = 0; i < W; ++i) {
            int ti = repeat ? i % n : std::min(i, n - 1);
            out[(size_t)j * W + i] = texels[(size_t)tj * n + ti];
        }
    }
}

void *orc_scene_create(const double *node_values, int nodes, double scale, double offset, int repeat, int build,
                       int n_tri, const double *tri_data)
{
    Scene *s = new Scene;
    s->nodes.assign(node_values, node_values + (size_t)nodes * nodes);
    s->grid = Grid{nodes, nodes, scale, s->nodes.data(), offset, repeat != 0};
    if (!s->pyr.build(s->nodes.data(), nodes, scale, build, offset)) {
        delete s;
        return nullptr;
    }
    s->tris.resize(n_tri);
    for (int k = 0; k < n_tri; ++k)
        if (!make_tri(tri_data + 24 * (size_t)k, s->tris[k])) {
            delete s;
            return nullptr;
        }
    s->its.resize(n_tri);
    s->boxes.resize(n_tri);
    s->h_lo.resize(n_tri);
    s->h_hi.resize(n_tri);
    for (int k = 0; k < n_tri; ++k) {
        Intersector &it = s->its[k];
        it.init(s->tris[k], s->pyr, s->grid);
        // triangle_height_range (node_bounds.cpp:238-249).
        int level;
        int64_t i0, j0;
        footprint_roots(it.domain, s->pyr.n_leaf, level, i0, j0);
        double lo = INFINITY, hi = -INFINITY;
        for (int ch = 0; ch < 4; ++ch)
            height_range_walk(it.domain, s->pyr, level, i0 + (ch & 1), j0 + (ch >> 1), lo, hi);
        s->h_lo[k] = lo;
        s->h_hi[k] = hi;
        s->boxes[k] = triangle_aabb(s->tris[k], lo, hi);
    }
    return s;
}

void orc_scene_destroy(void *scene) { delete (Scene *)scene; }

int orc_triangle_info(void *scene, int k, int *root_level, int *n_roots, int64_t *roots_ij, double *aabb,
                      double *h_range, double *residual_tolerance)
{
    Scene *s = (Scene *)scene;
    if (k < 0 || k >= (int)s->its.size())
        return -1;
    const Intersector &it = s->its[k];
    *root_level = it.root_level;
    *n_roots = it.n_roots;
    for (int r = 0; r < it.n_roots; ++r) {
        roots_ij[2 * r] = it.roots[r].i;
        roots_ij[2 * r + 1] = it.roots[r].j;
    }
    for (int a = 0; a < 3; ++a) {
        aabb[a] = s->boxes[k].lo[a];
        aabb[3 + a] = s->boxes[k].hi[a];
    }
    h_range[0] = s->h_lo[k];
    h_range[1] = s->h_hi[k];
    *residual_tolerance = it.residual_tolerance;
    return 0;
}

void orc_scene_intersect(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                         const double *tmax, int bound, int slab_max_level, int leaf, int nthreads, double *t,
                         double *st, double *bary, double *ng, int32_t *prim, int64_t *texel, int64_t *stats)
{
    const Scene *s = (const Scene *)scene;
    Options opt{bound, slab_max_level, leaf};
    parallel_rays(n, nthreads, [&](int64_t r) {
        Hit h;
        int tri = -1;
        Stats stt;
        bool hit = s->intersect(v3(o[3 * r], o[3 * r + 1], o[3 * r + 2]), v3(d[3 * r], d[3 * r + 1], d[3 * r + 2]),
                                tmin[r], tmax[r], opt, h, tri, stats ? &stt : nullptr, false);
        t[r] = hit ? h.t : INFINITY;
        if (st) {
            st[2 * r] = hit ? h.st.x : 0.0;
            st[2 * r + 1] = hit ? h.st.y : 0.0;
        }
        if (bary) {
            bary[2 * r] = hit ? h.bary.x : 0.0;
            bary[2 * r + 1] = hit ? h.bary.y : 0.0;
        }
        if (ng) {
            ng[3 * r] = hit ? h.ng.x : 0.0;
            ng[3 * r + 1] = hit ? h.ng.y : 0.0;
            ng[3 * r + 2] = hit ? h.ng.z : 0.0;
        }
        if (prim)
            prim[r] = hit ? tri : -1;
        if (texel) {
            texel[2 * r] = hit ? h.i : 0;
            texel[2 * r + 1] = hit ? h.j : 0;
        }
        if (stats)
            for (int q = 0; q < ORC_STATS; ++q)
                stats[ORC_STATS * r + q] = stt.v[q];
    });
}

void orc_scene_records(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                       const double *tmax, int bound, int slab_max_level, int leaf, int64_t capacity,
                       orc_node_record *out, int64_t *count)
{
    const Scene *s = (const Scene *)scene;
    Options opt{bound, slab_max_level, leaf};
    std::vector<orc_node_record> records;
    for (int64_t r = 0; r < n; ++r) {
        Hit h;
        int tri = -1;
        Stats stt;
        stt.records = &records;
        stt.ray = (int32_t)r;
        s->intersect(v3(o[3 * r], o[3 * r + 1], o[3 * r + 2]), v3(d[3 * r], d[3 * r + 1], d[3 * r + 2]), tmin[r],
                     tmax[r], opt, h, tri, &stt, false);
    }
    *count = (int64_t)records.size();
    for (int64_t q = 0; q < std::min(capacity, *count); ++q)
        out[q] = records[(size_t)q];
}

// ------------------------------------------------------------------ adjoint
//
// No counterpart in the reference. The closest hit is a root of
// S(s, t) = o + tau d with S = P + h N (displaced_position,
// displaced_surface.cpp:162-166) and h bilinear in the four nodes of the cell
// holding (s, t) (HeightGrid::h, displaced_surface.cpp:12-39). For a node
// value v the implicit-function theorem on the 3x3 system
// [S_s | S_t | -d] (ds, dt, dtau) = -dS/dv gives dtau/dv, and the normal
// ng = unit(S_s x S_t) (displaced_intersector.cpp:357-366) moves through
// dS_s = dS_s/dv + S_ss ds + S_st dt and dS_t alike. Everything is in tile
// coordinates (s, t) with the grid's own bilinear weights.

namespace
{
struct Geo
{
    V3 S, Ss, St, Sss, Sst, Stt, N, Nu, Nv;
    int ci = 0, cj = 0;
    double fs = 0.0, ft = 0.0;
};

// S and its first and second (s, t) derivatives. N = M / |M| with M affine
// in (s, t): N_u = (M_u - N (N . M_u)) / |M| (normal_frame_at,
// displaced_surface.cpp:99-108), and differentiating once more
// N_uu = -(2 N_u (N . M_u) + N (N_u . M_u)) / |M|, the mixed and vv terms alike.
Geo surface_geometry(const Tri &tri, const Grid &g, double s, double t)
{
    Geo q;
    g.cell(s, t, q.ci, q.cj, q.fs, q.ft);
    const double n1 = g.W - 1, fs = q.fs, ft = q.ft;
    const double v00 = g.value(q.cj, q.ci), v10 = g.value(q.cj, q.ci + 1), v01 = g.value(q.cj + 1, q.ci),
                 v11 = g.value(q.cj + 1, q.ci + 1);
    const double h = g.scale * ((1 - fs) * (1 - ft) * v00 + fs * (1 - ft) * v10 + (1 - fs) * ft * v01 + fs * ft * v11) +
                     g.offset;
    const double hs = g.scale * n1 * ((1 - ft) * (v10 - v00) + ft * (v11 - v01));
    const double ht = g.scale * n1 * ((1 - fs) * (v01 - v00) + fs * (v11 - v10));
    const double hst = g.scale * n1 * n1 * (v11 - v10 - v01 + v00);
    const V3 M = tri.M(s, t);
    const double len = norm(M);
    const V3 N = M / len;
    const double nu = dot(N, tri.Mu), nv = dot(N, tri.Mv);
    const V3 Nu = (tri.Mu - nu * N) / len, Nv = (tri.Mv - nv * N) / len;
    const V3 Nuu = -1.0 / len * (2.0 * nu * Nu + dot(Nu, tri.Mu) * N);
    const V3 Nuv = -1.0 / len * (nu * Nv + nv * Nu + dot(Nv, tri.Mu) * N);
    const V3 Nvv = -1.0 / len * (2.0 * nv * Nv + dot(Nv, tri.Mv) * N);
    q.N = N;
    q.Nu = Nu;
    q.Nv = Nv;
    q.S = tri.P(s, t) + h * N;
    q.Ss = tri.e1 + hs * N + h * Nu;
    q.St = tri.e2 + ht * N + h * Nv;
    q.Sss = 2.0 * hs * Nu + h * Nuu;
    q.Sst = hst * N + hs * Nv + ht * Nu + h * Nuv;
    q.Stt = 2.0 * ht * Nv + h * Nvv;
    return q;
}
} // namespace

void orc_scene_adjoint(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                       const double *tmax, const double *dl_dt, const double *dl_dn, int nthreads, double *grad_nodes,
                       double *t_out, double *ng_out)
{
    const Scene *sc = (const Scene *)scene;
    const Grid &g = sc->grid;
    Options opt{2, 1, 0};
    std::vector<int64_t> node(4 * (size_t)n, -1);
    std::vector<double> val(4 * (size_t)n, 0.0);
    parallel_rays(n, nthreads, [&](int64_t r) {
        const V3 O = v3(o[3 * r], o[3 * r + 1], o[3 * r + 2]), D = v3(d[3 * r], d[3 * r + 1], d[3 * r + 2]);
        Hit hit;
        int k = -1;
        if (t_out)
            t_out[r] = INFINITY;
        if (ng_out)
            ng_out[3 * r] = ng_out[3 * r + 1] = ng_out[3 * r + 2] = 0.0;
        if (!sc->intersect(O, D, tmin[r], tmax[r], opt, hit, k, nullptr, false))
            return;
        const Tri &tri = sc->tris[k];
        // Newton on F = ((S - o) . d1, (S - o) . d2) to the double root.
        const V3 dn = normalized(D);
        V3 d1, d2;
        plane_basis(dn, d1, d2);
        double s = hit.st.x, t = hit.st.y;
        for (int it = 0; it < 8; ++it) {
            Geo q = surface_geometry(tri, g, s, t);
            V3 rel = q.S - O;
            double F1 = dot(rel, d1), F2 = dot(rel, d2);
            double J00 = dot(d1, q.Ss), J01 = dot(d1, q.St), J10 = dot(d2, q.Ss), J11 = dot(d2, q.St);
            double det = J00 * J11 - J01 * J10;
            if (det == 0.0)
                break;
            double ds = (J11 * F1 - J01 * F2) / det, dt = (-J10 * F1 + J00 * F2) / det;
            s -= ds;
            t -= dt;
            if (std::abs(ds) + std::abs(dt) < 1e-17)
                break;
        }
        const Geo q = surface_geometry(tri, g, s, t);
        const double tau = dot(q.S - O, D) / dot(D, D);
        const V3 nr = cross(q.Ss, q.St);
        const double nlen = norm(nr);
        const V3 ng = nr / nlen;
        if (t_out)
            t_out[r] = tau;
        if (ng_out) {
            ng_out[3 * r] = ng.x;
            ng_out[3 * r + 1] = ng.y;
            ng_out[3 * r + 2] = ng.z;
        }
        const double lt = dl_dt ? dl_dt[r] : 0.0;
        const V3 ln = dl_dn ? v3(dl_dn[3 * r], dl_dn[3 * r + 1], dl_dn[3 * r + 2]) : V3{};
        M3 A;
        for (int row = 0; row < 3; ++row) {
            const double c[3][3] = {{q.Ss.x, q.St.x, -D.x}, {q.Ss.y, q.St.y, -D.y}, {q.Ss.z, q.St.z, -D.z}};
            for (int col = 0; col < 3; ++col)
                A.m[row][col] = c[row][col];
        }
        const M3 Ainv = inverse(A);
        const double n1 = g.W - 1;
        for (int c = 0; c < 4; ++c) {
            const int di = c & 1, dj = c >> 1;
            const double wu = di ? q.fs : 1.0 - q.fs, wv = dj ? q.ft : 1.0 - q.ft;
            const double w = wu * wv, ws = n1 * (di ? 1.0 : -1.0) * wv, wt = n1 * wu * (dj ? 1.0 : -1.0);
            const V3 dS = g.scale * w * q.N;
            const V3 dSs = g.scale * (ws * q.N + w * q.Nu), dSt = g.scale * (wt * q.N + w * q.Nv);
            const V3 x = Ainv.mul(-dS); // (ds, dt, dtau)
            const V3 DSs = dSs + x.x * q.Sss + x.y * q.Sst, DSt = dSt + x.x * q.Sst + x.y * q.Stt;
            const V3 dnr = cross(DSs, q.St) + cross(q.Ss, DSt);
            const V3 dng = (dnr - dot(ng, dnr) * ng) / nlen;
            node[4 * r + c] = (int64_t)(q.cj + dj) * g.W + (q.ci + di);
            val[4 * r + c] = lt * x.z + dot(ln, dng);
        }
    });
    for (int64_t q = 0; q < 4 * n; ++q)
        if (node[q] >= 0)
            grad_nodes[node[q]] += val[q];
}

void orc_scene_occluded(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                        const double *tmax, int bound, int slab_max_level, int leaf, int nthreads, uint8_t *occluded)
{
    const Scene *s = (const Scene *)scene;
    Options opt{bound, slab_max_level, leaf};
    parallel_rays(n, nthreads, [&](int64_t r) {
        Hit h;
        int tri = -1;
        occluded[r] = s->intersect(v3(o[3 * r], o[3 * r + 1], o[3 * r + 2]), v3(d[3 * r], d[3 * r + 1], d[3 * r + 2]),
                                   tmin[r], tmax[r], opt, h, tri, nullptr, true)
                          ? 1
                          : 0;
    });
}

int orc_triangle_intersect(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int bound,
                           int slab_max_level, int leaf, double *hit10, int64_t *stats)
{
    const Scene *s = (const Scene *)scene;
    Options opt{bound, slab_max_level, leaf};
    Hit h;
    Stats st;
    bool hit = s->its[k].intersect(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmin, tmax, opt, h, &st);
    if (hit)
        write_hit(h, hit10);
    if (stats)
        for (int q = 0; q < ORC_STATS; ++q)
            stats[q] = st.v[q];
    return hit ? 1 : 0;
}

int orc_leaf_newton(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                    int64_t j, double *hit10, int64_t *sub_squares)
{
    const Scene *s = (const Scene *)scene;
    Hit h;
    Stats st;
    bool hit = s->its[k].intersect_newton(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmin, tmax, i, j, h, &st);
    if (hit)
        write_hit(h, hit10);
    if (sub_squares)
        *sub_squares = st.v[ORC_ST_SUB];
    return hit ? 1 : 0;
}

int orc_newton_from(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                    int64_t j, double s0, double t0, double *hit10)
{
    const Scene *s = (const Scene *)scene;
    Hit h;
    bool hit = s->its[k].newton_from(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmin, tmax, i, j, v2(s0, t0), h,
                                     nullptr);
    if (hit)
        write_hit(h, hit10);
    return hit ? 1 : 0;
}

int orc_leaf_two_triangle(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                          int64_t j, double *hit10)
{
    const Scene *s = (const Scene *)scene;
    Hit h;
    bool hit = s->its[k].intersect_two_triangle(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), tmin, tmax, i, j, h);
    if (hit)
        write_hit(h, hit10);
    return hit ? 1 : 0;
}

void orc_node_bound(void *scene, int k, int level, int64_t i, int64_t j, double *box6, double *m6, double *centre2,
                    double *rect4, double *node8)
{
    const Scene *s = (const Scene *)scene;
    const Intersector &it = s->its[k];
    NodeBound nb = node_bound(*it.tri, it.frame, s->pyr, level, i, j);
    for (int a = 0; a < 3; ++a) {
        box6[a] = nb.box.lo[a];
        box6[3 + a] = nb.box.hi[a];
        Interval r = nb.m[a].interval();
        m6[2 * a] = r.lo;
        m6[2 * a + 1] = r.hi;
    }
    centre2[0] = nb.centre.x;
    centre2[1] = nb.centre.y;
    rect4[0] = nb.rect_lo.x;
    rect4[1] = nb.rect_lo.y;
    rect4[2] = nb.rect_hi.x;
    rect4[3] = nb.rect_hi.y;
    const Node &n = nb.node;
    double v[8] = {n.h0, n.gu, n.gv, n.r, n.ru, n.rv, n.h_min, n.h_max};
    std::memcpy(node8, v, sizeof v);
}

int orc_node_interval(void *scene, int k, int level, int64_t i, int64_t j, const double *o, const double *d,
                      double tmin, double tmax, int mode, double *t01)
{
    const Scene *s = (const Scene *)scene;
    const Intersector &it = s->its[k];
    NodeBound nb = node_bound(*it.tri, it.frame, s->pyr, level, i, j);
    TRay ray{it.frame.point(v3(o[0], o[1], o[2])), it.frame.direction(v3(d[0], d[1], d[2])), tmin, tmax};
    double t0 = 0.0, t1 = 0.0;
    bool ok = mode == 0 ? box_interval(nb, ray, t0, t1)
                        : mode == 1 ? slab_interval(nb, ray, t0, t1) : both_interval(nb, ray, t0, t1);
    t01[0] = t0;
    t01[1] = t1;
    return ok ? 1 : 0;
}

void orc_displaced_position(void *scene, int k, double s_, double t_, double *S3)
{
    const Scene *s = (const Scene *)scene;
    V3 S = displaced_position(s->tris[k], s->grid, s_, t_);
    S3[0] = S.x;
    S3[1] = S.y;
    S3[2] = S.z;
}

static V3 surface_normal(const Scene *s, int k, double s_, double t_)
{
    const Tri &tri = s->tris[k];
    NFrame f = normal_frame_at(tri, s_, t_);
    double h = s->grid.h(s_, t_);
    V2 gh = s->grid.grad(s_, t_);
    V3 Su = (tri.e1 + gh.x * f.N) + h * f.Nu;
    V3 Sv = (tri.e2 + gh.y * f.N) + h * f.Nv;
    return normalized(cross(Su, Sv));
}

void orc_surface_normal(void *scene, int k, double s_, double t_, double *n3)
{
    V3 n = surface_normal((const Scene *)scene, k, s_, t_);
    n3[0] = n.x;
    n3[1] = n.y;
    n3[2] = n.z;
}

void orc_surface_points(void *scene, int64_t n, const int32_t *tri, const double *st, double *S3, double *n3)
{
    const Scene *s = (const Scene *)scene;
    for (int64_t r = 0; r < n; ++r) {
        V3 S = displaced_position(s->tris[tri[r]], s->grid, st[2 * r], st[2 * r + 1]);
        V3 nn = surface_normal(s, tri[r], st[2 * r], st[2 * r + 1]);
        S3[3 * r] = S.x;
        S3[3 * r + 1] = S.y;
        S3[3 * r + 2] = S.z;
        n3[3 * r] = nn.x;
        n3[3 * r + 1] = nn.y;
        n3[3 * r + 2] = nn.z;
    }
}

void orc_field_eval(void *scene, int64_t n, const double *xy, double *h, double *grad2)
{
    const Scene *s = (const Scene *)scene;
    for (int64_t r = 0; r < n; ++r) {
        h[r] = s->grid.h(xy[2 * r], xy[2 * r + 1]);
        V2 g = s->grid.grad(xy[2 * r], xy[2 * r + 1]);
        grad2[2 * r] = g.x;
        grad2[2 * r + 1] = g.y;
    }
}

// AffineVec3::normalize (affine_form.cpp:87-95) on one affine vector. in12
// and out12 hold x, y, z as (c, a, b, k) each: c + a eu + b ev +- k.
void orc_affine_normalize(const double *in12, double *out12)
{
    AV3 v{AF{in12[0], in12[1], in12[2], in12[3]}, AF{in12[4], in12[5], in12[6], in12[7]},
          AF{in12[8], in12[9], in12[10], in12[11]}};
    normalize(v);
    const AF *f[3] = {&v.x, &v.y, &v.z};
    for (int a = 0; a < 3; ++a) {
        out12[4 * a] = f[a]->c;
        out12[4 * a + 1] = f[a]->a;
        out12[4 * a + 2] = f[a]->b;
        out12[4 * a + 3] = f[a]->k;
    }
}

// tangent_frame (node_bounds.cpp:13-23) of triangle k: origin (3) and the
// row-major to_tangent matrix (9).
void orc_tangent_frame(void *scene, int k, double *origin3, double *to_tangent9)
{
    const Frame &f = ((const Scene *)scene)->its[k].frame;
    origin3[0] = f.origin.x;
    origin3[1] = f.origin.y;
    origin3[2] = f.origin.z;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            to_tangent9[3 * r + c] = f.to_tangent.m[r][c];
}

int orc_classify_square(const double *t6, double cx, double cy, double s)
{
    UvTri u;
    u.init(v2(t6[0], t6[1]), v2(t6[2], t6[3]), v2(t6[4], t6[5]));
    return (int)classify_square(u, v2(cx, cy), s);
}

void orc_hit_record(void *scene, int k, double u, double v, const double *dir3, const double *uv_map, int twosided,
                    double *out20)
{
    // DisplacedGeometry::surface_point (displaced_geometry.cpp:126-152) with the object's uv map
    // (DisplacedObject::object_uv, scene_spec.cpp:382-388), then compute_intersection's geometric frame
    // (displaced_geometry.cpp:154-183) in single precision as there, in object space.
    const Scene *s = (const Scene *)scene;
    const Tri &tri = s->tris[k];
    const double su = uv_map[0], sv = uv_map[1], ou = uv_map[2], ov = uv_map[3];
    const bool flip = uv_map[4] != 0.0;
    V2 st = tri.param(u, v);
    NFrame f = normal_frame_at(tri, st.x, st.y);
    double h = s->grid.h(st.x, st.y);
    V2 gh = s->grid.grad(st.x, st.y);
    V3 Ss = (tri.e1 + gh.x * f.N) + h * f.Nu;
    V3 St = (tri.e2 + gh.y * f.N) + h * f.Nv;
    V3 p = tri.P(st.x, st.y) + h * f.N;
    V3 wind = cross(tri.p1 - tri.p0, tri.p2 - tri.p0);
    V3 ng = cross(Ss, St);
    if (dot(ng, wind) < 0.0)
        ng = -ng;
    ng = normalized(ng);
    if (!(std::isfinite(ng.x) && std::isfinite(ng.y) && std::isfinite(ng.z)))
        ng = normalized(wind);
    V3 dpdu = su * Ss, dpdv = (flip ? -sv : sv) * St;
    double uu = (st.x - ou) / su, vv = (st.y - ov) / sv;
    if (flip)
        vv = 1.0 - vv;
    // The frame in float.
    auto f3 = [](V3 a) { return std::array<float, 3>{(float)a.x, (float)a.y, (float)a.z}; };
    auto fcross = [](const std::array<float, 3> &a, const std::array<float, 3> &b) {
        return std::array<float, 3>{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    };
    auto fnorm = [](std::array<float, 3> a) {
        float l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        return std::array<float, 3>{a[0] / l, a[1] / l, a[2] / l};
    };
    auto finite = [](const std::array<float, 3> &a) {
        return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]);
    };
    std::array<float, 3> ngf = f3(ng), du = f3(dpdu), dv = f3(dpdv);
    std::array<float, 3> b = fnorm(fcross(ngf, du));
    if (!finite(b)) {
        std::array<float, 3> ax = std::abs(ngf[0]) < 0.9f ? std::array<float, 3>{1, 0, 0} : std::array<float, 3>{0, 1, 0};
        b = fnorm(fcross(ngf, ax));
    }
    std::array<float, 3> t = fnorm(fcross(b, ngf));
    if (twosided) {
        std::array<float, 3> dd = fnorm(std::array<float, 3>{(float)dir3[0], (float)dir3[1], (float)dir3[2]});
        float wo_ng = -(dd[0] * ngf[0] + dd[1] * ngf[1] + dd[2] * ngf[2]);
        if (wo_ng < 0.0f)
            for (int q = 0; q < 3; ++q) {
                ngf[q] = -ngf[q];
                du[q] = -du[q];
                t[q] = -t[q];
            }
    }
    double vals[20] = {p.x,   p.y,   p.z,   ngf[0], ngf[1], ngf[2], du[0], du[1], du[2], dv[0],
                       dv[1], dv[2], uu,    vv,     t[0],   t[1],   t[2],  b[0],  b[1],  b[2]};
    std::memcpy(out20, vals, sizeof vals);
}

void orc_surface_point(void *scene, int k, double u, double v, double *out12)
{
    // DisplacedGeometry::surface_point (displaced_geometry.cpp:126-152), uv scale 1, no flip.
    const Scene *s = (const Scene *)scene;
    const Tri &tri = s->tris[k];
    V2 st = tri.param(u, v);
    NFrame f = normal_frame_at(tri, st.x, st.y);
    double h = s->grid.h(st.x, st.y);
    V2 gh = s->grid.grad(st.x, st.y);
    V3 Ss = (tri.e1 + gh.x * f.N) + h * f.Nu;
    V3 St = (tri.e2 + gh.y * f.N) + h * f.Nv;
    V3 p = tri.P(st.x, st.y) + h * f.N;
    V3 ng = cross(Ss, St);
    V3 wind = cross(tri.p1 - tri.p0, tri.p2 - tri.p0);
    if (dot(ng, wind) < 0.0)
        ng = -ng;
    ng = normalized(ng);
    if (!(std::isfinite(ng.x) && std::isfinite(ng.y) && std::isfinite(ng.z)))
        ng = normalized(wind);
    double vals[12] = {p.x, p.y, p.z, ng.x, ng.y, ng.z, Ss.x, Ss.y, Ss.z, St.x, St.y, St.z};
    std::memcpy(out12, vals, sizeof vals);
}

} // extern "C"
