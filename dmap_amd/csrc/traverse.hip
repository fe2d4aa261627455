// Closest-hit and any-hit traversal of displaced triangles on gfx950
// (kernels #2 and #3): the stackless quadtree walk with the box / slab
// node tests and the certified Newton leaf, one ray per lane.
//
// Follows code/cpp/src/displaced_intersector.cpp (walk :85-121, :175-237;
// Newton leaf :239-388; two-triangle leaf :390-426; flat object loop
// :578-599) and node_bounds.cpp (:25-44, :83-211), restated for a
// wavefront:
//   * the per-lane state machine (next triangle / walk / leaf) runs the
//     walk until every lane of the wave has a leaf pending or is done, then
//     the leaves run together (while-while), so the expensive leaf search
//     is not serialised lane by lane;
//   * the leaf's sub-square search keeps its depth-first stack (at most 16
//     entries) in LDS, lane-interleaved;
//   * node data come from the Morton-ordered fp32 pyramid: a box test reads
//     one float2, a slab test one float4 more.
//
// Precision. The kernel is templated on the working type R. R = float is
// the product path; R = double is the FP64 device run of the same solve.
// Every interval the float path computes is widened by its rounding error
// (a few ulps of the magnitudes involved), so the float tests keep every
// node and sub-square the double tests keep. Quantities that cancel are
// formed in double once per walk or per leaf: the ray in tangent space,
// and the leaf's anchor projections d1 . (P0 - o), d2 . (P0 - o),
// dn . (P0 - o) about the texel corner P0; the Newton iteration then runs
// in texel-local coordinates (a, b) in [0, 1]^2, where its terms are small.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>

#include "kernels.h"
#include "layout.h"

namespace dmap_amd
{

template <class R> struct Num;
template <> struct Num<float>
{
    static constexpr float eps = 0x1p-24f;
    static constexpr float tol_rel = 16.0f * 0x1p-24f; // interval_tolerance's relative room
};
template <> struct Num<double>
{
    static constexpr double eps = 0x1p-53;
    static constexpr double tol_rel = 1e-9; // node_bounds.cpp:128
};

constexpr int KPAD = 8;        // ulps of padding per rounded interval
constexpr int KNEWTON = 16;    // ulps of the residual's magnitude accepted as converged
constexpr int SUBDIV_DEPTH = 5; // displaced_intersector.h:132
constexpr int MAX_ITERS = 20;   // displaced_intersector.h:129
constexpr int STACK = 16;       // 4 + 3 x 4 entries for a depth-5 search
// BVHs of at most this many leaves go to LDS (64 B per node + 4 B per leaf).
constexpr int LDS_BVH_LEAVES = 128;
constexpr size_t BVH_LDS_BYTES = (sizeof(BvhNode) + sizeof(int)) * LDS_BVH_LEAVES;
constexpr int BLOCK = 256;

template <class R> __device__ inline R pad_of(R magnitude) { return (R)KPAD * Num<R>::eps * magnitude; }
// max / min for the node test's bounds. In fp32 they propagate NaN
// (v_maximum3_f32 / v_minimum3_f32), which needs no quieting of the
// operands as fmax / fmin do. A NaN there comes only from an overflow, and
// every comparison it reaches is false, so the test passes: the bound only
// widens. In fp64 they are fmax / fmin.
template <class R> __device__ inline R nmax(R a, R b) { return fmax(a, b); }
template <class R> __device__ inline R nmin(R a, R b) { return fmin(a, b); }
template <> __device__ inline float nmax<float>(float a, float b) { return __builtin_elementwise_maximum(a, b); }
template <> __device__ inline float nmin<float>(float a, float b) { return __builtin_elementwise_minimum(a, b); }

template <class R> __device__ inline R tolerance(R t_scale) { return Num<R>::tol_rel * nmax((R)1, t_scale); }
// a / b for a bound that pad_of widens afterwards (the slab test's
// quotients, and the ray's inverse direction, which clip_axis and
// clip_centre use). In fp32 it is a times the hardware reciprocal
// (v_rcp_f32, within 1 ulp for a normal result): with the rounding of the
// numerator and of the product the result lies within 4 unit roundoffs of
// the exact quotient, inside the 8 of pad_of. A |b| of
// 2^126 or more (a denormal reciprocal) is scaled by 2^-64 first, exactly;
// a product that overflows where the quotient does not only widens the
// interval. In fp64 it is the division.
template <class R> __device__ inline R padded_div(R a, R b) { return a / b; }
template <> __device__ inline float padded_div<float>(float a, float b)
{
    const float k = fabsf(b) >= 0x1p126f ? 0x1p-64f : 1.0f;
    return (a * __builtin_amdgcn_rcpf(b * k)) * k;
}

// clip_axis for the slab [c - e, c + e] given by its centre and radius:
// the ray meets it over (c - o) / d -+ e / |d|, with no ordering of the two
// ends. Rounding room: pad (8 unit roundoffs of |o| + |c| + e) on the
// radius covers the reciprocal and the two products, and 8 unit roundoffs
// of |tm| + tr on the parameters cover the final sums. An overflow gives
// an infinite or NaN end: an infinite one is the exact end's rounding
// outward, and a NaN one passes the test (nmax).
template <class R> __device__ inline bool clip_centre(R o, R d, R inv, R c, R e, R &t0, R &t1)
{
    R pad = pad_of<R>(fabs(o) + fabs(c) + e);
    if (d != (R)0) {
        R tm = (c - o) * inv, tr = (e + pad) * fabs(inv);
        tr += pad_of<R>(fabs(tm) + tr);
        t0 = nmax(t0, tm - tr);
        t1 = nmin(t1, tm + tr);
    } else if (fabs(o - c) > e + pad) {
        return false;
    }
    return true;
}

// clip_centre for a node box whose centre c and radius e are themselves
// rounded results: one pad of 8 unit roundoffs of |o| + 2 (|c| + e) is
// the room for their rounding (8 of |c| + e) and for the clip's own (8 of
// |o| + |c| + e, which needs under 5), folded into one radius.
template <class R> __device__ inline bool clip_node_box(R o, R d, R inv, R c, R e, R &t0, R &t1)
{
    R E = e + pad_of<R>(fabs(o) + (R)2 * (fabs(c) + e));
    if (d != (R)0) {
        R tm = (c - o) * inv, tr = E * fabs(inv);
        tr += pad_of<R>(fabs(tm) + tr);
        t0 = nmax(t0, tm - tr);
        t1 = nmin(t1, tm + tr);
    } else if (fabs(o - c) > E) {
        return false;
    }
    return true;
}

// Aabb3::clip (node_bounds.cpp:25-44) on one axis, widened by rounding room.
template <class R> __device__ inline bool clip_axis(R o, R d, R inv, R lo, R hi, R &t0, R &t1)
{
    R pad = pad_of<R>(fabs(o) + nmax(fabs(lo), fabs(hi)));
    lo -= pad;
    hi += pad;
    if (d != (R)0) {
        R ta = (lo - o) * inv, tb = (hi - o) * inv;
        if (ta > tb) {
            R x = ta;
            ta = tb;
            tb = x;
        }
        ta -= pad_of<R>(fabs(ta));
        tb += pad_of<R>(fabs(tb));
        t0 = nmax(t0, ta);
        t1 = nmin(t1, tb);
    } else if (o < lo || o > hi) {
        return false;
    }
    return true;
}

template <class R> __device__ inline bool close_interval(R &t0, R &t1)
{
    if (t0 > t1 + tolerance<R>(nmax(fabs(t0), fabs(t1))))
        return false;
    if (t0 > t1) {
        R x = t0;
        t0 = t1;
        t1 = x;
    }
    return true;
}

template <class R> struct IvR
{
    R lo, hi;
};

// The normalized normal's affine form in tangent space over one cell:
// centre mc, symbol coefficients ma, mb, remainder mk per axis.
template <class R> struct MForm
{
    R c[3], a[3], b[3], k[3];
};

template <class R> struct AF
{
    R c, a, b, k;
};
template <class R> __device__ inline R af_dev(const AF<R> &f) { return fabs(f.a) + fabs(f.b) + f.k; }
// Product (affine_form.cpp:37-47) with the rounding room of the float path.
template <class R> __device__ inline AF<R> af_mul(const AF<R> &x, const AF<R> &r)
{
    R u = af_dev(x), v = af_dev(r);
    AF<R> o;
    o.c = x.c * r.c;
    o.a = x.c * r.a + r.c * x.a;
    o.b = x.c * r.b + r.c * x.b;
    o.k = fabs(r.c) * x.k + fabs(x.c) * r.k + u * v;
    o.k += pad_of<R>(fabs(o.c) + fabs(o.a) + fabs(o.b) + o.k);
    return o;
}

// AffineVec3::normalize (affine_form.cpp:78-95: sq_length, Min-Range
// rsqrt, product) with the rounding room of the float path. The input
// carries that room as remainders k_q, so the squared length bounds its
// cross term 2 sum_q c_q k_q e_q by 2 sum_q |c_q| k_q (affine_form.cpp:83
// writes 2 |c . k|, a bound only when every c_q k_q has one sign; the
// reference's own inputs have k = 0).
template <class R> __device__ void affine_normalize(AF<R> N[3])
{
    R dev[3], cc = 0, cu = 0, cv = 0, ck = 0, dd = 0;
    for (int q = 0; q < 3; ++q) {
        dev[q] = fabs(N[q].a) + fabs(N[q].b) + N[q].k;
        cc += N[q].c * N[q].c;
        cu += N[q].c * N[q].a;
        cv += N[q].c * N[q].b;
        ck += fabs(N[q].c) * N[q].k; // the cross term's bound over independent remainders
        dd += dev[q] * dev[q];
    }
    R offset = (R)0.5 * dd;
    AF<R> l{cc + offset, (R)2 * cu, (R)2 * cv, (R)2 * ck + offset};
    l.k += pad_of<R>(cc + fabs(l.a) + fabs(l.b) + l.k);
    R lo = l.c - af_dev(l), hi = l.c + af_dev(l);
    if (lo <= (R)0) {
        for (int q = 0; q < 3; ++q)
            N[q] = AF<R>{0, 0, 0, 1};
        return;
    }
    R glo = (R)1 / sqrt(lo), ghi = (R)1 / sqrt(hi);
    R alpha = (R)-0.5 * ghi * ghi * ghi;
    R beta = (R)0.5 * (glo + ghi - alpha * (lo + hi));
    R delta = (R)0.5 * fabs(glo - ghi - alpha * (lo - hi));
    delta += pad_of<R>(glo + ghi + fabs(alpha) * (lo + hi));
    AF<R> rs{alpha * l.c + beta, alpha * l.a, alpha * l.b, fabs(alpha) * l.k + delta};
    rs.k += pad_of<R>(fabs(alpha * l.c) + fabs(beta) + fabs(rs.a) + fabs(rs.b));
    for (int q = 0; q < 3; ++q)
        N[q] = af_mul(rs, N[q]);
}

// node_bound's normal form normalized (node_bounds.cpp:65-68, 100-103)
// over the rectangle (cx, cy) +- (hu, hv), mapped to tangent space. The
// general path, for a non-constant vertex normal.
template <class R> __device__ void normal_mform(const TriRecord &T, R cx, R cy, R hu, R hv, MForm<R> &out)
{
    AF<R> N[3];
    for (int q = 0; q < 3; ++q) {
        N[q].c = (R)T.m0[q] + cx * (R)T.Mu[q] + cy * (R)T.Mv[q];
        N[q].a = hu * (R)T.Mu[q];
        N[q].b = hv * (R)T.Mv[q];
        N[q].k = pad_of<R>(fabs((R)T.m0[q]) + fabs(cx * (R)T.Mu[q]) + fabs(cy * (R)T.Mv[q]));
    }
    affine_normalize<R>(N);
    // to_tangent * N (affine_form.cpp:100-106): rows of T.
    for (int r = 0; r < 3; ++r) {
        R c = 0, a = 0, b = 0, k = 0, mag = 0;
        for (int q = 0; q < 3; ++q) {
            R t = (R)T.T[3 * r + q];
            c += t * N[q].c;
            a += t * N[q].a;
            b += t * N[q].b;
            k += fabs(t) * N[q].k;
            mag += fabs(t) * (fabs(N[q].c) + fabs(N[q].a) + fabs(N[q].b));
        }
        out.c[r] = c;
        out.a[r] = a;
        out.b[r] = b;
        out.k[r] = k + pad_of<R>(mag);
    }
}

// The ray of one lane in a triangle's tangent space.
template <class R> struct Walk
{
    R o[3], d[3], inv[3];
    R omag, dmag;
    int si, sj; // the ray runs towards -i (-j): the far child along i (j) is met first
};

// classify_square (uv_clip.cpp:27-44), Outside only, with rounding room so
// that a cell the double test keeps is kept.
template <class R> __device__ inline bool square_outside(const TriRecord &T, R cx, R cy, R s)
{
    R m = pad_of<R>((R)1 + fabs(cx) + fabs(cy) + s);
    if (cx + s <= (R)T.dlo[0] - m || cx - s >= (R)T.dhi[0] + m || cy + s <= (R)T.dlo[1] - m ||
        cy - s >= (R)T.dhi[1] + m)
        return true;
    for (int k = 0; k < 3; ++k) {
        R nx = (R)T.dn[2 * k], ny = (R)T.dn[2 * k + 1];
        R proj = nx * cx + ny * cy;
        R ext = (fabs(nx) + fabs(ny)) * s;
        R mk = pad_of<R>(fabs(nx * cx) + fabs(ny * cy) + ext + fabs((R)T.dd[k]));
        if (proj + ext <= (R)T.dd[k] - mk)
            return true;
    }
    return false;
}
template <> __device__ inline bool square_outside<double>(const TriRecord &T, double cx, double cy, double s)
{
    double m = pad_of<double>(1.0 + fabs(cx) + fabs(cy) + s);
    if (cx + s <= T.dlo_d[0] - m || cx - s >= T.dhi_d[0] + m || cy + s <= T.dlo_d[1] - m || cy - s >= T.dhi_d[1] + m)
        return true;
    for (int k = 0; k < 3; ++k) {
        double nx = T.dn_d[2 * k], ny = T.dn_d[2 * k + 1];
        double proj = nx * cx + ny * cy;
        double ext = (fabs(nx) + fabs(ny)) * s;
        double mk = pad_of<double>(fabs(nx * cx) + fabs(ny * cy) + ext + fabs(T.dd_d[k]));
        if (proj + ext <= T.dd_d[k] - mk)
            return true;
    }
    return false;
}

// The square lies inside the triangle's domain with at least
// square_outside's rounding room to spare on every side, so square_outside
// is false for every square within it: a node group whose 2 x 2 block passes
// this skips its four children's tests.
template <class R> __device__ inline bool square_inside(const TriRecord &T, R cx, R cy, R s)
{
    R m = pad_of<R>((R)1 + fabs(cx) + fabs(cy) + s);
    if (!(cx - s >= (R)T.dlo[0] + m && cx + s <= (R)T.dhi[0] - m && cy - s >= (R)T.dlo[1] + m &&
          cy + s <= (R)T.dhi[1] - m))
        return false;
    for (int k = 0; k < 3; ++k) {
        R nx = (R)T.dn[2 * k], ny = (R)T.dn[2 * k + 1];
        R proj = nx * cx + ny * cy;
        R ext = (fabs(nx) + fabs(ny)) * s;
        R mk = pad_of<R>(fabs(nx * cx) + fabs(ny * cy) + ext + fabs((R)T.dd[k]));
        if (!(proj - ext >= (R)T.dd[k] + mk))
            return false;
    }
    return true;
}
template <> __device__ inline bool square_inside<double>(const TriRecord &T, double cx, double cy, double s)
{
    double m = pad_of<double>(1.0 + fabs(cx) + fabs(cy) + s);
    if (!(cx - s >= T.dlo_d[0] + m && cx + s <= T.dhi_d[0] - m && cy - s >= T.dlo_d[1] + m &&
          cy + s <= T.dhi_d[1] - m))
        return false;
    for (int k = 0; k < 3; ++k) {
        double nx = T.dn_d[2 * k], ny = T.dn_d[2 * k + 1];
        double proj = nx * cx + ny * cy;
        double ext = (fabs(nx) + fabs(ny)) * s;
        double mk = pad_of<double>(fabs(nx * cx) + fabs(ny * cy) + ext + fabs(T.dd_d[k]));
        if (!(proj - ext >= T.dd_d[k] + mk))
            return false;
    }
    return true;
}

// UvTriangle::inside (uv_clip.h:14-20).
template <class R> __device__ inline bool domain_inside(const TriRecord &T, R s, R t)
{
    for (int k = 0; k < 3; ++k)
        if ((R)T.dn_d[2 * k] * s + (R)T.dn_d[2 * k + 1] * t < (R)T.dd_d[k])
            return false;
    return true;
}

__device__ inline float2 fetch_mm(const PyramidView &P, int level, int i, int j)
{
    if (level >= P.n_levels)
        return make_float2(P.root_hmin, P.root_hmax);
    uint32_t m = (uint32_t)P.n_leaf >> level;
    uint32_t idx = level_base(m) + morton2((uint32_t)i & (m - 1), (uint32_t)j & (m - 1));
    return P.mm[idx];
}
__device__ inline float4 fetch_tp(const PyramidView &P, int level, int i, int j)
{
    if (level >= P.n_levels) {
        // Above the root: the root's min-max node (taylor_pyramid.cpp:261-266).
        return make_float4(0.5f * (P.root_hmin + P.root_hmax), 0.f, 0.f, 0.5f * (P.root_hmax - P.root_hmin));
    }
    uint32_t m = (uint32_t)P.n_leaf >> level;
    uint32_t idx = level_base(m) + morton2((uint32_t)i & (m - 1), (uint32_t)j & (m - 1));
    return P.tp[idx];
}

// One axis of a cell [k w, (k + 1) w] clipped to the triangle's texture
// bounding box [lo, hi] (node_bounds.cpp:94-99), as centre and half-width.
template <class R> __device__ inline void cell_span(int k, R w, R lo, R hi, R &c, R &h)
{
    R clo = (R)k * w;
    R rlo = nmax(clo, lo), rhi = nmin(clo + w, hi);
    c = (R)0.5 * (rlo + rhi);
    h = nmax((R)0.5 * (rhi - rlo), (R)0);
}

// node_bound + node_interval (node_bounds.cpp:83-109, 111-211) for cell
// (level, ci, cj), clipped to the triangle's texture bounding box as centre
// (cx, cy) and half-widths (hx, hy) (cell_span), with its min-max pair mm
// and, when the slab is read, its Taylor quad tp; bound 0 box, 1 slab,
// 2 both, the slab read when read_slab (else the box alone). Returns
// whether the ray can meet the surface over the cell within [t_min, t_max],
// and then the interval in [r0, r1].
template <class R, bool CN>
__device__ bool cell_interval(int bound, bool read_slab, const PyramidView &P, const TriRecord &T, const Walk<R> &wk,
                              int level, int ci, int cj, R cx, R cy, R hx, R hy, float2 mm, float4 tp, R t_min,
                              R t_max, R &r0, R &r1)
{
    R w = ldexp((R)1, level - (__ffs(P.n_leaf) - 1)); // 2^level / n_leaf, n_leaf a power of two

    MForm<R> m;
    if (CN || T.const_normal) {
        for (int q = 0; q < 3; ++q) {
            m.c[q] = sizeof(R) == 8 ? (R)T.mc_d[q] : (R)T.mc[q];
            m.a[q] = 0;
            m.b[q] = 0;
            m.k[q] = sizeof(R) == 8 ? (R)T.mk_d[q] : (R)T.mk[q];
        }
    } else {
        normal_mform<R>(T, cx, cy, hx, hy, m);
    }

    R hc = (R)0.5 * ((R)mm.x + (R)mm.y), hr = (R)0.5 * ((R)mm.y - (R)mm.x);
    // box = hull(p + h m): p = (cx +- hx, cy +- hy, 0), h = hc +- hr.
    R bc[3], be[3]; // the box: centre +- radius per axis
    const R pc[3] = {cx, cy, (R)0}, pa[3] = {hx, (R)0, (R)0}, pb[3] = {(R)0, hy, (R)0};
    for (int q = 0; q < 3; ++q) {
        // Under a constant normal (CN) the symbol coefficients are zero and
        // these are the same values without the zero terms.
        R dev_m = CN ? m.k[q] : fabs(m.a[q]) + fabs(m.b[q]) + m.k[q];
        // (z: the base point's component is 0.)
        R c = q == 2 ? hc * m.c[q] : pc[q] + hc * m.c[q];
        R a = CN ? pa[q] : pa[q] + hc * m.a[q];
        R b = CN ? pb[q] : pb[q] + hc * m.b[q];
        // In fp32 under a constant normal, hr (|mc| + mk) with the sum read
        // rounded up (TriRecord::mck): one product fewer, never smaller.
        R k = (CN && sizeof(R) == sizeof(float)) ? hr * (R)T.mck[q] + fabs(hc) * m.k[q]
                                                  : fabs(m.c[q]) * hr + fabs(hc) * m.k[q] + hr * dev_m;
        // (CN: |a| + |b| is hx, hy or 0 on x, y, z.)
        R e = CN ? (q == 2 ? k : (q == 0 ? hx : hy) + k) : fabs(a) + fabs(b) + k;
        bc[q] = c; // unpadded: clip_node_box adds the rounding room of c and e
        be[q] = e;
    }

    bool want_box = bound == 0 || bound == 2 || !read_slab;
    bool want_slab = bound != 0 && read_slab;

    // The box's x and y clips are the column's (slab_interval,
    // node_bounds.cpp:137-179): both start from [t_min, t_max], so the two
    // are clipped once and shared.
    R t0 = t_min, t1 = t_max;
    if (!clip_node_box<R>(wk.o[0], wk.d[0], wk.inv[0], bc[0], be[0], t0, t1) ||
        !clip_node_box<R>(wk.o[1], wk.d[1], wk.inv[1], bc[1], be[1], t0, t1))
        return false;
    R b0 = t0, b1 = t1;
    if (want_box) {
        if (!clip_node_box<R>(wk.o[2], wk.d[2], wk.inv[2], bc[2], be[2], b0, b1) || !close_interval<R>(b0, b1))
            return false;
        if (!want_slab) {
            r0 = b0;
            r1 = b1;
            return true;
        }
    }

    // slab_interval: the column.
    if (!close_interval<R>(t0, t1))
        return false;
    R anchor[3];
    for (int q = 0; q < 3; ++q)
        anchor[q] = wk.o[q] + t0 * wk.d[q];
    R span = t1 - t0;
    R h0 = tp.x, gu = tp.y, gv = tp.z, rr = tp.w;
    if (level >= P.n_levels) {
        gu = 0;
        gv = 0;
    }
    R ucx = ((R)ci + (R)0.5) * w, ucy = ((R)cj + (R)0.5) * w; // nb.centre
    // The interval products and sums in centre-radius form: s [c - r, c + r]
    // = [s c - |s| r, s c + |s| r] and sums add centres and radii, so in
    // exact arithmetic the intervals equal the endpoint forms, with no sign
    // selects; the rounding of either form is inside pad_of(err) below.
    const R xc = m.c[0], yc = m.c[1], zc = m.c[2];
    const R xr = CN ? m.k[0] : fabs(m.a[0]) + fabs(m.b[0]) + m.k[0],
            yr = CN ? m.k[1] : fabs(m.a[1]) + fabs(m.b[1]) + m.k[1],
            zr = CN ? m.k[2] : fabs(m.a[2]) + fabs(m.b[2]) + m.k[2];
    // Magnitudes: the largest |value| of each form, |c| + r (in fp32 under a
    // constant normal the per-triangle sums rounded up, TriRecord::mck).
    const bool pre = CN && sizeof(R) == sizeof(float);
    const R xm = pre ? (R)T.mck[0] : fabs(xc) + xr, ym = pre ? (R)T.mck[1] : fabs(yc) + yr,
            zm = pre ? (R)T.mck[2] : fabs(zc) + zr;
    const R Ac = (R)1 + gu * xc + gv * yc, Ar = fabs(gu) * xr + fabs(gv) * yr;
    const R Bc = -gu * zc, Br = fabs(gu) * zr, Cc = -gv * zc, Cr = fabs(gv) * zr;
    R dconst = -(h0 - gu * ucx - gv * ucy);
    const R Dc = dconst * zc, Dr = fabs(dconst) * zr;
    R Rr = rr * zm;
    const R F0c = anchor[2] * Ac + anchor[0] * Bc + anchor[1] * Cc + Dc;
    const R F0r = fabs(anchor[2]) * Ar + fabs(anchor[0]) * Br + fabs(anchor[1]) * Cr + Dr;
    const R F1c = wk.d[2] * Ac + wk.d[0] * Bc + wk.d[1] * Cc;
    const R F1r = fabs(wk.d[2]) * Ar + fabs(wk.d[0]) * Br + fabs(wk.d[1]) * Cr;
    const IvR<R> F0{F0c - F0r, F0c + F0r}, F1{F1c - F1r, F1c + F1r};
    // Rounding room of everything above, folded into the remainder R.
    R magA = (R)1 + fabs(gu) * xm + fabs(gv) * ym, magB = fabs(gu) * zm, magC = fabs(gv) * zm;
    R magD = (fabs(h0) + fabs(gu * ucx) + fabs(gv * ucy)) * zm;
    R err = (fabs(anchor[2]) + span * fabs(wk.d[2])) * magA + (fabs(anchor[0]) + span * fabs(wk.d[0])) * magB +
            (fabs(anchor[1]) + span * fabs(wk.d[1])) * magC + magD +
            (wk.omag + fabs(t0) * wk.dmag) * (magA + magB + magC) + Rr + (fabs(F0c) + F0r);
    Rr += pad_of<R>(err);
    R s0 = 0, s1 = span;
    if (F1.lo > (R)0)
        s1 = nmin(s1, padded_div<R>(Rr - F0.lo, F1.lo));
    else if (F1.lo < (R)0)
        s0 = nmax(s0, padded_div<R>(Rr - F0.lo, F1.lo));
    else if (F0.lo > Rr)
        return false;
    if (F1.hi > (R)0)
        s0 = nmax(s0, padded_div<R>(-Rr - F0.hi, F1.hi));
    else if (F1.hi < (R)0)
        s1 = nmin(s1, padded_div<R>(-Rr - F0.hi, F1.hi));
    else if (F0.hi < -Rr)
        return false;
    s0 -= pad_of<R>(fabs(s0));
    s1 += pad_of<R>(fabs(s1));
    if (s0 > s1 + tolerance<R>(fabs(t0) + span))
        return false;
    if (s0 > s1) {
        R x = s0;
        s0 = s1;
        s1 = x;
    }
    R st1 = t0 + s1, st0 = t0 + s0;
    if (!want_box) {
        r0 = st0;
        r1 = st1;
        return true;
    }
    // both_interval (node_bounds.cpp:181-197).
    R u0 = nmax(b0, st0), u1 = nmin(b1, st1);
    if (!close_interval<R>(u0, u1))
        return false;
    r0 = u0;
    r1 = u1;
    return true;
}

// node_interval with the options' bound mode: whether the ray can meet the
// surface over the cell within [t_min, t_max].
template <class R, bool CN>
__device__ inline bool cell_test(const TraceParams &p, const PyramidView &P, const TriRecord &T, const Walk<R> &wk,
                                 int level, int ci, int cj, R cx, R cy, R hx, R hy, float2 mm, float4 tp, R t_min,
                                 R t_max)
{
    R r0, r1;
    return cell_interval<R, CN>(p.bound, level <= p.slab_max_level, P, T, wk, level, ci, cj, cx, cy, hx, hy, mm, tp,
                                t_min, t_max, r0, r1);
}

// record_node_test (displaced_intersector.cpp:159-172) for one node test of
// the walk: the box's and the slab's decisions with the walk's t_max, and
// the three interval lengths with t_max = inf, the slab read at any level.
// Appended at the next slot of p.stats.records (counted past its capacity,
// not written).
template <class R, bool CN>
__device__ void record_node_test(const TraceParams &p, const PyramidView &P, const TriRecord &T, const Walk<R> &wk,
                                 int level, int ci, int cj, R cx, R cy, R hx, R hy, float2 mm, R t_min, R t_max,
                                 int64_t rid, int prim, int seq)
{
    const float4 tp = fetch_tp(P, level, ci, cj);
    NodeRecordDev rec;
    rec.ray = (int32_t)rid;
    rec.prim = prim;
    rec.level = level;
    rec.i = ci;
    rec.j = cj;
    rec.seq = seq;
    R r0, r1;
    rec.box_pass = cell_interval<R, CN>(0, true, P, T, wk, level, ci, cj, cx, cy, hx, hy, mm, tp, t_min, t_max, r0, r1);
    rec.slab_pass = cell_interval<R, CN>(1, true, P, T, wk, level, ci, cj, cx, cy, hx, hy, mm, tp, t_min, t_max, r0, r1);
    const R inf = (R)INFINITY;
    float len[3];
    for (int mode = 0; mode < 3; ++mode)
        len[mode] = cell_interval<R, CN>(mode, true, P, T, wk, level, ci, cj, cx, cy, hx, hy, mm, tp, t_min, inf, r0,
                                         r1)
                        ? (float)(r1 - r0)
                        : -1.f;
    rec.box_length = len[0];
    rec.slab_length = len[1];
    rec.both_length = len[2];
    rec.pad_ = 0.f;
    const unsigned long long slot = atomicAdd(p.stats.record_count, 1ull);
    if (slot < (unsigned long long)p.stats.record_capacity)
        p.stats.records[slot] = rec;
}

// ------------------------------------------------------------------ leaf

// Per-leaf constants of the certified Newton leaf, in texel-local
// coordinates (a, b) in [0, 1]^2: S(a, b) - o = (A - o) + a E1 + b E2 +
// (h(a, b) - shift) N(a, b), E1 = w e1, E2 = w e2, P0 = P(i w, j w), with the
// anchor A = P0 + h00 N and shift = h00 under a constant normal, A = P0 and
// shift = 0 otherwise. The anchor projections are formed in double.
template <class R> struct Leaf
{
    R c1, c2, cn;                     // d1, d2, dn . (P0 - o)
    R p1a, p1b, p2a, p2b, pna, pnb;   // d1, d2, dn . E1, E2
    R h00, ha, hb, tw;                // h - shift = h00 + ha a + hb b + tw a b (shift: the anchor height)
    R n1, n2, nn;                     // d1, d2, dn . N (constant normal)
    R a1, a2, an;                     // sum |d_x| |N_x|
    R k1, k2, kn;                     // sum |d_x| N remainder_x
    R E1[3], E2[3], N[3];             // object-space vectors for the normal
    R d1[3], d2[3], dn[3];            // ray basis (general normal path)
    R M0[3], wMu[3], wMv[3];          // interpolated normal over the texel (general path)
    R hmax, magF, tol2, d_len;
    R slo, tlo, w;
    int i, j;
};

// plane_basis (displaced_intersector.cpp:22-27) in double.
__device__ inline void plane_basis(const double n[3], double d1[3], double d2[3])
{
    double a[3] = {fabs(n[0]) < 0.9 ? 1.0 : 0.0, fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    double c[3] = {a[1] * n[2] - a[2] * n[1], a[2] * n[0] - a[0] * n[2], a[0] * n[1] - a[1] * n[0]};
    double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int q = 0; q < 3; ++q)
        d1[q] = c[q] / len;
    d2[0] = n[1] * d1[2] - n[2] * d1[1];
    d2[1] = n[2] * d1[0] - n[0] * d1[2];
    d2[2] = n[0] * d1[1] - n[1] * d1[0];
}

__device__ inline float node_value(const PyramidView &P, int i, int j)
{
    // HeightGrid::h at a node (displaced_surface.cpp:12-39): node (i, j) of
    // the closed grid, the texture wrapping (or clamping) outside the tile.
    int n = P.n_leaf;
    if (P.repeat) {
        i &= n - 1;
        j &= n - 1;
        // the closed grid's last row and column repeat the first, so the
        // wrapped index n never occurs: nodes (n, .) equal (0, .).
    } else {
        i = min(max(i, 0), n);
        j = min(max(j, 0), n);
    }
    return P.nodes[(size_t)j * (n + 1) + i];
}

template <class R>
__device__ void leaf_setup(const PyramidView &P, const TriRecord &T, const double o[3], const double d[3], int i,
                           int j, Leaf<R> &L)
{
    double w = 1.0 / P.n_leaf;
    double slo = i * w, tlo = j * w;
    // Corner heights h = scale * value + offset (HeightGrid::h at the corners).
    double h00 = P.scale_d * (double)node_value(P, i, j) + P.offset_d;
    double h10 = P.scale_d * (double)node_value(P, i + 1, j) + P.offset_d;
    double h01 = P.scale_d * (double)node_value(P, i, j + 1) + P.offset_d;
    double h11 = P.scale_d * (double)node_value(P, i + 1, j + 1) + P.offset_d;
    // The anchor: the displaced texel corner P0 + h00 N under a constant
    // normal (heights then enter relative to h00 and stay small), the base
    // corner P0 otherwise.
    double shift = T.const_normal ? h00 : 0.0;
    double P0[3], rel[3];
    for (int q = 0; q < 3; ++q) {
        P0[q] = T.q0[q] + slo * T.e1[q] + tlo * T.e2[q];
        rel[q] = P0[q] + (T.const_normal ? h00 * T.nhat_d[q] : 0.0) - o[q];
    }
    double d_len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double dn[3] = {d[0] / d_len, d[1] / d_len, d[2] / d_len}, d1[3], d2[3];
    plane_basis(dn, d1, d2);
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    double E1[3] = {w * T.e1[0], w * T.e1[1], w * T.e1[2]}, E2[3] = {w * T.e2[0], w * T.e2[1], w * T.e2[2]};
    L.c1 = (R)dot(d1, rel);
    L.c2 = (R)dot(d2, rel);
    L.cn = (R)dot(dn, rel);
    L.p1a = (R)dot(d1, E1);
    L.p1b = (R)dot(d1, E2);
    L.p2a = (R)dot(d2, E1);
    L.p2b = (R)dot(d2, E2);
    L.pna = (R)dot(dn, E1);
    L.pnb = (R)dot(dn, E2);
    for (int q = 0; q < 3; ++q) {
        L.E1[q] = (R)E1[q];
        L.E2[q] = (R)E2[q];
        L.d1[q] = (R)d1[q];
        L.d2[q] = (R)d2[q];
        L.dn[q] = (R)dn[q];
    }
    // h - shift = h00 + ha a + hb b + tw a b over the texel.
    L.h00 = (R)(h00 - shift);
    L.ha = (R)(h10 - h00);
    L.hb = (R)(h01 - h00);
    L.tw = (R)(((h11 - h10) - h01) + h00);
    L.hmax = (R)fmax(fmax(fabs(h00 - shift), fabs(h10 - shift)), fmax(fabs(h01 - shift), fabs(h11 - shift)));
    if (T.const_normal) {
        double N[3] = {T.nhat_d[0], T.nhat_d[1], T.nhat_d[2]};
        double nk = T.nhat_kd;
        L.n1 = (R)dot(d1, N);
        L.n2 = (R)dot(d2, N);
        L.nn = (R)dot(dn, N);
        L.a1 = (R)(fabs(d1[0] * N[0]) + fabs(d1[1] * N[1]) + fabs(d1[2] * N[2]));
        L.a2 = (R)(fabs(d2[0] * N[0]) + fabs(d2[1] * N[1]) + fabs(d2[2] * N[2]));
        L.an = (R)(fabs(dn[0] * N[0]) + fabs(dn[1] * N[1]) + fabs(dn[2] * N[2]));
        L.k1 = (R)(nk * (fabs(d1[0]) + fabs(d1[1]) + fabs(d1[2])));
        L.k2 = (R)(nk * (fabs(d2[0]) + fabs(d2[1]) + fabs(d2[2])));
        L.kn = (R)(nk * (fabs(dn[0]) + fabs(dn[1]) + fabs(dn[2])));
        for (int q = 0; q < 3; ++q)
            L.N[q] = (R)N[q];
    } else {
        for (int q = 0; q < 3; ++q) {
            L.M0[q] = (R)(T.m0[q] + slo * T.Mu[q] + tlo * T.Mv[q]);
            L.wMu[q] = (R)(w * T.Mu[q]);
            L.wMv[q] = (R)(w * T.Mv[q]);
        }
    }
    L.d_len = (R)d_len;
    L.slo = (R)slo;
    L.tlo = (R)tlo;
    L.w = (R)w;
    L.i = i;
    L.j = j;
    R magF = fmax(fabs(L.c1), fabs(L.c2)) + fabs(L.p1a) + fabs(L.p1b) + fabs(L.p2a) + fabs(L.p2b) + L.hmax;
    L.magF = magF;
    R tol = fmax((R)T.residual_tolerance, (R)KNEWTON * Num<R>::eps * magF);
    L.tol2 = tol * tol;
}

// The normal N(a, b), its a- and b-derivatives, over the texel (general
// normal path): N = M / |M|, N_a = (wMu - N (N . wMu)) / |M|
// (normal_frame_at, displaced_surface.cpp:99-108, in local units).
template <class R> __device__ inline void normal_at(const Leaf<R> &L, R a, R b, R N[3], R Na[3], R Nb[3])
{
    R M[3];
    for (int q = 0; q < 3; ++q)
        M[q] = L.M0[q] + a * L.wMu[q] + b * L.wMv[q];
    R len = sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
    for (int q = 0; q < 3; ++q)
        N[q] = M[q] / len;
    R nu = N[0] * L.wMu[0] + N[1] * L.wMu[1] + N[2] * L.wMu[2];
    R nv = N[0] * L.wMv[0] + N[1] * L.wMv[1] + N[2] * L.wMv[2];
    for (int q = 0; q < 3; ++q) {
        Na[q] = (L.wMu[q] - N[q] * nu) / len;
        Nb[q] = (L.wMv[q] - N[q] * nv) / len;
    }
}

// The normalized normal over the whole texel as an affine form in object
// space (intersect_newton, displaced_intersector.cpp:250-251): centre at the
// texel centre, symbol coefficients w Mu / 2 and w Mv / 2.
template <class R> __device__ void normal_texel_form(const Leaf<R> &L, MForm<R> &nf)
{
    AF<R> N[3];
    for (int q = 0; q < 3; ++q) {
        N[q].c = L.M0[q] + (R)0.5 * L.wMu[q] + (R)0.5 * L.wMv[q];
        N[q].a = (R)0.5 * L.wMu[q];
        N[q].b = (R)0.5 * L.wMv[q];
        N[q].k = pad_of<R>(fabs(L.M0[q]) + fabs(L.wMu[q]) + fabs(L.wMv[q]));
    }
    affine_normalize<R>(N);
    for (int q = 0; q < 3; ++q) {
        nf.c[q] = N[q].c;
        nf.a[q] = N[q].a;
        nf.b[q] = N[q].b;
        nf.k[q] = N[q].k;
    }
}

// sub_square_bound (displaced_intersector.cpp:266-290) on the sub-square
// with centre (a, b) and half-width rho (texel units). `inside`: the whole
// texel passes square_inside, so no sub-square of it is outside.
template <class R, bool CN>
__device__ bool sub_square_bound(const TriRecord &T, const Leaf<R> &L, R a, R b, R rho, R t_min, R t_max, R &t_lo,
                                 bool inside)
{
    if (!inside && square_outside<R>(T, L.slo + L.w * a, L.tlo + L.w * b, L.w * rho))
        return false;
    // h over the sub-square: affine up to the twist (rho eu)(rho ev).
    R hc = ((L.h00 + L.ha * a) + L.hb * b) + L.tw * a * b;
    R hA = (L.ha + L.tw * b) * rho, hB = (L.hb + L.tw * a) * rho, hk = fabs(L.tw) * rho * rho;
    R hmag = fabs(hc) + fabs(hA) + fabs(hB) + hk;
    R F1c, F1a, F1b, F1k, F2c, F2a, F2b, F2k, Gc, Ga, Gb, Gk;
    if (CN || T.const_normal) {
        F1c = L.c1 + a * L.p1a + b * L.p1b + hc * L.n1;
        F1a = rho * L.p1a + hA * L.n1;
        F1b = rho * L.p1b + hB * L.n1;
        F1k = hk * L.a1 + hmag * L.k1;
        F2c = L.c2 + a * L.p2a + b * L.p2b + hc * L.n2;
        F2a = rho * L.p2a + hA * L.n2;
        F2b = rho * L.p2b + hB * L.n2;
        F2k = hk * L.a2 + hmag * L.k2;
        Gc = L.cn + a * L.pna + b * L.pnb + hc * L.nn;
        Ga = rho * L.pna + hA * L.nn;
        Gb = rho * L.pnb + hB * L.nn;
        Gk = hk * L.an + hmag * L.kn;
    } else {
        // N_hat over the texel, restricted to the sub-square
        // (displaced_intersector.cpp:72-81, 250-251).
        MForm<R> nf;
        normal_texel_form(L, nf);
        AF<R> h{hc, hA, hB, hk};
        AF<R> S[3];
        R alpha = (R)2 * a - (R)1, beta = (R)2 * b - (R)1, r2 = (R)2 * rho;
        for (int q = 0; q < 3; ++q) {
            AF<R> nq{nf.c[q] + nf.a[q] * alpha + nf.b[q] * beta, nf.a[q] * r2, nf.b[q] * r2, nf.k[q]};
            AF<R> hn = af_mul(h, nq);
            S[q] = AF<R>{a * L.E1[q] + b * L.E2[q] + hn.c, rho * L.E1[q] + hn.a, rho * L.E2[q] + hn.b, hn.k};
        }
        auto proj = [&](const R *dv, R base, R &c, R &aa, R &bb, R &kk) {
            c = base + dv[0] * S[0].c + dv[1] * S[1].c + dv[2] * S[2].c;
            aa = dv[0] * S[0].a + dv[1] * S[1].a + dv[2] * S[2].a;
            bb = dv[0] * S[0].b + dv[1] * S[1].b + dv[2] * S[2].b;
            kk = fabs(dv[0]) * S[0].k + fabs(dv[1]) * S[1].k + fabs(dv[2]) * S[2].k;
        };
        proj(L.d1, L.c1, F1c, F1a, F1b, F1k);
        proj(L.d2, L.c2, F2c, F2a, F2b, F2k);
        proj(L.dn, L.cn, Gc, Ga, Gb, Gk);
    }
    R local = (fabs(a) + rho) * (fabs(L.p1a) + fabs(L.p1b) + fabs(L.p2a) + fabs(L.p2b) + fabs(L.pna) + fabs(L.pnb)) +
              hmag;
    R e1 = fabs(F1a) + fabs(F1b) + F1k + pad_of<R>(fabs(L.c1) + local);
    R e2 = fabs(F2a) + fabs(F2b) + F2k + pad_of<R>(fabs(L.c2) + local);
    if (!(F1c - e1 <= (R)0 && (R)0 <= F1c + e1) || !(F2c - e2 <= (R)0 && (R)0 <= F2c + e2))
        return false;
    R eg = fabs(Ga) + fabs(Gb) + Gk + pad_of<R>(fabs(L.cn) + local);
    t_lo = padded_div<R>(Gc - eg, L.d_len);
    t_lo -= pad_of<R>(fabs(t_lo));
    R t_hi = padded_div<R>(Gc + eg, L.d_len);
    t_hi += pad_of<R>(fabs(t_hi));
    return t_lo < t_max && t_hi > t_min;
}

// F = ((S - o) . d1, (S - o) . d2), the distance along the ray and the
// Jacobian of F at texel-local (a, b) (displaced_intersector.cpp:337-375).
template <class R, bool CN>
__device__ inline void newton_eval(const TriRecord &T, const Leaf<R> &L, R a, R b, R &F1, R &F2, R &along, R &J00,
                                   R &J01, R &J10, R &J11)
{
    R h = ((L.h00 + L.ha * a) + L.hb * b) + L.tw * a * b;
    R hda = L.ha + L.tw * b, hdb = L.hb + L.tw * a;
    if (CN || T.const_normal) {
        F1 = L.c1 + a * L.p1a + b * L.p1b + h * L.n1;
        F2 = L.c2 + a * L.p2a + b * L.p2b + h * L.n2;
        along = L.cn + a * L.pna + b * L.pnb + h * L.nn;
        J00 = L.p1a + hda * L.n1;
        J01 = L.p1b + hdb * L.n1;
        J10 = L.p2a + hda * L.n2;
        J11 = L.p2b + hdb * L.n2;
    } else {
        R N[3], Na[3], Nb[3];
        normal_at(L, a, b, N, Na, Nb);
        R Sl[3], Sa[3], Sb[3];
        for (int q = 0; q < 3; ++q) {
            Sl[q] = a * L.E1[q] + b * L.E2[q] + h * N[q];
            Sa[q] = L.E1[q] + hda * N[q] + h * Na[q];
            Sb[q] = L.E2[q] + hdb * N[q] + h * Nb[q];
        }
        auto dt = [](const R *x, const R *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
        F1 = L.c1 + dt(L.d1, Sl);
        F2 = L.c2 + dt(L.d2, Sl);
        along = L.cn + dt(L.dn, Sl);
        J00 = dt(L.d1, Sa);
        J01 = dt(L.d1, Sb);
        J10 = dt(L.d2, Sa);
        J11 = dt(L.d2, Sb);
    }
}

// Newton steps in double from an accepted root, for the reported values
// only: the fp32 iteration stops at a residual of a few ulps of its
// magnitudes, two double steps bring the root to the double solve's. A
// step that would leave the texel is not taken.
__device__ inline void polish_root(const TriRecord &T, const Leaf<double> &L, double &a, double &b)
{
    for (int it = 0; it < 2; ++it) {
        double F1, F2, along, J00, J01, J10, J11;
        newton_eval<double, false>(T, L, a, b, F1, F2, along, J00, J01, J10, J11);
        double det = J00 * J11 - J01 * J10;
        if (det == 0.0)
            return;
        double na = a - (J11 * F1 - J01 * F2) / det, nb = b - (-J10 * F1 + J00 * F2) / det;
        if (!(na >= 0.0 && na <= 1.0 && nb >= 0.0 && nb <= 1.0))
            return;
        a = na;
        b = nb;
    }
}

// newton_from (displaced_intersector.cpp:325-388) in texel-local
// coordinates. On success writes the root (a, b) and its t.
template <class R, bool CN>
__device__ bool newton_from(const TriRecord &T, const Leaf<R> &L, R a, R b, R t_min, R t_max, R &out_a, R &out_b,
                            R &out_t, int &iters)
{
    R prev_err2 = (R)INFINITY;
    int err_streak = 0, behind_streak = 0, invalid_streak = 0;
    for (int iter = 0; iter < MAX_ITERS; ++iter) {
        ++iters;
        R F1, F2, along, J00, J01, J10, J11;
        newton_eval<R, CN>(T, L, a, b, F1, F2, along, J00, J01, J10, J11);
        R err2 = F1 * F1 + F2 * F2;
        err_streak = err2 > prev_err2 ? err_streak + 1 : 0;
        behind_streak = along < (R)0 ? behind_streak + 1 : 0;
        if (err_streak >= 2 || behind_streak >= 2)
            return false;
        prev_err2 = err2;
        if (err2 < L.tol2) {
            if (along < (R)0 || !domain_inside<R>(T, L.slo + L.w * a, L.tlo + L.w * b))
                return false;
            R t_hit = along / L.d_len;
            if (t_hit <= t_min || t_hit >= t_max)
                return false;
            out_a = a;
            out_b = b;
            out_t = t_hit;
            return true;
        }
        R det = J00 * J11 - J01 * J10;
        if (det == (R)0)
            return false;
        a -= (J11 * F1 - J01 * F2) / det;
        b -= (-J10 * F1 + J00 * F2) / det;
        bool outside = a < (R)0 || a > (R)1 || b < (R)0 || b > (R)1 ||
                       !domain_inside<R>(T, L.slo + L.w * a, L.tlo + L.w * b);
        if (outside) {
            if (++invalid_streak >= 3)
                return false;
            a = fmin(fmax(a, (R)0), (R)1);
            b = fmin(fmax(b, (R)0), (R)1);
        } else {
            invalid_streak = 0;
        }
    }
    return false;
}

template <class R> struct StackSlot;
template <> struct StackSlot<float>
{
    using T = float;
};
template <> struct StackSlot<double>
{
    using T = double;
};

struct LaneStats
{
    int tri, node, leaf, sub, runs, iters;
    // Tests after the last hit, summed over the walks that hit; and those
    // since the last hit of the current triangle's walk, which hit if found.
    int node_after, leaf_after, node_since, leaf_since;
    bool found;
    // Ends the current triangle's walk (displaced_intersector.cpp:232-235).
    __device__ void end_walk()
    {
        if (found) {
            node_after += node_since;
            leaf_after += leaf_since;
        }
        node_since = leaf_since = 0;
        found = false;
    }
};

// search_leaf (displaced_intersector.cpp:292-323) as a depth-first search
// over an explicit stack; entries are (depth, qa, qb) packed in 16 bits and
// the sub-square's lower bound on t. Children are pushed farthest first,
// so they pop nearest first, as the recursion visits them.
template <class R, bool CN>
__device__ bool search_leaf(const TriRecord &T, const Leaf<R> &L, R t_min, R &t_max, R &hit_a, R &hit_b,
                            typename StackSlot<R>::T *s_t, uint16_t *s_code, LaneStats &st)
{
    const int lane = threadIdx.x, stride = blockDim.x;
    int sp = 0;
    bool found = false;
    uint32_t code = 0; // depth 0, the whole texel
    // A texel inside the triangle's domain with rounding room to spare
    // skips its sub-squares' domain tests (each would be false). Only in
    // fp32: the fp64 kernel runs at 2 waves with spills already, and the
    // flag's register made its leaf search 24% slower.
    const bool inside = sizeof(R) == sizeof(float) &&
                        square_inside<R>(T, L.slo + (R)0.5 * L.w, L.tlo + (R)0.5 * L.w, (R)0.5 * L.w);
    while (true) {
        int depth = code >> 10, qa = (code >> 5) & 31, qb = code & 31;
        if (depth == SUBDIV_DEPTH) {
            ++st.runs;
            R scale = (R)1 / (R)(1 << SUBDIV_DEPTH);
            R ra, rb, rt;
            if (newton_from<R, CN>(T, L, ((R)qa + (R)0.5) * scale, ((R)qb + (R)0.5) * scale, t_min, t_max, ra, rb, rt,
                               st.iters)) {
                hit_a = ra;
                hit_b = rb;
                t_max = rt;
                found = true;
            }
        } else {
            int cd = depth + 1;
            R inv = ldexp((R)1, -cd); // 1 / 2^cd, exactly
            R rho = (R)0.5 * inv;
            R key[4];
            uint32_t cc[4];
            for (int q = 0; q < 4; ++q) {
                int ca = 2 * qa + (q & 1), cb = 2 * qb + (q >> 1);
                cc[q] = ((uint32_t)cd << 10) | ((uint32_t)ca << 5) | (uint32_t)cb;
                ++st.sub;
                R tl;
                bool keep = sub_square_bound<R, CN>(T, L, ((R)ca + (R)0.5) * inv, ((R)cb + (R)0.5) * inv, rho, t_min,
                                               t_max, tl, inside);
                key[q] = keep ? tl : (R)INFINITY;
            }
            // Sort the four by key (a network that keeps equal keys in q order).
#define DMAP_CSWAP(x, y)                                                                                              \
    if (key[y] < key[x]) {                                                                                            \
        R tk = key[x];                                                                                                \
        key[x] = key[y];                                                                                              \
        key[y] = tk;                                                                                                  \
        uint32_t tc = cc[x];                                                                                          \
        cc[x] = cc[y];                                                                                                \
        cc[y] = tc;                                                                                                   \
    }
            DMAP_CSWAP(0, 1)
            DMAP_CSWAP(2, 3)
            DMAP_CSWAP(0, 2)
            DMAP_CSWAP(1, 3)
            DMAP_CSWAP(1, 2)
#undef DMAP_CSWAP
            for (int q = 3; q >= 0; --q)
                if (key[q] != (R)INFINITY) {
                    s_t[sp * stride + lane] = key[q];
                    s_code[sp * stride + lane] = (uint16_t)cc[q];
                    ++sp;
                }
        }
        // Pop the next sub-square that can still hold a nearer root.
        bool have = false;
        while (sp > 0) {
            --sp;
            R tl = s_t[sp * stride + lane];
            if (tl < t_max) {
                code = s_code[sp * stride + lane];
                have = true;
                break;
            }
        }
        if (!have)
            break;
    }
    return found;
}

// The two-triangle leaf (displaced_intersector.cpp:390-426, Moller-Trumbore
// :30-49) with corners relative to the ray origin.
template <class R>
__device__ bool leaf_two_triangle(const TraceParams &p, const TriRecord &T, const Leaf<R> &L, const double o[3],
                                  const double d[3], R t_min, R &t_max, R &hit_a, R &hit_b)
{
    // Corner positions S - o; rel = P0 - o is c = (c1, c2, cn) in the
    // (d1, d2, dn) basis, rebuilt in object space.
    R rel[3];
    for (int q = 0; q < 3; ++q)
        rel[q] = L.c1 * L.d1[q] + L.c2 * L.d2[q] + L.cn * L.dn[q];
    R hc[4] = {L.h00, L.h00 + L.ha, L.h00 + L.hb, L.h00 + L.ha + L.hb + L.tw};
    const R ca[4] = {0, 1, 0, 1}, cb[4] = {0, 0, 1, 1};
    R S[4][3];
    for (int c = 0; c < 4; ++c) {
        R N[3], Na[3], Nb[3];
        if (T.const_normal) {
            for (int q = 0; q < 3; ++q)
                N[q] = L.N[q];
        } else {
            normal_at(L, ca[c], cb[c], N, Na, Nb);
        }
        for (int q = 0; q < 3; ++q)
            S[c][q] = rel[q] + ca[c] * L.E1[q] + cb[c] * L.E2[q] + hc[c] * N[q];
    }
    R dd[3] = {(R)d[0], (R)d[1], (R)d[2]};
    // Faces (00, 10, 01) and (10, 11, 01) (tessellation_grid's split).
    const int A[2] = {0, 1}, B[2] = {1, 3}, C[2] = {2, 2};
    bool found = false;
    for (int f = 0; f < 2; ++f) {
        R e1[3], e2[3], pv[3], s[3], qv[3];
        for (int q = 0; q < 3; ++q) {
            e1[q] = S[B[f]][q] - S[A[f]][q];
            e2[q] = S[C[f]][q] - S[A[f]][q];
            s[q] = -S[A[f]][q];
        }
        pv[0] = dd[1] * e2[2] - dd[2] * e2[1];
        pv[1] = dd[2] * e2[0] - dd[0] * e2[2];
        pv[2] = dd[0] * e2[1] - dd[1] * e2[0];
        R det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
        if (det == (R)0)
            continue;
        R inv = (R)1 / det;
        R b1 = (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]) * inv;
        if (b1 < (R)0 || b1 > (R)1)
            continue;
        qv[0] = s[1] * e1[2] - s[2] * e1[1];
        qv[1] = s[2] * e1[0] - s[0] * e1[2];
        qv[2] = s[0] * e1[1] - s[1] * e1[0];
        R b2 = (dd[0] * qv[0] + dd[1] * qv[1] + dd[2] * qv[2]) * inv;
        if (b2 < (R)0 || b1 + b2 > (R)1)
            continue;
        R t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
        if (!(t > t_min && t < t_max))
            continue;
        R la = (1 - b1 - b2) * ca[A[f]] + b1 * ca[B[f]] + b2 * ca[C[f]];
        R lb = (1 - b1 - b2) * cb[A[f]] + b1 * cb[B[f]] + b2 * cb[C[f]];
        if (!domain_inside<R>(T, L.slo + L.w * la, L.tlo + L.w * lb))
            continue;
        hit_a = la;
        hit_b = lb;
        t_max = t;
        found = true;
    }
    return found;
}

// ------------------------------------------------------------------ kernel

// A triangle's box (lo, hi), clipped against [t_min, t_max].
template <class R>
__device__ inline bool ray_box(const R o[3], const R d[3], const R inv[3], const float lo[3], const float hi[3], R t_min,
                               R t_max)
{
    R t0 = t_min, t1 = t_max;
    return clip_axis<R>(o[0], d[0], inv[0], (R)lo[0], (R)hi[0], t0, t1) &&
           clip_axis<R>(o[1], d[1], inv[1], (R)lo[1], (R)hi[1], t0, t1) &&
           clip_axis<R>(o[2], d[2], inv[2], (R)lo[2], (R)hi[2], t0, t1) && close_interval<R>(t0, t1);
}

// A BVH box, stored as centre +- radius, clipped against [t_min, t_max].
template <class R>
__device__ inline bool ray_box_cr(const R o[3], const R d[3], const R inv[3], const float c[3], const float r[3],
                                  R t_min, R t_max)
{
    R t0 = t_min, t1 = t_max;
    return clip_centre<R>(o[0], d[0], inv[0], (R)c[0], (R)r[0], t0, t1) &&
           clip_centre<R>(o[1], d[1], inv[1], (R)c[1], (R)r[1], t0, t1) &&
           clip_centre<R>(o[2], d[2], inv[2], (R)c[2], (R)r[2], t0, t1) && close_interval<R>(t0, t1);
}

template <class R> __device__ inline R axis_of(const R v[3], int a) { return a == 0 ? v[0] : a == 1 ? v[1] : v[2]; }

// Traversal state of the heap-ordered BVH: `cur` is 0 before the first
// call, a leaf after a call that returned its triangle, BVH_DONE after the
// last; bit L of `trail` marks that the far child of the current path's
// node at depth L is still to be visited.
constexpr uint32_t BVH_DONE = 0xFFFFFFFFu;

__device__ __forceinline__ int depth_of(uint32_t node) { return 31 - __clz(node); }

// Advances to the next triangle whose box the ray meets within [t_min,
// t_max] and returns it, or -1 when the tree is exhausted: depth first, the
// near child (by the ray's direction along the node's split axis) first, so
// that a shrinking t_max culls the far boxes, which are tested again when
// the walk comes back to them. The closest hit does not depend on the order
// triangles are visited in (ObjectIntersector::intersect_flat keeps the
// nearest, displaced_intersector.cpp:578-599).
template <class R>
__device__ __forceinline__ int bvh_next(const BvhNode *N, const int *leaf, uint32_t P, uint32_t &cur,
                                        uint32_t &trail, const R o[3], const R d[3], const R inv[3], R t_min, R t_max,
                                        bool narrowed)
{
    if (cur == BVH_DONE)
        return -1;
    bool pop = true;
    if (cur == 0) {
        if (!ray_box_cr<R>(o, d, inv, N[0].ctr[0], N[0].rad[0], t_min, t_max)) {
            cur = BVH_DONE;
            return -1;
        }
        cur = 1;
        trail = 0;
        pop = false;
    }
    while (true) {
        if (pop) {
            if (trail == 0) {
                cur = BVH_DONE;
                return -1;
            }
            int L = depth_of(trail);
            trail &= ~(1u << L);
            int dc = depth_of(cur);
            uint32_t anc = cur >> (dc - L);
            uint32_t far = (cur >> (dc - L - 1)) ^ 1u;
            cur = far;
            int c = (int)(far & 1u);
            // The far child passed its box test when it was pushed; only a
            // t_max narrowed since (by a hit) can fail it now.
            if (narrowed && !ray_box_cr<R>(o, d, inv, N[anc].ctr[c], N[anc].rad[c], t_min, t_max))
                continue;
            pop = false;
        }
        if (cur >= P) {
            int tri = leaf[cur - P];
            if (tri >= 0)
                return tri;
            pop = true;
            continue;
        }
        const BvhNode &nd = N[cur];
        bool h0 = !(nd.empty & 1) && ray_box_cr<R>(o, d, inv, nd.ctr[0], nd.rad[0], t_min, t_max);
        bool h1 = !(nd.empty & 2) && ray_box_cr<R>(o, d, inv, nd.ctr[1], nd.rad[1], t_min, t_max);
        if (h0 && h1) {
            trail |= 1u << depth_of(cur);
            cur = 2 * cur + (axis_of(d, nd.axis) >= (R)0 ? 0u : 1u);
        } else if (h0 || h1) {
            cur = 2 * cur + (h1 ? 1u : 0u);
        } else {
            pop = true;
        }
    }
}

// The four children of cell (pi, pj) at `level` + 1 share one Morton group
// at `level`: their min-max pairs are one 32 B load, their Taylor quads one
// 64 B load.
struct Group
{
    float4 mm01, mm23;
    float4 tp[4];
};

__device__ inline void load_group(const PyramidView &P, int level, int pi, int pj, bool want_tp, Group &g)
{
    if (level >= P.n_levels) {
        float lo = P.root_hmin, hi = P.root_hmax;
        g.mm01 = make_float4(lo, hi, lo, hi);
        g.mm23 = g.mm01;
        float4 t = make_float4(0.5f * (lo + hi), 0.f, 0.f, 0.5f * (hi - lo));
        g.tp[0] = g.tp[1] = g.tp[2] = g.tp[3] = t;
        return;
    }
    uint32_t m = (uint32_t)P.n_leaf >> level;
    uint32_t base = level_base(m) + morton2((uint32_t)(2 * pi) & (m - 1), (uint32_t)(2 * pj) & (m - 1));
    const float4 *mm4 = reinterpret_cast<const float4 *>(P.mm + base);
    g.mm01 = mm4[0];
    g.mm23 = mm4[1];
    if (want_tp) {
        const float4 *tp4 = P.tp + base;
        g.tp[0] = tp4[0];
        g.tp[1] = tp4[1];
        g.tp[2] = tp4[2];
        g.tp[3] = tp4[3];
    }
}

__device__ inline float2 group_mm(const Group &g, int q)
{
    float4 v = q < 2 ? g.mm01 : g.mm23;
    return (q & 1) ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
}
__device__ inline float4 group_tp(const Group &g, int q)
{
    float4 a = (q & 1) ? g.tp[1] : g.tp[0], b = (q & 1) ? g.tp[3] : g.tp[2];
    return q < 2 ? a : b;
}

// Node tests of a 2 x 2 block of cells at `level` with lowest corner
// (bi, bj), skipping cells outside the domain (TFDM's texel discard, before
// any node is used) and cells not in `valid` (Morton bit q = io + 2 jo).
// An aligned block is the four children of one cell, one Morton group; the
// roots of a triangle need not be aligned and are loaded one by one.
// Returns the passing cells as a 4-bit mask in the order the ray meets
// them: bit k is cell (bi + ((k >> 1) ^ si), bj + ((k & 1) ^ sj)), the
// child order of cell_next (displaced_intersector.cpp:99-121).
// REC: the launch counts work (STATS); then, when p.stats.records is set,
// each node test is recorded for ray rid, triangle prim, its place in the
// ray's node tests counted from seq0.
template <class R, bool CN, bool REC>
__device__ uint32_t expand(const TraceParams &p, const PyramidView &P, const TriRecord &T, const Walk<R> &wk, int level,
                           int bi, int bj, bool aligned, uint32_t valid, bool &inside, R t_min, R t_max, int &node_tests,
                           int64_t rid, int prim, int seq0)
{
    bool slab = p.bound != 0 && level <= p.slab_max_level;
    Group g;
    if (aligned) {
        load_group(P, level, bi >> 1, bj >> 1, slab, g);
    } else {
        float2 m0 = fetch_mm(P, level, bi, bj), m1 = fetch_mm(P, level, bi + 1, bj);
        float2 m2 = fetch_mm(P, level, bi, bj + 1), m3 = fetch_mm(P, level, bi + 1, bj + 1);
        g.mm01 = make_float4(m0.x, m0.y, m1.x, m1.y);
        g.mm23 = make_float4(m2.x, m2.y, m3.x, m3.y);
        if (slab) {
            g.tp[0] = fetch_tp(P, level, bi, bj);
            g.tp[1] = fetch_tp(P, level, bi + 1, bj);
            g.tp[2] = fetch_tp(P, level, bi, bj + 1);
            g.tp[3] = fetch_tp(P, level, bi + 1, bj + 1);
        }
    }
    R w = ldexp((R)1, level - (__ffs(P.n_leaf) - 1)); // 2^level / n_leaf, n_leaf a power of two
    // `inside` comes in set when an ancestor block was inside (then so is
    // every block below it); otherwise it is tested here.
    if (!inside)
        inside = square_inside<R>(T, ((R)bi + (R)1) * w, ((R)bj + (R)1) * w, w);
    // The block's two columns and two rows, clipped once for its four cells.
    const R blo_x = sizeof(R) == 8 ? (R)T.dlo_d[0] : (R)T.dlo[0], bhi_x = sizeof(R) == 8 ? (R)T.dhi_d[0] : (R)T.dhi[0];
    const R blo_y = sizeof(R) == 8 ? (R)T.dlo_d[1] : (R)T.dlo[1], bhi_y = sizeof(R) == 8 ? (R)T.dhi_d[1] : (R)T.dhi[1];
    R sx_c[2], sx_h[2], sy_c[2], sy_h[2];
    for (int k = 0; k < 2; ++k) {
        cell_span<R>(bi + k, w, blo_x, bhi_x, sx_c[k], sx_h[k]);
        cell_span<R>(bj + k, w, blo_y, bhi_y, sy_c[k], sy_h[k]);
    }
    uint32_t mask = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int io = q & 1, jo = q >> 1;
        int ci = bi + io, cj = bj + jo;
        if (!((valid >> q) & 1u))
            continue;
        if (!inside && square_outside<R>(T, ((R)ci + (R)0.5) * w, ((R)cj + (R)0.5) * w, (R)0.5 * w))
            continue;
        if (REC && p.stats.records)
            record_node_test<R, CN>(p, P, T, wk, level, ci, cj, sx_c[io], sy_c[jo], sx_h[io], sy_h[jo],
                                    group_mm(g, q), t_min, t_max, rid, prim, seq0 + node_tests);
        ++node_tests;
        float4 tp = slab ? g.tp[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (cell_test<R, CN>(p, P, T, wk, level, ci, cj, sx_c[io], sy_c[jo], sx_h[io], sy_h[jo], group_mm(g, q), tp,
                             t_min, t_max))
            mask |= 1u << (((io ^ wk.si) << 1) | (jo ^ wk.sj));
    }
    return mask;
}

// Relative room of start_block's bounds, in units of the ray's magnitudes.
constexpr double START_ROOM = 1e-5;

// The finest 2 x 2 block of cells below a constant-normal triangle's root
// block that holds every leaf where the walk can accept a hit of the ray
// (tangent-space wk) within [t_min, t_max]. The surface over the footprint
// lies in the shell {(u, v, 0) + h m}: (u, v) in the domain, h in the
// footprint's height range [h_lo, h_hi], m the normal's tangent form
// (centre mc, remainder mk per axis). A hit's z is h m_z, which clips the
// ray to [t0, t1], and its (u, v) is the ray's (x, y) there less
// h (m_x, m_y). A leaf accepts a surface point within its residual
// tolerance of the ray (KNEWTON ulps of the residual's magnitude), so every
// bound is widened by E, START_ROOM times the ray's magnitudes: above that
// tolerance and the rounding of the few operations here (each within an
// ulp of the same magnitudes) by two orders or more.
// Returns 0 when the ray cannot meet the shell, 1 with the block set (its
// lowest corner (bi, bj) at `level`, `valid` its cells that meet the
// bounds, bit io + 2 jo), 2 when no block below the root block holds the
// bounds (or they are unbounded): the walk starts at the roots.
// The walk from the block meets every leaf where the walk from the roots
// accepts a hit: a cell outside the bounds holds none, and a cell test only
// prunes.
template <class R>
__device__ inline int start_block(const TriRecord &T, int n_leaf, const Walk<R> &wk, R t_min, R t_max, int &level,
                                  int &bi, int &bj, uint32_t &valid)
{
    // (nmax and nmin pass NaNs on: every value below is kept finite.)
    if (!isfinite(wk.inv[2]))
        return 2;
    const R room = (R)START_ROOM;
    const R hl = (R)T.h_lo, hh = (R)T.h_hi, ah = nmax(fabs(hl), fabs(hh));
    R lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
        const R mc = sizeof(R) == 8 ? (R)T.mc_d[q] : (R)T.mc[q], mk = sizeof(R) == 8 ? (R)T.mk_d[q] : (R)T.mk[q];
        const R a = hl * mc, b = hh * mc;
        lo[q] = nmin(a, b) - ah * mk;
        hi[q] = nmax(a, b) + ah * mk;
    }
    const R inv = wk.inv[2];
    R ta = (lo[2] - wk.o[2]) * inv, tb = (hi[2] - wk.o[2]) * inv;
    if (ta > tb) {
        R x = ta;
        ta = tb;
        tb = x;
    }
    const R zmag = fabs(lo[2]) + fabs(hi[2]);
    const R dt = room * (wk.omag + wk.dmag * nmax(fabs(ta), fabs(tb)) + zmag) * fabs(inv);
    const R t0 = nmax(t_min, ta - dt), t1 = nmin(t_max, tb + dt);
    if (!(t0 <= t1))
        return 0;
    if (!(isfinite(t0) && isfinite(t1)))
        return 2;
    const R E = room * (wk.omag + wk.dmag * nmax(fabs(t0), fabs(t1)) + zmag);
    // The root block's cells that meet the domain's bounding box.
    const R W = ldexp((R)1, T.root_level) / (R)n_leaf;
    int ri0 = T.root_i[0], ri1 = T.root_i[0], rj0 = T.root_j[0], rj1 = T.root_j[0];
    for (int r = 1; r < T.n_roots; ++r) {
        ri0 = min(ri0, T.root_i[r]);
        ri1 = max(ri1, T.root_i[r]);
        rj0 = min(rj0, T.root_j[r]);
        rj1 = max(rj1, T.root_j[r]);
    }
    const R xa = wk.o[0] + t0 * wk.d[0], xb = wk.o[0] + t1 * wk.d[0];
    const R ya = wk.o[1] + t0 * wk.d[1], yb = wk.o[1] + t1 * wk.d[1];
    const R blo_x = sizeof(R) == 8 ? (R)T.dlo_d[0] : (R)T.dlo[0], bhi_x = sizeof(R) == 8 ? (R)T.dhi_d[0] : (R)T.dhi[0];
    const R blo_y = sizeof(R) == 8 ? (R)T.dlo_d[1] : (R)T.dlo[1], bhi_y = sizeof(R) == 8 ? (R)T.dhi_d[1] : (R)T.dhi[1];
    const R u0 = nmax(nmax(nmin(xa, xb) - hi[0], blo_x) - E, (R)ri0 * W);
    const R u1 = nmin(nmin(nmax(xa, xb) - lo[0], bhi_x) + E, (R)(ri1 + 1) * W);
    const R v0 = nmax(nmax(nmin(ya, yb) - hi[1], blo_y) - E, (R)rj0 * W);
    const R v1 = nmin(nmin(nmax(ya, yb) - lo[1], bhi_y) + E, (R)(rj1 + 1) * W);
    if (!(u0 <= u1 && v0 <= v1))
        return 0;
    // An interval no longer than a cell crosses at most one cell boundary:
    // the level whose cells are at least the longer side fits, and the one
    // below it may.
    const R span = nmax(u1 - u0, v1 - v0) * (R)n_leaf; // in leaf texels
    int L = span < (R)1 ? 0 : max(0, ilogb(span));
    for (; L < T.root_level; ++L) {
        const R s = ldexp((R)n_leaf, -L); // cells per unit
        const R i0 = floor(u0 * s), i1 = nmax(i0, ceil(u1 * s) - (R)1);
        const R j0 = floor(v0 * s), j1 = nmax(j0, ceil(v1 * s) - (R)1);
        if (i1 - i0 <= (R)1 && j1 - j0 <= (R)1) {
            level = L;
            bi = (int)i0;
            bj = (int)j0;
            const bool wi = i1 > i0, wj = j1 > j0;
            valid = 1u | (wi ? 2u : 0u) | (wj ? 4u : 0u) | (wi && wj ? 8u : 0u);
            return 1;
        }
    }
    return 2;
}

enum
{
    ST_NEED = 0,   // bookkeeping: find the next block to expand or leaf to test
    ST_EXPAND = 1, // a block's node tests are pending
    ST_LEAF = 2,   // a leaf test is pending
    ST_DONE = 3,
    ST_SETUP = 4   // the set-up of the next triangle's walk is pending
};

// Leaves run when at least this share (in 1/64) of the wave's busy lanes
// have one pending, or when no lane has a block to expand.
constexpr int LEAF_BATCH = 56;
// A wave claims new rays when at least this many of its lanes are idle
// (both thresholds from a sweep on the incoherent 4M-ray workload: fewer,
// fuller refills and leaf batches keep more lanes active per instruction).
constexpr int REFILL = 48;
// The same two thresholds for multi-request (serving) launches, whose rays
// are not sorted.
constexpr int MULTI_LEAF_BATCH = 56;
constexpr int MULTI_REFILL = 48;
// Triangle set-ups (the ray in tangent space, the start block) run when at
// least this share (in 1/64) of the wave's busy lanes have one pending, or
// when no lane has a block to expand. Run as each lane's BVH step found its
// triangle, a set-up took a few lanes at a time through its dependent record
// reads and some 200 instructions: 21% of a wave's cycles on the incoherent
// 4M-ray workload. Batched, the trace took 2.45 ms against 2.55 ms (8 to 20
// gave the same within 0.5%, 24 was slower).
constexpr int SETUP_BATCH = 16;

// Closest or any hit, one ray per lane. The walk of each triangle is the
// reference's depth-first descent from its roots (DisplacedIntersector::
// intersect, displaced_intersector.cpp:175-237) with the four children of a
// passing cell tested together: `trail` keeps, four bits per level, the
// cells still to visit, in the ray's order. Each round of the outer loop
// first advances every lane's bookkeeping (pops, climbs, the next triangle
// from the BVH) to its next block of node tests, then runs those node
// tests for the whole wave at once; pending triangle set-ups and pending
// leaves are each run together once enough lanes have one.
// Waves per SIMD the trace kernel is compiled for: 3 for fp32 objects with
// constant normals (168 VGPRs; 2 and 4 measured slower); 2 for the
// general-normal fp32 kernel (it spills 136 B per lane at 3, and runs 4%
// faster at 2) and for fp64, whose doubled state would spill 400-1000 B per
// lane at 3 and runs 1.9x faster at 2.
template <class R, bool CN> constexpr int trace_waves() { return (sizeof(R) == sizeof(double) || !CN) ? 2 : 3; }

// The walk of k_trace (single-object launches, sd NULL) and of
// k_trace_multi (a serving step, sd its descriptor in the kernel's
// arguments).
template <class R, bool STATS, bool CN, bool TWO, bool MULTI>
__device__ __forceinline__ void trace_body(const TraceParams &p, const StepDesc *sd)
{

    // Multi-request launches: the workgroup's object group (uniform across
    // the workgroup); workgroups past the groups' shares have no work.
    int grp = 0;
    if (MULTI) {
        grp = -1;
        for (int g = 0; g < sd->n_grp; ++g)
            if ((int)blockIdx.x >= sd->grp[g].wg_lo && (int)blockIdx.x < sd->grp[g].wg_hi)
                grp = g;
        if (grp < 0)
            return;
    }
    // The object: the launch's, or the workgroup's group's in a serving
    // step, in the kernel arguments either way.
    const ObjectDev &OBJ = MULTI ? sd->grp[grp].obj : p.obj;

    // The leaf search stacks: STACK entries per thread, stride blockDim.x
    // (trace_lds_bytes per launch).
    extern __shared__ __align__(16) unsigned char s_lds[];
    typename StackSlot<R>::T *s_t = (typename StackSlot<R>::T *)s_lds;
    uint16_t *s_code = (uint16_t *)(s_t + STACK * blockDim.x);
    // A small BVH is copied into LDS after the stacks: the walk's dependent
    // node reads are the bulk of the per-triangle bookkeeping.
    const bool lds_bvh = OBJ.bvh_leaves <= LDS_BVH_LEAVES;
    unsigned char *bvh_lds = (unsigned char *)(s_code + STACK * blockDim.x);
    // Multi-request launches: the step's segments, copied from the kernel
    // arguments into LDS (after the BVH's room) for the refill's binary
    // search and the lanes' request records.
    Segment *s_seg = (Segment *)(bvh_lds + BVH_LDS_BYTES);
    if (MULTI) {
        const uint32_t *src = (const uint32_t *)sd->seg;
        uint32_t *dst = (uint32_t *)s_seg;
        const int words = sd->n_seg * (int)(sizeof(Segment) / 4);
        for (int q = threadIdx.x; q < words; q += blockDim.x)
            dst[q] = src[q];
        __syncthreads();
    }
    if (lds_bvh) {
        uint32_t *dst = (uint32_t *)bvh_lds;
        const uint32_t *src = (const uint32_t *)OBJ.bvh;
        const int node_words = OBJ.bvh_leaves * (int)(sizeof(BvhNode) / 4);
        for (int q = threadIdx.x; q < node_words; q += blockDim.x)
            dst[q] = src[q];
        const uint32_t *lsrc = (const uint32_t *)OBJ.bvh_leaf;
        uint32_t *ldst = (uint32_t *)(bvh_lds + sizeof(BvhNode) * LDS_BVH_LEAVES);
        for (int q = threadIdx.x; q < OBJ.bvh_leaves; q += blockDim.x)
            ldst[q] = lsrc[q];
        __syncthreads();
    }

    const int lane = threadIdx.x & 63;
    // Persistent waves: a lane that has finished its ray writes the result
    // and, once enough lanes of its wave are idle, the wave claims that many
    // new rays from the launch's queue (its group's, in a serving step); the
    // wave exits when the queue is drained and its lanes are done.
    bool drained = false, have = false;
    int64_t rid = 0;
    // Multi-request launches: the lane's request.
    const Segment *SG = nullptr;
    const int64_t n_total = MULTI ? sd->grp[grp].end - sd->grp[grp].start : p.n_rays;
    const int64_t slot0 = MULTI ? sd->grp[grp].start : 0;
    unsigned *const queue = MULTI ? p.queue + grp : p.queue;
#define DMAP_OBJ (OBJ)
#define DMAP_RAYS (MULTI ? segment_rays(*SG) : p.rays)
#define DMAP_HITS (MULTI ? segment_hits(*SG) : p.hits)
    R t_min = 0, t_best = 0;
    // The ray as given (fp32, so exact in R and in double).
    R o_f[3] = {0, 0, 0}, d_f[3] = {0, 0, 1}, inv_f[3] = {0, 0, 0};
    int state = ST_DONE;
    bool walking = false;
    uint32_t bvh_cur = 0, bvh_trail = 0;
    bool single_done = false;
    int k = -1;
    int hit_tri = -1, hit_i = 0, hit_j = 0;
    R hit_a = 0, hit_b = 0;
    LaneStats st{};
    Walk<R> wk;
    const TriRecord *T = OBJ.tris;
    int root_level = 0, rb_i = 0, rb_j = 0;
    int lvl = 0, pi = 0, pj = 0; // the group being walked: the cells at lvl below (pi, pj), or the root block
    uint64_t trail = 0;
    // The level of the block on the current path found inside the
    // triangle's domain (its cells and all below them are inside), or -1.
    int in_lvl = -1;
    int ex_level = 0, ex_i = 0, ex_j = 0; // the block to expand
    bool ex_aligned = true;
    uint32_t ex_valid = 15;
    int leaf_i = 0, leaf_j = 0;

    while (true) {
        if (have && state == ST_DONE) {
            // The result of the finished ray.
            if (p.any_hit) {
                p.occluded[rid] = hit_tri >= 0 ? 1 : 0;
            } else if (!MULTI && p.raw) {
                // The raw record, one whole sector; k_finalize turns it into the hit.
                float4 *dst = (float4 *)(p.raw + rid);
                dst[0] = make_float4(hit_tri >= 0 ? (float)t_best : INFINITY, __int_as_float(hit_tri),
                                     __int_as_float(hit_i), __int_as_float(hit_j));
                dst[1] = make_float4((float)hit_a, (float)hit_b, 0.f, 0.f);
            } else {
                // The raw record in the hit arrays; k_finalize turns it into the hit.
                const HitsDev &H = DMAP_HITS;
                H.t[rid] = hit_tri >= 0 ? (float)t_best : INFINITY;
                H.prim[rid] = hit_tri;
                H.texel_i[rid] = hit_i;
                H.texel_j[rid] = hit_j;
                H.s[rid] = (float)hit_a;
                H.tt[rid] = (float)hit_b;
            }
            if (STATS) {
                const StatsDev &S = p.stats;
                st.end_walk();
                if (S.node_tests_after_hit)
                    S.node_tests_after_hit[rid] = st.node_after;
                if (S.leaf_tests_after_hit)
                    S.leaf_tests_after_hit[rid] = st.leaf_after;
                if (S.triangle_tests)
                    S.triangle_tests[rid] = st.tri;
                if (S.node_tests)
                    S.node_tests[rid] = st.node;
                if (S.leaf_tests)
                    S.leaf_tests[rid] = st.leaf;
                if (S.sub_squares)
                    S.sub_squares[rid] = st.sub;
                if (S.newton_runs)
                    S.newton_runs[rid] = st.runs;
                if (S.newton_iters)
                    S.newton_iters[rid] = st.iters;
            }
            have = false;
        }
        if (!drained) {
            uint64_t idle = __ballot(!have);
            if (__popcll(idle) >= (MULTI ? MULTI_REFILL : REFILL) || idle == ~0ull) {
                int cnt = __popcll(idle);
                int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned base = 0;
                if (lane == leader)
                    base = atomicAdd(queue, (unsigned)cnt);
                base = __shfl(base, leader);
                if ((int64_t)base + cnt >= n_total)
                    drained = true;
                if (!have) {
                    int64_t slot = (int64_t)base + __popcll(idle & ((1ull << lane) - 1));
                    if (slot < n_total) {
                        have = true;
                        if (MULTI) {
                            // The request holding this step slot: a binary
                            // search of the group's segments' starts.
                            slot += slot0;
                            int lo = sd->grp[grp].seg_lo, hi = sd->grp[grp].seg_hi - 1;
                            while (lo < hi) {
                                int mid = (lo + hi + 1) >> 1;
                                if (s_seg[mid].start <= slot)
                                    lo = mid;
                                else
                                    hi = mid - 1;
                            }
                            SG = &s_seg[lo];
                            rid = slot - SG->start + SG->first;
                        } else {
                            rid = p.perm ? (int64_t)p.perm[slot] : slot;
                        }
                        const RaysDev &RY = DMAP_RAYS;
                        o_f[0] = RY.ox[rid];
                        o_f[1] = RY.oy[rid];
                        o_f[2] = RY.oz[rid];
                        d_f[0] = RY.dx[rid];
                        d_f[1] = RY.dy[rid];
                        d_f[2] = RY.dz[rid];
                        t_min = (R)RY.tmin[rid];
                        t_best = (R)RY.tmax[rid];
                        for (int q = 0; q < 3; ++q)
                            inv_f[q] = padded_div<R>((R)1, d_f[q]);
                        // A ray with a non-finite or zero component, or an empty
                        // (t_min, t_max), has no hit: the reference's walk finds none.
                        bool ok = isfinite(o_f[0]) && isfinite(o_f[1]) && isfinite(o_f[2]) && isfinite(d_f[0]) &&
                                  isfinite(d_f[1]) && isfinite(d_f[2]) &&
                                  (d_f[0] != (R)0 || d_f[1] != (R)0 || d_f[2] != (R)0) && t_min < t_best &&
                                  !isnan(t_best);
                        state = ok ? ST_NEED : ST_DONE;
                        walking = false;
                        bvh_cur = 0;
                        bvh_trail = 0;
                        single_done = false;
                        k = -1;
                        hit_tri = -1;
                        hit_i = hit_j = 0;
                        hit_a = hit_b = 0;
                        st = LaneStats{};
                        if (STATS && p.stats.node_tests_per_level)
                            for (int l = 0; l < STAT_LEVELS; ++l)
                                p.stats.node_tests_per_level[(int64_t)l * p.n_rays + rid] = 0;
                    }
                }
            }
        }
        if (!__any(have))
            break;
        // Bookkeeping, per lane, up to its next block of node tests or leaf.
        while (state == ST_NEED) {
            if (walking) {
                // Climb straight to the nearest level with cells left (no
                // trail bits sit above the root block's level); none left
                // ends this triangle's walk.
                uint64_t rest = trail >> (4 * lvl);
                if (rest == 0) {
                    walking = false;
                    continue;
                }
                int up = (__ffsll((unsigned long long)rest) - 1) >> 2;
                lvl += up;
                if (lvl > in_lvl)
                    in_lvl = -1; // the walk has left the inside block
                pi >>= up;
                pj >>= up;
                uint32_t m = (uint32_t)(trail >> (4 * lvl)) & 15u;
                int kk = __ffs(m) - 1;
                trail &= ~((uint64_t)1 << (4 * lvl + kk));
                int io = (kk >> 1) ^ wk.si, jo = (kk & 1) ^ wk.sj;
                int ci = lvl == root_level ? rb_i + io : 2 * pi + io;
                int cj = lvl == root_level ? rb_j + jo : 2 * pj + jo;
                if (lvl == 0) {
                    leaf_i = ci;
                    leaf_j = cj;
                    state = ST_LEAF;
                    break;
                }
                ex_level = lvl - 1;
                ex_i = 2 * ci;
                ex_j = 2 * cj;
                ex_aligned = true;
                ex_valid = 15;
                state = ST_EXPAND;
                break;
            }
            // The next triangle whose box the ray meets (intersect_flat,
            // displaced_intersector.cpp:578-599, through the BVH).
            if (p.prim_mask >= 0) {
                k = -1;
                if (!single_done) {
                    single_done = true;
                    const float *bx = DMAP_OBJ.boxes + 6 * p.prim_mask;
                    if (ray_box<R>(o_f, d_f, inv_f, bx, bx + 3, t_min, t_best))
                        k = p.prim_mask;
                }
            } else {
                if (lds_bvh)
                    k = bvh_next<R>((const BvhNode *)bvh_lds, (const int *)(bvh_lds + sizeof(BvhNode) * LDS_BVH_LEAVES),
                                    (uint32_t)OBJ.bvh_leaves, bvh_cur, bvh_trail, o_f, d_f, inv_f, t_min, t_best,
                                    hit_tri >= 0);
                else
                    k = bvh_next<R>(DMAP_OBJ.bvh, DMAP_OBJ.bvh_leaf, (uint32_t)DMAP_OBJ.bvh_leaves, bvh_cur,
                                    bvh_trail, o_f, d_f, inv_f, t_min, t_best, hit_tri >= 0);
            }
            if (k < 0) {
                state = ST_DONE;
                break;
            }
            if (STATS) {
                st.end_walk();
                ++st.tri;
            }
            state = ST_SETUP;
        }
        // The set-up of each pending triangle's walk, in a batch (the
        // ballots taken by the whole wave).
        const uint64_t setup_lanes = __ballot(state == ST_SETUP);
        const uint64_t expand_lanes = __ballot(state == ST_EXPAND);
        const uint64_t ready_lanes = __ballot(state != ST_DONE);
        if (state == ST_SETUP &&
            (expand_lanes == 0 || 64 * __popcll(setup_lanes) >= SETUP_BATCH * __popcll(ready_lanes))) {
            T = DMAP_OBJ.tris + k;
            // The ray in tangent space (DisplacedIntersector::intersect :178),
            // formed in double.
            double rel[3] = {(double)o_f[0] - T->q0[0], (double)o_f[1] - T->q0[1], (double)o_f[2] - T->q0[2]};
            R omag = 0, dmag = 0;
            for (int q = 0; q < 3; ++q) {
                const double *row = T->T + 3 * q;
                wk.o[q] = (R)(row[0] * rel[0] + row[1] * rel[1] + row[2] * rel[2]);
                wk.d[q] = (R)(row[0] * (double)d_f[0] + row[1] * (double)d_f[1] + row[2] * (double)d_f[2]);
                wk.inv[q] = padded_div<R>((R)1, wk.d[q]);
                omag = fmax(omag, fabs(wk.o[q]));
                dmag = fmax(dmag, fabs(wk.d[q]));
            }
            wk.omag = omag;
            wk.dmag = dmag;
            wk.si = wk.d[0] < (R)0;
            wk.sj = wk.d[1] < (R)0;
            root_level = T->root_level;
            // The roots: the cells of the footprint's 2 x 2 root block that
            // meet the domain's bounding box (displaced_intersector.cpp:139-149).
            rb_i = T->root_i[0];
            rb_j = T->root_j[0];
            uint32_t rv = 0;
            for (int r = 0; r < T->n_roots; ++r)
                rv |= 1u << ((T->root_i[r] - rb_i) + 2 * (T->root_j[r] - rb_j));
            int how = 2;
            if (T->const_normal && p.walk_start == 0) {
                // Under a constant normal the walk starts at the finest block
                // that holds the ray's possible hits.
                int sl = 0, sbi = 0, sbj = 0;
                uint32_t sv = 0;
                how = start_block<R>(*T, DMAP_OBJ.pyr.n_leaf, wk, t_min, t_best, sl, sbi, sbj, sv);
                if (how == 1) {
                    root_level = sl;
                    rb_i = sbi;
                    rb_j = sbj;
                    rv = sv;
                }
            }
            if (how == 0) {
                state = ST_NEED; // no hit on this triangle: on to the next
            } else {
                ex_level = root_level;
                ex_i = rb_i;
                ex_j = rb_j;
                ex_aligned = false;
                ex_valid = rv;
                in_lvl = -1;
                state = ST_EXPAND;
            }
        }
        if (state == ST_EXPAND) {
            int tests = 0;
            bool inside = ex_level < in_lvl;
            uint32_t mask = expand<R, CN, STATS>(p, DMAP_OBJ.pyr, *T, wk, ex_level, ex_i, ex_j, ex_aligned, ex_valid,
                                                 inside, t_min, t_best, tests, rid, k, STATS ? st.node : 0);
            if (inside && in_lvl < 0)
                in_lvl = ex_level;
            if (STATS) {
                st.node += tests;
                st.node_since += tests;
                if (p.stats.node_tests_per_level)
                    p.stats.node_tests_per_level[(int64_t)ex_level * p.n_rays + rid] += tests;
            }
            if (!ex_aligned) {
                // the root block
                lvl = root_level;
                trail = (uint64_t)mask << (4 * lvl);
                walking = true;
            } else if (mask) {
                lvl = ex_level;
                pi = ex_i >> 1;
                pj = ex_j >> 1;
                trail |= (uint64_t)mask << (4 * lvl);
            }
            state = ST_NEED;
        }
        uint64_t leaf_lanes = __ballot(state == ST_LEAF);
        uint64_t busy_lanes = __ballot(state != ST_DONE);
        if (leaf_lanes && (leaf_lanes == busy_lanes ||
                           64 * __popcll(leaf_lanes) >= (MULTI ? MULTI_LEAF_BATCH : LEAF_BATCH) * __popcll(busy_lanes))) {
            if (state == ST_LEAF) {
                if (STATS) {
                    ++st.leaf;
                    ++st.leaf_since;
                }
                const double o[3] = {(double)o_f[0], (double)o_f[1], (double)o_f[2]};
                const double d[3] = {(double)d_f[0], (double)d_f[1], (double)d_f[2]};
                Leaf<R> L;
                leaf_setup<R>(DMAP_OBJ.pyr, *T, o, d, leaf_i, leaf_j, L);
                R la = 0, lb = 0;
                bool found;
                if (!TWO)
                    found = search_leaf<R, CN>(*T, L, t_min, t_best, la, lb, s_t, s_code, st);
                else
                    found = leaf_two_triangle<R>(p, *T, L, o, d, t_min, t_best, la, lb);
                if (found) {
                    hit_tri = k;
                    hit_i = leaf_i;
                    hit_j = leaf_j;
                    hit_a = la;
                    hit_b = lb;
                    if (STATS) {
                        st.node_since = st.leaf_since = 0;
                        st.found = true;
                    }
                }
                state = (found && p.any_hit) ? ST_DONE : ST_NEED;
            }
        }
    }
}

#undef DMAP_OBJ
#undef DMAP_RAYS
#undef DMAP_HITS

// Closest hit / any hit of one object: persistent waves over the launch's
// ray queue.
template <class R, bool STATS, bool CN, bool TWO>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(trace_waves<R, CN>(), 8)))
k_trace(TraceParams p)
{
    trace_body<R, STATS, CN, TWO, false>(p, nullptr);
}

// One serving step: the requests of `sd`, each against its own object.
template <bool CN>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(trace_waves<float, CN>(), 8)))
k_trace_multi(TraceParams p, StepDesc sd)
{
    trace_body<float, false, CN, false, true>(p, &sd);
}

// The reported hit, in double: the root polished (Newton leaf) or the face
// crossing recomputed (two-triangle leaf), then st, the chart's
// barycentrics and ng = unit (Su x Sv) at the root
// (displaced_intersector.cpp:357-366, 413-421).
__device__ __forceinline__ void finalize_hit(const TraceParams &p, const PyramidView &P, const TriRecord *tris,
                                             const RaysDev &rays, const HitsDev &H, int64_t rid, int tri, float t_raw,
                                             float a_raw, float b_raw, int hi, int hj)
{
    H.prim[rid] = tri;
    if (tri < 0) {
        H.t[rid] = INFINITY;
        H.s[rid] = 0.f;
        H.tt[rid] = 0.f;
        H.texel_i[rid] = 0;
        H.texel_j[rid] = 0;
        if (H.bu)
            H.bu[rid] = 0.f;
        if (H.bv)
            H.bv[rid] = 0.f;
        if (H.nx)
            H.nx[rid] = 0.f;
        if (H.ny)
            H.ny[rid] = 0.f;
        if (H.nz)
            H.nz[rid] = 0.f;
        return;
    }
    H.texel_i[rid] = hi;
    H.texel_j[rid] = hj;
    double o[3] = {rays.ox[rid], rays.oy[rid], rays.oz[rid]};
    double d[3] = {rays.dx[rid], rays.dy[rid], rays.dz[rid]};
    double t_min = rays.tmin[rid];
    const TriRecord &TH = tris[tri];
    Leaf<double> L;
    leaf_setup<double>(P, TH, o, d, hi, hj, L);
    double a = a_raw, b = b_raw, t_hit = t_raw;
    double Sa[3], Sb[3];
    if (p.leaf == 1) {
        double ra, rb, tm = (double)INFINITY;
        if (leaf_two_triangle<double>(p, TH, L, o, d, t_min, tm, ra, rb)) {
            a = ra;
            b = rb;
            t_hit = tm;
        }
        bool upper = a + b > 1.0;
        double hc[4] = {L.h00, L.h00 + L.ha, L.h00 + L.hb, L.h00 + L.ha + L.hb + L.tw};
        double N[3], Na[3], Nb[3], Sc[4][3];
        const double ca[4] = {0, 1, 0, 1}, cb[4] = {0, 0, 1, 1};
        for (int c = 0; c < 4; ++c) {
            if (TH.const_normal) {
                for (int q = 0; q < 3; ++q)
                    N[q] = L.N[q];
            } else {
                normal_at(L, ca[c], cb[c], N, Na, Nb);
            }
            for (int q = 0; q < 3; ++q)
                Sc[c][q] = ca[c] * L.E1[q] + cb[c] * L.E2[q] + hc[c] * N[q];
        }
        int A = upper ? 1 : 0, B = upper ? 3 : 1, C = 2;
        for (int q = 0; q < 3; ++q) {
            Sa[q] = Sc[B][q] - Sc[A][q];
            Sb[q] = Sc[C][q] - Sc[A][q];
        }
    } else {
        if (p.polish) {
            polish_root(TH, L, a, b);
            double F1, F2, along, J00, J01, J10, J11;
            newton_eval<double, false>(TH, L, a, b, F1, F2, along, J00, J01, J10, J11);
            t_hit = along / L.d_len;
        }
        double h = ((L.h00 + L.ha * a) + L.hb * b) + L.tw * a * b;
        double hda = L.ha + L.tw * b, hdb = L.hb + L.tw * a;
        if (TH.const_normal) {
            for (int q = 0; q < 3; ++q) {
                Sa[q] = L.E1[q] + hda * L.N[q];
                Sb[q] = L.E2[q] + hdb * L.N[q];
            }
        } else {
            double N[3], Na[3], Nb[3];
            normal_at(L, a, b, N, Na, Nb);
            for (int q = 0; q < 3; ++q) {
                Sa[q] = L.E1[q] + hda * N[q] + h * Na[q];
                Sb[q] = L.E2[q] + hdb * N[q] + h * Nb[q];
            }
        }
    }
    double s = L.slo + L.w * a, t = L.tlo + L.w * b;
    double n[3] = {Sa[1] * Sb[2] - Sa[2] * Sb[1], Sa[2] * Sb[0] - Sa[0] * Sb[2], Sa[0] * Sb[1] - Sa[1] * Sb[0]};
    double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0.0)
        for (int q = 0; q < 3; ++q)
            n[q] /= len;
    H.t[rid] = (float)t_hit;
    H.s[rid] = (float)s;
    H.tt[rid] = (float)t;
    double ds = s - TH.t0[0], dt = t - TH.t0[1];
    if (H.bu)
        H.bu[rid] = (float)(TH.J[0] * ds + TH.J[1] * dt);
    if (H.bv)
        H.bv[rid] = (float)(TH.J[2] * ds + TH.J[3] * dt);
    if (H.nx)
        H.nx[rid] = (float)n[0];
    if (H.ny)
        H.ny[rid] = (float)n[1];
    if (H.nz)
        H.nz[rid] = (float)n[2];
}

// The same from the raw record the trace left in the hit arrays.
__device__ __forceinline__ void finalize_one(const TraceParams &p, const PyramidView &P, const TriRecord *tris,
                                             const RaysDev &rays, const HitsDev &H, int64_t rid)
{
    finalize_hit(p, P, tris, rays, H, rid, H.prim[rid], H.t[rid], H.s[rid], H.tt[rid], H.texel_i[rid],
                 H.texel_j[rid]);
}

// The hits of a closest-hit launch from the raw records its trace left.
__global__ void __launch_bounds__(256) k_finalize(TraceParams p)
{
    int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= p.n_rays)
        return;
    if (p.raw) {
        const float4 *src = (const float4 *)(p.raw + rid);
        const float4 r0 = src[0], r1 = src[1];
        finalize_hit(p, p.obj.pyr, p.obj.tris, p.rays, p.hits, rid, __float_as_int(r0.y), r0.x, r1.x, r1.y,
                     __float_as_int(r0.z), __float_as_int(r0.w));
    } else {
        finalize_one(p, p.obj.pyr, p.obj.tris, p.rays, p.hits, rid);
    }
}

// The same for a serving step: each workgroup finalizes its object group's
// step slots, grid-strided over the group's workgroups (the grid is fixed
// for graph replay), with the segments staged in LDS for the search. It
// runs while other slots' persistent traces hold most of the CUs, so its
// waves are kept small (one-wave workgroups, at most 168 VGPRs for 3 waves
// per SIMD): 256-thread workgroups at 182 VGPRs took twice as long there.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_finalize_multi(TraceParams p, StepDesc sd)
{
    int grp = -1;
    for (int g = 0; g < sd.n_grp; ++g)
        if ((int)blockIdx.x >= sd.grp[g].fwg_lo && (int)blockIdx.x < sd.grp[g].fwg_hi)
            grp = g;
    if (grp < 0)
        return;
    const StepGroup &GR = sd.grp[grp];
    __shared__ Segment s_seg[MAX_SEGMENTS];
    {
        const uint32_t *src = (const uint32_t *)sd.seg;
        uint32_t *dst = (uint32_t *)s_seg;
        const int words = sd.n_seg * (int)(sizeof(Segment) / 4);
        for (int q = threadIdx.x; q < words; q += blockDim.x)
            dst[q] = src[q];
    }
    __syncthreads();
    const int64_t stride = (int64_t)(GR.fwg_hi - GR.fwg_lo) * blockDim.x;
    for (int64_t slot = GR.start + (int64_t)(blockIdx.x - GR.fwg_lo) * blockDim.x + threadIdx.x; slot < GR.end;
         slot += stride) {
        int lo = GR.seg_lo, hi = GR.seg_hi - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (s_seg[mid].start <= slot)
                lo = mid;
            else
                hi = mid - 1;
        }
        const Segment &G = s_seg[lo];
        finalize_one(p, GR.obj.pyr, GR.obj.tris, segment_rays(G), segment_hits(G), slot - G.start + G.first);
    }
}


// ------------------------------------------------------------------ adjoint
//
// No counterpart in the reference. A hit is a root of F(a, b) = 0 with
// F = ((S - o) . d1, (S - o) . d2) over the texel (texel-local a, b), where
// S = P0 + a E1 + b E2 + h N, h = strength * bilinear(v00, v10, v01, v11)
// + offset and N the (possibly interpolated) unit normal. For a corner value
// v_c the implicit-function theorem gives (da, db) = -J^-1 dF/dv_c with
// dF/dv_c = strength w_c (d1 . N, d2 . N); t = (S - o) . dn / |d| and
// ng = unit(Sa x Sb) then move through S, Sa, Sb. The kernel forms the
// vector-Jacobian product per ray: with lt = dL/dt, u = (I - ng ng^T)
// dL/dng / |Sa x Sb|, A = Sb x u, B = u x Sa,
//   mu = (lt dn . Sa / |d| + A . Saa + B . Sab, lt dn . Sb / |d| + A . Sab + B . Sbb),
//   eta = J^-T mu,
//   dL/dv_c = strength [lt w_c dn . N / |d| + A . (w_c,a N + w_c Na)
//             + B . (w_c,b N + w_c Nb) - w_c (eta1 d1 . N + eta2 d2 . N)],
// where Saa = 2 h_a Na + h Naa, Sab = h_ab N + h_a Nb + h_b Na + h Nab,
// Sbb = 2 h_b Nb + h Nbb and, for N = M / |M| with M affine in (a, b),
// Naa = -(2 Na (N . Mu) + N (Na . Mu)) / |M| (Nab, Nbb alike).
// The root is the finalized hit's (s, t) brought back to the texel and
// polished in double, so the gradient is the double solve's.

__device__ inline int texel_of(int k, int n, bool repeat)
{
    // node k of the closed grid holds texel k mod n (repeat) or min(k, n - 1)
    return repeat ? (k & (n - 1)) : min(max(k, 0), n - 1);
}

__global__ void __launch_bounds__(256) k_adjoint(AdjointParams p)
{
    const int64_t rid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rid >= p.n_rays)
        return;
    const HitsDev &H = p.hits;
    const int tri = H.prim[rid];
    if (tri < 0)
        return;
    double lt = p.dl_dt ? (double)p.dl_dt[rid] : 0.0;
    double ln[3] = {0.0, 0.0, 0.0};
    if (p.dl_dn[0])
        for (int q = 0; q < 3; ++q)
            ln[q] = (double)p.dl_dn[q][rid];
    if (!isfinite(lt) || !isfinite(ln[0]) || !isfinite(ln[1]) || !isfinite(ln[2]))
        return;
    if (lt == 0.0 && ln[0] == 0.0 && ln[1] == 0.0 && ln[2] == 0.0)
        return;
    const PyramidView &P = p.obj.pyr;
    const TriRecord &T = p.obj.tris[tri];
    const RaysDev &RY = p.rays;
    const double o[3] = {RY.ox[rid], RY.oy[rid], RY.oz[rid]};
    const double d[3] = {RY.dx[rid], RY.dy[rid], RY.dz[rid]};
    const int i = H.texel_i[rid], j = H.texel_j[rid];
    Leaf<double> L;
    leaf_setup<double>(P, T, o, d, i, j, L);
    double a = fmin(fmax(((double)H.s[rid] - L.slo) / L.w, 0.0), 1.0);
    double b = fmin(fmax(((double)H.tt[rid] - L.tlo) / L.w, 0.0), 1.0);
    polish_root(T, L, a, b);
    polish_root(T, L, a, b);
    // The surface and its derivatives at the root.
    const double h = ((L.h00 + L.ha * a) + L.hb * b) + L.tw * a * b;
    const double ha = L.ha + L.tw * b, hb = L.hb + L.tw * a, hab = L.tw;
    double N[3], Na[3] = {0, 0, 0}, Nb[3] = {0, 0, 0}, Naa[3] = {0, 0, 0}, Nab[3] = {0, 0, 0}, Nbb[3] = {0, 0, 0};
    if (T.const_normal) {
        for (int q = 0; q < 3; ++q)
            N[q] = L.N[q];
    } else {
        normal_at(L, a, b, N, Na, Nb);
        double M[3];
        for (int q = 0; q < 3; ++q)
            M[q] = L.M0[q] + a * L.wMu[q] + b * L.wMv[q];
        const double len = sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
        const double nu = N[0] * L.wMu[0] + N[1] * L.wMu[1] + N[2] * L.wMu[2];
        const double nv = N[0] * L.wMv[0] + N[1] * L.wMv[1] + N[2] * L.wMv[2];
        const double au = Na[0] * L.wMu[0] + Na[1] * L.wMu[1] + Na[2] * L.wMu[2];
        const double bu = Nb[0] * L.wMu[0] + Nb[1] * L.wMu[1] + Nb[2] * L.wMu[2];
        const double bv = Nb[0] * L.wMv[0] + Nb[1] * L.wMv[1] + Nb[2] * L.wMv[2];
        for (int q = 0; q < 3; ++q) {
            Naa[q] = -(2.0 * nu * Na[q] + au * N[q]) / len;
            Nab[q] = -(nu * Nb[q] + nv * Na[q] + bu * N[q]) / len;
            Nbb[q] = -(2.0 * nv * Nb[q] + bv * N[q]) / len;
        }
    }
    double Sa[3], Sb[3], Saa[3], Sab[3], Sbb[3];
    for (int q = 0; q < 3; ++q) {
        Sa[q] = L.E1[q] + ha * N[q] + h * Na[q];
        Sb[q] = L.E2[q] + hb * N[q] + h * Nb[q];
        Saa[q] = 2.0 * ha * Na[q] + h * Naa[q];
        Sab[q] = hab * N[q] + ha * Nb[q] + hb * Na[q] + h * Nab[q];
        Sbb[q] = 2.0 * hb * Nb[q] + h * Nbb[q];
    }
    auto dot3 = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double J00 = dot3(L.d1, Sa), J01 = dot3(L.d1, Sb), J10 = dot3(L.d2, Sa), J11 = dot3(L.d2, Sb);
    const double det = J00 * J11 - J01 * J10;
    if (det == 0.0)
        return;
    // The normal term's vectors.
    double nr[3] = {Sa[1] * Sb[2] - Sa[2] * Sb[1], Sa[2] * Sb[0] - Sa[0] * Sb[2], Sa[0] * Sb[1] - Sa[1] * Sb[0]};
    const double nlen = sqrt(dot3(nr, nr));
    double u[3] = {0, 0, 0}, A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
    if (nlen > 0.0) {
        double ng[3] = {nr[0] / nlen, nr[1] / nlen, nr[2] / nlen};
        const double gl = dot3(ng, ln);
        for (int q = 0; q < 3; ++q)
            u[q] = (ln[q] - ng[q] * gl) / nlen;
        A[0] = Sb[1] * u[2] - Sb[2] * u[1];
        A[1] = Sb[2] * u[0] - Sb[0] * u[2];
        A[2] = Sb[0] * u[1] - Sb[1] * u[0];
        B[0] = u[1] * Sa[2] - u[2] * Sa[1];
        B[1] = u[2] * Sa[0] - u[0] * Sa[2];
        B[2] = u[0] * Sa[1] - u[1] * Sa[0];
    }
    const double ltd = lt / L.d_len;
    const double mua = ltd * dot3(L.dn, Sa) + dot3(A, Saa) + dot3(B, Sab);
    const double mub = ltd * dot3(L.dn, Sb) + dot3(A, Sab) + dot3(B, Sbb);
    const double eta1 = (J11 * mua - J10 * mub) / det, eta2 = (-J01 * mua + J00 * mub) / det;
    const double dnN = dot3(L.dn, N), d1N = dot3(L.d1, N), d2N = dot3(L.d2, N);
    const double AN = dot3(A, N), BN = dot3(B, N), ANa = dot3(A, Na), BNb = dot3(B, Nb);
    const double scale = P.scale_d;
    const int n = P.n_leaf;
    for (int c = 0; c < 4; ++c) {
        const int di = c & 1, dj = c >> 1;
        const double wu = di ? a : 1.0 - a, wv = dj ? b : 1.0 - b;
        const double w = wu * wv, wa = (di ? 1.0 : -1.0) * wv, wb = wu * (dj ? 1.0 : -1.0);
        const double g = scale * (ltd * w * dnN + AN * wa + w * ANa + BN * wb + w * BNb - w * (eta1 * d1N + eta2 * d2N));
        const int ti = texel_of(i + di, n, P.repeat), tj = texel_of(j + dj, n, P.repeat);
        unsafeAtomicAdd(p.grad + (size_t)tj * n + ti, g);
    }
}

hipError_t launch_adjoint(const AdjointParams &p, hipStream_t stream)
{
    if (p.n_rays <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_adjoint, dim3((unsigned)((p.n_rays + 255) / 256)), dim3(256), 0, stream, p);
    return hipGetLastError();
}

// ------------------------------------------------------------- hit record

// HeightGrid::h and grad (displaced_surface.cpp:12-47) in double over the
// closed node grid: height and (dh/ds, dh/dt) at tile coordinates (s, t).
__device__ void field_at(const PyramidView &P, double s, double t, double &h, double &hs, double &ht)
{
    const int W = P.n_leaf + 1;
    if (P.repeat) {
        s -= floor(s);
        t -= floor(t);
    } else {
        s = fmin(fmax(s, 0.0), 1.0);
        t = fmin(fmax(t, 0.0), 1.0);
    }
    const double fx = s * (W - 1), fy = t * (W - 1);
    const int i = min((int)fx, W - 2), j = min((int)fy, W - 2);
    const double a = fx - i, b = fy - j;
    const float *row0 = P.nodes + (size_t)j * W, *row1 = row0 + W;
    const double v00 = row0[i], v10 = row0[i + 1], v01 = row1[i], v11 = row1[i + 1];
    h = P.scale_d * ((1.0 - a) * (1.0 - b) * v00 + a * (1.0 - b) * v10 + (1.0 - a) * b * v01 + a * b * v11) +
        P.offset_d;
    hs = P.scale_d * (((1.0 - b) * (v10 - v00) + b * (v11 - v01)) * (W - 1));
    ht = P.scale_d * (((1.0 - a) * (v01 - v00) + a * (v11 - v10)) * (W - 1));
}

__global__ void __launch_bounds__(256) k_hit_record(ObjectDev obj, RaysDev rays, HitsDev hits, int64_t n, UvMap uv,
                                                    int twosided, Instance inst, RayDiffs diffs, float *rec)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const int prim = hits.prim[r];
    const int planes = diffs.on ? 30 : 20;
    if (prim < 0) {
        for (int q = 0; q < planes; ++q)
            rec[q * n + r] = 0.f;
        return;
    }
    const TriRecord &T = obj.tris[prim];
    // st = t0 + J^-1 (u, v) (BaseTriangle::param).
    const double u = hits.bu[r], v = hits.bv[r];
    const double detJ = T.J[0] * T.J[3] - T.J[1] * T.J[2];
    const double s = T.t0[0] + (T.J[3] * u - T.J[1] * v) / detJ;
    const double t = T.t0[1] + (-T.J[2] * u + T.J[0] * v) / detJ;
    // normal_frame_at (displaced_surface.cpp:99-108).
    double M[3], N[3], Nu[3], Nv[3];
    for (int q = 0; q < 3; ++q)
        M[q] = T.m0[q] + s * T.Mu[q] + t * T.Mv[q];
    const double len = sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
    for (int q = 0; q < 3; ++q)
        N[q] = M[q] / len;
    const double nu = N[0] * T.Mu[0] + N[1] * T.Mu[1] + N[2] * T.Mu[2];
    const double nv = N[0] * T.Mv[0] + N[1] * T.Mv[1] + N[2] * T.Mv[2];
    for (int q = 0; q < 3; ++q) {
        Nu[q] = (T.Mu[q] - nu * N[q]) / len;
        Nv[q] = (T.Mv[q] - nv * N[q]) / len;
    }
    double h, hs, ht;
    field_at(obj.pyr, s, t, h, hs, ht);
    double Ss[3], St[3], p[3];
    for (int q = 0; q < 3; ++q) {
        Ss[q] = (T.e1[q] + hs * N[q]) + h * Nu[q];
        St[q] = (T.e2[q] + ht * N[q]) + h * Nv[q];
        p[q] = T.q0[q] + s * T.e1[q] + t * T.e2[q] + h * N[q];
    }
    // Oriented along the base face's winding: (p1 - p0) x (p2 - p0) = det(J^-1) e1 x e2.
    double wind[3] = {T.e1[1] * T.e2[2] - T.e1[2] * T.e2[1], T.e1[2] * T.e2[0] - T.e1[0] * T.e2[2],
                      T.e1[0] * T.e2[1] - T.e1[1] * T.e2[0]};
    if (detJ < 0.0)
        for (int q = 0; q < 3; ++q)
            wind[q] = -wind[q];
    double ng[3] = {Ss[1] * St[2] - Ss[2] * St[1], Ss[2] * St[0] - Ss[0] * St[2], Ss[0] * St[1] - Ss[1] * St[0]};
    if (ng[0] * wind[0] + ng[1] * wind[1] + ng[2] * wind[2] < 0.0)
        for (int q = 0; q < 3; ++q)
            ng[q] = -ng[q];
    double nl = sqrt(ng[0] * ng[0] + ng[1] * ng[1] + ng[2] * ng[2]);
    if (!(nl > 0.0) || !isfinite(nl)) {
        for (int q = 0; q < 3; ++q)
            ng[q] = wind[q];
        nl = sqrt(ng[0] * ng[0] + ng[1] * ng[1] + ng[2] * ng[2]);
    }
    for (int q = 0; q < 3; ++q)
        ng[q] /= nl;
    double uu = (s - uv.offset_u) / uv.scale_u, vv = (t - uv.offset_v) / uv.scale_v;
    if (uv.flip_v)
        vv = 1.0 - vv;
    const double sv = uv.flip_v ? -uv.scale_v : uv.scale_v;
    // The derivatives with respect to the object's uv.
    double dpu[3], dpv[3];
    for (int q = 0; q < 3; ++q) {
        dpu[q] = Ss[q] * uv.scale_u;
        dpv[q] = St[q] * sv;
    }
    auto fcross = [](const float *a, const float *b, float *c) {
        c[0] = a[1] * b[2] - a[2] * b[1];
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto fnormalize = [](float *a) {
        const float l = sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        a[0] /= l;
        a[1] /= l;
        a[2] /= l;
        return isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]);
    };
    // compute_intersection's two-sided flip, decided in fp32 on the
    // object-space ray and normal.
    bool flip = false;
    if (twosided) {
        float dd[3] = {rays.dx[r], rays.dy[r], rays.dz[r]};
        const float n0[3] = {(float)ng[0], (float)ng[1], (float)ng[2]};
        fnormalize(dd);
        flip = -(dd[0] * n0[0] + dd[1] * n0[1] + dd[2] * n0[2]) < 0.f;
    }
    if (inst.on) {
        // To world space: points by m and t, tangents by m, the normal by
        // m^-T (renormalized).
        double pw[3], uw[3], vw[3], nw[3];
        for (int q = 0; q < 3; ++q) {
            const double *row = inst.m + 3 * q, *nrow = inst.nt + 3 * q;
            pw[q] = (row[0] * p[0] + row[1] * p[1] + row[2] * p[2]) + inst.t[q];
            uw[q] = row[0] * dpu[0] + row[1] * dpu[1] + row[2] * dpu[2];
            vw[q] = row[0] * dpv[0] + row[1] * dpv[1] + row[2] * dpv[2];
            nw[q] = nrow[0] * ng[0] + nrow[1] * ng[1] + nrow[2] * ng[2];
        }
        const double nwl = sqrt(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
        for (int q = 0; q < 3; ++q) {
            p[q] = pw[q];
            dpu[q] = uw[q];
            dpv[q] = vw[q];
            ng[q] = nw[q] / nwl;
        }
    }
    // The geometric frame in fp32, as compute_intersection builds it (in
    // world space when a transform is given), then the two-sided flip.
    float ngf[3], du[3], dv[3], pf[3];
    for (int q = 0; q < 3; ++q) {
        ngf[q] = (float)ng[q];
        du[q] = (float)dpu[q];
        dv[q] = (float)dpv[q];
        pf[q] = (float)p[q];
    }
    float bf[3], tf[3];
    fcross(ngf, du, bf);
    if (!fnormalize(bf)) {
        const float ax[3] = {fabsf(ngf[0]) < 0.9f ? 1.f : 0.f, fabsf(ngf[0]) < 0.9f ? 0.f : 1.f, 0.f};
        fcross(ngf, ax, bf);
        fnormalize(bf);
    }
    fcross(bf, ngf, tf);
    fnormalize(tf);
    if (flip)
        for (int q = 0; q < 3; ++q) {
            ngf[q] = -ngf[q];
            du[q] = -du[q];
            tf[q] = -tf[q];
        }
    const float vals[20] = {pf[0],  pf[1],  pf[2], ngf[0], ngf[1], ngf[2], du[0], du[1], du[2], dv[0],
                            dv[1],  dv[2],  (float)uu, (float)vv, tf[0], tf[1], tf[2], bf[0], bf[1], bf[2]};
    for (int q = 0; q < 20; ++q)
        rec[q * n + r] = vals[q];
    if (!diffs.on)
        return;
    // The uv partials from the ray differentials (the renderer's
    // compute_uv_partials, by the published tangent-plane method of PBRT,
    // 3rd ed., sec. 10.1.1): each offset ray meets the plane through p with
    // normal ng at px (py); dp/dx = px - p, and (du/dx, dv/dx) solves
    // dp/dx = du/dx dpdu + dv/dx dpdv on the two axes other than ng's
    // largest. An offset ray parallel to the plane, or a singular system,
    // gives zeros, as does a miss.
    double dpx[2][3] = {{0, 0, 0}, {0, 0, 0}}, duv[2][2] = {{0, 0}, {0, 0}};
    const double plane = ng[0] * p[0] + ng[1] * p[1] + ng[2] * p[2];
    int a0 = 0, a1 = 1;
    if (fabs(ng[0]) > fabs(ng[1]) && fabs(ng[0]) > fabs(ng[2])) {
        a0 = 1;
        a1 = 2;
    } else if (fabs(ng[1]) > fabs(ng[2])) {
        a0 = 0;
        a1 = 2;
    }
    // dpdu as the record reports it (turned by the two-sided flip).
    const double su = flip ? -1.0 : 1.0;
    const double A00 = su * dpu[a0], A01 = dpv[a0], A10 = su * dpu[a1], A11 = dpv[a1];
    const double det = A00 * A11 - A01 * A10;
    const bool solvable = fabs(det) > 1e-12 * (fabs(A00 * A11) + fabs(A01 * A10));
    for (int k = 0; k < 2; ++k) {
        double o[3], d[3];
        for (int q = 0; q < 3; ++q) {
            o[q] = diffs.p[6 * k + q][r];
            d[q] = diffs.p[6 * k + 3 + q][r];
        }
        const double nd = ng[0] * d[0] + ng[1] * d[1] + ng[2] * d[2];
        const double tx = (plane - (ng[0] * o[0] + ng[1] * o[1] + ng[2] * o[2])) / nd;
        if (!isfinite(tx))
            continue;
        for (int q = 0; q < 3; ++q)
            dpx[k][q] = (o[q] + tx * d[q]) - p[q];
        if (!solvable)
            continue;
        const double b0 = dpx[k][a0], b1 = dpx[k][a1];
        const double x0 = (A11 * b0 - A01 * b1) / det, x1 = (A00 * b1 - A10 * b0) / det;
        if (isfinite(x0) && isfinite(x1)) {
            duv[k][0] = x0;
            duv[k][1] = x1;
        }
    }
    const float dvals[10] = {(float)dpx[0][0], (float)dpx[0][1], (float)dpx[0][2], (float)dpx[1][0],
                             (float)dpx[1][1], (float)dpx[1][2], (float)duv[0][0], (float)duv[0][1],
                             (float)duv[1][0], (float)duv[1][1]};
    for (int q = 0; q < 10; ++q)
        rec[(20 + q) * n + r] = dvals[q];
}

hipError_t launch_hit_record(const ObjectDev &obj, const RaysDev &rays, const HitsDev &hits, int64_t n,
                             const UvMap &uv, int twosided, const Instance &inst, const RayDiffs &diffs,
                             float *record, hipStream_t stream)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_hit_record, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, obj, rays, hits, n, uv,
                       twosided, inst, diffs, record);
    return hipGetLastError();
}

// ------------------------------------------------------- height range

// classify_square in double (uv_clip.cpp:27-44): 0 outside, 1 inside, 2 straddle.
__device__ inline int classify_d(const TriRecord &T, double cx, double cy, double s)
{
    if (cx + s <= T.dlo_d[0] || cx - s >= T.dhi_d[0] || cy + s <= T.dlo_d[1] || cy - s >= T.dhi_d[1])
        return 0;
    bool all_inside = true;
    for (int k = 0; k < 3; ++k) {
        double nx = T.dn_d[2 * k], ny = T.dn_d[2 * k + 1];
        double proj = nx * cx + ny * cy;
        double ext = (fabs(nx) + fabs(ny)) * s;
        if (proj + ext <= T.dd_d[k])
            return 0;
        if (proj - ext < T.dd_d[k])
            all_inside = false;
    }
    return all_inside ? 1 : 2;
}

// height_range_walk / triangle_height_range (node_bounds.cpp:218-249).
// One workgroup per triangle. The reference's walk (height_range_walk,
// node_bounds.cpp:216-233: a cell inside the domain contributes its own
// min-max node, a straddling cell descends, down to the leaf level) is split
// at level L0 = root_level - D: each task is one cell of a root cell's
// subtree at L0. A task first classifies the cell's ancestors from the root
// down; an outside ancestor ends it, an inside one contributes its node and
// ends it (the serial walk stops there too), and a cell whose ancestors all
// straddle is walked depth first as the serial walk would. The tasks
// together visit exactly the serial walk's nodes, some of them more than
// once, which min and max do not mind.
constexpr int HR_BLOCK = 256;
constexpr int HR_SPLIT = 5; // levels below the root cells handed out as tasks (4^5 per root)

__global__ void __launch_bounds__(HR_BLOCK) k_height_range(PyramidView P, const TriRecord *tris, int n_tri,
                                                          float *h_range)
{
    const int k = blockIdx.x;
    if (k >= n_tri)
        return;
    const TriRecord &T = tris[k];
    const int L = T.root_level;
    const int D = L < HR_SPLIT ? L : HR_SPLIT;
    const int L0 = L - D;
    const int per_root = 1 << (2 * D);
    const int tasks = T.n_roots * per_root;
    float lo = INFINITY, hi = -INFINITY;
    int stack_l[64], stack_i[64], stack_j[64];
    for (int t = threadIdx.x; t < tasks; t += HR_BLOCK) {
        const int r = t / per_root, m = t % per_root;
        int di = 0, dj = 0;
        for (int b = 0; b < D; ++b) {
            di |= ((m >> (2 * b)) & 1) << b;
            dj |= ((m >> (2 * b + 1)) & 1) << b;
        }
        const int ci = T.root_i[r] * (1 << D) + di, cj = T.root_j[r] * (1 << D) + dj;
        bool done = false;
        for (int lv = L; lv > L0 && !done; --lv) {
            const int ai = ci >> (lv - L0), aj = cj >> (lv - L0);
            const double w = ldexp(1.0, lv) / P.n_leaf;
            const int ov = classify_d(T, (ai + 0.5) * w, (aj + 0.5) * w, 0.5 * w);
            if (ov == 0) {
                done = true;
            } else if (ov == 1) {
                const float2 mm = fetch_mm(P, lv, ai, aj);
                lo = fminf(lo, mm.x);
                hi = fmaxf(hi, mm.y);
                done = true;
            }
        }
        if (done)
            continue;
        int sp = 0;
        stack_l[sp] = L0;
        stack_i[sp] = ci;
        stack_j[sp] = cj;
        ++sp;
        while (sp > 0) {
            --sp;
            const int level = stack_l[sp], i = stack_i[sp], j = stack_j[sp];
            const double w = ldexp(1.0, level) / P.n_leaf;
            const int ov = classify_d(T, (i + 0.5) * w, (j + 0.5) * w, 0.5 * w);
            if (ov == 0)
                continue;
            if (ov == 1 || level == 0) {
                const float2 mm = fetch_mm(P, level, i, j);
                lo = fminf(lo, mm.x);
                hi = fmaxf(hi, mm.y);
                continue;
            }
            if (sp + 4 > 64)
                continue; // cannot happen below level 20
            for (int ch = 3; ch >= 0; --ch) {
                stack_l[sp] = level - 1;
                stack_i[sp] = 2 * i + (ch & 1);
                stack_j[sp] = 2 * j + (ch >> 1);
                ++sp;
            }
        }
    }
    // The workgroup's min and max.
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    __shared__ float s_lo[HR_BLOCK / 64], s_hi[HR_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < HR_BLOCK / 64; ++q) {
            lo = fminf(lo, s_lo[q]);
            hi = fmaxf(hi, s_hi[q]);
        }
        h_range[2 * k] = lo;
        h_range[2 * k + 1] = hi;
    }
}

hipError_t launch_height_range(const PyramidView &P, const TriRecord *tris, int n_tri, float *h_range,
                               hipStream_t stream)
{
    hipLaunchKernelGGL(k_height_range, dim3(n_tri), dim3(HR_BLOCK), 0, stream, P, tris, n_tri, h_range);
    return hipGetLastError();
}

// Zeroes blockDim.x queue heads (one for a launch, MAX_GROUPS for a step).
__global__ void k_queue_reset(unsigned *queue) { queue[threadIdx.x] = 0u; }

template <class R> static size_t trace_lds_bytes(unsigned block)
{
    return (size_t)STACK * block * (sizeof(typename StackSlot<R>::T) + sizeof(uint16_t));
}

// The LDS copy of a small BVH (trace_body's lds_bvh).
template <class R> static size_t bvh_lds_bytes(const ObjectDev &obj)
{
    return obj.bvh_leaves <= LDS_BVH_LEAVES ? BVH_LDS_BYTES : 0;
}

// Persistent single-object launches: workgroups per CU when the runtime
// cannot say how many stay resident.
constexpr int TRACE_GROUPS_PER_CU = 3;

// Multi-request launches use two-wave workgroups: a workgroup's resources
// free as soon as its waves are done, so the next step's waves fill the
// tail of a running step almost wave by wave, while two waves share one LDS
// copy of the object's BVH (one-wave workgroups with a copy each would hold
// 2.25 instead of 3 waves per SIMD).
constexpr unsigned MULTI_BLOCK = 128;
constexpr unsigned MULTI_FINALIZE_BLOCK = 64;

static int cu_count();

// Workgroups of a single-object trace kernel that fit on one CU at once
// (registers and LDS), asked of the runtime once per kernel and LDS size.
template <class R, bool STATS, bool CN, bool TWO> static int resident_groups(size_t lds, bool with_bvh)
{
    static std::atomic<int> cached[2];
    int n = cached[with_bvh ? 1 : 0].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace<R, STATS, CN, TWO>, BLOCK, lds) !=
                hipSuccess ||
            n <= 0)
            n = TRACE_GROUPS_PER_CU;
        cached[with_bvh ? 1 : 0].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Persistent waves: as many workgroups as stay resident together on every
// CU (a workgroup beyond that would start only when another ends, and the
// rays it took from the queue would finish last), fewer for small launches.
template <class R, bool STATS, bool CN, bool TWO>
static void launch_trace_k(const TraceParams &p, int64_t want, hipStream_t stream)
{
    const size_t bvh_lds = bvh_lds_bytes<R>(p.obj);
    const size_t lds = trace_lds_bytes<R>(BLOCK) + bvh_lds;
    const int64_t cap = (int64_t)cu_count() * resident_groups<R, STATS, CN, TWO>(lds, bvh_lds > 0);
    const unsigned blocks = (unsigned)(want < cap ? want : cap);
    hipLaunchKernelGGL((k_trace<R, STATS, CN, TWO>), dim3(blocks), dim3(BLOCK), lds, stream, p);
}

template <class R, bool STATS, bool TWO>
static void launch_trace_l(const TraceParams &p, int64_t want, hipStream_t stream)
{
    if (p.obj.const_normal)
        launch_trace_k<R, STATS, true, TWO>(p, want, stream);
    else
        launch_trace_k<R, STATS, false, TWO>(p, want, stream);
}

static int cu_count()
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

template <class R, bool STATS>
static void launch_trace_r(const TraceParams &p, int64_t want, hipStream_t stream)
{
    if (p.leaf == 1)
        launch_trace_l<R, STATS, true>(p, want, stream);
    else
        launch_trace_l<R, STATS, false>(p, want, stream);
}

hipError_t launch_trace(const TraceParams &p, int precision, hipStream_t stream, hipEvent_t ev_start,
                        hipEvent_t ev_end)
{
    if (p.n_rays <= 0)
        return hipSuccess;
    const int64_t want = (p.n_rays + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, stream, p.queue);
    bool stats = p.stats.node_tests || p.stats.leaf_tests || p.stats.sub_squares || p.stats.triangle_tests ||
                 p.stats.newton_runs || p.stats.newton_iters || p.stats.node_tests_after_hit ||
                 p.stats.leaf_tests_after_hit || p.stats.node_tests_per_level || p.stats.records;
    if (ev_start)
        (void)hipEventRecord(ev_start, stream);
    if (precision == 1) {
        if (stats)
            launch_trace_r<double, true>(p, want, stream);
        else
            launch_trace_r<double, false>(p, want, stream);
    } else {
        if (stats)
            launch_trace_r<float, true>(p, want, stream);
        else
            launch_trace_r<float, false>(p, want, stream);
    }
    hipError_t e = hipGetLastError();
    if (ev_end)
        (void)hipEventRecord(ev_end, stream);
    if (e != hipSuccess || p.any_hit)
        return e;
    unsigned fblocks = (unsigned)((p.n_rays + 255) / 256);
    hipLaunchKernelGGL(k_finalize, dim3(fblocks), dim3(256), 0, stream, p);
    return hipGetLastError();
}

// The step graph's launch shapes: as many trace waves as stay resident with
// the persistent single-object launch (3 per SIMD), in two-wave workgroups
// with the stacks, the BVH's room and the segments in LDS; a finalize grid
// of one-wave workgroups that strides over the step. Both grids are fixed
// for graph replay and hold room for the one workgroup each group gets at
// least.
static unsigned multi_trace_share() { return (unsigned)cu_count() * 12 / (MULTI_BLOCK / 64); }
static unsigned multi_trace_blocks() { return multi_trace_share() + MAX_GROUPS; }
static size_t multi_trace_lds()
{
    return trace_lds_bytes<float>(MULTI_BLOCK) + BVH_LDS_BYTES + sizeof(Segment) * MAX_SEGMENTS;
}
constexpr unsigned MULTI_FINALIZE_SHARE = 4096;
constexpr unsigned MULTI_FINALIZE_BLOCKS = MULTI_FINALIZE_SHARE + MAX_GROUPS;

void step_partition(StepDesc &d)
{
    const unsigned share[2] = {multi_trace_share(), MULTI_FINALIZE_SHARE};
    int next[2] = {0, 0};
    for (int g = 0; g < d.n_grp; ++g) {
        StepGroup &G = d.grp[g];
        const int64_t n = G.end - G.start;
        int cnt[2];
        for (int k = 0; k < 2; ++k) {
            int64_t c = d.total > 0 ? (int64_t)share[k] * n / d.total : 1;
            cnt[k] = (int)std::max<int64_t>(1, c);
        }
        G.wg_lo = next[0];
        G.wg_hi = next[0] += cnt[0];
        G.fwg_lo = next[1];
        G.fwg_hi = next[1] += cnt[1];
    }
}

static void step_node_params(const StepGraph &g, bool trace, void **args, hipKernelNodeParams &kp)
{
    std::memset(&kp, 0, sizeof kp);
    if (trace) {
        // A step holds objects of one kind only (constant normals or
        // interpolated ones), so it runs the specialisation a single-object
        // launch of those objects would.
        kp.func = g.const_normal ? (void *)k_trace_multi<true> : (void *)k_trace_multi<false>;
        kp.gridDim = dim3(multi_trace_blocks());
        kp.blockDim = dim3(MULTI_BLOCK);
        kp.sharedMemBytes = (unsigned)multi_trace_lds();
    } else {
        kp.func = (void *)k_finalize_multi;
        kp.gridDim = dim3(MULTI_FINALIZE_BLOCKS);
        kp.blockDim = dim3(MULTI_FINALIZE_BLOCK);
    }
    kp.kernelParams = args;
}

hipError_t step_graph_create(unsigned *queue, bool const_normal, StepGraph &g)
{
    g = StepGraph{};
    g.const_normal = const_normal;
    hipError_t e = hipGraphCreate(&g.graph, 0);
    if (e != hipSuccess)
        return e;
    // The queue reset takes the queue's address; the trace and finalize
    // nodes get each step's parameters at launch (an empty step here).
    hipGraphNode_t reset = nullptr;
    hipKernelNodeParams kp;
    std::memset(&kp, 0, sizeof kp);
    void *reset_args[1] = {&queue};
    kp.func = (void *)k_queue_reset;
    kp.gridDim = dim3(1);
    kp.blockDim = dim3(MAX_GROUPS);
    kp.kernelParams = reset_args;
    TraceParams p;
    std::memset(&p, 0, sizeof p);
    p.queue = queue;
    StepDesc desc;
    std::memset(&desc, 0, sizeof desc);
    void *args[2] = {&p, &desc};
    hipKernelNodeParams tp, fp;
    step_node_params(g, true, args, tp);
    step_node_params(g, false, args, fp);
    if ((e = hipGraphAddKernelNode(&reset, g.graph, nullptr, 0, &kp)) != hipSuccess ||
        (e = hipGraphAddKernelNode(&g.trace, g.graph, &reset, 1, &tp)) != hipSuccess ||
        (e = hipGraphAddKernelNode(&g.finalize, g.graph, &g.trace, 1, &fp)) != hipSuccess ||
        (e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0)) != hipSuccess) {
        step_graph_destroy(g);
        return e;
    }
    return hipSuccess;
}

hipError_t step_graph_launch(StepGraph &g, const TraceParams &p, const StepDesc &desc, hipStream_t stream)
{
    TraceParams pp = p;
    StepDesc dd = desc;
    void *args[2] = {&pp, &dd};
    hipKernelNodeParams tp, fp;
    step_node_params(g, true, args, tp);
    step_node_params(g, false, args, fp);
    hipError_t e;
    if ((e = hipGraphExecKernelNodeSetParams(g.exec, g.trace, &tp)) != hipSuccess ||
        (e = hipGraphExecKernelNodeSetParams(g.exec, g.finalize, &fp)) != hipSuccess)
        return e;
    return hipGraphLaunch(g.exec, stream);
}

void step_graph_destroy(StepGraph &g)
{
    if (g.exec)
        (void)hipGraphExecDestroy(g.exec);
    if (g.graph)
        (void)hipGraphDestroy(g.graph);
    g = StepGraph{};
}

} // namespace dmap_amd
