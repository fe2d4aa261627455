// Taylor pyramid build on gfx950 (kernel #1).
//
// The closed node grid (close_tile, texture_grid.cpp:118-132) and the
// eight-channel pyramid of taylor_pyramid.cpp: the exact leaf level
// (:45-67), the conservative 4-to-1 fold (:77-138) and the direct per-level
// build (:151-201). Arithmetic is double precision with contraction off
// (this file is compiled with -ffp-contract=off), in the reference's
// operation order, so the double node equals the reference's bit for bit;
// only the store rounds it to fp32, outward (see layout.h).
//
// The fold is HBM-bound: a 32 x 32 texel tile per workgroup reads its
// 33 x 33 nodes once, computes its leaves in registers and folds five
// levels in LDS, writing each level once. The top cell of every tile goes
// to a double scratch level that the next pass folds the same way.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "layout.h"

namespace dmap_amd
{

struct DNode
{
    double h0, gu, gv, r, ru, rv, hmin, hmax;
};

// The next float above f (not NaN, below +inf), by one step of its bit
// pattern: magnitudes are ordered as unsigned integers, so a positive f
// steps up, a negative one steps towards zero, and either zero goes to the
// smallest positive denormal: nextafterf(f, +inf), in five integer
// operations and no branch.
__device__ inline float next_up_f(float f)
{
    const uint32_t b = __float_as_uint(f);
    return __uint_as_float((b & 0x7fffffffu) == 0 ? 1u : (int32_t)b >= 0 ? b + 1 : b - 1);
}
__device__ inline float round_down_f(double x)
{
    const float f = (float)x;
    return (double)f > x ? -next_up_f(-f) : f;
}
__device__ inline float round_up_f(double x)
{
    const float f = (float)x;
    return (double)f < x ? next_up_f(f) : f;
}

// The fp32 store of one node: the outward rounding of layout.h. `half` is
// the cell's half-extent; the offset (taylor_pyramid.cpp:242-250) is added
// to h0, h_min and h_max here, after the double build, as the reference does.
__device__ void store_node(const PyramidOut &P, int level, int i, int j, const DNode &nd)
{
    double off = P.offset;
    double h0 = off != 0.0 ? nd.h0 + off : nd.h0;
    double hmin = off != 0.0 ? nd.hmin + off : nd.hmin;
    double hmax = off != 0.0 ? nd.hmax + off : nd.hmax;
    float h0f = (float)h0, guf = (float)nd.gu, gvf = (float)nd.gv;
    double half = ldexp(0.5, level - (__ffs(P.n_leaf) - 1)); // 0.5 * 2^level / n_leaf, exactly
    double egu = fabs(nd.gu - (double)guf), egv = fabs(nd.gv - (double)gvf);
    double r = nd.r + fabs(h0 - (double)h0f) + (egu + egv) * half;
    double ru = nd.ru + egu, rv = nd.rv + egv;
    // The sums above round too: a relative 2^-50 nudge covers them.
    const double bump = 1.0 + 0x1p-50;
    uint32_t m = (uint32_t)P.n_leaf >> level;
    size_t idx = level_base(m) + morton2((uint32_t)i, (uint32_t)j);
    P.mm[idx] = make_float2(round_down_f(hmin), round_up_f(hmax));
    P.tp[idx] = make_float4(h0f, guf, gvf, round_up_f(r * bump));
    P.rr[idx] = make_float2(round_up_f(ru * bump), round_up_f(rv * bump));
}

// leaf_level (taylor_pyramid.cpp:45-67) for one texel from its four scaled nodes.
// n is a power of two, so w = 1 / n is exact and dividing by w or w^2 is
// the exact scaling by n or n^2: the products below equal the reference's
// quotients bit for bit.
__device__ inline DNode leaf_node(double v00, double v10, double v01, double v11, int n)
{
    const double dn = (double)n, w = ldexp(1.0, -(__ffs(n) - 1)), s = 0.5 * w;
    DNode o;
    o.h0 = 0.25 * (v00 + v10 + v01 + v11);
    o.gu = 0.5 * ((v10 - v00) + (v11 - v01)) * dn;
    o.gv = 0.5 * ((v01 - v00) + (v11 - v10)) * dn;
    double c3 = (v00 - v10 - v01 + v11) * (dn * dn);
    o.r = fabs(c3) * s * s;
    o.ru = fabs(c3) * s;
    o.rv = fabs(c3) * s;
    o.hmin = fmin(fmin(v00, v10), fmin(v01, v11));
    o.hmax = fmax(fmax(v00, v10), fmax(v01, v11));
    return o;
}

// fold (taylor_pyramid.cpp:77-138) of the four children c[a * 2 + b]
// (a along v, b along u) of a parent on a level of m cells per side.
__device__ DNode fold_node(const DNode c[4], int m)
{
    double s = ldexp(0.5, -(__ffs(m) - 1)); // 0.5 / m, m a power of two
    const double off[2] = {-0.5 * s, 0.5 * s};
    double gu_hi = -INFINITY, gu_lo = INFINITY, gv_hi = -INFINITY, gv_lo = INFINITY;
    double h_min = INFINITY, h_max = -INFINITY;
    for (int q = 0; q < 4; ++q) {
        gu_hi = fmax(gu_hi, c[q].gu + c[q].ru);
        gu_lo = fmin(gu_lo, c[q].gu - c[q].ru);
        gv_hi = fmax(gv_hi, c[q].gv + c[q].rv);
        gv_lo = fmin(gv_lo, c[q].gv - c[q].rv);
        h_min = fmin(h_min, c[q].hmin);
        h_max = fmax(h_max, c[q].hmax);
    }
    DNode o;
    double gu_p = 0.5 * (gu_hi + gu_lo), gv_p = 0.5 * (gv_hi + gv_lo);
    o.gu = gu_p;
    o.ru = 0.5 * (gu_hi - gu_lo);
    o.gv = gv_p;
    o.rv = 0.5 * (gv_hi - gv_lo);
    o.hmin = h_min;
    o.hmax = h_max;
    double upper = -INFINITY, lower = INFINITY;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const DNode &k = c[a * 2 + b];
            double dx = off[b], dy = off[a];
            double du_g = k.gu - gu_p, dv_g = k.gv - gv_p;
            double base = k.h0 - k.gu * dx - k.gv * dy;
            for (int cu = 0; cu < 2; ++cu)
                for (int cv = 0; cv < 2; ++cv) {
                    double corner = base + du_g * (dx + off[cu]) + dv_g * (dy + off[cv]);
                    upper = fmax(upper, corner + k.r);
                    lower = fmin(lower, corner - k.r);
                }
        }
    o.h0 = 0.5 * (upper + lower);
    o.r = 0.5 * (upper - lower);
    return o;
}

// close_tile (texture_grid.cpp:118-132): n x n texels -> (n+1)^2 nodes.
__global__ void k_close_tile(const float *__restrict__ texels, float *__restrict__ nodes, int n, int repeat)
{
    int W = n + 1;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)W * W)
        return;
    int j = (int)(k / W), i = (int)(k % W);
    int tj = repeat ? j % n : min(j, n - 1);
    int ti = repeat ? i % n : min(i, n - 1);
    nodes[k] = texels[(size_t)tj * n + ti];
}

// In-block fold of the (cur x cur) double level held in buf (Morton order)
// up to one cell; each produced level is stored. Returns, in thread 0, the
// top cell.
__device__ void fold_in_lds(const PyramidOut &P, DNode *buf, int cur, int level, int ci0, int cj0, DNode &top)
{
    int t = threadIdx.x;
    while (cur > 1) {
        int half = cur / 2;
        int m_parent = P.n_leaf >> (level + 1);
        DNode parent;
        bool active = t < half * half;
        int x = active ? (int)compact1by1((uint32_t)t) : 0, y = active ? (int)compact1by1((uint32_t)t >> 1) : 0;
        if (active) {
            DNode c[4];
            for (int q = 0; q < 4; ++q)
                c[q] = buf[4 * t + q];
            parent = fold_node(c, m_parent);
        }
        __syncthreads();
        if (active) {
            store_node(P, level + 1, ci0 / 2 + x, cj0 / 2 + y, parent);
            buf[t] = parent;
        }
        __syncthreads();
        cur = half;
        ++level;
        ci0 /= 2;
        cj0 /= 2;
    }
    if (t == 0)
        top = buf[0];
}

// Leaves of a T x T texel tile (T = min(32, n)) and the fold of the tile to
// one cell. tile_top receives that cell (double, row-major over the
// (n / T)^2 tiles) for the next pass.
__global__ void __launch_bounds__(256) k_leaf_fold(PyramidOut P, const float *__restrict__ nodes, int T,
                                                   double *__restrict__ tile_top)
{
    __shared__ double V[33 * 33];
    __shared__ DNode buf[256];
    int n = P.n_leaf, W = n + 1;
    int tiles_x = n / T;
    int I0 = (blockIdx.x % tiles_x) * T, J0 = (blockIdx.x / tiles_x) * T;
    int t = threadIdx.x;
    for (int k = t; k < (T + 1) * (T + 1); k += blockDim.x) {
        int y = k / (T + 1), x = k % (T + 1);
        V[y * 33 + x] = (double)nodes[(size_t)(J0 + y) * W + (I0 + x)] * P.scale;
    }
    __syncthreads();
    if (T == 1) {
        if (t == 0) {
            DNode lf = leaf_node(V[0], V[1], V[33], V[34], n);
            store_node(P, 0, I0, J0, lf);
            tile_top[blockIdx.x] = lf.h0; // unused: a one-texel tile is the root
        }
        return;
    }
    // Thread t owns the 2 x 2 texel block at Morton position t, so its four
    // leaves are Morton entries 4t..4t+3 and a wave stores one contiguous run.
    int half = T / 2;
    DNode parent;
    bool active = t < half * half;
    int x = active ? (int)compact1by1((uint32_t)t) : 0, y = active ? (int)compact1by1((uint32_t)t >> 1) : 0;
    if (active) {
        DNode c[4];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                int ly = 2 * y + a, lx = 2 * x + b;
                c[a * 2 + b] = leaf_node(V[ly * 33 + lx], V[ly * 33 + lx + 1], V[(ly + 1) * 33 + lx],
                                         V[(ly + 1) * 33 + lx + 1], n);
                store_node(P, 0, I0 + lx, J0 + ly, c[a * 2 + b]);
            }
        parent = fold_node(c, n / 2);
        store_node(P, 1, I0 / 2 + x, J0 / 2 + y, parent);
        buf[t] = parent;
    }
    __syncthreads();
    DNode top;
    fold_in_lds(P, buf, half, 1, I0 / 2, J0 / 2, top);
    if (t == 0) {
        double *dst = tile_top + (size_t)blockIdx.x * 8;
        dst[0] = top.h0;
        dst[1] = top.gu;
        dst[2] = top.gv;
        dst[3] = top.r;
        dst[4] = top.ru;
        dst[5] = top.rv;
        dst[6] = top.hmin;
        dst[7] = top.hmax;
    }
}

// One fold pass over a double level (m0 x m0, row-major, 8 doubles per
// cell) at `level`: tiles of T x T cells (T = min(32, m0)) fold log2(T)
// levels; each tile's top cell goes to tile_top.
__global__ void __launch_bounds__(256) k_fold_pass(PyramidOut P, const double *__restrict__ src, int level, int m0,
                                                   int T, double *__restrict__ tile_top)
{
    __shared__ DNode buf[256];
    int tiles_x = m0 / T;
    int I0 = (blockIdx.x % tiles_x) * T, J0 = (blockIdx.x / tiles_x) * T;
    int t = threadIdx.x, half = T / 2;
    bool active = t < half * half;
    int x = active ? (int)compact1by1((uint32_t)t) : 0, y = active ? (int)compact1by1((uint32_t)t >> 1) : 0;
    if (active) {
        DNode c[4];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const double *s = src + ((size_t)(J0 + 2 * y + a) * m0 + (I0 + 2 * x + b)) * 8;
                c[a * 2 + b] = DNode{s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]};
            }
        DNode parent = fold_node(c, m0 / 2);
        store_node(P, level + 1, I0 / 2 + x, J0 / 2 + y, parent);
        buf[t] = parent;
    }
    __syncthreads();
    DNode top;
    fold_in_lds(P, buf, half, level + 1, I0 / 2, J0 / 2, top);
    if (t == 0) {
        double *dst = tile_top + (size_t)blockIdx.x * 8;
        dst[0] = top.h0;
        dst[1] = top.gu;
        dst[2] = top.gv;
        dst[3] = top.r;
        dst[4] = top.ru;
        dst[5] = top.rv;
        dst[6] = top.hmin;
        dst[7] = top.hmax;
    }
}

// ---------------------------------------------------------------- direct

// Texel coefficients c1, c2, c3 (taylor_pyramid.cpp:20-39) of texel (i, j);
// the divisions by w = 1 / n are the exact scalings by n and n^2.
__device__ inline void texel_c(const float *nodes, int W, double scale, int n, int i, int j, double &c1, double &c2,
                               double &c3)
{
    const double dn = (double)n;
    double v00 = (double)nodes[(size_t)j * W + i] * scale, v10 = (double)nodes[(size_t)j * W + i + 1] * scale;
    double v01 = (double)nodes[(size_t)(j + 1) * W + i] * scale;
    double v11 = (double)nodes[(size_t)(j + 1) * W + i + 1] * scale;
    c1 = 0.5 * ((v10 - v00) + (v11 - v01)) * dn;
    c2 = 0.5 * ((v01 - v00) + (v11 - v10)) * dn;
    c3 = (v00 - v10 - v01 + v11) * (dn * dn);
}

// direct_level (taylor_pyramid.cpp:151-201), one thread per cell (fine levels).
__global__ void k_direct_thread(PyramidOut P, const float *__restrict__ nodes, int level)
{
    int n = P.n_leaf, W = n + 1, k = 1 << level, m = n >> level;
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (int64_t)m * m)
        return;
    int J = (int)(id / m), I = (int)(id % m);
    double s_texel = 0.5 / n;
    double gu_hi = -INFINITY, gu_lo = INFINITY, gv_hi = -INFINITY, gv_lo = INFINITY;
    for (int b = 0; b < k; ++b)
        for (int a = 0; a < k; ++a) {
            double c1, c2, c3;
            texel_c(nodes, W, P.scale, n, I * k + a, J * k + b, c1, c2, c3);
            double spread = fabs(c3) * s_texel;
            gu_hi = fmax(gu_hi, c1 + spread);
            gu_lo = fmin(gu_lo, c1 - spread);
            gv_hi = fmax(gv_hi, c2 + spread);
            gv_lo = fmin(gv_lo, c2 - spread);
        }
    DNode o;
    double gu = 0.5 * (gu_hi + gu_lo), gv = 0.5 * (gv_hi + gv_lo);
    o.gu = gu;
    o.ru = 0.5 * (gu_hi - gu_lo);
    o.gv = gv;
    o.rv = 0.5 * (gv_hi - gv_lo);
    double upper = -INFINITY, lower = INFINITY, h_min = INFINITY, h_max = -INFINITY;
    for (int b = 0; b <= k; ++b) {
        double dv = (b - 0.5 * k) / n;
        for (int a = 0; a <= k; ++a) {
            double du = (a - 0.5 * k) / n;
            double value = (double)nodes[(size_t)(J * k + b) * W + (I * k + a)] * P.scale;
            double residual = value - gu * du - gv * dv;
            upper = fmax(upper, residual);
            lower = fmin(lower, residual);
            h_min = fmin(h_min, value);
            h_max = fmax(h_max, value);
        }
    }
    o.h0 = 0.5 * (upper + lower);
    o.r = 0.5 * (upper - lower);
    o.hmin = h_min;
    o.hmax = h_max;
    store_node(P, level, I, J, o);
}

__device__ inline double block_max(double v, double *red)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
        r = fmax(r, red[w]);
    __syncthreads();
    return r;
}

// direct_level, one 256-thread workgroup per cell (coarse levels).
__global__ void __launch_bounds__(256) k_direct_block(PyramidOut P, const float *__restrict__ nodes, int level)
{
    __shared__ double red[4];
    int n = P.n_leaf, W = n + 1, k = 1 << level, m = n >> level;
    int J = blockIdx.x / m, I = blockIdx.x % m;
    double s_texel = 0.5 / n;
    double gu_hi = -INFINITY, gu_lo = INFINITY, gv_hi = -INFINITY, gv_lo = INFINITY;
    for (int q = threadIdx.x; q < k * k; q += blockDim.x) {
        int b = q / k, a = q % k;
        double c1, c2, c3;
        texel_c(nodes, W, P.scale, n, I * k + a, J * k + b, c1, c2, c3);
        double spread = fabs(c3) * s_texel;
        gu_hi = fmax(gu_hi, c1 + spread);
        gu_lo = fmin(gu_lo, c1 - spread);
        gv_hi = fmax(gv_hi, c2 + spread);
        gv_lo = fmin(gv_lo, c2 - spread);
    }
    gu_hi = block_max(gu_hi, red);
    gu_lo = -block_max(-gu_lo, red);
    gv_hi = block_max(gv_hi, red);
    gv_lo = -block_max(-gv_lo, red);
    double gu = 0.5 * (gu_hi + gu_lo), gv = 0.5 * (gv_hi + gv_lo);
    double upper = -INFINITY, lower = INFINITY, h_min = INFINITY, h_max = -INFINITY;
    for (int q = threadIdx.x; q < (k + 1) * (k + 1); q += blockDim.x) {
        int b = q / (k + 1), a = q % (k + 1);
        double dv = (b - 0.5 * k) / n, du = (a - 0.5 * k) / n;
        double value = (double)nodes[(size_t)(J * k + b) * W + (I * k + a)] * P.scale;
        double residual = value - gu * du - gv * dv;
        upper = fmax(upper, residual);
        lower = fmin(lower, residual);
        h_min = fmin(h_min, value);
        h_max = fmax(h_max, value);
    }
    upper = block_max(upper, red);
    lower = -block_max(-lower, red);
    h_min = -block_max(-h_min, red);
    h_max = block_max(h_max, red);
    if (threadIdx.x == 0) {
        DNode o;
        o.gu = gu;
        o.ru = 0.5 * (gu_hi - gu_lo);
        o.gv = gv;
        o.rv = 0.5 * (gv_hi - gv_lo);
        o.h0 = 0.5 * (upper + lower);
        o.r = 0.5 * (upper - lower);
        o.hmin = h_min;
        o.hmax = h_max;
        store_node(P, level, I, J, o);
    }
}

// ----------------------------------------------------------------- launch

hipError_t launch_close_tile(const float *texels, float *nodes, int n, int repeat, hipStream_t stream)
{
    int64_t total = (int64_t)(n + 1) * (n + 1);
    int blocks = (int)((total + 255) / 256);
    hipLaunchKernelGGL(k_close_tile, dim3(blocks), dim3(256), 0, stream, texels, nodes, n, repeat);
    return hipGetLastError();
}

size_t fold_scratch_doubles(int n)
{
    int T = n < 32 ? n : 32;
    size_t tiles = (size_t)(n / T) * (n / T);
    return 2 * tiles * 8 + 16;
}

hipError_t launch_build_fold(const PyramidOut &P, const float *nodes, double *scratch, hipStream_t stream)
{
    int n = P.n_leaf;
    int T = n < 32 ? n : 32;
    int tiles = (n / T) * (n / T);
    size_t half = fold_scratch_doubles(n) / 2;
    double *bufs[2] = {scratch, scratch + half};
    hipLaunchKernelGGL(k_leaf_fold, dim3(tiles), dim3(256), 0, stream, P, nodes, T, bufs[0]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    int level = 0;
    while ((1 << level) < T)
        ++level; // the leaf pass folded log2(T) levels
    int m0 = n / T, which = 0;
    while (m0 > 1) {
        int T2 = m0 < 32 ? m0 : 32;
        int tiles2 = (m0 / T2) * (m0 / T2);
        hipLaunchKernelGGL(k_fold_pass, dim3(tiles2), dim3(256), 0, stream, P, bufs[which], level, m0, T2,
                           bufs[1 - which]);
        if ((e = hipGetLastError()) != hipSuccess)
            return e;
        int folded = 0;
        while ((1 << folded) < T2)
            ++folded;
        level += folded;
        m0 /= T2;
        which = 1 - which;
    }
    return hipSuccess;
}

hipError_t launch_build_direct(const PyramidOut &P, const float *nodes, hipStream_t stream)
{
    int n = P.n_leaf;
    for (int level = 0; level < P.n_levels; ++level) {
        int k = 1 << level, m = n >> level;
        if (k <= 8) {
            int64_t cells = (int64_t)m * m;
            hipLaunchKernelGGL(k_direct_thread, dim3((unsigned)((cells + 127) / 128)), dim3(128), 0, stream, P,
                               nodes, level);
        } else {
            hipLaunchKernelGGL(k_direct_block, dim3((unsigned)((int64_t)m * m)), dim3(256), 0, stream, P, nodes,
                               level);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

} // namespace dmap_amd
