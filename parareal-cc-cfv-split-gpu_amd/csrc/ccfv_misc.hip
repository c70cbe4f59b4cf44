// ccfv_misc.hip — element-wise kernels of the Parareal driver, reductions and
// the initial / exact cell averages.
#include "ccfv_line.h"

#include <algorithm>
#include <cstdint>

namespace ccfv {

int chunk_for(int n)
{
    int ch = 1;
    while (ch * 64 < n) ch *= 2;
    return ch;
}

bool supported_n(int n)
{
    if (n < 8 || n > 2048 || (n & 1)) return false;
    return n % chunk_for(n) == 0;
}

// ---------------------------------------------------------------------------
// element-wise kernels
// ---------------------------------------------------------------------------
// U = G_new + (F - G_old), with increment partials (T2S4_Parareal.cu:771-782)
__global__ __launch_bounds__(256) void correction_kernel(
    const double *g_new, const double *g_old, const double *f, double *u,
    size_t n, double *partial)
{
    double s0 = 0.0, s1 = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        double un = g_new[i] + (f[i] - g_old[i]);
        if (partial) {
            double d = un - u[i];
            s0 += d * d;
            s1 += un * un;
        }
        u[i] = un;
    }
    if (partial) block_reduce2(s0, s1, partial, blockIdx.x);
}

hipError_t launch_correction(const double *g_new, const double *g_old,
                             const double *f, double *u, size_t n,
                             double *partial, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(correction_kernel, dim3(nblocks), dim3(256), 0, s, g_new,
                       g_old, f, u, n, partial);
    return hipGetLastError();
}

// fixed-order sum of count (num, den) pairs
__global__ __launch_bounds__(256) void reduce_partials_kernel(
    const double *partial, int count, double *out2)
{
    __shared__ double r0[256], r1[256];
    double s0 = 0.0, s1 = 0.0;
    for (int i = threadIdx.x; i < count; i += 256) {
        s0 += partial[2 * i];
        s1 += partial[2 * i + 1];
    }
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[0] = r0[0];
        out2[1] = r1[0];
    }
}

hipError_t launch_reduce_partials(const double *partial, int count, double *out2,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s, partial,
                       count, out2);
    return hipGetLastError();
}

// one wavefront per slot: lane l sums partials l, l+64, ... in order, then a
// fixed butterfly across lanes (deterministic for a given segment length)
__global__ __launch_bounds__(64) void reduce_segments_kernel(
    const double *partial, int seg, int first, double *out)
{
    const int s = first + blockIdx.x, lane = threadIdx.x;
    const double *p = partial + 2 * (size_t)seg * s;
    double s0 = 0.0, s1 = 0.0;
    for (int i = lane; i < seg; i += 64) {
        s0 += p[2 * i];
        s1 += p[2 * i + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
    }
    if (lane == 0) {
        out[2 * s] = s0;
        out[2 * s + 1] = s1;
    }
}

hipError_t launch_reduce_segments(const double *partial, int seg, int first,
                                  int count, double *out, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_segments_kernel, dim3(count), dim3(64), 0, s,
                       partial, seg, first, out);
    return hipGetLastError();
}

// ragged elementwise correction: block b works for item i on element blocks
// b - prefix[i] of prefix[i+1] - prefix[i]
__global__ __launch_bounds__(256) void ragged_correction_kernel(
    const RaggedCorr *items, const int *prefix, int nitems)
{
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const RaggedCorr it = items[i];
    const int nb = prefix[i + 1] - prefix[i], lb = b - prefix[i];
    double s0 = 0.0, s1 = 0.0;
    for (size_t e = lb * (size_t)blockDim.x + threadIdx.x; e < it.n;
         e += (size_t)nb * blockDim.x) {
        double un = it.g_new[e] + (it.f[e] - it.g_old[e]);
        if (it.partial) {
            double d = un - it.u[e];
            s0 += d * d;
            s1 += un * un;
        }
        it.u[e] = un;
    }
    if (it.partial) block_reduce2(s0, s1, it.partial, lb);
}

hipError_t launch_ragged_correction(const RaggedCorr *items, const int *prefix,
                                    int nitems, int nblocks, hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(ragged_correction_kernel, dim3(nblocks), dim3(256), 0, s, items,
                       prefix, nitems);
    return hipGetLastError();
}

// ragged slot sums: block b reduces slot (b - prefix[i]) of item i
__global__ __launch_bounds__(64) void ragged_reduce_kernel(const RaggedReduce *items,
                                                           const int *prefix, int nitems)
{
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const RaggedReduce it = items[i];
    const int s = it.first + (b - prefix[i]), lane = threadIdx.x;
    const double *p = it.partial + 2 * (size_t)it.seg * s;
    double s0 = 0.0, s1 = 0.0;
    for (int k = lane; k < it.seg; k += 64) {
        s0 += p[2 * k];
        s1 += p[2 * k + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
    }
    if (lane == 0) {
        it.out[2 * s] = s0;
        it.out[2 * s + 1] = s1;
    }
}

hipError_t launch_ragged_reduce(const RaggedReduce *items, const int *prefix,
                                int nitems, int nblocks, hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(ragged_reduce_kernel, dim3(nblocks), dim3(64), 0, s, items,
                       prefix, nitems);
    return hipGetLastError();
}

// ragged copies: block b copies element blocks of item i
__global__ __launch_bounds__(256) void ragged_copy_kernel(const RaggedCopy *items,
                                                          const int *prefix, int nitems)
{
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const RaggedCopy it = items[i];
    const int nb = prefix[i + 1] - prefix[i], lb = b - prefix[i];
    for (size_t e = lb * (size_t)blockDim.x + threadIdx.x; e < it.n;
         e += (size_t)nb * blockDim.x)
        it.dst[e] = it.src[e];
}

hipError_t launch_ragged_copy(const RaggedCopy *items, const int *prefix, int nitems,
                              int nblocks, hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(ragged_copy_kernel, dim3(nblocks), dim3(256), 0, s, items, prefix,
                       nitems);
    return hipGetLastError();
}

__global__ void copy_kernel(const double *src, double *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// 16 bytes per work-item and pass (both pointers 16-byte aligned): n2 pairs,
// then the odd last element
__global__ void copy2_kernel(const double2 *src, double2 *dst, size_t n2, const double *s1,
                             double *d1, bool tail)
{
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (size_t i = i0; i < n2; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
    if (tail && i0 == 0) *d1 = *s1;
}

hipError_t launch_copy(const double *src, double *dst, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
        const size_t n2 = n / 2;
        int blocks = (int)std::min<size_t>((n2 + 255) / 256, 16384);
        if (blocks < 1) blocks = 1;
        hipLaunchKernelGGL(copy2_kernel, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<const double2 *>(src),
                           reinterpret_cast<double2 *>(dst), n2, src + (n - 1), dst + (n - 1),
                           (n & 1) != 0);
        return hipGetLastError();
    }
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

// face sines and cosines of the swirl, with the expressions of
// face_foot_swirl (so the table holds the values it would compute)
__global__ void fill_trig_kernel(double lo, double d, int n, double *out)
{
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n) return;
    const double s = lo + d * q;
    out[q] = sin(M_PI * s);
    out[n + 1 + q] = cos(M_PI * s);
}

hipError_t launch_fill_trig(double lo, double d, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(fill_trig_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, s, lo, d, n,
                       out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launch-rate floor of the one-field sweeps (instrumentation): the sweeps'
// grid with no work, and a copy of the field with a sweep's memory pattern
// and no arithmetic.  One wave per line, lanes of ch contiguous cells.
// ---------------------------------------------------------------------------
__global__ void floor_empty_kernel(int *never)
{
    if (never && threadIdx.x == 1u << 30) *never = 1;  // no store is ever made
}

// y pattern: line i is row i (contiguous); per lane CH cells, 16-byte moves.
// Every load of a lane is issued before its first store, as in the sweeps
// (in and out may be the same field).
template <int CH>
__global__ void floor_rows_kernel(const double *in, double *out, int nl, int n, int ld)
{
    const int line = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (line >= nl || lane * CH >= n) return;
    const double *src = in + (size_t)line * ld + lane * CH;
    double *dst = out + (size_t)line * ld + lane * CH;
    if constexpr (CH >= 2) {
        double v[CH];
#pragma unroll
        for (int k = 0; k < CH / 2; k++) {
            const double2 x = reinterpret_cast<const double2 *>(src)[k];
            v[2 * k] = x.x;
            v[2 * k + 1] = x.y;
        }
#pragma unroll
        for (int k = 0; k < CH / 2; k++)
            reinterpret_cast<double2 *>(dst)[k] = make_double2(v[2 * k], v[2 * k + 1]);
    } else {
        dst[0] = src[0];
    }
}

// x pattern: line j is column j (stride ld); per lane CH cells of the column;
// column tiles dealt to the XCDs as the x sweeps deal them (xcd_tile)
template <int CH>
__global__ void floor_cols_kernel(const double *in, double *out, int nl, int n, int ld)
{
    const int line = xcd_tile(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (line >= nl || lane * CH >= n) return;
    const size_t base = (size_t)(lane * CH) * ld + line;
    double v[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) v[k] = in[base + (size_t)k * ld];
#pragma unroll
    for (int k = 0; k < CH; k++) out[base + (size_t)k * ld] = v[k];
}

template <int CH>
void launch_floor_ch(int dir, const double *in, double *out, int nl, int n, int ld,
                     dim3 grid, dim3 block, hipStream_t s)
{
    if (dir == DIR_Y)
        hipLaunchKernelGGL(floor_rows_kernel<CH>, grid, block, 0, s, in, out, nl, n, ld);
    else hipLaunchKernelGGL(floor_cols_kernel<CH>, grid, block, 0, s, in, out, nl, n, ld);
}

hipError_t launch_floor(int kind, int dir, const double *in, double *out, int n, int nl,
                        int ld, Tile t, hipStream_t s)
{
    // one wave per line: nl / nw blocks of nw waves (the direct one-field
    // sweeps' shape, W == NW lines per block)
    const int nw = t.nw;
    const dim3 grid((unsigned)((nl + nw - 1) / nw)), block(64 * nw);
    if (kind == 0) {
        hipLaunchKernelGGL(floor_empty_kernel, grid, block, 0, s, nullptr);
        return hipGetLastError();
    }
    switch (chunk_for(n)) {
    case 1: launch_floor_ch<1>(dir, in, out, nl, n, ld, grid, block, s); break;
    case 2: launch_floor_ch<2>(dir, in, out, nl, n, ld, grid, block, s); break;
    case 4: launch_floor_ch<4>(dir, in, out, nl, n, ld, grid, block, s); break;
    case 8: launch_floor_ch<8>(dir, in, out, nl, n, ld, grid, block, s); break;
    case 16: launch_floor_ch<16>(dir, in, out, nl, n, ld, grid, block, s); break;
    case 32: launch_floor_ch<32>(dir, in, out, nl, n, ld, grid, block, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void nonfinite_kernel(const double *src, size_t n, int *flag)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        if (!isfinite(src[i])) *flag = 1;
}

hipError_t launch_nonfinite_check(const double *src, size_t n, int *flag,
                                  hipStream_t s)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(nonfinite_kernel, dim3(blocks), dim3(256), 0, s, src, n,
                       flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// initial / exact cell averages (T2S4_Parareal.cu:118-140)
// ---------------------------------------------------------------------------
struct ExactArgs {
    int nx, ny, bc, kind, ng;
    double x_lo, y_lo, dx, dy, Lx, Ly, nu, t;
    double p[16];
};

__device__ double rot_gauss(const ExactArgs &e, double x, double y)
{
    double x0 = e.p[0], y0 = e.p[1], s = e.p[2], om = e.p[3], cx = e.p[4],
           cy = e.p[5], t = e.t;
    double co = cos(om * t), si = sin(om * t);
    double ax = x - cx, ay = y - cy;
    double xs = cx + (ax * co + ay * si);
    double ys = cy + (-ax * si + ay * co);
    double den = s + 4 * e.nu * t;
    double ex = xs - x0, ey = ys - y0;
    return s / den * exp(-(ex * ex + ey * ey) / den);
}

__device__ double gaussians(const ExactArgs &e, double x, double y)
{
    double u = e.p[0], v = e.p[1], nu = e.p[2], t = e.t, acc = 0.0;
    for (int k = 0; k < e.ng && k < 3; k++) {
        const double *g = e.p + 3 + 4 * k;
        double w2 = g[2] * g[2], s2 = w2 + 2.0 * nu * t;
        double ex = x - g[0] - u * t, ey = y - g[1] - v * t;
        acc += g[3] * (w2 / s2) * exp(-(ex * ex + ey * ey) / (2.0 * s2));
    }
    return acc;
}

__device__ double cosine_bell(const ExactArgs &e, double x, double y)
{
    double ex = x - e.p[0], ey = y - e.p[1];
    double r = sqrt(ex * ex + ey * ey), R = e.p[2];
    return r < R ? 0.5 * (1.0 + cos(M_PI * r / R)) : 0.0;
}

__device__ double point_value(const ExactArgs &e, double x, double y)
{
    int img = e.bc == BC_PER ? 1 : 0;
    double acc = 0.0;
    for (int ia = -img; ia <= img; ia++)
        for (int ib = -img; ib <= img; ib++) {
            double xx = x + ia * e.Lx, yy = y + ib * e.Ly;
            double v;
            if (e.kind == CCFV_IC_ROT_GAUSSIAN) v = rot_gauss(e, xx, yy);
            else if (e.kind == CCFV_IC_GAUSSIANS) v = gaussians(e, xx, yy);
            else v = cosine_bell(e, xx, yy);
            acc += v;
        }
    return acc;
}

__global__ void exact_kernel(ExactArgs e, double *out)
{
    size_t total = (size_t)e.nx * e.ny;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx / e.ny), j = (int)(idx % e.ny);
        double x = e.x_lo + (i + 0.5) * e.dx, y = e.y_lo + (j + 0.5) * e.dy;
        double xm = x - 0.5 * e.dx, xp = x + 0.5 * e.dx;
        double ym = y - 0.5 * e.dy, yp = y + 0.5 * e.dy;
        double v;
        if (e.kind == CCFV_IC_SINSIN) {
            const double tp = 2.0 * M_PI;
            double u = e.p[0], w = e.p[1], nu = e.p[2], t = e.t;
            double ax = (cos(tp * (xm - u * t)) - cos(tp * (xp - u * t))) / (tp * e.dx);
            double ay = (cos(tp * (ym - w * t)) - cos(tp * (yp - w * t))) / (tp * e.dy);
            v = ax * ay * exp(-2.0 * tp * tp * nu * t);
        } else {
            v = 1.0 / 36.0 *
                (point_value(e, xm, ym) + point_value(e, xp, ym) +
                 point_value(e, xm, yp) + point_value(e, xp, yp) +
                 4.0 * (point_value(e, xm, y) + point_value(e, xp, y) +
                        point_value(e, x, ym) + point_value(e, x, yp)) +
                 16.0 * point_value(e, x, y));
        }
        out[idx] = v;
    }
}

hipError_t launch_fill_exact(const ccfv_problem &p, const ccfv_ic &ic, double t,
                             double *out, hipStream_t s)
{
    ExactArgs e;
    e.nx = p.nx;
    e.ny = p.ny;
    e.bc = p.bc;
    e.kind = ic.kind;
    e.ng = ic.n;
    e.x_lo = p.x_lo;
    e.y_lo = p.y_lo;
    e.dx = (p.x_hi - p.x_lo) / p.nx;
    e.dy = (p.y_hi - p.y_lo) / p.ny;
    e.Lx = p.x_hi - p.x_lo;
    e.Ly = p.y_hi - p.y_lo;
    e.nu = p.nu;
    e.t = t;
    for (int k = 0; k < 16; k++) e.p[k] = ic.p[k];
    size_t total = (size_t)p.nx * p.ny;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(exact_kernel, dim3(blocks), dim3(256), 0, s, e, out);
    return hipGetLastError();
}

}  // namespace ccfv
