// ccfv_line.h — device building blocks of one line sweep, shared by the
// general and the uniform-displacement sweep kernels.
//
// A line is one x- or y-line of the field staged in LDS; one wavefront sweeps
// it with `CH` contiguous cells per lane.  The stages restate
// T2S4_Parareal.cu: feet (:203-258), PPM coefficients (:260-304), remap
// (:306-438), foot derivative (:440-547), RHS (:564-613), compact operator
// (:616-654) and the tridiagonal implicit half (:657-719).
#pragma once

#include "ccfv_feet.h"
#include "ccfv_internal.h"

namespace ccfv {

constexpr int BC_REF = CCFV_BC_REFERENCE;
constexpr int BC_PER = CCFV_BC_PERIODIC;

// LDS position of element e of a line: one pad slot every 8 elements keeps
// both the chunked (lane*CH + k) and the strided (lane + 64k) access
// patterns free of bank conflicts.
__device__ __forceinline__ int pidx(int e) { return e + (e >> 3); }
__host__ __device__ inline int line_pitch(int n) { return n + (n >> 3) + 1; }
// ghost cells after the last cell of a line in the uniform-offset sweep of
// lanes with ch >= 8 cells: a window of ch + 5 cells starting up to 7 cells
// into a chunk reaches (ch + 11) / 8 chunks past the lane's first
__host__ __device__ inline int fast_ghost_cells(int ch) { return ch >= 8 ? 8 * ((ch + 11) >> 3) : 0; }

__device__ __forceinline__ int clampi(int v, int lo, int hi)
{
    return v < lo ? lo : (v > hi ? hi : v);
}

struct Times {
    double hi, lo;
};

// sampling times of the velocity for the sub-step of field f of a launch:
// x sub-step 1 over [tn, tn+h/2], y over [tn, tn+h], x sub-step 3 over
// [tn+h/2, tn+h]; tn = t0 + n*DT (+ j*dt inside the fine propagator, :981)
__device__ __forceinline__ Times sub_times(const SweepArgs &a, int f)
{
#pragma clang fp contract(off)
    double tn = a.t0 + (double)(a.field0 + f) * a.tDT;
    double tl = tn + a.jdt;
    Times t;
    if (a.substep == 1) {
        t.lo = tl;
        t.hi = tl + 0.5 * a.h;
    } else if (a.substep == 3) {
        t.lo = tl + 0.5 * a.h;
        t.hi = tl + a.h;
    } else {
        t.lo = tl;
        t.hi = tl + a.h;
    }
    return t;
}

// sampling times of the first x sub-step of the next step in a fused x
// launch: the arithmetic of sub_times with t0_2, field0_2 and jdt2
__device__ __forceinline__ Times sub_times_next(const SweepArgs &a, int f)
{
#pragma clang fp contract(off)
    double tn = a.t0_2 + (double)(a.field0_2 + f) * a.tDT;
    double tl = tn + a.jdt2;
    Times t;
    t.lo = tl;
    t.hi = tl + 0.5 * a.h;
    return t;
}

// a line's corner coordinates in the other direction and, for the swirl,
// its velocity constants
struct LineGeo {
    double o0, o1;
    SwirlLine sw;
};

__device__ __forceinline__ LineGeo line_geo(const SweepArgs &a, const Times &t, int line)
{
    LineGeo g;
    g.o0 = corner(a.olo, a.od, line);
    g.o1 = corner(a.olo, a.od, line + 1);
    if (a.vel_kind == CCFV_VEL_SWIRL) {
        VelParams v{a.vel_kind, a.v0, a.v1, a.v2};
        g.sw = swirl_line(v, g.o0, g.o1, t.hi, t.lo);
    }
    return g;
}

// foot of face q of the line
template <int DIR>
__device__ __forceinline__ double face_foot(const SweepArgs &a, const Times &t,
                                            int q, const LineGeo &lg)
{
    if (a.vel_kind == CCFV_VEL_SWIRL) {
        if (a.trig)
            return face_foot_swirl_tab<DIR>(lg.sw, a.lo, a.d, a.tau, q, a.trig[q],
                                            a.trig[a.n + 1 + q], a.trig_small != 0);
        return face_foot_swirl<DIR>(lg.sw, a.lo, a.d, a.tau, q);
    }
    VelParams v{a.vel_kind, a.v0, a.v1, a.v2};
    return face_foot<DIR>(v, a.lo, a.d, a.tau, t.hi, t.lo, q, lg.o0, lg.o1);
}

// cell_of with the division by d as a multiplication by 1/d (equal when d
// is a power of two)
__device__ __forceinline__ int cell_of(const SweepArgs &a, double x, double &xi)
{
#pragma clang fp contract(off)
    double v = floor((x - a.lo) * a.id);
    v = fmin(fmax(v, -4.0 * a.n), 4.0 * a.n);
    const int m = (int)v;
    xi = (x - (a.lo + a.d * m)) * a.id;
    return m;
}

// ---------------------------------------------------------------------------
// reconstruction
// ---------------------------------------------------------------------------
// zero-slope cubic ghosts {Cm2, Cm1, CN, CNp1} (T2S4_Parareal.cu:275-286)
__device__ __forceinline__ void ghosts(const double *cs, int n, double g[4])
{
    const double e = 1.0 / 11.0;
    double c0 = cs[pidx(0)], c1 = cs[pidx(1)], c2 = cs[pidx(2)];
    double a0 = cs[pidx(n - 3)], a1 = cs[pidx(n - 2)], a2 = cs[pidx(n - 1)];
    g[2] = e * (-a0 + 3 * a1 + 9 * a2);
    g[3] = e * (-15 * a0 + 56 * a1 - 30 * a2);
    g[1] = e * (-c2 + 3 * c1 + 9 * c0);
    g[0] = e * (-15 * c2 + 56 * c1 - 30 * c0);
}

// five cell averages around cell m: E[m-2..m+2]; m in [0,n) (periodic: m is
// already wrapped; the closed box substitutes the ghosts)
template <int BC>
__device__ __forceinline__ void window(const double *cs, int n,
                                       const double g[4], int m, double w[5])
{
#pragma unroll
    for (int s = 0; s < 5; s++) {
        int k = m - 2 + s;
        if (BC == BC_PER) {
            k = k < 0 ? k + n : (k >= n ? k - n : k);
            w[s] = cs[pidx(k)];
        } else {
            double v = cs[pidx(clampi(k, 0, n - 1))];
            w[s] = k < 0 ? (k == -1 ? g[1] : g[0])
                         : (k >= n ? (k == n ? g[2] : g[3]) : v);
        }
    }
}

struct Ppm {
    double cl, dl, c6;
};

// PPM parabola of the centre cell of w (T2S4_Parareal.cu:296-300)
__device__ __forceinline__ Ppm ppm(const double w[5])
{
    const double s12 = 7.0 / 12.0, o12 = 1.0 / 12.0;
    double cl = s12 * (w[2] + w[1]) - o12 * (w[3] + w[0]);
    double cr = s12 * (w[3] + w[2]) - o12 * (w[4] + w[1]);
    Ppm p;
    p.cl = cl;
    p.dl = cr - cl;
    p.c6 = 6.0 * (w[2] - 0.5 * (cl + cr));
    return p;
}

// antiderivative of the parabola (T2S4_Parareal.cu:63-65)
__device__ __forceinline__ double intc(const Ppm &p, double xi)
{
    return -1.0 / 3 * p.c6 * (xi * xi * xi) + 0.5 * (p.dl + p.c6) * (xi * xi) +
           p.cl * xi;
}

// derivative of the quartic through the five averages (T2S4_Parareal.cu:481-486)
__device__ __forceinline__ double deriv(const double w[5], double xi,
                                        double inv12d)
{
    double x2 = xi * xi, x3 = x2 * xi;
    return inv12d * ((2.0 * x3 - 6.0 * x2 + 3.0 * xi + 1.0) * w[0] +
                     (-8.0 * x3 + 18.0 * x2 + 6.0 * xi - 15.0) * w[1] +
                     (12.0 * x3 - 18.0 * x2 - 24.0 * xi + 15.0) * w[2] +
                     (-8.0 * x3 + 6.0 * x2 + 18.0 * xi - 1.0) * w[3] +
                     (2.0 * x3 - 3.0 * xi) * w[4]);
}

// k mod n in [0, n); feet within one period of the line need no division
__device__ __forceinline__ int wrapn(int k, int n)
{
    if (k < -n || k >= 2 * n) {
        k %= n;
        return k < 0 ? k + n : k;
    }
    return k < 0 ? k + n : (k >= n ? k - n : k);
}

struct Face {
    double xb;        // raw foot
    double xiL;       // offset of the (left-clipped) foot in its cell
    int mL, mR;       // cells of the foot as left / right end of a cell
    double ivL, icL;  // I(xi), I(1)-I(xi) as left end
    double ivR;       // I(xi) as right end
    double D;         // derivative at the foot
};

// ---------------------------------------------------------------------------
// source term f(x, y, t) (CCFV_SRC_*; T2S4_Parareal.cu:59-61, :143-165)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double src_point(const SweepArgs &a, double x, double y, double t)
{
#pragma clang fp contract(off)
    if (a.src_kind == CCFV_SRC_CONSTANT) return a.src[0];
    const double tp = 2.0 * M_PI, w = a.src[0];
    const double sx = sin(tp * x), sy = sin(tp * y);
    const double cx = cos(tp * x), cy = cos(tp * y);
    const double ct = cos(w * t), st = sin(w * t);
    return (a.src[1] * ct + a.src[2] * st) * sx * sy +
           st * (a.src[3] * cx * sy + a.src[4] * sx * cy);
}

// cell average of f over cell q of line L by the 3x3 Simpson rule, summed
// as T2S4_Parareal.cu:155-160 (x: along-line coordinate for DIR 0)
template <int DIR>
__device__ __forceinline__ double src_avg(const SweepArgs &a, int q, int L, double t)
{
#pragma clang fp contract(off)
    const double s = a.lo + (q + 0.5) * a.d, o = a.olo + (L + 0.5) * a.od;
    const double x = DIR == 0 ? s : o, y = DIR == 0 ? o : s;
    const double dx = DIR == 0 ? a.d : a.od, dy = DIR == 0 ? a.od : a.d;
    const double xm = x - 0.5 * dx, xp = x + 0.5 * dx;
    const double ym = y - 0.5 * dy, yp = y + 0.5 * dy;
    return 1.0 / 36.0 *
           (src_point(a, xm, ym, t) + src_point(a, xp, ym, t) + src_point(a, xm, yp, t) +
            src_point(a, xp, yp, t) +
            4.0 * (src_point(a, xm, y, t) + src_point(a, xp, y, t) + src_point(a, x, ym, t) +
                   src_point(a, x, yp, t)) +
            16.0 * src_point(a, x, y, t));
}

// times of the source's earlier (remapped) and later cell averages of a
// sub-step: x sub-step 1 [tn, tn+h/2], y [tn, tn+h], x sub-step 3 both tn+h
// (T2S4_Parareal.cu:839-845, :887-892, :935-940)
struct SrcTimes {
    double t0, t1;
};

__device__ __forceinline__ SrcTimes src_times(const SweepArgs &a, int f, bool next)
{
#pragma clang fp contract(off)
    const double tn = next ? a.t0_2 + (double)(a.field0_2 + f) * a.tDT + a.jdt2
                           : a.t0 + (double)(a.field0 + f) * a.tDT + a.jdt;
    const int sub = next ? 1 : a.substep;
    SrcTimes t;
    if (sub == 1) {
        t.t0 = tn;
        t.t1 = tn + 0.5 * a.h;
    } else if (sub == 3) {
        t.t0 = tn + a.h;
        t.t1 = tn + a.h;
    } else {
        t.t0 = tn;
        t.t1 = tn + a.h;
    }
    return t;
}

// Closed box foot derivative: zero outside the box, else clamp to
// [lo+1e-15, hi-1e-15] (T2S4_Parareal.cu:457-490, :514-543).
template <int BC>
__device__ __forceinline__ double deriv_ref(const SweepArgs &a,
                                            const double *src,
                                            const double g[4], double xb,
                                            double inv12d)
{
#pragma clang fp contract(off)
    if (!(xb >= a.lo && xb <= a.hi)) return 0.0;
    double xd = fmin(fmax(xb, a.lo + 1e-15), a.hi - 1e-15);
    double xi;
    int m = cell_of(a, xd, xi);
    int md = clampi(m, 0, a.n - 1);
    if (md != m) xi = (xd - (a.lo + a.d * md)) * a.id;
    double w[5];
    window<BC_REF>(src, a.n, g, md, w);
    if (a.quirk) {  // :476-479 — iL==Nx-2 takes CNp1, iL==1 takes Cm2
        if (md > a.n - 3) w[4] = g[3];
        if (md < 2) w[0] = g[0];
    }
    return deriv(w, xi, inv12d);
}

template <int DIR, int BC, bool POST>
__device__ __forceinline__ Face eval_face(const SweepArgs &a, const Times &tt,
                                          const double *cs, const double g[4],
                                          int q, const LineGeo &lg,
                                          double inv12d)
{
    Face f;
    const int n = a.n;
    if (BC == BC_PER) {
        // face n is face 0 one period on (exact periodicity of the fluxes)
        int qq = q == n ? 0 : q;
        f.xb = face_foot<DIR>(a, tt, qq, lg);
        double xi;
        int m = cell_of(a, f.xb, xi);
        int mw = wrapn(m, n);
        double w[5];
        window<BC_PER>(cs, n, g, mw, w);
        Ppm p = ppm(w);
        double iv = intc(p, xi);
        f.ivL = f.ivR = iv;
        f.icL = intc(p, 1.0) - iv;
        f.xiL = xi;
        f.mL = f.mR = q == n ? m + n : m;
        f.D = POST ? 0.0 : deriv(w, xi, inv12d);
    } else {
        f.xb = face_foot<DIR>(a, tt, q, lg);
        double xl, xr;
        {
#pragma clang fp contract(off)
            xl = fmax(f.xb, a.lo + a.eps);
            xr = fmin(f.xb, a.hi - a.eps);
        }
        double xi;
        int m = clampi(cell_of(a, xl, xi), 0, n - 1);
        {
#pragma clang fp contract(off)
            xi = (xl - (a.lo + a.d * m)) * a.id;
        }
        double w[5];
        window<BC_REF>(cs, n, g, m, w);
        Ppm p = ppm(w);
        f.mL = m;
        f.xiL = xi;
        f.ivL = intc(p, xi);
        f.icL = intc(p, 1.0) - f.ivL;
        if (xr == xl) {
            f.mR = m;
            f.ivR = f.ivL;
        } else {
            double xi2;
            int m2 = clampi(cell_of(a, xr, xi2), 0, n - 1);
            {
#pragma clang fp contract(off)
                xi2 = (xr - (a.lo + a.d * m2)) * a.id;
            }
            double w2[5];
            window<BC_REF>(cs, n, g, m2, w2);
            Ppm p2 = ppm(w2);
            f.mR = m2;
            f.ivR = intc(p2, xi2);
        }
        f.D = POST ? 0.0 : deriv_ref<BC>(a, cs, g, f.xb, inv12d);
    }
    return f;
}

// conservative remap of one cell between its feet (T2S4_Parareal.cu:319-344)
template <int BC>
__device__ __forceinline__ double remap_cell(const SweepArgs &a,
                                             const double *cs, const Face &L,
                                             const Face &R, bool &bad)
{
    const int n = a.n;
    if (BC == BC_REF) {
        if (R.xb < L.xb) bad = true;
        if (L.xb > a.hi || R.xb < a.lo || L.xb == R.xb) return 0.0;
        if (R.mR == L.mL) return R.ivR - L.ivL;
        double s = 0.0;
        for (int m = L.mL + 1; m <= R.mR - 1; m++) s += cs[pidx(m)];
        return L.icL + s + R.ivR;
    } else {
        int mL = L.mL, mR = R.mR;
        if (mR < mL || (mR == mL && R.xiL < L.xiL) || mR - mL > n) {
            bad = true;
            return 0.0;
        }
        if (mR == mL) return R.ivR - L.ivL;
        int w = wrapn(mL + 1, n);
        double s = 0.0;
        for (int m = mL + 1; m <= mR - 1; m++) {
            s += cs[pidx(w)];
            w = w + 1 == n ? 0 : w + 1;
        }
        return L.icL + s + R.ivR;
    }
}


// RHS b of the cells of this lane by the general per-face path; b goes to the
// lane's own slots of `aux` (an LDS line used as an indexable register file).
// The source's earlier cell averages remapped over the same feet, plus its
// later ones, weighted by theta*h: the source part of the RHS of cell q.
// fs: that line's earlier cell averages in LDS, fg its ghosts.
template <int DIR, int BC>
__device__ __forceinline__ double src_rhs(const SweepArgs &a, const Times &tt,
                                          const double *fs, const double fg[4],
                                          const LineGeo &lg, int line, int q,
                                          const Face &gl, const Face &gr, double t1,
                                          bool &bad)
{
    (void)tt;
    (void)lg;
    const double fbar = remap_cell<BC>(a, fs, gl, gr, bad);
    return a.src_w * (fbar + src_avg<DIR>(a, q, line, t1));
}

template <int DIR, int BC, bool SRC>
__device__ void rhs_general(const SweepArgs &a, const Times &tt, const double *cs,
                            const double g[4], double *aux, int base, int ch,
                            const LineGeo &lg, double inv12d, bool post,
                            const double *fs = nullptr, const double *fg = nullptr,
                            int line = 0, double t1 = 0.0)
{
    bool bad = false;
    Face fl = post ? eval_face<DIR, BC, true>(a, tt, cs, g, base, lg, inv12d)
                   : eval_face<DIR, BC, false>(a, tt, cs, g, base, lg, inv12d);
    // with the third-sub-step quirk the source joins after the derivative
    // of the remapped line (the caller adds it)
    const bool src = SRC && fs && !post;
    Face gl;
    if (src) gl = eval_face<DIR, BC, true>(a, tt, fs, fg, base, lg, inv12d);
#pragma unroll 1
    for (int k = 0; k < ch; k++) {
        Face fr = post ? eval_face<DIR, BC, true>(a, tt, cs, g, base + k + 1, lg, inv12d)
                       : eval_face<DIR, BC, false>(a, tt, cs, g, base + k + 1, lg, inv12d);
        double cbar = remap_cell<BC>(a, cs, fl, fr, bad);
        if (src) {
            // b = C + theta h (fbar0 + f1) + ... (T2S4_Parareal.cu:575, :607)
            Face gr = eval_face<DIR, BC, true>(a, tt, fs, fg, base + k + 1, lg, inv12d);
            cbar = cbar + src_rhs<DIR, BC>(a, tt, fs, fg, lg, line, base + k, gl, gr, t1, bad);
            gl = gr;
        }
        aux[pidx(base + k)] = post ? cbar : cbar + a.coef * (fr.D - fl.D);
        fl = fr;
    }
    if (bad) *a.flag = 1;
}

// The same RHS without the third-sub-step variant, straight into the lane's
// registers: b[k] for its CH cells.  The cells are visited in a loop (the
// face evaluation is too large to unroll CH times) that writes with a
// wave-uniform runtime index, in parts of at most 16 cells with an array of
// their own each: arrays of up to 16 doubles indexed that way stay in VGPRs
// (32 go to scratch).  Same arithmetic, in the same order, as rhs_general.
template <int DIR, int BC, bool SRC, int CH>
__device__ void rhs_general_reg(const SweepArgs &a, const Times &tt, const double *cs,
                                const double g[4], double (&b)[CH], int base,
                                const LineGeo &lg, double inv12d, const double *fs,
                                const double *fg, int line, double t1)
{
    constexpr int P = CH < 16 ? CH : 16, NP = CH / P;
    bool bad = false;
    Face fl = eval_face<DIR, BC, false>(a, tt, cs, g, base, lg, inv12d);
    const bool src = SRC && fs;
    Face gl;
    if (src) gl = eval_face<DIR, BC, true>(a, tt, fs, fg, base, lg, inv12d);
#pragma unroll
    for (int c = 0; c < NP; c++) {
        double part[P];
#pragma unroll 2
        for (int k = 0; k < P; k++) {
            const int e = base + c * P + k;
            Face fr = eval_face<DIR, BC, false>(a, tt, cs, g, e + 1, lg, inv12d);
            double cbar = remap_cell<BC>(a, cs, fl, fr, bad);
            if (src) {
                Face gr = eval_face<DIR, BC, true>(a, tt, fs, fg, e + 1, lg, inv12d);
                cbar = cbar + src_rhs<DIR, BC>(a, tt, fs, fg, lg, line, e, gl, gr, t1, bad);
                gl = gr;
            }
            part[k] = cbar + a.coef * (fr.D - fl.D);
            fl = fr;
        }
#pragma unroll
        for (int k = 0; k < P; k++) b[c * P + k] = part[k];
    }
    if (bad) *a.flag = 1;
}

// ---------------------------------------------------------------------------
// block reduction of two doubles into partial[2*slot ...]
// ---------------------------------------------------------------------------
__device__ inline void block_reduce2(double s0, double s1, double *partial, int slot)
{
    __shared__ double red[2][16];
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off);
        s1 += __shfl_down(s1, off);
    }
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[0][wave] = s0;
        red[1][wave] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        int nw = blockDim.x >> 6;
        for (int w = 0; w < nw; w++) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        partial[2 * slot] = t0;
        partial[2 * slot + 1] = t1;
    }
}


// item of block b in a ragged launch: the last i with prefix[i] <= b, so
// prefix[i] <= b < prefix[i+1] (prefix[0] = 0, non-decreasing).  Every
// wave searches on its own, 64 entries per round: the lanes test evenly
// spaced candidates of [lo, hi] at once and the ballot keeps the last one
// that is <= b, so 64 items take one dependent load and 4096 two (a binary
// search takes log2(nitems)).  Call with all 64 lanes of the wave active.
__device__ __forceinline__ int ragged_item(const int *prefix, int nitems, int b)
{
    const int lane = threadIdx.x & 63;
    int lo = 0, hi = nitems - 1;  // prefix[lo] <= b; the answer is in [lo, hi]
    while (hi > lo) {
        const int step = (hi - lo + 64) >> 6;  // ceil((hi - lo + 1) / 64)
        const int k = lo + lane * step;
        // candidates past hi fail; their address is clamped so that no lane
        // reads past the prefix array
        const bool ok = k <= hi && prefix[min(k, hi)] <= b;
        // the candidates are ascending and prefix is non-decreasing: the lanes
        // that pass are 0 .. last (lane 0 passes: k = lo)
        const unsigned long long m = __ballot(ok);
        const int last = 63 - __clzll(m);
        lo += last * step;
        hi = min(hi, lo + step - 1);
    }
    return lo;
}

// ---------------------------------------------------------------------------
// compact operator and implicit half, shared by both paths
// ---------------------------------------------------------------------------
// b: RHS of this lane's CH cells.  Tables: mneg = -a*den, den = Thomas
// pivots' inverses over 12 (the compact operator's 1/12 taken with them),
// pneg = -p, z = T^{-1} u (Sherman-Morrison), mf/mb = per-lane chunk
// products of mneg/pneg (identity 1 on idle lanes).  LDSTAB: mneg..z are LDS
// lines (pidx layout) rather than global arrays.  The result overwrites the
// lane's cells of cs.
template <int BC, int CH, bool LDSTAB>
__device__ __forceinline__ void solve_tail(const SweepArgs &a, double (&b)[CH],
                                           int lane, bool active,
                                           const double *mneg, const double *den,
                                           const double *pneg, const double *zt,
                                           const double *mf, const double *mb,
                                           double *cs)
{
    const int n = a.n, NL = n / CH, base = lane * CH;
#define CCFV_TAB(t, e) (LDSTAB ? (t)[pidx(e)] : (t)[e])
    // (b_{i-1} + 10 b_i + b_{i+1}) / 12, neighbours across lanes by shuffle;
    // zero ghosts in the closed box (:624-631), wrap-around when periodic
    double bprev = __shfl_up(b[CH - 1], 1);
    double bnext = __shfl_down(b[0], 1);
    double wrap_last = __shfl(b[CH - 1], NL - 1);
    double wrap_first = __shfl(b[0], 0);
    if (lane == 0) bprev = BC == BC_PER ? wrap_last : 0.0;
    if (lane == NL - 1) bnext = BC == BC_PER ? wrap_first : 0.0;
    double r[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        double bn = k + 1 < CH ? b[k + 1] : bnext;
        double bp = k > 0 ? b[k - 1] : bprev;
        r[k] = bn + 10.0 * b[k] + bp;  // the 1/12 is in den
    }

    // forward pass q_e = r_e den_e + mneg_e q_{e-1}: chunk pass, lane scan
    // (a.sf doubling steps reach every factor product above 1e-18), replay
    double M = mf[lane], Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            int e = base + k;
            double c = r[k] * CCFV_TAB(den, e);
            r[k] = c;
            Q = c + CCFV_TAB(mneg, e) * Q;
        }
    }
    for (int st = 0, off = 1; st < a.sf; st++, off <<= 1) {
        double Mp = __shfl_up(M, off), Qp = __shfl_up(Q, off);
        if (lane >= off) {
            Q = M * Qp + Q;
            M = M * Mp;
        }
    }
    double carry = __shfl_up(Q, 1);
    if (lane == 0) carry = 0.0;
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            carry = r[k] + CCFV_TAB(mneg, base + k) * carry;
            r[k] = carry;
        }
    }

    // back substitution x_e = q_e + pneg_e x_{e+1}
    M = mb[lane];
    Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) Q = r[k] + CCFV_TAB(pneg, base + k) * Q;
    }
    for (int st = 0, off = 1; st < a.sb; st++, off <<= 1) {
        double Mn = __shfl_down(M, off), Qn = __shfl_down(Q, off);
        if (lane + off < 64) {
            Q = M * Qn + Q;
            M = M * Mn;
        }
    }
    carry = __shfl_down(Q, 1);
    if (lane == 63) carry = 0.0;
    if (active) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) {
            carry = r[k] + CCFV_TAB(pneg, base + k) * carry;
            r[k] = carry;
        }
    }

    if (BC == BC_PER) {
        // Sherman-Morrison: x = y - ((y_0 + vl y_{n-1}) / (1 + v.z)) z
        double y0 = __shfl(r[0], 0);
        double yn = __shfl(r[CH - 1], NL - 1);
        double fac = (y0 + a.vl * yn) * a.smrcp;
        if (active) {
#pragma unroll
            for (int k = 0; k < CH; k++) r[k] = r[k] - fac * CCFV_TAB(zt, base + k);
        }
    }
#undef CCFV_TAB
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) cs[pidx(base + k)] = r[k];
    }
}

// ---------------------------------------------------------------------------
// tiles: global <-> LDS
// ---------------------------------------------------------------------------
// x lines are columns j0..j0+W-1 of the row-major field: rows of W doubles
// are read and written coalesced (double2) and transposed through LDS.
// W is a power of two (enforced by the host), so idx splits by shifts.
__device__ __forceinline__ int log2i(int v) { return 31 - __builtin_clz(v); }

// Column tile of an x-sweep block.  Workgroups are dealt round-robin to the
// eight XCDs, each with its own L2; give every XCD a contiguous range of
// column tiles so that the 128-byte lines its tiles touch are shared there
// instead of being fetched by several XCDs.
__device__ __forceinline__ int xcd_tile(int bx, int gx)
{
    if (gx & 7) return bx;
    return (bx & 7) * (gx >> 3) + (bx >> 3);
}

// The same for the blocks of one item of a ragged x launch: block `local` of
// an item whose nblk blocks start at launch block `first` sweeps the item's
// tile number (local mod 8) * nblk/8 + local/8 (tile number = field * tiles
// per field + tile), so each XCD takes a contiguous run of the item's
// column tiles.  Applies when first and nblk are multiples of 8 (then block
// local sits on XCD local mod 8); a bijection of [0, nblk) either way.
__device__ __forceinline__ int ragged_xcd_tile(int local, int first, int nblk)
{
    if ((first | nblk) & 7) return local;
    return (local & 7) * (nblk >> 3) + (local >> 3);
}

// Loads of a tile are issued TILE_UNROLL at a time per work-item before any
// of them is written to LDS, so each wave keeps several requests in flight.
constexpr int TILE_UNROLL = 8;

__device__ __forceinline__ void load_tile_x(const double *in, int ld, int n,
                                            int j0, int W, double *lds, int LP)
{
    const int hw = W >> 1, sh = log2i(hw);
    const int R = (int)blockDim.x >> sh;  // rows per pass of the block
    if ((R & 7) == 0 && n % R == 0) {
        // Each work-item keeps its column pair and moves rows i0, i0+R, ...:
        // global address = uniform row offset + fixed item offset, LDS
        // address = item base + (R + R/8) per pass (pidx of R rows later).
        const int c2 = threadIdx.x & (hw - 1), i0 = threadIdx.x >> sh;
        const unsigned lo = (unsigned)(i0 * ld + 2 * c2);
        double *d0 = lds + 2 * c2 * LP + pidx(i0);
        double *d1 = d0 + LP;
        const int nr = n / R, ls = R + (R >> 3);
        for (int p = 0; p < nr; p += TILE_UNROLL) {
            double2 v[TILE_UNROLL];
#pragma unroll
            for (int u = 0; u < TILE_UNROLL; u++)
                if (p + u < nr) {
                    const double *row = in + j0 + (size_t)(p + u) * R * ld;
                    v[u] = *reinterpret_cast<const double2 *>(row + lo);
                }
#pragma unroll
            for (int u = 0; u < TILE_UNROLL; u++)
                if (p + u < nr) {
                    d0[(p + u) * ls] = v[u].x;
                    d1[(p + u) * ls] = v[u].y;
                }
        }
        return;
    }
    const int total = hw * n, step = blockDim.x;
    for (int b0 = threadIdx.x; b0 < total; b0 += TILE_UNROLL * step) {
        double2 v[TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int idx = b0 + u * step;
            if (idx < total) {
                const int c2 = idx & (hw - 1), i = idx >> sh;
                v[u] = *reinterpret_cast<const double2 *>(in + (size_t)i * ld + j0 + 2 * c2);
            }
        }
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int idx = b0 + u * step;
            if (idx < total) {
                const int c2 = idx & (hw - 1), i = idx >> sh;
                lds[(2 * c2) * LP + pidx(i)] = v[u].x;
                lds[(2 * c2 + 1) * LP + pidx(i)] = v[u].y;
            }
        }
    }
}

__device__ __forceinline__ void store_tile_x(double *out, int ld, int n, int j0,
                                             int W, const double *lds, int LP)
{
    const int hw = W >> 1, sh = log2i(hw);
    const int R = (int)blockDim.x >> sh;
    if ((R & 7) == 0 && n % R == 0) {
        // the addressing of load_tile_x
        const int c2 = threadIdx.x & (hw - 1), i0 = threadIdx.x >> sh;
        const unsigned lo = (unsigned)(i0 * ld + 2 * c2);
        const double *d0 = lds + 2 * c2 * LP + pidx(i0);
        const double *d1 = d0 + LP;
        const int nr = n / R, ls = R + (R >> 3);
        for (int p = 0; p < nr; p++) {
            double2 v;
            v.x = d0[p * ls];
            v.y = d1[p * ls];
            double *row = out + j0 + (size_t)p * R * ld;
            *reinterpret_cast<double2 *>(row + lo) = v;
        }
        return;
    }
    for (int idx = threadIdx.x; idx < hw * n; idx += blockDim.x) {
        int c2 = idx & (hw - 1), i = idx >> sh;
        double2 v;
        v.x = lds[(2 * c2) * LP + pidx(i)];
        v.y = lds[(2 * c2 + 1) * LP + pidx(i)];
        *reinterpret_cast<double2 *>(out + (size_t)i * ld + j0 + 2 * c2) = v;
    }
}

// Parareal epilogue of the last x sweep of a coarse step: g is G(U_n^{k+1}).
//   EPI_INIT   : G cache <- g, U[n+1] <- g
//   EPI_CORRECT: U[n+1] <- g + (F - G_old) (T2S4_Parareal.cu:779),
//                G cache <- g, block partials of |dU|^2 and |U|^2
// The correction's three inputs for the first EPI_UNROLL elements of each
// work-item, loaded before the sweep (they do not depend on it) so that the
// epilogue of a one-field launch does not wait for memory.
constexpr int EPI_UNROLL = 4;
struct EpiInputs {
    double2 go[EPI_UNROLL], fv[EPI_UNROLL], uo[EPI_UNROLL];
};

__device__ __forceinline__ void epi_load(const SweepArgs &a, int n, int j0, int W, int b0,
                                         EpiInputs &in)
{
    const int hw = W >> 1, sh = log2i(hw);
    const int total = hw * n, step = blockDim.x;
#pragma unroll
    for (int u = 0; u < EPI_UNROLL; u++) {
        const int idx = b0 + u * step;
        if (idx < total) {
            const int c2 = idx & (hw - 1), i = idx >> sh;
            const size_t gi = (size_t)i * a.ld + j0 + 2 * c2;
            in.go[u] = *reinterpret_cast<const double2 *>(a.g_cache + gi);
            in.fv[u] = *reinterpret_cast<const double2 *>(a.f_in + gi);
            in.uo[u] = *reinterpret_cast<const double2 *>(a.u_next + gi);
        }
    }
}

// WB (fused launches): the corrected U[n+1] also replaces the tile in LDS,
// where the next step's first x sub-step reads it.  pre: the inputs of the
// first chunk, loaded by epi_load(..., threadIdx.x, ...) (EPI_CORRECT).
template <int EPI, bool WB = false>
__device__ __forceinline__ void epilogue_x(const SweepArgs &a, int n, int j0,
                                           int W, double *lds, int LP,
                                           int slot, const EpiInputs *pre = nullptr)
{
    const int hw = W >> 1, sh = log2i(hw);
    const int total = hw * n, step = blockDim.x;
    constexpr int U = EPI_UNROLL;
    double s0 = 0.0, s1 = 0.0;
    for (int b0 = threadIdx.x; b0 < total; b0 += U * step) {
        EpiInputs in;
        if (EPI == EPI_CORRECT) {
            if (pre && b0 == (int)threadIdx.x) in = *pre;
            else epi_load(a, n, j0, W, b0, in);
        }
        const double2 *go = in.go, *fv = in.fv, *uo = in.uo;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int idx = b0 + u * step;
            if (idx >= total) continue;
            const int c2 = idx & (hw - 1), i = idx >> sh;
            const size_t gi = (size_t)i * a.ld + j0 + 2 * c2;
            double2 g;
            g.x = lds[(2 * c2) * LP + pidx(i)];
            g.y = lds[(2 * c2 + 1) * LP + pidx(i)];
            if (EPI == EPI_INIT) {
                *reinterpret_cast<double2 *>(a.g_cache + gi) = g;
                *reinterpret_cast<double2 *>(a.u_next + gi) = g;
            } else {
                double2 un;
                un.x = g.x + (fv[u].x - go[u].x);
                un.y = g.y + (fv[u].y - go[u].y);
                *reinterpret_cast<double2 *>(a.g_cache + gi) = g;
                *reinterpret_cast<double2 *>(a.u_next + gi) = un;
                if (WB) {
                    lds[(2 * c2) * LP + pidx(i)] = un.x;
                    lds[(2 * c2 + 1) * LP + pidx(i)] = un.y;
                }
                double d0 = un.x - uo[u].x, d1 = un.y - uo[u].y;
                s0 += d0 * d0 + d1 * d1;
                s1 += un.x * un.x + un.y * un.y;
            }
        }
    }
    if (EPI == EPI_CORRECT) block_reduce2(s0, s1, a.partial, slot);
}

// y lines are rows i0..i0+W-1 (contiguous)
// row index and offset of element idx of a W x hn tile (shifts when hn is a
// power of two)
__device__ __forceinline__ void split_row(int idx, int hn, int &r, int &j2)
{
    if ((hn & (hn - 1)) == 0) {
        int sh = log2i(hn);
        r = idx >> sh;
        j2 = idx & (hn - 1);
    } else {
        r = idx / hn;
        j2 = idx - r * hn;
    }
}

__device__ __forceinline__ void load_tile_y(const double *in, int ld, int n,
                                            int W, double *lds, int LP)
{
    const int hn = n >> 1;
    const int total = W * hn, step = blockDim.x;
    if (hn % step == 0) {
        // Work-item t moves double2 t, t+step, ... of every line: global
        // address = uniform (line, pass) offset + 2t, LDS address = pidx(2t)
        // + uniform offset (2 step + step/4 per pass: step is a multiple of 64).
        const int QP = hn / step, tot = W * QP, t = threadIdx.x;
        double *dl = lds + pidx(2 * t);
        const int qs = 2 * step + (step >> 2);
        for (int e0 = 0; e0 < tot; e0 += TILE_UNROLL) {
            const int l0 = e0 / QP, q0 = e0 - l0 * QP;
            double2 v[TILE_UNROLL];
            int l = l0, q = q0;
#pragma unroll
            for (int u = 0; u < TILE_UNROLL; u++)
                if (e0 + u < tot) {
                    const double *row = in + (size_t)l * ld + 2 * q * step;
                    v[u] = *reinterpret_cast<const double2 *>(row + 2 * t);
                    if (++q == QP) {
                        q = 0;
                        ++l;
                    }
                }
            l = l0;
            q = q0;
#pragma unroll
            for (int u = 0; u < TILE_UNROLL; u++)
                if (e0 + u < tot) {
                    double *d = dl + l * LP + q * qs;
                    d[0] = v[u].x;
                    d[1] = v[u].y;
                    if (++q == QP) {
                        q = 0;
                        ++l;
                    }
                }
        }
        return;
    }
    for (int b0 = threadIdx.x; b0 < total; b0 += TILE_UNROLL * step) {
        double2 v[TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int idx = b0 + u * step;
            if (idx < total) {
                int r, j2;
                split_row(idx, hn, r, j2);
                v[u] = *reinterpret_cast<const double2 *>(in + (size_t)r * ld + 2 * j2);
            }
        }
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int idx = b0 + u * step;
            if (idx < total) {
                int r, j2;
                split_row(idx, hn, r, j2);
                lds[r * LP + pidx(2 * j2)] = v[u].x;
                lds[r * LP + pidx(2 * j2 + 1)] = v[u].y;
            }
        }
    }
}

__device__ __forceinline__ void store_tile_y(double *out, int ld, int n, int W,
                                             const double *lds, int LP)
{
    const int hn = n >> 1;
    const int step = blockDim.x;
    if (hn % step == 0) {
        // the addressing of load_tile_y
        const int QP = hn / step, t = threadIdx.x;
        const double *dl = lds + pidx(2 * t);
        const int qs = 2 * step + (step >> 2);
        for (int l = 0; l < W; l++)
            for (int q = 0; q < QP; q++) {
                const double *d = dl + l * LP + q * qs;
                double2 v;
                v.x = d[0];
                v.y = d[1];
                double *row = out + (size_t)l * ld + 2 * q * step;
                *reinterpret_cast<double2 *>(row + 2 * t) = v;
            }
        return;
    }
    for (int idx = threadIdx.x; idx < W * hn; idx += blockDim.x) {
        int r, j2;
        split_row(idx, hn, r, j2);
        double2 v;
        v.x = lds[r * LP + pidx(2 * j2)];
        v.y = lds[r * LP + pidx(2 * j2 + 1)];
        *reinterpret_cast<double2 *>(out + (size_t)r * ld + 2 * j2) = v;
    }
}

// ---------------------------------------------------------------------------
// panel layout (SweepArgs::in_panel / out_panel): column panels of PANEL_W
// columns, each stored row by row, so that the x tile of lines j0..j0+7 is
// one contiguous block of 8 n doubles and a y tile of W rows reads 8 W
// contiguous doubles (whole 128-byte lines for W >= 2) from every panel
// ---------------------------------------------------------------------------
// x tiles (W == PANEL_W, j0 a multiple of PANEL_W): the block of tile j0
// starts at j0 * n and holds row i at PANEL_W * i, which is the row-major
// addressing of load_tile_x / store_tile_x with row length PANEL_W and
// base f + j0 * (n - 1)
__device__ __forceinline__ const double *panel_x_base(const double *f, int n, int j0)
{
    return f + (size_t)j0 * (size_t)(n - 1);
}
__device__ __forceinline__ double *panel_x_base(double *f, int n, int j0)
{
    return f + (size_t)j0 * (size_t)(n - 1);
}

// y tiles of W rows i0.. (W a power of two, n a multiple of PANEL_W): ps =
// PANEL_W * rows of the field.  double2 q of the tile is pair e = q mod 4W
// of panel p = q / 4W: row r = e / 4, columns 2 (e mod 4) and +1, at
// p ps + PANEL_W (i0 + r) + 2 (e mod 4); in LDS line r, pidx(8 p + c) =
// 9 p + c.
__device__ __forceinline__ void load_tile_y_panel(const double *f, size_t ps, int n, int i0,
                                                  int W, double *lds, int LP)
{
    static_assert(PANEL_W == 8, "pidx chunks are panels");
    const int sw = log2i(W) + 2, em = (1 << sw) - 1;
    const int total = W * (n >> 1), step = blockDim.x;
    const double *base = f + (size_t)PANEL_W * i0;
    for (int b0 = threadIdx.x; b0 < total; b0 += TILE_UNROLL * step) {
        double2 v[TILE_UNROLL];
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int q = b0 + u * step;
            if (q < total)
                v[u] = *reinterpret_cast<const double2 *>(base + (size_t)(q >> sw) * ps +
                                                          2 * (q & em));
        }
#pragma unroll
        for (int u = 0; u < TILE_UNROLL; u++) {
            const int q = b0 + u * step;
            if (q < total) {
                const int e = q & em;
                double *d = lds + (e >> 2) * LP + 9 * (q >> sw) + 2 * (e & 3);
                d[0] = v[u].x;
                d[1] = v[u].y;
            }
        }
    }
}

__device__ __forceinline__ void store_tile_y_panel(double *f, size_t ps, int n, int i0,
                                                   int W, const double *lds, int LP)
{
    const int sw = log2i(W) + 2, em = (1 << sw) - 1;
    const int total = W * (n >> 1);
    double *base = f + (size_t)PANEL_W * i0;
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        const int e = q & em;
        const double *d = lds + (e >> 2) * LP + 9 * (q >> sw) + 2 * (e & 3);
        *reinterpret_cast<double2 *>(base + (size_t)(q >> sw) * ps + 2 * e) =
            make_double2(d[0], d[1]);
    }
}

}  // namespace ccfv
