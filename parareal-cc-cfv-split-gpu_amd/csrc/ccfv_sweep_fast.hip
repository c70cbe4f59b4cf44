// ccfv_sweep_fast.hip — line sweeps for a uniform foot offset per line.
//
// When the velocity component along a line does not depend on the
// along-line coordinate (solid-body rotation, constant velocity) and the
// boundary is periodic, every face foot of a line sits at the same offset xi
// in cell q+S.  The host locates face 0 of every line with the same
// contraction-free arithmetic as the general path and accepts the line only
// when xi is at least 1e-9 away from a cell edge, so each face's floor()
// agrees.  The PPM antiderivative, remap and foot derivative
// (T2S4_Parareal.cu:260-547) then collapse to
//     b_i = C_{i+S} + K_{i+S+1} - K_{i+S},   K_m = sum_s kappa_s C_{m-2+s},
// i.e. one 6-tap filter b_i = sum_j w_j C_{i+S-2+j}; the host stores S and
// w(xi) per line.  The compact operator and the tridiagonal solve follow.
#include "ccfv_line.h"

namespace ccfv {

// Thomas factors of this lane's fixed cell positions.  For the symmetric
// matrix -p_e == mneg_e for 0 < e < n-1, so one array serves both passes
// (pneg_0 is a scalar, pneg_{n-1} = 0).  Chunks of up to 16 cells keep them
// in registers for every line a wave sweeps (16-cell chunks neither in the
// coarse-chain epilogue launches, whose registers would then cost a wave per
// SIMD, nor in the panel-layout batches that run beside the chain: 1024^2
// pipelined Parareal 97 instead of 82 ms); larger chunks read the global
// table.
template <int CH, bool REG>
struct LaneTab;

template <int CH>
struct LaneTab<CH, true> {
    double m[CH], den[CH], z[CH];
    __device__ __forceinline__ void load(const SweepArgs &a, int base, bool active)
    {
        const int n = a.n;
#pragma unroll
        for (int k = 0; k < CH; k++) {
            int e = active ? base + k : 0;
            m[k] = a.tab[e];
            den[k] = a.tab[n + e];
            z[k] = a.tab[3 * n + e];
        }
    }
    __device__ __forceinline__ double M(int k, int) const { return m[k]; }
    __device__ __forceinline__ double D(int k, int) const { return den[k]; }
    __device__ __forceinline__ double Z(int k, int) const { return z[k]; }
};

template <int CH>
struct LaneTab<CH, false> {
    const double *t;
    int n;
    __device__ __forceinline__ void load(const SweepArgs &a, int, bool)
    {
        t = a.tab;
        n = a.n;
    }
    __device__ __forceinline__ double M(int, int e) const { return t[e]; }
    __device__ __forceinline__ double D(int, int e) const { return t[n + e]; }
    __device__ __forceinline__ double Z(int, int e) const { return t[3 * n + e]; }
};

// a line's cell shift and filter weights
struct LineK {
    int S;
    double w0, w1, w2, w3, w4, w5;
};

__device__ __forceinline__ LineK line_k(const SweepArgs &a, int line)
{
    LineK l;
    l.S = (int)a.lineS[line];
    const double *wp = a.lineK + 6 * (size_t)line;
    l.w0 = wp[0];
    l.w1 = wp[1];
    l.w2 = wp[2];
    l.w3 = wp[3];
    l.w4 = wp[4];
    l.w5 = wp[5];
    return l;
}

// b_q = sum_j w_j C_{q+j} over a window starting two cells before the shift
template <int CH>
__device__ __forceinline__ void filter6(const LineK &lk, const double (&cw)[CH + 5],
                                        double (&b)[CH])
{
#pragma unroll
    for (int q = 0; q < CH; q++)
        b[q] = lk.w0 * cw[q] + lk.w1 * cw[q + 1] + lk.w2 * cw[q + 2] + lk.w3 * cw[q + 3] +
               lk.w4 * cw[q + 4] + lk.w5 * cw[q + 5];
}

template <int CH, bool REG>
__device__ __forceinline__ void fast_line(const SweepArgs &a, double *cs,
                                          const LineK &lk, int lane,
                                          const LaneTab<CH, REG> &tb, double mfl,
                                          double mbl, double pneg0)
{
    const int n = a.n, NL = n / CH;
    const bool active = lane < NL;
    const int base = lane * CH;
    const int S = lk.S;
    double b[CH];
    if (active) {
        double cw[CH + 5];
        int k = base + S - 2;
        k %= n;
        k = k < 0 ? k + n : k;
#pragma unroll
        for (int t = 0; t < CH + 5; t++) {
            cw[t] = cs[pidx(k)];
            k = k + 1 == n ? 0 : k + 1;
        }
        filter6<CH>(lk, cw, b);
    } else {
#pragma unroll
        for (int q = 0; q < CH; q++) b[q] = 0.0;
    }

    // compact operator, periodic (T2S4_Parareal.cu:616-654 with wrap-around)
    double bprev = __shfl_up(b[CH - 1], 1);
    double bnext = __shfl_down(b[0], 1);
    double wrap_last = __shfl(b[CH - 1], NL - 1);
    double wrap_first = __shfl(b[0], 0);
    if (lane == 0) bprev = wrap_last;
    if (lane == NL - 1) bnext = wrap_first;
    double r[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        double bn = k + 1 < CH ? b[k + 1] : bnext;
        double bp = k > 0 ? b[k - 1] : bprev;
        r[k] = bn + 10.0 * b[k] + bp;  // the 1/12 is in den
    }

    // forward pass q_e = r_e den_e + mneg_e q_{e-1} (mneg_0 = 0)
    double M = mfl, Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            double c = r[k] * tb.D(k, base + k);
            r[k] = c;
            Q = c + tb.M(k, base + k) * Q;
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
            carry = r[k] + tb.M(k, base + k) * carry;
            r[k] = carry;
        }
    }

    // back substitution x_e = q_e + pneg_e x_{e+1}
    M = mbl;
    Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) {
            int e = base + k;
            double pn = e == 0 ? pneg0 : (e == n - 1 ? 0.0 : tb.M(k, e));
            Q = r[k] + pn * Q;
        }
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
            int e = base + k;
            double pn = e == 0 ? pneg0 : (e == n - 1 ? 0.0 : tb.M(k, e));
            carry = r[k] + pn * carry;
            r[k] = carry;
        }
    }

    // Sherman-Morrison: x = y - ((y_0 + vl y_{n-1}) / (1 + v.z)) z
    double y0 = __shfl(r[0], 0);
    double yn = __shfl(r[CH - 1], NL - 1);
    double fac = (y0 + a.vl * yn) * a.smrcp;
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) cs[pidx(base + k)] = r[k] - fac * tb.Z(k, base + k);
    }
}

// ---------------------------------------------------------------------------
// Lines of 8 or more cells per lane (n > 256).  A lane's cells are one or
// more whole pidx chunks of 8, so the window of every lane of a line starts
// at the same offset r within a chunk: element t of lane L's window sits at
// 9 Cb(L) + pidx(r + t), a per-lane base plus a wave-uniform offset.  Windows that run
// past the last cell read ghost chunks after it, refreshed from the line's
// first cells before each sweep, so no address wraps.  When the line fills
// the wave (n = 64 CH) the lane-neighbour moves are DPP wave shifts.
// ---------------------------------------------------------------------------
// lane i takes lane i-1 (shr) or i+1 (shl), 0 where that lane is off the
// wave; rotates wrap around the wave
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_ROL1 = 0x134;
constexpr int DPP_WAVE_SHR1 = 0x138, DPP_WAVE_ROR1 = 0x13C;

__device__ __forceinline__ double readlane_d(double x, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), l);
    return __hiloint2double(hi, lo);
}

// the lane's window: cells Cb(L) + pidx(r + t), r wave-uniform (a scalar
// offset added to one per-lane base)
template <int CH>
__device__ __forceinline__ void load_window(const double *p, int r, double (&cw)[CH + 5])
{
#pragma unroll
    for (int t = 0; t < CH + 5; t++) cw[t] = p[(r + t) + ((r + t) >> 3)];
}

// pf / pl: the back-substitution factor of the lane's first / last cell
// (pneg_0 on lane 0, 0 on the last lane, mneg elsewhere).  FULL: n = 64 CH,
// every lane holds cells.
// res: when given, the lane's CH results go there instead of back to cs
template <int CH, bool REG, bool FULL>
__device__ __forceinline__ void fast_line_chunked(const SweepArgs &a, double *cs,
                                                  const LineK &lk, int lane,
                                                  const LaneTab<CH, REG> &tb, double mfl,
                                                  double mbl, double pf, double pl,
                                                  double (*res)[CH] = nullptr)
{
    static_assert(CH >= 8, "chunk-aligned lanes");
    const int n = FULL ? 64 * CH : a.n, NL = FULL ? 64 : n / CH, NC = n >> 3;
    constexpr bool full = FULL;
    const bool active = FULL || lane < NL;
    const int base = lane * CH;

    // ghost chunks: cells [0, FAST_GHOST) again after cell n-1
    if (lane < fast_ghost_cells(CH)) cs[pidx(lane) + 9 * NC] = cs[pidx(lane)];

    int s2 = (lk.S - 2) % n;
    s2 = s2 < 0 ? s2 + n : s2;
    int cb = lane * (CH / 8) + (s2 >> 3);
    cb = cb >= NC ? cb - NC : cb;
    double b[CH];
    if (active) {
        double cw[CH + 5];
        load_window<CH>(cs + 9 * cb, s2 & 7, cw);
        filter6<CH>(lk, cw, b);
    } else {
#pragma unroll
        for (int q = 0; q < CH; q++) b[q] = 0.0;
    }

    // compact operator, periodic (T2S4_Parareal.cu:616-654 with wrap-around)
    double bprev, bnext;
    if constexpr (full) {
        bprev = dpp_d<DPP_WAVE_ROR1>(b[CH - 1]);
        bnext = dpp_d<DPP_WAVE_ROL1>(b[0]);
    } else {
        bprev = __shfl_up(b[CH - 1], 1);
        bnext = __shfl_down(b[0], 1);
        const double wrap_last = readlane_d(b[CH - 1], NL - 1);
        const double wrap_first = readlane_d(b[0], 0);
        if (lane == 0) bprev = wrap_last;
        if (lane == NL - 1) bnext = wrap_first;
    }
    double r[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        double bn = k + 1 < CH ? b[k + 1] : bnext;
        double bp = k > 0 ? b[k - 1] : bprev;
        r[k] = bn + 10.0 * b[k] + bp;  // the 1/12 is in den
    }

    // forward pass q_e = r_e den_e + mneg_e q_{e-1} (mneg_0 = 0)
    double M = mfl, Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            double c = r[k] * tb.D(k, base + k);
            r[k] = c;
            Q = c + tb.M(k, base + k) * Q;
        }
    }
    // lane scan of the chunk maps (Q, M).  A DPP source off the wave reads
    // (0, 0): Q is unchanged and M becomes 0, and such an M only ever scales
    // carries from before lane 0, which are 0, so the scan is exact.
    for (int st = 0, off = 1; st < a.sf; st++, off <<= 1) {
        if (full && off == 1) {
            const double Mp = dpp_d<DPP_WAVE_SHR1>(M);
            const double Qp = dpp_d<DPP_WAVE_SHR1>(Q);
            Q = M * Qp + Q;
            M = M * Mp;
        } else {
            double Mp = __shfl_up(M, off), Qp = __shfl_up(Q, off);
            if (lane >= off) {
                Q = M * Qp + Q;
                M = M * Mp;
            }
        }
    }
    double carry;
    if constexpr (full) {
        carry = dpp_d<DPP_WAVE_SHR1>(Q);
    } else {
        carry = __shfl_up(Q, 1);
        if (lane == 0) carry = 0.0;
    }
    if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) {
            carry = r[k] + tb.M(k, base + k) * carry;
            r[k] = carry;
        }
    }

    // back substitution x_e = q_e + pneg_e x_{e+1}
    M = mbl;
    Q = 0.0;
    if (active) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) {
            const double pn = k == 0 ? pf : (k == CH - 1 ? pl : tb.M(k, base + k));
            Q = r[k] + pn * Q;
        }
    }
    for (int st = 0, off = 1; st < a.sb; st++, off <<= 1) {
        if (full && off == 1) {
            const double Mn = dpp_d<DPP_WAVE_SHL1>(M);
            const double Qn = dpp_d<DPP_WAVE_SHL1>(Q);
            Q = M * Qn + Q;
            M = M * Mn;
        } else {
            double Mn = __shfl_down(M, off), Qn = __shfl_down(Q, off);
            if (lane + off < 64) {
                Q = M * Qn + Q;
                M = M * Mn;
            }
        }
    }
    if constexpr (full) {
        carry = dpp_d<DPP_WAVE_SHL1>(Q);
    } else {
        carry = __shfl_down(Q, 1);
        if (lane == 63) carry = 0.0;
    }
    if (active) {
#pragma unroll
        for (int k = CH - 1; k >= 0; k--) {
            const double pn = k == 0 ? pf : (k == CH - 1 ? pl : tb.M(k, base + k));
            carry = r[k] + pn * carry;
            r[k] = carry;
        }
    }

    // Sherman-Morrison: x = y - ((y_0 + vl y_{n-1}) / (1 + v.z)) z
    const double y0 = readlane_d(r[0], 0);
    const double yn = readlane_d(r[CH - 1], NL - 1);
    const double fac = (y0 + a.vl * yn) * a.smrcp;
    if (res) {
#pragma unroll
        for (int k = 0; k < CH; k++) (*res)[k] = r[k] - fac * tb.Z(k, base + k);
    } else if (active) {
#pragma unroll
        for (int k = 0; k < CH; k++) cs[pidx(base + k)] = r[k] - fac * tb.Z(k, base + k);
    }
}

// The wave's lines l0 = wave, wave+NW, ... of a tile whose first line
// starts at line index L0; each line's parameters are loaded while the
// previous line is swept.
template <int CH, bool REG>
__device__ __forceinline__ void sweep_lines(const SweepArgs &a, double *lds, int LP, int L0,
                                            int W, int NW, int wave, int lane,
                                            const LineK &first, const LaneTab<CH, REG> &tb,
                                            double mfl, double mbl, double pneg0)
{
    LineK cur = first;
    if constexpr (CH >= 8) {
        // back-substitution factors of the lane's end cells
        const int NL = a.n / CH;
        double pf = 0.0, pl = 0.0;
        if (lane < NL) {
            pf = lane == 0 ? pneg0 : tb.M(0, lane * CH);
            pl = lane == NL - 1 ? 0.0 : tb.M(CH - 1, lane * CH + CH - 1);
        }
        if (a.n == 64 * CH) {
            for (int l = wave; l < W; l += NW) {
                LineK nxt = cur;
                if (l + NW < W) nxt = line_k(a, L0 + l + NW);
                fast_line_chunked<CH, REG, true>(a, lds + l * LP, cur, lane, tb, mfl, mbl, pf,
                                                 pl);
                cur = nxt;
            }
        } else {
            for (int l = wave; l < W; l += NW) {
                LineK nxt = cur;
                if (l + NW < W) nxt = line_k(a, L0 + l + NW);
                fast_line_chunked<CH, REG, false>(a, lds + l * LP, cur, lane, tb, mfl, mbl, pf,
                                                  pl);
                cur = nxt;
            }
        }
    } else {
        for (int l = wave; l < W; l += NW) {
            LineK nxt = cur;
            if (l + NW < W) nxt = line_k(a, L0 + l + NW);
            fast_line<CH, REG>(a, lds + l * LP, cur, lane, tb, mfl, mbl, pneg0);
            cur = nxt;
        }
    }
}

// x lines of 64*CH cells (CH >= 8), one wave per line (W == NW, the
// one-field launches), no epilogue: each lane gathers its CH cells of the
// column into registers, runs the sweep (and, fused, the next step's first
// x sweep) through the wave's own LDS line, and scatters its results back.
// No barriers and no LDS round trip before the store.
template <int CH, bool FUSE>
__device__ __forceinline__ void fast_x_direct(const SweepArgs &a, int field, int tile,
                                             int W, double *lds)
{
    static_assert(CH >= 8, "full-wave lines of whole chunks");
    constexpr bool REG = CH <= 16;
    constexpr int n = 64 * CH;
    const int LP = line_pitch(n + fast_ghost_cells(CH));
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int j = tile * W + wave;
    double *cs = lds + wave * LP;
    const size_t row0 = (size_t)(lane * CH) * a.ld;
    const double *in = a.in + foff + j + row0;
    double v[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) v[k] = in[(size_t)k * a.ld];
    LaneTab<CH, REG> tb;
    tb.load(a, lane * CH, true);
    const double mfl = a.tab[4 * n + lane], mbl = a.tab[4 * n + 64 + lane];
    const double pneg0 = a.tab[2 * n];
    const double pf = lane == 0 ? pneg0 : tb.M(0, lane * CH);
    const double pl = lane == 63 ? 0.0 : tb.M(CH - 1, lane * CH + CH - 1);
    const LineK lk = line_k(a, j);
    // the lane's chunk sits at pidx(CH * lane), CH cells and CH/8 pads
    double *own = cs + pidx(CH * lane);
#pragma unroll
    for (int k = 0; k < CH; k++) own[k + (k >> 3)] = v[k];
    double res[CH];
    if (FUSE) {
        fast_line_chunked<CH, REG, true>(a, cs, lk, lane, tb, mfl, mbl, pf, pl);
        fast_line_chunked<CH, REG, true>(a, cs, lk, lane, tb, mfl, mbl, pf, pl, &res);
    } else {
        fast_line_chunked<CH, REG, true>(a, cs, lk, lane, tb, mfl, mbl, pf, pl, &res);
    }
    double *out = (FUSE ? a.out2 : a.out) + foff + j + row0;
#pragma unroll
    for (int k = 0; k < CH; k++) out[(size_t)k * a.ld] = res[k];
}

// FUSE: after the sweep (and its epilogue, whose result stays in LDS) the
// same lines take the next step's first x sub-step, with the same table
// (one step size) and the same per-line shift, and go to out2.
// direct: the launch sweeps one field alone (not a batch, not a ragged
// launch over many problems), where the per-lane direct access pays; a
// bandwidth-bound launch of many fields keeps the coalesced tile loads.
// LAYOUT: LAYOUT_IN / LAYOUT_OUT bits, input / output (out or out2) in the
// panel layout (tiles of PANEL_W lines)
constexpr int LAYOUT_IN = 1, LAYOUT_OUT = 2;
template <int CH, int EPI, bool FUSE, int LAYOUT = 0>
__device__ __forceinline__ void fast_x_body(const SweepArgs &a, int field, int tile,
                                           int W, int NW, double *lds, bool direct = false)
{
    static_assert(LAYOUT == 0 || EPI == EPI_NONE, "epilogues write row-major fields");
    constexpr bool pin = (LAYOUT & LAYOUT_IN) != 0, pout = (LAYOUT & LAYOUT_OUT) != 0;
    if constexpr (CH >= 8 && EPI == EPI_NONE && LAYOUT == 0) {
        if (direct && a.n == 64 * CH && W == NW) {
            fast_x_direct<CH, FUSE>(a, field, tile, W, lds);
            return;
        }
    }
    constexpr bool REG = CH <= 8 || (CH <= 16 && EPI == EPI_NONE && LAYOUT == 0);
    const int n = a.n, LP = line_pitch(n + fast_ghost_cells(CH));
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int j0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    LaneTab<CH, REG> tb;
    tb.load(a, lane * CH, lane < n / CH);
    const double mfl = a.tab[4 * n + lane], mbl = a.tab[4 * n + 64 + lane];
    const double pneg0 = a.tab[2 * n];
    // the wave's first line parameters (and the correction's inputs) travel
    // with the tile loads
    const LineK first = line_k(a, j0 + wave);
    EpiInputs pre;
    if (EPI == EPI_CORRECT) epi_load(a, n, j0, W, threadIdx.x, pre);
    if constexpr (pin) load_tile_x(panel_x_base(a.in + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
    else load_tile_x(a.in + foff, a.ld, n, j0, W, lds, LP);
    __syncthreads();
    sweep_lines<CH, REG>(a, lds, LP, j0, W, NW, wave, lane, first, tb, mfl, mbl, pneg0);
    // the second sweep of a fused launch without an epilogue touches only the
    // wave's own lines: no barrier before it
    if (!FUSE || EPI != EPI_NONE) __syncthreads();
    if (!FUSE) {
        if constexpr (EPI != EPI_NONE) epilogue_x<EPI>(a, n, j0, W, lds, LP, tile, &pre);
        else if constexpr (pout)
            store_tile_x(panel_x_base(a.out + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
        else store_tile_x(a.out + foff, a.ld, n, j0, W, lds, LP);
        return;
    }
    if (EPI != EPI_NONE) {
        epilogue_x<EPI, true>(a, n, j0, W, lds, LP, tile, &pre);
        __syncthreads();
    }
    sweep_lines<CH, REG>(a, lds, LP, j0, W, NW, wave, lane, first, tb, mfl, mbl, pneg0);
    __syncthreads();
    if constexpr (pout)
        store_tile_x(panel_x_base(a.out2 + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
    else store_tile_x(a.out2 + foff, a.ld, n, j0, W, lds, LP);
}

// y lines of 64*CH cells (CH >= 8), one wave per line (W == NW, the
// one-field launches): a y line is a contiguous row, so each lane moves its
// own CH cells between HBM and registers (CH/2 16-byte loads and stores) and
// the line passes through LDS only for the shifted windows.  No barriers,
// and no LDS round trip before the store: the launch's latency drops
// (one-field y sweep at 512^2 6.7 against 7.6 us).  Batches keep the
// coalesced tile loads: per-lane chunks touch four times as many cache
// lines per instruction, which costs more than this saves once the launch
// is bandwidth-bound (64 fields: 57 against 47 us).
template <int CH>
__device__ __forceinline__ void fast_y_direct(const SweepArgs &a, int field, int tile,
                                             int W, int NW, double *lds)
{
    static_assert(CH >= 8, "full-wave lines of whole chunks");
    constexpr bool REG = CH <= 16;
    constexpr int n = 64 * CH;
    const int LP = line_pitch(n + fast_ghost_cells(CH));
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int i0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    double *cs = lds + wave * LP;
    LaneTab<CH, REG> tb;
    tb.load(a, lane * CH, true);
    const double mfl = a.tab[4 * n + lane], mbl = a.tab[4 * n + 64 + lane];
    const double pneg0 = a.tab[2 * n];
    const double pf = lane == 0 ? pneg0 : tb.M(0, lane * CH);
    const double pl = lane == 63 ? 0.0 : tb.M(CH - 1, lane * CH + CH - 1);
    const double *in = a.in + foff + lane * CH;
    double *out = a.out + foff + lane * CH;
    double2 nxt[CH / 2];
    int l = wave;
    if (l < W) {
        const double2 *src = reinterpret_cast<const double2 *>(in + (size_t)(i0 + l) * a.ld);
#pragma unroll
        for (int k = 0; k < CH / 2; k++) nxt[k] = src[k];
    }
    for (; l < W; l += NW) {
        const LineK lk = line_k(a, i0 + l);
        // the lane's chunk sits at pidx(CH * lane) = 9 CH/8 * lane, contiguous
        double *own = cs + pidx(CH * lane);
#pragma unroll
        for (int k = 0; k < CH / 2; k++) {
            own[2 * k + (2 * k >> 3)] = nxt[k].x;
            own[2 * k + 1 + ((2 * k + 1) >> 3)] = nxt[k].y;
        }
        if (l + NW < W) {
            const double2 *src =
                reinterpret_cast<const double2 *>(in + (size_t)(i0 + l + NW) * a.ld);
#pragma unroll
            for (int k = 0; k < CH / 2; k++) nxt[k] = src[k];
        }
        double res[CH];
        fast_line_chunked<CH, REG, true>(a, cs, lk, lane, tb, mfl, mbl, pf, pl, &res);
        double2 *dst = reinterpret_cast<double2 *>(out + (size_t)(i0 + l) * a.ld);
#pragma unroll
        for (int k = 0; k < CH / 2; k++) dst[k] = make_double2(res[2 * k], res[2 * k + 1]);
    }
}

// The same for a field in the panel layout (in and out): row i of the field
// is 8 contiguous doubles at p * ps + 8 i in every panel p, so a lane's CH
// cells are CH / 8 runs of 8 (64 bytes) in consecutive panels.
template <int CH>
__device__ __forceinline__ void fast_y_direct_panel(const SweepArgs &a, int field, int tile,
                                                   int W, int NW, double *lds)
{
    static_assert(CH >= 8 && PANEL_W == 8, "full-wave lines of whole panels");
    constexpr bool REG = CH <= 16;
    constexpr int n = 64 * CH, NP = CH / 8;
    const int LP = line_pitch(n + fast_ghost_cells(CH));
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const size_t ps = (size_t)PANEL_W * (size_t)a.nl;
    const int i0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    double *cs = lds + wave * LP;
    LaneTab<CH, REG> tb;
    tb.load(a, lane * CH, true);
    const double mfl = a.tab[4 * n + lane], mbl = a.tab[4 * n + 64 + lane];
    const double pneg0 = a.tab[2 * n];
    const double pf = lane == 0 ? pneg0 : tb.M(0, lane * CH);
    const double pl = lane == 63 ? 0.0 : tb.M(CH - 1, lane * CH + CH - 1);
    // the lane's first panel
    const double *in = a.in + foff + (size_t)(lane * NP) * ps;
    double *out = a.out + foff + (size_t)(lane * NP) * ps;
    double2 nxt[CH / 2];
    auto fetch = [&](int row) {
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const double2 *src =
                reinterpret_cast<const double2 *>(in + (size_t)q * ps + (size_t)PANEL_W * row);
#pragma unroll
            for (int k = 0; k < 4; k++) nxt[4 * q + k] = src[k];
        }
    };
    int l = wave;
    if (l < W) fetch(i0 + l);
    for (; l < W; l += NW) {
        const LineK lk = line_k(a, i0 + l);
        double *own = cs + pidx(CH * lane);
#pragma unroll
        for (int k = 0; k < CH / 2; k++) {
            own[2 * k + (2 * k >> 3)] = nxt[k].x;
            own[2 * k + 1 + ((2 * k + 1) >> 3)] = nxt[k].y;
        }
        if (l + NW < W) fetch(i0 + l + NW);
        double res[CH];
        fast_line_chunked<CH, REG, true>(a, cs, lk, lane, tb, mfl, mbl, pf, pl, &res);
#pragma unroll
        for (int q = 0; q < NP; q++) {
            double2 *dst = reinterpret_cast<double2 *>(out + (size_t)q * ps +
                                                       (size_t)PANEL_W * (i0 + l));
#pragma unroll
            for (int k = 0; k < 4; k++)
                dst[k] = make_double2(res[8 * q + 2 * k], res[8 * q + 2 * k + 1]);
        }
    }
}

template <int CH, int LAYOUT = 0>
__device__ __forceinline__ void fast_y_body(const SweepArgs &a, int field, int tile,
                                           int W, int NW, double *lds, bool direct = false)
{
    constexpr bool pin = (LAYOUT & LAYOUT_IN) != 0, pout = (LAYOUT & LAYOUT_OUT) != 0;
    if constexpr (CH >= 8 && LAYOUT == 0) {
        if (direct && a.n == 64 * CH && W == NW) {
            fast_y_direct<CH>(a, field, tile, W, NW, lds);
            return;
        }
    }
    if constexpr (CH >= 8 && pin && pout) {
        if (direct && a.n == 64 * CH && W == NW) {
            fast_y_direct_panel<CH>(a, field, tile, W, NW, lds);
            return;
        }
    }
    constexpr bool REG = CH <= 8 || (CH <= 16 && LAYOUT == 0);
    const int n = a.n, LP = line_pitch(n + fast_ghost_cells(CH));
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int i0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    LaneTab<CH, REG> tb;
    tb.load(a, lane * CH, lane < n / CH);
    const double mfl = a.tab[4 * n + lane], mbl = a.tab[4 * n + 64 + lane];
    const double pneg0 = a.tab[2 * n];
    const LineK first = line_k(a, i0 + wave);
    // panel stride: PANEL_W doubles per row of the field (a.nl rows)
    const size_t ps = (size_t)PANEL_W * (size_t)a.nl;
    if constexpr (pin) load_tile_y_panel(a.in + foff, ps, n, i0, W, lds, LP);
    else load_tile_y(a.in + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
    __syncthreads();
    sweep_lines<CH, REG>(a, lds, LP, i0, W, NW, wave, lane, first, tb, mfl, mbl, pneg0);
    __syncthreads();
    if constexpr (pout) store_tile_y_panel(a.out + foff, ps, n, i0, W, lds, LP);
    else store_tile_y(a.out + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
}

// ---------------------------------------------------------------------------
// Split lines: a line of n = 64 CH WPL cells swept by WPL waves of one
// workgroup, wave s of the line's group holding segment s (global lane
// g = 64 s + lane, cells CH g .. CH g + CH - 1).  1024-cell lines take it
// with CH = 8: one wave per such line held 16 cells per lane, its Thomas
// factors in 96 registers (two waves per SIMD), and twice the serial chain
// of a 512-cell line.  The segments meet through a few LDS slots after the
// line (SplitSlots): the compact operator's neighbours at the segment ends,
// the forward and backward chunk maps of the segment end lanes, and the
// Sherman-Morrison end values; five workgroup barriers per sweep, which
// every wave of the workgroup reaches (each group sweeps W / G lines).
// The lane scans run within each wave for sfs / sbs doubling steps (the
// truncation of the one-wave sweep; lanes below the step keep their exact
// prefix), and a segment's carry-in folds the end maps of the segments
// before it exactly, so a line's solve is the one-wave solve up to the
// products the truncation neglects (below 1e-18).
// ---------------------------------------------------------------------------
template <int WPL>
struct SplitSlots {
    // per segment: b of its first / last cell, the forward map (Q, M) of
    // its last lane, the backward map of its first lane; then y_0, y_{n-1}
    static constexpr int BF = 0, BL = WPL, QF = 2 * WPL, MF = 3 * WPL, QB = 4 * WPL,
                         MB = 5 * WPL, Y0 = 6 * WPL, YN = 6 * WPL + 1, SIZE = 6 * WPL + 2;
};

// LDS doubles per line: the cells (pidx layout), the ghost chunks, the slots
template <int CH, int WPL>
__host__ __device__ constexpr int split_pitch()
{
    return 64 * CH * WPL + 8 * ((CH + 11) >> 3) + ((64 * CH * WPL + 8 * ((CH + 11) >> 3)) >> 3) +
           1 + SplitSlots<WPL>::SIZE;
}

// res: when given, the lane's CH results go there instead of back to cs
template <int CH, int WPL>
__device__ __forceinline__ void fast_line_split(const SweepArgs &a, double *cs,
                                                const LineK &lk, int seg, int lane,
                                                const LaneTab<CH, true> &tb, double mfl,
                                                double mbl, double pf, double pl,
                                                double (*res)[CH] = nullptr)
{
    static_assert(CH == 8, "split lines: lanes of one pidx chunk");
    using X = SplitSlots<WPL>;
    constexpr int n = 64 * CH * WPL, NC = n >> 3;
    double *xc = cs + line_pitch(n + fast_ghost_cells(CH));
    const int base = (seg * 64 + lane) * CH;

    // ghost chunks after cell n-1 (segment 0 holds cells [0, FAST_GHOST));
    // then every segment's cells are in place
    if (seg == 0 && lane < fast_ghost_cells(CH)) cs[pidx(lane) + 9 * NC] = cs[pidx(lane)];
    __syncthreads();

    int s2 = (lk.S - 2) % n;
    s2 = s2 < 0 ? s2 + n : s2;
    int cb = (seg * 64 + lane) * (CH / 8) + (s2 >> 3);
    cb = cb >= NC ? cb - NC : cb;
    double b[CH];
    {
        double cw[CH + 5];
        load_window<CH>(cs + 9 * cb, s2 & 7, cw);
        filter6<CH>(lk, cw, b);
    }

    // compact operator, periodic (T2S4_Parareal.cu:616-654 with wrap-around);
    // past this barrier no wave reads the line's cells again
    if (lane == 0) xc[X::BF + seg] = b[0];
    if (lane == 63) xc[X::BL + seg] = b[CH - 1];
    __syncthreads();
    double bprev = dpp_d<DPP_WAVE_SHR1>(b[CH - 1]);
    double bnext = dpp_d<DPP_WAVE_SHL1>(b[0]);
    if (lane == 0) bprev = xc[X::BL + (seg == 0 ? WPL - 1 : seg - 1)];
    if (lane == 63) bnext = xc[X::BF + (seg == WPL - 1 ? 0 : seg + 1)];
    double r[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
        double bn = k + 1 < CH ? b[k + 1] : bnext;
        double bp = k > 0 ? b[k - 1] : bprev;
        r[k] = bn + 10.0 * b[k] + bp;  // the 1/12 is in den
    }

    // forward pass q_e = r_e den_e + mneg_e q_{e-1} (mneg_0 = 0)
    double M = mfl, Q = 0.0;
#pragma unroll
    for (int k = 0; k < CH; k++) {
        double c = r[k] * tb.D(k, base + k);
        r[k] = c;
        Q = c + tb.M(k, base + k) * Q;
    }
    for (int st = 0, off = 1; st < a.sfs; st++, off <<= 1) {
        double Mp, Qp;
        if (off == 1) {
            Mp = dpp_d<DPP_WAVE_SHR1>(M);
            Qp = dpp_d<DPP_WAVE_SHR1>(Q);
        } else {
            Mp = __shfl_up(M, off);
            Qp = __shfl_up(Q, off);
        }
        if (lane >= off) {
            Q = M * Qp + Q;
            M = M * Mp;
        }
    }
    if (lane == 63) {
        xc[X::QF + seg] = Q;
        xc[X::MF + seg] = M;
    }
    __syncthreads();
    double cin = 0.0;
    for (int s = 0; s < seg; s++) cin = xc[X::QF + s] + xc[X::MF + s] * cin;
    double carry;
    {
        const double Qs = dpp_d<DPP_WAVE_SHR1>(Q), Ms = dpp_d<DPP_WAVE_SHR1>(M);
        carry = lane == 0 ? cin : Qs + Ms * cin;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) {
        carry = r[k] + tb.M(k, base + k) * carry;
        r[k] = carry;
    }

    // back substitution x_e = q_e + pneg_e x_{e+1}
    M = mbl;
    Q = 0.0;
#pragma unroll
    for (int k = CH - 1; k >= 0; k--) {
        const double pn = k == 0 ? pf : (k == CH - 1 ? pl : tb.M(k, base + k));
        Q = r[k] + pn * Q;
    }
    for (int st = 0, off = 1; st < a.sbs; st++, off <<= 1) {
        double Mn, Qn;
        if (off == 1) {
            Mn = dpp_d<DPP_WAVE_SHL1>(M);
            Qn = dpp_d<DPP_WAVE_SHL1>(Q);
        } else {
            Mn = __shfl_down(M, off);
            Qn = __shfl_down(Q, off);
        }
        if (lane + off < 64) {
            Q = M * Qn + Q;
            M = M * Mn;
        }
    }
    if (lane == 0) {
        xc[X::QB + seg] = Q;
        xc[X::MB + seg] = M;
    }
    __syncthreads();
    double cinb = 0.0;
    for (int s = WPL - 1; s > seg; s--) cinb = xc[X::QB + s] + xc[X::MB + s] * cinb;
    {
        const double Qs = dpp_d<DPP_WAVE_SHL1>(Q), Ms = dpp_d<DPP_WAVE_SHL1>(M);
        carry = lane == 63 ? cinb : Qs + Ms * cinb;
    }
#pragma unroll
    for (int k = CH - 1; k >= 0; k--) {
        const double pn = k == 0 ? pf : (k == CH - 1 ? pl : tb.M(k, base + k));
        carry = r[k] + pn * carry;
        r[k] = carry;
    }

    // Sherman-Morrison: x = y - ((y_0 + vl y_{n-1}) / (1 + v.z)) z
    if (seg == 0 && lane == 0) xc[X::Y0] = r[0];
    if (seg == WPL - 1 && lane == 63) xc[X::YN] = r[CH - 1];
    __syncthreads();
    const double fac = (xc[X::Y0] + a.vl * xc[X::YN]) * a.smrcp;
    if (res) {
#pragma unroll
        for (int k = 0; k < CH; k++) (*res)[k] = r[k] - fac * tb.Z(k, base + k);
    } else {
#pragma unroll
        for (int k = 0; k < CH; k++) cs[pidx(base + k)] = r[k] - fac * tb.Z(k, base + k);
    }
}

// a split-line lane's table values: Thomas factors of its cells, chunk
// products, back-substitution factors of its end cells
template <int CH, int WPL>
struct SplitLane {
    LaneTab<CH, true> tb;
    double mfl, mbl, pf, pl;
    __device__ __forceinline__ void load(const SweepArgs &a, int seg, int lane)
    {
        constexpr int n = 64 * CH * WPL, NL = 64 * WPL;
        const int g = seg * 64 + lane;
        tb.load(a, g * CH, true);
        mfl = a.tab[4 * n + 128 + g];
        mbl = a.tab[4 * n + 128 + NL + g];
        pf = g == 0 ? a.tab[2 * n] : tb.M(0, g * CH);
        pl = g == NL - 1 ? 0.0 : tb.M(CH - 1, g * CH + CH - 1);
    }
};

// lines group, group + G, ... of a tile whose first line is L0 (every group
// sweeps W / G lines, so every wave reaches the same barriers)
template <int CH, int WPL>
__device__ __forceinline__ void sweep_lines_split(const SweepArgs &a, double *lds, int LP,
                                                  int L0, int W, int G, int group, int seg,
                                                  int lane, const SplitLane<CH, WPL> &sl)
{
    LineK cur = line_k(a, L0 + group);
    for (int l = group; l < W; l += G) {
        LineK nxt = cur;
        if (l + G < W) nxt = line_k(a, L0 + l + G);
        fast_line_split<CH, WPL>(a, lds + l * LP, cur, seg, lane, sl.tb, sl.mfl, sl.mbl, sl.pf,
                                 sl.pl);
        cur = nxt;
    }
}

// one-field x launch of split lines, one line per group (W == G): each lane
// gathers its CH cells of the column, the line's segments meet in LDS, the
// results are scattered back from registers
template <int CH, int WPL, bool FUSE>
__device__ __forceinline__ void fast_x_direct_split(const SweepArgs &a, int field, int tile,
                                                   int W, double *lds)
{
    constexpr int LP = split_pitch<CH, WPL>();
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int group = wave / WPL, seg = wave - group * WPL;
    const int j = tile * W + group;
    double *cs = lds + group * LP;
    const int g = seg * 64 + lane;
    const size_t row0 = (size_t)(g * CH) * a.ld;
    const double *in = a.in + foff + j + row0;
    double v[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) v[k] = in[(size_t)k * a.ld];
    SplitLane<CH, WPL> sl;
    sl.load(a, seg, lane);
    const LineK lk = line_k(a, j);
    double *own = cs + pidx(CH * g);
#pragma unroll
    for (int k = 0; k < CH; k++) own[k + (k >> 3)] = v[k];
    double res[CH];
    if (FUSE) fast_line_split<CH, WPL>(a, cs, lk, seg, lane, sl.tb, sl.mfl, sl.mbl, sl.pf, sl.pl);
    fast_line_split<CH, WPL>(a, cs, lk, seg, lane, sl.tb, sl.mfl, sl.mbl, sl.pf, sl.pl, &res);
    double *out = (FUSE ? a.out2 : a.out) + foff + j + row0;
#pragma unroll
    for (int k = 0; k < CH; k++) out[(size_t)k * a.ld] = res[k];
}

// one-field y launch of split lines (row-major, W == G): each lane moves
// its CH contiguous cells between HBM and registers
template <int CH, int WPL>
__device__ __forceinline__ void fast_y_direct_split(const SweepArgs &a, int field, int tile,
                                                   int W, double *lds)
{
    constexpr int LP = split_pitch<CH, WPL>();
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int group = wave / WPL, seg = wave - group * WPL;
    const int i = tile * W + group;
    double *cs = lds + group * LP;
    const int g = seg * 64 + lane;
    const double2 *src =
        reinterpret_cast<const double2 *>(a.in + foff + (size_t)i * a.ld + g * CH);
    double2 v[CH / 2];
#pragma unroll
    for (int k = 0; k < CH / 2; k++) v[k] = src[k];
    SplitLane<CH, WPL> sl;
    sl.load(a, seg, lane);
    const LineK lk = line_k(a, i);
    double *own = cs + pidx(CH * g);
#pragma unroll
    for (int k = 0; k < CH / 2; k++) {
        own[2 * k + (2 * k >> 3)] = v[k].x;
        own[2 * k + 1 + ((2 * k + 1) >> 3)] = v[k].y;
    }
    double res[CH];
    fast_line_split<CH, WPL>(a, cs, lk, seg, lane, sl.tb, sl.mfl, sl.mbl, sl.pf, sl.pl, &res);
    double2 *dst = reinterpret_cast<double2 *>(a.out + foff + (size_t)i * a.ld + g * CH);
#pragma unroll
    for (int k = 0; k < CH / 2; k++) dst[k] = make_double2(res[2 * k], res[2 * k + 1]);
}

// the x launch of split lines: tiles of W lines, G groups of WPL waves (the
// host's tile.nw is G); the arrangement of fast_x_body otherwise
template <int CH, int WPL, int EPI, bool FUSE, int LAYOUT = 0>
__device__ __forceinline__ void fast_x_body_split(const SweepArgs &a, int field, int tile,
                                                 int W, int G, double *lds, bool direct)
{
    static_assert(LAYOUT == 0 || EPI == EPI_NONE, "epilogues write row-major fields");
    constexpr bool pin = (LAYOUT & LAYOUT_IN) != 0, pout = (LAYOUT & LAYOUT_OUT) != 0;
    if constexpr (EPI == EPI_NONE && LAYOUT == 0) {
        if (direct && W == G) {
            fast_x_direct_split<CH, WPL, FUSE>(a, field, tile, W, lds);
            return;
        }
    }
    constexpr int n = 64 * CH * WPL, LP = split_pitch<CH, WPL>();
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int j0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int group = wave / WPL, seg = wave - group * WPL;
    SplitLane<CH, WPL> sl;
    sl.load(a, seg, lane);
    EpiInputs pre;
    if (EPI == EPI_CORRECT) epi_load(a, n, j0, W, threadIdx.x, pre);
    if constexpr (pin) load_tile_x(panel_x_base(a.in + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
    else load_tile_x(a.in + foff, a.ld, n, j0, W, lds, LP);
    __syncthreads();
    sweep_lines_split<CH, WPL>(a, lds, LP, j0, W, G, group, seg, lane, sl);
    __syncthreads();
    if (!FUSE) {
        if constexpr (EPI != EPI_NONE) epilogue_x<EPI>(a, n, j0, W, lds, LP, tile, &pre);
        else if constexpr (pout)
            store_tile_x(panel_x_base(a.out + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
        else store_tile_x(a.out + foff, a.ld, n, j0, W, lds, LP);
        return;
    }
    if (EPI != EPI_NONE) {
        epilogue_x<EPI, true>(a, n, j0, W, lds, LP, tile, &pre);
        __syncthreads();
    }
    sweep_lines_split<CH, WPL>(a, lds, LP, j0, W, G, group, seg, lane, sl);
    __syncthreads();
    if constexpr (pout)
        store_tile_x(panel_x_base(a.out2 + foff, n, j0), PANEL_W, n, j0, W, lds, LP);
    else store_tile_x(a.out2 + foff, a.ld, n, j0, W, lds, LP);
}

template <int CH, int WPL, int LAYOUT = 0>
__device__ __forceinline__ void fast_y_body_split(const SweepArgs &a, int field, int tile,
                                                 int W, int G, double *lds, bool direct)
{
    constexpr bool pin = (LAYOUT & LAYOUT_IN) != 0, pout = (LAYOUT & LAYOUT_OUT) != 0;
    if constexpr (LAYOUT == 0) {
        if (direct && W == G) {
            fast_y_direct_split<CH, WPL>(a, field, tile, W, lds);
            return;
        }
    }
    constexpr int n = 64 * CH * WPL, LP = split_pitch<CH, WPL>();
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int i0 = tile * W;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int group = wave / WPL, seg = wave - group * WPL;
    SplitLane<CH, WPL> sl;
    sl.load(a, seg, lane);
    const size_t ps = (size_t)PANEL_W * (size_t)a.nl;
    if constexpr (pin) load_tile_y_panel(a.in + foff, ps, n, i0, W, lds, LP);
    else load_tile_y(a.in + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
    __syncthreads();
    sweep_lines_split<CH, WPL>(a, lds, LP, i0, W, G, group, seg, lane, sl);
    __syncthreads();
    if constexpr (pout) store_tile_y_panel(a.out + foff, ps, n, i0, W, lds, LP);
    else store_tile_y(a.out + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
}

template <int CH, int WPL, int EPI, bool FUSE, int LAYOUT = 0>
__global__ __launch_bounds__(256 * WPL) void fast_x_split_kernel(SweepArgs a, int W, int G)
{
    extern __shared__ double lds[];
    fast_x_body_split<CH, WPL, EPI, FUSE, LAYOUT>(a, blockIdx.y,
                                                  xcd_tile(blockIdx.x, gridDim.x), W, G, lds,
                                                  gridDim.y == 1);
}

template <int CH, int WPL, int LAYOUT = 0>
__global__ __launch_bounds__(256 * WPL) void fast_y_split_kernel(SweepArgs a, int W, int G)
{
    extern __shared__ double lds[];
    fast_y_body_split<CH, WPL, LAYOUT>(a, blockIdx.y, blockIdx.x, W, G, lds, gridDim.y == 1);
}

template <int CH, int WPL, int EPI, bool FUSE>
__global__ __launch_bounds__(256 * WPL) void ragged_fast_x_split(const RaggedItem *items,
                                                                 const int *prefix, int nitems,
                                                                 int W, int G)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const int local = ragged_xcd_tile(b - prefix[i], prefix[i], prefix[i + 1] - prefix[i]);
    const int nt = items[i].ntiles;
    const int field = local / nt;
    const SweepArgs a = items[i].a;
    fast_x_body_split<CH, WPL, EPI, FUSE>(a, field, local - field * nt, W, G, lds, false);
}

template <int CH, int WPL>
__global__ __launch_bounds__(256 * WPL) void ragged_fast_y_split(const RaggedItem *items,
                                                                 const int *prefix, int nitems,
                                                                 int W, int G)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const int local = b - prefix[i];
    const int nt = items[i].ntiles;
    const int field = local / nt;
    const SweepArgs a = items[i].a;
    fast_y_body_split<CH, WPL>(a, field, local - field * nt, W, G, lds, false);
}

template <int CH, int EPI, bool FUSE, int LAYOUT = 0>
__global__ __launch_bounds__(256) void fast_x_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    fast_x_body<CH, EPI, FUSE, LAYOUT>(a, blockIdx.y, xcd_tile(blockIdx.x, gridDim.x), W,
                                       NW, lds, gridDim.y == 1);
}

template <int CH, int LAYOUT = 0>
__global__ __launch_bounds__(256) void fast_y_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    fast_y_body<CH, LAYOUT>(a, blockIdx.y, blockIdx.x, W, NW, lds, gridDim.y == 1);
}

// ragged: one launch over many problems of the same line class
template <int CH, int EPI, bool FUSE>
__global__ __launch_bounds__(256) void ragged_fast_x(const RaggedItem *items,
                                                      const int *prefix, int nitems,
                                                      int W, int NW)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const int local = ragged_xcd_tile(b - prefix[i], prefix[i], prefix[i + 1] - prefix[i]);
    const int nt = items[i].ntiles;
    const int field = local / nt;
    const SweepArgs a = items[i].a;
    fast_x_body<CH, EPI, FUSE>(a, field, local - field * nt, W, NW, lds);
}

template <int CH>
__global__ __launch_bounds__(256) void ragged_fast_y(const RaggedItem *items,
                                                      const int *prefix, int nitems,
                                                      int W, int NW)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = __builtin_amdgcn_readfirstlane(ragged_item(prefix, nitems, b));
    const int local = b - prefix[i];
    const int nt = items[i].ntiles;
    const int field = local / nt;
    const SweepArgs a = items[i].a;
    fast_y_body<CH>(a, field, local - field * nt, W, NW, lds);
}

size_t sweep_lds_bytes_fast(int n, Tile tile)
{
    if (fast_wpl(n) == 2) return sizeof(double) * (size_t)split_pitch<8, 2>() * (size_t)tile.w;
    return sizeof(double) * (size_t)line_pitch(n + fast_ghost_cells(chunk_for(n))) *
           (size_t)tile.w;
}

template <typename K>
static hipError_t launch(K k, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                         const SweepArgs &a, Tile t)
{
    hipLaunchKernelGGL(k, grid, block, lds, s, a, t.w, t.nw);
    return hipGetLastError();
}

template <int CH, bool FUSE>
static hipError_t dispatch_x(Epi epi, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                             const SweepArgs &a, Tile t)
{
    const int layout = (a.in_panel ? LAYOUT_IN : 0) | (a.out_panel ? LAYOUT_OUT : 0);
    if (layout && (epi != EPI_NONE || t.w != PANEL_W)) return hipErrorInvalidValue;
    if (layout == LAYOUT_IN)
        return launch(fast_x_kernel<CH, EPI_NONE, FUSE, LAYOUT_IN>, grid, block, lds, s, a, t);
    if (layout == LAYOUT_OUT)
        return launch(fast_x_kernel<CH, EPI_NONE, FUSE, LAYOUT_OUT>, grid, block, lds, s, a, t);
    if (layout)
        return launch(fast_x_kernel<CH, EPI_NONE, FUSE, LAYOUT_IN | LAYOUT_OUT>, grid, block,
                      lds, s, a, t);
    if (epi == EPI_NONE)
        return launch(fast_x_kernel<CH, EPI_NONE, FUSE>, grid, block, lds, s, a, t);
    if (epi == EPI_INIT)
        return launch(fast_x_kernel<CH, EPI_INIT, FUSE>, grid, block, lds, s, a, t);
    return launch(fast_x_kernel<CH, EPI_CORRECT, FUSE>, grid, block, lds, s, a, t);
}

template <int CH>
static hipError_t dispatch(Dir dir, Epi epi, bool fuse, const SweepArgs &a, int nf,
                           Tile t, hipStream_t s)
{
    dim3 grid(a.nl / t.w, nf), block(64 * t.nw);
    size_t lds = sweep_lds_bytes_fast(a.n, t);
    if (dir == DIR_Y) {
        if (a.in_panel != a.out_panel || (a.in_panel && a.n % PANEL_W))
            return hipErrorInvalidValue;
        if (a.in_panel)
            return launch(fast_y_kernel<CH, LAYOUT_IN | LAYOUT_OUT>, grid, block, lds, s, a, t);
        return launch(fast_y_kernel<CH>, grid, block, lds, s, a, t);
    }
    return fuse ? dispatch_x<CH, true>(epi, grid, block, lds, s, a, t)
                : dispatch_x<CH, false>(epi, grid, block, lds, s, a, t);
}

template <int CH, int WPL, bool FUSE>
static hipError_t dispatch_x_split(Epi epi, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                                   const SweepArgs &a, Tile t)
{
    const int layout = (a.in_panel ? LAYOUT_IN : 0) | (a.out_panel ? LAYOUT_OUT : 0);
    if (layout && (epi != EPI_NONE || t.w != PANEL_W)) return hipErrorInvalidValue;
    if (layout == LAYOUT_IN)
        return launch(fast_x_split_kernel<CH, WPL, EPI_NONE, FUSE, LAYOUT_IN>, grid, block, lds,
                      s, a, t);
    if (layout == LAYOUT_OUT)
        return launch(fast_x_split_kernel<CH, WPL, EPI_NONE, FUSE, LAYOUT_OUT>, grid, block, lds,
                      s, a, t);
    if (layout)
        return launch(fast_x_split_kernel<CH, WPL, EPI_NONE, FUSE, LAYOUT_IN | LAYOUT_OUT>, grid,
                      block, lds, s, a, t);
    if (epi == EPI_NONE)
        return launch(fast_x_split_kernel<CH, WPL, EPI_NONE, FUSE>, grid, block, lds, s, a, t);
    if (epi == EPI_INIT)
        return launch(fast_x_split_kernel<CH, WPL, EPI_INIT, FUSE>, grid, block, lds, s, a, t);
    return launch(fast_x_split_kernel<CH, WPL, EPI_CORRECT, FUSE>, grid, block, lds, s, a, t);
}

// split lines: tile.nw groups of WPL waves
template <int CH, int WPL>
static hipError_t dispatch_split(Dir dir, Epi epi, bool fuse, const SweepArgs &a, int nf,
                                 Tile t, hipStream_t s)
{
    dim3 grid(a.nl / t.w, nf), block(64 * t.nw * WPL);
    size_t lds = sweep_lds_bytes_fast(a.n, t);
    if (dir == DIR_Y) {
        if (a.in_panel != a.out_panel || (a.in_panel && a.n % PANEL_W))
            return hipErrorInvalidValue;
        if (a.in_panel)
            return launch(fast_y_split_kernel<CH, WPL, LAYOUT_IN | LAYOUT_OUT>, grid, block, lds,
                          s, a, t);
        return launch(fast_y_split_kernel<CH, WPL>, grid, block, lds, s, a, t);
    }
    return fuse ? dispatch_x_split<CH, WPL, true>(epi, grid, block, lds, s, a, t)
                : dispatch_x_split<CH, WPL, false>(epi, grid, block, lds, s, a, t);
}

hipError_t launch_sweep_fast(Dir dir, Epi epi, const SweepArgs &a, int nfields,
                             Tile tile, hipStream_t s, bool fuse)
{
    if (fast_wpl(a.n) == 2) return dispatch_split<8, 2>(dir, epi, fuse, a, nfields, tile, s);
    switch (chunk_for(a.n)) {
    case 1: return dispatch<1>(dir, epi, fuse, a, nfields, tile, s);
    case 2: return dispatch<2>(dir, epi, fuse, a, nfields, tile, s);
    case 4: return dispatch<4>(dir, epi, fuse, a, nfields, tile, s);
    case 8: return dispatch<8>(dir, epi, fuse, a, nfields, tile, s);
    case 16: return dispatch<16>(dir, epi, fuse, a, nfields, tile, s);
    case 32: return dispatch<32>(dir, epi, fuse, a, nfields, tile, s);
    }
    return hipErrorInvalidValue;
}

template <typename K>
static hipError_t launch_r(K k, int nblocks, Tile t, size_t lds, hipStream_t s,
                           const RaggedItem *items, const int *prefix, int nitems,
                           int wpl = 1)
{
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * t.nw * wpl), lds, s, items, prefix,
                       nitems, t.w, t.nw);
    return hipGetLastError();
}

template <int CH, int WPL, bool FUSE>
static hipError_t dispatch_rx_split(Epi epi, const RaggedItem *items, const int *prefix,
                                    int nitems, int nblocks, Tile t, size_t lds, hipStream_t s)
{
    if (epi == EPI_NONE)
        return launch_r(ragged_fast_x_split<CH, WPL, EPI_NONE, FUSE>, nblocks, t, lds, s, items,
                        prefix, nitems, WPL);
    if (epi == EPI_INIT)
        return launch_r(ragged_fast_x_split<CH, WPL, EPI_INIT, FUSE>, nblocks, t, lds, s, items,
                        prefix, nitems, WPL);
    return launch_r(ragged_fast_x_split<CH, WPL, EPI_CORRECT, FUSE>, nblocks, t, lds, s, items,
                    prefix, nitems, WPL);
}

template <int CH, bool FUSE>
static hipError_t dispatch_rx(Epi epi, const RaggedItem *items, const int *prefix,
                              int nitems, int nblocks, Tile t, size_t lds, hipStream_t s)
{
    if (epi == EPI_NONE)
        return launch_r(ragged_fast_x<CH, EPI_NONE, FUSE>, nblocks, t, lds, s, items,
                        prefix, nitems);
    if (epi == EPI_INIT)
        return launch_r(ragged_fast_x<CH, EPI_INIT, FUSE>, nblocks, t, lds, s, items,
                        prefix, nitems);
    return launch_r(ragged_fast_x<CH, EPI_CORRECT, FUSE>, nblocks, t, lds, s, items,
                    prefix, nitems);
}

template <int CH>
static hipError_t dispatch_r(Dir dir, Epi epi, bool fuse, const RaggedItem *items,
                             const int *prefix, int nitems, int nblocks, Tile t,
                             size_t lds, hipStream_t s)
{
    if (dir == DIR_Y)
        return launch_r(ragged_fast_y<CH>, nblocks, t, lds, s, items, prefix, nitems);
    return fuse ? dispatch_rx<CH, true>(epi, items, prefix, nitems, nblocks, t, lds, s)
                : dispatch_rx<CH, false>(epi, items, prefix, nitems, nblocks, t, lds, s);
}

hipError_t launch_ragged_fast(Dir dir, Epi epi, int ch, int wpl, bool fuse,
                              const RaggedItem *items, const int *prefix, int nitems,
                              int nblocks, Tile t, size_t lds, hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    if (wpl == 2) {
        // split lines (1024 cells, 8 per lane)
        if (dir == DIR_Y)
            return launch_r(ragged_fast_y_split<8, 2>, nblocks, t, lds, s, items, prefix,
                            nitems, 2);
        return fuse ? dispatch_rx_split<8, 2, true>(epi, items, prefix, nitems, nblocks, t, lds,
                                                    s)
                    : dispatch_rx_split<8, 2, false>(epi, items, prefix, nitems, nblocks, t,
                                                     lds, s);
    }
    if (wpl != 1) return hipErrorInvalidValue;
    switch (ch) {
    case 1: return dispatch_r<1>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    case 2: return dispatch_r<2>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    case 4: return dispatch_r<4>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    case 8: return dispatch_r<8>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    case 16: return dispatch_r<16>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    case 32: return dispatch_r<32>(dir, epi, fuse, items, prefix, nitems, nblocks, t, lds, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ccfv
