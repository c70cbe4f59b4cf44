// ccfv_sweep_generic.hip — general line sweeps: every face foot traced and
// located on its own (closed reference box, time-dependent or non-uniform
// velocity, and any line the uniform path cannot take).
//
// One launch performs one sub-step of the Strang step (T2S4_Parareal.cu:
// 825-870, :873-917 or :920-963) for a batch of fields.  A block stages W
// lines in LDS; each wavefront sweeps whole lines; the tile is written back
// in place (every sweep is line-local).
//
// The file is compiled once per chunk size (cells per lane), with
// -DCCFV_GEN_CH=1, 2, 4, 8, 16 or 32: each object holds that chunk size's
// kernels and its launchers launch_sweep_generic_chN / launch_ragged_generic_chN.
// Compiled without CCFV_GEN_CH it holds only the dispatch over chunk sizes.
#include "ccfv_line.h"

namespace ccfv {

#define CCFV_CAT2(a, b) a##b
#define CCFV_CAT(a, b) CCFV_CAT2(a, b)

// launchers of one chunk size (defined by the object built for it)
#define CCFV_GEN_DECL(CHV)                                                              \
    hipError_t launch_sweep_generic_ch##CHV(Dir dir, int bc, bool post, Epi epi,        \
                                            const SweepArgs &a, int nfields, Tile tile,  \
                                            hipStream_t s, bool fuse);                  \
    hipError_t launch_ragged_generic_ch##CHV(Dir dir, int bc, bool post, Epi epi,        \
                                             bool fuse, const RaggedItem *items,         \
                                             const int *prefix, int nitems, int nblocks, \
                                             Tile t, size_t lds, hipStream_t s);
CCFV_GEN_DECL(1)
CCFV_GEN_DECL(2)
CCFV_GEN_DECL(4)
CCFV_GEN_DECL(8)
CCFV_GEN_DECL(16)
CCFV_GEN_DECL(32)

#ifndef CCFV_GEN_CH

size_t sweep_lds_bytes_generic(int n, Tile tile, bool src, bool aux)
{
    // W tile lines (+ one auxiliary line per wave for the third-sub-step
    // variant) (+ one source line per wave)
    return sizeof(double) * (size_t)line_pitch(n) *
           (size_t)(tile.w + tile.nw * ((aux ? 1 : 0) + (src ? 1 : 0)));
}

hipError_t launch_sweep_generic(Dir dir, int bc, bool post, Epi epi,
                                const SweepArgs &a, int nfields, Tile tile,
                                hipStream_t s, bool fuse)
{
    switch (chunk_for(a.n)) {
    case 1: return launch_sweep_generic_ch1(dir, bc, post, epi, a, nfields, tile, s, fuse);
    case 2: return launch_sweep_generic_ch2(dir, bc, post, epi, a, nfields, tile, s, fuse);
    case 4: return launch_sweep_generic_ch4(dir, bc, post, epi, a, nfields, tile, s, fuse);
    case 8: return launch_sweep_generic_ch8(dir, bc, post, epi, a, nfields, tile, s, fuse);
    case 16: return launch_sweep_generic_ch16(dir, bc, post, epi, a, nfields, tile, s, fuse);
    case 32: return launch_sweep_generic_ch32(dir, bc, post, epi, a, nfields, tile, s, fuse);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_ragged_generic(Dir dir, int bc, bool post, Epi epi, int ch, bool fuse,
                                 const RaggedItem *items, const int *prefix,
                                 int nitems, int nblocks, Tile t, size_t lds,
                                 hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    switch (ch) {
#define CCFV_RG(CHV)                                                                        \
    case CHV:                                                                               \
        return launch_ragged_generic_ch##CHV(dir, bc, post, epi, fuse, items, prefix, nitems, \
                                             nblocks, t, lds, s);
        CCFV_RG(1)
        CCFV_RG(2)
        CCFV_RG(4)
        CCFV_RG(8)
        CCFV_RG(16)
        CCFV_RG(32)
#undef CCFV_RG
    }
    return hipErrorInvalidValue;
}

#else  // CCFV_GEN_CH: the kernels of one chunk size

// cs: the line in LDS (pidx layout), overwritten with the result.
// aux: a line-sized LDS buffer owned by this wave, where the RHS is staged
// unless rhs_in_regs (and, POST, the remapped line whose derivative the
// third sub-step takes); not allocated when every sweep of the kernel keeps
// the RHS in registers.
// fs: a second line-sized LDS buffer of this wave, for the source's earlier
// cell averages (st: the source's times).  SRC: the problem has a source
// (a.src_kind != 0); the source-free instance carries none of its code.
template <int DIR, int BC, int CH, bool POST, bool SRC>
__device__ void process_line(const SweepArgs &a, const Times tt, double *cs, double *aux,
                             double *fs, const SrcTimes st, int line, int lane)
{
    const int n = a.n;
    const int NL = n / CH;
    const bool active = lane < NL;
    const int base = lane * CH;
    const double inv12d = 1.0 / (12.0 * a.d);
    const LineGeo lg = line_geo(a, tt, line);
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    if (BC == BC_REF) ghosts(cs, n, g);
    constexpr bool src = SRC;
    double fg[4] = {0.0, 0.0, 0.0, 0.0};
    if (src) {
        // the source's earlier cell averages of this line (T2S4_Parareal.cu:
        // 143-165), remapped below like the solution
        if (active)
            for (int k = 0; k < CH; k++) fs[pidx(base + k)] = src_avg<DIR>(a, base + k, line, st.t0);
        __syncthreads();
        if (BC == BC_REF) ghosts(fs, n, fg);
    }

    constexpr bool regs = rhs_in_regs(CH, BC, SRC, POST);
    double b[CH];
    if (regs && active)
        rhs_general_reg<DIR, BC, SRC, CH>(a, tt, cs, g, b, base, lg, inv12d,
                                          src ? fs : nullptr, fg, line, st.t1);
    if (!regs && active)
        rhs_general<DIR, BC, SRC>(a, tt, cs, g, aux, base, CH, lg, inv12d, POST,
                                  src ? fs : nullptr, fg, line, st.t1);
    if (POST) {
        // the reference's third sub-step takes the derivative of the remapped
        // line (T2S4_Parareal.cu:929-932); the source joins afterwards
        __syncthreads();
        double gs[4] = {0.0, 0.0, 0.0, 0.0};
        if (BC == BC_REF) ghosts(aux, n, gs);
        if (active) {
            double dl = deriv_ref<BC>(a, aux, gs, face_foot<DIR>(a, tt, base, lg),
                                      inv12d);
            Face gl;
            bool bad = false;
            if (src) gl = eval_face<DIR, BC, true>(a, tt, fs, fg, base, lg, inv12d);
#pragma unroll
            for (int k = 0; k < CH; k++) {
                double dr = deriv_ref<BC>(a, aux, gs,
                                          face_foot<DIR>(a, tt, base + k + 1, lg),
                                          inv12d);
                double v = aux[pidx(base + k)];
                if (src) {
                    Face gr = eval_face<DIR, BC, true>(a, tt, fs, fg, base + k + 1, lg, inv12d);
                    v = v + src_rhs<DIR, BC>(a, tt, fs, fg, lg, line, base + k, gl, gr, st.t1,
                                             bad);
                    gl = gr;
                }
                b[k] = v + a.coef * (dr - dl);
                dl = dr;
            }
            if (bad) *a.flag = 1;
        }
        __syncthreads();
    } else if (!regs && active) {
#pragma unroll
        for (int k = 0; k < CH; k++) b[k] = aux[pidx(base + k)];
    }
    if (!active) {
#pragma unroll
        for (int k = 0; k < CH; k++) b[k] = 0.0;
    }
    const double *t = a.tab;
    solve_tail<BC, CH, false>(a, b, lane, active, t, t + n, t + 2 * n, t + 3 * n,
                              t + 4 * n, t + 4 * n + 64, cs);
}

// one block: field `field` of the launch, column tile `tile` (W lines).
// FUSE: after the sweep (and its epilogue, whose result stays in LDS) the
// same lines take the next step's first x sub-step (times from field0_2 and
// jdt2, the same step size) and go to out2.
template <int BC, int CH, bool POST, int EPI, bool FUSE, bool SRC>
__device__ __forceinline__ void sweep_x_body(const SweepArgs &a, int field, int tile,
                                             int W, int NW, double *lds)
{
    const int n = a.n, LP = line_pitch(n);
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int j0 = tile * W;
    load_tile_x(a.in + foff, a.ld, n, j0, W, lds, LP);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // per wave: an auxiliary line (unless every sweep here keeps the RHS in
    // registers), then a source line (SRC only)
    constexpr bool aux_line = POST || !rhs_in_regs(CH, BC, SRC, false);
    double *aux = lds + (W + wave) * LP;
    double *fs = lds + (W + (aux_line ? NW : 0) + wave) * LP;
    const Times tt = sub_times(a, field);
    const SrcTimes st = src_times(a, field, false);
    for (int l = wave; l < W; l += NW)
        process_line<DIR_X, BC, CH, POST, SRC>(a, tt, lds + l * LP, aux, fs, st, j0 + l, lane);
    __syncthreads();
    if (!FUSE) {
        if (EPI == EPI_NONE) store_tile_x(a.out + foff, a.ld, n, j0, W, lds, LP);
        else epilogue_x<EPI>(a, n, j0, W, lds, LP, tile);
        return;
    }
    if (EPI != EPI_NONE) {
        epilogue_x<EPI, true>(a, n, j0, W, lds, LP, tile);
        __syncthreads();
    }
    const Times t2 = sub_times_next(a, field);
    const SrcTimes st2 = src_times(a, field, true);
    for (int l = wave; l < W; l += NW)
        process_line<DIR_X, BC, CH, false, SRC>(a, t2, lds + l * LP, aux, fs, st2, j0 + l,
                                                lane);
    __syncthreads();
    store_tile_x(a.out2 + foff, a.ld, n, j0, W, lds, LP);
}

template <int BC, int CH, bool SRC>
__device__ __forceinline__ void sweep_y_body(const SweepArgs &a, int field, int tile,
                                             int W, int NW, double *lds)
{
    const int n = a.n, LP = line_pitch(n);
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int i0 = tile * W;
    load_tile_y(a.in + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // per wave: an auxiliary line (unless the RHS stays in registers), then
    // a source line (SRC only)
    constexpr bool aux_line = !rhs_in_regs(CH, BC, SRC, false);
    double *aux = lds + (W + wave) * LP;
    double *fs = lds + (W + (aux_line ? NW : 0) + wave) * LP;
    const Times tt = sub_times(a, field);
    const SrcTimes st = src_times(a, field, false);
    for (int l = wave; l < W; l += NW)
        process_line<DIR_Y, BC, CH, false, SRC>(a, tt, lds + l * LP, aux, fs, st, i0 + l, lane);
    __syncthreads();
    store_tile_y(a.out + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
}

// SRC: problems with a source term take their own instances, so the
// source-free kernels carry none of its code or registers
template <int BC, int CH, bool POST, int EPI, bool FUSE, bool SRC>
__global__ __launch_bounds__(512) void sweep_x_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    sweep_x_body<BC, CH, POST, EPI, FUSE, SRC>(a, blockIdx.y,
                                               xcd_tile(blockIdx.x, gridDim.x), W, NW, lds);
}

template <int BC, int CH, bool SRC>
__global__ __launch_bounds__(512) void sweep_y_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    sweep_y_body<BC, CH, SRC>(a, blockIdx.y, blockIdx.x, W, NW, lds);
}

template <int BC, int CH, bool POST, int EPI, bool FUSE>
__global__ __launch_bounds__(512) void ragged_x_kernel(const RaggedItem *items,
                                                       const int *prefix, int nitems,
                                                       int W, int NW)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = ragged_item(prefix, nitems, b);
    const int local = ragged_xcd_tile(b - prefix[i], prefix[i], prefix[i + 1] - prefix[i]);
    const int nt = items[i].ntiles;
    const int field = local / nt;
    // the arguments stay in global memory (scalar-cached); a private copy of
    // the struct trips the gfx950 backend for this kernel
    const SweepArgs &a = items[i].a;
    // the items of a ragged launch may mix problems with and without a source
    if (a.src_kind != CCFV_SRC_NONE)
        sweep_x_body<BC, CH, POST, EPI, FUSE, true>(a, field, local - field * nt, W, NW, lds);
    else
        sweep_x_body<BC, CH, POST, EPI, FUSE, false>(a, field, local - field * nt, W, NW, lds);
}

template <int BC, int CH>
__global__ __launch_bounds__(512) void ragged_y_kernel(const RaggedItem *items,
                                                       const int *prefix, int nitems,
                                                       int W, int NW)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = ragged_item(prefix, nitems, b);
    const int local = b - prefix[i];
    const int nt = items[i].ntiles;
    const int field = local / nt;
    const SweepArgs &a = items[i].a;
    if (a.src_kind != CCFV_SRC_NONE)
        sweep_y_body<BC, CH, true>(a, field, local - field * nt, W, NW, lds);
    else
        sweep_y_body<BC, CH, false>(a, field, local - field * nt, W, NW, lds);
}

template <typename K>
static hipError_t launch(K k, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                         const SweepArgs &a, Tile t)
{
    hipLaunchKernelGGL(k, grid, block, lds, s, a, t.w, t.nw);
    return hipGetLastError();
}

template <int BC, int CH, bool POST, bool FUSE, bool SRC>
static hipError_t dispatch_x(Epi epi, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                             const SweepArgs &a, Tile t)
{
    if (epi == EPI_NONE)
        return launch(sweep_x_kernel<BC, CH, POST, EPI_NONE, FUSE, SRC>, grid, block, lds, s,
                      a, t);
    if (epi == EPI_INIT)
        return launch(sweep_x_kernel<BC, CH, POST, EPI_INIT, FUSE, SRC>, grid, block, lds, s,
                      a, t);
    return launch(sweep_x_kernel<BC, CH, POST, EPI_CORRECT, FUSE, SRC>, grid, block, lds, s,
                  a, t);
}

template <int BC, int CH, bool SRC>
static hipError_t dispatch_src(Dir dir, bool post, Epi epi, bool fuse, const SweepArgs &a,
                               int nf, Tile t, hipStream_t s)
{
    dim3 grid(a.nl / t.w, nf), block(64 * t.nw);
    size_t lds = sweep_lds_bytes_generic(
        a.n, t, SRC, (dir == DIR_X && BC == BC_REF && post) || !rhs_in_regs(CH, BC, SRC, false));
    if (dir == DIR_Y) return launch(sweep_y_kernel<BC, CH, SRC>, grid, block, lds, s, a, t);
    if (BC == BC_REF && post)
        return fuse ? dispatch_x<BC, CH, true, true, SRC>(epi, grid, block, lds, s, a, t)
                    : dispatch_x<BC, CH, true, false, SRC>(epi, grid, block, lds, s, a, t);
    return fuse ? dispatch_x<BC, CH, false, true, SRC>(epi, grid, block, lds, s, a, t)
                : dispatch_x<BC, CH, false, false, SRC>(epi, grid, block, lds, s, a, t);
}

template <int BC, int CH>
static hipError_t dispatch(Dir dir, bool post, Epi epi, bool fuse, const SweepArgs &a,
                           int nf, Tile t, hipStream_t s)
{
    if (a.src_kind != CCFV_SRC_NONE)
        return dispatch_src<BC, CH, true>(dir, post, epi, fuse, a, nf, t, s);
    return dispatch_src<BC, CH, false>(dir, post, epi, fuse, a, nf, t, s);
}

hipError_t CCFV_CAT(launch_sweep_generic_ch, CCFV_GEN_CH)(Dir dir, int bc, bool post, Epi epi,
                                                        const SweepArgs &a, int nfields,
                                                        Tile tile, hipStream_t s, bool fuse)
{
    if (bc == BC_PER)
        return dispatch<BC_PER, CCFV_GEN_CH>(dir, false, epi, fuse, a, nfields, tile, s);
    return dispatch<BC_REF, CCFV_GEN_CH>(dir, post, epi, fuse, a, nfields, tile, s);
}

template <typename K>
static hipError_t launch_r(K k, int nblocks, Tile t, size_t lds, hipStream_t s,
                           const RaggedItem *items, const int *prefix, int nitems)
{
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * t.nw), lds, s, items, prefix,
                       nitems, t.w, t.nw);
    return hipGetLastError();
}

template <int BC, int CH, bool POST, bool FUSE>
static hipError_t dispatch_rx(Epi epi, const RaggedItem *it, const int *pf, int ni, int nb,
                              Tile t, size_t lds, hipStream_t s)
{
    if (epi == EPI_NONE)
        return launch_r(ragged_x_kernel<BC, CH, POST, EPI_NONE, FUSE>, nb, t, lds, s, it, pf,
                        ni);
    if (epi == EPI_INIT)
        return launch_r(ragged_x_kernel<BC, CH, POST, EPI_INIT, FUSE>, nb, t, lds, s, it, pf,
                        ni);
    return launch_r(ragged_x_kernel<BC, CH, POST, EPI_CORRECT, FUSE>, nb, t, lds, s, it, pf,
                    ni);
}

template <int BC, int CH>
static hipError_t dispatch_r(Dir dir, bool post, Epi epi, bool fuse, const RaggedItem *it,
                             const int *pf, int ni, int nb, Tile t, size_t lds,
                             hipStream_t s)
{
    if (dir == DIR_Y) return launch_r(ragged_y_kernel<BC, CH>, nb, t, lds, s, it, pf, ni);
    if (BC == BC_REF && post)
        return fuse ? dispatch_rx<BC, CH, true, true>(epi, it, pf, ni, nb, t, lds, s)
                    : dispatch_rx<BC, CH, true, false>(epi, it, pf, ni, nb, t, lds, s);
    return fuse ? dispatch_rx<BC, CH, false, true>(epi, it, pf, ni, nb, t, lds, s)
                : dispatch_rx<BC, CH, false, false>(epi, it, pf, ni, nb, t, lds, s);
}

hipError_t CCFV_CAT(launch_ragged_generic_ch, CCFV_GEN_CH)(Dir dir, int bc, bool post,
                                                         Epi epi, bool fuse,
                                                         const RaggedItem *items,
                                                         const int *prefix, int nitems,
                                                         int nblocks, Tile t, size_t lds,
                                                         hipStream_t s)
{
    if (bc == BC_PER)
        return dispatch_r<BC_PER, CCFV_GEN_CH>(dir, false, epi, fuse, items, prefix, nitems,
                                               nblocks, t, lds, s);
    return dispatch_r<BC_REF, CCFV_GEN_CH>(dir, post, epi, fuse, items, prefix, nitems,
                                           nblocks, t, lds, s);
}

#endif  // CCFV_GEN_CH

}  // namespace ccfv
