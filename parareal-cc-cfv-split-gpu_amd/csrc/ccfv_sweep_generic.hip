// ccfv_sweep_generic.hip — general line sweeps: every face foot traced and
// located on its own (closed reference box, time-dependent or non-uniform
// velocity, and any line the uniform path cannot take).
//
// One launch performs one sub-step of the Strang step (T2S4_Parareal.cu:
// 825-870, :873-917 or :920-963) for a batch of fields.  A block stages W
// lines in LDS; each wavefront sweeps whole lines; the tile is written back
// in place (every sweep is line-local).
#include "ccfv_line.h"

namespace ccfv {

// cs: the line in LDS (pidx layout), overwritten with the result.
// aux: a line-sized LDS buffer owned by this wave.
// fs: a second line-sized LDS buffer of this wave, for the source's earlier
// cell averages (st: the source's times; used when a.src_kind != 0).
template <int DIR, int BC, int CH, bool POST>
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
    const bool src = a.src_kind != CCFV_SRC_NONE;
    double fg[4] = {0.0, 0.0, 0.0, 0.0};
    if (src) {
        // the source's earlier cell averages of this line (T2S4_Parareal.cu:
        // 143-165), remapped below like the solution
        if (active)
            for (int k = 0; k < CH; k++) fs[pidx(base + k)] = src_avg<DIR>(a, base + k, line, st.t0);
        __syncthreads();
        if (BC == BC_REF) ghosts(fs, n, fg);
    }

    double b[CH];
    if (active)
        rhs_general<DIR, BC>(a, tt, cs, g, aux, base, CH, lg, inv12d, POST,
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
    } else if (active) {
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
template <int BC, int CH, bool POST, int EPI, bool FUSE>
__device__ __forceinline__ void sweep_x_body(const SweepArgs &a, int field, int tile,
                                             int W, int NW, double *lds)
{
    const int n = a.n, LP = line_pitch(n);
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int j0 = tile * W;
    load_tile_x(a.in + foff, a.ld, n, j0, W, lds, LP);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *aux = lds + (W + wave) * LP;
    double *fs = lds + (W + NW + wave) * LP;  // allocated with a source only
    const Times tt = sub_times(a, field);
    const SrcTimes st = src_times(a, field, false);
    for (int l = wave; l < W; l += NW)
        process_line<DIR_X, BC, CH, POST>(a, tt, lds + l * LP, aux, fs, st, j0 + l, lane);
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
        process_line<DIR_X, BC, CH, false>(a, t2, lds + l * LP, aux, fs, st2, j0 + l, lane);
    __syncthreads();
    store_tile_x(a.out2 + foff, a.ld, n, j0, W, lds, LP);
}

template <int BC, int CH>
__device__ __forceinline__ void sweep_y_body(const SweepArgs &a, int field, int tile,
                                             int W, int NW, double *lds)
{
    const int n = a.n, LP = line_pitch(n);
    const size_t foff = (size_t)field * (size_t)a.fstride;
    const int i0 = tile * W;
    load_tile_y(a.in + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *aux = lds + (W + wave) * LP;
    double *fs = lds + (W + NW + wave) * LP;  // allocated with a source only
    const Times tt = sub_times(a, field);
    const SrcTimes st = src_times(a, field, false);
    for (int l = wave; l < W; l += NW)
        process_line<DIR_Y, BC, CH, false>(a, tt, lds + l * LP, aux, fs, st, i0 + l, lane);
    __syncthreads();
    store_tile_y(a.out + foff + (size_t)i0 * a.ld, a.ld, n, W, lds, LP);
}

template <int BC, int CH, bool POST, int EPI, bool FUSE>
__global__ __launch_bounds__(512) void sweep_x_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    sweep_x_body<BC, CH, POST, EPI, FUSE>(a, blockIdx.y, xcd_tile(blockIdx.x, gridDim.x),
                                          W, NW, lds);
}

template <int BC, int CH>
__global__ __launch_bounds__(512) void sweep_y_kernel(SweepArgs a, int W, int NW)
{
    extern __shared__ double lds[];
    sweep_y_body<BC, CH>(a, blockIdx.y, blockIdx.x, W, NW, lds);
}

template <int BC, int CH, bool POST, int EPI, bool FUSE>
__global__ __launch_bounds__(512) void ragged_x_kernel(const RaggedItem *items,
                                                       const int *prefix, int nitems,
                                                       int W, int NW)
{
    extern __shared__ double lds[];
    const int b = blockIdx.x;
    const int i = ragged_item(prefix, nitems, b);
    const int local = b - prefix[i];
    const int nt = items[i].ntiles;
    const int field = local / nt;
    // the arguments stay in global memory (scalar-cached); a private copy of
    // the struct trips the gfx950 backend for this kernel
    const SweepArgs &a = items[i].a;
    sweep_x_body<BC, CH, POST, EPI, FUSE>(a, field, local - field * nt, W, NW, lds);
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
    sweep_y_body<BC, CH>(a, field, local - field * nt, W, NW, lds);
}

size_t sweep_lds_bytes_generic(int n, Tile tile, bool src)
{
    // W tile lines + one auxiliary line per wave (+ one source line per wave)
    return sizeof(double) * (size_t)line_pitch(n) *
           (size_t)(tile.w + tile.nw * (src ? 2 : 1));
}

template <typename K>
static hipError_t launch(K k, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                         const SweepArgs &a, Tile t)
{
    static bool attr = false;  // one attribute call per instantiation
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(
            (const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k, grid, block, lds, s, a, t.w, t.nw);
    return hipGetLastError();
}

template <int BC, int CH, bool POST, bool FUSE>
static hipError_t dispatch_x(Epi epi, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                             const SweepArgs &a, Tile t)
{
    if (epi == EPI_NONE)
        return launch(sweep_x_kernel<BC, CH, POST, EPI_NONE, FUSE>, grid, block, lds, s, a,
                      t);
    if (epi == EPI_INIT)
        return launch(sweep_x_kernel<BC, CH, POST, EPI_INIT, FUSE>, grid, block, lds, s, a,
                      t);
    return launch(sweep_x_kernel<BC, CH, POST, EPI_CORRECT, FUSE>, grid, block, lds, s, a,
                  t);
}

template <int BC, int CH>
static hipError_t dispatch(Dir dir, bool post, Epi epi, bool fuse, const SweepArgs &a,
                           int nf, Tile t, hipStream_t s)
{
    dim3 grid(a.nl / t.w, nf), block(64 * t.nw);
    size_t lds = sweep_lds_bytes_generic(a.n, t, a.src_kind != CCFV_SRC_NONE);
    if (dir == DIR_Y) return launch(sweep_y_kernel<BC, CH>, grid, block, lds, s, a, t);
    if (BC == BC_REF && post)
        return fuse ? dispatch_x<BC, CH, true, true>(epi, grid, block, lds, s, a, t)
                    : dispatch_x<BC, CH, true, false>(epi, grid, block, lds, s, a, t);
    return fuse ? dispatch_x<BC, CH, false, true>(epi, grid, block, lds, s, a, t)
                : dispatch_x<BC, CH, false, false>(epi, grid, block, lds, s, a, t);
}

template <int BC>
static hipError_t dispatch_bc(Dir dir, bool post, Epi epi, bool fuse, const SweepArgs &a,
                              int nf, Tile t, hipStream_t s)
{
    switch (chunk_for(a.n)) {
    case 1: return dispatch<BC, 1>(dir, post, epi, fuse, a, nf, t, s);
    case 2: return dispatch<BC, 2>(dir, post, epi, fuse, a, nf, t, s);
    case 4: return dispatch<BC, 4>(dir, post, epi, fuse, a, nf, t, s);
    case 8: return dispatch<BC, 8>(dir, post, epi, fuse, a, nf, t, s);
    case 16: return dispatch<BC, 16>(dir, post, epi, fuse, a, nf, t, s);
    case 32: return dispatch<BC, 32>(dir, post, epi, fuse, a, nf, t, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sweep_generic(Dir dir, int bc, bool post, Epi epi,
                                const SweepArgs &a, int nfields, Tile tile,
                                hipStream_t s, bool fuse)
{
    if (bc == BC_PER)
        return dispatch_bc<BC_PER>(dir, false, epi, fuse, a, nfields, tile, s);
    return dispatch_bc<BC_REF>(dir, post, epi, fuse, a, nfields, tile, s);
}

template <typename K>
static hipError_t launch_r(K k, int nblocks, Tile t, size_t lds, hipStream_t s,
                           const RaggedItem *items, const int *prefix, int nitems)
{
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute(
            (const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
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

template <int BC>
static hipError_t dispatch_rbc(Dir dir, bool post, Epi epi, bool fuse, int ch,
                               const RaggedItem *it, const int *pf, int ni, int nb, Tile t,
                               size_t lds, hipStream_t s)
{
    switch (ch) {
    case 1: return dispatch_r<BC, 1>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    case 2: return dispatch_r<BC, 2>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    case 4: return dispatch_r<BC, 4>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    case 8: return dispatch_r<BC, 8>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    case 16: return dispatch_r<BC, 16>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    case 32: return dispatch_r<BC, 32>(dir, post, epi, fuse, it, pf, ni, nb, t, lds, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_ragged_generic(Dir dir, int bc, bool post, Epi epi, int ch, bool fuse,
                                 const RaggedItem *items, const int *prefix,
                                 int nitems, int nblocks, Tile t, size_t lds,
                                 hipStream_t s)
{
    if (nblocks <= 0 || nitems <= 0) return hipSuccess;
    if (bc == BC_PER)
        return dispatch_rbc<BC_PER>(dir, false, epi, fuse, ch, items, prefix, nitems,
                                    nblocks, t, lds, s);
    return dispatch_rbc<BC_REF>(dir, post, epi, fuse, ch, items, prefix, nitems, nblocks,
                                t, lds, s);
}

}  // namespace ccfv
