// ccfv_internal.h — types shared by the gfx950 kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/ccfv.h"

namespace ccfv {

enum Dir { DIR_X = 0, DIR_Y = 1 };

// Epilogue of an x sweep (the last sub-step of a coarse step inside the
// Parareal chain).
//   EPI_NONE   : store the swept field.
//   EPI_INIT   : G cache <- g, U[n+1] <- g                 (k = 0 coarse sweep)
//   EPI_CORRECT: U[n+1] <- g + (F - G_old), G cache <- g,  increment partials
enum Epi { EPI_NONE = 0, EPI_INIT = 1, EPI_CORRECT = 2 };

// Per-(direction, step size) constants of the implicit half: tridiagonal
// coefficients and the Thomas factors of the constant matrix (identical for
// every line of a sweep, so they are computed once on the host).
struct SolveTable {
    // [mneg(n) | den(n) | pneg(n) | z(n) | mf(64) | mb(64)]: Thomas factors
    // of the constant matrix and the per-lane chunk products of mneg / pneg;
    // for split lines (fast_wpl(n) > 1) then [mfs | mbs], the products over
    // the 64 wpl lanes of 8 cells (tab_doubles(n) in all)
    double *dev = nullptr;
    // uniform-displacement path: per line the cell shift S (as double) and the
    // five filter weights; fast_ok when every line passed the edge guard
    double *line_dev = nullptr;
    bool fast_ok = false;
    double ca = 0, cb = 0;  // off-diagonal, diagonal
    double diag0 = 0;       // first diagonal entry of the factored matrix
    double coef = 0;        // explicit diffusion factor (theta*nu)*(h/d)
    double vl = 0, smden = 1;  // Sherman-Morrison (periodic)
    int sf = 6, sb = 6;        // lane-scan doubling steps (forward, backward)
    int sfs = 6, sbs = 6;      // the same for the 8-cell lanes of split lines
};

// Uniform-offset sweeps of 1024-cell lines run as split lines: two waves of
// one workgroup per line, 8 cells per lane (every other line length: one
// wave per line).  Waves of a line per line length:
__host__ __device__ constexpr int fast_wpl(int n) { return n == 1024 ? 2 : 1; }
// doubles of a solve table of lines of n cells (SolveTable::dev)
__host__ __device__ constexpr int tab_doubles(int n)
{
    return 4 * n + 128 + (fast_wpl(n) > 1 ? 128 * fast_wpl(n) : 0);
}

// Everything one sweep kernel needs, passed by value.
struct SweepArgs {
    const double *in;   // field 0 input
    double *out;        // field 0 output (may alias in: sweeps are line-local)
    long long fstride;  // doubles between consecutive fields
    int n;              // cells along a line
    int nl;             // lines (cells in the other direction)
    int ld;             // row length of the field (ny)
    int quirk;          // reference ghost selection in the derivative stencil
    int sf, sb;         // lane-scan doubling steps, forward / backward
    int sfs, sbs;       // the same within each wave of a split line
    double lo, hi, d;   // along-line box and cell width
    double olo, od;     // other-direction box start and cell width
    double eps;         // remap clip epsilon (closed box)
    double coef, ca, diag0, vl;
    double smrcp;       // Sherman-Morrison: 1 / (1 + v.z)
    const double *tab;
    const double *lineS, *lineK;  // uniform path: shift and 6 filter weights per line
    int vel_kind;
    // panel layout of the input: the field as column panels of PANEL_W
    // columns, cell (i, j) at (j / PANEL_W) * PANEL_W * nx + i * PANEL_W +
    // j % PANEL_W.  The host picks the kernel instance by these flags (the
    // uniform-offset batch sweeps, x tiles of PANEL_W lines); kernels do not
    // read them
    int in_panel;
    double v0, v1, v2;  // velocity parameters
    // time: field f of the launch sits at tn = t0 + (field0 + f)*tDT, the
    // sub-step starts at tl = tn + jdt and the velocity is sampled at
    // [t_lo, t_hi] of the sub-step (only time-dependent velocities read it)
    double t0, tDT, jdt, h;
    int field0, substep;
    double tau;
    int *flag;          // device status word (set to 1 on crossed feet)
    // epilogue (x sweeps inside the Parareal coarse chain)
    double *g_cache;      // G(U_n) cache slot
    const double *f_in;   // F(U_n^k)
    double *u_next;       // U[n+1] (read old, write new)
    double *partial;      // 2 doubles per block: sum (du)^2, sum u^2
    // fused x launch: the last x sub-step of one Strang step followed, on
    // the same lines in LDS, by the first x sub-step of the next step (its
    // field offset and time offset), whose result goes to out2
    double *out2;
    int field0_2;
    double t0_2, jdt2;
    // source term (CCFV_SRC_*): theta*h (src_w) times the remapped earlier
    // and the later cell averages
    int src_kind;
    // panel layout of the output (out, or out2 when fused); see in_panel
    int out_panel;
    double src[5];
    double src_w;
    // 1 / d: the general sweeps locate feet by multiplying with it
    double id;
    // swirl: sin(pi s_q) for the n + 1 faces of a line, then cos(pi s_q)
    // (null: evaluated per face); trig_small: pi tau <= 1/4, the traced
    // sines come from the table by angle addition
    const double *trig;
    int trig_small;
};

// columns per panel of the panel layout
constexpr int PANEL_W = 8;

struct Tile {
    int w;   // lines per block
    int nw;  // wavefronts per block
};

// cells per lane for a line of n cells: smallest power of two with 64*ch >= n
int chunk_for(int n);
bool supported_n(int n);

// One problem's share of a ragged launch: blocks prefix[i]..prefix[i+1]-1 of
// the launch sweep nfields fields of ntiles line tiles each with args a.
struct RaggedItem {
    SweepArgs a;
    int nfields;
    int ntiles;
};

// general sweeps (ccfv_sweep_generic.hip)
hipError_t launch_sweep_generic(Dir dir, int bc, bool post, Epi epi,
                                const SweepArgs &a, int nfields, Tile tile,
                                hipStream_t s, bool fuse = false);
// Whether a general sweep keeps the RHS of a lane's cells in registers
// (rhs_general_reg) or stages it in an auxiliary LDS line per wave
// (rhs_general): registers for periodic, source-free sweeps, which then fit
// in VGPRs without spilling at every chunk size (the closed box's ghosts and
// the source's terms already fill the registers); never for the
// third-sub-step variant, which needs the whole remapped line.  Kernels and
// the host's LDS sizes follow the same rule.
__host__ __device__ constexpr bool rhs_in_regs([[maybe_unused]] int ch, int bc, bool src,
                                               bool post)
{
    return !post && bc == CCFV_BC_PERIODIC && !src;
}
// aux: the kernel stages line data in an auxiliary LDS line per wave (the
// third-sub-step variant, or an RHS kept out of registers: rhs_in_regs)
size_t sweep_lds_bytes_generic(int n, Tile tile, bool src, bool aux = true);
// uniform-displacement sweeps (ccfv_sweep_fast.hip); periodic only
hipError_t launch_sweep_fast(Dir dir, Epi epi, const SweepArgs &a, int nfields,
                             Tile tile, hipStream_t s, bool fuse = false);
size_t sweep_lds_bytes_fast(int n, Tile tile);
// ragged launches over many problems of one line class (same chunk CH,
// boundary mode and path); items and prefix are device arrays.  fuse: x
// launches that also run the next step's first x sub-step (SweepArgs.out2)
hipError_t launch_ragged_fast(Dir dir, Epi epi, int ch, int wpl, bool fuse,
                              const RaggedItem *items, const int *prefix, int nitems,
                              int nblocks, Tile tile, size_t lds, hipStream_t s);
hipError_t launch_ragged_generic(Dir dir, int bc, bool post, Epi epi, int ch, bool fuse,
                                 const RaggedItem *items, const int *prefix,
                                 int nitems, int nblocks, Tile tile, size_t lds,
                                 hipStream_t s);
// elementwise corrections of many problems: item i covers blocks
// prefix[i]..prefix[i+1]-1
struct RaggedCorr {
    const double *g_new, *g_old, *f;
    double *u;
    size_t n;
    double *partial;
};
hipError_t launch_ragged_correction(const RaggedCorr *items, const int *prefix,
                                    int nitems, int nblocks, hipStream_t s);
// per-slot sums of many problems: item i reduces slots first..first+count-1
struct RaggedReduce {
    const double *partial;
    double *out;
    int seg, first, count;
};
hipError_t launch_ragged_reduce(const RaggedReduce *items, const int *prefix,
                                int nitems, int nblocks, hipStream_t s);
// copies of many fields: item i copies n doubles src -> dst
struct RaggedCopy {
    const double *src;
    double *dst;
    size_t n;
};
hipError_t launch_ragged_copy(const RaggedCopy *items, const int *prefix, int nitems,
                              int nblocks, hipStream_t s);
// element-wise kernels (ccfv_misc.hip)
hipError_t launch_correction(const double *g_new, const double *g_old,
                             const double *f, double *u, size_t n,
                             double *partial, int nblocks, hipStream_t s);
hipError_t launch_reduce_partials(const double *partial, int count,
                                  double *out2, hipStream_t s);
// per-slot sums of `seg` (num, den) partial pairs, slots first..first+count-1
hipError_t launch_reduce_segments(const double *partial, int seg, int first,
                                  int count, double *out, hipStream_t s);
hipError_t launch_fill_exact(const ccfv_problem &p, const ccfv_ic &ic, double t,
                             double *out, hipStream_t s);
hipError_t launch_copy(const double *src, double *dst, size_t n, hipStream_t s);
hipError_t launch_nonfinite_check(const double *src, size_t n, int *flag,
                                  hipStream_t s);
// sin(pi s_q), then cos(pi s_q), for the n + 1 faces s_q = lo + d q
hipError_t launch_fill_trig(double lo, double d, int n, double *out, hipStream_t s);
// launch-rate floor (instrumentation): kind 0 an empty kernel, kind 1 a copy
// in -> out with the one-field sweep's memory pattern of direction dir; one
// wave per line of n cells, nl lines, row length ld, t.nw waves per block
hipError_t launch_floor(int kind, int dir, const double *in, double *out, int n, int nl,
                        int ld, Tile t, hipStream_t s);

}  // namespace ccfv
