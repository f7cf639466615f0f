// kernels.h - gfx950 kernels of the granular-media step (FP64 plain form, FP32 precision form).
//
// Device data layout (G = grains over all layers in layer order, P = nx*ny cells per layer):
//   grain vectors      component-major, v[c*G + g]                 (coalesced per component)
//   grain state        5 rows of G: m_x, m_y, m_z, |m| (data_s%mag), and the Ms scale the
//                      field sees (ms_scale, or clamp(|m|) with use_mag_p); the double
//                      buffer swaps move |m| with m
//   cell2grain         [k][y][x] -> global grain or -1               (get_mag scatter, data.f90:344-353)
//   spectra S1, S2     [k][c][y][kx], kx = 0..nx/2                   (x = halved axis, fftw.f90:66-84)
//   Hcell              [c][k][y][x]                                   (rx2d, ry2d, rz2d of demag.f90)
//   Tspec              [kx][q][slot][ky], q = tensor_block            (n2d of demag.f90:255 regrouped so
//                                                                     one column workgroup streams one
//                                                                     contiguous block)
#pragma once
#include "fft_lds.h"

#define MRMG_MAX_LAYERS 16

// Time function of an applied field or source after set_fld's normalisation: func 0 is
// the linear ramp ((s1 - s0) clamp((t - t0)/(t1 - t0)) + s0) * scale, otherwise
// (amp * funct(t) + dc) * scale (applied.f90:328-396, 866-874).
struct TFun {
    int func, pad_;
    double amp, gamma, phase, param, duty, moddepth, dc;
};

// funct (applied.f90:328-396): SIN, EXP, COS, the NEW pulse train and the STE trapezoid
__host__ __device__ inline double funct_eval(int func, double t, double gamm, double phase, double param, double duty,
                                             double moddepth)
{
    if (func == MRMG_FN_SIN) return sin(t * gamm + phase);
    if (func == MRMG_FN_EXP) return exp(t * gamm + phase);
    if (func == MRMG_FN_COS) return cos(t * gamm + phase);
    if (func == MRMG_FN_NEW || func == MRMG_FN_STE) {
        // the phase-shifted time folded into one period [0, gamm)
        double x = t + phase;
        if (x < 0.0) x = x + gamm * (1 + (int)fabs(x / gamm));
        x = fmod(x, gamm);
        if (func == MRMG_FN_NEW) {
            const double a = gamm / param, y = exp(duty * a), w = exp(a);
            if (x > duty * gamm) return 1.0 - moddepth + exp((gamm - x) / param) * (y - 1.0) * moddepth / (w - 1.0);
            return 1.0 + exp(-x / param) * (y - w) * moddepth / (w - 1.0);
        }
        if (x < 0.5 * param) return 2.0 * x / param;
        if (x < 0.5 * (gamm - param)) return 1.0;
        if (x < 0.5 * (gamm + param)) return (gamm - 2.0 * x) / param;
        if (x < gamm - 0.5 * param) return -1.0;
        return 2.0 * (x - gamm) / param;
    }
    return 0.0;
}

__host__ __device__ inline double tscale_eval(const TFun& f, double t0, double t1, double s0, double s1, double scale,
                                              double t)
{
    if (f.func != MRMG_FN_RAMP)
        return (f.amp * funct_eval(f.func, t, f.gamma, f.phase, f.param, f.duty, f.moddepth) + f.dc) * scale;
    double c = (t - t0) / (t1 - t0);
    c = c < 1.0 ? c : 1.0;
    c = c > 0.0 ? c : 0.0;
    return ((s1 - s0) * c + s0) * scale;
}

// A moving applied source of one layer (slot 0 field, slot 1 temperature).
struct SrcParams {
    int kind, nxp, nyp, expn;
    int poff_x, poff_y, voff;   // offsets of the node positions and values in the tables
    int pad_;
    double vel_x, vel_y, offset_x, offset_y, ref_x, ref_y, t_origin, isigx, isigy, dtemp;
    double t0, t1, s0, s1, scale;
    TFun fn;
};

struct LayerParams {
    double alpha, gamma;
    double uf_dir[3], uf_t0, uf_t1, uf_s0, uf_s1, uf_scale;
    double spin;          // S of the LLB mean field
    double tsigma;        // sqrt(2 kB thm_alpha / gamma) of get_thermal
    double floor_ms;
    double msexp, hkexp;  // temperature power laws (active when >= 0 and tscale)
    int has_uf, has_uni, has_cub, thermal;
    int off, ng;
    int use_mag;          // Ms scale = clamp(|m|) (use_mag_p with tc_nom <= 1e10)
    int tscale;           // tc_nom <= 1e10
    SrcParams src[2];
    TFun uf_fn;           // time function of the uniform field
    // READ_SCALE tables (0 Ms, 1 alpha, 2 Hk): n nodes of T/Tc at tabs + off and the
    // scalings at tabs + off + n; n = 0 without a table
    int tab_off[3], tab_n[3];
};

// the table branch of find_temp_scaling (temp_scaling.f90:397-416): linear between the
// bracketing nodes, the end values outside
__host__ __device__ inline double table_eval(const double* tt, const double* ss, int n, double x)
{
    if (x <= tt[0]) return ss[0];
    if (x >= tt[n - 1]) return ss[n - 1];
    int k1 = 0, k2 = n - 1;
    while (k2 - k1 > 1) {
        const int k = (k1 + k2) >> 1;
        if (x < tt[k]) k2 = k;
        else k1 = k;
    }
    double y = (x - tt[k1]) / (tt[k2] - tt[k1]);
    y = y < 1.0 ? (y > 0.0 ? y : 0.0) : 1.0;
    return ss[k1] + y * (ss[k2] - ss[k1]);
}
static_assert(sizeof(LayerParams) % 8 == 0, "layer parameters are copied to LDS in 8-byte words");

// Per-stage axis tables of the moving sources, [layer][slot][time][n], time 0 = the
// evaluation, 1 = the next evaluation:
//   tables:          xi, xa = set()'s node index (1-based) and weight along x;
//                    yi, ya, yb = index, s*w, s*(1-w) along y (s the time scale)
//   super-Gaussian:  xa = gx, ya = dtemp*s*gy, and xc [layer][slot][time][nx+1] the
//                    running sums of xa (xc[i] = xa[0] + ... + xa[i-1]), so that a run of
//                    cells i0 .. i0+len-1 in row j sums to ya[j] (xc[i0+len] - xc[i0])
struct SrcAxes {
    const int* xi; const double* xa; const double* xc;
    const int* yi; const double* ya; const double* yb;
    const double* pos;    // node positions
    const double* vals;   // table values
    int nx, ny;
};

// Index of the stored tensor block that maps layer ks's magnetisation onto layer ko's
// field at column kx of Tspec[kx][q][slot][ky]: the layer pair q = ko*nl + ks, or, when the
// tensor depends on the layers only through their z offset (equal thicknesses and spacings,
// no magnetic images), q = ko - ks + nl - 1 of the 2 nl - 1 offsets.
__host__ __device__ inline size_t tensor_block(int kx, int ko, int ks, int nl, int toep)
{
    return toep ? (size_t)kx * (2 * nl - 1) + (size_t)(ko - ks + nl - 1) : ((size_t)kx * nl + ko) * nl + ks;
}

// Runs: the cells of one grain that follow each other along x in one row of a layer.
// The row FFT kernels read the grains' cell magnetisation through them (get_mag,
// data.f90:344-353) and write the cell field's sums over them, which the grain kernels add
// up (get_magnetostatic, field.f90:463-468); no cell array goes to memory.  Runs are
// numbered grain by grain (each grain's runs contiguous, in the order of its cells) and
// listed a second time row by row for the row kernels.
struct RowRuns {
    const int* row_off;   // [nl*ny + 1]: the runs of row (k, y) are [row_off[k*ny + y], row_off[k*ny + y + 1])
    const int* geo;       // row by row: first x | length << 16
    const int* grain;     // row by row: global grain
    const int* gm;        // row by row: the run's number in the grain-by-grain order
    int nruns;
};

// The per-context operands of the demag kernels.  A kernel launched for one context takes
// them by value; a launch over a batch of contexts whose FFT grids and launch shapes are
// equal takes a device table of them, one entry per context, indexed by blockIdx.z.
template <typename T> struct DemagOps {
    const T* Mg;            // [3][G] grain cell values
    C2<T>* S1;              // spectra of M
    C2<T>* S2;              // spectra of H
    const C2<T>* Tspec;     // tensor spectra
    T* Hcell;               // cell field (mrmg_demag_field only)
    double* Hr;             // [3][nruns] run sums of the cell field
    RowRuns R;
    int G, pad_;
};

// rows of the grain state buffers
#define ST_MAG 3
#define ST_MSC 4
#define ST_ROWS 5

// Per-grain static data of the field assembly and push.
template <typename T> struct GrainData {
    int G;
    const int* layer_of;
    const T* ms;          // Ms
    const T* msf;         // Ms*ms_scale (scatter factor, data.f90:349)
    const T* ms_scale;
    const T* alpha_scale;
    const T* hk_scale;
    const T* hk1;
    const T* hk2;
    const T* uni;         // [3][G]
    const T* cub;         // [9][G], (r + 3c) = cubic_axis(r,c,n)
    const T* happ;        // [3][G] static per-grain applied field, or null
    const T* sigma;       // thermal std sqrt(x)*tsigma (field.f90:407-423)
    const T* vol;
    const int* gc_off;    // grain -> cells CSR (grain2gridindex): the cell count of grain g
                          // is gc_off[g+1] - gc_off[g]
    const int* run_off;   // grain -> runs CSR (RowRuns' grain-by-grain numbering)
    const int* run_geo;   //   first x | length << 16
    const int* run_row;   //   row y within the grain's layer
    int nruns;
    // exchange CSR, per grain the intralayer entries, then those to the layer above, then
    // to the layer below; x_nb/x_w carry one spare entry at the end
    const int* x_off; const int* x_nb; const T* x_w;
    const LayerParams* lp;
    int nl;
    // LLB (FP64 in both forms): Tc, lambda = lambdaMs (S+1)/S, grain temperature
    const double* tc;
    const double* lam;
    const double* temp;
    const double* tabs;   // READ_SCALE table nodes and values of every layer
    int any_src;          // some layer has a moving source
    SrcAxes ax;
};

// The grain kernel modes
enum GrainMode {
    GM_HEUN1 = 0,   // Heun predictor: field at ma, thermal draw, write mb, dm, hthm
    GM_HEUN2 = 1,   // Heun corrector: field at mb, write ma = norm(ma + (d + dm)/2), h
    GM_ROT1 = 2,    // rotational push, single stage or mid-point first half: ma -> mb
    GM_ROT2 = 3,    // rotational mid-point second half: field at mb, push from ma -> ma
    GM_FIELD1 = 4,  // adaptive dt, get_fields: field at ma + thermal draw, store h and hthm,
                    // reduce max |h + hthm dti|^2 per layer; no push
    GM_PUSH1 = 5,   // adaptive dt, Heun predictor from the stored h and hthm: ma -> mb, dm
};

// Diagnostics partials of one workgroup (doubles), per layer k < MRMG_MAX_LAYERS:
#define D_ERRMX 0                              // max |m x H|^2 / Ms^2
#define D_ERRMX1 (D_ERRMX + MRMG_MAX_LAYERS)   // max |dm|^2
#define D_MMAG (D_ERRMX1 + MRMG_MAX_LAYERS)    // sum |m| ms_scale vol
#define D_HMAG (D_MMAG + MRMG_MAX_LAYERS)      // max |h + hthm dti|^2
#define DIAG_STRIDE (D_HMAG + MRMG_MAX_LAYERS)

struct StepArgs {
    double t;          // time of the field evaluation of this stage
    double dt;         // dt0 of push_llg (full step)
    double dts;        // dt*dt_s, the push's time step
    double sqrt_dt;    // sqrt(dt0)
    int type;          // llg_type
    int noise;         // 0 none, 1 inject (normals[3][G] doubles), 2 philox
    unsigned long long seed, step;
    const double* normals;   // injected normals of this step, [G][3]
    int diag;          // write diagnostics partials
    double dti;        // 1/sqrt(previous dt * max_dmdt), get_fields' |H| estimate
    int llb;           // push_llb instead of push_llg
    double* tlast;     // adaptive path: grain temperature of the last field evaluation
    // device clock [t, step] of the current step (a step replayed from a graph): when set,
    // the evaluation time is clk[0] + t and the thermal counter clk[1] + step
    const double* clk;
};
