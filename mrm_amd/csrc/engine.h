// engine.h - host side of the granular-media time step: device buffers, the per-layer
// host copies, the context and the Engine that sequences the kernels.  The kernels and
// the Engine members that launch them live in demag.hip, grain.hip and tensor.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mrmg.h"
#include "kernels.h"

// ------------------------------------------------------------------------- errors
// the message of the calling thread's last failure (mrmg_last_error); returns code
int set_err(int code, const char* fmt, ...);

#define HIPCHK(x)                                                                             \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess)                                                                 \
            return set_err(MRMG_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,             \
                           hipGetErrorString(e_));                                            \
    } while (0)

#define RCHK(x)                        \
    do {                               \
        int r_ = (x);                  \
        if (r_ != MRMG_OK) return r_;  \
    } while (0)

// ------------------------------------------------------------------------- host side
// Stockham passes of the largest radices up to maxr: fewer, heavier passes for large
// batches, more and lighter ones when a workgroup's batch has few butterflies per pass
static inline bool factorize(int n, FFTPlan& p, int maxr = 16)
{
    p.n = n;
    p.nfac = 0;
    int m = n;
    while (maxr >= 16 && m % 16 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 16; m /= 16; }
    while (maxr >= 8 && m % 8 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 8; m /= 8; }
    while (m % 4 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nfac < MRMG_MAX_FAC) { p.fac[p.nfac++] = 5; m /= 5; }
    plan_passes(p);
    return m == 1;
}

template <typename T> struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void swap_with(DBuf& o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count)
    {
        release();
        if (count == 0) return MRMG_OK;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return set_err(MRMG_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
        }
        n = count;
        return MRMG_OK;
    }
    int upload(const T* h, size_t count, hipStream_t s)
    {
        RCHK(alloc(count));
        if (count) HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
        return MRMG_OK;
    }
    size_t bytes() const { return n * sizeof(T); }
};

struct HostLayer {
    bool set = false;
    int ng = 0, ncell = 0;
    std::vector<int> grid2grain, g2gi, g2g;
    std::vector<double> vol, ms, hk, uni, cub, ms_scale, alpha_scale, hk_scale, temp, happ;
    std::vector<int> hxi, hxg, uhxi, uhxg, lhxi, lhxg;
    std::vector<double> hx, uhx, lhx;
    double alpha = 1, gamma = 1.76e7, thm_alpha = 1, floor_ms = 0.01;
    int thermal = 0;
    bool has_uni = false, has_cub = false, has_happ = false, has_hx = false, has_uhx = false, has_lhx = false;
    mrmg_ufield uf{};
    bool has_uf = false;
    // LLB and temperature scaling: Tc, lambdaMs, S; use_mag = use_mag_p with tc_nom <= 1e10
    std::vector<double> tc, lambda_ms;
    double spin = 0.5, msexp = -1.0, hkexp = -1.0;
    bool use_mag = false, tscale = false;
    std::vector<double> tab_t[3], tab_s[3];   // READ_SCALE tables: 0 Ms, 1 alpha, 2 Hk
    // moving applied sources: slot 0 field table, slot 1 temperature (table or super-Gaussian)
    mrmg_source src[2] = {};
    std::vector<double> src_x[2], src_y[2], src_v[2];
    std::vector<double> m;     // (3,ng) host copy used before the device state exists
    std::vector<double> mag;   // (ng) |m|, likewise
};

// get_thermal's standard deviation per grain (field.f90:407-423): the LLB form (no_ms_p),
// else the |m| form when |m| != 1, else the ms_scale form
static inline double thermal_sigma(const HostLayer& H, int n, double mag, bool no_ms)
{
    if (!H.thermal) return 0.0;
    const double tsigma = sqrt(2.0 * 1.380662e-16 * H.thm_alpha / H.gamma);
    double x;
    if (no_ms) {
        x = H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n]);
    } else if (mag != 1.0) {
        x = H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n] * std::max(H.floor_ms, fabs(mag)));
    } else {
        const double msc = H.use_mag ? 1.0 : H.ms_scale[n];
        x = msc > 0.0 ? H.alpha_scale[n] * H.temp[n] / (H.ms[n] * H.vol[n] * std::max(H.floor_ms, msc)) : 0.0;
    }
    return sqrt(x) * tsigma;
}

// the Ms scale the field sees for |m| = mag (find_temp_scaling's use_mag_p form)
static inline double field_ms_scale(const HostLayer& H, int n, double mag)
{
    return H.use_mag ? std::min(1.0, std::max(0.0, fabs(mag))) : H.ms_scale[n];
}

// set_fld's normalisation of a time function (applied.f90:276-310): duty and moddepth
// clamped to [0, 1]; a function with amp == 0 or gamma == 0 becomes the constant
// amp * f(0) + dc; equal ramp ends fold into the scale
static inline TFun norm_tfun(int func, double amp, double gamma, double phase, double param, double duty, double moddepth,
                      double dc, double& s0, double& s1, double& scale)
{
    TFun f{};
    f.func = func; f.amp = amp; f.gamma = gamma; f.phase = phase; f.param = param; f.dc = dc;
    f.duty = std::min(1.0, std::max(0.0, duty));
    f.moddepth = std::min(1.0, std::max(0.0, moddepth));
    if (f.func != MRMG_FN_RAMP && (f.amp == 0.0 || f.gamma == 0.0)) {
        s0 = f.amp * funct_eval(f.func, 0.0, f.gamma, f.phase, f.param, f.duty, f.moddepth) + f.dc;
        s1 = s0;
        f.func = MRMG_FN_RAMP;
    }
    if (s0 == s1) { scale = scale * s0; s0 = 1.0; s1 = 1.0; }
    return f;
}

// checks of a time function's arguments
static inline int check_tfun(int func, double t_start, double t_end, double amp, double gamma, double param)
{
    if (func < MRMG_FN_RAMP || func > MRMG_FN_STE) return set_err(MRMG_ERR_ARG, "time function %d unknown", func);
    if (func == MRMG_FN_RAMP && t_end == t_start) return set_err(MRMG_ERR_ARG, "t_end == t_start");
    if ((func == MRMG_FN_NEW || func == MRMG_FN_STE) && amp != 0.0 && gamma != 0.0 && param == 0.0)
        return set_err(MRMG_ERR_ARG, "NEW and STE need param != 0");
    return MRMG_OK;
}

struct EngineBase {
    virtual ~EngineBase() {}
    virtual int build(std::vector<HostLayer>& L) = 0;
    virtual int set_tensor_planes(const double* d_real, bool toeplitz) = 0;
    virtual int set_state(int k, const double* m) = 0;
    virtual int get_state(int k, double* m, double* h) = 0;
    virtual int set_mag(int k, const double* mag) = 0;
    virtual int get_mag(int k, double* mag) = 0;
    virtual int update_ufield(std::vector<HostLayer>& L) = 0;
    virtual int update_sources(std::vector<HostLayer>& L) = 0;
    virtual int step(const mrmg_step_params* p, const double* noise, mrmg_diag* diag) = 0;
    virtual int demag_field(double* h_cell, double* h_grain) = 0;
    virtual size_t footprint() const = 0;
    virtual bool has_tensor() const = 0;
    virtual int profile(int enable) = 0;
    virtual int tune(const char* key, int v) = 0;
    virtual int get_tune(const char* key, int* v) = 0;
    virtual int kernel_stats(mrmg_kstat* out, int max, int* n) = 0;
    virtual int readback(const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos, double* out) = 0;
    virtual int readback_into(const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos, double* out,
                              double* plane, double* pos, double* dout) = 0;
    virtual int switch_arm(int enable, const double* dir, double thr) = 0;
    virtual int switch_read(int k, int64_t* step, double* t, double* hap) = 0;
    virtual int time_class(int cls, int reps, double* ms) = 0;
    // the tensor spectra of another context of the same form, grid and layer stack
    virtual int copy_tensor(EngineBase* src) = 0;
};

struct mrmg_ctx {
    // nx, ny: the media grid; fx, fy: the FFT grid of the demag, equal to it for periodic
    // boundaries, 2 nx x 2 ny (zero-padded) for open ones; nxh = fx/2 + 1
    int nx = 0, ny = 0, nl = 0, nxh = 0, precision = 0, demag = 1;
    int fx = 0, fy = 0;
    bool open = false;
    int device = 0;   // the device current at mrmg_create; the stream and buffers live there
    double dx = 0, dy = 0;
    std::vector<double> dz, il;
    hipStream_t stream = nullptr;
    FFTPlan px{}, py{};
    std::vector<HostLayer> layers;
    bool dirty = true;
    std::unique_ptr<EngineBase> eng;
};

static inline std::vector<C2<double>> twiddle_table(int n)
{
    std::vector<C2<double>> w(n);
    for (int k = 0; k < n; k++) {
        // exp(-2 pi i k/n), angle reduced by symmetry before the libm call
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        w[k] = {(double)cosl(a), (double)(-sinl(a))};
    }
    return w;
}

// The twiddle table of plan p in its per-pass layout (FFTPlan::toff): entry toff[i] +
// (r-1) ns[i] + k holds w_n^(k r tstep[i]) of twiddle_table; n entries, the tail unused.
static inline std::vector<C2<double>> plan_twiddles(const FFTPlan& p)
{
    const std::vector<C2<double>> w = twiddle_table(p.n);
    std::vector<C2<double>> t(p.n, C2<double>{0.0, 0.0});
    for (int i = 0; i < p.nfac; i++) {
        const int ns = p.ns[i], R = p.fac[i];
        if (ns == 1) continue;
        for (int r = 1; r < R; r++)
            for (int k = 0; k < ns; k++) t[p.toff[i] + (r - 1) * ns + k] = w[(size_t)k * r * p.tstep[i] % p.n];
    }
    return t;
}

// E: complex elements per thread of one pass, B*N <= E*threads, rounded up to 4..64
static inline int pick_e(int batch, const FFTPlan& p, int threads)
{
    const int need = (batch * p.n + threads - 1) / threads;
    for (int e = 4; e <= 64; e *= 2)
        if (need <= e) return e;
    return -1;
}

#define LAUNCH_E(EV, KERNEL, ...)                                                                     \
    do {                                                                                              \
        switch (EV) {                                                                                 \
        case 4: KERNEL<4>(__VA_ARGS__); break;                                                        \
        case 8: KERNEL<8>(__VA_ARGS__); break;                                                        \
        case 16: KERNEL<16>(__VA_ARGS__); break;                                                      \
        case 32: KERNEL<32>(__VA_ARGS__); break;                                                      \
        default: KERNEL<64>(__VA_ARGS__); break;                                                      \
        }                                                                                             \
    } while (0)

// T: the type of the demag path (cell fields, spectra, tensor, FFT).  The grain state, the
// grain constants and the field assembly are FP64 in both forms, as in the reference's
// build with single-precision FFTs (ffttr = 4, data.f90:326).

// ---- launchers of kernels that host code outside the engine starts (defined with their
// kernels: grain.hip, tensor.hip)
int launch_philox_dump(hipStream_t s, unsigned long long seed, unsigned long long step0, int nsteps, int G,
                       double* out);
int launch_copy16(hipStream_t s, unsigned grid, const uint4* a, uint4* b, size_t n);
// onx, ony > 0: open boundaries, the (nx, ny) = (2 onx, 2 ony) padded grid's displacements
int launch_plane_to_plane(hipStream_t s, int nx, int ny, double dx, double dy, double dzs, double dzo, double rz,
                          int nimg, double norm, double* blk, int mode, double fxy, double fz, int onx = 0,
                          int ony = 0);
int launch_tensor_div(hipStream_t s, double* blk, size_t n, double norm);
// z-offset (Toeplitz) and index-symmetry checks of a real tensor of (nx,ny) planes on the device
int tensor_toeplitz_l1(hipStream_t s, const double* R, size_t block_elems, int nl, double* d_out);
int tensor_sym_count(hipStream_t s, const double* R, size_t plane, int nblk, unsigned long long* d_cnt);

template <typename T>
struct Engine : EngineBase {
    using S = double;
    mrmg_ctx* c;
    int G = 0, Ctot = 0, nslot = 0;
    std::vector<int> off;
    // grain data
    DBuf<int> layer_of, gc_off, cell2grain;
    // runs (RowRuns): grain by grain (run_off per grain, run_geo, run_row) and row by row
    // (rrow_off per layer row, rrm_geo, rrm_grain, rrm_gm)
    DBuf<int> run_off, run_geo, run_row, rrow_off, rrm_geo, rrm_grain, rrm_gm;
    int nruns = 0;
    DBuf<int> x_off, x_nb;
    DBuf<S> x_w;
    DBuf<S> ms, msf, ms_scale, alpha_scale, hk_scale, hk1, hk2, uni, cub, happ, sigma, sigma_llb, vol;
    DBuf<double> tc, lam, temp, tabs;
    DBuf<LayerParams> lp;
    // state, ST_ROWS rows of G (mc: corrector output of the adaptive path until the step
    // is accepted)
    DBuf<S> ma, mb, mc, dm, hthm, h;
    // |m| on the device differs from the host copy the LLG thermal variance was built from
    bool mag_dirty = false;
    // moving sources: node positions and values of every table, per-stage axis tables,
    // and the grain temperature of the last field evaluation (adaptive path)
    DBuf<double> src_pos, src_vals, ax_xa, ax_xc, ax_ya, ax_yb, tlast;
    bool any_sg = false;   // a super-Gaussian temperature source on some layer
    DBuf<int> ax_xi, ax_yi;
    bool any_src = false;
    // demag
    DBuf<C2<T>> S1, S2, Tspec, twx, twy;
    // Mg: the cell value Ms ms_scale m of every grain, [3][G]; Hr: the cell field summed
    // over every run, [3][nruns], FP64; Hcell: the cell field [c][k][y][x], only for
    // mrmg_demag_field (cells_out)
    DBuf<T> Mg, Hcell;
    DBuf<double> Hr;
    bool cells_out = false;
    DBuf<int> slot2c;
    // launch shapes (mrmg_tune): threads of the row and column workgroups, observation-layer
    // groups per column workgroup, lanes per grain in the grain kernels
    int row_threads = 128, col_threads = 512, kg = 1, lpg = 8;
    // split column path (column-contiguous spectra): rows per row workgroup, columns per
    // column workgroup, and whether it is in use (forced by col_path or by the grid size)
    bool split = false, fused_ok = false, split_ok = false;
    bool tl = false;   // fused column kernel over the split path's column-contiguous spectra
    // path 3: the forward column FFTs as the split path's kernel (many workgroups), then the
    // fused kernel's product and inverse FFT, one observation layer group per workgroup
    bool pre = false;
    bool wv = false;   // wave-local FFTs in the transposing row kernels and the path-2 column kernel
    int TYr = 8, CBc = 1, maxb_rowT = 4, maxb_colT = 4;
    // FFT plans of the step's row and column transforms and their largest radix
    FFTPlan prow{}, pcol{};
    int row_radix = 16, col_radix = 16;
    size_t lds_rowT = 0, lds_colT = 0;
    DBuf<double> noise_buf, diag_part, stats;
    // switching record and readback buffers (allocated on first use)
    DBuf<long long> sw_step;
    DBuf<double> sw_t, sw_h, rb_plane, rb_pos, rb_out;
    bool sw_on = false;
    double sw_dir[3] = {0, 0, 0}, sw_thr = 0;
    // instrumentation: event pairs of the current mrmg_step call, per-class totals
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_class;
    size_t ev_used = 0;
    double kt_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int kt_n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool tensor_ok = false, has_happ = false;
    bool toep = false;   // Tspec holds the 2 nl - 1 z-offset blocks (tensor_block)
    bool cur_llb = false;   // the step in flight uses push_llb (selects the thermal variance)
    int maxb_row = 1, maxb_col = 1;
    size_t lds_row = 0, lds_col = 0;
    std::vector<LayerParams> hlp;

    explicit Engine(mrmg_ctx* ctx) : c(ctx) {}
    ~Engine() override
    {
        for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    }

    // kernel classes of the profile
    enum { K_ROW_FWD = 0, K_COL = 1, K_ROW_INV = 2, K_GRAIN1 = 3, K_GRAIN2 = 4, K_COLF = 5, K_PROD = 6, K_COLI = 7,
           K_NCLASS = 8 };
    int ev_begin(int cls)
    {
        if (!prof) return MRMG_OK;
        if (ev_used + 2 > ev_pool.size()) {
            for (int i = 0; i < 2; i++) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                ev_pool.push_back(e);
            }
            ev_class.resize(ev_pool.size() / 2);
        }
        ev_class[ev_used / 2] = cls;
        HIPCHK(hipEventRecord(ev_pool[ev_used], c->stream));
        return MRMG_OK;
    }
    int ev_end()
    {
        if (!prof) return MRMG_OK;
        HIPCHK(hipEventRecord(ev_pool[ev_used + 1], c->stream));
        ev_used += 2;
        return MRMG_OK;
    }
    // after the stream is synchronised: fold the recorded pairs into the totals
    int ev_collect()
    {
        for (size_t i = 0; i < ev_used; i += 2) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]));
            kt_ms[ev_class[i / 2]] += ms;
            kt_n[ev_class[i / 2]] += 1;
        }
        ev_used = 0;
        return MRMG_OK;
    }

    size_t footprint() const override
    {
        size_t b = 0;
        b += layer_of.bytes() + gc_off.bytes() + cell2grain.bytes();
        b += run_off.bytes() + run_geo.bytes() + run_row.bytes() + rrow_off.bytes() + rrm_geo.bytes() + rrm_grain.bytes() +
             rrm_gm.bytes() + Mg.bytes() + Hr.bytes();
        b += x_off.bytes() + x_nb.bytes() + x_w.bytes();
        b += ms.bytes() + msf.bytes() + ms_scale.bytes() + alpha_scale.bytes() + hk_scale.bytes() + hk1.bytes() +
             hk2.bytes() + uni.bytes() + cub.bytes() + happ.bytes() + sigma.bytes() + sigma_llb.bytes() + vol.bytes() +
             lp.bytes() + tc.bytes() + lam.bytes() + temp.bytes();
        b += ma.bytes() + mb.bytes() + mc.bytes() + dm.bytes() + hthm.bytes() + h.bytes();
        b += S1.bytes() + S2.bytes() + Tspec.bytes() + twx.bytes() + twy.bytes() + Hcell.bytes() +
             slot2c.bytes();
        b += noise_buf.bytes() + diag_part.bytes() + stats.bytes();
        b += src_pos.bytes() + src_vals.bytes() + ax_xa.bytes() + ax_xc.bytes() + ax_ya.bytes() + ax_yb.bytes() +
             ax_xi.bytes() + ax_yi.bytes() + tlast.bytes();
        b += sw_step.bytes() + sw_t.bytes() + sw_h.bytes() + rb_plane.bytes() + rb_pos.bytes() + rb_out.bytes();
        b += btab_buf.bytes() + clk_buf.bytes();
        return b;
    }
    bool has_tensor() const override { return tensor_ok; }

    int copy_tensor(EngineBase* srcb) override
    {
        Engine<T>* s = static_cast<Engine<T>*>(srcb);   // the caller checked the form
        if (!s->tensor_ok) return set_err(MRMG_ERR_STATE, "the source context has no tensor");
        HIPCHK(hipStreamSynchronize(s->c->stream));
        RCHK(set_slots(s->nslot == 6));
        RCHK(Tspec.alloc(s->Tspec.n));
        HIPCHK(hipMemcpyAsync(Tspec.p, s->Tspec.p, Tspec.bytes(), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        toep = s->toep;
        tensor_ok = true;
        return MRMG_OK;
    }

    int init_fft()
    {
        // the FFT grid (padded to 2 nx x 2 ny for open boundaries)
        const int nx = c->fx, ny = c->fy, nxh = c->nxh, nl = c->nl;
        prow = c->px;
        pcol = c->py;
        RCHK(upload_twiddles(prow, twx));
        RCHK(upload_twiddles(pcol, twy));
        // fused column path: all 3*n_lay columns of one kx in one workgroup's LDS
        lds_row = sizeof(C2<T>) * nx + sizeof(T) * 2 * LPN(3 * nx);
        lds_col = sizeof(C2<T>) * ny + sizeof(T) * 2 * LPN(3 * nl * ny);
        fused_ok = lds_row <= 160 * 1024 && lds_col <= 160 * 1024 && pick_e(3, c->px, 256) > 0 &&
                   pick_e(3 * nl, c->py, col_threads) > 0 && nl * ny <= 4 * col_threads;
        // split path: TY rows per row workgroup, CB columns per column workgroup
        const size_t lim = 160 * 1024;
        split_ok = false;
        for (int ty = 8; ty >= 2; ty /= 2) {
            lds_rowT = sizeof(T) * 2 * LPN((ty / 2) * nx);   // the rows only: twiddles come from L2
            if (lds_rowT <= lim && pick_e(ty / 2, c->px, 256) > 0 && (lds_rowT <= 64 * 1024 || ty == 2)) {
                TYr = ty;
                split_ok = true;
                break;
            }
        }
        CBc = std::max(1, std::min(16, 2048 / ny));
        CBc = std::min(CBc, 3 * nl * nxh);
        lds_colT = sizeof(T) * 2 * LPN(CBc * ny);
        split_ok = split_ok && lds_colT <= lim && pick_e(CBc, c->py, 256) > 0;
        if (!fused_ok && !split_ok)
            return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: a row or column FFT does not fit one workgroup", nx, ny,
                           nl);
        split = !fused_ok;
        // long rows run fastest with the widest row workgroups (measured at 2048 points)
        if (nx >= 1024) row_threads = 256;
        RCHK(shapes());
        RCHK(lds_attr());
        RCHK(S1.alloc((size_t)nl * 3 * ny * nxh));
        RCHK(S2.alloc((size_t)nl * 3 * ny * nxh));
        // open boundaries: the row r2c writes the media rows only; the padding rows of S1
        // stay zero for the fused column kernel (which does not write S1), and the split
        // path's column kernel reads them as zero
        if (c->open) HIPCHK(hipMemsetAsync(S1.p, 0, S1.bytes(), c->stream));
        kg = 1;
        return MRMG_OK;
    }

    // the twiddle table of plan q (plan_twiddles), rounded once to T, into d; the stream
    // is synchronised so that no kernel in flight still reads the old table
    int upload_twiddles(const FFTPlan& q, DBuf<C2<T>>& d)
    {
        const std::vector<C2<double>> w = plan_twiddles(q);
        std::vector<C2<T>> t(w.size());
        for (size_t i = 0; i < w.size(); i++) t[i] = {(T)w[i].x, (T)w[i].y};
        HIPCHK(hipStreamSynchronize(c->stream));
        RCHK(d.upload(t.data(), t.size(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));   // t is a local of this call
        return MRMG_OK;
    }

    // derived launch parameters of the current shapes
    int shapes()
    {
        const int nl = c->nl, ny = c->fy;
        // wave-local transforms: a lane's bound covers its wave's run of ceil(B / waves)
        auto row_e = [&]() {
            const int nw = row_threads / 64, b = TYr / 2;
            return wv ? pick_e((b + nw - 1) / nw, c->px, 64) : pick_e(b, c->px, row_threads);
        };
        if (split || tl) {
            maxb_rowT = row_e();
            if (maxb_rowT < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: row FFT too large for %d threads", c->fx, ny, nl,
                               row_threads);
        }
        if (split || pre) {
            maxb_colT = pick_e(CBc, c->py, 256);
            if (maxb_rowT < 0 || maxb_colT < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d: split-path FFT too large for %d threads", c->fx, ny,
                               nl, row_threads);
            if (split) return MRMG_OK;
        }
        maxb_row = pick_e(3, c->px, row_threads);
        const int ncw = col_threads / 64;
        maxb_col = tl && wv ? pick_e((3 * nl + ncw - 1) / ncw, c->py, 64) : pick_e(3 * nl, c->py, col_threads);
        if (maxb_row < 0 || maxb_col < 0)
            return set_err(MRMG_ERR_UNSUPPORTED, "grid %dx%dx%d too large for one-workgroup LDS FFTs (threads %d/%d)",
                           c->fx, ny, nl, row_threads, col_threads);
        const int per = (nl + kg - 1) / kg;
        if (col_threads > 512 && wv)
            return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 runs without wave-local FFTs");
        if (col_threads > 512 && maxb_col > 16)   // more elements per lane spill at 128 VGPRs
            return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 is for columns of up to 16 elements per lane");
        if (per * ny > (col_threads > 512 ? 2 : 4) * col_threads)
            return set_err(MRMG_ERR_UNSUPPORTED, "column workgroup of %d threads cannot hold %d products", col_threads,
                           per * ny);
        return MRMG_OK;
    }
    int tune(const char* key, int v) override
    {
        std::string k(key ? key : "");
        int old;
        if (k == "row_threads") {
            if (v != 64 && v != 128 && v != 256) return set_err(MRMG_ERR_ARG, "row_threads must be 64, 128 or 256");
            old = row_threads; row_threads = v;
            if (shapes() != MRMG_OK) { row_threads = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "col_threads") {
            if (v != 256 && v != 512 && v != 1024) return set_err(MRMG_ERR_ARG, "col_threads must be 256, 512 or 1024");
            if (v > 512 && nslot != 6)
                return set_err(MRMG_ERR_UNSUPPORTED, "col_threads 1024 needs the symmetric (6-slot) tensor");
            old = col_threads; col_threads = v;
            if (shapes() != MRMG_OK) { col_threads = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "col_groups") {
            if (v < 1 || v > c->nl) return set_err(MRMG_ERR_ARG, "col_groups must be in 1..n_lay");
            old = kg; kg = v;
            if (shapes() != MRMG_OK) { kg = old; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "lanes_per_grain") {
            if (v != 1 && v != 2 && v != 4 && v != 8 && v != 16)
                return set_err(MRMG_ERR_ARG, "lanes_per_grain must be 1, 2, 4, 8 or 16");
            lpg = v;
        } else if (k == "col_path") {
            if (v < 0 || v > 3)
                return set_err(MRMG_ERR_ARG, "col_path must be 0 (fused), 1 (split), 2 (fused, column-contiguous) or 3 "
                                             "(split forward column FFTs, fused product and inverse)");
            if (v != 1 && !fused_ok) return set_err(MRMG_ERR_UNSUPPORTED, "the fused column kernel does not fit this grid");
            if (v != 0 && !split_ok) return set_err(MRMG_ERR_UNSUPPORTED, "the transposing row kernels do not fit this grid");
            const bool os = split, ot = tl, op = pre;
            const int ok = kg;
            split = v == 1;
            tl = v == 2 || v == 3;
            pre = v == 3;
            if (pre && !op) kg = c->nl;        // one observation layer per workgroup
            if (!pre && op) kg = 1;
            if (shapes() != MRMG_OK) { split = os; tl = ot; pre = op; kg = ok; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "wave_fft") {
            if (v != 0 && v != 1) return set_err(MRMG_ERR_ARG, "wave_fft must be 0 or 1");
            const bool ow = wv;
            wv = v == 1;
            if (shapes() != MRMG_OK) { wv = ow; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else if (k == "row_radix" || k == "col_radix") {
            if (v != 4 && v != 8 && v != 16) return set_err(MRMG_ERR_ARG, "%s must be 4, 8 or 16", k.c_str());
            FFTPlan q;
            if (!factorize(k == "row_radix" ? c->fx : c->fy, q, v) || q.nfac > MRMG_MAX_FAC)
                return set_err(MRMG_ERR_UNSUPPORTED, "%s %d: no plan", k.c_str(), v);
            if (k == "row_radix") {
                RCHK(upload_twiddles(q, twx));
                prow = q;
                row_radix = v;
            } else {
                RCHK(upload_twiddles(q, twy));
                pcol = q;
                col_radix = v;
            }
        } else if (k == "row_block") {
            if (v != 2 && v != 4 && v != 8) return set_err(MRMG_ERR_ARG, "row_block must be 2, 4 or 8");
            const size_t l = sizeof(T) * 2 * LPN((v / 2) * c->fx);
            if (l > 160 * 1024 || pick_e(v / 2, c->px, 256) < 0)
                return set_err(MRMG_ERR_UNSUPPORTED, "row_block %d does not fit LDS", v);
            const int ot = TYr;
            const size_t ol = lds_rowT;
            TYr = v;
            lds_rowT = l;
            if (shapes() != MRMG_OK || lds_attr() != MRMG_OK) { TYr = ot; lds_rowT = ol; return shapes(), MRMG_ERR_UNSUPPORTED; }
        } else {
            return set_err(MRMG_ERR_ARG, "unknown tunable '%s'", k.c_str());
        }
        return MRMG_OK;
    }
    int get_tune(const char* key, int* v) override
    {
        std::string k(key ? key : "");
        if (k == "row_threads") *v = row_threads;
        else if (k == "col_threads") *v = col_threads;
        else if (k == "col_groups") *v = kg;
        else if (k == "lanes_per_grain") *v = lpg;
        else if (k == "col_path") *v = split ? 1 : pre ? 3 : tl ? 2 : 0;
        else if (k == "row_block") *v = TYr;
        else if (k == "wave_fft") *v = wv ? 1 : 0;
        else if (k == "row_radix") *v = row_radix;
        else if (k == "col_radix") *v = col_radix;
        else if (k == "tensor_offsets") *v = toep ? 1 : 0;
        else return set_err(MRMG_ERR_ARG, "unknown tunable '%s'", k.c_str());
        return MRMG_OK;
    }

    template <int MB> int lds_attr_mb();
    template <int MB> int lds_attr_col();
    template <int MB> int lds_attr_split();
    int lds_attr();

    int build(std::vector<HostLayer>& L) override
    {
        const int nl = c->nl, nx = c->nx, ny = c->ny;
        const size_t P = (size_t)nx * ny;
        // a switching record refers to the grains of the previous build (and is sized for
        // them): a rebuild ends it
        sw_on = false;
        sw_step.release();
        sw_t.release();
        sw_h.release();
        off.assign(nl + 1, 0);
        for (int k = 0; k < nl; k++) off[k + 1] = off[k] + L[k].ng;
        G = off[nl];
        Ctot = 0;
        for (int k = 0; k < nl; k++) Ctot += L[k].ncell;
        std::vector<int> h_layer(G), h_gco(G + 1), h_gci(Ctot), h_c2g(nl * P, -1);
        std::vector<double> h_ms(G), h_msf(G), h_mss(G), h_as(G), h_hks(G), h_hk1(G), h_hk2(G), h_uni(3 * (size_t)G, 0.0),
            h_cub(9 * (size_t)G, 0.0), h_happ(3 * (size_t)G, 0.0), h_sig(G, 0.0), h_sigl(G, 0.0), h_vol(G);
        std::vector<double> h_tc(G), h_lam(G), h_temp(G), h_tabs;
        std::vector<double> h_m((size_t)ST_ROWS * G);
        has_happ = false;
        hlp.assign(nl, LayerParams{});
        int q = 0;
        h_gco[0] = 0;
        for (int k = 0; k < nl; k++) {
            const HostLayer& H = L[k];
            for (int n = 0; n < H.ng; n++) {
                const int g = off[k] + n;
                h_layer[g] = k;
                for (int e = H.g2gi[n]; e < H.g2gi[n + 1]; e++) {
                    const int i = H.g2g[2 * e] - 1, j = H.g2g[2 * e + 1] - 1;
                    if (i < 0 || i >= nx || j < 0 || j >= ny)
                        return set_err(MRMG_ERR_ARG, "layer %d grain %d: cell (%d,%d) outside grid", k, n + 1, i + 1, j + 1);
                    h_gci[q++] = (int)(k * P + (size_t)j * nx + i);
                }
                h_gco[g + 1] = q;
                if (H.g2gi[n + 1] == H.g2gi[n]) return set_err(MRMG_ERR_ARG, "layer %d grain %d has no cells", k, n + 1);
                h_ms[g] = H.ms[n];
                h_mss[g] = H.ms_scale[n];
                h_msf[g] = H.ms[n] * H.ms_scale[n];
                h_as[g] = H.alpha_scale[n];
                h_hks[g] = H.hk_scale[n];
                h_hk1[g] = H.hk[2 * n];
                h_hk2[g] = H.hk[2 * n + 1];
                h_vol[g] = H.vol[n];
                if (H.has_uni) for (int cc = 0; cc < 3; cc++) h_uni[(size_t)cc * G + g] = H.uni[3 * n + cc];
                if (H.has_cub) for (int cc = 0; cc < 9; cc++) h_cub[(size_t)cc * G + g] = H.cub[9 * n + cc];
                if (H.has_happ) for (int cc = 0; cc < 3; cc++) h_happ[(size_t)cc * G + g] = H.happ[3 * n + cc];
                h_sig[g] = thermal_sigma(H, n, H.mag[n], false);
                h_sigl[g] = thermal_sigma(H, n, H.mag[n], true);
                h_tc[g] = H.tc[n];
                h_lam[g] = H.lambda_ms[n] * (H.spin + 1) / H.spin;
                h_temp[g] = H.temp[n];
                for (int cc = 0; cc < 3; cc++) h_m[(size_t)cc * G + g] = H.m[3 * n + cc];
                h_m[(size_t)ST_MAG * G + g] = H.mag[n];
                h_m[(size_t)ST_MSC * G + g] = field_ms_scale(H, n, H.mag[n]);
            }
            for (size_t p = 0; p < P; p++) {
                const int gr = H.grid2grain[p];
                if (gr < 0 || gr > H.ng) return set_err(MRMG_ERR_ARG, "layer %d grid2grain value %d out of range", k, gr);
                h_c2g[k * P + p] = gr > 0 ? off[k] + gr - 1 : -1;
            }
            has_happ |= H.has_happ;
            LayerParams& lpk = hlp[k];
            lpk.alpha = H.alpha;
            lpk.gamma = H.gamma;
            lpk.spin = H.spin;
            lpk.use_mag = H.use_mag;
            lpk.tsigma = sqrt(2.0 * 1.380662e-16 * H.thm_alpha / H.gamma);
            lpk.floor_ms = H.floor_ms;
            lpk.msexp = H.msexp;
            lpk.hkexp = H.hkexp;
            lpk.tscale = H.tscale;
            for (int i = 0; i < 3; i++) {
                lpk.tab_off[i] = (int)h_tabs.size();
                lpk.tab_n[i] = (int)H.tab_t[i].size();
                h_tabs.insert(h_tabs.end(), H.tab_t[i].begin(), H.tab_t[i].end());
                h_tabs.insert(h_tabs.end(), H.tab_s[i].begin(), H.tab_s[i].end());
            }
            lpk.has_uni = H.has_uni;
            lpk.has_cub = H.has_cub;
            lpk.thermal = H.thermal;
            lpk.off = off[k];
            lpk.ng = H.ng;
        }
        RCHK(fill_sources(L));
        RCHK(fill_ufield(L));
        // exchange CSR in global numbering: per grain the intralayer entries, then those to
        // the layer above, then to the layer below, and one spare entry at the end
        {
            std::vector<int> o(G + 1, 0), nb;
            std::vector<S> w;
            for (int k = 0; k < nl; k++) {
                const HostLayer& H = L[k];
                for (int n = 0; n < H.ng; n++) {
                    for (int which = 0; which < 3; which++) {
                        const std::vector<int>& ix = which == 0 ? H.hxi : which == 1 ? H.uhxi : H.lhxi;
                        const std::vector<int>& gx = which == 0 ? H.hxg : which == 1 ? H.uhxg : H.lhxg;
                        const std::vector<double>& vx = which == 0 ? H.hx : which == 1 ? H.uhx : H.lhx;
                        const bool has = which == 0 ? H.has_hx : which == 1 ? H.has_uhx : H.has_lhx;
                        const int tk = which == 0 ? k : which == 1 ? k - 1 : k + 1;
                        if (!has) continue;
                        if (tk < 0 || tk >= nl) return set_err(MRMG_ERR_ARG, "layer %d: interlayer CSR to missing layer", k);
                        for (int e = ix[n]; e < ix[n + 1]; e++) {
                            const int gg = gx[e];
                            if (gg < 1 || gg > L[tk].ng)
                                return set_err(MRMG_ERR_ARG, "layer %d: exchange neighbour %d out of range", k, gg);
                            nb.push_back(off[tk] + gg - 1);
                            w.push_back(vx[e]);
                        }
                    }
                    o[off[k] + n + 1] = (int)nb.size();
                }
            }
            nb.push_back(0);
            w.push_back(0.0);
            RCHK(x_off.upload(o.data(), o.size(), c->stream));
            RCHK(x_nb.upload(nb.data(), nb.size(), c->stream));
            RCHK(x_w.upload(w.data(), w.size(), c->stream));
        }

        auto upT = [&](DBuf<S>& d, const std::vector<double>& v) -> int { return d.upload(v.data(), v.size(), c->stream); };
        RCHK(layer_of.upload(h_layer.data(), G, c->stream));
        RCHK(gc_off.upload(h_gco.data(), G + 1, c->stream));
        RCHK(build_runs(h_gco, h_gci, h_layer));
        RCHK(cell2grain.upload(h_c2g.data(), h_c2g.size(), c->stream));
        RCHK(upT(ms, h_ms)); RCHK(upT(msf, h_msf)); RCHK(upT(ms_scale, h_mss)); RCHK(upT(alpha_scale, h_as));
        RCHK(upT(hk_scale, h_hks)); RCHK(upT(hk1, h_hk1)); RCHK(upT(hk2, h_hk2)); RCHK(upT(uni, h_uni));
        RCHK(upT(cub, h_cub)); RCHK(upT(sigma, h_sig)); RCHK(upT(sigma_llb, h_sigl)); RCHK(upT(vol, h_vol));
        RCHK(tc.upload(h_tc.data(), G, c->stream));
        RCHK(lam.upload(h_lam.data(), G, c->stream));
        RCHK(temp.upload(h_temp.data(), G, c->stream));
        h_tabs.push_back(0.0);   // never empty
        RCHK(tabs.upload(h_tabs.data(), h_tabs.size(), c->stream));
        if (has_happ) RCHK(upT(happ, h_happ)); else happ.release();
        RCHK(upT(ma, h_m));
        mag_dirty = false;
        RCHK(mb.alloc((size_t)ST_ROWS * G)); RCHK(mc.alloc((size_t)ST_ROWS * G)); RCHK(dm.alloc(3 * (size_t)G)); RCHK(hthm.alloc(3 * (size_t)G)); RCHK(h.alloc(3 * (size_t)G));
        HIPCHK(hipMemsetAsync(mb.p, 0, mb.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(dm.p, 0, dm.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(h.p, 0, h.bytes(), c->stream));
        RCHK(diag_part.alloc((size_t)(((size_t)G * 16 + 255) / 256) * DIAG_STRIDE));
        RCHK(stats.alloc((size_t)10 * nl));
        RCHK(Mg.alloc(3 * (size_t)G));
        HIPCHK(hipMemsetAsync(Mg.p, 0, Mg.bytes(), c->stream));
        RCHK(Hr.alloc(3 * (size_t)nruns));
        HIPCHK(hipMemsetAsync(Hr.p, 0, Hr.bytes(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    // Runs from the grains' cells (gci: global cell k*P + j*nx + i, grain by grain): each
    // grain's cells in (row, x) order, split where the row ends or x jumps.  A cell that
    // two grains claim is refused.
    int build_runs(const std::vector<int>& gco, const std::vector<int>& gci, const std::vector<int>& layer)
    {
        const int nx = c->nx, ny = c->ny, nl = c->nl;
        const long long P = (long long)nx * ny;
        std::vector<int> roff(G + 1, 0), geo, row;
        std::vector<long long> key;
        for (int g = 0; g < G; g++) {
            key.assign(gci.begin() + gco[g], gci.begin() + gco[g + 1]);
            std::sort(key.begin(), key.end());
            for (size_t a = 0; a < key.size();) {
                size_t b = a + 1;
                while (b < key.size() && key[b] == key[b - 1] + 1 && key[b] % nx != 0) b++;
                if (b < key.size() && key[b] == key[b - 1])
                    return set_err(MRMG_ERR_ARG, "layer %d grain %d lists cell %lld twice", layer[g], g + 1, key[b] % P);
                const int cell = (int)(key[a] % P);
                geo.push_back((cell % nx) | ((int)(b - a) << 16));
                row.push_back(cell / nx);
                a = b;
            }
            roff[g + 1] = (int)geo.size();
        }
        nruns = (int)geo.size();
        // the same runs row by row, each row's runs in grain order
        std::vector<int> rowoff((size_t)nl * ny + 1, 0);
        std::vector<int> grain_of(nruns);
        for (int g = 0; g < G; g++)
            for (int r = roff[g]; r < roff[g + 1]; r++) {
                grain_of[r] = g;
                rowoff[(size_t)layer[g] * ny + row[r] + 1]++;
            }
        for (size_t i = 1; i < rowoff.size(); i++) rowoff[i] += rowoff[i - 1];
        std::vector<int> cur(rowoff.begin(), rowoff.end() - 1), rgeo(nruns), rgrain(nruns), rgm(nruns);
        std::vector<char> seen((size_t)nl * P, 0);
        for (int r = 0; r < nruns; r++) {
            const int g = grain_of[r];
            const size_t rw = (size_t)layer[g] * ny + row[r];
            const int i0 = geo[r] & 0xffff, len = geo[r] >> 16;
            for (int i = i0; i < i0 + len; i++) {
                char& f = seen[rw * nx + i];
                if (f) return set_err(MRMG_ERR_ARG, "layer %d: cell (%d,%d) belongs to two grains", layer[g], i + 1, row[r] + 1);
                f = 1;
            }
            const int q = cur[rw]++;
            rgeo[q] = geo[r];
            rgrain[q] = g;
            rgm[q] = r;
        }
        geo.push_back(0);   // no array is empty
        row.push_back(0);
        rgeo.push_back(0); rgrain.push_back(0); rgm.push_back(0);
        RCHK(run_off.upload(roff.data(), roff.size(), c->stream));
        RCHK(run_geo.upload(geo.data(), geo.size(), c->stream));
        RCHK(run_row.upload(row.data(), row.size(), c->stream));
        RCHK(rrow_off.upload(rowoff.data(), rowoff.size(), c->stream));
        RCHK(rrm_geo.upload(rgeo.data(), rgeo.size(), c->stream));
        RCHK(rrm_grain.upload(rgrain.data(), rgrain.size(), c->stream));
        RCHK(rrm_gm.upload(rgm.data(), rgm.size(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));   // the host vectors go out of scope
        return MRMG_OK;
    }

    RowRuns runs() const
    {
        RowRuns R;
        R.row_off = rrow_off.p; R.geo = rrm_geo.p; R.grain = rrm_grain.p; R.gm = rrm_gm.p; R.nruns = nruns;
        return R;
    }

    // the moving sources of every layer into LayerParams and the device tables
    int fill_sources(std::vector<HostLayer>& L)
    {
        const int nl = c->nl, nx = c->nx, ny = c->ny;
        std::vector<double> pos, vals;
        any_src = false;
        for (int k = 0; k < nl; k++) {
            for (int sl = 0; sl < 2; sl++) {
                SrcParams& sp = hlp[k].src[sl];
                sp = SrcParams{};
                const mrmg_source& q = L[k].src[sl];
                sp.kind = q.kind;
                if (q.kind == MRMG_SRC_NONE) continue;
                any_src = true;
                sp.nxp = q.nxp; sp.nyp = q.nyp; sp.expn = q.expn;
                sp.vel_x = q.vel_x; sp.vel_y = q.vel_y; sp.offset_x = q.offset_x; sp.offset_y = q.offset_y;
                sp.ref_x = q.ref_x; sp.ref_y = q.ref_y; sp.t_origin = q.t_origin; sp.isigx = q.isigx; sp.isigy = q.isigy;
                sp.dtemp = q.dtemp;
                // set_temp/set_fld normalisation of a constant scale (applied.f90:307-310)
                sp.t0 = q.t_start; sp.t1 = q.t_end; sp.s0 = q.s_start; sp.s1 = q.s_end; sp.scale = q.scale;
                sp.fn = norm_tfun(q.func, q.amp, q.gamma, q.phase, q.param, q.duty, q.moddepth, q.dc, sp.s0, sp.s1,
                                  sp.scale);
                if (q.kind != MRMG_SRC_TEMP_SG) {
                    sp.poff_x = (int)pos.size();
                    pos.insert(pos.end(), L[k].src_x[sl].begin(), L[k].src_x[sl].end());
                    sp.poff_y = (int)pos.size();
                    pos.insert(pos.end(), L[k].src_y[sl].begin(), L[k].src_y[sl].end());
                    sp.voff = (int)vals.size();
                    vals.insert(vals.end(), L[k].src_v[sl].begin(), L[k].src_v[sl].end());
                }
            }
        }
        any_sg = false;
        for (int k = 0; k < nl; k++)
            for (int sl = 0; sl < 2; sl++) any_sg = any_sg || hlp[k].src[sl].kind == MRMG_SRC_TEMP_SG;
        if (!any_src) {
            src_pos.release(); src_vals.release(); ax_xa.release(); ax_xc.release(); ax_ya.release(); ax_yb.release();
            ax_xi.release(); ax_yi.release(); tlast.release();
            return MRMG_OK;
        }
        if (pos.empty()) pos.push_back(0.0);
        if (vals.empty()) vals.push_back(0.0);
        RCHK(src_pos.upload(pos.data(), pos.size(), c->stream));
        RCHK(src_vals.upload(vals.data(), vals.size(), c->stream));
        const size_t nax = (size_t)nl * 4 * nx, nay = (size_t)nl * 4 * ny;
        RCHK(ax_xa.alloc(nax)); RCHK(ax_xi.alloc(nax));
        if (any_sg) RCHK(ax_xc.alloc((size_t)nl * 4 * (nx + 1)));
        else ax_xc.release();
        RCHK(ax_ya.alloc(nay)); RCHK(ax_yb.alloc(nay)); RCHK(ax_yi.alloc(nay));
        HIPCHK(hipMemsetAsync(ax_xi.p, 0, ax_xi.bytes(), c->stream));
        HIPCHK(hipMemsetAsync(ax_yi.p, 0, ax_yi.bytes(), c->stream));
        RCHK(tlast.alloc(G));
        return MRMG_OK;
    }

    int src_axes(double t0, double t1);

    int fill_ufield(std::vector<HostLayer>& L)
    {
        for (int k = 0; k < c->nl; k++) {
            LayerParams& p = hlp[k];
            p.has_uf = L[k].has_uf;
            mrmg_ufield u = L[k].uf;
            p.uf_fn = norm_tfun(u.func, u.amp, u.gamma, u.phase, u.param, u.duty, u.moddepth, u.dc, u.s_start, u.s_end,
                                u.scale);
            for (int i = 0; i < 3; i++) p.uf_dir[i] = u.dir[i];
            p.uf_t0 = u.t_start; p.uf_t1 = u.t_end; p.uf_s0 = u.s_start; p.uf_s1 = u.s_end; p.uf_scale = u.scale;
        }
        RCHK(lp.upload(hlp.data(), hlp.size(), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    int update_ufield(std::vector<HostLayer>& L) override { return fill_ufield(L); }
    int update_sources(std::vector<HostLayer>& L) override
    {
        RCHK(fill_sources(L));
        return fill_ufield(L);
    }

    int set_tensor_planes(const double* d_real, bool toeplitz) override;

    int set_slots(bool sym)
    {
        nslot = sym ? 6 : 9;
        if (!sym && col_threads > 512) {   // the 1024-thread column kernel is built for 6 slots only
            col_threads = 512;
            RCHK(shapes());
        }
        const int s6[6] = {0, 4, 8, 1, 2, 5};   // xx yy zz xy xz yz as c = 3a+b
        int s9[9];
        for (int i = 0; i < 9; i++) s9[i] = i;
        return slot2c.upload(sym ? s6 : s9, nslot, c->stream);
    }

    int set_state(int k, const double* m) override
    {
        const int ng = c->layers[k].ng;
        std::vector<S> t(3 * (size_t)ng);
        for (int n = 0; n < ng; n++)
            for (int cc = 0; cc < 3; cc++) t[(size_t)cc * ng + n] = (S)m[3 * n + cc];
        for (int cc = 0; cc < 3; cc++)
            HIPCHK(hipMemcpyAsync(ma.p + (size_t)cc * G + off[k], t.data() + (size_t)cc * ng, sizeof(S) * ng,
                                  hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MRMG_OK;
    }

    int get_state(int k, double* m, double* hh) override
    {
        const int ng = c->layers[k].ng;
        std::vector<S> t(3 * (size_t)ng);
        for (int pass = 0; pass < 2; pass++) {
            double* out = pass == 0 ? m : hh;
            S* src = pass == 0 ? ma.p : h.p;
            if (!out) continue;
            for (int cc = 0; cc < 3; cc++)
                HIPCHK(hipMemcpyAsync(t.data() + (size_t)cc * ng, src + (size_t)cc * G + off[k], sizeof(S) * ng,
                                      hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int n = 0; n < ng; n++)
                for (int cc = 0; cc < 3; cc++) out[3 * n + cc] = (double)t[(size_t)cc * ng + n];
        }
        return MRMG_OK;
    }

    GrainData<S> gdata() const
    {
        GrainData<S> d;
        d.G = G;
        d.layer_of = layer_of.p;
        d.ms = ms.p; d.msf = msf.p; d.ms_scale = ms_scale.p; d.alpha_scale = alpha_scale.p; d.hk_scale = hk_scale.p;
        d.hk1 = hk1.p; d.hk2 = hk2.p; d.uni = uni.p; d.cub = cub.p; d.happ = has_happ ? happ.p : nullptr;
        d.sigma = sigma.p; d.vol = vol.p; d.gc_off = gc_off.p;
        d.run_off = run_off.p; d.run_geo = run_geo.p; d.run_row = run_row.p; d.nruns = nruns;
        d.x_off = x_off.p; d.x_nb = x_nb.p; d.x_w = x_w.p;
        d.nl = c->nl;
        d.lp = lp.p;
        d.tc = tc.p; d.lam = lam.p; d.temp = temp.p; d.tabs = tabs.p;
        if (cur_llb) d.sigma = sigma_llb.p;
        d.any_src = any_src;
        d.ax.xi = ax_xi.p; d.ax.xa = ax_xa.p; d.ax.xc = ax_xc.p; d.ax.yi = ax_yi.p; d.ax.ya = ax_ya.p;
        d.ax.yb = ax_yb.p;
        d.ax.pos = src_pos.p; d.ax.vals = src_vals.p; d.ax.nx = c->nx; d.ax.ny = c->ny;
        return d;
    }

    // |m| of the state: rows ST_MAG and ST_MSC of ma are the truth; the other buffers take
    // a copy before a step so that the buffer swaps of the LLG pushes (which leave |m|
    // alone) keep it
    int sync_aux()
    {
        const size_t off3 = (size_t)ST_MAG * G, n = 2 * (size_t)G;
        HIPCHK(hipMemcpyAsync(mb.p + off3, ma.p + off3, n * sizeof(S), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(mc.p + off3, ma.p + off3, n * sizeof(S), hipMemcpyDeviceToDevice, c->stream));
        return MRMG_OK;
    }
    // the LLG thermal variance depends on |m| when it is not 1: rebuild it from the device |m|
    int refresh_sigma()
    {
        std::vector<S> t(G);
        HIPCHK(hipMemcpyAsync(t.data(), ma.p + (size_t)ST_MAG * G, sizeof(S) * G, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        std::vector<S> sg(G);
        for (int k = 0; k < c->nl; k++) {
            HostLayer& H = c->layers[k];
            H.mag.resize(H.ng);
            for (int n = 0; n < H.ng; n++) {
                H.mag[n] = (double)t[off[k] + n];
                sg[off[k] + n] = (S)thermal_sigma(H, n, H.mag[n], false);
            }
        }
        RCHK(sigma.upload(sg.data(), G, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        mag_dirty = false;
        return MRMG_OK;
    }
    int set_mag(int k, const double* mg) override
    {
        const HostLayer& H = c->layers[k];
        std::vector<S> a(H.ng), b(H.ng);
        for (int n = 0; n < H.ng; n++) {
            a[n] = (S)mg[n];
            b[n] = (S)field_ms_scale(H, n, mg[n]);
        }
        HIPCHK(hipMemcpyAsync(ma.p + (size_t)ST_MAG * G + off[k], a.data(), sizeof(S) * H.ng, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ma.p + (size_t)ST_MSC * G + off[k], b.data(), sizeof(S) * H.ng, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        mag_dirty = true;
        return MRMG_OK;
    }
    int get_mag(int k, double* mg) override
    {
        const HostLayer& H = c->layers[k];
        std::vector<S> a(H.ng);
        HIPCHK(hipMemcpyAsync(a.data(), ma.p + (size_t)ST_MAG * G + off[k], sizeof(S) * H.ng, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int n = 0; n < H.ng; n++) mg[n] = (double)a[n];
        return MRMG_OK;
    }

    bool demag_on() const { return c->demag && tensor_ok; }

    int scatter(S* m, const double* t0 = nullptr);

    int launch_class(int cls);

    // get_demag: row r2c of the cell magnetisation, column FFT * tensor, row c2r
    int demag()
    {
        static const int fused[3] = {K_ROW_FWD, K_COL, K_ROW_INV};
        static const int splitc[5] = {K_ROW_FWD, K_COLF, K_PROD, K_COLI, K_ROW_INV};
        static const int prec[4] = {K_ROW_FWD, K_COLF, K_COL, K_ROW_INV};
        const int* seq = split ? splitc : pre ? prec : fused;
        const int n = split ? 5 : pre ? 4 : 3;
        for (int i = 0; i < n; i++) {
            RCHK(ev_begin(seq[i]));
            RCHK(launch_class(seq[i]));
            RCHK(ev_end());
        }
        return MRMG_OK;
    }

    // reps back-to-back launches of one demag kernel class between two events on the
    // context's stream: the kernel's average duration without a per-launch event pair.
    // Only the spectra and the run sums are written; the state is untouched.
    int time_class(int cls, int reps, double* ms) override
    {
        if (!demag_on()) return set_err(MRMG_ERR_STATE, "demag is off or has no tensor");
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev() { if (e) (void)hipEventDestroy(e); }
        } a, b;
        HIPCHK(hipEventCreate(&a.e));
        HIPCHK(hipEventCreate(&b.e));
        RCHK(launch_class(cls));   // warm
        HIPCHK(hipEventRecord(a.e, c->stream));
        for (int r = 0; r < reps; r++) RCHK(launch_class(cls));
        HIPCHK(hipEventRecord(b.e, c->stream));
        HIPCHK(hipEventSynchronize(b.e));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, a.e, b.e));
        *ms = (double)t / reps;
        return MRMG_OK;
    }
    // this context's operands of the demag kernels
    DemagOps<T> ops() const
    {
        DemagOps<T> o{};
        o.Mg = Mg.p; o.S1 = S1.p; o.S2 = S2.p; o.Tspec = Tspec.p; o.Hcell = Hcell.p; o.Hr = Hr.p;
        o.R = runs();
        o.G = G;
        return o;
    }
    // batched demag launches (serving): while btab is set, launch_class runs one launch
    // for the bn contexts of the device table btab (this context's shapes, their operands)
    const DemagOps<T>* btab = nullptr;
    int bn = 1;
    DBuf<DemagOps<T>> btab_buf;
    DBuf<double> clk_buf;   // device clocks [t, step] of a batch's members
    // The product kernel streams its operands (spectra in and out, the tensor) non-temporally
    // when together they exceed the 256 MB last-level cache of the device, which they then
    // could not stay in between launches anyway (2048x2048x8 in FP32: 2.3 GB, measured 391
    // against 454 us); smaller problems keep them cacheable.
    bool prod_streams() const { return S1.bytes() + S2.bytes() + Tspec.bytes() > (size_t)256 << 20; }
    // transposing row kernels: (ny/TY) * 3 n_lay blocks on a 1D grid padded to a multiple of 8
    unsigned rowT_blocks() const { return (unsigned)(((c->ny + TYr - 1) / TYr * 3 * c->nl + 7) / 8 * 8); }
    template <int MB> void launch_rowT_fwd(dim3 g, dim3 b);
    template <int MB> void launch_colT_fwd(dim3 g, dim3 b);
    template <int MB> void launch_colT_inv(dim3 g, dim3 b);
    template <int MB> void launch_rowT_inv(dim3 g, dim3 b);
    template <int MB> void launch_row_fwd(dim3 g, dim3 b);
    template <int MB> void launch_col_sym(int nb8);
    template <int MB> void launch_col_gen(int nb8);
    template <int MB> void launch_row_inv(dim3 g, dim3 b);

    template <int MODE, int LPG> void launch_grain(const StepArgs& a, S* mout);
    template <int MODE> int grain(const StepArgs& a, S* mout = nullptr, double t_next = 0.0);

    // per-layer reductions of the last diag-enabled grain launch
    struct LayerRed {
        std::vector<double> errmx, errmx1, mmag, hmag;
    };
    int read_partials(LayerRed& r)
    {
        const int nb = (int)(((size_t)G * lpg + 255) / 256), nl = c->nl;
        std::vector<double> part((size_t)nb * DIAG_STRIDE);
        HIPCHK(hipMemcpyAsync(part.data(), diag_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        r.errmx.assign(nl, 0.0); r.errmx1.assign(nl, 0.0); r.mmag.assign(nl, 0.0); r.hmag.assign(nl, 0.0);
        for (int b = 0; b < nb; b++) {
            const double* q = &part[(size_t)b * DIAG_STRIDE];
            for (int k = 0; k < nl; k++) {
                r.errmx[k] = std::max(r.errmx[k], q[D_ERRMX + k]);
                r.errmx1[k] = std::max(r.errmx1[k], q[D_ERRMX1 + k]);
                r.hmag[k] = std::max(r.hmag[k], q[D_HMAG + k]);
                r.mmag[k] += q[D_MMAG + k];
            }
        }
        return MRMG_OK;
    }
    int layer_stats();
    // the statistics into the diagnostics record, the layers combined as llg.f90:231-233 does
    int read_stats(mrmg_diag* d)
    {
        const int nl = c->nl;
        std::vector<double> o((size_t)10 * nl);
        HIPCHK(hipMemcpyAsync(o.data(), stats.p, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double x = 0.0;
        for (int k = 0; k < nl; k++) {
            const double* q = &o[(size_t)10 * k];
            const double x1 = q[0];
            x = x + x1;
            for (int cc = 0; cc < 3; cc++) {
                d->mave_all[cc] = d->mave_all[cc] + x1 / x * (q[1 + cc] - d->mave_all[cc]);
                d->have_all[cc] = d->have_all[cc] + x1 / x * (q[7 + cc] - d->have_all[cc]);
                d->mvar_all[cc] = d->mvar_all[cc] + x1 / x * (q[4 + cc] - d->mvar_all[cc]);
                if (k < 8) {
                    d->mave[k][cc] = q[1 + cc];
                    d->mvar[k][cc] = q[4 + cc];
                    d->have[k][cc] = q[7 + cc];
                }
            }
        }
        return MRMG_OK;
    }
    int zero_partials()
    {
        const int nb = (int)(((size_t)G * lpg + 255) / 256);
        HIPCHK(hipMemsetAsync(diag_part.p, 0, (size_t)nb * DIAG_STRIDE * sizeof(double), c->stream));
        return MRMG_OK;
    }

    // The llg() loop with get_fields' adaptive dt (llg.f90:320-388) and push_m's rejection
    // of steps that rotate a layer by more than max_dm (llg.f90:1195-1271), Heun only.
    int step_adaptive(const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
    {
        const int nl = c->nl;
        const size_t per = 3 * (size_t)G;
        const double max_dm = p->max_dm * acos(-1.0) / 180.0;
        if (p->noise == MRMG_NOISE_INJECT) RCHK(noise_buf.alloc(per));
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        if (demag_on()) RCHK(scatter(ma.p, &p->t0));
        double t = p->t0, dt = p->dt, max_dmdt = 1.0, max_dmdtnext = 1.0;
        int rejected = 0;
        LayerRed r;
        StepArgs a{};
        a.llb = p->llb;
        a.tlast = any_src ? tlast.p : nullptr;
        a.type = p->type;
        a.noise = p->noise;
        a.seed = p->seed;
        for (int it = 0; it < p->n_steps; it++) {
            a.step = p->step0 + it;
            a.normals = nullptr;
            if (p->noise == MRMG_NOISE_INJECT) {
                HIPCHK(hipMemcpyAsync(noise_buf.p, noise + per * it, sizeof(double) * per, hipMemcpyHostToDevice,
                                      c->stream));
                a.normals = noise_buf.p;
            }
            if (diag && it == p->n_steps - 1) RCHK(layer_stats());
            // get_fields(t)
            if (demag_on()) RCHK(demag());
            a.t = t;
            a.dti = 1.0 / sqrt(dt * max_dmdt);
            a.diag = 1;
            RCHK(zero_partials());
            RCHK(grain<GM_FIELD1>(a, nullptr, t));
            HIPCHK(hipStreamSynchronize(c->stream));
            RCHK(read_partials(r));
            double nd = 1.e22;
            for (int k = 0; k < nl; k++) {
                const HostLayer& H = c->layers[k];
                const double hmag = sqrt(r.hmag[k]);
                const double fact1 = (1.0 + H.alpha * H.alpha) / ((1.0 + H.alpha) * H.gamma) * p->dt_frac;
                const double fact0 = fact1 != 0.0 ? 0.0 : 1.0;
                const double cand = fact1 / std::max(1.0, hmag) + fact0 * nd;
                nd = std::min(std::min(nd, p->dt_max), cand);
            }
            dt = nd;
            // push_m(t)
            for (;;) {
                const double dtt = dt * max_dmdt;
                a.dt = dtt; a.dts = dtt; a.sqrt_dt = sqrt(dtt); a.diag = 0;
                RCHK(grain<GM_PUSH1>(a, nullptr, t + dtt));
                if (demag_on()) RCHK(demag());
                a.t = t + dtt; a.diag = 1;
                RCHK(zero_partials());
                RCHK(grain<GM_HEUN2>(a, mc.p, t + dtt));
                HIPCHK(hipStreamSynchronize(c->stream));
                RCHK(read_partials(r));
                bool too_far = false;
                max_dmdtnext = 1.e22;
                for (int k = 0; k < nl; k++) {
                    const double rot = acos(1.0 - 0.5 * r.errmx1[k]) / max_dm;
                    if (rot > 1.0) {
                        max_dmdt = 0.9 * max_dmdt / rot;
                        too_far = true;
                        break;
                    }
                    if (rot < 2.0 / 3.0) max_dmdtnext = std::min(max_dmdtnext, std::min(2.0, max_dmdt / (rot + 1.e-200)));
                    else max_dmdtnext = std::min(max_dmdtnext, max_dmdt);
                }
                if (!too_far) break;
                rejected++;
                if (rejected > 1000000) return set_err(MRMG_ERR_STATE, "push_m: no acceptable step");
            }
            ma.swap_with(mc);
            if (max_dmdtnext > 1.e20) max_dmdtnext = max_dmdt;
            dt = dt * max_dmdt;
            max_dmdt = max_dmdtnext;
            t = t + dt;
            if (sw_on) RCHK(switch_mark((long long)a.step, t));
        }
        if (diag) {
            memset(diag, 0, sizeof(*diag));
            fill_diag(diag, r, t, dt, p->n_steps);
            if (p->n_steps > 0) RCHK(read_stats(diag));
            diag->rejected = rejected;
            diag->max_dmdt = max_dmdt;
        }
        return MRMG_OK;
    }

    void fill_diag(mrmg_diag* d, const LayerRed& r, double t, double dt, int steps)
    {
        d->t = t; d->dt = dt; d->steps = steps; d->n_lay = c->nl;
        if (steps <= 0) return;
        double e = 0, e1 = 0;
        for (int k = 0; k < c->nl; k++) {
            e = std::max(e, sqrt(r.errmx[k]));
            e1 = std::max(e1, r.errmx1[k]);
            double sv = 0;
            for (int n = 0; n < c->layers[k].ng; n++) sv += c->layers[k].vol[n];
            d->mmag += r.mmag[k] / sv;
        }
        d->errmx = e;
        d->errmx1 = e1;
    }

    int step(const mrmg_step_params* p, const double* noise, mrmg_diag* diag) override
    {
        const int type = p->type;
        if (!(type == -3 || type == 2 || type == -2 || type == 1 || type == -1))
            return set_err(MRMG_ERR_UNSUPPORTED, "llg_type %d not implemented", type);
        if (p->dt <= 0.0) return set_err(MRMG_ERR_ARG, "dt must be > 0");
        if (p->noise == MRMG_NOISE_INJECT && !noise) return set_err(MRMG_ERR_ARG, "noise array is NULL");
        if (p->dt_frac != 0.0) {
            if (type != -3) return set_err(MRMG_ERR_UNSUPPORTED, "adaptive dt is implemented for Heun (llg_type -3) only");
            if (p->max_dm <= 0.0 || p->dt_max <= 0.0) return set_err(MRMG_ERR_ARG, "adaptive dt needs max_dm > 0, dt_max > 0");
        }
        if (p->llb) {
            for (int k = 0; k < c->nl; k++)
                if (!(c->layers[k].spin > 0.0))
                    return set_err(MRMG_ERR_ARG, "layer %d: LLB needs spin > 0", k);
        }
        cur_llb = p->llb != 0;
        RCHK(sync_aux());
        if (!cur_llb && mag_dirty) RCHK(refresh_sigma());
        const int rc = p->dt_frac != 0.0 ? step_adaptive(p, noise, diag) : step_fixed(p, noise, diag);
        if (cur_llb) mag_dirty = true;
        cur_llb = false;
        return rc;
    }

    int step_fixed(const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
    {
        const int type = p->type;
        // llg() start: init_applied zeroes hthm (field.f90:703)
        HIPCHK(hipMemsetAsync(hthm.p, 0, hthm.bytes(), c->stream));
        // the cells must hold the current state before the first field evaluation
        if (demag_on()) RCHK(scatter(ma.p, &p->t0));
        // injected normals go up in chunks of steps; Philox normals are drawn a chunk of
        // steps ahead by one grid-wide kernel (the same values the grain kernel would draw)
        // and then read like injected ones, which takes the draw off the grain kernel
        const bool pre = p->noise == MRMG_NOISE_PHILOX && p->n_steps > 1;
        const int chunk = pre ? 16 : 64;
        const size_t per = 3 * (size_t)G;
        if (p->noise == MRMG_NOISE_INJECT || pre) {
            // Philox: always a whole chunk, so that calls of any length reuse one buffer
            // (no allocation inside a caller's loop of short calls)
            const size_t need = per * (pre ? chunk : std::min(chunk, std::max(p->n_steps, 1)));
            if (noise_buf.n < need) RCHK(noise_buf.alloc(need));
        }
        double t = p->t0;
        // get_fields sets llgp%dt = min(dt, dt_max, dt) with a fixed step (fact1 = 0,
        // fact0 = 1, llg.f90:383-386); dt_max <= 0 means no bound here
        const double dt = p->dt_max > 0.0 ? std::min(p->dt, p->dt_max) : p->dt;
        StepArgs a{};
        a.llb = p->llb;
        a.dt = dt;
        a.sqrt_dt = sqrt(dt);
        a.type = type;
        a.noise = pre ? MRMG_NOISE_INJECT : p->noise;
        a.seed = p->seed;
        for (int it = 0; it < p->n_steps; it++) {
            const bool last = it == p->n_steps - 1;
            a.step = p->step0 + it;
            a.normals = nullptr;
            if (p->noise == MRMG_NOISE_INJECT || pre) {
                const int ci = it % chunk;
                if (ci == 0) {
                    const int n = std::min(chunk, p->n_steps - it);
                    if (pre) {
                        const size_t cnt = (size_t)n * G;
                        (void)cnt;
                        RCHK(launch_philox_dump(c->stream, p->seed, p->step0 + it, n, G, noise_buf.p));
                    } else {
                        HIPCHK(hipMemcpyAsync(noise_buf.p, noise + per * it, sizeof(double) * per * n,
                                              hipMemcpyHostToDevice, c->stream));
                    }
                }
                a.normals = noise_buf.p + per * ci;
            }
            if (last) RCHK(zero_partials());
            if (last && diag) RCHK(layer_stats());
            if (type == -3) {
                // stage 1 (get_fields at t, push iter=2)
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt; a.diag = 0;
                RCHK(grain<GM_HEUN1>(a, nullptr, t + dt));
                // stage 2 (get_field at t+dt on the predicted m, push iter=1)
                if (demag_on()) RCHK(demag());
                a.t = t + dt; a.diag = last;
                RCHK(grain<GM_HEUN2>(a, nullptr, t + dt));
            } else if (type < 0) {
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt * 0.5; a.diag = 0;
                RCHK(grain<GM_ROT1>(a, nullptr, t + dt * 0.5));
                if (demag_on()) RCHK(demag());
                a.t = t + dt * 0.5; a.dts = dt; a.diag = last;
                RCHK(grain<GM_ROT2>(a, nullptr, t + dt));
            } else {
                if (demag_on()) RCHK(demag());
                a.t = t; a.dts = dt; a.diag = last;
                RCHK(grain<GM_ROT1>(a, nullptr, t + dt));
                ma.swap_with(mb);
            }
            t = t + dt;
            if (sw_on) RCHK(switch_mark((long long)a.step, t));
            if (prof && ev_used >= 512) {
                HIPCHK(hipStreamSynchronize(c->stream));
                RCHK(ev_collect());
            }
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (prof) RCHK(ev_collect());
        if (diag) {
            memset(diag, 0, sizeof(*diag));
            LayerRed r;
            if (p->n_steps > 0) RCHK(read_partials(r));
            else { r.errmx.assign(c->nl, 0); r.errmx1.assign(c->nl, 0); r.mmag.assign(c->nl, 0); }
            fill_diag(diag, r, t, dt, p->n_steps);
            if (p->n_steps > 0) RCHK(read_stats(diag));
            diag->max_dmdt = 1.0;
        }
        return MRMG_OK;
    }

    int profile(int enable) override
    {
        prof = enable != 0;
        ev_used = 0;
        for (int i = 0; i < 8; i++) { kt_ms[i] = 0; kt_n[i] = 0; }
        return MRMG_OK;
    }
    int kernel_stats(mrmg_kstat* out, int max, int* n) override
    {
        static const char* names[K_NCLASS] = {"row_fwd", "col", "row_inv", "grain1", "grain2",
                                              "col_fwd", "prod", "col_inv"};
        *n = K_NCLASS;
        for (int i = 0; i < K_NCLASS && i < max; i++) {
            memset(out[i].name, 0, sizeof(out[i].name));
            strncpy(out[i].name, names[i], sizeof(out[i].name) - 1);
            out[i].launches = kt_n[i];
            out[i].total_ms = kt_ms[i];
        }
        return MRMG_OK;
    }

    int readback(const mrmg_sensor* sn, int n_pos, const double* xpos, const double* ypos, double* out) override
    {
        const size_t nxy = (size_t)c->nx * c->ny;
        if (rb_plane.n < nxy) RCHK(rb_plane.alloc(nxy));
        if (rb_pos.n < 2 * (size_t)n_pos) RCHK(rb_pos.alloc(2 * (size_t)n_pos));
        if (rb_out.n < (size_t)n_pos) RCHK(rb_out.alloc(n_pos));
        return readback_into(sn, n_pos, xpos, ypos, out, rb_plane.p, rb_pos.p, rb_out.p);
    }
    int readback_into(const mrmg_sensor* sn, int n_pos, const double* xpos, const double* ypos, double* out, double* plane, double* pos, double* dout) override;
    int switch_arm(int enable, const double* dir, double thr) override;
    int switch_mark(long long step, double t);
    int switch_read(int k, int64_t* step, double* t, double* hap) override
    {
        if (sw_step.n < (size_t)G) return set_err(MRMG_ERR_STATE, "switching record was never armed");
        const int ng = c->layers[k].ng;
        HIPCHK(hipStreamSynchronize(c->stream));
        if (step) {
            std::vector<long long> v(ng);
            HIPCHK(hipMemcpyAsync(v.data(), sw_step.p + off[k], sizeof(long long) * ng, hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int n = 0; n < ng; n++) step[n] = (int64_t)v[n];
        }
        if (t) {
            HIPCHK(hipMemcpyAsync(t, sw_t.p + off[k], sizeof(double) * ng, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (hap) {
            std::vector<double> v(3 * (size_t)ng);
            for (int cc = 0; cc < 3; cc++)
                HIPCHK(hipMemcpyAsync(v.data() + (size_t)cc * ng, sw_h.p + (size_t)cc * G + off[k],
                                      sizeof(double) * ng, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int n = 0; n < ng; n++)
                for (int cc = 0; cc < 3; cc++) hap[3 * (size_t)n + cc] = v[(size_t)cc * ng + n];
        }
        return MRMG_OK;
    }

    int demag_field(double* h_cell, double* h_grain) override
    {
        const size_t CP = (size_t)c->nl * c->nx * c->ny;
        // the cell field is written to memory for this call only (the step keeps run sums)
        if (Hcell.n < 3 * CP) RCHK(Hcell.alloc(3 * CP));
        if (demag_on()) {
            RCHK(scatter(ma.p));
            cells_out = true;
            const int rc = demag();
            cells_out = false;
            RCHK(rc);
        } else {
            HIPCHK(hipMemsetAsync(Hcell.p, 0, Hcell.bytes(), c->stream));
        }
        std::vector<T> t(3 * CP);
        HIPCHK(hipMemcpyAsync(t.data(), Hcell.p, t.size() * sizeof(T), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (h_cell)
            for (size_t i = 0; i < t.size(); i++) h_cell[i] = (double)t[i];
        if (h_grain) {
            int g = 0;
            for (int k = 0; k < c->nl; k++) {
                const HostLayer& H = c->layers[k];
                for (int n = 0; n < H.ng; n++, g++) {
                    double y[3] = {0, 0, 0};
                    for (int e = H.g2gi[n]; e < H.g2gi[n + 1]; e++) {
                        const size_t ci = k * (size_t)c->nx * c->ny + (size_t)(H.g2g[2 * e + 1] - 1) * c->nx + (H.g2g[2 * e] - 1);
                        for (int cc = 0; cc < 3; cc++) y[cc] += (double)t[cc * CP + ci];
                    }
                    const int cnt = H.g2gi[n + 1] - H.g2gi[n];
                    for (int cc = 0; cc < 3; cc++) h_grain[3 * g + cc] = y[cc] / cnt;
                }
            }
        }
        return MRMG_OK;
    }
};

// ---------------------------------------------------------------------------------------
// Serving: fixed-dt steps of a batch of contexts with equal grids and launch shapes.  The
// members' work runs on the first member's stream: per stage each member's grain kernel,
// and each demag kernel class once for all members (their operands in a device table,
// blockIdx.z the member).  The time and thermal step counter of every member live in a
// device clock, so one step is the same launch sequence every time; with graph set it is
// captured once as a hipGraph and replayed n_steps times.  Each member gets the values
// its own mrmg_step with Philox noise (or none) would give: the same kernels, the same
// thermal normals (drawn in the grain kernel from the member's seed and step) and the
// host loop's time arithmetic, done on the device.
int launch_clock_advance(hipStream_t s, double* clk, int n, double dt);

template <typename T>
int batch_step(std::vector<Engine<T>*>& M, const mrmg_step_params* P, int graph, double* t_end)
{
    const int B = (int)M.size();
    if (B < 1) return set_err(MRMG_ERR_ARG, "empty batch");
    Engine<T>& L = *M[0];
    const mrmg_step_params& p0 = P[0];
    const int type = p0.type;
    if (!(type == -3 || type == 2 || type == -2 || type == 1 || type == -1))
        return set_err(MRMG_ERR_UNSUPPORTED, "llg_type %d not implemented", type);
    if (graph && type > 0) return set_err(MRMG_ERR_UNSUPPORTED, "graph replay needs llg_type < 0 (no buffer swap)");
    if (p0.dt <= 0.0 || p0.n_steps < 0) return set_err(MRMG_ERR_ARG, "dt must be > 0 and n_steps >= 0");
    for (int b = 0; b < B; b++) {
        const Engine<T>& e = *M[b];
        const mrmg_step_params& p = P[b];
        if (p.n_steps != p0.n_steps || p.dt != p0.dt || p.type != type || p.llb != p0.llb || p.noise != p0.noise ||
            p.dt_max != p0.dt_max)
            return set_err(MRMG_ERR_ARG, "batch member %d: n_steps, dt, dt_max, type, llb and noise must be equal", b);
        if (p.dt_frac != 0.0) return set_err(MRMG_ERR_UNSUPPORTED, "batch member %d: adaptive dt", b);
        if (p.noise != MRMG_NOISE_NONE && p.noise != MRMG_NOISE_PHILOX)
            return set_err(MRMG_ERR_UNSUPPORTED, "batch member %d: noise must be none or Philox", b);
        if (e.c->nx != L.c->nx || e.c->ny != L.c->ny || e.c->nl != L.c->nl || e.c->fx != L.c->fx ||
            e.c->fy != L.c->fy || e.c->device != L.c->device)
            return set_err(MRMG_ERR_ARG, "batch member %d: grid or device differs", b);
        if (e.pre) return set_err(MRMG_ERR_UNSUPPORTED, "batch member %d: col_path 3 runs unbatched", b);
        if (e.demag_on() != L.demag_on() || e.split != L.split || e.tl != L.tl || e.wv || e.TYr != L.TYr ||
            e.CBc != L.CBc || e.row_threads != L.row_threads || e.col_threads != L.col_threads || e.kg != L.kg ||
            e.row_radix != L.row_radix || e.col_radix != L.col_radix || e.maxb_row != L.maxb_row ||
            e.maxb_col != L.maxb_col || e.maxb_rowT != L.maxb_rowT || e.maxb_colT != L.maxb_colT ||
            e.nslot != L.nslot || e.toep != L.toep || e.lds_row != L.lds_row || e.lds_col != L.lds_col ||
            e.lds_rowT != L.lds_rowT || e.lds_colT != L.lds_colT)
            return set_err(MRMG_ERR_ARG, "batch member %d: launch shapes or tensor storage differ", b);
        if (e.col_threads > 512) return set_err(MRMG_ERR_UNSUPPORTED, "batched column launches need col_threads <= 512");
        if (e.any_src || e.sw_on || e.prof)
            return set_err(MRMG_ERR_UNSUPPORTED, "batch member %d: moving sources, switching record or profiling", b);
        for (int k = 0; k < e.c->nl && p.llb; k++)
            if (!(e.c->layers[k].spin > 0.0)) return set_err(MRMG_ERR_ARG, "batch member %d: LLB needs spin > 0", b);
        for (int q = 0; q < b; q++)
            if (M[q] == M[b]) return set_err(MRMG_ERR_ARG, "batch member %d appears twice", b);
    }
    // every member's work goes on the lead stream for the call
    const hipStream_t s = L.c->stream;
    std::vector<hipStream_t> own(B);
    for (int b = 0; b < B; b++) {
        own[b] = M[b]->c->stream;
        if (own[b] != s) HIPCHK(hipStreamSynchronize(own[b]));
        M[b]->c->stream = s;
    }
    struct Restore {
        std::vector<Engine<T>*>& M;
        std::vector<hipStream_t>& own;
        ~Restore()
        {
            for (size_t b = 0; b < M.size(); b++) M[b]->c->stream = own[b];
        }
    } restore{M, own};
    const double dt = p0.dt_max > 0.0 ? std::min(p0.dt, p0.dt_max) : p0.dt;
    // the prologue of each member's llg() (step, step_fixed)
    std::vector<DemagOps<T>> tab(B);
    std::vector<double> clk(2 * (size_t)B);
    for (int b = 0; b < B; b++) {
        Engine<T>& e = *M[b];
        e.cur_llb = p0.llb != 0;
        RCHK(e.sync_aux());
        if (!e.cur_llb && e.mag_dirty) RCHK(e.refresh_sigma());
        HIPCHK(hipMemsetAsync(e.hthm.p, 0, e.hthm.bytes(), s));
        if (e.demag_on()) RCHK(e.scatter(e.ma.p, &P[b].t0));
        tab[b] = e.ops();
        clk[2 * b] = P[b].t0;
        clk[2 * b + 1] = (double)P[b].step0;
    }
    if (L.btab_buf.n < (size_t)B) RCHK(L.btab_buf.alloc(B));
    if (L.clk_buf.n < 2 * (size_t)B) RCHK(L.clk_buf.alloc(2 * (size_t)B));
    HIPCHK(hipMemcpyAsync(L.btab_buf.p, tab.data(), sizeof(DemagOps<T>) * B, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(L.clk_buf.p, clk.data(), sizeof(double) * 2 * B, hipMemcpyHostToDevice, s));
    std::vector<StepArgs> a(B);
    for (int b = 0; b < B; b++) {
        a[b] = StepArgs{};
        a[b].llb = p0.llb;
        a[b].dt = dt;
        a[b].sqrt_dt = sqrt(dt);
        a[b].type = type;
        a[b].noise = p0.noise;
        a[b].seed = P[b].seed;
        a[b].clk = L.clk_buf.p + 2 * b;
    }
    auto demag_all = [&]() -> int {
        if (!L.demag_on()) return MRMG_OK;
        L.btab = L.btab_buf.p;
        L.bn = B;
        const int rc = L.demag();
        L.btab = nullptr;
        L.bn = 1;
        return rc;
    };
    // one step of every member (step_fixed's loop body; stage times as offsets to the clock)
    auto one_step = [&]() -> int {
        if (type == -3) {
            RCHK(demag_all());
            for (int b = 0; b < B; b++) {
                a[b].t = 0.0; a[b].dts = dt;
                RCHK(M[b]->template grain<GM_HEUN1>(a[b], nullptr, 0.0));
            }
            RCHK(demag_all());
            for (int b = 0; b < B; b++) {
                a[b].t = dt;
                RCHK(M[b]->template grain<GM_HEUN2>(a[b], nullptr, 0.0));
            }
        } else if (type < 0) {
            RCHK(demag_all());
            for (int b = 0; b < B; b++) {
                a[b].t = 0.0; a[b].dts = dt * 0.5;
                RCHK(M[b]->template grain<GM_ROT1>(a[b], nullptr, 0.0));
            }
            RCHK(demag_all());
            for (int b = 0; b < B; b++) {
                a[b].t = dt * 0.5; a[b].dts = dt;
                RCHK(M[b]->template grain<GM_ROT2>(a[b], nullptr, 0.0));
            }
        } else {
            RCHK(demag_all());
            for (int b = 0; b < B; b++) {
                a[b].t = 0.0; a[b].dts = dt;
                RCHK(M[b]->template grain<GM_ROT1>(a[b], nullptr, 0.0));
                M[b]->ma.swap_with(M[b]->mb);
            }
        }
        return launch_clock_advance(s, L.clk_buf.p, B, dt);
    };
    int rc = MRMG_OK;
    if (graph && p0.n_steps > 0) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = one_step();
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc == MRMG_OK && ec != hipSuccess) rc = set_err(MRMG_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
        if (rc == MRMG_OK && hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess)
            rc = set_err(MRMG_ERR_HIP, "graph instantiation failed");
        for (int it = 0; rc == MRMG_OK && it < p0.n_steps; it++)
            if (hipGraphLaunch(ge, s) != hipSuccess) rc = set_err(MRMG_ERR_HIP, "graph launch failed");
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    } else {
        for (int it = 0; rc == MRMG_OK && it < p0.n_steps; it++) rc = one_step();
    }
    if (rc == MRMG_OK && t_end) {
        HIPCHK(hipMemcpyAsync(clk.data(), L.clk_buf.p, sizeof(double) * 2 * B, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    if (rc == MRMG_OK && t_end)
        for (int b = 0; b < B; b++) t_end[b] = clk[2 * b];
    for (int b = 0; b < B; b++) {
        if (M[b]->cur_llb) M[b]->mag_dirty = true;
        M[b]->cur_llb = false;
    }
    return rc;
}
