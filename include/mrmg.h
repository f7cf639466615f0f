/*
 * mrmg.h - C ABI of the MI355X (gfx950) granular-media time-stepping path.
 *
 * The library replaces the time loop of mrm's `llg_m` module together with the
 * field assembly of `field_m` and the FFT demag convolution of `demag_m`/`fftw_m`:
 *
 *   llg.f90:111-318   llg()        time loop, termination          -> mrmg_step (n_steps per call)
 *   llg.f90:320-388   get_fields() field + thermal draw + dt       -> stage-1 kernel
 *   llg.f90:1195-1271 push_m()     two-stage Heun control          -> mrmg_step
 *   llg.f90:1066-1155 push_llg()   Heun / Boris / quaternion push  -> grain stage kernels
 *   llg.f90:916-1065  push_llb()   LLB Heun / rotational push      -> grain stage kernels
 *   temp_scaling.f90:365-431 find_temp_scaling()                   -> mrmg_set_layer + grain kernels
 *   field.f90:160-273 get_field()  H = Happ + Hdemag + Hk + Hex    -> grain stage kernels
 *   field.f90:397-436 get_thermal()                                -> injected normals or Philox
 *   field.f90:438-517 get_magnetostatic() cell -> grain average    -> grain stage kernels
 *   data.f90:322-390  get_mag()    grain -> cell scatter           -> row r2c kernel
 *   demag.f90:33-135  get_demag()  r2c, k-space product, c2r       -> row/column FFT kernels
 *   demag.f90:200-286 init_demag() tensor FFT                      -> mrmg_set_tensor
 *   tensor.f90:82-216 calculate_tensor() (plane_to_plane images)   -> mrmg_build_tensor
 *
 * Conventions follow the reference's one C-ABI precedent (src/FP/fp_main.cpp:600-660,
 * src/FP/fp.f90:12-44): an opaque handle passed by value, a factory that owns it, no
 * exceptions across the boundary.  Every entry point returns an int status (0 = ok); a
 * non-zero status leaves a message readable with mrmg_last_error().
 *
 * Array arguments use the reference's own Fortran layouts and 1-based index values, so
 * the Fortran caller passes its `data_m`/`field_m` arrays unchanged:
 *   (3,ng) arrays are x,y,z-fastest; grid2grain is (nx,ny) x-fastest; grain2grid is
 *   (2,ncell) with 1-based cell coordinates; grain2gridindex is the (0:ng) CSR offset
 *   array; hxgrain/uhxgrain/lhxgrain hold 1-based grain numbers.
 * Layer 1 of the reference (the top layer) is layer index 0 here.
 * Units are the reference's CGS units (cm, emu/cc, Oe, K, s).
 */
#ifndef MRMG_H
#define MRMG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define MRMG_OK               0
#define MRMG_ERR_ARG          1  /* bad argument (null pointer, size out of range)      */
#define MRMG_ERR_HIP          2  /* HIP runtime error (no device, launch failure, ...)   */
#define MRMG_ERR_STATE        3  /* call order violated (e.g. step before set_layer)     */
#define MRMG_ERR_UNSUPPORTED  4  /* configuration outside what the kernels implement     */
#define MRMG_ERR_NOMEM        5  /* device allocation failed                             */

/* ---- precision forms -------------------------------------------------------- */
#define MRMG_FP64 0   /* plain form: every stage in FP64 (reference built with FFTD)      */
#define MRMG_FP32 1   /* precision form: cell fields, FFT and convolution in FP32; grain state,
                        field assembly and push in FP64 (the reference built with ffttr = 4)  */

/* ---- thermal noise sources --------------------------------------------------- */
#define MRMG_NOISE_NONE   0  /* no thermal field (thermal_p = .false.)                    */
#define MRMG_NOISE_INJECT 1  /* caller-supplied standard normals, (3, G, n_steps)         */
#define MRMG_NOISE_PHILOX 2  /* on-device Philox4x32-10 keyed by seed, counter = step     */

/* ---- integrator types (llg_type, llg.f90:179-199) ------------------------------ */
#define MRMG_TYPE_HEUN       (-3)  /* Heun predictor-corrector, llg.f90:1122-1139          */
#define MRMG_TYPE_BORIS        2   /* semi-implicit rotation, llg.f90:1109-1118            */
#define MRMG_TYPE_BORIS_MID  (-2)  /* Boris, mid-point two-stage form                       */
#define MRMG_TYPE_QUAT         1   /* analytic rotation with quaternions, llg.f90:1093-1107 */
#define MRMG_TYPE_QUAT_MID   (-1)

typedef struct mrmg_ctx mrmg_ctx;

/* Geometry: set_geometry (data.f90:407-413) + start_demag (demag.f90:184-198). */
typedef struct {
    int nx, ny, n_lay;        /* FFT grid = media grid (the moving window is not used) */
    int precision;            /* MRMG_FP64 or MRMG_FP32                                 */
    int demag;                /* demag_p: 0 skips the magnetostatic field entirely      */
    double dx, dy;            /* cell size (cm)                                         */
    const double *dz;         /* [n_lay] layer thickness (cm), layer index 0 = top      */
    const double *il_space;   /* [n_lay] spacing to the layer above (cm)                */
} mrmg_geom;

/* One layer's grains: the file_grain (data.f90:219-320) + file_field (field.f90:898-1008)
 * payload plus the per-layer llg/thermal parameters (data.f90:91-104). */
typedef struct {
    int ng;                          /* number of grains                                  */
    const int *grid2grain;           /* (nx,ny): 1-based grain number, 0 = non-magnetic   */
    const int *grain2gridindex;      /* (0:ng) CSR offsets into grain2grid                */
    const int *grain2grid;           /* (2,ncell): 1-based (i,j) of each grain's cells    */
    const double *grain_vol;         /* (ng) grain volume (cm^3)                           */
    const double *ms;                /* (ng) Ms (emu/cc)                                   */
    const double *hk;                /* (2,ng) Hk1, Hk2 (Oe)                               */
    const double *uni_axis;          /* (3,ng) uniaxial easy axis, or NULL                 */
    const double *cubic_axis;        /* (3,3,ng) cubic axes, or NULL                       */
    /* grain-boundary exchange, field.f90:326-363 (hx in Oe, already / (Ms_i A_i)) */
    const int *hxgrainindex;  const int *hxgrain;  const double *hx;   /* intralayer, or NULL */
    const int *uhxgrainindex; const int *uhxgrain; const double *uhx;  /* to layer index k-1  */
    const int *lhxgrainindex; const int *lhxgrain; const double *lhx;  /* to layer index k+1  */
    /* temperature scaling (temp_scaling.f90:365-431); NULL means 1 */
    const double *ms_scale, *alpha_scale, *hk_scale;
    const double *temp;              /* (ng) grain temperature (K); NULL means 0           */
    const double *h_app;             /* (3,ng) static per-grain applied field, or NULL     */
    double alpha;                    /* Gilbert damping of the push (llg namelist alpha)   */
    double gamma;                    /* gyromagnetic ratio (rad/Oe/s)                      */
    double thm_alpha;                /* alpha used by get_thermal (data_s%thm_alpha)       */
    double floor_ms;                 /* lower bound of |m| in the thermal variance         */
    int thermal;                     /* thermal_p                                          */
    /* LLB and temperature scaling (push_llb llg.f90:916-1065, temp_scaling.f90:365-431) */
    int use_mag;                     /* Ms scale = clamp(|m|, 0, 1) every field (use_mag_p) */
    const double *tc;                /* (ng) Curie temperature (K); NULL means 1e22         */
    const double *lambda_ms;         /* (ng) lambdaMs; NULL: 3 Tc / (g muB/kB (S+1))        */
    double spin;                     /* spin S of the Brillouin/Langevin mean field         */
    double tc_nom;                   /* temperature scaling on when tc_nom <= 1e10          */
    double msexp, hkexp;             /* Ms, Hk ~ (1 - T/Tc)^exp when >= 0                   */
    /* READ_SCALE tables of find_temp_scaling (temp_scaling.f90:365-431): n >= 2 nodes of
     * increasing T/Tc (tab_t) and the scaling there (tab_s), interpolated linearly and held
     * at the end values outside.  Ms uses its table when neither use_mag nor msexp >= 0 is
     * set, Hk when hkexp < 0; alpha has only the table.  n = 0: no table.  With tc_nom <=
     * 1e10 they set ms_scale, alpha_scale, hk_scale from the static temperature and, under a
     * moving temperature source, follow the grain temperature every field evaluation.     */
    int n_ms_tab, n_alpha_tab, n_hk_tab, pad3_;
    const double *ms_tab_t, *ms_tab_s;
    const double *alpha_tab_t, *alpha_tab_s;
    const double *hk_tab_t, *hk_tab_s;
} mrmg_layer;

/* Time function of an applied field or source (funct, applied.f90:328-396):
 *   MRMG_FN_RAMP (0)  the linear ramp of set_fld (applied.f90:210-320):
 *                     ((s_end - s_start) * clamp((t - t_start)/(t_end - t_start), 0, 1) + s_start) * scale
 *   otherwise         (amp * f(t) + dc) * scale with f = sin(gamma t + phase) (SIN),
 *                     exp(gamma t + phase) (EXP), cos(gamma t + phase) (COS), the
 *                     exponential pulse train of period gamma, time constant param, duty
 *                     and modulation depth (NEW), or the trapezoid wave of period gamma
 *                     and ramp width param (STE).
 * As set_fld does, a function with amp == 0 or gamma == 0 becomes the constant
 * amp * f(0) + dc, and duty and moddepth are clamped to [0, 1].  ROT is not carried. */
#define MRMG_FN_RAMP 0
#define MRMG_FN_SIN 1
#define MRMG_FN_EXP 2
#define MRMG_FN_COS 3
#define MRMG_FN_NEW 4
#define MRMG_FN_STE 5

/* Uniform applied field of one layer: the single uniform field branch of get_appl
 * (applied.f90:863-883), h(t) = scale(t) * dir with the time function above. */
typedef struct {
    double dir[3];
    double t_start, t_end, s_start, s_end, scale;
    int func;                        /* MRMG_FN_*                                             */
    int pad_;
    double amp, gamma, phase, param, duty, moddepth, dc;
} mrmg_ufield;

/* A moving applied source of one layer (get_appl, applied.f90:827-1034), sampled at the
 * cell centres every field evaluation and averaged over each grain's cells like
 * get_applied (field.f90:545-556):
 *   MRMG_SRC_FIELD_TABLE  vector field table vals (3,nxp,nyp) on increasing node
 *                         positions xpos (nxp), ypos (nyp) (cm), bilinear (set +
 *                         interpolate, applied.f90:1757-1776, 1471-1522)
 *   MRMG_SRC_TEMP_TABLE   scalar temperature table vals (nxp,nyp), likewise
 *   MRMG_SRC_TEMP_SG      super-Gaussian temperature dtemp * gx(x) * gy(y),
 *                         g(u) = exp(-((u * isig)^expn)) (set_sg/get_sg, applied.f90:1778-1811)
 * Position along each axis at time t: offset + vel*(t - t_origin) - ref (tables) and
 * ref - vel*(t - t_origin) - offset (super-Gaussian), as get_appl forms them.  The value
 * is scaled by the time function of mrmg_ufield (ramp or funct).
 * A grain's field is its static h_app + the source average + the uniform field; its
 * temperature is the source average + its static temp.  A temperature source makes the
 * temperature scalings (msexp, hkexp) and the thermal variance follow the temperature. */
#define MRMG_SRC_NONE 0
#define MRMG_SRC_FIELD_TABLE 1
#define MRMG_SRC_TEMP_TABLE 2
#define MRMG_SRC_TEMP_SG 3
typedef struct {
    int kind;
    int nxp, nyp;
    int expn;                        /* super-Gaussian power (the reference's 2*expn)        */
    const double *xpos, *ypos;       /* table nodes (cm)                                      */
    const double *vals;              /* table values, x fastest                               */
    double vel_x, vel_y;             /* cm/s                                                  */
    double offset_x, offset_y, ref_x, ref_y, t_origin;
    double isigx, isigy;             /* 1/sigma (cm^-1), 0 = flat along that axis           */
    double dtemp;                    /* super-Gaussian peak (K)                               */
    double t_start, t_end, s_start, s_end, scale;
    int func;                        /* time function, as in mrmg_ufield                      */
    int pad_;
    double amp, gamma, phase, param, duty, moddepth, dc;
} mrmg_source;

typedef struct {
    int type;                 /* MRMG_TYPE_*                                            */
    int n_steps;              /* accepted steps to run in this call                     */
    double t0;                /* simulation time at the first step (s)                  */
    double dt;                /* llgp%dt: the fixed step, or the adaptive step carried
                                 from the previous call (mrmg_diag.dt)                   */
    int noise;                /* MRMG_NOISE_*                                           */
    int pad_;
    uint64_t seed;            /* Philox key                                             */
    uint64_t step0;           /* Philox counter of the first step                       */
    double dt_frac;           /* 0: fixed dt (fixed_dt_p, llg.f90:1484-1486); else the
                                 gyro-period fraction of get_fields (llg.f90:384-386)    */
    double dt_max;            /* upper bound of dt (s): get_fields applies it to the
                                 adaptive step and to a fixed one (llg.f90:383-386);
                                 > 0 with dt_frac != 0, <= 0 means none with a fixed dt  */
    double max_dm;            /* max rotation per step (degrees) before push_m rejects
                                 and retries the step (llg.f90:1229-1236)               */
    int llb;                  /* llgp%llb: Landau-Lifshitz-Bloch push (Heun)              */
    int pad2_;
} mrmg_step_params;

/* Diagnostics of the call (llg.f90:1209-1257, 217-237).  mave, mvar and have are the
 * llg() loop-top statistics of the last step, as llgl(:)%mave etc. hold after llg()
 * returns: the state before that step's push and h from the field evaluation before it. */
typedef struct {
    double t;                 /* time after the last step                               */
    double dt;                /* llgp%dt after the last step (pass it to the next call)  */
    int steps;                /* steps completed in this call                           */
    int n_lay;
    double errmx;             /* max |m x H|/Ms over all grains (sqrt of errmx_l)        */
    double errmx1;            /* max |dm|^2                                              */
    double mmag;              /* sum over layers of volume-weighted <|m| ms_scale>       */
    double mave[8][3];        /* per-layer volume-weighted <m> (llgl%mave, up to 8)      */
    int rejected;             /* steps push_m rejected and retried in this call          */
    int pad_;
    double max_dmdt;          /* llgp%max_dmdt after the call                             */
    double mvar[8][3];        /* per-layer volume-weighted variance of m (llgl%mvar)      */
    double have[8][3];        /* per-layer volume-weighted <h> (llgl%have)                */
    double mave_all[3];       /* llgp%mave: layers combined with their total volumes     */
    double mvar_all[3];       /* llgp%mvar                                                */
    double have_all[3];       /* llgp%have                                                */
} mrmg_diag;

/* ---- lifecycle ----------------------------------------------------------------- */
const char *mrmg_last_error(void);
int mrmg_device_count(int *n);
int mrmg_create(const mrmg_geom *g, mrmg_ctx **out);
/* No counterpart in the reference, whose demag is periodic in x and y (demag.f90:113-127):
 * a context with open in-plane boundaries.  The demag is the aperiodic convolution over the
 * nx x ny media, computed with FFTs of the grid zero-padded to 2nx x 2ny (the media in the
 * first nx x ny cells, zeros elsewhere, every displacement of the media grid held once).
 * The tensor calls below then work on that padded grid: (2nx,2ny,9,n_lay,n_lay) arrays
 * divided by dx*dy*dz(ko)*2nx*2ny, num_images 0.  Everything else (layers, states, sources,
 * readback, which then uses plain instead of minimum-image distances) is on the media grid.
 * 2nx and 2ny must factor into 2, 3 and 5. */
int mrmg_create_open(const mrmg_geom *g, mrmg_ctx **out);
int mrmg_destroy(mrmg_ctx *ctx);

/* Demag tensor, real space, the (nx,ny,9,n_lay,n_lay) FP64 array of calculate_tensor
 * (tensor.f90:82-216) already divided by dx*dy*dz(ko)*nx*ny; component c = 3(a-1)+b maps
 * source component b to field component a (demag.f90:117-121).  Transformed on the
 * device in FP64 (demag.f90:259-267).  (2nx,2ny,...) in an open context. */
int mrmg_set_tensor(mrmg_ctx *ctx, const double *n_real);
/* Build the same tensor on the device in FP64 (tensor.f90:136-183 without SUL/head
 * images, plane_to_plane with num_images periodic images in x and y) and transform it.
 * When n_real_out is not NULL, the real-space tensor is also copied back to it.  In an
 * open context num_images must be 0 and the tensor is the padded grid's: at index i of
 * 2nx the displacement -i (i < nx) or 2nx - i (i > nx) of the media grid, zero at i = nx
 * (likewise in y). */
int mrmg_build_tensor(mrmg_ctx *ctx, int num_images, double *n_real_out);
/* As mrmg_build_tensor, with calculate_tensor's magnetic images (tensor.f90:103-200):
 * sul_mu > 1 adds each source layer mirrored in the soft underlayer, space_to_sul below the
 * bottom layer; head_mu > 1 adds its mirror in an infinite head, fly_height above the top
 * layer.  An image is scaled by (mu-1)/(mu+1) with its x and y rows negated (so the tensor
 * is no longer symmetric in its two indices).  num_images applies to both in-plane axes,
 * as in the reference (plane_to_plane reads n_imgxi for y too, tensor_element.f90:230).
 * mrmg_build_tensor is this with sul_mu = head_mu = 1. */
int mrmg_build_tensor_images(mrmg_ctx *ctx, int num_images, double sul_mu, double head_mu, double fly_height,
                             double space_to_sul, double *n_real_out);
/* No counterpart in the reference (its program builds one tensor per run): give dst the
 * tensor spectra of src, a context of the same precision, grid, boundaries, cell size and
 * layer stack on the same device, as a device-to-device copy instead of a rebuild (a
 * request loop starting many simulations of one grid). */
int mrmg_copy_tensor(mrmg_ctx *dst, const mrmg_ctx *src);

int mrmg_set_layer(mrmg_ctx *ctx, int k, const mrmg_layer *layer);
int mrmg_set_ufield(mrmg_ctx *ctx, int k, const mrmg_ufield *uf);
/* Set (or with kind MRMG_SRC_NONE clear) the field source (slot 0) or the temperature
 * source (slot 1) of layer k. */
int mrmg_set_source(mrmg_ctx *ctx, int k, int slot, const mrmg_source *src);

/* Magnetisation state of layer k: m (3,ng) unit vectors (the `in` slot of data_s%m). */
int mrmg_set_state(mrmg_ctx *ctx, int k, const double *m);
/* m (3,ng) and the last field h (3,ng) (either may be NULL). */
int mrmg_get_state(mrmg_ctx *ctx, int k, double *m, double *h);
/* |m| of layer k (ng): data_s%mag(in/3+1,:) (data.f90:198, set to 1 by grain.f90:61), the
 * LLB magnitude; push_llg leaves it alone.  set_mag writes both slots of the reference's
 * (2,ng) array, so the value holds whichever slot the next step reads. */
int mrmg_set_mag(mrmg_ctx *ctx, int k, const double *mag);
int mrmg_get_mag(mrmg_ctx *ctx, int k, double *mag);

/* Run p->n_steps steps of the llg loop (fixed dt when p->dt_frac == 0; otherwise the
 * adaptive dt of get_fields and the rejection loop of push_m, Heun only; each adaptive
 * step reads two small reductions back to the host).  noise: MRMG_NOISE_INJECT only, host
 * array of standard normals laid out (3, G, n_steps), G = total grains over all layers in
 * layer order, in the order get_thermal consumes them (field.f90:405-431).  diag may be
 * NULL.  The call returns after the device work is complete. */
int mrmg_step(mrmg_ctx *ctx, const mrmg_step_params *p, const double *noise, mrmg_diag *diag);

/* No counterpart in the reference (its program runs one simulation): fixed-dt steps of n
 * contexts together, for a request loop that serves many simulations on one GPU.  The
 * contexts must have equal grids (media and FFT), precision, layer count, launch shapes
 * and tensor storage, and each params[b] the same n_steps, dt, dt_max, llg_type, llb and
 * noise (MRMG_NOISE_NONE or MRMG_NOISE_PHILOX; seed, step0 and t0 are per member); no
 * adaptive dt, moving sources, switching record or profiling.  Each demag kernel runs once
 * for all members; with graph != 0 (llg_type < 0) one step is captured as a hipGraph and
 * replayed n_steps times.  Every member ends where its own mrmg_step with the same
 * parameters would end, bit for bit; t_end (n values, may be NULL) receives each member's
 * end time (its t0 plus n_steps additions of dt, the next call's t0).  The work runs on
 * ctxs[0]'s stream; no other call may use these contexts meanwhile. */
int mrmg_step_batch(mrmg_ctx *const *ctxs, const mrmg_step_params *params, int n, int graph, double *t_end);

/* The normals the Philox source draws for steps step0 .. step0+n_steps-1, same (3,G,n)
 * layout as injected noise, so a CPU path can be fed the identical thermal field. */
int mrmg_philox_normals(mrmg_ctx *ctx, uint64_t seed, uint64_t step0, int n_steps, double *out);

/* Demag field of the current state on the cell grid: H (nx,ny,n_lay,3) FP64 (component
 * slowest) and its grain average (3,G); either may be NULL.  Used by tests. */
int mrmg_demag_field(mrmg_ctx *ctx, double *h_cell, double *h_grain);

/* Device bytes held by the context. */
int mrmg_footprint(mrmg_ctx *ctx, size_t *bytes);

/* ---- outputs the reference does not compute: readback and the switching record.  Its
 * nearest relative is off_field.f90 (the stray field on a plane above the media,
 * plot_off_field 27-52), a plot-only path. ------------------------------------------- */
typedef struct {
    double sigma_x, sigma_y;  /* widths (cm) of the Gaussian sensitivity, > 0              */
    double dir[3];            /* direction of the head's field at the media (reciprocity)   */
    const double *w;          /* n_lay layer weights (e.g. a spacing loss per layer); NULL = 1 */
} mrmg_sensor;

/* Reciprocity readback of the current state at n_pos head positions:
 *   out[p] = sum over cells of w[k] dz[k] dx dy (dir . M) gx gy,
 *   gx = exp(-ddx^2 / (2 sigma_x^2)), gy = exp(-ddy^2 / (2 sigma_y^2)),
 * with M the cell magnetisation as the demag convolution sees it (Ms m, emu/cm^3; 0 in
 * non-magnetic cells) and ddx, ddy the periodic minimum-image distances from the cell
 * centre ((i+1/2) dx, (j+1/2) dy) to (xpos[p], ypos[p]).  out is in emu. */
int mrmg_readback(mrmg_ctx *ctx, const mrmg_sensor *s, int n_pos, const double *xpos, const double *ypos,
                  double *out);

/* mrmg_readback without device allocation.  It has no counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops and harnesses that forbid hipMalloc).  The result equals mrmg_readback's bit for
 * bit (the same kernels); the folded plane, the head positions and the outputs are staged
 * in one fixed device scratch array of MRMG_NOALLOC_PLANE_MAX + 3 * MRMG_NOALLOC_POS_MAX
 * doubles (33.5 MB on each GPU the library runs on, held until the process ends) instead
 * of buffers the library allocates.  Limits: nx * ny <= MRMG_NOALLOC_PLANE_MAX and
 * n_pos <= MRMG_NOALLOC_POS_MAX, else MRMG_ERR_ARG; the context must have been built by an
 * earlier call (mrmg_step, mrmg_readback, ...), else MRMG_ERR_STATE, because building it
 * allocates.  Every call on one GPU uses that GPU's one scratch array: calls that overlap
 * on one GPU (from several host threads) race, and the caller must serialise them. */
#define MRMG_NOALLOC_PLANE_MAX (4 * 1024 * 1024)
#define MRMG_NOALLOC_POS_MAX 65536
int mrmg_readback_noalloc(mrmg_ctx *ctx, const mrmg_sensor *s, int n_pos, const double *xpos, const double *ypos,
                          double *out);

/* Switching record.  enable != 0 clears it (every grain "not switched") and starts it:
 * after each later completed step of mrmg_step, every grain not yet marked whose m
 * satisfies dir . m > thr gets that step's counter (step0 + i), its end time and the
 * applied field there (the grain's static field plus the uniform field of
 * mrmg_set_ufield at that time; a moving source's field is not included).  enable 0
 * stops marking; the record stays readable until the next mrmg_set_layer, which ends
 * and discards it (the grains change).  E.g. dir = (0,0,-1), thr = 0 under a
 * reversing field gives each grain's first crossing of m_z = 0, and the switching-field
 * distribution from h_app. */
int mrmg_switch_arm(mrmg_ctx *ctx, int enable, const double dir[3], double thr);
/* The record of layer k: step (ng; -1 = not marked), t (ng) and h_app (3,ng); any may be
 * NULL.  t and h_app are meaningful only where step >= 0. */
int mrmg_switch_read(mrmg_ctx *ctx, int k, int64_t *step, double *t, double *h_app);

/* ---- instrumentation (no counterpart in the reference: its only timers are cpu_time
 * print-outs, demag.f90:280-283, main.f90:31-34) ------------------------------------- */
typedef struct {
    char name[32];          /* kernel class: row_fwd, col, row_inv, grain1, grain2        */
    int launches;           /* launches timed since profiling was enabled                 */
    double total_ms;        /* sum of HIP-event durations on the context's stream         */
} mrmg_kstat;

/* enable != 0 brackets every kernel launch of mrmg_step with HIP events on the
 * context's stream and accumulates per-class durations (resets the counters). */
int mrmg_profile(mrmg_ctx *ctx, int enable);
/* Launch shapes of the kernels, chosen by the caller or by mrm_amd's autotuner (the role
 * FFTW_EXHAUSTIVE planning plays for the reference, fftw.f90:25): "row_threads" (64, 128,
 * 256), "col_threads" (256, 512, or 1024 with the 6-slot tensor and without wave-local
 * FFTs), "col_groups" (observation-layer groups per column workgroup, 1..n_lay),
 * "col_path" (0 fused, 1 split, 2 fused on column-contiguous spectra, 3 split forward
 * column FFTs then the fused product and inverse, one observation layer per workgroup),
 * "row_block" (2, 4, 8 rows per transposing row workgroup), "row_radix" / "col_radix"
 * (largest Stockham radix: 4, 8, 16; radix 16 takes the compile-time passes for
 * power-of-two lengths from 256 to 4096), "wave_fft" (0, 1: wave-local FFTs),
 * "lanes_per_grain" (1, 2, 4, 8, 16).  The thread counts, "col_groups", "row_block" and
 * (with "wave_fft" 0) "col_path" leave results unchanged bit for bit; the others may
 * change the order of floating-point operations (FFT factorisation, gather sums) and so
 * the last bits.  A combination that does not fit the grid is refused. */
int mrmg_tune(mrmg_ctx *ctx, const char *key, int value);
/* The value of a tunable, or of the read-only key "tensor_offsets": 1 when the stored
 * tensor spectra are the 2 n_lay - 1 z-offset blocks (the tensor depends on the layer pair
 * only through ko - ks: equal thicknesses and spacings, no magnetic images above or
 * below), 0 when they are the n_lay^2 layer-pair blocks. */
int mrmg_get_tune(mrmg_ctx *ctx, const char *key, int *value);
/* Copies up to max entries; *n receives the number of kernel classes. */
int mrmg_kernel_stats(mrmg_ctx *ctx, mrmg_kstat *out, int max, int *n);
/* `reps` back-to-back launches of one demag kernel class ("row_fwd", "col", "row_inv" on
 * the fused paths; "row_fwd", "col_fwd", "prod", "col_inv", "row_inv" on the split path)
 * between two events on the context's stream; *ms receives the average duration, free of
 * the per-launch event pair that mrmg_profile adds.  Writes only the spectra and the
 * field's run sums, which every field evaluation rewrites before it reads them. */
int mrmg_time_kernel(mrmg_ctx *ctx, const char *name, int reps, double *ms);
/* Device-to-device copy of `bytes` bytes on the context's stream with the library's own
 * streaming copy kernel, `reps` times; *ms receives the average duration.  Used as the
 * known-size reference when scaling HBM traffic counters. */
int mrmg_copy_probe(mrmg_ctx *ctx, size_t bytes, int reps, double *ms);
/* The same copy by the runtime (hipMemcpyAsync device-to-device on the context's stream),
 * to cross-check the copy kernel's rate as the reachable bandwidth. */
int mrmg_memcpy_probe(mrmg_ctx *ctx, size_t bytes, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* MRMG_H */
