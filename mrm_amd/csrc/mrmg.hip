// mrmg.hip - the C ABI of include/mrmg.h over the MI355X (gfx950) granular-media time step.
// The Engine (engine.h) sequences the llg loop (llg.f90:111-318) and push_m's stages
// (llg.f90:1195-1271) over the kernels of three translation units:
//   demag.hip   get_mag's scatter into the row r2c FFT (data.f90:322-390) and get_demag's
//               column FFTs, k-space product and row c2r (demag.f90:33-135)
//   grain.hip   get_field (field.f90:160-273), get_thermal (397-436) and the push_llg /
//               push_llb pushes (llg.f90:916-1155) in one grain kernel per stage
//   tensor.hip  calculate_tensor / plane_to_plane (tensor.f90:82-216,
//               tensor_element.f90:13-261) in FP64 and init_demag's tensor FFT
//               (demag.f90:259-267)
#include "engine.h"

// ------------------------------------------------------------------------- errors
static thread_local std::string g_err;

int set_err(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// Scratch of the allocation-free readback: the plane, then the head positions (x, y), then
// the outputs.  Host code never passes this name to a kernel (it is a host-side stand-in);
// it uses the device address hipGetSymbolAddress returns for the current device.
#define RB_SCRATCH_DOUBLES ((size_t)MRMG_NOALLOC_PLANE_MAX + 3 * (size_t)MRMG_NOALLOC_POS_MAX)
static __device__ __align__(256) double scratch[RB_SCRATCH_DOUBLES];

// ------------------------------------------------------------------------- C ABI
// (the definitions take C linkage from their declarations in mrmg.h)

const char* mrmg_last_error(void) { return g_err.c_str(); }

// Device address of the readback scratch array on each device, taken once per device by
// mrmg_create with that device current.  Each device holds its own copy of the array.
#define SCRATCH_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

static int scratch_setup(int dev)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return set_err(MRMG_ERR_UNSUPPORTED, "device %d: too many devices", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return MRMG_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
    return MRMG_OK;
}

static double* scratch_of(int dev)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

int mrmg_device_count(int* n)
{
    if (!n) return set_err(MRMG_ERR_ARG, "null argument");
    HIPCHK(hipGetDeviceCount(n));
    return MRMG_OK;
}

static int create_ctx(const mrmg_geom* g, bool open, mrmg_ctx** out)
{
    if (!g || !out) return set_err(MRMG_ERR_ARG, "null argument");
    *out = nullptr;
    if (g->nx < 2 || g->ny < 1 || g->n_lay < 1 || g->n_lay > MRMG_MAX_LAYERS)
        return set_err(MRMG_ERR_ARG, "bad grid %d x %d x %d (n_lay <= %d)", g->nx, g->ny, g->n_lay, MRMG_MAX_LAYERS);
    if (g->precision != MRMG_FP64 && g->precision != MRMG_FP32) return set_err(MRMG_ERR_ARG, "bad precision");
    if (!g->dz) return set_err(MRMG_ERR_ARG, "dz is NULL");
    if (g->nx > 32767) return set_err(MRMG_ERR_UNSUPPORTED, "nx = %d: runs store x and their length in 16 bits", g->nx);
    std::unique_ptr<mrmg_ctx> c(new mrmg_ctx());
    c->nx = g->nx; c->ny = g->ny; c->nl = g->n_lay;
    c->open = open;
    c->fx = open ? 2 * g->nx : g->nx;
    c->fy = open ? 2 * g->ny : g->ny;
    c->nxh = c->fx / 2 + 1;
    c->precision = g->precision; c->demag = g->demag;
    c->dx = g->dx; c->dy = g->dy;
    c->dz.assign(g->dz, g->dz + g->n_lay);
    c->il.assign(g->n_lay, 0.0);
    if (g->il_space) c->il.assign(g->il_space, g->il_space + g->n_lay);
    if (!factorize(c->fx, c->px) || !factorize(c->fy, c->py))
        return set_err(MRMG_ERR_UNSUPPORTED, "FFT grid %d x %d must factor into 2, 3, 5", c->fx, c->fy);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return set_err(MRMG_ERR_HIP, "no HIP device");
    HIPCHK(hipGetDevice(&c->device));
    RCHK(scratch_setup(c->device));
    HIPCHK(hipStreamCreate(&c->stream));
    c->layers.resize(c->nl);
    if (c->precision == MRMG_FP64) c->eng.reset(new Engine<double>(c.get()));
    else c->eng.reset(new Engine<float>(c.get()));
    int rc = c->precision == MRMG_FP64 ? static_cast<Engine<double>*>(c->eng.get())->init_fft()
                                       : static_cast<Engine<float>*>(c->eng.get())->init_fft();
    if (rc != MRMG_OK) {
        c->eng.reset();
        (void)hipStreamDestroy(c->stream);
        return rc;
    }
    *out = c.release();
    return MRMG_OK;
}

int mrmg_create(const mrmg_geom* g, mrmg_ctx** out) { return create_ctx(g, false, out); }
int mrmg_create_open(const mrmg_geom* g, mrmg_ctx** out) { return create_ctx(g, true, out); }

int mrmg_destroy(mrmg_ctx* c)
{
    if (!c) return MRMG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->eng.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MRMG_OK;
}

static int ensure_built(mrmg_ctx* c)
{
    if (!c->dirty) return MRMG_OK;
    for (int k = 0; k < c->nl; k++)
        if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set (mrmg_set_layer)", k);
    RCHK(c->eng->build(c->layers));
    c->dirty = false;
    return MRMG_OK;
}

template <typename T>
static int set_tensor_T(mrmg_ctx* c, const double* d_blocks, bool sym, bool toep)
{
    Engine<T>* e = static_cast<Engine<T>*>(c->eng.get());
    RCHK(e->set_slots(sym));
    return e->set_tensor_planes(d_blocks, toep);
}

static int set_tensor_blocks(mrmg_ctx* c, const double* d_blocks, bool sym, bool toep)
{
    return c->precision == MRMG_FP64 ? set_tensor_T<double>(c, d_blocks, sym, toep)
                                     : set_tensor_T<float>(c, d_blocks, sym, toep);
}

// Is component (a,b) of every block of the real tensor equal to (b,a), bit for bit?
static int tensor_symmetric(mrmg_ctx* c, const double* d_blocks, int nblk, bool* sym)
{
    DBuf<unsigned long long> cnt;
    RCHK(cnt.alloc(1));
    RCHK(tensor_sym_count(c->stream, d_blocks, (size_t)c->fx * c->fy, nblk, cnt.p));
    unsigned long long h = 1;
    HIPCHK(hipMemcpyAsync(&h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *sym = h == 0;
    return MRMG_OK;
}

// z-offset representative of offset block q (d = q - (nl-1) = ko - ks): the pair
// (ko, ks) = (max(d,0), max(d,0) - d)
static void offset_pair(int q, int nl, int* ko, int* ks)
{
    const int d = q - (nl - 1);
    *ko = d > 0 ? d : 0;
    *ks = *ko - d;
}

// A full tensor of nl^2 layer-pair blocks on the device into the stored spectra.  When
// every pair block lies within 1e-14 (L1, relative) of the block of its z-offset
// representative, only the 2 nl - 1 offset blocks are kept: the tensor then depends on
// the layers through ko - ks alone (equal thicknesses and spacings, no images above or
// below), up to the rounding of the z positions it was built from.
static int tensor_from_full(mrmg_ctx* c, DBuf<double>& full)
{
    const int nl = c->nl;
    const size_t blk = (size_t)9 * c->fx * c->fy;
    bool sym = false, toep = false;
    RCHK(tensor_symmetric(c, full.p, nl * nl, &sym));
    if (nl > 1) {
        DBuf<double> l1;
        RCHK(l1.alloc(2 * (size_t)nl * nl));
        RCHK(tensor_toeplitz_l1(c->stream, full.p, blk, nl, l1.p));
        std::vector<double> h(2 * (size_t)nl * nl);
        HIPCHK(hipMemcpyAsync(h.data(), l1.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        toep = true;
        for (int p = 0; p < nl * nl; p++)
            if (!(h[2 * p] <= 1e-14 * h[2 * p + 1])) toep = false;
    }
    if (!toep) return set_tensor_blocks(c, full.p, sym, false);
    DBuf<double> comp;
    RCHK(comp.alloc((size_t)(2 * nl - 1) * blk));
    for (int q = 0; q < 2 * nl - 1; q++) {
        int ko, ks;
        offset_pair(q, nl, &ko, &ks);
        HIPCHK(hipMemcpyAsync(comp.p + q * blk, full.p + (size_t)(ko + nl * ks) * blk, blk * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    full.release();
    return set_tensor_blocks(c, comp.p, sym, true);
}

int mrmg_set_tensor(mrmg_ctx* c, const double* n_real)
{
    if (!c || !n_real) return set_err(MRMG_ERR_ARG, "null argument");
    const size_t count = (size_t)c->fx * c->fy * 9 * c->nl * c->nl;
    DBuf<double> d;
    RCHK(d.upload(n_real, count, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return tensor_from_full(c, d);
}

int mrmg_build_tensor(mrmg_ctx* c, int num_images, double* n_real_out)
{
    return mrmg_build_tensor_images(c, num_images, 1.0, 1.0, 8.e-7, 40.e-7, n_real_out);
}

int mrmg_build_tensor_images(mrmg_ctx* c, int num_images, double sul_mu, double head_mu, double fly_height,
                             double space_to_sul, double* n_real_out)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (num_images < 0) return set_err(MRMG_ERR_ARG, "num_images < 0");
    if (c->open && num_images != 0)
        return set_err(MRMG_ERR_ARG, "open boundaries take no periodic images (num_images %d)", num_images);
    // the blocks are built on the FFT grid: for open boundaries every displacement of the
    // media grid once, none wrapped (the grid is padded to twice the media)
    const int nx = c->fx, ny = c->fy, nl = c->nl;
    const int open_nx = c->open ? c->nx : 0, open_ny = c->open ? c->ny : 0;
    const size_t plane = (size_t)nx * ny, blk = plane * 9, count = blk * nl * nl;
    // image strengths (mu-1)/(mu+1) of the soft underlayer and the head (tensor.f90:125-126)
    const bool sul = sul_mu > 1.0, head = head_mu > 1.0;
    const double mus = sul ? (sul_mu - 1.0) / (sul_mu + 1.0) : 0.0;
    const double mush = head ? (head_mu - 1.0) / (head_mu + 1.0) : 0.0;
    // distance from the head plane to the centre of each layer (tensor.f90:141-147)
    std::vector<double> hdz(nl);
    hdz[0] = fly_height;
    for (int ks = 1; ks < nl; ks++) hdz[ks] = hdz[ks - 1] + c->dz[ks - 1] + c->il[ks];
    for (int ks = 0; ks < nl; ks++) hdz[ks] = hdz[ks] + 0.5 * c->dz[ks];
    // bottom of each layer above the soft underlayer, in calculate_tensor's layer loop order
    // (tensor.f90:148-200): zb[nl-1] = space_to_sul, zb[k-1] = zb[k] + dz[k] + il[k]
    std::vector<double> zb(nl);
    zb[nl - 1] = space_to_sul;
    for (int k = nl - 1; k > 0; k--) zb[k - 1] = zb[k] + c->dz[k] + c->il[k];
    // the block of source layer ks acting on observation layer ko (space_to_sul cancels in
    // the direct rz)
    auto block = [&](int ko, int ks, double* dst) -> int {
        const double zsb = zb[ks], zst = zsb + c->dz[ks], zob = zb[ko], zot = zob + c->dz[ko];
        const double rz = 0.5 * (zst - zot + zsb - zob);
        const double norm = c->dx * c->dy * c->dz[ko] * nx * ny;
        RCHK(launch_plane_to_plane(c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko], rz, num_images, norm, dst,
                                   (sul || head) ? 1 : 0, 0.0, 0.0, open_nx, open_ny));
        if (sul)
            RCHK(launch_plane_to_plane(c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko],
                                       0.5 * (-zst - zot - zsb - zob), num_images, norm, dst, 2, -mus, mus, open_nx,
                                       open_ny));
        if (head)
            RCHK(launch_plane_to_plane(c->stream, nx, ny, c->dx, c->dy, c->dz[ks], c->dz[ko], hdz[ks] + hdz[ko],
                                       num_images, norm, dst, 2, -mush, mush, open_nx, open_ny));
        if (sul || head) RCHK(launch_tensor_div(c->stream, dst, blk, norm));
        return MRMG_OK;
    };
    // equal thicknesses and spacings without images above or below: the block depends on
    // ko - ks alone, and only the 2 nl - 1 offsets are built (each with its representative
    // pair's z positions, which differ from the other pairs' by rounding only)
    bool uniform = nl > 1 && !sul && !head;
    for (int k = 1; k < nl && uniform; k++)
        if (c->dz[k] != c->dz[0] || (k > 1 && c->il[k] != c->il[1])) uniform = false;
    if (uniform) {
        const int nq = 2 * nl - 1;
        DBuf<double> comp;
        RCHK(comp.alloc((size_t)nq * blk));
        for (int q = 0; q < nq; q++) {
            int ko, ks;
            offset_pair(q, nl, &ko, &ks);
            RCHK(block(ko, ks, comp.p + q * blk));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        if (n_real_out) {
            std::vector<double> h((size_t)nq * blk);
            HIPCHK(hipMemcpyAsync(h.data(), comp.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int ko = 0; ko < nl; ko++)
                for (int ks = 0; ks < nl; ks++)
                    memcpy(n_real_out + (size_t)(ko + nl * ks) * blk, h.data() + (size_t)(ko - ks + nl - 1) * blk,
                           blk * sizeof(double));
        }
        bool sym = false;
        RCHK(tensor_symmetric(c, comp.p, nq, &sym));
        return set_tensor_blocks(c, comp.p, sym, true);
    }
    DBuf<double> d;
    RCHK(d.alloc(count));
    for (int ks = nl - 1; ks >= 0; ks--)
        for (int ko = nl - 1; ko >= 0; ko--) RCHK(block(ko, ks, d.p + (size_t)(ko + nl * ks) * blk));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n_real_out) {
        HIPCHK(hipMemcpyAsync(n_real_out, d.p, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return tensor_from_full(c, d);
}

template <typename V> static void copy_vec(std::vector<V>& dst, const V* src, size_t n)
{
    if (src) dst.assign(src, src + n);
    else dst.clear();
}

// before a rebuild: the device state of every layer back into the host copies
static int save_device_state(mrmg_ctx* c)
{
    if (c->dirty) return MRMG_OK;
    for (int q = 0; q < c->nl; q++) {
        HostLayer& H = c->layers[q];
        if (!H.set) continue;
        H.m.resize(3 * (size_t)H.ng);
        H.mag.resize(H.ng);
        RCHK(c->eng->get_state(q, H.m.data(), nullptr));
        RCHK(c->eng->get_mag(q, H.mag.data()));
    }
    return MRMG_OK;
}

int mrmg_set_layer(mrmg_ctx* c, int k, const mrmg_layer* in)
{
    if (!c || !in) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (in->ng < 1) return set_err(MRMG_ERR_ARG, "ng must be >= 1");
    if (!in->grid2grain || !in->grain2gridindex || !in->grain2grid || !in->grain_vol || !in->ms || !in->hk)
        return set_err(MRMG_ERR_ARG, "layer %d: required array is NULL", k);
    RCHK(save_device_state(c));
    HostLayer H;
    H.set = true;
    const int ng = in->ng;
    H.ng = ng;
    const size_t P = (size_t)c->nx * c->ny;
    copy_vec(H.grid2grain, in->grid2grain, P);
    copy_vec(H.g2gi, in->grain2gridindex, ng + 1);
    if (H.g2gi[0] != 0) return set_err(MRMG_ERR_ARG, "grain2gridindex(0) must be 0");
    for (int n = 0; n < ng; n++)
        if (H.g2gi[n + 1] < H.g2gi[n]) return set_err(MRMG_ERR_ARG, "grain2gridindex not monotone");
    H.ncell = H.g2gi[ng];
    if ((size_t)H.ncell > P) return set_err(MRMG_ERR_ARG, "more grain cells than grid cells");
    copy_vec(H.g2g, in->grain2grid, 2 * (size_t)H.ncell);
    copy_vec(H.vol, in->grain_vol, ng);
    copy_vec(H.ms, in->ms, ng);
    copy_vec(H.hk, in->hk, 2 * (size_t)ng);
    H.has_uni = in->uni_axis != nullptr;
    if (H.has_uni) copy_vec(H.uni, in->uni_axis, 3 * (size_t)ng);
    H.has_cub = in->cubic_axis != nullptr;
    if (H.has_cub) copy_vec(H.cub, in->cubic_axis, 9 * (size_t)ng);
    auto csr = [&](const int* idx, const int* gr, const double* w, std::vector<int>& vi, std::vector<int>& vg,
                   std::vector<double>& vw, bool& has) -> int {
        has = idx != nullptr;
        if (!has) return MRMG_OK;
        if (!gr || !w) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR incomplete", k);
        copy_vec(vi, idx, ng + 1);
        if (vi[0] != 0) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR offset(0) must be 0", k);
        for (int n = 0; n < ng; n++)
            if (vi[n + 1] < vi[n]) return set_err(MRMG_ERR_ARG, "layer %d: exchange CSR not monotone", k);
        copy_vec(vg, gr, vi[ng]);
        copy_vec(vw, w, vi[ng]);
        return MRMG_OK;
    };
    RCHK(csr(in->hxgrainindex, in->hxgrain, in->hx, H.hxi, H.hxg, H.hx, H.has_hx));
    RCHK(csr(in->uhxgrainindex, in->uhxgrain, in->uhx, H.uhxi, H.uhxg, H.uhx, H.has_uhx));
    RCHK(csr(in->lhxgrainindex, in->lhxgrain, in->lhx, H.lhxi, H.lhxg, H.lhx, H.has_lhx));
    H.ms_scale.assign(ng, 1.0); if (in->ms_scale) copy_vec(H.ms_scale, in->ms_scale, ng);
    H.alpha_scale.assign(ng, 1.0); if (in->alpha_scale) copy_vec(H.alpha_scale, in->alpha_scale, ng);
    H.hk_scale.assign(ng, 1.0); if (in->hk_scale) copy_vec(H.hk_scale, in->hk_scale, ng);
    H.temp.assign(ng, 0.0); if (in->temp) copy_vec(H.temp, in->temp, ng);
    H.has_happ = in->h_app != nullptr;
    if (H.has_happ) copy_vec(H.happ, in->h_app, 3 * (size_t)ng);
    H.alpha = in->alpha; H.gamma = in->gamma; H.thm_alpha = in->thm_alpha; H.floor_ms = in->floor_ms;
    H.thermal = in->thermal;
    // LLB parameters: Tc defaults to 1e22 (data.f90:201), lambdaMs to
    // 3 Tc / (g muB/kB (S+1)) (grain.f90:61)
    H.spin = in->spin;
    H.tc.assign(ng, 1.e22); if (in->tc) copy_vec(H.tc, in->tc, ng);
    H.lambda_ms.resize(ng);
    for (int n = 0; n < ng; n++)
        H.lambda_ms[n] = in->lambda_ms ? in->lambda_ms[n] : 3.0 * H.tc[n] / (1.343427768e-4 * (H.spin + 1.0));
    // find_temp_scaling (temp_scaling.f90:365-431): nothing when tc_nom > 1e10; the
    // power laws of the (static) grain temperature are applied here, the |m| form at
    // every field evaluation on the device
    H.use_mag = false;
    H.tscale = in->tc_nom <= 1.e10;
    H.msexp = in->msexp;
    H.hkexp = in->hkexp;
    {
        const int nt[3] = {in->n_ms_tab, in->n_alpha_tab, in->n_hk_tab};
        const double* tt[3] = {in->ms_tab_t, in->alpha_tab_t, in->hk_tab_t};
        const double* ts[3] = {in->ms_tab_s, in->alpha_tab_s, in->hk_tab_s};
        for (int i = 0; i < 3; i++) {
            if (nt[i] == 0) continue;
            if (nt[i] < 2 || !tt[i] || !ts[i]) return set_err(MRMG_ERR_ARG, "scaling table %d: need >= 2 nodes", i);
            for (int j = 1; j < nt[i]; j++)
                if (!(tt[i][j] > tt[i][j - 1])) return set_err(MRMG_ERR_ARG, "scaling table %d: T/Tc not increasing", i);
            H.tab_t[i].assign(tt[i], tt[i] + nt[i]);
            H.tab_s[i].assign(ts[i], ts[i] + nt[i]);
        }
    }
    if (in->tc_nom <= 1.e10) {
        H.use_mag = in->use_mag != 0;
        auto cl = [](double x) { return x < 1.0 ? (x > 0.0 ? x : 0.0) : 1.0; };
        auto tab = [&](int i, int n) {
            return table_eval(H.tab_t[i].data(), H.tab_s[i].data(), (int)H.tab_t[i].size(), H.temp[n] / H.tc[n]);
        };
        for (int n = 0; n < ng; n++) {
            if (!H.use_mag) {
                if (in->msexp >= 0.0) H.ms_scale[n] = pow(cl(1.0 - H.temp[n] / H.tc[n]), in->msexp);
                else if (!H.tab_t[0].empty()) H.ms_scale[n] = tab(0, n);
            }
            if (!H.tab_t[1].empty()) H.alpha_scale[n] = tab(1, n);
            if (in->hkexp >= 0.0) H.hk_scale[n] = pow(cl(1.0 - H.temp[n] / H.tc[n]), in->hkexp);
            else if (!H.tab_t[2].empty()) H.hk_scale[n] = tab(2, n);
        }
    }
    // init_data_one: m = +z (data.f90:201-202), |m| = 1 (grain.f90:61)
    H.m.assign(3 * (size_t)ng, 0.0);
    for (int n = 0; n < ng; n++) H.m[3 * n + 2] = 1.0;
    H.mag.assign(ng, 1.0);
    H.uf = c->layers[k].uf;
    H.has_uf = c->layers[k].has_uf;
    for (int sl = 0; sl < 2; sl++) {
        H.src[sl] = c->layers[k].src[sl];
        H.src_x[sl] = c->layers[k].src_x[sl];
        H.src_y[sl] = c->layers[k].src_y[sl];
        H.src_v[sl] = c->layers[k].src_v[sl];
    }
    c->layers[k] = std::move(H);
    c->dirty = true;
    return MRMG_OK;
}

int mrmg_set_ufield(mrmg_ctx* c, int k, const mrmg_ufield* uf)
{
    if (!c || !uf) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    RCHK(check_tfun(uf->func, uf->t_start, uf->t_end, uf->amp, uf->gamma, uf->param));
    c->layers[k].uf = *uf;
    c->layers[k].has_uf = true;
    if (!c->dirty) return c->eng->update_ufield(c->layers);
    return MRMG_OK;
}

int mrmg_set_source(mrmg_ctx* c, int k, int slot, const mrmg_source* src)
{
    if (!c || !src) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (slot != 0 && slot != 1) return set_err(MRMG_ERR_ARG, "source slot must be 0 (field) or 1 (temperature)");
    const int kind = src->kind;
    if (kind != MRMG_SRC_NONE) {
        if (slot == 0 && kind != MRMG_SRC_FIELD_TABLE)
            return set_err(MRMG_ERR_ARG, "slot 0 takes a field table (MRMG_SRC_FIELD_TABLE)");
        if (slot == 1 && kind != MRMG_SRC_TEMP_TABLE && kind != MRMG_SRC_TEMP_SG)
            return set_err(MRMG_ERR_ARG, "slot 1 takes a temperature table or super-Gaussian");
        RCHK(check_tfun(src->func, src->t_start, src->t_end, src->amp, src->gamma, src->param));
        if (kind == MRMG_SRC_TEMP_SG) {
            if (src->expn < 1) return set_err(MRMG_ERR_ARG, "super-Gaussian power must be >= 1");
        } else {
            if (src->nxp < 2 || src->nyp < 2 || !src->xpos || !src->ypos || !src->vals)
                return set_err(MRMG_ERR_ARG, "a table needs at least 2 x 2 nodes and its arrays");
            for (int i = 1; i < src->nxp; i++)
                if (!(src->xpos[i] > src->xpos[i - 1])) return set_err(MRMG_ERR_ARG, "table x nodes must increase");
            for (int i = 1; i < src->nyp; i++)
                if (!(src->ypos[i] > src->ypos[i - 1])) return set_err(MRMG_ERR_ARG, "table y nodes must increase");
        }
    }
    // a built context takes the new source without a rebuild: only the source tables and
    // the layer parameters change
    const bool built = !c->dirty;
    HostLayer& H = c->layers[k];
    H.src[slot] = *src;
    H.src_x[slot].clear(); H.src_y[slot].clear(); H.src_v[slot].clear();
    if (kind == MRMG_SRC_FIELD_TABLE || kind == MRMG_SRC_TEMP_TABLE) {
        const int nv = kind == MRMG_SRC_FIELD_TABLE ? 3 : 1;
        H.src_x[slot].assign(src->xpos, src->xpos + src->nxp);
        H.src_y[slot].assign(src->ypos, src->ypos + src->nyp);
        H.src_v[slot].assign(src->vals, src->vals + (size_t)nv * src->nxp * src->nyp);
    }
    H.src[slot].xpos = H.src[slot].ypos = H.src[slot].vals = nullptr;
    if (built) return c->eng->update_sources(c->layers);
    return MRMG_OK;
}

int mrmg_set_state(mrmg_ctx* c, int k, const double* m)
{
    if (!c || !m) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        c->layers[k].m.assign(m, m + 3 * (size_t)c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->set_state(k, m);
}

int mrmg_set_mag(mrmg_ctx* c, int k, const double* mag)
{
    if (!c || !mag) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        c->layers[k].mag.assign(mag, mag + c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->set_mag(k, mag);
}

int mrmg_get_mag(mrmg_ctx* c, int k, double* mag)
{
    if (!c || !mag) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        memcpy(mag, c->layers[k].mag.data(), sizeof(double) * c->layers[k].ng);
        return MRMG_OK;
    }
    return c->eng->get_mag(k, mag);
}

int mrmg_get_state(mrmg_ctx* c, int k, double* m, double* h)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    if (!c->layers[k].set) return set_err(MRMG_ERR_STATE, "layer %d not set", k);
    if (c->dirty) {
        const HostLayer& H = c->layers[k];
        if (m) memcpy(m, H.m.data(), sizeof(double) * 3 * H.ng);
        if (h) memset(h, 0, sizeof(double) * 3 * H.ng);
        return MRMG_OK;
    }
    return c->eng->get_state(k, m, h);
}

int mrmg_step(mrmg_ctx* c, const mrmg_step_params* p, const double* noise, mrmg_diag* diag)
{
    if (!c || !p) return set_err(MRMG_ERR_ARG, "null argument");
    if (p->n_steps < 0) return set_err(MRMG_ERR_ARG, "n_steps < 0");
    if (c->demag && !c->eng->has_tensor())
        return set_err(MRMG_ERR_STATE, "demag enabled but no tensor (mrmg_set_tensor / mrmg_build_tensor)");
    RCHK(ensure_built(c));
    return c->eng->step(p, noise, diag);
}

int mrmg_copy_tensor(mrmg_ctx* dst, const mrmg_ctx* src)
{
    if (!dst || !src) return set_err(MRMG_ERR_ARG, "null argument");
    if (dst == src) return MRMG_OK;
    if (dst->precision != src->precision || dst->nx != src->nx || dst->ny != src->ny || dst->nl != src->nl ||
        dst->open != src->open || dst->dx != src->dx || dst->dy != src->dy || dst->dz != src->dz ||
        dst->il != src->il || dst->device != src->device)
        return set_err(MRMG_ERR_ARG, "the two contexts differ in form, grid, layer stack or device");
    return dst->eng->copy_tensor(src->eng.get());
}

int mrmg_step_batch(mrmg_ctx* const* ctxs, const mrmg_step_params* params, int n, int graph, double* t_end)
{
    if (!ctxs || !params || n < 1) return set_err(MRMG_ERR_ARG, "null argument or empty batch");
    for (int b = 0; b < n; b++) {
        mrmg_ctx* c = ctxs[b];
        if (!c) return set_err(MRMG_ERR_ARG, "batch member %d is NULL", b);
        if (c->precision != ctxs[0]->precision) return set_err(MRMG_ERR_ARG, "batch member %d: precision differs", b);
        if (c->demag && !c->eng->has_tensor()) return set_err(MRMG_ERR_STATE, "batch member %d has no tensor", b);
        RCHK(ensure_built(c));
    }
    if (ctxs[0]->precision == MRMG_FP64) {
        std::vector<Engine<double>*> M(n);
        for (int b = 0; b < n; b++) M[b] = static_cast<Engine<double>*>(ctxs[b]->eng.get());
        return batch_step(M, params, graph, t_end);
    }
    std::vector<Engine<float>*> M(n);
    for (int b = 0; b < n; b++) M[b] = static_cast<Engine<float>*>(ctxs[b]->eng.get());
    return batch_step(M, params, graph, t_end);
}

int mrmg_philox_normals(mrmg_ctx* c, uint64_t seed, uint64_t step0, int n_steps, double* out)
{
    if (!c || !out) return set_err(MRMG_ERR_ARG, "null argument");
    RCHK(ensure_built(c));
    int G = 0;
    for (auto& L : c->layers) G += L.ng;
    const size_t n = (size_t)n_steps * G;
    DBuf<double> d;
    RCHK(d.alloc(3 * n));
    RCHK(launch_philox_dump(c->stream, seed, step0, n_steps, G, d.p));
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpyAsync(out, d.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MRMG_OK;
}

int mrmg_demag_field(mrmg_ctx* c, double* h_cell, double* h_grain)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (c->demag && !c->eng->has_tensor()) return set_err(MRMG_ERR_STATE, "no tensor");
    RCHK(ensure_built(c));
    return c->eng->demag_field(h_cell, h_grain);
}

int mrmg_readback(mrmg_ctx* c, const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos,
                  double* out)
{
    if (!c || !s) return set_err(MRMG_ERR_ARG, "null argument");
    if (n_pos < 0) return set_err(MRMG_ERR_ARG, "n_pos < 0");
    if (n_pos > 0 && (!xpos || !ypos || !out)) return set_err(MRMG_ERR_ARG, "null position or output array");
    if (!(s->sigma_x > 0.0) || !(s->sigma_y > 0.0)) return set_err(MRMG_ERR_ARG, "sensor widths must be > 0");
    if ((size_t)(c->nx + c->ny) * sizeof(double) > 64 * 1024 - 64)   // with the kernel's static reduction array
        return set_err(MRMG_ERR_UNSUPPORTED, "readback: nx + ny too large for one workgroup's table");
    RCHK(ensure_built(c));
    if (n_pos == 0) return MRMG_OK;
    return c->eng->readback(s, n_pos, xpos, ypos, out);
}

int mrmg_readback_noalloc(mrmg_ctx* c, const mrmg_sensor* s, int n_pos, const double* xpos, const double* ypos,
                          double* out)
{
    if (!c || !s) return set_err(MRMG_ERR_ARG, "null argument");
    if (n_pos < 0) return set_err(MRMG_ERR_ARG, "n_pos < 0");
    if (n_pos > 0 && (!xpos || !ypos || !out)) return set_err(MRMG_ERR_ARG, "null position or output array");
    if (!(s->sigma_x > 0.0) || !(s->sigma_y > 0.0)) return set_err(MRMG_ERR_ARG, "sensor widths must be > 0");
    if ((size_t)(c->nx + c->ny) * sizeof(double) > 64 * 1024 - 64)
        return set_err(MRMG_ERR_UNSUPPORTED, "readback: nx + ny too large for one workgroup's table");
    if ((size_t)c->nx * c->ny > (size_t)MRMG_NOALLOC_PLANE_MAX)
        return set_err(MRMG_ERR_ARG, "readback_noalloc: nx*ny = %zu exceeds the scratch plane of %d",
                       (size_t)c->nx * c->ny, MRMG_NOALLOC_PLANE_MAX);
    if (n_pos > MRMG_NOALLOC_POS_MAX)
        return set_err(MRMG_ERR_ARG, "readback_noalloc: n_pos = %d exceeds %d", n_pos, MRMG_NOALLOC_POS_MAX);
    // building the engine allocates: this entry point runs only on a built context
    if (c->dirty) return set_err(MRMG_ERR_STATE, "readback_noalloc: context not built yet (run mrmg_step first)");
    if (n_pos == 0) return MRMG_OK;
    double* sp = scratch_of(c->device);
    if (!sp) return set_err(MRMG_ERR_STATE, "readback_noalloc: no scratch address for device %d", c->device);
    double* plane = sp;
    double* pos = sp + MRMG_NOALLOC_PLANE_MAX;
    double* dout = pos + 2 * (size_t)MRMG_NOALLOC_POS_MAX;
    return c->eng->readback_into(s, n_pos, xpos, ypos, out, plane, pos, dout);
}

int mrmg_switch_arm(mrmg_ctx* c, int enable, const double* dir, double thr)
{
    if (!c || (enable && !dir)) return set_err(MRMG_ERR_ARG, "null argument");
    RCHK(ensure_built(c));
    return c->eng->switch_arm(enable, dir, thr);
}

int mrmg_switch_read(mrmg_ctx* c, int k, int64_t* step, double* t, double* h_app)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    if (k < 0 || k >= c->nl) return set_err(MRMG_ERR_ARG, "layer index %d out of range", k);
    RCHK(ensure_built(c));
    return c->eng->switch_read(k, step, t, h_app);
}

int mrmg_profile(mrmg_ctx* c, int enable)
{
    if (!c) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->profile(enable);
}

int mrmg_tune(mrmg_ctx* c, const char* key, int value)
{
    if (!c || !key) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->tune(key, value);
}

int mrmg_get_tune(mrmg_ctx* c, const char* key, int* value)
{
    if (!c || !key || !value) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->get_tune(key, value);
}

int mrmg_kernel_stats(mrmg_ctx* c, mrmg_kstat* out, int max, int* n)
{
    if (!c || !out || !n) return set_err(MRMG_ERR_ARG, "null argument");
    return c->eng->kernel_stats(out, max, n);
}

int mrmg_time_kernel(mrmg_ctx* c, const char* name, int reps, double* ms)
{
    if (!c || !name || !ms) return set_err(MRMG_ERR_ARG, "null argument");
    if (reps < 1) return set_err(MRMG_ERR_ARG, "reps must be >= 1");
    static const char* names[] = {"row_fwd", "col", "row_inv", "", "", "col_fwd", "prod", "col_inv"};
    int cls = -1;
    for (int i = 0; i < 8; i++)
        if (names[i][0] && strcmp(names[i], name) == 0) cls = i;
    if (cls < 0) return set_err(MRMG_ERR_ARG, "unknown demag kernel class '%s'", name);
    RCHK(ensure_built(c));
    return c->eng->time_class(cls, reps, ms);
}

int mrmg_copy_probe(mrmg_ctx* c, size_t bytes, int reps, double* ms)
{
    if (!c || !ms || reps < 1) return set_err(MRMG_ERR_ARG, "bad argument");
    const size_t n = bytes / 16;
    if (n == 0) return set_err(MRMG_ERR_ARG, "bytes < 16");
    DBuf<uint4> a, b;
    RCHK(a.alloc(n));
    RCHK(b.alloc(n));
    HIPCHK(hipMemsetAsync(a.p, 1, n * 16, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    RCHK(launch_copy16(c->stream, grid, a.p, b.p, n));   // warm
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) RCHK(launch_copy16(c->stream, grid, a.p, b.p, n));
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MRMG_OK;
}

int mrmg_memcpy_probe(mrmg_ctx* c, size_t bytes, int reps, double* ms)
{
    if (!c || !ms || reps < 1 || bytes == 0) return set_err(MRMG_ERR_ARG, "bad argument");
    DBuf<unsigned char> a, b;
    RCHK(a.alloc(bytes));
    RCHK(b.alloc(bytes));
    HIPCHK(hipMemsetAsync(a.p, 1, bytes, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, c->stream));   // warm
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; r++) HIPCHK(hipMemcpyAsync(b.p, a.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return MRMG_OK;
}

int mrmg_footprint(mrmg_ctx* c, size_t* bytes)
{
    if (!c || !bytes) return set_err(MRMG_ERR_ARG, "null argument");
    *bytes = c->eng->footprint();
    return MRMG_OK;
}
