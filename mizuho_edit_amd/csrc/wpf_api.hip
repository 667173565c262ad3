// wpf_api.hip -- the C ABI of libwpf_hip.so (declared in include/wpf.h).
//
// Each entry point mirrors one procedure of module wPFAW3D; the file:line of
// the reference procedure is given at each definition.  State lives on the
// device; host copies happen only in wpf_upload / wpf_download /
// wpf_get_pressure / wpf_diagnostics.
#include "wpf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

namespace wpf {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int reducer_init(Reducer& r, int slots, int64_t cap, int64_t ring_n)
{
    r.slots = slots;
    r.cap = cap;
    r.hring_n = ring_n;
    WPF_HIP(hipHostMalloc((void**)&r.hring, sizeof(double) * 2 * (size_t)ring_n, hipHostMallocCoherent));
    WPF_HIP(hipMalloc((void**)&r.partials, sizeof(double) * (size_t)cap));
    WPF_HIP(hipMalloc((void**)&r.counter, sizeof(unsigned) * 4));
    WPF_HIP(hipMemset(r.counter, 0, sizeof(unsigned) * 4));
    WPF_HIP(hipMalloc((void**)&r.result, sizeof(double) * slots));
    WPF_HIP(hipMemset(r.result, 0, sizeof(double) * slots));
    WPF_HIP(hipHostMalloc((void**)&r.host, sizeof(double) * slots, hipHostMallocDefault));
    return WPF_OK;
}

int64_t vector_block_capacity(const Grid& g) { return ((g.plane + 511) / 512) * g.nz; }

double finalize_on_host(const double* partials, int64_t nb)
{
    double total = 0.0;
    for (int w = 0; w < 16; w++) {
        double v[64];
        for (int l = 0; l < 64; l++) {
            double s = 0.0;
            for (int64_t b = 64 * w + l; b < nb; b += 1024) s += partials[b];
            v[l] = s;
        }
        for (int o = 32; o > 0; o >>= 1) {
            double u[64];
            for (int l = 0; l < 64; l++) u[l] = v[l] + v[l ^ o];
            for (int l = 0; l < 64; l++) v[l] = u[l];
        }
        total += v[0];
    }
    return total;
}

int64_t reduction_capacity(const Grid& g)
{
    // block counts: 64 x 4 x 1 blocks of the per-procedure stencils (the
    // fused step's and the scalar solver's tiles are coarser), the vector
    // z-march kernels (16 bytes per thread, so at least 2 cells, and at least
    // one plane per tile), and the capped 1-D grids
    const int64_t b64x4 = (int64_t)((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz;
    const int64_t bvec = vector_block_capacity(g);
    const int64_t nb = std::max<int64_t>({(int64_t)2048, b64x4, bvec});
    return nb * kMaxValuesPerBlock;
}

void reducer_free(Reducer& r)
{
    if (r.partials) (void)hipFree(r.partials);
    if (r.counter) (void)hipFree(r.counter);
    if (r.result) (void)hipFree(r.result);
    if (r.host) (void)hipHostFree(r.host);
    if (r.hring) (void)hipHostFree(r.hring);
    r = Reducer{};
}

static inline int64_t hidx(const wpf_ctx* c, int i, int j, int k)
{
    return (int64_t)i + (int64_t)(c->g.nx + 2) * ((int64_t)j + (int64_t)(c->g.ny + 2) * k);
}

static int dkey_host(double v, unsigned long long* k)
{
    unsigned long long u;
    memcpy(&u, &v, 8);
    *k = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return 0;
}

// streaming-bandwidth probe (wpf_stream_probe)
struct StreamPtrs {
    const float4* in[10];
    float4* out[3];
};
// Each thread moves kProbeU 16-byte vectors per array: a block covers a
// contiguous run of 256 * kProbeU vectors, every load of the run is issued
// before the first use, and the stores are non-temporal (the written arrays
// are not read back by the probe).
static constexpr int kProbeU = 4;
template <int NR, int NW>
__global__ void __launch_bounds__(256) k_stream_probe(StreamPtrs P, int64_t n4)
{
    const int64_t base = (int64_t)blockIdx.x * (256 * kProbeU) + threadIdx.x;
    float4 s[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; u++) s[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int r = 0; r < NR; r++) {
        float4 v[kProbeU];
#pragma unroll
        for (int u = 0; u < kProbeU; u++) {
            const int64_t i = base + 256 * u;
            v[u] = i < n4 ? P.in[r][i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < kProbeU; u++) {
            s[u].x += v[u].x;
            s[u].y += v[u].y;
            s[u].z += v[u].z;
            s[u].w += v[u].w;
        }
    }
#pragma unroll
    for (int w = 0; w < NW; w++)
#pragma unroll
        for (int u = 0; u < kProbeU; u++) {
            const int64_t i = base + 256 * u;
            if (i < n4) {
                float* o = reinterpret_cast<float*>(P.out[w] + i);
                __builtin_nontemporal_store(s[u].x + (float)w, o);
                __builtin_nontemporal_store(s[u].y, o + 1);
                __builtin_nontemporal_store(s[u].z, o + 2);
                __builtin_nontemporal_store(s[u].w, o + 3);
            }
        }
}
template <int NR>
static const void* probe_nw(int nw)
{
    switch (nw) {
    case 1: return (const void*)k_stream_probe<NR, 1>;
    case 2: return (const void*)k_stream_probe<NR, 2>;
    case 3: return (const void*)k_stream_probe<NR, 3>;
    }
    return nullptr;
}
static const void* probe_kernel(int nr, int nw)
{
    switch (nr) {
    case 0: return probe_nw<0>(nw);
    case 1: return probe_nw<1>(nw);
    case 2: return probe_nw<2>(nw);
    case 3: return probe_nw<3>(nw);
    case 4: return probe_nw<4>(nw);
    case 5: return probe_nw<5>(nw);
    case 6: return probe_nw<6>(nw);
    case 7: return probe_nw<7>(nw);
    case 8: return probe_nw<8>(nw);
    case 9: return probe_nw<9>(nw);
    case 10: return probe_nw<10>(nw);
    }
    return nullptr;
}
}  // namespace wpf

using namespace wpf;

extern "C" {

const char* wpf_last_error(void) { return g_err; }

const char* wpf_status_string(int32_t s)
{
    switch (s) {
    case WPF_OK: return "ok";
    case WPF_ERR_ARG: return "invalid argument";
    case WPF_ERR_HIP: return "HIP runtime error";
    case WPF_ERR_NOCONV: return "pressure solve did not converge";
    case WPF_ERR_NAN: return "non-finite residual";
    case WPF_ERR_NODEV: return "no GPU device";
    case WPF_ERR_STATE: return "call out of order";
    case WPF_ERR_NOMEM: return "device allocation failed";
    case WPF_ERR_IO: return "file error";
    default: return "unknown status";
    }
}

int wpf_device_count(int32_t* n)
{
    int d = 0;
    hipError_t e = hipGetDeviceCount(&d);
    if (e != hipSuccess) d = 0;
    if (n) *n = d;
    return d > 0 ? WPF_OK : WPF_ERR_NODEV;
}

// loadConfig's allocation and initial state (wPFAW3D.f90:440-483), the well
// index table (451-461) turned into a per-cell type mask.
int wpf_create(const wpf_params* p, const int32_t* lw_haloed, const wpf_well* wells, int32_t nwell,
               int32_t dev, wpf_ctx** out)
{
    WPF_CHECK(p && out, WPF_ERR_ARG, "wpf_create: null argument");
    WPF_CHECK(p->nx >= 1 && p->ny >= 1 && p->nz >= 1, WPF_ERR_ARG, "wpf_create: bad grid %d %d %d",
              p->nx, p->ny, p->nz);
    WPF_CHECK(nwell >= 0 && (nwell == 0 || wells), WPF_ERR_ARG, "wpf_create: bad well table");
    WPF_CHECK((int64_t)p->nx * p->ny * p->nz < ((int64_t)1 << 31), WPF_ERR_ARG,
              "wpf_create: grid too large for 32-bit well indices");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("wpf_create: no GPU visible");
        return WPF_ERR_NODEV;
    }
    if (dev >= 0) WPF_HIP(hipSetDevice(dev));
    wpf_ctx* c = new wpf_ctx();
    WPF_HIP(hipGetDevice(&c->dev));
    c->prm = *p;
    c->g = make_grid(p->nx, p->ny, p->nz);
    WPF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const Grid& g = c->g;
    const size_t nb = sizeof(double) * (size_t)g.n;
    c->wells.assign(wells, wells + nwell);

    // mask and sparse Dirichlet / injection lists from lw (interior cells)
    std::vector<uint8_t> mask(g.n, 0);
    std::vector<int32_t> didx, iidx;
    std::vector<double> dval, ival;
    if (lw_haloed) {
        c->lw_host.assign(lw_haloed, lw_haloed + (size_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2));
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) {
                    const int iw = lw_haloed[hidx(c, i + 1, j + 1, k + 1)];
                    if (iw <= 0) continue;
                    if (iw > nwell) {
                        delete c;
                        set_error("wpf_create: lw refers to well %d of %d", iw, nwell);
                        return WPF_ERR_ARG;
                    }
                    const int64_t cc = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
                    const wpf_well& w = wells[iw - 1];
                    uint8_t m = WPF_MASK_WELL;
                    if (w.pressure != 0.0) {  // onBoundary (1:1106)
                        m |= WPF_MASK_FIXED;
                        didx.push_back((int32_t)cc);
                        dval.push_back(w.pressure);
                    } else if (w.injection != 0.0) {  // onInlet (1:1140)
                        m |= WPF_MASK_INLET;
                        iidx.push_back((int32_t)cc);
                        // b = -q*(dx*dx/dx**3) (1:817); the SPD form carries +q*(...)
                        const double dx = p->dx;
                        ival.push_back(w.injection * (dx * dx / (dx * dx * dx)));
                    }
                    mask[cc] = m;
                }
    }
    int rc;
    WPF_HIP(hipMalloc((void**)&c->mask_raw, g.n + 2 * g.pad));
    WPF_HIP(hipMemset(c->mask_raw, 0, g.n + 2 * g.pad));
    c->mask = c->mask_raw + g.pad;
    WPF_HIP(hipMemcpy(c->mask, mask.data(), g.n, hipMemcpyHostToDevice));
    c->ndir = (int)didx.size();
    c->ninj = (int)iidx.size();
    if (c->ndir) {
        WPF_HIP(hipMalloc((void**)&c->dir_idx, 4 * c->ndir));
        WPF_HIP(hipMalloc((void**)&c->dir_val, 8 * c->ndir));
        WPF_HIP(hipMemcpy(c->dir_idx, didx.data(), 4 * c->ndir, hipMemcpyHostToDevice));
        WPF_HIP(hipMemcpy(c->dir_val, dval.data(), 8 * c->ndir, hipMemcpyHostToDevice));
    }
    if (c->ninj) {
        WPF_HIP(hipMalloc((void**)&c->inj_idx, 4 * c->ninj));
        WPF_HIP(hipMalloc((void**)&c->inj_val, 8 * c->ninj));
        WPF_HIP(hipMemcpy(c->inj_idx, iidx.data(), 4 * c->ninj, hipMemcpyHostToDevice));
        WPF_HIP(hipMemcpy(c->inj_val, ival.data(), 8 * c->ninj, hipMemcpyHostToDevice));
    }
    for (int b = 0; b < 2; b++) {
        WPF_HIP(hipMalloc((void**)&c->phi[b], nb));
        WPF_HIP(hipMalloc((void**)&c->conc[b], nb));
        WPF_HIP(hipMemset(c->phi[b], 0, nb));
        WPF_HIP(hipMemset(c->conc[b], 0, nb));
    }
    WPF_HIP(hipMalloc((void**)&c->krock, nb));
    WPF_HIP(hipMalloc((void**)&c->dphidt, nb));
    WPF_HIP(hipMalloc((void**)&c->dcdt, nb));
    WPF_HIP(hipMemset(c->krock, 0, nb));
    WPF_HIP(hipMemset(c->dphidt, 0, nb));
    WPF_HIP(hipMemset(c->dcdt, 0, nb));
    if ((rc = dev_alloc(c->p, g)) || (rc = dev_alloc(c->tx, g)) || (rc = dev_alloc(c->ty, g)) ||
        (rc = dev_alloc(c->tz, g)) || (rc = dev_alloc(c->invD, g)) || (rc = dev_alloc(c->r, g)) ||
        (rc = dev_alloc(c->z, g)) || (rc = dev_alloc(c->pp, g)) || (rc = dev_alloc(c->q, g)) ||
        (rc = dev_alloc(c->xprev, g))) {
        wpf_destroy(c);
        return rc;
    }
    // initial fields (1:463-483): phase -1, concent 0, pressure pini, wells'
    // own phase/concent; callers normally upload the loaded fields next
    {
        std::vector<double> h(g.n);
        std::fill(h.begin(), h.end(), p->pini);
        WPF_HIP(hipMemcpy(c->p.p, h.data(), nb, hipMemcpyHostToDevice));
        std::vector<double> ph(g.n, -1.0), cc(g.n, 0.0);
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) {
                    if (!lw_haloed) continue;
                    const int iw = lw_haloed[hidx(c, i + 1, j + 1, k + 1)];
                    if (iw <= 0) continue;
                    const int64_t q = i + (int64_t)g.nx * (j + (int64_t)g.ny * k);
                    ph[q] = wells[iw - 1].phase;
                    cc[q] = wells[iw - 1].concent;
                }
        WPF_HIP(hipMemcpy(c->phi[0], ph.data(), nb, hipMemcpyHostToDevice));
        WPF_HIP(hipMemcpy(c->conc[0], cc.data(), nb, hipMemcpyHostToDevice));
    }
    // scalars: time, dtime; taumin keys (3 for the fused step, 1 for the
    // per-procedure call, 5 for diagnostics); breakthrough step
    WPF_HIP(hipMalloc((void**)&c->scal, sizeof(double) * 8));
    WPF_HIP(hipMemset(c->scal, 0, sizeof(double) * 8));
    WPF_HIP(hipMalloc((void**)&c->tau_key, sizeof(unsigned long long) * 16));
    {
        unsigned long long k[16];
        for (int q = 0; q < 16; q++) dkey_host(1.0e6, &k[q]);
        WPF_HIP(hipMemcpy(c->tau_key, k, sizeof(k), hipMemcpyHostToDevice));
    }
    WPF_HIP(hipMalloc((void**)&c->bt_step, sizeof(int32_t)));
    {
        int32_t v = 0x7fffffff;
        WPF_HIP(hipMemcpy(c->bt_step, &v, 4, hipMemcpyHostToDevice));
    }
    c->bt_plane = g.nx >= 3 ? g.nx - 3 : -1;  // ix_g = nx-2
    if ((rc = noalloc_scratch_init(c->dev))) {
        wpf_destroy(c);
        return rc;
    }
    if ((rc = reducer_init(c->red, 64, reduction_capacity(g), vector_block_capacity(g)))) {
        wpf_destroy(c);
        return rc;
    }
    // FP32 operator scale: 1 / (largest possible face coefficient)
    {
        double kmax = p->permeab_hole;
        c->t_scale = (kmax > 0 && p->viscosity > 0) ? p->viscosity / kmax : 1.0;
    }
    *out = c;
    return WPF_OK;
}

int wpf_destroy(wpf_ctx* c)
{
    if (!c) return WPF_OK;
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_learned(c);
    free_solver(c);
    auto F = [](void* q) { if (q) (void)hipFree(q); };
    F(c->mask_raw); F(c->dir_idx); F(c->dir_val); F(c->inj_idx); F(c->inj_val);
    for (int b = 0; b < 2; b++) { F(c->phi[b]); F(c->conc[b]); }
    F(c->krock); F(c->dphidt); F(c->dcdt); F(c->vx); F(c->vy); F(c->vz);
    F(c->scal); F(c->tau_key); F(c->bt_step);
    F(c->cap_fields); F(c->cap_target); F(c->cap_guess);
    for (int j = 0; j < wpf_ctx::kFcapMax; j++) {
        F(c->fcap_fields[j]); F(c->fcap_x[j]); F(c->fcap_target[j]);
    }
    dev_free(c->p); dev_free(c->tx); dev_free(c->ty); dev_free(c->tz); dev_free(c->invD);
    dev_free(c->r); dev_free(c->z); dev_free(c->pp); dev_free(c->q); dev_free(c->xprev);
    reducer_free(c->red);
    for (auto& e : c->prof_ev) (void)hipEventDestroy(e);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    for (auto& e : c->ev_ring)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WPF_OK;
}

static double* field_ptr(wpf_ctx* c, int f)
{
    switch (f) {
    case WPF_FIELD_PERMEAB_ROCK: return c->krock;
    case WPF_FIELD_PRESSURE: return c->p.p;
    case WPF_FIELD_PHASE: return c->phi[c->cur];
    case WPF_FIELD_CONCENT: return c->conc[c->cur];
    case WPF_FIELD_DPHASEDT: return c->dphidt;
    case WPF_FIELD_DCONCENTDT: return c->dcdt;
    default: return nullptr;
    }
}

int wpf_upload(wpf_ctx* c, int32_t field, const double* h)
{
    WPF_CHECK(c && h, WPF_ERR_ARG, "wpf_upload: null argument");
    double* d = field_ptr(c, field);
    WPF_CHECK(d, WPF_ERR_ARG, "wpf_upload: field %d cannot be uploaded", field);
    WPF_HIP(hipSetDevice(c->dev));
    const Grid& g = c->g;
    std::vector<double> tmp(g.n);
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++) {
            const double* src = h + hidx(c, 1, j + 1, k + 1);
            memcpy(&tmp[(size_t)g.nx * (j + (size_t)g.ny * k)], src, sizeof(double) * g.nx);
        }
    WPF_HIP(hipMemcpyAsync(d, tmp.data(), sizeof(double) * g.n, hipMemcpyHostToDevice, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    if (field == WPF_FIELD_PRESSURE) {
        c->proj_count = 0;  // earlier solution changes no longer apply
        c->push_pending = false;
    }
    if (field == WPF_FIELD_PERMEAB_ROCK) {
        double m = 0.0;
        for (double v : tmp) m = std::max(m, v);
        c->krock_max = m;
    }
    if (field == WPF_FIELD_PHASE || field == WPF_FIELD_PERMEAB_ROCK) perm_stale(c);
    if (field == WPF_FIELD_PRESSURE) c->have_velocity = false;
    return WPF_OK;
}

// Fill a haloed host array from interior values with the clamp rule.
static void fill_haloed(const wpf_ctx* c, const double* interior, double* h)
{
    const Grid& g = c->g;
    for (int k = 0; k <= g.nz + 1; k++) {
        const int kk = std::min(std::max(k - 1, 0), g.nz - 1);
        for (int j = 0; j <= g.ny + 1; j++) {
            const int jj = std::min(std::max(j - 1, 0), g.ny - 1);
            const double* row = interior + (size_t)g.nx * (jj + (size_t)g.ny * kk);
            double* o = h + hidx(c, 0, j, k);
            o[0] = row[0];
            memcpy(o + 1, row, sizeof(double) * g.nx);
            o[g.nx + 1] = row[g.nx - 1];
        }
    }
}

int wpf_download(wpf_ctx* c, int32_t field, double* h)
{
    WPF_CHECK(c && h, WPF_ERR_ARG, "wpf_download: null argument");
    WPF_HIP(hipSetDevice(c->dev));
    const Grid& g = c->g;
    const size_t nb = sizeof(double) * (size_t)g.n;
    const int64_t nh = (int64_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    if (field == WPF_FIELD_FACE_T || field == WPF_FIELD_FACE_V) {
        // the device's own +x/+y/+z face arrays of the interior cells
        const bool t = field == WPF_FIELD_FACE_T;
        WPF_CHECK(t ? c->have_perm : c->have_velocity, WPF_ERR_STATE,
                  t ? "face coefficients before calcPermeability" : "face velocities before calcVelocity");
        const double* src[3] = {t ? c->tx.p : c->vx, t ? c->ty.p : c->vy, t ? c->tz.p : c->vz};
        std::vector<double> a(3 * (size_t)g.n);
        for (int d = 0; d < 3; d++)
            WPF_HIP(hipMemcpyAsync(a.data() + d * (size_t)g.n, src[d], nb, hipMemcpyDeviceToHost, c->stream));
        WPF_HIP(hipStreamSynchronize(c->stream));
        memset(h, 0, sizeof(double) * 3 * nh);
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) {
                    const size_t q = (size_t)i + (size_t)g.nx * (j + (size_t)g.ny * k);
                    const int64_t o = hidx(c, i + 1, j + 1, k + 1);
                    for (int d = 0; d < 3; d++) h[3 * o + d] = a[d * (size_t)g.n + q];
                }
        return WPF_OK;
    }
    if (field == WPF_FIELD_PERMEAB || field == WPF_FIELD_VELOCITY) {
        // permeab(0:3,...) and velocity(1:3,...) in the reference's AoS layout:
        // node value and faces 0..n, with the face below cell 1 and halo rows
        // filled as calcPermeability / calcVelocity compute them on halos.
        std::vector<double> phi(g.n), kr(g.n), p(g.n), tx(g.n), ty(g.n), tz(g.n);
        WPF_HIP(hipMemcpyAsync(phi.data(), c->phi[c->cur], nb, hipMemcpyDeviceToHost, c->stream));
        WPF_HIP(hipMemcpyAsync(kr.data(), c->krock, nb, hipMemcpyDeviceToHost, c->stream));
        WPF_HIP(hipMemcpyAsync(p.data(), c->p.p, nb, hipMemcpyDeviceToHost, c->stream));
        WPF_HIP(hipStreamSynchronize(c->stream));
        std::vector<double> ph(nh), krh(nh), ph_p(nh);
        fill_haloed(c, phi.data(), ph.data());
        fill_haloed(c, kr.data(), krh.data());
        fill_haloed(c, p.data(), ph_p.data());
        const double kwh = c->prm.permeab_hole, mu = c->prm.viscosity, dx = c->prm.dx;
        std::vector<double> k0(nh);
        for (int64_t q = 0; q < nh; q++) {
            const double f = 0.5 * (1.0 + ph[q]);
            k0[q] = kwh * f + krh[q] * (1.0 - f);
        }
        const int comps = field == WPF_FIELD_PERMEAB ? 4 : 3;
        memset(h, 0, sizeof(double) * comps * nh);
        auto hm = [](double a, double b) { return 2.0 / (1.0 / a + 1.0 / b); };
        for (int k = 0; k <= g.nz + 1; k++)
            for (int j = 0; j <= g.ny + 1; j++)
                for (int i = 0; i <= g.nx + 1; i++) {
                    const int64_t q = hidx(c, i, j, k);
                    if (field == WPF_FIELD_PERMEAB) h[4 * q] = k0[q];
                    if (i > g.nx || j > g.ny || k > g.nz) continue;
                    const double kx = hm(k0[q], k0[hidx(c, i + 1, j, k)]);
                    const double ky = hm(k0[q], k0[hidx(c, i, j + 1, k)]);
                    const double kz = hm(k0[q], k0[hidx(c, i, j, k + 1)]);
                    if (field == WPF_FIELD_PERMEAB) {
                        h[4 * q + 1] = kx; h[4 * q + 2] = ky; h[4 * q + 3] = kz;
                    } else {
                        const double p0 = ph_p[q];
                        h[3 * q + 0] = -(kx / mu * ((ph_p[hidx(c, i + 1, j, k)] - p0) / dx));
                        h[3 * q + 1] = -(ky / mu * ((ph_p[hidx(c, i, j + 1, k)] - p0) / dx));
                        h[3 * q + 2] = -(kz / mu * ((ph_p[hidx(c, i, j, k + 1)] - p0) / dx));
                    }
                }
        return WPF_OK;
    }
    double* d = field_ptr(c, field);
    WPF_CHECK(d, WPF_ERR_ARG, "wpf_download: unknown field %d", field);
    std::vector<double> tmp(g.n);
    WPF_HIP(hipMemcpyAsync(tmp.data(), d, nb, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    if (field == WPF_FIELD_DPHASEDT || field == WPF_FIELD_DCONCENTDT) {
        // derivatives are never set on halos (they stay 0 from allocation)
        memset(h, 0, sizeof(double) * nh);
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                memcpy(h + hidx(c, 1, j + 1, k + 1), &tmp[(size_t)g.nx * (j + (size_t)g.ny * k)],
                       sizeof(double) * g.nx);
        return WPF_OK;
    }
    fill_haloed(c, tmp.data(), h);
    return WPF_OK;
}

int wpf_calc_permeability(wpf_ctx* c)  // wPFAW3D.f90:561-599
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    return launch_perm(c, PERM_F64, nullptr);
}

int wpf_solve_pressure(wpf_ctx* c, const wpf_solver_opts* o, wpf_solve_stats* st)  // :603-1160
{
    WPF_CHECK(c && o, WPF_ERR_ARG, "null argument");
    return solve_pressure(c, o, st);
}

int wpf_calc_velocity(wpf_ctx* c)  // :1164-1193
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    return launch_velocity(c);
}

int wpf_calc_phase_deriv(wpf_ctx* c, double* taumin)  // :1197-1475
{
    WPF_CHECK(c && taumin, WPF_ERR_ARG, "null argument");
    return launch_phase_deriv(c, taumin);
}

int wpf_calc_concent_deriv(wpf_ctx* c)  // :1554-1624
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    return launch_concent_deriv(c);
}

int wpf_evolve_phase(wpf_ctx* c, double dtime)  // :1628-1632
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    int rc = launch_evolve(c, 0, dtime);
    perm_stale(c);
    c->dt_hist[1] = c->dt_hist[0];
    c->dt_hist[0] = dtime;
    return rc;
}

int wpf_evolve_concent(wpf_ctx* c, double dtime)  // :1636-1640
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    return launch_evolve(c, 1, dtime);
}

int wpf_update_halo(wpf_ctx* c, int32_t field)  // :1645-1657
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    WPF_CHECK(field >= 0 && field <= WPF_FIELD_DCONCENTDT, WPF_ERR_ARG, "unknown field %d", field);
    return WPF_OK;
}

int wpf_get_pressure(wpf_ctx* c, int32_t ix_g, int32_t iy_g, int32_t iz_g, double* p)  // :2151-2166
{
    WPF_CHECK(c && p, WPF_ERR_ARG, "null argument");
    const Grid& g = c->g;
    if (ix_g < 1 || ix_g > g.nx || iy_g < 1 || iy_g > g.ny || iz_g < 1 || iz_g > g.nz) {
        *p = 0.0;
        return WPF_OK;
    }
    const int64_t q = (ix_g - 1) + (int64_t)g.nx * ((iy_g - 1) + (int64_t)g.ny * (iz_g - 1));
    WPF_HIP(hipMemcpyAsync(p, c->p.p + q, 8, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    return WPF_OK;
}

int wpf_diagnostics(wpf_ctx* c, wpf_diag* d)
{
    WPF_CHECK(c && d, WPF_ERR_ARG, "null argument");
    return launch_diag(c, d);
}

// Loop body of wPFAW3D.f90:2225-2294.
int wpf_step(wpf_ctx* c, const wpf_solver_opts* o, int32_t itime, double* time, double* dtime,
             double* taumin, wpf_solve_stats* st)
{
    WPF_CHECK(c && o && time && dtime && taumin, WPF_ERR_ARG, "null argument");
    WPF_CHECK(itime >= 1, WPF_ERR_ARG, "itime must be >= 1");
    // dt rule (1:2227-2230) on the host copy of the previous taumin
    if (itime >= 3) {
        const wpf_params& p = c->prm;
        *dtime = p.tfac * 0.5 * (p.dx * p.dx) / (p.epsilon * p.epsilon) *
                 std::fmin((p.epsilon * p.epsilon) / p.diffusion, *taumin);
    }
    *time = *time + *dtime;
    // the mixed solver's FP32 operator comes out of the permeability pass
    wpf_solver_opts od;
    solver_defaults(o, &od);
    int rc;
    if (od.solver == WPF_SOLVER_MIXED) {
        PermOut32 P;
        int mode;
        if ((rc = op32_setup(c, &od, &P, &mode))) return rc;
        rc = launch_perm(c, mode, &P);
    } else {
        rc = launch_perm(c, PERM_F64, nullptr);
    }
    if (rc) return rc;
    rc = solve_pressure(c, o, st);
    if (rc && rc != WPF_ERR_NOCONV) return rc;
    const int solve_rc = rc;
    rc = launch_step_fields(c, itime, *dtime, false, (o->flags & WPF_FLAG_F32_STENCIL) != 0);
    if (rc) return rc;
    c->dt_hist[1] = c->dt_hist[0];
    c->dt_hist[0] = *dtime;
    // this step's taumin lives in key (itime % 3)
    unsigned long long k;
    WPF_HIP(hipMemcpyAsync(&k, c->tau_key + (itime % 3), 8, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    memcpy(taumin, &u, 8);
    perm_stale(c);
    c->have_velocity = false;
    return solve_rc;
}

int wpf_profile_start(wpf_ctx* c, int32_t kernel)
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    WPF_CHECK(kernel >= 0 && kernel <= WPF_KERNEL_CG_UPDATE, WPF_ERR_ARG, "unknown kernel %d", kernel);
    if (c->prof_ev.empty()) {
        c->prof_ev.resize(2 * 8192);
        for (auto& e : c->prof_ev) WPF_HIP(hipEventCreate(&e));
    }
    WPF_HIP(hipStreamSynchronize(c->stream));
    c->prof_target = kernel;
    c->prof_used = 0;
    c->prof_open = false;
    return WPF_OK;
}

int wpf_profile_stop(wpf_ctx* c, double* total_ms, int32_t* count)
{
    WPF_CHECK(c && total_ms && count, WPF_ERR_ARG, "null argument");
    WPF_HIP(hipStreamSynchronize(c->stream));
    double sum = 0.0;
    for (int i = 0; i < c->prof_used; i++) {
        float ms = 0.0f;
        WPF_HIP(hipEventElapsedTime(&ms, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]));
        sum += ms;
    }
    *total_ms = sum;
    *count = c->prof_used;
    c->prof_target = -1;
    c->prof_used = 0;
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// streaming-bandwidth probe
// ---------------------------------------------------------------------------

int wpf_stream_probe(int32_t device, int32_t nread, int32_t nwrite, int64_t n, int32_t reps, double* us)
{
    WPF_CHECK(us && nread >= 0 && nread <= 10 && nwrite >= 1 && nwrite <= 3 && n > 0 && n % 4 == 0 &&
                  reps > 0,
              WPF_ERR_ARG, "bad probe arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device");
        return WPF_ERR_NODEV;
    }
    WPF_CHECK(device >= 0 && device < ndev, WPF_ERR_ARG, "no device %d", device);
    WPF_HIP(hipSetDevice(device));
    StreamPtrs P{};
    std::vector<void*> bufs;
    int rc = WPF_OK;
    for (int a = 0; a < nread + nwrite && rc == WPF_OK; a++) {
        void* b = nullptr;
        if (hipMalloc(&b, sizeof(float) * (size_t)n) != hipSuccess ||
            hipMemset(b, 0, sizeof(float) * (size_t)n) != hipSuccess) {
            if (b) (void)hipFree(b);
            set_error("probe allocation failed");
            rc = WPF_ERR_HIP;
            break;
        }
        bufs.push_back(b);
        if (a < nread)
            P.in[a] = (const float4*)b;
        else
            P.out[a - nread] = (float4*)b;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == WPF_OK) {
        int64_t n4 = n / 4;
        const void* fn = probe_kernel(nread, nwrite);
        void* args[] = {&P, &n4};
        const dim3 grid((unsigned)((n4 + 256 * kProbeU - 1) / (256 * kProbeU))), block(256);
        float ms = 0.0f;
        bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
        for (int r = 0; ok && r < 3; r++) ok = hipLaunchKernel(fn, grid, block, args, 0, nullptr) == hipSuccess;
        ok = ok && hipEventRecord(e0, nullptr) == hipSuccess;
        for (int r = 0; ok && r < reps; r++) ok = hipLaunchKernel(fn, grid, block, args, 0, nullptr) == hipSuccess;
        ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok)
            *us = 1e3 * ms / reps;
        else {
            set_error("probe launch failed");
            rc = WPF_ERR_HIP;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void* b : bufs) (void)hipFree(b);
    return rc;
}

int wpf_capture_get(wpf_ctx* c, float* fields, float* target)
{
    WPF_CHECK(c && fields && target, WPF_ERR_ARG, "null argument");
    WPF_CHECK(c->have_capture, WPF_ERR_STATE, "no solve with WPF_FLAG_CAPTURE has run");
    const size_t n = (size_t)c->g.n;
    WPF_HIP(hipMemcpyAsync(fields, c->cap_fields, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipMemcpyAsync(target, c->cap_target, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    return WPF_OK;
}

int wpf_capture_fcg_count(wpf_ctx* c, int32_t* count)
{
    WPF_CHECK(c && count, WPF_ERR_ARG, "null argument");
    *count = c->fcap_n;
    return WPF_OK;
}

int wpf_capture_fcg_get(wpf_ctx* c, int32_t j, float* fields, float* target)
{
    WPF_CHECK(c && fields && target, WPF_ERR_ARG, "null argument");
    WPF_CHECK(j >= 0 && j < c->fcap_n, WPF_ERR_ARG, "capture %d of %d", j, c->fcap_n);
    WPF_HIP(hipSetDevice(c->dev));
    const size_t n = (size_t)c->g.n;
    WPF_HIP(hipMemcpyAsync(fields, c->fcap_fields[j], sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipMemcpyAsync(target, c->fcap_target[j], sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    return WPF_OK;
}

int wpf_breakthrough_step(wpf_ctx* c, int32_t* step)
{
    WPF_CHECK(c && step, WPF_ERR_ARG, "null argument");
    int rc = check_breakthrough(c, 0);
    if (rc) return rc;
    *step = c->bt_host;
    return WPF_OK;
}

int wpf_set_breakthrough_plane(wpf_ctx* c, int32_t ix_g)
{
    WPF_CHECK(c, WPF_ERR_ARG, "null context");
    WPF_CHECK(ix_g >= 1 && ix_g <= c->g.nx, WPF_ERR_ARG, "plane %d outside 1..%d", ix_g, c->g.nx);
    c->bt_plane = ix_g - 1;
    int32_t v = 0x7fffffff;
    WPF_HIP(hipMemcpy(c->bt_step, &v, 4, hipMemcpyHostToDevice));
    return WPF_OK;
}

}  // extern "C"
